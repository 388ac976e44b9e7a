"""Per-case parity table from a GPU test run's parity log.

    FDADI_PARITY_LOG=parity.jsonl python -m pytest tests -m gpu
    python tests/parity_table.py parity.jsonl > tests/parity_cases.md

Each line of the log is one comparison recorded by tests/parity.py: the GPU
grid's max error against the oracle, the oracle's own rounding-order floor,
the tolerance applied, their ratio and the relative error at the price node.
"""

import json
import sys


def main(path):
    rows = [json.loads(l) for l in open(path) if l.strip()]
    print("| case | grid err | oracle floor | err / floor | tolerance | price rel err |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        name = r.get("case") or r["test"].split("::", 1)[-1]
        if r.get("case"):
            name = "%s [%s]" % (r["test"].split("::", 1)[-1], r["case"])
        ratio = "%.2f" % r["ratio"] if r.get("ratio") is not None else "-"
        print("| %s | %.2e | %.2e | %s | %.2e | %.1e |" % (
            name, r["err"], r["floor"], ratio, r["tol"], r["price_rel"]))
    ratios = [r["ratio"] for r in rows if r.get("ratio") is not None and r["err"] > 1e-10]
    prel = [r["price_rel"] for r in rows if not r.get("case", "").startswith("label")]
    print()
    print("%d comparisons; largest err / floor where err > 1e-10: %.2f; largest price rel "
          "err (node prices): %.1e" % (len(rows), max(ratios) if ratios else 0.0, max(prel)))


if __name__ == "__main__":
    main(sys.argv[1])
