"""The price_one_option command line (argument handling on the CPU; one
small solve on the GPU)."""

import os
import pickle

import numpy as np
import pytest

from fd_adi_amd import price_one_option as cli


def test_arguments_and_defaults():
    opt = cli.read_args(["-nx", "40", "20", "-nt", "10", "--gpu"])
    assert opt.nx == [40, 20] and opt.nt == 10 and opt.scheme == "hv"
    assert opt.mean_variance == opt.variance == 0.04 and opt.strike == 99.0
    with pytest.raises(SystemExit):
        cli.read_args(["-nx", "40", "20", "-nt", "10", "--cpu"])
    with pytest.raises(SystemExit):
        cli.read_args(["-nx", "40", "20", "-nt", "10", "--top", "150"])


@pytest.mark.gpu
def test_cli_matches_engine(gpu, tmp_path, capsys):
    args = ["-nx", "40", "20", "-nt", "10", "--gpu", "--scheme", "d", "-p", "-0.3",
            "-o", "0.3", "--out-dir", str(tmp_path)]
    assert cli.main(args) == 0
    out = capsys.readouterr().out
    price = float(out.split("FD:")[1].split()[0])
    opt = cli.read_args(args)
    e, p2 = cli.run(opt)
    assert price == p2
    files = os.listdir(tmp_path)
    assert len(files) == 1 and files[0].startswith("gpu_moneyness_")
    assert files[0].endswith(".txt")
    # the reference's file: [grid.mesh, grid.domain[-1]], pickled (written above)
    with open(os.path.join(tmp_path, files[0]), "rb") as f:
        mesh, V = pickle.load(f)
    assert V.shape == (40, 20) and V[e.idx] == price
    # the strike (99) is not S0: S0 is snapped onto the mesh in place; v0 is
    # a node to 10 decimals (the reference's check, price_one_option.py:130-133)
    assert mesh[0][e.idx[0]] == 100.0 and abs(mesh[1][e.idx[1]] - 0.04) < 1e-10
    np.testing.assert_array_equal(mesh[0], e.grid.mesh[0])
    assert cli.main(args + ["--npz"]) == 0
    z = [f for f in os.listdir(tmp_path) if f.endswith(".npz")]
    assert len(z) == 1
    np.testing.assert_array_equal(np.load(os.path.join(tmp_path, z[0]))["V"], V)


def test_bench_rejects_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message (no GPU use)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
