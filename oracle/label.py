"""Read-out of a value grid at a point off the nodes -- TEST INFRASTRUCTURE ONLY.

The surrogate's labels are the ADI value grid read at (S0, v0), which lie
between nodes of the label meshes.  The read-out is tensor-product cubic
Lagrange interpolation on the 4 x 4 nodes around the point: along each axis
the stencil is nodes i-2 .. i+1 where i = searchsorted(mesh, x) (the first
node >= x), shifted to stay inside the mesh.  The reference has no such
read-out (its price is the value at a node, FiniteDifferenceEngineGPU.pyx:
727-738); this is the checker's statement of the one the product uses.

Pinned by: exactness for every bicubic polynomial on nonuniform meshes (the
4-point interpolant is the unique cubic through the four nodes, so
exactness on cubics and linearity fix it for any data), the value at a node,
and the stencil choice (tests/test_oracle_label.py).
"""

import numpy as np


def stencil_start(mesh, x):
    """First node of the 4-point stencil around x."""
    n = len(mesh)
    return int(min(max(np.searchsorted(mesh, x) - 2, 0), n - 4))


def lagrange4(x, nodes):
    """Weights of the cubic through the four ``nodes`` evaluated at x."""
    w = np.ones(4)
    for i in range(4):
        for j in range(4):
            if i != j:
                w[i] *= (x - nodes[j]) / (nodes[i] - nodes[j])
    return w


def bicubic_at(V, spots, vars_, S0, v0):
    """V [nS, nv] on meshes (spots, vars_) read at (S0, v0)."""
    V = np.asarray(V, dtype=float)
    i = stencil_start(spots, S0)
    j = stencil_start(vars_, v0)
    ws = lagrange4(S0, np.asarray(spots[i:i + 4], dtype=float))
    wv = lagrange4(v0, np.asarray(vars_[j:j + 4], dtype=float))
    return float(ws.dot(V[i:i + 4, j:j + 4]).dot(wv))
