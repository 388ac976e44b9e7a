"""Tensor grid holding the solution history (FiniteDifference/Grid.py:17-52)."""

import numpy as np


class Grid(object):
    """n-dimensional grid from monotonic 1-D meshes; ``domain[-1]`` is the
    current state with shape tuple(len(m) for m in mesh) (last axis contiguous)."""

    def __init__(self, mesh=(np.linspace(0, 1, 5), np.linspace(0, 1, 5)),
                 initializer=lambda x, y: x * 0):
        mesh = tuple(np.asarray(m, dtype=float) for m in mesh)
        for m in mesh:
            if len(m) > 1:
                s = np.sign(np.diff(m))
                if not np.all(s == s[0]):
                    raise AssertionError("mesh is not monotonic")
        self._mesh = mesh
        self.domain = []
        self.ndim = len(mesh)
        self.shape = tuple(len(m) for m in mesh)
        if len(mesh) == 1:
            self.domain.append(initializer(*mesh))
        else:
            self.domain.append(initializer(*(x.T for x in np.meshgrid(*mesh))))
        if np.isscalar(self.domain[-1]):
            self.domain[-1] = self.domain[-1] * np.ones(self.shape)
        self.domain[-1] = np.asarray(self.domain[-1], dtype=float)
        self.dx = [np.hstack((np.nan, np.diff(m))) for m in mesh]
        self.initializer = initializer

    @property
    def mesh(self):
        return self._mesh

    @property
    def size(self):
        return int(np.prod(self.shape))

    def reset(self):
        self.__init__(mesh=self.mesh, initializer=self.initializer)

    def copy(self):
        G = Grid(mesh=self._mesh, initializer=self.initializer)
        G.domain = [d.copy() for d in self.domain]
        return G

    def max(self, dim):
        return self.mesh[dim][-1]

    def min(self, dim):
        return self.mesh[dim][0]

    def __getitem__(self, idx):
        return self.domain[-1][idx]

    def __setitem__(self, idx, val):
        self.domain[-1][idx] = val
