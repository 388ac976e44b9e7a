"""Banded (DIA) finite-difference operator -- oracle restatement.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates FiniteDifference/BandedOperator.pyx for the central-difference,
second-order configuration the Heston GPU engine supports.  Storage follows
scipy's DIA convention, which the reference relies on for
``scipy.linalg.solve_banded``: ``data[k, j] == A[j - offsets[k], j]`` with the
offsets stored in descending order.

An operator acting on axis ``axis`` of a 2-D state is block diagonal: one block
per grid line of that axis, rows ordered so that the operator axis is the
fastest varying one (BandedOperator.pyx:331-361 transposes the state before
``D.dot``).
"""

import numpy as np
import scipy.linalg as spl
import scipy.sparse

FOLDED = "FOLDED"
CAN_FOLD = "CAN_FOLD"
CANNOT_FOLD = "CANNOT_FOLD"


def _index(offsets, o):
    """Row of ``data`` holding diagonal ``o`` (BandedOperator.pyx:853-858)."""
    for k, x in enumerate(offsets):
        if x == o:
            return k
    raise ValueError("Value not in array: %s" % o)


class BandedOperator(object):
    """U2 = L*U1 + R  (apply)  and  U2 = L^-1 (U1 - R)  (solve).

    BandedOperator.pyx:20-72.
    """

    def __init__(self, data, offsets, residual=None, axis=0, derivative=1,
                 deltas=None, order=2):
        data = np.array(data, dtype=float)
        offsets = tuple(int(o) for o in offsets)
        if data.shape[1] <= 3:
            raise AssertionError("Vector too short to use finite differencing.")
        n = data.shape[1]
        self.shape = (n, n)
        self.data = data
        self.offsets = offsets
        if residual is None:
            self.R = np.zeros(n)
        else:
            residual = np.asarray(residual, dtype=float)
            if residual.ndim != 1 or residual.shape[0] != n:
                raise ValueError("Residual vector has wrong shape")
            self.R = residual.copy()
        self.blocks = 1
        self.derivative = derivative
        self.order = order
        self.axis = axis
        self.deltas = np.array([np.nan]) if deltas is None else deltas
        self.dirichlet = [None, None]
        self.top_factors = None
        self.bottom_factors = None
        self.top_fold_status = CAN_FOLD if 2 in offsets else CANNOT_FOLD
        self.bottom_fold_status = CAN_FOLD if -2 in offsets else CANNOT_FOLD
        self.solve_banded_offsets = (abs(min(offsets)), abs(max(offsets)))
        self.is_mixed_derivative = False

    # -- helpers -----------------------------------------------------------
    @property
    def D(self):
        return scipy.sparse.dia_matrix((self.data, self.offsets), shape=self.shape)

    @property
    def block_len(self):
        return self.shape[0] // self.blocks

    def copy(self):
        """BandedOperator.pyx:170-178 (deep copy including metadata)."""
        B = BandedOperator.__new__(BandedOperator)
        B.__dict__.update(self.__dict__)
        B.data = self.data.copy()
        B.R = None if self.R is None else self.R.copy()
        B.dirichlet = list(self.dirichlet)
        B.top_factors = None if self.top_factors is None else self.top_factors.copy()
        B.bottom_factors = None if self.bottom_factors is None else self.bottom_factors.copy()
        return B

    def dense(self):
        return self.D.toarray()

    def is_folded(self):
        return FOLDED in (self.top_fold_status, self.bottom_fold_status)

    def is_foldable(self):
        return CAN_FOLD in (self.top_fold_status, self.bottom_fold_status)

    def is_tridiagonal(self):
        return tuple(self.offsets) == (1, 0, -1)

    def _block_starts(self):
        return np.arange(self.blocks) * self.block_len

    # -- folding (row elimination of the +-2 band entries) -------------------
    def fold_bottom(self, unfold=False):
        """Eliminate A[n-1, n-3] of every block with row n-2.

        BandedOperator.pyx:223-255.  f = -A[n-1,n-3] / A[n-2,n-3];
        row_{n-1} += f * row_{n-2}.
        """
        d = self.data
        m = _index(self.offsets, 0)
        for i in (1, 0, -1, -2):
            if self.offsets[m - i] != i:
                raise ValueError("Operator data is the wrong shape. Requires "
                                 "contiguous offsets (1, 0, -1, -2).")
        last = self._block_starts() + self.block_len - 1
        cn1 = (m - 1, last)          # A[n-2, n-1]
        bn = (m, last)               # A[n-1, n-1]
        bn1 = (m, last - 1)          # A[n-2, n-2]
        an = (m + 1, last - 1)       # A[n-1, n-2]
        an1 = (m + 1, last - 2)      # A[n-2, n-3]
        zn = (m + 2, last - 2)       # A[n-1, n-3]
        if unfold:
            f = self.bottom_factors
            d[zn] -= d[an1] * f
            d[an] -= d[bn1] * f
            d[bn] -= d[cn1] * f
            self.bottom_fold_status = CAN_FOLD
            self.bottom_factors = None
        else:
            a = d[an1]
            safe = np.where(a != 0, a, 1.0)
            f = np.where(a != 0, -d[zn] / safe, 0.0)
            self.bottom_factors = f
            d[zn] += d[an1] * f
            d[an] += d[bn1] * f
            d[bn] += d[cn1] * f
            self.bottom_fold_status = FOLDED

    def fold_top(self, unfold=False):
        """Eliminate A[0, 2] of every block with row 1.

        BandedOperator.pyx:258-290.  f = -A[0,2] / A[1,2]; row_0 += f * row_1.
        """
        d = self.data
        m = _index(self.offsets, 0)
        for i in (2, 1, 0, -1):
            if self.offsets[m - i] != i:
                raise ValueError("Operator data is the wrong shape. Requires "
                                 "contiguous offsets (2, 1, 0, -1).")
        first = self._block_starts()
        a1 = (m + 1, first)          # A[1, 0]
        b0 = (m, first)              # A[0, 0]
        b1 = (m, first + 1)          # A[1, 1]
        c0 = (m - 1, first + 1)      # A[0, 1]
        c1 = (m - 1, first + 2)      # A[1, 2]
        d0 = (m - 2, first + 2)      # A[0, 2]
        if unfold:
            f = self.top_factors
            d[d0] -= d[c1] * f
            d[c0] -= d[b1] * f
            d[b0] -= d[a1] * f
            self.top_fold_status = CAN_FOLD
            self.top_factors = None
        else:
            c = d[c1]
            safe = np.where(c != 0, c, 1.0)
            f = np.where(c != 0, -d[d0] / safe, 0.0)
            self.top_factors = f
            d[d0] += d[c1] * f
            d[c0] += d[b1] * f
            d[b0] += d[a1] * f
            self.top_fold_status = FOLDED

    def diagonalize(self):
        """BandedOperator.pyx:181-198: fold, then drop the +-2 bands."""
        top, bot = 0, len(self.offsets)
        if 2 in self.offsets:
            self.fold_top()
            top += 1
        if -2 in self.offsets:
            self.fold_bottom()
            bot -= 1
        self.data = self.data[top:bot].copy()
        self.offsets = tuple(self.offsets[top:bot])
        self.solve_banded_offsets = (1, 1)

    def undiagonalize(self):
        """BandedOperator.pyx:201-220: back to a 5-band matrix, unfold."""
        data = np.zeros((5, self.shape[0]))
        offsets = (2, 1, 0, -1, -2)
        for k, o in enumerate(self.offsets):
            data[_index(offsets, o)] += self.data[k]
        self.data = data
        self.offsets = offsets
        if self.top_factors is not None:
            self.fold_top(unfold=True)
        if self.bottom_factors is not None:
            self.fold_bottom(unfold=True)
        self.solve_banded_offsets = (2, 2)

    def fold_vector(self, v, unfold=False):
        """BandedOperator.pyx:293-311 (in place on the flat vector)."""
        sign = -1.0 if unfold else 1.0
        s = self._block_starts()
        e = s + self.block_len - 1
        if self.top_factors is not None:
            v[s] += sign * v[s + 1] * self.top_factors
        if self.bottom_factors is not None:
            v[e] += sign * v[e - 1] * self.bottom_factors
        return v

    # -- state <-> operator ordering ------------------------------------------
    def _to_lines(self, V):
        V = np.array(V, dtype=float, copy=True)
        if V.ndim == 1:
            return V, (lambda x: x)
        axes = list(range(V.ndim))
        axes.remove(self.axis)
        axes.append(self.axis)
        T = np.ascontiguousarray(np.transpose(V, axes))
        inv = np.argsort(axes)
        shp = T.shape
        return T, (lambda x: np.ascontiguousarray(np.transpose(x.reshape(shp), inv)))

    def _force_dirichlet(self, T):
        if self.dirichlet[0] is not None:
            T[..., 0] = np.reshape(self.dirichlet[0], T[..., 0].shape) \
                if np.ndim(self.dirichlet[0]) else self.dirichlet[0]
        if self.dirichlet[1] is not None:
            T[..., -1] = np.reshape(self.dirichlet[1], T[..., -1].shape) \
                if np.ndim(self.dirichlet[1]) else self.dirichlet[1]

    def apply(self, V):
        """y = A x + R with Dirichlet values forced into x first.

        BandedOperator.pyx:331-361.
        """
        T, back = self._to_lines(V)
        self._force_dirichlet(T)
        if getattr(self, "apply_order", None) == "reversed":
            ret = self._dot_reversed(T.ravel())
        else:
            ret = self.D.dot(T.ravel())
        if self.is_folded():
            ret = self.fold_vector(ret, unfold=True)
        if self.R is not None:
            ret = ret + self.R
        c = getattr(self, "post_scale", None)
        if c is not None:
            ret = c * ret
        return back(ret)

    def solve(self, V):
        """x = A^-1 (v - R) with Dirichlet forcing and RHS folding.

        BandedOperator.pyx:364-394 (LAPACK gbsv through solve_banded).
        """
        T, back = self._to_lines(V)
        self._force_dirichlet(T)
        rhs = T.ravel()
        if self.R is not None:
            rhs = rhs - self.R
        if self.is_folded():
            rhs = self.fold_vector(rhs)
        solver = getattr(self, "solver", "lapack")
        if solver == "thomas" and self.solve_banded_offsets == (1, 1):
            ret = self._thomas(rhs)
        elif solver == "chunked" and self.solve_banded_offsets == (1, 1):
            ret = self._chunked(rhs)
        else:
            ret = spl.solve_banded(self.solve_banded_offsets, self.data, rhs)
        return back(ret)

    def _dot_reversed(self, x):
        """A x summing the diagonals in the reverse of dia_matrix order (a
        rounding-order probe for tests, like _thomas)."""
        y = np.zeros_like(x)
        n = x.shape[0]
        for k in range(len(self.offsets) - 1, -1, -1):
            o = int(self.offsets[k])
            band = self.data[k]
            # y[i] += A[i, i+o] x[i+o]; DIA stores A[i, i+o] at column j = i + o
            if o >= 0:
                y[:n - o] += band[o:] * x[o:]
            else:
                y[-o:] += band[:n + o] * x[:n + o]
        return y

    def _thomas(self, rhs):
        """Per-block Thomas elimination (no pivoting) of the tridiagonal
        system, vectorised over the blocks.  Not the reference's algorithm: it
        exists so that tests can measure how much the rounding order of the
        implicit solves alone moves a result (the reference solves with LAPACK
        gbsv, whose operation order differs from any GPU solve)."""
        nb, L = self.blocks, self.block_len
        sup = self.data[0].reshape(nb, L)
        b = self.data[1].reshape(nb, L)
        sub = self.data[2].reshape(nb, L)
        d = rhs.reshape(nb, L)
        cp = np.zeros((nb, L))
        dp = np.zeros((nb, L))
        cp[:, 0] = sup[:, 1] / b[:, 0] if L > 1 else 0.0
        dp[:, 0] = d[:, 0] / b[:, 0]
        for i in range(1, L):
            a = sub[:, i - 1]
            m = b[:, i] - a * cp[:, i - 1]
            if i < L - 1:
                cp[:, i] = sup[:, i + 1] / m
            dp[:, i] = (d[:, i] - a * dp[:, i - 1]) / m
        x = np.empty((nb, L))
        x[:, L - 1] = dp[:, L - 1]
        for i in range(L - 2, -1, -1):
            x[:, i] = dp[:, i] - cp[:, i] * x[:, i + 1]
        return x.ravel()

    def _chunked(self, rhs, chunk=16):
        """The same elimination (no pivoting) associated the way a parallel
        line solve does it -- another rounding-order probe, like _thomas:
        pivots rd_i = 1 / (b_i - (a_i rd_{i-1}) c_{i-1}); the forward sweep in
        scaled form y_i = rd_i d_i + mu_i y_{i-1} (mu_i = -a_i rd_i) and the
        backward sweep x_i = y_i + n_i x_{i+1} (n_i = -c_i rd_i), each split
        into chunks of ``chunk`` rows: a chunk's affine map y -> E + M y is
        formed with a zero carry, the maps are composed chunk to chunk to give
        each chunk its incoming carry, and the chunk is run again from it."""
        nb, L = self.blocks, self.block_len
        c = self.data[0].reshape(nb, L)      # c[:, i] multiplies x[i] in row i-1
        b = self.data[1].reshape(nb, L)
        a = self.data[2].reshape(nb, L)      # a[:, i] multiplies x[i] in row i+1
        d = rhs.reshape(nb, L)
        sub = np.zeros((nb, L))
        sub[:, 1:] = a[:, :-1]
        sup = np.zeros((nb, L))
        sup[:, :-1] = c[:, 1:]
        rd = np.empty((nb, L))
        prev_rd = np.zeros(nb)
        prev_c = np.zeros(nb)
        for i in range(L):
            rd[:, i] = 1.0 / (b[:, i] - (sub[:, i] * prev_rd) * prev_c)
            prev_rd, prev_c = rd[:, i], sup[:, i]
        mu = -sub * rd
        nn = -sup * rd
        ds = rd * d

        def sweep(m, f, order):
            starts = list(range(0, L, chunk))
            carry = np.zeros(nb)
            out = np.empty((nb, L))
            for s0 in (starts if order > 0 else starts[::-1]):
                rows = list(range(s0, min(s0 + chunk, L)))
                if order < 0:
                    rows = rows[::-1]
                E, M = np.zeros(nb), np.ones(nb)
                for i in rows:
                    E = f[:, i] + m[:, i] * E
                    M = m[:, i] * M
                y = carry
                for i in rows:
                    y = f[:, i] + m[:, i] * y
                    out[:, i] = y
                carry = E + M * carry
            return out

        y = sweep(mu, ds, +1)
        x = sweep(nn, y, -1)
        return x.ravel()

    # -- boundaries --------------------------------------------------------------
    def applyboundary(self, boundary, mesh_deltas=None):
        """Write the boundary rows of a single-block operator.

        BandedOperator.pyx:398-524.  ``boundary`` is ((lower_type, lower_val),
        (upper_type, upper_val)); type 0 = Dirichlet, 1 = Von Neumann (value
        goes to R), None = free boundary (one-sided stencil; for the second
        derivative a known first derivative ``val`` is used when given).
        """
        B = self.data
        R = self.R
        m = _index(self.offsets, 0)
        d = self.deltas
        (lt, lv), (ut, uv) = boundary
        if lt == 0:
            B[m, 0] = 1
            self.dirichlet[0] = lv
        elif lt == 1:
            R[0] = lv
        elif lt is None and self.derivative == 1:
            B[m - 1, 1] = 1 / d[1]
            B[m, 0] = -1 / d[1]
        elif lt is None and self.derivative == 2:
            if lv is not None:
                x = d[1] * d[1]
                B[m - 1, 1] = 2 / x
                B[m, 0] = -2 / x
                R[0] = -lv * 2 / d[1]
            else:
                rd = 1.0 / (0.5 * (d[2] + d[1]) * d[2] * d[1])
                B[m - 2, 2] = d[1] * rd
                B[m - 1, 1] = -(d[2] + d[1]) * rd
                B[m, 0] = d[2] * rd
        else:
            raise NotImplementedError("boundary derivative order")
        if ut == 0:
            B[m, -1] = 1
            self.dirichlet[1] = uv
        elif ut == 1:
            R[-1] = uv
        elif ut is None and self.derivative == 1:
            B[m, -1] = 1.0 / d[-1]
            B[m + 1, -2] = -1.0 / d[-1]
        elif ut is None and self.derivative == 2:
            if uv is not None:
                x = d[-1] * d[-1]
                B[m, -1] = -2 / x
                B[m + 1, -2] = 2 / x
                R[-1] = (uv * 2) / d[-1]
            else:
                rd = 1.0 / (0.5 * (d[-2] + d[-1]) * d[-2] * d[-1])
                B[m + 2, -3] = d[-1] * rd
                B[m + 1, -2] = -(d[-2] + d[-1]) * rd
                B[m, -1] = d[-2] * rd
        else:
            raise NotImplementedError("boundary derivative order")

    # -- algebra -----------------------------------------------------------------
    def _interior_mask(self):
        """Rows that are not Dirichlet rows (main diagonal updates skip them,
        BandedOperator.pyx:684-695, 735-742)."""
        mask = np.ones(self.shape[0], dtype=bool)
        s = self._block_starts()
        if self.dirichlet[0] is not None:
            mask[s] = False
        if self.dirichlet[1] is not None:
            mask[s + self.block_len - 1] = False
        return mask

    def add_operator(self, other, inplace=False):
        """BandedOperator.pyx:618-710.  Residuals add; the other's main
        diagonal is not added on this operator's Dirichlet rows."""
        if self.axis != other.axis:
            raise ValueError("Both operators must operate on the same axis.")
        if self.shape[1] != other.shape[1]:
            raise ValueError("Both operators must have the same length")
        if self.is_folded() or other.is_folded():
            raise ValueError("Cannot add diagonalized operators.")
        if inplace:
            if tuple(self.offsets) != tuple(other.offsets):
                raise ValueError("Both operators must have (exactly) the"
                                 " same offsets to add in-place.")
            B = self
        else:
            offs = tuple(sorted(set(self.offsets) | set(other.offsets), reverse=True))
            data = np.zeros((len(offs), self.shape[1]))
            for k, o in enumerate(self.offsets):
                data[_index(offs, o)] += self.data[k]
            B = self.copy()
            B.data = data
            B.offsets = offs
            B.solve_banded_offsets = (abs(min(offs)), abs(max(offs)))
        mask = B._interior_mask()
        for k, o in enumerate(other.offsets):
            to = _index(B.offsets, o)
            if o == 0:
                B.data[to, mask] += other.data[k, mask]
            else:
                B.data[to] += other.data[k]
        if other.R is not None:
            B.R = other.R.copy() if B.R is None else B.R + other.R
        if CAN_FOLD in (other.top_fold_status, self.top_fold_status):
            B.top_fold_status = CAN_FOLD
        if CAN_FOLD in (other.bottom_fold_status, self.bottom_fold_status):
            B.bottom_fold_status = CAN_FOLD
        return B

    def add_scalar(self, val, inplace=False):
        """BandedOperator.pyx:713-745: main diagonal, Dirichlet rows skipped."""
        B = self if inplace else self.copy()
        m = _index(B.offsets, 0)
        B.data[m, B._interior_mask()] += val
        return B

    def add(self, val, inplace=False):
        if isinstance(val, BandedOperator):
            return self.add_operator(val, inplace)
        return self.add_scalar(val, inplace)

    def __add__(self, val):
        return self.add(val, False)

    def __iadd__(self, val):
        return self.add(val, True)

    def vectorized_scale(self, vector):
        """Row scaling by ``vector`` (BandedOperator.pyx:749-792).

        The vector is forced to 1 on Dirichlet rows; R is scaled as well.
        """
        n = self.shape[0]
        vector = np.array(vector, dtype=float, copy=True)
        if self.blocks > 1 and vector.shape[0] == self.block_len:
            vector = np.tile(vector, self.blocks)
        if vector.shape[0] != n:
            raise AssertionError("scaling vector has the wrong length")
        s = self._block_starts()
        if self.dirichlet[0] is not None:
            vector[s] = 1
        if self.dirichlet[1] is not None:
            vector[s + self.block_len - 1] = 1
        for k, o in enumerate(self.offsets):
            if o >= 0:
                self.data[k, o:] *= vector[:n - o]
            else:
                self.data[k, :n + o] *= vector[-o:]
        self.R = self.R * vector
        # Fold factors are the eliminated band entries' multipliers of a
        # row combination; scaling a folded operator is undefined upstream.

    def mul(self, val, inplace=False):
        B = self if inplace else self.copy()
        B.vectorized_scale(np.ones(self.shape[0]) * val)
        return B

    def __mul__(self, val):
        return self.mul(val, False)

    def __imul__(self, val):
        return self.mul(val, True)


def centercoeffs(deltas, derivative=1, force_bandwidth=None):
    """Central nonuniform stencils for interior rows 1..n-2.

    BandedOperator.pyx:970-1017.  With h- = d[i], h+ = d[i+1]:
      first:  sub = -h+/(h-(h-+h+)), mid = (h+-h-)/(h- h+), sup = h-/(h+(h-+h+))
      second: sub = 2/(h-(h-+h+)),   mid = -2/(h- h+),      sup = 2/(h+(h-+h+))
    """
    d = deltas
    if force_bandwidth is not None:
        lo, hi = (int(o) for o in force_bandwidth)
        offsets = tuple(range(hi, lo - 1, -1))
    else:
        offsets = (1, 0, -1)
    n = len(d)
    data = np.zeros((len(offsets), n))
    m = _index(offsets, 0)
    i = np.arange(1, n - 1)
    hm, hp = d[i], d[i + 1]
    x = hm + hp
    if derivative == 1:
        data[m - 1, i + 1] = hm / (hp * x)
        data[m, i] = (-hm + hp) / (hm * hp)
        data[m + 1, i - 1] = -hp / (hm * x)
    elif derivative == 2:
        data[m - 1, i + 1] = 2 / (hp * x)
        data[m, i] = -2 / (hm * hp)
        data[m + 1, i - 1] = 2 / (hm * x)
    else:
        raise NotImplementedError("Derivative must be 1 or 2")
    return data, offsets


def for_vector(vector, derivative=1, force_bandwidth=None, axis=0, residual=None):
    """BandedOperator.pyx:886-914 for the central scheme."""
    deltas = np.hstack((np.nan, np.diff(vector)))
    data, offsets = centercoeffs(deltas, derivative, force_bandwidth)
    B = BandedOperator(data, offsets, residual=residual, axis=axis,
                       derivative=derivative, deltas=deltas)
    return B


def flatten_blocks(mats):
    """Stack identical-offset single-block operators into one block-diagonal
    operator (FiniteDifferenceEngine.pyx:1022-1038)."""
    offs = tuple(mats[0].offsets)
    assert all(tuple(m.offsets) == offs for m in mats)
    B = mats[0].copy()
    B.data = np.hstack([m.data for m in mats])
    B.R = np.hstack([m.R for m in mats])
    n = B.data.shape[1]
    B.shape = (n, n)
    B.blocks = len(mats)
    if len(mats) > 1:
        B.dirichlet = [None if mats[0].dirichlet[k] is None
                       else tuple(m.dirichlet[k] for m in mats) for k in (0, 1)]
    return B


def mixed_operator(mesh0, mesh1):
    """Cross derivative d2/dx0 dx1 on the (n0, n1) grid, axis-1-fastest rows.

    The reference builds it from shifted block copies
    (FiniteDifference/FiniteDifferenceEngine.pyx:502-545); the result equals the
    Kronecker product of the two central first-derivative stencils restricted
    to interior rows, which is how the reference test writes its closed form
    (FiniteDifference/tests/test_BandedOperatorGPU.py:349-420).
    """
    n0, n1 = len(mesh0), len(mesh1)
    d0 = np.hstack((np.nan, np.diff(mesh0)))
    d1 = np.hstack((np.nan, np.diff(mesh1)))

    def weights(d, n):
        i = np.arange(1, n - 1)
        hm, hp = d[i], d[i + 1]
        x = hm + hp
        return (-hp / (hm * x), (-hm + hp) / (hm * hp), hm / (hp * x))

    w0 = weights(d0, n0)   # (sub, mid, sup) for rows 1..n0-2
    w1 = weights(d1, n1)
    N = n0 * n1
    offsets = []
    diags = []
    for a in (1, 0, -1):          # spot neighbour offset (row + a)
        for b in (1, 0, -1):      # var neighbour offset
            o = a * n1 + b
            band = np.zeros(N)
            ws = w0[a + 1]
            wv = w1[b + 1]
            rows = (np.arange(1, n0 - 1)[:, None] * n1 + np.arange(1, n1 - 1)[None, :])
            vals = ws[:, None] * wv[None, :]
            cols = rows + o
            band[cols.ravel()] = vals.ravel()   # DIA: data[k, col] = A[col-o, col]
            offsets.append(o)
            diags.append(band)
    order = np.argsort(offsets)[::-1]
    data = np.array(diags)[order]
    offs = tuple(int(offsets[k]) for k in order)
    B = BandedOperator.__new__(BandedOperator)
    B.shape = (N, N)
    B.data = data
    B.offsets = offs
    B.R = np.zeros(N)
    B.blocks = 1
    B.derivative = (0, 1)
    B.order = 2
    B.axis = 1
    B.deltas = np.array([np.nan])
    B.dirichlet = [None, None]
    B.top_factors = None
    B.bottom_factors = None
    B.top_fold_status = CANNOT_FOLD
    B.bottom_fold_status = CANNOT_FOLD
    B.solve_banded_offsets = None
    B.is_mixed_derivative = True
    return B
