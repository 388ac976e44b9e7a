// Tridiagonal banded operator on one axis of a 2-D (nS, nv) state.
//
// Replaces _TriBandedOperator (_TriBandedOperatorGPU.cu).  The operator is
// block diagonal with one block per grid line; bands are stored row-indexed
// (sup[r] = A[r, r+1], sub[r] = A[r, r-1]) with rows ordered line-major
// (r = line * block_len + pos), exactly the reference's layout, so immigrate
// and vectorized_scale keep their meaning.  Instead of transposing the state
// for axis 0 (the reference's 4 transposes per apply+solve pair), kernels
// address the state with strides: state index of (line, pos) is
//   axis 0: pos * blocks + line    (spot lines, stride nv: coalesced across lines)
//   axis 1: line * block_len + pos (variance lines, contiguous)
// The per-line solve is a Thomas sweep with no pivoting (as cuSPARSE gtsv),
// one thread per line with the line data staged through LDS for coalesced
// accesses; the +-2 band entries of CAN_FOLD operators are folded on the
// fly, FOLDED operators fold the right-hand side.
#include "common.h"

struct fdadi_tri {
  int axis = 0;
  int64_t rows = 0, blocks = 0, block_len = 0;
  int has_low = 0, has_high = 0, has_residual = 0;
  int top_status = FDADI_CANNOT_FOLD, bottom_status = FDADI_CANNOT_FOLD;
  double *bands = nullptr;  // [sup | mid | sub], rows each
  double *R = nullptr;      // rows
  double *meta = nullptr;   // [low | high | top | bottom], blocks each
  double *deltas = nullptr; // ndeltas
  int64_t ndeltas = 0;
  double *scratch = nullptr;  // rows, Thomas modified super-diagonal
  double *sup() const { return bands; }
  double *mid() const { return bands + rows; }
  double *sub() const { return bands + 2 * rows; }
  double *low() const { return meta; }
  double *high() const { return meta + blocks; }
  double *top() const { return meta + 2 * blocks; }
  double *bottom() const { return meta + 3 * blocks; }
};

namespace fdadi {
namespace {

constexpr int kBlock = 256;

struct TriView {
  const double *sup, *mid, *sub, *R, *low, *high, *top, *bottom;
  int64_t blocks, bl;
  int axis, has_low, has_high, has_residual, top_status, bottom_status;
};

TriView view_of(const fdadi_tri *op) {
  TriView v;
  v.sup = op->sup(); v.mid = op->mid(); v.sub = op->sub(); v.R = op->R;
  v.low = op->low(); v.high = op->high(); v.top = op->top(); v.bottom = op->bottom();
  v.blocks = op->blocks; v.bl = op->block_len; v.axis = op->axis;
  v.has_low = op->has_low; v.has_high = op->has_high; v.has_residual = op->has_residual;
  v.top_status = op->top_status; v.bottom_status = op->bottom_status;
  return v;
}

__device__ __forceinline__ int64_t sidx(const TriView &o, int64_t line, int64_t pos) {
  return o.axis == 0 ? pos * o.blocks + line : line * o.bl + pos;
}

// x with the Dirichlet values of the operator forced in.
__device__ __forceinline__ double xf(const TriView &o, const double *x, int64_t line,
                                     int64_t pos) {
  if (pos == 0 && o.has_low) return o.low[line];
  if (pos == o.bl - 1 && o.has_high) return o.high[line];
  return x[sidx(o, line, pos)];
}

// Row (line, pos) of A x, including unfolded +-2 entries.
__device__ double row_dot(const TriView &o, const double *x, int64_t line, int64_t pos) {
  const int64_t r = line * o.bl + pos;
  double acc = o.mid[r] * xf(o, x, line, pos);
  if (pos > 0) acc += o.sub[r] * xf(o, x, line, pos - 1);
  if (pos < o.bl - 1) acc += o.sup[r] * xf(o, x, line, pos + 1);
  if (pos == o.bl - 1 && o.bottom_status == FDADI_CAN_FOLD && o.bl >= 3)
    acc += o.bottom[line] * xf(o, x, line, pos - 2);
  if (pos == 0 && o.top_status == FDADI_CAN_FOLD && o.bl >= 3)
    acc += o.top[line] * xf(o, x, line, 2);
  return acc;
}

__global__ void tri_apply_kernel(TriView o, const double *__restrict__ x,
                                 double *__restrict__ y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= o.blocks * o.bl) return;
  const int64_t line = r / o.bl, pos = r % o.bl;
  double v = row_dot(o, x, line, pos);
  // Folded rows hold X*A; undo the elimination on the result (fold_vector
  // with unfold=true, _TriBandedOperatorGPU.cu:525-529).
  if (pos == o.bl - 1 && o.bottom_status == FDADI_FOLDED)
    v -= o.bottom[line] * row_dot(o, x, line, pos - 1);
  if (pos == 0 && o.top_status == FDADI_FOLDED)
    v -= o.top[line] * row_dot(o, x, line, 1);
  if (o.has_residual) v += o.R[r];
  y[sidx(o, line, pos)] = v;
}

// Thomas sweep (no pivoting), one thread per line, one wave per 64
// consecutive lines.  The bands, the right-hand side, the modified
// super-diagonal (`cp`, in scratch) and the forward-substituted values
// (kept in place in v) move between global memory and LDS in tiles of 64
// lines x kSolveChunk positions with coalesced accesses -- 128-byte runs of
// one line for line-major arrays, 64 consecutive lines of one position for
// the strided state of axis 0 -- and each thread runs its line's
// recurrences out of the LDS tile (odd pitch: no bank conflicts).  Rows 0
// and n-1 (with their fold entries) are read directly, once per line.
constexpr int kSolveChunk = 16;
constexpr int kSolvePitch = kSolveChunk + 1;

__device__ __forceinline__ void tile_index(const TriView &o, int64_t L0, int64_t p0, int e,
                                           bool state, int &l, int &p, int64_t &g, bool &ok) {
  const int64_t n = o.bl;
  if (state && o.axis == 0) {  // position-major: 64 lines of one position per load
    p = e / 64;
    l = e % 64;
  } else {  // line-major: runs of kSolveChunk positions of one line
    l = e / kSolveChunk;
    p = e % kSolveChunk;
  }
  const int64_t line = L0 + l, pos = p0 + p;
  ok = line < o.blocks && pos < n;
  g = ok ? (state ? sidx(o, line, pos) : line * n + pos) : 0;
}

__global__ void __launch_bounds__(64)
tri_solve_kernel(TriView o, double *__restrict__ v, double *__restrict__ cpg) {
  __shared__ double ta[64 * kSolvePitch], tb[64 * kSolvePitch], tc[64 * kSolvePitch],
      td[64 * kSolvePitch];
  const int t = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * 64, line = L0 + t, n = o.bl, base = line * n;
  const bool live = line < o.blocks;
  auto rhs_at = [&](int64_t p) {
    double d = xf(o, v, line, p);
    if (o.has_residual) d -= o.R[base + p];
    return d;
  };
  // Row 0 (with top fold) and the last row (with bottom fold), per line.
  double b0 = 0.0, c0 = 0.0, d0 = 0.0, an = 0.0, bn = 0.0, fbot = 0.0;
  if (live) {
    b0 = o.mid[base];
    c0 = o.sup[base];
    d0 = rhs_at(0);
    if (o.top_status == FDADI_FOLDED) {
      d0 += o.top[line] * rhs_at(1);
    } else if (o.top_status == FDADI_CAN_FOLD) {
      const double c1 = o.sup[base + 1];
      const double f = c1 == 0 ? 0.0 : -o.top[line] / c1;
      b0 += f * o.sub[base + 1];
      c0 += f * o.mid[base + 1];
      d0 += f * rhs_at(1);
    }
    an = o.sub[base + n - 1];
    bn = o.mid[base + n - 1];
    if (o.bottom_status == FDADI_FOLDED) {
      fbot = o.bottom[line];
    } else if (o.bottom_status == FDADI_CAN_FOLD) {
      const double a = o.sub[base + n - 2];
      fbot = a == 0 ? 0.0 : -o.bottom[line] / a;
      an += fbot * o.mid[base + n - 2];
      bn += fbot * o.sup[base + n - 2];
    }
  }
  // Forward elimination, chunk by chunk.
  double cprev = 0.0, dprev = 0.0, dnm2 = 0.0;
  for (int64_t p0 = 0; p0 < n; p0 += kSolveChunk) {
    for (int e = t; e < 64 * kSolveChunk; e += 64) {
      int l, p;
      int64_t g;
      bool ok;
      tile_index(o, L0, p0, e, false, l, p, g, ok);
      ta[l * kSolvePitch + p] = ok ? o.sub[g] : 0.0;
      tb[l * kSolvePitch + p] = ok ? o.mid[g] : 0.0;
      tc[l * kSolvePitch + p] = ok ? o.sup[g] : 0.0;
      if (o.has_residual) td[l * kSolvePitch + p] = ok ? -o.R[g] : 0.0;
    }
    __syncthreads();
    for (int e = t; e < 64 * kSolveChunk; e += 64) {
      int l, p;
      int64_t g;
      bool ok;
      tile_index(o, L0, p0, e, true, l, p, g, ok);
      const int64_t ln = L0 + l, pos = p0 + p;
      double x = 0.0;
      if (ok) {
        if (pos == 0 && o.has_low) x = o.low[ln];
        else if (pos == n - 1 && o.has_high) x = o.high[ln];
        else x = v[g];
      }
      // rhs = x - R (the R tile holds -R when the residual is on)
      double &slot = td[l * kSolvePitch + p];
      slot = o.has_residual ? x + slot : x;
    }
    __syncthreads();
    if (live) {
      for (int k = 0; k < kSolveChunk && p0 + k < n; ++k) {
        const int64_t pos = p0 + k;
        double *ck = &tc[t * kSolvePitch + k], *dk = &td[t * kSolvePitch + k];
        if (pos == 0) {
          cprev = c0 / b0;
          dprev = d0 / b0;
        } else {
          double a, b, c, d;
          if (pos == n - 1) {
            d = *dk + fbot * dnm2;
            a = an; b = bn; c = 0.0;
          } else {
            d = *dk;
            if (pos == n - 2) dnm2 = d;
            a = ta[t * kSolvePitch + k]; b = tb[t * kSolvePitch + k]; c = *ck;
          }
          const double m = 1.0 / (b - a * cprev);
          cprev = c * m;
          dprev = (d - a * dprev) * m;
        }
        *ck = cprev;
        *dk = dprev;
      }
    }
    __syncthreads();
    for (int e = t; e < 64 * kSolveChunk; e += 64) {
      int l, p;
      int64_t g;
      bool ok;
      tile_index(o, L0, p0, e, false, l, p, g, ok);
      if (ok) cpg[g] = tc[l * kSolvePitch + p];
      tile_index(o, L0, p0, e, true, l, p, g, ok);
      if (ok) v[g] = td[l * kSolvePitch + p];
    }
    __syncthreads();
  }
  // Back substitution, chunks from the last; x(n-1) = dprev.
  double x = dprev;
  const int64_t last0 = (n - 1) / kSolveChunk * kSolveChunk;
  for (int64_t p0 = last0; p0 >= 0; p0 -= kSolveChunk) {
    for (int e = t; e < 64 * kSolveChunk; e += 64) {
      int l, p;
      int64_t g;
      bool ok;
      tile_index(o, L0, p0, e, false, l, p, g, ok);
      tc[l * kSolvePitch + p] = ok ? cpg[g] : 0.0;
      tile_index(o, L0, p0, e, true, l, p, g, ok);
      td[l * kSolvePitch + p] = ok ? v[g] : 0.0;
    }
    __syncthreads();
    if (live) {
      for (int k = kSolveChunk - 1; k >= 0; --k) {
        const int64_t pos = p0 + k;
        if (pos > n - 2) continue;  // x(n-1) is dprev, already in place
        x = td[t * kSolvePitch + k] - tc[t * kSolvePitch + k] * x;
        td[t * kSolvePitch + k] = x;
      }
    }
    __syncthreads();
    for (int e = t; e < 64 * kSolveChunk; e += 64) {
      int l, p;
      int64_t g;
      bool ok;
      tile_index(o, L0, p0, e, true, l, p, g, ok);
      if (ok) v[g] = td[l * kSolvePitch + p];
    }
    __syncthreads();
  }
}

// Scale by vec tiled over the rows, or by the scalar `value` when vec is null.
__global__ void tri_scale_kernel(double *sup, double *mid, double *sub, double *R,
                                 const double *__restrict__ vec, double value, int64_t vsize,
                                 int64_t rows, int64_t bl, int has_low, int has_high) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const double s = vec ? vec[r % vsize] : value;
  const int64_t pos = r % bl;
  sup[r] *= s;
  sub[r] *= s;
  const bool dir = (pos == 0 && has_low) || (pos == bl - 1 && has_high);
  if (!dir) mid[r] *= s;
  R[r] *= s;
}

__global__ void tri_scale_factors_kernel(double *top, double *bottom,
                                         const double *__restrict__ vec, double value,
                                         int64_t vsize, int64_t blocks, int64_t bl,
                                         int scale_top, int scale_bottom) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= blocks) return;
  if (scale_top) top[b] *= vec ? vec[(b * bl) % vsize] : value;
  if (scale_bottom) bottom[b] *= vec ? vec[(b * bl + bl - 1) % vsize] : value;
}

__global__ void tri_add_kernel(double *sup, double *mid, double *sub, double *R,
                               const double *__restrict__ osup, const double *__restrict__ omid,
                               const double *__restrict__ osub, const double *__restrict__ oR,
                               int64_t rows, int64_t bl, int has_low, int has_high,
                               int add_R) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t pos = r % bl;
  sup[r] += osup[r];
  sub[r] += osub[r];
  const bool dir = (pos == 0 && has_low) || (pos == bl - 1 && has_high);
  if (!dir) mid[r] += omid[r];
  if (add_R) R[r] += oR[r];
}

__global__ void tri_add_scalar_kernel(double *mid, const double *__restrict__ dval, double hval,
                                      int64_t rows, int64_t bl, int has_low, int has_high) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t pos = r % bl;
  const bool dir = (pos == 0 && has_low) || (pos == bl - 1 && has_high);
  if (!dir) mid[r] += dval ? dval[0] : hval;
}

__global__ void vec_add_kernel(double *a, const double *__restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// fold_bottom / unfold (_TriBandedOperatorGPU.cu:953-1009): factor f =
// -z / A[n-2,n-3]; row_{n-1} += f * row_{n-2}; the factor replaces z.
__global__ void tri_fold_bottom_kernel(double *sup, double *mid, double *sub, double *bottom,
                                       int64_t blocks, int64_t bl, int unfold) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= blocks) return;
  const int64_t n1 = b * bl + bl - 1, n2 = n1 - 1;
  if (!unfold) {
    const double a = sub[n2];
    const double f = a == 0 ? 0.0 : -bottom[b] / a;
    sub[n1] += mid[n2] * f;
    mid[n1] += sup[n2] * f;
    bottom[b] = f;
  } else {
    const double f = bottom[b];
    sub[n1] -= mid[n2] * f;
    mid[n1] -= sup[n2] * f;
    bottom[b] = f * -sub[n2];
  }
}

// fold_top / unfold (_TriBandedOperatorGPU.cu:895-950): f = -z / A[1,2];
// row_0 += f * row_1.
__global__ void tri_fold_top_kernel(double *sup, double *mid, double *sub, double *top,
                                    int64_t blocks, int64_t bl, int unfold) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= blocks) return;
  const int64_t r0 = b * bl, r1 = r0 + 1;
  if (!unfold) {
    const double c = sup[r1];
    const double f = c == 0 ? 0.0 : -top[b] / c;
    sup[r0] += mid[r1] * f;
    mid[r0] += sub[r1] * f;
    top[b] = f;
  } else {
    const double f = top[b];
    sup[r0] -= mid[r1] * f;
    mid[r0] -= sub[r1] * f;
    top[b] = f * -sup[r1];
  }
}

// Heston 1-D central stencils with their boundary rows
// (_TriBandedOperatorGPU.cu:66-362, BandedOperator.pyx:432-523, 970-1017).
// Floating-point contraction is disabled so the bands equal the CPU
// construction bit for bit.
__global__ void tri_for_vector_kernel(const double *__restrict__ mesh, int64_t n,
                                      int64_t blocks, int derivative, int axis, int barrier,
                                      double *sup, double *mid, double *sub, double *R,
                                      double *bottom) {
#pragma clang fp contract(off)
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n * blocks) return;
  const int64_t p = r % n;
  double a = 0.0, b = 0.0, c = 0.0, res = 0.0;
  if (p > 0 && p < n - 1) {
    const double hm = mesh[p] - mesh[p - 1], hp = mesh[p + 1] - mesh[p];
    const double x = hm + hp;
    if (derivative == 1) {
      c = hm / (hp * x);
      b = (-hm + hp) / (hm * hp);
      a = -hp / (hm * x);
    } else {
      c = 2 / (hp * x);
      b = -2 / (hm * hp);
      a = 2 / (hm * x);
    }
  } else if (axis == 0) {
    if (p == 0 || barrier) {
      b = 1.0;  // Dirichlet row (value 0)
    } else if (derivative == 1) {
      res = 1.0;  // Von Neumann dU/dS = 1
    } else {
      const double d = mesh[n - 1] - mesh[n - 2];
      const double xx = d * d;  // second derivative with known first derivative 1
      b = -2 / xx;
      a = 2 / xx;
      res = (1.0 * 2) / d;
    }
  } else {
    if (derivative == 1) {
      if (p == 0) {
        const double d1 = mesh[1] - mesh[0];
        c = 1 / d1;
        b = -1 / d1;
      } else {
        const double d = mesh[n - 1] - mesh[n - 2];
        b = 1.0 / d;
        a = -1.0 / d;
      }
    } else {
      if (p == n - 1) {
        const double d1 = mesh[n - 1] - mesh[n - 2], d2 = mesh[n - 2] - mesh[n - 3];
        const double rd = 1.0 / (0.5 * (d2 + d1) * d2 * d1);
        const double z = d1 * rd;
        a = -(d2 + d1) * rd;
        b = d2 * rd;
        if (r / n < blocks) bottom[r / n] = z;
      }
      // p == 0: Von Neumann row with value 0: all zero.
    }
  }
  sup[r] = c;
  mid[r] = b;
  sub[r] = a;
  R[r] = res;
}

__global__ void deltas_kernel(const double *__restrict__ mesh, int64_t n, double *deltas) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  deltas[p] = p == 0 ? __longlong_as_double(0x7ff8000000000000LL) : mesh[p] - mesh[p - 1];
}

int alloc_tri(fdadi_tri *op) {
  const int64_t rows = op->rows, blocks = op->blocks;
  FDADI_TRY_HIP(hipMalloc((void **)&op->bands, sizeof(double) * 3 * rows));
  FDADI_TRY_HIP(hipMalloc((void **)&op->R, sizeof(double) * rows));
  FDADI_TRY_HIP(hipMalloc((void **)&op->meta, sizeof(double) * 4 * blocks));
  FDADI_TRY_HIP(hipMalloc((void **)&op->scratch, sizeof(double) * rows));
  if (op->ndeltas > 0)
    FDADI_TRY_HIP(hipMalloc((void **)&op->deltas, sizeof(double) * op->ndeltas));
  return FDADI_OK;
}

void free_tri(fdadi_tri *op) {
  if (!op) return;
  (void)hipFree(op->bands);
  (void)hipFree(op->R);
  (void)hipFree(op->meta);
  (void)hipFree(op->scratch);
  if (op->deltas) (void)hipFree(op->deltas);
  delete op;
}

bool dirichlet_ok(const fdadi_tri *op) { return op->block_len >= 3; }

}  // namespace
}  // namespace fdadi

using namespace fdadi;

extern "C" {

int fdadi_tri_create(const double *sup, const double *mid, const double *sub,
                     const double *R, const double *deltas, int64_t ndeltas,
                     const double *low_dirichlet, const double *high_dirichlet,
                     const double *top_factors, const double *bottom_factors,
                     int axis, int64_t rows, int64_t blocks, int has_low_dirichlet,
                     int has_high_dirichlet, int top_fold_status,
                     int bottom_fold_status, int has_residual, fdadi_tri **out) {
  FDADI_REQUIRE(out && sup && mid && sub, FDADI_ERR_ARG, "fdadi_tri_create: null argument");
  FDADI_REQUIRE(rows > 0 && blocks > 0 && rows % blocks == 0, FDADI_ERR_ARG,
                "fdadi_tri_create: rows must be a positive multiple of blocks");
  FDADI_REQUIRE(axis == 0 || axis == 1, FDADI_ERR_ARG, "fdadi_tri_create: axis must be 0 or 1");
  FDADI_REQUIRE(rows / blocks >= 3, FDADI_ERR_ARG,
                "fdadi_tri_create: vector too short to use finite differencing");
  FDADI_REQUIRE(top_fold_status >= 0 && top_fold_status <= 2 && bottom_fold_status >= 0 &&
                    bottom_fold_status <= 2,
                FDADI_ERR_ARG, "fdadi_tri_create: bad fold status");
  fdadi_tri *op = new fdadi_tri();
  op->axis = axis;
  op->rows = rows;
  op->blocks = blocks;
  op->block_len = rows / blocks;
  op->has_low = !!has_low_dirichlet;
  op->has_high = !!has_high_dirichlet;
  op->has_residual = !!has_residual;
  op->top_status = top_fold_status;
  op->bottom_status = bottom_fold_status;
  op->ndeltas = deltas ? ndeltas : 0;
  int st = alloc_tri(op);
  if (st != FDADI_OK) {
    free_tri(op);
    return st;
  }
  const size_t rb = sizeof(double) * rows, bb = sizeof(double) * blocks;
  std::vector<double> zeros((size_t)std::max(rows, blocks), 0.0);
  hipError_t e = hipSuccess;
  auto cp = [&](double *dst, const double *src, size_t bytes) {
    if (e != hipSuccess) return;
    e = hipMemcpy(dst, src ? src : zeros.data(), bytes, hipMemcpyHostToDevice);
  };
  cp(op->sup(), sup, rb);
  cp(op->mid(), mid, rb);
  cp(op->sub(), sub, rb);
  cp(op->R, R, rb);
  cp(op->low(), low_dirichlet, bb);
  cp(op->high(), high_dirichlet, bb);
  cp(op->top(), top_factors, bb);
  cp(op->bottom(), bottom_factors, bb);
  if (op->ndeltas) cp(op->deltas, deltas, sizeof(double) * op->ndeltas);
  if (e != hipSuccess) {
    free_tri(op);
    set_error(std::string("fdadi_tri_create: hipMemcpy: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_tri_for_vector(const double *mesh, int64_t n, int64_t blocks, int derivative,
                         int axis, int barrier, void *stream, fdadi_tri **out) {
  FDADI_REQUIRE(out && mesh, FDADI_ERR_ARG, "fdadi_tri_for_vector: null argument");
  FDADI_REQUIRE(n >= 4 && blocks >= 1, FDADI_ERR_ARG,
                "fdadi_tri_for_vector: vector too short to use finite differencing");
  FDADI_REQUIRE(derivative == 1 || derivative == 2, FDADI_ERR_ARG,
                "fdadi_tri_for_vector: derivative must be one of (1, 2)");
  FDADI_REQUIRE(axis == 0 || axis == 1, FDADI_ERR_ARG,
                "fdadi_tri_for_vector: axis must be one of (0, 1)");
  hipStream_t s = as_stream(stream);
  fdadi_tri *op = new fdadi_tri();
  op->axis = axis;
  op->rows = n * blocks;
  op->blocks = blocks;
  op->block_len = n;
  op->ndeltas = n;
  op->has_low = axis == 0;
  op->has_high = axis == 0 && barrier;
  op->has_residual = axis == 0;
  op->top_status = FDADI_CANNOT_FOLD;
  op->bottom_status = (axis == 1 && derivative == 2) ? FDADI_CAN_FOLD : FDADI_CANNOT_FOLD;
  int st = alloc_tri(op);
  if (st != FDADI_OK) {
    free_tri(op);
    return st;
  }
  double *d_mesh = nullptr;
  hipError_t e = hipMalloc((void **)&d_mesh, sizeof(double) * n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_mesh, mesh, sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(op->meta, 0, sizeof(double) * 4 * blocks, s);
  if (e != hipSuccess) {
    (void)hipFree(d_mesh);
    free_tri(op);
    set_error(std::string("fdadi_tri_for_vector: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  hipLaunchKernelGGL(tri_for_vector_kernel, dim3(grid_for(op->rows, kBlock)), dim3(kBlock), 0,
                     s, d_mesh, n, blocks, derivative, axis, barrier, op->sup(), op->mid(),
                     op->sub(), op->R, op->bottom());
  hipLaunchKernelGGL(deltas_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, d_mesh, n,
                     op->deltas);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_mesh);
  if (e != hipSuccess) {
    free_tri(op);
    set_error(std::string("fdadi_tri_for_vector: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_tri_copy(const fdadi_tri *src, void *stream, fdadi_tri **out) {
  FDADI_REQUIRE(src && out, FDADI_ERR_ARG, "fdadi_tri_copy: null argument");
  hipStream_t s = as_stream(stream);
  fdadi_tri *op = new fdadi_tri();
  op->axis = src->axis;
  op->rows = src->rows;
  op->blocks = src->blocks;
  op->block_len = src->block_len;
  op->has_low = src->has_low;
  op->has_high = src->has_high;
  op->has_residual = src->has_residual;
  op->top_status = src->top_status;
  op->bottom_status = src->bottom_status;
  op->ndeltas = src->ndeltas;
  int st = alloc_tri(op);
  if (st != FDADI_OK) {
    free_tri(op);
    return st;
  }
  hipError_t e = hipMemcpyAsync(op->bands, src->bands, sizeof(double) * 3 * op->rows,
                                hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(op->R, src->R, sizeof(double) * op->rows, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(op->meta, src->meta, sizeof(double) * 4 * op->blocks,
                       hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && op->ndeltas)
    e = hipMemcpyAsync(op->deltas, src->deltas, sizeof(double) * op->ndeltas,
                       hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    free_tri(op);
    set_error(std::string("fdadi_tri_copy: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_tri_destroy(fdadi_tri *op) {
  free_tri(op);
  return FDADI_OK;
}

int fdadi_tri_info(const fdadi_tri *op, int64_t *rows, int64_t *blocks, int *axis,
                   int *has_low_dirichlet, int *has_high_dirichlet, int *top_fold_status,
                   int *bottom_fold_status, int *has_residual, int64_t *ndeltas) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_info: null operator");
  if (rows) *rows = op->rows;
  if (blocks) *blocks = op->blocks;
  if (axis) *axis = op->axis;
  if (has_low_dirichlet) *has_low_dirichlet = op->has_low;
  if (has_high_dirichlet) *has_high_dirichlet = op->has_high;
  if (top_fold_status) *top_fold_status = op->top_status;
  if (bottom_fold_status) *bottom_fold_status = op->bottom_status;
  if (has_residual) *has_residual = op->has_residual;
  if (ndeltas) *ndeltas = op->ndeltas;
  return FDADI_OK;
}

int fdadi_tri_set_flags(fdadi_tri *op, int has_low_dirichlet, int has_high_dirichlet,
                        int has_residual) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_set_flags: null operator");
  if (has_low_dirichlet >= 0) op->has_low = !!has_low_dirichlet;
  if (has_high_dirichlet >= 0) op->has_high = !!has_high_dirichlet;
  if (has_residual >= 0) op->has_residual = !!has_residual;
  return FDADI_OK;
}

int fdadi_tri_read(const fdadi_tri *op, double *sup, double *mid, double *sub, double *R,
                   double *low_dirichlet, double *high_dirichlet, double *top_factors,
                   double *bottom_factors, double *deltas) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_read: null operator");
  FDADI_TRY_HIP(hipDeviceSynchronize());
  const size_t rb = sizeof(double) * op->rows, bb = sizeof(double) * op->blocks;
  auto get = [&](double *dst, const double *src, size_t bytes) -> hipError_t {
    if (!dst) return hipSuccess;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  FDADI_TRY_HIP(get(sup, op->sup(), rb));
  FDADI_TRY_HIP(get(mid, op->mid(), rb));
  FDADI_TRY_HIP(get(sub, op->sub(), rb));
  FDADI_TRY_HIP(get(R, op->R, rb));
  FDADI_TRY_HIP(get(low_dirichlet, op->low(), bb));
  FDADI_TRY_HIP(get(high_dirichlet, op->high(), bb));
  FDADI_TRY_HIP(get(top_factors, op->top(), bb));
  FDADI_TRY_HIP(get(bottom_factors, op->bottom(), bb));
  if (op->ndeltas) FDADI_TRY_HIP(get(deltas, op->deltas, sizeof(double) * op->ndeltas));
  return FDADI_OK;
}

int fdadi_tri_apply(const fdadi_tri *op, const double *d_x, double *d_y, int64_t n,
                    void *stream) {
  FDADI_REQUIRE(op && d_x && d_y, FDADI_ERR_ARG, "fdadi_tri_apply: null argument");
  FDADI_REQUIRE(n == op->rows, FDADI_ERR_ARG,
                "fdadi_tri_apply: vector size " + std::to_string(n) +
                    " != operator rows " + std::to_string(op->rows));
  FDADI_REQUIRE(d_x != d_y, FDADI_ERR_ARG, "fdadi_tri_apply: output must not alias input");
  hipLaunchKernelGGL(tri_apply_kernel, dim3(grid_for(op->rows, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), view_of(op), d_x, d_y);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_tri_solve(const fdadi_tri *op, double *d_v, int64_t n, void *stream) {
  FDADI_REQUIRE(op && d_v, FDADI_ERR_ARG, "fdadi_tri_solve: null argument");
  FDADI_REQUIRE(n == op->rows, FDADI_ERR_ARG,
                "fdadi_tri_solve: vector size " + std::to_string(n) +
                    " != operator rows " + std::to_string(op->rows));
  hipLaunchKernelGGL(tri_solve_kernel, dim3(grid_for(op->blocks, 64)), dim3(64), 0,
                     as_stream(stream), view_of(op), d_v, op->scratch);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_tri_add_operator(fdadi_tri *op, const fdadi_tri *other, void *stream) {
  FDADI_REQUIRE(op && other, FDADI_ERR_ARG, "fdadi_tri_add_operator: null operator");
  FDADI_REQUIRE(op->axis == other->axis, FDADI_ERR_ARG,
                "Both operators must operate on the same axis.");
  FDADI_REQUIRE(op->rows == other->rows && op->blocks == other->blocks, FDADI_ERR_ARG,
                "Both operators must have the same length");
  FDADI_REQUIRE(op->top_status != FDADI_FOLDED && op->bottom_status != FDADI_FOLDED &&
                    other->top_status != FDADI_FOLDED && other->bottom_status != FDADI_FOLDED,
                FDADI_ERR_STATE, "Cannot add folded (diagonalized) operators");
  hipStream_t s = as_stream(stream);
  const bool add_R = op->has_residual && other->has_residual;
  hipLaunchKernelGGL(tri_add_kernel, dim3(grid_for(op->rows, kBlock)), dim3(kBlock), 0, s,
                     op->sup(), op->mid(), op->sub(), op->R, other->sup(), other->mid(),
                     other->sub(), other->R, op->rows, op->block_len, op->has_low,
                     op->has_high, (int)add_R);
  FDADI_CHECK_LAUNCH();
  if (!op->has_residual && other->has_residual) {
    FDADI_TRY_HIP(hipMemcpyAsync(op->R, other->R, sizeof(double) * op->rows,
                                 hipMemcpyDeviceToDevice, s));
    op->has_residual = 1;
  }
  if (other->top_status == FDADI_CAN_FOLD) {
    if (op->top_status == FDADI_CAN_FOLD) {
      hipLaunchKernelGGL(vec_add_kernel, dim3(grid_for(op->blocks, kBlock)), dim3(kBlock), 0, s,
                         op->top(), other->top(), op->blocks);
      FDADI_CHECK_LAUNCH();
    } else {
      FDADI_TRY_HIP(hipMemcpyAsync(op->top(), other->top(), sizeof(double) * op->blocks,
                                   hipMemcpyDeviceToDevice, s));
      op->top_status = FDADI_CAN_FOLD;
    }
  }
  if (other->bottom_status == FDADI_CAN_FOLD) {
    if (op->bottom_status == FDADI_CAN_FOLD) {
      hipLaunchKernelGGL(vec_add_kernel, dim3(grid_for(op->blocks, kBlock)), dim3(kBlock), 0, s,
                         op->bottom(), other->bottom(), op->blocks);
      FDADI_CHECK_LAUNCH();
    } else {
      FDADI_TRY_HIP(hipMemcpyAsync(op->bottom(), other->bottom(), sizeof(double) * op->blocks,
                                   hipMemcpyDeviceToDevice, s));
      op->bottom_status = FDADI_CAN_FOLD;
    }
  }
  return FDADI_OK;
}

static int add_scalar_impl(fdadi_tri *op, const double *dval, double hval, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_add_scalar: null operator");
  hipLaunchKernelGGL(tri_add_scalar_kernel, dim3(grid_for(op->rows, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), op->mid(), dval, hval, op->rows, op->block_len,
                     op->has_low, op->has_high);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_tri_add_scalar(fdadi_tri *op, double value, void *stream) {
  return add_scalar_impl(op, nullptr, value, stream);
}

int fdadi_tri_add_scalar_dev(fdadi_tri *op, const double *d_value, void *stream) {
  FDADI_REQUIRE(d_value, FDADI_ERR_ARG, "fdadi_tri_add_scalar_dev: null value");
  return add_scalar_impl(op, d_value, 0.0, stream);
}

namespace {
// vectorized_scale with a device vector (d_vec) or a host scalar (d_vec null).
int scale_impl(fdadi_tri *op, const double *d_vec, double value, int64_t vsize,
               void *stream) {
  FDADI_REQUIRE(op->top_status != FDADI_FOLDED && op->bottom_status != FDADI_FOLDED,
                FDADI_ERR_STATE, "Cannot scale diagonalized operator.");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(tri_scale_kernel, dim3(grid_for(op->rows, kBlock)), dim3(kBlock), 0, s,
                     op->sup(), op->mid(), op->sub(), op->R, d_vec, value, vsize, op->rows,
                     op->block_len, op->has_low, op->has_high);
  FDADI_CHECK_LAUNCH();
  const int st = op->top_status == FDADI_CAN_FOLD && !op->has_low;
  const int sb = op->bottom_status == FDADI_CAN_FOLD && !op->has_high;
  if (st || sb) {
    hipLaunchKernelGGL(tri_scale_factors_kernel, dim3(grid_for(op->blocks, kBlock)),
                       dim3(kBlock), 0, s, op->top(), op->bottom(), d_vec, value, vsize,
                       op->blocks, op->block_len, st, sb);
    FDADI_CHECK_LAUNCH();
  }
  return FDADI_OK;
}
}  // namespace

int fdadi_tri_scale(fdadi_tri *op, const double *d_vec, int64_t vsize, void *stream) {
  FDADI_REQUIRE(op && d_vec, FDADI_ERR_ARG, "fdadi_tri_scale: null argument");
  FDADI_REQUIRE(vsize > 0, FDADI_ERR_ARG, "vsize == 0");
  FDADI_REQUIRE(op->rows % vsize == 0, FDADI_ERR_ARG,
                "Vector length does not divide evenly into operator size. Cannot scale. "
                "vsize(" + std::to_string(vsize) + ") operator_rows(" +
                    std::to_string(op->rows) + ")");
  return scale_impl(op, d_vec, 0.0, vsize, stream);
}

int fdadi_tri_scale_scalar(fdadi_tri *op, double value, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_scale_scalar: null operator");
  // mul_scalar_from_host (BandedOperatorGPU.pyx:488-501): the scalar travels
  // as a kernel argument (no staging copy, no synchronisation).
  return scale_impl(op, nullptr, value, 1, stream);
}

int fdadi_tri_fold_top(fdadi_tri *op, int unfold, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_fold_top: null operator");
  if (!unfold)
    FDADI_REQUIRE(op->top_status == FDADI_CAN_FOLD, FDADI_ERR_STATE,
                  "fold_top: operator has no top fold entry (not CAN_FOLD)");
  else
    FDADI_REQUIRE(op->top_status == FDADI_FOLDED, FDADI_ERR_STATE,
                  "fold_top(unfold): operator is not FOLDED");
  hipLaunchKernelGGL(tri_fold_top_kernel, dim3(grid_for(op->blocks, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), op->sup(), op->mid(), op->sub(), op->top(),
                     op->blocks, op->block_len, unfold);
  FDADI_CHECK_LAUNCH();
  op->top_status = unfold ? FDADI_CAN_FOLD : FDADI_FOLDED;
  return FDADI_OK;
}

int fdadi_tri_fold_bottom(fdadi_tri *op, int unfold, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_fold_bottom: null operator");
  if (!unfold)
    FDADI_REQUIRE(op->bottom_status == FDADI_CAN_FOLD, FDADI_ERR_STATE,
                  "fold_bottom: operator has no bottom fold entry (not CAN_FOLD)");
  else
    FDADI_REQUIRE(op->bottom_status == FDADI_FOLDED, FDADI_ERR_STATE,
                  "fold_bottom(unfold): operator is not FOLDED");
  hipLaunchKernelGGL(tri_fold_bottom_kernel, dim3(grid_for(op->blocks, kBlock)), dim3(kBlock),
                     0, as_stream(stream), op->sup(), op->mid(), op->sub(), op->bottom(),
                     op->blocks, op->block_len, unfold);
  FDADI_CHECK_LAUNCH();
  op->bottom_status = unfold ? FDADI_CAN_FOLD : FDADI_FOLDED;
  return FDADI_OK;
}

int fdadi_tri_diagonalize(fdadi_tri *op, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_diagonalize: null operator");
  if (op->bottom_status == FDADI_CAN_FOLD) {
    int st = fdadi_tri_fold_bottom(op, 0, stream);
    if (st) return st;
  }
  if (op->top_status == FDADI_CAN_FOLD) {
    int st = fdadi_tri_fold_top(op, 0, stream);
    if (st) return st;
  }
  return FDADI_OK;
}

int fdadi_tri_undiagonalize(fdadi_tri *op, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_tri_undiagonalize: null operator");
  if (op->bottom_status == FDADI_FOLDED) {
    int st = fdadi_tri_fold_bottom(op, 1, stream);
    if (st) return st;
  }
  if (op->top_status == FDADI_FOLDED) {
    int st = fdadi_tri_fold_top(op, 1, stream);
    if (st) return st;
  }
  return FDADI_OK;
}

}  // extern "C"
