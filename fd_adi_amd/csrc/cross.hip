// Cross-derivative operator d2/dS dv on the (n0, n1) grid and the Heston
// coefficient fields.
//
// Replaces _CSRBandedOperator (_CSRBandedOperatorGPU.cu) and
// _coefficients.cu.  The reference builds the 9-point stencil as COO, sorts
// it into CSR and multiplies with cusparseDcsrmv (int32 indices, ~12 B of
// index traffic per nonzero).  Here the stencil is stored as 9 row-indexed
// bands with the fixed offsets a*n1 + b (a, b in {1, 0, -1}), and apply is a
// direct 9-point gather: no indices, no sort.
#include "common.h"

struct fdadi_cross {
  int64_t n0 = 0, n1 = 0, N = 0;
  double *bands = nullptr;  // 9 * N, k = (1-a)*3 + (1-b)
  double *R = nullptr;      // N
};

namespace fdadi {
namespace {

constexpr int kBlock = 256;

// Central first-derivative weights (sub, mid, sup) of interior row p
// (BandedOperator.pyx:991-995).
__device__ void center_first(const double *m, int64_t p, double w[3]) {
#pragma clang fp contract(off)
  const double hm = m[p] - m[p - 1], hp = m[p + 1] - m[p];
  const double x = hm + hp;
  w[2] = hm / (hp * x);
  w[1] = (-hm + hp) / (hm * hp);
  w[0] = -hp / (hm * x);
}

// Interior rows get ws_a(i) * wv_b(j) (the variance weight scaled by the spot
// weight, the order of FiniteDifferenceEngine.pyx:520-532); boundary rows
// are zero.
__global__ void cross_for_vector_kernel(const double *__restrict__ m0, int64_t n0,
                                        const double *__restrict__ m1, int64_t n1,
                                        double *bands) {
#pragma clang fp contract(off)
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t N = n0 * n1;
  if (r >= N) return;
  const int64_t i = r / n1, j = r % n1;
  const bool interior = i > 0 && i < n0 - 1 && j > 0 && j < n1 - 1;
  double ws[3] = {0, 0, 0}, wv[3] = {0, 0, 0};
  if (interior) {
    center_first(m0, i, ws);
    center_first(m1, j, wv);
  }
  for (int a = 1; a >= -1; --a)
    for (int b = 1; b >= -1; --b) {
      const int k = (1 - a) * 3 + (1 - b);
      bands[k * N + r] = interior ? wv[b + 1] * ws[a + 1] : 0.0;
    }
}

__global__ void cross_apply_kernel(const double *__restrict__ bands, const double *__restrict__ R,
                                   int64_t n0, int64_t n1, const double *__restrict__ x,
                                   double *__restrict__ y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t N = n0 * n1;
  if (r >= N) return;
  const int64_t i = r / n1, j = r % n1;
  double acc = 0.0;
  for (int a = 1; a >= -1; --a) {
    const int64_t ii = i + a;
    if (ii < 0 || ii >= n0) continue;
    for (int b = 1; b >= -1; --b) {
      const int64_t jj = j + b;
      if (jj < 0 || jj >= n1) continue;
      const int k = (1 - a) * 3 + (1 - b);
      acc += bands[k * N + r] * x[ii * n1 + jj];
    }
  }
  y[r] = acc + R[r];
}

// Scale by vec tiled over the rows, or by the scalar `value` when vec is null.
__global__ void cross_scale_kernel(double *bands, double *R, const double *__restrict__ vec,
                                   double value, int64_t vsize, int64_t N) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= N) return;
  const double s = vec ? vec[r % vsize] : value;
  for (int k = 0; k < 9; ++k) bands[k * N + r] *= s;
  R[r] *= s;
}

// Heston PDE coefficient fields (heston.py:303-326, _coefficients.cu:122-338).
__global__ void heston_coef_kernel(int which, const double *__restrict__ S, int64_t nS,
                                   const double *__restrict__ v, int64_t nv, double r,
                                   double kappa, double theta, double sigma, double rho,
                                   double *out) {
#pragma clang fp contract(off)
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nS * nv) return;
  int64_t i, j;
  if (which <= 1) {  // spot axis fastest
    i = k % nS;
    j = k / nS;
  } else {  // variance axis fastest
    i = k / nv;
    j = k % nv;
  }
  const double s = S[i], x = v[j];
  double c;
  switch (which) {
    case 0: c = r * s; break;
    case 1: c = (0.5 * x) * (s * s); break;
    case 2: c = s == 0 ? 0.0 : kappa * (theta - x); break;
    case 3: c = s == 0 ? 0.0 : (0.5 * (sigma * sigma)) * x; break;
    default: c = ((rho * sigma) * s) * x; break;
  }
  out[k] = c;
}

void free_cross(fdadi_cross *op) {
  if (!op) return;
  (void)hipFree(op->bands);
  (void)hipFree(op->R);
  delete op;
}

int alloc_cross(fdadi_cross *op) {
  FDADI_TRY_HIP(hipMalloc((void **)&op->bands, sizeof(double) * 9 * op->N));
  FDADI_TRY_HIP(hipMalloc((void **)&op->R, sizeof(double) * op->N));
  return FDADI_OK;
}

}  // namespace
}  // namespace fdadi

using namespace fdadi;

extern "C" {

int fdadi_cross_for_vector(const double *mesh0, int64_t n0, const double *mesh1, int64_t n1,
                           void *stream, fdadi_cross **out) {
  FDADI_REQUIRE(out && mesh0 && mesh1, FDADI_ERR_ARG, "fdadi_cross_for_vector: null argument");
  FDADI_REQUIRE(n0 >= 3 && n1 >= 3, FDADI_ERR_ARG,
                "fdadi_cross_for_vector: both meshes need at least 3 points");
  hipStream_t s = as_stream(stream);
  fdadi_cross *op = new fdadi_cross();
  op->n0 = n0;
  op->n1 = n1;
  op->N = n0 * n1;
  int st = alloc_cross(op);
  if (st != FDADI_OK) {
    free_cross(op);
    return st;
  }
  double *d_m = nullptr;
  hipError_t e = hipMalloc((void **)&d_m, sizeof(double) * (n0 + n1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_m, mesh0, sizeof(double) * n0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_m + n0, mesh1, sizeof(double) * n1, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(op->R, 0, sizeof(double) * op->N, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(cross_for_vector_kernel, dim3(grid_for(op->N, kBlock)), dim3(kBlock), 0,
                       s, d_m, n0, d_m + n0, n1, op->bands);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_m);
  if (e != hipSuccess) {
    free_cross(op);
    set_error(std::string("fdadi_cross_for_vector: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_cross_create(const double *bands, const double *R, int64_t n0, int64_t n1,
                       fdadi_cross **out) {
  FDADI_REQUIRE(out && bands, FDADI_ERR_ARG, "fdadi_cross_create: null argument");
  FDADI_REQUIRE(n0 >= 3 && n1 >= 3, FDADI_ERR_ARG, "fdadi_cross_create: grid too small");
  fdadi_cross *op = new fdadi_cross();
  op->n0 = n0;
  op->n1 = n1;
  op->N = n0 * n1;
  int st = alloc_cross(op);
  if (st != FDADI_OK) {
    free_cross(op);
    return st;
  }
  hipError_t e = hipMemcpy(op->bands, bands, sizeof(double) * 9 * op->N, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (R) {
      e = hipMemcpy(op->R, R, sizeof(double) * op->N, hipMemcpyHostToDevice);
    } else {
      std::vector<double> z((size_t)op->N, 0.0);
      e = hipMemcpy(op->R, z.data(), sizeof(double) * op->N, hipMemcpyHostToDevice);
    }
  }
  if (e != hipSuccess) {
    free_cross(op);
    set_error(std::string("fdadi_cross_create: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_cross_copy(const fdadi_cross *src, void *stream, fdadi_cross **out) {
  FDADI_REQUIRE(src && out, FDADI_ERR_ARG, "fdadi_cross_copy: null argument");
  fdadi_cross *op = new fdadi_cross();
  op->n0 = src->n0;
  op->n1 = src->n1;
  op->N = src->N;
  int st = alloc_cross(op);
  if (st != FDADI_OK) {
    free_cross(op);
    return st;
  }
  hipStream_t s = as_stream(stream);
  hipError_t e = hipMemcpyAsync(op->bands, src->bands, sizeof(double) * 9 * op->N,
                                hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(op->R, src->R, sizeof(double) * op->N, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    free_cross(op);
    set_error(std::string("fdadi_cross_copy: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = op;
  return FDADI_OK;
}

int fdadi_cross_destroy(fdadi_cross *op) {
  free_cross(op);
  return FDADI_OK;
}

int fdadi_cross_info(const fdadi_cross *op, int64_t *n0, int64_t *n1) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_cross_info: null operator");
  if (n0) *n0 = op->n0;
  if (n1) *n1 = op->n1;
  return FDADI_OK;
}

int fdadi_cross_read(const fdadi_cross *op, double *bands, double *R) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_cross_read: null operator");
  FDADI_TRY_HIP(hipDeviceSynchronize());
  if (bands)
    FDADI_TRY_HIP(hipMemcpy(bands, op->bands, sizeof(double) * 9 * op->N, hipMemcpyDeviceToHost));
  if (R) FDADI_TRY_HIP(hipMemcpy(R, op->R, sizeof(double) * op->N, hipMemcpyDeviceToHost));
  return FDADI_OK;
}

int fdadi_cross_apply(const fdadi_cross *op, const double *d_x, double *d_y, int64_t n,
                      void *stream) {
  FDADI_REQUIRE(op && d_x && d_y, FDADI_ERR_ARG, "fdadi_cross_apply: null argument");
  FDADI_REQUIRE(n == op->N, FDADI_ERR_ARG,
                "fdadi_cross_apply: vector size " + std::to_string(n) + " != operator rows " +
                    std::to_string(op->N));
  FDADI_REQUIRE(d_x != d_y, FDADI_ERR_ARG, "fdadi_cross_apply: output must not alias input");
  hipLaunchKernelGGL(cross_apply_kernel, dim3(grid_for(op->N, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), op->bands, op->R, op->n0, op->n1, d_x, d_y);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_cross_scale(fdadi_cross *op, const double *d_vec, int64_t vsize, void *stream) {
  FDADI_REQUIRE(op && d_vec, FDADI_ERR_ARG, "fdadi_cross_scale: null argument");
  FDADI_REQUIRE(vsize > 0 && op->N % vsize == 0, FDADI_ERR_ARG,
                "Vector length does not divide evenly into operator size. Cannot scale.");
  hipLaunchKernelGGL(cross_scale_kernel, dim3(grid_for(op->N, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), op->bands, op->R, d_vec, 0.0, vsize, op->N);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_cross_scale_scalar(fdadi_cross *op, double value, void *stream) {
  FDADI_REQUIRE(op, FDADI_ERR_ARG, "fdadi_cross_scale_scalar: null operator");
  // the scalar travels as a kernel argument: no allocation, copy or sync
  hipLaunchKernelGGL(cross_scale_kernel, dim3(grid_for(op->N, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), op->bands, op->R, nullptr, value, 1, op->N);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_heston_coefficients(int which, const double *d_spots, int64_t nS,
                              const double *d_vars, int64_t nv, const double *p,
                              double *d_out, void *stream) {
  FDADI_REQUIRE(d_spots && d_vars && p && d_out, FDADI_ERR_ARG,
                "fdadi_heston_coefficients: null argument");
  FDADI_REQUIRE(which >= 0 && which <= 4, FDADI_ERR_ARG,
                "fdadi_heston_coefficients: unknown derivative");
  FDADI_REQUIRE(nS > 0 && nv > 0, FDADI_ERR_ARG, "fdadi_heston_coefficients: empty grid");
  hipLaunchKernelGGL(heston_coef_kernel, dim3(grid_for(nS * nv, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), which, d_spots, nS, d_vars, nv, p[0], p[1], p[2], p[3],
                     p[4], d_out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

}  // extern "C"
