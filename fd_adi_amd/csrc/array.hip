// Runtime helpers, device arrays, elementwise ops and the device arena.
//
// Replaces SizedArray<double> (VecArray.h:211-710, VecArray.cu:29-196):
// buffers are plain device pointers, the elementwise ops are grid-stride
// kernels on the caller's stream, and the transpose is an LDS-tiled kernel.
#include "common.h"

#include <mutex>
#include <vector>

namespace fdadi {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

constexpr int kBlock = 256;

// Grid-stride elementwise kernels; op: 0 add, 1 sub, 2 (a-b)/2, 3 mul.
__global__ void binop_kernel(double *__restrict__ a, const double *__restrict__ b,
                             int64_t n, int op) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double x = a[i], y = b[i];
    double r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = (x - y) * 0.5; break;
      default: r = x * y; break;
    }
    a[i] = r;
  }
}

__global__ void scalar_kernel(double *__restrict__ a, const double *__restrict__ sp,
                              double s, int64_t n, int op) {
  const double y = sp ? sp[0] : s;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double x = a[i];
    double r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = (x - y) * 0.5; break;
      default: r = x * y; break;
    }
    a[i] = r;
  }
}

__global__ void fill_kernel(double *__restrict__ a, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// (h, w) -> (w, h) through a 32x33 LDS tile (bank-conflict padding).
__global__ void transpose_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                 int64_t h, int64_t w) {
  __shared__ double tile[32][33];
  const int64_t x0 = (int64_t)blockIdx.x * 32, y0 = (int64_t)blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int64_t y = y0 + k, x = x0 + threadIdx.x;
    if (y < h && x < w) tile[k][threadIdx.x] = src[y * w + x];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int64_t x = x0 + k, y = y0 + threadIdx.x;  // dst row x, col y
    if (x < w && y < h) dst[x * h + y] = tile[threadIdx.x][k];
  }
}

unsigned ew_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace
}  // namespace fdadi

using namespace fdadi;

struct fdadi_arena {
  char *base = nullptr;
  int64_t bytes = 0;
  int64_t used = 0;
};

extern "C" {

const char *fdadi_last_error(void) { return g_last_error.c_str(); }

int fdadi_version(int *major, int *minor) {
  if (major) *major = 1;
  if (minor) *minor = 0;
  return FDADI_OK;
}

int fdadi_device_count(int *n) {
  FDADI_REQUIRE(n, FDADI_ERR_ARG, "fdadi_device_count: null output");
  FDADI_TRY_HIP(hipGetDeviceCount(n));
  return FDADI_OK;
}

int fdadi_set_device(int dev) {
  FDADI_TRY_HIP(hipSetDevice(dev));
  return FDADI_OK;
}

int fdadi_stream_create(void **stream) {
  FDADI_REQUIRE(stream, FDADI_ERR_ARG, "fdadi_stream_create: null output");
  hipStream_t s;
  FDADI_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return FDADI_OK;
}

int fdadi_stream_destroy(void *stream) {
  if (stream) FDADI_TRY_HIP(hipStreamDestroy(as_stream(stream)));
  return FDADI_OK;
}

int fdadi_stream_sync(void *stream) {
  FDADI_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
  return FDADI_OK;
}

int fdadi_device_sync(void) {
  FDADI_TRY_HIP(hipDeviceSynchronize());
  return FDADI_OK;
}

int fdadi_array_alloc(int64_t n, double **d_out) {
  FDADI_REQUIRE(d_out && n >= 0, FDADI_ERR_ARG, "fdadi_array_alloc: bad arguments");
  *d_out = nullptr;
  if (n == 0) return FDADI_OK;
  FDADI_TRY_HIP(hipMalloc((void **)d_out, (size_t)n * sizeof(double)));
  return FDADI_OK;
}

int fdadi_array_free(double *d_ptr) {
  if (d_ptr) FDADI_TRY_HIP(hipFree(d_ptr));
  return FDADI_OK;
}

int fdadi_array_h2d(double *d_dst, const double *src, int64_t n, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (d_dst && src)), FDADI_ERR_ARG,
                "fdadi_array_h2d: bad arguments");
  if (n == 0) return FDADI_OK;
  // Pageable host memory: the copy is staged and returns once src is consumed.
  FDADI_TRY_HIP(hipMemcpyAsync(d_dst, src, (size_t)n * sizeof(double),
                               hipMemcpyHostToDevice, as_stream(stream)));
  FDADI_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
  return FDADI_OK;
}

int fdadi_array_d2h(double *dst, const double *d_src, int64_t n, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (dst && d_src)), FDADI_ERR_ARG,
                "fdadi_array_d2h: bad arguments");
  if (n == 0) return FDADI_OK;
  FDADI_TRY_HIP(hipMemcpyAsync(dst, d_src, (size_t)n * sizeof(double),
                               hipMemcpyDeviceToHost, as_stream(stream)));
  FDADI_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
  return FDADI_OK;
}

int fdadi_array_copy(double *d_dst, const double *d_src, int64_t n, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (d_dst && d_src)), FDADI_ERR_ARG,
                "fdadi_array_copy: bad arguments");
  if (n == 0 || d_dst == d_src) return FDADI_OK;
  FDADI_TRY_HIP(hipMemcpyAsync(d_dst, d_src, (size_t)n * sizeof(double),
                               hipMemcpyDeviceToDevice, as_stream(stream)));
  return FDADI_OK;
}

int fdadi_array_fill(double *d_dst, double value, int64_t n, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || d_dst), FDADI_ERR_ARG, "fdadi_array_fill: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(ew_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     d_dst, value, n);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_array_binop(double *d_a, const double *d_b, int64_t n, int op, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (d_a && d_b)) && op >= 0 && op <= 3, FDADI_ERR_ARG,
                "fdadi_array_binop: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(binop_kernel, dim3(ew_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     d_a, d_b, n, op);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_array_scalar_op(double *d_a, double s, int64_t n, int op, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || d_a) && op >= 0 && op <= 3, FDADI_ERR_ARG,
                "fdadi_array_scalar_op: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(scalar_kernel, dim3(ew_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     d_a, (const double *)nullptr, s, n, op);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_array_scalar_op_dev(double *d_a, const double *d_s, int64_t n, int op,
                              void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (d_a && d_s)) && op >= 0 && op <= 3, FDADI_ERR_ARG,
                "fdadi_array_scalar_op_dev: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(scalar_kernel, dim3(ew_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     d_a, d_s, 0.0, n, op);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_array_transpose(double *d_dst, const double *d_src, int64_t h, int64_t w,
                          void *stream) {
  FDADI_REQUIRE(h >= 0 && w >= 0 && d_dst != d_src, FDADI_ERR_ARG,
                "fdadi_array_transpose: bad arguments (in-place not supported)");
  if (h == 0 || w == 0) return FDADI_OK;
  FDADI_REQUIRE(d_dst && d_src, FDADI_ERR_ARG, "fdadi_array_transpose: null pointer");
  dim3 grid((unsigned)((w + 31) / 32), (unsigned)((h + 31) / 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, as_stream(stream), d_dst, d_src,
                     h, w);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_arena_create(int64_t bytes, fdadi_arena **out) {
  FDADI_REQUIRE(out && bytes > 0, FDADI_ERR_ARG, "fdadi_arena_create: bad arguments");
  fdadi_arena *a = new fdadi_arena();
  hipError_t e = hipMalloc((void **)&a->base, (size_t)bytes);
  if (e != hipSuccess) {
    delete a;
    set_error(std::string("fdadi_arena_create: hipMalloc: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  a->bytes = bytes;
  *out = a;
  return FDADI_OK;
}

int fdadi_arena_alloc(fdadi_arena *a, int64_t n_doubles, double **d_out) {
  FDADI_REQUIRE(a && d_out && n_doubles >= 0, FDADI_ERR_ARG, "fdadi_arena_alloc: bad arguments");
  int64_t need = ((n_doubles * (int64_t)sizeof(double)) + 255) / 256 * 256;
  FDADI_REQUIRE(a->used + need <= a->bytes, FDADI_ERR_ARG,
                "fdadi_arena_alloc: arena exhausted (" + std::to_string(a->used + need) +
                    " > " + std::to_string(a->bytes) + " bytes)");
  *d_out = (double *)(a->base + a->used);
  a->used += need;
  return FDADI_OK;
}

int fdadi_arena_used(const fdadi_arena *a, int64_t *bytes) {
  FDADI_REQUIRE(a && bytes, FDADI_ERR_ARG, "fdadi_arena_used: bad arguments");
  *bytes = a->used;
  return FDADI_OK;
}

int fdadi_arena_reset(fdadi_arena *a) {
  FDADI_REQUIRE(a, FDADI_ERR_ARG, "fdadi_arena_reset: null arena");
  a->used = 0;
  return FDADI_OK;
}

int fdadi_arena_destroy(fdadi_arena *a) {
  if (!a) return FDADI_OK;
  if (a->base) FDADI_TRY_HIP(hipFree(a->base));
  delete a;
  return FDADI_OK;
}

}  // extern "C"
