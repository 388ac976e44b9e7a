// Device-resident serving of the price surrogate: the gate's regression
// trees, the ensemble combination, the gate and the fallback batch's inputs,
// all on the device so that a batch of queries in HBM is answered without a
// host round trip (only the count of fallback queries comes back).
//
// No counterpart in the reference (its only pricers are the ADI engines and
// the COS method, FiniteDifference/heston.py:490-693).
//
// Feature rows (8 doubles): kappa, theta, sigma, rho, v0, r, T, K/S0 (S0 = 100).
#include "common.h"

#include <vector>

namespace fdadi {
namespace {

// ---- regression trees ------------------------------------------------------------
// A fitted ensemble of binary regression trees (gradient boosting): every
// internal node sends x to `left` when x[feature] <= thr, else to `right`;
// the prediction is base + the sum of the reached leaves' values, tree by
// tree in order.  Children have larger indices than their parent (checked
// at creation), so every walk ends.
struct FNode {
  double thr, value;
  int32_t feature, left, right, leaf;
};

__global__ void forest_predict_kernel(const FNode *__restrict__ nodes,
                                      const int32_t *__restrict__ roots, int ntrees, double base,
                                      const double *__restrict__ X, int64_t n, int nfeat,
                                      double *__restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double *x = X + r * nfeat;
  double acc = base;
  for (int t = 0; t < ntrees; ++t) {
    int i = roots[t];
    FNode nd = nodes[i];
    while (!nd.leaf) {
      i = x[nd.feature] <= nd.thr ? nd.left : nd.right;
      nd = nodes[i];
    }
    acc += nd.value;
  }
  out[r] = acc;
}

// ---- ensemble combination and gate features ----------------------------------------
// price = baseline + mean of the M members' corrections (summed member by
// member); spread = their standard deviation (M - 1 in the denominator);
// the error model's 13 features: log of kappa, theta, sigma, v0, T, K/S0;
// rho and r as they are; log Feller ratio 2 kappa theta / sigma^2; log10 of
// spread + 1e-12; the mean correction; log10 of its magnitude + 1e-9; log10
// of the members' range (max - min) + 1e-12.  inside = the row lies in the
// box [lo, hi] and its Feller ratio is at least feller_min.
struct Box {
  double lo[8], hi[8];
};

constexpr int kGateFeatures = 13;  // columns of the error model's feature rows

__global__ void serve_combine_kernel(const double *__restrict__ P, int64_t n,
                                     const double *__restrict__ Y, int M,
                                     const double *__restrict__ base, Box box, double feller_min,
                                     double *__restrict__ price, double *__restrict__ X,
                                     int32_t *__restrict__ inside) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double *p = P + 8 * r;
  double sum = 0.0, ymin = Y[r], ymax = Y[r];
  for (int m = 0; m < M; ++m) {
    const double v = Y[m * n + r];
    sum += v;
    ymin = fmin(ymin, v);
    ymax = fmax(ymax, v);
  }
  const double mean = sum / M;
  double ss = 0.0;
  for (int m = 0; m < M; ++m) {
    const double d = Y[m * n + r] - mean;
    ss += d * d;
  }
  const double spread = M > 1 ? sqrt(ss / (M - 1)) : 0.0;
  price[r] = base[r] + mean;
  const double feller = 2.0 * p[0] * p[1] / (p[2] * p[2]);
  double *x = X + kGateFeatures * r;
  x[0] = log(p[0]);
  x[1] = log(p[1]);
  x[2] = log(p[2]);
  x[3] = p[3];
  x[4] = log(p[4]);
  x[5] = p[5];
  x[6] = log(p[6]);
  x[7] = log(p[7]);
  x[8] = log(feller);
  x[9] = log10(spread + 1e-12);
  x[10] = mean;
  x[11] = log10(fabs(mean) + 1e-9);
  x[12] = log10(ymax - ymin + 1e-12);
  bool in = feller >= feller_min;
  for (int k = 0; k < 8; ++k) in = in && p[k] >= box.lo[k] && p[k] <= box.hi[k];
  inside[r] = in ? 1 : 0;
}

// Rows of the semi-analytic pricer (S0, K, r, v0, T, kappa, theta, sigma,
// rho) from feature rows.
__global__ void cf_rows_kernel(const double *__restrict__ P, int64_t n, double *__restrict__ R) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double *p = P + 8 * r;
  double *q = R + 9 * r;
  q[0] = 100.0;
  q[1] = 100.0 * p[7];
  q[2] = p[5];
  q[3] = p[4];
  q[4] = p[6];
  q[5] = p[0];
  q[6] = p[1];
  q[7] = p[2];
  q[8] = p[3];
}

// ---- gate and the fallback list ----------------------------------------------------
// ok = inside && score <= tau.  The rejected rows are listed in order:
// per-block counts, one block's exclusive scan of them, then the scatter.
constexpr int kSelBlock = 256;

__global__ void gate_count_kernel(const double *__restrict__ score,
                                  const int32_t *__restrict__ inside, int64_t n, double tau,
                                  int32_t *__restrict__ ok, int32_t *__restrict__ counts) {
  __shared__ int wsum[kSelBlock / 64];
  const int64_t r = blockIdx.x * (int64_t)kSelBlock + threadIdx.x;
  int rej = 0;
  if (r < n) {
    const bool a = inside[r] != 0 && score[r] <= tau;
    ok[r] = a ? 1 : 0;
    rej = a ? 0 : 1;
  }
  const unsigned long long b = __ballot(rej);
  if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int w = 0; w < kSelBlock / 64; ++w) c += wsum[w];
    counts[blockIdx.x] = c;
  }
}

// Exclusive scan of the block counts (one work-group, any number of blocks)
// and the total at counts[nblocks].
__global__ void scan_counts_kernel(int32_t *counts, int nblocks) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nblocks + 1023) / 1024;
  int local = 0;
  for (int q = 0; q < per; ++q) {
    const int i = t * per + q;
    if (i < nblocks) local += counts[i];
  }
  part[t] = local;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int i = 0; i < 1024; ++i) {
      const int v = part[i];
      part[i] = run;
      run += v;
    }
    counts[nblocks] = run;
  }
  __syncthreads();
  int run = part[t];
  for (int q = 0; q < per; ++q) {
    const int i = t * per + q;
    if (i < nblocks) {
      const int v = counts[i];
      counts[i] = run;
      run += v;
    }
  }
}

__global__ void gate_scatter_kernel(const int32_t *__restrict__ ok, int64_t n,
                                    const int32_t *__restrict__ offsets, int32_t *__restrict__ idx) {
  __shared__ int wbase[kSelBlock / 64];
  const int64_t r = blockIdx.x * (int64_t)kSelBlock + threadIdx.x;
  const int rej = r < n && ok[r] == 0 ? 1 : 0;
  const unsigned long long b = __ballot(rej);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) wbase[w] = __popcll(b);
  __syncthreads();
  int before = offsets[blockIdx.x];
  for (int q = 0; q < w; ++q) before += wbase[q];
  const unsigned long long below = lane ? (b & ((1ull << lane) - 1ull)) : 0ull;
  if (rej) idx[before + __popcll(below)] = (int32_t)r;
}

// ---- fallback batch inputs -----------------------------------------------------------
// For rows idx[i] of P (i < m; rows m .. cap-1 repeat the last one): the
// engine parameters (r, kappa, theta, sigma, rho, T / nt), the spot mesh
// K * unit (K = 100 K/S0), and the strike, S0 and v0 of the read-out.
__global__ void label_inputs_kernel(const double *__restrict__ P, const int32_t *__restrict__ idx,
                                    int m, int cap, int nt, const double *__restrict__ unit,
                                    int nS, double *__restrict__ params,
                                    double *__restrict__ spots, double *__restrict__ K,
                                    double *__restrict__ S0, double *__restrict__ v0) {
  const int k = blockIdx.x;
  if (k >= cap) return;
  const int i = k < m ? k : m - 1;
  const int64_t row = idx ? idx[i] : i;
  const double *p = P + 8 * row;
  const double strike = 100.0 * p[7];
  for (int j = threadIdx.x; j < nS; j += blockDim.x) spots[(int64_t)k * nS + j] = strike * unit[j];
  if (threadIdx.x == 0) {
    double *q = params + 6 * (int64_t)k;
    q[0] = p[5];
    q[1] = p[0];
    q[2] = p[1];
    q[3] = p[2];
    q[4] = p[3];
    q[5] = p[6] / nt;
    K[k] = strike;
    S0[k] = 100.0;
    v0[k] = p[4];
  }
}

// out[r] = the largest of the k rows' values S[m * n + r] (the error models'
// estimates; exact, so host and device agree bit for bit).
__global__ void max_rows_kernel(const double *__restrict__ S, int k, int64_t n,
                                double *__restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v = S[r];
  for (int m = 1; m < k; ++m) v = fmax(v, S[(int64_t)m * n + r]);
  out[r] = v;
}

__global__ void scatter_kernel(const double *__restrict__ src, const int32_t *__restrict__ idx,
                               int m, double *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) dst[idx[i]] = src[i];
}

}  // namespace
}  // namespace fdadi

using namespace fdadi;

struct fdadi_forest {
  int ntrees = 0, nfeat = 0;
  int64_t nnodes = 0;
  double base = 0.0;
  FNode *d_nodes = nullptr;
  int32_t *d_roots = nullptr;
};

extern "C" {

int fdadi_forest_create(int32_t ntrees, const int64_t *tree_start, int32_t nfeat,
                        const int32_t *feature, const double *threshold, const int32_t *left,
                        const int32_t *right, const uint8_t *is_leaf, const double *value,
                        double base, fdadi_forest **out) {
  FDADI_REQUIRE(out && tree_start && feature && threshold && left && right && is_leaf && value &&
                    ntrees > 0 && nfeat > 0,
                FDADI_ERR_ARG, "fdadi_forest_create: bad arguments");
  const int64_t nn = tree_start[ntrees];
  FDADI_REQUIRE(nn > 0 && nn < (int64_t)1 << 30, FDADI_ERR_ARG,
                "fdadi_forest_create: node count out of range");
  std::vector<FNode> h((size_t)nn);
  std::vector<int32_t> roots((size_t)ntrees);
  for (int t = 0; t < ntrees; ++t) {
    const int64_t a = tree_start[t], b = tree_start[t + 1];
    FDADI_REQUIRE(a >= 0 && b > a && b <= nn, FDADI_ERR_ARG,
                  "fdadi_forest_create: tree " + std::to_string(t) + ": bad node range");
    roots[t] = (int32_t)a;
    for (int64_t i = a; i < b; ++i) {
      FNode nd{threshold[i], value[i], feature[i], 0, 0, is_leaf[i] ? 1 : 0};
      if (!nd.leaf) {
        // children local to the tree, after their parent: every walk ends
        const int64_t l = a + left[i], r = a + right[i];
        FDADI_REQUIRE(left[i] > i - a && right[i] > i - a && l < b && r < b &&
                          feature[i] >= 0 && feature[i] < nfeat,
                      FDADI_ERR_ARG,
                      "fdadi_forest_create: tree " + std::to_string(t) + ": bad node " +
                          std::to_string(i - a));
        nd.left = (int32_t)l;
        nd.right = (int32_t)r;
      }
      h[(size_t)i] = nd;
    }
  }
  fdadi_forest *f = new fdadi_forest();
  f->ntrees = ntrees;
  f->nfeat = nfeat;
  f->nnodes = nn;
  f->base = base;
  hipError_t e = hipMalloc((void **)&f->d_nodes, sizeof(FNode) * nn);
  if (e == hipSuccess) e = hipMalloc((void **)&f->d_roots, sizeof(int32_t) * ntrees);
  if (e == hipSuccess)
    e = hipMemcpy(f->d_nodes, h.data(), sizeof(FNode) * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f->d_roots, roots.data(), sizeof(int32_t) * ntrees, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(f->d_nodes);
    (void)hipFree(f->d_roots);
    delete f;
    set_error(std::string("fdadi_forest_create: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = f;
  return FDADI_OK;
}

int fdadi_forest_destroy(fdadi_forest *f) {
  if (!f) return FDADI_OK;
  (void)hipFree(f->d_nodes);
  (void)hipFree(f->d_roots);
  delete f;
  return FDADI_OK;
}

int fdadi_forest_predict(const fdadi_forest *f, const double *d_X, int64_t n, double *d_out,
                         void *stream) {
  FDADI_REQUIRE(f && n >= 0 && (n == 0 || (d_X && d_out)), FDADI_ERR_ARG,
                "fdadi_forest_predict: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(forest_predict_kernel, dim3(grid_for(n, 128)), dim3(128), 0,
                     as_stream(stream), f->d_nodes, f->d_roots, f->ntrees, f->base, d_X, n,
                     f->nfeat, d_out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_serve_combine(const double *d_P, int64_t n, const double *d_Y, int32_t M,
                        const double *d_base, const double *lo, const double *hi,
                        double feller_min, double *d_price, double *d_X, int32_t *d_inside,
                        void *stream) {
  FDADI_REQUIRE(n >= 0 && M >= 1 && lo && hi, FDADI_ERR_ARG, "fdadi_serve_combine: bad arguments");
  if (n == 0) return FDADI_OK;
  FDADI_REQUIRE(d_P && d_Y && d_base && d_price && d_X && d_inside, FDADI_ERR_ARG,
                "fdadi_serve_combine: null buffer");
  Box box;
  for (int k = 0; k < 8; ++k) {
    box.lo[k] = lo[k];
    box.hi[k] = hi[k];
  }
  hipLaunchKernelGGL(serve_combine_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     as_stream(stream), d_P, n, d_Y, M, d_base, box, feller_min, d_price, d_X,
                     d_inside);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_serve_cf_rows(const double *d_P, int64_t n, double *d_rows, void *stream) {
  FDADI_REQUIRE(n >= 0 && (n == 0 || (d_P && d_rows)), FDADI_ERR_ARG,
                "fdadi_serve_cf_rows: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(cf_rows_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     d_P, n, d_rows);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_serve_gate(const double *d_score, const int32_t *d_inside, int64_t n, double tau,
                     int32_t *d_ok, int32_t *d_idx, int32_t *d_work, int32_t *n_rejected,
                     void *stream) {
  FDADI_REQUIRE(n >= 0 && n < (int64_t)1 << 31 && n_rejected, FDADI_ERR_ARG,
                "fdadi_serve_gate: bad arguments");
  *n_rejected = 0;
  if (n == 0) return FDADI_OK;
  FDADI_REQUIRE(d_score && d_inside && d_ok && d_idx && d_work, FDADI_ERR_ARG,
                "fdadi_serve_gate: null buffer");
  hipStream_t s = as_stream(stream);
  const int nb = (int)grid_for(n, kSelBlock);
  hipLaunchKernelGGL(gate_count_kernel, dim3(nb), dim3(kSelBlock), 0, s, d_score, d_inside, n,
                     tau, d_ok, d_work);
  FDADI_CHECK_LAUNCH();
  hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, d_work, nb);
  FDADI_CHECK_LAUNCH();
  hipLaunchKernelGGL(gate_scatter_kernel, dim3(nb), dim3(kSelBlock), 0, s, d_ok, n, d_work,
                     d_idx);
  FDADI_CHECK_LAUNCH();
  int32_t total = 0;
  FDADI_TRY_HIP(hipMemcpyAsync(&total, d_work + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  FDADI_TRY_HIP(hipStreamSynchronize(s));
  *n_rejected = total;
  return FDADI_OK;
}

int fdadi_serve_label_inputs(const double *d_P, const int32_t *d_idx, int32_t m, int32_t cap,
                             int32_t nt, const double *d_unit, int32_t nS, double *d_params,
                             double *d_spots, double *d_K, double *d_S0, double *d_v0,
                             void *stream) {
  FDADI_REQUIRE(m >= 1 && cap >= m && nt >= 1 && nS >= 1 && d_P && d_unit && d_params &&
                    d_spots && d_K && d_S0 && d_v0,
                FDADI_ERR_ARG, "fdadi_serve_label_inputs: bad arguments");
  hipLaunchKernelGGL(label_inputs_kernel, dim3(cap), dim3(128), 0, as_stream(stream), d_P, d_idx,
                     m, cap, nt, d_unit, nS, d_params, d_spots, d_K, d_S0, d_v0);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_serve_max_rows(const double *d_S, int32_t k, int64_t n, double *d_out, void *stream) {
  FDADI_REQUIRE(k >= 1 && n >= 0 && (n == 0 || (d_S && d_out)), FDADI_ERR_ARG,
                "fdadi_serve_max_rows: bad arguments");
  if (n == 0) return FDADI_OK;
  hipLaunchKernelGGL(max_rows_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     d_S, k, n, d_out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_serve_scatter(const double *d_src, const int32_t *d_idx, int32_t m, double *d_dst,
                        void *stream) {
  FDADI_REQUIRE(m >= 0 && (m == 0 || (d_src && d_idx && d_dst)), FDADI_ERR_ARG,
                "fdadi_serve_scatter: bad arguments");
  if (m == 0) return FDADI_OK;
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(m, 256)), dim3(256), 0, as_stream(stream),
                     d_src, d_idx, m, d_dst);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

}  // extern "C"
