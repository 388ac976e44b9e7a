// Learned price surrogate: an MLP 8 -> 512 -> 512 -> 512 -> 512 -> 1 with
// SiLU, BF16 weights and activations, FP32 accumulation and FP32 master
// weights, trained with Adam by hand-written MFMA kernels (gfx950).
//
// GEMMs are "NT": C[m,n] = sum_k A[m,k] * B[n,k] with both operands
// row-major and K-contiguous, bf16.  Every product the network needs is put
// in that form by keeping transposed copies (written by the producing
// epilogue, never by a separate transpose):
//   forward     Z_l   = A_{l-1} W_l^T       A = A_{l-1} [B x in],  B = W_l   [out x in]
//   backward    dA_l  = dZ_{l+1} W_{l+1}    A = dZ      [B x out], B = W^T   [in x out]
//   weight grad dW_l  = dZ_l^T A_{l-1}      A = dZ^T    [out x B], B = A^T   [in x B]
// The input layer is zero-padded from 8 to 128 features so every dimension
// is a multiple of the 128 x 128 work-group tile.
#include "common.h"

#include <mutex>

#include <cmath>
#include <vector>

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

#define FDADI_RET(x)     \
  do {                   \
    int _st = (x);       \
    if (_st) return _st; \
  } while (0)

namespace fdadi {
namespace {

constexpr int kIn = 8;       // Heston features
constexpr int kMaxIn = 16;   // largest feature count of a network (fdadi_mlp_create_features)
constexpr int kInPad = 128;  // padded input width
constexpr int kH = 512;      // hidden width
constexpr int kLayers = 4;   // hidden layers (the 512 -> 1 head is separate)
constexpr int kTile = 128;
constexpr int kKStep = 64;   // K extent of one LDS stage of the GEMMs
constexpr int kMaxSplits = 16;  // split-K factor cap for the weight-gradient GEMMs
constexpr int kLdsPitch = kKStep + 8;  // bf16 elements per LDS row (16 B pad)
constexpr int kThreads = 512;          // GEMM work-group: 8 waves
constexpr int kWN = 4;                 // waves along N (2 along M)
constexpr int kBM = 4, kBN = 2;        // 16 x 16 MFMA blocks per wave: 64 x 32

enum Epi { EPI_FWD = 0, EPI_INF = 1, EPI_BWD = 2, EPI_F32 = 3 };

// Logistic sigmoid with the hardware reciprocal (about 1 ulp; every result
// that uses it is rounded to bf16).  silu(z) = z s(z), silu'(z) = s (1 + z (1 - s)).
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

struct EpiArgs {
  const float *bias;  // [N] (EPI_FWD/EPI_INF)
  bf16 *dT;           // [N x M] SiLU derivative at the pre-activations, transposed
                      // (EPI_FWD writes, EPI_BWD reads)
  bf16 *out;          // [M x N] bf16 result
  bf16 *outT;         // [N x M] bf16 transposed result (EPI_FWD, EPI_BWD)
  float *C;           // [M x N] fp32 result (EPI_F32); with split-K the
                      // [splits x M x N] partial products
  const bf16 *w4;     // EPI_FWD, last hidden layer: the head's weights [N] ...
  float *ypart;       // ... and its partial outputs [N/128 x M], one per column tile
};

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
constexpr int kEpPitch = kTile + 8;  // bf16 elements per row of an epilogue tile
// LDS of the GEMM: the K loop's two A and two B stages; the epilogue tiles
// reuse it (one 128 x 128 tile for EPI_INF, two for EPI_BWD, three for
// EPI_FWD).
constexpr int kLoopLds = 4 * kTile * kLdsPitch;
constexpr int kTileLds = kTile * kEpPitch;
constexpr int gemm_lds(int epi) {
  return epi == EPI_F32 ? kLoopLds
         : epi == EPI_INF ? (kLoopLds > kTileLds ? kLoopLds : kTileLds)
         : epi == EPI_BWD ? (kLoopLds > 2 * kTileLds ? kLoopLds : 2 * kTileLds)
                          : (kLoopLds > 3 * kTileLds ? kLoopLds : 3 * kTileLds);
}

// C = A B^T on MFMA 16x16x32 bf16.  512 threads = 8 waves in a 2 x 4 grid,
// each wave owns a 64 x 32 sub-tile = 4 x 2 MFMA blocks; two waves per SIMD
// hide each other's LDS and barrier latency.  The K loop stages 128 x 64
// tiles of A and B in LDS (two 32-deep MFMA steps per stage).
// MFMA operand layout (v_mfma_f32_16x16x32_bf16): lane l holds
//   A[i = l%16][k = 8*(l/16) .. +7],  B[k = 8*(l/16) .. +7][j = l%16],
//   D[i = 4*(l/16) + t][j = l%16] for t = 0..3.
// The grids here are about one work-group per CU, so the loads of a K stage
// must be in flight long before it: a ring of kDepth register stages keeps
// the next stage's global loads outstanding while the MFMAs of the current
// one run.  The K extent of a work-group is a multiple of kDepth stages, so
// the ring needs no remainder code.
constexpr int kDepth = 2;
constexpr int kKBlock = kKStep * kDepth;  // K granularity of the GEMMs
constexpr int kChunks = kTile * kKStep / 8 / kThreads;  // 16-byte loads per operand per stage

// One GEMM of a launch: operands, sizes, epilogue and its tile grid (gx x gy
// output tiles, gz splits of K).
struct GemmArgs {
  const bf16 *A, *B;
  int M, N, K, gx, gy, gz;
  EpiArgs ep;
};

// The work of work-group `Lin` (linear index within its GEMM's grid).
template <int EPI>
__device__ __forceinline__ void gemm_tile(bf16 *smem, const GemmArgs &g, int Lin) {
  const bf16 *__restrict__ A = g.A;
  const bf16 *__restrict__ B = g.B;
  const int M = g.M, N = g.N, K = g.K;
  const EpiArgs &ep = g.ep;
  // Two LDS stages: the MFMAs of stage s read buffer s % 2 while stage s + 1
  // is written to the other one, so one barrier per stage suffices.
  bf16 *const As[2] = {smem, smem + kTile * kLdsPitch};
  bf16 *const Bs[2] = {smem + 2 * kTile * kLdsPitch, smem + 3 * kTile * kLdsPitch};
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int wm = wave / kWN, wn = wave % kWN;
  // Work-groups are dealt to the 8 XCDs round-robin by linear id; renumber
  // them so that each XCD gets a contiguous run of tiles (x fastest): the
  // N tiles of one row block of A, and of one split, then share that XCD's L2.
  const int gx = g.gx, gy = g.gy, total = gx * gy * g.gz;
  int L = Lin;
  if (total % 8 == 0) L = (L % 8) * (total / 8) + L / 8;
  const int bx = L % gx, by = (L / gx) % gy, bz = L / (gx * gy);
  const int m0 = by * kTile, n0 = bx * kTile;
  // split-K: work-group z covers K columns [z*Ks, (z+1)*Ks) (gz = 1: all)
  const int Ks = K / g.gz, kbeg = bz * Ks, nsteps = Ks / kKStep;
  f32x4 acc[kBM][kBN];
#pragma unroll
  for (int a = 0; a < kBM; ++a)
#pragma unroll
    for (int b = 0; b < kBN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Global -> registers -> LDS: chunk q = tid + kThreads * p of a 128 x 64
  // tile is row q / 8, elements 8 (q % 8) .. +7.
  const int lrow = tid / (kKStep / 8), lk = (tid % (kKStep / 8)) * 8;
  constexpr int kRowStep = kThreads / (kKStep / 8);  // rows between a thread's chunks
  const bf16 *Ag = A + (size_t)(m0 + lrow) * K + kbeg + lk;
  const bf16 *Bg = B + (size_t)(n0 + lrow) * K + kbeg + lk;
  bf16x8 ra[kDepth][kChunks], rb[kDepth][kChunks];
  auto fetch = [&](bf16x8 *xa, bf16x8 *xb, int step) {
    const int k = step * kKStep;
#pragma unroll
    for (int p = 0; p < kChunks; ++p) {
      xa[p] = *reinterpret_cast<const bf16x8 *>(Ag + (size_t)kRowStep * p * K + k);
      xb[p] = *reinterpret_cast<const bf16x8 *>(Bg + (size_t)kRowStep * p * K + k);
    }
  };
  auto stage = [&](const bf16x8 *xa, const bf16x8 *xb, int buf) {
#pragma unroll
    for (int p = 0; p < kChunks; ++p) {
      const int r = lrow + kRowStep * p;
      *reinterpret_cast<bf16x8 *>(As[buf] + r * kLdsPitch + lk) = xa[p];
      *reinterpret_cast<bf16x8 *>(Bs[buf] + r * kLdsPitch + lk) = xb[p];
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < kKStep; kk += 32) {
      bf16x8 af[kBM], bfr[kBN];
      const int kseg = kk + (lane / 16) * 8;
#pragma unroll
      for (int a = 0; a < kBM; ++a) {
        const int row = wm * kBM * 16 + a * 16 + lane % 16;
        af[a] = *reinterpret_cast<const bf16x8 *>(As[buf] + row * kLdsPitch + kseg);
      }
#pragma unroll
      for (int b = 0; b < kBN; ++b) {
        const int col = wn * kBN * 16 + b * 16 + lane % 16;
        bfr[b] = *reinterpret_cast<const bf16x8 *>(Bs[buf] + col * kLdsPitch + kseg);
      }
#pragma unroll
      for (int a = 0; a < kBM; ++a)
#pragma unroll
        for (int b = 0; b < kBN; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
  };
#pragma unroll
  for (int j = 0; j < kDepth; ++j) fetch(ra[j], rb[j], j);
  stage(ra[0], rb[0], 0);
  __syncthreads();
  // Stage s computes from LDS buffer s % 2; its registers (ring slot s %
  // kDepth) were staged one stage earlier and now take stage s + kDepth.
  // The loop body is free of branches (past the end the loads repeat the
  // last stage and the store to LDS is never read), so the compiler's wait
  // counts keep the next stage's loads in flight.
  for (int s0 = 0; s0 < nsteps; s0 += kDepth) {
#pragma unroll
    for (int j = 0; j < kDepth; ++j) {
      const int st = s0 + j;
      fetch(ra[j], rb[j], min(st + kDepth, nsteps - 1));
      compute(j & 1);
      stage(ra[(j + 1) % kDepth], rb[(j + 1) % kDepth], (j + 1) & 1);
      __syncthreads();
    }
  }
  // Epilogue.  Lane l of wave (wm, wn) holds, for MFMA block (a, b), rows
  // ml = wm*64 + a*16 + 4*(l/16) + t (t = 0..3) of column nl = wn*32 + b*16 + l%16.
  auto row_of = [&](int a) { return wm * kBM * 16 + a * 16 + 4 * (lane / 16); };
  auto col_of = [&](int b) { return wn * kBN * 16 + b * 16 + lane % 16; };
  if (EPI == EPI_F32) {
    // split z writes its partial product to slice z of C
    float *Cz = ep.C + (size_t)bz * M * N;
#pragma unroll
    for (int a = 0; a < kBM; ++a)
#pragma unroll
      for (int b = 0; b < kBN; ++b) {
        const int n = n0 + col_of(b), mb = m0 + row_of(a);
#pragma unroll
        for (int t = 0; t < 4; ++t) Cz[(size_t)(mb + t) * N + n] = acc[a][b][t];
      }
    return;
  }
  // The bf16 results go through LDS tiles so that every global access is a
  // coalesced 16-byte one: T [ml][nl] (row-major) for out, TT [nl][ml] for
  // the transposed outT, DT [nl][ml] for dT.  The K loop ended with a
  // barrier, so its buffers are free.
  bf16 *const T = smem, *const TT = smem + kTileLds, *const DT = smem + 2 * kTileLds;
  // 2048 chunks of 8 elements per tile, 4 per thread: chunk q is row q / 16,
  // elements 8 (q % 16) .. +7 of that row (16 lanes cover one 256-byte row).
  constexpr int kTileChunks = kTile * kTile / 8 / kThreads;
  auto tile_to_global = [&](const bf16 *tile, bf16 *g, int ld, int r0, int c0) {
#pragma unroll
    for (int i = 0; i < kTileChunks; ++i) {
      const int q = tid + kThreads * i, r = q / 16, c = (q % 16) * 8;
      *reinterpret_cast<bf16x8 *>(g + (size_t)(r0 + r) * ld + c0 + c) =
          *reinterpret_cast<const bf16x8 *>(tile + r * kEpPitch + c);
    }
  };
  float r[kBM][kBN][4];
  if (EPI == EPI_BWD) {
    // r = (dZ_next W) * silu'(z): the derivative tile dT [n0.., m0..] into TT
#pragma unroll
    for (int i = 0; i < kTileChunks; ++i) {
      const int q = tid + kThreads * i, nl = q / 16, c = (q % 16) * 8;
      *reinterpret_cast<bf16x8 *>(TT + nl * kEpPitch + c) =
          *reinterpret_cast<const bf16x8 *>(ep.dT + (size_t)(n0 + nl) * M + m0 + c);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < kBM; ++a)
#pragma unroll
      for (int b = 0; b < kBN; ++b) {
        const bf16x4 d = *reinterpret_cast<const bf16x4 *>(TT + col_of(b) * kEpPitch + row_of(a));
#pragma unroll
        for (int t = 0; t < 4; ++t) r[a][b][t] = acc[a][b][t] * (float)d[t];
      }
    __syncthreads();
  } else {
#pragma unroll
    for (int b = 0; b < kBN; ++b) {
      const float bn = ep.bias[n0 + col_of(b)];
#pragma unroll
      for (int a = 0; a < kBM; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t) r[a][b][t] = acc[a][b][t] + bn;  // pre-activation z
    }
  }
  // activations (or back-propagated gradients) into T and TT; for the
  // forward pass also silu'(z) into DT
#pragma unroll
  for (int a = 0; a < kBM; ++a)
#pragma unroll
    for (int b = 0; b < kBN; ++b) {
      const int nl = col_of(b), ml = row_of(a);
      bf16x4 o, d;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float z = r[a][b][t];
        if (EPI == EPI_BWD) {
          o[t] = static_cast<bf16>(z);
        } else {
          const float sg = sigmoid(z);
          o[t] = static_cast<bf16>(z * sg);
          if (EPI == EPI_FWD) d[t] = static_cast<bf16>(sg * (1.0f + z * (1.0f - sg)));
        }
        T[(ml + t) * kEpPitch + nl] = o[t];
      }
      if (EPI != EPI_INF) *reinterpret_cast<bf16x4 *>(TT + nl * kEpPitch + ml) = o;
      if (EPI == EPI_FWD) *reinterpret_cast<bf16x4 *>(DT + nl * kEpPitch + ml) = d;
    }
  __syncthreads();
  if (EPI == EPI_FWD && ep.ypart) {
    // the head on this tile's 128 columns: four threads per row, 32 columns each
    const int row = tid / 4, h = (tid % 4) * 32;
    float acc_y = 0.f;
#pragma unroll
    for (int c = 0; c < 32; c += 8) {
      const bf16x8 av = *reinterpret_cast<const bf16x8 *>(T + row * kEpPitch + h + c);
      const bf16x8 wv = *reinterpret_cast<const bf16x8 *>(ep.w4 + n0 + h + c);
#pragma unroll
      for (int t = 0; t < 8; ++t) acc_y += (float)av[t] * (float)wv[t];
    }
    acc_y += __shfl_xor(acc_y, 1, 64);
    acc_y += __shfl_xor(acc_y, 2, 64);
    if (h == 0) ep.ypart[(size_t)bx * M + m0 + row] = acc_y;
  }
  tile_to_global(T, ep.out, N, m0, n0);
  if (EPI == EPI_INF) return;
  tile_to_global(TT, ep.outT, M, n0, m0);
  if (EPI == EPI_FWD) tile_to_global(DT, ep.dT, M, n0, m0);
}

template <int EPI>
__global__ void __launch_bounds__(kThreads) gemm_nt_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 smem[gemm_lds(EPI)];
  gemm_tile<EPI>(smem, g, blockIdx.x + g.gx * (blockIdx.y + g.gy * blockIdx.z));
}

// Two independent GEMMs in one launch (work-groups [0, n1) the first, the
// rest the second), so that their work-groups share the CUs: a layer's
// backward-data and weight-gradient products.  n1 is a multiple of 8, so
// each GEMM's work-groups keep their XCD order.
template <int E1, int E2>
__global__ void __launch_bounds__(kThreads) gemm_pair_kernel(GemmArgs g1, GemmArgs g2, int n1) {
  constexpr int kL1 = gemm_lds(E1), kL2 = gemm_lds(E2);
  __shared__ __attribute__((aligned(16))) bf16 smem[kL1 > kL2 ? kL1 : kL2];
  if ((int)blockIdx.x < n1)
    gemm_tile<E1>(smem, g1, blockIdx.x);
  else
    gemm_tile<E2>(smem, g2, blockIdx.x - n1);
}

// Features -> padded, normalised bf16 input rows:
// x_f = 2 (p_f - lo_f) / (hi_f - lo_f) - 1 for the nin features of a row,
// 0 for the padding; one thread per element, features fastest (coalesced).
__device__ __forceinline__ bf16 input_value(const double *__restrict__ P,
                                            const int *__restrict__ idx, int rows, int nin,
                                            const double *__restrict__ lo,
                                            const double *__restrict__ hi, int r, int f) {
  float v = 0.f;
  if (r < rows && f < nin) {
    const int src = idx ? idx[r] : r;
    v = (float)(2.0 * (P[(size_t)src * nin + f] - lo[f]) / (hi[f] - lo[f]) - 1.0);
  }
  return static_cast<bf16>(v);
}

// X [Mpad x 128] (every element) and, when XT is given, its transpose
// XT [128 x Mpad].  With F = 8 ceil(nin / 8) (the feature columns that can be
// non-zero) one launch covers three ranges of threads: (row, feature < F)
// computes the value once and stores it to X and XT; (row, 8-column chunk
// >= F) and (8-row chunk of an XT row >= F) store zeros, 16 bytes each.
// `zero`, when given, is cleared (the loss accumulator of a training step).
__global__ void prep_input_kernel(const double *__restrict__ P, const int *__restrict__ idx,
                                  int rows, int nin, const double *__restrict__ lo,
                                  const double *__restrict__ hi, bf16 *X, bf16 *XT, int Mpad,
                                  float *zero) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int F = (nin + 7) / 8 * 8;
  const int64_t n1 = (int64_t)Mpad * F, n2 = (int64_t)Mpad * (kInPad - F) / 8;
  if (zero && t == 0) *zero = 0.f;
  if (t < n1) {
    const int f = (int)(t / Mpad), r = (int)(t % Mpad);
    const bf16 v = input_value(P, idx, rows, nin, lo, hi, r, f);
    X[(size_t)r * kInPad + f] = v;
    if (XT) XT[(size_t)f * Mpad + r] = v;
    return;
  }
  const bf16x8 z = {};
  if (t < n1 + n2) {
    const int64_t u = t - n1;
    const int per = (kInPad - F) / 8;
    *reinterpret_cast<bf16x8 *>(X + (u / per) * kInPad + F + (u % per) * 8) = z;
    return;
  }
  const int64_t u = t - n1 - n2;
  if (XT && u < (int64_t)(kInPad - F) * Mpad / 8)
    *reinterpret_cast<bf16x8 *>(XT + (size_t)F * Mpad + 8 * u) = z;
}

// Threads of prep_input_kernel for Mpad rows (with the transpose or without).
inline int64_t prep_threads(int Mpad, int nin, bool transpose) {
  const int F = (nin + 7) / 8 * 8;
  return (int64_t)Mpad * F + (int64_t)Mpad * (kInPad - F) / 8 +
         (transpose ? (int64_t)(kInPad - F) * Mpad / 8 : 0);
}

// Head: y[r] = sum_n A[r,n] w[n] + b  (one wave per row, fp32).
__global__ void head_forward_kernel(const bf16 *__restrict__ Aact, const bf16 *__restrict__ w,
                                    const float *__restrict__ b, int rows, float *y) {
  const int r = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (r >= rows) return;
  float s = 0.f;
  for (int n = lane; n < kH; n += 64) s += (float)Aact[(size_t)r * kH + n] * (float)w[n];
  for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
  if (lane == 0) y[r] = s + b[0];
}

// Loss gradient and the head's backward on a 64 x 64 tile of (rows,
// features): y = b4 + the partial head outputs of the last hidden layer's
// column tiles, dy = 2 (y - t) / n per row (0 for padding rows >= valid),
// dZ = dy w4 silu'(z) with the derivative read from dT [features x rows];
// dZT is written along the rows, dZ along the features through an LDS tile,
// both with 16-byte accesses.  One loss partial per tile row block.
__global__ void __launch_bounds__(256)
head_backward_kernel(const float *__restrict__ ypart, const float *__restrict__ b4,
                     const double *__restrict__ T, const int *__restrict__ idx, int valid, int Mpad,
                     const bf16 *__restrict__ w, const bf16 *__restrict__ dT, float *dy,
                     bf16 *dZ, bf16 *dZT, float *loss_acc, double tmean, double tstd) {
  constexpr int kP = 72;  // tile pitch (bf16): rows of 144 bytes, 16-byte aligned
  __shared__ float gs[64];
  __shared__ float es[64];
  __shared__ __attribute__((aligned(16))) bf16 tile[64 * kP];
  const int r0 = blockIdx.y * 64, n0 = blockIdx.x * 64, tid = threadIdx.x;
  if (tid < 64) {
    const int r = r0 + tid;
    float g = 0.f, e2 = 0.f;
    if (r < valid) {
      // the head's output: bias + the column tiles' partial sums, in order
      float y = b4[0];
#pragma unroll
      for (int j = 0; j < kH / kTile; ++j) y += ypart[(size_t)j * Mpad + r];
      const float t = (float)((T[idx ? idx[r] : r] - tmean) / tstd);
      const float e = y - t;
      g = 2.f * e / (float)valid;
      e2 = e * e / (float)valid;
    }
    gs[tid] = g;
    es[tid] = e2;
    if (blockIdx.x == 0) dy[r] = g;
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid == 0) {
    float l = 0.f;
    for (int k = 0; k < 64; ++k) l += es[k];
    atomicAdd(loss_acc, l);
  }
  // 512 chunks of 8 rows of one feature, 2 per thread
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + 256 * i, nn = q / 8, rc = (q % 8) * 8;
    const size_t o = (size_t)(n0 + nn) * Mpad + r0 + rc;
    const bf16x8 d = *reinterpret_cast<const bf16x8 *>(dT + o);
    const float wn = (float)w[n0 + nn];
    bf16x8 v;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = static_cast<bf16>(gs[rc + t] * wn * (float)d[t]);
    *reinterpret_cast<bf16x8 *>(dZT + o) = v;
    *reinterpret_cast<bf16x8 *>(tile + nn * kP + rc) = v;
  }
  __syncthreads();
  // 512 chunks of 8 features of one row
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + 256 * i, rr = q / 8, nc = (q % 8) * 8;
    bf16x8 v;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = tile[(nc + t) * kP + rr];
    *reinterpret_cast<bf16x8 *>(dZ + (size_t)(r0 + rr) * kH + n0 + nc) = v;
  }
}

// Column reduction for the head: gw[n] = sum_r dy[r] A[r,n]; gb = sum_r dy[r].
// Reads the transposed activations AT[n][r] (contiguous in r), eight rows per
// 16-byte load (Mpad a multiple of 128).
__global__ void head_wgrad_kernel(const float *__restrict__ dy, const bf16 *__restrict__ AT,
                                  int Mpad, float *gw, float *gb) {
  const int n = blockIdx.x;  // 0..kH (kH -> bias)
  float s = 0.f;
  for (int q = threadIdx.x; q < Mpad / 8; q += blockDim.x) {
    const f32x4 d0 = *reinterpret_cast<const f32x4 *>(dy + 8 * q);
    const f32x4 d1 = *reinterpret_cast<const f32x4 *>(dy + 8 * q + 4);
    if (n < kH) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(AT + (size_t)n * Mpad + 8 * q);
#pragma unroll
      for (int t = 0; t < 4; ++t) s += d0[t] * (float)a[t] + d1[t] * (float)a[t + 4];
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) s += d0[t] + d1[t];
    }
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = blockDim.x / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (n < kH) gw[n] = red[0]; else gb[0] = red[0];
  }
}

// The weight and bias gradients of a layer, in one launch: work-groups
// [0, nsum) sum the split-K partials, out[i] = sum over z of part[z*n + i]
// in split order (deterministic, four elements per thread; none when the
// GEMM wrote the gradient itself); work-group nsum + j sums row j of
// dZ^T [kH x Mpad] into the bias gradient gb[j], eight elements per 16-byte
// load (Mpad a multiple of 128).
__global__ void __launch_bounds__(256)
layer_grad_kernel(const float *__restrict__ part, int splits, int64_t n, float *__restrict__ out,
                  int nsum, const bf16 *__restrict__ dZT, int Mpad, float *gb) {
  if ((int)blockIdx.x < nsum) {
    const int64_t i = 4 * (blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
    if (i >= n) return;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
    for (int z = 0; z < splits; ++z)
      acc += *reinterpret_cast<const f32x4 *>(part + (size_t)z * n + i);
    *reinterpret_cast<f32x4 *>(out + i) = acc;
    return;
  }
  const int j = (int)blockIdx.x - nsum;
  const bf16x8 *row = reinterpret_cast<const bf16x8 *>(dZT + (size_t)j * Mpad);
  float s = 0.f;
  for (int q = threadIdx.x; q < Mpad / 8; q += blockDim.x) {
    const bf16x8 v = row[q];
#pragma unroll
    for (int t = 0; t < 8; ++t) s += (float)v[t];
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = blockDim.x / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) gb[j] = red[0];
}

// Adam on the flat fp32 master parameters; the new value of a weight-matrix
// entry also goes to its bf16 copies (W and its transpose; the head's w4),
// which the next step's GEMMs read.
static_assert(kInPad == (1 << 7) && kH == (1 << 9), "adam_kernel indexes W_l by shifts");
struct CastMap {
  int64_t off[kLayers + 1];  // start of W_0..W_3 and of w4 in the flat array
  int lg_in[kLayers];        // log2 of the columns of W_l (128 or 512)
  bf16 *Wb[kLayers], *WbT[kLayers], *w4b;
};

__global__ void adam_kernel(float *p, const float *__restrict__ g, float *m, float *v, int64_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2,
                            CastMap cm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  const bf16 b = static_cast<bf16>(pi);
#pragma unroll
  for (int l = 0; l < kLayers; ++l) {
    const int64_t e = i - cm.off[l];
    if (e >= 0 && e < ((int64_t)kH << cm.lg_in[l])) {
      const int o = (int)e >> cm.lg_in[l], c = (int)e & ((1 << cm.lg_in[l]) - 1);
      cm.Wb[l][e] = b;
      cm.WbT[l][(size_t)c * kH + o] = b;
    }
  }
  if (i >= cm.off[kLayers] && i < cm.off[kLayers] + kH) cm.w4b[i - cm.off[kLayers]] = b;
}

// bf16 copies of a weight matrix W [out x in] and its transpose.
__global__ void cast_weights_kernel(const float *__restrict__ W, int out, int in, bf16 *Wb,
                                    bf16 *WbT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= out * in) return;
  const int o = i / in, c = i % in;
  const bf16 b = static_cast<bf16>(W[i]);
  Wb[i] = b;
  if (WbT) WbT[(size_t)c * out + o] = b;
}

__global__ void zero_kernel(float *p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

__global__ void unscale_kernel(const float *__restrict__ y, int rows, double tmean, double tstd,
                               double *out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) out[r] = (double)y[r] * tstd + tmean;
}

}  // namespace
}  // namespace fdadi

using namespace fdadi;

// Parameter layout in the flat fp32 master array:
//   W0 [kH x kInPad], b0 [kH], W1..W3 [kH x kH], b1..b3 [kH], w4 [kH], b4 [1].
struct fdadi_mlp {
  int max_rows = 0;  // capacity (multiple of 128)
  int64_t nparam = 0;
  int64_t off_W[kLayers] = {}, off_b[kLayers] = {}, off_w4 = 0, off_b4 = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // master, grad, Adam moments
  bf16 *Wb[kLayers] = {}, *WbT[kLayers] = {}, *w4b = nullptr;
  bf16 *X = nullptr, *XT = nullptr;
  bf16 *Act[kLayers] = {}, *ActT[kLayers] = {};
  bf16 *dT[kLayers] = {};  // silu'(pre-activations) [kH x rows], transposed
  bf16 *dZ = nullptr, *dZT = nullptr, *dZ2 = nullptr, *dZT2 = nullptr;
  float *y = nullptr, *dy = nullptr, *loss = nullptr;
  float *ypart = nullptr;  // training: head partial outputs [kH / 128 x rows]
  float *wpart = nullptr;  // split-K partial weight gradients [kMaxSplits x kH x kH]
  int n_in = kIn;          // features per input row
  double *lohi = nullptr;  // [lo(kMaxIn) | hi(kMaxIn)]
  double tmean = 0.0, tstd = 1.0;
  int64_t step = 0;
};

namespace fdadi {
namespace {

int in_dim(int l) { return l == 0 ? kInPad : kH; }

int gemm_args(int epi, const bf16 *A, const bf16 *B, int M, int N, int K, const EpiArgs &ep,
              int splits, GemmArgs *g) {
  FDADI_REQUIRE(M % kTile == 0 && N % kTile == 0 && K % kKBlock == 0, FDADI_ERR_ARG,
                "mlp gemm: dimensions must be multiples of 128 x 128 x 128");
  FDADI_REQUIRE(splits >= 1 && (splits == 1 || epi == EPI_F32) && K % (kKBlock * splits) == 0,
                FDADI_ERR_ARG, "mlp gemm: split-K needs the fp32 epilogue and K % (128 s) == 0");
  *g = GemmArgs{A, B, M, N, K, N / kTile, M / kTile, splits, ep};
  return FDADI_OK;
}

int gemm(int epi, const bf16 *A, const bf16 *B, int M, int N, int K, const EpiArgs &ep,
         hipStream_t s, int splits = 1) {
  GemmArgs g;
  FDADI_RET(gemm_args(epi, A, B, M, N, K, ep, splits, &g));
  dim3 grid(g.gx, g.gy, g.gz);
  switch (epi) {
    case EPI_FWD: hipLaunchKernelGGL(gemm_nt_kernel<EPI_FWD>, grid, dim3(kThreads), 0, s, g); break;
    case EPI_INF: hipLaunchKernelGGL(gemm_nt_kernel<EPI_INF>, grid, dim3(kThreads), 0, s, g); break;
    case EPI_BWD: hipLaunchKernelGGL(gemm_nt_kernel<EPI_BWD>, grid, dim3(kThreads), 0, s, g); break;
    default: hipLaunchKernelGGL(gemm_nt_kernel<EPI_F32>, grid, dim3(kThreads), 0, s, g); break;
  }
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

// A weight-gradient GEMM (fp32, split-K) and a backward-data GEMM in one
// launch.
int gemm_pair_f32_bwd(const GemmArgs &w, const GemmArgs &b, hipStream_t s) {
  const int n1 = w.gx * w.gy * w.gz, n2 = b.gx * b.gy * b.gz;
  FDADI_REQUIRE(n1 % 8 == 0, FDADI_ERR_ARG, "mlp gemm pair: first grid not a multiple of 8");
  hipLaunchKernelGGL((gemm_pair_kernel<EPI_F32, EPI_BWD>), dim3(n1 + n2), dim3(kThreads), 0, s, w,
                     b, n1);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int refresh_bf16(fdadi_mlp *m, hipStream_t s) {
  for (int l = 0; l < kLayers; ++l) {
    const int n = kH * in_dim(l);
    hipLaunchKernelGGL(cast_weights_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       m->p + m->off_W[l], kH, in_dim(l), m->Wb[l], m->WbT[l]);
  }
  hipLaunchKernelGGL(cast_weights_kernel, dim3(grid_for(kH, 256)), dim3(256), 0, s,
                     m->p + m->off_w4, 1, kH, m->w4b, (bf16 *)nullptr);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

// Forward over rows [0, Mpad) of the prepared input m->X.  Training keeps
// the transposed activations and SiLU derivatives for the backward pass and
// leaves the head's output as per-column-tile partial sums (m->ypart, summed
// by the head's backward); inference computes the head into m->y.
int forward(fdadi_mlp *m, int Mpad, bool train, hipStream_t s) {
  const bf16 *in = m->X;
  for (int l = 0; l < kLayers; ++l) {
    EpiArgs ep{};
    ep.bias = m->p + m->off_b[l];
    ep.dT = train ? m->dT[l] : nullptr;
    ep.out = m->Act[l];
    ep.outT = train ? m->ActT[l] : nullptr;
    if (train && l == kLayers - 1) {
      // the head's partial outputs come from this layer's epilogue
      ep.w4 = m->w4b;
      ep.ypart = m->ypart;
    }
    FDADI_RET(gemm(train ? EPI_FWD : EPI_INF, in, m->Wb[l], Mpad, kH, in_dim(l), ep, s));
    in = m->Act[l];
  }
  if (!train) {
    hipLaunchKernelGGL(head_forward_kernel, dim3((Mpad + 3) / 4), dim3(256), 0, s,
                       m->Act[kLayers - 1], m->w4b, m->p + m->off_b4, Mpad, m->y);
    FDADI_CHECK_LAUNCH();
  }
  return FDADI_OK;
}

// Inference forward pass into caller-given buffers: input X [Mpad x 128],
// hidden activations alternating between A0 and A1 [Mpad x 512], head y [Mpad].
int forward_inference(const fdadi_mlp *m, const bf16 *X, bf16 *A0, bf16 *A1, float *y, int Mpad,
                      hipStream_t s) {
  const bf16 *in = X;
  bf16 *bufs[2] = {A0, A1};
  for (int l = 0; l < kLayers; ++l) {
    EpiArgs ep{};
    ep.bias = m->p + m->off_b[l];
    ep.out = bufs[l & 1];
    FDADI_RET(gemm(EPI_INF, in, m->Wb[l], Mpad, kH, in_dim(l), ep, s));
    in = bufs[l & 1];
  }
  hipLaunchKernelGGL(head_forward_kernel, dim3((Mpad + 3) / 4), dim3(256), 0, s, in, m->w4b,
                     m->p + m->off_b4, Mpad, y);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

// ---- allocation-free inference ---------------------------------------------------
// fdadi_mlp_predict_noalloc keeps its activations in this file-scope device
// array instead of the model's buffers: X [kNoAllocRows x 128] bf16, two
// hidden-activation buffers [kNoAllocRows x 512] bf16, the head output
// [kNoAllocRows] fp32 -- 9 MiB + 16 KiB on every GPU the library runs on,
// held until the process ends.  In host code `noalloc_scratch` names a
// host-side stand-in: kernels get the device address resolved per device by
// hipGetSymbolAddress.
constexpr int kNoAllocRows = 4096;
constexpr size_t kNoAllocX = (size_t)kNoAllocRows * kInPad * sizeof(bf16);
constexpr size_t kNoAllocA = (size_t)kNoAllocRows * kH * sizeof(bf16);
constexpr size_t kNoAllocY = (size_t)kNoAllocRows * sizeof(float);
constexpr size_t kNoAllocBytes = kNoAllocX + 2 * kNoAllocA + kNoAllocY;
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
unsigned char *g_noalloc_ptr[kMaxDevices] = {};

// Device address of noalloc_scratch on the current device, resolved once per
// device with that device current.
int noalloc_scratch_ptr(unsigned char **out) {
  int dev = 0;
  FDADI_TRY_HIP(hipGetDevice(&dev));
  FDADI_REQUIRE(dev >= 0 && dev < kMaxDevices, FDADI_ERR_UNSUPPORTED,
                "fdadi_mlp_predict_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    FDADI_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    FDADI_REQUIRE(p != nullptr, FDADI_ERR_HIP,
                  "fdadi_mlp_predict_noalloc: no device address for the scratch array");
    g_noalloc_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_noalloc_ptr[dev];
  return FDADI_OK;
}

}  // namespace
}  // namespace fdadi

extern "C" {

int fdadi_mlp_predict_noalloc(const fdadi_mlp *m, const double *d_P, int64_t n, double *d_out,
                              void *stream) {
  FDADI_REQUIRE(m && n >= 0 && (n == 0 || (d_P && d_out)), FDADI_ERR_ARG,
                "fdadi_mlp_predict_noalloc: bad arguments");
  FDADI_REQUIRE(n <= kNoAllocRows, FDADI_ERR_ARG,
                "fdadi_mlp_predict_noalloc: " + std::to_string(n) +
                    " rows exceed the maximum of " + std::to_string(kNoAllocRows) +
                    " per call (split the batch, or use fdadi_mlp_predict)");
  if (n == 0) return FDADI_OK;
  unsigned char *base = nullptr;
  FDADI_RET(noalloc_scratch_ptr(&base));
  bf16 *X = reinterpret_cast<bf16 *>(base);
  bf16 *A0 = reinterpret_cast<bf16 *>(base + kNoAllocX);
  bf16 *A1 = reinterpret_cast<bf16 *>(base + kNoAllocX + kNoAllocA);
  float *y = reinterpret_cast<float *>(base + kNoAllocX + 2 * kNoAllocA);
  hipStream_t s = as_stream(stream);
  const int rows = (int)n;
  const int Mpad = (rows + kTile - 1) / kTile * kTile;
  // every row of X [0, Mpad) is written (padding rows as zeros), every
  // activation and head row of [0, Mpad) by the GEMMs: nothing carries over
  // from the previous call
  hipLaunchKernelGGL(prep_input_kernel, dim3(grid_for(prep_threads(Mpad, m->n_in, false), 256)),
                     dim3(256), 0, s, d_P, (const int *)nullptr, rows, m->n_in, m->lohi, m->lohi + kMaxIn, X,
                     (bf16 *)nullptr, Mpad, (float *)nullptr);
  FDADI_CHECK_LAUNCH();
  FDADI_RET(forward_inference(m, X, A0, A1, y, Mpad, s));
  hipLaunchKernelGGL(unscale_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, s, y, rows,
                     m->tmean, m->tstd, d_out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_mlp_create(int32_t max_rows, const float *init_params, int64_t nparam,
                     const double *lo, const double *hi, double tmean, double tstd,
                     fdadi_mlp **out) {
  return fdadi_mlp_create_features(kIn, max_rows, init_params, nparam, lo, hi, tmean, tstd, out);
}

int fdadi_mlp_create_features(int32_t n_in, int32_t max_rows, const float *init_params,
                              int64_t nparam, const double *lo, const double *hi, double tmean,
                              double tstd, fdadi_mlp **out) {
  FDADI_REQUIRE(out && init_params && lo && hi, FDADI_ERR_ARG, "fdadi_mlp_create: null argument");
  FDADI_REQUIRE(n_in >= 1 && n_in <= kMaxIn, FDADI_ERR_ARG,
                "fdadi_mlp_create: 1 to 16 input features");
  FDADI_REQUIRE(max_rows > 0 && max_rows % kTile == 0, FDADI_ERR_ARG,
                "fdadi_mlp_create: max_rows must be a positive multiple of 128");
  fdadi_mlp *m = new fdadi_mlp();
  m->max_rows = max_rows;
  m->n_in = n_in;
  int64_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    m->off_W[l] = o;
    o += (int64_t)kH * in_dim(l);
    m->off_b[l] = o;
    o += kH;
  }
  m->off_w4 = o;
  o += kH;
  m->off_b4 = o;
  o += 1;
  m->nparam = o;
  if (nparam != o) {
    delete m;
    set_error("fdadi_mlp_create: expected " + std::to_string(o) + " parameters, got " +
              std::to_string(nparam));
    return FDADI_ERR_ARG;
  }
  m->tmean = tmean;
  m->tstd = tstd;
  hipError_t e = hipSuccess;
  auto al = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  const size_t R = (size_t)max_rows;
  al((void **)&m->p, sizeof(float) * o);
  al((void **)&m->g, sizeof(float) * o);
  al((void **)&m->m, sizeof(float) * o);
  al((void **)&m->v, sizeof(float) * o);
  for (int l = 0; l < kLayers; ++l) {
    al((void **)&m->Wb[l], sizeof(bf16) * kH * in_dim(l));
    al((void **)&m->WbT[l], sizeof(bf16) * kH * in_dim(l));
    al((void **)&m->Act[l], sizeof(bf16) * R * kH);
    al((void **)&m->ActT[l], sizeof(bf16) * R * kH);
    al((void **)&m->dT[l], sizeof(bf16) * R * kH);
  }
  al((void **)&m->w4b, sizeof(bf16) * kH);
  al((void **)&m->X, sizeof(bf16) * R * kInPad);
  al((void **)&m->XT, sizeof(bf16) * R * kInPad);
  al((void **)&m->dZ, sizeof(bf16) * R * kH);
  al((void **)&m->dZT, sizeof(bf16) * R * kH);
  al((void **)&m->dZ2, sizeof(bf16) * R * kH);
  al((void **)&m->dZT2, sizeof(bf16) * R * kH);
  al((void **)&m->y, sizeof(float) * R);
  al((void **)&m->ypart, sizeof(float) * (kH / kTile) * R);
  al((void **)&m->dy, sizeof(float) * R);
  al((void **)&m->loss, sizeof(float));
  al((void **)&m->wpart, sizeof(float) * kMaxSplits * kH * kH);
  al((void **)&m->lohi, sizeof(double) * 2 * kMaxIn);
  if (e == hipSuccess) e = hipMemcpy(m->p, init_params, sizeof(float) * o, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->lohi, lo, sizeof(double) * n_in, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(m->lohi + kMaxIn, hi, sizeof(double) * n_in, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(o, 256)), dim3(256), 0, 0, m->m, o);
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(o, 256)), dim3(256), 0, 0, m->v, o);
    e = hipGetLastError();
  }
  if (e == hipSuccess && refresh_bf16(m, 0) != FDADI_OK) e = hipErrorUnknown;
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string("fdadi_mlp_create: ") + hipGetErrorString(e));
    // fall through to free what was allocated
    void *ptrs[] = {m->p, m->g, m->m, m->v, m->w4b, m->X, m->XT, m->dZ, m->dZT, m->dZ2,
                    m->dZT2, m->y, m->ypart, m->dy, m->loss, m->wpart, m->lohi};
    for (void *q : ptrs) (void)hipFree(q);
    for (int l = 0; l < kLayers; ++l) {
      (void)hipFree(m->Wb[l]); (void)hipFree(m->WbT[l]); (void)hipFree(m->Act[l]);
      (void)hipFree(m->ActT[l]); (void)hipFree(m->dT[l]);
    }
    delete m;
    return FDADI_ERR_HIP;
  }
  *out = m;
  return FDADI_OK;
}

int fdadi_mlp_destroy(fdadi_mlp *m) {
  if (!m) return FDADI_OK;
  void *ptrs[] = {m->p, m->g, m->m, m->v, m->w4b, m->X, m->XT, m->dZ, m->dZT, m->dZ2,
                  m->dZT2, m->y, m->ypart, m->dy, m->loss, m->wpart, m->lohi};
  for (void *q : ptrs) (void)hipFree(q);
  for (int l = 0; l < kLayers; ++l) {
    (void)hipFree(m->Wb[l]); (void)hipFree(m->WbT[l]); (void)hipFree(m->Act[l]);
    (void)hipFree(m->ActT[l]); (void)hipFree(m->dT[l]);
  }
  delete m;
  return FDADI_OK;
}

int fdadi_mlp_nparam(int64_t *n) {
  FDADI_REQUIRE(n, FDADI_ERR_ARG, "fdadi_mlp_nparam: null output");
  *n = (int64_t)kH * kInPad + kH + (kLayers - 1) * ((int64_t)kH * kH + kH) + kH + 1;
  return FDADI_OK;
}

int fdadi_mlp_get_params(const fdadi_mlp *m, float *params, float *grads) {
  FDADI_REQUIRE(m, FDADI_ERR_ARG, "fdadi_mlp_get_params: null model");
  FDADI_TRY_HIP(hipDeviceSynchronize());
  if (params) FDADI_TRY_HIP(hipMemcpy(params, m->p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  if (grads) FDADI_TRY_HIP(hipMemcpy(grads, m->g, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  return FDADI_OK;
}

int fdadi_mlp_set_params(fdadi_mlp *m, const float *params, void *stream) {
  FDADI_REQUIRE(m && params, FDADI_ERR_ARG, "fdadi_mlp_set_params: null argument");
  FDADI_TRY_HIP(hipMemcpy(m->p, params, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  FDADI_RET(refresh_bf16(m, as_stream(stream)));
  return FDADI_OK;
}

// Gradient of the mean squared error (targets standardised with tmean/tstd)
// for rows idx[0..rows) of (P [n x 8], T [n]); leaves it in the grad array.
static int grad_impl(fdadi_mlp *m, const double *d_P, const double *d_T, const int *d_idx,
                     int rows, hipStream_t s) {
  FDADI_REQUIRE(rows > 0 && rows <= m->max_rows, FDADI_ERR_ARG,
                "fdadi_mlp: batch rows exceed the model capacity");
  const int Mpad = (rows + kTile - 1) / kTile * kTile;
  hipLaunchKernelGGL(prep_input_kernel, dim3(grid_for(prep_threads(Mpad, m->n_in, true), 256)),
                     dim3(256), 0, s, d_P, d_idx, rows, m->n_in, m->lohi, m->lohi + kMaxIn, m->X, m->XT, Mpad,
                     m->loss);
  FDADI_CHECK_LAUNCH();
  FDADI_RET(forward(m, Mpad, true, s));
  hipLaunchKernelGGL(head_backward_kernel, dim3(kH / 64, Mpad / 64), dim3(256), 0, s, m->ypart,
                     m->p + m->off_b4, d_T, d_idx, rows, Mpad, m->w4b, m->dT[kLayers - 1], m->dy, m->dZ, m->dZT, m->loss,
                     m->tmean, m->tstd);
  hipLaunchKernelGGL(head_wgrad_kernel, dim3(kH + 1), dim3(256), 0, s, m->dy,
                     m->ActT[kLayers - 1], Mpad, m->g + m->off_w4, m->g + m->off_b4);
  FDADI_CHECK_LAUNCH();
  bf16 *dz = m->dZ, *dzt = m->dZT, *dz2 = m->dZ2, *dzt2 = m->dZT2;
  for (int l = kLayers - 1; l >= 0; --l) {
    // dW_l = dZ_l^T A_{l-1};  db_l = column sums of dZ_l
    const bf16 *aT = l == 0 ? m->XT : m->ActT[l - 1];
    // (kH x in_dim) output tiles are few; the long K = Mpad is split so that
    // the grid fills the GPU, and the partial products are summed in order
    int splits = 1;
    while (splits < kMaxSplits && Mpad % (kKBlock * splits * 2) == 0 &&
           Mpad / (splits * 2) >= 512)
      splits *= 2;
    EpiArgs ew{};
    ew.C = splits > 1 ? m->wpart : m->g + m->off_W[l];
    GemmArgs gw;
    FDADI_RET(gemm_args(EPI_F32, dzt, aT, kH, in_dim(l), Mpad, ew, splits, &gw));
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}), independent of dW_l: both in
      // one launch (the weight-gradient grid, 16 tiles x splits, is a
      // multiple of 8)
      EpiArgs eb{};
      eb.dT = m->dT[l - 1];
      eb.out = dz2;
      eb.outT = dzt2;
      GemmArgs gb;
      FDADI_RET(gemm_args(EPI_BWD, dz, m->WbT[l], Mpad, kH, kH, eb, 1, &gb));
      FDADI_RET(gemm_pair_f32_bwd(gw, gb, s));
    } else {
      FDADI_RET(gemm(EPI_F32, dzt, aT, kH, in_dim(l), Mpad, ew, s, splits));
    }
    const int64_t nw = (int64_t)kH * in_dim(l);
    const int nsum = splits > 1 ? (int)grid_for(nw / 4, 256) : 0;
    hipLaunchKernelGGL(layer_grad_kernel, dim3(nsum + kH), dim3(256), 0, s, m->wpart, splits, nw,
                       m->g + m->off_W[l], nsum, dzt, Mpad, m->g + m->off_b[l]);
    FDADI_CHECK_LAUNCH();
    if (l > 0) {
      std::swap(dz, dz2);
      std::swap(dzt, dzt2);
    }
  }
  return FDADI_OK;
}

int fdadi_mlp_grad(fdadi_mlp *m, const double *d_P, const double *d_T, const int *d_idx,
                   int32_t rows, float *loss, void *stream) {
  FDADI_REQUIRE(m && d_P && d_T, FDADI_ERR_ARG, "fdadi_mlp_grad: null argument");
  hipStream_t s = as_stream(stream);
  FDADI_RET(grad_impl(m, d_P, d_T, d_idx, rows, s));
  if (loss) {
    FDADI_TRY_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    FDADI_TRY_HIP(hipStreamSynchronize(s));
  }
  return FDADI_OK;
}

int fdadi_mlp_train_step(fdadi_mlp *m, const double *d_P, const double *d_T, const int *d_idx,
                         int32_t rows, float lr, void *stream) {
  FDADI_REQUIRE(m && d_P && d_T, FDADI_ERR_ARG, "fdadi_mlp_train_step: null argument");
  hipStream_t s = as_stream(stream);
  FDADI_RET(grad_impl(m, d_P, d_T, d_idx, rows, s));
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  CastMap cm;
  for (int l = 0; l < kLayers; ++l) {
    cm.off[l] = m->off_W[l];
    cm.lg_in[l] = in_dim(l) == kInPad ? 7 : 9;
    cm.Wb[l] = m->Wb[l];
    cm.WbT[l] = m->WbT[l];
  }
  cm.off[kLayers] = m->off_w4;
  cm.w4b = m->w4b;
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(m->nparam, 256)), dim3(256), 0, s, m->p, m->g,
                     m->m, m->v, m->nparam, lr, b1, b2, 1e-8f, c1, c2, cm);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_mlp_loss(const fdadi_mlp *m, float *loss) {
  FDADI_REQUIRE(m && loss, FDADI_ERR_ARG, "fdadi_mlp_loss: null argument");
  FDADI_TRY_HIP(hipDeviceSynchronize());
  FDADI_TRY_HIP(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  return FDADI_OK;
}

int fdadi_mlp_predict(fdadi_mlp *m, const double *d_P, int64_t n, double *d_out, void *stream) {
  FDADI_REQUIRE(m && d_P && d_out && n >= 0, FDADI_ERR_ARG, "fdadi_mlp_predict: bad arguments");
  hipStream_t s = as_stream(stream);
  for (int64_t r0 = 0; r0 < n; r0 += m->max_rows) {
    const int rows = (int)std::min<int64_t>(m->max_rows, n - r0);
    const int Mpad = (rows + kTile - 1) / kTile * kTile;
    hipLaunchKernelGGL(prep_input_kernel, dim3(grid_for(prep_threads(Mpad, m->n_in, false), 256)),
                       dim3(256), 0, s, d_P + r0 * m->n_in, (const int *)nullptr, rows,
                       m->n_in, m->lohi, m->lohi + kMaxIn, m->X, (bf16 *)nullptr, Mpad,
                       (float *)nullptr);
    FDADI_CHECK_LAUNCH();
    FDADI_RET(forward(m, Mpad, false, s));
    hipLaunchKernelGGL(unscale_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, s, m->y, rows,
                       m->tmean, m->tstd, d_out + r0);
    FDADI_CHECK_LAUNCH();
  }
  return FDADI_OK;
}

// Raw NT GEMM for tests: C (fp32, M x N) = A (bf16 bits, M x K) B^T (bf16 bits, N x K).
int fdadi_mlp_gemm_nt(const uint16_t *d_A, const uint16_t *d_B, int32_t M, int32_t N, int32_t K,
                      float *d_C, void *stream) {
  FDADI_REQUIRE(d_A && d_B && d_C, FDADI_ERR_ARG, "fdadi_mlp_gemm_nt: null argument");
  EpiArgs ep{};
  ep.C = d_C;
  return gemm(EPI_F32, reinterpret_cast<const bf16 *>(d_A), reinterpret_cast<const bf16 *>(d_B),
              M, N, K, ep, as_stream(stream));
}

}  // extern "C"
