// Fused, batched Heston ADI engine.
//
// Replaces the per-step operator-call loops of
// FiniteDifferenceEngineGPU.pyx:379-593 (and the CPU-only Craig-Sneyd loops
// of FiniteDifferenceEngine.pyx:778-898) with one launch per ADI stage for a
// whole batch of options with uneven grids.
//
// Layout.  Option k's state is an (nS, nv) row-major block at d_state+off[k]
// (variance contiguous, as the reference's Grid).  No 2-D operator is stored:
// every band of A0 (spot), A1 (variance) and A01 (cross) is generated in the
// kernels from 1-D tables built on the device from the meshes:
//   A0[i,j] band = P(i) + v_j Q(i)                 (mu_S d/dS + 1/2 v S^2 d2/dS2 - r/2)
//   A1[i,j] band = [S_i != 0] G(j) - r/2 on mid    (kappa(theta-v) d/dv + 1/2 sigma^2 v d2/dv2)
//   A01[i,j]     = rho sigma S_i v_j ws(i) x wv(j) (9-point)
// Boundary rows are those of _TriBandedOperatorGPU.cu:203-362: Dirichlet 0
// at S=0 (and at Smax for a barrier), dU/dS=1 at Smax, one-sided first
// derivative at both variance ends, d2/dv2 Von Neumann 0 at v=0 and the
// second-order backward row (with its -2 band entry, folded) at vmax.
//
// Sweeps.  A spot sweep is one workgroup per (option, tile of NL variance
// columns); a variance sweep one workgroup per (option, tile of NL spot
// rows).  The tile's right-hand side is formed (fusing every explicit term of
// the stage) straight into LDS, the NL tridiagonal lines are solved in LDS,
// and the stage output (fusing the next stage's explicit term) is written
// back: one read of each input vector and one write per stage, no
// transposes.  The line solve uses the time-invariant LU factorisation of
// I - theta dt A (built once per configure): each line is split into chunks
// owned by threads, and both triangular sweeps are first-order linear
// recurrences composed across chunks with a wave-level scan of affine maps.
#include "common.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace fdadi {
namespace {

enum SpotField {
  SP_PSUB, SP_PMID, SP_PSUP, SP_QSUB, SP_QMID, SP_QSUP, SP_PR, SP_QR,
  SP_WSUB, SP_WMID, SP_WSUP, SP_S, SP_DIR, SP_NF
};
enum VarField {
  VT_GSUB, VT_GMID, VT_GSUP, VT_WSUB, VT_WMID, VT_WSUP, VT_V, VT_M, VT_RD, VT_N, VT_NF
};

struct HOpt {
  int64_t nS, nv, off;  // grid and state offset
  int64_t so, vo;       // offsets into the spot / variance tables
  double r, kappa, theta, sigma, rho, dt;
  double tdt;           // weight of the implicit part (theta*dt, or dt for implicit)
  double th;            // the scheme's theta
  double fold_f;        // bottom fold factor of I - tdt A1 (lines with S != 0)
  double rdS0;          // 1 / (1 + tdt r/2): the S = 0 variance line
  double Gz;            // A1[nv-1, nv-3] for S != 0
  int32_t barrier, pad;
};

struct Tile {
  int32_t opt;
  int32_t start;  // first column (spot sweep) / first row (variance sweep)
};

enum StageMode {
  // spot sweeps
  SM_FIRST = 0,  // rhs = U + F U - Le0 U;               out = x - Le1 U
  SM_HV = 1,     // rhs = U + F U + 1/2 F^(Y2-U) - Le0 Y2; out = x - Le1 Y2
  SM_CS = 2,     // rhs = U + F U + dt/2 A01 (Y2-U) - Le0 U; out = x - Le1 U
  SM_MCS = 3,    // rhs = U + F U + th dt A01 D + (1/2-th) F D - Le0 U; out = x - Le1 U
  SM_IMPL = 4,   // rhs = W + dt R0;                     out = x
  // variance sweeps
  VM_PLAIN = 10, // rhs = W;                             out = x
  VM_IMPL = 11   // rhs = U + dt A01 U;                  out = x
};

struct Ctx {
  const HOpt *opts;
  const double *sp;  // spot tables, field-major, stride sp_n
  const double *vt;  // variance tables, field-major, stride vt_n
  int64_t sp_n, vt_n;
  const double *rd0;  // spot-line pivots 1/u, state layout
  const double *spr;  // spot table, one 16-double row per node
};

__device__ __forceinline__ double SP(const Ctx &c, const HOpt &o, int f, int64_t i) {
  return c.sp[f * c.sp_n + o.so + i];
}
__device__ __forceinline__ double VT(const Ctx &c, const HOpt &o, int f, int64_t j) {
  return c.vt[f * c.vt_n + o.vo + j];
}

// A0 band entry p + w2 * c00 as the reference's combined operator holds it:
// the (0,) band (with -r/2 on the diagonal) plus the (0,0) weight scaled by
// c00 = (1/2 v) S^2, one rounding each (no fused multiply-add), so that the
// discrete operator is bit-identical to the reference's.
__device__ __forceinline__ double band_add(double p, double w2, double c00) {
#pragma clang fp contract(off)
  const double t = w2 * c00;
  return p + t;
}

__device__ __forceinline__ double c00_of(double v, double S) {
#pragma clang fp contract(off)
  return (0.5 * v) * (S * S);
}

// (A01 x)(i,j): rho sigma S v times the 9-point central product stencil.
template <class F>
__device__ double a01_dot(const Ctx &c, const HOpt &o, int64_t i, int64_t j, F x) {
  if (o.rho == 0.0 || i == 0 || i == o.nS - 1 || j == 0 || j == o.nv - 1) return 0.0;
  const double ws[3] = {SP(c, o, SP_WSUB, i), SP(c, o, SP_WMID, i), SP(c, o, SP_WSUP, i)};
  const double wv[3] = {VT(c, o, VT_WSUB, j), VT(c, o, VT_WMID, j), VT(c, o, VT_WSUP, j)};
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double row = wv[0] * x(i + a - 1, j - 1) + wv[1] * x(i + a - 1, j) + wv[2] * x(i + a - 1, j + 1);
    acc += ws[a] * row;
  }
  const double coef = ((o.rho * o.sigma) * SP(c, o, SP_S, i)) * VT(c, o, VT_V, j);
  return coef * acc;
}

// Coefficients of the tridiagonal I - tdt*A0 on spot line j at row i.
__device__ __forceinline__ void li0_row(const Ctx &c, const HOpt &o, int64_t i, double v,
                                        double &a, double &b, double &cc) {
  if (SP(c, o, SP_DIR, i) != 0.0) {
    a = 0.0; b = 1.0; cc = 0.0;
    return;
  }
  const double c00 = c00_of(v, SP(c, o, SP_S, i));
  a = -o.tdt * band_add(SP(c, o, SP_PSUB, i), SP(c, o, SP_QSUB, i), c00);
  b = 1.0 - o.tdt * band_add(SP(c, o, SP_PMID, i), SP(c, o, SP_QMID, i), c00);
  cc = -o.tdt * band_add(SP(c, o, SP_PSUP, i), SP(c, o, SP_QSUP, i), c00);
  if (i == 0) a = 0.0;
  if (i == o.nS - 1) cc = 0.0;
}

// ---- wave-level scans of affine maps ------------------------------------------
// A map y -> E + M y is the pair (M, E); composing "earlier then later" gives
// (M_l M_e, E_l + M_l E_e).  The scans run on DPP lane moves (no LDS round
// trip): row_shr/row_shl within rows of 16 lanes, row_bcast:15/31 across rows
// for prefixes, wave_shr/wave_shl for the exclusive shift.  Lanes with no
// source keep the identity (M, E) = (1, 0).  Every lane of the wave must be
// active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_or(double x, double old) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(x), CTRL, ROWS,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(x), CTRL, ROWS,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}

// One prefix level: combine with the map of lane ln - D (ln = lane within a
// segment of W lanes); the DPP control already fetches from lane - D.
template <int CTRL, int ROWS, int D, int W>
__device__ __forceinline__ void prefix_level(double &M, double &E, int ln) {
  double Mp = dpp_or<CTRL, ROWS>(M, 1.0), Ep = dpp_or<CTRL, ROWS>(E, 0.0);
  if (W < 16 && ln < D) {
    Mp = 1.0;
    Ep = 0.0;
  }
  E = E + M * Ep;
  M = M * Mp;
}

// Inclusive prefix scan over segments of W lanes (W a power of two <= 64).
template <int W>
__device__ __forceinline__ void wave_prefix_scan(double &M, double &E, int ln) {
  if (W > 1) prefix_level<0x111, 0xF, 1, W>(M, E, ln);
  if (W > 2) prefix_level<0x112, 0xF, 2, W>(M, E, ln);
  if (W > 4) prefix_level<0x114, 0xF, 4, W>(M, E, ln);
  if (W > 8) prefix_level<0x118, 0xF, 8, W>(M, E, ln);
  if (W > 16) prefix_level<0x142, 0xA, 16, W>(M, E, ln);  // row_bcast:15 into rows 1, 3
  if (W > 32) prefix_level<0x143, 0xC, 32, W>(M, E, ln);  // row_bcast:31 into rows 2, 3
}

// Inclusive suffix scan over segments of W lanes: maps of higher lanes are
// applied first.  Within rows by DPP row_shl (after those levels a lane holds
// the maps from itself to the end of its row).  Across rows, mirroring
// row_bcast of the prefix scan: at D = 16 rows 0 and 2 take the first lane of
// the next row (which holds that whole row), at D = 32 rows 0 and 1 take lane
// 32 (which then holds rows 2-3); by ds_bpermute.
template <int CTRL, int D, int W>
__device__ __forceinline__ void suffix_level(double &M, double &E, int ln) {
  double Mn, En;
  bool take;
  if (D < 16) {
    Mn = dpp_or<CTRL, 0xF>(M, 1.0);
    En = dpp_or<CTRL, 0xF>(E, 0.0);
    take = ln + D < W;
  } else {
    const int lane = threadIdx.x % 64;
    const int src = D == 16 ? (lane | 15) + 1 : 32;
    Mn = __shfl(M, src & 63, 64);
    En = __shfl(E, src & 63, 64);
    take = D == 16 ? ((lane >> 4) & 1) == 0 : lane < 32;
  }
  if (!take) {
    Mn = 1.0;
    En = 0.0;
  }
  E = E + M * En;
  M = M * Mn;
}

template <int W>
__device__ __forceinline__ void wave_suffix_scan(double &M, double &E, int ln) {
  if (W > 1) suffix_level<0x101, 1, W>(M, E, ln);
  if (W > 2) suffix_level<0x102, 2, W>(M, E, ln);
  if (W > 4) suffix_level<0x104, 4, W>(M, E, ln);
  if (W > 8) suffix_level<0x108, 8, W>(M, E, ln);
  if (W > 16) suffix_level<0, 16, W>(M, E, ln);
  if (W > 32) suffix_level<0, 32, W>(M, E, ln);
}

// Value of lane ln - 1 of the segment (0 for its first lane).
template <int W>
__device__ __forceinline__ double from_prev_lane(double x, int ln) {
  const double v = dpp_or<0x138, 0xF>(x, 0.0);  // wave_shr:1
  return ln == 0 ? 0.0 : v;
}

// Value of lane ln + 1 of the segment (0 for its last lane).
template <int W>
__device__ __forceinline__ double from_next_lane(double x, int ln) {
  const double v = dpp_or<0x130, 0xF>(x, 0.0);  // wave_shl:1
  return ln == W - 1 ? 0.0 : v;
}

// Exclusive scan of affine maps y -> E + M y along the TPL chunks of each of
// the NL lines of a tile.  Entries live in LDS at [l*(TPL+1) + c] (line by
// line, odd pitch: the lanes of a line read consecutive entries and the
// writers of one chunk hit distinct banks).  Forward
// (REV=false): carry(0)=0, carry(c+1)=E[c]+M[c]*carry(c).  Backward
// (REV=true): the same from the last chunk down.  A line takes W = min(TPL,
// 64) lanes (K = TPL / W entries per lane), so a wave scans 64 / W lines at
// once with segmented scans.
template <int THREADS, int NL, bool REV>
__device__ void affine_scan(const double *cM, const double *cE, double *cIn) {
  constexpr int TPL = THREADS / NL;
  constexpr int TPLP = TPL + 1;
  constexpr int W = TPL < 64 ? TPL : 64;
  constexpr int K = TPL / W;
  constexpr int LPW = 64 / W;
  constexpr int NW = THREADS / 64;
  static_assert(W >= 16 && (W & (W - 1)) == 0, "16 to 64 lanes per line");
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int ln = lane % W;
  for (int l0 = wave * LPW; l0 < NL; l0 += NW * LPW) {
    const int l = l0 + lane / W;
    const bool ok = l < NL;
    // Compose this lane's K entries (in scan order).
    double Mloc = 1.0, Eloc = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int s = ln * K + q;
      const int cidx = REV ? TPL - 1 - s : s;
      const double m = ok ? cM[l * TPLP + cidx] : 1.0, e = ok ? cE[l * TPLP + cidx] : 0.0;
      Eloc = e + m * Eloc;
      Mloc = m * Mloc;
    }
    // Inclusive segmented scan of (M, E) (later maps applied after earlier ones).
    double M = Mloc, E = Eloc;
    wave_prefix_scan<W>(M, E, ln);
    // Exclusive: the carry entering this lane's first entry.
    double carry = from_prev_lane<W>(E, ln);
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int s = ln * K + q;
      const int cidx = REV ? TPL - 1 - s : s;
      if (ok) {
        cIn[l * TPLP + cidx] = carry;
        carry = cE[l * TPLP + cidx] + cM[l * TPLP + cidx] * carry;
      }
    }
  }
}

// ---- spot sweep ---------------------------------------------------------------
// One 128-byte row of per-node spot coefficients, built from the spot table
// with the boundary handling folded in:
// [psub' pmid psup' qsub' | qmid qsup' pr qr | S*wsub S*wmid S*wsup S^2 | S snz 0 0]
// (p*: first-derivative bands scaled by r S, with -r/2 on the diagonal; q*:
// raw second-derivative weights, and qr the raw residual, to be scaled by
// c00 = (1/2 v) S^2 in the kernel) where ' marks band entries zeroed when
// they multiply a Dirichlet neighbour (row 0 for row 1; row nS-1 for row nS-2
// under a barrier) -- the explicit apply forces those neighbours to 0, and in
// the solve the coupling to a Dirichlet row (whose value is 0) contributes
// nothing -- and snz = [S != 0].  Dirichlet rows have all-zero bands.  A sweep
// reads the first twelve fields of a row; S and snz only at row 0 (the spot
// mesh increases strictly from S >= 0, so only row 0 can have S = 0).
constexpr int kSpRow = 16;
constexpr int kSpFields = 12;

struct SpotRow {
  double psub, pmid, psup, qsub, qmid, qsup, pr, qr, wsub, wmid, wsup, S2;
};

// A1 on one variance column scaled by a step weight c (c = dt for the
// explicit operator dt A1, c = theta dt for theta dt A1): the entries the
// reference's scaled operator holds, fl(c * a), for lines with S != 0 (sub,
// mid, sup and, on the last column, z = the j-2 entry of the second-order
// backward row) and the diagonal c * (-r/2) of the S = 0 line.
struct ColOp {
  double sub, mid, sup, z, s0;
};

struct VarCol {  // variance coefficients of column j (constant along a spot line)
  double wv0, wv1, wv2;
  double c0h;     // v / 2 (c00 = c0h S^2)
  double rsv;     // rho * sigma * v
  ColOp f1, t1;   // dt A1 and theta dt A1 on this column
};

// A0 bands of row i at variance v (Dirichlet handling is in the bands).
struct Bands {
  double sub, mid, sup, r0;
};

// A0 band entry p + w2 * c00 with c00 = (1/2 v) S^2, one rounding each.
__device__ __forceinline__ Bands a0_bands(const SpotRow &R, double c0h) {
#pragma clang fp contract(off)
  const double c00 = c0h * R.S2;
  Bands b;
  b.sub = band_add(R.psub, R.qsub, c00);
  b.mid = band_add(R.pmid, R.qmid, c00);
  b.sup = band_add(R.psup, R.qsup, c00);
  b.r0 = band_add(R.pr, R.qr, c00);  // residual: zero except on the Smax row
  return b;
}

__device__ __forceinline__ ColOp col_op(double c, double gsub, double gmid, double gsup,
                                        double gz, double split) {
  return ColOp{c * gsub, c * gmid, c * gsup, c * gz, c * split};
}

// A state row at this lane's column j: columns j-1, j, j+1 (clamped to the
// grid) and j-2 on the grid's last column (0 elsewhere).
struct Row4 {
  double l, c, r, m2;
};

struct LaneCols {  // this lane's column indices within a row (clamped)
  unsigned jm, jc, jp, jm2, jx;
  bool last;
};

__device__ __forceinline__ Row4 load_row4(const double *__restrict__ f, unsigned rowoff,
                                          const LaneCols &lc) {
  const double *p = f + rowoff;
  Row4 w;
  w.l = p[lc.jm];
  w.c = p[lc.jc];
  w.r = p[lc.jp];
  w.m2 = lc.last ? p[lc.jm2] : 0.0;
  return w;
}

// A row as loaded for the window: column j and this lane's edge column jx
// (j - 1 on the tile's first lane, j + 1 on its last, j elsewhere).  The
// inner neighbours come from the adjacent lanes, which hold exactly those
// columns (a lane past the grid's last column loads that column, the clamped
// j + 1 of the last one); column j - 2 of the grid's last column comes from
// the lane two before it (the tile list puts the last column on a tile's last
// lane).
struct Raw2 {
  double e, c;
};

__device__ __forceinline__ Raw2 load_raw2(const double *__restrict__ f, unsigned rowoff,
                                          const LaneCols &lc) {
  const double *p = f + rowoff;
  return Raw2{p[lc.jx], p[lc.jc]};
}

// A double at a 32-bit byte offset from a work-group-uniform base: the global
// access takes the scalar base + 32-bit lane offset form, with no 64-bit
// address arithmetic per lane.
__device__ __forceinline__ double load_at(const double *__restrict__ base, unsigned byte_off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ void store_at(double *__restrict__ base, unsigned byte_off, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

// load_raw2 with the row given as a byte offset (row * nv * 8).
__device__ __forceinline__ Raw2 load_raw2_at(const double *__restrict__ f, unsigned rowb,
                                             const LaneCols &lc) {
  return Raw2{load_at(f, rowb + lc.jx * 8u), load_at(f, rowb + lc.jc * 8u)};
}

// Value of the lane before / after / two before within the tile's lanes
// (quad_perm for four columns; row_shr / row_shl within a DPP row of 16
// lanes for 8 and 16; wave_shr / wave_shl across the two DPP rows of a
// 32-column tile; the tile's first and last lanes use their edge load).  Two
// before serves the last column only, which sits on the tile's last lane
// (lane 31 of a 32-column tile, nv >= 32), so row_shr:2 stays in its row.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int NL>
__device__ __forceinline__ double from_prev_col(double x) {
  // quad_perm [0,0,1,2] / row_shr:1 / wave_shr:1
  return dpp_mov<NL == 4 ? 0x90 : (NL == 32 ? 0x138 : 0x111)>(x);
}
template <int NL>
__device__ __forceinline__ double from_next_col(double x) {
  // quad_perm [1,2,3,3] / row_shl:1 / wave_shl:1
  return dpp_mov<NL == 4 ? 0xF9 : (NL == 32 ? 0x130 : 0x101)>(x);
}
template <int NL>
__device__ __forceinline__ double from_prev2_col(double x) {
  return dpp_mov<NL == 4 ? 0x40 : 0x112>(x);  // quad_perm [0,0,0,1] / row_shr:2
}

template <int NL>
__device__ __forceinline__ Row4 expand_raw(const Raw2 &w, int jj, bool last) {
  const double m2 = from_prev2_col<NL>(w.c);
  if constexpr (NL == 16) {
    // a DPP row of 16 lanes is the tile: its first and last lanes, which have
    // no source lane for row_shr:1 / row_shl:1, keep the edge load
    return Row4{dpp_or<0x111, 0xF>(w.c, w.e), w.c, dpp_or<0x101, 0xF>(w.c, w.e), last ? m2 : 0.0};
  } else {
    const double l = from_prev_col<NL>(w.c), r = from_next_col<NL>(w.c);
    return Row4{jj == 0 ? w.e : l, w.c, jj == NL - 1 ? w.e : r, last ? m2 : 0.0};
  }
}
template <int NL>
__device__ __forceinline__ Row4 expand_raw(const Row4 &w, int, bool) {
  return w;
}

// Rows in flight: the edge-and-centre form (neighbours by DPP) or all three
// columns loaded.
template <bool NB_DPP>
struct RawRow {
  using type = Raw2;
  static __device__ __forceinline__ Raw2 load(const double *__restrict__ f, unsigned rowoff,
                                              const LaneCols &lc) {
    return load_raw2(f, rowoff, lc);
  }
  static __device__ __forceinline__ Raw2 load_b(const double *__restrict__ f, unsigned rowb,
                                                const LaneCols &lc) {
    return load_raw2_at(f, rowb, lc);
  }
  static __device__ __forceinline__ Raw2 zero() { return Raw2{0.0, 0.0}; }
};
template <>
struct RawRow<false> {
  using type = Row4;
  static __device__ __forceinline__ Row4 load(const double *__restrict__ f, unsigned rowoff,
                                              const LaneCols &lc) {
    return load_row4(f, rowoff, lc);
  }
  static __device__ __forceinline__ Row4 load_b(const double *__restrict__ f, unsigned rowb,
                                                const LaneCols &lc) {
    return load_row4(f, rowb / 8u, lc);
  }
  static __device__ __forceinline__ Row4 zero() { return Row4{0.0, 0.0, 0.0, 0.0}; }
};

// (c A1 f)(i,j) with the scaled entries K of this column over row i of f (K.z
// is zero except on the last column).
__device__ __forceinline__ double a1s(const ColOp &K, const Row4 &w) {
  return ((K.sup * w.r + K.mid * w.c) + K.sub * w.l) + K.z * w.m2;
}

// What the sliding window keeps of a row once it has entered: the value at
// column j and the row's share of the cross product, wv . (x_{j-1}, x_j,
// x_{j+1}), of U and of D = Y - U, Y at column j, and the A1 products the
// stage takes at this row (dt A1 U, theta dt A1 U, dt A1 D, theta dt A1 Y).
// Each is formed once per row, when the row enters.
struct WinRow {
  double uc, urs, dc, drs, yc;
  double f1u, t1u, f1d, t1y;
};

template <int MODE>
__device__ __forceinline__ void enter_row(const VarCol &q, const Row4 &u, const Row4 &y,
                                          WinRow &w) {
  constexpr bool HAS_Y = MODE == SM_HV || MODE == SM_CS || MODE == SM_MCS;
  w.uc = u.c;
  w.urs = q.wv0 * u.l + q.wv1 * u.c + q.wv2 * u.r;
  if constexpr (MODE != SM_IMPL) w.f1u = a1s(q.f1, u);
  if constexpr (MODE == SM_FIRST || MODE == SM_CS || MODE == SM_MCS) w.t1u = a1s(q.t1, u);
  if constexpr (HAS_Y) {
    Row4 d;
    d.l = y.l - u.l;
    d.c = y.c - u.c;
    d.r = y.r - u.r;
    d.m2 = y.m2 - u.m2;
    w.dc = d.c;
    w.drs = q.wv0 * d.l + q.wv1 * d.c + q.wv2 * d.r;
    w.yc = y.c;
    if constexpr (MODE == SM_HV || MODE == SM_MCS) w.f1d = a1s(q.f1, d);
    if constexpr (MODE == SM_HV) w.t1y = a1s(q.t1, y);
  }
}

// The S = 0 line (row 0 when the spot mesh starts at 0): A1 keeps only the
// -r/2 of the split, c * (-r/2) * f(i, j).
template <int MODE>
__device__ __forceinline__ void zero_line(const VarCol &q, WinRow &w) {
  if constexpr (MODE != SM_IMPL) w.f1u = q.f1.s0 * w.uc;
  if constexpr (MODE == SM_FIRST || MODE == SM_CS || MODE == SM_MCS) w.t1u = q.t1.s0 * w.uc;
  if constexpr (MODE == SM_HV || MODE == SM_MCS) w.f1d = q.f1.s0 * w.dc;
  if constexpr (MODE == SM_HV) w.t1y = q.t1.s0 * w.yc;
}

// Work-group index -> tile index so that consecutive tiles (which share cache
// lines of the state) run on the same XCD: dispatch places work-group k on
// XCD k mod 8.
__device__ __forceinline__ int xcd_tile(int k, int G) {
  const int q = G / 8, rem = G % 8, x = k % 8, y = k / 8;
  return x * q + (x < rem ? x : rem) + y;
}

// Broadcast lane Q of each group of four lanes (DPP quad_perm [Q,Q,Q,Q]).
template <int Q>
__device__ __forceinline__ double quad_bcast(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), Q * 0x55, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), Q * 0x55, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Four-column tiles: the quad's lanes share a table row; lane jj of the quad
// loads doubles [4jj, 4jj+4) (two 16-byte loads) and the fields are broadcast
// (every lane of the quad must execute quad_row_fields).
__device__ __forceinline__ void quad_row_load(const double *__restrict__ spr, int jj, double2 &a,
                                              double2 &b) {
  const double2 *p = reinterpret_cast<const double2 *>(spr + 4 * jj);
  a = p[0];
  b = p[1];
}

__device__ __forceinline__ SpotRow quad_row_fields(double2 a, double2 b) {
  SpotRow r;
  r.psub = quad_bcast<0>(a.x); r.pmid = quad_bcast<0>(a.y);
  r.psup = quad_bcast<0>(b.x); r.qsub = quad_bcast<0>(b.y);
  r.qmid = quad_bcast<1>(a.x); r.qsup = quad_bcast<1>(a.y);
  r.pr = quad_bcast<1>(b.x);   r.qr = quad_bcast<1>(b.y);
  r.wsub = quad_bcast<2>(a.x); r.wmid = quad_bcast<2>(a.y);
  r.wsup = quad_bcast<2>(b.x); r.S2 = quad_bcast<2>(b.y);
  return r;
}

// Wider tiles (8 or 16 columns): the work-group stages the table rows its
// chunks read into LDS, chunk by chunk (MAXL rows of 12 fields, 2 doubles of
// padding per chunk so that the chunks of a wave start on different banks),
// and each lane reads its row with six 16-byte LDS loads (the lanes of a
// chunk read the same address).
__device__ __forceinline__ SpotRow lds_row_fields(const double *__restrict__ p) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
  const double2 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
  return SpotRow{a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y, e.x, e.y, f.x, f.y};
}

// An empty asm that reads x: x has to be computed before this point.  The
// row walk keeps its results in registers (no store orders it), so without
// these fences the compiler hoists every row's loads to the top of the
// unrolled walk and runs out of registers.
__device__ __forceinline__ void pin(double x) { asm volatile("" ::"v"(x)); }

// A spot sweep: one work-group per (option, NL adjacent variance columns,
// NL = 4, 8 or 16); thread (jj, ch) = (tid % NL, tid / NL) owns rows
// [ch*L, ch*L+L) of column j0+jj through every phase.  The NL lanes of a
// chunk work on the same rows, so they share the spot-table rows.  Four
// columns per tile keep a large grid spread over every CU (2048 x 1024: 256
// tiles); wider tiles make each load of a wave touch fewer, longer row
// segments, which pays where the batch has tiles to spare.
//
// The row walk keeps a window of three rows (i-1, i, i+1) of what the stage
// needs from U and Y2, the loads of rows i+2 and i+3 in flight, and forms the
// fused right-hand side (and the next stage's explicit term) of row i.  The
// state rows are loaded with clamped indices; rows past the chunk are
// computed and discarded.
//
// The LU solve is run in scaled form: with pivots rd = 1/u, multipliers
// mu_i = -a_i rd_i and n_i = -c_i rd_i, the forward sweep computes
// y'_i = rd_i d_i + mu_i y'_{i-1} (= rd_i y_i) and the backward sweep
// x_i = y'_i + n_i x_{i+1}, so the pivots are needed only while the
// right-hand side is formed.  The pivots rd_i = 1 / (b_i - (a_i rd_{i-1})
// c_{i-1}) are recomputed down the chunk from the stored pivot of the row
// before it (the recurrence the factorisation ran).  mu, n and y' (then x)
// stay in registers; LDS holds the next stage's explicit term (thread-private
// slots [k*THREADS + tid]), the chunk carries of the scans and, for wide
// tiles, the table rows.
//
// Rows k >= L of a chunk belong to the next chunk: the uniform guard k < L
// keeps them out of both chunk maps.  In the last chunk, rows past the grid
// (i >= nS) have finite values and follow row nS - 1, whose super-diagonal is
// zero (n = 0), so they do not reach any valid row either way.
template <int MODE, int THREADS, int NL, int MAXL>
__global__ void __launch_bounds__(THREADS)
spot_sweep_kernel(Ctx c, const Tile *__restrict__ tiles, const double *__restrict__ U,
                  const double *__restrict__ Y, double *__restrict__ out) {
  static_assert(NL == 4 || NL == 8 || NL == 16 || NL == 32,
                "spot tiles are 4, 8, 16 or 32 columns wide");
  extern __shared__ double lds[];
  constexpr int TPL = THREADS / NL;
  constexpr bool HAS_Y = MODE == SM_HV || MODE == SM_CS || MODE == SM_MCS;
  constexpr bool HAS_EX = MODE != SM_IMPL;
  constexpr bool TAB_LDS = NL != 4;
  constexpr int CHP = MAXL * kSpFields + 2;  // LDS table: doubles per chunk
  // state rows in flight (two for the second-stage modes on long chunks,
  // whose window and registers are largest)
  constexpr int PF = (MODE == SM_HV || MODE == SM_MCS) && MAXL == 16 ? 2 : 4;
  constexpr bool NB_DPP = true;              // neighbour columns by DPP
  double *cM = lds;
  double *cE = cM + NL * (TPL + 1);
  double *cIn = cE + NL * (TPL + 1);
  double *eb = cIn + NL * (TPL + 1);  // next stage's explicit term, slot [k * THREADS + tid]
  double *tab = eb + (HAS_EX ? MAXL * THREADS : 0);  // 16-byte aligned
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const HOpt o = c.opts[t.opt];
  const int nS = (int)o.nS, nv = (int)o.nv, j0 = t.start;
  const int tid = threadIdx.x;
  const int jj = tid % NL, ch = tid / NL;
  const int j = j0 + jj;
  const bool col_ok = j < nv;
  const int jc = col_ok ? j : nv - 1;  // clamped column for loads
  const int L = (nS + TPL - 1) / TPL;
  const int s = ch * L;
  const double *u = U + o.off;
  const double *y2 = HAS_Y ? Y + o.off : nullptr;
  const double *rd = c.rd0 + o.off;
  const double *spr = c.spr + o.so * kSpRow;
  const unsigned unv = (unsigned)nv;
  auto clamp_row = [&](int i) { return (unsigned)(i < 0 ? 0 : (i >= nS ? nS - 1 : i)); };

  if constexpr (TAB_LDS) {
    // table rows of every chunk: TPL * MAXL slots (rows past the grid zero)
#pragma unroll
    for (int slot = tid; slot < TPL * MAXL; slot += THREADS) {
      const int cc = slot / MAXL, k = slot % MAXL;
      const int r = cc * L + k;
      const double2 *src = reinterpret_cast<const double2 *>(spr + (int64_t)(r < nS ? r : 0) * kSpRow);
      double2 *dst = reinterpret_cast<double2 *>(tab + cc * CHP + k * kSpFields);
      const bool in = r < nS;
#pragma unroll
      for (int part = 0; part < 6; ++part) {
        const double2 v = src[part];
        dst[part] = in ? v : make_double2(0.0, 0.0);
      }
    }
  }

  VarCol q;
  const bool last = jc == nv - 1;
  {
    const double v = VT(c, o, VT_V, jc);
    q.wv0 = VT(c, o, VT_WSUB, jc); q.wv1 = VT(c, o, VT_WMID, jc); q.wv2 = VT(c, o, VT_WSUP, jc);
    q.c0h = 0.5 * v;
    q.rsv = (o.rho * o.sigma) * v;
    const double gsub = VT(c, o, VT_GSUB, jc), gmid = VT(c, o, VT_GMID, jc),
                 gsup = VT(c, o, VT_GSUP, jc), gz = last ? o.Gz : 0.0, split = -o.r / 2.0;
    q.f1 = col_op(o.dt, gsub, gmid, gsup, gz, split);
    q.t1 = col_op(o.tdt, gsub, gmid, gsup, gz, split);
  }
  LaneCols lc;
  lc.jc = (unsigned)jc;
  lc.jm = jc > 0 ? lc.jc - 1 : 0u;
  lc.jp = jc < nv - 1 ? lc.jc + 1 : lc.jc;
  lc.jm2 = jc > 1 ? lc.jc - 2 : 0u;
  lc.jx = jj == 0 ? lc.jm : (jj == NL - 1 ? lc.jp : lc.jc);
  lc.last = last;
  const double tdt = o.tdt, dt = o.dt;
  const bool barrier = o.barrier != 0;
  // S = 0 line (row 0 when the mesh starts at 0): A1 keeps only the -r/2 split
  const bool s0_zero = s == 0 && spr[13] == 0.0;

  // Pivot recurrence state from the row before the chunk.
  double rdp = 0.0, cp = 0.0;
  if (s > 0) {
    const double *pr = spr + (int64_t)(s - 1) * kSpRow;
    const double c00 = q.c0h * pr[11];
    rdp = rd[(unsigned)(s - 1) * unv + lc.jc];
    cp = -tdt * band_add(pr[2], pr[5], c00);
  }

  // Window: rows s-1 (a) and s (b) entered; rows s+1 .. s+PF in flight
  // (row s+1+m in ring slot m % PF).
  WinRow wa, wb, wc;
  {
    const Row4 z{0.0, 0.0, 0.0, 0.0};
    const Row4 ua = load_row4(u, clamp_row(s - 1) * unv, lc);
    const Row4 ub = load_row4(u, clamp_row(s) * unv, lc);
    Row4 ya = z, yb = z;
    if constexpr (HAS_Y) {
      ya = load_row4(y2, clamp_row(s - 1) * unv, lc);
      yb = load_row4(y2, clamp_row(s) * unv, lc);
    }
    enter_row<MODE>(q, ua, ya, wa);
    enter_row<MODE>(q, ub, yb, wb);
    if (s0_zero) zero_line<MODE>(q, wb);
  }
  using RR = RawRow<NB_DPP>;
  typename RR::type ru[PF], ry[PF];
#pragma unroll
  for (int m = 0; m < PF; ++m) {
    const unsigned rr = clamp_row(s + 1 + m) * unv;
    ru[m] = RR::load(u, rr, lc);
    ry[m] = RR::zero();
    if constexpr (HAS_Y) ry[m] = RR::load(y2, rr, lc);
  }
  // four-column tiles: table rows i and i+1 in flight
  double2 ta0, tb0, ta1, tb1;
  if constexpr (!TAB_LDS) {
    quad_row_load(spr + clamp_row(s) * kSpRow, jj, ta0, tb0);
    quad_row_load(spr + clamp_row(s + 1) * kSpRow, jj, ta1, tb1);
  }
  if constexpr (TAB_LDS) __syncthreads();
  const double *tch = tab + ch * CHP;
  // byte offset of the row the walk loads next (row s + 1 + PF + k at step k),
  // clamped to the last row; nS <= 2048 and nv <= 1024 keep it far below 2^32
  const unsigned rstep = unv * 8u, rlast = (unsigned)(nS - 1) * rstep;
  unsigned rnext = (unsigned)(s + 1 + PF) * rstep;

  double mu[MAXL], nreg[MAXL], dv[MAXL];
#pragma unroll
  for (int k = 0; k < MAXL; ++k) {
    const int i = s + k;
    const int slot = k % PF;
    // the table row of row i
    SpotRow R;
    double2 ta2, tb2;
    if constexpr (TAB_LDS) {
      R = lds_row_fields(tch + k * kSpFields);
    } else {
      quad_row_load(spr + clamp_row(i + 2) * kSpRow, jj, ta2, tb2);
    }
    // row i + 1 enters the window; its slot takes the loads of row i + 1 + PF
    {
      const Row4 un = expand_raw<NL>(ru[slot], jj, last);
      Row4 yn{0.0, 0.0, 0.0, 0.0};
      if constexpr (HAS_Y) yn = expand_raw<NL>(ry[slot], jj, last);
      enter_row<MODE>(q, un, yn, wc);
      const unsigned rb = rnext < rlast ? rnext : rlast;
      rnext += rstep;
      ru[slot] = RR::load_b(u, rb, lc);
      if constexpr (HAS_Y) ry[slot] = RR::load_b(y2, rb, lc);
    }
    if constexpr (!TAB_LDS) R = quad_row_fields(ta0, tb0);

    const Bands B = a0_bands(R, q.c0h);
    // the entries of dt A0 and theta dt A0, fl(c * a), as the reference's
    // scaled operators hold them; applied in its band order (super, main, sub)
    const double tsub = tdt * B.sub, tmid = tdt * B.mid, tsup = tdt * B.sup;
    double rhs = 0.0, ex = 0.0;
    if constexpr (MODE == SM_IMPL) {
      rhs = wb.uc + dt * B.r0;
    } else {
      const double dsub = dt * B.sub, dmid = dt * B.mid, dsup = dt * B.sup;
      auto a0d = [&](double xa, double xb, double xc) {
        return (dsup * xc + dmid * xb) + dsub * xa;
      };
      auto a0t = [&](double xa, double xb, double xc) {
        return (tsup * xc + tmid * xb) + tsub * xa;
      };
      // (A01 f)(i, j) from the rows' shares
      auto x01 = [&](double ra, double rb, double rc) {
        return q.rsv * (R.wsub * ra + R.wmid * rb + R.wsup * rc);
      };
      // Y0 = U + dt A0 U + dt R0 + dt A01 U + dt A1 U, in the reference's
      // order of its explicit operators (spot, cross, variance)
      const double y0 = ((wb.uc + (a0d(wa.uc, wb.uc, wc.uc) + dt * B.r0)) +
                         dt * x01(wa.urs, wb.urs, wc.urs)) +
                        wb.f1u;
      if constexpr (MODE == SM_FIRST) {
        rhs = y0 - a0t(wa.uc, wb.uc, wc.uc);
        ex = wb.t1u;
      } else if constexpr (MODE == SM_HV) {
        // the second-stage corrections act on the difference of the states
        // (FiniteDifferenceEngine.pyx:744-765, 818-835, 881-894)
        // V = Y0 - 1/2 F (U - Y2) over the explicit operators (no residual)
        const double v1 = ((y0 + 0.5 * a0d(wa.dc, wb.dc, wc.dc)) +
                           0.5 * (dt * x01(wa.drs, wb.drs, wc.drs))) +
                          0.5 * wb.f1d;
        rhs = v1 - a0t(wa.yc, wb.yc, wc.yc);
        ex = wb.t1y;
      } else if constexpr (MODE == SM_CS) {
        rhs = (y0 + (0.5 * dt) * x01(wa.drs, wb.drs, wc.drs)) - a0t(wa.uc, wb.uc, wc.uc);
        ex = wb.t1u;
      } else {  // SM_MCS: (1/2 - theta) F (Y2 - U), residual included
        const double h = 0.5 - o.th;
        const double xd = x01(wa.drs, wb.drs, wc.drs);
        const double v1 = (((y0 + tdt * xd) + h * (a0d(wa.dc, wb.dc, wc.dc) + dt * B.r0)) +
                           h * (dt * xd)) +
                          h * wb.f1d;
        rhs = v1 - a0t(wa.uc, wb.uc, wc.uc);
        ex = wb.t1u;
      }
    }
    // the solve forces Dirichlet rows (0, and nS-1 under a barrier) to 0
    if ((k == 0 && s == 0) || (barrier && i == nS - 1)) rhs = 0.0;
    // I - tdt*A0 on this row (identity rows on Dirichlet rows and zero
    // off-diagonals at the line ends come from the bands), scaled by the
    // pivot.  The diagonal 1 - tdt*b is one fused operation, as the
    // factorisation forms it (li0_row), written out so that the compiler's
    // contraction choices cannot change how the recomputed pivots round.
    const double rd0 = 1.0 / (__builtin_fma(-tdt, B.mid, 1.0) - (-tsub * rdp) * cp);
    rdp = rd0;
    cp = -tsup;
    const double dsc = rd0 * rhs;
    mu[k] = tsub * rd0;
    nreg[k] = tsup * rd0;
    dv[k] = dsc;
    if constexpr (HAS_EX) eb[k * THREADS + tid] = ex;
    // slide the window
    wa = wb;
    wb = wc;
    if constexpr (!TAB_LDS) {
      ta0 = ta1; tb0 = tb1; ta1 = ta2; tb1 = tb2;
    }
    // this row's results before the next row's loads (registers)
    pin(dsc);
    pin(rdp);
    if constexpr (HAS_EX) pin(ex);
    __builtin_amdgcn_sched_barrier(0);
  }
  // Chunk maps and fix-ups.  Rows k >= L belong to the next chunk; L is
  // uniform, so full chunks (L == MAXL, the common case) take unguarded loops.
  const bool full = L == MAXL;
  const bool active = col_ok && s < nS;
  double M = 1.0, E = 0.0;
  if (full) {
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
      E = dv[k] + mu[k] * E;
      M = mu[k] * M;
    }
  } else {
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
      if (k < L) {
        E = dv[k] + mu[k] * E;
        M = mu[k] * M;
      }
    }
  }
  cM[jj * (TPL + 1) + ch] = active ? M : 1.0;
  cE[jj * (TPL + 1) + ch] = active ? E : 0.0;
  __syncthreads();
  affine_scan<THREADS, NL, false>(cM, cE, cIn);
  __syncthreads();
  // Forward fix-up, then the local backward recurrence x = y' + n x'.
  {
    double yv = cIn[jj * (TPL + 1) + ch];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
      yv = dv[k] + mu[k] * yv;
      dv[k] = yv;
    }
  }
  M = 1.0;
  E = 0.0;
  if (full) {
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
      E = dv[k] + nreg[k] * E;
      M = nreg[k] * M;
    }
  } else {
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
      if (k < L) {
        E = dv[k] + nreg[k] * E;
        M = nreg[k] * M;
      }
    }
  }
  __syncthreads();  // everyone done reading cIn of the forward scan
  cM[jj * (TPL + 1) + ch] = active ? M : 1.0;
  cE[jj * (TPL + 1) + ch] = active ? E : 0.0;
  __syncthreads();
  affine_scan<THREADS, NL, true>(cM, cE, cIn);
  __syncthreads();
  // Backward fix-up; stage output x - tdt*A1(f) written straight from the chunk.
  double *ob = out + o.off;
  double x = cIn[jj * (TPL + 1) + ch];
  // byte offset of (row s + k, column j), stepped down the chunk
  unsigned sb = (unsigned)(s + MAXL - 1) * rstep + (unsigned)j * 8u;
  if (full) {
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
      const int i = s + k;
      x = dv[k] + nreg[k] * x;
      if (col_ok && i < nS) store_at(ob, sb, HAS_EX ? x - eb[k * THREADS + tid] : x);
      sb -= rstep;
    }
  } else {
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
      const int i = s + k;
      if (k < L) {
        x = dv[k] + nreg[k] * x;
        if (col_ok && i < nS) store_at(ob, sb, HAS_EX ? x - eb[k * THREADS + tid] : x);
      }
      sb -= rstep;
    }
  }
}

// ---- variance sweep -------------------------------------------------------------
// Lines along the variance axis are contiguous.  A work-group of 256 threads
// handles G = 256/TPL adjacent spot rows of one option; lane ln of a line owns
// positions [ln*L, ln*L + L) with L = ceil(nv / TPL) <= MAXL.  The rows are
// staged in LDS in a slot layout: line g, chunk ln, slot k (k < MAXL) at
// g*LP + ln*(MAXL+1) + k (odd pitch: the TPL lanes reading slot k hit
// distinct banks).  Slots past a chunk's L positions or past the line end
// hold the identity of the two recurrences (right-hand side 0, multipliers
// 1, pivot 0), so every chunk is MAXL slots long and the loops carry no
// guards.  Staging walks the 256*MAXL slots of the tile with shifts only
// (G * TPL * MAXL = 256 * MAXL).  The factorisation tables (M, N, RD, shared
// by every S != 0 line of the option) are staged the same way.  Each lane
// then solves its chunk in registers; the chunk carries of both triangular
// sweeps are combined with segmented wave scans of affine maps, the result
// goes back to LDS and is stored with coalesced writes.  The S = 0 line is
// diagonal.
template <int MODE, int TPL, int MAXL>
__global__ void __launch_bounds__(256)
var_sweep_kernel(Ctx c, const Tile *__restrict__ tiles, const double *__restrict__ W,
                 double *__restrict__ out) {
  static_assert((MAXL & (MAXL - 1)) == 0 && (TPL & (TPL - 1)) == 0, "powers of two");
  extern __shared__ double lds[];
  constexpr int G = 256 / TPL;
  constexpr int P = MAXL + 1;         // slot pitch of a chunk
  constexpr int LP = TPL * P + 1;     // line pitch
  constexpr int SL = TPL * MAXL;      // slots per line
  constexpr int NI = G * SL / 256;    // slots per thread (= MAXL)
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const HOpt o = c.opts[t.opt];
  const int nS = (int)o.nS, nv = (int)o.nv, i0 = t.start;
  const int L = (nv + TPL - 1) / TPL;
  double *tM = lds, *tN = tM + LP, *tR = tN + LP, *lines = tR + LP;
  const int tid = threadIdx.x;
  const int rows = nS - i0 < G ? nS - i0 : G;
  const double *w = W + o.off;
  const unsigned unv = (unsigned)nv;

  // Stage the tables (one line of slots) ...
#pragma unroll
  for (int a = 0; a < (SL + 255) / 256; ++a) {
    const int sq = tid + 256 * a;
    if (SL >= 256 || sq < SL) {
      const int ln = sq / MAXL, k = sq % MAXL;
      const int j = ln * L + k;
      const bool valid = k < L && j < nv;
      const int jl = valid ? j : 0;
      const double m = VT(c, o, VT_M, jl), n = VT(c, o, VT_N, jl), r = VT(c, o, VT_RD, jl);
      const int slot = ln * P + k;
      tM[slot] = valid ? m : 1.0;
      tN[slot] = valid ? n : 1.0;
      tR[slot] = valid ? r : 0.0;
    }
  }
  // ... and the rows (all loads first, then the LDS writes).
  {
    double val[NI];
#pragma unroll
    for (int a = 0; a < NI; ++a) {
      const int q = tid + 256 * a;
      const int g = q / SL, sq = q % SL;
      const int ln = sq / MAXL, k = sq % MAXL;
      const int j = ln * L + k;
      const bool valid = g < rows && k < L && j < nv;
      const unsigned ii = (unsigned)(i0 + (g < rows ? g : 0));
      const unsigned jl = valid ? (unsigned)j : 0u;
      double v = w[ii * unv + jl];
      if (MODE == VM_IMPL) {
        auto xw = [&](int64_t x, int64_t y) { return w[x * nv + y]; };
        v += o.dt * a01_dot(c, o, (int64_t)ii, (int64_t)jl, xw);
      }
      val[a] = valid ? v : 0.0;
    }
#pragma unroll
    for (int a = 0; a < NI; ++a) {
      const int q = tid + 256 * a;
      const int g = q / SL, sq = q % SL;
      const int ln = sq / MAXL, k = sq % MAXL;
      lines[g * LP + ln * P + k] = val[a];
    }
  }
  __syncthreads();
  const int ln = tid % TPL, g = tid / TPL;
  const int i = i0 + g;
  const bool row_ok = g < rows;
  const bool diag = row_ok && SP(c, o, SP_S, i) == 0.0;
  const bool fact = row_ok && !diag;
  // Bottom fold of the right-hand side, d[nv-1] += f d[nv-2], by one lane
  // per line (the slots of both positions are staged).
  if (fact && ln == 0) {
    const int j1 = nv - 1, j2 = nv - 2;
    double *ln0 = lines + g * LP;
    ln0[(j1 / L) * P + j1 % L] += o.fold_f * ln0[(j2 / L) * P + j2 % L];
  }
  __syncthreads();
  double *lp = lines + g * LP + ln * P;
  double d[MAXL], mreg[MAXL], nreg[MAXL], rreg[MAXL];
#pragma unroll
  for (int k = 0; k < MAXL; ++k) {
    d[k] = lp[k];
    mreg[k] = tM[ln * P + k];
    nreg[k] = tN[ln * P + k];
    rreg[k] = tR[ln * P + k];
  }
  // Local forward recurrence y = d + m y.
  double M = 1.0, E = 0.0;
#pragma unroll
  for (int k = 0; k < MAXL; ++k) {
    E = d[k] + mreg[k] * E;
    M = mreg[k] * M;
  }
  // Segmented inclusive prefix scan, then the carry entering each chunk.
  wave_prefix_scan<TPL>(M, E, ln);
  const double carry = from_prev_lane<TPL>(E, ln);
  {
    double y = carry;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
      y = d[k] + mreg[k] * y;
      d[k] = y;
    }
  }
  // Local backward recurrence x = rd y + n x'.
  M = 1.0;
  E = 0.0;
#pragma unroll
  for (int k = MAXL - 1; k >= 0; --k) {
    E = rreg[k] * d[k] + nreg[k] * E;
    M = nreg[k] * M;
  }
  // Segmented inclusive suffix scan (maps of higher lanes applied first).
  wave_suffix_scan<TPL>(M, E, ln);
  const double after = from_next_lane<TPL>(E, ln);
  {
    double x = after;
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
      x = rreg[k] * d[k] + nreg[k] * x;
      d[k] = x;
    }
  }
  if (diag) {
    // the S = 0 line is diagonal: x = rdS0 * rhs (slots, not the solve above)
#pragma unroll
    for (int k = 0; k < MAXL; ++k) d[k] = o.rdS0 * lp[k];
  }
#pragma unroll
  for (int k = 0; k < MAXL; ++k) lp[k] = d[k];
  __syncthreads();
  double *ob = out + o.off + (int64_t)i0 * nv;
#pragma unroll
  for (int a = 0; a < NI; ++a) {
    const int q = tid + 256 * a;
    const int gq = q / SL, sq = q % SL;
    const int lq = sq / MAXL, k = sq % MAXL;
    const int j = lq * L + k;
    if (gq < rows && k < L && j < nv) ob[(unsigned)gq * unv + (unsigned)j] = lines[gq * LP + lq * P + k];
  }
}

// ---- set-up kernels ---------------------------------------------------------------
struct SetupOpt {
  int64_t nS, nv, so, vo;  // mesh / table offsets
};

// One thread per spot node of every option: the (0,) and (0,0) template rows
// (with boundary rows) scaled by r*S and 1/2*S^2, and the cross weights.
__global__ void spot_tables_kernel(const HOpt *opts, int nopt, const int64_t *row_opt,
                                   const double *__restrict__ spots, double *sp, int64_t sp_n,
                                   double *spr) {
#pragma clang fp contract(off)
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= sp_n) return;
  const int k = (int)row_opt[g];
  const HOpt o = opts[k];
  const int64_t i = g - o.so, n = o.nS;
  const double *m = spots + o.so;
  double w1[3] = {0, 0, 0}, w2[3] = {0, 0, 0}, ws[3] = {0, 0, 0};
  double R1 = 0.0, R2 = 0.0;
  const bool dir = (i == 0) || (o.barrier && i == n - 1);
  if (i > 0 && i < n - 1) {
    const double hm = m[i] - m[i - 1], hp = m[i + 1] - m[i], x = hm + hp;
    w1[2] = hm / (hp * x);
    w1[1] = (-hm + hp) / (hm * hp);
    w1[0] = -hp / (hm * x);
    w2[2] = 2 / (hp * x);
    w2[1] = -2 / (hm * hp);
    w2[0] = 2 / (hm * x);
    ws[0] = w1[0]; ws[1] = w1[1]; ws[2] = w1[2];
  } else if (i == n - 1 && !o.barrier) {
    const double d = m[n - 1] - m[n - 2];
    R1 = 1.0;                 // dU/dS = 1 (Von Neumann)
    w2[1] = -2 / (d * d);     // second derivative with known first derivative
    w2[0] = 2 / (d * d);
    R2 = 2 / d;
  }
  const double S = m[i];
  const double c0 = o.r * S;
  const double S2h = 0.5 * (S * S);
  double *col = sp + g;
  col[SP_PSUB * sp_n] = dir ? 0.0 : w1[0] * c0;
  col[SP_PMID * sp_n] = dir ? 0.0 : w1[1] * c0 + (-0.5 * o.r);
  col[SP_PSUP * sp_n] = dir ? 0.0 : w1[2] * c0;
  col[SP_QSUB * sp_n] = dir ? 0.0 : w2[0];
  col[SP_QMID * sp_n] = dir ? 0.0 : w2[1];
  col[SP_QSUP * sp_n] = dir ? 0.0 : w2[2];
  col[SP_PR * sp_n] = dir ? 0.0 : R1 * c0;
  col[SP_QR * sp_n] = dir ? 0.0 : R2;
  col[SP_WSUB * sp_n] = ws[0];
  col[SP_WMID * sp_n] = ws[1];
  col[SP_WSUP * sp_n] = ws[2];
  col[SP_S * sp_n] = S;
  col[SP_DIR * sp_n] = dir ? 1.0 : 0.0;
  // AoS row of the spot sweep (layout at kSpRow): neighbour-of-Dirichlet band
  // entries zeroed, cross weights premultiplied by S, S^2, S, [S != 0] flag.
  double *row = spr + g * 16;
  const bool zsub = i == 1;                       // row 0 is Dirichlet
  const bool zsup = o.barrier && i == n - 2;      // row n-1 is Dirichlet
  row[0] = zsub ? 0.0 : col[SP_PSUB * sp_n];
  row[1] = col[SP_PMID * sp_n];
  row[2] = zsup ? 0.0 : col[SP_PSUP * sp_n];
  row[3] = zsub ? 0.0 : col[SP_QSUB * sp_n];
  row[4] = col[SP_QMID * sp_n];
  row[5] = zsup ? 0.0 : col[SP_QSUP * sp_n];
  row[6] = col[SP_PR * sp_n];
  row[7] = col[SP_QR * sp_n];
  row[8] = S * ws[0];
  row[9] = S * ws[1];
  row[10] = S * ws[2];
  row[11] = S * S;
  row[12] = S;
  row[13] = S != 0.0 ? 1.0 : 0.0;
  row[14] = 0.0;
  row[15] = 0.0;
}

// One thread per variance node: A1 rows for S != 0 lines and cross weights.
__global__ void var_tables_kernel(const HOpt *opts, const int64_t *row_opt,
                                  const double *__restrict__ vars, double *vt, int64_t vt_n) {
#pragma clang fp contract(off)
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= vt_n) return;
  const int k = (int)row_opt[g];
  const HOpt o = opts[k];
  const int64_t j = g - o.vo, n = o.nv;
  const double *m = vars + o.vo;
  double u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0}, wv[3] = {0, 0, 0};
  if (j > 0 && j < n - 1) {
    const double hm = m[j] - m[j - 1], hp = m[j + 1] - m[j], x = hm + hp;
    u1[2] = hm / (hp * x);
    u1[1] = (-hm + hp) / (hm * hp);
    u1[0] = -hp / (hm * x);
    u2[2] = 2 / (hp * x);
    u2[1] = -2 / (hm * hp);
    u2[0] = 2 / (hm * x);
    wv[0] = u1[0]; wv[1] = u1[1]; wv[2] = u1[2];
  } else if (j == 0) {
    const double d1 = m[1] - m[0];
    u1[2] = 1 / d1;   // one-sided first derivative; d2/dv2 row is Von Neumann 0
    u1[1] = -1 / d1;
  } else {
    const double d1 = m[n - 1] - m[n - 2], d2 = m[n - 2] - m[n - 3];
    u1[1] = 1.0 / d1;
    u1[0] = -1.0 / d1;
    const double rd = 1.0 / (0.5 * (d2 + d1) * d2 * d1);
    u2[0] = -(d2 + d1) * rd;
    u2[1] = d2 * rd;
  }
  const double v = m[j];
  const double c1 = o.kappa * (o.theta - v);
  const double c11 = (0.5 * (o.sigma * o.sigma)) * v;
  double *col = vt + g;
  col[VT_GSUB * vt_n] = u1[0] * c1 + u2[0] * c11;
  // diagonal of the combined operator for S != 0 lines, in the reference's
  // order: ((1,) band + split) + (1,1) band
  col[VT_GMID * vt_n] = (u1[1] * c1 + (-o.r / 2.0)) + u2[1] * c11;
  col[VT_GSUP * vt_n] = u1[2] * c1 + u2[2] * c11;
  col[VT_WSUB * vt_n] = wv[0];
  col[VT_WMID * vt_n] = wv[1];
  col[VT_WSUP * vt_n] = wv[2];
  col[VT_V * vt_n] = v;
}

// Per option (one thread): Gz, the folded LU of I - tdt*A1 for S != 0 lines.
__global__ void var_factor_kernel(HOpt *opts, int nopt, const double *__restrict__ vars,
                                  double *vt, int64_t vt_n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nopt) return;
  HOpt o = opts[k];
  const int64_t n = o.nv;
  const double *m = vars + o.vo;
  const double *G = vt;
  auto g = [&](int f, int64_t j) { return G[f * vt_n + o.vo + j]; };
  {
    const double d1 = m[n - 1] - m[n - 2], d2 = m[n - 2] - m[n - 3];
    const double rd = 1.0 / (0.5 * (d2 + d1) * d2 * d1);
    o.Gz = (d1 * rd) * ((0.5 * (o.sigma * o.sigma)) * m[n - 1]);
  }
  const double t = o.tdt;
  auto row = [&](int64_t j, double &a, double &b, double &c) {
    a = j > 0 ? -t * g(VT_GSUB, j) : 0.0;
    b = 1.0 - t * g(VT_GMID, j);
    c = j < n - 1 ? -t * g(VT_GSUP, j) : 0.0;
  };
  // Fold of the last row: f = -zL / aL[n-2].
  double a2, b2, c2;
  row(n - 2, a2, b2, c2);
  const double zL = -t * o.Gz;
  const double f = a2 == 0.0 ? 0.0 : -zL / a2;
  o.fold_f = f;
  o.rdS0 = 1.0 / (1.0 - t * (-0.5 * o.r));
  double *M = vt + VT_M * vt_n + o.vo, *RD = vt + VT_RD * vt_n + o.vo,
         *N = vt + VT_N * vt_n + o.vo;
  double rdprev = 0.0, cprev = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    double a, b, c;
    row(j, a, b, c);
    if (j == n - 1) {
      a += f * b2;
      b += f * c2;
    }
    const double l = a * rdprev;
    const double rdj = 1.0 / (b - l * cprev);
    M[j] = -l;
    RD[j] = rdj;
    N[j] = -c * rdj;
    rdprev = rdj;
    cprev = c;
  }
  opts[k] = o;
}

// One thread per spot line (option, column): pivots of I - tdt*A0 down the line.
__global__ void spot_factor_kernel(Ctx c, const Tile *__restrict__ tiles, int ntiles,
                                   double *rd0) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t tix = g / 64;
  if (tix >= ntiles) return;
  const Tile t = tiles[tix];
  const HOpt o = c.opts[t.opt];
  const int64_t j = t.start + (g % 64);
  if (j >= o.nv) return;
  const double v = c.vt[VT_V * c.vt_n + o.vo + j];
  double *rd = rd0 + o.off;
  double rdprev = 0.0, cprev = 0.0;
  for (int64_t i = 0; i < o.nS; ++i) {
    double a, b, cc;
    li0_row(c, o, i, v, a, b, cc);
    const double rdi = 1.0 / (b - (a * rdprev) * cprev);
    rd[i * o.nv + j] = rdi;
    rdprev = rdi;
    cprev = cc;
  }
}

// Step sizes of every option from the scheme's theta and the dt multiplier,
// taken from device scalars (d_theta, d_scale) when given, else from the host
// values; the pair used is kept in `slots` for later refactorisations.
__global__ void step_scalars_kernel(HOpt *opts, int nopt, const double *__restrict__ base_dt,
                                    const double *d_theta, const double *d_scale, double theta,
                                    double scale, int implicit, double *slots) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nopt) return;
  const double th = d_theta ? *d_theta : theta;
  const double sc = d_scale ? *d_scale : scale;
  HOpt &o = opts[k];
  o.dt = base_dt[k] * sc;
  o.tdt = implicit ? o.dt : th * o.dt;
  o.th = implicit ? 1.0 : th;
  if (k == 0 && slots) {
    slots[0] = th;
    slots[1] = sc;
  }
}

// Parameters of every option from a device array of rows (r, kappa, theta,
// sigma, rho, dt); the shapes and offsets stay.
__global__ void set_params_kernel(HOpt *opts, int nopt, const double *__restrict__ rows,
                                  double *base_dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nopt) return;
  const double *q = rows + 6 * (int64_t)k;
  HOpt &o = opts[k];
  o.r = q[0];
  o.kappa = q[1];
  o.theta = q[2];
  o.sigma = q[3];
  o.rho = q[4];
  o.dt = q[5];
  o.barrier = 0;
  base_dt[k] = q[5];
}

// One work-group per spot row of every option: V(i, j) = max(S_i - K_k, 0).
__global__ void call_payoff_kernel(const HOpt *opts, const int64_t *row_opt,
                                   const double *__restrict__ spots, int64_t sp_n,
                                   const double *__restrict__ strikes, double *state) {
  const int64_t g = blockIdx.x;
  if (g >= sp_n) return;
  const int k = (int)row_opt[g];
  const HOpt o = opts[k];
  const int64_t i = g - o.so;
  const double v = fmax(spots[g] - strikes[k], 0.0);
  double *row = state + o.off + i * o.nv;
  for (int64_t j = threadIdx.x; j < o.nv; j += blockDim.x) row[j] = v;
}

// First node of the 4-point stencil around x on mesh m[0..n): nodes
// i-2 .. i+1 with i the first node >= x, shifted to stay inside the mesh.
__device__ __forceinline__ int64_t stencil4(const double *m, int64_t n, double x) {
  int64_t lo = 0, hi = n;  // lower bound: first index with m[i] >= x
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (m[mid] < x) lo = mid + 1; else hi = mid;
  }
  int64_t s = lo - 2;
  if (s > n - 4) s = n - 4;
  if (s < 0) s = 0;
  return s;
}

__device__ __forceinline__ void lagrange4w(double x, const double *nd, double w[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double p = 1.0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (a != b) p *= (x - nd[b]) / (nd[a] - nd[b]);
    w[a] = p;
  }
}

// One thread per option: the state read at (S0_k, v0_k) by tensor-product
// cubic Lagrange interpolation on the 4 x 4 nodes around the point.
__global__ void read_bicubic_kernel(const HOpt *opts, int nopt, const double *__restrict__ spots,
                                    const double *__restrict__ vars,
                                    const double *__restrict__ state, const double *S0,
                                    const double *v0, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nopt) return;
  const HOpt o = opts[k];
  const double *ms = spots + o.so, *mv = vars + o.vo;
  const int64_t i = stencil4(ms, o.nS, S0[k]), j = stencil4(mv, o.nv, v0[k]);
  double ws[4], wv[4];
  lagrange4w(S0[k], ms + i, ws);
  lagrange4w(v0[k], mv + j, wv);
  const double *V = state + o.off;
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double *row = V + (i + a) * o.nv + j;
    acc += ws[a] * (((row[0] * wv[0] + row[1] * wv[1]) + row[2] * wv[2]) + row[3] * wv[3]);
  }
  out[k] = acc;
}

}  // namespace
}  // namespace fdadi

using namespace fdadi;

struct fdadi_hb {
  int nopt = 0;
  std::vector<HOpt> h_opts;  // host copy (dt of each option before dt_scale)
  std::vector<double> base_dt;
  HOpt *d_opts = nullptr;
  double *d_spots = nullptr, *d_vars = nullptr;
  double *d_sp = nullptr, *d_vt = nullptr, *d_spr = nullptr;
  int64_t sp_n = 0, vt_n = 0, total = 0;
  int64_t nSmax = 0, nvmax = 0;
  double *d_state = nullptr, *d_w1 = nullptr, *d_y2 = nullptr, *d_rd0 = nullptr;
  int64_t *d_sp_row = nullptr, *d_vt_row = nullptr;
  Tile *d_stiles = nullptr, *d_vtiles = nullptr, *d_ftiles = nullptr;
  int nstiles = 0, nvtiles = 0, nftiles = 0;
  int64_t cap_stiles = 0, cap_vtiles = 0;  // tile-list capacity (every launch shape fits)
  fdadi_arena *arena = nullptr;            // the one device allocation behind every buffer
  int spot_threads = 0;
  int var_tpl = 1, var_rows = 256;  // variance sweep: lanes per line, rows per 256-thread tile
  std::vector<int> order;           // options, largest first
  int spot_maxl = 8;  // rows per thread in the spot sweep (8, or 16 for long lines)
  int spot_nl = 4;    // columns per spot-sweep tile (4, 8, 16 or 32)
  // Launch groups of each sweep: a shape and a range of its tile list.  One
  // group, unless the shape is chosen per option from its line length
  // (fdadi_hb_set_shape with threads = -1): then options of one shape share
  // a group, and a stage launches once per group.
  struct SweepGroup {
    int threads, maxl;  // spot: threads and rows per thread; variance: lanes per line, -
    int first, ntiles;
  };
  std::vector<SweepGroup> sgroups, vgroups;
  bool spot_split = false, var_split = false;
  int scheme = -1;
  double theta = 0.5, dt_scale = 1.0;
  bool dev_scalars = false;        // last configure took theta and dt from device memory
  bool opts_on_device = false;     // parameters last set by fdadi_hb_update_dev (the host
                                   // copy keeps only the shapes and offsets current)
  double *d_base_dt = nullptr;     // each option's step before the multiplier
  double *d_slots = nullptr;       // [theta, dt multiplier] of the configured scheme
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<hipEvent_t> events;  // pool for fdadi_hb_step_timed
};

namespace fdadi {
namespace {

constexpr int kNL = 8;     // variance sweep: rows per tile
constexpr int kMAXL = 16;  // variance sweep: positions per thread
constexpr int kMAXLs = 8;  // spot sweep: rows per thread
constexpr int kMAXLs3 = 12;  // spot sweep, 16-column tiles chosen per option: rows per thread

Ctx ctx_of(const fdadi_hb *hb) {
  Ctx c;
  c.opts = hb->d_opts;
  c.sp = hb->d_sp;
  c.vt = hb->d_vt;
  c.sp_n = hb->sp_n;
  c.vt_n = hb->vt_n;
  c.rd0 = hb->d_rd0;
  c.spr = hb->d_spr;
  return c;
}

#define FDADI_RET(x)         \
  do {                       \
    int _st = (x);           \
    if (_st) return _st;     \
  } while (0)

// Spot-sweep shape for lines of length n: `threads` per work-group (0 =
// default), `nl` columns per tile (4, 8, 16 or 32), and the rows per thread
// (8, or 16 for long lines).  A thread keeps four values per row in
// registers, so a work-group has at most 512 threads (two waves per SIMD).
int spot_shape(int64_t n, int threads, int nl, int *t_out, int *maxl_out) {
  if (nl != 4 && nl != 8 && nl != 16 && nl != 32) {
    set_error("fdadi_hb_set_shape: spot tiles are 4, 8, 16 or 32 columns wide");
    return FDADI_ERR_ARG;
  }
  if (threads == 0) {
    if (nl == 4 && n <= 512) {
      const int64_t need = (n + kMAXLs - 1) / kMAXLs * 4;
      int t = 64;
      while (t < need) t *= 2;
      *t_out = t;
      *maxl_out = kMAXLs;
      return FDADI_OK;
    }
    threads = 512;
  }
  const int tpl = threads / nl;
  const bool t_ok = threads >= 64 && threads <= 512 && (threads & (threads - 1)) == 0 &&
                    tpl >= 1 && (nl == 4 || threads >= 256) && (nl != 32 || threads == 512);
  if (t_ok && (int64_t)tpl * kMAXLs >= n) {
    *t_out = threads;
    *maxl_out = kMAXLs;
  } else if (t_ok && (int64_t)tpl * 2 * kMAXLs >= n && (threads == 256 || threads == 512)) {
    *t_out = threads;
    *maxl_out = 2 * kMAXLs;
  } else {
    set_error("fdadi_hb_set_shape: too few threads for the longest spot line");
    return FDADI_ERR_ARG;
  }
  return FDADI_OK;
}

// Spot shapes (threads, rows per thread) for nl-column tiles, by the spot
// lines they cover ((threads / nl) * rows), smallest first.
std::vector<std::pair<int, int>> spot_candidates(int nl) {
  std::vector<std::pair<int, int>> c;
  for (int t : {64, 128, 256, 512})
    for (int ml : {kMAXLs, kMAXLs3, 2 * kMAXLs}) {
      // the launch shapes spot_shape accepts, and 12 rows per thread for
      // 16-column tiles (the per-option groups' in-between size)
      if (t / nl < 1 || (nl != 4 && t < 256) || (nl == 32 && t != 512)) continue;
      if (ml != kMAXLs && t != 256 && t != 512) continue;
      if (ml == kMAXLs3 && nl != 16) continue;
      c.push_back({t, ml});
    }
  std::stable_sort(c.begin(), c.end(), [nl](const std::pair<int, int> &a, const std::pair<int, int> &b) {
    return a.first / nl * a.second < b.first / nl * b.second;
  });
  return c;
}

// The shape a split spot sweep gives option o: the smallest candidate whose
// lines cover its nS (-1 if none does).
int spot_group_of(const std::vector<std::pair<int, int>> &c, int nl, int64_t nS) {
  for (size_t g = 0; g < c.size(); ++g)
    if ((int64_t)(c[g].first / nl) * c[g].second >= nS) return (int)g;
  return -1;
}

// (Re)build the spot-sweep tile list (spot_nl columns per tile) and its
// launch groups.
int build_spot_tiles(fdadi_hb *hb) {
  std::vector<Tile> st;
  hb->sgroups.clear();
  auto add_tiles = [&](int k) {
    const HOpt &o = hb->h_opts[k];
    // the last tile ends on the last column (it may overlap the one before,
    // whose shared columns both compute identically): the kernel takes
    // column nv - 3 of the last column from the lane two before it
    for (int64_t j0 = 0; j0 < o.nv; j0 += hb->spot_nl) {
      const int64_t start = j0 + hb->spot_nl > o.nv && o.nv >= hb->spot_nl ? o.nv - hb->spot_nl : j0;
      st.push_back(Tile{k, (int32_t)start});
    }
  };
  if (!hb->spot_split) {
    for (int k : hb->order) add_tiles(k);
    hb->sgroups.push_back({hb->spot_threads, hb->spot_maxl, 0, (int)st.size()});
  } else {
    const auto c = spot_candidates(hb->spot_nl);
    for (size_t g = 0; g < c.size(); ++g) {
      const int first = (int)st.size();
      for (int k : hb->order)  // largest first within a group
        if (spot_group_of(c, hb->spot_nl, hb->h_opts[k].nS) == (int)g) add_tiles(k);
      if ((int)st.size() > first)
        hb->sgroups.push_back({c[g].first, c[g].second, first, (int)st.size() - first});
    }
  }
  FDADI_REQUIRE((int64_t)st.size() <= hb->cap_stiles, FDADI_ERR_STATE,
                "spot tile list exceeds its capacity");
  FDADI_TRY_HIP(hipMemcpy(hb->d_stiles, st.data(), sizeof(Tile) * st.size(),
                          hipMemcpyHostToDevice));
  hb->nstiles = (int)st.size();
  return FDADI_OK;
}

// Kernels using more than 64 KB of dynamic LDS must say so once.
int allow_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return FDADI_OK;
  FDADI_REQUIRE(bytes <= 160 * 1024, FDADI_ERR_UNSUPPORTED,
                "sweep tile needs " + std::to_string(bytes) + " bytes of LDS (> 160 KB)");
  FDADI_TRY_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return FDADI_OK;
}

// Lanes per variance line so that each lane's chunk fits kMAXL registers.
int var_tpl_for(int64_t nv) {
  int t = 1;
  while ((int64_t)t * kMAXL < nv && t < 64) t *= 2;
  return t;
}

// (Re)build the variance-sweep tile list (256 / lanes-per-line spot rows per
// tile) and its launch groups; split: each option gets the fewest lanes
// whose chunks of kMAXL positions cover its variance line.
int build_var_tiles(fdadi_hb *hb) {
  std::vector<Tile> vtl;
  hb->vgroups.clear();
  auto add_tiles = [&](int k, int rows) {
    const HOpt &o = hb->h_opts[k];
    for (int64_t i0 = 0; i0 < o.nS; i0 += rows) vtl.push_back(Tile{k, (int32_t)i0});
  };
  if (!hb->var_split) {
    for (int k : hb->order) add_tiles(k, hb->var_rows);
    hb->vgroups.push_back({hb->var_tpl, 0, 0, (int)vtl.size()});
  } else {
    for (int t = 1; t <= 64; t *= 2) {
      const int first = (int)vtl.size();
      for (int k : hb->order)
        if (var_tpl_for(hb->h_opts[k].nv) == t) add_tiles(k, 256 / t);
      if ((int)vtl.size() > first) hb->vgroups.push_back({t, 0, first, (int)vtl.size() - first});
    }
  }
  FDADI_REQUIRE((int64_t)vtl.size() <= hb->cap_vtiles, FDADI_ERR_STATE,
                "variance tile list exceeds its capacity");
  FDADI_TRY_HIP(hipMemcpy(hb->d_vtiles, vtl.data(), sizeof(Tile) * vtl.size(),
                          hipMemcpyHostToDevice));
  hb->nvtiles = (int)vtl.size();
  return FDADI_OK;
}

int threads_for(int64_t n, int nl = kNL, int maxl = kMAXL) {
  // Threads per workgroup so that each line's chunk fits MAXL registers.
  const int64_t need = (n + maxl - 1) / maxl * nl;
  int t = 64;
  while (t < need && t < 1024) t *= 2;
  return t;
}

template <int MODE, int T, int NL, int ML>
int launch_spot_t(const fdadi_hb *hb, const fdadi_hb::SweepGroup &g, const double *U,
                  const double *Y, double *out, hipStream_t s) {
  constexpr int TPL = T / NL;
  // chunk carries of the scans, the next stage's explicit term (MAXL per
  // thread); wide tiles add their table rows (MAXL rows of 12 fields + 2
  // doubles of padding per chunk)
  const size_t ex = MODE == SM_IMPL ? 0 : (size_t)ML * T;
  const size_t tab = NL == 4 ? 0 : (size_t)TPL * (ML * 12 + 2);
  const size_t lds = sizeof(double) * (3 * (size_t)NL * (TPL + 1) + ex + tab);
  FDADI_RET(allow_lds((const void *)spot_sweep_kernel<MODE, T, NL, ML>, lds));
  hipLaunchKernelGGL((spot_sweep_kernel<MODE, T, NL, ML>), dim3(g.ntiles), dim3(T), lds, s,
                     ctx_of(hb), hb->d_stiles + g.first, U, Y, out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

template <int MODE, int NL>
int launch_spot_wide(const fdadi_hb *hb, const fdadi_hb::SweepGroup &g, const double *U,
                     const double *Y, double *out, hipStream_t s) {
  const bool long_rows = g.maxl == 2 * kMAXLs;
  if constexpr (NL == 16) {
    if (g.maxl == kMAXLs3) {
      if (g.threads == 256) return launch_spot_t<MODE, 256, NL, kMAXLs3>(hb, g, U, Y, out, s);
      if (g.threads == 512) return launch_spot_t<MODE, 512, NL, kMAXLs3>(hb, g, U, Y, out, s);
    }
  }
  if constexpr (NL != 32) {  // 32-column tiles: 512 threads (16 lanes per line)
    switch (g.threads * 2 + (long_rows ? 1 : 0)) {
      case 512: return launch_spot_t<MODE, 256, NL, kMAXLs>(hb, g, U, Y, out, s);
      case 513: return launch_spot_t<MODE, 256, NL, 2 * kMAXLs>(hb, g, U, Y, out, s);
    }
  }
  switch (g.threads * 2 + (long_rows ? 1 : 0)) {
    case 1024: return launch_spot_t<MODE, 512, NL, kMAXLs>(hb, g, U, Y, out, s);
    case 1025: return launch_spot_t<MODE, 512, NL, 2 * kMAXLs>(hb, g, U, Y, out, s);
  }
  set_error("spot sweep: unsupported shape for 8-, 16- or 32-column tiles");
  return FDADI_ERR_UNSUPPORTED;
}

template <int MODE>
int launch_spot_group(const fdadi_hb *hb, const fdadi_hb::SweepGroup &g, const double *U,
                      const double *Y, double *out, hipStream_t s) {
  if (hb->spot_nl == 32) return launch_spot_wide<MODE, 32>(hb, g, U, Y, out, s);
  if (hb->spot_nl == 16) return launch_spot_wide<MODE, 16>(hb, g, U, Y, out, s);
  if (hb->spot_nl == 8) return launch_spot_wide<MODE, 8>(hb, g, U, Y, out, s);
  if (g.maxl == 2 * kMAXLs) {
    switch (g.threads) {
      case 256: return launch_spot_t<MODE, 256, 4, 2 * kMAXLs>(hb, g, U, Y, out, s);
      case 512: return launch_spot_t<MODE, 512, 4, 2 * kMAXLs>(hb, g, U, Y, out, s);
    }
    set_error("spot sweep: unsupported thread count for 16 rows per thread");
    return FDADI_ERR_UNSUPPORTED;
  }
  switch (g.threads) {
    case 64: return launch_spot_t<MODE, 64, 4, kMAXLs>(hb, g, U, Y, out, s);
    case 128: return launch_spot_t<MODE, 128, 4, kMAXLs>(hb, g, U, Y, out, s);
    case 256: return launch_spot_t<MODE, 256, 4, kMAXLs>(hb, g, U, Y, out, s);
    case 512: return launch_spot_t<MODE, 512, 4, kMAXLs>(hb, g, U, Y, out, s);
  }
  set_error("spot sweep: unsupported thread count");
  return FDADI_ERR_UNSUPPORTED;
}

// One spot-sweep stage: a launch per group.
template <int MODE>
int launch_spot(const fdadi_hb *hb, const double *U, const double *Y, double *out,
                hipStream_t s) {
  for (const fdadi_hb::SweepGroup &g : hb->sgroups)
    FDADI_RET(launch_spot_group<MODE>(hb, g, U, Y, out, s));
  return FDADI_OK;
}

template <int MODE, int TPL>
int launch_var_t(const fdadi_hb *hb, const fdadi_hb::SweepGroup &g, const double *W, double *out,
                 hipStream_t s) {
  // LDS: three tables and 256/TPL lines of TPL chunks (pitch kMAXL + 1) + 1.
  const size_t lds = sizeof(double) * (3 + 256 / TPL) * (TPL * (kMAXL + 1) + 1);
  FDADI_RET(allow_lds((const void *)var_sweep_kernel<MODE, TPL, kMAXL>, lds));
  hipLaunchKernelGGL((var_sweep_kernel<MODE, TPL, kMAXL>), dim3(g.ntiles), dim3(256), lds, s,
                     ctx_of(hb), hb->d_vtiles + g.first, W, out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

// One variance-sweep stage: a launch per group.
template <int MODE>
int launch_var(const fdadi_hb *hb, const double *W, double *out, hipStream_t s) {
  for (const fdadi_hb::SweepGroup &g : hb->vgroups) {
    int st = FDADI_ERR_UNSUPPORTED;
    switch (g.threads) {
      case 1: st = launch_var_t<MODE, 1>(hb, g, W, out, s); break;
      case 2: st = launch_var_t<MODE, 2>(hb, g, W, out, s); break;
      case 4: st = launch_var_t<MODE, 4>(hb, g, W, out, s); break;
      case 8: st = launch_var_t<MODE, 8>(hb, g, W, out, s); break;
      case 16: st = launch_var_t<MODE, 16>(hb, g, W, out, s); break;
      case 32: st = launch_var_t<MODE, 32>(hb, g, W, out, s); break;
      case 64: st = launch_var_t<MODE, 64>(hb, g, W, out, s); break;
      default: set_error("variance sweep: unsupported lanes per line");
    }
    FDADI_RET(st);
  }
  return FDADI_OK;
}

// Optional per-stage event pairs (fdadi_hb_step_timed); null = untimed.
struct StageEvents {
  hipEvent_t *ev;  // 2 per stage, in launch order
  int next;
};

int mark(StageEvents *t, hipStream_t s) {
  if (!t) return FDADI_OK;
  FDADI_TRY_HIP(hipEventRecord(t->ev[t->next++], s));
  return FDADI_OK;
}

#define STAGE(call)                 \
  do {                              \
    FDADI_RET(mark(tm, s));         \
    FDADI_RET(call);                \
    FDADI_RET(mark(tm, s));         \
  } while (0)

// One time step of the configured scheme (stage list of the module comment).
int step_once(fdadi_hb *hb, hipStream_t s, StageEvents *tm = nullptr) {
  double *U = hb->d_state, *W = hb->d_w1, *Y2 = hb->d_y2;
  switch (hb->scheme) {
    case FDADI_SCHEME_DOUGLAS:
      STAGE((launch_spot<SM_FIRST>(hb, U, nullptr, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, U, s)));
      return FDADI_OK;
    case FDADI_SCHEME_IMPLICIT:
      STAGE((launch_var<VM_IMPL>(hb, U, W, s)));
      STAGE((launch_spot<SM_IMPL>(hb, W, nullptr, U, s)));
      return FDADI_OK;
    case FDADI_SCHEME_HV:
      STAGE((launch_spot<SM_FIRST>(hb, U, nullptr, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, Y2, s)));
      STAGE((launch_spot<SM_HV>(hb, U, Y2, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, U, s)));
      return FDADI_OK;
    case FDADI_SCHEME_CS:
      STAGE((launch_spot<SM_FIRST>(hb, U, nullptr, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, Y2, s)));
      STAGE((launch_spot<SM_CS>(hb, U, Y2, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, U, s)));
      return FDADI_OK;
    case FDADI_SCHEME_MCS:
      STAGE((launch_spot<SM_FIRST>(hb, U, nullptr, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, Y2, s)));
      STAGE((launch_spot<SM_MCS>(hb, U, Y2, W, s)));
      STAGE((launch_var<VM_PLAIN>(hb, W, U, s)));
      return FDADI_OK;
  }
  set_error("fdadi_hb_step: engine not configured (call fdadi_hb_configure)");
  return FDADI_ERR_STATE;
}

// Spot and variance tables from the meshes and parameters on the device.
int build_tables(fdadi_hb *hb, hipStream_t s) {
  hipLaunchKernelGGL(spot_tables_kernel, dim3(grid_for(hb->sp_n, 256)), dim3(256), 0, s,
                     hb->d_opts, hb->nopt, hb->d_sp_row, hb->d_spots, hb->d_sp, hb->sp_n,
                     hb->d_spr);
  FDADI_CHECK_LAUNCH();
  hipLaunchKernelGGL(var_tables_kernel, dim3(grid_for(hb->vt_n, 256)), dim3(256), 0, s,
                     hb->d_opts, hb->d_vt_row, hb->d_vars, hb->d_vt, hb->vt_n);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

// Step sizes of the configured scheme and the time-invariant factorisations
// of I - tdt A0 and I - tdt A1, on the device.  theta and the dt multiplier
// come from the device scalars d_theta / d_scale when given (nothing is read
// back to the host and nothing synchronises), else from hb->theta /
// hb->dt_scale (synchronises).
int refactor(fdadi_hb *hb, hipStream_t s, const double *d_theta = nullptr,
             const double *d_scale = nullptr, bool capturing = false) {
  const bool dev = d_theta || d_scale;
  // the device copy of the options is current after fdadi_hb_update_dev (and
  // is made current before a capture, which records kernels only)
  if (!hb->opts_on_device && !capturing)
    FDADI_TRY_HIP(hipMemcpyAsync(hb->d_opts, hb->h_opts.data(), sizeof(HOpt) * hb->nopt,
                                 hipMemcpyHostToDevice, s));
  const bool reuse = d_theta == hb->d_slots;  // refactorisation with the kept pair
  hipLaunchKernelGGL(step_scalars_kernel, dim3(grid_for(hb->nopt, 64)), dim3(64), 0, s,
                     hb->d_opts, hb->nopt, hb->d_base_dt, d_theta, d_scale, hb->theta,
                     hb->dt_scale, (int)(hb->scheme == FDADI_SCHEME_IMPLICIT),
                     reuse ? nullptr : hb->d_slots);
  FDADI_CHECK_LAUNCH();
  hipLaunchKernelGGL(var_factor_kernel, dim3(grid_for(hb->nopt, 64)), dim3(64), 0, s, hb->d_opts,
                     hb->nopt, hb->d_vars, hb->d_vt, hb->vt_n);
  FDADI_CHECK_LAUNCH();
  hipLaunchKernelGGL(spot_factor_kernel, dim3(hb->nftiles), dim3(64), 0, s, ctx_of(hb),
                     hb->d_ftiles, hb->nftiles, hb->d_rd0);
  FDADI_CHECK_LAUNCH();
  if (!dev && !capturing) FDADI_TRY_HIP(hipStreamSynchronize(s));
  return FDADI_OK;
}

void drop_graph(fdadi_hb *hb) {
  if (hb->gexec) {
    (void)hipGraphExecDestroy(hb->gexec);
    hb->gexec = nullptr;
  }
  if (hb->graph) {
    (void)hipGraphDestroy(hb->graph);
    hb->graph = nullptr;
  }
}

void free_hb(fdadi_hb *hb) {
  if (!hb) return;
  if (hb->gexec) (void)hipGraphExecDestroy(hb->gexec);
  if (hb->graph) (void)hipGraphDestroy(hb->graph);
  for (hipEvent_t e : hb->events) (void)hipEventDestroy(e);
  (void)fdadi_arena_destroy(hb->arena);
  delete hb;
}

}  // namespace
}  // namespace fdadi

extern "C" {

int fdadi_hb_create(int32_t nopt, const fdadi_heston_option *opts, const double *spots,
                    const double *vars, void *stream, fdadi_hb **out) {
  FDADI_REQUIRE(out && opts && spots && vars && nopt > 0, FDADI_ERR_ARG,
                "fdadi_hb_create: bad arguments");
  hipStream_t s = as_stream(stream);
  fdadi_hb *hb = new fdadi_hb();
  hb->nopt = nopt;
  hb->h_opts.resize(nopt);
  hb->base_dt.resize(nopt);
  int64_t so = 0, vo = 0, off = 0;
  std::vector<int64_t> sp_row, vt_row;
  for (int k = 0; k < nopt; ++k) {
    const fdadi_heston_option &q = opts[k];
    if (q.nS < 4 || q.nv < 4) {
      free_hb(hb);
      set_error("fdadi_hb_create: option " + std::to_string(k) +
                ": grid needs at least 4 points per axis");
      return FDADI_ERR_ARG;
    }
    if (q.nS > 2048 || q.nv > 1024) {
      free_hb(hb);
      set_error("fdadi_hb_create: option " + std::to_string(k) +
                ": the fused engine supports nS <= 2048 and nv <= 1024");
      return FDADI_ERR_UNSUPPORTED;
    }
    HOpt &o = hb->h_opts[k];
    o = HOpt{};
    o.nS = q.nS;
    o.nv = q.nv;
    o.off = off;
    o.so = so;
    o.vo = vo;
    o.r = q.r;
    o.kappa = q.kappa;
    o.theta = q.theta;
    o.sigma = q.sigma;
    o.rho = q.rho;
    o.dt = q.dt;
    o.barrier = q.barrier ? 1 : 0;
    hb->base_dt[k] = q.dt;
    for (int64_t i = 0; i < q.nS; ++i) sp_row.push_back(k);
    for (int64_t j = 0; j < q.nv; ++j) vt_row.push_back(k);
    so += q.nS;
    vo += q.nv;
    off += q.nS * q.nv;
    hb->nSmax = std::max<int64_t>(hb->nSmax, q.nS);
    hb->nvmax = std::max<int64_t>(hb->nvmax, q.nv);
  }
  hb->sp_n = so;
  hb->vt_n = vo;
  hb->total = off;
  // Tiles, largest options first (the dispatcher then balances the tail).
  std::vector<int> order(nopt);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return hb->h_opts[a].nS * hb->h_opts[a].nv > hb->h_opts[b].nS * hb->h_opts[b].nv;
  });
  std::vector<Tile> ft;
  for (int k : order) {
    const HOpt &o = hb->h_opts[k];
    for (int64_t j0 = 0; j0 < o.nv; j0 += 64) ft.push_back(Tile{k, (int32_t)j0});
  }
  hb->nftiles = (int)ft.size();
  spot_shape(hb->nSmax, 0, 4, &hb->spot_threads, &hb->spot_maxl);
  hb->order = order;
  hb->var_tpl = var_tpl_for(hb->nvmax);
  hb->var_rows = 256 / hb->var_tpl;
  // Every device buffer of the engine is carved from one arena (one
  // hipMalloc); the tile lists get the capacity of their largest launch
  // shape (4-column spot tiles, variance tiles of 4 rows) so that changing
  // shapes later allocates nothing.
  for (int k = 0; k < nopt; ++k) {
    hb->cap_stiles += (hb->h_opts[k].nv + 3) / 4;
    hb->cap_vtiles += (hb->h_opts[k].nS + 3) / 4;
  }
  struct Part {
    void **p;
    int64_t bytes;
  };
  const Part parts[] = {
      {(void **)&hb->d_opts, (int64_t)sizeof(HOpt) * nopt},
      {(void **)&hb->d_spots, 8 * so},
      {(void **)&hb->d_vars, 8 * vo},
      {(void **)&hb->d_sp, 8 * SP_NF * so},
      {(void **)&hb->d_vt, 8 * VT_NF * vo},
      {(void **)&hb->d_spr, 8 * 16 * so},
      {(void **)&hb->d_state, 8 * off},
      {(void **)&hb->d_w1, 8 * off},
      {(void **)&hb->d_y2, 8 * off},
      {(void **)&hb->d_rd0, 8 * off},
      {(void **)&hb->d_sp_row, 8 * so},
      {(void **)&hb->d_vt_row, 8 * vo},
      {(void **)&hb->d_stiles, (int64_t)sizeof(Tile) * hb->cap_stiles},
      {(void **)&hb->d_vtiles, (int64_t)sizeof(Tile) * hb->cap_vtiles},
      {(void **)&hb->d_ftiles, (int64_t)sizeof(Tile) * (int64_t)ft.size()},
      {(void **)&hb->d_base_dt, 8 * (int64_t)nopt},
      {(void **)&hb->d_slots, 16},
  };
  int64_t total_bytes = 0;
  for (const Part &q : parts) total_bytes += (q.bytes + 255) / 256 * 256;
  if (fdadi_arena_create(total_bytes, &hb->arena) != FDADI_OK) {
    const std::string msg = fdadi_last_error();
    free_hb(hb);
    set_error("fdadi_hb_create: " + msg);
    return FDADI_ERR_HIP;
  }
  for (const Part &q : parts) {
    double *d = nullptr;
    (void)fdadi_arena_alloc(hb->arena, (q.bytes + 7) / 8, &d);  // fits by construction
    *q.p = d;
  }
  hipError_t e = hipSuccess;
  auto up = [&](void *dst, const void *src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  up(hb->d_opts, hb->h_opts.data(), sizeof(HOpt) * nopt);
  up(hb->d_spots, spots, sizeof(double) * so);
  up(hb->d_vars, vars, sizeof(double) * vo);
  up(hb->d_sp_row, sp_row.data(), sizeof(int64_t) * so);
  up(hb->d_vt_row, vt_row.data(), sizeof(int64_t) * vo);
  up(hb->d_ftiles, ft.data(), sizeof(Tile) * ft.size());
  up(hb->d_base_dt, hb->base_dt.data(), sizeof(double) * nopt);
  if (e == hipSuccess && build_spot_tiles(hb) != FDADI_OK) e = hipErrorUnknown;
  if (e == hipSuccess && build_var_tiles(hb) != FDADI_OK) e = hipErrorUnknown;
  if (e == hipSuccess && build_tables(hb, s) != FDADI_OK) e = hipErrorUnknown;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    free_hb(hb);
    set_error(std::string("fdadi_hb_create: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  *out = hb;
  return FDADI_OK;
}

int fdadi_hb_destroy(fdadi_hb *hb) {
  free_hb(hb);
  return FDADI_OK;
}

int fdadi_hb_configure(fdadi_hb *hb, int scheme, double theta, double dt_scale, void *stream) {
  FDADI_REQUIRE(hb, FDADI_ERR_ARG, "fdadi_hb_configure: null engine");
  FDADI_REQUIRE(scheme >= FDADI_SCHEME_IMPLICIT && scheme <= FDADI_SCHEME_HV, FDADI_ERR_ARG,
                "fdadi_hb_configure: unknown scheme");
  FDADI_REQUIRE(dt_scale > 0, FDADI_ERR_ARG, "fdadi_hb_configure: dt_scale must be positive");
  hipStream_t s = as_stream(stream);
  // A captured graph stays valid while the scheme (the kernel sequence) is
  // unchanged: theta and dt live in device memory and are re-read each step.
  const bool keep_graph = scheme == hb->scheme;
  hb->scheme = scheme;
  hb->theta = theta;
  hb->dt_scale = dt_scale;
  hb->dev_scalars = false;
  FDADI_RET(refactor(hb, s));
  if (!keep_graph) {
    if (hb->gexec) {
      (void)hipGraphExecDestroy(hb->gexec);
      hb->gexec = nullptr;
    }
    if (hb->graph) {
      (void)hipGraphDestroy(hb->graph);
      hb->graph = nullptr;
    }
  }
  return FDADI_OK;
}

int fdadi_hb_configure_dev(fdadi_hb *hb, int scheme, const double *d_theta,
                           const double *d_dt_scale, void *stream) {
  FDADI_REQUIRE(hb && d_theta && d_dt_scale, FDADI_ERR_ARG,
                "fdadi_hb_configure_dev: bad arguments");
  FDADI_REQUIRE(scheme >= FDADI_SCHEME_IMPLICIT && scheme <= FDADI_SCHEME_HV, FDADI_ERR_ARG,
                "fdadi_hb_configure_dev: unknown scheme");
  const bool keep_graph = scheme == hb->scheme;
  hb->scheme = scheme;
  hb->dev_scalars = true;
  FDADI_RET(refactor(hb, as_stream(stream), d_theta, d_dt_scale));
  if (!keep_graph) {
    if (hb->gexec) {
      (void)hipGraphExecDestroy(hb->gexec);
      hb->gexec = nullptr;
    }
    if (hb->graph) {
      (void)hipGraphDestroy(hb->graph);
      hb->graph = nullptr;
    }
  }
  return FDADI_OK;
}

int fdadi_hb_update(fdadi_hb *hb, int32_t nopt, const fdadi_heston_option *opts,
                    const double *spots, const double *vars, void *stream) {
  FDADI_REQUIRE(hb && opts && spots && vars, FDADI_ERR_ARG, "fdadi_hb_update: bad arguments");
  FDADI_REQUIRE(nopt == hb->nopt, FDADI_ERR_ARG,
                "fdadi_hb_update: option count differs from the engine's");
  for (int k = 0; k < nopt; ++k)
    FDADI_REQUIRE(opts[k].nS == hb->h_opts[k].nS && opts[k].nv == hb->h_opts[k].nv,
                  FDADI_ERR_ARG,
                  "fdadi_hb_update: option " + std::to_string(k) + ": grid shape differs");
  hipStream_t s = as_stream(stream);
  for (int k = 0; k < nopt; ++k) {
    HOpt &o = hb->h_opts[k];
    const fdadi_heston_option &q = opts[k];
    o.r = q.r;
    o.kappa = q.kappa;
    o.theta = q.theta;
    o.sigma = q.sigma;
    o.rho = q.rho;
    o.dt = q.dt;
    o.barrier = q.barrier ? 1 : 0;
    hb->base_dt[k] = q.dt;
  }
  // The host arrays may be pageable: synchronous copies, after the stream's
  // earlier work (which may still read the tables) has finished.
  FDADI_TRY_HIP(hipStreamSynchronize(s));
  hb->opts_on_device = false;
  FDADI_TRY_HIP(hipMemcpy(hb->d_opts, hb->h_opts.data(), sizeof(HOpt) * nopt,
                          hipMemcpyHostToDevice));
  FDADI_TRY_HIP(hipMemcpy(hb->d_spots, spots, sizeof(double) * hb->sp_n, hipMemcpyHostToDevice));
  FDADI_TRY_HIP(hipMemcpy(hb->d_vars, vars, sizeof(double) * hb->vt_n, hipMemcpyHostToDevice));
  FDADI_TRY_HIP(hipMemcpy(hb->d_base_dt, hb->base_dt.data(), sizeof(double) * nopt,
                          hipMemcpyHostToDevice));
  FDADI_RET(build_tables(hb, s));
  if (hb->scheme >= 0) {
    if (hb->dev_scalars)
      FDADI_RET(refactor(hb, s, hb->d_slots, hb->d_slots + 1));
    else
      FDADI_RET(refactor(hb, s));
  }
  FDADI_TRY_HIP(hipStreamSynchronize(s));
  return FDADI_OK;
}

int fdadi_hb_update_dev(fdadi_hb *hb, const double *d_params, const double *d_spots,
                        const double *d_vars, void *stream) {
  FDADI_REQUIRE(hb && d_params, FDADI_ERR_ARG, "fdadi_hb_update_dev: bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(set_params_kernel, dim3(grid_for(hb->nopt, 64)), dim3(64), 0, s,
                     hb->d_opts, hb->nopt, d_params, hb->d_base_dt);
  FDADI_CHECK_LAUNCH();
  if (d_spots)
    FDADI_TRY_HIP(hipMemcpyAsync(hb->d_spots, d_spots, sizeof(double) * hb->sp_n,
                                 hipMemcpyDeviceToDevice, s));
  if (d_vars)
    FDADI_TRY_HIP(hipMemcpyAsync(hb->d_vars, d_vars, sizeof(double) * hb->vt_n,
                                 hipMemcpyDeviceToDevice, s));
  hb->opts_on_device = true;
  FDADI_RET(build_tables(hb, s));
  // theta and the dt multiplier of the configured scheme stay in d_slots
  if (hb->scheme >= 0) FDADI_RET(refactor(hb, s, hb->d_slots, hb->d_slots + 1));
  return FDADI_OK;
}

int fdadi_hb_call_payoff(fdadi_hb *hb, const double *d_strikes, void *stream) {
  FDADI_REQUIRE(hb && d_strikes, FDADI_ERR_ARG, "fdadi_hb_call_payoff: bad arguments");
  hipLaunchKernelGGL(call_payoff_kernel, dim3(hb->sp_n), dim3(128), 0, as_stream(stream),
                     hb->d_opts, hb->d_sp_row, hb->d_spots, hb->sp_n, d_strikes, hb->d_state);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_hb_read_bicubic(const fdadi_hb *hb, const double *d_S0, const double *d_v0,
                          double *d_out, void *stream) {
  FDADI_REQUIRE(hb && d_S0 && d_v0 && d_out, FDADI_ERR_ARG,
                "fdadi_hb_read_bicubic: bad arguments");
  hipLaunchKernelGGL(read_bicubic_kernel, dim3(grid_for(hb->nopt, 64)), dim3(64), 0,
                     as_stream(stream), hb->d_opts, hb->nopt, hb->d_spots, hb->d_vars, hb->d_state,
                     d_S0, d_v0, d_out);
  FDADI_CHECK_LAUNCH();
  return FDADI_OK;
}

int fdadi_hb_state(fdadi_hb *hb, double **d_state, int64_t *total) {
  FDADI_REQUIRE(hb, FDADI_ERR_ARG, "fdadi_hb_state: null engine");
  if (d_state) *d_state = hb->d_state;
  if (total) *total = hb->total;
  return FDADI_OK;
}

int fdadi_hb_offsets(const fdadi_hb *hb, int64_t *offsets) {
  FDADI_REQUIRE(hb && offsets, FDADI_ERR_ARG, "fdadi_hb_offsets: bad arguments");
  for (int k = 0; k < hb->nopt; ++k) offsets[k] = hb->h_opts[k].off;
  return FDADI_OK;
}

int fdadi_hb_step(fdadi_hb *hb, int32_t nsteps, void *stream) {
  FDADI_REQUIRE(hb, FDADI_ERR_ARG, "fdadi_hb_step: null engine");
  FDADI_REQUIRE(nsteps >= 0, FDADI_ERR_ARG, "fdadi_hb_step: negative step count");
  hipStream_t s = as_stream(stream);
  for (int32_t n = 0; n < nsteps; ++n) FDADI_RET(step_once(hb, s));
  return FDADI_OK;
}

int fdadi_hb_step_timed(fdadi_hb *hb, int32_t nsteps, void *stream, double *stage_ms,
                        int32_t max_stages, int32_t *nstages) {
  FDADI_REQUIRE(hb && stage_ms && nstages && nsteps > 0, FDADI_ERR_ARG,
                "fdadi_hb_step_timed: bad arguments");
  // ADI stages per step (a stage is one launch per group of its sweep)
  FDADI_REQUIRE(hb->scheme >= 0, FDADI_ERR_STATE, "fdadi_hb_step_timed: engine not configured");
  const int per =
      hb->scheme == FDADI_SCHEME_DOUGLAS || hb->scheme == FDADI_SCHEME_IMPLICIT ? 2 : 4;
  FDADI_REQUIRE(per <= max_stages, FDADI_ERR_ARG, "fdadi_hb_step_timed: stage_ms too short");
  hipStream_t s = as_stream(stream);
  const size_t need = (size_t)2 * per * nsteps;
  while (hb->events.size() < need) {
    hipEvent_t e;
    FDADI_TRY_HIP(hipEventCreate(&e));
    hb->events.push_back(e);
  }
  StageEvents tm{hb->events.data(), 0};
  for (int32_t n = 0; n < nsteps; ++n) FDADI_RET(step_once(hb, s, &tm));
  FDADI_TRY_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < per; ++k) stage_ms[k] = 0.0;
  for (int32_t n = 0; n < nsteps; ++n)
    for (int k = 0; k < per; ++k) {
      float ms = 0.f;
      const size_t b = (size_t)2 * (n * per + k);
      FDADI_TRY_HIP(hipEventElapsedTime(&ms, hb->events[b], hb->events[b + 1]));
      stage_ms[k] += ms;
    }
  *nstages = per;
  return FDADI_OK;
}

int fdadi_hb_capture(fdadi_hb *hb, int32_t nsteps, void *stream) {
  FDADI_REQUIRE(hb, FDADI_ERR_ARG, "fdadi_hb_capture: null engine");
  FDADI_REQUIRE(nsteps > 0, FDADI_ERR_ARG, "fdadi_hb_capture: step count must be positive");
  FDADI_REQUIRE(hb->scheme >= 0, FDADI_ERR_STATE, "fdadi_hb_capture: engine not configured");
  FDADI_REQUIRE(stream, FDADI_ERR_ARG, "fdadi_hb_capture: needs a created stream");
  hipStream_t s = as_stream(stream);
  if (hb->gexec) {
    (void)hipGraphExecDestroy(hb->gexec);
    hb->gexec = nullptr;
  }
  if (hb->graph) {
    (void)hipGraphDestroy(hb->graph);
    hb->graph = nullptr;
  }
  FDADI_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int st = FDADI_OK;
  for (int32_t n = 0; n < nsteps && st == FDADI_OK; ++n) st = step_once(hb, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (st != FDADI_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) {
    set_error(std::string("fdadi_hb_capture: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  hb->graph = g;
  FDADI_TRY_HIP(hipGraphInstantiate(&hb->gexec, g, nullptr, nullptr, 0));
  return FDADI_OK;
}

int fdadi_hb_capture_solve(fdadi_hb *hb, int scheme, double theta, int32_t smoothing,
                           int32_t nsteps, void *stream) {
  FDADI_REQUIRE(hb, FDADI_ERR_ARG, "fdadi_hb_capture_solve: null engine");
  FDADI_REQUIRE(scheme >= FDADI_SCHEME_IMPLICIT && scheme <= FDADI_SCHEME_HV, FDADI_ERR_ARG,
                "fdadi_hb_capture_solve: unknown scheme");
  FDADI_REQUIRE(smoothing >= 0 && nsteps > 0, FDADI_ERR_ARG,
                "fdadi_hb_capture_solve: smoothing >= 0 and a positive step count");
  FDADI_REQUIRE(stream, FDADI_ERR_ARG, "fdadi_hb_capture_solve: needs a created stream");
  hipStream_t s = as_stream(stream);
  drop_graph(hb);
  // the device copy of the options made current here: the capture records
  // kernels only (no host copy, no synchronisation)
  if (!hb->opts_on_device) {
    FDADI_TRY_HIP(hipMemcpyAsync(hb->d_opts, hb->h_opts.data(), sizeof(HOpt) * hb->nopt,
                                 hipMemcpyHostToDevice, s));
    FDADI_TRY_HIP(hipStreamSynchronize(s));
    hb->opts_on_device = true;
  }
  const int prev_scheme = hb->scheme;
  const double prev_theta = hb->theta, prev_scale = hb->dt_scale;
  const bool prev_dev = hb->dev_scalars;
  hb->dev_scalars = false;
  FDADI_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int st = FDADI_OK;
  if (smoothing > 0) {
    // 2 * smoothing implicit steps of half the step
    hb->scheme = FDADI_SCHEME_IMPLICIT;
    hb->theta = 1.0;
    hb->dt_scale = 0.5;
    st = refactor(hb, s, nullptr, nullptr, true);
    for (int32_t n = 0; n < 2 * smoothing && st == FDADI_OK; ++n) st = step_once(hb, s);
  }
  if (st == FDADI_OK) {
    hb->scheme = scheme;
    hb->theta = theta;
    hb->dt_scale = 1.0;
    st = refactor(hb, s, nullptr, nullptr, true);
  }
  for (int32_t n = 0; n < nsteps && st == FDADI_OK; ++n) st = step_once(hb, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (st != FDADI_OK || e != hipSuccess) {
    // nothing ran: the engine keeps its previous configuration
    if (g) (void)hipGraphDestroy(g);
    hb->scheme = prev_scheme;
    hb->theta = prev_theta;
    hb->dt_scale = prev_scale;
    hb->dev_scalars = prev_dev;
    if (st != FDADI_OK) return st;
    set_error(std::string("fdadi_hb_capture_solve: ") + hipGetErrorString(e));
    return FDADI_ERR_HIP;
  }
  hb->graph = g;
  FDADI_TRY_HIP(hipGraphInstantiate(&hb->gexec, g, nullptr, nullptr, 0));
  // the capture recorded the refactorisations without running them: run the
  // last one now, so that eager steps after this call see `scheme` as well
  return refactor(hb, s);
}

int fdadi_hb_replay(fdadi_hb *hb, void *stream) {
  FDADI_REQUIRE(hb && hb->gexec, FDADI_ERR_STATE, "fdadi_hb_replay: no captured graph");
  FDADI_TRY_HIP(hipGraphLaunch(hb->gexec, as_stream(stream)));
  return FDADI_OK;
}

int fdadi_hb_set_shape(fdadi_hb *hb, int kind, int threads, int lines) {
  FDADI_REQUIRE(hb && (kind == 0 || kind == 1), FDADI_ERR_ARG, "fdadi_hb_set_shape: bad kind");
  FDADI_REQUIRE(kind == 0 || lines == 0, FDADI_ERR_UNSUPPORTED,
                "fdadi_hb_set_shape: variance tiles follow from the lanes per line (pass 0)");
  if (kind == 0) {
    FDADI_REQUIRE(threads == -1 || threads == 0 || threads == 64 || threads == 128 ||
                      threads == 256 || threads == 512,
                  FDADI_ERR_ARG,
                  "fdadi_hb_set_shape: threads must be a power of two in [64, 512], 0 or -1");
    const int nl = lines == 0 ? 4 : lines;
    if (nl == 32) {
      // the last column must fall on a tile's last lane (see from_prev2_col)
      for (const HOpt &o : hb->h_opts)
        FDADI_REQUIRE(o.nv >= 32, FDADI_ERR_ARG,
                      "fdadi_hb_set_shape: 32-column spot tiles need at least 32 variance nodes");
    }
    if (threads == -1) {
      int t = 0, ml = 0;
      FDADI_RET(spot_shape(1, 0, nl, &t, &ml));  // validates nl
      const auto c = spot_candidates(nl);
      for (const HOpt &o : hb->h_opts)
        FDADI_REQUIRE(spot_group_of(c, nl, o.nS) >= 0, FDADI_ERR_ARG,
                      "fdadi_hb_set_shape: no spot shape covers a line of " +
                          std::to_string(o.nS) + " nodes");
      hb->spot_split = true;
    } else {
      int t = 0, ml = 0;
      FDADI_RET(spot_shape(hb->nSmax, threads, nl, &t, &ml));
      hb->spot_split = false;
      hb->spot_threads = t;
      hb->spot_maxl = ml;
    }
    hb->spot_nl = nl;
    FDADI_RET(build_spot_tiles(hb));
  } else {
    // kind 1: `threads` = lanes per variance line (power of two <= 64), or
    // -1: per option, the fewest lanes that cover its line
    if (threads == -1) {
      hb->var_split = true;
    } else {
      const int t = threads == 0 ? var_tpl_for(hb->nvmax) : threads;
      FDADI_REQUIRE(t >= 1 && t <= 64 && (t & (t - 1)) == 0, FDADI_ERR_ARG,
                    "fdadi_hb_set_shape: lanes per variance line must be a power of two <= 64");
      FDADI_REQUIRE((int64_t)t * kMAXL >= hb->nvmax, FDADI_ERR_ARG,
                    "fdadi_hb_set_shape: too few lanes for the longest variance line");
      hb->var_split = false;
      hb->var_tpl = t;
      hb->var_rows = 256 / t;
    }
    FDADI_RET(build_var_tiles(hb));
  }
  if (hb->gexec) {
    (void)hipGraphExecDestroy(hb->gexec);
    hb->gexec = nullptr;
  }
  if (hb->graph) {
    (void)hipGraphDestroy(hb->graph);
    hb->graph = nullptr;
  }
  return FDADI_OK;
}

int fdadi_hb_get_shape(const fdadi_hb *hb, int kind, int *threads, int *lines) {
  FDADI_REQUIRE(hb && (kind == 0 || kind == 1), FDADI_ERR_ARG, "fdadi_hb_get_shape: bad kind");
  if (threads)
    *threads = kind == 0 ? (hb->spot_split ? -1 : hb->spot_threads)
                         : (hb->var_split ? -1 : hb->var_tpl);
  if (lines) *lines = kind == 0 ? hb->spot_nl : (hb->var_split ? 0 : hb->var_rows);
  return FDADI_OK;
}

int fdadi_hb_launches_per_step(const fdadi_hb *hb, int *n) {
  FDADI_REQUIRE(hb && n, FDADI_ERR_ARG, "fdadi_hb_launches_per_step: bad arguments");
  // one launch per group of each sweep stage
  const int pair = (int)(hb->sgroups.size() + hb->vgroups.size());
  switch (hb->scheme) {
    case FDADI_SCHEME_DOUGLAS:
    case FDADI_SCHEME_IMPLICIT: *n = pair; break;
    default: *n = 2 * pair; break;
  }
  return FDADI_OK;
}

}  // extern "C"
