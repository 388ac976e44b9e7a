/*
 * fdadi.h -- C ABI of the MI355X-native ADI finite-difference engine.
 *
 * Every function returns FDADI_OK (0) or a nonzero status; the message of the
 * last failure on the calling thread is returned by fdadi_last_error().
 * Pointers named d_* / *_dev are device pointers; all others are host
 * pointers.  `stream` is a hipStream_t passed as void* (NULL = the library's
 * default stream, which is the null stream).  Launches are asynchronous;
 * only the *_d2h / *_read functions synchronise.  All arithmetic is FP64.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference checkout, FiniteDifference/ omitted).
 */
#ifndef FDADI_H
#define FDADI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FDADI_OK = 0,
  FDADI_ERR_ARG = 1,          /* bad argument / size mismatch (reference: DIE -> domain_error) */
  FDADI_ERR_STATE = 2,        /* operation not allowed in this state (folded, mixed, ...) */
  FDADI_ERR_HIP = 3,          /* a HIP runtime call failed */
  FDADI_ERR_UNSUPPORTED = 4   /* configuration not supported */
};

/* Fold states: common.h:34-36 FOLDED / CAN_FOLD / CANNOT_FOLD strings. */
enum { FDADI_CANNOT_FOLD = 0, FDADI_CAN_FOLD = 1, FDADI_FOLDED = 2 };

/* Schemes of the fused engine (FiniteDifferenceEngine.pyx:616-1016). */
enum {
  FDADI_SCHEME_IMPLICIT = 0,  /* solve_implicit            :616-659  */
  FDADI_SCHEME_DOUGLAS = 1,   /* solve_douglas             :944-1006 */
  FDADI_SCHEME_CS = 2,        /* solve_craigsneyd          :842-898  */
  FDADI_SCHEME_MCS = 3,       /* solve_craigsneyd2         :778-839  */
  FDADI_SCHEME_HV = 4         /* solve_hundsdorferverwer   :699-775  */
};

/* ---- runtime ------------------------------------------------------------- */
/* Replaces the DIE()/std::domain_error convention, common.h:17-27. */
const char *fdadi_last_error(void);
int fdadi_version(int *major, int *minor);
int fdadi_device_count(int *n);
int fdadi_set_device(int dev);
int fdadi_stream_create(void **stream);
int fdadi_stream_destroy(void *stream);
int fdadi_stream_sync(void *stream);
int fdadi_device_sync(void);

/* ---- device arrays: SizedArray<double>, VecArray.h:211-710 ----------------- */
/* SizedArray ctors / device_malloc, VecArray.h:232-361. */
int fdadi_array_alloc(int64_t n, double **d_out);
int fdadi_array_free(double *d_ptr);
/* to_SizedArray H2D copy, SizedArrayPtr.pyx:320-326. */
int fdadi_array_h2d(double *d_dst, const double *src, int64_t n, void *stream);
/* from_SizedArray / to_numpy D2H copy, SizedArrayPtr.pyx:349-357 (synchronous). */
int fdadi_array_d2h(double *dst, const double *d_src, int64_t n, void *stream);
/* copy_from, VecArray.h:400-410. */
int fdadi_array_copy(double *d_dst, const double *d_src, int64_t n, void *stream);
int fdadi_array_fill(double *d_dst, double value, int64_t n, void *stream);
/* Elementwise a op= b: op 0 pluseq, 1 minuseq, 2 minuseq_over2 (a=(a-b)/2),
 * 3 timeseq (VecArray.cu:29-196). */
int fdadi_array_binop(double *d_a, const double *d_b, int64_t n, int op, void *stream);
/* Scalar variants (pluseq_scalar etc.), host value / device value
 * (VecArray.h:632-689 *_from_host and device-scalar forms). */
int fdadi_array_scalar_op(double *d_a, double s, int64_t n, int op, void *stream);
int fdadi_array_scalar_op_dev(double *d_a, const double *d_s, int64_t n, int op, void *stream);
/* transpose (h,w)->(w,h), VecArray.h:438-464 / _kernels.cu:7-37. */
int fdadi_array_transpose(double *d_dst, const double *d_src, int64_t h, int64_t w, void *stream);

/* ---- arena: one device allocation carved into many buffers --------------- */
/* Replaces the per-SizedArray device_malloc + tempspace pair (VecArray.h:237). */
typedef struct fdadi_arena fdadi_arena;
int fdadi_arena_create(int64_t bytes, fdadi_arena **out);
int fdadi_arena_alloc(fdadi_arena *a, int64_t n_doubles, double **d_out);
int fdadi_arena_used(const fdadi_arena *a, int64_t *bytes);
int fdadi_arena_reset(fdadi_arena *a);
int fdadi_arena_destroy(fdadi_arena *a);

/* ---- tridiagonal operator: _TriBandedOperator, _TriBandedOperatorGPU.cuh:17-87 */
typedef struct fdadi_tri fdadi_tri;
/* Constructor from host data, _TriBandedOperatorGPU.cuh:57-73 (emigrate_tri,
 * BandedOperatorGPU.pyx:264-314).  Bands are row-indexed (sup[i]=A[i,i+1],
 * sub[i]=A[i,i-1]), `rows` entries each; dirichlet/factor arrays have `blocks`
 * entries; deltas has ndeltas entries (may be 0). */
int fdadi_tri_create(const double *sup, const double *mid, const double *sub,
                     const double *R, const double *deltas, int64_t ndeltas,
                     const double *low_dirichlet, const double *high_dirichlet,
                     const double *top_factors, const double *bottom_factors,
                     int axis, int64_t rows, int64_t blocks, int has_low_dirichlet,
                     int has_high_dirichlet, int top_fold_status,
                     int bottom_fold_status, int has_residual, fdadi_tri **out);
/* for_vector, _TriBandedOperatorGPU.cu:1107-1182: Heston 1-D central stencil
 * with its boundary rows, built on the device from the host mesh. */
int fdadi_tri_for_vector(const double *mesh, int64_t n, int64_t blocks, int derivative,
                         int axis, int barrier, void *stream, fdadi_tri **out);
int fdadi_tri_copy(const fdadi_tri *src, void *stream, fdadi_tri **out);
int fdadi_tri_destroy(fdadi_tri *op);
int fdadi_tri_info(const fdadi_tri *op, int64_t *rows, int64_t *blocks, int *axis,
                   int *has_low_dirichlet, int *has_high_dirichlet, int *top_fold_status,
                   int *bottom_fold_status, int *has_residual, int64_t *ndeltas);
/* Public fields set from Python (has_low/high_dirichlet setters
 * BandedOperatorGPU.pyx:45-68, enable_residual :463-467). */
int fdadi_tri_set_flags(fdadi_tri *op, int has_low_dirichlet, int has_high_dirichlet,
                        int has_residual);
/* Public fields read back by immigrate_tri, BandedOperatorGPU.pyx:317-372.
 * Any output pointer may be NULL.  Synchronous. */
int fdadi_tri_read(const fdadi_tri *op, double *sup, double *mid, double *sub, double *R,
                   double *low_dirichlet, double *high_dirichlet, double *top_factors,
                   double *bottom_factors, double *deltas);
/* apply, _TriBandedOperatorGPU.cu:463-553: d_y = A x (+R) with the Dirichlet
 * values forced into x; the state is (nS,nv) row-major and the operator's axis
 * is addressed with strides (no transposes).  d_x is not modified; d_y must not
 * alias d_x.  A CAN_FOLD (unfolded) operator is applied with its +-2 entries. */
int fdadi_tri_apply(const fdadi_tri *op, const double *d_x, double *d_y, int64_t n,
                    void *stream);
/* solve, _TriBandedOperatorGPU.cu:719-770: d_v <- A^-1 (v - R) per grid line,
 * Dirichlet rows forced, RHS folded as needed (replaces
 * cusparseDgtsvStridedBatch and the transposes). */
int fdadi_tri_solve(const fdadi_tri *op, double *d_v, int64_t n, void *stream);
/* add_operator, _TriBandedOperatorGPU.cu:577-671. */
int fdadi_tri_add_operator(fdadi_tri *op, const fdadi_tri *other, void *stream);
/* add_scalar_from_host :674-689 and add_scalar (device value) :696-711. */
int fdadi_tri_add_scalar(fdadi_tri *op, double value, void *stream);
int fdadi_tri_add_scalar_dev(fdadi_tri *op, const double *d_value, void *stream);
/* vectorized_scale, _TriBandedOperatorGPU.cu:1012-1105: rows scaled by the
 * device vector tiled with period vsize (vsize divides rows; 1 = scalar). */
int fdadi_tri_scale(fdadi_tri *op, const double *d_vec, int64_t vsize, void *stream);
int fdadi_tri_scale_scalar(fdadi_tri *op, double value, void *stream);
/* fold_top / fold_bottom, _TriBandedOperatorGPU.cu:895-1009. */
int fdadi_tri_fold_top(fdadi_tri *op, int unfold, void *stream);
int fdadi_tri_fold_bottom(fdadi_tri *op, int unfold, void *stream);
/* diagonalize / undiagonalize, _TriBandedOperatorGPU.cu:836-867. */
int fdadi_tri_diagonalize(fdadi_tri *op, void *stream);
int fdadi_tri_undiagonalize(fdadi_tri *op, void *stream);

/* ---- cross-derivative operator: _CSRBandedOperator, _CSRBandedOperatorGPU.cuh:17-84 */
typedef struct fdadi_cross fdadi_cross;
/* mixed_for_vector, _CSRBandedOperatorGPU.cu:505-774: 9-point d2/dS dv. */
int fdadi_cross_for_vector(const double *mesh0, int64_t n0, const double *mesh1, int64_t n1,
                           void *stream, fdadi_cross **out);
/* Constructor from host data (emigrate_csr, BandedOperatorGPU.pyx:176-232):
 * bands[k*N + r] = A[r, r + off_k], off_k = a*n1 + b for a,b in (1,0,-1). */
int fdadi_cross_create(const double *bands, const double *R, int64_t n0, int64_t n1,
                       fdadi_cross **out);
int fdadi_cross_copy(const fdadi_cross *src, void *stream, fdadi_cross **out);
int fdadi_cross_destroy(fdadi_cross *op);
int fdadi_cross_info(const fdadi_cross *op, int64_t *n0, int64_t *n1);
int fdadi_cross_read(const fdadi_cross *op, double *bands, double *R);
/* apply, _CSRBandedOperatorGPU.cu:340-426 (replaces cusparseDcsrmv). */
int fdadi_cross_apply(const fdadi_cross *op, const double *d_x, double *d_y, int64_t n,
                      void *stream);
/* vectorized_scale, _CSRBandedOperatorGPU.cu:442-480. */
int fdadi_cross_scale(fdadi_cross *op, const double *d_vec, int64_t vsize, void *stream);
int fdadi_cross_scale_scalar(fdadi_cross *op, double value, void *stream);

/* ---- Heston coefficient fields: _coefficients.cu:122-338 --------------------
 * which: 0 = (0,) r*S, 1 = (0,0) 0.5*v*S^2, 2 = (1,) kappa*(theta-v)*[S!=0],
 * 3 = (1,1) 0.5*sigma^2*v*[S!=0], 4 = (0,1) rho*sigma*S*v.  Output ordered
 * with the derivative's axis fastest (spot-fastest for 0/1, var-fastest else).
 * p = {r, kappa, theta, sigma, rho}. */
int fdadi_heston_coefficients(int which, const double *d_spots, int64_t nS,
                              const double *d_vars, int64_t nv, const double *p,
                              double *d_out, void *stream);

/* ---- fused batched Heston engine ------------------------------------------------
 * One engine steps many options (uneven grids) with one launch per ADI stage;
 * it replaces the per-step operator-call loops of
 * FiniteDifferenceEngineGPU.pyx:379-593 and generates every coefficient and
 * boundary row on the device from the 1-D meshes. */
typedef struct fdadi_hb fdadi_hb;
typedef struct {
  double r, kappa, theta, sigma, rho; /* Heston parameters */
  double dt;                          /* step size of this option */
  int64_t nS, nv;                     /* grid size */
  int32_t barrier;                    /* nonzero: Dirichlet 0 at S = Smax (up-and-out) */
  int32_t pad;
} fdadi_heston_option;
/* spots/vars are the meshes of all options concatenated (host).  Each option
 * needs 4 <= nS <= 2048 and 4 <= nv <= 1024 (FDADI_ERR_UNSUPPORTED beyond). */
int fdadi_hb_create(int32_t nopt, const fdadi_heston_option *opts, const double *spots,
                    const double *vars, void *stream, fdadi_hb **out);
int fdadi_hb_destroy(fdadi_hb *hb);
/* Select the scheme, theta and a multiplier applied to every option's dt;
 * rebuilds the (time-invariant) factorisations on the device.  A graph
 * captured by fdadi_hb_capture survives a configure that keeps the scheme
 * (theta and dt are read from device memory at every step). */
int fdadi_hb_configure(fdadi_hb *hb, int scheme, double theta, double dt_scale,
                       void *stream);
/* fdadi_hb_configure with theta and the dt multiplier read from device
 * memory (one double each), as the reference's device variants take them
 * (SizedArrayPtr dt / theta, FiniteDifferenceEngineGPU.pyx:379,502,582);
 * asynchronous: nothing is read back to the host. */
int fdadi_hb_configure_dev(fdadi_hb *hb, int scheme, const double *d_theta,
                           const double *d_dt_scale, void *stream);
/* New parameters, step sizes and meshes for the engine's options, of the
 * same count and grid shapes as at creation (host arrays as for create); the
 * tables are rebuilt and, when configured, the factorisations too, so a
 * captured graph stays valid.  This is how one engine (and one graph) serves
 * batch after batch of the surrogate's fallback queries; the reference builds
 * a new engine per option (FiniteDifferenceEngineGPU.pyx:603-694, driven per
 * option by run_parameters.py:16-36).  Synchronises. */
int fdadi_hb_update(fdadi_hb *hb, int32_t nopt, const fdadi_heston_option *opts,
                    const double *spots, const double *vars, void *stream);
/* fdadi_hb_update from device memory, without synchronising: d_params holds
 * nopt rows (r, kappa, theta, sigma, rho, dt); d_spots / d_vars the packed
 * meshes (sp_n / vt_n doubles, same shapes), or NULL to keep the current
 * ones.  Barriers are cleared.  The labeller and the surrogate's fallback use
 * it to price a device-resident batch (no counterpart in the reference, which
 * builds a new engine per option from host values,
 * FiniteDifferenceEngineGPU.pyx:603-694). */
int fdadi_hb_update_dev(fdadi_hb *hb, const double *d_params, const double *d_spots,
                        const double *d_vars, void *stream);
/* European call payoff max(S - K_k, 0) on every option's grid (d_strikes:
 * device array of nopt strikes); the initial condition of
 * FiniteDifferenceEngineGPU.pyx:682. */
int fdadi_hb_call_payoff(fdadi_hb *hb, const double *d_strikes, void *stream);
/* Each option's state read at (S0_k, v0_k) (device arrays of nopt) by
 * tensor-product cubic Lagrange interpolation on the 4 x 4 nodes around the
 * point (stencil: nodes i-2 .. i+1 along each axis, i the first node >= the
 * coordinate, shifted inside the mesh) into d_out[nopt].  The reference
 * reads the value at a node instead (FiniteDifferenceEngineGPU.pyx:727-738);
 * the surrogate's labels need S0 and v0 off the nodes. */
int fdadi_hb_read_bicubic(const fdadi_hb *hb, const double *d_S0, const double *d_v0,
                          double *d_out, void *stream);
/* State of option k lives at d_state + offset[k] (layout (nS,nv) row-major). */
int fdadi_hb_state(fdadi_hb *hb, double **d_state, int64_t *total);
int fdadi_hb_offsets(const fdadi_hb *hb, int64_t *offsets);
/* Advance every option by nsteps steps of the configured scheme. */
int fdadi_hb_step(fdadi_hb *hb, int32_t nsteps, void *stream);
/* Advance nsteps steps with a HIP event pair around every ADI stage (the
 * launches of its sweep's groups) and return, per stage slot of a step
 * (launch order), the summed device time in milliseconds; *nstages = stages
 * per step (2 for Douglas and implicit, 4 otherwise).  Synchronises. */
int fdadi_hb_step_timed(fdadi_hb *hb, int32_t nsteps, void *stream, double *stage_ms,
                        int32_t max_stages, int32_t *nstages);
/* Capture nsteps steps into a HIP graph (replaces any previous one) and
 * replay it. */
int fdadi_hb_capture(fdadi_hb *hb, int32_t nsteps, void *stream);
int fdadi_hb_replay(fdadi_hb *hb, void *stream);
/* Capture one whole solve into a HIP graph (replaces any previous one; replay
 * with fdadi_hb_replay from the initial state): with smoothing = s > 0, first
 * 2s implicit steps of half the step (theta 1), then nsteps steps of
 * `scheme` with `theta`, each part preceded by its refactorisation -- the
 * sequence of the reference's smoothed solve
 * (FiniteDifferenceEngineGPU.pyx:565-593, solve_smooth_: 2s implicit half
 * steps, then HV with theta 0.6), without a host launch per stage or the
 * synchronisation of each configure.  Leaves the engine configured for
 * `scheme`. */
int fdadi_hb_capture_solve(fdadi_hb *hb, int scheme, double theta, int32_t smoothing,
                           int32_t nsteps, void *stream);
/* Launch shape per sweep kind, used by the autotuner (0 = library default):
 * kind 0 (spot sweep): `threads` per work-group (64..512) and `lines` =
 * variance columns per tile (4, 8, 16 or 32; 0 = 4; 8 and 16 need 256 or 512
 * threads, 32 needs 512 threads and at least 32 variance nodes in every
 * option);
 * kind 1 (variance sweep): `threads` = lanes per variance line (1..64, the
 * work-group has 256 threads), `lines` must be 0.
 * threads = -1 (either kind): a shape per option, the smallest whose chunks
 * cover its line (a batch of uneven grids then does no work for positions a
 * larger option's shape would add); options of one shape share a launch, so
 * a stage is one launch per shape in use.  get_shape reports threads (-1
 * when per option) and the columns (kind 0) / spot rows (kind 1, 0 when per
 * option) per tile.  Discards a captured graph. */
int fdadi_hb_set_shape(fdadi_hb *hb, int kind, int threads, int lines);
int fdadi_hb_get_shape(const fdadi_hb *hb, int kind, int *threads, int *lines);
/* Kernel launches per step: one per launch group of each stage's sweep (one
 * group unless shapes are chosen per option, fdadi_hb_set_shape). */
int fdadi_hb_launches_per_step(const fdadi_hb *hb, int *n);

/* ---- learned price surrogate ----------------------------------------------------
 * No counterpart in the reference (its only pricers are the ADI engines and
 * the COS method, heston.py:490-693).  An MLP 8 -> 512 x4 -> 1 (SiLU) maps
 * the 8 Heston features (kappa, theta, sigma, rho, v0, r, T, K/S0) to a
 * price; BF16 weights/activations, FP32 accumulation and master weights,
 * Adam.  Inputs are normalised with the box [lo, hi]; the target with
 * (tmean, tstd).  Parameter layout of the flat fp32 array: W0 [512 x 128]
 * (input zero-padded from 8 to 128 features), b0 [512], W1..W3 [512 x 512],
 * b1..b3, w4 [512], b4 [1]. */
typedef struct fdadi_mlp fdadi_mlp;
int fdadi_mlp_nparam(int64_t *n);
int fdadi_mlp_create(int32_t max_rows, const float *init_params, int64_t nparam,
                     const double *lo, const double *hi, double tmean, double tstd,
                     fdadi_mlp **out);
/* The same network on n_in (1..16) input features per row (d_P rows of n_in
 * doubles, lo/hi of n_in): the surrogate's error model reads the 8 features
 * plus derived ones.  fdadi_mlp_create is this with n_in = 8. */
int fdadi_mlp_create_features(int32_t n_in, int32_t max_rows, const float *init_params,
                              int64_t nparam, const double *lo, const double *hi, double tmean,
                              double tstd, fdadi_mlp **out);
int fdadi_mlp_destroy(fdadi_mlp *m);
int fdadi_mlp_get_params(const fdadi_mlp *m, float *params, float *grads);
int fdadi_mlp_set_params(fdadi_mlp *m, const float *params, void *stream);
/* Gradient of the mean squared (standardised) error over rows idx[0..rows)
 * of d_P [n x 8] / d_T [n] (idx NULL = the first rows); *loss if non-NULL. */
int fdadi_mlp_grad(fdadi_mlp *m, const double *d_P, const double *d_T, const int *d_idx,
                   int32_t rows, float *loss, void *stream);
/* grad + one Adam step (beta 0.9/0.999, eps 1e-8) at learning rate lr. */
int fdadi_mlp_train_step(fdadi_mlp *m, const double *d_P, const double *d_T, const int *d_idx,
                         int32_t rows, float lr, void *stream);
int fdadi_mlp_loss(const fdadi_mlp *m, float *loss);
/* Prices for d_P [n x 8] into d_out [n]. */
int fdadi_mlp_predict(fdadi_mlp *m, const double *d_P, int64_t n, double *d_out, void *stream);
/* fdadi_mlp_predict with the activations in a static device array instead
 * of the model's buffers: allocates nothing (no hipMalloc, no hipFree, no
 * allocating library call).  No counterpart in the reference: it is for
 * callers that cannot allocate -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation.  At most 4096 rows per call
 * (FDADI_ERR_ARG beyond; split the batch).  The array (9 MiB + 16 KiB) lives
 * on every GPU the library runs on until the process ends and is shared by
 * all calls on that GPU: calls that overlap on one GPU race. */
int fdadi_mlp_predict_noalloc(const fdadi_mlp *m, const double *d_P, int64_t n, double *d_out,
                              void *stream);
/* C (fp32) = A B^T with bf16 operands given as raw bits (test entry point);
 * M, N and K multiples of 128 (FDADI_ERR_ARG otherwise). */
int fdadi_mlp_gemm_nt(const uint16_t *d_A, const uint16_t *d_B, int32_t M, int32_t N, int32_t K,
                      float *d_C, void *stream);

/* ---- semi-analytic Heston call (csrc/cf.hip) ------------------------------------
 * Replaces the reference's semi-analytic checker HestonCos
 * (FiniteDifference/heston.py:490-693, used by
 * FiniteDifference/tests/test_HestonOptionGPU.py:74-112) with a device pricer:
 * Lewis's single-integral formula over the "little trap" characteristic
 * function, Gauss-Legendre quadrature (512 nodes on [0, a], 64 on the tail).
 * d_rows [n x 9] = S0, K, r, v0, T, kappa, theta, sigma, rho per option;
 * d_out [n] call prices (NaN for S0, K, T, sigma <= 0, v0 or theta < 0, or
 * |rho| >= 1).  Accurate to ~1e-9 for sigma >= 0.01; below sigma ~ 1e-3 the
 * kappa theta / sigma^2 terms cancel and accuracy is lost.  Asynchronous on
 * ``stream``; the node tables are copied to the current device on its first
 * call. */
int fdadi_heston_cf_call(const double *d_rows, int64_t n, double *d_out, void *stream);
/* Gauss-Legendre nodes x[n] (ascending) and weights w[n] on (-1, 1), host only
 * (the quadrature rule of fdadi_heston_cf_call). */
int fdadi_gauss_legendre(int32_t n, double *x, double *w);

/* ---- device-resident serving of the surrogate ----------------------------------
 * No counterpart in the reference (its pricers are the ADI engines and the COS
 * method, heston.py:490-693).  A batch of feature rows in device memory (8
 * doubles each: kappa, theta, sigma, rho, v0, r, T, K/S0; S0 = 100) is
 * answered without a host round trip; only the number of fallback rows comes
 * back.
 *
 * Regression trees (the gate's error model): ntrees trees, tree t holding
 * nodes [tree_start[t], tree_start[t+1]); per node the feature, threshold,
 * leaf flag, leaf value and the children as indices local to the tree (x goes
 * left when x[feature] <= threshold).  Children must come after their parent.
 * Prediction = base + the sum of the reached leaves, tree by tree in order
 * (the summation order of scikit-learn's gradient boosting). */
typedef struct fdadi_forest fdadi_forest;
int fdadi_forest_create(int32_t ntrees, const int64_t *tree_start, int32_t nfeat,
                        const int32_t *feature, const double *threshold, const int32_t *left,
                        const int32_t *right, const uint8_t *is_leaf, const double *value,
                        double base, fdadi_forest **out);
int fdadi_forest_destroy(fdadi_forest *f);
/* d_X: n rows of nfeat doubles; d_out[n]. */
int fdadi_forest_predict(const fdadi_forest *f, const double *d_X, int64_t n, double *d_out,
                         void *stream);
/* price = d_base + mean of the M member corrections d_Y (M x n, member-major);
 * d_X (n x 13) = the error model's features (log kappa, log theta, log sigma,
 * rho, log v0, r, log T, log K/S0, log Feller ratio 2 kappa theta / sigma^2,
 * log10(spread + 1e-12), mean, log10(|mean| + 1e-9), log10(max - min + 1e-12);
 * spread is the members' standard deviation with M - 1, mean, max and min
 * are over the members' corrections);
 * d_inside = row within the box [lo, hi] (host arrays of 8) with Feller ratio
 * >= feller_min. */
int fdadi_serve_combine(const double *d_P, int64_t n, const double *d_Y, int32_t M,
                        const double *d_base, const double *lo, const double *hi,
                        double feller_min, double *d_price, double *d_X, int32_t *d_inside,
                        void *stream);
/* Rows of fdadi_heston_cf_call (9 doubles each) from feature rows. */
int fdadi_serve_cf_rows(const double *d_P, int64_t n, double *d_rows, void *stream);
/* d_ok = inside && score <= tau; the rejected rows' indices in ascending order
 * into d_idx (room for n), their count into *n_rejected (synchronises).
 * d_work: ceil(n / 256) + 1 ints of scratch. */
int fdadi_serve_gate(const double *d_score, const int32_t *d_inside, int64_t n, double tau,
                     int32_t *d_ok, int32_t *d_idx, int32_t *d_work, int32_t *n_rejected,
                     void *stream);
/* Inputs of a fallback batch of cap options for rows d_idx[0..m) of d_P (d_idx
 * NULL: rows 0..m); rows m..cap-1 repeat the last one: d_params (cap x 6:
 * r, kappa, theta, sigma, rho, T / nt) for fdadi_hb_update_dev, d_spots (cap x
 * nS) = K * d_unit, the strikes, S0 and v0 of the payoff and the read-out. */
int fdadi_serve_label_inputs(const double *d_P, const int32_t *d_idx, int32_t m, int32_t cap,
                             int32_t nt, const double *d_unit, int32_t nS, double *d_params,
                             double *d_spots, double *d_K, double *d_S0, double *d_v0,
                             void *stream);
/* d_out[r] = max over m < k of d_S[m * n + r]: the largest of k error-model
 * estimates per row (the gate's score). */
int fdadi_serve_max_rows(const double *d_S, int32_t k, int64_t n, double *d_out, void *stream);
/* d_dst[d_idx[i]] = d_src[i], i < m. */
int fdadi_serve_scatter(const double *d_src, const int32_t *d_idx, int32_t m, double *d_dst,
                        void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FDADI_H */
