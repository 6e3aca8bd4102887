/* C ABI of libisogrid_hip.so — the MI355X (gfx950) kernels of the IsoGridDFT
 * Kohn-Sham inner loop.
 *
 * Conventions
 *   - Every pointer is a device pointer (HBM) unless stated otherwise; sizes are
 *     element counts; fields are C order (nx, ny, nz), z fastest.
 *   - Orbital blocks are orbital-major: orbital m occupies [m*ld, m*ld + N).
 *   - `stream` is a hipStream_t; every call is asynchronous on that stream.
 *   - Return value: 0 ok, -1 invalid argument, -2 unsupported, -3 HIP launch error.
 *     The Python layer maps -1/-2 to ValueError and -3 to RuntimeError, as the
 *     reference raises ValueError for invalid shapes/enums (ops/kinetic.py:41-58).
 *   - Output buffers must not alias inputs unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference's src/isogrid):
 *   see the comment above each entry point.
 */
#ifndef ISOGRID_HIP_H
#define ISOGRID_HIP_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ stencils */

/* Monitor-grid curvilinear operator on a block of nblk fields:
 *   out = scale * Lap(in) + veff * in,   Lap = J^-1 sum_a D_a c^{ab} D_b,
 * c^{ab} = J g^{ab} given as coef[6*N] = (cxx, cyy, czz, cxy, cxz, cyz), invj[N] = 1/J,
 * h* the uniform logical spacings, closure 0 = zero-ghost ("trial_fix"),
 * 1 = one-sided edge_order=2 ("production"); transpose=1 applies the exact
 * Euclidean transpose Lap^T (then veff multiplies the untransformed input);
 * interior_only=1 writes zero on boundary nodes; xchunk = planes per workgroup
 * along x (0 = all); tile = 0 (auto), 32 or 64 (z tile width).  veff may be NULL.
 * Replaces: ops/kinetic.py:212-358 (compute_monitor_grid_contravariant_flux_components,
 * apply_monitor_grid_laplacian_operator[_trial_boundary_fix],
 * apply_monitor_grid_kinetic_operator[_trial_boundary_fix]) with scale = -1/2,
 * ops/kinetic_jax.py:74-175, the block kernel ks/eigensolver_jax_cache.py:38-63
 * (with veff = V_eff), ks/eigensolver.py:729-784, and the Poisson matvec
 * poisson/poisson_jax.py:140-178 / open_boundary.py:1032-1038 (scale = -1,
 * interior_only = 1). */
int ig_monitor_stencil_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                           int nz, double hx, double hy, double hz, const double* coef,
                           const double* invj, const double* veff, double scale, int closure,
                           int transpose, int interior_only, int xchunk, int tile,
                           hipStream_t stream);
int ig_monitor_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                           int nz, double hx, double hy, double hz, const float* coef,
                           const float* invj, const float* veff, double scale, int closure,
                           int transpose, int interior_only, int xchunk, int tile,
                           hipStream_t stream);

/* Legacy separable 7-point operator:
 *   out = scale * sum_axes [a+ (f[i+1]-f[i]) - a- (f[i]-f[i-1])] + veff * in,
 * zero ghosts; coeffs = (a-_x[nx], a+_x[nx], a-_y[ny], a+_y[ny], a-_z[nz], a+_z[nz]).
 * Replaces: ops/kinetic.py:87-153 (apply_legacy_laplacian_operator /
 * apply_legacy_kinetic_operator, scale -1/2) and the legacy Poisson interior operator
 * poisson/open_boundary.py:990-1014, :1216-1226 (scale -1, interior_only 1). */
int ig_legacy_stencil_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                          int nz, const double* coeffs, const double* veff, double scale,
                          int transpose, int interior_only, hipStream_t stream);
int ig_legacy_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                          int nz, const float* coeffs, const float* veff, double scale,
                          int transpose, int interior_only, hipStream_t stream);

/* ------------------------------------------------------------------ reductions / Krylov */

/* Number of doubles of `work` needed by ig_dots_* / ig_bicgstab_x_* for length n. */
int ig_reduction_work_size(long long n);

/* out[o_i] = sum_t a_i[t] b_i[t] (w[t] if w != NULL), i < nd <= 3, FP64 accumulation.
 * Replaces: ops/kinetic.py:61-74 (integrate_field, weighted_l2_norm),
 * ops/reductions_jax.py:19-38, and the CG dot products of poisson/poisson_jax.py:779-895. */
int ig_dots_f64(int nd, const double* a0, const double* b0, const double* a1, const double* b1,
                const double* a2, const double* b2, const double* w, long long n, double* work,
                double* out, int o0, int o1, int o2, hipStream_t stream);
int ig_dots_f32(int nd, const float* a0, const float* b0, const float* a1, const float* b1,
                const float* a2, const float* b2, const float* w, long long n, double* work,
                double* out, int o0, int o1, int o2, hipStream_t stream);

/* Fused preconditioned-BiCGSTAB vector updates with device-resident scalars sc[9]
 * (rho, rho_prev, alpha, omega, (rhat,v), (t,s), (t,t), (r,r), breakdown flag);
 * dinv = Jacobi preconditioner (NULL = none).
 * Replaces: the SciPy bicgstab solves of poisson/open_boundary.py:1017-1064 (monitor)
 * and :1229-1276 (legacy, Jacobi-preconditioned). */
int ig_bicgstab_p_f64(long long n, const double* r, double* p, const double* v, const double* dinv,
                      double* phat, const double* sc, hipStream_t stream);
int ig_bicgstab_p_f32(long long n, const float* r, float* p, const float* v, const float* dinv,
                      float* phat, const double* sc, hipStream_t stream);
int ig_bicgstab_s_f64(long long n, const double* r, const double* v, const double* dinv,
                      double* s, double* shat, double* sc, hipStream_t stream);
int ig_bicgstab_s_f32(long long n, const float* r, const float* v, const float* dinv, float* s,
                      float* shat, double* sc, hipStream_t stream);
int ig_bicgstab_x_f64(long long n, double* x, const double* phat, const double* shat,
                      const double* s, const double* t, double* r, const double* rhat, double* sc,
                      double* work, hipStream_t stream);
int ig_bicgstab_x_f32(long long n, float* x, const float* phat, const float* shat, const float* s,
                      const float* t, float* r, const float* rhat, double* sc, double* work,
                      hipStream_t stream);

/* ------------------------------------------------------------------ Gram products */

/* Doubles of `work` required by ig_gram_* for an (m x n) product over N points. */
long long ig_gram_work_size(int m, int n, long long N);

/* G[m x n] (=|+=) X W Y^T, X (m, N) and Y (n, N) orbital-major, w[N]; FP32 inputs on
 * MFMA (v_mfma_f32_32x32x2f32) with FP64 split-K reduction; FP64 inputs on FP64 FMAs.
 * Replaces: ks/eigensolver.py:246-272 (weighted_overlap_matrix),
 * ops/reductions_jax.py:89-104, :150-160, ks/eigensolver_jax_cache.py:95-113 and the
 * projected-matrix builds of ks/eigensolver_jax.py:104-105, :163-176. */
int ig_gram_f32(const float* X, long long ldx, int m, const float* Y, long long ldy, int n,
                const float* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t stream);
int ig_gram_f64(const double* X, long long ldx, int m, const double* Y, long long ldy, int n,
                const double* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ISOGRID_HIP_H */
