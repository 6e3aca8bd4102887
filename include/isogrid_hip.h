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
 * c^{ab} = J g^{ab}, D the centred first difference (1/2h) on the uniform logical axes.
 * coef[6*N] holds the pre-scaled coefficients c^{ab} / (4 h_a h_b) in the order
 * (xx, yy, zz, xy, xz, yz); invj[N] = 1/J.  closure 0 = zero-ghost ("trial_fix"),
 * 1 = one-sided edge_order=2 ("production"); transpose=1 applies the exact Euclidean
 * transpose Lap^T (then veff multiplies the untransformed input); interior_only=1
 * writes zero on boundary nodes; xchunk = planes per workgroup along x (0 =
 * automatic, >= nx = no split); tile = 0 (automatic), 32 / 64 (z tile width) or
 * NB*10000 + TZ*100 + TY (NB orbitals per workgroup, TY x TZ tile).  veff may be NULL.
 * Replaces: ops/kinetic.py:212-358 (compute_monitor_grid_contravariant_flux_components,
 * apply_monitor_grid_laplacian_operator[_trial_boundary_fix],
 * apply_monitor_grid_kinetic_operator[_trial_boundary_fix]) with scale = -1/2,
 * ops/kinetic_jax.py:74-175, the block kernel ks/eigensolver_jax_cache.py:38-63
 * (with veff = V_eff), ks/eigensolver.py:729-784, and the Poisson matvec
 * poisson/poisson_jax.py:140-178 / open_boundary.py:1032-1038 (scale = -1,
 * interior_only = 1). */
int ig_monitor_stencil_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                           int nz, const double* coef, const double* invj, const double* veff,
                           double scale, int closure, int transpose, int interior_only,
                           int xchunk, int tile, hipStream_t stream);
int ig_monitor_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                           int nz, const float* coef, const float* invj, const float* veff,
                           double scale, int closure, int transpose, int interior_only,
                           int xchunk, int tile, hipStream_t stream);

/* The trial_fix forward operator above plus, from the same pass, the row dots
 *   dots[m] = sum_n w[n] in[m][n] out[m][n]   (FP64, fixed summation order)
 * of a CG step's (p, A p) products; work holds ig_monitor_stencil_dots_work_size(nblk,
 * nx, ny, nz) doubles of partial sums.  Returns -2 where the full-z kernel does not apply
 * (z rows wider than one wavefront: FP32 nz > 256, FP64 nz > 128; nz or ld not a multiple
 * of the 16-byte lane width; misaligned operands): the caller then applies the operator
 * and takes the dots separately.  Replaces: the kinetic applies of the preconditioned
 * iteration followed by weighted inner products, ops/kinetic.py:349-358 with
 * ops/reductions_jax.py:89-147 (the weighted inner product). */
long long ig_monitor_stencil_dots_work_size(int nblk, int nx, int ny, int nz);
int ig_monitor_stencil_dots_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                                int nz, const double* coef, const double* invj, const double* veff,
                                double scale, const double* w, double* work, long long work_elems,
                                double* dots, hipStream_t stream);
int ig_monitor_stencil_dots_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                                int nz, const float* coef, const float* invj, const float* veff,
                                double scale, const float* w, double* work, long long work_elems,
                                double* dots, hipStream_t stream);

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

/* ig_dots_f64 without a workspace argument and without any allocation: the block partials
 * go to a 24 KiB file-scope device array of the library (one per GPU, kept until the process
 * ends, its address looked up once per device on the first call there).  Any n is accepted
 * (at most 1024 blocks x 3 partials).  Calls that overlap on one GPU (concurrent streams or
 * host threads) share that array and race; serialise them.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight loops, harnesses that forbid hipMalloc). */
int ig_dots_noalloc_f64(int nd, const double* a0, const double* b0, const double* a1, const double* b1,
                        const double* a2, const double* b2, const double* w, long long n, double* out,
                        int o0, int o1, int o2, hipStream_t stream);

/* Fused preconditioned-BiCGSTAB vector updates with device-resident scalars sc[10]
 * (rho, rho_prev, alpha, omega, (rhat,v), (t,s), (t,t), (r,r), breakdown flag, and the
 * x step's own breakdown slot, which the x entry folds into the flag; zero sc[9] with the
 * rest before a solve);
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

/* Multigrid V-cycle of the monitor-grid Poisson preconditioner (the right preconditioner
   of the BiCGSTAB above).  Replaces nothing one to one: the reference runs the same
   interior solve unpreconditioned (poisson/open_boundary.py:1017-1064) or with a line
   preconditioner on its JAX path (poisson/poisson_jax.py:411-579).  Coarse spaces follow
   the parity split of the 2h-wide stencil: level 1 aggregates the interior nodes of one
   parity in a 4x4x4 block (fine node i -> parity p = i & 1, aggregate (i - p) >> 2 per
   axis), deeper levels 2x2x2 points of one parity.  A level vector of extents (cx, cy,
   cz) is laid out [p][Sx][Sy][Sz] (p = 4 px + 2 py + pz); its operator is coef[7][m],
   m = 8 cx cy cz, {diagonal, -x, +x, -y, +y, -z, +z}.  Level-1 extents are
   ceil(n / 4), each further level ceil(c / 2); other extents return -1. */
/* bc = sum over the interior children of (b - t) (t may be NULL: sum of b). */
int ig_mg_fine_restrict_f64(const double* b, const double* t, int nx, int ny, int nz, double* bc, int cx,
                            int cy, int cz, hipStream_t stream);
int ig_mg_fine_restrict_f32(const float* b, const float* t, int nx, int ny, int nz, float* bc, int cx, int cy,
                            int cz, hipStream_t stream);
/* x[i] += ec[aggregate of i] on interior nodes. */
int ig_mg_fine_prolong_f64(double* x, const double* ec, int nx, int ny, int nz, int cx, int cy, int cz,
                           hipStream_t stream);
int ig_mg_fine_prolong_f32(float* x, const float* ec, int nx, int ny, int nz, int cx, int cy, int cz,
                           hipStream_t stream);
/* x += dw (b - t) over n points (fine damped Jacobi, dw = omega / diagonal, 0 on faces). */
int ig_mg_smooth_f64(double* x, const double* b, const double* t, const double* dw, long long n,
                     hipStream_t stream);
int ig_mg_smooth_f32(float* x, const float* b, const float* t, const float* dw, long long n, hipStream_t stream);
/* xo = xi + omega (b - A xi) / diag on one level (xi NULL: xo = omega b / diag; xo != xi). */
int ig_mg7_jacobi_f64(double* xo, const double* xi, const double* b, const double* coef, double omega, int cx,
                      int cy, int cz, hipStream_t stream);
int ig_mg7_jacobi_f32(float* xo, const float* xi, const float* b, const float* coef, double omega, int cx, int cy,
                      int cz, hipStream_t stream);
/* bc (level extents d) = sum over the 2x2x2 children of (b - A x) on level extents c
   (x NULL: sum of b). */
int ig_mg7_restrict_f64(double* bc, const double* b, const double* x, const double* coef, int cx, int cy, int cz,
                        int dx, int dy, int dz, hipStream_t stream);
int ig_mg7_restrict_f32(float* bc, const float* b, const float* x, const float* coef, int cx, int cy, int cz,
                        int dx, int dy, int dz, hipStream_t stream);
/* Hierarchy set-up (FP64): the 28 probe vectors that read the fine same-parity entries
   off the operator.  Probe q (0..27) = 7 pair + colour holds the interior nodes of parity
   class c = 4 px + 2 py + pz with min(c, 7 - c) = pair and ((ix >> 1) + 2 (iy >> 1) +
   3 (iz >> 1)) mod 7 = colour; probe_read stores out[i] into ent[k][i] (ent: 7 fields of
   nx ny nz) for every interior node i whose offset-k partner (0, -+2 e_x, -+2 e_y,
   -+2 e_z) is in probe q. */
int ig_mg_probe_vector_f64(double* e, int q, int nx, int ny, int nz, hipStream_t stream);
int ig_mg_probe_read_f64(double* ent, const double* out, int q, int nx, int ny, int nz, hipStream_t stream);
/* x[p][s] += ec[p][s >> 1]. */
int ig_mg7_prolong_f64(double* x, const double* ec, int cx, int cy, int cz, int dx, int dy, int dz,
                       hipStream_t stream);
int ig_mg7_prolong_f32(float* x, const float* ec, int cx, int cy, int cz, int dx, int dy, int dz,
                       hipStream_t stream);

/* ------------------------------------------------------------------ Gram products */

/* Doubles of `work` required by ig_gram_* for an (m x n) product over N points. */
long long ig_gram_work_size(int m, int n, long long N);

/* G[m x n] (=|+=) X W Y^T, X (m, N) and Y (n, N) orbital-major, w[N] (NULL: W = I); FP32 inputs on
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

/* Y[M x N] = Q[M x K] X[K x N] on MFMA (FP32 v_mfma_f32_32x32x2f32, FP64
 * v_mfma_f64_16x16x4f64): Q row-major (row stride ldq), X and Y row-major with row strides
 * ldx / ldy.  Y may overlap X only row-aligned (ldy == ldx, offset by whole rows), e.g.
 * Y = X for an in-place rotation of the first M rows.  M <= 256.  ct: column tiles per
 * wavefront (0 automatic, 1, 2).  Replaces the basis rotations of the block iteration:
 * ks/eigensolver_jax.py:184-201 (coefficients^T [basis; residuals]) and
 * ks/eigensolver.py:916-932. */
int ig_rotate_f32(const float* Q, int ldq, const float* X, long long ldx, float* Y, long long ldy, int M,
                  int K, long long N, int ct, hipStream_t stream);
int ig_rotate_f64(const double* Q, int ldq, const double* X, long long ldx, double* Y, long long ldy, int M,
                  int K, long long N, int ct, hipStream_t stream);

/* Doubles of `work` that ig_gemm_nt_f32/f64 need for A (m x K) and B (n x K): one FP64
 * (m x n) partial per K panel. */
long long ig_gemm_nt_work_size(int m, int n, long long K);

/* C[m x n] (FP64, row stride ldc) = A B^T on MFMA (v_mfma_f32_16x16x4f32 /
 * v_mfma_f64_16x16x4f64), split-K over the long axis K with a deterministic FP64 sum of
 * the panel partials: A (m x K, row stride lda) and B (n x K, row stride ldb) row-major
 * with unit column stride and 16-byte aligned rows (-2 otherwise), `work` at least
 * ig_gemm_nt_work_size(m, n, K) doubles.  Replaces the projected-matrix products of the
 * block iteration: ops/reductions_jax.py:157-159 (Gram blocks), ks/eigensolver_jax.py:
 * 108-146 (basis overlap for the orthonormalisation), ks/eigensolver.py:246-272. */
int ig_gemm_nt_f32(const float* A, long long lda, int m, const float* B, long long ldb, int n, long long K,
                   double* work, long long work_size, double* C, long long ldc, hipStream_t stream);
int ig_gemm_nt_f64(const double* A, long long lda, int m, const double* B, long long ldb, int n, long long K,
                   double* work, long long work_size, double* C, long long ldc, hipStream_t stream);


/* ------------------------------------------------------------------ physics fields */

/* Doubles of `work` needed by the reductions below (grid-size independent). */
int ig_physics_work_size(long long N);

/* Local GTH potential of natoms atoms on the grid points xyz[3*N] (SoA x|y|z):
 * atoms = natoms x 9 doubles (x, y, z, Z_ion, r_loc, C1, C2, C3, C4), device memory.
 * out (=|+=) sum_A [-Z erf(r / (sqrt2 r_loc)) / r + exp(-r^2/2r_loc^2) sum_i C_i (r/r_loc)^(2i-2)].
 * Replaces: pseudo/local.py:138-180 (_screened_coulomb_term, _local_polynomial_term),
 * :303-357 (evaluate_monitor_grid/legacy_local_ionic_potential). */
int ig_vloc_gth_f64(const double* xyz, long long N, int natoms, const double* atoms, double* out,
                    int accumulate, hipStream_t stream);
/* out[3*natoms] = dE/dR_A of E = sum_p q_p V_loc(|r_p - R_A|) (q = rho * w), natoms <= 64.
 * No reference counterpart (the reference has no forces); used by the adjoint path. */
int ig_vloc_gth_grad_f64(const double* xyz, const double* q, long long N, int natoms,
                         const double* atoms, double* work, double* out, hipStream_t stream);

/* Spin-polarised Slater exchange + PW92 correlation per point: eps_xc (per electron),
 * v_up, v_down; *flag |= 1 if a density is below -1e-14 or not finite (the reference
 * raises ValueError there, xc/lsda.py:57-67).
 * Replaces: xc/lsda.py:70-129 (evaluate_lsda_terms; the reference's functional is
 * libxc 'lda,vwn', this kernel evaluates PW92). */
int ig_lsda_pw92_f64(const double* rho_up, const double* rho_down, long long N, double* eps,
                     double* v_up, double* v_down, int* flag, hipStream_t stream);
/* The same with VWN5 correlation (the functional behind the reference's libxc 'lda,vwn',
 * xc/lsda.py:22, :95-100): Vosko-Wilk-Nusair 1980 paramagnetic / ferromagnetic / spin-
 * stiffness fits with the VWN spin interpolation.  Replaces: xc/lsda.py:70-129. */
int ig_lsda_vwn_f64(const double* rho_up, const double* rho_down, long long N, double* eps,
                    double* v_up, double* v_down, int* flag, hipStream_t stream);

/* ------------------------------------------------------------------ nonlocal projectors */

/* sizeof of the projector descriptor {int l, m, i; double r_l, prefactor}. */
int ig_projector_desc_size(void);
/* P[nproj][box] = p_i^l(r) Y_lm(r) on the support box (deriv = 0) or its derivative with
 * respect to the atom position component deriv-1 (deriv = 1..3).  box (host pointer) =
 * {i0, j0, k0, ni, nj, nk}; atom_pos (host) = 3 doubles; descs (device) = nproj
 * descriptors.  Replaces: pseudo/nonlocal.py:99-206 (real harmonics, radial projector,
 * evaluate_atomic_projector_field) restricted to the support box. */
int ig_projector_fields_f64(const double* xyz, int nx, int ny, int nz, const int* box,
                            const double* atom_pos, int nproj, const void* descs, double* P,
                            int deriv, hipStream_t stream);
/* Doubles of `work` needed by ig_nonlocal_project_*. */
int ig_nonlocal_work_size(int nblk, int nproj);
/* coef[m][p] = <p | psi_m>_W over the box; d[m][p] = sum_q H[p][q] coef[m][q] (coef may be
 * NULL).  box is a host pointer.  Replaces: pseudo/nonlocal.py:254-266 (projection
 * coefficients and h-matrix resolution). */
int ig_nonlocal_project_f64(const double* psi, long long ld, int nblk, int nx, int ny, int nz,
                            const int* box, const double* P, int nproj, const double* w,
                            const double* H, double* work, double* coef, double* d,
                            hipStream_t stream);
int ig_nonlocal_project_f32(const float* psi, long long ld, int nblk, int nx, int ny, int nz,
                            const int* box, const double* P, int nproj, const double* w,
                            const double* H, double* work, double* coef, double* d,
                            hipStream_t stream);
/* out_m += sum_p d[m][p] P[p] on the box (host pointer box).
 * Replaces: pseudo/nonlocal.py:267-273 (atom_action accumulation). */
int ig_nonlocal_apply_f64(double* out, long long ld, int nblk, int nx, int ny, int nz,
                          const int* box, const double* P, int nproj, const double* d,
                          hipStream_t stream);
int ig_nonlocal_apply_f32(float* out, long long ld, int nblk, int nx, int ny, int nz,
                          const int* box, const double* P, int nproj, const double* d,
                          hipStream_t stream);

/* ------------------------------------------------------------------ orbital density */

/* rho (=|+=) sum_m occ[m] psi_m^2 over an orbital-major block (ld = orbital stride),
 * FP64 accumulation; occ is a device array of nblk doubles.
 * Replaces: scf/driver.py:663-676 (_build_density_from_occupied_orbitals),
 * ks/static_hamiltonian.py:199-235 (build_orbital_density / build_total_density). */
int ig_density_f64(const double* psi, long long ld, int nblk, long long N, const double* occ, double* rho,
                   int accumulate, hipStream_t stream);
int ig_density_f32(const float* psi, long long ld, int nblk, long long N, const double* occ, double* rho,
                   int accumulate, hipStream_t stream);

/* ------------------------------------------------------------------ multipoles */

/* Raw Cartesian moments sum_p rho w [frac] x^a y^b z^c of degree <= order (<= 4) about
 * center (host pointer, 3 doubles); out = nmono(order) doubles in the order
 * 1; x y z; xx xy xz yy yz zz; xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz; xxxx ... zzzz.
 * frac may be NULL.  Replaces: poisson/open_boundary.py:223-264, :267-296, :816-861. */
int ig_multipole_moments_f64(const double* rho, const double* w, const double* frac,
                             const double* xyz, long long N, const double* center, int order,
                             double* work, double* out, hipStream_t stream);
/* out = boundary potential sum_l Q_l(r)/r^(2l+1) on the six faces, 0 inside; qcoef
 * (device) holds the homogeneous polynomial coefficients of degree 0..order over the same
 * monomial order.  Replaces: poisson/open_boundary.py:467-495, :888-986. */
int ig_boundary_values_f64(const double* xyz, int nx, int ny, int nz, const double* center,
                           int order, const double* qcoef, double* out, hipStream_t stream);
/* Per selected cell (cells = ncells x (i, j, k), device int32): quadratic least-squares fit
 * of rho on the 4x4x4 node stencil -> coef[14] per cell (10 coefficients in centred,
 * scaled coordinates, the centroid and the inverse scale).
 * Replaces: poisson/open_boundary.py:299-342 (_cell_and_neighbor_slices,
 * _local_quadratic_fit_coefficients). */
int ig_cell_fit_f64(const double* rho, const double* xyz, int nx, int ny, int nz, const int* cells,
                    int ncells, double* coef, hipStream_t stream);
/* Raw moments (<= order) of the fitted density sampled at the 2x2x2 sub-cell midpoints,
 * weight rho_fit * J_trilinear * subvol.  Replaces: poisson/open_boundary.py:396-458. */
int ig_cell_moments_f64(const double* xyz, const double* jac, int nx, int ny, int nz,
                        const int* cells, int ncells, const double* coef, double subvol,
                        const double* center, int order, double* work, double* out,
                        hipStream_t stream);
/* Split correction of the monitor `legacy_split` boundary mode: with
 * d = (rho_fit - rho_trilinear) J_trilinear subvol at the 2x2x2 sub-cell midpoints of the
 * selected cells, center == NULL gives out[4] = [sum |d|, sum |d| x, sum |d| y,
 * sum |d| z] (the correction centre), otherwise the raw moments of d about center
 * (host pointer) up to order.  Replaces: poisson/open_boundary.py:559-719
 * (_monitor_grid_boundary_value_correction). */
int ig_cell_split_moments_f64(const double* rho, const double* xyz, const double* jac, int nx, int ny,
                              int nz, const int* cells, int ncells, const double* coef, double subvol,
                              const double* center, int order, double* work, double* out,
                              hipStream_t stream);
/* Adjoint of the nodal multipole moments: out (=|+=) w (1 - frac) sum_j q_j mono_j(x - c)
 * (frac may be NULL; q: nmono(order) host doubles, center: host).  No reference
 * counterpart (the reference has no derivatives): the transpose of ig_multipole_moments_f64,
 * used by the Poisson adjoint. */
int ig_moments_adjoint_f64(const double* w, const double* frac, const double* xyz, long long N,
                           const double* center, int order, const double* q, double* out, int accumulate,
                           hipStream_t stream);
/* Adjoint of the sub-cell quadratic-fit moments (ig_cell_fit_f64 + ig_cell_moments_f64):
 * out += sum_j q_j d moment_j / d rho, scattered onto the 4x4x4 fit stencils (atomic adds).
 * No reference counterpart: the transpose of the corrected-moment construction
 * (poisson/open_boundary.py:370-464), used by the Poisson adjoint. */
int ig_cell_moments_adjoint_f64(const double* xyz, const double* jac, int nx, int ny, int nz,
                                const int* cells, int ncells, double subvol, const double* center,
                                int order, const double* q, double* out, hipStream_t stream);

/* ---- near-core local-GTH patch (pseudo/local.py:379-703) ------------------------- */

/* Bytes of device workspace ig_knn_f64 needs for nq queries and k neighbours (k <= 32),
 * or -1 for invalid sizes. */
long long ig_knn_work_bytes(int nq, int k);
/* k nearest grid nodes of each query point (queries = [3, nq] device SoA, xyz = [3, N]):
 * idx[nq, k] (int32) and Euclidean dist[nq, k], ascending by (distance, node index).
 * Replaces: pseudo/local.py:416-440 (cKDTree(point_cloud).query(patch_points, k)); the
 * neighbour sets depend on the geometry only and are computed once per patch. */
int ig_knn_f64(const double* xyz, long long N, const double* queries, int nq, int k, void* work,
               long long work_bytes, int* idx, double* dist, hipStream_t stream);
/* values[t] = inverse-distance-squared interpolation of field at query t (exact hit when
 * dist[t, 0] <= 1e-12, weights 1 / max(d, 1e-12)^2 normalised), values may be NULL;
 * out[0] = sum_t values[t] vq[t].  work: ig_physics_work_size doubles.
 * Replaces: pseudo/local.py:442-458 and the patch quadrature of :489-492. */
int ig_idw_quadrature_f64(const double* field, const int* idx, const double* dist, int nq, int k,
                          const double* vq, double* values, double* work, double* out,
                          hipStream_t stream);
/* Main-grid quadrature over the sphere |r - R_A| <= radius of one atom (atom: host pointer
 * to 9 doubles x, y, z, Z_ion, r_loc, C1..C4): out[3] = {sum rho V_A w, sum rho w, node count}.
 * Replaces: pseudo/local.py:494-510 and the density weight of :572-581. */
int ig_patch_main_f64(const double* xyz, const double* rho, const double* w, long long N,
                      const double* atom, double radius, double* work, double* out,
                      hipStream_t stream);
/* out += scale * V_A (constant == 0) or out += scale (constant != 0) on the nodes with
 * |r - R_A| <= radius (atom: host pointer as above).  Replaces: pseudo/local.py:546-582
 * (_build_atomic_frozen_patch_correction_field). */
int ig_patch_field_f64(const double* xyz, long long N, const double* atom, double radius,
                       int constant, double scale, double* out, hipStream_t stream);

/* ---- monitor-grid generator (grid/monitor_geometry.py:51-577) --------------------- */

/* Doubles of `work` needed by the sweep and Jacobian reductions below. */
int ig_grid_work_size(void);
/* Atom-centred monitor 1 + sum_A cap (1 - exp(-raw_A / cap)) on the points xyz[3*N] (SoA);
 * atoms = natoms x 13 doubles (x, y, z, kinetic weight / radius / exponent, local weight /
 * radius / exponent, projector weight / radius / exponent, cap), device memory.
 * Replaces: grid/monitor_geometry.py:51-112 (evaluate_global_monitor_field, values). */
int ig_monitor_field_f64(const double* xyz, long long N, const double* atoms, int natoms, double* out,
                         hipStream_t stream);
/* out = max((1 - s) m + s avg7(m), 1) inside, max(m, 1) on the faces (s > 0); out = m for
 * s <= 0.  out must not alias m.  Replaces: grid/monitor_geometry.py:115-136. */
int ig_monitor_smooth_f64(const double* mon, int nx, int ny, int nz, double smoothing, double* out,
                          hipStream_t stream);
/* One damped Jacobi sweep of div(M grad X) = 0 for the SoA coordinates (faces copied):
 * out = (1 - relax) X + relax X_jacobi inside; *max_change (device) = max |out - X| over
 * the interior.  dx, dy, dz: logical spacings.  Replaces: one inner iteration of
 * grid/monitor_geometry.py:268-321 (_solve_weighted_harmonic_coordinates). */
int ig_harmonic_sweep_f64(const double* mon, const double* coords, double* out, int nx, int ny, int nz,
                          double dx, double dy, double dz, double relax, double* work, double* max_change,
                          hipStream_t stream);
/* *min_jac (device) = min det(d x / d xi) of the SoA coordinates (np.gradient, edge order 2).
 * Replaces: the check of grid/monitor_geometry.py:376-396 (_backtracking_update). */
int ig_mapping_min_jacobian_f64(const double* coords, int nx, int ny, int nz, double hx, double hy,
                                double hz, double* work, double* min_jac, hipStream_t stream);
/* Jacobian, inverse metric tensor (9 per point, row-major) and cell volumes J hx hy hz of a
 * mapped grid.  Replaces: grid/monitor_geometry.py:324-373. */
int ig_mapping_metric_f64(const double* coords, int nx, int ny, int nz, double hx, double hy, double hz,
                          double* jacobian, double* inverse_metric, double* cell_volumes,
                          hipStream_t stream);

/* ------------------------------------------------------------------ row-wise block CG */

/* Doubles of `work` needed by the row-wise entries below that return sums, for k rows. */
int ig_row_dots_work_size(int k);
/* out[i] = sum_n w[n] a[i][n] b[i][n] (w NULL: 1) for i < k, rows of length N at stride ld;
 * FP64 accumulation, deterministic.  No reference counterpart: the reference's block
 * iteration is unpreconditioned; these carry the kinetic preconditioner added here
 * (k conjugate-gradient solves of (T + c) x = r, one per residual row). */
int ig_row_dots_f64(const double* a, const double* b, const double* w, int k, long long N, long long ld,
                    double* work, double* out, hipStream_t stream);
int ig_row_dots_f32(const float* a, const float* b, const float* w, int k, long long N, long long ld,
                    double* work, double* out, hipStream_t stream);
/* Row i: x += alpha[i] p, r -= alpha[i] q, out[i] = sum_n w[n] dinv[n] r[i][n]^2 (alpha and
 * out FP64 device vectors of length k).  No reference counterpart (see above). */
int ig_cg_update_f64(double* x, double* r, const double* p, const double* q, const double* dinv, const double* w,
                     const double* alpha, int k, long long N, long long ld, double* work, double* out,
                     hipStream_t stream);
int ig_cg_update_f32(float* x, float* r, const float* p, const float* q, const float* dinv, const float* w,
                     const double* alpha, int k, long long N, long long ld, double* work, double* out,
                     hipStream_t stream);
/* Row i: p = dinv r + beta[i] p.  No reference counterpart (see above). */
int ig_cg_direction_f64(double* p, const double* r, const double* dinv, const double* beta, int k, long long N,
                        long long ld, hipStream_t stream);
int ig_cg_direction_f32(float* p, const float* r, const float* dinv, const double* beta, int k, long long N,
                        long long ld, hipStream_t stream);
/* Start of the row-wise CG from x0 = dinv isw b with q = A x0 already applied: row i:
 * r = isw b - q, p = dinv r, out[i] = sum_n w[n] r[i][n] p[i][n].  b has row stride ldb,
 * the other blocks ld; isw, dinv, w hold N values.  No reference counterpart (see above). */
int ig_cg_start_f64(const double* b, long long ldb, const double* isw, const double* q, const double* dinv,
                    const double* w, double* r, double* p, int k, long long N, long long ld, double* work,
                    double* out, hipStream_t stream);
int ig_cg_start_f32(const float* b, long long ldb, const float* isw, const float* q, const float* dinv,
                    const float* w, float* r, float* p, int k, long long N, long long ld, double* work, double* out,
                    hipStream_t stream);
/* Last CG step mapped back: row i: y = sw (x + alpha[i] p), y at row stride ldy (x is not
 * written).  No reference counterpart (see above). */
int ig_cg_finish_f64(const double* x, const double* p, const double* sw, const double* alpha, double* y,
                     long long ldy, int k, long long N, long long ld, hipStream_t stream);
int ig_cg_finish_f32(const float* x, const float* p, const float* sw, const double* alpha, float* y, long long ldy,
                     int k, long long N, long long ld, hipStream_t stream);

/* ------------------------------------------------------------------ Green identity */

/* Doubles of `work` needed by ig_green_identity_f64. */
int ig_green_identity_work_size(void);
/* Discrete Green-identity terms of a real field f on a monitor grid; out[4] =
 *   { <f, tf>_W,  1/2 sum w d_a f g^{ab} d_b f,  boundary flux term,  <f, f>_W }.
 * tf = T f from any closure (NULL: out[0] = 0); coef = c^{ab} = J g^{ab} [6, N] in the
 * order xx, yy, zz, xy, xz, yz; invj = 1/J; w = cell volumes.  The gradient form uses
 * np.gradient(edge_order=2) derivatives; the boundary term sums +-1/2 f F^a dS over the
 * face nodes that lie on no edge, with F^a = c^{ab} d_b f from `closure` (0 zero-ghost
 * centred, 1 np.gradient edge_order 2).  Deterministic fixed-order reduction.
 * Replaces: audit/h2_monitor_grid_kinetic_green_identity_audit.py:143-218 (gradient
 * density, boundary proxy) and :295-330 (operator / gradient / boundary energies). */
int ig_green_identity_f64(const double* f, const double* tf, const double* coef, const double* invj,
                          const double* w, int nx, int ny, int nz, double hx, double hy, double hz, int closure,
                          double* work, double* out, hipStream_t stream);

/* ------------------------------------------------------------------ calibration */

/* Streaming kernels that measure the HBM bandwidth a plain 128-bit copy / read gets on
 * this GPU (the practical ceiling quoted next to the stencil's) and scale the PMC byte
 * counters.  No reference counterpart: the reference has no GPU code.
 * ig_stream_copy_f32: out[i] = in[i], i < n, n a multiple of 4, 16-byte aligned.
 * ig_stream_read_f32: partial[b] = sum over workgroup b's range of in; partial holds
 * ig_stream_blocks(n) doubles. */
int ig_stream_blocks(long long n);
int ig_stream_copy_f32(const float* in, float* out, long long n, hipStream_t stream);
int ig_stream_read_f32(const float* in, long long n, double* partial, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ISOGRID_HIP_H */
