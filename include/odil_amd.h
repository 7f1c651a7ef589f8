/* C ABI of libodil_amd.so: MI355X (gfx950) kernels for the discrete-loss
 * hot path of ODIL -- fused finite-difference residuals, loss adjoints,
 * stencil Jacobians, matrix-free J^T J, and the geometric-multigrid
 * normal-equation solver of the Newton step.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a status code below and never
 *    throws; odil_last_error() returns the message of the last failure of
 *    the calling process.
 *  - All array arguments are device pointers owned by the caller, except
 *    the small host arrays documented as such (grid sizes `nw`, `reach`,
 *    `ent`/`shifts` outputs of *_jac_entries, `dq` and `ent_host` of the
 *    smoothers). No entry point allocates device memory.
 *  - Fields are C-order arrays of one grid (axis 0 slowest, e.g. t for the
 *    space-time operators, x for the cavity); "fields" arguments are arrays
 *    of per-field pointers (u, v, p for the cavity).
 *  - dtype selects the arithmetic: ODIL_F32 or ODIL_F64. Sums of squares
 *    and dot products are accumulated in FP64 in fixed order (deterministic).
 *  - `stream` is a hipStream_t (NULL = default stream). Work is queued
 *    asynchronously; reductions land in device memory.
 *
 * Reference paths are relative to the reference repository's code/
 * directory.
 */
#ifndef ODIL_AMD_H
#define ODIL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ODIL_F32 = 0, ODIL_F64 = 1 };
enum {
  ODIL_OK = 0,
  ODIL_ERR_ARG = 1,      /* invalid argument (Python: ValueError) */
  ODIL_ERR_LAUNCH = 2,   /* HIP launch/runtime failure (RuntimeError) */
  ODIL_ERR_CAPACITY = 3  /* size beyond a documented limit */
};

/* ---- runtime ----------------------------------------------------------- */

/* Message of the last failed call (thread-unsafe static buffer). Replaces
 * the Python exceptions the reference raises (util_op.py:539,
 * linsolver.py:105, multigrid.py:255,521). */
const char* odil_last_error(void);
/* Version of this ABI (incremented on incompatible changes). */
int odil_abi_version(void);
/* out[r] = sum_b partial[r * nb + b] for r < nrows (fixed order, FP64):
 * per-equation sums of squares from the per-block partials written by the
 * residual / loss_grad / jtf kernels (the sum inside mean(f^2) of
 * Problem._eval_loss, util_op.py:628-654). */
int odil_reduce_partials(const double* partial, int nrows, int nb,
                         double* out, void* stream);

/* ---- operators ----------------------------------------------------------
 * Per operator X in {cavity, wave, poisson, laplace, advection}:
 *  odil_X_residual   f = operator(fields) for every equation, plus per-block
 *                    partial sums of f^2 (`partial` [neq][nblocks(0)], may
 *                    be NULL; `res` may be NULL for the sums only).
 *                    Replaces Problem._eval_operator / eval_operator
 *                    (util_op.py:657-674,968-974) of the operator callback.
 *  odil_X_loss_grad  fused loss + adjoint: grad_k = sum_e (2/N) J_e,k^T f_e,
 *                    i.e. tape.gradient of loss = sum_e mean(f_e^2) with
 *                    frozen (stop-gradient) paths excluded, plus the
 *                    per-block f^2 partials [neq][nblocks(1)]. Replaces
 *                    Problem._eval_loss + eval_loss_grad (util_op.py:
 *                    628-654,939-966). `nblocks` (host, may be NULL)
 *                    receives the partial count.
 *  odil_X_jac        stencil Jacobian: coefficient fields
 *                    coef[q] = d f_e / d u_k(. + s) for the entries
 *                    (e, k, s) listed by odil_X_jac_entries, frozen paths
 *                    excluded, plus the residual (res may be NULL).
 *                    Replaces Problem._eval_grad0 (util_op.py:585-617, the
 *                    split_shifts tape Jacobian) feeding linearize
 *                    (util_op.py:859-937) and the SparseOperator rows
 *                    (util.py:544-566).
 *  odil_X_jac_entries host call: writes the entry table ent[3 q] =
 *                    (equation, field, shift index) and the shift table
 *                    shifts[2 s] (2D); returns the number of entries
 *                    (either array may be NULL to query the count).
 *  odil_X_nblocks    host call: partial-sum count of kind 0 (residual) or
 *                    1 (loss_grad / jtf) for the grid: one partial per
 *                    wavefront of the launch; for the cavity also kind 2
 *                    (normal_apply_dot): one per workgroup.
 *  odil_X_normal_apply  matrix-free N = J^T J with J recomputed from
 *                    `state`: mode 0 y = N x, 1 y = b - N x,
 *                    2 y = x + omega (b - N x) / d (weighted Jacobi),
 *                    3 y = diag(N). Replaces matr = sum_e mul_transpose_op
 *                    (util.py:578-585, multigrid.py:1035-1071) applied by
 *                    SparseOperator.mul_field / smoother_jacobi
 *                    (multigrid.py:958-969,543-552) on the finest level.
 *  odil_X_jtf        out = scale J^T f (scale = -1: the Newton right-hand
 *                    side rhs = -J^T const, util.py:572-577), plus f^2
 *                    partials.
 * ------------------------------------------------------------------------ */

/* Lid-driven cavity, steady incompressible Navier-Stokes on a collocated
 * grid with Rhie-Chow continuity and deferred-correction upwinding
 * (ns_steady.py:142-394): fields u, v, p on nx x ny cells, mu = 1/Re,
 * bc 0/1/2 and rhiechow as the reference flags. */
int odil_cavity_residual(int dtype, int nx, int ny, double dx, double dy,
                         double mu, int bc, int rhiechow,
                         const void* const* fields, void* const* res,
                         double* partial, void* stream);
int odil_cavity_jac(int dtype, int nx, int ny, double dx, double dy,
                    double mu, int bc, int rhiechow, const void* const* fields,
                    void* coef, void* const* res, void* stream);
int odil_cavity_jac_entries(int* ent, int* shifts);
int odil_cavity_loss_grad(int dtype, int nx, int ny, double dx, double dy,
                          double mu, int bc, int rhiechow,
                          const void* const* fields, void* const* grad,
                          double* partial, int* nblocks, void* stream);
int odil_cavity_nblocks(int kind, int dtype, int nx, int ny);
/* Allocation-free variant of odil_cavity_loss_grad. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). It takes no workspace: the
 * per-wavefront partial sums of f^2 live in a device array of the library
 * (3 x 65536 doubles per device, held until the process ends) and their
 * fixed-order sums go to ssq[0..2] (device, the values odil_reduce_partials
 * gives for odil_cavity_loss_grad's partials). Grids that need more than
 * 65536 partials per equation (square grids above 2022 x 2022 cells in FP64,
 * 2822 x 2822 in FP32) return ODIL_ERR_CAPACITY. Calls that overlap on one device race on the array. */
int odil_cavity_loss_grad_noalloc(int dtype, int nx, int ny, double dx,
                                  double dy, double mu, int bc, int rhiechow,
                                  const void* const* fields,
                                  void* const* grad, double* ssq,
                                  void* stream);
/* Extra loss terms of the cavity inverse problems (ns_steady.py:396-411),
 * appended to the cavity residual as six equations on (u, v, p):
 * mask (u - u_imp) cimp, mask (v - v_imp) cimp (imposed points, cimp =
 * alpha N / n_imp; mask/uimp/vimp NULL: no imposed points) and
 * kreg (u_xx, u_yy, v_xx, v_yy) (laplace_split of the halo-corrected live
 * stencils, ns_steady.py:213-217; kreg already annealed by the caller).
 * Residual, fused loss + adjoint (ADDS (2/N) J^T f to grad; partials as
 * odil_cavity_nblocks(0|1, ...) x 6) and Jacobian (14 coefficient fields,
 * entries from odil_cavity_extras_jac_entries). */
int odil_cavity_extras_residual(int dtype, int nx, int ny, double dx,
                                double dy, int bc, double kreg, double cimp,
                                const void* mask, const void* uimp,
                                const void* vimp, const void* const* fields,
                                void* const* res, double* partial,
                                void* stream);
int odil_cavity_extras_loss_grad(int dtype, int nx, int ny, double dx,
                                 double dy, int bc, double kreg, double cimp,
                                 const void* mask, const void* uimp,
                                 const void* vimp, const void* const* fields,
                                 void* const* grad, double* partial,
                                 int* nblocks, void* stream);
int odil_cavity_extras_jac(int dtype, int nx, int ny, double dx, double dy,
                           int bc, double kreg, double cimp, const void* mask,
                           const void* uimp, const void* vimp,
                           const void* const* fields, void* coef,
                           void* const* res, void* stream);
int odil_cavity_extras_jac_entries(int* ent, int* shifts);
int odil_cavity_extras_nblocks(int kind, int dtype, int nx, int ny);
int odil_cavity_normal_apply(int dtype, int nx, int ny, double dx, double dy,
                             double mu, int bc, int rhiechow,
                             const void* const* state, const void* const* x,
                             const void* const* b, const void* const* d,
                             void* const* y, double omega, int mode,
                             void* stream);
/* Mode 0 of odil_cavity_normal_apply, y = N x, and *dot = sum of x * y over
 * every cell and field (FP64, device scalar; partial holds
 * odil_cavity_nblocks(2, dtype, nx, ny) doubles, summed in a fixed order):
 * the product and the p.Np of a conjugate-gradient iteration in one launch.
 * No reference counterpart (its Krylov-accelerated solves are library
 * calls, pyamg with accel='cg', linsolver.py:77-88). */
int odil_cavity_normal_apply_dot(int dtype, int nx, int ny, double dx,
                                 double dy, double mu, int bc, int rhiechow,
                                 const void* const* state,
                                 const void* const* x, void* const* y,
                                 double* partial, double* dot, void* stream);
int odil_cavity_jtf(int dtype, int nx, int ny, double dx, double dy,
                    double mu, int bc, int rhiechow, const void* const* fields,
                    void* const* out, double scale, double* partial,
                    int* nblocks, void* stream);

/* 1D wave equation u_tt = u_xx on an nt x nx space-time grid
 * (wave.py:67-122, operator_fd). init_u, init_ut [nx] and left_u, right_u
 * [nt] are FP64 device arrays (initial value/velocity, boundary values). */
int odil_wave_residual(int dtype, int nt, int nx, double dt, double dx,
                       const double* init_u, const double* init_ut,
                       const double* left_u, const double* right_u,
                       const void* const* fields, void* const* res,
                       double* partial, void* stream);
int odil_wave_jac(int dtype, int nt, int nx, double dt, double dx,
                  const double* init_u, const double* init_ut,
                  const double* left_u, const double* right_u,
                  const void* const* fields, void* coef, void* const* res,
                  void* stream);
int odil_wave_jac_entries(int* ent, int* shifts);
int odil_wave_loss_grad(int dtype, int nt, int nx, double dt, double dx,
                        const double* init_u, const double* init_ut,
                        const double* left_u, const double* right_u,
                        const void* const* fields, void* const* grad,
                        double* partial, int* nblocks, void* stream);
int odil_wave_nblocks(int kind, int dtype, int nt, int nx);
int odil_wave_normal_apply(int dtype, int nt, int nx, double dt, double dx,
                           const double* init_u, const double* init_ut,
                           const double* left_u, const double* right_u,
                           const void* const* state, const void* const* x,
                           const void* const* b, const void* const* d,
                           void* const* y, double omega, int mode,
                           void* stream);
int odil_wave_jtf(int dtype, int nt, int nx, double dt, double dx,
                  const double* init_u, const double* init_ut,
                  const double* left_u, const double* right_u,
                  const void* const* fields, void* const* out, double scale,
                  double* partial, int* nblocks, void* stream);

/* Poisson residual f = lap u - src with Dirichlet zero through quadratic
 * half-cell extrapolation (the Laplacian of util_op.py:1058-1063 with
 * extrap_quadh walls, util_op.py:1017-1024; test_multigrid.py:275-296
 * poses the same problem). src has the field dtype. */
int odil_poisson_residual(int dtype, int n0, int n1, double dx, double dy,
                          const void* src, const void* const* fields,
                          void* const* res, double* partial, void* stream);
int odil_poisson_jac(int dtype, int n0, int n1, double dx, double dy,
                     const void* src, const void* const* fields, void* coef,
                     void* const* res, void* stream);
int odil_poisson_jac_entries(int* ent, int* shifts);
int odil_poisson_loss_grad(int dtype, int n0, int n1, double dx, double dy,
                           const void* src, const void* const* fields,
                           void* const* grad, double* partial, int* nblocks,
                           void* stream);
int odil_poisson_nblocks(int kind, int dtype, int n0, int n1);
int odil_poisson_normal_apply(int dtype, int n0, int n1, double dx,
                              double dy, const void* src,
                              const void* const* state, const void* const* x,
                              const void* const* b, const void* const* d,
                              void* const* y, double omega, int mode,
                              void* stream);
int odil_poisson_jtf(int dtype, int n0, int n1, double dx, double dy,
                     const void* src, const void* const* fields,
                     void* const* out, double scale, double* partial,
                     int* nblocks, void* stream);

/* operator_laplace (util_op.py:1038-1079): fields u, v; three equations
 * (lap u, lap v, and an identically zero one) with linear-extrapolation
 * walls. */
int odil_laplace_residual(int dtype, int n0, int n1, double dx, double dy,
                          const void* const* fields, void* const* res,
                          double* partial, void* stream);
int odil_laplace_jac(int dtype, int n0, int n1, double dx, double dy,
                     const void* const* fields, void* coef, void* const* res,
                     void* stream);
int odil_laplace_jac_entries(int* ent, int* shifts);
int odil_laplace_loss_grad(int dtype, int n0, int n1, double dx, double dy,
                           const void* const* fields, void* const* grad,
                           double* partial, int* nblocks, void* stream);
int odil_laplace_nblocks(int kind, int dtype, int n0, int n1);
int odil_laplace_normal_apply(int dtype, int n0, int n1, double dx,
                              double dy, const void* const* state,
                              const void* const* x, const void* const* b,
                              const void* const* d, void* const* y,
                              double omega, int mode, void* stream);
int odil_laplace_jtf(int dtype, int n0, int n1, double dx, double dy,
                     const void* const* fields, void* const* out,
                     double scale, double* partial, int* nblocks,
                     void* stream);

/* Inverse wave problem (speed recovery from sensor traces; no counterpart
 * in the reference, which has the forward wave.py case only). Loss =
 * mean(f_wave^2) + mean(f_data^2) + k_reg^2 mean((dc/dx)^2) with f_wave the
 * odil_wave residual with c(x)^2 u_xx (c [nx], field dtype) and
 * f_data = k_data (u - data[t][s]) on the columns with sensor_of[x] = s >= 0
 * (int32 [nx], -1 elsewhere; data [nt][nsensor], field dtype), the
 * reference's imposed-point equation form (ns_steady.py:396-411). Writes
 * grad_u [nt][nx], grad_c [nx] and split[3] (device FP64: the three loss
 * terms); `work` holds odil_wave_inverse_work(dtype, nt, nx) doubles. */
int odil_wave_inverse_work(int dtype, int nt, int nx);
int odil_wave_inverse_loss_grad(int dtype, int nt, int nx, double dt,
                                double dx, const double* init_u,
                                const double* init_ut, const double* left_u,
                                const double* right_u, const void* c,
                                const int* sensor_of, const void* data,
                                int nsensor, double k_data, double k_reg,
                                const void* u, void* grad_u, void* grad_c,
                                double* work, double* split, void* stream);
/* Solution of f_wave = 0 marched row by row (one workgroup): synthetic
 * sensor data of the inverse problem; c may be NULL (c = 1). */
int odil_wave_march(int dtype, int nt, int nx, double dt, double dx,
                    const double* init_u, const double* init_ut,
                    const double* left_u, const double* right_u,
                    const void* c, void* u, void* stream);

/* 2D linear advection p_t + a . grad p = 0 on an nt x nx x ny space-time
 * grid, periodic in x and y, first-order upwind on the sign of the constant
 * velocity (ax, ay), p at t = 0 from p_init [nx][ny] (advection_2d.py:
 * 103-223, problem 'direct'). One field, one equation. */
int odil_advection_nblocks(int nt, int nx, int ny);
int odil_advection_residual(int dtype, int nt, int nx, int ny, double dt,
                            double dx, double dy, double ax, double ay,
                            const void* p_init, const void* p, void* res,
                            double* partial, void* stream);
int odil_advection_loss_grad(int dtype, int nt, int nx, int ny, double dt,
                             double dx, double dy, double ax, double ay,
                             const void* p_init, const void* p, void* grad,
                             double* partial, void* stream);

/* Inverse advection problem (advection_2d.py:103-250, problem 'inverse',
 * central 0, implicit 0) on an nt x nx x ny space-time grid (nx, ny >= 3):
 * unknowns u, v, p (fields[3]); equations in the reference's order
 * fe = (p - p_exact) alpha m(t) (alpha already annealed by the caller,
 * get_anneal_factor :25-26; m = 1, or with onlyfinal 1 only at t = 0 and
 * nt - 1), fp = (p - p_prev)/dt + u Dx p + v Dy p (upwind on the sign of the
 * frozen velocity, p_prev = p_exact at t = 0), then kreg lap_xy(u),
 * kreg lap_xy(v) if kreg != 0, ktreg (u - u(t-1))/dt and the v form if
 * ktreg != 0 (0 at t = 0), divfree div_c(u, v) if divfree != 0. x, y
 * periodic (periodic != 0) or zero outside. p_exact [nt][nx][ny] in the
 * field dtype. *_neq gives the equation count, *_nblocks the partials per
 * equation (partial [neq][nblocks]); res may be NULL (sums only) and its
 * entries NULL. loss_grad writes grad[3] = (2/N) sum_e J_e^T f_e (the
 * reference's tape gradient: the upwind direction carries no derivative). */
int odil_advection_inverse_neq(double kreg, double ktreg, double divfree);
int odil_advection_inverse_nblocks(int nt, int nx, int ny);
int odil_advection_inverse_residual(int dtype, int nt, int nx, int ny,
                                    double dt, double dx, double dy,
                                    double alpha, int onlyfinal, double kreg,
                                    double ktreg, double divfree,
                                    int periodic, const void* p_exact,
                                    const void* const* fields,
                                    void* const* res, double* partial,
                                    void* stream);
int odil_advection_inverse_loss_grad(int dtype, int nt, int nx, int ny,
                                     double dt, double dx, double dy,
                                     double alpha, int onlyfinal, double kreg,
                                     double ktreg, double divfree,
                                     int periodic, const void* p_exact,
                                     const void* const* fields,
                                     void* const* grad, double* partial,
                                     void* stream);

/* Batched Poisson (many independent instances of uneven size in one
 * persistent launch; the reference solves one problem per process).
 * Instance i owns cells [off[i], off[i] + n0[i] n1[i]) of the packed u, src,
 * grad buffers (field dtype), spacings dx[i], dy[i] (device arrays:
 * off int64, n0/n1 int32, dx/dy FP64). `tiles` int32 [ntiles][4] =
 * (instance, x0, y0, 0) lists every tx x ty tile, grouped by instance with
 * tstart[i] (int32 [ninst + 1]) the first tile of instance i; nblocks is
 * the persistent grid size. Writes grad_i = (2 / N_i) J_i^T f_i,
 * loss[i] = mean(f_i^2) (FP64) and per-tile sums of f^2 in partial
 * [ntiles]. The per-instance residual is odil_poisson_residual's. */
int odil_poisson_batch_shapes(int dtype, int* shapes, int cap);
int odil_poisson_batch_loss_grad(int dtype, int ninst, const long long* off,
                                 const int* n0, const int* n1,
                                 const double* dx, const double* dy,
                                 const int* tiles, int ntiles,
                                 const int* tstart, int tx, int ty,
                                 int nblocks, const void* u, const void* src,
                                 void* grad, double* partial, double* loss,
                                 void* stream);

/* ---- block stencil operators and multigrid ------------------------------
 * A block stencil operator maps nvar_in fields on the node grid nw (2D or
 * 3D, host int[3], unused axes 1) to nvar_out fields on
 * nwh = (nw - 1) / stride + 1:
 *   (A x)_o[I] = sum_{e: out_e = o} coef_e[I] x_{in_e}[stride I + s_e],
 * zero outside the grid: SparseOperator (multigrid.py:877-1214) with
 * non-circular shifts (apply_shift, multigrid.py:942-947). coef is
 * [nent][nwh]; `ent` int32 [nent][5] = (out, in, s0, s1, s2) sorted by out,
 * `ostart` [nvar_out + 1] delimits each output field, `dent` [nvar_out] the
 * diagonal entry of each field (Jacobi). Index tables are device arrays.
 * ------------------------------------------------------------------------ */

/* mode 0 y = A x (SparseOperator.mul_field, multigrid.py:958-969);
 * 1 y = f - A x (MultigridOp.residual, multigrid.py:536-541);
 * 2 y = x + omega (f - A x) / diag (smoother_jacobi, multigrid.py:543-552).
 * max_entries = largest per-field entry count (<= 512). */
int odil_stencil_apply(int dtype, int ndim, const int* nw, int stride,
                       int nvar_in, int nvar_out, const int* ent,
                       const int* ostart, const int* dent, int max_entries,
                       const void* coef, const void* const* x, void* const* y,
                       const void* const* f, int mode, double omega,
                       void* stream);
/* Same modes for a symmetric square stride-1 operator stored as its upper
 * half (entries o < i, or o == i and s >= 0). `items` int32 [nitem][10]
 * lists per output field the direct and mirrored products
 * (q, var, cs_lin, xs_lin, cs0..2, xs0..2), `ostart` [nvar + 1], `reach`
 * host int[3] bounds the offsets per axis (at most 511), max_items <= 512,
 * nent the coefficient array count, lanes the threads per cell (1, 2, 4, 8,
 * 16; 0 picks by grid size; -1: one thread per 16 bytes of consecutive cells
 * with vector loads; -2: two interleaved cells per thread; x in one
 * allocation only for -1 and -2), per the 256-thread tiles one
 * workgroup sweeps (1..64; the autotuner chooses lanes and per for each
 * operator shape). Fields of x in one allocation (x_k = x_0 + k N) take
 * faster paths. `reach` may leave out the contiguous axis (with one more
 * line of margin on the axis before it) when the coefficients that couple
 * to outside points are zero (odil_stencil_zero_outside): reads past the
 * end of a grid line then meet zero coefficients. No reference counterpart
 * for the storage: it applies the same matrices as the reference's full
 * stencils (mul_transpose_op products and Galerkin levels are
 * symmetric). */
int odil_stencil_apply_sym(int dtype, int ndim, const int* nw, int nvar,
                           const int* items, const int* ostart,
                           const int* dent, const int* reach, int max_items,
                           int nent, int lanes, int per, const void* coef,
                           const void* const* x, void* const* y,
                           const void* const* f, int mode, double omega,
                           void* stream);
/* The same modes (0: y = A x, 1: y = f - A x, 2: one weighted-Jacobi
 * sweep, smoother_jacobi of multigrid.py:543-552) for a symmetric operator
 * in the storage of odil_stencil_apply_sym, on a 2D grid nw[2] with shifts
 * of at most 2 per axis (`reach`, host int[2 or 3]) and nvar in {1, 3, 4}
 * fields of x in one allocation. `ent` int32 [nent][5] = (o, i, s0, s1, s2)
 * is the stored entry table (device), `dent` the diagonal entry of each
 * field (device, mode 2). One workgroup per tile computes every output
 * field with the coefficient arrays streamed through LDS; `tile` in 1..6
 * picks the tile shape (odil_stencil_tile_lds gives its LDS bytes; shapes
 * above 64 KiB and more than 256 entries are refused with status 3). Same
 * reference call as odil_stencil_apply_sym. */
int odil_stencil_apply_sym_tiled(int dtype, const int* nw, int nvar,
                                 const int* ent, int nent, const int* dent,
                                 const int* reach, int tile, const void* coef,
                                 const void* const* x, void* const* y,
                                 const void* const* f, int mode, double omega,
                                 void* stream);
/* LDS bytes per workgroup of tile shape `tile` for nvar fields (-1 for an
 * unknown shape). Helper of odil_stencil_apply_sym_tiled. */
int odil_stencil_tile_lds(int dtype, int nvar, int tile);
/* Zero the coefficients of a stride-1 stencil operator that couple a node
 * to a point outside the grid (entry table `ent` int32 [nent][5] =
 * (o, i, s0, s1, s2), device; `reach` host int[3] bounds |s| per axis).
 * Under the zero fill of the reference's shifts (multigrid.py:941-947)
 * these coefficients never enter a product; with them zero the symmetric
 * sweeps need no per-product test along the contiguous axis away from the
 * other faces. No reference counterpart (a storage invariant). */
int odil_stencil_zero_outside(int dtype, int ndim, const int* nw,
                              const int* ent, int nent, const int* reach,
                              void* coef, void* stream);
/* y = A^T x or y += A^T x (mul_transpose_field, multigrid.py:1006-1033).
 * `tent` lists entry indices grouped by input field, `istart` delimits. */
int odil_stencil_apply_t(int dtype, int ndim, const int* nw, int stride,
                         int nvar_in, int nvar_out, const int* ent,
                         const int* tent, const int* istart, const void* coef,
                         const void* const* x, void* const* y, int accumulate,
                         void* stream);
/* Normal operator N = sum_e J_e^T J_e in stencil form from Jacobian
 * coefficients jc (mul_transpose_op + add_elementwise, multigrid.py:
 * 1035-1071,1138-1184, as assembled in util.py:578-585). `terms` int32
 * [nterm][5] = (a, b, sa0, sa1, sa2), grouped per output entry by
 * `tstart` [nout + 1]: N_o[m] = sum J_a[m - sa] J_b[m - sa]. FP64
 * accumulation; dtype_in/dtype_out in {F64->F64, F64->F32, F32->F32}. */
int odil_stencil_normal(int dtype_in, int dtype_out, int ndim, const int* nw,
                        const int* terms, const int* tstart, int nout,
                        const void* jc, void* out, void* stream);
/* Galerkin coarse operator on node grids (odd sizes):
 * sigma = +1 exact R A T (Multigrid.update_A, multigrid.py:116-157);
 * sigma = -1 the arithmetic of MultigridOp.coarsen_sparse_operator
 * (multigrid.py:391-427). `oent` int32 [nout][5] output entries, `terms`
 * int32 [nterm][9] = (q, r0, r1, r2, wt_id, s0, s1, s2, tr) per output
 * entry (tstart), rw [27] restriction weights, wtab interpolation weights;
 * rmode 0 injects at coarse boundary nodes (get_R, multigrid.py:497-534),
 * 1 keeps the weights (R = T^T / 2^dim). */
int odil_stencil_galerkin(int dtype, int ndim, const int* nw, const int* oent,
                          const int* terms, const int* tstart, int nout,
                          const double* rw, const double* wtab, const void* fc,
                          void* cc, int sigma, int rmode, void* stream);
/* coarse = R fine (multigrid.py:497-534 weights, rmode as above). */
int odil_mg_restrict(int dtype, int ndim, const int* nw, int nvar,
                     const double* rw, int rmode, const void* const* fine,
                     void* const* coarse, void* stream);
/* Cell-centred interpolation of one field from the grid nwc (host int[ndim],
 * ndim 1..3) to 2 nwc per axis, weights [1,3,3,1]/4 per axis over a halo
 * 2 SYMMETRIC - REFLECT (MultigridDecomp.interp_cell_field,
 * multigrid.py:1300-1360); accumulate != 0 adds to fine. The node-grid
 * interpolation (interp_node_field) is odil_mg_prolong_add. */
int odil_interp_cell(int dtype, int ndim, const int* nwc, const void* coarse,
                     void* fine, int accumulate, void* stream);
/* Its adjoint, coarse = I^T fine (the gradient of MultigridDecomp's
 * weights_to_field with respect to cell weights). */
int odil_interp_cell_t(int dtype, int ndim, const int* nwc, const void* fine,
                       void* coarse, void* stream);
/* odil_mg_restrict followed by the coarse level's first weighted-Jacobi
 * sweep from the zero initial guess (smoother_jacobi, multigrid.py:543-552,
 * on the zeros of MultigridOp.step, multigrid.py:590-593):
 * u_k = omega coarse_k / d_k with d_k the coarse operator's diagonal. */
int odil_mg_restrict_smooth(int dtype, int ndim, const int* nw, int nvar,
                            const double* rw, int rmode,
                            const void* const* fine, void* const* coarse,
                            const void* const* d, void* const* u,
                            double omega, void* stream);
/* fine += T coarse, multilinear interpolation (get_TT, multigrid.py:
 * 472-494, applied as in MultigridOp.step, multigrid.py:594-596). */
int odil_mg_prolong_add(int dtype, int ndim, const int* nw, int nvar,
                        const void* const* coarse, void* const* fine,
                        void* stream);
/* y_k = omega f_k / d_k on N values of nvar fields: smoother_jacobi
 * (multigrid.py:543-552) applied to the zero initial guess that each
 * coarse level of MultigridOp.step starts from (multigrid.py:590-593), with
 * d_k the diagonal of the level operator. fw != NULL also stores f_k
 * converted to the y precision (an FP32 preconditioner's copy of its FP64
 * input). dtype pairs (f, y): FP64/FP64, FP32/FP32, FP64/FP32. */
int odil_jacobi_zero(int dtype_f, int dtype_y, int nvar, long long N,
                     const void* const* f, const void* const* d,
                     void* const* y, void* const* fw, double omega,
                     void* stream);
/* Dense FP64 row-major matrix (zeroed by the caller) of a square stride-1
 * operator (sym != 0: upper-half storage) for the coarsest-level direct
 * solve (smoother_direct, multigrid.py:554-565; SparseOperator.tocsr,
 * multigrid.py:906-920). */
int odil_stencil_to_dense(int dtype, int ndim, const int* nw, int nvar,
                          const int* ent, int nent, const void* coef, int sym,
                          double* M, void* stream);
/* CSR of a stride-1 block stencil operator (J from odil_X_jac) in the layout
 * of Problem.linearize (util_op.py:877-894): row e N + c, column
 * k N + flat((c + s) mod shape) (np.roll wrap), int32 rowptr [nout N + 1],
 * col and val [nent N] with every row of field e holding its cnt_e entries
 * in entry order (explicit zeros kept). `ent`/`ostart` as above. */
int odil_stencil_to_csr(int dtype, int ndim, const int* nw, int nout,
                        const int* ent, const int* ostart, int nent,
                        const void* coef, int* rowptr, int* col, void* val,
                        void* stream);
/* One diagonal block of a right-looking blocked Cholesky factorisation of a
 * dense SPD FP64 matrix A (row-major, leading dimension lda, already
 * updated by the previous panels): A[k0:k0+nb, k0:k0+nb] = L L^T with L
 * written back (strict upper part of the block zeroed) and W = L^-1
 * (row-major, leading dimension odil_chol_block_size()); nb <= that size.
 * The panel and trailing updates are GEMMs of the caller. It replaces the
 * library factorisation of the coarsest level's Tikhonov solve
 * (smoother_direct, multigrid.py:554-565, solves the coarsest system with
 * a dense factorisation). */
int odil_chol_block_size(void);
int odil_chol_block(double* A, int lda, int k0, int nb, double* W,
                    void* stream);
/* y = M x with M dense FP64 [nvar N][nvar N] (the coarsest-level solve as
 * a precomputed solution operator). */
int odil_dense_gemv(int dtype_x, int dtype_y, int nvar, int N,
                    const double* M, const void* const* x, void* const* y,
                    void* stream);
/* Block-diagonal Jacobi, Multigrid.smoother_jacobi(full=True)
 * (multigrid.py:284-299): factor the per-cell nvar x nvar matrices of the
 * shift-0 coefficients of every block (`dq` host [nvar][nvar] index into
 * the nent coefficient arrays, -1 for none) by LU with partial pivoting into lu [nvar^2][N] and piv
 * int32 [nvar][N]; *singular (device int32, cleared by the caller) becomes 1
 * for an exactly zero pivot. odil_block_jacobi then does u += omega D^-1 r
 * per cell. */
int odil_block_diag_factor(int dtype, int nvar, int N, const int* dq,
                           int nent, const void* coef, void* lu, int* piv,
                           int* singular, void* stream);
int odil_block_jacobi(int dtype, int nvar, int N, const void* lu,
                      const int* piv, const void* const* r, void* const* u,
                      double omega, void* stream);
/* Multigrid.smoother_gauss_seidel (multigrid.py:301-305): one forward
 * Gauss-Seidel sweep u += omega (L^-1 (f - (A - L) u) - u), L the lower
 * triangle (diagonal included) of the square stride-1 operator in the flat
 * field-major C-order numbering; full (non-symmetric) storage, `ent` on the
 * device and the same table `ent_host` in host memory, x nvar fields of
 * working storage. */
int odil_stencil_gauss_seidel(int dtype, int ndim, const int* nw, int nvar,
                              const int* ent, const int* ent_host, int nent,
                              const void* coef, void* const* u,
                              const void* const* f, void* const* x,
                              double omega, void* stream);

/* ---- vector kernels of the Krylov / Newton drivers ----------------------
 * Scalars live in device memory so a whole solver iteration can be queued
 * (and captured in a HIP graph) without host round trips. */

/* Number of FP64 partials odil_dot needs in `partial`. */
int odil_dot_partials(void);
/* out[0] = sum_k x_k . y_k (FP64, fixed order); keep_prev != 0 copies the
 * old out[0] to out[1] first. */
int odil_dot(int dtype, int nvar, long long N, const void* const* x,
             const void* const* y, double* partial, double* out,
             int keep_prev, void* stream);
/* y_k += sign * (num / den) x_k (den == 0: no update). */
int odil_axpy_ratio(int dtype, int nvar, long long N, void* const* y,
                    const void* const* x, const double* num,
                    const double* den, double sign, void* stream);
/* x_k += (num / den) p_k and r_k -= (num / den) q_k in one pass (the two
 * conjugate-gradient updates sharing a step length; den == 0: none). */
int odil_axpy2_ratio(int dtype, int nvar, long long N, void* const* x,
                     const void* const* p, void* const* r,
                     const void* const* q, const double* num,
                     const double* den, void* stream);
/* odil_axpy2_ratio, then y_k = omega r_k / d_k with the updated r (the
 * arithmetic of odil_jacobi_zero, same dtype for every vector) in the same
 * pass: the conjugate-gradient updates and the first sweep from zero of the
 * preconditioner cycle applied to r next. No reference counterpart (the
 * reference solves with plain V-cycles, util.py:588-599). */
int odil_axpy2_jacobi_zero(int dtype, int nvar, long long N, void* const* x,
                           const void* const* p, void* const* r,
                           const void* const* q, const double* num,
                           const double* den, const void* const* d,
                           void* const* y, double omega, void* stream);
/* Flexible-CG pass after the preconditioner, one read of each vector:
 * z = zw converted to dtype (the Krylov precision of r and zo);
 * *rz_prev = *rz; *rz = r.z; *rzo = r.zo with zo the previous z (sums in
 * FP64); then zo = z. `partial` holds 2 * odil_dot_partials() FP64 values.
 * (dtype_w, dtype): FP64/FP64, FP32/FP64, FP32/FP32. */
int odil_fcg_z(int dtype_w, int dtype, int nvar, long long N,
               const void* const* zw, const void* const* r, void* const* zo,
               double* partial, double* rz, double* rz_prev, double* rzo,
               void* stream);
/* The sums of odil_fcg_z without the conversion and the copy (z, r, zo of
 * one dtype, none written): *rz_prev = *rz; *rz = r.z; *rzo = r.zo, the
 * same values odil_fcg_z gives, for a caller that keeps z and the previous
 * z in two buffers used in turn. */
int odil_fcg_dots(int dtype, int nvar, long long N, const void* const* z,
                  const void* const* r, const void* const* zo,
                  double* partial, double* rz, double* rz_prev, double* rzo,
                  void* stream);
/* p_k = z_k + ((num - sub) / den) p_k (sub may be NULL; den == 0: p = z). */
int odil_xpay_ratio(int dtype, int nvar, long long N, void* const* p,
                    const void* const* z, const double* num,
                    const double* sub, const double* den, void* stream);

/* ---- optimizer kernels -------------------------------------------------- */

/* One fused Adam step in place (AdamNativeOptimizer.run, optimizer.py:
 * 303-314): m += (g - m)(1 - beta1), v += (g^2 - v)(1 - beta2),
 * x -= alpha m / (sqrt(v) + eps), alpha = lr sqrt(1 - beta2^t) /
 * (1 - beta1^t) from the caller. */
int odil_adam_step(int dtype, long long N, void* x, const void* g, void* m,
                   void* v, double alpha, double beta1, double beta2,
                   double eps, void* stream);
/* out = x + a y (line-search trial points and gradient descent,
 * GdOptimizer, optimizer.py:241-268); out may alias x or y. */
int odil_waxpy(int dtype, long long N, void* out, const void* x,
               const void* y, double a, void* stream);
/* out = a x + b y (recurrences of the iterative normal-equation solvers
 * 'lsqr' and 'bicgstab', linsolver.py:59-101); out may alias x or y. */
int odil_axpby(int dtype, long long N, void* out, const void* x,
               const void* y, double a, double b, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ODIL_AMD_H */
