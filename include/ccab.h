/*
 * ccab.h -- C ABI of the MI355X (gfx950) collocation-BEM hot path.
 *
 * Drop-in boundary for the Rust crate dbartilson/cc_abstemious v1.2.0.  Each
 * entry point names the reference function it replaces (file:line in that
 * crate).  Conventions:
 *  - complex numbers are interleaved {re, im} (num-complex Complex<f64> /
 *    Complex<f32> layout, #[repr(C)]-compatible); matrices are column-major
 *    like nalgebra's DMatrix, with an explicit leading dimension;
 *  - pointers named d_* are device (HBM) pointers, h_* are host pointers;
 *  - every function returns CCAB_OK (0) or a negative status; the message of
 *    the last failure is ccab_last_error(ctx).  Nothing aborts or throws
 *    across the ABI;
 *  - a context owns one HIP stream (or borrows the caller's, see
 *    ccab_set_stream) and a device workspace; use one context per host thread.
 *  - collocation points are passed as the 8-double record of
 *    CollocationPoint (mesh_data.rs:71-77): x y z nx ny nz area wt.
 */
#ifndef CCAB_H
#define CCAB_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCAB_VERSION 10000

#define CCAB_OK 0
#define CCAB_EINVAL (-1)   /* invalid argument (sizes, enums, singleton rule) */
#define CCAB_EHIP (-2)     /* HIP runtime error */
#define CCAB_ENOMEM (-3)   /* device allocation failed */
#define CCAB_ENAN (-4)     /* NaN/Inf produced (zero-area element, origin source) */
#define CCAB_ESINGULAR (-5) /* exact zero pivot in LU */

typedef struct { double re, im; } ccab_c128;
typedef struct { float re, im; } ccab_c64;
typedef struct ccab_ctx ccab_ctx;
typedef struct ccab_hmatrix ccab_hmatrix;

/* input_data.rs:15-25 (ProblemType, MethodType), :34-51 (IncidentWaveInput, BCType) */
enum { CCAB_CLASSICAL = 0, CCAB_BURTON_MILLER = 1 };
enum { CCAB_EXTERIOR = 0, CCAB_INTERIOR = 1 };
enum { CCAB_BC_PRESSURE = 0, CCAB_BC_NORMAL_VELOCITY = 1, CCAB_BC_IMPEDANCE = 2 };
enum { CCAB_WAVE_PLANE = 0, CCAB_WAVE_SPHERICAL = 1 };
/* influence_matrix.rs:164-167 */
enum { CCAB_LHS = 0, CCAB_RHS = 1 };
/* storage / arithmetic precision of a matrix; CCAB_C32 (complex FP16, scaled
 * per system) is available as the GMRES operator of the ITERATIVE_MIXED sweep */
enum { CCAB_C128 = 0, CCAB_C64 = 1, CCAB_C32 = 2 };
/* solver kinds, input_data.rs:27-32 (+ the mixed-precision LU) */
enum { CCAB_SOLVER_DIRECT = 0, CCAB_SOLVER_ITERATIVE = 1, CCAB_SOLVER_HIERARCHICAL = 2,
       CCAB_SOLVER_DIRECT_MIXED = 3, CCAB_SOLVER_ITERATIVE_MIXED = 4 };
/* GMRES stopping rule */
enum { CCAB_STOP_BACKWARD_ERROR = 0, /* gmres.rs:142-144 */
       CCAB_STOP_RELATIVE_RESIDUAL = 1 };

/* Physics corrections (no reference counterpart; opt-in, 0 = the reference's
 * numerics bit for bit).  Each flag replaces one reference formula by the
 * textbook one:
 *   CCAB_FIX_BM_DH            Burton-Miller hypersingular kernel: the sign of
 *                             the (k^2 r + 3 f1)(e.n_x)(e.n_y) term
 *                             (influence_matrix.rs:30), so that H = dG/dn_y
 *   CCAB_FIX_FIELD_SIGN       field points: phi = phi_inc + M phi - L vn
 *                             (solve.rs:180-181 has the opposite sign)
 *   CCAB_FIX_IMPEDANCE_VN     impedance BC: vn = +i omega rho phi / Z, the
 *                             value the solved system implies (solve.rs:67,146
 *                             return -i omega rho phi / Z)
 *   CCAB_FIX_SPHERICAL_ORIGIN spherical wave: the source stays at the given
 *                             origin (incident_wave.rs:44,91 move it to the
 *                             unit sphere)
 *   CCAB_FIX_QUAD_AREA        ccab_collocation_points_ex: quad area = the
 *                             element's area with weight 1 (elements.rs:168
 *                             takes a quarter with weight 4)
 *   CCAB_FIX_PRESSURE_SIGN    pressure BC: G vn = rhs_inc + H p, the equation
 *                             the velocity branch (solve.rs:46-52) implies
 *                             (solve.rs:36-43 and :99-104 subtract H p, which
 *                             flips the sign of vn); also the RHS-side factors
 *                             of the row/column getter and the H matrix */
enum { CCAB_FIX_BM_DH = 1, CCAB_FIX_FIELD_SIGN = 2, CCAB_FIX_IMPEDANCE_VN = 4,
       CCAB_FIX_SPHERICAL_ORIGIN = 8, CCAB_FIX_QUAD_AREA = 16, CCAB_FIX_PRESSURE_SIGN = 32 };

/* The physics of one frequency (UserInput fields + PreData frequency,
 * input_data.rs:79-91, preprocess.rs:35-45). */
typedef struct {
    double frequency;     /* Hz; k = 2*pi*f / c */
    double sound_speed;
    double mass_density;
    int32_t method;       /* CCAB_CLASSICAL / CCAB_BURTON_MILLER */
    int32_t problem_type; /* CCAB_EXTERIOR / CCAB_INTERIOR */
    int32_t bc_type;      /* CCAB_BC_* */
    int32_t corrections;  /* CCAB_FIX_* bits; 0 for the reference's numerics */
    double bc_value[2];   /* SurfaceBoundaryCondition.value */
} ccab_physics;

/* IncidentWaveInput (input_data.rs:34-44) */
typedef struct {
    int32_t kind;          /* CCAB_WAVE_PLANE / CCAB_WAVE_SPHERICAL */
    int32_t pad_;
    double vec[3];         /* plane: direction; spherical: origin */
    double amplitude[2];   /* pressure amplitude */
} ccab_wave;

typedef struct {
    int64_t max_iterations;  /* <= 0 -> 1000, capped at n (gmres.rs:41-48) */
    double tolerance;        /* <= 0 -> 1e-5 */
    int32_t restart;         /* 0 -> ceil(sqrt(max_iterations)) (gmres.rs:45-46) */
    int32_t ref_mgs;         /* 1 -> Arnoldi coefficient qk1.dotc(q_i) (gmres.rs:168) */
    int32_t stop_rule;       /* CCAB_STOP_* */
    int32_t matrix_prec;     /* CCAB_C128 or CCAB_C64 operator */
} ccab_gmres_params;

typedef struct {
    int64_t num_mv;          /* matrix-vector products (gmres.rs:66) */
    int64_t iterations;      /* inner Arnoldi steps */
    int64_t restarts;
    int64_t ir_steps;        /* FP64 iterative-refinement corrections */
    double backward_error;   /* ||b-Ax|| / (||A||_F ||x|| + ||b||) at exit */
    double rel_residual;     /* ||b-Ax|| / ||b|| at exit (FP64 operator) */
    int32_t converged;
    int32_t pad_;
} ccab_solve_stats;

typedef struct {
    int32_t solver;          /* CCAB_SOLVER_* */
    int32_t leaf_size;       /* HIERARCHICAL: cluster leaf cardinality (<=0 -> 32, solve.rs:91,117) */
    ccab_gmres_params gmres; /* ITERATIVE(_MIXED), HIERARCHICAL */
    double ir_tolerance;     /* mixed: stop when ||dx||/||x|| < this (<=0 -> 1e-12) */
    int32_t ir_max_steps;    /* mixed: <=0 -> 10 */
    /* ITERATIVE_MIXED with the CCAB_C32 operator: the matrix of the FP64
     * residuals.  CCAB_C128 (0): the c128 matrix (16 B + 4 B stored per
     * entry).  CCAB_C32: the FP16 operator plus its complex-FP32 remainder
     * (4 B + 8 B per entry; the sum equals the FP64 entry to about 2^-35
     * relative).  No reference counterpart. */
    int32_t residual_matrix;
    double aca_tolerance;    /* HIERARCHICAL: ACA tolerance (<=0 -> 1e-4, solve.rs:92,118) */
    double eta;              /* HIERARCHICAL: admissibility (<=0 -> 4.0, h_matrix.rs:136) */
} ccab_solver_cfg;

/* ---- context ---------------------------------------------------------- */
int ccab_version(void);
int ccab_create(int device, ccab_ctx** out);
int ccab_destroy(ccab_ctx* ctx);
const char* ccab_last_error(const ccab_ctx* ctx);
/* Run on the caller's hipStream_t (NULL = the context's own stream). */
int ccab_set_stream(ccab_ctx* ctx, void* hip_stream);
int ccab_synchronize(ccab_ctx* ctx);
/* Order the context's stream after all work queued so far on hip_stream
 * (NULL = the legacy default stream): an event recorded there, waited on by
 * the context's stream.  For callers that produce inputs on another stream. */
int ccab_stream_wait(ccab_ctx* ctx, void* hip_stream);
/* The reverse order: work enqueued on hip_stream after this call waits for
 * everything enqueued on the context's stream so far (device-side, no host
 * synchronisation); with ccab_stream_wait before a call, a caller on another
 * stream needs no ccab_synchronize to read the results. */
int ccab_stream_join(ccab_ctx* ctx, void* hip_stream);

/* In-process kernel timing with HIP events on the context's stream (no
 * reference counterpart; the reference only logs).  Classes: 0 assembly
 * (work = bytes stored), 1 LU panel, 2 LU triangular update, 3 LU trailing
 * complex GEMM (work = flops), 4 row swaps, 5 diagonal-block inverses,
 * 6 triangular solves, 7 GEMV over c64 / c32 matrices (work = matrix
 * bytes), 8 Gram-Schmidt, 9 field evaluation (work = point pairs),
 * 10 H-matrix GEMV (work = block bytes), 11 surrogate network (work =
 * flops), 12 GEMV over c128 matrices (work = matrix bytes). */
int ccab_profile_enable(ccab_ctx* ctx, int on);
int ccab_profile_read(ccab_ctx* ctx, int cls, int64_t* count, double* ms, double* work,
                      int reset);

/* LU tuning (no reference counterpart): the complex-GEMM variant for
 * precision prec (0..7: the library's MFMA kernel with bit 0 XCD-aware tile
 * order, bit 1 LDS-staged epilogue, bit 2 128-row tiles; 8: the trailing
 * update C -= A B by rocBLAS's strided-batched complex GEMM, except inside a
 * captured HIP graph, where variant 3 runs) and the outer panel width (64
 * or 128) that sets the K of the trailing update. */
int ccab_set_tuning(ccab_ctx* ctx, int prec, int gemm_variant, int lu_outer);
/* LU panel placement (no reference counterpart): panels of at least
 * min_rows rows are factored by up to one workgroup per CU per system
 * (grid-synchronised per column) when the batch leaves CUs idle; 0 keeps
 * one workgroup per system.  Default 2048 (environment CCAB_LU_COOP_ROWS).
 * The pivot rule is the same either way; the factors agree to rounding. */
int ccab_set_lu_coop(ccab_ctx* ctx, int64_t min_rows);
/* HIP graphs (no reference counterpart): with on != 0 the blocked LU is
 * captured once per (shape, buffers, tuning) and replayed on later calls
 * (the 400-600 launches of a factorisation become one graph launch); off by
 * default, also set by the environment variable CCAB_GRAPHS=1.  Disabled
 * while ccab_profile_enable is on.  Captured graphs are dropped when a
 * workspace buffer is reallocated. */
int ccab_set_graphs(ccab_ctx* ctx, int on);
/* Warm starts of the batched GMRES-IR (no reference counterpart: the
 * reference always starts from zero, gmres.rs:38-39): with on != 0 (the
 * default; environment CCAB_WARM_AX=0 turns it off) ccab_sweep_warm and
 * ccab_solve_bodies with d_x0 form A x0 inside the assembly, each FP64
 * entry times the guess as it is formed, so the first residual b - A x0
 * reads no stored matrix; with 0 it is a product over the stored matrices
 * after the assembly.  The two agree to rounding. */
int ccab_set_warm_ax(ccab_ctx* ctx, int on);
/* Batched GMRES-IR (no reference counterpart: the reference restarts GMRES
 * from scratch, gmres.rs:176-190): with nvec > 0 (default 8; environment
 * CCAB_GIR_RECYCLE) every refinement cycle after a system's first is
 * deflated by nvec directions of the first cycle's Krylov space (GCRO:
 * C = A U from that cycle's Arnoldi relation, no extra matrix products):
 * its harmonic Ritz directions of smallest magnitude where the restart
 * length is at most 32 and nvec at most 8 (environment CCAB_GIR_RITZ=0:
 * never), else the first nvec directions of A K_k; 0 restarts every cycle
 * from scratch.  nvec <= 64; restart lengths above 64 and HIP-graph cycles
 * are never deflated. */
int ccab_set_gir_deflation(ccab_ctx* ctx, int nvec);
/* The autotuner: times the batched LU of `batch` synthetic n x n systems in
 * precision prec for every GEMM variant (0..8) and outer width (64, 128),
 * keeps the fastest in the context (as ccab_set_tuning) and returns it;
 * ms_table (optional, 18 doubles) receives the time of variant v at width 64
 * (index v) and 128 (index 9 + v).  Factors do not depend on variants 0..7;
 * variant 8 (rocBLAS) agrees with them to rounding. */
int ccab_autotune(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, int* best_variant,
                  int* best_outer, double* ms_table);

/* Sweep autotuner (no reference counterpart): times, on the caller's mesh and
 * frequencies (ph[nfreq], d_cpts), the batched sweep assembly for 8, 16 and
 * 32 systems per thread x 64, 128 and 256 columns per workgroup, and the
 * batched GEMV of the GMRES operator (precision op_prec: CCAB_C32, the
 * FP16 operator, or CCAB_C64) and of the refinement matrix (the FP32
 * remainder with CCAB_C32, c128 with CCAB_C64) for workgroup targets 256 ...
 * 4096 x ordinary / non-temporal loads; keeps the fastest in the context for
 * size n (ccab_sweep, ccab_sweep_warm and ccab_gmres_ir_batched of that size
 * use them).  best (optional, 6 ints): systems per thread, column chunk,
 * operator GEMV target, operator non-temporal, refinement GEMV target,
 * refinement non-temporal.  ms_table (optional, 29 doubles): assembly times
 * [kf index * 3 + chunk index], then 10 operator GEMV times [target index *
 * 2 + nt], then 10 refinement GEMV times.  The choices change only how the
 * work is split (the order of partial sums), so results agree to rounding. */
int ccab_autotune_sweep(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq,
                        const double* d_cpts, int64_t n, int op_prec, int* best,
                        double* ms_table);
/* Set the sweep tuning of size n by hand (cfg: the 6 ints of
 * ccab_autotune_sweep's best), or drop it (cfg = NULL: the defaults 16, 64
 * and the context's GEMV settings). */
int ccab_set_sweep_tuning(ccab_ctx* ctx, int op_prec, int64_t n, const int* cfg);

/* ---- geometry (host) --------------------------------------------------- */
/* One centroid collocation point per linear tri/quad element:
 * elements.rs:49-181 (TRIGP1 / QUADGP1, detj incl. the quad 0.25 factor at
 * elements.rs:168) + preprocess.rs:126-153.  conn: 4 node ids per element
 * (-1 padded), nverts 3 or 4.  h_cpts: 8*nelem doubles. */
int ccab_collocation_points(const double* h_nodes, const int64_t* h_conn,
                            const int32_t* h_nverts, int64_t nelem, double* h_cpts);
/* ccab_collocation_points with CCAB_FIX_QUAD_AREA (or 0) in corrections */
int ccab_collocation_points_ex(const double* nodes, const int64_t* conn, const int32_t* nverts,
                               int64_t nelem, int corrections, double* cpts);

/* ---- incident wave (device) ---------------------------------------------- */
/* incident_wave.rs:7-63 (get_incident_surface): d_rhs[n] */
int ccab_incident_surface(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w,
                          const double* d_cpts, int64_t n, ccab_c128* d_rhs);

/* ---- influence matrices (device) ------------------------------------------ */
/* influence_matrix.rs:64-98 (get_dense_surface_matrices) -> (H, G).  Writes
 * the n x n matrices (either pointer may be NULL) in precision prec
 * (CCAB_C128: FP64 arithmetic; CCAB_C64: FP32 arithmetic), leading dim ld. */
int ccab_dense_surface_matrices(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                int64_t n, int prec, void* d_H, void* d_G, int64_t ld);

/* Fused assembly + BC algebra of solve.rs:24-71 for `batch` frequencies on
 * one mesh: LHS_b = alpha_b H_b + beta_b G_b (get_lhs_factors,
 * influence_matrix.rs:130-146) written as c128 (d_lhs128) and/or c64
 * (d_lhs64, rounded from the same FP64 values), system b at offset b*stride,
 * and the right-hand side of the BC algebra into d_rhs[b*n..]:
 *   NormalVelocity: G v - rhs_inc, Pressure: rhs_inc - H p, Impedance: -rhs_inc.
 * d_rhs_inc may equal d_rhs.  ph[batch] are host structs. */
int ccab_assemble_system(ccab_ctx* ctx, const ccab_physics* ph, int64_t batch,
                         const double* d_cpts, int64_t n, ccab_c128* d_lhs128,
                         ccab_c64* d_lhs64, int64_t ld, int64_t stride,
                         const ccab_c128* d_rhs_inc, ccab_c128* d_rhs);

/* influence_matrix.rs:170-231 (get_surface_row_or_column): one row (ni == 1)
 * or one column (nj == 1) of alpha*H + beta*G for side CCAB_LHS / CCAB_RHS;
 * the diagonal shifts are added at the first position holding the diagonal.
 * Returns CCAB_EINVAL unless exactly one of ni, nj is 1 (the reference only
 * logs that error, influence_matrix.rs:174-176). */
int ccab_surface_row_or_column(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                               int64_t n, int side, const int64_t* d_i, int64_t ni,
                               const int64_t* d_j, int64_t nj, ccab_c128* d_out);

/* ---- dense LU (device) ------------------------------------------------------ */
/* solve.rs:166-171 (solve_lu): factor a copy of A (c128, FP64 MFMA trailing
 * updates) and overwrite d_x (= b on entry) with the solution. */
int ccab_solve_lu(ccab_ctx* ctx, int64_t n, const ccab_c128* d_A, int64_t lda, ccab_c128* d_x);

/* Batched right-looking partial-pivot LU in place, prec CCAB_C128 or CCAB_C64.
 * d_ipiv[b*n + i] = 0-based row swapped with row i (LAPACK order). */
int ccab_lu_factor_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, void* d_A,
                           int64_t lda, int64_t strideA, int32_t* d_ipiv);
/* Solve with the factors for nrhs right-hand sides per system (in place). */
int ccab_lu_solve_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch,
                          const void* d_LU, int64_t lda, int64_t strideA,
                          const int32_t* d_ipiv, void* d_B, int64_t ldb, int64_t strideB,
                          int64_t nrhs);

/* Mixed precision (no reference counterpart; replaces solve_lu for the
 * sweep): factor d_A64 in place in complex FP32 on MFMA, then FP64 iterative
 * refinement with residuals against d_A128 until ||dx||/||x|| < ir_tol.
 * d_x: b on entry, x on exit (n x batch, stride n).  stats[batch] or NULL. */
int ccab_solve_mixed_batched(ccab_ctx* ctx, int64_t n, int64_t batch,
                             const ccab_c128* d_A128, ccab_c64* d_A64, int64_t lda,
                             int64_t strideA, ccab_c128* d_x, double ir_tol,
                             int ir_max_steps, ccab_solve_stats* stats);

/* ---- GMRES (device) --------------------------------------------------------- */
/* gmres.rs:35-202 over a dense operator (c128 or c64 per p->matrix_prec).
 * d_x: b on entry, x on exit; d_x0: initial guess or NULL (zero start as the
 * reference, gmres.rs:38-39). */
int ccab_gmres_dense(ccab_ctx* ctx, int64_t n, const void* d_A, int64_t lda,
                     ccab_c128* d_x, const ccab_c128* d_x0, const ccab_gmres_params* p,
                     ccab_solve_stats* stats);

/* Batched GMRES with FP64 iterative refinement (no reference counterpart;
 * stands in for per-frequency solve_lu / GMRES::solve in the sweep).  Each
 * cycle: FP64 residual r = b - A x with d_A128 (c64 d_A64 when d_A128 is
 * NULL), one GMRES(m) cycle on the c64 operator for A d = r (FP64 Krylov
 * vectors and Gram-Schmidt) to ||r - A d|| <= inner_tol ||r||, x += d; stop at
 * ||b - A x|| <= tol ||b||.  Parameters from cfg: gmres.tolerance = tol (<=0 ->
 * 1e-10), gmres.restart = m (0 -> 50), ir_tolerance = inner_tol (<=0 -> 1e-4),
 * ir_max_steps = maximum cycles (<=0 -> 20), gmres.ref_mgs = Arnoldi
 * coefficient.  d_x: b on entry, x on exit (n x batch); d_x0 optional warm start,
 * first scaled by the complex alpha minimising the FP64 residual
 * ||b - alpha A x0|| (so a poor guess is never worse than the zero start). */
int ccab_gmres_ir_batched(ccab_ctx* ctx, int64_t n, int64_t batch, const ccab_c128* d_A128,
                          const ccab_c64* d_A64, int64_t lda, int64_t strideA, ccab_c128* d_x,
                          const ccab_c128* d_x0, const ccab_solver_cfg* cfg,
                          ccab_solve_stats* stats);

/* ---- field points (device) --------------------------------------------------- */
/* influence_matrix.rs:101-127 (get_dense_field_matrices) fused with
 * solve.rs:173-185 (get_field), incident_wave.rs:66-100 (get_incident_field)
 * and postprocess.rs:70-88 (convert to pressure), for nfreq frequencies of one
 * mesh; M and L are never stored.  d_fps: 3 doubles per field point; d_phi,
 * d_vn: n x nfreq surface solution; d_out[f*nfp + p] =
 *   s_f ((want_total ? phi_inc : 0) - sum_j M_pj phi_fj + sum_j L_pj vn_fj),
 * s_f = i omega_f rho if to_pressure else 1 (the reference's sign);
 * d_inc (optional, nfp x nfreq) = s_f phi_inc.  w may be NULL when neither
 * want_total nor d_inc asks for the incident field. */
int ccab_field(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
               const double* d_cpts, int64_t n, const double* d_fps, int64_t nfp,
               const ccab_c128* d_phi, const ccab_c128* d_vn, int want_total, int to_pressure,
               ccab_c128* d_out, ccab_c128* d_inc);
/* ccab_field without allocation (no reference counterpart): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * ccab_field; its scratch (per-frequency constants, a copy of the surface
 * points, per-chunk partial sums) lives in one 32 MiB array per device in
 * the library's code object instead of context buffers.  Limits: nfreq <=
 * 256 and 8 n doubles + 16 nfp nfreq bytes (one surface chunk) within the
 * 32 MiB, else CCAB_EINVAL; fewer surface chunks than ccab_field may be used
 * when the partial sums would not fit (same sums, other chunking).  Every
 * call on a device shares that device's array: calls that overlap on one GPU
 * (two contexts or streams) race.  The call waits for its stream before it
 * returns (the constants are staged from the caller's stack). */
int ccab_field_noalloc(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                       const double* d_cpts, int64_t n, const double* d_fps, int64_t nfp,
                       const ccab_c128* d_phi, const ccab_c128* d_vn, int want_total,
                       int to_pressure, ccab_c128* d_out, ccab_c128* d_inc);
/* incident_wave.rs:66-100 (get_incident_field): the incident potential at
 * nfp field points (d_out[nfp]) */
int ccab_incident_field(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w,
                        const double* d_fps, int64_t nfp, ccab_c128* d_out);
/* influence_matrix.rs:101-127 (get_dense_field_matrices) -> (M, L), nfp x n
 * column-major c128 with leading dimension ld (either pointer may be NULL);
 * for callers that keep the reference's separate get_field step. */
int ccab_dense_field_matrices(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                              int64_t n, const double* d_fps, int64_t nfp, ccab_c128* d_M,
                              ccab_c128* d_L, int64_t ld);

/* ---- hierarchical matrix (device) ---------------------------------------------- */
/* src/solve/h_matrix.rs:127-148 (HMatrix::new_from) of the side matrix
 * a H + b G (side CCAB_LHS / CCAB_RHS, the entries of
 * get_surface_row_or_column): cluster tree split at the bounding-box midpoint
 * of the longest side down to leaf_size points (src/solve/h_matrix/cluster.rs:19-110),
 * block tree with admissibility min(diam) <= eta * dist
 * (src/solve/h_matrix/block.rs:18-96), inadmissible blocks dense, admissible
 * blocks by ACA with partial pivoting and SVD recompression at aca_tolerance
 * (src/solve/aca.rs:57-179), all blocks of one kind in one batched launch.
 * An admissible block whose ACA would need more than 64 terms is stored dense
 * (counted in stats[4]).  The handle keeps its device memory until
 * ccab_hmatrix_destroy and is used on the device of the building context. */
int ccab_hmatrix_build(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts, int64_t n,
                       int side, int64_t leaf_size, double eta, double aca_tolerance,
                       ccab_hmatrix** out);
/* ccab_hmatrix_build with the block storage chosen (no reference counterpart):
 * prec CCAB_C128 (= ccab_hmatrix_build) or CCAB_C64, blocks built in FP64
 * (ACA, recompression, norms) and stored in complex FP32 (U and V packed to
 * their ranks), which halves the bytes of every matvec; the products still
 * accumulate in FP64.  The SOLVER_HIERARCHICAL sweep selects it with
 * gmres.matrix_prec = CCAB_C64. */
int ccab_hmatrix_build_ex(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts, int64_t n,
                          int side, int64_t leaf_size, double eta, double aca_tolerance, int prec,
                          ccab_hmatrix** out);
int ccab_hmatrix_destroy(ccab_hmatrix* h);
/* src/solve/h_matrix.rs:218-234 (HMatrix::gemv): d_y = alpha H d_x + beta d_y,
 * alpha and beta complex {re, im} (beta = 0: d_y is not read, as in BLAS);
 * block contributions are summed per row in a fixed order (dense blocks, then
 * low-rank blocks, in block-tree order). */
int ccab_hmatrix_gemv(ccab_ctx* ctx, const ccab_hmatrix* h, const double* alpha,
                      const ccab_c128* d_x, const double* beta, ccab_c128* d_y);
/* src/solve/h_matrix.rs:236-247 (to_full): the n x n matrix the H matrix stands for */
int ccab_hmatrix_to_full(ccab_ctx* ctx, const ccab_hmatrix* h, ccab_c128* d_A, int64_t lda);
/* src/solve/h_matrix.rs:189-216 (update_norm, print_stats): st[0] admissible
 * blocks, st[1] inadmissible blocks, st[2] stored entries, st[3] sum of the
 * ranks, st[4] admissible blocks stored dense; *norm = sum of the dense
 * Frobenius norms and the ACA norm estimates (the ||A|| of the GMRES stop). */
int ccab_hmatrix_stats(const ccab_hmatrix* h, int64_t* st, double* norm);
/* ccab_hmatrix_stats with more entries (no reference counterpart): st[0..4]
 * as there; st[5..8] admissible blocks by ACA rank before recompression
 * (<= 8, 9-16, 17-32, 33-64: the LDS sizes of the recompression kernels);
 * st[9] largest rank after recompression; st[10] rows, st[11] columns.
 * nst = entries wanted (at most 12 are written). */
int ccab_hmatrix_stats_ex(const ccab_hmatrix* h, int64_t* st, int nst, double* norm);
/* src/solve/h_matrix.rs:127-148 (HMatrix::new_from) over an explicit n x n
 * operator: its get_row_or_column closure reads the column-major c128 matrix
 * d_A (leading dimension lda, device), as in the reference's tests
 * (h_matrix.rs:282-298).  h_xyz (host, 3 doubles per point) holds the
 * coordinates of the CollocationPoints the cluster tree splits; point i is
 * equation i (an identity eqn_map).  leaf_size, eta (HMatrix::new_from fixes
 * 4.0, h_matrix.rs:136), aca_tolerance and prec (CCAB_C128, or CCAB_C64 block
 * storage) as ccab_hmatrix_build_ex.  The handle works with every
 * ccab_hmatrix_* function and ccab_gmres_hmatrix. */
int ccab_hmatrix_build_matrix(ccab_ctx* ctx, int64_t n, const ccab_c128* d_A, int64_t lda,
                              const double* h_xyz, int64_t leaf_size, double eta,
                              double aca_tolerance, int prec, ccab_hmatrix** out);
/* src/solve/aca.rs:22-33 (ACA::new(tol, m, n, get_row, get_column)) over an
 * explicit m x n matrix d_A (column-major c128, leading dimension lda,
 * device; the closures of aca.rs:303-304 read its rows and columns):
 * partial-pivot ACA with the reference's seeded reference rows
 * (aca.rs:57-136) and SVD recompression at tol (aca.rs:137-179).  The result
 * is an m x n ccab_hmatrix holding one low-rank block: ccab_hmatrix_gemv is
 * ACA::gemv (aca.rs:181-195; x has n entries, y has m), the norm of
 * ccab_hmatrix_stats is ACA::get_norm (aca.rs:212-241) and st[3] is
 * ACA::get_num_uv.  A matrix the ACA cannot represent with 64 terms is stored
 * dense (st[4] = 1).  ccab_hmatrix_to_full writes m x n (lda >= m). */
int ccab_aca_build(ccab_ctx* ctx, int64_t m, int64_t n, const ccab_c128* d_A, int64_t lda,
                   double tol, ccab_hmatrix** out);
/* src/solve/gmres.rs:35-202 with the H matrix as operator (src/solve.rs:119-125).
 * d_x: b on entry, x on exit; d_x0 initial guess or NULL. */
int ccab_gmres_hmatrix(ccab_ctx* ctx, const ccab_hmatrix* h, ccab_c128* d_x,
                       const ccab_c128* d_x0, const ccab_gmres_params* p,
                       ccab_solve_stats* stats);

/* ---- the surface solve (device) ---------------------------------------------- */
/* solve.rs:11-71 (solve_for_surface, dense branches): assembly + BC algebra +
 * LU / GMRES / mixed LU with FP64 refinement; outputs surface phi and vn.
 * CCAB_SOLVER_HIERARCHICAL follows get_surface_hmatrix (solve.rs:74-150):
 * RHS-side H matrix (when the BC value is nonzero) for the right-hand side,
 * LHS-side H matrix and GMRES over it. */
int ccab_solve_surface(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts, int64_t n,
                       const ccab_c128* d_rhs_inc, const ccab_solver_cfg* cfg,
                       ccab_c128* d_phi, ccab_c128* d_vn, ccab_solve_stats* stats);

/* analysis.rs:86-111 sweep loop (surface part) for nfreq frequencies at once:
 * batched assembly + batched solve.  d_phi, d_vn: n x nfreq.  stats[nfreq]. */
int ccab_sweep(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
               const double* d_cpts, int64_t n, const ccab_solver_cfg* cfg,
               ccab_c128* d_phi, ccab_c128* d_vn, ccab_solve_stats* stats);

/* ---- building block ------------------------------------------------------- */
/* C -= A B for complex column-major matrices (M x K times K x N) on MFMA; the
 * trailing-update kernel of the LU (replaces the rank-1 updates of nalgebra's
 * gauss_step behind solve.rs:169).  prec CCAB_C128 or CCAB_C64. */
int ccab_cgemm_update(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t K,
                      const void* d_A, int64_t lda, const void* d_B, int64_t ldb, void* d_C,
                      int64_t ldc);

/* ccab_sweep with solver CCAB_SOLVER_ITERATIVE_MIXED started from d_x0
 * (n x nfreq, the solved unknown of solve.rs:36-69 per frequency), e.g. a
 * learned initial guess (no reference counterpart: the reference always
 * starts GMRES from zero, gmres.rs:38-39). */
int ccab_sweep_warm(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                    const double* d_cpts, int64_t n, const ccab_solver_cfg* cfg,
                    const ccab_c128* d_x0, ccab_c128* d_phi, ccab_c128* d_vn,
                    ccab_solve_stats* stats);
/* solve_for_surface (solve.rs:11-71, with the incident RHS of
 * incident_wave.rs:7-63) for nbody bodies of uneven size in one call: body b
 * has n[b] (host array) collocation points, packed back to back in d_cpts
 * ((sum n) x 8), its own physics ph[b] (frequency; one BC type for all) and
 * incident wave w[b].  d_x0 (optional, packed like the outputs; needs the
 * ITERATIVE_MIXED or HIERARCHICAL solver) holds initial guesses; d_phi and
 * d_vn receive sum(n) values each, stats[b] per body.  With the
 * ITERATIVE_MIXED solver all bodies share one GMRES-IR: its matvecs (one
 * ragged batched GEMV over matrices packed back to back) and Arnoldi steps
 * are single launches over every body; other solvers run body by body.
 * Batches of bodies are formed the way the reference would loop over
 * analyses; the batch itself has no reference counterpart. */
int ccab_solve_bodies(ccab_ctx* ctx, int64_t nbody, const int64_t* n, const double* d_cpts,
                      const ccab_physics* ph, const ccab_wave* w, const ccab_solver_cfg* cfg,
                      const ccab_c128* d_x0, ccab_c128* d_phi, ccab_c128* d_vn,
                      ccab_solve_stats* stats);

/* ---- learned warm start (device; no reference counterpart) ------------------- */
/* A pointwise neural operator: per collocation point, 32 features of the
 * geometry (position about the body centroid, normal), the plane wave
 * (direction) and ka -> 4 linear layers (32 -> W -> W -> W -> 2, SiLU) ->
 * psi = phi / phi_inc, phi_inc = A/(i omega rho) e^{i k d.x}, normalised to
 * unit RMS per system (the shape of the solution).  The initial guess is
 * x0 = phi_inc psi for the rigid / normal-velocity unknown phi
 * (solve.rs:44-52); ccab_gmres_ir_batched fits its complex scale to the
 * FP64 residual.  BF16 GEMMs on MFMA with FP32 accumulation, FP32 Adam.
 * Parameters (FP32, flat): W1[W][32] b1[W] W2[W][W] b2[W] W3[W][W] b3[W]
 * W4[2][W] b4[2], weights row-major (out x in).  Features: 32 bf16 per point
 * (uint16 storage), point p = s*n + i for system s. */
typedef struct ccab_surrogate ccab_surrogate;
typedef struct {
    int32_t width;         /* hidden width: multiple of 64 (<=0 -> 256) */
    int32_t seed;          /* parameter initialisation */
    double lr;             /* Adam step (<=0 -> 1e-3) */
    double beta1, beta2;   /* <=0 -> 0.9, 0.999 */
    double eps;            /* <=0 -> 1e-8 */
    double weight_decay;   /* decoupled, weights only */
} ccab_surrogate_cfg;

int ccab_surrogate_create(ccab_ctx* ctx, const ccab_surrogate_cfg* cfg, ccab_surrogate** out);
int ccab_surrogate_destroy(ccab_surrogate* s);
int64_t ccab_surrogate_num_params(const ccab_surrogate* s);
int ccab_surrogate_get_params(ccab_ctx* ctx, const ccab_surrogate* s, float* h_params);
/* also resets the Adam moments and step count */
int ccab_surrogate_set_params(ccab_ctx* ctx, ccab_surrogate* s, const float* h_params);
int ccab_surrogate_set_lr(ccab_surrogate* s, double lr);
/* features of nsys plane-wave systems on one mesh (ph[s], w[s] per system)
 * into d_feat (nsys*n x 32 bf16); with d_phi (n x nsys, the solved surface
 * potential) also the targets psi / rms_s(psi) into d_target (nsys*n x 2
 * float). */
int ccab_surrogate_features(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w,
                            int64_t nsys, const double* d_cpts, int64_t n,
                            const ccab_c128* d_phi, uint16_t* d_feat, float* d_target);
/* mean over P points of |out - target|^2 and its gradient (no update) */
int ccab_surrogate_loss_grad(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat,
                             const float* d_target, int64_t P, double* loss, float* h_grad);
/* one Adam step on the rows d_index[0..batch) of (d_feat, d_target);
 * *loss (optional, synchronises) = the batch loss before the step */
int ccab_surrogate_train_step(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat,
                              const float* d_target, const int64_t* d_index, int64_t batch,
                              double* loss);
/* network output (P x 2 float) for P feature rows */
int ccab_surrogate_forward(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat, int64_t P,
                           float* d_out);
/* initial guesses x0 (n x nsys) for nsys plane-wave systems on one mesh, for
 * ccab_sweep_warm / ccab_gmres_ir_batched */
int ccab_surrogate_predict(ccab_ctx* ctx, ccab_surrogate* s, const ccab_physics* ph,
                           const ccab_wave* w, int64_t nsys, const double* d_cpts, int64_t n,
                           ccab_c128* d_x0);
/* initial guesses for bodies of uneven size in one pass (features, network
 * and guesses over all packed points), laid out as ccab_solve_bodies takes
 * them: body b has n[b] (host) points packed in d_cpts, physics ph[b] and
 * plane wave w[b]; d_x0 receives sum(n) values */
int ccab_surrogate_predict_bodies(ccab_ctx* ctx, ccab_surrogate* s, int64_t nbody,
                                  const int64_t* n, const double* d_cpts,
                                  const ccab_physics* ph, const ccab_wave* w, ccab_c128* d_x0);

#ifdef __cplusplus
}
#endif
#endif
