/*
 * ccab_oracle.h -- CPU FP64 restatement of the cc_abstemious collocation-BEM
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP product is compared against (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  Nothing in the product links or calls it.
 *
 * Every function restates one piece of the reference (Rust crate
 * dbartilson/cc_abstemious v1.2.0); the file:line it follows is given beside
 * each declaration.  All matrices are column-major complex128 (interleaved
 * re/im), like nalgebra's DMatrix<Complex<f64>>.
 *
 * Pinning (what makes this oracle trustworthy, not a retyped formula):
 *  - orc_solve_surface_dense + orc_field reproduce the reference's own
 *    end-to-end fixtures src/tests/verification.rs:117-119 (classical LU),
 *    :146-147 (GMRES), :202-204 (Burton-Miller LU);
 *  - orc_hmatrix_* reproduce verification.rs:174-175 (H-matrix/ACA);
 *  - orc_rng_* + orc_lu_solve reproduce src/solve.rs:216-217 (x[0] of the
 *    seeded 100x100 system), orc_gmres reproduces src/solve.rs:227-228;
 *  - cluster/block trees reproduce cluster.rs:138 and block.rs:123,145;
 *  - geometry, kernel, LU, GMRES, ACA are additionally checked against
 *    closed forms / LAPACK / truncated SVD in tests/test_oracle.py.
 */
#ifndef CCAB_ORACLE_H
#define CCAB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c128;

/* input_data.rs:47-51, 407-417 */
enum { ORC_CLASSICAL = 0, ORC_BURTON_MILLER = 1 };
enum { ORC_EXTERIOR = 0, ORC_INTERIOR = 1 };
enum { ORC_BC_PRESSURE = 0, ORC_BC_NORMAL_VELOCITY = 1, ORC_BC_IMPEDANCE = 2 };
enum { ORC_SOLVER_DIRECT = 0, ORC_SOLVER_ITERATIVE = 1, ORC_SOLVER_HIERARCHICAL = 2 };
enum { ORC_WAVE_PLANE = 0, ORC_WAVE_SPHERICAL = 1 };

typedef struct {
    double frequency;      /* Hz */
    double sound_speed;
    double mass_density;
    int32_t method;        /* ORC_CLASSICAL / ORC_BURTON_MILLER */
    int32_t problem_type;  /* ORC_EXTERIOR / ORC_INTERIOR */
    int32_t bc_type;       /* ORC_BC_* */
    int32_t pad_;
    double bc_value[2];
} orc_physics;

typedef struct {
    int32_t kind;          /* ORC_WAVE_PLANE / ORC_WAVE_SPHERICAL */
    int32_t pad_;
    double vec[3];         /* plane: direction, spherical: origin */
    double amplitude[2];   /* pressure amplitude (re, im) */
} orc_wave;

/* Collocation point record, 8 doubles: x y z nx ny nz area wt
 * (mesh_data.rs:71-77; wt*area is the element weight). */
#define ORC_CPT_STRIDE 8

/* elements.rs:49-181 + preprocess.rs:126-153: one centroid collocation point
 * per element.  conn: 4 node ids per element (-1 padding for triangles),
 * nverts: 3 (tri) or 4 (quad). */
int orc_collocation_points(const double* nodes, const int64_t* conn,
                           const int32_t* nverts, int64_t nelem, double* cpts);

/* influence_matrix.rs:12-37 (get_greens_functions), k real, beta = i/k. */
void orc_greens(double k, const double* x, const double* nx, const double* y,
                const double* ny, int bm, orc_c128 beta, orc_c128* g, orc_c128* h);

/* influence_matrix.rs:39-60 (get_gh_functions), no global diag shifts. */
void orc_gh_entry(const orc_physics* ph, const double* cpts, int64_t i, int64_t j,
                  orc_c128* g, orc_c128* h);

/* preprocess.rs:43,58,60-74 */
void orc_constants(const orc_physics* ph, double* k, orc_c128* beta,
                   orc_c128* hdiag, orc_c128* gdiag);

/* influence_matrix.rs:64-98.  H, G: n x n column-major.  nthreads<=0 means
 * max(nproc/2,2) (preprocess.rs:12-17); locked!=0 takes one global mutex per
 * entry as the reference does (influence_matrix.rs:87-90). */
int orc_dense_surface_matrices(const orc_physics* ph, const double* cpts, int64_t n,
                               orc_c128* H, orc_c128* G, int nthreads, int locked);

/* influence_matrix.rs:101-127: M (nfp x n), L (nfp x n), column-major. */
int orc_dense_field_matrices(const orc_physics* ph, const double* cpts, int64_t n,
                             const double* fps, int64_t nfp, orc_c128* M, orc_c128* L);

/* influence_matrix.rs:130-162 */
void orc_lhs_factors(const orc_physics* ph, orc_c128* alpha, orc_c128* beta);
void orc_rhs_factors(const orc_physics* ph, orc_c128* alpha, orc_c128* beta);

/* influence_matrix.rs:170-231.  side 0 = LHS, 1 = RHS.  Exactly one of ni, nj
 * must be 1; returns -1 otherwise. out has max(ni,nj) entries. */
int orc_surface_row_or_column(const orc_physics* ph, const double* cpts, int side,
                              const int64_t* ii, int64_t ni, const int64_t* jj,
                              int64_t nj, orc_c128* out);

/* incident_wave.rs:7-63 and 66-100 */
int orc_incident_surface(const orc_physics* ph, const orc_wave* w, const double* cpts,
                         int64_t n, orc_c128* rhs);
int orc_incident_field(const orc_physics* ph, const orc_wave* w, const double* fps,
                       int64_t nfp, orc_c128* phi);

/* solve.rs:166-171 (nalgebra 0.33.2 LU: unblocked right-looking, partial
 * pivoting on norm1 = |re|+|im|, full row swaps).  A is not modified; x holds
 * b on entry and the solution on exit.  Returns 1 if a zero pivot was met. */
int orc_lu_solve(int64_t n, const orc_c128* A, orc_c128* x);
/* Same factorisation in place (A <- LU, piv[i] = row swapped with i). */
int orc_lu_factor(int64_t n, orc_c128* A, int64_t lda, int64_t* piv);
void orc_lu_apply(int64_t n, const orc_c128* LU, int64_t lda, const int64_t* piv, orc_c128* x);

/* gmres.rs:35-202.  Dense operator.  ref_mgs=1 uses the reference's Arnoldi
 * coefficient qk1.dotc(q_i) = conj(q_i^H w) (gmres.rs:168); 0 = textbook.
 * x: b on entry, solution on exit.  stats (may be NULL): [0]=num_mv,
 * [1]=inner iterations, [2]=restarts used; err_out = last backward error. */
typedef void (*orc_matvec_fn)(void* user, const orc_c128* x, orc_c128* y); /* y = A x */
int orc_gmres(int64_t n, const orc_c128* A, orc_c128* x, int64_t max_it, double thresh,
              int ref_mgs, int64_t* stats, double* err_out);
int orc_gmres_op(int64_t n, orc_matvec_fn mv, void* user, double anorm, orc_c128* x,
                 const orc_c128* x0, int64_t max_it, double thresh, int ref_mgs,
                 int64_t* stats, double* err_out);

/* solve.rs:24-71 + 153-171: get_surface_dense.  solver: ORC_SOLVER_DIRECT or
 * ORC_SOLVER_ITERATIVE (max_it, tol). */
int orc_solve_surface_dense(const orc_physics* ph, const double* cpts, int64_t n,
                            const orc_c128* rhs_inc, int solver, int64_t max_it,
                            double tol, orc_c128* phi, orc_c128* vn, int64_t* stats);

/* solve.rs:173-185 + postprocess.rs:70-88 (p = i w rho phi when to_pressure). */
int orc_field(const orc_physics* ph, const double* cpts, int64_t n, const double* fps,
              int64_t nfp, const orc_c128* phi, const orc_c128* vn,
              const orc_c128* phi_inc_fp, int want_total, int to_pressure, orc_c128* out);

/* rand 0.9 StdRng (ChaCha12, rand_core PCG32 seed_from_u64), used by the
 * reference's tests (solve.rs:194-209) and ACA pivots (aca.rs:64-68). */
typedef struct { uint32_t key[8]; uint64_t counter; uint32_t buf[64]; int32_t idx; } orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint32_t orc_rng_u32(orc_rng* r);
uint64_t orc_rng_u64(orc_rng* r);
double orc_rng_f64(orc_rng* r);
/* Uniform::new_inclusive(low, high) for usize (rand 0.9 UniformUsize, 32-bit mode). */
uint64_t orc_rng_uniform_inclusive(orc_rng* r, uint64_t low, uint64_t high);
/* solve.rs:194-209 */
void orc_generate_random_ab(int64_t m, int64_t n, uint64_t seed, orc_c128* a, orc_c128* b);

/* Cluster / block trees (cluster.rs:19-110, block.rs:18-96).  Flattened block
 * list in the reference's depth-first order.  Output arrays are allocated by
 * the library and freed by orc_blocks_free. */
typedef struct {
    int64_t nblocks;
    int64_t* row_off;   /* nblocks+1 offsets into rows */
    int64_t* col_off;   /* nblocks+1 offsets into cols */
    int64_t* rows;
    int64_t* cols;
    int32_t* admissible;
    double root_diameter;
} orc_blocks;
int orc_build_blocks(const double* coords, int64_t stride, int64_t n, int64_t leaf,
                     double eta, orc_blocks* out);
void orc_blocks_free(orc_blocks* b);
/* block.rs:38-54 check of child path (block.rs:123 test): returns admissible flag
 * of the node reached by following child indices path[0..depth), -1 if absent. */
int orc_block_tree_admissible_at(const double* coords, int64_t stride, int64_t n,
                                 int64_t leaf, double eta, const int32_t* path, int depth);

/* ACA (aca.rs:57-268) over a row/column callback pair. */
typedef void (*orc_rowcol_fn)(void* user, int is_row, int64_t idx, orc_c128* out);
typedef struct {
    int64_t m, n, rank;
    orc_c128* u;  /* m x rank column-major */
    orc_c128* v;  /* n x rank column-major */
    double norm;
    int64_t rank_before_recompression;
} orc_aca;
int orc_aca_build(double tol, int64_t m, int64_t n, orc_rowcol_fn f, void* user, orc_aca* out);
void orc_aca_free(orc_aca* a);
/* aca.rs:181-195: b = alpha*(U V^T) x + beta*b (non-conjugating dot) */
void orc_aca_gemv(const orc_aca* a, orc_c128 alpha, const orc_c128* x, orc_c128 beta, orc_c128* b);

/* HMatrix (h_matrix.rs:127-247) over an explicit entry callback. */
typedef struct orc_hmatrix orc_hmatrix;
/* entry callback: fills out[0..cnt) with A(rows[r], cols[c]) where exactly one
 * of nr, nc is 1 (the reference's get_row_or_column closure). */
typedef void (*orc_rc_fn)(void* user, const int64_t* rows, int64_t nr,
                          const int64_t* cols, int64_t nc, orc_c128* out);
orc_hmatrix* orc_hmatrix_build(int64_t n, orc_rc_fn f, void* user, const double* coords,
                               int64_t stride, int64_t leaf, double tol, double eta);
void orc_hmatrix_free(orc_hmatrix* h);
void orc_hmatrix_gemv(const orc_hmatrix* h, orc_c128 alpha, const orc_c128* x,
                      orc_c128 beta, orc_c128* b);
double orc_hmatrix_norm(const orc_hmatrix* h);
/* stats: [0] admissible blocks, [1] inadmissible, [2] numerator of
 * compression ratio (sum r(m+n) + sum mn), [3] sum of ACA ranks */
void orc_hmatrix_stats(const orc_hmatrix* h, int64_t* stats);
void orc_hmatrix_to_full(const orc_hmatrix* h, orc_c128* A);
/* h_matrix for the BEM operator (solve.rs:74-150). */
int orc_solve_surface_hmatrix(const orc_physics* ph, const double* cpts, int64_t n,
                              const orc_c128* rhs_inc, int64_t max_it, double tol,
                              int64_t leaf, double aca_tol, double eta,
                              orc_c128* phi, orc_c128* vn, int64_t* stats);

#ifdef __cplusplus
}
#endif
#endif
