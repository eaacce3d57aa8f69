// Internal declarations shared by the HIP translation units of libccab.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/ccab.h"

// ---------------------------------------------------------------------------
// Complex arithmetic (interleaved re/im, same layout as ccab_c128 / ccab_c64)
// ---------------------------------------------------------------------------
template <typename T>
struct Cx {
    T re, im;
};
using c128 = Cx<double>;
using c64 = Cx<float>;
// complex FP16 storage (IEEE half bits), scaled per system: the GMRES-IR
// operator when half the c64 traffic is wanted
struct c32 {
    uint16_t re, im;
};

template <typename T>
__host__ __device__ __forceinline__ Cx<T> cx(T r, T i) {
    return Cx<T>{r, i};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator-(Cx<T> a, Cx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator-(Cx<T> a) {
    return {-a.re, -a.im};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator*(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator*(Cx<T> a, T s) {
    return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator*(T s, Cx<T> a) {
    return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T>& operator+=(Cx<T>& a, Cx<T> b) {
    a.re += b.re;
    a.im += b.im;
    return a;
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T>& operator-=(Cx<T>& a, Cx<T> b) {
    a.re -= b.re;
    a.im -= b.im;
    return a;
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> conj(Cx<T> a) {
    return {a.re, -a.im};
}
template <typename T>
__host__ __device__ __forceinline__ T norm_sqr(Cx<T> a) {
    return a.re * a.re + a.im * a.im;
}
// num-complex division: (a * conj(b)) / |b|^2
template <typename T>
__host__ __device__ __forceinline__ Cx<T> cdiv(Cx<T> a, Cx<T> b) {
    T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> crecip(Cx<T> b) {
    T d = b.re * b.re + b.im * b.im;
    return {b.re / d, -b.im / d};
}
template <typename To, typename Ti>
__host__ __device__ __forceinline__ Cx<To> cvt(Cx<Ti> a) {
    return {static_cast<To>(a.re), static_cast<To>(a.im)};
}

static_assert(sizeof(c128) == sizeof(ccab_c128), "layout");
static_assert(sizeof(c64) == sizeof(ccab_c64), "layout");

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
// Persistent per-slot device buffers.  A slot is owned by one routine; the
// buffer only grows (growing synchronises the stream first).  Returns
// nullptr and records the error on failure.
enum BufSlot {
    SLOT_SOA = 0, SLOT_PARAMS, SLOT_ROWSUM, SLOT_LU_LINV, SLOT_LU_UINV, SLOT_LU_PERM,
    SLOT_LU_INFO, SLOT_LU_Y, SLOT_GEMV_PART, SLOT_NORM, SLOT_MIXED_R, SLOT_MIXED_D,
    SLOT_MIXED_NORMS, SLOT_MIXED_B, SLOT_GMRES, SLOT_SOLVE_A, SLOT_SOLVE_A64, SLOT_SOLVE_RHS,
    SLOT_SOLVE_PIV, SLOT_SOLVE_X, SLOT_IDX, SLOT_GIR_B, SLOT_GIR_R, SLOT_GIR_Q, SLOT_GIR_H,
    SLOT_GIR_Y, SLOT_GIR_NRM, SLOT_GIR_ACT, SLOT_GMRES_NRM, SLOT_HMAT_X, SLOT_SUR_SYS,
    SLOT_SUR_LOSS, SLOT_FIELD_SYS, SLOT_FIELD_PART, SLOT_SOLVE_ISCALE, SLOT_LU_COOP,
    SLOT_RAGGED, SLOT_MGS_COOP, SLOT_GIR_GIV, SLOT_GIR_D, SLOT_INC_AMP, SLOT_HM_UA, SLOT_HM_VA, SLOT_HM_W,
    SLOT_HM_F, SLOT_HM_R, SLOT_BODY_DESC, SLOT_COOP_ABORT, SLOT_AX_PART, SLOT_AX, SLOT_GIR_U, SLOT_GIR_RC, SLOT_COUNT
};
// Sweep tuning (ccab_autotune_sweep / ccab_set_sweep_tuning), per size n:
// the batched assembly's systems per thread and column chunk for an
// operator precision, and the uniform GEMV's workgroup target and
// non-temporal loads for a matrix element precision.
struct AsmTune {
    int prec;          // CCAB_C32 (FP16 operator + FP32 remainder) or CCAB_C64 (c64 + c128)
    int64_t n;
    int kf;            // 8, 16 or 32 systems per thread
    int col_chunk;     // 64, 128 or 256 columns per workgroup
};
struct GemvTune {
    int prec;          // element precision of the matrix: CCAB_C32, CCAB_C64, CCAB_C128
    int64_t n;
    int wg_target;     // workgroups one product aims for
    int nt;            // non-temporal matrix loads
};
// identity of a captured launch sequence (shape, buffers, tuning)
struct GraphKey {
    int prec;
    int64_t n, batch, lda, stride;
    const void *a, *ipiv, *linv, *uinv, *perm, *info, *xbuf;
    int outer, variant;
    int64_t coop_rows;
    bool operator==(const GraphKey& o) const {
        return prec == o.prec && n == o.n && batch == o.batch && lda == o.lda &&
               stride == o.stride && a == o.a && ipiv == o.ipiv && linv == o.linv &&
               uinv == o.uinv && perm == o.perm && info == o.info && xbuf == o.xbuf &&
               outer == o.outer && variant == o.variant && coop_rows == o.coop_rows;
    }
};

// identity of a captured GMRES-IR cycle (see gmres_ir_run): sizes, every
// buffer and scalar its kernels take, and the settings that shape launches
struct GirKey {
    int64_t n, B, na;
    int m, op_prec, ref_mgs, mgs_coop, wg_target, nt, has_tune;
    const void *Aop, *inv_scale, *act, *Q, *Hd, *gv, *Yd, *X, *Zd, *Dv, *R, *nrm, *bn, *pinned,
        *rg_n, *rg_aoff;
    int64_t lda, sA;
    double inner_tol, tol;
    bool operator==(const GirKey& o) const {
        return n == o.n && B == o.B && na == o.na && m == o.m && op_prec == o.op_prec &&
               ref_mgs == o.ref_mgs && mgs_coop == o.mgs_coop && wg_target == o.wg_target &&
               nt == o.nt && has_tune == o.has_tune && Aop == o.Aop && inv_scale == o.inv_scale &&
               act == o.act && Q == o.Q && Hd == o.Hd && gv == o.gv && Yd == o.Yd && X == o.X &&
               Zd == o.Zd && Dv == o.Dv && R == o.R && nrm == o.nrm && bn == o.bn &&
               pinned == o.pinned && rg_n == o.rg_n && rg_aoff == o.rg_aoff && lda == o.lda &&
               sA == o.sA && inner_tol == o.inner_tol && tol == o.tol;
    }
};

struct ccab_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    std::string err;
    // grow-only device buffers, one per slot (see enum BufSlot)
    void* buf[SLOT_COUNT] = {};
    size_t buf_bytes[SLOT_COUNT] = {};
    // identity of the factors whose diagonal-block inverses and row
    // permutation are held in the LU slots
    const void* lu_key = nullptr;
    int lu_prec = -1;
    int64_t lu_n = 0, lu_batch = 0, lu_lda = 0, lu_stride = 0;
    // pinned host staging for small readbacks
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    int num_cus = 256;
    size_t lds_per_block = 65536;  // the device's largest workgroup LDS allocation
    // complex-GEMM variant per precision [c128, c64] (see lu.hip); set by
    // ccab_set_tuning / ccab_autotune, initialised from CCAB_GEMM_VARIANT
    int gemm_variant[2] = {3, 0};
    // rocBLAS handle of the library trailing update (GEMM variant 8), made on
    // first use; a void* so that this header needs no rocBLAS include
    void* blas = nullptr;
    // LU outer panel width (64 or 128; 128 halves the trailing-update traffic)
    int lu_outer = 128;
    // panels of at least this many rows are factored by many workgroups per
    // system when the batch is small (0: never); ccab_set_lu_coop
    int64_t lu_coop_rows = 2048;
    // Gram-Schmidt of few large systems over many workgroups per system
    // (environment CCAB_MGS_COOP=0 keeps one workgroup per system)
    bool mgs_coop = true;
    // the cooperative Gram-Schmidt's barrier headers: buffer and how many
    // system slots are known clear
    void* mgs_xbuf = nullptr;
    int64_t mgs_xbuf_ready = 0;
    // set (never cleared) by a cooperative Gram-Schmidt step whose grid
    // barrier timed out; the host reads and clears it when a NaN shows up
    uint64_t* coop_abort = nullptr;
    // GMRES-IR: the host applies the device's done flags this many Arnoldi
    // steps late (environment CCAB_GIR_LAG, 1-8)
    int gir_lag = 1;
    // GEMV: workgroups per launch the column split aims for
    int gemv_wg_target = 1024;
    // GEMV matrix loads non-temporal (the matrix streams through once per
    // product; x, the partial sums and the Krylov basis keep the caches);
    // environment CCAB_GEMV_NT=0 turns it off
    bool gemv_nt = true;
    int gemv_ragged_chunk = 512;  // columns per workgroup of ragged products (0: from the workgroup target)
    // H-matrix product: persistent waves per CU (0: one item per wave);
    // environment CCAB_HMV_WAVES_PER_CU
    int hmv_waves_per_cu = 0;
    // H-matrix build: columns per product segment (0: whole clusters) and
    // items by decreasing work (CCAB_HMV_NODE_BUDGET, CCAB_HMV_SORT)
    int64_t hmv_node_budget = 512;
    bool hmv_sort = true;
    // FP32 H-matrix storage: bit 0 two columns per 16-byte load in V^T x,
    // bit 1 two rows per 16-byte load in the product (CCAB_HMV_PAIRED)
    int hmv_paired = 3;
    // warm starts of the batched GMRES-IR (sweeps, uneven bodies): A x0 formed
    // by the assembly as it forms the entries (CCAB_WARM_AX=0: a product
    // over the stored matrices after the assembly)
    bool warm_ax = true;
    // batched GMRES-IR: Gram-Schmidt over basis vectors 8 at a time (classical
    // within a block, modified across blocks) on the multi-workgroup path;
    // 4, or 1: one at a time (CCAB_GS_BLOCK)
    int gs_block = 8;
    // batched GMRES-IR: refinement cycles after the first run GCRO, deflated
    // by this many directions of the first cycle's Krylov space (C = A U
    // from its Arnoldi relation; CCAB_GIR_RECYCLE, 0: every cycle restarts
    // from scratch)
    int gir_recycle = 8;
    // the deflation directions: harmonic Ritz vectors of the first cycle
    // (restart lengths up to 32, up to 8 vectors); 0: the first directions
    // of A K_k (CCAB_GIR_RITZ)
    bool gir_ritz = true;
    // FP32 H-matrix storage: ranks per pass of V^T x over paired loads, 4 or
    // 8 (CCAB_HMV_VT_RANKS)
    int hmv_vt_ranks = 8;
    // non-temporal loads of the H-matrix blocks in the product (CCAB_HMV_NT)
    bool hmv_nt = true;
    std::vector<AsmTune> asm_tune;
    std::vector<GemvTune> gemv_tune;
    // device bytes a sweep or a batch of bodies may hold at once (0: 70% of
    // the free memory, at most 96 GiB); environment CCAB_MEM_BUDGET_MB
    size_t mem_budget = 0;
    hipEvent_t wait_event = nullptr;  // ccab_stream_wait
    hipEvent_t join_event = nullptr;  // ccab_stream_join
    std::vector<hipEvent_t> gir_events;  // GMRES-IR: per-step convergence-flag copies
    hipStream_t side_stream = nullptr;  // default priority; work that overlaps the main stream
    hipEvent_t side_ev[2] = {nullptr, nullptr};
    // replay captured HIP graphs of fixed launch sequences (the LU)
    bool use_graphs = false;
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
    std::vector<std::pair<GirKey, hipGraphExec_t>> gir_graphs;  // GMRES-IR cycles
    // in-process kernel timing (HIP events on ctx->stream), see prof_begin
    bool prof_on = false;
    struct ProfRec {
        hipEvent_t a, b;
        int cls;
        double work;
    };
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> event_pool;
    double prof_ms[16] = {};
    double prof_work[16] = {};
    int64_t prof_count[16] = {};
};

// kernel classes of the profiler (ccab_profile_read); work is algorithmic:
// flops for the LU classes, bytes for HBM-bound classes
enum ProfClass {
    PROF_ASSEMBLY = 0,   // bytes stored
    PROF_LU_PANEL = 1,   // flops
    PROF_LU_TRSM = 2,    // flops
    PROF_LU_UPDATE = 3,  // flops (trailing complex GEMM)
    PROF_LU_LASWP = 4,
    PROF_LU_TRINV = 5,
    PROF_TRSV = 6,       // bytes of L/U read
    PROF_GEMV = 7,       // bytes of matrix read
    PROF_MGS = 8,
    PROF_FIELD = 9,      // (field point, surface point) pairs
    PROF_HMAT = 10,      // bytes of H-matrix blocks read
    PROF_SURROGATE = 11, // flops of the network
    PROF_GEMV128 = 12,   // bytes of c128 matrix read (FP64 residuals)
    PROF_NCLASS = 13
};
void prof_begin(ccab_ctx* ctx, hipEvent_t* ev);
void prof_end(ccab_ctx* ctx, hipEvent_t ev, int cls, double work);
// correct the algorithmic work of a class once it is known (launches whose
// systems turned out to be done on the device did none)
void prof_adjust(ccab_ctx* ctx, int cls, double dwork);
struct ProfScope {
    ccab_ctx* ctx;
    hipEvent_t ev = nullptr;
    int cls;
    double work;
    ProfScope(ccab_ctx* c, int k, double w) : ctx(c), cls(k), work(w) {
        if (ctx->prof_on) prof_begin(ctx, &ev);
    }
    ~ProfScope() {
        if (ctx->prof_on && ev) prof_end(ctx, ev, cls, work);
    }
};

int ctx_fail(ccab_ctx* ctx, int code, const char* fmt, ...);

#define CCAB_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ctx_fail(ctx, CCAB_EHIP, "%s failed: %s (%s:%d)", #call,             \
                            hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define CCAB_CHECK(cond, ...)                                                           \
    do {                                                                                \
        if (!(cond)) return ctx_fail(ctx, CCAB_EINVAL, __VA_ARGS__);                    \
    } while (0)

#define CCAB_TRY(expr)                                                                  \
    do {                                                                                \
        int rc_ = (expr);                                                               \
        if (rc_ != CCAB_OK) return rc_;                                                 \
    } while (0)

void* ctx_buf(ccab_ctx* ctx, int slot, size_t bytes);
// captured graphs: lookup (nullptr if absent) and insert (keeps the last 8)
hipGraphExec_t graph_find(ccab_ctx* ctx, const GraphKey& k);
void graph_store(ccab_ctx* ctx, const GraphKey& k, hipGraphExec_t e);
void graph_clear(ccab_ctx* ctx);
// lu.hip: destroys ctx->blas
void blas_release(ccab_ctx* ctx);
inline size_t ws_align(size_t b) { return (b + 255) & ~size_t(255); }
int pinned_reserve(ccab_ctx* ctx, size_t bytes);

// ---------------------------------------------------------------------------
// Physics constants per system (preprocess.rs:41-74, influence_matrix.rs:130-162)
// ---------------------------------------------------------------------------
struct SysParams {
    double k;            // wavenumber
    double omega;
    c128 beta;           // Burton-Miller coupling i/k
    c128 hdiag, gdiag;   // global diagonal shifts
    c128 lhs_a, lhs_b;   // LHS = a H + b G
    c128 rhs_a, rhs_b;   // RHS factors (get_rhs_factors)
    c128 bcv;            // BC value
    c128 rhs_scale;      // multiplies the row sum of the RHS-side matrix
    int32_t bm;
    int32_t rowsum_of;   // 0 none, 1 sum of G, 2 sum of H
    double rhs_inc_sign; // sign applied to rhs_inc
    double h_scale;      // factor of the FP16 (c32) copy of the LHS (power of two)
    double h_inv;        // 1 / h_scale (exact: set with h_scale by set_h_scale)
    double dh_sign;      // sign of the (k^2 r + 3 f1)(e.n_x)(e.n_y) term of dh
    c128 vn_factor;      // impedance BC: vn = vn_factor * phi
    int32_t corrections; // CCAB_FIX_* bits
    int32_t pad_;
    // lhs_a k, lhs_a beta.im, lhs_b beta.im (set by make_params): the LHS
    // entry lhs_a H + lhs_b G as one complex product per entry
    c128 lhs_ak, lhs_aik, lhs_bik;
};
SysParams make_params(const ccab_physics& ph);
inline void set_h_scale(SysParams& p, double s) {
    p.h_scale = s;
    p.h_inv = 1.0 / s;
}
int check_physics(ccab_ctx* ctx, const ccab_physics& ph);

constexpr double kPi = 3.14159265358979323846;
// GEMM variants 0..7 are lu.hip's MFMA kernel (tile order, epilogue, tile
// height); 8 sends the LU trailing update C -= A B to rocBLAS
constexpr int kGemmLibVariant = 8;

// ---------------------------------------------------------------------------
// Internal launchers (all enqueue on ctx->stream)
// ---------------------------------------------------------------------------
// assembly.hip
int launch_assemble(ccab_ctx* ctx, const SysParams* d_params, const SysParams* h_params,
                    int64_t batch, const double* d_soa, int64_t n, int prec_compute, int mode,
                    void* out0, void* out1, int64_t ld, int64_t stride, double* d_rowsum_part,
                    int64_t* nchunks_out, void* out2 = nullptr, const c128* d_x0 = nullptr,
                    double* d_ax_part = nullptr);
// ax[b n + i] = the sum over column chunks of the assembly's A x0 partials
int launch_ax_finish(ccab_ctx* ctx, int64_t batch, int64_t n, const double* d_ax_part,
                     int64_t nchunks, c128* d_ax);
int launch_rhs_finish(ccab_ctx* ctx, const SysParams* d_params, int64_t batch, int64_t n,
                      const double* d_rowsum_part, int64_t nchunks, const ccab_c128* d_rhs_inc,
                      ccab_c128* d_rhs);
int launch_cpts_to_soa(ccab_ctx* ctx, const double* d_cpts, int64_t n, double* d_soa);
// One body of a ragged batch (bodies of uneven size in one call): its
// unknowns, first point (collocation records and SoA, relative to the
// batch's first body), first matrix entry (n x n column-major, ld n), first
// assembly workgroup and first row-sum partial; one more entry ends the list
// (wg0 = total workgroups).  Vectors of body c sit at c * nmax.
struct BodyDesc {
    int64_t n, pts, aoff, wg0, part_off, pad_;
};
struct BodyWave {
    int32_t kind, pad_;
    double d[3];
    c128 amp;
};
// SoA points, incident right-hand sides (X, zero-filled, stride nmax) and
// optional guesses X0 of every body: three launches for the batch
// per system, the largest modulus of the LHS diagonal (d_out: nsys doubles)
// capi.hip: per-call staging of the points (SoA, context buffer) and of the
// systems' parameters (host copy hp, device copy *dp)
int stage_soa(ccab_ctx* ctx, const double* d_cpts, int64_t n, double** soa);
int stage_params(ccab_ctx* ctx, const ccab_physics* ph, int64_t batch, std::vector<SysParams>& hp,
                 SysParams** dp);
// tune.hip: the tuned entries for (prec, n), or null
const AsmTune* asm_tune_of(const ccab_ctx* ctx, int prec, int64_t n);
const GemvTune* gemv_tune_of(const ccab_ctx* ctx, int prec, int64_t n);
int launch_diag_max(ccab_ctx* ctx, int64_t nsys, const SysParams* d_params, const double* d_cpts,
                    const BodyDesc* d_desc, int64_t n, double* d_out);
int launch_bodies_setup(ccab_ctx* ctx, int64_t nbody, int64_t nmax, const BodyDesc* d_desc,
                        const double* d_cpts, double* soa, const SysParams* d_params,
                        const BodyWave* d_waves, ccab_c128* X, const ccab_c128* d_x0,
                        ccab_c128* X0);
// fused LHS assembly and BC algebra of every body in one launch, then their
// right-hand sides (row sums when the BC needs them): prec PREC_F64 or PREC_F64_H
int launch_assemble_bodies(ccab_ctx* ctx, int64_t nbody, int64_t nmax, const BodyDesc* d_desc,
                           const BodyDesc* h_desc, const SysParams* d_params,
                           const SysParams* h_params, const double* soa, int prec, void* out0,
                           void* out1, void* out2, double* part, ccab_c128* X,
                           const c128* X0 = nullptr, double* ax_part = nullptr,
                           c128* AX = nullptr);
// every frequency of a sweep in one launch (d_params: the device copy of h_params)
int launch_incident_sweep(ccab_ctx* ctx, const SysParams* d_params, const SysParams* h_params,
                          const ccab_physics* ph, int64_t nfreq, const ccab_wave& w,
                          const double* d_soa, int64_t n, ccab_c128* d_rhs);
int launch_incident(ccab_ctx* ctx, const SysParams* h_params, const ccab_physics& ph,
                    const ccab_wave& w, const double* d_soa, int64_t n, ccab_c128* d_rhs);
int launch_rowcol(ccab_ctx* ctx, const SysParams& p, const double* d_soa, int64_t n,
                  const int64_t* d_i, int64_t ni, const int64_t* d_j, int64_t nj, int side,
                  ccab_c128* d_out);

enum { ASM_HG = 0, ASM_LHS = 1 };
// assembly arithmetic / second-output type: FP64 (+ c64 copy), FP32 (c64),
// FP64 with a scaled FP16 (c32) copy of the LHS
enum { PREC_F64 = 0, PREC_F32 = 1, PREC_MIXED = 2, PREC_F64_H = 3 };

// lu.hip
int lu_factor(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, void* d_A, int64_t lda,
              int64_t strideA, int32_t* d_ipiv);
int lu_solve(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, const void* d_LU, int64_t lda,
             int64_t strideA, const int32_t* d_ipiv, void* d_B, int64_t ldb, int64_t strideB,
             int64_t nrhs);

// blas.hip
// y_b = alpha * A_b x_b + beta * y_b for each system b (A c128 or c64, x,y c128)
int gemv_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, const void* d_A,
                 int64_t lda, int64_t strideA, const c128* d_x, int64_t stridex, c128 alpha,
                 c128 beta, c128* d_y, int64_t stridey);
size_t gemv_workspace_bytes(int64_t n, int64_t batch);
// the same over the systems listed in d_active (device int32[nact])
// Systems of uneven size in one batch: system b is n[b] x n[b], column-major
// with leading dimension n[b], at element offset aoff[b] of the packed
// matrix array (device arrays); its vectors keep the uniform stride of the
// largest system, zero beyond n[b].  align: the largest of 4, 2, 1 that
// divides every n[b] and aoff[b] (it selects the 16-byte-load GEMV).
struct Ragged {
    const int64_t* n = nullptr;
    const int64_t* aoff = nullptr;
    int align = 1;
    int64_t n_min = 0;
    const int64_t* hn = nullptr;  // host copy of n
    double elems = 0.0;           // matrix entries one product reads (profiler work)
};

// prec CCAB_C32: d_inv_scale[b] multiplies system b's product (its matrix is
// stored scaled by 1 / d_inv_scale[b]).  rg (square products only): the
// matrices are ragged (lda and strideA unused); rows of y beyond n[b] get
// beta * y.
int gemv_active(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t nact,
                const int32_t* d_active, const void* d_A, int64_t lda, int64_t strideA,
                const c128* d_x, int64_t stridex, c128 alpha, c128 beta, c128* d_y,
                int64_t stridey, const double* d_inv_scale = nullptr,
                const Ragged* rg = nullptr, const c64* d_Alo = nullptr,
                const int32_t* d_skip = nullptr, int prof_cls = -1,
                bool reserve_only = false);  // only grow the workspace this call needs
int gemv_general(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t batch, const void* d_A,
                 int64_t lda, int64_t strideA, const c128* d_x, int64_t stridex, c128 alpha,
                 c128 beta, c128* d_y, int64_t stridey);
int nrm2_active(ccab_ctx* ctx, int64_t n, int64_t nact, const int32_t* d_active,
                const c128* d_v, int64_t stride, double* d_out);
// out[b] = ||v_b||_2 (FP64), written to device
int nrm2_batched(ccab_ctx* ctx, int64_t n, int64_t batch, const c128* d_v, int64_t stride,
                 double* d_out);
int frob_norm(ccab_ctx* ctx, int prec, int64_t n, const void* d_A, int64_t lda, double* d_out);
int convert_c128_to_c64(ccab_ctx* ctx, int64_t count, const c128* src, c64* dst);
int convert_c64_to_c128(ccab_ctx* ctx, int64_t count, const c64* src, c128* dst);

// solver.hip
// y = alpha A x + beta y for the operator GMRES iterates on
using MatvecFn = std::function<int(const c128* x, c128 alpha, c128 beta, c128* y)>;
int gmres_op_run(ccab_ctx* ctx, int64_t n, const MatvecFn& apply, double anorm, c128* d_x,
                 const c128* d_x0, const ccab_gmres_params& p, ccab_solve_stats* st);
// CCAB_EHIP with a message if a cooperative Gram-Schmidt barrier timed out
// since the last check (else CCAB_OK); called where a NaN is reported
int coop_abort_check(ccab_ctx* ctx);
// every workgroup of the multi-workgroup (grid-barrier) kernels fits on a CU
// beside nothing else: checked once per context, the paths are off otherwise
bool mgs_coop_fits();
bool lu_coop_fits();
// field.hip: the device address of ccab_field_noalloc's scratch array, looked
// up once per device (the device must be current)
void field_scratch_init(int device);

// hmatrix.hip
int hmatrix_build(ccab_ctx* ctx, const SysParams& p, const double* d_soa, const double* h_cpts,
                  int64_t n, int side, int64_t leaf, double eta, double tol, ccab_hmatrix** out,
                  int prec = CCAB_C128);  // CCAB_C64: blocks stored in complex FP32
int hmatrix_gemv(ccab_ctx* ctx, const ccab_hmatrix* H, c128 alpha, const c128* x, c128 beta,
                 c128* y);
int hmatrix_to_full(ccab_ctx* ctx, const ccab_hmatrix* H, c128* A, int64_t lda);
void hmatrix_stats(const ccab_hmatrix* H, int64_t* st, double* norm);
void hmatrix_stats_ex(const ccab_hmatrix* H, int64_t* st, int nst);
int hmatrix_build_matrix(ccab_ctx* ctx, const c128* d_A, int64_t lda, const double* h_xyz,
                         int64_t n, int64_t leaf, double eta, double tol, ccab_hmatrix** out,
                         int prec);
int aca_build_matrix(ccab_ctx* ctx, const c128* d_A, int64_t lda, int64_t m, int64_t ncol,
                     double tol, ccab_hmatrix** out, int prec);
int64_t hmatrix_ncol(const ccab_hmatrix* H);
void hmatrix_free(ccab_hmatrix* H);
int64_t hmatrix_n(const ccab_hmatrix* H);
int hmatrix_device(const ccab_hmatrix* H);
double hmatrix_norm(const ccab_hmatrix* H);
