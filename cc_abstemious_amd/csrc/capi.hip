// extern "C" entry points of libccab (declared in include/ccab.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "ccab_internal.h"
#include "greens.h"

int gmres_run(ccab_ctx* ctx, int64_t n, int prec, const void* d_A, int64_t lda, c128* d_x,
              const c128* d_x0, const ccab_gmres_params& p, ccab_solve_stats* st);
int mixed_run(ccab_ctx* ctx, int64_t n, int64_t batch, const c128* A128, c64* A64, int64_t lda,
              int64_t strideA, c128* d_x, double ir_tol, int ir_max, ccab_solve_stats* st);
int cgemm_update(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t K, const void* A,
                 int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc);
int gmres_ir_run(ccab_ctx* ctx, int64_t n, int64_t B, const c128* A128, const void* Aop,
                 int op_prec, const double* d_inv_scale, int64_t lda, int64_t sA, c128* X,
                 const c128* X0, int m, double inner_tol, double tol, int max_cycles,
                 bool ref_mgs, ccab_solve_stats* st, Ragged* rg = nullptr,
                 const c64* Alo = nullptr, const c128* AX0 = nullptr);

// ---------------------------------------------------------------------------
// context plumbing
// ---------------------------------------------------------------------------
int ctx_fail(ccab_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

hipGraphExec_t graph_find(ccab_ctx* ctx, const GraphKey& k) {
    for (auto& g : ctx->graphs)
        if (g.first == k) return g.second;
    return nullptr;
}

void graph_store(ccab_ctx* ctx, const GraphKey& k, hipGraphExec_t e) {
    if (ctx->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(ctx->graphs.front().second);
        ctx->graphs.erase(ctx->graphs.begin());
    }
    ctx->graphs.emplace_back(k, e);
}

void graph_clear(ccab_ctx* ctx) {
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    ctx->graphs.clear();
    for (auto& g : ctx->gir_graphs) (void)hipGraphExecDestroy(g.second);
    ctx->gir_graphs.clear();
}

void* ctx_buf(ccab_ctx* ctx, int slot, size_t bytes) {
    if (slot < 0 || slot >= SLOT_COUNT) return nullptr;
    bytes = ws_align(std::max<size_t>(bytes, 256));
    if (ctx->buf_bytes[slot] >= bytes) return ctx->buf[slot];
    if (ctx->buf[slot]) {
        (void)hipStreamSynchronize(ctx->stream);
        graph_clear(ctx);  // captured graphs may point into this buffer
        (void)hipFree(ctx->buf[slot]);
        ctx->buf[slot] = nullptr;
        ctx->buf_bytes[slot] = 0;
        if (slot == SLOT_LU_LINV || slot == SLOT_LU_UINV || slot == SLOT_LU_PERM)
            ctx->lu_key = nullptr;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        ctx_fail(ctx, CCAB_ENOMEM, "hipMalloc of %zu bytes failed (slot %d)", bytes, slot);
        return nullptr;
    }
    ctx->buf[slot] = p;
    ctx->buf_bytes[slot] = bytes;
    return p;
}

int pinned_reserve(ccab_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes >= bytes) return CCAB_OK;
    if (ctx->pinned) {
        (void)hipStreamSynchronize(ctx->stream);
        graph_clear(ctx);  // captured GMRES-IR cycles write flags into it
        (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    bytes = std::max<size_t>(bytes, 1 << 16);
    CCAB_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
    ctx->pinned_bytes = bytes;
    return CCAB_OK;
}

static hipEvent_t pool_get(ccab_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(ccab_ctx* ctx, hipEvent_t* ev) {
    *ev = pool_get(ctx);
    if (*ev) (void)hipEventRecord(*ev, ctx->stream);
}

void prof_end(ccab_ctx* ctx, hipEvent_t a, int cls, double work) {
    hipEvent_t b = pool_get(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof_pending.push_back({a, b, cls, work});
}

void prof_adjust(ccab_ctx* ctx, int cls, double dwork) {
    if (ctx->prof_on && cls >= 0 && cls < 16) ctx->prof_work[cls] += dwork;
}

static void prof_resolve(ccab_ctx* ctx) {
    if (ctx->prof_pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess && r.cls >= 0 && r.cls < 16) {
            ctx->prof_ms[r.cls] += ms;
            ctx->prof_work[r.cls] += r.work;
            ctx->prof_count[r.cls] += 1;
        }
        ctx->event_pool.push_back(r.a);
        ctx->event_pool.push_back(r.b);
    }
    ctx->prof_pending.clear();
}

SysParams make_params(const ccab_physics& ph) {
    SysParams p{};
    p.omega = 2.0 * kPi * ph.frequency;
    p.k = p.omega / ph.sound_speed;
    p.beta = c128{0.0, 1.0 / p.k};
    p.bm = ph.method == CCAB_BURTON_MILLER ? 1 : 0;
    p.hdiag = ph.problem_type == CCAB_EXTERIOR ? c128{-0.5, 0.0} : c128{0.0, 0.0};
    p.gdiag = p.bm ? c128{0.5 * p.beta.re, 0.5 * p.beta.im} : c128{0.0, 0.0};
    p.bcv = c128{ph.bc_value[0], ph.bc_value[1]};
    const bool bc_zero = ph.bc_value[0] == 0.0 && ph.bc_value[1] == 0.0;
    switch (ph.bc_type) {
    case CCAB_BC_PRESSURE: {
        // G vn = rhs_inc - H p (solve.rs:36-43); + H p with CCAB_FIX_PRESSURE_SIGN
        const double s = (ph.corrections & CCAB_FIX_PRESSURE_SIGN) ? 1.0 : -1.0;
        p.lhs_a = {0, 0}; p.lhs_b = {1, 0}; p.rhs_a = {s, 0}; p.rhs_b = {0, 0};
        p.rowsum_of = bc_zero ? 0 : 2;
        p.rhs_scale = c128{s * p.bcv.re, s * p.bcv.im};
        p.rhs_inc_sign = 1.0;
        break;
    }
    case CCAB_BC_NORMAL_VELOCITY:
        p.lhs_a = {1, 0}; p.lhs_b = {0, 0}; p.rhs_a = {0, 0}; p.rhs_b = {1, 0};
        p.rowsum_of = bc_zero ? 0 : 1;
        p.rhs_scale = p.bcv;
        p.rhs_inc_sign = -1.0;
        break;
    default:  // impedance: (H + f G) phi = -rhs_inc, f = -i w rho / Z
        p.lhs_a = {1, 0};
        p.lhs_b = cdiv(c128{0.0, -p.omega * ph.mass_density}, p.bcv);
        p.rhs_a = {0, 0}; p.rhs_b = {0, 0};
        p.rowsum_of = 0;
        p.rhs_scale = {0, 0};
        p.rhs_inc_sign = -1.0;
    }
    set_h_scale(p, 1.0);
    p.lhs_ak = c128{p.lhs_a.re * p.k, p.lhs_a.im * p.k};
    p.lhs_aik = c128{p.lhs_a.re * p.beta.im, p.lhs_a.im * p.beta.im};
    p.lhs_bik = c128{p.lhs_b.re * p.beta.im, p.lhs_b.im * p.beta.im};
    p.corrections = ph.corrections;
    p.dh_sign = (ph.corrections & CCAB_FIX_BM_DH) ? 1.0 : -1.0;
    p.vn_factor = (ph.corrections & CCAB_FIX_IMPEDANCE_VN) ? c128{-p.lhs_b.re, -p.lhs_b.im}
                                                           : p.lhs_b;
    return p;
}

int check_physics(ccab_ctx* ctx, const ccab_physics& ph) {
    CCAB_CHECK(ph.frequency > 0.0 && std::isfinite(ph.frequency), "frequency must be > 0");
    CCAB_CHECK(ph.sound_speed > 0.0 && std::isfinite(ph.sound_speed), "sound_speed must be > 0");
    CCAB_CHECK(std::isfinite(ph.mass_density), "mass_density must be finite");
    CCAB_CHECK(ph.method == CCAB_CLASSICAL || ph.method == CCAB_BURTON_MILLER, "bad method");
    CCAB_CHECK(ph.problem_type == CCAB_EXTERIOR || ph.problem_type == CCAB_INTERIOR,
               "bad problem type");
    CCAB_CHECK(ph.bc_type >= CCAB_BC_PRESSURE && ph.bc_type <= CCAB_BC_IMPEDANCE, "bad bc type");
    CCAB_CHECK((ph.corrections & ~63) == 0, "unknown correction bits");
    if (ph.bc_type == CCAB_BC_IMPEDANCE)
        CCAB_CHECK(ph.bc_value[0] != 0.0 || ph.bc_value[1] != 0.0, "impedance must be non-zero");
    return CCAB_OK;
}

namespace {

__global__ void fill_kernel(int64_t n, c128 v, ccab_c128* dst, int64_t stride) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[blockIdx.y * stride + i] = ccab_c128{v.re, v.im};
}

__global__ void negate_kernel(int64_t n, ccab_c128* v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = ccab_c128{-v[i].re, -v[i].im};
}

// host copy of the collocation points (the cluster tree is built on the host)
int host_cpts(ccab_ctx* ctx, const double* d_cpts, int64_t n, std::vector<double>& h) {
    h.resize(size_t(n) * 8);
    CCAB_HIP(hipMemcpyAsync(h.data(), d_cpts, size_t(n) * 64, hipMemcpyDeviceToHost, ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

struct HDel {
    void operator()(ccab_hmatrix* h) const { hmatrix_free(h); }
};
using HPtr = std::unique_ptr<ccab_hmatrix, HDel>;

// get_surface_hmatrix (solve.rs:74-150) for every frequency of a sweep; rhs
// holds rhs_inc per frequency on entry and the solved unknown on exit.
int hierarchical_sweep(ccab_ctx* ctx, const std::vector<SysParams>& hp, int bc,
                       const double* soa, const double* d_cpts, int64_t n, int64_t nfreq,
                       const ccab_solver_cfg* cfg, c128* rhs, const ccab_c128* d_x0,
                       ccab_solve_stats* stats) {
    std::vector<double> h;
    CCAB_TRY(host_cpts(ctx, d_cpts, n, h));
    const int64_t leaf = cfg->leaf_size > 0 ? cfg->leaf_size : 32;
    const double tol = cfg->aca_tolerance > 0.0 ? cfg->aca_tolerance : 1e-4;
    const double eta = cfg->eta > 0.0 ? cfg->eta : 4.0;
    c128* tmp = static_cast<c128*>(ctx_buf(ctx, SLOT_HMAT_X, size_t(n) * 16));
    if (!tmp) return CCAB_ENOMEM;
    const dim3 g((unsigned)((n + 255) / 256));
    const c128 one{1.0, 0.0}, mone{-1.0, 0.0};
    // block storage: c128 as the reference, or complex FP32 (matrix_prec C64)
    CCAB_CHECK(cfg->gmres.matrix_prec == CCAB_C128 || cfg->gmres.matrix_prec == CCAB_C64,
               "H-matrix storage is CCAB_C128 or CCAB_C64");
    const int hprec = cfg->gmres.matrix_prec;
    for (int64_t f = 0; f < nfreq; ++f) {
        const SysParams& p = hp[f];
        c128* x = rhs + f * n;
        ccab_c128* xc = reinterpret_cast<ccab_c128*>(x);
        const bool zero_bc = p.bcv.re == 0.0 && p.bcv.im == 0.0;
        if (bc == CCAB_BC_IMPEDANCE) {
            hipLaunchKernelGGL(negate_kernel, g, dim3(256), 0, ctx->stream, n, xc);
        } else if (zero_bc) {
            // the empty H matrix still scales by beta (h_matrix.rs:219-223)
            if (bc == CCAB_BC_NORMAL_VELOCITY)
                hipLaunchKernelGGL(negate_kernel, g, dim3(256), 0, ctx->stream, n, xc);
        } else {
            ccab_hmatrix* hr_raw = nullptr;
            CCAB_TRY(hmatrix_build(ctx, p, soa, h.data(), n, CCAB_RHS, leaf, eta, tol, &hr_raw, hprec));
            HPtr hr(hr_raw);
            hipLaunchKernelGGL(fill_kernel, g, dim3(256), 0, ctx->stream, n, p.bcv,
                               reinterpret_cast<ccab_c128*>(tmp), 0);
            // Pressure: rhs = (rhs_a H) phi + rhs_inc; NormalVelocity: rhs = G vn - rhs_inc
            CCAB_TRY(hmatrix_gemv(ctx, hr.get(), one, tmp,
                                  bc == CCAB_BC_PRESSURE ? one : mone, x));
            CCAB_HIP(hipStreamSynchronize(ctx->stream));
        }
        ccab_hmatrix* hl_raw = nullptr;
        CCAB_TRY(hmatrix_build(ctx, p, soa, h.data(), n, CCAB_LHS, leaf, eta, tol, &hl_raw, hprec));
        HPtr hl(hl_raw);
        MatvecFn op = [&](const c128* xin, c128 a, c128 b, c128* yout) -> int {
            return hmatrix_gemv(ctx, hl.get(), a, xin, b, yout);
        };
        ccab_solve_stats st{};
        const c128* x0 = d_x0 ? reinterpret_cast<const c128*>(d_x0 + f * n) : nullptr;
        CCAB_TRY(gmres_op_run(ctx, n, op, hmatrix_norm(hl.get()), x, x0, cfg->gmres, &st));
        CCAB_HIP(hipStreamSynchronize(ctx->stream));
        if (stats) stats[f] = st;
    }
    return CCAB_OK;
}

// device bytes the matrices of one chunk of systems may take
size_t mem_budget(ccab_ctx* ctx) {
    if (ctx->mem_budget) return ctx->mem_budget;
    size_t budget = size_t(96) << 30;
    size_t freemem = 0, total = 0;
    if (hipMemGetInfo(&freemem, &total) == hipSuccess) budget = std::min(budget, freemem / 10 * 7);
    return budget;
}

}  // namespace

// SoA copy of the collocation points (8 n doubles) for the current call
int stage_soa(ccab_ctx* ctx, const double* d_cpts, int64_t n, double** soa) {
    *soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(n) * 64));
    if (!*soa) return CCAB_ENOMEM;
    return launch_cpts_to_soa(ctx, d_cpts, n, *soa);
}

int stage_params(ccab_ctx* ctx, const ccab_physics* ph, int64_t batch, std::vector<SysParams>& hp,
                 SysParams** dp) {
    hp.resize(batch);
    for (int64_t b = 0; b < batch; ++b) {
        CCAB_TRY(check_physics(ctx, ph[b]));
        hp[b] = make_params(ph[b]);
    }
    *dp = static_cast<SysParams*>(ctx_buf(ctx, SLOT_PARAMS, sizeof(SysParams) * batch));
    if (!*dp) return CCAB_ENOMEM;
    CCAB_TRY(pinned_reserve(ctx, sizeof(SysParams) * batch));
    // stage through pinned memory; the copy completes before pinned is reused
    std::memcpy(ctx->pinned, hp.data(), sizeof(SysParams) * batch);
    CCAB_HIP(hipMemcpyAsync(*dp, ctx->pinned, sizeof(SysParams) * batch, hipMemcpyHostToDevice,
                            ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

namespace {

// phi / vn from the solved unknown x per BC (solve.rs:36-69)
// phi and vn of every system from the solution x (solve.rs:24-71): pressure
// BC phi = value, vn = x; velocity BC phi = x, vn = value; impedance BC phi =
// x, vn = vn_factor x (solve.rs:67, negated with CCAB_FIX_IMPEDANCE_VN).  x may alias phi (velocity, impedance) or vn (pressure).
__global__ void finish_surface_kernel(int64_t n, const SysParams* __restrict__ params, int bc,
                                      const ccab_c128* x, ccab_c128* phi, ccab_c128* vn) {
    const int64_t b = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ccab_c128 xv = x[b * n + i];
    const c128 bcv = params[b].bcv;
    if (bc == CCAB_BC_PRESSURE) {
        phi[b * n + i] = ccab_c128{bcv.re, bcv.im};
        vn[b * n + i] = xv;
    } else if (bc == CCAB_BC_NORMAL_VELOCITY) {
        phi[b * n + i] = xv;
        vn[b * n + i] = ccab_c128{bcv.re, bcv.im};
    } else {
        const c128 v = params[b].vn_factor * c128{xv.re, xv.im};
        phi[b * n + i] = xv;
        vn[b * n + i] = ccab_c128{v.re, v.im};
    }
}

int finish_surface(ccab_ctx* ctx, const SysParams* hp, const SysParams* dp, int bc, int64_t n,
                   int64_t batch, const ccab_c128* x, ccab_c128* phi, ccab_c128* vn) {
    (void)hp;
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = std::min<int64_t>(65535, batch - b0);
        hipLaunchKernelGGL(finish_surface_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nb),
                           dim3(256), 0, ctx->stream, n, dp + b0, bc, x + b0 * n, phi + b0 * n,
                           vn + b0 * n);
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

}  // namespace

extern "C" {

int ccab_version(void) { return CCAB_VERSION; }

int ccab_create(int device, ccab_ctx** out) {
    if (!out) return CCAB_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CCAB_EHIP;
    if (device < 0 || device >= ndev) return CCAB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return CCAB_EHIP;
    ccab_ctx* ctx = new ccab_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CCAB_EHIP;
    }
    ctx->stream = ctx->own_stream;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cus = prop.multiProcessorCount;
        ctx->lds_per_block = prop.sharedMemPerBlock;
    }
    if (const char* gv = getenv("CCAB_GEMM_VARIANT")) {
        const int v = atoi(gv);
        ctx->gemm_variant[0] = ctx->gemm_variant[1] = (v >= 0 && v <= kGemmLibVariant) ? v : 3;
    }
    if (const char* lo = getenv("CCAB_LU_OUTER")) ctx->lu_outer = atoi(lo) == 64 ? 64 : 128;
    if (const char* lc = getenv("CCAB_LU_COOP_ROWS")) ctx->lu_coop_rows = std::max(0, atoi(lc));
    if (const char* mc = getenv("CCAB_MGS_COOP")) ctx->mgs_coop = atoi(mc) != 0;
    if (const char* gl = getenv("CCAB_GIR_LAG")) ctx->gir_lag = std::max(1, std::min(8, atoi(gl)));
    if (const char* mb = getenv("CCAB_MEM_BUDGET_MB"))
        ctx->mem_budget = size_t(std::max(0L, atol(mb))) << 20;
    if (const char* gr = getenv("CCAB_GRAPHS")) ctx->use_graphs = atoi(gr) != 0;
    if (const char* nt = getenv("CCAB_GEMV_NT")) ctx->gemv_nt = atoi(nt) != 0;
    if (const char* rc = getenv("CCAB_GEMV_RAGGED_CHUNK")) ctx->gemv_ragged_chunk = std::max(0, atoi(rc));
    if (const char* hb = getenv("CCAB_HMV_NODE_BUDGET")) ctx->hmv_node_budget = std::max(0L, atol(hb));
    if (const char* hs = getenv("CCAB_HMV_SORT")) ctx->hmv_sort = atoi(hs) != 0;
    if (const char* hp = getenv("CCAB_HMV_PAIRED")) ctx->hmv_paired = atoi(hp) & 3;
    if (const char* hn = getenv("CCAB_HMV_NT")) ctx->hmv_nt = atoi(hn) != 0;
    if (const char* wa = getenv("CCAB_WARM_AX")) ctx->warm_ax = atoi(wa) != 0;
    if (const char* gc = getenv("CCAB_GIR_RECYCLE")) ctx->gir_recycle = std::max(0, std::min(64, atoi(gc)));
    if (const char* gz = getenv("CCAB_GIR_RITZ")) ctx->gir_ritz = atoi(gz) != 0;
    if (const char* gb = getenv("CCAB_GS_BLOCK")) {
        const int v = atoi(gb);
        ctx->gs_block = v == 1 ? 1 : v == 4 ? 4 : 8;
    }
    if (const char* hr = getenv("CCAB_HMV_VT_RANKS")) ctx->hmv_vt_ranks = atoi(hr) == 4 ? 4 : 8;
    if (const char* hw = getenv("CCAB_HMV_WAVES_PER_CU"))
        ctx->hmv_waves_per_cu = std::max(0, std::min(64, atoi(hw)));
    if (const char* gw = getenv("CCAB_GEMV_WG_TARGET"))
        ctx->gemv_wg_target = std::max(64, std::min(65536, atoi(gw)));
    // the grid-barrier kernels assume one resident workgroup per CU: without
    // room for one, their paths stay off
    if (!mgs_coop_fits()) ctx->mgs_coop = false;
    if (!lu_coop_fits()) ctx->lu_coop_rows = 0;
    field_scratch_init(device);
    *out = ctx;
    return CCAB_OK;
}

int ccab_profile_enable(ccab_ctx* ctx, int on) {
    if (!ctx) return CCAB_EINVAL;
    prof_resolve(ctx);
    ctx->prof_on = on != 0;
    return CCAB_OK;
}

int ccab_profile_read(ccab_ctx* ctx, int cls, int64_t* count, double* ms, double* work,
                      int reset) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(cls >= 0 && cls < PROF_NCLASS, "bad profile class %d", cls);
    prof_resolve(ctx);
    if (count) *count = ctx->prof_count[cls];
    if (ms) *ms = ctx->prof_ms[cls];
    if (work) *work = ctx->prof_work[cls];
    if (reset) {
        ctx->prof_count[cls] = 0;
        ctx->prof_ms[cls] = 0.0;
        ctx->prof_work[cls] = 0.0;
    }
    return CCAB_OK;
}

int ccab_set_tuning(ccab_ctx* ctx, int prec, int gemm_variant, int lu_outer) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "bad precision");
    CCAB_CHECK(gemm_variant >= 0 && gemm_variant <= kGemmLibVariant, "gemm variant must be 0..8");
    CCAB_CHECK(lu_outer == 64 || lu_outer == 128, "lu_outer must be 64 or 128");
    ctx->gemm_variant[prec == CCAB_C128 ? 0 : 1] = gemm_variant;
    ctx->lu_outer = lu_outer;
    return CCAB_OK;
}

int ccab_set_lu_coop(ccab_ctx* ctx, int64_t min_rows) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(min_rows >= 0, "min_rows must be >= 0");
    ctx->lu_coop_rows = min_rows;
    return CCAB_OK;
}

int ccab_stream_wait(ccab_ctx* ctx, void* hip_stream) {
    if (!ctx) return CCAB_EINVAL;
    hipStream_t other = static_cast<hipStream_t>(hip_stream);
    if (other == ctx->stream) return CCAB_OK;
    if (!ctx->wait_event) CCAB_HIP(hipEventCreateWithFlags(&ctx->wait_event, hipEventDisableTiming));
    CCAB_HIP(hipEventRecord(ctx->wait_event, other));
    CCAB_HIP(hipStreamWaitEvent(ctx->stream, ctx->wait_event, 0));
    return CCAB_OK;
}

int ccab_stream_join(ccab_ctx* ctx, void* hip_stream) {
    if (!ctx) return CCAB_EINVAL;
    hipStream_t other = static_cast<hipStream_t>(hip_stream);
    if (other == ctx->stream) return CCAB_OK;
    if (!ctx->join_event) CCAB_HIP(hipEventCreateWithFlags(&ctx->join_event, hipEventDisableTiming));
    CCAB_HIP(hipEventRecord(ctx->join_event, ctx->stream));
    CCAB_HIP(hipStreamWaitEvent(other, ctx->join_event, 0));
    return CCAB_OK;
}

int ccab_set_graphs(ccab_ctx* ctx, int on) {
    if (!ctx) return CCAB_EINVAL;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->use_graphs = on != 0;
    if (!ctx->use_graphs) graph_clear(ctx);
    return CCAB_OK;
}

int ccab_set_warm_ax(ccab_ctx* ctx, int on) {
    if (!ctx) return CCAB_EINVAL;
    ctx->warm_ax = on != 0;
    return CCAB_OK;
}

int ccab_set_gir_deflation(ccab_ctx* ctx, int nvec) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(nvec >= 0 && nvec <= 64, "deflation vectors must be 0..64");
    ctx->gir_recycle = nvec;
    return CCAB_OK;
}

int ccab_destroy(ccab_ctx* ctx) {
    if (!ctx) return CCAB_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    graph_clear(ctx);
    blas_release(ctx);
    if (ctx->wait_event) (void)hipEventDestroy(ctx->wait_event);
    if (ctx->join_event) (void)hipEventDestroy(ctx->join_event);
    for (hipEvent_t e : ctx->gir_events) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->side_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->side_stream) {
        (void)hipStreamSynchronize(ctx->side_stream);
        (void)hipStreamDestroy(ctx->side_stream);
    }
    prof_resolve(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (int s = 0; s < SLOT_COUNT; ++s)
        if (ctx->buf[s]) (void)hipFree(ctx->buf[s]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CCAB_OK;
}

const char* ccab_last_error(const ccab_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ccab_set_stream(ccab_ctx* ctx, void* hip_stream) {
    if (!ctx) return CCAB_EINVAL;
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
    return CCAB_OK;
}

int ccab_synchronize(ccab_ctx* ctx) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

// elements.rs:49-181 + preprocess.rs:126-153 (host)
int ccab_collocation_points_ex(const double* nodes, const int64_t* conn, const int32_t* nverts,
                               int64_t nelem, int corrections, double* cpts) {
    if (!nodes || !conn || !nverts || !cpts || nelem < 0) return CCAB_EINVAL;
    const bool quad_area = (corrections & CCAB_FIX_QUAD_AREA) != 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t* cn = conn + 4 * e;
        double* out = cpts + 8 * e;
        double x[3] = {0, 0, 0}, nu[3];
        if (nverts[e] == 3) {
            const double xi = 1.0 / 3.0, eta = 1.0 / 3.0;
            const double N[3] = {1.0 - xi - eta, xi, eta};
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) x[d] += N[a] * nodes[3 * cn[a] + d];
            const double* p0 = nodes + 3 * cn[0];
            const double* p1 = nodes + 3 * cn[1];
            const double* p2 = nodes + 3 * cn[2];
            const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            nu[0] = a[1] * b[2] - a[2] * b[1];
            nu[1] = a[2] * b[0] - a[0] * b[2];
            nu[2] = a[0] * b[1] - a[1] * b[0];
            const double nn = std::sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
            out[6] = 0.5 * nn;
            out[7] = 1.0;
            for (int d = 0; d < 3; ++d) out[3 + d] = nu[d] / nn;
        } else if (nverts[e] == 4) {
            // QUADGP1 with the reference's detj = 0.25 |dxi x deta| (elements.rs:168)
            const double dN[4][2] = {{-0.25, -0.25}, {0.25, -0.25}, {0.25, 0.25}, {-0.25, 0.25}};
            double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0};
            for (int a = 0; a < 4; ++a) {
                const double* p = nodes + 3 * cn[a];
                for (int d = 0; d < 3; ++d) {
                    x[d] += 0.25 * p[d];
                    dxi[d] += dN[a][0] * p[d];
                    deta[d] += dN[a][1] * p[d];
                }
            }
            nu[0] = dxi[1] * deta[2] - dxi[2] * deta[1];
            nu[1] = dxi[2] * deta[0] - dxi[0] * deta[2];
            nu[2] = dxi[0] * deta[1] - dxi[1] * deta[0];
            const double nn = std::sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
            // CCAB_FIX_QUAD_AREA: the one-point Gauss area 4 |dxi x deta| with
            // weight 1 (the reference's wt * area is a quarter of it)
            out[6] = quad_area ? 4.0 * nn : 0.25 * nn;
            out[7] = quad_area ? 1.0 : 4.0;
            for (int d = 0; d < 3; ++d) out[3 + d] = nu[d] / nn;
        } else {
            return CCAB_EINVAL;
        }
        if (!(out[6] > 0.0)) return CCAB_ENAN;  // degenerate element: b = 0 on the diagonal
        out[0] = x[0];
        out[1] = x[1];
        out[2] = x[2];
    }
    return CCAB_OK;
}

int ccab_collocation_points(const double* nodes, const int64_t* conn, const int32_t* nverts,
                            int64_t nelem, double* cpts) {
    return ccab_collocation_points_ex(nodes, conn, nverts, nelem, 0, cpts);
}

int ccab_incident_surface(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w,
                          const double* d_cpts, int64_t n, ccab_c128* d_rhs) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && w && d_cpts && d_rhs && n > 0, "null argument or n <= 0");
    CCAB_CHECK(w->kind == CCAB_WAVE_PLANE || w->kind == CCAB_WAVE_SPHERICAL, "bad wave kind");
    CCAB_TRY(check_physics(ctx, *ph));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    const SysParams p = make_params(*ph);
    return launch_incident(ctx, &p, *ph, *w, soa, n, d_rhs);
}

int ccab_dense_surface_matrices(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                int64_t n, int prec, void* d_H, void* d_G, int64_t ld) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && n > 0 && ld >= n, "invalid arguments");
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "bad precision");
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, 1, hp, &dp));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    return launch_assemble(ctx, dp, hp.data(), 1, soa, n, prec == CCAB_C128 ? PREC_F64 : PREC_F32,
                           ASM_HG, d_H, d_G, ld, 0, nullptr, nullptr);
}

int ccab_assemble_system(ccab_ctx* ctx, const ccab_physics* ph, int64_t batch,
                         const double* d_cpts, int64_t n, ccab_c128* d_lhs128, ccab_c64* d_lhs64,
                         int64_t ld, int64_t stride, const ccab_c128* d_rhs_inc,
                         ccab_c128* d_rhs) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && n > 0 && batch > 0 && ld >= n, "invalid arguments");
    CCAB_CHECK(batch == 1 || stride >= ld * n, "stride too small");
    CCAB_CHECK(d_lhs128 || d_lhs64, "no output matrix");
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, batch, hp, &dp));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    bool need_rowsum = false;
    for (auto& p : hp) need_rowsum |= p.rowsum_of != 0;
    const int64_t nchunks = (n + 63) / 64;
    double* part = nullptr;
    if (need_rowsum && d_rhs) {
        part = static_cast<double*>(ctx_buf(ctx, SLOT_ROWSUM, size_t(batch) * nchunks * n * 16));
        if (!part) return CCAB_ENOMEM;
    }
    const int pc = d_lhs128 ? PREC_F64 : PREC_F64;  // FP64 arithmetic for every fused output
    int64_t nch = 0;
    CCAB_TRY(launch_assemble(ctx, dp, hp.data(), batch, soa, n, pc, ASM_LHS, d_lhs128, d_lhs64, ld,
                             stride, part, &nch));
    if (d_rhs) {
        CCAB_CHECK(d_rhs_inc != nullptr, "rhs requested without rhs_inc");
        CCAB_TRY(launch_rhs_finish(ctx, dp, batch, n, part, nch, d_rhs_inc, d_rhs));
    }
    return CCAB_OK;
}

int ccab_surface_row_or_column(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                               int64_t n, int side, const int64_t* d_i, int64_t ni,
                               const int64_t* d_j, int64_t nj, ccab_c128* d_out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK((ni == 1) != (nj == 1), "Invalid call to get_surface_row_or_column: i or j must be singleton");
    CCAB_CHECK(ph && d_cpts && d_i && d_j && d_out && n > 0, "null argument");
    CCAB_CHECK(side == CCAB_LHS || side == CCAB_RHS, "bad side");
    CCAB_TRY(check_physics(ctx, *ph));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    return launch_rowcol(ctx, make_params(*ph), soa, n, d_i, ni, d_j, nj, side, d_out);
}

int ccab_lu_factor_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, void* d_A,
                           int64_t lda, int64_t strideA, int32_t* d_ipiv) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A && d_ipiv, "null argument");
    return lu_factor(ctx, prec, n, batch, d_A, lda, strideA, d_ipiv);
}

int ccab_lu_solve_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, const void* d_LU,
                          int64_t lda, int64_t strideA, const int32_t* d_ipiv, void* d_B,
                          int64_t ldb, int64_t strideB, int64_t nrhs) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_LU && d_ipiv && d_B, "null argument");
    return lu_solve(ctx, prec, n, batch, d_LU, lda, strideA, d_ipiv, d_B, ldb, strideB, nrhs);
}

int ccab_solve_lu(ccab_ctx* ctx, int64_t n, const ccab_c128* d_A, int64_t lda, ccab_c128* d_x) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A && d_x && n > 0 && lda >= n, "invalid arguments");
    // a_lu = a.clone().lu(): factor a private copy
    c128* W = static_cast<c128*>(ctx_buf(ctx, SLOT_SOLVE_A, size_t(n) * n * 16));
    int32_t* piv = static_cast<int32_t*>(ctx_buf(ctx, SLOT_SOLVE_PIV, size_t(n) * 4));
    if (!W || !piv) return CCAB_ENOMEM;
    CCAB_HIP(hipMemcpy2DAsync(W, n * 16, d_A, lda * 16, n * 16, n, hipMemcpyDeviceToDevice,
                              ctx->stream));
    CCAB_TRY(lu_factor(ctx, CCAB_C128, n, 1, W, n, n * n, piv));
    return lu_solve(ctx, CCAB_C128, n, 1, W, n, n * n, piv, d_x, n, n, 1);
}

int ccab_solve_mixed_batched(ccab_ctx* ctx, int64_t n, int64_t batch, const ccab_c128* d_A128,
                             ccab_c64* d_A64, int64_t lda, int64_t strideA, ccab_c128* d_x,
                             double ir_tol, int ir_max_steps, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A128 && d_A64 && d_x && n > 0 && batch > 0 && lda >= n, "invalid arguments");
    CCAB_CHECK(batch == 1 || strideA >= lda * n, "stride too small");
    return mixed_run(ctx, n, batch, reinterpret_cast<const c128*>(d_A128),
                     reinterpret_cast<c64*>(d_A64), lda, strideA, reinterpret_cast<c128*>(d_x),
                     ir_tol, ir_max_steps, stats);
}

int ccab_gmres_dense(ccab_ctx* ctx, int64_t n, const void* d_A, int64_t lda, ccab_c128* d_x,
                     const ccab_c128* d_x0, const ccab_gmres_params* p, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A && d_x && p && n > 0 && lda >= n, "invalid arguments");
    CCAB_CHECK(p->matrix_prec == CCAB_C128 || p->matrix_prec == CCAB_C64, "bad matrix precision");
    return gmres_run(ctx, n, p->matrix_prec, d_A, lda, reinterpret_cast<c128*>(d_x),
                     reinterpret_cast<const c128*>(d_x0), *p, stats);
}

}  // extern "C"

static int check_residual_matrix(ccab_ctx* ctx, const ccab_solver_cfg& cfg) {
    CCAB_CHECK(cfg.residual_matrix == CCAB_C128 ||
                   (cfg.residual_matrix == CCAB_C32 && cfg.solver == CCAB_SOLVER_ITERATIVE_MIXED &&
                    cfg.gmres.matrix_prec == CCAB_C32),
               "residual_matrix CCAB_C32 needs ITERATIVE_MIXED with the CCAB_C32 operator");
    return CCAB_OK;
}

// Shared by ccab_solve_surface (one frequency, given rhs_inc) and ccab_sweep
// (incident wave evaluated per frequency).
static int sweep_impl(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                      const ccab_c128* d_rhs_inc_given, const double* d_cpts, int64_t n,
                      const ccab_solver_cfg* cfg, ccab_c128* d_phi, ccab_c128* d_vn,
                      ccab_solve_stats* stats, const ccab_c128* d_x0 = nullptr) {
    CCAB_CHECK(ph && d_cpts && cfg && d_phi && d_vn && n > 0 && nfreq > 0, "invalid arguments");
    CCAB_TRY(check_residual_matrix(ctx, *cfg));
    const int bc = ph[0].bc_type;
    for (int64_t f = 1; f < nfreq; ++f) {
        CCAB_CHECK(ph[f].bc_type == bc && ph[f].method == ph[0].method,
                   "a sweep shares one BC type and method");
    }
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, nfreq, hp, &dp));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    // right-hand sides: rhs_inc per frequency, then the fused BC algebra
    c128* rhs = static_cast<c128*>(ctx_buf(ctx, SLOT_SOLVE_RHS, size_t(nfreq) * n * 16));
    if (!rhs) return CCAB_ENOMEM;
    ccab_c128* rhs_c = reinterpret_cast<ccab_c128*>(rhs);
    if (d_rhs_inc_given) {
        CCAB_HIP(hipMemcpyAsync(rhs, d_rhs_inc_given, size_t(n) * 16 * nfreq,
                                hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        CCAB_TRY(launch_incident_sweep(ctx, dp, hp.data(), ph, nfreq, *w, soa, n, rhs_c));
    }
    const int solver = cfg->solver;
    if (solver == CCAB_SOLVER_HIERARCHICAL) {
        CCAB_TRY(hierarchical_sweep(ctx, hp, bc, soa, d_cpts, n, nfreq, cfg, rhs, d_x0, stats));
        CCAB_TRY(finish_surface(ctx, hp.data(), dp, bc, n, nfreq, rhs_c, d_phi, d_vn));
        CCAB_HIP(hipGetLastError());
        return CCAB_OK;
    }
    const size_t n2 = size_t(n) * n;
    const bool half_op = solver == CCAB_SOLVER_ITERATIVE_MIXED && cfg->gmres.matrix_prec == CCAB_C32;
    // FP64 residuals against the FP16 operator + its FP32 remainder
    const bool split_res = half_op && cfg->residual_matrix == CCAB_C32;
    // process frequencies in chunks that fit a memory budget
    size_t per = 0;
    if (solver == CCAB_SOLVER_DIRECT) per = n2 * 16;
    else if (solver == CCAB_SOLVER_DIRECT_MIXED) per = n2 * 24;
    else if (solver == CCAB_SOLVER_ITERATIVE)
        per = n2 * (cfg->gmres.matrix_prec == CCAB_C64 ? 8 : 16);
    else if (solver == CCAB_SOLVER_ITERATIVE_MIXED)
        per = n2 * (split_res ? 12 : cfg->gmres.matrix_prec == CCAB_C32 ? 20 : 24);
    else return ctx_fail(ctx, CCAB_EINVAL, "solver %d not available on the dense path", solver);
    const size_t budget = mem_budget(ctx);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nfreq, int64_t(budget / per)));
    bool need_rowsum = false;
    for (auto& p : hp) need_rowsum |= p.rowsum_of != 0;
    const int64_t nchunks = (n + 63) / 64;
    double* part = nullptr;
    if (need_rowsum) {
        part = static_cast<double*>(ctx_buf(ctx, SLOT_ROWSUM, size_t(chunk) * nchunks * n * 16));
        if (!part) return CCAB_ENOMEM;
    }
    void* A = nullptr;
    c64* A64 = nullptr;
    c64* Alo = nullptr;
    if (split_res) {
        Alo = static_cast<c64*>(ctx_buf(ctx, SLOT_SOLVE_A, size_t(chunk) * n2 * 8));
        if (!Alo) return CCAB_ENOMEM;
    } else if (solver == CCAB_SOLVER_DIRECT || solver == CCAB_SOLVER_DIRECT_MIXED ||
               solver == CCAB_SOLVER_ITERATIVE_MIXED ||
               (solver == CCAB_SOLVER_ITERATIVE && cfg->gmres.matrix_prec == CCAB_C128)) {
        A = ctx_buf(ctx, SLOT_SOLVE_A, size_t(chunk) * n2 * 16);
        if (!A) return CCAB_ENOMEM;
    }
    if (solver == CCAB_SOLVER_DIRECT_MIXED || solver == CCAB_SOLVER_ITERATIVE_MIXED ||
        (solver == CCAB_SOLVER_ITERATIVE && cfg->gmres.matrix_prec == CCAB_C64)) {
        // c64 copy, or the scaled c32 copy for the half-precision GMRES operator
        A64 = static_cast<c64*>(ctx_buf(ctx, SLOT_SOLVE_A64, size_t(chunk) * n2 * (half_op ? 4 : 8)));
        if (!A64) return CCAB_ENOMEM;
    }
    double* inv_scale = nullptr;
    if (half_op) {
        // FP16 scale per system: the largest diagonal entry (which bounds the
        // near-singular entries of the row) maps to about 2^12, 16x below the
        // largest finite half
        // (the diagonal maxima on the device, one number per system back)
        inv_scale = static_cast<double*>(ctx_buf(ctx, SLOT_SOLVE_ISCALE, size_t(nfreq) * 8));
        if (!inv_scale) return CCAB_ENOMEM;
        CCAB_TRY(launch_diag_max(ctx, nfreq, dp, d_cpts, nullptr, n, inv_scale));
        std::vector<double> isc(nfreq);
        CCAB_HIP(hipMemcpyAsync(isc.data(), inv_scale, size_t(nfreq) * 8, hipMemcpyDeviceToHost,
                                ctx->stream));
        CCAB_HIP(hipStreamSynchronize(ctx->stream));
        for (int64_t f = 0; f < nfreq; ++f) {
            const double mx = isc[f];
            CCAB_CHECK(mx > 0.0 && std::isfinite(mx), "degenerate diagonal for the FP16 operator");
            set_h_scale(hp[f], std::ldexp(1.0, (int)std::floor(std::log2(4096.0 / mx))));
            isc[f] = 1.0 / hp[f].h_scale;
        }
        CCAB_HIP(hipMemcpy(dp, hp.data(), sizeof(SysParams) * nfreq, hipMemcpyHostToDevice));
        CCAB_HIP(hipMemcpy(inv_scale, isc.data(), size_t(nfreq) * 8, hipMemcpyHostToDevice));
    }
    int32_t* piv = nullptr;
    if (solver == CCAB_SOLVER_DIRECT) {
        piv = static_cast<int32_t*>(ctx_buf(ctx, SLOT_SOLVE_PIV, size_t(chunk) * n * 4));
        if (!piv) return CCAB_ENOMEM;
    }
    // a warm start's A x0 from the assembly: partials per column chunk, then
    // one vector per system
    // (with row sums as well: 8 systems per thread, 72 KiB of LDS per
    // workgroup, where the device allows it)
    const bool warm_ax = d_x0 && solver == CCAB_SOLVER_ITERATIVE_MIXED && ctx->warm_ax &&
                         (!need_rowsum || ctx->lds_per_block >= 73728);
    double* ax_part = nullptr;
    c128* AX = nullptr;
    if (warm_ax) {
        ax_part = static_cast<double*>(ctx_buf(ctx, SLOT_AX_PART, size_t(chunk) * nchunks * n * 16));
        AX = static_cast<c128*>(ctx_buf(ctx, SLOT_AX, size_t(chunk) * n * 16));
        if (!ax_part || !AX) return CCAB_ENOMEM;
    }
    for (int64_t f0 = 0; f0 < nfreq; f0 += chunk) {
        const int64_t nb = std::min(chunk, nfreq - f0);
        int64_t nch = 0;
        const c128* x0 = d_x0 ? reinterpret_cast<const c128*>(d_x0 + f0 * n) : nullptr;
        CCAB_TRY(launch_assemble(ctx, dp + f0, hp.data() + f0, nb, soa, n,
                                 half_op ? PREC_F64_H : PREC_F64, ASM_LHS, A, A64, n, int64_t(n2),
                                 part, &nch, Alo, warm_ax ? x0 : nullptr, ax_part));
        if (warm_ax) CCAB_TRY(launch_ax_finish(ctx, nb, n, ax_part, nch, AX));
        ccab_c128* x = rhs_c + f0 * n;
        CCAB_TRY(launch_rhs_finish(ctx, dp + f0, nb, n, part, nch, x, x));
        if (solver == CCAB_SOLVER_DIRECT) {
            CCAB_TRY(lu_factor(ctx, CCAB_C128, n, nb, A, n, int64_t(n2), piv));
            CCAB_TRY(lu_solve(ctx, CCAB_C128, n, nb, A, n, int64_t(n2), piv, x, n, n, 1));
            if (stats)
                for (int64_t b = 0; b < nb; ++b) {
                    stats[f0 + b] = ccab_solve_stats{};
                    stats[f0 + b].converged = 1;
                    stats[f0 + b].backward_error = -1.0;
                    stats[f0 + b].rel_residual = -1.0;
                }
        } else if (solver == CCAB_SOLVER_ITERATIVE_MIXED) {
            CCAB_TRY(gmres_ir_run(ctx, n, nb, static_cast<const c128*>(A), A64,
                                  half_op ? CCAB_C32 : CCAB_C64, half_op ? inv_scale + f0 : nullptr,
                                  n, int64_t(n2), reinterpret_cast<c128*>(x), x0, cfg->gmres.restart,
                                  cfg->ir_tolerance, cfg->gmres.tolerance, cfg->ir_max_steps,
                                  cfg->gmres.ref_mgs != 0, stats ? stats + f0 : nullptr, nullptr,
                                  Alo, warm_ax ? AX : nullptr));
        } else if (solver == CCAB_SOLVER_DIRECT_MIXED) {
            CCAB_TRY(mixed_run(ctx, n, nb, static_cast<const c128*>(A), A64, n, int64_t(n2),
                               reinterpret_cast<c128*>(x), cfg->ir_tolerance, cfg->ir_max_steps,
                               stats ? stats + f0 : nullptr));
        } else {
            const void* op = cfg->gmres.matrix_prec == CCAB_C64 ? static_cast<const void*>(A64) : A;
            const size_t esz = cfg->gmres.matrix_prec == CCAB_C64 ? 8 : 16;
            for (int64_t b = 0; b < nb; ++b) {
                ccab_solve_stats st{};
                CCAB_TRY(gmres_run(ctx, n, cfg->gmres.matrix_prec,
                                   static_cast<const char*>(op) + b * n2 * esz, n,
                                   reinterpret_cast<c128*>(x + b * n), nullptr, cfg->gmres, &st));
                if (stats) stats[f0 + b] = st;
            }
        }
        CCAB_TRY(finish_surface(ctx, hp.data() + f0, dp + f0, bc, n, nb, x, d_phi + f0 * n,
                                d_vn + f0 * n));
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

extern "C" int ccab_sweep(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                          const double* d_cpts, int64_t n, const ccab_solver_cfg* cfg,
                          ccab_c128* d_phi, ccab_c128* d_vn, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(w != nullptr, "null wave");
    CCAB_CHECK(w->kind == CCAB_WAVE_PLANE || w->kind == CCAB_WAVE_SPHERICAL, "bad wave kind");
    return sweep_impl(ctx, ph, nfreq, w, nullptr, d_cpts, n, cfg, d_phi, d_vn, stats);
}

extern "C" int ccab_sweep_warm(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq,
                               const ccab_wave* w, const double* d_cpts, int64_t n,
                               const ccab_solver_cfg* cfg, const ccab_c128* d_x0,
                               ccab_c128* d_phi, ccab_c128* d_vn, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(w != nullptr && d_x0 != nullptr, "null wave or initial guess");
    CCAB_CHECK(cfg && (cfg->solver == CCAB_SOLVER_ITERATIVE_MIXED ||
                       cfg->solver == CCAB_SOLVER_HIERARCHICAL),
               "a warm start needs the ITERATIVE_MIXED or HIERARCHICAL solver");
    return sweep_impl(ctx, ph, nfreq, w, nullptr, d_cpts, n, cfg, d_phi, d_vn, stats, d_x0);
}

// Bodies of uneven size (one system per body, each with its own mesh,
// frequency and incident wave) solved together.  ITERATIVE_MIXED: per body
// the incident RHS, the fused assembly (FP64 arithmetic; c128 matrix and the
// c64 or scaled-FP16 operator copy) and the BC algebra, packed back to back
// in HBM; then one GMRES-IR over all bodies whose matvecs and Arnoldi steps
// are single batched launches (ragged GEMV; vectors padded to the largest
// body with zeros).  Bodies are taken in chunks that fit the memory budget.
// Other solvers run body by body.
static int bodies_impl(ccab_ctx* ctx, int64_t nbody, const int64_t* nh, const double* d_cpts,
                       const ccab_physics* ph, const ccab_wave* w, const ccab_solver_cfg* cfg,
                       const ccab_c128* d_x0, ccab_c128* d_phi, ccab_c128* d_vn,
                       ccab_solve_stats* stats) {
    CCAB_CHECK(nbody > 0 && nh && d_cpts && ph && w && cfg && d_phi && d_vn, "invalid arguments");
    std::vector<int64_t> coff(nbody + 1, 0);
    for (int64_t b = 0; b < nbody; ++b) {
        CCAB_CHECK(nh[b] > 0 && nh[b] < (int64_t(1) << 31), "invalid body size");
        CCAB_CHECK(w[b].kind == CCAB_WAVE_PLANE || w[b].kind == CCAB_WAVE_SPHERICAL, "bad wave kind");
        CCAB_CHECK(ph[b].bc_type == ph[0].bc_type, "bodies share one BC type");
        coff[b + 1] = coff[b] + nh[b];
    }
    if (cfg->solver != CCAB_SOLVER_ITERATIVE_MIXED) {
        CCAB_CHECK(!d_x0 || cfg->solver == CCAB_SOLVER_HIERARCHICAL,
                   "a warm start needs the ITERATIVE_MIXED or HIERARCHICAL solver");
        for (int64_t b = 0; b < nbody; ++b)
            CCAB_TRY(sweep_impl(ctx, ph + b, 1, w + b, nullptr, d_cpts + 8 * coff[b], nh[b], cfg,
                                d_phi + coff[b], d_vn + coff[b], stats ? stats + b : nullptr,
                                d_x0 ? d_x0 + coff[b] : nullptr));
        return CCAB_OK;
    }
    CCAB_TRY(check_residual_matrix(ctx, *cfg));
    const int bc = ph[0].bc_type;
    const bool half_op = cfg->gmres.matrix_prec == CCAB_C32;
    const size_t esz_op = half_op ? 4 : 8;
    // FP64 residuals against the FP16 operator + its FP32 remainder
    const bool split_res = half_op && cfg->residual_matrix == CCAB_C32;
    const size_t esz_res = split_res ? 8 : 16;
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, nbody, hp, &dp));
    // chunks of bodies within the memory budget
    const size_t budget = mem_budget(ctx);
    for (int64_t b0 = 0; b0 < nbody;) {
        int64_t b1 = b0;
        size_t bytes = 0;
        while (b1 < nbody) {
            const size_t add = size_t(nh[b1]) * nh[b1] * (esz_res + esz_op);
            if (b1 > b0 && bytes + add > budget) break;
            bytes += add;
            ++b1;
        }
        const int64_t nc = b1 - b0;
        int64_t nmax = 0, nmin = INT64_MAX;
        std::vector<int64_t> hn(nc), haoff(nc);
        int64_t entries = 0;
        for (int64_t c = 0; c < nc; ++c) {
            hn[c] = nh[b0 + c];
            haoff[c] = entries;
            entries += hn[c] * hn[c];
            nmax = std::max(nmax, hn[c]);
            nmin = std::min(nmin, hn[c]);
        }
        int align = 4;
        for (int64_t c = 0; c < nc; ++c)
            while (align > 1 && (hn[c] % align != 0 || haoff[c] % align != 0)) align /= 2;
        const int64_t npts = coff[b1] - coff[b0];
        void* Ares = ctx_buf(ctx, SLOT_SOLVE_A, size_t(entries) * esz_res);
        c128* A128 = split_res ? nullptr : static_cast<c128*>(Ares);
        c64* Alo = split_res ? static_cast<c64*>(Ares) : nullptr;
        char* Aop = static_cast<char*>(ctx_buf(ctx, SLOT_SOLVE_A64, size_t(entries) * esz_op));
        c128* X = static_cast<c128*>(ctx_buf(ctx, SLOT_SOLVE_RHS, size_t(nc) * nmax * 16));
        double* soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(npts) * 64));
        int64_t* dr = static_cast<int64_t*>(ctx_buf(ctx, SLOT_RAGGED, size_t(nc) * 16));
        if (!Ares || !Aop || !X || !soa || !dr) return CCAB_ENOMEM;
        c128* X0 = nullptr;
        if (d_x0) {
            X0 = static_cast<c128*>(ctx_buf(ctx, SLOT_SOLVE_X, size_t(nc) * nmax * 16));
            if (!X0) return CCAB_ENOMEM;
        }
        bool need_rowsum = false;
        for (int64_t c = 0; c < nc; ++c) need_rowsum |= hp[b0 + c].rowsum_of != 0;
        double* part = nullptr;  // row-sum partials of every body (below)
        double* inv_scale = nullptr;
        {
            std::vector<int64_t> up(2 * nc);
            std::copy(hn.begin(), hn.end(), up.begin());
            std::copy(haoff.begin(), haoff.end(), up.begin() + nc);
            CCAB_HIP(hipMemcpy(dr, up.data(), size_t(nc) * 16, hipMemcpyHostToDevice));
        }
        // padded right-hand sides (zero beyond each body's size) and guesses
        hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((nc * nmax + 255) / 256)), dim3(256), 0,
                           ctx->stream, nc * nmax, c128{0.0, 0.0}, reinterpret_cast<ccab_c128*>(X), 0);
        if (X0)
            hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((nc * nmax + 255) / 256)), dim3(256), 0,
                               ctx->stream, nc * nmax, c128{0.0, 0.0},
                               reinterpret_cast<ccab_c128*>(X0), 0);
        // the chunk's bodies in one launch per stage: descriptors (points,
        // matrix, assembly tiles, row-sum partials) and incident waves
        std::vector<BodyDesc> desc(nc + 1);
        std::vector<BodyWave> bw(nc);
        int64_t wg = 0, poff = 0;
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t b = b0 + c, n = hn[c];
            desc[c] = BodyDesc{n, coff[b] - coff[b0], haoff[c], wg, poff, 0};
            wg += ((n + 255) / 256) * ((n + 63) / 64);
            poff += 2 * ((n + 63) / 64) * n;
            double d[3] = {w[b].vec[0], w[b].vec[1], w[b].vec[2]};
            const double nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (!(nn > 0.0)) return ctx_fail(ctx, CCAB_ENAN, "incident wave vector has zero length");
            // normalize(); the spherical origin stays put with CCAB_FIX_SPHERICAL_ORIGIN
            if (w[b].kind == CCAB_WAVE_PLANE || !(ph[b].corrections & CCAB_FIX_SPHERICAL_ORIGIN))
                for (double& v : d) v /= nn;
            bw[c] = BodyWave{w[b].kind, 0, {d[0], d[1], d[2]},
                             cdiv(c128{w[b].amplitude[0], w[b].amplitude[1]},
                                  c128{0.0, hp[b].omega * ph[b].mass_density})};
        }
        desc[nc] = BodyDesc{0, npts, entries, wg, poff, 0};
        const size_t desc_bytes = ws_align(size_t(nc + 1) * sizeof(BodyDesc));
        char* dd = static_cast<char*>(ctx_buf(ctx, SLOT_BODY_DESC, desc_bytes + size_t(nc) * sizeof(BodyWave)));
        if (!dd) return CCAB_ENOMEM;
        BodyDesc* d_desc = reinterpret_cast<BodyDesc*>(dd);
        BodyWave* d_waves = reinterpret_cast<BodyWave*>(dd + desc_bytes);
        CCAB_HIP(hipMemcpy(d_desc, desc.data(), size_t(nc + 1) * sizeof(BodyDesc), hipMemcpyHostToDevice));
        CCAB_HIP(hipMemcpy(d_waves, bw.data(), size_t(nc) * sizeof(BodyWave), hipMemcpyHostToDevice));
        if (half_op) {
            // FP16 scale per body, as in the sweep: the largest diagonal entry
            // maps to about 2^12 (the maxima on the device, one number per body back)
            inv_scale = static_cast<double*>(ctx_buf(ctx, SLOT_SOLVE_ISCALE, size_t(nc) * 8));
            if (!inv_scale) return CCAB_ENOMEM;
            CCAB_TRY(launch_diag_max(ctx, nc, dp + b0, d_cpts + 8 * coff[b0], d_desc, 0, inv_scale));
            std::vector<double> isc(nc);
            CCAB_HIP(hipMemcpyAsync(isc.data(), inv_scale, size_t(nc) * 8, hipMemcpyDeviceToHost,
                                    ctx->stream));
            CCAB_HIP(hipStreamSynchronize(ctx->stream));
            for (int64_t c = 0; c < nc; ++c) {
                SysParams& p = hp[b0 + c];
                const double mx = isc[c];
                CCAB_CHECK(mx > 0.0 && std::isfinite(mx), "degenerate diagonal for the FP16 operator");
                set_h_scale(p, std::ldexp(1.0, (int)std::floor(std::log2(4096.0 / mx))));
                isc[c] = 1.0 / p.h_scale;
            }
            CCAB_HIP(hipMemcpy(dp + b0, hp.data() + b0, sizeof(SysParams) * nc, hipMemcpyHostToDevice));
            CCAB_HIP(hipMemcpy(inv_scale, isc.data(), size_t(nc) * 8, hipMemcpyHostToDevice));
        }
        if (need_rowsum) {
            part = static_cast<double*>(ctx_buf(ctx, SLOT_ROWSUM, size_t(std::max<int64_t>(poff, 1)) * 8));
            if (!part) return CCAB_ENOMEM;
        }
        // a warm start's A x0 from the assembly (partials laid out as the
        // row sums', then one padded vector per body)
        double* ax_part = nullptr;
        c128* AX = nullptr;
        if (X0 && ctx->warm_ax) {
            ax_part = static_cast<double*>(ctx_buf(ctx, SLOT_AX_PART, size_t(std::max<int64_t>(poff, 1)) * 8));
            AX = static_cast<c128*>(ctx_buf(ctx, SLOT_AX, size_t(nc) * nmax * 16));
            if (!ax_part || !AX) return CCAB_ENOMEM;
        }
        CCAB_TRY(launch_bodies_setup(ctx, nc, nmax, d_desc, d_cpts + 8 * coff[b0], soa, dp + b0,
                                     d_waves, reinterpret_cast<ccab_c128*>(X),
                                     d_x0 ? d_x0 + coff[b0] : nullptr,
                                     reinterpret_cast<ccab_c128*>(X0)));
        CCAB_TRY(launch_assemble_bodies(ctx, nc, nmax, d_desc, desc.data(), dp + b0, hp.data() + b0,
                                        soa, half_op ? PREC_F64_H : PREC_F64, A128, Aop, Alo, part,
                                        reinterpret_cast<ccab_c128*>(X), ax_part ? X0 : nullptr,
                                        ax_part, AX));
        Ragged rg;
        rg.n = dr;
        rg.aoff = dr + nc;
        rg.align = align;
        rg.n_min = nmin;
        rg.hn = hn.data();
        CCAB_TRY(gmres_ir_run(ctx, nmax, nc, A128, Aop, half_op ? CCAB_C32 : CCAB_C64,
                              half_op ? inv_scale : nullptr, nmax, 0, X, X0, cfg->gmres.restart,
                              cfg->ir_tolerance, cfg->gmres.tolerance, cfg->ir_max_steps,
                              cfg->gmres.ref_mgs != 0, stats ? stats + b0 : nullptr, &rg, Alo,
                              AX));
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t b = b0 + c;
            CCAB_TRY(finish_surface(ctx, hp.data() + b, dp + b, bc, hn[c], 1,
                                    reinterpret_cast<const ccab_c128*>(X + c * nmax),
                                    d_phi + coff[b], d_vn + coff[b]));
        }
        b0 = b1;
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

extern "C" int ccab_solve_bodies(ccab_ctx* ctx, int64_t nbody, const int64_t* n,
                                 const double* d_cpts, const ccab_physics* ph, const ccab_wave* w,
                                 const ccab_solver_cfg* cfg, const ccab_c128* d_x0,
                                 ccab_c128* d_phi, ccab_c128* d_vn, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    return bodies_impl(ctx, nbody, n, d_cpts, ph, w, cfg, d_x0, d_phi, d_vn, stats);
}

extern "C" int ccab_gmres_ir_batched(ccab_ctx* ctx, int64_t n, int64_t batch,
                                     const ccab_c128* d_A128, const ccab_c64* d_A64, int64_t lda,
                                     int64_t strideA, ccab_c128* d_x, const ccab_c128* d_x0,
                                     const ccab_solver_cfg* cfg, ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A64 && d_x && cfg && n > 0 && batch > 0 && lda >= n, "invalid arguments");
    CCAB_CHECK(batch == 1 || strideA >= lda * n, "stride too small");
    return gmres_ir_run(ctx, n, batch, reinterpret_cast<const c128*>(d_A128), d_A64, CCAB_C64,
                        nullptr, lda, strideA,
                        reinterpret_cast<c128*>(d_x), reinterpret_cast<const c128*>(d_x0),
                        cfg->gmres.restart, cfg->ir_tolerance, cfg->gmres.tolerance,
                        cfg->ir_max_steps, cfg->gmres.ref_mgs != 0, stats);
}

extern "C" int ccab_solve_surface(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                  int64_t n, const ccab_c128* d_rhs_inc,
                                  const ccab_solver_cfg* cfg, ccab_c128* d_phi, ccab_c128* d_vn,
                                  ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_rhs_inc != nullptr, "null rhs_inc");
    return sweep_impl(ctx, ph, 1, nullptr, d_rhs_inc, d_cpts, n, cfg, d_phi, d_vn, stats);
}

extern "C" int ccab_cgemm_update(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t K,
                                 const void* d_A, int64_t lda, const void* d_B, int64_t ldb,
                                 void* d_C, int64_t ldc) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "bad precision");
    CCAB_CHECK(lda >= M && ldb >= K && ldc >= M, "bad leading dimensions");
    return cgemm_update(ctx, prec, M, N, K, d_A, lda, d_B, ldb, d_C, ldc);
}

// ---------------------------------------------------------------------------
// hierarchical matrix
// ---------------------------------------------------------------------------
extern "C" int ccab_hmatrix_build(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                  int64_t n, int side, int64_t leaf_size, double eta,
                                  double aca_tolerance, ccab_hmatrix** out) {
    return ccab_hmatrix_build_ex(ctx, ph, d_cpts, n, side, leaf_size, eta, aca_tolerance,
                                 CCAB_C128, out);
}

extern "C" int ccab_hmatrix_build_ex(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                     int64_t n, int side, int64_t leaf_size, double eta,
                                     double aca_tolerance, int prec, ccab_hmatrix** out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && out && n > 0, "invalid arguments");
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "H-matrix storage is CCAB_C128 or CCAB_C64");
    *out = nullptr;
    CCAB_CHECK(side == CCAB_LHS || side == CCAB_RHS, "bad side %d", side);
    CCAB_CHECK(leaf_size >= 1, "leaf_size must be positive");
    CCAB_CHECK(eta > 0.0 && aca_tolerance > 0.0, "eta and aca_tolerance must be positive");
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, 1, hp, &dp));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    std::vector<double> h;
    CCAB_TRY(host_cpts(ctx, d_cpts, n, h));
    return hmatrix_build(ctx, hp[0], soa, h.data(), n, side, leaf_size, eta, aca_tolerance, out,
                         prec);
}

extern "C" int ccab_hmatrix_build_matrix(ccab_ctx* ctx, int64_t n, const ccab_c128* d_A,
                                         int64_t lda, const double* h_xyz, int64_t leaf_size,
                                         double eta, double aca_tolerance, int prec,
                                         ccab_hmatrix** out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A && h_xyz && out && n > 0 && lda >= n, "invalid arguments");
    *out = nullptr;
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "H-matrix storage is CCAB_C128 or CCAB_C64");
    CCAB_CHECK(leaf_size >= 1, "leaf_size must be positive");
    CCAB_CHECK(eta > 0.0 && aca_tolerance > 0.0, "eta and aca_tolerance must be positive");
    return hmatrix_build_matrix(ctx, reinterpret_cast<const c128*>(d_A), lda, h_xyz, n, leaf_size,
                                eta, aca_tolerance, out, prec);
}

extern "C" int ccab_aca_build(ccab_ctx* ctx, int64_t m, int64_t n, const ccab_c128* d_A,
                              int64_t lda, double tol, ccab_hmatrix** out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(d_A && out && m > 0 && n > 0 && lda >= m, "invalid arguments");
    *out = nullptr;
    CCAB_CHECK(tol > 0.0, "tol must be positive");
    return aca_build_matrix(ctx, reinterpret_cast<const c128*>(d_A), lda, m, n, tol, out,
                            CCAB_C128);
}

extern "C" int ccab_hmatrix_destroy(ccab_hmatrix* h) {
    if (h) hmatrix_free(h);
    return CCAB_OK;
}

extern "C" int ccab_hmatrix_gemv(ccab_ctx* ctx, const ccab_hmatrix* h, const double* alpha,
                                 const ccab_c128* d_x, const double* beta, ccab_c128* d_y) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(h && alpha && beta && d_x && d_y, "invalid arguments");
    CCAB_CHECK(hmatrix_device(h) == ctx->device, "H matrix lives on another device");
    return hmatrix_gemv(ctx, h, c128{alpha[0], alpha[1]}, reinterpret_cast<const c128*>(d_x),
                        c128{beta[0], beta[1]}, reinterpret_cast<c128*>(d_y));
}

extern "C" int ccab_hmatrix_to_full(ccab_ctx* ctx, const ccab_hmatrix* h, ccab_c128* d_A,
                                    int64_t lda) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(h && d_A && lda >= hmatrix_n(h), "invalid arguments");
    CCAB_CHECK(hmatrix_device(h) == ctx->device, "H matrix lives on another device");
    return hmatrix_to_full(ctx, h, reinterpret_cast<c128*>(d_A), lda);
}

extern "C" int ccab_hmatrix_stats(const ccab_hmatrix* h, int64_t* st, double* norm) {
    if (!h || !st) return CCAB_EINVAL;
    hmatrix_stats(h, st, norm);
    return CCAB_OK;
}

extern "C" int ccab_hmatrix_stats_ex(const ccab_hmatrix* h, int64_t* st, int nst, double* norm) {
    if (!h || !st || nst < 0) return CCAB_EINVAL;
    hmatrix_stats_ex(h, st, nst);
    if (norm) *norm = hmatrix_norm(h);
    return CCAB_OK;
}

extern "C" int ccab_gmres_hmatrix(ccab_ctx* ctx, const ccab_hmatrix* h, ccab_c128* d_x,
                                  const ccab_c128* d_x0, const ccab_gmres_params* p,
                                  ccab_solve_stats* stats) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(h && d_x && p, "invalid arguments");
    CCAB_CHECK(hmatrix_device(h) == ctx->device, "H matrix lives on another device");
    CCAB_CHECK(hmatrix_ncol(h) == hmatrix_n(h), "GMRES needs a square operator");
    MatvecFn op = [&](const c128* xin, c128 a, c128 b, c128* yout) -> int {
        return hmatrix_gemv(ctx, h, a, xin, b, yout);
    };
    return gmres_op_run(ctx, hmatrix_n(h), op, hmatrix_norm(h), reinterpret_cast<c128*>(d_x),
                        reinterpret_cast<const c128*>(d_x0), *p, stats);
}
