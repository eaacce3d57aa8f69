// Autotuner of the blocked LU (no reference counterpart): times the batched
// factorisation of synthetic systems for every complex-GEMM variant (bit 0
// XCD-aware tile order, bit 1 LDS-staged epilogue, bit 2 128-row tiles) and
// both outer panel widths, and keeps the fastest in the context.  The GEMM
// variants change only the tile order, epilogue staging and tile height, not
// the order of the k-sums of any entry, so the factors do not depend on them.
#include <algorithm>
#include <cfloat>
#include <functional>
#include <vector>

#include "ccab_internal.h"

namespace {

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A_b(i, j) = u - 0.5 + i v (u, v uniform in [0, 1) from a hash of (b, i, j)),
// plus n/4 on the diagonal
template <typename T>
__global__ void tune_fill_kernel(Cx<T>* A, int64_t n, int64_t batch) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * n * batch) return;
    const uint64_t h = mix64(uint64_t(e) * 2654435761ull + 12345u);
    const double u = double(h >> 11) * (1.0 / 9007199254740992.0);
    const double v = double(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
    const int64_t ij = e % (n * n);
    const int64_t i = ij % n, j = ij / n;
    A[e] = Cx<T>{T(u - 0.5 + (i == j ? 0.25 * double(n) : 0.0)), T(v)};
}

}  // namespace

extern "C" int ccab_autotune(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, int* best_variant,
                             int* best_outer, double* ms_table) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(prec == CCAB_C128 || prec == CCAB_C64, "bad precision");
    CCAB_CHECK(n >= 64 && batch >= 1 && n < (int64_t(1) << 31), "invalid size");
    const size_t esz = prec == CCAB_C128 ? 16 : 8;
    const size_t bytes = size_t(n) * n * batch * esz;
    void *pristine = nullptr, *work = nullptr, *piv = nullptr;
    auto cleanup = [&]() {
        if (pristine) (void)hipFree(pristine);
        if (work) (void)hipFree(work);
        if (piv) (void)hipFree(piv);
    };
    if (hipMalloc(&pristine, bytes) != hipSuccess || hipMalloc(&work, bytes) != hipSuccess ||
        hipMalloc(&piv, size_t(n) * batch * 4) != hipSuccess) {
        cleanup();
        return ctx_fail(ctx, CCAB_ENOMEM, "autotune: cannot allocate %zu bytes", 2 * bytes);
    }
    hipStream_t s = ctx->stream;
    const int64_t total = n * n * batch;
    const dim3 g((unsigned)((total + 255) / 256));
    if (prec == CCAB_C128)
        hipLaunchKernelGGL(tune_fill_kernel<double>, g, dim3(256), 0, s, static_cast<c128*>(pristine), n, batch);
    else
        hipLaunchKernelGGL(tune_fill_kernel<float>, g, dim3(256), 0, s, static_cast<c64*>(pristine), n, batch);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return ctx_fail(ctx, CCAB_EHIP, "autotune: event creation failed");
    }
    const int pi = prec == CCAB_C128 ? 0 : 1;
    const int save_v = ctx->gemm_variant[pi], save_o = ctx->lu_outer;
    double best = DBL_MAX;
    int bv = save_v, bo = save_o;
    int rc = CCAB_OK;
    for (int oi = 0; oi < 2 && rc == CCAB_OK; ++oi) {
        const int outer = oi == 0 ? 64 : 128;
        for (int v = 0; v <= kGemmLibVariant && rc == CCAB_OK; ++v) {
            ctx->gemm_variant[pi] = v;
            ctx->lu_outer = outer;
            double tmin = DBL_MAX;
            for (int rep = 0; rep < 3 && rc == CCAB_OK; ++rep) {  // rep 0 warms up
                if (hipMemcpyAsync(work, pristine, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                    hipEventRecord(e0, s) != hipSuccess) {
                    rc = ctx_fail(ctx, CCAB_EHIP, "autotune: copy failed");
                    break;
                }
                rc = lu_factor(ctx, prec, n, batch, work, n, n * n, static_cast<int32_t*>(piv));
                if (rc != CCAB_OK) break;
                float ms = 0.f;
                if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = ctx_fail(ctx, CCAB_EHIP, "autotune: timing failed");
                    break;
                }
                if (rep > 0) tmin = std::min(tmin, double(ms));
            }
            if (ms_table) ms_table[oi * (kGemmLibVariant + 1) + v] = tmin;
            if (rc == CCAB_OK && tmin < best) {
                best = tmin;
                bv = v;
                bo = outer;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    if (rc != CCAB_OK) {
        ctx->gemm_variant[pi] = save_v;
        ctx->lu_outer = save_o;
        return rc;
    }
    ctx->gemm_variant[pi] = bv;
    ctx->lu_outer = bo;
    if (best_variant) *best_variant = bv;
    if (best_outer) *best_outer = bo;
    return CCAB_OK;
}

// ---------------------------------------------------------------------------
// Sweep autotuner: the batched assembly (systems per thread, column chunk)
// and the uniform GEMVs of the operator and of the refinement matrix
// (workgroup target, non-temporal loads), timed on the caller's mesh and
// frequencies and cached in the context per (precision, n).  None of the
// choices changes what is computed, only how the work is split: the chunking
// changes the order of partial sums, so results agree to rounding.
// ---------------------------------------------------------------------------
const AsmTune* asm_tune_of(const ccab_ctx* ctx, int prec, int64_t n) {
    for (const auto& t : ctx->asm_tune)
        if (t.prec == prec && t.n == n) return &t;
    return nullptr;
}

const GemvTune* gemv_tune_of(const ccab_ctx* ctx, int prec, int64_t n) {
    for (const auto& t : ctx->gemv_tune)
        if (t.prec == prec && t.n == n) return &t;
    return nullptr;
}

namespace {

constexpr int kTuneKf[3] = {8, 16, 32};
constexpr int kTuneCc[3] = {64, 128, 256};
constexpr int kTuneWg[5] = {256, 512, 1024, 2048, 4096};

void put_asm(ccab_ctx* ctx, int prec, int64_t n, int kf, int cc) {
    for (auto& t : ctx->asm_tune)
        if (t.prec == prec && t.n == n) {
            t.kf = kf;
            t.col_chunk = cc;
            return;
        }
    ctx->asm_tune.push_back(AsmTune{prec, n, kf, cc});
}

void put_gemv(ccab_ctx* ctx, int prec, int64_t n, int wg, int nt) {
    for (auto& t : ctx->gemv_tune)
        if (t.prec == prec && t.n == n) {
            t.wg_target = wg;
            t.nt = nt;
            return;
        }
    ctx->gemv_tune.push_back(GemvTune{prec, n, wg, nt});
}

void drop_tuning(ccab_ctx* ctx, int op_prec, int rem_prec, int64_t n) {
    auto& a = ctx->asm_tune;
    a.erase(std::remove_if(a.begin(), a.end(),
                           [&](const AsmTune& t) { return t.prec == op_prec && t.n == n; }),
            a.end());
    auto& g = ctx->gemv_tune;
    g.erase(std::remove_if(g.begin(), g.end(),
                           [&](const GemvTune& t) {
                               return (t.prec == op_prec || t.prec == rem_prec) && t.n == n;
                           }),
            g.end());
}

__global__ void tune_ones_kernel(int64_t count, double* v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) v[i] = 1.0;
}

// the refinement's matrix that goes with an operator precision
int rem_prec_of(int op_prec) { return op_prec == CCAB_C32 ? CCAB_C64 : CCAB_C128; }

}  // namespace

extern "C" int ccab_set_sweep_tuning(ccab_ctx* ctx, int op_prec, int64_t n, const int* cfg) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(op_prec == CCAB_C32 || op_prec == CCAB_C64, "operator precision CCAB_C32 or CCAB_C64");
    CCAB_CHECK(n > 0, "invalid size");
    const int rp = rem_prec_of(op_prec);
    drop_tuning(ctx, op_prec, rp, n);
    if (!cfg) return CCAB_OK;
    auto in = [](int v, const int* set, int cnt) { return std::find(set, set + cnt, v) != set + cnt; };
    CCAB_CHECK(in(cfg[0], kTuneKf, 3), "systems per thread must be 8, 16 or 32");
    CCAB_CHECK(in(cfg[1], kTuneCc, 3), "column chunk must be 64, 128 or 256");
    CCAB_CHECK(cfg[2] >= 64 && cfg[2] <= 65536 && cfg[4] >= 64 && cfg[4] <= 65536,
               "workgroup targets must lie in [64, 65536]");
    CCAB_CHECK((cfg[3] == 0 || cfg[3] == 1) && (cfg[5] == 0 || cfg[5] == 1), "nt flags are 0 or 1");
    put_asm(ctx, op_prec, n, cfg[0], cfg[1]);
    put_gemv(ctx, op_prec, n, cfg[2], cfg[3]);
    put_gemv(ctx, rp, n, cfg[4], cfg[5]);
    return CCAB_OK;
}

extern "C" int ccab_autotune_sweep(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq,
                                   const double* d_cpts, int64_t n, int op_prec, int* best,
                                   double* ms_table) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && n > 0 && nfreq >= 1 && nfreq <= 65535, "invalid arguments");
    CCAB_CHECK(op_prec == CCAB_C32 || op_prec == CCAB_C64, "operator precision CCAB_C32 or CCAB_C64");
    const int rp = rem_prec_of(op_prec);
    std::vector<SysParams> hp;
    SysParams* dp;
    CCAB_TRY(stage_params(ctx, ph, nfreq, hp, &dp));
    double* soa;
    CCAB_TRY(stage_soa(ctx, d_cpts, n, &soa));
    const size_t n2 = size_t(n) * size_t(n);
    const size_t op_b = op_prec == CCAB_C32 ? 4 : 8, rem_b = op_prec == CCAB_C32 ? 8 : 16;
    // as many systems as half the free memory holds (the launches of a
    // sweep chunk see the same batch)
    size_t freemem = 0, total = 0;
    int64_t batch = nfreq;
    if (hipMemGetInfo(&freemem, &total) == hipSuccess)
        batch = std::max<int64_t>(1, std::min<int64_t>(nfreq, int64_t(freemem / 2 / (n2 * (op_b + rem_b)))));
    void *op = nullptr, *rem = nullptr, *xy = nullptr;
    auto cleanup = [&]() {
        if (op) (void)hipFree(op);
        if (rem) (void)hipFree(rem);
        if (xy) (void)hipFree(xy);
    };
    const size_t vec = size_t(n) * size_t(batch) * 16;
    if (hipMalloc(&op, n2 * batch * op_b) != hipSuccess ||
        hipMalloc(&rem, n2 * batch * rem_b) != hipSuccess ||
        hipMalloc(&xy, 2 * vec + size_t(batch) * 8) != hipSuccess) {
        cleanup();
        return ctx_fail(ctx, CCAB_ENOMEM, "autotune: cannot allocate %zu bytes",
                        n2 * batch * (op_b + rem_b));
    }
    hipStream_t s = ctx->stream;
    c128* x = static_cast<c128*>(xy);
    c128* y = x + size_t(n) * batch;
    double* iscale = reinterpret_cast<double*>(y + size_t(n) * batch);
    const int64_t ones = 2 * int64_t(n) * batch;
    hipLaunchKernelGGL(tune_ones_kernel, dim3((unsigned)((ones + 255) / 256)), dim3(256), 0, s, ones,
                       reinterpret_cast<double*>(x));
    hipLaunchKernelGGL(tune_ones_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s,
                       int64_t(batch), iscale);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return ctx_fail(ctx, CCAB_EHIP, "autotune: event creation failed");
    }
    // best of two timed runs after one warm-up
    auto timed = [&](const std::function<int()>& run, double* out) -> int {
        double tmin = DBL_MAX;
        for (int rep = 0; rep < 3; ++rep) {
            CCAB_HIP(hipEventRecord(e0, s));
            CCAB_TRY(run());
            CCAB_HIP(hipEventRecord(e1, s));
            CCAB_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CCAB_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tmin = std::min(tmin, double(ms));
        }
        *out = tmin;
        return CCAB_OK;
    };
    const AsmTune* had_a = asm_tune_of(ctx, op_prec, n);
    const AsmTune save_a = had_a ? *had_a : AsmTune{op_prec, n, 16, 64};
    const bool keep_a = had_a != nullptr;
    int rc = CCAB_OK;
    int bkf = 16, bcc = 64, bwo = ctx->gemv_wg_target, bno = ctx->gemv_nt ? 1 : 0;
    int bwr = bwo, bnr = bno;
    double best_ms = DBL_MAX;
    int slot = 0;
    // assembly: ASM_LHS into the operator copy and the refinement matrix, as a sweep does
    for (int a = 0; a < 3 && rc == CCAB_OK; ++a)
        for (int c = 0; c < 3 && rc == CCAB_OK; ++c, ++slot) {
            put_asm(ctx, op_prec, n, kTuneKf[a], kTuneCc[c]);
            double ms = 0.0;
            rc = timed([&]() {
                return op_prec == CCAB_C32
                           ? launch_assemble(ctx, dp, hp.data(), batch, soa, n, PREC_F64_H, ASM_LHS,
                                             nullptr, op, n, int64_t(n2), nullptr, nullptr, rem)
                           : launch_assemble(ctx, dp, hp.data(), batch, soa, n, PREC_F64, ASM_LHS,
                                             rem, op, n, int64_t(n2), nullptr, nullptr);
            }, &ms);
            if (ms_table) ms_table[slot] = ms;
            if (rc == CCAB_OK && ms < best_ms) {
                best_ms = ms;
                bkf = kTuneKf[a];
                bcc = kTuneCc[c];
            }
        }
    // GEMVs of the operator and of the refinement matrix over the whole batch
    const c128 one{1.0, 0.0}, zero{0.0, 0.0};
    for (int which = 0; which < 2 && rc == CCAB_OK; ++which) {
        const int prec = which == 0 ? op_prec : rp;
        const void* A = which == 0 ? op : rem;
        const GemvTune* had = gemv_tune_of(ctx, prec, n);
        best_ms = DBL_MAX;
        int bw = ctx->gemv_wg_target, bn = ctx->gemv_nt ? 1 : 0;
        for (int w = 0; w < 5 && rc == CCAB_OK; ++w)
            for (int t = 0; t < 2 && rc == CCAB_OK; ++t, ++slot) {
                put_gemv(ctx, prec, n, kTuneWg[w], t);
                double ms = 0.0;
                rc = timed([&]() {
                    return gemv_active(ctx, prec, n, n, batch, nullptr, A, n, int64_t(n2), x, n, one,
                                       zero, y, n, prec == CCAB_C32 ? iscale : nullptr, nullptr,
                                       nullptr, nullptr, -1);
                }, &ms);
                if (ms_table) ms_table[slot] = ms;
                if (rc == CCAB_OK && ms < best_ms) {
                    best_ms = ms;
                    bw = kTuneWg[w];
                    bn = t;
                }
            }
        (void)had;
        if (which == 0) {
            bwo = bw;
            bno = bn;
        } else {
            bwr = bw;
            bnr = bn;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    if (rc != CCAB_OK) {
        drop_tuning(ctx, op_prec, rp, n);
        if (keep_a) put_asm(ctx, op_prec, n, save_a.kf, save_a.col_chunk);
        return rc;
    }
    put_asm(ctx, op_prec, n, bkf, bcc);
    put_gemv(ctx, op_prec, n, bwo, bno);
    put_gemv(ctx, rp, n, bwr, bnr);
    if (best) {
        best[0] = bkf;
        best[1] = bcc;
        best[2] = bwo;
        best[3] = bno;
        best[4] = bwr;
        best[5] = bnr;
    }
    return CCAB_OK;
}
