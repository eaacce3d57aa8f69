// Autotuner of the blocked LU (no reference counterpart): times the batched
// factorisation of synthetic systems for every complex-GEMM variant (bit 0
// XCD-aware tile order, bit 1 LDS-staged epilogue, bit 2 128-row tiles) and
// both outer panel widths, and keeps the fastest in the context.  The GEMM
// variants change only the tile order, epilogue staging and tile height, not
// the order of the k-sums of any entry, so the factors do not depend on them.
#include <algorithm>
#include <cfloat>
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
        for (int v = 0; v < 8 && rc == CCAB_OK; ++v) {
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
            if (ms_table) ms_table[oi * 8 + v] = tmin;
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
