// Field-point evaluation for a batch of frequencies:
// get_dense_field_matrices (influence_matrix.rs:101-127) fused with get_field
// (solve.rs:173-185), the incident field (incident_wave.rs:66-100) and the
// potential-to-pressure conversion (postprocess.rs:70-88).  M and L are never
// stored: each wavefront evaluates 64 field points against a chunk of the
// surface (surface geometry and phi, vn are wave-uniform loads), writes
// per-chunk partial sums, and a second kernel adds the chunks in a fixed
// order.  Work per (field point, surface point, frequency): one kernel pair
// (g, h) -- sqrt, reciprocal, sincos and ~30 flops, no matrix traffic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ccab_internal.h"
#include "greens.h"

namespace {

struct FieldSys {
    double k;
    c128 scale;   // i omega rho (pressure) or 1
    c128 amp;     // incident amplitude A / (i omega rho)
    double d[3];  // plane: unit direction; spherical: the source point
    double msign; // +1: -M phi + L vn (solve.rs:180-181); -1 with CCAB_FIX_FIELD_SIGN
};

constexpr int kFieldWave = 64;

// part[(f * nchunk + c) * nfp + p] = sum_{j in chunk c} (-M_pj phi_fj + L_pj vn_fj)
__global__ __launch_bounds__(kFieldWave) void field_partial_kernel(
    const FieldSys* __restrict__ sys, const double* __restrict__ soa_base, int64_t n,
    const double* __restrict__ fps, int64_t nfp, const ccab_c128* __restrict__ phi,
    const ccab_c128* __restrict__ vn, int64_t chunk, int nchunk, c128* part) {
    const int64_t p = int64_t(blockIdx.x) * kFieldWave + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t f = blockIdx.z;
    const Soa s = soa_of(soa_base, n);
    const double k = sys[f].k;
    const bool live = p < nfp;
    const double px = live ? fps[3 * p] : 0.0;
    const double py = live ? fps[3 * p + 1] : 0.0;
    const double pz = live ? fps[3 * p + 2] : 0.0;
    const int64_t j0 = int64_t(c) * chunk, j1 = std::min<int64_t>(n, j0 + chunk);
    const ccab_c128* ph = phi + f * n;
    const ccab_c128* v = vn + f * n;
    c128 acc{0.0, 0.0};
    for (int64_t j = j0; j < j1; ++j) {
        c128 g, h;
        // x = field point, y = surface point j with its normal (the field
        // point's normal is the reference's zero dummy, unused without BM)
        greens<double, false>(k, c128{0.0, 0.0}, px - s.x[j], py - s.y[j], pz - s.z[j], 0.0, 0.0,
                              0.0, s.nx[j], s.ny[j], s.nz[j], g, h);
        const double wa = s.wt[j] * s.area[j];
        const c128 pj{ph[j].re, ph[j].im}, vj{v[j].re, v[j].im};
        acc += (g * wa) * vj - (h * wa) * pj;
    }
    if (live) part[(f * nchunk + c) * nfp + p] = acc;
}

// out[f * nfp + p] = scale_f (inc_fp + sum_c part); inc_out = scale_f inc_fp
__global__ void field_finish_kernel(const FieldSys* __restrict__ sys, int kind,
                                    const double* __restrict__ fps, int64_t nfp, int nchunk,
                                    const c128* __restrict__ part, int want_total,
                                    int with_incident, ccab_c128* out, ccab_c128* inc_out) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t f = blockIdx.y;
    if (p >= nfp) return;
    const FieldSys q = sys[f];
    c128 inc{0.0, 0.0};
    if (with_incident) {
        const double x = fps[3 * p], y = fps[3 * p + 1], z = fps[3 * p + 2];
        double sn, cs;
        if (kind == CCAB_WAVE_PLANE) {
            sincos(q.k * (q.d[0] * x + q.d[1] * y + q.d[2] * z), &sn, &cs);
            inc = q.amp * c128{cs, sn};
        } else {
            const double rx = x - q.d[0], ry = y - q.d[1], rz = z - q.d[2];
            const double r = sqrt(rx * rx + ry * ry + rz * rz);
            sincos(q.k * r, &sn, &cs);
            inc = (q.amp * c128{cs, sn}) * (1.0 / (4.0 * kPi * r));
        }
    }
    c128 sc{0.0, 0.0};
    for (int c = 0; c < nchunk; ++c) sc += part[(f * nchunk + c) * nfp + p];
    const c128 acc = (want_total ? inc : c128{0.0, 0.0}) + sc * q.msign;
    const c128 o = q.scale * acc;
    out[f * nfp + p] = ccab_c128{o.re, o.im};
    if (inc_out) {
        const c128 io = q.scale * inc;
        inc_out[f * nfp + p] = ccab_c128{io.re, io.im};
    }
}

// M[p + j ld] = h_pj wt_j area_j, L[p + j ld] = g_pj wt_j area_j (influence_matrix.rs:108-125)
__global__ void field_matrices_kernel(double k, const double* __restrict__ soa_base, int64_t n,
                                      const double* __restrict__ fps, int64_t nfp, int64_t ld,
                                      ccab_c128* M, ccab_c128* L) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (p >= nfp) return;
    const Soa s = soa_of(soa_base, n);
    c128 g, h;
    greens<double, false>(k, c128{0.0, 0.0}, fps[3 * p] - s.x[j], fps[3 * p + 1] - s.y[j],
                          fps[3 * p + 2] - s.z[j], 0.0, 0.0, 0.0, s.nx[j], s.ny[j], s.nz[j], g, h);
    const double wa = s.wt[j] * s.area[j];
    if (M) M[p + j * ld] = ccab_c128{h.re * wa, h.im * wa};
    if (L) L[p + j * ld] = ccab_c128{g.re * wa, g.im * wa};
}

}  // namespace

extern "C" int ccab_dense_field_matrices(ccab_ctx* ctx, const ccab_physics* ph, const double* d_cpts,
                                         int64_t n, const double* d_fps, int64_t nfp,
                                         ccab_c128* d_M, ccab_c128* d_L, int64_t ld) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && d_fps && n > 0 && nfp > 0 && ld >= nfp, "invalid arguments");
    CCAB_CHECK(n <= 2147483647, "too many surface points");
    CCAB_TRY(check_physics(ctx, *ph));
    const SysParams p = make_params(*ph);
    double* soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(n) * 64));
    if (!soa) return CCAB_ENOMEM;
    CCAB_TRY(launch_cpts_to_soa(ctx, d_cpts, n, soa));
    hipLaunchKernelGGL(field_matrices_kernel, dim3((unsigned)((nfp + 255) / 256), (unsigned)n),
                       dim3(256), 0, ctx->stream, p.k, soa, n, d_fps, nfp, ld, d_M, d_L);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

namespace {

// Surface chunks of the field evaluation (grid.y): enough wavefronts to fill
// the chip, at most max_chunks; *chunk = surface points per chunk.
int field_chunks(const ccab_ctx* ctx, int64_t n, int64_t nfp, int64_t nfreq, int64_t max_chunks,
                 int64_t* chunk) {
    const int64_t pblocks = (nfp + kFieldWave - 1) / kFieldWave;
    const int64_t want_waves = 8 * std::max(ctx->num_cus, 1) * 4;
    int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(
        (n + 127) / 128, (want_waves + pblocks * nfreq - 1) / (pblocks * nfreq)));
    nchunk = std::max<int64_t>(1, std::min(nchunk, max_chunks));
    *chunk = (n + nchunk - 1) / nchunk;
    return (int)((n + *chunk - 1) / *chunk);
}

// per-frequency constants on the host (hs[nfreq])
int field_sys_host(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                   bool to_pressure, bool need_inc, FieldSys* hs) {
    if (w) CCAB_CHECK(w->kind == CCAB_WAVE_PLANE || w->kind == CCAB_WAVE_SPHERICAL, "bad wave kind");
    CCAB_CHECK(!need_inc || w, "the incident field needs a wave");
    double d[3] = {0.0, 0.0, 0.0};
    if (w) {
        const double nn =
            std::sqrt(w->vec[0] * w->vec[0] + w->vec[1] * w->vec[1] + w->vec[2] * w->vec[2]);
        if (need_inc && !(nn > 0.0))
            return ctx_fail(ctx, CCAB_ENAN, "incident wave vector has zero length");
        // normalize(); a spherical source moves to the unit sphere
        // (incident_wave.rs:91) unless CCAB_FIX_SPHERICAL_ORIGIN
        const bool keep = w->kind == CCAB_WAVE_SPHERICAL &&
                          (ph[0].corrections & CCAB_FIX_SPHERICAL_ORIGIN);
        for (int t = 0; t < 3; ++t) d[t] = keep ? w->vec[t] : (nn > 0.0 ? w->vec[t] / nn : 0.0);
    }
    for (int64_t f = 0; f < nfreq; ++f) {
        CCAB_TRY(check_physics(ctx, ph[f]));
        const SysParams p = make_params(ph[f]);
        FieldSys& q = hs[f];
        q.k = p.k;
        const double wr = p.omega * ph[f].mass_density;
        q.scale = to_pressure ? c128{0.0, wr} : c128{1.0, 0.0};
        q.amp = w ? cdiv(c128{w->amplitude[0], w->amplitude[1]}, c128{0.0, wr}) : c128{0.0, 0.0};
        for (int t = 0; t < 3; ++t) q.d[t] = d[t];
        q.msign = (ph[f].corrections & CCAB_FIX_FIELD_SIGN) ? -1.0 : 1.0;
    }
    return CCAB_OK;
}

// per-frequency constants on the device
int stage_field_sys(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                    bool to_pressure, bool need_inc, FieldSys** dsys) {
    std::vector<FieldSys> hs(nfreq);
    CCAB_TRY(field_sys_host(ctx, ph, nfreq, w, to_pressure, need_inc, hs.data()));
    *dsys = static_cast<FieldSys*>(ctx_buf(ctx, SLOT_FIELD_SYS, sizeof(FieldSys) * nfreq));
    if (!*dsys) return CCAB_ENOMEM;
    CCAB_TRY(pinned_reserve(ctx, sizeof(FieldSys) * nfreq));
    std::memcpy(ctx->pinned, hs.data(), sizeof(FieldSys) * nfreq);
    CCAB_HIP(hipMemcpyAsync(*dsys, ctx->pinned, sizeof(FieldSys) * nfreq, hipMemcpyHostToDevice,
                            ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));  // the pinned staging is reused by later calls
    return CCAB_OK;
}

}  // namespace

extern "C" int ccab_incident_field(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w,
                                   const double* d_fps, int64_t nfp, ccab_c128* d_out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && w && d_fps && d_out && nfp > 0, "invalid arguments");
    FieldSys* dsys;
    CCAB_TRY(stage_field_sys(ctx, ph, 1, w, false, true, &dsys));
    hipLaunchKernelGGL(field_finish_kernel, dim3((unsigned)((nfp + 255) / 256), 1u), dim3(256), 0,
                       ctx->stream, dsys, w->kind, d_fps, nfp, 0, nullptr, 1, 1, d_out, nullptr);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

extern "C" int ccab_field(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq, const ccab_wave* w,
                          const double* d_cpts, int64_t n, const double* d_fps, int64_t nfp,
                          const ccab_c128* d_phi, const ccab_c128* d_vn, int want_total,
                          int to_pressure, ccab_c128* d_out, ccab_c128* d_inc) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && d_fps && d_phi && d_vn && d_out && n > 0 && nfp > 0 && nfreq > 0,
               "invalid arguments");
    const bool need_inc = want_total || d_inc;
    FieldSys* dsys;
    CCAB_TRY(stage_field_sys(ctx, ph, nfreq, w, to_pressure != 0, need_inc, &dsys));
    double* soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(n) * 64));
    if (!soa) return CCAB_ENOMEM;
    CCAB_TRY(launch_cpts_to_soa(ctx, d_cpts, n, soa));
    const int64_t pblocks = (nfp + kFieldWave - 1) / kFieldWave;
    int64_t chunk;
    const int nchunk = field_chunks(ctx, n, nfp, nfreq, INT32_MAX, &chunk);
    c128* part = static_cast<c128*>(
        ctx_buf(ctx, SLOT_FIELD_PART, size_t(nfreq) * nchunk * nfp * sizeof(c128)));
    if (!part) return CCAB_ENOMEM;
    {
        ProfScope ps(ctx, PROF_FIELD, double(nfreq) * double(nfp) * double(n));
        hipLaunchKernelGGL(field_partial_kernel, dim3((unsigned)pblocks, (unsigned)nchunk, (unsigned)nfreq),
                           dim3(kFieldWave), 0, ctx->stream, dsys, soa, n, d_fps, nfp, d_phi, d_vn,
                           chunk, nchunk, part);
        hipLaunchKernelGGL(field_finish_kernel, dim3((unsigned)((nfp + 255) / 256), (unsigned)nfreq),
                           dim3(256), 0, ctx->stream, dsys, w ? w->kind : 0, d_fps, nfp, nchunk,
                           part, want_total ? 1 : 0, need_inc ? 1 : 0, d_out, d_inc);
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

// ---------------------------------------------------------------------------
// ccab_field_noalloc: ccab_field with every scratch buffer (the per-frequency
// constants, the SoA copy of the surface points and the per-chunk partial
// sums) in one array of the code object, so that a call allocates nothing.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kFieldScratchBytes = size_t(32) << 20;   // per device
constexpr int64_t kFieldNoAllocMaxFreq = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kFieldScratchBytes];

// the array's device address on every device, from hipGetSymbolAddress
// (the host-side name `scratch` is not a device address)
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Once per device, with that device current (ccab_create calls it).
void field_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
}

extern "C" int ccab_field_noalloc(ccab_ctx* ctx, const ccab_physics* ph, int64_t nfreq,
                                  const ccab_wave* w, const double* d_cpts, int64_t n,
                                  const double* d_fps, int64_t nfp, const ccab_c128* d_phi,
                                  const ccab_c128* d_vn, int want_total, int to_pressure,
                                  ccab_c128* d_out, ccab_c128* d_inc) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && d_cpts && d_fps && d_phi && d_vn && d_out && n > 0 && nfp > 0 && nfreq > 0,
               "invalid arguments");
    CCAB_CHECK(nfreq <= kFieldNoAllocMaxFreq, "ccab_field_noalloc: at most %lld frequencies",
               (long long)kFieldNoAllocMaxFreq);
    const int dev = ctx->device;
    CCAB_CHECK(dev >= 0 && dev < kMaxDevices, "ccab_field_noalloc: device %d out of range", dev);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_done[dev] || !g_scratch_ptr[dev])
            return ctx_fail(ctx, CCAB_EHIP, "ccab_field_noalloc: no scratch address on device %d (%s)",
                            dev, g_scratch_done[dev] ? hipGetErrorString(g_scratch_err[dev])
                                                     : "not initialised");
        base = g_scratch_ptr[dev];
    }
    // layout: [FieldSys x nfreq][SoA, 8 n doubles][partials, nfreq x nchunk x nfp c128]
    const size_t sys_bytes = align256(sizeof(FieldSys) * size_t(nfreq));
    const size_t soa_bytes = align256(size_t(n) * 64);
    const size_t part_unit = size_t(nfreq) * size_t(nfp) * sizeof(c128);  // one chunk
    if (sys_bytes + soa_bytes + part_unit > kFieldScratchBytes)
        return ctx_fail(ctx, CCAB_EINVAL,
                        "ccab_field_noalloc: n = %lld surface points, %lld field points and %lld "
                        "frequencies need more than its %zu MiB scratch (8 n doubles + 16 nfp "
                        "nfreq bytes per surface chunk); use ccab_field",
                        (long long)n, (long long)nfp, (long long)nfreq, kFieldScratchBytes >> 20);
    const int64_t max_chunks = int64_t((kFieldScratchBytes - sys_bytes - soa_bytes) / part_unit);
    const bool need_inc = want_total || d_inc;
    FieldSys hs[kFieldNoAllocMaxFreq];
    CCAB_TRY(field_sys_host(ctx, ph, nfreq, w, to_pressure != 0, need_inc, hs));
    unsigned char* b = static_cast<unsigned char*>(base);
    FieldSys* dsys = reinterpret_cast<FieldSys*>(b);
    double* soa = reinterpret_cast<double*>(b + sys_bytes);
    c128* part = reinterpret_cast<c128*>(b + sys_bytes + soa_bytes);
    CCAB_HIP(hipMemcpyAsync(dsys, hs, sizeof(FieldSys) * size_t(nfreq), hipMemcpyHostToDevice,
                            ctx->stream));
    CCAB_TRY(launch_cpts_to_soa(ctx, d_cpts, n, soa));
    const int64_t pblocks = (nfp + kFieldWave - 1) / kFieldWave;
    int64_t chunk;
    const int nchunk = field_chunks(ctx, n, nfp, nfreq, max_chunks, &chunk);
    // every partial slot of the nchunk chunks is written before it is read
    hipLaunchKernelGGL(field_partial_kernel, dim3((unsigned)pblocks, (unsigned)nchunk, (unsigned)nfreq),
                       dim3(kFieldWave), 0, ctx->stream, dsys, soa, n, d_fps, nfp, d_phi, d_vn,
                       chunk, nchunk, part);
    hipLaunchKernelGGL(field_finish_kernel, dim3((unsigned)((nfp + 255) / 256), (unsigned)nfreq),
                       dim3(256), 0, ctx->stream, dsys, w ? w->kind : 0, d_fps, nfp, nchunk,
                       part, want_total ? 1 : 0, need_inc ? 1 : 0, d_out, d_inc);
    CCAB_HIP(hipGetLastError());
    // the constants were copied from this stack frame
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}
