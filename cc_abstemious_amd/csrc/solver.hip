// Solvers over the dense operator:
//  - restarted GMRES following gmres.rs:35-202 (zero or warm start, restart
//    ceil(sqrt(max_it)), backward-error stop of gmres.rs:142-144, Givens on
//    the host, modified Gram-Schmidt on the device in FP64 with the
//    reference's coefficient qk1.dotc(q_i) or the textbook q_i^H w);
//  - complex-FP32 LU on MFMA with FP64 iterative refinement (no reference
//    counterpart: it stands in for solve_lu in the frequency sweep).
#include <cmath>
#include <complex>
#include <cstring>

#include "ccab_internal.h"
#include "coop.h"

int gemv_general(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t batch, const void* d_A,
                 int64_t lda, int64_t strideA, const c128* d_x, int64_t stridex, c128 alpha,
                 c128 beta, c128* d_y, int64_t stridey);

namespace {

constexpr int kMgsThreads = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Two-value block reduction over kMgsThreads lanes.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[2 * wave] = a;
        red[2 * wave + 1] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < kMgsThreads / 64; ++w) {
        a += red[2 * w];
        b += red[2 * w + 1];
    }
}

// Orthogonalise w = Q[:, k+1] against Q[:, 0..k] (modified Gram-Schmidt),
// h[0..k] = coefficients, h[k+1] = ||w||, then w /= h[k+1].
// REF: coefficient sum_j conj(w_j) q_ij (nalgebra qk1.dotc(q_i), gmres.rs:168).
template <bool REF>
__global__ __launch_bounds__(kMgsThreads) void mgs_kernel(int64_t n, int64_t k, c128* Q, int64_t ldq,
                                                          int64_t sQ, c128* h, int64_t sh,
                                                          const int32_t* __restrict__ active,
                                                          const int32_t* __restrict__ skip) {
    __shared__ double red[2 * kMgsThreads / 64];
    const int64_t b = active ? active[blockIdx.x] : blockIdx.x;
    if (skip && skip[b]) return;
    c128* q = Q + b * sQ;
    c128* w = q + (k + 1) * ldq;
    c128* hb = h + b * sh;
    for (int64_t i = 0; i <= k; ++i) {
        const c128* qi = q + i * ldq;
        double sr = 0.0, si = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) {
            const c128 a = qi[r], x = w[r];
            if (REF) {  // conj(w) * q
                sr += x.re * a.re + x.im * a.im;
                si += x.re * a.im - x.im * a.re;
            } else {    // conj(q) * w
                sr += a.re * x.re + a.im * x.im;
                si += a.re * x.im - a.im * x.re;
            }
        }
        block_sum2(sr, si, red);
        const c128 hi{sr, si};
        for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) w[r] = w[r] - hi * qi[r];
        if (threadIdx.x == 0) hb[i] = hi;
        __syncthreads();
    }
    double nn = 0.0, dummy = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) nn += w[r].re * w[r].re + w[r].im * w[r].im;
    block_sum2(nn, dummy, red);
    const double hn = sqrt(nn);
    if (threadIdx.x == 0) hb[k + 1] = c128{hn, 0.0};
    // scale = 1 / (hn + 0i) as a complex reciprocal
    const c128 sc = crecip(c128{hn, 0.0});
    for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) w[r] = w[r] * sc;
}

// The textbook step of mgs_kernel with the basis vectors taken B at a time
// (classical Gram-Schmidt within a block, modified across blocks, as
// mgs_block_coop_kernel): per block one pass over w and the block's vectors
// for the B coefficients, one block reduction, one pass for the update.
template <int B>
__global__ __launch_bounds__(kMgsThreads) void mgs_block_kernel(int64_t n, int64_t k, c128* Q,
                                                                int64_t ldq, int64_t sQ, c128* h,
                                                                int64_t sh,
                                                                const int32_t* __restrict__ active,
                                                                const int32_t* __restrict__ skip) {
    constexpr int NV = 2 * B;
    constexpr int NW = kMgsThreads / 64;
    __shared__ double red[NV][NW];
    const int64_t b = active ? active[blockIdx.x] : blockIdx.x;
    if (skip && skip[b]) return;
    c128* q = Q + b * sQ;
    c128* w = q + (k + 1) * ldq;
    c128* hb = h + b * sh;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t i0 = 0; i0 <= k; i0 += B) {
        const int nbk = (int)min<int64_t>(B, k + 1 - i0);
        double v[NV];
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) {
            const c128 x = w[r];
            // (vectors past k repeat vector k; their sums are not used)
#pragma unroll
            for (int bb = 0; bb < B; ++bb) {
                const c128 a = q[min<int64_t>(i0 + bb, k) * ldq + r];
                v[2 * bb] += a.re * x.re + a.im * x.im;
                v[2 * bb + 1] += a.re * x.im - a.im * x.re;
            }
        }
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] = wave_sum(v[t]);
        if (lane == 0)
#pragma unroll
            for (int t = 0; t < NV; ++t) red[t][wave] = v[t];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NV; ++t) {
            v[t] = 0.0;
            for (int ww = 0; ww < NW; ++ww) v[t] += red[t][ww];
        }
        __syncthreads();  // red is reused by the next block
        for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) {
            c128 x = w[r];
            // (static trip counts keep v in registers)
#pragma unroll
            for (int bb = 0; bb < B; ++bb)
                if (bb < nbk) x = x - c128{v[2 * bb], v[2 * bb + 1]} * q[(i0 + bb) * ldq + r];
            w[r] = x;
        }
        if (threadIdx.x == 0)
#pragma unroll
            for (int bb = 0; bb < B; ++bb)
                if (bb < nbk) hb[i0 + bb] = c128{v[2 * bb], v[2 * bb + 1]};
        __syncthreads();
    }
    double nn = 0.0, dummy = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) nn += w[r].re * w[r].re + w[r].im * w[r].im;
    block_sum2(nn, dummy, reinterpret_cast<double*>(red));
    const double hn = sqrt(nn);
    if (threadIdx.x == 0) hb[k + 1] = c128{hn, 0.0};
    const c128 sc = crecip(c128{hn, 0.0});
    for (int64_t r = threadIdx.x; r < n; r += kMgsThreads) w[r] = w[r] * sc;
}

// The same step for few large systems: G workgroups per system (grid.x), each
// holding rows g*256*RQ .. of w in registers.  Per basis vector every
// workgroup publishes its partial dot product, one grid barrier, and all
// workgroups add the G partials in the same fixed order (so they agree on
// the coefficient bit for bit), then update their rows.  k + 2 barriers per
// step; Q is read once.  After a failed barrier (see coop.h) h[k+1] is NaN
// and *sticky is set: the caller's Givens step finds the NaN and
// coop_abort_check reports the timeout.
constexpr int kMgsCoopThreads = 256;

template <bool REF, int RQ>
__global__ __launch_bounds__(kMgsCoopThreads) void mgs_coop_kernel(
    int64_t n, int64_t k, c128* Q, int64_t ldq, int64_t sQ, c128* h, int64_t sh,
    const int32_t* __restrict__ active, uint64_t* __restrict__ xbuf, int64_t sys_words,
    unsigned spin_limit, const int32_t* __restrict__ skip, uint64_t* __restrict__ sticky) {
    constexpr int NW = kMgsCoopThreads / 64;
    __shared__ double red[2 * NW];
    __shared__ double tot[2];
    __shared__ int s_ok;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t z = blockIdx.y;
    const int64_t b = active ? active[z] : z;
    // every workgroup of a skipped system leaves before its first barrier
    if (skip && skip[b]) return;
    uint64_t* hdr = xbuf + z * sys_words;
    c128* q = Q + b * sQ;
    c128* w = q + (k + 1) * ldq;
    c128* hb = h + b * sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = int64_t(g) * kMgsCoopThreads * RQ;
    c128 wv[RQ];
#pragma unroll
    for (int u = 0; u < RQ; ++u) {
        const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
        wv[u] = row < n ? w[row] : c128{0.0, 0.0};
    }
    // the grid-wide sum of (a, b) over the workgroups' partials, barrier
    // number nb; false after a failed barrier
    auto grid_sum = [&](double& a, double& bb, int64_t nb) -> bool {
        a = wave_sum(a);
        bb = wave_sum(bb);
        if (lane == 0) {
            red[2 * wave] = a;
            red[2 * wave + 1] = bb;
        }
        __syncthreads();
        uint64_t* par = hdr + kCoopHeaderWords + (nb & 1) * 2 * int64_t(G);
        if (tid == 0) {
            double pa = 0.0, pb = 0.0;
            for (int t = 0; t < NW; ++t) {
                pa += red[2 * t];
                pb += red[2 * t + 1];
            }
            __hip_atomic_store(par + 2 * g, uint64_t(__double_as_longlong(pa)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(par + 2 * g + 1, uint64_t(__double_as_longlong(pb)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!coop_barrier(hdr, uint64_t(nb + 1) * uint64_t(G), spin_limit, &s_ok)) return false;
        if (wave == 0) {
            double sa = 0.0, sb = 0.0;
            for (int t = lane; t < G; t += 64) {
                sa += __longlong_as_double((long long)__hip_atomic_load(par + 2 * t, __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT));
                sb += __longlong_as_double((long long)__hip_atomic_load(
                    par + 2 * t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            }
            sa = wave_sum(sa);
            sb = wave_sum(sb);
            if (lane == 0) {
                tot[0] = sa;
                tot[1] = sb;
            }
        }
        __syncthreads();
        a = tot[0];
        bb = tot[1];
        __syncthreads();  // tot and red are reused by the next sum
        return true;
    };
    // the last workgroup of the system to leave clears its header (barrier
    // counter, abort flag, leave counter at word 16) for the next launch
    auto leave = [&]() {
        __syncthreads();
        if (tid == 0) {
            const uint64_t before = __hip_atomic_fetch_add(hdr + 16, uint64_t(1), __ATOMIC_ACQ_REL,
                                                           __HIP_MEMORY_SCOPE_AGENT);
            if (before == uint64_t(G - 1)) {
                __hip_atomic_store(hdr, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(hdr + 8, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(hdr + 16, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    };
    bool ok = true;
    int64_t nb = 0;
    for (int64_t i = 0; i <= k && ok; ++i) {
        const c128* qi = q + i * ldq;
        c128 qv[RQ];
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int u = 0; u < RQ; ++u) {
            const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
            qv[u] = row < n ? qi[row] : c128{0.0, 0.0};
            const c128 a = qv[u], x = wv[u];
            if (REF) {  // conj(w) * q
                sr += x.re * a.re + x.im * a.im;
                si += x.re * a.im - x.im * a.re;
            } else {    // conj(q) * w
                sr += a.re * x.re + a.im * x.im;
                si += a.re * x.im - a.im * x.re;
            }
        }
        ok = grid_sum(sr, si, nb++);
        if (!ok) break;
        const c128 hi{sr, si};
#pragma unroll
        for (int u = 0; u < RQ; ++u) wv[u] = wv[u] - hi * qv[u];
        if (g == 0 && tid == 0) hb[i] = hi;
    }
    double nn = 0.0, dummy = 0.0;
    if (ok) {
#pragma unroll
        for (int u = 0; u < RQ; ++u) nn += wv[u].re * wv[u].re + wv[u].im * wv[u].im;
        ok = grid_sum(nn, dummy, nb++);
    }
    if (!ok) {
        if (g == 0 && tid == 0) {
            hb[k + 1] = c128{__builtin_nan(""), 0.0};
            __hip_atomic_store(sticky, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        leave();
        return;
    }
    const double hn = sqrt(nn);
    if (g == 0 && tid == 0) hb[k + 1] = c128{hn, 0.0};
    const c128 sc = crecip(c128{hn, 0.0});
#pragma unroll
    for (int u = 0; u < RQ; ++u) {
        const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
        if (row < n) w[row] = wv[u] * sc;
    }
    leave();
}

// The textbook step (coefficient conj(q_i) w) of mgs_coop_kernel with the
// basis vectors taken B at a time: the B coefficients of a block are formed
// against the same w (classical Gram-Schmidt within a block, modified across
// blocks), so a step waits at ceil((k + 1) / B) + 1 grid barriers instead of
// k + 2 and loads B vectors at a time (each read twice: for the coefficients
// and for the update).  The Arnoldi relation A q_k = Q h holds to rounding
// either way; within a block the basis vectors are orthonormal to working
// accuracy, so orthogonality is kept to the same order.  Used by the
// batched GMRES-IR (its inner solves stop at 1e-3 relative).
template <int RQ, int B>
__global__ __launch_bounds__(kMgsCoopThreads) void mgs_block_coop_kernel(
    int64_t n, int64_t k, c128* Q, int64_t ldq, int64_t sQ, c128* h, int64_t sh,
    const int32_t* __restrict__ active, uint64_t* __restrict__ xbuf, int64_t sys_words,
    unsigned spin_limit, const int32_t* __restrict__ skip, uint64_t* __restrict__ sticky) {
    constexpr int NW = kMgsCoopThreads / 64;
    constexpr int NV = 2 * B;
    __shared__ double red[NV][NW];
    __shared__ double tot[NV];
    __shared__ int s_ok;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t z = blockIdx.y;
    const int64_t b = active ? active[z] : z;
    // every workgroup of a skipped system leaves before its first barrier
    if (skip && skip[b]) return;
    uint64_t* hdr = xbuf + z * sys_words;
    c128* q = Q + b * sQ;
    c128* w = q + (k + 1) * ldq;
    c128* hb = h + b * sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = int64_t(g) * kMgsCoopThreads * RQ;
    c128 wv[RQ];
#pragma unroll
    for (int u = 0; u < RQ; ++u) {
        const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
        wv[u] = row < n ? w[row] : c128{0.0, 0.0};
    }
    // the grid-wide sums of v[0..NV) over the workgroups' partials, barrier
    // number nb; every workgroup adds the G partials in the same order
    auto grid_sum = [&](double (&v)[NV], int64_t nb) -> bool {
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] = wave_sum(v[t]);
        if (lane == 0)
#pragma unroll
            for (int t = 0; t < NV; ++t) red[t][wave] = v[t];
        __syncthreads();
        uint64_t* par = hdr + kCoopHeaderWords + (nb & 1) * NV * int64_t(G);
        if (tid < NV) {
            double pa = 0.0;
            for (int t = 0; t < NW; ++t) pa += red[tid][t];
            __hip_atomic_store(par + NV * int64_t(g) + tid, uint64_t(__double_as_longlong(pa)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!coop_barrier(hdr, uint64_t(nb + 1) * uint64_t(G), spin_limit, &s_ok)) return false;
        if (tid < NV) {
            double sa = 0.0;
            for (int t = 0; t < G; ++t)
                sa += __longlong_as_double((long long)__hip_atomic_load(
                    par + NV * int64_t(t) + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            tot[tid] = sa;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] = tot[t];
        __syncthreads();  // tot and red are reused by the next sum
        return true;
    };
    auto leave = [&]() {
        __syncthreads();
        if (tid == 0) {
            const uint64_t before = __hip_atomic_fetch_add(hdr + 16, uint64_t(1), __ATOMIC_ACQ_REL,
                                                           __HIP_MEMORY_SCOPE_AGENT);
            if (before == uint64_t(G - 1)) {
                __hip_atomic_store(hdr, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(hdr + 8, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(hdr + 16, uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    };
    bool ok = true;
    int64_t nb = 0;
    for (int64_t i0 = 0; i0 <= k && ok; i0 += B) {
        const int nbk = (int)min<int64_t>(B, k + 1 - i0);
        double v[NV];
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] = 0.0;
        // (vectors past k repeat vector k; their sums are not used)
#pragma unroll
        for (int bb = 0; bb < B; ++bb) {
            const c128* qi = q + min<int64_t>(i0 + bb, k) * ldq;
#pragma unroll
            for (int u = 0; u < RQ; ++u) {
                const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
                const c128 a = row < n ? qi[row] : c128{0.0, 0.0};
                const c128 x = wv[u];
                v[2 * bb] += a.re * x.re + a.im * x.im;
                v[2 * bb + 1] += a.re * x.im - a.im * x.re;
            }
        }
        ok = grid_sum(v, nb++);
        if (!ok) break;
        // (a static trip count keeps v in registers)
#pragma unroll
        for (int bb = 0; bb < B; ++bb) {
            if (bb >= nbk) break;
            const c128 hi{v[2 * bb], v[2 * bb + 1]};
            const c128* qi = q + (i0 + bb) * ldq;
#pragma unroll
            for (int u = 0; u < RQ; ++u) {
                const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
                if (row < n) wv[u] = wv[u] - hi * qi[row];
            }
            if (g == 0 && tid == 0) hb[i0 + bb] = hi;
        }
    }
    double nv[NV];
#pragma unroll
    for (int t = 0; t < NV; ++t) nv[t] = 0.0;
    if (ok) {
#pragma unroll
        for (int u = 0; u < RQ; ++u) nv[0] += wv[u].re * wv[u].re + wv[u].im * wv[u].im;
        ok = grid_sum(nv, nb++);
    }
    if (!ok) {
        if (g == 0 && tid == 0) {
            hb[k + 1] = c128{__builtin_nan(""), 0.0};
            __hip_atomic_store(sticky, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        leave();
        return;
    }
    const double hn = sqrt(nv[0]);
    if (g == 0 && tid == 0) hb[k + 1] = c128{hn, 0.0};
    const c128 sc = crecip(c128{hn, 0.0});
#pragma unroll
    for (int u = 0; u < RQ; ++u) {
        const int64_t row = r0 + tid + int64_t(u) * kMgsCoopThreads;
        if (row < n) w[row] = wv[u] * sc;
    }
    leave();
}

// Modified Gram-Schmidt step of nact systems (active list, or all when
// null): one workgroup per system, or, for few large systems that would
// leave most CUs idle, mgs_coop_kernel over up to one workgroup per CU.
int launch_mgs(ccab_ctx* ctx, bool ref, int64_t n, int64_t k, c128* Q, int64_t ldq, int64_t sQ,
               c128* h, int64_t sh, const int32_t* act, int64_t nact, const int32_t* skip = nullptr,
               bool reserve_only = false, int block = 1) {
    hipStream_t s = ctx->stream;
    if (ctx->mgs_coop && n >= 4096) {
        int rq = 0;
        int64_t G = 0;
        for (int cand : {1, 2, 4, 8}) {
            G = (n + int64_t(kMgsCoopThreads) * cand - 1) / (int64_t(kMgsCoopThreads) * cand);
            if (G * nact <= ctx->num_cus) {
                rq = cand;
                break;
            }
        }
        if (rq) {
            // (room for the exchange of mgs_block_coop_kernel<., 8>: two
            // buffers of 16 values per workgroup; the same layout for every
            // launch of the context, whose headers sit at multiples of it)
            const int64_t sys_words = kCoopHeaderWords + 32 * int64_t(std::max(ctx->num_cus, 1));
            uint64_t* xbuf = static_cast<uint64_t*>(
                ctx_buf(ctx, SLOT_MGS_COOP, size_t(nact) * sys_words * 8));
            if (!xbuf) return CCAB_ENOMEM;
            // headers clear themselves at the end of every launch; a new or
            // grown buffer is cleared once
            if (xbuf != ctx->mgs_xbuf || nact > ctx->mgs_xbuf_ready) {
                hipLaunchKernelGGL(coop_reset_kernel, dim3((unsigned)((nact + 63) / 64)), dim3(64), 0,
                                   s, xbuf, sys_words, nact);
                ctx->mgs_xbuf = xbuf;
                ctx->mgs_xbuf_ready = nact;
            }
            if (!ctx->coop_abort) {
                uint64_t* st = static_cast<uint64_t*>(ctx_buf(ctx, SLOT_COOP_ABORT, 8));
                if (!st) return CCAB_ENOMEM;
                hipLaunchKernelGGL(coop_reset_kernel, dim3(1), dim3(64), 0, s, st, int64_t(1),
                                   int64_t(1));
                ctx->coop_abort = st;
            }
            if (reserve_only) return CCAB_OK;
            const dim3 grid((unsigned)G, (unsigned)nact);
#define CCAB_MGS_COOP(R)                                                                        \
    do {                                                                                        \
        if (!ref && block == 4)                                                                 \
            hipLaunchKernelGGL((mgs_block_coop_kernel<R, 4>), grid, dim3(kMgsCoopThreads), 0, s, \
                               n, k, Q, ldq, sQ, h, sh, act, xbuf, sys_words, kCoopSpinLimit,    \
                               skip, ctx->coop_abort);                                 \
        else if (!ref && block == 8)                                                            \
            hipLaunchKernelGGL((mgs_block_coop_kernel<R, 8>), grid, dim3(kMgsCoopThreads), 0, s, \
                               n, k, Q, ldq, sQ, h, sh, act, xbuf, sys_words, kCoopSpinLimit,    \
                               skip, ctx->coop_abort);                                 \
        else if (ref)                                                                           \
            hipLaunchKernelGGL((mgs_coop_kernel<true, R>), grid, dim3(kMgsCoopThreads), 0, s, n, \
                               k, Q, ldq, sQ, h, sh, act, xbuf, sys_words, kCoopSpinLimit, skip, \
                               ctx->coop_abort);                                       \
        else                                                                                    \
            hipLaunchKernelGGL((mgs_coop_kernel<false, R>), grid, dim3(kMgsCoopThreads), 0, s,   \
                               n, k, Q, ldq, sQ, h, sh, act, xbuf, sys_words, kCoopSpinLimit, skip, \
                               ctx->coop_abort);                                       \
    } while (0)
            switch (rq) {
            case 1: CCAB_MGS_COOP(1); break;
            case 2: CCAB_MGS_COOP(2); break;
            case 4: CCAB_MGS_COOP(4); break;
            default: CCAB_MGS_COOP(8); break;
            }
#undef CCAB_MGS_COOP
            CCAB_HIP(hipGetLastError());
            return CCAB_OK;
        }
    }
    if (reserve_only) return CCAB_OK;
    if (!ref && block == 8)
        hipLaunchKernelGGL(mgs_block_kernel<8>, dim3((unsigned)nact), dim3(kMgsThreads), 0, s, n, k,
                           Q, ldq, sQ, h, sh, act, skip);
    else if (!ref && block == 4)
        hipLaunchKernelGGL(mgs_block_kernel<4>, dim3((unsigned)nact), dim3(kMgsThreads), 0, s, n, k,
                           Q, ldq, sQ, h, sh, act, skip);
    else if (ref)
        hipLaunchKernelGGL(mgs_kernel<true>, dim3((unsigned)nact), dim3(kMgsThreads), 0, s, n, k, Q,
                           ldq, sQ, h, sh, act, skip);
    else
        hipLaunchKernelGGL(mgs_kernel<false>, dim3((unsigned)nact), dim3(kMgsThreads), 0, s, n, k,
                           Q, ldq, sQ, h, sh, act, skip);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

}  // namespace

int coop_abort_check(ccab_ctx* ctx) {
    if (!ctx->coop_abort) return CCAB_OK;
    uint64_t v = 0;
    CCAB_HIP(hipMemcpyAsync(&v, ctx->coop_abort, 8, hipMemcpyDeviceToHost, ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    if (!v) return CCAB_OK;
    hipLaunchKernelGGL(coop_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->coop_abort,
                       int64_t(1), int64_t(1));
    return ctx_fail(ctx, CCAB_EHIP,
                    "Gram-Schmidt grid barrier timed out (a workgroup of the multi-workgroup "
                    "step was not resident; CCAB_MGS_COOP=0 keeps one workgroup per system)");
}

bool mgs_coop_fits() {
    auto fits = [](const void* f) {
        int nb = 0;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kMgsCoopThreads, 0) ==
                   hipSuccess && nb >= 1;
    };
    return fits(reinterpret_cast<const void*>(&mgs_block_coop_kernel<1, 4>)) &&
           fits(reinterpret_cast<const void*>(&mgs_block_coop_kernel<8, 4>)) &&
           fits(reinterpret_cast<const void*>(&mgs_block_coop_kernel<1, 8>)) &&
           fits(reinterpret_cast<const void*>(&mgs_block_coop_kernel<8, 8>)) &&
           fits(reinterpret_cast<const void*>(&mgs_coop_kernel<true, 1>)) &&
           fits(reinterpret_cast<const void*>(&mgs_coop_kernel<true, 8>)) &&
           fits(reinterpret_cast<const void*>(&mgs_coop_kernel<false, 1>)) &&
           fits(reinterpret_cast<const void*>(&mgs_coop_kernel<false, 8>));
}

namespace {

__global__ void scale_kernel(int64_t n, const c128* __restrict__ src, c128 s, c128* dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] * s;
}

__global__ void fill_zero_kernel(int64_t n, c128* dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = c128{0.0, 0.0};
}

__global__ void copy_c128_kernel(int64_t n, const c128* __restrict__ src, c128* dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void add_c128_kernel(int64_t n, const c128* __restrict__ d, c128* x) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + d[i];
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// dst_b = src_b / nrm[b] for the active systems (q0 = r / ||r||)
__global__ void scale_inv_norm_kernel(int64_t n, const c128* __restrict__ src, int64_t ss,
                                      const double* __restrict__ nrm, c128* dst, int64_t sd,
                                      const int32_t* __restrict__ active) {
    const int64_t b = active[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[b * sd + i] = src[b * ss + i] * crecip(c128{nrm[b], 0.0});
}

// dst_b = src_b for the active systems
__global__ void copy_active_kernel(int64_t n, const c128* __restrict__ src, c128* dst,
                                   const int32_t* __restrict__ active) {
    const int64_t b = active[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[b * n + i] = src[b * n + i];
}

// dst_b -= src_b for the active systems
__global__ void sub_active_kernel(int64_t n, const c128* __restrict__ src, c128* dst,
                                  const int32_t* __restrict__ active) {
    const int64_t b = active[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[b * n + i] = dst[b * n + i] - src[b * n + i];
}

// Warm start judged by the residual: with u = A x0 = b - r (r the FP64
// residual of x0), alpha = (u^H b) / (u^H u) minimises ||b - alpha A x0||,
// so the scaled guess is never worse than the zero start (alpha = 0 when u
// is orthogonal to b).  Then x0 <- alpha x0, r <- b - alpha u.  One
// workgroup per system, fixed-order reduction.
__global__ __launch_bounds__(256) void warm_scale_kernel(int64_t n, const c128* __restrict__ Bv,
                                                         c128* R, c128* X,
                                                         const int32_t* __restrict__ active) {
    __shared__ double red[3][4];
    __shared__ c128 s_alpha;
    const int64_t b = active[blockIdx.x];
    const c128* bb = Bv + b * n;
    c128* rr = R + b * n;
    c128* xx = X + b * n;
    double ure = 0.0, uim = 0.0, uu = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const c128 u = bb[i] - rr[i];
        const c128 d = conj(u) * bb[i];
        ure += d.re;
        uim += d.im;
        uu += norm_sqr(u);
    }
    for (int off = 32; off > 0; off >>= 1) {
        ure += __shfl_xor(ure, off, 64);
        uim += __shfl_xor(uim, off, 64);
        uu += __shfl_xor(uu, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = ure;
        red[1][threadIdx.x >> 6] = uim;
        red[2][threadIdx.x >> 6] = uu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const double c = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        const double d = red[2][0] + red[2][1] + red[2][2] + red[2][3];
        s_alpha = d > 0.0 ? c128{a / d, c / d} : c128{0.0, 0.0};
    }
    __syncthreads();
    const c128 al = s_alpha;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const c128 u = bb[i] - rr[i];
        xx[i] = al * xx[i];
        rr[i] = bb[i] - al * u;
    }
}

// ---------------------------------------------------------------------------
// Device side of the batched GMRES cycle (gmres.rs:158-200 for many systems):
// the Givens rotations, the rotated right-hand side beta and the stop test
// run next to the data, so the host never waits on an Arnoldi step.  Per
// system b: H (m+1) x m rotated Hessenberg columns, cs/sn the rotations, be =
// beta, kdone[b] = number of Arnoldi steps once the system's cycle is done
// (0 while it runs).  One thread per active system; same arithmetic as the
// host loop of gmres.rs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cabs_d(c128 z) { return hypot(z.re, z.im); }

__global__ void gir_cycle_init_kernel(int m, const int32_t* __restrict__ act, int nact,
                                      int32_t* kdone, c128* be, const double* __restrict__ rn) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nact) return;
    const int64_t b = act[z];
    kdone[b] = 0;
    be[b * (m + 1)] = c128{rn[b], 0.0};
}

// Column k from the Gram-Schmidt step (h[0..k+1], stride sh per system)
__global__ void gir_givens_kernel(int k, int m, const int32_t* __restrict__ act, int nact,
                                  const c128* __restrict__ hd, int64_t sh, c128* __restrict__ Hg,
                                  c128* __restrict__ csd, c128* __restrict__ snd,
                                  c128* __restrict__ bed, int32_t* __restrict__ kdone,
                                  const double* __restrict__ rn, const double* __restrict__ bn,
                                  double inner_tol, double tol, c128* __restrict__ Hbar,
                                  int32_t* __restrict__ flags_out, int hoff = 0,
                                  c128* __restrict__ Bsave = nullptr) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nact) return;
    const int64_t b = act[z];
    const int32_t kd = kdone[b];
    if (kd != 0) {
        if (flags_out) {
            flags_out[b] = kd;
            __threadfence_system();
        }
        return;
    }
    // (a deflated cycle: the first hoff entries are the coefficients against
    // the deflation vectors, kept in Bsave; the Hessenberg column follows)
    const c128* h = hd + b * sh + hoff;
    if (Bsave) {
        c128* bs = Bsave + b * int64_t(m) * m + int64_t(k) * m;
        for (int i = 0; i < hoff; ++i) bs[i] = hd[b * sh + i];
    }
    c128* col = Hg + b * int64_t(m + 1) * m + int64_t(k) * (m + 1);
    if (Hbar) {  // the unrotated column, for A Q_k = Q_(k+1) Hbar_k
        c128* hb = Hbar + b * int64_t(m + 1) * m + int64_t(k) * (m + 1);
        for (int i = 0; i <= k + 1; ++i) hb[i] = h[i];
    }
    c128* cs = csd + b * m;
    c128* sn = snd + b * m;
    c128* be = bed + b * (m + 1);
    // the previous rotations applied to the new column: entry i+1 enters
    // rotation i unrotated, entry i arrives rotated (carried in registers,
    // so the loads of cs, sn and h do not wait on the chain)
    c128 carry = h[0];
#pragma unroll 4
    for (int i = 0; i < k; ++i) {
        const c128 c = cs[i], sv = sn[i], nx = h[i + 1];
        col[i] = c * carry + sv * nx;
        carry = c * nx - conj(sv) * carry;
    }
    const c128 v1 = carry, v2 = h[k + 1];
    const double a1 = cabs_d(v1);
    const c128 signf = (v1.re == 0.0 && v1.im == 0.0) ? c128{1.0, 0.0} : c128{v1.re / a1, v1.im / a1};
    const double rr = sqrt(norm_sqr(v1) + norm_sqr(v2));
    const c128 ck{a1 / rr, 0.0};
    const c128 t = signf * conj(v2);
    const c128 sk{t.re / rr, t.im / rr};
    cs[k] = ck;
    sn[k] = sk;
    col[k] = ck * v1 + sk * v2;
    col[k + 1] = c128{0.0, 0.0};
    const c128 bk1 = -(conj(sk) * be[k]);
    be[k + 1] = bk1;
    be[k] = ck * be[k];
    const double res = cabs_d(bk1);
    const double est = res / rn[b];
    const bool done = est <= inner_tol || k == m - 1 || res <= 0.5 * tol * bn[b] || !isfinite(est);
    if (done) kdone[b] = k + 1;
    // the host's copy of the flag (pinned, device-visible memory)
    if (flags_out) {
        flags_out[b] = done ? k + 1 : 0;
        __threadfence_system();
    }
}

// y = R^{-1} beta over the system's kdone[b] steps (R the rotated, upper
// triangular H), zero beyond; column-oriented back substitution.  With Hbar:
// also z = Hbar y (kdone[b] + 1 entries), so that A d = Q z for d = Q y.
__global__ void gir_solve_y_kernel(int m, const int32_t* __restrict__ act, int nact,
                                   const c128* __restrict__ Hg, const c128* __restrict__ bed,
                                   const int32_t* __restrict__ kdone, c128* Yd,
                                   const c128* __restrict__ Hbar, c128* Zd) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nact) return;
    const int64_t b = act[z];
    const int kk = kdone[b];
    const c128* H = Hg + b * int64_t(m + 1) * m;
    const c128* be = bed + b * (m + 1);
    c128* y = Yd + b * (m + 1);
    for (int i = 0; i <= m; ++i) y[i] = i < kk ? be[i] : c128{0.0, 0.0};
    for (int j = kk - 1; j >= 0; --j) {
        const c128 yj = cdiv(y[j], H[j + int64_t(j) * (m + 1)]);
        y[j] = yj;
        for (int r = 0; r < j; ++r) y[r] = y[r] - yj * H[r + int64_t(j) * (m + 1)];
    }
    if (!Hbar) return;
    const c128* Hb = Hbar + b * int64_t(m + 1) * m;
    c128* zz = Zd + b * (m + 1);
    for (int i = 0; i <= m; ++i) {
        c128 acc{0.0, 0.0};
        if (i <= kk)  // column j holds rows 0..j+1
            for (int j = (i > 0 ? i - 1 : 0); j < kk; ++j) acc += Hb[i + int64_t(j) * (m + 1)] * y[j];
        zz[i] = acc;
    }
}

// The same for restart lengths m <= kSolveYLdsM: one wave per system, H (and
// Hbar) staged in LDS by all lanes at once, then the column-oriented back
// substitution with the row updates spread over the lanes, and z row-parallel.
constexpr int kSolveYLdsM = 50;

__global__ __launch_bounds__(64) void gir_solve_y_wave_kernel(
    int m, const int32_t* __restrict__ act, const c128* __restrict__ Hg,
    const c128* __restrict__ bed, const int32_t* __restrict__ kdone, c128* Yd,
    const c128* __restrict__ Hbar, c128* Zd) {
    __shared__ c128 sH[kSolveYLdsM * (kSolveYLdsM + 1)];
    __shared__ c128 sy[kSolveYLdsM + 1];
    const int lane = threadIdx.x;
    const int64_t b = act[blockIdx.x];
    const int kk = kdone[b];
    const int ld = m + 1;
    const c128* H = Hg + b * int64_t(ld) * m;
    for (int e = lane; e < ld * kk; e += 64) sH[e] = H[e];
    for (int i = lane; i <= m; i += 64) sy[i] = i < kk ? bed[b * ld + i] : c128{0.0, 0.0};
    __syncthreads();
    for (int j = kk - 1; j >= 0; --j) {
        if (lane == 0) sy[j] = cdiv(sy[j], sH[j + j * ld]);
        __syncthreads();
        const c128 yj = sy[j];
        for (int r = lane; r < j; r += 64) sy[r] = sy[r] - yj * sH[r + j * ld];
        __syncthreads();
    }
    c128* y = Yd + b * ld;
    for (int i = lane; i <= m; i += 64) y[i] = sy[i];
    if (!Hbar) return;
    // z = Hbar y: column j of Hbar holds rows 0..j+1
    const c128* Hb = Hbar + b * int64_t(ld) * m;
    c128* zz = Zd + b * ld;
    for (int i = lane; i <= m; i += 64) {
        c128 acc{0.0, 0.0};
        if (i <= kk)
            for (int j = (i > 0 ? i - 1 : 0); j < kk; ++j) acc += Hb[i + int64_t(j) * ld] * sy[j];
        zz[i] = acc;
    }
}

// x_b += Q_b y_b over the system's own kdone[b] basis vectors only (columns
// beyond them may hold anything: never written, or left from another cycle).
// With Zd: also d_b = Q_b y_b into D and r_b -= Q_b z_b (kdone[b] + 1
// vectors), the operator's share of r - A d.
__global__ void gir_update_x_kernel(int64_t n, int m, const int32_t* __restrict__ act,
                                    const int32_t* __restrict__ kdone, const c128* __restrict__ Q,
                                    int64_t sq, const c128* __restrict__ Yd, c128* X,
                                    const c128* __restrict__ Zd, c128* D, c128* R) {
    const int64_t b = act[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int kk = kdone[b];
    const c128* q = Q + b * sq + i;
    const c128* y = Yd + b * (m + 1);
    c128 acc{0.0, 0.0};
    for (int j = 0; j < kk; ++j) acc += q[int64_t(j) * n] * y[j];
    X[b * n + i] = X[b * n + i] + acc;
    if (!Zd) return;
    D[b * n + i] = acc;
    const c128* z = Zd + b * (m + 1);
    c128 az{0.0, 0.0};
    for (int j = 0; j <= kk; ++j)  // (q_kk is not finite after an exact breakdown, where z_kk = 0)
        if (z[j].re != 0.0 || z[j].im != 0.0) az += q[int64_t(j) * n] * z[j];
    R[b * n + i] = R[b * n + i] - az;
}

// ---- deflated refinement cycles (GCRO over the first cycle's space) -------
// After a system's first GMRES-IR cycle of k steps, A V_k = V_(k+1) Hbar with
// Hbar = Omega^H [R; 0] (Omega the cycle's Givens rotations, R in Hg).  Then
// C = V_(k+1) Qh, Qh = Omega^H [I_k; 0], has orthonormal columns and
// A (V_k R^-1) = C.  Later cycles keep C as the first basis vectors: each new
// Arnoldi vector is orthogonalised against C as well (coefficients B), the
// least-squares problem over the new vectors is the usual one, and the
// correction is d = V y + U (c0 - B y) with U = V_k R^-1 and c0 = C^H r.
constexpr int kRcMaxM = 64;   // largest restart length with deflated cycles
constexpr int kRcRows = 32;   // rows per workgroup of gir_rc_basis_kernel

// Qh ((m+1) x m per system, column c at c (m+1)) and R^-1 (m x m, column c
// at c m, rows 0..c) for the first kdone[b] steps; one wavefront per
// system, a lane per column.  Column c of Qh is the unit vector e_c through
// the inverse rotations, [[c, -s], [conj(s), c]] on rows (r, r + 1), from
// the last to the first: rotation c makes rows c, c + 1 nonzero, each later
// one (r = c - 1 .. 0) moves the value of row r + 1 up one row.
__global__ __launch_bounds__(64) void gir_rc_factors_kernel(int m, const int32_t* __restrict__ act,
                                                            const int32_t* __restrict__ kdone,
                                                            const c128* __restrict__ Hg,
                                                            const c128* __restrict__ csd,
                                                            const c128* __restrict__ snd, c128* Qh,
                                                            c128* Ri) {
    const int64_t b = act[blockIdx.x];
    const int k = kdone[b];
    const int ld = m + 1;
    const c128* H = Hg + b * int64_t(ld) * m;  // R(i, j) = H[i + j ld], i <= j
    const c128* cs = csd + b * m;
    const c128* sn = snd + b * m;
    for (int c = threadIdx.x; c < k; c += 64) {
        c128* q = Qh + b * int64_t(ld) * m + int64_t(c) * ld;
        for (int i = c + 2; i <= k; ++i) q[i] = c128{0.0, 0.0};
        if (c + 1 <= k) q[c + 1] = conj(sn[c]);
        c128 carry = cs[c];
        for (int r = c - 1; r >= 0; --r) {
            q[r + 1] = cs[r] * carry;
            carry = -(sn[r] * carry);
        }
        q[0] = carry;
        c128* ri = Ri + b * int64_t(m) * m + int64_t(c) * m;
        for (int i = c + 1; i < k; ++i) ri[i] = c128{0.0, 0.0};
        ri[c] = crecip(H[c + c * ld]);
        for (int i = c - 1; i >= 0; --i) {
            c128 acc{0.0, 0.0};
            for (int j = i + 1; j <= c; ++j) acc += H[i + j * ld] * ri[j];
            ri[i] = -cdiv(acc, H[i + i * ld]);
        }
    }
}

// Harmonic Ritz directions of a system's first cycle, for deflation by
// C = A Y rather than by the first directions of A K_k.  With
// Hbar = Qh R (gir_rc_factors_kernel), the harmonic Ritz pairs of the cycle,
// Hbar^H Hbar g = theta H_k^H g, are the eigenpairs of W = R^-1 Qh_top^H with
// eigenvalues 1 / theta (Qh_top: the first k rows of Qh).  The s = min(8, k)
// of smallest |theta| span the dominant invariant subspace of W, found by
// kRitzIters steps of subspace iteration (Z <- orth(W Z) from a fixed
// pseudo-random start).  With P = Z and R P = Q2 R2 (thin QR), the
// deflation space is C = V_(k+1) (Qh Q2) (orthonormal) and U = V_k (P R2^-1),
// A U = C.  Writes Cq = Qh Q2 ((m+1) x m layout) and Uq = P R2^-1 (m x m
// layout) for columns 0..s-1.  One wavefront per system, a lane per row;
// k <= kRitzMaxK.
constexpr int kRitzVecs = 8;
constexpr int kRitzMaxK = 32;
constexpr int kRitzIters = 20;

__device__ __forceinline__ c128 wave_csum(c128 v) { return c128{wave_sum(v.re), wave_sum(v.im)}; }

// Modified Gram-Schmidt of the s columns held one row per lane (rows >= k
// hold zeros); r (may be null) receives the s x s triangular factor.
// Returns false (the same on every lane: lane 0's norms decide) when a
// column loses more than all but 1e-10 of its norm.
__device__ __forceinline__ bool wave_mgs(c128 (&z)[kRitzVecs], int s, c128* r) {
    bool ok = true;
#pragma unroll
    for (int c = 0; c < kRitzVecs; ++c) {
        if (c >= s) break;
        const double n0 = __shfl(sqrt(wave_sum(norm_sqr(z[c]))), 0, 64);
#pragma unroll
        for (int c2 = 0; c2 < kRitzVecs; ++c2) {
            if (c2 >= c) break;
            const c128 d = wave_csum(conj(z[c2]) * z[c]);
            z[c] = z[c] - d * z[c2];
            if (r && threadIdx.x == 0) r[c2 + c * kRitzVecs] = d;
        }
        const double nn = __shfl(sqrt(wave_sum(norm_sqr(z[c]))), 0, 64);
        if (!(nn > 1e-10 * n0) || !(nn < 1e300)) {
            ok = false;
            z[c] = c128{0.0, 0.0};
            if (r && threadIdx.x == 0) r[c + c * kRitzVecs] = c128{1.0, 0.0};
            continue;
        }
        z[c] = z[c] * crecip(c128{nn, 0.0});
        if (r && threadIdx.x == 0) r[c + c * kRitzVecs] = c128{nn, 0.0};
    }
    return ok;
}

__global__ __launch_bounds__(64) void gir_rc_ritz_kernel(int m, int nvec,
                                                         const int32_t* __restrict__ act,
                                                         const int32_t* __restrict__ kdone,
                                                         const c128* __restrict__ Hg,
                                                         const c128* __restrict__ Qh,
                                                         const c128* __restrict__ Ri, c128* Cq,
                                                         c128* Uq) {
    // column-major in LDS ([column][row]): lane i reads row i of a column
    // (consecutive addresses), a fixed entry is read by every lane (broadcast)
    __shared__ c128 sA[kRitzMaxK][kRitzMaxK];       // R^-1, later R
    __shared__ c128 sQ[kRitzMaxK][kRitzMaxK + 1];   // Qh, rows 0..k
    __shared__ c128 sW[kRitzMaxK][kRitzMaxK];
    __shared__ c128 sZ[kRitzMaxK][kRitzVecs];       // [row][column]
    __shared__ c128 sR2[kRitzVecs * kRitzVecs];
    const int64_t b = act[blockIdx.x];
    const int k = kdone[b];
    const int s = min(min(nvec, kRitzVecs), k);
    const int i = threadIdx.x;
    const int ld = m + 1;
    const c128* H = Hg + b * int64_t(ld) * m;      // R(i, j) = H[i + j ld]
    const c128* q = Qh + b * int64_t(ld) * m;      // Qh(i, c) = q[i + c ld]
    const c128* ri = Ri + b * int64_t(m) * m;      // R^-1(i, c) = ri[i + c m]
    for (int c = 0; c < k; ++c) {
        if (i < k) sA[c][i] = ri[i + c * m];
        if (i <= k) sQ[c][i] = q[i + c * ld];
    }
    __syncthreads();
    // W = R^-1 Qh_top^H: W(i, j) = sum_l R^-1(i, l) conj(Qh(j, l)), l >= i
    if (i < k)
        for (int j = 0; j < k; ++j) {
            c128 acc{0.0, 0.0};
            for (int l = i; l < k; ++l) acc += sA[l][i] * conj(sQ[l][j]);
            sW[j][i] = acc;
        }
    c128 z[kRitzVecs];
#pragma unroll
    for (int c = 0; c < kRitzVecs; ++c) {
        // a fixed pseudo-random start (integer hash of row and column)
        uint32_t h = uint32_t(i) * 2654435761u ^ uint32_t(c + 1) * 40503u;
        h ^= h >> 13;
        h *= 0x5bd1e995u;
        h ^= h >> 15;
        const double a = double(h & 0xffff) / 32768.0 - 1.0, bb = double(h >> 16) / 32768.0 - 1.0;
        z[c] = (i < k && c < s) ? c128{a, bb} : c128{0.0, 0.0};
    }
    bool ok = wave_mgs(z, s, nullptr);
    __syncthreads();   // sW complete
    // (two products per orthonormalisation: the wave reductions of the
    // Gram-Schmidt steps, not the products, set this kernel's time)
    for (int it = 0; it < kRitzIters && ok; ++it) {
        if (i < k)
#pragma unroll
            for (int c = 0; c < kRitzVecs; ++c) sZ[i][c] = z[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c) z[c] = c128{0.0, 0.0};
        if (i < k)
            for (int j = 0; j < k; ++j) {
                const c128 w = sW[j][i];
#pragma unroll
                for (int c = 0; c < kRitzVecs; ++c) z[c] += w * sZ[j][c];
            }
        __syncthreads();
        if (it % 2 == 1 || it == kRitzIters - 1) ok = wave_mgs(z, s, nullptr);
    }
    // P = z; R P (rows i < k), then its thin QR: Q2 and R2
    if (i < k)
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c) sZ[i][c] = z[c];
    for (int c = 0; c < k; ++c)
        if (i < k) sA[c][i] = i <= c ? H[i + c * ld] : c128{0.0, 0.0};
    __syncthreads();
    c128 p[kRitzVecs];
#pragma unroll
    for (int c = 0; c < kRitzVecs; ++c) {
        p[c] = z[c];
        z[c] = c128{0.0, 0.0};
    }
    if (i < k)
        for (int l = i; l < k; ++l) {
            const c128 w = sA[l][i];
#pragma unroll
            for (int c = 0; c < kRitzVecs; ++c) z[c] += w * sZ[l][c];
        }
    if (ok) ok = wave_mgs(z, s, sR2);   // z = Q2
    __syncthreads();
    c128* cq = Cq + b * int64_t(ld) * m;
    c128* uq = Uq + b * int64_t(m) * m;
    if (!ok) {
        // (degenerate: the first s directions of A K_k, Qh and R^-1 with its
        // zero lower triangle, as without harmonic Ritz vectors)
        for (int c = 0; c < s; ++c) {
            if (i <= k) cq[i + c * ld] = q[i + c * ld];
            if (i < k) uq[i + c * m] = i <= c ? ri[i + c * m] : c128{0.0, 0.0};
        }
        return;
    }
    if (i < k)
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c) sZ[i][c] = z[c];
    __syncthreads();
    // Cq = Qh Q2 (rows 0..k), Uq = P R2^-1 (rows 0..k-1)
    if (i <= k) {
        c128 acc[kRitzVecs];
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c) acc[c] = c128{0.0, 0.0};
        for (int l = 0; l < k; ++l) {
            const c128 w = sQ[l][i];
#pragma unroll
            for (int c = 0; c < kRitzVecs; ++c) acc[c] += w * sZ[l][c];
        }
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c)
            if (c < s) cq[i + c * ld] = acc[c];
    }
    if (i < k) {
        c128 u[kRitzVecs];
#pragma unroll
        for (int c = 0; c < kRitzVecs; ++c) {
            if (c >= s) break;
            c128 v = p[c];
#pragma unroll
            for (int l = 0; l < kRitzVecs; ++l)
                if (l < c) v = v - u[l] * sR2[l + c * kRitzVecs];
            u[c] = cdiv(v, sR2[c + c * kRitzVecs]);
            uq[i + c * m] = u[c];
        }
    }
}

// The first s columns of C = V_(k+1) Cq into basis slots m - s .. m - 1 and
// of U = V_k Uq into U (stride su per system); columns k .. s - 1 of both
// zero where k < s.  (Cq, Uq) = (Qh, R^-1) for the first directions of
// A K_k, or the harmonic Ritz coefficients of gir_rc_ritz_kernel.  The Krylov basis V starts at slot m; each workgroup
// stages the k + 1 entries of its 32 rows in LDS, two lanes per row.
__global__ __launch_bounds__(64) void gir_rc_basis_kernel(int64_t n, int m, int s,
                                                          const int32_t* __restrict__ act,
                                                          const int32_t* __restrict__ kdone,
                                                          c128* Q, int64_t sq,
                                                          const c128* __restrict__ Qh,
                                                          const c128* __restrict__ Ri, c128* U,
                                                          int64_t su) {
    __shared__ c128 v[kRcMaxM + 1][kRcRows];
    const int64_t b = act[blockIdx.y];
    const int k = kdone[b];
    const int r = threadIdx.x % kRcRows, half = threadIdx.x / kRcRows;
    const int64_t row0 = int64_t(blockIdx.x) * kRcRows;
    c128* q = Q + b * sq;
    const c128* V = q + int64_t(m) * n;
    for (int e = threadIdx.x; e < (k + 1) * kRcRows; e += 64) {
        const int j = e / kRcRows, rr = e % kRcRows;
        const int64_t i = row0 + rr;
        v[j][rr] = i < n ? V[int64_t(j) * n + i] : c128{0.0, 0.0};
    }
    __syncthreads();
    const int64_t i = row0 + r;
    if (i >= n) return;
    const c128* qh = Qh + b * int64_t(m + 1) * m;
    const c128* ri = Ri + b * int64_t(m) * m;
    c128* C = q + int64_t(m - s) * n;
    c128* u = U + b * su;
    for (int c = half; c < s; c += 2) {
        c128 ac{0.0, 0.0}, au{0.0, 0.0};
        if (c < k) {
            for (int j = 0; j <= k; ++j) {
                // (row k of Qh is zero after an exact breakdown, where v_k is not finite)
                const c128 w = qh[j + c * (m + 1)];
                if (w.re != 0.0 || w.im != 0.0) ac += v[j][r] * w;
            }
            for (int j = 0; j < k; ++j) au += v[j][r] * ri[j + c * m];
        }
        C[int64_t(c) * n + i] = ac;
        u[int64_t(c) * n + i] = au;
    }
}

// Start of a deflated cycle, after the Gram-Schmidt step of the scaled
// residual r / rn against C: h[0..s) = C^H r / rn, h[s] = |(I - C C^H) r| / rn.
// c0 = C^H r; the rotated right-hand side starts at |(I - C C^H) r| (the
// stop test stays relative to |r|).
__global__ void gir_rc_start_kernel(int m, int s, const int32_t* __restrict__ act, int nact,
                                    int32_t* kdone, c128* be, const double* __restrict__ rn,
                                    const c128* __restrict__ hd, int64_t sh, c128* c0) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nact) return;
    const int64_t b = act[z];
    const c128* h = hd + b * sh;
    const c128 sc{rn[b], 0.0};
    for (int i = 0; i < s; ++i) c0[b * m + i] = h[i] * sc;
    kdone[b] = 0;
    be[b * (m + 1)] = c128{h[s].re * rn[b], 0.0};
}

// z = c0 - B y over the cycle's kdone[b] steps (column j of B: C^H A v_j,
// m entries per column, saved by gir_givens_kernel)
__global__ void gir_rc_z_kernel(int m, int s, const int32_t* __restrict__ act, int nact,
                                const int32_t* __restrict__ kdone, const c128* __restrict__ Bs,
                                const c128* __restrict__ Yd, const c128* __restrict__ c0,
                                c128* Zc) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nact) return;
    const int64_t b = act[z];
    const int kk = kdone[b];
    const c128* y = Yd + b * (m + 1);
    const c128* B = Bs + b * int64_t(m) * m;
    for (int i = 0; i < s; ++i) {
        c128 acc = c0[b * m + i];
        for (int j = 0; j < kk; ++j) acc -= B[int64_t(j) * m + i] * y[j];
        Zc[b * m + i] = acc;
    }
}

// x += d, d = V y + U z over the system's kdone[b] Krylov vectors and the s
// deflation vectors; with Zd also D = d and r -= V (Hbar y) + C c0 (the
// operator's share of r - A d: A d = C c0 + V Hbar y).  Terms with a zero
// coefficient are skipped (a basis vector after an exact breakdown is not
// finite).
__global__ void gir_update_x_rc_kernel(int64_t n, int m, int s, const int32_t* __restrict__ act,
                                       const int32_t* __restrict__ kdone,
                                       const c128* __restrict__ Q, int64_t sq,
                                       const c128* __restrict__ Yd, const c128* __restrict__ U,
                                       int64_t su, const c128* __restrict__ Zc,
                                       const c128* __restrict__ c0, c128* X,
                                       const c128* __restrict__ Zd, c128* D, c128* R) {
    const int64_t b = act[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int kk = kdone[b];
    const c128* V = Q + b * sq + int64_t(m) * n + i;
    const c128* C = Q + b * sq + int64_t(m - s) * n + i;
    const c128* u = U + b * su + i;
    const c128* y = Yd + b * (m + 1);
    const c128* zc = Zc + b * m;
    c128 acc{0.0, 0.0};
    for (int j = 0; j < kk; ++j)
        if (y[j].re != 0.0 || y[j].im != 0.0) acc += V[int64_t(j) * n] * y[j];
    for (int c = 0; c < s; ++c)
        if (zc[c].re != 0.0 || zc[c].im != 0.0) acc += u[int64_t(c) * n] * zc[c];
    X[b * n + i] = X[b * n + i] + acc;
    if (!Zd) return;
    D[b * n + i] = acc;
    const c128* z = Zd + b * (m + 1);
    const c128* cb = c0 + b * m;
    c128 az{0.0, 0.0};
    for (int j = 0; j <= kk; ++j)
        if (z[j].re != 0.0 || z[j].im != 0.0) az += V[int64_t(j) * n] * z[j];
    for (int c = 0; c < s; ++c)
        if (cb[c].re != 0.0 || cb[c].im != 0.0) az += C[int64_t(c) * n] * cb[c];
    R[b * n + i] = R[b * n + i] - az;
}

using cd = std::complex<double>;

// Host complex LU solve (partial pivoting on |re|+|im|) for the small
// Hessenberg system of get_x (gmres.rs:146-156).
void host_lu_solve(int64_t k, std::vector<cd>& A, std::vector<cd>& x) {
    std::vector<int64_t> piv(k);
    for (int64_t i = 0; i < k; ++i) {
        int64_t p = i;
        double best = std::abs(A[i + i * k].real()) + std::abs(A[i + i * k].imag());
        for (int64_t r = i + 1; r < k; ++r) {
            const double v = std::abs(A[r + i * k].real()) + std::abs(A[r + i * k].imag());
            if (v > best) {
                best = v;
                p = r;
            }
        }
        piv[i] = p;
        const cd d = A[p + i * k];
        if (d == cd(0.0, 0.0)) continue;
        if (p != i)
            for (int64_t c = 0; c < k; ++c) std::swap(A[i + c * k], A[p + c * k]);
        const cd inv = 1.0 / d;
        for (int64_t r = i + 1; r < k; ++r) A[r + i * k] *= inv;
        for (int64_t c = i + 1; c < k; ++c)
            for (int64_t r = i + 1; r < k; ++r) A[r + c * k] -= A[i + c * k] * A[r + i * k];
    }
    for (int64_t i = 0; i < k; ++i)
        if (piv[i] != i) std::swap(x[i], x[piv[i]]);
    for (int64_t j = 0; j < k; ++j)
        for (int64_t r = j + 1; r < k; ++r) x[r] -= x[j] * A[r + j * k];
    for (int64_t j = k - 1; j >= 0; --j) {
        x[j] /= A[j + j * k];
        for (int64_t r = 0; r < j; ++r) x[r] -= x[j] * A[r + j * k];
    }
}

}  // namespace

// GMRES over any operator y = alpha A x + beta y (gmres.rs:35-202); anorm is
// the norm the backward error uses (dense: Frobenius; H matrix: the sum of
// its block norms, h_matrix.rs:189-196).
int gmres_op_run(ccab_ctx* ctx, int64_t n, const MatvecFn& apply, double anorm, c128* d_x,
                 const c128* d_x0, const ccab_gmres_params& p, ccab_solve_stats* st) {
    int64_t max_it = p.max_iterations <= 0 ? 1000 : p.max_iterations;
    max_it = std::min<int64_t>(max_it, n);
    const double thresh = p.tolerance <= 0.0 ? 1.0e-5 : p.tolerance;
    const int64_t m = p.restart > 0 ? std::min<int64_t>(p.restart, max_it)
                                    : (int64_t)std::ceil(std::sqrt((double)max_it));
    const int64_t max_restarts = std::max<int64_t>(1, max_it / m);
    // device buffers: b, r, Q (n x (m+1)), h (m+2), y (m+1), norms
    const size_t nb = size_t(n) * 16;
    const size_t need = 2 * nb + nb * (m + 1) + 16 * (m + 2) + 16 * (m + 1) + 64;
    char* base = static_cast<char*>(ctx_buf(ctx, SLOT_GMRES, need + 4096));
    if (!base) return CCAB_ENOMEM;
    c128* b = reinterpret_cast<c128*>(base);
    c128* r = b + n;
    c128* Q = r + n;
    c128* h = Q + n * (m + 1);
    c128* y = h + (m + 2);
    double* nrm = reinterpret_cast<double*>(y + (m + 1));
    CCAB_TRY(pinned_reserve(ctx, 16 * (m + 2) + 64));
    double* hp = static_cast<double*>(ctx->pinned);
    hipStream_t s = ctx->stream;
    auto read_norm = [&](const c128* v, double* out) -> int {
        CCAB_TRY(nrm2_batched(ctx, n, 1, v, n, nrm));
        CCAB_HIP(hipMemcpyAsync(hp, nrm, 8, hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        *out = hp[0];
        return CCAB_OK;
    };
    int64_t num_mv = 0, iters = 0, restarts = 0;
    auto matvec = [&](const c128* xin, c128 alpha, c128 beta, c128* yout) -> int {
        ++num_mv;
        return apply(xin, alpha, beta, yout);
    };
    hipLaunchKernelGGL(copy_c128_kernel, grid1(n), dim3(256), 0, s, n, d_x, b);
    if (d_x0)
        hipLaunchKernelGGL(copy_c128_kernel, grid1(n), dim3(256), 0, s, n, d_x0, d_x);
    else
        hipLaunchKernelGGL(fill_zero_kernel, grid1(n), dim3(256), 0, s, n, d_x);
    double bnorm = 0.0;
    CCAB_TRY(read_norm(b, &bnorm));
    const bool rel = p.stop_rule == CCAB_STOP_RELATIVE_RESIDUAL;
    auto err_of = [&](double rn, double xn) { return rel ? rn / bnorm : rn / (anorm * xn + bnorm); };

    std::vector<cd> cs(m), sn(m), beta(m + 1), H((m + 1) * m), hk(m + 2);
    double error = 0.0, rn = 0.0, xn = 0.0;
    bool converged = false;
    const c128 one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
    for (int64_t rs = 0; rs < max_restarts && !converged; ++rs) {
        ++restarts;
        hipLaunchKernelGGL(copy_c128_kernel, grid1(n), dim3(256), 0, s, n, b, r);
        CCAB_TRY(matvec(d_x, mone, one, r));
        CCAB_TRY(read_norm(r, &rn));
        CCAB_TRY(read_norm(d_x, &xn));
        error = err_of(rn, xn);
        if (!(rn > 0.0)) {  // exact solution (or b = 0)
            converged = error < thresh || rn == 0.0;
            break;
        }
        std::fill(cs.begin(), cs.end(), cd(0));
        std::fill(sn.begin(), sn.end(), cd(0));
        std::fill(beta.begin(), beta.end(), cd(0));
        std::fill(H.begin(), H.end(), cd(0));
        const cd inv_rn = cd(1.0, 0.0) / cd(rn, 0.0);
        hipLaunchKernelGGL(scale_kernel, grid1(n), dim3(256), 0, s, n, r,
                           c128{inv_rn.real(), inv_rn.imag()}, Q);
        beta[0] = cd(rn, 0.0);
        for (int64_t k = 0; k < m; ++k) {
            ++iters;
            CCAB_TRY(matvec(Q + k * n, one, zero, Q + (k + 1) * n));
            ProfScope ps(ctx, PROF_MGS, double(k + 2) * 2.0 * double(n) * 16.0);
            CCAB_TRY(launch_mgs(ctx, p.ref_mgs != 0, n, k, Q, n, 0, h, 0, nullptr, 1));
            CCAB_HIP(hipMemcpyAsync(hp, h, 16 * (k + 2), hipMemcpyDeviceToHost, s));
            CCAB_HIP(hipStreamSynchronize(s));
            for (int64_t i = 0; i < k + 2; ++i) hk[i] = cd(hp[2 * i], hp[2 * i + 1]);
            // apply_givens_rotation (gmres.rs:178-194)
            for (int64_t i = 0; i < k; ++i) {
                const cd temp = cs[i] * hk[i] + sn[i] * hk[i + 1];
                hk[i + 1] = -std::conj(sn[i]) * hk[i] + cs[i] * hk[i + 1];
                hk[i] = temp;
            }
            {
                const cd v1 = hk[k], v2 = hk[k + 1];
                const cd signf = (v1 == cd(0.0, 0.0)) ? cd(1.0, 0.0) : v1 / std::abs(v1);
                const double rr = std::sqrt(std::norm(v1) + std::norm(v2));
                cs[k] = cd(std::abs(v1) / rr, 0.0);
                sn[k] = signf * std::conj(v2) / rr;
            }
            hk[k] = cs[k] * hk[k] + sn[k] * hk[k + 1];
            hk[k + 1] = cd(0.0, 0.0);
            for (int64_t i = 0; i < k + 2; ++i) H[i + k * (m + 1)] = hk[i];
            beta[k + 1] = -std::conj(sn[k]) * beta[k];
            beta[k] *= cs[k];
            error = err_of(std::abs(beta[k + 1]), xn);
            if (!std::isfinite(error)) {
                CCAB_TRY(coop_abort_check(ctx));
                return ctx_fail(ctx, CCAB_ENAN, "GMRES error is NaN");
            }
            if (error < thresh || k == m - 1) {
                const int64_t kk = k + 1;
                std::vector<cd> Hk(kk * kk), yv(kk);
                for (int64_t c = 0; c < kk; ++c)
                    for (int64_t q = 0; q < kk; ++q) Hk[q + c * kk] = H[q + c * (m + 1)];
                for (int64_t i = 0; i < kk; ++i) yv[i] = beta[i];
                host_lu_solve(kk, Hk, yv);
                for (int64_t i = 0; i < kk; ++i) {
                    hp[2 * i] = yv[i].real();
                    hp[2 * i + 1] = yv[i].imag();
                }
                CCAB_HIP(hipMemcpyAsync(y, hp, 16 * kk, hipMemcpyHostToDevice, s));
                // x += Q[:, :kk] y
                CCAB_TRY(gemv_general(ctx, CCAB_C128, n, kk, 1, Q, n, 0, y, 0, one, one, d_x, 0));
                hipLaunchKernelGGL(copy_c128_kernel, grid1(n), dim3(256), 0, s, n, b, r);
                CCAB_TRY(matvec(d_x, mone, one, r));
                CCAB_TRY(read_norm(r, &rn));
                CCAB_TRY(read_norm(d_x, &xn));
                error = err_of(rn, xn);
                if (error < thresh) {
                    converged = true;
                    break;
                }
            }
        }
    }
    if (st) {
        st->num_mv = num_mv;
        st->iterations = iters;
        st->restarts = restarts;
        st->backward_error = rel ? rn / (anorm * xn + bnorm) : error;
        st->rel_residual = rn / bnorm;
        st->converged = converged ? 1 : 0;
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int gmres_run(ccab_ctx* ctx, int64_t n, int prec, const void* d_A, int64_t lda, c128* d_x,
              const c128* d_x0, const ccab_gmres_params& p, ccab_solve_stats* st) {
    double* nrm = static_cast<double*>(ctx_buf(ctx, SLOT_GMRES_NRM, 64));
    if (!nrm) return CCAB_ENOMEM;
    CCAB_TRY(pinned_reserve(ctx, 64));
    double* hp = static_cast<double*>(ctx->pinned);
    CCAB_TRY(frob_norm(ctx, prec, n, d_A, lda, nrm));
    CCAB_HIP(hipMemcpyAsync(hp, nrm, 8, hipMemcpyDeviceToHost, ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    const double anorm = hp[0];
    MatvecFn dense = [&](const c128* xin, c128 alpha, c128 beta, c128* yout) -> int {
        return gemv_general(ctx, prec, n, n, 1, d_A, lda, 0, xin, 0, alpha, beta, yout, 0);
    };
    return gmres_op_run(ctx, n, dense, anorm, d_x, d_x0, p, st);
}

// Mixed precision: x = (LU_c64)^{-1} b, then x += (LU_c64)^{-1} (b - A128 x)
int mixed_run(ccab_ctx* ctx, int64_t n, int64_t batch, const c128* A128, c64* A64, int64_t lda,
              int64_t strideA, c128* d_x, double ir_tol, int ir_max, ccab_solve_stats* st) {
    if (ir_tol <= 0.0) ir_tol = 1e-12;
    if (ir_max <= 0) ir_max = 10;
    hipStream_t s = ctx->stream;
    const int64_t tot = n * batch;
    c128* B = static_cast<c128*>(ctx_buf(ctx, SLOT_MIXED_B, size_t(tot) * 16));
    c128* R = static_cast<c128*>(ctx_buf(ctx, SLOT_MIXED_R, size_t(tot) * 16));
    c64* D = static_cast<c64*>(ctx_buf(ctx, SLOT_MIXED_D, size_t(tot) * 8));
    double* nrm = static_cast<double*>(ctx_buf(ctx, SLOT_MIXED_NORMS, size_t(batch) * 3 * 8));
    int32_t* piv = static_cast<int32_t*>(ctx_buf(ctx, SLOT_SOLVE_PIV, size_t(tot) * 4));
    if (!B || !R || !D || !nrm || !piv) return CCAB_ENOMEM;
    CCAB_TRY(pinned_reserve(ctx, size_t(batch) * 3 * 8));
    hipLaunchKernelGGL(copy_c128_kernel, grid1(tot), dim3(256), 0, s, tot, d_x, B);
    CCAB_TRY(lu_factor(ctx, CCAB_C64, n, batch, A64, lda, strideA, piv));
    CCAB_TRY(convert_c128_to_c64(ctx, tot, B, D));
    CCAB_TRY(lu_solve(ctx, CCAB_C64, n, batch, A64, lda, strideA, piv, D, n, n, 1));
    CCAB_TRY(convert_c64_to_c128(ctx, tot, D, d_x));
    const c128 one{1.0, 0.0}, mone{-1.0, 0.0};
    int steps = 0;
    std::vector<double> dn(batch), xn(batch);
    for (; steps < ir_max;) {
        hipLaunchKernelGGL(copy_c128_kernel, grid1(tot), dim3(256), 0, s, tot, B, R);
        CCAB_TRY(gemv_general(ctx, CCAB_C128, n, n, batch, A128, lda, strideA, d_x, n, mone, one,
                              R, n));
        CCAB_TRY(convert_c128_to_c64(ctx, tot, R, D));
        CCAB_TRY(lu_solve(ctx, CCAB_C64, n, batch, A64, lda, strideA, piv, D, n, n, 1));
        CCAB_TRY(convert_c64_to_c128(ctx, tot, D, R));
        hipLaunchKernelGGL(add_c128_kernel, grid1(tot), dim3(256), 0, s, tot, R, d_x);
        ++steps;
        CCAB_TRY(nrm2_batched(ctx, n, batch, R, n, nrm));
        CCAB_TRY(nrm2_batched(ctx, n, batch, d_x, n, nrm + batch));
        double* hp = static_cast<double*>(ctx->pinned);
        CCAB_HIP(hipMemcpyAsync(hp, nrm, size_t(batch) * 16, hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        bool done = true;
        for (int64_t b = 0; b < batch; ++b) {
            if (!std::isfinite(hp[b]) || !std::isfinite(hp[batch + b]))
                return ctx_fail(ctx, CCAB_ENAN, "refinement produced NaN in system %lld",
                                (long long)b);
            if (hp[b] > ir_tol * hp[batch + b]) done = false;
        }
        if (done) break;
    }
    if (st) {
        // final FP64 residual
        hipLaunchKernelGGL(copy_c128_kernel, grid1(tot), dim3(256), 0, s, tot, B, R);
        CCAB_TRY(gemv_general(ctx, CCAB_C128, n, n, batch, A128, lda, strideA, d_x, n, mone, one,
                              R, n));
        CCAB_TRY(nrm2_batched(ctx, n, batch, R, n, nrm));
        CCAB_TRY(nrm2_batched(ctx, n, batch, B, n, nrm + batch));
        CCAB_TRY(nrm2_batched(ctx, n, batch, d_x, n, nrm + 2 * batch));
        double* hp = static_cast<double*>(ctx->pinned);
        CCAB_HIP(hipMemcpyAsync(hp, nrm, size_t(batch) * 24, hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        for (int64_t b = 0; b < batch; ++b) {
            st[b] = ccab_solve_stats{};
            st[b].ir_steps = steps;
            st[b].rel_residual = hp[b] / hp[batch + b];
            st[b].backward_error = -1.0;  // needs ||A||_F: not computed on this path
            st[b].converged = 1;
        }
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

// ---------------------------------------------------------------------------
// Batched GMRES with FP64 iterative refinement (no reference counterpart: it
// replaces per-frequency LU/GMRES in the sweep).  Every cycle computes the
// FP64 residual r = b - A x with the c128 matrix (with the FP16 operator plus
// its complex-FP32 remainder Alo when Alo is given; with the c64 operator
// when neither is), then runs one GMRES(m) cycle on the c64 operator for A d = r (FP64
// Krylov vectors, modified Gram-Schmidt in FP64) until ||r - A d|| <=
// inner_tol ||r||, and updates x += d.  All systems advance together: one
// batched c64 matvec and one batched MGS launch per Arnoldi step, the Givens
// rotations of every system on the host; converged systems leave the active
// set.  Stops when ||b - A x|| <= tol ||b|| (FP64).  AX0 (with X0): A x0
// formed by the caller (the assembly multiplies each FP64 entry by the guess
// as it forms it), so the warm start's first residual b - A x0 reads no
// matrix.  The cycles after a system's first are deflated (GCRO) by a few
// directions of the first cycle's Krylov space (gir_rc_* kernels,
// ccab_set_gir_deflation), except in HIP-graph cycles.
// ---------------------------------------------------------------------------
int gmres_ir_run(ccab_ctx* ctx, int64_t n, int64_t B, const c128* A128, const void* Aop,
                 int op_prec, const double* d_inv_scale, int64_t lda, int64_t sA, c128* X,
                 const c128* X0, int m, double inner_tol, double tol, int max_cycles,
                 bool ref_mgs, ccab_solve_stats* st, Ragged* rg, const c64* Alo,
                 const c128* AX0) {
    if (AX0 && !X0) return ctx_fail(ctx, CCAB_EINVAL, "A x0 without the guess x0");
    if (op_prec == CCAB_C32 && (!(A128 || Alo) || !d_inv_scale))
        return ctx_fail(ctx, CCAB_EINVAL,
                        "the FP16 operator needs scales and the c128 matrix or its FP32 remainder");
    if (Alo && op_prec != CCAB_C32)
        return ctx_fail(ctx, CCAB_EINVAL, "the FP32 remainder goes with the FP16 operator");
    if (m <= 0) m = 50;
    // a Krylov space cannot exceed the smallest system
    m = (int)std::min<int64_t>(m, rg ? rg->n_min : n);
    if (inner_tol <= 0) inner_tol = 1e-4;
    if (tol <= 0) tol = 1e-10;
    if (max_cycles <= 0) max_cycles = 20;
    hipStream_t s = ctx->stream;
    const size_t nb = size_t(n) * B;
    // HIP graphs (ccab_set_graphs / CCAB_GRAPHS, not while profiling): a
    // cycle over na systems is captured once per key and replayed
    const bool graph_cycles = ctx->use_graphs && !ctx->prof_on && ctx->stream;
    // deflated cycles after the first (gir_rc_*): the basis buffer holds m
    // slots for the deflation vectors C, then the Krylov basis V (m + 1)
    const bool defl = ctx->gir_recycle > 0 && !graph_cycles && m <= kRcMaxM;
    const int64_t qslots = defl ? 2 * int64_t(m) + 1 : int64_t(m) + 1;
    const int64_t sh = defl ? 2 * int64_t(m) + 2 : int64_t(m) + 2;  // Gram-Schmidt column stride
    c128* Bv = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_B, nb * 16));
    c128* R = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_R, nb * 16));
    c128* Q = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_Q, nb * qslots * 16));
    c128* Hd = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_H, size_t(B) * sh * 16));
    c128* Yd = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_Y, size_t(B) * (m + 1) * 16));
    double* nrm = static_cast<double*>(ctx_buf(ctx, SLOT_GIR_NRM, size_t(B) * 8));
    int32_t* act = static_cast<int32_t*>(ctx_buf(ctx, SLOT_GIR_ACT, size_t(B) * 4 * 2));
    if (!Bv || !R || !Q || !Hd || !Yd || !nrm || !act) return CCAB_ENOMEM;
    const int64_t sq = n * qslots;
    c128* Qv = Q + (defl ? int64_t(m) * n : 0);  // the Krylov basis
    // deflation state: U (m vectors per system); per system Qh ((m+1) x m),
    // R^-1 (m x m), B (m x m), c0 and z (m each)
    c128 *Ub = nullptr, *rcQh = nullptr, *rcRi = nullptr, *rcB = nullptr, *rcC0 = nullptr,
         *rcZ = nullptr, *rcCq = nullptr, *rcUq = nullptr;
    if (defl) {
        Ub = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_U, nb * size_t(m) * 16));
        const size_t per = 2 * size_t(m + 1) * m + 3 * size_t(m) * m + 2 * size_t(m);
        rcQh = static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_RC, size_t(B) * per * 16));
        if (!Ub || !rcQh) return CCAB_ENOMEM;
        rcRi = rcQh + size_t(B) * (m + 1) * m;
        rcB = rcRi + size_t(B) * m * m;
        rcC0 = rcB + size_t(B) * m * m;
        rcZ = rcC0 + size_t(B) * m;
        rcCq = rcZ + size_t(B) * m;
        rcUq = rcCq + size_t(B) * (m + 1) * m;
    }
    // harmonic Ritz deflation directions (gir_rc_ritz_kernel)
    const bool ritz = defl && ctx->gir_ritz && m <= kRitzMaxK && ctx->gir_recycle <= kRitzVecs;
    int s_rc = 0;               // deflation vectors in use (0: none yet)
    std::vector<int> k1(B, 0);  // steps of each system's first cycle
    // device Givens state (see gir_givens_kernel)
    const size_t g_h = ws_align(size_t(B) * (m + 1) * m * 16), g_cs = ws_align(size_t(B) * m * 16);
    const size_t g_be = ws_align(size_t(B) * (m + 1) * 16), g_i = ws_align(size_t(B) * 4);
    const size_t g_all = g_h + 2 * g_cs + g_be + g_i + ws_align(size_t(B) * 8);
    // with the FP32 remainder the residual follows r <- r - Q (Hbar y) - A_l d
    // (A_h d = Q_(k+1) Hbar y by the Arnoldi relation of the FP16 operator), so
    // each refinement residual reads the 8-byte remainder instead of 12 bytes
    const bool recur = Alo != nullptr;
    char* gv = static_cast<char*>(ctx_buf(ctx, SLOT_GIR_GIV, g_all + (recur ? g_h + g_be : 0)));
    if (!gv) return CCAB_ENOMEM;
    c128* Hbar = recur ? reinterpret_cast<c128*>(gv + g_all) : nullptr;
    c128* Zd = recur ? reinterpret_cast<c128*>(gv + g_all + g_h) : nullptr;
    c128* Dv = recur ? static_cast<c128*>(ctx_buf(ctx, SLOT_GIR_D, nb * 16)) : nullptr;
    if (recur && !Dv) return CCAB_ENOMEM;
    c128* Hg = reinterpret_cast<c128*>(gv);
    c128* csd = reinterpret_cast<c128*>(gv + g_h);
    c128* snd = reinterpret_cast<c128*>(gv + g_h + g_cs);
    c128* bed = reinterpret_cast<c128*>(gv + g_h + 2 * g_cs);
    int32_t* kdone_d = reinterpret_cast<int32_t*>(gv + g_h + 2 * g_cs + g_be);
    double* bn_d = reinterpret_cast<double*>(gv + g_h + 2 * g_cs + g_be + g_i);
    // pinned: the synchronous active-list staging, norms, and per Arnoldi
    // step one active-list slot and one done-flag slot (asynchronous copies)
    const size_t pin_a = ws_align(size_t(B) * 4), pin_n = ws_align(size_t(B) * 8);
    CCAB_TRY(pinned_reserve(ctx, pin_a + pin_n + 2 * size_t(m + 1) * pin_a));
    char* pin = static_cast<char*>(ctx->pinned);
    int32_t* hAct = reinterpret_cast<int32_t*>(pin);
    double* hN = reinterpret_cast<double*>(pin + pin_a);
    auto ring_act = [&](int k) { return reinterpret_cast<int32_t*>(pin + pin_a + pin_n + size_t(k) * pin_a); };
    auto ring_done = [&](int k) {
        return reinterpret_cast<int32_t*>(pin + pin_a + pin_n + size_t(m + 1 + k) * pin_a);
    };
    // the Givens kernel writes the step's done flags straight into the pinned
    // slot (device address of the same memory)
    char* pin_dev = nullptr;
    CCAB_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&pin_dev), ctx->pinned, 0));
    auto ring_done_dev = [&](int k) {
        return reinterpret_cast<int32_t*>(pin_dev + pin_a + pin_n + size_t(m + 1 + k) * pin_a);
    };
    // an event per Arnoldi step marks its done-flag copy
    while ((int)ctx->gir_events.size() < m + 1) {
        hipEvent_t e;
        CCAB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->gir_events.push_back(e);
    }
    // the host learns which systems finished kLag steps late (it waits for
    // step k - kLag after enqueuing step k, so the GPU always has work)
    const int kLag = ctx->gir_lag;
    const c128 one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
    // residual matrix: c128; or the FP16 operator + FP32 remainder; or the
    // c64 operator itself
    const void* Ares = Alo ? Aop : A128 ? static_cast<const void*>(A128) : Aop;
    const int pres = Alo ? CCAB_C32 : A128 ? CCAB_C128 : CCAB_C64;

    hipLaunchKernelGGL(copy_c128_kernel, grid1(nb), dim3(256), 0, s, nb, X, Bv);
    if (X0) hipLaunchKernelGGL(copy_c128_kernel, grid1(nb), dim3(256), 0, s, nb, X0, X);
    else hipLaunchKernelGGL(fill_zero_kernel, grid1(nb), dim3(256), 0, s, nb, X);

    std::vector<int32_t> active(B);
    for (int64_t b = 0; b < B; ++b) active[b] = int32_t(b);
    // the device copy of the active list changes only when a system
    // finishes: skip the upload (and its synchronisation) otherwise
    std::vector<int32_t> uploaded;
    auto upload_active = [&](const std::vector<int32_t>& a, int32_t* dst) -> int {
        if (a == uploaded) return CCAB_OK;
        std::memcpy(hAct, a.data(), a.size() * 4);
        CCAB_HIP(hipMemcpyAsync(dst, hAct, a.size() * 4, hipMemcpyHostToDevice, s));
        CCAB_HIP(hipStreamSynchronize(s));   // hAct is reused right away
        uploaded = a;
        return CCAB_OK;
    };
    auto read_norms = [&](const c128* V, int64_t stride, const std::vector<int32_t>& a,
                          std::vector<double>& out) -> int {
        CCAB_TRY(nrm2_active(ctx, n, (int64_t)a.size(), act, V, stride, nrm));
        CCAB_HIP(hipMemcpyAsync(hN, nrm, size_t(B) * 8, hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        for (int32_t b : a) out[b] = hN[b];
        return CCAB_OK;
    };
    // matrix entries a product over a set of systems reads (profiler work)
    auto set_elems = [&](const std::vector<int32_t>& a) {
        if (!rg) return;
        double e = 0.0;
        for (int32_t b : a) e += double(rg->hn[b]) * double(rg->hn[b]);
        rg->elems = e;
    };
    // one Arnoldi step over the first ni systems of the device's active list
    auto arnoldi_step = [&](int k, int ni) -> int {
        CCAB_TRY(gemv_active(ctx, op_prec, n, n, ni, act, Aop, lda, sA, Qv + int64_t(k) * n, sq, one,
                             zero, Qv + int64_t(k + 1) * n, sq, d_inv_scale, rg, nullptr, kdone_d));
        {
            // (a deflated cycle orthogonalises against C, the s_rc slots below V, too)
            ProfScope ps(ctx, PROF_MGS, double(k + 2 + s_rc) * 2.0 * double(n) * 16.0 * ni);
            CCAB_TRY(launch_mgs(ctx, ref_mgs, n, k + s_rc, Qv - int64_t(s_rc) * n, n, sq, Hd, sh,
                                act, ni, kdone_d, false, ctx->gs_block));
        }
        hipLaunchKernelGGL(gir_givens_kernel, dim3((ni + 63) / 64), dim3(64), 0, s, k, m, act, ni,
                           Hd, sh, Hg, csd, snd, bed, kdone_d, nrm, bn_d, inner_tol, tol,
                           Hbar, ring_done_dev(k), s_rc, s_rc ? rcB : nullptr);
        CCAB_HIP(hipGetLastError());
        return CCAB_OK;
    };
    // y = R^{-1} beta and x += Q y over each system's own steps
    auto finish_cycle = [&](int na_c) -> int {
        if (m <= kSolveYLdsM)
            hipLaunchKernelGGL(gir_solve_y_wave_kernel, dim3((unsigned)na_c), dim3(64), 0, s, m, act,
                               Hg, bed, kdone_d, Yd, Hbar, Zd);
        else
            hipLaunchKernelGGL(gir_solve_y_kernel, dim3((unsigned)((na_c + 63) / 64)), dim3(64), 0, s,
                               m, act, na_c, Hg, bed, kdone_d, Yd, Hbar, Zd);
        if (s_rc) {
            hipLaunchKernelGGL(gir_rc_z_kernel, dim3((unsigned)((na_c + 63) / 64)), dim3(64), 0, s, m,
                               s_rc, act, na_c, kdone_d, rcB, Yd, rcC0, rcZ);
            hipLaunchKernelGGL(gir_update_x_rc_kernel,
                               dim3((unsigned)((n + 255) / 256), (unsigned)na_c), dim3(256), 0, s, n,
                               m, s_rc, act, kdone_d, Q, sq, Yd, Ub, n * int64_t(m), rcZ, rcC0, X,
                               Zd, Dv, R);
        } else {
            hipLaunchKernelGGL(gir_update_x_kernel,
                               dim3((unsigned)((n + 255) / 256), (unsigned)na_c), dim3(256), 0, s, n,
                               m, act, kdone_d, Qv, sq, Yd, X, Zd, Dv, R);
        }
        CCAB_HIP(hipGetLastError());
        return CCAB_OK;
    };
    auto run_cycle_graph = [&](int na_c) -> int {
        const GemvTune* tn = gemv_tune_of(ctx, op_prec, n);
        GirKey key{n, B, na_c, m, op_prec, ref_mgs ? 1 : 0, ctx->mgs_coop ? 1 : 0,
                   tn ? tn->wg_target : ctx->gemv_wg_target, tn ? tn->nt : (ctx->gemv_nt ? 1 : 0),
                   tn ? 1 : 0, Aop, d_inv_scale, act, Q, Hd, gv, Yd, X, Zd, Dv, R, nrm, bn_d,
                   ctx->pinned, rg ? rg->n : nullptr, rg ? rg->aoff : nullptr, lda, sA,
                   inner_tol, tol};
        hipGraphExec_t exec = nullptr;
        for (auto& g : ctx->gir_graphs)
            if (g.first == key) exec = g.second;
        if (!exec) {
            // grow every workspace the captured launches use before capturing
            // (a growing buffer would drop the graphs and allocate mid-capture)
            CCAB_TRY(gemv_active(ctx, op_prec, n, n, na_c, act, Aop, lda, sA, Qv, sq, one, zero,
                                 Qv + n, sq, d_inv_scale, rg, nullptr, kdone_d, -1, true));
            CCAB_TRY(launch_mgs(ctx, ref_mgs, n, 0, Qv, n, sq, Hd, sh, act, na_c, kdone_d, true,
                                ctx->gs_block));
            // (the active list and every buffer key.* names stay valid for the graph)
            CCAB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = CCAB_OK;
            hipLaunchKernelGGL(scale_inv_norm_kernel, dim3((unsigned)((n + 255) / 256),
                               (unsigned)na_c), dim3(256), 0, s, n, R, n, nrm, Qv, sq, act);
            hipLaunchKernelGGL(gir_cycle_init_kernel, dim3((unsigned)((na_c + 63) / 64)), dim3(64), 0,
                               s, m, act, na_c, kdone_d, bed, nrm);
            for (int k = 0; k < m && rc == CCAB_OK; ++k) rc = arnoldi_step(k, na_c);
            if (rc == CCAB_OK) rc = finish_cycle(na_c);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &graph);
            if (rc != CCAB_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            CCAB_HIP(ec);
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            CCAB_HIP(ei);
            if (ctx->gir_graphs.size() >= 16) {
                (void)hipGraphExecDestroy(ctx->gir_graphs.front().second);
                ctx->gir_graphs.erase(ctx->gir_graphs.begin());
            }
            ctx->gir_graphs.emplace_back(key, exec);
        }
        CCAB_HIP(hipGraphLaunch(exec, s));
        return CCAB_OK;
    };
    std::vector<double> bn(B), rn(B), rel(B, 1.0);
    std::vector<int64_t> iters(B, 0), nmv(B, 0), cycles(B, 0);
    std::vector<char> conv(B, 0);
    CCAB_TRY(upload_active(active, act));
    CCAB_TRY(read_norms(Bv, n, active, bn));
    CCAB_HIP(hipMemcpy(bn_d, bn.data(), size_t(B) * 8, hipMemcpyHostToDevice));
    for (int cyc = 0; cyc <= max_cycles; ++cyc) {
        // FP64 residual of the active systems
        CCAB_TRY(upload_active(active, act));
        const unsigned na = (unsigned)active.size();
        if (recur && cyc > 0) {
            // R = r_old - Q z already; subtract the remainder's share A_l d
            set_elems(active);
            CCAB_TRY(gemv_active(ctx, CCAB_C64, n, n, na, act, Alo, lda, sA, Dv, n, mone, one, R, n,
                                 nullptr, rg, nullptr, nullptr, PROF_GEMV128));
            for (int32_t b : active) nmv[b]++;
        } else {
            hipLaunchKernelGGL(copy_active_kernel, dim3((unsigned)((n + 255) / 256), na), dim3(256),
                               0, s, n, Bv, R, act);
            if (cyc == 0 && AX0) {
                hipLaunchKernelGGL(sub_active_kernel, dim3((unsigned)((n + 255) / 256), na),
                                   dim3(256), 0, s, n, AX0, R, act);
            } else if (cyc > 0 || X0) {
                set_elems(active);
                // (a warm start's first residual with the FP32 remainder reads
                // operator + remainder: leaving A_l x0 to a later cycle costs
                // that cycle more steps than it saves, measured)
                CCAB_TRY(gemv_active(ctx, pres, n, n, na, act, Ares, lda, sA, X, n, mone, one, R, n,
                                     Alo ? d_inv_scale : nullptr, rg, Alo));
                for (int32_t b : active) nmv[b]++;
            }
        }
        if (cyc == 0 && X0)
            hipLaunchKernelGGL(warm_scale_kernel, dim3(na), dim3(256), 0, s, n, Bv, R, X, act);
        CCAB_TRY(read_norms(R, n, active, rn));
        std::vector<int32_t> still;
        for (int32_t b : active) {
            rel[b] = bn[b] > 0 ? rn[b] / bn[b] : 0.0;
            if (!std::isfinite(rel[b])) {
                CCAB_TRY(coop_abort_check(ctx));
                return ctx_fail(ctx, CCAB_ENAN, "GMRES-IR residual is NaN in system %d", b);
            }
            if (rel[b] <= tol || rn[b] == 0.0) conv[b] = 1;
            else still.push_back(b);
        }
        active.swap(still);
        if (active.empty() || cyc == max_cycles) break;
        CCAB_TRY(upload_active(active, act));
        // one GMRES(m) cycle for A64 d = r on the active systems
        const int na_c = (int)active.size();
        for (int32_t b : active) cycles[b]++;
        std::vector<int> launched(B, 0);  // Arnoldi steps enqueued per system (profiler)
        if (graph_cycles) {
            // the whole cycle (m Arnoldi steps over the cycle's systems, the
            // back substitution and the update) as one captured graph: the
            // device's done flags make finished systems skip the later steps,
            // so the launch sequence does not depend on the data
            CCAB_TRY(run_cycle_graph(na_c));
        } else {
            if (defl && cyc > 0 && s_rc == 0) {
                // the second cycle: C and U from the first cycle's basis and
                // rotations (still in Qv, Hg, csd, snd; kdone_d holds its steps)
                // (the first directions of A K_k: 8 of them deflate about as
                // much as all k, measured on rigid-sphere sweeps, at one
                // Gram-Schmidt block per step)
                for (int32_t b : active) s_rc = std::max(s_rc, std::min(k1[b], ctx->gir_recycle));
                hipLaunchKernelGGL(gir_rc_factors_kernel, dim3((unsigned)na_c), dim3(64), 0, s, m,
                                   act, kdone_d, Hg, csd, snd, rcQh, rcRi);
                if (ritz)
                    hipLaunchKernelGGL(gir_rc_ritz_kernel, dim3((unsigned)na_c), dim3(64), 0, s, m,
                                       ctx->gir_recycle, act, kdone_d, Hg, rcQh, rcRi, rcCq, rcUq);
                hipLaunchKernelGGL(gir_rc_basis_kernel,
                                   dim3((unsigned)((n + kRcRows - 1) / kRcRows), (unsigned)na_c),
                                   dim3(64), 0, s, n, m, s_rc, act, kdone_d, Q, sq,
                                   ritz ? rcCq : rcQh, ritz ? rcUq : rcRi, Ub, n * int64_t(m));
                CCAB_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(scale_inv_norm_kernel, dim3((unsigned)((n + 255) / 256),
                               (unsigned)na_c), dim3(256), 0, s, n, R, n, nrm, Qv, sq, act);
            if (s_rc) {
                // r / |r| against C: c0 = C^H r and the cycle's first vector
                CCAB_TRY(launch_mgs(ctx, ref_mgs, n, s_rc - 1, Qv - int64_t(s_rc) * n, n, sq, Hd, sh,
                                    act, na_c, nullptr, false, ctx->gs_block));
                hipLaunchKernelGGL(gir_rc_start_kernel, dim3((unsigned)((na_c + 63) / 64)), dim3(64),
                                   0, s, m, s_rc, act, na_c, kdone_d, bed, nrm, Hd, sh, rcC0);
            } else {
                hipLaunchKernelGGL(gir_cycle_init_kernel, dim3((unsigned)((na_c + 63) / 64)),
                                   dim3(64), 0, s, m, act, na_c, kdone_d, bed, nrm);
            }
            std::vector<int32_t> inner = active;
            int known = -1;  // the latest step whose done flags the host has applied
            for (int k = 0; k < m && !inner.empty(); ++k) {
                if (inner != uploaded) {  // a fresh pinned slot per step: no host wait
                    std::memcpy(ring_act(k), inner.data(), inner.size() * 4);
                    CCAB_HIP(hipMemcpyAsync(act, ring_act(k), inner.size() * 4,
                                            hipMemcpyHostToDevice, s));
                    uploaded = inner;
                }
                const unsigned ni = (unsigned)inner.size();
                set_elems(inner);
                for (int32_t b : inner) launched[b]++;
                // systems already done on the device skip the product, the
                // Gram-Schmidt step and the rotation
                CCAB_TRY(arnoldi_step(k, (int)ni));
                CCAB_HIP(hipEventRecord(ctx->gir_events[k], s));
                // newest finished flags: wait for step k - kLag, take a later one if ready
                int target = -1;
                if (k - kLag > known) {
                    CCAB_HIP(hipEventSynchronize(ctx->gir_events[k - kLag]));
                    target = k - kLag;
                }
                for (int t = k; t > std::max(target, known); --t) {
                    const hipError_t q = hipEventQuery(ctx->gir_events[t]);
                    if (q == hipSuccess) {
                        target = t;
                        break;
                    }
                    if (q != hipErrorNotReady) CCAB_HIP(q);
                }
                if (target > known) {
                    known = target;
                    const int32_t* fl = ring_done(target);
                    std::vector<int32_t> next;
                    for (int32_t b : inner)
                        if (fl[b] == 0) next.push_back(b);
                    inner.swap(next);
                }
            }
            // every system of the cycle is done on the device (step m - 1 ends
            // the cycle): y = R^{-1} beta, then x += Q y
            CCAB_TRY(upload_active(active, act));
            CCAB_TRY(finish_cycle(na_c));
        }
        CCAB_HIP(hipMemcpyAsync(hAct, kdone_d, size_t(B) * 4, hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        uploaded.clear();  // hAct now holds flags, not the list
        double gemv_over = 0.0, mgs_over = 0.0;
        const double esz = op_prec == CCAB_C32 ? 4.0 : 8.0;
        for (int32_t b : active) {
            if (hAct[b] <= 0 || hAct[b] > m)
                return ctx_fail(ctx, CCAB_EHIP, "GMRES-IR cycle of system %d did not finish", b);
            if (cyc == 0) k1[b] = hAct[b];
            iters[b] += hAct[b];
            nmv[b] += hAct[b];
            // steps enqueued after the system was done did no work
            const double nb = rg ? double(rg->hn[b]) : double(n);
            for (int k = hAct[b]; k < launched[b]; ++k) {
                gemv_over += nb * nb * esz;
                mgs_over += double(k + 2 + s_rc) * 2.0 * double(n) * 16.0;
            }
        }
        prof_adjust(ctx, PROF_GEMV, -gemv_over);
        prof_adjust(ctx, PROF_MGS, -mgs_over);
    }
    if (st) {
        for (int64_t b = 0; b < B; ++b) {
            st[b] = ccab_solve_stats{};
            st[b].num_mv = nmv[b];
            st[b].iterations = iters[b];
            st[b].restarts = cycles[b];
            st[b].ir_steps = cycles[b];
            st[b].rel_residual = rel[b];
            st[b].backward_error = -1.0;
            st[b].converged = conv[b];
        }
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}
