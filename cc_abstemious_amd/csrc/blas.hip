// Level-1/2 kernels of the solvers (FP64 accumulation throughout):
// batched GEMV over a c128 or c64 column-major matrix (the GMRES matvec of
// gmres.rs:64-69 and the FP64 refinement residual), vector norms, the
// Frobenius norm of gmres.rs:75-79, precision conversions and the
// modified Gram-Schmidt step of gmres.rs:158-176.
//
// The GEMV is HBM-bound: n*n*sizeof(entry) bytes per product.  A workgroup
// of 256 lanes owns 256 rows; lanes walk a chunk of columns whose x[j] is
// wave-uniform.  Column chunks are spread over grid.y so that even a single
// system fills the 256 CUs; partial sums are combined in fixed order.
#include <cstdint>

#include "ccab_internal.h"

namespace {

template <typename TA>
__device__ __forceinline__ c128 ld_c(const TA* p, int64_t off);
template <>
__device__ __forceinline__ c128 ld_c<c128>(const c128* p, int64_t off) {
    return p[off];
}
template <>
__device__ __forceinline__ c128 ld_c<c64>(const c64* p, int64_t off) {
    const c64 v = p[off];
    return c128{double(v.re), double(v.im)};
}
template <>
__device__ __forceinline__ c128 ld_c<c32>(const c32* p, int64_t off) {
    const c32 v = p[off];
    return c128{double(__half2float(__ushort_as_half(v.re))),
                double(__half2float(__ushort_as_half(v.im)))};
}

// Ragged batches (rn != nullptr): system b is rn[b] x rn[b] at A + raoff[b]
// with leading dimension rn[b]; rows beyond rn[b] get zero partial sums.
// acc += a * x (complex) as four FMAs (the plain expression compiles to a
// multiply, an FMA and an add per component)
__device__ __forceinline__ void cmac(double& re, double& im, double ar, double ai, const c128& x) {
    re = fma(ar, x.re, re);
    re = fma(-ai, x.im, re);
    im = fma(ar, x.im, im);
    im = fma(ai, x.re, im);
}

template <bool NT, typename TA>
__device__ __forceinline__ c128 ld_m(const TA* p, int64_t off) {
    if constexpr (NT && sizeof(TA) == 16) {
        typedef double f64x2_t __attribute__((ext_vector_type(2)));
        const f64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f64x2_t*>(p + off));
        return c128{v.x, v.y};
    } else {
        return ld_c(p, off);
    }
}

template <typename TA, bool NT = false>
__global__ __launch_bounds__(256) void gemv_part_kernel(int64_t M, int64_t N, const TA* __restrict__ A,
                                                        int64_t lda, int64_t sA,
                                                        const c128* __restrict__ x, int64_t sx,
                                                        int64_t chunk, c128* __restrict__ part,
                                                        int64_t nsplit,
                                                        const int32_t* __restrict__ active,
                                                        const int64_t* __restrict__ rn,
                                                        const int64_t* __restrict__ raoff,
                                                        const int32_t* __restrict__ skip) {
    const int64_t z = blockIdx.z;
    const int64_t b = active ? active[z] : z;
    if (skip && skip[b]) return;
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t nb = rn ? rn[b] : N;
    const int64_t j0 = int64_t(blockIdx.y) * chunk;
    const int64_t j1 = (rn && i >= nb) ? j0 : min(j0 + chunk, nb);
    // ragged: rows past the body and chunks past its columns write nothing
    // (the finish reads only the body's own rows and chunks)
    if (i >= M || (rn && (i >= nb || j0 >= nb))) return;
    if (rn) lda = nb;
    const TA* a = (rn ? A + raoff[b] : A + b * sA) + i;
    const c128* xb = x + b * sx;
    double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
    int64_t j = j0;
    for (; j + 3 < j1; j += 4) {
        c128 av[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            av[u] = ld_m<NT>(a, (j + u) * lda);
            xv[u] = xb[j + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cmac(re[u], im[u], av[u].re, av[u].im, xv[u]);
        }
    }
    for (; j < j1; ++j) {
        const c128 a0 = ld_m<NT>(a, j * lda);
        const c128 x0 = xb[j];
        cmac(re[0], im[0], a0.re, a0.im, x0);
    }
    part[(z * nsplit + blockIdx.y) * M + i] =
        c128{(re[0] + re[1]) + (re[2] + re[3]), (im[0] + im[1]) + (im[2] + im[3])};
}

// R consecutive rows per lane with one 16-byte load per column (R = 2 for
// c64, 4 for c32); needs lda, sA and the row start multiples of R
template <typename TA, int R>
struct Vec16;
template <>
struct Vec16<c64, 2> {
    using type = float4;
};
template <>
struct Vec16<c32, 4> {
    using type = uint4;
};

template <int R>
__device__ __forceinline__ void unpack16(const float4& v, c128* a) {
    a[0] = c128{double(v.x), double(v.y)};
    a[1] = c128{double(v.z), double(v.w)};
}
__device__ __forceinline__ double hf(uint32_t bits16) {
    return double(__half2float(__ushort_as_half(uint16_t(bits16))));
}
template <int R>
__device__ __forceinline__ void unpack16(const uint4& v, c128* a) {
    a[0] = c128{hf(v.x & 0xffffu), hf(v.x >> 16)};
    a[1] = c128{hf(v.y & 0xffffu), hf(v.y >> 16)};
    a[2] = c128{hf(v.z & 0xffffu), hf(v.z >> 16)};
    a[3] = c128{hf(v.w & 0xffffu), hf(v.w >> 16)};
}

// one 16-byte matrix load; NT: non-temporal (the matrix streams through once
// per product, x and the partial sums are the data worth caching)
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
template <bool NT, typename V>
__device__ __forceinline__ V ld16(const void* p) {
    if constexpr (NT) {
        const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
        V out;
        __builtin_memcpy(&out, &v, 16);
        return out;
    } else {
        return *reinterpret_cast<const V*>(p);
    }
}

// columns whose loads a lane has in flight at once (64-system n = 5120 sweep:
// 4 -> 38.5 ms of matvecs, 8 -> 37.9 ms, 16 -> 39.1 ms)
constexpr int kVecCols = 8;

template <typename TA, int R, bool NT = false>
__global__ __launch_bounds__(256) void gemv_part_vec_kernel(int64_t M, int64_t N, const TA* __restrict__ A,
                                                            int64_t lda, int64_t sA,
                                                            const c128* __restrict__ x, int64_t sx,
                                                            int64_t chunk, c128* __restrict__ part,
                                                            int64_t nsplit,
                                                            const int32_t* __restrict__ active,
                                                            const int64_t* __restrict__ rn,
                                                            const int64_t* __restrict__ raoff,
                                                            const int32_t* __restrict__ skip) {
    using V = typename Vec16<TA, R>::type;
    const int64_t z = blockIdx.z;
    const int64_t b = active ? active[z] : z;
    if (skip && skip[b]) return;
    const int64_t i0 = (int64_t(blockIdx.x) * 256 + threadIdx.x) * R;
    const int64_t nb = rn ? rn[b] : N;     // columns (and, ragged, rows) of this system
    const int64_t mb = rn ? nb : M;
    const int64_t j0 = int64_t(blockIdx.y) * chunk;
    const int64_t j1 = i0 >= mb ? j0 : min(j0 + chunk, nb);
    if (i0 >= M || (rn && (i0 >= mb || j0 >= nb))) return;  // (as above)
    if (rn) lda = nb;
    const TA* a = (rn ? A + raoff[b] : A + b * sA) + i0;
    const c128* xb = x + b * sx;
    double re[R], im[R];
#pragma unroll
    for (int r = 0; r < R; ++r) re[r] = im[r] = 0.0;
    if (i0 + R <= mb) {
        int64_t j = j0;
        for (; j + kVecCols - 1 < j1; j += kVecCols) {
            V v[kVecCols];
#pragma unroll
            for (int u = 0; u < kVecCols; ++u) v[u] = ld16<NT, V>(a + (j + u) * lda);
#pragma unroll
            for (int u = 0; u < kVecCols; ++u) {
                c128 av[R];
                unpack16<R>(v[u], av);
                const c128 xv = xb[j + u];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cmac(re[r], im[r], av[r].re, av[r].im, xv);
                }
            }
        }
        for (; j < j1; ++j) {
            c128 av[R];
            unpack16<R>(ld16<NT, V>(a + j * lda), av);
            const c128 xv = xb[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cmac(re[r], im[r], av[r].re, av[r].im, xv);
            }
        }
    } else {  // last rows of the matrix
        for (int64_t j = j0; j < j1; ++j) {
            const c128 xv = xb[j];
            for (int r = 0; r < R && i0 + r < mb; ++r) {
                const c128 av = ld_c(a + r, j * lda);
                cmac(re[r], im[r], av.re, av.im, xv);
            }
        }
    }
    c128* o = part + (z * nsplit + blockIdx.y) * M + i0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (i0 + r < M) o[r] = c128{re[r], im[r]};
}

// Product with a matrix held as a scaled complex-FP16 copy H plus a
// complex-FP32 remainder L: entry = H * inv_scale[b] + L (FP64).  VEC: four
// consecutive rows per lane, one 16-byte load of H and two of L per column
// (needs lda, sA, the ragged alignment and the row start multiples of 4);
// otherwise one row per lane.  Same partial-sum layout as the kernels above.
template <bool VEC>
__global__ __launch_bounds__(256) void gemv_split_kernel(int64_t M, int64_t N,
                                                         const c32* __restrict__ Ah,
                                                         const c64* __restrict__ Al,
                                                         const double* __restrict__ inv_scale,
                                                         int64_t lda, int64_t sA,
                                                         const c128* __restrict__ x, int64_t sx,
                                                         int64_t chunk, c128* __restrict__ part,
                                                         int64_t nsplit,
                                                         const int32_t* __restrict__ active,
                                                         const int64_t* __restrict__ rn,
                                                         const int64_t* __restrict__ raoff,
                                                         const int32_t* __restrict__ skip) {
    constexpr int R = VEC ? 4 : 1;
    const int64_t z = blockIdx.z;
    const int64_t b = active ? active[z] : z;
    if (skip && skip[b]) return;
    const int64_t i0 = (int64_t(blockIdx.x) * 256 + threadIdx.x) * R;
    const int64_t nb = rn ? rn[b] : N;
    const int64_t mb = rn ? nb : M;
    const int64_t j0 = int64_t(blockIdx.y) * chunk;
    const int64_t j1 = i0 >= mb ? j0 : min(j0 + chunk, nb);
    if (i0 >= M || (rn && (i0 >= mb || j0 >= nb))) return;  // (as above)
    if (rn) lda = nb;
    const int64_t base = (rn ? raoff[b] : b * sA) + i0;
    const c32* ah = Ah + base;
    const c64* al = Al + base;
    const double isc = inv_scale[b];
    const c128* xb = x + b * sx;
    double re[R], im[R];
#pragma unroll
    for (int r = 0; r < R; ++r) re[r] = im[r] = 0.0;
    if (VEC && i0 + R <= mb) {
        auto column = [&](const uint4& vh, const float4& v0, const float4& v1, const c128& xv) {
            c128 hv[4];
            unpack16<4>(vh, hv);
            const double lr[4] = {double(v0.x), double(v0.z), double(v1.x), double(v1.z)};
            const double li[4] = {double(v0.y), double(v0.w), double(v1.y), double(v1.w)};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double er = fma(hv[r].re, isc, lr[r]);
                const double ei = fma(hv[r].im, isc, li[r]);
                cmac(re[r], im[r], er, ei, xv);
            }
        };
        int64_t j = j0;
        for (; j + 1 < j1; j += 2) {
            uint4 vh[2];
            float4 v0[2], v1[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                vh[u] = *reinterpret_cast<const uint4*>(ah + (j + u) * lda);
                v0[u] = *reinterpret_cast<const float4*>(al + (j + u) * lda);
                v1[u] = *reinterpret_cast<const float4*>(al + (j + u) * lda + 2);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) column(vh[u], v0[u], v1[u], xb[j + u]);
        }
        if (j < j1)
            column(*reinterpret_cast<const uint4*>(ah + j * lda),
                   *reinterpret_cast<const float4*>(al + j * lda),
                   *reinterpret_cast<const float4*>(al + j * lda + 2), xb[j]);
    } else {
        for (int64_t j = j0; j < j1; ++j) {
            const c128 xv = xb[j];
            for (int r = 0; r < R && i0 + r < mb; ++r) {
                const c128 h = ld_c(ah + r, j * lda);
                const c128 l = ld_c(al + r, j * lda);
                const double er = fma(h.re, isc, l.re);
                const double ei = fma(h.im, isc, l.im);
                cmac(re[r], im[r], er, ei, xv);
            }
        }
    }
    c128* o = part + (z * nsplit + blockIdx.y) * M + i0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (i0 + r < M) o[r] = c128{re[r], im[r]};
}

__global__ void gemv_finish_kernel(int64_t M, const c128* __restrict__ part, int64_t nsplit,
                                   c128 alpha, c128 beta, c128* y, int64_t sy,
                                   const int32_t* __restrict__ active,
                                   const double* __restrict__ inv_scale,
                                   const int32_t* __restrict__ skip,
                                   const int64_t* __restrict__ rn, int64_t chunk) {
    const int64_t z = blockIdx.y;
    const int64_t b = active ? active[z] : z;
    if (skip && skip[b]) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M) return;
    // ragged: the body's chunks that hold columns; rows past the body are 0
    int64_t np = nsplit;
    if (rn) np = i < rn[b] ? min(nsplit, (rn[b] + chunk - 1) / chunk) : 0;
    c128 s{0.0, 0.0};
    for (int64_t p = 0; p < np; ++p) s += part[(z * nsplit + p) * M + i];
    if (inv_scale) s = s * inv_scale[b];
    c128* yy = y + b * sy + i;
    const bool zero_beta = beta.re == 0.0 && beta.im == 0.0;
    *yy = alpha * s + (zero_beta ? c128{0.0, 0.0} : beta * *yy);
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < BS / 64; ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(256) void nrm2_kernel(int64_t n, const c128* __restrict__ v,
                                                   int64_t stride, double* out,
                                                   const int32_t* __restrict__ active) {
    __shared__ double red[4];
    const int64_t sys = active ? active[blockIdx.x] : blockIdx.x;
    const c128* p = v + sys * stride;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += p[i].re * p[i].re + p[i].im * p[i].im;
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[sys] = sqrt(s);
}

template <typename TA>
__global__ __launch_bounds__(256) void sumsq_kernel(int64_t n, const TA* __restrict__ A,
                                                    int64_t lda, double* part) {
    __shared__ double red[4];
    // one column per block row
    const int64_t j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const c128 z = ld_c(A, i + j * lda);
        s += z.re * z.re + z.im * z.im;
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) part[j] = s;
}

__global__ __launch_bounds__(256) void sum_sqrt_kernel(int64_t n, const double* part, double* out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) *out = sqrt(s);
}

__global__ void c128_to_c64_kernel(int64_t count, const c128* __restrict__ s, c64* __restrict__ d) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) d[i] = c64{float(s[i].re), float(s[i].im)};
}
__global__ void c64_to_c128_kernel(int64_t count, const c64* __restrict__ s, c128* __restrict__ d) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) d[i] = c128{double(s[i].re), double(s[i].im)};
}

}  // namespace

size_t gemv_workspace_bytes(int64_t n, int64_t batch) { return 0; }

int gemv_active(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t nact,
                const int32_t* d_active, const void* d_A, int64_t lda, int64_t strideA,
                const c128* d_x, int64_t stridex, c128 alpha, c128 beta, c128* d_y,
                int64_t stridey, const double* d_inv_scale, const Ragged* rg, const c64* d_Alo,
                const int32_t* d_skip, int prof_cls, bool reserve_only) {
    if (M <= 0 || nact <= 0) return CCAB_OK;
    if (prec == CCAB_C32 && !d_inv_scale) return ctx_fail(ctx, CCAB_EINVAL, "c32 GEMV needs scales");
    if (d_Alo && prec != CCAB_C32)
        return ctx_fail(ctx, CCAB_EINVAL, "the FP32 remainder goes with the c32 matrix");
    if (rg && (!rg->n || !rg->aoff || M != N))
        return ctx_fail(ctx, CCAB_EINVAL, "ragged GEMV needs sizes, offsets and square systems");
    const int64_t* rn = rg ? rg->n : nullptr;
    const int64_t* raoff = rg ? rg->aoff : nullptr;
    // rows per lane: one 16-byte load per column where the layout allows it
    int R = prec == CCAB_C32 ? 4 : prec == CCAB_C64 ? 2 : 1;
    if ((rg ? rg->align % R != 0 : (lda % R != 0 || strideA % R != 0)) ||
        (reinterpret_cast<uintptr_t>(d_A) % 16) != 0 ||
        (d_Alo && reinterpret_cast<uintptr_t>(d_Alo) % 16 != 0))
        R = 1;
    const int64_t rowblocks = (M + 256 * R - 1) / (256 * R);
    const int64_t finish_blocks = (M + 255) / 256;
    // ragged batches: finer column chunks balance bodies of uneven size
    // (measured: 4x the uniform target, 4.0 -> 4.7 TB/s on 256 bodies of
    // 1280-5120 unknowns)
    // uniform products of a tuned size (ccab_autotune_sweep) take its
    // workgroup target and load kind
    const GemvTune* tn = (!rg && !d_Alo) ? gemv_tune_of(ctx, prec, N) : nullptr;
    const int64_t target = tn ? tn->wg_target : ctx->gemv_wg_target * (rg ? 4 : 1);
    int64_t nsplit = (target + rowblocks * nact - 1) / (rowblocks * nact);
    nsplit = std::max<int64_t>(1, std::min<int64_t>(nsplit, (N + 63) / 64));
    int64_t chunk = N > 0 ? (N + nsplit - 1) / nsplit : 1;
    // ragged: a fixed column chunk, so that a large body's workgroups stream
    // no more than a small body's and the launch has no long tail
    if (rg && ctx->gemv_ragged_chunk > 0 && N > 0) {
        chunk = std::min<int64_t>(N, ctx->gemv_ragged_chunk);
        // the iteration tail (few bodies left): halve the chunk (down to 64
        // columns) while the launch would have fewer than 4 workgroups per CU
        const int64_t want = 4 * int64_t(std::max(ctx->num_cus, 1));
        while (chunk > 64 && rowblocks * ((N + chunk - 1) / chunk) * nact < want) chunk /= 2;
    }
    nsplit = N > 0 ? (N + chunk - 1) / chunk : 1;
    c128* part = static_cast<c128*>(ctx_buf(ctx, SLOT_GEMV_PART, size_t(nact) * nsplit * M * 16));
    if (!part) return CCAB_ENOMEM;
    if (reserve_only) return CCAB_OK;
    dim3 grid((unsigned)rowblocks, (unsigned)nsplit, (unsigned)nact);
    ProfScope ps(ctx, prof_cls >= 0 ? prof_cls : (prec == CCAB_C128 || d_Alo) ? PROF_GEMV128 : PROF_GEMV,
                 (rg ? rg->elems : double(M) * double(N) * double(nact)) *
                     (d_Alo ? 12.0 : prec == CCAB_C128 ? 16.0 : prec == CCAB_C32 ? 4.0 : 8.0));
    if (N > 0) {
        const bool nt = tn ? tn->nt != 0 : ctx->gemv_nt;
        // (the split product keeps ordinary loads: non-temporal ones measured
        // slower for its paired 16-byte remainder loads)
        if (d_Alo && R == 4)
            hipLaunchKernelGGL(gemv_split_kernel<true>, grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c32*>(d_A), d_Alo, d_inv_scale, lda, strideA, d_x,
                               stridex, chunk, part, nsplit, d_active, rn, raoff, d_skip);
        else if (d_Alo)
            hipLaunchKernelGGL(gemv_split_kernel<false>, grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c32*>(d_A), d_Alo, d_inv_scale, lda, strideA, d_x,
                               stridex, chunk, part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C128 && nt)
            hipLaunchKernelGGL((gemv_part_kernel<c128, true>), grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c128*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C128)
            hipLaunchKernelGGL(gemv_part_kernel<c128>, grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c128*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C32 && R == 4 && nt)
            hipLaunchKernelGGL((gemv_part_vec_kernel<c32, 4, true>), grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c32*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C32 && R == 4)
            hipLaunchKernelGGL((gemv_part_vec_kernel<c32, 4>), grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c32*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C64 && R == 2 && nt)
            hipLaunchKernelGGL((gemv_part_vec_kernel<c64, 2, true>), grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c64*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C64 && R == 2)
            hipLaunchKernelGGL((gemv_part_vec_kernel<c64, 2>), grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c64*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else if (prec == CCAB_C32)
            hipLaunchKernelGGL(gemv_part_kernel<c32>, grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c32*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
        else
            hipLaunchKernelGGL(gemv_part_kernel<c64>, grid, dim3(256), 0, ctx->stream, M, N,
                               static_cast<const c64*>(d_A), lda, strideA, d_x, stridex, chunk,
                               part, nsplit, d_active, rn, raoff, d_skip);
    } else {
        nsplit = 0;
    }
    hipLaunchKernelGGL(gemv_finish_kernel, dim3((unsigned)finish_blocks, (unsigned)nact), dim3(256),
                       0, ctx->stream, M, part, nsplit, alpha, beta, d_y, stridey, d_active,
                       (prec == CCAB_C32 && !d_Alo) ? d_inv_scale : nullptr, d_skip, rn, chunk);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int gemv_general(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t batch, const void* d_A,
                 int64_t lda, int64_t strideA, const c128* d_x, int64_t stridex, c128 alpha,
                 c128 beta, c128* d_y, int64_t stridey) {
    return gemv_active(ctx, prec, M, N, batch, nullptr, d_A, lda, strideA, d_x, stridex, alpha,
                       beta, d_y, stridey);
}

int gemv_batched(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, const void* d_A, int64_t lda,
                 int64_t strideA, const c128* d_x, int64_t stridex, c128 alpha, c128 beta,
                 c128* d_y, int64_t stridey) {
    return gemv_general(ctx, prec, n, n, batch, d_A, lda, strideA, d_x, stridex, alpha, beta, d_y,
                        stridey);
}

int nrm2_active(ccab_ctx* ctx, int64_t n, int64_t nact, const int32_t* d_active,
                const c128* d_v, int64_t stride, double* d_out) {
    if (nact <= 0) return CCAB_OK;
    hipLaunchKernelGGL(nrm2_kernel, dim3((unsigned)nact), dim3(256), 0, ctx->stream, n, d_v,
                       stride, d_out, d_active);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int nrm2_batched(ccab_ctx* ctx, int64_t n, int64_t batch, const c128* d_v, int64_t stride,
                 double* d_out) {
    hipLaunchKernelGGL(nrm2_kernel, dim3((unsigned)batch), dim3(256), 0, ctx->stream, n, d_v,
                       stride, d_out, nullptr);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int frob_norm(ccab_ctx* ctx, int prec, int64_t n, const void* d_A, int64_t lda, double* d_out) {
    double* part = static_cast<double*>(ctx_buf(ctx, SLOT_NORM, size_t(n + 1) * 8));
    if (!part) return CCAB_ENOMEM;
    if (prec == CCAB_C128)
        hipLaunchKernelGGL(sumsq_kernel<c128>, dim3((unsigned)n), dim3(256), 0, ctx->stream, n,
                           static_cast<const c128*>(d_A), lda, part);
    else
        hipLaunchKernelGGL(sumsq_kernel<c64>, dim3((unsigned)n), dim3(256), 0, ctx->stream, n,
                           static_cast<const c64*>(d_A), lda, part);
    hipLaunchKernelGGL(sum_sqrt_kernel, dim3(1), dim3(256), 0, ctx->stream, n, part, d_out);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int convert_c128_to_c64(ccab_ctx* ctx, int64_t count, const c128* src, c64* dst) {
    hipLaunchKernelGGL(c128_to_c64_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, count, src, dst);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}
int convert_c64_to_c128(ccab_ctx* ctx, int64_t count, const c64* src, c128* dst) {
    hipLaunchKernelGGL(c64_to_c128_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, count, src, dst);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}
