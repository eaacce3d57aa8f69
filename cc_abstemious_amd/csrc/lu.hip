// Batched dense complex LU with partial pivoting on gfx950.
//
// Replaces nalgebra's LU behind solve.rs:166-171 (solve_lu).  Right-looking,
// blocked with panel width NB = 64:
//   1. panel_kernel    one workgroup per system factors the (n-k0) x NB panel
//                      column by column; pivot = first row maximising
//                      |re|+|im| (nalgebra icamax on norm1); an all-zero
//                      column is skipped as nalgebra does.
//   2. trinv_kernel    inverts the NB x NB unit-lower L11 and upper U11 (kept
//                      for the blocked triangular solves).
//   3. laswp_kernel    applies the panel's row swaps to the other columns.
//   4. cgemm (overwrite) U12 = inv(L11) A12.
//   5. cgemm (update)  A22 -= L21 U12  -- the O(n^3) part, on MFMA.
// The complex GEMM runs as a real GEMM of twice the size on the 16x16x4 MFMA
// (f32 for c64, f64 for c128):  [C_re; C_im] interleaved = A' B' where B' and
// C' are the interleaved memory itself and A' is the 2x2 real expansion
// [[Re, -Im], [Im, Re]] built on the fly from LDS.
#include <rocblas/rocblas.h>

#include "ccab_internal.h"
#include "coop.h"

namespace {

constexpr int NB = 64;         // panel width
constexpr int kPanelThreads = 1024;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc_t = floatx4;
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<double> {
    using acc_t = doublex4;
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ float norm1(c64 z) { return fabsf(z.re) + fabsf(z.im); }
__device__ __forceinline__ double norm1(c128 z) { return fabs(z.re) + fabs(z.im); }

// ---------------------------------------------------------------------------
// Panel factorisation: one workgroup per system.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void panel_kernel(Cx<T>* __restrict__ A, int64_t lda,
                                                               int64_t strideA, int64_t n,
                                                               int64_t k0, int nb,
                                                               int32_t* __restrict__ ipiv,
                                                               int32_t* __restrict__ perm,
                                                               int32_t* __restrict__ info) {
    using C = Cx<T>;
    const int64_t b = blockIdx.x;
    C* a = A + b * strideA;
    int32_t* piv = ipiv + b * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    __shared__ T red_v[kPanelThreads / 64];
    __shared__ int64_t red_i[kPanelThreads / 64];
    __shared__ int64_t s_piv;
    __shared__ C s_row[NB];
    __shared__ C s_inv;
    __shared__ int s_zero;

    for (int jj = 0; jj < nb; ++jj) {
        const int64_t c = k0 + jj;
        // --- pivot search: max norm1 over rows c..n-1, first index on ties
        T best = T(-1);
        int64_t bi = n;
        for (int64_t r = c + tid; r < n; r += kPanelThreads) {
            const T v = norm1(a[r + c * lda]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(best, off, 64);
            const int64_t oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = best;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            T bv = red_v[0];
            int64_t bx = red_i[0];
            for (int w = 1; w < kPanelThreads / 64; ++w) {
                if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bx)) {
                    bv = red_v[w];
                    bx = red_i[w];
                }
            }
            const bool zero = !(bv > T(0));
            s_zero = zero ? 1 : 0;
            s_piv = zero ? c : bx;
            piv[c] = int32_t(zero ? c : bx);
            if (zero) {
                info[b] = 1;
            } else if (bx != c) {
                int32_t* pm = perm + b * n;
                const int32_t t = pm[c];
                pm[c] = pm[bx];
                pm[bx] = t;
            }
        }
        __syncthreads();
        if (s_zero) continue;  // nalgebra skips an all-zero column
        const int64_t p = s_piv;
        // --- swap rows c and p within the panel columns
        if (p != c) {
            for (int t = tid; t < nb; t += kPanelThreads) {
                const int64_t col = k0 + t;
                const C x = a[c + col * lda];
                a[c + col * lda] = a[p + col * lda];
                a[p + col * lda] = x;
            }
        }
        __syncthreads();
        // --- pivot row of the remaining panel columns to LDS, reciprocal
        for (int t = tid; t < nb; t += kPanelThreads) s_row[t] = a[c + (k0 + t) * lda];
        if (tid == 0) s_inv = crecip(a[c + c * lda]);
        __syncthreads();
        const C inv = s_inv;
        // --- scale column c below the diagonal, rank-1 update of the panel
        for (int64_t r = c + 1 + tid; r < n; r += kPanelThreads) {
            const C l = a[r + c * lda] * inv;
            a[r + c * lda] = l;
            for (int t = jj + 1; t < nb; ++t) {
                C& dst = a[r + (k0 + t) * lda];
                dst = dst - l * s_row[t];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Inverses of the NB x NB diagonal blocks, column-major, zero padded when
// nb < NB: Linv (unit lower, needed right after each panel for the U12
// update) and Uinv (upper, only for the triangular solves, so all blocks are
// inverted in one launch after the factorisation).  The block is staged in
// LDS; lane c solves for column c with four partial sums per dot product.
// ---------------------------------------------------------------------------
template <typename T, bool UPPER>
__global__ __launch_bounds__(NB) void trinv_kernel(const Cx<T>* __restrict__ A, int64_t lda,
                                                   int64_t strideA, int64_t n, int64_t kb0,
                                                   Cx<T>* __restrict__ Inv, int64_t inv_stride) {
    using C = Cx<T>;
    const int64_t b = blockIdx.x;
    const int64_t kb = kb0 + blockIdx.y;
    const int64_t k0 = kb * NB;
    const int nb = (int)min<int64_t>(NB, n - k0);
    const C* a = A + b * strideA + k0 + k0 * lda;
    C* out = Inv + b * inv_stride + kb * NB * NB;
    __shared__ C blk[NB][NB + 1];  // blk[r][t] = A(k0 + r, k0 + t)
    __shared__ C X[NB][NB + 1];    // X[r][c]
    const int c = threadIdx.x;
    const C zero{T(0), T(0)}, one{T(1), T(0)};
    for (int t = 0; t < NB; ++t) blk[c][t] = (c < nb && t < nb) ? a[c + t * lda] : zero;
    __syncthreads();
    if (!UPPER) {
        // column c of L^{-1}: x_c = 1, x_r = -sum_{t=c}^{r-1} L[r][t] x_t
        for (int r = 0; r < NB; ++r) {
            C x = zero;
            if (c < nb && r < nb && r >= c) {
                if (r == c) {
                    x = one;
                } else {
                    C s0 = zero, s1 = zero, s2 = zero, s3 = zero;
                    int t = c;
                    for (; t + 3 < r; t += 4) {
                        s0 += blk[r][t] * X[t][c];
                        s1 += blk[r][t + 1] * X[t + 1][c];
                        s2 += blk[r][t + 2] * X[t + 2][c];
                        s3 += blk[r][t + 3] * X[t + 3][c];
                    }
                    for (; t < r; ++t) s0 += blk[r][t] * X[t][c];
                    x = -((s0 + s1) + (s2 + s3));
                }
            }
            X[r][c] = x;
        }
    } else {
        // column c of U^{-1}: rows c down to 0
        for (int r = NB - 1; r >= 0; --r) {
            C y = zero;
            if (c < nb && r <= c) {
                C s0 = (r == c) ? one : zero, s1 = zero, s2 = zero, s3 = zero;
                int t = r + 1;
                for (; t + 3 <= c; t += 4) {
                    s0 -= blk[r][t] * X[t][c];
                    s1 -= blk[r][t + 1] * X[t + 1][c];
                    s2 -= blk[r][t + 2] * X[t + 2][c];
                    s3 -= blk[r][t + 3] * X[t + 3][c];
                }
                for (; t <= c; ++t) s0 -= blk[r][t] * X[t][c];
                const C d = blk[r][r];
                y = (d.re == T(0) && d.im == T(0)) ? zero : cdiv((s0 + s1) + (s2 + s3), d);
            }
            X[r][c] = y;
        }
    }
    __syncthreads();
    for (int e = c; e < NB * NB; e += NB) out[e] = X[e % NB][e / NB];
}

// ---------------------------------------------------------------------------
// Slab factorisation: W columns c0..c0+W of rows c0..n-1 held in registers
// (row r on lane (r-c0) % kSlabThreads, slot (r-c0) / kSlabThreads).  W
// pivot steps (block arg-max of |re|+|im|, first index on ties, row exchange
// through LDS, scale, rank-1 update), write back, then
//   - the slab's row swaps on the other columns of the 64-wide panel,
//   - U rows of the panel columns right of the slab: L_ss^{-1} A(c0:c0+W, .)
// The rest of the panel is updated by the MFMA GEMM afterwards, so all
// O(n NB^2) panel work runs on the matrix cores.
// ---------------------------------------------------------------------------
constexpr int kSlabThreads = 1024;

template <typename T, int W, int R>
__global__ __launch_bounds__(kSlabThreads) void slab_factor_kernel(Cx<T>* __restrict__ A, int64_t lda,
                                                                   int64_t strideA, int64_t n,
                                                                   int64_t k0, int nb, int c0o,
                                                                   int32_t* __restrict__ ipiv,
                                                                   int32_t* __restrict__ perm,
                                                                   int32_t* __restrict__ info) {
    using C = Cx<T>;
    constexpr int NW = kSlabThreads / 64;
    const int64_t b = blockIdx.x;
    C* a = A + b * strideA;
    int32_t* piv = ipiv + b * n;
    int32_t* pm = perm + b * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const C zero{T(0), T(0)};
    __shared__ C srow[2][W];
    __shared__ C sL[W][W];
    __shared__ T red_v[NW];
    __shared__ int64_t red_i[NW];
    __shared__ int64_t s_p;
    __shared__ int s_zero;
    __shared__ int64_t s_piv[W];

    const int64_t c0 = k0 + c0o;
    const int weff = min(W, nb - c0o);
    C v[R][W];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t row = c0 + tid + int64_t(q) * kSlabThreads;
#pragma unroll
        for (int w = 0; w < W; ++w) v[q][w] = (row < n && w < weff) ? a[row + (c0 + w) * lda] : zero;
    }
#pragma unroll
    for (int jj = 0; jj < W; ++jj) {
        if (jj < weff) {
            const int64_t dr = c0 + jj;
            T best = T(-1);
            int64_t bi = n;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int64_t row = c0 + tid + int64_t(q) * kSlabThreads;
                if (row >= dr && row < n) {
                    const T x = norm1(v[q][jj]);
                    if (x > best) {
                        best = x;
                        bi = row;
                    }
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const T ov = __shfl_xor(best, off, 64);
                const int64_t oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) {
                    best = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                red_v[wave] = best;
                red_i[wave] = bi;
            }
            __syncthreads();
            if (tid == 0) {
                T bv = red_v[0];
                int64_t bx = red_i[0];
                for (int w = 1; w < NW; ++w)
                    if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bx)) {
                        bv = red_v[w];
                        bx = red_i[w];
                    }
                const bool zcol = !(bv > T(0));
                s_zero = zcol ? 1 : 0;
                const int64_t p = zcol ? dr : bx;
                s_p = p;
                s_piv[jj] = p;
                piv[dr] = int32_t(p);
                if (zcol) {
                    info[b] = 1;
                } else if (p != dr) {
                    const int32_t t = pm[dr];
                    pm[dr] = pm[p];
                    pm[p] = t;
                }
            }
            __syncthreads();
            if (!s_zero) {
                const int64_t p = s_p;
                const int tp = int((p - c0) % kSlabThreads);
                const int qp = int((p - c0) / kSlabThreads);
                if (p != dr) {
                    if (tid == tp) {
#pragma unroll
                        for (int q = 0; q < R; ++q)
                            if (q == qp) {
#pragma unroll
                                for (int w = 0; w < W; ++w) srow[0][w] = v[q][w];
                            }
                    }
                    if (tid == jj) {
#pragma unroll
                        for (int w = 0; w < W; ++w) srow[1][w] = v[0][w];
                    }
                    __syncthreads();
                    if (tid == tp) {
#pragma unroll
                        for (int q = 0; q < R; ++q)
                            if (q == qp) {
#pragma unroll
                                for (int w = 0; w < W; ++w) v[q][w] = srow[1][w];
                            }
                    }
                    if (tid == jj) {
#pragma unroll
                        for (int w = 0; w < W; ++w) v[0][w] = srow[0][w];
                    }
                } else if (tid == jj) {
#pragma unroll
                    for (int w = 0; w < W; ++w) srow[0][w] = v[0][w];
                }
                __syncthreads();
                C prow[W];
#pragma unroll
                for (int w = 0; w < W; ++w) prow[w] = srow[0][w];
                const C inv = crecip(prow[jj]);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int64_t row = c0 + tid + int64_t(q) * kSlabThreads;
                    if (row > dr && row < n) {
                        const C l = v[q][jj] * inv;
                        v[q][jj] = l;
#pragma unroll
                        for (int w = jj + 1; w < W; ++w) v[q][w] -= l * prow[w];
                    }
                }
            }
            __syncthreads();
        } else if (tid == 0) {
            s_piv[jj] = c0 + jj;
        }
    }
    // write back; the top W x W of the slab (L_ss below its diagonal) to LDS
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t row = c0 + tid + int64_t(q) * kSlabThreads;
        if (row < n) {
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (w < weff) a[row + (c0 + w) * lda] = v[q][w];
        }
    }
    if (tid < W) {
#pragma unroll
        for (int w = 0; w < W; ++w) sL[tid][w] = v[0][w];
    }
    __syncthreads();
    // the slab's swaps on the panel columns outside the slab (global, L2)
    for (int cc = tid; cc < nb; cc += kSlabThreads) {
        if (cc >= c0o && cc < c0o + weff) continue;
        C* col = a + (k0 + cc) * lda;
        for (int jj = 0; jj < weff; ++jj) {
            const int64_t r = c0 + jj;
            const int64_t p = s_piv[jj];
            if (p != r) {
                const C x = col[r];
                col[r] = col[p];
                col[p] = x;
            }
        }
        // U rows of the panel columns right of the slab: unit-L forward solve
        if (cc >= c0o + weff) {
            C u[W];
#pragma unroll
            for (int w = 0; w < W; ++w) u[w] = w < weff ? col[c0 + w] : zero;
#pragma unroll
            for (int w = 1; w < W; ++w)
#pragma unroll
                for (int t = 0; t < w; ++t) u[w] -= sL[w][t] * u[t];
#pragma unroll
            for (int w = 1; w < W; ++w)
                if (w < weff) col[c0 + w] = u[w];
        }
    }
}

// ---------------------------------------------------------------------------
// The same slab factorisation spread over G workgroups of one system, for
// small batches whose tall panels would otherwise run on one CU each.
// Workgroup g holds rows c0 + g*RW .. c0 + (g+1)*RW - 1 (RW = 256*RQ) of the
// slab in registers.  Per column, every workgroup publishes its best pivot
// candidate (largest norm1, first row on ties) together with that row, and
// the owner of the diagonal row publishes the diagonal row; after one grid
// barrier all workgroups choose the same pivot by the same rule, the owners
// of the diagonal and pivot rows exchange them, and every workgroup scales
// and updates its own rows.  The pivot rule is the one of panel_kernel and
// slab_factor_kernel; the slab is wider (16 c128 / 32 c64 columns against
// 1-8), so more of the panel's updates are rank-1 steps in registers rather
// than MFMA GEMMs, and the factors agree with those paths to rounding.
//
// Exchange memory per system (64-bit words): a 32-word header (word 0: the
// barrier counter, word 8: the abort flag), then two parities (column jj
// uses jj & 1, so a workgroup still reading column jj never sees column jj+1
// overwrite it) of: G candidates (value, row), G candidate rows of W
// entries, the diagonal row.  Exchange data moves with agent-scope atomic
// loads and stores; the barrier counter is released and acquired at agent
// scope.  A barrier whose wait exceeds spin_limit sleeps sets the abort flag
// and info = -2 instead of hanging; every later barrier then passes at once,
// so the kernel always exits.
// ---------------------------------------------------------------------------
constexpr int kCoopThreads = 256;

template <typename T>
struct CoopWords;
template <>
struct CoopWords<double> {
    static constexpr int N = 2;  // 64-bit words per complex entry
    static __device__ __forceinline__ void put(uint64_t* p, c128 z) {
        __hip_atomic_store(p, uint64_t(__double_as_longlong(z.re)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 1, uint64_t(__double_as_longlong(z.im)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ __forceinline__ c128 get(uint64_t* p) {
        const uint64_t a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return c128{__longlong_as_double((long long)a), __longlong_as_double((long long)b)};
    }
};
template <>
struct CoopWords<float> {
    static constexpr int N = 1;
    static __device__ __forceinline__ void put(uint64_t* p, c64 z) {
        const uint64_t w = uint64_t(__float_as_uint(z.re)) | (uint64_t(__float_as_uint(z.im)) << 32);
        __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ __forceinline__ c64 get(uint64_t* p) {
        const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return c64{__uint_as_float(uint32_t(w)), __uint_as_float(uint32_t(w >> 32))};
    }
};

// words of one parity block and of one system's exchange area
inline int64_t coop_parity_words(int G, int W, int wpc) {
    const int64_t w = 2 * int64_t(G) + (int64_t(G) + 1) * W * wpc;
    return (w + 31) / 32 * 32;
}
inline int64_t coop_system_words(int G, int W, int wpc) {
    return kCoopHeaderWords + 2 * coop_parity_words(G, W, wpc);
}

template <typename T, int W, int RQ>
__global__ __launch_bounds__(kCoopThreads) void coop_slab_kernel(
    Cx<T>* __restrict__ A, int64_t lda, int64_t strideA, int64_t n, int64_t k0, int nb, int c0o,
    int32_t* __restrict__ ipiv, int32_t* __restrict__ perm, int32_t* __restrict__ info,
    uint64_t* __restrict__ xbuf, int64_t sys_words, int64_t par_words, unsigned spin_limit) {
    using C = Cx<T>;
    using X = CoopWords<T>;
    constexpr int NW = kCoopThreads / 64;
    constexpr int64_t RW = int64_t(kCoopThreads) * RQ;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t b = blockIdx.y;
    C* a = A + b * strideA;
    int32_t* piv = ipiv + b * n;
    int32_t* pm = perm + b * n;
    uint64_t* hdr = xbuf + b * sys_words;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const C zero{T(0), T(0)};
    __shared__ C prow_s[W];
    __shared__ T red_v[NW];
    __shared__ int64_t red_i[NW];
    __shared__ int64_t s_p;
    __shared__ int s_zero;
    __shared__ int s_ok;

    const int64_t c0 = k0 + c0o;
    const int weff = min(W, nb - c0o);
    const int64_t rbase = c0 + int64_t(g) * RW;
    C v[RQ][W];
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        const int64_t row = rbase + tid + int64_t(q) * kCoopThreads;
#pragma unroll
        for (int w = 0; w < W; ++w) v[q][w] = (row < n && w < weff) ? a[row + (c0 + w) * lda] : zero;
    }
    // The column loop is not unrolled (its body holds a grid barrier); column
    // jj of the register slab is reached by selects over the unrolled w.
    int jj = 0;
#pragma unroll 1
    for (; jj < weff; ++jj) {
        const int64_t dr = c0 + jj;
        uint64_t* par = hdr + kCoopHeaderWords + (jj & 1) * par_words;
        uint64_t* cand = par;                       // 2 words per workgroup
        uint64_t* rows = par + 2 * int64_t(G);      // W entries per workgroup
        uint64_t* drow = rows + int64_t(G) * W * X::N;
        // --- this workgroup's candidate
        T best = T(-1);
        int64_t bi = n;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            const int64_t row = rbase + tid + int64_t(q) * kCoopThreads;
            C vj = zero;
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (w == jj) vj = v[q][w];
            if (row >= dr && row < n) {
                const T x = norm1(vj);
                if (x > best) {
                    best = x;
                    bi = row;
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(best, off, 64);
            const int64_t oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = best;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            T bv = red_v[0];
            int64_t bx = red_i[0];
            for (int w = 1; w < NW; ++w)
                if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bx)) {
                    bv = red_v[w];
                    bx = red_i[w];
                }
            __hip_atomic_store(cand + 2 * g, uint64_t(__double_as_longlong(double(bv))),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(cand + 2 * g + 1, uint64_t(bx), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            s_p = bx;
        }
        __syncthreads();
        {
            // the candidate row, and the diagonal row, from their owners
            const int64_t wr = s_p;
            const int64_t ow = wr - rbase, od = dr - rbase;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                if (wr < n && ow == tid + int64_t(q) * kCoopThreads) {
#pragma unroll
                    for (int w = 0; w < W; ++w) X::put(rows + (int64_t(g) * W + w) * X::N, v[q][w]);
                }
                if (od == tid + int64_t(q) * kCoopThreads) {
#pragma unroll
                    for (int w = 0; w < W; ++w) X::put(drow + int64_t(w) * X::N, v[q][w]);
                }
            }
        }
        if (!coop_barrier(hdr, uint64_t(jj + 1) * uint64_t(G), spin_limit, &s_ok)) break;
        // --- the global pivot: every workgroup reduces the same candidates
        if (wave == 0) {
            double bv = -1.0;
            int64_t bx = n;
            for (int t = lane; t < G; t += 64) {
                const double x = __longlong_as_double((long long)__hip_atomic_load(
                    cand + 2 * t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                const int64_t i = int64_t(__hip_atomic_load(cand + 2 * t + 1, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT));
                if (x > bv || (x == bv && i < bx)) {
                    bv = x;
                    bx = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off, 64);
                const int64_t oi = __shfl_xor(bx, off, 64);
                if (ov > bv || (ov == bv && oi < bx)) {
                    bv = ov;
                    bx = oi;
                }
            }
            if (lane == 0) {
                const bool zcol = !(bv > 0.0);
                s_zero = zcol ? 1 : 0;
                s_p = zcol ? dr : bx;
            }
        }
        __syncthreads();
        const int64_t p = s_p;
        if (p < dr || p >= n) break;  // only after a failed barrier: never index with it
        if (!s_zero) {
            const int gw = int((p - c0) / RW);
            if (tid < W) prow_s[tid] = X::get(rows + (int64_t(gw) * W + tid) * X::N);
            __syncthreads();
            if (p != dr) {
                const int64_t op = p - rbase, od = dr - rbase;
#pragma unroll
                for (int q = 0; q < RQ; ++q) {
                    if (od == tid + int64_t(q) * kCoopThreads) {
#pragma unroll
                        for (int w = 0; w < W; ++w) v[q][w] = prow_s[w];
                    }
                    if (op == tid + int64_t(q) * kCoopThreads) {
#pragma unroll
                        for (int w = 0; w < W; ++w) v[q][w] = X::get(drow + int64_t(w) * X::N);
                    }
                }
            }
            // scale column jj below the diagonal, rank-1 update of the columns
            // right of it (the pivot row stays in LDS: broadcast reads)
            C pj = zero;
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (w == jj) pj = prow_s[w];
            const C inv = crecip(pj);
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int64_t row = rbase + tid + int64_t(q) * kCoopThreads;
                if (row > dr && row < n) {
                    C vj = zero;
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        if (w == jj) vj = v[q][w];
                    const C l = vj * inv;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        if (w == jj) v[q][w] = l;
                        else if (w > jj) v[q][w] -= l * prow_s[w];
                    }
                }
            }
        }
        if (g == 0 && tid == 0) {
            piv[dr] = int32_t(p);
            if (s_zero) {
                info[b] = 1;
            } else if (p != dr) {
                const int32_t t = pm[dr];
                pm[dr] = pm[p];
                pm[p] = t;
            }
        }
        __syncthreads();
    }
    if (tid == 0 && __hip_atomic_load(hdr + 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        info[b] = -2;
    // after a failed barrier: identity pivots for the columns not done, so
    // every later kernel that reads ipiv stays in range (the result is
    // reported as an error)
    if (g == 0 && tid == 0)
        for (int j = jj; j < weff; ++j) piv[c0 + j] = int32_t(c0 + j);
    // write back this workgroup's rows of the slab
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        const int64_t row = rbase + tid + int64_t(q) * kCoopThreads;
        if (row < n) {
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (w < weff) a[row + (c0 + w) * lda] = v[q][w];
        }
    }
}

// After coop_slab_kernel: the slab's swaps on the other panel columns and
// the U rows right of the slab (L_ss^{-1} A(c0:c0+W, .)), from the pivots and
// the slab's top W x W block in memory; one lane per panel column, the same
// operations in the same order as the tail of slab_factor_kernel.
template <typename T, int W>
__global__ __launch_bounds__(64) void coop_finish_kernel(Cx<T>* __restrict__ A, int64_t lda,
                                                         int64_t strideA, int64_t n, int64_t k0,
                                                         int nb, int c0o,
                                                         const int32_t* __restrict__ ipiv) {
    using C = Cx<T>;
    const int64_t b = blockIdx.x;
    C* a = A + b * strideA;
    const int32_t* piv = ipiv + b * n;
    const int tid = threadIdx.x;
    const C zero{T(0), T(0)};
    const int64_t c0 = k0 + c0o;
    const int weff = min(W, nb - c0o);
    __shared__ C sL[W][W];
    __shared__ int64_t s_piv[W];
    __shared__ C su[W][64];
    for (int e = tid; e < W * W; e += 64) {
        const int r = e % W, c = e / W;
        sL[r][c] = (r < weff && c < weff) ? a[(c0 + r) + (c0 + c) * lda] : zero;
    }
    if (tid < W) {
        const int64_t pv = tid < weff ? int64_t(piv[c0 + tid]) : c0 + tid;
        s_piv[tid] = (pv >= c0 + tid && pv < n) ? pv : c0 + tid;
    }
    __syncthreads();
    for (int cc = tid; cc < nb; cc += 64) {
        if (cc >= c0o && cc < c0o + weff) continue;
        C* col = a + (k0 + cc) * lda;
        for (int jj = 0; jj < weff; ++jj) {
            const int64_t r = c0 + jj;
            const int64_t p = s_piv[jj];
            if (p != r) {
                const C x = col[r];
                col[r] = col[p];
                col[p] = x;
            }
        }
        if (cc >= c0o + weff) {
            // unit-lower forward solve, column by column: entry w still sees
            // its subtractions in the order t = 0, 1, .., w-1
            for (int w = 0; w < weff; ++w) su[w][tid] = col[c0 + w];
            for (int t = 0; t < weff; ++t) {
                const C ut = su[t][tid];
                for (int w = t + 1; w < weff; ++w) su[w][tid] -= sL[w][t] * ut;
            }
            for (int w = 1; w < weff; ++w) col[c0 + w] = su[w][tid];
        }
    }
}

// ---------------------------------------------------------------------------
// Row swaps of pivots r0..r0+np-1 (LAPACK order) applied to the columns in
// [c_lo, c_hi) outside [x_lo, x_hi); one lane per column.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void laswp_kernel(Cx<T>* __restrict__ A, int64_t lda, int64_t strideA, int64_t n,
                             int64_t r0, int np, const int32_t* __restrict__ ipiv, int64_t c_lo,
                             int64_t c_hi, int64_t x_lo, int64_t x_hi) {
    const int64_t b = blockIdx.y;
    const int64_t col = c_lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= c_hi || (col >= x_lo && col < x_hi)) return;
    Cx<T>* a = A + b * strideA + col * lda;
    const int32_t* piv = ipiv + b * n;
    for (int j = 0; j < np; ++j) {
        const int64_t r = r0 + j;
        const int64_t p = piv[r];
        if (p != r) {
            const Cx<T> x = a[r];
            a[r] = a[p];
            a[p] = x;
        }
    }
}

constexpr int TN = 64;
constexpr int kGroupM = 8;   // row tiles per group of the XCD-aware tile order
constexpr int kXcd = 8;      // MI355X: workgroups go round-robin to 8 XCDs
constexpr int KC = 16;       // complex k per LDS chunk

// Map the linear workgroup id to (system, row tile, column tile).
// XCD = true: workgroup w runs on XCD w % 8; each XCD gets a contiguous range
// of the tile sequence, and the sequence walks groups of kGroupM row tiles
// column by column, so an XCD keeps a few A row-panels and B column-panels
// in its own L2.  XCD = false: plain row-tile-fastest order.
template <bool XCD>
__device__ __forceinline__ bool tile_of(int64_t tm, int64_t tn, int64_t batch, int64_t& b,
                                        int64_t& pm, int64_t& pn) {
    const int64_t total = tm * tn * batch;
    const int64_t w = blockIdx.x;
    int64_t t = w;
    if (XCD) {
        const int64_t per = (total + kXcd - 1) / kXcd;
        t = (w % kXcd) * per + w / kXcd;
    }
    if (t >= total) return false;
    b = t / (tm * tn);
    const int64_t r = t % (tm * tn);
    if (XCD) {
        const int64_t in_group = kGroupM * tn;
        const int64_t g = r / in_group;
        const int64_t first = g * kGroupM;
        const int64_t gsz = min<int64_t>(tm - first, kGroupM);
        pm = first + (r % in_group) % gsz;
        pn = (r % in_group) / gsz;
    } else {
        pm = r % tm;
        pn = r / tm;
    }
    return true;
}

// Complex GEMM on MFMA.  MODE 0: C -= A B.  MODE 1: C = A B.
// Workgroup tile TM (complex rows) x 64 (cols), 4 wavefronts in 2x2; a wave
// owns TM/2 complex rows (= TM real rows, TM/16 MFMA tiles) x 32 columns.
// K is walked in chunks of KC complex through LDS.  EPI_LDS: the C tile goes
// through LDS half by half so every global access is a contiguous column
// segment; otherwise lanes update C directly.  With MODE 1 and M <= TM, C may
// alias B: a workgroup reads its whole B column block before its first store.
template <typename T, int MODE, int TM, bool XCD, bool EPI_LDS>
__global__ __launch_bounds__(256) void cgemm_kernel(int64_t M, int64_t N, int64_t K,
                                                    const Cx<T>* __restrict__ A, int64_t lda,
                                                    int64_t sA, const Cx<T>* B, int64_t ldb,
                                                    int64_t sB, Cx<T>* C, int64_t ldc,
                                                    int64_t sC, int64_t batch) {
    using Cxt = Cx<T>;
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    constexpr int RT = TM / 16;  // MFMA row tiles per wave
    const int64_t tm = (M + TM - 1) / TM, tn = (N + TN - 1) / TN;
    int64_t bz, pm, pn;
    if (!tile_of<XCD>(tm, tn, batch, bz, pm, pn)) return;
    const int64_t m0 = pm * TM;
    const int64_t n0 = pn * TN;
    const Cxt* a = A + bz * sA;
    const Cxt* bmat = B + bz * sB;
    Cxt* c = C + bz * sC;

    constexpr size_t kA = sizeof(Cxt) * KC * (TM + 1);
    constexpr size_t kB = sizeof(Cxt) * KC * (TN + 1);
    constexpr size_t kCs = EPI_LDS ? sizeof(Cxt) * (TN / 2) * (TM + 1) : 0;
    constexpr size_t kS = (kA + kB) > kCs ? (kA + kB) : kCs;
    __shared__ __attribute__((aligned(16))) char smem[kS];
    Cxt(*As)[TM + 1] = reinterpret_cast<Cxt(*)[TM + 1]>(smem);        // [k][row]
    Cxt(*Bs)[TN + 1] = reinterpret_cast<Cxt(*)[TN + 1]>(smem + kA);   // [k][col]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave & 1;   // row half
    const int wc = wave >> 1;  // 32-col half
    const int l16 = lane & 15;
    const int lq = lane >> 4;  // 0..3: real k index within the MFMA step
    // MFMA row l16 -> real row of the tile.  The f32 instruction returns row
    // 4*(lane/16) + v in register v, the f64 one row 4*v + lane/16; feeding
    // the f64 A rows transposed within each 16 makes both precisions return
    // real rows 4*lq + v, i.e. the re/im pairs of two complex rows.
    const int mrow = sizeof(T) == 8 ? 4 * (l16 & 3) + (l16 >> 2) : l16;

    acc_t acc[RT][2];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = MF::zero();

    for (int64_t kc0 = 0; kc0 < K; kc0 += KC) {
        for (int e = tid; e < TM * KC; e += 256) {
            const int r = e % TM, kk = e / TM;
            const int64_t gr = m0 + r, gk = kc0 + kk;
            As[kk][r] = (gr < M && gk < K) ? a[gr + gk * lda] : Cxt{T(0), T(0)};
        }
        for (int e = tid; e < TN * KC; e += 256) {
            const int kk = e % KC, cc = e / KC;
            const int64_t gk = kc0 + kk, gc = n0 + cc;
            Bs[kk][cc] = (gk < K && gc < N) ? bmat[gk + gc * ldb] : Cxt{T(0), T(0)};
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2 * KC; ks += 4) {
            const int rk = ks + lq;     // real k of this lane
            const int lk = rk >> 1;     // complex k
            const int tk = rk & 1;      // 0: multiplies Re(B), 1: multiplies Im(B)
            T av[RT], bv[2];
#pragma unroll
            for (int tr = 0; tr < RT; ++tr) {
                const int rrow = wr * TM + tr * 16 + mrow;
                const Cxt z = As[lk][rrow >> 1];
                const int sgn = rrow & 1;
                // A'[2i+s][2l+t]: s=0: (Re, -Im), s=1: (Im, Re)
                av[tr] = sgn == 0 ? (tk == 0 ? z.re : -z.im) : (tk == 0 ? z.im : z.re);
            }
#pragma unroll
            for (int tc = 0; tc < 2; ++tc) {
                const Cxt z = Bs[lk][wc * 32 + tc * 16 + l16];
                bv[tc] = tk == 0 ? z.re : z.im;
            }
#pragma unroll
            for (int tr = 0; tr < RT; ++tr)
#pragma unroll
                for (int tc = 0; tc < 2; ++tc) acc[tr][tc] = MF::mma(av[tr], bv[tc], acc[tr][tc]);
        }
        __syncthreads();
    }
    if constexpr (EPI_LDS) {
        Cxt(*Cs)[TM + 1] = reinterpret_cast<Cxt(*)[TM + 1]>(smem);   // [col][row]
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            if (wc == half) {
#pragma unroll
                for (int tr = 0; tr < RT; ++tr)
#pragma unroll
                    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int cr = (wr * TM + tr * 16) / 2 + 2 * lq + h;
                            Cs[tc * 16 + l16][cr] = Cxt{acc[tr][tc][2 * h], acc[tr][tc][2 * h + 1]};
                        }
            }
            __syncthreads();
            for (int e = tid; e < TM * (TN / 2); e += 256) {
                const int cr = e % TM, cc = e / TM;
                const int64_t row = m0 + cr, col = n0 + half * (TN / 2) + cc;
                if (row < M && col < N) {
                    Cxt* dst = c + row + col * ldc;
                    if (MODE == 0) *dst = *dst - Cs[cc][cr];
                    else *dst = Cs[cc][cr];
                }
            }
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int tr = 0; tr < RT; ++tr)
#pragma unroll
            for (int tc = 0; tc < 2; ++tc) {
                const int64_t col = n0 + wc * 32 + tc * 16 + l16;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int64_t row = m0 + (wr * TM + tr * 16) / 2 + 2 * lq + h;
                    if (row < M && col < N) {
                        const Cxt v{acc[tr][tc][2 * h], acc[tr][tc][2 * h + 1]};
                        Cxt* dst = c + row + col * ldc;
                        if (MODE == 0) *dst = *dst - v;
                        else *dst = v;
                    }
                }
            }
    }
}

// Variant table: bit 0 = XCD-aware order, bit 1 = LDS epilogue, bit 2 = TM 128.
constexpr int kGemmVariants = 8;

template <typename T, int MODE, int V>
void gemm_variant(ccab_ctx* ctx, int64_t grid, int64_t M, int64_t N, int64_t K, const Cx<T>* A,
                  int64_t lda, int64_t sA, const Cx<T>* B, int64_t ldb, int64_t sB, Cx<T>* C,
                  int64_t ldc, int64_t sC, int64_t batch) {
    constexpr int TMv = (V & 4) ? 128 : 64;
    hipLaunchKernelGGL((cgemm_kernel<T, MODE, TMv, (V & 1) != 0, (V & 2) != 0>),
                       dim3((unsigned)grid), dim3(256), 0, ctx->stream, M, N, K, A, lda, sA, B,
                       ldb, sB, C, ldc, sC, batch);
}

// Variant 8: the trailing update C -= A B as rocBLAS's strided-batched
// complex GEMM on the context's stream (column-major, as the LU stores its
// systems).  It is never captured into a HIP graph: under capture, and for
// C = A B (whose C may alias B), the MFMA kernel's variant 3 runs instead.
rocblas_handle blas_of(ccab_ctx* ctx) {
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
        ctx->blas = h;
    }
    return static_cast<rocblas_handle>(ctx->blas);
}

template <typename T>
int gemm_library(ccab_ctx* ctx, int64_t M, int64_t N, int64_t K, const Cx<T>* A, int64_t lda,
                 int64_t sA, const Cx<T>* B, int64_t ldb, int64_t sB, Cx<T>* C, int64_t ldc,
                 int64_t sC, int64_t batch) {
    constexpr int64_t kMax = 0x7fffffff;
    CCAB_CHECK(M <= kMax && N <= kMax && K <= kMax && lda <= kMax && ldb <= kMax && ldc <= kMax &&
                   batch <= kMax,
               "GEMM dimensions beyond rocBLAS's 32-bit sizes");
    rocblas_handle h = blas_of(ctx);
    if (!h) return ctx_fail(ctx, CCAB_EHIP, "rocblas_create_handle failed");
    if (rocblas_set_stream(h, ctx->stream) != rocblas_status_success)
        return ctx_fail(ctx, CCAB_EHIP, "rocblas_set_stream failed");
    rocblas_status st;
    if constexpr (sizeof(T) == 8) {
        const rocblas_double_complex alpha{-1.0, 0.0}, beta{1.0, 0.0};
        st = rocblas_zgemm_strided_batched(
            h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)M, (rocblas_int)N,
            (rocblas_int)K, &alpha, reinterpret_cast<const rocblas_double_complex*>(A),
            (rocblas_int)lda, sA, reinterpret_cast<const rocblas_double_complex*>(B),
            (rocblas_int)ldb, sB, &beta, reinterpret_cast<rocblas_double_complex*>(C),
            (rocblas_int)ldc, sC, (rocblas_int)batch);
    } else {
        const rocblas_float_complex alpha{-1.0f, 0.0f}, beta{1.0f, 0.0f};
        st = rocblas_cgemm_strided_batched(
            h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)M, (rocblas_int)N,
            (rocblas_int)K, &alpha, reinterpret_cast<const rocblas_float_complex*>(A),
            (rocblas_int)lda, sA, reinterpret_cast<const rocblas_float_complex*>(B),
            (rocblas_int)ldb, sB, &beta, reinterpret_cast<rocblas_float_complex*>(C),
            (rocblas_int)ldc, sC, (rocblas_int)batch);
    }
    if (st != rocblas_status_success)
        return ctx_fail(ctx, CCAB_EHIP, "rocBLAS complex GEMM failed (status %d)", (int)st);
    return CCAB_OK;
}

template <typename T, int MODE>
int gemm_launch(ccab_ctx* ctx, int64_t M, int64_t N, int64_t K, const Cx<T>* A, int64_t lda,
                int64_t sA, const Cx<T>* B, int64_t ldb, int64_t sB, Cx<T>* C, int64_t ldc,
                int64_t sC, int64_t batch) {
    if (M <= 0 || N <= 0 || K <= 0) return CCAB_OK;
    const int pi = sizeof(T) == 8 ? 0 : 1;
    int v = ctx->gemm_variant[pi];
    if (v == kGemmLibVariant) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        CCAB_HIP(hipStreamIsCapturing(ctx->stream, &cs));
        if (MODE == 0 && cs == hipStreamCaptureStatusNone)
            return gemm_library<T>(ctx, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
        v = 3;
    }
    if (MODE == 1 || M <= 64) v &= ~4;  // in-place U12 needs one row tile
    const int64_t tmv = (v & 4) ? 128 : 64;
    const int64_t total = ((M + tmv - 1) / tmv) * ((N + TN - 1) / TN) * batch;
    const int64_t grid = (v & 1) ? ((total + kXcd - 1) / kXcd) * kXcd : total;
    CCAB_CHECK(grid < (int64_t(1) << 31), "GEMM grid too large");
    switch (v) {
    case 0: gemm_variant<T, MODE, 0>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 1: gemm_variant<T, MODE, 1>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 2: gemm_variant<T, MODE, 2>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 3: gemm_variant<T, MODE, 3>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 4: gemm_variant<T, MODE, 4>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 5: gemm_variant<T, MODE, 5>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    case 6: gemm_variant<T, MODE, 6>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    default: gemm_variant<T, MODE, 7>(ctx, grid, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch); break;
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

// ---------------------------------------------------------------------------
// Blocked triangular solves with the stored diagonal-block inverses.
// Forward (unit L): X[kb] = Linv_kb * Y[kb]; Y[rows below] -= L[below, kb] X[kb]
// Backward (U):     X[kb] = Uinv_kb * Y[kb]; Y[rows above] -= U[above, kb] X[kb]
// Y is read-only for block kb in its step, so every workgroup recomputes
// X[kb] from it and workgroup 0 stores it.
// ---------------------------------------------------------------------------
template <typename T, bool FWD>
__global__ __launch_bounds__(256) void trsv_step_kernel(const Cx<T>* __restrict__ LU, int64_t lda,
                                                        int64_t strideA, int64_t n, int64_t k0,
                                                        int nb, const Cx<T>* __restrict__ inv,
                                                        int64_t inv_stride, int64_t blk,
                                                        Cx<T>* Y, Cx<T>* X, int64_t ldb,
                                                        int64_t strideB, int64_t nrhs) {
    using C = Cx<T>;
    const int64_t b = blockIdx.y / nrhs;
    const int64_t rhs = blockIdx.y % nrhs;
    const C* a = LU + b * strideA;
    const C* iv = inv + b * inv_stride + blk * NB * NB;
    C* y = Y + b * strideB + rhs * ldb;
    C* x = X + b * strideB + rhs * ldb;
    __shared__ C ys[NB];
    __shared__ C xs[NB];
    const int tid = threadIdx.x;
    if (tid < NB) ys[tid] = tid < nb ? y[k0 + tid] : C{T(0), T(0)};
    __syncthreads();
    if (tid < NB) {
        C s{T(0), T(0)};
        for (int t = 0; t < nb; ++t) s += iv[tid + t * NB] * ys[t];
        xs[tid] = s;
        if (blockIdx.x == 0 && tid < nb) x[k0 + tid] = s;
    }
    __syncthreads();
    // rows to update: FWD -> [k0+nb, n), BWD -> [0, k0)
    const int64_t r0 = FWD ? k0 + nb : 0;
    const int64_t r1 = FWD ? n : k0;
    const int64_t r = r0 + int64_t(blockIdx.x) * 256 + tid;
    if (r < r1) {
        C s{T(0), T(0)};
        for (int t = 0; t < nb; ++t) s += a[r + (k0 + t) * lda] * xs[t];
        y[r] = y[r] - s;
    }
}

template <typename T>
__global__ void gather_rows_kernel(const int32_t* __restrict__ perm, int64_t n,
                                   const Cx<T>* __restrict__ B, int64_t ldb, int64_t strideB,
                                   Cx<T>* __restrict__ Y, int64_t nrhs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y / nrhs, r = blockIdx.y % nrhs;
    if (i < n) Y[(b * nrhs + r) * n + i] = B[b * strideB + r * ldb + perm[b * n + i]];
}

template <typename T>
__global__ void copy_kernel(const Cx<T>* __restrict__ src, int64_t lds, int64_t ss, Cx<T>* dst,
                            int64_t ldd, int64_t sd, int64_t n, int64_t nrhs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y / nrhs, r = blockIdx.y % nrhs;
    if (i < n) dst[b * sd + r * ldd + i] = src[b * ss + r * lds + i];
}

__global__ void init_perm_kernel(int32_t* perm, int64_t n, int32_t* info) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) perm[blockIdx.y * n + i] = int32_t(i);
    if (i == 0) info[blockIdx.y] = 0;
}

// perm from LAPACK-order pivots (for factors produced elsewhere)
__global__ void perm_from_ipiv_kernel(const int32_t* __restrict__ ipiv, int64_t n, int32_t* perm,
                                      int64_t batch) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    int32_t* pm = perm + b * n;
    const int32_t* pv = ipiv + b * n;
    for (int64_t i = 0; i < n; ++i) pm[i] = int32_t(i);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t p = pv[i];
        if (p != i) {
            const int32_t t = pm[i];
            pm[i] = pm[p];
            pm[p] = t;
        }
    }
}

inline int64_t nblocks_of(int64_t n) { return (n + NB - 1) / NB; }

template <typename T, int W, int R>
void slab_factor(ccab_ctx* ctx, Cx<T>* A, int64_t lda, int64_t sA, int64_t n, int64_t k0, int nb,
                 int c0o, int32_t* ipiv, int32_t* perm, int32_t* info, int64_t batch) {
    hipLaunchKernelGGL((slab_factor_kernel<T, W, R>), dim3((unsigned)batch), dim3(kSlabThreads), 0,
                       ctx->stream, A, lda, sA, n, k0, nb, c0o, ipiv, perm, info);
}

template <typename T, int MODE>
int gemm_launch(ccab_ctx* ctx, int64_t M, int64_t N, int64_t K, const Cx<T>* A, int64_t lda,
                int64_t sA, const Cx<T>* B, int64_t ldb, int64_t sB, Cx<T>* C, int64_t ldc,
                int64_t sC, int64_t batch);

// Right-looking over slabs of W columns: register pivoting of the slab, then
// the MFMA GEMM update of the panel columns right of it.  Rows per lane
// R = ceil((n - k0) / kSlabThreads) sets W so that the R x W register block
// stays within the 128 VGPRs a 1024-lane workgroup allows.  Panels taller
// than 16 rows per lane use the generic column-by-column kernel.
template <typename T, int W, int R>
int panel_slabs(ccab_ctx* ctx, Cx<T>* A, int64_t lda, int64_t sA, int64_t n, int64_t k0, int nb,
                int32_t* ipiv, int32_t* perm, int32_t* info, int64_t batch) {
    for (int c0o = 0; c0o < nb; c0o += W) {
        slab_factor<T, W, R>(ctx, A, lda, sA, n, k0, nb, c0o, ipiv, perm, info, batch);
        const int w = std::min(W, nb - c0o);
        const int64_t c1 = k0 + c0o + w;         // first column right of the slab
        const int64_t rest = k0 + nb - c1;       // panel columns still to update
        if (rest > 0 && n - c1 > 0) {
            CCAB_TRY((gemm_launch<T, 0>(ctx, n - c1, rest, w, A + c1 + (c1 - w) * lda, lda, sA,
                                        A + (c1 - w) + c1 * lda, lda, sA, A + c1 + c1 * lda, lda,
                                        sA, batch)));
        }
    }
    return CCAB_OK;
}

// Width of the slab a coop_slab_kernel holds in registers: 16 complex
// entries per row and lane for c128, 32 for c64.
template <typename T, int RQ>
constexpr int coop_width() { return (sizeof(T) == 8 ? 16 : 32) / RQ; }

// Exchange words per system: sized for the largest grid (ctx->num_cus).
inline int64_t coop_words_per_system(const ccab_ctx* ctx) {
    return coop_system_words(std::max(ctx->num_cus, 1), 32, 2);
}

template <typename T, int RQ>
int coop_panel_slabs(ccab_ctx* ctx, Cx<T>* A, int64_t lda, int64_t sA, int64_t n, int64_t k0,
                     int nb, int32_t* ipiv, int32_t* perm, int32_t* info, int64_t batch,
                     uint64_t* xbuf) {
    constexpr int W = coop_width<T, RQ>();
    constexpr int64_t RW = int64_t(kCoopThreads) * RQ;
    const int64_t sys_words = coop_words_per_system(ctx);
    for (int c0o = 0; c0o < nb; c0o += W) {
        const int64_t c0 = k0 + c0o;
        const int G = int((n - c0 + RW - 1) / RW);
        const int64_t par_words = coop_parity_words(G, W, CoopWords<T>::N);
        hipLaunchKernelGGL(coop_reset_kernel, dim3((unsigned)((batch + 63) / 64)), dim3(64), 0,
                           ctx->stream, xbuf, sys_words, batch);
        hipLaunchKernelGGL((coop_slab_kernel<T, W, RQ>), dim3((unsigned)G, (unsigned)batch),
                           dim3(kCoopThreads), 0, ctx->stream, A, lda, sA, n, k0, nb, c0o, ipiv,
                           perm, info, xbuf, sys_words, par_words, kCoopSpinLimit);
        hipLaunchKernelGGL((coop_finish_kernel<T, W>), dim3((unsigned)batch), dim3(64), 0,
                           ctx->stream, A, lda, sA, n, k0, nb, c0o, ipiv);
        const int w = std::min(W, nb - c0o);
        const int64_t c1 = c0 + w;
        const int64_t rest = k0 + nb - c1;
        if (rest > 0 && n - c1 > 0) {
            CCAB_TRY((gemm_launch<T, 0>(ctx, n - c1, rest, w, A + c1 + (c1 - w) * lda, lda, sA,
                                        A + (c1 - w) + c1 * lda, lda, sA, A + c1 + c1 * lda, lda,
                                        sA, batch)));
        }
    }
    return CCAB_OK;
}

template <typename T>
int launch_panel(ccab_ctx* ctx, Cx<T>* A, int64_t lda, int64_t sA, int64_t n, int64_t k0, int nb,
                 int32_t* ipiv, int32_t* perm, int32_t* info, int64_t batch, uint64_t* xbuf) {
    const int64_t m = n - k0;
    // small batches of tall panels: spread each system's slab over many CUs
    // (a grid of at most num_cus workgroups, so all of them are resident)
    if (xbuf && ctx->lu_coop_rows > 0 && m >= ctx->lu_coop_rows) {
        if ((m + kCoopThreads - 1) / kCoopThreads * batch <= ctx->num_cus)
            return coop_panel_slabs<T, 1>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch, xbuf);
        if ((m + 2 * kCoopThreads - 1) / (2 * kCoopThreads) * batch <= ctx->num_cus)
            return coop_panel_slabs<T, 2>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch, xbuf);
    }
    const int64_t R = (m + kSlabThreads - 1) / kSlabThreads;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int WA = F32 ? 8 : 4;   // R <= 4
    constexpr int WB = F32 ? 4 : 2;   // R <= 8
    constexpr int WC = F32 ? 2 : 1;   // R <= 16
    switch (R) {
    case 1: return panel_slabs<T, WA, 1>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 2: return panel_slabs<T, WA, 2>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 3: return panel_slabs<T, WA, 3>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 4: return panel_slabs<T, WA, 4>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 5: case 6: return panel_slabs<T, WB, 6>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 7: case 8: return panel_slabs<T, WB, 8>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 9: case 10: case 11: case 12:
        return panel_slabs<T, WC, 12>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    case 13: case 14: case 15: case 16:
        return panel_slabs<T, WC, 16>(ctx, A, lda, sA, n, k0, nb, ipiv, perm, info, batch);
    default:
        hipLaunchKernelGGL(panel_kernel<T>, dim3((unsigned)batch), dim3(kPanelThreads), 0,
                           ctx->stream, A, lda, sA, n, k0, nb, ipiv, perm, info);
        return CCAB_OK;
    }
}

template <typename T>
int lu_aux(ccab_ctx* ctx, int64_t n, int64_t batch, Cx<T>** linv, Cx<T>** uinv, int32_t** perm,
           int32_t** info, uint64_t** xbuf = nullptr) {
    const size_t inv = size_t(batch) * nblocks_of(n) * NB * NB * sizeof(Cx<T>);
    *linv = static_cast<Cx<T>*>(ctx_buf(ctx, SLOT_LU_LINV, inv));
    *uinv = static_cast<Cx<T>*>(ctx_buf(ctx, SLOT_LU_UINV, inv));
    *perm = static_cast<int32_t*>(ctx_buf(ctx, SLOT_LU_PERM, size_t(batch) * n * 4));
    *info = static_cast<int32_t*>(ctx_buf(ctx, SLOT_LU_INFO, size_t(batch) * 4));
    if (!*linv || !*uinv || !*perm || !*info) return CCAB_ENOMEM;
    if (xbuf) {
        // exchange memory of the multi-workgroup panel, when it can be used
        *xbuf = nullptr;
        if (ctx->lu_coop_rows > 0 && n >= ctx->lu_coop_rows && batch <= ctx->num_cus) {
            *xbuf = static_cast<uint64_t*>(
                ctx_buf(ctx, SLOT_LU_COOP, size_t(batch) * coop_words_per_system(ctx) * 8));
            if (!*xbuf) return CCAB_ENOMEM;
        }
    }
    return CCAB_OK;
}

// The whole factorisation as a fixed launch sequence (no allocation, no host
// synchronisation): what a HIP graph captures.
template <typename T>
int lu_factor_launch(ccab_ctx* ctx, int64_t n, int64_t batch, Cx<T>* A, int64_t lda,
                     int64_t strideA, int32_t* ipiv, Cx<T>* linv, Cx<T>* uinv, int32_t* perm,
                     int32_t* info, uint64_t* xbuf) {
    const int64_t inv_stride = nblocks_of(n) * NB * NB;
    hipLaunchKernelGGL(init_perm_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)batch),
                       dim3(256), 0, ctx->stream, perm, n, info);
    const double B = double(batch);
    auto laswp = [&](int64_t r0, int np, int64_t c_lo, int64_t c_hi, int64_t x_lo, int64_t x_hi) {
        if (c_hi <= c_lo) return;
        ProfScope ps(ctx, PROF_LU_LASWP, 0.0);
        hipLaunchKernelGGL(laswp_kernel<T>,
                           dim3((unsigned)((c_hi - c_lo + 255) / 256), (unsigned)batch),
                           dim3(256), 0, ctx->stream, A, lda, strideA, n, r0, np, ipiv, c_lo, c_hi,
                           x_lo, x_hi);
    };
    // U rows j0..j0+nb of columns [c, c+w): in place, U = Linv(j0) A
    auto trsm = [&](int64_t j0, int nb, int64_t c, int64_t w) -> int {
        if (w <= 0) return CCAB_OK;
        ProfScope ps(ctx, PROF_LU_TRSM, 8.0 * nb * double(w) * nb * B);
        const int64_t kb = j0 / NB;
        return gemm_launch<T, 1>(ctx, nb, w, nb, linv + kb * NB * NB, NB, inv_stride,
                                 A + j0 + c * lda, lda, strideA, A + j0 + c * lda, lda, strideA,
                                 batch);
    };
    // A[r:r+m, c:c+w] -= A[r:r+m, k:k+kk] A[k:k+kk, c:c+w]
    auto update = [&](int64_t r, int64_t m, int64_t c, int64_t w, int64_t k, int64_t kk,
                      int cls) -> int {
        if (m <= 0 || w <= 0) return CCAB_OK;
        ProfScope ps(ctx, cls, 8.0 * double(m) * double(w) * double(kk) * B);
        return gemm_launch<T, 0>(ctx, m, w, kk, A + r + k * lda, lda, strideA, A + k + c * lda,
                                 lda, strideA, A + r + c * lda, lda, strideA, batch);
    };
    // Outer panels of NBO = lu_outer columns (64 or 128) made of 64-wide inner
    // panels; the trailing update then has K = NBO, which halves the C traffic
    // of the O(n^3) part at NBO = 128.
    const int NBO = ctx->lu_outer;
    for (int64_t k0 = 0; k0 < n; k0 += NBO) {
        const int64_t nbo = std::min<int64_t>(NBO, n - k0);
        for (int64_t j0 = k0; j0 < k0 + nbo; j0 += NB) {
            const int nb = (int)std::min<int64_t>(NB, k0 + nbo - j0);
            const int64_t kb = j0 / NB;
            double pflops = 0.0;
            for (int j = 0; j < nb; ++j) pflops += 8.0 * double(n - j0 - j - 1) * double(nb - j - 1);
            {
                ProfScope ps(ctx, PROF_LU_PANEL, pflops * B);
                CCAB_TRY(launch_panel<T>(ctx, A, lda, strideA, n, j0, nb, ipiv, perm, info, batch,
                                         xbuf));
            }
            {
                ProfScope ps(ctx, PROF_LU_TRINV, 0.0);
                hipLaunchKernelGGL((trinv_kernel<T, false>), dim3((unsigned)batch, 1), dim3(NB),
                                   0, ctx->stream, A, lda, strideA, n, kb, linv, inv_stride);
            }
            // this inner panel's swaps on the rest of the outer panel
            laswp(j0, nb, k0, k0 + nbo, j0, j0 + nb);
            const int64_t c1 = j0 + nb, w = k0 + nbo - c1;
            if (w > 0) {
                CCAB_TRY(trsm(j0, nb, c1, w));
                CCAB_TRY(update(c1, n - c1, c1, w, j0, nb, PROF_LU_UPDATE));
            }
        }
        // all swaps of the outer panel on the columns outside it
        laswp(k0, (int)nbo, 0, n, k0, k0 + nbo);
        const int64_t rest = n - k0 - nbo;
        if (rest > 0) {
            const int64_t c1 = k0 + nbo;
            for (int64_t j0 = k0; j0 < k0 + nbo; j0 += NB) {
                const int nb = (int)std::min<int64_t>(NB, k0 + nbo - j0);
                CCAB_TRY(trsm(j0, nb, c1, rest));
                const int64_t r1 = j0 + nb;
                if (r1 < k0 + nbo) CCAB_TRY(update(r1, k0 + nbo - r1, c1, rest, j0, nb, PROF_LU_TRSM));
            }
            CCAB_TRY(update(c1, rest, c1, rest, k0, nbo, PROF_LU_UPDATE));
        }
    }
    {
        ProfScope ps(ctx, PROF_LU_TRINV, 0.0);
        hipLaunchKernelGGL((trinv_kernel<T, true>), dim3((unsigned)batch, (unsigned)nblocks_of(n)),
                           dim3(NB), 0, ctx->stream, A, lda, strideA, n, int64_t(0), uinv,
                           inv_stride);
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

template <typename T>
int lu_factor_t(ccab_ctx* ctx, int64_t n, int64_t batch, Cx<T>* A, int64_t lda, int64_t strideA,
                int32_t* ipiv) {
    Cx<T>*linv, *uinv;
    int32_t *perm, *info;
    uint64_t* xbuf;
    CCAB_TRY(lu_aux<T>(ctx, n, batch, &linv, &uinv, &perm, &info, &xbuf));
    const int prec = sizeof(T) == 8 ? CCAB_C128 : CCAB_C64;
    if (ctx->use_graphs && !ctx->prof_on && ctx->stream) {
        // replay a captured factorisation for this shape, these buffers and
        // this tuning; capture it on first use
        GraphKey key{prec, n, batch, lda, strideA, A, ipiv, linv, uinv, perm, info, xbuf,
                     ctx->lu_outer, ctx->gemm_variant[prec == CCAB_C128 ? 0 : 1],
                     ctx->lu_coop_rows};
        hipGraphExec_t exec = graph_find(ctx, key);
        if (!exec) {
            CCAB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const int rc = lu_factor_launch<T>(ctx, n, batch, A, lda, strideA, ipiv, linv, uinv,
                                               perm, info, xbuf);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
            if (rc != CCAB_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            CCAB_HIP(ec);
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            CCAB_HIP(ei);
            graph_store(ctx, key, exec);
        }
        CCAB_HIP(hipGraphLaunch(exec, ctx->stream));
    } else {
        CCAB_TRY(lu_factor_launch<T>(ctx, n, batch, A, lda, strideA, ipiv, linv, uinv, perm, info,
                                     xbuf));
    }
    ctx->lu_key = A;
    ctx->lu_prec = sizeof(T) == 8 ? CCAB_C128 : CCAB_C64;
    ctx->lu_n = n;
    ctx->lu_batch = batch;
    ctx->lu_lda = lda;
    ctx->lu_stride = strideA;
    // singular check (one small readback)
    CCAB_TRY(pinned_reserve(ctx, size_t(batch) * 4));
    CCAB_HIP(hipMemcpyAsync(ctx->pinned, info, size_t(batch) * 4, hipMemcpyDeviceToHost,
                            ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    const int32_t* hi = static_cast<const int32_t*>(ctx->pinned);
    for (int64_t b = 0; b < batch; ++b) {
        if (hi[b] == -2)
            return ctx_fail(ctx, CCAB_EHIP, "LU panel grid barrier timed out (system %lld)",
                            (long long)b);
        if (hi[b]) return ctx_fail(ctx, CCAB_ESINGULAR, "zero pivot in system %lld", (long long)b);
    }
    return CCAB_OK;
}

template <typename T>
int lu_solve_t(ccab_ctx* ctx, int64_t n, int64_t batch, const Cx<T>* LU, int64_t lda,
               int64_t strideA, const int32_t* ipiv, Cx<T>* B, int64_t ldb, int64_t strideB,
               int64_t nrhs) {
    Cx<T>*linv, *uinv;
    int32_t *perm, *info;
    const bool cached = ctx->lu_key == LU && ctx->lu_n == n && ctx->lu_batch == batch &&
                        ctx->lu_lda == lda && ctx->lu_stride == strideA &&
                        ctx->lu_prec == (sizeof(T) == 8 ? CCAB_C128 : CCAB_C64);
    CCAB_TRY(lu_aux<T>(ctx, n, batch, &linv, &uinv, &perm, &info));
    const int64_t inv_stride = nblocks_of(n) * NB * NB;
    const int64_t nblk = nblocks_of(n);
    if (!cached) {
        // factors from elsewhere: rebuild the diagonal inverses and permutation
        hipLaunchKernelGGL(perm_from_ipiv_kernel, dim3((unsigned)((batch + 63) / 64)), dim3(64), 0,
                           ctx->stream, ipiv, n, perm, batch);
        hipLaunchKernelGGL((trinv_kernel<T, false>), dim3((unsigned)batch, (unsigned)nblk),
                           dim3(NB), 0, ctx->stream, LU, lda, strideA, n, int64_t(0), linv,
                           inv_stride);
        hipLaunchKernelGGL((trinv_kernel<T, true>), dim3((unsigned)batch, (unsigned)nblk),
                           dim3(NB), 0, ctx->stream, LU, lda, strideA, n, int64_t(0), uinv,
                           inv_stride);
        ctx->lu_key = LU;
        ctx->lu_prec = sizeof(T) == 8 ? CCAB_C128 : CCAB_C64;
        ctx->lu_n = n;
        ctx->lu_batch = batch;
        ctx->lu_lda = lda;
        ctx->lu_stride = strideA;
    }
    const int64_t ystride = nrhs * n;
    Cx<T>* Y = static_cast<Cx<T>*>(ctx_buf(ctx, SLOT_LU_Y, size_t(batch) * ystride * sizeof(Cx<T>)));
    if (!Y) return CCAB_ENOMEM;
    const unsigned nbr = (unsigned)(batch * nrhs);
    const dim3 vgrid((unsigned)((n + 255) / 256), nbr);
    hipLaunchKernelGGL(gather_rows_kernel<T>, vgrid, dim3(256), 0, ctx->stream, perm, n, B, ldb,
                       strideB, Y, nrhs);
    // forward: B <- L^{-1} Y  (block by block)
    for (int64_t kb = 0; kb < nblk; ++kb) {
        const int64_t k0 = kb * NB;
        const int nb = (int)std::min<int64_t>(NB, n - k0);
        const int64_t rows = n - k0 - nb;
        const unsigned gx = (unsigned)std::max<int64_t>(1, (rows + 255) / 256);
        ProfScope ps(ctx, PROF_TRSV, double(rows) * nb * sizeof(Cx<T>) * nbr);
        hipLaunchKernelGGL((trsv_step_kernel<T, true>), dim3(gx, nbr), dim3(256), 0, ctx->stream,
                           LU, lda, strideA, n, k0, nb, linv, inv_stride, kb, Y, B, ldb, strideB,
                           nrhs);
    }
    // Y <- B, backward: B <- U^{-1} Y
    hipLaunchKernelGGL(copy_kernel<T>, vgrid, dim3(256), 0, ctx->stream, B, ldb, strideB, Y, n,
                       ystride, n, nrhs);
    for (int64_t kb = nblk - 1; kb >= 0; --kb) {
        const int64_t k0 = kb * NB;
        const int nb = (int)std::min<int64_t>(NB, n - k0);
        const int64_t rows = k0;
        const unsigned gx = (unsigned)std::max<int64_t>(1, (rows + 255) / 256);
        ProfScope ps(ctx, PROF_TRSV, double(rows) * nb * sizeof(Cx<T>) * nbr);
        hipLaunchKernelGGL((trsv_step_kernel<T, false>), dim3(gx, nbr), dim3(256), 0,
                           ctx->stream, LU, lda, strideA, n, k0, nb, uinv, inv_stride, kb, Y, B,
                           ldb, strideB, nrhs);
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

}  // namespace

int lu_factor(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, void* d_A, int64_t lda,
              int64_t strideA, int32_t* d_ipiv) {
    CCAB_CHECK(n > 0 && batch > 0 && lda >= n && (batch == 1 || strideA >= lda * n),
               "invalid LU dimensions");
    CCAB_CHECK(n < (int64_t(1) << 31), "n too large for 32-bit pivots");
    if (prec == CCAB_C128)
        return lu_factor_t<double>(ctx, n, batch, static_cast<c128*>(d_A), lda, strideA, d_ipiv);
    if (prec == CCAB_C64)
        return lu_factor_t<float>(ctx, n, batch, static_cast<c64*>(d_A), lda, strideA, d_ipiv);
    return ctx_fail(ctx, CCAB_EINVAL, "unknown precision %d", prec);
}

int lu_solve(ccab_ctx* ctx, int prec, int64_t n, int64_t batch, const void* d_LU, int64_t lda,
             int64_t strideA, const int32_t* d_ipiv, void* d_B, int64_t ldb, int64_t strideB,
             int64_t nrhs) {
    CCAB_CHECK(n > 0 && batch > 0 && nrhs > 0 && ldb >= n, "invalid LU solve dimensions");
    if (prec == CCAB_C128)
        return lu_solve_t<double>(ctx, n, batch, static_cast<const c128*>(d_LU), lda, strideA,
                                  d_ipiv, static_cast<c128*>(d_B), ldb, strideB, nrhs);
    if (prec == CCAB_C64)
        return lu_solve_t<float>(ctx, n, batch, static_cast<const c64*>(d_LU), lda, strideA,
                                 d_ipiv, static_cast<c64*>(d_B), ldb, strideB, nrhs);
    return ctx_fail(ctx, CCAB_EINVAL, "unknown precision %d", prec);
}


// exposed for the GEMM unit test entry point
int cgemm_update(ccab_ctx* ctx, int prec, int64_t M, int64_t N, int64_t K, const void* A,
                 int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc) {
    if (prec == CCAB_C128)
        return gemm_launch<double, 0>(ctx, M, N, K, static_cast<const c128*>(A), lda, 0,
                                      static_cast<const c128*>(B), ldb, 0,
                                      static_cast<c128*>(C), ldc, 0, 1);
    return gemm_launch<float, 0>(ctx, M, N, K, static_cast<const c64*>(A), lda, 0,
                                 static_cast<const c64*>(B), ldb, 0, static_cast<c64*>(C), ldc,
                                 0, 1);
}

bool lu_coop_fits() {
    auto fits = [](const void* f) {
        int nb = 0;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kCoopThreads, 0) ==
                   hipSuccess && nb >= 1;
    };
    return fits(reinterpret_cast<const void*>(&coop_slab_kernel<double, coop_width<double, 1>(), 1>)) &&
           fits(reinterpret_cast<const void*>(&coop_slab_kernel<double, coop_width<double, 2>(), 2>)) &&
           fits(reinterpret_cast<const void*>(&coop_slab_kernel<float, coop_width<float, 1>(), 1>)) &&
           fits(reinterpret_cast<const void*>(&coop_slab_kernel<float, coop_width<float, 2>(), 2>));
}

void blas_release(ccab_ctx* ctx) {
    if (ctx->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
    ctx->blas = nullptr;
}
