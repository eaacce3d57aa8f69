// Hierarchical matrix of the BEM operator on gfx950: cluster / block trees
// (cluster.rs, block.rs; host C++), dense near-field blocks and batched ACA
// low-rank blocks (aca.rs; one workgroup per admissible block), recompression
// (QR + SVD per block, aca.rs:137-179), the H matrix-vector product
// (h_matrix.rs:218-234) and GMRES over it (solve.rs:74-150).
//
// Block storage in HBM: row / column index lists back to back; dense blocks
// column-major back to back (c128); low-rank blocks as U (nr x r) and V
// (nc x r) with A ~ U V^T (non-conjugated, as the reference's ACA::gemv).
// The matvec (no atomics, fixed summation order): t_b = V_b^T x for every
// low-rank block; then one wave per 64-row tile of every row cluster adds,
// over the cluster's blocks (dense first, then low-rank, in block order),
// D_b x_cols and U_b t_b into one slot per (cluster, row); then every row sums
// the slots of the clusters that contain it (about one per tree level).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <vector>

#include "ccab_internal.h"
#include "greens.h"

namespace {

constexpr int RMAX = 64;        // rank capacity of a low-rank block (ACA terms before recompression)
constexpr int kAcaThreads = 256;
constexpr int kRngTable = 1024; // ChaCha12 words available to one block's draws

struct HBlkDev {
    int64_t r_off, nr, c_off, nc;  // into rows[] / cols[]
    int64_t u_off, v_off;          // into U, V (during the ACA: the capacity buffers, nr x cap
                                   // and nc x cap; after recompression: the packed factors)
    int64_t d_off;                 // into D (dense blocks)
    int64_t p_off;                 // low-rank: offset of V^T x in the matvec's t buffer
    int64_t w_off;                 // ACA workspace (vectors), in c128 units
    int64_t f_off;                 // ACA exclusion flags, in bytes
    int64_t u2_off, v2_off;        // recompressed U, V (packed to the ACA rank)
    int64_t clo;                   // first position of the column cluster in the
                                   // depth-first order of its tree (the matvec's x)
    int64_t s_off;                 // low-rank: U in the product's storage (the blocks of a
                                   // segment one after another)
    int64_t ts_off;                // low-rank: its summed V^T x in the matvec's t buffer
    int32_t adm, rank;
    int32_t status;                // ACA: 0 ok, 1 rank capacity reached, 2 rng exhausted
    int32_t vtc;                   // matvec: column chunks of V^T x (partials at p_off + c
                                   // rank when there are several, else at ts_off)
    double norm;                   // Frobenius norm estimate / dense Frobenius norm
};

// ---------------------------------------------------------------------------
// rand 0.9 StdRng words (ChaCha12, seed_from_u64 via PCG32), host side: the
// reference seeds every ACA block with 10 (aca.rs:64), so all blocks consume
// the same word stream; it is tabulated once and read on the device.
// ---------------------------------------------------------------------------
inline uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

void chacha12_words(uint64_t seed, uint32_t* out, int nwords) {
    uint32_t key[8];
    uint64_t st = seed;
    for (int c = 0; c < 8; ++c) {
        st = st * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = uint32_t(((st >> 18) ^ st) >> 27);
        const uint32_t rot = uint32_t(st >> 59);
        key[c] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    for (int blk = 0; blk * 16 < nwords; ++blk) {
        uint32_t s0[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                           key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                           uint32_t(blk), 0u, 0u, 0u};
        uint32_t x[16];
        std::memcpy(x, s0, sizeof(x));
        auto qr = [&](int a, int b, int c, int d) {
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 16);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 12);
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 8);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 7);
        };
        for (int r = 0; r < 12; r += 2) {
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
        }
        for (int i = 0; i < 16 && blk * 16 + i < nwords; ++i) out[blk * 16 + i] = x[i] + s0[i];
    }
}

// ---------------------------------------------------------------------------
// Host cluster tree (cluster.rs:35-110) and block tree (block.rs:38-96)
// ---------------------------------------------------------------------------
struct HNode {
    double lo[3], hi[3], diam;
    std::vector<int64_t> idx;
    int son[2] = {-1, -1};
};

int build_cluster(std::vector<HNode>& T, const double* xyz, int64_t stride,
                  std::vector<int64_t>&& idx, int64_t leaf) {
    const int id = (int)T.size();
    T.emplace_back();
    {
        HNode& c = T.back();
        c.idx = std::move(idx);
        for (int d = 0; d < 3; ++d) {
            c.lo[d] = INFINITY;
            c.hi[d] = -INFINITY;
        }
        for (int64_t t : c.idx)
            for (int d = 0; d < 3; ++d) {
                c.lo[d] = std::fmin(xyz[stride * t + d], c.lo[d]);
                c.hi[d] = std::fmax(xyz[stride * t + d], c.hi[d]);
            }
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += (c.hi[d] - c.lo[d]) * (c.hi[d] - c.lo[d]);
        c.diam = std::sqrt(s);
    }
    if ((int64_t)T[id].idx.size() <= leaf) return id;
    int jmax = 0;
    double diff = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double e = T[id].hi[d] - T[id].lo[d];
        if (e > diff) {
            jmax = d;
            diff = e;
        }
    }
    const double gamma = 0.5 * (T[id].lo[jmax] + T[id].hi[jmax]);
    std::vector<int64_t> a, b;
    for (int64_t t : T[id].idx) (xyz[stride * t + jmax] <= gamma ? a : b).push_back(t);
    if (a.empty() || b.empty()) return id;  // coincident points cannot be split
    const int s0 = build_cluster(T, xyz, stride, std::move(a), leaf);
    const int s1 = build_cluster(T, xyz, stride, std::move(b), leaf);
    T[id].son[0] = s0;
    T[id].son[1] = s1;
    return id;
}

double cluster_dist(const HNode& a, const HNode& b) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double t1 = std::fmax(0.0, a.lo[d] - b.hi[d]);
        const double t2 = std::fmax(0.0, b.lo[d] - a.hi[d]);
        s += t1 * t1;
        s += t2 * t2;
    }
    return std::sqrt(s);
}

struct HostBlock {
    int r, c;
    bool adm;
};

void build_blocks(const std::vector<HNode>& T, int r, int c, double eta,
                  std::vector<HostBlock>& out) {
    const bool adm = std::fmin(T[r].diam, T[c].diam) <= eta * cluster_dist(T[r], T[c]);
    const bool rl = T[r].son[0] < 0, cl = T[c].son[0] < 0;
    if (!(adm || rl || cl)) {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) build_blocks(T, T[r].son[a], T[c].son[b], eta, out);
    } else {
        out.push_back({r, c, adm});
    }
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double cabs(c128 z) { return hypot(z.re, z.im); }

// Entry sources of the blocks: the BEM side matrix a H + b G
// (get_surface_row_or_column, influence_matrix.rs:170-231), or an explicit
// column-major matrix in HBM (the closures of the reference's ACA and HMatrix
// tests, aca.rs:303-304, h_matrix.rs:282-298)
struct BemEntry {
    SysParams p;
    const double* soa;
    int64_t n;
    c128 fa, fb;
    __device__ __forceinline__ c128 operator()(int64_t i, int64_t j) const {
        return side_entry(p, soa_of(soa, n), i, j, fa, fb);
    }
};

struct MatEntry {
    const c128* A;
    int64_t lda;
    __device__ __forceinline__ c128 operator()(int64_t i, int64_t j) const { return A[i + j * lda]; }
};

// dense near-field blocks: D[i + j nr] = entry(rows[i], cols[j])
template <typename E>
__global__ void dense_block_kernel(E entry, const HBlkDev* __restrict__ blks,
                                   const int32_t* __restrict__ list, const int64_t* __restrict__ rows,
                                   const int64_t* __restrict__ cols, c128* __restrict__ D) {
    const HBlkDev bk = blks[list[blockIdx.x]];
    const int64_t cnt = bk.nr * bk.nc;
    for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cnt;
         e += int64_t(blockDim.x) * gridDim.y) {
        const int64_t i = e % bk.nr, j = e / bk.nr;
        D[bk.d_off + e] = entry(rows[bk.r_off + i], cols[bk.c_off + j]);
    }
}

// Frobenius norm of each dense block (h_matrix.rs:189-196 sums them)
__global__ void dense_norm_kernel(HBlkDev* blks, const int32_t* __restrict__ list,
                                  const c128* __restrict__ D) {
    __shared__ double red[4];
    HBlkDev& bk = blks[list[blockIdx.x]];
    const int64_t cnt = bk.nr * bk.nc;
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < cnt; e += 256) s += norm_sqr(D[bk.d_off + e]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) bk.norm = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// --- ACA (aca.rs:57-136), one workgroup of kAcaThreads per admissible block ---
template <typename E>
struct AcaCtx {
    E entry;
    const int64_t* rows;
    const int64_t* cols;
    int64_t nr, nc;
    c128* U;  // nr x RMAX
    c128* V;  // nc x RMAX
};

// r[j] = A(rows[i], cols[j]) - sum_k U[i,k] V[j,k]   (residual row)
template <typename E>
__device__ void aca_row(const AcaCtx<E>& a, int64_t i, int rank, c128* r) {
    const int64_t gi = a.rows[i];
    for (int64_t j = threadIdx.x; j < a.nc; j += kAcaThreads) {
        c128 v = a.entry(gi, a.cols[j]);
        for (int k = 0; k < rank; ++k) v -= a.U[i + k * a.nr] * a.V[j + k * a.nc];
        r[j] = v;
    }
}

// r[i] = A(rows[i], cols[j]) - sum_k V[j,k] U[i,k]   (residual column)
template <typename E>
__device__ void aca_col(const AcaCtx<E>& a, int64_t j, int rank, c128* r) {
    const int64_t gj = a.cols[j];
    for (int64_t i = threadIdx.x; i < a.nr; i += kAcaThreads) {
        c128 v = a.entry(a.rows[i], gj);
        for (int k = 0; k < rank; ++k) v -= a.V[j + k * a.nc] * a.U[i + k * a.nr];
        r[i] = v;
    }
}

// index of the max |r| outside the excluded set, first index on ties
// (aca.rs:256-268); len if all excluded
__device__ int64_t argmax_excl(const c128* r, int64_t len, const uint8_t* excl, double* red_v,
                               int64_t* red_i) {
    double best = -1.0;
    int64_t bi = len;
    for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) {
        if (excl[t]) continue;
        const double v = cabs(r[t]);
        if (best < v) {
            best = v;
            bi = t;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int64_t oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
        }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red_v[threadIdx.x >> 6] = best;
        red_i[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    double bv = red_v[0];
    int64_t bx = red_i[0];
    for (int w = 1; w < kAcaThreads / 64; ++w)
        if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bx)) {
            bv = red_v[w];
            bx = red_i[w];
        }
    return bx;
}

// sum over the block of f(t) for t < len (complex, all threads get the sum)
template <typename F>
__device__ c128 block_csum(int64_t len, F f, double* red) {
    double re = 0.0, im = 0.0;
    for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) {
        const c128 v = f(t);
        re += v.re;
        im += v.im;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = re;
        red[2 * (threadIdx.x >> 6) + 1] = im;
    }
    __syncthreads();
    c128 s{0.0, 0.0};
    for (int w = 0; w < kAcaThreads / 64; ++w) s += c128{red[2 * w], red[2 * w + 1]};
    return s;
}

// Uniform::new_inclusive(0, hi).sample(rng) for usize in 32-bit mode
// (rand 0.9 UniformUsize): Lemire's widening multiply with rejection.
__device__ int64_t rng_uniform(const uint32_t* words, int* pos, uint32_t range, bool* ok) {
    const uint32_t thresh = (uint32_t)(0u - range) % range;
    while (*pos < kRngTable) {
        const uint64_t w = uint64_t(words[(*pos)++]) * range;
        if (uint32_t(w) >= thresh) return int64_t(w >> 32);
    }
    *ok = false;
    return 0;
}

template <typename E>
__global__ __launch_bounds__(kAcaThreads) void aca_kernel(E entry, HBlkDev* blks,
                                                          const int32_t* __restrict__ list,
                                                          const int64_t* __restrict__ rows,
                                                          const int64_t* __restrict__ cols,
                                                          c128* U, c128* V, c128* W, uint8_t* F,
                                                          const uint32_t* __restrict__ rng,
                                                          double tol) {
    __shared__ double red_v[kAcaThreads / 64];
    __shared__ int64_t red_i[kAcaThreads / 64];
    __shared__ double red[2 * kAcaThreads / 64];
    __shared__ double dots[RMAX][4][kAcaThreads / 64];
    __shared__ int64_t s_ref[2];
    __shared__ int s_status;
    HBlkDev& bk = blks[list[blockIdx.x]];
    AcaCtx<E> a{entry, rows + bk.r_off, cols + bk.c_off, bk.nr, bk.nc, U + bk.u_off, V + bk.v_off};
    const int64_t nr = bk.nr, nc = bk.nc;
    c128* riref = W + bk.w_off;       // nc
    c128* rjref = riref + nc;         // nr
    c128* ristar = rjref + nr;        // nc
    c128* rjstar = ristar + nc;       // nr
    uint8_t* exI = F + bk.f_off;      // nr
    uint8_t* exJ = exI + nr;          // nc
    for (int64_t t = threadIdx.x; t < nr + nc; t += kAcaThreads) exI[t] = 0;
    const int64_t nmin = nr < nc ? nr : nc;
    if (threadIdx.x == 0) {
        int pos = 0;
        bool ok = true;
        s_ref[0] = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
        s_ref[1] = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
        s_status = pos | (ok ? 0 : (1 << 30));
    }
    __syncthreads();
    int64_t iref = s_ref[0], jref = s_ref[1];
    int pos = s_status & 0xffff;
    int status = (s_status >> 30) ? 2 : 0;
    aca_row(a, iref, 0, riref);
    aca_col(a, jref, 0, rjref);
    __syncthreads();
    double total = 0.0;
    int rank = 0;
    for (int64_t k = 0; k < nmin; ++k) {
        if (rank == RMAX) {
            status = 1;
            break;
        }
        // pivots (aca.rs:83-101); an index equal to the length means every
        // candidate was NaN, which stops the block (stored dense)
        int64_t jstar = argmax_excl(riref, nc, exJ, red_v, red_i);
        int64_t istar = argmax_excl(rjref, nr, exI, red_v, red_i);
        if (jstar >= nc || istar >= nr) {
            status = 3;
            break;
        }
        if (cabs(rjref[istar]) > cabs(riref[jstar])) {
            aca_row(a, istar, rank, ristar);
            __syncthreads();
            jstar = argmax_excl(ristar, nc, exJ, red_v, red_i);
            if (jstar >= nc) {
                status = 3;
                break;
            }
            aca_col(a, jstar, rank, rjstar);
        } else {
            aca_col(a, jstar, rank, rjstar);
            __syncthreads();
            istar = argmax_excl(rjstar, nr, exI, red_v, red_i);
            if (istar >= nr) {
                status = 3;
                break;
            }
            aca_row(a, istar, rank, ristar);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            exI[istar] = 1;
            exJ[jstar] = 1;
        }
        // u = rjstar, v = ristar / ristar[jstar]
        const c128 factor = crecip(ristar[jstar]);
        for (int64_t t = threadIdx.x; t < nc; t += kAcaThreads) a.V[t + rank * nc] = ristar[t] * factor;
        for (int64_t t = threadIdx.x; t < nr; t += kAcaThreads) a.U[t + rank * nr] = rjstar[t];
        __syncthreads();
        // norm estimate (aca.rs:224-241): |u|^2 |v|^2 + sum_i 2 |u_i.u| |v_i.v|
        const c128* uk = a.U + rank * nr;
        const c128* vk = a.V + rank * nc;
        const c128 un = block_csum(nr, [&](int64_t t) { return c128{norm_sqr(uk[t]), 0.0}; }, red);
        const c128 vn = block_csum(nc, [&](int64_t t) { return c128{norm_sqr(vk[t]), 0.0}; }, red);
        const double nk2 = un.re * vn.re;
        total += nk2;
        // cross terms 2 |u_i . u_k| |v_i . v_k| for the previous terms: wave
        // partials of all pairs, one barrier, the same ordered sum in every thread
        for (int i = 0; i < rank; ++i) {
            double ur = 0, ui = 0, vr = 0, vi = 0;
            const c128* u0 = a.U + i * nr;
            const c128* v0 = a.V + i * nc;
            for (int64_t t = threadIdx.x; t < nr; t += kAcaThreads) {
                const c128 z = u0[t] * uk[t];
                ur += z.re;
                ui += z.im;
            }
            for (int64_t t = threadIdx.x; t < nc; t += kAcaThreads) {
                const c128 z = v0[t] * vk[t];
                vr += z.re;
                vi += z.im;
            }
            ur = wave_sum(ur);
            ui = wave_sum(ui);
            vr = wave_sum(vr);
            vi = wave_sum(vi);
            if ((threadIdx.x & 63) == 0) {
                const int w = threadIdx.x >> 6;
                dots[i][0][w] = ur;
                dots[i][1][w] = ui;
                dots[i][2][w] = vr;
                dots[i][3][w] = vi;
            }
        }
        __syncthreads();
        for (int i = 0; i < rank; ++i) {
            double d[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q = 0; q < 4; ++q)
                for (int w = 0; w < kAcaThreads / 64; ++w) d[q] += dots[i][q][w];
            total += 2.0 * hypot(d[0], d[1]) * hypot(d[2], d[3]);
        }
        ++rank;
        const double step = sqrt(nk2) / sqrt(total);
        if (step < tol / 50.0) break;
        // next reference row / column (aca.rs:116-132)
        if (iref == istar) {
            __syncthreads();
            if (threadIdx.x == 0) {
                bool ok = true;
                int64_t ii = iref;
                if (exI[ii]) {
                    ii = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
                    for (int64_t t = 0; t < nmin - 1; ++t) {
                        ii = (ii + 1) % nmin;
                        if (!exI[ii]) break;
                    }
                }
                s_ref[0] = ii;
                s_status = pos | (ok ? 0 : (1 << 30));
            }
            __syncthreads();
            iref = s_ref[0];
            pos = s_status & 0xffff;
            if (s_status >> 30) status = 2;
            aca_row(a, iref, rank, riref);
        } else {
            const c128 ui = a.U[iref + (rank - 1) * nr];
            for (int64_t t = threadIdx.x; t < nc; t += kAcaThreads) riref[t] -= ui * a.V[t + (rank - 1) * nc];
        }
        if (jref == jstar) {
            __syncthreads();
            if (threadIdx.x == 0) {
                bool ok = true;
                int64_t jj = jref;
                if (exJ[jj]) {
                    jj = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
                    for (int64_t t = 0; t < nmin - 1; ++t) {
                        jj = (jj + 1) % nmin;
                        if (!exJ[jj]) break;
                    }
                }
                s_ref[1] = jj;
                s_status = pos | (ok ? 0 : (1 << 30));
            }
            __syncthreads();
            jref = s_ref[1];
            pos = s_status & 0xffff;
            if (s_status >> 30) status = 2;
            aca_col(a, jref, rank, rjref);
        } else {
            const c128 vj = a.V[jref + (rank - 1) * nc];
            for (int64_t t = threadIdx.x; t < nr; t += kAcaThreads) rjref[t] -= vj * a.U[t + (rank - 1) * nr];
        }
        __syncthreads();
        if (status) break;
    }
    if (threadIdx.x == 0) {
        bk.rank = rank;
        bk.status = status;
        bk.norm = sqrt(total);
    }
}

// The ACA of aca_kernel for blocks of at most 64 PL rows and 64 PL columns
// (PL = 1, 2 or 4): one wavefront per block, four blocks per workgroup.
// Lane l holds entries l, l + 64, ... of the residual rows (column indices)
// and of the residual columns (row indices) in registers, the exclusion sets
// are bit masks and every reduction is a wave reduction, so the steps wait
// on no workgroup barrier and on no workspace round trip.  The steps, pivot
// rules (first index of the largest |r| on ties) and expressions are those
// of aca_kernel; only the sums of the norm estimate are formed in another
// order.
__device__ __forceinline__ c128 shfl_c(c128 v, int src) {
    return c128{__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
}

template <int PL>
struct LaneVec {
    c128 v[PL];
    // entry t (wave-uniform t < 64 PL) to every lane
    __device__ __forceinline__ c128 get(int64_t t) const {
        c128 r = v[0];
#pragma unroll
        for (int p = 1; p < PL; ++p)
            if (t >> 6 == p) r = v[p];
        return shfl_c(r, int(t & 63));
    }
};

template <int PL>
struct LaneMask {
    uint64_t w[PL];
    __device__ __forceinline__ bool test(int64_t t) const {
        uint64_t x = w[0];
#pragma unroll
        for (int p = 1; p < PL; ++p)
            if (t >> 6 == p) x = w[p];
        return (x >> (t & 63)) & 1;
    }
    __device__ __forceinline__ void set(int64_t t) {
#pragma unroll
        for (int p = 0; p < PL; ++p)
            if (t >> 6 == p) w[p] |= uint64_t(1) << (t & 63);
    }
};

template <typename E, int PL>
__global__ __launch_bounds__(256) void aca_wave_kernel(E entry, HBlkDev* blks,
                                                       const int32_t* __restrict__ list,
                                                       int64_t cnt,
                                                       const int64_t* __restrict__ rows,
                                                       const int64_t* __restrict__ cols,
                                                       c128* U, c128* V,
                                                       const uint32_t* __restrict__ rng,
                                                       double tol) {
    using Vec = LaneVec<PL>;
    const int lane = threadIdx.x & 63;
    const int64_t w = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (w >= cnt) return;  // the whole wavefront
    HBlkDev& bk = blks[list[w]];
    const int64_t nr = bk.nr, nc = bk.nc;
    const int64_t* R = rows + bk.r_off;
    const int64_t* C = cols + bk.c_off;
    c128* Ub = U + bk.u_off;  // nr x capacity
    c128* Vb = V + bk.v_off;  // nc x capacity
    int64_t gi[PL], gj[PL];
#pragma unroll
    for (int p = 0; p < PL; ++p) {
        const int64_t t = lane + 64 * p;
        gi[p] = t < nr ? R[t] : 0;
        gj[p] = t < nc ? C[t] : 0;
    }
    const int64_t nmin = nr < nc ? nr : nc;
    // residual row i over the columns, residual column j over the rows
    // (aca_row, aca_col)
    auto res_row = [&](int64_t i, int rank) {
        Vec r;
        const int64_t g = R[i];
#pragma unroll
        for (int p = 0; p < PL; ++p) {
            const int64_t t = lane + 64 * p;
            c128 v{0.0, 0.0};
            if (t < nc) {
                v = entry(g, gj[p]);
                for (int k = 0; k < rank; ++k) v -= Ub[i + k * nr] * Vb[t + k * nc];
            }
            r.v[p] = v;
        }
        return r;
    };
    auto res_col = [&](int64_t j, int rank) {
        Vec r;
        const int64_t g = C[j];
#pragma unroll
        for (int p = 0; p < PL; ++p) {
            const int64_t t = lane + 64 * p;
            c128 v{0.0, 0.0};
            if (t < nr) {
                v = entry(gi[p], g);
                for (int k = 0; k < rank; ++k) v -= Vb[j + k * nc] * Ub[t + k * nr];
            }
            r.v[p] = v;
        }
        return r;
    };
    // argmax_excl: len when every candidate is excluded or NaN
    auto argmax = [&](const Vec& r, int64_t len, const LaneMask<PL>& excl, double& bv) {
        double best = -1.0;
        int64_t bi = len;
#pragma unroll
        for (int p = 0; p < PL; ++p) {
            const int64_t t = lane + 64 * p;
            if (t < len && !((excl.w[p] >> lane) & 1)) {
                const double c = cabs(r.v[p]);
                if (best < c) {
                    best = c;
                    bi = t;
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(best, off, 64);
            const int64_t oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        bv = best;
        return bi;
    };
    int pos = 0;
    int status = 0;
    int64_t iref, jref;
    {
        bool ok = true;
        iref = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
        jref = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
        if (!ok) status = 2;
    }
    Vec riref = res_row(iref, 0), rjref = res_col(jref, 0);
    LaneMask<PL> exI, exJ;
#pragma unroll
    for (int p = 0; p < PL; ++p) exI.w[p] = exJ.w[p] = 0;
    double total = 0.0;
    int rank = 0;
    for (int64_t k = 0; k < nmin; ++k) {
        if (rank == RMAX) {
            status = 1;
            break;
        }
        double bvj, bvi;
        int64_t jstar = argmax(riref, nc, exJ, bvj);
        int64_t istar = argmax(rjref, nr, exI, bvi);
        if (jstar >= nc || istar >= nr) {
            status = 3;
            break;
        }
        Vec ristar, rjstar;
        if (bvi > bvj) {  // |rjref[istar]| > |riref[jstar]|
            ristar = res_row(istar, rank);
            jstar = argmax(ristar, nc, exJ, bvj);
            if (jstar >= nc) {
                status = 3;
                break;
            }
            rjstar = res_col(jstar, rank);
        } else {
            rjstar = res_col(jstar, rank);
            istar = argmax(rjstar, nr, exI, bvi);
            if (istar >= nr) {
                status = 3;
                break;
            }
            ristar = res_row(istar, rank);
        }
        exI.set(istar);
        exJ.set(jstar);
        // u = rjstar, v = ristar / ristar[jstar]
        const c128 factor = crecip(ristar.get(jstar));
        Vec vk;
        const Vec& uk = rjstar;
        double un = 0.0, vn = 0.0;
#pragma unroll
        for (int p = 0; p < PL; ++p) {
            const int64_t t = lane + 64 * p;
            vk.v[p] = ristar.v[p] * factor;
            if (t < nc) {
                Vb[t + rank * nc] = vk.v[p];
                vn += norm_sqr(vk.v[p]);
            }
            if (t < nr) {
                Ub[t + rank * nr] = uk.v[p];
                un += norm_sqr(uk.v[p]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // norm estimate (aca.rs:224-241): |u|^2 |v|^2 + sum_i 2 |u_i.u| |v_i.v|
        un = wave_sum(un);
        vn = wave_sum(vn);
        const double nk2 = un * vn;
        total += nk2;
        for (int i = 0; i < rank; ++i) {
            double ur = 0.0, ui = 0.0, vr = 0.0, vi = 0.0;
#pragma unroll
            for (int p = 0; p < PL; ++p) {
                const int64_t t = lane + 64 * p;
                if (t < nr) {
                    const c128 z = Ub[t + i * nr] * uk.v[p];
                    ur += z.re;
                    ui += z.im;
                }
                if (t < nc) {
                    const c128 z = Vb[t + i * nc] * vk.v[p];
                    vr += z.re;
                    vi += z.im;
                }
            }
            ur = wave_sum(ur);
            ui = wave_sum(ui);
            vr = wave_sum(vr);
            vi = wave_sum(vi);
            total += 2.0 * hypot(ur, ui) * hypot(vr, vi);
        }
        ++rank;
        const double step = sqrt(nk2) / sqrt(total);
        if (step < tol / 50.0) break;
        // next reference row / column (aca.rs:116-132)
        if (iref == istar) {
            bool ok = true;
            int64_t ii = iref;
            if (exI.test(ii)) {
                ii = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
                for (int64_t t = 0; t < nmin - 1; ++t) {
                    ii = (ii + 1) % nmin;
                    if (!exI.test(ii)) break;
                }
            }
            iref = ii;
            if (!ok) status = 2;
            riref = res_row(iref, rank);
        } else {
            const c128 ui = uk.get(iref);  // U[iref, rank - 1]
#pragma unroll
            for (int p = 0; p < PL; ++p) riref.v[p] -= ui * vk.v[p];
        }
        if (jref == jstar) {
            bool ok = true;
            int64_t jj = jref;
            if (exJ.test(jj)) {
                jj = rng_uniform(rng, &pos, uint32_t(nmin), &ok);
                for (int64_t t = 0; t < nmin - 1; ++t) {
                    jj = (jj + 1) % nmin;
                    if (!exJ.test(jj)) break;
                }
            }
            jref = jj;
            if (!ok) status = 2;
            rjref = res_col(jref, rank);
        } else {
            const c128 vj = vk.get(jref);  // V[jref, rank - 1]
#pragma unroll
            for (int p = 0; p < PL; ++p) rjref.v[p] -= vj * uk.v[p];
        }
        if (status) break;
    }
    if (lane == 0) {
        bk.rank = rank;
        bk.status = status;
        bk.norm = sqrt(total);
    }
}

// One-sided complex Jacobi SVD of the r x r matrix W (LDS, [row][col]),
// rotations accumulated in Vs (identity on entry), with the parallel
// (round-robin) ordering: the r columns (one padding column when r is odd)
// meet in r' - 1 rounds of r'/2 disjoint pairs, so NTHR / SW pairs rotate at
// once, SW lanes (one per row, SW >= r, a power of two <= 64) per pair, the
// norms and the inner product of a pair summed over its SW lanes.  The
// rotation of a pair is that of aca.rs:137-179's SVD in one-sided form
// (tau = (a_qq - a_pp) / 2|a_pq|); sweeps until every pair is below 1e-15
// relative (at most 60).  NTHR threads take part (tid < NTHR); with
// NTHR > 64 the caller's whole workgroup runs it (barriers between rounds).
template <int RM, int SW, int NTHR>
__device__ __forceinline__ void jacobi_rr(c128 (*Wm)[RM + 1], c128 (*Vs)[RM + 1], int r, int tid,
                                          double* red_off) {
    constexpr int NSLOT = NTHR / SW;
    const int slot = tid / SW, row = tid % SW;
    const int m = r + (r & 1);         // columns incl. the padding one
    const int npair = m / 2;
    auto sync = [&]() {
        if constexpr (NTHR > 64) __syncthreads();
        else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    };
    auto ssum = [](double v) {
#pragma unroll
        for (int off = SW / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    };
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int k = 0; k < m - 1; ++k) {
            for (int p0 = 0; p0 < npair; p0 += NSLOT) {
                const int pi = p0 + slot;
                // circle method: column 0 fixed, the others rotate by k
                int a = 0, b = 0;
                if (pi < npair) {
                    auto rot = [&](int j) { return 1 + (j + k) % (m - 1); };
                    a = pi == 0 ? 0 : rot(pi);
                    b = pi == 0 ? rot(0) : rot(m - 1 - pi);
                }
                const int p = a < b ? a : b, q = a < b ? b : a;
                const bool live = pi < npair && q < r;  // no padding column
                const bool mine = live && row < r;
                const c128 wp = mine ? Wm[row][p] : c128{0.0, 0.0};
                const c128 wq = mine ? Wm[row][q] : c128{0.0, 0.0};
                const double app = ssum(norm_sqr(wp));
                const double aqq = ssum(norm_sqr(wq));
                const c128 z = conj(wp) * wq;
                const c128 apq{ssum(z.re), ssum(z.im)};
                const double mag = hypot(apq.re, apq.im);
                if (live && mag != 0.0 && mag > 1e-15 * sqrt(app * aqq)) {
                    off = fmax(off, mag / sqrt(app * aqq));
                    const c128 ph{apq.re / mag, apq.im / mag};
                    const double tau = (aqq - app) / (2.0 * mag);
                    const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                    if (row < r) {
                        Wm[row][p] = wp * cs - (conj(ph) * wq) * sn;
                        Wm[row][q] = (ph * wp) * sn + wq * cs;
                        const c128 vp = Vs[row][p], vq = Vs[row][q];
                        Vs[row][p] = vp * cs - (conj(ph) * vq) * sn;
                        Vs[row][q] = (ph * vp) * sn + vq * cs;
                    }
                }
            }
            sync();  // the next round pairs the columns differently
        }
        // the largest off-diagonal measure of the sweep over all threads
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) off = fmax(off, __shfl_xor(off, o, 64));
        if constexpr (NTHR > 64) {
            if ((tid & 63) == 0) red_off[tid >> 6] = off;
            __syncthreads();
            off = 0.0;
            for (int w = 0; w < NTHR / 64; ++w) off = fmax(off, red_off[w]);
            __syncthreads();
        }
        if (off < 1e-15) break;
    }
}

// --- recompression (aca.rs:137-179): QR(U), QR(V) by CGS2, SVD of R_u R_v^T
//     by one-sided Jacobi (one wavefront, LDS), truncation, new U and V -------
template <int RM>
__global__ __launch_bounds__(kAcaThreads) void recompress_kernel(HBlkDev* blks,
                                                                 const int32_t* __restrict__ list,
                                                                 c128* U, c128* V, c128* U2,
                                                                 c128* V2, c128* Rws, double tol) {
    __shared__ c128 Wm[RM][RM + 1];  // [row][col]
    __shared__ c128 Vs[RM][RM + 1];
    __shared__ double red[2 * kAcaThreads / 64];
    __shared__ double sv[RM];
    __shared__ int ord[RM];
    __shared__ double s_tmp[RM][2][kAcaThreads / 64];
    HBlkDev& bk = blks[list[blockIdx.x]];
    if (bk.status != 0) return;
    const int r = bk.rank;  // <= RM (the host buckets blocks by rank)
    const int64_t nr = bk.nr, nc = bk.nc;
    c128* Ru = Rws + int64_t(blockIdx.x) * 2 * RM * RM;  // [c][i] = R[i][c]
    c128* Rv = Ru + RM * RM;
    for (int side = 0; side < 2; ++side) {
        c128* Q = side == 0 ? U + bk.u_off : V + bk.v_off;
        const int64_t len = side == 0 ? nr : nc;
        c128* R = side == 0 ? Ru : Rv;
        for (int e = threadIdx.x; e < RM * RM; e += kAcaThreads) R[e] = c128{0.0, 0.0};
        __syncthreads();
        for (int c = 0; c < r; ++c) {
            c128* qc = Q + c * len;
            for (int pass = 0; pass < 2; ++pass) {
                // h_i = q_i^H q_c for all i < c (wave partials, one barrier)
                for (int i = 0; i < c; ++i) {
                    const c128* qi = Q + i * len;
                    double hr = 0, hi = 0;
                    for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) {
                        const c128 z = conj(qi[t]) * qc[t];
                        hr += z.re;
                        hi += z.im;
                    }
                    hr = wave_sum(hr);
                    hi = wave_sum(hi);
                    if ((threadIdx.x & 63) == 0) {
                        s_tmp[i][0][threadIdx.x >> 6] = hr;
                        s_tmp[i][1][threadIdx.x >> 6] = hi;
                    }
                }
                __syncthreads();
                for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) {
                    c128 v = qc[t];
                    for (int i = 0; i < c; ++i) {
                        c128 h{0.0, 0.0};
                        for (int w = 0; w < kAcaThreads / 64; ++w)
                            h += c128{s_tmp[i][0][w], s_tmp[i][1][w]};
                        v -= h * Q[i * len + t];
                    }
                    qc[t] = v;
                }
                if (threadIdx.x < c) {
                    c128 h{0.0, 0.0};
                    for (int w = 0; w < kAcaThreads / 64; ++w)
                        h += c128{s_tmp[threadIdx.x][0][w], s_tmp[threadIdx.x][1][w]};
                    R[c * RM + threadIdx.x] += h;
                }
                __syncthreads();
            }
            double nn = 0.0;
            for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) nn += norm_sqr(qc[t]);
            nn = wave_sum(nn);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nn;
            __syncthreads();
            const double rn = sqrt(red[0] + red[1] + red[2] + red[3]);
            const double inv = rn > 1e-300 ? 1.0 / rn : 0.0;
            for (int64_t t = threadIdx.x; t < len; t += kAcaThreads) qc[t] = qc[t] * inv;
            if (threadIdx.x == 0) R[c * RM + c] = c128{rn, 0.0};
            __syncthreads();
        }
    }
    // W = R_u R_v^T (r x r)
    for (int e = threadIdx.x; e < r * r; e += kAcaThreads) {
        const int i = e % r, c = e / r;
        c128 sum{0.0, 0.0};
        for (int t = 0; t < r; ++t) sum += Ru[t * RM + i] * Rv[t * RM + c];
        Wm[i][c] = sum;
        Vs[i][c] = c128{i == c ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    // one-sided Jacobi on the columns of Wm, all four wavefronts
    jacobi_rr<RM, (RM <= 16 ? 16 : RM <= 32 ? 32 : 64), kAcaThreads>(Wm, Vs, r, threadIdx.x, red);
    // singular values = column norms of the rotated W
    if (threadIdx.x < r) {
        double sq = 0.0;
        for (int i = 0; i < r; ++i) sq += norm_sqr(Wm[i][threadIdx.x]);
        sv[threadIdx.x] = sqrt(sq);
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // descending order, stable
        for (int c = 0; c < r; ++c) ord[c] = c;
        for (int x = 1; x < r; ++x) {
            const int t = ord[x];
            int y = x - 1;
            while (y >= 0 && sv[ord[y]] < sv[t]) {
                ord[y + 1] = ord[y];
                --y;
            }
            ord[y + 1] = t;
        }
    }
    __syncthreads();
    // truncation: first sigma_i < tol * sum_{<=i} sigma, keep >= 1
    int index = 0;
    {
        double sum = 0.0;
        for (int t = 0; t < r; ++t) {
            sum += sv[ord[t]];
            if (sv[ord[t]] < tol * sum) break;
            ++index;
        }
        index = index + 1 < r ? index + 1 : r;
        if (index < 1) index = 1;
    }
    // with W Y = X Sigma (Y = the accumulated rotations): U2[:, t] = Q_u (W Y)[:, o_t]
    // = sigma Q_u x, V2[:, t] = Q_v conj(Y[:, o_t]), so U2 V2^T = Q_u W Q_v^T truncated
    const c128* Qu = U + bk.u_off;
    const c128* Qv = V + bk.v_off;
    for (int t = 0; t < index; ++t) {
        const int o = ord[t];
        for (int64_t i = threadIdx.x; i < nr; i += kAcaThreads) {
            c128 acc{0.0, 0.0};
            for (int c = 0; c < r; ++c) acc += Qu[i + c * nr] * Wm[c][o];
            U2[bk.u2_off + i + t * nr] = acc;
        }
        for (int64_t j = threadIdx.x; j < nc; j += kAcaThreads) {
            c128 acc{0.0, 0.0};
            for (int c = 0; c < r; ++c) acc += Qv[j + c * nc] * conj(Vs[c][o]);
            V2[bk.v2_off + j + t * nc] = acc;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) bk.rank = index;
}

// The same recompression with one wavefront per block and four blocks per
// workgroup, for blocks of rank <= RM <= 16 (the bulk of them: a few hundred
// rows per factor, where a workgroup per block waits mostly at its barriers):
// every reduction a wave_sum, no workgroup barrier.  R factors in the same
// scratch, one slot per block of the list.
template <int RM>
__global__ __launch_bounds__(256) void recompress_wave_kernel(HBlkDev* blks,
                                                              const int32_t* __restrict__ list,
                                                              int64_t cnt, c128* U, c128* V,
                                                              c128* U2, c128* V2, c128* Rws,
                                                              double tol) {
    constexpr int NWV = 4;
    __shared__ c128 Wm_all[NWV][RM][RM + 1];
    __shared__ c128 Vs_all[NWV][RM][RM + 1];
    __shared__ double sv_all[NWV][RM];
    __shared__ int ord_all[NWV][RM];
    __shared__ c128 hs_all[NWV][RM];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t idx = int64_t(blockIdx.x) * NWV + wv;
    if (idx >= cnt) return;  // (no workgroup barrier below)
    HBlkDev& bk = blks[list[idx]];
    if (bk.status != 0) return;
    c128(&Wm)[RM][RM + 1] = Wm_all[wv];
    c128(&Vs)[RM][RM + 1] = Vs_all[wv];
    double* sv = sv_all[wv];
    int* ord = ord_all[wv];
    c128* hs = hs_all[wv];
    const int r = bk.rank;  // <= RM (bucketed by the host)
    const int64_t nr = bk.nr, nc = bk.nc;
    c128* Ru = Rws + idx * 2 * RM * RM;  // [c][i] = R[i][c]
    c128* Rv = Ru + RM * RM;
    for (int side = 0; side < 2; ++side) {
        c128* Q = side == 0 ? U + bk.u_off : V + bk.v_off;
        const int64_t len = side == 0 ? nr : nc;
        c128* R = side == 0 ? Ru : Rv;
        for (int e = lane; e < RM * RM; e += 64) R[e] = c128{0.0, 0.0};
        for (int c = 0; c < r; ++c) {
            c128* qc = Q + c * len;
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 0; i < c; ++i) {  // hs_i = q_i^H q_c
                    const c128* qi = Q + i * len;
                    double hr = 0, hi = 0;
                    for (int64_t t = lane; t < len; t += 64) {
                        const c128 z = conj(qi[t]) * qc[t];
                        hr += z.re;
                        hi += z.im;
                    }
                    hr = wave_sum(hr);
                    hi = wave_sum(hi);
                    if (lane == 0) hs[i] = c128{hr, hi};
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                for (int64_t t = lane; t < len; t += 64) {
                    c128 v = qc[t];
                    for (int i = 0; i < c; ++i) v -= hs[i] * Q[i * len + t];
                    qc[t] = v;
                }
                if (lane < c) R[c * RM + lane] += hs[lane];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
            double nn = 0.0;
            for (int64_t t = lane; t < len; t += 64) nn += norm_sqr(qc[t]);
            const double rn = sqrt(wave_sum(nn));
            const double inv = rn > 1e-300 ? 1.0 / rn : 0.0;
            for (int64_t t = lane; t < len; t += 64) qc[t] = qc[t] * inv;
            if (lane == 0) R[c * RM + c] = c128{rn, 0.0};
        }
    }
    // R written by single lanes is read by all lanes below
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // W = R_u R_v^T (r x r)
    for (int e = lane; e < r * r; e += 64) {
        const int i = e % r, c = e / r;
        c128 sum{0.0, 0.0};
        for (int t = 0; t < r; ++t) sum += Ru[t * RM + i] * Rv[t * RM + c];
        Wm[i][c] = sum;
        Vs[i][c] = c128{i == c ? 1.0 : 0.0, 0.0};
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // one-sided Jacobi on the columns of Wm, four column pairs at once
    jacobi_rr<RM, 16, 64>(Wm, Vs, r, lane, nullptr);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (lane < r) {  // singular values = column norms of the rotated W
        double sq = 0.0;
        for (int i = 0; i < r; ++i) sq += norm_sqr(Wm[i][lane]);
        sv[lane] = sqrt(sq);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (lane == 0) {  // descending order, stable
        for (int c = 0; c < r; ++c) ord[c] = c;
        for (int x = 1; x < r; ++x) {
            const int t = ord[x];
            int y = x - 1;
            while (y >= 0 && sv[ord[y]] < sv[t]) {
                ord[y + 1] = ord[y];
                --y;
            }
            ord[y + 1] = t;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // truncation: first sigma_i < tol * sum_{<=i} sigma, keep >= 1
    int index = 0;
    {
        double sum = 0.0;
        for (int t = 0; t < r; ++t) {
            sum += sv[ord[t]];
            if (sv[ord[t]] < tol * sum) break;
            ++index;
        }
        index = index + 1 < r ? index + 1 : r;
        if (index < 1) index = 1;
    }
    const c128* Qu = U + bk.u_off;
    const c128* Qv = V + bk.v_off;
    for (int t = 0; t < index; ++t) {
        const int o = ord[t];
        for (int64_t i = lane; i < nr; i += 64) {
            c128 acc{0.0, 0.0};
            for (int c = 0; c < r; ++c) acc += Qu[i + c * nr] * Wm[c][o];
            U2[bk.u2_off + i + t * nr] = acc;
        }
        for (int64_t j = lane; j < nc; j += 64) {
            c128 acc{0.0, 0.0};
            for (int c = 0; c < r; ++c) acc += Qv[j + c * nc] * conj(Vs[c][o]);
            V2[bk.v2_off + j + t * nc] = acc;
        }
    }
    if (lane == 0) bk.rank = index;
}

// norm estimate of the recompressed factors ab initio (aca.rs:212-222)
__global__ __launch_bounds__(kAcaThreads) void aca_norm_kernel(HBlkDev* blks,
                                                               const int32_t* __restrict__ list,
                                                               const c128* __restrict__ U,
                                                               const c128* __restrict__ V) {
    __shared__ double red[2 * kAcaThreads / 64];
    HBlkDev& bk = blks[list[blockIdx.x]];
    if (bk.status != 0) return;  // stored dense instead
    const int r = bk.rank;
    const int64_t nr = bk.nr, nc = bk.nc;
    const c128* Ub = U + bk.u_off;
    const c128* Vb = V + bk.v_off;
    double tot = 0.0;
    for (int i = 0; i < r; ++i) {
        const c128 un = block_csum(nr, [&](int64_t t) { return c128{norm_sqr(Ub[t + i * nr]), 0.0}; }, red);
        const c128 vn = block_csum(nc, [&](int64_t t) { return c128{norm_sqr(Vb[t + i * nc]), 0.0}; }, red);
        tot += un.re * vn.re;
        for (int j = 0; j < i; ++j) {
            const c128 du = block_csum(nr, [&](int64_t t) { return Ub[t + i * nr] * Ub[t + j * nr]; }, red);
            const c128 dv = block_csum(nc, [&](int64_t t) { return Vb[t + i * nc] * Vb[t + j * nc]; }, red);
            tot += 2.0 * hypot(du.re, du.im) * hypot(dv.re, dv.im);
        }
    }
    if (threadIdx.x == 0) bk.norm = sqrt(tot);
}

// --- matvec ------------------------------------------------------------------
// stored block entries (c128, or c64 with FP32 storage) as FP64
__device__ __forceinline__ c128 ldx(const c128* p) { return *p; }
__device__ __forceinline__ c128 ldx(const c64* p) {
    const c64 v = *p;
    return c128{double(v.re), double(v.im)};
}

// block entries streamed once per product: non-temporal loads (NT) keep x,
// t and the slots in the caches
typedef double hf64x2_t __attribute__((ext_vector_type(2)));
typedef float hf32x2_t __attribute__((ext_vector_type(2)));
typedef float hf32x4_t __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ c128 ldm(const c128* p) {
    if constexpr (NT) {
        const hf64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const hf64x2_t*>(p));
        return c128{v.x, v.y};
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ c128 ldm(const c64* p) {
    if constexpr (NT) {
        const hf32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const hf32x2_t*>(p));
        return c128{double(v.x), double(v.y)};
    } else {
        return ldx(p);
    }
}
// two consecutive complex-FP32 entries (16-byte aligned)
template <bool NT>
__device__ __forceinline__ float4 ldm2(const c64* p) {
    if constexpr (NT) {
        const hf32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const hf32x4_t*>(p));
        return float4{v.x, v.y, v.z, v.w};
    } else {
        return *reinterpret_cast<const float4*>(p);
    }
}

// The matvec's two passes work on balanced items: V^T x per (low-rank block,
// chunk of at most kVtChunk columns), and per (segment of a row cluster's
// blocks worth at most kNodeBudget columns, 64-row tile); each item writes
// its own partial (t, or a slot), so sums keep a fixed order.  The segment
// budget (CCAB_HMV_NODE_BUDGET columns, 0 = a whole cluster) and the item
// order (CCAB_HMV_SORT: by decreasing work, or tree order) are set when the
// H matrix is built; a launch gives each wave one item, or with a cap on
// resident waves (CCAB_HMV_WAVES_PER_CU) wave w takes items w, w + W, ...
// (persistent waves).
constexpr int64_t kVtChunk = 512;

// A V^T x item with everything its wave needs (one load, no chain through
// the block descriptors): the chunk's first V entry and x entry, its t slot
struct VtItem {
    int64_t v;    // V offset of the chunk's first column (rank 0)
    int64_t x;    // xp offset of the chunk's first column
    int64_t t;    // t offset of the chunk's rank partials
    int64_t ld;   // rows of V (the block's column count)
    int32_t len;  // columns in the chunk
    int32_t rank;
};

// The wave's sums of N complex accumulators (N = 2, 4 or 8) by recursive
// halving: the exchanges between lanes 32, 16, ... apart (log2 N of them)
// each pass on half of what a lane holds, so that afterwards lane l holds
// accumulator wave_sum_index<N>(l) summed over the 64 / N lanes that share
// its low bits; the remaining xor exchanges sum those.  For N = 8: 40 lane
// exchanges of 32 bits instead of 192 for eight wave_sums of the real and
// imaginary parts.  All 64 / N lanes of an index hold the same sum.
__device__ __forceinline__ c128 shfl_xor_c(c128 v, int off) {
    return c128{__shfl_xor(v.re, off, 64), __shfl_xor(v.im, off, 64)};
}
template <int N>
__device__ __forceinline__ int wave_sum_index(int lane) {
    int idx = 0;
#pragma unroll
    for (int off = 32, m = N; m > 1; m >>= 1, off >>= 1) idx = 2 * idx + ((lane & off) ? 1 : 0);
    return idx;
}
template <int N>
__device__ __forceinline__ c128 wave_sum_many(const c128 (&a)[N], int lane) {
    c128 cur[N];
#pragma unroll
    for (int u = 0; u < N; ++u) cur[u] = a[u];
    int off = 32;
#pragma unroll
    for (int m = N; m > 1; m >>= 1, off >>= 1) {
        const bool b = lane & off;
#pragma unroll
        for (int u = 0; u < m / 2; ++u)
            cur[u] = (b ? cur[u + m / 2] : cur[u]) + shfl_xor_c(b ? cur[u] : cur[u + m / 2], off);
    }
#pragma unroll
    for (; off > 0; off >>= 1) cur[0] += shfl_xor_c(cur[0], off);
    return cur[0];
}

// f(std::integral_constant<int, n>) for a wave-uniform count 1 <= n <= MAXN:
// the loops over n loads become straight-line code without branches
template <int K, int MAXN, typename F>
__device__ __forceinline__ void with_count(int n, F& f) {
    if constexpr (K <= MAXN) {
        if (n == K) f(std::integral_constant<int, K>{});
        else with_count<K + 1, MAXN>(n, f);
    }
}
template <int NU>
constexpr int pow2_ceil() { return NU <= 1 ? 1 : NU <= 2 ? 2 : NU <= 4 ? 4 : 8; }

// t[item.t + k] = sum over the chunk's columns j of V[item.v + j + k ld] xp[item.x + j]
// Up to RB ranks per pass (8 for c128 storage; 4 or 8 for complex FP32, whose
// paired loads carry two columns each): the pass over NU ranks is compiled
// for each NU, so a column's NU loads are issued together without branches,
// and its NU sums are formed by one wave_sum_many.
template <typename T, bool NT, int RB>
__global__ __launch_bounds__(256) void hmv_vt_kernel(const VtItem* __restrict__ items, int64_t cnt,
                                                     const T* __restrict__ V,
                                                     const c128* __restrict__ xp, c128* t,
                                                     int paired) {
    const int lane = threadIdx.x & 63;
    const int wpg = blockDim.x >> 6;
    const int64_t W = int64_t(gridDim.x) * wpg;
    for (int64_t it = int64_t(blockIdx.x) * wpg + (threadIdx.x >> 6); it < cnt; it += W) {
        const VtItem q = items[it];
        const int r = q.rank;
        const T* Vb = V + q.v;
        const c128* xc = xp + q.x;  // the column cluster's x, contiguous
        int k0 = 0;
        // complex FP32 with even offsets and even column counts: columns
        // (j, j + 1) of a rank are one aligned 16-byte load, 128 columns per
        // pass
        const bool pair = sizeof(T) == 8 && paired && (q.ld & 1) == 0 && (q.v & 1) == 0 &&
                          (q.len & 1) == 0;
        auto pass = [&](auto nuc) {
            constexpr int NU = decltype(nuc)::value, N = pow2_ceil<NU>();
            c128 acc[N] = {};
            if constexpr (sizeof(T) == 8) {
                if (pair) {
                    for (int j = 2 * lane; j < q.len; j += 128) {
                        const c128 x0 = xc[j], x1 = xc[j + 1];
                        float4 v[NU];
#pragma unroll
                        for (int u = 0; u < NU; ++u) v[u] = ldm2<NT>(Vb + j + int64_t(k0 + u) * q.ld);
#pragma unroll
                        for (int u = 0; u < NU; ++u)
                            acc[u] += c128{v[u].x, v[u].y} * x0 + c128{v[u].z, v[u].w} * x1;
                    }
                }
            }
            if (!pair) {
                for (int j = lane; j < q.len; j += 64) {
                    const c128 xj = xc[j];
                    c128 v[NU];
#pragma unroll
                    for (int u = 0; u < NU; ++u) v[u] = ldm<NT>(Vb + j + int64_t(k0 + u) * q.ld);
#pragma unroll
                    for (int u = 0; u < NU; ++u) acc[u] += v[u] * xj;
                }
            }
            const c128 s = wave_sum_many<N>(acc, lane);
            const int su = wave_sum_index<N>(lane);
            if ((lane & (64 / N - 1)) == 0 && su < NU) t[q.t + k0 + su] = s;
        };
        for (; k0 < r; k0 += RB) with_count<1, RB>(min(r - k0, RB), pass);
    }
}

// One product work entry per dense block and per low-rank run of a segment
// (a segment lists its dense blocks, then at most one run): the segment's
// entries are read with one vector load (a lane per entry) and broadcast,
// instead of a dependent chain of scalar loads per block.  A run is the
// segment's low-rank blocks stored one after another: U of all of them one
// nr x R column-major matrix, their V^T x one vector of R values in t, so
// the product reads it as it reads a dense block (16 columns in flight per
// lane, no per-block rank loop).
struct HmvWork {
    int64_t off;   // dense: d_off (D); run: its first U entry (U)
    int64_t xoff;  // dense: clo (xp); run: its first value in t
    int32_t cnt;   // dense: columns; run: the sum of its blocks' ranks
    int32_t kind;  // 0: dense; 1: low-rank run
};

// The rank partials of a low-rank block's V^T x chunks summed in chunk order
// into the block's place in t (one thread per rank of a block with more than
// one chunk), so that the product reads one t per rank: summed inside the
// product, every 64-row tile of the block repeated the sum as a chain of
// dependent scalar loads.
struct TsumItem {
    int64_t p;       // t offset of the rank's first-chunk partial
    int64_t dst;     // t offset of the summed value
    int32_t stride;  // the block's rank (distance between its chunks' partials)
    int32_t cnt;     // chunks
};
__global__ void hmv_tsum_kernel(const TsumItem* __restrict__ items, int64_t cnt, c128* t) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const TsumItem q = items[i];
    c128 s = t[q.p];
    for (int ch = 1; ch < q.cnt; ++ch) s += t[q.p + int64_t(ch) * q.stride];
    t[q.dst] = s;
}

__device__ __forceinline__ int64_t bcast64(int64_t v, int src) {
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), src);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), src);
    return (int64_t(hi) << 32) | int64_t(uint32_t(lo));
}

// Segment s of a row cluster (work entries seg_e0[s] .. seg_e1[s]), rows
// r0 .. r0 + 63 of the cluster (seg_nr[s] rows): acc = sum of D_b xp over its
// dense blocks and U_run t_run over its low-rank run, slot nacc[seg_off[s] +
// row] = alpha acc.  Both are the same loop: a column-major nr x c matrix
// times c values (xp of the block's column cluster, or the run's t).
template <typename T, bool NT>
__global__ __launch_bounds__(256) void hmv_node_kernel(
    const int32_t* __restrict__ item_seg, const int32_t* __restrict__ item_r0, int64_t cnt,
    const int32_t* __restrict__ seg_nr, const int32_t* __restrict__ seg_e0,
    const int32_t* __restrict__ seg_e1, const int64_t* __restrict__ seg_off,
    const HmvWork* __restrict__ work, const T* __restrict__ D, const T* __restrict__ U,
    const c128* __restrict__ t, const c128* __restrict__ xp, c128 alpha, c128* nacc,
    int paired) {
    const int lane = threadIdx.x & 63;
    const int wpg = blockDim.x >> 6;
    const int64_t W = int64_t(gridDim.x) * wpg;
    for (int64_t it = int64_t(blockIdx.x) * wpg + (threadIdx.x >> 6); it < cnt; it += W) {
        const int sg = item_seg[it];
        const int64_t p = item_r0[it] + lane;
        const int64_t e0 = seg_e0[sg], e1 = seg_e1[sg];
        const int64_t nr = seg_nr[sg];
        if constexpr (sizeof(T) == 8) {
            // complex FP32, an even row count (offsets are even): lane l of
            // half h takes rows r0 + 2(l mod 32) and the next (one aligned
            // 16-byte load per column), columns congruent to h mod 2; the
            // halves are added at the end
            if (paired && (nr & 1) == 0) {
                const int h = lane >> 5;
                const int64_t p0 = item_r0[it] + 2 * (lane & 31);
                const bool on2 = p0 < nr;
                const int64_t q0 = on2 ? p0 : 0;
                c128 a0{0.0, 0.0}, a1{0.0, 0.0};
                for (int64_t eb = e0; eb < e1; eb += 64) {
                    const int nq = (int)min<int64_t>(64, e1 - eb);
                    HmvWork mine{0, 0, 0, 0};
                    if (lane < nq) mine = work[eb + lane];
                    for (int qi = 0; qi < nq; ++qi) {
                        const int64_t off = bcast64(mine.off, qi);
                        const int64_t xoff = bcast64(mine.xoff, qi);
                        const int c = __builtin_amdgcn_readlane(mine.cnt, qi);
                        const bool run = __builtin_amdgcn_readlane(mine.kind, qi) != 0;
                        const T* d = (run ? U : D) + off + q0 + int64_t(h) * nr;
                        // wave-uniform: both parities loaded, one kept
                        const c128* xc = (run ? t : xp) + xoff;
                        int jb = 0;
                        for (; jb + 15 < c; jb += 16) {
                            float4 v[8];
                            c128 xv[8];
#pragma unroll
                            for (int u = 0; u < 8; ++u) {
                                v[u] = ldm2<NT>(d + int64_t(jb + 2 * u) * nr);
                                const c128 xa = xc[jb + 2 * u], xb = xc[jb + 2 * u + 1];
                                xv[u] = h ? xb : xa;
                            }
#pragma unroll
                            for (int u = 0; u < 8; ++u) {
                                a0 += c128{v[u].x, v[u].y} * xv[u];
                                a1 += c128{v[u].z, v[u].w} * xv[u];
                            }
                        }
                        for (; jb < c; jb += 2) {
                            if (jb + h < c) {
                                const float4 v = ldm2<NT>(d + int64_t(jb) * nr);
                                const c128 xv = xc[jb + h];
                                a0 += c128{v.x, v.y} * xv;
                                a1 += c128{v.z, v.w} * xv;
                            }
                        }
                    }
                }
                a0.re += __shfl_xor(a0.re, 32, 64);
                a0.im += __shfl_xor(a0.im, 32, 64);
                a1.re += __shfl_xor(a1.re, 32, 64);
                a1.im += __shfl_xor(a1.im, 32, 64);
                if (h == 0 && on2) {
                    nacc[seg_off[sg] + p0] = alpha * a0;
                    nacc[seg_off[sg] + p0 + 1] = alpha * a1;
                }
                continue;
            }
        }
        const bool on = p < nr;
        const int64_t pp = on ? p : 0;          // lanes past the cluster read row 0
        c128 acc{0.0, 0.0};
        for (int64_t eb = e0; eb < e1; eb += 64) {
            const int nq = (int)min<int64_t>(64, e1 - eb);
            HmvWork mine{0, 0, 0, 0};
            if (lane < nq) mine = work[eb + lane];
            for (int q = 0; q < nq; ++q) {
                const int64_t off = bcast64(mine.off, q);
                const int64_t xoff = bcast64(mine.xoff, q);
                const int c = __builtin_amdgcn_readlane(mine.cnt, q);
                const bool run = __builtin_amdgcn_readlane(mine.kind, q) != 0;
                const T* d = (run ? U : D) + off + pp;
                const c128* xc = (run ? t : xp) + xoff;  // wave-uniform, contiguous
                int j = 0;
                for (; j + 15 < c; j += 16) {
                    c128 a[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) a[u] = ldm<NT>(d + int64_t(j + u) * nr);
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc += a[u] * xc[j + u];
                }
                for (; j + 3 < c; j += 4) {
                    c128 a[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) a[u] = ldm<NT>(d + int64_t(j + u) * nr);
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc += a[u] * xc[j + u];
                }
                for (; j < c; ++j) acc += ldm<NT>(d + int64_t(j) * nr) * xc[j];
            }
        }
        if (on) nacc[seg_off[sg] + p] = alpha * acc;
    }
}

// y[i] = beta y[i] + sum of the contributions of row i (fixed order)
__global__ void hmv_reduce_kernel(int64_t n, const int64_t* __restrict__ rowptr,
                                  const int64_t* __restrict__ ent, const c128* __restrict__ part,
                                  c128 beta, c128* y) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    c128 s{0.0, 0.0};
    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += part[ent[e]];
    // BLAS convention: beta = 0 does not read y (which may be uninitialised)
    if (beta.re == 0.0 && beta.im == 0.0) y[i] = s;
    else if (beta.re == 1.0 && beta.im == 0.0) y[i] = y[i] + s;
    else y[i] = beta * y[i] + s;
}

// A[rows[i] + cols[j] lda] = block entry (the blocks partition the matrix,
// so every entry is written once; h_matrix.rs:235-247)
template <typename T>
__global__ void hmat_full_kernel(const HBlkDev* __restrict__ blks, const int64_t* __restrict__ rows,
                                 const int64_t* __restrict__ cols, const T* __restrict__ U,
                                 const T* __restrict__ V, const T* __restrict__ D, c128* A,
                                 int64_t lda) {
    const HBlkDev bk = blks[blockIdx.x];
    const int64_t cnt = bk.nr * bk.nc;
    for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cnt;
         e += int64_t(blockDim.x) * gridDim.y) {
        const int64_t i = e % bk.nr, j = e / bk.nr;
        c128 v{0.0, 0.0};
        if (bk.adm) {
            for (int k = 0; k < bk.rank; ++k)
                v += ldx(U + bk.u_off + i + k * bk.nr) * ldx(V + bk.v_off + j + k * bk.nc);
        } else {
            v = ldx(D + bk.d_off + e);
        }
        A[rows[bk.r_off + i] + cols[bk.c_off + j] * lda] = v;
    }
}

// low-rank factors to the depth-first row order of their clusters:
// row i of a factor moves to pos[list offset + i]
__global__ void reorder_lr_kernel(const HBlkDev* __restrict__ blks, const int32_t* __restrict__ list,
                                  const int64_t* __restrict__ pos, const c128* __restrict__ U,
                                  const c128* __restrict__ V, c128* U2, c128* V2) {
    const HBlkDev bk = blks[list[blockIdx.x]];
    const int64_t cu = bk.nr * bk.rank, cv = bk.nc * bk.rank;
    for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cu;
         e += int64_t(blockDim.x) * gridDim.y) {
        const int64_t i = e % bk.nr, k = e / bk.nr;
        U2[bk.u_off + pos[bk.r_off + i] + k * bk.nr] = U[bk.u_off + e];
    }
    for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cv;
         e += int64_t(blockDim.x) * gridDim.y) {
        const int64_t j = e % bk.nc, k = e / bk.nc;
        V2[bk.v_off + pos[bk.c_off + j] + k * bk.nc] = V[bk.v_off + e];
    }
}

// U of every low-rank block (nr x rank, contiguous) copied to its place in
// the runs' storage (s_off)
__global__ void repack_u_kernel(const HBlkDev* __restrict__ blks, const int32_t* __restrict__ list,
                                const c128* __restrict__ U, c128* U3) {
    const HBlkDev bk = blks[list[blockIdx.x]];
    const int64_t cu = bk.nr * bk.rank;
    for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cu;
         e += int64_t(blockDim.x) * gridDim.y)
        U3[bk.s_off + e] = U[bk.u_off + e];
}

// xp[p] = x[perm[p]]
__global__ void hmv_gather_kernel(int64_t n, const int64_t* __restrict__ perm,
                                  const c128* __restrict__ x, c128* __restrict__ xp) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) xp[p] = x[perm[p]];
}

// FP32 storage: the first rank columns of every low-rank block's U and V
// (contiguous, nr * rank and nc * rank entries) packed at the new offsets
// held in w_off / f_off, and every dense block at the offset held in u2_off,
// as c64
__global__ void hmat_to_c64_kernel(const HBlkDev* __restrict__ blks, const c128* __restrict__ U,
                                   const c128* __restrict__ V, const c128* __restrict__ D,
                                   c64* U32, c64* V32, c64* D32) {
    const HBlkDev bk = blks[blockIdx.x];
    auto cvt = [](c128 v) { return c64{static_cast<float>(v.re), static_cast<float>(v.im)}; };
    if (bk.adm) {
        const int64_t cu = bk.nr * bk.rank, cv = bk.nc * bk.rank;
        for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cu;
             e += int64_t(blockDim.x) * gridDim.y)
            U32[bk.w_off + e] = cvt(U[bk.u_off + e]);
        for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cv;
             e += int64_t(blockDim.x) * gridDim.y)
            V32[bk.f_off + e] = cvt(V[bk.v_off + e]);
    } else {
        const int64_t cd = bk.nr * bk.nc;
        for (int64_t e = threadIdx.x + int64_t(blockIdx.y) * blockDim.x; e < cd;
             e += int64_t(blockDim.x) * gridDim.y)
            D32[bk.u2_off + e] = cvt(D[bk.d_off + e]);
    }
}

template <typename T>
T* dalloc(std::vector<void*>& owned, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
    owned.push_back(p);
    return static_cast<T*>(p);
}

}  // namespace

struct ccab_hmatrix {
    int64_t n = 0;     // rows
    int64_t ncol = 0;  // columns (= n except for a single ACA block)
    int device = 0;
    std::vector<void*> owned;  // device allocations
    std::vector<HBlkDev> blk;  // host copy (ranks, norms after build)
    HBlkDev* d_blk = nullptr;
    int64_t* d_rows = nullptr;
    int64_t* d_cols = nullptr;
    c128* d_U = nullptr;
    c128* d_V = nullptr;
    c128* d_D = nullptr;
    // FP32 storage (the blocks built in FP64, then stored as c64; U, V packed
    // to their ranks): the c128 arrays above are released
    bool fp32 = false;
    c64* d_U32 = nullptr;
    c64* d_V32 = nullptr;
    c64* d_D32 = nullptr;
    // matvec plan: xp = x in the depth-first order of the column tree
    // (xp[p] = x[perm[p]]), t = V^T xp per low-rank block, one slot per
    // (cluster, row)
    int64_t* d_perm = nullptr;
    c128* d_xp = nullptr;
    c128* d_t = nullptr;
    c128* d_nacc = nullptr;
    VtItem* d_vt_items = nullptr;    // V^T x items: (low-rank block, column chunk)
    int64_t n_vt = 0;
    TsumItem* d_tsum = nullptr;      // chunk sums: (block of several chunks, rank)
    int64_t n_tsum = 0;
    int32_t* d_item_seg = nullptr;   // product items: (segment, 64-row tile)
    int32_t* d_item_r0 = nullptr;
    int64_t n_items = 0;
    int32_t* d_seg_nr = nullptr;     // segment -> rows of its cluster, entries [e0, e1) of d_work
    int32_t* d_seg_e0 = nullptr;
    int32_t* d_seg_e1 = nullptr;
    int64_t* d_seg_off = nullptr;    // segment -> first slot
    HmvWork* d_work = nullptr;       // clusters' blocks (dense first, then low-rank)
    int64_t* d_rowptr = nullptr;     // row -> its slots (CSR over d_ent)
    int64_t* d_ent = nullptr;
    int32_t* d_dense_list = nullptr;
    int32_t* d_lr_list = nullptr;
    int64_t n_dense = 0, n_lr = 0;
    double norm = 0.0;
    double mv_bytes = 0.0;  // block data one matvec reads
    int64_t aca_fallbacks = 0;
    int64_t aca_bucket[4] = {0, 0, 0, 0};  // admissible blocks by ACA rank: <= 8, 16, 32, 64
    ~ccab_hmatrix() {
        for (void* p : owned) (void)hipFree(p);
    }
};

namespace {
// phase times of the build on stderr when CCAB_TRACE_HBUILD is set (each mark
// waits for the stream, so traced builds are slower)
struct BuildTrace {
    hipStream_t s;
    bool on;
    std::chrono::steady_clock::time_point t0, t;
    explicit BuildTrace(hipStream_t st) : s(st), on(std::getenv("CCAB_TRACE_HBUILD") != nullptr) {
        t0 = t = std::chrono::steady_clock::now();
    }
    void mark(const char* what) {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[hbuild] %-28s %8.2f ms (total %8.2f ms)\n", what,
                     std::chrono::duration<double, std::milli>(now - t).count(),
                     std::chrono::duration<double, std::milli>(now - t0).count());
        t = now;
    }
};

// The blocks of a block list (bl, over the clusters T) built on the device
// from an entry source: batched ACA + recompression of the admissible
// blocks, dense inadmissible blocks, norms and the matvec plan.  The operator
// has nrow rows and ncol columns (square for the cluster-tree H matrices; one
// rectangular admissible block for a plain ACA).
template <typename E>
int hmatrix_from_blocks(ccab_ctx* ctx, const E& entry, std::vector<HNode> T,
                        const std::vector<HostBlock>& bl, int64_t nrow, int64_t ncol, double tol,
                        ccab_hmatrix** out, int prec, BuildTrace& trace) {
    hipStream_t s = ctx->stream;
    std::unique_ptr<ccab_hmatrix> owner(new ccab_hmatrix());
    ccab_hmatrix* H = owner.get();
    H->n = nrow;
    H->ncol = ncol;
    H->device = ctx->device;
    auto fail = [&](int code, const char* msg) { return ctx_fail(ctx, code, "%s", msg); };
    const int64_t nb = (int64_t)bl.size();
    const int64_t n = nrow;
    // every cluster's index list stored once (n x tree depth entries); a
    // block's rows / columns are its clusters' lists, so rows[] and cols[]
    // are the same array
    std::vector<int64_t> node_off(T.size());
    int64_t nidx = 0;
    for (size_t c = 0; c < T.size(); ++c) {
        node_off[c] = nidx;
        nidx += (int64_t)T[c].idx.size();
    }
    std::vector<int64_t> rows(nidx);
    for (size_t c = 0; c < T.size(); ++c)
        std::copy(T[c].idx.begin(), T[c].idx.end(), rows.begin() + node_off[c]);
    H->blk.resize(nb);
    int64_t u_off = 0, v_off = 0, d_off = 0, w_off = 0, f_off = 0;
    std::vector<int32_t> dl, ll;
    for (int64_t b = 0; b < nb; ++b) {
        HBlkDev& d = H->blk[b];
        std::memset(&d, 0, sizeof(d));
        d.r_off = node_off[bl[b].r];
        d.nr = (int64_t)T[bl[b].r].idx.size();
        d.c_off = node_off[bl[b].c];
        d.nc = (int64_t)T[bl[b].c].idx.size();
        d.adm = bl[b].adm ? 1 : 0;
        if (d.adm) {
            // the ACA takes at most min(nr, nc) terms
            const int64_t cap = std::min<int64_t>(std::min(d.nr, d.nc), RMAX);
            d.u_off = u_off;
            d.v_off = v_off;
            u_off += d.nr * cap;
            v_off += d.nc * cap;
            d.w_off = w_off;
            w_off += 2 * (d.nr + d.nc);
            d.f_off = f_off;
            f_off += d.nr + d.nc;
            ll.push_back((int32_t)b);
        } else {
            d.d_off = d_off;
            d_off += d.nr * d.nc;
            dl.push_back((int32_t)b);
        }
    }
    trace.mark("block layout");
    H->d_rows = dalloc<int64_t>(H->owned, rows.size());
    H->d_cols = H->d_rows;
    H->d_blk = dalloc<HBlkDev>(H->owned, nb);
    // the ACA's factors at full rank capacity and its vectors and flags live
    // in context buffers reused by every build (tens of GB at n = 81920:
    // allocating them per build made build times erratic); the H matrix owns
    // only the recompressed factors, packed to the ACA ranks
    c128* Ua = static_cast<c128*>(ctx_buf(ctx, SLOT_HM_UA, size_t(std::max<int64_t>(u_off, 1)) * 16));
    c128* Va = static_cast<c128*>(ctx_buf(ctx, SLOT_HM_VA, size_t(std::max<int64_t>(v_off, 1)) * 16));
    c128* Wws = static_cast<c128*>(ctx_buf(ctx, SLOT_HM_W, size_t(std::max<int64_t>(w_off, 1)) * 16));
    uint8_t* Fws = static_cast<uint8_t*>(ctx_buf(ctx, SLOT_HM_F, size_t(std::max<int64_t>(f_off, 1))));
    uint32_t* d_rng = dalloc<uint32_t>(H->owned, kRngTable);
    // fallbacks move blocks from the low-rank to the dense list: size both for every block
    H->d_dense_list = dalloc<int32_t>(H->owned, nb + 1);
    H->d_lr_list = dalloc<int32_t>(H->owned, nb + 1);
    if (!H->d_rows || !H->d_cols || !H->d_blk || !Ua || !Va || !Wws || !Fws || !d_rng ||
        !H->d_dense_list || !H->d_lr_list)
        return fail(CCAB_ENOMEM, "H-matrix allocation failed");
    std::vector<uint32_t> words(kRngTable);
    chacha12_words(10, words.data(), kRngTable);
    CCAB_HIP(hipMemcpy(H->d_rows, rows.data(), rows.size() * 8, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(d_rng, words.data(), kRngTable * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_lr_list, ll.data(), ll.size() * 4, hipMemcpyHostToDevice));
    trace.mark("allocation and upload");
    // batched ACA + recompression of the admissible blocks
    if (!ll.empty()) {
        // blocks of at most 64 PL rows and columns (PL = 1, 2, 4 up to
        // CCAB_HM_WAVE_ACA, default 2; 0: none): one wavefront each; the
        // larger ones a workgroup each (n = 81920 ACA: 56 ms with a workgroup
        // per block, 51.5 with wavefronts up to 64 or 128, 54 up to 256,
        // where the registers leave one wavefront per SIMD)
        const char* wac = std::getenv("CCAB_HM_WAVE_ACA");
        const int wave_pl = wac ? std::atoi(wac) : 2;
        std::vector<int32_t> grp[4];  // PL = 1, 2, 4; workgroup
        for (int32_t b : ll) {
            const int64_t m = std::max(H->blk[b].nr, H->blk[b].nc);
            const int g = m <= 64 && wave_pl >= 1 ? 0 : m <= 128 && wave_pl >= 2 ? 1
                        : m <= 256 && wave_pl >= 4 ? 2 : 3;
            grp[g].push_back(b);
        }
        int32_t* d_aca = dalloc<int32_t>(H->owned, ll.size());
        if (!d_aca) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        std::vector<int32_t> aca_order;
        for (auto& g : grp) aca_order.insert(aca_order.end(), g.begin(), g.end());
        CCAB_HIP(hipMemcpy(d_aca, aca_order.data(), aca_order.size() * 4, hipMemcpyHostToDevice));
        int64_t off = 0;
        for (int g = 0; g < 4; ++g) {
            const int64_t c = (int64_t)grp[g].size();
            if (c == 0) continue;
            const dim3 wg((unsigned)((c + 3) / 4));
            if (g == 0)
                hipLaunchKernelGGL((aca_wave_kernel<E, 1>), wg, dim3(256), 0, s, entry, H->d_blk,
                                   d_aca + off, c, H->d_rows, H->d_cols, Ua, Va, d_rng, tol);
            else if (g == 1)
                hipLaunchKernelGGL((aca_wave_kernel<E, 2>), wg, dim3(256), 0, s, entry, H->d_blk,
                                   d_aca + off, c, H->d_rows, H->d_cols, Ua, Va, d_rng, tol);
            else if (g == 2)
                hipLaunchKernelGGL((aca_wave_kernel<E, 4>), wg, dim3(256), 0, s, entry, H->d_blk,
                                   d_aca + off, c, H->d_rows, H->d_cols, Ua, Va, d_rng, tol);
            else
                hipLaunchKernelGGL(aca_kernel<E>, dim3((unsigned)c), dim3(kAcaThreads), 0, s, entry,
                                   H->d_blk, d_aca + off, H->d_rows, H->d_cols, Ua, Va, Wws, Fws,
                                   d_rng, tol);
            off += c;
        }
        // recompression with LDS sized by rank: blocks bucketed by their ACA
        // rank (most are far below RMAX, and the RMAX-sized kernel holds one
        // workgroup per CU)
        CCAB_HIP(hipMemcpyAsync(H->blk.data(), H->d_blk, nb * sizeof(HBlkDev), hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        trace.mark("ACA");
        constexpr int kBuckets = 4;
        const int cap[kBuckets] = {8, 16, 32, RMAX};
        std::vector<int32_t> bucket[kBuckets];
        for (int32_t b : ll) {
            const HBlkDev& d = H->blk[b];
            if (d.status != 0) continue;  // becomes dense below
            int k = 0;
            while (k < kBuckets - 1 && d.rank > cap[k]) ++k;
            bucket[k].push_back(b);
        }
        std::vector<int32_t> order;
        for (auto& v : bucket) order.insert(order.end(), v.begin(), v.end());
        for (int k = 0; k < kBuckets; ++k) H->aca_bucket[k] = (int64_t)bucket[k].size();
        if (trace.on) {
            for (int k = 0; k < kBuckets; ++k) {
                double rows = 0.0;
                for (int32_t b : bucket[k]) rows += double(H->blk[b].nr + H->blk[b].nc);
                std::fprintf(stderr, "[hbuild] rank <= %2d: %6zu blocks, mean nr + nc %.0f\n", cap[k],
                             bucket[k].size(), bucket[k].empty() ? 0.0 : rows / double(bucket[k].size()));
            }
        }
        // recompressed factors packed to the ACA ranks (the recompressed rank
        // is at most the ACA rank); R factors sized per bucket (the bucket
        // launches run one after another)
        int64_t cu = 0, cv = 0;
        size_t rws = 1;
        for (int k = 0; k + 1 < kBuckets; ++k)
            rws = std::max(rws, bucket[k].size() * 2 * size_t(cap[k]) * size_t(cap[k]));
        // the largest-rank bucket (few blocks, one workgroup per CU each) runs
        // on the side stream beside the others, in R space of its own
        const size_t rws_big = rws;
        rws += bucket[kBuckets - 1].size() * 2 * size_t(RMAX) * size_t(RMAX);
        for (int32_t b : order) {
            HBlkDev& d = H->blk[b];
            d.u2_off = cu;
            cu += d.nr * d.rank;
            d.v2_off = cv;
            cv += d.nc * d.rank;
        }
        H->d_U = dalloc<c128>(H->owned, cu);
        H->d_V = dalloc<c128>(H->owned, cv);
        c128* Rws = static_cast<c128*>(ctx_buf(ctx, SLOT_HM_R, rws * 16));
        if (!H->d_U || !H->d_V || !Rws) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
        if (!order.empty()) {
            int32_t* d_order = dalloc<int32_t>(H->owned, order.size());
            if (!d_order) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
            CCAB_HIP(hipMemcpy(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice));
            c128* U2 = H->d_U;
            c128* V2 = H->d_V;
            size_t off = 0;
            hipStream_t sb = s;  // stream of the largest-rank bucket
            if (!bucket[kBuckets - 1].empty()) {
                if (!ctx->side_stream) {
                    CCAB_HIP(hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking));
                    for (hipEvent_t& e : ctx->side_ev)
                        CCAB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                }
                sb = ctx->side_stream;
                CCAB_HIP(hipEventRecord(ctx->side_ev[0], s));
                CCAB_HIP(hipStreamWaitEvent(sb, ctx->side_ev[0], 0));
            }
            // blocks of rank <= 16: one wavefront each (CCAB_HM_WAVE_RECOMPRESS=0: a workgroup each)
            const char* wrc = std::getenv("CCAB_HM_WAVE_RECOMPRESS");
            const bool wave_rc = !wrc || std::atoi(wrc) != 0;
            for (int k = 0; k < kBuckets; ++k) {
                const unsigned cnt = (unsigned)bucket[k].size();
                if (cnt == 0) continue;
                const int32_t* lst = d_order + off;
                switch (k) {
                case 0:
                    if (wave_rc)
                        hipLaunchKernelGGL(recompress_wave_kernel<8>, dim3((cnt + 3) / 4), dim3(256), 0, s, H->d_blk, lst, int64_t(cnt), Ua, Va, U2, V2, Rws, tol);
                    else
                        hipLaunchKernelGGL(recompress_kernel<8>, dim3(cnt), dim3(kAcaThreads), 0, s, H->d_blk, lst, Ua, Va, U2, V2, Rws, tol);
                    break;
                case 1:
                    if (wave_rc)
                        hipLaunchKernelGGL(recompress_wave_kernel<16>, dim3((cnt + 3) / 4), dim3(256), 0, s, H->d_blk, lst, int64_t(cnt), Ua, Va, U2, V2, Rws, tol);
                    else
                        hipLaunchKernelGGL(recompress_kernel<16>, dim3(cnt), dim3(kAcaThreads), 0, s, H->d_blk, lst, Ua, Va, U2, V2, Rws, tol);
                    break;
                case 2: hipLaunchKernelGGL(recompress_kernel<32>, dim3(cnt), dim3(kAcaThreads), 0, s, H->d_blk, lst, Ua, Va, U2, V2, Rws, tol); break;
                default: hipLaunchKernelGGL(recompress_kernel<RMAX>, dim3(cnt), dim3(kAcaThreads), 0, sb, H->d_blk, lst, Ua, Va, U2, V2, Rws + rws_big, tol); break;
                }
                off += cnt;
            }
            CCAB_HIP(hipGetLastError());
            if (sb != s) {  // the main stream waits for the side stream's bucket
                CCAB_HIP(hipEventRecord(ctx->side_ev[1], sb));
                CCAB_HIP(hipStreamWaitEvent(s, ctx->side_ev[1], 0));
            }
        }
        // the recompressed ranks come back; the factors now sit at the packed offsets
        CCAB_HIP(hipMemcpyAsync(H->blk.data(), H->d_blk, nb * sizeof(HBlkDev), hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
        for (int32_t b : order) {
            HBlkDev& d = H->blk[b];
            d.u_off = d.u2_off;
            d.v_off = d.v2_off;
        }
        CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
        trace.mark("recompression");
        hipLaunchKernelGGL(aca_norm_kernel, dim3((unsigned)ll.size()), dim3(kAcaThreads), 0, s,
                           H->d_blk, H->d_lr_list, H->d_U, H->d_V);
        CCAB_HIP(hipGetLastError());
        CCAB_HIP(hipMemcpyAsync(H->blk.data(), H->d_blk, nb * sizeof(HBlkDev), hipMemcpyDeviceToHost, s));
        CCAB_HIP(hipStreamSynchronize(s));
    }
    // blocks whose ACA hit the rank capacity become dense
    std::vector<int32_t> ll2;
    for (int32_t b : ll) {
        HBlkDev& d = H->blk[b];
        if (d.status != 0) {
            d.adm = 0;
            d.d_off = d_off;
            d_off += d.nr * d.nc;
            dl.push_back(b);
            H->aca_fallbacks++;
        } else {
            ll2.push_back(b);
        }
    }
    std::sort(dl.begin(), dl.end());
    H->n_dense = (int64_t)dl.size();
    H->n_lr = (int64_t)ll2.size();
    // V^T x of every low-rank block: one rank-vector of partials per chunk
    // of kVtChunk columns (offsets with the matvec plan below)
    for (int32_t b : ll2) {
        HBlkDev& d = H->blk[b];
        d.vtc = (int32_t)((d.nc + kVtChunk - 1) / kVtChunk);
    }
    H->d_D = dalloc<c128>(H->owned, d_off);
    if (!H->d_D) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
    CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_dense_list, dl.data(), dl.size() * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_lr_list, ll2.data(), ll2.size() * 4, hipMemcpyHostToDevice));
    trace.mark("norms, dense storage");
    // Storage order of the matvec: every cluster's points in the depth-first
    // leaf order of its tree, so that a cluster is one contiguous range of
    // positions (the product permutes x once and reads it contiguously).  The
    // ACA above ran on the cluster lists in the order of cluster.rs (its
    // pivots depend on it); the low-rank factors' rows move to the new order
    // here and the dense blocks are built in it.
    {
        std::vector<int> parent(T.size(), -1);
        for (size_t c = 0; c < T.size(); ++c)
            for (int q : T[c].son)
                if (q >= 0) parent[q] = (int)c;
        std::vector<int64_t> lo(T.size(), 0);
        std::vector<int> root_of(T.size(), 0);
        std::vector<std::vector<int64_t>> perm_of;   // per root: position -> point
        std::vector<std::vector<int64_t>> inv_of;    // per root: point -> position
        std::vector<int> roots;
        for (size_t c = 0; c < T.size(); ++c)
            if (parent[c] < 0) roots.push_back((int)c);
        for (size_t ri = 0; ri < roots.size(); ++ri) {
            std::vector<int64_t> perm;
            // depth-first, son 0 before son 1; lo of a cluster = its first leaf's start
            std::function<void(int)> walk = [&](int c) {
                root_of[c] = (int)ri;
                lo[c] = (int64_t)perm.size();
                if (T[c].son[0] < 0) {
                    perm.insert(perm.end(), T[c].idx.begin(), T[c].idx.end());
                } else {
                    walk(T[c].son[0]);
                    walk(T[c].son[1]);
                }
            };
            walk(roots[ri]);
            int64_t mx = 0;
            for (int64_t g : perm) mx = std::max(mx, g + 1);
            std::vector<int64_t> inv(mx, -1);
            for (size_t p = 0; p < perm.size(); ++p) inv[perm[p]] = (int64_t)p;
            perm_of.push_back(std::move(perm));
            inv_of.push_back(std::move(inv));
        }
        // position of every list entry in its cluster's new order, then the new lists
        std::vector<int64_t> pos(nidx);
        for (size_t c = 0; c < T.size(); ++c) {
            const auto& inv = inv_of[root_of[c]];
            for (size_t i = 0; i < T[c].idx.size(); ++i)
                pos[node_off[c] + i] = inv[T[c].idx[i]] - lo[c];
            const auto& perm = perm_of[root_of[c]];
            std::copy(perm.begin() + lo[c], perm.begin() + lo[c] + (int64_t)T[c].idx.size(),
                      T[c].idx.begin());
            std::copy(T[c].idx.begin(), T[c].idx.end(), rows.begin() + node_off[c]);
        }
        for (int64_t b = 0; b < nb; ++b) H->blk[b].clo = lo[bl[b].c];
        const int col_root = root_of[bl.empty() ? 0 : bl[0].c];
        H->d_perm = dalloc<int64_t>(H->owned, perm_of[col_root].size());
        H->d_xp = dalloc<c128>(H->owned, perm_of[col_root].size());
        int64_t* d_pos = dalloc<int64_t>(H->owned, nidx);
        if (!H->d_perm || !H->d_xp || !d_pos) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        CCAB_HIP(hipMemcpy(H->d_perm, perm_of[col_root].data(), perm_of[col_root].size() * 8,
                           hipMemcpyHostToDevice));
        CCAB_HIP(hipMemcpy(d_pos, pos.data(), pos.size() * 8, hipMemcpyHostToDevice));
        CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
        trace.mark("depth-first order (host)");
        if (!ll2.empty()) {
            int64_t cu = 0, cv = 0;
            for (int32_t b : ll2) {
                cu = std::max(cu, H->blk[b].u_off + H->blk[b].nr * H->blk[b].rank);
                cv = std::max(cv, H->blk[b].v_off + H->blk[b].nc * H->blk[b].rank);
            }
            c128* U2 = dalloc<c128>(H->owned, cu);
            c128* V2 = dalloc<c128>(H->owned, cv);
            if (!U2 || !V2) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
            hipLaunchKernelGGL(reorder_lr_kernel, dim3((unsigned)ll2.size(), 4), dim3(256), 0, s,
                               H->d_blk, H->d_lr_list, d_pos, H->d_U, H->d_V, U2, V2);
            CCAB_HIP(hipGetLastError());
            CCAB_HIP(hipStreamSynchronize(s));
            trace.mark("factor rows reordered");
            for (void* old : {static_cast<void*>(H->d_U), static_cast<void*>(H->d_V)}) {
                auto it = std::find(H->owned.begin(), H->owned.end(), old);
                if (it != H->owned.end()) {
                    (void)hipFree(*it);
                    H->owned.erase(it);
                }
            }
            H->d_U = U2;
            H->d_V = V2;
        }
        CCAB_HIP(hipMemcpy(H->d_rows, rows.data(), rows.size() * 8, hipMemcpyHostToDevice));
        auto it = std::find(H->owned.begin(), H->owned.end(), static_cast<void*>(d_pos));
        if (it != H->owned.end()) {
            (void)hipFree(*it);
            H->owned.erase(it);
        }
        trace.mark("depth-first storage order");
    }
    if (!dl.empty()) {
        hipLaunchKernelGGL(dense_block_kernel<E>, dim3((unsigned)dl.size(), 4), dim3(256), 0, s,
                           entry, H->d_blk, H->d_dense_list, H->d_rows, H->d_cols, H->d_D);
        hipLaunchKernelGGL(dense_norm_kernel, dim3((unsigned)dl.size()), dim3(256), 0, s, H->d_blk,
                           H->d_dense_list, H->d_D);
    }
    CCAB_HIP(hipGetLastError());
    CCAB_HIP(hipMemcpyAsync(H->blk.data(), H->d_blk, nb * sizeof(HBlkDev), hipMemcpyDeviceToHost, s));
    CCAB_HIP(hipStreamSynchronize(s));
    trace.mark("norms and dense blocks");
    // norm = sum of dense Frobenius norms, then the low-rank estimates (h_matrix.rs:189-196)
    double nd = 0.0, nl = 0.0;
    for (int32_t b : dl) nd += H->blk[b].norm;
    for (int32_t b : ll2) nl += H->blk[b].norm;
    H->norm = nd + nl;
    for (const auto& d : H->blk)
        H->mv_bytes += 16.0 * (d.adm ? double(d.rank) * double(d.nr + d.nc) : double(d.nr) * double(d.nc));
    // matvec plan: every cluster's blocks (dense first, then low-rank, each in
    // block order), 64-row tiles of the clusters that own blocks, and for
    // every row the slots of those clusters (cluster order)
    const int64_t nn = (int64_t)T.size();
    std::vector<int64_t> nptr(nn + 1, 0);
    for (int64_t b = 0; b < nb; ++b) nptr[bl[b].r + 1]++;
    for (int64_t c = 0; c < nn; ++c) nptr[c + 1] += nptr[c];
    std::vector<int32_t> nlist(nb);
    {
        std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
        for (int pass = 0; pass < 2; ++pass)
            for (int64_t b = 0; b < nb; ++b)
                if ((pass == 0) == (H->blk[b].adm == 0)) nlist[fill[bl[b].r]++] = (int32_t)b;
    }
    // segments: runs of a cluster's blocks (in nlist order) worth at most
    // kNodeBudget columns each (at least one block), one slot per (segment,
    // row); items: (segment, 64-row tile)
    std::vector<int32_t> seg_node, seg_e0, seg_e1, item_seg, item_r0;
    std::vector<int64_t> seg_off, seg_work;
    int64_t nslots = 0;
    const int64_t budget = ctx->hmv_node_budget > 0 ? ctx->hmv_node_budget : INT64_MAX;
    for (int64_t c = 0; c < nn; ++c) {
        int64_t e = nptr[c];
        while (e < nptr[c + 1]) {
            const int64_t ea = e;
            int64_t w = 0;
            do {
                const HBlkDev& d = H->blk[nlist[e]];
                w += d.adm ? d.rank : d.nc;
                ++e;
            } while (e < nptr[c + 1] && w + (H->blk[nlist[e]].adm ? H->blk[nlist[e]].rank
                                                                    : H->blk[nlist[e]].nc) <=
                                            budget);
            seg_node.push_back((int32_t)c);
            seg_e0.push_back((int32_t)ea);
            seg_e1.push_back((int32_t)e);
            seg_off.push_back(nslots);
            seg_work.push_back(w);
            nslots += (int64_t)T[c].idx.size();
        }
    }
    const int64_t nseg = (int64_t)seg_node.size();
    std::vector<int32_t> seg_nr(nseg);
    for (int64_t sg = 0; sg < nseg; ++sg) seg_nr[sg] = (int32_t)T[seg_node[sg]].idx.size();
    // Low-rank runs: the low-rank blocks of a segment (its last entries; a
    // cluster lists its dense blocks first) are stored one after another, so
    // that their U form one nr x R column-major matrix (R the sum of their
    // ranks; the segment's blocks share the row cluster) and their summed
    // V^T x one vector of R values: the product treats the run as one dense
    // block over t.  Work entries: a dense block (>= 0) or a run (-1 - its
    // first block, with R in wk_cnt); every run's U at an even offset.
    std::vector<int64_t> wk_blk;
    std::vector<int32_t> wk_cnt, seg_w0(nseg), seg_w1(nseg);
    int64_t s_total = 0, ts_total = 0;
    auto even = [](int64_t v) { return (v + 1) & ~int64_t(1); };
    for (int64_t sg = 0; sg < nseg; ++sg) {
        seg_w0[sg] = (int32_t)wk_blk.size();
        int32_t first = -1;
        int64_t R = 0;
        for (int64_t e = seg_e0[sg]; e < seg_e1[sg]; ++e) {
            const int32_t b = nlist[e];
            HBlkDev& d = H->blk[b];
            if (!d.adm) {
                wk_blk.push_back(b);
                wk_cnt.push_back(0);
                continue;
            }
            if (first < 0) {
                first = b;
                s_total = even(s_total);
            }
            d.s_off = s_total;
            s_total += d.nr * d.rank;
            d.ts_off = ts_total;
            ts_total += d.rank;
            R += d.rank;
        }
        if (first >= 0) {
            wk_blk.push_back(-1 - int64_t(first));
            wk_cnt.push_back((int32_t)R);
        }
        seg_w1[sg] = (int32_t)wk_blk.size();
    }
    // t: the summed V^T x of every block (ts_off), then the chunk partials
    // of the blocks with several chunks (p_off)
    int64_t t_total = ts_total;
    for (int32_t b : ll2) {
        HBlkDev& d = H->blk[b];
        if (d.vtc > 1) {
            d.p_off = t_total;
            t_total += int64_t(d.vtc) * d.rank;
        } else {
            d.p_off = d.ts_off;
        }
    }
    H->d_t = dalloc<c128>(H->owned, std::max<int64_t>(t_total, 1));
    if (!H->d_t) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
    for (int64_t sg = 0; sg < nseg; ++sg)
        for (int64_t r0 = 0; r0 < (int64_t)T[seg_node[sg]].idx.size(); r0 += 64) {
            item_seg.push_back((int32_t)sg);
            item_r0.push_back((int32_t)r0);
        }
    // every row's slots, in segment order (a fixed summation order)
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t sg = 0; sg < nseg; ++sg)
        for (int64_t g : T[seg_node[sg]].idx) cnt[g + 1]++;
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int64_t> ent(cnt[n]);
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t sg = 0; sg < nseg; ++sg) {
            const auto& idx = T[seg_node[sg]].idx;
            for (int64_t q = 0; q < (int64_t)idx.size(); ++q) ent[fill[idx[q]]++] = seg_off[sg] + q;
        }
    }
    // items by decreasing work (stable: equal work keeps the tree order)
    if (ctx->hmv_sort) {
        std::vector<size_t> ord(item_seg.size());
        for (size_t q = 0; q < ord.size(); ++q) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
            return seg_work[item_seg[x]] > seg_work[item_seg[y]];
        });
        std::vector<int32_t> a2(ord.size()), b2(ord.size());
        for (size_t q = 0; q < ord.size(); ++q) {
            a2[q] = item_seg[ord[q]];
            b2[q] = item_r0[ord[q]];
        }
        item_seg.swap(a2);
        item_r0.swap(b2);
    }
    std::vector<int32_t> vt_blk, vt_chunk;
    std::vector<int64_t> vt_work;
    for (int32_t b : ll2) {
        const HBlkDev& d = H->blk[b];
        for (int32_t ch = 0; ch < d.vtc; ++ch) {
            vt_blk.push_back(b);
            vt_chunk.push_back(ch);
            vt_work.push_back(std::min<int64_t>(kVtChunk, d.nc - int64_t(ch) * kVtChunk) * d.rank);
        }
    }
    if (ctx->hmv_sort) {
        std::vector<size_t> ord(vt_blk.size());
        for (size_t q = 0; q < ord.size(); ++q) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](size_t x, size_t y) { return vt_work[x] > vt_work[y]; });
        std::vector<int32_t> a2(ord.size()), b2(ord.size());
        for (size_t q = 0; q < ord.size(); ++q) {
            a2[q] = vt_blk[ord[q]];
            b2[q] = vt_chunk[ord[q]];
        }
        vt_blk.swap(a2);
        vt_chunk.swap(b2);
    }
    H->n_vt = (int64_t)vt_blk.size();
    H->n_items = (int64_t)item_seg.size();
    std::vector<TsumItem> tsum;
    for (int32_t b : ll2) {
        const HBlkDev& d = H->blk[b];
        if (d.vtc > 1)
            for (int32_t k = 0; k < d.rank; ++k)
                tsum.push_back(TsumItem{d.p_off + k, d.ts_off + k, d.rank, d.vtc});
    }
    H->n_tsum = (int64_t)tsum.size();
    if (H->n_tsum) {
        H->d_tsum = dalloc<TsumItem>(H->owned, tsum.size());
        if (!H->d_tsum) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        CCAB_HIP(hipMemcpy(H->d_tsum, tsum.data(), tsum.size() * sizeof(TsumItem),
                           hipMemcpyHostToDevice));
    }
    H->d_nacc = dalloc<c128>(H->owned, nslots);
    H->d_vt_items = dalloc<VtItem>(H->owned, vt_blk.size());
    H->d_item_seg = dalloc<int32_t>(H->owned, item_seg.size());
    H->d_item_r0 = dalloc<int32_t>(H->owned, item_r0.size());
    H->d_seg_nr = dalloc<int32_t>(H->owned, nseg);
    H->d_work = dalloc<HmvWork>(H->owned, std::max<size_t>(wk_blk.size(), 1));
    H->d_seg_e0 = dalloc<int32_t>(H->owned, nseg);
    H->d_seg_e1 = dalloc<int32_t>(H->owned, nseg);
    H->d_seg_off = dalloc<int64_t>(H->owned, nseg);
    H->d_rowptr = dalloc<int64_t>(H->owned, n + 1);
    H->d_ent = dalloc<int64_t>(H->owned, ent.size());
    if (!H->d_nacc || !H->d_vt_items || !H->d_item_seg || !H->d_item_r0 ||
        !H->d_seg_nr || !H->d_work || !H->d_seg_e0 || !H->d_seg_e1 || !H->d_seg_off ||
        !H->d_rowptr || !H->d_ent)
        return fail(CCAB_ENOMEM, "H-matrix allocation failed");
    CCAB_HIP(hipMemcpy(H->d_item_seg, item_seg.data(), item_seg.size() * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_item_r0, item_r0.data(), item_r0.size() * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_seg_nr, seg_nr.data(), nseg * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_seg_e0, seg_w0.data(), nseg * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_seg_e1, seg_w1.data(), nseg * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_seg_off, seg_off.data(), nseg * 8, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_rowptr, cnt.data(), (n + 1) * 8, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(H->d_ent, ent.data(), ent.size() * 8, hipMemcpyHostToDevice));
    trace.mark("matvec plan");
    std::vector<void*> release;
    if (prec == CCAB_C64) {
        // FP32 storage: U and V packed to their ranks at new offsets (staged
        // in w_off / f_off for the conversion), dense blocks at new offsets
        // (staged in u2_off); every offset even, so that two consecutive
        // complex-FP32 entries of a column are one aligned 16-byte load
        int64_t cu = s_total, cv = 0, cd = 0;
        for (int32_t b : ll2) {
            HBlkDev& d = H->blk[b];
            d.w_off = d.s_off;
            d.f_off = cv;
            cv = even(cv + d.nc * d.rank);
        }
        for (int32_t b : dl) {
            HBlkDev& d = H->blk[b];
            d.u2_off = cd;
            cd = even(cd + d.nr * d.nc);
        }
        H->d_U32 = dalloc<c64>(H->owned, cu);
        H->d_V32 = dalloc<c64>(H->owned, cv);
        H->d_D32 = dalloc<c64>(H->owned, cd);
        if (!H->d_U32 || !H->d_V32 || !H->d_D32) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
        if (nb)
            hipLaunchKernelGGL(hmat_to_c64_kernel, dim3((unsigned)nb, 4), dim3(256), 0, s, H->d_blk,
                               H->d_U, H->d_V, H->d_D, H->d_U32, H->d_V32, H->d_D32);
        CCAB_HIP(hipGetLastError());
        for (int32_t b : ll2) {
            HBlkDev& d = H->blk[b];
            d.u_off = d.w_off;
            d.v_off = d.f_off;
        }
        for (int32_t b : dl) H->blk[b].d_off = H->blk[b].u2_off;
        CCAB_HIP(hipMemcpyAsync(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice, s));
        CCAB_HIP(hipStreamSynchronize(s));
        release.insert(release.end(), {H->d_U, H->d_V, H->d_D});
        H->d_U = H->d_V = H->d_D = nullptr;
        H->fp32 = true;
        H->mv_bytes *= 0.5;
        trace.mark("FP32 storage");
    } else if (!ll2.empty()) {
        // c128 storage: U moved to the runs' order
        c128* U3 = dalloc<c128>(H->owned, std::max<int64_t>(s_total, 1));
        if (!U3) return fail(CCAB_ENOMEM, "H-matrix allocation failed");
        CCAB_HIP(hipMemcpy(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(repack_u_kernel, dim3((unsigned)ll2.size(), 4), dim3(256), 0, s, H->d_blk,
                           H->d_lr_list, H->d_U, U3);
        CCAB_HIP(hipGetLastError());
        for (int32_t b : ll2) H->blk[b].u_off = H->blk[b].s_off;
        CCAB_HIP(hipMemcpyAsync(H->d_blk, H->blk.data(), nb * sizeof(HBlkDev), hipMemcpyHostToDevice, s));
        CCAB_HIP(hipStreamSynchronize(s));
        release.push_back(H->d_U);
        H->d_U = U3;
        trace.mark("low-rank runs");
    }
    // release the build workspace
    for (void* w : release) {
        auto it = std::find(H->owned.begin(), H->owned.end(), w);
        if (it != H->owned.end()) {
            (void)hipFree(*it);
            H->owned.erase(it);
        }
    }
    // the product's work entries, with the final storage offsets (FP32
    // storage repacks U)
    {
        std::vector<HmvWork> work(wk_blk.size());
        for (size_t e = 0; e < wk_blk.size(); ++e) {
            if (wk_blk[e] >= 0) {
                const HBlkDev& d = H->blk[wk_blk[e]];
                work[e] = HmvWork{d.d_off, d.clo, (int32_t)d.nc, 0};
            } else {
                const HBlkDev& d = H->blk[-1 - wk_blk[e]];
                work[e] = HmvWork{d.u_off, d.ts_off, wk_cnt[e], 1};
            }
        }
        if (!work.empty())
            CCAB_HIP(hipMemcpy(H->d_work, work.data(), work.size() * sizeof(HmvWork),
                               hipMemcpyHostToDevice));
        std::vector<VtItem> items(vt_blk.size());
        for (size_t q = 0; q < items.size(); ++q) {
            const HBlkDev& d = H->blk[vt_blk[q]];
            const int64_t j0 = int64_t(vt_chunk[q]) * kVtChunk;
            items[q] = VtItem{d.v_off + j0, d.clo + j0, d.p_off + int64_t(vt_chunk[q]) * d.rank, d.nc,
                              (int32_t)std::min<int64_t>(kVtChunk, d.nc - j0), d.rank};
        }
        CCAB_HIP(hipMemcpy(H->d_vt_items, items.data(), items.size() * sizeof(VtItem),
                           hipMemcpyHostToDevice));
    }
    trace.mark("workspace release");
    *out = owner.release();
    return CCAB_OK;
}

// cluster tree of the n points (stride doubles apart, coordinates first) and
// its block tree (index order as cluster.rs / block.rs)
void trees(const double* xyz, int64_t stride, int64_t n, int64_t leaf, double eta,
           std::vector<HNode>& T, std::vector<HostBlock>& bl) {
    std::vector<int64_t> all(n);
    for (int64_t i = 0; i < n; ++i) all[i] = i;
    build_cluster(T, xyz, stride, std::move(all), leaf);
    build_blocks(T, 0, 0, eta, bl);
}
}  // namespace

int hmatrix_build(ccab_ctx* ctx, const SysParams& p, const double* d_soa, const double* h_cpts,
                  int64_t n, int side, int64_t leaf, double eta, double tol, ccab_hmatrix** out,
                  int prec) {
    BuildTrace trace(ctx->stream);
    std::vector<HNode> T;
    std::vector<HostBlock> bl;
    trees(h_cpts, 8, n, leaf, eta, T, bl);
    trace.mark("cluster and block trees");
    const BemEntry entry{p, d_soa, n, side == CCAB_LHS ? p.lhs_a : p.rhs_a,
                         side == CCAB_LHS ? p.lhs_b : p.rhs_b};
    return hmatrix_from_blocks(ctx, entry, std::move(T), bl, n, n, tol, out, prec, trace);
}

int hmatrix_build_matrix(ccab_ctx* ctx, const c128* d_A, int64_t lda, const double* h_xyz,
                         int64_t n, int64_t leaf, double eta, double tol, ccab_hmatrix** out,
                         int prec) {
    BuildTrace trace(ctx->stream);
    std::vector<HNode> T;
    std::vector<HostBlock> bl;
    trees(h_xyz, 3, n, leaf, eta, T, bl);
    trace.mark("cluster and block trees");
    return hmatrix_from_blocks(ctx, MatEntry{d_A, lda}, std::move(T), bl, n, n, tol, out, prec, trace);
}

int aca_build_matrix(ccab_ctx* ctx, const c128* d_A, int64_t lda, int64_t m, int64_t ncol,
                     double tol, ccab_hmatrix** out, int prec) {
    BuildTrace trace(ctx->stream);
    // two "clusters", the rows and the columns, and one admissible block
    std::vector<HNode> T(2);
    T[0].idx.resize(m);
    T[1].idx.resize(ncol);
    for (int64_t i = 0; i < m; ++i) T[0].idx[i] = i;
    for (int64_t j = 0; j < ncol; ++j) T[1].idx[j] = j;
    const std::vector<HostBlock> bl{{0, 1, true}};
    return hmatrix_from_blocks(ctx, MatEntry{d_A, lda}, std::move(T), bl, m, ncol, tol, out, prec, trace);
}

int hmatrix_gemv(ccab_ctx* ctx, const ccab_hmatrix* H, c128 alpha, const c128* x, c128 beta,
                 c128* y) {
    hipStream_t s = ctx->stream;
    ProfScope ps(ctx, PROF_HMAT, H->mv_bytes);
    hipLaunchKernelGGL(hmv_gather_kernel, dim3((unsigned)((H->ncol + 255) / 256)), dim3(256), 0, s,
                       H->ncol, H->d_perm, x, H->d_xp);
    // one item per wave (64-thread workgroups), or persistent waves capped
    // per CU (CCAB_HMV_WAVES_PER_CU, four per workgroup)
    const int64_t cap = int64_t(std::max(ctx->num_cus, 1)) * std::max(ctx->hmv_waves_per_cu, 0);
    auto grid_of = [&](int64_t items) {
        return cap > 0 ? dim3((unsigned)((std::min(items, cap) + 3) / 4)) : dim3((unsigned)items);
    };
    const dim3 blk(cap > 0 ? 256 : 64);
    const int pv = ctx->hmv_paired & 1, pn = (ctx->hmv_paired >> 1) & 1;
#define CCAB_HMV_LAUNCH(T, NTF, Vp, Dp, Up, PV, PN)                                              \
    do {                                                                                       \
        if (H->n_vt && (sizeof(T) == 16 || ctx->hmv_vt_ranks == 8))                            \
            hipLaunchKernelGGL((hmv_vt_kernel<T, NTF, 8>), grid_of(H->n_vt), blk, 0, s,        \
                               H->d_vt_items, H->n_vt, Vp, H->d_xp, H->d_t, PV);               \
        else if (H->n_vt)                                                                      \
            hipLaunchKernelGGL((hmv_vt_kernel<T, NTF, 4>), grid_of(H->n_vt), blk, 0, s,        \
                               H->d_vt_items, H->n_vt, Vp, H->d_xp, H->d_t, PV);               \
        if (H->n_tsum)                                                                         \
            hipLaunchKernelGGL(hmv_tsum_kernel, dim3((unsigned)((H->n_tsum + 255) / 256)),     \
                               dim3(256), 0, s, H->d_tsum, H->n_tsum, H->d_t);                 \
        if (H->n_items)                                                                        \
            hipLaunchKernelGGL((hmv_node_kernel<T, NTF>), grid_of(H->n_items), blk, 0, s,      \
                               H->d_item_seg, H->d_item_r0, H->n_items, H->d_seg_nr,           \
                               H->d_seg_e0, H->d_seg_e1, H->d_seg_off, H->d_work, Dp, Up,      \
                               H->d_t, H->d_xp, alpha, H->d_nacc, PN);                         \
    } while (0)
    if (H->fp32) {
        if (ctx->hmv_nt) CCAB_HMV_LAUNCH(c64, true, H->d_V32, H->d_D32, H->d_U32, pv, pn);
        else CCAB_HMV_LAUNCH(c64, false, H->d_V32, H->d_D32, H->d_U32, pv, pn);
    } else {
        if (ctx->hmv_nt) CCAB_HMV_LAUNCH(c128, true, H->d_V, H->d_D, H->d_U, 0, 0);
        else CCAB_HMV_LAUNCH(c128, false, H->d_V, H->d_D, H->d_U, 0, 0);
    }
#undef CCAB_HMV_LAUNCH
    hipLaunchKernelGGL(hmv_reduce_kernel, dim3((unsigned)((H->n + 255) / 256)), dim3(256), 0, s,
                       H->n, H->d_rowptr, H->d_ent, H->d_nacc, beta, y);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

void hmatrix_stats(const ccab_hmatrix* H, int64_t* st, double* norm) {
    int64_t adm = 0, iad = 0, storage = 0, rsum = 0;
    for (const auto& d : H->blk) {
        if (d.adm) {
            adm++;
            storage += int64_t(d.rank) * (d.nr + d.nc);
            rsum += d.rank;
        } else {
            iad++;
            storage += d.nr * d.nc;
        }
    }
    st[0] = adm;
    st[1] = iad;
    st[2] = storage;
    st[3] = rsum;
    st[4] = H->aca_fallbacks;
    if (norm) *norm = H->norm;
}

void hmatrix_stats_ex(const ccab_hmatrix* H, int64_t* st, int nst) {
    int64_t all[12];
    hmatrix_stats(H, all, nullptr);
    for (int k = 0; k < 4; ++k) all[5 + k] = H->aca_bucket[k];
    int64_t rmax = 0;
    for (const auto& d : H->blk)
        if (d.adm) rmax = std::max<int64_t>(rmax, d.rank);
    all[9] = rmax;
    all[10] = H->n;
    all[11] = H->ncol;
    for (int k = 0; k < nst && k < 12; ++k) st[k] = all[k];
}

int hmatrix_to_full(ccab_ctx* ctx, const ccab_hmatrix* H, c128* A, int64_t lda) {
    const int64_t nb = (int64_t)H->blk.size();
    if (nb && H->fp32)
        hipLaunchKernelGGL(hmat_full_kernel<c64>, dim3((unsigned)nb, 4), dim3(256), 0, ctx->stream,
                           H->d_blk, H->d_rows, H->d_cols, H->d_U32, H->d_V32, H->d_D32, A, lda);
    else if (nb)
        hipLaunchKernelGGL(hmat_full_kernel<c128>, dim3((unsigned)nb, 4), dim3(256), 0, ctx->stream,
                           H->d_blk, H->d_rows, H->d_cols, H->d_U, H->d_V, H->d_D, A, lda);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

void hmatrix_free(ccab_hmatrix* H) { delete H; }
int hmatrix_device(const ccab_hmatrix* H) { return H->device; }

int64_t hmatrix_n(const ccab_hmatrix* H) { return H->n; }
int64_t hmatrix_ncol(const ccab_hmatrix* H) { return H->ncol; }
double hmatrix_norm(const ccab_hmatrix* H) { return H->norm; }
