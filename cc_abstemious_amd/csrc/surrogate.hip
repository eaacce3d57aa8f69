// Learned warm start for the surface solve (no counterpart in the reference,
// whose GMRES always starts from zero, gmres.rs:38-39).
//
// A pointwise neural operator: per collocation point, 32 geometric / wave
// features -> 4 linear layers (32 -> W -> W -> W -> 2, SiLU between) ->
// psi = phi / phi_inc, the surface potential relative to the incident
// plane wave.  Training and inference run on this file's kernels:
//   * GEMMs in BF16 on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
//     accumulation and fused epilogues (bias + SiLU; SiLU'; weight-gradient
//     partials), 64x64 tiles staged through LDS;
//   * the 2-wide output layer, loss, bias/weight-gradient reductions and
//     Adam in FP32.
// Storage points in BF16: features, every W1..W3 copy used by a GEMM, the
// pre-activations Z and activations A of the hidden layers, and the
// back-propagated dZ.  Master parameters, W4, biases, gradients and Adam
// moments are FP32.
//
// Parameter layout (FP32, flat): W1[W][32] b1[W] W2[W][W] b2[W] W3[W][W] b3[W]
// W4[2][W] b4[2]; W row-major (out x in).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "ccab_internal.h"

namespace {

constexpr int kFeat = 32;
constexpr int GT = 64;          // GEMM tile (rows and columns)
constexpr int GK = 32;          // GEMM k step (one MFMA)
constexpr int LDSK = GK + 8;    // padded LDS row, in bf16
constexpr int kWgradSplit = 32; // reduction splits of the weight-gradient GEMMs
constexpr int kSumSplit = 256;  // row blocks of the column sums (bias gradients, output layer)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
    return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    return __builtin_bit_cast(float, uint32_t(h) << 16);
}
__device__ inline float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// per system: wave number, unit direction, amplitude A/(i omega rho), body
// centroid and RMS radius (the length scale a of ka)
struct SurSys {
    double k;
    double d[3];
    c128 amp;
    double c0[3];
    double a;
    double mean_area;
};

// ---------------------------------------------------------------------------
// features (and targets psi = phi / phi_inc)
// ---------------------------------------------------------------------------
// the 32 BF16 features of one collocation point (position, normal, area)
__device__ inline void feat_point(const SurSys& q, double x, double y, double z, double nx,
                                  double ny, double nz, double area, uint16_t* o) {
    const double rx = x - q.c0[0], ry = y - q.c0[1], rz = z - q.c0[2];
    const double rl = sqrt(rx * rx + ry * ry + rz * rz);
    const double inv = rl > 0.0 ? 1.0 / rl : 0.0;
    const double hx = rx * inv, hy = ry * inv, hz = rz * inv;
    const double ka = q.k * q.a;
    const double dr = q.d[0] * hx + q.d[1] * hy + q.d[2] * hz;
    const double dn = q.d[0] * nx + q.d[1] * ny + q.d[2] * nz;
    const double nr = nx * hx + ny * hy + nz * hz;
    const double th = acos(fmin(1.0, fmax(-1.0, dr)));
    const double dx = q.d[0] * rx + q.d[1] * ry + q.d[2] * rz;
    double f[kFeat];
    f[0] = ka / 8.0;
    f[1] = f[0] * f[0];
    f[2] = dr;
    f[3] = dn;
    f[4] = nr;
    f[5] = rl / q.a - 1.0;
    f[6] = hx;
    f[7] = hy;
    f[8] = hz;
    f[9] = nx;
    f[10] = ny;
    f[11] = nz;
    f[12] = th / kPi;
    f[13] = cos(ka * dr);
    f[14] = sin(ka * dr);
    f[15] = cos(ka * th);
    f[16] = sin(ka * th);
    f[17] = cos(ka * (kPi - th));
    f[18] = sin(ka * (kPi - th));
    f[19] = cos(2.0 * ka * dr);
    f[20] = sin(2.0 * ka * dr);
    f[21] = cos(ka);
    f[22] = sin(ka);
    f[23] = cos(q.k * (rl - dx));
    f[24] = sin(q.k * (rl - dx));
    f[25] = 1.0 / (1.0 + ka);
    f[26] = area / q.mean_area - 1.0;
#pragma unroll
    for (int t = 27; t < kFeat; ++t) f[t] = 0.0;
#pragma unroll
    for (int t = 0; t < kFeat; ++t) o[t] = f2bf(float(f[t]));
}

// x0 at a point x: amp e^{i k d.x} (out0 + i out1)
__device__ inline ccab_c128 guess_point(const SurSys& q, double x, double y, double z,
                                        const float* out2) {
    double sn, cs;
    sincos(q.k * (q.d[0] * x + q.d[1] * y + q.d[2] * z), &sn, &cs);
    const c128 v = (q.amp * c128{cs, sn}) * c128{double(out2[0]), double(out2[1])};
    return ccab_c128{v.re, v.im};
}

__global__ void feat_kernel(const SurSys* __restrict__ sys, int64_t nsys, const double* __restrict__ soa,
                            int64_t n, const ccab_c128* __restrict__ phi, uint16_t* feat,
                            float* target) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nsys * n) return;
    const int64_t s = p / n, i = p - s * n;
    const SurSys q = sys[s];
    const double x = soa[i], y = soa[n + i], z = soa[2 * n + i];
    feat_point(q, x, y, z, soa[3 * n + i], soa[4 * n + i], soa[5 * n + i], soa[6 * n + i],
               feat + p * kFeat);
    if (phi && target) {
        // psi = phi / (amp e^{i k d.x}), x absolute as in the incident wave
        double sn, cs;
        sincos(q.k * (q.d[0] * x + q.d[1] * y + q.d[2] * z), &sn, &cs);
        const c128 inc = q.amp * c128{cs, sn};
        const c128 ph{phi[p].re, phi[p].im};
        const c128 psi = cdiv(ph, inc);
        target[2 * p] = float(psi.re);
        target[2 * p + 1] = float(psi.im);
    }
}

// targets of system s divided by their RMS over the n points: the network
// learns the shape of psi; the solver fits amplitude and phase of the guess
// to the residual
__global__ __launch_bounds__(256) void normalize_targets_kernel(int64_t n, float* target) {
    __shared__ double red[4];
    float* t = target + int64_t(blockIdx.x) * n * 2;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < 2 * n; i += 256) s += double(t[i]) * double(t[i]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    const double tot = red[0] + red[1] + red[2] + red[3];
    const float inv = tot > 0.0 ? float(1.0 / sqrt(tot / double(n))) : 0.0f;
    for (int64_t i = threadIdx.x; i < 2 * n; i += 256) t[i] *= inv;
}

// x0 = amp e^{i k d.x} (out0 + i out1)
__global__ void predict_finish_kernel(const SurSys* __restrict__ sys, int64_t nsys,
                                      const double* __restrict__ soa, int64_t n,
                                      const float* __restrict__ out, ccab_c128* x0) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nsys * n) return;
    const int64_t s = p / n, i = p - s * n;
    x0[p] = guess_point(sys[s], soa[i], soa[n + i], soa[2 * n + i], out + 2 * p);
}

// body geometry on the device, one workgroup per body: centroid, RMS
// radius and mean element area from the packed records (x y z nx ny nz area
// wt), the same sums in the same order for every call
// (n_all > 0: every system is on the one mesh of points 0 .. n_all-1)
__global__ __launch_bounds__(256) void body_geometry_kernel(SurSys* sys,
                                                            const int64_t* __restrict__ coff,
                                                            int64_t n_all,
                                                            const double* __restrict__ cpts) {
    __shared__ double red[4][4];
    __shared__ double c0s[3];
    const int64_t b = blockIdx.x;
    const int64_t p0 = n_all > 0 ? 0 : coff[b], n = n_all > 0 ? n_all : coff[b + 1] - coff[b];
    const double* r = cpts + 8 * p0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        acc[0] += r[8 * i];
        acc[1] += r[8 * i + 1];
        acc[2] += r[8 * i + 2];
        acc[3] += r[8 * i + 6];
    }
    for (int t = 0; t < 4; ++t) {
        for (int off = 32; off > 0; off >>= 1) acc[t] += __shfl_xor(acc[t], off, 64);
        if (lane == 0) red[t][wave] = acc[t];
    }
    __syncthreads();
    if (threadIdx.x < 3)
        c0s[threadIdx.x] = (red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                            red[threadIdx.x][3]) / double(n);
    const double area = (red[3][0] + red[3][1] + red[3][2] + red[3][3]) / double(n);
    __syncthreads();
    double a2 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double dx = r[8 * i] - c0s[0], dy = r[8 * i + 1] - c0s[1], dz = r[8 * i + 2] - c0s[2];
        a2 += dx * dx + dy * dy + dz * dz;
    }
    for (int off = 32; off > 0; off >>= 1) a2 += __shfl_xor(a2, off, 64);
    __syncthreads();
    if (lane == 0) red[0][wave] = a2;
    __syncthreads();
    if (threadIdx.x == 0) {
        SurSys& q = sys[b];
        q.c0[0] = c0s[0];
        q.c0[1] = c0s[1];
        q.c0[2] = c0s[2];
        q.a = sqrt((red[0][0] + red[0][1] + red[0][2] + red[0][3]) / double(n));
        q.mean_area = area;
    }
}

// bodies of uneven size, points packed back to back: body s owns points
// coff[s] .. coff[s+1]-1 (records x y z nx ny nz area wt)
__device__ inline int64_t body_of(const int64_t* __restrict__ coff, int64_t nbody, int64_t p) {
    int64_t lo = 0, hi = nbody - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (coff[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void feat_bodies_kernel(const SurSys* __restrict__ sys, const int64_t* __restrict__ coff,
                                   int64_t nbody, const double* __restrict__ cpts, uint16_t* feat) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= coff[nbody]) return;
    const double* r = cpts + 8 * p;
    feat_point(sys[body_of(coff, nbody, p)], r[0], r[1], r[2], r[3], r[4], r[5], r[6],
               feat + p * kFeat);
}

__global__ void predict_finish_bodies_kernel(const SurSys* __restrict__ sys,
                                             const int64_t* __restrict__ coff, int64_t nbody,
                                             const double* __restrict__ cpts,
                                             const float* __restrict__ out, ccab_c128* x0) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= coff[nbody]) return;
    const double* r = cpts + 8 * p;
    x0[p] = guess_point(sys[body_of(coff, nbody, p)], r[0], r[1], r[2], out + 2 * p);
}

__global__ void gather_kernel(const uint16_t* __restrict__ feat, const float* __restrict__ target,
                              const int64_t* __restrict__ idx, int64_t B, uint16_t* X, float* T) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= B * kFeat) return;
    const int64_t b = t / kFeat, f = t - b * kFeat;
    const int64_t src = idx[b];
    X[t] = feat[src * kFeat + f];
    if (f < 2) T[2 * b + f] = target[2 * src + f];
}

// ---------------------------------------------------------------------------
// BF16 GEMM on MFMA: C(m, n) = sum_k A(m, k) B(k, n), FP32 accumulation
// ---------------------------------------------------------------------------
enum { EPI_FWD = 0, EPI_BWD = 1, EPI_WGRAD = 2 };

struct GemmArgs {
    const uint16_t* A;
    int64_t sam, sak;  // A(m, k) = A[m sam + k sak]
    const uint16_t* B;
    int64_t sbk, sbn;  // B(k, n) = B[k sbk + n sbn]
    int64_t M, N, K;
    int64_t ldc;
    const float* bias;  // FWD: + bias[n]
    uint16_t* Z;        // FWD: pre-activation (may be null)
    uint16_t* Act;      // FWD: SiLU(z)
    const uint16_t* Zp; // BWD: pre-activation multiplying SiLU'
    uint16_t* dZ;       // BWD: output
    float* part;        // WGRAD: partial sums [split][M][ldc]
    int64_t kchunk;     // WGRAD: reduction range per split (multiple of GK)
};

// 8 consecutive elements along a line with stride st starting at e0 (the
// first `valid` are inside the operand), into dst[0..7]
__device__ inline void load_line8(const uint16_t* __restrict__ base, int64_t e0, int64_t st,
                                  int valid, uint16_t* dst) {
    if (st == 1 && valid >= 8 && ((e0 & 7) == 0)) {
        const uint4 v = *reinterpret_cast<const uint4*>(base + e0);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[j] = h[j];
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[j] = j < valid ? base[e0 + j * st] : uint16_t(0);
    }
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[GT][LDSK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GT][LDSK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t m0 = int64_t(blockIdx.x) * GT, n0 = int64_t(blockIdx.y) * GT;
    int64_t kb = 0, ke = g.K;
    if (EPI == EPI_WGRAD) {
        kb = int64_t(blockIdx.z) * g.kchunk;
        ke = std::min<int64_t>(g.K, kb + g.kchunk);
    }
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bool a_kc = g.sak == 1;  // A contiguous along k
    const bool b_kc = g.sbk == 1;  // B contiguous along k
    for (int64_t k0 = kb; k0 < ke; k0 += GK) {
        __attribute__((aligned(16))) uint16_t v[8];
        // A tile: As[m][k]
        if (a_kc) {
            const int row = t >> 2, ks = (t & 3) * 8;
            const int64_t gm = m0 + row, gk = k0 + ks;
            const int valid = gm < g.M ? int(std::max<int64_t>(0, std::min<int64_t>(8, ke - gk))) : 0;
            load_line8(g.A, gm * g.sam + gk, 1, valid, v);
            *reinterpret_cast<uint4*>(&As[row][ks]) = *reinterpret_cast<const uint4*>(v);
        } else {
            const int kk = t >> 3, ms = (t & 7) * 8;
            const int64_t gk = k0 + kk, gm = m0 + ms;
            const int valid = gk < ke ? int(std::max<int64_t>(0, std::min<int64_t>(8, g.M - gm))) : 0;
            load_line8(g.A, gm * g.sam + gk * g.sak, g.sam, valid, v);
#pragma unroll
            for (int j = 0; j < 8; ++j) As[ms + j][kk] = v[j];
        }
        // B tile: Bs[n][k]
        if (b_kc) {
            const int row = t >> 2, ks = (t & 3) * 8;
            const int64_t gn = n0 + row, gk = k0 + ks;
            const int valid = gn < g.N ? int(std::max<int64_t>(0, std::min<int64_t>(8, ke - gk))) : 0;
            load_line8(g.B, gn * g.sbn + gk, 1, valid, v);
            *reinterpret_cast<uint4*>(&Bs[row][ks]) = *reinterpret_cast<const uint4*>(v);
        } else {
            const int kk = t >> 3, ns = (t & 7) * 8;
            const int64_t gk = k0 + kk, gn = n0 + ns;
            const int valid = gk < ke ? int(std::max<int64_t>(0, std::min<int64_t>(8, g.N - gn))) : 0;
            load_line8(g.B, gk * g.sbk + gn * g.sbn, g.sbn, valid, v);
#pragma unroll
            for (int j = 0; j < 8; ++j) Bs[ns + j][kk] = v[j];
        }
        __syncthreads();
        const int lr = lane & 15, lk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + lr][lk]);
            b[i] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * i + lr][lk]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // epilogue: lane holds rows 4 (lane / 16) + v, column lane % 16 of each 16x16 block
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int64_t row = m0 + wm + 16 * i + 4 * (lane >> 4) + vv;
                const int64_t col = n0 + wn + 16 * j + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                const float c = acc[i][j][vv];
                const int64_t o = row * g.ldc + col;
                if (EPI == EPI_FWD) {
                    const float z = c + g.bias[col];
                    if (g.Z) g.Z[o] = f2bf(z);
                    g.Act[o] = f2bf(z * sigm(z));
                } else if (EPI == EPI_BWD) {
                    const float zp = bf2f(g.Zp[o]);
                    const float s = sigm(zp);
                    g.dZ[o] = f2bf(c * s * (1.0f + zp * (1.0f - s)));
                } else {
                    g.part[int64_t(blockIdx.z) * g.M * g.ldc + o] = c;
                }
            }
}

// ---------------------------------------------------------------------------
// output layer (W -> 2, FP32), loss, its backward, reductions, Adam
// ---------------------------------------------------------------------------
// out[p] = W4 a3[p] + b4; with a target: dout = 2 (out - t) / B and the
// per-block partial of sum |out - t|^2.  One wavefront per point.
__global__ __launch_bounds__(256) void last_fwd_kernel(const uint16_t* __restrict__ A3, int64_t P,
                                                       int W, const float* __restrict__ W4,
                                                       const float* __restrict__ b4, float* out,
                                                       const float* __restrict__ target,
                                                       float* dout, float inv_b, double* loss_part) {
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double lsum = 0.0;
    for (int64_t p = int64_t(blockIdx.x) * 4 + w; p < P; p += int64_t(gridDim.x) * 4) {
        float s0 = 0.f, s1 = 0.f;
        for (int k = lane; k < W; k += 64) {
            const float a = bf2f(A3[p * W + k]);
            s0 += a * W4[k];
            s1 += a * W4[W + k];
        }
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off, 64);
            s1 += __shfl_xor(s1, off, 64);
        }
        s0 += b4[0];
        s1 += b4[1];
        if (lane == 0) {
            out[2 * p] = s0;
            out[2 * p + 1] = s1;
            if (target) {
                const float e0 = s0 - target[2 * p], e1 = s1 - target[2 * p + 1];
                dout[2 * p] = 2.0f * e0 * inv_b;
                dout[2 * p + 1] = 2.0f * e1 * inv_b;
                lsum += double(e0) * e0 + double(e1) * e1;
            }
        }
    }
    if (loss_part) {
        if (lane == 0) red[w] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

// dZ3 = (dout W4) * SiLU'(Z3)
__global__ void last_bwd_kernel(const float* __restrict__ dout, const float* __restrict__ W4,
                                const uint16_t* __restrict__ Z3, int64_t P, int W, uint16_t* dZ3) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= P * W) return;
    const int64_t p = e / W;
    const int k = int(e - p * W);
    const float g = dout[2 * p] * W4[k] + dout[2 * p + 1] * W4[W + k];
    const float z = bf2f(Z3[e]);
    const float s = sigm(z);
    dZ3[e] = f2bf(g * s * (1.0f + z * (1.0f - s)));
}

// partial column sums over a row chunk: bias gradients of the hidden layers
// (colsum of dZ) or, with A3, the output layer's weight gradient.
// out layout per split: [ncols] (bias) or [2][W] + [2] (output layer)
// column sums of dZ (P x W BF16, row-major) over the rows of row block
// blockIdx.x (chunk rows), one thread per column, 16 rows in flight
__global__ void colsum_kernel(const uint16_t* __restrict__ dZ, int64_t P, int W, int64_t chunk,
                              float* part) {
    const int k = threadIdx.x;
    const int64_t p0 = int64_t(blockIdx.x) * chunk, p1 = std::min(P, p0 + chunk);
    float s = 0.f;
    int64_t p = p0;
    for (; p + 15 < p1; p += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = bf2f(dZ[(p + u) * W + k]);
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; p < p1; ++p) s += bf2f(dZ[p * W + k]);
    part[int64_t(blockIdx.x) * W + k] = s;
}

// output layer over the rows of row block blockIdx.x: dW4[c][k] partials
// sum_p dout[p][c] A3[p][k] and db4[c] partials sum_p dout[p][c]
__global__ void last_wgrad_kernel(const float* __restrict__ dout, const uint16_t* __restrict__ A3,
                                  int64_t P, int W, int64_t chunk, float* part) {
    const int k = threadIdx.x;
    const int64_t p0 = int64_t(blockIdx.x) * chunk, p1 = std::min(P, p0 + chunk);
    float s0 = 0.f, s1 = 0.f, d0 = 0.f, d1 = 0.f;
    int64_t p = p0;
    for (; p + 7 < p1; p += 8) {
        float a[8];
        float2 g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = bf2f(A3[(p + u) * W + k]);
            g[u] = reinterpret_cast<const float2*>(dout)[p + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            s0 += g[u].x * a[u];
            s1 += g[u].y * a[u];
            d0 += g[u].x;
            d1 += g[u].y;
        }
    }
    for (; p < p1; ++p) {
        const float a = bf2f(A3[p * W + k]);
        s0 += dout[2 * p] * a;
        s1 += dout[2 * p + 1] * a;
        d0 += dout[2 * p];
        d1 += dout[2 * p + 1];
    }
    float* o = part + int64_t(blockIdx.x) * (2 * W + 2);
    o[k] = s0;
    o[W + k] = s1;
    if (k == 0) {
        o[2 * W] = d0;
        o[2 * W + 1] = d1;
    }
}

// dst[e] = the sum of the nsplit partials part[q count + e]: one workgroup
// per entry, a fixed-order tree
__global__ __launch_bounds__(256) void reduce_splits_tree_kernel(const float* __restrict__ part,
                                                                 int nsplit, int64_t count,
                                                                 float* dst) {
    __shared__ float red[4];
    const int64_t e = blockIdx.x;
    float s = 0.f;
    for (int q = threadIdx.x; q < nsplit; q += 256) s += part[int64_t(q) * count + e];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) dst[e] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void reduce_splits_kernel(const float* __restrict__ part, int nsplit, int64_t count,
                                     float* dst) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float s = 0.f;
    for (int q = 0; q < nsplit; ++q) s += part[int64_t(q) * count + e];
    dst[e] = s;
}

__global__ void sum_doubles_kernel(const double* __restrict__ part, int cnt, double* out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < cnt; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam (decoupled weight decay on weights only) and the BF16 weight copies
__global__ void adam_kernel(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ grad, uint16_t* __restrict__ prm16,
                            const uint8_t* __restrict__ is_weight, int64_t count, float lr,
                            float b1, float b2, float eps, float wd, float c1, float c2) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    float p = prm[i];
    if (is_weight[i]) p -= lr * wd * p;
    p -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    prm[i] = p;
    prm16[i] = f2bf(p);
}

__global__ void to_bf16_kernel(const float* __restrict__ src, int64_t count, uint16_t* dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = f2bf(src[i]);
}

inline dim3 grid1d(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

struct ccab_surrogate {
    int device = 0;
    int W = 256;
    ccab_surrogate_cfg cfg{};
    int64_t nparam = 0;
    int64_t off_w[4], off_b[4];
    int64_t step = 0;
    std::vector<void*> owned;
    float* prm = nullptr;     // FP32 master
    uint16_t* prm16 = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    uint8_t* is_weight = nullptr;
    // activation workspace (grows)
    int64_t cap = 0;
    uint16_t *X = nullptr, *Z[3] = {}, *A[3] = {}, *dZa = nullptr, *dZb = nullptr;
    float *T = nullptr, *out = nullptr, *dout = nullptr;
    float* part = nullptr;
    int64_t part_cap = 0;
    double* loss_part = nullptr;
    double* loss = nullptr;
    int64_t* idx = nullptr;
    int64_t idx_cap = 0;
    ~ccab_surrogate() {
        for (void* p : owned) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        owned.push_back(p);
        return static_cast<T*>(p);
    }
    void release(void* p) {
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) {
            (void)hipFree(*it);
            owned.erase(it);
        }
    }
};

namespace {

int sur_reserve(ccab_ctx* ctx, ccab_surrogate* S, int64_t P) {
    if (P <= S->cap) return CCAB_OK;
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    for (void* p : {static_cast<void*>(S->X), static_cast<void*>(S->Z[0]), static_cast<void*>(S->Z[1]),
                    static_cast<void*>(S->Z[2]), static_cast<void*>(S->A[0]), static_cast<void*>(S->A[1]),
                    static_cast<void*>(S->A[2]), static_cast<void*>(S->dZa), static_cast<void*>(S->dZb),
                    static_cast<void*>(S->T), static_cast<void*>(S->out), static_cast<void*>(S->dout),
                    static_cast<void*>(S->loss_part)})
        if (p) S->release(p);
    const int64_t cap = std::max<int64_t>(P, 4096);
    const int W = S->W;
    S->X = S->alloc<uint16_t>(cap * kFeat);
    for (int l = 0; l < 3; ++l) {
        S->Z[l] = S->alloc<uint16_t>(cap * W);
        S->A[l] = S->alloc<uint16_t>(cap * W);
    }
    S->dZa = S->alloc<uint16_t>(cap * W);
    S->dZb = S->alloc<uint16_t>(cap * W);
    S->T = S->alloc<float>(cap * 2);
    S->out = S->alloc<float>(cap * 2);
    S->dout = S->alloc<float>(cap * 2);
    S->loss_part = S->alloc<double>(4096);
    if (!S->X || !S->Z[2] || !S->A[2] || !S->dZa || !S->dZb || !S->T || !S->out || !S->dout ||
        !S->loss_part)
        return ctx_fail(ctx, CCAB_ENOMEM, "surrogate workspace allocation failed");
    S->cap = cap;
    return CCAB_OK;
}

int gemm(ccab_ctx* ctx, int epi, const GemmArgs& g) {
    const int64_t nsplit = epi == EPI_WGRAD ? (g.K + g.kchunk - 1) / g.kchunk : 1;
    dim3 grid((unsigned)((g.M + GT - 1) / GT), (unsigned)((g.N + GT - 1) / GT), (unsigned)nsplit);
    if (epi == EPI_FWD) hipLaunchKernelGGL(gemm_bf16_kernel<EPI_FWD>, grid, dim3(256), 0, ctx->stream, g);
    else if (epi == EPI_BWD) hipLaunchKernelGGL(gemm_bf16_kernel<EPI_BWD>, grid, dim3(256), 0, ctx->stream, g);
    else hipLaunchKernelGGL(gemm_bf16_kernel<EPI_WGRAD>, grid, dim3(256), 0, ctx->stream, g);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

// forward over P rows of S->X; target/dout/loss when T != null
int sur_forward(ccab_ctx* ctx, ccab_surrogate* S, int64_t P, const float* T, float inv_b,
                int nloss_blocks) {
    const int W = S->W;
    const uint16_t* in = S->X;
    int64_t K = kFeat;
    for (int l = 0; l < 3; ++l) {
        GemmArgs g{};
        g.A = in;
        g.sam = K;
        g.sak = 1;
        g.B = S->prm16 + S->off_w[l];  // B(k, n) = W[n][k]
        g.sbk = 1;
        g.sbn = K;
        g.M = P;
        g.N = W;
        g.K = K;
        g.ldc = W;
        g.bias = S->prm + S->off_b[l];
        g.Z = S->Z[l];
        g.Act = S->A[l];
        CCAB_TRY(gemm(ctx, EPI_FWD, g));
        in = S->A[l];
        K = W;
    }
    hipLaunchKernelGGL(last_fwd_kernel, dim3(nloss_blocks), dim3(256), 0, ctx->stream, S->A[2], P, W,
                       S->prm + S->off_w[3], S->prm + S->off_b[3], S->out, T, S->dout, inv_b,
                       T ? S->loss_part : nullptr);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

// backward over P rows after sur_forward with a target: S->grad
int sur_backward(ccab_ctx* ctx, ccab_surrogate* S, int64_t P) {
    const int W = S->W;
    const int64_t chunk = (P + kWgradSplit - 1) / kWgradSplit;
    const int64_t schunk = (P + kSumSplit - 1) / kSumSplit;
    const int64_t need = std::max(int64_t(kWgradSplit) * W * W, int64_t(kSumSplit) * (2 * W + 2));
    if (S->part_cap < need) {
        CCAB_HIP(hipStreamSynchronize(ctx->stream));
        if (S->part) S->release(S->part);
        S->part = S->alloc<float>(need);
        if (!S->part) return ctx_fail(ctx, CCAB_ENOMEM, "surrogate workspace allocation failed");
        S->part_cap = need;
    }
    hipStream_t s = ctx->stream;
    // output layer: dW4, db4; dZ3
    hipLaunchKernelGGL(last_wgrad_kernel, dim3(kSumSplit), dim3(W), 0, s, S->dout, S->A[2], P, W,
                       schunk, S->part);
    hipLaunchKernelGGL(reduce_splits_tree_kernel, dim3(unsigned(2 * W + 2)), dim3(256), 0, s,
                       S->part, kSumSplit, int64_t(2 * W + 2), S->grad + S->off_w[3]);
    uint16_t* dZ = S->dZa;
    uint16_t* dZn = S->dZb;
    hipLaunchKernelGGL(last_bwd_kernel, grid1d(P * W), dim3(256), 0, s, S->dout,
                       S->prm + S->off_w[3], S->Z[2], P, W, dZ);
    for (int l = 2; l >= 0; --l) {
        const int64_t K = l == 0 ? kFeat : W;
        const uint16_t* in = l == 0 ? S->X : S->A[l - 1];
        // dW_l[n][k] = sum_p dZ[p][n] in[p][k]
        GemmArgs g{};
        g.A = dZ;
        g.sam = 1;
        g.sak = W;
        g.B = in;
        g.sbk = K;
        g.sbn = 1;
        g.M = W;
        g.N = K;
        g.K = P;
        g.ldc = K;
        g.part = S->part;
        g.kchunk = ((chunk + GK - 1) / GK) * GK;
        CCAB_TRY(gemm(ctx, EPI_WGRAD, g));
        const int nsp = int((P + g.kchunk - 1) / g.kchunk);
        hipLaunchKernelGGL(reduce_splits_kernel, grid1d(W * K), dim3(256), 0, s, S->part, nsp,
                           int64_t(W) * K, S->grad + S->off_w[l]);
        // db_l = column sums of dZ (the part buffer is free again after the reduction)
        hipLaunchKernelGGL(colsum_kernel, dim3(kSumSplit), dim3(W), 0, s, dZ, P, W, schunk, S->part);
        hipLaunchKernelGGL(reduce_splits_tree_kernel, dim3(unsigned(W)), dim3(256), 0, s, S->part,
                           kSumSplit, int64_t(W), S->grad + S->off_b[l]);
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1})
            GemmArgs h{};
            h.A = dZ;
            h.sam = W;
            h.sak = 1;
            h.B = S->prm16 + S->off_w[l];  // B(k = n_out, n = k_in) = W[n_out][k_in]
            h.sbk = K;
            h.sbn = 1;
            h.M = P;
            h.N = K;
            h.K = W;
            h.ldc = K;
            h.Zp = S->Z[l - 1];
            h.dZ = dZn;
            CCAB_TRY(gemm(ctx, EPI_BWD, h));
            std::swap(dZ, dZn);
        }
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int loss_blocks(int64_t P) { return (int)std::min<int64_t>(4096, (P + 3) / 4); }

int make_sys(ccab_ctx* ctx, const ccab_physics& ph, const ccab_wave& w, const double c0[3],
             double a, double area, SurSys& q) {
    CCAB_TRY(check_physics(ctx, ph));
    CCAB_CHECK(w.kind == CCAB_WAVE_PLANE, "the surrogate models plane-wave incidence");
    const SysParams p = make_params(ph);
    double d[3] = {w.vec[0], w.vec[1], w.vec[2]};
    const double nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    CCAB_CHECK(nn > 0.0, "incident wave vector has zero length");
    q.k = p.k;
    for (int t = 0; t < 3; ++t) {
        q.d[t] = d[t] / nn;
        q.c0[t] = c0[t];
    }
    q.amp = cdiv(c128{w.amplitude[0], w.amplitude[1]}, c128{0.0, p.omega * ph.mass_density});
    q.a = a;
    q.mean_area = area;
    return CCAB_OK;
}

int upload_sys(ccab_ctx* ctx, const std::vector<SurSys>& hs, SurSys** d_sys) {
    *d_sys = static_cast<SurSys*>(ctx_buf(ctx, SLOT_SUR_SYS, sizeof(SurSys) * hs.size()));
    if (!*d_sys) return CCAB_ENOMEM;
    CCAB_HIP(hipMemcpyAsync(*d_sys, hs.data(), sizeof(SurSys) * hs.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

// per-system constants on the device (nsys systems on one mesh): wave and
// physics from the host, the mesh's centroid, RMS radius and mean element
// area by body_geometry_kernel
int stage_sys(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w, int64_t nsys,
              const double* d_cpts, int64_t n, SurSys** d_sys) {
    const double none[3] = {0.0, 0.0, 0.0};
    std::vector<SurSys> hs(nsys);
    for (int64_t s = 0; s < nsys; ++s) CCAB_TRY(make_sys(ctx, ph[s], w[s], none, 1.0, 1.0, hs[s]));
    CCAB_TRY(upload_sys(ctx, hs, d_sys));
    hipLaunchKernelGGL(body_geometry_kernel, dim3((unsigned)nsys), dim3(256), 0, ctx->stream,
                       *d_sys, static_cast<const int64_t*>(nullptr), n, d_cpts);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

}  // namespace

extern "C" {

int ccab_surrogate_create(ccab_ctx* ctx, const ccab_surrogate_cfg* cfg, ccab_surrogate** out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(cfg && out, "invalid arguments");
    *out = nullptr;
    const int W = cfg->width <= 0 ? 256 : cfg->width;
    CCAB_CHECK(W % 64 == 0 && W <= 1024, "width must be a multiple of 64, at most 1024");
    std::unique_ptr<ccab_surrogate> S(new ccab_surrogate());
    S->device = ctx->device;
    S->W = W;
    S->cfg = *cfg;
    S->cfg.width = W;
    int64_t o = 0;
    const int64_t fan_in[4] = {kFeat, W, W, W}, fan_out[4] = {W, W, W, 2};
    for (int l = 0; l < 4; ++l) {
        S->off_w[l] = o;
        o += fan_out[l] * fan_in[l];
        S->off_b[l] = o;
        o += fan_out[l];
    }
    S->nparam = o;
    S->prm = S->alloc<float>(o);
    S->prm16 = S->alloc<uint16_t>(o);
    S->m = S->alloc<float>(o);
    S->v = S->alloc<float>(o);
    S->grad = S->alloc<float>(o);
    S->is_weight = S->alloc<uint8_t>(o);
    if (!S->prm || !S->prm16 || !S->m || !S->v || !S->grad || !S->is_weight)
        return ctx_fail(ctx, CCAB_ENOMEM, "surrogate allocation failed");
    // initialisation: weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0
    std::vector<float> h(o, 0.0f);
    std::vector<uint8_t> isw(o, 0);
    std::mt19937_64 rng(uint64_t(cfg->seed) * 0x9E3779B97F4A7C15ull + 12345u);
    for (int l = 0; l < 4; ++l) {
        const double bound = 1.0 / std::sqrt(double(fan_in[l]));
        std::uniform_real_distribution<double> U(-bound, bound);
        for (int64_t e = 0; e < fan_out[l] * fan_in[l]; ++e) {
            h[S->off_w[l] + e] = float(U(rng));
            isw[S->off_w[l] + e] = 1;
        }
    }
    CCAB_HIP(hipMemcpy(S->prm, h.data(), o * 4, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemcpy(S->is_weight, isw.data(), o, hipMemcpyHostToDevice));
    CCAB_HIP(hipMemset(S->m, 0, o * 4));
    CCAB_HIP(hipMemset(S->v, 0, o * 4));
    hipLaunchKernelGGL(to_bf16_kernel, grid1d(o), dim3(256), 0, ctx->stream, S->prm, o, S->prm16);
    CCAB_HIP(hipGetLastError());
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    *out = S.release();
    return CCAB_OK;
}

int ccab_surrogate_destroy(ccab_surrogate* s) {
    delete s;
    return CCAB_OK;
}

int64_t ccab_surrogate_num_params(const ccab_surrogate* s) { return s ? s->nparam : -1; }

int ccab_surrogate_get_params(ccab_ctx* ctx, const ccab_surrogate* s, float* h_params) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && h_params, "invalid arguments");
    CCAB_HIP(hipMemcpyAsync(h_params, s->prm, s->nparam * 4, hipMemcpyDeviceToHost, ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

int ccab_surrogate_set_params(ccab_ctx* ctx, ccab_surrogate* s, const float* h_params) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && h_params, "invalid arguments");
    CCAB_HIP(hipMemcpyAsync(s->prm, h_params, s->nparam * 4, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(to_bf16_kernel, grid1d(s->nparam), dim3(256), 0, ctx->stream, s->prm,
                       s->nparam, s->prm16);
    CCAB_HIP(hipMemsetAsync(s->m, 0, s->nparam * 4, ctx->stream));
    CCAB_HIP(hipMemsetAsync(s->v, 0, s->nparam * 4, ctx->stream));
    s->step = 0;
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    return CCAB_OK;
}

int ccab_surrogate_features(ccab_ctx* ctx, const ccab_physics* ph, const ccab_wave* w, int64_t nsys,
                            const double* d_cpts, int64_t n, const ccab_c128* d_phi,
                            uint16_t* d_feat, float* d_target) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(ph && w && d_cpts && d_feat && n > 0 && nsys > 0, "invalid arguments");
    CCAB_CHECK(!d_phi == !d_target, "d_phi and d_target go together");
    SurSys* sys;
    CCAB_TRY(stage_sys(ctx, ph, w, nsys, d_cpts, n, &sys));
    double* soa;
    soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(n) * 64));
    if (!soa) return CCAB_ENOMEM;
    CCAB_TRY(launch_cpts_to_soa(ctx, d_cpts, n, soa));
    hipLaunchKernelGGL(feat_kernel, grid1d(nsys * n), dim3(256), 0, ctx->stream, sys, nsys, soa, n,
                       d_phi, d_feat, d_target);
    if (d_target)
        hipLaunchKernelGGL(normalize_targets_kernel, dim3((unsigned)nsys), dim3(256), 0, ctx->stream,
                           n, d_target);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int ccab_surrogate_loss_grad(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat,
                             const float* d_target, int64_t P, double* loss, float* h_grad) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && d_feat && d_target && P > 0, "invalid arguments");
    CCAB_CHECK(s->device == ctx->device, "surrogate lives on another device");
    CCAB_TRY(sur_reserve(ctx, s, P));
    CCAB_HIP(hipMemcpyAsync(s->X, d_feat, size_t(P) * kFeat * 2, hipMemcpyDeviceToDevice, ctx->stream));
    const int nb = loss_blocks(P);
    CCAB_TRY(sur_forward(ctx, s, P, d_target, 1.0f / float(P), nb));
    CCAB_TRY(sur_backward(ctx, s, P));
    double* dl = static_cast<double*>(ctx_buf(ctx, SLOT_SUR_LOSS, 64));
    if (!dl) return CCAB_ENOMEM;
    hipLaunchKernelGGL(sum_doubles_kernel, dim3(1), dim3(256), 0, ctx->stream, s->loss_part, nb, dl);
    double hl = 0.0;
    CCAB_HIP(hipMemcpyAsync(&hl, dl, 8, hipMemcpyDeviceToHost, ctx->stream));
    if (h_grad)
        CCAB_HIP(hipMemcpyAsync(h_grad, s->grad, s->nparam * 4, hipMemcpyDeviceToHost, ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = hl / double(P);
    return CCAB_OK;
}

int ccab_surrogate_train_step(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat,
                              const float* d_target, const int64_t* d_index, int64_t batch,
                              double* loss) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && d_feat && d_target && d_index && batch > 0, "invalid arguments");
    CCAB_CHECK(s->device == ctx->device, "surrogate lives on another device");
    CCAB_TRY(sur_reserve(ctx, s, batch));
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(gather_kernel, grid1d(batch * kFeat), dim3(256), 0, st, d_feat, d_target,
                       d_index, batch, s->X, s->T);
    const int nb = loss_blocks(batch);
    CCAB_TRY(sur_forward(ctx, s, batch, s->T, 1.0f / float(batch), nb));
    CCAB_TRY(sur_backward(ctx, s, batch));
    s->step++;
    const ccab_surrogate_cfg& c = s->cfg;
    const float b1 = float(c.beta1 > 0 ? c.beta1 : 0.9), b2 = float(c.beta2 > 0 ? c.beta2 : 0.999);
    const float c1 = float(1.0 - std::pow(double(b1), double(s->step)));
    const float c2 = float(1.0 - std::pow(double(b2), double(s->step)));
    hipLaunchKernelGGL(adam_kernel, grid1d(s->nparam), dim3(256), 0, st, s->prm, s->m, s->v, s->grad,
                       s->prm16, s->is_weight, s->nparam, float(c.lr > 0 ? c.lr : 1e-3), b1, b2,
                       float(c.eps > 0 ? c.eps : 1e-8), float(c.weight_decay), c1, c2);
    CCAB_HIP(hipGetLastError());
    if (loss) {
        double* dl = static_cast<double*>(ctx_buf(ctx, SLOT_SUR_LOSS, 64));
        if (!dl) return CCAB_ENOMEM;
        hipLaunchKernelGGL(sum_doubles_kernel, dim3(1), dim3(256), 0, st, s->loss_part, nb, dl);
        double hl = 0.0;
        CCAB_HIP(hipMemcpyAsync(&hl, dl, 8, hipMemcpyDeviceToHost, st));
        CCAB_HIP(hipStreamSynchronize(st));
        *loss = hl / double(batch);
    }
    return CCAB_OK;
}

int ccab_surrogate_set_lr(ccab_surrogate* s, double lr) {
    if (!s || !(lr > 0.0)) return CCAB_EINVAL;
    s->cfg.lr = lr;
    return CCAB_OK;
}

int ccab_surrogate_forward(ccab_ctx* ctx, ccab_surrogate* s, const uint16_t* d_feat, int64_t P,
                           float* d_out) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && d_feat && d_out && P > 0, "invalid arguments");
    CCAB_CHECK(s->device == ctx->device, "surrogate lives on another device");
    CCAB_TRY(sur_reserve(ctx, s, P));
    CCAB_HIP(hipMemcpyAsync(s->X, d_feat, size_t(P) * kFeat * 2, hipMemcpyDeviceToDevice, ctx->stream));
    CCAB_TRY(sur_forward(ctx, s, P, nullptr, 0.0f, loss_blocks(P)));
    CCAB_HIP(hipMemcpyAsync(d_out, s->out, size_t(P) * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return CCAB_OK;
}

int ccab_surrogate_predict_bodies(ccab_ctx* ctx, ccab_surrogate* s, int64_t nbody,
                                  const int64_t* n, const double* d_cpts, const ccab_physics* ph,
                                  const ccab_wave* w, ccab_c128* d_x0) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && nbody > 0 && n && d_cpts && ph && w && d_x0, "invalid arguments");
    CCAB_CHECK(s->device == ctx->device, "surrogate lives on another device");
    std::vector<int64_t> coff(nbody + 1, 0);
    for (int64_t b = 0; b < nbody; ++b) {
        CCAB_CHECK(n[b] > 0, "invalid body size");
        coff[b + 1] = coff[b] + n[b];
    }
    const int64_t P = coff[nbody];
    CCAB_TRY(sur_reserve(ctx, s, P));
    // wave and physics on the host; body geometry on the device
    const double none[3] = {0.0, 0.0, 0.0};
    std::vector<SurSys> hs(nbody);
    for (int64_t b = 0; b < nbody; ++b) CCAB_TRY(make_sys(ctx, ph[b], w[b], none, 1.0, 1.0, hs[b]));
    SurSys* sys;
    CCAB_TRY(upload_sys(ctx, hs, &sys));
    int64_t* dco = static_cast<int64_t*>(ctx_buf(ctx, SLOT_RAGGED, size_t(nbody + 1) * 8));
    if (!dco) return CCAB_ENOMEM;
    CCAB_HIP(hipMemcpy(dco, coff.data(), size_t(nbody + 1) * 8, hipMemcpyHostToDevice));
    hipStream_t st = ctx->stream;
    const double W = s->W;
    ProfScope ps(ctx, PROF_SURROGATE, 2.0 * double(P) * (kFeat * W + 2.0 * W * W + 2.0 * W));
    hipLaunchKernelGGL(body_geometry_kernel, dim3((unsigned)nbody), dim3(256), 0, st, sys, dco,
                       int64_t(0), d_cpts);
    hipLaunchKernelGGL(feat_bodies_kernel, grid1d(P), dim3(256), 0, st, sys, dco, nbody, d_cpts,
                       s->X);
    CCAB_TRY(sur_forward(ctx, s, P, nullptr, 0.0f, loss_blocks(P)));
    hipLaunchKernelGGL(predict_finish_bodies_kernel, grid1d(P), dim3(256), 0, st, sys, dco, nbody,
                       d_cpts, s->out, d_x0);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int ccab_surrogate_predict(ccab_ctx* ctx, ccab_surrogate* s, const ccab_physics* ph,
                           const ccab_wave* w, int64_t nsys, const double* d_cpts, int64_t n,
                           ccab_c128* d_x0) {
    if (!ctx) return CCAB_EINVAL;
    CCAB_CHECK(s && ph && w && d_cpts && d_x0 && n > 0 && nsys > 0, "invalid arguments");
    CCAB_CHECK(s->device == ctx->device, "surrogate lives on another device");
    const int64_t P = nsys * n;
    CCAB_TRY(sur_reserve(ctx, s, P));
    SurSys* sys;
    CCAB_TRY(stage_sys(ctx, ph, w, nsys, d_cpts, n, &sys));
    double* soa = static_cast<double*>(ctx_buf(ctx, SLOT_SOA, size_t(n) * 64));
    if (!soa) return CCAB_ENOMEM;
    CCAB_TRY(launch_cpts_to_soa(ctx, d_cpts, n, soa));
    hipStream_t st = ctx->stream;
    const double W = s->W;
    ProfScope ps(ctx, PROF_SURROGATE, 2.0 * double(P) * (kFeat * W + 2.0 * W * W + 2.0 * W));
    hipLaunchKernelGGL(feat_kernel, grid1d(P), dim3(256), 0, st, sys, nsys, soa, n, nullptr, s->X,
                       nullptr);
    CCAB_TRY(sur_forward(ctx, s, P, nullptr, 0.0f, loss_blocks(P)));
    hipLaunchKernelGGL(predict_finish_kernel, grid1d(P), dim3(256), 0, st, sys, nsys, soa, n, s->out,
                       d_x0);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

}  // extern "C"
