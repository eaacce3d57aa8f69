// Influence-matrix assembly on gfx950.
//
// Implements the Helmholtz kernel of influence_matrix.rs:12-60 (one-point
// centroid collocation, the flat-disk singular diagonal, Burton-Miller
// coupling with the reference's dh sign at influence_matrix.rs:30) and the
// dense assembly of influence_matrix.rs:64-98, fused with the BC algebra of
// solve.rs:24-71.
//
// Layout: matrices are column-major (entry (i,j) at i + j*ld).  A workgroup
// of 256 lanes owns 256 consecutive rows i and a chunk of CJ columns j; each
// lane keeps its row's collocation point in registers and walks the columns,
// whose geometry is wave-uniform (scalar loads).  Every store of a wavefront
// is 64 consecutive entries of one column: 512 B (c64) / 1 KiB (c128), full
// 128-byte lines.  The kernel is HBM-store bound: 8 B (c64) or 16 B (c128)
// per stored entry and matrix.
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "ccab_internal.h"
#include "greens.h"

namespace {

constexpr int kRows = 256;  // rows per workgroup (4 wavefronts)
constexpr int kColChunk = 64;

// FP16 with saturation at the largest finite half
__device__ __forceinline__ uint16_t to_half_sat(double v) {
    const float f = fminf(fmaxf(float(v), -65504.0f), 65504.0f);
    return __half_as_ushort(__float2half_rn(f));
}

template <typename T>
__device__ __forceinline__ void store(void* base, int64_t off, c128 v, double h_scale = 1.0) {
    if constexpr (sizeof(T) == 16) {
        reinterpret_cast<c128*>(base)[off] = v;
    } else if constexpr (sizeof(T) == 8) {
        reinterpret_cast<c64*>(base)[off] = c64{static_cast<float>(v.re), static_cast<float>(v.im)};
    } else {
        reinterpret_cast<c32*>(base)[off] = c32{to_half_sat(v.re * h_scale), to_half_sat(v.im * h_scale)};
    }
}

// MODE ASM_HG: out0 = H, out1 = G (either may be null), element type OutT.
// MODE ASM_LHS: out0 = LHS c128 (or null), out1 = LHS as OutT (c64, or c32
// scaled by p.h_scale; or null).
template <typename T, bool BM, int MODE, typename OutT>
__global__ __launch_bounds__(kRows) void assemble_kernel(const SysParams* __restrict__ params,
                                                         const double* __restrict__ soa_base,
                                                         int64_t n, void* out0, void* out1,
                                                         int64_t ld, int64_t stride,
                                                         double* __restrict__ rowsum_part,
                                                         int64_t nchunks) {
    const int64_t b = blockIdx.z;
    const SysParams& p = params[b];
    const Soa s = soa_of(soa_base, n);
    const int64_t i = int64_t(blockIdx.x) * kRows + threadIdx.x;
    const int64_t j0 = int64_t(blockIdx.y) * kColChunk;
    const int64_t j1 = min(j0 + kColChunk, n);
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    const double yx = s.x[ic], yy = s.y[ic], yz = s.z[ic];
    const T nyx = T(s.nx[ic]), nyy = T(s.ny[ic]), nyz = T(s.nz[ic]);
    const T k = T(p.k);
    const Cx<T> beta{T(p.beta.re), T(p.beta.im)};
    const T dh_sign = T(p.dh_sign);
    const int64_t sys_off = b * stride;
    double rs_re = 0.0, rs_im = 0.0;
    for (int64_t j = j0; j < j1; ++j) {
        const double wa = s.wt[j] * s.area[j];
        c128 G, H;
        if (j == ic) {
            diag_entry(p, s.area[j], s.wt[j], G, H);
        } else {
            Cx<T> g, h;
            if constexpr (std::is_same<T, double>::value)
                greens_fast<BM>(k, beta, s.x[j] - yx, s.y[j] - yy, s.z[j] - yz, s.nx[j], s.ny[j],
                                s.nz[j], nyx, nyy, nyz, g, h, dh_sign);
            else
                greens<T, BM>(k, beta, T(s.x[j] - yx), T(s.y[j] - yy), T(s.z[j] - yz),
                              T(s.nx[j]), T(s.ny[j]), T(s.nz[j]), nyx, nyy, nyz, g, h, dh_sign);
            G = c128{double(g.re) * wa, double(g.im) * wa};
            H = c128{double(h.re) * wa, double(h.im) * wa};
        }
        if (!active) continue;
        const int64_t off = sys_off + i + j * ld;
        if constexpr (MODE == ASM_HG) {
            if (out0) store<OutT>(out0, off, H);
            if (out1) store<OutT>(out1, off, G);
        } else {
            const c128 L = p.lhs_a * H + p.lhs_b * G;
            if (out0) store<c128>(out0, off, L);
            if (out1) store<OutT>(out1, off, L, p.h_scale);
            if (p.rowsum_of == 1) {
                rs_re += G.re;
                rs_im += G.im;
            } else if (p.rowsum_of == 2) {
                rs_re += H.re;
                rs_im += H.im;
            }
        }
    }
    if (MODE == ASM_LHS && rowsum_part != nullptr && active && p.rowsum_of != 0) {
        double* dst = rowsum_part + 2 * ((b * nchunks + blockIdx.y) * n + i);
        dst[0] = rs_re;
        dst[1] = rs_im;
    }
}

// With beta = i/k (Burton-Miller) the pair of influence_matrix.rs:12-37
// times the weight w = wa / (4 pi r) is
//   H = w e^{ikr} (a0, k a1 + a2 / k),  G = w e^{ikr} (g0, g1 / k)
// with E = n_y.n_x - 3 rr, rr = (e.n_x)(e.n_y) dh_sign, a0 = (e.n_y + E) / r,
// a1 = rr - e.n_y, a2 = E / r^2, g0 = 1 - e.n_x, g1 = -(e.n_x) / r
// (classical: H = w e^{ikr} (e.n_y)(1/r, -k), G = w e^{ikr}); only the
// phasor and k, 1/k change with the frequency.
template <bool BM>
__device__ __forceinline__ void pair_coeffs(const PairGeom<double>& q, double dh_sign, double& a0,
                                            double& a1, double& a2, double& g0, double& g1) {
    if constexpr (BM) {
        const double rr = q.rdotx * q.rdoty * dh_sign;
        const double E = q.nn - 3.0 * rr;
        a0 = q.inv_r * (q.rdoty + E);
        a1 = rr - q.rdoty;
        a2 = q.inv_r * q.inv_r * E;
        g0 = 1.0 - q.rdotx;
        g1 = -q.inv_r * q.rdotx;
    } else {
        a0 = q.inv_r * q.rdoty;
        a1 = -q.rdoty;
        a2 = 0.0;
        g0 = 1.0;
        g1 = 0.0;
    }
}

// c with lhs_a H + lhs_b G = w e^{ikr} c off the diagonal, from the
// coefficients of pair_coeffs and the per-system products lhs_a k,
// lhs_a / k, lhs_b / k of SysParams: one complex product per entry instead of
// forming H and G (used where no row sum of G or H is needed)
template <bool BM>
__device__ __forceinline__ c128 lhs_coef(const SysParams& p, double a0, double a1, double a2,
                                         double g0, double g1) {
    if constexpr (BM)
        return c128{p.lhs_a.re * a0 - p.lhs_ak.im * a1 - p.lhs_aik.im * a2 + p.lhs_b.re * g0 -
                        p.lhs_bik.im * g1,
                    p.lhs_a.im * a0 + p.lhs_ak.re * a1 + p.lhs_aik.re * a2 + p.lhs_b.im * g0 +
                        p.lhs_bik.re * g1};
    else  // a2 = 0, g0 = 1, g1 = 0
        return c128{p.lhs_a.re * a0 - p.lhs_ak.im * a1 + p.lhs_b.re,
                    p.lhs_a.im * a0 + p.lhs_ak.re * a1 + p.lhs_b.im};
}

// lhs_coef for lhs_a = 1, lhs_b = 0 (H alone: the normal-velocity BC)
template <bool BM>
__device__ __forceinline__ c128 lhs_coef_h(const SysParams& p, double a0, double a1, double a2) {
    if constexpr (BM) return c128{a0, fma(p.lhs_aik.re, a2, p.lhs_ak.re * a1)};
    else return c128{a0, p.lhs_ak.re * a1};
}

// one LHS entry L to the outputs: out0 c128; out1 OutT (c64, or c32 scaled by
// h_scale); with OutT = c32 and out2, out2 the complex-FP32 remainder
// L - fp16(L h_scale) / h_scale
template <typename OutT>
__device__ __forceinline__ void store_lhs(const SysParams& p, int64_t off, c128 L, void* out0,
                                          void* out1, void* out2) {
    if (out0) store<c128>(out0, off, L);
    if constexpr (sizeof(OutT) == 4) {
        if (out2) {
            const uint16_t hr = to_half_sat(L.re * p.h_scale);
            const uint16_t hi = to_half_sat(L.im * p.h_scale);
            reinterpret_cast<c32*>(out1)[off] = c32{hr, hi};
            // h_inv = 1 / h_scale, a power of two: the products are exact
            const double br = double(__half2float(__ushort_as_half(hr))) * p.h_inv;
            const double bi = double(__half2float(__ushort_as_half(hi))) * p.h_inv;
            reinterpret_cast<c64*>(out2)[off] =
                c64{static_cast<float>(L.re - br), static_cast<float>(L.im - bi)};
        } else if (out1) {
            store<OutT>(out1, off, L, p.h_scale);
        }
    } else {
        if (out1) store<OutT>(out1, off, L, p.h_scale);
    }
}

// ASM_LHS assembly of a frequency sweep over one mesh, FP64 arithmetic, KF
// systems per thread and cc columns per workgroup (defaults 16 and 64, or
// the tuned values of ccab_autotune_sweep): the pair geometry (r, 1/r, the
// normal projections) is computed once for the KF systems (measured 6.6 ms
// at 8, 5.7 at 16, 5.7 at 32 for 64 frequencies at n = 5120).  With equally spaced wavenumbers (step
// dk != 0) the phasor e^{i k r} comes from sincos for the first system of the
// group and from e^{i k_f r} = e^{i k_(f-1) r} e^{i dk r} for the others (two
// sincos per kF entries instead of kF); otherwise from sincos per system.
// Same layout, outputs and row-sum partials as assemble_kernel; with OutT =
// c32 and out2 set, out2 also receives the complex-FP32 remainder
// L - fp16(L h_scale) / h_scale (the FP16 copy plus the remainder stand for
// the FP64 matrix to about 2^-35 relative per entry).  With ax_part set,
// every entry L_ij is also multiplied by x0_j as it is formed: the partial
// sums of A x0 per column chunk (the first residual of a warm start, b - A
// x0, in FP64 from the FP64 entries, without reading the stored matrix again).
constexpr int kF = 16;

template <bool BM, typename OutT, int KF>
__global__ __launch_bounds__(kRows) void assemble_sweep_kernel(
    const SysParams* __restrict__ params, const double* __restrict__ soa_base, int64_t n,
    int64_t batch, void* out0, void* out1, void* out2, int64_t ld, int64_t stride,
    double* __restrict__ rowsum_part, int64_t nchunks, double dk, int cc,
    const c128* __restrict__ x0, double* __restrict__ ax_part) {
    const int64_t b0 = int64_t(blockIdx.z) * KF;
    const int nf = (int)min<int64_t>(KF, batch - b0);
    const Soa s = soa_of(soa_base, n);
    const int64_t i = int64_t(blockIdx.x) * kRows + threadIdx.x;
    const int64_t j0 = int64_t(blockIdx.y) * cc;
    const int64_t j1 = min(j0 + cc, n);
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    const double yx = s.x[ic], yy = s.y[ic], yz = s.z[ic];
    const double nyx = s.nx[ic], nyy = s.ny[ic], nyz = s.nz[ic];
    const double k0 = params[b0].k;
    const double dh_sign = params[b0].dh_sign;  // one value per batch (checked by the launcher)
    // per-lane row-sum accumulators of the kF systems (private LDS slots;
    // dynamic LDS of kF * kRows * 16 bytes, launched only with row sums)
    extern __shared__ double2 rs_lds[];
    auto rs = [&](int f) -> double2& { return rs_lds[f * kRows + threadIdx.x]; };
    if (rowsum_part)
        for (int f = 0; f < KF; ++f) rs(f) = double2{0.0, 0.0};
    // per-lane accumulators of (A x0)_i of the kF systems, in the dynamic
    // LDS after the row sums' slots, then the guess over the workgroup's
    // columns (kF x cc, staged once: read per entry, it would be a scalar
    // load the loop waits for)
    double2* ax_lds = rs_lds + (rowsum_part ? KF * kRows : 0);
    c128* x0s = reinterpret_cast<c128*>(ax_lds + KF * kRows);
    auto axs = [&](int f) -> double2& { return ax_lds[f * kRows + threadIdx.x]; };
    if (ax_part) {
        for (int f = 0; f < KF; ++f) axs(f) = double2{0.0, 0.0};
        const int ncol = (int)(j1 - j0);
        for (int t = threadIdx.x; t < nf * cc; t += kRows) {
            const int f = t / cc, jj = t - f * cc;
            x0s[t] = jj < ncol ? x0[(b0 + f) * n + j0 + jj] : c128{0.0, 0.0};
        }
        __syncthreads();
    }
    auto add_ax = [&](int f, int64_t j, const c128& L) {
        const c128 xv = x0s[f * cc + (j - j0)];
        double2& a = axs(f);
        a.x += L.re * xv.re - L.im * xv.im;
        a.y += L.re * xv.im + L.im * xv.re;
    };
    // column loop variants: 0 general, 1 every system of the group H alone
    // (lhs_a = 1, lhs_b = 0, no row sums): only H's coefficient is formed
    auto columns = [&](auto variant) {
    constexpr int VAR = decltype(variant)::value;
    for (int64_t j = j0; j < j1; ++j) {
        const double wa = s.wt[j] * s.area[j];
        const bool diag = j == ic;
        // the diagonal lane evaluates a unit-distance pair and discards it
        const double rx = diag ? 1.0 : s.x[j] - yx;
        const PairGeom<double> q = pair_geom_fast<BM>(rx, s.y[j] - yy, s.z[j] - yz, s.nx[j],
                                                      s.ny[j], s.nz[j], nyx, nyy, nyz);
        const double w = wa * q.inv_r * (1.0 / (4.0 * kPi));
        double a0, a1, a2, g0, g1;
        pair_coeffs<BM>(q, dh_sign, a0, a1, a2, g0, g1);
        double sn, cs, sd = 0.0, cd = 1.0;
        sincos_fast(k0 * q.rd, &sn, &cs);
        if (dk != 0.0) sincos_fast(dk * q.rd, &sd, &cd);
        // the phasor carries the weight
        cs *= w;
        sn *= w;
#pragma unroll 1
        for (int f = 0; f < nf; ++f) {
            const SysParams& p = params[b0 + f];
            if (f > 0) {
                if (dk != 0.0) {
                    const double c2 = cs * cd - sn * sd;
                    sn = sn * cd + cs * sd;
                    cs = c2;
                } else {
                    sincos_fast(p.k * q.rd, &sn, &cs);
                    cs *= w;
                    sn *= w;
                }
            }
            const c128 ph{cs, sn};
            if constexpr (VAR == 1) {
                c128 L;
                if (diag) {
                    c128 G, H;
                    diag_entry(p, s.area[j], s.wt[j], G, H);
                    L = p.lhs_a * H + p.lhs_b * G;
                } else {
                    L = ph * lhs_coef_h<BM>(p, a0, a1, a2);
                }
                if (active) store_lhs<OutT>(p, (b0 + f) * stride + i + j * ld, L, out0, out1, out2);
                if (ax_part) add_ax(f, j, L);
                continue;
            }
            if (rowsum_part && p.rowsum_of != 0) {
                c128 G, H;
                if (diag) {
                    diag_entry(p, s.area[j], s.wt[j], G, H);
                } else {
                    const double ik = p.beta.im;  // 1/k (Burton-Miller; unused otherwise)
                    H = ph * c128{a0, p.k * a1 + a2 * ik};
                    G = ph * c128{g0, g1 * ik};
                }
                const c128 L = p.lhs_a * H + p.lhs_b * G;
                if (active) store_lhs<OutT>(p, (b0 + f) * stride + i + j * ld, L, out0, out1, out2);
                if (ax_part) add_ax(f, j, L);
                const c128 a = p.rowsum_of == 1 ? G : H;
                double2& r = rs(f);
                r.x += a.re;
                r.y += a.im;
            } else {
                c128 L;
                if (diag) {
                    c128 G, H;
                    diag_entry(p, s.area[j], s.wt[j], G, H);
                    L = p.lhs_a * H + p.lhs_b * G;
                } else {
                    L = ph * lhs_coef<BM>(p, a0, a1, a2, g0, g1);
                }
                if (active) store_lhs<OutT>(p, (b0 + f) * stride + i + j * ld, L, out0, out1, out2);
                if (ax_part) add_ax(f, j, L);
            }
        }
    }
    };
    bool h_group = true;
    for (int f = 0; f < nf; ++f) {
        const SysParams& p = params[b0 + f];
        h_group = h_group && p.lhs_b.re == 0.0 && p.lhs_b.im == 0.0 && p.lhs_a.re == 1.0 &&
                  p.lhs_a.im == 0.0 && !(rowsum_part && p.rowsum_of != 0);
    }
    if (h_group)
        columns(std::integral_constant<int, 1>{});
    else
        columns(std::integral_constant<int, 0>{});
    if (ax_part && active)
        for (int f = 0; f < nf; ++f) {
            double* dst = ax_part + 2 * (((b0 + f) * nchunks + blockIdx.y) * n + i);
            dst[0] = axs(f).x;
            dst[1] = axs(f).y;
        }
    if (rowsum_part == nullptr || !active) return;
    for (int f = 0; f < nf; ++f) {
        if (params[b0 + f].rowsum_of == 0) continue;
        double* dst = rowsum_part + 2 * (((b0 + f) * nchunks + blockIdx.y) * n + i);
        dst[0] = rs(f).x;
        dst[1] = rs(f).y;
    }
}

// rhs = rhs_inc_sign * rhs_inc + rhs_scale * rowsum   (deterministic chunk order)
__global__ void rhs_finish_kernel(const SysParams* __restrict__ params, int64_t n,
                                  const double* __restrict__ part, int64_t nchunks,
                                  const ccab_c128* __restrict__ rhs_inc, ccab_c128* rhs) {
    const int64_t b = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SysParams& p = params[b];
    c128 acc{0.0, 0.0};
    if (p.rowsum_of != 0 && part != nullptr) {
        for (int64_t c = 0; c < nchunks; ++c) {
            const double* src = part + 2 * ((b * nchunks + c) * n + i);
            acc.re += src[0];
            acc.im += src[1];
        }
        acc = p.rhs_scale * acc;
    }
    const ccab_c128 ri = rhs_inc[b * n + i];
    rhs[b * n + i] = ccab_c128{acc.re + p.rhs_inc_sign * ri.re, acc.im + p.rhs_inc_sign * ri.im};
}

// ax[b n + i] = sum of the chunk partials of (A x0)_i (fixed chunk order)
__global__ void ax_finish_kernel(int64_t n, const double* __restrict__ part, int64_t nchunks,
                                 c128* ax) {
    const int64_t b = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    c128 acc{0.0, 0.0};
    for (int64_t c = 0; c < nchunks; ++c) {
        const double* src = part + 2 * ((b * nchunks + c) * n + i);
        acc.re += src[0];
        acc.im += src[1];
    }
    ax[b * n + i] = acc;
}

__global__ void cpts_to_soa_kernel(const double* __restrict__ aos, int64_t n, double* soa) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int f = 0; f < 8; ++f) soa[f * n + i] = aos[8 * i + f];
}

// incident_wave.rs:7-63
__device__ __forceinline__ c128 incident_value(const SysParams& p, int kind, double d0, double d1,
                                               double d2, c128 amp, const Soa& s, int64_t i) {
    const double cx_ = s.x[i], cy = s.y[i], cz = s.z[i];
    c128 out;
    if (kind == CCAB_WAVE_PLANE) {
        double sn, cs;
        sincos(p.k * (d0 * cx_ + d1 * cy + d2 * cz), &sn, &cs);
        const c128 phi = amp * c128{cs, sn};
        out = phi;
        if (p.bm) {
            const double dn = d0 * s.nx[i] + d1 * s.ny[i] + d2 * s.nz[i];
            const c128 vn = (phi * c128{0.0, p.k}) * dn;
            out += p.beta * vn;
        }
    } else {
        const double rx = cx_ - d0, ry = cy - d1, rz = cz - d2;
        const double r = sqrt(rx * rx + ry * ry + rz * rz);
        double sn, cs;
        sincos(p.k * r, &sn, &cs);
        const double inv = 1.0 / (4.0 * kPi * r);
        const c128 phi = (amp * c128{cs, sn}) * inv;
        out = phi;
        if (p.bm) {
            const double en = (rx * s.nx[i] + ry * s.ny[i] + rz * s.nz[i]) / r;
            const c128 vn = (phi * c128{-1.0 / r, p.k}) * en;
            out += p.beta * vn;
        }
    }
    return out;
}

__global__ void incident_kernel(SysParams p, int kind, double d0, double d1, double d2,
                                c128 amp, const double* __restrict__ soa_base, int64_t n,
                                ccab_c128* rhs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const c128 v = incident_value(p, kind, d0, d1, d2, amp, soa_of(soa_base, n), i);
    rhs[i] = ccab_c128{v.re, v.im};
}

// one wave, every frequency of a sweep (grid.y): system b's amplitude amps[b]
__global__ void incident_sweep_kernel(const SysParams* __restrict__ params,
                                      const c128* __restrict__ amps, int kind, double d0,
                                      double d1, double d2, const double* __restrict__ soa_base,
                                      int64_t n, ccab_c128* rhs) {
    const int64_t b = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const c128 v = incident_value(params[b], kind, d0, d1, d2, amps[b], soa_of(soa_base, n), i);
    rhs[b * n + i] = ccab_c128{v.re, v.im};
}

// influence_matrix.rs:170-231.  Diagonal shift at the first position t whose
// (i, j) pair is diagonal.
template <bool BM>
__global__ void rowcol_kernel(SysParams p, const double* __restrict__ soa_base, int64_t n,
                              const int64_t* __restrict__ di, int64_t ni,
                              const int64_t* __restrict__ dj, int64_t nj, int side,
                              ccab_c128* out) {
    const int64_t cnt = ni == 1 ? nj : ni;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const Soa s = soa_of(soa_base, n);
    const int64_t i = ni == 1 ? di[0] : di[t];
    const int64_t j = ni == 1 ? dj[t] : dj[0];
    c128 G, H;
    if (i == j) {
        // values without the shift, then shift only at the first diagonal hit
        diag_entry(p, s.area[j], s.wt[j], G, H);
        G -= p.gdiag;
        H -= p.hdiag;
        bool first = true;
        for (int64_t u = 0; u < t; ++u) {
            const int64_t iu = ni == 1 ? di[0] : di[u];
            const int64_t ju = ni == 1 ? dj[u] : dj[0];
            if (iu == ju) {
                first = false;
                break;
            }
        }
        if (first) {
            G += p.gdiag;
            H += p.hdiag;
        }
    } else {
        c128 g, h;
        greens_fast<BM>(p.k, p.beta, s.x[j] - s.x[i], s.y[j] - s.y[i], s.z[j] - s.z[i],
                        s.nx[j], s.ny[j], s.nz[j], s.nx[i], s.ny[i], s.nz[i], g, h, p.dh_sign);
        const double wa = s.wt[j] * s.area[j];
        G = g * wa;
        H = h * wa;
    }
    const c128 a = side == CCAB_LHS ? p.lhs_a : p.rhs_a;
    const c128 bb = side == CCAB_LHS ? p.lhs_b : p.rhs_b;
    const c128 v = bb * G + a * H;
    out[t] = ccab_c128{v.re, v.im};
}

template <typename T, bool BM, int MODE, typename OutT>
void launch_one(dim3 grid, hipStream_t st, const SysParams* d_params, const double* soa,
                int64_t n, void* o0, void* o1, int64_t ld, int64_t stride, double* part,
                int64_t nchunks) {
    hipLaunchKernelGGL((assemble_kernel<T, BM, MODE, OutT>), grid, dim3(kRows), 0, st, d_params,
                       soa, n, o0, o1, ld, stride, part, nchunks);
}

// --- ragged batches of bodies --------------------------------------------------
__device__ __forceinline__ int body_of_wg(const BodyDesc* __restrict__ desc, int nbody, int64_t wg) {
    int lo = 0, hi = nbody;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (desc[mid].wg0 <= wg) lo = mid;
        else hi = mid;
    }
    return lo;
}

// grid (rows / 256, bodies): SoA copy of body c's points
__global__ void bodies_soa_kernel(const BodyDesc* __restrict__ desc, const double* __restrict__ aos,
                                  double* soa) {
    const BodyDesc d = desc[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const double* src = aos + 8 * (d.pts + i);
    double* dst = soa + 8 * d.pts;
#pragma unroll
    for (int f = 0; f < 8; ++f) dst[f * d.n + i] = src[f];
}

// grid (rows / 256, bodies): incident right-hand side and optional guess of body c
__global__ void bodies_incident_kernel(const BodyDesc* __restrict__ desc,
                                       const SysParams* __restrict__ params,
                                       const BodyWave* __restrict__ waves,
                                       const double* __restrict__ soa, int64_t nmax, ccab_c128* X,
                                       const ccab_c128* __restrict__ x0, ccab_c128* X0) {
    const int64_t c = blockIdx.y;
    const BodyDesc d = desc[c];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const BodyWave w = waves[c];
    const c128 v = incident_value(params[c], w.kind, w.d[0], w.d[1], w.d[2], w.amp,
                                  soa_of(soa + 8 * d.pts, d.n), i);
    X[c * nmax + i] = ccab_c128{v.re, v.im};
    if (X0) X0[c * nmax + i] = x0[d.pts + i];
}

// Fused LHS assembly of every body (one system each, FP64 arithmetic): the
// 1-D grid is every body's (row block of 256, column chunk of 64) tiles back
// to back; the pair as in the sweep kernel, with the phasor from sincos_fast.
// A body whose right-hand side needs no row sum (rowsum_of 0) takes the
// column loop that forms lhs_a H + lhs_b G as one complex product per entry
// (H alone for lhs_a = 1, lhs_b = 0).  With ax_part set, the partial sums of
// A x0 per column chunk as in the sweep kernel (x0 of body c at c nmax, laid
// out as the row-sum partials).
template <bool BM, typename OutT>
__global__ __launch_bounds__(kRows) void assemble_bodies_kernel(
    const SysParams* __restrict__ params, const BodyDesc* __restrict__ desc, int nbody,
    const double* __restrict__ soa_all, void* out0, void* out1, void* out2,
    double* __restrict__ rowsum_part, const c128* __restrict__ x0, int64_t nmax,
    double* __restrict__ ax_part) {
    const int c = body_of_wg(desc, nbody, blockIdx.x);
    const BodyDesc d = desc[c];
    const SysParams p = params[c];  // a private copy: not reloaded after each store
    const int64_t n = d.n;
    const int64_t rowblocks = (n + kRows - 1) / kRows;
    const int64_t local = int64_t(blockIdx.x) - d.wg0;
    const int64_t rb = local % rowblocks, cb = local / rowblocks;
    const Soa s = soa_of(soa_all + 8 * d.pts, n);
    const int64_t i = rb * kRows + threadIdx.x;
    const int64_t j0 = cb * kColChunk;
    const int64_t j1 = min(j0 + kColChunk, n);
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    const double yx = s.x[ic], yy = s.y[ic], yz = s.z[ic];
    const double nyx = s.nx[ic], nyy = s.ny[ic], nyz = s.nz[ic];
    const double ik = p.beta.im;  // 1/k (Burton-Miller; unused otherwise)
    double rs_re = 0.0, rs_im = 0.0;
    c128 ax{0.0, 0.0};
    // the guess over the workgroup's columns, staged once in LDS
    __shared__ c128 x0s[kColChunk];
    if (x0) {
        if (threadIdx.x < kColChunk)
            x0s[threadIdx.x] = j0 + threadIdx.x < j1 ? x0[int64_t(c) * nmax + j0 + threadIdx.x]
                                                     : c128{0.0, 0.0};
        __syncthreads();
    }
    // column loop variants: 0 fused lhs_a H + lhs_b G, 1 H alone, 2 H and G
    // separately with row sums
    auto columns = [&](auto variant) {
        constexpr int VAR = decltype(variant)::value;
        constexpr bool RS = VAR == 2;
        for (int64_t j = j0; j < j1; ++j) {
            const double wa = s.wt[j] * s.area[j];
            const bool diag = j == ic;
            const double rx = diag ? 1.0 : s.x[j] - yx;
            const PairGeom<double> q = pair_geom_fast<BM>(rx, s.y[j] - yy, s.z[j] - yz, s.nx[j],
                                                          s.ny[j], s.nz[j], nyx, nyy, nyz);
            c128 L, G, H;
            if (diag) {
                diag_entry(p, s.area[j], s.wt[j], G, H);
                L = p.lhs_a * H + p.lhs_b * G;
            } else {
                double a0, a1, a2, g0, g1;
                pair_coeffs<BM>(q, p.dh_sign, a0, a1, a2, g0, g1);
                const double w = wa * q.inv_r * (1.0 / (4.0 * kPi));
                double sn, cs;
                sincos_fast(p.k * q.rd, &sn, &cs);
                const c128 ph{cs * w, sn * w};
                if constexpr (RS) {
                    H = ph * c128{a0, p.k * a1 + a2 * ik};
                    G = ph * c128{g0, g1 * ik};
                    L = p.lhs_a * H + p.lhs_b * G;
                } else if constexpr (VAR == 1) {
                    L = ph * lhs_coef_h<BM>(p, a0, a1, a2);
                } else {
                    L = ph * lhs_coef<BM>(p, a0, a1, a2, g0, g1);
                }
            }
            if (active) store_lhs<OutT>(p, d.aoff + i + j * n, L, out0, out1, out2);
            if (x0) ax += L * x0s[j - j0];
            if constexpr (RS) {
                const c128 a = p.rowsum_of == 1 ? G : H;
                rs_re += a.re;
                rs_im += a.im;
            }
        }
    };
    const bool rs_on = rowsum_part != nullptr && p.rowsum_of != 0;
    const bool h_only = p.lhs_b.re == 0.0 && p.lhs_b.im == 0.0 && p.lhs_a.re == 1.0 &&
                        p.lhs_a.im == 0.0;
    if (rs_on)
        columns(std::integral_constant<int, 2>{});
    else if (h_only)
        columns(std::integral_constant<int, 1>{});
    else
        columns(std::integral_constant<int, 0>{});
    if (rs_on && active) {
        double* dst = rowsum_part + d.part_off + 2 * (cb * n + i);
        dst[0] = rs_re;
        dst[1] = rs_im;
    }
    if (ax_part && active) {
        double* dst = ax_part + d.part_off + 2 * (cb * n + i);
        dst[0] = ax.re;
        dst[1] = ax.im;
    }
}

// grid (rows / 256, bodies): AX[c nmax + i] = the sum of body c's A x0
// partials (fixed chunk order), zero past the body's size
__global__ void ax_finish_bodies_kernel(const BodyDesc* __restrict__ desc,
                                        const double* __restrict__ part, int64_t nmax, c128* AX) {
    const int64_t c = blockIdx.y;
    const BodyDesc d = desc[c];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nmax) return;
    c128 acc{0.0, 0.0};
    if (i < d.n) {
        const int64_t nch = (d.n + kColChunk - 1) / kColChunk;
        for (int64_t k = 0; k < nch; ++k) {
            const double* src = part + d.part_off + 2 * (k * d.n + i);
            acc.re += src[0];
            acc.im += src[1];
        }
    }
    AX[c * nmax + i] = acc;
}

// grid (rows / 256, bodies): X = rhs_inc_sign X + rhs_scale sum of the row
// partials (fixed chunk order), as rhs_finish_kernel per body
__global__ void rhs_finish_bodies_kernel(const BodyDesc* __restrict__ desc,
                                         const SysParams* __restrict__ params,
                                         const double* __restrict__ part, int64_t nmax,
                                         ccab_c128* X) {
    const int64_t c = blockIdx.y;
    const BodyDesc d = desc[c];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const SysParams& p = params[c];
    c128 acc{0.0, 0.0};
    if (p.rowsum_of != 0 && part != nullptr) {
        const int64_t nch = (d.n + kColChunk - 1) / kColChunk;
        for (int64_t k = 0; k < nch; ++k) {
            const double* src = part + d.part_off + 2 * (k * d.n + i);
            acc.re += src[0];
            acc.im += src[1];
        }
        acc = p.rhs_scale * acc;
    }
    const ccab_c128 ri = X[c * nmax + i];
    X[c * nmax + i] = ccab_c128{acc.re + p.rhs_inc_sign * ri.re, acc.im + p.rhs_inc_sign * ri.im};
}

// Largest |lhs_a H_ii + lhs_b G_ii| over the diagonal of each system, one
// workgroup per system (the FP16 operator's scale follows from it).  desc
// null: every system has the points 0..n-1 of cpts (a sweep); otherwise
// system s has desc[s].n points from point desc[s].pts.  cpts: 8 doubles
// per point, area and weight at 6 and 7.
__global__ __launch_bounds__(256) void diag_max_kernel(const SysParams* __restrict__ params,
                                                       const double* __restrict__ cpts,
                                                       const BodyDesc* __restrict__ desc, int64_t n,
                                                       double* __restrict__ out) {
    __shared__ double red[256];
    const int64_t s = blockIdx.x;
    const SysParams& p = params[s];
    const int64_t ns = desc ? desc[s].n : n;
    const double* c = cpts + 8 * (desc ? desc[s].pts : 0);
    double mx = 0.0;
    for (int64_t i = threadIdx.x; i < ns; i += 256) {
        c128 G, H;
        diag_entry(p, c[8 * i + 6], c[8 * i + 7], G, H);
        const c128 L = p.lhs_a * H + p.lhs_b * G;
        mx = fmax(mx, sqrt(norm_sqr(L)));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

}  // namespace

int launch_diag_max(ccab_ctx* ctx, int64_t nsys, const SysParams* d_params, const double* d_cpts,
                    const BodyDesc* d_desc, int64_t n, double* d_out) {
    if (nsys <= 0) return CCAB_OK;
    hipLaunchKernelGGL(diag_max_kernel, dim3((unsigned)nsys), dim3(256), 0, ctx->stream, d_params,
                       d_cpts, d_desc, n, d_out);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_bodies_setup(ccab_ctx* ctx, int64_t nbody, int64_t nmax, const BodyDesc* d_desc,
                        const double* d_cpts, double* soa, const SysParams* d_params,
                        const BodyWave* d_waves, ccab_c128* X, const ccab_c128* d_x0,
                        ccab_c128* X0) {
    CCAB_CHECK(nbody >= 1 && nbody <= 65535, "bodies per launch out of range");
    const dim3 g((unsigned)((nmax + 255) / 256), (unsigned)nbody);
    hipLaunchKernelGGL(bodies_soa_kernel, g, dim3(256), 0, ctx->stream, d_desc, d_cpts, soa);
    hipLaunchKernelGGL(bodies_incident_kernel, g, dim3(256), 0, ctx->stream, d_desc, d_params,
                       d_waves, soa, nmax, X, d_x0, X0);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_assemble_bodies(ccab_ctx* ctx, int64_t nbody, int64_t nmax, const BodyDesc* d_desc,
                           const BodyDesc* h_desc, const SysParams* d_params,
                           const SysParams* h_params, const double* soa, int prec, void* out0,
                           void* out1, void* out2, double* part, ccab_c128* X, const c128* X0,
                           double* ax_part, c128* AX) {
    CCAB_CHECK((X0 == nullptr) == (ax_part == nullptr) && (AX == nullptr) == (ax_part == nullptr),
               "A x0 needs the guess, its partials and the output");
    CCAB_CHECK(nbody >= 1 && nbody <= 65535, "bodies per launch out of range");
    CCAB_CHECK(prec == PREC_F64 || prec == PREC_F64_H, "bodies assemble in FP64");
    CCAB_CHECK(out2 == nullptr || (prec == PREC_F64_H && out1), "the FP32 remainder goes with the FP16 copy");
    const bool bm = h_params[0].bm != 0;
    double entries = 0.0;
    for (int64_t c = 0; c < nbody; ++c) {
        CCAB_CHECK((h_params[c].bm != 0) == bm, "all bodies of a batch share one method");
        CCAB_CHECK(!bm || (h_params[c].beta.re == 0.0 &&
                           std::fabs(h_params[c].beta.im * h_params[c].k - 1.0) < 1e-14),
                   "Burton-Miller coupling beta = i/k expected");
        entries += double(h_desc[c].n) * double(h_desc[c].n);
    }
    const int64_t wgs = h_desc[nbody].wg0;
    CCAB_CHECK(wgs >= 1 && wgs < (int64_t(1) << 31), "assembly grid out of range");
    const double bytes = (out0 ? 16.0 : 0.0) + (out1 ? (prec == PREC_F64_H ? 4.0 : 8.0) : 0.0) +
                         (out2 ? 8.0 : 0.0);
    hipStream_t st = ctx->stream;
    {
        ProfScope ps(ctx, PROF_ASSEMBLY, entries * bytes);
        const dim3 grid((unsigned)wgs);
        const int nb = (int)nbody;
        if (prec == PREC_F64_H) {
            if (bm) hipLaunchKernelGGL((assemble_bodies_kernel<true, c32>), grid, dim3(kRows), 0, st, d_params, d_desc, nb, soa, out0, out1, out2, part, X0, nmax, ax_part);
            else hipLaunchKernelGGL((assemble_bodies_kernel<false, c32>), grid, dim3(kRows), 0, st, d_params, d_desc, nb, soa, out0, out1, out2, part, X0, nmax, ax_part);
        } else {
            if (bm) hipLaunchKernelGGL((assemble_bodies_kernel<true, c64>), grid, dim3(kRows), 0, st, d_params, d_desc, nb, soa, out0, out1, out2, part, X0, nmax, ax_part);
            else hipLaunchKernelGGL((assemble_bodies_kernel<false, c64>), grid, dim3(kRows), 0, st, d_params, d_desc, nb, soa, out0, out1, out2, part, X0, nmax, ax_part);
        }
    }
    hipLaunchKernelGGL(rhs_finish_bodies_kernel, dim3((unsigned)((nmax + 255) / 256), (unsigned)nbody),
                       dim3(256), 0, st, d_desc, d_params, part, nmax, X);
    if (ax_part)
        hipLaunchKernelGGL(ax_finish_bodies_kernel, dim3((unsigned)((nmax + 255) / 256), (unsigned)nbody),
                           dim3(256), 0, st, d_desc, ax_part, nmax, AX);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_cpts_to_soa(ccab_ctx* ctx, const double* d_cpts, int64_t n, double* d_soa) {
    hipLaunchKernelGGL(cpts_to_soa_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       d_cpts, n, d_soa);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_assemble(ccab_ctx* ctx, const SysParams* d_params, const SysParams* h_params,
                    int64_t batch, const double* d_soa, int64_t n, int prec_compute, int mode,
                    void* out0, void* out1, int64_t ld, int64_t stride, double* d_rowsum_part,
                    int64_t* nchunks_out, void* out2, const c128* d_x0, double* d_ax_part) {
    const int64_t nchunks = (n + kColChunk - 1) / kColChunk;
    CCAB_CHECK((d_x0 == nullptr) == (d_ax_part == nullptr), "A x0 needs the guess and its partials");
    if (nchunks_out) *nchunks_out = nchunks;
    CCAB_CHECK(batch >= 1 && batch <= 65535, "batch out of range");
    const bool bm = h_params[0].bm != 0;
    for (int64_t b = 1; b < batch; ++b)
        CCAB_CHECK((h_params[b].bm != 0) == bm, "all systems of a batch share one method");
    dim3 grid((unsigned)((n + kRows - 1) / kRows), (unsigned)nchunks, (unsigned)batch);
    hipStream_t st = ctx->stream;
    double bytes_per_entry = 0.0;
    if (mode == ASM_HG) {
        const double e = prec_compute == PREC_F64 ? 16.0 : 8.0;
        bytes_per_entry = (out0 ? e : 0.0) + (out1 ? e : 0.0);
    } else {
        bytes_per_entry = (out0 ? 16.0 : 0.0) + (out1 ? (prec_compute == PREC_F64_H ? 4.0 : 8.0) : 0.0) +
                          (out2 ? 8.0 : 0.0);
    }
    CCAB_CHECK(out2 == nullptr || (mode == ASM_LHS && prec_compute == PREC_F64_H && out1),
               "the FP32 remainder goes with the FP16 copy of the LHS");
    ProfScope ps(ctx, PROF_ASSEMBLY, double(n) * double(n) * double(batch) * bytes_per_entry);
    static const bool sweep_kernel = [] {
        const char* e = std::getenv("CCAB_ASM_SWEEP");
        return !(e && e[0] == '0');
    }();
    CCAB_CHECK(!d_ax_part || (mode == ASM_LHS && prec_compute != PREC_F32),
               "A x0 is formed by the FP64 LHS assembly");
    if (mode == ASM_LHS && prec_compute != PREC_F32 &&
        ((batch >= 2 && sweep_kernel) || out2 || d_ax_part)) {
        // equally spaced wavenumbers (to 4 ulp of the largest) let the sweep
        // kernel step the phasor instead of calling sincos per system
        const double k0 = h_params[0].k;
        const double step = batch >= 2 ? (h_params[batch - 1].k - k0) / double(batch - 1) : 0.0;
        double kmax = 0.0;
        for (int64_t b = 0; b < batch; ++b) kmax = std::max(kmax, std::fabs(h_params[b].k));
        bool uniform = step != 0.0 && std::isfinite(step);
        for (int64_t b = 0; b < batch && uniform; ++b)
            uniform = std::fabs(h_params[b].k - (k0 + double(b) * step)) <= 4.0 * DBL_EPSILON * kmax;
        const double dk = uniform ? step : 0.0;
        for (int64_t b = 0; b < batch; ++b) {
            CCAB_CHECK(h_params[b].dh_sign == h_params[0].dh_sign,
                       "the systems of a batch share one dh convention");
            CCAB_CHECK(!bm || (h_params[b].beta.re == 0.0 &&
                               std::fabs(h_params[b].beta.im * h_params[b].k - 1.0) < 1e-14),
                       "Burton-Miller coupling beta = i/k expected");
        }
        // tuned systems per thread and column chunk (ccab_autotune_sweep); a
        // chunk is never narrower than kColChunk, so callers size the row-sum
        // partials for (n + kColChunk - 1) / kColChunk chunks
        const AsmTune* tn = asm_tune_of(ctx, prec_compute == PREC_F64_H ? CCAB_C32 : CCAB_C64, n);
        // (32 systems with row sums or A x0 would need 128 KiB of dynamic
        // LDS: 16 then; 8 with both)
        // (A x0: 8 systems per thread, 32 KiB of accumulators, so that the
        // LDS does not lower the occupancy the registers allow; measured
        // faster than 16)
        static const int ax_kf = [] {
            const char* e = std::getenv("CCAB_ASM_AX_KF");
            return e && atoi(e) == 16 ? 16 : 8;
        }();
        int kf = tn ? tn->kf : kF;
        if (d_ax_part) kf = std::min(kf, d_rowsum_part ? 8 : ax_kf);
        else if (d_rowsum_part) kf = std::min(kf, 16);
        // (the staged guess takes kf x cc x 16 bytes more: 64 columns then)
        const int cc = d_ax_part ? kColChunk
                                 : tn ? std::max(tn->col_chunk, kColChunk) : kColChunk;
        const int64_t nch = (n + cc - 1) / cc;
        if (nchunks_out) *nchunks_out = nch;
        dim3 g2(grid.x, (unsigned)nch, (unsigned)((batch + kf - 1) / kf));
        // row sums and A x0 in dynamic LDS, 16 B per frequency and lane each
        // (64 KiB at kf = 16)
        const size_t lds = size_t((d_rowsum_part ? 1 : 0) + (d_ax_part ? 1 : 0)) * kf * kRows *
                               sizeof(double2) +
                           (d_ax_part ? size_t(kf) * cc * sizeof(double2) : 0);
        CCAB_CHECK(lds <= ctx->lds_per_block, "assembly LDS (%zu bytes) exceeds the device's %zu",
                   lds, ctx->lds_per_block);
#define CCAB_ASM_SWEEP_LAUNCH(KF)                                                                  \
    do {                                                                                           \
        if (prec_compute == PREC_F64_H) {                                                          \
            if (bm) hipLaunchKernelGGL((assemble_sweep_kernel<true, c32, KF>), g2, dim3(kRows), lds, st, d_params, d_soa, n, batch, out0, out1, out2, ld, stride, d_rowsum_part, nch, dk, cc, d_x0, d_ax_part); \
            else hipLaunchKernelGGL((assemble_sweep_kernel<false, c32, KF>), g2, dim3(kRows), lds, st, d_params, d_soa, n, batch, out0, out1, out2, ld, stride, d_rowsum_part, nch, dk, cc, d_x0, d_ax_part); \
        } else {                                                                                   \
            if (bm) hipLaunchKernelGGL((assemble_sweep_kernel<true, c64, KF>), g2, dim3(kRows), lds, st, d_params, d_soa, n, batch, out0, out1, out2, ld, stride, d_rowsum_part, nch, dk, cc, d_x0, d_ax_part); \
            else hipLaunchKernelGGL((assemble_sweep_kernel<false, c64, KF>), g2, dim3(kRows), lds, st, d_params, d_soa, n, batch, out0, out1, out2, ld, stride, d_rowsum_part, nch, dk, cc, d_x0, d_ax_part); \
        }                                                                                          \
    } while (0)
        if (kf == 8) CCAB_ASM_SWEEP_LAUNCH(8);
        else if (kf == 32) CCAB_ASM_SWEEP_LAUNCH(32);
        else CCAB_ASM_SWEEP_LAUNCH(16);
#undef CCAB_ASM_SWEEP_LAUNCH
    } else if (mode == ASM_HG) {
        if (prec_compute == PREC_F64) {
            if (bm) launch_one<double, true, ASM_HG, c128>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, nullptr, nchunks);
            else launch_one<double, false, ASM_HG, c128>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, nullptr, nchunks);
        } else {
            if (bm) launch_one<float, true, ASM_HG, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, nullptr, nchunks);
            else launch_one<float, false, ASM_HG, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, nullptr, nchunks);
        }
    } else {
        if (prec_compute == PREC_F32) {
            CCAB_CHECK(out0 == nullptr, "FP32 assembly writes only the c64 LHS");
            if (bm) launch_one<float, true, ASM_LHS, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
            else launch_one<float, false, ASM_LHS, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
        } else if (prec_compute == PREC_F64_H) {
            if (bm) launch_one<double, true, ASM_LHS, c32>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
            else launch_one<double, false, ASM_LHS, c32>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
        } else {
            if (bm) launch_one<double, true, ASM_LHS, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
            else launch_one<double, false, ASM_LHS, c64>(grid, st, d_params, d_soa, n, out0, out1, ld, stride, d_rowsum_part, nchunks);
        }
    }
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_ax_finish(ccab_ctx* ctx, int64_t batch, int64_t n, const double* d_ax_part,
                     int64_t nchunks, c128* d_ax) {
    hipLaunchKernelGGL(ax_finish_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)batch),
                       dim3(256), 0, ctx->stream, n, d_ax_part, nchunks, d_ax);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_rhs_finish(ccab_ctx* ctx, const SysParams* d_params, int64_t batch, int64_t n,
                      const double* d_rowsum_part, int64_t nchunks, const ccab_c128* d_rhs_inc,
                      ccab_c128* d_rhs) {
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(rhs_finish_kernel, grid, dim3(256), 0, ctx->stream, d_params, n,
                       d_rowsum_part, nchunks, d_rhs_inc, d_rhs);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_incident_sweep(ccab_ctx* ctx, const SysParams* dp, const SysParams* hp,
                          const ccab_physics* ph, int64_t nfreq, const ccab_wave& w,
                          const double* d_soa, int64_t n, ccab_c128* d_rhs) {
    bool same = true;
    for (int64_t f = 1; f < nfreq; ++f)
        same &= (ph[f].corrections & CCAB_FIX_SPHERICAL_ORIGIN) ==
                (ph[0].corrections & CCAB_FIX_SPHERICAL_ORIGIN);
    if (!same || nfreq == 1 || nfreq > 65535) {
        for (int64_t f = 0; f < nfreq; ++f)
            CCAB_TRY(launch_incident(ctx, &hp[f], ph[f], w, d_soa, n, d_rhs + f * n));
        return CCAB_OK;
    }
    double d[3] = {w.vec[0], w.vec[1], w.vec[2]};
    const double nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(nn > 0.0)) return ctx_fail(ctx, CCAB_ENAN, "incident wave vector has zero length");
    if (w.kind == CCAB_WAVE_PLANE || !(ph[0].corrections & CCAB_FIX_SPHERICAL_ORIGIN))
        for (double& v : d) v /= nn;
    const c128 pamp{w.amplitude[0], w.amplitude[1]};
    std::vector<c128> amps(nfreq);
    for (int64_t f = 0; f < nfreq; ++f)
        amps[f] = cdiv(pamp, c128{0.0, hp[f].omega * ph[f].mass_density});
    c128* d_amps = static_cast<c128*>(ctx_buf(ctx, SLOT_INC_AMP, size_t(nfreq) * 16));
    if (!d_amps) return CCAB_ENOMEM;
    CCAB_HIP(hipMemcpyAsync(d_amps, amps.data(), size_t(nfreq) * 16, hipMemcpyHostToDevice,
                            ctx->stream));
    CCAB_HIP(hipStreamSynchronize(ctx->stream));  // amps is a host temporary
    hipLaunchKernelGGL(incident_sweep_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nfreq),
                       dim3(256), 0, ctx->stream, dp, d_amps, w.kind, d[0], d[1], d[2], d_soa, n,
                       d_rhs);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_incident(ccab_ctx* ctx, const SysParams* hp, const ccab_physics& ph,
                    const ccab_wave& w, const double* d_soa, int64_t n, ccab_c128* d_rhs) {
    double d[3] = {w.vec[0], w.vec[1], w.vec[2]};
    const double nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(nn > 0.0)) return ctx_fail(ctx, CCAB_ENAN, "incident wave vector has zero length");
    // normalize(); the spherical origin moves to the unit sphere unless
    // CCAB_FIX_SPHERICAL_ORIGIN keeps it where it is
    if (w.kind == CCAB_WAVE_PLANE || !(ph.corrections & CCAB_FIX_SPHERICAL_ORIGIN))
        for (double& v : d) v /= nn;
    const c128 pamp{w.amplitude[0], w.amplitude[1]};
    const c128 amp = cdiv(pamp, c128{0.0, hp->omega * ph.mass_density});
    hipLaunchKernelGGL(incident_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, *hp,
                       w.kind, d[0], d[1], d[2], amp, d_soa, n, d_rhs);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}

int launch_rowcol(ccab_ctx* ctx, const SysParams& p, const double* d_soa, int64_t n,
                  const int64_t* d_i, int64_t ni, const int64_t* d_j, int64_t nj, int side,
                  ccab_c128* d_out) {
    const int64_t cnt = ni == 1 ? nj : ni;
    if (cnt == 0) return CCAB_OK;
    dim3 grid((unsigned)((cnt + 255) / 256));
    if (p.bm)
        hipLaunchKernelGGL(rowcol_kernel<true>, grid, dim3(256), 0, ctx->stream, p, d_soa, n, d_i,
                           ni, d_j, nj, side, d_out);
    else
        hipLaunchKernelGGL(rowcol_kernel<false>, grid, dim3(256), 0, ctx->stream, p, d_soa, n,
                           d_i, ni, d_j, nj, side, d_out);
    CCAB_HIP(hipGetLastError());
    return CCAB_OK;
}
