// Device code shared by the assembly and H-matrix kernels: SoA collocation
// points, the Helmholtz kernel pair (influence_matrix.rs:12-37) and the
// singular diagonal (influence_matrix.rs:39-50).
#pragma once
#include "ccab_internal.h"

struct Soa {
    const double *x, *y, *z, *nx, *ny, *nz, *area, *wt;
};

__device__ __forceinline__ Soa soa_of(const double* base, int64_t n) {
    return Soa{base, base + n, base + 2 * n, base + 3 * n,
               base + 4 * n, base + 5 * n, base + 6 * n, base + 7 * n};
}

__device__ __forceinline__ void sincos_t(double a, double* s, double* c) { sincos(a, s, c); }
__device__ __forceinline__ void sincos_t(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ double sqrt_t(double a) { return sqrt(a); }
__device__ __forceinline__ float sqrt_t(float a) { return sqrtf(a); }

// Geometry of the pair (source point j at x with normal n_x, collocation
// point i at y with normal n_y), r = x - y: everything of influence_matrix.rs:
// 12-37 that does not depend on the wavenumber.
template <typename T>
struct PairGeom {
    T rd, inv_r, rdoty, rdotx, nn;
};

template <typename T, bool BM>
__device__ __forceinline__ PairGeom<T> pair_geom(T rx, T ry, T rz, T nxx, T nxy, T nxz, T nyx,
                                                 T nyy, T nyz) {
    PairGeom<T> q;
    q.rd = sqrt_t(rx * rx + ry * ry + rz * rz);
    q.inv_r = T(1) / q.rd;
    const T erx = rx * q.inv_r, ery = ry * q.inv_r, erz = rz * q.inv_r;
    q.rdoty = erx * nyx + ery * nyy + erz * nyz;
    q.rdotx = BM ? erx * nxx + ery * nxy + erz * nxz : T(0);
    q.nn = BM ? nyx * nxx + nyy * nxy + nyz * nxz : T(0);
    return q;
}

// pair_geom in FP64 for the assembly kernels, with 1/r from the hardware
// reciprocal square root (relative error e0 below 2^-22) and one third-order
// Newton correction y <- y (1 + e/2 + 3e^2/8), e = 1 - r^2 y^2, which leaves
// about 2.5 e0^3 (under an FP64 ulp) and takes 7 instructions instead of the
// 26 of an IEEE sqrt and divide; then r = r^2 y.  Needs r > 0 (the diagonal
// lane of the assembly kernels evaluates a unit-distance pair).
template <bool BM>
__device__ __forceinline__ PairGeom<double> pair_geom_fast(double rx, double ry, double rz,
                                                           double nxx, double nxy, double nxz,
                                                           double nyx, double nyy, double nyz) {
    PairGeom<double> q;
    const double r2 = fma(rx, rx, fma(ry, ry, rz * rz));
    double y = __builtin_amdgcn_rsq(r2);
    const double e = fma(-(r2 * y), y, 1.0);
    y = fma(y, fma(e, 0.375, 0.5) * e, y);
    q.inv_r = y;
    q.rd = r2 * y;
    q.rdoty = fma(rx, nyx, fma(ry, nyy, rz * nyz)) * y;
    q.rdotx = BM ? fma(rx, nxx, fma(ry, nxy, rz * nxz)) * y : 0.0;
    q.nn = BM ? fma(nyx, nxx, fma(nyy, nxy, nyz * nxz)) : 0.0;
    return q;
}

// sin and cos of a: a = q pi/2 + r by Cody and Waite's reduction with
// pi/2 = P1 + P2 + P3 (three doubles; each product is exact inside the fused
// multiply-add, so r carries an absolute error of about 1.5e-16 for
// |a| <= 2^16), then sin r = r + r z S(z) and cos r = 1 - z/2 + z^2 C(z),
// z = r^2, with S and C of degree 5 from Chebyshev interpolation of
// (sin(sqrt z)/sqrt z - 1)/z and (cos(sqrt z) - 1 + z/2)/z^2 on
// 0 <= z <= (pi/4)^2 (interpolation error 2e-17 and 1.3e-18; measured
// against long double: 1.3e-16 absolute for |a| <= 1e5), and the quadrant
// q mod 4.  About 40 instructions against about 70 for the library sincos,
// which takes |a| > 2^16 and non-finite a.
__device__ __forceinline__ void sincos_fast(double a, double* s, double* c) {
    if (!(fabs(a) <= 65536.0)) {
        sincos(a, s, c);
        return;
    }
    const double q = rint(a * 0.6366197723675814);  // 2/pi
    double r = fma(-q, 0x1.921fb54442d18p+0, a);    // P1 = pi/2 rounded to a double
    r = fma(-q, 0x1.1a62633145c07p-54, r);          // P2 = pi/2 - P1 rounded
    r = fma(-q, -0x1.f1976b7ed8fbcp-110, r);        // P3 = pi/2 - P1 - P2 rounded
    const double z = r * r;
    double ps = 1.5918115263265974e-10;
    ps = fma(ps, z, -2.505113165023518e-08);
    ps = fma(ps, z, 2.7557316101617874e-06);
    ps = fma(ps, z, -0.00019841269836756774);
    ps = fma(ps, z, 0.008333333333330948);
    ps = fma(ps, z, -0.16666666666666666);
    const double sr = fma(r * z, ps, r);
    double pc = -1.1382623647474604e-11;
    pc = fma(pc, z, 2.087614614655861e-09);
    pc = fma(pc, z, -2.7557317271145144e-07);
    pc = fma(pc, z, 2.480158729876456e-05);
    pc = fma(pc, z, -0.0013888888888887398);
    pc = fma(pc, z, 0.041666666666666664);
    const double cr = fma(z, fma(z, pc, -0.5), 1.0);
    const int n = int(q);
    const double sv = (n & 1) ? cr : sr;
    const double cv = (n & 1) ? sr : cr;
    *s = (n & 2) ? -sv : sv;
    *c = ((n + 1) & 2) ? -cv : cv;
}

// Kernel pair (g, h) at wavenumber k from the geometry and the phasor
// (c, s) = (cos kr, sin kr).  dh_sign: -1 as the reference
// (influence_matrix.rs:30), +1 for d^2 G / dn_x dn_y.
template <typename T, bool BM>
__device__ __forceinline__ void greens_at(const PairGeom<T>& q, T k, Cx<T> beta, T c, T s,
                                          Cx<T>& g, Cx<T>& h, T dh_sign) {
    const T gs = q.inv_r * T(1.0 / (4.0 * kPi));
    g = Cx<T>{c * gs, s * gs};
    const Cx<T> f1{-q.inv_r, k};
    const Cx<T> gf1 = g * f1;
    h = gf1 * (-q.rdoty);
    if (BM) {
        const Cx<T> dg = gf1 * q.rdotx;
        const T rr = q.rdotx * q.rdoty * dh_sign;
        // -f1 (n_y.n_x) - (k^2 r + 3 f1)(e.n_x)(e.n_y)  (dh_sign = -1)
        const Cx<T> term{-f1.re * q.nn + (k * k * q.rd + T(3) * f1.re) * rr,
                         -f1.im * q.nn + T(3) * f1.im * rr};
        const Cx<T> dh = (g * q.inv_r) * term;
        g += beta * dg;
        h += beta * dh;
    }
}

// Off-diagonal kernel pair (g, h) for source point j (x, n_x) and collocation
// point i (y, n_y), r = x - y.  influence_matrix.rs:12-37.
template <typename T, bool BM>
__device__ __forceinline__ void greens(T k, Cx<T> beta, T rx, T ry, T rz, T nxx, T nxy, T nxz,
                                       T nyx, T nyy, T nyz, Cx<T>& g, Cx<T>& h,
                                       T dh_sign = T(-1)) {
    const PairGeom<T> q = pair_geom<T, BM>(rx, ry, rz, nxx, nxy, nxz, nyx, nyy, nyz);
    T s, c;
    sincos_t(k * q.rd, &s, &c);
    greens_at<T, BM>(q, k, beta, c, s, g, h, dh_sign);
}

// greens in FP64 with pair_geom_fast and sincos_fast (the assembly, row /
// column and H-matrix block kernels; r > 0)
template <bool BM>
__device__ __forceinline__ void greens_fast(double k, c128 beta, double rx, double ry, double rz,
                                            double nxx, double nxy, double nxz, double nyx,
                                            double nyy, double nyz, c128& g, c128& h,
                                            double dh_sign = -1.0) {
    const PairGeom<double> q = pair_geom_fast<BM>(rx, ry, rz, nxx, nxy, nxz, nyx, nyy, nyz);
    double s, c;
    sincos_fast(k * q.rd, &s, &c);
    greens_at<double, BM>(q, k, beta, c, s, g, h, dh_sign);
}

// Singular diagonal (influence_matrix.rs:42-50) incl. the global shifts of
// get_dense_surface_matrices (influence_matrix.rs:78-79).  Always FP64.
__host__ __device__ __forceinline__ void diag_entry(const SysParams& p, double area, double wt,
                                                    c128& g, c128& h) {
    const double b = sqrt(area / kPi);
    const double wa = wt * area;
    g = c128{1.0 / (2.0 * kPi * b) * wa, 0.0};
    h = c128{0.0, 0.0};
    if (p.bm) {
        const double s = -1.0 / (2.0 * kPi * (b * b * b));
        h = c128{p.beta.re * s * wa, p.beta.im * s * wa};
    }
    g += p.gdiag;
    h += p.hdiag;
}


// One entry of the side matrix a H + b G (LHS or RHS factors,
// influence_matrix.rs:130-162) with the global diagonal shifts at i == j.
__device__ __forceinline__ c128 side_entry(const SysParams& p, const Soa& s, int64_t i, int64_t j,
                                           c128 fa, c128 fb) {
    c128 G, H;
    if (i == j) {
        diag_entry(p, s.area[j], s.wt[j], G, H);
    } else {
        c128 g, h;
        if (p.bm)
            greens_fast<true>(p.k, p.beta, s.x[j] - s.x[i], s.y[j] - s.y[i], s.z[j] - s.z[i],
                              s.nx[j], s.ny[j], s.nz[j], s.nx[i], s.ny[i], s.nz[i], g, h,
                              p.dh_sign);
        else
            greens_fast<false>(p.k, p.beta, s.x[j] - s.x[i], s.y[j] - s.y[i], s.z[j] - s.z[i],
                               s.nx[j], s.ny[j], s.nz[j], s.nx[i], s.ny[i], s.nz[i], g, h);
        const double wa = s.wt[j] * s.area[j];
        G = g * wa;
        H = h * wa;
    }
    return fb * G + fa * H;
}
