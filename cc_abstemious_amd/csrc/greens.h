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
            greens<double, true>(p.k, p.beta, s.x[j] - s.x[i], s.y[j] - s.y[i], s.z[j] - s.z[i],
                                 s.nx[j], s.ny[j], s.nz[j], s.nx[i], s.ny[i], s.nz[i], g, h,
                                 p.dh_sign);
        else
            greens<double, false>(p.k, p.beta, s.x[j] - s.x[i], s.y[j] - s.y[i], s.z[j] - s.z[i],
                                  s.nx[j], s.ny[j], s.nz[j], s.nx[i], s.ny[i], s.nz[i], g, h);
        const double wa = s.wt[j] * s.area[j];
        G = g * wa;
        H = h * wa;
    }
    return fb * G + fa * H;
}
