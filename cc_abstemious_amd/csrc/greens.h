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

// Off-diagonal kernel pair (g, h) for source point j (x, n_x) and collocation
// point i (y, n_y), r = x - y.  influence_matrix.rs:12-37.  dh_sign: -1 as
// the reference (influence_matrix.rs:30), +1 for d^2 G / dn_x dn_y.
template <typename T, bool BM>
__device__ __forceinline__ void greens(T k, Cx<T> beta, T rx, T ry, T rz, T nxx, T nxy, T nxz,
                                       T nyx, T nyy, T nyz, Cx<T>& g, Cx<T>& h,
                                       T dh_sign = T(-1)) {
    const T rd = sqrt_t(rx * rx + ry * ry + rz * rz);
    const T inv_r = T(1) / rd;
    const T erx = rx * inv_r, ery = ry * inv_r, erz = rz * inv_r;
    T s, c;
    sincos_t(k * rd, &s, &c);
    const T gs = inv_r * T(1.0 / (4.0 * kPi));
    g = Cx<T>{c * gs, s * gs};
    const Cx<T> f1{-inv_r, k};
    const T rdoty = erx * nyx + ery * nyy + erz * nyz;
    const Cx<T> gf1 = g * f1;
    h = gf1 * (-rdoty);
    if (BM) {
        const T rdotx = erx * nxx + ery * nxy + erz * nxz;
        const Cx<T> dg = gf1 * rdotx;
        const T nn = nyx * nxx + nyy * nxy + nyz * nxz;
        const T rr = rdotx * rdoty * dh_sign;
        // -f1 (n_y.n_x) - (k^2 r + 3 f1)(e.n_x)(e.n_y)  (dh_sign = -1)
        const Cx<T> term{-f1.re * nn + (k * k * rd + T(3) * f1.re) * rr,
                         -f1.im * nn + T(3) * f1.im * rr};
        const Cx<T> dh = (g * inv_r) * term;
        g += beta * dg;
        h += beta * dh;
    }
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
