/*
 * ccab_oracle.c -- CPU FP64 restatement of the cc_abstemious BEM hot path.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * ccab_oracle.h for the reference file:line each function follows and for
 * what pins it.
 *
 * Built with -fcx-limited-range so complex * and / are the textbook
 * formulas, as num-complex 0.4.6 implements them.
 */
#define _GNU_SOURCE
#include "ccab_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef double complex cplx;
#define C(p) ((cplx*)(p))
#define CC(p) ((const cplx*)(p))

static const double PI = 3.14159265358979323846;

static inline double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm3(const double* a) { return sqrt(dot3(a, a)); }
/* num-complex norm() is hypot(re, im) */
static inline double cabs_h(cplx z) { return hypot(creal(z), cimag(z)); }
static inline double cnorm_sqr(cplx z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }
/* nalgebra vector norm: sqrt(sum |z|^2) */
static double vnorm(const cplx* v, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += cnorm_sqr(v[i]);
    return sqrt(s);
}
/* num-complex Complex::exp for a purely imaginary argument i*t */
static inline cplx cexpi(double t) { return cos(t) + I * sin(t); }

/* ------------------------------------------------------------------ */
/* Geometry: elements.rs:49-181, preprocess.rs:126-153                 */
/* ------------------------------------------------------------------ */
int orc_collocation_points(const double* nodes, const int64_t* conn,
                           const int32_t* nverts, int64_t nelem, double* cpts) {
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t* cn = conn + 4 * e;
        double* out = cpts + ORC_CPT_STRIDE * e;
        double x[3] = {0, 0, 0}, nu[3];
        if (nverts[e] == 3) {
            /* TRIGP1 at (1/3, 1/3), weight 1 (elements.rs:11,75-77) */
            const double xi = 1. / 3., eta = 1. / 3.;
            const double N[3] = {1.0 - xi - eta, xi, eta};
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) x[d] += N[a] * nodes[3 * cn[a] + d];
            const double* p0 = nodes + 3 * cn[0];
            const double* p1 = nodes + 3 * cn[1];
            const double* p2 = nodes + 3 * cn[2];
            double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            nu[0] = a[1] * b[2] - a[2] * b[1];
            nu[1] = a[2] * b[0] - a[0] * b[2];
            nu[2] = a[0] * b[1] - a[1] * b[0];
            double nn = norm3(nu);
            out[6] = 0.5 * nn; /* detj (elements.rs:167) */
            out[7] = 1.0;
            for (int d = 0; d < 3; ++d) out[3 + d] = nu[d] / nn;
        } else if (nverts[e] == 4) {
            /* QUADGP1 at (0,0), weight 4 (elements.rs:29,80-83,98-102,168) */
            const double dN[4][2] = {{-0.25, -0.25}, {0.25, -0.25}, {0.25, 0.25}, {-0.25, 0.25}};
            double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0};
            for (int a = 0; a < 4; ++a) {
                const double* p = nodes + 3 * cn[a];
                const double N = 0.25;
                for (int d = 0; d < 3; ++d) {
                    x[d] += N * p[d];
                    dxi[d] += dN[a][0] * p[d];
                    deta[d] += dN[a][1] * p[d];
                }
            }
            nu[0] = dxi[1] * deta[2] - dxi[2] * deta[1];
            nu[1] = dxi[2] * deta[0] - dxi[0] * deta[2];
            nu[2] = dxi[0] * deta[1] - dxi[1] * deta[0];
            double nn = norm3(nu);
            out[6] = 0.25 * nn; /* the reference's extra 1/4 (elements.rs:168) */
            out[7] = 4.0;
            for (int d = 0; d < 3; ++d) out[3 + d] = nu[d] / nn;
        } else {
            return -1;
        }
        out[0] = x[0];
        out[1] = x[1];
        out[2] = x[2];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Constants: preprocess.rs:41-74                                      */
/* ------------------------------------------------------------------ */
void orc_constants(const orc_physics* ph, double* k, orc_c128* beta, orc_c128* hdiag,
                   orc_c128* gdiag) {
    double omega = 2.0 * PI * ph->frequency;
    double kk = omega / ph->sound_speed;
    cplx b = 0.0 + I * (1.0 / kk);
    if (k) *k = kk;
    if (beta) *C(beta) = b;
    if (hdiag) *C(hdiag) = ph->problem_type == ORC_EXTERIOR ? -0.5 : 0.0;
    if (gdiag) *C(gdiag) = ph->method == ORC_BURTON_MILLER ? 0.5 * b : 0.0;
}

/* ------------------------------------------------------------------ */
/* Kernel: influence_matrix.rs:12-60                                   */
/* ------------------------------------------------------------------ */
static void greens(double k, const double* x, const double* n_x, const double* y,
                   const double* n_y, int bm, cplx beta, cplx* gout, cplx* hout) {
    double nxn = norm3(n_x), nyn = norm3(n_y);
    double e_nx[3] = {n_x[0] / nxn, n_x[1] / nxn, n_x[2] / nxn};
    double e_ny[3] = {n_y[0] / nyn, n_y[1] / nyn, n_y[2] / nyn};
    double r[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double rdist = norm3(r);
    double e_r[3] = {r[0] / rdist, r[1] / rdist, r[2] / rdist};
    cplx g = cexpi(k * rdist) / (4.0 * PI * rdist);
    cplx f1 = -1.0 / rdist + I * k;
    double rdoty = dot3(e_r, e_ny);
    cplx h = g * f1 * (-rdoty);
    if (bm) {
        double rdotx = dot3(e_r, e_nx);
        cplx dg = g * f1 * rdotx;
        /* sign of the (k^2 r + 3 f1) term exactly as influence_matrix.rs:30 */
        cplx dh = (1.0 / rdist) * g *
                  (-f1 * dot3(e_ny, e_nx) - (k * k * rdist + 3.0 * f1) * rdotx * rdoty);
        g += beta * dg;
        h += beta * dh;
    }
    *gout = g;
    *hout = h;
}

void orc_greens(double k, const double* x, const double* nx, const double* y,
                const double* ny, int bm, orc_c128 beta, orc_c128* g, orc_c128* h) {
    greens(k, x, nx, y, ny, bm, *CC(&beta), C(g), C(h));
}

static void gh_entry(double k, int bm, cplx beta, const double* cpts, int64_t i, int64_t j,
                     cplx* g, cplx* h) {
    const double* cj = cpts + ORC_CPT_STRIDE * j;
    double area = cj[6], wt = cj[7];
    if (i == j) {
        double b = sqrt(area / PI);
        cplx gg = 1.0 / (2.0 * PI * b);
        cplx hh = 0.0;
        if (bm) hh += beta * (-1.0 / (2.0 * PI * (b * b * b)));
        *g = gg * wt * area;
        *h = hh * wt * area;
        return;
    }
    const double* ci = cpts + ORC_CPT_STRIDE * i;
    cplx gg, hh;
    greens(k, cj, cj + 3, ci, ci + 3, bm, beta, &gg, &hh);
    *g = gg * wt * area;
    *h = hh * wt * area;
}

void orc_gh_entry(const orc_physics* ph, const double* cpts, int64_t i, int64_t j,
                  orc_c128* g, orc_c128* h) {
    double k;
    orc_c128 beta;
    orc_constants(ph, &k, &beta, NULL, NULL);
    gh_entry(k, ph->method == ORC_BURTON_MILLER, *C(&beta), cpts, i, j, C(g), C(h));
}

/* ------------------------------------------------------------------ */
/* Dense assembly: influence_matrix.rs:64-98 (threaded, per-row tasks) */
/* ------------------------------------------------------------------ */
typedef struct {
    double k;
    int bm;
    cplx beta;
    const double* cpts;
    int64_t n;
    cplx* H;
    cplx* G;
    int locked;
    pthread_mutex_t* mtx;   /* H, then G: the two locks of influence_matrix.rs:87-88 */
    pthread_mutex_t* mtx_g;
    int64_t next_row;
    pthread_mutex_t* qmtx;
} asm_job;

static void* asm_worker(void* arg) {
    asm_job* jb = (asm_job*)arg;
    for (;;) {
        pthread_mutex_lock(jb->qmtx);
        int64_t i = jb->next_row++;
        pthread_mutex_unlock(jb->qmtx);
        if (i >= jb->n) break;
        for (int64_t j = 0; j < jb->n; ++j) {
            cplx g, h;
            gh_entry(jb->k, jb->bm, jb->beta, jb->cpts, i, j, &g, &h);
            if (jb->locked) {
                pthread_mutex_lock(jb->mtx);
                pthread_mutex_lock(jb->mtx_g);
            }
            jb->H[i + j * jb->n] += h;
            jb->G[i + j * jb->n] += g;
            if (jb->locked) {
                pthread_mutex_unlock(jb->mtx_g);
                pthread_mutex_unlock(jb->mtx);
            }
        }
    }
    return NULL;
}

static int default_threads(void) {
    long p = sysconf(_SC_NPROCESSORS_ONLN);
    int t = (int)(p / 2);
    return t < 2 ? 2 : t;
}

int orc_dense_surface_matrices(const orc_physics* ph, const double* cpts, int64_t n,
                               orc_c128* Hc, orc_c128* Gc, int nthreads, int locked) {
    double k;
    orc_c128 beta, hd, gd;
    orc_constants(ph, &k, &beta, &hd, &gd);
    cplx* H = C(Hc);
    cplx* G = C(Gc);
    for (int64_t p = 0; p < n * n; ++p) H[p] = G[p] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        H[i + i * n] = *C(&hd);
        G[i + i * n] = *C(&gd);
    }
    if (nthreads <= 0) nthreads = default_threads();
    pthread_mutex_t mtx = PTHREAD_MUTEX_INITIALIZER, mtx_g = PTHREAD_MUTEX_INITIALIZER,
                    qmtx = PTHREAD_MUTEX_INITIALIZER;
    asm_job jb = {k, ph->method == ORC_BURTON_MILLER, *C(&beta), cpts, n, H, G, locked,
                  &mtx, &mtx_g, 0, &qmtx};
    if (nthreads == 1) {
        asm_worker(&jb);
        return 0;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, asm_worker, &jb);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    return 0;
}

/* influence_matrix.rs:101-127 */
int orc_dense_field_matrices(const orc_physics* ph, const double* cpts, int64_t n,
                             const double* fps, int64_t nfp, orc_c128* Mc, orc_c128* Lc) {
    double k;
    orc_constants(ph, &k, NULL, NULL, NULL);
    cplx* M = C(Mc);
    cplx* L = C(Lc);
    const double nx_dummy[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
        const double* cj = cpts + ORC_CPT_STRIDE * j;
        for (int64_t p = 0; p < nfp; ++p) {
            cplx g, h;
            greens(k, fps + 3 * p, nx_dummy, cj, cj + 3, 0, 0.0, &g, &h);
            M[p + j * nfp] = 0.0 + h * cj[7] * cj[6];
            L[p + j * nfp] = 0.0 + g * cj[7] * cj[6];
        }
    }
    return 0;
}

/* influence_matrix.rs:130-162 */
void orc_lhs_factors(const orc_physics* ph, orc_c128* a, orc_c128* b) {
    switch (ph->bc_type) {
    case ORC_BC_PRESSURE: *C(a) = 0.0; *C(b) = 1.0; break;
    case ORC_BC_NORMAL_VELOCITY: *C(a) = 1.0; *C(b) = 0.0; break;
    default: {
        cplx z = ph->bc_value[0] + I * ph->bc_value[1];
        double omega = 2.0 * PI * ph->frequency;
        *C(a) = 1.0;
        *C(b) = (0.0 - I * (omega * ph->mass_density)) / z;
    }
    }
}
void orc_rhs_factors(const orc_physics* ph, orc_c128* a, orc_c128* b) {
    switch (ph->bc_type) {
    case ORC_BC_PRESSURE: *C(a) = -1.0; *C(b) = 0.0; break;
    case ORC_BC_NORMAL_VELOCITY: *C(a) = 0.0; *C(b) = 1.0; break;
    default: *C(a) = 0.0; *C(b) = 0.0;
    }
}

/* influence_matrix.rs:170-231 */
int orc_surface_row_or_column(const orc_physics* ph, const double* cpts, int side,
                              const int64_t* ii, int64_t ni, const int64_t* jj, int64_t nj,
                              orc_c128* outc) {
    if (!((ni == 1) ^ (nj == 1))) return -1;
    double k;
    orc_c128 beta, hd, gd, al, be;
    orc_constants(ph, &k, &beta, &hd, &gd);
    if (side == 0) orc_lhs_factors(ph, &al, &be);
    else orc_rhs_factors(ph, &al, &be);
    int bm = ph->method == ORC_BURTON_MILLER;
    cplx* out = C(outc);
    int64_t cnt = ni == 1 ? nj : ni;
    /* the diagonal shift is added once, at the first position holding the
     * diagonal index (Iterator::position, influence_matrix.rs:206,225) */
    int64_t diag = -1;
    for (int64_t t = 0; t < cnt; ++t) {
        int64_t i = ni == 1 ? ii[0] : ii[t];
        int64_t j = ni == 1 ? jj[t] : jj[0];
        if (i == j) { diag = t; break; }
    }
    for (int64_t t = 0; t < cnt; ++t) {
        int64_t i = ni == 1 ? ii[0] : ii[t];
        int64_t j = ni == 1 ? jj[t] : jj[0];
        cplx g, h;
        gh_entry(k, bm, *C(&beta), cpts, i, j, &g, &h);
        cplx hh = 0.0 + h, gg = 0.0 + g;
        if (t == diag) {
            hh += *C(&hd);
            gg += *C(&gd);
        }
        /* h.axpy(beta, g, alpha): h = beta*g + alpha*h */
        out[t] = *C(&be) * gg + *C(&al) * hh;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Incident waves: incident_wave.rs:7-100                              */
/* ------------------------------------------------------------------ */
static cplx wave_amp(const orc_physics* ph, const orc_wave* w) {
    cplx p_amp = w->amplitude[0] + I * w->amplitude[1];
    double omega = 2.0 * PI * ph->frequency;
    return p_amp / (0.0 + I * (omega * ph->mass_density));
}

int orc_incident_surface(const orc_physics* ph, const orc_wave* w, const double* cpts,
                         int64_t n, orc_c128* rhsc) {
    double k;
    orc_c128 beta;
    orc_constants(ph, &k, &beta, NULL, NULL);
    int bm = ph->method == ORC_BURTON_MILLER;
    cplx* rhs = C(rhsc);
    cplx amp = wave_amp(ph, w);
    double vn = norm3(w->vec);
    double d[3] = {w->vec[0] / vn, w->vec[1] / vn, w->vec[2] / vn};
    for (int64_t i = 0; i < n; ++i) {
        const double* c = cpts + ORC_CPT_STRIDE * i;
        if (w->kind == ORC_WAVE_PLANE) {
            cplx phi = amp * cexpi(k * dot3(d, c));
            rhs[i] = phi;
            if (bm) {
                cplx vninc = phi * (0.0 + I * k) * dot3(d, c + 3);
                rhs[i] += *C(&beta) * vninc;
            }
        } else {
            /* origin normalised to unit length (incident_wave.rs:44) */
            double rv[3] = {c[0] - d[0], c[1] - d[1], c[2] - d[2]};
            double r = norm3(rv);
            double e_r[3] = {rv[0] / r, rv[1] / r, rv[2] / r};
            cplx phi = amp * cexpi(k * r) / (4.0 * PI * r);
            rhs[i] = phi;
            if (bm) {
                cplx vninc = phi * (-1.0 / r + I * k) * dot3(e_r, c + 3);
                rhs[i] += *C(&beta) * vninc;
            }
        }
    }
    return 0;
}

int orc_incident_field(const orc_physics* ph, const orc_wave* w, const double* fps,
                       int64_t nfp, orc_c128* phic) {
    double k;
    orc_constants(ph, &k, NULL, NULL, NULL);
    cplx* phi = C(phic);
    cplx amp = wave_amp(ph, w);
    double vn = norm3(w->vec);
    double d[3] = {w->vec[0] / vn, w->vec[1] / vn, w->vec[2] / vn};
    for (int64_t p = 0; p < nfp; ++p) {
        const double* x = fps + 3 * p;
        if (w->kind == ORC_WAVE_PLANE) {
            phi[p] = amp * cexpi(k * dot3(d, x));
        } else {
            double rv[3] = {x[0] - d[0], x[1] - d[1], x[2] - d[2]};
            double r = norm3(rv);
            phi[p] = amp * cexpi(k * r) / (4.0 * PI * r);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* LU: nalgebra 0.33.2 (solve.rs:166-171)                              */
/* ------------------------------------------------------------------ */
static inline double norm1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

int orc_lu_factor(int64_t n, orc_c128* Ac, int64_t lda, int64_t* piv) {
    cplx* A = C(Ac);
    int singular = 0;
    for (int64_t i = 0; i < n; ++i) {
        /* icamax on norm1 over rows i.. of column i; first max wins */
        int64_t p = i;
        double best = norm1(A[i + i * lda]);
        for (int64_t r = i + 1; r < n; ++r) {
            double v = norm1(A[r + i * lda]);
            if (v > best) { best = v; p = r; }
        }
        piv[i] = p;
        cplx diag = A[p + i * lda];
        if (diag == 0.0) { singular = 1; continue; }
        if (p != i) {
            for (int64_t c = 0; c < n; ++c) {
                cplx t = A[i + c * lda];
                A[i + c * lda] = A[p + c * lda];
                A[p + c * lda] = t;
            }
        }
        cplx inv = 1.0 / diag;
        for (int64_t r = i + 1; r < n; ++r) A[r + i * lda] *= inv;
        for (int64_t c = i + 1; c < n; ++c) {
            cplx u = A[i + c * lda];
            cplx* col = A + c * lda;
            const cplx* l = A + i * lda;
            for (int64_t r = i + 1; r < n; ++r) col[r] += -u * l[r];
        }
    }
    return singular;
}

void orc_lu_apply(int64_t n, const orc_c128* LUc, int64_t lda, const int64_t* piv,
                  orc_c128* xc) {
    const cplx* A = CC(LUc);
    cplx* x = C(xc);
    for (int64_t i = 0; i < n; ++i) {
        int64_t p = piv[i];
        if (p != i) { cplx t = x[i]; x[i] = x[p]; x[p] = t; }
    }
    /* unit lower forward substitution (column oriented) */
    for (int64_t j = 0; j < n; ++j) {
        cplx xj = x[j];
        for (int64_t r = j + 1; r < n; ++r) x[r] += -xj * A[r + j * lda];
    }
    /* upper back substitution */
    for (int64_t j = n - 1; j >= 0; --j) {
        x[j] = x[j] / A[j + j * lda];
        cplx xj = x[j];
        for (int64_t r = 0; r < j; ++r) x[r] += -xj * A[r + j * lda];
    }
}

int orc_lu_solve(int64_t n, const orc_c128* A, orc_c128* x) {
    orc_c128* W = (orc_c128*)malloc(sizeof(orc_c128) * n * n);
    int64_t* piv = (int64_t*)malloc(sizeof(int64_t) * n);
    memcpy(W, A, sizeof(orc_c128) * n * n);
    int s = orc_lu_factor(n, W, n, piv);
    orc_lu_apply(n, W, n, piv, x);
    free(W);
    free(piv);
    return s;
}

/* ------------------------------------------------------------------ */
/* GMRES: gmres.rs:35-202                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    const cplx* A;
    int64_t n;
} dense_op;

static void dense_mv(void* user, const orc_c128* xc, orc_c128* yc) {
    dense_op* d = (dense_op*)user;
    const cplx* x = CC(xc);
    cplx* y = C(yc);
    int64_t n = d->n;
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        cplx xj = x[j];
        const cplx* col = d->A + j * n;
        for (int64_t i = 0; i < n; ++i) y[i] += xj * col[i];
    }
}

static double backward_error(double r, double alpha, const cplx* x, int64_t n, double beta) {
    return r / (alpha * vnorm(x, n) + beta);
}

static void givens(cplx v1, cplx v2, cplx* cs, cplx* sn) {
    cplx signf = (v1 == 0.0) ? 1.0 : v1 / cabs_h(v1);
    double r = sqrt(cnorm_sqr(v1) + cnorm_sqr(v2));
    *cs = cabs_h(v1) / r;
    *sn = signf * conj(v2) / r;
}

int orc_gmres_op(int64_t n, orc_matvec_fn mv, void* user, double anorm, orc_c128* xc,
                 const orc_c128* x0, int64_t max_it, double thresh, int ref_mgs,
                 int64_t* stats, double* err_out) {
    cplx* x = C(xc);
    cplx* b = (cplx*)malloc(sizeof(cplx) * n);
    memcpy(b, x, sizeof(cplx) * n);
    if (x0) memcpy(x, x0, sizeof(cplx) * n);
    else for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
    if (max_it <= 0) max_it = 1000;
    if (max_it > n) max_it = n;
    if (thresh <= 0.0) thresh = 1.0e-5;
    int64_t m = (int64_t)ceil(sqrt((double)max_it));
    int64_t max_restarts = max_it / m;
    int64_t num_mv = 0, iters = 0, restarts = 0;
    double b_norm = vnorm(b, n);
    /* the operator may compute y = A x + 0 * y (h_matrix.rs:219-223 scales b
     * by beta): start every output vector finite, as the reference's
     * DVector allocations are */
    cplx* r = (cplx*)calloc(n, sizeof(cplx));
    cplx* tmp = (cplx*)calloc(n, sizeof(cplx));
    cplx* q = (cplx*)malloc(sizeof(cplx) * n * (m + 1));
    cplx* qk1 = (cplx*)calloc(n, sizeof(cplx));
    cplx* sn = (cplx*)malloc(sizeof(cplx) * m);
    cplx* cs = (cplx*)malloc(sizeof(cplx) * m);
    cplx* bet = (cplx*)malloc(sizeof(cplx) * (m + 1));
    cplx* h = (cplx*)malloc(sizeof(cplx) * (m + 1) * m);
    cplx* hk1 = (cplx*)malloc(sizeof(cplx) * (m + 1));
    double error = 0.0;
    int converged = 0;
    for (int64_t rs = 0; rs < max_restarts && !converged; ++rs) {
        restarts++;
        /* r = b - A x */
        mv(user, (orc_c128*)x, (orc_c128*)tmp);
        num_mv++;
        for (int64_t i = 0; i < n; ++i) r[i] = -tmp[i] + b[i];
        double alpha = anorm;
        double rn = vnorm(r, n);
        error = backward_error(rn, alpha, x, n, b_norm);
        for (int64_t i = 0; i < m; ++i) sn[i] = cs[i] = 0.0;
        for (int64_t i = 0; i < (m + 1) * m; ++i) h[i] = 0.0;
        for (int64_t i = 0; i < m + 1; ++i) bet[i] = 0.0;
        for (int64_t i = 0; i < n * (m + 1); ++i) q[i] = 0.0;
        for (int64_t i = 0; i < n; ++i) q[i] = r[i] / rn;
        bet[0] = rn;
        for (int64_t k = 0; k < m; ++k) {
            iters++;
            /* arnoldi (gmres.rs:158-176) */
            mv(user, (orc_c128*)(q + k * n), (orc_c128*)qk1);
            num_mv++;
            for (int64_t i = 0; i < m + 1; ++i) hk1[i] = 0.0;
            for (int64_t i = 0; i <= k; ++i) {
                const cplx* qi = q + i * n;
                cplx s = 0.0;
                if (ref_mgs) for (int64_t j = 0; j < n; ++j) s += conj(qk1[j]) * qi[j];
                else for (int64_t j = 0; j < n; ++j) s += conj(qi[j]) * qk1[j];
                hk1[i] = s;
                for (int64_t j = 0; j < n; ++j) qk1[j] += -s * qi[j];
            }
            hk1[k + 1] = vnorm(qk1, n);
            cplx scale = 1.0 / hk1[k + 1];
            for (int64_t j = 0; j < n; ++j) qk1[j] *= scale;
            /* apply_givens_rotation (gmres.rs:178-194) */
            for (int64_t i = 0; i < k; ++i) {
                cplx temp = cs[i] * hk1[i] + sn[i] * hk1[i + 1];
                hk1[i + 1] = -conj(sn[i]) * hk1[i] + cs[i] * hk1[i + 1];
                hk1[i] = temp;
            }
            givens(hk1[k], hk1[k + 1], &cs[k], &sn[k]);
            hk1[k] = cs[k] * hk1[k] + sn[k] * hk1[k + 1];
            hk1[k + 1] = 0.0;
            for (int64_t i = 0; i < k + 2; ++i) h[i + k * (m + 1)] = hk1[i];
            memcpy(q + (k + 1) * n, qk1, sizeof(cplx) * n);
            bet[k + 1] = -conj(sn[k]) * bet[k];
            bet[k] *= cs[k];
            error = backward_error(cabs_h(bet[k + 1]), alpha, x, n, b_norm);
            if (error < thresh || k == m - 1) {
                /* get_x (gmres.rs:146-156): LU of leading (k+1)x(k+1) of h */
                int64_t kk = k + 1;
                orc_c128* hkk = (orc_c128*)malloc(sizeof(orc_c128) * kk * kk);
                for (int64_t c = 0; c < kk; ++c)
                    for (int64_t rr = 0; rr < kk; ++rr) *C(&hkk[rr + c * kk]) = h[rr + c * (m + 1)];
                orc_c128* y = (orc_c128*)malloc(sizeof(orc_c128) * kk);
                for (int64_t i = 0; i < kk; ++i) *C(&y[i]) = bet[i];
                orc_lu_solve(kk, hkk, y);
                for (int64_t c = 0; c < kk; ++c) {
                    cplx yc = *C(&y[c]);
                    for (int64_t j = 0; j < n; ++j) x[j] += yc * q[j + c * n];
                }
                free(hkk);
                free(y);
                mv(user, (orc_c128*)x, (orc_c128*)tmp);
                num_mv++;
                for (int64_t i = 0; i < n; ++i) r[i] = -tmp[i] + b[i];
                error = backward_error(vnorm(r, n), alpha, x, n, b_norm);
                if (error < thresh) { converged = 1; break; }
            }
        }
    }
    if (stats) { stats[0] = num_mv; stats[1] = iters; stats[2] = restarts; }
    if (err_out) *err_out = error;
    free(b); free(r); free(tmp); free(q); free(qk1); free(sn); free(cs);
    free(bet); free(h); free(hk1);
    return converged ? 0 : 1;
}

int orc_gmres(int64_t n, const orc_c128* A, orc_c128* x, int64_t max_it, double thresh,
              int ref_mgs, int64_t* stats, double* err_out) {
    dense_op d = {CC(A), n};
    double anorm = vnorm(CC(A), n * n); /* Frobenius norm (gmres.rs:75-79) */
    return orc_gmres_op(n, dense_mv, &d, anorm, x, NULL, max_it, thresh, ref_mgs, stats, err_out);
}

/* ------------------------------------------------------------------ */
/* BC algebra + solve: solve.rs:24-71, 153-171                         */
/* ------------------------------------------------------------------ */
int orc_solve_surface_dense(const orc_physics* ph, const double* cpts, int64_t n,
                            const orc_c128* rhs_incc, int solver, int64_t max_it, double tol,
                            orc_c128* phic, orc_c128* vnc, int64_t* stats) {
    const cplx* rhs_inc = CC(rhs_incc);
    cplx* phi = C(phic);
    cplx* vn = C(vnc);
    cplx* H = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* G = (cplx*)malloc(sizeof(cplx) * n * n);
    orc_dense_surface_matrices(ph, cpts, n, (orc_c128*)H, (orc_c128*)G, 0, 0);
    cplx* lhs = NULL;
    cplx* x = NULL;
    cplx bcv = ph->bc_value[0] + I * ph->bc_value[1];
    double omega = 2.0 * PI * ph->frequency;
    for (int64_t i = 0; i < n; ++i) phi[i] = vn[i] = 0.0;
    if (ph->bc_type == ORC_BC_PRESSURE) {
        for (int64_t i = 0; i < n; ++i) { phi[i] = bcv; vn[i] = rhs_inc[i]; }
        /* vn = -H*phi + vn */
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) vn[i] += -phi[j] * H[i + j * n];
        lhs = G;
        x = vn;
    } else if (ph->bc_type == ORC_BC_NORMAL_VELOCITY) {
        for (int64_t i = 0; i < n; ++i) { vn[i] = bcv; phi[i] = -rhs_inc[i]; }
        /* phi = G*vn - rhs_inc */
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) phi[i] += vn[j] * G[i + j * n];
        lhs = H;
        x = phi;
    } else {
        cplx factor = (0.0 - I * (omega * ph->mass_density)) / bcv;
        for (int64_t p = 0; p < n * n; ++p) H[p] += factor * G[p];
        for (int64_t i = 0; i < n; ++i) phi[i] = -rhs_inc[i];
        lhs = H;
        x = phi;
    }
    int rc = 0;
    if (solver == ORC_SOLVER_DIRECT) {
        rc = orc_lu_solve(n, (orc_c128*)lhs, (orc_c128*)x);
    } else {
        rc = orc_gmres(n, (orc_c128*)lhs, (orc_c128*)x, max_it, tol, 1, stats, NULL);
    }
    if (ph->bc_type == ORC_BC_IMPEDANCE) {
        cplx factor = (0.0 - I * (omega * ph->mass_density)) / bcv;
        for (int64_t i = 0; i < n; ++i) vn[i] = factor * phi[i];
    }
    free(H);
    free(G);
    return rc;
}

/* solve.rs:173-185, postprocess.rs:70-88 */
int orc_field(const orc_physics* ph, const double* cpts, int64_t n, const double* fps,
              int64_t nfp, const orc_c128* phic, const orc_c128* vnc,
              const orc_c128* phi_inc_fpc, int want_total, int to_pressure, orc_c128* outc) {
    cplx* M = (cplx*)malloc(sizeof(cplx) * n * nfp);
    cplx* L = (cplx*)malloc(sizeof(cplx) * n * nfp);
    orc_dense_field_matrices(ph, cpts, n, fps, nfp, (orc_c128*)M, (orc_c128*)L);
    const cplx* phi = CC(phic);
    const cplx* vn = CC(vnc);
    cplx* out = C(outc);
    for (int64_t p = 0; p < nfp; ++p) out[p] = want_total ? CC(phi_inc_fpc)[p] : 0.0;
    /* phi_fp = -M phi + phi_fp ; phi_fp = L vn + phi_fp (sign as solve.rs:180-181) */
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = 0; p < nfp; ++p) out[p] += -phi[j] * M[p + j * nfp];
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = 0; p < nfp; ++p) out[p] += vn[j] * L[p + j * nfp];
    if (to_pressure) {
        cplx s = 0.0 + I * (2.0 * PI * ph->frequency * ph->mass_density);
        for (int64_t p = 0; p < nfp; ++p) out[p] *= s;
    }
    free(M);
    free(L);
    return 0;
}

/* ------------------------------------------------------------------ */
/* rand 0.9 StdRng = ChaCha12Rng, seed_from_u64 via PCG32 (rand_core)  */
/* ------------------------------------------------------------------ */
static inline uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define QR(a, b, c, d)                                                                    \
    a += b; d ^= a; d = rotl32(d, 16);                                                    \
    c += d; b ^= c; b = rotl32(b, 12);                                                    \
    a += b; d ^= a; d = rotl32(d, 8);                                                     \
    c += d; b ^= c; b = rotl32(b, 7);

static void chacha_block(const uint32_t key[8], uint64_t ctr, uint32_t out[16], int rounds) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                      (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int i = 0; i < rounds; i += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

static void rng_refill(orc_rng* r) {
    for (int b = 0; b < 4; ++b) chacha_block(r->key, r->counter + (uint64_t)b, r->buf + 16 * b, 12);
    r->counter += 4;
}

void orc_rng_seed(orc_rng* r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int c = 0; c < 8; ++c) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[c] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    r->counter = 0;
    r->idx = 64;
}

uint32_t orc_rng_u32(orc_rng* r) {
    if (r->idx >= 64) { rng_refill(r); r->idx = 0; }
    return r->buf[r->idx++];
}

uint64_t orc_rng_u64(orc_rng* r) {
    const int len = 64;
    int index = r->idx;
    if (index < len - 1) {
        r->idx += 2;
        return ((uint64_t)r->buf[index + 1] << 32) | r->buf[index];
    } else if (index >= len) {
        rng_refill(r);
        r->idx = 2;
        return ((uint64_t)r->buf[1] << 32) | r->buf[0];
    } else {
        uint64_t x = r->buf[len - 1];
        rng_refill(r);
        r->idx = 1;
        uint64_t y = r->buf[0];
        return (y << 32) | x;
    }
}

double orc_rng_f64(orc_rng* r) {
    uint64_t v = orc_rng_u64(r);
    return (double)(v >> 11) * (1.0 / (double)(1ull << 53));
}

uint64_t orc_rng_uniform_inclusive(orc_rng* r, uint64_t low, uint64_t high) {
    uint32_t range = (uint32_t)high - (uint32_t)low + 1u;
    if (range == 0) return orc_rng_u32(r);
    uint32_t thresh = (uint32_t)(-range) % range;
    for (;;) {
        uint64_t w = (uint64_t)orc_rng_u32(r) * (uint64_t)range;
        uint32_t hi = (uint32_t)(w >> 32), lo = (uint32_t)w;
        if (lo >= thresh) return low + hi;
    }
}

void orc_generate_random_ab(int64_t m, int64_t n, uint64_t seed, orc_c128* ac, orc_c128* bc) {
    cplx* a = C(ac);
    cplx* b = C(bc);
    for (int64_t p = 0; p < m * n; ++p) a[p] = 0.0;
    for (int64_t i = 0; i < (m < n ? m : n); ++i) a[i + i * m] = (double)n;
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (int64_t p = 0; p < m * n; ++p) {
        double re = orc_rng_f64(&r);
        double im = orc_rng_f64(&r);
        a[p] = (creal(a[p]) + re) + I * (cimag(a[p]) + im);
    }
    for (int64_t i = 0; i < n; ++i) {
        double re = orc_rng_f64(&r);
        double im = orc_rng_f64(&r);
        b[i] = re + I * im;
    }
}

/* ------------------------------------------------------------------ */
/* Cluster and block trees: cluster.rs:35-110, block.rs:38-96          */
/* ------------------------------------------------------------------ */
typedef struct cnode {
    double lo[3], hi[3], diam;
    int64_t* idx;
    int64_t cnt;
    struct cnode* son[2];
} cnode;

static cnode* cluster_new(const double* xyz, int64_t stride, int64_t* idx, int64_t cnt,
                          int64_t leaf) {
    cnode* c = (cnode*)calloc(1, sizeof(cnode));
    c->idx = idx;
    c->cnt = cnt;
    for (int d = 0; d < 3; ++d) { c->lo[d] = INFINITY; c->hi[d] = -INFINITY; }
    for (int64_t t = 0; t < cnt; ++t)
        for (int d = 0; d < 3; ++d) {
            double v = xyz[stride * idx[t] + d];
            c->lo[d] = fmin(v, c->lo[d]);
            c->hi[d] = fmax(v, c->hi[d]);
        }
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += (c->hi[d] - c->lo[d]) * (c->hi[d] - c->lo[d]);
    c->diam = sqrt(s);
    if (cnt <= leaf) return c;
    int jmax = 0;
    double diff = 0.0;
    for (int d = 0; d < 3; ++d) {
        double e = c->hi[d] - c->lo[d];
        if (e > diff) { jmax = d; diff = e; }
    }
    double gamma = 0.5 * (c->lo[jmax] + c->hi[jmax]);
    int64_t* a = (int64_t*)malloc(sizeof(int64_t) * cnt);
    int64_t* b = (int64_t*)malloc(sizeof(int64_t) * cnt);
    int64_t na = 0, nb = 0;
    for (int64_t t = 0; t < cnt; ++t) {
        if (xyz[stride * idx[t] + jmax] <= gamma) a[na++] = idx[t];
        else b[nb++] = idx[t];
    }
    c->son[0] = cluster_new(xyz, stride, a, na, leaf);
    c->son[1] = cluster_new(xyz, stride, b, nb, leaf);
    return c;
}

static void cluster_free(cnode* c) {
    if (!c) return;
    cluster_free(c->son[0]);
    cluster_free(c->son[1]);
    free(c->idx);
    free(c);
}

static double cluster_dist(const cnode* a, const cnode* b) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        double t1 = fmax(0.0, a->lo[d] - b->hi[d]);
        double t2 = fmax(0.0, b->lo[d] - a->hi[d]);
        s += t1 * t1;
        s += t2 * t2;
    }
    return sqrt(s);
}

typedef struct {
    int64_t n, cap;
    const cnode** r;
    const cnode** c;
    int32_t* adm;
} blist;

static void blist_push(blist* L, const cnode* r, const cnode* c, int adm) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->r = (const cnode**)realloc((void*)L->r, sizeof(void*) * L->cap);
        L->c = (const cnode**)realloc((void*)L->c, sizeof(void*) * L->cap);
        L->adm = (int32_t*)realloc(L->adm, sizeof(int32_t) * L->cap);
    }
    L->r[L->n] = r;
    L->c[L->n] = c;
    L->adm[L->n] = adm;
    L->n++;
}

static void block_recurse(const cnode* r, const cnode* c, double eta, blist* L) {
    int adm = fmin(r->diam, c->diam) <= eta * cluster_dist(r, c);
    int r_leaf = r->son[0] == NULL, c_leaf = c->son[0] == NULL;
    if (!(adm || r_leaf || c_leaf)) {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) block_recurse(r->son[a], c->son[b], eta, L);
    } else {
        blist_push(L, r, c, adm);
    }
}

static cnode* root_cluster(const double* coords, int64_t stride, int64_t n, int64_t leaf) {
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    return cluster_new(coords, stride, idx, n, leaf);
}

int orc_build_blocks(const double* coords, int64_t stride, int64_t n, int64_t leaf, double eta,
                     orc_blocks* out) {
    cnode* root = root_cluster(coords, stride, n, leaf);
    blist L = {0, 0, NULL, NULL, NULL};
    block_recurse(root, root, eta, &L);
    out->nblocks = L.n;
    out->row_off = (int64_t*)malloc(sizeof(int64_t) * (L.n + 1));
    out->col_off = (int64_t*)malloc(sizeof(int64_t) * (L.n + 1));
    out->admissible = (int32_t*)malloc(sizeof(int32_t) * (L.n ? L.n : 1));
    int64_t tr = 0, tc = 0;
    for (int64_t b = 0; b < L.n; ++b) { tr += L.r[b]->cnt; tc += L.c[b]->cnt; }
    out->rows = (int64_t*)malloc(sizeof(int64_t) * (tr ? tr : 1));
    out->cols = (int64_t*)malloc(sizeof(int64_t) * (tc ? tc : 1));
    out->row_off[0] = out->col_off[0] = 0;
    for (int64_t b = 0; b < L.n; ++b) {
        memcpy(out->rows + out->row_off[b], L.r[b]->idx, sizeof(int64_t) * L.r[b]->cnt);
        memcpy(out->cols + out->col_off[b], L.c[b]->idx, sizeof(int64_t) * L.c[b]->cnt);
        out->row_off[b + 1] = out->row_off[b] + L.r[b]->cnt;
        out->col_off[b + 1] = out->col_off[b] + L.c[b]->cnt;
        out->admissible[b] = L.adm[b];
    }
    out->root_diameter = root->diam;
    free((void*)L.r);
    free((void*)L.c);
    free(L.adm);
    cluster_free(root);
    return 0;
}

void orc_blocks_free(orc_blocks* b) {
    free(b->row_off); free(b->col_off); free(b->rows); free(b->cols); free(b->admissible);
    memset(b, 0, sizeof(*b));
}

int orc_block_tree_admissible_at(const double* coords, int64_t stride, int64_t n, int64_t leaf,
                                 double eta, const int32_t* path, int depth) {
    cnode* root = root_cluster(coords, stride, n, leaf);
    const cnode* r = root;
    const cnode* c = root;
    int result = -1;
    for (int d = 0;; ++d) {
        int adm = fmin(r->diam, c->diam) <= eta * cluster_dist(r, c);
        if (d == depth) { result = adm; break; }
        int divisible = !(adm || r->son[0] == NULL || c->son[0] == NULL);
        if (!divisible) break;
        int ch = path[d];
        r = r->son[ch / 2];
        c = c->son[ch % 2];
    }
    cluster_free(root);
    return result;
}

/* ------------------------------------------------------------------ */
/* Complex Householder QR (thin) and one-sided Jacobi SVD              */
/* ------------------------------------------------------------------ */
/* A (m x r, column-major, m >= r) -> Q (m x r), R (r x r) */
static void qr_thin(int64_t m, int64_t r, const cplx* A, cplx* Q, cplx* R) {
    cplx* W = (cplx*)malloc(sizeof(cplx) * m * r);
    memcpy(W, A, sizeof(cplx) * m * r);
    cplx* vs = (cplx*)calloc((size_t)m * r, sizeof(cplx));
    cplx* taus = (cplx*)calloc(r, sizeof(cplx));
    for (int64_t k = 0; k < r; ++k) {
        double nrm = 0.0;
        for (int64_t i = k; i < m; ++i) nrm += cnorm_sqr(W[i + k * m]);
        nrm = sqrt(nrm);
        cplx* v = vs + k * m;
        if (nrm == 0.0) { taus[k] = 0.0; continue; }
        cplx x0 = W[k + k * m];
        cplx ph = (x0 == 0.0) ? 1.0 : x0 / cabs_h(x0);
        cplx alpha = -ph * nrm;
        for (int64_t i = k; i < m; ++i) v[i] = W[i + k * m];
        v[k] -= alpha;
        double vn2 = 0.0;
        for (int64_t i = k; i < m; ++i) vn2 += cnorm_sqr(v[i]);
        if (vn2 == 0.0) { taus[k] = 0.0; continue; }
        taus[k] = 2.0 / vn2;
        for (int64_t c = k; c < r; ++c) {
            cplx s = 0.0;
            for (int64_t i = k; i < m; ++i) s += conj(v[i]) * W[i + c * m];
            s *= taus[k];
            for (int64_t i = k; i < m; ++i) W[i + c * m] -= s * v[i];
        }
    }
    for (int64_t c = 0; c < r; ++c)
        for (int64_t i = 0; i < r; ++i) R[i + c * r] = i <= c ? W[i + c * m] : 0.0;
    /* Q = H_0 ... H_{r-1} applied to the first r columns of I */
    for (int64_t c = 0; c < r; ++c)
        for (int64_t i = 0; i < m; ++i) Q[i + c * m] = (i == c) ? 1.0 : 0.0;
    for (int64_t k = r - 1; k >= 0; --k) {
        if (taus[k] == 0.0) continue;
        cplx* v = vs + k * m;
        for (int64_t c = 0; c < r; ++c) {
            cplx s = 0.0;
            for (int64_t i = k; i < m; ++i) s += conj(v[i]) * Q[i + c * m];
            s *= taus[k];
            for (int64_t i = k; i < m; ++i) Q[i + c * m] -= s * v[i];
        }
    }
    free(W);
    free(vs);
    free(taus);
}

/* One-sided Jacobi SVD of a square r x r matrix: A = U diag(s) V^H, s sorted
 * descending (nalgebra's svd() sorts, Matrix::svd in nalgebra >= 0.31). */
static void svd_jacobi(int64_t r, const cplx* A, cplx* U, double* s, cplx* V) {
    cplx* W = (cplx*)malloc(sizeof(cplx) * r * r);
    memcpy(W, A, sizeof(cplx) * r * r);
    for (int64_t c = 0; c < r; ++c)
        for (int64_t i = 0; i < r; ++i) V[i + c * r] = (i == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int64_t p = 0; p < r - 1; ++p)
            for (int64_t q = p + 1; q < r; ++q) {
                double app = 0.0, aqq = 0.0;
                cplx apq = 0.0;
                for (int64_t i = 0; i < r; ++i) {
                    app += cnorm_sqr(W[i + p * r]);
                    aqq += cnorm_sqr(W[i + q * r]);
                    apq += conj(W[i + p * r]) * W[i + q * r];
                }
                double mag = cabs_h(apq);
                if (mag == 0.0 || mag <= 1e-15 * sqrt(app * aqq)) continue;
                off = fmax(off, mag / sqrt(app * aqq));
                cplx ph = apq / mag;
                double tau = (aqq - app) / (2.0 * mag);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                for (int64_t i = 0; i < r; ++i) {
                    cplx wp = W[i + p * r], wq = W[i + q * r];
                    W[i + p * r] = cs * wp - sn * conj(ph) * wq;
                    W[i + q * r] = sn * ph * wp + cs * wq;
                    cplx vp = V[i + p * r], vq = V[i + q * r];
                    V[i + p * r] = cs * vp - sn * conj(ph) * vq;
                    V[i + q * r] = sn * ph * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * r);
    double* sv = (double*)malloc(sizeof(double) * r);
    for (int64_t c = 0; c < r; ++c) {
        double nn = 0.0;
        for (int64_t i = 0; i < r; ++i) nn += cnorm_sqr(W[i + c * r]);
        sv[c] = sqrt(nn);
        ord[c] = c;
    }
    for (int64_t a = 1; a < r; ++a) { /* stable insertion sort, descending */
        int64_t t = ord[a];
        int64_t b = a - 1;
        while (b >= 0 && sv[ord[b]] < sv[t]) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = t;
    }
    cplx* V2 = (cplx*)malloc(sizeof(cplx) * r * r);
    for (int64_t c = 0; c < r; ++c) {
        int64_t o = ord[c];
        s[c] = sv[o];
        for (int64_t i = 0; i < r; ++i) {
            U[i + c * r] = sv[o] > 0 ? W[i + o * r] / sv[o] : ((i == c) ? 1.0 : 0.0);
            V2[i + c * r] = V[i + o * r];
        }
    }
    memcpy(V, V2, sizeof(cplx) * r * r);
    free(V2);
    free(ord);
    free(sv);
    free(W);
}

/* ------------------------------------------------------------------ */
/* ACA: aca.rs:57-268                                                  */
/* ------------------------------------------------------------------ */
static cplx dotu(const cplx* a, const cplx* b, int64_t n) {
    cplx s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

typedef struct {
    int64_t m, n, cnt, cap;
    cplx* u; /* m x cap */
    cplx* v; /* n x cap */
} uvlist;

static void uv_push(uvlist* L, const cplx* u, const cplx* v) {
    if (L->cnt == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 8;
        L->u = (cplx*)realloc(L->u, sizeof(cplx) * L->m * L->cap);
        L->v = (cplx*)realloc(L->v, sizeof(cplx) * L->n * L->cap);
    }
    memcpy(L->u + L->cnt * L->m, u, sizeof(cplx) * L->m);
    memcpy(L->v + L->cnt * L->n, v, sizeof(cplx) * L->n);
    L->cnt++;
}

static void residual_row(const uvlist* L, orc_rowcol_fn f, void* user, int64_t i, cplx* r) {
    f(user, 1, i, (orc_c128*)r);
    for (int64_t k = 0; k < L->cnt; ++k) {
        cplx a = -L->u[k * L->m + i];
        const cplx* v = L->v + k * L->n;
        for (int64_t j = 0; j < L->n; ++j) r[j] += a * v[j];
    }
}

static void residual_col(const uvlist* L, orc_rowcol_fn f, void* user, int64_t j, cplx* r) {
    f(user, 0, j, (orc_c128*)r);
    for (int64_t k = 0; k < L->cnt; ++k) {
        cplx a = -L->v[k * L->n + j];
        const cplx* u = L->u + k * L->m;
        for (int64_t i = 0; i < L->m; ++i) r[i] += a * u[i];
    }
}

static int64_t index_max_exclude(const cplx* a, int64_t len, const char* excl) {
    int64_t index = len;
    double best = -1.0;
    for (int64_t i = 0; i < len; ++i) {
        if (excl[i]) continue;
        double v = cabs_h(a[i]);
        if (best < v) { index = i; best = v; }
    }
    return index;
}

static void random_index_exclude(int64_t* i, int64_t n, const char* excl, orc_rng* rng) {
    if (excl[*i]) {
        *i = (int64_t)orc_rng_uniform_inclusive(rng, 0, (uint64_t)(n - 1));
        for (int64_t t = 0; t < n - 1; ++t) {
            *i = (*i + 1) % n;
            if (!excl[*i]) break;
        }
    }
}

static double norm_estimate_update(const uvlist* L, double* total) {
    int64_t last = L->cnt - 1;
    const cplx* u = L->u + last * L->m;
    const cplx* v = L->v + last * L->n;
    double nu = vnorm(u, L->m), nv = vnorm(v, L->n);
    double k2 = (nu * nu) * (nv * nv);
    *total += k2;
    for (int64_t i = 0; i < last; ++i)
        *total += 2.0 * cabs_h(dotu(L->u + i * L->m, u, L->m)) *
                  cabs_h(dotu(L->v + i * L->n, v, L->n));
    return k2;
}

int orc_aca_build(double tol, int64_t m, int64_t n, orc_rowcol_fn f, void* user, orc_aca* out) {
    uvlist L = {m, n, 0, 0, NULL, NULL};
    int64_t nr = m < n ? m : n;
    orc_rng rng;
    orc_rng_seed(&rng, 10);
    int64_t iref = (int64_t)orc_rng_uniform_inclusive(&rng, 0, (uint64_t)(nr - 1));
    int64_t jref = (int64_t)orc_rng_uniform_inclusive(&rng, 0, (uint64_t)(nr - 1));
    /* exclusion sets sized to the larger dimension: the random reference
     * indices are drawn from [0, min(m,n)-1] (aca.rs:63-65) */
    int64_t mx = m > n ? m : n;
    char* istar_set = (char*)calloc(mx, 1);
    char* jstar_set = (char*)calloc(mx, 1);
    cplx* riref = (cplx*)malloc(sizeof(cplx) * n);
    cplx* rjref = (cplx*)malloc(sizeof(cplx) * m);
    cplx* ristar = (cplx*)malloc(sizeof(cplx) * n);
    cplx* rjstar = (cplx*)malloc(sizeof(cplx) * m);
    residual_row(&L, f, user, iref, riref);
    residual_col(&L, f, user, jref, rjref);
    double total = 0.0;
    for (int64_t k = 0; k < nr; ++k) {
        int64_t jstar = index_max_exclude(riref, n, jstar_set);
        double jv = cabs_h(riref[jstar]);
        int64_t istar = index_max_exclude(rjref, m, istar_set);
        double iv = cabs_h(rjref[istar]);
        if (iv > jv) {
            residual_row(&L, f, user, istar, ristar);
            jstar = index_max_exclude(ristar, n, jstar_set);
            residual_col(&L, f, user, jstar, rjstar);
        } else {
            residual_col(&L, f, user, jstar, rjstar);
            istar = index_max_exclude(rjstar, m, istar_set);
            residual_row(&L, f, user, istar, ristar);
        }
        istar_set[istar] = 1;
        jstar_set[jstar] = 1;
        cplx factor = 1.0 / ristar[jstar];
        for (int64_t j = 0; j < n; ++j) ristar[j] *= factor;
        uv_push(&L, rjstar, ristar);
        double k2 = norm_estimate_update(&L, &total);
        double step = sqrt(k2) / sqrt(total);
        if (step < tol / 50.0) break;
        if (iref == istar) {
            random_index_exclude(&iref, nr, istar_set, &rng);
            residual_row(&L, f, user, iref, riref);
        } else {
            cplx a = -L.u[k * m + iref];
            for (int64_t j = 0; j < n; ++j) riref[j] += a * L.v[k * n + j];
        }
        if (jref == jstar) {
            random_index_exclude(&jref, nr, jstar_set, &rng);
            residual_col(&L, f, user, jref, rjref);
        } else {
            cplx a = -L.v[k * n + jref];
            for (int64_t i = 0; i < m; ++i) rjref[i] += a * L.u[k * m + i];
        }
    }
    out->m = m;
    out->n = n;
    out->rank_before_recompression = L.cnt;
    /* svd_recompression (aca.rs:137-179) */
    int64_t r = L.cnt;
    cplx* Qu = (cplx*)malloc(sizeof(cplx) * m * r);
    cplx* Ru = (cplx*)malloc(sizeof(cplx) * r * r);
    cplx* Qv = (cplx*)malloc(sizeof(cplx) * n * r);
    cplx* Rv = (cplx*)malloc(sizeof(cplx) * r * r);
    qr_thin(m, r, L.u, Qu, Ru);
    qr_thin(n, r, L.v, Qv, Rv);
    cplx* RR = (cplx*)calloc((size_t)r * r, sizeof(cplx));
    for (int64_t c = 0; c < r; ++c)
        for (int64_t i = 0; i < r; ++i) {
            cplx s = 0.0;
            for (int64_t t = 0; t < r; ++t) s += Ru[i + t * r] * Rv[c + t * r]; /* Rv^T */
            RR[i + c * r] = s;
        }
    cplx* W = (cplx*)malloc(sizeof(cplx) * r * r);
    cplx* Vs = (cplx*)malloc(sizeof(cplx) * r * r);
    double* sig = (double*)malloc(sizeof(double) * r);
    svd_jacobi(r, RR, W, sig, Vs);
    int64_t index = 0;
    double sum = 0.0;
    for (int64_t t = 0; t < r; ++t) {
        sum += sig[t];
        if (sig[t] < tol * sum) break;
        index++;
    }
    index = index + 1 < r ? index + 1 : r;
    if (index < 1) index = 1;
    out->rank = index;
    out->u = (orc_c128*)malloc(sizeof(orc_c128) * m * index);
    out->v = (orc_c128*)malloc(sizeof(orc_c128) * n * index);
    cplx* U = C(out->u);
    cplx* V = C(out->v);
    for (int64_t t = 0; t < index; ++t) {
        for (int64_t i = 0; i < m; ++i) {
            cplx s = 0.0;
            for (int64_t c = 0; c < r; ++c) s += Qu[i + c * m] * W[c + t * r];
            U[i + t * m] = sig[t] * s;
        }
        for (int64_t j = 0; j < n; ++j) {
            cplx s = 0.0;
            for (int64_t c = 0; c < r; ++c) s += Qv[j + c * n] * conj(Vs[c + t * r]);
            V[j + t * n] = s;
        }
    }
    /* calculate_norm_estimate (aca.rs:212-222) */
    double tot = 0.0;
    for (int64_t i = 0; i < index; ++i) {
        double nu = vnorm(U + i * m, m), nv = vnorm(V + i * n, n);
        tot += (nu * nu) * (nv * nv);
        for (int64_t j = 0; j < i; ++j)
            tot += 2.0 * cabs_h(dotu(U + i * m, U + j * m, m)) *
                   cabs_h(dotu(V + i * n, V + j * n, n));
    }
    out->norm = sqrt(tot);
    free(Qu); free(Ru); free(Qv); free(Rv); free(RR); free(W); free(Vs); free(sig);
    free(istar_set); free(jstar_set); free(riref); free(rjref); free(ristar); free(rjstar);
    free(L.u); free(L.v);
    return 0;
}

void orc_aca_free(orc_aca* a) {
    free(a->u);
    free(a->v);
    memset(a, 0, sizeof(*a));
}

void orc_aca_gemv(const orc_aca* a, orc_c128 alpha, const orc_c128* xc, orc_c128 beta,
                  orc_c128* bc) {
    const cplx* x = CC(xc);
    cplx* b = C(bc);
    const cplx* U = CC(a->u);
    const cplx* V = CC(a->v);
    for (int64_t k = 0; k < a->rank; ++k) {
        cplx d = *C(&alpha) * dotu(V + k * a->n, x, a->n);
        cplx bb = k == 0 ? *C(&beta) : 1.0;
        for (int64_t i = 0; i < a->m; ++i) b[i] = d * U[i + k * a->m] + bb * b[i];
    }
}

/* ------------------------------------------------------------------ */
/* HMatrix: h_matrix.rs:127-247                                        */
/* ------------------------------------------------------------------ */
struct orc_hmatrix {
    int64_t n;
    double norm;
    orc_blocks blocks;
    orc_aca* aca;      /* per block (admissible) */
    cplx** dense;      /* per block (inadmissible), rows x cols col-major */
};

typedef struct {
    orc_rc_fn f;
    void* user;
    const int64_t* rows;
    int64_t nr;
    const int64_t* cols;
    int64_t nc;
} blk_ctx;

static void blk_rowcol(void* user, int is_row, int64_t idx, orc_c128* out) {
    blk_ctx* b = (blk_ctx*)user;
    if (is_row) b->f(b->user, b->rows + idx, 1, b->cols, b->nc, out);
    else b->f(b->user, b->rows, b->nr, b->cols + idx, 1, out);
}

orc_hmatrix* orc_hmatrix_build(int64_t n, orc_rc_fn f, void* user, const double* coords,
                               int64_t stride, int64_t leaf, double tol, double eta) {
    orc_hmatrix* h = (orc_hmatrix*)calloc(1, sizeof(orc_hmatrix));
    h->n = n;
    orc_build_blocks(coords, stride, n, leaf, eta, &h->blocks);
    int64_t nb = h->blocks.nblocks;
    h->aca = (orc_aca*)calloc(nb, sizeof(orc_aca));
    h->dense = (cplx**)calloc(nb, sizeof(cplx*));
    double norm_iad = 0.0, norm_ad = 0.0;
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t* rows = h->blocks.rows + h->blocks.row_off[b];
        const int64_t* cols = h->blocks.cols + h->blocks.col_off[b];
        int64_t nr = h->blocks.row_off[b + 1] - h->blocks.row_off[b];
        int64_t nc = h->blocks.col_off[b + 1] - h->blocks.col_off[b];
        if (h->blocks.admissible[b]) {
            blk_ctx bc = {f, user, rows, nr, cols, nc};
            orc_aca_build(tol, nr, nc, blk_rowcol, &bc, &h->aca[b]);
        } else {
            cplx* D = (cplx*)malloc(sizeof(cplx) * nr * nc);
            cplx* row = (cplx*)malloc(sizeof(cplx) * nc);
            for (int64_t i = 0; i < nr; ++i) {
                f(user, rows + i, 1, cols, nc, (orc_c128*)row);
                for (int64_t j = 0; j < nc; ++j) D[i + j * nr] = row[j];
            }
            free(row);
            h->dense[b] = D;
        }
    }
    /* update_norm: inadmissible Frobenius norms first, then ACA estimates */
    for (int64_t b = 0; b < nb; ++b)
        if (!h->blocks.admissible[b]) {
            int64_t nr = h->blocks.row_off[b + 1] - h->blocks.row_off[b];
            int64_t nc = h->blocks.col_off[b + 1] - h->blocks.col_off[b];
            norm_iad += vnorm(h->dense[b], nr * nc);
        }
    for (int64_t b = 0; b < nb; ++b)
        if (h->blocks.admissible[b]) norm_ad += h->aca[b].norm;
    h->norm = norm_iad + norm_ad;
    return h;
}

void orc_hmatrix_free(orc_hmatrix* h) {
    if (!h) return;
    for (int64_t b = 0; b < h->blocks.nblocks; ++b) {
        if (h->blocks.admissible[b]) orc_aca_free(&h->aca[b]);
        free(h->dense[b]);
    }
    free(h->aca);
    free(h->dense);
    orc_blocks_free(&h->blocks);
    free(h);
}

double orc_hmatrix_norm(const orc_hmatrix* h) { return h->norm; }

void orc_hmatrix_gemv(const orc_hmatrix* h, orc_c128 alphac, const orc_c128* xc,
                      orc_c128 betac, orc_c128* bc) {
    cplx alpha = *C(&alphac), beta = *C(&betac);
    const cplx* x = CC(xc);
    cplx* b = C(bc);
    if (beta != 1.0)
        for (int64_t i = 0; i < h->n; ++i) b[i] *= beta;
    int64_t nb = h->blocks.nblocks;
    cplx* x1 = (cplx*)malloc(sizeof(cplx) * (h->n + 1));
    cplx* b1 = (cplx*)malloc(sizeof(cplx) * (h->n + 1));
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t blk = 0; blk < nb; ++blk) {
            int adm = h->blocks.admissible[blk];
            if ((pass == 0) == (adm != 0)) continue; /* inadmissible first */
            const int64_t* rows = h->blocks.rows + h->blocks.row_off[blk];
            const int64_t* cols = h->blocks.cols + h->blocks.col_off[blk];
            int64_t nr = h->blocks.row_off[blk + 1] - h->blocks.row_off[blk];
            int64_t nc = h->blocks.col_off[blk + 1] - h->blocks.col_off[blk];
            for (int64_t j = 0; j < nc; ++j) x1[j] = x[cols[j]];
            for (int64_t i = 0; i < nr; ++i) b1[i] = 0.0;
            if (adm) {
                orc_aca_gemv(&h->aca[blk], *(orc_c128*)&alpha, (orc_c128*)x1, (orc_c128){0, 0},
                             (orc_c128*)b1);
            } else {
                const cplx* D = h->dense[blk];
                for (int64_t j = 0; j < nc; ++j) {
                    cplx s = alpha * x1[j];
                    for (int64_t i = 0; i < nr; ++i) b1[i] += s * D[i + j * nr];
                }
            }
            for (int64_t i = 0; i < nr; ++i) b[rows[i]] += b1[i];
        }
    free(x1);
    free(b1);
}

void orc_hmatrix_stats(const orc_hmatrix* h, int64_t* st) {
    st[0] = st[1] = st[2] = st[3] = 0;
    for (int64_t b = 0; b < h->blocks.nblocks; ++b) {
        int64_t nr = h->blocks.row_off[b + 1] - h->blocks.row_off[b];
        int64_t nc = h->blocks.col_off[b + 1] - h->blocks.col_off[b];
        if (h->blocks.admissible[b]) {
            st[0]++;
            st[2] += h->aca[b].rank * (nr + nc);
            st[3] += h->aca[b].rank;
        } else {
            st[1]++;
            st[2] += nr * nc;
        }
    }
}

void orc_hmatrix_to_full(const orc_hmatrix* h, orc_c128* Ac) {
    cplx* A = C(Ac);
    int64_t n = h->n;
    for (int64_t p = 0; p < n * n; ++p) A[p] = 0.0;
    for (int64_t b = 0; b < h->blocks.nblocks; ++b) {
        const int64_t* rows = h->blocks.rows + h->blocks.row_off[b];
        const int64_t* cols = h->blocks.cols + h->blocks.col_off[b];
        int64_t nr = h->blocks.row_off[b + 1] - h->blocks.row_off[b];
        int64_t nc = h->blocks.col_off[b + 1] - h->blocks.col_off[b];
        if (h->blocks.admissible[b]) {
            const orc_aca* a = &h->aca[b];
            for (int64_t k = 0; k < a->rank; ++k)
                for (int64_t i = 0; i < nr; ++i)
                    for (int64_t j = 0; j < nc; ++j)
                        A[rows[i] + cols[j] * n] += CC(a->u)[i + k * nr] * CC(a->v)[j + k * nc];
        } else {
            for (int64_t i = 0; i < nr; ++i)
                for (int64_t j = 0; j < nc; ++j) A[rows[i] + cols[j] * n] += h->dense[b][i + j * nr];
        }
    }
}

typedef struct {
    const orc_physics* ph;
    const double* cpts;
    int side;
} bem_rc;

static void bem_rc_fn(void* user, const int64_t* rows, int64_t nr, const int64_t* cols,
                      int64_t nc, orc_c128* out) {
    bem_rc* b = (bem_rc*)user;
    orc_surface_row_or_column(b->ph, b->cpts, b->side, rows, nr, cols, nc, out);
}

static void hmat_mv(void* user, const orc_c128* x, orc_c128* y) {
    orc_hmatrix* h = (orc_hmatrix*)user;
    orc_hmatrix_gemv(h, (orc_c128){1, 0}, x, (orc_c128){0, 0}, y);
}

int orc_solve_surface_hmatrix(const orc_physics* ph, const double* cpts, int64_t n,
                              const orc_c128* rhs_incc, int64_t max_it, double tol,
                              int64_t leaf, double aca_tol, double eta, orc_c128* phic,
                              orc_c128* vnc, int64_t* stats) {
    cplx* phi = C(phic);
    cplx* vn = C(vnc);
    cplx* rhs = (cplx*)malloc(sizeof(cplx) * n);
    memcpy(rhs, rhs_incc, sizeof(cplx) * n);
    cplx bcv = ph->bc_value[0] + I * ph->bc_value[1];
    for (int64_t i = 0; i < n; ++i) phi[i] = vn[i] = 0.0;
    orc_hmatrix* hr = NULL;
    if (!(ph->bc_value[0] == 0.0 && ph->bc_value[1] == 0.0)) {
        bem_rc rc = {ph, cpts, 1};
        hr = orc_hmatrix_build(n, bem_rc_fn, &rc, cpts, ORC_CPT_STRIDE, leaf, aca_tol, eta);
    }
    if (ph->bc_type == ORC_BC_PRESSURE) {
        for (int64_t i = 0; i < n; ++i) phi[i] = bcv;
        if (hr) orc_hmatrix_gemv(hr, (orc_c128){1, 0}, phic, (orc_c128){1, 0}, (orc_c128*)rhs);
    } else if (ph->bc_type == ORC_BC_NORMAL_VELOCITY) {
        for (int64_t i = 0; i < n; ++i) vn[i] = bcv;
        if (hr) orc_hmatrix_gemv(hr, (orc_c128){1, 0}, vnc, (orc_c128){-1, 0}, (orc_c128*)rhs);
        else for (int64_t i = 0; i < n; ++i) rhs[i] *= -1.0;
    } else {
        for (int64_t i = 0; i < n; ++i) rhs[i] = -CC(rhs_incc)[i];
    }
    orc_hmatrix_free(hr);
    bem_rc rcl = {ph, cpts, 0};
    orc_hmatrix* hl = orc_hmatrix_build(n, bem_rc_fn, &rcl, cpts, ORC_CPT_STRIDE, leaf, aca_tol, eta);
    int rc = orc_gmres_op(n, hmat_mv, hl, hl->norm, (orc_c128*)rhs, NULL, max_it, tol, 1, stats, NULL);
    if (stats) orc_hmatrix_stats(hl, stats + 3);
    orc_hmatrix_free(hl);
    if (ph->bc_type == ORC_BC_PRESSURE) {
        memcpy(vn, rhs, sizeof(cplx) * n);
    } else if (ph->bc_type == ORC_BC_NORMAL_VELOCITY) {
        memcpy(phi, rhs, sizeof(cplx) * n);
    } else {
        memcpy(phi, rhs, sizeof(cplx) * n);
        double omega = 2.0 * PI * ph->frequency;
        cplx factor = (0.0 - I * (omega * ph->mass_density)) / bcv;
        for (int64_t i = 0; i < n; ++i) vn[i] = factor * phi[i];
    }
    free(rhs);
    return rc;
}
