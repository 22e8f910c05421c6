/*
 * efie_oracle.c -- CPU FP64 restatement of the reference EFIE path
 * (mahmut26/Matlab-cuda-efief-vs-cpu).
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.
 * It is never part of the product library.  Each function cites the
 * reference file:line it restates; loop and summation orders follow the
 * MATLAB code so that results agree with it to rounding.
 *
 * Parity pins (see tests/test_oracle_*.py): closed-form mesh/RWG counts and
 * brute-force tables on tiny grids, exact monomial integrals for every
 * quadrature rule, an independent numpy restatement (oracle/mirror_np.py)
 * on small meshes, a high-order tensor rule for well-separated pairs, the
 * static (k -> 0) limit, symmetry Z = Z^T, and LAPACK zgetrf/zgetrs via
 * scipy for the solve.
 */
#include "efie_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.141592653589793238462643383279502884

typedef struct { double re, im; } cx;

static inline cx cmk(double r, double i) { cx z = {r, i}; return z; }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) {
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cscale(cx a, double s) { return cmk(a.re * s, a.im * s); }
static inline cx cdiv(cx a, cx b) {
    double d = b.re * b.re + b.im * b.im;
    return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
/* principal square root */
static cx csqrt_p(cx z) {
    double r = hypot(z.re, z.im);
    if (r == 0.0) return cmk(0.0, 0.0);
    if (z.re >= 0.0) {
        double t = sqrt(0.5 * (r + z.re));
        return cmk(t, z.im / (2.0 * t));
    } else {
        double t = sqrt(0.5 * (r - z.re));
        return cmk(fabs(z.im) / (2.0 * t), z.im >= 0.0 ? t : -t);
    }
}

static inline double norm3(const double* a) {
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

/* exp(-1j*k*R)/(4*pi*R) as the MATLAB expression evaluates it
 * (rwg_pair_integrals_regular.m:22): -1j*k = (ki, -kr), times R, exp. */
static inline cx green(double kr, double ki, double R) {
    double x = ki * R, y = -kr * R;
    double e = exp(x);
    cx g = cmk(e * cos(y), e * sin(y));
    double d = 4.0 * PI_D * R;
    return cmk(g.re / d, g.im / d);
}

/* ---------------------------------------------------------------- mesh -- */

void orc_mesh_counts(int npx, int npy, int ngx, int ngy, int* Nv, int* Nt) {
    *Nv = (npx + 1) * (npy + 1) + (ngx + 1) * (ngy + 1);
    *Nt = 2 * (npx * npy + ngx * ngy);
}

/* MATLAB linspace(d1,d2,n): d1 + (0:n1).*(d2-d1)./n1 with y(end)=d2 */
static void linspace(double d1, double d2, int n, double* y) {
    int n1 = n - 1;
    for (int i = 0; i <= n1; ++i) y[i] = d1 + ((double)i * (d2 - d1)) / (double)n1;
    y[n1] = d2;
}

/* make_patch_with_ground_mesh.m:1-46.  Vertices come from
 * meshgrid(x,y)(:), i.e. y runs fastest (:5-7, :23-25); faces use the
 * (Np_x+1) row stride of :11-20 / :30-39 exactly as written. */
void orc_make_mesh(double Lp, double Wp, double Lg, double Wg, double h,
                   int npx, int npy, int ngx, int ngy, double* V, int* F) {
    int Nv, Nt;
    orc_mesh_counts(npx, npy, ngx, ngy, &Nv, &Nt);
    double* xs = (double*)malloc(sizeof(double) * (size_t)(npx + ngx + 2));
    double* ys = (double*)malloc(sizeof(double) * (size_t)(npy + ngy + 2));
    int v = 0;
    /* patch layer */
    linspace(-Wp / 2, Wp / 2, npx + 1, xs);
    linspace(-Lp / 2, Lp / 2, npy + 1, ys);
    for (int c = 0; c <= npx; ++c)
        for (int r = 0; r <= npy; ++r) {
            V[v + 0 * Nv] = xs[c];
            V[v + 1 * Nv] = ys[r];
            V[v + 2 * Nv] = (+h / 2) * 1.0;
            ++v;
        }
    int nvp = v;
    linspace(-Wg / 2, Wg / 2, ngx + 1, xs);
    linspace(-Lg / 2, Lg / 2, ngy + 1, ys);
    for (int c = 0; c <= ngx; ++c)
        for (int r = 0; r <= ngy; ++r) {
            V[v + 0 * Nv] = xs[c];
            V[v + 1 * Nv] = ys[r];
            V[v + 2 * Nv] = (-h / 2) * 1.0;
            ++v;
        }
    int t = 0;
    for (int j = 1; j <= npy; ++j)
        for (int i = 1; i <= npx; ++i) {
            int v1 = (j - 1) * (npx + 1) + i, v2 = v1 + 1;
            int v3 = v1 + (npx + 1), v4 = v3 + 1;
            F[t + 0 * Nt] = v1; F[t + 1 * Nt] = v2; F[t + 2 * Nt] = v4; ++t;
            F[t + 0 * Nt] = v1; F[t + 1 * Nt] = v4; F[t + 2 * Nt] = v3; ++t;
        }
    for (int j = 1; j <= ngy; ++j)
        for (int i = 1; i <= ngx; ++i) {
            int v1 = nvp + (j - 1) * (ngx + 1) + i, v2 = v1 + 1;
            int v3 = v1 + (ngx + 1), v4 = v3 + 1;
            F[t + 0 * Nt] = v1; F[t + 1 * Nt] = v2; F[t + 2 * Nt] = v4; ++t;
            F[t + 0 * Nt] = v1; F[t + 1 * Nt] = v4; F[t + 2 * Nt] = v3; ++t;
        }
    free(xs);
    free(ys);
}

/* ---------------------------------------------------------------- RWG -- */

typedef struct { int a, b, raw; } rawedge;

static int cmp_raw(const void* x, const void* y) {
    const rawedge* p = (const rawedge*)x;
    const rawedge* q = (const rawedge*)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return p->raw < q->raw ? -1 : (p->raw > q->raw);
}

/* rawKey/rawDir of build_rwg_interior.m:5-23, then unique(...,'rows') (:25):
 * a lexicographic sort; ties keep raw order, which is the order of
 * find(ic==eOld) (:47). */
static rawedge* sorted_raw_edges(const int* F, int Nt) {
    rawedge* R = (rawedge*)malloc(sizeof(rawedge) * (size_t)Nt * 3);
    for (int t = 0; t < Nt; ++t) {
        int tri[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
        for (int le = 0; le < 3; ++le) {
            int a = tri[le], b = tri[(le + 1) % 3];
            rawedge e = {a < b ? a : b, a < b ? b : a, 3 * t + le};
            R[3 * t + le] = e;
        }
    }
    qsort(R, (size_t)Nt * 3, sizeof(rawedge), cmp_raw);
    return R;
}

int orc_rwg_count(const int* F, int Nt, int Nv) {
    for (int i = 0; i < 3 * Nt; ++i)
        if (F[i] < 1 || F[i] > Nv) return -1;
    rawedge* R = sorted_raw_edges(F, Nt);
    int ne = 0, n = 3 * Nt;
    for (int i = 0; i < n;) {
        int j = i;
        while (j < n && R[j].a == R[i].a && R[j].b == R[i].b) ++j;
        if (j - i == 2) ++ne; /* isInt = counts==2 (:29-31) */
        i = j;
    }
    free(R);
    return ne;
}

/* opposite_vertex_and_area (build_rwg_interior.m:103-113) */
static int opposite_vertex_and_area(const double* V, int Nv, const int* F,
                                    int Nt, int t, int va, int vb, int* opp,
                                    double* A) {
    int ids[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
    int cnt = 0;
    for (int i = 0; i < 3; ++i)
        if (ids[i] != va && ids[i] != vb) { *opp = ids[i]; ++cnt; }
    if (cnt != 1) return -1;
    double r1[3], r2[3], r3[3];
    for (int d = 0; d < 3; ++d) {
        r1[d] = V[(ids[0] - 1) + d * Nv];
        r2[d] = V[(ids[1] - 1) + d * Nv];
        r3[d] = V[(ids[2] - 1) + d * Nv];
    }
    double e1[3] = {r2[0] - r1[0], r2[1] - r1[1], r2[2] - r1[2]};
    double e2[3] = {r3[0] - r1[0], r3[1] - r1[1], r3[2] - r1[2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    *A = 0.5 * norm3(c);
    return 0;
}

/* build_rwg_interior.m:1-101 */
int orc_build_rwg(const double* V, int Nv, const int* F, int Nt, int Ne,
                  int* edge, int* plusTri, int* minusTri, int* plusSign,
                  int* minusSign, double* rp, double* rm, double* Ap,
                  double* Am, double* len, double* center, int* zSign) {
    rawedge* R = sorted_raw_edges(F, Nt);
    int n = 3 * Nt, e = 0;
    for (int i = 0; i < n;) {
        int j = i;
        while (j < n && R[j].a == R[i].a && R[j].b == R[i].b) ++j;
        if (j - i == 2 && e < Ne) {
            int a = R[i].a, b = R[i].b;
            edge[e] = a;
            edge[e + Ne] = b;
            for (int s = 0; s < 2; ++s) {
                int raw = R[i + s].raw, t = raw / 3, le = raw % 3;
                int tri[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
                int da = tri[le], db = tri[(le + 1) % 3];
                int sg = (da == a && db == b) ? +1 : -1; /* :53-54, :59-60 */
                if (s == 0) { plusTri[e] = t + 1; plusSign[e] = sg; }
                else { minusTri[e] = t + 1; minusSign[e] = sg; }
            }
            ++e;
        }
        i = j;
    }
    free(R);
    if (e != Ne) return -2;
    double zmax = -INFINITY, zmin = INFINITY;
    for (int v = 0; v < Nv; ++v) {
        double z = V[v + 2 * Nv];
        if (z > zmax) zmax = z;
        if (z < zmin) zmin = z;
    }
    double zmid = 0.5 * (zmax + zmin);
    for (e = 0; e < Ne; ++e) {
        int va = edge[e], vb = edge[e + Ne];
        double ra[3], rb[3], d[3];
        for (int k = 0; k < 3; ++k) {
            ra[k] = V[(va - 1) + k * Nv];
            rb[k] = V[(vb - 1) + k * Nv];
            d[k] = rb[k] - ra[k];
        }
        len[e] = norm3(d);
        for (int k = 0; k < 3; ++k) center[e + k * Ne] = 0.5 * (ra[k] + rb[k]);
        int oppP, oppM;
        double aP, aM;
        if (opposite_vertex_and_area(V, Nv, F, Nt, plusTri[e] - 1, va, vb, &oppP, &aP)) return -1;
        if (opposite_vertex_and_area(V, Nv, F, Nt, minusTri[e] - 1, va, vb, &oppM, &aM)) return -1;
        for (int k = 0; k < 3; ++k) {
            rp[e + k * Ne] = V[(oppP - 1) + k * Nv];
            rm[e + k * Ne] = V[(oppM - 1) + k * Nv];
        }
        Ap[e] = aP;
        Am[e] = aM;
        zSign[e] = center[e + 2 * Ne] > zmid ? +1 : -1; /* :88-90 */
    }
    return 0;
}

/* --------------------------------------------------------- quadrature -- */

/* tri_quad_rule.m:1-23; order 7 adds Radon's 7-point degree-5 rule
 * (centroid + two orbits, a = (6 -+ sqrt(15))/21), weights scaled to the
 * reference area 1/2. */
int orc_quad_rule(int order, double* qp, double* qw) {
    if (order == 1) {
        qp[0] = 1.0 / 3; qp[1] = 1.0 / 3;
        qw[0] = 1.0 / 2;
        return 1;
    }
    if (order == 2) {
        double u[3] = {1.0 / 6, 2.0 / 3, 1.0 / 6}, v[3] = {1.0 / 6, 1.0 / 6, 2.0 / 3};
        for (int i = 0; i < 3; ++i) { qp[i] = u[i]; qp[i + 3] = v[i]; qw[i] = (1.0 / 6) * 1.0; }
        return 3;
    }
    if (order == 7) {
        double s15 = sqrt(15.0);
        double a1 = (6.0 - s15) / 21.0, a2 = (6.0 + s15) / 21.0;
        double w0 = 9.0 / 80.0;
        double w1 = (155.0 - s15) / 2400.0, w2 = (155.0 + s15) / 2400.0;
        double u[7] = {1.0 / 3, a1, 1 - 2 * a1, a1, a2, 1 - 2 * a2, a2};
        double v[7] = {1.0 / 3, a1, a1, 1 - 2 * a1, a2, a2, 1 - 2 * a2};
        double w[7] = {w0, w1, w1, w1, w2, w2, w2};
        for (int i = 0; i < 7; ++i) { qp[i] = u[i]; qp[i + 7] = v[i]; qw[i] = w[i]; }
        return 7;
    }
    {
        double a = 0.445948490915965, b = 0.108103018168070;
        double w1 = 0.111690794839005, w2 = 0.054975871827661;
        double u[6] = {a, a, 1 - 2 * a, b, b, 1 - 2 * b};
        double v[6] = {a, 1 - 2 * a, a, b, 1 - 2 * b, b};
        double w[6] = {w1, w1, w1, w2, w2, w2};
        for (int i = 0; i < 6; ++i) { qp[i] = u[i]; qp[i + 6] = v[i]; qw[i] = w[i]; }
        return 6;
    }
}

/* ---------------------------------------------------- triangle helpers -- */

typedef struct { double r1[3], r2[3], r3[3], area2, cent[3]; } trigeom;

/* precompute_tri_geom.m:1-18 for one triangle (0-based t) */
static void tri_geom(const double* V, int Nv, const int* F, int Nt, int t, trigeom* g) {
    int ids[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
    for (int d = 0; d < 3; ++d) {
        g->r1[d] = V[(ids[0] - 1) + d * Nv];
        g->r2[d] = V[(ids[1] - 1) + d * Nv];
        g->r3[d] = V[(ids[2] - 1) + d * Nv];
        g->cent[d] = (g->r1[d] + g->r2[d] + g->r3[d]) / 3.0;
    }
    double e1[3] = {g->r2[0] - g->r1[0], g->r2[1] - g->r1[1], g->r2[2] - g->r1[2]};
    double e2[3] = {g->r3[0] - g->r1[0], g->r3[1] - g->r1[1], g->r3[2] - g->r1[2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    g->area2 = norm3(c);
}

/* map_ref_to_vertices.m / map_ref_to_tri.m: A + u(B-A) + v(C-A) */
static void map_ref(const double* A, const double* B, const double* C,
                    const double* qp, int Nq, double* r /* Nq x 3 row-major */) {
    for (int i = 0; i < Nq; ++i) {
        double u = qp[i], v = qp[i + Nq];
        for (int d = 0; d < 3; ++d)
            r[3 * i + d] = A[d] + u * (B[d] - A[d]) + v * (C[d] - A[d]);
    }
}

/* subdivide_triangles.m:1-20: returns 4^depth children, 9 doubles each
 * (A,B,C), child order [A,AB,CA],[AB,B,BC],[CA,BC,C],[AB,BC,CA]. */
static double* subdivide(const double* A, const double* B, const double* C, int depth, int* Ns) {
    int n = 1;
    double* cur = (double*)malloc(sizeof(double) * 9);
    memcpy(cur, A, 24); memcpy(cur + 3, B, 24); memcpy(cur + 6, C, 24);
    for (int d = 0; d < depth; ++d) {
        double* nw = (double*)malloc(sizeof(double) * 9 * 4 * (size_t)n);
        int t = 0;
        for (int s = 0; s < n; ++s) {
            const double *A0 = cur + 9 * s, *B0 = A0 + 3, *C0 = A0 + 6;
            double AB[3], BC[3], CA[3];
            for (int k = 0; k < 3; ++k) {
                AB[k] = 0.5 * (A0[k] + B0[k]);
                BC[k] = 0.5 * (B0[k] + C0[k]);
                CA[k] = 0.5 * (C0[k] + A0[k]);
            }
            const double* kids[4][3] = {{A0, AB, CA}, {AB, B0, BC}, {CA, BC, C0}, {AB, BC, CA}};
            for (int c = 0; c < 4; ++c, ++t)
                for (int v = 0; v < 3; ++v) memcpy(nw + 9 * t + 3 * v, kids[c][v], 24);
        }
        free(cur);
        cur = nw;
        n *= 4;
    }
    *Ns = n;
    return cur;
}

static double tri_area2(const double* A, const double* B, const double* C) {
    double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    return norm3(c);
}

/* rwg_eval_on_triangle.m:1-29 (m, t 1-based). f row-major Nq x 3. */
static void rwg_eval(const orc_problem* P, int m, int t, const double* r, int Nq,
                     double* f, double* divf) {
    int e = m - 1;
    if (t == P->plusTri[e]) {
        double s = P->plusSign[e], l = P->len[e], A = P->Ap[e];
        double sc = s * (l / (2 * A));
        for (int i = 0; i < Nq; ++i) {
            for (int d = 0; d < 3; ++d) f[3 * i + d] = sc * (r[3 * i + d] - P->rp[e + d * P->Ne]);
            divf[i] = s * (l / A);
        }
        return;
    }
    if (t == P->minusTri[e]) {
        double s = P->minusSign[e], l = P->len[e], A = P->Am[e];
        double sc = s * (l / (2 * A));
        for (int i = 0; i < Nq; ++i) {
            for (int d = 0; d < 3; ++d) f[3 * i + d] = sc * (P->rm[e + d * P->Ne] - r[3 * i + d]);
            divf[i] = s * (-l / A);
        }
        return;
    }
    for (int i = 0; i < Nq; ++i) { f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0; divf[i] = 0.0; }
}

int orc_triangles_are_near(const double* V, int Nv, const int* F, int Nt, int t1,
                           int t2, double eps_scale) {
    trigeom g1, g2;
    tri_geom(V, Nv, F, Nt, t1 - 1, &g1);
    tri_geom(V, Nv, F, Nt, t2 - 1, &g2);
    double d3[3] = {g1.cent[0] - g2.cent[0], g1.cent[1] - g2.cent[1], g1.cent[2] - g2.cent[2]};
    double d = norm3(d3);
    double A1 = 0.5 * g1.area2, A2 = 0.5 * g2.area2;
    double l1 = sqrt(4 * A1 / sqrt(3.0)), l2 = sqrt(4 * A2 / sqrt(3.0));
    double h = 0.5 * (l1 + l2);
    return d < eps_scale * h;
}

/* rwg_pair_integrals_regular.m:1-29 */
static void pair_regular(const orc_problem* P, int m, int n, const trigeom* g1,
                         const trigeom* g2, int t1, int t2, const double* qp,
                         const double* qw, int Nq, double kr, double ki, cx* Ivec, cx* Isca) {
    double r[3 * 16], rq[3 * 16], fm[3 * 16], fn[3 * 16], dm[16], dn[16];
    map_ref(g1->r1, g1->r2, g1->r3, qp, Nq, r);
    map_ref(g2->r1, g2->r2, g2->r3, qp, Nq, rq);
    double J1 = g1->area2, J2 = g2->area2;
    rwg_eval(P, m, t1, r, Nq, fm, dm);
    rwg_eval(P, n, t2, rq, Nq, fn, dn);
    cx iv = cmk(0, 0), is = cmk(0, 0);
    for (int i = 0; i < Nq; ++i)
        for (int j = 0; j < Nq; ++j) {
            double d[3] = {r[3 * i] - rq[3 * j], r[3 * i + 1] - rq[3 * j + 1], r[3 * i + 2] - rq[3 * j + 2]};
            double R = norm3(d);
            if (R < 1e-15) R = 1e-15;
            cx G = green(kr, ki, R);
            double w = qw[i] * qw[j] * J1 * J2;
            double dot = fm[3 * i] * fn[3 * j] + fm[3 * i + 1] * fn[3 * j + 1] + fm[3 * i + 2] * fn[3 * j + 2];
            iv = cadd(iv, cscale(cscale(G, dot), w));
            is = cadd(is, cscale(cscale(G, dm[i] * dn[j]), w));
        }
    *Ivec = iv;
    *Isca = is;
}

/* Shared body of rwg_pair_integrals_near_singular.m:24-53 (self == 0) and
 * rwg_pair_integrals_singular.m:20-61 (self == 1): both triangles are
 * subdivided, each sub-triangle pair contributes sum(sum(W.*X.*G)) with
 * column sums first. */
static void pair_subdivided(const orc_problem* P, int m, int n, const trigeom* g1,
                            const trigeom* g2, int t1, int t2, int self,
                            const orc_opts* o, double kr, double ki, cx* Ivec, cx* Isca) {
    double qp2[32], qw2[16];
    int Nq2 = orc_quad_rule(o->near_order, qp2, qw2);
    int depth = self ? o->self_subdiv : o->near_subdiv;
    int Ns1, Ns2;
    double* sub1 = subdivide(g1->r1, g1->r2, g1->r3, depth, &Ns1);
    double* sub2 = subdivide(g2->r1, g2->r2, g2->r3, depth, &Ns2);
    double a_soft_near = 0.0;
    if (!self) {
        double At1 = 0.5 * g1->area2, At2 = 0.5 * g2->area2;
        double a1 = sqrt(At1 / PI_D), a2 = sqrt(At2 / PI_D);
        double aeq = a1 < a2 ? a1 : a2; /* min(...) :18 */
        a_soft_near = o->near_alpha * aeq;
    }
    double r[48], rq[48], fm[48], fn[48], dm[16], dn[16], wp[16], wq[16];
    cx iv = cmk(0, 0), is = cmk(0, 0);
    for (int p = 0; p < Ns1; ++p) {
        const double *Ap_ = sub1 + 9 * p, *Bp = Ap_ + 3, *Cp = Ap_ + 6;
        double a2p = tri_area2(Ap_, Bp, Cp);
        map_ref(Ap_, Bp, Cp, qp2, Nq2, r);
        for (int i = 0; i < Nq2; ++i) wp[i] = qw2[i] * a2p;
        rwg_eval(P, m, t1, r, Nq2, fm, dm);
        double a_soft_self = 0.0;
        if (self) {
            double Asub = 0.5 * a2p;
            double aeq = sqrt(Asub / PI_D);
            a_soft_self = o->self_alpha * aeq;
        }
        for (int q = 0; q < Ns2; ++q) {
            const double *Aq = sub2 + 9 * q, *Bq = Aq + 3, *Cq = Aq + 6;
            double a2q = tri_area2(Aq, Bq, Cq);
            map_ref(Aq, Bq, Cq, qp2, Nq2, rq);
            for (int j = 0; j < Nq2; ++j) wq[j] = qw2[j] * a2q;
            rwg_eval(P, n, t2, rq, Nq2, fn, dn);
            cx bv = cmk(0, 0), bs = cmk(0, 0);
            for (int j = 0; j < Nq2; ++j) {
                cx cv = cmk(0, 0), cs = cmk(0, 0);
                for (int i = 0; i < Nq2; ++i) {
                    double dx = r[3 * i] - rq[3 * j];
                    double dy = r[3 * i + 1] - rq[3 * j + 1];
                    double dz = r[3 * i + 2] - rq[3 * j + 2];
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    if (!self) {
                        R = sqrt(R * R + a_soft_near * a_soft_near);
                    } else if (p == q) {
                        R = sqrt(R * R + a_soft_self * a_soft_self);
                    } else {
                        R = R > 1e-15 ? R : 1e-15;
                    }
                    cx G = green(kr, ki, R);
                    double ff = fm[3 * i] * fn[3 * j] + fm[3 * i + 1] * fn[3 * j + 1] + fm[3 * i + 2] * fn[3 * j + 2];
                    double dd = dm[i] * dn[j];
                    double W = wp[i] * wq[j];
                    cv = cadd(cv, cscale(G, W * ff));
                    cs = cadd(cs, cscale(G, W * dd));
                }
                bv = cadd(bv, cv);
                bs = cadd(bs, cs);
            }
            iv = cadd(iv, bv);
            is = cadd(is, bs);
        }
    }
    free(sub1);
    free(sub2);
    *Ivec = iv;
    *Isca = is;
}

/* ------------------------------------------- singularity extraction -- */

static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Closed-form potentials, edge by edge (edge i from v_i to v_{i+1},
 * counter-clockwise about the normal n): with l^-, l^+ the in-line
 * coordinates of the edge ends, P0 the signed distance of the projection
 * rho from the edge line (positive inside), d the height of r above the
 * plane, R0^2 = P0^2 + d^2 and R^+- the distances to the edge ends,
 *   f_i  = ln((R^+ + l^+)/(R^- + l^-))
 *   b_i  = atan(P0 l^+ / (R0^2 + |d| R^+)) - atan(P0 l^- / (R0^2 + |d| R^-))
 *   int 1/R          = sum_i  P0_i f_i - |d| b_i
 *   int (r'-rho)/R   = 1/2 sum_i u_i (R0_i^2 f_i + l^+ R^+ - l^- R^-)
 * (u_i the outward in-plane edge normal).  f_i uses the algebraically equal
 * ln((R^- - l^-)/(R^+ - l^+)) when both l are negative, and
 * ln((R^+ + l^+)(R^- - l^-)/R0^2) when rho projects inside the edge, to
 * avoid cancellation; an edge whose line passes through r contributes no
 * P0 f term (P0 f -> 0). */
int orc_tri_potentials(const double* r, const double* v1, const double* v2, const double* v3,
                       double* out) {
    const double* v[3] = {v1, v2, v3};
    double e1[3], e2[3], nn[3];
    for (int d = 0; d < 3; ++d) { e1[d] = v2[d] - v1[d]; e2[d] = v3[d] - v1[d]; }
    cross3(e1, e2, nn);
    double nl = norm3(nn);
    for (int d = 0; d < 3; ++d) nn[d] /= nl;
    double rv[3] = {r[0] - v1[0], r[1] - v1[1], r[2] - v1[2]};
    double h = dot3(rv, nn);
    double ah = fabs(h);
    double rho[3] = {r[0] - h * nn[0], r[1] - h * nn[1], r[2] - h * nn[2]};
    double scale = norm3(e1) + norm3(e2);
    double I0 = 0.0, I1[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const double *a = v[i], *b = v[(i + 1) % 3];
        double s[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double L = norm3(s);
        double lh[3] = {s[0] / L, s[1] / L, s[2] / L};
        double u[3];
        cross3(lh, nn, u);
        double am[3] = {a[0] - rho[0], a[1] - rho[1], a[2] - rho[2]};
        double bm[3] = {b[0] - rho[0], b[1] - rho[1], b[2] - rho[2]};
        double lm = dot3(am, lh), lp = dot3(bm, lh);
        double P0 = dot3(am, u);
        double R02 = P0 * P0 + h * h;
        double Rm = sqrt(R02 + lm * lm), Rp = sqrt(R02 + lp * lp);
        double f;
        if (R02 <= 1e-28 * scale * scale) {
            f = 0.0; /* r on the edge line: P0 f and R0^2 f vanish */
        } else if (lm >= 0.0) {
            f = log((Rp + lp) / (Rm + lm));
        } else if (lp <= 0.0) {
            f = log((Rm - lm) / (Rp - lp));
        } else {
            /* rho projects inside the edge: R^- + l^- = R0^2 / (R^- - l^-) */
            f = log((Rp + lp) * (Rm - lm) / R02);
        }
        double beta = 0.0;
        if (ah > 0.0)
            beta = atan(P0 * lp / (R02 + ah * Rp)) - atan(P0 * lm / (R02 + ah * Rm));
        I0 += P0 * f - ah * beta;
        double c = 0.5 * (R02 * f + lp * Rp - lm * Rm);
        for (int d = 0; d < 3; ++d) I1[d] += c * u[d];
    }
    out[0] = I0;
    out[1] = I1[0]; out[2] = I1[1]; out[3] = I1[2];
    return 0;
}

/* (exp(-jkR) - 1)/(4 pi R), with its limit -jk/(4 pi) at R = 0; a Taylor
 * series for |kR| < 1e-3. */
static cx green_smooth(double kr, double ki, double R) {
    double ar = kr * R, ai = ki * R;
    if (sqrt(ar * ar + ai * ai) < 1e-3) {
        /* sum_{q>=1} (-jk)^q R^(q-1)/q! with -jk = (ki, -kr) */
        cx mjk = cmk(ki, -kr);
        cx term = mjk, acc = mjk;
        for (int q = 2; q <= 6; ++q) {
            term = cscale(cmul(term, mjk), R / q);
            acc = cadd(acc, term);
        }
        return cscale(acc, 1.0 / (4.0 * PI_D));
    }
    cx g = green(kr, ki, R);
    return cmk(g.re - 1.0 / (4.0 * PI_D * R), g.im);
}

/* f_n = c (r' - v) on triangle t of RWG n, div f_n = dv */
static void rwg_affine(const orc_problem* P, int n, int t, double* c, double* v, double* dv) {
    int e = n - 1;
    if (t == P->plusTri[e]) {
        double s = P->plusSign[e], l = P->len[e], A = P->Ap[e];
        *c = s * (l / (2 * A));
        for (int d = 0; d < 3; ++d) v[d] = P->rp[e + d * P->Ne];
        *dv = s * (l / A);
    } else if (t == P->minusTri[e]) {
        double s = P->minusSign[e], l = P->len[e], A = P->Am[e];
        *c = -s * (l / (2 * A));
        for (int d = 0; d < 3; ++d) v[d] = P->rm[e + d * P->Ne];
        *dv = s * (-l / A);
    } else {
        *c = 0.0; *dv = 0.0; v[0] = v[1] = v[2] = 0.0;
    }
}

/* Outer triangle t1 (RWG m), inner t2 (RWG n), t1 <= t2. */
static void pair_extract_ordered(const orc_problem* P, int m, int n, const trigeom* g1, const trigeom* g2,
                                 int t1, int t2, const orc_opts* o, double kr, double ki, cx* Ivec, cx* Isca) {
    double qp2[32], qw2[16];
    int Nq2 = orc_quad_rule(o->near_order, qp2, qw2);
    int depth = t1 == t2 ? o->self_subdiv : o->near_subdiv;
    int Ns1, Ns2;
    double* sub1 = subdivide(g1->r1, g1->r2, g1->r3, depth, &Ns1);
    double* sub2 = subdivide(g2->r1, g2->r2, g2->r3, depth, &Ns2);
    double cn, vn[3], dn;
    rwg_affine(P, n, t2, &cn, vn, &dn);
    double r[48], rq[48], fm[48], fn[48], dm[16], dq[16], wp[16], wq[16];
    cx iv = cmk(0, 0), is = cmk(0, 0);
    const double inv4pi = 1.0 / (4.0 * PI_D);
    for (int p = 0; p < Ns1; ++p) {
        const double *Ap_ = sub1 + 9 * p, *Bp = Ap_ + 3, *Cp = Ap_ + 6;
        double a2p = tri_area2(Ap_, Bp, Cp);
        map_ref(Ap_, Bp, Cp, qp2, Nq2, r);
        for (int i = 0; i < Nq2; ++i) wp[i] = qw2[i] * a2p;
        rwg_eval(P, m, t1, r, Nq2, fm, dm);
        for (int i = 0; i < Nq2; ++i) {
            const double* ri = r + 3 * i;
            /* static part: inner integral in closed form */
            double pot[4];
            orc_tri_potentials(ri, g2->r1, g2->r2, g2->r3, pot);
            double e1[3], e2[3], nn[3];
            for (int d = 0; d < 3; ++d) { e1[d] = g2->r2[d] - g2->r1[d]; e2[d] = g2->r3[d] - g2->r1[d]; }
            cross3(e1, e2, nn);
            double nl = norm3(nn);
            for (int d = 0; d < 3; ++d) nn[d] /= nl;
            double h = (ri[0] - g2->r1[0]) * nn[0] + (ri[1] - g2->r1[1]) * nn[1] + (ri[2] - g2->r1[2]) * nn[2];
            double vec[3];
            for (int d = 0; d < 3; ++d) vec[d] = cn * (pot[1 + d] + (ri[d] - h * nn[d] - vn[d]) * pot[0]);
            double ffs = fm[3 * i] * vec[0] + fm[3 * i + 1] * vec[1] + fm[3 * i + 2] * vec[2];
            cx cv = cmk(inv4pi * ffs, 0.0);
            cx cs = cmk(inv4pi * dn * pot[0], 0.0);
            /* smooth part by quadrature */
            for (int q = 0; q < Ns2; ++q) {
                const double *Aq = sub2 + 9 * q, *Bq = Aq + 3, *Cq = Aq + 6;
                double a2q = tri_area2(Aq, Bq, Cq);
                map_ref(Aq, Bq, Cq, qp2, Nq2, rq);
                for (int j = 0; j < Nq2; ++j) wq[j] = qw2[j] * a2q;
                rwg_eval(P, n, t2, rq, Nq2, fn, dq);
                for (int j = 0; j < Nq2; ++j) {
                    double dx = ri[0] - rq[3 * j], dy = ri[1] - rq[3 * j + 1], dz = ri[2] - rq[3 * j + 2];
                    cx gs = green_smooth(kr, ki, sqrt(dx * dx + dy * dy + dz * dz));
                    double ff = fm[3 * i] * fn[3 * j] + fm[3 * i + 1] * fn[3 * j + 1] + fm[3 * i + 2] * fn[3 * j + 2];
                    cv = cadd(cv, cscale(gs, wq[j] * ff));
                    cs = cadd(cs, cscale(gs, wq[j] * dq[j]));
                }
            }
            iv = cadd(iv, cscale(cv, wp[i]));
            is = cadd(is, cscale(cs, wp[i] * dm[i]));
        }
    }
    free(sub1);
    free(sub2);
    *Ivec = iv;
    *Isca = is;
}

/* The outer triangle is the lower-numbered one, and on a self pair the
 * outer function is the lower-numbered RWG (the integrals are symmetric
 * under exchange of the two sides, their discretisation is not). */
static void pair_extract(const orc_problem* P, int m, int n, const trigeom* g1, const trigeom* g2, int t1,
                         int t2, const orc_opts* o, double kr, double ki, cx* Ivec, cx* Isca) {
    if (t1 < t2 || (t1 == t2 && m <= n)) pair_extract_ordered(P, m, n, g1, g2, t1, t2, o, kr, ki, Ivec, Isca);
    else pair_extract_ordered(P, n, m, g2, g1, t2, t1, o, kr, ki, Ivec, Isca);
}

int orc_pair_integral(const orc_problem* P, int kind, int m, int n, int t1, int t2,
                      double k_re, double k_im, const orc_opts* o, double* out) {
    trigeom g1, g2;
    tri_geom(P->V, P->Nv, P->F, P->Nt, t1 - 1, &g1);
    tri_geom(P->V, P->Nv, P->F, P->Nt, t2 - 1, &g2);
    cx iv, is;
    if (kind == 0) {
        if (t1 == t2) return -1; /* rwg_pair_integrals_regular.m:2-4 */
        double qp[32], qw[16];
        int Nq = orc_quad_rule(o->quad_order, qp, qw);
        pair_regular(P, m, n, &g1, &g2, t1, t2, qp, qw, Nq, k_re, k_im, &iv, &is);
    } else if (kind == 1) {
        pair_subdivided(P, m, n, &g1, &g2, t1, t2, 0, o, k_re, k_im, &iv, &is);
    } else if (kind == 3) {
        pair_extract(P, m, n, &g1, &g2, t1, t2, o, k_re, k_im, &iv, &is);
    } else {
        if (t1 != t2) return -1; /* rwg_pair_integrals_singular.m:3-5 */
        pair_subdivided(P, m, n, &g1, &g2, t1, t2, 1, o, k_re, k_im, &iv, &is);
    }
    out[0] = iv.re; out[1] = iv.im; out[2] = is.re; out[3] = is.im;
    return 0;
}

/* ----------------------------------------------------------- assembly -- */

/* assemble_ZV_EFIE_RWG_DIFF_CPU.m:3-36 for the listed rows */
int orc_assemble_z_rows(const orc_problem* P, double k_re, double k_im, double omega,
                        double mu, double eps_re, double eps_im, const orc_opts* o,
                        const int* rows, int nrows, int threads, double* Zrows) {
    int Ne = P->Ne, Nt = P->Nt;
    double qp[32], qw[16];
    int Nq = orc_quad_rule(o->quad_order, qp, qw);
    trigeom* G = (trigeom*)malloc(sizeof(trigeom) * (size_t)Nt);
    for (int t = 0; t < Nt; ++t) tri_geom(P->V, P->Nv, P->F, Nt, t, &G[t]);
    cx jwmu = cmk(0.0, omega * mu);                   /* 1j*omega*mu */
    cx jwe = cmul(cmk(0.0, omega), cmk(eps_re, eps_im)); /* 1j*omega*eps */
    cx beta = cdiv(cmk(1.0, 0.0), jwe);
    int bad = 0;
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : bad)
#endif
    for (int rr = 0; rr < nrows; ++rr) {
        int m = rows[rr];
        if (m < 1 || m > Ne) { bad |= 1; continue; }
        int tm[2] = {P->plusTri[m - 1], P->minusTri[m - 1]};
        for (int n = 1; n <= Ne; ++n) {
            int tn[2] = {P->plusTri[n - 1], P->minusTri[n - 1]};
            cx Zmn = cmk(0, 0);
            for (int im = 0; im < 2; ++im)
                for (int in = 0; in < 2; ++in) {
                    int t1 = tm[im], t2 = tn[in];
                    if (t1 == 0 || t2 == 0) continue;
                    cx iv, is;
                    const trigeom *g1 = &G[t1 - 1], *g2 = &G[t2 - 1];
                    if (t1 == t2) {
                        if (o->extract) pair_extract(P, m, n, g1, g2, t1, t2, o, k_re, k_im, &iv, &is);
                        else pair_subdivided(P, m, n, g1, g2, t1, t2, 1, o, k_re, k_im, &iv, &is);
                    } else {
                        double d3[3] = {g1->cent[0] - g2->cent[0], g1->cent[1] - g2->cent[1],
                                        g1->cent[2] - g2->cent[2]};
                        double d = norm3(d3);
                        double A1 = 0.5 * g1->area2, A2 = 0.5 * g2->area2;
                        double l1 = sqrt(4 * A1 / sqrt(3.0)), l2 = sqrt(4 * A2 / sqrt(3.0));
                        double h = 0.5 * (l1 + l2);
                        if (d < o->near_sing_eps * h && o->extract)
                            pair_extract(P, m, n, g1, g2, t1, t2, o, k_re, k_im, &iv, &is);
                        else if (d < o->near_sing_eps * h)
                            pair_subdivided(P, m, n, g1, g2, t1, t2, 0, o, k_re, k_im, &iv, &is);
                        else
                            pair_regular(P, m, n, g1, g2, t1, t2, qp, qw, Nq, k_re, k_im, &iv, &is);
                    }
                    /* Zmn + 1j*omega*mu*Ivec + (1/(1j*omega*eps))*Isca  (:30) */
                    Zmn = cadd(cadd(Zmn, cmul(jwmu, iv)), cmul(beta, is));
                }
            Zrows[2 * ((size_t)rr + (size_t)nrows * (n - 1))] = Zmn.re;
            Zrows[2 * ((size_t)rr + (size_t)nrows * (n - 1)) + 1] = Zmn.im;
        }
    }
    free(G);
    return bad ? -1 : 0;
}

/* kernel_Z_centroid (patch_sweep_cuda.cu:94-153) with its host precompute
 * (:244-326): one point per triangle at the centroid, weight = area,
 * self softening a = self_alpha*sqrt(A/pi) on equal triangles. */
int orc_assemble_z_centroid(const orc_problem* P, double k_re, double k_im, double alpha_re,
                            double alpha_im, double beta_re, double beta_im,
                            double self_alpha, double* Z) {
    int Ne = P->Ne, Nt = P->Nt;
    double* cent = (double*)malloc(sizeof(double) * 3 * (size_t)Nt);
    double* asoft = (double*)malloc(sizeof(double) * (size_t)Nt);
    for (int t = 0; t < Nt; ++t) {
        int i1 = P->F[t] - 1, i2 = P->F[t + Nt] - 1, i3 = P->F[t + 2 * Nt] - 1;
        double r1[3], r2[3], r3[3];
        for (int d = 0; d < 3; ++d) {
            r1[d] = P->V[i1 + d * P->Nv]; r2[d] = P->V[i2 + d * P->Nv]; r3[d] = P->V[i3 + d * P->Nv];
            cent[3 * t + d] = (r1[d] + r2[d] + r3[d]) / 3.0;
        }
        double area2 = tri_area2(r1, r2, r3);
        double A = 0.5 * area2;
        asoft[t] = self_alpha * sqrt(A / PI_D);
    }
    /* side data per RWG: pos, f, div, w, t */
    double* pos = (double*)calloc(6 * (size_t)Ne, sizeof(double));
    double* f = (double*)calloc(6 * (size_t)Ne, sizeof(double));
    double* dv = (double*)calloc(2 * (size_t)Ne, sizeof(double));
    double* w = (double*)calloc(2 * (size_t)Ne, sizeof(double));
    for (int e = 0; e < Ne; ++e) {
        int tp = P->plusTri[e], tmn = P->minusTri[e];
        double l = P->len[e];
        if (tp > 0) for (int d = 0; d < 3; ++d) pos[6 * e + d] = cent[3 * (tp - 1) + d];
        if (tmn > 0) for (int d = 0; d < 3; ++d) pos[6 * e + 3 + d] = cent[3 * (tmn - 1) + d];
        w[2 * e] = tp > 0 ? P->Ap[e] : 0.0;
        w[2 * e + 1] = tmn > 0 ? P->Am[e] : 0.0;
        if (tp > 0 && P->Ap[e] > 0) {
            double s = P->plusSign[e], A = P->Ap[e], sc = s * (l / (2.0 * A));
            for (int d = 0; d < 3; ++d) f[6 * e + d] = sc * (pos[6 * e + d] - P->rp[e + d * Ne]);
            dv[2 * e] = s * (l / A);
        }
        if (tmn > 0 && P->Am[e] > 0) {
            double s = P->minusSign[e], A = P->Am[e], sc = s * (l / (2.0 * A));
            for (int d = 0; d < 3; ++d) f[6 * e + 3 + d] = sc * (P->rm[e + d * Ne] - pos[6 * e + 3 + d]);
            dv[2 * e + 1] = s * (-(l / A));
        }
    }
    cx alpha = cmk(alpha_re, alpha_im), beta = cmk(beta_re, beta_im);
    for (int n = 0; n < Ne; ++n)
        for (int m = 0; m < Ne; ++m) {
            cx Zmn = cmk(0, 0);
            for (int sa = 0; sa < 2; ++sa) {
                int tm = sa == 0 ? P->plusTri[m] : P->minusTri[m];
                double wm = w[2 * m + sa];
                if (wm == 0.0 || tm == 0) continue;
                for (int sb = 0; sb < 2; ++sb) {
                    int tn = sb == 0 ? P->plusTri[n] : P->minusTri[n];
                    double wn = w[2 * n + sb];
                    if (wn == 0.0 || tn == 0) continue;
                    double d[3];
                    for (int k = 0; k < 3; ++k) d[k] = pos[6 * m + 3 * sa + k] - pos[6 * n + 3 * sb + k];
                    double R = norm3(d);
                    if (tm == tn) { double a = asoft[tm - 1]; R = sqrt(R * R + a * a); }
                    if (R < 1e-15) R = 1e-15;
                    /* green(): amp*(cos, -sin) * 1/(4 pi R) (patch_sweep_cuda.cu:76-88) */
                    double amp = exp(k_im * R), ph = k_re * R, inv = 1.0 / (4.0 * PI_D * R);
                    cx G = cmk(amp * cos(ph) * inv, -amp * sin(ph) * inv);
                    double dotff = f[6 * m + 3 * sa] * f[6 * n + 3 * sb] + f[6 * m + 3 * sa + 1] * f[6 * n + 3 * sb + 1] +
                                   f[6 * m + 3 * sa + 2] * f[6 * n + 3 * sb + 2];
                    double divprod = dv[2 * m + sa] * dv[2 * n + sb];
                    double ww = wm * wn;
                    cx tv = cscale(G, dotff * ww), ts = cscale(G, divprod * ww);
                    Zmn = cadd(Zmn, cadd(cmul(alpha, tv), cmul(beta, ts)));
                }
            }
            Z[2 * ((size_t)m + (size_t)n * Ne)] = Zmn.re;
            Z[2 * ((size_t)m + (size_t)n * Ne) + 1] = Zmn.im;
        }
    free(cent); free(asoft); free(pos); free(f); free(dv); free(w);
    return 0;
}

/* ---------------------------------------------------------- RHS, port -- */

/* local_gauss_on_layer (assemble_ZV_EFIE_RWG_DIFF_CPU.m:53-71) */
static int gauss_on_layer(int Ne, const double* center, const int* zSign, int edge0,
                          double amp, double spread, double* v) {
    int layer = zSign[edge0 - 1];
    double r0[3] = {center[edge0 - 1], center[edge0 - 1 + Ne], center[edge0 - 1 + 2 * Ne]};
    double sw = 0.0;
    for (int e = 0; e < Ne; ++e) {
        double d3[3] = {center[e] - r0[0], center[e + Ne] - r0[1], center[e + 2 * Ne] - r0[2]};
        double d = norm3(d3);
        double t = d / spread;
        double w = exp(-(t * t)) * (zSign[e] == layer ? 1.0 : 0.0);
        v[e] = w;
        sw += w;
    }
    if (sw < 1e-15) return -1;
    for (int e = 0; e < Ne; ++e) v[e] = amp * (v[e] / sw);
    return 0;
}

int orc_rhs(int Ne, const double* center, const int* zSign, int patchEdge, int groundEdge,
            double V0, double spread, double* Vrhs) {
    if (patchEdge < 1 || patchEdge > Ne || groundEdge < 1 || groundEdge > Ne) return -2;
    double* a = (double*)malloc(sizeof(double) * (size_t)Ne);
    double* b = (double*)malloc(sizeof(double) * (size_t)Ne);
    int rc = gauss_on_layer(Ne, center, zSign, patchEdge, +V0 / 2, spread, a);
    if (!rc) rc = gauss_on_layer(Ne, center, zSign, groundEdge, -V0 / 2, spread, b);
    if (!rc)
        for (int e = 0; e < Ne; ++e) {
            double v = (0.0 + a[e]) + b[e]; /* V = 0 + gauss(+) + gauss(-) */
            Vrhs[2 * e] = v;
            Vrhs[2 * e + 1] = 0.0;
        }
    free(a);
    free(b);
    return rc;
}

int orc_pick_port(int Ne, const double* center, const int* zSign, const double* p, int layerSign) {
    int best = -1;
    double bd = INFINITY;
    for (int e = 0; e < Ne; ++e) {
        if (zSign[e] != layerSign) continue;
        double d3[3] = {center[e] - p[0], center[e + Ne] - p[1], center[e + 2 * Ne] - p[2]};
        double d = norm3(d3);
        if (best < 0 || d < bd) { bd = d; best = e + 1; } /* first min on ties */
    }
    return best;
}

/* port_weights_gaussian.m:1-29 */
int orc_port_weights(const double* V, int Nv, const int* edge, int Ne, int patchEdge,
                     int groundEdge, double spread, double* wPatch, double* wGnd) {
    double* mid = (double*)malloc(sizeof(double) * 3 * (size_t)Ne);
    double zmin = INFINITY, zmax = -INFINITY;
    for (int e = 0; e < Ne; ++e) {
        int v1 = edge[e] - 1, v2 = edge[e + Ne] - 1;
        for (int d = 0; d < 3; ++d) mid[3 * e + d] = 0.5 * (V[v1 + d * Nv] + V[v2 + d * Nv]);
        if (mid[3 * e + 2] < zmin) zmin = mid[3 * e + 2];
        if (mid[3 * e + 2] > zmax) zmax = mid[3 * e + 2];
    }
    double zmid = 0.5 * (zmin + zmax);
    int edges[2] = {patchEdge, groundEdge};
    double* outs[2] = {wPatch, wGnd};
    for (int s = 0; s < 2; ++s) {
        int e0 = edges[s] - 1;
        int lay0 = mid[3 * e0 + 2] > zmid ? 1 : -1;
        double sum = 0.0;
        for (int e = 0; e < Ne; ++e) {
            int lay = mid[3 * e + 2] > zmid ? 1 : -1;
            double d3[3] = {mid[3 * e] - mid[3 * e0], mid[3 * e + 1] - mid[3 * e0 + 1], mid[3 * e + 2] - mid[3 * e0 + 2]};
            double t = norm3(d3) / spread;
            double w = exp(-(t * t)) * (lay == lay0 ? 1.0 : 0.0);
            outs[s][e] = w;
            sum += w;
        }
        for (int e = 0; e < Ne; ++e) outs[s][e] /= sum;
    }
    free(mid);
    return 0;
}

/* ------------------------------------------------------------------ LU -- */

int orc_zgetrf(int n, double* A, int* ipiv) {
    int info = 0;
    for (int j = 0; j < n; ++j) {
        /* izamax on |re|+|im|, first index on ties */
        int p = j;
        double best = -1.0;
        for (int i = j; i < n; ++i) {
            double* a = A + 2 * ((size_t)i + (size_t)j * n);
            double v = fabs(a[0]) + fabs(a[1]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p + 1;
        double* piv = A + 2 * ((size_t)p + (size_t)j * n);
        if (piv[0] == 0.0 && piv[1] == 0.0) {
            if (!info) info = j + 1;
            continue;
        }
        if (p != j)
            for (int c = 0; c < n; ++c) {
                double* x = A + 2 * ((size_t)j + (size_t)c * n);
                double* y = A + 2 * ((size_t)p + (size_t)c * n);
                double t0 = x[0], t1 = x[1];
                x[0] = y[0]; x[1] = y[1]; y[0] = t0; y[1] = t1;
            }
        cx d = cmk(A[2 * ((size_t)j + (size_t)j * n)], A[2 * ((size_t)j + (size_t)j * n) + 1]);
        cx inv = cdiv(cmk(1, 0), d);
        for (int i = j + 1; i < n; ++i) {
            double* a = A + 2 * ((size_t)i + (size_t)j * n);
            cx v = cmul(cmk(a[0], a[1]), inv);
            a[0] = v.re; a[1] = v.im;
        }
        for (int c = j + 1; c < n; ++c) {
            double* u = A + 2 * ((size_t)j + (size_t)c * n);
            cx uj = cmk(u[0], u[1]);
            if (uj.re == 0.0 && uj.im == 0.0) continue;
            double* col = A + 2 * ((size_t)c * n);
            double* lc = A + 2 * ((size_t)j * n);
            for (int i = j + 1; i < n; ++i) {
                cx l = cmk(lc[2 * i], lc[2 * i + 1]);
                cx pr = cmul(l, uj);
                col[2 * i] -= pr.re;
                col[2 * i + 1] -= pr.im;
            }
        }
    }
    return info;
}

void orc_zgetrs(int n, const double* LU, const int* ipiv, double* b) {
    for (int i = 0; i < n; ++i) {
        int p = ipiv[i] - 1;
        if (p != i) {
            double t0 = b[2 * i], t1 = b[2 * i + 1];
            b[2 * i] = b[2 * p]; b[2 * i + 1] = b[2 * p + 1];
            b[2 * p] = t0; b[2 * p + 1] = t1;
        }
    }
    for (int j = 0; j < n; ++j) { /* L y = b, unit diagonal */
        cx bj = cmk(b[2 * j], b[2 * j + 1]);
        for (int i = j + 1; i < n; ++i) {
            const double* l = LU + 2 * ((size_t)i + (size_t)j * n);
            cx pr = cmul(cmk(l[0], l[1]), bj);
            b[2 * i] -= pr.re; b[2 * i + 1] -= pr.im;
        }
    }
    for (int j = n - 1; j >= 0; --j) { /* U x = y */
        const double* u = LU + 2 * ((size_t)j + (size_t)j * n);
        cx xj = cdiv(cmk(b[2 * j], b[2 * j + 1]), cmk(u[0], u[1]));
        b[2 * j] = xj.re; b[2 * j + 1] = xj.im;
        for (int i = 0; i < j; ++i) {
            const double* a = LU + 2 * ((size_t)i + (size_t)j * n);
            cx pr = cmul(cmk(a[0], a[1]), xj);
            b[2 * i] -= pr.re; b[2 * i + 1] -= pr.im;
        }
    }
}

/* -------------------------------------------------------------- sweep -- */

int orc_sweep(const orc_problem* P, const double* center, const int* zSign, const double* freqs,
              int Nf, double mu, double eps_re, double eps_im, int patchEdge, int groundEdge,
              double V0, double spread, double Z0, const orc_opts* o, int centroid,
              double self_alpha, int threads, double* Zin, double* S11) {
    int Ne = P->Ne;
    if (patchEdge < 1 || patchEdge > Ne || groundEdge < 1 || groundEdge > Ne) return -2;
    double* Z = (double*)malloc(sizeof(double) * 2 * (size_t)Ne * Ne);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)Ne);
    double* b0 = (double*)malloc(sizeof(double) * 2 * (size_t)Ne);
    double* wP = (double*)malloc(sizeof(double) * (size_t)Ne);
    double* wG = (double*)malloc(sizeof(double) * (size_t)Ne);
    int* ipiv = (int*)malloc(sizeof(int) * (size_t)Ne);
    int* rows = (int*)malloc(sizeof(int) * (size_t)Ne);
    for (int i = 0; i < Ne; ++i) rows[i] = i + 1;
    int rc = orc_rhs(Ne, center, zSign, patchEdge, groundEdge, V0, spread, b0);
    if (rc) goto done;
    {
        /* port_weights_gaussian.m:5-28 on the edge midpoints, which are
         * exactly rwg.center (0.5*(ra+rb), build_rwg_interior.m:71) */
        double zmin = INFINITY, zmax = -INFINITY;
        for (int e = 0; e < Ne; ++e) {
            double z = center[e + 2 * Ne];
            if (z < zmin) zmin = z;
            if (z > zmax) zmax = z;
        }
        double zmid = 0.5 * (zmin + zmax);
        int edges[2] = {patchEdge, groundEdge};
        double* outs[2] = {wP, wG};
        for (int s = 0; s < 2; ++s) {
            int e0 = edges[s] - 1;
            int lay0 = center[e0 + 2 * Ne] > zmid ? 1 : -1;
            double sum = 0.0;
            for (int e = 0; e < Ne; ++e) {
                int lay = center[e + 2 * Ne] > zmid ? 1 : -1;
                double d3[3] = {center[e] - center[e0], center[e + Ne] - center[e0 + Ne],
                                center[e + 2 * Ne] - center[e0 + 2 * Ne]};
                double t = norm3(d3) / spread;
                double w = exp(-(t * t)) * (lay == lay0 ? 1.0 : 0.0);
                outs[s][e] = w;
                sum += w;
            }
            for (int e = 0; e < Ne; ++e) outs[s][e] /= sum;
        }
    }
    for (int fi = 0; fi < Nf; ++fi) {
        double omega = 2 * PI_D * freqs[fi];
        cx k = cscale(csqrt_p(cmul(cmk(mu, 0.0), cmk(eps_re, eps_im))), omega);
        cx alpha = cmk(0.0, omega * mu);
        cx beta = cdiv(cmk(1.0, 0.0), cmul(cmk(0.0, omega), cmk(eps_re, eps_im)));
        if (centroid)
            orc_assemble_z_centroid(P, k.re, k.im, alpha.re, alpha.im, beta.re, beta.im, self_alpha, Z);
        else
            orc_assemble_z_rows(P, k.re, k.im, omega, mu, eps_re, eps_im, o, rows, Ne, threads, Z);
        memcpy(b, b0, sizeof(double) * 2 * (size_t)Ne);
        int info = orc_zgetrf(Ne, Z, ipiv);
        if (info) { rc = -3; goto done; }
        orc_zgetrs(Ne, Z, ipiv, b);
        cx Iport;
        if (centroid) {
            Iport = csub(cmk(b[2 * (patchEdge - 1)], b[2 * (patchEdge - 1) + 1]),
                         cmk(b[2 * (groundEdge - 1)], b[2 * (groundEdge - 1) + 1]));
        } else {
            /* sum((wP.*I).*len) - sum((wG.*I).*len)  (main_patch_efie_rwg.m:336-338) */
            cx ip = cmk(0, 0), ig = cmk(0, 0);
            for (int e = 0; e < Ne; ++e) {
                cx I = cmk(b[2 * e], b[2 * e + 1]);
                ip = cadd(ip, cscale(cscale(I, wP[e]), P->len[e]));
                ig = cadd(ig, cscale(cscale(I, wG[e]), P->len[e]));
            }
            Iport = csub(ip, ig);
        }
        cx zin = cdiv(cmk(V0, 0.0), Iport);
        cx s11 = cdiv(csub(zin, cmk(Z0, 0)), cadd(zin, cmk(Z0, 0)));
        Zin[2 * fi] = zin.re; Zin[2 * fi + 1] = zin.im;
        S11[2 * fi] = s11.re; S11[2 * fi + 1] = s11.im;
    }
done:
    free(Z); free(b); free(b0); free(wP); free(wG); free(ipiv); free(rows);
    return rc;
}
