/*
 * efie_oracle.h -- CPU FP64 restatement of the reference EFIE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP product is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may use it.
 *
 * Conventions follow the reference MATLAB code: every matrix is column-major,
 * every index table is 1-based int32, complex numbers are interleaved
 * (re, im) pairs of doubles.
 */
#ifndef EFIE_ORACLE_H
#define EFIE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Numerical options; field names and defaults follow the reference `opts`
 * struct (rwg_pair_integrals_near_singular.m:3-5, rwg_pair_integrals_singular.m:7-9,
 * main_patch_efie_rwg.m:281-289). quad_order: 1 -> 1 point, 2 -> 3 points,
 * 7 -> 7-point degree-5 rule (not in the reference), anything else -> the
 * reference's 6-point rule (tri_quad_rule.m:3-22). */
typedef struct {
    int quad_order;
    double near_sing_eps;
    int self_subdiv;
    double self_alpha;
    int near_subdiv;
    double near_alpha;
    int near_order;
    /* 1: self and near pairs by singularity extraction (orc_pair_extract)
     * instead of the reference's softened subdivision; not in the
     * reference (its Z is only defined for extract = 0). */
    int extract;
} orc_opts;

/* Mesh sizes of make_patch_with_ground_mesh.m:1-46. */
void orc_mesh_counts(int npx, int npy, int ngx, int ngy, int* Nv, int* Nt);
/* make_patch_with_ground_mesh.m:1-46: V [Nv x 3], F [Nt x 3] (1-based). */
void orc_make_mesh(double Lp, double Wp, double Lg, double Wg, double h,
                   int npx, int npy, int ngx, int ngy, double* V, int* F);

/* Number of interior edges (RWG unknowns) of build_rwg_interior.m:25-35.
 * Returns -1 if a triangle index is out of range. */
int orc_rwg_count(const int* F, int Nt, int Nv);
/* build_rwg_interior.m:1-113.  Returns 0, or -1 when an opposite vertex is
 * missing (build_rwg_interior.m:108-110). center [Ne x 3], zSign [Ne]. */
int orc_build_rwg(const double* V, int Nv, const int* F, int Nt, int Ne,
                  int* edge, int* plusTri, int* minusTri, int* plusSign,
                  int* minusSign, double* rp, double* rm, double* Ap,
                  double* Am, double* len, double* center, int* zSign);

/* tri_quad_rule.m:1-23 (+ the 7-point rule).  qp [Nq x 2] col-major. */
int orc_quad_rule(int order, double* qp, double* qw);

/* triangles_are_near.m:1-12 (1-based triangle indices). */
int orc_triangles_are_near(const double* V, int Nv, const int* F, int Nt,
                           int t1, int t2, double eps_scale);

/* One triangle-pair integral [Ivec, Isca] as rwg_pair_integrals_regular.m,
 * rwg_pair_integrals_near_singular.m or rwg_pair_integrals_singular.m
 * (kind 0,1,2). m, n, t1, t2 are 1-based.  out[4] = Ivec re/im, Isca re/im. */
typedef struct {
    const double* V; int Nv;
    const int* F; int Nt;
    int Ne;
    const int* plusTri; const int* minusTri;
    const int* plusSign; const int* minusSign;
    const double* rp; const double* rm;
    const double* Ap; const double* Am; const double* len;
} orc_problem;

/* Static potentials of a flat triangle (v1, v2, v3) at the point r, in
 * closed form (Wilton, Rao, Glisson, Schaubert, Al-Bundak and Butler, IEEE
 * TAP 32(3), 1984; Graglia, IEEE TAP 41(10), 1993):
 *   out[0]   = int_T 1/|r - r'| dS'
 *   out[1:4] = int_T (r' - rho)/|r - r'| dS'  (rho = projection of r onto
 *              the triangle's plane; the vector lies in that plane).
 * Not in the reference.  Returns 0. */
int orc_tri_potentials(const double* r, const double* v1, const double* v2, const double* v3,
                       double* out);

/* kind 3 of orc_pair_integral: singularity extraction for a self or near
 * pair, G = 1/(4 pi R) + (exp(-jkR) - 1)/(4 pi R).  The outer integral runs
 * over the lower-numbered triangle, subdivided 4^depth times (self_subdiv
 * for t1 == t2, near_subdiv otherwise) with the near_order rule; the inner
 * static part is orc_tri_potentials over the whole other triangle, the
 * inner smooth part the same subdivided rule.  Not in the reference. */

int orc_pair_integral(const orc_problem* P, int kind, int m, int n, int t1,
                      int t2, double k_re, double k_im, const orc_opts* o,
                      double* out);

/* assemble_ZV_EFIE_RWG_DIFF_CPU.m:1-36 restricted to the listed (1-based)
 * rows: Zrows[r + nrows*(n-1)] = Z(rows[r], n), interleaved complex.
 * Rows run in parallel over `threads` OpenMP threads (the reference's
 * parfor-over-rows variant, assemble_ZV_EFIE_RWG_DIFF_CPUp.m:12). */
int orc_assemble_z_rows(const orc_problem* P, double k_re, double k_im,
                        double omega, double mu, double eps_re, double eps_im,
                        const orc_opts* o, const int* rows, int nrows,
                        int threads, double* Zrows);

/* Centroid one-point fill of the reference MEX kernel_Z_centroid
 * (patch_sweep_cuda.cu:94-153 with the host precompute at :244-326). */
int orc_assemble_z_centroid(const orc_problem* P, double k_re, double k_im,
                            double alpha_re, double alpha_im, double beta_re,
                            double beta_im, double self_alpha, double* Z);

/* RHS of assemble_ZV_EFIE_RWG_DIFF_CPU.m:43-71 (uses rwg.center/zSign).
 * Returns -1 when a layer has no weight (the reference's error). */
int orc_rhs(int Ne, const double* center, const int* zSign, int patchEdge,
            int groundEdge, double V0, double spread, double* Vrhs);

/* pick_port_edge_by_point_and_layer.m:1-11; returns 1-based index or -1. */
int orc_pick_port(int Ne, const double* center, const int* zSign,
                  const double* p, int layerSign);

/* port_weights_gaussian.m:1-29. */
int orc_port_weights(const double* V, int Nv, const int* edge, int Ne,
                     int patchEdge, int groundEdge, double spread,
                     double* wPatch, double* wGnd);

/* LAPACK-style zgetrf/zgetrs with partial pivoting on |re|+|im| (the pivot
 * rule of izamax), A column-major interleaved complex, ipiv 1-based.
 * Returns info (0 ok, j>0 zero pivot at column j). */
int orc_zgetrf(int n, double* A, int* ipiv);
void orc_zgetrs(int n, const double* LU, const int* ipiv, double* b);

/* Frequency sweep of main_patch_efie_rwg.m:314-341 (quadrature fill, solve,
 * Zin) with S11 as patch_sweep_fastquad_cuda.cu:793-799.  centroid != 0
 * selects the centroid MEX model with Iport = x[p] - x[g]
 * (patch_sweep_cuda.cu:483-494).  Zin, S11 interleaved [Nf]. */
int orc_sweep(const orc_problem* P, const double* center, const int* zSign,
              const double* freqs, int Nf, double mu, double eps_re,
              double eps_im, int patchEdge, int groundEdge, double V0,
              double spread, double Z0, const orc_opts* o, int centroid,
              double self_alpha, int threads, double* Zin, double* S11);

#ifdef __cplusplus
}
#endif
#endif
