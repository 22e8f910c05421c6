/*
 * efie.h -- C ABI of libefie_hip.so, the MI355X (gfx950) implementation of
 * the method-of-moments EFIE hot path of mahmut26/Matlab-cuda-efief-vs-cpu:
 * RWG impedance-matrix fill Z, port excitation V, per-frequency complex LU
 * solve and the Zin/S11 epilogue of the microstrip-patch sweep.
 *
 * Plain C types only.  Conventions are the reference MEX conventions:
 *   - index tables are 1-based int32 (0 in plusTri/minusTri means "no
 *     triangle", patch_sweep_fastquad_cuda.cu:375,379),
 *   - matrices are column-major ([Nv x 3] V, [Nt x 3] F, [Ne x 2] edge,
 *     [Ne x 3] rp/rm, [Ne x Ne] Z),
 *   - complex numbers are interleaved (re, im) doubles.
 * Every entry point returns an efie_status; on failure efie_last_error()
 * and efie_last_error_id() give the message and the reference MEX error
 * identifier ("mex:port", "mex:rhs", ...), and no output is written.
 * A context is bound to one HIP device and one stream; calls on one context
 * are not re-entrant across threads.
 */
#ifndef EFIE_H
#define EFIE_H

#ifdef __cplusplus
extern "C" {
#endif

#define EFIE_ABI_VERSION 1

typedef enum {
    EFIE_OK = 0,
    EFIE_E_ARGS = 1,     /* bad argument (sizes, ranges, options)            */
    EFIE_E_MESH = 2,     /* mesh / RWG tables inconsistent                   */
    EFIE_E_PORT = 3,     /* port edge out of range                           */
    EFIE_E_RHS = 4,      /* spread radius leaves a layer without weight      */
    EFIE_E_SINGULAR = 5, /* exactly zero pivot in the LU                     */
    EFIE_E_HIP = 6,      /* HIP runtime error                                */
    EFIE_E_NOMEM = 7,    /* device or host allocation failed                 */
    EFIE_E_STATE = 8     /* mesh or RWG tables not set on the context        */
} efie_status;

enum { EFIE_MODEL_QUADRATURE = 0, EFIE_MODEL_CENTROID = 1 };
enum { EFIE_PREC_FP64 = 0, EFIE_PREC_MIXED = 1 };

/* Numerical options.  The first seven fields carry the names and meaning of
 * the reference `opts` struct (main_patch_efie_rwg.m:281-289); quad_order 7
 * selects a 7-point degree-5 rule that the reference does not have.
 * model CENTROID reproduces the reference kernel_Z_centroid
 * (patch_sweep_cuda.cu:94-153) and its Iport = x[p]-x[g] epilogue. */
typedef struct {
    int quad_order;
    double near_sing_eps;
    int self_subdiv;
    double self_alpha;
    int near_subdiv;
    double near_alpha;
    int near_order;
    int model;          /* EFIE_MODEL_*                                       */
    int precision;      /* EFIE_PREC_*                                        */
    int batch;          /* frequencies factored together (0 = automatic: a   *
                         * batch of 64-256 matrices, more for smaller n, the  *
                         * buffers within 60% of free memory)                 */
    double ir_tol;      /* mixed: stop refinement at ||V-Zx||/||V|| <= ir_tol  *
                         * or when a correction no longer halves the residual */
    int ir_max_iter;    /* mixed: refinement iteration cap                    */
    int near_mode;      /* EFIE_NEAR_*: how near and self pairs are integrated */
    int profile;        /* 1: time every kernel class into efie_stats          */
    int lu_kernel;      /* EFIE_LU_*: matrix-core form of the FP32 LU update   */
} efie_opts;

/* lu_kernel (mixed precision only): FP32 runs the pivoted LU's trailing
 * update on the FP32 matrix cores; BF16X6 splits every FP32 operand into
 * three bf16 parts and accumulates six part products in FP32 on the bf16
 * matrix cores (experimental); SYM uses that Z is complex symmetric: the
 * FP32 factorisation runs without pivoting as A = L D L^T, every step
 * reading the lower triangle only (U = D L^T), so the updates cover only the
 * lower triangle (about half the flops) and the pivot search and row
 * interchanges disappear; the FP64 refinement corrects its larger rounding
 * growth.  A frequency that does not reach max(ir_tol, 1e-11) (a zero or
 * tiny pivot) is solved again with the pivoted FP32 factorisation
 * (efie_stats.n_lu_retry), and one the pivoted FP32 forms do not refine to
 * that bound with the FP64 LU (efie_stats.n_fp64_retry), so every result of
 * a mixed sweep is refined to max(ir_tol, 1e-11) or comes from the FP64 LU.
 * AUTO picks SYM in efie_sweep (Z is complex symmetric) and FP32 in
 * efie_zgesv (general matrices). */
enum { EFIE_LU_AUTO = 0, EFIE_LU_FP32 = 1, EFIE_LU_BF16X6 = 2, EFIE_LU_SYM = 3 };

/* near_mode: DIRECT evaluates every triangle pair's quadrature (regular,
 * subdivided near, subdivided self) at every frequency as the reference
 * does; MOMENTS expands the Green's function of every pair about its
 * centroid distance R0, exp(-jkR)/R = exp(-jkR0) sum_q (-jk)^q (R-R0)^q /
 * (q! R), integrates the frequency-independent coefficients of the same
 * quadrature rules once per mesh and forms Z of a whole batch of
 * frequencies from them in one pass (series tail below 1e-17 relative);
 * when |k| times the pair size needs more than 22 terms, or the series
 * would not fit in a third of the free device memory, the regular pairs
 * are integrated per frequency and only the near and self pairs use a
 * moment series (falling back to DIRECT when that is too long too).
 * EXTRACT replaces the
 * reference's softened scheme on near and self pairs by singularity
 * extraction: the static 1/(4 pi R) part of the inner integral in closed
 * form over the whole source triangle, the smooth remainder
 * (exp(-jkR) - 1)/(4 pi R) by the subdivided near/self rule (as a moment
 * series).  It has no counterpart in the reference.  SURROGATE (declared
 * with its training entry points in efie_surrogate.h) replaces the
 * subdivision quadrature of near pairs by the regular product rule plus a
 * learned correction, with the quadrature as fallback outside the trained
 * envelope; self pairs as MOMENTS. */
enum { EFIE_NEAR_DIRECT = 0, EFIE_NEAR_MOMENTS = 1, EFIE_NEAR_EXTRACT = 2, EFIE_NEAR_SURROGATE = 3 };

/* Kernel classes timed when efie_opts.profile is set. */
enum {
    EFIE_K_REGULAR = 0,  /* regular triangle-pair records                     */
    EFIE_K_NEAR = 1,     /* near-pair records (direct quadrature)             */
    EFIE_K_SELF = 2,     /* self-pair records (direct quadrature)             */
    EFIE_K_POLY = 3,     /* near+self records from moments                    */
    EFIE_K_COMBINE = 4,  /* Z(m,n) from records                               */
    EFIE_K_TSLU = 5,     /* LU panel pivot selection                          */
    EFIE_K_LASWP = 6,    /* LU row interchanges                               */
    EFIE_K_PANEL = 7,    /* LU panel triangular solves                        */
    EFIE_K_GEMM = 8,     /* LU trailing update (matrix cores)                 */
    EFIE_K_TRSV = 9,     /* triangular solves                                 */
    EFIE_K_OTHER = 10,   /* conversions, residuals, epilogue                  */
    EFIE_K_SURROGATE = 11, /* near-pair network corrections                   */
    EFIE_K_SERIES = 12,  /* Z of a batch from the all-pairs series (MOMENTS)  */
    EFIE_NKCLASS = 13
};

/* Per-call timings (HIP events on the context stream) and counters. */
typedef struct {
    double ms_total;        /* whole call, device work only                  */
    double ms_fill;         /* Z fill (all fill kernels)                     */
    double ms_factor;       /* LU factorisation                              */
    double ms_solve;        /* triangular solves + refinement + epilogue     */
    long long n_near_pairs; /* unordered near triangle pairs                 */
    long long n_green;      /* Green-function evaluations per frequency      */
    int ir_iters_max;       /* mixed: most refinement iterations used         */
    double ir_resid_max;    /* mixed: largest final ||V-Zx||/||V||            */
    int moment_terms;       /* series length P+1 used by MOMENTS (0: direct)  */
    double ms_moments;      /* one-off moment integration (when it ran)       */
    /* profile == 1: summed device time and launch count per kernel class  */
    double ms_kernel[EFIE_NKCLASS];
    int n_kernel[EFIE_NKCLASS];
    /* near_mode SURROGATE: near pairs given by the network, and those
     * integrated by the quadrature because they left its envelope        */
    long long n_sur_pairs;
    long long n_sur_fallback;
    /* lu_kernel SYM: frequencies solved again with pivoting because the
     * unpivoted factorisation met a zero pivot or did not refine to
     * max(ir_tol, 1e-11)                                                  */
    int n_lu_retry;
    /* mixed precision: systems solved again by the FP64 LU because the
     * pivoted FP32 factorisation met a zero pivot or its refinement did not
     * reach max(ir_tol, 1e-11) (stalled or at ir_max_iter); their results
     * have the FP64 path's accuracy                                        */
    int n_fp64_retry;
    /* mixed precision: refinement corrections summed over every system
     * (ir_iters_total / frequencies is the average per system)             */
    long long ir_iters_total;
} efie_stats;

/* "accurate" preset (main_patch_efie_rwg.m:281-289), FP64, quadrature. */
void efie_opts_default(efie_opts* o);
/* name: "accurate", "fast" (main_patch_gpu_mex_sweep_l.m:50-56),
 * "centroid" (the reference MEX centroid model). */
efie_status efie_opts_preset(efie_opts* o, const char* name);

const char* efie_last_error(void);
const char* efie_last_error_id(void);

/* ---- host-side helpers: the MATLAB functions that build the MEX inputs -- */

/* make_patch_with_ground_mesh.m:1-46 */
efie_status efie_mesh_counts(int npx, int npy, int ngx, int ngy, int* Nv, int* Nt);
efie_status efie_make_mesh(double Lp, double Wp, double Lg, double Wg, double h,
                           int npx, int npy, int ngx, int ngy,
                           double* V, int* F);
/* build_rwg_interior.m:1-113 (two calls: count, then fill) */
efie_status efie_rwg_count(const int* F, int Nt, int Nv, int* Ne);
efie_status efie_build_rwg(const double* V, int Nv, const int* F, int Nt, int Ne,
                           int* edge, int* plusTri, int* minusTri,
                           int* plusSign, int* minusSign, double* rp,
                           double* rm, double* Ap, double* Am, double* len,
                           double* center, int* zSign);
/* pick_port_edge_by_point_and_layer.m:1-11; *edge is 1-based */
efie_status efie_pick_port(int Ne, const double* center, const int* zSign,
                           const double p[3], int layerSign, int* edge);

/* ---- device context ------------------------------------------------------ */

typedef struct efie_ctx efie_ctx;

efie_status efie_create(efie_ctx** ctx, int device);
void efie_destroy(efie_ctx* ctx);

/* ---- the same preprocessing on the device, from a context's mesh -------- */

/* build_rwg_interior.m:1-113 on the device from the mesh given to
 * efie_set_mesh: *Ne = the number of interior edges. */
efie_status efie_rwg_count_device(efie_ctx* ctx, int* Ne);
/* build_rwg_interior.m:1-113 on the device: the tables of efie_build_rwg
 * (bit-identical to it), copied to the host arrays and installed as the
 * context's RWG tables, as efie_set_rwg would. */
efie_status efie_build_rwg_device(efie_ctx* ctx, int Ne, int* edge, int* plusTri,
                                  int* minusTri, int* plusSign, int* minusSign,
                                  double* rp, double* rm, double* Ap, double* Am,
                                  double* len, double* center, int* zSign);
/* pick_port_edge_by_point_and_layer.m:1-11 on the device, over the edge
 * centres and layer signs of the last efie_build_rwg_device; *edge is
 * 1-based (the lowest index among equal distances, as efie_pick_port). */
efie_status efie_pick_port_device(efie_ctx* ctx, const double p[3], int layerSign,
                                  int* edge);

/* Mesh and RWG tables: arguments 1-12 of both reference MEX functions
 * (patch_sweep_cuda.cu:4-16, gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:127-139).
 * Copied to HBM once; frequency-independent geometry (quadrature points,
 * near-pair lists) is derived lazily and cached until the tables or the
 * geometry options change. */
efie_status efie_set_mesh(efie_ctx* ctx, const double* V, int Nv, const int* F, int Nt);
efie_status efie_set_rwg(efie_ctx* ctx, const int* edge, const int* plusTri,
                         const int* minusTri, const int* plusSign,
                         const int* minusSign, const double* rp,
                         const double* rm, const double* Ap, const double* Am,
                         const double* len, int Ne);

/* [Z, Vrhs] of gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST (arguments 13-17,
 * gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:140-148): k, omega, mu, eps and
 * exc = [patchEdge, groundEdge, V0, spreadRadius, selfAlpha].  Z is
 * [Ne x Ne] column-major interleaved complex, Vrhs [Ne] interleaved complex
 * (imaginary parts 0).  Z is computed with the quadrature scheme of
 * assemble_ZV_EFIE_RWG_DIFF_CPU.m (opts), or the centroid model when
 * opts->model is CENTROID (then exc[4] is the self-softening factor). */
efie_status efie_assemble_zv(efie_ctx* ctx, double k_re, double k_im,
                             double omega, double mu, double eps_re,
                             double eps_im, const double* exc, int n_exc,
                             const efie_opts* opts, double* Z, double* Vrhs);

/* [Zin, S11] of patch_sweep_cuda (arguments 13-20, patch_sweep_cuda.cu:17-27):
 * per frequency k = omega*sqrt(mu*eps), fill Z, LU-solve Z x = V, then
 * Iport = sum((wPatch - wGnd) .* x .* len) (main_patch_efie_rwg.m:336-338),
 * Zin = V0/Iport, S11 = (Zin - Z0)/(Zin + Z0).  Zin, S11: [Nf] interleaved
 * complex.  stats may be NULL. */
efie_status efie_sweep(efie_ctx* ctx, const double* freqs, int Nf, double mu,
                       double eps_re, double eps_im, int patchEdge,
                       int groundEdge, double V0, double spread, double Z0,
                       const efie_opts* opts, double* Zin, double* S11,
                       efie_stats* stats);

/* ---- many sweeps in one call (geometries of uneven size) ----------------- */

/* One sweep of efie_sweep_many: the context holding its problem (mesh and
 * RWG tables set by efie_set_mesh + efie_set_rwg, on the engine's device),
 * arguments 13-20 of patch_sweep_cuda (patch_sweep_cuda.cu:17-27) but for
 * the shared mu, eps, and the two outputs. */
typedef struct {
    efie_ctx* problem;
    const double* freqs;
    int Nf;
    int patchEdge, groundEdge;
    double V0, spread, Z0;
    double* Zin;   /* [Nf] interleaved complex                               */
    double* S11;   /* [Nf] interleaved complex                               */
    int release;   /* 1: the problem's frequency-independent series
                    * coefficients (near_mode MOMENTS) are freed once its last
                    * batch is filled (bounded memory over many problems)     */
} efie_sweep_item;

/* No counterpart in the reference, whose drivers sweep one geometry per
 * patch_sweep_cuda call and solve each frequency alone (fill, sync,
 * Zgetrf, sync: patch_sweep_fastquad_cuda.cu:713-772).  The sweeps of
 * `count` problems, each with the results and error behaviour of
 * efie_sweep(items[i].problem, ...) under the shared mu, eps and opts, run
 * as one pipeline on the engine context's streams and buffers:
 *   - the frequencies of all problems are cut into batches; problems of
 *     similar order share a batch, their matrices padded to the largest
 *     order with the identity (the padded system's solution is the
 *     unpadded one exactly; the factorisation skips the padding);
 *   - the next batch is filled while one is factored and the one before is
 *     refined, also across problems;
 *   - a problem's one-off work (geometry, near lists, series) runs while
 *     earlier batches are factored.
 * The engine need not hold a problem itself; one problem context may
 * appear in several items.  The call uses the problem contexts (their
 * streams and caches): no other thread may use them meanwhile.  stats
 * (may be NULL) sums the whole call. */
efie_status efie_sweep_many(efie_ctx* engine, const efie_sweep_item* items, int count,
                            double mu, double eps_re, double eps_im,
                            const efie_opts* opts, efie_stats* stats);

/* ---- precomputed quadrature caches (the prepack struct) ------------------ */

/* The quadrature caches of efie_precompute.m:1-97 as packed by
 * efie_precompute_pack.m:1-59 (the `prepack` argument of
 * patch_sweep_fastquad_cuda.cu:463-557, ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu
 * :542-641 and assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:277-392), in MATLAB's
 * column-major layouts:
 *   r_reg [Nq x 3 x Nt], w_reg [Nq x Nt]   regular points, weights x 2A
 *   nearMat [Nt x Nt] (uint8)              near pairs (the diagonal is not read)
 *   self_r [Nq2 x 3 x NsS x Nt], self_w [Nq2 x NsS x Nt], self_a [NsS x Nt]
 *   near_r [Nq2 x 3 x NsN x Nt], near_w [Nq2 x NsN x Nt]
 *   aeqTri [Nt]                            sqrt(A/pi), near softening length
 * The kernels support Nq and Nq2 in {1, 3, 6, 7} and NsS, NsN in
 * {1, 4, 16, 64}; nearMat must be symmetric. */
typedef struct {
    int Nq, Nq2, NsS, NsN;
    const double* r_reg;
    const double* w_reg;
    const unsigned char* nearMat;
    const double* self_r;
    const double* self_w;
    const double* self_a;
    const double* near_r;
    const double* near_w;
    const double* aeqTri;
    double near_alpha;
} efie_prepack;

/* Install the caches on the context's problem (set by efie_set_mesh +
 * efie_set_rwg, or efie_set_rwg_tables): the quadrature-model fills that
 * follow (efie_assemble_zv, efie_assemble_z, efie_sweep) integrate with
 * these points, weights, softening lengths and near pairs, and take
 * near_alpha from the pack; the rule options quad_order, near_order,
 * near_subdiv, self_subdiv, self_alpha, near_sing_eps and near_alpha of
 * efie_opts are not used.  near_mode DIRECT or MOMENTS.  pk = NULL returns
 * to the rule-derived tables; new tables (efie_set_mesh, efie_set_rwg,
 * efie_set_rwg_tables) drop the pack. */
efie_status efie_set_prepack(efie_ctx* ctx, const efie_prepack* pk);

/* A problem without the mesh arrays, as assemble_ZV_EFIE_RWG_DIFF_GPU_FAST
 * (8 inputs, .cu:277-339) receives it: the RWG tables of the rwg struct
 * and Nt triangles.  Only efie_assemble_z with a prepack can fill it: the
 * triangles' centroids and areas come from the pack's regular points
 * (sum w r / sum w and 2 sum w: the rules integrate linear functions
 * exactly and their weights sum to 1/2). */
efie_status efie_set_rwg_tables(efie_ctx* ctx, int Nt, const int* plusTri, const int* minusTri,
                                const int* plusSign, const int* minusSign, const double* rp,
                                const double* rm, const double* Ap, const double* Am,
                                const double* len, int Ne);

/* Z alone (no excitation) for k, omega, mu, eps: the Z output of
 * assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:483-560.  Z is [Ne x Ne]
 * column-major interleaved complex; quadrature model only. */
efie_status efie_assemble_z(efie_ctx* ctx, double k_re, double k_im, double omega, double mu,
                            double eps_re, double eps_im, const efie_opts* opts, double* Z);

/* Dense complex solve A x = b for `batch` independent systems (A: n x n
 * column-major interleaved complex, batch after batch; b, x: n each), the
 * cusolverDnZgetrf + cusolverDnZgetrs pair of the reference
 * (patch_sweep_cuda.cu:473-480).  precision FP64: blocked LU with
 * tournament pivoting on the matrix cores; MIXED: complex FP32 LU plus FP64
 * iterative refinement (ir_tol 1e-14, at most 16 corrections); a system
 * whose FP32 factorisation meets a zero pivot or whose refinement stalls
 * above 1e-11 is solved again by the FP64 LU (efie_stats.n_fp64_retry), and
 * only a zero pivot there returns EFIE_E_SINGULAR. */
efie_status efie_zgesv(efie_ctx* ctx, int n, int batch, const double* A,
                       const double* b, int precision, double* x,
                       efie_stats* stats);

/* No counterpart in the reference.  efie_zgesv for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments, results, error
 * behaviour and FP64 fallback as efie_zgesv; it calls no hipMalloc,
 * hipFree or other allocating routine: every device buffer of the solve is
 * carved from one static device array of EFIE_NOALLOC_SCRATCH_BYTES per
 * GPU, which each GPU the library runs on holds until the process ends.
 * The limit: batch <= 64 and the solve's buffers (about 16 n (n + 32)
 * batch bytes for FP64, 24 n (n + 32) batch for MIXED, plus O(n batch))
 * within that array, so n up to about 1900 for one FP64 system; larger
 * calls return EFIE_E_ARGS.  All calls on one GPU share the array and it
 * keeps the last call's contents (the solve clears what must start at
 * zero): calls that overlap on one GPU, from any context, race. */
#define EFIE_NOALLOC_SCRATCH_BYTES (64u << 20)
efie_status efie_zgesv_noalloc(efie_ctx* ctx, int n, int batch, const double* A,
                               const double* b, int precision, double* x,
                               efie_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
