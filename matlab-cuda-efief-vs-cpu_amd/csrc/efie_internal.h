// efie_internal.h -- device data layout shared by the host orchestration
// (efie_api.cpp) and the kernels (efie_fill.hip, efie_lu.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace efie {

// Quadrature point relative to its triangle's centroid, with its full
// weight (rule weight times twice the (sub-)triangle area).
struct Pt {
    double x, y, z, w;
};

// Per-triangle data: centroid and twice the area (precompute_tri_geom.m).
struct TriGeo {
    double cx, cy, cz, area2;
};

// One side of an RWG basis function restricted to one triangle:
// f(r) = c * (r - v) and div f = 2c with c = +s*l/(2A) on the plus
// triangle and -s*l/(2A) on the minus triangle (rwg_eval_on_triangle.m).
// u = v - centroid(tri).
struct Side {
    double ux, uy, uz, c;
};

// Moments of one unordered triangle pair {t1 <= t2} for one frequency,
// with rho = r - centroid(t1), rho' = r' - centroid(t2):
//   S0 = sum W G,  S1 = sum W G rho,  S1p = sum W G rho',
//   S2 = sum W G (rho . rho').
// Stored as 16 doubles in the order S0, S1x, S1y, S1z, S1px, S1py, S1pz, S2
// (re, im interleaved); 128 bytes, one cache line.
struct Rec {
    double v[16];
};

// Packed upper-triangular index of the record of pair (t1 <= t2),
// column-major over t2 so that consecutive t1 are consecutive records.
__host__ __device__ inline size_t rec_index(int t1, int t2) {
    return (size_t)t2 * (size_t)(t2 + 1) / 2 + (size_t)t1;
}

// Scalars of one frequency.
struct Freq {
    double kr, ki;       // k = kr + j ki
    double ar, ai;       // alpha = j omega mu
    double br, bi;       // beta = 1/(j omega eps)
    double hr, hi;       // -alpha / (4 pi k^2): the series fill's factor
};

// --------------------------------------------------------- launchers ----
// All launchers enqueue on `s` and return hipGetLastError().

// Near pairs (triangles_are_near.m) by cell binning: cubic cells of side h
// (>= eps * max(l)) from (x0, y0, z0), nx * ny * nz of them; cellOf[t] the
// cell of triangle t, cellTri the triangles cell by cell from
// cellStart[cell] (ncell + 1 offsets).  list == nullptr: rowCount[t1] =
// near t2 > t1; otherwise those pairs, sorted by t2, to list[rowStart[t1]..].
// RWG tables on the device (efie_rwg.hip).  Count pass: cnt/fill/nInt
// [Nv], start/eStart [Nv + 1] (exclusive scans; eStart[Nv] = Ne), slot
// [3 Nt].  Tables pass: column-major outputs with leading dimension Ne
// (edge, tri = plus|minus, sgn = plus|minus, rpm = rp|rm, Apm = Ap|Am,
// len, center, zSign); *err = lowest 1-based edge without a unique
// opposite vertex, INT_MAX when there is none.
hipError_t launch_rwg_count(const int* F, int Nt, int Nv, int* cnt, int2* slot, int* fill, int* nInt, int* start,
                            int* eStart, hipStream_t s);
hipError_t launch_rwg_tables(const double* V, int Nv, const int* F, int Nt, const int* start, const int2* slot,
                             const int* eStart, int Ne, double zmid, int* edge, int* tri, int* sgn, double* rpm,
                             double* Apm, double* len, double* center, int* zSign, int* err, hipStream_t s);
// Port pick over the edges of one layer: *out = 1-based nearest edge (lowest
// index among equal distances), 0 when the layer has none; pd/pi hold
// port_pick_blocks(Ne) partials.
int port_pick_blocks(int Ne);
hipError_t launch_port_pick(const double* center, const int* zSign, int Ne, const double p[3], int layer, double* pd,
                            int* pi, int* out, hipStream_t s);

// Quadrature rules of the rule-derived geometry (reference coordinates
// u, v and weights w): the regular rule and the rule on subdivided
// triangles.
struct QuadRules {
    double ru[16], rv[16], rw[16], su[16], sv[16], sw[16];
    int nReg, nSub;
};
// Points of every triangle relative to its centroid from the absolute
// vertices tv (9 per triangle): regular [Nt][nReg], near children
// [Nt][4^depthN][nSub], self children [Nt][4^depthS][nSub], the self
// softening radii [Nt][4^depthS], rad[t] and mx[4] = {near radius, self
// radius, max softening radius, max rad}.
hipError_t launch_quad_points(const double* tv, const TriGeo* tg, int Nt, const QuadRules& R, int depthN,
                              int depthS, double selfAlpha, Pt* regPts, Pt* nearPts, Pt* selfPts, double* aSelf,
                              double* rad, double* mx, hipStream_t s);
hipError_t launch_near_bins(const TriGeo* tg, int Nt, double eps, double x0, double y0, double z0, double h, int nx,
                            int ny, int nz, const int* cellOf, const int* cellStart, const int* cellTri,
                            int* rowCount, const int* rowStart, int2* list, hipStream_t s);
hipError_t launch_near_list(const uint8_t* nearMat, const int* rowStart, int Nt,
                            int2* list, hipStream_t s);

// Regular (non-near, t1 < t2) records; near pairs are those of nearMat
// when it is given, else those the near test with eps finds.
// rad[t]: largest distance from the centroid of t to its vertices (bounds
// every quadrature point of t).
hipError_t launch_rec_regular(const Pt* regPts, int Nq, const TriGeo* tg, const double* rad, int Nt,
                              const uint8_t* nearMat, double eps, const Freq* fr, Rec* recs,
                              hipStream_t s);
// Subdivided near-pair records (rwg_pair_integrals_near_singular.m).
hipError_t launch_rec_near(const Pt* subPts, int Np, const TriGeo* tg, const double* rad,
                           const double* aeq, double near_alpha, const int2* list, int nPairs,
                           const Freq* fr, Rec* recs, hipStream_t s);
// Self records (rwg_pair_integrals_singular.m): Np points per triangle in
// Ns sub-triangles of Nq2 points; aSelf[t*Ns + p] is the softening length.
hipError_t launch_rec_self(const Pt* selfPts, int Nq2, int Ns, const double* aSelf,
                           const double* rad, double aSelfMax, int Nt, const Freq* fr, Rec* recs,
                           hipStream_t s);

// Frequency-independent moments Q[pair][q][8], q = 0..P, of the near pairs
// (Qn) and the self pairs (Qs), and the per-frequency records built from
// them (near pairs and self pairs).
// With triV (vertices relative to the centroid, [Nt][9]) non-null: the
// singularity-extraction moments (unsoftened q >= 1 tail plus the
// closed-form static q = 0 slot).
hipError_t launch_moments(const Pt* subPts, int Np, const TriGeo* tg, const double* aeq,
                          double near_alpha, const int2* list, int nPairs, const Pt* selfPts,
                          int Nq2, int Ns, const double* aSelf, int Nt, int P, double* Qn,
                          double* Qs, const double* triV, hipStream_t s);
hipError_t launch_poly_records(const double* Qn, const int2* list, int nNear, const double* Qs,
                               int Nt, int P, const Freq* fr, Rec* recs, hipStream_t s);

// Frequency-independent series of every triangle pair {t1 <= t2} about its
// centroid distance R0 (0 for t1 = t2): Q[rec_index(t1,t2)][q][8] =
// sum W X (R - R0)^q / (q! R), q = 0..P (P + 1 <= 32), with the regular rule
// for the pairs that are not near (nearMat, or the near test with eps), the
// subdivided near rule for the nPairs pairs of list and the self rule for
// t1 = t2.
hipError_t launch_series(const Pt* regPts, int Nq, const TriGeo* tg, int Nt, const uint8_t* nearMat, double eps,
                         const Pt* subPts, int Np, int Nq2, const double* aeq, double near_alpha, const int2* list,
                         int nPairs, const Pt* selfPts, int Ns, const double* aSelf, int P, double* Q,
                         hipStream_t s);
// Number of series coefficients per entry (P + 3, rounded up to a size the
// fill kernel is built for: 10..20, even) or 0 if P is too large.
int series_nt(int P);
// Per-entry coefficients of the upper-triangular 16 x 16 tiles of Z from
// the pair series Q (P = NT - 3): NT coefficients and the entry's phase
// centre, series_coeff_bytes(Ne, NT) bytes.
size_t series_coeff_bytes(int Ne, int NT);
hipError_t launch_series_coeffs(const double* Q, int NT, const TriGeo* tg, const int* sideTri, const Side* side,
                                int Ne, double* C, hipStream_t s);
// Z of nf frequencies from the coefficients: matrix f at Z + f*zStride,
// element (m,n) at 2*(m*ld + n), and its complex64 copy at
// Z32 + f*z32Stride when Z32 is set.  lossy: some k has a nonzero
// imaginary part.  maxBlocks > 0 caps the grid (tiles in a grid-stride
// loop; a multiple of 8 keeps each XCD on its run of tiles).  tri (Z32
// required): only what the complex-symmetric mixed solve reads, the FP64
// matrix on and above the diagonal plus its diagonal 64 x 64 blocks (the
// residual) and the FP32 copy on and below the diagonal (the
// factorisation); the other entries keep what they held.
hipError_t launch_series_fill(const double* C, int NT, bool lossy, int Ne, const Freq* fr, int nf, double* Z,
                              size_t ld, size_t zStride, float* Z32, size_t ld32, size_t z32Stride, hipStream_t s,
                              int maxBlocks = 0, bool tri = false);
// blocks of a series fill that runs beside a factorisation
constexpr int kFillBlocksShared = 256;

// Z(m,n) = sum over the four side pairs of alpha*Ivec + beta*Isca.
// Output element (m,n) goes to Z[m*ld + n] (interleaved complex; Z is
// symmetric, so this is also the column-major matrix with leading
// dimension ld); with Z32 non-null also a complex64 copy at Z32[m*ld32 + n].
hipError_t launch_combine(const Rec* recs, const int* sideTri, const Side* side, int Ne,
                          const Freq* fr, double* Z, size_t ld, float* Z32, size_t ld32, hipStream_t s);

// Reference centroid model (kernel_Z_centroid).
hipError_t launch_centroid(const double* pos, const double* f, const double* dv,
                           const double* w, const int* tsd, const double* asoft, int Ne,
                           const Freq* fr, double* Z, size_t ldr, size_t ldc, hipStream_t s);

// Gaussian edge-strip weights on the edge midpoints of one layer,
// normalised (port_weights_gaussian.m): out[e] = w_e / sum(w), and
// sums[which] = sum(w).
hipError_t launch_port_weights(const double* mid, int Ne, int edge0, double spread,
                               double* out, double* sums, int which, hipStream_t s);
// V = +V0/2 wP - V0/2 wG (interleaved complex, imag 0) and
// wdiff = (wP - wG) .* len for the Zin epilogue.
hipError_t launch_rhs(const double* wP, const double* wG, const double* len, int Ne,
                      double V0, double* Vrhs, double* wdiff, hipStream_t s);

// --------------------------------------------------------------- LU -----
// Batched in-place LU with tournament (CALU) pivoting of nb batch matrices
// stored row-major with leading dimension lda (A[b*strideA + i*lda + j]),
// interleaved complex of type T (double or float).  Column n (< lda) holds
// the right-hand side and is carried through all row operations, so after
// the factorisation it holds L^{-1} P b.
struct LuWork {
    int* cand;      // tournament candidates
    int* counters;  // tree arrival counters
    int* ipiv;      // [batch][n] 1-based LAPACK-style pivots
    int* info;      // [batch] first zero pivot (1-based) or 0
    int* swaps;     // [batch][lu_swap_stride()]: row moves of the current block's panels
    // look-ahead: the next outer block's panels are factored on `panel`
    // while the trailing update runs on the main stream (panel may be the
    // main stream itself); the two events order them
    hipStream_t panel;
    hipEvent_t evPanel, evAhead;
    // FP32 factorisation: non-null = the trailing updates run on the bf16
    // matrix cores from a split of L21 and U12 (lu_split_bytes(n) per batch
    // element)
    void* splitPlanes;
    // complex-symmetric A: factor without pivoting, U = D L^T (lu_factor_sym)
    bool noPivot;
    // lu_factor_sym of a packed batch: matrix b holds a problem of order
    // nElem[b] in its leading block and the identity beyond (null: order n);
    // steps wholly inside the padding are skipped
    const int* nElem = nullptr;
    // lu_factor_sym with look-ahead: the first block's panels run on this
    // stream, which the caller has ordered after the matrices are complete
    // (the fill stream), so they can overlap the previous factorisation's
    // last steps on the main and panel streams; null: on the panel stream
    hipStream_t head = nullptr;
    // with a head stream: the first block's update (and the second block's
    // panels) also run there, after headWait (the previous factorisation's
    // tailMark: its last kTailCols columns have begun, the latency-bound
    // steps these fill); tailMark is recorded on the main stream at that
    // point of this factorisation (either may be null)
    hipEvent_t headWait = nullptr, tailMark = nullptr;
};
// columns of a factorisation's end that the next one's first update may
// overlap
constexpr int kTailCols = 2560;
size_t lu_split_bytes(int n);
// ints per batch element of LuWork::swaps
int lu_swap_stride();

// Tournament tree nodes of an n-row panel; cand needs
// batch * nodes * lu_cand_stride() ints, counters batch * nodes (zeroed once).
int lu_tree_nodes(int n);
int lu_cand_stride();
template <typename T>
hipError_t lu_factor(T* A, size_t strideA, int n, int lda, int batch, LuWork w,
                     hipStream_t s, const struct Prof* prof);
// Work buffers of the single-launch triangular solve: flags [batch][nblk]
// (zeroed once), tickets [batch], err [1]; epoch grows by one per launch.
struct TrsvWork {
    int* flags;
    int* tickets;
    int* err;
    int epoch;
    // blocks per matrix (each takes row blocks in order until none is
    // left); 0: one per row block.  A few blocks leave the CUs to work that
    // runs beside the solve (the next batch's factorisation).
    int rows = 0;
};
constexpr int kTrsvRowsShared = 4;

// Back substitution U x = y in place on column n of A (single-launch
// solve; tw holds its flags, tickets and error word).
template <typename T>
hipError_t lu_backsolve(T* A, size_t strideA, int n, int lda, int batch, hipStream_t s, TrsvWork* tw);
// Triangular solve with the factors in A on a vector x (element e of batch
// b at X[b*strideX + 2*e*xs]): unit lower L (lower) or upper U; batch
// elements with skip[b] != 0 (skip may be null) are not touched.
template <typename T>
hipError_t lu_trsv(const T* A, size_t strideA, int n, int lda, T* X, size_t strideX, size_t xs,
                   int batch, bool lower, const int* skip, hipStream_t s, TrsvWork* tw);
// Row permutation equivalent to the sequential interchanges of ipiv.
hipError_t lu_make_perm(const int* ipiv, int n, int batch, int* perm, hipStream_t s);
// The identity permutation (a factorisation without interchanges).
hipError_t lu_ident_perm(int n, int batch, int* perm, hipStream_t s);
size_t lu_perm_lds_limit();

// ------------------------------------------- allocation-free solve ------
// One static device array per GPU (efie_noalloc.hip) from which
// efie_zgesv_noalloc carves every buffer of its solve.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // EFIE_NOALLOC_SCRATCH_BYTES of efie.h
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocMaxBatch = 64;
// Device address of `device`'s copy of the array (looked up once per device
// with hipGetSymbolAddress, then cached).
hipError_t noalloc_scratch(int device, void** ptr);
// In-place transpose of the leading n x n block of a row-major complex
// (interleaved double) matrix with leading dimension ld.
hipError_t launch_transpose_square(double* A, int n, size_t ld, hipStream_t s);

// ------------------------------------------------ mixed precision -------
hipError_t launch_to_f32(const double* A64, float* A32, size_t complexElems, hipStream_t s);
// Column n of the nb matrices (complex elements zElems apart, row-major,
// leading dimension lda) := V (matrix b's at V + b*vStride doubles); also in
// Z32 when non-null.
hipError_t launch_set_rhs(const double* V, double* Z, float* Z32, int n, int lda, size_t zElems, int nb,
                          hipStream_t s, size_t vStride = 0);
// Packed batches (matrices of problems of order n <= np at order np):
// rows and columns n..np-1 of nf matrices (zElems complex elements apart)
// := the identity's, in Z and, when non-null, Z32; and nf copies of the
// right-hand side V[n] padded with zeros to np (out[f] at out + 2 np f).
hipError_t launch_pad_identity(double* Z, float* Z32, size_t zElems, int n, int np, int lda, int nf, hipStream_t s);
hipError_t launch_pack_rhs(const double* V, int n, int np, int nf, double* out, hipStream_t s);
hipError_t launch_col_to_vec(const float* A32, size_t strideA, int n, int lda, double* X, int batch,
                             hipStream_t s);
// R = V - A64 X per batch element (V[b] at V + b*vStride), rn2[b] += ||R||^2
hipError_t launch_residual(const double* A64, size_t strideA, int n, int lda, const double* V,
                           size_t vStride, const double* X, double* R, double* rn2, const int* done,
                           int batch, hipStream_t s);
// The same for a complex symmetric A64 (A64 = A64^T bit-exact), reading
// only its upper triangle; P holds residual_sym_bytes(n, batch) of partial
// sums.  xZero: X = 0 (R = V, A64 and X not read).  shared: the residual
// runs beside a factorisation (narrow strips, few registers per wave).
hipError_t launch_residual_sym(const double* A64, size_t strideA, int n, int lda, const double* V, size_t vStride,
                               const double* X, double* R, double* P, double* rn2, const int* done, int batch,
                               int xZero, bool shared, hipStream_t s);
size_t residual_sym_bytes(int n, int batch);
// remaining > 0 (a retry follows): also stop a system whose last reduction
// factor, kept for the remaining steps, would not reach tol.
hipError_t launch_ir_check(double* rn2, const double* vn2, double tol, int batch, int* done,
                           int* iters, double* resid, int firstPass, int lastPass, int remaining, hipStream_t s);
hipError_t launch_perm_f32(const double* R, const int* perm, int n, int batch, float* R32,
                           const int* done, hipStream_t s);
hipError_t launch_add_f32(double* X, const float* D32, int n, int batch, const int* done,
                          hipStream_t s);
hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s);
hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_fill_dbl(double* p, double v, int n, hipStream_t s);

// Optional per-kernel-class timing hook (efie_opts.profile).
// Event marks around kernel launches of one class; end() names how many
// kernels the region launched (-1 marks a region start).
struct Prof {
    void* user = nullptr;
    void (*mark)(void* user, int cls, int launches) = nullptr;
    void begin(int cls) const {
        if (mark) mark(user, cls, -1);
    }
    void end(int cls, int launches = 1) const {
        if (mark) mark(user, cls, launches);
    }
};

// ------------------------------------------------ near-pair surrogate ---
// Network 13 -> 64 -> 64 -> 64 -> 16 (GELU), FP32, parameters in one flat
// array W0[64][13] b0 W1[64][64] b1 W2[64][64] b2 W3[16][64] b3.
constexpr int kMlpIn = 13, kMlpHidden = 64, kMlpOut = 16;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + 2 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpOut * kMlpHidden + kMlpOut;
// Standardisation stored after the parameters in the device copy of a
// trained model: xmean[13] xinv[13] (1/std) ymean[16] ystd[16].
constexpr int kMlpNorm = 2 * kMlpIn + 2 * kMlpOut;

// A near pair (t1 < t2) in the test triangle's frame: origin at the
// centroid of t1, rows of Rm = (e_x along its first edge, e_y, normal),
// lengths in units of L = sqrt(4 A1 / sqrt(3)); d is the centroid distance.
// feat: the 12 normalised coordinates (x0, y0, x1 of t1's first two
// vertices; all of t2's three vertices).
struct SurPair {
    int t1, t2;
    float feat[12];
    double Rm[9];
    double L, d;
};

// Labels of the surrogate: for each pair (18 normalised vertex coordinates,
// t1's three then t2's three, and the normalised wavenumber kt = k L) the
// difference between the subdivided near-pair moments (softened, as
// rwg_pair_integrals_near_singular.m) and the plain product rule of the
// regular pairs, times exp(+j kt dt) 4 pi: T[p][16] (re, im of S0, S1,
// S1', S2); nrm[p] (if non-null) = 4 pi |near-pair moments|.  With fr
// non-null, kt = fr->kr * pairs[p].L instead of kt[p].  subBary/regBary:
// rule points of the subdivided / regular rule as barycentric weights of
// the parent's vertices and weight fractions of twice its area ([n][4]).
hipError_t launch_sur_labels(const double* tri18, const double* kt, const Freq* fr, const SurPair* pairs, int n,
                             const double* subBary, int Np, const double* regBary, int Nq, double near_alpha,
                             double* T, double* nrm, hipStream_t s);
// MLP kernels (plain arrays: X [N][13], Y and targets [N][16]).
hipError_t launch_mlp_forward(const float* X, const float* params, float* Y, int N, hipStream_t s);
// Gradient of the loss sum((y - t)^2) / (16 N) over the N samples; work
// needs mlp_grad_work_floats(N) floats; lossOut gets the loss (one float).
size_t mlp_grad_work_floats(int N);
hipError_t launch_mlp_loss_grad(const float* X, const float* T, const float* params, int N, float* work,
                                float* grad, float* lossOut, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, int step, hipStream_t s);
// Records of the n near pairs in the network's envelope at wavenumber
// fr->kr, each written once: the plain product rule of the regular pairs
// (the rule's points regBary [Nq][4] on the pair's normalised vertices
// tri18 [n][18]) plus the network's correction (model: params then the
// kMlpNorm standardisation), or with Tbuf non-null the correction taken
// from Tbuf (labels, [n][16]).
hipError_t launch_sur_records(const SurPair* pairs, int n, const float* model, const double* Tbuf,
                              const double* tri18, const double* regBary, int Nq, const Freq* fr, Rec* recs,
                              hipStream_t s);

// Zin epilogue of batch element b: x[e] = X[b*strideX + e*xinc] (complex),
// Iport = sum(wP.*len.*x) - sum(wG.*len.*x) (or x[p]-x[g] in the centroid
// model), Zin = V0/Iport, S11 = (Zin-Z0)/(Zin+Z0); ZinS11[4b..4b+3].
hipError_t launch_epilogue(const double* X, size_t strideX, size_t xinc, const double* wdiff,
                           int Ne, int patchEdge0, int groundEdge0, int centroidMode, double V0,
                           double Z0, int batch, double* ZinS11, hipStream_t s);

}  // namespace efie
