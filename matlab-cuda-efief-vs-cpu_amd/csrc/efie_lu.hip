// efie_lu.hip -- batched complex LU solve for gfx950.
//
// Replaces the per-frequency cusolverDnZgetrf/Zgetrs pair of the reference
// (patch_sweep_cuda.cu:473-480, patch_sweep_fastquad_cuda.cu:765-771).
//
// Layout: a batch of matrices, each row-major with leading dimension lda,
// interleaved complex of real type T (double, or float for the mixed
// precision factorisation).  Column n carries the right-hand side through
// every row operation, so the factorisation leaves L^{-1} P b there and
// only the back substitution remains.
//
// Right-looking blocked algorithm, panel width NB:
//   1. tournament pivoting (CALU; Grigori, Demmel and Xiang 2011) selects
//      the NB pivot rows of the panel: every leaf block of rows runs
//      partial pivoting in LDS and proposes NB candidates, candidate sets
//      are merged pairwise up a binary tree; the last workgroup to arrive
//      at a node merges it, so the whole tree is one launch without any
//      spin-wait;
//   2. the selected rows are swapped into place (LAPACK ipiv order);
//   3. A11 = L11 U11 without further pivoting, L21 = A21 U11^-1 and
//      U12 = L11^-1 A12 (all in one launch);
//   4. A22 -= L21 U12 on the matrix cores (complex product as four real
//      MFMA products).
// Pivot search uses |re|+|im| like LAPACK's izamax.
#include "efie_internal.h"

#include <math.h>
#include <type_traits>

namespace efie {

constexpr int NB = 32;    // panel width
constexpr int BR = 128;   // leaf rows of the tournament

template <typename T>
struct cpx {
    T r, i;
};

template <typename T>
__device__ __forceinline__ cpx<T> ld(const T* p) {
    if constexpr (std::is_same<T, double>::value) {
        double2 v = *reinterpret_cast<const double2*>(p);
        return {v.x, v.y};
    } else {
        float2 v = *reinterpret_cast<const float2*>(p);
        return {v.x, v.y};
    }
}
template <typename T>
__device__ __forceinline__ void st(T* p, cpx<T> v) {
    if constexpr (std::is_same<T, double>::value)
        *reinterpret_cast<double2*>(p) = make_double2(v.r, v.i);
    else
        *reinterpret_cast<float2*>(p) = make_float2(v.r, v.i);
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T>
__device__ __forceinline__ cpx<T> cinv(cpx<T> a) {
    T d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}

// ------------------------------------------------------ GEPP in LDS ------
// Partial pivoting on an r x w block (row stride w) held in LDS, rows
// labelled by lab[].  Afterwards lab[0..min(r,w)) are the pivot rows in
// order.  Returns (block-uniformly) the first column with an exactly zero
// pivot, or -1.  Requires blockDim.x == 256.
template <typename T>
__device__ int gepp_lds(cpx<T>* a, int* lab, int r, int w, T* redv, int* redi) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    int zero_col = -1;
    const int steps = r < w ? r : w;
    for (int c = 0; c < steps; ++c) {
        // pivot search: max |re|+|im|, first index on ties
        T best = (T)-1;
        int bi = 0x7fffffff;
        for (int i = c + tid; i < r; i += 256) {
            cpx<T> v = a[i * w + c];
            T m = fabs(v.r) + fabs(v.i);
            if (m > best) { best = m; bi = i; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            T ob = __shfl_xor(best, o, 64);
            int oi = __shfl_xor(bi, o, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) { redv[wv] = best; redi[wv] = bi; }
        __syncthreads();
        if (tid == 0) {
            T b = redv[0];
            int ii = redi[0];
            for (int k = 1; k < 4; ++k)
                if (redv[k] > b || (redv[k] == b && redi[k] < ii)) { b = redv[k]; ii = redi[k]; }
            redv[4] = b;
            redi[4] = ii;
        }
        __syncthreads();
        const int p = redi[4];
        const bool zp = !(redv[4] > (T)0);
        // swap rows c and p
        if (p != c) {
            for (int k = tid; k < w; k += 256) {
                cpx<T> t = a[c * w + k];
                a[c * w + k] = a[p * w + k];
                a[p * w + k] = t;
            }
            if (tid == 0) { int t = lab[c]; lab[c] = lab[p]; lab[p] = t; }
        }
        __syncthreads();
        if (zp) {
            if (zero_col < 0) zero_col = c;
            continue;
        }
        const cpx<T> inv = cinv(a[c * w + c]);
        // scale column c, then rank-1 update of the trailing block
        for (int i = c + 1 + tid; i < r; i += 256) a[i * w + c] = cmul(a[i * w + c], inv);
        __syncthreads();
        const int rr = r - c - 1, cc = w - c - 1;
        for (int e = tid; e < rr * cc; e += 256) {
            int i = c + 1 + e / cc, k = c + 1 + e % cc;
            cpx<T> l = a[i * w + c], u = a[c * w + k];
            cpx<T> v = a[i * w + k];
            v.r -= l.r * u.r - l.i * u.i;
            v.i -= l.r * u.i + l.i * u.r;
            a[i * w + k] = v;
        }
        __syncthreads();
    }
    return zero_col;
}

// ------------------------------------------------------ tournament -------

struct TreeGeom {
    int nLevels;
    int levelBase[24];
    int levelCount[24];
    int nNodes;
};

static TreeGeom tree_geom(int M) {
    TreeGeom g{};
    int L = (M + BR - 1) / BR;
    int base = 0, lv = 0;
    while (true) {
        g.levelBase[lv] = base;
        g.levelCount[lv] = L;
        base += L;
        ++lv;
        if (L == 1) break;
        L = (L + 1) / 2;
    }
    g.nLevels = lv;
    g.nNodes = base;
    return g;
}

__device__ __forceinline__ int ald(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// cand layout per batch: [nNodes][NB + 1] (count followed by labels)
template <typename T>
__global__ __launch_bounds__(256) void k_tslu(T* A, size_t strideA, int n, int lda, int j0, int w,
                                              TreeGeom g, int nodeStride, int* cand, int* counters,
                                              int* ipiv, int* info) {
    __shared__ cpx<T> a[(BR > 2 * NB ? BR : 2 * NB) * NB];
    __shared__ int lab[(BR > 2 * NB ? BR : 2 * NB)];
    __shared__ T redv[8];
    __shared__ int redi[8];
    __shared__ int s_go;
    const int b = blockIdx.y;
    T* Ab = A + (size_t)b * strideA;
    int* cb = cand + (size_t)b * nodeStride * (NB + 1);
    int* ctr = counters + (size_t)b * nodeStride;
    const int tid = threadIdx.x;
    const int M = n - j0;

    int level = 0, node = blockIdx.x;
    // leaf rows
    int r = min(BR, M - node * BR);
    for (int i = tid; i < r; i += 256) lab[i] = j0 + node * BR + i;
    __syncthreads();
    while (true) {
        // load the labelled rows of the panel into LDS
        for (int e = tid; e < r * w; e += 256) {
            int i = e / w, k = e % w;
            a[i * w + k] = ld(Ab + 2 * ((size_t)lab[i] * lda + j0 + k));
        }
        __syncthreads();
        int zc = gepp_lds<T>(a, lab, r, w, redv, redi);
        const int keep = r < w ? r : w;
        const bool root = (g.levelCount[level] == 1);
        if (root) {
            if (tid == 0) {
                if (zc >= 0 && info[b] == 0) info[b] = j0 + zc + 1;
                // simulate the sequential swaps: ipiv[j0+i] = current position
                // of the i-th selected row
                int rowAt[2 * NB], posOf[2 * NB], nt = 0;
                // rows touched: positions j0..j0+keep-1 and the pivot rows
                for (int i = 0; i < keep; ++i) {
                    int target = lab[i];
                    int q = target;  // current position of target
                    for (int k = 0; k < nt; ++k)
                        if (rowAt[k] == target) { q = posOf[k]; break; }
                    int pos = j0 + i;
                    ipiv[(size_t)b * n + pos] = q + 1;
                    // row currently at pos
                    int cur = pos;
                    bool found = false;
                    for (int k = 0; k < nt; ++k)
                        if (posOf[k] == pos) { cur = rowAt[k]; found = true; break; }
                    (void)found;
                    // move cur to q, target to pos
                    bool fc = false, ft = false;
                    for (int k = 0; k < nt; ++k) {
                        if (rowAt[k] == cur && !fc) { posOf[k] = q; fc = true; }
                        else if (rowAt[k] == target && !ft) { posOf[k] = pos; ft = true; }
                    }
                    if (!fc) { rowAt[nt] = cur; posOf[nt] = q; ++nt; }
                    if (!ft && target != cur) { rowAt[nt] = target; posOf[nt] = pos; ++nt; }
                }
            }
            return;
        }
        // publish this node's candidates
        int* my = cb + (size_t)(g.levelBase[level] + node) * (NB + 1);
        for (int i = tid; i < keep; i += 256) ast(my + 1 + i, lab[i]);
        if (tid == 0) ast(my, keep);
        __threadfence();
        __syncthreads();
        const int parent = node >> 1;
        const int sib = node ^ 1;
        const bool hasSib = sib < g.levelCount[level];
        if (tid == 0) {
            int go = 1;
            if (hasSib) {
                int* c = ctr + g.levelBase[level + 1] + parent;
                int old = atomicAdd(c, 1);
                if (old == 0) go = 0;
                else ast(c, 0);  // last arrival: reset for the next panel
            }
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;
        __threadfence();
        // gather children candidates, left child first
        int r2 = 0;
        for (int ch = 0; ch < 2; ++ch) {
            int cn = 2 * parent + ch;
            if (cn >= g.levelCount[level]) break;
            const int* cp = cb + (size_t)(g.levelBase[level] + cn) * (NB + 1);
            int cnt = ald(cp);
            for (int i = tid; i < cnt; i += 256) lab[r2 + i] = ald(cp + 1 + i);
            r2 += cnt;
        }
        __syncthreads();
        r = r2;
        level += 1;
        node = parent;
    }
}

// ------------------------------------------------------------ laswp -------
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* A, size_t strideA, int n, int lda, int j0, int w,
                                               const int* ipiv) {
    const int b = blockIdx.y;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c > n) return;
    T* Ab = A + (size_t)b * strideA;
    const int* ip = ipiv + (size_t)b * n;
    for (int i = 0; i < w; ++i) {
        int r1 = j0 + i, r2 = ip[j0 + i] - 1;
        if (r1 == r2) continue;
        cpx<T> x = ld(Ab + 2 * ((size_t)r1 * lda + c));
        cpx<T> y = ld(Ab + 2 * ((size_t)r2 * lda + c));
        st(Ab + 2 * ((size_t)r1 * lda + c), y);
        st(Ab + 2 * ((size_t)r2 * lda + c), x);
    }
}

// ----------------------------------------------------- panel finish ------
// k_panel_diag: A11 (w x w, after the interchanges) = L11 U11 in place,
// without pivoting (the tournament already ordered the pivot rows).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_diag(T* A, size_t strideA, int lda, int j0, int w, int* info) {
    __shared__ cpx<T> a11[NB * NB];
    const int b = blockIdx.x;
    T* Ab = A + (size_t)b * strideA;
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        int i = e / w, k = e % w;
        a11[i * NB + k] = ld(Ab + 2 * ((size_t)(j0 + i) * lda + j0 + k));
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
        const cpx<T> d = a11[c * NB + c];
        const bool zp = (d.r == (T)0 && d.i == (T)0);
        if (!zp) {
            const cpx<T> inv = cinv(d);
            for (int i = c + 1 + tid; i < w; i += 256) a11[i * NB + c] = cmul(a11[i * NB + c], inv);
        } else if (tid == 0) {
            atomicCAS(info + b, 0, j0 + c + 1);
        }
        __syncthreads();
        if (!zp) {
            const int rr = w - c - 1;
            for (int e = tid; e < rr * rr; e += 256) {
                int i = c + 1 + e / rr, k = c + 1 + e % rr;
                cpx<T> l = a11[i * NB + c], u = a11[c * NB + k], v = a11[i * NB + k];
                v.r -= l.r * u.r - l.i * u.i;
                v.i -= l.r * u.i + l.i * u.r;
                a11[i * NB + k] = v;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < w * w; e += 256) {
        int i = e / w, k = e % w;
        st(Ab + 2 * ((size_t)(j0 + i) * lda + j0 + k), a11[i * NB + k]);
    }
}

// k_panel: with L11 U11 in place, blocks [0, nRowBlocks) solve 256 rows of
// L21 = A21 U11^-1 each and the others 256 columns of U12 = L11^-1 A12
// (columns up to and including the rhs column n).
template <typename T>
__global__ __launch_bounds__(256) void k_panel(T* A, size_t strideA, int n, int lda, int j0, int w,
                                               int nRowBlocks) {
    __shared__ cpx<T> a11[NB * NB];
    const int b = blockIdx.y;
    T* Ab = A + (size_t)b * strideA;
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        int i = e / w, k = e % w;
        a11[i * NB + k] = ld(Ab + 2 * ((size_t)(j0 + i) * lda + j0 + k));
    }
    __syncthreads();
    if ((int)blockIdx.x < nRowBlocks) {
        // rows of L21: x U11 = a
        int row = j0 + w + blockIdx.x * 256 + tid;
        if (row >= n) return;
        T* p = Ab + 2 * ((size_t)row * lda + j0);
        cpx<T> x[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < w) x[c] = ld(p + 2 * c);
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c >= w) break;
            cpx<T> s = x[c];
            for (int k = 0; k < c; ++k) {
                cpx<T> u = a11[k * NB + c];
                s.r -= x[k].r * u.r - x[k].i * u.i;
                s.i -= x[k].r * u.i + x[k].i * u.r;
            }
            cpx<T> d = a11[c * NB + c];
            x[c] = (d.r == (T)0 && d.i == (T)0) ? s : cmul(s, cinv(d));
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < w) st(p + 2 * c, x[c]);
    } else {
        // columns of U12: L11 y = a, unit diagonal
        int col = j0 + w + (blockIdx.x - nRowBlocks) * 256 + tid;
        if (col > n) return;
        cpx<T> y[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < w) y[c] = ld(Ab + 2 * ((size_t)(j0 + c) * lda + col));
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c >= w) break;
            cpx<T> s = y[c];
            for (int k = 0; k < c; ++k) {
                cpx<T> l = a11[c * NB + k];
                s.r -= l.r * y[k].r - l.i * y[k].i;
                s.i -= l.r * y[k].i + l.i * y[k].r;
            }
            y[c] = s;
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < w) st(Ab + 2 * ((size_t)(j0 + c) * lda + col), y[c]);
    }
}

// -------------------------------------------------- trailing update ------
// C (rows r0.., cols c0..) -= L (rows r0.., cols j0..j0+w) * U (rows
// j0..j0+w, cols c0..).  Block tile 64 x 64 complex, four waves of 32 x 32,
// each 2 x 2 MFMA 16x16 tiles with separate real and imaginary accumulators.
// MFMA 16x16x4 operand layout: A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k.  The 16x16 result: lane l element v holds D[4(l/16)+v][l%16]
// for the f32 instruction and D[4v + l/16][l%16] for the f64 instruction
// (both measured on gfx950).

template <typename T>
struct mf;
template <>
struct mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * v + lk; }
};
template <>
struct mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_update(T* A, size_t strideA, int n, int lda, int j0,
                                                     int w) {
    typedef typename mf<T>::v4 v4;
    __shared__ cpx<T> Ls[64][NB + 1];
    __shared__ cpx<T> Us[NB][64 + 1];
    const int b = blockIdx.z;
    T* Ab = A + (size_t)b * strideA;
    const int r0 = j0 + w + blockIdx.y * 64;
    const int c0 = j0 + w + blockIdx.x * 64;
    const int tid = threadIdx.x;
    // stage L (64 x w) and U (w x 64), zero padded
    for (int e = tid; e < 64 * NB; e += 256) {
        int i = e / NB, k = e % NB;
        int row = r0 + i;
        cpx<T> v = {0, 0};
        if (row < n && k < w) v = ld(Ab + 2 * ((size_t)row * lda + j0 + k));
        Ls[i][k] = v;
    }
    for (int e = tid; e < NB * 64; e += 256) {
        int k = e / 64, j = e % 64;
        int col = c0 + j;
        cpx<T> v = {0, 0};
        if (col <= n && k < w) v = ld(Ab + 2 * ((size_t)(j0 + k) * lda + col));
        Us[k][j] = v;
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4 accR[2][2], accI[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accR[x][y] = v4{0, 0, 0, 0};
            accI[x][y] = v4{0, 0, 0, 0};
        }
    const int kmax = (w + 3) & ~3;
    for (int k0 = 0; k0 < kmax; k0 += 4) {
        cpx<T> av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) av[x] = Ls[wr + 16 * x + li][k0 + lk];
#pragma unroll
        for (int y = 0; y < 2; ++y) bv[y] = Us[k0 + lk][wc + 16 * y + li];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                accR[x][y] = mf<T>::mfma(av[x].r, bv[y].r, accR[x][y]);
                accR[x][y] = mf<T>::mfma(-av[x].i, bv[y].i, accR[x][y]);
                accI[x][y] = mf<T>::mfma(av[x].r, bv[y].i, accI[x][y]);
                accI[x][y] = mf<T>::mfma(av[x].i, bv[y].r, accI[x][y]);
            }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int row = r0 + wr + 16 * x + mf<T>::row(lk, v);
                int col = c0 + wc + 16 * y + li;
                if (row < n && col <= n) {
                    T* p = Ab + 2 * ((size_t)row * lda + col);
                    cpx<T> c = ld(p);
                    c.r -= accR[x][y][v];
                    c.i -= accI[x][y][v];
                    st(p, c);
                }
            }
}

// ------------------------------------------------- triangular solves -----
// Blocked triangular solve on a vector x (element e of batch b at
// X[b*strideX + 2*e*xs]), upper (U x = y, non-unit diagonal, bottom block
// first) or unit lower (L y = r, top block first): the diagonal block is
// solved by one wave in LDS, then the rows not yet solved are updated by a
// GEMV.  Batch elements with skip[b] != 0 are left untouched.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_diag(const T* A, size_t strideA, int lda, T* X,
                                                   size_t strideX, size_t xs, int r0, int bs,
                                                   const int* skip) {
    __shared__ cpx<T> u[64][65];
    __shared__ cpx<T> x[64];
    const int b = blockIdx.x;
    if (skip && skip[b]) return;
    const T* Ab = A + (size_t)b * strideA;
    T* Xb = X + (size_t)b * strideX;
    const int tid = threadIdx.x;
    for (int e = tid; e < bs * bs; e += 256) {
        int i = e / bs, k = e % bs;
        u[i][k] = ld(Ab + 2 * ((size_t)(r0 + i) * lda + r0 + k));
    }
    for (int i = tid; i < bs; i += 256) x[i] = ld(Xb + 2 * (size_t)(r0 + i) * xs);
    __syncthreads();
    if (tid < 64) {
        for (int jj = 0; jj < bs; ++jj) {
            const int j = LOWER ? jj : bs - 1 - jj;
            if (!LOWER && tid == 0) {
                cpx<T> d = u[j][j];
                x[j] = (d.r == (T)0 && d.i == (T)0) ? x[j] : cmul(x[j], cinv(d));
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            const cpx<T> xj = x[j];
            const bool act = LOWER ? (tid > j && tid < bs) : (tid < j);
            if (act) {
                cpx<T> l = u[tid][j];
                cpx<T> v = x[tid];
                v.r -= l.r * xj.r - l.i * xj.i;
                v.i -= l.r * xj.i + l.i * xj.r;
                x[tid] = v;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
    __syncthreads();
    for (int i = tid; i < bs; i += 256) st(Xb + 2 * (size_t)(r0 + i) * xs, x[i]);
}

// rows [rlo, rhi) -= A[rows, r0:r0+bs] * x[r0:r0+bs]; one wave per row
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_update(const T* A, size_t strideA, int lda, T* X,
                                                     size_t strideX, size_t xs, int r0, int bs,
                                                     int rlo, int rhi, const int* skip) {
    const int b = blockIdx.y;
    if (skip && skip[b]) return;
    const T* Ab = A + (size_t)b * strideA;
    T* Xb = X + (size_t)b * strideX;
    const int lane = threadIdx.x & 63;
    const int row = rlo + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rhi) return;
    cpx<T> s = {0, 0};
    if (lane < bs) {
        cpx<T> u = ld(Ab + 2 * ((size_t)row * lda + r0 + lane));
        cpx<T> xv = ld(Xb + 2 * (size_t)(r0 + lane) * xs);
        s.r = u.r * xv.r - u.i * xv.i;
        s.i = u.r * xv.i + u.i * xv.r;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s.r += __shfl_xor(s.r, o, 64);
        s.i += __shfl_xor(s.i, o, 64);
    }
    if (lane == 0) {
        T* y = Xb + 2 * (size_t)row * xs;
        cpx<T> v = ld(y);
        v.r -= s.r;
        v.i -= s.i;
        st(y, v);
    }
}

// perm[b][i] = source row of position i after the sequential swaps of ipiv
// (applying ipiv to a vector v gives v_new[i] = v[perm[i]]).
__global__ __launch_bounds__(256) void k_make_perm(const int* ipiv, int n, int* perm) {
    extern __shared__ int sh[];
    const int b = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sh[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int* ip = ipiv + (size_t)b * n;
        for (int i = 0; i < n; ++i) {
            int q = ip[i] - 1;
            int t = sh[i];
            sh[i] = sh[q];
            sh[q] = t;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[(size_t)b * n + i] = sh[i];
}

// ---------------------------------------------------------- drivers ------

int lu_tree_nodes(int n) { return tree_geom(n).nNodes; }
int lu_cand_stride() { return NB + 1; }

template <typename T>
hipError_t lu_factor(T* A, size_t strideA, int n, int lda, int batch, LuWork wk, hipStream_t s,
                     const Prof* pf) {
    Prof none;
    if (!pf) pf = &none;
    // candidate and counter slots per batch element are sized for the first
    // (largest) panel; every counter is back at zero when a launch ends
    const int nodeStride = tree_geom(n).nNodes;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int w = n - j0 < NB ? n - j0 : NB;
        TreeGeom g = tree_geom(n - j0);
        pf->begin(5);
        hipLaunchKernelGGL(k_tslu<T>, dim3(g.levelCount[0], batch), dim3(256), 0, s, A, strideA, n,
                           lda, j0, w, g, nodeStride, wk.cand, wk.counters, wk.ipiv, wk.info);
        pf->end(5);
        pf->begin(6);
        hipLaunchKernelGGL(k_laswp<T>, dim3((n + 1 + 255) / 256, batch), dim3(256), 0, s, A,
                           strideA, n, lda, j0, w, wk.ipiv);
        pf->end(6);
        const int rem = n - j0 - w;
        const int nRowBlocks = (rem + 255) / 256;
        const int nColBlocks = (rem + 1 + 255) / 256;
        pf->begin(7);
        hipLaunchKernelGGL(k_panel_diag<T>, dim3(batch), dim3(256), 0, s, A, strideA, lda, j0, w, wk.info);
        hipLaunchKernelGGL(k_panel<T>, dim3(nRowBlocks + nColBlocks, batch), dim3(256), 0, s, A,
                           strideA, n, lda, j0, w, nRowBlocks);
        pf->end(7);
        if (rem > 0) {
            dim3 gg((rem + 1 + 63) / 64, (rem + 63) / 64, batch);
            pf->begin(8);
            hipLaunchKernelGGL(k_gemm_update<T>, gg, dim3(256), 0, s, A, strideA, n, lda, j0, w);
            pf->end(8);
        }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t lu_trsv(const T* A, size_t strideA, int n, int lda, T* X, size_t strideX, size_t xs,
                   int batch, bool lower, const int* skip, hipStream_t s) {
    const int nblk = (n + 63) / 64;
    for (int q = 0; q < nblk; ++q) {
        const int bk = lower ? q : nblk - 1 - q;
        const int r0 = bk * 64;
        const int bs = n - r0 < 64 ? n - r0 : 64;
        if (lower)
            hipLaunchKernelGGL((k_trsv_diag<T, true>), dim3(batch), dim3(256), 0, s, A, strideA, lda, X, strideX,
                               xs, r0, bs, skip);
        else
            hipLaunchKernelGGL((k_trsv_diag<T, false>), dim3(batch), dim3(256), 0, s, A, strideA, lda, X, strideX,
                               xs, r0, bs, skip);
        const int rlo = lower ? r0 + bs : 0, rhi = lower ? n : r0;
        if (rhi > rlo)
            hipLaunchKernelGGL(k_trsv_update<T>, dim3((rhi - rlo + 3) / 4, batch), dim3(256), 0, s, A, strideA,
                               lda, X, strideX, xs, r0, bs, rlo, rhi, skip);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t lu_backsolve(T* A, size_t strideA, int n, int lda, int batch, hipStream_t s) {
    return lu_trsv<T>(A, strideA, n, lda, A + 2 * (size_t)n, strideA, (size_t)lda, batch, false, nullptr, s);
}

hipError_t lu_make_perm(const int* ipiv, int n, int batch, int* perm, hipStream_t s) {
    hipLaunchKernelGGL(k_make_perm, dim3(batch), dim3(256), sizeof(int) * (size_t)n, s, ipiv, n, perm);
    return hipGetLastError();
}

template hipError_t lu_factor<double>(double*, size_t, int, int, int, LuWork, hipStream_t, const Prof*);
template hipError_t lu_factor<float>(float*, size_t, int, int, int, LuWork, hipStream_t, const Prof*);
template hipError_t lu_backsolve<double>(double*, size_t, int, int, int, hipStream_t);
template hipError_t lu_backsolve<float>(float*, size_t, int, int, int, hipStream_t);
template hipError_t lu_trsv<double>(const double*, size_t, int, int, double*, size_t, size_t, int, bool,
                                    const int*, hipStream_t);
template hipError_t lu_trsv<float>(const float*, size_t, int, int, float*, size_t, size_t, int, bool,
                                   const int*, hipStream_t);

}  // namespace efie
