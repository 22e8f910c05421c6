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
constexpr int WB = 8 * NB;  // outer block: the trailing update runs with K = WB
constexpr int kSwapSlots = WB / NB;  // interchange lists kept per block
constexpr int lu_swap_stride_c = 1 + 4 * NB;  // count + up to 2 NB (dst, src) pairs


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
// two consecutive complex values from a 16-byte-aligned address, as one
// 16-byte load (float) or two (double)
template <typename T>
__device__ __forceinline__ void ld_pair(const T* p, cpx<T>& a, cpx<T>& b) {
    if constexpr (std::is_same<T, double>::value) {
        const double2 u = reinterpret_cast<const double2*>(p)[0], v = reinterpret_cast<const double2*>(p)[1];
        a = {u.x, u.y};
        b = {v.x, v.y};
    } else {
        const float4 u = *reinterpret_cast<const float4*>(p);
        a = {u.x, u.y};
        b = {u.z, u.w};
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
// a - l*u as four fused multiply-adds
template <typename T>
__device__ __forceinline__ cpx<T> cfms(cpx<T> a, cpx<T> l, cpx<T> u) {
    return {fma(-l.r, u.r, fma(l.i, u.i, a.r)), fma(-l.r, u.i, fma(-l.i, u.r, a.i))};
}
template <typename T>
__device__ __forceinline__ cpx<T> cinv(cpx<T> a) {
    T d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------ tournament -------
// Tree of candidate sets over the rows of the panel.  Every node is one
// wave: lane l holds one panel row (w <= NB complex values in registers)
// and the node runs partial pivoting on its rows (pivot = largest
// |re|+|im|, first position on ties, as LAPACK's izamax), keeping the
// pivot rows in pivot order as its candidates.  Leaves hold TL = 64 rows.

constexpr int TL = 64;

struct TreeGeom {
    int nLevels;
    int levelBase[24];
    int levelCount[24];
    int nNodes;
};

static TreeGeom tree_geom(int M) {
    TreeGeom g{};
    int L = (M + TL - 1) / TL;
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

// Cross-workgroup exchange of small data: agent-scope relaxed atomics read
// and write the coherent copy directly, and the writer's release fence
// orders them before the flag or counter that publishes them; on the
// reading side the thread that observes the flag takes an agent-scope
// acquire fence (once per wait, not per datum).
__device__ __forceinline__ int ald(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ cpx<T> ald_c(const T* p) {
    cpx<T> v;
    v.r = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v.i = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}
template <typename T>
__device__ __forceinline__ void ast_c(T* p, cpx<T> v) {
    __hip_atomic_store(p, v.r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, v.i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// orders this thread's earlier memory operations before later ones
// (release to the agent; no cache invalidation)
__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
// taken by the one thread that observed a flag or counter published by
// another workgroup (possibly on another XCD): the agent-scope acquire
// orders every later read after that observation; the workgroup barrier
// that follows, with acquire_wg in the other threads, extends the order to
// the whole block
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void acquire_wg() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// Largest 64-bit key of the wave, returned wave-uniformly: DPP steps
// quad_perm [1,0,3,2] and [2,3,0,1] (groups of 4), row_half_mirror (8),
// row_mirror (16), then row_bcast:15 and row_bcast:31 leave the maximum in
// lane 63, which is read back with readlane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    const int lo = (int)(unsigned)x, hi = (int)(unsigned)(x >> 32);
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
    return ((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) {
    return a > b ? a : b;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
    k = umax64(k, dpp_u64<0xB1, 0xF>(k));
    k = umax64(k, dpp_u64<0x4E, 0xF>(k));
    k = umax64(k, dpp_u64<0x141, 0xF>(k));
    k = umax64(k, dpp_u64<0x140, 0xF>(k));
    k = umax64(k, dpp_u64<0x142, 0xA>(k));
    k = umax64(k, dpp_u64<0x143, 0xC>(k));
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)k, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(k >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWMASK, 0xF, false);
}
// Largest value of the wave (lanes a DPP pattern cannot reach compare with
// themselves), broadcast from lane 63.
__device__ __forceinline__ int wave_max_i32(int v) {
    v = max(v, dpp_i32<0xB1, 0xF>(v));
    v = max(v, dpp_i32<0x4E, 0xF>(v));
    v = max(v, dpp_i32<0x141, 0xF>(v));
    v = max(v, dpp_i32<0x140, 0xF>(v));
    v = max(v, dpp_i32<0x142, 0xA>(v));
    v = max(v, dpp_i32<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_i32(int v) {
    v = min(v, dpp_i32<0xB1, 0xF>(v));
    v = min(v, dpp_i32<0x4E, 0xF>(v));
    v = min(v, dpp_i32<0x141, 0xF>(v));
    v = min(v, dpp_i32<0x140, 0xF>(v));
    v = min(v, dpp_i32<0x142, 0xA>(v));
    v = min(v, dpp_i32<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}
// 1/x from the hardware reciprocal (one Newton step for double); used only
// to rank pivot candidates, never for the factors themselves
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_fast(double x) {
    double r = __builtin_amdgcn_rcp(x);
    return r * (2.0 - x * r);
}

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane) {
    if constexpr (std::is_same<T, double>::value) {
        unsigned long long u = __double_as_longlong(v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    }
}

// Wave-level partial pivoting of up to 64 rows (one per lane, label myLab,
// -1 for none) of the w columns j0.. of A.  Returns in `chosen` (lane i)
// the label of the i-th pivot row; zc = first column with a zero pivot.
// The pivot of a column is the row of largest |re|+|im| (compared in
// single precision), the first position among equals.  Must be called by
// all 64 threads of a one-wave block; prow is NB complex values of LDS.
//
// The step is a chain of dependent operations, so it is kept short: a
// float DPP max (positions are compared only when magnitudes tie), the
// pivot row through LDS, and the multipliers from a hardware reciprocal.
// The multipliers here only rank the candidates of later columns; the
// factors themselves are computed by the panel kernels.  Columns past w
// are zero in every row, so the updates run over all columns unguarded.
template <typename T>
__device__ __forceinline__ void gepp_wave(const T* Ab, int lda, int j0, int w, int r, int myLab, int& chosen,
                                          int& zc, cpx<T>* prow) {
    const int lane = threadIdx.x;
    // all NB loads are issued unconditionally (row clamped to j0, columns
    // past w read in-bounds padding) and masked afterwards
    cpx<T> a[NB];
    const T* rowp = Ab + 2 * ((size_t)(myLab >= 0 ? myLab : j0) * lda + j0);
#pragma unroll
    for (int k = 0; k < NB; ++k) a[k] = ld(rowp + 2 * k);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const bool ok = k < w && myLab >= 0;
        a[k].r = ok ? a[k].r : (T)0;
        a[k].i = ok ? a[k].i : (T)0;
    }
    int pos = lane;
    bool act = lane < r;
    chosen = -1;
    zc = -1;
    const int steps = r < w ? r : w;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < steps) {
            // magnitudes compared as the integer bits of non-negative floats;
            // inactive rows carry -1 and never win (some row is active at
            // every step)
            const int mag = act ? __float_as_int((float)(fabs(a[c].r) + fabs(a[c].i))) : -1;
            const int mx = wave_max_i32(mag);
            const unsigned long long m = __ballot(act && mag == mx);
            int p = __ffsll((long long)m) - 1;
            if (__popcll(m) > 1) {
                // equal magnitudes: the first position wins
                const bool cand = (m >> lane) & 1ull;
                const int pmin = wave_min_i32(cand ? pos : 0x7fffffff);
                p = __ffsll((long long)__ballot(cand && pos == pmin)) - 1;
            }
            const bool nonzero = mx > 0;
            const int bpos = __builtin_amdgcn_readlane(pos, p);
            const int labp = __builtin_amdgcn_readlane(myLab, p);
            if (lane == p) {
#pragma unroll
                for (int k = c; k < NB; ++k) prow[k] = a[k];
            }
            if (pos == c) pos = bpos;  // the row at position c takes the pivot row's place
            if (lane == p) {
                pos = c;
                act = false;
            }
            if (lane == c) chosen = labp;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            if (!nonzero) {
                if (zc < 0) zc = c;
            } else {
                const cpx<T> d = prow[c];
                const T rn = rcp_fast(d.r * d.r + d.i * d.i);
                const cpx<T> inv = {d.r * rn, -d.i * rn};
                // rows already chosen (and the pivot row) get multiplier 0
                cpx<T> l = cmul(a[c], inv);
                if (!act) l = {(T)0, (T)0};
#pragma unroll
                for (int k = c + 1; k < NB; ++k) a[k] = cfms(a[k], l, prow[k]);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
}

// cand layout per batch: [nNodes][NB + 1] (count followed by labels)
template <typename T>
__global__ __launch_bounds__(64) void k_tslu(T* A, size_t strideA, int n, int lda, int j0, int w,
                                             TreeGeom g, int nodeStride, int* cand, int* counters,
                                             int* ipiv, int* info, int* swaps, int slot) {
    __shared__ int s_go;
    __shared__ cpx<T> prow[NB];
    const int b = blockIdx.y;
    const T* Ab = A + (size_t)b * strideA;
    int* cb = cand + (size_t)b * nodeStride * (NB + 1);
    int* ctr = counters + (size_t)b * nodeStride;
    const int lane = threadIdx.x;
    const int M = n - j0;

    int level = 0, node = blockIdx.x;
    int r = min(TL, M - node * TL);
    int myLab = lane < r ? j0 + node * TL + lane : -1;
    while (true) {
        int chosen, zc;
        gepp_wave<T>(Ab, lda, j0, w, r, myLab, chosen, zc, prow);
        const int keep = r < w ? r : w;
        if (g.levelCount[level] == 1) {
            // Root: turn the chosen rows (lane i holds the i-th) into LAPACK
            // interchanges ipiv[j0+i] = current position of the i-th row.
            // Lane k < w tracks the row at position j0+k; lane k < nd tracks
            // the k-th row pushed out of the top block and its position.
            if (lane == 0 && zc >= 0 && info[b] == 0) info[b] = j0 + zc + 1;
            int topRow = j0 + lane;
            int dRow = -1, dPos = -1, nd = 0;
            for (int i = 0; i < keep; ++i) {
                const int t = __shfl(chosen, i, 64);
                const unsigned long long mTop = __ballot(lane < keep && lane >= i && topRow == t);
                const unsigned long long mDisp = __ballot(lane < nd && dRow == t);
                int q = t, kd = -1;
                if (mTop) q = j0 + (__ffsll((long long)mTop) - 1);
                else if (mDisp) {
                    kd = __ffsll((long long)mDisp) - 1;
                    q = __shfl(dPos, kd, 64);
                }
                if (lane == 0) ipiv[(size_t)b * n + j0 + i] = q + 1;
                const int cur = __shfl(topRow, i, 64);  // row now at position j0+i
                if (q - j0 < keep) {
                    if (lane == q - j0) topRow = cur;
                } else if (kd >= 0) {
                    if (lane == kd) dRow = cur;          // position q now holds cur
                } else {
                    if (lane == nd) { dRow = cur; dPos = q; }
                    ++nd;
                }
                if (lane == i) topRow = t;
            }
            // The same interchanges as a gather list for k_laswp: position
            // j0+lane now holds row topRow (lane < keep), position dPos holds
            // dRow (lane < nd); unchanged positions are left out.
            int* sw = swaps + ((size_t)b * kSwapSlots + slot) * lu_swap_stride_c;
            const bool e1 = lane < keep && topRow != j0 + lane;
            const bool e2 = lane < nd && dRow != dPos;
            const unsigned long long m1 = __ballot(e1), m2 = __ballot(e2);
            const unsigned long long below = (1ull << lane) - 1ull;
            const int n1 = __popcll(m1);
            if (e1) {
                const int k = __popcll(m1 & below);
                sw[1 + 2 * k] = j0 + lane;
                sw[2 + 2 * k] = topRow;
            }
            if (e2) {
                const int k = n1 + __popcll(m2 & below);
                sw[1 + 2 * k] = dPos;
                sw[2 + 2 * k] = dRow;
            }
            if (lane == 0) sw[0] = n1 + __popcll(m2);
            return;
        }
        // publish this node's candidates
        int* my = cb + (size_t)(g.levelBase[level] + node) * (NB + 1);
        if (lane < keep) ast(my + 1 + lane, chosen);
        if (lane == 0) ast(my, keep);
        release_agent();
        __syncthreads();
        const int parent = node >> 1;
        const bool hasSib = (node ^ 1) < g.levelCount[level];
        if (threadIdx.x == 0) {
            int go = 1;
            if (hasSib) {
                int* c = ctr + g.levelBase[level + 1] + parent;
                const int old = atomicAdd(c, 1);
                if (old == 0) go = 0;
                else ast(c, 0);  // last arrival resets the counter for the next panel
            }
            if (go) acquire_agent();
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;
        acquire_wg();
        // candidates of both children, left child first, one per lane
        const int* c0 = cb + (size_t)(g.levelBase[level] + 2 * parent) * (NB + 1);
        const int n0 = ald(c0);
        int n1 = 0;
        const int* c1 = c0 + (NB + 1);
        if (2 * parent + 1 < g.levelCount[level]) n1 = ald(c1);
        myLab = -1;
        if (lane < n0) myLab = ald(c0 + 1 + lane);
        else if (lane < n0 + n1) myLab = ald(c1 + 1 + (lane - n0));
        r = n0 + n1;
        level += 1;
        node = parent;
    }
}

// ------------------------------------------------------------ laswp -------
// Row interchanges as gathers: the tournament root lists the m <= 2 NB
// positions whose row changes (destination, source).  Each thread owns one
// column, loads all m sources first (unconditionally, unused slots read
// row j0) and then stores, so no round trip waits on another.
template <typename T>
__device__ __forceinline__ void swap_column(T* Ab, int lda, int c, const int* L) {
    const int m = L[0];
    if (m == 0) return;
    const int fallback = L[2];
    cpx<T> v[2 * NB];
#pragma unroll
    for (int e = 0; e < 2 * NB; ++e) {
        const int src = e < m ? L[2 + 2 * e] : fallback;
        v[e] = ld(Ab + 2 * ((size_t)src * lda + c));
    }
#pragma unroll
    for (int e = 0; e < 2 * NB; ++e)
        if (e < m) st(Ab + 2 * ((size_t)L[1 + 2 * e] * lda + c), v[e]);
}

// One panel's interchanges (list `slot`) on columns [clo, chi).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* A, size_t strideA, int lda, const int* swaps, int slot, int clo,
                                               int chi) {
    __shared__ int L[lu_swap_stride_c];
    const int b = blockIdx.y;
    const int* sw = swaps + ((size_t)b * kSwapSlots + slot) * lu_swap_stride_c;
    if (threadIdx.x < lu_swap_stride_c) L[threadIdx.x] = sw[threadIdx.x];
    __syncthreads();
    const int c = clo + blockIdx.x * 256 + threadIdx.x;
    if (c >= chi) return;
    swap_column(A + (size_t)b * strideA, lda, c, L);
}

// The interchanges of nslots panels, in order, on columns [lo1, hi1) and
// [lo2, hi2) (the columns outside the block that produced them).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_rest(T* A, size_t strideA, int lda, const int* swaps, int nslots,
                                                    int lo1, int hi1, int lo2, int hi2) {
    __shared__ int L[kSwapSlots * lu_swap_stride_c];
    const int b = blockIdx.y;
    const int* sw = swaps + (size_t)b * kSwapSlots * lu_swap_stride_c;
    for (int i = threadIdx.x; i < nslots * lu_swap_stride_c; i += 256) L[i] = sw[i];
    __syncthreads();
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int n1 = hi1 - lo1;
    const int c = t < n1 ? lo1 + t : lo2 + (t - n1);
    if (t >= n1 + (hi2 - lo2)) return;
    T* Ab = A + (size_t)b * strideA;
    for (int k = 0; k < nslots; ++k) swap_column(Ab, lda, c, L + k * lu_swap_stride_c);
}

// ------------------------------------------ symmetric factorisation ------
// For a complex symmetric A (A = A^T) elimination without pivoting gives
// U = D L^T with D the diagonal of U, so the block row U12 of an outer block
// need not be eliminated: k_sym_u12 writes U[k][j] = U[k][k] L[j][k] for
// rows k in [k0, k1) and columns j in [c0, c1), through a 64 x 32 LDS tile
// (coalesced on both sides).  It also leaves the final U in place for the
// triangular solves.
template <typename T>
__global__ __launch_bounds__(256) void k_sym_u12(T* A, size_t strideA, int lda, int k0, int k1, int c0, int c1,
                                                 const int* __restrict__ nElem) {
    __shared__ cpx<T> t[64][33];
    __shared__ cpx<T> dg[32];
    const int b = blockIdx.z;
    T* Ab = A + (size_t)b * strideA;
    const int jb = c0 + blockIdx.x * 64, kb = k0 + blockIdx.y * 32;
    // padded matrix: L rows and D entries beyond its order give zero U
    if (nElem && (jb >= nElem[b] || kb >= nElem[b])) return;
    const int tid = threadIdx.x;
    // read L[j][k], j in [jb, jb+64), k in [kb, kb+32): 32 consecutive k per row
    for (int e = tid; e < 64 * 32; e += 256) {
        const int jj = e / 32, kk = e % 32, j = jb + jj, k = kb + kk;
        cpx<T> v{(T)0, (T)0};
        if (j < c1 && k < k1) v = ld(Ab + 2 * ((size_t)j * lda + k));
        t[jj][kk] = v;
    }
    if (tid < 32) {
        const int k = kb + tid;
        cpx<T> v{(T)0, (T)0};
        if (k < k1) v = ld(Ab + 2 * ((size_t)k * lda + k));
        dg[tid] = v;
    }
    __syncthreads();
    // write U[k][j] = d_k L[j][k]: 64 consecutive j per row
    for (int e = tid; e < 32 * 64; e += 256) {
        const int kk = e / 64, jj = e % 64, k = kb + kk, j = jb + jj;
        if (k < k1 && j < c1) st(Ab + 2 * ((size_t)k * lda + j), cmul(dg[kk], t[jj][kk]));
    }
}

// ----------------------------------------------------- panel finish ------
// k_panel_diag: A11 (w x w, after the interchanges) = L11 U11 in place,
// without pivoting (the tournament already ordered the pivot rows).
// SYM (complex-symmetric A11, no interchanges): every step reads the lower
// triangle only.  Column c of the active matrix (d_c L[k][c] for k > c)
// stands in for row c, and lane c stores it as its U row, so U11 = D L11^T
// exactly.  The rounding of the two triangles then cannot drift apart;
// with a drift, U = D L^T of the outer blocks would not be the U of the
// elimination, and near resonance small pivots amplify that mismatch into a
// backward error of order 1e-4 in FP32.
// The body of k_panel_diag for a block of width w; inlined once with w =
// NB (the full sub-panels: every bound a constant, so the column's LDS
// reads of a step issue together) and once for a narrower last one.
template <typename T, bool SYM>
__device__ __forceinline__ void panel_diag_body(T* Ab, int lda, int j0, int w, int* infob, cpx<T>* prow) {
    const int lane = threadIdx.x;
    cpx<T> a[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        a[k].r = (T)0;
        a[k].i = (T)0;
        if (k < w && lane < w) a[k] = ld(Ab + 2 * ((size_t)(j0 + lane) * lda + j0 + k));
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < w) {
            if constexpr (SYM) {
                if (lane >= c && lane < w) prow[lane] = a[c];
            } else if (lane == c) {
#pragma unroll
                for (int k = c; k < NB; ++k)
                    if (k < w) prow[k] = a[k];
            }
            __syncthreads();
            const cpx<T> d = prow[c];
            const bool zp = (d.r == (T)0 && d.i == (T)0);
            if (zp) {
                if (lane == 0) atomicCAS(infob, 0, j0 + c + 1);
            } else if (lane > c && lane < w) {
                const cpx<T> l = cmul(a[c], cinv(d));
                a[c] = l;
#pragma unroll
                for (int k = c + 1; k < NB; ++k)
                    if (k < w) a[k] = cfms(a[k], l, prow[k]);
            } else if (SYM && lane == c) {
#pragma unroll
                for (int k = c + 1; k < NB; ++k)
                    if (k < w) a[k] = prow[k];
            }
            __syncthreads();
        }
    }
    if (lane < w) {
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k < w) st(Ab + 2 * ((size_t)(j0 + lane) * lda + j0 + k), a[k]);
    }
}

template <typename T, bool SYM = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void k_panel_diag(T* A, size_t strideA, int lda, int j0, int w, int* info,
                                                   const int* __restrict__ nElem) {
    // one wave, lane k holds row k of the block; row c (SYM: column c) is
    // broadcast through LDS
    __shared__ cpx<T> prow[NB];
    const int b = blockIdx.x;
    // padded matrix: a diagonal block in the identity padding is its own
    // factorisation
    if (nElem && j0 >= nElem[b]) return;
    T* Ab = A + (size_t)b * strideA;
    if (w == NB)
        panel_diag_body<T, SYM>(Ab, lda, j0, NB, info + b, prow);
    else
        panel_diag_body<T, SYM>(Ab, lda, j0, w, info + b, prow);
}

// With L11 U11 of the w x w diagonal block at (d, d) in place:
// k_panel_rows solves rows [rlo, rhi) of L21 = A21 U11^-1 (columns d..d+w),
// k_panel_cols solves columns [clo, chi) of U12 = L11^-1 A12 (rows d..d+w).
// Row solve x U11 = a of k_panel_rows for a block of width w (x in place
// of a).  Right-looking: once x_k is final it is taken out of every later
// column, so the updates of a step are independent of each other; each
// x_c still receives its terms in the order k = 0, 1, ..., c-1 and is then
// multiplied by the reciprocal of U11[c][c] (dinv, zero pivots left
// undivided), as a left-looking substitution would.
template <typename T>
__device__ __forceinline__ void panel_rows_body(cpx<T>* x, const cpx<T>* a11, const cpx<T>* dinv, int w) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k < w) {
            x[k] = cmul(x[k], dinv[k]);
#pragma unroll
            for (int c = k + 1; c < NB; ++c) {
                if (c < w) {
                    const cpx<T> u = a11[k * NB + c];
                    x[c].r -= x[k].r * u.r - x[k].i * u.i;
                    x[c].i -= x[k].r * u.i + x[k].i * u.r;
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void k_panel_rows(T* A, size_t strideA, int lda, int d, int w, int rlo,
                                                    int rhi, const int* __restrict__ nElem) {
    __shared__ cpx<T> a11[NB * NB];
    __shared__ cpx<T> dinv[NB];
    const int b = blockIdx.y;
    // padded matrix: rows in the padding are zero left of the diagonal
    if (nElem && rlo + (int)blockIdx.x * 256 >= nElem[b]) return;
    T* Ab = A + (size_t)b * strideA;
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        int i = e / w, k = e % w;
        a11[i * NB + k] = ld(Ab + 2 * ((size_t)(d + i) * lda + d + k));
    }
    if (tid < w) {
        // 1 / U11[c][c]; a zero pivot multiplies by one
        const cpx<T> dd = ld(Ab + 2 * ((size_t)(d + tid) * lda + d + tid));
        dinv[tid] = (dd.r == (T)0 && dd.i == (T)0) ? cpx<T>{(T)1, (T)0} : cinv(dd);
    }
    __syncthreads();
    const int row = rlo + blockIdx.x * 256 + tid;
    if (row >= rhi) return;
    T* p = Ab + 2 * ((size_t)row * lda + d);
    cpx<T> x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) x[c] = ld(p + 2 * c);
    if (w == NB)
        panel_rows_body<T>(x, a11, dinv, NB);
    else
        panel_rows_body<T>(x, a11, dinv, w);
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) st(p + 2 * c, x[c]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_panel_cols(T* A, size_t strideA, int lda, int d, int w, int clo,
                                                    int chi) {
    __shared__ cpx<T> a11[NB * NB];
    const int b = blockIdx.y;
    T* Ab = A + (size_t)b * strideA;
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        int i = e / w, k = e % w;
        a11[i * NB + k] = ld(Ab + 2 * ((size_t)(d + i) * lda + d + k));
    }
    __syncthreads();
    const int col = clo + blockIdx.x * 256 + tid;
    if (col >= chi) return;
    cpx<T> y[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) y[c] = ld(Ab + 2 * ((size_t)(d + c) * lda + col));
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < w) {
            cpx<T> s = y[c];
#pragma unroll
            for (int k = 0; k < c; ++k) {
                cpx<T> l = a11[c * NB + k];
                s.r -= l.r * y[k].r - l.i * y[k].i;
                s.i -= l.r * y[k].i + l.i * y[k].r;
            }
            y[c] = s;
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) st(Ab + 2 * ((size_t)(d + c) * lda + col), y[c]);
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

// C[r0:r1, c0:c1) -= A[r0:r1, k0:k0+K) * A[k0:k0+K, c0:c1) for every batch
// matrix (all three blocks live in the same matrix).
//
// One 256-thread block computes a BM x BN tile of C; its four waves own
// (BM/2) x (BN/2) quadrants as TM x TN tiles of 16 x 16, each complex tile
// kept as two real accumulators (re, im) of a 16x16x4 MFMA.  The product
// is accumulated from zero and subtracted from C once at the end: starting
// the accumulators at C would round every partial sum at the magnitude of
// C, which in FP32 costs the LU enough accuracy to stall the refinement on
// ill-conditioned EFIE matrices.  K goes through LDS in chunks of KC complex
// values (128 bytes of a row), double-buffered: the next chunk is loaded
// into registers while the current one feeds the matrix cores.
//
// The reduction index inside a chunk is permuted so that lane (i, g) of an
// MFMA operand (i = lane % 16, g = lane / 16) supplies k = J g + j to the
// j-th of J = KC/4 consecutive MFMAs; A rows and B columns (stored
// transposed in LDS) then give each lane its J operands as one contiguous
// 32-byte read.  Rows are padded by 16 bytes so the eight lanes served per
// LDS cycle hit distinct banks.
//
// Blocks are numbered so that each of the 8 XCDs (which receive block ids
// round-robin) works through a contiguous run of tiles, sharing its A row
// panels and B column panels in its own L2.
template <typename T>
struct gemm_vec;
template <>
struct gemm_vec<float> {
    typedef float4 v;
};
template <>
struct gemm_vec<double> {
    typedef double2 v;
};

// FP64: 158 VGPRs + 96 accumulation registers would leave two waves per
// SIMD; asking for three (a few spilled registers) makes the c2_6k FP64
// factorisation's updates 20 % faster (532 -> 426 ms per 51 matrices).
// FP32: four waves (124 registers, against 132 and three waves by
// default) make the updates 2 % faster alone but the mixed sweep no faster,
// its refinement kernels sharing the CUs; the default is kept.
template <typename T, int BM, int BN, bool M3 = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 1))) void k_gemm(
    T* A, size_t strideA, int lda, int r0, int r1, int c0,
                                              int c1, int k0, int K, int gx, int gy, int nblk, int skipUpper,
                                              const int* __restrict__ nElem) {
    typedef typename mf<T>::v4 v4;
    typedef typename gemm_vec<T>::v vec;
    constexpr int CB = (int)sizeof(cpx<T>);   // bytes per complex value
    constexpr int KC = 128 / CB;              // complex values per chunk row
    constexpr int J = KC / 4;                 // MFMAs per chunk per component
    constexpr int LDK = KC + 16 / CB;         // padded LDS row (complex values)
    constexpr int TM = BM / 32, TN = BN / 32;
    constexpr int VPR = 128 / 16;             // 16-byte vectors per chunk row of A
    constexpr int AV = BM * VPR / 256;        // A vectors per thread
    constexpr int BE = KC * BN / 256;         // B complex values per thread
    constexpr int CPV = 16 / CB;              // complex values per vector
    __shared__ cpx<T> As[2][BM][LDK];
    __shared__ cpx<T> Bs[2][BN][LDK];

    // XCD-aware tile order
    int id = blockIdx.x;
    {
        const int q = nblk >> 3, rr = nblk & 7, x = id & 7, loc = id >> 3;
        id = x < rr ? x * (q + 1) + loc : rr * (q + 1) + (x - rr) * q + loc;
    }
    const int per = gx * gy;
    const int b = id / per;
    int t = id - b * per;
    int bx, by;
    {
        // bands of 8 tile rows, column-major inside a band: the tiles an
        // XCD runs at once share a few A row panels and B column panels in
        // its L2 (3 % faster updates than row-major runs)
        const int band = t / (8 * gx), tb = t - band * (8 * gx);
        const int rows = min(8, gy - band * 8);
        by = band * 8 + tb % rows;
        bx = tb / rows;
    }
    T* Ab = A + (size_t)b * strideA;
    const int rb = r0 + by * BM;
    const int cb = c0 + bx * BN;
    // symmetric factorisation: tiles strictly above the diagonal are never
    // read again (U comes from L, and the panels read the lower triangle)
    if (skipUpper && rb + BM <= cb) return;
    // a padded batch matrix (order nElem[b] in the leading block, identity
    // beyond): a tile with its rows or its columns in the padding multiplies
    // zeros and keeps its values
    if (nElem) {
        const int ne = nElem[b];
        if (rb >= ne || cb >= ne) return;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = (wv >> 1) * (BM / 2), wc = (wv & 1) * (BN / 2);
    const int li = lane & 15, lg = lane >> 4;

    // 4M: accR = sum ar br - ai bi, accI = sum ar bi + ai br.
    // 3M: accR = sum ar br, accI = sum ai bi, acc3 = sum (ar+ai)(br+bi);
    // re = accR - accI, im = acc3 - accR - accI.
    v4 accR[TM][TN], accI[TM][TN], acc3[M3 ? TM : 1][M3 ? TN : 1];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) {
            accR[x][y] = v4{0, 0, 0, 0};
            accI[x][y] = v4{0, 0, 0, 0};
            if constexpr (M3) acc3[x][y] = v4{0, 0, 0, 0};
        }

    // Operand loads are unconditional, from addresses clamped into the
    // operand blocks; out-of-range values are zeroed when the registers go
    // to LDS.  A load inside a divergent branch, or a select right after
    // it, makes the compiler wait for the load before the matrix-core
    // work, which would serialise the prefetch.
    vec ra[AV];
    cpx<T> rbv[BE];
    auto gload = [&](int kc) {
#pragma unroll
        for (int u = 0; u < AV; ++u) {
            const int e = tid + 256 * u;
            const int i = e / VPR, k = (e % VPR) * CPV;
            const int row = rb + i, kk = kc + k;
            const int rowc = row < r1 ? row : r1 - 1;
            ra[u] = *reinterpret_cast<const vec*>(Ab + 2 * ((size_t)rowc * lda + k0 + (kk < K ? kk : 0)));
        }
#pragma unroll
        for (int u = 0; u < BE; ++u) {
            const int e = tid + 256 * u;
            const int k = e / BN, j = e % BN;
            const int col = cb + j, kk = kc + k;
            rbv[u] = ld(Ab + 2 * ((size_t)(k0 + (kk < K ? kk : K - 1)) * lda + (col < c1 ? col : c1 - 1)));
        }
    };
    auto sstore = [&](int buf, int kc) {
#pragma unroll
        for (int u = 0; u < AV; ++u) {
            const int e = tid + 256 * u;
            const int i = e / VPR, k = (e % VPR) * CPV;
            const int kk = kc + k;
            const bool ok = rb + i < r1 && kk < K;
            vec v = ra[u];
            if constexpr (CPV == 2) {
                const bool ok2 = ok && kk + 1 < K;
                v.x = ok ? v.x : 0.f;
                v.y = ok ? v.y : 0.f;
                v.z = ok2 ? v.z : 0.f;
                v.w = ok2 ? v.w : 0.f;
            } else {
                v.x = ok ? v.x : 0.0;
                v.y = ok ? v.y : 0.0;
            }
            *reinterpret_cast<vec*>(&As[buf][i][k]) = v;
        }
#pragma unroll
        for (int u = 0; u < BE; ++u) {
            const int e = tid + 256 * u;
            const int k = e / BN, j = e % BN;
            const bool ok = cb + j < c1 && kc + k < K;
            cpx<T> v = rbv[u];
            v.r = ok ? v.r : (T)0;
            v.i = ok ? v.i : (T)0;
            Bs[buf][j][k] = v;
        }
    };

    const int nch = (K + KC - 1) / KC;
    gload(0);
    sstore(0, 0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nch) gload((ch + 1) * KC);
        cpx<T> bf[TN][J];
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int j = 0; j < J; ++j) bf[y][j] = Bs[buf][wc + 16 * y + li][J * lg + j];
#pragma unroll
        for (int x = 0; x < TM; ++x) {
            cpx<T> af[J];
#pragma unroll
            for (int j = 0; j < J; ++j) af[j] = As[buf][wr + 16 * x + li][J * lg + j];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                if constexpr (M3) {
                    const T as = af[j].r + af[j].i;
#pragma unroll
                    for (int y = 0; y < TN; ++y) {
                        accR[x][y] = mf<T>::mfma(af[j].r, bf[y][j].r, accR[x][y]);
                        accI[x][y] = mf<T>::mfma(af[j].i, bf[y][j].i, accI[x][y]);
                        acc3[x][y] = mf<T>::mfma(as, bf[y][j].r + bf[y][j].i, acc3[x][y]);
                    }
                } else {
                    const T ni = -af[j].i;
#pragma unroll
                    for (int y = 0; y < TN; ++y) {
                        accR[x][y] = mf<T>::mfma(af[j].r, bf[y][j].r, accR[x][y]);
                        accR[x][y] = mf<T>::mfma(ni, bf[y][j].i, accR[x][y]);
                        accI[x][y] = mf<T>::mfma(af[j].r, bf[y][j].i, accI[x][y]);
                        accI[x][y] = mf<T>::mfma(af[j].i, bf[y][j].r, accI[x][y]);
                    }
                }
            }
        }
        if (ch + 1 < nch) sstore(buf ^ 1, (ch + 1) * KC);
        __syncthreads();
    }

    // C -= product, one row of tiles at a time: all loads of the row are
    // issued (clamped, unconditionally) before the first store
#pragma unroll
    for (int x = 0; x < TM; ++x) {
        cpx<T> cv[TN][4];
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb + wr + 16 * x + mf<T>::row(lg, v);
                const int col = cb + wc + 16 * y + li;
                cv[y][v] = ld(Ab + 2 * ((size_t)(row < r1 ? row : r1 - 1) * lda + (col < c1 ? col : c1 - 1)));
            }
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb + wr + 16 * x + mf<T>::row(lg, v);
                const int col = cb + wc + 16 * y + li;
                T pr = accR[x][y][v], pi = accI[x][y][v];
                if constexpr (M3) {
                    const T t1 = accR[x][y][v], t2 = accI[x][y][v];
                    pr = t1 - t2;
                    pi = acc3[x][y][v] - t1 - t2;
                }
                if (row < r1 && col < c1)
                    st(Ab + 2 * ((size_t)row * lda + col), cpx<T>{cv[y][v].r - pr, cv[y][v].i - pi});
            }
    }
}

// ------------------------------------------- split-bf16 trailing update --
// FP32 C -= A B on the bf16 matrix cores.  Before the trailing update of an
// outer block, k_split6 splits every FP32 value of L21 and U12 once into
// three bf16 parts by round-to-nearest, x = hi + mid + lo + e with
// |mid| <= 2^-8 |x|, |lo| <= 2^-16 |x| and |e| <= 2^-25 |x|, stored as six
// planes (re/im x hi/mid/lo) of [outer index][k] with k contiguous: rows of
// L21, columns of U12 (transposed), both zero-padded to whole tiles and
// 32-wide k chunks.  k_gemm_p6 then accumulates, for each of the four real
// products of the complex product (RR = ar br, II = ai bi, RI = ar bi,
// IR = ai br), the six part products down to 2^-16 relative (hh, hm, mh,
// hl, mm, lh) in FP32 with v_mfma_f32_16x16x32_bf16; the dropped ones are
// below 2^-23 relative and of random sign, the size of FP32 rounding, so
// the update is as accurate as the FP32 MFMA one at half its matrix-core
// time.  MFMA operand layout (measured on gfx950): A[i][k] in lane
// i + 16 (k / 8), vector slot k % 8; B[k][j] in lane j + 16 (k / 8); the
// result as for the f32 16x16 instructions.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int P6_KC = 32;               // k per chunk (one MFMA k step)
constexpr int P6_LDK = 40;              // bf16 per LDS row: 32 + 8 (80-byte rows spread the banks)
constexpr int P6_PL = 64 * P6_LDK;      // bf16 per LDS plane

// round-to-nearest-even FP32 -> bf16 bits (finite inputs)
__device__ __forceinline__ uint32_t bf16_rn(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_f(uint32_t b) { return __uint_as_float(b << 16); }

// Plane layout of one batch element: 6 planes of [nOuter][ldp] bf16,
// nOuter a multiple of 64, ldp (= padded K) a multiple of 32.
// T = false: plane row o holds A[r0 + o][k0 .. k0 + K) (rows of L21);
// T = true:  plane row o holds A[k0 .. k0 + K)[c0 + o]  (columns of U12).
// One thread per 8 consecutive k of one plane row; values outside
// o < nOut, k < K are zero.
template <bool T>
__global__ __launch_bounds__(256) void k_split6(const float* A, size_t strideA, int lda, int base, int nOut,
                                                int k0, int K, uint16_t* planes, size_t batchStride, int nOuter,
                                                int ldp) {
    const int b = blockIdx.y;
    const int gpr = ldp / 8;  // k groups per plane row
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nOuter * gpr) return;
    const int o = e / gpr, kk = (e % gpr) * 8;
    const float* Ab = A + (size_t)b * strideA;
    uint32_t hr[8], mr[8], lr[8], hi[8], mi[8], li[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = kk + s;
        cpx<float> v{0.0f, 0.0f};
        if (o < nOut && k < K)
            v = T ? ld(Ab + 2 * ((size_t)(k0 + k) * lda + base + o)) : ld(Ab + 2 * ((size_t)(base + o) * lda + k0 + k));
        hr[s] = bf16_rn(v.r);
        const float r1 = v.r - bf16_f(hr[s]);
        mr[s] = bf16_rn(r1);
        lr[s] = bf16_rn(r1 - bf16_f(mr[s]));
        hi[s] = bf16_rn(v.i);
        const float i1 = v.i - bf16_f(hi[s]);
        mi[s] = bf16_rn(i1);
        li[s] = bf16_rn(i1 - bf16_f(mi[s]));
    }
    auto pk = [](const uint32_t* q) {
        return make_uint4(q[0] | (q[1] << 16), q[2] | (q[3] << 16), q[4] | (q[5] << 16), q[6] | (q[7] << 16));
    };
    const size_t pl = (size_t)nOuter * ldp;
    uint16_t* P = planes + (size_t)b * batchStride + (size_t)o * ldp + kk;
    *reinterpret_cast<uint4*>(P + 0 * pl) = pk(hr);
    *reinterpret_cast<uint4*>(P + 1 * pl) = pk(mr);
    *reinterpret_cast<uint4*>(P + 2 * pl) = pk(lr);
    *reinterpret_cast<uint4*>(P + 3 * pl) = pk(hi);
    *reinterpret_cast<uint4*>(P + 4 * pl) = pk(mi);
    *reinterpret_cast<uint4*>(P + 5 * pl) = pk(li);
}

__device__ __forceinline__ mf<float>::v4 mfma_bf(const uint4& a, const uint4& b, mf<float>::v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

// C[r0:r1, c0:c1) -= L21 U12 from the planes: PA rows r - r0 (nOuterA
// rows per plane), PB rows c - cBase (nOuterB), ldp = padded K.  Tile
// 64 x 64 per 256-thread block, four waves of 32 x 32 (2 x 2 MFMA tiles),
// a chunk of 32 k staged in LDS (single buffer, the next chunk's 16-byte
// plane runs prefetched into registers); XCD-aware block order as k_gemm.
__global__ __launch_bounds__(256, 2) void k_gemm_p6(float* A, size_t strideA, int lda, int r0, int r1, int c0, int c1,
                                                    const uint16_t* PA, const uint16_t* PB, size_t batchStride,
                                                    int nOuterA, int nOuterB, int cBase, int ldp, int gx, int gy,
                                                    int nblk) {
    typedef mf<float>::v4 v4;
    __shared__ uint4 As[6][P6_PL / 8];
    __shared__ uint4 Bs[6][P6_PL / 8];
    int id = blockIdx.x;
    {
        const int q = nblk >> 3, rr = nblk & 7, x = id & 7, loc = id >> 3;
        id = x < rr ? x * (q + 1) + loc : rr * (q + 1) + (x - rr) * q + loc;
    }
    const int bx = id % gx;
    id /= gx;
    const int by = id % gy;
    const int b = id / gy;
    float* Ab = A + (size_t)b * strideA;
    const int rb = r0 + by * 64, cb = c0 + bx * 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    const int li = lane & 15, lg = lane >> 4;
    // loader: plane row tid / 4 of the tile, k run 8 (tid % 4)
    const int lrow = tid >> 2, lk = (tid & 3) * 8;
    const size_t plA = (size_t)nOuterA * ldp, plB = (size_t)nOuterB * ldp;
    const uint16_t* pa = PA + (size_t)b * batchStride + (size_t)(rb - r0 + lrow) * ldp + lk;
    const uint16_t* pb = PB + (size_t)b * batchStride + (size_t)(cb - cBase + lrow) * ldp + lk;

    v4 acc[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[x][y][c] = v4{0, 0, 0, 0};

    constexpr int PAi[6] = {0, 0, 1, 0, 1, 2}, PBi[6] = {0, 1, 0, 2, 1, 0};
    const int nch = ldp / P6_KC;
    const int so = (lrow * P6_LDK + lk) / 8;
    uint4 ra0 = *reinterpret_cast<const uint4*>(pa), ra1 = *reinterpret_cast<const uint4*>(pa + plA),
          ra2 = *reinterpret_cast<const uint4*>(pa + 2 * plA), ra3 = *reinterpret_cast<const uint4*>(pa + 3 * plA),
          ra4 = *reinterpret_cast<const uint4*>(pa + 4 * plA), ra5 = *reinterpret_cast<const uint4*>(pa + 5 * plA);
    uint4 rb0 = *reinterpret_cast<const uint4*>(pb), rb1 = *reinterpret_cast<const uint4*>(pb + plB),
          rb2 = *reinterpret_cast<const uint4*>(pb + 2 * plB), rb3 = *reinterpret_cast<const uint4*>(pb + 3 * plB),
          rb4 = *reinterpret_cast<const uint4*>(pb + 4 * plB), rb5 = *reinterpret_cast<const uint4*>(pb + 5 * plB);
    for (int ch = 0; ch < nch; ++ch) {
        __syncthreads();  // the previous chunk's operands have been read
        As[0][so] = ra0; As[1][so] = ra1; As[2][so] = ra2; As[3][so] = ra3; As[4][so] = ra4; As[5][so] = ra5;
        Bs[0][so] = rb0; Bs[1][so] = rb1; Bs[2][so] = rb2; Bs[3][so] = rb3; Bs[4][so] = rb4; Bs[5][so] = rb5;
        __syncthreads();
        {
            // the next chunk (the last one again at the end: loads stay
            // unconditional and in bounds)
            const int kc = (ch + 1 < nch ? ch + 1 : ch) * P6_KC;
            ra0 = *reinterpret_cast<const uint4*>(pa + kc);
            ra1 = *reinterpret_cast<const uint4*>(pa + plA + kc);
            ra2 = *reinterpret_cast<const uint4*>(pa + 2 * plA + kc);
            ra3 = *reinterpret_cast<const uint4*>(pa + 3 * plA + kc);
            ra4 = *reinterpret_cast<const uint4*>(pa + 4 * plA + kc);
            ra5 = *reinterpret_cast<const uint4*>(pa + 5 * plA + kc);
            rb0 = *reinterpret_cast<const uint4*>(pb + kc);
            rb1 = *reinterpret_cast<const uint4*>(pb + plB + kc);
            rb2 = *reinterpret_cast<const uint4*>(pb + 2 * plB + kc);
            rb3 = *reinterpret_cast<const uint4*>(pb + 3 * plB + kc);
            rb4 = *reinterpret_cast<const uint4*>(pb + 4 * plB + kc);
            rb5 = *reinterpret_cast<const uint4*>(pb + 5 * plB + kc);
        }
        uint4 af[2][6];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int p = 0; p < 6; ++p) af[x][p] = As[p][((wr + 16 * x + li) * P6_LDK + 8 * lg) / 8];
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            uint4 bf[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) bf[p] = Bs[p][((wc + 16 * y + li) * P6_LDK + 8 * lg) / 8];
#pragma unroll
            for (int q = 0; q < 6; ++q)
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    acc[x][y][0] = mfma_bf(af[x][PAi[q]], bf[PBi[q]], acc[x][y][0]);          // ar br
                    acc[x][y][1] = mfma_bf(af[x][3 + PAi[q]], bf[3 + PBi[q]], acc[x][y][1]);  // ai bi
                    acc[x][y][2] = mfma_bf(af[x][PAi[q]], bf[3 + PBi[q]], acc[x][y][2]);      // ar bi
                    acc[x][y][3] = mfma_bf(af[x][3 + PAi[q]], bf[PBi[q]], acc[x][y][3]);      // ai br
                }
        }
    }

    // C -= product, one row of tiles at a time (loads first, clamped)
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        cpx<float> cv[2][4];
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb + wr + 16 * x + 4 * lg + v, col = cb + wc + 16 * y + li;
                cv[y][v] = ld(Ab + 2 * ((size_t)(row < r1 ? row : r1 - 1) * lda + (col < c1 ? col : c1 - 1)));
            }
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb + wr + 16 * x + 4 * lg + v, col = cb + wc + 16 * y + li;
                const float pr = acc[x][y][0][v] - acc[x][y][1][v], pi = acc[x][y][2][v] + acc[x][y][3][v];
                if (row < r1 && col < c1)
                    st(Ab + 2 * ((size_t)row * lda + col), cpx<float>{cv[y][v].r - pr, cv[y][v].i - pi});
            }
    }
}

// ------------------------------------------------- triangular solves -----
// Blocked triangular solve on a vector x (element e of batch b at
// X[b*strideX + 2*e*xs]), upper (U x = y, non-unit diagonal, bottom block
// first) or unit lower (L y = r, top block first), in 64-row blocks.
//
// Single-launch blocked triangular solve.  Workgroups take tickets in
// dispatch order; ticket v solves block v (lower, top first) or block
// nblk-1-v (upper, bottom first).  The grid is (batch, nblk), so dispatch
// walks the matrices first: the resident workgroups hold the leading blocks
// of every matrix and all chains advance together (with blocks first, the
// chip filled with the blocks of a few matrices, most of them waiting).
// A workgroup subtracts the contribution of the earlier blocks as they are
// published: it waits for the next one it needs, then takes the run of
// blocks already published after it (flags are set in ticket order), up to
// kTrsvRun at a time, loads their x into LDS with one round of loads and
// applies them; then it solves its diagonal block and publishes its x with
// a flag.  It only waits on workgroups with smaller tickets, which were
// dispatched before it, so the waits always end; a bounded spin reports an
// error instead of waiting forever.
//
// x and the flags are only read and written through agent-scope relaxed
// atomics (global_load / global_store with sc1 on gfx950), which are
// coherent across the XCDs' L2 caches without cache maintenance.  The
// order between a block's x and its flag comes from the waits alone: every
// thread waits until its x stores have completed (s_waitcnt 0) before the
// barrier after which thread 0 stores the flag, and a reader issues its x
// loads only after its flag load has returned the epoch (the spin waits for
// each load) and the barrier that follows.  Agent-scope fences would add an
// L2 writeback per published block and an L2 invalidate per wait; the
// invalidates also throw away the cached operands of the trailing-update
// GEMM that runs beside the solve on the same XCD (without them: c2_6k sweep
// 3.5 %, c4 4.5 % faster, results bit-identical).  The L values are
// written by earlier kernels and read with plain loads.
constexpr long long kSpinLimit = 1LL << 28;
constexpr int kTrsvRun = 8;

// x from the lane the DPP control CTRL names within the 16-lane row
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// sum over the 16 lanes of a row, the same bits in every lane: swap
// neighbours (quad_perm 1,0,3,2), pairs (quad_perm 2,3,0,1), mirror the
// half rows, mirror the row; each step adds two equal partial sums in
// either order
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    return v;
}

// element j of a wave-held vector, read from lane j (j wave-uniform)
__device__ __forceinline__ float lane_read(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_read(double v, int j) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// SHARED: the solve runs beside a factorisation, whose update GEMM holds
// 128 registers per wave; the FP32 solve then keeps to 128 too (a few
// spills), so the GEMM keeps three waves on the SIMDs it shares.
template <typename T, bool LOWER, bool SHARED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && SHARED ? 4 : 1))) void k_trsv_fused(const T* A, size_t strideA, int lda, int n, T* X,
                                                    size_t strideX, size_t xs, int* flags, int* tickets,
                                                    int epoch, const int* skip, int* err) {
    __shared__ cpx<T> u[64][65];
    __shared__ cpx<T> xb[64];
    __shared__ cpx<T> xc[kTrsvRun][64];
    __shared__ cpx<T> part[2][64];
    __shared__ int s_vid, s_avail;
    const int b = blockIdx.x;
    if (skip && skip[b]) return;
    const int tid = threadIdx.x;
    const int nblk = (n + 63) / 64;
    // row blocks are taken in order (tickets), one after another, until
    // none is left: a row block waits only on earlier ones, which running
    // blocks hold, so the solve drains with any number of blocks
    for (;;) {
        if (tid == 0) s_vid = atomicAdd(tickets + b, 1);
        __syncthreads();
        const int vid = __builtin_amdgcn_readfirstlane(s_vid);
        if (vid >= nblk) break;
        const int bk = LOWER ? vid : nblk - 1 - vid;
        const int r0 = bk * 64;
        const int bs = n - r0 < 64 ? n - r0 : 64;
        const T* Ab = A + (size_t)b * strideA;
        T* Xb = X + (size_t)b * strideX;
        int* fl = flags + (size_t)b * nblk;
        // diagonal block: all 16 loads per thread in flight at once (clamped
        // into the block, kept only where inside)
        {
            // thread (row group g, column k) takes rows g + 4 u2
            cpx<T> v[16];
            const int g = __builtin_amdgcn_readfirstlane(tid >> 6), k = tid & 63, kc = k < bs ? k : bs - 1;
    #pragma unroll
            for (int u2 = 0; u2 < 16; ++u2) {
                const int ii = g + 4 * u2;
                v[u2] = ld(Ab + 2 * ((size_t)(r0 + (ii < bs ? ii : bs - 1)) * lda + r0 + kc));
            }
    #pragma unroll
            for (int u2 = 0; u2 < 16; ++u2) {
                const int ii = g + 4 * u2;
                if (ii < bs && k < bs) u[ii][k] = v[u2];
            }
        }
        for (int i = tid; i < bs; i += 256) xb[i] = ld(Xb + 2 * (size_t)(r0 + i) * xs);
        // thread (wave w, half h, pair p) accumulates, for its 8 rows
        // 16w + 2k + h, the products of columns 2p and 2p+1 of every earlier
        // block column with x; one load instruction of a wave then covers two
        // contiguous 64-column row segments.  The 32 lanes of a half are summed
        // once, after the last block column.
        const int w = __builtin_amdgcn_readfirstlane(tid >> 6), h = (tid >> 5) & 1, p = tid & 31;
        cpx<T> acc[8];
    #pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = {0, 0};
        auto blk = [&](int st) { return LOWER ? st : nblk - 1 - st; };
        // columns 2p, 2p+1 of this thread's rows in block column blk(st); the
        // pair start is clamped into the block's columns (the row has lda > n
        // entries, so the pair's second column stays in the row).  Columns past
        // the block are zeroed where the values are used, so no select follows
        // the loads (that would make the prefetch wait for them).
        auto load_m = [&](int st, cpx<T>* m) {
            // an opaque zero: the lane offsets below are formed here, at
            // each call, and not hoisted out of the loops into registers
            // held for the whole kernel
            int z0 = 0;
            asm volatile("" : "+s"(z0));
            const int c0 = blk(st) * 64;
            const int cs = n - c0 < 64 ? n - c0 : 64;
            const int jp = 2 * p < cs ? 2 * p : ((cs - 1) & ~1);
            if (bs == 64) {
                // full block: a wave-uniform base (scalar registers) and one
                // 32-bit lane offset per row pair
                const char* rb = reinterpret_cast<const char*>(Ab + 2 * ((size_t)(r0 + 16 * w) * lda + c0));
                const unsigned o = (unsigned)(2 * (h * lda + jp + z0)) * (unsigned)sizeof(T);
                const size_t rs = (size_t)(4 * lda) * sizeof(T);
    #pragma unroll
                for (int k = 0; k < 8; ++k)
                    ld_pair(reinterpret_cast<const T*>((rb + k * rs) + o), m[2 * k], m[2 * k + 1]);
            } else {
                // last row block: rows past it read its last row; the lane's
                // first row and how far its rows may step are all it keeps
                const int r1 = 16 * w + h + z0 < bs ? 16 * w + h + z0 : bs - 1;
                const int lim = bs - 1 - r1;
                const T* rl = Ab + 2 * ((size_t)(r0 + r1) * lda + c0 + jp);
    #pragma unroll
                for (int k = 0; k < 8; ++k)
                    ld_pair(rl + 2 * (size_t)(2 * k < lim ? 2 * k : lim) * lda, m[2 * k], m[2 * k + 1]);
            }
        };
        cpx<T> m[16], mn[16];
        for (int step = 0; step < vid;) {
            load_m(step, m);  // in flight during the wait
            if (tid == 0) {
                long long spins = 0;
                while (__hip_atomic_load(fl + blk(step), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kSpinLimit) {
                        atomicExch(err, 1);
                        break;
                    }
                }
                int av = step + 1;
                while (av < vid && av < step + kTrsvRun &&
                       __hip_atomic_load(fl + blk(av), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
                    ++av;
                // the x loads below are coherent atomics issued after this
                // flag load returned: no cache invalidation needed
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                s_avail = av;
            }
            __syncthreads();
            acquire_wg();
            const int av = __builtin_amdgcn_readfirstlane(s_avail);
            // x of the published run, one round of loads
            for (int e = tid; e < (av - step) * 64; e += 256) {
                const int c0 = blk(step + (e >> 6)) * 64, j = e & 63;
                cpx<T> v = {0, 0};
                if (c0 + j < n) v = ald_c(Xb + 2 * (size_t)(c0 + j) * xs);
                xc[e >> 6][j] = v;
            }
            __syncthreads();
            for (int st = step; st < av; ++st) {
                if (st + 1 < av) load_m(st + 1, mn);  // next block's values before this one's products
                const cpx<T> x0 = xc[st - step][2 * p], x1 = xc[st - step][2 * p + 1];
                const int cs = n - blk(st) * 64;
                if (cs < 64) {
                    // last block column: whatever the padding columns hold
    #pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        if (2 * p >= cs) m[2 * k] = {0, 0};
                        if (2 * p + 1 >= cs) m[2 * k + 1] = {0, 0};
                    }
                }
    #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const cpx<T> a = m[2 * k], c = m[2 * k + 1];
                    acc[k].r += a.r * x0.r - a.i * x0.i + c.r * x1.r - c.i * x1.i;
                    acc[k].i += a.r * x0.i + a.i * x0.r + c.r * x1.i + c.i * x1.r;
                }
                if (st + 1 < av) {
    #pragma unroll
                    for (int k = 0; k < 16; ++k) m[k] = mn[k];
                }
            }
            __syncthreads();  // xc is refilled by the next run
            step = av;
        }
        // each 16-lane row of pairs sums its partials in registers; the two
        // rows of a half meet in part
    #pragma unroll
        for (int k = 0; k < 8; ++k) {
            acc[k].r = row16_sum(acc[k].r);
            acc[k].i = row16_sum(acc[k].i);
        }
        if ((p & 15) == 0) {
    #pragma unroll
            for (int k = 0; k < 8; ++k) part[p >> 4][16 * w + 2 * k + h] = acc[k];
        }
        __syncthreads();
        // diagonal solve by wave 0 with x in registers (lane k holds x_k);
        // x_j reaches every lane by readlane, so the steps need no LDS round
        // trip or barrier.  The sequential chain of a step is the readlane
        // and the update: each lane's diagonal reciprocal (upper) is formed
        // before the steps, and a lane's matrix entries are read from LDS
        // eight steps at a time, the next eight in flight during these.
        if (tid < 64) {
            cpx<T> v = {0, 0};
            if (tid < bs) {
                v = xb[tid];
                v.r -= part[0][tid].r + part[1][tid].r;
                v.i -= part[0][tid].i + part[1][tid].i;
            }
            cpx<T> dinv = {(T)1, (T)0};
            bool dz = true;
            if (!LOWER && tid < bs) {
                const cpx<T> d = u[tid][tid];
                dz = d.r == (T)0 && d.i == (T)0;
                if (!dz) dinv = cinv(d);
            }
            // column of step jj (clamped where jj is past the block: unused);
            // the opaque zero keeps these LDS reads from being hoisted above
            // the update loop, where their registers would be held
            int z1 = 0;
            asm volatile("" : "+s"(z1));
            auto col = [&](int jj) { return (jj < bs ? (LOWER ? jj : bs - 1 - jj) : 0) + z1; };
            cpx<T> lc[8], ln[8];
    #pragma unroll
            for (int q = 0; q < 8; ++q) lc[q] = u[tid][col(q)];
            for (int c0 = 0; c0 < bs; c0 += 8) {
                if (c0 + 8 < bs) {
    #pragma unroll
                    for (int q = 0; q < 8; ++q) ln[q] = u[tid][col(c0 + 8 + q)];
                }
    #pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (c0 + q < bs) {
                        const int j = col(c0 + q);
                        cpx<T> xj;
                        if (LOWER) {
                            xj.r = lane_read(v.r, j);
                            xj.i = lane_read(v.i, j);
                        } else {
                            // x_j = v_j / d_j, formed in lane j
                            const cpx<T> wv = dz ? v : cmul(v, dinv);
                            xj.r = lane_read(wv.r, j);
                            xj.i = lane_read(wv.i, j);
                            if (tid == j) v = xj;
                        }
                        const bool act = LOWER ? (tid > j && tid < bs) : (tid < j);
                        if (act) {
                            const cpx<T> l = lc[q];
                            v.r -= l.r * xj.r - l.i * xj.i;
                            v.i -= l.r * xj.i + l.i * xj.r;
                        }
                    }
                }
    #pragma unroll
                for (int q = 0; q < 8; ++q) lc[q] = ln[q];
            }
            if (tid < bs) xb[tid] = v;
        }
        __syncthreads();
        for (int k = tid; k < bs; k += 256) ast_c(Xb + 2 * (size_t)(r0 + k) * xs, xb[k]);
        // this thread's x stores complete before the barrier and the flag
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(fl + bk, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();  // LDS and s_vid are reused by the next row block
    }
}

// perm[b][i] = source row of position i after the sequential swaps of ipiv
// (applying ipiv to a vector v gives v_new[i] = v[perm[i]]).  The swaps run
// in LDS (n ints, up to the device's per-block limit); INLDS false swaps in
// place in perm itself (one thread, global memory: slower, any n).
template <bool INLDS>
__global__ __launch_bounds__(256) void k_make_perm(const int* ipiv, int n, int* perm) {
    extern __shared__ int sh[];
    const int b = blockIdx.x;
    int* w = INLDS ? sh : perm + (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int* ip = ipiv + (size_t)b * n;
        for (int i = 0; i < n; ++i) {
            int q = ip[i] - 1;
            int t = w[i];
            w[i] = w[q];
            w[q] = t;
        }
    }
    if (INLDS) {
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) perm[(size_t)b * n + i] = sh[i];
    }
}

// ---------------------------------------------------------- drivers ------

int lu_tree_nodes(int n) { return tree_geom(n).nNodes; }
int lu_cand_stride() { return NB + 1; }
int lu_swap_stride() { return kSwapSlots * lu_swap_stride_c; }


template <typename T, int BM, int BN>
static void gemm_launch(T* A, size_t strideA, int lda, int batch, int r0, int r1, int c0, int c1, int k0, int K,
                        hipStream_t s, int skipUpper = 0, const int* nElem = nullptr) {
    const int gx = (c1 - c0 + BN - 1) / BN, gy = (r1 - r0 + BM - 1) / BM;
    const int nblk = gx * gy * batch;
    hipLaunchKernelGGL((k_gemm<T, BM, BN, true>), dim3(nblk), dim3(256), 0, s, A, strideA, lda, r0, r1, c0, c1, k0,
                       K, gx, gy, nblk, skipUpper, nElem);
}

// 64 x 64 tiles: three blocks per CU (FP32: 84 VGPRs + 48 accumulation
// registers) keep 12 waves resident, which hides the operand and C latency
// better than larger tiles at lower occupancy (measured at n = 6009,
// K = 128: 112 TFLOP/s FP32 against 96 for 128 x 128).  A fourth block
// (the registers capped at 128) gained about 1 % on the update and
// nothing on the sweep.  The complex product uses three real products (3M:
// ar br, ai bi, (ar+ai)(br+bi)); with K = 256 the FP32 update runs at
// 145 TFLOP/s counted as 8 M N K.
// Returns the number of kernels launched (0 for an empty update).
template <typename T>
static int gemm(T* A, size_t strideA, int lda, int batch, int r0, int r1, int c0, int c1, int k0, int K,
                hipStream_t s, int skipUpper = 0, const int* nElem = nullptr) {
    if (r1 <= r0 || c1 <= c0 || K <= 0) return 0;
    gemm_launch<T, 64, 64>(A, strideA, lda, batch, r0, r1, c0, c1, k0, K, s, skipUpper, nElem);
    return 1;
}

// Plane geometry of the split-bf16 update for an n x n factorisation:
// per batch element the L21 planes (rows padded to 64) then the U12 planes
// (n + 1 columns padded to 64), 6 planes each of ldp = WB bf16 per row.
static int p6_outer(int n) { return (n + 1 + 63) / 64 * 64; }
size_t lu_split_bytes(int n) { return (size_t)2 * 6 * p6_outer(n) * WB * sizeof(uint16_t); }

// Trailing update A[r0:n, c0:c1) -= L21 U12 of the outer block j0..j0+wb
// from the planes (split by split_block).
static int gemm_p6(float* A, size_t strideA, int lda, int batch, int r0, int r1, int c0, int c1,
                   const LuWork& wk, int n, int jend, hipStream_t s) {
    if (r1 <= r0 || c1 <= c0) return 0;
    const int no = p6_outer(n);
    const size_t bs = lu_split_bytes(n) / sizeof(uint16_t);
    const uint16_t* PA = static_cast<const uint16_t*>(wk.splitPlanes);
    const uint16_t* PB = PA + (size_t)6 * no * WB;
    const int gx = (c1 - c0 + 63) / 64, gy = (r1 - r0 + 63) / 64;
    const int nblk = gx * gy * batch;
    hipLaunchKernelGGL(k_gemm_p6, dim3(nblk), dim3(256), 0, s, A, strideA, lda, r0, r1, c0, c1, PA, PB, bs, no, no,
                       jend, WB, gx, gy, nblk);
    return 1;
}

// Planes of L21 = A[jend:n, j0:jend) and U12 = A[j0:jend, jend:n+1).
static int split_block(float* A, size_t strideA, int lda, int batch, int n, int j0, int jend, const LuWork& wk,
                       hipStream_t s) {
    const int no = p6_outer(n), wb = jend - j0;
    const size_t bs = lu_split_bytes(n) / sizeof(uint16_t);
    uint16_t* PA = static_cast<uint16_t*>(wk.splitPlanes);
    uint16_t* PB = PA + (size_t)6 * no * WB;
    const int nA = (n - jend + 63) / 64 * 64, nB = (n + 1 - jend + 63) / 64 * 64;
    const int tA = nA * (WB / 8), tB = nB * (WB / 8);
    hipLaunchKernelGGL(k_split6<false>, dim3((tA + 255) / 256, batch), dim3(256), 0, s, A, strideA, lda, jend,
                       n - jend, j0, wb, PA, bs, no, WB);
    hipLaunchKernelGGL(k_split6<true>, dim3((tB + 255) / 256, batch), dim3(256), 0, s, A, strideA, lda, jend,
                       n + 1 - jend, j0, wb, PB, bs, no, WB);
    return 2;
}

template <typename T>
static int panel_cols(T* A, size_t strideA, int lda, int batch, int d, int w, int clo, int chi, hipStream_t s) {
    if (chi <= clo) return 0;
    hipLaunchKernelGGL(k_panel_cols<T>, dim3((chi - clo + 255) / 256, batch), dim3(256), 0, s, A, strideA, lda, d,
                       w, clo, chi);
    return 1;
}

// Columns [c0, c1) of an outer block of the symmetric factorisation, all
// updates from columns left of c0 applied: halves split at a multiple of
// NB, the left half factored, then U = D L^T of its rows for the right
// half's columns and the right half's update with K = the left half's
// width, then the right half.  Against NB-wide steps each followed by a
// K = NB update of the rest of the block, the flops are the same (with 64
// x 64 update tiles) and the update passes over the block's columns halve.
template <typename T>
static void sym_block(T* A, size_t strideA, int n, int lda, int batch, const LuWork& wk, int c0, int c1,
                      hipStream_t sp, const Prof* pf) {
    if (c1 - c0 <= NB) {
        const int w = c1 - c0;
        pf->begin(7);
        hipLaunchKernelGGL((k_panel_diag<T, true>), dim3(batch), dim3(64), 0, sp, A, strideA, lda, c0, w, wk.info,
                           wk.nElem);
        int np = 1;
        if (c1 < n) {
            hipLaunchKernelGGL(k_panel_rows<T>, dim3((n - c1 + 255) / 256, batch), dim3(256), 0, sp, A, strideA, lda,
                               c0, w, c1, n, wk.nElem);
            ++np;
        }
        pf->end(7, np);
        return;
    }
    const int mid = c0 + (c1 - c0 + 2 * NB - 1) / (2 * NB) * NB;
    sym_block<T>(A, strideA, n, lda, batch, wk, c0, mid, sp, pf);
    pf->begin(7);
    hipLaunchKernelGGL(k_sym_u12<T>, dim3((c1 - mid + 63) / 64, (mid - c0 + 31) / 32, batch), dim3(256), 0, sp, A,
                       strideA, lda, c0, mid, mid, c1, wk.nElem);
    pf->end(7);
    pf->begin(8);
    pf->end(8, gemm<T>(A, strideA, lda, batch, mid, n, mid, c1, c0, mid - c0, sp, 1, wk.nElem));
    sym_block<T>(A, strideA, n, lda, batch, wk, mid, c1, sp, pf);
}

// Panels of the outer block [j0, jend): for every NB-wide sub-panel the
// tournament, its interchanges on the block's own columns, the diagonal
// block, L21, the block-local U12 and the update of the rest of the block.
// The symmetric factorisation takes the block by halves (sym_block).
template <typename T>
static void panel_phase(T* A, size_t strideA, int n, int lda, int batch, const LuWork& wk, int j0, int jend,
                        int nodeStride, hipStream_t sp, const Prof* pf) {
    if (wk.noPivot) {
        sym_block<T>(A, strideA, n, lda, batch, wk, j0, jend, sp, pf);
        return;
    }
    for (int c0 = j0; c0 < jend; c0 += NB) {
        const int w = jend - c0 < NB ? jend - c0 : NB;
        const int slot = (c0 - j0) / NB;
        if (!wk.noPivot) {
            TreeGeom g = tree_geom(n - c0);
            pf->begin(5);
            hipLaunchKernelGGL(k_tslu<T>, dim3(g.levelCount[0], batch), dim3(64), 0, sp, A, strideA, n, lda, c0, w,
                               g, nodeStride, wk.cand, wk.counters, wk.ipiv, wk.info, wk.swaps, slot);
            pf->end(5);
            pf->begin(6);
            hipLaunchKernelGGL(k_laswp<T>, dim3((jend - j0 + 255) / 256, batch), dim3(256), 0, sp, A, strideA, lda,
                               wk.swaps, slot, j0, jend);
            pf->end(6);
        }
        pf->begin(7);
        int np = 1;
        if (wk.noPivot)
            hipLaunchKernelGGL((k_panel_diag<T, true>), dim3(batch), dim3(64), 0, sp, A, strideA, lda, c0, w,
                               wk.info, wk.nElem);
        else
            hipLaunchKernelGGL((k_panel_diag<T, false>), dim3(batch), dim3(64), 0, sp, A, strideA, lda, c0, w,
                               wk.info, nullptr);
        if (c0 + w < n) {
            hipLaunchKernelGGL(k_panel_rows<T>, dim3((n - c0 - w + 255) / 256, batch), dim3(256), 0, sp, A, strideA,
                               lda, c0, w, c0 + w, n, wk.noPivot ? wk.nElem : nullptr);
            ++np;
        }
        if (!wk.noPivot) {
            np += panel_cols<T>(A, strideA, lda, batch, c0, w, c0 + w, jend, sp);
        } else if (c0 + w < jend) {
            // the block-local U12 from L21 as well (the upper triangle is
            // never read in the symmetric factorisation)
            hipLaunchKernelGGL(k_sym_u12<T>, dim3((jend - c0 - w + 63) / 64, (w + 31) / 32, batch), dim3(256), 0,
                               sp, A, strideA, lda, c0, c0 + w, c0 + w, jend, wk.nElem);
            ++np;
        }
        pf->end(7, np);
        pf->begin(8);
        pf->end(8, gemm<T>(A, strideA, lda, batch, c0 + w, n, c0 + w, jend, c0, w, sp, wk.noPivot ? 1 : 0,
                            wk.noPivot ? wk.nElem : nullptr));
    }
}

// ipiv = identity (the symmetric factorisation does not pivot)
__global__ void k_ident_ipiv(int* ipiv, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[(size_t)blockIdx.y * n + i] = i + 1;
}

// perm = identity: the row permutation of a factorisation without
// interchanges (what k_make_perm makes of an identity ipiv, in parallel)
__global__ void k_ident_perm(int* perm, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[(size_t)blockIdx.y * n + i] = i;
}

// The first failure of a sequence of stream and event calls (launch
// errors surface through hipGetLastError at the end of the sequence).
struct HipFirst {
    hipError_t e = hipSuccess;
    void operator()(hipError_t r) {
        if (e == hipSuccess && r != hipSuccess) e = r;
    }
};

// Complex-symmetric factorisation without pivoting: A = L U with U = D L^T,
// L stored below the diagonal and U on and above it.  Outer blocks and
// look-ahead as in lu_factor; the panels of a block are panel_phase without
// the tournament and interchanges; the block row U12 comes from L21 by
// k_sym_u12 instead of elimination; the updates skip the 64 x 64 tiles
// strictly above the diagonal, which are never read again.  The rhs column
// is not carried (the refinement starts from x = 0); the pivots are the
// identity.
template <typename T>
static hipError_t lu_factor_sym(T* A, size_t strideA, int n, int lda, int batch, const LuWork& wk, hipStream_t s,
                                const Prof* pf) {
    HipFirst chk;
    hipStream_t sp = wk.panel ? wk.panel : s;
    // the first block's panels: on the head stream when given (already
    // ordered after the matrices' fill), else on the panel stream after s
    hipStream_t sh = wk.panel && wk.head ? wk.head : sp;
    hipLaunchKernelGGL(k_ident_ipiv, dim3((n + 255) / 256, batch), dim3(256), 0, s, wk.ipiv, n);
    if (sh == sp && sp != s) {
        chk(hipEventRecord(wk.evAhead, s));
        chk(hipStreamWaitEvent(sp, wk.evAhead, 0));
    }
    panel_phase<T>(A, strideA, n, lda, batch, wk, 0, n < WB ? n : WB, 0, sh, pf);
    // with a head stream the first block's update and the second block's
    // panels run there too, in order, once the previous factorisation has
    // reached its last (latency-bound) steps on s
    int j0 = 0;
    if (sh != sp && n > WB) {
        const int jend = WB, jnext = n - jend < WB ? n : jend + WB;
        if (wk.headWait) chk(hipStreamWaitEvent(sh, wk.headWait, 0));
        pf->begin(7);
        hipLaunchKernelGGL(k_sym_u12<T>, dim3((n - jend + 63) / 64, (WB + 31) / 32, batch), dim3(256), 0, sh, A,
                           strideA, lda, 0, jend, jend, n, wk.nElem);
        pf->end(7);
        pf->begin(8);
        pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jend, jnext, 0, WB, sh, 1, wk.nElem));
        panel_phase<T>(A, strideA, n, lda, batch, wk, jend, jnext, 0, sh, pf);
        pf->begin(8);
        pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jnext, n, 0, WB, sh, 1, wk.nElem));
        j0 = WB;
    }
    if (sh != s) chk(hipEventRecord(wk.evPanel, sh));
    bool marked = false;
    for (; j0 < n; j0 += WB) {
        const int wb = n - j0 < WB ? n - j0 : WB;
        const int jend = j0 + wb;
        const int jnext = n - jend < WB ? n : jend + WB;
        if ((j0 == 0 || (j0 == WB && sh != sp) ? sh : sp) != s) chk(hipStreamWaitEvent(s, wk.evPanel, 0));
        if (wk.tailMark && !marked && n - j0 <= kTailCols) {
            chk(hipEventRecord(wk.tailMark, s));
            marked = true;
        }
        if (jend >= n) break;
        pf->begin(7);
        hipLaunchKernelGGL(k_sym_u12<T>, dim3((n - jend + 63) / 64, (wb + 31) / 32, batch), dim3(256), 0, s, A,
                           strideA, lda, j0, jend, jend, n, wk.nElem);
        pf->end(7);
        pf->begin(8);
        pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jend, jnext, j0, wb, s, 1, wk.nElem));
        if (sp != s) {
            chk(hipEventRecord(wk.evAhead, s));
            chk(hipStreamWaitEvent(sp, wk.evAhead, 0));
        }
        panel_phase<T>(A, strideA, n, lda, batch, wk, jend, jnext, 0, sp, pf);
        if (sp != s) chk(hipEventRecord(wk.evPanel, sp));
        pf->begin(8);
        pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jnext, n, j0, wb, s, 1, wk.nElem));
    }
    if (wk.tailMark && !marked) chk(hipEventRecord(wk.tailMark, s));
    return chk.e != hipSuccess ? chk.e : hipGetLastError();
}

// Blocked right-looking LU with look-ahead.  Outer blocks of WB columns are
// factored as NB-wide sub-panels (panel_phase) whose interchanges touch
// only the block's columns at first.  For every block: apply its
// interchanges to all other columns (incl. the rhs column n), U12 = L^-1
// A12 over the columns right of it, then the trailing update A22 -= L21
// U12 with K = WB, split in two: the next block's columns first, after
// which that block's panels run on the panel stream while the main stream
// updates the remaining columns.  The two never touch the same columns:
// the panels stay inside their block, and their interchanges reach the
// other columns only after the update that reads those rows has finished.
template <typename T>
hipError_t lu_factor(T* A, size_t strideA, int n, int lda, int batch, LuWork wk, hipStream_t s,
                     const Prof* pf) {
    Prof none;
    if (!pf) pf = &none;
    if (wk.noPivot) return lu_factor_sym<T>(A, strideA, n, lda, batch, wk, s, pf);
    HipFirst chk;
    hipStream_t sp = wk.panel ? wk.panel : s;
    // candidate and counter slots per batch element are sized for the first
    // (largest) panel; every counter is back at zero when a launch ends
    const int nodeStride = tree_geom(n).nNodes;
    if (sp != s) {
        // the panel stream starts after everything already queued on s
        chk(hipEventRecord(wk.evAhead, s));
        chk(hipStreamWaitEvent(sp, wk.evAhead, 0));
    }
    panel_phase<T>(A, strideA, n, lda, batch, wk, 0, n < WB ? n : WB, nodeStride, sp, pf);
    if (sp != s) {
        chk(hipEventRecord(wk.evPanel, sp));
    }
    for (int j0 = 0; j0 < n; j0 += WB) {
        const int wb = n - j0 < WB ? n - j0 : WB;
        const int jend = j0 + wb;
        const int jnext = n - jend < WB ? n : jend + WB;
        if (sp != s) chk(hipStreamWaitEvent(s, wk.evPanel, 0));
        // this block's interchanges on the columns outside it
        const int nrest = j0 + (n + 1 - jend);
        if (nrest > 0) {
            pf->begin(6);
            hipLaunchKernelGGL(k_laswp_rest<T>, dim3((nrest + 255) / 256, batch), dim3(256), 0, s, A, strideA, lda,
                               wk.swaps, (wb + NB - 1) / NB, 0, j0, jend, n + 1);
            pf->end(6);
        }
        // U12 of the block rows over columns jend..n (rhs included)
        for (int c0 = j0; c0 < jend; c0 += NB) {
            const int w = jend - c0 < NB ? jend - c0 : NB;
            pf->begin(7);
            pf->end(7, panel_cols<T>(A, strideA, lda, batch, c0, w, jend, n + 1, s));
            pf->begin(8);
            pf->end(8, gemm<T>(A, strideA, lda, batch, c0 + w, jend, jend, n + 1, c0, w, s));
        }
        if (jend >= n) break;
        // trailing update: the next block's columns, then (while its panels
        // are factored on the panel stream) the remaining columns
        bool p6 = false;
        if constexpr (std::is_same<T, float>::value) p6 = wk.splitPlanes != nullptr;
        pf->begin(8);
        if constexpr (std::is_same<T, float>::value) {
            if (p6) {
                const int ns = split_block(A, strideA, lda, batch, n, j0, jend, wk, s);
                pf->end(8, ns + gemm_p6(A, strideA, lda, batch, jend, n, jend, jnext, wk, n, jend, s));
            }
        }
        if (!p6) pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jend, jnext, j0, wb, s));
        if (sp != s) {
            chk(hipEventRecord(wk.evAhead, s));
            chk(hipStreamWaitEvent(sp, wk.evAhead, 0));
        }
        panel_phase<T>(A, strideA, n, lda, batch, wk, jend, jnext, nodeStride, sp, pf);
        if (sp != s) chk(hipEventRecord(wk.evPanel, sp));
        pf->begin(8);
        if constexpr (std::is_same<T, float>::value) {
            if (p6) pf->end(8, gemm_p6(A, strideA, lda, batch, jend, n, jnext, n + 1, wk, n, jend, s));
        }
        if (!p6) pf->end(8, gemm<T>(A, strideA, lda, batch, jend, n, jnext, n + 1, j0, wb, s));
    }
    return chk.e != hipSuccess ? chk.e : hipGetLastError();
}

template <typename T>
hipError_t lu_trsv(const T* A, size_t strideA, int n, int lda, T* X, size_t strideX, size_t xs,
                   int batch, bool lower, const int* skip, hipStream_t s, TrsvWork* tw) {
    const int nblk = (n + 63) / 64;
    // the kernel loads column pairs with 16-byte loads (lda even) and may
    // touch column n of a row (lda > n)
    if (!tw || (lda & 1) || lda <= n) return hipErrorInvalidValue;
    ++tw->epoch;
    hipLaunchKernelGGL(k_zero_int, dim3((batch + 255) / 256), dim3(256), 0, s, tw->tickets, batch);
    const int nb = tw->rows > 0 && tw->rows < nblk ? tw->rows : nblk;
    if constexpr (std::is_same<T, float>::value) {
        if (tw->rows > 0) {
            if (lower)
                hipLaunchKernelGGL((k_trsv_fused<T, true, true>), dim3(batch, nb), dim3(256), 0, s, A, strideA, lda,
                                   n, X, strideX, xs, tw->flags, tw->tickets, tw->epoch, skip, tw->err);
            else
                hipLaunchKernelGGL((k_trsv_fused<T, false, true>), dim3(batch, nb), dim3(256), 0, s, A, strideA, lda,
                                   n, X, strideX, xs, tw->flags, tw->tickets, tw->epoch, skip, tw->err);
            return hipGetLastError();
        }
    }
    if (lower)
        hipLaunchKernelGGL((k_trsv_fused<T, true, false>), dim3(batch, nb), dim3(256), 0, s, A, strideA, lda, n, X,
                           strideX, xs, tw->flags, tw->tickets, tw->epoch, skip, tw->err);
    else
        hipLaunchKernelGGL((k_trsv_fused<T, false, false>), dim3(batch, nb), dim3(256), 0, s, A, strideA, lda, n, X,
                           strideX, xs, tw->flags, tw->tickets, tw->epoch, skip, tw->err);
    return hipGetLastError();
}

template <typename T>
hipError_t lu_backsolve(T* A, size_t strideA, int n, int lda, int batch, hipStream_t s, TrsvWork* tw) {
    return lu_trsv<T>(A, strideA, n, lda, A + 2 * (size_t)n, strideA, (size_t)lda, batch, false, nullptr, s, tw);
}



// Largest dynamic LDS one block of k_make_perm may take on the current
// device (the opt-in limit, 160 KiB on gfx950; 64 KiB if it cannot be read).
size_t lu_perm_lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || v <= 0)
        return 65536;
    return (size_t)v;
}

hipError_t lu_ident_perm(int n, int batch, int* perm, hipStream_t s) {
    hipLaunchKernelGGL(k_ident_perm, dim3((n + 255) / 256, batch), dim3(256), 0, s, perm, n);
    return hipGetLastError();
}

hipError_t lu_make_perm(const int* ipiv, int n, int batch, int* perm, hipStream_t s) {
    const size_t lds = sizeof(int) * (size_t)n;
    if (lds <= lu_perm_lds_limit()) {
        if (lds > 65536) {
            // above the default dynamic-LDS limit the kernel opts in
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_make_perm<true>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_make_perm<true>, dim3(batch), dim3(256), lds, s, ipiv, n, perm);
    } else {
        hipLaunchKernelGGL(k_make_perm<false>, dim3(batch), dim3(256), 0, s, ipiv, n, perm);
    }
    return hipGetLastError();
}

template hipError_t lu_factor<double>(double*, size_t, int, int, int, LuWork, hipStream_t, const Prof*);
template hipError_t lu_factor<float>(float*, size_t, int, int, int, LuWork, hipStream_t, const Prof*);
template hipError_t lu_backsolve<double>(double*, size_t, int, int, int, hipStream_t, TrsvWork*);
template hipError_t lu_backsolve<float>(float*, size_t, int, int, int, hipStream_t, TrsvWork*);
template hipError_t lu_trsv<double>(const double*, size_t, int, int, double*, size_t, size_t, int, bool,
                                    const int*, hipStream_t, TrsvWork*);
template hipError_t lu_trsv<float>(const float*, size_t, int, int, float*, size_t, size_t, int, bool,
                                   const int*, hipStream_t, TrsvWork*);

}  // namespace efie
