// efie_rwg.hip -- RWG tables and the port pick on the device.
//
// build_rwg_interior.m:1-113 from the mesh arrays: every triangle edge
// (raw index 3t + le, le over (v1,v2), (v2,v3), (v3,v1)) is keyed by its
// sorted vertex pair (a < b).  The keys are bucketed by a (a counting sort:
// counts, one scan, scatter), each bucket is ordered by (b, raw) by one
// thread, and a run of exactly two equal keys is an interior edge.  Edges
// are numbered in the lexicographic order of (a, b), which is the order of
// unique(rawKey, 'rows') (:25); the first incidence in raw order is the
// plus side (:47-63).  The geometry (length, centre, opposite vertices,
// areas, layer sign) repeats the host builder's operations in the same
// order with contraction off, so every table is bit-identical to it.
//
// pick_port_edge_by_point_and_layer.m:1-11: the edge of the given layer
// nearest to a point, the lowest index among equal distances.
#include "efie_internal.h"

namespace efie {

namespace {

// raw edge le of triangle t: its two vertices in triangle order
__device__ __forceinline__ void raw_edge(const int* F, int Nt, int raw, int& d0, int& d1) {
    const int t = raw / 3, le = raw % 3;
    const int v[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
    d0 = v[le];
    d1 = v[le == 2 ? 0 : le + 1];
}

__global__ __launch_bounds__(256) void k_rwg_count(const int* __restrict__ F, int Nt, int* __restrict__ cnt) {
    const int raw = blockIdx.x * blockDim.x + threadIdx.x;
    if (raw >= 3 * Nt) return;
    int d0, d1;
    raw_edge(F, Nt, raw, d0, d1);
    atomicAdd(cnt + (min(d0, d1) - 1), 1);
}

// Exclusive scan of n ints in one block of 1024 threads (each thread a
// contiguous chunk); out[n] = total.  `out` may alias `in`.
__global__ __launch_bounds__(1024) void k_scan1(const int* in, int n, int* out) {
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int i0 = min(n, tid * chunk), i1 = min(n, i0 + chunk);
    int s = 0;
    for (int i = i0; i < i1; ++i) s += in[i];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int i = i0; i < i1; ++i) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
    if (tid == 1023) out[n] = part[1023];
}

// raw edges into their buckets (fill[] counts the slots taken)
__global__ __launch_bounds__(256) void k_rwg_scatter(const int* __restrict__ F, int Nt, const int* __restrict__ start,
                                                     int* __restrict__ fill, int2* __restrict__ slot) {
    const int raw = blockIdx.x * blockDim.x + threadIdx.x;
    if (raw >= 3 * Nt) return;
    int d0, d1;
    raw_edge(F, Nt, raw, d0, d1);
    const int a = min(d0, d1) - 1;
    const int p = start[a] + atomicAdd(fill + a, 1);
    slot[p] = make_int2(max(d0, d1), raw);
}

// one thread per bucket: order by (b, raw), count the runs of length 2
__global__ __launch_bounds__(256) void k_rwg_runs(const int* __restrict__ start, int Nv, int2* __restrict__ slot,
                                                  int* __restrict__ nInt) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= Nv) return;
    const int s0 = start[a], s1 = start[a + 1];
    for (int i = s0 + 1; i < s1; ++i) {
        const int2 v = slot[i];
        int j = i;
        while (j > s0 && (slot[j - 1].x > v.x || (slot[j - 1].x == v.x && slot[j - 1].y > v.y))) {
            slot[j] = slot[j - 1];
            --j;
        }
        slot[j] = v;
    }
    int n = 0;
    for (int i = s0; i < s1;) {
        int j = i;
        while (j < s1 && slot[j].x == slot[i].x) ++j;
        n += (j - i == 2) ? 1 : 0;
        i = j;
    }
    nInt[a] = n;
}

struct DV {
    double x, y, z;
};

__device__ __forceinline__ DV vtx(const double* V, int Nv, int v) {
    return {V[v - 1], V[v - 1 + Nv], V[v - 1 + 2 * Nv]};
}

__device__ __forceinline__ double dnorm(double x, double y, double z) {
#pragma clang fp contract(off)
    return sqrt(x * x + y * y + z * z);
}

// one thread per bucket: the interior edges of bucket a, numbered from
// eStart[a]; the tables are column-major with leading dimension Ne
__global__ __launch_bounds__(256) void k_rwg_tables(const double* __restrict__ V, int Nv, const int* __restrict__ F,
                                                    int Nt, const int* __restrict__ start,
                                                    const int2* __restrict__ slot, const int* __restrict__ eStart,
                                                    int Ne, double zmid, int* __restrict__ edge,
                                                    int* __restrict__ tri, int* __restrict__ sgn,
                                                    double* __restrict__ rpm, double* __restrict__ Apm,
                                                    double* __restrict__ len, double* __restrict__ center,
                                                    int* __restrict__ zSign, int* __restrict__ err) {
#pragma clang fp contract(off)
    const int a0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (a0 >= Nv) return;
    const int a = a0 + 1, s0 = start[a0], s1 = start[a0 + 1];
    int e = eStart[a0];
    for (int i = s0; i < s1;) {
        int j = i;
        while (j < s1 && slot[j].x == slot[i].x) ++j;
        if (j - i == 2) {
            const int b = slot[i].x;
            edge[e] = a;
            edge[e + Ne] = b;
            const DV ra = vtx(V, Nv, a), rb = vtx(V, Nv, b);
            len[e] = dnorm(rb.x - ra.x, rb.y - ra.y, rb.z - ra.z);
            const DV c = {0.5 * (ra.x + rb.x), 0.5 * (ra.y + rb.y), 0.5 * (ra.z + rb.z)};
            center[e] = c.x;
            center[e + Ne] = c.y;
            center[e + 2 * Ne] = c.z;
            zSign[e] = c.z > zmid ? 1 : -1;
            for (int s = 0; s < 2; ++s) {
                const int raw = slot[i + s].y, t = raw / 3;
                int d0, d1;
                raw_edge(F, Nt, raw, d0, d1);
                tri[e + s * Ne] = t + 1;
                sgn[e + s * Ne] = (d0 == a && d1 == b) ? 1 : -1;
                const int ids[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
                int opp = 0, cnt = 0;
                for (int k = 0; k < 3; ++k)
                    if (ids[k] != a && ids[k] != b) {
                        opp = ids[k];
                        ++cnt;
                    }
                if (cnt != 1) {
                    atomicMin(err, e + 1);
                    opp = ids[0];
                }
                const DV o = vtx(V, Nv, opp);
                double* rr = rpm + (size_t)s * 3 * Ne;
                rr[e] = o.x;
                rr[e + Ne] = o.y;
                rr[e + 2 * Ne] = o.z;
                const DV p0 = vtx(V, Nv, ids[0]), p1 = vtx(V, Nv, ids[1]), p2 = vtx(V, Nv, ids[2]);
                const double ux = p1.x - p0.x, uy = p1.y - p0.y, uz = p1.z - p0.z;
                const double wx = p2.x - p0.x, wy = p2.y - p0.y, wz = p2.z - p0.z;
                Apm[e + (size_t)s * Ne] =
                    0.5 * dnorm(uy * wz - uz * wy, uz * wx - ux * wz, ux * wy - uy * wx);
            }
            ++e;
        }
        i = j;
    }
}

// port pick, pass 1: per block the (distance, index) minimum over the
// edges of the layer; pass 2: one block over the partials.  Distances are
// non-negative, so (bits of d, index) compares as the pair.
struct DI {
    double d;
    int i;
};

__device__ __forceinline__ DI di_min(DI x, DI y) { return (y.d < x.d || (y.d == x.d && y.i < x.i)) ? y : x; }

__device__ DI block_min(DI v) {
    __shared__ double sd[256];
    __shared__ int si[256];
    sd[threadIdx.x] = v.d;
    si[threadIdx.x] = v.i;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const DI w = di_min({sd[threadIdx.x], si[threadIdx.x]}, {sd[threadIdx.x + o], si[threadIdx.x + o]});
            sd[threadIdx.x] = w.d;
            si[threadIdx.x] = w.i;
        }
        __syncthreads();
    }
    return {sd[0], si[0]};
}

constexpr int kNoEdge = 0x7fffffff;

__global__ __launch_bounds__(256) void k_port_pass1(const double* __restrict__ center, const int* __restrict__ zSign,
                                                    int Ne, double px, double py, double pz, int layer,
                                                    double* __restrict__ pd, int* __restrict__ pi) {
#pragma clang fp contract(off)
    DI v = {__builtin_huge_val(), kNoEdge};
    for (int e = blockIdx.x * 256 + threadIdx.x; e < Ne; e += gridDim.x * 256) {
        if (zSign[e] != layer) continue;
        const double d = dnorm(center[e] - px, center[e + Ne] - py, center[e + 2 * Ne] - pz);
        v = di_min(v, {d, e});
    }
    v = block_min(v);
    if (threadIdx.x == 0) {
        pd[blockIdx.x] = v.d;
        pi[blockIdx.x] = v.i;
    }
}

__global__ __launch_bounds__(256) void k_port_pass2(const double* __restrict__ pd, const int* __restrict__ pi, int n,
                                                    int* __restrict__ out) {
    DI v = {__builtin_huge_val(), kNoEdge};
    for (int k = threadIdx.x; k < n; k += 256) v = di_min(v, {pd[k], pi[k]});
    v = block_min(v);
    if (threadIdx.x == 0) out[0] = v.i == kNoEdge ? 0 : v.i + 1;
}

}  // namespace

hipError_t launch_rwg_count(const int* F, int Nt, int Nv, int* cnt, int2* slot, int* fill, int* nInt, int* start,
                            int* eStart, hipStream_t s) {
    if (Nt < 1 || Nv < 1) return hipErrorInvalidValue;
    const unsigned gr = (unsigned)((3LL * Nt + 255) / 256), gv = (unsigned)((Nv + 255) / 256);
    hipError_t e = launch_fill_int(cnt, 0, Nv, s);
    if (e == hipSuccess) e = launch_fill_int(fill, 0, Nv, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rwg_count, dim3(gr), dim3(256), 0, s, F, Nt, cnt);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, cnt, Nv, start);
    hipLaunchKernelGGL(k_rwg_scatter, dim3(gr), dim3(256), 0, s, F, Nt, start, fill, slot);
    hipLaunchKernelGGL(k_rwg_runs, dim3(gv), dim3(256), 0, s, start, Nv, slot, nInt);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, nInt, Nv, eStart);
    return hipGetLastError();
}

hipError_t launch_rwg_tables(const double* V, int Nv, const int* F, int Nt, const int* start, const int2* slot,
                             const int* eStart, int Ne, double zmid, int* edge, int* tri, int* sgn, double* rpm,
                             double* Apm, double* len, double* center, int* zSign, int* err, hipStream_t s) {
    hipError_t e = launch_fill_int(err, 0x7fffffff, 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rwg_tables, dim3((unsigned)((Nv + 255) / 256)), dim3(256), 0, s, V, Nv, F, Nt, start, slot,
                       eStart, Ne, zmid, edge, tri, sgn, rpm, Apm, len, center, zSign, err);
    return hipGetLastError();
}

int port_pick_blocks(int Ne) { return std::max(1, std::min(1024, (Ne + 255) / 256)); }

hipError_t launch_port_pick(const double* center, const int* zSign, int Ne, const double p[3], int layer, double* pd,
                            int* pi, int* out, hipStream_t s) {
    const int nb = port_pick_blocks(Ne);
    hipLaunchKernelGGL(k_port_pass1, dim3(nb), dim3(256), 0, s, center, zSign, Ne, p[0], p[1], p[2], layer, pd, pi);
    hipLaunchKernelGGL(k_port_pass2, dim3(1), dim3(256), 0, s, pd, pi, nb, out);
    return hipGetLastError();
}

}  // namespace efie
