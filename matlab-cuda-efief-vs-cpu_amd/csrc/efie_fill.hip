// efie_fill.hip -- impedance-matrix fill, excitation and epilogue kernels
// for gfx950.
//
// The reference fills Z one RWG pair at a time and recomputes every
// triangle-pair integral for each of the up to nine RWG pairs that share it
// (assemble_ZV_EFIE_RWG_DIFF_CPU.m:9-36, patch_sweep_fastquad_cuda.cu:347-412).
// Here the fill is split in two:
//   1. triangle-pair moments ("records"): for every unordered pair t1 <= t2
//      the four Green-weighted moments S0, S1, S1', S2 of the quadrature
//      rule the reference would use for that pair (regular, subdivided near
//      or subdivided self), computed once per frequency;
//   2. combine: Z(m,n) = sum over the four side pairs of alpha*Ivec +
//      beta*Isca, where with f = c (r - v) on each side
//        Ivec = cm cn [S2 - S1.un - S1'.um + S0 (um.un)],  Isca = 4 cm cn S0,
//      which is the reference integrand sum regrouped (rwg_eval_on_triangle.m
//      gives f and div f = 2c on both sides).
// Coordinates inside a pair are relative to each triangle's centroid so the
// regrouping does not cancel digits.
#include "efie_internal.h"
#include "efie_math.h"

#include <math.h>

namespace efie {

static constexpr double kPi = 3.141592653589793238462643383279502884;

// G = exp(-j k R) / (4 pi R) for complex k = kr + j ki
// (rwg_pair_integrals_regular.m:22; exp(ki R)(cos(kr R) - j sin(kr R))) is
// evaluated per triangle pair by PairGreen / green_eval below.

// 1/sqrt(x) for x > 0: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// Green's function of one triangle pair, factored about the centroid
// distance R0: G(R) = e0 * e^{-jk(R - R0)} / R with e0 = e^{-jk R0}/(4 pi).
// The inner sums accumulate only e^{-jk(R - R0)}/R; the caller multiplies
// the finished moments by e0 once per pair.  |R - R0| <= dmax (the sum of
// the two triangles' point radii) bounds the residual phase, which picks
// the cheapest series that stays below 1e-17 relative (mode 0: |k dmax| <=
// 0.2, mode 1: <= 0.75) or the full range-reduced sine/cosine (mode 2).
struct PairGreen {
    double R0, e0r, e0i, kr, ki;
    int mode;
};

__device__ __forceinline__ PairGreen pair_green(double kr, double ki, double R0, double dmax) {
    PairGreen g;
    g.R0 = R0;
    g.kr = kr;
    g.ki = ki;
    const double x = fabs(kr) * dmax;
    const bool lossOk = fabs(ki) * dmax <= 1e-3;
    g.mode = (x <= 0.2 && lossOk) ? 0 : (x <= 0.75 && lossOk) ? 1 : 2;
    double s, c;
    sincos_cw(kr * R0, &s, &c);
    const double amp = ((ki == 0.0) ? 1.0 : exp(ki * R0)) * (1.0 / (4.0 * kPi));
    g.e0r = amp * c;
    g.e0i = -amp * s;
    return g;
}

// e^{-jk(R - R0)} / R = af (c - j s) for R^2 = R2 (R2 > 0) with the phase
// series MODE (any mode at or above the pair's own is as accurate) and
// LOSSY = (ki != 0).
template <int MODE, bool LOSSY>
__device__ __forceinline__ void green_cs(const PairGreen& g, double R2, double& c, double& s, double& af) {
    const double invR = rsqrt64(R2);
    const double d = R2 * invR - g.R0;
    af = invR;
    if constexpr (MODE <= 1) {
        if constexpr (MODE == 0) sincos_small_short(g.kr * d, &s, &c);
        else sincos_small(g.kr * d, &s, &c);
        if constexpr (LOSSY) {
            // e^{ki d} = 1 + ki d + (ki d)^2/2 + ... for |ki d| <= 1e-3
            const double z = g.ki * d;
            af *= 1.0 + z * (1.0 + z * (0.5 + z * (1.0 / 6.0 + z * (1.0 / 24.0))));
        }
    } else {
        sincos_cw(g.kr * d, &s, &c);
        if constexpr (LOSSY) af = exp(g.ki * d) * invR;
    }
}

template <int MODE, bool LOSSY>
__device__ __forceinline__ void green_t(const PairGreen& g, double R2, double& gr, double& gi) {
    double c, s, af;
    green_cs<MODE, LOSSY>(g, R2, c, s, af);
    gr = c * af;
    gi = -s * af;
}

__device__ __forceinline__ void green_rel(const PairGreen& g, double R2, double& gr, double& gi) {
    if (g.ki != 0.0) {
        if (g.mode == 0) green_t<0, true>(g, R2, gr, gi);
        else if (g.mode == 1) green_t<1, true>(g, R2, gr, gi);
        else green_t<2, true>(g, R2, gr, gi);
    } else {
        if (g.mode == 0) green_t<0, false>(g, R2, gr, gi);
        else if (g.mode == 1) green_t<1, false>(g, R2, gr, gi);
        else green_t<2, false>(g, R2, gr, gi);
    }
}

// Near test of triangles_are_near.m:1-12 in the reference's operation
// order, with contraction disabled so every rounding matches the C oracle.
__device__ bool near_test(const TriGeo a, const TriGeo b, double eps) {
#pragma clang fp contract(off)
    double dx = a.cx - b.cx, dy = a.cy - b.cy, dz = a.cz - b.cz;
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    double A1 = 0.5 * a.area2, A2 = 0.5 * b.area2;
    double s3 = sqrt(3.0);
    double l1 = sqrt(4 * A1 / s3), l2 = sqrt(4 * A2 / s3);
    double h = 0.5 * (l1 + l2);
    return d < eps * h;
}

// ---------------------------------------------------- quadrature points --

// Quadrature points of every triangle, relative to its centroid, built on
// the device from the vertices (efie_precompute.m's tri_quad / subdivided
// rules): the regular rule on the triangle, and the subdivided rule on each
// of the 4^depth midpoint children for near and self pairs.  The operations
// are the host generator's (map_ref_to_vertices.m, subdivide_triangles.m)
// in the same order with contraction off, so the points are bit-identical
// to it.  Child p of the subdivision is found digit by digit from the top
// level (base 4, most significant first): the children of (a, b, c) are
// (a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca).
struct QV {
    double x, y, z;
};

__device__ __forceinline__ QV qv_mid(QV a, QV b) {
#pragma clang fp contract(off)
    return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
}

__device__ __forceinline__ QV qv_map(QV A, QV B, QV C, double u, double v) {
#pragma clang fp contract(off)
    return {A.x + u * (B.x - A.x) + v * (C.x - A.x), A.y + u * (B.y - A.y) + v * (C.y - A.y),
            A.z + u * (B.z - A.z) + v * (C.z - A.z)};
}

__device__ __forceinline__ double qv_area2(QV A, QV B, QV C) {
#pragma clang fp contract(off)
    const double bx = B.x - A.x, by = B.y - A.y, bz = B.z - A.z;
    const double cx = C.x - A.x, cy = C.y - A.y, cz = C.z - A.z;
    const double x = by * cz - bz * cy, y = bz * cx - bx * cz, z = bx * cy - by * cx;
    return sqrt(x * x + y * y + z * z);
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// one thread per (triangle, unit): unit 0 the regular rule and the
// vertices, units 1..NsN the near children, then the NsS self children.
// rad[t] (max distance of any point or vertex from the centroid) and
// mx = {near radius, self radius, max self softening, max rad} are maxima
// of non-negative doubles, kept as their bit patterns.
__global__ __launch_bounds__(256) void k_quad_points(const double* __restrict__ tv, const TriGeo* __restrict__ tg,
                                                     int Nt, QuadRules R, int depthN, int depthS, double selfAlpha,
                                                     Pt* __restrict__ regPts, Pt* __restrict__ nearPts,
                                                     Pt* __restrict__ selfPts, double* __restrict__ aSelf,
                                                     unsigned long long* __restrict__ rad,
                                                     unsigned long long* __restrict__ mx) {
#pragma clang fp contract(off)
    const int NsN = 1 << (2 * depthN), NsS = 1 << (2 * depthS);
    const int units = 1 + NsN + NsS;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = id < (long long)Nt * units;
    const int t = live ? (int)(id / units) : 0, u = live ? (int)(id % units) : 0;
    double r = 0.0, rN = 0.0, rS = 0.0, aS = 0.0;
    if (live) {
        const QV A{tv[9 * (size_t)t], tv[9 * (size_t)t + 1], tv[9 * (size_t)t + 2]};
        const QV B{tv[9 * (size_t)t + 3], tv[9 * (size_t)t + 4], tv[9 * (size_t)t + 5]};
        const QV C{tv[9 * (size_t)t + 6], tv[9 * (size_t)t + 7], tv[9 * (size_t)t + 8]};
        const TriGeo g = tg[t];
        auto dist = [&](QV q) {
            const double x = q.x - g.cx, y = q.y - g.cy, z = q.z - g.cz;
            return sqrt(x * x + y * y + z * z);
        };
        if (u == 0) {
            r = fmax(dist(A), fmax(dist(B), dist(C)));
            for (int i = 0; i < R.nReg; ++i) {
                const QV q = qv_map(A, B, C, R.ru[i], R.rv[i]);
                const Pt p = {q.x - g.cx, q.y - g.cy, q.z - g.cz, R.rw[i] * g.area2};
                regPts[(size_t)t * R.nReg + i] = p;
                r = fmax(r, sqrt(p.x * p.x + p.y * p.y + p.z * p.z));
            }
        } else {
            const bool self = u > NsN;
            const int depth = self ? depthS : depthN, Ns = self ? NsS : NsN;
            const int child = self ? u - 1 - NsN : u - 1;
            QV a = A, b = B, c = C;
            for (int d = depth - 1; d >= 0; --d) {
                const int k = (child >> (2 * d)) & 3;
                const QV ab = qv_mid(a, b), bc = qv_mid(b, c), ca = qv_mid(c, a);
                if (k == 0) { b = ab; c = ca; }
                else if (k == 1) { a = ab; c = bc; }
                else if (k == 2) { a = ca; b = bc; }
                else { a = ab; b = bc; c = ca; }
            }
            const double a2p = qv_area2(a, b, c);
            if (self) {
                aS = selfAlpha * sqrt((0.5 * a2p) / kPi);
                aSelf[(size_t)t * Ns + child] = aS;
            }
            Pt* out = (self ? selfPts : nearPts) + ((size_t)t * Ns + child) * R.nSub;
            for (int i = 0; i < R.nSub; ++i) {
                const QV q = qv_map(a, b, c, R.su[i], R.sv[i]);
                const Pt p = {q.x - g.cx, q.y - g.cy, q.z - g.cz, R.sw[i] * a2p};
                out[i] = p;
                const double d = sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
                r = fmax(r, d);
                if (self) rS = fmax(rS, d);
                else rN = fmax(rN, d);
            }
        }
        atomicMax(rad + t, (unsigned long long)__double_as_longlong(r));
    }
    rN = wave_max(rN);
    rS = wave_max(rS);
    aS = wave_max(aS);
    r = wave_max(r);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(mx + 0, (unsigned long long)__double_as_longlong(rN));
        atomicMax(mx + 1, (unsigned long long)__double_as_longlong(rS));
        atomicMax(mx + 2, (unsigned long long)__double_as_longlong(aS));
        atomicMax(mx + 3, (unsigned long long)__double_as_longlong(r));
    }
}

// ------------------------------------------------------------ near map --

// Near pairs by cell binning: triangle centroids are binned into cubic
// cells of side h >= near_sing_eps * max(l), which bounds the distance of
// every near pair, so a triangle's near partners lie in its own cell or
// the 26 around it.  `cellTri` lists the triangles cell by cell (offsets
// `cellStart`, ncell + 1 entries), `cellOf` is each triangle's cell.
struct CellGrid {
    double x0, y0, z0, h;
    int nx, ny, nz;
};

__device__ __forceinline__ void cell_coords(const CellGrid& g, int c, int& ix, int& iy, int& iz) {
    ix = c % g.nx;
    iy = (c / g.nx) % g.ny;
    iz = c / (g.nx * g.ny);
}

// rowCount[t1] = number of near t2 > t1 (one thread per row).
__global__ __launch_bounds__(256) void k_near_count(const TriGeo* __restrict__ tg, int Nt, double eps, CellGrid g,
                                                    const int* __restrict__ cellOf, const int* __restrict__ cellStart,
                                                    const int* __restrict__ cellTri, int* __restrict__ rowCount) {
    const int t1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t1 >= Nt) return;
    const TriGeo a = tg[t1];
    int ix, iy, iz;
    cell_coords(g, cellOf[t1], ix, iy, iz);
    int cnt = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = iz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = iy + dy;
            if (y < 0 || y >= g.ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = ix + dx;
                if (x < 0 || x >= g.nx) continue;
                const int c = (z * g.ny + y) * g.nx + x;
                for (int k = cellStart[c]; k < cellStart[c + 1]; ++k) {
                    const int t2 = cellTri[k];
                    cnt += (t2 > t1 && near_test(a, tg[t2], eps)) ? 1 : 0;
                }
            }
        }
    }
    rowCount[t1] = cnt;
}

// The near t2 > t1 of row t1 to list[rowStart[t1]..], sorted by t2 (the
// order of the dense scan over t2 that the reference's near matrix
// implies).  One wave per row: the lanes test 64 candidates of a cell at a
// time and append the near ones to the row's LDS buffer (ballot prefix
// counts); the row is then written in order, each entry at its rank.  A
// row longer than the buffer is built by lane 0 alone, by insertion into
// the list.
constexpr int kNearRowCap = 512;

__global__ __launch_bounds__(256) void k_near_rows(const TriGeo* __restrict__ tg, int Nt, double eps, CellGrid g,
                                                   const int* __restrict__ cellOf, const int* __restrict__ cellStart,
                                                   const int* __restrict__ cellTri, const int* __restrict__ rowStart,
                                                   int2* __restrict__ list) {
    __shared__ int buf[4][kNearRowCap];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t1 = blockIdx.x * 4 + wv;
    if (t1 >= Nt) return;
    int* rb = buf[wv];
    const TriGeo a = tg[t1];
    int ix, iy, iz;
    cell_coords(g, cellOf[t1], ix, iy, iz);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int cnt = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = iz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = iy + dy;
            if (y < 0 || y >= g.ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = ix + dx;
                if (x < 0 || x >= g.nx) continue;
                const int c = (z * g.ny + y) * g.nx + x;
                const int k1 = cellStart[c + 1];
                for (int k0 = cellStart[c]; k0 < k1; k0 += 64) {
                    const int k = k0 + lane;
                    const int t2 = k < k1 ? cellTri[k] : -1;
                    const bool nr = t2 > t1 && near_test(a, tg[t2], eps);
                    const unsigned long long m = __ballot(nr);
                    const int at = cnt + __popcll(m & below);
                    if (nr && at < kNearRowCap) rb[at] = t2;
                    cnt += __popcll(m);
                }
            }
        }
    }
    int2* row = list + rowStart[t1];
    if (cnt <= kNearRowCap) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int e = lane; e < cnt; e += 64) {
            const int v = rb[e];
            int rank = 0;
            for (int j = 0; j < cnt; ++j) rank += rb[j] < v ? 1 : 0;
            row[rank] = make_int2(t1, v);
        }
        return;
    }
    if (lane != 0) return;
    int n = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = iz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = iy + dy;
            if (y < 0 || y >= g.ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = ix + dx;
                if (x < 0 || x >= g.nx) continue;
                const int c = (z * g.ny + y) * g.nx + x;
                for (int k = cellStart[c]; k < cellStart[c + 1]; ++k) {
                    const int t2 = cellTri[k];
                    if (t2 <= t1 || !near_test(a, tg[t2], eps)) continue;
                    int p = n;
                    while (p > 0 && row[p - 1].y > t2) {
                        row[p] = row[p - 1];
                        --p;
                    }
                    row[p] = make_int2(t1, t2);
                    ++n;
                }
            }
        }
    }
}

// Ordered compaction of the near pairs (t1 < t2) of one row.
__global__ __launch_bounds__(256) void k_near_list(const uint8_t* __restrict__ nearMat,
                                                   const int* __restrict__ rowStart, int Nt,
                                                   int2* __restrict__ list) {
    __shared__ int scan[256];
    int t1 = blockIdx.x;
    int out = rowStart[t1];
    for (int base = t1 + 1; base < Nt; base += 256) {
        int t2 = base + (int)threadIdx.x;
        int f = (t2 < Nt) ? nearMat[(size_t)t1 * Nt + t2] : 0;
        scan[threadIdx.x] = f;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            int v = ((int)threadIdx.x >= o) ? scan[threadIdx.x - o] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (f) list[out + scan[threadIdx.x] - 1] = make_int2(t1, t2);
        int tot = scan[255];
        __syncthreads();
        out += tot;
    }
}

// ---------------------------------------------------- record helpers ----

struct Acc {
    double s0r, s0i, s1xr, s1xi, s1yr, s1yi, s1zr, s1zi;
    double s1pxr, s1pxi, s1pyr, s1pyi, s1pzr, s1pzi, s2r, s2i;
};

__device__ __forceinline__ void acc_zero(Acc& a) {
    a.s0r = a.s0i = a.s1xr = a.s1xi = a.s1yr = a.s1yi = a.s1zr = a.s1zi = 0.0;
    a.s1pxr = a.s1pxi = a.s1pyr = a.s1pyi = a.s1pzr = a.s1pzi = a.s2r = a.s2i = 0.0;
}

// Fold the inner sums of one test point i into the pair moments:
// g = sum_j W_j G_ij, h = sum_j W_j G_ij rho'_j.
__device__ __forceinline__ void acc_fold(Acc& a, double w, double x, double y, double z,
                                         double gr, double gi, double hxr, double hxi,
                                         double hyr, double hyi, double hzr, double hzi) {
    double wgr = w * gr, wgi = w * gi;
    a.s0r += wgr;
    a.s0i += wgi;
    a.s1xr += wgr * x; a.s1xi += wgi * x;
    a.s1yr += wgr * y; a.s1yi += wgi * y;
    a.s1zr += wgr * z; a.s1zi += wgi * z;
    a.s1pxr += w * hxr; a.s1pxi += w * hxi;
    a.s1pyr += w * hyr; a.s1pyi += w * hyi;
    a.s1pzr += w * hzr; a.s1pzi += w * hzi;
    a.s2r += w * (x * hxr + y * hyr + z * hzr);
    a.s2i += w * (x * hxi + y * hyi + z * hzi);
}

// Multiply the eight complex moments by e0 (the pair's reference factor).
__device__ __forceinline__ void acc_scale(Acc& a, double er, double ei) {
    double* v = &a.s0r;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double x = v[2 * k], y = v[2 * k + 1];
        v[2 * k] = x * er - y * ei;
        v[2 * k + 1] = x * ei + y * er;
    }
}

__device__ __forceinline__ void acc_store(Rec* r, const Acc& a) {
    double2* p = reinterpret_cast<double2*>(r->v);
    p[0] = make_double2(a.s0r, a.s0i);
    p[1] = make_double2(a.s1xr, a.s1xi);
    p[2] = make_double2(a.s1yr, a.s1yi);
    p[3] = make_double2(a.s1zr, a.s1zi);
    p[4] = make_double2(a.s1pxr, a.s1pxi);
    p[5] = make_double2(a.s1pyr, a.s1pyi);
    p[6] = make_double2(a.s1pzr, a.s1pzi);
    p[7] = make_double2(a.s2r, a.s2i);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------ regular pairs ---
// Tile of 64 test triangles (one per lane) by 64 source triangles (16 per
// wave); one block per upper-triangular tile pair.  The test points stay in
// registers, the source points are wave-uniform loads.

constexpr int kRegTile = 64;
// blocks per 64 x 64 tile pair (each takes kRegTile / kRegSplit source
// triangles): a finer grid shortens the tail of the last dispatch round
constexpr int kRegSplit = 2;

// The NQ x NQ product-rule sums of one regular pair.  MODE and LOSSY are
// wave-uniform, so the loop body has no branches and the NQ source points
// of one test point are evaluated side by side.
template <int NQ, int MODE, bool LOSSY>
__device__ __forceinline__ void regular_sums(const Pt* __restrict__ p1, const Pt* __restrict__ p2, double ox,
                                             double oy, double oz, const PairGreen& pg, Acc& a) {
    acc_zero(a);
#pragma unroll 1
    for (int i = 0; i < NQ; ++i) {
        const Pt pi = p1[i];
        const double xi = pi.x + ox, yi = pi.y + oy, zi = pi.z + oz;
        double gr = 0, gi = 0, hxr = 0, hxi = 0, hyr = 0, hyi = 0, hzr = 0, hzi = 0;
        auto point = [&](int j) {
            const Pt q2 = p2[j];
            const double dx = xi - q2.x, dy = yi - q2.y, dz = zi - q2.z;
            // R < 1e-15 -> 1e-15 (R2 >= 0, so a max is the same clamp)
            const double R2 = fmax(dx * dx + dy * dy + dz * dz, 1e-30);
            double c, s, af;
            green_cs<MODE, LOSSY>(pg, R2, c, s, af);
            // W G = (W af) (c - j s); the minus sign goes into the sums
            const double wa = q2.w * af;
            const double wr = c * wa, ws = s * wa;
            gr += wr; gi -= ws;
            hxr += wr * q2.x; hxi -= ws * q2.x;
            hyr += wr * q2.y; hyi -= ws * q2.y;
            hzr += wr * q2.z; hzi -= ws * q2.z;
        };
        if constexpr (MODE == 2) {
            // the long range reduction: rolled, so it does not set the
            // register budget of the short-series loops
#pragma unroll 1
            for (int j = 0; j < NQ; ++j) point(j);
        } else {
#pragma unroll
            for (int j = 0; j < NQ; ++j) point(j);
        }
        acc_fold(a, pi.w, pi.x, pi.y, pi.z, gr, gi, hxr, hxi, hyr, hyi, hzr, hzi);
    }
}

template <int NQ, bool LOSSY>
__device__ __forceinline__ void regular_sums_mode(int mode, const Pt* __restrict__ p1, const Pt* __restrict__ p2,
                                                  double ox, double oy, double oz, const PairGreen& pg, Acc& a) {
    if (mode == 0) regular_sums<NQ, 0, LOSSY>(p1, p2, ox, oy, oz, pg, a);
    else if (mode == 1) regular_sums<NQ, 1, LOSSY>(p1, p2, ox, oy, oz, pg, a);
    else regular_sums<NQ, 2, LOSSY>(p1, p2, ox, oy, oz, pg, a);
}

// One block per (64 x 64 tile pair, part): lane = test triangle t1 of tile
// ti, each wave walks its share of the part's source triangles t2 of tile
// tj, one pair per lane at a time.  The wave takes the largest phase-series
// mode its active lanes need.
template <int NQ>
__global__ __launch_bounds__(256, 3) void k_rec_regular(const Pt* __restrict__ pts,
                                                     const TriGeo* __restrict__ tg,
                                                     const double* __restrict__ rad, int Nt,
                                                     const uint8_t* __restrict__ nearMat, double eps,
                                                     const Freq* __restrict__ fr,
                                                     Rec* __restrict__ recs) {
    // linear block id -> tile pair (ti <= tj)
    long long b = blockIdx.x;
    int tj = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((long long)(tj + 1) * (tj + 2) / 2 <= b) ++tj;
    while ((long long)tj * (tj + 1) / 2 > b) --tj;
    int ti = (int)(b - (long long)tj * (tj + 1) / 2);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t1 = ti * kRegTile + lane;
    const double kr = fr->kr, ki = fr->ki;
    const bool lossy = ki != 0.0;

    // the block's 64 x NQ test points are staged in LDS (a register array
    // indexed by the rolled loop below would live in scratch)
    __shared__ Pt sp[kRegTile * NQ];
    for (int e = threadIdx.x; e < kRegTile * NQ; e += 256) {
        const int t = ti * kRegTile + e / NQ;
        sp[e] = pts[(size_t)(t < Nt ? t : Nt - 1) * NQ + e % NQ];
    }
    __syncthreads();
    const int t1c = t1 < Nt ? t1 : Nt - 1;
    const TriGeo g1 = tg[t1c];
    const double rad1 = rad[t1c];
    const Pt* p1 = sp + lane * NQ;

    constexpr int perWave = kRegTile / (4 * kRegSplit);
    const int t2base = tj * kRegTile + (int)blockIdx.y * (kRegTile / kRegSplit) + wave * perWave;
    for (int q = 0; q < perWave; ++q) {
        const int t2 = t2base + q;
        if (t2 >= Nt) break;
        const TriGeo g2 = tg[t2];
        // near pairs: the given map (a prepack, the surrogate's fallback
        // pairs) or the reference's near test itself
        const bool active = (t1 < t2) && !(nearMat ? nearMat[(size_t)t2 * Nt + t1] != 0 : near_test(g1, g2, eps));
        const Pt* p2 = pts + (size_t)t2 * NQ;
        const double ox = g1.cx - g2.cx, oy = g1.cy - g2.cy, oz = g1.cz - g2.cz;
        const PairGreen pg = pair_green(kr, ki, sqrt(ox * ox + oy * oy + oz * oz), rad1 + rad[t2]);
        const unsigned long long need2 = __ballot(active && pg.mode == 2);
        const unsigned long long need1 = __ballot(active && pg.mode == 1);
        const int mode = need2 ? 2 : (need1 ? 1 : 0);
        if (!active) continue;
        Acc a;
        if (lossy) regular_sums_mode<NQ, true>(mode, p1, p2, ox, oy, oz, pg, a);
        else regular_sums_mode<NQ, false>(mode, p1, p2, ox, oy, oz, pg, a);
        acc_scale(a, pg.e0r, pg.e0i);
        acc_store(recs + rec_index(t1, t2), a);
    }
}

// ------------------------------------------- subdivided (near / self) ---
// One wave per triangle pair.  The Np x Np point pairs are covered by lanes
// that each own one test point i and a contiguous range of source points j:
// rows are taken 64 at a time; when r < 64 rows remain, floor(64/r) lanes
// share a row and split its j range.  Partial moments are reduced across
// the wave at the end.

// The wave-reduced moments end up in res[16] of every lane.
template <bool SELF>
__device__ __forceinline__ void subdiv_moments(const Pt* __restrict__ P1, const Pt* __restrict__ P2, int Np,
                                               int Nq2, double ox, double oy, double oz, double aNear2,
                                               const double* __restrict__ aSelf, double kr, double ki,
                                               double dmax, double* res) {
    const int lane = threadIdx.x & 63;
    // the softened distance stays within dmax of sqrt(|o|^2 + a^2)
    const PairGreen pg = pair_green(kr, ki, sqrt(ox * ox + oy * oy + oz * oz + aNear2), dmax);
    Acc a;
    acc_zero(a);
    for (int base = 0; base < Np; base += 64) {
        const int r = min(64, Np - base);
        const int parts = 64 / r;
        const int part = lane / r;
        const int i = base + lane % r;
        if (part < parts) {
            const int j0 = (Np * part) / parts, j1 = (Np * (part + 1)) / parts;
            const Pt pi = P1[i];
            const double xi = pi.x + ox, yi = pi.y + oy, zi = pi.z + oz;
            const int psub = i / Nq2;
            double aS2 = 0.0;
            if (SELF) {
                double as = aSelf[psub];
                aS2 = as * as;
            }
            double gr = 0, gi = 0, hxr = 0, hxi = 0, hyr = 0, hyi = 0, hzr = 0, hzi = 0;
            for (int j = j0; j < j1; ++j) {
                const Pt q2 = P2[j];
                double dx = xi - q2.x, dy = yi - q2.y, dz = zi - q2.z;
                double R2 = dx * dx + dy * dy + dz * dz;
                if (!SELF) {
                    R2 = R2 + aNear2;            // near_singular.m:41-42
                } else if (j / Nq2 == psub) {
                    R2 = R2 + aS2;               // singular.m:46-47
                } else {
                    R2 = R2 > 1e-30 ? R2 : 1e-30;  // max(R, 1e-15), singular.m:48-49
                }
                double Gr, Gi;
                green_rel(pg, R2, Gr, Gi);
                double wr = q2.w * Gr, wi = q2.w * Gi;
                gr += wr; gi += wi;
                hxr += wr * q2.x; hxi += wi * q2.x;
                hyr += wr * q2.y; hyi += wi * q2.y;
                hzr += wr * q2.z; hzi += wi * q2.z;
            }
            acc_fold(a, pi.w, pi.x, pi.y, pi.z, gr, gi, hxr, hxi, hyr, hyi, hzr, hzi);
        }
    }
    acc_scale(a, pg.e0r, pg.e0i);
    double* v = &a.s0r;
#pragma unroll
    for (int k = 0; k < 16; ++k) res[k] = wave_sum(v[k]);
}

template <bool SELF>
__device__ void subdiv_pair(const Pt* __restrict__ P1, const Pt* __restrict__ P2, int Np,
                            int Nq2, double ox, double oy, double oz, double aNear2,
                            const double* __restrict__ aSelf, double kr, double ki, double dmax,
                            Rec* __restrict__ out) {
    double res[16];
    subdiv_moments<SELF>(P1, P2, Np, Nq2, ox, oy, oz, aNear2, aSelf, kr, ki, dmax, res);
    if ((threadIdx.x & 63) == 0) {
        double2* p = reinterpret_cast<double2*>(out->v);
#pragma unroll
        for (int k = 0; k < 8; ++k) p[k] = make_double2(res[2 * k], res[2 * k + 1]);
    }
}

__global__ __launch_bounds__(256) void k_rec_near(const Pt* __restrict__ subPts, int Np, int Nq2,
                                                  const TriGeo* __restrict__ tg,
                                                  const double* __restrict__ rad,
                                                  const double* __restrict__ aeq,
                                                  double near_alpha,
                                                  const int2* __restrict__ list, int nPairs,
                                                  const Freq* __restrict__ fr,
                                                  Rec* __restrict__ recs) {
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= nPairs) return;
    const int2 tp = list[pair];
    const TriGeo g1 = tg[tp.x], g2 = tg[tp.y];
    // a = near_alpha * min(sqrt(A1/pi), sqrt(A2/pi))  (near_singular.m:16-19)
    double a1 = aeq[tp.x], a2 = aeq[tp.y];
    double as = near_alpha * (a1 < a2 ? a1 : a2);
    subdiv_pair<false>(subPts + (size_t)tp.x * Np, subPts + (size_t)tp.y * Np, Np, Nq2,
                       g1.cx - g2.cx, g1.cy - g2.cy, g1.cz - g2.cz, as * as, nullptr,
                       fr->kr, fr->ki, rad[tp.x] + rad[tp.y], recs + rec_index(tp.x, tp.y));
}

__global__ __launch_bounds__(256) void k_rec_self(const Pt* __restrict__ selfPts, int Nq2, int Ns,
                                                  const double* __restrict__ aSelf,
                                                  const double* __restrict__ rad, double aSelfMax, int Nt,
                                                  const Freq* __restrict__ fr,
                                                  Rec* __restrict__ recs) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= Nt) return;
    const int Np = Ns * Nq2;
    const Pt* P = selfPts + (size_t)t * Np;
    // self pair: R0 = 0 and |R - R0| <= 2 rad + a
    subdiv_pair<true>(P, P, Np, Nq2, 0.0, 0.0, 0.0, 0.0, aSelf + (size_t)t * Ns, fr->kr, fr->ki,
                      2.0 * rad[t] + aSelfMax, recs + rec_index(t, t));
}

// ------------------------------------------------ surrogate labels ----
// One wave per pair of triangles given by 18 normalised vertex coordinates
// (test triangle t1 first).  The rule points of both rules are formed in
// LDS from barycentric tables, the subdivided softened near-pair moments
// (as k_rec_near) and the plain product-rule moments (as k_rec_regular) are
// integrated by the same wave code, and the label is their difference
// times exp(+j kt dt) 4 pi, dt being the centroid distance.

__global__ __launch_bounds__(256) void k_sur_labels(const double* __restrict__ tri18, const double* __restrict__ kt,
                                                    const Freq* __restrict__ fr, const SurPair* __restrict__ pairs,
                                                    int n, const double* __restrict__ subBary, int Np,
                                                    const double* __restrict__ regBary, int Nq, double near_alpha,
                                                    double* __restrict__ T, double* __restrict__ nrm) {
    extern __shared__ Pt lp[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pair = blockIdx.x * 4 + wave;
    if (pair >= n) return;
    Pt* s1 = lp + (size_t)wave * 2 * (Np + Nq);
    Pt* s2 = s1 + Np;
    Pt* r1 = s2 + Np;
    Pt* r2 = r1 + Nq;
    double v[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) v[i] = tri18[(size_t)pair * 18 + i];
    double c[6], a2[2], rad[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const double* w = v + 9 * t;
#pragma unroll
        for (int d = 0; d < 3; ++d) c[3 * t + d] = (w[d] + w[3 + d] + w[6 + d]) / 3.0;
        const double ex = w[3] - w[0], ey = w[4] - w[1], ez = w[5] - w[2];
        const double fx = w[6] - w[0], fy = w[7] - w[1], fz = w[8] - w[2];
        const double nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
        a2[t] = sqrt(nx * nx + ny * ny + nz * nz);
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double dx = w[3 * q] - c[3 * t], dy = w[3 * q + 1] - c[3 * t + 1], dz = w[3 * q + 2] - c[3 * t + 2];
            r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz));
        }
        rad[t] = r;
    }
    // rule points relative to each centroid, weight = fraction * twice the area
    for (int e = lane; e < 2 * (Np + Nq); e += 64) {
        const bool sub = e < 2 * Np;
        const int idx = sub ? e % Np : (e - 2 * Np) % Nq;
        const int t = sub ? e / Np : (e - 2 * Np) / Nq;
        const double* b = (sub ? subBary : regBary) + 4 * idx;
        const double* w = v + 9 * t;
        Pt q;
        q.x = b[0] * w[0] + b[1] * w[3] + b[2] * w[6] - c[3 * t];
        q.y = b[0] * w[1] + b[1] * w[4] + b[2] * w[7] - c[3 * t + 1];
        q.z = b[0] * w[2] + b[1] * w[5] + b[2] * w[8] - c[3 * t + 2];
        q.w = b[3] * a2[t];
        lp[(size_t)wave * 2 * (Np + Nq) + e] = q;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const double ox = c[0] - c[3], oy = c[1] - c[4], oz = c[2] - c[5];
    const double ae1 = sqrt(0.5 * a2[0] / kPi), ae2 = sqrt(0.5 * a2[1] / kPi);
    const double as = near_alpha * (ae1 < ae2 ? ae1 : ae2);
    const double k = fr ? fr->kr * pairs[pair].L : kt[pair];
    double resN[16], resB[16];
    subdiv_moments<false>(s1, s2, Np, Np, ox, oy, oz, as * as, nullptr, k, 0.0, rad[0] + rad[1], resN);
    subdiv_moments<false>(r1, r2, Nq, Nq, ox, oy, oz, 0.0, nullptr, k, 0.0, rad[0] + rad[1], resB);
    double sn, cs;
    sincos_cw(k * sqrt(ox * ox + oy * oy + oz * oz), &sn, &cs);
    cs *= 4.0 * kPi;
    sn *= 4.0 * kPi;
    if (lane < 8) {
        double dr = 0.0, di = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q == lane) {
                dr = resN[2 * q] - resB[2 * q];
                di = resN[2 * q + 1] - resB[2 * q + 1];
            }
        T[(size_t)pair * 16 + 2 * lane] = dr * cs - di * sn;
        T[(size_t)pair * 16 + 2 * lane + 1] = dr * sn + di * cs;
    }
    if (nrm && lane == 0) {
        double q = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) q += resN[i] * resN[i];
        nrm[pair] = 4.0 * kPi * sqrt(q);
    }
}

// ------------------------------------ frequency-independent moments ---
// With G(R) = (1/4pi) sum_n (-jk)^n R^(n-1) / n!, every record of a near or
// self pair is S_X(k) = (1/4pi) sum_{n=0..P} ((-jk)^n/n!) Q_X[n] where
//   Q_X[q] = sum_ij W_ij X_ij R_ij^(q-1),   X in {1, rho, rho', rho.rho'},
// R being the (softened) distance of the reference scheme.  The Q depend
// only on the mesh, so the 9216 Green evaluations of a near pair are done
// once per mesh, and each frequency costs (P+1) complex multiply-adds per
// moment.  P is chosen on the host so that the series tail is below 1e-17.
// Layout: Q[pair][q][8] with X order 1, rho x/y/z, rho' x/y/z, rho.rho'.

constexpr int kMomChunk = 4;

template <bool SELF, bool EXTRACT = false>
__device__ void moment_pair(const Pt* __restrict__ P1, const Pt* __restrict__ P2, int Np, int Nq2,
                            double ox, double oy, double oz, double aNear2,
                            const double* __restrict__ aSelf, int P, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    for (int q0 = 0; q0 <= P; q0 += kMomChunk) {
        double acc[kMomChunk][8];
#pragma unroll
        for (int c = 0; c < kMomChunk; ++c)
#pragma unroll
            for (int x = 0; x < 8; ++x) acc[c][x] = 0.0;
        for (int base = 0; base < Np; base += 64) {
            const int r = min(64, Np - base);
            const int parts = 64 / r;
            const int part = lane / r;
            const int i = base + lane % r;
            if (part >= parts) continue;
            const int j0 = (Np * part) / parts, j1 = (Np * (part + 1)) / parts;
            const Pt pi = P1[i];
            const double xi = pi.x + ox, yi = pi.y + oy, zi = pi.z + oz;
            const int psub = i / Nq2;
            double aS2 = 0.0;
            if (SELF) {
                double as = aSelf[psub];
                aS2 = as * as;
            }
            double g[kMomChunk], hx[kMomChunk], hy[kMomChunk], hz[kMomChunk];
#pragma unroll
            for (int c = 0; c < kMomChunk; ++c) g[c] = hx[c] = hy[c] = hz[c] = 0.0;
            for (int j = j0; j < j1; ++j) {
                const Pt q2 = P2[j];
                double dx = xi - q2.x, dy = yi - q2.y, dz = zi - q2.z;
                double R2 = dx * dx + dy * dy + dz * dz;
                double R;
                if (EXTRACT) {
                    R = sqrt(R2);  // unsoftened; only q >= 1 (R^0, R^1, ...) is summed
                } else if (!SELF) {
                    R = sqrt(R2 + aNear2);
                } else if (j / Nq2 == psub) {
                    R = sqrt(R2 + aS2);
                } else {
                    R = sqrt(R2);
                    R = R > 1e-15 ? R : 1e-15;
                }
                if (EXTRACT) {
                    // R^(q-1) for q >= 1; the q = 0 slot (1/R) is the closed-form
                    // static part written by k_static_extract
                    double rp = 1.0;
                    for (int e = 1; e < q0; ++e) rp *= R;
#pragma unroll
                    for (int c = 0; c < kMomChunk; ++c) {
                        const double t = (q0 + c == 0) ? 0.0 : q2.w * rp;
                        g[c] += t;
                        hx[c] += t * q2.x;
                        hy[c] += t * q2.y;
                        hz[c] += t * q2.z;
                        if (q0 + c >= 1) rp *= R;
                    }
                } else {
                    // R^(q0-1)
                    double rp = 1.0 / R;
                    for (int e = 0; e < q0; ++e) rp *= R;
#pragma unroll
                    for (int c = 0; c < kMomChunk; ++c) {
                        const double t = q2.w * rp;
                        g[c] += t;
                        hx[c] += t * q2.x;
                        hy[c] += t * q2.y;
                        hz[c] += t * q2.z;
                        rp *= R;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kMomChunk; ++c) {
                const double w = pi.w;
                acc[c][0] += w * g[c];
                acc[c][1] += w * g[c] * pi.x;
                acc[c][2] += w * g[c] * pi.y;
                acc[c][3] += w * g[c] * pi.z;
                acc[c][4] += w * hx[c];
                acc[c][5] += w * hy[c];
                acc[c][6] += w * hz[c];
                acc[c][7] += w * (pi.x * hx[c] + pi.y * hy[c] + pi.z * hz[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < kMomChunk; ++c)
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                double v = wave_sum(acc[c][x]);
                if (lane == 0 && q0 + c <= P) out[(size_t)(q0 + c) * 8 + x] = v;
            }
    }
}

template <bool EXTRACT>
__global__ __launch_bounds__(256) void k_moment_near(const Pt* __restrict__ subPts, int Np,
                                                     const TriGeo* __restrict__ tg,
                                                     const double* __restrict__ aeq,
                                                     double near_alpha,
                                                     const int2* __restrict__ list, int nPairs,
                                                     int P, double* __restrict__ Q) {
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= nPairs) return;
    const int2 tp = list[pair];
    const TriGeo g1 = tg[tp.x], g2 = tg[tp.y];
    double a1 = aeq[tp.x], a2 = aeq[tp.y];
    double as = near_alpha * (a1 < a2 ? a1 : a2);
    moment_pair<false, EXTRACT>(subPts + (size_t)tp.x * Np, subPts + (size_t)tp.y * Np, Np, Np,
                                g1.cx - g2.cx, g1.cy - g2.cy, g1.cz - g2.cz, as * as, nullptr, P,
                                Q + (size_t)pair * (P + 1) * 8);
}

template <bool EXTRACT>
__global__ __launch_bounds__(256) void k_moment_self(const Pt* __restrict__ selfPts, int Nq2,
                                                     int Ns, const double* __restrict__ aSelf,
                                                     int Nt, int P, double* __restrict__ Q) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= Nt) return;
    const int Np = Ns * Nq2;
    const Pt* Pp = selfPts + (size_t)t * Np;
    moment_pair<true, EXTRACT>(Pp, Pp, Np, Nq2, 0.0, 0.0, 0.0, 0.0, aSelf + (size_t)t * Ns, P,
                               Q + (size_t)t * (P + 1) * 8);
}

// ------------------------------------------- singularity extraction -----
// Mode "extract": G = 1/(4 pi R) + (exp(-jkR) - 1)/(4 pi R).  The smooth
// part is the q >= 1 tail of the moment series above (unsoftened); the
// static part replaces the q = 0 moments: for every outer quadrature point
// r of the lower triangle, the inner integrals over the whole other
// triangle T2 are taken in closed form,
//   int_T2 1/R dS'            = I0(r)
//   int_T2 (r' - c2)/R dS'    = I1(r) + (rho - c2) I0(r)
// (rho = projection of r onto T2's plane), and summed with the outer
// weights into the eight moments.
//
// Closed form (Wilton, Rao, Glisson, Schaubert, Al-Bundak and Butler, IEEE
// TAP 32(3), 1984; Graglia, IEEE TAP 41(10), 1993), edge by edge (edge i
// from v_i to v_i+1, counter-clockwise about the unit normal n): l-, l+
// in-line coordinates of the edge ends, P0 signed distance of rho from the
// edge line, h height of r above the plane, R0^2 = P0^2 + h^2, R-, R+ the
// distances to the edge ends, u the outward in-plane edge normal:
//   f = ln((R+ + l+)/(R- + l-)),  b = atan(P0 l+/(R0^2 + |h| R+)) - atan(P0 l-/(R0^2 + |h| R-))
//   I0 = sum P0 f - |h| b,   I1 = 1/2 sum u (R0^2 f + l+ R+ - l- R-).
// f is evaluated without cancellation: ln((R- - l-)/(R+ - l+)) when both l
// are negative, ln((R+ + l+)(R- - l-)/R0^2) when rho projects inside the
// edge; an edge whose line passes through r contributes no f term.
__device__ void tri_potentials(double rx, double ry, double rz, const double* __restrict__ v, double& I0,
                               double& I1x, double& I1y, double& I1z, double& hx, double& hy, double& hz) {
    const double e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    const double e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double nl = sqrt(nx * nx + ny * ny + nz * nz);
    nx /= nl;
    ny /= nl;
    nz /= nl;
    const double h = (rx - v[0]) * nx + (ry - v[1]) * ny + (rz - v[2]) * nz;
    const double ah = fabs(h);
    const double px = rx - h * nx, py = ry - h * ny, pz = rz - h * nz;  // rho
    const double scale = sqrt(e1x * e1x + e1y * e1y + e1z * e1z) + sqrt(e2x * e2x + e2y * e2y + e2z * e2z);
    I0 = 0.0;
    I1x = I1y = I1z = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double* a = v + 3 * i;
        const double* b = v + 3 * ((i + 1) % 3);
        const double sx = b[0] - a[0], sy = b[1] - a[1], sz = b[2] - a[2];
        const double L = sqrt(sx * sx + sy * sy + sz * sz);
        const double lx = sx / L, ly = sy / L, lz = sz / L;
        const double ux = ly * nz - lz * ny, uy = lz * nx - lx * nz, uz = lx * ny - ly * nx;
        const double amx = a[0] - px, amy = a[1] - py, amz = a[2] - pz;
        const double lm = amx * lx + amy * ly + amz * lz;
        const double lp = (b[0] - px) * lx + (b[1] - py) * ly + (b[2] - pz) * lz;
        const double P0 = amx * ux + amy * uy + amz * uz;
        const double R02 = P0 * P0 + h * h;
        const double Rm = sqrt(R02 + lm * lm), Rp = sqrt(R02 + lp * lp);
        double f;
        if (R02 <= 1e-28 * scale * scale) f = 0.0;
        else if (lm >= 0.0) f = log((Rp + lp) / (Rm + lm));
        else if (lp <= 0.0) f = log((Rm - lm) / (Rp - lp));
        else f = log((Rp + lp) * (Rm - lm) / R02);
        double beta = 0.0;
        if (ah > 0.0) beta = atan(P0 * lp / (R02 + ah * Rp)) - atan(P0 * lm / (R02 + ah * Rm));
        I0 += P0 * f - ah * beta;
        const double c = 0.5 * (R02 * f + lp * Rp - lm * Rm);
        I1x += c * ux;
        I1y += c * uy;
        I1z += c * uz;
    }
    hx = px;
    hy = py;
    hz = pz;
}

// Static moments of one pair (one wave): outer points P1 relative to
// centroid c1, inner triangle vertices v2 relative to c2, o = c1 - c2.
__device__ void static_pair(const Pt* __restrict__ P1, int Np, const double* __restrict__ v2, double ox,
                            double oy, double oz, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = lane; i < Np; i += 64) {
        const Pt p = P1[i];
        double I0, I1x, I1y, I1z, rhx, rhy, rhz;
        tri_potentials(p.x + ox, p.y + oy, p.z + oz, v2, I0, I1x, I1y, I1z, rhx, rhy, rhz);
        // int (r' - c2)/R = I1 + (rho - c2) I0, rho relative to c2 already
        const double vx = I1x + rhx * I0, vy = I1y + rhy * I0, vz = I1z + rhz * I0;
        acc[0] += p.w * I0;
        acc[1] += p.w * I0 * p.x;
        acc[2] += p.w * I0 * p.y;
        acc[3] += p.w * I0 * p.z;
        acc[4] += p.w * vx;
        acc[5] += p.w * vy;
        acc[6] += p.w * vz;
        acc[7] += p.w * (p.x * vx + p.y * vy + p.z * vz);
    }
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const double v = wave_sum(acc[x]);
        if (lane == 0) out[x] = v;
    }
}

// Near pairs (SELF false: list entries, outer = the lower triangle) or
// self pairs (SELF true: triangle t with itself); writes Q[pair][0][*].
template <bool SELF>
__global__ __launch_bounds__(256) void k_static_extract(const Pt* __restrict__ pts, int Np,
                                                        const TriGeo* __restrict__ tg,
                                                        const double* __restrict__ triV,
                                                        const int2* __restrict__ list, int nPairs, int P,
                                                        double* __restrict__ Q) {
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= nPairs) return;
    const int2 tp = SELF ? make_int2(pair, pair) : list[pair];
    const TriGeo g1 = tg[tp.x], g2 = tg[tp.y];
    static_pair(pts + (size_t)tp.x * Np, Np, triV + 9 * (size_t)tp.y, g1.cx - g2.cx, g1.cy - g2.cy,
                g1.cz - g2.cz, Q + (size_t)pair * (P + 1) * 8);
}

// Records of near pairs (threads 0..nNear-1) and self pairs (the next Nt)
// at one frequency from the moments.
__global__ __launch_bounds__(256) void k_poly_records(const double* __restrict__ Qn,
                                                      const int2* __restrict__ list, int nNear,
                                                      const double* __restrict__ Qs, int Nt,
                                                      int P, const Freq* __restrict__ fr,
                                                      Rec* __restrict__ recs) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nNear + Nt) return;
    const double* q;
    size_t ri;
    if (id < nNear) {
        const int2 tp = list[id];
        q = Qn + (size_t)id * (P + 1) * 8;
        ri = rec_index(tp.x, tp.y);
    } else {
        const int t = id - nNear;
        q = Qs + (size_t)t * (P + 1) * 8;
        ri = rec_index(t, t);
    }
    // c_n = (-jk)^n / n! / (4 pi);  -jk = ki - j kr
    const double mr = fr->ki, mi = -fr->kr;
    double cr = 1.0 / (4.0 * kPi), ci = 0.0;
    double sr[8], si[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) { sr[x] = 0.0; si[x] = 0.0; }
    for (int n = 0; n <= P; ++n) {
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const double v = q[(size_t)n * 8 + x];
            sr[x] += cr * v;
            si[x] += ci * v;
        }
        const double inv = 1.0 / (double)(n + 1);
        const double tr = (cr * mr - ci * mi) * inv, ti = (cr * mi + ci * mr) * inv;
        cr = tr;
        ci = ti;
    }
    double2* o = reinterpret_cast<double2*>(recs[ri].v);
#pragma unroll
    for (int x = 0; x < 8; ++x) o[x] = make_double2(sr[x], si[x]);
}

// ------------------------------------------------------------ combine ---

// Z is complex symmetric (Z(m,n) = Z(n,m): the four side pairs of (n,m)
// are those of (m,n) with the roles swapped), so one block computes a
// 16 x 16 tile (I, J) with I <= J and writes it twice, the mirror through
// LDS so both stores are row-contiguous (on a diagonal tile the entries
// with m <= n, mirrored to the others).  With Z32 set, the complex64 copy
// of the mixed-precision solve is written in the same pass.
//
// Eight lanes per entry.  Lane q of a group holds moment q of each of the
// entry's four records, so a group reads whole 128-byte records and a wave
// load instruction touches 8 cache lines (with a lane per entry: 64).
// Every moment enters the entry with a real coefficient (the side vectors,
// their dot product, 1, and 4 for the scalar term), so each lane sums its
// weighted moment over the four side pairs and the group adds its 8 lanes
// in registers (DPP).  The block takes its tile in 8 passes of 32 entries,
// the tile's side tables staged in LDS.  Record loads are unconditional
// (triangle indices clamped, missing triangles weighted 0).
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// sum over the 8 lanes of an aligned group, the same bits in every lane
// (swap neighbours, swap pairs, mirror the group)
__device__ __forceinline__ double oct_sum(double v) {
    v += dpp_mov_d<0xB1>(v);
    v += dpp_mov_d<0x4E>(v);
    v += dpp_mov_d<0x141>(v);
    return v;
}

__global__ __launch_bounds__(256) void k_combine(const Rec* __restrict__ recs, const int* __restrict__ sideTri,
                                                  const Side* __restrict__ side, int Ne,
                                                  const Freq* __restrict__ fr, double* __restrict__ Z, size_t ld,
                                                  float* __restrict__ Z32, size_t ld32, int nblk) {
    __shared__ double2 tile[16][17];
    __shared__ int sT[2][16][2];
    __shared__ Side sS[2][16][2];
    // XCD-aware order (each XCD takes a contiguous run of tiles), then the
    // linear tile index -> (I <= J)
    long long id = blockIdx.x;
    {
        const int q = nblk >> 3, rr = nblk & 7, x = (int)(id & 7), loc = (int)(id >> 3);
        id = x < rr ? (long long)x * (q + 1) + loc : (long long)rr * (q + 1) + (long long)(x - rr) * q + loc;
    }
    int J = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((long long)(J + 1) * (J + 2) / 2 <= id) ++J;
    while ((long long)J * (J + 1) / 2 > id) --J;
    const int I = (int)(id - (long long)J * (J + 1) / 2);
    if (threadIdx.x < 64) {
        // [0]: the tile's 16 rows (edges 16I..), [1]: its 16 columns
        const int which = threadIdx.x >> 5, e = (threadIdx.x & 31) >> 1, sd = threadIdx.x & 1;
        const int edge = 16 * (which ? J : I) + e;
        const int ec = edge < Ne ? edge : Ne - 1;
        sT[which][e][sd] = sideTri[2 * ec + sd];
        sS[which][e][sd] = side[2 * ec + sd];
    }
    const Freq f = *fr;
    __syncthreads();
    const int g = threadIdx.x >> 3, q = threadIdx.x & 7;
    const int d = (q - 1) % 3;  // component of the side vector for q = 1..6
#pragma unroll 2
    for (int ps = 0; ps < 8; ++ps) {
        const int e = ps * 32 + g, r = e >> 4, c = e & 15;
        int tm[2], tn[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            tm[k] = sT[0][r][k];
            tn[k] = sT[1][c][k];
        }
        double2 R[2][2];
#pragma unroll
        for (int sa = 0; sa < 2; ++sa)
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) {
                const int t1 = tm[sa] < 0 ? 0 : tm[sa], t2 = tn[sb] < 0 ? 0 : tn[sb];
                R[sa][sb] = reinterpret_cast<const double2*>(
                    recs[t1 > t2 ? rec_index(t2, t1) : rec_index(t1, t2)].v)[q];
            }
        double vr = 0.0, vi = 0.0, sr = 0.0, si = 0.0;
#pragma unroll
        for (int sa = 0; sa < 2; ++sa)
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) {
                const Side a = sS[0][r][sa], b = sS[1][c][sb];
                // S2 - S1.un - S1'.um + S0 (um.un), S1 and S1' exchanged
                // when the record is stored the other way round
                const bool sw = tm[sa] > tn[sb];
                const double ua = d == 0 ? a.ux : (d == 1 ? a.uy : a.uz);
                const double ub = d == 0 ? b.ux : (d == 1 ? b.uy : b.uz);
                double coef;
                if (q == 0) coef = a.ux * b.ux + a.uy * b.uy + a.uz * b.uz;
                else if (q < 4) coef = -(sw ? ua : ub);
                else if (q < 7) coef = -(sw ? ub : ua);
                else coef = 1.0;
                // a missing triangle (index -1) contributes nothing
                const double w = (tm[sa] >= 0 && tn[sb] >= 0) ? a.c * b.c : 0.0;
                const double wc = w * coef;
                vr += wc * R[sa][sb].x;
                vi += wc * R[sa][sb].y;
                sr += w * R[sa][sb].x;  // the scalar term's factor 4 is applied
                si += w * R[sa][sb].y;  // once below (exact: a power of two)
            }
        vr = oct_sum(vr);
        vi = oct_sum(vi);
        sr *= 4.0;
        si *= 4.0;
        if (q == 0)
            tile[r][c] = make_double2((f.ar * vr - f.ai * vi) + (f.br * sr - f.bi * si),
                                      (f.ar * vi + f.ai * vr) + (f.br * si + f.bi * sr));
    }
    __syncthreads();
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
    const int m = 16 * I + r, n = 16 * J + c;
    if (m < Ne && n < Ne && (I != J || r <= c)) {
        const double2 v = tile[r][c];
        *reinterpret_cast<double2*>(Z + 2 * ((size_t)m * ld + n)) = v;
        if (Z32) *reinterpret_cast<float2*>(Z32 + 2 * ((size_t)m * ld32 + n)) = make_float2((float)v.x, (float)v.y);
    }
    const int mm = 16 * J + r, nn = 16 * I + c;
    if (mm < Ne && nn < Ne && (I != J || r > c)) {
        const double2 v = tile[c][r];
        *reinterpret_cast<double2*>(Z + 2 * ((size_t)mm * ld + nn)) = v;
        if (Z32) *reinterpret_cast<float2*>(Z32 + 2 * ((size_t)mm * ld32 + nn)) = make_float2((float)v.x, (float)v.y);
    }
}

// --------------------------------------- frequency-independent series ----
// The fill of near_mode MOMENTS, for every triangle pair.  With R0 the
// centroid distance of the pair (0 for a triangle with itself) and
// z = -jk, the Green's function factors as
//   e^{-jkR}/R = e^{-jk R0} sum_q z^q (R - R0)^q / (q! R),
// so the eight moments of the pair's record are
//   e^{-jk R0}/(4 pi) sum_q z^q Q[q][x],  Q[q][x] = sum W X (R - R0)^q/(q! R)
// with real, frequency-independent Q (X = 1, rho, rho', rho.rho'; R the
// same (softened) distance the reference rule uses for the pair).  Q is
// integrated once per mesh for q = 0..P, P chosen so the series tail is
// below 1e-17 relative for the sweep's largest |k|.
//
// With alpha = j omega mu = -eta z and beta = 1/(j omega eps) = -eta/z
// (eta = sqrt(mu/eps)), one pair's share of Z(m,n) is
//   alpha Ivec + beta Isca = -eta e^{-jkR0}/(4 pi) sum_{s=-1}^{P+1} z^s D[s],
//   D[s] = V[s-1] + S[s+1],
// where V[q] and S[q] are the vector and scalar combinations of Q[q] with
// the two RWG sides (the coefficients of the record combine below).  The
// four pairs of an entry are brought to one phase centre R0e (their mean
// R0): e^{-jk R0} = e^{-jk R0e} e^{z (R0 - R0e)}, and multiplying a series
// by e^{z d} is the exact Cauchy product with d^i/i!.  Per frequency an
// entry then costs one phase and one polynomial in z.


// (R - R0)^q / q! / R for q = q0.. is t, t*d/(q0+1), ...: 1/n for n = 1..32
__constant__ double kInvN[33] = {0.0,        1.0,        1.0 / 2,    1.0 / 3,    1.0 / 4,    1.0 / 5,    1.0 / 6,
                                 1.0 / 7,    1.0 / 8,    1.0 / 9,    1.0 / 10,   1.0 / 11,   1.0 / 12,   1.0 / 13,
                                 1.0 / 14,   1.0 / 15,   1.0 / 16,   1.0 / 17,   1.0 / 18,   1.0 / 19,   1.0 / 20,
                                 1.0 / 21,   1.0 / 22,   1.0 / 23,   1.0 / 24,   1.0 / 25,   1.0 / 26,   1.0 / 27,
                                 1.0 / 28,   1.0 / 29,   1.0 / 30,   1.0 / 31,   1.0 / 32};

// Orders summed per pass over the point pairs: regular pairs (49 point
// pairs, many pairs per block) and subdivided pairs (one wave per pair).
constexpr int kSerReg = 4;
constexpr int kSerSub = 8;

constexpr double inv_factorial(int q) {
    double r = 1.0;
    for (int i = 2; i <= q; ++i) r /= i;
    return r;
}
// 1/q! for q = 0..32
__constant__ double kInvFact[33] = {
    inv_factorial(0),  inv_factorial(1),  inv_factorial(2),  inv_factorial(3),  inv_factorial(4),
    inv_factorial(5),  inv_factorial(6),  inv_factorial(7),  inv_factorial(8),  inv_factorial(9),
    inv_factorial(10), inv_factorial(11), inv_factorial(12), inv_factorial(13), inv_factorial(14),
    inv_factorial(15), inv_factorial(16), inv_factorial(17), inv_factorial(18), inv_factorial(19),
    inv_factorial(20), inv_factorial(21), inv_factorial(22), inv_factorial(23), inv_factorial(24),
    inv_factorial(25), inv_factorial(26), inv_factorial(27), inv_factorial(28), inv_factorial(29),
    inv_factorial(30), inv_factorial(31), inv_factorial(32)};

// d^q0 / q0! times w/R, q0 a multiple of the pass width C (a power of 2):
// (d^C)^(q0/C) / q0!
template <int C>
__device__ __forceinline__ double series_head(double wInvR, double d, int q0) {
    static_assert((C & (C - 1)) == 0, "pass width must be a power of 2");
    double t = wInvR * kInvFact[q0];
    double dc = d;
#pragma unroll
    for (int k = 1; k < C; k <<= 1) dc *= dc;
    for (int e = 0; e < q0; e += C) t *= dc;
    return t;
}

// Regular pairs: tile of 64 test triangles (one per lane) by 64 source
// triangles, as the per-frequency regular kernel; each pair's NQ x NQ
// product rule is summed in chunks of kSerReg orders.
template <int NQ>
__global__ __launch_bounds__(256) void k_series_regular(const Pt* __restrict__ pts, const TriGeo* __restrict__ tg,
                                                        int Nt, const uint8_t* __restrict__ nearMat, double eps,
                                                        int P, double* __restrict__ Q) {
    long long b = blockIdx.x;
    int tj = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((long long)(tj + 1) * (tj + 2) / 2 <= b) ++tj;
    while ((long long)tj * (tj + 1) / 2 > b) --tj;
    int ti = (int)(b - (long long)tj * (tj + 1) / 2);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t1 = ti * kRegTile + lane;
    __shared__ Pt sp[kRegTile * NQ];
    for (int e = threadIdx.x; e < kRegTile * NQ; e += 256) {
        const int t = ti * kRegTile + e / NQ;
        sp[e] = pts[(size_t)(t < Nt ? t : Nt - 1) * NQ + e % NQ];
    }
    __syncthreads();
    const int t1c = t1 < Nt ? t1 : Nt - 1;
    const TriGeo g1 = tg[t1c];
    const Pt* p1 = sp + lane * NQ;
    constexpr int perWave = kRegTile / (4 * kRegSplit);
    const int t2base = tj * kRegTile + (int)blockIdx.y * (kRegTile / kRegSplit) + wave * perWave;
    const size_t stride = (size_t)(P + 1) * 8;
    for (int qq = 0; qq < perWave; ++qq) {
        const int t2 = t2base + qq;
        if (t2 >= Nt) break;
        const TriGeo g2 = tg[t2];
        if (!(t1 < t2) || (nearMat ? nearMat[(size_t)t2 * Nt + t1] != 0 : near_test(g1, g2, eps))) continue;
        const Pt* p2 = pts + (size_t)t2 * NQ;
        const double ox = g1.cx - g2.cx, oy = g1.cy - g2.cy, oz = g1.cz - g2.cz;
        const double R0 = sqrt(ox * ox + oy * oy + oz * oz);
        double* out = Q + rec_index(t1, t2) * stride;
        for (int q0 = 0; q0 <= P; q0 += kSerReg) {
            double acc[kSerReg][8];
#pragma unroll
            for (int c = 0; c < kSerReg; ++c)
#pragma unroll
                for (int x = 0; x < 8; ++x) acc[c][x] = 0.0;
#pragma unroll 1
            for (int i = 0; i < NQ; ++i) {
                const Pt pi = p1[i];
                const double xi = pi.x + ox, yi = pi.y + oy, zi = pi.z + oz;
                double g[kSerReg], hx[kSerReg], hy[kSerReg], hz[kSerReg];
#pragma unroll
                for (int c = 0; c < kSerReg; ++c) g[c] = hx[c] = hy[c] = hz[c] = 0.0;
#pragma unroll
                for (int j = 0; j < NQ; ++j) {
                    const Pt q2 = p2[j];
                    const double dx = xi - q2.x, dy = yi - q2.y, dz = zi - q2.z;
                    // R < 1e-15 -> 1e-15, as the regular rule
                    const double R2 = fmax(dx * dx + dy * dy + dz * dz, 1e-30);
                    const double invR = rsqrt64(R2);
                    const double d = R2 * invR - R0;
                    double t = series_head<kSerReg>(q2.w * invR, d, q0);
#pragma unroll
                    for (int c = 0; c < kSerReg; ++c) {
                        g[c] += t;
                        hx[c] += t * q2.x;
                        hy[c] += t * q2.y;
                        hz[c] += t * q2.z;
                        t *= d * kInvN[q0 + c + 1];
                    }
                }
#pragma unroll
                for (int c = 0; c < kSerReg; ++c) {
                    const double w = pi.w;
                    acc[c][0] += w * g[c];
                    acc[c][1] += w * g[c] * pi.x;
                    acc[c][2] += w * g[c] * pi.y;
                    acc[c][3] += w * g[c] * pi.z;
                    acc[c][4] += w * hx[c];
                    acc[c][5] += w * hy[c];
                    acc[c][6] += w * hz[c];
                    acc[c][7] += w * (pi.x * hx[c] + pi.y * hy[c] + pi.z * hz[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < kSerReg; ++c)
                if (q0 + c <= P) {
                    double2* o2 = reinterpret_cast<double2*>(out + (size_t)(q0 + c) * 8);
#pragma unroll
                    for (int x = 0; x < 4; ++x) o2[x] = make_double2(acc[c][2 * x], acc[c][2 * x + 1]);
                }
        }
    }
}

// Subdivided pairs (one wave per pair): near pairs with the softened
// distance sqrt(R^2 + a^2) on every point pair (near_singular.m:41-42),
// self pairs (R0 = 0) softened on coincident sub-triangles and clamped at
// 1e-15 elsewhere (singular.m:46-50).  Lanes split the Np x Np point pairs
// as the per-frequency subdivided kernel does.
template <bool SELF>
__device__ void series_subdiv(const Pt* __restrict__ P1, const Pt* __restrict__ P2, int Np, int Nq2, double ox,
                              double oy, double oz, double R0, double aNear2, const double* __restrict__ aSelf, int P,
                              double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    for (int q0 = 0; q0 <= P; q0 += kSerSub) {
        double acc[kSerSub][8];
#pragma unroll
        for (int c = 0; c < kSerSub; ++c)
#pragma unroll
            for (int x = 0; x < 8; ++x) acc[c][x] = 0.0;
        for (int base = 0; base < Np; base += 64) {
            const int r = min(64, Np - base);
            const int parts = 64 / r;
            const int part = lane / r;
            const int i = base + lane % r;
            if (part >= parts) continue;
            const int j0 = (Np * part) / parts, j1 = (Np * (part + 1)) / parts;
            const Pt pi = P1[i];
            const double xi = pi.x + ox, yi = pi.y + oy, zi = pi.z + oz;
            const int psub = i / Nq2;
            double aS2 = 0.0;
            if (SELF) {
                const double as = aSelf[psub];
                aS2 = as * as;
            }
            double g[kSerSub], hx[kSerSub], hy[kSerSub], hz[kSerSub];
#pragma unroll
            for (int c = 0; c < kSerSub; ++c) g[c] = hx[c] = hy[c] = hz[c] = 0.0;
            for (int j = j0; j < j1; ++j) {
                const Pt q2 = P2[j];
                const double dx = xi - q2.x, dy = yi - q2.y, dz = zi - q2.z;
                double R2 = dx * dx + dy * dy + dz * dz;
                if (!SELF) {
                    R2 = R2 + aNear2;
                } else if (j / Nq2 == psub) {
                    R2 = R2 + aS2;
                } else {
                    R2 = R2 > 1e-30 ? R2 : 1e-30;  // max(R, 1e-15)
                }
                const double invR = rsqrt64(R2);
                const double d = R2 * invR - R0;
                double t = series_head<kSerSub>(q2.w * invR, d, q0);
#pragma unroll
                for (int c = 0; c < kSerSub; ++c) {
                    g[c] += t;
                    hx[c] += t * q2.x;
                    hy[c] += t * q2.y;
                    hz[c] += t * q2.z;
                    t *= d * kInvN[q0 + c + 1];
                }
            }
#pragma unroll
            for (int c = 0; c < kSerSub; ++c) {
                const double w = pi.w;
                acc[c][0] += w * g[c];
                acc[c][1] += w * g[c] * pi.x;
                acc[c][2] += w * g[c] * pi.y;
                acc[c][3] += w * g[c] * pi.z;
                acc[c][4] += w * hx[c];
                acc[c][5] += w * hy[c];
                acc[c][6] += w * hz[c];
                acc[c][7] += w * (pi.x * hx[c] + pi.y * hy[c] + pi.z * hz[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < kSerSub; ++c)
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const double v = wave_sum(acc[c][x]);
                if (lane == 0 && q0 + c <= P) out[(size_t)(q0 + c) * 8 + x] = v;
            }
    }
}

__global__ __launch_bounds__(256) void k_series_near(const Pt* __restrict__ subPts, int Np, int Nq2,
                                                     const TriGeo* __restrict__ tg, const double* __restrict__ aeq,
                                                     double near_alpha, const int2* __restrict__ list, int nPairs,
                                                     int P, double* __restrict__ Q) {
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= nPairs) return;
    const int2 tp = list[pair];
    const TriGeo g1 = tg[tp.x], g2 = tg[tp.y];
    // a = near_alpha * min(sqrt(A1/pi), sqrt(A2/pi))  (near_singular.m:16-19)
    const double a1 = aeq[tp.x], a2 = aeq[tp.y];
    const double as = near_alpha * (a1 < a2 ? a1 : a2);
    const double ox = g1.cx - g2.cx, oy = g1.cy - g2.cy, oz = g1.cz - g2.cz;
    series_subdiv<false>(subPts + (size_t)tp.x * Np, subPts + (size_t)tp.y * Np, Np, Nq2, ox, oy, oz,
                         sqrt(ox * ox + oy * oy + oz * oz), as * as, nullptr, P,
                         Q + rec_index(tp.x, tp.y) * (size_t)(P + 1) * 8);
}

__global__ __launch_bounds__(256) void k_series_self(const Pt* __restrict__ selfPts, int Nq2, int Ns,
                                                     const double* __restrict__ aSelf, int Nt, int P,
                                                     double* __restrict__ Q) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= Nt) return;
    const int Np = Ns * Nq2;
    const Pt* Pp = selfPts + (size_t)t * Np;
    series_subdiv<true>(Pp, Pp, Np, Nq2, 0.0, 0.0, 0.0, 0.0, 0.0, aSelf + (size_t)t * Ns, P,
                        Q + rec_index(t, t) * (size_t)(P + 1) * 8);
}

// Per-entry coefficients of the series (once per mesh and series length):
// one block per 16 x 16 upper-triangular tile of Z, one entry per thread.
// Each thread combines its entry's four pair series with the RWG side
// coefficients into D[0..NT-1] (NT = P + 3 orders z^-1 .. z^(P+1)) about
// the common centre R0e (the middle of the four centroid distances) and
// stores them, with R0e, as
// C[(tile * (NT + 1) + j) * 256 + thread] (coefficient-major per tile, so
// the fill reads them with coalesced loads).
//
// The tile's entries draw on few triangle pairs: the distinct triangles of
// its 16 row and 16 column RWGs (at most 32 each, usually ~20) form every
// pair any entry needs.  Those pairs are staged through LDS one order at a
// time (8 lanes per pair record, coalesced; the next order in flight in
// registers), so each record is fetched once per tile instead of once per
// entry that uses it, and every entry shifts each order's contribution to
// the common centre as it arrives (D[q + k] += c d^k / k!).  A tile with
// more than kCoefSlots pairs reads the records directly.
constexpr int kCoefSlots = 512;

template <int NT>
__global__ __launch_bounds__(256) void k_series_coeffs(const double* __restrict__ Q, const TriGeo* __restrict__ tg,
                                                       const int* __restrict__ sideTri,
                                                       const Side* __restrict__ side, int Ne,
                                                       double* __restrict__ C) {
    constexpr int P = NT - 3;
    constexpr size_t stride = (size_t)(P + 1) * 8;
    __shared__ double qs[kCoefSlots][8];
    __shared__ size_t slotRec[kCoefSlots];
    __shared__ int rl[32], cl[32], cnt[2];
    const long long id = blockIdx.x;
    int J = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((long long)(J + 1) * (J + 2) / 2 <= id) ++J;
    while ((long long)J * (J + 1) / 2 > id) --J;
    const int I = (int)(id - (long long)J * (J + 1) / 2);
    const int r = threadIdx.x >> 4, cI = threadIdx.x & 15;
    const int m = 16 * I + r, n = 16 * J + cI;
    const int mc = m < Ne ? m : Ne - 1, nc = n < Ne ? n : Ne - 1;

    int tm[2], tn[2];
    Side sa[2], sb[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        tm[k] = sideTri[2 * mc + k];
        tn[k] = sideTri[2 * nc + k];
        sa[k] = side[2 * mc + k];
        sb[k] = side[2 * nc + k];
    }
    // phase centres of the four pairs and the middle of their range (each
    // then lies within half the range, at most the centroid distance of one
    // RWG's two triangles, of the common centre)
    double R0p[2][2], lo = 1e300, hi = -1e300;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            R0p[a][b] = 0.0;
            if (tm[a] >= 0 && tn[b] >= 0) {
                const TriGeo g1 = tg[tm[a]], g2 = tg[tn[b]];
                const double dx = g1.cx - g2.cx, dy = g1.cy - g2.cy, dz = g1.cz - g2.cz;
                R0p[a][b] = sqrt(dx * dx + dy * dy + dz * dz);
                lo = fmin(lo, R0p[a][b]);
                hi = fmax(hi, R0p[a][b]);
            }
        }
    const double R0e = hi >= lo ? 0.5 * (lo + hi) : 0.0;

    // distinct triangles of the tile's row and column RWGs (wave 0: lanes
    // 0-31 the rows' sides, 32-63 the columns'; first occurrences kept)
    if (threadIdx.x < 64) {
        const int l = threadIdx.x, k = l & 31;
        const int e = (l < 32 ? 16 * I : 16 * J) + (k >> 1);
        const int t = e < Ne ? sideTri[2 * e + (k & 1)] : -1;
        bool first = t >= 0;
        for (int q = 0; q < 32; ++q) {
            const int u = __shfl(t, (l & 32) + q);
            if (q < k && u == t) first = false;
        }
        const unsigned long long mk = __ballot(first);
        const unsigned long long half = l < 32 ? (mk & 0xffffffffull) : (mk >> 32);
        const int rank = __popcll(half & ((1ull << k) - 1ull));
        if (first) (l < 32 ? rl : cl)[rank] = t;
        if (k == 0) cnt[l >> 5] = __popcll(half);
    }
    __syncthreads();
    const int nR = cnt[0], nC = cnt[1], nP = nR * nC;
    const bool live = m < Ne && n < Ne;

    double D[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) D[j] = 0.0;
    if (nP <= kCoefSlots) {
        for (int s = threadIdx.x; s < nP; s += 256) {
            const int t1 = rl[s / nC], t2 = cl[s % nC];
            slotRec[s] = (t1 <= t2 ? rec_index(t1, t2) : rec_index(t2, t1)) * stride;
        }
        // this entry's four pairs: slot, side coefficients, shift
        int slot[4];
        double k0[4], k1x[4], k1y[4], k1z[4], k4x[4], k4y[4], k4z[4], cab[4], dl[4];
#pragma unroll
        for (int ab = 0; ab < 4; ++ab) {
            const int a = ab >> 1, b = ab & 1;
            slot[ab] = -1;
            k0[ab] = k1x[ab] = k1y[ab] = k1z[ab] = k4x[ab] = k4y[ab] = k4z[ab] = cab[ab] = dl[ab] = 0.0;
            if (!live || tm[a] < 0 || tn[b] < 0) continue;
            int ri = 0, ci = 0;
            for (int i = 0; i < nR; ++i) ri = rl[i] == tm[a] ? i : ri;
            for (int i = 0; i < nC; ++i) ci = cl[i] == tn[b] ? i : ci;
            slot[ab] = ri * nC + ci;
            const bool sw = tm[a] > tn[b];
            const Side A = sa[a], B = sb[b];
            const double c2 = A.c * B.c;
            // S2 - S1.un - S1'.um + S0 (um.un); S1 and S1' exchanged when
            // the pair is stored the other way round
            cab[ab] = c2;
            k0[ab] = c2 * (A.ux * B.ux + A.uy * B.uy + A.uz * B.uz);
            k1x[ab] = -c2 * (sw ? A.ux : B.ux);
            k1y[ab] = -c2 * (sw ? A.uy : B.uy);
            k1z[ab] = -c2 * (sw ? A.uz : B.uz);
            k4x[ab] = -c2 * (sw ? B.ux : A.ux);
            k4y[ab] = -c2 * (sw ? B.uy : A.uy);
            k4z[ab] = -c2 * (sw ? B.uz : A.uz);
            dl[ab] = R0p[a][b] - R0e;
        }
        __syncthreads();  // slotRec
        constexpr int NU = kCoefSlots * 8 / 256;
        double pre[NU];
        auto load = [&](int qq) {
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u, sl = e >> 3;
                pre[u] = sl < nP ? Q[slotRec[sl] + (size_t)qq * 8 + (e & 7)] : 0.0;
            }
        };
        load(0);
#pragma unroll
        for (int qq = 0; qq <= P; ++qq) {
            __syncthreads();  // the previous order is consumed
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u, sl = e >> 3;
                if (sl < nP) qs[sl][e & 7] = pre[u];
            }
            __syncthreads();
            if (qq < P) load(qq + 1);
#pragma unroll
            for (int ab = 0; ab < 4; ++ab) {
                if (slot[ab] < 0) continue;
                const double2* v = reinterpret_cast<const double2*>(qs[slot[ab]]);
                const double2 v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
                // D[j] (order z^(j-1)) = V[j-2] + S[j], shifted by e^{z d}
                const double a0 = 4.0 * cab[ab] * v0.x;
                const double V = k0[ab] * v0.x + k1x[ab] * v0.y + k1y[ab] * v1.x + k1z[ab] * v1.y +
                                 k4x[ab] * v2.x + k4y[ab] * v2.y + k4z[ab] * v3.x + cab[ab] * v3.y;
                double t = 1.0;
#pragma unroll
                for (int k = 0; qq + k < NT; ++k) {
                    D[qq + k] += a0 * t;
                    if (qq + 2 + k < NT) D[qq + 2 + k] += V * t;
                    t *= dl[ab] * kInvN[k + 1];
                }
            }
        }
    } else {
        // unrolled, so the side and pair arrays stay in registers
    #pragma unroll
        for (int ab = 0; ab < 4; ++ab) {
            const int a = ab >> 1, b = ab & 1;
            if (tm[a] < 0 || tn[b] < 0) continue;
            const bool sw = tm[a] > tn[b];
            const double* q = Q + (sw ? rec_index(tn[b], tm[a]) : rec_index(tm[a], tn[b])) * stride;
            const Side A = sa[a], B = sb[b];
            const double cab = A.c * B.c;
            // S2 - S1.un - S1'.um + S0 (um.un); S1 and S1' exchanged when the
            // pair is stored the other way round
            const double k0 = cab * (A.ux * B.ux + A.uy * B.uy + A.uz * B.uz);
            const double k1x = -cab * (sw ? A.ux : B.ux), k1y = -cab * (sw ? A.uy : B.uy), k1z = -cab * (sw ? A.uz : B.uz);
            const double k4x = -cab * (sw ? B.ux : A.ux), k4y = -cab * (sw ? B.uy : A.uy), k4z = -cab * (sw ? B.uz : A.uz);
            const double ks = 4.0 * cab;
            // D[j] (order z^(j-1)) = V[j-2] + S[j]
            double Dp[NT];
    #pragma unroll
            for (int j = 0; j < NT; ++j) Dp[j] = 0.0;
    #pragma unroll
            for (int qq = 0; qq <= P; ++qq) {
                const double2* v = reinterpret_cast<const double2*>(q + (size_t)qq * 8);
                const double2 v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
                const double V = k0 * v0.x + k1x * v0.y + k1y * v1.x + k1z * v1.y + k4x * v2.x + k4y * v2.y +
                                 k4z * v3.x + cab * v3.y;
                Dp[qq] += ks * v0.x;
                Dp[qq + 2] += V;
            }
            // shift to the common centre: times e^{z d}, d = R0 - R0e
            const double d = R0p[a][b] - R0e;
            double e[NT];
            e[0] = 1.0;
    #pragma unroll
            for (int j = 1; j < NT; ++j) e[j] = e[j - 1] * d * kInvN[j];
    #pragma unroll
            for (int j = 0; j < NT; ++j) {
                double s = 0.0;
    #pragma unroll
                for (int i = 0; i <= j; ++i) s += Dp[i] * e[j - i];
                D[j] += s;
            }
        }
    }

    double* out = C + (size_t)id * (NT + 1) * 256 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < NT; ++j) out[(size_t)j * 256] = D[j];
    out[(size_t)NT * 256] = R0e;
}

// Z of a batch of nf frequencies from the per-entry coefficients C (above):
// one block per 16 x 16 upper-triangular tile, one entry per thread; for
// every frequency
//   Z(m,n) = -eta/(4 pi) e^{-jk R0e} z^{-1} sum_j z^j D[j]
//          = h e^{-jk R0e} sum_j z^j D[j],  h = -alpha/(4 pi k^2) (Freq.hr/hi),
// written with its mirror (through LDS) to matrix f of the batch (FP64, and
// FP32 when Z32 is set).  It runs beside the previous batch's factorisation,
// so it keeps to 128 registers per wave (the update GEMM's share: the GEMM
// then keeps three waves on the SIMDs it shares).
template <int NT, bool LOSSY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_series_fill(const double* __restrict__ C, int Ne,
                                                     const Freq* __restrict__ fr, int nf, double* __restrict__ Z,
                                                     size_t ld, size_t zStride, float* __restrict__ Z32,
                                                     size_t ld32, size_t z32Stride, int nblk, int tri) {
    __shared__ double2 tile[16][17];
    // tiles in a grid-stride loop: the grid may be smaller than the tile
    // count (a fill that runs beside a factorisation leaves it the CUs)
    for (long long t = blockIdx.x; t < nblk; t += gridDim.x) {
        long long id = t;
        {
            const int q = nblk >> 3, rr = nblk & 7, x = (int)(id & 7), loc = (int)(id >> 3);
            id = x < rr ? (long long)x * (q + 1) + loc : (long long)rr * (q + 1) + (long long)(x - rr) * q + loc;
        }
        int J = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
        while ((long long)(J + 1) * (J + 2) / 2 <= id) ++J;
        while ((long long)J * (J + 1) / 2 > id) --J;
        const int I = (int)(id - (long long)J * (J + 1) / 2);
        const int r = threadIdx.x >> 4, cI = threadIdx.x & 15;
        const int m = 16 * I + r, n = 16 * J + cI;
        const double* in = C + (size_t)id * (NT + 1) * 256 + threadIdx.x;
        double D[NT];
    #pragma unroll
        for (int j = 0; j < NT; ++j) D[j] = in[(size_t)j * 256];
        const double R0e = in[(size_t)NT * 256];

        const bool upper = m < Ne && n < Ne && (I != J || r <= cI);
        const int mm = 16 * J + r, nn = 16 * I + cI;
        const bool lower = mm < Ne && nn < Ne && (I != J || r > cI);
        // tri: only what the symmetric solve reads -- FP64 above the
        // diagonal and the whole diagonal 64 x 64 blocks (the residual), FP32
        // below it and on the diagonal (the factorisation)
        const bool up32 = !tri || I == J, lo64 = !tri || (I >> 2) == (J >> 2);
        for (int f = 0; f < nf; ++f) {
            const Freq F = fr[f];
            // p = sum_j z^j D[j], z = -jk = ki - j kr
            double pr, pi;
            if (!LOSSY) {
                // z = -j kr: p = A - j kr B with A, B the even and odd parts,
                // polynomials in u = -kr^2
                const double u = -F.kr * F.kr;
                double A = D[NT - 1 - ((NT - 1) & 1)], B = D[NT - 1 - (NT & 1)];
    #pragma unroll
                for (int j = NT - 3 - ((NT - 1) & 1); j >= 0; j -= 2) A = A * u + D[j];
    #pragma unroll
                for (int j = NT - 3 - (NT & 1); j >= 1; j -= 2) B = B * u + D[j];
                pr = A;
                pi = -F.kr * B;
            } else {
                const double zr = F.ki, zi = -F.kr;
                pr = D[NT - 1];
                pi = 0.0;
    #pragma unroll
                for (int j = NT - 2; j >= 0; --j) {
                    const double tr = pr * zr - pi * zi + D[j];
                    pi = pr * zi + pi * zr;
                    pr = tr;
                }
            }
            // e^{-jk R0e} = e^{ki R0e} (cos - j sin)(kr R0e)
            double s, c;
            sincos_cw(F.kr * R0e, &s, &c);
            if (LOSSY) {
                const double amp = exp(F.ki * R0e);
                s *= amp;
                c *= amp;
            }
            const double wr = pr * c + pi * s, wi = pi * c - pr * s;
            // -eta/(4 pi) z^{-1} = -alpha/(4 pi k^2) = h
            const double2 val = make_double2(F.hr * wr - F.hi * wi, F.hr * wi + F.hi * wr);
            double* Zf = Z + zStride * f;
            float* Z32f = Z32 ? Z32 + z32Stride * f : nullptr;
            tile[r][cI] = val;
            if (upper) {
                *reinterpret_cast<double2*>(Zf + 2 * ((size_t)m * ld + n)) = val;
                if (Z32f && up32)
                    *reinterpret_cast<float2*>(Z32f + 2 * ((size_t)m * ld32 + n)) = make_float2((float)val.x, (float)val.y);
            }
            __syncthreads();
            if (lower) {
                const double2 v = tile[cI][r];
                if (lo64) *reinterpret_cast<double2*>(Zf + 2 * ((size_t)mm * ld + nn)) = v;
                if (Z32f)
                    *reinterpret_cast<float2*>(Z32f + 2 * ((size_t)mm * ld32 + nn)) = make_float2((float)v.x, (float)v.y);
            }
            __syncthreads();
        }
    }
}

// ----------------------------------------------- reference centroid fill --
// kernel_Z_centroid (patch_sweep_cuda.cu:94-153): one point per triangle.

template <bool ROWFAST>
__global__ __launch_bounds__(256) void k_centroid(const double* __restrict__ pos,
                                                  const double* __restrict__ f,
                                                  const double* __restrict__ dv,
                                                  const double* __restrict__ w,
                                                  const int* __restrict__ tsd,
                                                  const double* __restrict__ asoft, int Ne,
                                                  const Freq* __restrict__ fr,
                                                  double* __restrict__ Z, size_t ldr, size_t ldc) {
    int a = threadIdx.x & 15, b = threadIdx.x >> 4;
    int m = blockIdx.x * 16 + (ROWFAST ? a : b);
    int n = blockIdx.y * 16 + (ROWFAST ? b : a);
    if (m >= Ne || n >= Ne) return;
    const double kr = fr->kr, ki = fr->ki;
    double zr = 0.0, zi = 0.0;
    for (int sa = 0; sa < 2; ++sa) {
        const double wm = w[2 * m + sa];
        const int tm = tsd[2 * m + sa];
        if (wm == 0.0 || tm == 0) continue;
        for (int sb = 0; sb < 2; ++sb) {
            const double wn = w[2 * n + sb];
            const int tn = tsd[2 * n + sb];
            if (wn == 0.0 || tn == 0) continue;
            const double* pm = pos + 6 * m + 3 * sa;
            const double* pn = pos + 6 * n + 3 * sb;
            double dx = pm[0] - pn[0], dy = pm[1] - pn[1], dz = pm[2] - pn[2];
            double R = sqrt(dx * dx + dy * dy + dz * dz);
            if (tm == tn) {
                double as = asoft[tm - 1];
                R = sqrt(R * R + as * as);
            }
            if (R < 1e-15) R = 1e-15;
            double s, c;
            sincos(kr * R, &s, &c);
            double amp = exp(ki * R);
            double inv = 1.0 / (4.0 * kPi * R);
            double Gr = amp * c * inv, Gi = -amp * s * inv;
            const double* fm = f + 6 * m + 3 * sa;
            const double* fn = f + 6 * n + 3 * sb;
            double dotff = fm[0] * fn[0] + fm[1] * fn[1] + fm[2] * fn[2];
            double divprod = dv[2 * m + sa] * dv[2 * n + sb];
            double ww = wm * wn;
            double tvr = Gr * (dotff * ww), tvi = Gi * (dotff * ww);
            double tsr = Gr * (divprod * ww), tsi = Gi * (divprod * ww);
            double ur = fr->ar * tvr - fr->ai * tvi, ui = fr->ar * tvi + fr->ai * tvr;
            double vr = fr->br * tsr - fr->bi * tsi, vi = fr->br * tsi + fr->bi * tsr;
            zr += ur + vr;
            zi += ui + vi;
        }
    }
    double* o = Z + 2 * ((size_t)m * ldr + (size_t)n * ldc);
    *reinterpret_cast<double2*>(o) = make_double2(zr, zi);
}

// ------------------------------------------------- excitation / port ----

__device__ double block_sum_1024(double v, double* sh) {
    v = wave_sum(v);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
        sh[0] = t;
    }
    __syncthreads();
    t = sh[0];
    __syncthreads();
    return t;
}

// port_weights_gaussian.m:5-28: layer of every edge midpoint from the
// midpoint z range, Gaussian in the distance to edge0, normalised.
__global__ __launch_bounds__(1024) void k_port_weights(const double* __restrict__ mid, int Ne,
                                                       int edge0, double spread,
                                                       double* __restrict__ out,
                                                       double* __restrict__ sums, int which) {
    __shared__ double sh[32];
    __shared__ double zr[2];
    double zmin = INFINITY, zmax = -INFINITY;
    for (int e = threadIdx.x; e < Ne; e += blockDim.x) {
        double z = mid[3 * e + 2];
        zmin = fmin(zmin, z);
        zmax = fmax(zmax, z);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        zmin = fmin(zmin, __shfl_xor(zmin, o, 64));
        zmax = fmax(zmax, __shfl_xor(zmax, o, 64));
    }
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[w] = zmin; sh[16 + w] = zmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = INFINITY, b = -INFINITY;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) { a = fmin(a, sh[k]); b = fmax(b, sh[16 + k]); }
        zr[0] = a;
        zr[1] = b;
    }
    __syncthreads();
    const double zmid = 0.5 * (zr[0] + zr[1]);
    const double x0 = mid[3 * edge0], y0 = mid[3 * edge0 + 1], z0 = mid[3 * edge0 + 2];
    const int lay0 = z0 > zmid ? 1 : -1;
    double part = 0.0;
    for (int e = threadIdx.x; e < Ne; e += blockDim.x) {
        double dx = mid[3 * e] - x0, dy = mid[3 * e + 1] - y0, dz = mid[3 * e + 2] - z0;
        double t = sqrt(dx * dx + dy * dy + dz * dz) / spread;
        int lay = mid[3 * e + 2] > zmid ? 1 : -1;
        double wt = exp(-(t * t)) * (lay == lay0 ? 1.0 : 0.0);
        out[e] = wt;
        part += wt;
    }
    double tot = block_sum_1024(part, sh);
    for (int e = threadIdx.x; e < Ne; e += blockDim.x) out[e] = out[e] / tot;
    if (threadIdx.x == 0) sums[which] = tot;
}

__global__ void k_rhs(const double* __restrict__ wP, const double* __restrict__ wG,
                      const double* __restrict__ len, int Ne, double V0,
                      double* __restrict__ Vrhs, double* __restrict__ wdiff) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Ne) return;
    // V = (0 + (+V0/2) wP) + (-V0/2) wG   (assemble_ZV_EFIE_RWG_DIFF_CPU.m:43-51)
    double v = (0.0 + (V0 / 2) * wP[e]) + (-V0 / 2) * wG[e];
    Vrhs[2 * e] = v;
    Vrhs[2 * e + 1] = 0.0;
    wdiff[2 * e] = wP[e] * len[e];
    wdiff[2 * e + 1] = wG[e] * len[e];
}

// Zin epilogue (main_patch_efie_rwg.m:336-340, patch_sweep_fastquad_cuda.cu:
// 787-799).  One block per batch element; X holds the solution column.
__global__ __launch_bounds__(1024) void k_epilogue(const double* __restrict__ X, size_t strideX,
                                                   size_t xinc, const double* __restrict__ wl, int Ne,
                                                   int pe, int ge, int centroidMode, double V0,
                                                   double Z0, double* __restrict__ out) {
    __shared__ double sh[32];
    const double* x = X + (size_t)blockIdx.x * strideX;
    double ipr, ipi;
    if (centroidMode) {
        // Iport = x[p] - x[g]  (patch_sweep_cuda.cu:489-491)
        ipr = x[xinc * pe] - x[xinc * ge];
        ipi = x[xinc * pe + 1] - x[xinc * ge + 1];
    } else {
        double pr = 0, pim = 0, gr = 0, gim = 0;
        for (int e = threadIdx.x; e < Ne; e += blockDim.x) {
            double xr = x[xinc * e], xi = x[xinc * e + 1];
            pr += wl[2 * e] * xr; pim += wl[2 * e] * xi;
            gr += wl[2 * e + 1] * xr; gim += wl[2 * e + 1] * xi;
        }
        pr = block_sum_1024(pr, sh);
        pim = block_sum_1024(pim, sh);
        gr = block_sum_1024(gr, sh);
        gim = block_sum_1024(gim, sh);
        ipr = pr - gr;
        ipi = pim - gim;
    }
    if (threadIdx.x == 0) {
        double den = ipr * ipr + ipi * ipi;
        double zr = V0 * (ipr / den), zi = -V0 * (ipi / den);
        double a_r = zr - Z0, a_i = zi, b_r = zr + Z0, b_i = zi;
        double bden = b_r * b_r + b_i * b_i;
        double* o = out + 4 * blockIdx.x;
        o[0] = zr;
        o[1] = zi;
        o[2] = (a_r * b_r + a_i * b_i) / bden;
        o[3] = (a_i * b_r - a_r * b_i) / bden;
    }
}

// ------------------------------------------------------------ launchers --

hipError_t launch_quad_points(const double* tv, const TriGeo* tg, int Nt, const QuadRules& R, int depthN,
                              int depthS, double selfAlpha, Pt* regPts, Pt* nearPts, Pt* selfPts, double* aSelf,
                              double* rad, double* mx, hipStream_t s) {
    if (Nt < 1 || depthN < 0 || depthS < 0 || depthN > 4 || depthS > 4 || R.nReg < 1 || R.nReg > 16 ||
        R.nSub < 1 || R.nSub > 16)
        return hipErrorInvalidValue;
    const long long n = (long long)Nt * (1 + (1 << (2 * depthN)) + (1 << (2 * depthS)));
    hipError_t e = launch_fill_dbl(rad, 0.0, Nt, s);
    if (e == hipSuccess) e = launch_fill_dbl(mx, 0.0, 4, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_quad_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tv, tg, Nt, R, depthN,
                       depthS, selfAlpha, regPts, nearPts, selfPts, aSelf,
                       reinterpret_cast<unsigned long long*>(rad), reinterpret_cast<unsigned long long*>(mx));
    return hipGetLastError();
}

hipError_t launch_near_bins(const TriGeo* tg, int Nt, double eps, double x0, double y0, double z0, double h, int nx,
                            int ny, int nz, const int* cellOf, const int* cellStart, const int* cellTri,
                            int* rowCount, const int* rowStart, int2* list, hipStream_t s) {
    const CellGrid g{x0, y0, z0, h, nx, ny, nz};
    if (!list)
        hipLaunchKernelGGL(k_near_count, dim3((Nt + 255) / 256), dim3(256), 0, s, tg, Nt, eps, g, cellOf, cellStart,
                           cellTri, rowCount);
    else
        hipLaunchKernelGGL(k_near_rows, dim3((Nt + 3) / 4), dim3(256), 0, s, tg, Nt, eps, g, cellOf, cellStart,
                           cellTri, rowStart, list);
    return hipGetLastError();
}

hipError_t launch_near_list(const uint8_t* nearMat, const int* rowStart, int Nt, int2* list,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_near_list, dim3(Nt), dim3(256), 0, s, nearMat, rowStart, Nt, list);
    return hipGetLastError();
}

hipError_t launch_rec_regular(const Pt* regPts, int Nq, const TriGeo* tg, const double* rad, int Nt,
                              const uint8_t* nearMat, double eps, const Freq* fr, Rec* recs, hipStream_t s) {
    long long T = (Nt + kRegTile - 1) / kRegTile;
    long long nb = T * (T + 1) / 2;
    dim3 g((unsigned)nb, kRegSplit), blk(256);
    switch (Nq) {
        case 1: hipLaunchKernelGGL(k_rec_regular<1>, g, blk, 0, s, regPts, tg, rad, Nt, nearMat, eps, fr, recs); break;
        case 3: hipLaunchKernelGGL(k_rec_regular<3>, g, blk, 0, s, regPts, tg, rad, Nt, nearMat, eps, fr, recs); break;
        case 6: hipLaunchKernelGGL(k_rec_regular<6>, g, blk, 0, s, regPts, tg, rad, Nt, nearMat, eps, fr, recs); break;
        case 7: hipLaunchKernelGGL(k_rec_regular<7>, g, blk, 0, s, regPts, tg, rad, Nt, nearMat, eps, fr, recs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rec_near(const Pt* subPts, int Np, const TriGeo* tg, const double* rad, const double* aeq,
                           double near_alpha, const int2* list, int nPairs, const Freq* fr,
                           Rec* recs, hipStream_t s) {
    if (nPairs <= 0) return hipSuccess;
    // Nq2 only matters for the self kernel; pass Np so psub is unused
    hipLaunchKernelGGL(k_rec_near, dim3((nPairs + 3) / 4), dim3(256), 0, s, subPts, Np, Np, tg, rad, aeq,
                       near_alpha, list, nPairs, fr, recs);
    return hipGetLastError();
}

hipError_t launch_rec_self(const Pt* selfPts, int Nq2, int Ns, const double* aSelf, const double* rad,
                           double aSelfMax, int Nt, const Freq* fr, Rec* recs, hipStream_t s) {
    hipLaunchKernelGGL(k_rec_self, dim3((Nt + 3) / 4), dim3(256), 0, s, selfPts, Nq2, Ns, aSelf, rad, aSelfMax,
                       Nt, fr, recs);
    return hipGetLastError();
}

hipError_t launch_sur_labels(const double* tri18, const double* kt, const Freq* fr, const SurPair* pairs, int n,
                             const double* subBary, int Np, const double* regBary, int Nq, double near_alpha,
                             double* T, double* nrm, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const size_t lds = sizeof(Pt) * 4 * 2 * (size_t)(Np + Nq);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sur_labels, dim3((n + 3) / 4), dim3(256), lds, s, tri18, kt, fr, pairs, n, subBary, Np,
                       regBary, Nq, near_alpha, T, nrm);
    return hipGetLastError();
}

hipError_t launch_moments(const Pt* subPts, int Np, const TriGeo* tg, const double* aeq,
                          double near_alpha, const int2* list, int nPairs, const Pt* selfPts,
                          int Nq2, int Ns, const double* aSelf, int Nt, int P, double* Qn,
                          double* Qs, const double* triV, hipStream_t s) {
    if (!triV) {
        if (nPairs > 0)
            hipLaunchKernelGGL(k_moment_near<false>, dim3((nPairs + 3) / 4), dim3(256), 0, s, subPts, Np, tg, aeq,
                               near_alpha, list, nPairs, P, Qn);
        hipLaunchKernelGGL(k_moment_self<false>, dim3((Nt + 3) / 4), dim3(256), 0, s, selfPts, Nq2, Ns, aSelf, Nt,
                           P, Qs);
        return hipGetLastError();
    }
    // singularity extraction: smooth tail, then the static q = 0 slot
    if (nPairs > 0) {
        hipLaunchKernelGGL(k_moment_near<true>, dim3((nPairs + 3) / 4), dim3(256), 0, s, subPts, Np, tg, aeq, 0.0,
                           list, nPairs, P, Qn);
        hipLaunchKernelGGL(k_static_extract<false>, dim3((nPairs + 3) / 4), dim3(256), 0, s, subPts, Np, tg, triV,
                           list, nPairs, P, Qn);
    }
    hipLaunchKernelGGL(k_moment_self<true>, dim3((Nt + 3) / 4), dim3(256), 0, s, selfPts, Nq2, Ns, aSelf, Nt, P,
                       Qs);
    hipLaunchKernelGGL(k_static_extract<true>, dim3((Nt + 3) / 4), dim3(256), 0, s, selfPts, Ns * Nq2, tg, triV,
                       nullptr, Nt, P, Qs);
    return hipGetLastError();
}

hipError_t launch_poly_records(const double* Qn, const int2* list, int nNear, const double* Qs, int Nt,
                               int P, const Freq* fr, Rec* recs, hipStream_t s) {
    const int n = nNear + Nt;
    hipLaunchKernelGGL(k_poly_records, dim3((n + 255) / 256), dim3(256), 0, s, Qn, list, nNear, Qs, Nt, P, fr,
                       recs);
    return hipGetLastError();
}

hipError_t launch_series(const Pt* regPts, int Nq, const TriGeo* tg, int Nt, const uint8_t* nearMat, double eps,
                         const Pt* subPts, int Np, int Nq2, const double* aeq, double near_alpha, const int2* list,
                         int nPairs, const Pt* selfPts, int Ns, const double* aSelf, int P, double* Q,
                         hipStream_t s) {
    if (P < 0 || P + 1 > 32) return hipErrorInvalidValue;
    long long T = (Nt + kRegTile - 1) / kRegTile;
    long long nb = T * (T + 1) / 2;
    dim3 g((unsigned)nb, kRegSplit), blk(256);
    switch (Nq) {
        case 1: hipLaunchKernelGGL(k_series_regular<1>, g, blk, 0, s, regPts, tg, Nt, nearMat, eps, P, Q); break;
        case 3: hipLaunchKernelGGL(k_series_regular<3>, g, blk, 0, s, regPts, tg, Nt, nearMat, eps, P, Q); break;
        case 6: hipLaunchKernelGGL(k_series_regular<6>, g, blk, 0, s, regPts, tg, Nt, nearMat, eps, P, Q); break;
        case 7: hipLaunchKernelGGL(k_series_regular<7>, g, blk, 0, s, regPts, tg, Nt, nearMat, eps, P, Q); break;
        default: return hipErrorInvalidValue;
    }
    if (nPairs > 0)
        hipLaunchKernelGGL(k_series_near, dim3((nPairs + 3) / 4), dim3(256), 0, s, subPts, Np, Nq2, tg, aeq,
                           near_alpha, list, nPairs, P, Q);
    hipLaunchKernelGGL(k_series_self, dim3((Nt + 3) / 4), dim3(256), 0, s, selfPts, Nq2, Ns, aSelf, Nt, P, Q);
    return hipGetLastError();
}

int series_nt(int P) {
    for (int nt : {10, 12, 14, 16, 18, 20, 22, 24})
        if (P + 3 <= nt) return nt;
    return 0;
}

size_t series_coeff_bytes(int Ne, int NT) {
    const long long T = (Ne + 15) / 16;
    return sizeof(double) * (size_t)(T * (T + 1) / 2) * (size_t)(NT + 1) * 256;
}

hipError_t launch_series_coeffs(const double* Q, int NT, const TriGeo* tg, const int* sideTri, const Side* side,
                                int Ne, double* C, hipStream_t s) {
    const int T = (Ne + 15) / 16;
    const long long nb = (long long)T * (T + 1) / 2;
    const dim3 g((unsigned)nb), blk(256);
    switch (NT) {
        case 10: hipLaunchKernelGGL(k_series_coeffs<10>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 12: hipLaunchKernelGGL(k_series_coeffs<12>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 14: hipLaunchKernelGGL(k_series_coeffs<14>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 16: hipLaunchKernelGGL(k_series_coeffs<16>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 18: hipLaunchKernelGGL(k_series_coeffs<18>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 20: hipLaunchKernelGGL(k_series_coeffs<20>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 22: hipLaunchKernelGGL(k_series_coeffs<22>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        case 24: hipLaunchKernelGGL(k_series_coeffs<24>, g, blk, 0, s, Q, tg, sideTri, side, Ne, C); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_series_fill(const double* C, int NT, bool lossy, int Ne, const Freq* fr, int nf, double* Z,
                              size_t ld, size_t zStride, float* Z32, size_t ld32, size_t z32Stride, hipStream_t s,
                              int maxBlocks, bool tri) {
    if (tri && !Z32) return hipErrorInvalidValue;
    const int T = (Ne + 15) / 16;
    const long long nb = (long long)T * (T + 1) / 2;
    const dim3 g((unsigned)(maxBlocks > 0 && maxBlocks < nb ? maxBlocks : nb)), blk(256);
#define EFIE_SER_CASE(N)                                                                                         \
    case N:                                                                                                       \
        if (lossy)                                                                                                \
            hipLaunchKernelGGL((k_series_fill<N, true>), g, blk, 0, s, C, Ne, fr, nf, Z, ld, zStride, Z32, ld32,  \
                               z32Stride, (int)nb, (int)tri);                                                     \
        else                                                                                                      \
            hipLaunchKernelGGL((k_series_fill<N, false>), g, blk, 0, s, C, Ne, fr, nf, Z, ld, zStride, Z32, ld32, \
                               z32Stride, (int)nb, (int)tri);                                                     \
        break;
    switch (NT) {
        EFIE_SER_CASE(10)
        EFIE_SER_CASE(12)
        EFIE_SER_CASE(14)
        EFIE_SER_CASE(16)
        EFIE_SER_CASE(18)
        EFIE_SER_CASE(20)
        EFIE_SER_CASE(22)
        EFIE_SER_CASE(24)
        default: return hipErrorInvalidValue;
    }
#undef EFIE_SER_CASE
    return hipGetLastError();
}

hipError_t launch_combine(const Rec* recs, const int* sideTri, const Side* side, int Ne,
                          const Freq* fr, double* Z, size_t ld, float* Z32, size_t ld32, hipStream_t s) {
    const int T = (Ne + 15) / 16;
    const long long nb = (long long)T * (T + 1) / 2;
    hipLaunchKernelGGL(k_combine, dim3((unsigned)nb), dim3(256), 0, s, recs, sideTri, side, Ne, fr, Z, ld, Z32,
                       ld32, (int)nb);
    return hipGetLastError();
}

hipError_t launch_centroid(const double* pos, const double* f, const double* dv, const double* w,
                           const int* tsd, const double* asoft, int Ne, const Freq* fr, double* Z,
                           size_t ldr, size_t ldc, hipStream_t s) {
    dim3 g((Ne + 15) / 16, (Ne + 15) / 16), blk(256);
    if (ldr == 1)
        hipLaunchKernelGGL(k_centroid<true>, g, blk, 0, s, pos, f, dv, w, tsd, asoft, Ne, fr, Z, ldr, ldc);
    else
        hipLaunchKernelGGL(k_centroid<false>, g, blk, 0, s, pos, f, dv, w, tsd, asoft, Ne, fr, Z, ldr, ldc);
    return hipGetLastError();
}

hipError_t launch_port_weights(const double* mid, int Ne, int edge0, double spread, double* out,
                               double* sums, int which, hipStream_t s) {
    hipLaunchKernelGGL(k_port_weights, dim3(1), dim3(1024), 0, s, mid, Ne, edge0, spread, out, sums,
                       which);
    return hipGetLastError();
}

hipError_t launch_rhs(const double* wP, const double* wG, const double* len, int Ne, double V0,
                      double* Vrhs, double* wdiff, hipStream_t s) {
    hipLaunchKernelGGL(k_rhs, dim3((Ne + 255) / 256), dim3(256), 0, s, wP, wG, len, Ne, V0, Vrhs, wdiff);
    return hipGetLastError();
}

hipError_t launch_epilogue(const double* X, size_t strideX, size_t xinc, const double* wdiff,
                           int Ne, int patchEdge0, int groundEdge0, int centroidMode, double V0,
                           double Z0, int batch, double* ZinS11, hipStream_t s) {
    hipLaunchKernelGGL(k_epilogue, dim3(batch), dim3(1024), 0, s, X, strideX, xinc, wdiff, Ne, patchEdge0,
                       groundEdge0, centroidMode, V0, Z0, ZinS11);
    return hipGetLastError();
}

}  // namespace efie
