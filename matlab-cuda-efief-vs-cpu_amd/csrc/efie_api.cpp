// efie_api.cpp -- C ABI of libefie_hip.so: argument checking, host-side
// table builders, device context and the frequency-sweep driver.
#include "../../include/efie.h"
#include "../../include/efie_diag.h"
#include "efie_internal.h"
#include "efie_sur.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace efie;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

thread_local std::string g_err = "";
thread_local std::string g_err_id = "";

efie_status fail(efie_status st, const char* id, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
efie_status fail(efie_status st, const char* id, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    g_err_id = id;
    return st;
}

efie_status ok() {
    g_err.clear();
    g_err_id.clear();
    return EFIE_OK;
}

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(EFIE_E_HIP, "efie:hip", "%s: %s", #call, hipGetErrorString(e_));     \
    } while (0)

// Device buffer that frees itself and only grows.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// ------------------------------------------------------ quadrature rules --
// tri_quad_rule.m:3-22 (order 1, 2, otherwise the 6-point rule with the
// reference's 15-digit literals) plus Radon's 7-point degree-5 rule for 7.
int quad_rule(int order, double* u, double* v, double* w) {
    if (order == 1) {
        u[0] = 1.0 / 3; v[0] = 1.0 / 3; w[0] = 1.0 / 2;
        return 1;
    }
    if (order == 2) {
        const double uu[3] = {1.0 / 6, 2.0 / 3, 1.0 / 6}, vv[3] = {1.0 / 6, 1.0 / 6, 2.0 / 3};
        for (int i = 0; i < 3; ++i) { u[i] = uu[i]; v[i] = vv[i]; w[i] = (1.0 / 6) * 1.0; }
        return 3;
    }
    if (order == 7) {
        const double s15 = std::sqrt(15.0);
        const double a1 = (6.0 - s15) / 21.0, a2 = (6.0 + s15) / 21.0;
        const double w0 = 9.0 / 80.0, w1 = (155.0 - s15) / 2400.0, w2 = (155.0 + s15) / 2400.0;
        const double uu[7] = {1.0 / 3, a1, 1 - 2 * a1, a1, a2, 1 - 2 * a2, a2};
        const double vv[7] = {1.0 / 3, a1, a1, 1 - 2 * a1, a2, a2, 1 - 2 * a2};
        const double ww[7] = {w0, w1, w1, w1, w2, w2, w2};
        for (int i = 0; i < 7; ++i) { u[i] = uu[i]; v[i] = vv[i]; w[i] = ww[i]; }
        return 7;
    }
    const double a = 0.445948490915965, b = 0.108103018168070;
    const double w1 = 0.111690794839005, w2 = 0.054975871827661;
    const double uu[6] = {a, a, 1 - 2 * a, b, b, 1 - 2 * b};
    const double vv[6] = {a, 1 - 2 * a, a, b, 1 - 2 * b, b};
    const double ww[6] = {w1, w1, w1, w2, w2, w2};
    for (int i = 0; i < 6; ++i) { u[i] = uu[i]; v[i] = vv[i]; w[i] = ww[i]; }
    return 6;
}

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 mid(V3 a, V3 b) { return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)}; }

// A + u(B-A) + v(C-A)  (map_ref_to_vertices.m)
inline V3 map_ref(V3 A, V3 B, V3 C, double u, double v) {
    return {A.x + u * (B.x - A.x) + v * (C.x - A.x), A.y + u * (B.y - A.y) + v * (C.y - A.y),
            A.z + u * (B.z - A.z) + v * (C.z - A.z)};
}

// 4^depth midpoint children in the order of subdivide_triangles.m:13-16.
void subdivide(V3 A, V3 B, V3 C, int depth, std::vector<V3>& out) {
    out.assign({A, B, C});
    for (int d = 0; d < depth; ++d) {
        std::vector<V3> nw;
        nw.reserve(out.size() * 4);
        for (size_t s = 0; s < out.size(); s += 3) {
            V3 a = out[s], b = out[s + 1], c = out[s + 2];
            V3 ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const V3 kids[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
            nw.insert(nw.end(), kids, kids + 12);
        }
        out.swap(nw);
    }
}

struct GeoKey {
    int quad_order = -1, near_order = -1, near_subdiv = -1, self_subdiv = -1;
    double near_sing_eps = -1, self_alpha = -1;
    int pack = 0;  // install count of the prepack in use, 0: rule-derived tables
    bool operator==(const GeoKey& o) const {
        return quad_order == o.quad_order && near_order == o.near_order &&
               near_subdiv == o.near_subdiv && self_subdiv == o.self_subdiv &&
               near_sing_eps == o.near_sing_eps && self_alpha == o.self_alpha && pack == o.pack;
    }
};

}  // namespace

namespace efie {

efie_status sur_fail(efie_status st, const char* id, const char* msg) { return fail(st, id, "%s", msg); }

void sur_rule_tables(int quad_order, int near_order, int near_subdiv, std::vector<double>& sub,
                     std::vector<double>& reg) {
    // rules mapped onto the unit vectors: the mapped points are barycentric
    double u[16], v[16], w[16];
    const V3 A{1, 0, 0}, B{0, 1, 0}, C{0, 0, 1};
    reg.clear();
    sub.clear();
    const int nq = quad_rule(quad_order, u, v, w);
    for (int i = 0; i < nq; ++i) {
        const V3 b = map_ref(A, B, C, u[i], v[i]);
        reg.insert(reg.end(), {b.x, b.y, b.z, w[i]});
    }
    const int n2 = quad_rule(near_order, u, v, w);
    std::vector<V3> kids;
    subdivide(A, B, C, near_subdiv, kids);
    const double frac = 1.0 / (double)(1 << (2 * near_subdiv));
    for (size_t p = 0; p < kids.size(); p += 3)
        for (int i = 0; i < n2; ++i) {
            const V3 b = map_ref(kids[p], kids[p + 1], kids[p + 2], u[i], v[i]);
            sub.insert(sub.end(), {b.x, b.y, b.z, w[i] * frac});
        }
}

}  // namespace efie

namespace {
struct RawE {
    int a, b, raw;
};
std::vector<RawE> sorted_raw(const int* F, int Nt) {
    std::vector<RawE> r((size_t)Nt * 3);
    for (int t = 0; t < Nt; ++t) {
        const int tri[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
        for (int le = 0; le < 3; ++le) {
            int a = tri[le], b = tri[(le + 1) % 3];
            r[3 * t + le] = {std::min(a, b), std::max(a, b), 3 * t + le};
        }
    }
    std::stable_sort(r.begin(), r.end(), [](const RawE& x, const RawE& y) {
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    });
    return r;
}
}  // namespace

struct efie_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream: the next batch of a sweep is filled on it while the
    // current batch is factored on `stream`
    hipStream_t fillStream = nullptr;
    hipEvent_t evFill[2] = {}, evFillStart[2] = {};
    // third stream: LU look-ahead (the next block's panels)
    hipStream_t panelStream = nullptr;
    hipEvent_t evPanel = nullptr, evAhead = nullptr;
    // a pipelined factorisation's tail mark, per buffer (LuWork::tailMark)
    hipEvent_t evTail[2] = {nullptr, nullptr};
    // fourth stream: refinement / back substitution of a batch while the
    // next batch is factored
    hipStream_t irStream = nullptr;
    hipEvent_t evLU[2] = {}, evLUStart[2] = {}, evIR[2] = {}, evIRStart[2] = {};
    // pinned host slots for the per-batch results (info, iterations,
    // residuals, wait-limit flag), one per buffer
    void* hStat = nullptr;
    size_t hStatBytes = 0;
    hipEvent_t ev[8] = {};

    // host copies of the tables (column-major, 1-based as given)
    int Nv = 0, Nt = 0, Ne = 0;
    std::vector<double> V;
    std::vector<int> F;
    std::vector<int> edge, plusTri, minusTri, plusSign, minusSign;
    std::vector<double> rp, rm, Ap, Am, len;
    bool haveMesh = false, haveRwg = false, midValid = false;
    // efie_set_rwg_tables: RWG tables without V, F and edges
    bool meshless = false;

    // prepacked quadrature caches (efie_set_prepack) in the internal
    // layouts, positions absolute; packGen counts installs
    bool havePack = false;
    int packGen = 0, pkNq = 0, pkNq2 = 0, pkNsS = 0, pkNsN = 0;
    double pkNearAlpha = 0;
    std::vector<Pt> pkReg, pkNear, pkSelf;
    std::vector<double> pkASelf, pkAeq;
    std::vector<uint8_t> pkNearMat;  // [t1][t2], diagonal 0

    // geometry cache (quadrature tables etc.)
    bool geoValid = false;
    GeoKey key;
    int Nq = 0, NpNear = 0, Nq2s = 0, NsSelf = 0;
    long long nNear = 0;
    double rNearMax = 0, rSelfMax = 0, lMax = 0, aNearMax = 0, aSelfMax = 0;
    double dNearMax = 0;  // bound on the centroid distance of a near pair
    // moment cache (near_mode MOMENTS or EXTRACT), valid for series length
    // momP + 1 of the scheme momExtract says
    int momP = -1;
    double nearAlphaMom = -1;
    bool momExtract = false;
    bool momSelfOnly = false;  // only the self-pair moments were computed
    DBuf dTriV;  // triangle vertices relative to the centroid [Nt][9]
    DBuf dTriAbs;  // triangle vertices [Nt][9], then 4 radius maxima of the point generator
    // RWG tables built on the device: mesh copies, count-pass work, tables,
    // port-pick partials; rwgDevNe >= 0 once counted for the current mesh,
    // portValid once the tables (centres, layer signs) are built
    DBuf dMeshV, dMeshF, dRwgWork, dRwgTab, dPortPart;
    int rwgDevNe = -1;
    bool portValid = false;
    // frequency-independent series of every triangle pair (near_mode
    // MOMENTS), valid for P = serP and near_alpha = serAlpha
    DBuf dQall, dSerC;  // pair series (transient) and per-entry coefficients
    int serP = -1;
    double serAlpha = -1;
    double radMax = 0;  // largest point radius of a triangle about its centroid
    double dcMax = 0;   // largest centroid distance of an RWG's two triangles

    // near-pair surrogate: the installed model and, per mesh, the pairs it
    // takes (dSurPairs, nSurIn, with their normalised vertices dSurTri18)
    // and the pairs left to the quadrature (dSurFb, nSurFb)
    std::vector<float> surBlob;
    bool surExact = false, surValid = false;
    double surAlpha = -1, surLmax = 0, surDmax = 0;
    int nSurIn = 0, nSurFb = 0, surNp = 0, surNq = 0;
    DBuf dSurModel, dSurPairs, dSurFb, dSurTri18, dSurT, dSurSub, dSurReg;
    DBuf dMlpA, dMlpB, dMlpC, dMlpD, dMlpE;  // scratch of the network entry points

    DBuf dTri, dReg, dNearPts, dSelfPts, dASelf, dAeq, dNearMat, dRowCnt, dList;
    DBuf dCellOf, dCellStart, dCellTri;  // centroid bins of the near search
    double nearEps = 0;  // near test of the regular kernels (0: the map dNearMat)
    DBuf dSideTri, dSide, dMid, dLen, dQn, dQs, dRad;
    // centroid model tables
    DBuf dCPos, dCF, dCDv, dCW, dCT, dCAsoft;
    double cAlpha = -1;
    // per-call buffers
    DBuf dRecs, dFreq, dZ, dZ32, dWp, dWg, dSums, dV, dWl, dOut;
    DBuf dCand, dCnt, dIpiv, dInfo, dSwp, dPerm, dX, dR, dR32, dRn2, dBn2, dDone, dIters, dResid, dB, dP;
    DBuf dFlags, dTickets, dErr;
    // efie_sweep_many: per buffer the batch's right-hand sides padded to its
    // order and the order of each matrix's problem; per item its right-hand
    // side and port weights
    DBuf dVp, dNElem, dItemV, dItemWl;
    // and the scratch of its series builds: the pair series (one build at a
    // time) and the coefficient slots of released problems
    DBuf dSerQ;
    std::vector<std::unique_ptr<DBuf>> dSerPool;
    DBuf dSplit;  // bf16 planes of the split-bf16 LU update (lu_kernel BF16X6)
    TrsvWork tw{nullptr, nullptr, nullptr, 0};
    // this GPU's copy of the allocation-free solve's scratch array
    void* scratch = nullptr;

    // profiling (opts.profile): event pairs per kernel class
    bool prof = false;
    std::vector<hipEvent_t> pool;
    size_t poolUsed = 0;
    struct Mark {
        int cls, launches;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    int openCls = -1;
    hipEvent_t openEv = nullptr;
};

namespace {
// Back to the rule-derived quadrature tables.  The converted pack stays
// stored, so that installing the same pack again (efie_set_prepack) finds
// the geometry built from it still valid; the geometry key (pack install
// number or rule options) decides which tables a cached geometry belongs to.
void drop_pack(efie_ctx* c) { c->havePack = false; }

template <typename T>
bool same_table(const std::vector<T>& have, const T* in, size_t n) {
    return have.size() == n && (n == 0 || std::memcmp(have.data(), in, sizeof(T) * n) == 0);
}

// process-wide counts of geometry and moment builds (efie_diag_cache_counts)
std::atomic<long long> g_geoBuilds{0}, g_momBuilds{0};
}  // namespace

extern "C" {

const char* efie_last_error(void) { return g_err.c_str(); }
const char* efie_last_error_id(void) { return g_err_id.c_str(); }

void efie_opts_default(efie_opts* o) {
    // accurate preset (main_patch_efie_rwg.m:281-289)
    o->quad_order = 4;
    o->near_sing_eps = 2.5;
    o->self_subdiv = 2;
    o->self_alpha = 0.20;
    o->near_subdiv = 2;
    o->near_alpha = 0.15;
    o->near_order = 4;
    o->model = EFIE_MODEL_QUADRATURE;
    o->precision = EFIE_PREC_FP64;
    o->batch = 0;
    o->ir_tol = 1e-14;
    o->ir_max_iter = 16;
    o->near_mode = EFIE_NEAR_DIRECT;
    o->profile = 0;
    o->lu_kernel = EFIE_LU_AUTO;
}

efie_status efie_opts_preset(efie_opts* o, const char* name) {
    if (!o || !name) return fail(EFIE_E_ARGS, "efie:opts", "null argument");
    efie_opts_default(o);
    if (!std::strcmp(name, "accurate")) return ok();
    if (!std::strcmp(name, "fast")) {
        // main_patch_gpu_mex_sweep_l.m:50-56
        o->quad_order = 2;
        o->near_sing_eps = 1.2;
        o->self_subdiv = 2;
        o->self_alpha = 0.2;
        o->near_subdiv = 1;
        o->near_alpha = 0.15;
        o->near_order = 2;
        return ok();
    }
    if (!std::strcmp(name, "centroid")) {
        o->model = EFIE_MODEL_CENTROID;
        return ok();
    }
    return fail(EFIE_E_ARGS, "efie:opts", "unknown preset '%s'", name);
}

// ------------------------------------------------------------- mesh ------

efie_status efie_mesh_counts(int npx, int npy, int ngx, int ngy, int* Nv, int* Nt) {
    if (npx < 1 || npy < 1 || ngx < 1 || ngy < 1 || !Nv || !Nt)
        return fail(EFIE_E_ARGS, "efie:mesh", "cell counts must be >= 1");
    *Nv = (npx + 1) * (npy + 1) + (ngx + 1) * (ngy + 1);
    *Nt = 2 * (npx * npy + ngx * ngy);
    return ok();
}

// make_patch_with_ground_mesh.m:1-46.  Vertices in meshgrid(x,y)(:) order
// (y fastest), faces with the (N_x+1) stride exactly as the reference
// writes them; for N_x != N_y that stride does not match the vertex order,
// which the reference does as well.
efie_status efie_make_mesh(double Lp, double Wp, double Lg, double Wg, double h, int npx, int npy,
                           int ngx, int ngy, double* V, int* F) {
    int Nv, Nt;
    efie_status st = efie_mesh_counts(npx, npy, ngx, ngy, &Nv, &Nt);
    if (st) return st;
    if (!V || !F) return fail(EFIE_E_ARGS, "efie:mesh", "null output");
    auto lin = [](double d1, double d2, int n) {
        std::vector<double> y(n);
        const int n1 = n - 1;
        for (int i = 0; i <= n1; ++i) y[i] = d1 + ((double)i * (d2 - d1)) / (double)n1;
        y[n1] = d2;
        return y;
    };
    int v = 0, t = 0;
    auto layer = [&](double W, double L, int nx, int ny, double z) {
        const int off = v;
        std::vector<double> xs = lin(-W / 2, W / 2, nx + 1), ys = lin(-L / 2, L / 2, ny + 1);
        for (int c = 0; c <= nx; ++c)
            for (int r = 0; r <= ny; ++r, ++v) {
                V[v] = xs[c];
                V[v + Nv] = ys[r];
                V[v + 2 * Nv] = z * 1.0;
            }
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int v1 = off + (j - 1) * (nx + 1) + i, v2 = v1 + 1, v3 = v1 + (nx + 1), v4 = v3 + 1;
                F[t] = v1; F[t + Nt] = v2; F[t + 2 * Nt] = v4; ++t;
                F[t] = v1; F[t + Nt] = v4; F[t + 2 * Nt] = v3; ++t;
            }
    };
    layer(Wp, Lp, npx, npy, +h / 2);
    layer(Wg, Lg, ngx, ngy, -h / 2);
    return ok();
}

// --------------------------------------------------------------- RWG -----
// build_rwg_interior.m: unique undirected edges in lexicographic order of
// the sorted vertex pair (:25); an edge is interior when exactly two
// triangle edges map to it (:29-31); its first incidence in raw order is
// the plus side (:47-63).


efie_status efie_rwg_count(const int* F, int Nt, int Nv, int* Ne) {
    if (!F || !Ne || Nt < 1 || Nv < 3) return fail(EFIE_E_ARGS, "efie:rwg", "bad mesh sizes");
    for (long long i = 0; i < 3LL * Nt; ++i)
        if (F[i] < 1 || F[i] > Nv) return fail(EFIE_E_MESH, "mex:F", "F has invalid indices.");
    std::vector<RawE> r = sorted_raw(F, Nt);
    int ne = 0;
    for (size_t i = 0; i < r.size();) {
        size_t j = i;
        while (j < r.size() && r[j].a == r[i].a && r[j].b == r[i].b) ++j;
        if (j - i == 2) ++ne;
        i = j;
    }
    *Ne = ne;
    return ok();
}

efie_status efie_build_rwg(const double* V, int Nv, const int* F, int Nt, int Ne, int* edge,
                           int* plusTri, int* minusTri, int* plusSign, int* minusSign, double* rp,
                           double* rm, double* Ap, double* Am, double* len, double* center,
                           int* zSign) {
    int ne = 0;
    efie_status st = efie_rwg_count(F, Nt, Nv, &ne);
    if (st) return st;
    if (ne != Ne) return fail(EFIE_E_ARGS, "efie:rwg", "Ne=%d but the mesh has %d interior edges", Ne, ne);
    std::vector<RawE> r = sorted_raw(F, Nt);
    auto vx = [&](int vi) { return V3{V[vi - 1], V[vi - 1 + Nv], V[vi - 1 + 2 * Nv]}; };
    int e = 0;
    for (size_t i = 0; i < r.size();) {
        size_t j = i;
        while (j < r.size() && r[j].a == r[i].a && r[j].b == r[i].b) ++j;
        if (j - i == 2) {
            const int a = r[i].a, b = r[i].b;
            edge[e] = a;
            edge[e + Ne] = b;
            for (int s = 0; s < 2; ++s) {
                const int raw = r[i + s].raw, t = raw / 3, le = raw % 3;
                const int tri[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
                const int sg = (tri[le] == a && tri[(le + 1) % 3] == b) ? 1 : -1;
                (s == 0 ? plusTri : minusTri)[e] = t + 1;
                (s == 0 ? plusSign : minusSign)[e] = sg;
            }
            ++e;
        }
        i = j;
    }
    double zmax = -INFINITY, zmin = INFINITY;
    for (int v = 0; v < Nv; ++v) {
        zmax = std::max(zmax, V[v + 2 * Nv]);
        zmin = std::min(zmin, V[v + 2 * Nv]);
    }
    const double zmid = 0.5 * (zmax + zmin);
    for (e = 0; e < Ne; ++e) {
        const int va = edge[e], vb = edge[e + Ne];
        const V3 ra = vx(va), rb = vx(vb);
        len[e] = norm(rb - ra);
        const V3 c = mid(ra, rb);
        center[e] = c.x; center[e + Ne] = c.y; center[e + 2 * Ne] = c.z;
        for (int s = 0; s < 2; ++s) {
            const int t = (s == 0 ? plusTri : minusTri)[e] - 1;
            const int ids[3] = {F[t], F[t + Nt], F[t + 2 * Nt]};
            int opp = 0, cnt = 0;
            for (int k = 0; k < 3; ++k)
                if (ids[k] != va && ids[k] != vb) { opp = ids[k]; ++cnt; }
            if (cnt != 1) return fail(EFIE_E_MESH, "efie:rwg", "opposite vertex not found for edge %d", e + 1);
            const V3 o = vx(opp);
            double* rr = s == 0 ? rp : rm;
            rr[e] = o.x; rr[e + Ne] = o.y; rr[e + 2 * Ne] = o.z;
            (s == 0 ? Ap : Am)[e] = 0.5 * norm(cross(vx(ids[1]) - vx(ids[0]), vx(ids[2]) - vx(ids[0])));
        }
        zSign[e] = c.z > zmid ? 1 : -1;
    }
    return ok();
}

efie_status efie_pick_port(int Ne, const double* center, const int* zSign, const double p[3],
                           int layerSign, int* edgeOut) {
    if (!center || !zSign || !p || !edgeOut || Ne < 1) return fail(EFIE_E_ARGS, "efie:port", "bad arguments");
    int best = -1;
    double bd = INFINITY;
    for (int e = 0; e < Ne; ++e) {
        if (zSign[e] != layerSign) continue;
        const double d = norm(V3{center[e] - p[0], center[e + Ne] - p[1], center[e + 2 * Ne] - p[2]});
        if (best < 0 || d < bd) { bd = d; best = e + 1; }
    }
    if (best < 0) return fail(EFIE_E_PORT, "efie:port", "no edge on layer %d", layerSign);
    *edgeOut = best;
    return ok();
}

// ------------------------------------------------------------ context ----

efie_status efie_create(efie_ctx** out, int device) {
    if (!out) return fail(EFIE_E_ARGS, "efie:args", "null context pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(EFIE_E_ARGS, "efie:args", "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    efie_ctx* c = new efie_ctx();
    c->device = device;
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(EFIE_E_HIP, "efie:hip", "hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (hipStream_t* st : {&c->fillStream, &c->panelStream, &c->irStream}) {
        e = hipStreamCreate(st);
        if (e != hipSuccess) {
            efie_destroy(c);
            return fail(EFIE_E_HIP, "efie:hip", "hipStreamCreate: %s", hipGetErrorString(e));
        }
    }
    for (hipEvent_t* ev : {&c->ev[0], &c->ev[1], &c->ev[2], &c->ev[3], &c->ev[4], &c->ev[5], &c->ev[6], &c->ev[7],
                           &c->evFill[0], &c->evFill[1], &c->evFillStart[0], &c->evFillStart[1], &c->evPanel,
                           &c->evAhead, &c->evLU[0], &c->evLU[1], &c->evLUStart[0], &c->evLUStart[1], &c->evIR[0],
                           &c->evIR[1], &c->evIRStart[0], &c->evIRStart[1]}) {
        e = hipEventCreate(ev);
        if (e != hipSuccess) {
            efie_destroy(c);
            return fail(EFIE_E_HIP, "efie:hip", "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    for (hipEvent_t* ev : {&c->evTail[0], &c->evTail[1]}) {
        e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            efie_destroy(c);
            return fail(EFIE_E_HIP, "efie:hip", "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    e = noalloc_scratch(device, &c->scratch);
    if (e != hipSuccess) {
        efie_destroy(c);
        return fail(EFIE_E_HIP, "efie:hip", "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    *out = c;
    return ok();
}

void efie_destroy(efie_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fillStream) (void)hipStreamSynchronize(c->fillStream);
    if (c->panelStream) (void)hipStreamSynchronize(c->panelStream);
    if (c->irStream) (void)hipStreamSynchronize(c->irStream);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : {c->evFill[0], c->evFill[1], c->evFillStart[0], c->evFillStart[1], c->evPanel, c->evAhead,
                          c->evLU[0], c->evLU[1], c->evLUStart[0], c->evLUStart[1], c->evIR[0], c->evIR[1],
                          c->evIRStart[0], c->evIRStart[1], c->evTail[0], c->evTail[1]})
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->pool) (void)hipEventDestroy(ev);
    if (c->hStat) (void)hipHostFree(c->hStat);
    if (c->fillStream) (void)hipStreamDestroy(c->fillStream);
    if (c->panelStream) (void)hipStreamDestroy(c->panelStream);
    if (c->irStream) (void)hipStreamDestroy(c->irStream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;  // DBuf destructors free device memory
}

efie_status efie_set_mesh(efie_ctx* c, const double* V, int Nv, const int* F, int Nt) {
    if (!c || !V || !F) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (Nv < 3 || Nt < 1) return fail(EFIE_E_ARGS, "mex:V", "V must be [Nv x 3] double.");
    for (long long i = 0; i < 3LL * Nt; ++i)
        if (F[i] < 1 || F[i] > Nv) return fail(EFIE_E_MESH, "mex:F", "F has invalid indices.");
    // the same mesh again (a MEX call per frequency or per sweep): the
    // derived geometry, moments and centroid tables stay valid; the RWG
    // tables must still follow (and are compared in turn)
    if (c->haveMesh && !c->meshless && same_table(c->V, V, 3 * (size_t)Nv) && same_table(c->F, F, 3 * (size_t)Nt) &&
        Nv == c->Nv && Nt == c->Nt) {
        c->haveRwg = false;
        drop_pack(c);
        return ok();
    }
    c->Nv = Nv;
    c->Nt = Nt;
    c->V.assign(V, V + 3 * (size_t)Nv);
    c->F.assign(F, F + 3 * (size_t)Nt);
    c->rwgDevNe = -1;
    c->portValid = false;
    c->haveMesh = true;
    c->haveRwg = false;
    c->midValid = false;
    c->geoValid = false;
    c->cAlpha = -1;
    c->meshless = false;
    drop_pack(c);
    return ok();
}

efie_status efie_set_rwg(efie_ctx* c, const int* edge, const int* plusTri, const int* minusTri,
                         const int* plusSign, const int* minusSign, const double* rp,
                         const double* rm, const double* Ap, const double* Am, const double* len,
                         int Ne) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    if (!c->haveMesh) return fail(EFIE_E_STATE, "efie:state", "efie_set_mesh must come first");
    if (!edge || !plusTri || !minusTri || !plusSign || !minusSign || !rp || !rm || !Ap || !Am || !len)
        return fail(EFIE_E_ARGS, "efie:args", "null table");
    if (Ne < 1) return fail(EFIE_E_ARGS, "mex:edge", "edge must be [Ne x 2] int32.");
    for (int e = 0; e < Ne; ++e) {
        if (plusTri[e] < 0 || plusTri[e] > c->Nt)
            return fail(EFIE_E_MESH, "mex:plusTri", "plusTri must be int32 [Ne x 1].");
        if (minusTri[e] < 0 || minusTri[e] > c->Nt)
            return fail(EFIE_E_MESH, "mex:minusTri", "minusTri must be int32 [Ne x 1].");
        if (edge[e] < 1 || edge[e] > c->Nv || edge[e + Ne] < 1 || edge[e + Ne] > c->Nv)
            return fail(EFIE_E_MESH, "mex:edge", "edge has invalid vertex indices.");
    }
    if (!c->meshless && Ne == c->Ne && same_table(c->edge, edge, 2 * (size_t)Ne) &&
        same_table(c->plusTri, plusTri, Ne) && same_table(c->minusTri, minusTri, Ne) &&
        same_table(c->plusSign, plusSign, Ne) && same_table(c->minusSign, minusSign, Ne) &&
        same_table(c->rp, rp, 3 * (size_t)Ne) && same_table(c->rm, rm, 3 * (size_t)Ne) && same_table(c->Ap, Ap, Ne) &&
        same_table(c->Am, Am, Ne) && same_table(c->len, len, Ne)) {
        c->haveRwg = true;  // unchanged tables keep every cache
        drop_pack(c);
        return ok();
    }
    c->Ne = Ne;
    c->edge.assign(edge, edge + 2 * (size_t)Ne);
    c->plusTri.assign(plusTri, plusTri + Ne);
    c->minusTri.assign(minusTri, minusTri + Ne);
    c->plusSign.assign(plusSign, plusSign + Ne);
    c->minusSign.assign(minusSign, minusSign + Ne);
    c->rp.assign(rp, rp + 3 * (size_t)Ne);
    c->rm.assign(rm, rm + 3 * (size_t)Ne);
    c->Ap.assign(Ap, Ap + Ne);
    c->Am.assign(Am, Am + Ne);
    c->len.assign(len, len + Ne);
    c->haveRwg = true;
    c->midValid = false;
    c->geoValid = false;
    c->cAlpha = -1;
    drop_pack(c);
    return ok();
}

// ------------------------------------------- RWG tables on the device --
// build_rwg_interior.m:1-113 and pick_port_edge_by_point_and_layer.m:1-11
// from the context's mesh (efie_rwg.hip).

namespace {

// Work layout of the count pass in dRwgWork (ints, then the int2 slots).
struct RwgWork {
    int *cnt, *fill, *nInt, *start, *eStart;
    int2* slot;
};

RwgWork rwg_work(efie_ctx* c) {
    const size_t Nv = c->Nv;
    int* w = c->dRwgWork.as<int>();
    RwgWork r;
    r.cnt = w;
    r.fill = w + Nv;
    r.nInt = w + 2 * Nv;
    r.start = w + 3 * Nv;
    r.eStart = w + 4 * Nv + 1;
    const size_t ints = 5 * Nv + 2;
    r.slot = reinterpret_cast<int2*>(w + ((ints + 1) & ~(size_t)1));
    return r;
}

efie_status rwg_count_device(efie_ctx* c) {
    if (c->rwgDevNe >= 0) return ok();
    const size_t Nv = c->Nv, Nt = c->Nt;
    hipStream_t s = c->stream;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(c->dMeshV.ensure(sizeof(double) * 3 * Nv));
    HIPCHK(c->dMeshF.ensure(sizeof(int) * 3 * Nt));
    const size_t ints = ((5 * Nv + 2) + 1) & ~(size_t)1;
    HIPCHK(c->dRwgWork.ensure(sizeof(int) * ints + sizeof(int2) * 3 * Nt));
    HIPCHK(hipMemcpyAsync(c->dMeshV.p, c->V.data(), sizeof(double) * 3 * Nv, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dMeshF.p, c->F.data(), sizeof(int) * 3 * Nt, hipMemcpyHostToDevice, s));
    const RwgWork w = rwg_work(c);
    HIPCHK(launch_rwg_count(c->dMeshF.as<int>(), c->Nt, c->Nv, w.cnt, w.slot, w.fill, w.nInt, w.start, w.eStart, s));
    int ne = 0;
    HIPCHK(hipMemcpyAsync(&ne, w.eStart + Nv, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->rwgDevNe = ne;
    return ok();
}

}  // namespace

efie_status efie_rwg_count_device(efie_ctx* c, int* Ne) {
    if (!c || !Ne) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (!c->haveMesh || c->meshless) return fail(EFIE_E_STATE, "efie:state", "efie_set_mesh must come first");
    efie_status st = rwg_count_device(c);
    if (st) return st;
    *Ne = c->rwgDevNe;
    return ok();
}

efie_status efie_build_rwg_device(efie_ctx* c, int Ne, int* edge, int* plusTri, int* minusTri, int* plusSign,
                                  int* minusSign, double* rp, double* rm, double* Ap, double* Am, double* len,
                                  double* center, int* zSign) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    if (!edge || !plusTri || !minusTri || !plusSign || !minusSign || !rp || !rm || !Ap || !Am || !len || !center ||
        !zSign)
        return fail(EFIE_E_ARGS, "efie:args", "null output");
    if (!c->haveMesh || c->meshless) return fail(EFIE_E_STATE, "efie:state", "efie_set_mesh must come first");
    efie_status st = rwg_count_device(c);
    if (st) return st;
    if (Ne != c->rwgDevNe)
        return fail(EFIE_E_ARGS, "efie:rwg", "Ne=%d but the mesh has %d interior edges", Ne, c->rwgDevNe);
    if (Ne < 1) return fail(EFIE_E_MESH, "efie:rwg", "the mesh has no interior edge");
    const size_t n = Ne;
    hipStream_t s = c->stream;
    // tables: ints (edge, tri, sgn: 2 Ne each; zSign, err), then doubles
    // (rpm 6 Ne, Apm 2 Ne, len Ne, center 3 Ne)
    const size_t nInts = ((7 * n + 1) + 1) & ~(size_t)1;
    HIPCHK(c->dRwgTab.ensure(sizeof(int) * nInts + sizeof(double) * 12 * n));
    int* ti = c->dRwgTab.as<int>();
    int *dEdge = ti, *dTri = ti + 2 * n, *dSgn = ti + 4 * n, *dZ = ti + 6 * n, *dErr = ti + 7 * n;
    double* td = reinterpret_cast<double*>(ti + nInts);
    double *dRpm = td, *dApm = td + 6 * n, *dLen = td + 8 * n, *dCen = td + 9 * n;
    double zmax = -INFINITY, zmin = INFINITY;
    for (int v = 0; v < c->Nv; ++v) {
        zmax = std::max(zmax, c->V[v + 2 * (size_t)c->Nv]);
        zmin = std::min(zmin, c->V[v + 2 * (size_t)c->Nv]);
    }
    const RwgWork w = rwg_work(c);
    HIPCHK(launch_rwg_tables(c->dMeshV.as<double>(), c->Nv, c->dMeshF.as<int>(), c->Nt, w.start, w.slot, w.eStart, Ne,
                             0.5 * (zmax + zmin), dEdge, dTri, dSgn, dRpm, dApm, dLen, dCen, dZ, dErr, s));
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, dErr, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(edge, dEdge, sizeof(int) * 2 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plusTri, dTri, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(minusTri, dTri + n, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plusSign, dSgn, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(minusSign, dSgn + n, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rp, dRpm, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rm, dRpm + 3 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(Ap, dApm, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(Am, dApm + n, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(len, dLen, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(center, dCen, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(zSign, dZ, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err != 0x7fffffff) return fail(EFIE_E_MESH, "efie:rwg", "opposite vertex not found for edge %d", err);
    // the context's RWG tables from here on (unchanged tables keep every cache)
    st = efie_set_rwg(c, edge, plusTri, minusTri, plusSign, minusSign, rp, rm, Ap, Am, len, Ne);
    if (st) return st;
    c->portValid = true;
    return ok();
}

efie_status efie_pick_port_device(efie_ctx* c, const double p[3], int layerSign, int* edgeOut) {
    if (!c || !p || !edgeOut) return fail(EFIE_E_ARGS, "efie:port", "bad arguments");
    if (!c->portValid || c->rwgDevNe < 1)
        return fail(EFIE_E_STATE, "efie:state", "efie_build_rwg_device must come first");
    const int Ne = c->rwgDevNe;
    const size_t n = Ne, nInts = ((7 * n + 1) + 1) & ~(size_t)1;
    int* ti = c->dRwgTab.as<int>();
    const double* dCen = reinterpret_cast<double*>(ti + nInts) + 9 * n;
    const int nb = port_pick_blocks(Ne);
    HIPCHK(c->dPortPart.ensure(sizeof(double) * nb + sizeof(int) * (nb + 1)));
    double* pd = c->dPortPart.as<double>();
    int* pi = reinterpret_cast<int*>(pd + nb);
    hipStream_t s = c->stream;
    HIPCHK(launch_port_pick(dCen, ti + 6 * n, Ne, p, layerSign, pd, pi, pi + nb, s));
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, pi + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (e < 1) return fail(EFIE_E_PORT, "efie:port", "no edge on layer %d", layerSign);
    *edgeOut = e;
    return ok();
}

efie_status efie_set_rwg_tables(efie_ctx* c, int Nt, const int* plusTri, const int* minusTri, const int* plusSign,
                                const int* minusSign, const double* rp, const double* rm, const double* Ap,
                                const double* Am, const double* len, int Ne) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    if (!plusTri || !minusTri || !plusSign || !minusSign || !rp || !rm || !Ap || !Am || !len)
        return fail(EFIE_E_ARGS, "efie:args", "null table");
    if (Nt < 1) return fail(EFIE_E_ARGS, "mex:pre", "r_reg third dim must be Nt >= 1");
    if (Ne < 1) return fail(EFIE_E_ARGS, "mex:rwg", "Missing tri/sign fields");
    for (int e = 0; e < Ne; ++e) {
        if (plusTri[e] < 0 || plusTri[e] > Nt) return fail(EFIE_E_MESH, "mex:rwg", "plusTri out of range");
        if (minusTri[e] < 0 || minusTri[e] > Nt) return fail(EFIE_E_MESH, "mex:rwg", "minusTri out of range");
    }
    if (c->meshless && c->haveRwg && Nt == c->Nt && Ne == c->Ne && same_table(c->plusTri, plusTri, Ne) &&
        same_table(c->minusTri, minusTri, Ne) && same_table(c->plusSign, plusSign, Ne) &&
        same_table(c->minusSign, minusSign, Ne) && same_table(c->rp, rp, 3 * (size_t)Ne) &&
        same_table(c->rm, rm, 3 * (size_t)Ne) && same_table(c->Ap, Ap, Ne) && same_table(c->Am, Am, Ne) &&
        same_table(c->len, len, Ne)) {
        drop_pack(c);  // unchanged tables keep every cache
        return ok();
    }
    c->Nv = 0;
    c->Nt = Nt;
    c->V.clear();
    c->F.clear();
    c->edge.clear();
    c->Ne = Ne;
    c->plusTri.assign(plusTri, plusTri + Ne);
    c->minusTri.assign(minusTri, minusTri + Ne);
    c->plusSign.assign(plusSign, plusSign + Ne);
    c->minusSign.assign(minusSign, minusSign + Ne);
    c->rp.assign(rp, rp + 3 * (size_t)Ne);
    c->rm.assign(rm, rm + 3 * (size_t)Ne);
    c->Ap.assign(Ap, Ap + Ne);
    c->Am.assign(Am, Am + Ne);
    c->len.assign(len, len + Ne);
    c->haveMesh = false;
    c->meshless = true;
    c->haveRwg = true;
    c->midValid = false;
    c->geoValid = false;
    c->cAlpha = -1;
    drop_pack(c);
    return ok();
}

efie_status efie_set_prepack(efie_ctx* c, const efie_prepack* pk) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    if (!pk) {
        drop_pack(c);
        return ok();
    }
    if (!c->haveRwg) return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
    if (!pk->r_reg || !pk->w_reg || !pk->nearMat || !pk->self_r || !pk->self_w || !pk->self_a || !pk->near_r ||
        !pk->near_w || !pk->aeqTri)
        return fail(EFIE_E_ARGS, "mex:prepack", "Missing field in prepack.");
    auto rule = [](int n) { return n == 1 || n == 3 || n == 6 || n == 7; };
    auto subs = [](int n) { return n == 1 || n == 4 || n == 16 || n == 64; };
    if (!rule(pk->Nq)) return fail(EFIE_E_ARGS, "mex:r_reg", "r_reg: Nq must be 1, 3, 6 or 7 points.");
    if (!rule(pk->Nq2)) return fail(EFIE_E_ARGS, "mex:self_r", "self_r: Nq2 must be 1, 3, 6 or 7 points.");
    if (!subs(pk->NsS)) return fail(EFIE_E_ARGS, "mex:self_r", "self_r: NsS must be 1, 4, 16 or 64.");
    if (!subs(pk->NsN)) return fail(EFIE_E_ARGS, "mex:near_r", "near_r: NsN must be 1, 4, 16 or 64.");
    const int Nt = c->Nt, Nq = pk->Nq, Nq2 = pk->Nq2, NsS = pk->NsS, NsN = pk->NsN;
    std::vector<uint8_t> nm((size_t)Nt * Nt);
    for (int t1 = 0; t1 < Nt; ++t1)
        for (int t2 = 0; t2 < Nt; ++t2) {
            const bool a = pk->nearMat[(size_t)t1 + (size_t)Nt * t2] != 0;
            if (a != (pk->nearMat[(size_t)t2 + (size_t)Nt * t1] != 0))
                return fail(EFIE_E_ARGS, "mex:nearMat", "nearMat must be symmetric.");
            nm[(size_t)t1 * Nt + t2] = (a && t1 != t2) ? 1 : 0;
        }
    // efie_precompute_pack.m layouts: x(i, d, p, t) at i + Nq2 (d + 3 (p + Ns t))
    auto cvt = [&](const double* r, const double* w, int Ns, std::vector<Pt>& out) {
        out.resize((size_t)Nt * Ns * Nq2);
        for (int t = 0; t < Nt; ++t)
            for (int p = 0; p < Ns; ++p)
                for (int i = 0; i < Nq2; ++i) {
                    const size_t b = (size_t)i + (size_t)Nq2 * 3 * ((size_t)p + (size_t)Ns * t);
                    out[((size_t)t * Ns + p) * Nq2 + i] = {r[b], r[b + Nq2], r[b + 2 * (size_t)Nq2],
                                                           w[(size_t)i + (size_t)Nq2 * ((size_t)p + (size_t)Ns * t)]};
                }
    };
    std::vector<Pt> reg((size_t)Nt * Nq);
    for (int t = 0; t < Nt; ++t)
        for (int i = 0; i < Nq; ++i) {
            const size_t b = (size_t)i + (size_t)Nq * 3 * t;
            reg[(size_t)t * Nq + i] = {pk->r_reg[b], pk->r_reg[b + Nq], pk->r_reg[b + 2 * (size_t)Nq],
                                       pk->w_reg[(size_t)i + (size_t)Nq * t]};
        }
    std::vector<Pt> near, self;
    cvt(pk->near_r, pk->near_w, NsN, near);
    cvt(pk->self_r, pk->self_w, NsS, self);
    std::vector<double> aSelf((size_t)Nt * NsS);
    for (int t = 0; t < Nt; ++t)
        for (int p = 0; p < NsS; ++p) aSelf[(size_t)t * NsS + p] = pk->self_a[(size_t)p + (size_t)NsS * t];
    // the pack installed last, again: the geometry built from it stays valid
    const bool same = c->packGen > 0 && c->pkNq == Nq && c->pkNq2 == Nq2 && c->pkNsS == NsS && c->pkNsN == NsN &&
                      c->pkNearAlpha == pk->near_alpha && same_table(c->pkReg, reg.data(), reg.size()) &&
                      same_table(c->pkNear, near.data(), near.size()) &&
                      same_table(c->pkSelf, self.data(), self.size()) &&
                      same_table(c->pkASelf, aSelf.data(), aSelf.size()) &&
                      same_table(c->pkAeq, pk->aeqTri, (size_t)Nt) && same_table(c->pkNearMat, nm.data(), nm.size());
    if (same) {
        c->havePack = true;
        return ok();
    }
    c->pkReg.swap(reg);
    c->pkNear.swap(near);
    c->pkSelf.swap(self);
    c->pkASelf.swap(aSelf);
    c->pkAeq.assign(pk->aeqTri, pk->aeqTri + Nt);
    c->pkNearMat.swap(nm);
    c->pkNq = Nq;
    c->pkNq2 = Nq2;
    c->pkNsS = NsS;
    c->pkNsN = NsN;
    c->pkNearAlpha = pk->near_alpha;
    c->havePack = true;
    ++c->packGen;
    c->geoValid = false;
    return ok();
}

}  // extern "C"

namespace {

// ------------------------------------------------------------ profiling --

void prof_mark(void* user, int cls, int launches) {
    efie_ctx* c = static_cast<efie_ctx*>(user);
    auto next = [c]() {
        if (c->poolUsed == c->pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return (hipEvent_t) nullptr;
            c->pool.push_back(e);
        }
        return c->pool[c->poolUsed++];
    };
    hipEvent_t e = next();
    if (!e) return;
    (void)hipEventRecord(e, c->stream);
    if (launches < 0) {
        c->openCls = cls;
        c->openEv = e;
    } else if (c->openCls == cls && c->openEv) {
        c->marks.push_back({cls, launches, c->openEv, e});
        c->openEv = nullptr;
        c->openCls = -1;
    }
}

Prof make_prof(efie_ctx* c) {
    Prof p;
    if (c->prof) {
        p.user = c;
        p.mark = prof_mark;
    }
    return p;
}

// Sum the recorded event pairs into stats (after a stream synchronize).
void prof_collect(efie_ctx* c, efie_stats* st) {
    for (auto& m : c->marks) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess && st) {
            st->ms_kernel[m.cls] += ms;
            st->n_kernel[m.cls] += m.launches;
        }
    }
    c->marks.clear();
    c->poolUsed = 0;
}

// Triangle geometry of precompute_tri_geom.m:11-17 for triangle t (0-based).
void tri_vertices(const efie_ctx* c, int t, V3& A, V3& B, V3& C) {
    const int Nv = c->Nv, Nt = c->Nt;
    auto vx = [&](int vi) { return V3{c->V[vi - 1], c->V[vi - 1 + Nv], c->V[vi - 1 + 2 * Nv]}; };
    A = vx(c->F[t]);
    B = vx(c->F[t + Nt]);
    C = vx(c->F[t + 2 * Nt]);
}

efie_status check_opts(const efie_opts* o) {
    if (o->precision != EFIE_PREC_FP64 && o->precision != EFIE_PREC_MIXED)
        return fail(EFIE_E_ARGS, "efie:opts", "unknown precision");
    if (o->lu_kernel < EFIE_LU_AUTO || o->lu_kernel > EFIE_LU_SYM)
        return fail(EFIE_E_ARGS, "efie:opts", "unknown lu_kernel");
    if (o->ir_max_iter < 0) return fail(EFIE_E_ARGS, "efie:opts", "ir_max_iter must be >= 0");
    if (!(o->ir_tol > 0.0)) return fail(EFIE_E_ARGS, "efie:opts", "ir_tol must be > 0");
    if (o->model == EFIE_MODEL_CENTROID) return ok();
    if (o->model != EFIE_MODEL_QUADRATURE) return fail(EFIE_E_ARGS, "efie:opts", "unknown model");
    if (o->near_subdiv < 0 || o->near_subdiv > 3 || o->self_subdiv < 0 || o->self_subdiv > 3)
        return fail(EFIE_E_ARGS, "efie:opts", "subdivision depth must be 0..3");
    if (o->near_mode != EFIE_NEAR_DIRECT && o->near_mode != EFIE_NEAR_MOMENTS && o->near_mode != EFIE_NEAR_EXTRACT &&
        o->near_mode != EFIE_NEAR_SURROGATE)
        return fail(EFIE_E_ARGS, "efie:opts", "unknown near_mode");
    return ok();
}

// Edge midpoints (port_weights_gaussian.m:7-11) and lengths for the RHS
// and the Zin epilogue, needed by both models.
efie_status ensure_edge_mid(efie_ctx* c) {
    if (c->midValid) return ok();
    const int Ne = c->Ne;
    std::vector<double> midp(3 * (size_t)Ne);
    for (int e = 0; e < Ne; ++e) {
        const int a = c->edge[e] - 1, b = c->edge[e + Ne] - 1;
        for (int d = 0; d < 3; ++d)
            midp[3 * e + d] = 0.5 * (c->V[a + d * (size_t)c->Nv] + c->V[b + d * (size_t)c->Nv]);
    }
    HIPCHK(c->dMid.ensure(sizeof(double) * midp.size()));
    HIPCHK(c->dLen.ensure(sizeof(double) * Ne));
    HIPCHK(hipMemcpyAsync(c->dMid.p, midp.data(), sizeof(double) * midp.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dLen.p, c->len.data(), sizeof(double) * Ne, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->midValid = true;
    return ok();
}

// Centroids binned into cubic cells of side h (just above `reach`, the
// largest centroid distance of a near pair), cell by cell: the near search
// then visits a triangle's own cell and the 26 around it.  Extents with
// more cells than 8 per triangle get larger cells (any h >= reach is
// correct).
struct CellBins {
    double x0 = 0, y0 = 0, z0 = 0, h = 1;
    int nx = 1, ny = 1, nz = 1;
    std::vector<int> cellOf, start, tri;
};

CellBins bin_centroids(const std::vector<TriGeo>& tg, double reach) {
    CellBins b;
    const int Nt = (int)tg.size();
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const TriGeo& g : tg) {
        const double c[3] = {g.cx, g.cy, g.cz};
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], c[d]);
            hi[d] = std::max(hi[d], c[d]);
        }
    }
    double h = std::max(reach * (1.0 + 1e-9), 1e-300);
    auto dims = [&](double hh, int* n) {
        double cells = 1;
        for (int d = 0; d < 3; ++d) {
            n[d] = (int)std::min(1e6, std::floor((hi[d] - lo[d]) / hh) + 1.0);
            cells *= n[d];
        }
        return cells;
    };
    int n[3];
    while (dims(h, n) > 8.0 * Nt + 64) h *= 2;
    b.x0 = lo[0];
    b.y0 = lo[1];
    b.z0 = lo[2];
    b.h = h;
    b.nx = n[0];
    b.ny = n[1];
    b.nz = n[2];
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    b.cellOf.resize(Nt);
    b.start.assign(ncell + 1, 0);
    for (int t = 0; t < Nt; ++t) {
        const double c[3] = {tg[t].cx, tg[t].cy, tg[t].cz};
        int ix[3];
        for (int d = 0; d < 3; ++d) ix[d] = std::min(n[d] - 1, std::max(0, (int)((c[d] - lo[d]) / h)));
        b.cellOf[t] = (ix[2] * n[1] + ix[1]) * n[0] + ix[0];
        ++b.start[b.cellOf[t] + 1];
    }
    for (size_t i = 0; i < ncell; ++i) b.start[i + 1] += b.start[i];
    b.tri.resize(Nt);
    std::vector<int> fill(b.start.begin(), b.start.end() - 1);
    for (int t = 0; t < Nt; ++t) b.tri[fill[b.cellOf[t]]++] = t;
    return b;
}

// Frequency-independent data of the quadrature model, cached per options.
efie_status ensure_geometry(efie_ctx* c, const efie_opts* o) {
    GeoKey k;
    if (c->havePack) {
        k.pack = c->packGen;  // the rule options do not matter
    } else {
        if (c->meshless)
            return fail(EFIE_E_STATE, "efie:state", "RWG tables without a mesh need a prepack (efie_set_prepack)");
        k.quad_order = o->quad_order;
        k.near_order = o->near_order;
        k.near_subdiv = o->near_subdiv;
        k.self_subdiv = o->self_subdiv;
        k.near_sing_eps = o->near_sing_eps;
        k.self_alpha = o->self_alpha;
    }
    if (c->geoValid && c->key == k) return ok();
    const int Nt = c->Nt, Ne = c->Ne;
    hipStream_t s = c->stream;

    std::vector<TriGeo> tg(Nt);
    std::vector<V3> tv;
    double lmax = 0, aeqMax = 0;
    if (!c->meshless) {
        tv.resize(3 * (size_t)Nt);
        for (int t = 0; t < Nt; ++t) {
            V3 A, B, C;
            tri_vertices(c, t, A, B, C);
            tv[3 * t] = A; tv[3 * t + 1] = B; tv[3 * t + 2] = C;
            tg[t].cx = (A.x + B.x + C.x) / 3.0;
            tg[t].cy = (A.y + B.y + C.y) / 3.0;
            tg[t].cz = (A.z + B.z + C.z) / 3.0;
            tg[t].area2 = norm(cross(B - A, C - A));
        }
    } else {
        // centroid and 2A from the pack's regular points: the rules
        // integrate linear functions exactly and their weights sum to 1/2
        const int Nq = c->pkNq;
        for (int t = 0; t < Nt; ++t) {
            double sw = 0, sx = 0, sy = 0, sz = 0;
            for (int i = 0; i < Nq; ++i) {
                const Pt q = c->pkReg[(size_t)t * Nq + i];
                sw += q.w;
                sx += q.w * q.x;
                sy += q.w * q.y;
                sz += q.w * q.z;
            }
            if (!(sw > 0)) return fail(EFIE_E_ARGS, "mex:pre", "w_reg of a triangle must have a positive sum");
            tg[t] = {sx / sw, sy / sw, sz / sw, 2 * sw};
        }
    }
    for (int t = 0; t < Nt; ++t) {
        lmax = std::max(lmax, std::sqrt(4 * (0.5 * tg[t].area2) / std::sqrt(3.0)));
        aeqMax = std::max(aeqMax, std::sqrt((0.5 * tg[t].area2) / kPi));
    }
    int Nq, Nq2, NsN, NsS;
    QuadRules rules{};
    std::vector<Pt> reg, nearPts, selfPts;
    std::vector<double> aSelf, aeq(Nt), rad(Nt, 0.0);
    double rmaxN = 0, rmaxS = 0, aSelfMax = 0;
    // distance of a point from its triangle's centroid, folded into rad
    auto rel = [&](const Pt& q, int t, double& rmax) {
        Pt r = {q.x - tg[t].cx, q.y - tg[t].cy, q.z - tg[t].cz, q.w};
        const double d = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
        rmax = std::max(rmax, d);
        rad[t] = std::max(rad[t], d);
        return r;
    };
    if (c->havePack) {
        Nq = c->pkNq;
        Nq2 = c->pkNq2;
        NsN = c->pkNsN;
        NsS = c->pkNsS;
        double rmaxR = 0;
        reg.resize(c->pkReg.size());
        for (size_t i = 0; i < reg.size(); ++i) reg[i] = rel(c->pkReg[i], (int)(i / Nq), rmaxR);
        nearPts.resize(c->pkNear.size());
        for (size_t i = 0; i < nearPts.size(); ++i) nearPts[i] = rel(c->pkNear[i], (int)(i / ((size_t)NsN * Nq2)), rmaxN);
        selfPts.resize(c->pkSelf.size());
        for (size_t i = 0; i < selfPts.size(); ++i) selfPts[i] = rel(c->pkSelf[i], (int)(i / ((size_t)NsS * Nq2)), rmaxS);
        aSelf = c->pkASelf;
        for (double a : aSelf) aSelfMax = std::max(aSelfMax, a);
        aeq = c->pkAeq;
        aeqMax = 0;
        for (double a : aeq) aeqMax = std::max(aeqMax, a);
    } else {
        // rule-derived points: built on the device from the vertices
        // (launch_quad_points, below)
        double su[16], sv[16], sw[16];
        Nq = quad_rule(o->quad_order, rules.ru, rules.rv, rules.rw);
        Nq2 = quad_rule(o->near_order, su, sv, sw);
        for (int i = 0; i < Nq2; ++i) {
            rules.su[i] = su[i];
            rules.sv[i] = sv[i];
            rules.sw[i] = sw[i];
        }
        rules.nReg = Nq;
        rules.nSub = Nq2;
        NsN = 1 << (2 * o->near_subdiv);
        NsS = 1 << (2 * o->self_subdiv);
        for (int t = 0; t < Nt; ++t) aeq[t] = std::sqrt((0.5 * tg[t].area2) / kPi);
    }
    // radius bound of each triangle about its centroid: every given point
    // and, with mesh arrays, its vertices (the device generator folds in
    // both for rule-derived points)
    for (int t = 0; t < Nt && c->havePack && !c->meshless; ++t) {
        const V3 cc{tg[t].cx, tg[t].cy, tg[t].cz};
        rad[t] = std::max(rad[t], std::max(norm(tv[3 * t] - cc), std::max(norm(tv[3 * t + 1] - cc),
                                                                         norm(tv[3 * t + 2] - cc))));
    }

    // RWG sides: f = c (r - v); rwg_eval_on_triangle.m tests t == plusTri first
    std::vector<int> sideTri(2 * (size_t)Ne);
    std::vector<Side> side(2 * (size_t)Ne);
    for (int e = 0; e < Ne; ++e) {
        for (int sd = 0; sd < 2; ++sd) {
            const int t1 = (sd == 0 ? c->plusTri[e] : c->minusTri[e]);
            if (t1 == 0) {
                sideTri[2 * e + sd] = -1;
                side[2 * e + sd] = {0, 0, 0, 0};
                continue;
            }
            const bool plus = (t1 == c->plusTri[e]);
            const double sg = plus ? c->plusSign[e] : c->minusSign[e];
            const double l = c->len[e], A = plus ? c->Ap[e] : c->Am[e];
            const double* vv = plus ? c->rp.data() : c->rm.data();
            const double cc = plus ? sg * (l / (2 * A)) : -(sg * (l / (2 * A)));
            const TriGeo g = tg[t1 - 1];
            sideTri[2 * e + sd] = t1 - 1;
            side[2 * e + sd] = {vv[e] - g.cx, vv[e + Ne] - g.cy, vv[e + 2 * Ne] - g.cz, cc};
        }
    }

    // vertices relative to the centroid (singularity extraction; zero
    // without mesh arrays, where extraction is not offered)
    std::vector<double> triV(9 * (size_t)Nt, 0.0);
    for (int t = 0; t < Nt && !c->meshless; ++t)
        for (int q = 0; q < 3; ++q) {
            const V3 v = tv[3 * t + q];
            triV[9 * (size_t)t + 3 * q] = v.x - tg[t].cx;
            triV[9 * (size_t)t + 3 * q + 1] = v.y - tg[t].cy;
            triV[9 * (size_t)t + 3 * q + 2] = v.z - tg[t].cz;
        }
    HIPCHK(c->dTriV.ensure(sizeof(double) * triV.size()));
    HIPCHK(hipMemcpyAsync(c->dTriV.p, triV.data(), sizeof(double) * triV.size(), hipMemcpyHostToDevice, s));
    HIPCHK(c->dTri.ensure(sizeof(TriGeo) * Nt));
    const size_t nReg = (size_t)Nt * Nq, nNearPts = (size_t)Nt * NsN * Nq2, nSelfPts = (size_t)Nt * NsS * Nq2;
    HIPCHK(c->dReg.ensure(sizeof(Pt) * nReg));
    HIPCHK(c->dNearPts.ensure(sizeof(Pt) * nNearPts));
    HIPCHK(c->dSelfPts.ensure(sizeof(Pt) * nSelfPts));
    HIPCHK(c->dASelf.ensure(sizeof(double) * (size_t)Nt * NsS));
    HIPCHK(c->dAeq.ensure(sizeof(double) * Nt));
    HIPCHK(c->dRad.ensure(sizeof(double) * Nt));
    HIPCHK(c->dSideTri.ensure(sizeof(int) * sideTri.size()));
    HIPCHK(c->dSide.ensure(sizeof(Side) * side.size()));
    HIPCHK(c->dRowCnt.ensure(sizeof(int) * Nt));
    HIPCHK(hipMemcpyAsync(c->dTri.p, tg.data(), sizeof(TriGeo) * Nt, hipMemcpyHostToDevice, s));
    double radMax = 0;
    if (c->havePack) {
        HIPCHK(hipMemcpyAsync(c->dReg.p, reg.data(), sizeof(Pt) * nReg, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dNearPts.p, nearPts.data(), sizeof(Pt) * nNearPts, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dSelfPts.p, selfPts.data(), sizeof(Pt) * nSelfPts, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dASelf.p, aSelf.data(), sizeof(double) * aSelf.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dRad.p, rad.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, s));
        for (int t = 0; t < Nt; ++t) radMax = std::max(radMax, rad[t]);
    } else {
        // every point on the device from the vertices (9 doubles per
        // triangle), with the radius maxima
        std::vector<double> tvAbs(9 * (size_t)Nt);
        for (int t = 0; t < Nt; ++t)
            for (int q = 0; q < 3; ++q) {
                tvAbs[9 * (size_t)t + 3 * q] = tv[3 * t + q].x;
                tvAbs[9 * (size_t)t + 3 * q + 1] = tv[3 * t + q].y;
                tvAbs[9 * (size_t)t + 3 * q + 2] = tv[3 * t + q].z;
            }
        HIPCHK(c->dTriAbs.ensure(sizeof(double) * (tvAbs.size() + 4)));
        double* dMx = c->dTriAbs.as<double>() + tvAbs.size();
        HIPCHK(hipMemcpyAsync(c->dTriAbs.p, tvAbs.data(), sizeof(double) * tvAbs.size(), hipMemcpyHostToDevice, s));
        HIPCHK(launch_quad_points(c->dTriAbs.as<double>(), c->dTri.as<TriGeo>(), Nt, rules, o->near_subdiv,
                                  o->self_subdiv, o->self_alpha, c->dReg.as<Pt>(), c->dNearPts.as<Pt>(),
                                  c->dSelfPts.as<Pt>(), c->dASelf.as<double>(), c->dRad.as<double>(), dMx, s));
        double mx[4];
        HIPCHK(hipMemcpyAsync(mx, dMx, sizeof(mx), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        rmaxN = mx[0];
        rmaxS = mx[1];
        aSelfMax = mx[2];
        radMax = mx[3];
    }

    HIPCHK(hipMemcpyAsync(c->dAeq.p, aeq.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dSideTri.p, sideTri.data(), sizeof(int) * sideTri.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dSide.p, side.data(), sizeof(Side) * side.size(), hipMemcpyHostToDevice, s));

    // near pairs: the pack's map (kept on the device: the pack defines
    // them), or triangles_are_near.m on the device over binned centroids
    // (no Nt x Nt map); then the row counts of t2 > t1 and the pair list
    // ordered by t1, then t2
    std::vector<int> cnt(Nt, 0);
    double dNearMax = 0;
    CellBins bins;
    if (c->havePack) {
        for (int t1 = 0; t1 < Nt; ++t1)
            for (int t2 = t1 + 1; t2 < Nt; ++t2)
                if (c->pkNearMat[(size_t)t1 * Nt + t2]) {
                    ++cnt[t1];
                    const double dx = tg[t1].cx - tg[t2].cx, dy = tg[t1].cy - tg[t2].cy, dz = tg[t1].cz - tg[t2].cz;
                    dNearMax = std::max(dNearMax, std::sqrt(dx * dx + dy * dy + dz * dz));
                }
        HIPCHK(c->dNearMat.ensure((size_t)Nt * Nt));
        HIPCHK(hipMemcpyAsync(c->dNearMat.p, c->pkNearMat.data(), (size_t)Nt * Nt, hipMemcpyHostToDevice, s));
    } else {
        c->dNearMat.release();
        dNearMax = o->near_sing_eps * lmax;
        bins = bin_centroids(tg, dNearMax);
        HIPCHK(c->dCellOf.ensure(sizeof(int) * Nt));
        HIPCHK(c->dCellStart.ensure(sizeof(int) * bins.start.size()));
        HIPCHK(c->dCellTri.ensure(sizeof(int) * Nt));
        HIPCHK(hipMemcpyAsync(c->dCellOf.p, bins.cellOf.data(), sizeof(int) * Nt, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dCellStart.p, bins.start.data(), sizeof(int) * bins.start.size(),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dCellTri.p, bins.tri.data(), sizeof(int) * Nt, hipMemcpyHostToDevice, s));
        HIPCHK(launch_near_bins(c->dTri.as<TriGeo>(), Nt, o->near_sing_eps, bins.x0, bins.y0, bins.z0, bins.h,
                                bins.nx, bins.ny, bins.nz, c->dCellOf.as<int>(), c->dCellStart.as<int>(),
                                c->dCellTri.as<int>(), c->dRowCnt.as<int>(), nullptr, nullptr, s));
        HIPCHK(hipMemcpyAsync(cnt.data(), c->dRowCnt.p, sizeof(int) * Nt, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> start(Nt);
    long long tot = 0;
    for (int t = 0; t < Nt; ++t) {
        start[t] = (int)tot;
        tot += cnt[t];
    }
    HIPCHK(c->dList.ensure(sizeof(int2) * (size_t)std::max(1LL, tot)));
    HIPCHK(hipMemcpyAsync(c->dRowCnt.p, start.data(), sizeof(int) * Nt, hipMemcpyHostToDevice, s));
    if (c->havePack)
        HIPCHK(launch_near_list(c->dNearMat.as<uint8_t>(), c->dRowCnt.as<int>(), Nt, c->dList.as<int2>(), s));
    else
        HIPCHK(launch_near_bins(c->dTri.as<TriGeo>(), Nt, o->near_sing_eps, bins.x0, bins.y0, bins.z0, bins.h,
                                bins.nx, bins.ny, bins.nz, c->dCellOf.as<int>(), c->dCellStart.as<int>(),
                                c->dCellTri.as<int>(), nullptr, c->dRowCnt.as<int>(), c->dList.as<int2>(), s));
    HIPCHK(hipStreamSynchronize(s));
    c->nearEps = c->havePack ? 0.0 : o->near_sing_eps;

    c->Nq = Nq;
    c->NpNear = NsN * Nq2;
    c->Nq2s = Nq2;
    c->NsSelf = NsS;
    c->nNear = tot;
    c->rNearMax = rmaxN;
    c->rSelfMax = rmaxS;
    c->lMax = lmax;
    c->dNearMax = dNearMax;
    c->aNearMax = aeqMax;
    c->aSelfMax = aSelfMax;
    double dcMax = 0;
    for (int e = 0; e < Ne; ++e) {
        const int tp = c->plusTri[e], tm = c->minusTri[e];
        if (tp > 0 && tm > 0) {
            const double dx = tg[tp - 1].cx - tg[tm - 1].cx, dy = tg[tp - 1].cy - tg[tm - 1].cy,
                         dz = tg[tp - 1].cz - tg[tm - 1].cz;
            dcMax = std::max(dcMax, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
    }
    c->radMax = radMax;
    c->dcMax = dcMax;
    c->key = k;
    c->geoValid = true;
    g_geoBuilds.fetch_add(1);
    c->serP = -1;  // series and moments belong to the old geometry
    c->momP = -1;
    c->surValid = false;
    return ok();
}

// Series length for the moment expansion: the smallest P with
// (x^(P+1)/(P+1)!) e^x <= 1e-17, x = |k|max * Rmax, where Rmax bounds the
// (softened) distance of every near and self point pair.  -1 if P > 60.
int moment_terms(const efie_ctx* c, const efie_opts* o, double kmax) {
    const bool extract = o->near_mode == EFIE_NEAR_EXTRACT;  // no softening
    const double dN = c->dNearMax;  // centroid distance of a near pair
    const double aN = extract ? 0.0 : o->near_alpha * c->aNearMax;
    const double aS = extract ? 0.0 : c->aSelfMax;
    const double rN = std::sqrt((dN + 2 * c->rNearMax) * (dN + 2 * c->rNearMax) + aN * aN);
    const double rS = std::sqrt(4 * c->rSelfMax * c->rSelfMax + aS * aS);
    const double x = kmax * std::max(rN, rS) * 1.0001;
    double term = std::exp(x);  // e^x x^n / n!
    for (int n = 1; n <= 61; ++n) {
        term *= x / n;
        if (term <= 1e-17) return n - 1;
    }
    return -1;
}

// Moments of the near and self pairs (selfOnly: self pairs only) for
// series length P+1 (cached; a longer series than cached recomputes them).
efie_status ensure_moments(efie_ctx* c, const efie_opts* o, int P, efie_stats* st, bool selfOnly) {
    const bool extract = o->near_mode == EFIE_NEAR_EXTRACT;
    if (c->momP >= P && c->momExtract == extract && (extract || c->nearAlphaMom == o->near_alpha) &&
        (selfOnly || !c->momSelfOnly))
        return ok();
    hipStream_t s = c->stream;
    const size_t per = (size_t)(P + 1) * 8;
    HIPCHK(c->dQn.ensure(sizeof(double) * per * (size_t)std::max(1LL, c->nNear)));
    HIPCHK(c->dQs.ensure(sizeof(double) * per * (size_t)c->Nt));
    HIPCHK(hipEventRecord(c->ev[6], s));
    HIPCHK(launch_moments(c->dNearPts.as<Pt>(), c->NpNear, c->dTri.as<TriGeo>(), c->dAeq.as<double>(), o->near_alpha,
                          c->dList.as<int2>(), selfOnly ? 0 : (int)c->nNear, c->dSelfPts.as<Pt>(), c->Nq2s, c->NsSelf,
                          c->dASelf.as<double>(), c->Nt, P, c->dQn.as<double>(), c->dQs.as<double>(),
                          extract ? c->dTriV.as<double>() : nullptr, s));
    HIPCHK(hipEventRecord(c->ev[7], s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    if (st) st->ms_moments += ms;
    g_momBuilds.fetch_add(1);
    c->momP = P;
    c->nearAlphaMom = o->near_alpha;
    c->momExtract = extract;
    c->momSelfOnly = selfOnly;
    return ok();
}

// Series length of the all-pairs fill (near_mode MOMENTS): the smallest P
// with (x^(P+1)/(P+1)!) e^x <= 1e-17, x = |k|max * b, where b bounds
// |R - R0e| for every point pair of every entry: the distance of a point
// pair from its own pair's centroid distance (two point radii plus the
// near or self softening length) plus the distance of that centroid
// distance from the entry's centre R0e, the middle of the four pairs'
// (half their range, which the two RWGs' own centroid distances bound:
// at most dcMax).  -1 if P would exceed 21.
int series_terms(const efie_ctx* c, const efie_opts* o, double kmax) {
    const double soft = std::max(o->near_alpha * c->aNearMax, c->aSelfMax);
    const double x = kmax * (2 * c->radMax + soft + c->dcMax) * 1.0001;
    double term = std::exp(x);
    for (int n = 1; n <= 22; ++n) {
        term *= x / n;
        if (term <= 1e-17) return n - 1;
    }
    return -1;
}

// Frequency-independent series of all triangle pairs for at least P + 1
// terms (rounded up to a size the fill kernel is built for).  *ready is
// false when the series is too long or would take more than a third of
// the free device memory; the caller then fills per frequency.
efie_status ensure_series(efie_ctx* c, const efie_opts* o, double kmax, efie_stats* st, bool* ready) {
    *ready = false;
    const int P0 = series_terms(c, o, kmax);
    if (P0 < 0) return ok();
    if (c->serP >= P0 && c->serAlpha == o->near_alpha && c->dSerC.p) {
        *ready = true;
        return ok();
    }
    const int NT = series_nt(P0);
    if (NT == 0) return ok();
    const int P = NT - 3;
    const size_t qBytes = sizeof(double) * 8 * (size_t)(P + 1) * rec_index(0, c->Nt);
    const size_t cBytes = series_coeff_bytes(c->Ne, NT);
    size_t freeB = 0, totB = 0;
    HIPCHK(hipMemGetInfo(&freeB, &totB));
    const size_t have = c->dSerC.bytes;
    if (qBytes + cBytes > have && qBytes + cBytes - have > freeB / 3) return ok();
    hipStream_t s = c->stream;
    c->serP = -1;
    HIPCHK(c->dSerC.ensure(cBytes));
    HIPCHK(c->dQall.ensure(qBytes));
    HIPCHK(hipEventRecord(c->ev[6], s));
    HIPCHK(launch_series(c->dReg.as<Pt>(), c->Nq, c->dTri.as<TriGeo>(), c->Nt,
                         c->havePack ? c->dNearMat.as<uint8_t>() : nullptr, c->nearEps,
                         c->dNearPts.as<Pt>(), c->NpNear, c->Nq2s, c->dAeq.as<double>(), o->near_alpha,
                         c->dList.as<int2>(), (int)c->nNear, c->dSelfPts.as<Pt>(), c->NsSelf, c->dASelf.as<double>(),
                         P, c->dQall.as<double>(), s));
    HIPCHK(launch_series_coeffs(c->dQall.as<double>(), NT, c->dTri.as<TriGeo>(), c->dSideTri.as<int>(),
                                c->dSide.as<Side>(), c->Ne, c->dSerC.as<double>(), s));
    HIPCHK(hipEventRecord(c->ev[7], s));
    HIPCHK(hipStreamSynchronize(s));
    c->dQall.release();  // only the per-entry coefficients are kept
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    if (st) st->ms_moments += ms;
    g_momBuilds.fetch_add(1);
    c->serP = P;
    c->serAlpha = o->near_alpha;
    *ready = true;
    return ok();
}

// Host precompute of the centroid model (patch_sweep_cuda.cu:244-326).
efie_status ensure_centroid(efie_ctx* c, double selfAlpha) {
    if (c->cAlpha == selfAlpha && c->dCPos.p) return ok();
    const int Nt = c->Nt, Ne = c->Ne;
    std::vector<V3> cent(Nt);
    std::vector<double> asoft(Nt);
    for (int t = 0; t < Nt; ++t) {
        V3 A, B, C;
        tri_vertices(c, t, A, B, C);
        cent[t] = {(A.x + B.x + C.x) / 3.0, (A.y + B.y + C.y) / 3.0, (A.z + B.z + C.z) / 3.0};
        const double area2 = norm(cross(B - A, C - A));
        asoft[t] = selfAlpha * std::sqrt((0.5 * area2) / kPi);
    }
    std::vector<double> pos(6 * (size_t)Ne, 0.0), f(6 * (size_t)Ne, 0.0), dv(2 * (size_t)Ne, 0.0),
        w(2 * (size_t)Ne, 0.0);
    std::vector<int> tsd(2 * (size_t)Ne);
    for (int e = 0; e < Ne; ++e) {
        const int tp = c->plusTri[e], tm = c->minusTri[e];
        tsd[2 * e] = tp;
        tsd[2 * e + 1] = tm;
        const double l = c->len[e];
        if (tp > 0) { pos[6 * e] = cent[tp - 1].x; pos[6 * e + 1] = cent[tp - 1].y; pos[6 * e + 2] = cent[tp - 1].z; }
        if (tm > 0) { pos[6 * e + 3] = cent[tm - 1].x; pos[6 * e + 4] = cent[tm - 1].y; pos[6 * e + 5] = cent[tm - 1].z; }
        w[2 * e] = tp > 0 ? c->Ap[e] : 0.0;
        w[2 * e + 1] = tm > 0 ? c->Am[e] : 0.0;
        if (tp > 0 && c->Ap[e] > 0) {
            const double sc = (double)c->plusSign[e] * (l / (2.0 * c->Ap[e]));
            for (int d = 0; d < 3; ++d) f[6 * e + d] = sc * (pos[6 * e + d] - c->rp[e + d * (size_t)Ne]);
            dv[2 * e] = (double)c->plusSign[e] * (l / c->Ap[e]);
        }
        if (tm > 0 && c->Am[e] > 0) {
            const double sc = (double)c->minusSign[e] * (l / (2.0 * c->Am[e]));
            for (int d = 0; d < 3; ++d) f[6 * e + 3 + d] = sc * (c->rm[e + d * (size_t)Ne] - pos[6 * e + 3 + d]);
            dv[2 * e + 1] = (double)c->minusSign[e] * (-(l / c->Am[e]));
        }
    }
    hipStream_t s = c->stream;
    HIPCHK(c->dCPos.ensure(sizeof(double) * pos.size()));
    HIPCHK(c->dCF.ensure(sizeof(double) * f.size()));
    HIPCHK(c->dCDv.ensure(sizeof(double) * dv.size()));
    HIPCHK(c->dCW.ensure(sizeof(double) * w.size()));
    HIPCHK(c->dCT.ensure(sizeof(int) * tsd.size()));
    HIPCHK(c->dCAsoft.ensure(sizeof(double) * Nt));
    HIPCHK(hipMemcpyAsync(c->dCPos.p, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dCF.p, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dCDv.p, dv.data(), sizeof(double) * dv.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dCW.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dCT.p, tsd.data(), sizeof(int) * tsd.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dCAsoft.p, asoft.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->cAlpha = selfAlpha;
    return ok();
}

// h = -alpha / (4 pi k^2): with z = -jk, alpha = -eta z and the series
// fill's -eta/(4 pi) e^{-jkR0} z^{-1} p(z) is h e^{-jkR0} p(z).
void set_series_factor(Freq& f) {
    const std::complex<double> k(f.kr, f.ki), a(f.ar, f.ai);
    const std::complex<double> h = -a / (4.0 * kPi * k * k);
    f.hr = h.real();
    f.hi = h.imag();
}

Freq make_freq(std::complex<double> k, double omega, double mu, std::complex<double> eps) {
    Freq f;
    f.kr = k.real();
    f.ki = k.imag();
    // alpha = 1j*omega*mu, beta = 1/(1j*omega*eps)  (assemble_ZV_EFIE_RWG_DIFF_CPU.m:30)
    f.ar = 0.0;
    f.ai = omega * mu;
    std::complex<double> jwe = std::complex<double>(0.0, omega) * eps;
    double d = jwe.real() * jwe.real() + jwe.imag() * jwe.imag();
    f.br = jwe.real() / d;
    f.bi = -jwe.imag() / d;
    set_series_factor(f);
    return f;
}

// How the near and self pairs of one call are integrated.
struct NearPlan {
    bool momNear = false;  // near-pair records from the cached moments
    bool momSelf = false;  // self-pair records from the cached moments
    bool sur = false;      // near pairs in the surrogate's envelope: regular rule + network
    bool series = false;   // every pair from the frequency-independent series (one kernel per batch)
    int serNT = 0;         // its coefficients per entry
    bool lossy = false;    // some k of the call is complex
};

// Fill one Z (interleaved complex, element (m,n) at Z[2*(m*ld + n)], which
// for the symmetric Z is also the column-major matrix).  With Z32
// non-null the quadrature model also writes the complex64 copy
// (Z32[2*(m*ld + n)]) and sets *wrote32.
efie_status fill_z(efie_ctx* c, const efie_opts* o, const NearPlan& plan, const Freq* dfr, double* Z, size_t ld,
                   float* Z32 = nullptr, bool* wrote32 = nullptr, hipStream_t s = nullptr) {
    if (!s) s = c->stream;
    Prof pf = make_prof(c);
    if (wrote32) *wrote32 = false;
    if (o->model == EFIE_MODEL_CENTROID) {
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_centroid(c->dCPos.as<double>(), c->dCF.as<double>(), c->dCDv.as<double>(),
                               c->dCW.as<double>(), c->dCT.as<int>(), c->dCAsoft.as<double>(), c->Ne, dfr,
                               Z, ld, 1, s));
        pf.end(EFIE_K_OTHER);
        return ok();
    }
    if (plan.series) {
        pf.begin(EFIE_K_SERIES);
        HIPCHK(launch_series_fill(c->dSerC.as<double>(), plan.serNT, plan.lossy, c->Ne, dfr, 1, Z, ld, 0, Z32, ld,
                                  0, s));
        pf.end(EFIE_K_SERIES);
        if (wrote32) *wrote32 = Z32 != nullptr;
        return ok();
    }
    const int Nt = c->Nt;
    HIPCHK(c->dRecs.ensure(sizeof(Rec) * rec_index(0, Nt)));
    Rec* recs = c->dRecs.as<Rec>();
    if (plan.momSelf) {
        pf.begin(EFIE_K_POLY);
        HIPCHK(launch_poly_records(c->dQn.as<double>(), c->dList.as<int2>(), plan.momNear ? (int)c->nNear : 0,
                                   c->dQs.as<double>(), Nt, c->momP, dfr, recs, s));
        pf.end(EFIE_K_POLY);
    } else {
        pf.begin(EFIE_K_SELF);
        HIPCHK(launch_rec_self(c->dSelfPts.as<Pt>(), c->Nq2s, c->NsSelf, c->dASelf.as<double>(), c->dRad.as<double>(),
                               c->aSelfMax, Nt, dfr, recs, s));
        pf.end(EFIE_K_SELF);
    }
    const int nDirect = plan.momNear ? 0 : plan.sur ? c->nSurFb : (int)c->nNear;
    if (nDirect > 0) {
        pf.begin(EFIE_K_NEAR);
        HIPCHK(launch_rec_near(c->dNearPts.as<Pt>(), c->NpNear, c->dTri.as<TriGeo>(), c->dRad.as<double>(),
                               c->dAeq.as<double>(), o->near_alpha,
                               plan.sur ? c->dSurFb.as<int2>() : c->dList.as<int2>(), nDirect, dfr, recs, s));
        pf.end(EFIE_K_NEAR);
    }
    pf.begin(EFIE_K_REGULAR);
    HIPCHK(launch_rec_regular(c->dReg.as<Pt>(), c->Nq, c->dTri.as<TriGeo>(), c->dRad.as<double>(), Nt,
                              c->havePack ? c->dNearMat.as<uint8_t>() : nullptr, c->nearEps, dfr, recs, s));
    pf.end(EFIE_K_REGULAR);
    if (plan.sur) {
        // the records of the near pairs inside the envelope: product rule
        // and network correction in one kernel
        pf.begin(EFIE_K_SURROGATE);
        const double* Tb = nullptr;
        if (c->surExact) {
            HIPCHK(launch_sur_labels(c->dSurTri18.as<double>(), nullptr, dfr, c->dSurPairs.as<SurPair>(), c->nSurIn,
                                     c->dSurSub.as<double>(), c->surNp, c->dSurReg.as<double>(), c->surNq,
                                     o->near_alpha, c->dSurT.as<double>(), nullptr, s));
            Tb = c->dSurT.as<double>();
        }
        HIPCHK(launch_sur_records(c->dSurPairs.as<SurPair>(), c->nSurIn, c->dSurModel.as<float>(), Tb,
                                  c->dSurTri18.as<double>(), c->dSurReg.as<double>(), c->surNq, dfr, recs, s));
        pf.end(EFIE_K_SURROGATE, c->surExact ? 2 : 1);
    }
    pf.begin(EFIE_K_COMBINE);
    HIPCHK(launch_combine(recs, c->dSideTri.as<int>(), c->dSide.as<Side>(), c->Ne, dfr, Z, ld, Z32, ld, s));
    pf.end(EFIE_K_COMBINE);
    if (wrote32) *wrote32 = Z32 != nullptr;
    return ok();
}

// Port weights and V on the device; fails with mex:rhs like the reference
// when a layer gets no weight.
efie_status make_rhs(efie_ctx* c, int pe, int ge, double V0, double spread) {
    if (c->meshless) return fail(EFIE_E_STATE, "efie:state", "the excitation needs the mesh arrays (efie_set_mesh)");
    hipStream_t s = c->stream;
    const int Ne = c->Ne;
    HIPCHK(c->dWp.ensure(sizeof(double) * Ne));
    HIPCHK(c->dWg.ensure(sizeof(double) * Ne));
    HIPCHK(c->dSums.ensure(sizeof(double) * 2));
    HIPCHK(c->dV.ensure(sizeof(double) * 2 * Ne));
    HIPCHK(c->dWl.ensure(sizeof(double) * 2 * Ne));
    HIPCHK(launch_port_weights(c->dMid.as<double>(), Ne, pe - 1, spread, c->dWp.as<double>(), c->dSums.as<double>(), 0, s));
    HIPCHK(launch_port_weights(c->dMid.as<double>(), Ne, ge - 1, spread, c->dWg.as<double>(), c->dSums.as<double>(), 1, s));
    double sums[2];
    HIPCHK(hipMemcpyAsync(sums, c->dSums.p, sizeof sums, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (!(sums[0] > 1e-30) || !(sums[1] > 1e-30))
        return fail(EFIE_E_RHS, "mex:rhs", "spreadRadius too small (no weights).");
    HIPCHK(launch_rhs(c->dWp.as<double>(), c->dWg.as<double>(), c->dLen.as<double>(), Ne, V0, c->dV.as<double>(),
                      c->dWl.as<double>(), s));
    return ok();
}

// With a prepack the near softening is the pack's, and only the near modes
// that integrate the pack's points apply; without mesh arrays only the
// quadrature model can run.
efie_status apply_pack(efie_ctx* c, efie_opts* o) {
    if (c->meshless && o->model != EFIE_MODEL_QUADRATURE)
        return fail(EFIE_E_STATE, "efie:state", "the centroid model needs the mesh arrays");
    if (!c->havePack || o->model != EFIE_MODEL_QUADRATURE) return ok();
    if (o->near_mode != EFIE_NEAR_DIRECT && o->near_mode != EFIE_NEAR_MOMENTS)
        return fail(EFIE_E_ARGS, "efie:opts", "with a prepack near_mode must be DIRECT or MOMENTS");
    o->near_alpha = c->pkNearAlpha;
    return ok();
}

efie_status prepare(efie_ctx* c, efie_opts* o, double selfAlphaCentroid) {
    if (!(c->haveMesh || c->meshless) || !c->haveRwg)
        return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
    HIPCHK(hipSetDevice(c->device));
    efie_status st = check_opts(o);
    if (st) return st;
    st = apply_pack(c, o);
    if (st) return st;
    if (!c->meshless) {
        st = ensure_edge_mid(c);
        if (st) return st;
    }
    if (o->model == EFIE_MODEL_CENTROID) return ensure_centroid(c, selfAlphaCentroid);
    return ensure_geometry(c, o);
}

// Z of one frequency into the host array Z ([Ne x Ne] column-major
// interleaved complex), prepare() done: alpha = j omega mu, beta =
// 1/(j omega eps) (gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:341-351).
efie_status fill_single(efie_ctx* c, const efie_opts& o, double k_re, double k_im, double omega, double mu,
                        double eps_re, double eps_im, double* Z);

// Per-mesh data of the surrogate: the near pairs inside the installed
// model's envelope (frames, features), the rest, and the near map without
// the former.  Quadrature options other than the trained ones leave every
// pair to the quadrature.
efie_status ensure_surrogate(efie_ctx* c, const efie_opts* o) {
    if (c->surValid && c->surAlpha == o->near_alpha) return ok();
    hipStream_t s = c->stream;
    const float* env = c->surBlob.data() + kSurModel;
    const float* meta = env + 2 * kMlpIn;
    const bool optsOk = (float)o->near_alpha == meta[kMetaNearAlpha] && o->near_subdiv == (int)meta[kMetaNearSubdiv] &&
                        o->near_order == (int)meta[kMetaNearOrder] && o->quad_order == (int)meta[kMetaQuadOrder];
    const long long nN = c->nNear;
    std::vector<int2> list((size_t)nN);
    if (nN > 0) {
        HIPCHK(hipMemcpyAsync(list.data(), c->dList.p, sizeof(int2) * (size_t)nN, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    std::vector<SurPair> in;
    std::vector<int2> fb;
    std::vector<double> t18;
    double lmax = 0, dmax = 0;
    for (long long i = 0; i < nN; ++i) {
        const int t1 = list[i].x, t2 = list[i].y;
        V3 v[6];
        tri_vertices(c, t1, v[0], v[1], v[2]);
        tri_vertices(c, t2, v[3], v[4], v[5]);
        double w[18], tt[18];
        for (int q = 0; q < 6; ++q) {
            w[3 * q] = v[q].x;
            w[3 * q + 1] = v[q].y;
            w[3 * q + 2] = v[q].z;
        }
        SurPair p;
        const double dl = sur_frame(w, w + 9, &p, tt);
        bool inside = optsOk && dl >= meta[kMetaDlMin] && dl <= meta[kMetaDlMax];
        for (int f = 0; f < 12 && inside; ++f) inside = p.feat[f] >= env[f] && p.feat[f] <= env[kMlpIn + f];
        if (!inside) {
            fb.push_back(list[i]);
            continue;
        }
        p.t1 = t1;
        p.t2 = t2;
        in.push_back(p);
        t18.insert(t18.end(), tt, tt + 18);
        lmax = std::max(lmax, p.L);
        dmax = std::max(dmax, p.d);
    }
    std::vector<double> sub, reg;
    sur_rule_tables(o->quad_order, o->near_order, o->near_subdiv, sub, reg);
    const size_t nIn = in.size(), nFb = fb.size();
    HIPCHK(c->dSurPairs.ensure(sizeof(SurPair) * std::max<size_t>(1, nIn)));
    HIPCHK(c->dSurTri18.ensure(sizeof(double) * 18 * std::max<size_t>(1, nIn)));
    HIPCHK(c->dSurT.ensure(sizeof(double) * 16 * std::max<size_t>(1, nIn)));
    HIPCHK(c->dSurFb.ensure(sizeof(int2) * std::max<size_t>(1, nFb)));
    HIPCHK(c->dSurModel.ensure(sizeof(float) * kSurModel));
    HIPCHK(c->dSurSub.ensure(sizeof(double) * sub.size()));
    HIPCHK(c->dSurReg.ensure(sizeof(double) * reg.size()));
    if (nIn) {
        HIPCHK(hipMemcpyAsync(c->dSurPairs.p, in.data(), sizeof(SurPair) * nIn, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->dSurTri18.p, t18.data(), sizeof(double) * t18.size(), hipMemcpyHostToDevice, s));
    }
    if (nFb) HIPCHK(hipMemcpyAsync(c->dSurFb.p, fb.data(), sizeof(int2) * nFb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dSurModel.p, c->surBlob.data(), sizeof(float) * kSurModel, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dSurSub.p, sub.data(), sizeof(double) * sub.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dSurReg.p, reg.data(), sizeof(double) * reg.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->nSurIn = (int)nIn;
    c->nSurFb = (int)nFb;
    c->surNp = (int)(sub.size() / 4);
    c->surNq = (int)(reg.size() / 4);
    c->surLmax = lmax;
    c->surDmax = dmax;
    c->surAlpha = o->near_alpha;
    c->surValid = true;
    return ok();
}

// Decide the near/self integration for a set of wavenumbers (largest
// |k| kmax; lossy: some k is complex): the moment series when asked for
// and short enough, otherwise direct quadrature; near_mode SURROGATE takes
// the near pairs inside the model's envelope when every k of the call is
// inside it too.
efie_status choose_near(efie_ctx* c, const efie_opts* o, double kmax, bool lossy, efie_stats* st,
                        NearPlan* plan) {
    *plan = NearPlan();
    plan->lossy = lossy;
    if (o->model != EFIE_MODEL_QUADRATURE || o->near_mode == EFIE_NEAR_DIRECT) return ok();
    if (o->near_mode == EFIE_NEAR_MOMENTS) {
        bool ready = false;
        efie_status e = ensure_series(c, o, kmax, st, &ready);
        if (e) return e;
        if (ready) {
            plan->series = true;
            plan->serNT = c->serP + 3;
            if (st) st->moment_terms = c->serP + 1;
            return ok();
        }
    }
    const bool sur = o->near_mode == EFIE_NEAR_SURROGATE;
    if (sur) {
        if (c->surBlob.empty())
            return fail(EFIE_E_STATE, "efie:surrogate", "near_mode surrogate needs a model (efie_sur_train/efie_sur_set)");
        efie_status e = ensure_surrogate(c, o);
        if (e) return e;
        const float* env = c->surBlob.data() + kSurModel;
        const float* meta = env + 2 * kMlpIn;
        plan->sur = c->nSurIn > 0 && !lossy && kmax * c->surLmax <= meta[kMetaKlMax] &&
                    kmax * c->surDmax <= env[kMlpIn + 12];
        if (st) {
            st->n_sur_pairs = plan->sur ? c->nSurIn : 0;
            st->n_sur_fallback = plan->sur ? c->nSurFb : c->nNear;
        }
    }
    const int P = moment_terms(c, o, kmax);
    if (P < 0) {
        // the extraction exists only in series form
        if (o->near_mode == EFIE_NEAR_EXTRACT)
            return fail(EFIE_E_ARGS, "efie:opts", "near_mode extract: |k| times the near-pair size is too large");
        return ok();
    }
    efie_status e = ensure_moments(c, o, P, st, sur);
    if (e) return e;
    plan->momSelf = true;
    plan->momNear = !sur;
    if (st) st->moment_terms = c->momP + 1;
    return ok();
}

efie_status fill_single(efie_ctx* c, const efie_opts& o, double k_re, double k_im, double omega, double mu,
                        double eps_re, double eps_im, double* Z) {
    NearPlan plan;
    efie_status st = choose_near(c, &o, std::hypot(k_re, k_im), k_im != 0.0, nullptr, &plan);
    if (st) return st;
    const int Ne = c->Ne;
    hipStream_t s = c->stream;
    Freq fr;
    fr.kr = k_re;
    fr.ki = k_im;
    fr.ar = 0.0;
    fr.ai = omega * mu;
    {
        double dr = -omega * eps_im, di = omega * eps_re, d2 = dr * dr + di * di;
        fr.br = dr / d2;
        fr.bi = -di / d2;
    }
    set_series_factor(fr);
    HIPCHK(c->dFreq.ensure(sizeof(Freq)));
    HIPCHK(hipMemcpyAsync(c->dFreq.p, &fr, sizeof fr, hipMemcpyHostToDevice, s));
    HIPCHK(c->dZ.ensure(sizeof(double) * 2 * (size_t)Ne * Ne));
    st = fill_z(c, &o, plan, c->dFreq.as<Freq>(), c->dZ.as<double>(), (size_t)Ne);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(Z, c->dZ.p, sizeof(double) * 2 * (size_t)Ne * Ne, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ok();
}


// Frequencies factored together.  The latency-bound steps of the LU (pivot
// selection, panel solves, triangular-solve chains) cost per launch, not
// per matrix, so a batch should hold enough matrix elements to bury them:
// about kBatchElems (64 matrices at n = 6009, 256 below n ~ 3000), never
// fewer than 64.  A sweep that fits one batch needs one buffer of
// matrices; a longer one is pipelined over two (one filled while the other
// is factored), and its batches are split evenly (201 frequencies at
// n = 6009 -> 4 batches of 51).  Either way the buffers stay within 60% of
// the free device memory.
constexpr double kBatchElems = 2.4e9;

int batch_size(const efie_opts& o, size_t zElems, int Nf) {
    const size_t per = zElems * (o.precision == EFIE_PREC_MIXED ? 24 : 16);
    int B = o.batch;
    if (B <= 0) B = (int)std::min(256.0, std::max(64.0, kBatchElems / (double)std::max<size_t>(zElems, 1)));
    B = std::max(1, std::min(B, std::max(Nf, 1)));
    size_t freeB = 0, totB = 0;
    if (hipMemGetInfo(&freeB, &totB) != hipSuccess) return B;
    auto cap = [&](int nbuf) {
        return (size_t)(0.6 * (double)freeB) / std::max<size_t>(per * (size_t)nbuf, 1);
    };
    if (B >= Nf && cap(1) >= (size_t)B) return B;  // the whole sweep, one buffer
    if (cap(2) < (size_t)B) B = (int)std::max<size_t>(1, cap(2));
    if (Nf > 0 && o.batch <= 0) {
        const int nbatch = (Nf + B - 1) / B;
        B = (Nf + nbatch - 1) / nbatch;
    }
    return std::max(1, B);
}

// Host bytes of one batch's results: info and iterations (int), final
// relative residuals (double), the wait-limit flag (int).
constexpr size_t stat_slot_bytes(int B) { return (size_t)B * (2 * sizeof(int) + sizeof(double)) + 64; }

// Sizes of the batched solver's buffers: per buffer of matrices the complex
// elements (zBuf) and the batch-element-by-row vectors (vec, the largest
// B n); the residual's partial sums (pBytes), the pivot tree nodes of all
// matrices, the bf16 planes, the triangular-solve flags and the most
// matrices in a batch (B).  A batch of B matrices of order n (leading
// dimension lda) needs solver_dims(n, lda, B); a sequence of batches of
// different orders the largest of each.
struct SolverDims {
    size_t zBuf = 0, vec = 0, pBytes = 0, nodes = 0, split = 0, flags = 0;
    int B = 0;
    void grow(const SolverDims& d) {
        zBuf = std::max(zBuf, d.zBuf);
        vec = std::max(vec, d.vec);
        pBytes = std::max(pBytes, d.pBytes);
        nodes = std::max(nodes, d.nodes);
        split = std::max(split, d.split);
        flags = std::max(flags, d.flags);
        B = std::max(B, d.B);
    }
};

SolverDims solver_dims(int n, int lda, int B) {
    SolverDims d;
    d.zBuf = (size_t)n * lda * B;
    d.vec = (size_t)n * B;
    d.pBytes = residual_sym_bytes(n, B);
    d.nodes = (size_t)lu_tree_nodes(n) * B;
    d.split = lu_split_bytes(n) * B;
    d.flags = (size_t)B * ((n + 63) / 64);
    d.B = B;
    return d;
}

// Device buffers of the batched solver; nbuf buffers of matrices (and of
// their permutations, first-zero-pivot flags and ||V||^2).
efie_status ensure_solver_dims(efie_ctx* c, const efie_opts& o, const SolverDims& d, int nbuf) {
    hipStream_t s = c->stream;
    const int B = d.B;
    HIPCHK(c->dZ.ensure(sizeof(double) * 2 * d.zBuf * nbuf));
    if (o.precision == EFIE_PREC_MIXED) {
        HIPCHK(c->dZ32.ensure(sizeof(float) * 2 * d.zBuf * nbuf));
        HIPCHK(c->dPerm.ensure(sizeof(int) * d.vec * nbuf));
        HIPCHK(c->dX.ensure(sizeof(double) * 2 * d.vec));
        HIPCHK(c->dR.ensure(sizeof(double) * 2 * d.vec));
        HIPCHK(c->dR32.ensure(sizeof(float) * 2 * d.vec));
        HIPCHK(c->dP.ensure(d.pBytes));
        HIPCHK(c->dRn2.ensure(sizeof(double) * B));
        HIPCHK(c->dBn2.ensure(sizeof(double) * B * nbuf));
        HIPCHK(c->dDone.ensure(sizeof(int) * B));
        HIPCHK(c->dIters.ensure(sizeof(int) * B));
        HIPCHK(c->dResid.ensure(sizeof(double) * B));
        if (o.lu_kernel == EFIE_LU_BF16X6) HIPCHK(c->dSplit.ensure(d.split));
    }
    HIPCHK(c->dCand.ensure(sizeof(int) * d.nodes * lu_cand_stride()));
    const size_t cntBytes = sizeof(int) * d.nodes;
    if (c->dCnt.bytes < cntBytes || !c->dCnt.p) {
        HIPCHK(c->dCnt.ensure(cntBytes));
        HIPCHK(launch_fill_int(c->dCnt.as<int>(), 0, (int)(c->dCnt.bytes / sizeof(int)), s));
    }
    HIPCHK(c->dIpiv.ensure(sizeof(int) * d.vec));
    HIPCHK(c->dInfo.ensure(sizeof(int) * B * nbuf));
    const size_t statBytes = 2 * stat_slot_bytes(B);
    if (c->hStatBytes < statBytes) {
        if (c->hStat) HIPCHK(hipHostFree(c->hStat));
        c->hStat = nullptr;
        c->hStatBytes = 0;
        HIPCHK(hipHostMalloc(&c->hStat, statBytes));
        c->hStatBytes = statBytes;
    }
    HIPCHK(c->dSwp.ensure(sizeof(int) * (size_t)B * lu_swap_stride()));
    // single-launch triangular solve: flags hold the epoch of the last
    // launch that finished each block, so they are zeroed only when new
    const size_t flagBytes = sizeof(int) * d.flags;
    if (c->dFlags.bytes < flagBytes || !c->dFlags.p) {
        HIPCHK(c->dFlags.ensure(flagBytes));
        HIPCHK(launch_fill_int(c->dFlags.as<int>(), 0, (int)(c->dFlags.bytes / sizeof(int)), s));
        c->tw.epoch = 0;
    }
    HIPCHK(c->dTickets.ensure(sizeof(int) * B));
    if (!c->dErr.p) {
        HIPCHK(c->dErr.ensure(sizeof(int)));
        HIPCHK(launch_fill_int(c->dErr.as<int>(), 0, 1, s));
    }
    c->tw.flags = c->dFlags.as<int>();
    c->tw.tickets = c->dTickets.as<int>();
    c->tw.err = c->dErr.as<int>();
    return ok();
}

// Device buffers of the batched solver for n unknowns and B matrices.
efie_status ensure_solver(efie_ctx* c, const efie_opts& o, int n, int lda, int B, int nbuf = 1) {
    return ensure_solver_dims(c, o, solver_dims(n, lda, B), nbuf);
}

// Device work buffers of the batched solver besides the matrices: pivot
// tournament (candidates, arrival counters, pivots, swap lists), the
// split-bf16 planes, refinement vectors and per-matrix state, the
// triangular solve's flags.
struct SolverWork {
    int *cand, *cnt, *ipiv, *swp, *err;
    void* split;
    double *X, *R, *P, *Rn2, *Bn2, *resid;
    float* R32;
    int *done, *iters;
    TrsvWork* tw;
};

SolverWork solver_work(efie_ctx* c) {
    SolverWork w;
    w.cand = c->dCand.as<int>();
    w.cnt = c->dCnt.as<int>();
    w.ipiv = c->dIpiv.as<int>();
    w.swp = c->dSwp.as<int>();
    w.err = c->dErr.as<int>();
    w.split = c->dSplit.p;
    w.X = c->dX.as<double>();
    w.R = c->dR.as<double>();
    w.P = c->dP.as<double>();
    w.Rn2 = c->dRn2.as<double>();
    w.Bn2 = c->dBn2.as<double>();
    w.resid = c->dResid.as<double>();
    w.R32 = c->dR32.as<float>();
    w.done = c->dDone.as<int>();
    w.iters = c->dIters.as<int>();
    w.tw = &c->tw;
    return w;
}

// Buffers of one batch: matrices (row-major, rhs in column n), the FP32
// copy (mixed precision), first-zero-pivot flags, row permutations, the
// solver's work buffers and the host slot its results are copied to.
struct BatchBufs {
    double* Z;
    float* Z32;
    int* info;
    int* perm;
    SolverWork w;
    char* host;
};

// Buffer `buf` of a solver sized by d.
BatchBufs batch_bufs_dims(efie_ctx* c, const efie_opts& o, const SolverDims& d, int buf) {
    BatchBufs bb;
    bb.Z = c->dZ.as<double>() + 2 * d.zBuf * buf;
    bb.Z32 = o.precision == EFIE_PREC_MIXED ? c->dZ32.as<float>() + 2 * d.zBuf * buf : nullptr;
    bb.info = c->dInfo.as<int>() + (size_t)d.B * buf;
    bb.perm = o.precision == EFIE_PREC_MIXED ? c->dPerm.as<int>() + d.vec * buf : nullptr;
    bb.w = solver_work(c);
    bb.host = static_cast<char*>(c->hStat) + stat_slot_bytes(d.B) * buf;
    return bb;
}

BatchBufs batch_bufs(efie_ctx* c, const efie_opts& o, int n, int lda, int B, int buf) {
    return batch_bufs_dims(c, o, solver_dims(n, lda, B), buf);
}

// LU of the nb matrices of a batch on the context stream (look-ahead
// panels on the panel stream).  FP64: in place; MIXED: of the FP32 copy
// (converted here unless the fill wrote it), and the row permutation of
// its pivots for the refinement.
// nElem (packed batches, symmetric factorisation): the order of each
// matrix's problem, the identity padding beyond it skipped.
efie_status factor_batch(efie_ctx* c, const efie_opts& o, int n, int lda, int nb, const Prof& pf, bool z32Ready,
                         const BatchBufs& bb, const int* nElem = nullptr, hipStream_t head = nullptr,
                         hipEvent_t headWait = nullptr, hipEvent_t tailMark = nullptr) {
    hipStream_t s = c->stream;
    const size_t zElems = (size_t)n * lda;
    // look-ahead on the panel stream, except when per-kernel profiling
    // needs every kernel on one stream
    LuWork wk{bb.w.cand, bb.w.cnt, bb.w.ipiv, bb.info,
              bb.w.swp, c->prof ? nullptr : c->panelStream, c->evPanel, c->evAhead,
              (o.precision == EFIE_PREC_MIXED && o.lu_kernel == EFIE_LU_BF16X6) ? bb.w.split : nullptr,
              o.precision == EFIE_PREC_MIXED && o.lu_kernel == EFIE_LU_SYM};
    if (wk.noPivot) {
        wk.nElem = nElem;
        wk.head = head;
        wk.headWait = head ? headWait : nullptr;
        wk.tailMark = head ? tailMark : nullptr;
    }
    HIPCHK(launch_fill_int(bb.info, 0, nb, s));
    if (o.precision != EFIE_PREC_MIXED) {
        HIPCHK(lu_factor<double>(bb.Z, 2 * zElems, n, lda, nb, wk, s, &pf));
        return ok();
    }
    if (!z32Ready) {
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_to_f32(bb.Z, bb.Z32, zElems * nb, s));
        pf.end(EFIE_K_OTHER);
    }
    HIPCHK(lu_factor<float>(bb.Z32, 2 * zElems, n, lda, nb, wk, s, &pf));
    pf.begin(EFIE_K_OTHER);
    if (wk.noPivot)
        HIPCHK(lu_ident_perm(n, nb, bb.perm, s));
    else
        HIPCHK(lu_make_perm(bb.w.ipiv, n, nb, bb.perm, s));
    pf.end(EFIE_K_OTHER);
    return ok();
}

// Solution of a factored batch on stream si.  FP64: back substitution on
// column n.  MIXED: back substitution in FP32 (symmetric factorisation:
// x = 0), then FP64 iterative refinement against the FP64 matrix with
// right-hand sides V + b*vStride; symZ (the matrix is complex symmetric
// bit-exact) lets the residual read only its upper triangle.  Returns where
// the FP64 solutions are.
efie_status refine_batch(efie_ctx* c, const efie_opts& o, int n, int lda, int nb, const double* V, size_t vStride,
                         bool symZ, const Prof& pf, const BatchBufs& bb, hipStream_t si, const double** X,
                         size_t* xStride, size_t* xinc) {
    const size_t zElems = (size_t)n * lda;
    const SolverWork& w = bb.w;
    if (o.precision != EFIE_PREC_MIXED) {
        pf.begin(EFIE_K_TRSV);
        HIPCHK(lu_backsolve<double>(bb.Z, 2 * zElems, n, lda, nb, si, w.tw));
        pf.end(EFIE_K_TRSV);
        *X = bb.Z + 2 * n;
        *xStride = 2 * zElems;
        *xinc = 2 * (size_t)lda;
        return ok();
    }
    float* Z32 = bb.Z32;
    if (o.lu_kernel == EFIE_LU_SYM) {
        // no rhs column: the first refinement step from x = 0 is the solve
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_fill_dbl(w.X, 0.0, 2 * n * nb, si));
        pf.end(EFIE_K_OTHER);
    } else {
        pf.begin(EFIE_K_TRSV);
        HIPCHK(lu_backsolve<float>(Z32, 2 * zElems, n, lda, nb, si, w.tw));
        pf.end(EFIE_K_TRSV);
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_col_to_vec(Z32, 2 * zElems, n, lda, w.X, nb, si));
        pf.end(EFIE_K_OTHER);
    }
    pf.begin(EFIE_K_OTHER);
    HIPCHK(launch_fill_int(w.done, 0, nb, si));
    HIPCHK(launch_fill_int(w.iters, 0, nb, si));
    HIPCHK(launch_fill_dbl(w.Rn2, 0.0, nb, si));
    HIPCHK(launch_fill_dbl(w.resid, 1.0, nb, si));  // never a stale value from an older call
    pf.end(EFIE_K_OTHER);
    for (int it = 0; it <= o.ir_max_iter; ++it) {
        pf.begin(EFIE_K_OTHER);
        if (symZ)
            HIPCHK(launch_residual_sym(bb.Z, 2 * zElems, n, lda, V, vStride, w.X, w.R, w.P, w.Rn2, w.done, nb,
                                       it == 0 && o.lu_kernel == EFIE_LU_SYM, w.tw->rows > 0, si));
        else
            HIPCHK(launch_residual(bb.Z, 2 * zElems, n, lda, V, vStride, w.X, w.R,
                                   w.Rn2, w.done, nb, si));
        HIPCHK(launch_ir_check(w.Rn2, w.Bn2, o.ir_tol, nb, w.done,
                               w.iters, w.resid, it == 0, it == o.ir_max_iter,
                               o.lu_kernel == EFIE_LU_SYM ? o.ir_max_iter - it : 0, si));
        pf.end(EFIE_K_OTHER);
        if (it == o.ir_max_iter) break;
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_perm_f32(w.R, bb.perm, n, nb, w.R32, w.done, si));
        pf.end(EFIE_K_OTHER);
        pf.begin(EFIE_K_TRSV);
        HIPCHK(lu_trsv<float>(Z32, 2 * zElems, n, lda, w.R32, 2 * (size_t)n, 1, nb, true,
                              w.done, si, w.tw));
        HIPCHK(lu_trsv<float>(Z32, 2 * zElems, n, lda, w.R32, 2 * (size_t)n, 1, nb, false,
                              w.done, si, w.tw));
        pf.end(EFIE_K_TRSV);
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_add_f32(w.X, w.R32, n, nb, w.done, si));
        pf.end(EFIE_K_OTHER);
    }
    *X = w.X;
    *xStride = 2 * (size_t)n;
    *xinc = 2;
    return ok();
}

// Queue the copy of a batch's results into host slot `slot` (pinned, so
// the copy is asynchronous) on stream si.
efie_status stage_stats(const efie_opts& o, int nb, int B, const BatchBufs& bb, hipStream_t si) {
    char* h = bb.host;
    int* info = reinterpret_cast<int*>(h);
    int* iters = info + B;
    double* resid = reinterpret_cast<double*>(h + 2 * sizeof(int) * (size_t)B);
    int* werr = reinterpret_cast<int*>(h + (2 * sizeof(int) + sizeof(double)) * (size_t)B);
    HIPCHK(hipMemcpyAsync(info, bb.info, sizeof(int) * nb, hipMemcpyDeviceToHost, si));
    HIPCHK(hipMemcpyAsync(werr, bb.w.err, sizeof(int), hipMemcpyDeviceToHost, si));
    if (o.precision == EFIE_PREC_MIXED) {
        HIPCHK(hipMemcpyAsync(iters, bb.w.iters, sizeof(int) * nb, hipMemcpyDeviceToHost, si));
        HIPCHK(hipMemcpyAsync(resid, bb.w.resid, sizeof(double) * nb, hipMemcpyDeviceToHost, si));
    }
    return ok();
}

// After a batch's slot event has completed: zero pivots become
// efie:singular, refinement counters go into the stats, profile marks are
// summed.
// retry (mixed precision): systems whose FP32 factorisation hit a zero
// pivot or whose refinement did not reach max(ir_tol, 1e-11) (1e-11: above
// the FP64 residual floor of these systems) are appended (as offsets into
// the batch) instead of failing; the caller solves them again, the
// unpivoted symmetric form with the pivoted FP32 LU, the pivoted forms with
// the FP64 LU.
efie_status finish_batch(efie_ctx* c, const efie_opts& o, int nb, const BatchBufs& bb, int B, efie_stats* S,
                         const double* freqs, int* retry = nullptr, int* nretry = nullptr) {
    if (nretry) *nretry = 0;
    const char* h = bb.host;
    const int* info = reinterpret_cast<const int*>(h);
    const int* iters = info + B;
    const double* resid = reinterpret_cast<const double*>(h + 2 * sizeof(int) * (size_t)B);
    const int werr = *reinterpret_cast<const int*>(h + (2 * sizeof(int) + sizeof(double)) * (size_t)B);
    const bool mixed = o.precision == EFIE_PREC_MIXED;
    prof_collect(c, S);
    if (werr) {
        HIPCHK(launch_fill_int(bb.w.err, 0, 1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return fail(EFIE_E_HIP, "efie:hip", "triangular solve: a block wait exceeded its limit");
    }
    for (int i = 0; i < nb; ++i) {
        if (retry && mixed && (info[i] != 0 || !(resid[i] <= std::max(o.ir_tol, 1e-11)))) {
            retry[(*nretry)++] = i;
            continue;
        }
        if (info[i] != 0) {
            if (freqs)
                return fail(EFIE_E_SINGULAR, "efie:singular", "Z is singular at f=%g Hz (zero pivot in column %d)",
                            freqs[i], info[i]);
            return fail(EFIE_E_SINGULAR, "efie:singular", "matrix %d is singular (zero pivot in column %d)", i,
                        info[i]);
        }
        if (mixed) {
            S->ir_iters_max = std::max(S->ir_iters_max, iters[i]);
            S->ir_resid_max = std::max(S->ir_resid_max, resid[i]);
            S->ir_iters_total += iters[i];
        }
    }
    return ok();
}


// ------------------------------------------------------- efie_sweep_many --
// Frequencies f0..f0+nf-1 of item `item` are matrices e0..e0+nf-1 of a
// packed batch.
struct PackSeg {
    int item, f0, nf, e0;
};

// A batch of efie_sweep_many: nb matrices of order np (leading dimension
// lda), each holding one problem's matrix in its leading block.
struct PackBatch {
    int np = 0, lda = 0, nb = 0;
    std::vector<PackSeg> seg;
};

// Problems of order at least kPackRatio times a batch's order may share it
// (the padding's factorisation steps are skipped; its triangular solves and
// residual still read order np).
constexpr double kPackRatio = 0.75;
// Matrices per packed batch at most.
constexpr int kPackMaxBatch = 512;

// Cut the frequencies of the items, taken in `order` (largest order
// first), into batches.  A batch of order np holds about kBatchElems matrix
// elements (at least 64 matrices, at most kPackMaxBatch; opts.batch when
// set), two batches within 60 % of the free device memory.  An item with
// more frequencies left than one batch holds is split evenly; otherwise the
// batch takes the rest of it and then the next items while their order is
// at least kPackRatio np.
std::vector<PackBatch> plan_packs(const std::vector<int>& order, const std::vector<int>& nOf,
                                  const std::vector<int>& nfOf, const efie_opts& o, size_t freeB) {
    const double per = o.precision == EFIE_PREC_MIXED ? 24.0 : 16.0;
    std::vector<PackBatch> out;
    size_t oi = 0;
    int done = 0;  // frequencies of order[oi] already placed
    while (oi < order.size()) {
        const int g = order[oi];
        PackBatch b;
        b.np = nOf[g];
        b.lda = ((b.np + 1) + 31) & ~31;  // rhs in column np
        const double zE = (double)b.np * b.lda;
        double cap = o.batch > 0 ? o.batch : std::min<double>(kPackMaxBatch, std::max(64.0, kBatchElems / zE));
        cap = std::min(cap, std::max(1.0, std::floor(0.6 * (double)freeB / (2.0 * per * zE))));
        const int icap = std::max(1, (int)cap);
        const int rem = nfOf[g] - done;
        if (rem > icap) {
            const int parts = (rem + icap - 1) / icap;
            const int take = o.batch > 0 ? icap : (rem + parts - 1) / parts;
            b.seg.push_back({g, done, take, 0});
            b.nb = take;
            done += take;
            out.push_back(std::move(b));
            continue;
        }
        while (oi < order.size() && b.nb < icap) {
            const int h = order[oi];
            if (b.nb > 0 && (double)nOf[h] < kPackRatio * b.np) break;
            const int take = std::min(nfOf[h] - done, icap - b.nb);
            b.seg.push_back({h, done, take, b.nb});
            b.nb += take;
            done += take;
            if (done == nfOf[h]) {
                ++oi;
                done = 0;
            }
        }
        out.push_back(std::move(b));
    }
    return out;
}

}  // namespace

extern "C" {

efie_status efie_assemble_zv(efie_ctx* c, double k_re, double k_im, double omega, double mu,
                             double eps_re, double eps_im, const double* exc, int n_exc,
                             const efie_opts* opts, double* Z, double* Vrhs) {
    if (!c || !exc || !Z || !Vrhs) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (n_exc < 5)
        return fail(EFIE_E_ARGS, "mex:exc",
                    "exc must be double vector len>=5: [patchEdge, groundEdge, V0, spreadRadius, selfAlpha].");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    const int pe = (int)exc[0], ge = (int)exc[1];
    if (!c->haveRwg) return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
    if (pe < 1 || pe > c->Ne) return fail(EFIE_E_PORT, "mex:port", "patchEdge out of range.");
    if (ge < 1 || ge > c->Ne) return fail(EFIE_E_PORT, "mex:port", "groundEdge out of range.");
    efie_status st = prepare(c, &o, exc[4]);
    if (st) return st;
    c->prof = false;
    st = make_rhs(c, pe, ge, exc[2], exc[3]);
    if (st) return st;
    st = fill_single(c, o, k_re, k_im, omega, mu, eps_re, eps_im, Z);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(Vrhs, c->dV.p, sizeof(double) * 2 * (size_t)c->Ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ok();
}

efie_status efie_assemble_z(efie_ctx* c, double k_re, double k_im, double omega, double mu, double eps_re,
                            double eps_im, const efie_opts* opts, double* Z) {
    if (!c || !Z) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    if (!c->haveRwg) return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
    if (o.model != EFIE_MODEL_QUADRATURE)
        return fail(EFIE_E_ARGS, "efie:opts", "efie_assemble_z computes the quadrature model only");
    efie_status st = prepare(c, &o, 0.20);
    if (st) return st;
    c->prof = false;
    return fill_single(c, o, k_re, k_im, omega, mu, eps_re, eps_im, Z);
}

efie_status efie_sweep(efie_ctx* c, const double* freqs, int Nf, double mu, double eps_re,
                       double eps_im, int patchEdge, int groundEdge, double V0, double spread,
                       double Z0, const efie_opts* opts, double* Zin, double* S11,
                       efie_stats* stats) {
    if (!c || (!freqs && Nf > 0) || (!Zin && Nf > 0) || (!S11 && Nf > 0))
        return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (Nf < 0) return fail(EFIE_E_ARGS, "mex:freqs", "freqs must be double vector.");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    if (!c->haveRwg) return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
    if (patchEdge < 1 || patchEdge > c->Ne) return fail(EFIE_E_PORT, "mex:port", "patchEdge out of range.");
    if (groundEdge < 1 || groundEdge > c->Ne) return fail(EFIE_E_PORT, "mex:port", "groundEdge out of range.");
    // Z is complex symmetric: the mixed path factors it without pivoting
    // unless asked otherwise (frequencies that need pivoting are redone)
    if (o.lu_kernel == EFIE_LU_AUTO && o.precision == EFIE_PREC_MIXED) o.lu_kernel = EFIE_LU_SYM;
    // the centroid sweep MEX hard-codes self_alpha = 0.20 (patch_sweep_cuda.cu:269)
    efie_status st = prepare(c, &o, 0.20);
    if (st) return st;
    efie_stats local;
    efie_stats* S = stats ? stats : &local;
    std::memset(S, 0, sizeof *S);
    c->prof = o.profile != 0;
    c->marks.clear();
    c->poolUsed = 0;
    st = make_rhs(c, patchEdge, groundEdge, V0, spread);
    if (st) return st;
    if (Nf == 0) return ok();

    std::complex<double> eps(eps_re, eps_im);
    std::vector<std::complex<double>> ks(Nf);
    double kmax = 0;
    for (int f = 0; f < Nf; ++f) {
        const double omega = 2.0 * kPi * freqs[f];
        ks[f] = omega * std::sqrt(std::complex<double>(mu, 0.0) * eps);
        kmax = std::max(kmax, std::abs(ks[f]));
    }
    bool lossy = false;
    for (int f = 0; f < Nf; ++f) lossy = lossy || ks[f].imag() != 0.0;
    NearPlan plan;
    st = choose_near(c, &o, kmax, lossy, S, &plan);
    if (st) return st;

    const int n = c->Ne;
    const int lda = ((n + 1) + 31) & ~31;  // rhs in column n
    const size_t zElems = (size_t)n * lda;
    const int B = batch_size(o, zElems, Nf);
    const int nBatch = (Nf + B - 1) / B;
    // batches are pipelined over two streams (fill of the next batch while
    // this one is factored) unless per-kernel profiling asks for one stream
    const bool overlap = nBatch > 1 && !c->prof;
    const int nbuf = overlap ? 2 : 1;
    hipStream_t s = c->stream;
    hipStream_t sf = overlap ? c->fillStream : s;
    st = ensure_solver(c, o, n, lda, B, nbuf);
    if (st) return st;
    HIPCHK(c->dFreq.ensure(sizeof(Freq) * B * nbuf));
    HIPCHK(c->dOut.ensure(sizeof(double) * 4 * (size_t)Nf));
    Prof pf = make_prof(c);
    if (o.precision == EFIE_PREC_MIXED) {
        // ||V||^2 of the shared right-hand side, for every batch element
        std::vector<double> hv(2 * (size_t)n);
        HIPCHK(hipMemcpyAsync(hv.data(), c->dV.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double vn2 = 0;
        for (double v : hv) vn2 += v * v;
        HIPCHK(launch_fill_dbl(c->dBn2.as<double>(), vn2, B, s));
    }

    std::vector<Freq> fr[2] = {std::vector<Freq>(B), std::vector<Freq>(B)};
    bool have32[2] = {false, false};
    const bool mixed = o.precision == EFIE_PREC_MIXED;
    // the symmetric mixed solve reads the FP64 matrix's upper triangle (and
    // diagonal blocks) and the FP32 copy's lower one: the series fill writes
    // only those
    const bool triFill = mixed && o.lu_kernel == EFIE_LU_SYM && o.model != EFIE_MODEL_CENTROID;
    auto zbuf = [&](int buf) { return c->dZ.as<double>() + 2 * zElems * (size_t)B * buf; };
    auto z32buf = [&](int buf) { return mixed ? c->dZ32.as<float>() + 2 * zElems * (size_t)B * buf : nullptr; };
    // fill batch bi into buffer bi % nbuf on stream sf (Z, its complex64
    // copy and the rhs column); evFill[buf] marks its end
    auto enqueue_fill = [&](int bi) -> efie_status {
        const int f0 = bi * B, nb = std::min(B, Nf - f0), buf = bi % nbuf;
        for (int i = 0; i < nb; ++i) fr[buf][i] = make_freq(ks[f0 + i], 2.0 * kPi * freqs[f0 + i], mu, eps);
        Freq* dfr = c->dFreq.as<Freq>() + (size_t)B * buf;
        HIPCHK(hipEventRecord(c->evFillStart[buf], sf));
        HIPCHK(hipMemcpyAsync(dfr, fr[buf].data(), sizeof(Freq) * nb, hipMemcpyHostToDevice, sf));
        double* Zb = zbuf(buf);
        float* Z32b = z32buf(buf);
        bool h32 = Z32b != nullptr;
        if (plan.series) {
            // every frequency of the batch in one pass over the series
            Prof pff = make_prof(c);
            pff.begin(EFIE_K_SERIES);
            // every fill after the first runs beside a factorisation: a
            // capped grid leaves that the CUs
            HIPCHK(launch_series_fill(c->dSerC.as<double>(), plan.serNT, plan.lossy, n, dfr, nb, Zb, (size_t)lda,
                                      2 * zElems, Z32b, (size_t)lda, 2 * zElems, sf,
                                      overlap && bi > 0 ? kFillBlocksShared : 0, triFill));
            pff.end(EFIE_K_SERIES);
        } else {
            for (int i = 0; i < nb; ++i) {
                bool w32 = false;
                efie_status e = fill_z(c, &o, plan, dfr + i, Zb + 2 * zElems * i, (size_t)lda,
                                       Z32b ? Z32b + 2 * zElems * i : nullptr, &w32, sf);
                if (e) return e;
                h32 = h32 && w32;
            }
        }
        // rhs into column n (of the complex64 copy too when the fill wrote it)
        HIPCHK(launch_set_rhs(c->dV.as<double>(), Zb, h32 ? Z32b : nullptr, n, lda, zElems, nb, sf));
        HIPCHK(hipEventRecord(c->evFill[buf], sf));
        have32[buf] = h32;
        return ok();
    };

    // Pipeline over batches (overlap): fill on sf, LU on s (panels on the
    // panel stream), back substitution / refinement / Zin on si.  Batch bi
    // uses buffer bi % 2.  Once batch bi is queued, the host waits for the
    // refinement of batch bi-1 (which ran beside the LU of batch bi), reads
    // its results and queues the fill of batch bi+1 into its buffer.
    hipStream_t si = overlap ? c->irStream : s;
    double msFill = 0, msFac = 0, msSolve = 0;
    HIPCHK(hipEventRecord(c->ev[0], s));
    if (overlap) {
        HIPCHK(hipStreamWaitEvent(sf, c->ev[0], 0));
        HIPCHK(hipStreamWaitEvent(si, c->ev[0], 0));
    }
    // mixed precision: systems the FP32 factorisation plus refinement did
    // not solve are redone after the sweep
    const bool symRetry = mixed;
    std::vector<int> retry;
    auto finish = [&](int bj) -> efie_status {
        const int f0 = bj * B, nb = std::min(B, Nf - f0), buf = bj % nbuf;
        HIPCHK(hipEventSynchronize(c->evIR[buf]));
        float t1, t2, t3;
        HIPCHK(hipEventElapsedTime(&t1, c->evFillStart[buf], c->evFill[buf]));
        HIPCHK(hipEventElapsedTime(&t2, c->evLUStart[buf], c->evLU[buf]));
        HIPCHK(hipEventElapsedTime(&t3, c->evIRStart[buf], c->evIR[buf]));
        msFill += t1;
        msFac += t2;
        msSolve += t3;
        const BatchBufs bbj = batch_bufs(c, o, n, lda, B, buf);
        if (!symRetry) return finish_batch(c, o, nb, bbj, B, S, freqs + f0);
        std::vector<int> rb(nb);
        int nrb = 0;
        efie_status e = finish_batch(c, o, nb, bbj, B, S, freqs + f0, rb.data(), &nrb);
        for (int i = 0; i < nrb; ++i) retry.push_back(f0 + rb[i]);
        return e;
    };
    st = enqueue_fill(0);
    if (st) return st;
    for (int bi = 0; bi < nBatch; ++bi) {
        const int f0 = bi * B, nb = std::min(B, Nf - f0), buf = bi % nbuf;
        const BatchBufs bb = batch_bufs(c, o, n, lda, B, buf);
        HIPCHK(hipStreamWaitEvent(s, c->evFill[buf], 0));
        HIPCHK(hipEventRecord(c->evLUStart[buf], s));
        // the first panels of a batch on the fill stream, behind its fill:
        // beside the previous batch's last factorisation steps
        st = factor_batch(c, o, n, lda, nb, pf, have32[buf], bb, nullptr, overlap ? sf : nullptr,
                          bi > 0 ? c->evTail[(bi - 1) & 1] : nullptr, c->evTail[bi & 1]);
        if (st) return st;
        HIPCHK(hipEventRecord(c->evLU[buf], s));
        HIPCHK(hipStreamWaitEvent(si, c->evLU[buf], 0));
        HIPCHK(hipEventRecord(c->evIRStart[buf], si));
        const double* X = nullptr;
        size_t xStride = 0, xinc = 0;
        // beside the next batch's factorisation the triangular solves run
        // with a few blocks per matrix (the update keeps its CUs: sweep
        // 309 -> 316 freq/s at c2_6k); alone, with one per row block
        c->tw.rows = overlap && bi + 1 < nBatch ? kTrsvRowsShared : 0;
        st = refine_batch(c, o, n, lda, nb, c->dV.as<double>(), 0, o.model != EFIE_MODEL_CENTROID, pf, bb, si, &X,
                           &xStride, &xinc);
        c->tw.rows = 0;
        if (st) return st;
        pf.begin(EFIE_K_OTHER);
        HIPCHK(launch_epilogue(X, xStride, xinc, c->dWl.as<double>(), n, patchEdge - 1, groundEdge - 1,
                               o.model == EFIE_MODEL_CENTROID, V0, Z0, nb, c->dOut.as<double>() + 4 * (size_t)f0,
                               si));
        pf.end(EFIE_K_OTHER);
        st = stage_stats(o, nb, B, bb, si);
        if (st) return st;
        HIPCHK(hipEventRecord(c->evIR[buf], si));
        if (overlap) {
            // batch bi-1's results (its refinement ran beside this LU),
            // then the next fill into its buffer
            if (bi >= 1) {
                st = finish(bi - 1);
                if (st) return st;
            }
            if (bi + 1 < nBatch) {
                if (bi >= 1) HIPCHK(hipStreamWaitEvent(sf, c->evIR[(bi + 1) % nbuf], 0));
                st = enqueue_fill(bi + 1);
                if (st) return st;
            }
        } else {
            st = finish(bi);
            if (st) return st;
            if (bi + 1 < nBatch) {
                st = enqueue_fill(bi + 1);
                if (st) return st;
            }
        }
    }
    if (overlap) {
        st = finish(nBatch - 1);
        if (st) return st;
        // the context stream carries the remaining work of the call
        HIPCHK(hipEventRecord(c->ev[1], si));
        HIPCHK(hipStreamWaitEvent(s, c->ev[1], 0));
    }
    HIPCHK(hipEventRecord(c->ev[5], s));
    std::vector<double> out(4 * (size_t)Nf);
    HIPCHK(hipMemcpyAsync(out.data(), c->dOut.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int f = 0; f < Nf; ++f) {
        Zin[2 * f] = out[4 * f];
        Zin[2 * f + 1] = out[4 * f + 1];
        S11[2 * f] = out[4 * f + 2];
        S11[2 * f + 1] = out[4 * f + 3];
    }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[5]));
    S->ms_total = tot;
    S->ms_fill = msFill;
    S->ms_factor = msFac;
    S->ms_solve = msSolve;
    if (!retry.empty()) {
        // the frequencies the mixed solve could not finish, again in a sweep
        // of their own: after the unpivoted symmetric factorisation with the
        // pivoted FP32 LU (whose own failures go on to FP64), after a
        // pivoted FP32 LU with the FP64 LU
        const int k = (int)retry.size();
        std::vector<double> fr2(k), zin2(2 * (size_t)k), s112(2 * (size_t)k);
        for (int i = 0; i < k; ++i) fr2[i] = freqs[retry[i]];
        efie_opts o2 = o;
        const bool toPivoted = o.lu_kernel == EFIE_LU_SYM;
        if (toPivoted) o2.lu_kernel = EFIE_LU_FP32;
        else o2.precision = EFIE_PREC_FP64;
        efie_stats S2;
        st = efie_sweep(c, fr2.data(), k, mu, eps_re, eps_im, patchEdge, groundEdge, V0, spread, Z0, &o2,
                        zin2.data(), s112.data(), &S2);
        if (st) return st;
        for (int i = 0; i < k; ++i) {
            Zin[2 * retry[i]] = zin2[2 * i];
            Zin[2 * retry[i] + 1] = zin2[2 * i + 1];
            S11[2 * retry[i]] = s112[2 * i];
            S11[2 * retry[i] + 1] = s112[2 * i + 1];
        }
        S->ms_total += S2.ms_total;
        S->ms_fill += S2.ms_fill;
        S->ms_factor += S2.ms_factor;
        S->ms_solve += S2.ms_solve;
        S->ir_iters_max = std::max(S->ir_iters_max, S2.ir_iters_max);
        S->ir_iters_total += S2.ir_iters_total;
        S->ir_resid_max = std::max(S->ir_resid_max, S2.ir_resid_max);
        for (int q = 0; q < EFIE_NKCLASS; ++q) {
            S->ms_kernel[q] += S2.ms_kernel[q];
            S->n_kernel[q] += S2.n_kernel[q];
        }
        if (toPivoted) {
            S->n_lu_retry = k;
            S->n_fp64_retry = S2.n_fp64_retry;
        } else {
            S->n_fp64_retry = k;
        }
    }
    S->n_near_pairs = c->nNear;
    if (o.model == EFIE_MODEL_QUADRATURE) {
        const long long Nt = c->Nt, np = (long long)c->NpNear, ns = (long long)c->NsSelf * c->Nq2s;
        const long long reg = Nt * (Nt - 1) / 2 - c->nNear;
        const long long nearDirect = plan.momNear ? 0 : plan.sur ? c->nSurFb : c->nNear;
        S->n_green = plan.series ? 0
                                 : (reg + (plan.sur ? c->nSurIn : 0)) * c->Nq * c->Nq + nearDirect * np * np +
                                       (plan.momSelf ? 0 : Nt * ns * ns);
    }
    c->prof = false;
    return ok();
}

efie_status efie_sweep_many(efie_ctx* c, const efie_sweep_item* items, int count, double mu, double eps_re,
                            double eps_im, const efie_opts* opts, efie_stats* stats) {
    if (!c || count < 0 || (!items && count > 0)) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    efie_status st = check_opts(&o);
    if (st) return st;
    // Z is complex symmetric: the mixed path factors it without pivoting
    // unless asked otherwise (as efie_sweep)
    if (o.lu_kernel == EFIE_LU_AUTO && o.precision == EFIE_PREC_MIXED) o.lu_kernel = EFIE_LU_SYM;
    for (int g = 0; g < count; ++g) {
        const efie_sweep_item& it = items[g];
        if (!it.problem) return fail(EFIE_E_ARGS, "efie:args", "item %d: null problem context", g);
        if (it.problem->device != c->device)
            return fail(EFIE_E_ARGS, "efie:args", "item %d: problem on device %d, engine on device %d", g,
                        it.problem->device, c->device);
        if (it.Nf < 0) return fail(EFIE_E_ARGS, "mex:freqs", "freqs must be double vector.");
        if (it.Nf > 0 && (!it.freqs || !it.Zin || !it.S11))
            return fail(EFIE_E_ARGS, "efie:args", "item %d: null argument", g);
        if (!it.problem->haveRwg) return fail(EFIE_E_STATE, "efie:state", "mesh and RWG tables must be set first");
        if (it.patchEdge < 1 || it.patchEdge > it.problem->Ne)
            return fail(EFIE_E_PORT, "mex:port", "patchEdge out of range.");
        if (it.groundEdge < 1 || it.groundEdge > it.problem->Ne)
            return fail(EFIE_E_PORT, "mex:port", "groundEdge out of range.");
    }
    HIPCHK(hipSetDevice(c->device));
    efie_stats local;
    efie_stats* S = stats ? stats : &local;
    std::memset(S, 0, sizeof *S);
    const bool mixed = o.precision == EFIE_PREC_MIXED;
    // the unpivoted symmetric factorisation carries no rhs column, and of
    // the series fill it needs only the triangles it reads
    const bool rhsCol = !(mixed && o.lu_kernel == EFIE_LU_SYM);
    const bool triFill = !rhsCol && o.model != EFIE_MODEL_CENTROID;

    // wavenumbers; per problem context the largest |k| of its items (items
    // sharing a context share its series) and whether all its items
    // release the series
    const std::complex<double> eps(eps_re, eps_im);
    std::vector<std::vector<std::complex<double>>> ks(count);
    std::vector<int> nOf(count), nfOf(count), ctxOf(count);
    std::vector<efie_ctx*> ctxs;
    std::vector<double> kmaxCtx;
    std::vector<char> relCtx;
    bool lossy = false;
    for (int g = 0; g < count; ++g) {
        const efie_sweep_item& it = items[g];
        nOf[g] = it.problem->Ne;
        nfOf[g] = it.Nf;
        const auto pos = std::find(ctxs.begin(), ctxs.end(), it.problem);
        ctxOf[g] = (int)(pos - ctxs.begin());
        if (pos == ctxs.end()) {
            ctxs.push_back(it.problem);
            kmaxCtx.push_back(0.0);
            relCtx.push_back(1);
        }
        if (!it.release) relCtx[ctxOf[g]] = 0;
        ks[g].resize(it.Nf);
        for (int f = 0; f < it.Nf; ++f) {
            const double omega = 2.0 * kPi * it.freqs[f];
            ks[g][f] = omega * std::sqrt(std::complex<double>(mu, 0.0) * eps);
            kmaxCtx[ctxOf[g]] = std::max(kmaxCtx[ctxOf[g]], std::abs(ks[g][f]));
            lossy = lossy || ks[g][f].imag() != 0.0;
        }
    }
    const int nCtx = (int)ctxs.size();
    std::vector<int> order;
    for (int g = 0; g < count; ++g)
        if (nfOf[g] > 0) order.push_back(g);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nOf[a] > nOf[b]; });

    // One-off work of every item before the pipeline (all host round trips
    // happen here): geometry, excitation (its errors come before any
    // sweep work), the near plan.  Series that are not cached are only
    // sized here; they are integrated on the fill stream right before the
    // item's first batch, into the context's own buffer (kept) or a pool
    // slot (released items), so the pipeline holds no host wait and no
    // device free.
    std::vector<size_t> vOff(count), outOff(count);
    size_t vTot = 0, outTot = 0;
    for (int g = 0; g < count; ++g) {
        vOff[g] = vTot;
        vTot += 2 * (size_t)nOf[g];
        outOff[g] = outTot;
        outTot += (size_t)nfOf[g];
    }
    HIPCHK(c->dItemV.ensure(sizeof(double) * vTot));
    HIPCHK(c->dItemWl.ensure(sizeof(double) * vTot));
    std::vector<NearPlan> plans(count);
    std::vector<efie_opts> optsOf(count, o);
    std::vector<double> vn2(count, 0.0);
    // per context: series coefficients to read, whether to build them (and
    // with how many terms), built into the context's own buffer or a slot
    std::vector<const double*> coefC(nCtx, nullptr);
    std::vector<int> serNT(nCtx, 0), buildP(nCtx, -1), slotOf(nCtx, -1);
    std::vector<char> planned(nCtx, 0), freeAfter(nCtx, 0);
    std::vector<NearPlan> planCtx(nCtx);
    std::vector<double> alphaCtx(nCtx, 0.0);
    std::vector<size_t> cBytesOf(nCtx, 0);
    size_t qMax = 0, cMaxSlot = 0;
    for (int g = 0; g < count; ++g) {
        const efie_sweep_item& it = items[g];
        efie_ctx* p = it.problem;
        const int ci = ctxOf[g];
        st = prepare(p, &optsOf[g], 0.20);
        if (st) return st;
        st = make_rhs(p, it.patchEdge, it.groundEdge, it.V0, it.spread);
        if (st) return st;
        if (nfOf[g] == 0) continue;
        const size_t nv = 2 * (size_t)nOf[g];
        std::vector<double> hv(nv);
        HIPCHK(hipMemcpyAsync(c->dItemV.as<double>() + vOff[g], p->dV.p, sizeof(double) * nv,
                              hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(c->dItemWl.as<double>() + vOff[g], p->dWl.p, sizeof(double) * nv,
                              hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(hv.data(), p->dV.p, sizeof(double) * nv, hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        for (double v : hv) vn2[g] += v * v;
        const efie_opts& og = optsOf[g];
        if (planned[ci]) {
            // a second item of the context: the same plan (same options and |k|)
            plans[g] = planCtx[ci];
            continue;
        }
        planned[ci] = 1;
        alphaCtx[ci] = og.near_alpha;
        S->n_near_pairs += p->nNear;
        if (og.model == EFIE_MODEL_QUADRATURE && og.near_mode == EFIE_NEAR_MOMENTS) {
            const int P0 = series_terms(p, &og, kmaxCtx[ci]);
            const int NT = P0 >= 0 ? series_nt(P0) : 0;
            if (NT > 0) {
                NearPlan pl;
                pl.lossy = lossy;
                pl.series = true;
                if (p->serP >= P0 && p->serAlpha == og.near_alpha && p->dSerC.p) {
                    pl.serNT = p->serP + 3;
                    coefC[ci] = p->dSerC.as<double>();
                    freeAfter[ci] = relCtx[ci];
                } else {
                    pl.serNT = NT;
                    buildP[ci] = NT - 3;
                    cBytesOf[ci] = series_coeff_bytes(p->Ne, NT);
                    qMax = std::max(qMax, sizeof(double) * 8 * (size_t)(NT - 2) * rec_index(0, p->Nt));
                    if (relCtx[ci]) {
                        cMaxSlot = std::max(cMaxSlot, cBytesOf[ci]);
                    } else {
                        p->serP = -1;  // valid once built
                        HIPCHK(p->dSerC.ensure(cBytesOf[ci]));
                        coefC[ci] = p->dSerC.as<double>();
                    }
                }
                S->moment_terms = std::max(S->moment_terms, pl.serNT - 2);
                plans[g] = planCtx[ci] = pl;
                continue;
            }
        }
        const int mt = S->moment_terms;
        st = choose_near(p, &optsOf[g], kmaxCtx[ci], lossy, S, &plans[g]);
        if (st) return st;
        planCtx[ci] = plans[g];
        S->moment_terms = std::max(mt, S->moment_terms);
    }
    if (order.empty()) return ok();

    // the pair-series scratch of the builds (one at a time, on the fill
    // stream); like the engine's other buffers it is kept for later calls
    HIPCHK(c->dSerQ.ensure(std::max<size_t>(qMax, 16)));
    size_t freeB = 0, totB = 0;
    HIPCHK(hipMemGetInfo(&freeB, &totB));
    // the engine's matrix buffers are reused (grown if need be), so they
    // count as free for the batch plan
    freeB += c->dZ.bytes + c->dZ32.bytes;
    // room for a few pool slots before the batches take their share
    int nRelBuild = 0;
    for (int ci = 0; ci < nCtx; ++ci) nRelBuild += buildP[ci] >= 0 && relCtx[ci];
    size_t poolEst = 0;  // beyond the slots the engine holds from earlier calls
    for (int k = 0; k < std::min(nRelBuild, 8); ++k) {
        const size_t have = k < (int)c->dSerPool.size() ? c->dSerPool[k]->bytes : 0;
        poolEst += cMaxSlot > have ? cMaxSlot - have : 0;
    }
    const std::vector<PackBatch> batches =
        plan_packs(order, nOf, nfOf, o, freeB > poolEst ? freeB - poolEst : freeB / 2);
    const int nBatch = (int)batches.size();
    std::vector<int> firstCtx(nCtx, -1), lastCtx(nCtx, -1);
    for (int bi = 0; bi < nBatch; ++bi)
        for (const PackSeg& sg : batches[bi].seg) {
            const int ci = ctxOf[sg.item];
            if (firstCtx[ci] < 0) firstCtx[ci] = bi;
            lastCtx[ci] = bi;
        }
    // pool slots: a released context's slot is taken again by a context
    // first filled after its last batch (the build is queued on the fill
    // stream behind that batch's fill, which is the last reader)
    {
        std::vector<int> slotFreeFrom;  // batch from which the slot is free
        for (int bi = 0; bi < nBatch; ++bi)
            for (const PackSeg& sg : batches[bi].seg) {
                const int ci = ctxOf[sg.item];
                if (firstCtx[ci] != bi || buildP[ci] < 0 || !relCtx[ci] || slotOf[ci] >= 0) continue;
                int sl = -1;
                for (size_t k = 0; k < slotFreeFrom.size() && sl < 0; ++k)
                    if (slotFreeFrom[k] <= bi) sl = (int)k;
                if (sl < 0) {
                    sl = (int)slotFreeFrom.size();
                    slotFreeFrom.push_back(0);
                }
                slotFreeFrom[sl] = lastCtx[ci] + 1;
                slotOf[ci] = sl;
            }
        while (c->dSerPool.size() < slotFreeFrom.size()) c->dSerPool.emplace_back(new DBuf());
        for (size_t k = 0; k < slotFreeFrom.size(); ++k) HIPCHK(c->dSerPool[k]->ensure(cMaxSlot));
        for (int ci = 0; ci < nCtx; ++ci)
            if (slotOf[ci] >= 0) coefC[ci] = c->dSerPool[slotOf[ci]]->as<double>();
    }

    c->prof = o.profile != 0;
    c->marks.clear();
    c->poolUsed = 0;
    // batches are pipelined over two buffers unless per-kernel profiling
    // asks for one stream
    const bool overlap = nBatch > 1 && !c->prof;
    const int nbuf = overlap ? 2 : 1;
    hipStream_t s = c->stream;
    hipStream_t sf = overlap ? c->fillStream : s;
    hipStream_t si = overlap ? c->irStream : s;
    SolverDims dims;
    for (const PackBatch& b : batches) dims.grow(solver_dims(b.np, b.lda, b.nb));
    st = ensure_solver_dims(c, o, dims, nbuf);
    if (st) return st;
    HIPCHK(c->dFreq.ensure(sizeof(Freq) * dims.B * nbuf));
    HIPCHK(c->dNElem.ensure(sizeof(int) * dims.B * nbuf));
    HIPCHK(c->dVp.ensure(sizeof(double) * 2 * dims.vec * nbuf));
    HIPCHK(c->dOut.ensure(sizeof(double) * 4 * outTot));
    Prof pf = make_prof(c);
    // the series builds' timing (one event pair each, destroyed on return)
    struct EvPairs {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> v;
        ~EvPairs() {
            for (auto& e : v) {
                if (e.first) (void)hipEventDestroy(e.first);
                if (e.second) (void)hipEventDestroy(e.second);
            }
        }
    } buildEv;
    std::vector<char> built(nCtx, 0);

    std::vector<Freq> frH[2];
    std::vector<int> neH[2];
    std::vector<double> fqH[2];
    auto zbuf = [&](int buf) { return c->dZ.as<double>() + 2 * dims.zBuf * (size_t)buf; };
    auto z32buf = [&](int buf) { return mixed ? c->dZ32.as<float>() + 2 * dims.zBuf * (size_t)buf : nullptr; };
    auto vpbuf = [&](int buf) { return c->dVp.as<double>() + 2 * dims.vec * (size_t)buf; };
    auto nebuf = [&](int buf) { return c->dNElem.as<int>() + (size_t)dims.B * buf; };
    auto bufs = [&](int buf) {
        BatchBufs bb = batch_bufs_dims(c, o, dims, buf);
        if (mixed) bb.w.Bn2 = c->dBn2.as<double>() + (size_t)dims.B * buf;
        return bb;
    };
    // fill batch bi into buffer bi % nbuf on stream sf: the series of the
    // contexts it starts, every segment's matrices (FP64 and the complex64
    // copy), the identity padding, the padded right-hand sides and their
    // norms; evFill[buf] marks its end
    auto enqueue_fill = [&](int bi) -> efie_status {
        const PackBatch& b = batches[bi];
        const int buf = bi % nbuf;
        HIPCHK(hipEventRecord(c->evFillStart[buf], sf));
        for (const PackSeg& sg : b.seg) {
            const int ci = ctxOf[sg.item];
            if (buildP[ci] < 0 || built[ci]) continue;
            built[ci] = 1;
            efie_ctx* p = ctxs[ci];
            buildEv.v.push_back({nullptr, nullptr});
            HIPCHK(hipEventCreate(&buildEv.v.back().first));
            HIPCHK(hipEventCreate(&buildEv.v.back().second));
            HIPCHK(hipEventRecord(buildEv.v.back().first, sf));
            const int P = buildP[ci];
            HIPCHK(launch_series(p->dReg.as<Pt>(), p->Nq, p->dTri.as<TriGeo>(), p->Nt,
                                 p->havePack ? p->dNearMat.as<uint8_t>() : nullptr, p->nearEps,
                                 p->dNearPts.as<Pt>(), p->NpNear, p->Nq2s, p->dAeq.as<double>(),
                                 optsOf[sg.item].near_alpha, p->dList.as<int2>(), (int)p->nNear,
                                 p->dSelfPts.as<Pt>(), p->NsSelf, p->dASelf.as<double>(), P, c->dSerQ.as<double>(),
                                 sf));
            HIPCHK(launch_series_coeffs(c->dSerQ.as<double>(), P + 3, p->dTri.as<TriGeo>(), p->dSideTri.as<int>(),
                                        p->dSide.as<Side>(), p->Ne, const_cast<double*>(coefC[ci]), sf));
            HIPCHK(hipEventRecord(buildEv.v.back().second, sf));
        }
        frH[buf].assign(b.nb, Freq());
        neH[buf].assign(b.nb, 0);
        fqH[buf].assign(b.nb, 0.0);
        for (const PackSeg& sg : b.seg)
            for (int i = 0; i < sg.nf; ++i) {
                const int f = sg.f0 + i;
                const double fq = items[sg.item].freqs[f];
                frH[buf][sg.e0 + i] = make_freq(ks[sg.item][f], 2.0 * kPi * fq, mu, eps);
                neH[buf][sg.e0 + i] = nOf[sg.item];
                fqH[buf][sg.e0 + i] = fq;
            }
        Freq* dfr = c->dFreq.as<Freq>() + (size_t)dims.B * buf;
        HIPCHK(hipMemcpyAsync(dfr, frH[buf].data(), sizeof(Freq) * b.nb, hipMemcpyHostToDevice, sf));
        HIPCHK(hipMemcpyAsync(nebuf(buf), neH[buf].data(), sizeof(int) * b.nb, hipMemcpyHostToDevice, sf));
        const size_t zE = (size_t)b.np * b.lda;
        double* Zb = zbuf(buf);
        float* Z32b = z32buf(buf);
        double* Vp = vpbuf(buf);
        for (const PackSeg& sg : b.seg) {
            const int g = sg.item, n = nOf[g];
            efie_ctx* p = items[g].problem;
            const NearPlan& plan = plans[g];
            double* Zs = Zb + 2 * zE * sg.e0;
            float* Z32s = Z32b ? Z32b + 2 * zE * sg.e0 : nullptr;
            bool w32 = Z32s != nullptr;
            if (plan.series) {
                pf.begin(EFIE_K_SERIES);
                // beside a factorisation a capped grid leaves it the CUs
                HIPCHK(launch_series_fill(coefC[ctxOf[g]], plan.serNT, plan.lossy, n, dfr + sg.e0, sg.nf, Zs,
                                          (size_t)b.lda, 2 * zE, Z32s, (size_t)b.lda, 2 * zE, sf,
                                          overlap && bi > 0 ? kFillBlocksShared : 0, triFill));
                pf.end(EFIE_K_SERIES);
            } else {
                for (int i = 0; i < sg.nf; ++i) {
                    bool w = false;
                    efie_status e = fill_z(p, &optsOf[g], plan, dfr + sg.e0 + i, Zs + 2 * zE * i, (size_t)b.lda,
                                           Z32s ? Z32s + 2 * zE * i : nullptr, &w, sf);
                    if (e) return e;
                    w32 = w32 && w;
                }
            }
            pf.begin(EFIE_K_OTHER);
            if (Z32s && !w32) HIPCHK(launch_to_f32(Zs, Z32s, zE * sg.nf, sf));
            HIPCHK(launch_pad_identity(Zs, Z32s, zE, n, b.np, b.lda, sg.nf, sf));
            HIPCHK(launch_pack_rhs(c->dItemV.as<double>() + vOff[g], n, b.np, sg.nf, Vp + 2 * (size_t)b.np * sg.e0,
                                   sf));
            if (mixed) HIPCHK(launch_fill_dbl(c->dBn2.as<double>() + (size_t)dims.B * buf + sg.e0, vn2[g], sg.nf, sf));
            pf.end(EFIE_K_OTHER);
        }
        if (rhsCol) HIPCHK(launch_set_rhs(Vp, Zb, Z32b, b.np, b.lda, zE, b.nb, sf, 2 * (size_t)b.np));
        HIPCHK(hipEventRecord(c->evFill[buf], sf));
        return ok();
    };

    double msFill = 0, msFac = 0, msSolve = 0;
    HIPCHK(hipEventRecord(c->ev[0], s));
    if (overlap) {
        HIPCHK(hipStreamWaitEvent(sf, c->ev[0], 0));
        HIPCHK(hipStreamWaitEvent(si, c->ev[0], 0));
    }
    // mixed precision: systems the FP32 factorisation plus refinement did
    // not solve, as (item, frequency), redone after the pipeline
    std::vector<std::pair<int, int>> retry;
    auto finish = [&](int bj) -> efie_status {
        const PackBatch& b = batches[bj];
        const int buf = bj % nbuf;
        HIPCHK(hipEventSynchronize(c->evIR[buf]));
        float t1, t2, t3;
        HIPCHK(hipEventElapsedTime(&t1, c->evFillStart[buf], c->evFill[buf]));
        HIPCHK(hipEventElapsedTime(&t2, c->evLUStart[buf], c->evLU[buf]));
        HIPCHK(hipEventElapsedTime(&t3, c->evIRStart[buf], c->evIR[buf]));
        msFill += t1;
        msFac += t2;
        msSolve += t3;
        std::vector<int> rb(b.nb);
        int nrb = 0;
        efie_status e = finish_batch(c, o, b.nb, bufs(buf), dims.B, S, fqH[buf].data(), mixed ? rb.data() : nullptr,
                                     mixed ? &nrb : nullptr);
        for (int i = 0; i < nrb; ++i)
            for (const PackSeg& sg : b.seg)
                if (rb[i] >= sg.e0 && rb[i] < sg.e0 + sg.nf) retry.push_back({sg.item, sg.f0 + rb[i] - sg.e0});
        return e;
    };
    st = enqueue_fill(0);
    if (st) return st;
    for (int bi = 0; bi < nBatch; ++bi) {
        const PackBatch& b = batches[bi];
        const int buf = bi % nbuf;
        const BatchBufs bb = bufs(buf);
        HIPCHK(hipStreamWaitEvent(s, c->evFill[buf], 0));
        HIPCHK(hipEventRecord(c->evLUStart[buf], s));
        st = factor_batch(c, o, b.np, b.lda, b.nb, pf, true, bb, nebuf(buf), overlap ? sf : nullptr,
                          bi > 0 ? c->evTail[(bi - 1) & 1] : nullptr, c->evTail[bi & 1]);
        if (st) return st;
        HIPCHK(hipEventRecord(c->evLU[buf], s));
        HIPCHK(hipStreamWaitEvent(si, c->evLU[buf], 0));
        HIPCHK(hipEventRecord(c->evIRStart[buf], si));
        const double* X = nullptr;
        size_t xStride = 0, xinc = 0;
        // beside the next batch's factorisation the triangular solves run
        // with a few blocks per matrix
        c->tw.rows = overlap && bi + 1 < nBatch ? kTrsvRowsShared : 0;
        st = refine_batch(c, o, b.np, b.lda, b.nb, vpbuf(buf), 2 * (size_t)b.np, o.model != EFIE_MODEL_CENTROID, pf,
                          bb, si, &X, &xStride, &xinc);
        c->tw.rows = 0;
        if (st) return st;
        pf.begin(EFIE_K_OTHER);
        for (const PackSeg& sg : b.seg) {
            const efie_sweep_item& it = items[sg.item];
            HIPCHK(launch_epilogue(X + xStride * sg.e0, xStride, xinc, c->dItemWl.as<double>() + vOff[sg.item],
                                   nOf[sg.item], it.patchEdge - 1, it.groundEdge - 1,
                                   o.model == EFIE_MODEL_CENTROID, it.V0, it.Z0, sg.nf,
                                   c->dOut.as<double>() + 4 * (outOff[sg.item] + sg.f0), si));
        }
        pf.end(EFIE_K_OTHER, (int)b.seg.size());
        st = stage_stats(o, b.nb, dims.B, bb, si);
        if (st) return st;
        HIPCHK(hipEventRecord(c->evIR[buf], si));
        if (overlap) {
            if (bi >= 1) {
                st = finish(bi - 1);
                if (st) return st;
            }
            if (bi + 1 < nBatch) {
                if (bi >= 1) HIPCHK(hipStreamWaitEvent(sf, c->evIR[(bi + 1) % nbuf], 0));
                st = enqueue_fill(bi + 1);
                if (st) return st;
            }
        } else {
            st = finish(bi);
            if (st) return st;
            if (bi + 1 < nBatch) {
                st = enqueue_fill(bi + 1);
                if (st) return st;
            }
        }
    }
    if (overlap) {
        st = finish(nBatch - 1);
        if (st) return st;
        HIPCHK(hipEventRecord(c->ev[1], si));
        HIPCHK(hipStreamWaitEvent(s, c->ev[1], 0));
    }
    HIPCHK(hipEventRecord(c->ev[5], s));
    std::vector<double> out(4 * outTot);
    HIPCHK(hipMemcpyAsync(out.data(), c->dOut.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipStreamSynchronize(sf));
    for (int g = 0; g < count; ++g)
        for (int f = 0; f < nfOf[g]; ++f) {
            const double* r = out.data() + 4 * (outOff[g] + f);
            items[g].Zin[2 * f] = r[0];
            items[g].Zin[2 * f + 1] = r[1];
            items[g].S11[2 * f] = r[2];
            items[g].S11[2 * f + 1] = r[3];
        }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[5]));
    S->ms_total = tot;
    S->ms_fill = msFill;
    S->ms_factor = msFac;
    S->ms_solve = msSolve;
    for (auto& e : buildEv.v) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        S->ms_moments += ms;
        g_momBuilds.fetch_add(1);
    }
    // series built into a context's own buffer stay cached there; released
    // ones (in the pool, or cached before) are gone
    for (int ci = 0; ci < nCtx; ++ci) {
        efie_ctx* p = ctxs[ci];
        if (buildP[ci] >= 0 && !relCtx[ci]) {
            p->serP = buildP[ci];
            p->serAlpha = alphaCtx[ci];
        }
        if (freeAfter[ci]) {
            p->dSerC.release();
            p->serP = -1;
        }
    }
    c->prof = false;
    if (!retry.empty()) {
        // the frequencies the mixed solve could not finish, again in a call
        // of their own: after the unpivoted symmetric factorisation with the
        // pivoted FP32 LU (whose own failures go on to FP64), after a
        // pivoted FP32 LU with the FP64 LU
        std::vector<std::vector<int>> fi(count);
        for (const auto& r : retry) fi[r.first].push_back(r.second);
        std::vector<std::vector<double>> fr2(count), zin2(count), s112(count);
        std::vector<efie_sweep_item> sub;
        for (int g = 0; g < count; ++g) {
            if (fi[g].empty()) continue;
            const size_t k = fi[g].size();
            for (int f : fi[g]) fr2[g].push_back(items[g].freqs[f]);
            zin2[g].resize(2 * k);
            s112[g].resize(2 * k);
            efie_sweep_item it = items[g];
            it.freqs = fr2[g].data();
            it.Nf = (int)k;
            it.Zin = zin2[g].data();
            it.S11 = s112[g].data();
            sub.push_back(it);
        }
        efie_opts o2 = o;
        const bool toPivoted = o.lu_kernel == EFIE_LU_SYM;
        if (toPivoted) o2.lu_kernel = EFIE_LU_FP32;
        else o2.precision = EFIE_PREC_FP64;
        efie_stats S2;
        st = efie_sweep_many(c, sub.data(), (int)sub.size(), mu, eps_re, eps_im, &o2, &S2);
        if (st) return st;
        for (int g = 0; g < count; ++g)
            for (size_t i = 0; i < fi[g].size(); ++i) {
                const int f = fi[g][i];
                items[g].Zin[2 * f] = zin2[g][2 * i];
                items[g].Zin[2 * f + 1] = zin2[g][2 * i + 1];
                items[g].S11[2 * f] = s112[g][2 * i];
                items[g].S11[2 * f + 1] = s112[g][2 * i + 1];
            }
        S->ms_total += S2.ms_total;
        S->ms_fill += S2.ms_fill;
        S->ms_factor += S2.ms_factor;
        S->ms_solve += S2.ms_solve;
        S->ms_moments += S2.ms_moments;
        S->ir_iters_max = std::max(S->ir_iters_max, S2.ir_iters_max);
        S->ir_iters_total += S2.ir_iters_total;
        S->ir_resid_max = std::max(S->ir_resid_max, S2.ir_resid_max);
        for (int q = 0; q < EFIE_NKCLASS; ++q) {
            S->ms_kernel[q] += S2.ms_kernel[q];
            S->n_kernel[q] += S2.n_kernel[q];
        }
        if (toPivoted) {
            S->n_lu_retry = (int)retry.size();
            S->n_fp64_retry = S2.n_fp64_retry;
        } else {
            S->n_fp64_retry = (int)retry.size();
        }
    }
    return ok();
}

efie_status efie_zgesv(efie_ctx* c, int n, int batch, const double* A, const double* b, int precision,
                       double* x, efie_stats* stats) {
    if (!c || !A || !b || !x) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (n < 1 || batch < 1) return fail(EFIE_E_ARGS, "efie:args", "n and batch must be >= 1");
    if (precision != EFIE_PREC_FP64 && precision != EFIE_PREC_MIXED)
        return fail(EFIE_E_ARGS, "efie:opts", "unknown precision");
    HIPCHK(hipSetDevice(c->device));
    efie_opts o;
    efie_opts_default(&o);
    o.precision = precision;
    o.lu_kernel = EFIE_LU_FP32;  // general matrices: pivoted
    efie_stats local;
    efie_stats* S = stats ? stats : &local;
    std::memset(S, 0, sizeof *S);
    const int lda = ((n + 1) + 31) & ~31;
    const size_t zElems = (size_t)n * lda;
    hipStream_t s = c->stream;
    efie_status st = ensure_solver(c, o, n, lda, batch);
    if (st) return st;
    // row-major copies with the right-hand side in column n
    std::vector<double> h(2 * zElems, 0.0);
    std::vector<double> bn2(batch, 0.0);
    HIPCHK(c->dB.ensure(sizeof(double) * 2 * (size_t)n * batch));
    HIPCHK(hipMemcpyAsync(c->dB.p, b, sizeof(double) * 2 * (size_t)n * batch, hipMemcpyHostToDevice, s));
    for (int q = 0; q < batch; ++q) {
        const double* Aq = A + 2 * (size_t)n * n * q;
        const double* bq = b + 2 * (size_t)n * q;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) {
                h[2 * ((size_t)i * lda + j)] = Aq[2 * ((size_t)i + (size_t)j * n)];
                h[2 * ((size_t)i * lda + j) + 1] = Aq[2 * ((size_t)i + (size_t)j * n) + 1];
            }
            h[2 * ((size_t)i * lda + n)] = bq[2 * i];
            h[2 * ((size_t)i * lda + n) + 1] = bq[2 * i + 1];
            bn2[q] += bq[2 * i] * bq[2 * i] + bq[2 * i + 1] * bq[2 * i + 1];
        }
        HIPCHK(hipMemcpyAsync(c->dZ.as<double>() + 2 * zElems * q, h.data(), sizeof(double) * h.size(),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (precision == EFIE_PREC_MIXED)
        HIPCHK(hipMemcpyAsync(c->dBn2.p, bn2.data(), sizeof(double) * batch, hipMemcpyHostToDevice, s));
    Prof pf = make_prof(c);
    const BatchBufs bb = batch_bufs(c, o, n, lda, batch, 0);
    HIPCHK(hipEventRecord(c->ev[2], s));
    const double* X = nullptr;
    size_t xStride = 0, xinc = 0;
    st = factor_batch(c, o, n, lda, batch, pf, false, bb);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[3], s));
    st = refine_batch(c, o, n, lda, batch, c->dB.as<double>(), 2 * (size_t)n, false, pf, bb, s, &X, &xStride,
                       &xinc);
    if (st) return st;
    st = stage_stats(o, batch, batch, bb, s);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> redo(batch);
    int nredo = 0;
    st = finish_batch(c, o, batch, bb, batch, S, nullptr, precision == EFIE_PREC_MIXED ? redo.data() : nullptr,
                      &nredo);
    redo.resize(nredo);
    if (st) return st;
    for (int q = 0; q < batch; ++q)
        HIPCHK(hipMemcpy2DAsync(x + 2 * (size_t)n * q, sizeof(double) * 2, X + xStride * q, sizeof(double) * xinc,
                                sizeof(double) * 2, n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float t2 = 0, t3 = 0;
    HIPCHK(hipEventElapsedTime(&t2, c->ev[2], c->ev[3]));
    HIPCHK(hipEventElapsedTime(&t3, c->ev[3], c->ev[4]));
    S->ms_factor = t2;
    S->ms_solve = t3;
    S->ms_total = t2 + t3;
    // systems whose FP32 factorisation met a zero pivot or whose
    // refinement stalled above max(ir_tol, 1e-11): solved again by the FP64 LU
    for (int q : redo) {
        efie_stats S2;
        st = efie_zgesv(c, n, 1, A + 2 * (size_t)n * n * q, b + 2 * (size_t)n * q, EFIE_PREC_FP64,
                        x + 2 * (size_t)n * q, &S2);
        if (st) return st;
        S->ms_factor += S2.ms_factor;
        S->ms_solve += S2.ms_solve;
        S->ms_total += S2.ms_total;
    }
    S->n_fp64_retry = (int)redo.size();
    return ok();
}

efie_status efie_zgesv_noalloc(efie_ctx* c, int n, int batch, const double* A, const double* b, int precision,
                               double* x, efie_stats* stats) {
    if (!c || !A || !b || !x) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (n < 1 || batch < 1) return fail(EFIE_E_ARGS, "efie:args", "n and batch must be >= 1");
    if (batch > kNoAllocMaxBatch)
        return fail(EFIE_E_ARGS, "efie:args", "batch %d exceeds the allocation-free solve's %d", batch,
                    kNoAllocMaxBatch);
    if (precision != EFIE_PREC_FP64 && precision != EFIE_PREC_MIXED)
        return fail(EFIE_E_ARGS, "efie:opts", "unknown precision");
    if (!c->scratch) return fail(EFIE_E_STATE, "efie:state", "no scratch array on this context's device");
    HIPCHK(hipSetDevice(c->device));
    efie_opts o;
    efie_opts_default(&o);
    o.precision = precision;
    o.lu_kernel = EFIE_LU_FP32;  // general matrices: pivoted
    const bool mixed = precision == EFIE_PREC_MIXED;
    efie_stats local;
    efie_stats* S = stats ? stats : &local;
    std::memset(S, 0, sizeof *S);
    const int lda = ((n + 1) + 31) & ~31;
    const size_t zElems = (size_t)n * lda;
    const size_t nb = (size_t)batch, nn = (size_t)n;
    // every buffer of the solve, carved from the scratch array (256-byte aligned)
    char* base = static_cast<char*>(c->scratch);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        char* p = base + off;
        off += bytes;
        return static_cast<void*>(p);
    };
    const int nodes = lu_tree_nodes(n);
    double* Z = static_cast<double*>(take(16 * zElems * nb));
    float* Z32 = mixed ? static_cast<float*>(take(8 * zElems * nb)) : nullptr;
    double* dB = static_cast<double*>(take(16 * nn * nb));
    SolverWork w{};
    w.cand = static_cast<int*>(take(sizeof(int) * nb * nodes * lu_cand_stride()));
    w.cnt = static_cast<int*>(take(sizeof(int) * nb * nodes));
    w.ipiv = static_cast<int*>(take(sizeof(int) * nb * nn));
    int* info = static_cast<int*>(take(sizeof(int) * nb));
    w.swp = static_cast<int*>(take(sizeof(int) * nb * lu_swap_stride()));
    int* perm = mixed ? static_cast<int*>(take(sizeof(int) * nb * nn)) : nullptr;
    w.X = static_cast<double*>(take(16 * nb * nn));
    w.R = static_cast<double*>(take(16 * nb * nn));
    w.R32 = static_cast<float*>(take(8 * nb * nn));
    w.Rn2 = static_cast<double*>(take(sizeof(double) * nb));
    w.Bn2 = static_cast<double*>(take(sizeof(double) * nb));
    w.resid = static_cast<double*>(take(sizeof(double) * nb));
    w.done = static_cast<int*>(take(sizeof(int) * nb));
    w.iters = static_cast<int*>(take(sizeof(int) * nb));
    const size_t nflags = nb * (size_t)((n + 63) / 64);
    int* flags = static_cast<int*>(take(sizeof(int) * nflags));
    int* tickets = static_cast<int*>(take(sizeof(int) * nb));
    w.err = static_cast<int*>(take(sizeof(int)));
    if (off > kNoAllocScratchBytes)
        return fail(EFIE_E_ARGS, "efie:args",
                    "n = %d, batch = %d needs %zu bytes; the allocation-free solve has %zu (use efie_zgesv)", n,
                    batch, off, kNoAllocScratchBytes);
    // the array keeps the previous call's contents: clear what must start
    // at zero (matrix padding, tournament counters, solve flags, error word)
    hipStream_t s = c->stream;
    HIPCHK(launch_fill_dbl(Z, 0.0, (int)(2 * zElems * nb), s));
    HIPCHK(launch_fill_int(w.cnt, 0, (int)(nb * nodes), s));
    HIPCHK(launch_fill_int(flags, 0, (int)nflags, s));
    HIPCHK(launch_fill_int(w.err, 0, 1, s));
    TrsvWork tw{flags, tickets, w.err, 0};
    w.tw = &tw;
    // column-major A (A^T row-major) into the rows of Z, transposed in
    // place; the right-hand side into column n
    double bn2[kNoAllocMaxBatch];
    for (int q = 0; q < batch; ++q) {
        double* Zq = Z + 2 * zElems * q;
        HIPCHK(hipMemcpy2DAsync(Zq, 16 * (size_t)lda, A + 2 * nn * nn * q, 16 * nn, 16 * nn, nn,
                                hipMemcpyHostToDevice, s));
        HIPCHK(launch_transpose_square(Zq, n, (size_t)lda, s));
        HIPCHK(hipMemcpy2DAsync(Zq + 2 * nn, 16 * (size_t)lda, b + 2 * nn * q, 16, 16, nn, hipMemcpyHostToDevice, s));
        double v = 0;
        for (size_t i = 0; i < 2 * nn; ++i) v += b[2 * nn * q + i] * b[2 * nn * q + i];
        bn2[q] = v;
    }
    HIPCHK(hipMemcpyAsync(dB, b, 16 * nn * nb, hipMemcpyHostToDevice, s));
    if (mixed) HIPCHK(hipMemcpyAsync(w.Bn2, bn2, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    alignas(16) char host[stat_slot_bytes(kNoAllocMaxBatch)];
    BatchBufs bb{Z, Z32, info, perm, w, host};
    Prof pf = make_prof(c);
    HIPCHK(hipEventRecord(c->ev[2], s));
    const double* X = nullptr;
    size_t xStride = 0, xinc = 0;
    efie_status st = factor_batch(c, o, n, lda, batch, pf, false, bb);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[3], s));
    st = refine_batch(c, o, n, lda, batch, dB, 2 * nn, false, pf, bb, s, &X, &xStride, &xinc);
    if (st) return st;
    st = stage_stats(o, batch, batch, bb, s);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    int redo[kNoAllocMaxBatch];
    int nredo = 0;
    st = finish_batch(c, o, batch, bb, batch, S, nullptr, mixed ? redo : nullptr, &nredo);
    if (st) return st;
    for (int q = 0; q < batch; ++q)
        HIPCHK(hipMemcpy2DAsync(x + 2 * nn * q, sizeof(double) * 2, X + xStride * q, sizeof(double) * xinc,
                                sizeof(double) * 2, nn, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float t2 = 0, t3 = 0;
    HIPCHK(hipEventElapsedTime(&t2, c->ev[2], c->ev[3]));
    HIPCHK(hipEventElapsedTime(&t3, c->ev[3], c->ev[4]));
    S->ms_factor = t2;
    S->ms_solve = t3;
    S->ms_total = t2 + t3;
    // systems the FP32 factorisation could not solve: the FP64 LU, in the
    // same scratch array (the other results are already out)
    for (int i = 0; i < nredo; ++i) {
        const int q = redo[i];
        efie_stats S2;
        st = efie_zgesv_noalloc(c, n, 1, A + 2 * nn * nn * q, b + 2 * nn * q, EFIE_PREC_FP64, x + 2 * nn * q, &S2);
        if (st) return st;
        S->ms_factor += S2.ms_factor;
        S->ms_solve += S2.ms_solve;
        S->ms_total += S2.ms_total;
    }
    S->n_fp64_retry = nredo;
    return ok();
}

// ------------------------------------------------------------ surrogate --

void efie_sur_cfg_default(efie_sur_cfg* g) {
    if (!g) return;
    g->n_pairs = 2000000;
    g->n_val = 20000;
    g->epochs = 60;
    g->batch = 4096;
    g->lr = 2e-3;
    g->kl_max = 0.35;
    g->seed = 1;
}

int efie_sur_blob_floats(void) { return kSurBlob; }
int efie_mlp_params(void) { return kMlpParams; }

efie_status efie_sur_train(efie_ctx* c, const efie_opts* opts, const efie_sur_cfg* cfg, efie_sur_report* rep) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    HIPCHK(hipSetDevice(c->device));
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    efie_sur_cfg g;
    if (cfg) g = *cfg;
    else efie_sur_cfg_default(&g);
    std::vector<float> blob;
    efie_status st = sur_train(c->stream, &o, &g, blob, rep);
    if (st) return st;
    c->surBlob.swap(blob);
    c->surValid = false;
    return ok();
}

efie_status efie_sur_get(efie_ctx* c, float* blob) {
    if (!c || !blob) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (c->surBlob.empty()) return fail(EFIE_E_STATE, "efie:surrogate", "no model installed");
    std::copy(c->surBlob.begin(), c->surBlob.end(), blob);
    return ok();
}

efie_status efie_sur_set(efie_ctx* c, const float* blob) {
    if (!c || !blob) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    if (blob[kSurModel + 2 * kMlpIn + kMetaVersion] != kSurVersion)
        return fail(EFIE_E_ARGS, "efie:surrogate", "not a surrogate model blob of this version");
    c->surBlob.assign(blob, blob + kSurBlob);
    c->surValid = false;
    return ok();
}

efie_status efie_sur_set_exact(efie_ctx* c, int on) {
    if (!c) return fail(EFIE_E_ARGS, "efie:args", "null context");
    c->surExact = on != 0;
    return ok();
}

efie_status efie_mlp_forward(efie_ctx* c, const float* params, const float* X, int N, float* Y) {
    if (!c || !params || !X || !Y || N < 1) return fail(EFIE_E_ARGS, "efie:args", "bad arguments");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    HIPCHK(c->dMlpA.ensure(sizeof(float) * kMlpParams));
    HIPCHK(c->dMlpB.ensure(sizeof(float) * kMlpIn * (size_t)N));
    HIPCHK(c->dMlpC.ensure(sizeof(float) * kMlpOut * (size_t)N));
    HIPCHK(hipMemcpyAsync(c->dMlpA.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dMlpB.p, X, sizeof(float) * kMlpIn * (size_t)N, hipMemcpyHostToDevice, s));
    HIPCHK(launch_mlp_forward(c->dMlpB.as<float>(), c->dMlpA.as<float>(), c->dMlpC.as<float>(), N, s));
    HIPCHK(hipMemcpyAsync(Y, c->dMlpC.p, sizeof(float) * kMlpOut * (size_t)N, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ok();
}

efie_status efie_mlp_loss_grad(efie_ctx* c, const float* params, const float* X, const float* T, int N, float* loss,
                               float* grad) {
    if (!c || !params || !X || !T || !loss || !grad || N < 1) return fail(EFIE_E_ARGS, "efie:args", "bad arguments");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    HIPCHK(c->dMlpA.ensure(sizeof(float) * (kMlpParams + 1)));
    HIPCHK(c->dMlpB.ensure(sizeof(float) * kMlpIn * (size_t)N));
    HIPCHK(c->dMlpC.ensure(sizeof(float) * kMlpOut * (size_t)N));
    HIPCHK(c->dMlpD.ensure(sizeof(float) * mlp_grad_work_floats(N)));
    HIPCHK(c->dMlpE.ensure(sizeof(float) * kMlpParams));
    HIPCHK(hipMemcpyAsync(c->dMlpA.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dMlpB.p, X, sizeof(float) * kMlpIn * (size_t)N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->dMlpC.p, T, sizeof(float) * kMlpOut * (size_t)N, hipMemcpyHostToDevice, s));
    float* dLoss = c->dMlpA.as<float>() + kMlpParams;
    HIPCHK(launch_mlp_loss_grad(c->dMlpB.as<float>(), c->dMlpC.as<float>(), c->dMlpA.as<float>(), N,
                                c->dMlpD.as<float>(), c->dMlpE.as<float>(), dLoss, s));
    HIPCHK(hipMemcpyAsync(grad, c->dMlpE.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(loss, dLoss, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ok();
}

efie_status efie_diag_cache_counts(long long* geometry_builds, long long* moment_builds) {
    if (!geometry_builds || !moment_builds) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    *geometry_builds = g_geoBuilds.load();
    *moment_builds = g_momBuilds.load();
    return ok();
}

efie_status efie_diag_near_pairs(efie_ctx* c, const efie_opts* opts, long long* n, int* pairs, long long cap) {
    if (!c || !n || (cap > 0 && !pairs)) return fail(EFIE_E_ARGS, "efie:args", "null argument");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    o.model = EFIE_MODEL_QUADRATURE;
    efie_status st = prepare(c, &o, 0.20);
    if (st) return st;
    *n = c->nNear;
    const long long k = std::min(cap, c->nNear);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(pairs, c->dList.p, sizeof(int2) * (size_t)k, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return ok();
}

efie_status efie_diag_quad_points(efie_ctx* c, const efie_opts* opts, int set, long long* n, double* out,
                                  long long cap) {
    if (!c || !n || (cap > 0 && !out) || set < 0 || set > 4)
        return fail(EFIE_E_ARGS, "efie:args", "bad arguments");
    efie_opts o;
    if (opts) o = *opts;
    else efie_opts_default(&o);
    o.model = EFIE_MODEL_QUADRATURE;
    efie_status st = prepare(c, &o, 0.20);
    if (st) return st;
    const long long Nt = c->Nt;
    const DBuf* src[5] = {&c->dReg, &c->dNearPts, &c->dSelfPts, &c->dASelf, &c->dRad};
    const long long len[5] = {4 * Nt * c->Nq, 4 * Nt * c->NpNear, 4 * Nt * c->NsSelf * c->Nq2s,
                              Nt * c->NsSelf, Nt};
    *n = len[set];
    const long long k = std::min(cap, len[set]);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(out, src[set]->p, sizeof(double) * (size_t)k, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return ok();
}

efie_status efie_diag_pack_plan(const int* n, const int* nf, int count, int precision, int batch,
                                double free_bytes, int* out, int cap, int* nseg) {
    if (count < 0 || (count > 0 && (!n || !nf)) || !nseg || cap < 0 || (cap > 0 && !out))
        return fail(EFIE_E_ARGS, "efie:args", "bad arguments");
    std::vector<int> nOf(n, n + count), nfOf(nf, nf + count), order;
    for (int g = 0; g < count; ++g) {
        if (nOf[g] < 1 || nfOf[g] < 0) return fail(EFIE_E_ARGS, "efie:args", "item %d: bad order or count", g);
        if (nfOf[g] > 0) order.push_back(g);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nOf[a] > nOf[b]; });
    efie_opts o;
    efie_opts_default(&o);
    o.precision = precision;
    o.batch = batch;
    const std::vector<PackBatch> bs = plan_packs(order, nOf, nfOf, o, (size_t)std::max(0.0, free_bytes));
    int k = 0;
    for (size_t b = 0; b < bs.size(); ++b)
        for (const PackSeg& sg : bs[b].seg) {
            if (k < cap) {
                int* r = out + 6 * (size_t)k;
                r[0] = (int)b;
                r[1] = sg.item;
                r[2] = sg.f0;
                r[3] = sg.nf;
                r[4] = sg.e0;
                r[5] = bs[b].np;
            }
            ++k;
        }
    *nseg = k;
    return ok();
}

efie_status efie_diag_copy(efie_ctx* c, size_t bytes, int reps, double* ms) {
    if (!c || !ms || bytes == 0 || bytes % 16 || reps < 1) return fail(EFIE_E_ARGS, "efie:args", "bad arguments");
    HIPCHK(hipSetDevice(c->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(EFIE_E_NOMEM, "efie:hip", "hipMalloc: %s", hipGetErrorString(e));
    }
    hipStream_t s = c->stream;
    e = hipMemsetAsync(a, 0, bytes, s);
    if (e == hipSuccess) e = launch_copy16(a, b, bytes, s);
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy16(a, b, bytes, s);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(EFIE_E_HIP, "efie:hip", "copy: %s", hipGetErrorString(e));
    *ms = t / reps;
    return ok();
}

}  // extern "C"
