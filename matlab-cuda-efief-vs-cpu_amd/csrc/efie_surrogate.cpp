// efie_surrogate.cpp -- training of the near-pair surrogate: sampling of
// near triangle pairs from seeded random patch meshes, labels from the
// subdivided near-pair quadrature on the GPU, and the Adam loop over the
// network kernels (efie_mlp.hip).  No counterpart in the reference.
#include "efie_sur.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace efie {

namespace {

constexpr double kPiS = 3.141592653589793238462643383279502884;

// splitmix64 (Steele, Lea and Flood, 2014)
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    int irange(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
};

// Device buffer freed on scope exit.
struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 16); }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

#define SUR_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return sur_fail(EFIE_E_HIP, "efie:hip", hipGetErrorString(e_)); \
    } while (0)

// Training pairs: normalised vertices (18), features (12), d / L.
struct PairSet {
    std::vector<double> tri18;
    std::vector<float> feat;
    std::vector<double> dl;
    size_t size() const { return dl.size(); }
};

// Near pairs (t1 < t2, triangles_are_near.m criterion) of one random mesh of
// the make_patch_with_ground_mesh family with d / L in the accepted range.
void sample_mesh(Rng& g, const efie_opts* o, PairSet& out) {
    const int np = g.irange(4, 12);
    const int ng = np + g.irange(2, 14);
    const double cx = 1.0, cy = g.uni(0.6, 1.6), cm = std::sqrt(cx * cy);
    const double gc = g.uni(0.7, 1.8) * cm;
    const double h = g.uni(0.4, 2.5) * cm;
    const double jit = g.uni() < 0.3 ? 0.0 : g.uni(0.0, 0.12);
    int Nv = 0, Nt = 0;
    if (efie_mesh_counts(np, np, ng, ng, &Nv, &Nt)) return;
    std::vector<double> V(3 * (size_t)Nv);
    std::vector<int> F(3 * (size_t)Nt);
    if (efie_make_mesh(np * cy, np * cx, ng * gc, ng * gc, h, np, np, ng, ng, V.data(), F.data())) return;
    for (int v = 0; v < Nv; ++v) {
        const bool patch = V[v + 2 * (size_t)Nv] > 0;
        const double sx = patch ? cx : gc, sy = patch ? cy : gc;
        V[v] += jit * sx * g.uni(-1.0, 1.0);
        V[v + Nv] += jit * sy * g.uni(-1.0, 1.0);
    }
    std::vector<double> tv(9 * (size_t)Nt), cen(3 * (size_t)Nt), leq(Nt);
    for (int t = 0; t < Nt; ++t) {
        for (int q = 0; q < 3; ++q)
            for (int d = 0; d < 3; ++d) tv[9 * (size_t)t + 3 * q + d] = V[(F[t + q * (size_t)Nt] - 1) + d * (size_t)Nv];
        const double* w = &tv[9 * (size_t)t];
        for (int d = 0; d < 3; ++d) cen[3 * t + d] = (w[d] + w[3 + d] + w[6 + d]) / 3.0;
        const double ex = w[3] - w[0], ey = w[4] - w[1], ez = w[5] - w[2];
        const double fx = w[6] - w[0], fy = w[7] - w[1], fz = w[8] - w[2];
        const double nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
        const double A = 0.5 * std::sqrt(nx * nx + ny * ny + nz * nz);
        leq[t] = std::sqrt(4 * A / std::sqrt(3.0));
    }
    double t18[18];
    SurPair p;
    for (int t1 = 0; t1 < Nt; ++t1)
        for (int t2 = t1 + 1; t2 < Nt; ++t2) {
            const double dx = cen[3 * t1] - cen[3 * t2], dy = cen[3 * t1 + 1] - cen[3 * t2 + 1],
                         dz = cen[3 * t1 + 2] - cen[3 * t2 + 2];
            if (!(std::sqrt(dx * dx + dy * dy + dz * dz) < o->near_sing_eps * 0.5 * (leq[t1] + leq[t2]))) continue;
            const double dl = sur_frame(&tv[9 * (size_t)t1], &tv[9 * (size_t)t2], &p, t18);
            if (dl < kSurDlMin || dl > kSurDlMax) continue;
            out.tri18.insert(out.tri18.end(), t18, t18 + 18);
            out.feat.insert(out.feat.end(), p.feat, p.feat + 12);
            out.dl.push_back(dl);
        }
}

// n pairs from meshes drawn with g, in random order
PairSet sample_pairs(Rng& g, const efie_opts* o, size_t n) {
    PairSet all;
    while (all.size() < n) sample_mesh(g, o, all);
    // Fisher-Yates over the pairs, then keep n
    const size_t m = all.size();
    std::vector<size_t> perm(m);
    for (size_t i = 0; i < m; ++i) perm[i] = i;
    for (size_t i = m - 1; i > 0; --i) std::swap(perm[i], perm[g.next() % (i + 1)]);
    PairSet out;
    out.tri18.resize(18 * n);
    out.feat.resize(12 * n);
    out.dl.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t j = perm[i];
        std::memcpy(&out.tri18[18 * i], &all.tri18[18 * j], 18 * sizeof(double));
        std::memcpy(&out.feat[12 * i], &all.feat[12 * j], 12 * sizeof(float));
        out.dl[i] = all.dl[j];
    }
    return out;
}

// Labels T [n][16] (and record norms) of a pair set with wavenumbers kt.
efie_status label_pairs(hipStream_t s, const PairSet& P, const std::vector<double>& kt, const efie_opts* o,
                        std::vector<double>& T, std::vector<double>& nrm, double* ms) {
    std::vector<double> sub, reg;
    sur_rule_tables(o->quad_order, o->near_order, o->near_subdiv, sub, reg);
    const int Np = (int)(sub.size() / 4), Nq = (int)(reg.size() / 4);
    const size_t n = P.size(), chunk = 65536;
    Dev dSub, dReg, dTri, dKt, dT, dN;
    SUR_HIP(dSub.alloc(sizeof(double) * sub.size()));
    SUR_HIP(dReg.alloc(sizeof(double) * reg.size()));
    SUR_HIP(dTri.alloc(sizeof(double) * 18 * chunk));
    SUR_HIP(dKt.alloc(sizeof(double) * chunk));
    SUR_HIP(dT.alloc(sizeof(double) * 16 * chunk));
    SUR_HIP(dN.alloc(sizeof(double) * chunk));
    SUR_HIP(hipMemcpyAsync(dSub.p, sub.data(), sizeof(double) * sub.size(), hipMemcpyHostToDevice, s));
    SUR_HIP(hipMemcpyAsync(dReg.p, reg.data(), sizeof(double) * reg.size(), hipMemcpyHostToDevice, s));
    T.resize(16 * n);
    nrm.resize(n);
    hipEvent_t e0, e1;
    SUR_HIP(hipEventCreate(&e0));
    SUR_HIP(hipEventCreate(&e1));
    float tot = 0;
    for (size_t c0 = 0; c0 < n; c0 += chunk) {
        const int m = (int)std::min(chunk, n - c0);
        SUR_HIP(hipMemcpyAsync(dTri.p, &P.tri18[18 * c0], sizeof(double) * 18 * m, hipMemcpyHostToDevice, s));
        SUR_HIP(hipMemcpyAsync(dKt.p, &kt[c0], sizeof(double) * m, hipMemcpyHostToDevice, s));
        SUR_HIP(hipEventRecord(e0, s));
        SUR_HIP(launch_sur_labels(dTri.as<double>(), dKt.as<double>(), nullptr, nullptr, m, dSub.as<double>(), Np,
                                  dReg.as<double>(), Nq, o->near_alpha, dT.as<double>(), dN.as<double>(), s));
        SUR_HIP(hipEventRecord(e1, s));
        SUR_HIP(hipMemcpyAsync(&T[16 * c0], dT.p, sizeof(double) * 16 * m, hipMemcpyDeviceToHost, s));
        SUR_HIP(hipMemcpyAsync(&nrm[c0], dN.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        SUR_HIP(hipStreamSynchronize(s));
        float t = 0;
        SUR_HIP(hipEventElapsedTime(&t, e0, e1));
        tot += t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = tot;
    return EFIE_OK;
}

}  // namespace

double sur_frame(const double* v1, const double* v2, SurPair* p, double* tri18) {
    const double c1[3] = {(v1[0] + v1[3] + v1[6]) / 3.0, (v1[1] + v1[4] + v1[7]) / 3.0,
                          (v1[2] + v1[5] + v1[8]) / 3.0};
    const double c2[3] = {(v2[0] + v2[3] + v2[6]) / 3.0, (v2[1] + v2[4] + v2[7]) / 3.0,
                          (v2[2] + v2[5] + v2[8]) / 3.0};
    double e[3] = {v1[3] - v1[0], v1[4] - v1[1], v1[5] - v1[2]};
    const double f[3] = {v1[6] - v1[0], v1[7] - v1[1], v1[8] - v1[2]};
    double nn[3] = {e[1] * f[2] - e[2] * f[1], e[2] * f[0] - e[0] * f[2], e[0] * f[1] - e[1] * f[0]};
    const double a2 = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    const double le = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (int d = 0; d < 3; ++d) {
        e[d] /= le;
        nn[d] /= a2;
    }
    const double ey[3] = {nn[1] * e[2] - nn[2] * e[1], nn[2] * e[0] - nn[0] * e[2], nn[0] * e[1] - nn[1] * e[0]};
    const double L = std::sqrt(2.0 * a2 / std::sqrt(3.0));  // sqrt(4 A / sqrt(3)), A = a2 / 2
    for (int d = 0; d < 3; ++d) {
        p->Rm[d] = e[d];
        p->Rm[3 + d] = ey[d];
        p->Rm[6 + d] = nn[d];
    }
    double loc[18];
    for (int q = 0; q < 6; ++q) {
        const double* v = q < 3 ? v1 + 3 * q : v2 + 3 * (q - 3);
        const double r[3] = {v[0] - c1[0], v[1] - c1[1], v[2] - c1[2]};
        for (int a = 0; a < 3; ++a)
            loc[3 * q + a] = (p->Rm[3 * a] * r[0] + p->Rm[3 * a + 1] * r[1] + p->Rm[3 * a + 2] * r[2]) / L;
    }
    if (tri18) std::memcpy(tri18, loc, sizeof loc);
    p->feat[0] = (float)loc[0];
    p->feat[1] = (float)loc[1];
    p->feat[2] = (float)loc[3];
    for (int i = 0; i < 9; ++i) p->feat[3 + i] = (float)loc[9 + i];
    const double dx = c2[0] - c1[0], dy = c2[1] - c1[1], dz = c2[2] - c1[2];
    p->L = L;
    p->d = std::sqrt(dx * dx + dy * dy + dz * dz);
    return p->d / L;
}

efie_status sur_train(hipStream_t s, const efie_opts* o, const efie_sur_cfg* cfg, std::vector<float>& blob,
                      efie_sur_report* rep) {
    efie_sur_report R;
    std::memset(&R, 0, sizeof R);
    const size_t nTr = (size_t)cfg->n_pairs, nVa = (size_t)cfg->n_val;
    const int B = cfg->batch;
    if (nTr < (size_t)B || B < 256 || B % 256 || cfg->epochs < 1 || nVa < 1 || !(cfg->lr > 0) ||
        !(cfg->kl_max > 0))
        return sur_fail(EFIE_E_ARGS, "efie:surrogate",
                        "need batch a multiple of 256, n_pairs >= batch, n_val >= 1, epochs >= 1, lr > 0, kl_max > 0");
    if (o->model != EFIE_MODEL_QUADRATURE || o->near_subdiv < 0 || o->near_subdiv > 2)
        return sur_fail(EFIE_E_ARGS, "efie:surrogate", "the surrogate covers the quadrature model, near_subdiv 0..2");

    // ---- sample (host) -------------------------------------------------
    auto t0 = std::chrono::steady_clock::now();
    Rng gTr(cfg->seed * 0x2545F4914F6CDD1DULL + 1), gVa(cfg->seed * 0x2545F4914F6CDD1DULL + 0x5bd1e995ULL);
    PairSet tr = sample_pairs(gTr, o, nTr), va = sample_pairs(gVa, o, nVa);
    std::vector<double> ktTr(nTr), ktVa(nVa);
    for (auto& k : ktTr) k = gTr.uni(0.0, cfg->kl_max);
    for (auto& k : ktVa) k = gVa.uni(0.0, cfg->kl_max);
    R.ms_sample = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    // ---- labels (device) -----------------------------------------------
    std::vector<double> Ttr, Tva, nTrm, nVam;
    double ms1 = 0, ms2 = 0;
    efie_status st = label_pairs(s, tr, ktTr, o, Ttr, nTrm, &ms1);
    if (st) return st;
    st = label_pairs(s, va, ktVa, o, Tva, nVam, &ms2);
    if (st) return st;
    R.ms_label = ms1 + ms2;

    // ---- standardisation -----------------------------------------------
    auto xrow = [](const PairSet& P, const std::vector<double>& kt, size_t i, double* x) {
        for (int c = 0; c < 12; ++c) x[c] = P.feat[12 * i + c];
        x[12] = kt[i] * P.dl[i];  // k d = (k L)(d / L)
    };
    double xm[kMlpIn] = {0}, xs[kMlpIn] = {0}, xmin[kMlpIn], xmax[kMlpIn], ym[kMlpOut] = {0}, ys[kMlpOut] = {0};
    for (int c = 0; c < kMlpIn; ++c) {
        xmin[c] = 1e300;
        xmax[c] = -1e300;
    }
    for (size_t i = 0; i < nTr; ++i) {
        double x[kMlpIn];
        xrow(tr, ktTr, i, x);
        for (int c = 0; c < kMlpIn; ++c) {
            xm[c] += x[c];
            xs[c] += x[c] * x[c];
            xmin[c] = std::min(xmin[c], x[c]);
            xmax[c] = std::max(xmax[c], x[c]);
        }
        for (int c = 0; c < kMlpOut; ++c) {
            ym[c] += Ttr[16 * i + c];
            ys[c] += Ttr[16 * i + c] * Ttr[16 * i + c];
        }
    }
    double ysMax = 0;
    for (int c = 0; c < kMlpIn; ++c) {
        xm[c] /= (double)nTr;
        xs[c] = std::sqrt(std::max(0.0, xs[c] / (double)nTr - xm[c] * xm[c]));
        if (!(xs[c] > 1e-12)) xs[c] = 1.0;  // a constant input
    }
    for (int c = 0; c < kMlpOut; ++c) {
        ym[c] /= (double)nTr;
        ys[c] = std::sqrt(std::max(0.0, ys[c] / (double)nTr - ym[c] * ym[c]));
        ysMax = std::max(ysMax, ys[c]);
    }
    // outputs that vanish by symmetry (up to rounding) are not scaled up
    for (int c = 0; c < kMlpOut; ++c) ys[c] = std::max(ys[c], 1e-9 * ysMax);

    auto pack = [&](const PairSet& P, const std::vector<double>& kt, const std::vector<double>& T, size_t n,
                    std::vector<float>& X, std::vector<float>& Y) {
        X.resize(kMlpIn * n);
        Y.resize(kMlpOut * n);
        for (size_t i = 0; i < n; ++i) {
            double x[kMlpIn];
            xrow(P, kt, i, x);
            for (int c = 0; c < kMlpIn; ++c) X[kMlpIn * i + c] = (float)((x[c] - xm[c]) / xs[c]);
            for (int c = 0; c < kMlpOut; ++c) Y[kMlpOut * i + c] = (float)((T[16 * i + c] - ym[c]) / ys[c]);
        }
    };
    std::vector<float> Xtr, Ytr, Xva, Yva;
    pack(tr, ktTr, Ttr, nTr, Xtr, Ytr);
    pack(va, ktVa, Tva, nVa, Xva, Yva);

    // ---- parameters: Glorot-uniform weights, zero biases ---------------
    std::vector<float> params(kMlpParams, 0.0f);
    {
        Rng gi(cfg->seed * 0x9E3779B97F4A7C15ULL + 7);
        const int fin[4] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
        const int fout[4] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
        size_t off = 0;
        for (int l = 0; l < 4; ++l) {
            const double a = std::sqrt(6.0 / (fin[l] + fout[l]));
            for (int i = 0; i < fin[l] * fout[l]; ++i) params[off + i] = (float)gi.uni(-a, a);
            off += (size_t)fin[l] * fout[l] + fout[l];
        }
    }

    // ---- training (device) ---------------------------------------------
    Dev dX, dY, dP, dG, dM, dV, dW, dL, dXv, dYv;
    SUR_HIP(dX.alloc(sizeof(float) * Xtr.size()));
    SUR_HIP(dY.alloc(sizeof(float) * Ytr.size()));
    SUR_HIP(dP.alloc(sizeof(float) * kMlpParams));
    SUR_HIP(dG.alloc(sizeof(float) * kMlpParams));
    SUR_HIP(dM.alloc(sizeof(float) * kMlpParams));
    SUR_HIP(dV.alloc(sizeof(float) * kMlpParams));
    SUR_HIP(dW.alloc(sizeof(float) * mlp_grad_work_floats(B)));
    const int steps = (int)(nTr / (size_t)B);
    SUR_HIP(dL.alloc(sizeof(float) * steps));
    SUR_HIP(dXv.alloc(sizeof(float) * Xva.size()));
    SUR_HIP(dYv.alloc(sizeof(float) * Yva.size()));
    SUR_HIP(hipMemcpyAsync(dX.p, Xtr.data(), sizeof(float) * Xtr.size(), hipMemcpyHostToDevice, s));
    SUR_HIP(hipMemcpyAsync(dY.p, Ytr.data(), sizeof(float) * Ytr.size(), hipMemcpyHostToDevice, s));
    SUR_HIP(hipMemcpyAsync(dXv.p, Xva.data(), sizeof(float) * Xva.size(), hipMemcpyHostToDevice, s));
    SUR_HIP(hipMemcpyAsync(dP.p, params.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
    SUR_HIP(hipMemsetAsync(dM.p, 0, sizeof(float) * kMlpParams, s));
    SUR_HIP(hipMemsetAsync(dV.p, 0, sizeof(float) * kMlpParams, s));
    hipEvent_t e0, e1;
    SUR_HIP(hipEventCreate(&e0));
    SUR_HIP(hipEventCreate(&e1));
    SUR_HIP(hipEventRecord(e0, s));
    const long long total = (long long)steps * cfg->epochs;
    const long long warm = std::max(1LL, total / 50);
    const double lr0 = cfg->lr, lr1 = cfg->lr / 100.0;
    Rng go(cfg->seed + 99);
    std::vector<int> order(steps);
    std::vector<float> lossH(steps);
    long long t = 0;
    for (int ep = 0; ep < cfg->epochs; ++ep) {
        for (int i = 0; i < steps; ++i) order[i] = i;
        for (int i = steps - 1; i > 0; --i) std::swap(order[i], order[go.next() % (uint64_t)(i + 1)]);
        for (int i = 0; i < steps; ++i, ++t) {
            const size_t off = (size_t)order[i] * B;
            SUR_HIP(launch_mlp_loss_grad(dX.as<float>() + kMlpIn * off, dY.as<float>() + kMlpOut * off, dP.as<float>(),
                                         B, dW.as<float>(), dG.as<float>(), dL.as<float>() + i, s));
            double lr = t < warm ? lr0 * (double)(t + 1) / (double)warm
                                 : lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(kPiS * (double)(t - warm) /
                                                                              (double)std::max(1LL, total - warm)));
            SUR_HIP(launch_adam(dP.as<float>(), dG.as<float>(), dM.as<float>(), dV.as<float>(), kMlpParams,
                                (float)lr, (int)(t + 1), s));
        }
    }
    SUR_HIP(hipEventRecord(e1, s));
    SUR_HIP(hipMemcpyAsync(lossH.data(), dL.p, sizeof(float) * steps, hipMemcpyDeviceToHost, s));
    SUR_HIP(hipMemcpyAsync(params.data(), dP.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
    SUR_HIP(launch_mlp_forward(dXv.as<float>(), dP.as<float>(), dYv.as<float>(), (int)nVa, s));
    std::vector<float> Yp(Yva.size());
    SUR_HIP(hipMemcpyAsync(Yp.data(), dYv.p, sizeof(float) * Yp.size(), hipMemcpyDeviceToHost, s));
    SUR_HIP(hipStreamSynchronize(s));
    float msT = 0;
    SUR_HIP(hipEventElapsedTime(&msT, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    R.ms_train = msT;
    R.steps = (int)total;
    double lsum = 0;
    for (float l : lossH) lsum += l;
    R.train_loss = lsum / steps;

    // ---- held-out errors -----------------------------------------------
    double vl = 0, se = 0, sb = 0, me = 0, mb = 0;
    for (size_t i = 0; i < nVa; ++i) {
        double e2 = 0, b2 = 0;
        for (int c = 0; c < kMlpOut; ++c) {
            const double dy = (double)Yp[kMlpOut * i + c] - (double)Yva[kMlpOut * i + c];
            vl += dy * dy;
            const double pred = (double)Yp[kMlpOut * i + c] * ys[c] + ym[c];
            const double tv = Tva[16 * i + c];
            e2 += (pred - tv) * (pred - tv);
            b2 += tv * tv;
        }
        const double re = std::sqrt(e2) / nVam[i], rb = std::sqrt(b2) / nVam[i];
        se += re * re;
        sb += rb * rb;
        me = std::max(me, re);
        mb = std::max(mb, rb);
    }
    R.val_loss = vl / ((double)nVa * kMlpOut);
    R.val_rel_rms = std::sqrt(se / (double)nVa);
    R.val_rel_max = me;
    R.base_rel_rms = std::sqrt(sb / (double)nVa);
    R.base_rel_max = mb;
    R.n_train = (long long)nTr;
    R.n_val = (long long)nVa;

    // ---- blob ----------------------------------------------------------
    blob.assign(kSurBlob, 0.0f);
    std::copy(params.begin(), params.end(), blob.begin());
    float* nm = blob.data() + kMlpParams;
    for (int c = 0; c < kMlpIn; ++c) {
        nm[c] = (float)xm[c];
        nm[kMlpIn + c] = (float)(1.0 / xs[c]);
    }
    for (int c = 0; c < kMlpOut; ++c) {
        nm[2 * kMlpIn + c] = (float)ym[c];
        nm[2 * kMlpIn + kMlpOut + c] = (float)ys[c];
    }
    float* env = blob.data() + kSurModel;
    for (int c = 0; c < kMlpIn; ++c) {
        env[c] = (float)xmin[c];
        env[kMlpIn + c] = (float)xmax[c];
    }
    float* meta = env + 2 * kMlpIn;
    meta[kMetaVersion] = kSurVersion;
    meta[kMetaNearAlpha] = (float)o->near_alpha;
    meta[kMetaNearSubdiv] = (float)o->near_subdiv;
    meta[kMetaNearOrder] = (float)o->near_order;
    meta[kMetaQuadOrder] = (float)o->quad_order;
    meta[kMetaKlMax] = (float)cfg->kl_max;
    meta[kMetaDlMin] = (float)kSurDlMin;
    meta[kMetaDlMax] = (float)kSurDlMax;
    if (rep) *rep = R;
    return EFIE_OK;
}

}  // namespace efie
