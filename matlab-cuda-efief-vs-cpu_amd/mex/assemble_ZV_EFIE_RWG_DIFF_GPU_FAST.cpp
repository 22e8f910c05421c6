// assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -- MEX entry point with the
// argument list of the reference assemble_ZV_EFIE_RWG_DIFF_GPU_FAST
// (assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:277-392): [Z, Vrhs] of one
// frequency from the rwg struct (RWG tables, no mesh arrays) and the
// quadrature caches of efie_precompute_pack.m, computed on MI355X.
//
//   [Z, Vrhs] = assemble_ZV_EFIE_RWG_DIFF_GPU_FAST(rwg, prepack, k, omega,
//               mu, eps, exc, opts)
//
// rwg: plusTri, minusTri, plusSign, minusSign (int32), len, Ap, Am, rp, rm,
// center, zSign; exc: patchEdge, groundEdge, V0, spreadRadius; opts:
// near_alpha (default 0.15; the near softening of this entry point comes
// from opts, not from the prepack) and the extension fields of read_opts.
// The self softening lengths are read from prepack.self_a (the field
// efie_precompute_pack.m writes) or, failing that, prepack.self_aSoft (the
// name the reference source reads).  The triangles' centroids and areas
// come from the prepack's regular points.  Vrhs is the Gaussian strip over
// rwg.center on the port's zSign layer (.cu:584-623).
//
// Build: mex -R2018a assemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -I<repo>/include -L<lib dir> -lefie_hip
#include <math.h>

#include <vector>

#include "efie_mex_common.h"

using namespace efie_mex;

namespace {

const mxArray* field(const mxArray* s, const char* name) { return mxIsStruct(s) ? mxGetField(s, 0, name) : nullptr; }

// double or int32 vector element e
double elem(const mxArray* a, mwSize e) {
    if (mxIsInt32(a)) return (double)mxGetInt32s(a)[e];
    return mxGetDoubles(a)[e];
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    check(nrhs == 8, "mex:args", "Expected 8 inputs: (rwg, prepack, k, omega, mu, eps, exc, opts)");
    check(nlhs == 2, "mex:args", "Expected 2 outputs: [Z, Vrhs]");
    const mxArray *rwgS = prhs[0], *preS = prhs[1], *excS = prhs[6], *optS = prhs[7];
    double kr, ki, er, ei;
    complex_scalar(prhs[2], &kr, &ki);
    const double omega = mxGetScalar(prhs[3]);
    const double mu = mxGetScalar(prhs[4]);
    complex_scalar(prhs[5], &er, &ei);
    efie_opts o;
    read_opts(mxIsStruct(optS) ? optS : nullptr, &o);

    const mxArray *pT = field(rwgS, "plusTri"), *mT = field(rwgS, "minusTri");
    const mxArray *pS = field(rwgS, "plusSign"), *mS = field(rwgS, "minusSign");
    check(pT && mT && pS && mS, "mex:rwg", "Missing tri/sign fields");
    check(mxIsInt32(pT) && mxIsInt32(mT), "mex:rwg", "plusTri/minusTri must be int32");
    check(mxIsInt32(pS) && mxIsInt32(mS), "mex:rwg", "plusSign/minusSign must be int32");
    const mwSize Ne = mxGetNumberOfElements(pT);
    const mxArray *lenMx = field(rwgS, "len"), *ApMx = field(rwgS, "Ap"), *AmMx = field(rwgS, "Am");
    const mxArray *rpMx = field(rwgS, "rp"), *rmMx = field(rwgS, "rm");
    const mxArray *cMx = field(rwgS, "center"), *zMx = field(rwgS, "zSign");
    check(lenMx && ApMx && AmMx && rpMx && rmMx && cMx && zMx, "mex:rwg",
          "rwg must have len, Ap, Am, rp, rm, center, zSign");
    check(mxGetNumberOfElements(mT) == Ne && mxGetNumberOfElements(pS) == Ne && mxGetNumberOfElements(mS) == Ne,
          "mex:rwg", "Missing tri/sign fields");
    check(mxIsDouble(rpMx) && mxGetN(rpMx) == 3 && mxGetM(rpMx) == Ne, "mex:rwg", "rp must be Ne x 3");
    check(mxIsDouble(rmMx) && mxGetN(rmMx) == 3 && mxGetM(rmMx) == Ne, "mex:rwg", "rm must be Ne x 3");
    check(mxIsDouble(lenMx) && mxGetNumberOfElements(lenMx) == Ne && mxIsDouble(ApMx) &&
              mxGetNumberOfElements(ApMx) == Ne && mxIsDouble(AmMx) && mxGetNumberOfElements(AmMx) == Ne,
          "mex:rwg", "len, Ap, Am must be Ne x 1 double");
    check(mxIsDouble(cMx) && mxGetM(cMx) == Ne && mxGetN(cMx) == 3, "mex:rwg", "center must be Ne x 3");
    check((mxIsDouble(zMx) || mxIsInt32(zMx)) && mxGetNumberOfElements(zMx) == Ne, "mex:rwg",
          "zSign must be Ne x 1");

    const mxArray *nmMx = field(preS, "nearMat"), *rregMx = field(preS, "r_reg"), *wregMx = field(preS, "w_reg");
    const mxArray* aeqMx = field(preS, "aeqTri");
    check(nmMx && rregMx && wregMx && aeqMx, "mex:pre", "Missing prepack basic fields");
    check(mxIsUint8(nmMx), "mex:pre", "nearMat must be uint8");
    check(mxIsDouble(rregMx) && mxIsDouble(wregMx), "mex:pre", "r_reg/w_reg must be double");
    check(mxGetNumberOfDimensions(rregMx) == 3, "mex:pre", "r_reg must be Nq x 3 x Nt");
    const mwSize* dr = mxGetDimensions(rregMx);
    check(dr[1] == 3, "mex:pre", "r_reg second dim must be 3");
    const mwSize Nt = dr[2];
    check(mxGetM(wregMx) == dr[0], "mex:pre", "w_reg rows must match Nq");
    check(mxGetN(wregMx) == Nt, "mex:pre", "w_reg cols must be Nt");
    check(field(preS, "self_r") && field(preS, "self_w") && (field(preS, "self_a") || field(preS, "self_aSoft")) &&
              field(preS, "near_r") && field(preS, "near_w"),
          "mex:pre", "Missing self/near cached fields.");
    efie_prepack pk;
    read_prepack(preS, Nt, false, &pk);
    pk.near_alpha = o.near_alpha;

    const char* excMsg = "exc must have patchEdge, groundEdge, V0, spreadRadius";
    const mxArray *peMx = field(excS, "patchEdge"), *geMx = field(excS, "groundEdge");
    const mxArray *v0Mx = field(excS, "V0"), *spMx = field(excS, "spreadRadius");
    check(peMx && geMx && v0Mx && spMx, "mex:exc", excMsg);
    const int pe = (int)mxGetScalar(peMx), ge = (int)mxGetScalar(geMx);
    const double V0 = mxGetScalar(v0Mx), spread = mxGetScalar(spMx);
    check(pe >= 1 && pe <= (int)Ne, "mex:port", "patchEdge out of range.");
    check(ge >= 1 && ge <= (int)Ne, "mex:port", "groundEdge out of range.");

    // Vrhs: +V0/2 and -V0/2 Gaussian strips over rwg.center on the port's
    // zSign layer, each normalised to unit sum
    const double* cen = mxGetDoubles(cMx);
    std::vector<double> vr(Ne, 0.0), w(Ne);
    for (int k = 0; k < 2; ++k) {
        const int e0 = (k == 0 ? pe : ge) - 1;
        const double amp = (k == 0 ? 0.5 : -0.5) * V0;
        const int layer0 = elem(zMx, e0) > 0 ? 1 : -1;
        double sumw = 0.0;
        for (mwSize e = 0; e < Ne; ++e) {
            const int layer = elem(zMx, e) > 0 ? 1 : -1;
            if (layer != layer0) {
                w[e] = 0.0;
                continue;
            }
            const double dx = cen[e] - cen[e0], dy = cen[e + Ne] - cen[e0 + Ne], dz = cen[e + 2 * Ne] - cen[e0 + 2 * Ne];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            w[e] = exp(-(d / spread) * (d / spread));
            sumw += w[e];
        }
        check(sumw >= 1e-30, "mex:rhs", "spreadRadius too small (no weights)");
        for (mwSize e = 0; e < Ne; ++e) vr[e] += amp * (w[e] / sumw);
    }

    efie_ctx* c = context();
    status(efie_set_rwg_tables(c, (int)Nt, mxGetInt32s(pT), mxGetInt32s(mT), mxGetInt32s(pS), mxGetInt32s(mS),
                               mxGetDoubles(rpMx), mxGetDoubles(rmMx), mxGetDoubles(ApMx), mxGetDoubles(AmMx),
                               mxGetDoubles(lenMx), (int)Ne));
    status(efie_set_prepack(c, &pk));
    plhs[0] = mxCreateDoubleMatrix(Ne, Ne, mxCOMPLEX);
    status(efie_assemble_z(c, kr, ki, omega, mu, er, ei, &o, reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0]))));
    plhs[1] = mxCreateDoubleMatrix(Ne, 1, mxCOMPLEX);
    mxComplexDouble* V = mxGetComplexDoubles(plhs[1]);
    for (mwSize e = 0; e < Ne; ++e) {
        V[e].real = vr[e];
        V[e].imag = 0.0;
    }
}
