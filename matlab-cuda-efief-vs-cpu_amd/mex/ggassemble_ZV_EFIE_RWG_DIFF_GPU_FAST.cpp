// ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -- MEX entry point with the
// argument list of the reference ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST
// (ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:542-641): [Z, Vrhs] of one
// frequency with the quadrature caches of efie_precompute_pack.m passed in
// as the struct `prepack` (argument 13), computed on MI355X.
//
//   [Z, Vrhs] = ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST(V, F, E, plusTri,
//               minusTri, plusSign, minusSign, rp, rm, Ap, Am, len, prepack,
//               k, omega, mu, eps, excvec [, opts])
//
// excvec = [patchEdge, groundEdge, V0, spreadRadius]; Vrhs is the Gaussian
// edge-strip excitation over the edge midpoints (.cu:495-540).  opts may
// set near_mode (DIRECT or MOMENTS).
//
// Build: mex -R2018a ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -I<repo>/include -L<lib dir> -lefie_hip
#include "efie_mex_common.h"

using namespace efie_mex;

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    check(nrhs == 18 || nrhs == 19, "mex:args",
          "Expected 18 inputs: V,F,E,plusTri,minusTri,plusSign,minusSign,rp,rm,Ap,Am,len,prepack,k,omega,mu,eps,"
          "excvec");
    check(nlhs == 2, "mex:args", "Expected 2 outputs [Z,Vrhs].");
    efie_opts o;
    read_opts(nrhs == 19 ? prhs[18] : nullptr, &o);
    const mxArray* prepackMx = prhs[12];
    double kr, ki, er, ei;
    complex_scalar(prhs[13], &kr, &ki);
    const double omega = mxGetScalar(prhs[14]);
    const double mu = mxGetScalar(prhs[15]);
    complex_scalar(prhs[16], &er, &ei);
    const mxArray* excMx = prhs[17];
    check_tables(prhs, "mex:E", "E must be [Ne x 2] int32.");
    const mwSize Nt = mxGetM(prhs[1]), Ne = mxGetM(prhs[2]);
    check(mxIsStruct(prepackMx), "mex:prepack", "prepack must be a struct from efie_precompute_pack.");
    check(mxIsDouble(excMx) && mxGetNumberOfElements(excMx) >= 4, "mex:exc",
          "excvec must be [patchEdge, groundEdge, V0, spreadRadius].");
    const double* exc = mxGetDoubles(excMx);
    check((int)exc[0] >= 1 && (int)exc[0] <= (int)Ne, "mex:port", "patchEdge out of range.");
    check((int)exc[1] >= 1 && (int)exc[1] <= (int)Ne, "mex:port", "groundEdge out of range.");
    efie_prepack pk;
    read_prepack(prepackMx, Nt, true, &pk);
    const double dr = -omega * ei, di = omega * er;
    check(dr * dr + di * di > 0.0, "mex:eps", "eps causes denom=0.");
    efie_ctx* c = context();
    upload_tables(c, prhs);
    status(efie_set_prepack(c, &pk));
    // the self-softening slot of exc belongs to the centroid model only
    const double exc5[5] = {exc[0], exc[1], exc[2], exc[3], 0.20};
    plhs[0] = mxCreateDoubleMatrix(Ne, Ne, mxCOMPLEX);
    plhs[1] = mxCreateDoubleMatrix(Ne, 1, mxCOMPLEX);
    status(efie_assemble_zv(c, kr, ki, omega, mu, er, ei, exc5, 5, &o,
                            reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])),
                            reinterpret_cast<double*>(mxGetComplexDoubles(plhs[1]))));
}
