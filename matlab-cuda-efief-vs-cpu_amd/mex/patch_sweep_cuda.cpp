// patch_sweep_cuda.cpp -- MEX entry point with the argument list of the
// reference patch_sweep_cuda (patch_sweep_cuda.cu:4-27), computed on
// MI355X through libefie_hip.so.
//
//   [Zin, S11] = patch_sweep_cuda(V, F, edge, plusTri, minusTri, plusSign,
//                                 minusSign, rp, rm, Ap, Am, len, freqs, mu,
//                                 eps, patchEdge, groundEdge, V0,
//                                 spreadRadius, Z0 [, opts])
//
// Z is filled with the quadrature scheme of assemble_ZV_EFIE_RWG_DIFF_CPU.m
// (accurate preset unless opts says otherwise); opts.model = 1 selects the
// reference's centroid kernel and its Iport = x(p) - x(g) epilogue.
//
// Build: mex -R2018a patch_sweep_cuda.cpp -I<repo>/include -L<lib dir> -lefie_hip
#include "efie_mex_common.h"

using namespace efie_mex;

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    check(nrhs == 20 || nrhs == 21, "mex:args", "Expected 20 inputs.");
    check(nlhs == 2, "mex:args", "Expected 2 outputs [Zin,S11].");
    efie_opts o;
    read_opts(nrhs == 21 ? prhs[20] : nullptr, &o);
    const mxArray* freqsMx = prhs[12];
    const double mu = mxGetScalar(prhs[13]);
    double er, ei;
    complex_scalar(prhs[14], &er, &ei);
    const int patchEdge = (int)mxGetScalar(prhs[15]);
    const int groundEdge = (int)mxGetScalar(prhs[16]);
    const double V0 = mxGetScalar(prhs[17]);
    const double spread = mxGetScalar(prhs[18]);
    const double Z0 = mxGetScalar(prhs[19]);
    check_tables(prhs, "mex:edge", "edge must be [Ne x 2] int32.");
    check(mxIsDouble(freqsMx), "mex:freqs", "freqs must be double vector.");
    const mwSize Nf = mxGetNumberOfElements(freqsMx);
    const mwSize Ne = mxGetM(prhs[2]);
    check(patchEdge >= 1 && patchEdge <= (int)Ne, "mex:port", "patchEdge out of range.");
    check(groundEdge >= 1 && groundEdge <= (int)Ne, "mex:port", "groundEdge out of range.");
    efie_ctx* c = context();
    upload_tables(c, prhs);
    plhs[0] = mxCreateDoubleMatrix(Nf, 1, mxCOMPLEX);
    plhs[1] = mxCreateDoubleMatrix(Nf, 1, mxCOMPLEX);
    status(efie_sweep(c, mxGetDoubles(freqsMx), (int)Nf, mu, er, ei, patchEdge, groundEdge, V0, spread, Z0, &o,
                      reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])),
                      reinterpret_cast<double*>(mxGetComplexDoubles(plhs[1])), nullptr));
}
