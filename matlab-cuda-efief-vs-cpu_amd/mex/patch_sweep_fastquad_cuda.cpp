// patch_sweep_fastquad_cuda.cpp -- MEX entry point with the argument list of
// the reference patch_sweep_fastquad_cuda (patch_sweep_fastquad_cuda.cu:
// 463-557): the patch_sweep_cuda sweep with the quadrature caches of
// efie_precompute_pack.m passed in as the struct `prepack` (argument 13),
// computed on MI355X through libefie_hip.so.
//
//   [Zin, S11] = patch_sweep_fastquad_cuda(V, F, E, plusTri, minusTri,
//                plusSign, minusSign, rp, rm, Ap, Am, len, prepack, freqs,
//                mu, eps, patchEdge, groundEdge, V0, spreadRadius, Z0
//                [, opts])
//
// Z is integrated with the prepack's points, weights, softening lengths,
// near matrix and near_alpha (the kernel of patch_sweep_fastquad_cuda.cu:
// 336-411); opts may set precision, batch, ir_tol, ir_max_iter, lu_kernel
// and near_mode (DIRECT or MOMENTS).
//
// Build: mex -R2018a patch_sweep_fastquad_cuda.cpp -I<repo>/include -L<lib dir> -lefie_hip
#include "efie_mex_common.h"

using namespace efie_mex;

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    check(nrhs == 21 || nrhs == 22, "mex:args", "Expected 21 inputs.");
    check(nlhs == 2, "mex:args", "Expected 2 outputs [Zin,S11].");
    efie_opts o;
    read_opts(nrhs == 22 ? prhs[21] : nullptr, &o);
    const mxArray* prepackMx = prhs[12];
    const mxArray* freqsMx = prhs[13];
    const double mu = mxGetScalar(prhs[14]);
    double er, ei;
    complex_scalar(prhs[15], &er, &ei);
    const int patchEdge = (int)mxGetScalar(prhs[16]);
    const int groundEdge = (int)mxGetScalar(prhs[17]);
    const double V0 = mxGetScalar(prhs[18]);
    const double spread = mxGetScalar(prhs[19]);
    const double Z0 = mxGetScalar(prhs[20]);
    check_tables(prhs, "mex:E", "E must be [Ne x 2] int32.");
    const mwSize Nt = mxGetM(prhs[1]), Ne = mxGetM(prhs[2]);
    check(mxIsDouble(freqsMx), "mex:freqs", "freqs must be double vector.");
    const mwSize Nf = mxGetNumberOfElements(freqsMx);
    check(mxIsStruct(prepackMx), "mex:prepack", "prepack must be struct.");
    check(patchEdge >= 1 && patchEdge <= (int)Ne, "mex:port", "patchEdge out of range.");
    check(groundEdge >= 1 && groundEdge <= (int)Ne, "mex:port", "groundEdge out of range.");
    efie_prepack pk;
    read_prepack(prepackMx, Nt, true, &pk);
    efie_ctx* c = context();
    upload_tables(c, prhs);
    status(efie_set_prepack(c, &pk));
    plhs[0] = mxCreateDoubleMatrix(Nf, 1, mxCOMPLEX);
    plhs[1] = mxCreateDoubleMatrix(Nf, 1, mxCOMPLEX);
    status(efie_sweep(c, mxGetDoubles(freqsMx), (int)Nf, mu, er, ei, patchEdge, groundEdge, V0, spread, Z0, &o,
                      reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])),
                      reinterpret_cast<double*>(mxGetComplexDoubles(plhs[1])), nullptr));
}
