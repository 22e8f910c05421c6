// gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -- MEX entry point with the
// argument list of the reference gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST
// (gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:127-148), computed on MI355X.
//
//   [Z, Vrhs] = gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST(V, F, E, plusTri,
//               minusTri, plusSign, minusSign, rp, rm, Ap, Am, len, k,
//               omega, mu, eps, exc [, opts])
//
// exc = [patchEdge, groundEdge, V0, spreadRadius, selfAlpha]; selfAlpha is
// the self-softening factor of the centroid model (opts.model = 1).
//
// Build: mex -R2018a gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cpp -I<repo>/include -L<lib dir> -lefie_hip
#include "efie_mex_common.h"

using namespace efie_mex;

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    check(nrhs == 17 || nrhs == 18, "mex:args", "Expected 17 inputs.");
    check(nlhs == 2, "mex:args", "Expected 2 outputs [Z,Vrhs].");
    efie_opts o;
    read_opts(nrhs == 18 ? prhs[17] : nullptr, &o);
    double kr, ki, er, ei;
    complex_scalar(prhs[12], &kr, &ki);
    const double omega = mxGetScalar(prhs[13]);
    const double mu = mxGetScalar(prhs[14]);
    complex_scalar(prhs[15], &er, &ei);
    const mxArray* excMx = prhs[16];
    check_tables(prhs, "mex:E", "E must be [Ne x 2] int32.");
    check(mxIsDouble(excMx) && mxGetNumberOfElements(excMx) >= 5, "mex:exc",
          "exc must be double vector len>=5: [patchEdge, groundEdge, V0, spreadRadius, selfAlpha].");
    const mwSize Ne = mxGetM(prhs[2]);
    const double* exc = mxGetDoubles(excMx);
    check((int)exc[0] >= 1 && (int)exc[0] <= (int)Ne, "mex:port", "patchEdge out of range.");
    check((int)exc[1] >= 1 && (int)exc[1] <= (int)Ne, "mex:port", "groundEdge out of range.");
    efie_ctx* c = context();
    upload_tables(c, prhs);
    plhs[0] = mxCreateDoubleMatrix(Ne, Ne, mxCOMPLEX);
    plhs[1] = mxCreateDoubleMatrix(Ne, 1, mxCOMPLEX);
    status(efie_assemble_zv(c, kr, ki, omega, mu, er, ei, exc, (int)mxGetNumberOfElements(excMx), &o,
                            reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])),
                            reinterpret_cast<double*>(mxGetComplexDoubles(plhs[1]))));
}
