/* mex.h -- minimal stand-in for the MATLAB MEX/mx C API, written for this
 * project's tests (no MathWorks or Octave header is available here).  It
 * declares only the functions the EFIE MEX entry points use, with the
 * interleaved-complex (R2018a) semantics, so the entry points can run
 * outside MATLAB.  Implemented in mxshim.c. */
#ifndef EFIE_MXSHIM_MEX_H
#define EFIE_MXSHIM_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef int32_t mxInt32;
typedef uint8_t mxUint8;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct {
    double real, imag;
} mxComplexDouble;
typedef struct mxArray_tag mxArray;

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mwSize mxGetM(const mxArray* a);
mwSize mxGetN(const mxArray* a);
mwSize mxGetNumberOfElements(const mxArray* a);
int mxIsDouble(const mxArray* a);
int mxIsInt32(const mxArray* a);
int mxIsComplex(const mxArray* a);
int mxIsStruct(const mxArray* a);
int mxIsUint8(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
mxUint8* mxGetUint8s(const mxArray* a);
double mxGetScalar(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
mxInt32* mxGetInt32s(const mxArray* a);
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));
/* the entry point every MEX file defines (C linkage) */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif
#endif
