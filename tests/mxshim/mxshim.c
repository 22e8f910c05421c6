/* mxshim.c -- implementation of the minimal mx API of mex.h for tests.
 * Errors raised with mexErrMsgIdAndTxt unwind to shim_call() through
 * longjmp, as MATLAB does, and are reported as (id, message). */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

enum { CLS_DOUBLE = 1, CLS_INT32 = 2, CLS_STRUCT = 3, CLS_UINT8 = 4 };

/* dims[] holds every dimension (MATLAB drops trailing singletons beyond
 * the second); m = dims[0], n = the product of the others */
struct mxArray_tag {
    int cls;
    size_t m, n;
    int cplx;
    void* data;
    int nfields;
    char** names;
    mxArray** values;
    size_t ndim;
    size_t dims[8];
};

static jmp_buf g_jmp;
static char g_id[256], g_msg[1024];
static void (*g_exit_fns[16])(void);
static int g_n_exit;

static mxArray* alloc(int cls, size_t m, size_t n, int cplx, size_t esz) {
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = cls;
    a->m = m;
    a->n = n;
    a->cplx = cplx;
    a->ndim = 2;
    a->dims[0] = m;
    a->dims[1] = n;
    size_t cnt = m * n;
    a->data = calloc(cnt > 0 ? cnt : 1, esz * (cplx ? 2 : 1));
    return a;
}

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) { return alloc(CLS_DOUBLE, m, n, c == mxCOMPLEX, 8); }
mwSize mxGetM(const mxArray* a) { return a->m; }
mwSize mxGetN(const mxArray* a) { return a->n; }
mwSize mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
int mxIsDouble(const mxArray* a) { return a->cls == CLS_DOUBLE; }
int mxIsInt32(const mxArray* a) { return a->cls == CLS_INT32; }
int mxIsComplex(const mxArray* a) { return a->cplx; }
int mxIsStruct(const mxArray* a) { return a->cls == CLS_STRUCT; }
int mxIsUint8(const mxArray* a) { return a->cls == CLS_UINT8; }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->ndim; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }
mxUint8* mxGetUint8s(const mxArray* a) { return a->cls == CLS_UINT8 ? (mxUint8*)a->data : NULL; }
double mxGetScalar(const mxArray* a) {
    if (a->m * a->n == 0) return 0.0;
    if (a->cls == CLS_INT32) return (double)((int32_t*)a->data)[0];
    if (a->cls == CLS_DOUBLE) return ((double*)a->data)[0];
    if (a->cls == CLS_UINT8) return (double)((uint8_t*)a->data)[0];
    return 0.0;
}
double* mxGetDoubles(const mxArray* a) { return (a->cls == CLS_DOUBLE && !a->cplx) ? (double*)a->data : NULL; }
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
    return (a->cls == CLS_DOUBLE && a->cplx) ? (mxComplexDouble*)a->data : NULL;
}
mxInt32* mxGetInt32s(const mxArray* a) { return a->cls == CLS_INT32 ? (mxInt32*)a->data : NULL; }
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name) {
    (void)i;
    if (a->cls != CLS_STRUCT) return NULL;
    for (int k = 0; k < a->nfields; ++k)
        if (!strcmp(a->names[k], name)) return a->values[k];
    return NULL;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    snprintf(g_id, sizeof g_id, "%s", id);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}
int mexAtExit(void (*fn)(void)) {
    if (g_n_exit < 16) g_exit_fns[g_n_exit++] = fn;
    return 0;
}

/* ---- test driver side ---- */
mxArray* shim_double(size_t m, size_t n, int cplx, const double* src) {
    mxArray* a = alloc(CLS_DOUBLE, m, n, cplx, 8);
    if (src) memcpy(a->data, src, m * n * 8 * (cplx ? 2 : 1));
    return a;
}
mxArray* shim_int32(size_t m, size_t n, const int32_t* src) {
    mxArray* a = alloc(CLS_INT32, m, n, 0, 4);
    if (src) memcpy(a->data, src, m * n * 4);
    return a;
}
/* an N-d array (ndim <= 8) of class cls (1 double, 2 int32, 4 uint8),
 * column-major data; trailing singleton dimensions beyond the second are
 * dropped as MATLAB does */
mxArray* shim_nd(int cls, size_t ndim, const size_t* dims, const void* src) {
    size_t nd = ndim, cnt = 1;
    while (nd > 2 && dims[nd - 1] == 1) --nd;
    for (size_t k = 0; k < ndim; ++k) cnt *= dims[k];
    const size_t esz = cls == CLS_DOUBLE ? 8 : (cls == CLS_INT32 ? 4 : 1);
    mxArray* a = alloc(cls, dims[0], ndim > 0 ? cnt / (dims[0] ? dims[0] : 1) : 1, 0, esz);
    a->ndim = nd < 2 ? 2 : nd;
    for (size_t k = 0; k < 8; ++k) a->dims[k] = k < ndim ? dims[k] : 1;
    if (src) memcpy(a->data, src, cnt * esz);
    return a;
}
mxArray* shim_struct(int nf, const char** names, mxArray** values) {
    mxArray* a = alloc(CLS_STRUCT, 1, 1, 0, 1);
    a->nfields = nf;
    a->names = (char**)calloc(nf ? nf : 1, sizeof(char*));
    a->values = (mxArray**)calloc(nf ? nf : 1, sizeof(mxArray*));
    for (int k = 0; k < nf; ++k) {
        size_t L = strlen(names[k]) + 1;
        a->names[k] = (char*)malloc(L);
        memcpy(a->names[k], names[k], L);
        a->values[k] = values[k];
    }
    return a;
}
void* shim_data(const mxArray* a) { return a->data; }
int shim_is_complex(const mxArray* a) { return a->cplx; }
void shim_free(mxArray* a) {
    if (!a) return;
    for (int k = 0; k < a->nfields; ++k) {
        free(a->names[k]);
        shim_free(a->values[k]);
    }
    free(a->names);
    free(a->values);
    free(a->data);
    free(a);
}
typedef void (*mexfn)(int, mxArray**, int, const mxArray**);
/* returns 0 on success, 1 when the entry point raised an error */
int shim_call(mexfn f, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs) {
    g_id[0] = g_msg[0] = 0;
    if (setjmp(g_jmp)) return 1;
    f(nlhs, plhs, nrhs, prhs);
    return 0;
}
const char* shim_err_id(void) { return g_id; }
const char* shim_err_msg(void) { return g_msg; }
void shim_run_exit(void) {
    while (g_n_exit > 0) g_exit_fns[--g_n_exit]();
}
