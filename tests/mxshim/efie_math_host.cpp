// Host build of the device sine/cosine (efie_math.h) for its error-bound test.
#include "../../matlab-cuda-efief-vs-cpu_amd/csrc/efie_math.h"

extern "C" void efie_math_sincos(const double* x, int n, double* s, double* c) {
    for (int i = 0; i < n; ++i) efie::sincos_cw(x[i], s + i, c + i);
}

// which = 0: sincos_small (|t| <= 0.75), 1: sincos_small_short (|t| <= 0.2)
extern "C" void efie_math_sincos_small(const double* x, int n, double* s, double* c, int which) {
    for (int i = 0; i < n; ++i) {
        if (which == 0) efie::sincos_small(x[i], s + i, c + i);
        else efie::sincos_small_short(x[i], s + i, c + i);
    }
}
