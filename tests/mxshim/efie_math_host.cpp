// Host build of the device sine/cosine (efie_math.h) for its error-bound test.
#include "../../matlab-cuda-efief-vs-cpu_amd/csrc/efie_math.h"

extern "C" void efie_math_sincos(const double* x, int n, double* s, double* c) {
    for (int i = 0; i < n; ++i) efie::sincos_cw(x[i], s + i, c + i);
}
