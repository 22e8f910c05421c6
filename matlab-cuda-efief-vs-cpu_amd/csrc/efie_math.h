// efie_math.h -- sine/cosine for the Green's function phase, usable from
// device and host code (the host build is what the error-bound test runs).
//
// Method: Cody-Waite reduction x = n (pi/2) + r with pi/2 split into three
// parts (30 + 30 significant bits + the rest, derived from 60 digits of pi),
// so n*C1 and n*C2 are exact for |n| < 2^23, then the Taylor series of sin
// and cos on |r| <= pi/4 through r^21 / r^22 (truncation below 2e-22), and
// the quadrant from n mod 4.  Intended for |x| < 2^20 (the phase k R of
// this code stays below a few hundred); the error is a few ulp.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define EFIE_HD __host__ __device__
#else
#define EFIE_HD
#endif

namespace efie {

EFIE_HD inline void sincos_cw(double x, double* s, double* c) {
    const double kTwoOverPi = 0.6366197723675814;
    const double C1 = 1.5707963276654482, C2 = -8.705515692000731e-10, C3 = -3.50343439808993e-19;
    const double n = rint(x * kTwoOverPi);
    const double r = ((x - n * C1) - n * C2) - n * C3;
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ...), cos r = 1 - r^2/2! + r^4/4! - ...
    double ps = -1.9572941063391263e-20;
    ps = ps * r2 + 8.22063524662433e-18;
    ps = ps * r2 - 2.8114572543455206e-15;
    ps = ps * r2 + 7.647163731819816e-13;
    ps = ps * r2 - 1.6059043836821613e-10;
    ps = ps * r2 + 2.505210838544172e-08;
    ps = ps * r2 - 2.7557319223985893e-06;
    ps = ps * r2 + 0.0001984126984126984;
    ps = ps * r2 - 0.008333333333333333;
    ps = ps * r2 + 0.16666666666666666;
    const double sr = r - r * r2 * ps;
    double pc = 8.896791392450574e-22;
    pc = pc * r2 - 4.110317623312165e-19;
    pc = pc * r2 + 1.5619206968586225e-16;
    pc = pc * r2 - 4.779477332387385e-14;
    pc = pc * r2 + 1.1470745597729725e-11;
    pc = pc * r2 - 2.08767569878681e-09;
    pc = pc * r2 + 2.755731922398589e-07;
    pc = pc * r2 - 2.48015873015873e-05;
    pc = pc * r2 + 0.001388888888888889;
    pc = pc * r2 - 0.041666666666666664;
    pc = pc * r2 + 0.5;
    const double cr = 1.0 - r2 * pc;
    const int q = ((int)n) & 3;
    // x = n pi/2 + r:  q=0 (s,c)=(sr,cr), 1 (cr,-sr), 2 (-sr,-cr), 3 (-cr,sr)
    const double ss = (q & 1) ? cr : sr;
    const double cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// cos t and sin t for |t| <= 0.75 by their Taylor series through t^18 and
// t^19 (truncation below 1e-18).
EFIE_HD inline void sincos_small(double t, double* s, double* c) {
    const double t2 = t * t;
    double pc = -1.5619206968586225e-16;
    pc = pc * t2 + 4.779477332387385e-14;
    pc = pc * t2 - 1.1470745597729725e-11;
    pc = pc * t2 + 2.08767569878681e-09;
    pc = pc * t2 - 2.755731922398589e-07;
    pc = pc * t2 + 2.48015873015873e-05;
    pc = pc * t2 - 0.001388888888888889;
    pc = pc * t2 + 0.041666666666666664;
    pc = pc * t2 - 0.5;
    *c = pc * t2 + 1.0;
    double ps = -8.22063524662433e-18;
    ps = ps * t2 + 2.8114572543455206e-15;
    ps = ps * t2 - 7.647163731819816e-13;
    ps = ps * t2 + 1.6059043836821613e-10;
    ps = ps * t2 - 2.505210838544172e-08;
    ps = ps * t2 + 2.7557319223985893e-06;
    ps = ps * t2 - 0.0001984126984126984;
    ps = ps * t2 + 0.008333333333333333;
    ps = ps * t2 - 0.16666666666666666;
    *s = t + t * t2 * ps;
}

// cos t and sin t for |t| <= 0.2 by their Taylor series through t^10 and
// t^11 (truncation below 1e-17).
EFIE_HD inline void sincos_small_short(double t, double* s, double* c) {
    const double t2 = t * t;
    double pc = -2.755731922398589e-07;
    pc = pc * t2 + 2.48015873015873e-05;
    pc = pc * t2 - 0.001388888888888889;
    pc = pc * t2 + 0.041666666666666664;
    pc = pc * t2 - 0.5;
    *c = pc * t2 + 1.0;
    double ps = -2.505210838544172e-08;
    ps = ps * t2 + 2.7557319223985893e-06;
    ps = ps * t2 - 0.0001984126984126984;
    ps = ps * t2 + 0.008333333333333333;
    ps = ps * t2 - 0.16666666666666666;
    *s = t + t * t2 * ps;
}

}  // namespace efie
