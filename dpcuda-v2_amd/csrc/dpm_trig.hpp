// dpm_trig.hpp -- sine and cosine used for rigid-body rotations, compiled into both
// the gfx950 kernels and the CPU oracle so that rotated vertex positions carry the same
// bits on both sides (the library sin/cos of the device and of the host differ in the
// last ulp, which FIRE and NVE amplify).
//
// Method.  Cody-Waite argument reduction (W. J. Cody and W. Waite, "Software Manual for
// the Elementary Functions", 1980): k = rint(x * 2/pi) and
// r = ((x - k*P1) - k*P2) - k*P3 with pi/2 = P1 + P2 + P3, where P1 and P2 carry 33
// significant bits, so k*P1 and k*P2 are exact for |k| < 2^20.  sin r and cos r come
// from their Taylor series about 0 truncated after r^17 and r^18 (on |r| <= pi/4 the
// first omitted terms are below 2^-62 relative), evaluated by Horner's rule; cos r adds
// the rounding error of 1 - r^2/2 back in.  The quadrant k mod 4 selects and signs the
// pair.  Constants: pi/2 from Machin's formula split into three doubles, 1/n! rounded
// to double.  Measured against long-double sinl/cosl: at most 0.75 ulp for |x| <= pi/4,
// 1.5 ulp for |x| <= 100, 2.5 ulp for |x| <= 1e6.
//
// Build with -ffp-contract=off on both sides (no fused multiply-adds).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define DPM_TRIG_HD __host__ __device__
#else
#define DPM_TRIG_HD
#endif

DPM_TRIG_HD inline void dpm_sincos(double x, double* sin_out, double* cos_out) {
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb544p+0;
    const double P2 = 0x1.0b4611a6p-34;
    const double P3 = 0x1.3198a2e037073p-69;
    double k = rint(x * TWO_OVER_PI);
    double r = ((x - k * P1) - k * P2) - k * P3;
    double r2 = r * r;
    // sin r = r + r^3 * (-1/3! + r^2/5! - r^4/7! + ... + r^14/17!)
    double ps = 0x1.952c77030ad4ap-49;    // 1/17!
    ps = ps * r2 - 0x1.ae7f3e733b81fp-41;  // 1/15!
    ps = ps * r2 + 0x1.6124613a86d09p-33;  // 1/13!
    ps = ps * r2 - 0x1.ae64567f544e4p-26;  // 1/11!
    ps = ps * r2 + 0x1.71de3a556c734p-19;  // 1/9!
    ps = ps * r2 - 0x1.a01a01a01a01ap-13;  // 1/7!
    ps = ps * r2 + 0x1.1111111111111p-7;   // 1/5!
    ps = ps * r2 - 0x1.5555555555555p-3;   // 1/3!
    double sr = r + (r * r2) * ps;
    // cos r = 1 - r^2/2 + r^4 * (1/4! - r^2/6! + ... - r^14/18!)
    double pc = -0x1.6827863b97d97p-53;    // 1/18!
    pc = pc * r2 + 0x1.ae7f3e733b81fp-45;  // 1/16!
    pc = pc * r2 - 0x1.93974a8c07c9dp-37;  // 1/14!
    pc = pc * r2 + 0x1.1eed8eff8d898p-29;  // 1/12!
    pc = pc * r2 - 0x1.27e4fb7789f5cp-22;  // 1/10!
    pc = pc * r2 + 0x1.a01a01a01a01ap-16;  // 1/8!
    pc = pc * r2 - 0x1.6c16c16c16c17p-10;  // 1/6!
    pc = pc * r2 + 0x1.5555555555555p-5;   // 1/4!
    double hz = 0.5 * r2;
    double w = 1.0 - hz;                   // hz <= 0.31: 1 - w is exact
    double cr = w + (((1.0 - w) - hz) + (r2 * r2) * pc);
    long long q = (long long)k & 3;
    double s, c;
    if (q == 0) {
        s = sr;
        c = cr;
    } else if (q == 1) {
        s = cr;
        c = -sr;
    } else if (q == 2) {
        s = -sr;
        c = -cr;
    } else {
        s = -cr;
        c = sr;
    }
    *sin_out = x == 0.0 ? x : s;  // sin(-0) = -0
    *cos_out = c;
}
