/*
 * mc_fmath.h - deterministic single-precision elementary functions.
 *
 * The transport path needs log (optical-depth draw), sin and cos (isotropic
 * and Henyey-Greenstein directions).  Vendor libraries differ in the last
 * bit between host and device, which would make a packet history computed
 * on the GPU diverge from the same history computed on the CPU.  These
 * routines take a float, evaluate the function in IEEE double precision
 * with basic operations only (+ - * /, no fused multiply-add), and round
 * once to float, so host and device produce identical bits and the result
 * is the correctly rounded float except in rare hard cases.
 *
 * Methods (textbook, no vendor constants):
 *   log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
 *        s = (m-1)/(m+1), summed as the odd series 2(s + s^3/3 + s^5/5 ...)
 *        to s^25 (|s| <= 0.1716 so the truncation is below 1e-20).
 *   sin/cos: Cody-Waite reduction by pi/2 with a 33-bit head constant
 *        (so k*head is exact for |k| < 2^20) and a 53-bit tail, then the
 *        Taylor series to r^19 / r^20 on |r| <= pi/4.
 *   atan: |t| > 1 folded by atan(t) = pi/2 - atan(1/t), then two half-angle
 *        steps u -> u / (1 + sqrt(1 + u^2)) to |u| <= tan(pi/16), and the
 *        odd Taylor series to u^25 (truncation below 1e-18); IEEE sqrt is
 *        the one operation beyond + - * /.
 *   acos: acos(x) = 2 atan(sqrt((1-x)/(1+x))) on (-1, 1]; acos(-1) = pi.
 *
 * Must be compiled with floating-point contraction disabled.
 */
#ifndef MC_FMATH_H
#define MC_FMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define MCF_FN __host__ __device__ static inline
#else
#define MCF_FN static inline
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

typedef union { float f; uint32_t u; } mcf_fu;

MCF_FN float mcf_logf(float xf)
{
    mcf_fu v;
    v.f = xf;
    if (!(xf > 0.0f)) {                     /* zero, negative or NaN */
        if (xf == 0.0f) { v.u = 0xff800000u; return v.f; }   /* -inf */
        v.u = 0x7fc00000u; return v.f;                      /* NaN */
    }
    if (v.u >= 0x7f800000u) return xf;       /* +inf */
    int e = 0;
    if (v.u < 0x00800000u) {                 /* subnormal: scale by 2^24 */
        v.f = xf * 16777216.0f;
        e = -24;
    }
    e += (int)((v.u >> 23) & 0xffu) - 127;
    mcf_fu mv;
    mv.u = (v.u & 0x007fffffu) | 0x3f800000u;   /* m in [1, 2) */
    double m = (double)mv.f;
    if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 25.0;
    p = p * z + 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z;
    double t = s + s;
    double lm = t + t * p;
    double r = (double)e * 0.6931471805599453 + lm;
    return (float)r;
}

/* sin and cos of a float argument with |a| < 2^19. */
MCF_FN void mcf_sincosf(float af, float *sp, float *cp)
{
    const double pio2_head = 1.57079632673412561417e+00;  /* 33 significant bits */
    const double pio2_tail = 6.07710050650619224932e-11;  /* pi/2 - head */
    double a = (double)af;
    double kd = a * 0.63661977236758134308 + 0.5;
    double kf = (double)(int64_t)kd;
    if (kf > kd) kf = kf - 1.0;                             /* floor */
    int k = (int)kf;
    double r = (a - kf * pio2_head) - kf * pio2_tail;
    double z = r * r;
    double sp_ = 1.0 / 121645100408832000.0;               /* 1/19! */
    sp_ = -sp_ * z + 1.0 / 355687428096000.0;               /* 1/17! */
    sp_ = sp_ * z - 1.0 / 1307674368000.0;                  /* 1/15! */
    sp_ = sp_ * z + 1.0 / 6227020800.0;                     /* 1/13! */
    sp_ = sp_ * z - 1.0 / 39916800.0;                       /* 1/11! */
    sp_ = sp_ * z + 1.0 / 362880.0;                         /* 1/9!  */
    sp_ = sp_ * z - 1.0 / 5040.0;                           /* 1/7!  */
    sp_ = sp_ * z + 1.0 / 120.0;                            /* 1/5!  */
    sp_ = sp_ * z - 1.0 / 6.0;                              /* 1/3!  */
    double sinr = r + (r * z) * sp_;
    double cq = 1.0 / 2432902008176640000.0;                /* 1/20! */
    cq = -cq * z + 1.0 / 6402373705728000.0;                /* 1/18! */
    cq = cq * z - 1.0 / 20922789888000.0;                   /* 1/16! */
    cq = cq * z + 1.0 / 87178291200.0;                      /* 1/14! */
    cq = cq * z - 1.0 / 479001600.0;                        /* 1/12! */
    cq = cq * z + 1.0 / 3628800.0;                          /* 1/10! */
    cq = cq * z - 1.0 / 40320.0;                            /* 1/8!  */
    cq = cq * z + 1.0 / 720.0;                              /* 1/6!  */
    cq = cq * z - 1.0 / 24.0;                               /* 1/4!  */
    cq = cq * z + 0.5;                                      /* 1/2!  */
    double cosr = 1.0 - z * cq;
    double s, c;
    switch (k & 3) {
    case 0: s = sinr;  c = cosr;  break;
    case 1: s = cosr;  c = -sinr; break;
    case 2: s = -sinr; c = -cosr; break;
    default: s = -cosr; c = sinr; break;
    }
    *sp = (float)s;
    *cp = (float)c;
}

MCF_FN double mcf_atan_d(double t)
{
    if (t != t) return t;
    const double sgn = t < 0.0 ? -1.0 : 1.0;
    double a = t * sgn;
    const int inv = a > 1.0;
    if (inv) a = 1.0 / a;
    a = a / (1.0 + sqrt(1.0 + a * a));
    a = a / (1.0 + sqrt(1.0 + a * a));
    const double z = a * a;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 - p * z;
    p = 1.0 / 21.0 - p * z;
    p = 1.0 / 19.0 - p * z;
    p = 1.0 / 17.0 - p * z;
    p = 1.0 / 15.0 - p * z;
    p = 1.0 / 13.0 - p * z;
    p = 1.0 / 11.0 - p * z;
    p = 1.0 / 9.0 - p * z;
    p = 1.0 / 7.0 - p * z;
    p = 1.0 / 5.0 - p * z;
    p = 1.0 / 3.0 - p * z;
    p = 1.0 - p * z;
    double r = 4.0 * (a * p);
    if (inv) r = 1.5707963267948966 - r;
    return sgn * r;
}

MCF_FN float mcf_atanf(float x) { return (float)mcf_atan_d((double)x); }

MCF_FN float mcf_acosf(float xf)
{
    const double x = (double)xf;
    if (!(x >= -1.0 && x <= 1.0)) return (float)((x - x) / (x - x));   /* NaN */
    if (x == -1.0) return (float)3.141592653589793;
    return (float)(2.0 * mcf_atan_d(sqrt((1.0 - x) / (1.0 + x))));
}

#endif /* MC_FMATH_H */
