/*
 * ctsm_math.h -- deterministic binary64 log / exp / pow / log10 / atan.
 *
 * One implementation, compiled for the gfx950 kernels and for host code, so
 * both produce the same bits for the same arguments: only IEEE operations
 * that are exactly specified (+ - * / with round-to-nearest, fma(), frexp,
 * ldexp, rint) are used, and callers compile with -ffp-contract=off.  The
 * physics the library reproduces is sensitive to the last bit of x**y in
 * ill-conditioned columns (Clapp-Hornberger suction reaches 1e15 mm), so
 * the GPU and the CPU check must agree on the elementary functions.
 *
 * Method (textbook argument reduction with double-double intermediates):
 *   log(x): x = 2^e m, m in [sqrt(1/2), sqrt(2)); f = (m-1)/(m+1) in
 *           double-double; log m = 2 atanh f = 2f + 2f^3/3 + f^5 sum_k
 *           2 f^(2k)/(2k+5), the first two terms in double-double, the tail
 *           (|f|^5 <= 1.5e-4) by Horner in double.  Relative error ~2^-64.
 *   exp(z): z = k ln2 + r (Cody-Waite split of ln2 into 32 + 53 bits),
 *           |r| <= ln2/2, e^r - 1 by a degree-14 Taylor polynomial, scaled by
 *           2^k.  The low word of a double-double z is folded into r.
 *   pow(x,y) = exp(y * log x) with y*log x formed in double-double.
 *   atan(x):  table of atan(k/8) plus the odd series of a reduced argument
 *           (see cm_atan).
 * Results are within about 1 ulp of the true value (checked against the host
 * libm in tests/test_math.py); they are not claimed correctly rounded.
 * Constants are ln 2, 1/ln 2, 1/ln 10, 2/3, pi/2, atan(k/8), 1/n! and
 * 1/(2k+1) to binary64 (computed at 60 digits with Python's decimal and
 * mpmath modules).
 */
#ifndef CTSM_MATH_H
#define CTSM_MATH_H

#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define CM_FN __host__ __device__ static inline
#else
#define CM_FN static inline
#endif

/* ln 2 = CM_LN2_H + CM_LN2_L (double-double) */
#define CM_LN2_H 0x1.62e42fefa39efp-1
#define CM_LN2_L 0x1.abc9e3b39803fp-56
/* ln 2 = CM_LN2_HI32 (32 significant bits, so k*CM_LN2_HI32 is exact) + CM_LN2_LO32 */
#define CM_LN2_HI32 0x1.62e42fee00000p-1
#define CM_LN2_LO32 0x1.a39ef35793c76p-33
#define CM_INV_LN2 0x1.71547652b82fep+0
#define CM_INV_LN10_H 0x1.bcb7b1526e50ep-2
#define CM_INV_LN10_L 0x1.95355baaafad3p-57
#define CM_TWO_THIRDS_H 0x1.5555555555555p-1
#define CM_TWO_THIRDS_L 0x1.5555555555555p-55
#define CM_SQRT_HALF 0x1.6a09e667f3bcdp-1

/* error-free transformations */
CM_FN void cm_two_sum(double a, double b, double *s, double *e)
{
    double ss = a + b;
    double bb = ss - a;
    *e = (a - (ss - bb)) + (b - bb);
    *s = ss;
}
CM_FN void cm_fast_two_sum(double a, double b, double *s, double *e)   /* |a| >= |b| */
{
    double ss = a + b;
    *e = b - (ss - a);
    *s = ss;
}
/* (ah,al) + (bh,bl) */
CM_FN void cm_dd_add(double ah, double al, double bh, double bl, double *h, double *l)
{
    double s, e;
    cm_two_sum(ah, bh, &s, &e);
    e = e + (al + bl);
    cm_fast_two_sum(s, e, h, l);
}
/* (ah,al) * (bh,bl) */
CM_FN void cm_dd_mul(double ah, double al, double bh, double bl, double *h, double *l)
{
    double p = ah * bh;
    double e = fma(ah, bh, -p);
    e = e + (ah * bl + al * bh);
    cm_fast_two_sum(p, e, h, l);
}

/* natural log of x > 0 (finite, normal or subnormal) as a double-double */
CM_FN void cm_log_dd(double x, double *lh, double *ll)
{
    int e;
    double m = frexp(x, &e);                 /* m in [0.5, 1) */
    if (m < CM_SQRT_HALF) { m = m * 2.0; e = e - 1; }   /* m in [sqrt(1/2), sqrt(2)) */
    const double num = m - 1.0;              /* exact (Sterbenz) */
    double dh, dl;
    cm_two_sum(m, 1.0, &dh, &dl);            /* m + 1 exactly */
    /* f = num / (dh + dl) in double-double */
    const double q1 = num / dh;
    double r = fma(-q1, dh, num);            /* exact remainder */
    r = r - q1 * dl;
    const double q2 = r / dh;
    double fh, fl;
    cm_fast_two_sum(q1, q2, &fh, &fl);
    /* s = f^2, t = f^3 in double-double */
    double sh, sl, th, tl;
    cm_dd_mul(fh, fl, fh, fl, &sh, &sl);
    cm_dd_mul(fh, fl, sh, sl, &th, &tl);
    /* u = (2/3) f^3 */
    double uh, ul;
    cm_dd_mul(CM_TWO_THIRDS_H, CM_TWO_THIRDS_L, th, tl, &uh, &ul);
    /* tail: f^5 * sum_{k>=0} 2 s^k / (2k+5), |s| <= 0.0295 */
    const double s = sh;
    double p = 0x1.0842108421084p-4;         /* 2/31 */
    p = 0x1.1a7b9611a7b96p-4 + s * p;        /* 2/29 */
    p = 0x1.2f684bda12f68p-4 + s * p;        /* 2/27 */
    p = 0x1.47ae147ae147bp-4 + s * p;        /* 2/25 */
    p = 0x1.642c8590b2164p-4 + s * p;        /* 2/23 */
    p = 0x1.8618618618618p-4 + s * p;        /* 2/21 */
    p = 0x1.af286bca1af28p-4 + s * p;        /* 2/19 */
    p = 0x1.e1e1e1e1e1e1ep-4 + s * p;        /* 2/17 */
    p = 0x1.1111111111111p-3 + s * p;        /* 2/15 */
    p = 0x1.3b13b13b13b14p-3 + s * p;        /* 2/13 */
    p = 0x1.745d1745d1746p-3 + s * p;        /* 2/11 */
    p = 0x1.c71c71c71c71cp-3 + s * p;        /* 2/9  */
    p = 0x1.2492492492492p-2 + s * p;        /* 2/7  */
    p = 0x1.999999999999ap-2 + s * p;        /* 2/5  */
    const double v = (th * sh) * p;          /* f^5 * tail */
    /* log m = 2f + u + v */
    double ah, al;
    cm_dd_add(2.0 * fh, 2.0 * fl, uh, ul, &ah, &al);
    cm_dd_add(ah, al, v, 0.0, &ah, &al);
    /* + e ln 2 */
    double kh, kl;
    cm_dd_mul((double)e, 0.0, CM_LN2_H, CM_LN2_L, &kh, &kl);
    cm_dd_add(kh, kl, ah, al, lh, ll);
}

/* exp of the double-double zh + zl */
CM_FN double cm_exp_dd(double zh, double zl)
{
    if (zh != zh) return zh;
    if (zh > 709.8) return HUGE_VAL;
    if (zh < -745.2) return 0.0;
    const double k = rint(zh * CM_INV_LN2);
    double r = (zh - k * CM_LN2_HI32) - k * CM_LN2_LO32;
    r = r + zl;
    /* e^r - 1, Taylor to r^14 (|r| <= 0.35) */
    double p = 0x1.93974a8c07c9dp-37;        /* 1/14! */
    p = 0x1.6124613a86d09p-33 + r * p;       /* 1/13! */
    p = 0x1.1eed8eff8d898p-29 + r * p;       /* 1/12! */
    p = 0x1.ae64567f544e4p-26 + r * p;       /* 1/11! */
    p = 0x1.27e4fb7789f5cp-22 + r * p;       /* 1/10! */
    p = 0x1.71de3a556c734p-19 + r * p;       /* 1/9!  */
    p = 0x1.a01a01a01a01ap-16 + r * p;       /* 1/8!  */
    p = 0x1.a01a01a01a01ap-13 + r * p;       /* 1/7!  */
    p = 0x1.6c16c16c16c17p-10 + r * p;       /* 1/6!  */
    p = 0x1.1111111111111p-7 + r * p;        /* 1/5!  */
    p = 0x1.5555555555555p-5 + r * p;        /* 1/4!  */
    p = 0x1.5555555555555p-3 + r * p;        /* 1/3!  */
    p = 0.5 + r * p;                         /* 1/2!  */
    const double em1 = r + (r * r) * p;
    const int ki = (int)k;
    /* 2^k * (1 + em1); split the scaling so large |k| cannot overflow early */
    if (ki > 1000 || ki < -1000) {
        const int k1 = ki / 2;
        return ldexp(ldexp(1.0 + em1, k1), ki - k1);
    }
    return ldexp(1.0 + em1, ki);
}

CM_FN double cm_exp(double x) { return cm_exp_dd(x, 0.0); }

CM_FN double cm_log(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL) return x;
    double h, l;
    cm_log_dd(x, &h, &l);
    return h + l;
}

CM_FN double cm_log10(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL) return x;
    double h, l, rh, rl;
    cm_log_dd(x, &h, &l);
    cm_dd_mul(h, l, CM_INV_LN10_H, CM_INV_LN10_L, &rh, &rl);
    return rh + rl;
}

/* arctangent.  |x| > 1 is folded to pi/2 - atan(1/|x|) with 1/|x| kept as a
 * double-double; the reduced argument a is split as a = c + (a - c) with
 * c = k/8 nearest, atan a = atan c + atan t, t = (a - c) / (1 + a c) formed in
 * double-double, atan c from a table (60-digit values rounded to
 * double-double) and atan t - t by its odd Taylor series (|t| <= 1/16). */
CM_FN double cm_atan(double x)
{
    if (x != x || x == 0.0) return x;
    const double a = fabs(x);
    if (a == HUGE_VAL) return x < 0.0 ? -0x1.921fb54442d18p+0 : 0x1.921fb54442d18p+0;
    double ih, il;
    int inv = 0;
    if (a > 1.0) {
        ih = 1.0 / a;
        il = fma(-ih, a, 1.0) / a;
        inv = 1;
    } else {
        ih = a;
        il = 0.0;
    }
    static const double AT_H[9] = {
        0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2,
        0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};
    static const double AT_L[9] = {
        0.0, -0x1.cd37686760c17p-59, 0x1.8ab6e3cf7afbdp-57, -0x1.c63aae6f6e918p-56, 0x1.a2b7f222f65e2p-56,
        -0x1.928df287a668fp-58, 0x1.2419a87f2a458p-56, -0x1.8c34d25aadef6p-56, 0x1.1a62633145c07p-55};
    const int k = (int)rint(ih * 8.0);       /* 0..8 */
    const double c = (double)k * 0.125;
    /* numerator (ih - c) + il, exact difference (Sterbenz) */
    double nh, nl;
    cm_two_sum(ih - c, il, &nh, &nl);
    /* denominator 1 + (ih + il) c */
    const double p = ih * c;
    const double pe = fma(ih, c, -p) + il * c;
    double dh, dl;
    cm_two_sum(1.0, p, &dh, &dl);
    dl = dl + pe;
    const double q1 = nh / dh;
    double r = fma(-q1, dh, nh);
    r = (r + nl) - q1 * dl;
    const double q2 = r / dh;
    const double s = q1 * q1;
    double poly = 0x1.e1e1e1e1e1e1ep-5;                /*  1/17 */
    poly = -0x1.1111111111111p-4 + s * poly;           /* -1/15 */
    poly = 0x1.3b13b13b13b14p-4 + s * poly;            /*  1/13 */
    poly = -0x1.745d1745d1746p-4 + s * poly;           /* -1/11 */
    poly = 0x1.c71c71c71c71cp-4 + s * poly;            /*  1/9  */
    poly = -0x1.2492492492492p-3 + s * poly;           /* -1/7  */
    poly = 0x1.999999999999ap-3 + s * poly;            /*  1/5  */
    poly = -0x1.5555555555555p-2 + s * poly;           /* -1/3  */
    const double tail = (q1 * s) * poly;
    double s1, e1, rh, rl;
    cm_two_sum(AT_H[k], q1, &s1, &e1);
    cm_fast_two_sum(s1, e1 + ((AT_L[k] + q2) + tail), &rh, &rl);
    double res;
    if (inv) {
        double h, l;
        cm_dd_add(0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54, -rh, -rl, &h, &l);   /* pi/2 - atan(1/a) */
        res = h + l;
    } else {
        res = rh + rl;
    }
    return x < 0.0 ? -res : res;
}

/* x**y for x > 0 (the only case the physics uses); x == 0 and x < 0 follow C. */
CM_FN double cm_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (!(x > 0.0) || x == HUGE_VAL || y != y) return pow(x, y);
    if (x == 1.0) return 1.0;
    double lh, ll;
    cm_log_dd(x, &lh, &ll);
    const double zh = y * lh;
    const double zl = fma(y, lh, -zh) + y * ll;
    return cm_exp_dd(zh, zl);
}

#endif /* CTSM_MATH_H */
