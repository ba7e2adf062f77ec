/*
 * ctsm_math.h -- deterministic binary64 log / exp / pow / log10 / atan / sin.
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
 *   log(x): x = 2^e m, m in [sqrt(1/2), sqrt(2)); m = F (1 + u) with F the
 *           nearest k/128 (table of log F in double-double), u in
 *           double-double from an fma remainder (no division), log(1+u) by
 *           its series to u^9.  Relative error ~2^-64.
 *   exp(z): z = m ln2 + r (Cody-Waite split of ln2 into 32 + 53 bits),
 *           |r| <= ln2/2, e^r - 1 by its Taylor polynomial to r^13 (no
 *           table), 1 + r carried exactly, scaled by 2^m.  The low word of a
 *           double-double z is folded into r.
 *   pow(x,y) = exp(y * log x) with y*log x formed in double-double.
 *   atan(x):  table of atan(k/8) plus the odd series of a reduced argument
 *           (see cm_atan).
 *   sin(x):   three-part reduction by pi/2, Taylor polynomials (see cm_sin).
 * Results are within about 1 ulp of the true value (checked against the host
 * libm in tests/test_math.py); they are not claimed correctly rounded.
 * Constants and tables (ln 2, 1/ln 10, log(k/128), pi/2, atan(k/8), 1/n,
 * 1/n!) are the binary64 / double-double roundings of
 * 60-digit values computed with Python's mpmath.
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
#define CM_INV_LN10_H 0x1.bcb7b1526e50ep-2
#define CM_INV_LN10_L 0x1.95355baaafad3p-57
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

/* natural log of x > 0 (finite, normal or subnormal) as a double-double.
 * x = 2^e m, m in [sqrt(1/2), sqrt(2)); F = k/128 nearest m (table of log F
 * in double-double and 1/F), u = (m - F)/F in double-double without a
 * division (f = m - F is exact, the remainder of f * (1/F) comes from an fma),
 * log(1 + u) = u - u^2/2 + ... + u^9/9 (|u| <= 2^-7.5, truncation < 2^-66 u). */
#define CM_LOGT_K0 91
CM_FN void cm_log_dd(double x, double *lh, double *ll)
{
    static const double CM_LOGT[91][3] = {
        {-0x1.5d5bddf595f30p-2, 0x1.6541148cbb8a2p-56, 0x1.6816816816817p+0},
        {-0x1.522ae0738a3d8p-2, 0x1.8f7e9b38a6979p-57, 0x1.642c8590b2164p+0},
        {-0x1.4718dc271c41bp-2, -0x1.8fb4c14c56eefp-60, 0x1.6058160581606p+0},
        {-0x1.3c25277333184p-2, 0x1.2ad27e50a8ec6p-56, 0x1.5c9882b931057p+0},
        {-0x1.314f1e1d35ce4p-2, 0x1.3d69909e5c3dcp-56, 0x1.58ed2308158edp+0},
        {-0x1.269621134db92p-2, -0x1.e0efadd9db02bp-56, 0x1.5555555555555p+0},
        {-0x1.1bf99635a6b95p-2, 0x1.12aeb84249223p-57, 0x1.51d07eae2f815p+0},
        {-0x1.1178e8227e47cp-2, 0x1.0e63a5f01c691p-57, 0x1.4e5e0a72f0539p+0},
        {-0x1.07138604d5862p-2, -0x1.cdb16ed4e9138p-56, 0x1.4afd6a052bf5bp+0},
        {-0x1.f991c6cb3b379p-3, -0x1.f665066f980a2p-57, 0x1.47ae147ae147bp+0},
        {-0x1.e530effe71012p-3, -0x1.2276041f43042p-59, 0x1.446f86562d9fbp+0},
        {-0x1.d1037f2655e7bp-3, -0x1.60629242471a2p-57, 0x1.4141414141414p+0},
        {-0x1.bd087383bd8adp-3, -0x1.dd355f6a516d7p-60, 0x1.3e22cbce4a902p+0},
        {-0x1.a93ed3c8ad9e3p-3, -0x1.bcafa9de97203p-57, 0x1.3b13b13b13b14p+0},
        {-0x1.95a5adcf7017fp-3, -0x1.142c507fb7a3dp-58, 0x1.3813813813814p+0},
        {-0x1.823c16551a3c2p-3, 0x1.1232ce70be781p-57, 0x1.3521cfb2b78c1p+0},
        {-0x1.6f0128b756abcp-3, 0x1.8de59c21e166cp-57, 0x1.323e34a2b10bfp+0},
        {-0x1.5bf406b543db2p-3, 0x1.1f5b44c0df7e7p-61, 0x1.2f684bda12f68p+0},
        {-0x1.4913d8333b561p-3, 0x1.0d5604930f135p-58, 0x1.2c9fb4d812ca0p+0},
        {-0x1.365fcb0159016p-3, -0x1.7d411a5b944adp-58, 0x1.29e4129e4129ep+0},
        {-0x1.23d712a49c202p-3, 0x1.6e38161051d69p-57, 0x1.27350b8812735p+0},
        {-0x1.1178e8227e47cp-3, 0x1.0e63a5f01c691p-58, 0x1.2492492492492p+0},
        {-0x1.fe89139dbd566p-4, 0x1.ac9f4215f9393p-58, 0x1.21fb78121fb78p+0},
        {-0x1.da727638446a2p-4, -0x1.401fa71733019p-58, 0x1.1f7047dc11f70p+0},
        {-0x1.b6ac88dad5b1cp-4, 0x1.0057eed1ca59fp-59, 0x1.1cf06ada2811dp+0},
        {-0x1.9335e5d594989p-4, 0x1.478a85704ccb7p-58, 0x1.1a7b9611a7b96p+0},
        {-0x1.700d30aeac0e1p-4, 0x1.72566212cdd05p-61, 0x1.1811811811812p+0},
        {-0x1.4d3115d207eacp-4, -0x1.769f42c7842ccp-58, 0x1.15b1e5f75270dp+0},
        {-0x1.2aa04a44717a5p-4, 0x1.d15d38d2fa3f7p-58, 0x1.135c81135c811p+0},
        {-0x1.08598b59e3a07p-4, 0x1.dd7009902bf32p-58, 0x1.1111111111111p+0},
        {-0x1.ccb73cdddb2ccp-5, 0x1.e48fb0500efd4p-59, 0x1.0ecf56be69c90p+0},
        {-0x1.894aa149fb343p-5, -0x1.a8be97660a23dp-60, 0x1.0c9714fbcda3bp+0},
        {-0x1.466aed42de3eap-5, 0x1.cdd6f7f4a137ep-59, 0x1.0a6810a6810a7p+0},
        {-0x1.0415d89e74444p-5, -0x1.c05cf1d753622p-59, 0x1.0842108421084p+0},
        {-0x1.8492528c8cabfp-6, 0x1.d192d0619fa67p-60, 0x1.0624dd2f1a9fcp+0},
        {-0x1.0205658935847p-6, -0x1.27c8e8416e71fp-60, 0x1.0410410410410p+0},
        {-0x1.010157588de71p-7, -0x1.46662d417ced0p-62, 0x1.0204081020408p+0},
        {0.0, 0.0, 0x1.0000000000000p+0},
        {0x1.fe02a6b106789p-8, -0x1.e44b7e3711ebfp-67, 0x1.fc07f01fc07f0p-1},
        {0x1.fc0a8b0fc03e4p-7, -0x1.83092c59642a1p-62, 0x1.f81f81f81f820p-1},
        {0x1.7b91b07d5b11bp-6, -0x1.5b602ace3a510p-60, 0x1.f44659e4a4271p-1},
        {0x1.f829b0e783300p-6, 0x1.33e3f04f1ef23p-60, 0x1.f07c1f07c1f08p-1},
        {0x1.39e87b9febd60p-5, -0x1.5bfa937f551bbp-59, 0x1.ecc07b301ecc0p-1},
        {0x1.77458f632dcfcp-5, 0x1.18d3ca87b9296p-59, 0x1.e9131abf0b767p-1},
        {0x1.b42dd711971bfp-5, -0x1.eb9759c130499p-60, 0x1.e573ac901e574p-1},
        {0x1.f0a30c01162a6p-5, 0x1.85f325c5bbacdp-59, 0x1.e1e1e1e1e1e1ep-1},
        {0x1.16536eea37ae1p-4, -0x1.79da3e8c22cdap-60, 0x1.de5d6e3f8868ap-1},
        {0x1.341d7961bd1d1p-4, -0x1.b599f227becbbp-58, 0x1.dae6076b981dbp-1},
        {0x1.51b073f06183fp-4, 0x1.a49e39a1a8be4p-58, 0x1.d77b654b82c34p-1},
        {0x1.6f0d28ae56b4cp-4, -0x1.906d99184b992p-58, 0x1.d41d41d41d41dp-1},
        {0x1.8c345d6319b21p-4, -0x1.4a697ab3424a9p-61, 0x1.d0cb58f6ec074p-1},
        {0x1.a926d3a4ad563p-4, 0x1.942f48aa70ea9p-58, 0x1.cd85689039b0bp-1},
        {0x1.c5e548f5bc743p-4, 0x1.5d617ef8161b1p-60, 0x1.ca4b3055ee191p-1},
        {0x1.e27076e2af2e6p-4, -0x1.61578001e0162p-60, 0x1.c71c71c71c71cp-1},
        {0x1.fec9131dbeabbp-4, -0x1.5746b9981b36cp-58, 0x1.c3f8f01c3f8f0p-1},
        {0x1.0d77e7cd08e59p-3, 0x1.9a5dc5e9030acp-57, 0x1.c0e070381c0e0p-1},
        {0x1.1b72ad52f67a0p-3, 0x1.483023472cd74p-58, 0x1.bdd2b899406f7p-1},
        {0x1.29552f81ff523p-3, 0x1.301771c407dbfp-57, 0x1.bacf914c1bad0p-1},
        {0x1.371fc201e8f74p-3, 0x1.de6cb62af18a0p-58, 0x1.b7d6c3dda338bp-1},
        {0x1.44d2b6ccb7d1ep-3, 0x1.9f4f6543e1f88p-57, 0x1.b4e81b4e81b4fp-1},
        {0x1.526e5e3a1b438p-3, -0x1.746ff8a470d3ap-57, 0x1.b2036406c80d9p-1},
        {0x1.5ff3070a793d4p-3, -0x1.bc60efafc6f6ep-58, 0x1.af286bca1af28p-1},
        {0x1.6d60fe719d21dp-3, -0x1.caae268ecd179p-57, 0x1.ac5701ac5701bp-1},
        {0x1.7ab890210d909p-3, 0x1.be36b2d6a0608p-59, 0x1.a98ef606a63bep-1},
        {0x1.87fa06520c911p-3, -0x1.bf7fdbfa08d9ap-57, 0x1.a6d01a6d01a6dp-1},
        {0x1.9525a9cf456b4p-3, 0x1.d904c1d4e2e26p-57, 0x1.a41a41a41a41ap-1},
        {0x1.a23bc1fe2b563p-3, 0x1.93711b07a998cp-59, 0x1.a16d3f97a4b02p-1},
        {0x1.af3c94e80bff3p-3, -0x1.398cff3641985p-58, 0x1.9ec8e951033d9p-1},
        {0x1.bc286742d8cd6p-3, 0x1.4fce744870f55p-58, 0x1.9c2d14ee4a102p-1},
        {0x1.c8ff7c79a9a22p-3, -0x1.4f689f8434012p-57, 0x1.999999999999ap-1},
        {0x1.d5c216b4fbb91p-3, 0x1.6e443597e4d40p-57, 0x1.970e4f80cb872p-1},
        {0x1.e27076e2af2e6p-3, -0x1.61578001e0162p-59, 0x1.948b0fcd6e9e0p-1},
        {0x1.ef0adcbdc5936p-3, 0x1.48637950dc20dp-57, 0x1.920fb49d0e229p-1},
        {0x1.fb9186d5e3e2bp-3, -0x1.caaae64f21acbp-57, 0x1.8f9c18f9c18fap-1},
        {0x1.0402594b4d041p-2, -0x1.28ec217a5022dp-57, 0x1.8d3018d3018d3p-1},
        {0x1.0a324e27390e3p-2, 0x1.7dcfde8061c03p-56, 0x1.8acb90f6bf3aap-1},
        {0x1.1058bf9ae4ad5p-2, 0x1.89fa0ab4cb31dp-58, 0x1.886e5f0abb04ap-1},
        {0x1.1675cababa60ep-2, 0x1.ce63eab883717p-61, 0x1.8618618618618p-1},
        {0x1.1c898c16999fbp-2, -0x1.0e5c62aff1c44p-60, 0x1.83c977ab2beddp-1},
        {0x1.22941fbcf7966p-2, -0x1.76f5eb09628afp-56, 0x1.8181818181818p-1},
        {0x1.2895a13de86a3p-2, 0x1.7ad24c13f040ep-56, 0x1.7f405fd017f40p-1},
        {0x1.2e8e2bae11d31p-2, -0x1.8f4cdb95ebdf9p-56, 0x1.7d05f417d05f4p-1},
        {0x1.347dd9a987d55p-2, -0x1.4dd4c580919f8p-57, 0x1.7ad2208e0ecc3p-1},
        {0x1.3a64c556945eap-2, -0x1.c68651945f97cp-57, 0x1.78a4c8178a4c8p-1},
        {0x1.404308686a7e4p-2, -0x1.0bcfb6082ce6dp-56, 0x1.767dce434a9b1p-1},
        {0x1.4618bc21c5ec2p-2, 0x1.f42decdeccf1dp-56, 0x1.745d1745d1746p-1},
        {0x1.4be5f957778a1p-2, -0x1.259b35b04813dp-57, 0x1.724287f46debcp-1},
        {0x1.51aad872df82dp-2, 0x1.3927ac19f55e3p-59, 0x1.702e05c0b8170p-1},
        {0x1.5767717455a6cp-2, 0x1.526adb283660cp-56, 0x1.6e1f76b4337c7p-1},
        {0x1.5d1bdbf5809cap-2, 0x1.4236383dc7fe1p-56, 0x1.6c16c16c16c17p-1},
        {0x1.62c82f2b9c795p-2, 0x1.7b7af915300e5p-57, 0x1.6a13cd1537290p-1},
    };
    int e;
    double m = frexp(x, &e);                 /* m in [0.5, 1) */
    if (m < CM_SQRT_HALF) { m = m * 2.0; e = e - 1; }   /* m in [sqrt(1/2), sqrt(2)) */
    const int k = (int)rint(m * 128.0);      /* 91..181 */
    const double F = (double)k * 0.0078125;  /* exact */
    const double f = m - F;                  /* exact (Sterbenz), |f| <= 2^-8 */
    const double *t = CM_LOGT[k - CM_LOGT_K0];
    const double q1 = f * t[2];
    const double r = fma(-q1, F, f);         /* f - q1 F exactly */
    const double q2 = r * t[2];
    const double u = q1;
    double p = 0x1.c71c71c71c71cp-4;         /*  1/9 */
    p = -0x1.0000000000000p-3 + u * p;       /* -1/8 */
    p = 0x1.2492492492492p-3 + u * p;        /*  1/7 */
    p = -0x1.5555555555555p-3 + u * p;       /* -1/6 */
    p = 0x1.999999999999ap-3 + u * p;        /*  1/5 */
    p = -0x1.0000000000000p-2 + u * p;       /* -1/4 */
    p = 0x1.5555555555555p-2 + u * p;        /*  1/3 */
    p = -0.5 + u * p;                        /* -1/2 */
    p = (u * u) * p;                         /* log(1+u) - u */
    const double ed = (double)e;
    double s1, e1, s2, e2;
    cm_two_sum(ed * CM_LN2_HI32, t[0], &s1, &e1);   /* e ln2 (high part exact) + log F */
    cm_two_sum(s1, q1, &s2, &e2);
    const double lo = ((e1 + e2) + (ed * CM_LN2_LO32 + t[1])) + ((q2 - q1 * q2) + p);
    cm_fast_two_sum(s2, lo, lh, ll);
}

/* 1 / ln 2 */
#define CM_INV_LN2 0x1.71547652b82fep+0

/* exp of the double-double zh + zl, without a table (a table lookup is a
 * memory access per call on the GPU, where the kernels call exp in their
 * iteration loops): z = m ln2 + r (Cody-Waite, ln 2 split into 32 + 53 bits so
 * m * CM_LN2_HI32 is exact), |r| <= ln2/2 (+ the low word zl folded in),
 * e^r = 1 + r + r^2 q(r) with q the Taylor polynomial 1/2 + r/6 + ... + r^11/13!
 * (truncation below 2^-60 relative), the rounding of 1 + r carried by an
 * error-free sum, scaled by 2^m. */
CM_FN double cm_exp_dd(double zh, double zl)
{
    if (zh != zh) return zh;
    if (zh > 709.8) return HUGE_VAL;
    if (zh < -745.2) return 0.0;
    const double kd = rint(zh * CM_INV_LN2);
    const int m = (int)kd;
    double r = (zh - kd * CM_LN2_HI32) - kd * CM_LN2_LO32;
    r = r + zl;
    double q = 0x1.6124613a86d09p-33;        /* 1/13! */
    q = 0x1.1eed8eff8d898p-29 + r * q;       /* 1/12! */
    q = 0x1.ae64567f544e4p-26 + r * q;       /* 1/11! */
    q = 0x1.27e4fb7789f5cp-22 + r * q;       /* 1/10! */
    q = 0x1.71de3a556c734p-19 + r * q;       /* 1/9!  */
    q = 0x1.a01a01a01a01ap-16 + r * q;       /* 1/8!  */
    q = 0x1.a01a01a01a01ap-13 + r * q;       /* 1/7!  */
    q = 0x1.6c16c16c16c17p-10 + r * q;       /* 1/6!  */
    q = 0x1.1111111111111p-7 + r * q;        /* 1/5!  */
    q = 0x1.5555555555555p-5 + r * q;        /* 1/4!  */
    q = 0x1.5555555555555p-3 + r * q;        /* 1/3!  */
    q = 0.5 + r * q;                         /* 1/2   */
    const double p = (r * r) * q;            /* e^r - 1 - r */
    double s, e;
    cm_fast_two_sum(1.0, r, &s, &e);         /* 1 + r = s + e exactly (|r| < 1) */
    const double res = s + (e + p);
    /* 2^m * res; split the scaling so large |m| cannot overflow early */
    if (m > 1000 || m < -1000) {
        const int m1 = m / 2;
        return ldexp(ldexp(res, m1), m - m1);
    }
    return ldexp(res, m);
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

/* x**y for x > 0 (the only case the physics uses); x == 0 and x < 0 follow C;
 * x**0 = 1 and x**1 = x exactly, as a C pow gives them. */
CM_FN double cm_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    if (!(x > 0.0) || x == HUGE_VAL || y != y) return pow(x, y);
    if (x == 1.0) return 1.0;
    double lh, ll;
    cm_log_dd(x, &lh, &ll);
    const double zh = y * lh;
    const double zl = fma(y, lh, -zh) + y * ll;
    return cm_exp_dd(zh, zl);
}

/* sin(x) for |x| < 2^20 (the physics takes it of slopes in radians, 0..pi/2).
 * x = k pi/2 + r with k the nearest integer of x 2/pi and pi/2 split in three
 * parts (33, 33 and 53 significant bits, so k P1 and k P2 are exact), r kept
 * as r + r_lo, |r| <= pi/4; then sin r = r + r^3 S(r^2) or
 * cos r = 1 - r^2/2 + r^4 C(r^2), the Taylor series to r^17 / r^18
 * (truncation below 1e-19 for |r| <= pi/4), with the tail r_lo and the
 * rounding of 1 - r^2/2 carried as in fdlibm's kernel_sin / kernel_cos
 * (Sun Microsystems, 1993). */
#define CM_PIO2_1 0x1.921fb54400000p+0
#define CM_PIO2_2 0x1.0b4611a600000p-34
#define CM_PIO2_3 0x1.3198a2e037073p-69
#define CM_2_PI 0x1.45f306dc9c883p-1
/* x = k pi/2 + r + rl; returns k mod 4 */
CM_FN int cm__reduce_pio2(double x, double *r, double *rl)
{
    const double kd = rint(x * CM_2_PI);
    const double t = (x - kd * CM_PIO2_1) - kd * CM_PIO2_2;
    const double p3 = kd * CM_PIO2_3;
    *r = t - p3;
    *rl = (t - *r) - p3;
    return ((int)kd) & 3;
}

/* sin(r + rl) and cos(r + rl) for |r| <= pi/4 */
CM_FN double cm__sin_r(double r, double rl)
{
    const double z = r * r;
    const double S = -0x1.5555555555555p-3 + z * (0x1.1111111111111p-7 + z * (-0x1.a01a01a01a01ap-13
                      + z * (0x1.71de3a556c734p-19 + z * (-0x1.ae64567f544e4p-26 + z * (0x1.6124613a86d09p-33
                      + z * (-0x1.ae7f3e733b81fp-41 + z * 0x1.952c77030ad4ap-49))))));
    return r + (r * (z * S) + rl);
}

CM_FN double cm__cos_r(double r, double rl)
{
    const double z = r * r;
    const double C4 = 0x1.5555555555555p-5 + z * (-0x1.6c16c16c16c17p-10 + z * (0x1.a01a01a01a01ap-16
                       + z * (-0x1.27e4fb7789f5cp-22 + z * (0x1.1eed8eff8d898p-29 + z * (-0x1.93974a8c07c9dp-37
                       + z * (0x1.ae7f3e733b81fp-45 + z * -0x1.6827863b97d97p-53))))));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * (z * C4) - r * rl));
}

CM_FN double cm_sin(double x)
{
    if (x != x || fabs(x) == HUGE_VAL) return x - x;
    if (x == 0.0) return x;
    if (!(fabs(x) < 0x1.0p+20)) return sin(x);
    double r, rl;
    const int q = cm__reduce_pio2(x, &r, &rl);
    const double v = (q & 1) == 0 ? cm__sin_r(r, rl) : cm__cos_r(r, rl);
    return (q & 2) ? -v : v;
}

/* tan(x) for |x| < 2^20 (the drainage takes it of slopes in radians): the
 * quotient of the reduced sine and cosine above, sin r / cos r or
 * -cos r / sin r by quadrant, the division's rounding corrected once (within
 * 2.5 ulp: the reduced sine and cosine each carry up to half an ulp). */
CM_FN double cm_tan(double x)
{
    if (x != x || fabs(x) == HUGE_VAL) return x - x;
    if (x == 0.0) return x;
    if (!(fabs(x) < 0x1.0p+20)) return tan(x);
    double r, rl;
    const int q = cm__reduce_pio2(x, &r, &rl);
    const double s = cm__sin_r(r, rl), c = cm__cos_r(r, rl);
    const double num = (q & 1) == 0 ? s : c, den = (q & 1) == 0 ? c : s;
    double t = num / den;
    t = t + fma(-t, den, num) / den;    /* the quotient's own rounding corrected */
    return (q & 1) == 0 ? t : -t;
}

#endif /* CTSM_MATH_H */
