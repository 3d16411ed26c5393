/* fcfd_math.h -- exp, log, log10 and pow in FP64, written once and compiled unchanged for the GPU
 * (HIP, gfx950) and for the host (C11 or C++).
 *
 * Why: the rate constants (Arrhenius factors, equilibrium constants, fall-off blending), the TRANLIB
 * property fits and the integrator's step-size controller all go through these four functions.  The
 * GPU's math library and the host's differ in the last bit for some arguments, so two evaluations of
 * the same formula drift apart by ulps that a stiff integration amplifies.  Evaluating the SAME
 * sequence of IEEE operations on both sides (+, -, *, /, fma, all correctly rounded on both; no
 * contraction: every translation unit that includes this file is compiled with -ffp-contract=off)
 * gives bitwise equal results.
 *
 * Methods (textbook argument reduction + series, no tables, so the device code needs no memory):
 *   exp(x): Cody & Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split in two so that
 *           x - k*LN2_HI is exact (one fused multiply-add); exp(r) by its Taylor series to r^14
 *           (truncation below 2^-62), 1 + r added as an exact two-term sum; scaling by 2^k in two
 *           exact power-of-two factors.  Error below 0.55 ulp.
 *   log(x): x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f) carried as
 *           a two-term value (the division's remainder is exact under a fused multiply-add);
 *           log(1 + f) = 2 atanh(s) = 2s + s^3 (2/3 + 2/5 s^2 + ... + 2/25 s^22) (truncation below
 *           2^-62); e ln2 added with an exact product (fma) and an exact two-term sum (Knuth's
 *           TwoSum).  The two-term result is accurate to about 2^-60.
 *   log10(x) = log(x) * (1/ln 10) with both factors as two-term values.
 *   pow(x, y) = exp(y log x), the product y*log x carried as a two-term value into exp's
 *           reduction; exp10(y) the same with a stored two-term ln 10.
 * Special values follow C99 Annex F for the cases the kinetics and transport code meets (NaN in,
 * NaN out; overflow to +inf; underflow to 0; log(0) = -inf; log of a negative number NaN; pow with a
 * zero base or a zero exponent; pow of a negative base with an integer exponent).
 */
#ifndef FCFD_MATH_H
#define FCFD_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FM_FN __host__ __device__ static inline
#else
#define FM_FN static inline
#endif

#define FM_LN2_HI 0x1.62e42fefa39efp-1    /* ln 2 rounded to double */
#define FM_LN2_LO 0x1.abc9e3b39803fp-56   /* ln 2 - FM_LN2_HI */
#define FM_INV_LN2 0x1.71547652b82fep+0   /* 1 / ln 2 */
#define FM_INV_LN10_HI 0x1.bcb7b1526e50ep-2   /* 1 / ln 10 rounded to double */
#define FM_INV_LN10_LO 0x1.95355baaafad3p-57  /* 1 / ln 10 - FM_INV_LN10_HI */
#define FM_SHIFT 0x1.8p52                 /* adding and subtracting it rounds to an integer */

FM_FN double fm_from_bits(uint64_t b) {
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}

FM_FN uint64_t fm_to_bits(double d) {
    uint64_t b;
    __builtin_memcpy(&b, &d, 8);
    return b;
}

/* 2^m for -1022 <= m <= 1023 */
FM_FN double fm_pow2i(int m) { return fm_from_bits((uint64_t)(m + 1023) << 52); }

/* 1/2! + r/3! + ... + r^12/14!, the Taylor series of (exp(r) - 1 - r) / r^2 */
FM_FN double fm_exp_series(double rh) {
    double p = 0x1.93974a8c07c9dp-37;            /* 1/14! */
    p = __builtin_fma(p, rh, 0x1.6124613a86d09p-33);  /* 1/13! */
    p = __builtin_fma(p, rh, 0x1.1eed8eff8d898p-29);  /* 1/12! */
    p = __builtin_fma(p, rh, 0x1.ae64567f544e4p-26);  /* 1/11! */
    p = __builtin_fma(p, rh, 0x1.27e4fb7789f5cp-22);  /* 1/10! */
    p = __builtin_fma(p, rh, 0x1.71de3a556c734p-19);  /* 1/9! */
    p = __builtin_fma(p, rh, 0x1.a01a01a01a01ap-16);  /* 1/8! */
    p = __builtin_fma(p, rh, 0x1.a01a01a01a01ap-13);  /* 1/7! */
    p = __builtin_fma(p, rh, 0x1.6c16c16c16c17p-10);  /* 1/6! */
    p = __builtin_fma(p, rh, 0x1.1111111111111p-7);   /* 1/5! */
    p = __builtin_fma(p, rh, 0x1.5555555555555p-5);   /* 1/4! */
    p = __builtin_fma(p, rh, 0x1.5555555555555p-3);   /* 1/3! */
    p = __builtin_fma(p, rh, 0.5);                    /* 1/2! */
    return p;
}

/* exp(xh + xl), |xl| small against ulp(xh) * 2^10.  Branch-free: the reduction and series run on
 * the argument clamped to [-745.2, 709.79] (unchanged inside that range) and the results for
 * arguments outside it (0, +inf, NaN) are selected at the end, so that on the GPU the lanes of a
 * wave never split over the special cases. */
FM_FN double fm_exp_dd(double xh, double xl) {
    const double xc = __builtin_fmin(__builtin_fmax(xh, -745.2), 709.79);
    const double kd = (xc * FM_INV_LN2 + FM_SHIFT) - FM_SHIFT;
    /* |xc| >= 1/2 or kd = 0: xc - kd*LN2_HI is exactly representable, so the fma is exact */
    const double r0 = __builtin_fma(-kd, FM_LN2_HI, xc);
    /* r = rh + rl = r0 + xl - kd*LN2_LO with rh the rounded sum and rl its rounding error */
    const double r1 = xl - kd * FM_LN2_LO;
    const double rh = r0 + r1;
    const double rl = r1 - (rh - r0);
    /* q = rh^2 (1/2! + rh/3! + ... + rh^12/14!) */
    const double p = fm_exp_series(rh);
    const double q = (rh * rh) * p;
    /* exp(r) = (1 + rh + q)(1 + rl) to first order in rl (|rl| <= ulp(rh) / 2); 1 + rh as an exact
     * two-term sum (|rh| < 1) */
    const double s = 1.0 + rh;
    const double e = rh - (s - 1.0);
    const double v = s + (e + (rl + (rh * rl + q)));
    const int k = (int)kd;                /* -1075 <= k <= 1024: both factors below are normal */
    const int k1 = k / 2;
    double res = (v * fm_pow2i(k1)) * fm_pow2i(k - k1);
    res = xh > 709.79 ? __builtin_inf() : res;
    res = xh < -745.2 ? 0.0 : res;
    return xh != xh ? xh + xh : res;
}

FM_FN double fm_exp(double x) { return fm_exp_dd(x, 0.0); }

/* exp(x) - 1 without the cancellation of forming exp(x) first: the same reduction x = k ln2 + r,
 * e^r - 1 = r + r^2 P(r) kept apart from the 1, then 2^k (e^r - 1) + (2^k - 1), the second term
 * exact for |k| <= 53.  Error below 2 ulp. */
FM_FN double fm_expm1(double x) {
    if (x != x) return x + x;
    if (x > 36.0) return fm_exp(x) - 1.0;      /* exp(x) > 2^51: the -1 is a rounding of it */
    if (x < -38.0) return -1.0;                /* exp(x) < 2^-54: -1 + exp(x) rounds to -1 */
    if (x > -0x1.62e42fefa39efp-2 && x < 0x1.62e42fefa39efp-2)   /* |x| < ln2 / 2: k = 0 */
        return x + (x * x) * fm_exp_series(x);
    const double kd = (x * FM_INV_LN2 + FM_SHIFT) - FM_SHIFT;
    const double r0 = __builtin_fma(-kd, FM_LN2_HI, x);
    const double r1 = -kd * FM_LN2_LO;
    const double rh = r0 + r1;
    const double rl = r1 - (rh - r0);
    const double t = rh + (rl + (rh * rl + (rh * rh) * fm_exp_series(rh)));   /* e^r - 1 */
    const double sk = fm_pow2i((int)kd);
    return t * sk + (sk - 1.0);
}

/* tanh(x) = -expm1(-2|x|) / (2 + expm1(-2|x|)) with the sign of x; 1 for |x| > 22 (the quotient
 * is then 1 - 2e^-44, which rounds to 1).  Error below 3 ulp. */
FM_FN double fm_tanh(double x) {
    if (x != x) return x + x;
    const double a = x < 0.0 ? -x : x;
    double r = 1.0;
    if (a <= 22.0) {
        const double em = fm_expm1(-2.0 * a);
        r = -em / (2.0 + em);
    }
    return x < 0.0 ? -r : r;
}

/* log(x) as a two-term value (*lo carries the low part); x finite and > 0 */
FM_FN double fm_log_dd(double x, double *lo) {
    int e = 0;
    uint64_t b = fm_to_bits(x);
    if ((b >> 52) == 0) {                  /* subnormal: scale into the normal range */
        x = x * 0x1p54;
        b = fm_to_bits(x);
        e = -54;
    }
    e += (int)(b >> 52) - 1023;
    b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;   /* m in [1, 2) */
    if (b > 0x3ff6a09e667f3bcdULL) {       /* m > sqrt(2): use m/2 */
        b -= 0x0010000000000000ULL;
        e += 1;
    }
    const double f = fm_from_bits(b) - 1.0;    /* exact */
    const double u = 2.0 + f;
    const double ul = f - (u - 2.0);            /* 2 + f = u + ul exactly */
    const double sh = f / u;
    const double sl = (__builtin_fma(-sh, u, f) - sh * ul) / u;
    const double z = sh * sh;
    double p = 2.0 / 25.0;
    p = __builtin_fma(p, z, 2.0 / 23.0);
    p = __builtin_fma(p, z, 2.0 / 21.0);
    p = __builtin_fma(p, z, 2.0 / 19.0);
    p = __builtin_fma(p, z, 2.0 / 17.0);
    p = __builtin_fma(p, z, 2.0 / 15.0);
    p = __builtin_fma(p, z, 2.0 / 13.0);
    p = __builtin_fma(p, z, 2.0 / 11.0);
    p = __builtin_fma(p, z, 2.0 / 9.0);
    p = __builtin_fma(p, z, 2.0 / 7.0);
    p = __builtin_fma(p, z, 2.0 / 5.0);
    p = __builtin_fma(p, z, 2.0 / 3.0);
    /* small terms: 2 sl, the series tail s^3 P(s^2), and the tail's first-order change in sl */
    const double small = 2.0 * sl + ((sh * z) * p + 2.0 * z * sl);
    const double ed = (double)e;
    const double ph = ed * FM_LN2_HI;
    const double pl = __builtin_fma(ed, FM_LN2_HI, -ph);   /* e*LN2_HI = ph + pl exactly */
    const double b2 = 2.0 * sh;
    /* TwoSum(ph, b2) */
    const double t = ph + b2;
    const double bv = t - ph;
    const double err = (ph - (t - bv)) + (b2 - bv);
    const double tl = err + (pl + (ed * FM_LN2_LO + small));
    const double hi = t + tl;
    *lo = tl - (hi - t);
    return hi;
}

FM_FN double fm_log(double x) {
    if (x != x) return x + x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    double lo;
    return fm_log_dd(x, &lo);
}

FM_FN double fm_log10(double x) {
    if (x != x) return x + x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    double lo;
    const double hi = fm_log_dd(x, &lo);
    const double ph = hi * FM_INV_LN10_HI;
    const double pl = __builtin_fma(hi, FM_INV_LN10_HI, -ph) + (hi * FM_INV_LN10_LO + lo * FM_INV_LN10_HI);
    return ph + pl;
}

#define FM_LN10_HI 0x1.26bb1bbb55516p+1   /* ln 10 rounded to double */
#define FM_LN10_LO -0x1.f48ad494ea3e9p-53 /* ln 10 - FM_LN10_HI */

/* 10^y = exp(y ln 10) with y ln 10 carried as a two-term value (the fall-off blending of CKRATX
 * raises 10 to a power; a stored ln 10 saves the logarithm pow would take of its base) */
FM_FN double fm_exp10(double y) {
    if (y != y) return y + y;
    const double ph = y * FM_LN10_HI;
    const double pl = __builtin_fma(y, FM_LN10_HI, -ph) + y * FM_LN10_LO;
    return fm_exp_dd(ph, pl);
}

FM_FN double fm_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    double sign = 1.0;
    if (x < 0.0) {
        const double yi = (y + FM_SHIFT) - FM_SHIFT;
        if (!(yi == y) || !(y < 0x1p52 && y > -0x1p52)) return __builtin_nan("");
        const double h = y * 0.5;
        if (!((h + FM_SHIFT) - FM_SHIFT == h)) sign = -1.0;                          /* odd y */
        x = -x;
    }
    if (x == 0.0) return y > 0.0 ? 0.0 * sign : sign * __builtin_inf();
    if (x == __builtin_inf()) return y > 0.0 ? sign * x : 0.0 * sign;
    if (x == 1.0) return sign;
    double lo;
    const double lh = fm_log_dd(x, &lo);
    const double ph = y * lh;
    const double pl = __builtin_fma(y, lh, -ph) + y * lo;
    return sign * fm_exp_dd(ph, pl);
}

#endif /* FCFD_MATH_H */
