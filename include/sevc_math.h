/*
 * sevc_math.h -- arithmetic shared bit-for-bit by the host (gcc) and the
 * device (hipcc, gfx950) sides of the vertex-model time step.
 *
 * Everything here is written with +, -, *, /, sqrt, floor, ldexp and
 * comparisons only, so that a translation unit compiled with
 * -ffp-contract=off yields identical bits on x86-64 and on CDNA4.
 *
 *  - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 *    "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Replaces the
 *    cuRAND XORWOW states of the reference (nucleation2.cu:12-16,243,298).
 *  - atan / atan2 / sin / cos / exp following the Cephes Math Library
 *    (S. L. Moshier, "Methods and Programs for Mathematical Functions",
 *    1989; atan.c, sin.c, exp.c): range reduction plus rational / polynomial
 *    minimax approximations.  They stand in for the libm calls of the
 *    reference (geometry.h:282-284 atan2/cos/sin, geometry.h:549 cos/pow,
 *    nucleation2.cu:422 exp).  Accuracy is ~1-2 ulp against glibc.
 */
#ifndef SEVC_MATH_H
#define SEVC_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define SEVC_HD __host__ __device__ inline
#else
#define SEVC_HD static inline
#endif

#define SEVC_PI    3.14159265358979323846
#define SEVC_PIO2  1.57079632679489661923
#define SEVC_PIO4  0.78539816339744830962

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                        */
/* ------------------------------------------------------------------ */
#define SEVC_PHILOX_M0 0xD2511F53u
#define SEVC_PHILOX_M1 0xCD9E8D57u
#define SEVC_PHILOX_W0 0x9E3779B9u
#define SEVC_PHILOX_W1 0xBB67AE85u

SEVC_HD void sevc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)SEVC_PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)SEVC_PHILOX_M1 * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += SEVC_PHILOX_W0; k1 += SEVC_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform double in (0, 1] from 64 random bits (53 significant bits).
 * The half-open-at-zero range matches curand_uniform_double's documented
 * range, which the reference's formulas assume (nucleation2.cu:262-263). */
SEVC_HD double sevc_u01(uint32_t hi, uint32_t lo) {
    uint64_t k = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;   /* 0 .. 2^53-1 */
    return (double)(k + 1) * (1.0 / 9007199254740992.0);        /* (k+1) 2^-53 */
}

/* Two uniforms for (key=seed, counter=(a, b, c, d)). */
SEVC_HD void sevc_uniform2(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                           double* u0, double* u1) {
    uint32_t ctr[4] = {a, b, c, d};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    sevc_philox4x32_10(ctr, key, o);
    *u0 = sevc_u01(o[0], o[1]);
    *u1 = sevc_u01(o[2], o[3]);
}

/* ------------------------------------------------------------------ */
/* Cephes-style elementary functions                                   */
/* ------------------------------------------------------------------ */

/* atan(x), Cephes atan.c: reduction to [0, 0.66] by tan(3pi/8) and
 * tan(pi/8) breakpoints, then x + x^3 P(x^2)/Q(x^2). */
SEVC_HD double sevc_atan(double x) {
    const double T3P8 = 2.41421356237309504880;   /* tan(3 pi / 8) */
    const double MOREBITS = 6.123233995736765886130e-17;
    double y, z;
    int sign = 0, flag = 0;
    if (x == 0.0) return x;
    if (x < 0.0) { sign = 1; x = -x; }
    /* the three reductions as one division: -(1/x) = (-1)/x and x/1 = x
     * exactly, so each branch keeps Cephes's value */
    double num = x, den = 1.0;
    if (x > T3P8) {
        y = SEVC_PIO2; flag = 1; num = -1.0; den = x;
    } else if (x <= 0.66) {
        y = 0.0;
    } else {
        y = SEVC_PIO4; flag = 2; num = x - 1.0; den = x + 1.0;
    }
    x = num / den;
    z = x * x;
    {
        double p = -8.750608600031904122785e-1;
        p = p * z + -1.615753718733365076637e1;
        p = p * z + -7.500855792314704667340e1;
        p = p * z + -1.228866684490136173410e2;
        p = p * z + -6.485021904942025371773e1;
        double q = z + 2.485846490142306297962e1;
        q = q * z + 1.650270098316988542046e2;
        q = q * z + 4.328810604912902668951e2;
        q = q * z + 4.853903996359136964868e2;
        q = q * z + 1.945506571482613964425e2;
        z = z * p / q;
    }
    z = x * z + x;
    if (flag == 2) z += 0.5 * MOREBITS;
    else if (flag == 1) z += MOREBITS;
    y = y + z;
    return sign ? -y : y;
}

/* atan2(y, x), Cephes atan2 quadrant logic. */
SEVC_HD double sevc_atan2(double y, double x) {
    int code = 0;
    double w;
    if (x < 0.0) code = 2;
    if (y < 0.0) code |= 1;
    if (x == 0.0) {
        if (code & 1) return -SEVC_PIO2;
        if (y == 0.0) return 0.0;
        return SEVC_PIO2;
    }
    if (y == 0.0) {
        if (code & 2) return SEVC_PI;
        return 0.0;
    }
    switch (code) {
        case 2: w = SEVC_PI; break;
        case 3: w = -SEVC_PI; break;
        default: w = 0.0; break;
    }
    return w + sevc_atan(y / x);
}

/* Shared octant reduction of Cephes sin.c / cos.c: returns the reduced
 * argument z in [-pi/4, pi/4] and the octant index j in 0..3 plus a sign
 * flip.  Extended-precision pi/4 split DP1 + DP2 + DP3. */
SEVC_HD double sevc__reduce_pio4(double x, int* jout) {
    const double DP1 = 7.85398125648498535156e-1;
    const double DP2 = 3.77489470793079817668e-8;
    const double DP3 = 2.69515142907905952645e-15;
    const double FOPI = 1.27323954473516268615;   /* 4 / pi */
    double y = floor(x * FOPI);
    double z = y - 16.0 * floor(y * 0.0625);       /* y mod 16, exact */
    int j = (int)z;
    if (j & 1) { j += 1; y += 1.0; }
    *jout = j & 7;
    return ((x - y * DP1) - y * DP2) - y * DP3;
}

SEVC_HD double sevc__sin_poly(double z, double zz) {
    double p = 1.58962301576546568060e-10;
    p = p * zz + -2.50507477628578072866e-8;
    p = p * zz + 2.75573136213857245213e-6;
    p = p * zz + -1.98412698295895385996e-4;
    p = p * zz + 8.33333333332211858878e-3;
    p = p * zz + -1.66666666666666307295e-1;
    return z + z * (zz * p);
}

SEVC_HD double sevc__cos_poly(double zz) {
    double p = -1.13585365213876817300e-11;
    p = p * zz + 2.08757008419747316778e-9;
    p = p * zz + -2.75573141792967388112e-7;
    p = p * zz + 2.48015872888517045348e-5;
    p = p * zz + -1.38888888888730564116e-3;
    p = p * zz + 4.16666666666665929218e-2;
    return 1.0 - 0.5 * zz + zz * zz * p;
}

SEVC_HD double sevc_sin(double x) {
    int sign = 0, j;
    if (x == 0.0) return x;
    if (x < 0.0) { sign = 1; x = -x; }
    double z = sevc__reduce_pio4(x, &j);
    if (j > 3) { sign = !sign; j -= 4; }
    double zz = z * z;
    double y = (j == 1 || j == 2) ? sevc__cos_poly(zz) : sevc__sin_poly(z, zz);
    return sign ? -y : y;
}

SEVC_HD double sevc_cos(double x) {
    int sign = 0, j;
    if (x < 0.0) x = -x;
    double z = sevc__reduce_pio4(x, &j);
    if (j > 3) { j -= 4; sign = !sign; }
    if (j > 1) sign = !sign;
    double zz = z * z;
    double y = (j == 1 || j == 2) ? sevc__sin_poly(z, zz) : sevc__cos_poly(zz);
    return sign ? -y : y;
}

/* sin(x) and cos(x) of one argument with one octant reduction: each value
 * is the one sevc_sin / sevc_cos returns (the same reduction of |x|, the
 * same polynomials and the same octant and sign rules). */
SEVC_HD void sevc_sincos(double x, double* s, double* c) {
    int ssign = 0, csign = 0, j;
    double ax = x;
    if (x < 0.0) { ssign = 1; ax = -x; }
    double z = sevc__reduce_pio4(ax, &j);
    if (j > 3) { j -= 4; ssign = !ssign; csign = !csign; }
    if (j > 1) csign = !csign;
    double zz = z * z;
    double ps = sevc__sin_poly(z, zz), pc = sevc__cos_poly(zz);
    double ys = (j == 1 || j == 2) ? pc : ps;
    double yc = (j == 1 || j == 2) ? ps : pc;
    *s = (x == 0.0) ? x : (ssign ? -ys : ys);
    *c = csign ? -yc : yc;
}

/* exp(x), Cephes exp.c: x = n ln2 + r with a two-part ln2, then the Pade
 * form 1 + 2 r P(r^2) / (Q(r^2) - r P(r^2)), scaled by 2^n. */
SEVC_HD double sevc_exp(double x) {
    const double LOG2E = 1.4426950408889634073599;
    const double C1 = 6.93145751953125e-1;
    const double C2 = 1.42860682030941723212e-6;
    if (x != x) return x;
    if (x > 709.78271289338399) return HUGE_VAL;
    if (x < -745.13321910194110) return 0.0;
    double px = floor(LOG2E * x + 0.5);
    int n = (int)px;
    x = x - px * C1;
    x = x - px * C2;
    double xx = x * x;
    double p = 1.26177193074810590878e-4;
    p = p * xx + 3.02994407707441961300e-2;
    p = p * xx + 9.99999999999999999910e-1;
    p = x * p;
    double q = 3.00198505138664455042e-6;
    q = q * xx + 2.52448340349684104192e-3;
    q = q * xx + 2.27265548208155028766e-1;
    q = q * xx + 2.00000000000000000009e0;
    x = p / (q - p);
    x = 1.0 + 2.0 * x;
    return ldexp(x, n);
}

#endif /* SEVC_MATH_H */
