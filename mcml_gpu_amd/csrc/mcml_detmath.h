/*
 * mcml_detmath.h - deterministic arithmetic shared by the HIP transport kernels and any CPU
 * code that must reproduce their photon trajectories bit for bit.
 *
 * Every function here is built only from IEEE-754 binary32 operations that are correctly
 * rounded on both x86-64 (SSE/FMA) and gfx950 (VALU): +, -, *, /, sqrtf, fmaf, rintf,
 * integer arithmetic and bit moves.  Compile every user with -ffp-contract=off and without
 * fast-math, so that no a*b+c is silently fused on one side and not on the other.
 *
 * Contents
 *   - multiply-with-carry uniform generator (Marsaglia 2003) with per-photon seeding by the
 *     SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014);
 *   - the step length -ln(u) and the azimuth's sin/cos straight from a 32-bit draw, both
 *     table-driven (tables in mcml_tables.h, generated by gen_tables.py; the callers pass
 *     them, the GPU kernels from LDS): Tang's logarithm (ACM TOMS 16(4), 1990) with a
 *     Cody-Waite split of ln 2, and angle addition from 256 tabulated angles;
 *   - sin/cos of 2*pi*u for launch angles: exact quarter-turn reduction + near-minimax
 *     polynomials on |a| <= pi/4 (fit_sincos.py; reduction after Cody & Waite, "Software
 *     Manual for the Elementary Functions", 1980);
 *   - acos: asin Taylor series (A&S 4.4.40) on [0, 1/2] and acos(x) = 2 asin(sqrt((1-x)/2)).
 * The asin coefficients are the exact rational Taylor coefficients rounded to binary32; the
 * launch sin/cos coefficients come from fit_sincos.py.
 * None of these reproduce any vendor math library; they are accurate to a few ulp and,
 * more importantly, give identical bits on the host and on the GPU.
 */
#ifndef MCML_DETMATH_H
#define MCML_DETMATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#define MCML_HD __host__ __device__ static inline
#else
#define MCML_HD static inline
#endif

/* ---------------------------------------------------------------- bit moves */
MCML_HD uint32_t mcml_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
MCML_HD float mcml_u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* ---------------------------------------------------------------- RNG */
/* One photon's random stream: lag-1 multiply-with-carry, base 2^32.
 * state = carry * 2^32 + x; next state = a * x + carry (fits in 64 bits for a < 2^32). */
typedef struct {
    uint64_t s;
    uint32_t a;
} mcml_mwc;

MCML_HD uint32_t mcml_mwc_next(mcml_mwc* r)
{
    r->s = (uint64_t)r->a * (uint64_t)(uint32_t)r->s + (r->s >> 32);
    return (uint32_t)r->s;
}

/* Uniform in (0, 1]: the low 23 bits k of a draw become 1 - k 2^-23, computed exactly as 2
 * minus the float 1.k in [1, 2) built from the bits (one and-or and one subtraction, no
 * conversion).  Never 0, so -log(u) is finite (the same half-open interval as the
 * reference's generator); the smallest value is 2^-23.  Every bit of a multiply-with-carry
 * output is a digit bit of the expansion of a fraction with a prime denominator, the low
 * ones as good as the high ones (tests/test_rng.py checks the uniforms' moments, lag
 * correlation and chi-square). */
/* 1 + k 2^-23 in [1, 2) from the low 23 bits k of a draw: the uniform below is 2 minus it. */
MCML_HD float mcml_one_plus(uint32_t x) { return mcml_u2f(0x3F800000u | (x & 0x007FFFFFu)); }

MCML_HD float mcml_uniform(mcml_mwc* r)
{
    return 2.0f - mcml_one_plus(mcml_mwc_next(r));
}

MCML_HD uint64_t mcml_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#define MCML_GOLDEN 0x9E3779B97F4A7C15ull

/* Stream of photon `k` of parameter set `set` under run seed `seed`.  The stream depends
 * on (seed, set, k) only, so it does not matter which lane, thread or pass runs it.
 * `table` holds MCML_MWC_TABLE_SIZE multipliers a in (2^31, 2^32) with a*2^32-1 a safe
 * prime; the table size must be a power of two no larger than 4096. */
MCML_HD mcml_mwc mcml_stream(uint64_t seed, uint32_t set, uint64_t k, const uint32_t* table,
                             uint32_t table_size)
{
    mcml_mwc r;
    uint64_t h = mcml_mix64(seed + MCML_GOLDEN * ((uint64_t)set + 1u));
    h = mcml_mix64(h ^ (k * 0xD1B54A32D192ED03ull + MCML_GOLDEN));
    uint32_t a = table[(uint32_t)(h >> 52) & (table_size - 1u)];
    uint64_t h2 = mcml_mix64(h + MCML_GOLDEN);
    uint32_t x = (uint32_t)h2;
    uint32_t c = (uint32_t)(h2 >> 32);
    if (c >= a) c -= a; /* a > 2^31, so one subtraction reduces c into [0, a) */
    if (c == 0u && x == 0u) x = 1u;                   /* absorbing zero state */
    if (c == a - 1u && x == 0xFFFFFFFFu) x = 0xFFFFFFFEu; /* the other fixed point */
    r.s = ((uint64_t)c << 32) | x;
    r.a = a;
    return r;
}

/* ---------------------------------------------------------------- step length */
/* One table entry: two floats read together (one 64-bit load from LDS on the GPU). */
typedef struct __attribute__((aligned(8))) {
    float a, b;
} mcml_pair;

#define MCML_LN2 0.693147182f /* ln 2 rounded to binary32 */

/* 0x4B000000 | (b >> 23): the funnel shift (0x00258000:b) >> 23 on the GPU (one v_alignbit),
 * the same bits on the host */
MCML_HD uint32_t mcml_exp_magic(uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(0x00258000u, b, 23u);
#else
    return 0x4B000000u | (b >> 23);
#endif
}

/* -ln(u), the exponential step of the photon walk, for the uniform u = (2k + 1) 2^-24 drawn
 * from the low 23 bits k of one 32-bit draw x: the midpoint of one of 2^23 equal cells of
 * (0, 1), as curand_uniform's (x + 1/2) 2^-32 is at 32 bits, formed exactly as
 * 1.k - (1 - 2^-24).  So 2^-24 <= u <= 1 - 2^-24 and every step is positive.  Table-driven
 * (Tang, ACM TOMS 16(4), 1990): u = 2^e m with m in [1, 2) in cell j of 128 (its top seven
 * fraction bits); with the tabulated c_j ~ 1 / (1 + j/128) and L_j = -ln(c_j) (mcml_tables.h,
 * `tab` = mcml_lg_table), r = m c_j - 1 (one fma, 0 <= r < 1/128 up to rounding) and
 * ln u = e ln 2 + L_j + ln(1 + r),  ln(1 + r) = r (1 + r (r/3 - 1/2))  (error < r^4/4 < 1e-9),
 * evaluated as three fmas ending on L_j.  e ln 2 is one fma with ln 2 rounded to binary32:
 * -24 <= e <= -1, so the constant adds at most 4.6e-8 where -ln u is near 16.6.  Over all
 * 2^23 inputs the error is below 6.3e-8 absolute where -ln u < 1 and 8.7e-8 relative above,
 * and the smallest step is 5.96e-8 mean free paths (tests/test_detmath.py checks every k). */
MCML_HD float mcml_neglog_u32(uint32_t x, const mcml_pair* tab)
{
    const float u = mcml_one_plus(x) - 0x1.fffffep-1f; /* exact */
    const uint32_t b = mcml_f2u(u);
    const uint32_t j = (b >> 16) & 127u;
    const float m = mcml_u2f((b & 0x007FFFFFu) | 0x3F800000u);
    /* e = (biased exponent) - 127, exactly: 2^23 + biased exponent as a float (its bits
     * 0x4B000000 | b >> 23, one funnel shift on the GPU) minus 2^23 + 127 */
    const float le = mcml_u2f(mcml_exp_magic(b)) - 8388735.0f;
    const mcml_pair t = tab[j];
    const float c = t.a, L = t.b;
    const float r = fmaf(m, c, -1.0f);
    const float q = fmaf(r, 0.333333343f, -0.5f);
    const float lg = fmaf(le, MCML_LN2, fmaf(r, fmaf(q, r, 1.0f), L));
    return -lg;
}

/* ---------------------------------------------------------------- sin / cos */
#define MCML_TWO_PI 6.28318530717958647692f
#define MCML_INV_TWO_PI 0.159154943091895335768f

/* sin(2*pi*t) and cos(2*pi*t) for 0 <= t <= 1 (any |t| < 2^22 works); used at launch. */
MCML_HD void mcml_sincos_turns(float t, float* sn, float* cs)
{
    float q = rintf(t * 4.0f);          /* nearest quarter turn; t*4 is exact */
    float f = t - q * 0.25f;            /* exact: |f| <= 1/8 */
    int32_t k = ((int32_t)q) & 3;
    float a = f * MCML_TWO_PI;          /* |a| <= pi/4 */
    float z = a * a;
    /* near-minimax on |a| <= pi/4 (fit_sincos.py): sin a = a + a^3 P(a^2), cos a = 1 + a^2 Q(a^2),
     * errors 2.6e-9 and 7.1e-11 before float rounding */
    float ps = fmaf(z, -1.951813902e-04f, 8.332200348e-03f);
    ps = fmaf(z, ps, -1.666665524e-01f);
    float s = fmaf(a * z, ps, a);
    float pc = fmaf(z, 2.440824028e-05f, -1.388699282e-03f);
    pc = fmaf(z, pc, 4.166663438e-02f);
    pc = fmaf(z, pc, -0.5f);
    float c = fmaf(z, pc, 1.0f);
    /* quadrant k: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s) -- as selects and sign
     * flips, with no data-dependent branch */
    const int32_t odd = k & 1;
    float rs = odd ? c : s;
    float rc = odd ? s : c;
    rs = mcml_u2f(mcml_f2u(rs) ^ ((uint32_t)(k & 2) << 30));
    rc = mcml_u2f(mcml_f2u(rc) ^ ((uint32_t)((k + 1) & 2) << 30));
    *sn = rs;
    *cs = rc;
}

/* sin and cos of the azimuth 2 pi x / 2^32 for one 32-bit draw x.  Table-driven: the top
 * eight bits j pick (c, s) = (cos, sin) of 2 pi j / 256 (mcml_tables.h, `tab` =
 * mcml_sc_table), the low 24 bits rho give the remaining angle d = K rho < 2 pi / 256 =
 * 0.0246 (K = 2 pi / 2^32), and the angle addition reads
 *   cos(a + d) = c - s sin d - c (1 - cos d),  sin(a + d) = s + c sin d - s (1 - cos d)
 * with sin d = rho (K - (K^3/6) rho^2) (error < d^5/120 = 7.4e-11) and 1 - cos d =
 * (K^2/2) rho^2 (error < d^4/24 = 1.5e-8): four operations for the two polynomials and two
 * fmas for each result.  Both results are within 1.1e-7 of float64 (tests/test_detmath.py). */
#define MCML_SC_K1 0x1.921fb6p-30f /* 2 pi / 2^32 */
#define MCML_SC_K2 0x1.3bd3ccp-60f /* (2 pi / 2^32)^2 / 2 */
#define MCML_SC_K3 0x1.4abbcep-91f /* (2 pi / 2^32)^3 / 6 */
MCML_HD void mcml_sincos_u32(uint32_t x, const mcml_pair* tab, float* sn, float* cs)
{
    const uint32_t j = x >> 24;
    const float rho = (float)(x & 0x00FFFFFFu);
    const float r2 = rho * rho;
    const float sd = rho * fmaf(-MCML_SC_K3, r2, MCML_SC_K1); /* sin d */
    const float vd = MCML_SC_K2 * r2;                         /* 1 - cos d */
    const mcml_pair t = tab[j];
    const float ca = t.a, sa = t.b;
    *sn = fmaf(ca, sd, fmaf(-sa, vd, sa));
    *cs = fmaf(-sa, sd, fmaf(-ca, vd, ca));
}

/* sin and cos of an angle in radians, |a| <= 4 (used for launch angles only). */
MCML_HD void mcml_sincos(float a, float* sn, float* cs)
{
    mcml_sincos_turns(a * MCML_INV_TWO_PI, sn, cs);
}

/* ---------------------------------------------------------------- acos */
/* asin(x) for |x| <= 1/2 from its Taylor series through x^23 (truncation < 2e-9 rel). */
MCML_HD float mcml_asin_small(float x)
{
    float z = x * x;
    float p = fmaf(z, 7.312525874e-03f, 8.390335810e-03f);
    p = fmaf(z, p, 9.761609529e-03f);
    p = fmaf(z, p, 1.155180090e-02f);
    p = fmaf(z, p, 1.396484375e-02f);
    p = fmaf(z, p, 1.735276442e-02f);
    p = fmaf(z, p, 2.237215909e-02f);
    p = fmaf(z, p, 3.038194444e-02f);
    p = fmaf(z, p, 4.464285714e-02f);
    p = fmaf(z, p, 7.500000000e-02f);
    p = fmaf(z, p, 1.666666667e-01f);
    return fmaf(x * z, p, x);
}

#define MCML_PIO2_HI 1.57079637050628662109f
#define MCML_PIO2_LO -4.37113900018624283e-8f
#define MCML_PI_HI 3.14159274101257324219f
#define MCML_PI_LO -8.74227800037248566e-8f

/* acos(t) for |t| <= 1; NaN outside [-1, 1] and for NaN input, like acosf. */
MCML_HD float mcml_acos(float t)
{
    float a = fabsf(t);
    if (!(a <= 1.0f)) return mcml_u2f(0x7FC00000u);
    if (a <= 0.5f) return MCML_PIO2_HI - (mcml_asin_small(t) - MCML_PIO2_LO);
    float w = sqrtf((1.0f - a) * 0.5f);
    float r = 2.0f * mcml_asin_small(w); /* acos(|t|) */
    if (t > 0.0f) return r;
    return MCML_PI_HI - (r - MCML_PI_LO);
}

#endif /* MCML_DETMATH_H */
