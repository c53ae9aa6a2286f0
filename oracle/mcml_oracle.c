/*
 * mcml_oracle.c - CPU restatement of the reference photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared with;
 * only the tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * No product code links or calls it.
 *
 * It restates, one function per reference routine, the semantics of
 *   transport.cu:335-1000  (MCd, MCd_replay, LaunchPhoton, Spin, Reflect, PhotonSurvive,
 *                           detect, detect_replay, the A_rz/A0_z bookkeeping)
 *   transport.cu:274-332   (host reductions)
 *   io.cpp:179-237         (derived quantities)
 * with three deliberate substitutions, made identically in the HIP kernels:
 *   1. cuRAND XORWOW (transport.cu:51-55, 733) -> a per-photon multiply-with-carry stream
 *      keyed by (seed, set, photon index) (mcml_stream in mcml_detmath.h);
 *   2. CUDA fast intrinsics (__logf, __sincosf, __fdividef, rsqrtf) -> correctly rounded
 *      IEEE operations and the deterministic log/sincos/acos of mcml_detmath.h;
 *   3. acos(|u|) <= theta_c (transport.cu:887, 904, 911) -> |u| >= cos(theta_c), the same
 *      test by monotonicity of acos, with cos(theta_c) computed once per detector.
 * Everything else (operation order, weight arithmetic in integer-valued float "W-units",
 * the RNG draw order, quirks listed in the function comments) follows the reference.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off, no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/mcml.h"
#include "../mcml_gpu_amd/csrc/mcml_detmath.h"
#include "../mcml_gpu_amd/csrc/mwc_multipliers.h"
#include "../mcml_gpu_amd/csrc/mcml_tables.h"

#define O_PI 3.141592654f  /* header.h:40 */
#define O_RPI 0.318309886f /* header.h:41 */

typedef struct {
    float x, y, z, dx, dy, dz, w;
    int32_t layer;
    int32_t steps;        /* p.scatter_event (first pass) / f_r.scatter_event (replay) */
    int32_t first_scatter;
    mcml_mwc rng;
} oracle_photon;

/* ------------------------------------------------------------------ derivation */
/* read_mua_mus, io.cpp:179-237 (per simulation), plus per-probe trigonometry. */
int oracle_derive(const mcml_sim* s, mcml_derived* d)
{
    memset(d, 0, sizeof(*d));
    int L = s->number_layers, S = s->num_sds;
    if (L < 1 || L > MCML_MAX_LAYERS || S < 1 || S > MCML_MAX_SDS) return MCML_ERR_LIMIT;
    d->number_photons = s->number_photons;
    d->num_layers = L;
    d->num_sds = S;
    d->detector_reflectance = s->detector_reflectance;
    d->source_oblique = s->source_probe_oblique;
    d->detector_oblique = s->detector_probe_oblique;

    d->layer[0].n = s->upper_n; /* io.cpp:204 */
    float zmin = 0.0f, zmax = 0.0f;
    for (int l = 1; l <= L; l++) { /* io.cpp:210-221 */
        float h = (l < L) ? s->layers[l - 1].thickness : 0.0f; /* last thickness unread (io.cpp:146) */
        if (!(h >= 0.0f)) return MCML_ERR_INPUT; /* as mcml_derive */
        zmax += h;
        d->layer[l].n = s->layers[l - 1].n;
        d->layer[l].mua = s->layers[l - 1].mua;
        d->layer[l].g = s->layers[l - 1].g;
        d->layer[l].z_min = zmin;
        d->layer[l].z_max = zmax;
        d->layer[l].mutr = 1.0f / (s->layers[l - 1].mua + s->layers[l - 1].mus);
        zmin += h;
    }
    if (d->layer[L].z_max < 10.0f) d->layer[L].z_max = 10.0f; /* io.cpp:223-225 */
    d->layer[L + 1].n = s->lower_n;                            /* io.cpp:228 */

    double n1 = d->layer[0].n, n2 = d->layer[1].n; /* io.cpp:231-237 */
    double r = (n1 - n2) / (n1 + n2);
    r = r * r;
    d->start_weight = (float)(unsigned int)((double)0xffffffffu * (1 - r));

    /* source (io.cpp:181-184) */
    float a = s->source.angle_deg * O_PI / 180;
    d->src_radius = s->source.radius;
    d->src_na = s->source.NA;
    d->src_theta_max = asinf(s->source.NA / d->layer[0].n); /* transport.cu:703 */
    d->src_sin_a = (float)sin((double)a);
    d->src_cos_a = (float)cos((double)a);
    for (int i = 1; i <= S; i++) { /* io.cpp:186-197 */
        const mcml_probe* p = &s->detectors[i - 1];
        float ang = p->angle_deg * O_PI / 180;
        d->det_pos[i] = p->pos;
        d->det_radius[i] = p->radius;
        d->det_r2[i] = p->radius * p->radius;
        d->det_cos_a[i] = (float)cos((double)ang);
        d->det_sin_a[i] = (float)sin((double)ang);
        d->det_crit[i] = asinf(p->NA / s->upper_n);
        d->det_cos_crit[i] = (float)cos((double)d->det_crit[i]);
    }
    d->grid_nr = s->grid_nr > 0 ? s->grid_nr : MCML_DEFAULT_NR;
    d->grid_nz = s->grid_nz > 0 ? s->grid_nz : MCML_DEFAULT_NZ;
    d->grid_dr = s->grid_dr > 0 ? s->grid_dr : MCML_DEFAULT_DR;
    d->grid_dz = s->grid_dz > 0 ? s->grid_dz : MCML_DEFAULT_DZ;
    return MCML_OK;
}

/* ------------------------------------------------------------------ photon life */
static float urand(oracle_photon* p) { return mcml_uniform(&p->rng); }

/* LaunchPhoton, transport.cu:624-719.  Quirks kept: oblique source ignores NA/radius;
 * launch radius is linear in u (not sqrt u); NA cone is uniform in theta and not refracted. */
void oracle_launch(const mcml_derived* d, uint64_t seed, uint32_t set, uint64_t k, oracle_photon* p)
{
    p->rng = mcml_stream(seed, set, k, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    if (d->source_oblique) {
        p->x = 0.0f; p->y = 0.0f; p->z = 0.0f;
        p->dx = d->src_sin_a; p->dy = 0.0f; p->dz = d->src_cos_a;
    } else {
        if (d->src_radius == 0.0f) {
            p->x = 0.0f; p->y = 0.0f; p->z = 0.0f;
        } else {
            float u_pos = urand(p);
            float u_az = urand(p);
            float rho = d->src_radius * u_pos;
            float sa, ca;
            mcml_sincos_turns(u_az, &sa, &ca);
            p->x = rho * sa;
            p->y = rho * ca;
            p->z = 0.0f;
        }
        if (d->src_na == 0.0f) {
            p->dx = 0.0f; p->dy = 0.0f; p->dz = 1.0f;
        } else {
            float u_dir = urand(p);
            float u_rot = urand(p);
            float th = d->src_theta_max * u_dir;
            float st, ct, sr, cr;
            mcml_sincos(th, &st, &ct);
            mcml_sincos_turns(u_rot, &sr, &cr);
            p->dz = ct;
            p->dx = st * cr;
            p->dy = st * sr;
        }
    }
    p->layer = 1;
    p->first_scatter = 1;
    p->steps = 0;
    p->w = d->start_weight;
}

/* Spin, transport.cu:742-786: Henyey-Greenstein (or isotropic for g == 0) deflection,
 * uniform azimuth, MCML direction update with the |dz| == 1 special case, renormalise.
 * Same mathematics, fewer divisions and square roots than the reference's expression:
 *   cos(theta) = c0 - a2 (1/D)^2, c0 = (1 + g^2)(1/2g), a2 = (1 - g^2)^2 (1/2g)
 *                                                         (= (1 + g^2 - t^2) / 2g,
 *                                                          t = (1 - g^2) / D, the
 *                                                          reference's expression)
 *   P = sqrt((1 - cos^2)(1 - dz^2)), Q = P / (1 - dz^2)    (= sin(theta) * sqrt(1-dz^2) and
 *                                                           sin(theta) / sqrt(1-dz^2))
 *   with one reciprocal r = 1 / (D (1 - dz^2)), D = 1 - g + 2gu, for both divisions
 *   (D formed from v = 2 - u, the float 1.k the uniform is built from):
 *   1/D = (1 - dz^2) r and Q = P (D r); P = sqrtf(x), the
 *   reciprocal and the square root both correctly rounded; at u = 0 or 1 the inversion can
 *   round cos(theta) past +-1, where the reference's sqrtf(1 - cos^2) is NaN: here
 *   sin(theta) is 0 (transport.cu:755-759)
 *   renormalisation by 1 - d/2, d = |u|^2 - 1               (first-order rsqrt about 1:
 *                                                           |d| stays at a few ulp, so
 *                                                           3d^2/8 is below 1e-13)
 * Returns cos(theta) for the tests. */
/* The spin's square roots: sqrtf(x) for x > 0, else 0 (sin(theta) = 0 where cos(theta)
 * rounded to or past +-1). */
static float spin_sqrt(float x) { return x > 0.0f ? sqrtf(x) : 0.0f; }

float oracle_spin(oracle_photon* p, float g)
{
    /* u = 2 - v with v = 1.k in [1, 2) (mcml_uniform): the HG denominator 1 - g + 2gu is
     * evaluated as (2(2g) + (1 - g)) - (2g) v, and 2u - 1 as 3 - 2v (exact) */
    const float v = mcml_one_plus(mcml_mwc_next(&p->rng));
    const int iso = g == 0.0f;
    const float D = fmaf(-(2.0f * g), v, 2.0f * (2.0f * g) + (1.0f - g));
    /* cos(theta) = (1 + g^2 - ((1 - g^2) / D)^2) / (2g) = c0 - a2 / D^2, with the per-layer
     * constants formed as the library forms them (runtime.cpp fill_devset) */
    const float inv2g = 1.0f / (2.0f * g);
    const float hg_a = 1.0f - g * g;
    const float a2 = (hg_a * hg_a) * inv2g, c0 = (1.0f + g * g) * inv2g;
    float sinp, cosp;
    mcml_sincos_u32(mcml_mwc_next(&p->rng), mcml_sc_table, &sinp, &cosp);
    float b = fmaf(-p->dz, p->dz, 1.0f);
    float cost;
    if (b == 0.0f) { /* travelling along z */
        cost = fmaf(-2.0f, v, 3.0f);
        if (!iso) {
            float t = 1.0f / D;
            cost = fmaf(-(t * t), a2, c0);
        }
        float sint = spin_sqrt(fmaf(-cost, cost, 1.0f)); /* cos(theta) may round past +-1 */
        p->dx = sint * cosp;
        p->dy = sint * sinp;
        p->dz = copysignf(cost, p->dz * cost);
    } else {
        /* one division serves the HG inversion (1/D = b r) and the rotation (1/b = D r) */
        float r = 1.0f / (D * b);
        cost = fmaf(-2.0f, v, 3.0f);
        if (!iso) {
            float br = b * r; /* 1 / D */
            cost = fmaf(-(br * br), a2, c0);
        }
        float ab = fmaf(-cost, cost, 1.0f) * b;
        float P = spin_sqrt(ab); /* cos(theta) may round past +-1: sin(theta) = 0 */
        float Q = P * (D * r);
        /* MCML's update (Wang et al. 1995) with the shared bracket cos + uz Q cos(psi):
         * ux' = ux (cos + uz Q cos psi) - uy Q sin psi, uy' = uy (...) + ux Q sin psi */
        float Qc = Q * cosp, Qs = Q * sinp;
        float k = fmaf(p->dz, Qc, cost);
        float cxp = fmaf(p->dx, k, -(Qs * p->dy));
        float cyp = fmaf(p->dy, k, Qs * p->dx);
        p->dz = fmaf(-P, cosp, p->dz * cost);
        p->dx = cxp;
        p->dy = cyp;
    }
    float d = fmaf(p->dx, p->dx, fmaf(p->dy, p->dy, fmaf(p->dz, p->dz, -1.0f)));
    float r = fmaf(d, -0.5f, 1.0f);
    p->dx = p->dx * r;
    p->dy = p->dy * r;
    p->dz = p->dz * r;
    return cost;
}

/* Reflect, transport.cu:790-853.  1 = reflected (dz mirrored, layer kept),
 * 0 = transmitted (layer updated, direction refracted). */
int oracle_reflect(const mcml_derived* d, oracle_photon* p, int new_layer)
{
    float n1 = d->layer[p->layer].n;
    float n2 = d->layer[new_layer].n;
    float r;
    float ci = fabsf(p->dz);
    if (n1 == n2) {
        p->layer = new_layer;
        return 0;
    }
    if (n1 > n2 && n2 * n2 < n1 * n1 * (1.0f - ci * ci)) { /* total internal reflection */
        p->dz = -p->dz;
        return 1;
    }
    if (ci == 1.0f) { /* normal incidence */
        r = (n1 - n2) / (n1 + n2);
        if (urand(p) <= r * r) {
            p->dz = -p->dz;
            return 1;
        }
        p->layer = new_layer;
        return 0;
    }
    float e = (n1 * n1) / (n2 * n2) * (1.0f - ci * ci);
    /* Just inside the critical angle the test above can round to "not total" while e, the
     * squared sine of the refracted angle, rounds above 1: the reference's sqrtf of the
     * negative arguments here and below is then NaN (transport.cu:831, 848) and the photon
     * is lost as NaN.  Taking them as 0 gives reflectance 1 to rounding (sin = 1, total
     * reflection) and a grazing direction otherwise. */
    r = 2.0f * sqrtf(fmaxf((1.0f - ci * ci) * (1.0f - e) * e * ci * ci, 0.0f));
    e = e + (ci * ci) * (1.0f - 2.0f * e);
    r = e * ((1.0f - e - r) / ((1.0f - e + r) * (e + r)));
    if (urand(p) <= r) {
        p->dz = -p->dz;
        return 1;
    }
    r = n1 / n2;
    e = r * r * (1.0f - ci * ci);
    p->dx = p->dx * r;
    p->dy = p->dy * r;
    p->dz = copysignf(sqrtf(fmaxf(1.0f - e, 0.0f)), p->dz);
    p->layer = new_layer;
    return 0;
}

/* PhotonSurvive, transport.cu:855-870.  The step cap applies to the first pass only:
 * MCd_replay never advances p.scatter_event (transport.cu:528). */
int oracle_survive(oracle_photon* p, int step_cap)
{
    if (step_cap && p->steps >= MCML_MAX_SCATTER) return 0;
    if (p->w > (float)MCML_WEIGHTI) return 1;
    if (p->w == 0.0f) return 0;
    if (urand(p) < MCML_CHANCE) {
        p->w = rintf(p->w / MCML_CHANCE);
        return 1;
    }
    return 0;
}

/* Ring-detector arc fraction weight factor acos(t)/pi, transport.cu:915-922 (t clamped
 * at 1; NaN when t < -1 or pos == 0, as in the reference). */
float oracle_ring_weight(float w, float dist, float pos, float radius)
{
    float t = (dist * dist + pos * pos - radius * radius) / (2.0f * dist * pos);
    if (t > 1.0f) t = 1.0f;
    return w * mcml_acos(t) * O_RPI;
}

/* detect / detect_replay, transport.cu:872-1000.  Tests detector `only` (1-based) or all
 * detectors when only == 0.  Returns the bit mask (bit i = SDS i) and fills wout[i]. */
uint32_t oracle_detect(const mcml_derived* d, const oracle_photon* p, int only, float* wout)
{
    uint32_t mask = 0;
    int i0 = only ? only : 1, i1 = only ? only : d->num_sds;
    if (d->detector_oblique) { /* fiber mode: ellipse footprint + tilted NA cone */
        for (int i = i0; i <= i1; i++) {
            float t = (p->x - d->det_pos[i]) * d->det_cos_a[i];
            if (t * t + p->y * p->y <= d->det_r2[i]) {
                float uz = p->dx * d->det_sin_a[i] + fabsf(p->dz * d->det_cos_a[i]);
                if (fabsf(uz) >= d->det_cos_crit[i]) {
                    mask |= 1u << i;
                    wout[i] = p->w;
                }
            }
        }
    } else { /* ring mode: annulus + NA cone, arc-fraction weight */
        float adz = fabsf(p->dz);
        float dist = sqrtf(p->x * p->x + p->y * p->y);
        for (int i = i0; i <= i1; i++) {
            if (adz >= d->det_cos_crit[i]) {
                if (dist >= d->det_pos[i] - d->det_radius[i] && dist <= d->det_pos[i] + d->det_radius[i]) {
                    mask |= 1u << i;
                    wout[i] = oracle_ring_weight(p->w, dist, d->det_pos[i], d->det_radius[i]);
                }
            }
        }
    }
    return mask;
}

/* Bins of the absorbance grid, transport.cu:583, 588 (trunc then clamp to the last bin).
 * The reference's __fdividef(z, dz) is a product with a reciprocal; here the reciprocal is
 * the correctly rounded 1/dz. */
static int zbin_of(const mcml_derived* d, float z)
{
    float q = z * (1.0f / d->grid_dz);
    q = fminf(fmaxf(q, 0.0f), (float)(d->grid_nz - 1));
    return (int)q;
}
static int rbin_of(const mcml_derived* d, float x, float y)
{
    float q = sqrtf(x * x + y * y) * (1.0f / d->grid_dr);
    q = fminf(fmaxf(q, 0.0f), (float)(d->grid_nr - 1));
    return (int)q;
}

typedef struct {
    uint64_t steps, boundary, interaction;
} oracle_counts;

/* Optional record of every interaction of one photon (test hook): per interaction
 * ORACLE_TRACE_COLS floats (x, y, z, drop, first, layer), taken where the reference books
 * the drop (transport.cu:573-600: after the hop, before the spin). */
#define ORACLE_TRACE_COLS 6
typedef struct {
    float* buf;
    int64_t cap; /* rows */
    int64_t n;   /* rows seen (may exceed cap) */
} oracle_trace_buf;

/* One photon through MCd's loop body (transport.cu:370-453) when target == 0, or through
 * MCd_replay's (transport.cu:506-610) when target is an SDS number.  Returns the detection
 * mask (0 = not detected); wout[i] holds the detected weight per SDS.  In replay mode
 * `pl` (L floats) receives the per-layer pathlength and the grids (if non-NULL) the drops. */
static uint32_t photon_life(const mcml_derived* d, oracle_photon* p, int target, float* wout, float* pl,
                            uint64_t* A_rz, uint64_t* A0_z, oracle_counts* cnt, oracle_trace_buf* tr);

uint32_t oracle_photon_life(const mcml_derived* d, oracle_photon* p, int target, float* wout, float* pl,
                            uint64_t* A_rz, uint64_t* A0_z, oracle_counts* cnt)
{
    return photon_life(d, p, target, wout, pl, A_rz, A0_z, cnt, NULL);
}

/* The distance to the boundary a hop is cut at, a / dz: the quotient formed from the correctly
 * rounded reciprocal y = 1/b as q = a y, then q + (a - b q) y (Markstein), as the kernel forms
 * it.  It is the IEEE quotient a / b whenever neither over- nor underflows (Markstein's
 * theorem; no difference on 4e8 random pairs of the hop's range), so this is the reference's
 * division (transport.cu:382, 388) in another form. */
static float cut_div(float a, float b)
{
    const float y = 1.0f / b;
    const float q = a * y;
    return fmaf(fmaf(-b, q, a), y, q);
}

static uint32_t photon_life(const mcml_derived* d, oracle_photon* p, int target, float* wout, float* pl,
                            uint64_t* A_rz, uint64_t* A0_z, oracle_counts* cnt, oracle_trace_buf* tr)
{
    const int replay = target != 0;
    uint32_t detected = 0;
    for (;;) {
        const mcml_layer_derived* Ly = &d->layer[p->layer];
        float s = mcml_neglog_u32(mcml_mwc_next(&p->rng), mcml_lg_table) * Ly->mutr; /* transport.cu:373 */
        int new_layer = p->layer;
        if (fmaf(s, p->dz, p->z) < Ly->z_min) { /* transport.cu:379-386 */
            new_layer--;
            s = cut_div(Ly->z_min - p->z, p->dz);
        }
        if (fmaf(s, p->dz, p->z) > Ly->z_max) {
            new_layer++;
            s = cut_div(Ly->z_max - p->z, p->dz);
        }
        p->x = fmaf(p->dx, s, p->x);
        p->y = fmaf(p->dy, s, p->y);
        p->z = fmaf(p->dz, s, p->z);
        p->steps++;
        if (replay) pl[p->layer - 1] += s; /* transport.cu:529 */
        if (p->z > Ly->z_max) p->z = Ly->z_max;
        if (p->z < Ly->z_min) p->z = Ly->z_min;
        if (cnt) cnt->steps++;

        if (new_layer != p->layer) { /* transport.cu:397-428 */
            int temp_layer = p->layer;
            s = 0.0f;
            if (cnt) cnt->boundary++;
            if (oracle_reflect(d, p, new_layer) == 0) {
                if (new_layer == 0) {
                    uint32_t m = oracle_detect(d, p, target, wout);
                    if (m) {
                        detected = m;
                        p->w = 0.0f;
                        if (replay) break;
                    } else if (urand(p) > d->detector_reflectance) {
                        p->w = 0.0f;
                        if (replay) break;
                    } else { /* detector reflects: back into the tissue, refracted direction kept */
                        p->dz = -p->dz;
                        p->z = -p->z;
                        p->layer = temp_layer;
                    }
                }
                if (new_layer > d->num_layers) {
                    p->w = 0.0f;
                    if (replay) break;
                }
            }
        }
        if (s > 0.0f) { /* transport.cu:430-437 / 573-603 */
            const mcml_layer_derived* Lc = &d->layer[p->layer];
            float w_temp = rintf(Lc->mua * Lc->mutr * p->w);
            p->w -= w_temp;
            if (cnt) cnt->interaction++;
            if (tr) {
                if (tr->n < tr->cap) {
                    float* row = tr->buf + tr->n * ORACLE_TRACE_COLS;
                    row[0] = p->x; row[1] = p->y; row[2] = p->z;
                    row[3] = w_temp; row[4] = (float)p->first_scatter; row[5] = (float)p->layer;
                }
                tr->n++;
            }
            if (replay && A0_z && A_rz) {
                int zb = zbin_of(d, p->z);
                if (p->first_scatter) {
                    A0_z[zb] += (uint64_t)w_temp;
                } else {
                    int rb = rbin_of(d, p->x, p->y);
                    A_rz[(size_t)zb * d->grid_nr + rb] += (uint64_t)w_temp;
                }
            }
            p->first_scatter = 0;
            oracle_spin(p, Lc->g);
        }
        if (!oracle_survive(p, !replay)) break;
    }
    return detected;
}

/* ------------------------------------------------------------------ exported drivers */

/* First pass over photons [k0, k1) of one set (MCd + calculate_reflectance,
 * transport.cu:114-147, 299-311).  Records come out in photon order, ascending SDS within
 * a photon.  Returns the record count (may exceed cap: then only cap were written), or a
 * negative error.  counts[3] = {steps, boundary, interaction}.  nthreads <= 0: OpenMP
 * default. */
int64_t oracle_first_pass(const mcml_derived* d, uint64_t seed, uint32_t set, uint64_t k0, uint64_t k1,
                          mcml_record* out, int64_t cap, uint64_t* counts, int nthreads)
{
    if (k1 < k0) return MCML_ERR_ARG;
    int64_t n = (int64_t)(k1 - k0);
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    if (nt > n && n > 0) nt = (int)n;
    if (nt < 1) nt = 1;
    /* per-thread contiguous photon ranges keep the merged output in photon order */
    int64_t* tcount = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
    mcml_record** tbuf = (mcml_record**)calloc((size_t)nt, sizeof(mcml_record*));
    int64_t* tcap = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
    oracle_counts* tc = (oracle_counts*)calloc((size_t)nt, sizeof(oracle_counts));
    int failed = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        float wout[MCML_MAX_SDS + 1];
        for (int64_t i = lo; i < hi; i++) {
            oracle_photon p;
            uint64_t k = k0 + (uint64_t)i;
            oracle_launch(d, seed, set, k, &p);
            uint32_t m = oracle_photon_life(d, &p, 0, wout, NULL, NULL, NULL, &tc[t]);
            for (int s = 1; m && s <= d->num_sds; s++) {
                if (!(m & (1u << s))) continue;
                if (tcount[t] == tcap[t]) {
                    int64_t nc = tcap[t] ? tcap[t] * 2 : 1024;
                    mcml_record* nb = (mcml_record*)realloc(tbuf[t], (size_t)nc * sizeof(mcml_record));
                    if (!nb) { failed = 1; break; }
                    tbuf[t] = nb;
                    tcap[t] = nc;
                }
                mcml_record* r = &tbuf[t][tcount[t]++];
                r->key = ((uint64_t)set << 40) | k;
                r->sds = (uint32_t)s;
                r->weight = wout[s];
            }
        }
    }
    int64_t total = 0;
    for (int t = 0; t < nt; t++) {
        for (int64_t i = 0; i < tcount[t]; i++) {
            if (out && total < cap) out[total] = tbuf[t][i];
            total++;
        }
        if (counts) {
            counts[0] += tc[t].steps;
            counts[1] += tc[t].boundary;
            counts[2] += tc[t].interaction;
        }
        free(tbuf[t]);
    }
    free(tbuf);
    free(tcount);
    free(tcap);
    free(tc);
    return failed ? MCML_ERR_NOMEM : total;
}

/* Replay of photon k of `set` for detector `sds` (MCd_replay + calculate_reflectance_replay,
 * transport.cu:473-622, 315-332).  rec receives [weight, PL_1..PL_L, steps].  A_rz
 * ([nz*nr]) and A0_z ([nz]) of this SDS are accumulated when non-NULL.  Returns 1 when the
 * photon is detected by `sds` again, 0 otherwise. */
int oracle_replay(const mcml_derived* d, uint64_t seed, uint32_t set, uint64_t k, int sds, float* rec,
                  uint64_t* A_rz, uint64_t* A0_z)
{
    oracle_photon p;
    float wout[MCML_MAX_SDS + 1];
    float pl[MCML_MAX_LAYERS];
    memset(pl, 0, sizeof(pl));
    oracle_launch(d, seed, set, k, &p);
    uint32_t m = oracle_photon_life(d, &p, sds, wout, pl, A_rz, A0_z, NULL);
    int L = d->num_layers;
    rec[0] = m ? wout[sds] : 0.0f;
    for (int l = 0; l < L; l++) rec[1 + l] = pl[l];
    rec[L + 1] = (float)p.steps;
    return m ? 1 : 0;
}

/* ------------------------------------------------------------------ test hooks */
/* Replay of photon k for detector `sds` (as oracle_replay, grids optional) that also
 * records each interaction into trace[cap][ORACLE_TRACE_COLS]; returns the number of
 * interactions (only the first cap are written), and *hit = 1 when detected again. */
int64_t oracle_trace(const mcml_derived* d, uint64_t seed, uint32_t set, uint64_t k, int sds, float* trace,
                     int64_t cap, float* rec, uint64_t* A_rz, uint64_t* A0_z, int* hit)
{
    oracle_photon p;
    float wout[MCML_MAX_SDS + 1];
    float pl[MCML_MAX_LAYERS];
    memset(pl, 0, sizeof(pl));
    oracle_trace_buf tr = {trace, cap, 0};
    oracle_launch(d, seed, set, k, &p);
    uint32_t m = photon_life(d, &p, sds, wout, pl, A_rz, A0_z, NULL, &tr);
    int L = d->num_layers;
    rec[0] = m ? wout[sds] : 0.0f;
    for (int l = 0; l < L; l++) rec[1 + l] = pl[l];
    rec[L + 1] = (float)p.steps;
    *hit = m ? 1 : 0;
    return tr.n;
}
/* The absorbance cell of a drop at (x, y, z). */
void oracle_bin(const mcml_derived* d, float x, float y, float z, int* zb, int* rb)
{
    *zb = zbin_of(d, z);
    *rb = rbin_of(d, x, y);
}
size_t oracle_photon_size(void) { return sizeof(oracle_photon); }
void oracle_uniforms(uint64_t seed, uint32_t set, uint64_t k, float* out, int n)
{
    mcml_mwc r = mcml_stream(seed, set, k, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    for (int i = 0; i < n; i++) out[i] = mcml_uniform(&r);
}
void oracle_stream_state(uint64_t seed, uint32_t set, uint64_t k, uint64_t* state, uint32_t* mult)
{
    mcml_mwc r = mcml_stream(seed, set, k, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    *state = r.s;
    *mult = r.a;
}
const uint32_t* oracle_multiplier_table(uint32_t* n)
{
    *n = MCML_MWC_TABLE_SIZE;
    return mcml_mwc_multipliers;
}
void oracle_neglog_u32_n(const uint32_t* x, float* y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) y[i] = mcml_neglog_u32(x[i], mcml_lg_table);
}
void oracle_sincos_u32_n(const uint32_t* x, float* s, float* c, int64_t n)
{
    for (int64_t i = 0; i < n; i++) mcml_sincos_u32(x[i], mcml_sc_table, &s[i], &c[i]);
}
void oracle_tables(const float** sc, int* n_sc, const float** lg, int* n_lg)
{
    *sc = &mcml_sc_table[0].a; /* pairs of floats, contiguous */
    *n_sc = MCML_SC_N;
    *lg = &mcml_lg_table[0].a;
    *n_lg = MCML_LG_N;
}
void oracle_sincos_turns_n(const float* t, float* s, float* c, int64_t n)
{
    for (int64_t i = 0; i < n; i++) mcml_sincos_turns(t[i], &s[i], &c[i]);
}
void oracle_acos_n(const float* x, float* y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) y[i] = mcml_acos(x[i]);
}
/* Spin n photons travelling along d0 with anisotropy g; writes cos(theta) drawn and the
 * new directions (n x 3). */
void oracle_spin_n(float g, const float* d0, uint64_t seed, float* cost, float* dirs, int64_t n)
{
    oracle_photon p;
    memset(&p, 0, sizeof(p));
    p.rng = mcml_stream(seed, 0, 0, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    for (int64_t i = 0; i < n; i++) {
        p.dx = d0[3 * i]; p.dy = d0[3 * i + 1]; p.dz = d0[3 * i + 2];
        cost[i] = oracle_spin(&p, g);
        dirs[3 * i] = p.dx; dirs[3 * i + 1] = p.dy; dirs[3 * i + 2] = p.dz;
    }
}
/* Reflect one photon (direction d, in layer `layer`) toward new_layer, n times with
 * independent draws; returns the fraction reflected and the transmitted direction of the
 * last transmitted trial in dout. */
double oracle_reflect_fraction(const mcml_derived* d, int layer, int new_layer, const float* dir, uint64_t seed,
                               int64_t n, float* dout)
{
    oracle_photon p;
    memset(&p, 0, sizeof(p));
    p.rng = mcml_stream(seed, 0, 1, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    int64_t nref = 0;
    for (int64_t i = 0; i < n; i++) {
        p.layer = layer;
        p.dx = dir[0]; p.dy = dir[1]; p.dz = dir[2];
        if (oracle_reflect(d, &p, new_layer)) nref++;
        else if (dout) { dout[0] = p.dx; dout[1] = p.dy; dout[2] = p.dz; }
    }
    return (double)nref / (double)n;
}
/* Launch n photons; positions and directions (n x 6). */
void oracle_launch_n(const mcml_derived* d, uint64_t seed, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        oracle_photon p;
        oracle_launch(d, seed, 0, (uint64_t)i, &p);
        float* o = out + 6 * i;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.dx; o[4] = p.dy; o[5] = p.dz;
    }
}
/* Roulette n times from weight w; returns the mean surviving weight (0 for the dead). */
double oracle_roulette_mean(float w, uint64_t seed, int64_t n)
{
    oracle_photon p;
    memset(&p, 0, sizeof(p));
    p.rng = mcml_stream(seed, 0, 2, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    double sum = 0;
    for (int64_t i = 0; i < n; i++) {
        p.w = w;
        p.steps = 0;
        if (oracle_survive(&p, 1)) sum += p.w;
    }
    return sum / (double)n;
}
/* Draw n step lengths in layer `layer` (s = -ln(u) * mutr). */
void oracle_steps_n(const mcml_derived* d, int layer, uint64_t seed, float* out, int64_t n)
{
    oracle_photon p;
    memset(&p, 0, sizeof(p));
    p.rng = mcml_stream(seed, 0, 3, mcml_mwc_multipliers, MCML_MWC_TABLE_SIZE);
    for (int64_t i = 0; i < n; i++) out[i] = mcml_neglog_u32(mcml_mwc_next(&p.rng), mcml_lg_table) * d->layer[layer].mutr;
}
/* Detection test on an explicit exit state (x, y, dx, dy, dz, w); returns the mask. */
uint32_t oracle_detect_state(const mcml_derived* d, const float* st, int only, float* wout)
{
    oracle_photon p;
    memset(&p, 0, sizeof(p));
    p.x = st[0]; p.y = st[1]; p.z = 0.0f; p.dx = st[2]; p.dy = st[3]; p.dz = st[4]; p.w = st[5];
    return oracle_detect(d, &p, only, wout);
}
