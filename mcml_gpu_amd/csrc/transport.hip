/*
 * transport.hip - gfx950 kernels of the photon-transport hot path.
 *
 *   mcml_first_pass  persistent hop-drop-spin kernel (reference: MCd + LaunchPhoton +
 *                    Spin + Reflect + PhotonSurvive + detect, transport.cu:341-471,
 *                    624-938).  Every lane owns one photon at a time; a wavefront claims
 *                    photon tickets in chunks from one global counter and refills idle
 *                    lanes in place, so lanes whose photon died or escaped are reused
 *                    without leaving the kernel.  Tickets span all parameter sets of a
 *                    batch, so one launch runs many sets with uneven budgets.
 *   mcml_replay      re-runs detected photons from their (seed, set, photon) stream for
 *                    one target detector, accumulating per-layer pathlength, step count
 *                    and the absorbance grids (transport.cu:473-622, 941-1000).
 *   mcml_dump        packs the replay outputs into the on-disk record layout
 *                    [weight, PL_1..PL_L, steps] (io.cpp:302-310) with coalesced stores.
 *   mcml_refshape_*  the same photon step in the reference's own launch structure (one
 *                    photon per thread, 50 000-step relaunches), kept as a measured
 *                    comparison for the persistent kernel.
 *
 * Arithmetic: only correctly rounded IEEE binary32 operations, the deterministic
 * functions of mcml_detmath.h and explicit operation order, compiled with
 * -ffp-contract=off, so each photon's trajectory is a function of (seed, set, photon)
 * alone and is bit-identical to a CPU evaluation of the same steps.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mcml_detmath.h"
#include "mcml_device.h"
#include "mwc_multipliers.h"
#include "mcml_tables.h"

#define RPI_F 0.318309886f /* 1/pi as the reference writes it (header.h:41) */

namespace {

/* The step's two lookup tables (azimuth sin/cos, logarithm cells), copied into LDS by every
 * workgroup of a kernel that moves photons before its first step. */
__shared__ mcml_pair g_tab[MCML_SC_N + MCML_LG_N];
#define SC_TAB (g_tab)
#define LG_TAB (g_tab + MCML_SC_N)

__device__ __forceinline__ void stage_tables()
{
    for (int i = threadIdx.x; i < MCML_SC_N; i += blockDim.x) g_tab[i] = mcml_sc_table[i];
    for (int i = threadIdx.x; i < MCML_LG_N; i += blockDim.x) g_tab[MCML_SC_N + i] = mcml_lg_table[i];
    __syncthreads();
}

/* ----------------------------------------------------------------- wave helpers */
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rank_in(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t bcast_first(uint64_t v)
{
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, off, 64);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

/* ----------------------------------------------------------------- photon state */
struct Photon {
    float x, y, z, dx, dy, dz, w;
    int32_t layer;
    int32_t steps;
    int32_t first;
    mcml_mwc rng;
    /* the current layer's parameters, reloaded only when the layer changes */
    float zmin, zmax, mutr, albedo, hg_a2, hg_e, hg_cn, hg_c0; /* hg_cn = -2g: 0 = isotropic */
    /* running pathlength in the current layer; the per-layer totals live in LDS and are
     * exchanged with this register only when the photon changes layer */
    float pl_cur;
    int32_t pl_layer; /* the tissue layer pl_cur belongs to */
};

__device__ __forceinline__ void load_layer(Photon& p, const DevSet* S)
{
    const float4* L = reinterpret_cast<const float4*>(&S->layer[p.layer]);
    const float4 a = L[0], b = L[1], c = L[2];
    p.zmin = a.x;
    p.zmax = a.y;
    p.mutr = a.z;
    p.albedo = a.w;
    p.hg_a2 = b.y;
    p.hg_e = b.z;
    p.hg_cn = b.w;
    p.hg_c0 = c.x;
}

/* Launch photon k of set `set` (LaunchPhoton, transport.cu:624-719). */
__device__ __forceinline__ void launch(Photon& p, const DevSet* S, uint64_t seed, uint32_t set, uint64_t k,
                                       const uint32_t* mult)
{
    p.rng = mcml_stream(seed, set, k, mult, MCML_MWC_TABLE_SIZE);
    if (S->src_oblique) {
        p.x = 0.0f; p.y = 0.0f; p.z = 0.0f;
        p.dx = S->src_sin_a; p.dy = 0.0f; p.dz = S->src_cos_a;
    } else {
        if (S->src_radius == 0.0f) {
            p.x = 0.0f; p.y = 0.0f; p.z = 0.0f;
        } else {
            float u_pos = mcml_uniform(&p.rng);
            float u_az = mcml_uniform(&p.rng);
            float rho = S->src_radius * u_pos;
            float sa, ca;
            mcml_sincos_turns(u_az, &sa, &ca);
            p.x = rho * sa;
            p.y = rho * ca;
            p.z = 0.0f;
        }
        if (S->src_na == 0.0f) {
            p.dx = 0.0f; p.dy = 0.0f; p.dz = 1.0f;
        } else {
            float u_dir = mcml_uniform(&p.rng);
            float u_rot = mcml_uniform(&p.rng);
            float th = S->src_theta_max * u_dir;
            float st, ct, sr, cr;
            mcml_sincos(th, &st, &ct);
            mcml_sincos_turns(u_rot, &sr, &cr);
            p.dz = ct;
            p.dx = st * cr;
            p.dy = st * sr;
        }
    }
    p.layer = 1;
    p.first = 1;
    p.steps = 0;
    p.w = S->start_weight;
    p.pl_cur = 0.0f;
    p.pl_layer = 1;
    load_layer(p, S);
}

/* Correctly rounded 1/x and sqrt(x) from the hardware units.  The reciprocal unit's result
 * (within 1 ulp) refined by one Newton step e = 1 - x y, y + y e (two fmas) is the correctly
 * rounded 1/x: Markstein's theorem leaves only significands of all ones in doubt, and
 * mcml_selftest_math checks every significand at 16 exponents on the device.  The square
 * root comes from the reciprocal square root unit, y ~ 1/sqrt(x), with one Markstein
 * correction: s = x y, h = y / 2, r = x - s^2 (exact by fma), s + r h.  On gfx950 this is
 * the correctly rounded sqrt for every float from 2^-96 up (checked exhaustively; none of
 * the bare v_sqrt_f32, v_rcp_f32 results is correctly rounded everywhere: 15 % and 11 % of
 * the significands differ), and mcml_selftest_math checks every normal float at context
 * creation; below 2^-96, where r underflows, the compiler's IEEE sequence takes over.  So
 * the transport computes exactly 1.0f / x and sqrtf(x), as its CPU restatement does, in 3
 * and 5 VALU instructions (one of them a transcendental each) instead of 6 and 11 for
 * Newton iterations from a bit-pattern seed, or 9 for the square root unit's result and the
 * test of its two neighbours. */
__device__ __forceinline__ float rcp_rn(float x)
{
    const float y = __builtin_amdgcn_rcpf(x);
    const float e = fmaf(-x, y, 1.0f);
    return fmaf(e, y, y);
}

__device__ __forceinline__ float sqrt_rn(float x)
{
    /* below 2^-96 the IEEE sequence, with a negative x taken as 0 */
    if (__builtin_expect(x < 0x1p-96f, 0)) return sqrtf(fmaxf(x, 0.0f));
    const float y = __builtin_amdgcn_rsqf(x);
    const float s = x * y, h = 0.5f * y;
    const float r = fmaf(-s, s, x);
    return fmaf(r, h, s);
}

/* a / b from the correctly rounded reciprocal y = rcp_rn(b): q = a y, then q + (a - b q) y
 * with the residual exact by fma.  By Markstein's theorem (y within half an ulp of 1/b, q
 * within an ulp of a/b) this is the IEEE quotient unless it over- or underflows; the oracle
 * forms the hop's cut distance the same way, so the two agree in every case.  6 instructions
 * (one transcendental) against 11 for the IEEE division sequence. */
__device__ __forceinline__ float div_rn(float a, float b)
{
    const float y = rcp_rn(b);
    const float q = a * y;
    const float r = fmaf(-b, q, a);
    return fmaf(r, y, q);
}

/* sqrt_rn for the spin's arguments, (1 - cos^2) and (1 - cos^2)(1 - dz^2).  Each factor is
 * 1 - c^2 formed by one fma from a float c: 0 at |c| = 1 and otherwise at least 2^-24 in
 * magnitude (the floats next to 1 are 1 - 2^-24 and 1 + 2^-23), so the argument is 0 or at
 * least 2^-48 in magnitude, never in (0, 2^-96) where sqrt_rn takes its IEEE branch.  A
 * non-positive argument (c rounded to or past +-1) gives sin(theta) = 0 by a select. */
__device__ __forceinline__ float sqrt_spin(float x)
{
    const float y = __builtin_amdgcn_rsqf(x);
    const float s = x * y, h = 0.5f * y;
    const float r = fmaf(-s, s, x);
    return x > 0.0f ? fmaf(r, h, s) : 0.0f;
}

/* Henyey-Greenstein deflection + azimuth + renormalisation (Spin, transport.cu:742-786),
 * written with one division and one square root on the common path:
 * P = sin(theta) sqrt(1-dz^2), Q = sin(theta) / sqrt(1-dz^2), and a first-order
 * rsqrt about 1 for the renormalisation (the norm never drifts more than a few ulp, so the
 * second-order term 3d^2/8 is below 1e-13, far under half an ulp of 1). */
template <bool ISO>
__device__ __forceinline__ void spin(Photon& p)
{
    /* the HG uniform u = 2 - v, v = 1.k in [1, 2): D = 1 - g + 2gu = (1 + 3g) - 2gv, and the
     * isotropic cos(theta) = 2u - 1 = 3 - 2v exactly */
    const float v = mcml_one_plus(mcml_mwc_next(&p.rng));
    const bool iso = ISO && p.hg_cn == 0.0f; /* g == 0 (ISO false: no layer of the launch has g = 0) */
    const float D = fmaf(p.hg_cn, v, p.hg_e); /* exactly 1 for g = 0 */
    float sinp, cosp;
    mcml_sincos_u32(mcml_mwc_next(&p.rng), SC_TAB, &sinp, &cosp);
    const float b = fmaf(-p.dz, p.dz, 1.0f);
    /* cos(theta) = fma(X, hg_a2, hg_c0): X = -1/D^2, or for an isotropic layer X = v with
     * hg_a2 = -2, hg_c0 = 3 (3 - 2v, one rounding) */
    const float dz0 = p.dz;
    {
        /* one division serves the HG inversion (1/D = b r) and the rotation (1/b = D r) */
        const float r = rcp_rn(D * b);
        const float br = b * r; /* 1 / D */
        const float cost = fmaf(iso ? v : -(br * br), p.hg_a2, p.hg_c0);
        const float ab = fmaf(-cost, cost, 1.0f) * b;
        const float P = sqrt_spin(ab);
        const float Q = P * (D * r);
        /* ux' = ux (cos + uz Q cos psi) - uy Q sin psi, uy' = uy (cos + uz Q cos psi) +
         * ux Q sin psi: the bracket is shared */
        const float Qc = Q * cosp, Qs = Q * sinp;
        const float k = fmaf(p.dz, Qc, cost);
        const float cxp = fmaf(p.dx, k, -(Qs * p.dy));
        const float cyp = fmaf(p.dy, k, Qs * p.dx);
        p.dz = fmaf(-P, cosp, p.dz * cost);
        p.dx = cxp;
        p.dy = cyp;
    }
    if (b == 0.0f) { /* travelling along z: the lines above divided by 0, replace their result */
        const float t = 1.0f / D;
        const float cost = fmaf(iso ? v : -(t * t), p.hg_a2, p.hg_c0);
        const float sint = sqrt_spin(fmaf(-cost, cost, 1.0f));
        p.dx = sint * cosp;
        p.dy = sint * sinp;
        p.dz = copysignf(cost, dz0 * cost);
    }
    const float d = fmaf(p.dx, p.dx, fmaf(p.dy, p.dy, fmaf(p.dz, p.dz, -1.0f)));
    const float r = fmaf(d, -0.5f, 1.0f);
    p.dx = p.dx * r;
    p.dy = p.dy * r;
    p.dz = p.dz * r;
}

/* Fresnel boundary (Reflect, transport.cu:790-853): true = reflected.  The interface
 * constants come precomputed from the layer table (same float expressions). */
__device__ __forceinline__ bool reflect(Photon& p, const DevSet* S, int new_layer)
{
    const DevLayer& Lc = S->layer[p.layer];
    const bool up = new_layer < p.layer;
    const float4 B = reinterpret_cast<const float4*>(&Lc)[up ? 3 : 4];
    const float n1 = Lc.n;
    const float n2 = B.x;
    const float ci = fabsf(p.dz);
    float r;
    if (n1 == n2) {
        p.layer = new_layer;
        return false;
    }
    if (n1 > n2 && n2 * n2 < n1 * n1 * (1.0f - ci * ci)) {
        p.dz = -p.dz;
        return true;
    }
    if (ci == 1.0f) {
        r = B.w;
        if (mcml_uniform(&p.rng) <= r * r) {
            p.dz = -p.dz;
            return true;
        }
        p.layer = new_layer;
        return false;
    }
    float e = B.z * (1.0f - ci * ci);
    r = 2.0f * sqrt_rn((1.0f - ci * ci) * (1.0f - e) * e * ci * ci);
    e = e + (ci * ci) * (1.0f - 2.0f * e);
    r = e * ((1.0f - e - r) / ((1.0f - e + r) * (e + r)));
    if (mcml_uniform(&p.rng) <= r) {
        p.dz = -p.dz;
        return true;
    }
    r = B.y;
    e = r * r * (1.0f - ci * ci);
    p.dx = p.dx * r;
    p.dy = p.dy * r;
    p.dz = copysignf(sqrt_rn(1.0f - e), p.dz);
    p.layer = new_layer;
    return false;
}

/* Detectors accepting the escaping photon (detect / detect_replay, transport.cu:872-1000),
 * as a bit mask (bit i = SDS i).  Fiber mode: tilted elliptical footprint + NA cone; ring
 * mode: annulus + NA cone.  `target` = 0 tests every detector, otherwise that one only.
 * The set's reject box skips the per-detector tests for photons no detector can accept. */
__device__ __forceinline__ uint32_t detect_mask(const Photon& p, const DevSet* S, int target)
{
    const int i0 = target ? target : 1, i1 = target ? target : S->num_sds;
    uint32_t m = 0;
    if (S->det_oblique) {
        if (p.y * p.y > S->y2max) return 0u;
        if (S->xlo <= S->xhi && (p.x < S->xlo || p.x > S->xhi)) return 0u;
        for (int i = i0; i <= i1; i++) {
            const DevDet D = S->det[i];
            const float t = (p.x - D.pos) * D.cos_a;
            if (t * t + p.y * p.y <= D.r2) {
                const float uz = p.dx * D.sin_a + fabsf(p.dz * D.cos_a);
                if (fabsf(uz) >= D.cos_crit) m |= 1u << i;
            }
        }
        return m;
    }
    const float adz = fabsf(p.dz);
    if (adz < S->cosmin) return 0u;
    const float dist = sqrt_rn(p.x * p.x + p.y * p.y);
    if (dist < S->dlo || dist > S->dhi) return 0u;
    for (int i = i0; i <= i1; i++) {
        const DevDet D = S->det[i];
        if (adz >= D.cos_crit && dist >= D.pos - D.radius && dist <= D.pos + D.radius) m |= 1u << i;
    }
    return m;
}

/* Weight credited to detector i (ring mode applies the arc fraction acos(t)/pi). */
__device__ __forceinline__ float credited(const Photon& p, const DevSet* S, int i)
{
    if (S->det_oblique) return p.w;
    const DevDet D = S->det[i];
    const float dist = sqrt_rn(p.x * p.x + p.y * p.y);
    float t = (dist * dist + D.pos * D.pos - D.radius * D.radius) / (2.0f * dist * D.pos);
    if (t > 1.0f) t = 1.0f;
    return p.w * mcml_acos(t) * RPI_F;
}

enum StepOutcome { ST_ALIVE = 0, ST_DEAD = 1, ST_DETECTED = 2 };

/* Hop (transport.cu:372-395): sample the step, stop it at the layer boundary, move, count
 * the step, clamp z.  Returns true when the hop ended on a boundary: cut at z_min (moving up)
 * or at z_max (moving down), the layer beyond being beyond_layer(p).  A hop cut at z_min
 * lands on z_min and cannot pass z_max, so at most one cut happens; the reference's layer
 * index would move by -1 and +1 if both did, which "exactly one" restates. */
template <bool REPLAY>
__device__ __forceinline__ bool hop(Photon& p, float& s)
{
    s = mcml_neglog_u32(mcml_mwc_next(&p.rng), LG_TAB) * p.mutr;
    /* the uncut landing passes at most one of the two (zmin <= zmax): one branch for both cuts */
    const float zn = fmaf(s, p.dz, p.z);
    const bool lo = zn < p.zmin;
    bool hi = zn > p.zmax;
    if (lo || hi) {
        s = div_rn((lo ? p.zmin : p.zmax) - p.z, p.dz);
        /* a hop cut at zmin is tested against zmax again (transport.cu:387), which only a
         * layer thinner than the rounding of the cut can fail */
        if (lo && fmaf(s, p.dz, p.z) > p.zmax) {
            hi = true;
            s = div_rn(p.zmax - p.z, p.dz);
        }
    }
    p.x = fmaf(p.dx, s, p.x);
    p.y = fmaf(p.dy, s, p.y);
    /* transport.cu:394-395 as one median: an uncut step already lies in [zmin, zmax] */
    p.z = __builtin_amdgcn_fmed3f(fmaf(p.dz, s, p.z), p.zmin, p.zmax);
    p.steps++;
    if (REPLAY) p.pl_cur += s; /* transport.cu:529: pathlength of the layer the hop started in */
    return lo != hi;
}

/* The layer beyond the boundary a pending hop stopped at.  A hop is cut at z_min only when
 * it moves up (dz < 0) and at z_max only when it moves down (z stays in [z_min, z_max]
 * between hops), and a waiting lane does not move, so the sign of dz tells which. */
__device__ __forceinline__ int beyond_layer(const Photon& p) { return p.layer + (p.dz > 0.0f ? 1 : -1); }

/* PhotonSurvive (transport.cu:855-870).  The reference applies the step cap in the first
 * pass only; a replayed photon was detected within the cap, so applying it in replay too
 * changes nothing for detected photons and bounds every replay loop. */
__device__ __forceinline__ int survive(Photon& p)
{
    if (p.steps >= MCML_MAX_SCATTER) return ST_DEAD;
    if (p.w > (float)MCML_WEIGHTI) return ST_ALIVE;
    if (p.w == 0.0f) return ST_DEAD;
    if (mcml_uniform(&p.rng) < MCML_CHANCE) {
        p.w = rintf(p.w / MCML_CHANCE);
        return ST_ALIVE;
    }
    return ST_DEAD;
}

/* survive() for a photon at the step cap or at or below WEIGHTI: true = it ends. */
__device__ __forceinline__ bool survive_ends(Photon& p)
{
    if (p.steps >= MCML_MAX_SCATTER || p.w == 0.0f) return true;
    if (mcml_uniform(&p.rng) < MCML_CHANCE) {
        p.w = rintf(p.w / MCML_CHANCE);
        return false;
    }
    return true;
}

/* Absorbance bookkeeping of one replayed photon (transport.cu:580-600): the first drop goes
 * to A0_z, later ones to A_rz, merged while consecutive drops stay in one cell.
 *
 * Each wavefront owns an LDS tile of the hot corner of one (set, SDS) group's grids (z bins
 * < TILE_Z, and r bins < TILE_R for A_rz), accumulated with u64 LDS atomics and flushed to
 * the global grids with u64 atomics when the wavefront moves on to another group and at the
 * end.  A corner drop of a lane whose photon belongs to another group than the tile's (a
 * wavefront moving between groups holds photons of both for a while) goes to the wavefront's
 * staging copy of the corner in global memory (ReplayArgs.stage, one of up to 64 copies, so
 * few wavefronts share a word), added into the grids by mcml_stage_reduce after the
 * replay; drops outside the corner go straight to the global grids.  u64 sums are exact and
 * associative, so the grids do not depend on where a drop was first added (the reference's
 * own accumulation type, transport.cu:335-339; FP32 tiles would round).
 *
 * The tile is kept in TILE_COPIES copies, lane i adding to copy i % TILE_COPIES: the drops
 * crowd into a few cells near the source (on the HW8_A example 7 % of the runs land in cell
 * (0, 0) and every first drop in A0_z[0]), and lanes adding to one LDS word in the same
 * instruction are served one after another.  With a 32 x 32 corner (all of HW8_A's runs,
 * against 94 % for 16 x 16; the others go to contended global atomics) and two copies the
 * HW8_A -R -A replay took 6.3 ms instead of 10.0; four copies cost more in occupancy than
 * they saved, u32 cells were slower.  Half of those 6.3 ms were the other-group drops, as
 * global atomics on the grids' hottest cells that every wavefront hits: with the staging
 * copies the replay takes 3.5 ms (without any grid atomics, 2.4 ms). */
#define TILE_Z 32
#define TILE_R 32
#define TILE_COPIES 2
#define TILE_WORDS (TILE_Z * TILE_R + TILE_Z)

struct GridLane {
    unsigned long long* a_rz;   /* this lane's (set, SDS) slabs */
    unsigned long long* a0_z;
    unsigned long long* tile;   /* the wavefront's LDS tile (NULL without grids) */
    unsigned long long* stage;  /* this lane's group in the wavefront's staging copy (or NULL) */
    uint64_t rl_key;            /* run in progress: z bin << 32 | r bin */
    uint64_t rl_w;
    uint32_t grp;               /* this lane's (set, SDS) group: set * S + sds - 1 */
    uint32_t tile_grp;          /* the group the tile holds (wave-uniform; ~0 = none) */
};

__device__ __forceinline__ void flush_run(GridLane& gl, int nr)
{
    if (gl.rl_w == 0) return;
    const uint32_t zb = (uint32_t)(gl.rl_key >> 32), rb = (uint32_t)gl.rl_key;
    const bool corner = zb < TILE_Z && rb < TILE_R;
    if (corner && gl.grp == gl.tile_grp)
        atomicAdd(&gl.tile[zb * TILE_R + rb], (unsigned long long)gl.rl_w);
    else if (corner && gl.stage != nullptr)
        atomicAdd(&gl.stage[zb * TILE_R + rb], (unsigned long long)gl.rl_w);
    else
        atomicAdd(&gl.a_rz[(uint64_t)zb * (uint64_t)nr + rb], (unsigned long long)gl.rl_w);
    gl.rl_w = 0;
}

/* Interaction (transport.cu:430-437 / 573-603): drop weight and scatter when the hop did
 * not end on a boundary, then the survival test.  The reference tests s > 0, with s set to
 * 0 on a boundary; an uncut hop's s = -ln(u) mutr is at least 5.96e-8 mutr (u < 1), positive
 * for any layer with mua + mus below 4e37 /cm, so "not on a boundary" is the same test. */
template <bool REPLAY>
__device__ __forceinline__ int interact(Photon& p, GridLane& gl, const ReplayArgs* ra, uint32_t& n_int)
{
    n_int++;
    const float w_temp = rintf(p.albedo * p.w);
    p.w -= w_temp;
    if (REPLAY && gl.a_rz != nullptr) {
        float qz = p.z * ra->rdz;
        qz = fminf(fmaxf(qz, 0.0f), (float)(ra->nz - 1));
        const int zb = (int)qz;
        if (p.first) {
            if (zb < TILE_Z && gl.grp == gl.tile_grp)
                atomicAdd(&gl.tile[TILE_Z * TILE_R + zb], (unsigned long long)w_temp);
            else if (zb < TILE_Z && gl.stage != nullptr)
                atomicAdd(&gl.stage[TILE_Z * TILE_R + zb], (unsigned long long)w_temp);
            else
                atomicAdd(&gl.a0_z[zb], (unsigned long long)w_temp);
        } else {
            float qr = sqrt_rn(p.x * p.x + p.y * p.y) * ra->rdr;
            qr = fminf(fmaxf(qr, 0.0f), (float)(ra->nr - 1));
            const uint64_t key = ((uint64_t)(uint32_t)zb << 32) | (uint32_t)(int)qr;
            if (key == gl.rl_key) {
                gl.rl_w += (uint64_t)w_temp;
            } else {
                flush_run(gl, ra->nr);
                gl.rl_key = key;
                gl.rl_w = (uint64_t)w_temp;
            }
        }
    }
    p.first = 0;
    spin<true>(p);
    return survive(p);
}

/* Boundary (transport.cu:397-428): Fresnel, then escape through the top (detection, or
 * the detector face's reflection back into the tissue), loss through the bottom, or entry
 * into the next tissue layer; then the survival test.  `target` = 0 tests every detector
 * (first pass), otherwise that detector only (replay).  On ST_DETECTED *mask holds the
 * accepting detectors (bit i = SDS i). */
template <bool PL>
__device__ __forceinline__ int cross(Photon& p, int nl, const DevSet* S, int target, uint32_t& mask, float* pl_lane,
                                     int pl_stride)
{
    const int temp_layer = p.layer;
    if (!reflect(p, S, nl)) {
        if (nl == 0) {
            const uint32_t m = detect_mask(p, S, target);
            if (m) {
                mask = m;
                return ST_DETECTED;
            }
            if (mcml_uniform(&p.rng) > S->det_reflectance) return ST_DEAD;
            p.dz = -p.dz; /* the detector face reflects the photon, refracted direction kept */
            p.z = -p.z;
            p.layer = temp_layer;
        } else if (nl > S->num_layers) {
            return ST_DEAD;
        } else {
            if (PL) { /* the same sequence of float additions per layer as pl[layer-1] += s */
                pl_lane[(p.pl_layer - 1) * pl_stride] = p.pl_cur;
                p.pl_cur = pl_lane[(nl - 1) * pl_stride];
                p.pl_layer = nl;
            }
            load_layer(p, S);
        }
    }
    return survive(p);
}

__device__ __forceinline__ uint32_t set_of_ticket(const uint64_t* set_begin, int nsets, uint64_t t)
{
    int lo = 0, hi = nsets;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (set_begin[mid] <= t) lo = mid;
        else hi = mid;
    }
    return (uint32_t)lo;
}

/* Wavefront work queue: lanes named in `need` take consecutive ids from the wavefront's
 * private range [q_next, q_end), which is refilled `chunk` ids at a time from the global
 * counter.  Returns this lane's id, or ~0 when it gets none.
 *
 * Pause: once `*n_rec` (records appended so far) reaches `stop_at`, the wavefront claims no
 * further chunk and runs dry as if the counter were exhausted; it still finishes the
 * photons it holds and the rest of its private range.  The host then grows the record
 * buffer and launches again; the counter carries on where the paused launch left it, so
 * every ticket is run exactly once.  stop_at = ~0 disables the check. */
struct WaveQueue {
    uint64_t next, end;
    bool dry;       /* no further chunk will be claimed (exhausted or paused) */
    bool exhausted; /* the counter passed the last ticket */
};

__device__ __forceinline__ uint64_t take_ids(WaveQueue& q, uint64_t need, bool mine, unsigned long long* counter,
                                             uint64_t chunk, uint64_t total, uint64_t& claimed,
                                             const unsigned long long* n_rec, uint64_t stop_at)
{
    claimed = ~0ull;
    const int n = __popcll(need);
    const uint32_t rank = rank_in(need);
    const uint64_t avail = q.end - q.next;
    uint64_t id = ~0ull;
    if (mine && rank < avail) id = q.next + rank;
    if ((uint64_t)n > avail) {
        const uint32_t first = (uint32_t)(__ffsll((long long)__ballot(1)) - 1);
        bool stop = false;
        if (stop_at != ~0ull) { /* wave-uniform */
            unsigned long long r = 0;
            if (lane_id() == first) r = __hip_atomic_load(n_rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            stop = bcast_first(r) >= stop_at;
        }
        unsigned long long base = total;
        if (!stop) {
            if (lane_id() == first) base = atomicAdd(counter, chunk);
            base = bcast_first(base);
        }
        uint64_t cend = base + chunk;
        if (cend > total) cend = total;
        if (base >= total) {
            q.dry = true;
            q.exhausted = !stop;
            cend = base;
        } else {
            claimed = base;
        }
        if (mine && rank >= avail) {
            const uint64_t t = base + (rank - avail);
            if (t < cend) id = t;
        }
        const uint64_t used = base + ((uint64_t)n - avail);
        q.next = used < cend ? used : cend;
        q.end = cend;
    } else {
        q.next += (uint64_t)n;
    }
    return id;
}

/* Should the wavefront resolve its pending boundary hits now?  Yes once `bnd_min` lanes
 * wait, or when every live lane waits (no lane could make progress otherwise). */
__device__ __forceinline__ bool resolve_now(bool pending, bool alive, int bnd_min)
{
    const uint64_t pend = __ballot(pending);
    if (pend == 0) return false;
    const int np = __popcll(pend);
    return np >= bnd_min || pend == __ballot(alive);
}

/* Photon steps a wavefront takes between its refill and boundary decisions.  Each decision
 * costs a few ballots and selects; making them every sixth step instead of every step
 * took the cfg1 first pass from 215 to 197 ms (3: 200, 4: 199); with the step as flat
 * conditions, twelve gain another 1 % over six (8: 0.5 %). */
#define STEPS_PER_CHECK 12

/* Set owning ticket t, using the wavefront's cached set range [clo, chi) first. */
__device__ __forceinline__ uint32_t set_for(const uint64_t* set_begin, int nsets, uint64_t t, uint32_t cset,
                                            uint64_t clo, uint64_t chi)
{
    if (nsets == 1) return 0u;
    if (t >= clo && t < chi) return cset;
    return set_of_ticket(set_begin, nsets, t);
}

/* Add a detected weight to its group's accumulator slot (see MCML_ACC_SLOTS): the float's
 * exact value as a fixed-point integer with 64 fraction bits, in two words with the carry of
 * the low word propagated by the caller that produced it, so the total is exact in any order.
 * Weights are below 2^32 W-units, so 2^32 records cannot overflow 128 bits; values below
 * 2^-41 W-units lose the bits beyond 2^-64. */
__device__ __forceinline__ void acc_weight(unsigned long long* e, float w)
{
    const uint32_t b = __float_as_uint(w);
    const uint32_t ex = b >> 23;
    atomicAdd(&e[ACC_COUNT], 1ull);
    if ((b >> 31) != 0u || ex == 255u) { /* NaN, infinity or negative */
        atomicAdd(&e[ACC_BAD], 1ull);
        return;
    }
    if (ex == 0u) return; /* subnormal: below 2^-126, nothing to add at 2^-64 resolution */
    const uint64_t m = (uint64_t)((b & 0x7FFFFFu) | 0x800000u);
    const int sh = (int)ex - 86; /* value * 2^64 = m * 2^(ex - 150 + 64) */
    uint64_t lo = 0, hi = 0;
    if (sh >= 64) hi = m << (sh - 64);
    else if (sh > 40) { lo = m << sh; hi = m >> (64 - sh); }
    else if (sh >= 0) lo = m << sh;
    else if (sh > -24) lo = m >> (-sh);
    if (lo != 0) {
        const unsigned long long old = atomicAdd(&e[ACC_LO], (unsigned long long)lo);
        if (old + lo < old) hi++; /* the low word wrapped: carry into the high word */
    }
    if (hi != 0) atomicAdd(&e[ACC_HI], (unsigned long long)hi);
}

/* First pass over all photon tickets (MCd + LaunchPhoton_Global + calculate_reflectance,
 * transport.cu:341-471, 722-739, 299-311).  PL = also accumulate per-layer pathlength and
 * step count, so that each detection record is already the replay's output row.  SUMS =
 * no records: each detection is added to its (set, SDS) group's accumulators instead. */
/* Book one drop of a lane replaying its detected photon for group `grp` (the replay's
 * bookkeeping of transport.cu:580-600, as in interact<true>): the first drop to A0_z, later
 * ones to A_rz, merged while consecutive drops stay in one cell.  The run is kept in two
 * registers: its cell as z bin << 16 | r bin (grids of fewer than 65536 bins a side; ~0 =
 * none) and its u32 weight, flushed before it would wrap.  Cells are addressed by 32-bit
 * byte offsets from the grids' base (grids below 4 GiB), one u64 atomic per run.  A drop in
 * the grids' corner goes to the wavefront's staging copy instead when there is one: the
 * corner's few hottest cells take the drops of every replay, and atomics serialised on them
 * (as in mcml_replay). */
__device__ __forceinline__ unsigned long long* drop_stage(const FirstPassArgs& a, uint32_t grp)
{
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    return a.stage + ((uint64_t)(wave % (uint32_t)a.stage_k) * (uint64_t)a.stage_groups + grp) * TILE_WORDS;
}

__device__ __forceinline__ void flush_drop_run(const FirstPassArgs& a, uint32_t grp, uint32_t rl_key, uint32_t rl_w)
{
    if (rl_w == 0) return;
    const uint32_t zb = rl_key >> 16, rb = rl_key & 0xFFFFu;
    if (a.stage != nullptr && zb < TILE_Z && rb < TILE_R) {
        atomicAdd(drop_stage(a, grp) + zb * TILE_R + rb, (unsigned long long)rl_w);
        return;
    }
    const uint32_t cell = (grp * (uint32_t)a.nz + zb) * (uint32_t)a.nr + rb;
    atomicAdd(reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(a.A_rz) + cell * 8u),
              (unsigned long long)rl_w);
}

__device__ __forceinline__ void book_drop(const FirstPassArgs& a, const Photon& p, uint32_t grp, float w_temp,
                                          uint32_t& rl_key, uint32_t& rl_w)
{
    float qz = p.z * a.rdz;
    qz = fminf(fmaxf(qz, 0.0f), (float)(a.nz - 1));
    const uint32_t zb = (uint32_t)(int)qz;
    const uint32_t wd = (uint32_t)w_temp; /* a drop is at most the weight, below 2^32 */
    if (p.first) {
        if (a.stage != nullptr && zb < TILE_Z) {
            atomicAdd(drop_stage(a, grp) + TILE_Z * TILE_R + zb, (unsigned long long)wd);
            return;
        }
        const uint32_t cell = grp * (uint32_t)a.nz + zb;
        atomicAdd(reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(a.A0_z) + cell * 8u),
                  (unsigned long long)wd);
        return;
    }
    float qr = sqrt_rn(p.x * p.x + p.y * p.y) * a.rdr;
    qr = fminf(fmaxf(qr, 0.0f), (float)(a.nr - 1));
    const uint32_t k = (zb << 16) | (uint32_t)(int)qr;
    if (k == rl_key && rl_w + wd >= rl_w) {
        rl_w += wd;
        return;
    }
    flush_drop_run(a, grp, rl_key, rl_w);
    rl_key = k;
    rl_w = wd;
}

/* LDS_SETS: the (small) set table is staged in LDS and read there with LDS instructions;
 * otherwise it is read from global memory (a large batch).  Both are compiled into each
 * kernel and chosen by a.lds_sets, so that no access goes through flat addressing. */
template <bool PL, bool SUMS, bool STATS, bool GRIDS, bool LDS_SETS, bool ISO>
__device__ __forceinline__ void first_pass_body(const FirstPassArgs& a, unsigned char* lds_raw)
{
    const DevSet* sets = a.sets;
    size_t lds_off = 0;
    if (LDS_SETS) {
        const uint4* src = reinterpret_cast<const uint4*>(a.sets);
        uint4* dst = reinterpret_cast<uint4*>(lds_raw);
        const int n16 = (int)(sizeof(DevSet) / 16) * a.nsets;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
        sets = reinterpret_cast<const DevSet*>(lds_raw);
        lds_off = sizeof(DevSet) * (size_t)a.nsets;
    }
    float* pl_lane = reinterpret_cast<float*>(lds_raw + lds_off) + threadIdx.x;
    const int pl_stride = blockDim.x;
    const int L = a.num_layers;
    Photon p;
    bool alive = false, pending = false;
    uint32_t set = 0;
    uint64_t key = 0; /* set << 45 | photon index */
    WaveQueue q = {0, 0, false, false};
    uint32_t cset = 0;
    uint64_t clo = 0, chi = 0;
    /* per-lane work counters (statistics only; the results do not depend on them).  A
     * lane's steps run one after another, about a microsecond each at full occupancy, so
     * its totals stay below 2^32 unless one launch runs for more than an hour.  A hop
     * either ends on a boundary or interacts (its step is never 0), so interactions are
     * counted as steps - boundary hops. */
    uint32_t c_photons = 0, c_steps = 0, c_bnd = 0;
    const int refill_min = a.refill_min < 1 ? 1 : a.refill_min;
    const int bnd_min = a.bnd_min < 1 ? 1 : a.bnd_min;
    uint32_t s_step = 0, s_pend = 0, s_refill = 0, s_dry = 0; /* lane-blocks by state (STATS) */
    /* GRIDS: detectors still to replay the lane's detected photon for (bit i = SDS i; the
     * lowest is the replay in progress, none = the lane runs its first-pass photon), and the
     * drop run of the replay */
    uint32_t rp_mask = 0;
    uint32_t rl_key = ~0u, rl_w = 0;
    if (STATS && lane_id() == 0) atomicMax(&a.stats[STAT_T_START], ~(unsigned long long)wall_clock64());

    for (;;) {
        const uint64_t dead = __ballot(!alive);
        const int ndead = __popcll(dead);
        if (!q.dry && (ndead >= refill_min || ndead == 64)) {
            uint64_t claimed;
            const uint64_t t = take_ids(q, dead, !alive, a.queue, (uint64_t)a.chunk, a.total, claimed, a.n_records,
                                        a.stop_at);
            if (STATS && q.dry && lane_id() == 0) /* this wavefront found the queue dry */
                atomicMax(&a.stats[STAT_T_DRY], ~(unsigned long long)wall_clock64());
            if (claimed != ~0ull && a.nsets > 1) { /* wave-uniform: cache the new chunk's set */
                cset = set_of_ticket(a.set_begin, a.nsets, claimed);
                clo = a.set_begin[cset];
                chi = a.set_begin[cset + 1];
            }
            if (GRIDS && t != ~0ull && a.resume && t >= a.resume_photons) { /* a replay job */
                const uint64_t rk = a.replay_keys[t - a.resume_photons];
                set = (uint32_t)(rk >> MCML_KEY_SET_SHIFT);
                key = rk & ~(31ull << MCML_KEY_SDS_SHIFT);
                launch(p, sets + set, a.seed, set, rk & MCML_KEY_K_MASK, a.mult);
                rp_mask = 1u << (uint32_t)((rk >> MCML_KEY_SDS_SHIFT) & 31u);
                rl_key = ~0u;
                rl_w = 0;
                alive = true;
            } else if (t != ~0ull && a.resume) { /* continue a spilled photon */
                const float* col = a.resume_buf + t;
                const uint64_t cs = a.resume_cap;
                p.x = col[SP_X * cs];
                p.y = col[SP_Y * cs];
                p.z = col[SP_Z * cs];
                p.dx = col[SP_DX * cs];
                p.dy = col[SP_DY * cs];
                p.dz = col[SP_DZ * cs];
                p.w = col[SP_W * cs];
                p.layer = __float_as_int(col[SP_LAYER * cs]);
                p.steps = __float_as_int(col[SP_STEPS * cs]);
                p.first = __float_as_int(col[SP_FIRST * cs]);
                p.rng.s = ((uint64_t)__float_as_uint(col[SP_RNG_HI * cs]) << 32) | __float_as_uint(col[SP_RNG_LO * cs]);
                p.rng.a = __float_as_uint(col[SP_RNG_A * cs]);
                p.pl_cur = col[SP_PL_CUR * cs];
                p.pl_layer = __float_as_int(col[SP_PL_LAYER * cs]);
                key = ((uint64_t)__float_as_uint(col[SP_KEY_HI * cs]) << 32) | __float_as_uint(col[SP_KEY_LO * cs]);
                const uint32_t pb = __float_as_uint(col[SP_PENDING * cs]);
                pending = (pb & 1u) != 0u;
                if (GRIDS) { /* a replay in progress continues where it was spilled */
                    rp_mask = __float_as_uint(col[SP_RP_MASK * cs]);
                    rl_key = __float_as_uint(col[SP_RL_KEY * cs]);
                    rl_w = __float_as_uint(col[SP_RL_W * cs]);
                }
                if (PL)
                    for (int l = 0; l < L; l++) pl_lane[l * pl_stride] = col[(SP_PL0 + l) * cs];
                set = (uint32_t)(key >> MCML_KEY_SET_SHIFT);
                load_layer(p, sets + set);
                alive = true;
            } else if (t != ~0ull) {
                set = set_for(a.set_begin, a.nsets, t, cset, clo, chi);
                const uint64_t k = t - a.set_begin[set];
                key = ((uint64_t)set << MCML_KEY_SET_SHIFT) | k;
                launch(p, sets + set, a.seed, set, k, a.mult);
                if (PL)
                    for (int l = 0; l < L; l++) pl_lane[l * pl_stride] = 0.0f;
                alive = true;
                c_photons++;
            }
        }
        if (__ballot(alive) == 0) break;
        if (a.spill != nullptr && q.exhausted) { /* wave-uniform */
            const uint64_t live = __ballot(alive);
            if (__popcll(live) <= a.spill_below) { /* spill the remaining photons and stop */
                unsigned long long base = 0;
                if (lane_id() == (uint32_t)(__ffsll((long long)__ballot(1)) - 1))
                    base = atomicAdd(a.n_spill, (unsigned long long)__popcll(live));
                base = bcast_first(base);
                if (alive) {
                    float* col = a.spill + base + rank_in(live);
                    const uint64_t cs = a.spill_cap;
                    col[SP_X * cs] = p.x;
                    col[SP_Y * cs] = p.y;
                    col[SP_Z * cs] = p.z;
                    col[SP_DX * cs] = p.dx;
                    col[SP_DY * cs] = p.dy;
                    col[SP_DZ * cs] = p.dz;
                    col[SP_W * cs] = p.w;
                    col[SP_LAYER * cs] = __int_as_float(p.layer);
                    col[SP_STEPS * cs] = __int_as_float(p.steps);
                    col[SP_FIRST * cs] = __int_as_float(p.first);
                    col[SP_RNG_LO * cs] = __uint_as_float((uint32_t)p.rng.s);
                    col[SP_RNG_HI * cs] = __uint_as_float((uint32_t)(p.rng.s >> 32));
                    col[SP_RNG_A * cs] = __uint_as_float(p.rng.a);
                    col[SP_PL_CUR * cs] = p.pl_cur;
                    col[SP_PL_LAYER * cs] = __int_as_float(p.pl_layer);
                    col[SP_KEY_LO * cs] = __uint_as_float((uint32_t)key);
                    col[SP_KEY_HI * cs] = __uint_as_float((uint32_t)(key >> 32));
                    col[SP_PENDING * cs] = __uint_as_float(pending ? 1u : 0u);
                    /* written by every variant (no replay: mask 0), as a drain launch of
                     * mcml_first_pass_grids may continue what the main launch spilled */
                    col[SP_RP_MASK * cs] = __uint_as_float(rp_mask);
                    col[SP_RL_KEY * cs] = __uint_as_float(rl_key);
                    col[SP_RL_W * cs] = __uint_as_float(rl_w);
                    if (PL)
                        for (int l = 0; l < L; l++) col[(SP_PL0 + l) * cs] = pl_lane[l * pl_stride];
                    alive = false; /* its steps count where it ends */
                }
                break;
            }
        }
        if (STATS) {
            s_step += (alive && !pending) ? 1u : 0u;
            s_pend += (alive && pending) ? 1u : 0u;
            s_refill += (!alive && !q.dry) ? 1u : 0u;
            s_dry += (!alive && q.dry) ? 1u : 0u;
        }
        bool ended = false; /* weight or step cap, a lane mask */
        uint32_t mask = 0;
        /* several steps between wavefront-level decisions: a lane that reaches a boundary
         * or ends its photon sits out the rest of the block */
#pragma unroll
        for (int kk = 0; kk < STEPS_PER_CHECK; kk++) {
            /* interact<false> written as flat conditions: one level of divergence around
             * the spin instead of three, so the compiler keeps the direction in place
             * rather than copying it in and out of each level */
            const bool act = alive && !pending && !ended;
            float s = 0.0f;
            bool bnd = false;
            if (act) bnd = hop<PL>(p, s);
            if (act && bnd) pending = true; /* the layer beyond follows from dz (beyond_layer) */
            if (act && !bnd) { /* interact<false>: s > 0 on every uncut hop */
                const float w_temp = rintf(p.albedo * p.w);
                p.w -= w_temp;
                if (GRIDS && rp_mask != 0) {
                    book_drop(a, p, set * (uint32_t)sets[set].num_sds + (uint32_t)(__ffs(rp_mask) - 2), w_temp,
                              rl_key, rl_w);
                    p.first = 0;
                }
                spin<ISO>(p);
                /* survive() only where its first tests can fail: one branch on the common path */
                if (p.steps >= MCML_MAX_SCATTER || !(p.w > (float)MCML_WEIGHTI)) ended = survive_ends(p);
            }
        }
        int st = ended ? ST_DEAD : ST_ALIVE;
        if (resolve_now(pending, alive, bnd_min) && pending) {
            pending = false;
            if (!GRIDS || rp_mask == 0) c_bnd++;
            st = cross<PL>(p, beyond_layer(p), sets + set, GRIDS && rp_mask != 0 ? __ffs(rp_mask) - 1 : 0, mask,
                           pl_lane, pl_stride);
        }
        if (GRIDS && rp_mask != 0 && st != ST_ALIVE) { /* a replay ended: flush its run */
            const int sds = __ffs(rp_mask) - 1;
            flush_drop_run(a, set * (uint32_t)sets[set].num_sds + (uint32_t)(sds - 1), rl_key, rl_w);
            if (st != ST_DETECTED) atomicAdd(a.replay_missed, 1ull);
            rp_mask &= rp_mask - 1u; /* done with this detector */
            alive = false;
        } else if (st != ST_ALIVE) {
            if (GRIDS && st == ST_DETECTED) rp_mask = mask;
            alive = false;
            if (PL) pl_lane[(p.pl_layer - 1) * pl_stride] = p.pl_cur;
            c_steps += (uint32_t)p.steps;
            if (SUMS && st == ST_DETECTED) {
                const DevSet* S = sets + set;
                const int ns = S->num_sds;
                unsigned long long* slot = a.accum + (uint64_t)(blockIdx.x & (MCML_ACC_SLOTS - 1)) * MCML_ACC_WORDS;
                for (int i = 1; i <= ns; i++) {
                    if (!(mask & (1u << i))) continue;
                    const uint64_t grp = (uint64_t)set * (uint64_t)ns + (uint64_t)(i - 1);
                    acc_weight(slot + grp * (MCML_ACC_SLOTS * MCML_ACC_WORDS), credited(p, S, i));
                }
            } else if (st == ST_DETECTED) {
                const unsigned long long base = atomicAdd(a.n_records, (unsigned long long)__popc(mask));
                unsigned long long idx = base;
                const DevSet* S = sets + set;
                const int ns = S->num_sds;
                for (int i = 1; i <= ns; i++) {
                    if (!(mask & (1u << i))) continue;
                    if (idx < a.cap) {
                        a.rec_key[idx] = key | ((uint64_t)i << MCML_KEY_SDS_SHIFT);
                        float* row = a.rec_row + idx * (uint64_t)a.row_stride;
                        row[0] = credited(p, S, i);
                        if (PL) {
                            for (int l = 0; l < L; l++) row[1 + l] = pl_lane[l * pl_stride];
                            row[L + 1] = (float)p.steps;
                        }
                    }
                    idx++;
                }
            }
        }
        if (GRIDS && st != ST_ALIVE && rp_mask != 0) { /* replay the detected photon for its next detector */
            launch(p, sets + set, a.seed, set, key & MCML_KEY_K_MASK, a.mult);
            rl_key = ~0u;
            rl_w = 0;
            alive = true;
        }
    }
    /* A photon's boundary hops count in the launch that resolves them and its steps in the
     * launch that ends it, so one launch's steps - boundary hops can be negative: the
     * counters add modulo 2^64, and their sums over a batch's launches are exact. */
    const uint64_t s0 = wave_sum(c_photons), s1 = wave_sum(c_steps), s2 = wave_sum(c_bnd), s3 = s1 - s2;
    if (lane_id() == 0) { /* wave_sum widens to 64 bits before adding */
        atomicAdd(&a.counters[CNT_PHOTONS], (unsigned long long)s0);
        atomicAdd(&a.counters[CNT_STEPS], (unsigned long long)s1);
        atomicAdd(&a.counters[CNT_BOUNDARY], (unsigned long long)s2);
        atomicAdd(&a.counters[CNT_INTERACTION], (unsigned long long)s3);
    }
    if (STATS) {
        const uint64_t b0 = wave_sum(s_step), b1 = wave_sum(s_pend), b2 = wave_sum(s_refill), b3 = wave_sum(s_dry);
        if (lane_id() == 0) {
            atomicAdd(&a.stats[STAT_STEP], (unsigned long long)b0);
            atomicAdd(&a.stats[STAT_PENDING], (unsigned long long)b1);
            atomicAdd(&a.stats[STAT_REFILL], (unsigned long long)b2);
            atomicAdd(&a.stats[STAT_DRY], (unsigned long long)b3);
            atomicMax(&a.stats[STAT_T_END], (unsigned long long)wall_clock64());
        }
    }
}

template <bool PL, bool SUMS, bool STATS = false, bool GRIDS = false>
__device__ __forceinline__ void first_pass_entry(const FirstPassArgs& a, unsigned char* lds_raw)
{
    /* wave-uniform: one body per launch, so each launch's code is that of one body */
    if (a.lds_sets && a.any_iso)
        first_pass_body<PL, SUMS, STATS, GRIDS, true, true>(a, lds_raw);
    else if (a.lds_sets)
        first_pass_body<PL, SUMS, STATS, GRIDS, true, false>(a, lds_raw);
    else if (a.any_iso)
        first_pass_body<PL, SUMS, STATS, GRIDS, false, true>(a, lds_raw);
    else
        first_pass_body<PL, SUMS, STATS, GRIDS, false, false>(a, lds_raw);
}

} // namespace

/* ----------------------------------------------------------------- first pass */
/* Eight wavefronts per SIMD: the transport is VALU bound and needs every resident wave to
 * hide the latency of its dependent chains; the few uniform values that do not fit in 96
 * SGPRs are kept in VGPR lanes. */
#define FIRST_PASS_ATTRS __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))

extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<false, false>(a, lds_raw);
}

/* Statistics variants: the same passes counting lane-blocks by state and stamping the
 * real-time clock at the start, when the queue first runs dry and at the end. */
extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass_stats(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<false, false, true>(a, lds_raw);
}

extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass_pl_stats(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<true, false, true>(a, lds_raw);
}

/* The drain launch of a -R -A pass: it continues the main launch's spilled photons, takes the
 * main launch's detections as replay jobs into the absorbance grids, and replays its own
 * detections in the lanes that made them (records with their pathlength rows as in
 * mcml_first_pass_pl). */
extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass_grids(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<true, false, false, true>(a, lds_raw);
}

/* The same pass adding each detection to its group's accumulators (no records). */
extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass_sums(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<false, true>(a, lds_raw);
}

/* The same pass recording per-layer pathlength (replay without absorbance maps needs no
 * second pass: a detected photon's first-pass trajectory is its replay trajectory). */
extern "C" __global__ void FIRST_PASS_ATTRS mcml_first_pass_pl(FirstPassArgs a)
{
    stage_tables();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    first_pass_entry<true, false>(a, lds_raw);
}

/* ------------------------------------------------------- reference launch shape */
/* The reference's own launch structure (transport.cu:100-147, 341-471) around this
 * library's photon step, kept to measure what that structure costs on this GPU beside the
 * persistent wavefront kernel: one photon per thread, no refill compaction and no boundary
 * batching (a boundary hit and an interaction diverge within the step), each termination
 * claiming the next photon with its own atomic, at most `steps` steps per launch with the
 * photon kept in global memory between launches, and detections written to a per-thread
 * buffer that the host copies back and reduces after every launch.  Photon k draws from
 * stream (seed, 0, k) as in the persistent kernel, so the detections are the same. */
namespace {
struct RefSlot {
    Photon p;
    uint64_t k; /* photon index: the reference keeps the photon's initial curand state */
};
} // namespace

/* LaunchPhoton_Global (transport.cu:722-739): photon k starts in slot k. */
extern "C" __global__ void __launch_bounds__(256) mcml_refshape_init(RefShapeArgs a)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.threads) return;
    const bool on = tid < a.total;
    if (on) {
        RefSlot* slot = static_cast<RefSlot*>(a.state) + tid;
        Photon p;
        launch(p, a.set, a.seed, 0u, tid, a.mult);
        slot->p = p;
        slot->k = tid;
    }
    a.active[tid] = on ? 1u : 0u;
}

/* MCd (transport.cu:341-471) with this library's hop / interaction / boundary steps. */
extern "C" __global__ void __launch_bounds__(256) mcml_refshape_pass(RefShapeArgs a)
{
    stage_tables();
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.threads) return;
    RefFibers* f = a.fibers + tid;
    int nf = 0;
    uint32_t c_photons = 0, c_steps = 0, c_bnd = 0, n_int = 0;
    if (a.active[tid]) {
        RefSlot* slot = static_cast<RefSlot*>(a.state) + tid;
        Photon p = slot->p;
        uint64_t k = slot->k;
        const DevSet* S = a.set;
        bool live = true;
        for (int ii = 0; ii < a.steps; ii++) {
            float s;
            uint32_t mask = 0;
            int st = ST_ALIVE;
            /* the persistent kernel's step, flat conditions included */
            const bool bnd = hop<false>(p, s);
            if (bnd) {
                c_bnd++;
                st = cross<false>(p, beyond_layer(p), S, 0, mask, nullptr, 0);
            }
            if (!bnd) { /* interact<false>: s > 0 on every uncut hop */
                n_int++;
                p.w -= rintf(p.albedo * p.w);
                spin<true>(p);
            }
            if (!bnd) st = survive(p);
            if (st == ST_ALIVE) continue;
            c_steps += (uint32_t)p.steps;
            if (st == ST_DETECTED) { /* detect(), transport.cu:872-938 */
                for (int i = 1; i <= S->num_sds; i++) {
                    if (!(mask & (1u << i))) continue;
                    const float w = credited(p, S, i);
                    const uint64_t key = k | ((uint64_t)i << MCML_KEY_SDS_SHIFT);
                    if (nf < MCML_REF_SLOTS) {
                        f->sds[nf] = i;
                        f->w[nf] = w;
                        f->key[nf] = key;
                        nf++;
                    } else {
                        const unsigned long long idx = atomicAdd(a.n_spill, 1ull);
                        if (idx < a.spill_cap) {
                            a.spill_key[idx] = key;
                            a.spill_w[idx] = w;
                        }
                    }
                }
            }
            /* transport.cu:439-452: relaunch while photons remain, else retire the slot */
            const unsigned long long t = atomicAdd(a.terminated, 1ull);
            if (a.threads + t < a.total) {
                k = a.threads + t;
                launch(p, S, a.seed, 0u, k, a.mult);
                c_photons++;
            } else {
                live = false;
                break;
            }
        }
        slot->p = p;
        slot->k = k;
        if (!live) a.active[tid] = 0u;
    }
    f->count = nf;
    if (c_photons) atomicAdd(&a.counters[CNT_PHOTONS], (unsigned long long)c_photons);
    if (c_steps) atomicAdd(&a.counters[CNT_STEPS], (unsigned long long)c_steps);
    if (c_bnd) atomicAdd(&a.counters[CNT_BOUNDARY], (unsigned long long)c_bnd);
    if (n_int) atomicAdd(&a.counters[CNT_INTERACTION], (unsigned long long)n_int);
}

/* ----------------------------------------------------------------- replay */
#define REPLAY_BLOCK 256

/* Replay job queue of one wavefront.  Jobs are claimed in chunks from per-(set, SDS)
 * group counters, so a wavefront stays with one group -- and with the group its LDS
 * absorbance tile holds -- until that group runs dry, then moves to the next group with
 * jobs left.  (One counter over all jobs hands a wavefront chunks from wherever the sweep
 * has reached: with a few jobs per lane its lanes soon hold several groups, and the
 * deposits of all but one go to global atomics.) */
struct JobQueue {
    uint64_t next, end; /* the wavefront's claimed, not yet handed out, job range */
    int32_t g, g_start; /* current group; the group it started from */
    bool dry;
};

__device__ __forceinline__ uint64_t take_jobs(JobQueue& q, uint64_t need, bool mine, const ReplayArgs& a)
{
    const uint64_t n = (uint64_t)__popcll(need);
    const uint64_t rank = rank_in(need);
    const uint64_t avail = q.end - q.next;
    uint64_t id = ~0ull;
    if (mine && rank < avail) id = q.next + rank;
    if (n <= avail) {
        q.next += n;
        return id;
    }
    q.next = q.end;
    const uint32_t first = (uint32_t)(__ffsll((long long)__ballot(1)) - 1);
    while (!q.dry) { /* wave-uniform */
        const uint64_t gb = a.grp_begin[q.g], ge = a.grp_begin[q.g + 1];
        unsigned long long base = 0;
        if (lane_id() == first) base = atomicAdd(&a.grp_queue[q.g], (unsigned long long)a.chunk);
        base = bcast_first(base);
        if (gb + base < ge) {
            q.next = gb + base;
            q.end = gb + base + (uint64_t)a.chunk < ge ? gb + base + (uint64_t)a.chunk : ge;
            break;
        }
        q.g = q.g + 1 == a.ngroups ? 0 : q.g + 1;
        if (q.g == q.g_start) q.dry = true;
    }
    if (q.dry) return id;
    if (mine && rank >= avail) {
        const uint64_t t = q.next + (rank - avail);
        if (t < q.end) id = t;
    }
    const uint64_t take = n - avail, left = q.end - q.next;
    q.next += take < left ? take : left;
    return id;
}

/* Add a wavefront's tile (all copies; `tile` is the calling lane's copy) into its group's global
 * grids and clear it (wave-cooperative). */
__device__ __forceinline__ void flush_tile(unsigned long long* tile, uint32_t grp, const ReplayArgs& a)
{
    if (grp == ~0u) return;
    unsigned long long* rz = a.A_rz + (uint64_t)grp * (uint64_t)a.nz * (uint64_t)a.nr;
    unsigned long long* z0 = a.A0_z + (uint64_t)grp * (uint64_t)a.nz;
    unsigned long long* base = tile - (lane_id() % TILE_COPIES) * TILE_WORDS; /* copy 0 */
    for (int i = lane_id(); i < TILE_WORDS; i += 64) {
        unsigned long long v = 0ull;
        for (int c = 0; c < TILE_COPIES; c++) {
            v += base[c * TILE_WORDS + i];
            base[c * TILE_WORDS + i] = 0ull;
        }
        if (v == 0ull) continue;
        if (i < TILE_Z * TILE_R)
            atomicAdd(&rz[(uint64_t)(i / TILE_R) * (uint64_t)a.nr + (uint64_t)(i % TILE_R)], v);
        else
            atomicAdd(&z0[i - TILE_Z * TILE_R], v);
    }
}


extern "C" __global__ void __launch_bounds__(REPLAY_BLOCK) mcml_replay(ReplayArgs a)
{
    stage_tables();
    /* per-lane pathlength accumulators, layer-major so lanes never share a bank word */
    __shared__ float pl_lds[MCML_MAX_LAYERS * REPLAY_BLOCK];
    __shared__ unsigned long long grid_tiles[REPLAY_BLOCK / 64][TILE_WORDS * TILE_COPIES];
    float* pl_lane = pl_lds + threadIdx.x;
    const int L = a.num_layers;
    const int bnd_min = a.bnd_min < 1 ? 1 : a.bnd_min;
    Photon p;
    bool alive = false, pending = false;
    uint64_t job = 0;
    uint32_t jset = 0, jsds = 1;
    const DevSet* S = a.sets;
    GridLane gl = {nullptr, nullptr, nullptr, nullptr, ~0ull, 0, 0, ~0u};
    /* this wavefront's staging copy: wavefront index mod stage_k */
    unsigned long long* stage_copy =
        a.stage == nullptr ? nullptr
                           : a.stage + (((uint64_t)blockIdx.x * (REPLAY_BLOCK / 64) + (threadIdx.x >> 6)) %
                                        (uint64_t)a.stage_k) * (uint64_t)a.ngroups * TILE_WORDS;
    JobQueue jq = {0, 0, 0, 0, a.ngroups < 1};
    if (!jq.dry) { /* start in the group holding this wavefront's share of the job order */
        const uint64_t w = ((uint64_t)blockIdx.x * REPLAY_BLOCK + threadIdx.x) >> 6;
        const uint64_t nw = ((uint64_t)gridDim.x * REPLAY_BLOCK) >> 6;
        const uint64_t j = w * a.njobs / nw;
        int lo = 0, hi = a.ngroups;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.grp_begin[mid] <= j) lo = mid;
            else hi = mid;
        }
        jq.g = jq.g_start = lo;
    }
    if (a.A_rz != nullptr) {
        gl.tile = grid_tiles[threadIdx.x >> 6] + (lane_id() % TILE_COPIES) * TILE_WORDS;
        for (int i = lane_id(); i < TILE_WORDS * TILE_COPIES; i += 64) grid_tiles[threadIdx.x >> 6][i] = 0ull;
    }
    uint32_t n_int = 0;

    for (;;) {
        const uint64_t idle = __ballot(!alive);
        if (!jq.dry && idle != 0) {
            const uint64_t t = take_jobs(jq, idle, !alive, a);
            if (t != ~0ull) {
                job = t;
                const uint64_t key = a.keys[job];
                jset = (uint32_t)(key >> MCML_KEY_SET_SHIFT);
                jsds = (uint32_t)((key >> MCML_KEY_SDS_SHIFT) & 31u);
                S = a.sets + jset;
                launch(p, S, a.seed, jset, key & MCML_KEY_K_MASK, a.mult);
                for (int l = 0; l < L; l++) pl_lane[l * REPLAY_BLOCK] = 0.0f;
                gl.rl_key = ~0ull;
                gl.rl_w = 0;
                if (a.A_rz != nullptr) {
                    gl.grp = jset * (uint32_t)a.num_sds + (jsds - 1);
                    gl.a_rz = a.A_rz + (uint64_t)gl.grp * (uint64_t)a.nz * (uint64_t)a.nr;
                    gl.a0_z = a.A0_z + (uint64_t)gl.grp * (uint64_t)a.nz;
                    gl.stage = stage_copy == nullptr ? nullptr : stage_copy + (uint64_t)gl.grp * TILE_WORDS;
                }
                alive = true;
            }
            /* wave-uniform: when the tile's group holds under a quarter of the live lanes,
             * hand the tile to the group of the first other live lane if that group holds
             * more (a wavefront moving on to a new group keeps a few long photons of the
             * old one for a while) */
            if (gl.tile != nullptr) {
                const uint64_t live = __ballot(alive);
                const uint64_t held = __ballot(alive && gl.grp == gl.tile_grp);
                if (live != 0 && 4 * __popcll(held) < __popcll(live)) {
                    const uint64_t others = live & ~held;
                    const uint32_t cand = (uint32_t)__shfl((int)gl.grp, __ffsll((long long)others) - 1, 64);
                    if (__popcll(__ballot(alive && gl.grp == cand)) > __popcll(held)) {
                        flush_tile(gl.tile, gl.tile_grp, a);
                        gl.tile_grp = cand;
                    }
                }
            }
        }
        if (__ballot(alive) == 0) break;
        int st = ST_ALIVE;
        uint32_t mask = 0;
#pragma unroll
        for (int kk = 0; kk < STEPS_PER_CHECK; kk++) { /* as in the first pass */
            const bool act = alive && !pending && st == ST_ALIVE;
            float s = 0.0f;
            bool bnd = false;
            if (act) bnd = hop<true>(p, s);
            if (act && bnd) pending = true;
            if (act && !bnd) st = interact<true>(p, gl, &a, n_int);
        }
        if (resolve_now(pending, alive, bnd_min) && pending) {
            pending = false;
            st = cross<true>(p, beyond_layer(p), S, (int)jsds, mask, pl_lane, REPLAY_BLOCK);
        }
        if (st != ST_ALIVE) {
            alive = false;
            pl_lane[(p.pl_layer - 1) * REPLAY_BLOCK] = p.pl_cur;
            if (gl.a_rz != nullptr) flush_run(gl, a.nr);
            float* row = a.rows + job * (uint64_t)(L + 2);
            if (st != ST_DETECTED) atomicAdd(a.missed, 1ull);
            row[0] = (st == ST_DETECTED) ? credited(p, S, (int)jsds) : 0.0f;
            for (int l = 0; l < L; l++) row[1 + l] = pl_lane[l * REPLAY_BLOCK];
            row[L + 1] = (float)p.steps;
        }
    }
    if (gl.tile != nullptr) flush_tile(gl.tile, gl.tile_grp, a);
}

/* Add staging copies of the grids' corner ([k][ngroups][TILE_WORDS]) into the grids: one
 * thread per (group, corner word) sums the k copies in a fixed order.  It runs after the
 * kernels that fill them, on the same stream and alone on its cells, so a plain add
 * suffices; the sums are exact either way. */
extern "C" __global__ void __launch_bounds__(256) mcml_stage_reduce(const unsigned long long* stage, int k, int ngroups,
                                                                    unsigned long long* A_rz, unsigned long long* A0_z,
                                                                    int nz, int nr)
{
    const uint64_t n = (uint64_t)ngroups * TILE_WORDS;
    for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (uint64_t)gridDim.x * 256) {
        unsigned long long v = 0ull;
        for (int c = 0; c < k; c++) v += stage[(uint64_t)c * n + t];
        if (v == 0ull) continue;
        const uint64_t g = t / TILE_WORDS, i = t % TILE_WORDS;
        if (i < TILE_Z * TILE_R)
            A_rz[(g * (uint64_t)nz + i / TILE_R) * (uint64_t)nr + i % TILE_R] += v;
        else
            A0_z[g * (uint64_t)nz + (i - TILE_Z * TILE_R)] += v;
    }
}

/* ----------------------------------------------------------------- sorted records */
/* Pathlength dump into file order, in tiles of DUMP_TILE records.  The tile's source
 * indices are read once into LDS; then thread t copies float t of the tile's record image
 * (record t / stride, field t % stride, the division by a multiply-shift that is exact in
 * this range), so both the reads of each source row and the stores are contiguous across
 * lanes.  Sorted weights and public records are written one record per thread. */
#define DUMP_TILE 256
extern "C" __global__ void __launch_bounds__(DUMP_TILE) mcml_dump(GatherArgs a)
{
    __shared__ uint32_t src_idx[DUMP_TILE];
    const uint32_t so = (uint32_t)a.stride_out;
    /* e / so == (e * magic) >> 16 for e < DUMP_TILE * 12 and so <= 12 */
    const uint32_t magic = so > 0 ? (65536u + so - 1u) / so : 0u;
    for (uint64_t base = (uint64_t)blockIdx.x * DUMP_TILE; base < a.n; base += (uint64_t)gridDim.x * DUMP_TILE) {
        const uint64_t left = a.n - base;
        const uint32_t cnt = left < DUMP_TILE ? (uint32_t)left : DUMP_TILE;
        const uint64_t i = base + threadIdx.x;
        if (threadIdx.x < cnt) {
            const uint32_t src = a.order != nullptr ? a.order[i] : (uint32_t)i;
            src_idx[threadIdx.x] = src;
            const float w = a.rows_in[(uint64_t)src * (uint64_t)a.stride_in];
            a.weight_out[i] = w;
            if (a.records_out != nullptr) {
                const uint64_t key = a.sorted_keys[i];
                mcml_record r;
                r.key = ((key >> MCML_KEY_SET_SHIFT) << 40) | (key & MCML_KEY_K_MASK);
                r.sds = (uint32_t)((key >> MCML_KEY_SDS_SHIFT) & 31u);
                r.weight = w;
                a.records_out[i] = r;
            }
        }
        if (so > 0) {
            __syncthreads();
            float* out = a.rows_out + base * (uint64_t)so;
            const uint32_t total = cnt * so;
            for (uint32_t e = threadIdx.x; e < total; e += DUMP_TILE) {
                const uint32_t rr = (e * magic) >> 16;
                const uint32_t f = e - rr * so;
                out[e] = a.rows_in[(uint64_t)src_idx[rr] * (uint64_t)a.stride_in + f];
            }
        }
        __syncthreads();
    }
}

/* Start offset of every (set, SDS) group in the sorted keys (lower bound per group). */
extern "C" __global__ void __launch_bounds__(256) mcml_group_offsets(const uint64_t* keys, uint64_t n, int nsets,
                                                                      int nsds, uint64_t* offsets)
{
    const int groups = nsets * nsds;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= groups; g += gridDim.x * blockDim.x) {
        uint64_t target;
        if (g == groups) target = ~0ull;
        else {
            const uint64_t set = (uint64_t)(g / nsds), sds = (uint64_t)(g % nsds) + 1u;
            target = (set << MCML_KEY_SET_SHIFT) | (sds << MCML_KEY_SDS_SHIFT);
        }
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (keys[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        offsets[g] = (g == groups) ? n : lo;
    }
}

/* Sum of detected weight per (set, SDS) group, in FP64 with a fixed reduction order (one
 * workgroup per group), so the result depends neither on scheduling nor on the other groups
 * of the batch.  Each thread keeps GROUP_ILP independent loads and sums in flight, so that a
 * group of millions of records (one workgroup reading it alone) is not bound by the latency
 * of one load at a time. */
#define GROUP_BLOCK 1024
#define GROUP_ILP 8
extern "C" __global__ void __launch_bounds__(GROUP_BLOCK) mcml_group_sums(const float* w, const uint64_t* offsets,
                                                                          double* sums)
{
    __shared__ double part[GROUP_BLOCK];
    const int g = blockIdx.x;
    const uint64_t b = offsets[g], e = offsets[g + 1];
    constexpr uint64_t T = GROUP_BLOCK;
    double a[GROUP_ILP];
#pragma unroll
    for (int j = 0; j < GROUP_ILP; j++) a[j] = 0.0;
    uint64_t i = b + threadIdx.x;
    for (; i + (GROUP_ILP - 1) * T < e; i += GROUP_ILP * T) {
        float v[GROUP_ILP];
#pragma unroll
        for (int j = 0; j < GROUP_ILP; j++) v[j] = w[i + j * T];
#pragma unroll
        for (int j = 0; j < GROUP_ILP; j++) a[j] += (double)v[j];
    }
#pragma unroll
    for (int j = 0; j < GROUP_ILP - 1; j++) /* fewer than GROUP_ILP left */
        if (i + j * T < e) a[j] += (double)w[i + j * T];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < GROUP_ILP; j++) acc += a[j];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = GROUP_BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[g] = part[0];
}

/* Per-group totals of the sums sink: the slots' 128-bit sums added exactly, then rounded
 * to double; NaN when a record's weight was not a finite non-negative number. */
extern "C" __global__ void __launch_bounds__(256) mcml_sums_finalize(const unsigned long long* accum, int groups,
                                                                      double* reflectance, uint64_t* detected)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    const unsigned long long* e = accum + (uint64_t)g * (MCML_ACC_SLOTS * MCML_ACC_WORDS);
    uint64_t lo = 0, hi = 0, n = 0, bad = 0;
    for (int k = 0; k < MCML_ACC_SLOTS; k++, e += MCML_ACC_WORDS) {
        const uint64_t l = e[ACC_LO];
        lo += l;
        hi += e[ACC_HI] + (lo < l ? 1u : 0u);
        n += e[ACC_COUNT];
        bad += e[ACC_BAD];
    }
    reflectance[g] = bad ? __builtin_nan("") : (double)hi + (double)lo * 0x1p-64;
    detected[g] = n;
}

/* Zero n words with a kernel (no memset node, so a caller may capture it in a graph). */
extern "C" __global__ void __launch_bounds__(256) mcml_clear_words(unsigned long long* p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

/* Self-test of rcp_rn and sqrt_rn against IEEE division and square root: kind 0 checks
 * every significand of x at 16 exponents from 2^-126 to 2^125, kind 1 every normal float
 * (bit patterns [SELFTEST_SQRT_LO, SELFTEST_SQRT_HI): 2^-126 up to the largest finite
 * float).  Mismatches are counted in *bad. */
#define SELFTEST_SQRT_LO 0x00800000ull
#define SELFTEST_SQRT_HI 0x7F800000ull
extern "C" __global__ void __launch_bounds__(256) mcml_selftest_kernel(int kind, unsigned long long* bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool same = true;
    if (kind == 0) {
        if (i >= (16ull << 23)) return;
        const int exps[16] = {-126, -100, -60, -30, -10, -2, -1, 0, 1, 2, 10, 30, 60, 100, 120, 125};
        const float x = __uint_as_float(((uint32_t)(exps[i >> 23] + 127) << 23) | (uint32_t)(i & 0x7FFFFFu));
        same = __float_as_uint(rcp_rn(x)) == __float_as_uint(1.0f / x);
    } else {
        if (i >= SELFTEST_SQRT_HI - SELFTEST_SQRT_LO) return;
        const float x = __uint_as_float((uint32_t)(SELFTEST_SQRT_LO + i));
        same = __float_as_uint(sqrt_rn(x)) == __float_as_uint(sqrtf(x));
    }
    if (!same) atomicAdd(bad, 1ull);
}

/* ----------------------------------------------------------------- counters reset */
/* Zero the control block [queue, n_records, counters...] with a kernel rather than a
 * memset, so the same sequence can be recorded in a HIP graph. */
extern "C" __global__ void mcml_reset_ctrl(unsigned long long* ctrl, int n)
{
    if ((int)threadIdx.x < n) ctrl[threadIdx.x] = 0ull;
}

/* ----------------------------------------------------------------- host launchers */
#include "mcml_launch.h"

hipError_t mcml_launch_reset(unsigned long long* ctrl, int n, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_reset_ctrl, dim3(1), dim3(64), 0, st, ctrl, n);
    return hipGetLastError();
}

/* First pass as a graph: reset kernel -> transport kernel -> read-back of the control
 * block into pinned host memory.  Rebuilt when the launch shape or a buffer changes,
 * otherwise only the two kernel nodes' arguments are replaced before each launch. */
static size_t first_pass_lds(const FirstPassArgs& a, int block)
{
    size_t lds = a.lds_sets ? sizeof(DevSet) * (size_t)a.nsets : 0;
    if (a.row_stride > 1) lds += sizeof(float) * (size_t)a.num_layers * (size_t)block;
    return lds;
}

hipError_t mcml_graph_first_pass(FirstPassGraph& G, const FirstPassArgs& a, int grid, int block,
                                 unsigned long long* ctrl, int nctrl, void* host_ctrl, hipStream_t st)
{
    const size_t lds = first_pass_lds(a, block);
    hipError_t e;
    FirstPassArgs args = a;
    unsigned long long* ctrl_arg = ctrl;
    int n_arg = nctrl;
    void* fp_params[] = {(void*)&args};
    void* rs_params[] = {(void*)&ctrl_arg, (void*)&n_arg};
    hipKernelNodeParams kp = {};
    kp.func = a.A_rz != nullptr ? (void*)mcml_first_pass_grids
                                : (a.row_stride > 1 ? (void*)mcml_first_pass_pl : (void*)mcml_first_pass);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(block);
    kp.sharedMemBytes = (unsigned int)lds;
    kp.kernelParams = fp_params;
    kp.extra = nullptr;
    hipKernelNodeParams rp = {};
    rp.func = (void*)mcml_reset_ctrl;
    rp.gridDim = dim3(1);
    rp.blockDim = dim3(64);
    rp.sharedMemBytes = 0;
    rp.kernelParams = rs_params;
    rp.extra = nullptr;
    const bool same = G.exec != nullptr && G.grid == grid && G.block == block && G.lds == lds && G.ctrl == ctrl &&
                      G.host_ctrl == host_ctrl && G.nctrl == nctrl && G.func == kp.func;
    if (!same) {
        mcml_graph_release(G);
        if ((e = hipGraphCreate(&G.graph, 0)) != hipSuccess) return e;
        if ((e = hipGraphAddKernelNode(&G.reset, G.graph, nullptr, 0, &rp)) != hipSuccess) return e;
        if ((e = hipGraphAddKernelNode(&G.kernel, G.graph, &G.reset, 1, &kp)) != hipSuccess) return e;
        if ((e = hipGraphAddMemcpyNode1D(&G.readback, G.graph, &G.kernel, 1, host_ctrl, ctrl,
                                         sizeof(unsigned long long) * (size_t)nctrl, hipMemcpyDeviceToHost)) !=
            hipSuccess)
            return e;
        if ((e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0)) != hipSuccess) return e;
        G.grid = grid;
        G.block = block;
        G.lds = lds;
        G.ctrl = ctrl;
        G.host_ctrl = host_ctrl;
        G.nctrl = nctrl;
        G.func = kp.func;
    } else {
        if ((e = hipGraphExecKernelNodeSetParams(G.exec, G.reset, &rp)) != hipSuccess) return e;
        if ((e = hipGraphExecKernelNodeSetParams(G.exec, G.kernel, &kp)) != hipSuccess) return e;
    }
    return hipGraphLaunch(G.exec, st);
}

void mcml_graph_release(FirstPassGraph& G)
{
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G = FirstPassGraph();
}

/* Device scratch of the allocation-free entry point (mcml_run_batch_sums_noalloc): the
 * control block, the set table, the ticket prefix and the group accumulators.  Each GPU
 * holds its own copy from the code object's load until the process ends; host code reaches
 * it only through the address mcml_scratch_address returns for the current device. */
static __device__ __align__(256) unsigned char scratch[MCML_NOALLOC_SCRATCH_BYTES];

hipError_t mcml_scratch_address(void** p)
{
    *p = nullptr;
    return hipGetSymbolAddress(p, HIP_SYMBOL(scratch));
}

hipError_t mcml_launch_selftest(unsigned long long* bad, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_selftest_kernel, dim3((unsigned)((16ull << 23) / 256)), dim3(256), 0, st, 0, bad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mcml_selftest_kernel, dim3((unsigned)((SELFTEST_SQRT_HI - SELFTEST_SQRT_LO + 255) / 256)), dim3(256), 0,
                       st, 1, bad + 1);
    return hipGetLastError();
}

hipError_t mcml_launch_first_pass_sums(const FirstPassArgs& a, int grid, int block, hipStream_t st)
{
    const size_t lds = first_pass_lds(a, block);
    hipLaunchKernelGGL(mcml_first_pass_sums, dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

hipError_t mcml_occupancy_first_pass_sums(const FirstPassArgs& a, int block, int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_first_pass_sums, block,
                                                         first_pass_lds(a, block));
}

hipError_t mcml_launch_clear(unsigned long long* p, uint64_t n, hipStream_t st)
{
    const uint64_t grid = std::min<uint64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(mcml_clear_words, dim3((unsigned)std::max<uint64_t>(grid, 1)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t mcml_launch_sums_finalize(const unsigned long long* accum, int groups, double* reflectance,
                                     uint64_t* detected, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_sums_finalize, dim3((groups + 255) / 256), dim3(256), 0, st, accum, groups, reflectance,
                       detected);
    return hipGetLastError();
}

hipError_t mcml_launch_first_pass(const FirstPassArgs& a, int grid, int block, hipStream_t st)
{
    const size_t lds = first_pass_lds(a, block);
    if (a.A_rz != nullptr) {
        hipLaunchKernelGGL(mcml_first_pass_grids, dim3(grid), dim3(block), lds, st, a);
    } else if (a.stats != nullptr) {
        if (a.row_stride > 1) hipLaunchKernelGGL(mcml_first_pass_pl_stats, dim3(grid), dim3(block), lds, st, a);
        else hipLaunchKernelGGL(mcml_first_pass_stats, dim3(grid), dim3(block), lds, st, a);
    } else if (a.row_stride > 1) {
        hipLaunchKernelGGL(mcml_first_pass_pl, dim3(grid), dim3(block), lds, st, a);
    } else {
        hipLaunchKernelGGL(mcml_first_pass, dim3(grid), dim3(block), lds, st, a);
    }
    return hipGetLastError();
}

hipError_t mcml_launch_replay(const ReplayArgs& a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_replay, dim3(grid), dim3(REPLAY_BLOCK), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.stage == nullptr) return e;
    return mcml_launch_stage_reduce(a.stage, a.stage_k, a.ngroups, a.A_rz, a.A0_z, a.nz, a.nr, st);
}

hipError_t mcml_launch_stage_reduce(const unsigned long long* stage, int k, int ngroups, unsigned long long* A_rz,
                                    unsigned long long* A0_z, int nz, int nr, hipStream_t st)
{
    const uint64_t n = (uint64_t)ngroups * TILE_WORDS;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(mcml_stage_reduce, dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(256), 0, st, stage, k,
                       ngroups, A_rz, A0_z, nz, nr);
    return hipGetLastError();
}

size_t mcml_replay_stage_words() { return TILE_WORDS; }

hipError_t mcml_launch_dump(const GatherArgs& a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_dump, dim3(grid), dim3(DUMP_TILE), 0, st, a);
    return hipGetLastError();
}

hipError_t mcml_launch_groups(const uint64_t* keys, uint64_t n, int nsets, int nsds, uint64_t* offsets,
                              const float* w, double* sums, hipStream_t st)
{
    const int groups = nsets * nsds;
    hipLaunchKernelGGL(mcml_group_offsets, dim3((groups + 1 + 255) / 256), dim3(256), 0, st, keys, n, nsets, nsds,
                       offsets);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mcml_group_sums, dim3(groups), dim3(GROUP_BLOCK), 0, st, w, (const uint64_t*)offsets, sums);
    return hipGetLastError();
}

hipError_t mcml_occupancy_first_pass(const FirstPassArgs& a, int block, int* blocks_per_cu)
{
    const size_t lds = first_pass_lds(a, block);
    if (a.A_rz != nullptr)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_first_pass_grids, block, lds);
    if (a.row_stride > 1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_first_pass_pl, block, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_first_pass, block, lds);
}

hipError_t mcml_occupancy_replay(int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_replay, REPLAY_BLOCK, 0);
}

int mcml_replay_block() { return REPLAY_BLOCK; }

size_t mcml_refshape_slot_bytes() { return sizeof(RefSlot); }

hipError_t mcml_launch_refshape_init(const RefShapeArgs& a, int block, hipStream_t st)
{
    const uint64_t grid = (a.threads + (uint64_t)block - 1) / (uint64_t)block;
    hipLaunchKernelGGL(mcml_refshape_init, dim3((unsigned)grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

hipError_t mcml_launch_refshape_pass(const RefShapeArgs& a, int block, hipStream_t st)
{
    const uint64_t grid = (a.threads + (uint64_t)block - 1) / (uint64_t)block;
    hipLaunchKernelGGL(mcml_refshape_pass, dim3((unsigned)grid), dim3(block), 0, st, a);
    return hipGetLastError();
}
