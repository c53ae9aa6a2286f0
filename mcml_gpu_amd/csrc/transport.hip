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
 *
 * Arithmetic: only correctly rounded IEEE binary32 operations, the deterministic
 * functions of mcml_detmath.h and explicit operation order, compiled with
 * -ffp-contract=off, so each photon's trajectory is a function of (seed, set, photon)
 * alone and is bit-identical to a CPU evaluation of the same steps.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcml_detmath.h"
#include "mcml_device.h"
#include "mwc_multipliers.h"

#define RPI_F 0.318309886f /* 1/pi as the reference writes it (header.h:41) */

namespace {

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
    float zmin, zmax, mutr, albedo, g, hg_a, hg_b, hg_c, hg_d, hg_inv2g, n;
};

__device__ __forceinline__ void load_layer(Photon& p, const DevSet* S)
{
    const float4* L = reinterpret_cast<const float4*>(&S->layer[p.layer]);
    const float4 a = L[0], b = L[1], c = L[2];
    p.zmin = a.x;
    p.zmax = a.y;
    p.mutr = a.z;
    p.albedo = a.w;
    p.g = b.x;
    p.hg_a = b.y;
    p.hg_b = b.z;
    p.hg_c = b.w;
    p.hg_d = c.x;
    p.hg_inv2g = c.y;
    p.n = c.z;
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
    load_layer(p, S);
}

/* Henyey-Greenstein deflection + azimuth + renormalisation (Spin, transport.cu:742-786),
 * written with one division and one square root on the common path:
 * P = sin(theta) sqrt(1-dz^2), Q = sin(theta) / sqrt(1-dz^2), and a second-order
 * rsqrt about 1 for the renormalisation (the norm never drifts more than a few ulp). */
__device__ __forceinline__ void spin(Photon& p)
{
    float cost;
    if (p.g == 0.0f) {
        cost = 2.0f * mcml_uniform(&p.rng) - 1.0f;
    } else {
        const float temp = p.hg_a / (p.hg_b + p.hg_c * mcml_uniform(&p.rng));
        cost = (p.hg_d - temp * temp) * p.hg_inv2g;
    }
    const float a = 1.0f - cost * cost;
    float sinp, cosp;
    mcml_sincos_turns(mcml_uniform(&p.rng), &sinp, &cosp);
    const float b = 1.0f - p.dz * p.dz;
    if (b == 0.0f) {
        const float sint = sqrtf(a);
        p.dx = sint * cosp;
        p.dy = sint * sinp;
        p.dz = copysignf(cost, p.dz * cost);
    } else {
        const float P = sqrtf(a * b);
        const float Q = P / b;
        const float cxp = Q * (p.dx * p.dz * cosp - p.dy * sinp) + p.dx * cost;
        const float cyp = Q * (p.dy * p.dz * cosp + p.dx * sinp) + p.dy * cost;
        p.dz = -P * cosp + p.dz * cost;
        p.dx = cxp;
        p.dy = cyp;
    }
    const float d = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz - 1.0f;
    const float r = fmaf(d, fmaf(d, 0.375f, -0.5f), 1.0f);
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
    const float n1 = p.n;
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
    r = 2.0f * sqrtf((1.0f - ci * ci) * (1.0f - e) * e * ci * ci);
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
    p.dz = copysignf(sqrtf(1.0f - e), p.dz);
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
    const float dist = sqrtf(p.x * p.x + p.y * p.y);
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
    const float dist = sqrtf(p.x * p.x + p.y * p.y);
    float t = (dist * dist + D.pos * D.pos - D.radius * D.radius) / (2.0f * dist * D.pos);
    if (t > 1.0f) t = 1.0f;
    return p.w * mcml_acos(t) * RPI_F;
}

enum StepOutcome { ST_ALIVE = 0, ST_DEAD = 1, ST_DETECTED = 2 };

/* Hop (transport.cu:372-395): sample the step, stop it at the layer boundary, move, count
 * the step, clamp z.  Returns true when the hop ended on a boundary (*nl = layer beyond). */
template <bool REPLAY>
__device__ __forceinline__ bool hop(Photon& p, float& s, int& nl, float* pl_lane, int pl_stride)
{
    s = -mcml_log_unit(mcml_uniform(&p.rng)) * p.mutr;
    nl = p.layer;
    if (p.z + s * p.dz < p.zmin) {
        nl--;
        s = (p.zmin - p.z) / p.dz;
    }
    if (p.z + s * p.dz > p.zmax) {
        nl++;
        s = (p.zmax - p.z) / p.dz;
    }
    p.x += p.dx * s;
    p.y += p.dy * s;
    p.z += p.dz * s;
    p.steps++;
    if (REPLAY) pl_lane[(p.layer - 1) * pl_stride] += s; /* transport.cu:529 */
    if (p.z > p.zmax) p.z = p.zmax;
    if (p.z < p.zmin) p.z = p.zmin;
    return nl != p.layer;
}

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

/* Absorbance bookkeeping of one replayed photon (transport.cu:580-600): the first drop goes
 * to A0_z, later ones to A_rz, merged while consecutive drops stay in one cell. */
struct GridLane {
    unsigned long long* a_rz;
    unsigned long long* a0_z;
    uint64_t rl_idx;
    uint64_t rl_w;
};

/* Interaction (transport.cu:430-437 / 573-603): drop weight and scatter when the hop did
 * not end on a boundary and s > 0, then the survival test. */
template <bool REPLAY>
__device__ __forceinline__ int interact(Photon& p, float s, GridLane& gl, const ReplayArgs* ra, uint32_t& n_int)
{
    if (s > 0.0f) {
        n_int++;
        const float w_temp = rintf(p.albedo * p.w);
        p.w -= w_temp;
        if (REPLAY && gl.a_rz != nullptr) {
            float qz = p.z / ra->dz;
            qz = fminf(fmaxf(qz, 0.0f), (float)(ra->nz - 1));
            const int zb = (int)qz;
            if (p.first) {
                atomicAdd(&gl.a0_z[zb], (unsigned long long)w_temp);
            } else {
                float qr = sqrtf(p.x * p.x + p.y * p.y) / ra->dr;
                qr = fminf(fmaxf(qr, 0.0f), (float)(ra->nr - 1));
                const uint64_t idx = (uint64_t)zb * (uint64_t)ra->nr + (uint64_t)(int)qr;
                if (idx == gl.rl_idx) {
                    gl.rl_w += (uint64_t)w_temp;
                } else {
                    if (gl.rl_w != 0) atomicAdd(&gl.a_rz[gl.rl_idx], (unsigned long long)gl.rl_w);
                    gl.rl_idx = idx;
                    gl.rl_w = (uint64_t)w_temp;
                }
            }
        }
        p.first = 0;
        spin(p);
    }
    return survive(p);
}

/* Boundary (transport.cu:397-428): Fresnel, then escape through the top (detection, or
 * the detector face's reflection back into the tissue), loss through the bottom, or entry
 * into the next tissue layer; then the survival test.  `target` = 0 tests every detector
 * (first pass), otherwise that detector only (replay).  On ST_DETECTED *mask holds the
 * accepting detectors (bit i = SDS i). */
__device__ __forceinline__ int cross(Photon& p, int nl, const DevSet* S, int target, uint32_t& mask)
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
 * counter.  Returns this lane's id, or ~0 when it gets none. */
struct WaveQueue {
    uint64_t next, end;
    bool dry;
};

__device__ __forceinline__ uint64_t take_ids(WaveQueue& q, uint64_t need, bool mine, unsigned long long* counter,
                                             uint64_t chunk, uint64_t total)
{
    const int n = __popcll(need);
    const uint32_t rank = rank_in(need);
    const uint64_t avail = q.end - q.next;
    uint64_t id = ~0ull;
    if (mine && rank < avail) id = q.next + rank;
    if ((uint64_t)n > avail) {
        unsigned long long base = 0;
        if (lane_id() == (uint32_t)(__ffsll((long long)__ballot(1)) - 1)) base = atomicAdd(counter, chunk);
        base = bcast_first(base);
        uint64_t cend = base + chunk;
        if (cend > total) cend = total;
        if (base >= total) {
            q.dry = true;
            cend = base;
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

} // namespace

/* ----------------------------------------------------------------- first pass */
extern "C" __global__ void __launch_bounds__(256) mcml_first_pass(FirstPassArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const DevSet* sets = a.sets;
    if (a.lds_sets) { /* stage the (small) set table in LDS; reads go through one pointer */
        const uint4* src = reinterpret_cast<const uint4*>(a.sets);
        uint4* dst = reinterpret_cast<uint4*>(lds_raw);
        const int n16 = (int)(sizeof(DevSet) / 16) * a.nsets;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
        sets = reinterpret_cast<const DevSet*>(lds_raw);
    }
    Photon p;
    bool alive = false, pending = false;
    int pend_layer = 0;
    const DevSet* S = sets;
    uint64_t key = 0;
    WaveQueue q = {0, 0, false};
    uint64_t c_photons = 0, c_steps = 0, c_bnd = 0, c_int = 0;
    uint32_t n_int = 0;
    GridLane gl = {nullptr, nullptr, ~0ull, 0};
    const int refill_min = a.refill_min < 1 ? 1 : a.refill_min;
    const int bnd_min = a.bnd_min < 1 ? 1 : a.bnd_min;

    for (;;) {
        const uint64_t dead = __ballot(!alive);
        const int ndead = __popcll(dead);
        if (!q.dry && (ndead >= refill_min || ndead == 64)) {
            const uint64_t t = take_ids(q, dead, !alive, a.queue, (uint64_t)a.chunk, a.total);
            if (t != ~0ull) {
                const uint32_t set = a.nsets > 1 ? set_of_ticket(a.set_begin, a.nsets, t) : 0u;
                S = sets + set;
                const uint64_t k = t - a.set_begin[set];
                key = ((uint64_t)set << 40) | k;
                launch(p, S, a.seed, set, k, a.mult);
                alive = true;
                c_photons++;
            }
        }
        if (__ballot(alive) == 0) break;
        int st = ST_ALIVE;
        uint32_t mask = 0;
        if (alive && !pending) {
            float s;
            int nl;
            if (hop<false>(p, s, nl, nullptr, 0)) {
                pending = true;
                pend_layer = nl;
            } else {
                st = interact<false>(p, s, gl, nullptr, n_int);
            }
        }
        if (resolve_now(pending, alive, bnd_min) && pending) {
            pending = false;
            c_bnd++;
            st = cross(p, pend_layer, S, 0, mask);
        }
        if (st != ST_ALIVE) {
            alive = false;
            c_steps += (uint64_t)p.steps;
            c_int += n_int;
            n_int = 0;
            if (st == ST_DETECTED) {
                const unsigned long long base = atomicAdd(a.n_records, (unsigned long long)__popc(mask));
                unsigned long long idx = base;
                const int ns = S->num_sds;
                for (int i = 1; i <= ns; i++) {
                    if (!(mask & (1u << i))) continue;
                    if (idx < a.cap) {
                        mcml_record r;
                        r.key = key;
                        r.sds = (uint32_t)i;
                        r.weight = credited(p, S, i);
                        a.records[idx] = r;
                    }
                    idx++;
                }
            }
        }
    }
    const uint64_t s0 = wave_sum(c_photons), s1 = wave_sum(c_steps), s2 = wave_sum(c_bnd), s3 = wave_sum(c_int);
    if (lane_id() == 0) {
        atomicAdd(&a.counters[CNT_PHOTONS], (unsigned long long)s0);
        atomicAdd(&a.counters[CNT_STEPS], (unsigned long long)s1);
        atomicAdd(&a.counters[CNT_BOUNDARY], (unsigned long long)s2);
        atomicAdd(&a.counters[CNT_INTERACTION], (unsigned long long)s3);
    }
}

/* ----------------------------------------------------------------- replay */
#define REPLAY_BLOCK 256

extern "C" __global__ void __launch_bounds__(REPLAY_BLOCK) mcml_replay(ReplayArgs a)
{
    /* per-lane pathlength accumulators, layer-major so lanes never share a bank word */
    __shared__ float pl_lds[MCML_MAX_LAYERS * REPLAY_BLOCK];
    float* pl_lane = pl_lds + threadIdx.x;
    const int L = a.num_layers;
    const int bnd_min = a.bnd_min < 1 ? 1 : a.bnd_min;
    WaveQueue q = {0, 0, false};
    Photon p;
    bool alive = false, pending = false;
    int pend_layer = 0;
    uint64_t job = 0;
    ReplayJob J = {0, 0, 0};
    const DevSet* S = a.sets;
    GridLane gl = {nullptr, nullptr, ~0ull, 0};
    uint32_t n_int = 0;

    for (;;) {
        const uint64_t idle = __ballot(!alive);
        if (!q.dry && idle != 0) {
            const uint64_t t = take_ids(q, idle, !alive, a.queue, (uint64_t)a.chunk, a.njobs);
            if (t != ~0ull) {
                job = t;
                J = a.jobs[job];
                S = a.sets + J.set;
                launch(p, S, a.seed, J.set, J.k, a.mult);
                for (int l = 0; l < L; l++) pl_lane[l * REPLAY_BLOCK] = 0.0f;
                gl.rl_idx = ~0ull;
                gl.rl_w = 0;
                if (a.A_rz != nullptr) {
                    const uint64_t slab = (uint64_t)J.set * (uint64_t)a.num_sds + (uint64_t)(J.sds - 1);
                    gl.a_rz = a.A_rz + slab * (uint64_t)a.nz * (uint64_t)a.nr;
                    gl.a0_z = a.A0_z + slab * (uint64_t)a.nz;
                }
                alive = true;
            }
        }
        if (__ballot(alive) == 0) break;
        int st = ST_ALIVE;
        uint32_t mask = 0;
        if (alive && !pending) {
            float s;
            int nl;
            if (hop<true>(p, s, nl, pl_lane, REPLAY_BLOCK)) {
                pending = true;
                pend_layer = nl;
            } else {
                st = interact<true>(p, s, gl, &a, n_int);
            }
        }
        if (resolve_now(pending, alive, bnd_min) && pending) {
            pending = false;
            st = cross(p, pend_layer, S, (int)J.sds, mask);
        }
        if (st != ST_ALIVE) {
            alive = false;
            if (gl.a_rz != nullptr && gl.rl_w != 0) atomicAdd(&gl.a_rz[gl.rl_idx], (unsigned long long)gl.rl_w);
            a.weight[job] = (st == ST_DETECTED) ? credited(p, S, (int)J.sds) : 0.0f;
            a.steps[job] = (float)p.steps;
            for (int l = 0; l < L; l++) a.pl[(uint64_t)l * a.njobs + job] = pl_lane[l * REPLAY_BLOCK];
        }
    }
}

/* ----------------------------------------------------------------- dump */
/* Thread t writes float t of the record image: record t / (L+2), field t % (L+2). */
extern "C" __global__ void __launch_bounds__(256) mcml_dump(DumpArgs a)
{
    const uint64_t stride = (uint64_t)a.num_layers + 2u;
    const uint64_t total = a.njobs * stride;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t rec = t / stride;
        const uint32_t f = (uint32_t)(t - rec * stride);
        float v;
        if (f == 0) v = a.weight[rec];
        else if (f <= (uint32_t)a.num_layers) v = a.pl[(uint64_t)(f - 1) * a.njobs + rec];
        else v = a.steps[rec];
        a.out[t] = v;
    }
}

/* ----------------------------------------------------------------- host launchers */
#include "mcml_launch.h"

hipError_t mcml_launch_first_pass(const FirstPassArgs& a, int grid, int block, hipStream_t st)
{
    const size_t lds = a.lds_sets ? sizeof(DevSet) * (size_t)a.nsets : 0;
    hipLaunchKernelGGL(mcml_first_pass, dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

hipError_t mcml_launch_replay(const ReplayArgs& a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_replay, dim3(grid), dim3(REPLAY_BLOCK), 0, st, a);
    return hipGetLastError();
}

hipError_t mcml_launch_dump(const DumpArgs& a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_dump, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t mcml_occupancy_first_pass(int block, size_t lds_bytes, int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_first_pass, block, lds_bytes);
}

hipError_t mcml_occupancy_replay(int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, mcml_replay, REPLAY_BLOCK, 0);
}

int mcml_replay_block() { return REPLAY_BLOCK; }
