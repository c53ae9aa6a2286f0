/*
 * mcml_device.h - device-resident descriptors shared by the HIP kernels and the host
 * runtime that fills them.  Internal to libmcml; not part of the C ABI.
 */
#ifndef MCML_DEVICE_H
#define MCML_DEVICE_H

#include <stdint.h>
#include "../../include/mcml.h"

/* One layer of one parameter set.  Index 0 is the upper ambient medium, 1..L tissue,
 * L+1 the lower ambient medium (only n is used for the ambient entries).  The first 48 B
 * are what a photon keeps in registers while it stays in the layer; the boundary block
 * holds, for the interface toward layer-1 ("up") and layer+1 ("dn"), the refractive
 * index beyond it and the ratios the Fresnel test needs.  The host forms every entry with
 * the same float expression the transport would evaluate inline. */
struct alignas(16) DevLayer {
    float z_min, z_max, mutr, albedo;  /* albedo = mua * mutr */
    float g, hg_a2, hg_e, hg_cn;       /* ((1 - g*g) * (1 - g*g)) * (1 / (2*g)),
                                          2*(2*g) + (1 - g), -(2*g): the HG denominator
                                          1 - g + 2gu is D = hg_e + hg_cn * (2 - u), and
                                          cos(theta) = hg_c0 - hg_a2 / D^2; for g = 0
                                          hg_a2 = -2, hg_c0 = 3 (cos(theta) = 3 - 2v) */
    float hg_c0, hg_inv2g, n, mua;     /* (1 + g*g) * (1 / (2*g)), 1 / (2*g) */
    float up_n, up_ratio, up_ratio2, up_rnorm; /* n2, n1/n2, (n1*n1)/(n2*n2), (n1-n2)/(n1+n2) */
    float dn_n, dn_ratio, dn_ratio2, dn_rnorm;
};

/* One detector (index 1..S; entry 0 unused), 32 B. */
struct DevDet {
    float pos, radius, r2, cos_a, sin_a, cos_crit, pad0, pad1;
};

/* One tissue parameter set with its photon budget.  Photon k of set j is global queue
 * ticket ticket_begin + k. */
struct alignas(16) DevSet {
    DevLayer layer[MCML_MAX_LAYERS + 2];
    DevDet det[MCML_MAX_SDS + 1];
    uint64_t ticket_begin;
    uint64_t n_photons;
    int32_t num_layers, num_sds;
    float start_weight, det_reflectance;
    int32_t src_oblique, det_oblique;
    float src_radius, src_na, src_theta_max, src_sin_a, src_cos_a;
    /* conservative reject box of all detectors (never rejects an accepted photon):
     * fiber mode: y*y > y2max, or x outside [xlo, xhi] (xlo > xhi disables the x test);
     * ring mode: dist outside [dlo, dhi] or |dz| < cosmin */
    float y2max, xlo, xhi, dlo, dhi, cosmin;
    int32_t pad[2];
};

/* Work counters accumulated by the first-pass kernel (one atomic per lane at exit). */
enum { CNT_PHOTONS = 0, CNT_STEPS = 1, CNT_BOUNDARY = 2, CNT_INTERACTION = 3, CNT_N = 4 };

/* Lane statistics of the statistics variant of the first pass (mcml_opts.lane_stats):
 * lane-blocks (one lane for one block of STEPS_PER_CHECK steps) by what the lane was doing
 * when the block began, and real-time clock stamps (the two minima stored complemented, so
 * that all three are atomic maxima over a zeroed block). */
enum { STAT_STEP = 0, STAT_PENDING = 1, STAT_REFILL = 2, STAT_DRY = 3, STAT_T_START = 4, STAT_T_DRY = 5,
       STAT_T_END = 6, STAT_N = 7 };

/* Detection-record sort key: set (19 bits) | SDS (5 bits) | photon index (40 bits), so
 * that ascending keys group records by (set, SDS) in photon order -- the order of the
 * pathlength_SDS_# files. */
#define MCML_KEY_SET_SHIFT 45
#define MCML_KEY_SDS_SHIFT 40
#define MCML_KEY_K_MASK ((1ull << 40) - 1ull)
#define MCML_MAX_SETS (1 << 19)

struct FirstPassArgs {
    const DevSet* sets;
    const uint64_t* set_begin; /* [nsets+1] ticket prefix: set j owns [set_begin[j], set_begin[j+1]) */
    int32_t nsets;
    int32_t chunk;             /* tickets a wavefront claims per queue access (>= 64) */
    int32_t refill_min;        /* refill once this many lanes of the wavefront are idle */
    int32_t bnd_min;           /* resolve boundary hits once this many lanes wait on one */
    int32_t lds_sets;          /* 1: stage the set table in LDS (dynamic shared memory) */
    int32_t num_layers;        /* L (pathlength recording) */
    int32_t any_iso;           /* 1: some tissue layer of some set has g = 0 (isotropic) */
    int32_t pad_iso;
    uint64_t total;            /* total tickets */
    uint64_t seed;
    const uint32_t* mult;      /* MWC multiplier table */
    unsigned long long* queue;     /* global ticket counter */
    uint64_t* rec_key;             /* [cap] sort keys of detection records */
    float* rec_row;                /* [cap * row_stride]: weight, or [weight, PL_1..PL_L, steps] */
    int32_t row_stride;            /* 1, or L + 2 when the kernel records pathlengths */
    int32_t pad;
    unsigned long long* n_records; /* records appended (never beyond cap: see stop_at) */
    uint64_t cap;
    unsigned long long* counters;  /* [CNT_N] */
    uint64_t stop_at;              /* claim no new tickets once n_records reaches this (~0 = never);
                                      cap - stop_at covers every photon that can still be run */
    unsigned long long* accum;     /* sums sink (mcml_first_pass_sums): MCML_ACC_SLOTS entries of
                                      MCML_ACC_WORDS words per (set, SDS) group instead of records */
    unsigned long long* stats;     /* [STAT_N] lane statistics (statistics variant only) */
    /* Drain compaction.  Once the ticket queue is exhausted, a wavefront with at most
     * spill_below lanes still holding photons writes those photons to `spill` and exits,
     * so that sparse wavefronts stop competing for their SIMDs; the host continues the
     * spilled photons in a launch with resume = 1, where tickets index `resume_buf` instead
     * of new photons (and which may spill again). */
    float* spill;                  /* [SPILL_FIELDS(L) x spill_cap] SoA photon states (NULL = no spill) */
    unsigned long long* n_spill;   /* photons written to `spill` */
    uint64_t spill_cap;            /* column stride of `spill`: at least every lane of the launch */
    const float* resume_buf;       /* resume = 1: the photons to continue, SoA as `spill` */
    uint64_t resume_cap;           /* column stride of resume_buf */
    int32_t spill_below;
    int32_t resume;
    /* Absorbance replays in the drain launch (mcml_first_pass_grids; NULL otherwise).  Its
     * tickets past the resume_photons spilled photons are replay jobs, the record keys
     * replay_keys[t - resume_photons] of the main launch; a lane runs the job's photon again
     * for its detector, and a lane whose own photon is detected replays it for each detector
     * that recorded it, booking the drops into that (set, SDS) group's slabs of nz*nr / nz
     * cells. */
    const uint64_t* replay_keys;
    uint64_t resume_photons;
    unsigned long long* A_rz;
    unsigned long long* A0_z;
    unsigned long long* replay_missed; /* replays that ended without the detection (never expected) */
    int32_t nr, nz;
    float rdr, rdz;                    /* 1/dr, 1/dz: bins are trunc(r * rdr), trunc(z * rdz) */
    /* staging copies of the grids' corner for the replays' drops (NULL = none), laid out as
     * ReplayArgs.stage: [stage_k][stage_groups][corner words], copy = wavefront index mod
     * stage_k, added into the grids by mcml_launch_stage_reduce after the last drain launch */
    unsigned long long* stage;
    int32_t stage_k, stage_groups;
};

/* Fields of one spilled photon (each a 32-bit column of the SoA spill arrays): position,
 * direction, weight, layer, steps, first-scatter flag, the RNG state (3 words), the running
 * pathlength and its layer, the record key (2 words), the pending boundary, the replay state
 * of mcml_first_pass_grids (detectors still to replay for, the drop run's cell and weight;
 * no replay in the other variants), then the per-layer pathlength totals. */
enum { SP_X = 0, SP_Y, SP_Z, SP_DX, SP_DY, SP_DZ, SP_W, SP_LAYER, SP_STEPS, SP_FIRST, SP_RNG_LO, SP_RNG_HI, SP_RNG_A,
       SP_PL_CUR, SP_PL_LAYER, SP_KEY_LO, SP_KEY_HI, SP_PENDING, SP_RP_MASK, SP_RL_KEY, SP_RL_W, SP_PL0 };
#define SPILL_FIELDS(L) (SP_PL0 + (L))

/* Per-(set, SDS) accumulators of the sums sink.  A group's detected weight is summed exactly
 * as a 128-bit fixed-point integer with 64 fraction bits (words lo, hi), so the total does
 * not depend on the order of the additions; `count` counts its records and `bad` those
 * whose weight is NaN (the reference's pos = 0 ring quirk), which make the sum NaN.  Each
 * group has MCML_ACC_SLOTS copies, one 64-byte line each, picked by workgroup, to spread the
 * atomics. */
#define MCML_ACC_SLOTS 16
#define MCML_ACC_WORDS 8
enum { ACC_LO = 0, ACC_HI = 1, ACC_COUNT = 2, ACC_BAD = 3 };

struct ReplayArgs {
    const DevSet* sets;
    const uint64_t* keys;      /* [njobs] sorted record keys: each is one replay job */
    uint64_t njobs;
    uint64_t seed;
    const uint32_t* mult;
    /* jobs of (set, SDS) group g are [grp_begin[g], grp_begin[g+1]); grp_queue[g] counts
     * the jobs of group g handed out so far (zero at launch) */
    const uint64_t* grp_begin;
    unsigned long long* grp_queue;
    int32_t ngroups;
    int32_t chunk;
    int32_t num_layers;
    int32_t bnd_min;
    int32_t pad2;
    float* rows;               /* [njobs * (L+2)]: [weight, PL_1..PL_L, steps] */
    /* absorbance grids (NULL = off): per (set, sds) slab of nz*nr / nz cells */
    unsigned long long* A_rz;
    unsigned long long* A0_z;
    int32_t num_sds, nr, nz, pad;
    float rdr, rdz;            /* 1/dr, 1/dz: bins are trunc(r * rdr), trunc(z * rdz) */
    unsigned long long* missed; /* jobs whose replay ended without the detection (never expected) */
    /* staging copies of the grids' tile corner (NULL = none): [stage_k][ngroups][tile words],
     * zero at launch; a drop in the corner that the wavefront's LDS tile cannot take goes to
     * copy (wavefront index mod stage_k) of its group, and mcml_stage_reduce adds the
     * copies into the grids after the replay */
    unsigned long long* stage;
    int32_t stage_k, stage_pad;
};

/* Reference launch shape (transport.cu:57-147, 341-471): one photon per thread, at most
 * MCML_REF_STEPS steps per launch, photon state in global memory between launches.
 * RefFibers is one thread's detection buffer for one launch (Fibers, header.h:127-133):
 * the photon key stands in for the reference's saved curand state. */
#define MCML_REF_STEPS 50000 /* NUMSTEPS_GPU, header.h:39 */
#define MCML_REF_SLOTS 20    /* SDS_detected_temp_size, header.h:61 */
struct RefFibers {
    int32_t count;
    int32_t sds[MCML_REF_SLOTS];
    float w[MCML_REF_SLOTS];
    uint64_t key[MCML_REF_SLOTS];
};

struct RefShapeArgs {
    const DevSet* set;         /* the one parameter set of the run */
    uint64_t total;            /* photons */
    uint64_t threads;          /* T photon slots; photon k < T starts in slot k */
    uint64_t seed;
    const uint32_t* mult;
    void* state;               /* [T] slot state between launches (mcml_refshape_slot_bytes each) */
    uint32_t* active;          /* [T] 1 while the slot still has a photon */
    RefFibers* fibers;         /* [T] this launch's detections */
    unsigned long long* terminated; /* photons ended so far; photon T + t is the t-th relaunch */
    uint64_t* spill_key;       /* detections beyond a full RefFibers (the reference drops them) */
    float* spill_w;
    unsigned long long* n_spill;
    uint64_t spill_cap;
    unsigned long long* counters; /* [CNT_N] */
    int32_t steps;             /* step budget of one launch */
    int32_t pad;
};

/* Records in sorted order: gather rows, build the public records, reduce per (set, SDS). */
struct GatherArgs {
    const uint64_t* sorted_keys;
    const uint32_t* order;     /* source index of each sorted record (NULL = identity) */
    const float* rows_in;
    int32_t stride_in;         /* floats per source row */
    int32_t stride_out;        /* floats per gathered row (0 = do not gather rows) */
    uint64_t n;
    float* rows_out;
    float* weight_out;         /* [n] sorted weights */
    mcml_record* records_out;  /* [n] public records (NULL = skip) */
};

#endif /* MCML_DEVICE_H */
