/*
 * mcml.h - C ABI of the MI355X photon-transport library (libmcml.so).
 *
 * The reference program has no library interface: its operator surface is the process
 * `MCML_GPU sim_set.json input.txt output.txt [-R] [-A] [-P] [-AP] [-B] [-G n] [-LS] [-h]`
 * (main.cu:17-145) and one internal seam,
 *     void DoOneSimulation(SimulationStruct*, int, char*, SimOptions, GPUInfo**)
 * (transport.cu:57), called once per input line by main.cu:130-135.  This header replaces
 * that seam with plain-C entry points (POD structs, pointers and sizes; no C++ or torch
 * types), and exposes the file-level surface (input parsing, output writers, the CLI main)
 * so that the drop-in executable is a thin shell over it.
 *
 * Conventions
 *   - Every function returns 0 (MCML_OK) or a negative MCML_ERR_* code; no exceptions cross
 *     the ABI.  A context keeps the last error message (mcml_last_error).
 *   - Result buffers are allocated by the library and released by mcml_result_free.
 *   - Calls are synchronous.  A context is used by one host thread at a time.
 *   - Indices that the reference prints or writes 1-based (SDS numbers) stay 1-based.
 */
#ifndef MCML_H
#define MCML_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCML_MAX_LAYERS 10 /* header.h:38 PRESET_NUM_LAYER */
#define MCML_MAX_SDS 30    /* header.h:44 PRESET_NUM_DETECTOR */

/* Error codes. */
#define MCML_OK 0
#define MCML_ERR_ARG -1     /* invalid argument or option combination */
#define MCML_ERR_INPUT -2   /* sim_set.json / input.txt content rejected (io.cpp:53-160) */
#define MCML_ERR_IO -3      /* file could not be read or written */
#define MCML_ERR_DEVICE -4  /* HIP runtime error */
#define MCML_ERR_NOMEM -5   /* host or device allocation failed */
#define MCML_ERR_NODEV -6   /* no usable GPU */
#define MCML_ERR_LIMIT -7   /* size beyond a documented maximum */

/* Reference constants kept as defaults (header.h:57-71). */
#define MCML_WEIGHTI 429497u          /* roulette threshold, W-units */
#define MCML_CHANCE 0.1f              /* roulette survival probability */
#define MCML_MAX_SCATTER 10000        /* first-pass step cap */
#define MCML_DEFAULT_NR 1000
#define MCML_DEFAULT_NZ 500
#define MCML_DEFAULT_DR 0.01f
#define MCML_DEFAULT_DZ 0.01f
#define MCML_MUA_CHANGE_RATIO 0.001   /* average-pathlength perturbation */

/* One tissue layer as written on a line of input.txt: "h mua mus n g" (io.cpp:141-146).
 * The thickness of the last layer is not read by the reference and is ignored here. */
typedef struct {
    float thickness; /* cm */
    float mua;       /* 1/cm */
    float mus;       /* 1/cm */
    float n;
    float g;
} mcml_layer;

/* One probe of sim_set.json "probes" (io.cpp:181-191). */
typedef struct {
    float pos;       /* cm from the source axis (0 for the source) */
    float NA;
    float radius;    /* cm */
    float angle_deg; /* degrees; source toward +x, detectors toward -x */
} mcml_probe;

/* One simulation = one line of input.txt under one sim_set.json (SimulationStruct,
 * header.h:113-125, before derivation). */
typedef struct {
    uint64_t number_photons;
    int32_t number_layers;
    int32_t num_sds;
    float detector_reflectance;
    float upper_n;
    float lower_n;
    int32_t source_probe_oblique;
    int32_t detector_probe_oblique;
    mcml_probe source;
    mcml_probe detectors[MCML_MAX_SDS];
    mcml_layer layers[MCML_MAX_LAYERS];
    /* Absorbance grid (header.h:65-68 defaults; optional sim_set.json keys
     * "grid_nr", "grid_nz", "grid_dr", "grid_dz"). */
    int32_t grid_nr;
    int32_t grid_nz;
    float grid_dr;
    float grid_dz;
} mcml_sim;

/* Quantities derived from an mcml_sim exactly as read_mua_mus does (io.cpp:179-237),
 * plus the per-probe trigonometry the transport needs.  Index 0 of `layer` is the upper
 * ambient medium, 1..L the tissue, L+1 the lower ambient medium; detector arrays are
 * indexed 1..S (index 0 unused), like detInfo (header.h:105-111). */
typedef struct {
    float z_min, z_max, mutr, mua, g, n;
} mcml_layer_derived;

typedef struct {
    uint64_t number_photons;
    int32_t num_layers;
    int32_t num_sds;
    float start_weight;        /* W-units, float((uint)(0xffffffff*(1-Rsp))) */
    float detector_reflectance;
    int32_t source_oblique;
    int32_t detector_oblique;
    mcml_layer_derived layer[MCML_MAX_LAYERS + 2];
    float src_radius;
    float src_na;
    float src_theta_max;       /* asin(NA/upper_n) */
    float src_sin_a, src_cos_a;
    float det_pos[MCML_MAX_SDS + 1];
    float det_radius[MCML_MAX_SDS + 1];
    float det_r2[MCML_MAX_SDS + 1];
    float det_cos_a[MCML_MAX_SDS + 1];
    float det_sin_a[MCML_MAX_SDS + 1];
    float det_crit[MCML_MAX_SDS + 1];      /* asin(NA/upper_n) (io.cpp:196) */
    float det_cos_crit[MCML_MAX_SDS + 1];  /* cos(det_crit) */
    int32_t grid_nr, grid_nz;
    float grid_dr, grid_dz;
} mcml_derived;

/* Run options: the reference's SimOptions (header.h:195-203) plus launch knobs. */
typedef struct {
    int32_t replay;          /* -R  */
    int32_t absorbance;      /* -A  */
    int32_t pathlength;      /* -P  */
    int32_t avg_pathlength;  /* -AP */
    int32_t binary;          /* -B  (writers only) */
    int32_t block_size;      /* threads per workgroup; 0 = tuned default */
    int32_t blocks_per_cu;   /* resident workgroups per CU; 0 = occupancy maximum */
    int32_t chunk;           /* photon ids a wavefront claims per queue access; 0 = tuned.
                                With reference_shape: steps per launch (0 = 50 000) */
    int32_t refill_min;      /* refill when at least this many lanes are idle; 0 = tuned */
    int32_t boundary_min;    /* resolve boundary hits when at least this many lanes wait on
                                one; 0 = tuned (results never depend on the launch knobs) */
    int32_t reference_shape; /* 1 = run the first pass in the reference's launch structure
                                (transport.cu:76-147): one photon per thread in workgroups of
                                block_size threads (0 = 32), 50 000-step relaunches, per-launch
                                host read-back and reduction of per-thread detection buffers;
                                replay follows as a separate pass even without -A.  One set
                                per run.  Same records and sums as 0; kept as a measured
                                comparison, it is several times slower */
    int32_t use_graph;       /* 1 = run the first pass through a cached HIP graph (reset
                                kernel + transport kernel + counter read-back), updating
                                only its kernel arguments between calls */
    int32_t skip_records;    /* 1 = do not return the per-photon detection records (sums,
                                counts and pathlength rows are still returned) */
    uint64_t record_cap;     /* initial capacity of the device record buffer, in records;
                                0 = the most records the run's photons and detectors can
                                produce (while that fits in a quarter of free device memory).
                                A first pass that fills a smaller buffer pauses, the buffer
                                grows and the pass resumes; the records do not depend on it */
    int32_t lane_stats;      /* 1 = run the first pass's statistics variant and fill
                                mcml_result.lane_blocks, ms_drain and ms_clock (a few percent
                                slower; not with use_graph) */
    int32_t spill_below;     /* drain compaction: once the photon queue is exhausted, a
                                wavefront left with at most this many photons hands them to a
                                compacted follow-up launch; 0 = default (32), -1 = off */
    int32_t separate_replay; /* -A: 1 = replay the detected photons in a pass of their own after
                                the first pass (as the reference does) instead of within it.
                                Also taken when the grids exceed 4 GiB or 65535 bins a side,
                                and with lane_stats or reference_shape */
} mcml_opts;

/* First-pass detection record: photon `key` = (set << 40) | photon index, detected by
 * detector `sds` (1-based) with weight `weight` (W-units; ring mode includes the arc
 * fraction).  One photon may produce records for several detectors (transport.cu:878-929). */
typedef struct {
    uint64_t key;
    uint32_t sds;
    float weight;
} mcml_record;

typedef struct {
    int32_t nsets;
    int32_t num_sds;            /* S, equal for all sets of a run */
    int32_t num_layers;         /* L, equal for all sets of a run */
    int32_t grid_nr, grid_nz;
    /* first pass */
    double* reflectance;        /* [nsets*S] sum of detected weight, W-units */
    uint64_t* detected;         /* [nsets*S] detected photon count */
    uint64_t n_records;
    mcml_record* records;       /* sorted by (set, sds, photon index); NULL with skip_records */
    /* replay (opts.replay) */
    int32_t replayed;
    double* replay_reflectance; /* [nsets*S] */
    uint64_t* pl_offset;        /* [nsets*S+1] record offsets into `pathlength` */
    float* pathlength;          /* records of L+2 floats [weight, PL_1..PL_L, steps],
                                   for set j / SDS s at pl_offset[j*S+s-1], photon order */
    uint64_t* A_rz;             /* [nsets*S*nz*nr] u64 W-units, z-major (opts.absorbance) */
    uint64_t* A0_z;             /* [nsets*S*nz] */
    /* work counters (first pass) */
    uint64_t photons;           /* photons launched */
    uint64_t steps;             /* all hop steps */
    uint64_t steps_boundary;    /* steps that hit a layer boundary */
    uint64_t steps_interaction; /* steps with absorption + scattering */
    /* timings, milliseconds: kernels by HIP events on the library stream, walls by
     * steady clock around the whole phase (allocation, copies, host reductions) */
    double ms_first_pass_kernel;  /* HIP events around all the first pass's launches */
    double ms_replay_kernel;
    double ms_dump_kernel;
    double ms_first_pass_wall;
    double ms_replay_wall;
    int32_t first_pass_launches; /* kernel launches the first pass needed (more than one only
                                    when the record buffer filled and the pass resumed) */
    int32_t block_size, grid_size, chunk, refill_min, boundary_min; /* launch shape used */
    /* opts.lane_stats only.  lane_blocks: lane-blocks (one lane for one block of photon steps
     * between two wavefront decisions) by the lane's state when the block began: [0] stepping,
     * [1] waiting at a boundary, [2] waiting for a refill, [3] idle because the ticket queue
     * had run dry.  ms_drain: from the first wavefront finding the queue dry to the last
     * wavefront's exit; ms_clock: from the first wavefront's start to the last one's exit
     * (the GPU's real-time clock, first launch of the pass). */
    uint64_t lane_blocks[4];
    double ms_drain;
    double ms_clock;
    int32_t drain_launches;     /* follow-up launches that continued spilled photons */
} mcml_result;

typedef struct mcml_ctx mcml_ctx;

/* Open the library on HIP device `device` (main.cu:120 cudaSetDevice).  *err receives a
 * code; NULL is returned on failure. */
mcml_ctx* mcml_create(int device, int* err);
void mcml_destroy(mcml_ctx* ctx);
const char* mcml_last_error(const mcml_ctx* ctx);
/* Device name as reported by the runtime (summary.json "sim_GPU", io.cpp:364). */
const char* mcml_device_name(const mcml_ctx* ctx);
int mcml_device_count(int* count);

/* Self-test of the device arithmetic the transport relies on (no counterpart in the
 * reference): the hardware reciprocal and reciprocal square root with their one-step
 * corrections must give exactly IEEE 1.0f / x (every significand at 16 exponents) and
 * sqrtf(x) (every normal float), as the CPU restatement computes them.  mismatches[0], [1]
 * receive the counts.
 * mcml_create runs it and fails with MCML_ERR_DEVICE on any mismatch. */
int mcml_selftest_math(mcml_ctx* ctx, uint64_t* mismatches);

/* Derive the per-simulation constants (io.cpp:179-237).  A negative or NaN layer thickness
 * is rejected with MCML_ERR_INPUT (the reference reads it unchecked). */
int mcml_derive(const mcml_sim* sim, mcml_derived* out);

/* One simulation: replaces DoOneSimulation (transport.cu:57-272) minus the file output.
 * Photon k of the run draws from the stream (seed, set 0, k). */
int mcml_run(mcml_ctx* ctx, const mcml_sim* sim, const mcml_opts* opts, uint64_t seed,
             mcml_result* out);

/* Many parameter sets with their own photon budgets in one persistent launch; set j uses
 * streams (seed, j, k).  All sets must share number_layers and num_sds.  Replaces the
 * sequential per-line loop of main.cu:130-135. */
int mcml_run_batch(mcml_ctx* ctx, const mcml_sim* sims, int nsets, const mcml_opts* opts,
                   uint64_t seed, mcml_result* out);

void mcml_result_free(mcml_result* res);

/* mcml_run_batch's per-(set, SDS) detected weight and detected count (its `reflectance` and
 * `detected` arrays) computed without allocating: the call makes no hipMalloc, hipFree or
 * host allocation.  No counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 *   d_reflectance [nsets*S] doubles and d_detected [nsets*S] uint64 are DEVICE pointers on
 *   the context's GPU, written at [set * S + sds - 1] like the result arrays.
 *   The device work runs in one fixed scratch array per GPU (set table, counters, per-group
 *   accumulators; MCML_NOALLOC_SCRATCH_MIB MiB, held until the process ends): at most
 *   MCML_NOALLOC_MAX_SETS sets and MCML_NOALLOC_MAX_GROUPS (set, SDS) groups per call, else
 *   MCML_ERR_LIMIT.  Only the launch knobs of opts apply; replay, absorbance, pathlength,
 *   average pathlength, reference_shape and use_graph are rejected with MCML_ERR_ARG.
 *   Detected weights are summed exactly (128-bit fixed point) and rounded to double once, so
 *   the sums match mcml_run_batch's FP64 sums to their rounding and do not depend on the
 *   launch shape; the counts are equal.
 *   Every call on one GPU uses that GPU's scratch: calls that overlap on one GPU (from two
 *   threads or two contexts) race.  The call is synchronous like the others. */
#define MCML_NOALLOC_MAX_SETS 1024
#define MCML_NOALLOC_MAX_GROUPS 4096
#define MCML_NOALLOC_SCRATCH_MIB 7
int mcml_run_batch_sums_noalloc(mcml_ctx* ctx, const mcml_sim* sims, int nsets, const mcml_opts* opts,
                                uint64_t seed, double* d_reflectance, uint64_t* d_detected);

/* Pick the launch shape (block_size, chunk, refill_min, boundary_min) for these parameter
 * sets: times the first pass of a reduced copy of the batch (about `trial_photons` photons in
 * total) for each candidate shape and writes the fastest into *opts (other fields
 * untouched; a boundary_min already set in *opts is kept).  No
 * counterpart in the reference, which sizes its grid from compute-capability tables
 * (io.cpp:471-493). */
int mcml_autotune(mcml_ctx* ctx, const mcml_sim* sims, int nsets, uint64_t trial_photons, mcml_opts* opts);

/* Average pathlength per SDS and layer, ln(sum w / sum w e^{-dmua L}) / dmua with
 * dmua = 0.001 mua (transport.cu:274-297).  out: [S*L] (SDS-major). */
int mcml_average_pathlength(const mcml_sim* sim, const mcml_result* res, int set, double* out);

/* White Monte Carlo rescaling of a replayed result (res->replayed): for every set j and
 * each of its m new absorption vectors mua_new[(j*m + i)*L + l] (1/cm, layer l+1), the
 * detected weight per SDS the same photons would carry, w * exp(-sum_l (mua'_l - mua_l) L_l),
 * summed per (set, SDS) into out[(j*m + i)*S + sds-1] (W-units, like res->reflectance).
 * sims are the sets the result was simulated with.  The analogue of the reference's
 * perturbation estimate in calculate_average_pathlength (transport.cu:285-292), here for
 * arbitrary mua.  out_sq (optional, same shape) receives the sums of squared rescaled
 * weights, from which a label's standard error follows. */
int mcml_white_rescale(mcml_ctx* ctx, const mcml_sim* sims, const mcml_result* res, const float* mua_new, int m,
                       double* out, double* out_sq);

/* ---------------------------------------------------------------------------------
 * Reflectance surrogate network (no counterpart in the reference).  A fully connected
 * network with widths[0] inputs, GELU hidden layers widths[1..n-2] and widths[n-1] linear
 * outputs; matrix products on the matrix cores in bf16 with FP32 accumulation, FP32 master
 * weights, Adam.  Parameters: per layer W (out x in, row-major) then b (out).
 * --------------------------------------------------------------------------------- */
typedef struct mcml_mlp mcml_mlp;
mcml_mlp* mcml_mlp_create(int device, const int* widths, int nwidths, uint64_t seed, int* err);
void mcml_mlp_destroy(mcml_mlp* mlp);
int64_t mcml_mlp_num_params(const mcml_mlp* mlp);
/* enable = 1/0 allows/forbids the fused inference kernel (-1: leave as is); returns 1 when
 * predict will use it, 0 when it runs layer by layer, < 0 on error. */
int mcml_mlp_use_fused(mcml_mlp* mlp, int enable);
/* write = 0: copy parameters out; write = 1: load them (and reset the Adam state). */
int mcml_mlp_params(mcml_mlp* mlp, float* host, int64_t n, int write);
/* y[B x out] = network(x[B x in]); *ms = kernel time of the forward pass (HIP events).
 * Networks of <= 32 inputs, 1..8 GELU layers of 256 and <= 16 outputs run as one fused
 * kernel (weights staged in LDS, activations in registers); others layer by layer. */
int mcml_mlp_predict(mcml_mlp* mlp, const float* x, int B, float* y, double* ms);
/* Gradient of mean_b mean_o w_bo (y_bo - t_bo)^2 at the current parameters; w holds one
 * weight per target [B x out] and may be NULL (all ones).  *loss = that mean. */
int mcml_mlp_gradient(mcml_mlp* mlp, const float* x, const float* t, const float* w, int B, float* grad, float* loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8; L2 weight decay wd added to the gradient) with
 * gradient grad. */
int mcml_mlp_adam_step(mcml_mlp* mlp, const float* grad, float lr, float wd);
/* Mini-batch Adam over n samples (per-target weights w [n x out] or NULL) for `epochs`
 * epochs of n / batch steps, cosine learning-rate decay from lr to lr_final, reshuffled
 * every epoch from seed; epoch_loss[e] = loss of epoch e's last batch (may be NULL);
 * *ms = device time of the whole run. */
int mcml_mlp_train(mcml_mlp* mlp, const float* x, const float* t, const float* w, int n, int epochs, int batch,
                   float lr, float lr_final, float wd, uint64_t seed, float* epoch_loss, double* ms);

/* Read sim_set.json + input.txt (read_mua_mus, io.cpp:35-256).  On success *sims holds
 * *nsims simulations (free with mcml_free_sims).  On rejection the reference's message is
 * copied to msg. */
int mcml_read_input(const char* sim_set_json, const char* input_txt, mcml_sim** sims,
                    int* nsims, char* msg, size_t msglen);
void mcml_free_sims(mcml_sim* sims);

/* Timing block written to summary.json (SummaryStruct, header.h:168-174). */
typedef struct {
    double sim_time_s;    /* first pass */
    double replay_time_s; /* replay + outputs */
    const char* gpu_name;
} mcml_summary_info;

/* Write the outputs of set `set` of a result into directory `outdir` (NULL = current
 * directory) and append to `output_txt`, exactly as DoOneSimulation's tail does
 * (transport.cu:156-251 with io.cpp:11-33, 281-469). */
int mcml_write_outputs(const mcml_sim* sim, const mcml_opts* opts, const mcml_result* res,
                       int set, const char* output_txt, const char* outdir,
                       const mcml_summary_info* info);

/* The drop-in executable: argv as for MCML_GPU (main.cu:17-145), plus "-S seed". */
int mcml_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* MCML_H */
