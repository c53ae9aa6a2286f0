/*
 * runtime.cpp - host orchestration behind the C ABI (include/mcml.h).
 *
 * Replaces DoOneSimulation (transport.cu:57-272) and its helpers:
 *   - derivation of per-simulation constants (read_mua_mus, io.cpp:179-237);
 *   - the first pass: one persistent kernel launch per run instead of the reference's
 *     50 000-step relaunch loop with per-launch host copies and reductions
 *     (transport.cu:114-147, fiber.cu:3-13); detections come back as compact records;
 *   - the replay: one launch over all detected photons of all detectors instead of
 *     per-detector batches of at most T photons (transport.cu:160-222);
 *   - the host reductions (transport.cu:274-332) over the returned records.
 * Device buffers live in the context and only grow, so repeated runs do not allocate.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include <sys/mman.h>

#include "../../include/mcml.h"
#include "mcml_device.h"
#include "mcml_launch.h"
#include "mcml_stream.h"
#include "mwc_multipliers.h"
#include "white.h"

namespace {

constexpr float kPI = 3.141592654f; /* header.h:40 */
/* launch defaults (results never depend on them: each photon's stream is fixed) */
constexpr int kDefaultChunk = 128;
constexpr int kDefaultRefill = 4;
constexpr int kDefaultBoundary = 4;
constexpr size_t kLdsTableMax = 16384;
/* The control block: ticket counter, records appended, work counters, lane statistics, the
 * drain's spill counter and resume counter (the words a first pass resets and reads back),
 * then two words for the math self-test. */
constexpr int kCtrlQueue = 0, kCtrlRecords = 1, kCtrlCounters = 2;
constexpr int kCtrlStats = kCtrlCounters + CNT_N;
constexpr int kCtrlSpill = kCtrlStats + STAT_N;
constexpr int kCtrlResume = kCtrlSpill + 1;
constexpr int kCtrlMissed = kCtrlResume + 1; /* replays that ended without their detection */
constexpr int kCtrlFirstPass = kCtrlMissed + 1;
constexpr int kCtrlWords = 24;
static_assert(kCtrlFirstPass + 2 <= kCtrlWords, "two words for the math self-test");
constexpr int kDefaultSpillBelow = 32; /* drain: a wavefront with at most this many photons left spills them */
constexpr uint64_t kFinalDrain = 16384; /* a drain launch of at most this many photons runs them to the end */
constexpr size_t kStageBytes = 128u << 20; /* pinned staging buffer for large read-backs */
constexpr size_t kSmallStageBytes = 16u << 20; /* ... and the one for mid-sized read-backs */
constexpr size_t kDirectCopyBytes = 64u << 10;  /* read-backs this small go straight to pageable memory */
constexpr size_t kStreamChunkBytes = 16u << 20; /* read-back chunk when the caller is told of each one */
/* replay: at most this many staging copies of the grids' corner, in at most this many bytes */
constexpr size_t kReplayStageCopies = 64;
constexpr size_t kReplayStageBytes = 64u << 20;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

} // namespace

struct mcml_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    int clock_khz = 0; /* the real-time clock the statistics variant stamps */
    char name[256] = {0};
    char err[1024] = {0};
    uint32_t* d_mult = nullptr;
    /* control block: [0] first-pass ticket counter, [1] records appended, [2..5] work
     * counters; mirrored into pinned host memory */
    unsigned long long* ctrl = nullptr;
    unsigned long long* h_ctrl = nullptr;
    FirstPassGraph graph;
    DevBuf sets, set_begin;
    DevBuf rec_key, rec_row;          /* detection records as appended by the first pass */
    DevBuf skey, sidx, stemp;         /* sorted keys, their source indices, sort scratch */
    DevBuf rows, weights, precs;      /* sorted rows (file image), weights, public records */
    DevBuf offsets, sums;             /* per (set, SDS) group: start offsets, FP64 sums */
    DevBuf rrows, arz, a0z;           /* replay rows and absorbance grids */
    DevBuf rgrp, rgq;                 /* replay: job range and hand-out counter per (set, SDS) */
    DevBuf rstage;                    /* replay: staging copies of the grids' tile corner */
    DevBuf wrows, woff, wold, wnew, wout; /* White Monte Carlo inputs and output */
    DevBuf ref_state, ref_active, ref_fibers, ref_spill_key, ref_spill_w; /* reference launch shape */
    DevBuf spill[2];                  /* drain compaction: photons handed from one launch to the next */
    void* stage[2] = {nullptr, nullptr}; /* pinned read-back staging, allocated on first use */
    void* stage_small = nullptr;         /* pinned staging for read-backs up to kSmallStageBytes */
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    /* allocation-free entry point: this device's scratch array (its device address, taken
     * at creation) and a pinned host staging area for its set table and ticket prefix */
    void* scratch = nullptr;
    void* na_stage = nullptr;
};

static_assert(sizeof(DevSet) == 2048, "the scratch layout assumes 2048-byte set entries");
static_assert(MCML_NOALLOC_SCRATCH_BYTES <= ((size_t)MCML_NOALLOC_SCRATCH_MIB << 20), "scratch size as documented");

namespace {

void set_err(mcml_ctx* c, const char* fmt, ...)
{
    if (!c) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_err(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));                     \
            return MCML_ERR_DEVICE;                                                          \
        }                                                                                    \
    } while (0)

int ensure(mcml_ctx* c, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return MCML_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) {
            set_err(c, "hipFree failed: %s", hipGetErrorString(e));
            return MCML_ERR_DEVICE;
        }
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        set_err(c, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return MCML_ERR_NOMEM;
    }
    b.bytes = bytes;
    return MCML_OK;
}

/* Grow b to `bytes`, keeping its first `keep` bytes (a device-to-device copy). */
int ensure_keep(mcml_ctx* c, DevBuf& b, size_t bytes, size_t keep)
{
    if (b.bytes >= bytes) return MCML_OK;
    if (keep == 0 || b.p == nullptr) return ensure(c, b, bytes);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_err(c, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return MCML_ERR_NOMEM;
    }
    if ((e = hipMemcpyAsync(p, b.p, keep, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        (void)hipFree(p);
        set_err(c, "record buffer copy failed: %s", hipGetErrorString(e));
        return MCML_ERR_DEVICE;
    }
    (void)hipFree(b.p);
    b.p = p;
    b.bytes = bytes;
    return MCML_OK;
}

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

template <class T>
T* as(DevBuf& b) { return static_cast<T*>(b.p); }

/* Host array of a result.  Large ones are 2 MB aligned and advised for transparent huge
 * pages, so that their first touch (the read-back) faults once per 2 MB rather than once
 * per 4 KB.  Every result array is released with free(). */
void* host_alloc(size_t bytes)
{
    if (bytes < (32u << 20)) return malloc(bytes);
    void* p = nullptr;
    if (posix_memalign(&p, 2u << 20, bytes) != 0) return nullptr;
    (void)madvise(p, bytes, MADV_HUGEPAGE);
    return p;
}

/* memcpy on up to eight host threads (pinned staging buffer -> pageable result). */
void parallel_memcpy(char* dst, const char* src, size_t len)
{
    const size_t slice = 8u << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({8, hw, (len + slice - 1) / slice});
    if (nt <= 1) {
        memcpy(dst, src, len);
        return;
    }
    const size_t per = (len + nt - 1) / nt;
    std::vector<std::thread> th;
    size_t done = per; /* slices [0, done) are copied by this thread or a started one */
    try {
        for (size_t t = 1; t < nt && done < len; t++, done += per) {
            const size_t off = t * per;
            th.emplace_back([=] { memcpy(dst + off, src + off, std::min(per, len - off)); });
        }
    } catch (...) { /* no thread to spare: this thread copies the rest */
    }
    memcpy(dst, src, std::min(per, len));
    if (done < len) memcpy(dst + done, src + done, len - done);
    for (std::thread& t : th) t.join();
}

/* Device-to-host read-back of a result array.  Small copies go straight to `dst`.  Mid-sized
 * ones (an absorbance grid of a few MB) are one DMA into a pinned buffer and a memcpy: a
 * direct copy into pageable memory moved a 3.2 MB grid at about 2.5 GB/s.  Large ones run
 * through two pinned staging buffers, the DMA of one chunk overlapping the host threads' copy
 * of the previous one into the pageable destination.  With `landed`, the caller is told the
 * bytes [0, b) that are in `dst` after every chunk, and chunks are 16 MB instead of 128 MB so
 * that it can start on the first ones early. */
int copy_d2h(mcml_ctx* c, void* dst, const void* src, size_t bytes,
             const std::function<void(size_t)>* landed = nullptr)
{
    if (bytes == 0) return MCML_OK;
    if (bytes <= kDirectCopyBytes) {
        HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (landed) (*landed)(bytes);
        return MCML_OK;
    }
    if (bytes <= kSmallStageBytes) {
        if (!c->stage_small) HIP_TRY(c, hipHostMalloc(&c->stage_small, kSmallStageBytes, hipHostMallocDefault));
        HIP_TRY(c, hipMemcpyAsync(c->stage_small, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        parallel_memcpy(static_cast<char*>(dst), static_cast<const char*>(c->stage_small), bytes);
        if (landed) (*landed)(bytes);
        return MCML_OK;
    }
    for (int i = 0; i < 2; i++) {
        if (!c->stage[i]) HIP_TRY(c, hipHostMalloc(&c->stage[i], kStageBytes, hipHostMallocDefault));
        if (!c->stage_ev[i]) HIP_TRY(c, hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
    }
    const size_t chunk = landed ? kStreamChunkBytes : kStageBytes;
    const size_t n = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t off = i * chunk, len = std::min(chunk, bytes - off);
        hipError_t e = hipMemcpyAsync(c->stage[i & 1], static_cast<const char*>(src) + off, len,
                                      hipMemcpyDeviceToHost, c->stream);
        return e != hipSuccess ? e : hipEventRecord(c->stage_ev[i & 1], c->stream);
    };
    HIP_TRY(c, issue(0));
    if (n > 1) HIP_TRY(c, issue(1));
    for (size_t i = 0; i < n; i++) {
        HIP_TRY(c, hipEventSynchronize(c->stage_ev[i & 1]));
        const size_t off = i * chunk, len = std::min(chunk, bytes - off);
        parallel_memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(c->stage[i & 1]), len);
        if (i + 2 < n) HIP_TRY(c, issue(i + 2)); /* into the buffer just drained */
        if (landed) (*landed)(off + len);
    }
    return MCML_OK;
}

double ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* 1 when a tissue layer of set S scatters isotropically (g = 0): the transport then keeps
 * the isotropic cos(theta) select in its step. */
int devset_iso(const DevSet& S, int num_layers)
{
    for (int l = 1; l <= num_layers; l++)
        if (S.layer[l].hg_cn == 0.0f) return 1;
    return 0;
}

void fill_devset(const mcml_derived& d, DevSet& S)
{
    memset(&S, 0, sizeof(S));
    for (int l = 0; l < d.num_layers + 2; l++) {
        const mcml_layer_derived& in = d.layer[l];
        DevLayer& o = S.layer[l];
        o.z_min = in.z_min;
        o.z_max = in.z_max;
        o.mutr = in.mutr;
        o.mua = in.mua;
        o.albedo = in.mua * in.mutr;
        o.g = in.g;
        o.hg_inv2g = 1.0f / (2.0f * in.g);
        const float hg_a = 1.0f - in.g * in.g;
        o.hg_a2 = (hg_a * hg_a) * o.hg_inv2g;
        o.hg_e = 2.0f * (2.0f * in.g) + (1.0f - in.g);
        o.hg_cn = -(2.0f * in.g);
        o.hg_c0 = (1.0f + in.g * in.g) * o.hg_inv2g;
        if (in.g == 0.0f) { /* isotropic: cos(theta) = hg_c0 + hg_a2 * v = 3 - 2v for the draw's 1.k */
            o.hg_a2 = -2.0f;
            o.hg_c0 = 3.0f;
        }
        o.n = in.n;
        if (l >= 1 && l <= d.num_layers) {
            const float n1 = in.n;
            const float nu = d.layer[l - 1].n, nd = d.layer[l + 1].n;
            o.up_n = nu;
            o.up_ratio = n1 / nu;
            o.up_ratio2 = (n1 * n1) / (nu * nu);
            o.up_rnorm = (n1 - nu) / (n1 + nu);
            o.dn_n = nd;
            o.dn_ratio = n1 / nd;
            o.dn_ratio2 = (n1 * n1) / (nd * nd);
            o.dn_rnorm = (n1 - nd) / (n1 + nd);
        }
    }
    for (int i = 1; i <= d.num_sds; i++) {
        S.det[i].pos = d.det_pos[i];
        S.det[i].radius = d.det_radius[i];
        S.det[i].r2 = d.det_r2[i];
        S.det[i].cos_a = d.det_cos_a[i];
        S.det[i].sin_a = d.det_sin_a[i];
        S.det[i].cos_crit = d.det_cos_crit[i];
    }
    S.n_photons = d.number_photons;
    S.num_layers = d.num_layers;
    S.num_sds = d.num_sds;
    S.start_weight = d.start_weight;
    S.det_reflectance = d.detector_reflectance;
    S.src_oblique = d.source_oblique;
    S.det_oblique = d.detector_oblique;
    S.src_radius = d.src_radius;
    S.src_na = d.src_na;
    S.src_theta_max = d.src_theta_max;
    S.src_sin_a = d.src_sin_a;
    S.src_cos_a = d.src_cos_a;
    /* detector reject box; the x extent of a tilted fibre's footprint is r / cos(angle),
     * widened by 1% so float rounding can never move an accepted photon outside it */
    S.y2max = 0.0f;
    S.xlo = 3.0e38f;
    S.xhi = -3.0e38f;
    S.dlo = 3.0e38f;
    S.dhi = -3.0e38f;
    S.cosmin = 3.0e38f;
    bool x_ok = true;
    for (int i = 1; i <= d.num_sds; i++) {
        S.y2max = std::max(S.y2max, d.det_r2[i]);
        const float ca = fabsf(d.det_cos_a[i]);
        if (ca < 0.01f) x_ok = false;
        else {
            const float half = 1.01f * d.det_radius[i] / ca + 1e-6f;
            S.xlo = std::min(S.xlo, d.det_pos[i] - half);
            S.xhi = std::max(S.xhi, d.det_pos[i] + half);
        }
        S.dlo = std::min(S.dlo, d.det_pos[i] - d.det_radius[i]);
        S.dhi = std::max(S.dhi, d.det_pos[i] + d.det_radius[i]);
        S.cosmin = std::min(S.cosmin, d.det_cos_crit[i]);
    }
    if (!x_ok) {
        S.xlo = 1.0f;
        S.xhi = -1.0f;
    }
}


/* The most detectors of one set that can record the same photon: photons are killed at
 * their first detection (transport.cu:405-409), so a photon yields at most this many
 * records.  A fibre accepts only exit points with |x - pos| <= radius / |cos(angle)|, a
 * ring only distances within [pos - radius, pos + radius]; the answer is the largest
 * number of these intervals sharing a point (1 when the footprints are disjoint). */
int max_overlap(const mcml_derived& d)
{
    const int S = d.num_sds;
    std::vector<std::pair<double, int>> ev; /* (coordinate, +1 start / -1 end) */
    for (int i = 1; i <= S; i++) {
        double lo, hi;
        if (d.detector_oblique) {
            const double ca = fabs((double)d.det_cos_a[i]);
            if (ca < 0.01) return S;
            const double half = 1.01 * d.det_radius[i] / ca + 1e-6;
            lo = d.det_pos[i] - half;
            hi = d.det_pos[i] + half;
        } else {
            lo = (double)d.det_pos[i] - d.det_radius[i] - 1e-6;
            hi = (double)d.det_pos[i] + d.det_radius[i] + 1e-6;
        }
        if (!(lo <= hi)) return S; /* NaN or an inverted interval: no bound but S */
        ev.push_back({lo, +1});
        ev.push_back({hi, -1});
    }
    std::sort(ev.begin(), ev.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
        return a.first < b.first || (a.first == b.first && a.second > b.second); /* starts first */
    });
    int cur = 0, best = 1;
    for (const auto& e : ev) {
        cur += e.second;
        best = std::max(best, cur);
    }
    return std::min(best, S);
}

} // namespace

/* ====================================================================== derivation */
extern "C" int mcml_derive(const mcml_sim* s, mcml_derived* d)
{
    if (!s || !d) return MCML_ERR_ARG;
    memset(d, 0, sizeof(*d));
    const int L = s->number_layers, S = s->num_sds;
    if (L < 1 || L > MCML_MAX_LAYERS || S < 1 || S > MCML_MAX_SDS) return MCML_ERR_LIMIT;
    d->number_photons = s->number_photons;
    d->num_layers = L;
    d->num_sds = S;
    d->detector_reflectance = s->detector_reflectance;
    d->source_oblique = s->source_probe_oblique;
    d->detector_oblique = s->detector_probe_oblique;
    /* layer table, io.cpp:204-228: cumulative float z, mutr = 1/(mua+mus), the last
     * tissue layer reaches at least 10 cm */
    d->layer[0].n = s->upper_n;
    float z_lo = 0.0f, z_hi = 0.0f;
    for (int l = 1; l <= L; l++) {
        const mcml_layer& in = s->layers[l - 1];
        const float h = (l < L) ? in.thickness : 0.0f;
        if (!(h >= 0.0f)) return MCML_ERR_INPUT; /* the transport relies on z_min <= z_max */
        z_hi += h;
        mcml_layer_derived& o = d->layer[l];
        o.z_min = z_lo;
        o.z_max = z_hi;
        o.mutr = 1.0f / (in.mua + in.mus);
        o.mua = in.mua;
        o.g = in.g;
        o.n = in.n;
        z_lo += h;
    }
    if (d->layer[L].z_max < 10.0f) d->layer[L].z_max = 10.0f;
    d->layer[L + 1].n = s->lower_n;
    /* specular loss at entry, io.cpp:231-237 */
    const double n0 = d->layer[0].n, n1 = d->layer[1].n;
    double rsp = (n0 - n1) / (n0 + n1);
    rsp = rsp * rsp;
    d->start_weight = (float)(unsigned int)((double)0xffffffffu * (1 - rsp));
    /* probes: angles in degrees -> radians with the reference's float PI (io.cpp:184, 190) */
    const float a_src = s->source.angle_deg * kPI / 180;
    d->src_radius = s->source.radius;
    d->src_na = s->source.NA;
    d->src_theta_max = asinf(s->source.NA / d->layer[0].n);
    d->src_sin_a = (float)sin((double)a_src);
    d->src_cos_a = (float)cos((double)a_src);
    for (int i = 1; i <= S; i++) {
        const mcml_probe& p = s->detectors[i - 1];
        const float a = p.angle_deg * kPI / 180;
        d->det_pos[i] = p.pos;
        d->det_radius[i] = p.radius;
        d->det_r2[i] = p.radius * p.radius;
        d->det_cos_a[i] = (float)cos((double)a);
        d->det_sin_a[i] = (float)sin((double)a);
        d->det_crit[i] = asinf(p.NA / s->upper_n); /* io.cpp:196 */
        d->det_cos_crit[i] = (float)cos((double)d->det_crit[i]);
    }
    d->grid_nr = s->grid_nr > 0 ? s->grid_nr : MCML_DEFAULT_NR;
    d->grid_nz = s->grid_nz > 0 ? s->grid_nz : MCML_DEFAULT_NZ;
    d->grid_dr = s->grid_dr > 0 ? s->grid_dr : MCML_DEFAULT_DR;
    d->grid_dz = s->grid_dz > 0 ? s->grid_dz : MCML_DEFAULT_DZ;
    return MCML_OK;
}

/* ====================================================================== context */
extern "C" int mcml_device_count(int* count)
{
    if (!count) return MCML_ERR_ARG;
    *count = 0;
    if (hipGetDeviceCount(count) != hipSuccess) {
        *count = 0;
        return MCML_ERR_NODEV;
    }
    return MCML_OK;
}

extern "C" mcml_ctx* mcml_create(int device, int* err)
{
    int dummy;
    if (!err) err = &dummy;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        *err = MCML_ERR_NODEV;
        return nullptr;
    }
    if (device < 0 || device >= n) {
        *err = MCML_ERR_ARG;
        return nullptr;
    }
    mcml_ctx* c = new mcml_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        *err = MCML_ERR_DEVICE;
        mcml_destroy(c);
        return nullptr;
    }
    c->cu_count = prop.multiProcessorCount;
    if (hipDeviceGetAttribute(&c->clock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->clock_khz = 0;
    snprintf(c->name, sizeof(c->name), "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    if (mcml_scratch_address(&c->scratch) != hipSuccess || c->scratch == nullptr ||
        hipHostMalloc(&c->na_stage, MCML_NOALLOC_OFF_ACC - MCML_NOALLOC_OFF_SETS, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&c->ctrl, sizeof(unsigned long long) * kCtrlWords) != hipSuccess ||
        hipHostMalloc((void**)&c->h_ctrl, sizeof(unsigned long long) * kCtrlWords, hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void**)&c->d_mult, sizeof(uint32_t) * MCML_MWC_TABLE_SIZE) != hipSuccess ||
        hipMemcpy(c->d_mult, mcml_mwc_multipliers, sizeof(uint32_t) * MCML_MWC_TABLE_SIZE,
                  hipMemcpyHostToDevice) != hipSuccess) {
        *err = MCML_ERR_DEVICE;
        mcml_destroy(c);
        return nullptr;
    }
    /* the transport's reciprocal and square root must be the correctly rounded ones on this
     * device, or photon trajectories would differ from their CPU restatement */
    uint64_t bad[2] = {0, 0};
    if (mcml_selftest_math(c, bad) != MCML_OK || bad[0] != 0 || bad[1] != 0) {
        *err = MCML_ERR_DEVICE;
        mcml_destroy(c);
        return nullptr;
    }
    *err = MCML_OK;
    return c;
}

extern "C" int mcml_selftest_math(mcml_ctx* c, uint64_t* mismatches)
{
    if (!c || !mismatches) return MCML_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    unsigned long long* bad = c->ctrl + kCtrlFirstPass; /* two words past the first pass's */
    HIP_TRY(c, mcml_launch_reset(bad, 2, c->stream));
    HIP_TRY(c, mcml_launch_selftest(bad, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->h_ctrl + kCtrlFirstPass, bad, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    mismatches[0] = c->h_ctrl[kCtrlFirstPass];
    mismatches[1] = c->h_ctrl[kCtrlFirstPass + 1];
    return MCML_OK;
}

extern "C" void mcml_destroy(mcml_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->sets,  &c->set_begin, &c->rec_key, &c->rec_row, &c->skey,  &c->sidx, &c->stemp,
                      &c->rows,  &c->weights,   &c->precs,   &c->offsets, &c->sums,  &c->rrows, &c->arz,
                      &c->a0z,   &c->wrows,     &c->woff,    &c->wold,    &c->wnew,  &c->wout,
                      &c->ref_state, &c->ref_active, &c->ref_fibers, &c->ref_spill_key, &c->ref_spill_w,
                      &c->rgrp,      &c->rgq,        &c->spill[0],   &c->spill[1], &c->rstage};
    for (DevBuf* b : bufs) release(*b);
    mcml_graph_release(c->graph);
    for (int i = 0; i < 2; i++) {
        if (c->stage[i]) (void)hipHostFree(c->stage[i]);
        if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
    }
    if (c->stage_small) (void)hipHostFree(c->stage_small);
    if (c->na_stage) (void)hipHostFree(c->na_stage);
    if (c->d_mult) (void)hipFree(c->d_mult);
    if (c->ctrl) (void)hipFree(c->ctrl);
    if (c->h_ctrl) (void)hipHostFree(c->h_ctrl);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* mcml_last_error(const mcml_ctx* c) { return c ? c->err : "no context"; }
extern "C" const char* mcml_device_name(const mcml_ctx* c) { return c ? c->name : ""; }

extern "C" void mcml_result_free(mcml_result* r)
{
    if (!r) return;
    free(r->reflectance);
    free(r->detected);
    free(r->records);
    free(r->replay_reflectance);
    free(r->pl_offset);
    free(r->pathlength);
    free(r->A_rz);
    free(r->A0_z);
    memset(r, 0, sizeof(*r));
}

/* ====================================================================== runs */
namespace {

/* Reports why mcml_derive rejected set j and returns its code. */
int derive_error(mcml_ctx* c, int j, int rc)
{
    if (rc == MCML_ERR_INPUT)
        set_err(c, "set %d: a layer thickness is negative or NaN", j);
    else
        set_err(c, "set %d: layers must be 1..%d and SDS 1..%d", j, MCML_MAX_LAYERS, MCML_MAX_SDS);
    return rc;
}

int validate_batch(mcml_ctx* c, const mcml_sim* sims, int nsets, std::vector<mcml_derived>& der)
{
    if (nsets < 1) {
        set_err(c, "nsets must be >= 1");
        return MCML_ERR_ARG;
    }
    if (nsets >= MCML_MAX_SETS) {
        set_err(c, "too many parameter sets (%d, at most %d)", nsets, MCML_MAX_SETS - 1);
        return MCML_ERR_LIMIT;
    }
    der.resize((size_t)nsets);
    for (int j = 0; j < nsets; j++) {
        int rc = mcml_derive(&sims[j], &der[(size_t)j]);
        if (rc != MCML_OK) return derive_error(c, j, rc);
        if (der[(size_t)j].num_layers != der[0].num_layers || der[(size_t)j].num_sds != der[0].num_sds) {
            set_err(c, "set %d: all sets of a batch must share number_layers and num_SDS", j);
            return MCML_ERR_ARG;
        }
        if (sims[j].number_photons >= (1ull << 40)) {
            set_err(c, "set %d: number_photons must be below 2^40", j);
            return MCML_ERR_LIMIT;
        }
    }
    return MCML_OK;
}

/* Staging copies of the grids' corner for `groups` (set, SDS) groups (ReplayArgs.stage): as
 * many copies as fit kReplayStageCopies and kReplayStageBytes, cleared by a kernel.  A drop
 * in the corner whose wavefront cannot add it to an LDS tile goes to its wavefront's copy
 * rather than to the grids, whose few hottest cells every replay hits: atomics serialised
 * there (half of the HW8_A replay).  *k = 0 (no copies) when fewer than two fit. */
int stage_copies(mcml_ctx* c, size_t groups, unsigned long long** stage, int32_t* k)
{
    *stage = nullptr;
    *k = 0;
    const size_t per_copy = groups * mcml_replay_stage_words() * sizeof(unsigned long long);
    const size_t n = std::min<size_t>(kReplayStageCopies, kReplayStageBytes / std::max<size_t>(per_copy, 1));
    if (n < 2) return MCML_OK;
    int rc;
    if ((rc = ensure(c, c->rstage, n * per_copy))) return rc;
    HIP_TRY(c, mcml_launch_clear(as<unsigned long long>(c->rstage), n * per_copy / 8, c->stream));
    *stage = as<unsigned long long>(c->rstage);
    *k = (int32_t)n;
    return MCML_OK;
}

/* Absorbance grids the first pass books its replays into (-R -A). */
struct GridSpec {
    unsigned long long* A_rz;
    unsigned long long* A0_z;
    int nr, nz;
    float rdr, rdz;
};

/* First pass: leaves n_rec records in c->rec_key / c->rec_row (row_stride floats each). */
int first_pass(mcml_ctx* c, const std::vector<mcml_derived>& der, const mcml_opts& o, uint64_t seed, int row_stride,
               mcml_result* out, uint64_t& n_rec, const GridSpec* grids = nullptr, bool* grids_done = nullptr)
{
    const int nsets = (int)der.size();
    std::vector<DevSet> hs((size_t)nsets);
    std::vector<uint64_t> begin((size_t)nsets + 1, 0);
    for (int j = 0; j < nsets; j++) {
        fill_devset(der[(size_t)j], hs[(size_t)j]);
        hs[(size_t)j].ticket_begin = begin[(size_t)j];
        begin[(size_t)j + 1] = begin[(size_t)j] + der[(size_t)j].number_photons;
    }
    const uint64_t total = begin[(size_t)nsets];
    int rc;
    if ((rc = ensure(c, c->sets, sizeof(DevSet) * (size_t)nsets)) ||
        (rc = ensure(c, c->set_begin, sizeof(uint64_t) * ((size_t)nsets + 1))))
        return rc;
    HIP_TRY(c, hipMemcpyAsync(c->sets.p, hs.data(), sizeof(DevSet) * (size_t)nsets, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->set_begin.p, begin.data(), sizeof(uint64_t) * ((size_t)nsets + 1),
                              hipMemcpyHostToDevice, c->stream));

    const int block = o.block_size > 0 ? o.block_size : 256;
    if (block % 64 != 0 || block > 256) {
        set_err(c, "block_size must be a multiple of 64 and at most 256");
        return MCML_ERR_ARG;
    }
    FirstPassArgs a;
    memset(&a, 0, sizeof(a));
    a.sets = as<DevSet>(c->sets);
    a.set_begin = as<uint64_t>(c->set_begin);
    a.nsets = nsets;
    a.num_layers = der[0].num_layers;
    a.row_stride = row_stride;
    /* a small set table is staged in LDS by every workgroup; a large batch reads it from
     * global memory (it is only read at launch, layer changes and boundary hits) */
    a.lds_sets = sizeof(DevSet) * (size_t)nsets <= kLdsTableMax ? 1 : 0;
    for (int j = 0; j < nsets && !a.any_iso; j++) a.any_iso = devset_iso(hs[(size_t)j], der[0].num_layers);
    int occ = 0;
    HIP_TRY(c, mcml_occupancy_first_pass(a, block, &occ));
    if (occ < 1) occ = 1;
    const int bpc = o.blocks_per_cu > 0 ? std::min(o.blocks_per_cu, occ) : occ;
    const int grid = c->cu_count * bpc;
    a.chunk = o.chunk > 0 ? std::max(o.chunk, 64) : kDefaultChunk;
    a.refill_min = o.refill_min > 0 ? std::min(o.refill_min, 64) : kDefaultRefill;
    a.bnd_min = o.boundary_min > 0 ? std::min(o.boundary_min, 64) : kDefaultBoundary;
    a.total = total;
    a.seed = seed;
    a.mult = c->d_mult;
    a.queue = c->ctrl + kCtrlQueue;
    a.n_records = c->ctrl + kCtrlRecords;
    a.counters = c->ctrl + kCtrlCounters;
    a.stats = o.lane_stats ? c->ctrl + kCtrlStats : nullptr;
    /* drain compaction: spill arrays hold one photon per lane of the launch */
    const int spill_below = o.spill_below < 0 ? 0 : (o.spill_below > 0 ? std::min(o.spill_below, 63) : kDefaultSpillBelow);
    const uint64_t spill_cap = (uint64_t)grid * (uint64_t)block;
    const size_t spill_bytes = sizeof(float) * (size_t)SPILL_FIELDS(a.num_layers) * (size_t)spill_cap;
    if (spill_below > 0) {
        if ((rc = ensure(c, c->spill[0], spill_bytes)) || (rc = ensure(c, c->spill[1], spill_bytes))) return rc;
        a.spill = as<float>(c->spill[0]);
        a.n_spill = c->ctrl + kCtrlSpill;
        a.spill_cap = spill_cap;
        a.spill_below = spill_below;
    }

    /* Record capacity.  Each photon yields at most max_overlap records, so `bound` records
     * can never overflow; it is used whenever it fits in a quarter of the free device memory
     * (or the caller's record_cap is smaller).  A smaller buffer makes the kernel pause:
     * once `stop_at` records exist it claims no new tickets, and cap - stop_at covers every
     * photon that can still run then -- one per lane, the private ranges of the wavefronts
     * and one more chunk per wavefront claimed as the limit was crossed.  The buffer then
     * grows and the pass resumes from the ticket counter.  Records never depend on it:
     * each photon's stream is fixed. */
    uint64_t bound = 0;
    int overlap = 1;
    for (const mcml_derived& d : der) {
        const int ov = max_overlap(d);
        overlap = std::max(overlap, ov);
        bound += d.number_photons * (uint64_t)ov;
    }
    const uint64_t lanes = (uint64_t)grid * (uint64_t)block, waves = lanes / 64;
    const uint64_t reserve = (lanes + 2 * waves * (uint64_t)a.chunk) * (uint64_t)overlap;
    const uint64_t per_rec = sizeof(uint64_t) + sizeof(float) * (uint64_t)row_stride;
    uint64_t cap;
    if (o.record_cap > 0) {
        cap = o.record_cap;
    } else {
        size_t mem_free = 0, mem_total = 0;
        HIP_TRY(c, hipMemGetInfo(&mem_free, &mem_total));
        const uint64_t held = std::min(c->rec_key.bytes / sizeof(uint64_t),
                                       c->rec_row.bytes / (sizeof(float) * (size_t)row_stride));
        cap = std::max<uint64_t>({(uint64_t)(mem_free / 4) / per_rec, held, 1ull << 20});
    }
    cap = std::max<uint64_t>(cap, reserve + 4096);
    if (cap >= bound) cap = std::max<uint64_t>(bound, 1);
    unsigned long long cnt[CNT_N] = {0, 0, 0, 0};
    int launches = 0;
    n_rec = 0;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream)); /* the pass: every launch and the gaps between them */
    for (;;) {
        if ((rc = ensure_keep(c, c->rec_key, sizeof(uint64_t) * (size_t)cap, sizeof(uint64_t) * (size_t)n_rec)) ||
            (rc = ensure_keep(c, c->rec_row, sizeof(float) * (size_t)cap * (size_t)row_stride,
                              sizeof(float) * (size_t)n_rec * (size_t)row_stride)))
            return rc;
        a.rec_key = as<uint64_t>(c->rec_key);
        a.rec_row = as<float>(c->rec_row);
        a.cap = cap;
        a.stop_at = cap >= bound ? ~0ull : cap - reserve;
        /* a pass that may pause for records does not spill: its spill counter would carry
         * photons across the pause */
        if (a.stop_at != ~0ull) a.spill = nullptr;
        if (launches == 0 && o.use_graph && !o.lane_stats) {
            HIP_TRY(c, mcml_graph_first_pass(c->graph, a, grid, block, c->ctrl, kCtrlFirstPass, c->h_ctrl, c->stream));
        } else {
            if (launches == 0) HIP_TRY(c, mcml_launch_reset(c->ctrl, kCtrlFirstPass, c->stream));
            HIP_TRY(c, mcml_launch_first_pass(a, grid, block, c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->h_ctrl, c->ctrl, sizeof(unsigned long long) * kCtrlFirstPass,
                                      hipMemcpyDeviceToHost, c->stream));
        }
        launches++;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        const uint64_t claimed = c->h_ctrl[kCtrlQueue];
        n_rec = c->h_ctrl[kCtrlRecords];
        if (n_rec > cap) {
            set_err(c, "first pass appended %llu records into a buffer of %llu", (unsigned long long)n_rec,
                    (unsigned long long)cap);
            return MCML_ERR_DEVICE;
        }
        if (claimed >= total) break;
        if (claimed == 0 || launches > 64) {
            set_err(c, "first pass made no progress (%llu of %llu tickets after %d launches)",
                    (unsigned long long)claimed, (unsigned long long)total, launches);
            return MCML_ERR_DEVICE;
        }
        /* paused: grow toward the records the remaining tickets are expected to add */
        const double est = (double)n_rec * (double)total / (double)claimed;
        uint64_t next = std::max<uint64_t>(2 * cap, (uint64_t)(1.25 * est)) + reserve;
        cap = next >= bound ? std::max<uint64_t>(bound, 1) : next;
    }
    /* the drain: photons spilled by sparse wavefronts continue in compacted launches; each
     * spills again while more than kFinalDrain photons are left, the last runs them out.
     * With absorbance grids the drain launches are mcml_first_pass_grids: the first also
     * replays every detection of the main launch into the grids, and each replays its own
     * detections as they happen (a replay in progress spills and continues like any photon),
     * so that the replays of a few long trajectories overlap the drain rather than follow it.
     * Many detections (more than the launch has lanes) are left to the replay pass, whose
     * LDS tiles absorb their drops. */
    uint64_t n_sp = a.spill != nullptr ? c->h_ctrl[kCtrlSpill] : 0;
    int drains = 0, from = 0;
    if (grids) *grids_done = n_rec <= (uint64_t)grid * (uint64_t)block;
    FirstPassArgs d = a; /* what every drain launch shares */
    d.resume = 1;
    d.resume_cap = spill_cap;
    d.queue = c->ctrl + kCtrlResume;
    d.chunk = 64;
    d.stop_at = ~0ull; /* the record reserve covers every photon that was in flight */
    const bool grid_drains = grids && *grids_done;
    if (grid_drains) {
        d.A_rz = grids->A_rz;
        d.A0_z = grids->A0_z;
        d.nr = grids->nr;
        d.nz = grids->nz;
        d.rdr = grids->rdr;
        d.rdz = grids->rdz;
        d.replay_missed = c->ctrl + kCtrlMissed;
        d.stage_groups = (int32_t)(der.size() * (size_t)der[0].num_sds);
        if ((rc = stage_copies(c, (size_t)d.stage_groups, &d.stage, &d.stage_k))) return rc;
    }
    uint64_t n_jobs = grid_drains ? n_rec : 0; /* the main launch's detections, replayed by the first drain */
    while (n_sp + n_jobs > 0) {
        const uint64_t total_d = n_sp + n_jobs;
        if (n_sp > spill_cap || drains > 64) {
            set_err(c, "drain: %llu spilled photons for %llu slots after %d launches", (unsigned long long)n_sp,
                    (unsigned long long)spill_cap, drains);
            return MCML_ERR_DEVICE;
        }
        FirstPassArgs r = d;
        r.resume_buf = as<float>(c->spill[from]);
        r.resume_photons = n_sp; /* tickets past the spilled photons are replay jobs */
        r.replay_keys = as<uint64_t>(c->rec_key);
        r.total = total_d;
        const bool again = a.spill != nullptr && total_d > kFinalDrain;
        r.spill = again ? as<float>(c->spill[from ^ 1]) : nullptr;
        const int rgrid = (int)std::min<uint64_t>((uint64_t)grid, (total_d + (uint64_t)block - 1) / (uint64_t)block);
        HIP_TRY(c, mcml_launch_clear(c->ctrl + kCtrlSpill, 2, c->stream)); /* spill and resume counters */
        HIP_TRY(c, mcml_launch_first_pass(r, rgrid, block, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->h_ctrl, c->ctrl, sizeof(unsigned long long) * kCtrlFirstPass,
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        drains++;
        n_rec = c->h_ctrl[kCtrlRecords];
        if (n_rec > cap || c->h_ctrl[kCtrlResume] < total_d) {
            set_err(c, "drain launch %d: %llu records for %llu slots, %llu of %llu tickets taken", drains,
                    (unsigned long long)n_rec, (unsigned long long)cap, (unsigned long long)c->h_ctrl[kCtrlResume],
                    (unsigned long long)total_d);
            return MCML_ERR_DEVICE;
        }
        n_sp = again ? c->h_ctrl[kCtrlSpill] : 0;
        n_jobs = 0;
        from ^= 1;
    }
    if (grid_drains && d.stage != nullptr)
        HIP_TRY(c, mcml_launch_stage_reduce(d.stage, d.stage_k, d.stage_groups, d.A_rz, d.A0_z, d.nz, d.nr, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    const double ms_kernel = ms_between(c->ev0, c->ev1);
    if (grids && *grids_done && c->h_ctrl[kCtrlMissed] != 0) {
        set_err(c, "%llu replays ended without the detection they replay", (unsigned long long)c->h_ctrl[kCtrlMissed]);
        return MCML_ERR_DEVICE;
    }
    for (int i = 0; i < CNT_N; i++) cnt[i] = c->h_ctrl[kCtrlCounters + i];
    if (o.lane_stats) {
        const unsigned long long* st = c->h_ctrl + kCtrlStats;
        for (int i = 0; i < 4; i++) out->lane_blocks[i] = st[STAT_STEP + i];
        const unsigned long long t0 = ~st[STAT_T_START], td = ~st[STAT_T_DRY], te = st[STAT_T_END];
        if (c->clock_khz > 0 && te >= t0) {
            out->ms_clock = (double)(te - t0) / c->clock_khz;
            out->ms_drain = te >= td ? (double)(te - td) / c->clock_khz : 0.0;
        }
    }
    out->drain_launches = drains;
    if (n_rec >= (1ull << 32)) {
        set_err(c, "more than 2^32 detection records");
        return MCML_ERR_LIMIT;
    }
    out->photons = cnt[CNT_PHOTONS];
    out->steps = cnt[CNT_STEPS];
    out->steps_boundary = cnt[CNT_BOUNDARY];
    out->steps_interaction = cnt[CNT_INTERACTION];
    out->ms_first_pass_kernel = ms_kernel;
    out->first_pass_launches = launches;
    out->block_size = block;
    out->grid_size = grid;
    out->chunk = a.chunk;
    out->refill_min = a.refill_min;
    out->boundary_min = a.bnd_min;
    if (out->photons != total) {
        set_err(c, "first pass launched %llu photons, expected %llu", (unsigned long long)out->photons,
                (unsigned long long)total);
        return MCML_ERR_DEVICE;
    }
    return MCML_OK;
}

/* First pass in the reference's launch structure (transport.cu:76-147, io.cpp:487-526):
 * T = maxThreadsPerMultiProcessor x CUs threads in workgroups of `block` (the reference
 * picks 32), one photon per thread, relaunched every MCML_REF_STEPS steps (o.chunk
 * when set, so that tests can force many relaunches) until no thread
 * is active; after each launch the host clears and uploads the per-thread detection
 * buffers, reads back the active flags and the buffers and reduces them (fiber.cu:3-13,
 * transport.cu:299-311).  The records end in c->rec_key / c->rec_row like first_pass's;
 * host_count and host_sum hold the per-detector counts and weight sums of the host
 * reduction. */
int first_pass_refshape(mcml_ctx* c, const mcml_derived& d, const mcml_opts& o, uint64_t seed, mcml_result* out,
                        uint64_t& n_rec, std::vector<uint64_t>& host_count, std::vector<double>& host_sum)
{
    const int S = d.num_sds;
    DevSet hs;
    fill_devset(d, hs);
    hs.ticket_begin = 0;
    const uint64_t total = d.number_photons;
    int rc;
    if ((rc = ensure(c, c->sets, sizeof(DevSet)))) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->sets.p, &hs, sizeof(DevSet), hipMemcpyHostToDevice, c->stream));
    const int block = o.block_size > 0 ? o.block_size : 32;
    if (block % 32 != 0 || block > 256) {
        set_err(c, "block_size must be a multiple of 32 and at most 256 in the reference launch shape");
        return MCML_ERR_ARG;
    }
    int max_mp = 0;
    HIP_TRY(c, hipDeviceGetAttribute(&max_mp, hipDeviceAttributeMaxThreadsPerMultiProcessor, c->device));
    const uint64_t blocks = (uint64_t)std::max(1, max_mp / block) * (uint64_t)c->cu_count;
    const uint64_t T = blocks * (uint64_t)block;
    if ((rc = ensure(c, c->ref_state, mcml_refshape_slot_bytes() * (size_t)T)) ||
        (rc = ensure(c, c->ref_active, sizeof(uint32_t) * (size_t)T)) ||
        (rc = ensure(c, c->ref_fibers, sizeof(RefFibers) * (size_t)T)))
        return rc;
    /* detections beyond a thread's MCML_REF_SLOTS in one launch: kept (the reference drops
     * them, transport.cu:899-904), so the records equal the persistent kernel's */
    uint64_t spill_cap = std::min<uint64_t>(total * (uint64_t)S, std::max<uint64_t>(1ull << 16, total / 64 + 1024));
    std::vector<RefFibers> hf((size_t)T); /* pageable, as HostMem.f (transport.cu:1059) */
    std::vector<uint32_t> hact((size_t)T);
    std::vector<uint64_t> hkey;
    std::vector<float> hw;
    std::vector<double>& hsum = host_sum;
    hsum.assign((size_t)S, 0.0);
    double ms_kernel = 0;
    int launches = 0;
    uint64_t n_spill = 0;
    for (;;) {
        if ((rc = ensure(c, c->ref_spill_key, sizeof(uint64_t) * (size_t)spill_cap)) ||
            (rc = ensure(c, c->ref_spill_w, sizeof(float) * (size_t)spill_cap)))
            return rc;
        RefShapeArgs a;
        memset(&a, 0, sizeof(a));
        a.set = as<DevSet>(c->sets);
        a.total = total;
        a.threads = T;
        a.seed = seed;
        a.mult = c->d_mult;
        a.state = c->ref_state.p;
        a.active = as<uint32_t>(c->ref_active);
        a.fibers = as<RefFibers>(c->ref_fibers);
        a.terminated = c->ctrl + 0;
        a.n_spill = c->ctrl + 1;
        a.counters = c->ctrl + 2;
        a.spill_key = as<uint64_t>(c->ref_spill_key);
        a.spill_w = as<float>(c->ref_spill_w);
        a.spill_cap = spill_cap;
        a.steps = o.chunk > 0 ? o.chunk : MCML_REF_STEPS;
        hkey.clear();
        hw.clear();
        std::fill(hsum.begin(), hsum.end(), 0.0);
        host_count.assign((size_t)S, 0);
        ms_kernel = 0;
        launches = 0;
        HIP_TRY(c, mcml_launch_reset(c->ctrl, kCtrlFirstPass, c->stream));
        HIP_TRY(c, mcml_launch_refshape_init(a, block, c->stream));
        uint64_t active_total = 1;
        while (active_total > 0) {
            memset(hf.data(), 0, sizeof(RefFibers) * hf.size()); /* fiber_initialization */
            HIP_TRY(c, hipMemcpyAsync(a.fibers, hf.data(), sizeof(RefFibers) * hf.size(), hipMemcpyHostToDevice,
                                      c->stream));
            HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
            HIP_TRY(c, mcml_launch_refshape_pass(a, block, c->stream));
            HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
            HIP_TRY(c, hipMemcpyAsync(hact.data(), a.active, sizeof(uint32_t) * hact.size(), hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            ms_kernel += ms_between(c->ev0, c->ev1);
            launches++;
            active_total = 0;
            for (uint32_t v : hact) active_total += v;
            HIP_TRY(c, hipMemcpyAsync(hf.data(), a.fibers, sizeof(RefFibers) * hf.size(), hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            for (const RefFibers& f : hf) /* calculate_reflectance */
                for (int k = 0; k < f.count; k++) {
                    const int s = f.sds[k];
                    if (s < 1 || s > S) {
                        set_err(c, "reference launch shape: detector index %d out of range", s);
                        return MCML_ERR_DEVICE;
                    }
                    hsum[(size_t)s - 1] += (double)f.w[k];
                    host_count[(size_t)s - 1]++;
                    hkey.push_back(f.key[k]);
                    hw.push_back(f.w[k]);
                }
        }
        HIP_TRY(c, hipMemcpyAsync(c->h_ctrl, c->ctrl, sizeof(unsigned long long) * kCtrlFirstPass,
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        n_spill = c->h_ctrl[1];
        if (n_spill <= spill_cap) break;
        spill_cap = n_spill + n_spill / 8 + 1024; /* rerun: each photon's stream is fixed */
    }
    if (c->h_ctrl[0] != total) {
        set_err(c, "reference launch shape ended %llu photons, expected %llu", (unsigned long long)c->h_ctrl[0],
                (unsigned long long)total);
        return MCML_ERR_DEVICE;
    }
    n_rec = n_spill + (uint64_t)hkey.size();
    if (n_rec >= (1ull << 32)) {
        set_err(c, "more than 2^32 detection records");
        return MCML_ERR_LIMIT;
    }
    if ((rc = ensure(c, c->rec_key, sizeof(uint64_t) * (size_t)n_rec)) ||
        (rc = ensure(c, c->rec_row, sizeof(float) * (size_t)n_rec)))
        return rc;
    if (n_spill) {
        HIP_TRY(c, hipMemcpyAsync(c->rec_key.p, c->ref_spill_key.p, sizeof(uint64_t) * (size_t)n_spill,
                                  hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->rec_row.p, c->ref_spill_w.p, sizeof(float) * (size_t)n_spill,
                                  hipMemcpyDeviceToDevice, c->stream));
    }
    if (!hkey.empty()) {
        HIP_TRY(c, hipMemcpyAsync(as<uint64_t>(c->rec_key) + n_spill, hkey.data(), sizeof(uint64_t) * hkey.size(),
                                  hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(as<float>(c->rec_row) + n_spill, hw.data(), sizeof(float) * hw.size(),
                                  hipMemcpyHostToDevice, c->stream));
    }
    if (n_spill) { /* spilled records count toward the per-detector totals as well */
        std::vector<uint64_t> sk((size_t)n_spill);
        HIP_TRY(c, hipMemcpyAsync(sk.data(), c->ref_spill_key.p, sizeof(uint64_t) * (size_t)n_spill,
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        std::vector<float> sw((size_t)n_spill);
        HIP_TRY(c, hipMemcpyAsync(sw.data(), c->ref_spill_w.p, sizeof(float) * (size_t)n_spill, hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < sk.size(); i++) {
            const uint64_t s = (sk[i] >> MCML_KEY_SDS_SHIFT) & 31u;
            if (s >= 1 && s <= (uint64_t)S) {
                host_count[(size_t)s - 1]++;
                hsum[(size_t)s - 1] += (double)sw[i];
            }
        }
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    out->photons = std::min<uint64_t>(T, total) + c->h_ctrl[2 + CNT_PHOTONS];
    out->steps = c->h_ctrl[2 + CNT_STEPS];
    out->steps_boundary = c->h_ctrl[2 + CNT_BOUNDARY];
    out->steps_interaction = c->h_ctrl[2 + CNT_INTERACTION];
    out->ms_first_pass_kernel = ms_kernel;
    out->first_pass_launches = launches;
    out->block_size = block;
    out->grid_size = (int)blocks;
    if (out->photons != total) {
        set_err(c, "reference launch shape launched %llu photons, expected %llu", (unsigned long long)out->photons,
                (unsigned long long)total);
        return MCML_ERR_DEVICE;
    }
    return MCML_OK;
}

/* Sort the records by (set, SDS, photon) on the device. */
int sort_records(mcml_ctx* c, uint64_t n)
{
    int rc;
    if ((rc = ensure(c, c->skey, sizeof(uint64_t) * (size_t)n)) ||
        (rc = ensure(c, c->sidx, sizeof(uint32_t) * (size_t)n)))
        return rc;
    if (n == 0) return MCML_OK;
    size_t temp = 0;
    HIP_TRY(c, mcml_sort_keys(nullptr, temp, as<uint64_t>(c->rec_key), as<uint64_t>(c->skey), as<uint32_t>(c->sidx), n,
                              c->stream));
    if ((rc = ensure(c, c->stemp, temp))) return rc;
    HIP_TRY(c, mcml_sort_keys(c->stemp.p, temp, as<uint64_t>(c->rec_key), as<uint64_t>(c->skey), as<uint32_t>(c->sidx),
                              n, c->stream));
    return MCML_OK;
}

/* Gather rows/weights into sorted order (the dump), build public records, and reduce per
 * (set, SDS) group; sums and offsets land in host arrays. */
int gather_and_reduce(mcml_ctx* c, uint64_t n, int nsets, int S, const float* rows_in, int stride_in,
                      const uint32_t* order, int stride_out, bool records, double* sums_host,
                      std::vector<uint64_t>& offsets_host, double* dump_ms)
{
    int rc;
    const size_t groups = (size_t)nsets * (size_t)S;
    if ((rc = ensure(c, c->weights, sizeof(float) * (size_t)n)) ||
        (rc = ensure(c, c->offsets, sizeof(uint64_t) * (groups + 1))) ||
        (rc = ensure(c, c->sums, sizeof(double) * groups)))
        return rc;
    if (stride_out > 0 && (rc = ensure(c, c->rows, sizeof(float) * (size_t)n * (size_t)stride_out))) return rc;
    if (records && (rc = ensure(c, c->precs, sizeof(mcml_record) * (size_t)n))) return rc;
    if (n > 0) {
        GatherArgs g;
        memset(&g, 0, sizeof(g));
        g.sorted_keys = as<uint64_t>(c->skey);
        g.order = order;
        g.rows_in = rows_in;
        g.stride_in = stride_in;
        g.stride_out = stride_out;
        g.n = n;
        g.rows_out = stride_out > 0 ? as<float>(c->rows) : nullptr;
        g.weight_out = as<float>(c->weights);
        g.records_out = records ? as<mcml_record>(c->precs) : nullptr;
        /* one 256-record tile per workgroup pass; the grid covers the chip several times over */
        const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)c->cu_count * 16));
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        HIP_TRY(c, mcml_launch_dump(g, (int)grid, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    }
    HIP_TRY(c, mcml_launch_groups(as<uint64_t>(c->skey), n, nsets, S, as<uint64_t>(c->offsets), as<float>(c->weights),
                                  as<double>(c->sums), c->stream));
    offsets_host.resize(groups + 1);
    HIP_TRY(c, hipMemcpyAsync(offsets_host.data(), c->offsets.p, sizeof(uint64_t) * (groups + 1), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipMemcpyAsync(sums_host, c->sums.p, sizeof(double) * groups, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (dump_ms && n > 0) *dump_ms = ms_between(c->ev0, c->ev1);
    return MCML_OK;
}

/* The replay of every detected photon (MCd_replay, transport.cu:473-622) in a pass of its
 * own: jobs are the sorted record keys, grouped by (set, SDS) at `off`; rows go to c->rrows,
 * drops to the grids of gs (none when gs.A_rz is NULL).  Buffers rrows, rgrp and rgq must
 * hold n rows and the groups. */
int replay_pass(mcml_ctx* c, const std::vector<mcml_derived>& der, const mcml_opts& o, uint64_t seed, uint64_t n,
                const std::vector<uint64_t>& off, const GridSpec& gs, mcml_result* out)
{
    const int S = der[0].num_sds;
    const size_t groups = der.size() * (size_t)S;
    HIP_TRY(c, hipMemcpyAsync(c->rgrp.p, off.data(), sizeof(uint64_t) * (groups + 1), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipMemsetAsync(c->rgq.p, 0, sizeof(unsigned long long) * groups, c->stream));
    ReplayArgs a;
    memset(&a, 0, sizeof(a));
    a.sets = as<DevSet>(c->sets);
    a.keys = as<uint64_t>(c->skey);
    a.njobs = n;
    a.seed = seed;
    a.mult = c->d_mult;
    a.grp_begin = as<uint64_t>(c->rgrp);
    a.grp_queue = as<unsigned long long>(c->rgq);
    a.ngroups = (int32_t)groups;
    a.chunk = 64;
    a.num_layers = der[0].num_layers;
    a.bnd_min = o.boundary_min > 0 ? std::min(o.boundary_min, 64) : kDefaultBoundary;
    a.rows = as<float>(c->rrows);
    a.A_rz = gs.A_rz;
    a.A0_z = gs.A0_z;
    a.num_sds = S;
    a.nr = gs.nr;
    a.nz = gs.nz;
    a.rdr = gs.rdr;
    a.rdz = gs.rdz;
    a.missed = c->ctrl + kCtrlMissed;
    if (gs.A_rz != nullptr) { /* corner drops of a group other than the wavefront's tile's */
        const int rc = stage_copies(c, groups, &a.stage, &a.stage_k);
        if (rc) return rc;
    }
    HIP_TRY(c, mcml_launch_reset(c->ctrl + kCtrlMissed, 1, c->stream));
    int occ = 0;
    HIP_TRY(c, mcml_occupancy_replay(&occ));
    if (occ < 1) occ = 1;
    const uint64_t per_block = (uint64_t)mcml_replay_block();
    const uint64_t grid =
        std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cu_count * occ, (n + per_block - 1) / per_block));
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, mcml_launch_replay(a, (int)grid, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->h_ctrl + kCtrlMissed, c->ctrl + kCtrlMissed, sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    out->ms_replay_kernel = ms_between(c->ev0, c->ev1);
    if (c->h_ctrl[kCtrlMissed] != 0) {
        set_err(c, "%llu replays ended without the detection they replay", (unsigned long long)c->h_ctrl[kCtrlMissed]);
        return MCML_ERR_DEVICE;
    }
    return MCML_OK;
}

} // namespace

int mcml_run_batch_sink(mcml_ctx* c, const mcml_sim* sims, int nsets, const mcml_opts* opts, uint64_t seed,
                        mcml_result* out, const mcml_readback_sink* sink)
{
    if (!c || !sims || !out) return MCML_ERR_ARG;
    memset(out, 0, sizeof(*out));
    mcml_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if ((o.pathlength || o.absorbance || o.avg_pathlength) && !o.replay) {
        set_err(c, "-P, -A and -AP need -R (main.cu:77-88)");
        return MCML_ERR_ARG;
    }
    std::vector<mcml_derived> der;
    int rc = validate_batch(c, sims, nsets, der);
    if (rc) return rc;
    if (o.reference_shape && nsets != 1) {
        set_err(c, "the reference launch shape runs one parameter set at a time");
        return MCML_ERR_ARG;
    }
    if (o.replay && o.absorbance)
        for (int j = 1; j < nsets; j++)
            if (der[(size_t)j].grid_nr != der[0].grid_nr || der[(size_t)j].grid_nz != der[0].grid_nz ||
                der[(size_t)j].grid_dr != der[0].grid_dr || der[(size_t)j].grid_dz != der[0].grid_dz) {
                set_err(c, "set %d: all sets of a batch must share the absorbance grid (-A)", j);
                return MCML_ERR_ARG;
            }
    HIP_TRY(c, hipSetDevice(c->device));
    const int S = der[0].num_sds, L = der[0].num_layers;
    const int nr = der[0].grid_nr, nz = der[0].grid_nz;
    out->nsets = nsets;
    out->num_sds = S;
    out->num_layers = L;
    out->grid_nr = nr;
    out->grid_nz = nz;
    const size_t groups = (size_t)nsets * (size_t)S;
    out->reflectance = (double*)calloc(groups, sizeof(double));
    out->detected = (uint64_t*)calloc(groups, sizeof(uint64_t));
    if (!out->reflectance || !out->detected) {
        mcml_result_free(out);
        return MCML_ERR_NOMEM;
    }
    /* Replay: record the pathlengths in the first pass itself, and with -A let each lane
     * replay its detected photons into the grids within the same pass.  The reference launch
     * shape (and the lane-statistics variant with -A) replays as a separate pass, as the
     * reference does. */
    const bool grids_inline = o.replay && o.absorbance && !o.reference_shape && !o.lane_stats && !o.separate_replay &&
                              nr < 65536 && nz < 65536 && groups * (size_t)nz * (size_t)nr * 8 < (1ull << 32);
    const bool inline_pl = o.replay && (!o.absorbance || grids_inline) && !o.reference_shape;
    const int stride = inline_pl ? L + 2 : 1;
    const size_t rz_cells = o.absorbance ? groups * (size_t)nz * (size_t)nr : 0;
    const size_t z_cells = o.absorbance ? groups * (size_t)nz : 0;
    const double t0 = now_ms();
    GridSpec gs = {nullptr, nullptr, nr, nz, 1.0f / der[0].grid_dr, 1.0f / der[0].grid_dz};
    if (grids_inline) {
        if ((rc = ensure(c, c->arz, sizeof(unsigned long long) * rz_cells)) ||
            (rc = ensure(c, c->a0z, sizeof(unsigned long long) * z_cells))) {
            mcml_result_free(out);
            return rc;
        }
        HIP_TRY(c, hipMemsetAsync(c->arz.p, 0, sizeof(unsigned long long) * rz_cells, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->a0z.p, 0, sizeof(unsigned long long) * z_cells, c->stream));
        gs.A_rz = as<unsigned long long>(c->arz);
        gs.A0_z = as<unsigned long long>(c->a0z);
    }
    uint64_t n = 0;
    std::vector<uint64_t> off, host_count;
    std::vector<double> host_sum;
    double dump_ms = 0;
    bool grids_done = false; /* the first pass replayed into the grids itself */
    rc = o.reference_shape ? first_pass_refshape(c, der[0], o, seed, out, n, host_count, host_sum)
                           : first_pass(c, der, o, seed, stride, out, n, grids_inline ? &gs : nullptr, &grids_done);
    if (rc || (rc = sort_records(c, n)) ||
        (rc = gather_and_reduce(c, n, nsets, S, as<float>(c->rec_row), stride, as<uint32_t>(c->sidx),
                                inline_pl ? L + 2 : 0, !o.skip_records, out->reflectance, off, &dump_ms))) {
        mcml_result_free(out);
        return rc;
    }
    for (size_t g = 0; g < groups; g++) out->detected[g] = off[g + 1] - off[g];
    if (o.reference_shape) /* the host reduction (calculate_reflectance) agrees with the device's */
        for (size_t g = 0; g < groups; g++)
            if (out->detected[g] != host_count[g] ||
                fabs(out->reflectance[g] - host_sum[g]) > 1e-9 * fabs(out->reflectance[g])) {
                set_err(c, "reference launch shape: SDS %d host reduction %llu records / %.17g, the device %llu / %.17g",
                        (int)g + 1, (unsigned long long)host_count[g], host_sum[g],
                        (unsigned long long)out->detected[g], out->reflectance[g]);
                mcml_result_free(out);
                return MCML_ERR_DEVICE;
            }
    out->n_records = n;
    if (!o.skip_records) {
        out->records = (mcml_record*)host_alloc(sizeof(mcml_record) * (size_t)std::max<uint64_t>(n, 1));
        if (!out->records) {
            mcml_result_free(out);
            return MCML_ERR_NOMEM;
        }
        if (n)
            rc = copy_d2h(c, out->records, c->precs.p, sizeof(mcml_record) * (size_t)n);
        if (rc) {
            mcml_result_free(out);
            return rc;
        }
    }
    const double t1 = now_ms();
    out->ms_first_pass_wall = t1 - t0;
    /* the caller hears of the pathlength rows as they land (whole rows only) */
    std::function<void(size_t)> rows_fn;
    const std::function<void(size_t)>* rows_cb = nullptr;
    bool reported = false;
    if (sink && sink->rows_landed) {
        const size_t row_bytes = sizeof(float) * (size_t)(L + 2);
        rows_fn = [=, &reported](size_t b) {
            reported = true;
            sink->rows_landed(sink->user, out, (uint64_t)(b / row_bytes));
        };
        rows_cb = &rows_fn;
    }
    auto fail = [&](int code) { /* the caller stops reading the rows before they are freed */
        if (reported && sink->release) sink->release(sink->user);
        mcml_result_free(out);
        return code;
    };
    if (o.replay) {
        out->pl_offset = (uint64_t*)malloc(sizeof(uint64_t) * (groups + 1));
        out->replay_reflectance = (double*)calloc(groups, sizeof(double));
        out->pathlength = (float*)host_alloc(sizeof(float) * (size_t)std::max<uint64_t>(n, 1) * (size_t)(L + 2));
        if (!out->pl_offset || !out->replay_reflectance || !out->pathlength) {
            return fail(MCML_ERR_NOMEM);
        }
        memcpy(out->pl_offset, off.data(), sizeof(uint64_t) * (groups + 1));
        const bool grids = o.absorbance != 0;
        if (inline_pl) { /* the first pass already produced the replay rows, in file order */
            memcpy(out->replay_reflectance, out->reflectance, sizeof(double) * groups);
            out->ms_dump_kernel = dump_ms;
            if ((rc = copy_d2h(c, out->pathlength, c->rows.p, sizeof(float) * (size_t)n * (size_t)(L + 2), rows_cb))) {
                return fail(rc);
            }
        }
        if (grids && inline_pl && !grids_done && n > 0) { /* the replay pass books the grids only */
            if ((rc = ensure(c, c->rrows, sizeof(float) * (size_t)n * (size_t)(L + 2))) ||
                (rc = ensure(c, c->rgrp, sizeof(uint64_t) * (groups + 1))) ||
                (rc = ensure(c, c->rgq, sizeof(unsigned long long) * groups))) {
                return fail(rc);
            }
            if ((rc = replay_pass(c, der, o, seed, n, off, gs, out))) {
                return fail(rc);
            }
        } else if (!inline_pl) { /* absorbance maps replayed in a pass of their own (transport.cu:160-222) */
            if ((rc = ensure(c, c->rrows, sizeof(float) * (size_t)n * (size_t)(L + 2))) ||
                (rc = ensure(c, c->arz, sizeof(unsigned long long) * rz_cells)) ||
                (rc = ensure(c, c->a0z, sizeof(unsigned long long) * z_cells))) {
                return fail(rc);
            }
            if (grids) {
                HIP_TRY(c, hipMemsetAsync(c->arz.p, 0, sizeof(unsigned long long) * rz_cells, c->stream));
                HIP_TRY(c, hipMemsetAsync(c->a0z.p, 0, sizeof(unsigned long long) * z_cells, c->stream));
                gs.A_rz = as<unsigned long long>(c->arz);
                gs.A0_z = as<unsigned long long>(c->a0z);
            }
            if (n) {
                if ((rc = ensure(c, c->rgrp, sizeof(uint64_t) * (groups + 1))) ||
                    (rc = ensure(c, c->rgq, sizeof(unsigned long long) * groups)) ||
                    (rc = replay_pass(c, der, o, seed, n, off, gs, out))) {
                    return fail(rc);
                }
            }
            std::vector<uint64_t> off2;
            if ((rc = gather_and_reduce(c, n, nsets, S, as<float>(c->rrows), L + 2, nullptr, L + 2, false,
                                        out->replay_reflectance, off2, &out->ms_dump_kernel))) {
                return fail(rc);
            }
            if ((rc = copy_d2h(c, out->pathlength, c->rows.p, sizeof(float) * (size_t)n * (size_t)(L + 2), rows_cb))) {
                return fail(rc);
            }
        }
        if (grids) {
            out->A_rz = (uint64_t*)host_alloc(sizeof(uint64_t) * std::max<size_t>(rz_cells, 1));
            out->A0_z = (uint64_t*)malloc(sizeof(uint64_t) * std::max<size_t>(z_cells, 1));
            if (!out->A_rz || !out->A0_z) {
                return fail(MCML_ERR_NOMEM);
            }
            if ((rc = copy_d2h(c, out->A_rz, c->arz.p, sizeof(uint64_t) * rz_cells)) ||
                (rc = copy_d2h(c, out->A0_z, c->a0z.p, sizeof(uint64_t) * z_cells))) {
                return fail(rc);
            }
        }
        out->replayed = 1;
        out->ms_replay_wall = now_ms() - t1;
    }
    return MCML_OK;
}

extern "C" int mcml_run_batch(mcml_ctx* c, const mcml_sim* sims, int nsets, const mcml_opts* opts, uint64_t seed,
                              mcml_result* out)
{
    return mcml_run_batch_sink(c, sims, nsets, opts, seed, out, nullptr);
}

extern "C" int mcml_run(mcml_ctx* c, const mcml_sim* sim, const mcml_opts* opts, uint64_t seed, mcml_result* out)
{
    return mcml_run_batch(c, sim, 1, opts, seed, out);
}

/* mcml_run_batch's reflectance and detected arrays with nothing allocated: the set table
 * goes through the context's pinned staging area into this device's scratch array, the
 * transport adds every detection to per-group accumulators there (no records, no sort),
 * and a last kernel writes the caller's device arrays.  The accumulators and the control
 * block are cleared by a kernel at the start of every call (the scratch keeps the previous
 * call's contents). */
extern "C" int mcml_run_batch_sums_noalloc(mcml_ctx* c, const mcml_sim* sims, int nsets, const mcml_opts* opts,
                                           uint64_t seed, double* d_reflectance, uint64_t* d_detected)
{
    if (!c || !sims || !d_reflectance || !d_detected) return MCML_ERR_ARG;
    mcml_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.replay || o.absorbance || o.pathlength || o.avg_pathlength || o.reference_shape || o.use_graph) {
        set_err(c, "the allocation-free first pass returns per-(set, SDS) sums only: replay, absorbance, "
                   "pathlength, reference_shape and use_graph are not available");
        return MCML_ERR_ARG;
    }
    if (nsets < 1) {
        set_err(c, "nsets must be >= 1");
        return MCML_ERR_ARG;
    }
    if (nsets > MCML_NOALLOC_MAX_SETS) {
        set_err(c, "%d parameter sets: the allocation-free call takes at most %d", nsets, MCML_NOALLOC_MAX_SETS);
        return MCML_ERR_LIMIT;
    }
    if (!c->scratch || !c->na_stage) {
        set_err(c, "no scratch array on this device");
        return MCML_ERR_DEVICE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    DevSet* hs = static_cast<DevSet*>(c->na_stage);
    uint64_t* begin = reinterpret_cast<uint64_t*>(static_cast<char*>(c->na_stage) +
                                                  (MCML_NOALLOC_OFF_BEGIN - MCML_NOALLOC_OFF_SETS));
    begin[0] = 0;
    int L = 0, S = 0;
    for (int j = 0; j < nsets; j++) {
        mcml_derived d;
        const int rc = mcml_derive(&sims[j], &d);
        if (rc != MCML_OK) return derive_error(c, j, rc);
        if (j == 0) {
            L = d.num_layers;
            S = d.num_sds;
        } else if (d.num_layers != L || d.num_sds != S) {
            set_err(c, "set %d: all sets of a batch must share number_layers and num_SDS", j);
            return MCML_ERR_ARG;
        }
        if (sims[j].number_photons >= (1ull << 40)) {
            set_err(c, "set %d: number_photons must be below 2^40", j);
            return MCML_ERR_LIMIT;
        }
        fill_devset(d, hs[j]);
        hs[j].ticket_begin = begin[j];
        begin[j + 1] = begin[j] + d.number_photons;
    }
    const int groups = nsets * S;
    if (groups > MCML_NOALLOC_MAX_GROUPS) {
        set_err(c, "%d (set, SDS) groups: the allocation-free call takes at most %d", groups, MCML_NOALLOC_MAX_GROUPS);
        return MCML_ERR_LIMIT;
    }
    const int block = o.block_size > 0 ? o.block_size : 256;
    if (block % 64 != 0 || block > 256) {
        set_err(c, "block_size must be a multiple of 64 and at most 256");
        return MCML_ERR_ARG;
    }
    char* scr = static_cast<char*>(c->scratch);
    unsigned long long* ctrl = reinterpret_cast<unsigned long long*>(scr + MCML_NOALLOC_OFF_CTRL);
    DevSet* d_sets = reinterpret_cast<DevSet*>(scr + MCML_NOALLOC_OFF_SETS);
    uint64_t* d_begin = reinterpret_cast<uint64_t*>(scr + MCML_NOALLOC_OFF_BEGIN);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(scr + MCML_NOALLOC_OFF_ACC);
    const uint64_t acc_words = (uint64_t)groups * MCML_ACC_SLOTS * MCML_ACC_WORDS;
    HIP_TRY(c, hipMemcpyAsync(d_sets, hs, sizeof(DevSet) * (size_t)nsets, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_begin, begin, sizeof(uint64_t) * ((size_t)nsets + 1), hipMemcpyHostToDevice,
                              c->stream));
    FirstPassArgs a;
    memset(&a, 0, sizeof(a));
    a.sets = d_sets;
    a.set_begin = d_begin;
    a.nsets = nsets;
    a.num_layers = L;
    a.row_stride = 1;
    a.lds_sets = sizeof(DevSet) * (size_t)nsets <= kLdsTableMax ? 1 : 0;
    for (int j = 0; j < nsets && !a.any_iso; j++) a.any_iso = devset_iso(hs[j], L);
    int occ = 0;
    HIP_TRY(c, mcml_occupancy_first_pass_sums(a, block, &occ));
    if (occ < 1) occ = 1;
    const int bpc = o.blocks_per_cu > 0 ? std::min(o.blocks_per_cu, occ) : occ;
    const int grid = c->cu_count * bpc;
    a.chunk = o.chunk > 0 ? std::max(o.chunk, 64) : kDefaultChunk;
    a.refill_min = o.refill_min > 0 ? std::min(o.refill_min, 64) : kDefaultRefill;
    a.bnd_min = o.boundary_min > 0 ? std::min(o.boundary_min, 64) : kDefaultBoundary;
    a.total = begin[nsets];
    a.seed = seed;
    a.mult = c->d_mult;
    a.queue = ctrl + kCtrlQueue;
    a.n_records = ctrl + kCtrlRecords;
    a.counters = ctrl + kCtrlCounters;
    a.stop_at = ~0ull;
    a.accum = acc;
    HIP_TRY(c, mcml_launch_clear(ctrl, kCtrlWords, c->stream));
    HIP_TRY(c, mcml_launch_clear(acc, acc_words, c->stream));
    HIP_TRY(c, mcml_launch_first_pass_sums(a, grid, block, c->stream));
    HIP_TRY(c, mcml_launch_sums_finalize(acc, groups, d_reflectance, d_detected, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->h_ctrl, ctrl, sizeof(unsigned long long) * kCtrlFirstPass, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->h_ctrl[2 + CNT_PHOTONS] != begin[nsets]) {
        set_err(c, "allocation-free first pass launched %llu photons, expected %llu",
                (unsigned long long)c->h_ctrl[2 + CNT_PHOTONS], (unsigned long long)begin[nsets]);
        return MCML_ERR_DEVICE;
    }
    return MCML_OK;
}

/* calculate_average_pathlength, transport.cu:274-297 */
extern "C" int mcml_average_pathlength(const mcml_sim* sim, const mcml_result* res, int set, double* outv)
{
    if (!sim || !res || !outv || !res->replayed || set < 0 || set >= res->nsets) return MCML_ERR_ARG;
    mcml_derived d;
    int rc = mcml_derive(sim, &d);
    if (rc) return rc;
    const int S = res->num_sds, L = res->num_layers;
    const double inverse_start_weight = double(1 / d.start_weight);
    /* one (SDS, layer) pair: its two sums run over the records in order, as in the reference */
    auto pair = [&](int s, int l) {
        const uint64_t b = res->pl_offset[(size_t)set * S + s], e = res->pl_offset[(size_t)set * S + s + 1];
        double sum_r = 0, sum_c = 0;
        const double changed_mua = MCML_MUA_CHANGE_RATIO * d.layer[l].mua;
        for (uint64_t i = b; i < e; i++) {
            const float* rec = res->pathlength + i * (uint64_t)(L + 2);
            double energy = rec[0] * inverse_start_weight;
            sum_r += energy;
            energy *= exp(-1 * changed_mua * rec[l]);
            sum_c += energy;
        }
        outv[s * L + l - 1] = log(sum_r / sum_c) / changed_mua;
    };
    /* The S x L pairs are independent; with many records they run on up to 16 host threads
     * (each pair's sums keep their order, so the values do not depend on the threads). */
    const int npairs = S * L;
    const uint64_t nrec = res->pl_offset[(size_t)set * S + S] - res->pl_offset[(size_t)set * S];
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = nrec < 100000 ? 1 : (int)std::min<unsigned>({16u, hw, (unsigned)npairs});
    std::atomic<int> next(0);
    auto work = [&] {
        for (int q; (q = next.fetch_add(1)) < npairs;) pair(q / L, q % L + 1);
    };
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < nt; t++) th.emplace_back(work);
    } catch (...) { /* no thread to spare: this thread takes the rest */
    }
    work();
    for (std::thread& t : th) t.join();
    return MCML_OK;
}

/* ====================================================================== autotuning */
extern "C" int mcml_autotune(mcml_ctx* c, const mcml_sim* sims, int nsets, uint64_t trial_photons, mcml_opts* opts)
{
    if (!c || !sims || nsets < 1 || !opts || trial_photons < 64) return MCML_ERR_ARG;
    uint64_t total = 0;
    for (int j = 0; j < nsets; j++) total += sims[j].number_photons;
    if (total == 0) return MCML_ERR_ARG;
    /* the same sets with budgets scaled to about trial_photons in total */
    std::vector<mcml_sim> trial(sims, sims + nsets);
    const double scale = std::min(1.0, (double)trial_photons / (double)total);
    for (mcml_sim& s : trial) s.number_photons = std::max<uint64_t>(1, (uint64_t)((double)s.number_photons * scale));
    std::vector<mcml_derived> der;
    int rc = validate_batch(c, trial.data(), nsets, der);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    const int blocks[] = {128, 256};
    /* (refill threshold, boundary batch) pairs: idle lanes to wait for before a refill, and
     * lanes at a boundary to wait for before resolving boundaries together */
    const int rb[][2] = {{2, 1}, {4, 1}, {2, 4}, {4, 4}, {4, 8}};
    const int chunks[] = {64, 128, 256};
    double best = 1e300;
    mcml_opts pick = *opts;
    bool warmed = false;
    for (int b : blocks)
        for (const auto& p : rb)
            for (int ch : chunks) {
                const int r = p[0];
                mcml_opts o;
                memset(&o, 0, sizeof(o));
                o.block_size = b;
                o.refill_min = r;
                o.chunk = ch;
                o.boundary_min = opts->boundary_min > 0 ? opts->boundary_min : p[1];
                for (int rep = warmed ? 0 : -1; rep < 1; rep++) { /* one untimed warm-up run */
                    mcml_result res;
                    memset(&res, 0, sizeof(res));
                    uint64_t nrec = 0;
                    rc = first_pass(c, der, o, 12345, 1, &res, nrec);
                    if (rc) return rc;
                    if (rep == 0 && res.ms_first_pass_kernel < best) {
                        best = res.ms_first_pass_kernel;
                        pick.block_size = b;
                        pick.refill_min = r;
                        pick.chunk = ch;
                        pick.boundary_min = o.boundary_min;
                    }
                }
                warmed = true;
            }
    *opts = pick;
    return MCML_OK;
}

/* ====================================================================== White Monte Carlo */
extern "C" int mcml_white_rescale(mcml_ctx* c, const mcml_sim* sims, const mcml_result* res, const float* mua_new,
                                  int m, double* out, double* out_sq)
{
    if (!c || !sims || !res || !mua_new || !out || m < 1 || !res->replayed) return MCML_ERR_ARG;
    const int nsets = res->nsets, S = res->num_sds, L = res->num_layers;
    if (L > 16) return MCML_ERR_LIMIT;
    if (m > 65535 || (int64_t)nsets * S > 65535) {
        set_err(c, "at most 65535 mua vectors and 65535 (set, SDS) groups per call");
        return MCML_ERR_LIMIT;
    }
    const size_t groups = (size_t)nsets * (size_t)S;
    const uint64_t n = res->pl_offset[groups];
    std::vector<float> old((size_t)nsets * (size_t)L);
    for (int j = 0; j < nsets; j++)
        for (int l = 0; l < L; l++) old[(size_t)j * L + l] = sims[j].layers[l].mua;
    int rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = ensure(c, c->wrows, sizeof(float) * (size_t)std::max<uint64_t>(n, 1) * (size_t)(L + 2))) ||
        (rc = ensure(c, c->woff, sizeof(uint64_t) * (groups + 1))) ||
        (rc = ensure(c, c->wold, sizeof(float) * old.size())) ||
        (rc = ensure(c, c->wnew, sizeof(float) * (size_t)nsets * (size_t)m * (size_t)L)) ||
        (rc = ensure(c, c->wout, 2 * sizeof(double) * (size_t)nsets * (size_t)m * (size_t)S)))
        return rc;
    const size_t nout = (size_t)nsets * (size_t)m * (size_t)S;
    if (n)
        HIP_TRY(c, hipMemcpyAsync(c->wrows.p, res->pathlength, sizeof(float) * (size_t)n * (size_t)(L + 2),
                                  hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->woff.p, res->pl_offset, sizeof(uint64_t) * (groups + 1), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->wold.p, old.data(), sizeof(float) * old.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->wnew.p, mua_new, sizeof(float) * (size_t)nsets * (size_t)m * (size_t)L,
                              hipMemcpyHostToDevice, c->stream));
    WhiteArgs a;
    memset(&a, 0, sizeof(a));
    a.rows = as<float>(c->wrows);
    a.offsets = as<uint64_t>(c->woff);
    a.mua_old = as<float>(c->wold);
    a.mua_new = as<float>(c->wnew);
    a.out = as<double>(c->wout);
    a.out_sq = out_sq ? as<double>(c->wout) + nout : nullptr;
    a.m = m;
    a.num_layers = L;
    a.num_sds = S;
    HIP_TRY(c, mcml_launch_white(a, nsets, c->stream));
    HIP_TRY(c, hipMemcpyAsync(out, c->wout.p, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream));
    if (out_sq)
        HIP_TRY(c, hipMemcpyAsync(out_sq, as<double>(c->wout) + nout, sizeof(double) * nout, hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MCML_OK;
}
