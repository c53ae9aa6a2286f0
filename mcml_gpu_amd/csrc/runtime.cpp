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
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mcml.h"
#include "mcml_device.h"
#include "mcml_launch.h"
#include "mwc_multipliers.h"

namespace {

constexpr float kPI = 3.141592654f; /* header.h:40 */
/* launch defaults (results never depend on them: each photon's stream is fixed) */
constexpr int kDefaultChunk = 128;
constexpr int kDefaultRefill = 2;
constexpr int kDefaultBoundary = 1;
constexpr size_t kLdsTableMax = 16384;

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
    char name[256] = {0};
    char err[1024] = {0};
    uint32_t* d_mult = nullptr;
    DevBuf sets, set_begin, queue, records, nrec, counters, jobs, rw, rs, rpl, dump, arz, a0z;
};

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

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

template <class T>
T* as(DevBuf& b) { return static_cast<T*>(b.p); }

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
        o.hg_a = 1.0f - in.g * in.g;
        o.hg_b = 1.0f - in.g;
        o.hg_c = 2.0f * in.g;
        o.hg_d = 1.0f + in.g * in.g;
        o.hg_inv2g = 1.0f / (2.0f * in.g);
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

bool rec_less(const mcml_record& a, const mcml_record& b)
{
    return a.key < b.key || (a.key == b.key && a.sds < b.sds);
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
    snprintf(c->name, sizeof(c->name), "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    if (hipMalloc((void**)&c->d_mult, sizeof(uint32_t) * MCML_MWC_TABLE_SIZE) != hipSuccess ||
        hipMemcpy(c->d_mult, mcml_mwc_multipliers, sizeof(uint32_t) * MCML_MWC_TABLE_SIZE,
                  hipMemcpyHostToDevice) != hipSuccess) {
        *err = MCML_ERR_DEVICE;
        mcml_destroy(c);
        return nullptr;
    }
    *err = MCML_OK;
    return c;
}

extern "C" void mcml_destroy(mcml_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->sets, &c->set_begin, &c->queue, &c->records, &c->nrec, &c->counters, &c->jobs,
                      &c->rw, &c->rs, &c->rpl, &c->dump, &c->arz, &c->a0z};
    for (DevBuf* b : bufs) release(*b);
    if (c->d_mult) (void)hipFree(c->d_mult);
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

int validate_batch(mcml_ctx* c, const mcml_sim* sims, int nsets, std::vector<mcml_derived>& der)
{
    if (nsets < 1) {
        set_err(c, "nsets must be >= 1");
        return MCML_ERR_ARG;
    }
    if ((uint64_t)nsets >= (1ull << 23)) {
        set_err(c, "too many parameter sets (%d)", nsets);
        return MCML_ERR_LIMIT;
    }
    der.resize((size_t)nsets);
    for (int j = 0; j < nsets; j++) {
        int rc = mcml_derive(&sims[j], &der[(size_t)j]);
        if (rc != MCML_OK) {
            set_err(c, "set %d: layers must be 1..%d and SDS 1..%d", j, MCML_MAX_LAYERS, MCML_MAX_SDS);
            return rc;
        }
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

int first_pass(mcml_ctx* c, const std::vector<mcml_derived>& der, const mcml_opts& o, uint64_t seed,
               mcml_result* out, std::vector<mcml_record>& recs)
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
        (rc = ensure(c, c->set_begin, sizeof(uint64_t) * ((size_t)nsets + 1))) ||
        (rc = ensure(c, c->queue, sizeof(unsigned long long))) ||
        (rc = ensure(c, c->nrec, sizeof(unsigned long long))) ||
        (rc = ensure(c, c->counters, sizeof(unsigned long long) * CNT_N)))
        return rc;
    HIP_TRY(c, hipMemcpyAsync(c->sets.p, hs.data(), sizeof(DevSet) * (size_t)nsets, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->set_begin.p, begin.data(), sizeof(uint64_t) * ((size_t)nsets + 1),
                              hipMemcpyHostToDevice, c->stream));

    const int block = o.block_size > 0 ? o.block_size : 256;
    if (block % 64 != 0 || block > 256) {
        set_err(c, "block_size must be a multiple of 64 and at most 256");
        return MCML_ERR_ARG;
    }
    /* a small set table is staged in LDS by every workgroup; a large batch reads it from
     * global memory (it is only read at launch, layer changes and boundary hits) */
    const size_t table_bytes = sizeof(DevSet) * (size_t)nsets;
    const int lds_sets = table_bytes <= kLdsTableMax ? 1 : 0;
    int occ = 0;
    HIP_TRY(c, mcml_occupancy_first_pass(block, lds_sets ? table_bytes : 0, &occ));
    if (occ < 1) occ = 1;
    const int bpc = o.blocks_per_cu > 0 ? std::min(o.blocks_per_cu, occ) : occ;
    const int grid = c->cu_count * bpc;
    const int chunk = o.chunk > 0 ? std::max(o.chunk, 64) : kDefaultChunk;
    const int refill_min = o.refill_min > 0 ? std::min(o.refill_min, 64) : kDefaultRefill;
    const int bnd_min = o.boundary_min > 0 ? std::min(o.boundary_min, 64) : kDefaultBoundary;

    /* record capacity: detection is rare for small fibers; grow and rerun on overflow
     * (the result does not depend on the capacity, each photon's stream is fixed) */
    uint64_t cap = std::max<uint64_t>(1u << 16, std::min<uint64_t>(total / 16 + 1024, 1ull << 26));
    uint64_t n_rec = 0;
    unsigned long long cnt[CNT_N];
    double ms_kernel = 0;
    int launches = 0;
    for (;;) {
        if ((rc = ensure(c, c->records, sizeof(mcml_record) * (size_t)cap))) return rc;
        HIP_TRY(c, hipMemsetAsync(c->queue.p, 0, sizeof(unsigned long long), c->stream));
        HIP_TRY(c, hipMemsetAsync(c->nrec.p, 0, sizeof(unsigned long long), c->stream));
        HIP_TRY(c, hipMemsetAsync(c->counters.p, 0, sizeof(unsigned long long) * CNT_N, c->stream));
        FirstPassArgs a;
        memset(&a, 0, sizeof(a));
        a.sets = as<DevSet>(c->sets);
        a.set_begin = as<uint64_t>(c->set_begin);
        a.nsets = nsets;
        a.chunk = chunk;
        a.refill_min = refill_min;
        a.bnd_min = bnd_min;
        a.lds_sets = lds_sets;
        a.total = total;
        a.seed = seed;
        a.mult = c->d_mult;
        a.queue = as<unsigned long long>(c->queue);
        a.records = as<mcml_record>(c->records);
        a.n_records = as<unsigned long long>(c->nrec);
        a.cap = cap;
        a.counters = as<unsigned long long>(c->counters);
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        HIP_TRY(c, mcml_launch_first_pass(a, grid, block, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        launches++;
        unsigned long long nr = 0;
        HIP_TRY(c, hipMemcpyAsync(&nr, c->nrec.p, sizeof(nr), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        ms_kernel += ms_between(c->ev0, c->ev1);
        n_rec = nr;
        if (n_rec <= cap) break;
        cap = n_rec + n_rec / 8 + 1024;
    }
    recs.resize((size_t)n_rec);
    if (n_rec)
        HIP_TRY(c, hipMemcpy(recs.data(), c->records.p, sizeof(mcml_record) * (size_t)n_rec, hipMemcpyDeviceToHost));
    std::sort(recs.begin(), recs.end(), rec_less);
    out->photons = cnt[CNT_PHOTONS];
    out->steps = cnt[CNT_STEPS];
    out->steps_boundary = cnt[CNT_BOUNDARY];
    out->steps_interaction = cnt[CNT_INTERACTION];
    out->ms_first_pass_kernel = ms_kernel;
    out->first_pass_launches = launches;
    out->block_size = block;
    out->grid_size = grid;
    out->chunk = chunk;
    out->refill_min = refill_min;
    out->boundary_min = bnd_min;
    if (out->photons != total) {
        set_err(c, "first pass launched %llu photons, expected %llu", (unsigned long long)out->photons,
                (unsigned long long)total);
        return MCML_ERR_DEVICE;
    }
    return MCML_OK;
}

int replay(mcml_ctx* c, const std::vector<mcml_derived>& der, const mcml_opts& o, uint64_t seed,
           const std::vector<mcml_record>& recs, mcml_result* out)
{
    const int nsets = (int)der.size();
    const int S = der[0].num_sds, L = der[0].num_layers;
    const int nr = der[0].grid_nr, nz = der[0].grid_nz;
    /* jobs grouped by (set, sds), photon order inside a group: each group is one
     * pathlength_SDS_# file image */
    std::vector<ReplayJob> jobs(recs.size());
    for (size_t i = 0; i < recs.size(); i++) {
        jobs[i].k = recs[i].key & ((1ull << 40) - 1);
        jobs[i].set = (uint32_t)(recs[i].key >> 40);
        jobs[i].sds = recs[i].sds;
    }
    std::stable_sort(jobs.begin(), jobs.end(), [](const ReplayJob& a, const ReplayJob& b) {
        return a.set < b.set || (a.set == b.set && (a.sds < b.sds || (a.sds == b.sds && a.k < b.k)));
    });
    const uint64_t nj = jobs.size();
    const size_t groups = (size_t)nsets * (size_t)S;
    out->pl_offset = (uint64_t*)calloc(groups + 1, sizeof(uint64_t));
    out->replay_reflectance = (double*)calloc(groups, sizeof(double));
    out->pathlength = (float*)malloc(sizeof(float) * (size_t)std::max<uint64_t>(nj, 1) * (size_t)(L + 2));
    if (!out->pl_offset || !out->replay_reflectance || !out->pathlength) {
        set_err(c, "host allocation failed");
        return MCML_ERR_NOMEM;
    }
    for (const ReplayJob& j : jobs) out->pl_offset[(size_t)j.set * S + (j.sds - 1) + 1]++;
    for (size_t g = 0; g < groups; g++) out->pl_offset[g + 1] += out->pl_offset[g];

    const bool grids = o.absorbance != 0;
    const size_t rz_cells = grids ? groups * (size_t)nz * (size_t)nr : 0;
    const size_t z_cells = grids ? groups * (size_t)nz : 0;
    int rc;
    if ((rc = ensure(c, c->jobs, sizeof(ReplayJob) * (size_t)nj)) || (rc = ensure(c, c->rw, sizeof(float) * (size_t)nj)) ||
        (rc = ensure(c, c->rs, sizeof(float) * (size_t)nj)) ||
        (rc = ensure(c, c->rpl, sizeof(float) * (size_t)nj * (size_t)L)) ||
        (rc = ensure(c, c->dump, sizeof(float) * (size_t)nj * (size_t)(L + 2))))
        return rc;
    if (grids) {
        if ((rc = ensure(c, c->arz, sizeof(unsigned long long) * rz_cells)) ||
            (rc = ensure(c, c->a0z, sizeof(unsigned long long) * z_cells)))
            return rc;
        HIP_TRY(c, hipMemsetAsync(c->arz.p, 0, sizeof(unsigned long long) * rz_cells, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->a0z.p, 0, sizeof(unsigned long long) * z_cells, c->stream));
    }
    if (nj) {
        HIP_TRY(c, hipMemcpyAsync(c->jobs.p, jobs.data(), sizeof(ReplayJob) * (size_t)nj, hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, hipMemsetAsync(c->queue.p, 0, sizeof(unsigned long long), c->stream));
        ReplayArgs a;
        memset(&a, 0, sizeof(a));
        a.sets = as<DevSet>(c->sets);
        a.jobs = as<ReplayJob>(c->jobs);
        a.njobs = nj;
        a.seed = seed;
        a.mult = c->d_mult;
        a.queue = as<unsigned long long>(c->queue);
        a.chunk = 64;
        a.num_layers = L;
        a.bnd_min = o.boundary_min > 0 ? std::min(o.boundary_min, 64) : kDefaultBoundary;
        a.weight = as<float>(c->rw);
        a.steps = as<float>(c->rs);
        a.pl = as<float>(c->rpl);
        a.A_rz = grids ? as<unsigned long long>(c->arz) : nullptr;
        a.A0_z = grids ? as<unsigned long long>(c->a0z) : nullptr;
        a.num_sds = S;
        a.nr = nr;
        a.nz = nz;
        a.dr = der[0].grid_dr;
        a.dz = der[0].grid_dz;
        int occ = 0;
        HIP_TRY(c, mcml_occupancy_replay(&occ));
        if (occ < 1) occ = 1;
        const uint64_t waves_needed = (nj + 63) / 64;
        const int per_block = mcml_replay_block() / 64;
        uint64_t grid = std::min<uint64_t>((uint64_t)c->cu_count * occ, (waves_needed + per_block - 1) / per_block);
        if (grid < 1) grid = 1;
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        HIP_TRY(c, mcml_launch_replay(a, (int)grid, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        out->ms_replay_kernel = ms_between(c->ev0, c->ev1);

        DumpArgs d;
        memset(&d, 0, sizeof(d));
        d.weight = a.weight;
        d.steps = a.steps;
        d.pl = a.pl;
        d.njobs = nj;
        d.num_layers = L;
        d.out = as<float>(c->dump);
        const uint64_t elems = nj * (uint64_t)(L + 2);
        const uint64_t dgrid = std::min<uint64_t>((elems + 255) / 256, (uint64_t)c->cu_count * 16);
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        HIP_TRY(c, mcml_launch_dump(d, (int)std::max<uint64_t>(dgrid, 1), c->stream));
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipMemcpyAsync(out->pathlength, c->dump.p, sizeof(float) * (size_t)elems, hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        out->ms_dump_kernel = ms_between(c->ev0, c->ev1);
    }
    for (size_t g = 0; g < groups; g++) {
        double acc = 0;
        for (uint64_t i = out->pl_offset[g]; i < out->pl_offset[g + 1]; i++) acc += out->pathlength[i * (L + 2)];
        out->replay_reflectance[g] = acc;
    }
    if (grids) {
        out->A_rz = (uint64_t*)malloc(sizeof(uint64_t) * std::max<size_t>(rz_cells, 1));
        out->A0_z = (uint64_t*)malloc(sizeof(uint64_t) * std::max<size_t>(z_cells, 1));
        if (!out->A_rz || !out->A0_z) {
            set_err(c, "host allocation failed");
            return MCML_ERR_NOMEM;
        }
        HIP_TRY(c, hipMemcpy(out->A_rz, c->arz.p, sizeof(uint64_t) * rz_cells, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(out->A0_z, c->a0z.p, sizeof(uint64_t) * z_cells, hipMemcpyDeviceToHost));
    }
    out->replayed = 1;
    return MCML_OK;
}

} // namespace

extern "C" int mcml_run_batch(mcml_ctx* c, const mcml_sim* sims, int nsets, const mcml_opts* opts, uint64_t seed,
                              mcml_result* out)
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
    HIP_TRY(c, hipSetDevice(c->device));
    const int S = der[0].num_sds;
    out->nsets = nsets;
    out->num_sds = S;
    out->num_layers = der[0].num_layers;
    out->grid_nr = der[0].grid_nr;
    out->grid_nz = der[0].grid_nz;
    const size_t groups = (size_t)nsets * (size_t)S;
    out->reflectance = (double*)calloc(groups, sizeof(double));
    out->detected = (uint64_t*)calloc(groups, sizeof(uint64_t));
    if (!out->reflectance || !out->detected) {
        mcml_result_free(out);
        return MCML_ERR_NOMEM;
    }
    std::vector<mcml_record> recs;
    double t0 = now_ms();
    rc = first_pass(c, der, o, seed, out, recs);
    if (rc) {
        mcml_result_free(out);
        return rc;
    }
    for (const mcml_record& r : recs) { /* calculate_reflectance, transport.cu:299-311 */
        const size_t g = (size_t)(r.key >> 40) * S + (r.sds - 1);
        out->reflectance[g] += (double)r.weight;
        out->detected[g]++;
    }
    out->n_records = recs.size();
    out->records = (mcml_record*)malloc(sizeof(mcml_record) * std::max<size_t>(recs.size(), 1));
    if (!out->records) {
        mcml_result_free(out);
        return MCML_ERR_NOMEM;
    }
    if (!recs.empty()) memcpy(out->records, recs.data(), sizeof(mcml_record) * recs.size());
    double t1 = now_ms();
    out->ms_first_pass_wall = t1 - t0;
    if (o.replay) {
        rc = replay(c, der, o, seed, recs, out);
        if (rc) {
            mcml_result_free(out);
            return rc;
        }
        out->ms_replay_wall = now_ms() - t1;
    }
    return MCML_OK;
}

extern "C" int mcml_run(mcml_ctx* c, const mcml_sim* sim, const mcml_opts* opts, uint64_t seed, mcml_result* out)
{
    return mcml_run_batch(c, sim, 1, opts, seed, out);
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
    for (int s = 0; s < S; s++) {
        const uint64_t b = res->pl_offset[(size_t)set * S + s], e = res->pl_offset[(size_t)set * S + s + 1];
        for (int l = 1; l <= L; l++) {
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
        }
    }
    return MCML_OK;
}
