/*
 * mlp_host.cpp - host side of the reflectance surrogate (C ABI in include/mcml.h).
 *
 * Parameters are one flat FP32 array, layer by layer: W_l (out_l x in_l, row-major) then
 * b_l (out_l).  Hidden layers use GELU; the last layer is linear.  Training minimises the
 * (optionally per-sample weighted) mean squared error with Adam; mini-batches are drawn
 * from a device-resident copy of the training set by a per-epoch shuffle.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <math.h>
#include <random>
#include <vector>

#include "../../include/mcml.h"
#include "mlp.h"

namespace {

struct Layer {
    int in, out;
    int64_t w_off, b_off;
};

struct Buf {
    float* p = nullptr;
    size_t n = 0;
};

int grow(Buf& b, size_t n)
{
    if (n == 0) n = 1;
    if (b.n >= n) return MCML_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    if (hipMalloc((void**)&b.p, n * sizeof(float)) != hipSuccess) return MCML_ERR_NOMEM;
    b.n = n;
    return MCML_OK;
}

} // namespace

struct mcml_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<Layer> layers;
    int64_t nparams = 0;
    int64_t step = 0;
    Buf params, grad, m1, m2, part;
    std::vector<Buf> z, a;   /* per hidden layer: pre-activation and activation */
    Buf x, y, dy, loss, dz0, dz1;
    Buf data_x, data_y, data_w;
    Buf bt, bw;              /* mini-batch targets and weights */
    /* fused inference: packed bf16 weights, refreshed when the parameters change */
    bool fused_ok = false, fused_on = true;
    int64_t version = 0, packed_version = -1;
    uint16_t* packed = nullptr;
    Buf packed_b;
    MlpFusedArgs fa;
    int32_t* idx = nullptr;
    size_t idx_n = 0;
    int cap = 0;             /* batch capacity of the activation buffers */
};

namespace {

#define MLP_TRY(expr)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return MCML_ERR_DEVICE;                                                                  \
    } while (0)

int ensure_batch(mcml_mlp* m, int B)
{
    if (B <= m->cap) return MCML_OK;
    const size_t nh = m->layers.size() - 1;
    m->z.resize(nh);
    m->a.resize(nh);
    int rc;
    for (size_t l = 0; l < nh; l++) {
        if ((rc = grow(m->z[l], (size_t)B * m->layers[l].out)) || (rc = grow(m->a[l], (size_t)B * m->layers[l].out)))
            return rc;
    }
    int widest = 0;
    for (const Layer& L : m->layers) widest = std::max(widest, std::max(L.in, L.out));
    const int outs = m->layers.back().out;
    if ((rc = grow(m->x, (size_t)B * m->layers[0].in)) || (rc = grow(m->y, (size_t)B * outs)) ||
        (rc = grow(m->dy, (size_t)B * outs)) || (rc = grow(m->loss, (size_t)B)) ||
        (rc = grow(m->dz0, (size_t)B * widest)) || (rc = grow(m->dz1, (size_t)B * widest)))
        return rc;
    m->cap = B;
    return MCML_OK;
}

/* Pack the parameters for the fused kernel if they changed since the last pack. */
int pack_fused(mcml_mlp* m)
{
    if (m->packed_version == m->version) return MCML_OK;
    const size_t nl = m->layers.size();
    int64_t w = 0, b = 0;
    for (size_t l = 0; l < nl; l++) {
        const Layer& L = m->layers[l];
        const int Kp = l == 0 ? 32 : MLP_FUSED_H, Np = l + 1 < nl ? MLP_FUSED_H : 16;
        m->fa.w_off[l] = w;
        m->fa.b_off[l] = b;
        MLP_TRY(mcml_mlp_launch_pack(m->params.p + L.w_off, m->params.p + L.b_off, L.in, L.out, Kp, Np, m->packed + w,
                                     m->packed_b.p + b, m->stream));
        w += (int64_t)Kp * Np;
        b += Np;
    }
    m->packed_version = m->version;
    return MCML_OK;
}

/* Forward pass of B rows held in m->x; keep_z stores pre-activations for backward. */
int forward(mcml_mlp* m, int B, bool keep_z)
{
    const size_t nl = m->layers.size();
    if (!keep_z && m->fused_ok && m->fused_on) {
        int rc = pack_fused(m);
        if (rc) return rc;
        m->fa.x = m->x.p;
        m->fa.y = m->y.p;
        m->fa.packed = m->packed;
        m->fa.bias = m->packed_b.p;
        m->fa.B = B;
        m->fa.in = m->layers[0].in;
        m->fa.outs = m->layers.back().out;
        m->fa.hidden = (int)nl - 1;
        MLP_TRY(mcml_mlp_launch_fused(m->fa, m->stream));
        return MCML_OK;
    }
    const float* in = m->x.p;
    for (size_t l = 0; l < nl; l++) {
        const Layer& L = m->layers[l];
        MlpGemmArgs g;
        memset(&g, 0, sizeof(g));
        g.A = in;
        g.sam = L.in;
        g.sak = 1;
        g.B = m->params.p + L.w_off; /* B(k, n) = W[n][k] */
        g.sbk = 1;
        g.sbn = L.in;
        g.bias = m->params.p + L.b_off;
        g.M = B;
        g.N = L.out;
        g.K = L.in;
        if (l + 1 < nl) {
            g.C = m->a[l].p;
            g.Z = keep_z ? m->z[l].p : nullptr;
            g.ldc = L.out;
            g.epilogue = MLP_EPI_BIAS_GELU;
            in = m->a[l].p;
        } else {
            g.C = m->y.p;
            g.ldc = L.out;
            g.epilogue = MLP_EPI_BIAS;
        }
        MLP_TRY(mcml_mlp_launch_gemm(g, 1, m->stream));
    }
    return MCML_OK;
}

/* dW = dZ^T A_in over B rows (split over the batch, reduced in order) and db = colsum(dZ). */
int weight_grad(mcml_mlp* m, const Layer& L, const float* dz, const float* a_in, int B)
{
    const int splits = std::max(1, std::min(64, B / 256));
    const int64_t plane = (int64_t)L.out * L.in;
    /* weight partials [splits][out][in], then bias partials [splits][out] */
    if (grow(m->part, (size_t)(plane + L.out) * splits)) return MCML_ERR_NOMEM;
    float* bias_part = m->part.p + plane * splits;
    MlpGemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = dz; /* A(o, b) = dZ[b][o] */
    g.sam = 1;
    g.sak = L.out;
    g.B = a_in; /* B(b, i) = A_in[b][i] */
    g.sbk = L.in;
    g.sbn = 1;
    g.C = m->part.p;
    g.ldc = L.in;
    g.split_stride = plane;
    g.M = L.out;
    g.N = L.in;
    g.K = B;
    g.epilogue = MLP_EPI_SPLIT;
    MLP_TRY(mcml_mlp_launch_gemm(g, splits, m->stream));
    MLP_TRY(mcml_mlp_launch_reduce(m->part.p, splits, plane, plane, m->grad.p + L.w_off, m->stream));
    MLP_TRY(mcml_mlp_launch_colsum(dz, B, L.out, splits, bias_part, m->stream));
    MLP_TRY(mcml_mlp_launch_reduce(bias_part, splits, L.out, L.out, m->grad.p + L.b_off, m->stream));
    return MCML_OK;
}

/* Gradient of the weighted MSE for B rows in m->x with targets t (device) and weights w. */
int backward(mcml_mlp* m, int B, const float* t, const float* w)
{
    const int nl = (int)m->layers.size();
    const int outs = m->layers.back().out;
    MLP_TRY(mcml_mlp_launch_loss(m->y.p, t, w, B, outs, m->dy.p, m->loss.p, m->stream));
    const float* dcur = m->dy.p; /* gradient w.r.t. the current layer's output (pre-activation) */
    Buf* bufs[2] = {&m->dz0, &m->dz1};
    int which = 0;
    for (int l = nl - 1; l >= 0; l--) {
        const Layer& L = m->layers[l];
        const float* a_in = l > 0 ? m->a[l - 1].p : m->x.p;
        int rc = weight_grad(m, L, dcur, a_in, B);
        if (rc) return rc;
        if (l == 0) break;
        /* dZ_{l-1} = (dcur W_l) * GELU'(Z_{l-1}) */
        MlpGemmArgs g;
        memset(&g, 0, sizeof(g));
        g.A = dcur;
        g.sam = L.out;
        g.sak = 1;
        g.B = m->params.p + L.w_off; /* B(o, i) = W[o][i] */
        g.sbk = L.in;
        g.sbn = 1;
        g.C = bufs[which]->p;
        g.ldc = L.in;
        g.Z = m->z[l - 1].p;
        g.M = B;
        g.N = L.in;
        g.K = L.out;
        g.epilogue = MLP_EPI_DGELU;
        MLP_TRY(mcml_mlp_launch_gemm(g, 1, m->stream));
        dcur = bufs[which]->p;
        which ^= 1;
    }
    return MCML_OK;
}

int adam(mcml_mlp* m, float lr, float wd)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->step++;
    const double c1 = 1.0 - std::pow((double)b1, (double)m->step), c2 = 1.0 - std::pow((double)b2, (double)m->step);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    m->version++;
    MLP_TRY(mcml_mlp_launch_adam(m->params.p, m->grad.p, m->m1.p, m->m2.p, m->nparams, lr_t, b1, b2, eps, wd,
                                 m->stream));
    return MCML_OK;
}

} // namespace

extern "C" mcml_mlp* mcml_mlp_create(int device, const int* widths, int nwidths, uint64_t seed, int* err)
{
    int dummy;
    if (!err) err = &dummy;
    if (!widths || nwidths < 2) {
        *err = MCML_ERR_ARG;
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        *err = MCML_ERR_NODEV;
        return nullptr;
    }
    mcml_mlp* m = new mcml_mlp();
    m->device = device;
    int64_t off = 0;
    for (int l = 0; l + 1 < nwidths; l++) {
        Layer L;
        L.in = widths[l];
        L.out = widths[l + 1];
        L.w_off = off;
        off += (int64_t)L.in * L.out;
        L.b_off = off;
        off += L.out;
        m->layers.push_back(L);
    }
    m->nparams = off;
    /* the fused inference kernel takes <= 32 inputs, 1..8 GELU layers of 256, <= 16 outputs */
    m->fused_ok = nwidths >= 3 && nwidths - 2 <= MLP_FUSED_MAX_HIDDEN && widths[0] <= 32 && widths[nwidths - 1] <= 16;
    for (int l = 1; l + 1 < nwidths; l++) m->fused_ok = m->fused_ok && widths[l] == MLP_FUSED_H;
    memset(&m->fa, 0, sizeof(m->fa));
    /* He-style initialisation (normal, std sqrt(2/in)), zero biases, from the seed */
    std::vector<float> init((size_t)off, 0.0f);
    std::mt19937_64 rng(seed);
    for (const Layer& L : m->layers) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)L.in));
        for (int64_t i = 0; i < (int64_t)L.in * L.out; i++) init[(size_t)(L.w_off + i)] = nd(rng);
    }
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&m->e0) != hipSuccess || hipEventCreate(&m->e1) != hipSuccess || grow(m->params, off) ||
        grow(m->grad, off) || grow(m->m1, off) || grow(m->m2, off) ||
        hipMemcpy(m->params.p, init.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1.p, 0, sizeof(float) * off) != hipSuccess ||
        hipMemset(m->m2.p, 0, sizeof(float) * off) != hipSuccess ||
        (m->fused_ok && (hipMalloc((void**)&m->packed, sizeof(uint16_t) * (size_t)(32 * MLP_FUSED_H + (nwidths - 3) *
                                                                                   MLP_FUSED_H * MLP_FUSED_H +
                                                                               MLP_FUSED_H * 16)) != hipSuccess ||
                         grow(m->packed_b, (size_t)(nwidths - 2) * MLP_FUSED_H + 16)))) {
        *err = MCML_ERR_DEVICE;
        mcml_mlp_destroy(m);
        return nullptr;
    }
    *err = MCML_OK;
    return m;
}

extern "C" void mcml_mlp_destroy(mcml_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    Buf* bufs[] = {&m->params, &m->grad, &m->m1,  &m->m2,     &m->part,   &m->x,     &m->y,  &m->dy,
                   &m->loss,   &m->dz0,  &m->dz1, &m->data_x, &m->data_y, &m->data_w, &m->bt, &m->bw};
    for (Buf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (Buf& b : m->z)
        if (b.p) (void)hipFree(b.p);
    for (Buf& b : m->a)
        if (b.p) (void)hipFree(b.p);
    if (m->idx) (void)hipFree(m->idx);
    if (m->packed) (void)hipFree(m->packed);
    if (m->packed_b.p) (void)hipFree(m->packed_b.p);
    if (m->e0) (void)hipEventDestroy(m->e0);
    if (m->e1) (void)hipEventDestroy(m->e1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" int64_t mcml_mlp_num_params(const mcml_mlp* m) { return m ? m->nparams : 0; }

extern "C" int mcml_mlp_use_fused(mcml_mlp* m, int enable)
{
    if (!m) return MCML_ERR_ARG;
    if (enable >= 0) m->fused_on = enable != 0;
    return m->fused_ok && m->fused_on ? 1 : 0;
}

extern "C" int mcml_mlp_params(mcml_mlp* m, float* host, int64_t n, int write)
{
    if (!m || !host || n != m->nparams) return MCML_ERR_ARG;
    (void)hipSetDevice(m->device);
    const hipMemcpyKind k = write ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (write) MLP_TRY(hipMemcpy(m->params.p, host, sizeof(float) * n, k));
    else MLP_TRY(hipMemcpy(host, m->params.p, sizeof(float) * n, k));
    if (write) {
        m->version++;
        MLP_TRY(hipMemset(m->m1.p, 0, sizeof(float) * n));
        MLP_TRY(hipMemset(m->m2.p, 0, sizeof(float) * n));
        m->step = 0;
    }
    return MCML_OK;
}

extern "C" int mcml_mlp_predict(mcml_mlp* m, const float* x, int B, float* y, double* ms)
{
    if (!m || !x || !y || B < 1) return MCML_ERR_ARG;
    (void)hipSetDevice(m->device);
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    const int in = m->layers[0].in, outs = m->layers.back().out;
    MLP_TRY(hipMemcpyAsync(m->x.p, x, sizeof(float) * (size_t)B * in, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipEventRecord(m->e0, m->stream));
    if ((rc = forward(m, B, false))) return rc;
    MLP_TRY(hipEventRecord(m->e1, m->stream));
    MLP_TRY(hipMemcpyAsync(y, m->y.p, sizeof(float) * (size_t)B * outs, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (ms) {
        float t = 0.0f;
        MLP_TRY(hipEventElapsedTime(&t, m->e0, m->e1));
        *ms = t;
    }
    return MCML_OK;
}

extern "C" int mcml_mlp_gradient(mcml_mlp* m, const float* x, const float* t, const float* w, int B, float* grad,
                                 float* loss)
{
    if (!m || !x || !t || !grad || B < 1) return MCML_ERR_ARG;
    (void)hipSetDevice(m->device);
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    const int in = m->layers[0].in, outs = m->layers.back().out;
    if ((rc = grow(m->data_y, (size_t)B * outs)) || (rc = grow(m->data_w, (size_t)B * outs))) return rc;
    MLP_TRY(hipMemcpyAsync(m->x.p, x, sizeof(float) * (size_t)B * in, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->data_y.p, t, sizeof(float) * (size_t)B * outs, hipMemcpyHostToDevice, m->stream));
    if (w)
        MLP_TRY(hipMemcpyAsync(m->data_w.p, w, sizeof(float) * (size_t)B * outs, hipMemcpyHostToDevice, m->stream));
    if ((rc = forward(m, B, true)) || (rc = backward(m, B, m->data_y.p, w ? m->data_w.p : nullptr))) return rc;
    MLP_TRY(hipMemcpyAsync(grad, m->grad.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->stream));
    std::vector<float> rl((size_t)B);
    MLP_TRY(hipMemcpyAsync(rl.data(), m->loss.p, sizeof(float) * B, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0;
        for (float v : rl) s += v;
        *loss = (float)(s / B);
    }
    return MCML_OK;
}

extern "C" int mcml_mlp_adam_step(mcml_mlp* m, const float* grad, float lr, float wd)
{
    if (!m || !grad) return MCML_ERR_ARG;
    (void)hipSetDevice(m->device);
    MLP_TRY(hipMemcpyAsync(m->grad.p, grad, sizeof(float) * m->nparams, hipMemcpyHostToDevice, m->stream));
    int rc = adam(m, lr, wd);
    if (rc) return rc;
    MLP_TRY(hipStreamSynchronize(m->stream));
    return MCML_OK;
}

extern "C" int mcml_mlp_train(mcml_mlp* m, const float* x, const float* t, const float* w, int n, int epochs, int batch,
                              float lr, float lr_final, float wd, uint64_t seed, float* epoch_loss, double* ms)
{
    if (!m || !x || !t || n < 1 || epochs < 1 || batch < 1) return MCML_ERR_ARG;
    (void)hipSetDevice(m->device);
    const int in = m->layers[0].in, outs = m->layers.back().out;
    batch = std::min(batch, n);
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = grow(m->data_x, (size_t)n * in)) ||
        (rc = grow(m->data_y, (size_t)n * outs)) || (rc = grow(m->data_w, (size_t)n * outs)))
        return rc;
    if (m->idx_n < (size_t)n) {
        if (m->idx) (void)hipFree(m->idx);
        m->idx = nullptr;
        m->idx_n = 0;
        MLP_TRY(hipMalloc((void**)&m->idx, sizeof(int32_t) * n));
        m->idx_n = (size_t)n;
    }
    Buf& bt = m->bt;
    Buf& bw = m->bw;
    if ((rc = grow(bt, (size_t)batch * outs)) || (rc = grow(bw, (size_t)batch * outs))) return rc;
    MLP_TRY(hipMemcpyAsync(m->data_x.p, x, sizeof(float) * (size_t)n * in, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->data_y.p, t, sizeof(float) * (size_t)n * outs, hipMemcpyHostToDevice, m->stream));
    if (w)
        MLP_TRY(hipMemcpyAsync(m->data_w.p, w, sizeof(float) * (size_t)n * outs, hipMemcpyHostToDevice, m->stream));
    std::vector<int32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    const int steps_per_epoch = n / batch;
    const int64_t total_steps = (int64_t)steps_per_epoch * epochs;
    int64_t s_idx = 0;
    std::vector<float> rl((size_t)batch);
    MLP_TRY(hipEventRecord(m->e0, m->stream));
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_TRY(hipMemcpyAsync(m->idx, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
        double acc = 0.0;
        for (int b = 0; b < steps_per_epoch; b++, s_idx++) {
            const int32_t* ib = m->idx + (size_t)b * batch;
            MLP_TRY(mcml_mlp_launch_gather(m->data_x.p, ib, batch, in, m->x.p, m->stream));
            MLP_TRY(mcml_mlp_launch_gather(m->data_y.p, ib, batch, outs, bt.p, m->stream));
            if (w) MLP_TRY(mcml_mlp_launch_gather(m->data_w.p, ib, batch, outs, bw.p, m->stream));
            if ((rc = forward(m, batch, true)) || (rc = backward(m, batch, bt.p, w ? bw.p : nullptr))) return rc;
            /* cosine decay from lr to lr_final over all steps */
            const double f = 0.5 * (1.0 + std::cos(M_PI * (double)s_idx / (double)std::max<int64_t>(1, total_steps)));
            if ((rc = adam(m, (float)(lr_final + (lr - lr_final) * f), wd))) return rc;
            if (epoch_loss && b + 1 == steps_per_epoch) { /* loss of the epoch's last batch */
                MLP_TRY(hipMemcpyAsync(rl.data(), m->loss.p, sizeof(float) * batch, hipMemcpyDeviceToHost, m->stream));
                MLP_TRY(hipStreamSynchronize(m->stream));
                for (float v : rl) acc += v;
                epoch_loss[e] = (float)(acc / batch);
            }
        }
    }
    MLP_TRY(hipEventRecord(m->e1, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (ms) {
        float tms = 0.0f;
        MLP_TRY(hipEventElapsedTime(&tms, m->e0, m->e1));
        *ms = tms;
    }
    return MCML_OK;
}
