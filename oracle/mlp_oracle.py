"""CPU restatement of the surrogate network's arithmetic (TEST INFRASTRUCTURE ONLY).

Used only by the tests to check the HIP kernels of mcml_gpu_amd/csrc/mlp.hip: forward,
backward and Adam in float64, with the operands of every matrix product rounded to bf16
exactly as the matrix-core kernels round them (round to nearest even).  With bf16=False it
is the plain float64 network, whose gradient the tests check against finite differences.
"""
import numpy as np


def to_bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


GELU_K = np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def gelu(z):
    """GELU in its tanh form: 0.5 z (1 + tanh(k (z + c z^3))), k = sqrt(2/pi), c = 0.044715."""
    return 0.5 * z * (1.0 + np.tanh(GELU_K * z * (1.0 + GELU_C * z * z)))


def gelu_grad(z):
    t = np.tanh(GELU_K * z * (1.0 + GELU_C * z * z))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_C * z * z)


def unpack(params, widths):
    """[(W [out x in], b [out])] views of a flat parameter vector."""
    out, off = [], 0
    for i, o in zip(widths[:-1], widths[1:]):
        W = params[off:off + i * o].reshape(o, i)
        off += i * o
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def _mm(a, b, bf16):
    if bf16:
        return to_bf16(a) @ to_bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(params, widths, x, bf16=True):
    """Returns (y, zs, acts): outputs, pre-activations and activations per hidden layer.
    The kernels keep activations in float32 between layers; so does this restatement."""
    layers = unpack(np.asarray(params, np.float64), widths)
    a = np.asarray(x, np.float32 if bf16 else np.float64)
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = _mm(a, W.T.astype(np.float32) if bf16 else W.T, bf16) + b
        if li + 1 < len(layers):
            if bf16:
                z = z.astype(np.float32)
            zs.append(z)
            a = gelu(z.astype(np.float64))
            if bf16:
                a = a.astype(np.float32)
            acts.append(a)
        else:
            return (z.astype(np.float32) if bf16 else z), zs, acts


def gradient(params, widths, x, t, w=None, bf16=True):
    """(flat gradient, loss) of mean_b mean_o w_bo (y_bo - t_bo)^2 (w per target, or None)."""
    y, zs, acts = forward(params, widths, x, bf16)
    B, outs = y.shape
    wv = np.ones((B, outs)) if w is None else np.asarray(w, np.float64).reshape(B, outs)
    d = y.astype(np.float64) - np.asarray(t, np.float64)
    loss = float(np.mean(wv * d * d))
    dcur = (2.0 * wv * d / (B * outs))
    if bf16:
        dcur = dcur.astype(np.float32)
    layers = unpack(np.asarray(params, np.float64), widths)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        a_in = acts[li]
        dW = _mm(np.asarray(dcur).T, a_in, bf16)
        db = np.asarray(dcur, np.float64).sum(axis=0)
        grads.append((dW, db))
        if li == 0:
            break
        da = _mm(dcur, W.astype(np.float32) if bf16 else W, bf16)
        dcur = da * gelu_grad(np.asarray(zs[li - 1], np.float64))
        if bf16:
            dcur = dcur.astype(np.float32)
    flat = []
    for dW, db in reversed(grads):
        flat.append(dW.ravel())
        flat.append(db)
    return np.concatenate(flat), loss


def adam_step(p, g, m1, m2, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), bias correction folded into the step size."""
    g = g + wd * p
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    lr_t = lr * np.sqrt(1 - b2 ** step) / (1 - b1 ** step)
    return p - lr_t * m1 / (np.sqrt(m2) + eps), m1, m2
