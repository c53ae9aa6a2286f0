"""The learned reflectance surrogate: network wrapper, features, data, evaluation.

Network: 14 inputs (mua, mus, g, n of three layers + the two upper thicknesses) ->
4 x 256 GELU -> 3 outputs (log reflectance per SDS), trained and evaluated by the HIP
kernels of csrc/mlp.hip through the C ABI (mcml_mlp_*).  Labels come from the batched
Monte Carlo (mcml_run_batch) of the cfg3 parameter sets, expanded over absorption by
White Monte Carlo rescaling of the recorded pathlengths (mcml_white_rescale).
"""
import ctypes as C

import numpy as np

from . import MCMLError, batch, load_library, make_opts

WIDTHS = (14, 256, 256, 256, 256, 3)
W_UNIT = 4294967295.0  # reflectance normaliser per photon (io.cpp:16)


def _bind(L):
    if getattr(L, "_mlp_bound", False):
        return L
    P = C.POINTER
    L.mcml_mlp_create.argtypes = [C.c_int, P(C.c_int), C.c_int, C.c_uint64, P(C.c_int)]
    L.mcml_mlp_create.restype = C.c_void_p
    L.mcml_mlp_destroy.argtypes = [C.c_void_p]
    L.mcml_mlp_num_params.argtypes = [C.c_void_p]
    L.mcml_mlp_num_params.restype = C.c_int64
    L.mcml_mlp_use_fused.argtypes = [C.c_void_p, C.c_int]
    L.mcml_mlp_params.argtypes = [C.c_void_p, P(C.c_float), C.c_int64, C.c_int]
    L.mcml_mlp_predict.argtypes = [C.c_void_p, P(C.c_float), C.c_int, P(C.c_float), P(C.c_double)]
    L.mcml_mlp_gradient.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int, P(C.c_float),
                                    P(C.c_float)]
    L.mcml_mlp_adam_step.argtypes = [C.c_void_p, P(C.c_float), C.c_float, C.c_float]
    L.mcml_mlp_train.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int, C.c_int, C.c_int,
                                 C.c_float, C.c_float, C.c_float, C.c_uint64, P(C.c_float), P(C.c_double)]
    L._mlp_bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The network on one GPU (mcml_mlp_* of libmcml)."""

    def __init__(self, widths=WIDTHS, seed=0, device=0):
        self.lib = _bind(load_library())
        self.widths = tuple(int(w) for w in widths)
        arr = (C.c_int * len(self.widths))(*self.widths)
        err = C.c_int(0)
        self.h = self.lib.mcml_mlp_create(device, arr, len(self.widths), seed, C.byref(err))
        if not self.h:
            raise MCMLError(err.value, "mcml_mlp_create")
        self.n = self.lib.mcml_mlp_num_params(self.h)

    def _check(self, rc, what):
        if rc:
            raise MCMLError(rc, what)

    def use_fused(self, enable=None):
        """Allow (True) or forbid (False) the fused inference kernel; returns whether predict
        uses it."""
        rc = self.lib.mcml_mlp_use_fused(self.h, -1 if enable is None else int(bool(enable)))
        if rc < 0:
            raise MCMLError(rc, "use_fused")
        return bool(rc)

    def get_params(self):
        p = np.zeros(self.n, np.float32)
        self._check(self.lib.mcml_mlp_params(self.h, _fp(p), self.n, 0), "params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.mcml_mlp_params(self.h, _fp(p), self.n, 1), "params")

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.widths[-1]), np.float32)
        ms = C.c_double(0)
        self._check(self.lib.mcml_mlp_predict(self.h, _fp(x), len(x), _fp(y), C.byref(ms)), "predict")
        self.last_ms = ms.value
        return y

    def gradient(self, x, t, w=None):
        """(flat gradient, loss) of the weighted mean squared error; w per target or None."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        wk = None if w is None else np.ascontiguousarray(np.broadcast_to(w, t.shape), np.float32)
        g = np.zeros(self.n, np.float32)
        loss = C.c_float(0)
        self._check(self.lib.mcml_mlp_gradient(self.h, _fp(x), _fp(t), None if wk is None else _fp(wk), len(x), _fp(g),
                                               C.byref(loss)), "gradient")
        return g, loss.value

    def adam_step(self, g, lr, wd=0.0):
        g = np.ascontiguousarray(g, np.float32)
        self._check(self.lib.mcml_mlp_adam_step(self.h, _fp(g), lr, wd), "adam")

    def train(self, x, t, w=None, epochs=10, batch=1024, lr=1e-3, lr_final=1e-5, wd=0.0, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        wk = None if w is None else np.ascontiguousarray(np.broadcast_to(w, t.shape), np.float32)
        hist = np.zeros(epochs, np.float32)
        ms = C.c_double(0)
        self._check(self.lib.mcml_mlp_train(self.h, _fp(x), _fp(t), _fp(wk) if wk is not None else None, len(x),
                                            epochs, batch, lr, lr_final, wd, seed, _fp(hist), C.byref(ms)), "train")
        self.train_ms = ms.value
        return hist

    def close(self):
        if self.h:
            self.lib.mcml_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------- features
LO = np.array([batch.RANGES["mua"][0]] * 3 + [batch.RANGES["mus"][0]] * 3 + [batch.RANGES["g"][0]] * 3 +
              [batch.RANGES["n"][0]] * 3 + [batch.RANGES["h"][0]] * 2)
HI = np.array([batch.RANGES["mua"][1]] * 3 + [batch.RANGES["mus"][1]] * 3 + [batch.RANGES["g"][1]] * 3 +
              [batch.RANGES["n"][1]] * 3 + [batch.RANGES["h"][1]] * 2)
LOG_FEATURES = np.r_[0:6]  # absorption and scattering enter on a log scale


def features(params):
    """Map the 14 optical parameters to [-1, 1] (log scale for mua and mus)."""
    p = np.asarray(params, np.float64).copy()
    lo, hi = LO.copy(), HI.copy()
    p[:, LOG_FEATURES] = np.log(p[:, LOG_FEATURES])
    lo[LOG_FEATURES] = np.log(lo[LOG_FEATURES])
    hi[LOG_FEATURES] = np.log(hi[LOG_FEATURES])
    return (2.0 * (p - lo) / (hi - lo) - 1.0).astype(np.float32)


# ------------------------------------------------------------------------- data
def simulate(ctx, params, budgets, seed, replay=True):
    """Run the parameter sets as one batch.  Returns (sims, result, R [nsets, S]) with R the
    detected weight per launched photon (the quantity output.txt holds, io.cpp:16-33)."""
    sims = [batch.sim_from_params(params[j], int(budgets[j])) for j in range(len(params))]
    res = ctx.run_batch(sims, make_opts(replay=replay, pathlength=replay, skip_records=True, use_graph=True),
                        seed=seed)
    R = res.reflectance / (W_UNIT * np.asarray(budgets, np.float64)[:, None])
    return sims, res, R


def relative_variance(sum_w, sum_w2, n):
    """Relative variance of the per-photon mean of n photons' detected weights (undetected
    photons count as zero) from the sums of the weights and of their squares."""
    sum_w = np.asarray(sum_w, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        rv = np.asarray(sum_w2, np.float64) / (sum_w * sum_w) - 1.0 / np.asarray(n, np.float64)
    return np.where(sum_w > 0, np.maximum(rv, 0.0), np.inf)


def white_dataset(ctx, params, budgets, m, seed):
    """Labels for the sets `params` (n_sets x 14) at their own absorption and at m - 1 further
    absorption vectors drawn uniformly from the mua range, by White Monte Carlo rescaling of
    one batched run.  Returns (x [N, 14], R [N, S], relvar [N, S]) with N = n_sets * m."""
    sims, res, _ = simulate(ctx, params, budgets, seed)
    nsets = len(params)
    rng = np.random.default_rng(seed)
    mua = rng.uniform(batch.RANGES["mua"][0], batch.RANGES["mua"][1], (nsets, m, 3)).astype(np.float32)
    mua[:, 0, :] = params[:, 0:3]
    W, W2 = ctx.white_rescale(sims, res, mua, with_squares=True)  # [nsets, m, S], W-units
    n = np.asarray(budgets, np.float64)[:, None, None]
    R = W / (W_UNIT * n)
    rv = relative_variance(W, W2, n)
    P = np.repeat(np.asarray(params, np.float64)[:, None, :], m, axis=1).copy()
    P[:, :, 0:3] = mua
    S = R.shape[-1]
    return P.reshape(-1, 14), R.reshape(-1, S), rv.reshape(-1, S)


def training_set(ctx, n_sets=512, m=391, seed=1, budget_lo=1e5, budget_hi=1e7):
    """The cfg3 batch (n_sets sets, log-uniform budgets) expanded m-fold over absorption."""
    _, params, budgets = batch.cfg3_sets(n_sets, seed=seed, budget_lo=budget_lo, budget_hi=budget_hi)
    return white_dataset(ctx, params, budgets, m, seed)


def reference_set(ctx, n_sets, photons, seed, chunk_photons=2e9, progress=None):
    """Held-out sets drawn from their own seed, each simulated with `photons` photons (plain
    MC, no rescaling).  Returns (params, R, relvar).  Runs in batches of about chunk_photons
    photons; progress(i, n) is called after each."""
    params = batch.draw_parameters(n_sets, seed)
    per = max(1, int(chunk_photons // photons))
    R = np.zeros((n_sets, 3))
    rv = np.zeros((n_sets, 3))
    for i in range(0, n_sets, per):
        p = params[i:i + per]
        budgets = np.full(len(p), int(photons))
        sims, res, Ri = simulate(ctx, p, budgets, seed + 7919 * (i + 1), replay=True)
        R[i:i + len(p)] = Ri
        mua = p[:, None, 0:3].astype(np.float32)
        W, W2 = ctx.white_rescale(sims, res, mua, with_squares=True)
        rv[i:i + len(p)] = relative_variance(W[:, 0], W2[:, 0], budgets[:, None])
        if progress:
            progress(min(i + per, n_sets), n_sets)
    return params, R, rv


# ------------------------------------------------------------------------- training
def fit(x_params, R, relvar=None, epochs=200, batch_size=1024, lr=2e-3, lr_final=2e-5, wd=0.0, seed=0,
        max_relvar=0.25, floor=1e-4, mlp=None):
    """Train the network on standardised log R.  Samples with an undetected SDS, or with a
    label noisier than max_relvar, are left out.  Each target is weighted by the inverse of
    its variance in standardised units, sd^2 / (relvar + floor) (the variance of log R is the
    relative variance of R; `floor` bounds the weight of near-noiseless labels), normalised
    to mean 1.  Returns (net, norm, loss history)."""
    ok = np.all(R > 0, axis=1)
    if relvar is not None:
        ok &= np.all(relvar <= max_relvar, axis=1)
    X = features(x_params[ok])
    Y = np.log(R[ok])
    mu, sd = Y.mean(axis=0), Y.std(axis=0)
    T = ((Y - mu) / sd).astype(np.float32)
    w = None
    if relvar is not None:
        w = sd * sd / (relvar[ok] + floor)
        w = (w / w.mean()).astype(np.float32)
    net = mlp or MLP(seed=seed)
    hist = net.train(X, T, w, epochs=epochs, batch=batch_size, lr=lr, lr_final=lr_final, wd=wd, seed=seed)
    return net, (mu, sd), hist


def predict_R(net, norm, params):
    mu, sd = norm
    y = net.predict(features(params)).astype(np.float64)
    return np.exp(y * sd + mu)


def relative_error(R_pred, R_ref):
    return np.abs(np.asarray(R_pred, np.float64) / np.asarray(R_ref, np.float64) - 1.0)


def photons_for_error(relvar_per_photon, target=0.02):
    """Photons one plain-MC evaluation needs for a relative standard error of `target` at
    every SDS, given each SDS's relative variance for a single photon."""
    return np.max(np.asarray(relvar_per_photon, np.float64), axis=-1) / (target * target)


class Surrogate:
    """Trained network plus the plain-MC fallback: the network answers only while its
    validated median relative error against high-photon MC is within `bound`."""

    def __init__(self, net, norm, bound=0.02):
        self.net, self.norm, self.bound = net, norm, bound
        self.validated_error = None

    def validate(self, params, R_ref):
        err = relative_error(predict_R(self.net, self.norm, params), R_ref)
        self.validated_error = float(np.median(err))
        return self.validated_error

    @property
    def usable(self):
        return self.validated_error is not None and self.validated_error <= self.bound

    def reflectance(self, params, ctx=None, photons=None, seed=0):
        """(R [n, S], source): the network's answer when usable, else plain MC with
        `photons` per set on ctx."""
        params = np.atleast_2d(np.asarray(params, np.float64))
        if self.usable:
            return predict_R(self.net, self.norm, params), "surrogate"
        if ctx is None or not photons:
            raise MCMLError(-1, "surrogate not validated within its bound; plain MC needs ctx and photons")
        _, _, R = simulate(ctx, params, np.full(len(params), int(photons)), seed, replay=False)
        return R, "mc"
