"""Independent pins of the oracle's absorbance binning and of the detector face's
reflection back into the tissue (CPU).

* Bins (transport.cu:580-600, 583, 588): known cells of the bin function at chosen points;
  the grids one replay books against grids rebuilt in float64 NumPy from the replay's own
  interaction trace (first drop to A0_z by depth, later drops to A_rz z-major with r
  fastest, r about the source axis, both clamped to the last bin); and the first drop of a
  normal-incidence pencil beam, whose depth follows from the photon's first 32-bit draw
  alone.
* Detector-face reflection (transport.cu:410-420): an undetected escaping photon re-enters
  with probability detector_reflectance, mirrored at the surface with its refracted
  direction.  Compared with a separately written float64 NumPy Monte Carlo of the same
  slab and fibres, per detector, within 4 standard errors.
"""
import ctypes as C

import numpy as np
import pytest

import fixtures as F
import oracle_ffi as O
from mcml_gpu_amd import abi


# --------------------------------------------------------------------- bin function
def _grid_sim(nr, nz, dr, dz, photons=1000, **kw):
    sim = F.hemisphere(photons, **kw)
    sim.grid_nr, sim.grid_nz, sim.grid_dr, sim.grid_dz = nr, nz, dr, dz
    return sim


def test_bin_known_cells():
    """The absorbance grid of the 5-layer absorbance workload (nr 200 x dr 0.02 cm, nz 400
    x dz 0.005 cm): cells of points chosen away from bin edges, beyond the grid, and off
    the axis in every quadrant."""
    d = O.derive(_grid_sim(200, 400, 0.02, 0.005))
    z_cases = [(0.0, 0), (0.0049, 0), (0.0051, 1), (0.0126, 2), (0.0174, 3), (1.0025, 200),
               (1.9949, 398), (1.9999, 399), (2.0, 399), (7.3, 399), (9.99, 399)]
    for z, zb in z_cases:
        assert O.bin_of(d, 0.0, 0.0, z)[0] == zb, (z, zb)
    r_cases = [((0.0, 0.0), 0), ((0.0199, 0.0), 0), ((0.0201, 0.0), 1), ((0.03, 0.04), 2),
               ((-0.03, -0.04), 2), ((0.04, -0.03), 2), ((0.0, 2.751), 137), ((0.0, 3.999), 199),
               ((3.0, 4.0), 199), ((-30.0, 0.0), 199)]
    for (x, y), rb in r_cases:
        assert O.bin_of(d, x, y, 0.3)[1] == rb, (x, y, rb)
    # depth and radius use their own cell sizes: swapping dr and dz moves these cells
    d2 = O.derive(_grid_sim(50, 80, 0.1, 0.01))
    assert O.bin_of(d2, 0.25, 0.0, 0.25) == (25, 2)
    assert O.bin_of(d2, 0.0, 0.55, 0.05) == (5, 5)


def _bins64(v, cell, n):
    q = np.asarray(v, np.float64) / float(np.float32(cell))
    b = np.clip(np.floor(q), 0, n - 1).astype(np.int64)
    near_edge = (np.abs(q - np.round(q)) < 1e-4) & (np.round(q) >= 1) & (np.round(q) <= n - 1)
    return b, near_edge


@pytest.mark.parametrize("grid", [(30, 20, 0.01, 0.01), (7, 64, 0.05, 0.002)])
def test_replay_grids_equal_rebinned_trace(grid):
    """Every replayed photon's grids equal the grids rebuilt from its interaction trace:
    the trace's first drop in A0_z[z bin], every later one in A_rz[z bin * nr + r bin],
    with float64 bins (and, within 1e-4 of a cell edge, either neighbouring cell), clamps at
    the last bin included (these grids are smaller than the photons' reach)."""
    nr, nz, dr, dz = grid
    n = 2000
    sim = _grid_sim(nr, nz, dr, dz, photons=n, mua=1.0, mus=100.0, g=0.9, n=1.4)
    d = O.derive(sim)
    rec, _ = O.first_pass(d, 31, 0, n)
    assert len(rec) > 200
    arz = np.zeros(nz * nr, np.uint64)
    a0 = np.zeros(nz, np.uint64)
    clamped_z = clamped_r = edges = 0
    for key in rec["key"][:150]:
        hit, r, tr = O.trace(d, 31, int(key), 1, grids=(arz, a0))
        assert hit == 1 and len(tr) > 0
        first = tr[:, 4]
        assert first[0] == 1 and not first[1:].any()  # first_scatter only on the first drop
        drops = tr[:, 3].astype(np.uint64)
        zb, ez = _bins64(tr[:, 2], dz, nz)
        rb, er = _bins64(np.hypot(tr[:, 0].astype(np.float64), tr[:, 1].astype(np.float64)), dr, nr)
        amb = ez | er
        if amb.any():  # take the oracle's cell where float64 is ambiguous, if it is adjacent
            for i in np.nonzero(amb)[0]:
                ozb, orb = O.bin_of(d, float(tr[i, 0]), float(tr[i, 1]), float(tr[i, 2]))
                assert abs(ozb - zb[i]) <= 1 and abs(orb - rb[i]) <= 1
                zb[i], rb[i] = ozb, orb
            edges += int(amb.sum())
        exp_a0 = np.zeros(nz, np.uint64)
        exp_arz = np.zeros(nz * nr, np.uint64)
        exp_a0[zb[0]] += drops[0]
        np.add.at(exp_arz, zb[1:] * nr + rb[1:], drops[1:])
        assert np.array_equal(a0, exp_a0)
        assert np.array_equal(arz, exp_arz)
        clamped_z += int((zb == nz - 1).sum())
        clamped_r += int((rb[1:] == nr - 1).sum())
    assert clamped_z > 0 and clamped_r > 0  # the clamps were exercised
    assert edges < 50


def _mwc_first_draw(seed, k):
    s, a = C.c_uint64(), C.c_uint32()
    O.lib().oracle_stream_state(seed, 0, k, C.byref(s), C.byref(a))
    t = a.value * (s.value & 0xFFFFFFFF) + (s.value >> 32)
    return t & 0xFFFFFFFF


def test_pencil_beam_first_drop_depth():
    """A normal-incidence pencil beam into an index-matched medium hops straight down by
    s1 = -ln(u) / (mua + mus), u = (2 (x1 & 0x7FFFFF) + 1) 2^-24 from the photon's first draw x1,
    and books its first drop there: A0_z[floor(s1 / dz)] holds mua / (mua + mus) of the
    start weight and no other A0_z cell is touched."""
    mua, mus, dz = 1.0, 99.0, 0.001
    sim = _grid_sim(10, 1000, 0.01, dz, photons=1, mua=mua, mus=mus, g=0.8, n=1.0)
    d = O.derive(sim)
    assert d.src_cos_a == 1.0 and d.layer[0].n == d.layer[1].n
    w0 = float(d.start_weight)
    arz = np.zeros(10 * 1000, np.uint64)
    a0 = np.zeros(1000, np.uint64)
    checked = 0
    for k in range(300):
        x1 = _mwc_first_draw(5, k)
        s1 = -np.log((2 * (x1 & 0x7FFFFF) + 1) / 2.0 ** 24) / (mua + mus)
        q = s1 / dz
        if abs(q - round(q)) < 1e-4:
            continue
        O.trace(d, 5, k, 1, grids=(arz, a0))
        nz_cells = np.nonzero(a0)[0]
        assert list(nz_cells) == [int(np.floor(q))], (k, s1, nz_cells)
        assert abs(float(a0[nz_cells[0]]) - w0 * mua / (mua + mus)) <= 1e-6 * w0
        checked += 1
    assert checked > 250


# --------------------------------------------------------------------- detector reflection
def numpy_slab_fibres(n, d, mua, mus, g, n_slab, n_out, fibres, rho, seed):
    """Float64 Monte Carlo of a slab (index n_slab between media of index n_out) under a
    normal pencil beam, with vertical fibres at the top surface: fibre i = (pos, radius,
    cos_crit) accepts an escaping photon whose exit point lies in its disk and whose
    refracted direction is within its acceptance cone.  An escaping photon no fibre accepts
    is reflected back by the probe face with probability rho, keeping its refracted
    direction (mirrored); otherwise it is lost.  Roulette at 1e-4 of the start weight.
    Returns the detected weight per fibre and photon, in units of the incident weight."""
    rng = np.random.default_rng(seed)
    mut = mua + mus
    rsp = ((n_out - n_slab) / (n_out + n_slab)) ** 2
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    ux, uy, uz = np.zeros(n), np.zeros(n), np.ones(n)
    w = np.full(n, 1 - rsp)
    det = np.zeros((len(fibres), n))
    alive = np.ones(n, bool)
    ratio = n_slab / n_out
    while alive.any():
        idx = np.nonzero(alive)[0]
        s = -np.log(rng.random(idx.size)) / mut
        zn = z[idx] + s * uz[idx]
        hit = (zn < 0) | (zn > d)
        hb = idx[hit]
        if hb.size:  # move to the surface the step crosses
            top = uz[hb] < 0
            zt = np.where(top, 0.0, d)
            sb = (zt - z[hb]) / uz[hb]
            x[hb] += sb * ux[hb]
            y[hb] += sb * uy[hb]
            z[hb] = zt
            ci = np.abs(uz[hb])
            st2 = ratio ** 2 * (1 - ci * ci)
            tir = st2 >= 1
            ct = np.sqrt(np.clip(1 - st2, 0, 1))
            rs = (n_slab * ci - n_out * ct) / (n_slab * ci + n_out * ct)
            rp = (n_slab * ct - n_out * ci) / (n_slab * ct + n_out * ci)
            R = np.where(tir, 1.0, 0.5 * (rs * rs + rp * rp))
            refl = rng.random(hb.size) <= R
            uz[hb[refl]] = -uz[hb[refl]]
            esc, top_esc, ct_esc = hb[~refl], top[~refl], ct[~refl]
            alive[esc[~top_esc]] = False  # out through the bottom
            up, ct_up = esc[top_esc], ct_esc[top_esc]
            if up.size:
                tx, ty = ux[up] * ratio, uy[up] * ratio  # refracted direction, upward
                caught = np.zeros(up.size, bool)
                for i, (pos, rad, cos_c) in enumerate(fibres):
                    acc = ((x[up] - pos) ** 2 + y[up] ** 2 <= rad * rad) & (ct_up >= cos_c)
                    det[i, up[acc]] += w[up[acc]]
                    caught |= acc
                alive[up[caught]] = False
                rest, tx, ty, ct_r = up[~caught], tx[~caught], ty[~caught], ct_up[~caught]
                back = rng.random(rest.size) <= rho
                alive[rest[~back]] = False
                b = rest[back]  # the probe face mirrors the refracted direction back down
                ux[b], uy[b], uz[b] = tx[back], ty[back], ct_r[back]
        hi = idx[~hit]
        if hi.size:
            x[hi] += s[~hit] * ux[hi]
            y[hi] += s[~hit] * uy[hi]
            z[hi] = zn[~hit]
            w[hi] *= mus / mut
            u = rng.random(hi.size)
            tmp = (1 - g * g) / (1 - g + 2 * g * u)
            ct_ = (1 + g * g - tmp * tmp) / (2 * g)
            st_ = np.sqrt(np.clip(1 - ct_ * ct_, 0, 1))
            psi = 2 * np.pi * rng.random(hi.size)
            cp, sp = np.cos(psi), np.sin(psi)
            ox, oy, oz = ux[hi], uy[hi], uz[hi]
            den = np.sqrt(np.clip(1 - oz * oz, 0, 1))
            small = den < 1e-10
            dd = np.where(small, 1.0, den)
            nx = np.where(small, st_ * cp, st_ * (ox * oz * cp - oy * sp) / dd + ox * ct_)
            ny = np.where(small, st_ * sp, st_ * (oy * oz * cp + ox * sp) / dd + oy * ct_)
            nz = np.where(small, np.sign(oz) * ct_, -st_ * cp * den + oz * ct_)
            ux[hi], uy[hi], uz[hi] = nx, ny, nz
            low = hi[w[hi] < 1e-4]
            if low.size:
                survive = rng.random(low.size) < 0.1
                w[low[survive]] *= 10
                w[low[~survive]] = 0
                alive[low[~survive]] = False
    return det


@pytest.mark.parametrize("rho", [0.35, 1.0])
def test_detector_face_reflection_matches_independent_numpy_mc(rho):
    """A slab (d 0.02 cm, mua 1, mus 60, g 0.75, index 1.5 in air) read by three vertical
    fibres (NA 0.3 on the axis, NA 0.9 at 0.06 cm, NA 1 at 0.13 cm) whose faces reflect
    with probability rho.  Per fibre, the oracle's reflectance equals the NumPy Monte
    Carlo's within 4 combined standard errors, and the three deviations together stay
    within chi-square(3) at p = 1e-4.  The outer fibres see far more weight than without
    reflection (so the re-entry branch is what is compared); re-entering with the incident
    instead of the refracted direction moves them by about 5 standard errors each, which
    the chi-square bound rejects."""
    n = 1000000
    fib = [(0.0, 0.3, 0.02), (0.06, 0.9, 0.03), (0.13, 1.0, 0.04)]  # (pos, NA, radius)
    prb = F.probes((0, 0, 0), [(p, na, r, 0.0) for (p, na, r) in fib])
    layers = [(0.02, 1.0, 60.0, 1.5, 0.75), (0.0, 0.0, 1e-6, 1.0, 0.0)]

    def oracle_R(refl):
        sim = abi.make_sim(F.base_cfg(n, 2, 1.0, 1.0, 1, 1, prb, refl=refl), layers)
        rec, _ = O.first_pass(O.derive(sim), 404, 0, n)
        out = np.zeros((len(fib), n))
        for s in range(1, len(fib) + 1):
            m = rec["sds"] == s
            np.add.at(out[s - 1], (rec["key"][m] & ((1 << 40) - 1)).astype(np.int64),
                      rec["weight"][m].astype(np.float64) / 4294967295.0)
        return out

    got = oracle_R(rho)
    fibres = [(p, r, float(np.cos(np.arcsin(na / 1.0)))) for (p, na, r) in fib]
    ref = numpy_slab_fibres(n, 0.02, 1.0, 60.0, 0.75, 1.5, 1.0, fibres, rho, 8)
    none = oracle_R(0.0)
    se = np.hypot(got.std(axis=1), ref.std(axis=1)) / np.sqrt(n)
    z = (got.mean(axis=1) - ref.mean(axis=1)) / se
    assert np.all(np.abs(z) < 4), (rho, z, got.mean(axis=1), ref.mean(axis=1))
    assert (z * z).sum() < 21.1, (rho, z)
    gain = (got.mean(axis=1) - none.mean(axis=1)) / se
    assert gain[1] > 4 and gain[2] > 4, (rho, gain)
