"""Pins of the oracle functions that the closed forms of test_oracle.py leave open (CPU only).

Each expected value below is computed in float64 numpy from the reference's own formula at
the cited lines, vectorised and structured independently of the oracle's C code. Every
case is built so that each term of the formula moves the result, and each test also
evaluates plausible slips of the formula (a dropped term, a wrong constant, a swapped index,
skipped instead of clamped neighbours ...) and asserts that the oracle is far from every
one of them, so a test that passes shows that the oracle has none of those slips.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from speedball_amd import scenegen as sg
from speedball_amd.scene import SceneBuilder, directional_light, point_light, spot_light, uber_material

LUMA = np.array([0.2126, 0.7152, 0.0722])
INTERIOR = [(ly + 1) * 8 + lx + 1 for ly in range(6) for lx in range(6)]


def _field(res=(3, 3, 3), rays=64, depth_oct=6, prec="64"):
    w = sg.room_boxes(res=res, rays=rays)
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, res, rays, depth_oct, prec=prec)
    return w, o


def _far(expected, got, rel):
    """max relative distance between a slip's prediction and the oracle's output"""
    e = np.asarray(expected, np.float64)
    g = np.asarray(got, np.float64)
    return float(np.max(np.abs(e - g) / np.maximum(np.abs(e), 1e-12))) > rel


# ------------------------------------------------------------- upload bilateral filter
def _ref_filter(rgb, m2, strength=1.0, smooth=0.5, k_scale=1.0, rel_scale=1.0, eps_scale=1.0, slip=None):
    """uploadKernel's 3x3 filter of one 6x6 interior (js/gi_probes.js:2755-2790):
    Gaussian exp(-(jx^2+jy^2)/2) times the luma edge stop exp(-dL^2 / max(var kEff +
    max(EPS, L^2 relEff), 1e-8)), kEff = 8 smW, relEff = .0225 smW, smW = 1 + 6 smooth,
    neighbours clamped to the interior; mix(centre, filtered, strength)."""
    smw = 1.0 + smooth * 6.0
    if slip == "smooth_factor":
        smw = 1.0 + smooth
    k_eff, rel_eff, eps = 8.0 * smw * k_scale, 0.0225 * smw * rel_scale, 1e-3 * eps_scale
    if slip == "eps_dropped":
        eps = 0.0
    if slip == "k_rel_swapped":
        k_eff, rel_eff = rel_eff, k_eff
    lum = rgb @ LUMA
    var = np.maximum(m2 - lum * lum, 0.0)
    if slip == "no_variance":
        var = np.zeros_like(var)
    out = np.zeros_like(rgb)
    for sy in range(6):
        for sx in range(6):
            den = max(var[sy, sx] * k_eff + max(eps, lum[sy, sx] ** 2 * rel_eff), 1e-8)
            acc, ws = np.zeros(3), 0.0
            for jy in (-1, 0, 1):
                for jx in (-1, 0, 1):
                    gw = math.exp(-(jx * jx + jy * jy) * (1.0 if slip == "gauss_sigma" else 0.5))
                    nx, ny = sx + jx, sy + jy
                    if slip == "skip_border" and not (0 <= nx < 6 and 0 <= ny < 6):
                        continue
                    nx, ny = min(max(nx, 0), 5), min(max(ny, 0), 5)
                    dl = lum[ny, nx] - lum[sy, sx]
                    wgt = gw * math.exp(-dl * dl / den)
                    acc += rgb[ny, nx] * wgt
                    ws += wgt
            filt = acc / max(ws, 1e-4)
            t = (1.0 - strength) if slip == "mix_reversed" else strength
            out[sy, sx] = rgb[sy, sx] + (filt - rgb[sy, sx]) * t
    return out


def _upload_tile(o, p, rgb, m2, probe=0):
    """Write one probe's interior irradiance state, run the oracle's upload, return the
    probe's 6x6 interior from the atlas (rows = ly)."""
    irr = o.irr_view()
    irr[:] = 0.0
    for t, idx in enumerate(INTERIOR):
        irr[probe, idx, :3] = rgb[t // 6, t % 6]
        irr[probe, idx, 3] = m2[t // 6, t % 6]
    o.upload(p)
    a = o.atlas_view()
    rx, ry = o.res[0], o.res[1]
    col, row = probe % rx, (probe // (rx * ry)) * ry + (probe // rx) % ry
    return a[row * 8 + 1:row * 8 + 7, col * 8 + 1:col * 8 + 7, :3].copy()


COMMON_SLIPS = ("k_rel_swapped", "gauss_sigma", "skip_border", "mix_reversed")
FILTER_REGIMES = {
    # variance-dominated edge stop: var kEff ~ 1.6 >> L^2 relEff
    "variance": (lambda rng: (rng.uniform(0.5, 1.5, (6, 6, 3)), 0.05), COMMON_SLIPS + ("smooth_factor", "no_variance")),
    # relative-luma-dominated: no variance, L ~ 1..3
    "relative": (lambda rng: (rng.uniform(0.6, 3.0, (6, 6, 3)), 0.0), COMMON_SLIPS + ("smooth_factor",)),
    # epsilon-dominated: L ~ 0.03, L^2 relEff ~ 1e-4 < EPS
    "epsilon": (lambda rng: (rng.uniform(0.0, 0.06, (6, 6, 3)), 0.0), COMMON_SLIPS + ("eps_dropped",)),
}


@pytest.mark.parametrize("regime", sorted(FILTER_REGIMES))
def test_filter_default_strength_closed_form(regime):
    """The default filter (filterStrength 1, filterSmooth 0.5) on a coloured tile whose
    luma contrast makes the Gaussian, the edge stop and the clamped borders all matter, in
    the three regimes of the edge-stop denominator. The oracle's FP16 store must equal the
    formula's value rounded to FP16 (one binary16 step allowed at a rounding tie), and
    must be at least 4 steps away from every listed slip at some texel."""
    rng = np.random.default_rng({"variance": 1, "relative": 2, "epsilon": 3}[regime])
    make, slips = FILTER_REGIMES[regime]
    rgb, extra_var = make(rng)
    lum = rgb @ LUMA
    m2 = lum * lum + extra_var
    _, o = _field()
    p = o.params(frame_jitter=0.1)
    got = _upload_tile(o, p, rgb, m2, probe=13)
    want = _ref_filter(rgb, m2)
    half = np.float16(want).astype(np.float64)
    ulp = np.abs(np.spacing(np.float16(want)).astype(np.float64))
    assert np.all(np.abs(got - half) <= ulp)
    assert not np.allclose(want, rgb, rtol=1e-2)  # the filter does act on this tile
    for slip in slips:
        bad = _ref_filter(rgb, m2, slip=slip)
        assert _far(bad, got, 4 * 2.0 ** -10), slip


def test_filter_partial_strength_and_debug_scales():
    """filterStrength 0.4 mixes 40 % of the filtered value in; the debug K / REL / EPS
    scales multiply their constants."""
    rng = np.random.default_rng(8)
    rgb = rng.uniform(0.2, 2.0, (6, 6, 3))
    lum = rgb @ LUMA
    m2 = lum * lum + 0.02
    _, o = _field()
    p = o.params(frame_jitter=0.1)
    p.filter_strength, p.filter_smooth = 0.4, 0.25
    p.dbg_filter_k, p.dbg_filter_rel, p.dbg_filter_eps = 0.5, 3.0, 10.0
    got = _upload_tile(o, p, rgb, m2, probe=5)
    want = _ref_filter(rgb, m2, strength=0.4, smooth=0.25, k_scale=0.5, rel_scale=3.0, eps_scale=10.0)
    ulp = np.abs(np.spacing(np.float16(want)).astype(np.float64))
    assert np.all(np.abs(got - np.float16(want).astype(np.float64)) <= ulp)
    assert _far(_ref_filter(rgb, m2, strength=0.4, smooth=0.25), got, 4 * 2.0 ** -10)


def test_filter_keeps_a_luma_step():
    """Two flat regions whose lumas differ by 100x: across the step the edge stop is
    exp(-(9.9)^2 / (100 x .09)) ~ 2e-5 of the Gaussian, so each side keeps its own value
    (a filter without the edge stop would blur the step by tens of percent)."""
    rgb = np.zeros((6, 6, 3))
    rgb[:, :3] = 0.1
    rgb[:, 3:] = 10.0
    m2 = (rgb @ LUMA) ** 2
    _, o = _field()
    got = _upload_tile(o, o.params(), rgb, m2)
    np.testing.assert_allclose(got[:, :3], np.float16(0.1), rtol=2e-3)
    np.testing.assert_allclose(got[:, 3:], 10.0, rtol=2e-3)


# ------------------------------------------------------------- bounce tile addressing
def _ref_sample_atlas(atlas, res, gmin, gsize, pos, nrm, slip=None):
    """sampleAtlas (js/gi_probes.js:1941-1974): 8 trilinear probe taps with w + 1e-4, each
    a texel-centred bilinear fetch (clamp to edge) at tile (col = px, row = pz ry + py),
    texel (1 + 6 u, 1 + 6 v) with (u, v) = octEncode(n)."""
    res = np.asarray(res, np.float64)
    cell = np.asarray(gsize, np.float64) / np.maximum(1.0, res - 1.0)
    gf = (pos - np.asarray(gmin, np.float64)) / np.maximum(cell, 1e-6)
    base = np.clip(np.floor(gf), 0.0, np.maximum(res - 2.0, 0.0))
    fr = np.clip(gf - base, 0.0, 1.0)
    n = nrm / np.linalg.norm(nrm)
    s = np.abs(n).sum()
    ox, oy = n[0] / s, n[1] / s
    if n[2] < 0:
        ox, oy = (1 - abs(oy)) * (1 if ox >= 0 else -1), (1 - abs(ox)) * (1 if oy >= 0 else -1)
    u, v = ox * 0.5 + 0.5, oy * 0.5 + 0.5
    h, w = atlas.shape[:2]
    acc, ws = np.zeros(3), 0.0
    for i in range(8):
        d = np.array([i & 1, (i >> 1) & 1, (i >> 2) & 1], np.float64)
        pr = base + d
        wt = np.prod(np.where(d > 0, fr, 1.0 - fr)) + (0.0 if slip == "no_eps" else 1e-4)
        col, row = pr[0], pr[2] * res[1] + pr[1]
        if slip == "row_swap":
            row = pr[1] * res[2] + pr[2]
        border = 0.0 if slip == "no_border" else 1.0
        x = col * 8 + border + u * 6 - 0.5
        y = row * 8 + border + v * 6 - 0.5
        if slip == "uv_swap":
            x, y = col * 8 + border + v * 6 - 0.5, row * 8 + border + u * 6 - 0.5
        x0, y0 = math.floor(x), math.floor(y)
        fx, fy = x - x0, y - y0
        xa, xb = min(max(x0, 0), w - 1), min(max(x0 + 1, 0), w - 1)
        ya, yb = min(max(y0, 0), h - 1), min(max(y0 + 1, 0), h - 1)
        t = (atlas[ya, xa] * (1 - fx) * (1 - fy) + atlas[ya, xb] * fx * (1 - fy) +
             atlas[yb, xa] * (1 - fx) * fy + atlas[yb, xb] * fx * fy)
        acc += t * wt
        ws += wt
    return acc / max(ws, 1e-4)


def _hit_frames(w, o, scene):
    """Hit point (with its normal offset) and face-forwarded normal of every hit ray, from
    the triangle and t (js/gi_probes.js:1993-2017; one identity instance)."""
    rays, hits = o.rays_view(), o.hits_view()
    v = scene.vertex_data.reshape(-1, 8)[:, :3].astype(np.float64)
    ti = scene.tri_index.reshape(-1, 3)
    res = np.asarray(w.res)
    cell = min(np.asarray(w.grid_size, np.float64) / np.maximum(res - 1, 1))
    bias = max(max(cell, 1e-4) * 0.005, 1e-3)
    out = []
    for slot in range(hits.shape[0]):
        ix, iy, iz = slot % res[0], (slot // res[0]) % res[1], slot // (res[0] * res[1])
        ro = np.asarray(w.grid_min) + np.array([ix, iy, iz]) / np.maximum(res - 1, 1) * np.asarray(w.grid_size)
        for k in range(hits.shape[1]):
            tri = hits[slot, k, 0]
            if tri < 0:
                continue
            rd = oracle.ray_dir(k, w.rays, w.frame_jitter)
            p0, p1, p2 = v[ti[tri]]
            n = np.cross(p1 - p0, p2 - p0)
            n /= np.linalg.norm(n)
            if n @ rd >= 0:
                n = -n
            out.append((slot, k, ro + rd * rays[slot, k, 3] + n * bias, n, tri))
    return out


def test_bounce_reads_the_right_probe_tile_and_texel():
    """With no lights and a black sky every hit's radiance is the bounce kd/pi x
    sampleAtlas(previous atlas) (roll = 1 below radianceClamp). The previous atlas holds a
    distinct random value in every texel of every probe tile, so the tile (col = ix,
    row = iz ry + iy), the texel offset (border 1, 6 texels), the trilinear corner weights
    (+1e-4 each) and the bilinear weights are all pinned: each listed slip picks up other
    texels' values on most hits."""
    import dataclasses
    w = sg.room_boxes(res=(4, 3, 5), rays=64)
    scene = w.scene
    dark = dataclasses.replace(scene, lights=np.zeros(0, np.float32), _view=None)
    o = oracle.OracleField(dark, w.grid_min, w.grid_size, w.res, w.rays, 6, prec="64")
    rng = np.random.default_rng(21)
    a = o.atlas_view()
    a[..., :3] = rng.uniform(0.0, 1.0, a[..., :3].shape)
    p = o.params(frame_jitter=w.frame_jitter)
    o.trace(p)
    atlas = a[..., :3].astype(np.float64)
    mats = scene.materials.reshape(-1)
    frames = _hit_frames(w, o, scene)
    assert len(frames) > 2000
    assert len({tuple(np.round(f[3], 3)) for f in frames}) > 20  # walls, floor, ceiling, rotated boxes
    rays = o.rays_view()
    slips = dict.fromkeys(("no_border", "row_swap", "uv_swap", "no_eps"), 0)
    checked = 0
    for slot, k, pos, n, tri in frames[::5]:
        m = int(scene.tri_material[tri])
        kd = np.clip(mats[m * 28:m * 28 + 3].astype(np.float64) * (1 - mats[m * 28 + 4]) * (1 - mats[m * 28 + 5]),
                     0.0, 0.95)
        want = _ref_sample_atlas(atlas, w.res, w.grid_min, w.grid_size, pos, n) * kd / math.pi
        np.testing.assert_allclose(rays[slot, k, :3], want, rtol=1e-9, atol=1e-12)
        checked += 1
        for slip in slips:
            bad = _ref_sample_atlas(atlas, w.res, w.grid_min, w.grid_size, pos, n, slip=slip) * kd / math.pi
            slips[slip] += int(np.max(np.abs(bad - rays[slot, k, :3]) / np.maximum(want, 1e-12)) > 1e-7)
    for slip, n_off in slips.items():
        assert n_off > checked // 2, (slip, n_off, checked)


# ------------------------------------------------------------- depth moments
@pytest.mark.parametrize("depth_oct", [6, 14])
def test_depth_moments_use_the_sharpened_cosine_weight(depth_oct):
    """First update: meanR = sum dw d / sum dw and meanR2 = sum dw d^2 / sum dw with
    dw = cw^depthSharpness and cw = max(dir . r, 0)^cosinePower, a miss (hitT -1) counting
    as maxDist = |gridSize| (js/gi_probes.js:2366-2385, 2425-2432, 3209). cosinePower 2 and
    depthSharpness 5 separate cw^5 from cd^5, and the irradiance normaliser from the depth
    normaliser."""
    w, o = _field(res=(3, 2, 2), depth_oct=depth_oct)
    P, R = 12, 64
    rng = np.random.default_rng(31 + depth_oct)
    o.ray_data = np.zeros(P * R * 4, np.float64)
    d = rng.uniform(0.2, 9.0, (P, R))
    miss = rng.uniform(size=(P, R)) < 0.25
    d[miss] = -1.0
    o.ray_data[3::4] = d.reshape(-1)
    o.ray_data[0::4] = 0.5
    p = o.params(frame_jitter=0.35)
    p.dbg_cosine_power, p.depth_sharpness = 2.0, 5.0
    o.blend(p)
    # the maxDist uniform is a float: |gridSize| rounded to f32
    max_dist = float(np.float32(np.linalg.norm(np.asarray(w.grid_size, np.float32).astype(np.float64))))
    dirs = np.array([oracle.ray_dir(k, R, 0.35) for k in range(R)])
    dist = np.where(miss, max_dist, d)
    n = depth_oct
    tile = 8 if n == 6 else n + 2
    got = o.depth_view()
    checked = 0
    for t in range(n * n):
        tx, ty = t % n, t // n
        td = oracle.oct_decode((tx + 0.5) / n, (ty + 0.5) / n)
        cd = np.maximum(dirs @ td, 0.0)
        cw = cd ** 2
        dw = cw ** 5
        m1 = (dist * dw).sum(1) / max(dw.sum(), 1e-4)
        m2 = (dist * dist * dw).sum(1) / max(dw.sum(), 1e-4)
        g = got[:, (ty + 1) * tile + tx + 1]
        np.testing.assert_allclose(g[:, 0], m1, rtol=1e-11)
        np.testing.assert_allclose(g[:, 1], m2, rtol=1e-11)
        # slips: the plain cosine as depth weight, misses as -1, the raw cd to the power
        for bad in ((dist * cw).sum(1) / cw.sum(), (np.where(miss, -1.0, d) * dw).sum(1) / dw.sum(),
                    (dist * cd ** 5).sum(1) / (cd ** 5).sum()):
            checked += int(np.max(np.abs(bad - g[:, 0]) / m1) > 1e-3)
    assert checked >= n * n  # every texel rules out at least one slip; most rule out all


@pytest.mark.parametrize("depth_oct", [6, 14])
def test_miss_depth_is_the_grid_diagonal(depth_oct):
    """A probe field far from all geometry: every ray misses, so every depth texel's
    moments are (maxDist, maxDist^2) with maxDist = |gridSize| (js/gi_probes.js:3209), on a
    non-cubic grid where the longest side (12) or the side sum (19) would differ from the
    diagonal (13)."""
    w = sg.room_boxes(res=(2, 2, 2), rays=64)
    o = oracle.OracleField(w.scene, (500.0, 500.0, 500.0), (3.0, 4.0, 12.0), (2, 2, 2), 64, depth_oct, prec="64")
    p = o.params(frame_jitter=0.2)
    o.update(p)
    assert np.all(o.hits_view()[..., 0] < 0)
    n = depth_oct
    tile = 8 if n == 6 else n + 2
    idx = [(ly + 1) * tile + lx + 1 for ly in range(n) for lx in range(n)]
    dv = o.depth_view()[:, idx]
    np.testing.assert_allclose(dv[..., 0], 13.0, rtol=1e-12)
    np.testing.assert_allclose(dv[..., 1], 169.0, rtol=1e-12)


# ------------------------------------------------------------- temporal policy
def _ref_policy(prev, prev_m2, mean, cur_m2, hyst, hexp=1.0, sigma1=2.5, h_drop=0.30, clamp_sigma=6.0,
                noise_boost=0.25, sigma0=0.75, min_change_h=0.55, var_eps=2.5e-5, var_rel=2.5e-3, slip=None):
    """The irradiance temporal policy of one texel (js/gi_probes.js:2436-2493): returns
    (blended rgb, blended M2). The knobs are float uniforms, so they are rounded to f32."""
    f32 = lambda x: float(np.float32(x))
    hyst, hexp, sigma1, h_drop, clamp_sigma = f32(hyst), f32(hexp), f32(sigma1), f32(h_drop), f32(clamp_sigma)
    noise_boost, sigma0, min_change_h = f32(noise_boost), f32(sigma0), f32(min_change_h)
    var_eps, var_rel = f32(var_eps), f32(var_rel)
    prev, mean = np.asarray(prev, np.float64), np.asarray(mean, np.float64)
    was_black = prev.sum() < 1e-6
    pl, cl = prev @ LUMA, mean @ LUMA
    pv, cv = max(prev_m2 - pl * pl, 0.0), max(cur_m2 - cl * cl, 0.0)
    lref = max(cl, pl, 0.0)
    floor = var_eps + lref * lref * var_rel
    sigma = math.sqrt(pv + cv + floor)
    ad = abs(cl - pl)
    s0, s1 = sigma * sigma0, sigma * max(sigma1, sigma0 + 0.01)
    cwt = min(max((ad - s0) / max(s1 - s0, 1e-6), 0.0), 1.0)
    noise_h = hyst + (1 - hyst) * noise_boost
    change_h = max(hyst - h_drop, min_change_h)
    raw = noise_h * (1 - cwt) + change_h * cwt
    if slip == "no_change_weight":
        raw = noise_h
    h = 0.0 if was_black else min(max(raw, 1e-6), 1.0) ** hexp
    band_var = pv + cv if slip == "band_with_cur_var" else pv
    band = math.sqrt(band_var + floor) * clamp_sigma
    cs = min(1.0, band / max(ad, 1e-6))
    if slip == "no_clamp":
        cs = 1.0
    cand = mean if was_black else prev + (mean - prev) * cs
    out = cand * (1 - h) + prev * h
    return out, cur_m2 * (1 - h) + prev_m2 * h, cwt, cs


POLICY_CASES = {
    # a large jump: full change weight, the firefly clamp cuts the step to its band
    "jump": dict(prev=(0.1, 0.1, 0.1), prev_var=0.0, cur=(1.0, 1.0, 1.0), hyst=0.97, hexp=1.0),
    # a coloured jump with prior variance: clamp active, per channel on one luma scale
    "colour": dict(prev=(0.2, 0.4, 0.1), prev_var=0.004, cur=(1.6, 0.9, 1.2), hyst=0.9, hexp=1.0),
    # a moderate change: 0 < changeW < 1, clamp inactive
    "moderate": dict(prev=(0.5, 0.5, 0.5), prev_var=0.01, cur=(0.65, 0.65, 0.65), hyst=0.97, hexp=1.0),
    # gated default hysteresis 0.6: the change retention floors at minChangeH 0.55, exponent 2
    "floor": dict(prev=(0.3, 0.2, 0.25), prev_var=0.0, cur=(2.0, 2.0, 2.0), hyst=0.6, hexp=2.0),
}


@pytest.mark.parametrize("case", sorted(POLICY_CASES))
def test_temporal_policy_change_weight_and_firefly_clamp(case):
    """Second update of a field whose previous state is set directly: uniform current
    radiance (curVar 0) against a previous value with a given prior variance, through the
    change-weighted retention and the firefly clamp banded by the PRIOR variance only. The
    depth moments follow with h_d = clamp(hyst, 0, .999)^exp."""
    c = POLICY_CASES[case]
    w, o = _field(res=(3, 3, 3))
    P, R = 27, 64
    cur = np.asarray(c["cur"], np.float64)
    o.ray_data = np.zeros(P * R * 4, np.float64)
    for ch in range(3):
        o.ray_data[ch::4] = cur[ch]
    o.ray_data[3::4] = 4.0
    prev = np.asarray(c["prev"], np.float64)
    pl = prev @ LUMA
    prev_m2 = pl * pl + c["prev_var"]
    irr = o.irr_view()
    irr[:, INTERIOR, :3] = prev
    irr[:, INTERIOR, 3] = prev_m2
    dep = o.depth_view()
    dep[:, INTERIOR, 0] = 2.0
    dep[:, INTERIOR, 1] = 5.0
    p = o.params(frame_jitter=0.4, hysteresis=c["hyst"])
    p.hysteresis_exponent = c["hexp"]
    o.blend(p)
    cl = cur @ LUMA
    want, want_m2, cwt, cs = _ref_policy(prev, prev_m2, cur, cl * cl, c["hyst"], c["hexp"])
    got = o.irr_view()[:, INTERIOR]
    np.testing.assert_allclose(got[..., :3], np.broadcast_to(want, got[..., :3].shape), rtol=1e-12)
    np.testing.assert_allclose(got[..., 3], want_m2, rtol=1e-12)
    dh = min(max(float(np.float32(c["hyst"])), 0.0), 0.999) ** c["hexp"]
    np.testing.assert_allclose(o.depth_view()[:, INTERIOR, 0], 4.0 * (1 - dh) + 2.0 * dh, rtol=1e-12)
    np.testing.assert_allclose(o.depth_view()[:, INTERIOR, 1], 16.0 * (1 - dh) + 5.0 * dh, rtol=1e-12)
    # the case exercises the branch it names, and every slip would move the result
    if case in ("jump", "colour", "floor"):
        assert cwt == 1.0 and cs < 0.9
    if case == "moderate":
        assert 0.05 < cwt < 0.95 and cs == 1.0
    slips = ["no_change_weight"] + (["no_clamp"] if cs < 1 else [])
    for slip in slips:
        bad, _, _, _ = _ref_policy(prev, prev_m2, cur, cl * cl, c["hyst"], c["hexp"], slip=slip)
        assert _far(bad, got[0, 0, :3], 1e-4), slip


def test_temporal_policy_clamp_ignores_current_variance():
    """The firefly band uses the prior variance only (js/gi_probes.js:2475-2480): a noisy
    current update (large curVar) does not widen its own clamp."""
    w, o = _field(res=(3, 3, 3))
    P, R = 27, 64
    rng = np.random.default_rng(41)
    o.ray_data = np.zeros(P * R * 4, np.float64)
    g = np.where(rng.uniform(size=(P, R)) < 0.5, 0.2, 3.0)
    for ch in range(3):
        o.ray_data[ch::4] = g.reshape(-1)
    o.ray_data[3::4] = 4.0
    prev = np.array([0.3, 0.3, 0.3])
    irr = o.irr_view()
    irr[:, INTERIOR, :3] = prev
    irr[:, INTERIOR, 3] = 0.09
    o.depth_view()[:, INTERIOR, 0] = 1.0
    p = o.params(frame_jitter=0.4, hysteresis=0.95)
    o.blend(p)
    dirs = np.array([oracle.ray_dir(k, R, 0.4) for k in range(R)])
    got = o.irr_view()
    n_far = 0
    for t, idx in enumerate(INTERIOR):
        td = oracle.oct_decode((t % 6 + 0.5) / 6, (t // 6 + 0.5) / 6)
        cw = np.maximum(dirs @ td, 0.0)
        for probe in (0, 11, 26):
            mean = (g[probe] * cw).sum() / cw.sum()
            m2 = (g[probe] ** 2 * LUMA.sum() ** 2 * cw).sum() / cw.sum()
            want, want_m2, _, _ = _ref_policy(prev, 0.09, np.full(3, mean), m2, 0.95)
            np.testing.assert_allclose(got[probe, idx, :3], want, rtol=1e-11)
            np.testing.assert_allclose(got[probe, idx, 3], want_m2, rtol=1e-11)
            bad, _, _, _ = _ref_policy(prev, 0.09, np.full(3, mean), m2, 0.95, slip="band_with_cur_var")
            n_far += int(_far(bad, got[probe, idx, :3], 1e-4))
    assert n_far > 50


# ------------------------------------------------------------- NEE light types
def _floor_scene(light, kd=0.5):
    b = SceneBuilder()
    m = b.add_material(uber_material((kd, kd, kd), side=2))
    v = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    b.add_instance(b.add_mesh(v, t, tri_material=np.full(2, m, np.uint32)))
    b.add_light(light)
    return b.finalize()


def _floor_hits(light, res=(3, 2, 3), gmin=(-2.0, 0.5, -2.0), gsize=(4.0, 1.0, 4.0), rays=256):
    """Trace a probe grid above a double-sided floor lit by one light (FP64 oracle); return
    per floor hit (hit point with its normal offset, radiance)."""
    scene = _floor_scene(light)
    o = oracle.OracleField(scene, gmin, gsize, res, rays, 6, prec="64")
    o.trace(o.params(frame_jitter=0.15))
    res_a = np.asarray(res)
    cell = min(np.asarray(gsize) / np.maximum(res_a - 1, 1))
    bias = max(cell * 0.005, 1e-3)
    pts, rad = [], []
    for slot in range(int(np.prod(res))):
        ix, iy, iz = slot % res[0], (slot // res[0]) % res[1], slot // (res[0] * res[1])
        ro = np.asarray(gmin) + np.array([ix, iy, iz]) / np.maximum(res_a - 1, 1) * np.asarray(gsize)
        for k in range(rays):
            if o.hits_view()[slot, k, 0] < 0:
                continue
            rd = oracle.ray_dir(k, rays, 0.15)
            pts.append(ro + rd * o.rays_view()[slot, k, 3] + np.array([0.0, 1.0, 0.0]) * bias)
            rad.append(o.rays_view()[slot, k, :3].copy())
    return np.array(pts), np.array(rad)


def test_nee_directional_light():
    """Directional light: toLight = -dir, no distance falloff (js/gi_probes.js:2146-2170):
    every floor hit has radiance kd/pi (n . -dir) I."""
    rec = directional_light((0.0, 10.0, 0.0), (1.5, 0.0, -0.9), color=(1.0, 0.5, 0.25), intensity=2.0)
    pts, rad = _floor_hits(rec)
    d = rec[4:7].astype(np.float64)
    ndl = max(0.0, -d[1] / np.linalg.norm(d))
    want = 0.5 / math.pi * ndl * rec[7:10].astype(np.float64)
    assert len(pts) > 500
    np.testing.assert_allclose(rad, np.broadcast_to(want, rad.shape), rtol=1e-9)


def test_nee_spot_light_cone_and_penumbra():
    """Spot light (unranged, decay 2): falloff 1/d^2 times the linear cone
    clamp((cos_axis - cosAngle) / (cosPenumbra - cosAngle), 0, 1) (js/gi_probes.js:2163-2170);
    the hits cover the full cone, the penumbra ramp and the outside."""
    pos = np.array([0.3, 3.0, -0.2])
    rec = spot_light(pos, (0.6, 0.0, 0.1), color=(0.9, 0.8, 0.7), intensity=12.0, angle=0.5, penumbra=0.4)
    pts, rad = _floor_hits(rec)
    lp = rec[1:4].astype(np.float64)
    axis = rec[4:7].astype(np.float64)
    col = rec[7:10].astype(np.float64)
    ca, cp = float(rec[12]), float(rec[13])
    tl = lp - pts
    dist = np.linalg.norm(tl, axis=1)
    wi = tl / dist[:, None]
    ndl = np.maximum(wi[:, 1], 0.0)
    cos_axis = -(wi @ axis)
    spot = np.clip((cos_axis - ca) / max(cp - ca, 1e-4), 0.0, 1.0)
    want = 0.5 / math.pi * (ndl * spot / np.maximum(dist ** 2, 0.01))[:, None] * col
    np.testing.assert_allclose(rad, want, rtol=1e-9, atol=1e-15)
    assert (spot >= 1).sum() > 50 and ((spot > 0) & (spot < 1)).sum() > 50 and (spot <= 0).sum() > 50
    # the cone is not the smoothstep of three's renderer: it would move the penumbra hits
    ss = np.clip((cos_axis - ca) / (cp - ca), 0, 1)
    ss = ss * ss * (3 - 2 * ss)
    pen = (spot > 0.1) & (spot < 0.9)
    assert np.all(np.abs(ss[pen] - spot[pen]) > 1e-3)


def test_nee_ranged_point_light_window_and_gate():
    """Ranged point light (distance 3, decay 1.5): falloff 1/max(d^decay, .01) times the
    window (clamp(1 - (d/r)^4))^2, and no contribution beyond the range
    (js/gi_probes.js:2155-2162)."""
    rec = point_light((0.2, 1.4, 0.1), color=(1.0, 0.7, 0.4), intensity=5.0, distance=3.0, decay=1.5)
    pts, rad = _floor_hits(rec)
    lp = rec[1:4].astype(np.float64)
    tl = lp - pts
    dist = np.linalg.norm(tl, axis=1)
    ndl = np.maximum(tl[:, 1] / dist, 0.0)
    win = np.clip(1 - (dist / 3.0) ** 4, 0.0, 1.0)
    want = 0.5 / math.pi * (ndl * win * win / np.maximum(dist ** 1.5, 0.01))[:, None] * rec[7:10].astype(np.float64)
    want[dist >= 3.0] = 0.0
    np.testing.assert_allclose(rad, want, rtol=1e-9, atol=1e-15)
    assert (dist < 2.0).sum() > 50 and ((dist > 2.5) & (dist < 3.0)).sum() > 20 and (dist >= 3.0).sum() > 50
    # decay 2 or no window would give other values inside the range
    inside = dist < 2.9
    alt = 0.5 / math.pi * (ndl / np.maximum(dist ** 2, 0.01))[:, None] * rec[7:10].astype(np.float64)
    assert np.all(np.abs(alt[inside] - rad[inside]).max(axis=1) > 1e-6)


def test_nee_visibility_mask_follows_an_occluder():
    """Per-ray light visibility (OracleField.trace(vis=True)): a point light above a
    double-sided floor and a square occluder between them. A floor hit's bit is set exactly
    when the segment from its offset hit point to the light misses the square (computed
    here analytically); a hit on the square's underside sees the light from behind
    (n . l < 0), so its bit is clear, and a miss has no bit."""
    b = SceneBuilder()
    m = b.add_material(uber_material((0.5, 0.5, 0.5), side=2))
    fv = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    a = 0.6
    sv = np.array([[-a, 1.5, -a], [a, 1.5, -a], [a, 1.5, a], [-a, 1.5, a]], np.float32)
    quad = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    b.add_instance(b.add_mesh(np.concatenate([fv, sv]), np.concatenate([quad, quad + 4]),
                              tri_material=np.full(4, m, np.uint32)))
    lp = np.array([0.2, 3.0, -0.1])
    b.add_light(point_light(tuple(lp), intensity=4.0))
    scene = b.finalize()
    res, gmin, gsize, rays = (3, 2, 3), (-2.0, 0.5, -2.0), (4.0, 0.5, 4.0), 256
    o = oracle.OracleField(scene, gmin, gsize, res, rays, 6, prec="64")
    o.trace(o.params(frame_jitter=0.15), vis=True)
    vis = o.vis.reshape(-1, rays)
    hits, rv = o.hits_view(), o.rays_view()
    res_a = np.asarray(res)
    bias = max(min(np.asarray(gsize) / np.maximum(res_a - 1, 1)) * 0.005, 1e-3)
    n_shadow = n_lit = n_under = 0
    for slot in range(int(np.prod(res))):
        ix, iy, iz = slot % res[0], (slot // res[0]) % res[1], slot // (res[0] * res[1])
        ro = np.asarray(gmin) + np.array([ix, iy, iz]) / np.maximum(res_a - 1, 1) * np.asarray(gsize)
        for k in range(rays):
            rd = oracle.ray_dir(k, rays, 0.15)
            if hits[slot, k, 0] < 0:
                assert vis[slot, k] == 0
                continue
            if rd[1] > 0:  # the square, from below
                assert vis[slot, k] == 0
                n_under += 1
                continue
            h = ro + rd * rv[slot, k, 3] + np.array([0.0, bias, 0.0])
            s = (1.5 - h[1]) / (lp[1] - h[1])
            q = h + s * (lp - h)
            margin = min(abs(q[0]) - a, abs(q[2]) - a) if max(abs(q[0]), abs(q[2])) < a else \
                max(abs(q[0]) - a, abs(q[2]) - a)
            if abs(margin) < 1e-6:
                continue
            blocked = max(abs(q[0]), abs(q[2])) < a
            assert vis[slot, k] == (0 if blocked else 1), (slot, k, q)
            n_shadow += int(blocked)
            n_lit += int(not blocked)
    assert n_shadow > 50 and n_lit > 200 and n_under > 50


# ------------------------------------------------------------- watertight triangle test
def _mesh_scene(v, t, side=2):
    b = SceneBuilder()
    m = b.add_material(uber_material((0.5, 0.5, 0.5), side=side))
    b.add_instance(b.add_mesh(v, t, tri_material=np.full(len(t), m, np.uint32)))
    return b.finalize()


def _closest(scene, ro, rd, prec, wt):
    oracle.set_watertight(wt, prec)
    try:
        return oracle.brute_closest(scene, ro, rd, prec)
    finally:
        oracle.set_watertight(False, prec)


@pytest.mark.parametrize("prec", ["64", "32"])
def test_watertight_agrees_with_moller_trumbore_away_from_edges(prec):
    """Rays through a triangle's interior (barycentrics >= 0.05 from every edge) and rays
    missing it by the same margin: the watertight test and Moller-Trumbore agree on hit or
    miss, on t (1e-12 in FP64, 5e-5 in FP32; rays within 3 degrees of the plane, where t is
    ill-conditioned in either form, are left out) and on the sign of the determinant that the
    side rules read (front faces of a counter-clockwise triangle positive)."""
    rng = np.random.default_rng(51)
    n_hit = 0
    for trial in range(300):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        if np.linalg.norm(n) < 0.2:
            continue
        scene = _mesh_scene(tri, np.array([[0, 1, 2]], np.uint32), side=2 if trial % 2 else 0)
        bary = rng.dirichlet((1.0, 1.0, 1.0))
        inside = bary.min() >= 0.05
        if not inside and bary.min() > -0.05:
            continue
        target = bary @ tri.astype(np.float64)
        if trial % 3 == 0:  # a point beyond an edge: b0 < -0.05
            target = (-0.1) * tri[0] + 0.55 * tri[1] + 0.55 * tri[2]
            inside = False
        ro = target + rng.normal(size=3) * 2.0
        rd = (target - ro) / np.linalg.norm(target - ro)
        if abs(rd @ n) < 0.05 * np.linalg.norm(n):
            continue  # grazing: t is ill-conditioned in either form
        a = _closest(scene, ro, rd, prec, False)
        b = _closest(scene, ro, rd, prec, True)
        assert (a[1] >= 0) == (b[1] >= 0), trial
        if a[1] >= 0:
            np.testing.assert_allclose(b[0], a[0], rtol=1e-12 if prec == "64" else 5e-5)
            n_hit += 1
        # single-sided triangles (even trials) are seen only from the front, by both tests
        if trial % 2 == 0 and inside:
            assert (a[1] >= 0) == (rd @ n < 0)
    assert n_hit > 80


def _edge_and_vertex_rays(v, t, centre):
    """Rays from an interior point aimed at every vertex and at the f32 midpoint of every
    edge of a closed mesh."""
    pts = [v[i] for i in range(len(v))]
    edges = set()
    for a, b, c in t:
        for x, y in ((a, b), (b, c), (c, a)):
            edges.add((min(x, y), max(x, y)))
    for x, y in sorted(edges):
        pts.append((v[x] + (v[y] - v[x]) * np.float32(0.5)).astype(np.float32))
    dirs = np.array(pts, np.float64) - centre
    return dirs / np.linalg.norm(dirs, axis=1, keepdims=True)


@pytest.mark.parametrize("shape", ["sphere", "box"])
def test_watertight_closed_mesh_never_leaks(shape):
    """From inside a closed mesh, a ray aimed exactly at a shared vertex or at a shared
    edge must hit the surface: the watertight test (FP32, as the GPU evaluates it) leaks no
    ray. Moller-Trumbore's count of leaking rays on the same set is reported beside it."""
    if shape == "sphere":
        v, t = sg.uv_sphere(24, 48, radius=1.3)
        centre = np.array([0.013, -0.021, 0.007])
    else:
        v, t = sg.box_mesh((0.0, 0.0, 0.0), (1.1, 0.7, 0.9), 0.37)
        centre = np.array([0.05, -0.02, 0.03])
    v, t = np.asarray(v, np.float32), np.asarray(t, np.uint32)
    scene = _mesh_scene(v, t)
    dirs = _edge_and_vertex_rays(v, t, centre)
    leaks_wt = leaks_mt = 0
    for rd in dirs:
        leaks_wt += int(_closest(scene, centre, rd, "32", True)[1] < 0)
        leaks_mt += int(_closest(scene, centre, rd, "32", False)[1] < 0)
    print(f"\n{shape}: {len(dirs)} vertex and edge rays, leaks: watertight {leaks_wt}, Moller-Trumbore {leaks_mt}")
    assert len(dirs) > 20
    assert leaks_wt == 0
