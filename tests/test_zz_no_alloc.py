"""sbgi_query_rays_noalloc against sbgi_query_rays (run on an MI355X with -m gpu).

The no-allocation variant runs the same walk kernels with its staging and fetch counters in
the library's static device scratch, so its results must equal the allocating entry point's
bit for bit: on two calls in a row, the second with fewer rays (nothing of the first call
may leak into it), in all three query modes. A call above the scratch's ray limit is
refused before any device work.
"""
from __future__ import annotations

import numpy as np
import pytest

from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField

pytestmark = pytest.mark.gpu


def _rays(rng, w, n):
    lo = np.asarray(w.grid_min, np.float64)
    hi = lo + np.asarray(w.grid_size, np.float64)
    ro = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    rd = rng.normal(size=(n, 3))
    rd = (rd / np.linalg.norm(rd, axis=1, keepdims=True)).astype(np.float32)
    mode = rng.integers(0, 3, n).astype(np.float32)        # closest, any hit, closest-as-blocked
    tmax = rng.uniform(0.5, 30.0, n).astype(np.float32)
    return ro, rd, tmax, mode


def test_query_rays_noalloc_equals_query_rays():
    w = sg.instanced(res=(8, 4, 6), rays=64, total_tris=24576, n_instances=32, n_lights=4)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, traversal=N.TRAV_MT_ORDERED)
    f.set_scene(w.scene)
    rng = np.random.default_rng(77)
    try:
        for n in (12000, 3000):
            ro, rd, tmax, mode = _rays(rng, w, n)
            print(f"case: {n} query rays, allocating then no-allocation entry point", flush=True)
            want = f.query(ro, rd, tmax, mode)
            got = f.query(ro, rd, tmax, mode, noalloc=True)
            for a, b in zip(want, got):
                assert np.array_equal(a, b)
            hits = int((want[1] >= 0).sum())
            blocked = int(want[3].sum())
            print(f"{n} rays: {hits} hits, {blocked} blocked, identical")
            assert 0 < hits < n and 0 < blocked < n
        ro, rd, tmax, mode = _rays(rng, w, 16385)
        with pytest.raises(RuntimeError, match="16384"):
            f.query(ro, rd, tmax, mode, noalloc=True)
    finally:
        f.close()
