"""Batched discs (dfargo_batch_*) and the launch-shape autotuner.

A batch advances several discs of uneven grids and parameters through one
graph launch per sub-step, each disc with its own device-side CFL step and
loop condition (src/SourceEuler.c:398-453 per disc).  Every disc must come
out bit for bit as its solo run: the same kernels, the same order within a
disc, nothing shared between discs.  The autotuner's knobs change launch
shapes only, so a tuned disc must equal an untuned one bit for bit too."""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Batch, Disc, DfargoError
from dustyfargoadsg_amd.workloads import batch_discs, gas_planet

pytestmark = pytest.mark.gpu

DT = 2 * math.pi / 20


def _make(cases, prec):
    discs = []
    for p, pl in cases:
        d = Disc(p, prec)
        d.initial_state(pl)
        discs.append(d)
    return discs


@pytest.mark.parametrize("prec,mode", [(_capi.FP64, Batch.STREAMS), (_capi.FP32, Batch.STREAMS),
                                       (_capi.FP64, Batch.GRAPH)])
def test_batch_equals_solo_runs(prec, mode):
    cases = batch_discs(n=5, seed=7, nrads=(16, 32, 64))
    print(f"case batch of {[p.nrad for p, _ in cases]} prec={prec} mode={mode}", flush=True)
    batched, solo = _make(cases, prec), _make(cases, prec)
    b = Batch(batched)
    b.set_mode(mode)
    dts = [DT * (1.0 + 0.1 * k) for k in range(len(cases))]
    for _ in range(2):
        b.algogas(dts)
        for d, dt in zip(solo, dts):
            d.algogas(dt)
    for k, (x, y) in enumerate(zip(batched, solo)):
        assert x.status().substeps_total == y.status().substeps_total, k
        assert x.status().physical_time == y.status().physical_time, k
        sx, sy = x.state(), y.state()
        for f in ("rho", "vr", "vt"):
            assert np.array_equal(sx[f], sy[f]), (k, f)
    # the discs took different numbers of sub-steps
    assert len({d.status().substeps_total for d in batched}) > 1


def test_batch_rejects_a_changed_disc():
    cases = batch_discs(n=2, seed=3, nrads=(16,))
    discs = _make(cases, _capi.FP64)
    b = Batch(discs)
    b.algogas(DT)
    discs[1].set_tuning(azi_waves=2)  # a new launch shape: a new sub-step graph
    with pytest.raises(DfargoError):
        b.algogas(DT)


def test_autotune_changes_launch_shapes_not_results():
    p, pl = gas_planet(nrad=64, nsec=192)
    print("case autotune 64x192", flush=True)
    a, b = Disc(p), Disc(p)
    a.initial_state(pl)
    b.initial_state(pl)
    t = a.autotune(substeps=3)
    assert t["xcd_remap"] in (0, 1) and t["azi_waves"] in (1, 2, 4) and t["finish_block"] in (64, 128, 256)
    assert a.tuning() == t
    # every shape, not only the one picked, gives the untuned results
    for shape in (dict(xcd_remap=0, azi_waves=1, finish_block=64), t):
        c = Disc(p)
        c.initial_state(pl)
        c.set_tuning(**shape)
        c.algogas(DT)
        if shape is t:
            a.algogas(DT)
        b2 = Disc(p)
        b2.initial_state(pl)
        b2.algogas(DT)
        for f in ("rho", "vr", "vt"):
            assert np.array_equal(c.state()[f], b2.state()[f]), (shape, f)
    assert np.array_equal(a.state()["rho"], b2.state()["rho"])
    # the same shape again comes from the process's cache
    d = Disc(p)
    d.initial_state(pl)
    assert d.autotune(substeps=3) == t
