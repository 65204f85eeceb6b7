"""The per-DT diagnostics on the GPU (dfargo_gas_budgets, dfargo_update_log)
against the FP64 oracle after AlgoGas calls of the same DT.

Bars: FP64 storage, the budgets within 1e-12 relative and the torques and
powers within 1e-10 of the largest of a planet's eight columns (the device
sums cells in a fixed tree per ring, the oracle sequentially); FP32 storage,
the budgets within 1e-5 relative.
"""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

DT = 2 * math.pi / 20


def pair(p, pl, precision=_capi.FP64, calls=2):
    print(f"case {p.nrad}x{p.nsec} prec={precision}", flush=True)
    o, d = Oracle(p), Disc(p, precision)
    o.initial_state(pl)
    d.initial_state(pl)
    for _ in range(calls):
        o.algogas(DT)
        d.algogas(DT)
    return o, d


def check_budgets(o, d, tol):
    bo, bg = o.budgets(), d.gas_budgets()
    for q in range(3):
        assert abs(bg[q] - bo[q]) <= tol * abs(bo[q]), (q, bg[q], bo[q])
    assert abs(bg[3] - bo[3]) <= 1e-14 * max(abs(bo[3]), 1e-300)


def check_log(o, d, tol):
    to, io = o.update_log()
    tg, ig = d.update_log()
    assert tg.shape == to.shape
    for k in range(to.shape[0]):
        scale = np.abs(to[k]).max()
        assert np.abs(tg[k] - to[k]).max() <= tol * scale, (k, tg[k], to[k])
        assert abs(ig[k] - io[k]) <= tol * max(abs(io[k]), 1e-30)


@pytest.mark.parametrize("nr,ns", [(32, 96), (16, 1536)])
def test_planet_disc_fp64(nr, ns):
    p, pl = gas_planet(nrad=nr, nsec=ns)
    o, d = pair(p, pl)
    check_budgets(o, d, 1e-12)
    check_log(o, d, 1e-10)


def test_feel_disk_hill_exclusion_fp64():
    """A planet that feels the disc, with the Hill-sphere exclusion."""
    p, pl = gas_planet(nrad=32, nsec=96, exclude_hill=1, forced_circular=0)
    pl["feel_disk"] = [1]
    o, d = pair(p, pl)
    check_budgets(o, d, 1e-12)
    check_log(o, d, 1e-10)
    to, _ = o.update_log()
    assert np.all(to[0, 2:4] != to[0, 0:2])


def test_energy_disc_budgets_fp64():
    """Adiabatic self-gravitating disc: the internal energy enters the total."""
    p, pl = energy_sg(nrad=32, nsec=96)
    o, d = pair(p, pl)
    check_budgets(o, d, 1e-12)
    assert d.update_log()[0].shape == (0, 8)


def test_budgets_fp32_storage():
    p, pl = gas_planet(nrad=32, nsec=96)
    o, d = pair(p, pl, precision=_capi.FP32)
    check_budgets(o, d, 1e-5)
