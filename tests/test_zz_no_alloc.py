"""dfargo_update_log_noalloc (the torque log with its scratch in the
library's static device array) against dfargo_update_log (scratch allocated
on first use), through the C ABI: bitwise equal results on two calls in a
row, the second on a smaller disc, so a stale slot left by the larger disc
would show.  Also the size limit of the scratch."""
import ctypes

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc, DfargoError
from dustyfargoadsg_amd.workloads import gas_planet

pytestmark = pytest.mark.gpu


def _log(d, fn):
    tq, ind, nb = np.zeros(64), np.zeros(8), ctypes.c_int()
    rc = fn(d.ctx, 8, tq.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ind.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nb))
    if rc != _capi.OK:
        raise DfargoError(d.lib.dfargo_last_error().decode(), rc)
    return tq[: 8 * nb.value], ind[: nb.value]


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.FP32])
def test_noalloc_torque_log_matches(prec):
    lib = _capi.load()
    discs = []
    for nr, ns in ((128, 384), (32, 96)):
        p, pl = gas_planet(nrad=nr, nsec=ns, forced_circular=0)
        pl = dict(dist=[1.0, 1.6], mass=[1e-3, 3e-4], acc=[0.0, 0.0], feel_disk=[1, 1], feel_others=[1, 1])
        print(f"case no-alloc torque log {nr}x{ns} prec={prec}", flush=True)
        d = Disc(p, prec)
        d.initial_state(pl)
        d.algogas(0.3)
        discs.append(d)
    for d in discs:  # larger disc first, then the smaller one
        want = _log(d, lib.dfargo_update_log)
        got = _log(d, lib.dfargo_update_log_noalloc)
        assert len(got[0]) == 16 and np.any(got[0] != 0.0)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # once more on the larger disc after the smaller one
    want = _log(discs[0], lib.dfargo_update_log)
    got = _log(discs[0], lib.dfargo_update_log_noalloc)
    assert np.array_equal(got[0], want[0])


def test_noalloc_rejects_too_many_planets():
    lib = _capi.load()
    p, pl = gas_planet(nrad=32, nsec=96)
    d = Disc(p)
    d.initial_state(pl)
    tq, ind, nb = np.zeros(8), np.zeros(1), ctypes.c_int()
    rc = lib.dfargo_update_log_noalloc(d.ctx, 0, tq.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       ind.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nb))
    assert rc == _capi.ERR_ARG
