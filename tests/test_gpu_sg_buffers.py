"""The self-gravity transforms in one LDS buffer per transform (in place,
half the LDS) against the two-buffer Stockham form (ping-pong): the
butterflies and the spectral product do the same arithmetic in the same
order, so the accelerations and a sub-step's state are equal bit for bit;
so are the row transforms taken two rings per workgroup.
The grids take radix 8, 4, 3 and 2 passes (rows of 48 to 192 cells,
radial transforms of 32 to 256; self-gravity needs NRAD a power of two)."""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc, DfargoError
from dustyfargoadsg_amd.workloads import energy_sg

pytestmark = pytest.mark.gpu

DT = 2 * math.pi / 20


def make(p, pl, prec, bufs, pair=1):
    d = Disc(p, prec)
    d.initial_state(pl)
    if bufs is not None:
        d.set_tuning(sg_fft_buffers=bufs, sg_rows_pair=pair)
    return d


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.FP32])
@pytest.mark.parametrize("nr,ns", [(16, 48), (32, 96), (64, 144), (128, 192)])
def test_one_buffer_equals_two(prec, nr, ns):
    print(f"case sg buffers {nr}x{ns} prec={prec}", flush=True)
    p, pl = energy_sg(nrad=nr, nsec=ns)
    two = make(p, pl, prec, 2)
    assert two.tuning()["sg_fft_buffers"] == 2 and two.tuning()["sg_rows_pair"] == 1
    g2 = two.compute_selfgravity()
    two.algogas(DT)
    s2 = two.state()
    # one buffer per transform; two rings per row workgroup (in place)
    for bufs, pair in ((1, 1), (2, 2)):
        one = make(p, pl, prec, bufs, pair)
        for a, b in zip(one.compute_selfgravity(), g2):
            assert np.array_equal(a, b), (bufs, pair)
        one.algogas(DT)
        s1 = one.state()
        for k in ("rho", "vr", "vt", "energy"):
            assert np.array_equal(s1[k], s2[k]), (bufs, pair, k)
        for f in ("sg_accr", "sg_acct"):
            assert np.array_equal(one.field(f), two.field(f)), (bufs, pair, f)


def test_default_at_full_size():
    """512x1536 with FP64 storage: by default the radial transforms take
    one buffer (four workgroups share a CU instead of two); the result is
    that of two buffers."""
    print("case sg buffers default 512x1536 FP64", flush=True)
    p, pl = energy_sg(nrad=512, nsec=1536)
    auto, two = make(p, pl, _capi.FP64, None), make(p, pl, _capi.FP64, 2)
    assert auto.tuning()["sg_fft_buffers"] == 0
    for a, b in zip(auto.compute_selfgravity(), two.compute_selfgravity()):
        assert np.array_equal(a, b)


def test_tuning_range():
    p, pl = energy_sg(nrad=32, nsec=96)
    d = make(p, pl, _capi.FP64, None)
    with pytest.raises(DfargoError):
        d.set_tuning(sg_fft_buffers=3)
    with pytest.raises(DfargoError):
        d.set_tuning(sg_fft_buffers=-1)
    with pytest.raises(DfargoError):
        d.set_tuning(sg_rows_pair=3)
