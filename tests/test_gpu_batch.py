"""Batched multi-shift solves: every system of a batch (pseudofermion fields
and lattices of different sizes, 25 and 12 shifts) equals the one-at-a-time
chs_qmrherm_dev solve bit for bit, whatever the launch parameters."""
import numpy as np
import pytest

from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import batch, remez  # noqa: E402

R4, R2G = remez.load("remez4"), remez.load("remez2g")


def make_systems(seed=3):
    rs = np.random.RandomState(seed)
    specs = [((8, 8), R4, 2), ((4, 4, 6), R4, 1), ((6, 12), R2G, 1), ((10, 6), R2G, 1), ((5, 4), R2G, 1)]
    lats, systems = [], []
    for dims, coef, npf in specs:
        lat = fm.Lattice(*dims[:2], Lt=dims[2] if len(dims) > 2 else None)
        lats.append(lat)
        chi = fm.DeviceArray.from_numpy(lat, phi_field(rs, lat.kvol))
        anum, aden = coef[0], coef[1]
        for _ in range(npf):
            phi = fm.spinor_dev(lat, spinor(rs, lat.shape))
            x = fm.spinor_dev(lat, None)
            systems.append(batch.System(lat, phi, chi, x, anum, aden, res=1e-9))
    return lats, systems


def single(s, precision="fp64"):
    x = fm.spinor_dev(s.lat, None)
    it = fm.qmrherm_prec_dev(s.lat, s.phi, s.res, s.anum, s.aden, s.chi, x, precision)
    s.lat.synchronize()
    return fm.to_host(s.lat, x), it


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("tune", [None, {"ept": 1, "chunk": 2, "graph": 0}, {"ept": 4, "chunk": 6, "graph": 1},
                                  {"ept": 2, "chunk": 16, "graph": 1}])
def test_batch_equals_single(tune, precision):
    lats, systems = make_systems()
    b = batch.Batch()
    st = b.qmrherm(systems, tune, precision)
    print(st)
    for s in systems:
        want, it = single(s, precision)
        got = fm.to_host(s.lat, s.x)
        assert s.itercg == it and it > 10
        assert np.array_equal(got, want), np.abs(got - want).max()
    assert st["iterations"] >= max(s.itercg for s in systems)


def test_batch_reuse_and_tuned_cache():
    """A second call on the same batch shape reuses the tuned parameters and
    the captured graphs and still gives the one-at-a-time results."""
    lats, systems = make_systems(seed=5)
    b = batch.Batch()
    st1 = b.qmrherm(systems)
    st2 = b.qmrherm(systems)
    assert st2["tune"] == st1["tune"]
    for s in systems:
        want, it = single(s)
        assert s.itercg == it
        assert np.array_equal(fm.to_host(s.lat, s.x), want)


def test_batch_fp32_accuracy():
    """FP32 storage against the FP64 solve of the same system: the solution
    error is stated and bounded (the scalars and x stay FP64)."""
    lats, systems = make_systems(seed=9)
    b = batch.Batch()
    b.qmrherm(systems, None, "fp32")
    worst = 0.0
    for s in systems:
        want, _ = single(s, "fp64")
        got = fm.to_host(s.lat, s.x)
        worst = max(worst, np.linalg.norm(got - want) / np.linalg.norm(want))
    print("FP32-storage batch: worst relative solution error vs FP64 %.2e" % worst)
    assert worst < 1e-5
