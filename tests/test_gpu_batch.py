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


def make_force_systems(seed=21):
    """A batch mixing the force's solves (iflag = 1, remez2g a-set as in
    rhmc/cheisdwf.f:505-513) with heatbath-type ones (iflag = 0), on lattices
    of different sizes; the force arrays start from random values, to which
    the solves add."""
    rs = np.random.RandomState(seed)
    # (dims, remez set, iflag, fields, imass, hpin): both wall couplings and the pinning field
    specs = [((8, 8), R2G, 1, 2, 1, 0.0), ((6, 12), R2G, 1, 1, 3, 0.0), ((4, 4, 6), R4, 0, 1, 1, 0.5),
             ((10, 6), R2G, 1, 1, 1, 0.5), ((5, 4), R2G, 0, 1, 3, 0.0), ((6, 3), R2G, 1, 1, 1, 0.0)]
    lats, systems, F0 = [], [], []
    for dims, coef, iflag, npf, imass, hpin in specs:
        lat = fm.Lattice(*dims[:2], Lt=dims[2] if len(dims) > 2 else None)
        lats.append(lat)
        chi = fm.DeviceArray.from_numpy(lat, phi_field(rs, lat.kvol))
        for _ in range(npf):
            phi = fm.spinor_dev(lat, spinor(rs, lat.shape))
            x = fm.spinor_dev(lat, None)
            f0 = rs.randn(3, lat.kvol) if iflag else None
            F = fm.DeviceArray.from_numpy(lat, f0) if iflag else None
            systems.append(batch.System(lat, phi, chi, x, coef[0], coef[1], res=1e-9, hpin=hpin, imass=imass,
                                        iflag=iflag, F=F))
            F0.append(f0)
    return lats, systems, F0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("tune", [None, {"ept": 2, "chunk": 4, "graph": 1}])
def test_batch_force_solves_equal_single(tune, precision):
    """iflag = 1 systems in a batch give the one-at-a-time solve's x (a0 Phi),
    force contributions and itercg bit for bit, beside iflag = 0 systems."""
    lats, systems, F0 = make_force_systems()
    b = batch.Batch()
    b.qmrherm(systems, tune, precision)
    for s, f0 in zip(systems, F0):
        x = fm.spinor_dev(s.lat, None)
        F = fm.DeviceArray.from_numpy(s.lat, f0) if s.iflag else None
        it = fm.qmrherm_prec_dev(s.lat, s.phi, s.res, s.anum, s.aden, s.chi, x, precision, hpin=s.hpin,
                                 imass=s.imass, iflag=s.iflag, F_dev=F)
        s.lat.synchronize()
        assert s.itercg == it and it > 10, (s.itercg, it)
        assert np.array_equal(fm.to_host(s.lat, s.x), fm.to_host(s.lat, x))
        if s.iflag:
            got, want = s.F.download(np.float64, (3, s.lat.kvol)), F.download(np.float64, (3, s.lat.kvol))
            assert np.array_equal(got, want), np.abs(got - want).max()
            assert np.abs(got - f0).max() > 0


def test_batch_force_matches_oracle():
    """The batched force solves against the oracle's qmrherm with iflag = 1
    (the CPU restatement pinned to the reference's golden vectors): force
    contributions to 1e-9 relative at res 1e-10, iteration counts within 5."""
    import oracle as orc
    rs = np.random.RandomState(8)
    cases = [(6, 8, None), (4, 4, 6), (8, 6, None)]
    lats, systems, hosts = [], [], []
    for L, Ls, Lt in cases:
        lat = fm.Lattice(L, Ls, Lt)
        lats.append(lat)
        chi = phi_field(rs, lat.kvol)
        for _ in range(2):
            ph = spinor(rs, lat.shape)
            F = fm.DeviceArray.from_numpy(lat, np.zeros((3, lat.kvol)))
            systems.append(batch.System(lat, fm.spinor_dev(lat, ph), fm.DeviceArray.from_numpy(lat, chi),
                                        fm.spinor_dev(lat, None), R2G[0], R2G[1], res=1e-10, iflag=1, F=F))
            hosts.append((orc.Lattice(L, Ls, Lt), ph, chi))
    b = batch.Batch()
    b.qmrherm(systems)
    for s, (olat, ph, chi) in zip(systems, hosts):
        Fo = np.zeros((3, olat.kvol))
        _, ito = orc.qmrherm(olat, ph, 1e-10, R2G[0], R2G[1], chi, iflag=1, dsdchi=Fo)
        got = s.F.download(np.float64, (3, olat.kvol))
        err = np.abs(got - Fo).max() / np.abs(Fo).max()
        assert err < 1e-9 and abs(s.itercg - ito) <= 5, (olat.L, olat.Ls, err, s.itercg, ito)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_forces_equal_chs_force(precision):
    """Batch.force -- the two multi-shift stages of several fields' forces,
    each stage one batch -- gives chs_force's dSdchi for every field bit for
    bit (FP32: with the context's force precision set to FP32 storage)."""
    rs = np.random.RandomState(12)
    fields, hosts = [], []
    for L, Ls, npf in ((8, 8, 2), (6, 12, 1), (10, 6, 1)):
        lat = fm.Lattice(L, Ls)
        fm.set_force_precision(lat, precision)
        chi = phi_field(rs, lat.kvol)
        chid = fm.DeviceArray.from_numpy(lat, chi)
        for _ in range(npf):
            ph = spinor(rs, lat.shape)
            fields.append((lat, fm.spinor_dev(lat, ph), chid, chi))
            hosts.append(ph)
    b = batch.Batch()
    Fs, _, (st1, st2) = b.force(fields, 0.475, 1e-6, R2G, precision)
    for (lat, _, _, chi), ph, F in zip(fields, hosts, Fs):
        want, _, _ = fm.force(lat, ph, 1e-6, chi, 0.475, R2G)
        got = F.download(np.float64, (3, lat.kvol))
        assert np.array_equal(got, want), np.abs(got - want).max()
