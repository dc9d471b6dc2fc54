"""Shift freezing of the multi-shift solve (chs_set_shift_freeze): a shift's
vectors stop being updated once its own term of the reference's stopping test
(rhmc/cheisdwf.f:697-715) is below f times the bound.  Checked: the stopping
iteration is the reference's, solutions and force contributions stay within
the oracle comparisons of the unfrozen solve, freezing does change the work
(and f = 0 restores the reference's results bit for bit), and batched solves
equal one-at-a-time solves under the same setting."""
import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import batch, remez  # noqa: E402
from chiralheisenberg_amd._lib import ChsError  # noqa: E402

R4, R2G = remez.load("remez4"), remez.load("remez2g")


@pytest.mark.parametrize("imass", [1, 3])
def test_freeze_iflag0_matches_oracle(imass):
    lat, olat = fm.Lattice(8, 8), orc.Lattice(8, 8)
    rs = np.random.RandomState(70 + imass)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R4[0], R4[1]
    x0, i0 = fm.qmrherm(lat, phi, 1e-9, a, ad, chi, 0.0, 0.3, imass, 0)
    xc, ic = orc.qmrherm(olat, phi, 1e-9, a, ad, chi, 0.3, imass, 0)
    for f in (0.01, 1.0):
        lat.set_shift_freeze(f)
        xf, itf = fm.qmrherm(lat, phi, 1e-9, a, ad, chi, 0.0, 0.3, imass, 0)
        err = np.abs(xf - xc).max() / np.abs(xc).max()
        diff = np.abs(xf - x0).max() / np.abs(x0).max()
        print("imass %d, f = %.2f: itercg %d (unfrozen %d, oracle %d), error vs oracle %.2e, change %.2e" % (
            imass, f, itf, i0, ic, err, diff))
        assert itf == i0
        # f = 0.01: the unfrozen solve's bound; f = 1 (each shift stopped at
        # the reference's own bound): within ten times the solver tolerance
        assert err < (1e-9 if f < 0.1 else 1e-8), err
        if f == 1.0:
            assert diff > 0    # some shift was frozen
    lat.set_shift_freeze(0.0)
    x1, i1 = fm.qmrherm(lat, phi, 1e-9, a, ad, chi, 0.0, 0.3, imass, 0)
    assert i1 == i0 and np.array_equal(x1, x0)
    lat.close()


def test_freeze_force_matches_oracle():
    """The force's solve (iflag = 1, remez2g a-set) at res 1e-12: force
    contributions to 1e-10 relative of the oracle's, as without freezing."""
    lat, olat = fm.Lattice(8, 8), orc.Lattice(8, 8)
    rs = np.random.RandomState(79)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R2G[0], R2G[1]
    F0 = rs.randn(3, lat.kvol)
    Fc = F0.copy()
    orc.qmrherm(olat, phi, 1e-12, a, ad, chi, 0.0, 1, 1, Fc)
    Fu = F0.copy()
    _, iu = fm.qmrherm(lat, phi, 1e-12, a, ad, chi, 0.0, 0.0, 1, 1, Fu)
    for f in (0.01, 1.0):
        lat.set_shift_freeze(f)
        Fg = F0.copy()
        _, ig = fm.qmrherm(lat, phi, 1e-12, a, ad, chi, 0.0, 0.0, 1, 1, Fg)
        err = np.abs(Fg - Fc).max() / np.abs(Fc - F0).max()
        diff = np.abs(Fg - Fu).max() / np.abs(Fu - F0).max()
        print("f = %.2f: itercg %d (unfrozen %d), force error vs oracle %.2e, change %.2e" % (f, ig, iu, err, diff))
        assert ig == iu
        assert err < 1e-10, err
    lat.close()


def test_freeze_chs_force_and_hamilton():
    """chs_force (both multi-shift stages) and hamilton at production
    tolerances (1e-6 force, 1e-9 Hamiltonian): frozen against unfrozen,
    force within 1e-6 relative, H within 1e-9 relative, same iterations."""
    lat = fm.Lattice(8, 8)
    rs = np.random.RandomState(80)
    Phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    R2 = remez.load("remez2")
    out = {}
    for f in (0.0, 1.0):
        lat.set_shift_freeze(f)
        F, _, itf = fm.force(lat, Phi, 1e-6, chi, 0.475, R2G, hpin=0.0, imass=1)
        h, _, ith = fm.hamilton(lat, Phi, 1e-9, chi, np.zeros((3, lat.kvol)), 0.475, R2, hpin=0.0, imass=1)
        out[f] = (F, itf, h, ith)
    (F0, i0, h0, j0), (F1, i1, h1, j1) = out[0.0], out[1.0]
    ef = np.abs(F1 - F0).max() / np.abs(F0).max()
    eh = abs(h1[0] - h0[0]) / abs(h0[0])
    print("force change %.2e, H change %.2e, iterations %s / %s" % (ef, eh, (i0, j0), (i1, j1)))
    assert (i0, j0) == (i1, j1)
    assert ef < 1e-6 and eh < 1e-9
    lat.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_with_freeze_equals_single(precision):
    """Each system's context sets its freezing; the batch reproduces the
    one-at-a-time solves (x, force, itercg) bit for bit, frozen and unfrozen
    systems mixed."""
    rs = np.random.RandomState(81)
    specs = [((8, 8), R2G, 1, 1.0), ((6, 12), R4, 0, 0.01), ((4, 4, 6), R4, 0, 0.0), ((6, 4), R2G, 1, 0.0)]
    lats, systems, F0 = [], [], []
    for dims, coef, iflag, f in specs:
        lat = fm.Lattice(*dims[:2], Lt=dims[2] if len(dims) > 2 else None)
        lat.set_shift_freeze(f)
        lats.append(lat)
        chi = fm.DeviceArray.from_numpy(lat, phi_field(rs, lat.kvol))
        phi = fm.spinor_dev(lat, spinor(rs, lat.shape))
        x = fm.spinor_dev(lat, None)
        f0 = rs.randn(3, lat.kvol) if iflag else None
        F = fm.DeviceArray.from_numpy(lat, f0) if iflag else None
        systems.append(batch.System(lat, phi, chi, x, coef[0], coef[1], res=1e-9, iflag=iflag, F=F))
        F0.append(f0)
    b = batch.Batch()
    b.qmrherm(systems, None, precision)
    for s, f0 in zip(systems, F0):
        x = fm.spinor_dev(s.lat, None)
        F = fm.DeviceArray.from_numpy(s.lat, f0) if s.iflag else None
        it = fm.qmrherm_prec_dev(s.lat, s.phi, s.res, s.anum, s.aden, s.chi, x, precision, iflag=s.iflag, F_dev=F)
        s.lat.synchronize()
        assert s.itercg == it and it > 10
        assert np.array_equal(fm.to_host(s.lat, s.x), fm.to_host(s.lat, x))
        if s.iflag:
            assert np.array_equal(s.F.download(np.float64, (3, s.lat.kvol)), F.download(np.float64, (3, s.lat.kvol)))
    for lat in lats:
        lat.close()


def test_freeze_factor_is_checked():
    lat = fm.Lattice(4, 4)
    for f in (-0.1, 1.5, float("nan")):
        with pytest.raises(ChsError):
            lat.set_shift_freeze(f)
    lat.close()


def test_driver_run_with_freezing(tmp_path):
    """rhmc.run at production tolerances (rescga 1e-9, rescgg 1e-6) with
    shift freezing f = 0.01 against the reference's solves: the same
    trajectory lengths and decisions, dH within the bound the GPU and oracle
    runs are compared to at these tolerances (1e-6 max(1, |dH|); measured
    1.7e-7)."""
    from chiralheisenberg_amd import checkpoint, rhmc
    L, Ls, iter2 = 6, 8, 4
    rs = np.random.RandomState(82)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.475)
    rows = {}
    for f in (0.0, 0.01):
        d = tmp_path / ("f%g" % f)
        d.mkdir()
        open(d / "midout", "w").write("0.1 0.475 1.0 0.00 0.0 1 10 %d\ndt beta am3 am h0 imass iterl iter2\n" % iter2)
        checkpoint.write_con(d / "con", chi0, 122110675517440.0)
        rhmc.run(str(d), L, Ls, iprint=100, shift_freeze=f)
        rows[f] = [l.split() for l in open(d / "control")]
    a, b = rows[0.0], rows[0.01]
    print("dH: %s / %s" % ([r[0] for r in a], [r[0] for r in b]))
    assert [r[2] for r in a] == [r[2] for r in b]
    for x, y in zip(a, b):
        assert abs(float(x[0]) - float(y[0])) <= 1e-6 * max(1.0, abs(float(x[0])))
    fa = [l.split()[0] for l in open(tmp_path / "f0" / "fort.11")]
    fb = [l.split()[0] for l in open(tmp_path / "f0.01" / "fort.11")]
    assert fa == fb


def test_driver_run_fast_forms(tmp_path):
    """rhmc.run with every opt-in fast form at once -- FP32 force solves,
    shift freezing f = 0.01, binary16-form propagator solves -- against the
    reference's solves: the same trajectory lengths and decisions, dH within
    the FP32 force's error (1e-3 max(1, |dH|)), correlators within 1e-6."""
    from chiralheisenberg_amd import checkpoint, rhmc
    L, Ls, iter2 = 6, 8, 4
    rs = np.random.RandomState(83)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.475)
    runs = {}
    for tag, kw in (("ref", {}), ("fast", dict(force_precision="fp32", shift_freeze=0.01, fermion_precision="fp16"))):
        d = tmp_path / tag
        d.mkdir()
        open(d / "midout", "w").write("0.1 0.475 1.0 0.00 0.0 1 10 %d\ndt beta am3 am h0 imass iterl iter2\n" % iter2)
        checkpoint.write_con(d / "con", chi0, 122110675517440.0)
        rhmc.run(str(d), L, Ls, iprint=2, **kw)
        runs[tag] = d
    a = [l.split() for l in open(runs["ref"] / "control")]
    b = [l.split() for l in open(runs["fast"] / "control")]
    print("dH reference %s, fast forms %s" % ([x[0] for x in a], [x[0] for x in b]))
    assert [x[2] for x in a] == [x[2] for x in b]
    for x, y in zip(a, b):
        assert abs(float(x[0]) - float(y[0])) <= 1e-3 * max(1.0, abs(float(x[0])))
    for f in ("fort.5001", "fort.5011"):
        p, q = np.loadtxt(runs["ref"] / f)[:, 1:], np.loadtxt(runs["fast"] / f)[:, 1:]
        assert p.shape == q.shape
        assert np.abs(p - q).max() <= 1e-6 * np.abs(p).max()
