"""GPU parity of the solvers, force and Hamiltonian against the CPU oracle.

Tolerances: solutions of the multi-shift solve and CG agree to the solver
tolerance; iteration counts agree to within a few iterations (the stopping
tests are thresholds on FP64 sums whose summation order differs between the
GPU tree reductions and the oracle's chunked sums); forces and H to 1e-10
relative when both sides run the force solves at a tight tolerance."""
import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import remez  # noqa: E402

R4, R2, R2G = remez.load("remez4"), remez.load("remez2"), remez.load("remez2g")
_lats = {}


def lattices(L, Ls, Lt=None):
    key = (L, Ls, Lt)
    if key not in _lats:
        _lats[key] = (fm.Lattice(L, Ls, Lt), orc.Lattice(L, Ls, Lt))
    return _lats[key]


@pytest.mark.parametrize("L,Ls", [(4, 4), (8, 8)])
@pytest.mark.parametrize("imass", [1, 3])
def test_qmrherm_iflag0(L, Ls, imass):
    lat, olat = lattices(L, Ls)
    rs = np.random.RandomState(20 + L + imass)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R4[0], R4[1]
    xg, ig = fm.qmrherm(lat, phi, 1e-9, a, ad, chi, 0.0, 0.3, imass, 0)
    xc, ic = orc.qmrherm(olat, phi, 1e-9, a, ad, chi, 0.3, imass, 0)
    assert abs(ig - ic) <= 5, (ig, ic)
    err = np.abs(xg - xc).max() / np.abs(xc).max()
    assert err < 1e-9, err


def test_qmrherm_iflag1_force_accumulates():
    lat, olat = lattices(8, 8)
    rs = np.random.RandomState(31)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R2G[0], R2G[1]
    F0 = rs.randn(3, lat.kvol)
    Fg, Fc = F0.copy(), F0.copy()
    xg, ig = fm.qmrherm(lat, phi, 1e-12, a, ad, chi, 0.0, 0.0, 1, 1, Fg)
    xc, ic = orc.qmrherm(olat, phi, 1e-12, a, ad, chi, 0.0, 1, 1, Fc)
    assert np.allclose(xg, a[0] * phi, rtol=0, atol=1e-15)
    err = np.abs(Fg - Fc).max() / np.abs(Fc - F0).max()
    assert err < 1e-10, err


def test_congrad_reference_rule():
    lat, olat = lattices(8, 8)
    rs = np.random.RandomState(41)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    x0 = 0.1 * spinor(rs, lat.shape)
    xg, ig = fm.congrad(lat, phi, 1e-9, chi, x0)
    xc, ic = orc.congrad(olat, phi, 1e-9, chi, x0)
    assert abs(ig - ic) <= 3, (ig, ic)
    err = np.abs(xg - xc).max() / np.abs(xc).max()
    assert err < 1e-6, err


@pytest.mark.parametrize("imass", [1, 3])
def test_force_parity(imass):
    """Force solves at res 1e-12 so that the solver tolerance is far below
    the 1e-10 comparison."""
    lat, olat = lattices(6, 6)
    rs = np.random.RandomState(51 + imass)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    Fg, xg, itg = fm.force(lat, phi, 1e-12, chi, 0.475, R2G, 0.0, 0.2, imass)
    Fc, xc, itc = orc.force(olat, phi, 1e-12, chi, 0.475, R2G, 0.2, imass)
    err = np.abs(Fg - Fc).max() / np.abs(Fc).max()
    assert err < 1e-10, (err, itg, itc)


@pytest.mark.parametrize("L,Ls,Lt", [(5, 6, 6), (6, 3, 4)])
@pytest.mark.parametrize("imass", [1, 3])
def test_force_hamilton_other_shapes(L, Ls, Lt, imass):
    """An odd L, Lt != L and an odd Ls: force and Hamiltonian (all three
    parts) against the oracle to 1e-10 relative at solver tolerance 1e-12,
    with the pinning field on."""
    lat, olat = lattices(L, Ls, Lt)
    rs = np.random.RandomState(100 + L + Ls + imass)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    Fg, _, itg = fm.force(lat, phi, 1e-12, chi, 0.475, R2G, 0.0, 0.2, imass)
    Fc, _, itc = orc.force(olat, phi, 1e-12, chi, 0.475, R2G, 0.2, imass)
    err = np.abs(Fg - Fc).max() / np.abs(Fc).max()
    assert err < 1e-10, (err, itg, itc)
    pp = rs.randn(3, lat.kvol)
    hg, _, _ = fm.hamilton(lat, phi, 1e-12, chi, pp, 0.475, R2, hpin=0.2, imass=imass)
    hc, _, _ = orc.hamilton(olat, phi, 1e-12, chi, pp, 0.475, R2, 0.2, imass)
    for a, b in zip(hg, hc):
        assert abs(a - b) <= 1e-10 * abs(b), (hg, hc)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp16"])
@pytest.mark.parametrize("L,Ls,Lt", [(5, 6, 6), (6, 3, 4)])
def test_solve_mdagm_other_shapes(prec, L, Ls, Lt):
    """Every form of the M^dagger M solve on an odd L (the one-site FP32
    stencil instead of the paired one), Lt != L and an odd Ls: the FP64 true
    residual, recomputed with the oracle's operator, meets 1e-10."""
    lat, olat = lattices(L, Ls, Lt)
    rs = np.random.RandomState(110 + L + Ls)
    b, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    bd = fm.spinor_dev(lat, b)
    xd = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    st = fm.solve_mdagm(lat, bd, xd, fm.DeviceArray.from_numpy(lat, chi), 1e-10, prec)
    x = fm.to_host(lat, xd)
    r = b - orc.dslash(olat, orc.dslash(olat, x, chi, 0.0, 1, False), chi, 0.0, 1, True)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert rel <= 1.05e-10, (prec, rel, st)


@pytest.mark.parametrize("imass", [1, 3])
def test_force_fp32_solves(imass):
    """Force with FP32-storage multi-shift solves (FP64 Lanczos scalars,
    reductions, derivs bilinears and sums) against the FP64 oracle at the
    reference's force tolerance 1e-6 and at 1e-9: the error of the fermion
    part is stated and bounded by 1e-4 relative (measured 1.4e-5 at 1e-6,
    L=8, Ls=8, imass=1)."""
    lat = fm.Lattice(8, 8)
    olat = orc.Lattice(8, 8)
    rs = np.random.RandomState(71 + imass)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    fm.set_force_precision(lat, "fp32")
    for res in (1e-6, 1e-9):
        Fg, _, itg = fm.force(lat, phi, res, chi, 0.475, R2G, 0.0, 0.2, imass)
        Fc, _, itc = orc.force(olat, phi, res, chi, 0.475, R2G, 0.2, imass)
        fg, fc = Fg - 2 * 0.475 * chi, Fc - 2 * 0.475 * chi
        err = np.linalg.norm(fg - fc) / np.linalg.norm(fc)
        print("imass %d res %g: FP32-solve force error %.2e (iterations %s vs %s)" % (imass, res, err, itg, itc))
        assert err < 1e-4, err
        assert abs(itg[0] - itc[0]) <= 0.3 * itc[0] + 5
    lat.close()


def test_hamilton_parity():
    lat, olat = lattices(8, 8)
    rs = np.random.RandomState(61)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    pp = rs.randn(3, lat.kvol)
    hg, _, itg = fm.hamilton(lat, phi, 1e-11, chi, pp, 0.475, R2)
    hc, _, itc = orc.hamilton(olat, phi, 1e-11, chi, pp, 0.475, R2)
    for a, b in zip(hg, hc):
        assert abs(a - b) <= 1e-10 * abs(b), (hg, hc)


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_solve_mdagm_deterministic(prec, monkeypatch):
    """The reductions sum in a fixed order, so a solve is reproducible bit for
    bit: twice in one context (iteration chunks replayed from a hipGraph) and
    once in a context that launches them one by one (CHS_GRAPHS=0)."""
    rs = np.random.RandomState(73)
    b, chi = spinor(rs, (2, 4, 8, 8, 8, 8)), phi_field(rs, 512)
    xs, its = [], []

    def solve(lat):
        bd = fm.spinor_dev(lat, b)
        xd = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
        st = fm.solve_mdagm(lat, bd, xd, fm.DeviceArray.from_numpy(lat, chi), 1e-10, prec)
        xs.append(fm.to_host(lat, xd))
        its.append((st["iterations"], st["reliable_updates"]))

    monkeypatch.setenv("CHS_GRAPHS", "1")
    lat = fm.Lattice(8, 8)
    assert lat.shape == b.shape
    solve(lat)
    solve(lat)
    lat.close()
    monkeypatch.setenv("CHS_GRAPHS", "0")
    lat = fm.Lattice(8, 8)
    solve(lat)
    lat.close()
    assert its[0] == its[1] == its[2], its
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_solve_mdagm_large_lattice(prec):
    """L=32, Ls=20: 5120 workgroup partials per stencil, more than the 4096
    the reductions take in their first batch of loads, and a 5.2M-element
    spinor; the FP64 true residual is checked with the oracle's operator."""
    lat = fm.Lattice(32, 20)
    olat = orc.Lattice(32, 20)
    rs = np.random.RandomState(74)
    b, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    bd = fm.spinor_dev(lat, b)
    xd = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    st = fm.solve_mdagm(lat, bd, xd, fm.DeviceArray.from_numpy(lat, chi), 1e-10, prec)
    x = fm.to_host(lat, xd)
    lat.close()
    r = b - orc.dslash(olat, orc.dslash(olat, x, chi, 0.0, 1, False), chi, 0.0, 1, True)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert rel <= 1.05e-10, (prec, rel, st)


@pytest.mark.parametrize("scale", [1e-20, 1e20])
def test_solve_mdagm_fp16_scaled_source(scale):
    """The binary16 form stores p and M p divided by the residual's rms: a
    source 1e-20 or 1e+20 times a unit one (far outside binary16's 6e-8 ..
    65504) reaches the same FP64 true relative residual, and the solution is
    the scaled solution of the unit source (to what two 1e-10 solves
    allow)."""
    lat, olat = lattices(8, 8)
    rs = np.random.RandomState(72)
    b, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    chid = fm.DeviceArray.from_numpy(lat, chi)
    xs = []
    for sc in (1.0, scale):
        bd = fm.spinor_dev(lat, sc * b)
        xd = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
        st = fm.solve_mdagm(lat, bd, xd, chid, 1e-10, "fp16")
        x = fm.to_host(lat, xd)
        r = sc * b - orc.dslash(olat, orc.dslash(olat, x, chi, 0.0, 1, False), chi, 0.0, 1, True)
        rel = np.linalg.norm(r) / np.linalg.norm(sc * b)
        assert rel <= 1.05e-10, (sc, rel, st)
        xs.append(x / sc)
    assert np.linalg.norm(xs[1] - xs[0]) <= 5e-6 * np.linalg.norm(xs[0])


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp16"])
def test_solve_mdagm_true_residual(prec):
    """Every form reaches an FP64 true relative residual <= 1e-10, checked
    here independently with the oracle's operator."""
    lat, olat = lattices(8, 8)
    rs = np.random.RandomState(71)
    b, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    bd = fm.spinor_dev(lat, b)
    xd = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    chid = fm.DeviceArray.from_numpy(lat, chi)
    st = fm.solve_mdagm(lat, bd, xd, chid, 1e-10, prec)
    x = fm.to_host(lat, xd)
    r = b - orc.dslash(olat, orc.dslash(olat, x, chi, 0.0, 1, False), chi, 0.0, 1, True)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert rel <= 1.05e-10, (prec, rel, st)
    assert abs(st["true_resid"] - rel) <= 1e-3 * rel + 1e-16


def test_qmrherm_rejects_iflag_other_than_0_or_1():
    from chiralheisenberg_amd._lib import ChsError
    lat, _ = lattices(4, 4)
    rs = np.random.RandomState(93)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    for iflag in (2, -1):
        with pytest.raises(ChsError):
            fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi, 0.0, 0.0, 1, iflag, np.zeros((3, lat.kvol)))
