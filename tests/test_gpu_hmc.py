"""The HMC sibling (mh/cheisdwf_hmc.f) on the GPU: force with its
chronological guess and the Hamiltonian against the reference's own build
(golden vectors) and the oracle; trajectories against an oracle-driven loop in
the same order of random draws; the driver's files."""
import importlib.util
import os

import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import checkpoint, hmc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_inputs():
    spec = importlib.util.spec_from_file_location("mkh", os.path.join(HERE, "golden", "make_golden_hmc.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    g = np.load(os.path.join(HERE, "golden", "ref_hmc_l12_ls8.npz"))
    return g, mk.inputs(int(g["seed"]))


def test_hmc_force_hamilton_vs_reference_golden():
    g, (phi, chi, pp, x0, idx) = golden_inputs()
    lat = fm.Lattice(12, 8)
    for imass in (1, 3):
        xf = x0.copy()
        F, it = fm.hmc_force(lat, phi, xf, float(g["res"]), chi, float(g["beta"]), imass)
        Fr = g["force_m%d" % imass]
        assert np.abs(F - Fr).max() <= 1e-9 * np.abs(Fr).max(), np.abs(F - Fr).max()
        assert np.abs(xf.reshape(-1)[idx] - g["force_x0_m%d_s" % imass]).max() <= 1e-10 * np.abs(xf).max()
        xh = x0.copy()
        h, _ = fm.hmc_hamilton(lat, phi, xh, float(g["res"]), chi, pp, float(g["beta"]), imass)
        assert np.allclose(h, g["ham_m%d" % imass], rtol=1e-12, atol=0), (h, g["ham_m%d" % imass])
        assert np.abs(xh.reshape(-1)[idx] - g["ham_x0_m%d_s" % imass]).max() <= 1e-10 * np.abs(xh).max()
    lat.close()


def test_hmc_force_vs_oracle_other_size():
    lat, olat = fm.Lattice(6, 4, 8), orc.Lattice(6, 4, 8)
    rs = np.random.RandomState(12)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    x0 = 0.1 * spinor(rs, lat.shape)
    xg = x0.copy()
    F, it = fm.hmc_force(lat, phi, xg, 1e-11, chi, 0.45, 3)
    Fo, xo, ito = orc.hmc_force(olat, phi, x0, 1e-11, chi, 0.45, 3)
    assert np.abs(F - Fo).max() <= 1e-9 * np.abs(Fo).max()
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert abs(it - ito) <= 5
    lat.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_hmc_force_reduced_precision_vs_oracle(precision):
    """The force's CG as the reliable-update CG (FP32 or binary16 search
    directions, FP64 true residual at the congrad rule's threshold) from the
    same guess: force and chronological guess against the oracle's FP64
    congrad."""
    lat, olat = fm.Lattice(6, 4, 8), orc.Lattice(6, 4, 8)
    rs = np.random.RandomState(12)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    x0 = 0.1 * spinor(rs, lat.shape)
    Fo, xo, ito = orc.hmc_force(olat, phi, x0, 1e-11, chi, 0.45, 3)
    fm.set_force_precision(lat, precision)
    xg = x0.copy()
    F, it = fm.hmc_force(lat, phi, xg, 1e-11, chi, 0.45, 3)
    ef = np.abs(F - Fo).max() / np.abs(Fo).max()
    ex = np.abs(xg - xo).max() / np.abs(xo).max()
    print("%s force CG: force error %.2e, guess error %.2e, iterations %d (oracle congrad %d)" % (
        precision, ef, ex, it, ito))
    assert ef <= 1e-8 and ex <= 1e-8
    lat.close()


def test_hmc_trajectories_fp16_force_match_oracle():
    """Trajectories with the binary16-form force CG against the oracle's
    FP64 loop: the same lengths and decisions, dH within 1e-6."""
    L, Ls, ntraj = 4, 4, 6
    par = dict(dt=0.06, beta=0.45, am=0.0, imass=1, iterl=3, itermax=1000, rescga=1e-11, rescgg=1e-10)
    rs = np.random.RandomState(41)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.45)
    seed = 693764523745.0
    lat = fm.Lattice(L, Ls)
    fm.set_force_precision(lat, "fp16")
    chi_g = chi0.copy()
    rng = hmc.RandomStream(seed)
    gpu = [hmc.trajectory(lat, chi_g, rng, i + 1, par) for i in range(ntraj)]
    olat = orc.Lattice(L, Ls)
    chi_c = chi0.copy()
    orng = orc.Rng(seed)
    cpu = [orc.hmc_trajectory(olat, chi_c, orng, par) for _ in range(ntraj)]
    ddh = max(abs(g["dH"] - c["dH"]) for g, c in zip(gpu, cpu))
    print("fp16 force CG: accepted %s, max |dH - dH_oracle| %.2e" % ([g["accepted"] for g in gpu], ddh))
    assert [g["md_steps"] for g in gpu] == [c["md_steps"] for c in cpu]
    assert [g["accepted"] for g in gpu] == [c["accepted"] for c in cpu]
    assert ddh <= 1e-6
    lat.close()


@pytest.mark.parametrize("imass", [1, 3])
def test_hmc_trajectories_match_oracle(imass):
    L, Ls, ntraj = 4, 4, 6
    par = dict(dt=0.06, beta=0.45, am=0.0, imass=imass, iterl=3, itermax=1000, rescga=1e-11, rescgg=1e-10)
    rs = np.random.RandomState(40 + imass)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.45)
    seed = 693764523745.0
    lat = fm.Lattice(L, Ls)
    chi_g = chi0.copy()
    rng = hmc.RandomStream(seed)
    gpu = [hmc.trajectory(lat, chi_g, rng, i + 1, par) for i in range(ntraj)]
    olat = orc.Lattice(L, Ls)
    chi_c = chi0.copy()
    orng = orc.Rng(seed)
    cpu = [orc.hmc_trajectory(olat, chi_c, orng, par) for _ in range(ntraj)]
    print("accepted", [g["accepted"] for g in gpu], "dH", [round(c["dH"], 4) for c in cpu])
    assert [g["md_steps"] for g in gpu] == [c["md_steps"] for c in cpu]
    assert [g["accepted"] for g in gpu] == [c["accepted"] for c in cpu]
    for g, c in zip(gpu, cpu):
        assert abs(g["dH"] - c["dH"]) <= 1e-7 * max(1.0, abs(c["dH"]))   # H ~ 1e3, solves at 1e-11
    assert np.abs(chi_g - chi_c).max() <= 1e-9 * np.abs(chi_c).max()
    assert rng.seed() == orng.get()
    lat.close()


def test_hmc_run_files(tmp_path):
    open(tmp_path / "midout", "w").write("0.3 0.45 1.0 0.00 1 3 3\ndt beta am3 am imass iterl iter2\n")
    rs = np.random.RandomState(3)
    checkpoint.write_con(tmp_path / "con", rs.randn(3, 64) / np.sqrt(0.45), 693764523745.0)
    out = hmc.run(str(tmp_path), 4, 4, iprint=1)
    assert out["trajectories"] == 3
    assert len(open(tmp_path / "control").readlines()) == 3
    assert len(open(tmp_path / "fort.14").readlines()) == 3
    chi, seed, _ = checkpoint.read_con(tmp_path / "con", kvol=64)
    assert np.all(np.isfinite(chi)) and seed != 693764523745.0
    # a second run from the con the first wrote: the sibling's build starts
    # control and fort.11-14 afresh (no position='append', mh/cheisdwf_hmc.f:74)
    hmc.run(str(tmp_path), 4, 4, iprint=1, iter2=2)
    assert len(open(tmp_path / "control").readlines()) == 2
    assert len(open(tmp_path / "fort.11").readlines()) == 2


def _hmc_md_start(seed):
    """L=4, Ls=8 (the congrad cap Ls*kvol = 512 lies above the ~290
    iterations these solves take), chi ~ N(0, 1/beta), Gaussian momenta,
    Phi = M^dagger R."""
    rs = np.random.RandomState(seed)
    lat = fm.Lattice(4, 8)
    chi0 = rs.randn(3, lat.kvol) / np.sqrt(0.45)
    pp0 = rs.randn(3, lat.kvol)
    Phi = fm.dslashd(lat, spinor(rs, lat.shape), chi0, 0.0, 0.0, 1)
    return lat, chi0, pp0, Phi


def test_hmc_leapfrog_reversible():
    """The sibling's MD evolution is time-reversible: 8 steps forward,
    momenta reversed, 8 back (the chronological guess carried on) return chi
    and the momenta to their start, to what the force solves' tolerance
    (1e-13) allows (the oracle: 8e-15 and 6e-14)."""
    lat, chi0, pp0, Phi = _hmc_md_start(9)
    par = dict(hmc.DEFAULTS, dt=0.06, rescgg=1e-13)
    x0 = Phi.copy()
    chit, pp = chi0.copy(), pp0.copy()
    assert hmc.leapfrog(lat, Phi, x0, chit, pp, par, lambda s: s == 8)[0] == 8
    moved = np.abs(chit - chi0).max()
    pp = -pp
    assert hmc.leapfrog(lat, Phi, x0, chit, pp, par, lambda s: s == 8)[0] == 8
    back = np.abs(chit - chi0).max() / np.abs(chi0).max()
    pback = np.abs(pp + pp0).max() / np.abs(pp0).max()
    print("moved %.3f, return error chi %.2e, momenta %.2e" % (moved, back, pback))
    assert moved > 0.1
    assert back <= 1e-11 and pback <= 1e-11
    lat.close()


def test_hmc_leapfrog_dH_second_order():
    """|dH| over a trajectory of fixed length 0.4 falls as dt^2: halving dt
    from 0.05 to 0.025 to 0.0125 divides it by ~4 (the oracle: 4.61, 4.14).
    The sibling's action has no rational approximation, so nothing but the
    solver tolerance puts a floor under dH."""
    lat, chi0, pp0, Phi = _hmc_md_start(9)

    def dH(dt):
        n = int(round(0.4 / dt))
        par = dict(hmc.DEFAULTS, dt=dt, rescgg=1e-12)
        chit, pp, x0 = chi0.copy(), pp0.copy(), Phi.copy()
        (h0, *_), _ = fm.hmc_hamilton(lat, Phi, x0.copy(), 1e-12, chit, pp, 0.45, 1)
        assert hmc.leapfrog(lat, Phi, x0, chit, pp, par, lambda s: s == n)[0] == n
        (h1, *_), _ = fm.hmc_hamilton(lat, Phi, x0.copy(), 1e-12, chit, pp, 0.45, 1)
        return h1 - h0

    d = [dH(dt) for dt in (0.05, 0.025, 0.0125)]
    print("dH at dt 0.05, 0.025, 0.0125:", d)
    assert 3.5 < d[0] / d[1] < 5.0 and 3.5 < d[1] / d[2] < 5.0, d
    lat.close()
