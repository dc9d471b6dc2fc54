"""RHMC trajectories (program dwf3d, rhmc/cheisdwf.f:207-413) with every
fermion routine on the GPU against the oracle's own trajectory loop, from
the same seed and the same chi: identical trajectory lengths and
accept/reject decisions, and H0, H1, dH in agreement.  The solves run at
tight tolerances (1e-11 / 1e-10) so that the comparison is not limited by
the solver tolerance; the reference's production tolerances (1e-9, 1e-6)
are the defaults of both drivers."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import remez, rhmc  # noqa: E402


@pytest.mark.parametrize("L,Ls,ntraj,imass,dt", [(4, 4, 8, 1, 0.45), (4, 6, 4, 3, 0.3)])
def test_trajectories_match_oracle(L, Ls, ntraj, imass, dt):
    """dt = 0.45 at L=4, Ls=4 gives a mix of accepted and rejected proposals."""
    par = dict(dt=dt, beta=0.475, hpin=0.0, imass=imass, iterl=3, itermax=1000, iprint=2,
               rescga=1e-11, rescgg=1e-10, am=0.0)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    rs = np.random.RandomState(100 + L)
    chi0 = rs.randn(3, L * L * L) / np.sqrt(0.475)
    seed = 693764523745.0

    lat = fm.Lattice(L, Ls)
    chi_g = chi0.copy()
    rng = rhmc.RandomStream(seed)
    gpu = [rhmc.trajectory(lat, chi_g, rng, isweep=i + 1, par=par, sets=sets) for i in range(ntraj)]

    olat = orc.Lattice(L, Ls)
    chi_c = chi0.copy()
    orng = orc.Rng(seed)
    cpu = orc.rhmc(olat, par, orng, chi_c, 1, ntraj, sets)

    assert [g["md_steps"] for g in gpu] == [c["md_steps"] for c in cpu]
    assert [g["accepted"] for g in gpu] == [c["accepted"] for c in cpu]
    for g, c in zip(gpu, cpu):
        for k in ("H0", "H1"):
            assert abs(g[k] - c[k]) <= 1e-10 * abs(c[k]), (k, g[k], c[k])
        assert abs(g["dH"] - c["dH"]) <= 1e-8 * max(1.0, abs(c["dH"])), (g["dH"], c["dH"])
    assert np.abs(chi_g - chi_c).max() <= 1e-9 * np.abs(chi_c).max()
    assert rng.seed() == orng.get()
    lat.close()


def test_trajectories_fp32_force_solves():
    """Production tolerances (1e-9 action, 1e-6 force) with the force's
    multi-shift solves in FP32 storage against the FP64 oracle: the same
    trajectory lengths and accept/reject decisions; the dH difference is
    stated and bounded."""
    L, Ls, ntraj = 4, 4, 8
    par = dict(dt=0.45, beta=0.475, hpin=0.0, imass=1, iterl=3, itermax=1000, iprint=2,
               rescga=1e-9, rescgg=1e-6, am=0.0)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    rs = np.random.RandomState(104)
    chi0 = rs.randn(3, L * L * L) / np.sqrt(0.475)
    seed = 693764523745.0
    lat = fm.Lattice(L, Ls)
    fm.set_force_precision(lat, "fp32")
    chi_g = chi0.copy()
    rng = rhmc.RandomStream(seed)
    gpu = [rhmc.trajectory(lat, chi_g, rng, isweep=i + 1, par=par, sets=sets) for i in range(ntraj)]
    olat = orc.Lattice(L, Ls)
    chi_c = chi0.copy()
    cpu = orc.rhmc(olat, par, orc.Rng(seed), chi_c, 1, ntraj, sets)
    ddh = [abs(g["dH"] - c["dH"]) for g, c in zip(gpu, cpu)]
    print("accepted %s; |dH(fp32 force) - dH(fp64 oracle)| max %.2e; dH %s" % (
        [g["accepted"] for g in gpu], max(ddh), [round(c["dH"], 4) for c in cpu]))
    assert [g["md_steps"] for g in gpu] == [c["md_steps"] for c in cpu]
    assert [g["accepted"] for g in gpu] == [c["accepted"] for c in cpu]
    assert max(ddh) <= 1e-4
    lat.close()


def test_driver_run_matches_oracle(tmp_path):
    """rhmc.run (program dwf3d: midout + con in, control / fort.11-14 /
    fort.50xx / con / output out, measurements every iprint trajectories)
    against the oracle's trajectory loop and fermion measurement driven in
    the same order of random draws."""
    from chiralheisenberg_amd import checkpoint
    L, Ls, iprint, iter2 = 4, 4, 2, 4
    open(tmp_path / "midout", "w").write("0.3 0.475 1.0 0.00 0.2 1 3 %d\ndt beta am3 am h0 imass iterl iter2\n" % iter2)
    rs = np.random.RandomState(7)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.475)
    seed = 122110675517440.0
    checkpoint.write_con(tmp_path / "con", chi0, seed)
    out = rhmc.run(str(tmp_path), L, Ls, iprint=iprint, rescgm=1e-10)

    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    par = dict(dt=0.3, beta=0.475, hpin=0.2, imass=1, iterl=3, itermax=1000, iprint=0, rescga=1e-9, rescgg=1e-6,
               am=0.0)
    olat = orc.Lattice(L, Ls)
    chi = chi0.copy()
    orng = orc.Rng(seed)
    orng.rano(0)
    control, corr = [], []
    for isweep in range(1, iter2 + 1):
        c = orc.rhmc(olat, par, orng, chi, isweep, 1, sets)[0]
        control.append((c["dH"], c["accepted"]))
        if isweep % iprint == 0:
            ittt0 = int(L * orng.rano(0)) + 1
            corr.append(orc.fermion(olat, chi, 1e-10, ittt0, 0.2, 1))
    got = [l.split() for l in open(tmp_path / "control")]
    assert [int(g[2]) for g in got] == [a for _, a in control]
    for g, (dh, _) in zip(got, control):
        assert abs(float(g[0]) - dh) <= 1e-6 * max(1.0, abs(dh))   # production tolerances 1e-9 / 1e-6
    c5001 = np.loadtxt(tmp_path / "fort.5001").reshape(len(corr), L, 3)
    c5012 = np.loadtxt(tmp_path / "fort.5012").reshape(len(corr), L, 3)
    for m, (o1, o2, _) in enumerate(corr):
        assert np.abs(c5001[m, :, 1] + 1j * c5001[m, :, 2] - o1[0]).max() <= 1e-7 * np.abs(o1).max()
        assert np.abs(c5012[m, :, 1] + 1j * c5012[m, :, 2] - o2[1]).max() <= 1e-7 * np.abs(o2).max()
    chi_end, seed_end, _ = checkpoint.read_con(tmp_path / "con", kvol=L ** 3)
    assert np.abs(chi_end - chi).max() <= 1e-9 * np.abs(chi).max()
    assert seed_end == orng.get()
    assert out["acceptances"] == sum(a for _, a in control)
    assert len(open(tmp_path / "fort.14").readlines()) == iter2 // iprint
    # a second run from the con the first wrote: dwf3d appends to control,
    # fort.11-14 and the correlators (position='append', rhmc/cheisdwf.f:100-113)
    # and starts output afresh (two seed lines in its header, one at the end)
    n5001 = len(open(tmp_path / "fort.5001").readlines())
    rhmc.run(str(tmp_path), L, Ls, iter2=iprint, iprint=iprint, rescgm=1e-10)
    assert len(open(tmp_path / "control").readlines()) == iter2 + iprint
    assert len(open(tmp_path / "fort.11").readlines()) == iter2 + iprint
    assert len(open(tmp_path / "fort.5001").readlines()) == n5001 + L
    assert open(tmp_path / "output").read().count("seed:") == 3


def _md_start(L, Ls, seed):
    """A lattice, chi ~ N(0, 1/beta), Gaussian momenta and a heatbath Phi."""
    from conftest import spinor
    rs = np.random.RandomState(seed)
    lat = fm.Lattice(L, Ls)
    chi0 = rs.randn(3, lat.kvol) / np.sqrt(0.475)
    pp0 = rs.randn(3, lat.kvol)
    a4, ad4, _, _ = remez.load("remez4")
    Phi, _ = fm.qmrherm(lat, spinor(rs, lat.shape), 1e-12, a4, ad4, chi0, 0.0, 0.0, 1, 0)
    return lat, chi0, pp0, Phi


def test_leapfrog_reversible():
    """The MD evolution (rhmc/cheisdwf.f:261-315) is time-reversible: n steps
    forward, momenta reversed, n steps back return chi and the momenta to
    their start, to what the force solves' tolerance (1e-13) allows."""
    lat, chi0, pp0, Phi = _md_start(4, 4, 7)
    par = dict(rhmc.DEFAULTS, dt=0.1, rescgg=1e-13)
    r2g = remez.load("remez2g")
    chit, pp = chi0.copy(), pp0.copy()
    assert rhmc.leapfrog(lat, Phi, chit, pp, par, r2g, lambda it: it == 8) == 8
    moved = np.abs(chit - chi0).max()
    pp = -pp
    assert rhmc.leapfrog(lat, Phi, chit, pp, par, r2g, lambda it: it == 8) == 8
    back = np.abs(chit - chi0).max() / np.abs(chi0).max()
    pback = np.abs(pp + pp0).max() / np.abs(pp0).max()
    print("moved %.3f, return error chi %.2e, momenta %.2e" % (moved, back, pback))
    assert moved > 0.1
    assert back <= 1e-10 and pback <= 1e-10
    lat.close()


def test_leapfrog_dH_second_order():
    """|dH| over a trajectory of fixed length 0.4 falls as dt^2 (leapfrog is
    second order): halving dt divides it by ~4.  The force uses the remez2
    set here, the Hamiltonian's own, so that no mismatch of two rational
    approximations puts a floor under dH."""
    lat, chi0, pp0, Phi = _md_start(4, 4, 8)
    r2 = remez.load("remez2")

    def dH(dt):
        n = int(round(0.4 / dt))
        par = dict(rhmc.DEFAULTS, dt=dt, rescgg=1e-12)
        chit, pp = chi0.copy(), pp0.copy()
        (h0, *_), _, _ = fm.hamilton(lat, Phi, 1e-12, chit, pp, 0.475, r2, 0.0, 0.0, 1, 1)
        assert rhmc.leapfrog(lat, Phi, chit, pp, par, r2, lambda it: it == n) == n
        (h1, *_), _, _ = fm.hamilton(lat, Phi, 1e-12, chit, pp, 0.475, r2, 0.0, 0.0, 1, 1)
        return h1 - h0

    d = [dH(dt) for dt in (0.1, 0.05, 0.025)]
    print("dH at dt 0.1, 0.05, 0.025:", d)
    assert 3.0 < d[0] / d[1] < 5.0 and 3.0 < d[1] / d[2] < 5.0, d
    lat.close()
