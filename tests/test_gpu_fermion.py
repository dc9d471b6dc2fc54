"""Wall-source fermion correlators (rhmc/cheisdwf.f:897-1058) on the GPU:
against the reference's own routine (golden vectors on its thermalised
configuration rhmc/_con, L=12, Ls=8) and against the oracle at other sizes;
the FP32 reliable-update solves to the same stopping residual."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import phi_field

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import checkpoint  # noqa: E402
from chiralheisenberg_amd._lib import ChsError  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fermion_vs_reference_golden():
    """imass 1 and 3 with the antiperiodic t boundary, and imass 1 periodic
    (ibound = +1)."""
    from test_oracle import fermion_cases
    g = np.load(os.path.join(HERE, "golden", "ref_fermion_l12_ls8.npz"))
    chi, _, _ = checkpoint.read_con(os.path.join(HERE, "golden", "con_l12_fp64"), kvol=1728)
    cases = list(fermion_cases(g))
    assert len(cases) == 3
    for k, ibound, tol in cases:
        lat = fm.Lattice(12, 8, ibound=ibound)
        rng = orc.Rng(float(g["rng_seed"]))   # the reference generator, restated
        rng.rano(0)
        ittt0 = int(12 * rng.rano(0)) + 1
        c1, c2, it = fm.fermion(lat, chi, float(g["case%d_res" % k]), ittt0, float(g["case%d_hpin" % k]),
                                int(g["case%d_imass" % k]))
        r1, r2 = g["case%d_cferm1" % k], g["case%d_cferm2" % k]
        e1 = np.abs(c1 - r1).max() / np.abs(r1).max()
        e2 = np.abs(c2 - r2).max() / np.abs(r2).max()
        print("case %d: cferm1 %.2e cferm2 %.2e, aviter %.1f vs %.1f" % (k, e1, e2, it / 10,
                                                                          float(g["case%d_aviter" % k])))
        assert e1 <= tol and e2 <= tol
        assert abs(it / 10 - float(g["case%d_aviter" % k])) <= 3
        lat.close()


@pytest.mark.parametrize("L,Ls,Lt,imass,ittt0", [(6, 4, 10, 1, 3), (8, 6, 8, 3, 8), (5, 3, 6, 1, 6)])
def test_fermion_vs_oracle(L, Ls, Lt, imass, ittt0):
    lat = fm.Lattice(L, Ls, Lt)
    olat = orc.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(L * 10 + Ls)
    chi = phi_field(rs, lat.kvol)
    c1, c2, it = fm.fermion(lat, chi, 1e-11, ittt0, 0.3, imass)
    o1, o2, oit = orc.fermion(olat, chi, 1e-11, ittt0, 0.3, imass)
    assert np.abs(c1 - o1).max() <= 1e-8 * np.abs(o1).max()
    assert np.abs(c2 - o2).max() <= 1e-8 * np.abs(o2).max()
    assert abs(it - oit) <= 20
    # FP32 and scaled-FP16 reliable-update solves to the same stopping residual
    for prec in ("fp32", "fp16"):
        f1, f2, fit = fm.fermion(lat, chi, 1e-11, ittt0, 0.3, imass, precision=prec)
        e = max(np.abs(f1 - o1).max() / np.abs(o1).max(), np.abs(f2 - o2).max() / np.abs(o2).max())
        print("L=%d Ls=%d: %s correlator error %.2e, iterations %d vs %d" % (L, Ls, prec, e, fit, it))
        assert e <= 1e-6
    lat.close()


def test_fermion_rejects_bad_source_slice():
    lat = fm.Lattice(4, 4)
    chi = np.zeros((3, lat.kvol))
    with pytest.raises(ChsError):
        fm.fermion(lat, chi, 1e-8, 0)
    with pytest.raises(ChsError):
        fm.fermion(lat, chi, 1e-8, 5)
    lat.close()


def test_fermion_congrad_cap_carries_on():
    """L=Lt=2, Ls=2: congrad's cap Ls*kvol = 16 lies below what res 1e-11
    needs, so solves stop there.  The library keeps their last iterates, as
    the reference's congrad does (rhmc/cheisdwf.f:1081,1151), writes the
    correlators and reports how many solves stopped (NoConvergence, its
    message's leading count, the results attached); they equal the oracle's,
    which carries on the same way."""
    from chiralheisenberg_amd._lib import NoConvergence
    lat, olat = fm.Lattice(2, 2), orc.Lattice(2, 2)
    chi = phi_field(np.random.RandomState(3), lat.kvol)
    with pytest.raises(NoConvergence) as ei:
        fm.fermion(lat, chi, 1e-11, 1, 0.3, 1)
    c1, c2, it = ei.value.result
    ncap = int(ei.value.msg.split()[0])
    o1, o2, oit = orc.fermion(olat, chi, 1e-11, 1, 0.3, 1)
    print("%d of 10 solves at the cap; iterations %d (oracle %d)" % (ncap, it, oit))
    assert 1 <= ncap <= 10 and it <= 10 * 16
    assert np.abs(c1 - o1).max() <= 1e-10 * np.abs(o1).max()
    assert np.abs(c2 - o2).max() <= 1e-10 * np.abs(o2).max()
    lat.close()


def test_drivers_write_the_congrad_message(tmp_path):
    """rhmc.run and hmc.run at L=Lt=2, Ls=2, where congrad stops at its cap:
    both runs end normally, and their output holds the programs' message
    ' # iterations of congrad exceeds niterc' between the header and the
    averages, once per capped solve."""
    from chiralheisenberg_amd import checkpoint, hmc, outfile, rhmc
    rs = np.random.RandomState(4)
    chi0 = rs.randn(3, 8) / np.sqrt(0.475)
    for name, run, mid in (("rhmc", rhmc.run, "0.1 0.475 1.0 0.00 0.0 1 3 1\n"),
                           ("hmc", hmc.run, "0.1 0.475 1.0 0.00 1 3 1\n")):
        d = tmp_path / name
        d.mkdir()
        (d / "midout").write_text(mid)
        checkpoint.write_con(d / "con", chi0, 122110675517440.0)
        run(str(d), 2, 2, iprint=1)
        lines = open(d / "output").read().split("\n")
        k = [i for i, l in enumerate(lines) if l == outfile.CONGRAD_CAP]
        first_avg = next(i for i, l in enumerate(lines) if "averages for last" in l)
        print("%s: %d congrad messages" % (name, len(k)))
        assert k and 9 <= min(k) and max(k) < first_avg
