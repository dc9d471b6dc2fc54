"""The drop-in boundary's conventions on the GPU:

* derivs (rhmc/cheisdwf.f:747-849) through chs_derivs against the oracle,
  whose derivs is pinned to the reference build (tests/test_oracle.py);
* the multi-shift iteration cap (niterc, rhmc/cheisdwf.f:603,639) is reported
  as CHS_ENOCONV with the results written -- and those results equal the
  oracle's at the same cap -- and a non-finite residual halts the solve;
* device mirrors of read-only host spinors: a repeat call with unchanged
  inputs uploads no spinor, a touched array is uploaded again, and results do
  not depend on whether a mirror is used."""
import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import remez  # noqa: E402
from chiralheisenberg_amd._lib import NoConvergence  # noqa: E402

R4, R2, R2G = remez.load("remez4"), remez.load("remez2"), remez.load("remez2g")


@pytest.fixture(scope="module")
def lats():
    lat, olat = fm.Lattice(6, 4), orc.Lattice(6, 4)
    yield lat, olat
    lat.close()


@pytest.mark.parametrize("imass", [1, 3])
def test_derivs_vs_oracle(lats, imass):
    lat, olat = lats
    rs = np.random.RandomState(70 + imass)
    R, X2 = spinor(rs, lat.shape), spinor(rs, lat.shape)
    F0 = rs.randn(3, lat.kvol)
    G = orc.derivs(olat, F0.copy(), R, X2, 0.41, imass)
    F = fm.derivs(lat, R, X2, 0.41, imass, F0.copy())
    assert np.abs(F - G).max() <= 1e-13 * np.abs(G - F0).max()


def test_derivs_of_solution_matches_force_contribution(lats):
    """derivs(x_k, M x_k, a_k) summed over the remez2g shifts is the iflag = 1
    force of qmrherm (rhmc/cheisdwf.f:733-740)."""
    lat, olat = lats
    rs = np.random.RandomState(75)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R2G[0], R2G[1]
    F_ms = np.zeros((3, lat.kvol))
    fm.qmrherm(lat, phi, 1e-12, a, ad, chi, iflag=1, dSdchi=F_ms)
    _, _, xk = orc.qmrherm(olat, phi, 1e-12, a, ad, chi, 0.0, 1, 1, None, want_shifts=True)
    F = np.zeros((3, lat.kvol))
    for k in range(len(ad)):
        z = orc.dslash(olat, xk[k], chi, 0.0, 1, False)
        fm.derivs(lat, xk[k], z, a[k + 1], 1, F)
    assert np.abs(F - F_ms).max() <= 1e-9 * np.abs(F_ms).max()


@pytest.mark.parametrize("iflag", [0, 1])
def test_qmrherm_cap_reports_and_keeps_results(lats, iflag):
    lat, olat = lats
    rs = np.random.RandomState(80 + iflag)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a, ad = R4[0], R4[1]
    cap = 17
    lat.set_lanczos_cap(cap)
    orc.set_lanczos_cap(cap)
    try:
        F = np.zeros((3, lat.kvol))
        with pytest.raises(NoConvergence, match="cap of 17") as ei:
            fm.qmrherm(lat, phi, 1e-10, a, ad, chi, iflag=iflag, dSdchi=F if iflag else None)
        x, it = ei.value.result
        Fo = np.zeros((3, lat.kvol))
        xo, ito = orc.qmrherm(olat, phi, 1e-10, a, ad, chi, 0.0, 1, iflag, Fo)
        assert it == ito == cap
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
        if iflag:
            assert np.abs(F - Fo).max() <= 1e-11 * np.abs(Fo).max()
    finally:
        lat.set_lanczos_cap(0)
        orc.set_lanczos_cap(0)
    # the default cap is back: the same solve converges
    x, it = fm.qmrherm(lat, phi, 1e-10, a, ad, chi)
    assert it < 5000


def test_qmrherm_nonfinite_source_halts(lats):
    lat, _ = lats
    rs = np.random.RandomState(85)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    phi.reshape(-1)[123] = np.nan
    with pytest.raises(NoConvergence, match="non-finite"):
        fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
    # the context is usable afterwards
    phi.reshape(-1)[123] = 0.5
    _, it = fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
    assert 0 < it < 5000


def test_mirror_uploads_phi_once(lats):
    lat, _ = lats
    rs = np.random.RandomState(90)
    phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    pp = rs.randn(3, lat.kvol)
    V = 16 * lat.kferm
    x_plain, it_plain = fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
    F_plain, _, _ = fm.force(lat, phi, 1e-7, chi, 0.475, R2G)
    h_plain, _, _ = fm.hamilton(lat, phi, 1e-9, chi, pp, 0.475, R2)
    lat.mirror(phi)
    try:
        u0, _ = lat.transfer_bytes()
        x1, it1 = fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
        u1, _ = lat.transfer_bytes()
        assert u1 - u0 == V                      # first use uploads Phi
        x2, it2 = fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
        F, _, _ = fm.force(lat, phi, 1e-7, chi, 0.475, R2G)
        h, _, _ = fm.hamilton(lat, phi, 1e-9, chi, pp, 0.475, R2)
        u2, _ = lat.transfer_bytes()
        assert u2 == u1                          # unchanged Phi: no spinor moved to the device
        assert it1 == it2 == it_plain
        assert np.array_equal(x1, x_plain) and np.array_equal(x2, x_plain)
        assert np.array_equal(F, F_plain) and h == h_plain
        # the caller writes Phi and says so: the next call sees the new values
        phi *= 2.0
        lat.touch(phi)
        x3, _ = fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
        u3, _ = lat.transfer_bytes()
        assert u3 - u2 == V
        assert np.abs(x3 - 2.0 * x_plain).max() <= 1e-8 * np.abs(x_plain).max()
    finally:
        lat.release(phi)
    # released: every call uploads again
    u4, _ = lat.transfer_bytes()
    fm.qmrherm(lat, phi, 1e-9, R4[0], R4[1], chi)
    assert lat.transfer_bytes()[0] - u4 == V


def test_mirror_dslash_moves_only_the_result(lats):
    lat, olat = lats
    rs = np.random.RandomState(91)
    R, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    lat.mirror(R)
    try:
        out1 = fm.dslash(lat, R, chi, 0.0, 0.2, 1)
        u1, d1 = lat.transfer_bytes()
        out2 = fm.dslashd(lat, R, chi, 0.0, 0.2, 1)
        u2, d2 = lat.transfer_bytes()
        assert u2 == u1 and d2 - d1 == 16 * lat.kferm
    finally:
        lat.release(R)
    for out, dag in ((out1, False), (out2, True)):
        ref = orc.dslash(olat, R, chi, 0.2, 1, dag)
        assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_trajectory_carries_on_at_the_lanczos_cap():
    """With the multi-shift cap lowered below what the solves need, the
    trajectory runs to its end on the last iterates, as the reference's does
    (rhmc/cheisdwf.f:639: no message, no stop), and matches the oracle's
    trajectory under the same cap: the same steps and decision, dH within
    1e-6."""
    from chiralheisenberg_amd import rhmc
    L, Ls, cap, seed = 4, 4, 40, 693764523745.0
    lat, olat = fm.Lattice(L, Ls), orc.Lattice(L, Ls)
    par = dict(dt=0.1, beta=0.475, am=0.0, hpin=0.0, imass=1, iterl=3, itermax=1000, iprint=5, rescga=1e-9,
               rescgg=1e-6)
    sets = {"remez2": R2, "remez4": R4, "remez2g": R2G}
    chi0 = np.random.RandomState(9).randn(3, L ** 3) / np.sqrt(par["beta"])
    lat.set_lanczos_cap(cap)
    orc.set_lanczos_cap(cap)
    try:
        chi_g, chi_c = chi0.copy(), chi0.copy()
        g = rhmc.trajectory(lat, chi_g, rhmc.RandomStream(seed), 1, par, sets)
        c = orc.rhmc(olat, par, orc.Rng(seed), chi_c, 1, 1, sets)[0]
    finally:
        lat.set_lanczos_cap(0)
        orc.set_lanczos_cap(0)
        lat.close()
    print("capped at %d: heatbath %d iterations, steps %d / %d, dH %.8f / %.8f" % (
        cap, g["it_pf"], g["md_steps"], c["md_steps"], g["dH"], c["dH"]))
    assert g["it_pf"] == cap                      # the heatbath solve did stop at the cap
    assert g["md_steps"] == c["md_steps"] and g["accepted"] == c["accepted"]
    assert abs(g["dH"] - c["dH"]) <= 1e-6 * max(1.0, abs(c["dH"]))
