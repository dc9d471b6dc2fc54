"""GPU parity of dslash/dslashd (rhmc/cheisdwf.f:2634-3501) against the CPU
oracle: <= 1e-13 relative (max-norm) in FP64; FP32 and BF16 storage within
their rounding (stated per test)."""
import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")

_lats = {}


def lattices(L, Ls, Lt):
    key = (L, Ls, Lt)
    if key not in _lats:
        _lats[key] = (fm.Lattice(L, Ls, Lt), orc.Lattice(L, Ls, Lt))
    return _lats[key]


SIZES = [(2, 2, 2), (4, 4, 4), (8, 8, 8), (6, 5, 4), (12, 8, 12), (12, 24, 12)]


@pytest.mark.parametrize("L,Ls,Lt", SIZES)
@pytest.mark.parametrize("imass", [1, 3, 2])
def test_dslash_fp64_parity(L, Ls, Lt, imass):
    lat, olat = lattices(L, Ls, Lt)
    rs = np.random.RandomState(1000 * L + 10 * Ls + imass)
    R = spinor(rs, lat.shape)
    chi = phi_field(rs, lat.kvol)
    for dag in (False, True):
        for hpin in (0.0, 0.5):
            g = (fm.dslashd if dag else fm.dslash)(lat, R, chi, 0.01, hpin, imass)
            c = orc.dslash(olat, R, chi, hpin, imass, dag)
            err = np.abs(g - c).max() / np.abs(c).max()
            assert err <= 1e-13, (L, Ls, Lt, imass, dag, hpin, err)


@pytest.mark.parametrize("ibound", [1, 0])
def test_dslash_periodic_t(ibound):
    """Any ibound other than -1 makes the t direction periodic, as in the
    reference's routines (rhmc/cheisdwf.f:2831, the dagger at :3265): FP64
    to 1e-13 and the paired FP32 stencil to its rounding, against the oracle
    with the same boundary."""
    lat = fm.Lattice(6, 4, 8, ibound=ibound)
    olat = orc.Lattice(6, 4, 8, ibound=ibound)
    rs = np.random.RandomState(9 + ibound)
    R, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    chi_d = fm.DeviceArray.from_numpy(lat, chi)
    for imass in (1, 3):
        for dag in (False, True):
            c = orc.dslash(olat, R, chi, 0.4, imass, dag)
            g = (fm.dslashd if dag else fm.dslash)(lat, R, chi, 0.0, 0.4, imass)
            assert np.abs(g - c).max() <= 1e-13 * np.abs(c).max(), (ibound, imass, dag)
            rin = fm.spinor_dev(lat, R, "fp32")
            out = fm.spinor_dev(lat, None, "fp32")
            fm.dslash_dev(lat, out, rin, chi_d, 0.4, imass, dag)
            g = fm.to_host(lat, out)
            assert np.linalg.norm(g - c) <= 2e-6 * np.linalg.norm(c), (ibound, imass, dag)
    # the boundary term is there: the antiperiodic operator differs
    c = orc.dslash(orc.Lattice(6, 4, 8), R, chi, 0.4, 1, False)
    assert np.abs(fm.dslash(lat, R, chi, 0.0, 0.4, 1) - c).max() > 0.1
    lat.close()


def test_dslash_adjoint_gpu():
    lat, _ = lattices(8, 8, 8)
    rs = np.random.RandomState(5)
    X, Y = spinor(rs, lat.shape), spinor(rs, lat.shape)
    chi = phi_field(rs, lat.kvol)
    for imass in (1, 3):
        lhs = np.vdot(Y, fm.dslash(lat, X, chi, 0.0, 0.7, imass))
        rhs = np.vdot(fm.dslashd(lat, Y, chi, 0.0, 0.7, imass), X)
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_am_has_no_effect():
    lat, _ = lattices(4, 4, 4)
    rs = np.random.RandomState(6)
    R, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    a = fm.dslash(lat, R, chi, 0.0, 0.0, 1)
    b = fm.dslash(lat, R, chi, 0.37, 0.0, 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16", 2e-2), ("fp16", 2e-3)])
@pytest.mark.parametrize("dagger", [False, True])
def test_dslash_low_precision(prec, tol, dagger):
    """FP32 storage/compute: within a few FP32 ulps of the operator norm;
    BF16 storage (8-bit mantissa) with FP32 accumulation: within ~2^-7
    relative in the 2-norm; FP16 storage (11-bit significand): within
    ~2^-10."""
    lat, olat = lattices(12, 24, 12)
    rs = np.random.RandomState(7)
    R, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    chi_d = fm.DeviceArray.from_numpy(lat, chi)
    rin = fm.spinor_dev(lat, R, prec)
    out = fm.spinor_dev(lat, None, prec)
    fm.dslash_dev(lat, out, rin, chi_d, 0.0, 1, dagger)
    g = fm.to_host(lat, out)
    c = orc.dslash(olat, R, chi, 0.0, 1, dagger)
    err = np.linalg.norm(g - c) / np.linalg.norm(c)
    assert err <= tol, (prec, err)


def test_dslash_rejects_alias():
    lat, _ = lattices(4, 4, 4)
    d = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    chi = fm.DeviceArray.from_numpy(lat, np.zeros((3, lat.kvol)))
    with pytest.raises(fm._lib.ChsError):
        fm.dslash_dev(lat, d, d, chi)


def test_dslash_vs_reference_golden_gpu():
    """GPU dslash/dslashd against the reference's own outputs
    (tests/golden/ref_l12_ls8.npz, rhmc/cheisdwf.f compiled unmodified)."""
    from test_oracle import GOLD, gold_inputs
    lat, _ = lattices(12, 8, 12)
    R, chi, *_ = gold_inputs()
    idx = GOLD["idx"]
    for imass in (1, 3):
        for dag in (0, 1):
            out = (fm.dslashd if dag else fm.dslash)(lat, R, chi, 0.0, 0.5, imass).reshape(-1)
            s = GOLD["dslash_m%d_d%d_s" % (imass, dag)]
            assert np.abs(out[idx] - s).max() <= 1e-14 * np.abs(s).max()
            n = float(GOLD["dslash_m%d_d%d_n" % (imass, dag)])
            assert abs(np.linalg.norm(out) - n) <= 1e-14 * n


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp16"])
def test_native_layout_round_trip(prec):
    """reference -> native -> reference reorders exactly (bit-exact in FP64;
    in FP32/BF16 equal to the rounded input)."""
    lat, _ = lattices(6, 5, 4)
    rs = np.random.RandomState(8)
    R = spinor(rs, lat.shape)
    back = fm.to_host(lat, fm.spinor_dev(lat, R, prec))
    if prec == "fp64":
        assert np.array_equal(back, R)
    elif prec == "fp32":
        assert np.array_equal(back, R.astype(np.complex64).astype(np.complex128))
    elif prec == "fp16":
        # round to nearest even, either directly from FP64 (what the compiler
        # makes of the FP64 -> FP32 -> binary16 conversion) or via FP32
        for part in (np.real, np.imag):
            v, got = part(R), part(back)
            one = v.astype(np.float16).astype(np.float64)
            two = v.astype(np.float32).astype(np.float16).astype(np.float64)
            assert np.all((got == one) | (got == two))
    else:
        assert np.abs(back - R).max() <= 2 ** -8 * np.abs(R).max()


@pytest.mark.parametrize("prec", ["fp32"])
@pytest.mark.parametrize("L,Ls,Lt", [(2, 2, 2), (4, 3, 6), (8, 8, 8), (6, 5, 3), (12, 24, 12)])
def test_pair_stencil_equals_one_site(prec, L, Ls, Lt, monkeypatch):
    """The paired-site FP32 stencil (two x-neighbours per thread) gives
    bit for bit the one-site stencil's output, with the wall terms (both
    imass, hpin at site 0) and the antiperiodic t boundary."""
    monkeypatch.setenv("CHS_PAIRS", "0")
    one = fm.Lattice(L, Ls, Lt)
    monkeypatch.setenv("CHS_PAIRS", "1")
    two = fm.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(31 * L + Ls)
    R, chi = spinor(rs, one.shape), phi_field(rs, one.kvol)
    for imass in (1, 3):
        for dagger in (False, True):
            outs = []
            for lat in (one, two):
                chi_d = fm.DeviceArray.from_numpy(lat, chi)
                rin = fm.spinor_dev(lat, R, prec)
                out = fm.spinor_dev(lat, None, prec)
                fm.dslash_dev(lat, out, rin, chi_d, 0.7, imass, dagger)
                outs.append(fm.to_host(lat, out))
            assert np.array_equal(outs[0], outs[1]), (prec, imass, dagger, np.abs(outs[0] - outs[1]).max())


@pytest.mark.parametrize("L,Ls,Lt", [(2, 2, 2), (5, 3, 4), (8, 8, 8), (24, 4, 6), (32, 4, 4)])
def test_plane_stencil_matches_one_site(L, Ls, Lt, monkeypatch):
    """The plain FP64 operator stages xy-planes in LDS (L^2 <= 1024); with
    CHS_PLANES=0 it runs the one-site stencil.  The plane kernel sums the two
    t hops apart and adds them after the x and y hops: outputs agree to
    rounding (1e-15 relative), with the wall terms (both imass, hpin at site 0)
    and the antiperiodic t boundary, up to L = 32 (1024-thread workgroups,
    64 KB of LDS)."""
    monkeypatch.setenv("CHS_PLANES", "0")
    one = fm.Lattice(L, Ls, Lt)
    monkeypatch.setenv("CHS_PLANES", "1")
    plane = fm.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(17 * L + Ls)
    R, chi = spinor(rs, one.shape), phi_field(rs, one.kvol)
    for imass in (1, 3):
        for dag in (False, True):
            f = fm.dslashd if dag else fm.dslash
            a, b = f(one, R, chi, 0.0, 0.7, imass), f(plane, R, chi, 0.0, 0.7, imass)
            assert np.abs(a - b).max() <= 1e-15 * np.abs(a).max(), (L, Ls, Lt, imass, dag)
    one.close()
    plane.close()


def test_odd_L_low_precision_uses_one_site_stencil():
    lat, olat = lattices(5, 4, 6)
    rs = np.random.RandomState(8)
    R, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    chi_d = fm.DeviceArray.from_numpy(lat, chi)
    rin = fm.spinor_dev(lat, R, "fp32")
    out = fm.spinor_dev(lat, None, "fp32")
    fm.dslash_dev(lat, out, rin, chi_d, 0.0, 1, False)
    c = orc.dslash(olat, R, chi, 0.0, 1, False)
    assert np.linalg.norm(fm.to_host(lat, out) - c) <= 2e-6 * np.linalg.norm(c)
