"""The Fortran drop-in binding (chiralheisenberg_amd/fortran/chs_shim.F90),
driven the way the reference's main program drives its own routines: by
name, with the reference's argument lists, through its COMMON blocks
(rhmc/cheisdwf.f:37-53).  Built for ksize = ksizet = kthird = 8 with
-fdefault-real-8; results are compared with the CPU oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

from chiralheisenberg_amd import build, remez  # noqa: E402

L = LT = LS = 8
KVOL = L * L * LT
KFERM = 8 * LS * KVOL
SHAPE = (2, 4, LT, L, L, LS)
D = ctypes.c_double


@pytest.fixture(scope="module")
def shim():
    path = build.shim_lib_path((L, LT, LS))
    if not os.path.exists(path):
        pytest.fail("Fortran shim library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)

    class Para(ctypes.Structure):
        _fields_ = [("beta", D), ("am3", D), ("ibound", ctypes.c_int)]

    s = type("S", (), {})()
    s.lib = lib
    s.para = Para.in_dll(lib, "para_")
    trial = (D * (6 * KVOL)).in_dll(lib, "trial_")
    s.chit = np.ctypeslib.as_array(trial)[:3 * KVOL].reshape(3, KVOL)
    s.pp = np.ctypeslib.as_array(trial)[3 * KVOL:].reshape(3, KVOL)
    s.F = np.ctypeslib.as_array((D * (3 * KVOL)).in_dll(lib, "sforce_")).reshape(3, KVOL)
    s.x = np.ctypeslib.as_array((D * (2 * KFERM)).in_dll(lib, "vector_")).view(np.complex128).reshape(SHAPE)
    s.r2 = np.ctypeslib.as_array((D * (4 * 25 + 2)).in_dll(lib, "remez2_"))
    s.r2g = np.ctypeslib.as_array((D * (4 * 12 + 2)).in_dll(lib, "remez2g_"))
    s.param = np.ctypeslib.as_array((D * 4).in_dll(lib, "param_"))
    s.para.beta, s.para.am3, s.para.ibound = 0.475, 1.0, -1
    a, ad, b, bd = remez.load("remez2")
    s.r2[:] = np.concatenate([a, ad, b, bd])
    a, ad, b, bd = remez.load("remez2g")
    s.r2g[:] = np.concatenate([a, ad, b, bd])
    return s


def byref(v, t=D):
    return ctypes.byref(t(v))


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


OLAT = orc.Lattice(L, LS, LT)


def test_shim_dslash_and_dslashd(shim):
    rs = np.random.RandomState(1)
    R, chi = spinor(rs, SHAPE), phi_field(rs, KVOL)
    for name, dag in (("dslash_", False), ("dslashd_", True)):
        for imass in (1, 3):
            out = np.empty(SHAPE, np.complex128)
            getattr(shim.lib, name)(P(out), P(R), P(chi), byref(0.01), byref(0.4), byref(imass, ctypes.c_int))
            ref = orc.dslash(OLAT, R, chi, 0.4, imass, dag)
            assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_shim_qmrherm_uses_common_blocks(shim):
    rs = np.random.RandomState(2)
    phi, chi = spinor(rs, SHAPE), phi_field(rs, KVOL)
    a, ad, _, _ = remez.load("remez4")
    shim.chit[:] = chi
    it = ctypes.c_int(0)
    shim.lib.qmrherm_(P(phi), byref(1e-9), ctypes.byref(it), byref(0.0), byref(0.0), byref(1, ctypes.c_int),
                      P(a), P(ad), byref(25, ctypes.c_int), byref(0, ctypes.c_int))
    xc, ic = orc.qmrherm(OLAT, phi, 1e-9, a, ad, chi)
    assert abs(it.value - ic) <= 5
    assert np.abs(shim.x - xc).max() <= 1e-9 * np.abs(xc).max()


def test_shim_force_hamilton_congrad(shim):
    rs = np.random.RandomState(3)
    phi, chi, pp = spinor(rs, SHAPE), phi_field(rs, KVOL), rs.randn(3, KVOL)
    shim.chit[:] = chi
    shim.pp[:] = pp
    shim.param[:] = 0
    shim.lib.force_(P(phi), byref(1e-12), byref(0.0), byref(0.0), byref(1, ctypes.c_int))
    Fc, _, itc = orc.force(OLAT, phi, 1e-12, chi, 0.475, remez.load("remez2g"))
    assert np.abs(shim.F - Fc).max() <= 1e-10 * np.abs(Fc).max()
    assert shim.param[0] > 0 and shim.param[2] > 0   # ancg, ancgf counters as the reference keeps them
    h = [D() for _ in range(4)]
    shim.lib.hamilton_(P(phi), *[ctypes.byref(v) for v in h], byref(1e-11), byref(1, ctypes.c_int), byref(0.0),
                       byref(0.0), byref(1, ctypes.c_int))
    hc, _, _ = orc.hamilton(OLAT, phi, 1e-11, chi, pp, 0.475, remez.load("remez2"))
    for a, b in zip([v.value for v in h], hc):
        assert abs(a - b) <= 1e-10 * abs(b)
    x0 = 0.1 * spinor(rs, SHAPE)
    shim.x[:] = x0
    it = ctypes.c_int(0)
    shim.lib.congrad_(P(phi), byref(1e-9), ctypes.byref(it), byref(0.0), byref(0.0), byref(1, ctypes.c_int))
    xc, ic = orc.congrad(OLAT, phi, 1e-9, chi, x0)
    assert abs(it.value - ic) <= 3
    assert np.abs(shim.x - xc).max() <= 1e-6 * np.abs(xc).max()


@pytest.mark.parametrize("imass", [1, 3])
def test_shim_derivs_accumulates_into_sforce(shim, imass):
    rs = np.random.RandomState(4 + imass)
    R, X2, chi = spinor(rs, SHAPE), spinor(rs, SHAPE), phi_field(rs, KVOL)
    out = np.empty(SHAPE, np.complex128)
    # derivs has no lattice arguments: the context exists once any other routine ran
    shim.lib.dslash_(P(out), P(R), P(chi), byref(0.0), byref(0.0), byref(1, ctypes.c_int))
    F0 = rs.randn(3, KVOL)
    shim.F[:] = F0
    shim.lib.derivs_(P(R), P(X2), byref(0.29), byref(imass, ctypes.c_int))
    G = orc.derivs(OLAT, F0.copy(), R, X2, 0.29, imass)
    assert np.abs(shim.F - G).max() <= 1e-13 * np.abs(G - F0).max()


def test_shim_options_reduced_force(shim):
    """chs_options (no counterpart in the reference) switches the shim's
    context to FP32 force solves with shift freezing: the force stays within
    the FP32 form's error of the oracle's (field-dependent: for this field
    1.6e-4 without freezing, 2.0e-4 with it; bound 5e-4); chs_options(0, 0)
    restores the reference's solves (the force to 1e-8 again)."""
    rs = np.random.RandomState(6)
    phi, chi = spinor(rs, SHAPE), phi_field(rs, KVOL)
    shim.chit[:] = chi
    Fc, _, _ = orc.force(OLAT, phi, 1e-9, chi, 0.475, remez.load("remez2g"))
    shim.lib.chs_options_(byref(1, ctypes.c_int), byref(0.01))
    shim.lib.force_(P(phi), byref(1e-9), byref(0.0), byref(0.0), byref(1, ctypes.c_int))
    fr = Fc - 2 * 0.475 * chi
    e = np.linalg.norm(shim.F - Fc) / np.linalg.norm(fr)
    print("shim force with FP32 solves and freezing: error %.2e of the fermion part (largest element %.2e)" % (
        e, np.abs(shim.F - Fc).max() / np.abs(fr).max()))
    assert 0 < e <= 5e-4
    shim.lib.chs_options_(byref(0, ctypes.c_int), byref(0.0))
    shim.lib.force_(P(phi), byref(1e-9), byref(0.0), byref(0.0), byref(1, ctypes.c_int))
    assert np.abs(shim.F - Fc).max() <= 1e-8 * np.abs(fr).max()
