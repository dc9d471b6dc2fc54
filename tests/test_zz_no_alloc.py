"""chs_congrad_noalloc_dev -- CGNR with its work buffers in the library's
static device scratch, for callers that cannot allocate -- against
chs_congrad_dev on the same inputs, through the C-ABI: the same iteration
count and the same solution bit for bit, on two calls in a row (the second on
a smaller lattice, so the scratch holds the first call's larger contents),
and the oracle's CGNR (rhmc/cheisdwf.f:1064-1157) to the solver tolerance."""
import ctypes

import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd._lib import CHS_EINVAL, CHS_OK, lib  # noqa: E402


def _solve(lat, fn, phi_d, chi_d, x0, res):
    x_d = fm.spinor_dev(lat, x0)
    it = ctypes.c_int(0)
    rc = fn(lat.h, phi_d.ptr, res, ctypes.byref(it), 0.25, 1, chi_d.ptr, x_d.ptr)
    lat.synchronize()
    return rc, it.value, fm.to_host(lat, x_d)


def test_noalloc_congrad_equals_congrad_dev():
    for L, Ls, seed in ((8, 8, 1), (6, 4, 2)):          # second call smaller than the first
        print("no-alloc congrad L=%d Ls=%d" % (L, Ls), flush=True)
        lat = fm.Lattice(L, Ls)
        rs = np.random.RandomState(seed)
        phi, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
        x0 = 0.1 * spinor(rs, lat.shape)
        phi_d, chi_d = fm.spinor_dev(lat, phi), fm.DeviceArray.from_numpy(lat, chi)
        rc_a, it_a, xa = _solve(lat, lib().chs_congrad_dev, phi_d, chi_d, x0, 1e-9)
        rc_b, it_b, xb = _solve(lat, lib().chs_congrad_noalloc_dev, phi_d, chi_d, x0, 1e-9)
        assert rc_a == rc_b == CHS_OK
        assert it_a == it_b > 0
        assert np.array_equal(xa, xb)
        xo, ito = orc.congrad(orc.Lattice(L, Ls), phi, 1e-9, chi, x0, 0.25, 1)
        assert abs(it_b - ito) <= 3
        assert np.abs(xb - xo).max() <= 1e-6 * np.abs(xo).max()
        lat.close()


def test_noalloc_congrad_rejects_a_lattice_beyond_its_scratch():
    lat = fm.Lattice(24, 32)                              # kferm = 3,538,944 > 2^20
    it = ctypes.c_int(0)
    rc = lib().chs_congrad_noalloc_dev(lat.h, None, 1e-9, ctypes.byref(it), 0.0, 1, None, None)
    assert rc == CHS_EINVAL
    assert b"exceeds the static scratch" in lib().chs_last_error(lat.h)
    lat.close()
