"""ctypes access to the reference's HMC sibling mh/cheisdwf_hmc.f, compiled
unmodified into oracle/_ref/libref_mh.so by `make -C oracle ref` (FP64 kinds:
-fdefault-real-8 -fdefault-double-8).

TEST INFRASTRUCTURE ONLY: tests/ use it (through committed golden vectors) to
pin the oracle's HMC force and Hamiltonian.  Sizes are compile-time
(ksize = ksizet = 12, kthird = 8); the implicit inputs live in COMMON blocks
(mh/cheisdwf_hmc.f:37-47, 399-405, 565-571).
"""
import ctypes
import os

import numpy as np

from refbind import KFERM, KVOL, L, LS, LT, _run_big_stack

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libref_mh.so")


def available():
    return os.path.exists(PATH)


class RefMH:
    def __init__(self):
        self.lib = ctypes.CDLL(PATH)
        D = ctypes.c_double

        class Para(ctypes.Structure):
            _fields_ = [("beta", D), ("am3", D), ("ibound", ctypes.c_int)]

        self.para = Para.in_dll(self.lib, "para_")
        tr = (ctypes.c_double * (6 * KVOL)).in_dll(self.lib, "trial_")
        arr = np.ctypeslib.as_array(tr)
        self.chit = arr[:3 * KVOL].reshape(3, KVOL)
        self.pp = arr[3 * KVOL:].reshape(3, KVOL)
        self.x1 = np.ctypeslib.as_array((ctypes.c_double * (2 * KFERM)).in_dll(self.lib, "vector_")).view(
            np.complex128).reshape(2, 4, LT, L, L, LS)
        self.lib.init_(ctypes.byref(ctypes.c_int(-1)))   # gamma tables only (mh/cheisdwf_hmc.f:1030-1140)

    def set_params(self, beta=0.45, am3=1.0, ibound=-1):
        self.para.beta, self.para.am3, self.para.ibound = beta, am3, ibound

    @staticmethod
    def _d(v):
        return ctypes.byref(ctypes.c_double(v))

    def dslash(self, r, chi, imass=1, dagger=False, am=0.0):
        r = np.ascontiguousarray(r, dtype=np.complex128)
        chi = np.ascontiguousarray(chi, dtype=np.float64)
        out = np.empty_like(r)
        fn = self.lib.dslashd_ if dagger else self.lib.dslash_
        fn(out.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
           chi.ctypes.data_as(ctypes.c_void_p), self._d(am), ctypes.byref(ctypes.c_int(imass)))
        return out

    def force(self, phi, x0, res, chi, imass=1, am=0.0):
        """force(Phi,X0,dSdchi,iflag,res1,am,imass): returns (dSdchi, X0 after)."""
        self.chit[:] = chi
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        x0 = np.array(x0, dtype=np.complex128, copy=True)
        F = np.zeros((3, KVOL))
        _run_big_stack(self.lib.force_, phi.ctypes.data_as(ctypes.c_void_p), x0.ctypes.data_as(ctypes.c_void_p),
                       F.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_int(0)), self._d(res), self._d(am),
                       ctypes.byref(ctypes.c_int(imass)))
        return F, x0

    def hamilton(self, phi, x0, res, chi, pp, imass=1, am=0.0, isweep=1):
        """hamilton(Phi,X0,h,hg,hp,s,res2,isweep,am,imass): returns
        ((h, hg, hp, s), X0 after)."""
        self.chit[:] = chi
        self.pp[:] = pp
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        x0 = np.array(x0, dtype=np.complex128, copy=True)
        h, hg, hp, s = (ctypes.c_double() for _ in range(4))
        _run_big_stack(self.lib.hamilton_, phi.ctypes.data_as(ctypes.c_void_p), x0.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(h), ctypes.byref(hg), ctypes.byref(hp), ctypes.byref(s), self._d(res),
                       ctypes.byref(ctypes.c_int(isweep)), self._d(am), ctypes.byref(ctypes.c_int(imass)))
        return (h.value, hg.value, hp.value, s.value), x0

    def ranset(self, seed):
        self.lib.ranset_(ctypes.byref(ctypes.c_double(seed)))

    def rano(self, y, idum):
        self.lib.rano_.restype = ctypes.c_double
        yy, ii = ctypes.c_double(y), ctypes.c_int(idum)
        v = self.lib.rano_(ctypes.byref(yy), ctypes.byref(ii))
        return v, yy.value, ii.value
