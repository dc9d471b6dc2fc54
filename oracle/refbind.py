"""ctypes access to the reference's own rhmc/cheisdwf.f, compiled unmodified
into oracle/_ref/libref_{fp64,shipped}.so by `make -C oracle ref`.

TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle.  The reference
sizes are compile-time: ksize = ksizet = 12, kthird = 8
(rhmc/cheisdwf.f:31).  Arguments are passed by reference, Fortran style, and
the implicit inputs live in COMMON blocks (rhmc/cheisdwf.f:37-53), which are
plain global symbols of the shared library.

Flavours:
  fp64    -- built with -fdefault-real-8 -fdefault-double-8: every default
             REAL is 8 bytes (the parity target).
  shipped -- default kinds: dSdchi, pp, beta, am3, hpin, res are REAL*4 and chi
             is rounded to FP32 inside dslash (the shipped semantics).
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
L, LT, LS = 12, 12, 8
KVOL = L * L * LT
KFERM = 8 * LS * KVOL


def path(flavour):
    return os.path.join(_HERE, "_ref", "libref_%s.so" % flavour)


def available(flavour="fp64"):
    return os.path.exists(path(flavour))


def _run_big_stack(fn, *a):
    """The reference keeps its solver work arrays on the stack (tens of MB)."""
    out = {}

    def body():
        try:
            out["v"] = fn(*a)
        except BaseException as e:  # pragma: no cover - re-raised below
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=body)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("v")


class Ref:
    def __init__(self, flavour="fp64"):
        self.flavour = flavour
        self.lib = ctypes.CDLL(path(flavour))
        self.real = ctypes.c_double if flavour == "fp64" else ctypes.c_float
        self.npreal = np.float64 if flavour == "fp64" else np.float32
        R = self.real

        class Para(ctypes.Structure):
            _fields_ = [("beta", R), ("am3", R), ("ibound", ctypes.c_int)]

        self.para = Para.in_dll(self.lib, "para_")
        # /trial/ chit(kvol,3) real*8, pp(kvol,3) real
        tr = ctypes.c_char * (KVOL * 3 * 8 + KVOL * 3 * ctypes.sizeof(R))
        self._trial = tr.in_dll(self.lib, "trial_")
        base = ctypes.addressof(self._trial)
        self.chit = np.ctypeslib.as_array((ctypes.c_double * (3 * KVOL)).from_address(base)).reshape(3, KVOL)
        self.pp = np.ctypeslib.as_array((R * (3 * KVOL)).from_address(base + 3 * KVOL * 8)).reshape(3, KVOL)
        self.dsdchi = np.ctypeslib.as_array((R * (3 * KVOL)).in_dll(self.lib, "sforce_")).reshape(3, KVOL)
        self.x1 = np.ctypeslib.as_array((ctypes.c_double * (2 * KFERM)).in_dll(self.lib, "vector_")).view(
            np.complex128).reshape(2, 4, LT, L, L, LS)
        # remez common blocks: anum(0:n), aden(n), bnum(0:n), bden(n), real*8
        self.remez = {}
        for nm, n in (("remez2_", 25), ("remez4_", 25), ("remez2g_", 12)):
            arr = np.ctypeslib.as_array((ctypes.c_double * (4 * n + 2)).in_dll(self.lib, nm))
            self.remez[nm] = arr
        self.tpi = R.in_dll(self.lib, "trans_")

    def set_params(self, beta=0.45, am3=1.0, ibound=-1):
        self.para.beta = beta
        self.para.am3 = am3
        self.para.ibound = ibound

    def set_remez(self, name, anum, aden, bnum, bden):
        n = len(aden)
        arr = self.remez[name + "_"]
        arr[0:n + 1] = anum
        arr[n + 1:2 * n + 1] = aden
        arr[2 * n + 1:3 * n + 2] = bnum
        arr[3 * n + 2:4 * n + 2] = bden

    def _r(self, v):
        return ctypes.byref(self.real(v))

    def init_tables(self):
        """init(nc) with nc < 0 (rhmc/cheisdwf.f:1491-1597): neighbour tables
        and the gamma tables /dirac/ that dslash_v1 reads; chi untouched."""
        self.lib.init_(ctypes.byref(ctypes.c_int(-1)))

    def dslash(self, r, chi, hpin=0.0, imass=1, dagger=False, am=0.0, v1=False):
        """dslash / dslashd (rhmc/cheisdwf.f:2634-3501), or with v1=True the
        legacy dslash_v1 / dslashd_v1 (:1804-2632; call init_tables first)."""
        r = np.ascontiguousarray(r, dtype=np.complex128)
        chi = np.ascontiguousarray(chi, dtype=np.float64)
        out = np.empty_like(r)
        if v1:
            fn = self.lib.dslashd_v1_ if dagger else self.lib.dslash_v1_
        else:
            fn = self.lib.dslashd_ if dagger else self.lib.dslash_
        fn(out.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
           chi.ctypes.data_as(ctypes.c_void_p), self._r(am), self._r(hpin), ctypes.byref(ctypes.c_int(imass)))
        return out

    def qmrherm(self, phi, res, anum, aden, chi, hpin=0.0, imass=1, iflag=0, dsdchi=None, am=0.0):
        """Sets /trial/ chit = chi; returns (x, itercg); dsdchi accumulates for iflag=1."""
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        self.chit[:] = chi
        anum = np.ascontiguousarray(anum, dtype=np.float64)
        aden = np.ascontiguousarray(aden, dtype=np.float64)
        if dsdchi is not None:
            self.dsdchi[:] = dsdchi
        it = ctypes.c_int(0)
        args = (phi.ctypes.data_as(ctypes.c_void_p), self._r(res), ctypes.byref(it), self._r(am), self._r(hpin),
                ctypes.byref(ctypes.c_int(imass)), anum.ctypes.data_as(ctypes.c_void_p),
                aden.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_int(len(aden))),
                ctypes.byref(ctypes.c_int(iflag)))
        _run_big_stack(self.lib.qmrherm_, *args)
        if dsdchi is not None:
            dsdchi[:] = self.dsdchi
        return self.x1.copy(), it.value

    def derivs(self, r, x2, anum, imass, dsdchi):
        """derivs(R, X2, anum, imass) (rhmc/cheisdwf.f:747-849): /sforce/
        dSdchi, set from dsdchi first, accumulates the wall bilinears of R
        and X2.  Returns the updated dSdchi as FP64."""
        r = np.ascontiguousarray(r, dtype=np.complex128)
        x2 = np.ascontiguousarray(x2, dtype=np.complex128)
        self.dsdchi[:] = dsdchi
        self.lib.derivs_(r.ctypes.data_as(ctypes.c_void_p), x2.ctypes.data_as(ctypes.c_void_p),
                         ctypes.byref(ctypes.c_double(anum)), ctypes.byref(ctypes.c_int(imass)))
        return np.array(self.dsdchi, dtype=np.float64)

    def congrad(self, phi, res, chi, x0, hpin=0.0, imass=1, am=0.0):
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        self.chit[:] = chi
        self.x1[:] = x0
        it = ctypes.c_int(0)
        args = (phi.ctypes.data_as(ctypes.c_void_p), self._r(res), ctypes.byref(it), self._r(am), self._r(hpin),
                ctypes.byref(ctypes.c_int(imass)))
        _run_big_stack(self.lib.congrad_, *args)
        return self.x1.copy(), it.value

    def force(self, phi, res1, chi, hpin=0.0, imass=1, am=0.0):
        """Uses /remez2g/ as set by set_remez; returns (dSdchi, x)."""
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        self.chit[:] = chi
        args = (phi.ctypes.data_as(ctypes.c_void_p), self._r(res1), self._r(am), self._r(hpin),
                ctypes.byref(ctypes.c_int(imass)))
        _run_big_stack(self.lib.force_, *args)
        return self.dsdchi.astype(np.float64), self.x1.copy()

    def hamilton(self, phi, res2, chi, pp, hpin=0.0, imass=1, am=0.0, isweep=1):
        """Uses /remez2/; returns (h, hg, hp, s)."""
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        self.chit[:] = chi
        self.pp[:] = pp
        R8 = ctypes.c_double
        h, hg, hp, s = R8(), R8(), R8(), R8()
        args = (phi.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h), ctypes.byref(hg), ctypes.byref(hp),
                ctypes.byref(s), self._r(res2), ctypes.byref(ctypes.c_int(isweep)), self._r(am), self._r(hpin),
                ctypes.byref(ctypes.c_int(imass)))
        _run_big_stack(self.lib.hamilton_, *args)
        return h.value, hg.value, hp.value, s.value

    def set_dirac(self):
        """COMMON /dirac/ gamval(6,4), gamin(6,4) as init sets them
        (rhmc/cheisdwf.f:1519-1595, values times akappa = 0.5)."""
        one, zi = 1.0 + 0j, 1j
        val = {1: [-zi, -zi, zi, zi], 2: [-one, one, one, -one], 3: [-zi, zi, zi, -zi], 4: [one, one, -one, -one],
               5: [-one, -one, -one, -one], 6: [-one, -one, one, one]}
        col = {1: [4, 3, 2, 1], 2: [4, 3, 2, 1], 3: [3, 4, 1, 2], 4: [1, 2, 3, 4], 5: [3, 4, 1, 2], 6: [3, 4, 1, 2]}
        blk = (ctypes.c_char * (6 * 4 * 16 + 6 * 4 * 4)).in_dll(self.lib, "dirac_")
        base = ctypes.addressof(blk)
        gv = np.ctypeslib.as_array((ctypes.c_double * 48).from_address(base)).view(np.complex128)
        gi = np.ctypeslib.as_array((ctypes.c_int * 24).from_address(base + 384))
        for mu in range(1, 7):
            for d in range(1, 5):
                gv[(mu - 1) + 6 * (d - 1)] = 0.5 * val[mu][d - 1]
                gi[(mu - 1) + 6 * (d - 1)] = col[mu][d - 1]

    def fermion(self, chi, res, workdir, hpin=0.0, imass=1, am=0.0):
        """The reference's own fermion(res,aviter,am,hpin,imass), run in
        workdir (it writes fort.5001/5011/5002/5012 there).  /trial/ chi,
        /para/ and /ranseed/ must be set.  Returns (cferm1, cferm2) complex
        (2 flavours, Lt) and aviter."""
        self.chit[:] = np.asarray(chi, dtype=np.float64)
        self.set_dirac()
        av = self.real(0)
        old = os.getcwd()
        os.chdir(workdir)
        try:
            _run_big_stack(self.lib.fermion_, self._r(res), ctypes.byref(av), self._r(am), self._r(hpin),
                           ctypes.byref(ctypes.c_int(imass)))
            out = []
            for unit in (5001, 5011, 5002, 5012):
                rows = np.loadtxt(os.path.join(workdir, "fort.%d" % unit))
                out.append(rows[:, 1] + 1j * rows[:, 2])
        finally:
            os.chdir(old)
        c1 = np.stack([out[0], out[2]])
        c2 = np.stack([out[1], out[3]])
        return c1, c2, av.value

    # random numbers --------------------------------------------------------
    def rranset(self, seed):
        self.lib.rranset_(ctypes.byref(ctypes.c_double(seed)))

    def rranget(self):
        s = ctypes.c_double()
        self.lib.rranget_(ctypes.byref(s))
        return s.value

    def rranf(self):
        self.lib.rranf_.restype = self.real
        return self.lib.rranf_()

    def rano(self, y, idum):
        """Returns (value, new y, new idum)."""
        self.lib.rano_.restype = self.real
        yy = self.real(y)
        ii = ctypes.c_int(idum)
        v = self.lib.rano_(ctypes.byref(yy), ctypes.byref(ii))
        return v, yy.value, ii.value

    def set_ranseed(self, yran, idum):
        class RS(ctypes.Structure):
            _fields_ = [("yran", self.real), ("idum", ctypes.c_int)]
        rs = RS.in_dll(self.lib, "ranseed_")
        rs.yran = yran
        rs.idum = idum

    def set_tpi(self):
        self.tpi.value = 2.0 * np.arccos(-1.0)

    def gauss0(self):
        ps = np.zeros((2, KVOL), dtype=self.npreal)
        self.lib.gauss0_(ps.ctypes.data_as(ctypes.c_void_p))
        return ps

    def gaussp(self):
        ps = np.zeros((2, KVOL), dtype=self.npreal)
        self.lib.gaussp_(ps.ctypes.data_as(ctypes.c_void_p))
        return ps
