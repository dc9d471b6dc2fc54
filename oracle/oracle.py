"""ctypes bindings of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product path (chiralheisenberg_amd)
never imports it.

Arrays are numpy complex128 in the reference layout: shape (2, 4, Lt, L, L, Ls)
in C order, which is Phi(Ls, L, L, Lt, 4, 2) of rhmc/cheisdwf.f:2644
(s fastest, then x, y, t, Dirac, flavour).  chi is float64 shape (3, kvol).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcLat(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("Lt", ctypes.c_int), ("Ls", ctypes.c_int),
                ("ibound", ctypes.c_int), ("am3", ctypes.c_double)]


class OrcRng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64), ("v", ctypes.c_double * 97),
                ("idum", ctypes.c_int), ("yran", ctypes.c_double),
                ("yrun", ctypes.c_double)]


class OrcRhmcPar(ctypes.Structure):
    _P = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("dt", ctypes.c_double), ("beta", ctypes.c_double), ("hpin", ctypes.c_double),
                ("imass", ctypes.c_int), ("iterl", ctypes.c_int), ("itermax", ctypes.c_int),
                ("iprint", ctypes.c_int), ("rescga", ctypes.c_double), ("rescgg", ctypes.c_double),
                ("anum2", _P), ("aden2", _P), ("bnum2", _P), ("bden2", _P), ("n2", ctypes.c_int),
                ("anum4", _P), ("aden4", _P), ("bnum4", _P), ("bden4", _P), ("n4", ctypes.c_int),
                ("anum2g", _P), ("aden2g", _P), ("bnum2g", _P), ("bden2g", _P), ("n2g", ctypes.c_int)]


class OrcTrajOut(ctypes.Structure):
    _fields_ = [("H0", ctypes.c_double), ("H1", ctypes.c_double), ("S0", ctypes.c_double),
                ("S1", ctypes.c_double), ("dH", ctypes.c_double), ("accepted", ctypes.c_int),
                ("md_steps", ctypes.c_int), ("it_pf", ctypes.c_int), ("it_h0", ctypes.c_int),
                ("it_h1", ctypes.c_int), ("it_force_total", ctypes.c_int)]


def lib():
    """Load liboracle.so, building it with make if it is absent."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
        _LIB = ctypes.CDLL(path)
        d = ctypes.c_double
        p = ctypes.c_void_p
        i = ctypes.c_int
        L = ctypes.POINTER(OrcLat)
        _LIB.orc_dslash.argtypes = [L, p, p, p, d, i, i]
        _LIB.orc_derivs.argtypes = [L, p, p, p, d, i]
        _LIB.orc_qmrherm_shifts.argtypes = [L, p, d, p, p, i, i, p, d, i, p, p, p]
        _LIB.orc_qmrherm_shifts.restype = i
        _LIB.orc_set_lanczos_cap.argtypes = [i]
        _LIB.orc_set_lanczos_cap.restype = None
        _LIB.orc_congrad.argtypes = [L, p, d, p, d, i, p]
        _LIB.orc_congrad.restype = i
        _LIB.orc_cg_iterations.argtypes = [L, p, p, d, i, i, p]
        _LIB.orc_force.argtypes = [L, p, d, p, d, p, p, p, p, i, d, i, p, p, p]
        _LIB.orc_hamilton.argtypes = [L, p, d, p, p, d, p, p, i, d, i, p, p]
        _LIB.orc_hamilton.restype = i
        _LIB.orc_fermion.argtypes = [L, p, d, d, i, i, p, p]
        _LIB.orc_fermion.restype = i
        _LIB.orc_hmc_force.argtypes = [L, p, p, d, p, d, i, p]
        _LIB.orc_hmc_force.restype = i
        _LIB.orc_hmc_hamilton.argtypes = [L, p, p, d, p, p, d, i, p]
        _LIB.orc_hmc_hamilton.restype = i
        R = ctypes.POINTER(OrcRng)
        _LIB.orc_rng_set.argtypes = [R, d]
        _LIB.orc_rng_get.argtypes = [R]
        _LIB.orc_rng_get.restype = d
        _LIB.orc_rranf.argtypes = [R]
        _LIB.orc_rranf.restype = d
        _LIB.orc_rano.argtypes = [R, i]
        _LIB.orc_rano.restype = d
        _LIB.orc_gauss0.argtypes = [R, p, i]
        _LIB.orc_gaussp.argtypes = [R, p, i]
        _LIB.orc_rhmc.argtypes = [L, ctypes.POINTER(OrcRhmcPar), R, p, i, i, ctypes.POINTER(OrcTrajOut)]
        _LIB.orc_norm2.argtypes = [ctypes.c_long, p]
        _LIB.orc_norm2.restype = d
        _LIB.orc_set_threads.argtypes = [i]
        _LIB.orc_get_threads.restype = i
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Lattice:
    """Run-time lattice geometry; mirrors the reference's compile-time
    parameter(ksize, ksizet, kthird) (rhmc/cheisdwf.f:31)."""

    def __init__(self, L, Ls, Lt=None, am3=1.0, ibound=-1):
        self.L, self.Ls = int(L), int(Ls)
        self.Lt = int(Lt if Lt is not None else L)
        self.am3, self.ibound = float(am3), int(ibound)
        self.kvol = self.L * self.L * self.Lt
        self.kferm = 8 * self.Ls * self.kvol
        self.c = OrcLat(self.L, self.Lt, self.Ls, self.ibound, self.am3)

    @property
    def shape(self):
        return (2, 4, self.Lt, self.L, self.L, self.Ls)

    def zeros(self):
        return np.zeros(self.shape, dtype=np.complex128)


def _c128(a, lat):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    assert a.size == lat.kferm
    return a


def dslash(lat, r, chi, hpin=0.0, imass=1, dagger=False):
    r = _c128(r, lat)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    out = np.empty_like(r)
    lib().orc_dslash(ctypes.byref(lat.c), _ptr(out), _ptr(r), _ptr(chi), hpin, imass, int(bool(dagger)))
    return out


def derivs(lat, dsdchi, y, z, anum, imass=1):
    lib().orc_derivs(ctypes.byref(lat.c), _ptr(dsdchi), _ptr(_c128(y, lat)), _ptr(_c128(z, lat)),
                     float(anum), imass)
    return dsdchi


def qmrherm(lat, phi, res, anum, aden, chi, hpin=0.0, imass=1, iflag=0, dsdchi=None,
            want_shifts=False):
    """Returns (x, itercg[, x_k]); dsdchi (float64 (3,kvol)) is accumulated for iflag=1."""
    phi = _c128(phi, lat)
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    n = len(aden)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    x = np.empty_like(phi)
    if dsdchi is None:
        dsdchi = np.zeros((3, lat.kvol))
    xk = np.empty((n,) + phi.shape, dtype=np.complex128) if want_shifts else None
    it = lib().orc_qmrherm_shifts(ctypes.byref(lat.c), _ptr(phi), float(res), _ptr(anum), _ptr(aden),
                                  n, iflag, _ptr(chi), float(hpin), imass, _ptr(x), _ptr(dsdchi),
                                  _ptr(xk) if want_shifts else None)
    return (x, it, xk) if want_shifts else (x, it)


def set_lanczos_cap(niterc):
    """Iteration cap of qmrherm (0: the reference's 5000)."""
    lib().orc_set_lanczos_cap(int(niterc))


def congrad(lat, phi, res, chi, x0=None, hpin=0.0, imass=1):
    phi = _c128(phi, lat)
    x = np.zeros_like(phi) if x0 is None else np.array(x0, dtype=np.complex128, copy=True)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    it = lib().orc_congrad(ctypes.byref(lat.c), _ptr(phi), float(res), _ptr(chi), float(hpin), imass, _ptr(x))
    return x, it


def fermion(lat, chi, res, ittt0, hpin=0.0, imass=1):
    """Wall-source correlators (rhmc/cheisdwf.f:897-1058) for the first source
    slice ittt0 (1-based, the reference's int(Lt*rano)+1).  Returns (cferm1,
    cferm2) complex (2 flavours, Lt) and the total congrad iterations."""
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    c1 = np.zeros((2, lat.c.Lt), dtype=np.complex128)
    c2 = np.zeros_like(c1)
    it = lib().orc_fermion(ctypes.byref(lat.c), _ptr(chi), float(res), float(hpin), imass, int(ittt0), _ptr(c1),
                           _ptr(c2))
    return c1, c2, it


def hmc_force(lat, phi, x0, res, chi, beta, imass=1):
    """HMC sibling force (mh/cheisdwf_hmc.f:393-557).  Returns (dSdchi,
    x0 after the chronological update, iterations)."""
    phi = _c128(phi, lat)
    x0 = np.array(x0, dtype=np.complex128, copy=True)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    F = np.zeros((3, lat.kvol))
    it = lib().orc_hmc_force(ctypes.byref(lat.c), _ptr(phi), _ptr(x0), float(res), _ptr(chi), float(beta), imass,
                             _ptr(F))
    return F, x0, it


def hmc_hamilton(lat, phi, x0, res, chi, pp, beta, imass=1):
    """HMC sibling hamilton (mh/cheisdwf_hmc.f:561-620).  Returns ((h, hg, hp,
    s), x0 after, iterations)."""
    phi = _c128(phi, lat)
    x0 = np.array(x0, dtype=np.complex128, copy=True)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    out = np.zeros(4)
    it = lib().orc_hmc_hamilton(ctypes.byref(lat.c), _ptr(phi), _ptr(x0), float(res), _ptr(chi), _ptr(pp),
                                float(beta), imass, _ptr(out))
    return tuple(out), x0, it


def cg_iterations(lat, phi, chi, niter, hpin=0.0, imass=1):
    phi = _c128(phi, lat)
    x = np.empty_like(phi)
    lib().orc_cg_iterations(ctypes.byref(lat.c), _ptr(phi), _ptr(np.ascontiguousarray(chi, dtype=np.float64)),
                            float(hpin), imass, int(niter), _ptr(x))
    return x


def force(lat, phi, res1, chi, beta, remez2g, hpin=0.0, imass=1):
    """Returns (dSdchi (3,kvol), x (the /vector/ contents), (it_guidance, it_force))."""
    phi = _c128(phi, lat)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    a, ad, b, bd = (np.ascontiguousarray(v, dtype=np.float64) for v in remez2g)
    F = np.empty((3, lat.kvol))
    x = np.empty_like(phi)
    its = np.zeros(2, dtype=np.int32)
    lib().orc_force(ctypes.byref(lat.c), _ptr(phi), float(res1), _ptr(chi), float(beta), _ptr(a), _ptr(ad),
                    _ptr(b), _ptr(bd), len(ad), float(hpin), imass, _ptr(F), _ptr(x), _ptr(its))
    return F, x, (int(its[0]), int(its[1]))


def hamilton(lat, phi, res2, chi, pp, beta, remez2, hpin=0.0, imass=1):
    """Returns ((h, hg, hp, s), x, itercg)."""
    phi = _c128(phi, lat)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    _, _, b, bd = (np.ascontiguousarray(v, dtype=np.float64) for v in remez2)
    out = np.zeros(4)
    x = np.empty_like(phi)
    it = lib().orc_hamilton(ctypes.byref(lat.c), _ptr(phi), float(res2), _ptr(chi), _ptr(pp), float(beta),
                            _ptr(b), _ptr(bd), len(bd), float(hpin), imass, _ptr(out), _ptr(x))
    return tuple(out), x, it


class Rng:
    """RRANF/rano/gauss0/gaussp restatement (rhmc/cheisdwf.f:1678-1801), FP64."""

    def __init__(self, seed):
        self.c = OrcRng()
        lib().orc_rng_set(ctypes.byref(self.c), float(seed))

    def rranf(self):
        return lib().orc_rranf(ctypes.byref(self.c))

    def rano(self, stream=0):
        return lib().orc_rano(ctypes.byref(self.c), stream)

    def get(self):
        return lib().orc_rng_get(ctypes.byref(self.c))

    def gauss0(self, kvol):
        ps = np.empty((2, kvol))
        lib().orc_gauss0(ctypes.byref(self.c), _ptr(ps), kvol)
        return ps

    def gaussp(self, kvol):
        ps = np.empty((2, kvol))
        lib().orc_gaussp(ctypes.byref(self.c), _ptr(ps), kvol)
        return ps


def rhmc(lat, params, rng, chi, isweep0, ntraj, remez):
    """Run ntraj trajectories; chi (float64 (3,kvol)) is updated in place.
    params: dict(dt, beta, hpin, imass, iterl, itermax, iprint, rescga, rescgg).
    remez: dict name -> (anum, aden, bnum, bden) for 'remez2', 'remez4', 'remez2g'."""
    keep = []

    def P(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    par = OrcRhmcPar()
    for k in ("dt", "beta", "hpin", "rescga", "rescgg"):
        setattr(par, k, float(params[k]))
    for k in ("imass", "iterl", "itermax", "iprint"):
        setattr(par, k, int(params[k]))
    for nm, suf in (("remez2", "2"), ("remez4", "4"), ("remez2g", "2g")):
        a, ad, b, bd = remez[nm]
        setattr(par, "anum" + suf, P(a))
        setattr(par, "aden" + suf, P(ad))
        setattr(par, "bnum" + suf, P(b))
        setattr(par, "bden" + suf, P(bd))
        setattr(par, "n" + suf, len(ad))
    outs = (OrcTrajOut * ntraj)()
    assert chi.dtype == np.float64 and chi.flags.c_contiguous
    lib().orc_rhmc(ctypes.byref(lat.c), ctypes.byref(par), ctypes.byref(rng.c), _ptr(chi), isweep0, ntraj, outs)
    return [{f: getattr(o, f) for f, _ in OrcTrajOut._fields_} for o in outs]


def hmc_trajectory(olat, chi, rng, par):
    """One trajectory of the HMC sibling's program (mh/cheisdwf_hmc.f:140-330)
    on the oracle: gauss0 sources in flavour, Dirac, s order, Phi = M^dagger R,
    the chronological guess X0 = Phi, momenta from two gaussp calls, H0, a
    leapfrog of random length (one yrun draw per step after its force, end
    with probability 1/iterl), H1, and the Metropolis draw only when dH < 0
    (accepted unless the draw exceeds exp(dH)).  chi is replaced in place on
    acceptance.  Returns dH, dS, accepted, md_steps, H0, H1 and the gauge and
    momentum parts of both Hamiltonians."""
    kv = olat.kvol
    chit = chi.copy()
    R = np.empty(olat.shape, np.complex128)
    Rv = R.reshape(2, 4, kv, olat.Ls)
    for f in range(2):
        for d in range(4):
            for s in range(olat.Ls):
                ps = rng.gauss0(kv)
                Rv[f, d, :, s] = ps[0] + 1j * ps[1]
    Phi = dslash(olat, R, chit, 0.0, par["imass"], True)
    x0 = Phi.copy()
    pp = np.empty((3, kv))
    ps = rng.gaussp(kv)
    pp[0], pp[1] = ps[0], ps[1]
    pp[2] = rng.gaussp(kv)[0]
    (H0, hg0, hp0, S0), x0, _ = hmc_hamilton(olat, Phi, x0, par["rescga"], chit, pp, par["beta"], par["imass"])
    F, x0, _ = hmc_force(olat, Phi, x0, par["rescgg"], chit, par["beta"], par["imass"])
    pp -= 0.5 * par["dt"] * F
    steps = 0
    for step in range(1, par["itermax"] + 1):
        chit += par["dt"] * pp
        F, x0, _ = hmc_force(olat, Phi, x0, par["rescgg"], chit, par["beta"], par["imass"])
        steps = step
        if rng.rano(1) < 1.0 / par["iterl"]:
            pp -= 0.5 * par["dt"] * F
            break
        pp -= par["dt"] * F
    (H1, hg1, hp1, S1), x0, _ = hmc_hamilton(olat, Phi, x0, par["rescga"], chit, pp, par["beta"], par["imass"])
    dH = H0 - H1
    acc = True
    if dH < 0:
        acc = not (rng.rano(1) > np.exp(dH))
    if acc:
        chi[...] = chit
    return dict(dH=dH, dS=S0 - S1, accepted=int(acc), md_steps=steps, H0=H0, H1=H1, hg0=hg0, hp0=hp0, hg1=hg1,
                hp1=hp1)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()
