"""Host-side mirror of the reference's fermion routines (rhmc/cheisdwf.f),
running on the MI355X through libchs.

The routine names, argument meanings and order follow the reference:
  dslash(Phi,R,chi,am,hpin,imass)                          rhmc/cheisdwf.f:2634
  dslashd(Phi,R,chi,am,hpin,imass)                         rhmc/cheisdwf.f:3071
  qmrherm(Phi,res,itercg,am,hpin,imass,anum,aden,ndiag,iflag)  rhmc/cheisdwf.f:599
  congrad(Phi,res,itercg,am,hpin,imass)                    rhmc/cheisdwf.f:1064
  force(Phi,res1,am,hpin,imass)                            rhmc/cheisdwf.f:470
  hamilton(Phi,h,hg,hp,s,res2,isweep,am,hpin,imass)        rhmc/cheisdwf.f:522
  derivs(R,X2,anum,imass)                                  rhmc/cheisdwf.f:747
The reference's COMMON inputs (/trial/ chit and pp, /para/ beta, the Remez
blocks) become explicit arguments and its outputs (/vector/, /sforce/)
return values.  Arrays are numpy complex128 of shape (2, 4, Lt, L, L, Ls) in C
order, i.e. the reference's Phi(Ls, L, L, Lt, 4, 2); chi, pp and dSdchi are
float64 (3, kvol).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, CHS_FP64, CHS_FP32, CHS_BF16, PRECISIONS, BYTES_PER_COMPLEX

_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


def _check_keep(lat, rc, result):
    """check(), with the results a solve wrote before stopping at its
    iteration cap attached to the NoConvergence it raises (.result)."""
    try:
        check(lat.h, rc)
    except _lib.NoConvergence as e:
        e.result = result
        raise


class Lattice:
    """One lattice geometry and one HIP stream on one GPU (a libchs context).
    Replaces the compile-time parameter(ksize, ksizet, kthird) and /para/ am3,
    ibound of the reference (rhmc/cheisdwf.f:31,40,85)."""

    def __init__(self, L, Ls, Lt=None, am3=1.0, ibound=-1, device=0):
        self.L, self.Ls = int(L), int(Ls)
        self.Lt = int(Lt if Lt is not None else L)
        self.am3, self.ibound = float(am3), int(ibound)
        h = lib().chs_ctx_create(self.L, self.Lt, self.Ls, self.ibound, self.am3, int(device))
        if not h:
            raise _lib.ChsError(_lib.CHS_EINVAL, lib().chs_create_error().decode())
        self.h = h
        self.kvol = self.L * self.L * self.Lt
        self.kferm = 8 * self.Ls * self.kvol

    @property
    def shape(self):
        return (2, 4, self.Lt, self.L, self.L, self.Ls)

    def close(self):
        if getattr(self, "h", None):
            lib().chs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(self.h, lib().chs_synchronize(self.h))

    # ---- device mirrors of read-only host spinors (include/chs.h) ----------
    def mirror(self, a):
        """Keep a device copy of the host spinor a (C-contiguous complex128,
        passed as is to later calls): it is uploaded again only after
        touch(a).  release(a) before a is freed."""
        if a.dtype != np.complex128 or not a.flags.c_contiguous or a.size != self.kferm:
            raise ValueError("a mirrored spinor must be C-contiguous complex128 with kferm elements")
        check(self.h, lib().chs_mirror_register(self.h, _p(a)))
        return a

    def touch(self, a):
        """a (mirrored) was written: its next use uploads it."""
        check(self.h, lib().chs_mirror_touch(self.h, _p(a)))

    def release(self, a):
        check(self.h, lib().chs_mirror_release(self.h, _p(a)))

    def transfer_bytes(self):
        """(host -> device, device -> host) spinor bytes moved so far."""
        h2d, d2h = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
        check(self.h, lib().chs_transfer_bytes(self.h, ctypes.byref(h2d), ctypes.byref(d2h)))
        return h2d.value, d2h.value

    def set_lanczos_cap(self, niterc):
        """Iteration cap of the multi-shift solves (0: the reference's 5000)."""
        check(self.h, lib().chs_set_lanczos_cap(self.h, int(niterc)))

    def set_shift_freeze(self, f):
        """Shift freezing of the multi-shift solves: 0 (default) updates every
        shift as the reference does; 0 < f <= 1 stops updating shift k's
        vectors once |phimod rho_k| < f * 2 resid."""
        check(self.h, lib().chs_set_shift_freeze(self.h, float(f)))

    def _spinor(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.size != self.kferm:
            raise ValueError("spinor has %d elements, lattice needs %d" % (a.size, self.kferm))
        return a

    def _field(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != 3 * self.kvol:
            raise ValueError("field has %d elements, lattice needs %d" % (a.size, 3 * self.kvol))
        return a


# ---- operators --------------------------------------------------------------
def dslash(lat, R, chi, am=0.0, hpin=0.0, imass=1):
    """Phi = M R (rhmc/cheisdwf.f:2634-3069)."""
    R = lat._spinor(R)
    phi = np.empty_like(R)
    check(lat.h, lib().chs_dslash(lat.h, _p(phi), _p(R), _p(lat._field(chi)), am, hpin, imass))
    return phi.reshape(lat.shape)


def dslashd(lat, R, chi, am=0.0, hpin=0.0, imass=1):
    """Phi = M^dagger R (rhmc/cheisdwf.f:3071-3501)."""
    R = lat._spinor(R)
    phi = np.empty_like(R)
    check(lat.h, lib().chs_dslashd(lat.h, _p(phi), _p(R), _p(lat._field(chi)), am, hpin, imass))
    return phi.reshape(lat.shape)


def qmrherm(lat, Phi, res, anum, aden, chi, am=0.0, hpin=0.0, imass=1, iflag=0, dSdchi=None):
    """Multi-shift Lanczos solve of (M^dagger M + aden_k) x_k = Phi
    (rhmc/cheisdwf.f:599-745).  Returns (x, itercg): x is /vector/ (a0 Phi +
    sum a_k x_k for iflag=0, a0 Phi for iflag=1).  For iflag=1 the force
    contributions are added to dSdchi in place.  A solve that reaches the
    iteration cap (5000) or a non-finite residual raises NoConvergence, whose
    .result holds (x, itercg) as the reference would return them."""
    Phi = lat._spinor(Phi)
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    if len(anum) != len(aden) + 1:
        raise ValueError("anum needs ndiag+1 entries")
    x = np.empty_like(Phi)
    it = ctypes.c_int(0)
    F = None
    if iflag == 1:
        if dSdchi is None or dSdchi.dtype != np.float64 or not dSdchi.flags.c_contiguous:
            raise ValueError("iflag=1 needs a C-contiguous float64 dSdchi to accumulate into")
        F = dSdchi
    rc = lib().chs_qmrherm(lat.h, _p(Phi), res, ctypes.byref(it), am, hpin, imass, _p(anum), _p(aden), len(aden),
                           iflag, _p(lat._field(chi)), _p(x), _p(F) if F is not None else None)
    _check_keep(lat, rc, (x.reshape(lat.shape), it.value))
    return x.reshape(lat.shape), it.value


def congrad(lat, Phi, res, chi, x0=None, am=0.0, hpin=0.0, imass=1):
    """CGNR for M^dagger M x = Phi from the initial guess x0
    (rhmc/cheisdwf.f:1064-1157).  Returns (x, itercg); raises NoConvergence
    (carrying no result) when the reference would print 'iterations of
    congrad exceeds niterc'."""
    Phi = lat._spinor(Phi)
    x = np.zeros_like(Phi) if x0 is None else np.array(lat._spinor(x0), copy=True)
    it = ctypes.c_int(0)
    check(lat.h, lib().chs_congrad(lat.h, _p(Phi), res, ctypes.byref(it), am, hpin, imass, _p(lat._field(chi)),
                                   _p(x)))
    return x.reshape(lat.shape), it.value


def force(lat, Phi, res1, chi, beta, remez2g, am=0.0, hpin=0.0, imass=1):
    """dSdchi = 2 beta chi + pseudofermion force (rhmc/cheisdwf.f:470-518).
    remez2g = (anum, aden, bnum, bden).  Returns (dSdchi, x, (it_guidance,
    it_force))."""
    Phi = lat._spinor(Phi)
    a, ad, b, bd = (np.ascontiguousarray(v, dtype=np.float64) for v in remez2g)
    F = np.empty((3, lat.kvol))
    x = np.empty_like(Phi)
    its = (ctypes.c_int * 2)()
    rc = lib().chs_force(lat.h, _p(Phi), res1, am, hpin, imass, _p(lat._field(chi)), beta, _p(a), _p(ad), _p(b),
                         _p(bd), len(ad), _p(F), _p(x), its)
    _check_keep(lat, rc, (F, x.reshape(lat.shape), (its[0], its[1])))
    return F, x.reshape(lat.shape), (its[0], its[1])


def fermion(lat, chi, res, ittt0, hpin=0.0, imass=1, precision="fp64"):
    """Wall-source correlators (rhmc/cheisdwf.f:897-1058) for the first source
    slice ittt0 (1..Lt, the reference's int(Lt*rano)+1).  precision: the
    solver form of the ten propagator solves, "fp64" (the reference CGNR),
    "fp32", "fp16" or "learned" (flexible CG with the learned preconditioner
    of learn.Preconditioner).  Returns (cferm1, cferm2) complex arrays
    (2 flavours, Lt) and the total solver iterations (aviter = iterations /
    10)."""
    c1 = np.zeros((2, lat.Lt), dtype=np.complex128)
    c2 = np.zeros_like(c1)
    it = ctypes.c_int(0)
    form = _lib.CHS_LEARNED if precision == "learned" else PRECISIONS[precision]
    rc = lib().chs_fermion(lat.h, _p(lat._field(chi)), res, hpin, imass, int(ittt0), form, _p(c1), _p(c2),
                           ctypes.byref(it))
    _check_keep(lat, rc, (c1, c2, it.value))
    return c1, c2, it.value


def hmc_force(lat, Phi, x0, res, chi, beta, imass=1, am=0.0):
    """HMC sibling force (mh/cheisdwf_hmc.f:393-557).  x0 (the chronological
    guess, reference layout) is updated in place to 2 X1 - x0.  Returns
    (dSdchi, iterations)."""
    Phi = lat._spinor(Phi)
    assert x0.dtype == np.complex128 and x0.flags.c_contiguous and x0.shape == lat.shape
    F = np.empty((3, lat.kvol))
    it = ctypes.c_int(0)
    rc = lib().chs_hmc_force(lat.h, _p(Phi), _p(x0), res, am, imass, _p(lat._field(chi)), beta, _p(F),
                             ctypes.byref(it))
    _check_keep(lat, rc, (F, it.value))
    return F, it.value


def hmc_hamilton(lat, Phi, x0, res, chi, pp, beta, imass=1, am=0.0):
    """HMC sibling hamilton (mh/cheisdwf_hmc.f:561-620); x0 receives X1.
    Returns ((h, hg, hp, s), iterations)."""
    Phi = lat._spinor(Phi)
    assert x0.dtype == np.complex128 and x0.flags.c_contiguous and x0.shape == lat.shape
    h4 = np.empty(4)
    it = ctypes.c_int(0)
    rc = lib().chs_hmc_hamilton(lat.h, _p(Phi), _p(x0), res, am, imass, _p(lat._field(chi)), _p(lat._field(pp)),
                                beta, _p(h4), ctypes.byref(it))
    _check_keep(lat, rc, (tuple(h4), it.value))
    return tuple(h4), it.value


def set_force_precision(lat, precision):
    """Precision of the solves inside the force for this lattice context:
    the multi-shift solves of force() in "fp64" or "fp32" storage ("fp16"
    runs them in FP32); the CG of hmc_force() as the reference CGNR ("fp64")
    or the reliable-update CG in "fp32" or "fp16" form.  hamilton stays
    FP64."""
    check(lat.h, lib().chs_set_force_precision(lat.h, PRECISIONS[precision]))


def hamilton(lat, Phi, res2, chi, pp, beta, remez2, am=0.0, hpin=0.0, imass=1, isweep=1):
    """H = beta sum chi^2 + 1/2 sum p^2 + Re Phi^dagger r_b(M^dagger M) Phi
    (rhmc/cheisdwf.f:522-589).  Returns ((h, hg, hp, s), x, itercg)."""
    Phi = lat._spinor(Phi)
    _, _, b, bd = (np.ascontiguousarray(v, dtype=np.float64) for v in remez2)
    out = [ctypes.c_double() for _ in range(4)]
    x = np.empty_like(Phi)
    it = ctypes.c_int(0)
    rc = lib().chs_hamilton(lat.h, _p(Phi), *[ctypes.byref(o) for o in out], res2, isweep, am, hpin, imass,
                            _p(lat._field(chi)), _p(lat._field(pp)), beta, _p(b), _p(bd), len(bd), _p(x),
                            ctypes.byref(it))
    _check_keep(lat, rc, (tuple(o.value for o in out), x.reshape(lat.shape), it.value))
    return tuple(o.value for o in out), x.reshape(lat.shape), it.value


def derivs(lat, R, X2, anum, imass, dSdchi):
    """dSdchi += wall bilinears of R and X2 = M R (rhmc/cheisdwf.f:747-849)."""
    if dSdchi.dtype != np.float64 or not dSdchi.flags.c_contiguous or dSdchi.size != 3 * lat.kvol:
        raise ValueError("dSdchi must be C-contiguous float64 (3, kvol)")
    check(lat.h, lib().chs_derivs(lat.h, _p(lat._spinor(R)), _p(lat._spinor(X2)), float(anum), imass,
                                  _p(dSdchi)))
    return dSdchi


# ---- device-resident data ----------------------------------------------------
class DeviceArray:
    """A buffer in HBM owned by the library (hipMalloc on the context's GPU)."""

    def __init__(self, lat, nbytes):
        self.lat, self.nbytes = lat, int(nbytes)
        self.ptr = lib().chs_alloc(lat.h, self.nbytes)
        if not self.ptr:
            check(lat.h, _lib.CHS_ENOMEM)

    @classmethod
    def from_numpy(cls, lat, a):
        a = np.ascontiguousarray(a)
        d = cls(lat, a.nbytes)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        check(self.lat.h, lib().chs_upload(self.lat.h, self.ptr, _p(a), self.nbytes))

    def download(self, dtype, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(self.lat.h, lib().chs_download(self.lat.h, _p(out), self.ptr, self.nbytes))
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().chs_free(self.lat.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.lat.h:
                self.free()
        except Exception:
            pass


def spinor_dev(lat, a=None, precision="fp64"):
    """Device spinor in the native layout and the given storage precision,
    optionally filled from a complex128 host array in the reference layout
    (reordered and converted on the device)."""
    prec = PRECISIONS[precision]
    d = DeviceArray(lat, lat.kferm * BYTES_PER_COMPLEX[prec])
    if a is not None:
        a = lat._spinor(a)
        ref = DeviceArray.from_numpy(lat, a)
        if prec == CHS_FP64:
            check(lat.h, lib().chs_layout_dev(lat.h, d.ptr, ref.ptr, CHS_FP64, 1))
        else:
            nat = DeviceArray(lat, lat.kferm * 16)
            check(lat.h, lib().chs_layout_dev(lat.h, nat.ptr, ref.ptr, CHS_FP64, 1))
            check(lat.h, lib().chs_convert_dev(lat.h, d.ptr, prec, nat.ptr, CHS_FP64, lat.kferm))
            lat.synchronize()
            nat.free()
        lat.synchronize()
        ref.free()
    d.prec = prec
    return d


def to_host(lat, d):
    """complex128 host copy, in the reference layout, of a native device
    spinor of any storage precision."""
    prec = getattr(d, "prec", CHS_FP64)
    nat = d
    if prec != CHS_FP64:
        nat = DeviceArray(lat, lat.kferm * 16)
        check(lat.h, lib().chs_convert_dev(lat.h, nat.ptr, CHS_FP64, d.ptr, prec, lat.kferm))
    ref = DeviceArray(lat, lat.kferm * 16)
    check(lat.h, lib().chs_layout_dev(lat.h, ref.ptr, nat.ptr, CHS_FP64, 0))
    out = ref.download(np.complex128, lat.shape)
    ref.free()
    if nat is not d:
        nat.free()
    return out


def dslash_dev(lat, out, inp, chi_dev, hpin=0.0, imass=1, dagger=False):
    prec = getattr(inp, "prec", CHS_FP64)
    check(lat.h, lib().chs_dslash_dev(lat.h, out.ptr, inp.ptr, chi_dev.ptr, hpin, imass, int(bool(dagger)), prec))


def solve_mdagm(lat, b_dev, x_dev, chi_dev, tol=1e-10, precision="fp64", hpin=0.0, imass=1, delta=0.0,
                maxiter=0):
    """M^dagger M x = b on device arrays to FP64 true relative residual tol.
    Returns a dict of solve statistics."""
    st = _lib.SolveStats()
    check(lat.h, lib().chs_solve_mdagm_dev(lat.h, b_dev.ptr, x_dev.ptr, chi_dev.ptr, hpin, imass, tol,
                                           PRECISIONS[precision], delta, maxiter, ctypes.byref(st)))
    return {f: getattr(st, f) for f, _ in _lib.SolveStats._fields_}


def qmrherm_prec_dev(lat, phi_dev, res, anum, aden, chi_dev, x_dev, precision="fp64", hpin=0.0, imass=1, iflag=0,
                     F_dev=None):
    """qmrherm_dev with the spinors stored in precision ("fp64" or "fp32")."""
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    it = ctypes.c_int(0)
    check(lat.h, lib().chs_qmrherm_prec_dev(lat.h, phi_dev.ptr, res, ctypes.byref(it), hpin, imass, _p(anum),
                                            _p(aden), len(aden), iflag, chi_dev.ptr, x_dev.ptr,
                                            F_dev.ptr if F_dev is not None else None, PRECISIONS[precision]))
    return it.value


def qmrherm_shifts(lat, Phi, res, aden, chi, hpin=0.0, imass=1):
    """Every shifted solution x_k = (M^dagger M + aden[k])^{-1} Phi of the
    multi-shift solve (the reference's stopping rule).  Returns (x_k complex128
    (ndiag,) + lat.shape in the reference layout, itercg)."""
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    nd, n = len(aden), lat.kferm
    phi_dev = spinor_dev(lat, Phi)
    chi_dev = DeviceArray.from_numpy(lat, lat._field(chi))
    xk = DeviceArray(lat, nd * n * 16)
    it = ctypes.c_int(0)
    check(lat.h, lib().chs_qmrherm_shifts_dev(lat.h, phi_dev.ptr, res, ctypes.byref(it), hpin, imass, _p(aden), nd,
                                              chi_dev.ptr, xk.ptr))
    out = np.empty((nd,) + lat.shape, np.complex128)
    ref = DeviceArray(lat, n * 16)
    for k in range(nd):
        check(lat.h, lib().chs_layout_dev(lat.h, ref.ptr, ctypes.c_void_p(xk.ptr + k * n * 16), CHS_FP64, 0))
        out[k] = ref.download(np.complex128, lat.shape)
    for d in (ref, xk, chi_dev, phi_dev):
        d.free()
    return out, it.value


def qmrherm_dev(lat, phi_dev, res, anum, aden, chi_dev, x_dev, hpin=0.0, imass=1, iflag=0, F_dev=None):
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    it = ctypes.c_int(0)
    check(lat.h, lib().chs_qmrherm_dev(lat.h, phi_dev.ptr, res, ctypes.byref(it), hpin, imass, _p(anum), _p(aden),
                                       len(aden), iflag, chi_dev.ptr, x_dev.ptr,
                                       F_dev.ptr if F_dev is not None else None))
    return it.value


def time_dslash(lat, precision="fp64", dagger=False, nbuf=4, iters=100, hpin=0.0, imass=1):
    """Average ms per dslash launch, cycling nbuf buffer pairs (HIP events)."""
    ms = ctypes.c_double()
    check(lat.h, lib().chs_time_dslash(lat.h, PRECISIONS[precision], int(bool(dagger)), nbuf, iters, hpin, imass,
                                       ctypes.byref(ms)))
    return ms.value


def time_copy(lat, precision="fp64", nbuf=4, iters=100):
    """Average ms per plain copy of one spinor (2V bytes, random contents,
    non-temporal stores), cycling nbuf buffer pairs (HIP events)."""
    ms = ctypes.c_double()
    check(lat.h, lib().chs_time_copy(lat.h, PRECISIONS[precision], nbuf, iters, ctypes.byref(ms)))
    return ms.value


def time_copy_grid(lat, grid, nbuf=2, iters=40):
    """Average ms per FP64 copy of one spinor by `grid` workgroups striding
    over it (a persistent grid; grid <= 0: one element per thread)."""
    ms = ctypes.c_double()
    check(lat.h, lib().chs_time_copy_grid(lat.h, nbuf, iters, int(grid), ctypes.byref(ms)))
    return ms.value


def time_cg_stencil(lat, precision="fp32", nbuf=4, iters=100, hpin=0.0, imass=1):
    """Average ms per launch of the solve's fused CG stencil (HIP events)."""
    ms = ctypes.c_double()
    check(lat.h, lib().chs_time_cg_stencil(lat.h, PRECISIONS[precision], nbuf, iters, hpin, imass, ctypes.byref(ms)))
    return ms.value
