"""RHMC trajectories of the reference's main program (program dwf3d,
rhmc/cheisdwf.f:207-413) with every fermion routine on the GPU.

Per trajectory, in the reference's order of random-number consumption:
  1. pseudofermion heatbath: complex Gaussians (gauss0, E|z|^2 = 1) in
     flavour, Dirac, s order, Phi = (M^dagger M)^{1/4} R via the remez4 a-set
     (rhmc/cheisdwf.f:221-245);
  2. momenta from two gaussp calls (rhmc/cheisdwf.f:249-257);
  3. H0 = hamilton, then leapfrog with a random length: after every step one
     draw from the yrun stream ends the trajectory with probability 1/iterl
     (rhmc/cheisdwf.f:261-315);
  4. H1 = hamilton; accept if rano(yrun) < exp(H0 - H1) (rhmc/cheisdwf.f:320-345);
  5. every iprint trajectories the measurement routine takes one draw from
     the yran stream (rhmc/cheisdwf.f:359, :929).
All reals are FP64 (the reference's -fdefault-real-8 semantics).
"""
import ctypes

import numpy as np

from . import fermion as fm
from . import remez as rz
from ._lib import Rng, lib

DEFAULTS = dict(dt=0.05, beta=0.475, am=0.0, hpin=0.0, imass=1, iterl=20, itermax=1000, iprint=5,
                rescga=1e-9, rescgg=1e-6)


class RandomStream:
    """The reference generator (RRANF + rano shuffle + Box-Muller), exact."""

    def __init__(self, seed=693764523745.0):
        self.c = Rng()
        lib().chs_rng_set(ctypes.byref(self.c), float(seed))

    def rano(self, stream=0):
        return lib().chs_rano(ctypes.byref(self.c), stream)

    def seed(self):
        return lib().chs_rng_get(ctypes.byref(self.c))

    def gauss0(self, kvol):
        ps = np.empty((2, kvol))
        lib().chs_gauss0(ctypes.byref(self.c), ps.ctypes.data_as(ctypes.c_void_p), kvol)
        return ps

    def gaussp(self, kvol):
        ps = np.empty((2, kvol))
        lib().chs_gaussp(ctypes.byref(self.c), ps.ctypes.data_as(ctypes.c_void_p), kvol)
        return ps


def heatbath_source(lat, rng):
    """R(s, i, d, f) = complex Gaussian, drawn in flavour, Dirac, s order."""
    R = np.empty(lat.shape, dtype=np.complex128)
    Rv = R.reshape(2, 4, lat.kvol, lat.Ls)
    for f in range(2):
        for d in range(4):
            for s in range(lat.Ls):
                ps = rng.gauss0(lat.kvol)
                Rv[f, d, :, s] = ps[0] + 1j * ps[1]
    return R


def trajectory(lat, chi, rng, isweep, par=None, sets=None):
    """One RHMC trajectory; chi (float64 (3, kvol)) is replaced in place when
    the proposal is accepted.  Returns a dict of H0, H1, dH, accepted and
    iteration counts."""
    p = dict(DEFAULTS)
    if par:
        p.update(par)
    sets = sets or {k: rz.load(k) for k in ("remez2", "remez4", "remez2g")}
    a4, ad4, _, _ = sets["remez4"]
    kv = lat.kvol
    chit = chi.copy()
    R = heatbath_source(lat, rng)
    Phi, it_pf = fm.qmrherm(lat, R, p["rescga"], a4, ad4, chit, p["am"], p["hpin"], p["imass"], 0)
    pp = np.empty((3, kv))
    ps = rng.gaussp(kv)
    pp[0], pp[1] = ps[0], ps[1]
    ps = rng.gaussp(kv)
    pp[2] = ps[0]
    (H0, _, _, S0), _, it_h0 = fm.hamilton(lat, Phi, p["rescga"], chit, pp, p["beta"], sets["remez2"], p["am"],
                                           p["hpin"], p["imass"], isweep)
    F, _, _ = fm.force(lat, Phi, p["rescgg"], chit, p["beta"], sets["remez2g"], p["am"], p["hpin"], p["imass"])
    pp -= 0.5 * p["dt"] * F
    proby = 1.0 / p["iterl"]
    steps = 0
    for it in range(1, p["itermax"] + 1):
        chit += p["dt"] * pp
        F, _, _ = fm.force(lat, Phi, p["rescgg"], chit, p["beta"], sets["remez2g"], p["am"], p["hpin"],
                           p["imass"])
        steps = it
        if rng.rano(1) < proby:
            pp -= 0.5 * p["dt"] * F
            break
        pp -= p["dt"] * F
    (H1, _, _, S1), _, it_h1 = fm.hamilton(lat, Phi, p["rescga"], chit, pp, p["beta"], sets["remez2"], p["am"],
                                           p["hpin"], p["imass"], isweep)
    dH = H0 - H1
    accepted = rng.rano(1) < np.exp(dH)
    if accepted:
        chi[...] = chit
    if p["iprint"] > 0 and isweep % p["iprint"] == 0:
        rng.rano(0)
    return dict(H0=H0, H1=H1, S0=S0, S1=S1, dH=dH, accepted=int(accepted), md_steps=steps, it_pf=it_pf,
                it_h0=it_h0, it_h1=it_h1)
