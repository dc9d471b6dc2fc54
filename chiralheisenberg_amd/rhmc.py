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
import os

import numpy as np

from . import fermion as fm
from ._lib import NoConvergence
from . import outfile
from .fortfmt import list_directed as LD
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


def carry_on(call):
    """The reference's multi-shift solve ends at its iteration cap with
    the last iterate and no message (rhmc/cheisdwf.f:603,639): run call()
    and, where the library reports the cap (NoConvergence with the results
    it kept), go on with those results; any other error stands."""
    try:
        return call()
    except NoConvergence as e:
        if getattr(e, "result", None) is None:
            raise
        return e.result


def leapfrog(lat, Phi, chit, pp, par, remez2g, stop, its=None):
    """The molecular-dynamics evolution of rhmc/cheisdwf.f:261-315, in place
    on chit and pp: a half momentum step, then steps of chit += dt pp and the
    force; after each step's force stop(step) decides the end, and the last
    momentum step is a half step (a full one when itermax ends the loop).
    The reference's stop is one rano(yrun) < 1/iterl draw per step.  Returns
    the number of steps; its (a list of two, if given) receives the force's
    multi-shift iterations, guidance and force solves (the reference's ancg
    and ancgf, rhmc/cheisdwf.f:507,515)."""
    dt = par["dt"]
    args = (par["rescgg"], chit, par["beta"], remez2g, par["am"], par["hpin"], par["imass"])

    def force():
        F, _, (i1, i2) = carry_on(lambda: fm.force(lat, Phi, *args))
        if its is not None:
            its[0] += i1
            its[1] += i2
        return F
    F = force()
    pp -= 0.5 * dt * F
    for it in range(1, par["itermax"] + 1):
        chit += dt * pp
        F = force()
        if stop(it):
            pp -= 0.5 * dt * F
            return it
        pp -= dt * F
    return par["itermax"]


def trajectory(lat, chi, rng, isweep, par=None, sets=None, measure_draw=True):
    """One RHMC trajectory; chi (float64 (3, kvol)) is replaced in place when
    the proposal is accepted.  Returns a dict of H0, H1, dH, accepted, the
    gauge/momentum parts of both Hamiltonians, the mean square velocity and
    iteration counts.  measure_draw: take the measurement's one yran draw
    here (run() takes it inside the measurement instead)."""
    p = dict(DEFAULTS)
    if par:
        p.update(par)
    sets = sets or {k: rz.load(k) for k in ("remez2", "remez4", "remez2g")}
    a4, ad4, _, _ = sets["remez4"]
    kv = lat.kvol
    chit = chi.copy()
    R = heatbath_source(lat, rng)
    Phi, it_pf = carry_on(lambda: fm.qmrherm(lat, R, p["rescga"], a4, ad4, chit, p["am"], p["hpin"], p["imass"], 0))
    pp = np.empty((3, kv))
    ps = rng.gaussp(kv)
    pp[0], pp[1] = ps[0], ps[1]
    ps = rng.gaussp(kv)
    pp[2] = ps[0]
    (H0, hg0, hp0, S0), _, it_h0 = carry_on(lambda: fm.hamilton(lat, Phi, p["rescga"], chit, pp, p["beta"],
                                                                sets["remez2"], p["am"], p["hpin"], p["imass"], isweep))
    proby = 1.0 / p["iterl"]
    its = [0, 0]
    steps = leapfrog(lat, Phi, chit, pp, p, sets["remez2g"], lambda it: rng.rano(1) < proby, its)
    (H1, hg1, hp1, S1), _, it_h1 = carry_on(lambda: fm.hamilton(lat, Phi, p["rescga"], chit, pp, p["beta"],
                                                                sets["remez2"], p["am"], p["hpin"], p["imass"], isweep))
    dH = H0 - H1
    accepted = rng.rano(1) < np.exp(dH)
    if accepted:
        chi[...] = chit
    if measure_draw and p["iprint"] > 0 and isweep % p["iprint"] == 0:
        rng.rano(0)
    return dict(H0=H0, H1=H1, S0=S0, S1=S1, dH=dH, dS=S0 - S1, accepted=int(accepted), md_steps=steps,
                hg0=hg0, hp0=hp0, hg1=hg1, hp1=hp1, vel2=float(np.sum(pp * pp) / kv), it_pf=it_pf,
                it_h0=it_h0, it_h1=it_h1, it_guidance=its[0], it_force=its[1])


def read_midout(path):
    """midout: dt, beta, am3, am, hpin, imass, iterl, iter2 on the first line
    (rhmc/cheisdwf.f:116)."""
    v = open(path).readline().split()
    dt, beta, am3, am, hpin = (float(t) for t in v[:5])
    imass, iterl, iter2 = (int(t) for t in v[5:8])
    return dict(dt=dt, beta=beta, am3=am3, am=am, hpin=hpin, imass=imass, iterl=iterl, iter2=iter2)


def aux_observables(chi):
    """cond(a) = <chi_a>, susc(a) = (sum chi_a)^2 / kvol, condmod = |cond|,
    susctot = sum susc (rhmc/cheisdwf.f:366-384)."""
    kvol = chi.shape[1]
    tot = chi.sum(axis=1)
    cond = tot / kvol
    susc = tot * tot / kvol
    return cond, susc, float(np.sqrt(np.sum(cond * cond))), float(susc.sum())


def run(workdir, L, Ls, Lt=None, par=None, iter2=None, iprint=5, icheck=None, iread=1, iwrite=1, iseed=0,
        rescgm=1e-9, force_precision="fp64", fermion_precision="fp64", sets=None, learned=None, shift_freeze=0.0):
    """The main program dwf3d (rhmc/cheisdwf.f:85-460) with every fermion
    routine on the GPU.  workdir holds the inputs (midout unless par gives
    the parameters, con when iread = 1, and the rational approximations
    remez2 / remez4 / remez2g as the program reads them there, the bundled
    copies where they are absent, unless sets gives them) and receives the
    reference's outputs:
    fort.11 (isweep, gauge and momentum action per site), control (dH, dS,
    accepted), fort.12/13/14 (cond, susc, condmod and susctot every iprint
    trajectories), fort.5001/5011/5002/5012 (fermion correlators), con (every
    icheck trajectories and at the end) and output (header, seeds and
    averages in the program's own formats 9001/9004/9022).  The other files
    hold one list-directed record per line, as the program's WRITE(unit,*)
    prints it (fortfmt.list_directed).  Returns the averages.

    shift_freeze = f > 0 freezes converged shifts in every multi-shift solve
    (fermion.Lattice.set_shift_freeze; 0, the default, is the reference's
    solve).

    learned = dict(pairs=N, epochs=E, lr=lr, seed=s) switches on the learned
    preconditioner for the measurement's propagator solves: the
    (source, solution) pairs of the first measurements' plain-form solves
    (fermion_precision; ten per measurement) are stored on the device until
    there are N, the network is trained on them (E epochs of Adam), and every
    later measurement solves with flexible CG and the network.  The Markov
    chain does not depend on it (the measurement draws the same random
    number either way); the returned averages then hold "learned" (pairs,
    per-epoch losses, the trajectory after which it was trained)."""
    from . import checkpoint
    if learned and int(learned.get("pairs", 0)) < 1:
        raise ValueError("learned: 'pairs' must be >= 1")
    p = dict(DEFAULTS)
    mid = os.path.join(workdir, "midout")
    if par is None and os.path.exists(mid):
        par = read_midout(mid)
    p.update(par or {})
    p["iprint"] = iprint
    iter2 = int(iter2 if iter2 is not None else p.get("iter2", 1))
    icheck = icheck or iprint
    lat = fm.Lattice(L, Ls, Lt, am3=p.get("am3", 1.0))
    fm.set_force_precision(lat, force_precision)
    lat.set_shift_freeze(shift_freeze)
    kv = lat.kvol
    con = os.path.join(workdir, "con")
    if iread == 1:
        chi, seed = checkpoint.sread(con, kv, p["beta"])
    else:
        chi, seed = np.full((3, kv), 0.1 / p["beta"]), checkpoint.DEFAULT_SEED
    if iseed != 0:
        seed = checkpoint.DEFAULT_SEED
    rng = RandomStream(seed)
    rng.rano(0)                      # y = rano(yran, idum) with idum = -1 (rhmc/cheisdwf.f:127-129)
    outfile.header(os.path.join(workdir, "output"), seed, rng.seed(), L, lat.Lt, Ls, p, iter2, rhmc=True)
    sets = sets or rz.load_sets(workdir)
    files = {k: open(os.path.join(workdir, k), "a") for k in
             ("fort.11", "fort.12", "fort.13", "fort.14", "control", "fort.5001", "fort.5011", "fort.5002",
              "fort.5012")}
    acc = dict(action=0.0, vel2=0.0, condmod=0.0, susc=0.0, cond=np.zeros(3), ancgm=0.0, yav=0.0, yyav=0.0,
               naccp=0, ipbp=0, itot=0, it_pf=0, it_h=0, it_g=0, it_f=0)
    gaction = paction = action = 0.0
    pc, linfo = None, None
    if learned:
        from . import learn
        pc = learn.Preconditioner(lat, seed=int(learned.get("seed", 1)))
        pc.collect(int(learned["pairs"]))
        linfo = dict(pairs=0, losses=[], trained_after=None)
    try:
        for isweep in range(1, iter2 + 1):
            t = trajectory(lat, chi, rng, isweep, p, sets, measure_draw=False)
            if isweep == 1:
                gaction = t["hg0"] / kv
            action, paction = t["S0"] / kv, t["hp0"] / kv
            y = np.exp(t["dH"])
            acc["yav"] += y
            acc["yyav"] += y * y
            if t["accepted"]:
                acc["naccp"] += 1
                action, gaction, paction = t["S1"] / kv, t["hg1"] / kv, t["hp1"] / kv
            files["control"].write(LD(t["dH"], t["dS"], int(t["accepted"])) + "\n")
            files["fort.11"].write(LD(isweep, gaction, paction) + "\n")
            acc["action"] += action
            acc["vel2"] += t["vel2"]
            acc["itot"] += t["md_steps"]
            acc["it_pf"] += t["it_pf"]
            acc["it_h"] += t["it_h0"] + t["it_h1"]
            acc["it_g"] += t["it_guidance"]
            acc["it_f"] += t["it_force"]
            if isweep % iprint == 0:
                cond, susc, condmod, susctot = aux_observables(chi)
                acc["condmod"] += condmod
                acc["susc"] += susctot
                acc["cond"] += cond
                files["fort.12"].write(LD(*cond) + "\n")
                files["fort.13"].write(LD(*susc) + "\n")
                files["fort.14"].write(LD(condmod, susctot) + "\n")
                ittt0 = int(lat.Lt * rng.rano(0)) + 1
                form = "learned" if linfo and linfo["trained_after"] is not None else fermion_precision
                try:
                    c1, c2, its = fm.fermion(lat, chi, rescgm, ittt0, p["hpin"], p["imass"], form)
                except NoConvergence as e:   # congrad at its cap: the message, then the last iterates
                    if getattr(e, "result", None) is None:
                        raise
                    outfile.congrad_capped(os.path.join(workdir, "output"), e)
                    c1, c2, its = e.result
                if linfo and linfo["trained_after"] is None and pc.num_collected() >= int(learned["pairs"]):
                    linfo["pairs"] = pc.num_collected()
                    linfo["losses"] = pc.train_collected(int(learned.get("epochs", 1)), float(learned.get("lr", 1e-3)),
                                                         p["hpin"], p["imass"]).tolist()
                    linfo["trained_after"] = isweep
                    pc.collect(0)
                for it in range(lat.Lt):
                    for f, (u1, u2) in enumerate(((5001, 5011), (5002, 5012))):
                        files["fort.%d" % u1].write(LD(it, c1[f, it].real, c1[f, it].imag) + "\n")
                        files["fort.%d" % u2].write(LD(it, c2[f, it].real, c2[f, it].imag) + "\n")
                acc["ancgm"] += its / 10.0
                acc["ipbp"] += 1
            if isweep % icheck == 0:
                checkpoint.write_con(con, chi, rng.seed())
    finally:
        for f in files.values():
            f.close()
    n = iter2
    ip = max(acc["ipbp"], 1)
    yav = acc["yav"] / n
    out = dict(trajectories=n, acceptances=acc["naccp"], av_traj_length=p["dt"] * acc["itot"] / n,
               exp_mdH=yav, exp_mdH_err=float(np.sqrt(max(acc["yyav"] / n - yav * yav, 0.0) / max(n - 1, 1))),
               it_pseudofermion=acc["it_pf"] / n, it_acceptance=acc["it_h"] / (2 * n),
               it_measurement=acc["ancgm"] / ip, condmod=acc["condmod"] / ip, susc=acc["susc"] / ip,
               cond=(acc["cond"] / ip).tolist(), vel2=acc["vel2"] / n, action=acc["action"] / n)
    # psibarpsi: the program never sets pbp (its estimator call is commented
    # out, rhmc/cheisdwf.f:392); its build prints NaN for the average, and so
    # does this file
    outfile.averages(os.path.join(workdir, "output"),
                     dict(naccp=acc["naccp"], itot=acc["itot"], yav=acc["yav"], yyav=acc["yyav"],
                          guidance=acc["it_g"], force=acc["it_f"], acceptance=acc["it_h"],
                          pseudofermion=acc["it_pf"], measurement=acc["ancgm"], ipbp=acc["ipbp"],
                          condmod=acc["condmod"], susc=acc["susc"], cond=acc["cond"], pbp=float("nan"), vel2=acc["vel2"],
                          action=acc["action"]),
                     n, p["dt"], seed_end=rng.seed() if iwrite == 1 else None, rhmc=True)
    if linfo is not None:
        out["learned"] = linfo
    if iwrite == 1:
        checkpoint.write_con(con, chi, rng.seed())
    lat.close()
    return out
