"""The reference's HMC sibling (program in mh/cheisdwf_hmc.f) with every
fermion routine on the GPU: two flavours through Phi = M^dagger R, a
congrad-based force and Hamiltonian sharing one chronological initial guess.

Per trajectory, in the reference's order of random-number consumption
(mh/cheisdwf_hmc.f:140-330):
  1. R complex Gaussian (gauss0 per s, Dirac, flavour as in the RHMC), Phi =
     M^dagger R, the guess X0 = Phi;
  2. momenta from two gaussp calls;
  3. H0 = hamilton (X0 <- X1), force with the guess update, leapfrog of random
     length (one yrun draw per step, end with probability 1/iterl);
  4. H1 = hamilton; accepted if dH = H0 - H1 >= 0, else if rano(yrun) <= exp(dH).
The operator is M with hpin = 0 (the sibling has no pinning term); all reals
FP64.
"""
import os

import numpy as np

from . import fermion as fm
from . import outfile
from ._lib import NoConvergence
from .fortfmt import list_directed as LD
from .rhmc import RandomStream, aux_observables, heatbath_source

DEFAULTS = dict(dt=0.06, beta=0.45, am=0.0, imass=1, iterl=16, itermax=1000, rescga=1e-9, rescgg=1e-6)


def _congrad(call, output):
    """The sibling's congrad keeps its last iterate at the iteration cap
    and writes its message to output (mh/cheisdwf_hmc.f:705-707): run
    call() and, where the library reports the cap, go on with the results
    it kept, the message appended to output when given."""
    try:
        return call()
    except NoConvergence as e:
        if getattr(e, "result", None) is None:
            raise
        if output:
            outfile.congrad_capped(output, e)
        return e.result


def leapfrog(lat, Phi, x0, chit, pp, par, stop, output=None):
    """The sibling's molecular-dynamics evolution (mh/cheisdwf_hmc.f:140-330),
    in place on chit, pp and the chronological guess x0: a half momentum
    step, then steps of chit += dt pp and the force; after each step's force
    stop(step) decides the end, and the last momentum step is a half step (a
    full one when itermax ends the loop).  The sibling's stop is one
    rano(yrun) < 1/iterl draw per step.  Returns (steps, force iterations)."""
    dt = par["dt"]
    def force():
        return _congrad(lambda: fm.hmc_force(lat, Phi, x0, par["rescgg"], chit, par["beta"], par["imass"]), output)
    F, its = force()
    pp -= 0.5 * dt * F
    for step in range(1, par["itermax"] + 1):
        chit += dt * pp
        F, it = force()
        its += it
        if stop(step):
            pp -= 0.5 * dt * F
            return step, its
        pp -= dt * F
    return par["itermax"], its


def trajectory(lat, chi, rng, isweep, par=None, output=None):
    """One HMC trajectory; chi (float64 (3, kvol)) is replaced in place when
    the proposal is accepted.  Returns a dict of H0, H1, dH, accepted, the
    Hamiltonian parts, the mean square velocity and iteration counts.
    output: the file that receives the congrad message of a solve at its
    iteration cap (run() passes its output file)."""
    p = dict(DEFAULTS)
    if par:
        p.update(par)
    kv = lat.kvol
    chit = chi.copy()
    R = heatbath_source(lat, rng)
    Phi = fm.dslashd(lat, R, chit, p["am"], 0.0, p["imass"])
    x0 = Phi.copy()
    pp = np.empty((3, kv))
    ps = rng.gaussp(kv)
    pp[0], pp[1] = ps[0], ps[1]
    ps = rng.gaussp(kv)
    pp[2] = ps[0]
    (H0, hg0, hp0, S0), it_h0 = _congrad(lambda: fm.hmc_hamilton(lat, Phi, x0, p["rescga"], chit, pp, p["beta"],
                                                                  p["imass"]), output)
    proby = 1.0 / p["iterl"]
    steps, its = leapfrog(lat, Phi, x0, chit, pp, p, lambda step: rng.rano(1) < proby, output)
    (H1, hg1, hp1, S1), it_h1 = _congrad(lambda: fm.hmc_hamilton(lat, Phi, x0, p["rescga"], chit, pp, p["beta"],
                                                                  p["imass"]), output)
    dH = H0 - H1
    accepted = True
    if dH < 0.0:
        accepted = not (rng.rano(1) > np.exp(dH))
    if accepted:
        chi[...] = chit
    return dict(H0=H0, H1=H1, S0=S0, S1=S1, dH=dH, dS=S0 - S1, accepted=int(accepted), md_steps=steps,
                hg0=hg0, hp0=hp0, hg1=hg1, hp1=hp1, vel2=float(np.sum(pp * pp) / kv), it_force=its,
                it_h0=it_h0, it_h1=it_h1)


def read_midout(path):
    """midout of the sibling: dt, beta, am3, am, imass, iterl, iter2
    (mh/cheisdwf_hmc.f:78)."""
    v = open(path).readline().split()
    dt, beta, am3, am = (float(t) for t in v[:4])
    imass, iterl, iter2 = (int(t) for t in v[4:7])
    return dict(dt=dt, beta=beta, am3=am3, am=am, imass=imass, iterl=iterl, iter2=iter2)


def run(workdir, L, Ls, Lt=None, par=None, iter2=None, iprint=5, icheck=100, iread=1, iwrite=1, iseed=0,
        force_precision="fp64"):
    """The sibling's main program: midout (unless par) and con in; control
    (dH, dS), fort.11 (isweep, gauge and momentum action per site), fort.12/13/14
    every iprint trajectories, con every icheck trajectories and at the end,
    output (header, seeds and averages in the sibling's own formats); the
    data files hold list-directed records as the sibling writes them and,
    like output, start afresh on every run (dwf3d's rhmc.run appends).
    force_precision: the force's CG, "fp64" (the
    reference CGNR) or the reliable-update CG in "fp32" or "fp16" form;
    hamilton stays FP64.  Returns the averages."""
    from . import checkpoint
    p = dict(DEFAULTS)
    mid = os.path.join(workdir, "midout")
    if par is None and os.path.exists(mid):
        par = read_midout(mid)
    p.update(par or {})
    iter2 = int(iter2 if iter2 is not None else p.get("iter2", 1))
    lat = fm.Lattice(L, Ls, Lt, am3=p.get("am3", 1.0))
    fm.set_force_precision(lat, force_precision)
    kv = lat.kvol
    con = os.path.join(workdir, "con")
    if iread == 1:
        chi, seed = checkpoint.sread(con, kv, p["beta"])
    else:
        chi, seed = np.full((3, kv), 0.1 / p["beta"]), checkpoint.DEFAULT_SEED
    if iseed != 0:
        seed = 4587658465895.0      # mh/cheisdwf_hmc.f:82
    rng = RandomStream(seed)
    rng.rano(0)                      # y = rano(yran, idum) with idum = -1
    outfile.header(os.path.join(workdir, "output"), seed, rng.seed(), L, lat.Lt, Ls, p, iter2, rhmc=False)
    # the sibling opens control with status='unknown' and units 11-14
    # implicitly, without position='append' (mh/cheisdwf_hmc.f:74): its flang
    # build starts these files afresh on every run (dwf3d appends to its own)
    files = {k: open(os.path.join(workdir, k), "w") for k in ("fort.11", "fort.12", "fort.13", "fort.14", "control")}
    acc = dict(action=0.0, vel2=0.0, condmod=0.0, susc=0.0, cond=np.zeros(3), yav=0.0, yyav=0.0, naccp=0, ipbp=0,
               itot=0, it_f=0, it_h=0)
    gaction = paction = action = 0.0
    try:
        for isweep in range(1, iter2 + 1):
            t = trajectory(lat, chi, rng, isweep, p, os.path.join(workdir, "output"))
            if isweep == 1:
                action, gaction, paction = t["S0"] / kv, t["hg0"] / kv, t["hp0"] / kv
            y = np.exp(t["dH"])
            acc["yav"] += y
            acc["yyav"] += y * y
            files["control"].write(LD(t["dH"], t["dS"]) + "\n")
            if t["accepted"]:
                acc["naccp"] += 1
                action, gaction, paction = t["S1"] / kv, t["hg1"] / kv, t["hp1"] / kv
            files["fort.11"].write(LD(isweep, gaction, paction) + "\n")
            acc["action"] += action
            acc["vel2"] += t["vel2"]
            acc["itot"] += t["md_steps"]
            acc["it_f"] += t["it_force"]
            acc["it_h"] += t["it_h0"] + t["it_h1"]
            if isweep % iprint == 0:
                cond, susc, condmod, susctot = aux_observables(chi)
                acc["condmod"] += condmod
                acc["susc"] += susctot
                acc["cond"] += cond
                acc["ipbp"] += 1
                files["fort.12"].write(LD(*cond) + "\n")
                files["fort.13"].write(LD(*susc) + "\n")
                files["fort.14"].write(LD(condmod, susctot) + "\n")
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
               it_force=acc["it_f"] / max(acc["itot"], 1), it_acceptance=acc["it_h"] / (2 * n),
               condmod=acc["condmod"] / ip, susc=acc["susc"] / ip, cond=(acc["cond"] / ip).tolist(),
               vel2=acc["vel2"] / n, action=acc["action"] / n)
    # measurement and psibarpsi: the sibling never sets ancgm and pbp (its
    # estimator call is commented out, mh/cheisdwf_hmc.f:321); its build
    # prints NaN for their averages, and so does this file
    nan = float("nan")
    outfile.averages(os.path.join(workdir, "output"),
                     dict(naccp=acc["naccp"], itot=acc["itot"], yav=acc["yav"], yyav=acc["yyav"],
                          guidance=acc["it_f"], acceptance=acc["it_h"], measurement=nan, ipbp=acc["ipbp"],
                          condmod=acc["condmod"], susc=acc["susc"], cond=acc["cond"], pbp=nan, vel2=acc["vel2"],
                          action=acc["action"]),
                     n, p["dt"], seed_end=rng.seed() if iwrite == 1 else None, rhmc=False)
    if iwrite == 1:
        checkpoint.write_con(con, chi, rng.seed())
    lat.close()
    return out
