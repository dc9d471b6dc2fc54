"""Writes tests/golden/oracle_traj_l12_ls24.npz: 20 seeded RHMC trajectories
(program dwf3d, rhmc/cheisdwf.f:207-413) of the CPU oracle at L=Lt=12,
Ls=24, beta=0.475, imass=1, hpin=0, dt=0.05, mean length 20 steps (iterl=20),
the reference's production tolerances (rescga=1e-9, rescgg=1e-6), chi drawn
N(0, 1/beta) from numpy seed 12, generator seed 693764523745.  The oracle
needs about a minute of CPU per molecular-dynamics step here (the twenty
trajectories hold 267 steps), far too long for a test, so its outputs are
kept as a fixture for the GPU test, which runs the same trajectories.  The
file is rewritten after every trajectory and holds the first n of them
(chi_final and seed_final after trajectory n); the test compares those n.

    python tests/golden/make_oracle_traj_l12_ls24.py [threads]

With --tight it writes tests/golden/oracle_traj_l12_ls24_tight.npz instead:
two trajectories from the same start with the solver tolerances tightened to
rescga = 1e-12 (hamilton, heatbath) and rescgg = 1e-11 (force) and a mean
length of 4 steps (iterl = 4), where H, dH and the forces are fixed to ~1e-12
by the mathematics rather than by where two solvers happen to stop.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import oracle as orc  # noqa: E402
from chiralheisenberg_amd import remez  # noqa: E402

L, LS, NTRAJ = 12, 24, 20
SEED = 693764523745.0
PAR = dict(dt=0.05, beta=0.475, hpin=0.0, imass=1, iterl=20, itermax=5000, iprint=5, rescga=1e-9, rescgg=1e-6,
           am=0.0)


PAR_TIGHT = dict(PAR, rescga=1e-12, rescgg=1e-11, iterl=4)
NTRAJ_TIGHT = 2


def chi0():
    return np.random.RandomState(12).randn(3, L * L * L) / np.sqrt(PAR["beta"])


def main():
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    tight = "--tight" in sys.argv
    if args:
        orc.set_threads(int(args[0]))
    par, ntraj = (PAR_TIGHT, NTRAJ_TIGHT) if tight else (PAR, NTRAJ)
    name = "oracle_traj_l12_ls24_tight.npz" if tight else "oracle_traj_l12_ls24.npz"
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    olat = orc.Lattice(L, LS)
    chi = chi0()
    rng = orc.Rng(SEED)
    out = []
    t0 = time.time()
    for i in range(ntraj):
        out += orc.rhmc(olat, par, rng, chi, i + 1, 1, sets)
        print("trajectory %d: steps %d accepted %d dH %.6e (%.0f s)" % (i + 1, out[-1]["md_steps"],
                                                                       out[-1]["accepted"], out[-1]["dH"],
                                                                       time.time() - t0), flush=True)
        keys = ("H0", "H1", "S0", "S1", "dH", "accepted", "md_steps")
        path = os.path.join(HERE, name)
        np.savez_compressed(path + ".tmp.npz", **{k: np.array([o[k] for o in out]) for k in keys},
                            chi_final=chi, seed_final=np.array(rng.get()))
        os.replace(path + ".tmp.npz", path)


if __name__ == "__main__":
    main()
