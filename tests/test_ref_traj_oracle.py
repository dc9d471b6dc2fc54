"""Pins the oracle's trajectory drivers (orc_rhmc in oracle/chs_oracle.c and
oracle.hmc_trajectory) to complete runs of the reference's own programs
(tests/golden/ref_traj_*.npz, tests/golden/make_ref_traj.py) -- CPU only.

1. The random stream's consumption, in closed form: the seed each program
   left in its final con equals seed0 * M^N mod 2^48 with N the number of
   draws the drivers are written to take -- 196 to start the shuffle
   (rano's 97 discarded draws, 97 table entries, its pointer, and the refill
   of the first call; rhmc/cheisdwf.f:127-129,1722-1751), per trajectory
   2*kvol per gauss0 call (radii, then angles) for each of the 2*4*Ls
   (flavour, Dirac, s) sources, 2 gaussp calls of 2*kvol, one draw per MD
   step, and the Metropolis draw (every trajectory in dwf3d, only when
   dH < 0 in the HMC sibling), plus one draw per measurement (every fifth
   trajectory, dwf3d's fermion).  The MD-step total comes from the
   program's own output.
2. The oracle's drivers run the same trajectories from the same con files
   and reproduce the programs' dH, decisions and actions (and, for the HMC
   sibling, all five trajectories: final chi and seed).  At L = 12, Ls = 8
   the oracle needs about 95 s for dwf3d's first trajectory (11 MD steps; dH
   -0.5748023034 against the program's -0.5748022946, same decision) and 20 s
   for the sibling's five, on 8 cores (with other OpenMP jobs on the same
   cores it takes many times longer).  On the GPU the product's drivers run
   every trajectory of every fixture (tests/test_gpu_ref_traj.py).

dH tolerances: the sibling's CG force at the program's res = 1e-6 moves by
3.8e-8 relative under a 1e-15 relative change of its source (finite-
precision CG amplifies rounding), which over a trajectory moves dH by ~1e-5
between two correct implementations (measured 2.1e-5); the multi-shift
forces of dwf3d are far less sensitive (1e-7)."""
import os

import numpy as np
import pytest

import oracle as orc
from chiralheisenberg_amd import remez

HERE = os.path.dirname(os.path.abspath(__file__))
MUL, MOD = 2651554 * 2 ** 24 + 15184245, 2 ** 48
KVOL, LS = 12 ** 3, 8


def fixture(case):
    path = os.path.join(HERE, "golden", "ref_traj_%s.npz" % case)
    if not os.path.exists(path):
        pytest.skip("fixture %s not written" % path)
    g = np.load(path)
    par = {k[4:]: g[k].item() for k in g.files if k.startswith("par_")}
    for k in ("imass", "iterl", "iter2"):
        par[k] = int(par[k])
    return g, par


@pytest.mark.parametrize("case", ["rhmc_imass1", "rhmc_imass3", "hmc"])
def test_random_stream_consumption(case):
    g, par = fixture(case)
    n = par["iter2"]
    per_traj = 2 * 4 * LS * 2 * KVOL + 2 * 2 * KVOL
    if case == "hmc":
        metropolis = int((g["control"][:, 0] < 0).sum())
        measurements = 0
    else:
        metropolis = n
        measurements = n // 5
    N = 196 + n * per_traj + int(g["itot"]) + metropolis + measurements
    assert int(g["seed0"]) * pow(MUL, N, MOD) % MOD == int(g["seed1"])


def test_oracle_rhmc_first_trajectory_vs_reference_program():
    """orc_rhmc against program dwf3d's first trajectory (imass = 1, hpin 0.5,
    dt 0.125, the reference's thermalised rhmc/_con)."""
    g, par = fixture("rhmc_imass1")
    lat = orc.Lattice(12, LS)
    chi = np.ascontiguousarray(g["chi0"])
    rng = orc.Rng(float(g["seed0"]))
    rng.rano(0)                                   # y = rano(yran, idum) with idum = -1
    p = dict(par, itermax=1000, iprint=5, rescga=1e-9, rescgg=1e-6)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    out = orc.rhmc(lat, p, rng, chi, 1, 1, sets)[0]
    ctl = g["control"][0]
    print("dH oracle %.10f reference %.10f, steps %d" % (out["dH"], ctl[0], out["md_steps"]), flush=True)
    assert abs(out["dH"] - ctl[0]) <= 1e-7
    assert out["accepted"] == int(ctl[2])
    assert abs((out["S0"] - out["S1"]) - ctl[1]) <= 1e-6
    # fort.11 after the first trajectory: paction = hp/kvol of the accepted state
    hp = (out["H1"] - out["S1"]) if out["accepted"] else (out["H0"] - out["S0"])
    assert abs(hp / KVOL - g["fort_11"][0, 2]) <= 1e-9 * abs(g["fort_11"][0, 2])


def test_oracle_hmc_run_vs_reference_program():
    """oracle.hmc_trajectory over the whole run of the sibling's program: every
    dH and dS, the actions of fort.11, the final chi and seed."""
    g, par = fixture("hmc")
    lat = orc.Lattice(12, LS)
    chi = np.array(g["chi0"])
    rng = orc.Rng(float(g["seed0"]))
    rng.rano(0)
    p = dict(par, itermax=1000, rescga=1e-9, rescgg=1e-6)
    steps, pact = 0, None
    for t in range(par["iter2"]):
        o = orc.hmc_trajectory(lat, chi, rng, p)
        steps += o["md_steps"]
        ref = g["control"][t]
        print("traj %d dH oracle %.10f reference %.10f" % (t + 1, o["dH"], ref[0]), flush=True)
        assert abs(o["dH"] - ref[0]) <= 1e-4
        assert abs(o["dS"] - ref[1]) <= 1e-3
        # the sibling sets the actions from H0 on its first trajectory and
        # from H1 on every accepted one (mh/cheisdwf_hmc.f:194-198,262-275)
        if o["accepted"]:
            pact = o["hp1"] / KVOL
        elif t == 0:
            pact = o["hp0"] / KVOL
        assert abs(pact - g["fort_11"][t, 2]) <= 1e-7 * abs(g["fort_11"][t, 2])
    assert steps == int(g["itot"])
    assert rng.get() == float(g["seed1"])
    assert np.abs(chi - g["chi1"]).max() <= 1e-4
