"""Twenty seeded RHMC trajectories at L=12, Ls=24 on the GPU against the CPU
oracle's outputs for the same inputs (a fixture: the oracle needs about an
hour for them)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import remez, rhmc  # noqa: E402


def test_trajectories_l12_ls24_match_oracle_fixture():
    """Seeded trajectories at L=12, Ls=24 (beta 0.475, dt 0.05, mean length
    20 steps, the reference's production tolerances 1e-9 / 1e-6) with every
    fermion routine on the GPU: the same trajectory lengths and accept/reject
    decisions as the CPU oracle's loop, H0 and H1 within 1e-9 relative, dH
    within 5e-8 (measured 2.6e-8), the chi and generator state after the last
    one.  The oracle
    takes about a minute per molecular-dynamics step here, so its outputs are
    the fixture tests/golden/oracle_traj_l12_ls24.npz (the first n of twenty
    trajectories), written by tests/golden/make_oracle_traj_l12_ls24.py with
    the same inputs."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_oracle_traj_l12_ls24 as gen
    ref = np.load(os.path.join(here, "golden", "oracle_traj_l12_ls24.npz"))
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    lat = fm.Lattice(gen.L, gen.LS)
    chi = gen.chi0()
    rng = rhmc.RandomStream(gen.SEED)
    n = len(ref["dH"])
    gpu = [rhmc.trajectory(lat, chi, rng, isweep=i + 1, par=gen.PAR, sets=sets) for i in range(n)]
    lat.close()
    dh = max(abs(g["dH"] - ref["dH"][i]) for i, g in enumerate(gpu))
    hh = max(abs(g[k] - ref[k][i]) / abs(ref[k][i]) for i, g in enumerate(gpu) for k in ("H0", "H1"))
    print("L=12, Ls=24, %d trajectories: max |dH - dH_oracle| = %.2e, max relative H error %.2e" % (n, dh, hh),
          flush=True)
    assert [g["md_steps"] for g in gpu] == list(ref["md_steps"])
    assert [int(g["accepted"]) for g in gpu] == [int(a) for a in ref["accepted"]]
    for i, g in enumerate(gpu):
        for k in ("H0", "H1"):
            assert abs(g[k] - ref[k][i]) <= 1e-9 * abs(ref[k][i]), (i, k, g[k], ref[k][i])
        assert abs(g["dH"] - ref["dH"][i]) <= 5e-8, (i, g["dH"], ref["dH"][i])
    assert np.abs(chi - ref["chi_final"]).max() <= 1e-6 * np.abs(ref["chi_final"]).max()
    assert rng.seed() == float(ref["seed_final"])


def test_trajectories_l12_ls24_tight_tolerances():
    """The same start at solver tolerances where H and dH are fixed by the
    mathematics rather than by where a solver stops: rescga = 1e-12 (heatbath,
    hamilton), rescgg = 1e-11 (force), mean length 4 steps.  H0, H1 and dH
    against the oracle's (fixture tests/golden/oracle_traj_l12_ls24_tight.npz,
    written by make_oracle_traj_l12_ls24.py --tight): H to 1e-12 relative and
    dH to 1e-9 absolute, i.e. 3e-15 of H ~ 3.4e5 (measured 1.75e-10 on the
    first trajectory: the FP64 rounding of sums over 3e5 terms of size 1, in
    different orders on the GPU and in the oracle).  At the production
    tolerances above the same comparison gives 2.6e-8."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "golden", "oracle_traj_l12_ls24_tight.npz")
    if not os.path.exists(path):
        pytest.skip("tight-tolerance fixture not written")
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_oracle_traj_l12_ls24 as gen
    ref = np.load(path)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    lat = fm.Lattice(gen.L, gen.LS)
    chi = gen.chi0()
    rng = rhmc.RandomStream(gen.SEED)
    n = len(ref["dH"])
    gpu = [rhmc.trajectory(lat, chi, rng, isweep=i + 1, par=gen.PAR_TIGHT, sets=sets) for i in range(n)]
    lat.close()
    for i, g in enumerate(gpu):
        print("tight trajectory %d: %d steps, H0 %.14e (oracle %.14e), dH %.12e (oracle %.12e)" % (
            i + 1, g["md_steps"], g["H0"], ref["H0"][i], g["dH"], ref["dH"][i]), flush=True)
    assert [g["md_steps"] for g in gpu] == list(ref["md_steps"])
    assert [int(g["accepted"]) for g in gpu] == [int(a) for a in ref["accepted"]]
    for i, g in enumerate(gpu):
        for k in ("H0", "H1"):
            assert abs(g[k] - ref[k][i]) <= 1e-12 * abs(ref[k][i]), (i, k, g[k], ref[k][i])
        assert abs(g["dH"] - ref["dH"][i]) <= 1e-9, (i, g["dH"], ref["dH"][i])
    assert np.abs(chi - ref["chi_final"]).max() <= 1e-10 * np.abs(ref["chi_final"]).max()
    assert rng.seed() == float(ref["seed_final"])


def test_trajectories_l12_ls24_fp32_force_same_decisions():
    """The same trajectories with the force's multi-shift solves in FP32
    storage (Lanczos scalars, shift sums and derivs in FP64; hamilton and the
    Metropolis test FP64): the same lengths and accept/reject decisions as the
    all-FP64 CPU oracle, dH within 1e-3 of it."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_oracle_traj_l12_ls24 as gen
    ref = np.load(os.path.join(here, "golden", "oracle_traj_l12_ls24.npz"))
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    lat = fm.Lattice(gen.L, gen.LS)
    fm.set_force_precision(lat, "fp32")
    chi = gen.chi0()
    rng = rhmc.RandomStream(gen.SEED)
    n = len(ref["dH"])
    gpu = [rhmc.trajectory(lat, chi, rng, isweep=i + 1, par=gen.PAR, sets=sets) for i in range(n)]
    lat.close()
    dh = [abs(g["dH"] - ref["dH"][i]) for i, g in enumerate(gpu)]
    print("FP32 force: max |dH - dH_oracle| = %.2e over %d trajectories" % (max(dh), n))
    assert [g["md_steps"] for g in gpu] == list(ref["md_steps"])
    assert [int(g["accepted"]) for g in gpu] == [int(a) for a in ref["accepted"]]
    assert max(dh) <= 1e-3
