"""Writes tests/golden/ref_traj_*.npz: complete runs of the reference's own
programs, compiled unmodified by `make -C oracle ref` (compile-time sizes
ksize = ksizet = 12, kthird = 8, FP64 kinds):

  rhmc_imass1  program dwf3d of rhmc/cheisdwf.f with the run parameters of
               the reference's own rhmc/midout (dt 0.125, beta 0.45, am3 1,
               hpin 0.5, imass 1, iterl 8), ten trajectories from its
               thermalised rhmc/_con; measurements (fermion) at trajectories
               5 and 10;
  rhmc_imass3  the same program in its imass = 3 production mode with the
               parameters of rhmc/Ls_8/b_0.25_m3/midout (dt 0.11, beta 0.25,
               iterl 9; hpin 0), five trajectories from that directory's
               FP32 _con widened to FP64 as rhmc/sp_to_dp.f does;
  hmc          the HMC sibling's program (mh/cheisdwf_hmc.f) with the
               parameters of mh/midout (dt 0.06, beta 0.45, imass 1,
               iterl 16), five trajectories from mh/con (FP32, widened).

Each program runs in a scratch directory holding con, midout and the Remez
files, as the reference's own batch scripts lay it out (rhmc/scan_in_beta.sh);
the fixture keeps what the program wrote: control (dH, dS and, for dwf3d,
accepted), fort.11 (isweep, gauge and momentum action per site), fort.12/13/14
(cond, susc, condmod and susctot), fort.5001/5011/5002/5012 (correlators),
the totals of `output` (acceptances, mean trajectory length) and its whole text
(output_text: the run's header, residuals, seeds and averages in the
program's own formats), the data files' text as written (text_control,
text_fort_11, ...: the programs' list-directed records) and the final con
(chi and the generator's seed), plus the inputs: the start con's chi and seed
and the run parameters.  Nothing else is read from the programs: their
trajectory lengths follow from the mean length, and the random stream's
consumption from the final seed.

Run in the build container:  python tests/golden/make_ref_traj.py [case ...]
(OpenMP threads per program: CHS_REF_THREADS, default 8).  The programs'
OpenMP reductions combine their partial sums in an order that changes from
run to run, so two runs of a case agree to the rounding level, not bit for
bit (dwf3d, imass 3, two runs on 3 threads: dH within 2.2e-7, chi within
6.8e-9; on 3 against 8 threads alike), with the same decisions, step counts
and seed.
"""
import os
import resource
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"
REFBIN = os.path.join(ROOT, "oracle", "_ref")

CASES = {
    "rhmc_imass1": dict(prog="dwf3d_fp64", con=REF + "/rhmc/_con", remez=REF + "/rhmc",
                        par=dict(dt=0.125, beta=0.45, am3=1.0, am=0.0, hpin=0.5, imass=1, iterl=8, iter2=10)),
    "rhmc_imass3": dict(prog="dwf3d_fp64", con=REF + "/rhmc/Ls_8/b_0.25_m3/_con", remez=REF + "/rhmc",
                        par=dict(dt=0.11, beta=0.25, am3=1.0, am=0.0, hpin=0.0, imass=3, iterl=9, iter2=5)),
    "hmc": dict(prog="hmc_fp64", con=REF + "/mh/con", remez=None,
                par=dict(dt=0.06, beta=0.45, am3=1.0, am=0.0, imass=1, iterl=16, iter2=5)),
}


def _numbers(path):
    if not os.path.exists(path):
        return np.zeros((0,))
    rows = [[float(t) for t in line.split()] for line in open(path) if line.strip()]
    return np.array(rows, dtype=np.float64)


def _output_totals(path):
    """'# of acceptances: N average trajectory length= T' of the output file."""
    naccp, atraj = None, None
    for line in open(path):
        if "# of acceptances:" in line:
            t = line.split()
            naccp = int(t[t.index("acceptances:") + 1])
            atraj = float(t[t.index("length=") + 1])
    return naccp, atraj


def _limits():
    # the programs keep their work arrays on the stack (the reference's own
    # HPC build uses -heap-arrays instead)
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_case(name, threads=None):
    from chiralheisenberg_amd import checkpoint
    c = CASES[name]
    p = c["par"]
    threads = threads or int(os.environ.get("CHS_REF_THREADS", "8"))
    with tempfile.TemporaryDirectory() as d:
        chi0, seed0, prec = checkpoint.read_con(c["con"], kvol=12 ** 3)
        checkpoint.write_con(os.path.join(d, "con"), chi0, seed0, "fp64")
        if c["remez"]:
            for f in ("remez2", "remez4", "remez2g"):
                with open(os.path.join(c["remez"], f)) as src, open(os.path.join(d, f), "w") as dst:
                    dst.write(src.read())
        with open(os.path.join(d, "midout"), "w") as f:
            if c["prog"] == "hmc_fp64":
                f.write("%g %g %g %g %d %d %d\n" % (p["dt"], p["beta"], p["am3"], p["am"], p["imass"], p["iterl"],
                                                    p["iter2"]))
            else:
                f.write("%g %g %g %g %g %d %d %d\n" % (p["dt"], p["beta"], p["am3"], p["am"], p["hpin"], p["imass"],
                                                       p["iterl"], p["iter2"]))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="512M")
        t0 = time.time()
        with open(os.path.join(d, "stdout"), "w") as so:
            subprocess.run([os.path.join(REFBIN, c["prog"])], cwd=d, env=env, stdout=so, stderr=subprocess.STDOUT,
                           preexec_fn=_limits, check=True)
        secs = time.time() - t0
        chi1, seed1, _ = checkpoint.read_con(os.path.join(d, "con"), kvol=12 ** 3, precision="fp64")
        naccp, atraj = _output_totals(os.path.join(d, "output"))
        output_text = np.array(open(os.path.join(d, "output")).read())
        out = dict(chi0=chi0, seed0=seed0, con_precision=prec, chi1=chi1, seed1=seed1, naccp=naccp, atraj=atraj,
                   output_text=output_text,
                   itot=int(round(atraj * p["iter2"] / p["dt"])), seconds=secs, threads=threads,
                   **{"par_" + k: v for k, v in p.items()})
        for f in ("control", "fort.11", "fort.12", "fort.13", "fort.14", "fort.5001", "fort.5011", "fort.5002",
                  "fort.5012"):
            out[f.replace(".", "_")] = _numbers(os.path.join(d, f))
            fp = os.path.join(d, f)
            out["text_" + f.replace(".", "_")] = np.array(open(fp).read() if os.path.exists(fp) else "")
    dst = os.path.join(HERE, "ref_traj_%s.npz" % name)
    np.savez_compressed(dst, **out)
    print("%s: %d trajectories, %d accepted, mean length %.4f (%d MD steps), %.0f s; wrote %s" % (
        name, p["iter2"], naccp, atraj, out["itot"], secs, dst), flush=True)


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(CASES)):
        run_case(n)
