"""The `output` file of the drivers in the reference programs' own formats:
the start seed, the run header (format 9001), the stopping residuals (9004)
and the generator's seed (rhmc/cheisdwf.f:126-182, mh/cheisdwf_hmc.f:84-117),
then after the run the averages (9022) and the final seed
(rhmc/cheisdwf.f:415-462, mh/cheisdwf_hmc.f:344-385)."""
import numpy as np

from . import fortfmt as F

RESPBP = 1e-6   # the estimator's residual, printed by both programs (rhmc/cheisdwf.f:32)


def _seed(v):
    return " seed:  " + F.list_real(v)


def header(path, seed_in, seed_now, L, Lt, Ls, p, iter2, rhmc=True):
    """Start the file (as the programs open it: overwritten)."""
    lines = [_seed(seed_in),
             " ksize=" + F.i(L, 3) + " ksizet=" + F.i(Lt, 3),
             " kthird=" + F.i(Ls, 3),
             " time step: dt=" + F.f(p["dt"], 6, 4) + " trajectory length=" + F.f(p["iterl"] * p["dt"], 9, 6),
             " # trajectories=" + F.i(iter2, 6) + " beta=" + F.f(p["beta"], 9, 6),
             " am3=" + F.f(p.get("am3", 1.0), 6, 4) + " am=" + F.f(p["am"], 6, 4)
             + (" hpin=" + F.f(p["hpin"], 7, 4) if rhmc else ""),
             " imass=" + F.i(p["imass"], 2),
             " Stopping residuals: guidance: " + F.e(p["rescgg"], 9, 4) + " acceptance: " + F.e(p["rescga"], 9, 4)
             + " estimator: " + F.e(RESPBP, 9, 4),
             _seed(seed_now)]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


CONGRAD_CAP = " # iterations of congrad exceeds niterc"   # format 1000 of both programs' congrad


def congrad_capped(path, err):
    """Append the programs' congrad message (rhmc/cheisdwf.f:1151-1153,
    mh/cheisdwf_hmc.f:705-707) once for each solve that a NoConvergence
    from the library reports at its iteration cap (the message's leading
    count)."""
    msg = getattr(err, "msg", "") or ""
    head = msg.split()[0] if msg.split() else ""
    n = int(head) if head.isdigit() else 1
    with open(path, "a") as fh:
        fh.write((CONGRAD_CAP + "\n") * n)


def _div(a, b):
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))


def averages(path, a, iter2, dt, seed_end=None, rhmc=True):
    """Append the averages of the run.  a: sums over the run -- naccp, itot
    (MD steps), yav, yyav (sums of exp(-dH) and its square), guidance and
    force (the force's two solves' iterations; the sibling: its CG),
    acceptance (the Hamiltonians'), pseudofermion (the heatbath's),
    measurement (the measurements' aviter), ipbp (measurements), condmod,
    susc, cond (3), pbp, vel2, action -- divided here as the programs do."""
    n = iter2
    yav = _div(a["yav"], n)
    with np.errstate(divide="ignore", invalid="ignore"):
        yyav = float(np.sqrt(np.float64(_div(a["yyav"], n) - yav * yav) / np.float64(n - 1)))
    ip = a["ipbp"]
    cond = [_div(c, ip) for c in a["cond"]]
    lines = [" averages for last " + F.i(n, 6) + " trajectories",
             "  # of acceptances: " + F.i(a["naccp"], 6) + " average trajectory length= "
             + F.f(_div(dt * a["itot"], n), 8, 3),
             "  <exp-dH>=" + F.e(yav, 11, 4) + " +/-" + F.e(yyav, 10, 3),
             "  av. # %s itr. guidance: " % ("QMR" if rhmc else "congrad") + F.f(_div(a["guidance"], a["itot"]), 9, 3)
             + "; acceptance: " + F.f(_div(a["acceptance"], 2 * n), 9, 3)]
    if rhmc:
        lines += ["  force: " + F.f(_div(a["force"], a["itot"]), 9, 3) + "; pseudofermion: "
                  + F.f(_div(a["pseudofermion"], n), 9, 3), ""]
    lines += ["  measurement: " + F.f(_div(a["measurement"], ip), 9, 3), "",
              "  condmod=" + F.e(_div(a["condmod"], ip), 11, 3) + " susc=" + F.e(_div(a["susc"], ip), 11, 3),
              " cond components=" + "".join(F.e(c, 11, 3) for c in cond),
              "  psibarpsi=" + F.e(_div(a["pbp"], ip), 11, 3),
              "  mean square velocity=" + F.e(_div(a["vel2"], n), 10, 3) + "; action per site="
              + F.e(_div(a["action"], n), 10, 3),
              "", "", "", ""]
    if seed_end is not None:
        lines.append(_seed(seed_end))
    with open(path, "a") as fh:
        fh.write("\n".join(lines) + "\n")
