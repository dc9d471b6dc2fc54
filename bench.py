"""Benchmark of the fermion hot path on one MI355X.

Workload (one "step"): one M^dagger M solve of the domain-wall operator of
rhmc/cheisdwf.f at L = Lt = 24, Ls = 32, beta = 0.475, M5 = 1, imass = 1,
antiperiodic in t, to FP64 true relative residual <= 1e-10, with phi_a ~
N(0, 1/beta) and a unit-variance complex Gaussian source drawn by the
reference's own generator (seeded), inputs resident in HBM.

The timed form is the fastest one the library has: CG with the search
direction and its image stored in binary16 (scaled by the residual's rms),
the residual and x increment in FP32, and FP64 reliable updates and true
residual.  The plain FP64 GPU solve and the FP32-storage form of the same
system are timed beside it, and the dslash kernel's HBM rate is measured
with HIP events over back-to-back launches that cycle through 4 buffer pairs
(> the 256 MB Infinity Cache) of random values.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M†M solve time to 1e-10 true resid, FP64 vs mixed+learned; dslash HBM GB/s"
HBM_PEAK = 8.0e12
# The timing loops cycle through enough buffer sets that their inputs exceed
# three times the 256 MB Infinity Cache (a memory-side cache in front of HBM):
# over fewer sets the inputs are re-read from it, not from HBM (measured at
# L=24, Ls=32: the FP64 dslash takes 25.0 us over 4 sets of 56.6 MB and 35.2 us
# over 16; a plain copy 16.9 and 20.9 us).
MALL_BYTES = 256 << 20


def cold_nbuf(input_bytes):
    return int(max(4, min(64, -(-3 * MALL_BYTES // int(input_bytes)))))


def make_inputs(lat, beta, seed=693764523745.0):
    from chiralheisenberg_amd.rhmc import RandomStream, heatbath_source
    rng = RandomStream(seed)
    ps1, ps2 = rng.gaussp(lat.kvol), rng.gaussp(lat.kvol)
    chi = np.stack([ps1[0], ps1[1], ps2[0]]) / np.sqrt(beta)
    b = heatbath_source(lat, rng)
    return chi, b


def measured_traffic(name, L, Ls):
    """HBM bytes per launch of a kernel from the FETCH_SIZE / WRITE_SIZE
    counter runs recorded in profiles/<name>_traffic.json (scaled by a plain
    copy of known size), for the workload they were measured on; None
    otherwise."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % name)
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    if rec.get("L") != L or rec.get("Ls") != Ls:
        return None
    return rec["traffic_bytes"]


def host_cpu():
    """(model name, logical CPUs of the host, CPUs this process may use)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, os.cpu_count(), allowed


def cpu_baseline(L, Ls, beta, gpu_fp64, full=False):
    """The CPU oracle (C/OpenMP, FP64; the reference's algorithm restated) on
    the host cores, threads bound one per core (OMP_PROC_BIND=close,
    OMP_PLACES=cores, as the reference's own batch script runs it), timed
    without extrapolation on the same workload as the GPU:
      - one complete CG solve of M^dagger M x = b at L, Ls (the bench's system,
        from x = 0, stopping rule |r|^2 < kferm res^2 with res = 1e-10, i.e. a
        relative residual of 1e-10 for a unit Gaussian source): `value`;
      - 100 dslash and 100 dslashd applications at L, Ls;
      - with full=True (bench.py --cpu-full; minutes on the CPU) also the
        multi-shift solve at L = 12, Ls = 24 (remez4, res 1e-10) and the
        force there (remez2g, res 1e-6), and one trajectory at L = 8, Ls = 8
        (dt = 0.05, iterl = 20, seeded) by the oracle's driver beside the
        same trajectory by the GPU driver.
    Threads: OMP_NUM_THREADS when the environment sets it (the host's share
    for one GPU), else every CPU this process may use."""
    model, ncpu, allowed = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or allowed)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from chiralheisenberg_amd import remez
    orc.set_threads(threads)
    print("bench: cpu baseline on %d threads (%s; %d logical CPUs, %d allowed)" % (threads, model, ncpu, allowed),
          file=sys.stderr, flush=True)
    olat = orc.Lattice(L, Ls)
    kvol = olat.kvol
    V, C = olat.kferm * 16, 3 * kvol * 8
    chi, b = make_inputs(olat, beta)     # the same system as the GPU's (the oracle lattice has its shape)
    out = {}
    t0 = time.perf_counter()
    for _ in range(100):
        orc.dslash(olat, b, chi, 0.0, 1, False)
    t1 = time.perf_counter()
    for _ in range(100):
        orc.dslash(olat, b, chi, 0.0, 1, True)
    t2 = time.perf_counter()
    out["dslash_ms"] = (t1 - t0) * 10.0
    out["dslashd_ms"] = (t2 - t1) * 10.0
    out["dslash_GB_per_s"] = (2 * V + C) / ((t1 - t0) / 100) / 1e9
    print("bench: cpu dslash %.1f ms, dslashd %.1f ms" % (out["dslash_ms"], out["dslashd_ms"]), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    x, its = orc.congrad(olat, b, 1e-10, chi)
    cg_s = time.perf_counter() - t0
    r = b - orc.dslash(olat, orc.dslash(olat, x, chi, 0.0, 1, False), chi, 0.0, 1, True)
    rel = float(np.linalg.norm(r) / np.linalg.norm(b))
    print("bench: cpu CG solve %.1f s, %d iterations, true resid %.2e" % (cg_s, its, rel), file=sys.stderr,
          flush=True)
    out["cg_solve"] = {"s": cg_s, "iterations": int(its), "true_resid": rel}
    if not full:
        return baseline_line(out, L, Ls, its, rel, threads, model, ncpu, cg_s, gpu_fp64)
    # the multi-shift solve and force at L = 12, Ls = 24
    l2 = orc.Lattice(12, 24)
    rs = np.random.RandomState(2)
    chi2 = rs.randn(3, l2.kvol) / np.sqrt(beta)
    phi2 = (rs.randn(*l2.shape) + 1j * rs.randn(*l2.shape)) / np.sqrt(2)
    a4, ad4, _, _ = remez.load("remez4")
    t0 = time.perf_counter()
    _, it_ms = orc.qmrherm(l2, phi2, 1e-10, a4, ad4, chi2)
    ms_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, it_f = orc.force(l2, phi2, 1e-6, chi2, beta, remez.load("remez2g"))
    f_s = time.perf_counter() - t0
    print("bench: cpu L=12 Ls=24 multi-shift %.1f s (%d its), force %.1f s" % (ms_s, it_ms, f_s), file=sys.stderr,
          flush=True)
    out["l12_ls24"] = {"multishift_remez4_s": ms_s, "multishift_iterations": int(it_ms), "force_s": f_s,
                      "force_iterations": list(it_f) if hasattr(it_f, "__len__") else it_f}
    out["trajectory_l8_ls8"] = trajectory_pair()
    print("bench: L=8 Ls=8 trajectory %s" % json.dumps(out["trajectory_l8_ls8"]), file=sys.stderr, flush=True)
    return baseline_line(out, L, Ls, its, rel, threads, model, ncpu, cg_s, gpu_fp64)


def trajectory_pair(L=8, Ls=8, seed=693764523745.0):
    """One RHMC trajectory at L = Ls = 8 (beta 0.475, dt 0.05,
    iterl 20, the reference's tolerances) from the same seeded phi and
    generator, by the CPU oracle's driver and by the GPU driver: seconds,
    MD steps, dH of each."""
    import oracle as orc
    from chiralheisenberg_amd import fermion as fm, remez, rhmc
    par = dict(dt=0.05, beta=0.475, am=0.0, hpin=0.0, imass=1, iterl=20, itermax=1000, iprint=5, rescga=1e-9,
               rescgg=1e-6)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    chi0 = np.random.RandomState(8).randn(3, L ** 3) / np.sqrt(par["beta"])
    lat = fm.Lattice(L, Ls)
    chi_g = chi0.copy()
    t0 = time.perf_counter()
    g = rhmc.trajectory(lat, chi_g, rhmc.RandomStream(seed), 1, par, sets)
    t_gpu = time.perf_counter() - t0
    lat.close()
    olat = orc.Lattice(L, Ls)
    chi_c = chi0.copy()
    t0 = time.perf_counter()
    c = orc.rhmc(olat, par, orc.Rng(seed), chi_c, 1, 1, sets)[0]
    t_cpu = time.perf_counter() - t0
    return {"L": L, "Ls": Ls, "md_steps": int(c["md_steps"]), "cpu_s": t_cpu, "gpu_s": t_gpu,
            "dH_cpu": float(c["dH"]), "dH_gpu": float(g["dH"]), "same_decision": bool(c["accepted"] == g["accepted"])}


def baseline_line(out, L, Ls, its, rel, threads, model, ncpu, cg_s, gpu_fp64):
    return dict({"value": cg_s * 1e3, "unit": "ms per solve", "cores": threads, "kind": "port",
                 "cpu": model, "host_logical_cpus": ncpu,
                 "sample": "one complete FP64 CG solve of the bench's system at L=%d Ls=%d (%d iterations, "
                           "true relative residual %.2e) by the CPU oracle, %d OpenMP threads bound to cores; "
                           "GPU FP64 form on the same system: %.1f ms" % (L, Ls, its, rel, threads, gpu_fp64),
                 "speedup_gpu_fp64_vs_cpu": cg_s * 1e3 / gpu_fp64}, **out)


BATCH_MIXES = {"full": ((8, 8), (12, 16), (16, 32), (24, 32)), "small": ((8, 8), (12, 8), (12, 16), (16, 16))}


def force_batch_bench(mix="full", npf=4, res=1e-6, seed=13, precision="fp64", beta=0.475):
    """The pseudofermion force (both multi-shift stages, remez2g, the
    reference's force tolerance 1e-6) of npf fields on every lattice of the
    mix: each stage one batch (iflag = 0, then iflag = 1) versus the same two
    solves per field one at a time on the device.  Returns forces/s of both
    and whether the forces are identical."""
    from chiralheisenberg_amd import fermion as fm, batch, remez
    rz = remez.load("remez2g")
    anum, aden, bnum, bden = rz[:4]
    rs = np.random.RandomState(seed)
    fields, lats = [], []
    for L, Ls in BATCH_MIXES[mix]:
        lat = fm.Lattice(L, Ls)
        lats.append(lat)
        chi = rs.randn(3, lat.kvol) / np.sqrt(beta)
        chid = fm.DeviceArray.from_numpy(lat, chi)
        for _ in range(npf):
            phi = fm.spinor_dev(lat, (rs.randn(*lat.shape) + 1j * rs.randn(*lat.shape)) / np.sqrt(2))
            fields.append((lat, phi, chid, chi))

    def single():
        Fs, its = [], []
        for lat, phi, chid, chi in fields:
            X = fm.spinor_dev(lat, None)
            y = fm.spinor_dev(lat, None)
            i1 = fm.qmrherm_prec_dev(lat, phi, res, bnum, bden, chid, X, precision)
            F = fm.DeviceArray.from_numpy(lat, 2.0 * beta * chi)
            i2 = fm.qmrherm_prec_dev(lat, X, res, anum, aden, chid, y, precision, iflag=1, F_dev=F)
            Fs.append(F)
            its.append((i1, i2))
        for lat in lats:
            lat.synchronize()
        return Fs, its

    single()   # warm-up (allocations)
    t0 = time.perf_counter()
    F1, its = single()
    t_single = time.perf_counter() - t0
    b = batch.Batch()
    b.force(fields, beta, res, rz, precision)   # tunes and captures
    t0 = time.perf_counter()
    F2, _, (st1, st2) = b.force(fields, beta, res, rz, precision)
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a.download(np.float64), c.download(np.float64)) for a, c in zip(F1, F2))
    b.close()
    return {"workload": "pseudofermion force x %d: remez2g (%d shifts, both stages) x %d fields x L/Ls %s, %s "
                        "storage, res %g" % (len(fields), len(aden), npf,
                                             ", ".join("%d/%d" % m for m in BATCH_MIXES[mix]), precision.upper(),
                                             res),
            "itercg": its, "one_at_a_time": {"seconds": t_single, "forces_per_s": len(fields) / t_single},
            "batched": {"seconds": t_batch, "forces_per_s": len(fields) / t_batch,
                        "tune": [st1["tune"], st2["tune"]]},
            "speedup": t_single / t_batch, "identical_results": bool(same)}


def batch_bench(mix="full", npf=4, res=1e-9, seed=11, precision="fp64"):
    """Batched multi-shift solves (all 12 remez2g shifts x npf pseudofermion
    fields x the lattices L/Ls of the mix, FP64) in one grouped,
    graph-replayed solve versus the same solves one at a time.  Returns a
    dict with aggregate solves/s and algorithmic HBM GB/s of both."""
    from chiralheisenberg_amd import fermion as fm, batch, remez
    anum, aden = remez.load("remez2g")[:2]
    n = len(aden)
    rs = np.random.RandomState(seed)
    systems, lats = [], []
    for L, Ls in BATCH_MIXES[mix]:
        lat = fm.Lattice(L, Ls)
        lats.append(lat)
        chi = fm.DeviceArray.from_numpy(lat, rs.randn(3, lat.kvol) / np.sqrt(0.475))
        for _ in range(npf):
            phi = fm.spinor_dev(lat, (rs.randn(*lat.shape) + 1j * rs.randn(*lat.shape)) / np.sqrt(2))
            systems.append(batch.System(lat, phi, chi, fm.spinor_dev(lat, None), anum, aden, res))

    def bytes_moved(its):
        # per iteration: two stencils (3V each incl. q / dot), the shift update
        # reading w, q, q_prev and every p_k, writing r and every p_k (V in the
        # storage precision), and reading and writing the FP64 x
        tot = 0
        bc = 16 if precision == "fp64" else 8
        for s, it in zip(systems, its):
            V, C, X = s.lat.kferm * bc, 3 * s.lat.kvol * 8, s.lat.kferm * 16
            tot += it * ((10 + 2 * n) * V + 2 * X + 2 * C)
        return tot

    # one at a time (the single-system path, chunked launches)
    fm.qmrherm_prec_dev(lats[0], systems[0].phi, res, anum, aden, systems[0].chi, systems[0].x, precision)
    lats[0].synchronize()
    t0 = time.perf_counter()
    its1 = []
    for s in systems:
        its1.append(fm.qmrherm_prec_dev(s.lat, s.phi, res, anum, aden, s.chi, s.x, precision))
    for lat in lats:
        lat.synchronize()
    t_single = time.perf_counter() - t0
    ref = [fm.to_host(s.lat, s.x) for s in systems]
    # batched: the first call tunes and captures, the second is timed
    b = batch.Batch()
    st_first = b.qmrherm(systems, None, precision)
    t0 = time.perf_counter()
    st = b.qmrherm(systems, None, precision)
    t_batch = time.perf_counter() - t0
    its = [s.itercg for s in systems]
    same = all(np.array_equal(fm.to_host(s.lat, s.x), r) for s, r in zip(systems, ref)) and its == its1
    B = bytes_moved(its)
    out = {"workload": "qmrherm x %d: remez2g (%d shifts) x %d pseudofermion fields x L/Ls %s, %s storage, res %g" % (
        len(systems), n, npf, ", ".join("%d/%d" % m for m in BATCH_MIXES[mix]), precision.upper(), res),
           "solves": len(systems), "itercg": its,
           "one_at_a_time": {"seconds": t_single, "solves_per_s": len(systems) / t_single,
                             "GB_per_s": B / t_single / 1e9},
           "batched": {"seconds": t_batch, "solves_per_s": len(systems) / t_batch, "GB_per_s": B / t_batch / 1e9,
                       "tune": st["tune"], "first_call_seconds": st_first["seconds"],
                       "graph_replays": st["graph_replays"]},
           "speedup": t_single / t_batch, "identical_results": bool(same)}
    b.close()
    return out


def sweep_bench(tol=1e-10, Ls=32):
    """Every solve form at L in {16, 24}, Ls=32 and beta in {0.45, 0.475,
    0.5}: iterations, reliable updates, device time to the FP64 true relative
    residual tol, and each mixed form's solution error against the FP64 GPU
    solve of the same system (relative 2-norm)."""
    from chiralheisenberg_amd import fermion as fm
    rows = []
    for L in (16, 24):
        lat = fm.Lattice(L, Ls)
        for beta in (0.45, 0.475, 0.5):
            chi, b = make_inputs(lat, beta)
            chi_d = fm.DeviceArray.from_numpy(lat, chi)
            b_d = fm.spinor_dev(lat, b)
            x_d = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
            sol = {}
            for form in ("fp64", "fp32", "fp16", "bf16"):
                try:
                    x_d.upload(np.zeros(lat.shape, np.complex128))
                    fm.solve_mdagm(lat, b_d, x_d, chi_d, tol, form)   # warm-up (allocations, graph)
                    x_d.upload(np.zeros(lat.shape, np.complex128))
                    st = fm.solve_mdagm(lat, b_d, x_d, chi_d, tol, form)
                except fm._lib.NoConvergence as e:
                    rows.append({"L": L, "Ls": Ls, "beta": beta, "form": form, "converged": False, "error": str(e)})
                    print("bench: sweep L=%d beta=%g %s: %s" % (L, beta, form, e), file=sys.stderr, flush=True)
                    continue
                sol[form] = fm.to_host(lat, x_d)
                err = np.linalg.norm(sol[form] - sol["fp64"]) / np.linalg.norm(sol["fp64"])
                rows.append({"L": L, "Ls": Ls, "beta": beta, "form": form, "converged": True,
                             "iterations": st["iterations"],
                             "reliable_updates": st["reliable_updates"], "ms": st["seconds"] * 1e3,
                             "true_resid": st["true_resid"], "solution_error_vs_fp64": float(err)})
                print("bench: sweep L=%d beta=%g %s: %d its, %d updates, %.1f ms, resid %.2e, error %.2e" % (
                    L, beta, form, st["iterations"], st["reliable_updates"], st["seconds"] * 1e3,
                    st["true_resid"], err), file=sys.stderr, flush=True)
            for a in (chi_d, b_d, x_d):
                a.free()
        lat.close()
    return {"workload": "M^dagger M solves to FP64 true relative residual %g, L in {16, 24}, Ls=%d, beta in "
                        "{0.45, 0.475, 0.5}, every storage form" % (tol, Ls), "rows": rows}


def learned_bench(fm, lat, chi_d, b_d, tol, npairs=200, epochs=2, lr=1e-3):
    """The learned preconditioner on the bench system: npairs (Gaussian source,
    solution) pairs from the plain binary16 solver, Adam training on the GPU
    (relative A-norm loss), then flexible CG (FP32, FP64 reliable updates,
    plain-CG fallback) with and without the network, beside the plain
    binary16 CG of the timed form."""
    from chiralheisenberg_amd import learn
    t0 = time.perf_counter()
    src, sol, solver_s = learn.pairs_dev(
        lat, chi_d, npairs, rng_seed=3, tol=tol, precision="fp16",
        progress=lambda i, s: print("bench: learned pairs %d / %d (solver %.1f s)" % (i, npairs, s), file=sys.stderr,
                                    flush=True))
    t1 = time.perf_counter()
    pc = learn.Preconditioner(lat, hidden=16, seed=1)
    losses = pc.train(src, sol, npairs, chi_d, epochs=epochs, lr=lr)
    lat.synchronize()
    t2 = time.perf_counter()
    src.free()
    sol.free()
    res = {}
    for use in (True, False):
        x = fm.spinor_dev(lat, None)
        x.upload(np.zeros(lat.shape, np.complex128))
        res[use] = learn.solve_fcg(lat, b_d, x, chi_d, tol, use_nn=use)
        x.free()
    x = fm.spinor_dev(lat, None)
    x.upload(np.zeros(lat.shape, np.complex128))
    h = fm.solve_mdagm(lat, b_d, x, chi_d, tol, "fp16")
    x.free()
    nn, pl = res[True], res[False]
    return {"pairs": npairs, "pairs_why": "2000 pairs take ~10 min of binary16 solves; the bench "
                                          "keeps to a few minutes",
            "epochs": epochs, "pair_generation_s": t1 - t0, "pair_solver_s": solver_s, "training_s": t2 - t1,
            "loss_first_last": [float(losses[0]), float(losses[-1])],
            "fcg_with_network": {"iterations": nn["iterations"], "ms": nn["seconds"] * 1e3,
                                 "fallbacks": nn["restarts"], "true_resid": nn["true_resid"]},
            "fcg_without_network": {"iterations": pl["iterations"], "ms": pl["seconds"] * 1e3,
                                    "true_resid": pl["true_resid"]},
            "plain_fp16_cg": {"iterations": h["iterations"], "ms": h["seconds"] * 1e3, "true_resid": h["true_resid"]},
            "iteration_gain_vs_fcg": pl["iterations"] / max(nn["iterations"], 1),
            "iteration_gain_vs_plain_fp16": h["iterations"] / max(nn["iterations"], 1),
            "time_gain_vs_plain_fp16": h["seconds"] / max(nn["seconds"], 1e-12)}


def force_bench(fm, lat, chi, beta, res=1e-6, seed=5, forms=("fp64", "fp32"), freezes=(0.0, 0.01)):
    """chs_force (remez2g, the reference's force tolerance 1e-6) with FP64 and
    with FP32-storage multi-shift solves: time and the relative difference of
    the fermion part."""
    from chiralheisenberg_amd import remez
    r2g = remez.load("remez2g")
    rs = np.random.RandomState(seed)
    phi = (rs.randn(*lat.shape) + 1j * rs.randn(*lat.shape)) / np.sqrt(2)
    out = {}
    F = {}
    for prec in forms:
        for f in freezes:
            key = prec if f == 0.0 else "%s_freeze%g" % (prec, f)
            fm.set_force_precision(lat, prec)
            lat.set_shift_freeze(f)
            fm.force(lat, phi, res, chi, beta, r2g)            # warm-up (allocations)
            t0 = time.perf_counter()
            F[key], _, its = fm.force(lat, phi, res, chi, beta, r2g)
            out[key] = {"ms": (time.perf_counter() - t0) * 1e3, "iterations": list(its)}
    fm.set_force_precision(lat, "fp64")
    lat.set_shift_freeze(0.0)
    if "fp64" in F:
        f64 = F["fp64"] - 2 * beta * chi
        for key in F:
            if key != "fp64":
                fk = F[key] - 2 * beta * chi
                out[key]["rel_error"] = float(np.linalg.norm(fk - f64) / np.linalg.norm(f64))
                out[key]["speedup"] = out["fp64"]["ms"] / out[key]["ms"]
    if "fp32" in out and "rel_error" in out["fp32"]:
        out["fp32_rel_error"] = out["fp32"]["rel_error"]
        out["speedup"] = out["fp32"]["speedup"]
    return out


TRAJ_FORMS = {"reference": dict(force_precision="fp64", shift_freeze=0.0),
              "fast": dict(force_precision="fp32", shift_freeze=0.01)}


def trajectory_bench(L, Ls, beta, forms=("reference", "fast"), seed=693764523745.0):
    """One RHMC trajectory (rhmc.trajectory: heatbath, two Hamiltonians, a
    force per MD step; dt 0.05, iterl 20, the reference's tolerances rescga
    1e-9 / rescgg 1e-6, imass 1) at L, Ls from the same unthermalised field
    (chi ~ N(0, 1/beta), seeded) and generator seed, in the reference's solves
    (FP64 everywhere) and in the opt-in fast forms (FP32 force storage, shift
    freezing f = 0.01): seconds, MD steps, H0, dH, decision of each."""
    from chiralheisenberg_amd import fermion as fm, remez, rhmc
    par = dict(dt=0.05, beta=beta, am=0.0, hpin=0.0, imass=1, iterl=20, itermax=1000, iprint=5, rescga=1e-9,
               rescgg=1e-6)
    sets = {k: remez.load(k) for k in ("remez2", "remez4", "remez2g")}
    out = {"L": L, "Ls": Ls, "dt": par["dt"], "iterl": par["iterl"]}
    for name in forms:
        kw = TRAJ_FORMS[name]
        lat = fm.Lattice(L, Ls)
        fm.set_force_precision(lat, kw["force_precision"])
        lat.set_shift_freeze(kw["shift_freeze"])
        chi = np.random.RandomState(8).randn(3, lat.kvol) / np.sqrt(beta)
        print("bench: trajectory L=%d Ls=%d, %s forms %s" % (L, Ls, name, kw), file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        t = rhmc.trajectory(lat, chi, rhmc.RandomStream(seed), 1, par, sets)
        lat.synchronize()
        secs = time.perf_counter() - t0
        lat.close()
        out[name] = dict(kw, seconds=secs, md_steps=int(t["md_steps"]), H0=float(t["H0"]), dH=float(t["dH"]),
                         accepted=int(t["accepted"]))
        print("bench: %s %.1f s, %d MD steps, dH %.6f" % (name, secs, t["md_steps"], t["dH"]), file=sys.stderr,
              flush=True)
    if "reference" in out and "fast" in out:
        out["speedup"] = out["reference"]["seconds"] / out["fast"]["seconds"]
    return out


def shard_bench(L, Lt_full, Ls, shards):
    """The hot kernels at the shape of one t-slab shard of an L x L x
    Lt_full lattice split over `shards` GPUs (L x L x Lt_full/shards, the
    same Ls), against the whole lattice on one GPU, in the same process: the
    FP64 dslash, the binary16 dslash, the binary16 fused CG stencil and a
    plain FP64 copy, each over enough buffer sets to run from HBM where the
    sets can exceed the Infinity Cache.  The halo a shard would exchange
    per stencil and face (Ls L^2 8 complex FP64) and its time on one xGMI
    link are stated beside it; no second GPU is used."""
    from chiralheisenberg_amd import fermion as fm
    if Lt_full % shards:
        raise ValueError("Lt %d does not split into %d slabs" % (Lt_full, shards))
    out = {"L": L, "Lt": Lt_full, "Ls": Ls, "shards": shards}
    for name, Lt in (("whole", Lt_full), ("shard", Lt_full // shards)):
        lat = fm.Lattice(L, Ls, Lt=Lt)
        kvol = L * L * Lt
        V = kvol * Ls * 8 * 16
        V32, V16, C = V // 2, V // 4, 3 * kvol * 8
        r = {"Lt": Lt, "spinor_MB": V / 1e6}
        for key, prec, nbytes, inb in (("dslash_fp64", "fp64", 2 * V + C, V), ("dslash_fp16", "fp16", 2 * V16 + C, V16)):
            ms = fm.time_dslash(lat, prec, False, nbuf=cold_nbuf(inb), iters=200)
            r[key] = {"us": ms * 1e3, "GB_per_s": nbytes / (ms * 1e-3) / 1e9, "frac": nbytes / (ms * 1e-3) / HBM_PEAK}
        ms = fm.time_cg_stencil(lat, "fp16", nbuf=cold_nbuf(V16 + V32), iters=200)
        nb = V16 + 2 * V32 + C
        r["cg_stencil_fp16"] = {"us": ms * 1e3, "GB_per_s": nb / (ms * 1e-3) / 1e9, "frac": nb / (ms * 1e-3) / HBM_PEAK}
        ms = fm.time_copy(lat, "fp64", nbuf=cold_nbuf(V), iters=200)
        r["copy_fp64"] = {"us": ms * 1e3, "GB_per_s": 2 * V / (ms * 1e-3) / 1e9, "frac": 2 * V / (ms * 1e-3) / HBM_PEAK}
        lat.close()
        out[name] = r
        print("bench: %s (L=%d Lt=%d Ls=%d): dslash fp64 %.2f us, fp16 %.2f us, cg stencil fp16 %.2f us, copy %.2f us"
              % (name, L, Lt, Ls, r["dslash_fp64"]["us"], r["dslash_fp16"]["us"], r["cg_stencil_fp16"]["us"],
                 r["copy_fp64"]["us"]), file=sys.stderr, flush=True)
    for k in ("dslash_fp64", "dslash_fp16", "cg_stencil_fp16", "copy_fp64"):
        # the shard's share of the whole lattice's time over the shard's time
        out[k + "_efficiency"] = out["whole"][k]["us"] / shards / out["shard"][k]["us"]
    halo = Ls * L * L * 8 * 16
    out["halo_per_face"] = {"bytes": halo, "us_one_xgmi_link": halo / 153e9 * 1e6,
                            "what": "Ls L^2 8 complex FP64 per face and stencil, at ~153 GB/s per link"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--Ls", type=int, default=32)
    ap.add_argument("--beta", type=float, default=0.475)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--form", default="fp16", choices=["fp64", "fp32", "bf16", "fp16"])
    ap.add_argument("--delta", type=float, default=0.0, help="reliable-update threshold (0: library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline also times the L=12 Ls=24 multi-shift solve and force, and an L=8 trajectory (minutes)")
    ap.add_argument("--counters", action="store_true",
                    help="only the roofline kernels and calibration copies (for the counter runs)")
    ap.add_argument("--batch", action="store_true", help="only the batched multi-shift workload (one JSON line)")
    ap.add_argument("--sweep", action="store_true",
                    help="only the solve forms at L in {16, 24}, Ls=32, beta in {0.45, 0.475, 0.5} (one JSON line)")
    ap.add_argument("--no-extras", action="store_true", help="skip the learned-preconditioner and force figures")
    ap.add_argument("--force-only", action="store_true",
                    help="only the pseudofermion force figures (FP64 / FP32, with and without shift freezing)")
    ap.add_argument("--force-forms", default="fp64,fp32", help="with --force-only: storage forms, comma-separated")
    ap.add_argument("--force-freeze", default="0,0.01", help="with --force-only: shift-freezing factors")
    ap.add_argument("--shard", type=int, default=0,
                    help="only the hot kernels at the shape of one of N t-slab shards of the --L lattice (one JSON line)")
    ap.add_argument("--trajectory", action="store_true",
                    help="only one RHMC trajectory at --L/--Ls, reference and fast forms (one JSON line)")
    ap.add_argument("--traj-forms", default="reference,fast", help="with --trajectory: reference, fast")
    ap.add_argument("--learned-pairs", type=int, default=200,
                    help="training pairs of the learned-preconditioner figure")
    ap.add_argument("--learned-epochs", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this deployment has one MI355X, and the path "
              "shards by independent lattice/chain, so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.sweep:
        print(json.dumps(sweep_bench(args.tol, args.Ls)), flush=True)
        return
    if args.shard:
        print(json.dumps({"shard": shard_bench(args.L, args.L, args.Ls, args.shard)}), flush=True)
        return
    if args.trajectory:
        forms = tuple(args.traj_forms.split(","))
        print(json.dumps({"trajectory": trajectory_bench(args.L, args.Ls, args.beta, forms)}), flush=True)
        return
    if args.batch:
        out = {"%s_%s" % (m, p): batch_bench(m, precision=p) for m in BATCH_MIXES for p in ("fp64", "fp32")}
        for m in BATCH_MIXES:
            for p in ("fp64", "fp32"):
                out["force_%s_%s" % (m, p)] = force_batch_bench(m, precision=p)
        print(json.dumps(out), flush=True)
        return
    from chiralheisenberg_amd import fermion as fm
    lat = fm.Lattice(args.L, args.Ls)
    if args.force_only:
        chi, _ = make_inputs(lat, args.beta)     # the default run's field
        forms = tuple(args.force_forms.split(","))
        freezes = tuple(float(f) for f in args.force_freeze.split(","))
        print(json.dumps({"force": force_bench(fm, lat, chi, args.beta, forms=forms, freezes=freezes)}),
              flush=True)
        lat.close()
        return
    if args.counters:
        # counter runs: plain copies of an FP64, an FP32 and a BF16 spinor
        # (k_convert: 16 / 8 / 4 B read and written per complex number) calibrate FETCH_SIZE /
        # WRITE_SIZE against known bytes; then 20 launches of each roofline
        # kernel over the timing loops' rotation (inputs from HBM)
        for prec in (0, 1, 2):
            a = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128), ("fp64", "fp32", "bf16")[prec])
            b2 = fm.spinor_dev(lat, None, ("fp64", "fp32", "bf16")[prec])
            fm.check(lat.h, fm.lib().chs_convert_dev(lat.h, b2.ptr, prec, a.ptr, prec, lat.kferm))
            lat.synchronize()
        nb = cold_nbuf(lat.kferm * 16)
        print("bench: dslash fp64 %.2f us" % (fm.time_dslash(lat, "fp64", False, nbuf=nb, iters=20) * 1e3),
              flush=True)
        for prec in ("fp32", "fp16"):
            nbc = cold_nbuf(lat.kferm * (16 if prec == "fp32" else 12))
            print("bench: cg stencil %s %.2f us" % (prec, fm.time_cg_stencil(lat, prec, nbuf=nbc, iters=20) * 1e3),
                  flush=True)
        # the low-precision plain stencils (FP32 pairs, BF16, binary16)
        for prec, bc in (("fp32", 8), ("bf16", 4), ("fp16", 4)):
            nbl = cold_nbuf(lat.kferm * bc)
            print("bench: dslash %s %.2f us" % (prec, fm.time_dslash(lat, prec, False, nbuf=nbl, iters=20) * 1e3),
                  flush=True)
        return
    chi, b = make_inputs(lat, args.beta)
    chi_d = fm.DeviceArray.from_numpy(lat, chi)
    b_d = fm.spinor_dev(lat, b)
    x_d = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    zero = np.zeros(lat.shape, np.complex128)

    def solve(form):
        x_d.upload(zero)
        lat.synchronize()
        return fm.solve_mdagm(lat, b_d, x_d, chi_d, args.tol, form, delta=args.delta if form != "fp64" else 0.0)

    print("bench: L=%d Ls=%d form=%s warmup=%d steps=%d" % (args.L, args.Ls, args.form, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    # plain FP64 GPU form, once
    st64 = solve("fp64")
    print("bench: fp64 solve %.1f ms, %d its, resid %.2e" % (st64["seconds"] * 1e3, st64["iterations"],
                                                              st64["true_resid"]), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        solve(args.form)
    lat.synchronize()
    t0 = time.perf_counter()
    stats = [solve(args.form) for _ in range(args.steps)]
    lat.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    st = stats[-1]
    print("bench: %s solve %.1f ms (wall %.1f ms), %d its, %d reliable updates, resid %.2e" % (
        args.form, st["seconds"] * 1e3, wall * 1e3, st["iterations"], st["reliable_updates"], st["true_resid"]),
        file=sys.stderr, flush=True)
    # the other mixed forms, once each, for comparison
    others = {}
    for form in ("fp32", "fp16", "bf16"):
        if form != args.form:
            so = solve(form)
            others[form] = {"ms": so["seconds"] * 1e3, "iterations": so["iterations"],
                            "reliable_updates": so["reliable_updates"], "true_resid": so["true_resid"]}
            print("bench: %s solve %.1f ms, %d its" % (form, so["seconds"] * 1e3, so["iterations"]), file=sys.stderr,
                  flush=True)

    # rooflines: the FP64 dslash (2 V64 + C), the kernel the metric names, and
    # the dominant kernel of the timed solve, its fused CG stencil (M^dagger
    # M p with r -= alpha M^dagger M p and |r|^2).
    kvol = lat.kvol
    V = lat.kferm * 16
    V32 = lat.kferm * 8
    V16 = lat.kferm * 4
    C = 3 * kvol * 8
    # the fused CG stencil of the timed form: reads M p (storage precision)
    # and r, writes r (r in FP64 for the FP64 form, FP32 otherwise)
    cg_form = args.form
    cg_bytes = {"fp64": 3 * V + C, "fp32": 3 * V32 + C, "bf16": V16 + 2 * V32 + C, "fp16": V16 + 2 * V32 + C}[cg_form]
    cg_kernel = {"fp64": "k_dsl<F64,F64,true,EPI_CG>", "fp32": "k_dsl_pair<F32,F32,true,EPI_CG>",
                 "bf16": "k_dsl<B16,F32,true,EPI_CG>", "fp16": "k_dsl<H16,F32,true,EPI_CG>"}[cg_form]
    # inputs per buffer set: M p and r for the CG stencil, the input spinor
    # for the dslash and the copy
    cg_in = {"fp64": 2 * V, "fp32": 2 * V32, "bf16": V16 + V32, "fp16": V16 + V32}[cg_form]
    nb_cg, nb64 = cold_nbuf(cg_in), cold_nbuf(V)
    ms_cg = fm.time_cg_stencil(lat, cg_form, nbuf=nb_cg, iters=200)
    achieved_cg = cg_bytes / (ms_cg * 1e-3)
    ms = fm.time_dslash(lat, "fp64", False, nbuf=nb64, iters=200)
    achieved = (2 * V + C) / (ms * 1e-3)
    ms32 = fm.time_dslash(lat, "fp32", False, nbuf=cold_nbuf(V32), iters=200)
    ms16 = fm.time_dslash(lat, "bf16", False, nbuf=cold_nbuf(V16), iters=200)
    msh = fm.time_dslash(lat, "fp16", False, nbuf=cold_nbuf(V16), iters=200)
    # the streaming ceiling on the same kind of data: a plain copy of one FP64
    # spinor (2V) with the stencil's non-temporal stores
    ms_copy = fm.time_copy(lat, "fp64", nbuf=nb64, iters=200)
    copy_rate = 2 * V / (ms_copy * 1e-3)
    # the streaming ceiling at >= 1 GB: an FP64 spinor of L = Lt = 56, Ls = 32
    # (719 MB, so one copy moves 1.44 GB), with one element per thread and on
    # persistent grids of 1024-8192 workgroups
    big = fm.Lattice(56, 32)
    big_bytes = 2 * big.kferm * 16
    large_copy = {"bytes": big_bytes, "what": "plain FP64 copy of a 719 MB spinor (L=Lt=56, Ls=32), 2 buffer "
                                              "pairs of random values, 40 launches"}
    for grid in (0, 1024, 2048, 4096, 8192):
        msb = fm.time_copy_grid(big, grid, nbuf=2, iters=40)
        large_copy["full_grid" if grid == 0 else "persistent_%d" % grid] = {
            "us": msb * 1e3, "GB_per_s": big_bytes / (msb * 1e-3) / 1e9,
            "frac_of_peak": big_bytes / (msb * 1e-3) / HBM_PEAK}
    big.close()
    # the same two over 4 sets, whose inputs stay in the Infinity Cache
    ms_warm = fm.time_dslash(lat, "fp64", False, nbuf=4, iters=200)
    ms_copy_warm = fm.time_copy(lat, "fp64", nbuf=4, iters=200)
    print("bench: cg stencil %s %.2f us (%.0f GB/s); dslash fp64 %.2f us (%.0f GB/s), fp32 %.2f us, bf16 %.2f us, "
          "fp16 %.2f us; copy fp64 %.2f us (%.0f GB/s)" % (cg_form, ms_cg * 1e3, achieved_cg / 1e9, ms * 1e3,
                                                          achieved / 1e9, ms32 * 1e3, ms16 * 1e3, msh * 1e3,
                                                          ms_copy * 1e3, copy_rate / 1e9), file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": wall * 1e3,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16": "bf16+f32+f64", "fp16": "f16+f32+f64"}[args.form],
        "data": "synthetic (seeded phi ~ N(0,1/beta), complex Gaussian source, reference RNG)",
        "config": {"workload": "M^dagger M solve, domain-wall operator, L=Lt=%d Ls=%d beta=%g M5=1 imass=1, "
                               "to FP64 true relative residual %g" % (args.L, args.Ls, args.beta, args.tol),
                   "form": args.form, "L": args.L, "Ls": args.Ls, "tol": args.tol,
                   "parallelism": "single GPU"},
        "solve": {"form": args.form, "iterations": st["iterations"], "reliable_updates": st["reliable_updates"],
                  "true_resid": st["true_resid"], "device_ms": st["seconds"] * 1e3},
        "fp64_solve": {"ms": st64["seconds"] * 1e3, "iterations": st64["iterations"],
                       "true_resid": st64["true_resid"]},
        "speedup_vs_fp64": st64["seconds"] / max(wall, 1e-12),
        "other_forms": others,
        "dslash_us": {"fp64": ms * 1e3, "fp32": ms32 * 1e3, "bf16": ms16 * 1e3, "fp16": msh * 1e3},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": measured_traffic("dslash", args.L, args.Ls),
                     "kernel": ("k_dsl_plane<F64,false>"
                                if args.L * args.L <= 1024 and os.environ.get("CHS_PLANES", "1") != "0" else
                                "k_dsl<F64,F64,false,EPI_STORE>"), "us": ms * 1e3, "bytes_per_launch": 2 * V + C,
                     "inputs": "random values (all-zero buffers run ~15%% faster on this hardware), cycled over "
                               "%d buffer sets so that they come from HBM, not the 256 MB Infinity Cache" % nb64,
                     "copy_ceiling": {"GB_per_s": copy_rate / 1e9, "frac_of_peak": copy_rate / HBM_PEAK,
                                      "us": ms_copy * 1e3, "what": "plain copy of one FP64 spinor (2V), "
                                      "non-temporal stores, same random-data buffers rotation"},
                     "frac_of_copy": achieved / copy_rate,
                     "large_copy_ceiling": large_copy,
                     "infinity_cache_resident": {"nbuf": 4, "us": ms_warm * 1e3,
                                                 "GB_per_s": (2 * V + C) / (ms_warm * 1e-3) / 1e9,
                                                 "copy_us": ms_copy_warm * 1e3,
                                                 "what": "the same kernels over 4 buffer sets, whose 226 MB of "
                                                         "inputs stay in the Infinity Cache: not an HBM rate"}},
        "roofline_solve_kernel": {"bound": "hbm", "achieved": achieved_cg / 1e9, "peak": HBM_PEAK / 1e9,
                                  "unit": "GB/s", "frac": achieved_cg / HBM_PEAK,
                                  "traffic": measured_traffic("cg_stencil_" + cg_form, args.L, args.Ls),
                                  "kernel": cg_kernel, "us": ms_cg * 1e3,
                                  "bytes_per_launch": cg_bytes,
                                  "note": "isolated, inputs cycled over %d buffer sets (from HBM), alpha read "
                                          "from memory; the solve runs the same stencil with alpha formed inside "
                                          "the launch (EPI_CGW), and its working set stays in the 256 MB Infinity "
                                          "Cache" % nb_cg},
    }
    if not args.no_extras:
        out["learned"] = learned_bench(fm, lat, chi_d, b_d, args.tol, npairs=args.learned_pairs,
                                       epochs=args.learned_epochs)
        print("bench: learned %s" % json.dumps(out["learned"]), file=sys.stderr, flush=True)
        out["force"] = force_bench(fm, lat, chi, args.beta)
        print("bench: force %s" % json.dumps(out["force"]), file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.L, args.Ls, args.beta, st64["seconds"] * 1e3, args.cpu_full)
    else:
        out["cpu_baseline"] = None
    lat.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
