"""Writes tests/golden/ref_fermion_l12_ls8.npz: the wall-source correlators
of the reference's own fermion routine (rhmc/cheisdwf.f:897-1058), compiled
unmodified by `make -C oracle ref` (ksize = ksizet = 12, kthird = 8), on the
reference's thermalised configuration rhmc/_con (tests/golden/con_l12_fp64).

The routine writes fort.5001/5011/5002/5012 and keeps those units open, so
every case runs in a process of its own, in a scratch directory.

Run in the build container:  python tests/golden/make_golden_fermion.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

RNG_SEED = 693764523745.0
# the third case runs with ibound = +1: periodic in t, and the propagator's
# wrap-around sign isign = ibound (rhmc/cheisdwf.f:1021) is +1
CASES = [dict(imass=1, hpin=0.0, res=1e-11), dict(imass=3, hpin=0.5, res=1e-8),
         dict(imass=1, hpin=0.3, res=1e-11, ibound=1)]


def run_case(case, out_json):
    import refbind as RB
    from chiralheisenberg_amd import checkpoint
    chi, _, _ = checkpoint.read_con(os.path.join(HERE, "con_l12_fp64"), kvol=RB.KVOL)
    ref = RB.Ref("fp64")
    ref.set_params(0.45, 1.0, case.get("ibound", -1))
    # the main stream after one draw, as the driver leaves it before a
    # measurement (the measurement draws the next number)
    ref.rranset(RNG_SEED)
    _, y, idum = ref.rano(0.0, -1)
    ref.set_ranseed(y, idum)
    with tempfile.TemporaryDirectory() as d:
        c1, c2, av = ref.fermion(chi, case["res"], d, case["hpin"], case["imass"])
    json.dump(dict(c1r=c1.real.tolist(), c1i=c1.imag.tolist(), c2r=c2.real.tolist(), c2i=c2.imag.tolist(),
                   aviter=av), open(out_json, "w"))


def main():
    if len(sys.argv) > 2:
        run_case(json.loads(sys.argv[1]), sys.argv[2])
        return
    out = {"rng_seed": RNG_SEED}
    for k, case in enumerate(CASES):
        with tempfile.TemporaryDirectory() as d:
            js = os.path.join(d, "out.json")
            subprocess.check_call([sys.executable, __file__, json.dumps(case), js])
            r = json.load(open(js))
        out["case%d_imass" % k] = case["imass"]
        out["case%d_hpin" % k] = case["hpin"]
        out["case%d_res" % k] = case["res"]
        out["case%d_ibound" % k] = case.get("ibound", -1)
        out["case%d_cferm1" % k] = np.array(r["c1r"]) + 1j * np.array(r["c1i"])
        out["case%d_cferm2" % k] = np.array(r["c2r"]) + 1j * np.array(r["c2i"])
        out["case%d_aviter" % k] = r["aviter"]
        print("case", case, "aviter", r["aviter"], flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_fermion_l12_ls8.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_fermion_l12_ls8.npz"))


if __name__ == "__main__":
    main()
