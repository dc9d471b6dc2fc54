"""The drivers' `output` writer (chiralheisenberg_amd/outfile.py over the
edit descriptors of fortfmt.py) against the output files of the reference's
own programs (tests/golden/ref_traj_rhmc_imass3.npz: dwf3d, imass 3;
ref_traj_hmc.npz: the HMC sibling), CPU only: the header from the runs'
parameters and seeds, and the averages block from the averages the programs
printed, turned back into the sums the writer divides, must come out as the
programs' text, character for character."""
import os
import re

import numpy as np
import pytest

from chiralheisenberg_amd import fortfmt as F
from chiralheisenberg_amd import outfile

HERE = os.path.dirname(os.path.abspath(__file__))
_NUM = re.compile(r"NaN|[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:E[-+]\d+)?")


def test_edit_descriptors():
    assert F.e(1e-6, 9, 4) == ".1000E-05"            # the leading zero dropped to fit
    assert F.e(1.061, 11, 4) == " 0.1061E+01"
    assert F.e(-0.000591, 11, 3) == " -0.591E-03"
    assert F.e(0.0, 10, 3) == " 0.000E+00"
    assert F.e(float("nan"), 11, 3) == "        NaN"
    assert F.f(0.06, 6, 4) == "0.0600" and F.f(247.61, 9, 3) == "  247.610"
    assert F.f(0.5, 4, 3) == ".500" and F.f(12345.0, 6, 2) == "******"
    assert F.i(7, 3) == "  7" and F.i(1234, 3) == "***"
    assert F.list_real(3390816911360.0) == "3390816911360."
    # list-directed REAL*8 as the reference's flang build prints these values:
    # the middle of the shortest candidates, not the correctly rounded digits
    # (exact values ...8945236, ...4976542764, 2**54), and positional
    # exactly for 0.095 <= |v| < 9.5e14
    assert F.list_real(21.872546690894524) == "21.872546690894523"
    assert F.list_real(-0.00020601407094976543) == "-2.0601407094976543E-04"
    assert F.list_real(2.0 ** 54) == "1.8014398509481984E+16"
    assert F.list_real(0.095) == ".095" and F.list_real(0.0949999999) == "9.49999999E-02"
    assert F.list_real(949999990000000.0) == "949999990000000." and F.list_real(9.5e14) == "9.5E+14"
    assert F.list_real(0.045) == "4.5E-02" and F.list_real(0.0) == "0." and F.list_real(1e-5) == "1.E-05"


@pytest.mark.parametrize("case,rhmc", [("rhmc_imass3", True), ("hmc", False)])
def test_output_file_as_the_programs_write_it(case, rhmc, tmp_path):
    g = np.load(os.path.join(HERE, "golden", "ref_traj_%s.npz" % case), allow_pickle=False)
    ref = str(g["output_text"]).split("\n")
    par = {k[4:]: g[k].item() for k in g.files if k.startswith("par_")}
    par.update(rescgg=1e-6, rescga=1e-9)
    seeds = [float(_NUM.findall(line)[0]) for line in ref if "seed:" in line]
    n, itot, ip = int(par["iter2"]), int(g["itot"]), 1
    path = str(tmp_path / "output")
    outfile.header(path, seeds[0], seeds[1], 12, 12, 8, par, n, rhmc=rhmc)
    block = "\n".join(ref[9:])
    nums = [float(x) for x in _NUM.findall(block)]
    # averages for last n / acceptances, length / exp-dH, error / guidance, acceptance [/ force, pseudofermion]
    # / measurement / condmod, susc / cond (3) / psibarpsi / vel2, action / seed
    k = 3
    yav, err = nums[k], nums[k + 1]
    k += 2
    guid, acc_h = nums[k], nums[k + 1]
    k += 2
    force = pf = 0.0
    if rhmc:
        force, pf = nums[k], nums[k + 1]
        k += 2
    meas, condmod, susc = nums[k], nums[k + 1], nums[k + 2]
    cond = nums[k + 3:k + 6]
    pbp, vel2, action = nums[k + 6], nums[k + 7], nums[k + 8]
    a = dict(naccp=int(nums[1]), itot=itot, yav=yav * n, yyav=n * (err * err * (n - 1) + yav * yav),
             guidance=guid * itot, force=force * itot, acceptance=acc_h * 2 * n, pseudofermion=pf * n,
             measurement=meas * ip, ipbp=ip, condmod=condmod, susc=susc, cond=cond, pbp=pbp, vel2=vel2 * n,
             action=action * n)
    outfile.averages(path, a, n, par["dt"], seed_end=seeds[2], rhmc=rhmc)
    mine = open(path).read().split("\n")
    assert mine == ref


_INT = re.compile(r"^-?\d+$")


@pytest.mark.parametrize("case", ["rhmc_imass1", "rhmc_imass3", "hmc"])
def test_data_files_as_the_programs_write_them(case):
    """Every record of the programs' data files (control, fort.11-14,
    fort.50xx: WRITE(unit,*) lists of integers and REAL*8), read back to its
    numbers and written again by list_directed, is the program's line
    character for character."""
    g = np.load(os.path.join(HERE, "golden", "ref_traj_%s.npz" % case), allow_pickle=False)
    keys = [k for k in g.files if k.startswith("text_")]
    assert keys
    n = 0
    for k in keys:
        for line in str(g[k]).splitlines():
            vals = [int(t) if _INT.match(t) else float(t) for t in line.split()]
            assert F.list_directed(*vals) == line, (k, line)
            n += 1
    assert n >= 5


FLANG = "/opt/rocm/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_list_real_against_flang(tmp_path):
    """list_real against the flang runtime itself (the compiler the
    reference is built with): a five-line program reads each value and
    writes it list-directed; 20,000 values over 60 decades, the powers of
    two, integers, and the edges of the positional range."""
    import subprocess
    rs = np.random.RandomState(5)
    vals = list(rs.randn(12000) * 10.0 ** rs.randint(-30, 30, 12000)) + list(rs.rand(4000)) + \
        list(rs.rand(4000) * 100) + [2.0 ** k for k in range(-70, 70)] + \
        list(rs.randint(-10 ** 6, 10 ** 6, 200).astype(float)) + \
        [0.1, 0.095, 0.0949999999, 9.5e14, 949999990000000.0, 1e15, 1e16, 2.0 ** 54, 5e-324, 1.7976931348623157e308,
         np.nextafter(0.095, 0.0), np.nextafter(9.5e14, 0.0), -0.0, 0.0]
    src = tmp_path / "ld.f"
    src.write_text("      program ld\n      real*8 v\n      integer k, n\n      read(5,*) n\n"
                   "      do k = 1, n\n        read(5,*) v\n        write(6,*) v\n      enddo\n      end\n")
    exe = tmp_path / "ld"
    subprocess.run([FLANG, "-fdefault-real-8", str(src), "-o", str(exe)], check=True, capture_output=True)
    inp = "%d\n" % len(vals) + "".join("%r\n" % float(v) for v in vals)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.splitlines()
    assert len(out) == len(vals)
    bad = [(v, o, F.list_real(v)) for v, o in zip(vals, out) if o != " " + F.list_real(v)]
    assert not bad, bad[:5]
