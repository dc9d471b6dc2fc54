"""The con checkpoint format (rhmc/cheisdwf.f:1455-1492, rhmc/sp_to_dp.f)
against the reference's own files: tests/golden/con_l12_fp64 is
rhmc/_con, tests/golden/con_l12_fp32 is rhmc/Ls_8/test/con (the FP64 file is
the FP32 one widened by sp_to_dp)."""
import os

import numpy as np
import pytest

from chiralheisenberg_amd import checkpoint as ck

G = os.path.join(os.path.dirname(__file__), "golden")


def test_read_reference_files():
    c64, s64, p64 = ck.read_con(os.path.join(G, "con_l12_fp64"), kvol=12 ** 3)
    c32, s32, p32 = ck.read_con(os.path.join(G, "con_l12_fp32"), kvol=12 ** 3)
    assert (p64, p32) == ("fp64", "fp32")
    assert s64 == s32 == 122110675517440.0
    assert c64.shape == (3, 1728)
    # sp_to_dp widens exactly
    assert np.array_equal(c64, c32)
    assert np.array_equal(c64, c64.astype(np.float32).astype(np.float64))


def test_round_trip_bit_exact(tmp_path):
    c64, s64, _ = ck.read_con(os.path.join(G, "con_l12_fp64"), kvol=1728)
    p = tmp_path / "con"
    ck.write_con(p, c64, s64)
    assert open(p, "rb").read() == open(os.path.join(G, "con_l12_fp64"), "rb").read()
    c32, s32, _ = ck.read_con(os.path.join(G, "con_l12_fp32"), precision="fp32")
    ck.write_con(p, c32, s32, "fp32")
    assert open(p, "rb").read() == open(os.path.join(G, "con_l12_fp32"), "rb").read()


def test_sp_to_dp(tmp_path):
    dst = tmp_path / "condp"
    ck.sp_to_dp(os.path.join(G, "con_l12_fp32"), dst)
    assert open(dst, "rb").read() == open(os.path.join(G, "con_l12_fp64"), "rb").read()


def test_sread_default_and_errors(tmp_path):
    chi, seed = ck.sread(tmp_path / "missing", 64, 0.475)
    assert seed == ck.DEFAULT_SEED and np.all(chi == 0.1 / 0.475) and chi.shape == (3, 64)
    bad = tmp_path / "bad"
    open(bad, "wb").write(b"\x10\x00\x00\x00" + b"\x00" * 16 + b"\x11\x00\x00\x00")
    with pytest.raises(ValueError):
        ck.read_con(bad)
    with pytest.raises(ValueError):
        ck.read_con(os.path.join(G, "con_l12_fp64"), kvol=8 ** 3)
    with pytest.raises(ValueError):   # 20,736 bytes: FP64 chi(864,3) or FP32 chi(1728,3)
        ck.read_con(os.path.join(G, "con_l12_fp32"))


def test_run_reads_the_callers_remez_files(tmp_path):
    """rhmc.run takes remez2 / remez4 / remez2g from its working directory
    when they are there, as dwf3d does (rhmc/cheisdwf.f:114-116), and the
    bundled copies otherwise."""
    from chiralheisenberg_amd import remez
    own = remez.load("remez2g")
    text = "%.17e\n" % (2 * own[0][0]) + "".join("%.17e, %.17e\n" % (a, d) for a, d in zip(own[0][1:], own[1]))
    text += "%.17e\n" % own[2][0] + "".join("%.17e, %.17e\n" % (b, d) for b, d in zip(own[2][1:], own[3]))
    (tmp_path / "remez2g").write_text(text)
    sets = remez.load_sets(str(tmp_path))
    assert sets["remez2g"][0][0] == 2 * own[0][0]
    assert np.array_equal(sets["remez2g"][1], own[1]) and np.array_equal(sets["remez2g"][3], own[3])
    for k in ("remez2", "remez4"):
        assert all(np.array_equal(x, y) for x, y in zip(sets[k], remez.load(k)))
    assert all(np.array_equal(x, y) for x, y in zip(remez.load_sets(None)["remez2g"], own))
