"""The C-ABI library loads on a host without a GPU and exports every entry
point include/chs.h declares; the product fails loudly (no CPU fallback)
when no GPU is present.  CPU only: no compute calls."""
import ctypes
import os
import re

import pytest

from chiralheisenberg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "chs.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(chs_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_reference_routines():
    names = declared()
    for n in ("chs_dslash", "chs_dslashd", "chs_qmrherm", "chs_congrad", "chs_force", "chs_hamilton",
              "chs_derivs", "chs_ctx_create", "chs_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    L = _lib.lib()
    for n in declared():
        assert getattr(L, n).argtypes is not None, n


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from chiralheisenberg_amd import fermion as fm
    with pytest.raises(_lib.ChsError):
        fm.Lattice(4, 4)


@pytest.mark.parametrize("dims,msg", [((1, 4, None), "extents"), ((4, 1, None), "extents"), ((4, 4, 1), "extents"),
                                      ((1024, 256, None), "too large")])
def test_context_rejects_bad_extents(dims, msg):
    """chs_ctx_create checks the extents before it touches a device: an
    extent below 2, and a lattice whose 8 Ls kvol components overflow the
    kernels' 32-bit indexing."""
    from chiralheisenberg_amd import fermion as fm
    L, Ls, Lt = dims
    with pytest.raises(_lib.ChsError, match=msg):
        fm.Lattice(L, Ls, Lt=Lt)


def test_rng_matches_oracle_on_host():
    """The product's host generator (libchs) and the oracle draw the same
    stream (both pinned to the reference by test_oracle)."""
    import numpy as np
    import oracle as orc
    from chiralheisenberg_amd.rhmc import RandomStream
    a, b = RandomStream(693764523745.0), orc.Rng(693764523745.0)
    assert [a.rano(i % 2) for i in range(500)] == [b.rano(i % 2) for i in range(500)]
    assert np.array_equal(a.gauss0(1000), b.gauss0(1000))
    assert np.array_equal(a.gaussp(1000), b.gaussp(1000))
    assert a.seed() == b.get()


def test_library_has_no_oracle_dependency():
    """The product library links no part of the oracle."""
    out = os.popen("nm -D --defined-only %s" % _lib.LIB_PATH).read()
    assert "orc_" not in out
    deps = os.popen("ldd %s" % _lib.LIB_PATH).read()
    assert "oracle" not in deps and "libref" not in deps


def test_driver_options_checked_before_any_gpu_work(tmp_path):
    """rhmc.run rejects a learned-preconditioner request without pairs before
    it creates a lattice context (host logic; no GPU needed)."""
    from chiralheisenberg_amd import rhmc
    with pytest.raises(ValueError):
        rhmc.run(str(tmp_path), 4, 4, learned=dict(pairs=0))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so,
    before it touches a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
