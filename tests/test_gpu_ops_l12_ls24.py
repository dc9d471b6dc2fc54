"""Single operations at L=Lt=12, Ls=24 (beta 0.475, imass 1, seeded
Gaussian phi_a and spinors: the FP64 drop-in workload) on the
GPU against the CPU oracle's results for the same inputs, kept as the fixture
tests/golden/oracle_ops_l12_ls24.npz (written by
make_oracle_ops_l12_ls24.py: the oracle needs minutes for them):
dslash / dslashd within 1e-13, the remez4 multi-shift solution at res 1e-10
within 1e-9 with the iteration count within 5, the force within 1e-10 and
the Hamiltonian within 1e-10."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import remez  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_ops_l12_ls24 as gen  # noqa: E402


@pytest.fixture(scope="module")
def setup():
    ref = np.load(os.path.join(HERE, "golden", "oracle_ops_l12_ls24.npz"))
    lat = fm.Lattice(gen.L, gen.LS)
    yield lat, ref, gen.inputs()
    lat.close()


def test_dslash_l12_ls24(setup):
    lat, ref, (chi, R, Phi, pp, idx) = setup
    for dag in (0, 1):
        y = (fm.dslashd if dag else fm.dslash)(lat, R, chi, 0.0, 0.5, 1).reshape(-1)
        s = ref["dslash_d%d_s" % dag]
        assert np.abs(y[idx] - s).max() <= 1e-13 * np.abs(s).max()
        n = float(ref["dslash_d%d_n" % dag])
        assert abs(np.linalg.norm(y) - n) <= 1e-13 * n


def test_qmrherm_remez4_l12_ls24(setup):
    lat, ref, (chi, R, Phi, pp, idx) = setup
    a4, ad4 = remez.load("remez4")[:2]
    x, it = fm.qmrherm(lat, Phi, 1e-10, a4, ad4, chi, 0.0, 0.0, 1, 0)
    x = x.reshape(-1)
    s = ref["qmr_s"]
    assert abs(it - int(ref["qmr_it"])) <= 5, (it, int(ref["qmr_it"]))
    assert np.abs(x[idx] - s).max() <= 1e-9 * np.abs(s).max()
    n = float(ref["qmr_n"])
    assert abs(np.linalg.norm(x) - n) <= 1e-9 * n


def test_force_l12_ls24(setup):
    lat, ref, (chi, R, Phi, pp, idx) = setup
    F, _, _ = fm.force(lat, Phi, 1e-12, chi, gen.BETA, remez.load("remez2g"), 0.0, 0.0, 1)
    Fc = ref["force"]
    assert np.abs(F - Fc).max() <= 1e-10 * np.abs(Fc).max()


def test_hamilton_l12_ls24(setup):
    lat, ref, (chi, R, Phi, pp, idx) = setup
    h, _, _ = fm.hamilton(lat, Phi, 1e-11, chi, pp, gen.BETA, remez.load("remez2"))
    for a, b in zip(h, ref["ham"]):
        assert abs(a - b) <= 1e-10 * abs(b), (h, ref["ham"])


def test_qmrherm_shifts_l12_ls24(setup):
    """Every shift's solution of the remez4 multi-shift solve (25 shifts) at
    L=12, Ls=24, res 1e-10 with the reference's stopping rule: against the
    oracle's x_k and each shift's FP64 true residual |Phi - (M^dagger M +
    aden_k) x_k| / |Phi| (the GPU's FP64 operator, pinned to the oracle's
    by the dslash test above)."""
    import oracle as orc
    lat, ref, (chi, R, Phi, pp, idx) = setup
    a4, ad4 = remez.load("remez4")[:2]
    xk, it = fm.qmrherm_shifts(lat, Phi, 1e-10, ad4, chi)
    assert abs(it - int(ref["qmr_it"])) <= 5
    # the combination of the shifts is the iflag = 0 solution
    x = a4[0] * Phi + np.tensordot(a4[1:], xk, axes=1)
    s = ref["qmr_s"]
    assert np.abs(x.reshape(-1)[idx] - s).max() <= 1e-9 * np.abs(s).max()
    worst = 0.0
    pn = np.linalg.norm(Phi)
    for k in range(len(ad4)):
        Ax = fm.dslashd(lat, fm.dslash(lat, xk[k], chi, 0.0, 0.0, 1), chi, 0.0, 0.0, 1)
        rk = np.linalg.norm(Phi - Ax - ad4[k] * xk[k]) / pn
        worst = max(worst, rk)
    print("remez4 at L=12, Ls=24: itercg %d, largest per-shift true residual %.2e" % (it, worst))
    assert worst <= 1e-10
    # the oracle's shift solutions on an L=8, Ls=8 lattice
    lat1, olat1 = fm.Lattice(8, 8), orc.Lattice(8, 8)
    rs = np.random.RandomState(90)
    phi1 = (rs.randn(*lat1.shape) + 1j * rs.randn(*lat1.shape)) / np.sqrt(2)
    chi1 = rs.randn(3, lat1.kvol) / np.sqrt(gen.BETA)
    xg, ig = fm.qmrherm_shifts(lat1, phi1, 1e-10, ad4, chi1)
    _, ic, xc = orc.qmrherm(olat1, phi1, 1e-10, a4, ad4, chi1, want_shifts=True)
    assert abs(ig - ic) <= 5
    for k in range(len(ad4)):
        assert np.abs(xg[k] - xc[k]).max() <= 1e-9 * np.abs(xc[k]).max(), k
    lat1.close()
