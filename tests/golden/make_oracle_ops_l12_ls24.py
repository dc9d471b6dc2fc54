"""Writes tests/golden/oracle_ops_l12_ls24.npz: the CPU oracle's results
for single operations at L=Lt=12, Ls=24, beta=0.475, imass=1 on seeded
inputs (chi ~ N(0, 1/beta), complex Gaussian spinors, Gaussian momenta):
dslash and dslashd (hpin 0.5; 4096 sampled entries and the 2-norm of each),
qmrherm over all 25 remez4 shifts to res 1e-10 (iflag 0; sampled entries,
norm, iterations), the force with remez2g at res 1e-12, and hamilton with
remez2 at res 1e-11.  The oracle needs a few minutes for these, so they are
a fixture for the GPU test, which rebuilds the same inputs with inputs().

    python tests/golden/make_oracle_ops_l12_ls24.py [threads]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import oracle as orc  # noqa: E402
from chiralheisenberg_amd import remez  # noqa: E402

L, LS, BETA = 12, 24, 0.475
SHAPE = (2, 4, L, L, L, LS)
KVOL = L ** 3


def inputs():
    """chi (3, kvol), R and Phi (complex spinors in the reference layout),
    pp (3, kvol) and the sample indices into a flattened spinor."""
    rs = np.random.RandomState(2012)
    chi = rs.randn(3, KVOL) / np.sqrt(BETA)
    R = (rs.randn(*SHAPE) + 1j * rs.randn(*SHAPE)) / np.sqrt(2.0)
    Phi = (rs.randn(*SHAPE) + 1j * rs.randn(*SHAPE)) / np.sqrt(2.0)
    pp = rs.randn(3, KVOL)
    idx = np.sort(rs.choice(R.size, 4096, replace=False))
    return chi, R, Phi, pp, idx


def main():
    if len(sys.argv) > 1:
        orc.set_threads(int(sys.argv[1]))
    chi, R, Phi, pp, idx = inputs()
    olat = orc.Lattice(L, LS)
    out = {"idx": idx}
    for dag in (0, 1):
        y = orc.dslash(olat, R, chi, 0.5, 1, bool(dag)).reshape(-1)
        out["dslash_d%d_s" % dag] = y[idx]
        out["dslash_d%d_n" % dag] = np.array(np.linalg.norm(y))
        print("dslash dagger=%d done" % dag, flush=True)
    a4, ad4 = remez.load("remez4")[:2]
    x, it = orc.qmrherm(olat, Phi, 1e-10, a4, ad4, chi, 0.0, 1, 0)
    x = x.reshape(-1)
    out["qmr_s"], out["qmr_n"], out["qmr_it"] = x[idx], np.array(np.linalg.norm(x)), np.array(it)
    print("qmrherm done (%d iterations)" % it, flush=True)
    F, _, itf = orc.force(olat, Phi, 1e-12, chi, BETA, remez.load("remez2g"), 0.0, 1)
    out["force"], out["force_it"] = F, np.array(itf)
    print("force done (%s iterations)" % (itf,), flush=True)
    h, _, ith = orc.hamilton(olat, Phi, 1e-11, chi, pp, BETA, remez.load("remez2"), 0.0, 1)
    out["ham"], out["ham_it"] = np.array(h), np.array(ith)
    print("hamilton done (%d iterations)" % ith, flush=True)
    np.savez_compressed(os.path.join(HERE, "oracle_ops_l12_ls24.npz"), **out)


if __name__ == "__main__":
    main()
