"""Writes tests/golden/ref_hmc_l12_ls8.npz: outputs of the reference's HMC
sibling mh/cheisdwf_hmc.f (force with its chronological guess, hamilton),
compiled unmodified by `make -C oracle ref` (ksize = ksizet = 12,
kthird = 8, FP64 kinds).  Inputs are regenerated by the tests from numpy's
legacy RandomState with the seed stored here; the fixture keeps the outputs
(full forces, samples of the updated guess at fixed indices plus 2-norms,
Hamiltonian parts).

Run in the build container:  python tests/golden/make_golden_hmc.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SEED = 20241203
BETA = 0.45
RES = 1e-11


def inputs(seed=SEED):
    rs = np.random.RandomState(seed)
    shape = (2, 4, 12, 12, 12, 8)
    phi = (rs.randn(*shape) + 1j * rs.randn(*shape)) / np.sqrt(2.0)
    chi = rs.randn(3, 1728) / np.sqrt(BETA)
    pp = rs.randn(3, 1728)
    x0 = 0.1 * (rs.randn(*shape) + 1j * rs.randn(*shape))
    idx = rs.choice(phi.size, 512, replace=False)
    return phi, chi, pp, x0, idx


def main():
    import refbind_mh as RM
    if not RM.available():
        sys.exit("oracle/_ref/libref_mh.so missing: run make -C oracle ref")
    mh = RM.RefMH()
    mh.set_params(BETA, 1.0, -1)
    phi, chi, pp, x0, idx = inputs()
    out = {"seed": SEED, "beta": BETA, "res": RES, "idx": idx}
    for imass in (1, 3):
        F, x0f = mh.force(phi, x0, RES, chi, imass)
        out["force_m%d" % imass] = F
        out["force_x0_m%d_s" % imass] = x0f.reshape(-1)[idx]
        out["force_x0_m%d_n" % imass] = np.linalg.norm(x0f)
        h, x0h = mh.hamilton(phi, x0, RES, chi, pp, imass)
        out["ham_m%d" % imass] = np.array(h)
        out["ham_x0_m%d_s" % imass] = x0h.reshape(-1)[idx]
    np.savez_compressed(os.path.join(HERE, "ref_hmc_l12_ls8.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_hmc_l12_ls8.npz"))


if __name__ == "__main__":
    main()
