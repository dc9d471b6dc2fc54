"""The batched workload at full size: multi-shift solves (qmrherm, rhmc/cheisdwf.f:
599-745, iflag = 0) of all 12 remez2g shifts x 4 pseudofermion fields over
the lattice mix L/Ls = 8/8, 12/16, 16/32, 24/32 in one grouped,
graph-replayed solve per batch, FP64 and FP32 storage.

Checked three ways:
* against the oracle's multi-shift solve (the CPU restatement pinned to the
  reference) on the systems the CPU finishes in seconds (L = 8 and 12):
  solutions to the solver tolerance, iteration counts within a few;
* against the one-at-a-time GPU solve of one field per lattice, bit for bit
  (the full mix shares no state between systems);
* on every system, through what the rational functions fix: the remez2g
  b-set is the reciprocal of its a-set (r_a(x) r_b(x) = 1 to 2e-13 for every
  x > 0), so a batch of r_a(M^dagger M) Phi followed by a batch of r_b of its
  results returns Phi, whatever the spectrum, to the solvers' tolerance
  (measured: 1.8e-13 in FP64 at res 1e-10).  FP32 storage runs at the force's
  tolerance 1e-6."""
import numpy as np
import pytest

import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import batch, remez  # noqa: E402

MIX = ((8, 8), (12, 16), (16, 32), (24, 32))
NPF = 4
R2G = remez.load("remez2g")


def build(seed=4, res=1e-10):
    rs = np.random.RandomState(seed)
    lats, systems, hosts = [], [], []
    for L, Ls in MIX:
        lat = fm.Lattice(L, Ls)
        lats.append(lat)
        chi = phi_field(rs, lat.kvol)
        chi_d = fm.DeviceArray.from_numpy(lat, chi)
        for _ in range(NPF):
            phi = spinor(rs, lat.shape)
            systems.append(batch.System(lat, fm.spinor_dev(lat, phi), chi_d, fm.spinor_dev(lat, None), R2G[0], R2G[1],
                                        res=res))
            hosts.append((phi, chi))
    return lats, systems, hosts


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_mix_l8_to_l24(precision):
    # FP32 storage serves the force's solves (the reference's rescgg = 1e-6):
    # its Lanczos vectors cannot carry a residual of 1e-10 (at L=24, Ls=32 it
    # stops at the 5000-iteration cap, reported as CHS_ENOCONV)
    res = 1e-10 if precision == "fp64" else 1e-6
    lats, systems, hosts = build(res=res)
    b = batch.Batch()
    print("batched mix (%s): %d systems" % (precision, len(systems)), flush=True)
    st = b.qmrherm(systems, None, precision)
    print("  r_a batch: %.2f s, %d lock-step iterations, tune %s" % (st["seconds"], st["iterations"], st["tune"]),
          flush=True)
    ya = [fm.to_host(s.lat, s.x) for s in systems]
    its = [s.itercg for s in systems]
    # oracle on the small lattices (FP64 storage only: the FP32 form is not
    # the oracle's arithmetic)
    if precision == "fp64":
        for s, (phi, chi), y, it in zip(systems, hosts, ya, its):
            if s.lat.L > 12:
                continue
            xo, ito = orc.qmrherm(orc.Lattice(s.lat.L, s.lat.Ls), phi, 1e-10, R2G[0], R2G[1], chi)
            assert abs(it - ito) <= 5
            assert np.abs(y - xo).max() <= 1e-9 * np.abs(xo).max()
        # one field per lattice one at a time: bit for bit
        for k in range(0, len(systems), NPF):
            s = systems[k]
            x1 = fm.spinor_dev(s.lat, None)
            it1 = fm.qmrherm_prec_dev(s.lat, s.phi, s.res, s.anum, s.aden, s.chi, x1, precision)
            assert it1 == its[k]
            assert np.array_equal(fm.to_host(s.lat, x1), ya[k])
    # r_b(M^dagger M) r_a(M^dagger M) Phi = Phi
    second = []
    for s in systems:
        y = fm.spinor_dev(s.lat, None)
        fm.check(s.lat.h, fm.lib().chs_convert_dev(s.lat.h, y.ptr, 0, s.x.ptr, 0, s.lat.kferm))
        s.lat.synchronize()
        second.append(batch.System(s.lat, y, s.chi, fm.spinor_dev(s.lat, None), R2G[2], R2G[3], res=res))
    st2 = b.qmrherm(second, None, precision)
    print("  r_b batch: %.2f s" % st2["seconds"], flush=True)
    worst = 0.0
    for s, (phi, _) in zip(second, hosts):
        z = fm.to_host(s.lat, s.x)
        worst = max(worst, np.linalg.norm(z - phi) / np.linalg.norm(phi))
    print("  |r_b(A) r_a(A) Phi - Phi| / |Phi|: worst %.2e over %d systems" % (worst, len(systems)), flush=True)
    assert worst <= (1e-8 if precision == "fp64" else 1e-5)
    b.close()
    for lat in lats:
        lat.close()
