"""The learned preconditioner: GPU network (BF16 MFMA kernels) against the
CPU oracle oracle/nn_oracle.py (same BF16 rounding points, FP64
accumulation), its symmetry and positivity, flexible CG with and without the
network, and training -- at a small lattice and at L = 24, Ls = 32."""
import ctypes

import numpy as np
import pytest

import nn_oracle as nno
import oracle as orc
from conftest import phi_field, spinor

pytestmark = pytest.mark.gpu

fm = pytest.importorskip("chiralheisenberg_amd.fermion")
from chiralheisenberg_amd import learn  # noqa: E402
from chiralheisenberg_amd._lib import ChsError  # noqa: E402


def net(L, Ls, Lt, seed=5, scale=0.3, theta=-1.0):
    lat = fm.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(seed)
    pc = learn.Preconditioner(lat, hidden=16, seed=7)
    w = (rs.randn(pc.nweights) * scale).astype(np.float32)
    w[-1] = theta
    pc.set_weights(w)
    return lat, pc, w, rs


def to_ref(lat, d):
    return fm.to_host(lat, d)


@pytest.mark.parametrize("L,Ls,Lt", [(4, 2, 4), (6, 4, 8), (5, 3, 3)])
def test_network_forward_matches_oracle(L, Ls, Lt):
    """P(r) on the GPU against the oracle, with bands of fewer than 4 rows
    (L = 6, 5) and partial 16-site tiles (L = 5)."""
    lat, pc, w, rs = net(L, Ls, Lt)
    src = spinor(rs, lat.shape).astype(np.complex64).astype(np.complex128)
    r = fm.spinor_dev(lat, src, "fp32")
    z = fm.spinor_dev(lat, None, "fp32")
    pc.apply(r, z)
    got = to_ref(lat, z)
    Z, _ = nno.forward(src, w, Ls)
    want = nno.reference_from_native(nno.channels_to_spinor(Z, Ls))
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 2e-3, err
    lat.close()


def test_network_is_symmetric_positive_definite():
    """<a, P b> = <P a, b> up to the BF16 rounding of the operands, and
    <a, P a> >= c |a|^2 (P's input is a / rms(a), so both sides are taken
    with that scaling)."""
    lat, pc, w, rs = net(6, 4, 6)
    a = spinor(rs, lat.shape).astype(np.complex64).astype(np.complex128)
    b = spinor(rs, lat.shape).astype(np.complex64).astype(np.complex128)

    def P(v):
        d, z = fm.spinor_dev(lat, v, "fp32"), fm.spinor_dev(lat, None, "fp32")
        pc.apply(d, z)
        rms = np.sqrt(np.mean(np.abs(v) ** 2) / 2.0)
        return to_ref(lat, z) * rms
    Pa, Pb = P(a), P(b)
    lhs, rhs = np.vdot(a, Pb).real, np.vdot(Pa, b).real
    assert abs(lhs - rhs) <= 2e-2 * np.sqrt(np.vdot(a, Pa).real * np.vdot(b, Pb).real), (lhs, rhs)
    assert np.vdot(a, Pa).real >= 0.99 * np.exp(w[-1]) * np.vdot(a, a).real
    lat.close()


def test_network_gradient_matches_oracle():
    """Loss (relative A-norm of P's error, A = M^dagger M by the FP32 stencil)
    and the gradients of every layer and of theta against the oracle (A in
    FP64 by the CPU oracle's operator)."""
    L, Ls, Lt = 4, 2, 6
    lat, pc, w, rs = net(L, Ls, Lt, scale=0.2)
    olat = orc.Lattice(L, Ls, Lt)
    chi = phi_field(rs, lat.kvol)
    chid = fm.DeviceArray.from_numpy(lat, chi)
    src = spinor(rs, lat.shape).astype(np.complex64).astype(np.complex128)
    sol = orc.congrad(olat, src, 1e-10, chi)[0].astype(np.complex64).astype(np.complex128)
    loss, g = pc.loss_grad(fm.spinor_dev(lat, src, "fp32"), fm.spinor_dev(lat, sol, "fp32"), chid)

    def A(v):
        return orc.dslash(olat, orc.dslash(olat, v, chi, 0.0, 1, False), chi, 0.0, 1, True)
    loss_o, g_o = nno.loss_and_grad(src, sol, w, Ls, A)
    assert abs(loss - loss_o) <= 1e-3 * abs(loss_o), (loss, loss_o)
    n1, n2 = 16 * 16 * Ls, 27 * 16 * 16
    for name, sl in (("W1", slice(0, n1)), ("W2", slice(n1, n1 + n2))):
        e = np.linalg.norm(g[sl] - g_o[sl]) / np.linalg.norm(g_o[sl])
        assert e < 2e-2, (name, e)
    assert abs(g[-1] - g_o[-1]) <= 1e-2 * abs(g_o[-1]), (g[-1], g_o[-1])
    lat.close()


def test_fcg_without_network_is_cg():
    lat = fm.Lattice(8, 8)
    rs = np.random.RandomState(9)
    b, chi = spinor(rs, lat.shape), phi_field(rs, lat.kvol)
    chid = fm.DeviceArray.from_numpy(lat, chi)
    bd = fm.spinor_dev(lat, b)
    x1 = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    st_f = learn.solve_fcg(lat, bd, x1, chid, 1e-10, use_nn=False)
    x2 = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    st_c = fm.solve_mdagm(lat, bd, x2, chid, 1e-10, "fp32")
    assert st_f["true_resid"] <= 1e-10 and st_c["true_resid"] <= 1e-10
    assert abs(st_f["iterations"] - st_c["iterations"]) <= 0.05 * st_c["iterations"] + 3, (st_f, st_c)
    lat.close()


def gain_run(L, Ls, npairs, epochs, lr, seed=13):
    """Train on npairs (Gaussian source, solution) pairs of the plain solver,
    then solve a fresh system with flexible CG with and without the network
    and with the plain mixed-precision CG (binary16 form); returns the
    statistics."""
    lat = fm.Lattice(L, Ls)
    rs = np.random.RandomState(seed)
    chi = phi_field(rs, lat.kvol)
    chid = fm.DeviceArray.from_numpy(lat, chi)
    src, sol, _ = learn.pairs_dev(lat, chid, npairs, rng_seed=3, precision="fp16")
    pc = learn.Preconditioner(lat, hidden=16, seed=1)
    losses = pc.train(src, sol, npairs, chid, epochs=epochs, lr=lr)
    src.free()
    sol.free()
    b = spinor(rs, lat.shape)
    bd = fm.spinor_dev(lat, b)
    out = {"losses": losses}
    for key, use in (("nn", True), ("fcg", False)):
        x = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
        out[key] = learn.solve_fcg(lat, bd, x, chid, 1e-10, use_nn=use)
        x.free()
    x = fm.spinor_dev(lat, np.zeros(lat.shape, np.complex128))
    out["fp16"] = fm.solve_mdagm(lat, bd, x, chid, 1e-10, "fp16")
    lat.close()
    return out


def report(tag, r):
    nn, fc, h = r["nn"], r["fcg"], r["fp16"]
    print("%s: loss %.4f -> %.4f; iterations / ms: FCG with network %d / %.1f (fallbacks %d), without %d / %.1f, "
          "plain binary16 CG %d / %.1f; iteration gain %.3fx over FCG, %.3fx over the plain CG; time gain %.3fx" % (
              tag, r["losses"][0], r["losses"][-1], nn["iterations"], nn["seconds"] * 1e3, nn["restarts"],
              fc["iterations"], fc["seconds"] * 1e3, h["iterations"], h["seconds"] * 1e3,
              fc["iterations"] / nn["iterations"], h["iterations"] / nn["iterations"],
              h["seconds"] / nn["seconds"]), flush=True)


def test_trained_network_gain_small():
    """Training lowers the loss; flexible CG with the trained network reaches
    the FP64 true residual, and the test records the gain over the same
    solver without it (measured: below 1, see the network's design notes)."""
    r = gain_run(8, 8, 16, 20, 2e-3)
    report("L=8 Ls=8", r)
    assert r["losses"][-1] < r["losses"][0], r["losses"]
    for k in ("nn", "fcg", "fp16"):
        assert r[k]["true_resid"] <= 1e-10, (k, r[k])
    assert r["nn"]["iterations"] <= 2.0 * r["fcg"]["iterations"] + 64 * r["nn"]["restarts"]


def test_trained_network_gain_l24_ls32():
    """The learned preconditioner's benchmark configuration (L = 24, Ls = 32,
    beta = 0.475, mass 0.01) with 32 training pairs (the bench trains on
    more): every solve reaches the FP64 true residual 1e-10; the gain over
    flexible CG without the network and over the plain binary16 CG is
    recorded, whatever its sign."""
    r = gain_run(24, 32, 32, 4, 1e-3)
    report("L=24 Ls=32", r)
    for k in ("nn", "fcg", "fp16"):
        assert r[k]["true_resid"] <= 1e-10, (k, r[k])


def test_fermion_learned_form_matches_oracle():
    """The propagator solves of fermion.fermion in the learned form (flexible
    CG with the network: untrained, then with random weights) give the
    oracle's correlators to the bound of the FP32 / FP16 forms."""
    L, Ls, Lt, imass, ittt0 = 6, 4, 10, 1, 3
    lat, olat = fm.Lattice(L, Ls, Lt), orc.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(64)
    chi = phi_field(rs, lat.kvol)
    o1, o2, oit = orc.fermion(olat, chi, 1e-11, ittt0, 0.3, imass)
    pc = learn.Preconditioner(lat, hidden=16, seed=3)
    w = (rs.randn(pc.nweights) * 0.3).astype(np.float32)
    w[-1] = -1.0
    for tag, weights in (("untrained", None), ("random weights", w)):
        if weights is not None:
            pc.set_weights(weights)
        f1, f2, fit = fm.fermion(lat, chi, 1e-11, ittt0, 0.3, imass, precision="learned")
        e = max(np.abs(f1 - o1).max() / np.abs(o1).max(), np.abs(f2 - o2).max() / np.abs(o2).max())
        print("learned form, %s network: correlator error %.2e, iterations %d (oracle congrad %d)" % (
            tag, e, fit, oit))
        assert e <= 1e-6
    lat.close()


def test_fermion_learned_form_needs_a_network():
    lat = fm.Lattice(4, 2)
    with pytest.raises(ChsError):
        fm.fermion(lat, np.zeros((3, lat.kvol)), 1e-8, 1, precision="learned")
    lat.close()


def test_pairs_collected_from_propagator_solves():
    """The pair store fills from the plain-form propagator solves, ten per
    fermion call, up to its size, each pair with its chi; the stored source
    is M^dagger of the reference's first wall source and the stored solution
    solves M^dagger M x = source (both against the oracle)."""
    L, Ls, Lt, hpin, imass, ittt0 = 6, 4, 10, 0.3, 1, 3
    lat, olat = fm.Lattice(L, Ls, Lt), orc.Lattice(L, Ls, Lt)
    rs = np.random.RandomState(65)
    chi_a, chi_b = phi_field(rs, lat.kvol), phi_field(rs, lat.kvol)
    pc = learn.Preconditioner(lat, seed=2)
    pc.collect(14)
    assert pc.num_collected() == 0
    fm.fermion(lat, chi_a, 1e-11, ittt0, hpin, imass)
    assert pc.num_collected() == 10
    fm.fermion(lat, chi_b, 1e-11, ittt0, hpin, imass, precision="fp16")
    assert pc.num_collected() == 14
    fm.fermion(lat, chi_b, 1e-11, ittt0, hpin, imass)
    assert pc.num_collected() == 14
    # pair 0: source 1 (slice (ittt0 + Lt/5) mod Lt + 1), Dirac 3, flavour 1, 1/L^2 at s = 1
    ittt = (ittt0 + Lt // 5) % Lt + 1
    xi = np.zeros(lat.shape, np.complex128)
    xi[0, 2, ittt - 1, :, :, 0] = 1.0 / (L * L)
    phi = orc.dslash(olat, xi, chi_a, hpin, imass, True)
    x = orc.congrad(olat, phi, 1e-12, chi_a, hpin=hpin, imass=imass)[0]
    src, sol, chi0 = pc.pair(0)
    assert np.array_equal(chi0, chi_a)
    es = np.linalg.norm(src - phi) / np.linalg.norm(phi)
    ex = np.linalg.norm(sol - x) / np.linalg.norm(x)
    print("pair 0: source error %.2e, solution error %.2e" % (es, ex))
    assert es <= 1e-7 and ex <= 1e-5
    assert np.array_equal(pc.pair(13)[2], chi_b)
    with pytest.raises(ChsError):
        pc.pair(14)
    pc.collect(0)
    assert pc.num_collected() == 0
    lat.close()


def test_train_collected_equals_train_dev():
    """Training on the store (each pair with its own chi) makes the same
    updates as chs_nn_train_dev on the same pairs when they share one chi:
    weights and losses bit for bit."""
    from chiralheisenberg_amd import _lib
    lat = fm.Lattice(4, 2, 6)
    rs = np.random.RandomState(66)
    chi = phi_field(rs, lat.kvol)
    pc = learn.Preconditioner(lat, seed=5)
    pc.collect(10)
    fm.fermion(lat, chi, 1e-10, 2, 0.0, 1, precision="fp32")
    n = lat.kferm
    src, sol = fm.DeviceArray(lat, 10 * n * 8), fm.DeviceArray(lat, 10 * n * 8)
    for i in range(10):
        ps, px, pchi = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(lat.h, _lib.lib().chs_nn_pair_dev(lat.h, i, ctypes.byref(ps), ctypes.byref(px),
                                                      ctypes.byref(pchi)))
        for dst, p in ((src, ps), (sol, px)):
            _lib.check(lat.h, _lib.lib().chs_convert_dev(lat.h, ctypes.c_void_p(dst.ptr + i * n * 8),
                                                          _lib.CHS_FP32, p, _lib.CHS_FP32, n))
    l_c = pc.train_collected(epochs=2, lr=1e-3)
    w_c = pc.get_weights()
    pc2 = learn.Preconditioner(lat, seed=5)      # the same initial weights, fresh Adam state
    chid = fm.DeviceArray.from_numpy(lat, chi)
    l_d = pc2.train(src, sol, 10, chid, epochs=2, lr=1e-3)
    assert np.array_equal(l_c, l_d), (l_c, l_d)
    assert np.array_equal(w_c, pc2.get_weights())
    assert l_c[1] < l_c[0]
    lat.close()


def test_driver_run_with_learned_propagators(tmp_path):
    """rhmc.run with the learned preconditioner: the first measurement's ten
    FP64 propagator solves fill the store, the network is trained, the second
    measurement solves with it.  The Markov chain (control, fort.11) is the
    plain run's exactly; the first measurement's correlators are identical,
    the second's agree to the solver tolerance."""
    from chiralheisenberg_amd import checkpoint, rhmc
    L, Ls, iprint, iter2 = 4, 4, 2, 4
    rs = np.random.RandomState(7)
    chi0 = rs.randn(3, L ** 3) / np.sqrt(0.475)
    runs = {}
    for tag, learned in (("plain", None), ("learned", dict(pairs=10, epochs=2, lr=1e-3, seed=4))):
        d = tmp_path / tag
        d.mkdir()
        open(d / "midout", "w").write("0.3 0.475 1.0 0.00 0.2 1 3 %d\ndt beta am3 am h0 imass iterl iter2\n" % iter2)
        checkpoint.write_con(d / "con", chi0, 122110675517440.0)
        out = rhmc.run(str(d), L, Ls, iprint=iprint, rescgm=1e-10, learned=learned)
        runs[tag] = (d, out)
    (dp, _), (dl, out) = runs["plain"], runs["learned"]
    info = out["learned"]
    print("learned: %s" % info)
    assert info["pairs"] == 10 and info["trained_after"] == 2 and len(info["losses"]) == 2
    assert all(np.isfinite(info["losses"]))
    for f in ("control", "fort.11"):
        assert open(dp / f).read() == open(dl / f).read()
    for f in ("fort.5001", "fort.5011", "fort.5002", "fort.5012"):
        a = np.loadtxt(dp / f)[:, 1:]
        b = np.loadtxt(dl / f)[:, 1:]
        assert a.shape == b.shape == (2 * L, 2)
        assert np.array_equal(a[:L], b[:L])
        assert np.abs(a[L:] - b[L:]).max() <= 1e-6 * np.abs(a[L:]).max()
