"""Learned preconditioner for M^dagger M solves (host side, over libchs).

The network lives on the GPU (chs_nn_* entry points): a linear, symmetric
positive definite 3-D convolutional net on the Ls-stacked spinor,
P(r) = c u + C1^T C2^T C2 C1 u (u = r / rms(r); C1 1x1x1, 16 Ls -> 16
channels; C2 3x3x3 over x, y, t, 16 -> 16), BF16 MFMA kernels for inference
and training.  Training data are (source, solution) pairs of M^dagger M x = b
produced by the plain solver, the loss the A-norm of P's error; the trained
network is the preconditioner of flexible CG, whose answer is accepted on the
FP64 true residual like every other solve (plain CG remains the fallback).

Pairs come either from pairs_dev (Gaussian sources at one chi) or from the
propagator solves themselves: Preconditioner.collect opens a device store
that every plain-form fermion.fermion call fills (each pair with the chi it
was solved with), train_collected trains on it, and fermion.fermion(...,
precision="learned") then solves with the network.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from . import fermion as fm


class Preconditioner:
    def __init__(self, lat, hidden=16, seed=1234):
        self.lat = lat
        check(lat.h, lib().chs_nn_init(lat.h, hidden, seed))
        self.nweights = lib().chs_nn_num_weights(lat.h)

    def get_weights(self):
        w = np.empty(self.nweights, dtype=np.float32)
        check(self.lat.h, lib().chs_nn_get_weights(self.lat.h, w.ctypes.data_as(ctypes.c_void_p), self.nweights))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nweights
        check(self.lat.h, lib().chs_nn_set_weights(self.lat.h, w.ctypes.data_as(ctypes.c_void_p), self.nweights))

    def apply(self, r_dev, z_dev):
        """z = P(r) on FP32 native device spinors."""
        check(self.lat.h, lib().chs_nn_apply_dev(self.lat.h, r_dev.ptr, z_dev.ptr))

    def loss_grad(self, src_dev, sol_dev, chi_dev, hpin=0.0, imass=1):
        """Relative A-norm loss of one pair and the gradient of all parameters
        (A = M^dagger M of the operator given by chi_dev, hpin, imass)."""
        loss = ctypes.c_double()
        g = np.empty(self.nweights, dtype=np.float32)
        check(self.lat.h, lib().chs_nn_loss_grad_dev(self.lat.h, src_dev.ptr, sol_dev.ptr, chi_dev.ptr, float(hpin),
                                                      int(imass), ctypes.byref(loss),
                                                      g.ctypes.data_as(ctypes.c_void_p)))
        return loss.value, g

    def train(self, src_dev, sol_dev, npairs, chi_dev, epochs=1, lr=1e-3, hpin=0.0, imass=1):
        """Adam over npairs pairs stored contiguously (FP32 native spinors);
        returns the mean loss of each epoch."""
        out = np.zeros(epochs)
        check(self.lat.h, lib().chs_nn_train_dev(self.lat.h, src_dev.ptr, sol_dev.ptr, int(npairs), int(epochs),
                                                  float(lr), chi_dev.ptr, float(hpin), int(imass),
                                                  out.ctypes.data_as(ctypes.c_void_p)))
        return out


    def collect(self, max_pairs):
        """Open a device store for max_pairs (source, solution) pairs of the
        propagator solves of fermion.fermion in a plain form (emptying any
        earlier one); 0 frees it."""
        check(self.lat.h, lib().chs_nn_collect(self.lat.h, int(max_pairs)))

    def num_collected(self):
        return lib().chs_nn_num_collected(self.lat.h)

    def pair(self, i):
        """Stored pair i as host arrays: (source, solution) complex128 in the
        reference layout (FP32 values) and chi (3, kvol)."""
        ps, px, pc = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(self.lat.h, lib().chs_nn_pair_dev(self.lat.h, int(i), ctypes.byref(ps), ctypes.byref(px),
                                                 ctypes.byref(pc)))
        out = [fm.to_host(self.lat, _Fp32View(p.value)) for p in (ps, px)]
        chi = np.empty((3, self.lat.kvol))
        check(self.lat.h, lib().chs_download(self.lat.h, chi.ctypes.data_as(ctypes.c_void_p), pc.value, chi.nbytes))
        return out[0], out[1], chi

    def train_collected(self, epochs=1, lr=1e-3, hpin=0.0, imass=1):
        """Adam over the stored pairs, each with its own chi; returns the mean
        loss of each epoch."""
        out = np.zeros(epochs)
        check(self.lat.h, lib().chs_nn_train_collected(self.lat.h, int(epochs), float(lr), float(hpin), int(imass),
                                                        out.ctypes.data_as(ctypes.c_void_p)))
        return out


class _Fp32View:
    """A non-owning FP32 native spinor at a device address (for fm.to_host)."""

    def __init__(self, ptr):
        self.ptr, self.prec = ptr, _lib.CHS_FP32


def solve_fcg(lat, b_dev, x_dev, chi_dev, tol=1e-10, use_nn=True, hpin=0.0, imass=1, delta=0.1, maxiter=0):
    """Flexible CG on M^dagger M x = b (FP64 native device arrays) to FP64 true
    relative residual tol."""
    st = _lib.SolveStats()
    check(lat.h, lib().chs_solve_mdagm_fcg_dev(lat.h, b_dev.ptr, x_dev.ptr, chi_dev.ptr, hpin, imass, tol, delta,
                                                maxiter, int(bool(use_nn)), ctypes.byref(st)))
    return {f: getattr(st, f) for f, _ in _lib.SolveStats._fields_}


def pairs_dev(lat, chi_dev, npairs, rng_seed=11, tol=1e-10, precision="fp32", kind="solve", hpin=0.0, imass=1,
              progress=None):
    """Generate npairs (source, solution) pairs of M^dagger M x = b.
    kind "solve": Gaussian b, x from the plain solver; kind "apply": Gaussian
    x, b = M^dagger M x (exact, no solve).  Returns (src, sol) FP32 native
    device arrays (npairs spinors each) and the solver time in seconds.
    progress(i, seconds): called after every 100th pair, if given."""
    if kind not in ("solve", "apply"):
        raise ValueError("kind must be 'solve' or 'apply'")
    n = lat.kferm
    src = fm.DeviceArray(lat, npairs * n * 8)
    sol = fm.DeviceArray(lat, npairs * n * 8)
    rs = np.random.RandomState(rng_seed)
    x = fm.spinor_dev(lat, None)
    t = fm.spinor_dev(lat, None)
    zero = np.zeros(lat.shape, np.complex128)
    seconds = 0.0
    for i in range(npairs):
        gh = (rs.randn(*lat.shape) + 1j * rs.randn(*lat.shape)) / np.sqrt(2.0)
        if kind == "solve":
            bd = fm.spinor_dev(lat, gh)
            x.upload(zero)
            st = fm.solve_mdagm(lat, bd, x, chi_dev, tol, precision, hpin=hpin, imass=imass)
            seconds += st["seconds"]
        else:
            bd = fm.spinor_dev(lat, None)
            x.free()
            x = fm.spinor_dev(lat, gh)
            fm.dslash_dev(lat, t, x, chi_dev, hpin, imass, dagger=False)
            fm.dslash_dev(lat, bd, t, chi_dev, hpin, imass, dagger=True)
        for dst, srcv in ((src, bd), (sol, x)):
            check(lat.h, lib().chs_convert_dev(lat.h, ctypes.c_void_p(dst.ptr + i * n * 8), _lib.CHS_FP32, srcv.ptr,
                                               _lib.CHS_FP64, n))
        bd.free()
        if progress is not None and (i + 1) % 100 == 0:
            progress(i + 1, seconds)
    lat.synchronize()
    return src, sol, seconds
