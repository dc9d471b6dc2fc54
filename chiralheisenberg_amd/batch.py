"""Batched multi-shift solves (host side, over libchs's chs_qmrherm_batch_dev).

Many qmrherm solves (rhmc/cheisdwf.f:599-745, iflag = 0 and the force's
iflag = 1 mixed) -- several pseudofermion fields, lattices of different
sizes -- run in lock step, one grouped launch per Lanczos stage, chunks of
steps replayed from hipGraphs, launch parameters chosen by the library's
autotuner.  Each result equals the one-at-a-time chs_qmrherm_dev result bit
for bit."""
import ctypes

import numpy as np

from . import _lib
from ._lib import ChsError, lib


class System:
    """One solve: lat (fermion.Lattice), phi / chi / x device arrays (FP64
    native layout), remez coefficients (anum, aden) and qmrherm's res.  With
    iflag = 1 (the force's solve) F is a (3, kvol) FP64 device array to which
    the shifts' derivs contributions are added, as by qmrherm_dev."""

    def __init__(self, lat, phi, chi, x, anum, aden, res, hpin=0.0, imass=1, iflag=0, F=None):
        self.lat, self.phi, self.chi, self.x = lat, phi, chi, x
        self.anum = np.ascontiguousarray(anum, dtype=np.float64)
        self.aden = np.ascontiguousarray(aden, dtype=np.float64)
        assert self.anum.size == self.aden.size + 1
        if iflag not in (0, 1) or (iflag == 1 and F is None):
            raise ValueError("iflag must be 0, or 1 with F")
        self.res, self.hpin, self.imass = float(res), float(hpin), int(imass)
        self.iflag, self.F = int(iflag), F
        self.itercg = None


class Batch:
    def __init__(self, device=0):
        self.h = lib().chs_batch_create(device)
        if not self.h:
            raise ChsError(_lib.CHS_EHIP, lib().chs_create_error().decode())

    def _check(self, rc):
        if rc != _lib.CHS_OK:
            raise ChsError(rc, lib().chs_batch_last_error(self.h).decode())

    def qmrherm(self, systems, tune=None, precision="fp64"):
        """Solve every system with its spinors stored in precision ("fp64" or
        "fp32"); sets System.itercg and returns the batch statistics (dict;
        'tune' holds the launch parameters used)."""
        arr = (_lib.MsSystem * len(systems))()
        for a, s in zip(arr, systems):
            a.ctx, a.phi, a.chi, a.x = s.lat.h, s.phi.ptr, s.chi.ptr, s.x.ptr
            a.hpin, a.imass, a.ndiag, a.res = s.hpin, s.imass, s.aden.size, s.res
            a.anum = s.anum.ctypes.data_as(ctypes.c_void_p)
            a.aden = s.aden.ctypes.data_as(ctypes.c_void_p)
            a.iflag = s.iflag
            a.dsdchi = s.F.ptr if s.F is not None else None
        t = None
        if tune is not None:
            t = _lib.BatchTune(int(tune["ept"]), int(tune["chunk"]), int(tune["graph"]), 0.0)
        st = _lib.BatchStats()
        self._check(lib().chs_qmrherm_batch_dev(self.h, arr, len(systems), _lib.PRECISIONS[precision],
                                                 ctypes.byref(t) if t else None, ctypes.byref(st)))
        for a, s in zip(arr, systems):
            s.itercg = a.itercg
        tu = st.tune
        return {"iterations": st.iterations, "graph_replays": st.graph_replays, "launches": st.launches,
                "seconds": st.seconds,
                "tune": {"ept": tu.ept, "chunk": tu.chunk, "graph": tu.graph, "us_per_iter": tu.us_per_iter}}

    def force(self, fields, beta, res1, rz, precision="fp64", tune=None):
        """The pseudofermion force (rhmc/cheisdwf.f:470-518) of several fields
        and lattices at once: X = r_b(M^dagger M) Phi for every field in one
        batch (iflag = 0, the b-set of rz = remez.load("remez2g")), then the
        a-set solves from X with iflag = 1 in a second batch, adding to
        F = 2 beta chi.  fields: (lat, phi_dev, chi_dev, chi_host) tuples.
        Returns ([F device arrays], [X device arrays], (stats1, stats2)); each
        F equals chs_force's for the same field bit for bit."""
        from . import fermion as fm
        anum, aden, bnum, bden = rz[:4]
        xs, s1 = [], []
        for lat, phi, chid, _ in fields:
            X = fm.spinor_dev(lat, None)
            s1.append(System(lat, phi, chid, X, bnum, bden, res1))
            xs.append(X)
        st1 = self.qmrherm(s1, tune, precision)
        Fs, s2 = [], []
        for (lat, _, chid, chih), X in zip(fields, xs):
            F = fm.DeviceArray.from_numpy(lat, 2.0 * beta * np.asarray(chih, dtype=np.float64))
            Fs.append(F)
            s2.append(System(lat, X, chid, fm.spinor_dev(lat, None), anum, aden, res1, iflag=1, F=F))
        st2 = self.qmrherm(s2, tune, precision)
        return Fs, xs, (st1, st2)

    def close(self):
        if self.h:
            lib().chs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
