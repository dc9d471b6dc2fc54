"""The reference's configuration checkpoint `con` (rhmc/cheisdwf.f:1455-1492,
sread / swrite; FP32 -> FP64 conversion rhmc/sp_to_dp.f).

The file is one Fortran unformatted sequential record: a 4-byte
little-endian length, chi(kvol, 3) (column-major, so the three components
one after the other) followed by the real*8 random-number seed, and the same
4-byte length again.  chi is real*8 in the FP64 build (record 24*kvol + 8
bytes, 41,488-byte file at L = 12) and real*4 in the shipped single-precision
build (12*kvol + 8 bytes, 20,752 bytes at L = 12); the seed is real*8 in both.
"""
import os
import struct

import numpy as np

DEFAULT_SEED = 693764523745.0   # sread's value when no con file exists (rhmc/cheisdwf.f:1468)


def read_con(path, kvol=None, precision=None):
    """Returns (chi float64 (3, kvol), seed, precision "fp64" or "fp32").
    The record length fixes 24*kvol (FP64) or 12*kvol (FP32); when both
    readings are possible (12*kvol is a multiple of 24 for even kvol), kvol or
    precision must say which."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 16:
        raise ValueError("%s: too short for a con record" % path)
    (n,) = struct.unpack("<i", data[:4])
    (m,) = struct.unpack("<i", data[-4:])
    if n != m or n != len(data) - 8:
        raise ValueError("%s: record markers %d/%d do not frame %d bytes" % (path, n, m, len(data) - 8))
    body = n - 8
    fits = [p for p, w in (("fp64", 24), ("fp32", 12))
            if body % w == 0 and (kvol is None or body == w * kvol) and precision in (None, p)]
    if not fits:
        raise ValueError("%s: record of %d bytes fits no chi(%s, 3) in %s" % (path, n, kvol, precision or "fp64/fp32"))
    if len(fits) > 1:
        raise ValueError("%s: record of %d bytes is FP64 chi(%d,3) or FP32 chi(%d,3); give kvol or precision" % (
            path, n, body // 24, body // 12))
    prec = fits[0]
    w, dt = (24, "<f8") if prec == "fp64" else (12, "<f4")
    chi = np.frombuffer(data[4:4 + body], dtype=dt).astype(np.float64).reshape(3, body // w)
    (seed,) = struct.unpack("<d", data[4 + body:4 + body + 8])
    return chi, seed, prec


def write_con(path, chi, seed, precision="fp64"):
    """swrite (FP64) or the shipped FP32 layout."""
    chi = np.asarray(chi, dtype=np.float64)
    if chi.ndim != 2 or chi.shape[0] != 3:
        raise ValueError("chi must have shape (3, kvol)")
    dt = {"fp64": "<f8", "fp32": "<f4"}[precision]
    body = np.ascontiguousarray(chi, dtype=dt).tobytes() + struct.pack("<d", float(seed))
    rec = struct.pack("<i", len(body))
    with open(path, "wb") as f:
        f.write(rec + body + rec)


def sread(path, kvol, beta):
    """sread: the checkpoint if it exists, else chi = 0.1/beta everywhere and
    the default seed (rhmc/cheisdwf.f:1455-1470)."""
    if os.path.exists(path):
        chi, seed, _ = read_con(path, kvol)
        return chi, seed
    return np.full((3, kvol), 0.1 / beta), DEFAULT_SEED


def sp_to_dp(src, dst):
    """rhmc/sp_to_dp.f: rewrite an FP32 con file as FP64 (the values widen
    exactly)."""
    chi, seed, _ = read_con(src, precision="fp32")
    write_con(dst, chi, seed, "fp64")
