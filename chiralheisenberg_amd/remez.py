"""Rational (Remez) approximations: partial fractions a0 + sum_k a_k/(x + s_k)
for x^{+p} (the 'a' set) and x^{-p} (the 'b' set), in the reference's text
format: a0, then n lines 'a_k, s_k', then b0, then n lines 'b_k, s_k'
(loader rhmc/cheisdwf.f:140-159).  remez2 and remez4 hold 25 poles for
p = 1/2 and p = 1/4; remez2g holds 12 poles for p = 1/2 (the force's
'guidance' set).  The files are the reference's own data (rhmc/remez*)."""
import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
POWERS = {"remez2": 0.5, "remez4": 0.25, "remez2g": 0.5}


def load(name_or_path):
    """Returns (anum, aden, bnum, bden) as float64 arrays (anum, bnum have n+1)."""
    path = name_or_path if os.path.sep in name_or_path else os.path.join(DATA, name_or_path)
    vals = [float(t) for t in open(path).read().replace(",", " ").split()]
    if (len(vals) - 2) % 4:
        raise ValueError("%s: malformed coefficient file" % path)
    n = (len(vals) - 2) // 4
    a0, rest = vals[0], vals[1:]
    a, ad = rest[0:2 * n:2], rest[1:2 * n:2]
    b0 = rest[2 * n]
    b, bd = rest[2 * n + 1::2], rest[2 * n + 2::2]
    return (np.array([a0] + a), np.array(ad), np.array([b0] + b), np.array(bd))


def load_sets(workdir=None):
    """The three sets of a run: the caller's own files remez2, remez4 and
    remez2g in workdir when they are there -- dwf3d reads them from its
    working directory (rhmc/cheisdwf.f:114-116, :140-159) -- else the copies
    bundled here."""
    out = {}
    for k in ("remez2", "remez4", "remez2g"):
        p = os.path.join(workdir, k) if workdir else None
        out[k] = load(p if p and os.path.exists(p) else k)
    return out


def evaluate(coef, x):
    """Value of the partial fraction a0 + sum a_k / (x + s_k) at x."""
    num, den = coef
    x = np.asarray(x, dtype=np.float64)[..., None]
    return num[0] + np.sum(num[1:] / (x + den), axis=-1)
