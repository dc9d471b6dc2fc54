"""Fortran edit descriptors for the drivers' `output` files, which the
reference programs write with formatted WRITE statements
(rhmc/cheisdwf.f:167-182, :439-462; mh/cheisdwf_hmc.f:103-117, :364-385):
Iw, Fw.d, Ew.d, nX and the list-directed output of a REAL*8 as the flang
build of the reference prints it."""
import math


def i(v, w):
    """Iw: right-justified integer; '*' * w when it does not fit."""
    s = "%d" % v
    return s.rjust(w) if len(s) <= w else "*" * w


def _special(v, w):
    if math.isnan(v):
        return "NaN".rjust(w)
    return ("Inf" if v > 0 else "-Inf").rjust(w)


def f(v, w, d):
    """Fw.d: fixed point with d decimals; the zero before the point is
    dropped when only that makes the field fit."""
    if not math.isfinite(v):
        return _special(v, w)
    s = "%.*f" % (d, v)
    if len(s) > w and s.startswith(("0.", "-0.")):
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def e(v, w, d):
    """Ew.d: 0.d1..dd E+xx (the exponent field as E+xx up to 99, +xxx
    beyond); the zero before the point is dropped when only that makes the
    field fit."""
    if not math.isfinite(v):
        return _special(v, w)
    if v == 0.0:
        mant, ex = "0" * d, 0
    else:
        m, x = ("%.*e" % (d - 1, abs(v))).split("e")
        mant, ex = m.replace(".", ""), int(x) + 1
    exs = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
    sign = "-" if v < 0 or (v == 0.0 and math.copysign(1.0, v) < 0) else ""
    s = sign + "0." + mant + exs
    if len(s) > w:
        s = sign + "." + mant + exs
    return s.rjust(w) if len(s) <= w else "*" * w


def list_real(v):
    """A REAL*8 in list-directed output: the shortest digits that read back
    to v, an integral value with a trailing point (' 3390816911360.')."""
    if math.isfinite(v) and v == int(v) and abs(v) < 1e16:
        return "%d." % int(v)
    r = repr(float(v))
    if "e" in r:
        m, x = r.split("e")
        return "%sE%+03d" % (m, int(x))
    return r
