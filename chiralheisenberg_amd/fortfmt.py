"""Fortran edit descriptors for the drivers' `output` files, which the
reference programs write with formatted WRITE statements
(rhmc/cheisdwf.f:167-182, :439-462; mh/cheisdwf_hmc.f:103-117, :364-385):
Iw, Fw.d, Ew.d, nX and the list-directed output of a REAL*8 as the flang
build of the reference prints it."""
import math
from fractions import Fraction


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


def _shortest_digits(a):
    """Digits K and exponent x (a ~ K * 10**x) of a positive finite double
    as the flang runtime's list-directed output chooses them: the shortest
    length n at which some n-digit numbers lie strictly between the
    midpoints to the neighbouring doubles, lo and hi, and of those the middle
    one (the lower of two middles).  This is not always the correctly rounded
    shortest form: 21.872546690894523 where the exact value is ...5236."""
    v = Fraction(a)
    lo = (v + Fraction(math.nextafter(a, 0.0))) / 2
    hi = v + (v - lo) if math.isinf(math.nextafter(a, math.inf)) else (v + Fraction(math.nextafter(a, math.inf))) / 2
    e = len(str(hi.numerator // hi.denominator)) - 1 if hi >= 1 else -len(str(hi.denominator // hi.numerator))
    while Fraction(10) ** e > hi:
        e -= 1
    while Fraction(10) ** (e + 1) <= hi:
        e += 1
    for n in range(1, 40):
        q = Fraction(10) ** (e - n + 1)
        first, last = lo // q + 1, -((-hi) // q) - 1      # the n-digit numbers strictly inside
        if first <= last:
            k, x = (first + last) // 2, e - n + 1
            while k % 10 == 0:
                k //= 10
                x += 1
            return k, x
    raise ValueError(a)


def list_real(v):
    """A REAL*8 as the flang build's list-directed output prints it: the
    shortest digits that read back to v (chosen as _shortest_digits does);
    positional for 0.1 <= |v| < 1e15 without the zero before the point
    ('.15', '15.', '3604080.59'), else d.dddE+xx ('1.5E-08', '1.E-05');
    zero as '0.'."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    a = abs(v)
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    if a == 0.0:
        return sign + "0."
    k, x = _shortest_digits(a)
    d = str(k)
    if Fraction(95, 1000) <= Fraction(a) < Fraction(95 * 10 ** 13):
        point = len(d) + x                     # digits before the decimal point
        if point <= 0:
            return sign + "." + "0" * (-point) + d
        if point >= len(d):
            return sign + d + "0" * (point - len(d)) + "."
        return sign + d[:point] + "." + d[point:]
    return "%s%s.%sE%+03d" % (sign, d[0], d[1:], len(d) - 1 + x)


def list_directed(*items):
    """One record of list-directed output (flang): every item after one
    blank; integers in their digits, REAL*8 by list_real."""
    out = []
    for it in items:
        if isinstance(it, (int,)) or (hasattr(it, "dtype") and getattr(it.dtype, "kind", "") in "iu"):
            out.append(" %d" % int(it))
        else:
            out.append(" " + list_real(it))
    return "".join(out)
