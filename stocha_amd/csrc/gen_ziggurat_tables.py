#!/usr/bin/env python3
"""Generate the N=256 ziggurat tables for the standard normal distribution.

The table defines which normal a given 64-bit draw maps to, so it has to be the
same table the reference embeds (src/dist/ziggurat_tables.rs:6-526, produced by
scripts/gen_ziggurat_table.py:17-90).  This script restates the published
construction (Marsaglia & Tsang 2000, "The Ziggurat Method for Generating
Random Variables", J. Stat. Software 5(8)) in 50-digit mpmath:

  * f(x) = exp(-x^2/2), every layer has area V = r f(r) + int_r^inf f.
  * x_0 = r,  x_{i} = f^{-1}(f(x_{i-1}) + V / x_{i-1})  for i = 1..N-1,  x_N = 0.
  * r is fixed by bisection on [2, 5] (200 halvings) with the rule the
    reference uses: a layer that climbs past f(0)=1 means r is too small,
    and x_{N-1} > 1e-20 means r is too large.  That rule is why the last
    non-zero edge is x_255 ~ 1e-20 and layer 255 always rejects.
  * y_i = f(x_i).

Every value is emitted as a C99 hex-float literal, so the header holds the
exact binary64 values (no decimal round trip).

It also emits, per layer l, the integer acceptance threshold T[l]: the
reference's fast path (src/dist/normal.rs:21-29) accepts a 53-bit draw m when
fl(fl(m * 2^-53) * X[l]) < X[l+1].  The left side is monotone in m, so the
accepted set is exactly m < T[l]; T[l] is found here by binary search with
Python floats (IEEE binary64, round-to-nearest-even, like Rust's f64), which
makes the integer test bit-for-bit equivalent to the floating-point one and
lets the FP32 kernels consume the stream exactly as the FP64 ones do.

Usage:  python gen_ziggurat_tables.py OUT.h [OUT2.h ...]
"""
from __future__ import annotations

import sys

from mpmath import erfc, exp, log, mp, mpf, pi, sqrt

N = 256


def _f(x):
    return exp(-x * x / 2)


def _f_inv(y):
    return sqrt(-2 * log(y))


def _tail_area(r):
    return sqrt(2 * pi) / 2 * erfc(r / sqrt(2))


def _climb(r, v, n):
    """Return the layer edges x_0..x_{n-1} for a trial r, or None if a layer
    overshoots the peak (f value above 1)."""
    xs = [r]
    for _ in range(1, n):
        fi = _f(xs[-1]) + v / xs[-1]
        if fi > 1:
            return None
        xs.append(_f_inv(fi))
    return xs


def solve_r(n: int = N):
    mp.dps = 50
    lo, hi = mpf(2), mpf(5)
    for _ in range(200):
        r = (lo + hi) / 2
        v = r * _f(r) + _tail_area(r)
        xs = _climb(r, v, n)
        if xs is None:
            lo = r
        elif xs[n - 1] > mpf("1e-20"):
            hi = r
        else:
            lo = r
    return (lo + hi) / 2


def build_tables(n: int = N):
    """Return (r, X[0..n], Y[0..n]) as Python floats (binary64)."""
    mp.dps = 50
    r = solve_r(n)
    v = r * _f(r) + _tail_area(r)
    xs = [r]
    for _ in range(1, n):
        xs.append(_f_inv(_f(xs[-1]) + v / xs[-1]))
    xs.append(mpf(0))
    ys = [_f(x) for x in xs]
    return float(r), [float(x) for x in xs], [float(y) for y in ys]


def fast_thresholds(X):
    """T[l] = number of 53-bit m with fl(fl(m*2^-53)*X[l]) < X[l+1]."""
    scale = 1.0 / float(1 << 53)
    out = []
    for l in range(len(X) - 1):
        xl, xn = X[l], X[l + 1]

        def accepted(m):
            return (float(m) * scale) * xl < xn

        # accepted() is true on [0, T) and false on [T, 2^53): find T.
        lo, hi = 0, 1 << 53
        while lo < hi:
            mid = (lo + hi) // 2
            if accepted(mid):
                lo = mid + 1
            else:
                hi = mid
        out.append(lo)
    return out


def render(r, X, Y, T) -> str:
    lines = [
        "/* Ziggurat tables for N(0,1), 256 layers.  GENERATED by",
        " * gen_ziggurat_tables.py -- do not edit.  Values are exact binary64",
        " * hex floats; ZIG_T[l] is the integer form of the fast-path test",
        " * fl(fl(m*2^-53)*X[l]) < X[l+1]  <=>  m < ZIG_T[l]  (m = u >> 11). */",
        "#ifndef STOCHA_ZIG_TABLES_H",
        "#define STOCHA_ZIG_TABLES_H",
        "#define ZIG_LAYERS 256",
        f"#define ZIG_R_VALUE {r.hex()}",
        "#define ZIG_X_VALUES \\",
    ]
    for i, x in enumerate(X):
        lines.append(f"    {x.hex()}{',' if i < len(X) - 1 else ''} \\")
    lines.append("")
    lines.append("#define ZIG_Y_VALUES \\")
    for i, y in enumerate(Y):
        lines.append(f"    {y.hex()}{',' if i < len(Y) - 1 else ''} \\")
    lines.append("")
    lines.append("#define ZIG_T_VALUES \\")
    for i, t in enumerate(T):
        lines.append(f"    {t}ull{',' if i < len(T) - 1 else ''} \\")
    lines.append("")
    lines.append("#endif")
    return "\n".join(lines) + "\n"


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    r, X, Y = build_tables()
    T = fast_thresholds(X)
    text = render(r, X, Y, T)
    for path in argv[1:]:
        with open(path, "w") as fh:
            fh.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
