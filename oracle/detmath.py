# -*- coding: utf-8 -*-
"""Deterministic pow and log shared by the HIP kernels and the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): this is the oracle's
copy; wdsgpu_amd/csrc/detmath.h is the kernels' copy of the same sequence of
IEEE-754 double operations.

The reference evaluates its power laws with the platform libm (glibc or
msvcrt pow/log, dgga/solver.py:58-69), whose last bit differs between
platforms and between CPU and GPU libraries; the 1-ulp difference is then
amplified by the Newton iteration.  This build fixes one algorithm for both
of its own paths instead, so that the oracle and the GPU compute the same
link coefficients bit for bit:

  log x   x = m 2^k with m in [sqrt(1/2), sqrt(2)) (exact), s = (m - 1) /
          (m + 1) in double-double, log m = 2 atanh s = 2 s + 2 s^3 P(s^2)
          with P the atanh series to s^22, k ln 2 with ln 2 split in a
          32-bit head and a tail (Cody & Waite, Software Manual for the
          Elementary Functions, 1980), all in double-double (Dekker 1971:
          TwoSum, TwoProduct).
  exp z   z in double-double, n = rint(z / ln 2), r = z - n ln 2 (Cody-Waite,
          exact head), e^r = 1 + r + r^2/2 + r^3/6 + r^4 R(r) with the first
          four terms in double-double and the Taylor tail R to r^14 in
          double, times 2^n.
  pow     x^y = exp(y log x) for x > 0 with y log x in double-double;
          y == 1 returns x and y == 0 returns 1 exactly.

Every rounding is a plain IEEE double operation in a fixed order (no fused
multiply-add outside TwoProduct, whose exact error term is the same whether
it comes from an FMA, as in the kernels, or from Veltkamp splitting, as
here), so the two copies agree bit for bit.  Accuracy: within one ulp of
the correctly rounded value (tested against libm)."""
from __future__ import annotations

import numpy as np

LN2_HI = 6.93147180369123816490e-01      # 0x3fe62e42fee00000 (32 significant bits)
LN2_LO = 1.90821492927058770002e-10      # 0x3dea39ef35793c76
INV_LN2 = 1.44269504088896338700e+00     # 0x3ff71547652b82fe
SPLIT = 134217729.0                      # 2^27 + 1 (Veltkamp)
# atanh series 1/3, 1/5, ..., 1/23 (double roundings of the exact reciprocals)
ATANH = [1.0 / (2 * i + 3) for i in range(11)]
# Taylor tail of e^r: 1/3!, 1/4!, ..., 1/14!
EXPC = []
_f = 6.0
for _i in range(3, 15):
    EXPC.append(1.0 / _f)
    _f *= _i + 1


def _two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _two_prod(a, b):
    p = a * b
    t = SPLIT * a
    ah = t - (t - a)
    al = a - ah
    t = SPLIT * b
    bh = t - (t - b)
    bl = b - bh
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def log_dd(x):
    """log x for x > 0 finite as a double-double (hi, lo)."""
    x = np.asarray(x, dtype=np.float64)
    m, k = np.frexp(x)                    # x = m 2^k, m in [0.5, 1)
    m = m * 2.0
    k = k - 1
    low = m < 1.4142135623730951          # sqrt(2): keep m in [sqrt(1/2), sqrt(2))
    m = np.where(low, m, m * 0.5)
    k = np.where(low, k, k + 1).astype(np.float64)
    f = m - 1.0                           # exact (Sterbenz)
    u_hi, u_lo = _two_sum(m, 1.0)         # m + 1
    s_hi = f / u_hi
    p, pe = _two_prod(s_hi, u_hi)
    s_lo = (((f - p) - pe) - s_hi * u_lo) / u_hi
    z = s_hi * s_hi
    P = ATANH[10]
    for c in reversed(ATANH[:10]):
        P = P * z + c
    tail = (2.0 * s_hi) * z * P           # 2 s^3 P(s^2)
    # k ln2_hi + 2 s (both exact heads) + the small terms
    a_hi, a_lo = _two_sum(k * LN2_HI, 2.0 * s_hi)
    lo = ((k * LN2_LO + 2.0 * s_lo) + tail) + a_lo
    return _two_sum(a_hi, lo)


def det_log(x):
    hi, lo = log_dd(x)
    return hi + lo


def exp_dd(z_hi, z_lo):
    """e^(z_hi + z_lo), |z_hi| <= 700."""
    n = np.rint(z_hi * INV_LN2)
    a = z_hi - n * LN2_HI                 # exact (Cody-Waite head)
    r_hi, r_lo = _two_sum(a, z_lo - n * LN2_LO)
    h_hi, h_lo = _two_prod(r_hi, r_hi)    # r^2
    h_lo = h_lo + 2.0 * r_hi * r_lo
    h_hi, h_lo = h_hi * 0.5, h_lo * 0.5   # r^2 / 2
    c_hi, c_lo = _two_prod(r_hi, h_hi)    # r^3 / 2
    c_lo = c_lo + r_hi * h_lo
    t_hi = c_hi / 3.0                     # r^3 / 6 in double-double
    p, pe = _two_prod(t_hi, 3.0)
    t_lo = (((c_hi - p) - pe) + c_lo) / 3.0
    R = EXPC[-1]
    for c in reversed(EXPC[1:-1]):
        R = R * r_hi + c
    t4 = (h_hi * h_hi) * (4.0 * R)        # r^4 R(r), R = 1/4! + r/5! + ...
    # 1 + r + r^2/2 + r^3/6 + r^4 R
    s_hi, s_lo = _two_sum(r_hi, h_hi)
    u_hi, u_lo = _two_sum(s_hi, t_hi)
    s_lo = ((((s_lo + u_lo) + r_lo) + h_lo) + t_lo) + t4
    s_hi, s_lo = _two_sum(u_hi, s_lo)
    o_hi, o_lo = _two_sum(1.0, s_hi)
    return np.ldexp(o_hi + (o_lo + s_lo), n.astype(np.int64))


def det_pow(x, y):
    """x^y for x >= 0 (finite); y == 1 gives x, y == 0 gives 1, and a zero
    base gives 0 (y > 0) or inf (y < 0)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xs = np.where(x > 0.0, x, 1.0)
    l_hi, l_lo = log_dd(xs)
    p_hi, p_lo = _two_prod(y, l_hi)
    p_lo = p_lo + y * l_lo
    z_hi, z_lo = _two_sum(p_hi, p_lo)
    out = exp_dd(z_hi, z_lo)
    out = np.where(x == 0.0, np.where(y > 0.0, 0.0, np.inf), out)
    out = np.where(y == 1.0, x, out)
    return np.where(y == 0.0, 1.0, out)
