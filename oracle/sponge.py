"""ORACLE (test infrastructure only): the sponge-layer coefficients restated.

Only tests/ use this module, as the checker of the library's
gcomcot_sponge_coef_; the product never imports it.  SPONGE_COEF
(boundaries.f90:250-330) with H_CALC (:334-365) in REAL: the mean wet depth
summed in the Fortran loop order (I outer, J inner), the characteristic
wave length and frequency, and the Kirby et al. (1998) damping profile over
a band of twice the wave length from each edge.  COEFX depends on X(I) only
and COEFY on Y(J) only.  EXP is taken in FP64 and rounded to REAL.  For
spherical layers the Fortran measures the band in degrees but divides by a
width in metres, so every coefficient is 0; the restatement keeps that.
Pinned bit for bit against the flang-built SPONGE_COEF in
tests/test_boundaries_oracle.py.
"""
import math

import numpy as np

F = np.float32
GX = F(1.0e-5)
EPS = F(1.0e-10)
GRAV = F(9.807)
PI = F(3.14159265358979)
R_EARTH = F(6378000.0)


def _exp(v):
    return F(math.exp(float(v)))


def sponge_coef(laycord, h, x, y):
    """(cx[nx], cy[ny]) for a layer of depth h (ny, nx) and coordinates x, y."""
    h = np.asarray(h, F)
    wet = h.T[h.T > GX]                 # I outer, J inner
    k = wet.size
    h_mean = np.cumsum(wet, dtype=F)[-1] / F(k) if k > 0 else F(0.0)
    depth = F(h_mean)
    if depth <= GX:
        depth = GX
    wavelength = F(20.0) * depth
    width = F(2.0) * wavelength
    alpha_c = F(10.0)
    t = wavelength / F(math.sqrt(float(GRAV * depth)))
    omega = F(2.0) * PI / t
    sponge_width = (width / R_EARTH * F(180.0) / PI) if laycord == 0 else width
    r_max = width
    x = np.asarray(x, F)
    y = np.asarray(y, F)
    e1 = _exp(F(1.0)) - F(1.0)

    def coef(v, lo, hi):
        r = F(0.0)
        if v <= lo:
            r = -(v - lo)
        elif v >= hi:
            r = v - hi
        rel = r
        if rel <= F(0.0):
            rel = F(0.0)
        if rel >= r_max:
            rel = r_max
        q = rel / r_max
        c = alpha_c * omega * (_exp(q * q) - F(1.0)) / e1
        return F(0.0) if c <= EPS else F(c)

    xs, ys = x[0] + sponge_width, y[0] + sponge_width
    xe, ye = x[-1] - sponge_width, y[-1] - sponge_width
    cx = np.array([coef(v, xs, xe) for v in x], F)
    cy = np.array([coef(v, ys, ye) for v in y], F)
    return cx, cy
