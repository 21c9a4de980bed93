"""ORACLE (test infrastructure only): the reference's fault-source model restated.

Only tests/ use this module, as the checker of the library's device
deformation; the product never imports it.  It restates, vectorised with
numpy:

- STEREO_PROJECTION (deform.f90:494-588), the oblique stereographic map
  of (lon, lat) onto the plane tangent at the epicentre, WGS84 constants;
- DEFORM_OKADA (deform.f90:74-191) with STRIKE_SLIP (:194-217) and
  DIP_SLIP (:221-238): the vertical seafloor displacement of one
  rectangular fault (Okada 1985) at every cell of a layer;
- CELL_INTERP (all_grids.f90:859-934): the bilinear spreading of a
  parent's deformation onto a nested layer, as ININTERP (:805-855) applies
  it level by level; CELL_INTERP_SC (:938-992) for a Cartesian child with
  SC_OPTION = 1, through its POS/CXY.

The fault model is evaluated in FP64 with the reference's REAL constants
(each literal rounded to float first, as the Fortran holds it; the
projection's NU0 truncated to an integer, since implicit typing makes it
INTEGER).  The reference evaluates it in FP32, so the two agree to the
reference's own rounding error; tests/test_okada_oracle.py pins this
restatement against the flang-built DEFORM_OKADA within a bound derived
from that error, and checks exact properties (linearity in the slip, zero
slip, the strike-slip/dip-slip split).  CELL_INTERP is REAL arithmetic
restated operation by operation and is pinned bit for bit against the
reference's own CELL_INTERP.
"""
import numpy as np


def _f(v):
    """A REAL literal as the Fortran holds it."""
    return float(np.float32(v))


RAD_DEG = _f(0.01745329252)
EPS = _f(1.0e-10)
PI = _f(3.14159265358979)
GX = _f(1.0e-5)
PI_OKADA = _f(3.14159)   # the literal in STRIKE_SLIP / DIP_SLIP
FLT_MAX = float(np.finfo(np.float32).max)

# WGS84 constants of STEREO_PROJECTION (deform.f90:540-548)
A_AX = _f(6378137.0000)
E_ECC = _f(0.08181919092891)
F2_ECC = _f(0.00669438000426)
K0 = _f(0.9996)


def stereo_projection(lon, lat, lon0, lat0):
    """(x, y) in metres east/north of (lon0, lat0) (deform.f90:494-588)."""
    lon = np.asarray(lon, np.float64)
    lat = np.asarray(lat, np.float64)
    pole = PI / 2.0 - EPS
    la = np.clip(lat * RAD_DEG, -pole, pole)
    lo = lon * RAD_DEG
    lt0 = min(max(float(lat0) * RAD_DEG, -pole), pole)
    ln0 = float(lon0) * RAD_DEG
    sn = np.sin(la)
    cs0, sn0 = np.cos(lt0), np.sin(lt0)
    tmp0 = np.sqrt(1.0 - F2_ECC * sn0 ** 2)
    rho0 = A_AX * (1.0 - F2_ECC) / tmp0 ** 3
    nu0 = float(int(A_AX / tmp0))   # INTEGER by implicit typing
    r = np.sqrt(rho0 * nu0)
    n = np.sqrt(1.0 + F2_ECC * cs0 ** 4 / (1.0 - F2_ECC))
    s1 = (1.0 + sn0) / (1.0 - sn0)
    s2 = (1.0 - E_ECC * sn0) / (1.0 + E_ECC * sn0)
    w1 = (s1 * s2 ** E_ECC) ** n
    sn_xi0 = (w1 - 1.0) / (w1 + 1.0)
    c = (n + sn0) * (1.0 - sn_xi0) / (n - sn0) / (1.0 + sn_xi0)
    w2 = c * w1
    sa = (1.0 + sn) / (1.0 - sn)
    sb = (1.0 - E_ECC * sn) / (1.0 + E_ECC * sn)
    w = c * (sa * sb ** E_ECC) ** n
    xi0 = np.arcsin((w2 - 1.0) / (w2 + 1.0))
    lm0 = ln0
    lm = n * (lo - lm0) + lm0
    xi = np.arcsin((w - 1.0) / (w + 1.0))
    beta = 1.0 + np.sin(xi) * np.sin(xi0) + np.cos(xi) * np.cos(xi0) * np.cos(lm - lm0)
    y = 2.0 * r * K0 * (np.sin(xi) * np.cos(xi0) - np.cos(xi) * np.sin(xi0) * np.cos(lm - lm0)) / beta
    x = 2.0 * r * K0 * np.cos(xi) * np.sin(lm - lm0) / beta
    return x, y


def strike_slip(x2, y1, y2, dp, dd):
    """STRIKE_SLIP (deform.f90:194-217); x2, y1, y2 arrays."""
    sn, cs = np.sin(dp), np.cos(dp)
    q = x2 * sn - dd * cs
    d_bar = y2 * sn - q * cs
    r = np.sqrt(y1 * y1 + y2 * y2 + q * q)
    t1 = np.where(r + d_bar <= EPS, EPS, r + d_bar)
    t2 = np.where(r + y2 <= EPS, EPS, r + y2)
    a4 = 0.5 * 1 / cs * (np.log(t1) - sn * np.log(t2))
    return -(d_bar * q / r / (r + y2) + q * sn / (r + y2) + a4 * sn) / 2 / PI_OKADA


def dip_slip(x2, y1, y2, dp, dd):
    """DIP_SLIP (deform.f90:221-238)."""
    sn, cs = np.sin(dp), np.cos(dp)
    q = x2 * sn - dd * cs
    d_bar = y2 * sn - q * cs
    r = np.sqrt(y1 * y1 + y2 * y2 + q * q)
    xx = np.sqrt(y1 * y1 + q * q)
    a5 = 0.5 * 2 / cs * np.arctan((y2 * (xx + q * cs) + xx * (r + xx) * sn) / y1 / (r + xx) / cs)
    return -(d_bar * q / r / (r + y1) + sn * np.arctan(y1 * y2 / q / r) - a5 * sn * cs) / 2 / PI_OKADA


class Fault:
    """One FAULT record (type_module.f90:130-147): depth HH of the top edge,
    length L, width W, slip D (m), strike TH, dip DL, rake RD (degrees),
    grid origin (XO, YO) and epicentre (X0, Y0) in degrees; switch 0 puts
    the epicentre at the plane's centre, 9 at the centre of its top edge."""

    def __init__(self, hh, l, w, d, th, dl, rd, yo=0.0, xo=0.0, y0=0.0, x0=0.0, switch=0):
        self.hh, self.l, self.w, self.d = hh, l, w, d
        self.th, self.dl, self.rd = th, dl, rd
        self.yo, self.xo, self.y0, self.x0 = yo, xo, y0, x0
        self.switch = switch

    def params(self):
        """The 11 REALs in FAULT order HH, L, W, D, TH, DL, RD, YO, XO, Y0, X0."""
        return np.array([self.hh, self.l, self.w, self.d, self.th, self.dl, self.rd, self.yo,
                         self.xo, self.y0, self.x0], np.float32)


def deform_okada(fault, nx, ny, laycord, x=None, y=None, dx=0.0, dy=0.0):
    """Vertical seafloor displacement (ny, nx) of one fault (deform.f90:74-191),
    FP64, returned rounded to REAL.  Spherical layers (laycord 0) take the
    cell longitudes x (nx) and latitudes y (ny); Cartesian ones the spacing."""
    p = [float(v) for v in fault.params()]
    hh0, l, w, d, th, dl, rd, yo, xo, y0, x0 = p
    ang_l, ang_r, ang_t = RAD_DEG * dl, RAD_DEG * rd, RAD_DEG * th
    halfl = 0.5 * l
    if fault.switch == 9:
        hh = hh0 + w * np.sin(ang_l)
        del_x = w * np.cos(ang_l) * np.cos(ang_t)
        del_y = w * np.cos(ang_l) * np.sin(ang_t)
    else:
        hh = hh0 + 0.5 * w * np.sin(ang_l)
        del_x = 0.5 * w * np.cos(ang_l) * np.cos(ang_t)
        del_y = 0.5 * w * np.cos(ang_l) * np.sin(ang_t)
    ds = d * np.cos(ang_r)
    dd = d * np.sin(ang_r)
    sn, cs = np.sin(ang_l), np.cos(ang_l)
    if laycord == 0:
        lon = np.asarray(x, np.float32).astype(np.float64)[None, :]
        lat = np.asarray(y, np.float32).astype(np.float64)[:, None]
        xs, ys = stereo_projection(lon, lat, x0, y0)
        xs, ys = xs - del_x, ys + del_y
    else:
        jj, ii = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64),
                             indexing="ij")
        yy = float(np.float32(dy)) * jj
        xx = float(np.float32(dx)) * ii
        xl, yl = stereo_projection(xo, yo, x0, y0)
        xs = xx - float(xl) - del_x
        ys = yy - float(yl) + del_y
    xs, ys = np.broadcast_arrays(xs, ys)
    with np.errstate(all="ignore"):
        x1 = xs * np.sin(ang_t) + ys * np.cos(ang_t)
        x2 = -(xs * np.cos(ang_t) - ys * np.sin(ang_t))
        pp = x2 * cs + hh * sn
        f1 = strike_slip(x2, x1 + halfl, pp, ang_l, hh)
        f2 = strike_slip(x2, x1 + halfl, pp - w, ang_l, hh)
        f3 = strike_slip(x2, x1 - halfl, pp, ang_l, hh)
        f4 = strike_slip(x2, x1 - halfl, pp - w, ang_l, hh)
        g1 = dip_slip(x2, x1 + halfl, pp, ang_l, hh)
        g2 = dip_slip(x2, x1 + halfl, pp - w, ang_l, hh)
        g3 = dip_slip(x2, x1 - halfl, pp, ang_l, hh)
        g4 = dip_slip(x2, x1 - halfl, pp - w, ang_l, hh)
        us = (f1 - f2 - f3 + f4) * ds
        ud = (g1 - g2 - g3 + g4) * dd
    out = []
    for u in (us, ud):
        u = np.where(np.abs(u) >= FLT_MAX, 0.0, u)
        u = np.where(np.isnan(u), 0.0, u)
        u = np.where(np.abs(u) <= GX, 0.0, u)
        out.append(u)
    return (out[0] + out[1]).astype(np.float32)


def cell_interp(parent, child_shape, corners, rel_size):
    """CELL_INTERP (all_grids.f90:859-934) in REAL: the child's deformation
    (cny, cnx) from the parent's (pny, pnx); the child's ghost row/column
    and anything the parent cells do not cover stay 0."""
    P = np.asarray(parent, np.float32)
    cny, cnx = child_shape
    out = np.zeros((cny, cnx), np.float32)
    IS, IE, JS, JE = corners
    ir = rel_size
    nxf = np.float32(2 * ir)
    four = np.float32(4.0)
    one = np.float32(1.0)
    d = lambda i, j: P[j - 1, i - 1]  # Fortran (I, J), 1-based
    for i in range(IS, IE + 1):
        for j in range(JS, JE + 1):
            tl = (((d(i, j) + d(i, j + 1)) + d(i - 1, j)) + d(i - 1, j + 1)) / four
            tr = (((d(i, j) + d(i, j + 1)) + d(i + 1, j)) + d(i + 1, j + 1)) / four
            bl = (((d(i, j) + d(i, j - 1)) + d(i - 1, j)) + d(i - 1, j - 1)) / four
            br = (((d(i, j) + d(i, j - 1)) + d(i + 1, j)) + d(i + 1, j - 1)) / four
            for k in range(1, ir + 1):
                for m in range(1, ir + 1):
                    km = np.float32(2 * m) - one     # CELL(1, 2M), CELL(NX+1, 2M)
                    cl = km * (tl - bl) / nxf + bl
                    cr = km * (tr - br) / nxf + br
                    ki = np.float32(2 * k) - one     # CELL(2K, 2M)
                    v = ki * (cr - cl) / nxf + cl
                    ci = (i - IS) * ir + 1 + k
                    cj = (j - JS) * ir + 1 + m
                    if ci <= cnx and cj <= cny:
                        out[cj - 1, ci - 1] = v
    return out


def cell_interp_sc(parent, pos, cxy, cnx, cny):
    """CELL_INTERP_SC (all_grids.f90:938-992) in REAL: each child cell takes
    the CXY-weighted sum of the four parent cells at POS (0 where POS lies
    outside the parent)."""
    P = np.asarray(parent, np.float32)
    pny, pnx = P.shape
    pos = np.asarray(pos, np.int32).reshape(2, cny, cnx)
    cxy = np.asarray(cxy, np.float32).reshape(4, cny, cnx)
    out = np.zeros((cny, cnx), np.float32)
    for j in range(cny):
        for i in range(cnx):
            ki, kj = int(pos[0, j, i]), int(pos[1, j, i])
            if 1 <= ki < pnx and 1 <= kj < pny:
                z1 = P[kj - 1, ki - 1] * cxy[0, j, i]
                z2 = P[kj - 1, ki] * cxy[1, j, i]
                z3 = P[kj, ki - 1] * cxy[2, j, i]
                z4 = P[kj, ki] * cxy[3, j, i]
                out[j, i] = ((z1 + z2) + z3) + z4
    return out
