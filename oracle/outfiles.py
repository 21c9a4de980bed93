"""ORACLE (test infrastructure only): the reference's output files restated.

Only tests/ use this module, as the checker of the library's file writer;
the product never imports it.  It restates, from the reference's
output.f90 and comcot.f90:

- DATA_PRT (output.f90:758-851): which files a layer writes at step K,
  their names, the dry-land mask and tide correction of the elevation;
- MAX_AMP (output.f90:855-940): when the maximum file of the outermost
  layer is written and under which name, and its mask;
- the main loop's model time (comcot.f90:304, 323), a REAL accumulated
  one DT per root step;
- the record layout of a Fortran unformatted sequential file holding one
  whole (NX, NY) REAL array: a 4-byte length, the column-major data, the
  same 4-byte length.

Pinned in tests/test_outputs_oracle.py against the flang-built reference's
own DATA_PRT and MAX_AMP writing files, byte for byte, on stepped nested
states with dry land, with and without a tide correction, and over a sweep
of model times covering several half-hour marks and the end of the run.
"""
import math
import struct

import numpy as np

GX = np.float32(1.0e-5)  # COMMON /CONS/ GX (comcot.f90 BLOCK DATA)


def fort_int(v, w):
    """Fortran Iw.w: zero-padded to w digits, w asterisks when it does not fit."""
    v = int(v)
    s = "%0*d" % (w, v)
    if v < 0 or len(s) > w:
        return "*" * w
    return s


def shown_elevation(z, h, tide=0.0):
    """Elevation as written (output.f90:781-793 for DATA_PRT, :918-931 for
    MAX_AMP): 0 on dry land (Z + H <= GX where H < 0), else Z plus the tide
    level when |TIDE_LEVEL| > GX.  All in REAL."""
    z = np.asarray(z, np.float32)
    h = np.asarray(h, np.float32)
    tide = np.float32(tide)
    dry = ((z + h) <= GX) & (h < np.float32(0.0))
    wet = z + tide if abs(tide) > GX else z
    return np.where(dry, np.float32(0.0), wet).astype(np.float32)


def fortran_record(plane):
    """One unformatted sequential record of a REAL (NX, NY) array given as
    a (ny, nx) row-major array (the same bytes as column-major (NX, NY))."""
    body = np.ascontiguousarray(plane, dtype="<f4").tobytes()
    mark = struct.pack("<i", len(body))
    return mark + body + mark


def snapshot_names(lid, k, fluxswitch):
    """File names DATA_PRT writes for layer lid at step k (output.f90:812-845)."""
    tail = "%s_%s.dat" % (fort_int(lid, 2), fort_int(k, 6))
    names = []
    if fluxswitch <= 1 or fluxswitch == 9:
        names.append("z_" + tail)
    if fluxswitch == 0:
        names += ["m_" + tail, "n_" + tail]
    return names


def snapshot_files(lid, k, fluxswitch, z, m, n, h, tide=0.0):
    """{name: bytes} of DATA_PRT for one layer at step k, from its current
    elevation z and fluxes m, n, (ny, nx) arrays."""
    out = {}
    for name in snapshot_names(lid, k, fluxswitch):
        if name.startswith("z_"):
            out[name] = fortran_record(shown_elevation(z, h, tide))
        elif name.startswith("m_"):
            out[name] = fortran_record(m)
        else:
            out[name] = fortran_record(n)
    return out


def model_times(dt, kmax):
    """TIME after root steps 0..kmax: REAL, starting at 0 (ISTART = 1) and
    raised by DT each step (comcot.f90:304, 323)."""
    dt = np.float32(dt)
    t = np.float32(0.0)
    out = [t]
    for _ in range(kmax):
        t = np.float32(t + dt)
        out.append(t)
    return out


def _nint(x):
    """Fortran NINT of a REAL: nearest integer, halves away from zero."""
    x = float(x)
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def maxamp_name(time, dt, tend):
    """The file MAX_AMP writes at model time `time` (output.f90:912-934),
    or None.  time, dt, tend are REAL."""
    time, dt, tend = np.float32(time), np.float32(dt), np.float32(tend)
    k1 = math.floor(np.float32(np.float32(time - dt) / np.float32(1800.0)))
    if k1 <= 0:
        k1 = 0
    k2 = math.floor(np.float32(time / np.float32(1800.0)))
    last = time >= np.float32(tend - np.float32(np.float32(2.0) * dt))
    if not (k2 > k1 or last):
        return None
    if last:
        return "zmax_layer01.dat"
    k = _nint(np.float32(time / np.float32(60.0)))
    return "zmax_layer%s_%smins.dat" % (fort_int(1, 2), fort_int(k, 4))


def maxamp_file(zraw, h, tide=0.0):
    """Body of a MAX_AMP file from the recorder maximum of layer 1."""
    return fortran_record(shown_elevation(zraw, h, tide))


def schedule(dt, tend, iprt, out_opt, k_from, k_to):
    """[(k, snapshot?, zmax name or None)] for root steps k_from..k_to of the
    main loop (comcot.f90:308-310 at k = 0, 415-423 after each step)."""
    times = model_times(dt, k_to)
    out = []
    for k in range(k_from, k_to + 1):
        snap = iprt > 0 and (k == 0 or k % iprt == 0)
        zname = None
        if k >= 1 and out_opt in (0, 2):
            zname = maxamp_name(times[k], dt, tend)
        if snap or zname:
            out.append((k, snap, zname))
    return out
