"""Fault sources (the reference's FAULT record, type_module.f90:130-147).

A Fault is handed to DeviceRun.fault / Batch.fault, which compute its
Okada (1985) seafloor deformation on the device and apply it to every
layer (deform.f90:2-71).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Fault:
    """Depth of the top edge hh (m), length l, width w (m), slip d (m),
    strike th, dip dl, rake rd (degrees), grid origin (xo, yo) for Cartesian
    layers and epicentre (x0, y0), in degrees; switch 0 puts the epicentre at
    the plane's centre, 9 at the centre of its top edge."""
    hh: float
    l: float
    w: float
    d: float
    th: float
    dl: float
    rd: float
    yo: float = 0.0
    xo: float = 0.0
    y0: float = 0.0
    x0: float = 0.0
    switch: int = 0

    def params(self):
        return np.array([self.hh, self.l, self.w, self.d, self.th, self.dl, self.rd, self.yo,
                         self.xo, self.y0, self.x0], np.float32)
