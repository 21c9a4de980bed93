"""Load the golden fixtures of tests/golden/ back into scenarios."""
import glob
import os

import numpy as np

from gcomcot.scenarios import LayerSpec, Scenario

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix=""):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def scenario(f):
    layers = []
    k = 1
    while f"L{k}_int" in f:
        iv, fv = f[f"L{k}_int"], f[f"L{k}_flt"]
        layers.append(LayerSpec(id=int(iv[0]), parent=int(iv[1]), level=int(iv[2]),
                                laycord=int(iv[3]), laygov=int(iv[4]), nx=int(iv[5]),
                                ny=int(iv[6]), fric_switch=int(iv[7]), rel_size=int(iv[8]),
                                rel_time=int(iv[9]), corners=tuple(int(v) for v in iv[10:14]),
                                dx=float(fv[0]), dy=float(fv[1]), dt=float(fv[2]),
                                fric_coef=float(fv[3]), h=f[f"L{k}_h"], z0=f[f"L{k}_z0"],
                                y=f.get(f"L{k}_y"), dely=f.get(f"L{k}_dely")))
        k += 1
    return Scenario("golden", layers, int(f["steps"][0]) if "steps" in f else 0)
