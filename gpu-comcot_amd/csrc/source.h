// Seafloor deformation of a fault plane applied to a layer tree on the
// device (GET_FLOOR_DEFORM's fault-model branch, deform.f90:2-71).
#pragma once

#include <hip/hip_runtime.h>

#include "tree.h"

namespace gc {

// params: the FAULT record's HH, L, W, D, TH, DL, RD, YO, XO, Y0, X0
// (type_module.f90:130-147); fswitch 0 = epicentre at the plane's centre,
// 9 = at the centre of its top edge.  Spherical outermost layers need their
// coordinates (Tree::xs/ys).  Every layer's level-1 elevation rises and its
// depth falls by its deformation, kept in LayerRec::deform.
int tree_apply_fault(Tree &T, hipStream_t st, const float *params, int fswitch);

}  // namespace gc
