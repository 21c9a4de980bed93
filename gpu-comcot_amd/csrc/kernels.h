// Device-side layer view and kernel launchers of the COMCOT step.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace gc {

// What one step kernel needs of a layer.  Planes are column-major NX*NY
// float (index (J-1)*NX + I-1), time level 1 (in) and 2 (out) bound per
// launch from the layer's current parity.
struct DevLayer {
  int nx, ny;
  int is_root;     // outermost layer: loop ranges of ID==1 and the OPEN ring
  int modscm;      // 0: modified-scheme (SCM) term on
  int linchk;      // 1: convection terms on
  int fric;        // 1: Manning friction on (FRIC_SWITCH != 1)
  float rx, ry, grx, gry, dt;
  float kfric;     // (0.5*DT*GRAV)*FM*FM, the Fortran's CONST*FM*FM
  const float *h;
  const float *coef;   // per-row table (comcot_consts.h), spherical only
  const float *r24;    // optional R2|R4 planes (drop-in fallback), else null
  const float *z1, *m1, *n1;
  float *z2, *m2, *n2;
  float *dzp;          // DZ(:,:,2) plane, kept for nonlinear parents, else null
  // wet-only max and first arrival, fused into the step when zmax != null
  float *zmax;
  int *arrival;
  int *ctl;            // root-step counters: read ctl[ctl_in], write ctl[1-ctl_in]
  int ctl_in, advance; // advance: this launch writes the next counter
  float arr_thr;
};

enum StepMode { kFused = 0, kMassOnly = 1, kMomentOnly = 2 };

// One pass over all tiles of a layer (tile shape chosen by variant).
hipError_t launch_step(const DevLayer &L, int mode, bool nl, bool sph, int variant,
                       hipStream_t st);
int num_tile_variants();
void tile_variant_shape(int v, int *tx, int *ty);

// Reference-semantics maximum (no wet test) plus the wet max/arrival.
hipError_t launch_max(const DevLayer &L, float *zmax_raw, hipStream_t st);

// Nesting: child boundary fluxes for sub-step t (1 = JNQ, >1 = NEWQ)
struct DevNest {
  // parent
  int pnx, pny, p_sph, p_nl, p_r24;
  const float *ph, *pz2, *pm1, *pn1, *pcoef, *pr24;
  float pgrx, pgry;
  float *xflux, *yflux;  // parent cache (NX,2), (NY,2)
  // child
  int cnx, cny, rel_size, rel_time;
  int is, ie, js, je;    // footprint in the parent, 1-based
  const float *ch, *cz1, *cz2;
  float *cm1, *cn1;      // child level-1 fluxes, boundary written
};
hipError_t launch_jnq(const DevNest &N, int t, hipStream_t st);
hipError_t launch_jnz(const DevNest &N, float *pz2_out, hipStream_t st);

}  // namespace gc
