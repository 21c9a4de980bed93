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
  // MAX_AMP's maximum (GPUOutput.cu:46-53: no wet test, NaN replaces),
  // fused into the step of the outermost layer when != null
  float *zraw;
  // SPONGE_LAYER's coefficients COEFX per column I and COEFY per row J
  // (boundaries.f90:206-246), applied after the momentum pass when != null
  const float *sponge_cx, *sponge_cy;
  // 1: the single-layer step kernels (launch_step, launch_step2) write the
  // new level-2 planes Z2, M2, N2 of interior tiles with nontemporal
  // stores.  For a layer whose one step moves more bytes than the chip's
  // last-level cache holds, the next step finds none of them in a cache
  // anyway, and streaming them keeps the depth and maximum planes there;
  // for a layer small enough to stay in L2, streaming spares each kernel
  // end the write-back of the whole new level.  The batch kernels always
  // store through the caches.
  int stream_out;
};

// kMomentMass: the momentum half of root step k of a linear outermost
// layer with nested layers, with the continuity (+ open boundary) of step
// k+1 in the same launch.  Nothing runs between the two in the root step
// order: momentum(k) reads Z(k) (after JNZ) and the level-1 fluxes and
// writes M(k), N(k) to level 2; continuity(k+1) reads Z(k), M(k), N(k)
// and writes Z(k+1) into the buffer that held Z(k-1), which linear
// momentum does not read.  The view's z1 is Z(k), its z2 that buffer.
enum StepMode { kFused = 0, kMassOnly = 1, kMomentOnly = 2, kMomentMass = 3 };

// One pass over all tiles of a layer (tile shape chosen by variant).
hipError_t launch_step(const DevLayer &L, int mode, bool nl, bool sph, int variant,
                       hipStream_t st);
int num_tile_variants();
void tile_variant_shape(int v, int *tx, int *ty);

// Two root steps of a linear outermost layer without children in one
// launch (temporal blocking); shape 0..6 = tiles of 64x8, 32x16, 64x16 with
// 256 threads, 64x16 with 512 and 1024, 32x16 with 512, 64x32 with 1024.
// Reads level 1, writes level 2 (two steps later); with zmax set, the
// wet maximum and arrival take both steps and the counter advances by 2.
hipError_t launch_step2(const DevLayer &L, bool sph, int shape, hipStream_t st);
constexpr int kStep2Shapes = 7;
// tile of two-step shape `shape`
void step2_shape(int shape, int *tx, int *ty);

// Reference-semantics maximum (no wet test) plus the wet max/arrival.
hipError_t launch_max(const DevLayer &L, float *zmax_raw, hipStream_t st);

// Output planes of one layer for the reference's file writers, in one pass:
// oz = elevation with DATA_PRT's dry-land mask and tide (output.f90:781-793),
// om/on = copies of the fluxes, ozmax = MAX_AMP's masked maximum
// (output.f90:918-931).  Any output may be null.
struct OutPack {
  const float *z, *h, *m, *n, *zraw;
  float *oz, *om, *on, *ozmax;
  float tide;
  long long count;
};
hipError_t launch_out_pack(const OutPack &p, hipStream_t st);

// One fault plane of the Okada (1985) source as DEFORM_OKADA evaluates it
// (deform.f90:74-191), in FP64: constants of the fault and, for spherical
// layers, the stereographic projection split into per-row and per-column
// factors (deform.f90:494-588 depends on the latitude and the longitude
// separately until the final combination).
struct FaultDev {
  int nx, ny, sph;
  double sin_t, cos_t, halfl, w, hh, del_x, del_y, sn, cs, ds, dd;
  double two_rk0, sxi0, cxi0;  // spherical: 2 R K0, sin/cos of XI0
  const double *row;           // spherical: {sin XI, cos XI} per row J
  const double *col;           // spherical: {sin, cos}(LM - LM0) per column I
  double dx, dy, xl, yl;       // Cartesian: spacing, projected grid origin
};
// deform = vertical seafloor displacement of every cell of the layer
hipError_t launch_okada(const FaultDev &f, float *deform, hipStream_t st);
// CELL_INTERP (all_grids.f90:859-934): a nested layer's deformation from its
// parent's, in REAL; corners are the child's IS, IE, JS, JE in the parent
hipError_t launch_cell_interp(const float *parent, int pnx, float *child, int cnx, int cny,
                              const int *corners, int rel_size, hipStream_t st);
// Z1 += D and H -= D (deform.f90:55-64, UPDATE_BATH initialization.f90:
// 2410-2460); zraw != null also restarts MAX_AMP's maximum at the new Z1
hipError_t launch_apply_deform(float *z1, float *h, const float *deform, float *zraw, size_t n,
                               hipStream_t st);

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
  float *pz2_out;        // parent level-2 Z, footprint written by JNZ
};
hipError_t launch_jnq(const DevNest &N, int t, hipStream_t st);
hipError_t launch_jnz(const DevNest &N, hipStream_t st);

// SC_OPTION = 1 coupling of a Cartesian child to its spherical parent
// (all_grids.f90:996-1425): boundary fluxes bilinear in the parent's fluxes
// at the positions POS/CXY that SUBGRID_MATCHING's reverse UTM projection
// gives every child cell, and the parent's elevation averaged from the wet
// child cells that fall in each parent cell.
struct DevSc {
  // parent (spherical)
  int pnx, pny, p_nl, p_modscm, p_r24;
  const float *ph, *pz2, *pm1, *pn1, *pcoef, *pr24;
  float *pm2, *pn2;    // parent level-2 fluxes: NEWQ_SC's band forecasts
  float *pz2_out;      // parent level-2 elevation, averaged by JNZ_SC
  // child (Cartesian)
  int cnx, cny, rel_time;
  float half;          // REAL(REL_SIZE**2)/2
  const int *pos;      // POS(NX,NY,2), component-major
  const float *cxy;    // CXY(NX,NY,4), component-major
  const float *ch, *cz_now, *cz2;  // H, newest Z (Fortran level 2 after UPDATE), level-2 buffer
  float *cm1, *cn1;    // child level-1 fluxes, boundary written
  const int *steps_done;  // root steps completed: the child's first sub-step ever sees Z2 = 0
  int band[4][4];      // M bands (left, right edge), N bands (top, bottom): IS, IE, JS, JE
  int corners[4];      // the POS extent
  const int *jnz_off, *jnz_idx;  // per parent cell of the extent: its child cells in Fortran order
};
hipError_t launch_sc_edges(const DevSc &S, int t, hipStream_t st);
hipError_t launch_sc_jnz(const DevSc &S, hipStream_t st);
// CELL_INTERP_SC (all_grids.f90:938-992): a child's deformation from its parent's
hipError_t launch_sc_cell_interp(const float *parent, int pnx, int pny, float *child, int cnx,
                                 int cny, const int *pos, const float *cxy, hipStream_t st);

// ---- batch engine: one persistent launch over the tiles of many layers
struct BatchDesc {
  DevLayer L;
  int tiles_x, ntiles;
};
struct alignas(8) BatchItem {
  int slot, tile;  // descriptor index, tile index within that layer
};
struct BatchLaunch {
  const BatchDesc *descs;
  const BatchItem *items;
  const int *qstart;  // 9 entries: items of XCD queue q are [qstart[q], qstart[q+1])
  int *heads;         // 8 queue heads, zero before the launch
  int nitems;
};
// Fused launches take tile shape `shape` (batch_shape); continuity and
// momentum halves of nested parents use 64 x 8.  prepare: only look up
// the kernel's resident-workgroup count (do it before a graph capture).
int batch_num_shapes();
void batch_shape(int k, int *tx, int *ty);
hipError_t launch_batch(const BatchLaunch &B, int mode, bool nl, bool sph, int shape,
                        hipStream_t st, bool prepare = false);
// Two root steps of single-layer linear layers, one workgroup per item;
// items tile each layer with step2_shape(shape) tiles.
hipError_t launch_batch2(const BatchLaunch &B, bool sph, int shape, hipStream_t st);
// JNQ/NEWQ and JNZ of several nests in one launch: nest n owns blocks
// [start[n], start[n+1]) of 128 threads; *_block_count gives each one's share.
int jnq_block_count(const DevNest &N);
int jnz_block_count(const DevNest &N);
hipError_t launch_jnq_multi(const DevNest *nests, const int *start, int nnest, int nblocks,
                            int t, hipStream_t st);
hipError_t launch_jnz_multi(const DevNest *nests, const int *start, int nnest, int nblocks,
                            hipStream_t st);
hipError_t launch_fill_i32(int *p, int n, int v, hipStream_t st);
// ---- gauges: bilinear samples of the elevation after each root step
struct GaugeDev {
  int lid, ki, kj, nx;   // layer, 1-based lower-left stencil cell, layer width
  float cx, cy, tide;    // GET_TS weights and tide-level offset
  int pad;
};
struct ZNow {
  const float *z[20];    // current elevation plane of each layer id (index = id)
};
// Sample n gauges into ring[g * cap + (step - 1) % cap], step = *step_now
// (the root steps completed), with GET_TS's arithmetic (output.f90:1085-1094).
hipError_t launch_gauges(const GaugeDev *g, int n, const ZNow &z, const int *step_now,
                         float *ring, int cap, hipStream_t st);

// Streaming copy of bytes (a multiple of 16) from src to dst.
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
// Self-check of the short divisions and square root (fastdiv.h) against
// the IEEE operations: out[6g..6g+5] = lane g's mismatches of div12 and
// sqrt_inrange (over its slice of all floats) and of div_inrange (over
// per_lane random in-range pairs), then its first mismatching x, pair
// (a << 32 | b) and square-root argument; 256*blocks lanes.
// powf_manning(x[q], 2.3333333f) -> pw[q] and expf_det(x[q]) -> ex[q], q < n.
hipError_t launch_detmath_eval(const float *x, long long n, float *pw, float *ex,
                               hipStream_t st);
hipError_t launch_division_check(unsigned long long seed, int per_lane, unsigned long long *out,
                                 int blocks, hipStream_t st);

}  // namespace gc
