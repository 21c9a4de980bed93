/* C restatement of COMCOT's staggered-leapfrog shallow-water step — the ORACLE.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker (or the
 * timed CPU baseline); the product path never calls it.
 *
 * Every function of oracle.cpp restates one Fortran routine of the
 * reference and cites it.  The state is kept exactly like the Fortran CPU
 * path: two time levels per plane, level-2 -> level-1 copies at the end of
 * a step (comcot.f90:487-489, all_grids.f90:323-328), column-major planes
 * (I fastest, index (J-1)*NX + (I-1)).
 *
 * Precision: 32 restates the FP32 (REAL) Fortran; 64 is the same algorithm
 * in double, the FP64 yardstick the FP32 runs are measured against.
 * pow_mode: 0 = libm powf/expf (what the flang-built reference calls);
 *           1 = the deterministic pow/exp of the HIP kernels (detmath.h),
 *               so the GPU path can be compared bit for bit.
 *
 * Arrays crossing this API are double (exact for FP32 values).
 */
#ifndef COMCOT_ORACLE_H
#define COMCOT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

void *orc_new(int precision, int pow_mode);
void orc_free(void *s);

/* Add a layer.  id 1 = root (parent -1).  Spherical layers take dx in arc
 * minutes, y[] the latitude of each row (degrees) and dely[] the row
 * spacing (arc minutes); Cartesian layers take dx, dy in metres.  Setup
 * mirrors ALLOC + PQ_DEPTH + PARA.  corners: IS,IE,JS,JE (1-based) in the
 * parent.  Returns 0 on success. */
int orc_add_layer(void *s, int id, int parent, int laycord, int laygov,
                  int nx, int ny, double dx, double dy, double dt,
                  const double *y, const double *dely,
                  const double *h, const double *z0,
                  int fric_switch, double fric_coef,
                  int rel_size, int rel_time, const int *corners);

/* Root steps: MASS, OPEN, ALL_GRID, MOMENT, max/arrival, level copy. */
void orc_step(void *s, int nsteps);

/* One routine: 1 MASS, 2 OPEN, 3 MOMENT, 4 JNQ, 5 NEWQ(t), 6 JNZ,
 * 7 UPDATE, 8 ALL_GRID, 9 root level copy, 10 max/arrival of layer id */
void orc_call(void *s, int op, int id, int t);

/* State planes, each 2*nx*ny (level 1 then level 2).  which 0 get, 1 set. */
void orc_state_io(void *s, int id, int which, double *z, double *m,
                  double *n, double *dz);

/* HP, HQ planes; R0,R1,R11,R2,R21,R22,R3,R4,R5 planes (9*nx*ny); R6 (ny);
 * RX, RY, GRX, GRY, DT */
void orc_setup_io(void *s, int id, double *hp, double *hq, double *rpl,
                  double *r6, double *sc);

/* Interface flux cache XFLUX(nx,2), YFLUX(ny,2) */
void orc_flux_io(void *s, int id, double *xflux, double *yflux);

/* Max/arrival fields of a layer: wet-only max (output.f90:876-885 form,
 * starts at 0), raw max (GPUOutput.cu:46-53 form, starts at the initial
 * Z), first-arrival step (-1 = never). */
void orc_max_io(void *s, int id, double *zmax_wet, double *zmax_raw,
                int *arrival);
void orc_set_arrival_threshold(void *s, double thr);

/* Time-step check (initialization.f90:1136-1307) for a root of nx*ny with
 * the given geometry: returns the adjusted dt. */
double orc_cr_dt(int laycord, int laygov, int nx, int ny, const double *h,
                 double dx, double dy, double lat_first, double lat_last,
                 double dt_in);

#ifdef __cplusplus
}
#endif
#endif
