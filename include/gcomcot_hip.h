/* gcomcot_hip — MI355X (gfx950) HIP implementation of COMCOT's staggered
 * leapfrog shallow-water step behind the C ABI that the COMCOT Fortran
 * driver already calls.
 *
 * Part 1 is the drop-in boundary: the 19 external symbols that the
 * reference Fortran objects leave undefined for the GPU half
 * (comcot.f90:281-298,362,367,390,399,421,509; output.f90:777-778,918,
 * 1085,1140; all_grids.f90:429-446), with the reference's signatures:
 * lowercase with a trailing underscore, every argument by reference,
 * arrays as base addresses of contiguous column-major (I fastest) data,
 * layer id 1 = outermost layer, ids 1..19.  Every call returns after the
 * device work it issued is complete, as the reference's calls do.
 * Errors print a message and exit(EXIT_FAILURE), as the reference does
 * (GPUHeader.h:29-44), unless gcomcot_set_error_mode_(0) selected status
 * mode (then calls return early and gcomcot_last_error_ reports).
 *
 * Part 2 extends the boundary with what the reference cannot express:
 * nonlinear and Cartesian layers, a whole-run device loop, the wet-only
 * maximum and first-arrival fields, and the scenario batch engine.
 */
#ifndef GCOMCOT_HIP_H
#define GCOMCOT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Part 1: reference entry points                                      */
/* ------------------------------------------------------------------ */

/* replaces gcomcot_init_gpu_ (GPUcomcot.cu:192-211): select device 0,
 * create the library's stream, clear the layer table */
void gcomcot_init_gpu_(void);

/* replaces gcomcot_init_layer_ (GPUcomcot.cu:214-251 + cuda_alloc_layer
 * :36-142).  Spherical linear layer (the only kind the reference's GPU
 * path runs, GPUall_grids.cu:683,687).  R1..R5, R11 are the (NX,NY)
 * planes of PARA, R6 its (NY) vector; corners = IS,IE,JS,JE in the parent
 * (1-based).  Row factors are taken from column I=1 as the reference does
 * (GPUcomcot.cu:62-68); R2/R4 are reduced to row factors of max(HP,0) when
 * every cell of the row reproduces bit for bit, otherwise kept as planes. */
void gcomcot_init_layer_(int *layerid, int *parent, int *level, int *rel_size, int *rel_time,
                         int *corners, float *grx, float *gry, float *R1, float *R2, float *R3,
                         float *R4, float *R5, float *R6, float *R11, float *H, float *Z,
                         int *nx, int *ny);

/* replaces gcomcot_show_grids_ (GPUall_grids.cu:26-30): print the tree */
void gcomcot_show_grids_(void);

/* replaces mass_launch_ (GPUMass_s.cu:9-37): continuity for layer *lid
 * into level 2.  For the outermost layer the radiation boundary of OPEN
 * (boundaries.f90:22-202) is computed in the same kernel.  Z1/Z2/H are
 * ignored (the reference uses them only in debug builds). */
void mass_launch_(const float *Z1, float *Z2, const float *H, const int *lid);

/* replaces openbd_launch_ (GPUOpen_BD.cu:13-37): nothing left to do — the
 * boundary ring was produced by mass_launch_ (same inputs, same values) */
void openbd_launch_(float *Z2);

/* replaces momt_launch_ (GPUMoment_s.cu:10-56): momentum for layer *lid
 * into level 2.  Arrays ignored, as in the reference. */
void momt_launch_(float *M2, float *N2, float *Z2, const int *lid);

/* replaces all_grid_launch_ (GPUall_grids.cu:701-705): every nested layer
 * below the root for one root step (JNQ/NEWQ, mass, recursion, momentum,
 * JNZ at the half step, level rotation), all_grids.f90:4-312 order */
void all_grid_launch_(void);

/* replaces maxamp_launch_ (GPUOutput.cu:11-27): Zmax = max(Zmax, Z2) with
 * no wet test (reference semantics); also advances the wet-only maximum
 * and the first-arrival field of the extension */
void maxamp_launch_(const int *lid);

/* replaces cuda_update_layer_ (GPUcomcot.cu:144-155): swap time levels */
void cuda_update_layer_(int *lid);

/* replace the accessors of GPUOutput.cu:55-103: level 2 (z, mn), level 1
 * (z1, mn1) and the raw maximum, whole planes, host <-> device */
void cuda_getz_(float *Z, int *lid);
void cuda_setz_(float *Z, int *lid);
void cuda_getz1_(float *Z, int *lid);
void cuda_setz1_(float *Z, int *lid);
void cuda_getmn_(float *M, float *N, int *lid);
void cuda_setmn_(float *M, float *N, int *lid);
void cuda_getmn1_(float *M, float *N, int *lid);
void cuda_setmn1_(float *M, float *N, int *lid);
void cuda_getzmax_(float *Zmax, int *lid);

/* replaces gcomcot_get_z_ (GPUOutput.cu:29-44): level-2 Z at 1-based
 * (i,j),(i+1,j),(i,j+1),(i+1,j+1) for gauge interpolation */
void gcomcot_get_z_(int *lid, float *l1, float *l2, float *l3, float *l4, int *i, int *j);

/* replaces cuda_shutdown_ (GPUcomcot.cu:253-256, empty there): frees every
 * device buffer of the library */
void cuda_shutdown_(void);

/* The reference's differential checks, called by its Fortran only when
 * built with DEBUG_CORE / DEBUG_ALL_GRID (debug_option.h), right after the
 * host computed the same routine from the state SET_ALL copied to the
 * device.  Both run the device's own routine on that state and print every
 * value that differs from the host's by 1e-4 or more, in the reference's
 * format.
 * replaces edgeinterp_dbglaunch_ (GPUall_grids.cu:299-394; called at
 * all_grids.f90:490 and :650): JNQ (*t = 1) or NEWQ sub-step *t of child
 * *idA from parent *idO, then compares the child's level-1 boundary fluxes
 * (left/right columns of mA, bottom/top rows of nA) and, at *t = 2, the
 * parent's flux cache YFLUX(NY,2)/XFLUX(NX,2) over the child's extent.
 * mO/nO are unused, as in the reference. */
void edgeinterp_dbglaunch_(float *mO, float *nO, int *idO, float *mA, float *nA, int *idA, int *t,
                           float *yflux, float *xflux);
/* replaces jnz_dbglaunch_ (GPUall_grids.cu:626-643; called at
 * all_grids.f90:417): JNZ of child *inid into parent *outid, then compares
 * the parent's whole level-2 elevation with Z_f. */
void jnz_dbglaunch_(float *Z_f, int *outid, int *inid);

/* ------------------------------------------------------------------ */
/* Part 2: extensions (no reference counterpart)                       */
/* ------------------------------------------------------------------ */

/* Per-row coefficient table of PARA (comcot.f90:694-804) for a spherical
 * layer: coef[k*ny + j], k = R0,R1,R11,F2,R21,R22,R3,F4,R5,R6 (10 rows of
 * ny).  R2 = F2*max(HP,0) and R4 = F4*max(HQ,0) for linear layers; F2, F4
 * for nonlinear ones. */
void gcomcot_para_rows_(const int *laygov, const int *ny, const float *dt, const float *dx,
                        const float *y, const float *dely, float *coef);

/* CR_CHECK (initialization.f90:1163-1196): dt of a first-level layer */
float gcomcot_cr_dt_(const int *laycord, const int *laygov, const int *nx, const int *ny,
                     const float *h, const float *dx, const float *dy, const float *lat_first,
                     const float *lat_last, const float *dt_in);

/* CR_CHECK child part (initialization.f90:1206-1247): REL_TIME and DT */
void gcomcot_child_dt_(const int *laycord, const int *laygov, const int *nx, const int *ny,
                       const float *h, const float *dx, const float *dy, const float *lat_first,
                       const float *lat_last, const float *dt_parent, int *rel_time,
                       float *dt_child);

/* Any layer: laycord 0 spherical / 1 Cartesian; laygov 0 linear / 1
 * nonlinear (9 and, for Cartesian, 5 accepted as nonlinear, as MASS and
 * MOMENT dispatch them, mass.f90:19-52); fric_switch 0 Manning (constant
 * fric_coef) / 1 none.  coef: per-row table of gcomcot_para_rows_ for
 * spherical layers (ignored for Cartesian).  dx, dy: arc minutes
 * (spherical) or metres (Cartesian).  Returns 0 or an error code. */
int gcomcot_init_layer_ex_(int *layerid, int *parent, int *level, int *laycord, int *laygov,
                           int *rel_size, int *rel_time, int *corners, int *fric_switch,
                           float *fric_coef, float *dt, float *dx, float *dy, float *coef,
                           float *H, float *Z, int *nx, int *ny);

/* Run *nsteps complete root steps on the device (mass+open, nested
 * layers, momentum, wet-only max and arrival, level rotation) as one
 * enqueued sequence (captured in a HIP graph when *use_graph != 0).
 * Returns 0 or an error code. */
int gcomcot_run_steps_(int *nsteps, int *use_graph);

/* Same work as gcomcot_run_steps_ without a graph, timed with HIP events
 * recorded on the library's stream: *ms_total from the first launch to
 * the end of the last step; ms_each[k] (if not NULL) the duration of root
 * step k, from the event before its first kernel to the event after its
 * last one. */
int gcomcot_time_steps_(int *nsteps, float *ms_total, float *ms_each);

/* Wet-only maximum (starts at 0; cells with H+Z > GX, as output.f90:
 * 876-885) and first-arrival root step (-1 = never) of a layer */
void gcomcot_get_max_(int *lid, float *zmax_wet, int *arrival);
void gcomcot_set_arrival_threshold_(float *thr);

/* Interface flux cache of a parent layer: XFLUX(NX,2), YFLUX(NY,2)
 * (all_grids.f90:720,790) */
void gcomcot_get_flux_cache_(int *lid, float *xflux, float *yflux);

/* Layer size query: returns 0 if the id exists */
int gcomcot_layer_dims_(int *lid, int *nx, int *ny);

/* Error handling: mode 1 = print and exit (reference), 0 = record and
 * return; last error text copied into buf (returns its code, 0 = none) */
void gcomcot_set_error_mode_(int *fatal);
int gcomcot_last_error_(char *buf, int *len);

/* Launch-shape override for the step kernels (autotuner hook): tile
 * variant index, see gcomcot_tile_variants_.  Returns 0 if valid. */
int gcomcot_set_tile_variant_(int *variant);
int gcomcot_tile_variants_(int *count, int *tx, int *ty);

/* Gauges on the device (reference: GET_TS, output.f90:1035-1155, which
 * calls gcomcot_get_z_ and copies a 2x2 stencil to the host per gauge per
 * step).  gcomcot_set_layer_coords_ gives a layer's LAYER%X(NX), %Y(NY);
 * gcomcot_gauges_set_ places *n gauges at (x0[k], y0[k]) on layers lid[k]
 * (stencil search and weights as GET_TS, inner layers from index 2) with a
 * ring of *capacity samples each (*n = 0 removes them).  The device loop
 * (gcomcot_run_steps_, gcomcot_time_steps_) then samples every gauge
 * after every root step on the device, GET_TS's arithmetic bit for bit;
 * gcomcot_gauges_read_ copies out the last min(steps, capacity,
 * *max_steps) samples in time order, series[k * *max_steps + s], count in
 * *nout.  With gauges set, linear layers run one step per launch. */
int gcomcot_set_layer_coords_(int *lid, const float *x, const float *y);
int gcomcot_gauges_set_(int *n, const float *x0, const float *y0, const int *lid, int *capacity);
int gcomcot_gauges_read_(float *series, int *max_steps, int *nout);

/* A Cartesian layer nested in the spherical layer 1 with SC_OPTION = 1, the
 * coupling the reference selects for such a layer (initialization.f90:
 * 362-370; all_grids.f90:996-1425: EDGE_INTERP_SC, NEWQ_SC, JNZ_SC,
 * CELL_INTERP_SC).  Arguments as gcomcot_init_layer_ex_ for a Cartesian
 * layer, with corners = LA%CORNERS as SUBGRID_MATCHING leaves them (the
 * extent of POS), plus its POS (NX,NY,2) and CXY (NX,NY,4) arrays
 * (initialization.f90:1062-1124).  EDGE_INTERP_SC's wet tests index H and
 * Z with stale loop variables in the reference (undefined values); here
 * they test the boundary cell itself, as NEWQ_SC does.  Returns 0 or an
 * error code. */
int gcomcot_init_layer_sc_(int *layerid, int *parent, int *level, int *laygov, int *rel_size,
                           int *rel_time, int *corners, int *fric_switch, float *fric_coef,
                           float *dt, float *dx, float *dy, float *H, float *Z, int *nx, int *ny,
                           int *pos, float *cxy);

/* Exact checkpoints of the device state (the reference's HOT_START,
 * hotstart.f90:2-166, saves Z, M, N and the maxima as 10F12.5 text, which
 * rounds them; the GPU build does not support it, comcot.f90:328-332).
 * gcomcot_checkpoint_save_ writes every plane of every layer, the time-level
 * parity and the root-step count to the file path[0..*pathlen);
 * gcomcot_checkpoint_load_ restores it into a tree built with the same
 * layers (same ids, parents, sizes, kinds and corners; refused otherwise
 * before anything changes, also when the file is cut short).  Continuing
 * after a load gives the same bits as continuing without it.  Scheduled
 * ruptures are not part of a checkpoint: a load drops those whose call lies
 * at or before the restored step (their deformation is in the state) and
 * keeps later ones.  The batch forms do the same for one scenario and
 * restore its step count, so a served request can be preempted and resumed. */
int gcomcot_checkpoint_save_(const char *path, int *pathlen);
int gcomcot_checkpoint_load_(const char *path, int *pathlen);
int gcomcot_batch_checkpoint_save_(int *sid, const char *path, int *pathlen);
int gcomcot_batch_checkpoint_load_(int *sid, const char *path, int *pathlen);

/* The reference's hot-start files (HOT_START / HOTSTART_DATA,
 * hotstart.f90:3-166, called at K = START_STEP in comcot.f90:330-333,
 * where the GPU build prints that it is unsupported).  Five text files per
 * layer in directory dir[0..*dirlen) (the working directory when empty):
 * z1_LL_SSSSSS.dat, m1_, n1_, zmax1_, zmin1_ with LL the layer id and
 * SSSSSS the start step, each holding cells I, J >= 1 of the outermost and
 * I, J >= 2 of a nested layer, one '(10F12.5)' record group per row J.
 * gcomcot_hotstart_write_ is START_TYPE 0: it writes level-1 Z, M, N of
 * every layer, the maximum cuda_getzmax_ returns and zeros for Z_MIN (the
 * reference never updates it).  The snapshot is meant for a state after
 * START_STEP - 1 root steps, as the reference's.  gcomcot_hotstart_read_
 * is START_TYPE 1 (the outermost layer) or 2 (every layer): it loads Z, M,
 * N of level 1 and that maximum (Z_MIN is read and dropped: the device
 * keeps none), with HOTSTART_DATA's fallback of a zero plane for a file
 * that cannot be opened (*nmissing counts them), and sets the root-step
 * count to START_STEP - 1 so that the loop continues at K = START_STEP.
 * Every file is read and checked before any plane changes; a malformed or
 * short file returns 92/93 and changes nothing.  The wet maximum and the
 * arrival field are not in the files and keep their values.  The values
 * are rounded to 5 decimals, so a resumed run differs from the run that
 * wrote them; exact resumption is gcomcot_checkpoint_save_/load_. */
int gcomcot_hotstart_write_(int *start_step, const char *dir, int *dirlen);
int gcomcot_hotstart_read_(int *start_type, int *start_step, const char *dir, int *dirlen,
                           int *nmissing);
/* One such file and a host plane of *nx x *ny cells (column-major), for
 * cells I, J >= *first (1 or 2); host only.  The reader returns 96 when the
 * file cannot be opened. */
int gcomcot_hotstart_file_write_(const char *path, int *pathlen, const float *plane, int *nx,
                                 int *ny, int *first);
int gcomcot_hotstart_file_read_(const char *path, int *pathlen, float *plane, int *nx, int *ny,
                                int *first);

/* Boundary treatments of the outermost layer.
 * gcomcot_sponge_coef_ is SPONGE_COEF (boundaries.f90:250-330, with H_CALC
 * :334-365) on the host, in REAL: cx[NX] = COEFX(I, .), cy[NY] = COEFY(., J)
 * (each depends on one coordinate only) from the layer's H (NX*NY,
 * column-major) and coordinates X(NX), Y(NY).
 * gcomcot_sponge_set_(*on) makes the device loop damp the fluxes of layer 1
 * after every momentum pass as SPONGE_LAYER does (boundaries.f90:206-246:
 * M2 += COEFX*DT*M1, N2 += COEFY*DT*N1 where either coefficient is > 0),
 * with coefficients from its current depth and its coordinates
 * (gcomcot_set_layer_coords_); the reference's GPU build does not run it
 * (comcot.f90:403-406).  Linear layers then take one step per launch.
 * gcomcot_bc_wall_ is BC_WALL (boundaries.f90:369-401) on a host depth
 * array: the two outermost rows and columns become land (H = -999), as the
 * reference applies it before the layers are created (comcot.f90:221). */
void gcomcot_sponge_coef_(const int *laycord, const int *nx, const int *ny, const float *h,
                          const float *x, const float *y, float *cx, float *cy);
int gcomcot_sponge_set_(int *on);
void gcomcot_bc_wall_(float *h, const int *nx, const int *ny);

/* Seafloor deformation of one fault plane, computed on the device
 * (reference: GET_FLOOR_DEFORM's fault-model branch, deform.f90:2-71, with
 * DEFORM_OKADA :74-191, STRIKE_SLIP/DIP_SLIP :194-238, STEREO_PROJECTION
 * :494-588, ININTERP/CELL_INTERP all_grids.f90:805-934, UPDATE_BATH
 * initialization.f90:2410-2460).  params = the FAULT record's 11 REALs HH,
 * L, W, D, TH, DL, RD, YO, XO, Y0, X0 (type_module.f90:130-147); *fswitch 0
 * puts the epicentre at the plane's centre, 9 at the centre of its top edge.
 * The Okada field of layer 1 is evaluated in FP64 (the reference: REAL),
 * spread to nested layers in REAL as CELL_INTERP does, then every layer's
 * level-1 elevation rises and its depth falls by it.  Call it before the
 * first step for a source at T = 0 (the maximum then starts from the raised
 * surface), or between gcomcot_run_steps_ calls at a later rupture time (the
 * reference applies those on the host only, deform.f90:29-31).  A spherical
 * layer 1 needs gcomcot_set_layer_coords_ first.  gcomcot_get_deform_ reads
 * a layer's last deformation, gcomcot_get_depth_ its still-water depth H. */
int gcomcot_fault_deform_(const float *params, int *fswitch);
/* A fault that ruptures at model time *t0 inside gcomcot_run_steps_: it is
 * applied, on the device, at the call of GET_FLOOR_DEFORM that the
 * reference's loop would make with TIME <= T0 < TIME + DT (before the loop
 * with TIME = 0, else before root step k with TIME = k DT accumulated in
 * REAL; deform.f90:29, comcot.f90:323, 341, initialization.f90:476).  A new
 * layer 1 clears the schedule; gcomcot_fault_pending_ counts faults still
 * waiting. */
int gcomcot_fault_schedule_(const float *params, int *fswitch, float *t0);
int gcomcot_fault_pending_(void);
int gcomcot_get_deform_(int *lid, float *deform);
int gcomcot_get_depth_(int *lid, float *h);

/* Reference-format output files of the device loop (reference: ALL_PRT /
 * DATA_PRT, output.f90:737-851, at K = 0 and every IPRT root steps,
 * comcot.f90:308-310, 415-417; MAX_AMP, output.f90:855-1031, after every
 * root step when OUT_OPT is 0 or 2, comcot.f90:420-423).
 * gcomcot_output_set_ turns them on for gcomcot_run_steps_: files go to the
 * directory dir[0..*dirlen) (0 = the working directory, which must exist);
 * *iprt <= 0 writes no snapshots; *out_opt 0 or 2 writes the maximum files
 * and makes the step kernel keep MAX_AMP's maximum of layer 1 (the one
 * cuda_getzmax_ returns); *tend is the run's TEND in seconds and *tide the
 * TIDE_LEVEL.  Snapshots z_LL_KKKKKK.dat (dry land 0, tide added) and, for
 * FLUXSWITCH 0, m_/n_LL_KKKKKK.dat; maxima zmax_layer01_MMMMmins.dat at each
 * half hour of model time and zmax_layer01.dat from TEND - 2 DT on; each file
 * one Fortran unformatted sequential record (4-byte length, the (NX,NY)
 * REAL plane, 4-byte length), as the reference writes them.
 * gcomcot_output_layer_ sets a layer's FLUXSWITCH (0: z, m, n; 1 or 9: z;
 * others: none; default 1).  Files are written by a thread of the library
 * from pinned host copies while the loop continues; gcomcot_output_flush_
 * waits until every file is on disk and reports how many were written
 * (cuda_shutdown_ and a new layer 1 flush too).  A file that could not be
 * written does not stop the loop: the next flush returns its error, once.
 * Each returns 0 or an error code. */
int gcomcot_output_set_(const char *dir, int *dirlen, int *iprt, int *out_opt, float *tend,
                        float *tide);
int gcomcot_output_layer_(int *lid, int *fluxswitch);
int gcomcot_output_flush_(long long *nfiles);

/* Autotuner: for an outermost layer without nested layers, time every
 * tile variant of the single-step kernel and, for linear layers, every
 * shape of the two-step kernel, *reps launches each, on the live layer
 * with the maximum and the step counter switched off (such launches read
 * level 1 and write only level 2, so the state is unchanged); keep the
 * fastest for gcomcot_run_steps_ / gcomcot_time_steps_.  Reports the
 * choice and its time per root step.  In a tree with nested layers each
 * layer without nested layers of its own (the ones that run fused steps)
 * gets its own fastest tile variant the same way; *variant is then that of
 * the layer with the most cell updates per root step, *ms_per_step the sum
 * of those layers' fused launches per root step.  gcomcot_set_tile_variant_
 * clears the per-layer choices.  Returns 0 or an error code. */
int gcomcot_autotune_(int *reps, int *variant, int *blocking, float *ms_per_step);

/* Two root steps per launch (temporal blocking) in gcomcot_run_steps_ for
 * a linear outermost layer without nested layers: *mode 0 = off, 1..7 =
 * on with tiles of 64x8, 32x16, 64x16 (256 threads), 64x16 (512 and 1024
 * threads), 32x16 (512 threads), 64x32 (1024 threads) (default 1; the
 * autotuner picks among them).  Results are bitwise
 * those of single steps; afterwards level 2 holds the state two steps
 * before level 1 when the last launch was a two-step one.  Returns 0 if
 * valid. */
int gcomcot_set_temporal_blocking_(int *mode);

/* For an outermost linear layer with nested layers: the momentum launch of
 * each root step but the last of a gcomcot_run_steps_ / gcomcot_time_steps_
 * call also computes the next root step's continuity and open boundary
 * (*on = 1, the default), which the root step order allows (nothing runs
 * between the two), saving one launch and one pass over the layer per
 * step; *on = 0: separate continuity and momentum launches.  The same bits
 * either way; between calls the state is always that of a whole step.
 * gcomcot_batch_set_fused_halves_ is the same switch for the batch engine.
 * Returns 0. */
int gcomcot_set_fused_halves_(int *on);

/* How the step launches of the device loop and the drop-in entry points
 * store the new time level (the batch engine's launches always store
 * through the caches): *mode = 1 (the default) streams it past the caches
 * (nontemporal stores) for layers of more than about 5.6 M cells, whose
 * one step moves more bytes than the 256 MB last-level cache holds, and
 * for layers of at most about 0.93 M cells, whose step fits in the L2s
 * (streaming spares every kernel end the write-back of the new level),
 * and stores it through the caches for the layers between, whose next
 * step or next phase finds it in the last-level cache; 0 = always through
 * the caches, 2 = always streamed.  The same bits every way.
 * The batch engine rebuilds its plan at its next run call.  Returns 0,
 * or 46 for another *mode. */
int gcomcot_set_store_policy_(int *mode);

/* Measurement aid: *reps copies of a *mib MiB buffer with a plain
 * streaming kernel; *ms_per_copy = mean duration (HIP events).  Gives the
 * chip's practical HBM bandwidth and the scale of rocprofv3's memory
 * counters for a transfer of known size. */
int gcomcot_hbm_copy_(int *mib, int *reps, float *ms_per_copy);

/* Self-check of the short division and square root sequences of the
 * nonlinear momentum kernels against the IEEE operations, on the device:
 * x / 12 for every float (*bad_div12 = mismatches where |x| >= 2^-124 or
 * x = 0; *example_x = the bits of one), sqrt(x) for every x = +-0 or
 * x >= 2^-96 (*bad_sqrt, *example_s), and a / b for about *pairs
 * pseudo-random operand pairs drawn from seed *seed (*bad_inrange =
 * mismatches among those inside the range the sequence is claimed for:
 * both normal, |b| within 2^+-120, |a| >= 2^-103, exponent gap below 96,
 * normal quotient; *example_a/_b = the bits of one; *tested = pairs
 * drawn).  All counts are 0 on a correct build. */
int gcomcot_division_check_(long long *pairs, unsigned long long *seed, long long *bad_div12,
                            long long *bad_inrange, long long *bad_sqrt, long long *tested,
                            unsigned *example_x, unsigned *example_a, unsigned *example_b,
                            unsigned *example_s);

/* The Manning depth power x^2.3333333 and the breaking exponential e^x as
 * the nonlinear momentum kernels compute them (moment.f90:459-460 and
 * :1323 call libm's powf and expf there), evaluated on the device for the
 * n arguments x[] into pow_out[] and exp_out[] (host arrays).  A check of
 * the device arithmetic against the same functions compiled for the host. */
int gcomcot_detmath_eval_(long long *n, const float *x, float *pow_out, float *exp_out);

/* gcomcot_detmath_eval_ without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The device
 * copies of x and of both results live in a static scratch array of the
 * library (12 MiB per device, held until the process ends), so n is at
 * most 1048576 (2^20); a larger n returns 37 and touches nothing.  Calls on
 * one device share that array: calls that overlap on one device race.
 * Same results and the other return codes as gcomcot_detmath_eval_. */
int gcomcot_detmath_eval_noalloc_(long long *n, const float *x, float *pow_out,
                                  float *exp_out);

/* ------------------------------------------------------------------ */
/* Part 3: scenario batch engine (no reference counterpart)            */
/* ------------------------------------------------------------------ */
/* Many independent scenarios, each an outermost layer with at most one
 * level of nested layers and its own number of root steps, advanced
 * together: every phase of a root step is one launch over all active
 * scenarios (a persistent kernel draining per-XCD tile queues), two root
 * steps are replayed from a HIP graph, and fused tile shapes are picked
 * by an on-line autotuner.  Each scenario's planes are bit-identical to a
 * run of the same layers alone through Part 1/2.  Scenarios can be added
 * and removed between calls (continuous batching).  All calls return 0 or
 * an error code (see gcomcot_set_error_mode_ / gcomcot_last_error_). */

/* drop every scenario and its device memory */
int gcomcot_batch_reset_(void);
/* new, empty scenario that will run *steps root steps; id in *sid */
int gcomcot_batch_new_scenario_(long long *steps, int *sid);
/* layer of scenario *sid; arguments as gcomcot_init_layer_ex_ (layer 1
 * first, then children with *parent == 1) */
int gcomcot_batch_add_layer_(int *sid, int *layerid, int *parent, int *laycord, int *laygov,
                             int *rel_size, int *rel_time, int *corners, int *fric_switch,
                             float *fric_coef, float *dt, float *dx, float *dy, float *coef,
                             float *H, float *Z, int *nx, int *ny);
/* a Cartesian child of scenario *sid's spherical layer 1 with SC_OPTION = 1;
 * arguments as gcomcot_init_layer_sc_ */
int gcomcot_batch_add_layer_sc_(int *sid, int *layerid, int *parent, int *laygov,
                                int *rel_size, int *rel_time, int *corners, int *fric_switch,
                                float *fric_coef, float *dt, float *dx, float *dy, float *H,
                                float *Z, int *nx, int *ny, int *pos, float *cxy);
/* advance every unfinished scenario by up to *nsteps root steps (each
 * stops at its own count); *nfinished = scenarios finished and not yet
 * removed.  Returns after the device work is complete. */
int gcomcot_batch_run_(long long *nsteps, int *nfinished);
/* ids of finished scenarios still held (up to *cap written); returns how
 * many there are */
int gcomcot_batch_finished_(int *sids, int *cap);
int gcomcot_batch_progress_(int *sid, long long *done, long long *total);
/* level-1 (current) Z, M, N of layer *lid of scenario *sid (NULL skips) */
int gcomcot_batch_get_state_(int *sid, int *lid, float *z, float *m, float *n);
/* wet-only maximum and first-arrival root step of layer *lid */
int gcomcot_batch_get_max_(int *sid, int *lid, float *zmax_wet, int *arrival);
int gcomcot_batch_remove_(int *sid);
/* device bytes held by scenario *sid (0: all scenarios) */
long long gcomcot_batch_device_bytes_(int *sid);
/* device bytes one layer of nx x ny cells takes (admission control) */
long long gcomcot_layer_footprint_(int *nx, int *ny, int *nonlinear);
void gcomcot_batch_set_arrival_threshold_(float *thr);
/* *autotune != 0: time the launch forms and tile shapes on the live batch;
 * *shape >= 0 forces one for the fused launches: 0/1/2 = persistent kernel
 * with 64x8 / 32x16 / 64x16 tiles, 3/4/5 = one workgroup per tile with
 * the same tiles */
int gcomcot_batch_set_tuning_(int *autotune, int *shape);

/* A fault source for a batched scenario (see gcomcot_fault_deform_), applied
 * to its layers on the device; spherical roots need their coordinates
 * first.  Call before the scenario's first step for a source at T = 0. */
int gcomcot_batch_set_coords_(int *sid, int *lid, const float *x, const float *y);
int gcomcot_batch_fault_deform_(int *sid, const float *params, int *fswitch);
/* *on == 0: enqueue every root step instead of replaying HIP graphs (for
 * tools that cannot follow graph launches) */
int gcomcot_batch_set_graphs_(int *on);
/* *on != 0 (the default): scenarios whose only layer is linear may advance
 * two root steps per launch (the two-step kernel of gcomcot_set_temporal_
 * blocking_, tile shape *shape = its mode - 1); plans then cover four root
 * steps.  shape NULL (the default): the autotuner times every two-step
 * shape on those layers and takes the fastest if it beats two one-step
 * launches, else they keep one step per launch.  Results are bitwise those
 * of one step per launch either way. */
int gcomcot_batch_set_two_step_(int *on, int *shape);
/* Streams of a plan: *n = 3 (the default): the launches of single-layer
 * scenarios run on a second stream and the nested scenarios are split into
 * two chains of launches (balanced by cell updates), one on the first
 * stream and one on a third, all forked at the plan's start and joined at
 * its end (parallel branches of the plan's graph); *n = 2: the nested
 * scenarios in one chain; *n = 1: everything on one stream.  The same bits
 * every way.  Returns 0, or 45 for another *n. */
int gcomcot_batch_set_streams_(int *n);
/* Nested linear roots in the batch: each momentum launch also computes the
 * next root step's continuity (*on = 1, the default; see
 * gcomcot_set_fused_halves_).  Same bits.  Returns 0. */
int gcomcot_batch_set_fused_halves_(int *on);
int gcomcot_batch_info_(long long *launches, long long *graph_replays, long long *plans,
                        long long *cells_per_step, int *tuned_count);

#ifdef __cplusplus
}
#endif
#endif
