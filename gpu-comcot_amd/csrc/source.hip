// Seafloor deformation of a fault on the device: GET_FLOOR_DEFORM's
// fault-model branch (deform.f90:2-71) for one fault plane.
//
//   DEFORM_OKADA (deform.f90:74-191)  k_okada on the outermost layer, FP64;
//     the stereographic projection (deform.f90:494-588) is split into
//     per-row and per-column factors computed here, on the host, in FP64
//   ININTERP / CELL_INTERP (all_grids.f90:805-934)  k_cell_interp, parent
//     to child, level by level, in REAL (CELL_INTERP_SC, :938-992, for
//     SC_OPTION = 1 children)
//   UPDATE_BATH (initialization.f90:2410-2460) and Z(:,:,1) += DEFORM
//     (deform.f90:55-64)  k_apply_deform on every layer
//
// Constants are the reference's REAL literals (each rounded to float), so
// the FP64 evaluation is the reference's formula without its FP32
// intermediate rounding.
#include <cmath>
#include <vector>

#include "source.h"

namespace gc {
namespace {

double f32(double v) { return (double)(float)v; }

const double kRadDeg = f32(0.01745329252);
const double kEpsD = f32(1.0e-10);
const double kPiD = f32(3.14159265358979);
const double kA = f32(6378137.0000);
const double kE = f32(0.08181919092891);
const double kF2 = f32(0.00669438000426);
const double kK0 = f32(0.9996);

// The projection's quantities that depend on the tangent point only.
struct ProjOrigin {
  double r, n, c, lm0, sxi0, cxi0, two_rk0, pole;
};

ProjOrigin proj_origin(double lon0, double lat0) {
  ProjOrigin o;
  o.pole = kPiD / 2.0 - kEpsD;
  double lt0 = lat0 * kRadDeg;
  if (lt0 > o.pole) lt0 = o.pole;
  if (lt0 < -o.pole) lt0 = -o.pole;
  const double ln0 = lon0 * kRadDeg;
  const double cs0 = std::cos(lt0), sn0 = std::sin(lt0);
  const double tmp0 = std::sqrt(1.0 - kF2 * std::pow(sn0, 2.0));
  const double rho0 = kA * (1.0 - kF2) / std::pow(tmp0, 3.0);
  const double nu0 = (double)(long long)(kA / tmp0);  // INTEGER by implicit typing
  o.r = std::sqrt(rho0 * nu0);
  o.n = std::sqrt(1.0 + kF2 * std::pow(cs0, 4.0) / (1.0 - kF2));
  const double s1 = (1.0 + sn0) / (1.0 - sn0);
  const double s2 = (1.0 - kE * sn0) / (1.0 + kE * sn0);
  const double w1 = std::pow(s1 * std::pow(s2, kE), o.n);
  const double sn_xi0 = (w1 - 1.0) / (w1 + 1.0);
  o.c = (o.n + sn0) * (1.0 - sn_xi0) / (o.n - sn0) / (1.0 + sn_xi0);
  const double w2 = o.c * w1;
  const double xi0 = std::asin((w2 - 1.0) / (w2 + 1.0));
  o.lm0 = ln0;
  o.sxi0 = std::sin(xi0);
  o.cxi0 = std::cos(xi0);
  o.two_rk0 = 2.0 * o.r * kK0;
  return o;
}

// XI of a latitude (degrees)
double proj_xi(const ProjOrigin &o, double lat) {
  double la = lat * kRadDeg;
  if (la > o.pole) la = o.pole;
  if (la < -o.pole) la = -o.pole;
  const double sn = std::sin(la);
  const double sa = (1.0 + sn) / (1.0 - sn);
  const double sb = (1.0 - kE * sn) / (1.0 + kE * sn);
  const double w = o.c * std::pow(sa * std::pow(sb, kE), o.n);
  return std::asin((w - 1.0) / (w + 1.0));
}

// LM - LM0 of a longitude (degrees)
double proj_dlm(const ProjOrigin &o, double lon) {
  const double lm = o.n * (lon * kRadDeg - o.lm0) + o.lm0;
  return lm - o.lm0;
}

// STEREO_PROJECTION of one point
void proj_point(const ProjOrigin &o, double lon, double lat, double *x, double *y) {
  const double xi = proj_xi(o, lat), dl = proj_dlm(o, lon);
  const double beta = 1.0 + std::sin(xi) * o.sxi0 + std::cos(xi) * o.cxi0 * std::cos(dl);
  *y = o.two_rk0 * (std::sin(xi) * o.cxi0 - std::cos(xi) * o.sxi0 * std::cos(dl)) / beta;
  *x = o.two_rk0 * std::cos(xi) * std::sin(dl) / beta;
}

int ensure_deform(LayerRec &L) {
  if (!L.deform) HIPCHK(hipMalloc(&L.deform, L.nxy() * sizeof(float)));
  return 0;
}

}  // namespace

int tree_apply_fault(Tree &T, hipStream_t st, const float *params, int fswitch) {
  T.mass_ahead = false;  // level 2 is no longer the next continuity
  LayerRec *R = T.layer(1);
  if (!R) return fail(8, "fault: no root layer");
  if (fswitch != 0 && fswitch != 9) return fail(60, "fault: epicentre switch %d not 0 or 9", fswitch);
  for (int k = 0; k < 11; ++k)
    if (!std::isfinite(params[k])) return fail(61, "fault: parameter %d is not finite", k + 1);
  const double hh0 = params[0], l = params[1], w = params[2], d = params[3];
  const double th = params[4], dl = params[5], rd = params[6];
  const double yo = params[7], xo = params[8], y0 = params[9], x0 = params[10];
  if (R->sph && ((int)T.xs[1].size() != R->nx || (int)T.ys[1].size() != R->ny))
    return fail(42, "fault: coordinates of layer 1 not set");
  // nested layers read one parent cell around their footprint (CELL_INTERP)
  for (int id = 2; id < kMaxLayers; ++id) {
    const LayerRec *C = T.layer(id);
    if (!C) continue;
    const LayerRec &P = T.layers[C->parent];
    if (C->sc) continue;  // CELL_INTERP_SC reads the parent through POS
    if (C->corners[0] < 2 || C->corners[1] + 1 > P.nx || C->corners[2] < 2 ||
        C->corners[3] + 1 > P.ny)
      return fail(62, "fault: layer %d touches its parent's edge; deformation cannot be spread", id);
  }

  FaultDev f{};
  f.nx = R->nx;
  f.ny = R->ny;
  f.sph = R->sph;
  const double ang_l = kRadDeg * dl, ang_r = kRadDeg * rd, ang_t = kRadDeg * th;
  f.halfl = 0.5 * l;
  if (fswitch == 9) {  // epicentre at the centre of the top edge
    f.hh = hh0 + w * std::sin(ang_l);
    f.del_x = w * std::cos(ang_l) * std::cos(ang_t);
    f.del_y = w * std::cos(ang_l) * std::sin(ang_t);
  } else {             // epicentre at the centre of the plane
    f.hh = hh0 + 0.5 * w * std::sin(ang_l);
    f.del_x = 0.5 * w * std::cos(ang_l) * std::cos(ang_t);
    f.del_y = 0.5 * w * std::cos(ang_l) * std::sin(ang_t);
  }
  f.w = w;
  f.ds = d * std::cos(ang_r);
  f.dd = d * std::sin(ang_r);
  f.sn = std::sin(ang_l);
  f.cs = std::cos(ang_l);
  f.sin_t = std::sin(ang_t);
  f.cos_t = std::cos(ang_t);
  const ProjOrigin o = proj_origin(x0, y0);
  std::vector<double> tab;
  if (R->sph) {
    tab.resize(2 * ((size_t)R->ny + R->nx));
    for (int j = 0; j < R->ny; ++j) {
      const double xi = proj_xi(o, (double)T.ys[1][j]);
      tab[2 * j] = std::sin(xi);
      tab[2 * j + 1] = std::cos(xi);
    }
    double *col = tab.data() + 2 * (size_t)R->ny;
    for (int i = 0; i < R->nx; ++i) {
      const double dlm = proj_dlm(o, (double)T.xs[1][i]);
      col[2 * i] = std::sin(dlm);
      col[2 * i + 1] = std::cos(dlm);
    }
    f.two_rk0 = o.two_rk0;
    f.sxi0 = o.sxi0;
    f.cxi0 = o.cxi0;
  } else {
    f.dx = (double)R->dx;
    f.dy = (double)R->dy;
    proj_point(o, xo, yo, &f.xl, &f.yl);
  }

  double *dtab = nullptr;
  if (!tab.empty()) {
    HIPCHK(hipMalloc(&dtab, tab.size() * sizeof(double)));
    HIPCHK(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    f.row = dtab;
    f.col = dtab + 2 * (size_t)R->ny;
  }
  int rc = ensure_deform(*R);
  hipError_t e = hipSuccess;
  if (!rc) e = launch_okada(f, R->deform, st);
  // ININTERP: every parent before its children (breadth first from layer 1)
  std::vector<int> order{1};
  for (size_t k = 0; k < order.size(); ++k)
    for (int cid : T.layers[order[k]].children) order.push_back(cid);
  for (size_t k = 1; !rc && e == hipSuccess && k < order.size(); ++k) {
    LayerRec &C = T.layers[order[k]];
    rc = ensure_deform(C);
    const LayerRec &P = T.layers[C.parent];
    if (!rc && C.sc)  // CELL_INTERP_SC (all_grids.f90:938-992)
      e = launch_sc_cell_interp(P.deform, P.nx, P.ny, C.deform, C.nx, C.ny, C.sc_pos, C.sc_cxy, st);
    else if (!rc)
      e = launch_cell_interp(P.deform, P.nx, C.deform, C.nx, C.ny, C.corners, C.rel_size, st);
  }
  // before the first step the maximum starts from the deformed surface, as
  // when the source is part of the initial state
  int done = 0;
  if (!rc && e == hipSuccess && T.ctl) {
    e = hipMemcpyAsync(&done, T.ctl + T.ctl_in, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  for (int id = 1; !rc && e == hipSuccess && id < kMaxLayers; ++id) {
    LayerRec *L = T.layer(id);
    if (!L) continue;
    e = launch_apply_deform(L->z[L->cur], L->h, L->deform, done == 0 ? L->zmax_raw : nullptr,
                            L->nxy(), st);
  }
  const hipError_t es = hipStreamSynchronize(st);
  if (dtab) (void)hipFree(dtab);
  if (rc) return rc;
  if (e != hipSuccess) return fail(206, "fault kernels failed: %s", hipGetErrorString(e));
  if (es != hipSuccess) return fail(206, "fault kernels failed: %s", hipGetErrorString(es));
  return 0;
}

}  // namespace gc
