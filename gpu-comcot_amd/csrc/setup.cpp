// Host-side layer setup: per-row spherical coefficients and the time step.
//
// PARA (comcot.f90:694-804) fills nine (NX,NY) coefficient planes; all of
// them are constant along I except R2/R4 of linear layers, which are a row
// factor times max(HP,0) / max(HQ,0).  The device keeps only the row
// factors (kNumCoef floats per row) and rebuilds R2/R4 from H, so the step
// reads 4 planes instead of 13.  Values are computed in FP32 with the same
// operation order as the Fortran, so R(I,J) = coef[.][J] bit for bit.
#include <cmath>

#include "comcot_consts.h"
#include "gcomcot_hip_internal.h"

using namespace gc;

extern "C" {

void gcomcot_para_rows_(const int *laygov, const int *ny, const float *dt, const float *dx,
                        const float *y, const float *dely, float *coef) {
  const int NY = *ny;
  const float RR = *dt / kREarth;
  const float RS = kGrav * RR;
  const float RT = 0.5f * *dt * kCoriW;
  (void)laygov;  // F2/F4 are the same factor for LAYGOV 0 and 1
  for (int j = 0; j < NY; ++j) {
    float ang_m = y[j] * kRadDeg;
    float ang_n = (y[j] + 0.5f * dely[j] * kOSixty) * kRadDeg;
    float cos_m = cosf(ang_m), cos_n = cosf(ang_n);
    float sin_m = sinf(ang_m), sin_n = sinf(ang_n);
    coef[kR0 * NY + j] = RR / (dely[j] * kRadMin);
    coef[kR1 * NY + j] = RR / cos_m / (*dx * kRadMin);
    coef[kR11 * NY + j] = RR / cos_m / (dely[j] * kRadMin);
    coef[kF2 * NY + j] = RS / cos_m / (*dx * kRadMin);
    coef[kR21 * NY + j] = RR / cos_m / (*dx * kRadMin);
    coef[kR22 * NY + j] = RR / cos_n / (*dx * kRadMin);
    coef[kR3 * NY + j] = RT * sin_m;
    coef[kF4 * NY + j] = RS / (dely[j] * kRadMin);
    coef[kR5 * NY + j] = RT * sin_n;
    coef[kR6 * NY + j] = cos_n;
  }
}

// CR_CHECK for a first-level layer (initialization.f90:1163-1196): the time
// step shrunk to Courant number 0.5 (0.35 for nonlinear layers) if needed.
float gcomcot_cr_dt_(const int *laycord, const int *laygov, const int *nx, const int *ny,
                     const float *h, const float *dx, const float *dy, const float *lat_first,
                     const float *lat_last, const float *dt_in) {
  float h_max = kGx;
  const long n = (long)*nx * *ny;
  for (long k = 0; k < n; ++k)
    if (h[k] > h_max) h_max = h[k];
  float DX, DY;
  if (*laycord == 0) {
    float lat_max = fmaxf(fabsf(*lat_first), fabsf(*lat_last)) * kRadDeg;
    DX = kREarth * cosf(lat_max) * (*dx * kRadMin);
    DY = kREarth * (*dy * kRadMin);
  } else {
    DX = *dx;
    DY = *dy;
  }
  float del_x = fminf(DX, DY);
  float dt = *dt_in;
  float cr = *dt_in / (del_x / sqrtf(kGrav * h_max));
  float cr_limit = (*laygov == 1 || *laygov == 3) ? 0.35f : 0.5f;
  if (cr > cr_limit) dt = cr_limit * del_x / sqrtf(kGrav * h_max);
  return dt <= *dt_in ? dt : *dt_in;
}

// Child time step (initialization.f90:1206-1247): REL_TIME and DT of a layer
// nested in a parent with time step dt_parent.
void gcomcot_child_dt_(const int *laycord, const int *laygov, const int *nx, const int *ny,
                       const float *h, const float *dx, const float *dy, const float *lat_first,
                       const float *lat_last, const float *dt_parent, int *rel_time,
                       float *dt_child) {
  float h_max = kGx;
  const long n = (long)*nx * *ny;
  for (long k = 0; k < n; ++k)
    if (h[k] > h_max) h_max = h[k];
  float DX, DY;
  if (*laycord == 0) {
    float lat_max = fmaxf(fabsf(*lat_first), fabsf(*lat_last)) * kRadDeg;
    DX = kREarth * cosf(lat_max) * (*dx * kRadMin);
    DY = kREarth * (*dy * kRadMin);
  } else {
    DX = *dx;
    DY = *dy;
  }
  float cr_limit = (*laygov == 1 || *laygov == 3) ? 0.35f : 0.5f;
  float del_x = fminf(DX, DY);
  float dt = cr_limit * del_x / sqrtf(kGrav * h_max);
  if (dt >= *dt_parent) {
    *dt_child = *dt_parent;
    *rel_time = 1;
  } else {
    int ir = (int)floorf(*dt_parent / dt) + 1;
    *rel_time = ir;
    *dt_child = *dt_parent / (float)ir;
  }
}

// SPONGE_COEF (boundaries.f90:250-330) with H_CALC (:334-365), in REAL:
// the damping coefficients of the sponge layer.  COEFX depends on X(I)
// only and COEFY on Y(J) only, so they come out as cx[NX] and cy[NY].
// The mean wet depth is summed in the Fortran loop order (I outer, J
// inner).  For spherical layers the Fortran compares a width in degrees
// with R_MAX in metres, which leaves every coefficient at 0; that is kept.
void gcomcot_sponge_coef_(const int *laycord, const int *nx, const int *ny, const float *h,
                          const float *x, const float *y, float *cx, float *cy) {
  const int NX = *nx, NY = *ny;
  float h_sum = 0.0f;
  int k = 0;
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY; ++j) {
      const float hv = h[(size_t)j * NX + i];
      if (hv > kGx) {
        h_sum = h_sum + hv;
        ++k;
      }
    }
  float depth = k > 0 ? h_sum / (float)k : 0.0f;
  if (depth <= kGx) depth = kGx;
  const float wavelength = 20.0f * depth;
  const float width = 2.0f * wavelength;
  const float alpha_c = 10.0f;
  const float pi = 3.14159265358979f;
  const float t = wavelength / sqrtf(kGrav * depth);
  const float omega = 2.0f * pi / t;
  const float sponge_width = *laycord == 0 ? width / kREarth * 180.0f / pi : width;
  const float r_max = width;
  const float xs = x[0] + sponge_width, ys = y[0] + sponge_width;
  const float xe = x[NX - 1] - sponge_width, ye = y[NY - 1] - sponge_width;
  const float e1 = expf(1.0f) - 1.0f;
  auto coef = [&](float v, float lo, float hi) {
    float r = 0.0f;
    if (v <= lo) r = -(v - lo);
    else if (v >= hi) r = v - hi;
    float rel = r;
    if (rel <= 0.0f) rel = 0.0f;
    if (rel >= r_max) rel = r_max;
    const float q = rel / r_max;
    float c = alpha_c * omega * (expf(q * q) - 1.0f) / e1;
    if (c <= kEps) c = 0.0f;
    return c;
  };
  for (int i = 0; i < NX; ++i) cx[i] = coef(x[i], xs, xe);
  for (int j = 0; j < NY; ++j) cy[j] = coef(y[j], ys, ye);
}

// BC_WALL (boundaries.f90:369-401, the branch without a wave maker): the
// two outermost rows and columns become land (H = -999).
void gcomcot_bc_wall_(float *h, const int *nx, const int *ny) {
  const int NX = *nx, NY = *ny;
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i)
      if (i < 2 || j < 2 || i >= NX - 2 || j >= NY - 2) h[(size_t)j * NX + i] = -999.0f;
}

}  // extern "C"
