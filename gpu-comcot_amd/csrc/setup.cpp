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

}  // extern "C"
