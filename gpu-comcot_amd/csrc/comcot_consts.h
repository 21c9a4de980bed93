// Physical and numerical constants of the COMCOT step, as the Fortran REAL
// (FP32) values of BLOCK DATA /CONS/ (comcot.f90:454-470), PARA's DATA
// statement (comcot.f90:709-710), OPEN's UB (boundaries.f90:30) and the
// literal constants of the momentum routines (moment.f90).
#pragma once

namespace gc {

constexpr float kElmax = -50.0f;    // ELMAX: land higher than this never floods
constexpr float kGrav = 9.807f;     // GRAV
constexpr float kREarth = 6378000.0f;
constexpr float kGx = 1.0e-5f;      // GX: dry threshold on total depth
constexpr float kEps = 1.0e-10f;    // EPS: values below are flushed to 0
constexpr float kVLimit = 20.0f;    // V_LIMIT: |flux| <= V_LIMIT * depth
constexpr float kRadDeg = 0.01745329252f;
constexpr float kRadMin = 0.000290888208665721f;
constexpr float kCoriW = 7.2722e-5f;
constexpr float kOSixty = 0.016666666667f;
constexpr float kUB = 99.0f;        // OPEN clamps |eta| > UB to 0
constexpr float kTwlvth = 0.08333333333333f;
constexpr float kManExp = 2.3333333f;  // Manning depth exponent
constexpr int kSpan = 10;           // convection buffer width (moment.f90:368, 835)

// Per-row coefficient table of a spherical layer, one float per row J,
// derived from PARA (comcot.f90:717-751).  The (I,J) planes of PARA are
// constant along I except R2/R4 of linear layers, which are
// F2[J]*max(HP,0) and F4[J]*max(HQ,0); the kernels rebuild them from H.
enum Coef {
  kR0 = 0,   // RR/(DEL_Y*RAD_MIN)
  kR1,       // RR/COS_M/(DX*RAD_MIN)
  kR11,      // RR/COS_M/(DEL_Y*RAD_MIN)
  kF2,       // RS/COS_M/(DX*RAD_MIN)       (R2 = F2*max(HP,0) if linear, F2 if nonlinear)
  kR21,      // RR/COS_M/(DX*RAD_MIN)
  kR22,      // RR/COS_N/(DX*RAD_MIN)
  kR3,       // RT*SIN_M
  kF4,       // RS/(DEL_Y*RAD_MIN)          (R4 analogous)
  kR5,       // RT*SIN_N
  kR6,       // COS_N
  kNumCoef
};

}  // namespace gc
