// HIP kernels of the COMCOT shallow-water step for gfx950 (MI355X).
//
// One layer-step is one kernel (kFused): each 256-thread workgroup owns a
// TX x TY tile of cells, stages Z1, M1, N1, H of the tile plus a two-cell
// halo in LDS (each HBM byte read once; the halo comes from neighbouring
// tiles, which the XCD-aware tile order keeps in the same L2), computes
// the new elevation Z2 (continuity, with the radiation boundary on the
// outer ring of the outermost layer) for the tile plus the halo it needs,
// then the new fluxes M2, N2 and the wet maximum / first arrival, and
// writes Z2, M2, N2 once.  28 B/cell-update of HBM traffic (36 with the
// maximum), against 68-80 for the reference's 8 launches.
//
// Layers with nested children need the parent's Z2 between continuity and
// momentum (the child writes the footprint, JNZ), so they run kMassOnly,
// the child sub-steps, then kMomentOnly.
//
// Arithmetic follows the reference Fortran operation by operation (no FMA
// contraction: built with -ffp-contract=off; IEEE division and sqrt), so
// results equal the FP32 Fortran bit for bit; pow/exp use detmath.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "comcot_consts.h"
#include "detmath.h"
#include "fastdiv.h"
#include "kernels.h"

namespace gc {

namespace {

// Store of a new level-2 plane value; ST: streamed past the caches
// (DevLayer::stream_out).  ST is a template argument of the single-layer
// kernels (k_step, k_step2), chosen at launch: a run-time test around the
// store lets the compiler merge the two stores and drop the hint, and a
// third tile body in every kernel costs the register allocation of the
// others (more spills).  Edge tiles and the batch kernels store through
// the caches.
template <bool ST>
__device__ __forceinline__ void st_plane(float *p, float v) {
  if constexpr (ST)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

constexpr int kThreads = 256;

template <int P>
struct Lds {
  float *z1, *m1, *n1, *h, *z2, *dz;
  int gi0, gj0;  // global (1-based) index of LDS element 0
  const gcdet::Tables *tab;  // pow/exp tables, LDS copy
  const float *crow;         // per-row coefficients of rows gj0.., LDS copy (kNumCoef x R)
  int R;
  __device__ __forceinline__ int o(int i, int j) const { return (j - gj0) * P + (i - gi0); }
  __device__ __forceinline__ float CF(const DevLayer &, int k, int j) const {
    return crow[k * R + (j - gj0)];
  }
  __device__ __forceinline__ float Z1(int i, int j) const { return z1[o(i, j)]; }
  __device__ __forceinline__ float Z2(int i, int j) const { return z2[o(i, j)]; }
  __device__ __forceinline__ float M(int i, int j) const { return m1[o(i, j)]; }
  __device__ __forceinline__ float N(int i, int j) const { return n1[o(i, j)]; }
  __device__ __forceinline__ float H(int i, int j) const { return h[o(i, j)]; }
  __device__ __forceinline__ float DZ(int i, int j) const { return dz[o(i, j)]; }
};

// PQ_DEPTH (initialization.f90:2512-2523).  IN: the cell is known to lie
// away from the grid edges, so no index clamp can apply.
template <bool IN, class A>
__device__ __forceinline__ float hp_at(const DevLayer &L, const A &s, int i, int j) {
  const float h = s.H(i, j);
  if (h > kGx) {
    const int ip1 = (IN || i + 1 < L.nx) ? i + 1 : L.nx;
    return 0.5f * (h + s.H(ip1, j));
  }
  return 0.0f;
}
template <bool IN, class A>
__device__ __forceinline__ float hq_at(const DevLayer &L, const A &s, int i, int j) {
  const float h = s.H(i, j);
  if (h > kGx) {
    const int jp1 = (IN || j + 1 < L.ny) ? j + 1 : L.ny;
    return 0.5f * (h + s.H(i, jp1));
  }
  return 0.0f;
}

__device__ __forceinline__ size_t gidx(const DevLayer &L, int i, int j) {
  return (size_t)(j - 1) * L.nx + (i - 1);
}

// OPEN (boundaries.f90:22-202) for one ring cell of the outermost layer.
// The Fortran reads level-2 fluxes, which equal level 1 at that point of
// the step (CHANGE copied them); the ring is written once per cell.
template <class A>
__device__ float open_cell(const DevLayer &L, const A &s, int i, int j) {
  const int nx = L.nx, ny = L.ny;
  if ((i == 1 || i == nx) && (j == 1 || j == ny)) {
    float qx, qy;
    bool neg;
    if (i == 1 && j == 1) {
      qx = s.M(1, 1); qy = s.N(1, 1); neg = qx > 0.0f || qy > 0.0f;
    } else if (i == nx && j == 1) {
      qx = s.M(nx - 1, 1); qy = s.N(nx, 1); neg = qx < 0.0f || qy > 0.0f;
    } else if (i == 1) {
      qx = s.M(1, ny); qy = s.N(1, ny - 1); neg = qx > 0.0f || qy < 0.0f;
    } else {
      qx = s.M(nx - 1, ny); qy = s.N(nx, ny - 1); neg = qx < 0.0f || qy < 0.0f;
    }
    if (s.H(i, j) > kGx) {
      float cc = sqrtf(kGrav * s.H(i, j));
      float uh = sqrtf(qx * qx + qy * qy);
      float zz = uh / cc;
      if (neg) zz = -zz;
      if (fabsf(zz) > kUB) zz = 0.0f;
      return zz;
    }
    return 0.0f;
  }
  if (!(s.H(i, j) > kGx)) return 0.0f;
  const float cc = sqrtf(kGrav * s.H(i, j));
  float uh, uu;
  bool neg;
  if (j == 1) {
    uh = 0.5f * (s.M(i, j) + s.M(i - 1, j));
    float v = s.N(i, j);
    uu = sqrtf(uh * uh + v * v);
    neg = v > 0.0f;
  } else if (j == ny) {
    uh = 0.5f * (s.M(i, j) + s.M(i - 1, j));
    float v = s.N(i, j - 1);
    uu = sqrtf(uh * uh + v * v);
    neg = v < 0.0f;
  } else if (i == 1) {
    if (s.H(i, j - 1) > kGx) uh = 0.5f * (s.N(i, j) + s.N(i, j - 1));
    else uh = s.N(i, j);
    float u = s.M(i, j);
    uu = sqrtf(uh * uh + u * u);
    neg = u > 0.0f;
  } else {
    uh = 0.5f * (s.N(i, j) + s.N(i, j - 1));
    float u = s.M(i - 1, j);
    uu = sqrtf(uh * uh + u * u);
    neg = u < 0.0f;
  }
  float zz = uu / cc;
  if (neg) zz = -zz;
  if (fabsf(zz) > kUB) zz = 0.0f;
  return zz;
}

// Continuity for one cell: MASS_S (mass.f90:61-100) / MASS_C (:103-157),
// CONMASS_S (:227-277) / CONMASS (:162-224) when NL.  dz2 is the total
// depth DZ(:,:,2) of the nonlinear forms (0 where not computed).  RING =
// false: the caller leaves the outermost layer's open-boundary ring to
// ring_cells below.
template <bool NL, bool SPH, bool IN, bool RING = true, class A>
__device__ __forceinline__ void mass_cell(const DevLayer &L, const A &s, int i, int j,
                                          float &z2, float &dz2) {
  if (RING && !IN && L.is_root && (i == 1 || j == 1 || i == L.nx || j == L.ny)) {
    z2 = open_cell(L, s, i, j);
    dz2 = 0.0f;
    return;
  }
  if (!IN && (i < 2 || j < 2)) {  // ghost row/column of an inner layer
    z2 = 0.0f;
    dz2 = 0.0f;
    return;
  }
  const float h = s.H(i, j);
  float zz;
  if (SPH) {
    const float r1 = s.CF(L, kR1, j), r11 = s.CF(L, kR11, j);
    const float r6 = s.CF(L, kR6, j), r6m = s.CF(L, kR6, j - 1);
    zz = s.Z1(i, j) - r1 * (s.M(i, j) - s.M(i - 1, j)) -
         r11 * (s.N(i, j) * r6 - s.N(i, j - 1) * r6m);
  } else {
    zz = s.Z1(i, j) - L.rx * (s.M(i, j) - s.M(i - 1, j)) - L.ry * (s.N(i, j) - s.N(i, j - 1));
  }
  if (fabsf(zz) < kEps) zz = 0.0f;
  if (!NL) {
    if (h > kGx) {
      if ((zz + h) <= kEps) zz = -h;
      z2 = zz;
    } else {
      z2 = 0.0f;
    }
    dz2 = 0.0f;
  } else {
    const bool active = SPH ? (h >= kElmax) : (h > kElmax);
    if (active) {
      float dd = zz + h;
      if (dd >= kGx) { dz2 = dd; z2 = zz; }
      else { dz2 = 0.0f; z2 = -h; }
    } else {
      dz2 = 0.0f;
      z2 = -h;
    }
  }
}

// On the outermost layer's open-boundary ring (I = 1 or NX, J = 1 or NY)?
__device__ __forceinline__ bool on_ring(const DevLayer &L, int i, int j) {
  return L.is_root && (i == 1 || j == 1 || i == L.nx || j == L.ny);
}

// The ring cells of the rectangle [x0, x0 + CX) x [y0, y0 + CY) of the
// outermost layer, one per lane: the columns I = 1 and I = NX, then the
// rows J = 1 and J = NY without those columns.  The tile loops skip the
// ring and this pass computes OPEN there (put(i, j) stores a cell), so
// only the ring's lanes run OPEN's square roots and division: spread over
// a tile's flat order, every wave of an edge tile would.
template <int NT, class F>
__device__ __forceinline__ void ring_cells(const DevLayer &L, int x0, int y0, int CX, int CY,
                                           int tid, F put) {
  const int nx = L.nx, ny = L.ny;
  const int xa = max(x0, 1), xb = min(x0 + CX - 1, nx);
  const int ya = max(y0, 1), yb = min(y0 + CY - 1, ny);
  if (!L.is_root || xa > xb || ya > yb) return;
  const bool c0 = xa == 1, c1 = xb == nx && nx > 1, r0 = ya == 1, r1 = yb == ny && ny > 1;
  const int ia = xa + (c0 ? 1 : 0), ib = xb - (c1 ? 1 : 0);
  const int rows = yb - ya + 1, cols = ib >= ia ? ib - ia + 1 : 0;
  const int n0 = c0 ? rows : 0, n1 = c1 ? rows : 0, n2 = r0 ? cols : 0, n3 = r1 ? cols : 0;
  for (int q = tid; q < n0 + n1 + n2 + n3; q += NT) {
    int r = q, i, j;
    if (r < n0) {
      i = 1;
      j = ya + r;
    } else if ((r -= n0) < n1) {
      i = nx;
      j = ya + r;
    } else if ((r -= n1) < n2) {
      i = ia + r;
      j = 1;
    } else {
      i = ia + (r - n2);
      j = ny;
    }
    put(i, j);
  }
}

// ---------------------------------------------------------------- linear
// MOMT_S x (moment.f90:86-115) / MOMT_C x (moment.f90:190-230)
template <bool SPH, bool IN, class A>
__device__ float lin_x(const DevLayer &L, const A &s, int i, int j) {
  if (!(s.H(i, j) > kGx && s.H(i + 1, j) > kGx)) return 0.0f;
  if (SPH) {
    int jm1 = j - 1;
    if (!IN && jm1 <= 1) jm1 = 1;
    float tot_n = s.N(i, j) + s.N(i + 1, j) + s.N(i, jm1) + s.N(i + 1, jm1);
    float r2 = L.r24 ? L.r24[gidx(L, i, j)] : s.CF(L, kF2, j) * fmaxf(hp_at<IN>(L, s, i, j), 0.0f);
    float xm = s.M(i, j) - r2 * (s.Z2(i + 1, j) - s.Z2(i, j)) + s.CF(L, kR3, j) * tot_n;
    if (fabsf(xm) < kEps) xm = 0.0f;
    return xm;
  } else {
    int jm1 = j - 1, jp1 = j + 1;
    if (!IN && jm1 < 1) jm1 = 1;
    if (!IN && jp1 > L.ny) jp1 = L.ny;
    float hm = hp_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(i + 1, j));
    float xm = s.M(i, j) - L.grx * hm * (s.Z2(i + 1, j) - s.Z2(i, j));
    float ff = 0.0f;
    if (L.fric) {
      float xqq = 0.25f * (s.N(i, j) + s.N(i + 1, j) + s.N(i, jm1) + s.N(i + 1, jm1));
      if (hm >= 1.0e-3f)
        ff = L.kfric * sqrtf(s.M(i, j) * s.M(i, j) + xqq * xqq) / gcdet::powf_manning(hm, kManExp, *s.tab);
      else
        ff = 0.0f;
      xm = xm - ff * s.M(i, j);
    }
    if (L.modscm == 0) {
      float scm = L.grx * kTwlvth * hm *
                  ((s.Z2(i + 1, jp1) - 2.0f * s.Z2(i + 1, j) + s.Z2(i + 1, jm1)) -
                   (s.Z2(i, jp1) - 2.0f * s.Z2(i, j) + s.Z2(i, jm1)));
      xm = xm - scm;
    }
    if (L.fric) xm = xm / (1.0f + ff);
    if (fabsf(xm) < kEps) xm = 0.0f;
    return xm;
  }
}

// MOMT_S y (moment.f90:117-146) / MOMT_C y (moment.f90:233-273)
template <bool SPH, bool IN, class A>
__device__ float lin_y(const DevLayer &L, const A &s, int i, int j) {
  if (!(s.H(i, j) > kGx && s.H(i, j + 1) > kGx)) return 0.0f;
  if (SPH) {
    int im1 = i - 1;
    if (!IN && im1 <= 1) im1 = 1;
    float tot_m = s.M(im1, j) + s.M(im1, j + 1) + s.M(i, j) + s.M(i, j + 1);
    float r4 = L.r24 ? L.r24[(size_t)L.nx * L.ny + gidx(L, i, j)]
                     : s.CF(L, kF4, j) * fmaxf(hq_at<IN>(L, s, i, j), 0.0f);
    float xn = s.N(i, j) - r4 * (s.Z2(i, j + 1) - s.Z2(i, j)) - s.CF(L, kR5, j) * tot_m;
    if (fabsf(xn) < kEps) xn = 0.0f;
    return xn;
  } else {
    int im1 = i - 1, ip1 = i + 1;
    if (!IN && im1 < 1) im1 = 1;
    if (!IN && ip1 >= L.nx) ip1 = L.nx;
    float hn = hq_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(i, j + 1));
    float xn = s.N(i, j) - L.gry * hn * (s.Z2(i, j + 1) - s.Z2(i, j));
    float ff = 0.0f;
    if (L.fric) {
      float xpp = 0.25f * (s.M(i, j) + s.M(i, j + 1) + s.M(im1, j) + s.M(im1, j + 1));
      if (hn >= 1.0e-5f)
        ff = L.kfric * sqrtf(s.N(i, j) * s.N(i, j) + xpp * xpp) / gcdet::powf_manning(hn, kManExp, *s.tab);
      else
        ff = 0.0f;
      xn = xn - ff * s.N(i, j);
    }
    if (L.modscm == 0) {
      float scm = L.gry * kTwlvth * hn *
                  ((s.Z2(ip1, j + 1) - 2.0f * s.Z2(i, j + 1) + s.Z2(im1, j + 1)) -
                   (s.Z2(ip1, j) - 2.0f * s.Z2(i, j) + s.Z2(im1, j)));
      xn = xn - scm;
    }
    if (L.fric) xn = xn / (1.0f + ff);
    if (fabsf(xn) < kEps) xn = 0.0f;
    return xn;
  }
}

// ------------------------------------------------------------- nonlinear
// Total depths at the flux points (moment.f90:345-366, 849-870); DZ at
// level 1 is identically zero in the reference (never refreshed after
// ALLOC), so DP(I,J,1) = 0.25*(DZ2(I)+DZ2(I+1)).
template <bool IN, class A>
__device__ __forceinline__ float dp1_at(const DevLayer &L, const A &s, int i, int j) {
  const int ip1 = (IN || i + 1 < L.nx) ? i + 1 : L.nx;
  float v = 0.25f * (s.DZ(i, j) + s.DZ(ip1, j));
  return v < kGx ? 0.0f : v;
}
template <bool IN, class A>
__device__ __forceinline__ float dp2_at(const DevLayer &L, const A &s, int i, int j) {
  const int ip1 = (IN || i + 1 < L.nx) ? i + 1 : L.nx;
  float v = 0.5f * (s.DZ(i, j) + s.DZ(ip1, j));
  return v < kGx ? 0.0f : v;
}
template <bool IN, class A>
__device__ __forceinline__ float dq1_at(const DevLayer &L, const A &s, int i, int j) {
  const int jp1 = (IN || j + 1 < L.ny) ? j + 1 : L.ny;
  float v = 0.25f * (s.DZ(i, j) + s.DZ(i, jp1));
  return v < kGx ? 0.0f : v;
}
template <bool IN, class A>
__device__ __forceinline__ float dq2_at(const DevLayer &L, const A &s, int i, int j) {
  const int jp1 = (IN || j + 1 < L.ny) ? j + 1 : L.ny;
  float v = 0.5f * (s.DZ(i, j) + s.DZ(i, jp1));
  return v < kGx ? 0.0f : v;
}

// WAVE_BREAKING (moment.f90:1272-1384): the factor applied to both
// convection terms where the surface rises faster than 0.08*sqrt(g*D),
// else 1 (multiplying by 1.0f leaves every value, NaN and infinities
// included, unchanged, so one factor per cell serves both directions).
// rdt = recip_refined(DT): |Z2 - Z1| / DT is div_by with it, the IEEE
// quotient wherever its value decides anything.  A wet cell's DZDT_F is at
// least 0.08*sqrt(g*GX) (8e-4 m/s), so a numerator that leaves div_inrange's
// range gives a quotient far below it either way; on a dry cell (DZ = 0,
// DZDT_F = 0) only the quotient's sign matters, and both forms give +0 for
// 0 and a positive value otherwise, whose size REF = DZDT/0 = +inf ignores.
template <class A>
__device__ __forceinline__ float breaking_coef(const DevLayer &L, const A &s, int i, int j,
                                               float rdt) {
  float c = sqrt_inrange(kGrav * s.DZ(i, j));
  float dzdt_f = 0.08f * c;
  float dzdt_i = 0.65f * c;
  float dzdt = div_by(fabsf(s.Z2(i, j) - s.Z1(i, j)), L.dt, rdt);
  if (dzdt > dzdt_f) {
    float ref = (dzdt - dzdt_f) / (dzdt_i - dzdt_f);
    return gcdet::expf_det(-(1.0f * ref), *s.tab);
  }
  return 1.0f;
}

// X flux of CONMOME_S (moment.f90:889-1066, SPH) / CONMOME (:393-577)
template <bool SPH, bool IN, class A>
__device__ float nl_x(const DevLayer &L, const A &s, int i, int j, float brk) {
  const int IX = L.nx, JY = L.ny;
  int ip1 = i + 1, im1 = i - 1;
  if (!IN && im1 <= 1) im1 = 1;
  if (!IN && ip1 >= IX) ip1 = IX;
  int jp1 = j + 1, jm1 = j - 1;
  if (!IN && jm1 <= 1) jm1 = 1;
  if (!IN && jp1 >= JY) jp1 = JY;
  if (s.H(i, j) <= kElmax || hp_at<IN>(L, s, i, j) <= kElmax) return 0.0f;
  float DD, DF;
  const float dzi = s.DZ(i, j), dzp = s.DZ(ip1, j);
  if (dzi > kGx && dzp > kGx) {
    DD = dp2_at<IN>(L, s, i, j);
    DF = dp1_at<IN>(L, s, i, j);
  } else if (dzi > kGx && dzp <= kGx && s.H(ip1, j) + s.Z2(i, j) > kGx) {
    DD = SPH ? hp_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(ip1, j)) : 0.5f * dzi;
    DF = DD;
  } else if (dzi <= kGx && dzp > kGx && s.H(i, j) + s.Z2(ip1, j) > kGx) {
    DD = SPH ? hp_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(ip1, j)) : 0.5f * dzp;
    DF = DD;
  } else {
    return 0.0f;
  }
  if (DD < kGx) return 0.0f;
  if (DF < 1.0e-3f) DF = 1.0e-3f;
  const float p = s.M(i, j);
  const float xqq = 0.25f * (s.N(i, j) + s.N(ip1, j) + s.N(i, jm1) + s.N(ip1, jm1));
  float FF = 0.0f;
  if (L.fric)
    FF = div_inrange(L.kfric * sqrt_inrange(p * p + xqq * xqq),
                     gcdet::powf_manning(DF, kManExp, *s.tab));
  float XP;
  if (SPH) {
    const float r2 = s.CF(L, kF2, j);
    if (L.modscm == 0 && (IN || j < JY)) {
      XP = p - r2 * DD * (s.Z2(ip1, j) - s.Z2(i, j)) -
           div12(r2 * DD) *
               ((s.Z2(ip1, jp1) - 2.0f * s.Z2(ip1, j) + s.Z2(ip1, jm1)) -
                (s.Z2(i, jp1) - 2.0f * s.Z2(i, j) + s.Z2(i, jm1)));
    } else {
      XP = p - r2 * DD * (s.Z2(ip1, j) - s.Z2(i, j));
    }
    if (L.fric) XP = XP - FF * p;
    XP = XP + 4.0f * s.CF(L, kR3, j) * xqq;
  } else {
    XP = p - L.grx * DD * (s.Z2(ip1, j) - s.Z2(i, j));
    if (L.modscm == 0 && (IN || j < JY)) {
      float scm = div12(L.grx * DD) *
                  ((s.Z2(ip1, jp1) - 2.0f * s.Z2(ip1, j) + s.Z2(ip1, jm1)) -
                   (s.Z2(i, jp1) - 2.0f * s.Z2(i, j) + s.Z2(i, jm1)));
      XP = XP - scm;
    }
    if (L.fric) XP = XP - FF * p;
  }
  if (L.linchk == 1 && dp1_at<IN>(L, s, i, j) >= kGx &&
      (IN || (i > kSpan && i < IX - kSpan && j > kSpan && j < JY - kSpan))) {
    const float fx = SPH ? s.CF(L, kR21, j) : L.rx;
    const float fy = SPH ? s.CF(L, kR0, j) : L.ry;
    const float d = dp1_at<IN>(L, s, i, j);
    // Upwind terms without divergent branches: each Fortran case is one of
    // fx*(-A), fx*A, fx*(B-A), fx*(A-B) of the same two quotients.
    const bool pn = p < 0.0f;
    const int ia = pn ? ip1 : im1;
    const float da = dp1_at<IN>(L, s, ia, j);
    const bool degx = pn ? (da < kGx || s.DZ(ip1, j) < kGx) : (da < kGx || s.DZ(i, j) < kGx);
    const float yd = recip_refined(d);  // shared by the two quotients over d
    const float ax = div_by(p * p, d, yd);
    const float pa = s.M(ia, j);
    const float bx = div_inrange(pa * pa, da);
    const float advx = fx * (degx ? (pn ? -ax : ax) : (pn ? bx - ax : ax - bx));
    const bool qn = xqq < 0.0f;
    const int jn = qn ? jp1 : jm1, jf = qn ? j : j - 2;
    const float dn = dp1_at<IN>(L, s, i, jn);
    const bool degy = s.DZ(i, jn) < kGx || s.DZ(ip1, jn) < kGx || dn < kGx;
    const float ay = div_by(p * xqq, d, yd);
    const float xqe = 0.25f * (s.N(i, jn) + s.N(ip1, jn) + s.N(i, jf) + s.N(ip1, jf));
    const float by = div_inrange(s.M(i, jn) * xqe, dn);
    float advy = fy * (degy ? (qn ? -ay : ay) : (qn ? by - ay : ay - by));
    XP = XP - brk * advx - brk * advy;
  }
  if (L.fric) XP = div_inrange(XP, 1.0f + FF);
  if (fabsf(XP) < kEps) XP = 0.0f;
  const float lim = kVLimit * DD;
  if (XP > lim) XP = lim;
  if (XP < -lim) XP = -lim;
  return XP;
}

// Y flux of CONMOME_S (moment.f90:1071-1258, SPH) / CONMOME (:582-770).
// CONMOME's Y equation subtracts its friction term from the X variable
// (moment.f90:672), so only the 1/(1+FF) factor reaches XQ there.
template <bool SPH, bool IN, class A>
__device__ float nl_y(const DevLayer &L, const A &s, int i, int j, float brk) {
  const int IX = L.nx, JY = L.ny;
  int ip1 = i + 1, im1 = i - 1;
  if (!IN && im1 <= 1) im1 = 1;
  if (!IN && ip1 >= IX) ip1 = IX;
  int jp1 = j + 1, jm1 = j - 1;
  if (!IN && jm1 <= 1) jm1 = 1;
  if (!IN && jp1 >= JY) jp1 = JY;
  if (s.H(i, j) <= kElmax || hq_at<IN>(L, s, i, j) <= kElmax) return 0.0f;
  float DD, DF;
  const float dzi = s.DZ(i, j), dzp = s.DZ(i, jp1);
  if (dzi > kGx && dzp > kGx) {
    DD = dq2_at<IN>(L, s, i, j);
    DF = dq1_at<IN>(L, s, i, j);
  } else if (dzi > kGx && dzp <= kGx && s.H(i, jp1) + s.Z2(i, j) > kGx) {
    DD = SPH ? hq_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(i, jp1)) : 0.5f * dzi;
    DF = DD;
  } else if (dzi <= kGx && dzp > kGx && s.H(i, j) + s.Z2(i, jp1) > kGx) {
    DD = SPH ? hq_at<IN>(L, s, i, j) + 0.5f * (s.Z2(i, j) + s.Z2(i, jp1)) : 0.5f * dzp;
    DF = DD;
  } else {
    return 0.0f;
  }
  if (DD < kGx) return 0.0f;
  if (DF < 1.0e-3f) DF = 1.0e-3f;
  const float q = s.N(i, j);
  const float xpp = 0.25f * (s.M(i, j) + s.M(i, jp1) + s.M(im1, j) + s.M(im1, jp1));
  float FF = 0.0f;
  if (L.fric)
    FF = div_inrange(L.kfric * sqrt_inrange(q * q + xpp * xpp),
                     gcdet::powf_manning(DF, kManExp, *s.tab));
  float XQ;
  if (SPH) {
    const float r4 = s.CF(L, kF4, j);
    if (L.modscm == 0 && (IN || i < IX)) {
      XQ = q - r4 * DD * (s.Z2(i, jp1) - s.Z2(i, j)) -
           div12(r4 * DD) *
               ((s.Z2(ip1, jp1) - 2.0f * s.Z2(i, jp1) + s.Z2(im1, jp1)) -
                (s.Z2(ip1, j) - 2.0f * s.Z2(i, j) + s.Z2(im1, j)));
    } else {
      XQ = q - r4 * DD * (s.Z2(i, jp1) - s.Z2(i, j));
    }
    if (L.fric) XQ = XQ - FF * q;
    XQ = XQ - 4.0f * s.CF(L, kR5, j) * xpp;
  } else {
    XQ = q - L.gry * DD * (s.Z2(i, jp1) - s.Z2(i, j));
    if (L.modscm == 0 && (IN || i < IX)) {
      float scm = div12(L.gry * DD) *
                  ((s.Z2(ip1, jp1) - 2.0f * s.Z2(i, jp1) + s.Z2(im1, jp1)) -
                   (s.Z2(ip1, j) - 2.0f * s.Z2(i, j) + s.Z2(im1, j)));
      XQ = XQ - scm;
    }
  }
  if (L.linchk == 1 && dq1_at<IN>(L, s, i, j) >= kGx &&
      (IN || (i > kSpan && i < IX - kSpan && j > kSpan && j < JY - kSpan))) {
    const float fy = SPH ? s.CF(L, kR0, j) : L.ry;
    const float fx = SPH ? s.CF(L, kR22, j) : L.rx;
    const float d = dq1_at<IN>(L, s, i, j);
    const bool qn = q < 0.0f;
    const int ja = qn ? jp1 : jm1;
    const float da = dq1_at<IN>(L, s, i, ja);
    const bool degy = qn ? (da < kGx || s.DZ(i, jp1) < kGx) : (da < kGx || s.DZ(i, j) < kGx);
    const float yd = recip_refined(d);  // shared by the two quotients over d
    const float ay = div_by(q * q, d, yd);
    const float qa = s.N(i, ja);
    const float by = div_inrange(qa * qa, da);
    float advy = fy * (degy ? (qn ? -ay : ay) : (qn ? by - ay : ay - by));
    const bool pn = xpp < 0.0f;
    const int in = pn ? ip1 : im1, ifar = pn ? i : i - 2;
    const float dn = dq1_at<IN>(L, s, pn ? ip1 : i - 1, j);
    const bool degx = pn ? (s.DZ(ip1, j) < kGx || s.DZ(ip1, jp1) < kGx || dn < kGx)
                         : (s.DZ(im1, j) < kGx || s.DZ(im1, jp1) < kGx || dn < kGx);
    const float ax = div_by(q * xpp, d, yd);
    const float xpe = 0.25f * (s.M(in, j) + s.M(in, jp1) + s.M(ifar, j) + s.M(ifar, jp1));
    const float bx = div_inrange(s.N(in, j) * xpe, dq1_at<IN>(L, s, in, j));
    float advx = fx * (degx ? (pn ? -ax : ax) : (pn ? bx - ax : ax - bx));
    XQ = XQ - brk * advx - brk * advy;
  }
  if (L.fric) XQ = div_inrange(XQ, 1.0f + FF);
  if (fabsf(XQ) < kEps) XQ = 0.0f;
  const float lim = kVLimit * DD;
  if (XQ > lim) XQ = lim;
  if (XQ < -lim) XQ = -lim;
  return XQ;
}

// A thread's walk over a region CX cells wide in steps of NT cells of the
// flat order (idx = tid, tid + NT, ...): column i and row j advance by
// constants with one wrap test instead of a division per cell.  idx <
// CX * CY exactly when j < CY.
template <int CX, int NT>
struct Walk {
  int i, j;
  __device__ __forceinline__ explicit Walk(int tid) : i(tid % CX), j(tid / CX) {}
  __device__ __forceinline__ void next() {
    i += NT % CX;
    j += NT / CX;
    if (i >= CX) {
      i -= CX;
      ++j;
    }
  }
};

// Copy the pow/exp tables into LDS (read after the first __syncthreads).
// The two halves are split so that a tile body can issue the table loads
// with its own staging loads (TabStage), before any LDS write.
template <int NT>
struct TabStage {
  static constexpr int n = sizeof(gcdet::Tables) / sizeof(double), K = (n + NT - 1) / NT;
  double v[K];
  __device__ __forceinline__ void load(int tid) {
    const double *src = reinterpret_cast<const double *>(&gcdet::kTab);
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = tid + q * NT < n ? src[tid + q * NT] : 0.0;
  }
  // dst: the workgroup's LDS copy, which the tile reads after its barrier
  __device__ __forceinline__ void store(const gcdet::Tables *dst, int tid) const {
    double *d = reinterpret_cast<double *>(const_cast<gcdet::Tables *>(dst));
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (tid + q * NT < n) d[tid + q * NT] = v[q];
  }
};
template <int NT = kThreads>
__device__ __forceinline__ void stage_tables(gcdet::Tables &dst, int tid) {
  TabStage<NT> t;  // all loads in flight before the first LDS write
  t.load(tid);
  t.store(&dst, tid);
}

// Tile -> workgroup order: workgroup b runs on XCD b%8; XCD x gets the
// contiguous run of tiles [x*per, (x+1)*per), so vertically adjacent tiles
// (whose halos overlap) share that XCD's L2.
__device__ __forceinline__ int tile_of_block(int per_xcd) {
  const int b = blockIdx.x;
  return (b & 7) * per_xcd + (b >> 3);
}

// One TX x TY tile of layer L (first cell ti0, tj0): stage, continuity,
// momentum, outputs.  sm holds 6 (TX+4)(TY+4) planes and the per-row
// coefficients of the TY+4 rows; stab the pow/exp
// tables, staged by the caller before the call (the first barrier below
// publishes them).  IN: the tile and its halo lie more than SPAN+2 cells
// inside the grid, so no index clamp, boundary ring, ghost row or range
// limit of the Fortran loops can apply and all of them compile away.
template <int MODE, bool NL, bool SPH, int TX, int TY, bool IN, bool ST, bool TAB>
__device__ __forceinline__ void tile_body(const DevLayer &L, int t, int ti0, int tj0, float *sm,
                                          const gcdet::Tables *stab) {
  constexpr int P = TX + 4, R = TY + 4, A = P * R;
  Lds<P> s{sm,     sm + A,     sm + 2 * A, sm + 3 * A, sm + 4 * A, sm + 5 * A, ti0 - 2,
           tj0 - 2, stab, sm + 6 * A, R};
  const int nx = L.nx, ny = L.ny, tid = threadIdx.x;

  // ---- per-row coefficients of the tile's rows (spherical layers) and
  // the inputs with a two-cell halo: every load of the thread is issued
  // before its first LDS write, so the tile waits for memory once rather
  // than once per pass of the thread over the staged rectangle
  {
    constexpr int KS = (A + kThreads - 1) / kThreads;
    constexpr int KC = SPH ? (kNumCoef * R + kThreads - 1) / kThreads : 0;
    constexpr bool Z1 = MODE != kMomentOnly || NL, Z2 = MODE == kMomentOnly;
    float vh[KS], vm[KS], vn[KS], vz1[KS], vz2[KS], vdz[KS], vc[KC > 0 ? KC : 1];
    TabStage<kThreads> tb;
    if constexpr (TAB) tb.load(tid);
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int idx = tid + k * kThreads, c = idx / R, j = tj0 - 2 + idx % R;
      vc[k] = idx < kNumCoef * R && (IN || (j >= 1 && j <= ny)) ? L.coef[c * ny + (j - 1)] : 0.0f;
    }
    Walk<P, kThreads> w(tid);
#pragma unroll
    for (int k = 0; k < KS; ++k, w.next()) {
      const int idx = tid + k * kThreads;
      const int gi = ti0 - 2 + w.i, gj = tj0 - 2 + w.j;
      const bool in = (KS * kThreads == A || idx < A) &&
                      (IN || (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny));
      const size_t g = in ? (size_t)(gj - 1) * nx + (gi - 1) : 0;
      vh[k] = in ? L.h[g] : 0.0f;
      vm[k] = in ? L.m1[g] : 0.0f;
      vn[k] = in ? L.n1[g] : 0.0f;
      if (Z1) vz1[k] = in ? L.z1[g] : 0.0f;
      if (Z2) vz2[k] = in ? L.z2[g] : 0.0f;
      if (Z2 && NL) vdz[k] = in && L.dzp ? L.dzp[g] : 0.0f;
    }
    if constexpr (TAB) tb.store(stab, tid);
    if (SPH) {
      float *crow = sm + 6 * A;
#pragma unroll
      for (int k = 0; k < KC; ++k)
        if (tid + k * kThreads < kNumCoef * R) crow[tid + k * kThreads] = vc[k];
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int idx = tid + k * kThreads;
      if (!(KS * kThreads == A || idx < A)) continue;
      const int gi = ti0 - 2 + idx % P, gj = tj0 - 2 + idx / P;
      const bool in = IN || (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny);
      s.h[idx] = vh[k];
      s.m1[idx] = vm[k];
      s.n1[idx] = vn[k];
      if (Z1) s.z1[idx] = vz1[k];
      if (Z2) {
        s.z2[idx] = vz2[k];
        if (NL) {
          float dz = 0.0f;
          if (in) {
            if (L.dzp) {
              dz = vdz[k];
            } else {
              const bool massed =
                  IN || (L.is_root ? (gi >= 2 && gi <= nx - 1 && gj >= 2 && gj <= ny - 1)
                                   : (gi >= 2 && gj >= 2));
              dz = massed ? vz2[k] + vh[k] : 0.0f;
            }
          }
          s.dz[idx] = dz;
        }
      } else {
        s.z2[idx] = 0.0f;
        s.dz[idx] = 0.0f;
      }
    }
  }
  __syncthreads();

  // ---- continuity (+ open boundary) on the tile and the halo it needs
  if (MODE != kMomentOnly) {
    constexpr int CX = MODE == kFused ? TX + 3 : TX;
    constexpr int CY = MODE == kFused ? TY + 3 : TY;
    const int ci0 = MODE == kFused ? ti0 - 1 : ti0;
    const int cj0 = MODE == kFused ? tj0 - 1 : tj0;
    for (Walk<CX, kThreads> w(tid); w.j < CY; w.next()) {
      const int i = ci0 + w.i, j = cj0 + w.j;
      if (!IN && (i < 1 || i > nx || j < 1 || j > ny || on_ring(L, i, j))) continue;
      float z2, dz2;
      mass_cell<NL, SPH, IN, false>(L, s, i, j, z2, dz2);
      s.z2[s.o(i, j)] = z2;
      if (NL) s.dz[s.o(i, j)] = dz2;
    }
    if (!IN)
      ring_cells<kThreads>(L, ci0, cj0, CX, CY, tid, [&](int i, int j) {
        s.z2[s.o(i, j)] = open_cell(L, s, i, j);
        if (NL) s.dz[s.o(i, j)] = 0.0f;
      });
    __syncthreads();
  }

  // ---- outputs of the tile
  int step = 0;
  if (MODE != kMassOnly && L.zmax) {
    step = L.ctl[L.ctl_in] + 1;
    if (L.advance && t == 0 && tid == 0) L.ctl[1 - L.ctl_in] = step;
  }
  const float rdt = recip_refined(L.dt);
#pragma unroll 1
  for (int idx = tid; idx < TX * TY; idx += kThreads) {
    const int i = ti0 + idx % TX, j = tj0 + idx / TX;
    if (!IN && (i > nx || j > ny)) continue;
    const size_t g = (size_t)(j - 1) * nx + (i - 1);
    if (MODE != kMomentOnly) {
      float zo;
      if (IN || L.is_root || (i >= 2 && j >= 2)) zo = s.Z2(i, j);
      else if (i == 1 && j == 1) zo = s.Z2(2, 2);  // UPDATE ghosts, all_grids.f90:323-324
      else if (i == 1) zo = s.Z2(2, j);
      else zo = s.Z2(i, 2);
      st_plane<ST>(L.z2 + g, zo);
      if (MODE == kMassOnly && NL && L.dzp) L.dzp[g] = s.DZ(i, j);
    }
    if (MODE != kMassOnly) {
      const bool r = L.is_root;
      float m2v, n2v;
      // Cells outside a routine's range are never written by the Fortran,
      // so its level 2 holds 0 there (ALLOC) and the level copy carries that
      // 0 into level 1; writing 0 keeps the swapped buffers identical to it.
      if (NL) {
        // CONMOME(_S) ranges (moment.f90:377-387, 583-593, 873-883, 1072-1082)
        const bool mx =
            IN || (r ? (i <= nx - 1 && j <= ny - 1) : (i >= 2 && i <= nx - 1 && j >= 2));
        const bool my =
            IN || (r ? (i <= nx - 1 && j <= ny - 1) : (i >= 2 && j >= 2 && j <= ny - 1));
        const float brk = L.linchk == 1 ? breaking_coef(L, s, i, j, rdt) : 1.0f;
        m2v = mx ? nl_x<SPH, IN>(L, s, i, j, brk) : 0.0f;
        n2v = my ? nl_y<SPH, IN>(L, s, i, j, brk) : 0.0f;
      } else {
        // MOMT_S / MOMT_C ranges (moment.f90:78-88, 117-119, 174-192, 233-235)
        const int lo = r ? 1 : 2;
        const bool mx = IN || (i >= lo && i <= nx - 1 && j >= lo);
        const bool my = IN || (j >= lo && j <= ny - 1 && i >= lo);
        m2v = mx ? lin_x<SPH, IN>(L, s, i, j) : 0.0f;
        n2v = my ? lin_y<SPH, IN>(L, s, i, j) : 0.0f;
      }
      if (L.sponge_cx) {  // SPONGE_LAYER (boundaries.f90:228-239)
        const float cx = L.sponge_cx[i - 1], cy = L.sponge_cy[j - 1];
        if (cx > 0.0f || cy > 0.0f) {
          m2v = m2v + cx * L.dt * s.M(i, j);
          n2v = n2v + cy * L.dt * s.N(i, j);
        }
      }
      st_plane<ST>(L.m2 + g, m2v);
      st_plane<ST>(L.n2 + g, n2v);
      if (L.zmax && (IN || r || (i >= 2 && j >= 2))) {
        const float z2 = s.Z2(i, j);
        if (s.H(i, j) + z2 > kGx) {
          const float zm = L.zmax[g];
          if (z2 >= L.arr_thr && !(zm >= L.arr_thr)) L.arrival[g] = step;
          if (z2 > zm) L.zmax[g] = z2;
        }
      }
      if (L.zraw) {  // outermost layer only: every cell is its own
        const float z2 = s.Z2(i, j);
        if (!(L.zraw[g] >= z2)) L.zraw[g] = z2;
      }
    }
  }
}

// kMomentMass for one tile of a linear outermost layer (see kernels.h).
// Stages Z(k), M(k-1), N(k-1), H with a two-cell halo; computes momentum
// (MOMT_S / MOMT_C, and SPONGE_LAYER) on the tile plus one cell to the
// left and below, into LDS; then, per tile cell, stores M(k), N(k), takes
// the wet maximum / arrival / MAX_AMP of Z(k) and computes continuity +
// open boundary of step k+1 from Z(k), M(k), N(k): the open boundary and
// the divergence read the fluxes of the cell and of its left and lower
// neighbours only.  Each value comes from the same functions and operands
// as in the momentum-only and continuity-only launches.
template <bool SPH, int TX, int TY, bool IN, bool ST, bool TAB>
__device__ __forceinline__ void tile_body_mm(const DevLayer &L, int t, int ti0, int tj0,
                                             float *sm, const gcdet::Tables *stab) {
  constexpr int P = TX + 4, R = TY + 4, A = P * R;
  float *pz = sm, *pm0 = sm + A, *pn0 = sm + 2 * A, *ph = sm + 3 * A;
  float *pm = sm + 4 * A, *pn = sm + 5 * A, *crow = sm + 6 * A;
  // momentum reads Z(k) as its level 2 and the level-1 fluxes; the next
  // continuity reads Z(k) as its level 1 and the new fluxes
  Lds<P> sa{pz, pm0, pn0, ph, pz, nullptr, ti0 - 2, tj0 - 2, stab, crow, R};
  Lds<P> sb{pz, pm, pn, ph, pz, nullptr, ti0 - 2, tj0 - 2, stab, crow, R};
  const int nx = L.nx, ny = L.ny, tid = threadIdx.x;
  {  // every load of the thread in flight before its first LDS write
    constexpr int KS = (A + kThreads - 1) / kThreads;
    constexpr int KC = SPH ? (kNumCoef * R + kThreads - 1) / kThreads : 0;
    float vh[KS], vz[KS], vm[KS], vn[KS], vc[KC > 0 ? KC : 1];
    TabStage<kThreads> tb;
    if constexpr (TAB) tb.load(tid);
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int idx = tid + k * kThreads, c = idx / R, j = tj0 - 2 + idx % R;
      vc[k] = idx < kNumCoef * R && (IN || (j >= 1 && j <= ny)) ? L.coef[c * ny + (j - 1)] : 0.0f;
    }
    Walk<P, kThreads> w(tid);
#pragma unroll
    for (int k = 0; k < KS; ++k, w.next()) {
      const int idx = tid + k * kThreads;
      const int gi = ti0 - 2 + w.i, gj = tj0 - 2 + w.j;
      const bool in = (KS * kThreads == A || idx < A) &&
                      (IN || (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny));
      const size_t g = in ? (size_t)(gj - 1) * nx + (gi - 1) : 0;
      vh[k] = in ? L.h[g] : 0.0f;
      vz[k] = in ? L.z1[g] : 0.0f;
      vm[k] = in ? L.m1[g] : 0.0f;
      vn[k] = in ? L.n1[g] : 0.0f;
    }
    if constexpr (TAB) tb.store(stab, tid);
    if (SPH) {
#pragma unroll
      for (int k = 0; k < KC; ++k)
        if (tid + k * kThreads < kNumCoef * R) crow[tid + k * kThreads] = vc[k];
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int idx = tid + k * kThreads;
      if (KS * kThreads == A || idx < A) {
        ph[idx] = vh[k];
        pz[idx] = vz[k];
        pm0[idx] = vm[k];
        pn0[idx] = vn[k];
      }
    }
  }
  __syncthreads();

  // momentum of step k on the tile + one cell left and below (MOMT ranges
  // of the outermost layer: M on I <= NX-1, N on J <= NY-1; 0 elsewhere)
  {
    constexpr int CX = TX + 1, CY = TY + 1;
    for (Walk<CX, kThreads> w(tid); w.j < CY; w.next()) {
      const int i = ti0 - 1 + w.i, j = tj0 - 1 + w.j;
      float m = 0.0f, n = 0.0f;
      if (IN || (i >= 1 && i <= nx && j >= 1 && j <= ny)) {
        if (IN || i <= nx - 1) m = lin_x<SPH, IN>(L, sa, i, j);
        if (IN || j <= ny - 1) n = lin_y<SPH, IN>(L, sa, i, j);
        if (L.sponge_cx) {  // SPONGE_LAYER (boundaries.f90:228-239)
          const float cx = L.sponge_cx[i - 1], cy = L.sponge_cy[j - 1];
          if (cx > 0.0f || cy > 0.0f) {
            m = m + cx * L.dt * sa.M(i, j);
            n = n + cy * L.dt * sa.N(i, j);
          }
        }
      }
      pm[sa.o(i, j)] = m;
      pn[sa.o(i, j)] = n;
    }
  }
  __syncthreads();

  int step = 0;
  if (L.zmax) {
    step = L.ctl[L.ctl_in] + 1;
    if (L.advance && t == 0 && tid == 0) L.ctl[1 - L.ctl_in] = step;
  }
#pragma unroll 1
  for (int idx = tid; idx < TX * TY; idx += kThreads) {
    const int i = ti0 + idx % TX, j = tj0 + idx / TX;
    if (!IN && (i > nx || j > ny)) continue;
    const size_t g = (size_t)(j - 1) * nx + (i - 1);
    st_plane<ST>(L.m2 + g, sb.M(i, j));
    st_plane<ST>(L.n2 + g, sb.N(i, j));
    const float zk = sa.Z2(i, j);
    if (L.zmax && sa.H(i, j) + zk > kGx) {
      const float zm = L.zmax[g];
      if (zk >= L.arr_thr && !(zm >= L.arr_thr)) L.arrival[g] = step;
      if (zk > zm) L.zmax[g] = zk;
    }
    if (L.zraw && !(L.zraw[g] >= zk)) L.zraw[g] = zk;
    if (!IN && on_ring(L, i, j)) continue;  // the ring pass below
    float z2, dz2;
    mass_cell<false, SPH, IN, false>(L, sb, i, j, z2, dz2);
    st_plane<ST>(L.z2 + g, z2);
  }
  if (!IN)
    ring_cells<kThreads>(L, ti0, tj0, TX, TY, tid, [&](int i, int j) {
      st_plane<ST>(L.z2 + (size_t)(j - 1) * nx + (i - 1), open_cell(L, sb, i, j));
    });
}

// TAB: the tile body also fills the workgroup's LDS copy of the pow/exp
// tables (its loads issued with the tile's staging loads).
template <int MODE, bool NL, bool SPH, int TX, int TY, bool ST = false, bool TAB = false>
__device__ __forceinline__ void tile_step(const DevLayer &L, int t, int tiles_x, float *sm,
                                          const gcdet::Tables *stab) {
  const int ti0 = (t % tiles_x) * TX + 1, tj0 = (t / tiles_x) * TY + 1;
  const bool interior = ti0 - 2 > kSpan + 2 && ti0 + TX + 2 < L.nx - kSpan &&
                        tj0 - 2 > kSpan + 2 && tj0 + TY + 2 < L.ny - kSpan;
  if constexpr (MODE == kMomentMass) {
    static_assert(!NL, "momentum + next continuity is built for linear layers");
    if (interior) tile_body_mm<SPH, TX, TY, true, ST, TAB>(L, t, ti0, tj0, sm, stab);
    else tile_body_mm<SPH, TX, TY, false, false, TAB>(L, t, ti0, tj0, sm, stab);
  } else {
    if (interior) tile_body<MODE, NL, SPH, TX, TY, true, ST, TAB>(L, t, ti0, tj0, sm, stab);
    else tile_body<MODE, NL, SPH, TX, TY, false, false, TAB>(L, t, ti0, tj0, sm, stab);
  }
}

// Waves per SIMD the register allocation of the step kernels aims for.
// The nonlinear tiles have long dependent chains (divisions, square roots,
// FP64 powers) that want many waves to hide latency: the spherical one needs
// 87 VGPRs unconstrained (5 waves) and is ~20% faster at 8 (64 VGPRs, a few
// values spilled); 7 is 3% and 6 is 11% slower than 8.  The Cartesian one is
// also fastest at 8.  Linear tiles aim at 7 (at 8 the 1024x768 linear layer
// and the 2048x1536 + 1801x1801 nest are 1% and 4% slower).  Register
// allocation never changes a result.
#ifndef GC_STEP_WAVES
#define GC_STEP_WAVES 7
#endif
#ifndef GC_STEP_WAVES_NL
#define GC_STEP_WAVES_NL 8
#endif
// Tiles whose LDS leaves room for fewer workgroups per CU (160 KiB each;
// one 256-thread workgroup is one wave per SIMD) aim at what fits.
constexpr int kLdsPerCu = 160 * 1024;
template <bool NL, bool SPH, int LDS_BYTES>
struct StepWaves {
  static constexpr int want = NL ? GC_STEP_WAVES_NL : GC_STEP_WAVES;
  static constexpr int fit = kLdsPerCu / LDS_BYTES;
  static constexpr int value = want < fit ? want : (fit > 0 ? fit : 1);
};
// LDS of the tile kernels (k_step, k_batch, k_batch_grid) and of k_stream
template <int TX, int TY>
constexpr int tile_lds_bytes() {
  return (int)((6 * (TX + 4) * (TY + 4) + kNumCoef * (TY + 4)) * sizeof(float) +
               sizeof(gcdet::Tables) + 2 * sizeof(int));
}

template <int MODE, bool NL, bool SPH, int TX, int TY, bool ST>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(StepWaves<NL, SPH, tile_lds_bytes<TX, TY>()>::value)))
    k_step(DevLayer L, int tiles_x, int ntiles, int per_xcd) {
  __shared__ float sm[6 * (TX + 4) * (TY + 4) + kNumCoef * (TY + 4)];
  __shared__ gcdet::Tables stab;
  const int t = tile_of_block(per_xcd);
  if (t >= ntiles) return;
  tile_step<MODE, NL, SPH, TX, TY, ST, NL || !SPH>(L, t, tiles_x, sm, &stab);
}

// ------------------------------------------------ two steps per launch
// Temporal blocking for a linear outermost layer without nested layers:
// one launch advances two root steps.  A tile stages Z, M, N, H with a
// four-cell halo, computes the first step's Z (continuity + open boundary)
// on the tile plus three cells and its M, N on the tile plus two cells,
// all in LDS, then the second step on the tile from those, and writes
// Z, M, N once.  Every cell value is produced by the same functions from
// the same operands as in two single-step launches, so the result is
// bitwise the same; HBM traffic and launches per step halve, at the cost
// of recomputing the halo ring (cheap for the linear equations).  The wet
// maximum and arrival see both steps in order.
template <bool SPH, int TX, int TY, bool IN, int NT, bool ST, bool TAB>
__device__ __forceinline__ void tile_body2(const DevLayer &L, int t, int ti0, int tj0,
                                           float *sm, const gcdet::Tables *stab) {
  constexpr int HW = 4, P = TX + 2 * HW, R = TY + 2 * HW, A = P * R;
  float *ph = sm, *pz0 = sm + A, *pm0 = sm + 2 * A, *pn0 = sm + 3 * A;
  float *pz1 = sm + 4 * A, *pm1 = sm + 5 * A, *pn1 = sm + 6 * A, *crow = sm + 7 * A;
  const int gi0 = ti0 - HW, gj0 = tj0 - HW;
  // first step: level 1 = staged inputs, new Z into pz1; second step:
  // level 1 = first step's planes, new Z into pz0 (dead after step one)
  Lds<P> s1{pz0, pm0, pn0, ph, pz1, nullptr, gi0, gj0, stab, crow, R};
  Lds<P> s2{pz1, pm1, pn1, ph, pz0, nullptr, gi0, gj0, stab, crow, R};
  const int nx = L.nx, ny = L.ny, tid = threadIdx.x;
  auto inside = [&](int i, int j) { return IN || (i >= 1 && i <= nx && j >= 1 && j <= ny); };

  // Staging: every load of the thread is issued before its first LDS
  // write, so the tile waits for memory once rather than once per pass of
  // the thread over the staged rectangle.
  {
    constexpr int KS = (A + NT - 1) / NT, KC = SPH ? (kNumCoef * R + NT - 1) / NT : 0;
    float vh[KS], vz[KS], vm[KS], vn[KS], vc[KC > 0 ? KC : 1];
    TabStage<NT> tb;
    if constexpr (TAB) tb.load(tid);
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int idx = tid + k * NT, c = idx / R, j = gj0 + idx % R;
      vc[k] = idx < kNumCoef * R && (IN || (j >= 1 && j <= ny)) ? L.coef[c * ny + (j - 1)] : 0.0f;
    }
    Walk<P, NT> w(tid);
#pragma unroll
    for (int k = 0; k < KS; ++k, w.next()) {
      const int idx = tid + k * NT;
      const int gi = gi0 + w.i, gj = gj0 + w.j;
      const bool in = idx < A && inside(gi, gj);
      const size_t g = in ? (size_t)(gj - 1) * nx + (gi - 1) : 0;
      vh[k] = in ? L.h[g] : 0.0f;
      vz[k] = in ? L.z1[g] : 0.0f;
      vm[k] = in ? L.m1[g] : 0.0f;
      vn[k] = in ? L.n1[g] : 0.0f;
    }
    if constexpr (TAB) tb.store(stab, tid);
#pragma unroll
    for (int k = 0; k < KC; ++k)
      if (tid + k * NT < kNumCoef * R) crow[tid + k * NT] = vc[k];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int idx = tid + k * NT;
      if (KS * NT == A || idx < A) {
        ph[idx] = vh[k];
        pz0[idx] = vz[k];
        pm0[idx] = vm[k];
        pn0[idx] = vn[k];
      }
    }
  }
  __syncthreads();

  // step one, continuity on the tile + 3
  {
    constexpr int CX = TX + 6, CY = TY + 6;
    for (Walk<CX, NT> w(tid); w.j < CY; w.next()) {
      const int i = ti0 - 3 + w.i, j = tj0 - 3 + w.j;
      if (!IN && inside(i, j) && on_ring(L, i, j)) continue;  // the ring pass below
      float z2 = 0.0f, dz2;
      if (inside(i, j)) mass_cell<false, SPH, IN, false>(L, s1, i, j, z2, dz2);
      pz1[s1.o(i, j)] = z2;
    }
    if (!IN)
      ring_cells<NT>(L, ti0 - 3, tj0 - 3, CX, CY, tid,
                     [&](int i, int j) { pz1[s1.o(i, j)] = open_cell(L, s1, i, j); });
  }
  __syncthreads();
  // step one, momentum on the tile + 2 (0 outside the MOMT ranges and the grid)
  {
    constexpr int CX = TX + 4, CY = TY + 4;
    for (Walk<CX, NT> w(tid); w.j < CY; w.next()) {
      const int i = ti0 - 2 + w.i, j = tj0 - 2 + w.j;
      float m = 0.0f, n = 0.0f;
      if (inside(i, j)) {
        const bool mx = IN || (i <= nx - 1);
        const bool my = IN || (j <= ny - 1);
        if (mx) m = lin_x<SPH, IN>(L, s1, i, j);
        if (my) n = lin_y<SPH, IN>(L, s1, i, j);
      }
      pm1[s1.o(i, j)] = m;
      pn1[s1.o(i, j)] = n;
    }
  }
  __syncthreads();
  // step two, continuity on the tile + 1
  {
    constexpr int CX = TX + 2, CY = TY + 2;
    for (Walk<CX, NT> w(tid); w.j < CY; w.next()) {
      const int i = ti0 - 1 + w.i, j = tj0 - 1 + w.j;
      if (!IN && inside(i, j) && on_ring(L, i, j)) continue;  // the ring pass below
      float z2 = 0.0f, dz2;
      if (inside(i, j)) mass_cell<false, SPH, IN, false>(L, s2, i, j, z2, dz2);
      pz0[s2.o(i, j)] = z2;
    }
    if (!IN)
      ring_cells<NT>(L, ti0 - 1, tj0 - 1, CX, CY, tid,
                     [&](int i, int j) { pz0[s2.o(i, j)] = open_cell(L, s2, i, j); });
  }
  __syncthreads();

  // step two, momentum and the outputs of the tile
  int step = 0;
  if (L.zmax) {
    step = L.ctl[L.ctl_in] + 1;
    if (L.advance && t == 0 && tid == 0) L.ctl[1 - L.ctl_in] = step + 1;
  }
#pragma unroll 1
  for (int idx = tid; idx < TX * TY; idx += NT) {
    const int i = ti0 + idx % TX, j = tj0 + idx / TX;
    if (!IN && (i > nx || j > ny)) continue;
    const size_t g = (size_t)(j - 1) * nx + (i - 1);
    const float za = pz1[s1.o(i, j)], zb = pz0[s2.o(i, j)];
    st_plane<ST>(L.z2 + g, zb);
    const bool mx = IN || (i <= nx - 1);
    const bool my = IN || (j <= ny - 1);
    st_plane<ST>(L.m2 + g, mx ? lin_x<SPH, IN>(L, s2, i, j) : 0.0f);
    st_plane<ST>(L.n2 + g, my ? lin_y<SPH, IN>(L, s2, i, j) : 0.0f);
    if (L.zmax) {
      const float h = ph[s1.o(i, j)];
      const bool wa = h + za > kGx, wb = h + zb > kGx;
      if (wa || wb) {  // the maximum is only read and written where wet
        const float z0 = L.zmax[g];
        float zm = z0;
        int arr = -2;
        if (wa) {
          if (za >= L.arr_thr && !(zm >= L.arr_thr)) arr = step;
          if (za > zm) zm = za;
        }
        if (wb) {
          if (zb >= L.arr_thr && !(zm >= L.arr_thr)) arr = step + 1;
          if (zb > zm) zm = zb;
        }
        if (arr != -2) L.arrival[g] = arr;
        if (zm != z0) L.zmax[g] = zm;
      }
    }
    if (L.zraw) {  // both steps in order, as two MAX_AMP launches would
      float zr = L.zraw[g];
      bool upd = false;
      if (!(zr >= za)) { zr = za; upd = true; }
      if (!(zr >= zb)) { zr = zb; upd = true; }
      if (upd) L.zraw[g] = zr;
    }
  }
}

// NT threads per tile: small grids leave the chip with few workgroups per
// CU, and more threads on the same LDS tile shorten each phase's chain of
// dependent LDS and memory accesses there; the arithmetic of every cell is
// the same whatever NT is.
template <bool SPH, int TX, int TY, int NT, bool ST>
__global__ void __launch_bounds__(NT)
    k_step2(DevLayer L, int tiles_x, int ntiles, int per_xcd) {
  constexpr int HW = 4;
  __shared__ float sm[7 * (TX + 2 * HW) * (TY + 2 * HW) + kNumCoef * (TY + 2 * HW)];
  __shared__ gcdet::Tables stab;
  const int t = tile_of_block(per_xcd);
  if (t >= ntiles) return;
  const int ti0 = (t % tiles_x) * TX + 1, tj0 = (t / tiles_x) * TY + 1;
  const bool interior = ti0 - HW > kSpan + 2 && ti0 + TX + HW < L.nx - kSpan &&
                        tj0 - HW > kSpan + 2 && tj0 + TY + HW < L.ny - kSpan;
  if (interior) tile_body2<SPH, TX, TY, true, NT, ST, !SPH>(L, t, ti0, tj0, sm, &stab);
  else tile_body2<SPH, TX, TY, false, NT, false, !SPH>(L, t, ti0, tj0, sm, &stab);
}

// ---------------------------------------------------------- strip kernel
// The fused step as a strip marching along J: a workgroup owns 252
// interior columns (256 lanes, two halo lanes each side) and a segment of
// rows.  Each iteration stores one prefetched input row into LDS row rings,
// computes continuity for the row two ahead, prefetches the next input row
// and computes the fluxes of the current row.  Halo overhead is ~1.6 % of
// columns and (SEG+4)/SEG of rows, against (TX+4)(TY+4)/(TX*TY) for a tile.
constexpr int kStreamW = 256;
constexpr int kStreamI = kStreamW - 4;

struct Ring {
  float *m1, *n1, *h, *z2, *dz;  // m1/n1/h: 6 rows, z2/dz: 5 rows of kStreamW
  int gi0;                       // global column of lane 0
  float z1own;                   // Z1 of the calling thread's own cell
  const gcdet::Tables *tab;      // pow/exp tables, LDS copy
  __device__ __forceinline__ int o6(int i, int j) const {
    return ((j + 12) % 6) * kStreamW + (i - gi0);
  }
  __device__ __forceinline__ int o5(int i, int j) const {
    return ((j + 10) % 5) * kStreamW + (i - gi0);
  }
  __device__ __forceinline__ float M(int i, int j) const { return m1[o6(i, j)]; }
  __device__ __forceinline__ float N(int i, int j) const { return n1[o6(i, j)]; }
  __device__ __forceinline__ float H(int i, int j) const { return h[o6(i, j)]; }
  __device__ __forceinline__ float Z2(int i, int j) const { return z2[o5(i, j)]; }
  __device__ __forceinline__ float DZ(int i, int j) const { return dz[o5(i, j)]; }
  // Z1 is only ever read at the thread's own cell (continuity, breaking)
  __device__ __forceinline__ float Z1(int, int) const { return z1own; }
  __device__ __forceinline__ float CF(const DevLayer &L, int k, int j) const {
    return L.coef[k * L.ny + (j - 1)];
  }
};

template <bool NL, bool SPH>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(
        StepWaves<NL, SPH, (int)((6 * 3 + 5 * 2) * kStreamW * sizeof(float) + sizeof(gcdet::Tables))>::value)))
    k_stream(DevLayer L, int nstrip, int nseg, int seg, int per_xcd) {
  __shared__ float sm[(6 * 3 + 5 * 2) * kStreamW];
  const int t = tile_of_block(per_xcd);
  if (t >= nstrip * nseg) return;
  const int strip = t / nseg, sg = t % nseg;
  const int nx = L.nx, ny = L.ny;
  const int c0 = strip * kStreamI + 1;
  const int j0 = sg * seg + 1;
  const int j1 = min(j0 + seg - 1, ny);
  const int lane = threadIdx.x;
  const int gi = c0 - 2 + lane;
  const bool col_in = gi >= 1 && gi <= nx;
  __shared__ gcdet::Tables stab;
  Ring r{sm, sm + 6 * kStreamW, sm + 12 * kStreamW, sm + 18 * kStreamW, sm + 23 * kStreamW,
         c0 - 2, 0.0f, &stab};
  if (NL || !SPH) stage_tables(stab, lane);

  auto load_row = [&](int j, float &h, float &m, float &n, float &z) {
    const bool in = col_in && j >= 1 && j <= ny;
    const size_t g = in ? (size_t)(j - 1) * nx + (gi - 1) : 0;
    h = in ? L.h[g] : 0.0f;
    m = in ? L.m1[g] : 0.0f;
    n = in ? L.n1[g] : 0.0f;
    z = in ? L.z1[g] : 0.0f;
  };
  auto store_row = [&](int j, float h, float m, float n) {
    const int o = r.o6(gi, j);
    r.h[o] = h;
    r.m1[o] = m;
    r.n1[o] = n;
  };
  // lane 0 has no left neighbour in LDS; its Z2 is never read
  auto mass_row = [&](int j, float z1) {
    float z2 = 0.0f, dz2 = 0.0f;
    if (col_in && lane >= 1 && j >= 1 && j <= ny) {
      r.z1own = z1;
      mass_cell<NL, SPH, false>(L, r, gi, j, z2, dz2);
    }
    const int o = r.o5(gi, j);
    r.z2[o] = z2;
    r.dz[o] = dz2;
  };

  float zp[4];
  {
    float h, m, n;
    for (int k = 0; k < 4; ++k) {
      load_row(j0 - 2 + k, h, m, n, zp[k]);
      store_row(j0 - 2 + k, h, m, n);
    }
  }
  __syncthreads();
  for (int k = 1; k < 4; ++k) mass_row(j0 - 2 + k, zp[k]);
  float zq0 = zp[2], zq1 = zp[3], zq2 = 0.0f;
  float ph, pm, pn, pz;
  load_row(j0 + 2, ph, pm, pn, pz);

  int step = 0;
  if (L.zmax) {
    step = L.ctl[L.ctl_in] + 1;
    if (L.advance && t == 0 && lane == 0) L.ctl[1 - L.ctl_in] = step;
  }
  const bool own = lane >= 2 && lane <= kStreamW - 3 && col_in;
  const float rdt = recip_refined(L.dt);
  for (int j = j0; j <= j1; ++j) {
    store_row(j + 2, ph, pm, pn);
    zq2 = pz;
    __syncthreads();
    mass_row(j + 2, zq2);
    __syncthreads();
    load_row(j + 3, ph, pm, pn, pz);
    if (own) {
      const int i = gi;
      const size_t g = (size_t)(j - 1) * nx + (i - 1);
      const bool rt = L.is_root;
      float zo;
      if (rt || (i >= 2 && j >= 2)) zo = r.Z2(i, j);
      else if (i == 1 && j == 1) zo = r.Z2(2, 2);  // UPDATE ghosts, all_grids.f90:323-324
      else if (i == 1) zo = r.Z2(2, j);
      else zo = r.Z2(i, 2);
      st_plane<false>(L.z2 + g, zo);
      r.z1own = zq0;
      float m2v, n2v;
      if (NL) {
        const bool mx = rt ? (i <= nx - 1 && j <= ny - 1) : (i >= 2 && i <= nx - 1 && j >= 2);
        const bool my = rt ? (i <= nx - 1 && j <= ny - 1) : (i >= 2 && j >= 2 && j <= ny - 1);
        const float brk = L.linchk == 1 ? breaking_coef(L, r, i, j, rdt) : 1.0f;
        m2v = mx ? nl_x<SPH, false>(L, r, i, j, brk) : 0.0f;
        n2v = my ? nl_y<SPH, false>(L, r, i, j, brk) : 0.0f;
      } else {
        const int lo = rt ? 1 : 2;
        const bool mx = i >= lo && i <= nx - 1 && j >= lo;
        const bool my = j >= lo && j <= ny - 1 && i >= lo;
        m2v = mx ? lin_x<SPH, false>(L, r, i, j) : 0.0f;
        n2v = my ? lin_y<SPH, false>(L, r, i, j) : 0.0f;
      }
      if (L.sponge_cx) {
        const float cx = L.sponge_cx[i - 1], cy = L.sponge_cy[j - 1];
        if (cx > 0.0f || cy > 0.0f) {
          m2v = m2v + cx * L.dt * r.M(i, j);
          n2v = n2v + cy * L.dt * r.N(i, j);
        }
      }
      st_plane<false>(L.m2 + g, m2v);
      st_plane<false>(L.n2 + g, n2v);
      if (L.zmax && (rt || (i >= 2 && j >= 2))) {
        const float z2 = r.Z2(i, j);
        if (r.H(i, j) + z2 > kGx) {
          const float zm = L.zmax[g];
          if (z2 >= L.arr_thr && !(zm >= L.arr_thr)) L.arrival[g] = step;
          if (z2 > zm) L.zmax[g] = z2;
        }
      }
      if (L.zraw) {
        const float z2 = r.Z2(i, j);
        if (!(L.zraw[g] >= z2)) L.zraw[g] = z2;
      }
    }
    zq0 = zq1;
    zq1 = zq2;
  }
}

template <bool NL, bool SPH>
hipError_t launch_stream(const DevLayer &L, int seg, hipStream_t st) {
  const int ns = (L.nx + kStreamI - 1) / kStreamI, ng = (L.ny + seg - 1) / seg;
  const int nt = ns * ng, per = (nt + 7) / 8;
  hipLaunchKernelGGL((k_stream<NL, SPH>), dim3(8 * per), dim3(kThreads), 0, st, L, ns, ng, seg,
                     per);
  return hipGetLastError();
}

// Reference maximum recorder (GPUOutput.cu:46-53 semantics: no wet test)
// plus the wet maximum / arrival of the extension, for maxamp_launch_.
__global__ void k_max(DevLayer L, float *zmax_raw) {
  const size_t n = (size_t)L.nx * L.ny;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int step = L.ctl[L.ctl_in] + 1;
  if (g == 0 && L.advance) L.ctl[1 - L.ctl_in] = step;
  if (g >= n) return;
  const int i = (int)(g % L.nx) + 1, j = (int)(g / L.nx) + 1;
  const float z2 = L.z2[g];
  if (!(zmax_raw[g] >= z2)) zmax_raw[g] = z2;  // a NaN replaces, as GPUOutput.cu:49-51
  if (L.zmax && (L.is_root || (i >= 2 && j >= 2)) && L.h[g] + z2 > kGx) {
    const float zm = L.zmax[g];
    if (z2 >= L.arr_thr && !(zm >= L.arr_thr)) L.arrival[g] = step;
    if (z2 > zm) L.zmax[g] = z2;
  }
}

// ---------------------------------------------------------------- nesting
__device__ __forceinline__ float pidx(const DevNest &N, const float *a, int i, int j) {
  return a[(size_t)(j - 1) * N.pnx + (i - 1)];
}

// Parent-side forecast of the interface flux at line index k (NEWQ at
// sub-step 2, all_grids.f90:706-720 and 778-790): one step of the
// parent's linear momentum, spherical (R2/R3 or R4/R5) or Cartesian
// (GRX*HM or GRY*HN).  side 0/1: lines I = IS-1 / IE (M); 2/3: J = JS-1 /
// JE (N).
__device__ float iface_forecast(const DevNest &N, int side, int k) {
  const bool vert = side < 2;
  float x;
  if (vert) {
    const int i = side == 0 ? N.is - 1 : N.ie, j = k, ip1 = i + 1, jm1 = j - 1;
    const float hpc = pidx(N, N.ph, i, j) > kGx
                          ? 0.5f * (pidx(N, N.ph, i, j) + pidx(N, N.ph, ip1 >= N.pnx ? N.pnx : ip1, j))
                          : 0.0f;
    if (N.p_sph) {
      float tot_n = pidx(N, N.pn1, i, j) + pidx(N, N.pn1, ip1, j) + pidx(N, N.pn1, i, jm1) +
                    pidx(N, N.pn1, ip1, jm1);
      const float f2 = N.pcoef[kF2 * N.pny + (j - 1)];
      float r2 = N.p_r24 ? pidx(N, N.pr24, i, j) : (N.p_nl ? f2 : f2 * fmaxf(hpc, 0.0f));
      x = pidx(N, N.pm1, i, j) - r2 * (pidx(N, N.pz2, ip1, j) - pidx(N, N.pz2, i, j)) +
          N.pcoef[kR3 * N.pny + (j - 1)] * tot_n;
    } else {
      float hm = hpc + 0.5f * (pidx(N, N.pz2, i, j) + pidx(N, N.pz2, ip1, j));
      x = pidx(N, N.pm1, i, j) - N.pgrx * hm * (pidx(N, N.pz2, ip1, j) - pidx(N, N.pz2, i, j));
    }
  } else {
    const int i = k, j = side == 2 ? N.js - 1 : N.je, jp1 = j + 1, im1 = i - 1;
    const float hqc = pidx(N, N.ph, i, j) > kGx
                          ? 0.5f * (pidx(N, N.ph, i, j) + pidx(N, N.ph, i, jp1 >= N.pny ? N.pny : jp1))
                          : 0.0f;
    if (N.p_sph) {
      float tot_m = pidx(N, N.pm1, im1, j) + pidx(N, N.pm1, im1, jp1) + pidx(N, N.pm1, i, j) +
                    pidx(N, N.pm1, i, jp1);
      const float f4 = N.pcoef[kF4 * N.pny + (j - 1)];
      float r4 = N.p_r24 ? N.pr24[(size_t)N.pnx * N.pny + (size_t)(j - 1) * N.pnx + (i - 1)]
                         : (N.p_nl ? f4 : f4 * fmaxf(hqc, 0.0f));
      x = pidx(N, N.pn1, i, j) - r4 * (pidx(N, N.pz2, i, jp1) - pidx(N, N.pz2, i, j)) -
          N.pcoef[kR5 * N.pny + (j - 1)] * tot_m;
    } else {
      float hn = hqc + 0.5f * (pidx(N, N.pz2, i, j) + pidx(N, N.pz2, i, jp1));
      x = pidx(N, N.pn1, i, j) - N.pgry * hn * (pidx(N, N.pz2, i, jp1) - pidx(N, N.pz2, i, j));
    }
  }
  if (fabsf(x) < kEps) x = 0.0f;
  return x;
}

// Both parent cells of the interface face at line index k wet in Z2?
__device__ __forceinline__ bool iface_wet(const DevNest &N, int side, int k) {
  const bool vert = side < 2;
  const int fi = vert ? (side == 0 ? N.is - 1 : N.ie) : k;
  const int fj = vert ? k : (side == 2 ? N.js - 1 : N.je);
  const int ai = vert ? fi + 1 : fi, aj = vert ? fj : fj + 1;
  return (pidx(N, N.ph, fi, fj) + pidx(N, N.pz2, fi, fj) > kGx) &&
         (pidx(N, N.ph, ai, aj) + pidx(N, N.pz2, ai, aj) > kGx);
}

// Parent flux on the interface line at index k for sub-step t: JNQ takes
// the parent's level-1 flux as is (all_grids.f90:484-487); NEWQ blends the
// cached forecast with it where both parent cells are wet and is 0
// elsewhere (all_grids.f90:703-727, 774-796).
__device__ float iface_flux(const DevNest &N, int side, int k, int t) {
  const bool vert = side < 2;
  const int fi = vert ? (side == 0 ? N.is - 1 : N.ie) : k;
  const int fj = vert ? k : (side == 2 ? N.js - 1 : N.je);
  const float f1 = vert ? pidx(N, N.pm1, fi, fj) : pidx(N, N.pn1, fi, fj);
  if (t == 1) return f1;
  if (!iface_wet(N, side, k)) return 0.0f;
  const float c1 = ((float)t - 1.0f) / (float)N.rel_time;
  const float c2 = 1.0f - c1;
  const float cached = t == 2 ? iface_forecast(N, side, k)
                              : (vert ? N.yflux[(side & 1) * N.pny + (k - 1)]
                                      : N.xflux[(side & 1) * N.pnx + (k - 1)]);
  return c1 * cached + c2 * f1;
}

// JNQ / NEWQ for one child: threads over the four interface lines.  A
// thread owns parent index k of one side: it writes the forecast cache at
// k (t == 2, k in the line's full range JS-2..JE+2 / IS-2..IE+2) and, for
// k in the footprint, the R child boundary fluxes of EDGEINTERP_VERT/HORI
// (all_grids.f90:502-604), from the interface flux at k-1, k, k+1.
__device__ __forceinline__ int jnq_threads(const DevNest &N) {
  return 2 * (N.je - N.js + 5) + 2 * (N.ie - N.is + 5);
}

__device__ void jnq_thread(const DevNest &N, int g, int t) {
  const int lv = N.je - N.js + 5, lh = N.ie - N.is + 5;  // k ranges incl. +-2
  int side;
  if (g < 2 * lv) { side = g / lv; g -= side * lv; }
  else { g -= 2 * lv; if (g >= 2 * lh) return; side = 2 + g / lh; g -= (side - 2) * lh; }
  const bool vert = side < 2;
  const int k = (vert ? N.js : N.is) - 2 + g;
  if (t == 2 && iface_wet(N, side, k)) {
    const float x = iface_forecast(N, side, k);
    if (vert) N.yflux[(side & 1) * N.pny + (k - 1)] = x;
    else N.xflux[(side & 1) * N.pnx + (k - 1)] = x;
  }
  const int lo = vert ? N.js : N.is, hi = vert ? N.je : N.ie;
  if (k < lo || k > hi) return;
  const float fm = iface_flux(N, side, k - 1, t);
  const float f0 = iface_flux(N, side, k, t);
  const float fp = iface_flux(N, side, k + 1, t);
  const float fs = 0.5f * (fm + f0);
  const float fe = 0.5f * (fp + f0);
  const int nn = 2 * N.rel_size;
  const float c1 = (fe - fs) / (float)nn;  // == REAL((fe-fs)/DBLE(nn)), all_grids.f90:538
  const float c2 = -c1 + fs;
  const int base = (k - lo) * N.rel_size + 1;
  for (int kk = 1; kk <= N.rel_size; ++kk) {
    const int c = base + kk;
    int ci, cj;
    if (vert) { ci = side == 0 ? 1 : N.cnx; cj = c; }
    else { ci = c; cj = side == 2 ? 1 : N.cny; }
    const size_t cg = (size_t)(cj - 1) * N.cnx + (ci - 1);
    float v = 2.0f * (float)kk * c1 + c2;
    if (N.ch[cg] + N.cz1[cg] <= kGx) v = 0.0f;
    if (vert) N.cm1[cg] = v;
    else N.cn1[cg] = v;
  }
}

__global__ void k_jnq(DevNest N, int t) {
  jnq_thread(N, blockIdx.x * blockDim.x + threadIdx.x, t);
}

// JNZ (all_grids.f90:342-421): parent Z2 over the footprint = mean of the
// wet child cells, summed in the Fortran order (KI outer, KJ inner).
// For R = 2..5 every load of the R x R child block is issued first (for
// one KI, KJ a wave reads one child row, R columns apart, and the R values
// of KI fill it in), then the sums run in that order; other R take the
// loop below.
template <int R>
__device__ __forceinline__ void jnz_block(const DevNest &N, int i, int j) {
  // 0-based first child cell: II = (I-IS) R + 1 + KI, KI = 1..R (1-based)
  const int i0 = (i - N.is) * R + 1, j0 = (j - N.js) * R + 1;
  const bool even = (N.rel_time % 2) == 0;
  float z2[R][R], hh[R][R], z1[R][R];
#pragma unroll
  for (int kj = 0; kj < R; ++kj)
#pragma unroll
    for (int ki = 0; ki < R; ++ki) {
      const size_t cg = (size_t)(j0 + kj) * N.cnx + (i0 + ki);
      z2[ki][kj] = N.cz2[cg];
      hh[ki][kj] = N.ch[cg];
      z1[ki][kj] = even ? N.cz1[cg] : 0.0f;
    }
  const float half = (float)(R * R) / 2.0f;
  float sum = 0.0f, count = 0.0f;
#pragma unroll
  for (int ki = 0; ki < R; ++ki)
#pragma unroll
    for (int kj = 0; kj < R; ++kj)
      if (hh[ki][kj] + z2[ki][kj] > kGx) {
        if (even) sum = sum + 0.5f * (z1[ki][kj] + z2[ki][kj]);
        else sum = sum + z2[ki][kj];
        count = count + 1.0f;
      }
  N.pz2_out[(size_t)(j - 1) * N.pnx + (i - 1)] = count > half ? sum / count : 0.0f;
}

__device__ void jnz_thread(const DevNest &N, int g) {
  const int w = N.ie - N.is + 1, hgt = N.je - N.js + 1;
  if (g >= w * hgt) return;
  const int i = N.is + g % w, j = N.js + g / w;
  const int ir = N.rel_size;
  switch (ir) {
    case 2: jnz_block<2>(N, i, j); return;
    case 3: jnz_block<3>(N, i, j); return;
    case 4: jnz_block<4>(N, i, j); return;
    case 5: jnz_block<5>(N, i, j); return;
    default: break;
  }
  const float half = (float)(ir * ir) / 2.0f;
  const int i0 = (i - N.is) * ir + 1, j0 = (j - N.js) * ir + 1;
  const bool even = (N.rel_time % 2) == 0;
  float sum = 0.0f, count = 0.0f;
  for (int ki = 1; ki <= ir; ++ki) {
    for (int kj = 1; kj <= ir; ++kj) {
      const size_t cg = (size_t)(j0 + kj - 1) * N.cnx + (i0 + ki - 1);
      const float z2 = N.cz2[cg];
      if (N.ch[cg] + z2 > kGx) {
        if (even) sum = sum + 0.5f * (N.cz1[cg] + z2);
        else sum = sum + z2;
        count = count + 1.0f;
      }
    }
  }
  N.pz2_out[(size_t)(j - 1) * N.pnx + (i - 1)] = count > half ? sum / count : 0.0f;
}

__global__ void k_jnz(DevNest N) { jnz_thread(N, blockIdx.x * blockDim.x + threadIdx.x); }

// ------------------------------------------------------------ batch forms
// Several nests in one launch: nest n owns blocks [start[n], start[n+1]).
__device__ __forceinline__ int nest_of_block(const int *start, int nnest) {
  int lo = 0, hi = nnest - 1;
  const int b = blockIdx.x;
  while (lo < hi) {  // last n with start[n] <= b
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_jnq_multi(const DevNest *__restrict__ nests, const int *__restrict__ start,
                            int nnest, int t) {
  const int n = __builtin_amdgcn_readfirstlane(nest_of_block(start, nnest));
  const DevNest N = nests[n];
  const int g = (blockIdx.x - start[n]) * blockDim.x + threadIdx.x;
  if (g < jnq_threads(N)) jnq_thread(N, g, t);
}

__global__ void k_jnz_multi(const DevNest *__restrict__ nests, const int *__restrict__ start,
                            int nnest) {
  const int n = __builtin_amdgcn_readfirstlane(nest_of_block(start, nnest));
  const DevNest N = nests[n];
  jnz_thread(N, (blockIdx.x - start[n]) * blockDim.x + threadIdx.x);
}

// Streaming copy, the reference for HBM bandwidth and for scaling the
// memory counters of rocprofv3: each workgroup copies one contiguous chunk
// of kCopyUnroll x 256 16-byte vectors, all its loads issued before its
// stores, non-temporal (streamed past the caches), and the hardware
// dispatcher hands chunks to CUs as they free up.  Of the shapes measured
// on MI355X (grid-stride loops with 1-8 loads in flight per lane and
// 1024-8192 workgroups, chunks of 1-8 vectors per lane) this one moves the
// most bytes per second.
constexpr int kCopyUnroll = 4;
typedef float CopyVec __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const CopyVec *__restrict__ src,
                                              CopyVec *__restrict__ dst, size_t n) {
  const size_t base = (size_t)blockIdx.x * (256 * kCopyUnroll) + threadIdx.x;
  CopyVec v[kCopyUnroll];
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) v[u] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) __builtin_nontemporal_store(v[u], dst + i);
  }
}

// Self-check of the short divisions and square root against the IEEE
// operations on the device.  Lane g owns a slice of the 2^32 floats for
// div12 (mismatches counted where |x| >= 2^-124 or x = 0) and sqrt_inrange
// (x = +-0 or x >= 2^-96), and `per_lane` pseudo-random pairs (a, b) for
// div_inrange, drawn with exponents spread over the range in which it
// claims equality (pairs outside it are skipped).  Lane g writes
// out[6g..6g+5]: the three counts and the bits of its first mismatching x,
// pair (a << 32 | b) and square-root argument; the host adds them up.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float f_of(unsigned u) { return __uint_as_float(u); }

__global__ void __launch_bounds__(256) k_division_check(unsigned long long seed, int per_lane,
                                                        unsigned long long *out) {
  const unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long lanes = (unsigned long long)gridDim.x * blockDim.x;
  unsigned long long bad12 = 0, bad = 0, ex12 = 0, ex = 0;
  const unsigned long long span = ((1ull << 32) + lanes - 1) / lanes;
  unsigned long long badsq = 0, exsq = 0;
  for (unsigned long long k = g * span; k < (g + 1) * span && k < (1ull << 32); ++k) {
    const float x = f_of((unsigned)k);
    if (x >= 0x1p-96f || x == 0.0f) {  // +-0, [2^-96, +inf]
      if (__float_as_uint(sqrtf(x)) != __float_as_uint(sqrt_inrange(x))) {
        if (!badsq) exsq = k;
        ++badsq;
      }
    }
    if (!(fabsf(x) <= 3.4028235e38f)) continue;  // NaN, infinities
    if (!(fabsf(x) >= 0x1p-124f) && x != 0.0f) continue;
    const float want = x / 12.0f, got = div12(x);
    if (__float_as_uint(want) != __float_as_uint(got)) {
      if (!bad12) ex12 = k;
      ++bad12;
    }
  }
  for (int k = 0; k < per_lane; ++k) {
    const unsigned long long r = mix64(seed + g * 0x9e3779b97f4a7c15ull + (unsigned long long)k);
    // exponents: b in [-kDivRange, kDivRange], a within 100 of it
    const int eb = (int)((r >> 32) % (2u * kDivRange + 1u)) - kDivRange;
    const int ea = eb + (int)(((r >> 40) & 0xffu) % 201u) - 100;
    if (ea < -126 || ea > 127) continue;
    const unsigned ma = (unsigned)(r & 0x7fffffu), mb = (unsigned)((r >> 8) & 0x7fffffu);
    const unsigned sa = (unsigned)((r >> 48) & 1u) << 31, sb = (unsigned)((r >> 49) & 1u) << 31;
    const float a = f_of(sa | ((unsigned)(ea + 127) << 23) | ma);
    const float b = f_of(sb | ((unsigned)(eb + 127) << 23) | mb);
    const float want = a / b;
    if (!div_inrange_applies(a, b, want)) continue;
    if (__float_as_uint(want) != __float_as_uint(div_inrange(a, b))) {
      if (!bad) ex = ((unsigned long long)__float_as_uint(a) << 32) | __float_as_uint(b);
      ++bad;
    }
  }
  out[6 * g] = bad12;
  out[6 * g + 1] = bad;
  out[6 * g + 2] = badsq;
  out[6 * g + 3] = ex12;
  out[6 * g + 4] = ex;
  out[6 * g + 5] = exsq;
}

// Dry land shows 0, else the elevation plus the tide correction when it is
// not negligible: DATA_PRT (output.f90:781-793) and MAX_AMP (:918-931).
__device__ __forceinline__ float shown_elevation(float z, float h, float tide) {
  if (z + h <= kGx && h < 0.0f) return 0.0f;
  return fabsf(tide) > kGx ? z + tide : z;
}

__global__ void __launch_bounds__(256) k_out_pack(OutPack p) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < (size_t)p.count;
       g += stride) {
    const float h = p.h[g];
    if (p.oz) p.oz[g] = shown_elevation(p.z[g], h, p.tide);
    if (p.om) p.om[g] = p.m[g];
    if (p.on) p.on[g] = p.n[g];
    if (p.ozmax) p.ozmax[g] = shown_elevation(p.zraw[g], h, p.tide);
  }
}

// ------------------------------------------------- SC_OPTION = 1 coupling
__device__ __forceinline__ float scp(const DevSc &S, const float *a, int i, int j) {
  return a[(size_t)(j - 1) * S.pnx + (i - 1)];
}
__device__ __forceinline__ bool in_rect(const int *r, int i, int j) {
  return i >= r[0] && i <= r[1] && j >= r[2] && j <= r[3];
}

// NEWQ_SC's forecast of the parent's linear flux at (i, j) (all_grids.f90:
// 1131-1271): MOMT_S with the SCM term where MODSCM = 0; x = M, else N
__device__ float sc_band_flux(const DevSc &S, bool x, int i, int j) {
  if (x) {
    const int ip1 = i + 1;
    if (!(scp(S, S.ph, i, j) > kGx && scp(S, S.ph, ip1, j) > kGx)) return 0.0f;
    int jm1 = j - 1, jp1 = j + 1;
    if (jm1 <= 1) jm1 = 1;
    if (jp1 >= S.pny) jp1 = S.pny;
    const float tot_n = scp(S, S.pn1, i, j) + scp(S, S.pn1, ip1, j) + scp(S, S.pn1, i, jm1) +
                        scp(S, S.pn1, ip1, jm1);
    const float hp = 0.5f * (scp(S, S.ph, i, j) + scp(S, S.ph, ip1 >= S.pnx ? S.pnx : ip1, j));
    const float f2 = S.pcoef[kF2 * S.pny + (j - 1)];
    const float r2 = S.p_r24 ? scp(S, S.pr24, i, j) : (S.p_nl ? f2 : f2 * fmaxf(hp, 0.0f));
    float xm = scp(S, S.pm1, i, j) - r2 * (scp(S, S.pz2, ip1, j) - scp(S, S.pz2, i, j)) +
               S.pcoef[kR3 * S.pny + (j - 1)] * tot_n;
    if (S.p_modscm == 0) {
      const float scm = r2 * kTwlvth *
                        ((scp(S, S.pz2, ip1, jp1) - 2.0f * scp(S, S.pz2, ip1, j) + scp(S, S.pz2, ip1, jm1)) -
                         (scp(S, S.pz2, i, jp1) - 2.0f * scp(S, S.pz2, i, j) + scp(S, S.pz2, i, jm1)));
      xm = xm - scm;
    }
    if (fabsf(xm) < kEps) xm = 0.0f;
    return xm;
  }
  const int jp1 = j + 1;
  if (!(scp(S, S.ph, i, j) > kGx && scp(S, S.ph, i, jp1) > kGx)) return 0.0f;
  int im1 = i - 1, ip1 = i + 1;
  if (im1 <= 1) im1 = 1;
  if (ip1 >= S.pnx) ip1 = S.pnx;
  const float tot_m = scp(S, S.pm1, im1, j) + scp(S, S.pm1, im1, jp1) + scp(S, S.pm1, i, j) +
                      scp(S, S.pm1, i, jp1);
  const float hq = 0.5f * (scp(S, S.ph, i, j) + scp(S, S.ph, i, jp1 >= S.pny ? S.pny : jp1));
  const float f4 = S.pcoef[kF4 * S.pny + (j - 1)];
  const float r4 = S.p_r24 ? S.pr24[(size_t)S.pnx * S.pny + (size_t)(j - 1) * S.pnx + (i - 1)]
                           : (S.p_nl ? f4 : f4 * fmaxf(hq, 0.0f));
  float xn = scp(S, S.pn1, i, j) - r4 * (scp(S, S.pz2, i, jp1) - scp(S, S.pz2, i, j)) -
             S.pcoef[kR5 * S.pny + (j - 1)] * tot_m;
  if (S.p_modscm == 0) {
    const float scm = r4 * kTwlvth *
                      ((scp(S, S.pz2, ip1, jp1) - 2.0f * scp(S, S.pz2, i, jp1) + scp(S, S.pz2, im1, jp1)) -
                       (scp(S, S.pz2, ip1, j) - 2.0f * scp(S, S.pz2, i, j) + scp(S, S.pz2, im1, j)));
    xn = xn - scm;
  }
  if (fabsf(xn) < kEps) xn = 0.0f;
  return xn;
}

// one thread per cell of the four bands (M bands first, then N bands)
__global__ void __launch_bounds__(256) k_sc_band(DevSc S) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  for (int b = 0; b < 4; ++b) {
    const int *r = S.band[b];
    const int w = r[1] - r[0] + 1, n = w * (r[3] - r[2] + 1);
    if (g < n) {
      const int i = r[0] + g % w, j = r[2] + g / w;
      const size_t q = (size_t)(j - 1) * S.pnx + (i - 1);
      if (b < 2) S.pm2[q] = sc_band_flux(S, true, i, j);
      else S.pn2[q] = sc_band_flux(S, false, i, j);
      return;
    }
    g -= n;
  }
}

// EDGE_INTERP_SC (t = 1, all_grids.f90:996-1091) and the edge part of
// NEWQ_SC (t > 1, :1274-1372): one thread per child boundary cell, left
// (I = 1), right (I = NX), bottom (J = 1), top (J = NY).  The parent's
// level 2 is the band forecast inside a band and level 1 elsewhere (where
// the Fortran's level 2 holds the copy CHANGE made).
__global__ void __launch_bounds__(256) k_sc_edges(DevSc S, int t) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int ny = S.cny, nx = S.cnx;
  if (g >= 2 * ny + 2 * nx) return;
  int side, ci, cj;
  if (g < ny) { side = 0; ci = 1; cj = g + 1; }
  else if (g < 2 * ny) { side = 1; ci = nx; cj = g - ny + 1; }
  else if (g < 2 * ny + nx) { side = 2; ci = g - 2 * ny + 1; cj = 1; }
  else { side = 3; ci = g - 2 * ny - nx + 1; cj = ny; }
  const bool x = side < 2;
  const size_t c = (size_t)(cj - 1) * nx + (ci - 1), cn = (size_t)nx * ny;
  float *out = x ? S.cm1 : S.cn1;
  // the child's level-2 elevation: 0 before its first sub-step ever (ALLOC)
  const float z2 = (t == 1 && *S.steps_done == 0) ? 0.0f : S.cz_now[c];
  if (!(S.ch[c] + z2 > kGx)) {
    out[c] = 0.0f;
    return;
  }
  const int ki = S.pos[c], kj = S.pos[cn + c];
  const float w1 = S.cxy[c], w2 = S.cxy[cn + c], w3 = S.cxy[2 * cn + c], w4 = S.cxy[3 * cn + c];
  const float *p1 = x ? S.pm1 : S.pn1;
  const float f1 = scp(S, p1, ki, kj) * w1 + scp(S, p1, ki + 1, kj) * w2 +
                   scp(S, p1, ki, kj + 1) * w3 + scp(S, p1, ki + 1, kj + 1) * w4;
  if (t == 1) {
    out[c] = f1;
    return;
  }
  const float *p2 = x ? S.pm2 : S.pn2;
  auto lev2 = [&](int i, int j) {
    const bool band = x ? (in_rect(S.band[0], i, j) || in_rect(S.band[1], i, j))
                        : (in_rect(S.band[2], i, j) || in_rect(S.band[3], i, j));
    return band ? scp(S, p2, i, j) : scp(S, p1, i, j);
  };
  const float f2 = lev2(ki, kj) * w1 + lev2(ki + 1, kj) * w2 + lev2(ki, kj + 1) * w3 +
                   lev2(ki + 1, kj + 1) * w4;
  const float c1 = ((float)t - 1.0f) / (float)S.rel_time;
  const float c2 = 1.0f - c1;
  // the top edge takes C1 for both terms (all_grids.f90:1365)
  out[c] = side == 3 ? c1 * f2 + c1 * f1 : c1 * f2 + c2 * f1;
}

// JNZ_SC (all_grids.f90:1376-1425): one thread per parent cell of the
// extent sums its wet child cells in the Fortran's order.  It runs before
// UPDATE, so the child's ghost column I = 1 and row J = 1, which continuity
// never computes (mass.f90:70-71), still hold the previous UPDATE's copies
// (all_grids.f90:323-324): the level-1 values here, whose level-2 slots the
// step has already refreshed; 0 before the first UPDATE ever (ALLOC).
__global__ void __launch_bounds__(256) k_sc_jnz(DevSc S) {
  const int w = S.corners[1] - S.corners[0] + 1;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= w * (S.corners[3] - S.corners[2] + 1)) return;
  const int i = S.corners[0] + g % w, j = S.corners[2] + g / w;
  // JNZ runs in sub-step REL_TIME/2 + 1: the first ever only when REL_TIME = 1
  const bool first = S.rel_time == 1 && *S.steps_done == 0;
  float z = 0.0f, cnt = 0.0f;
  for (int q = S.jnz_off[g]; q < S.jnz_off[g + 1]; ++q) {
    const int c = S.jnz_idx[q];
    const bool ghost = c < S.cnx || c % S.cnx == 0;
    const float zc = ghost ? (first ? 0.0f : S.cz_now[c]) : S.cz2[c];
    if (S.ch[c] + zc > kGx) {
      z = z + zc;
      cnt = cnt + 1.0f;
    }
  }
  if (cnt > S.half) S.pz2_out[(size_t)(j - 1) * S.pnx + (i - 1)] = z / cnt;
}

__global__ void __launch_bounds__(256) k_sc_cell_interp(const float *__restrict__ pd, int pnx,
                                                        int pny, float *cd, int cnx, int cny,
                                                        const int *__restrict__ pos,
                                                        const float *__restrict__ cxy) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)cnx * cny;
  if (g >= n) return;
  const int ki = pos[g], kj = pos[n + g];
  float v = 0.0f;
  if (ki >= 1 && ki < pnx && kj >= 1 && kj < pny) {
    auto P = [&](int a, int b) { return pd[(size_t)(b - 1) * pnx + (a - 1)]; };
    v = P(ki, kj) * cxy[g] + P(ki + 1, kj) * cxy[n + g] + P(ki, kj + 1) * cxy[2 * n + g] +
        P(ki + 1, kj + 1) * cxy[3 * n + g];
  }
  cd[g] = v;
}

// ---------------------------------------------------------------- source
// STRIKE_SLIP (deform.f90:194-217) and DIP_SLIP (:221-238) at one corner of
// the fault plane, FP64; sn, cs = sin/cos of the dip, dd = the depth.
__device__ __forceinline__ double okada_strike(double x2, double y1, double y2, double sn,
                                               double cs, double dd) {
  const double q = x2 * sn - dd * cs;
  const double d_bar = y2 * sn - q * cs;
  const double r = sqrt(y1 * y1 + y2 * y2 + q * q);
  double t1 = r + d_bar, t2 = r + y2;
  if (t1 <= (double)kEps) t1 = (double)kEps;
  if (t2 <= (double)kEps) t2 = (double)kEps;
  const double a4 = 0.5 * 1 / cs * (log(t1) - sn * log(t2));
  return -(d_bar * q / r / (r + y2) + q * sn / (r + y2) + a4 * sn) / 2 / (double)3.14159f;
}

__device__ __forceinline__ double okada_dip(double x2, double y1, double y2, double sn,
                                            double cs, double dd) {
  const double q = x2 * sn - dd * cs;
  const double d_bar = y2 * sn - q * cs;
  const double r = sqrt(y1 * y1 + y2 * y2 + q * q);
  const double xx = sqrt(y1 * y1 + q * q);
  const double a5 = 0.5 * 2 / cs * atan((y2 * (xx + q * cs) + xx * (r + xx) * sn) / y1 / (r + xx) / cs);
  return -(d_bar * q / r / (r + y1) + sn * atan(y1 * y2 / q / r) - a5 * sn * cs) / 2 /
         (double)3.14159f;
}

// DEFORM_OKADA's checks on each part (deform.f90:178-183), REAL limits
__device__ __forceinline__ double okada_clean(double u) {
  if (fabs(u) >= 3.4028234663852886e38) return 0.0;  // HUGE(REAL)
  if (u != u) return 0.0;
  if (fabs(u) <= (double)kGx) return 0.0;
  return u;
}

__global__ void __launch_bounds__(256) k_okada(FaultDev f, float *deform) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)f.nx * f.ny) return;
  const int i = (int)(g % f.nx), j = (int)(g / f.nx);  // 0-based I-1, J-1
  double xs, ys;
  if (f.sph) {
    const double sxi = f.row[2 * j], cxi = f.row[2 * j + 1];
    const double sdl = f.col[2 * i], cdl = f.col[2 * i + 1];
    const double beta = 1.0 + sxi * f.sxi0 + cxi * f.cxi0 * cdl;
    const double y = f.two_rk0 * (sxi * f.cxi0 - cxi * f.sxi0 * cdl) / beta;
    const double x = f.two_rk0 * cxi * sdl / beta;
    xs = x - f.del_x;
    ys = y + f.del_y;
  } else {
    const double yy = f.dy * (double)j, xx = f.dx * (double)i;
    xs = xx - f.xl - f.del_x;
    ys = yy - f.yl + f.del_y;
  }
  const double x1 = xs * f.sin_t + ys * f.cos_t;
  const double x2 = -(xs * f.cos_t - ys * f.sin_t);
  const double pp = x2 * f.cs + f.hh * f.sn;
  const double a = x1 + f.halfl, b = x1 - f.halfl, pw = pp - f.w;
  const double f1 = okada_strike(x2, a, pp, f.sn, f.cs, f.hh);
  const double f2 = okada_strike(x2, a, pw, f.sn, f.cs, f.hh);
  const double f3 = okada_strike(x2, b, pp, f.sn, f.cs, f.hh);
  const double f4 = okada_strike(x2, b, pw, f.sn, f.cs, f.hh);
  const double g1 = okada_dip(x2, a, pp, f.sn, f.cs, f.hh);
  const double g2 = okada_dip(x2, a, pw, f.sn, f.cs, f.hh);
  const double g3 = okada_dip(x2, b, pp, f.sn, f.cs, f.hh);
  const double g4 = okada_dip(x2, b, pw, f.sn, f.cs, f.hh);
  const double us = okada_clean((f1 - f2 - f3 + f4) * f.ds);
  const double ud = okada_clean((g1 - g2 - g3 + g4) * f.dd);
  deform[g] = (float)(us + ud);
}

// CELL_INTERP for one child cell (ci, cj), 1-based: the parent cell it lies
// in, the four corner means and the two linear interpolations, in the
// Fortran's REAL operation order; the ghost row/column stays 0.
__global__ void __launch_bounds__(256) k_cell_interp(const float *__restrict__ pd, int pnx,
                                                     float *cd, int cnx, int cny, int is, int ie,
                                                     int js, int je, int ir) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)cnx * cny) return;
  const int ci = (int)(g % cnx) + 1, cj = (int)(g / cnx) + 1;
  if (ci < 2 || cj < 2 || ci > (ie - is + 1) * ir + 1 || cj > (je - js + 1) * ir + 1) {
    cd[g] = 0.0f;
    return;
  }
  const int I = is + (ci - 2) / ir, K = (ci - 2) % ir + 1;
  const int J = js + (cj - 2) / ir, M = (cj - 2) % ir + 1;
  auto P = [&](int a, int b) { return pd[(size_t)(b - 1) * pnx + (a - 1)]; };
  const float tl = (P(I, J) + P(I, J + 1) + P(I - 1, J) + P(I - 1, J + 1)) / 4.0f;
  const float tr = (P(I, J) + P(I, J + 1) + P(I + 1, J) + P(I + 1, J + 1)) / 4.0f;
  const float bl = (P(I, J) + P(I, J - 1) + P(I - 1, J) + P(I - 1, J - 1)) / 4.0f;
  const float br = (P(I, J) + P(I, J - 1) + P(I + 1, J) + P(I + 1, J - 1)) / 4.0f;
  const float nxf = (float)(2 * ir);
  const float km = (float)(2 * M) - 1.0f;  // CELL(1, 2M) and CELL(NX+1, 2M)
  const float cl = km * (tl - bl) / nxf + bl;
  const float cr = km * (tr - br) / nxf + br;
  const float ki = (float)(2 * K) - 1.0f;  // CELL(2K, 2M)
  cd[g] = ki * (cr - cl) / nxf + cl;
}

__global__ void __launch_bounds__(256) k_apply_deform(float *z1, float *h,
                                                      const float *__restrict__ d, float *zraw,
                                                      long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float dv = d[g];
  h[g] = h[g] - dv;
  const float z = z1[g] + dv;
  z1[g] = z;
  if (zraw) zraw[g] = z;
}

// Gauge samples (GET_TS, output.f90:1085-1094): the 2x2 stencil of the
// current elevation weighted in the Fortran's operation order.
__global__ void k_gauges(const GaugeDev *__restrict__ gs, int n, ZNow zs,
                         const int *__restrict__ step_now, float *ring, int cap) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const GaugeDev G = gs[k];
  const float *z = zs.z[G.lid];
  const size_t o = (size_t)(G.kj - 1) * G.nx + (G.ki - 1);
  const float l1 = z[o], l2 = z[o + 1], l3 = z[o + G.nx], l4 = z[o + G.nx + 1];
  const float z1 = l1 * (1.0f - G.cx) * (1.0f - G.cy);
  const float z2 = l2 * G.cx * (1.0f - G.cy);
  const float z3 = l3 * (1.0f - G.cx) * G.cy;
  const float z4 = l4 * G.cx * G.cy;
  float dat = z1 + z2 + z3 + z4;
  if (fabsf(G.tide) > kGx) dat = dat + G.tide;
  const int step = *step_now;
  ring[(size_t)k * cap + (step - 1) % cap] = dat;
}

__global__ void k_fill_i32(int *p, int n, int v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) p[g] = v;
}

// Persistent step kernel of the batch engine: a grid no larger than the
// chip holds at once pulls (layer, tile) work items.  Items are split into
// 8 contiguous queues, one per XCD (workgroup b runs on XCD b % 8), each
// drained through an atomic head; a workgroup whose own queue is empty
// takes items from the others, so uneven scenario sizes still balance.
// Device-scope atomics on one address are serialised across the XCDs, so
// a claim takes a run of consecutive items (about three claims per
// workgroup and queue) and a queue is probed with a plain load before any
// claim on it.  Every workgroup leaves once all 8 heads have passed their
// ends.
template <int MODE, bool NL, bool SPH, int TX, int TY>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(StepWaves<NL, SPH, tile_lds_bytes<TX, TY>()>::value)))
    k_batch(const BatchDesc *__restrict__ descs, const BatchItem *__restrict__ items,
            const int *__restrict__ qstart, int *heads) {
  __shared__ float sm[6 * (TX + 4) * (TY + 4) + kNumCoef * (TY + 4)];
  __shared__ gcdet::Tables stab;
  __shared__ int s_item;
  if (NL || !SPH) stage_tables(stab, threadIdx.x);
  const int x = blockIdx.x & 7;
  for (int c = 0; c < 8; ++c) {
    const int q = (x + c) & 7;
    const int lo = qstart[q], hi = qstart[q + 1];
    if (hi <= lo) continue;
    const int nbq = ((int)gridDim.x + 7 - q) / 8;  // workgroups of that queue's XCD
    const int per = max(1, (hi - lo + 3 * nbq - 1) / (3 * nbq));
    for (;;) {
      __syncthreads();  // previous run done with LDS and s_item
      if (threadIdx.x == 0) {
        const int seen = __hip_atomic_load(&heads[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_item = lo + seen >= hi ? hi : lo + atomicAdd(&heads[q], per);
      }
      __syncthreads();
      const int i0 = __builtin_amdgcn_readfirstlane(s_item);
      if (i0 >= hi) break;
      const int i1 = min(i0 + per, hi);
      for (int it = i0; it < i1; ++it) {
        if (it > i0) __syncthreads();  // previous tile done with LDS
        const BatchItem w = items[it];
        const int slot = __builtin_amdgcn_readfirstlane(w.slot);
        const int tile = __builtin_amdgcn_readfirstlane(w.tile);
        const DevLayer L = descs[slot].L;
        tile_step<MODE, NL, SPH, TX, TY>(L, tile, descs[slot].tiles_x, sm, &stab);
      }
    }
  }
}

// The same work as one workgroup per item (the hardware dispatcher does
// the balancing): item (b % 8) * per_xcd + b / 8, so each XCD takes one
// contiguous eighth of the items, as the queues split them.
template <int MODE, bool NL, bool SPH, int TX, int TY>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(StepWaves<NL, SPH, tile_lds_bytes<TX, TY>()>::value)))
    k_batch_grid(const BatchDesc *__restrict__ descs, const BatchItem *__restrict__ items,
                 int nitems, int per_xcd) {
  __shared__ float sm[6 * (TX + 4) * (TY + 4) + kNumCoef * (TY + 4)];
  __shared__ gcdet::Tables stab;
  const int it = tile_of_block(per_xcd);
  if (it >= nitems) return;
  const BatchItem w = items[it];
  const int slot = __builtin_amdgcn_readfirstlane(w.slot);
  const int tile = __builtin_amdgcn_readfirstlane(w.tile);
  const DevLayer L = descs[slot].L;
  tile_step<MODE, NL, SPH, TX, TY, false, NL || !SPH>(L, tile, descs[slot].tiles_x, sm, &stab);
}


// Two root steps of single-layer linear scenarios, one workgroup per
// (layer, tile) item of a TX x TY tiling, NT threads: the tile body of
// k_step2 on each item's layer.
template <bool SPH, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT)
    k_batch2_grid(const BatchDesc *__restrict__ descs, const BatchItem *__restrict__ items,
                  int nitems, int per_xcd) {
  constexpr int HW = 4;
  __shared__ float sm[7 * (TX + 2 * HW) * (TY + 2 * HW) + kNumCoef * (TY + 2 * HW)];
  __shared__ gcdet::Tables stab;
  const int it = tile_of_block(per_xcd);
  if (it >= nitems) return;
  const BatchItem w = items[it];
  const int slot = __builtin_amdgcn_readfirstlane(w.slot);
  const int t = __builtin_amdgcn_readfirstlane(w.tile);
  const DevLayer L = descs[slot].L;
  const int tiles_x = descs[slot].tiles_x;
  const int ti0 = (t % tiles_x) * TX + 1, tj0 = (t / tiles_x) * TY + 1;
  const bool interior = ti0 - HW > kSpan + 2 && ti0 + TX + HW < L.nx - kSpan &&
                        tj0 - HW > kSpan + 2 && tj0 + TY + HW < L.ny - kSpan;
  if (interior) tile_body2<SPH, TX, TY, true, NT, false, !SPH>(L, t, ti0, tj0, sm, &stab);
  else tile_body2<SPH, TX, TY, false, NT, false, !SPH>(L, t, ti0, tj0, sm, &stab);
}

template <int MODE, bool NL, bool SPH, int TX, int TY>
hipError_t launch_t(const DevLayer &L, hipStream_t st) {
  const int tx = (L.nx + TX - 1) / TX, ty = (L.ny + TY - 1) / TY;
  const int nt = tx * ty, per = (nt + 7) / 8;
  if (L.stream_out)
    hipLaunchKernelGGL((k_step<MODE, NL, SPH, TX, TY, true>), dim3(8 * per), dim3(kThreads), 0,
                       st, L, tx, nt, per);
  else
    hipLaunchKernelGGL((k_step<MODE, NL, SPH, TX, TY, false>), dim3(8 * per), dim3(kThreads), 0,
                       st, L, tx, nt, per);
  return hipGetLastError();
}

template <int TX, int TY>
hipError_t launch_shape(const DevLayer &L, int mode, bool nl, bool sph, hipStream_t st) {
#define GC_CASE(MD, A, B)                                  \
  if (mode == MD && nl == A && sph == B) return launch_t<MD, A, B, TX, TY>(L, st);
  GC_CASE(kFused, false, true) GC_CASE(kFused, false, false)
  GC_CASE(kFused, true, true) GC_CASE(kFused, true, false)
  GC_CASE(kMassOnly, false, true) GC_CASE(kMassOnly, false, false)
  GC_CASE(kMassOnly, true, true) GC_CASE(kMassOnly, true, false)
  GC_CASE(kMomentOnly, false, true) GC_CASE(kMomentOnly, false, false)
  GC_CASE(kMomentOnly, true, true) GC_CASE(kMomentOnly, true, false)
#undef GC_CASE
  return hipErrorInvalidValue;
}

template <bool SPH, int TX, int TY, int NT = kThreads>
hipError_t launch2_t(const DevLayer &L, hipStream_t st) {
  const int tx = (L.nx + TX - 1) / TX, ty = (L.ny + TY - 1) / TY;
  const int nt = tx * ty, per = (nt + 7) / 8;
  if (L.stream_out)
    hipLaunchKernelGGL((k_step2<SPH, TX, TY, NT, true>), dim3(8 * per), dim3(NT), 0, st, L, tx,
                       nt, per);
  else
    hipLaunchKernelGGL((k_step2<SPH, TX, TY, NT, false>), dim3(8 * per), dim3(NT), 0, st, L, tx,
                       nt, per);
  return hipGetLastError();
}

// Fused-step launch shapes: LDS tiles TX x TY (variants 0-5; 0 is the
// default), then the strip kernel with SEG rows per segment (6-10).
const int kVariants[][3] = {{1, 64, 8},   {1, 64, 16},  {1, 32, 32},  {1, 128, 8},
                            {1, 64, 32},  {1, 32, 16},  {0, 256, 32}, {0, 256, 16},
                            {0, 256, 24}, {0, 256, 48}, {0, 256, 64}};

}  // namespace

int num_tile_variants() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
void tile_variant_shape(int v, int *tx, int *ty) {
  *tx = kVariants[v][0] ? kVariants[v][1] : -kVariants[v][1];  // negative: strip kernel
  *ty = kVariants[v][2];
}

hipError_t launch_step(const DevLayer &L, int mode, bool nl, bool sph, int variant,
                       hipStream_t st) {
  if (variant < 0 || variant >= num_tile_variants()) return hipErrorInvalidValue;
  if (mode == kFused && kVariants[variant][0] == 0) {
    const int seg = kVariants[variant][2];
    if (nl) return sph ? launch_stream<true, true>(L, seg, st) : launch_stream<true, false>(L, seg, st);
    return sph ? launch_stream<false, true>(L, seg, st) : launch_stream<false, false>(L, seg, st);
  }
  if (mode == kMomentMass) {  // outermost linear layers with nested layers: 64x8
    if (nl) return hipErrorInvalidValue;
    return sph ? launch_t<kMomentMass, false, true, 64, 8>(L, st)
               : launch_t<kMomentMass, false, false, 64, 8>(L, st);
  }
  // split continuity / momentum launches (parents of nested layers) use a 64x8 tile
  const int v = mode == kFused ? variant : 0;
  switch (kVariants[v][1] * 1000 + kVariants[v][2]) {
    case 64016: return launch_shape<64, 16>(L, mode, nl, sph, st);
    case 64008: return launch_shape<64, 8>(L, mode, nl, sph, st);
    case 128008: return launch_shape<128, 8>(L, mode, nl, sph, st);
    case 32032: return launch_shape<32, 32>(L, mode, nl, sph, st);
    case 64032: return launch_shape<64, 32>(L, mode, nl, sph, st);
    case 32016: return launch_shape<32, 16>(L, mode, nl, sph, st);
    default: return hipErrorInvalidValue;
  }
}

// Two-step shapes: tile TX x TY and threads NT (the same table for the
// single-tree and the batch launches).
#define GC_STEP2_SHAPES(X) \
  X(0, 64, 8, 256) X(1, 32, 16, 256) X(2, 64, 16, 256) X(3, 64, 16, 512) X(4, 64, 16, 1024) \
  X(5, 32, 16, 512) X(6, 64, 32, 1024)

hipError_t launch_step2(const DevLayer &L, bool sph, int shape, hipStream_t st) {
#define GC_S2(K, TX, TY, NT) \
  case K: return sph ? launch2_t<true, TX, TY, NT>(L, st) : launch2_t<false, TX, TY, NT>(L, st);
  switch (shape) {
    GC_STEP2_SHAPES(GC_S2)
    default: return hipErrorInvalidValue;
  }
#undef GC_S2
}

void step2_shape(int shape, int *tx, int *ty) {
#define GC_S2(K, TX, TY, NT) \
  case K: *tx = TX; *ty = TY; return;
  switch (shape) {
    GC_STEP2_SHAPES(GC_S2)
    default: *tx = 64; *ty = 16; return;
  }
#undef GC_S2
}

template <bool SPH, int TX, int TY, int NT>
hipError_t launch_batch2_t(const BatchLaunch &B, hipStream_t st) {
  const int per = (B.nitems + 7) / 8;
  hipLaunchKernelGGL((k_batch2_grid<SPH, TX, TY, NT>), dim3(8 * per), dim3(NT), 0, st, B.descs,
                     B.items, B.nitems, per);
  return hipGetLastError();
}

hipError_t launch_batch2(const BatchLaunch &B, bool sph, int shape, hipStream_t st) {
  if (B.nitems <= 0) return hipSuccess;
#define GC_S2(K, TX, TY, NT) \
  case K: return sph ? launch_batch2_t<true, TX, TY, NT>(B, st) : launch_batch2_t<false, TX, TY, NT>(B, st);
  switch (shape) {
    GC_STEP2_SHAPES(GC_S2)
    default: return hipErrorInvalidValue;
  }
#undef GC_S2
}

hipError_t launch_max(const DevLayer &L, float *zmax_raw, hipStream_t st) {
  const size_t n = (size_t)L.nx * L.ny;
  hipLaunchKernelGGL(k_max, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, zmax_raw);
  return hipGetLastError();
}

hipError_t launch_jnq(const DevNest &N, int t, hipStream_t st) {
  const int total = 2 * (N.je - N.js + 5) + 2 * (N.ie - N.is + 5);
  hipLaunchKernelGGL(k_jnq, dim3((total + 127) / 128), dim3(128), 0, st, N, t);
  return hipGetLastError();
}

hipError_t launch_jnz(const DevNest &N, hipStream_t st) {
  const int total = (N.ie - N.is + 1) * (N.je - N.js + 1);
  hipLaunchKernelGGL(k_jnz, dim3((total + 127) / 128), dim3(128), 0, st, N);
  return hipGetLastError();
}

// ------------------------------------------------------------ batch launchers
namespace {

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
      cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

template <int MODE, bool NL, bool SPH, int TX, int TY>
hipError_t launch_batch_t(const BatchLaunch &B, hipStream_t st, bool prepare, bool grid_form) {
  static int resident = 0;  // workgroups that fit on the chip at once
  if (!resident) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(&k_batch<MODE, NL, SPH, TX, TY>), kThreads, 0);
    if (e != hipSuccess) return e;
    resident = (per_cu > 0 ? per_cu : 1) * device_cus();
  }
  if (prepare) return hipSuccess;
  if (grid_form) {
    const int per = (B.nitems + 7) / 8;
    hipLaunchKernelGGL((k_batch_grid<MODE, NL, SPH, TX, TY>), dim3(8 * per), dim3(kThreads), 0,
                       st, B.descs, B.items, B.nitems, per);
    return hipGetLastError();
  }
  // at least four items per workgroup: fewer, and the claims cost more
  // than they balance
  int grid = resident;
  const int want = (B.nitems / 4 + 7) / 8 * 8;
  if (want < grid) grid = want;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL((k_batch<MODE, NL, SPH, TX, TY>), dim3(grid), dim3(kThreads), 0, st,
                     B.descs, B.items, B.qstart, B.heads);
  return hipGetLastError();
}

template <int TX, int TY>
hipError_t batch_shape_t(const BatchLaunch &B, int mode, bool nl, bool sph, hipStream_t st,
                         bool prepare, bool grid_form) {
#define GC_CASE(MD, A, C) \
  if (mode == MD && nl == A && sph == C) \
    return launch_batch_t<MD, A, C, TX, TY>(B, st, prepare, grid_form);
  GC_CASE(kFused, false, true) GC_CASE(kFused, false, false)
  GC_CASE(kFused, true, true) GC_CASE(kFused, true, false)
#undef GC_CASE
  return hipErrorInvalidValue;
}

// fused tile shapes; shape k < 3 runs the persistent kernel with tile k,
// k >= 3 the one-workgroup-per-tile form with tile k - 3
const int kBatchShapes[][2] = {{64, 8}, {32, 16}, {64, 16}};
constexpr int kNumBatchTiles = 3;

}  // namespace

int batch_num_shapes() { return 2 * kNumBatchTiles; }
void batch_shape(int k, int *tx, int *ty) {
  *tx = kBatchShapes[k % kNumBatchTiles][0];
  *ty = kBatchShapes[k % kNumBatchTiles][1];
}

hipError_t launch_batch(const BatchLaunch &B, int mode, bool nl, bool sph, int shape,
                        hipStream_t st, bool prepare) {
  if (B.nitems <= 0 && !prepare) return hipSuccess;
  if (shape < 0 || shape >= batch_num_shapes()) return hipErrorInvalidValue;
  const bool grid_form = shape >= kNumBatchTiles;
  if (mode != kFused) {  // continuity / momentum halves of nested parents: 64 x 8
#define GC_CASE(MD, A, C) \
    if (mode == MD && nl == A && sph == C) \
      return launch_batch_t<MD, A, C, 64, 8>(B, st, prepare, grid_form);
    GC_CASE(kMassOnly, false, true) GC_CASE(kMassOnly, false, false)
    GC_CASE(kMassOnly, true, true) GC_CASE(kMassOnly, true, false)
    GC_CASE(kMomentOnly, false, true) GC_CASE(kMomentOnly, false, false)
    GC_CASE(kMomentOnly, true, true) GC_CASE(kMomentOnly, true, false)
    GC_CASE(kMomentMass, false, true) GC_CASE(kMomentMass, false, false)
#undef GC_CASE
    return hipErrorInvalidValue;
  }
  switch (shape % kNumBatchTiles) {
    case 0: return batch_shape_t<64, 8>(B, mode, nl, sph, st, prepare, grid_form);
    case 1: return batch_shape_t<32, 16>(B, mode, nl, sph, st, prepare, grid_form);
    case 2: return batch_shape_t<64, 16>(B, mode, nl, sph, st, prepare, grid_form);
    default: return hipErrorInvalidValue;
  }
}

int jnq_block_count(const DevNest &N) {
  return (2 * (N.je - N.js + 5) + 2 * (N.ie - N.is + 5) + 127) / 128;
}
int jnz_block_count(const DevNest &N) {
  return ((N.ie - N.is + 1) * (N.je - N.js + 1) + 127) / 128;
}

hipError_t launch_jnq_multi(const DevNest *nests, const int *start, int nnest, int nblocks,
                            int t, hipStream_t st) {
  if (nnest <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jnq_multi, dim3(nblocks), dim3(128), 0, st, nests, start, nnest, t);
  return hipGetLastError();
}

hipError_t launch_jnz_multi(const DevNest *nests, const int *start, int nnest, int nblocks,
                            hipStream_t st) {
  if (nnest <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jnz_multi, dim3(nblocks), dim3(128), 0, st, nests, start, nnest);
  return hipGetLastError();
}

// The step's Manning power x^2.3333333 and breaking exponential e^x, as the
// momentum kernels evaluate them, for a list of arguments.
__global__ void __launch_bounds__(256)
    k_detmath_eval(const float *x, long long n, float *pw, float *ex) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    const float v = x[q];
    pw[q] = gcdet::powf_manning(v, kManExp, gcdet::kTab);
    ex[q] = gcdet::expf_det(v, gcdet::kTab);
  }
}

hipError_t launch_detmath_eval(const float *x, long long n, float *pw, float *ex,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long long blocks = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_detmath_eval, dim3((unsigned)blocks), dim3(256), 0, st, x, n, pw, ex);
  return hipGetLastError();
}

hipError_t launch_division_check(unsigned long long seed, int per_lane, unsigned long long *out,
                                 int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_division_check, dim3(blocks), dim3(256), 0, st, seed, per_lane, out);
  return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16, per = 256 * (size_t)kCopyUnroll;
  const size_t blocks = (n + per - 1) / per;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st,
                     static_cast<const CopyVec *>(src), static_cast<CopyVec *>(dst), n);
  return hipGetLastError();
}

hipError_t launch_out_pack(const OutPack &p, hipStream_t st) {
  if (p.count <= 0) return hipSuccess;
  const long long blocks = std::min<long long>((p.count + 255) / 256, 8192);
  hipLaunchKernelGGL(k_out_pack, dim3((unsigned)blocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_sc_edges(const DevSc &S, int t, hipStream_t st) {
  if (t > 1) {  // NEWQ_SC: the band forecasts first
    int n = 0;
    for (int b = 0; b < 4; ++b)
      n += (S.band[b][1] - S.band[b][0] + 1) * (S.band[b][3] - S.band[b][2] + 1);
    hipLaunchKernelGGL(k_sc_band, dim3((n + 255) / 256), dim3(256), 0, st, S);
    if (hipError_t e = hipGetLastError()) return e;
  }
  const int m = 2 * S.cnx + 2 * S.cny;
  hipLaunchKernelGGL(k_sc_edges, dim3((m + 255) / 256), dim3(256), 0, st, S, t);
  return hipGetLastError();
}

hipError_t launch_sc_jnz(const DevSc &S, hipStream_t st) {
  const int n = (S.corners[1] - S.corners[0] + 1) * (S.corners[3] - S.corners[2] + 1);
  hipLaunchKernelGGL(k_sc_jnz, dim3((n + 255) / 256), dim3(256), 0, st, S);
  return hipGetLastError();
}

hipError_t launch_sc_cell_interp(const float *parent, int pnx, int pny, float *child, int cnx,
                                 int cny, const int *pos, const float *cxy, hipStream_t st) {
  const long long n = (long long)cnx * cny;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sc_cell_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     parent, pnx, pny, child, cnx, cny, pos, cxy);
  return hipGetLastError();
}

hipError_t launch_okada(const FaultDev &f, float *deform, hipStream_t st) {
  const long long n = (long long)f.nx * f.ny;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_okada, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, deform);
  return hipGetLastError();
}

hipError_t launch_cell_interp(const float *parent, int pnx, float *child, int cnx, int cny,
                              const int *corners, int rel_size, hipStream_t st) {
  const long long n = (long long)cnx * cny;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, parent,
                     pnx, child, cnx, cny, corners[0], corners[1], corners[2], corners[3],
                     rel_size);
  return hipGetLastError();
}

hipError_t launch_apply_deform(float *z1, float *h, const float *deform, float *zraw, size_t n,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_deform, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, z1, h,
                     deform, zraw, (long long)n);
  return hipGetLastError();
}

hipError_t launch_gauges(const GaugeDev *g, int n, const ZNow &z, const int *step_now,
                         float *ring, int cap, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gauges, dim3((n + 63) / 64), dim3(64), 0, st, g, n, z, step_now, ring, cap);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int n, int v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace gc
