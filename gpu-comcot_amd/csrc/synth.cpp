// Seeded synthetic tsunami scenarios: bathymetry and initial surface.
//
// The reference reads bathymetry files and computes Okada sources
// (initialization.f90 READ_*_BATHY, deform.f90); neither is on the step's
// path, so scenarios here are generated from a seed: a 4000 m abyssal plain
// with Gaussian seamounts, a shelf ramp rising to dry land along the west
// edge (land strip, coastline, shelf ramp to the plain), optional
// embayments with 1:50 beaches cut into the land, and an elliptical Gaussian
// uplift or a dipole (uplift/subsidence pair) source.  The generator is a
// pure function of (seed, point), so any grid, parent or child, samples the
// same world.  Random numbers: SplitMix64.
#include <cmath>
#include <cstdint>

#include "gcomcot_synth.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
};

const double kPi = 3.14159265358979323846;
const double kKmPerDeg = 111.32;

}  // namespace

extern "C" {

void gcomcot_synth_world(unsigned long long seed, double lon0, double lat0, double lon1,
                         double lat1, int n_seamounts, int n_embayments, gc_world_t *w) {
  SplitMix64 rng(seed);
  w->seed = seed;
  w->lon0 = lon0;
  w->lat0 = lat0;
  w->lon1 = lon1;
  w->lat1 = lat1;
  w->shelf_frac = 0.12;
  w->land_height = -100.0;
  w->ocean_depth = 4000.0;
  w->beach_slope = 50.0;
  w->bay_floor = 30.0;
  w->coast_frac = 0.05;
  if (n_seamounts > GC_MAX_SEAMOUNTS) n_seamounts = GC_MAX_SEAMOUNTS;
  if (n_embayments > GC_MAX_EMBAYMENTS) n_embayments = GC_MAX_EMBAYMENTS;
  w->n_seamounts = n_seamounts;
  w->n_embayments = n_embayments;
  const double wx = lon1 - lon0, wy = lat1 - lat0;
  for (int k = 0; k < n_seamounts; ++k) {
    w->sm_lon[k] = lon0 + wx * rng.uni(w->coast_frac + w->shelf_frac + 0.05, 0.95);
    w->sm_lat[k] = lat0 + wy * rng.uni(0.05, 0.95);
    w->sm_height[k] = rng.uni(500.0, 2500.0);
    w->sm_sigma_km[k] = rng.uni(20.0, 60.0);
  }
  for (int k = 0; k < n_embayments; ++k) {
    double c = 0.2 + 0.6 * (k + 0.5) / n_embayments + 0.05 * (rng.uni() - 0.5);
    w->bay_lat[k] = lat0 + wy * c;
    w->bay_halfwidth_km[k] = rng.uni(25.0, 60.0);
    w->bay_depth_km[k] = rng.uni(30.0, 80.0);
  }
}

void gcomcot_synth_depth(const gc_world_t *w, long n, const double *lon, const double *lat,
                         float *h) {
  const double wx = w->lon1 - w->lon0;
  for (long p = 0; p < n; ++p) {
    const double cl = cos(lat[p] * kPi / 180.0);
    const double xkm = (lon[p] - w->lon0) * kKmPerDeg * cl;
    const double shelf_km = w->shelf_frac * wx * kKmPerDeg * cl;
    const double xc = w->coast_frac * wx * kKmPerDeg * cl;  // coastline
    // land rises west of the coast with the ramp's slope up to land_height
    double d = xkm < xc + shelf_km ? w->ocean_depth * (xkm - xc) / shelf_km : w->ocean_depth;
    if (d < w->land_height) d = w->land_height;
    const double ramp = d;
    for (int k = 0; k < w->n_seamounts; ++k) {
      double dx = (lon[p] - w->sm_lon[k]) * kKmPerDeg * cl;
      double dy = (lat[p] - w->sm_lat[k]) * kKmPerDeg;
      double s = w->sm_sigma_km[k];
      if (ramp > 50.0) d -= w->sm_height[k] * exp(-(dx * dx + dy * dy) / (2.0 * s * s));
    }
    if (ramp > 50.0 && d < 50.0) d = 50.0;
    // embayments cut into the land behind the coastline
    for (int k = 0; k < w->n_embayments; ++k) {
      double dykm = (lat[p] - w->bay_lat[k]) * kKmPerDeg;
      double hw = w->bay_halfwidth_km[k];
      if (fabs(dykm) >= hw) continue;
      double c = cos(0.5 * kPi * dykm / hw);
      double xb = xc - w->bay_depth_km[k] * c * c;
      double hb = (xkm - xb) * 1000.0 / w->beach_slope;
      if (hb > w->bay_floor) hb = w->bay_floor;
      if (hb < w->land_height) hb = w->land_height;
      if (hb > d) d = hb;
    }
    h[p] = (float)d;
  }
}

void gcomcot_synth_source(const gc_world_t *w, int kind, double peak, double sig_a_km,
                          double sig_b_km, double sep_km, double cx_frac, double cy_frac,
                          double strike, gc_source_t *s) {
  s->kind = kind;
  s->peak = peak;
  s->sig_a_km = sig_a_km;
  s->sig_b_km = sig_b_km;
  s->sep_km = sep_km;
  s->clon = w->lon0 + (w->lon1 - w->lon0) * cx_frac;
  s->clat = w->lat0 + (w->lat1 - w->lat0) * cy_frac;
  s->strike = strike;
}

void gcomcot_synth_source_draw(const gc_world_t *w, unsigned long long stream, int kind,
                               double peak, double sig_a_km, double sig_b_km, double sep_km,
                               gc_source_t *s) {
  SplitMix64 rng(w->seed ^ (0xA5A5A5A5ULL + stream * 0x9E3779B97F4A7C15ULL));
  double cx = rng.uni(0.35, 0.7), cy = rng.uni(0.3, 0.7), strike = rng.uni(0.0, kPi);
  gcomcot_synth_source(w, kind, peak, sig_a_km, sig_b_km, sep_km, cx, cy, strike, s);
}

void gcomcot_synth_eta(const gc_source_t *s, long n, const double *lon, const double *lat,
                       const float *h, float *z0) {
  const double cs = cos(s->strike), sn = sin(s->strike);
  for (long p = 0; p < n; ++p) {
    const double cl = cos(lat[p] * kPi / 180.0);
    double dx = (lon[p] - s->clon) * kKmPerDeg * cl;
    double dy = (lat[p] - s->clat) * kKmPerDeg;
    double u = dx * cs + dy * sn;   // along strike
    double v = -dx * sn + dy * cs;  // across strike
    double e;
    if (s->kind == 0) {
      e = s->peak * exp(-(u * u / (2.0 * s->sig_a_km * s->sig_a_km) +
                          v * v / (2.0 * s->sig_b_km * s->sig_b_km)));
    } else {
      double hs = 0.5 * s->sep_km, sg = s->sig_a_km;
      double vp = v - hs, vm = v + hs;
      e = s->peak * (exp(-(u * u + vp * vp) / (2.0 * sg * sg)) -
                     exp(-(u * u + vm * vm) / (2.0 * sg * sg)));
    }
    if (h[p] <= 1.0e-5f) e = 0.0;
    z0[p] = (float)e;
  }
}

}  // extern "C"
