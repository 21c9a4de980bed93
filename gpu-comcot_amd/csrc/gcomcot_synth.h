/* Seeded synthetic scenario generator (host only; see synth.cpp). */
#ifndef GCOMCOT_SYNTH_H
#define GCOMCOT_SYNTH_H

#ifdef __cplusplus
extern "C" {
#endif

#define GC_MAX_SEAMOUNTS 16
#define GC_MAX_EMBAYMENTS 8

typedef struct {
  unsigned long long seed;
  double lon0, lat0, lon1, lat1;   /* root domain, cell centres of the corner cells */
  double shelf_frac, land_height, ocean_depth;
  int n_seamounts, n_embayments;
  double sm_lon[GC_MAX_SEAMOUNTS], sm_lat[GC_MAX_SEAMOUNTS];
  double sm_height[GC_MAX_SEAMOUNTS], sm_sigma_km[GC_MAX_SEAMOUNTS];
  double bay_lat[GC_MAX_EMBAYMENTS], bay_halfwidth_km[GC_MAX_EMBAYMENTS];
  double bay_depth_km[GC_MAX_EMBAYMENTS];
  double beach_slope;              /* beaches rise 1 m per beach_slope m (50) */
  double bay_floor;                /* embayment floor depth, m (30) */
  double coast_frac;               /* coastline at this fraction of the width (0.05) */
} gc_world_t;

typedef struct {
  int kind;                        /* 0 elliptical Gaussian, 1 dipole */
  double peak, sig_a_km, sig_b_km, sep_km, clon, clat, strike;
} gc_source_t;

void gcomcot_synth_world(unsigned long long seed, double lon0, double lat0, double lon1,
                         double lat1, int n_seamounts, int n_embayments, gc_world_t *w);
void gcomcot_synth_depth(const gc_world_t *w, long n, const double *lon, const double *lat,
                         float *h);
void gcomcot_synth_source(const gc_world_t *w, int kind, double peak, double sig_a_km,
                          double sig_b_km, double sep_km, double cx_frac, double cy_frac,
                          double strike, gc_source_t *s);
void gcomcot_synth_source_draw(const gc_world_t *w, unsigned long long stream, int kind,
                               double peak, double sig_a_km, double sig_b_km, double sep_km,
                               gc_source_t *s);
void gcomcot_synth_eta(const gc_source_t *s, long n, const double *lon, const double *lat,
                       const float *h, float *z0);

#ifdef __cplusplus
}
#endif
#endif
