// Layer trees: registration, device planes and the step's enqueue order.
//
// Children are kept in registration order like the reference's sibling
// list (GPUcomcot.cu:157-190), which is also the order ALL_GRID visits
// them (all_grids.f90:34-35).  Each layer owns two time levels of Z, M, N
// on the device; a level rotation swaps them.
//
// Device memory per layer, one allocation: H, Z x2, M x2, N x2, raw max,
// wet max, arrival (11 planes of 4 B = 44 B/cell; +4 for the DZ plane of
// nonlinear layers), the per-row coefficient table and the interface
// flux cache.  Planes start on 256-byte boundaries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "comcot_consts.h"
#include "tree.h"

namespace gc {

// ------------------------------------------------------------ errors
int g_fatal = 1;
static int g_last_code = 0;
static char g_last_msg[512] = {0};

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_msg, sizeof(g_last_msg), fmt, ap);
  va_end(ap);
  g_last_code = code;
  if (g_fatal) {
    fprintf(stderr, "gcomcot_hip: %s\n", g_last_msg);
    exit(EXIT_FAILURE);
  }
  return code;
}

const char *last_message() { return g_last_msg; }

int take_last_code() {
  int c = g_last_code;
  g_last_code = 0;
  g_last_msg[0] = 0;
  return c;
}

// ------------------------------------------------------------ memory
namespace {

size_t pad(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

struct Carve {
  char *base;
  size_t off = 0;
  template <class T>
  T *take(size_t count) {
    T *p = reinterpret_cast<T *>(base + off);
    off += pad(count * sizeof(T));
    return p;
  }
};

}  // namespace

size_t layer_bytes(int nx, int ny, bool nl, bool r24) {
  const size_t n = (size_t)nx * ny;
  size_t b = 9 * pad(n * 4) + pad(n * 4);  // h, z x2, m x2, n x2, zmax x2; arrival
  if (nl) b += pad(n * 4);
  b += pad(2 * (size_t)nx * 4) + pad(2 * (size_t)ny * 4) + pad((size_t)kNumCoef * ny * 4);
  if (r24) b += pad(2 * n * 4);
  return b;
}

size_t Tree::device_bytes() const {
  size_t b = 0;
  for (const auto &L : layers)
    if (L.used) b += L.block_bytes + (L.r24 ? 2 * L.nxy() * 4 : 0);
  return b;
}

static void free_layer(LayerRec &L) {
  if (L.block) (void)hipFree(L.block);
  if (L.r24) (void)hipFree(L.r24);
  if (L.deform) (void)hipFree(L.deform);
  if (L.sponge) (void)hipFree(L.sponge);
  if (L.sc_pos) (void)hipFree(L.sc_pos);
  if (L.sc_cxy) (void)hipFree(L.sc_cxy);
  if (L.sc_jnz) (void)hipFree(L.sc_jnz);
  L = LayerRec();
}

void tree_free(Tree &T) {
  for (auto &L : T.layers)
    if (L.used) free_layer(L);
  if (T.gauges) (void)hipFree(T.gauges);
  if (T.ring) (void)hipFree(T.ring);
  T.gauges = nullptr;
  T.ring = nullptr;
  T.ngauges = 0;
  for (auto &v : T.xs) v.clear();
  for (auto &v : T.ys) v.clear();
  if (T.ctl) (void)hipFree(T.ctl);
  T.ctl = nullptr;
  T.ctl_in = 0;
}

int tree_validate_layer(Tree &T, int id, int parent, int rel_size, int rel_time,
                        const int *corners, int nx, int ny, bool sc) {
  if (id < 1 || id >= kMaxLayers) return fail(1, "invalid layer id %d (1..%d)", id, kMaxLayers - 1);
  if (nx < 3 || ny < 3) return fail(2, "layer %d: grid %dx%d too small", id, nx, ny);
  if (id == 1 && parent != -1) return fail(3, "layer 1 must be the root (parent -1)");
  if (id != 1) {
    LayerRec *P = T.layer(parent);
    if (!P) return fail(4, "layer %d: parent %d not initialised", id, parent);
    const int is = corners[0], ie = corners[1], js = corners[2], je = corners[3];
    if (sc) {  // the POS extent: NEWQ_SC's bands reach three parent cells beyond it
      if (P->id != 1 || P->laycord != 0)
        return fail(80, "layer %d: SC_OPTION 1 needs the spherical layer 1 as parent", id);
      if (is < 4 || js < 4 || ie > P->nx - 4 || je > P->ny - 4 || ie < is || je < js)
        return fail(5, "layer %d: SC extent I %d..%d J %d..%d must lie in 4..NX-4, 4..NY-4", id,
                    is, ie, js, je);
      if (rel_size < 1 || rel_time < 1)
        return fail(7, "layer %d: REL_SIZE %d / REL_TIME %d < 1", id, rel_size, rel_time);
      return 0;
    }
    if (is < 4 || js < 4 || ie > P->nx - 2 || je > P->ny - 2 || ie < is || je < js)
      return fail(5, "layer %d: footprint I %d..%d J %d..%d must lie in 4..NX-2, 4..NY-2 of its parent",
                  id, is, ie, js, je);
    if (rel_size < 1 || nx != (ie - is + 1) * rel_size + 1 || ny != (je - js + 1) * rel_size + 1)
      return fail(6, "layer %d: %dx%d does not match footprint x REL_SIZE %d + 1", id, nx, ny,
                  rel_size);
    if (rel_time < 1) return fail(7, "layer %d: REL_TIME %d < 1", id, rel_time);
  }
  return 0;
}

void tree_drop_layer(Tree &T, int id) {
  if (id < 2 || id >= kMaxLayers || !T.layers[id].used) return;
  if (LayerRec *op = T.layer(T.layers[id].parent)) {
    auto &ch = op->children;
    for (size_t q = 0; q < ch.size(); ++q)
      if (ch[q] == id) { ch.erase(ch.begin() + q); break; }
  }
  free_layer(T.layers[id]);
}

int tree_create_layer(Tree &T, hipStream_t st, int id, int parent, int level, int laycord,
                      int laygov, int rel_size, int rel_time, const int *corners,
                      int fric_switch, float fric_coef, float dt, float dx, float dy, int nx,
                      int ny, const float *H, const float *Z, bool sc) {
  if (int e = tree_validate_layer(T, id, parent, rel_size, rel_time, corners, nx, ny, sc))
    return e;
  if (!T.ctl) {
    HIPCHK(hipMalloc(&T.ctl, 2 * sizeof(int)));
    HIPCHK(hipMemsetAsync(T.ctl, 0, 2 * sizeof(int), st));
  }
  if (T.layers[id].used) {
    if (id == 1) {  // a new root starts a new layer tree
      for (int k = 1; k < kMaxLayers; ++k)
        if (T.layers[k].used) free_layer(T.layers[k]);
      if (T.gauges) (void)hipFree(T.gauges);
      if (T.ring) (void)hipFree(T.ring);
      T.gauges = nullptr;
      T.ring = nullptr;
      T.ngauges = 0;
      for (auto &v : T.xs) v.clear();
      for (auto &v : T.ys) v.clear();
    } else {
      tree_drop_layer(T, id);
    }
  }
  if (id == 1) {
    T.ctl_in = 0;
    HIPCHK(hipMemsetAsync(T.ctl, 0, 2 * sizeof(int), st));
  }
  T.mass_ahead = false;
  LayerRec &L = T.layers[id];
  L.used = true;
  L.id = id;
  L.parent = id == 1 ? -1 : parent;
  L.level = level;
  L.laycord = laycord;
  L.laygov = laygov;
  L.sph = laycord == 0;
  L.nl = L.sph ? (laygov == 1 || laygov == 9) : (laygov == 1 || laygov == 5 || laygov == 9);
  L.rel_size = id == 1 ? 1 : rel_size;
  L.rel_time = id == 1 ? 1 : rel_time;
  if (id == 1) {
    L.corners[0] = 1; L.corners[1] = nx; L.corners[2] = 1; L.corners[3] = ny;
  } else {
    for (int k = 0; k < 4; ++k) L.corners[k] = corners[k];
  }
  L.fric = fric_switch != 1;
  L.fric_coef = fric_coef;
  L.dt = dt;
  L.dx = dx;
  L.dy = dy;
  // ALLOC (comcot.f90:537-552, 602); CONMOME drops convection for LAYGOV 9
  L.rx = dt / dx;
  L.ry = dt / dy;
  L.grx = kGrav * L.rx;
  L.gry = kGrav * L.ry;
  L.modscm = id == 1 ? 0 : 1;
  if (laygov % 2 != 0) L.modscm = 1;
  L.linchk = laygov == 9 ? 0 : 1;
  const float cst = 0.5f * dt * kGrav;
  L.kfric = cst * fric_coef * fric_coef;
  L.nx = nx;
  L.ny = ny;
  const size_t n = L.nxy();
  L.block_bytes = layer_bytes(nx, ny, L.nl != 0, false);
  {
    hipError_t e = hipMalloc(&L.block, L.block_bytes);
    if (e != hipSuccess) {
      const size_t want = L.block_bytes;
      L = LayerRec();
      return fail(100 + (int)e, "layer %d: %zu bytes of device memory: %s", id, want,
                  hipGetErrorString(e));
    }
  }
  Carve c{static_cast<char *>(L.block)};
  L.h = c.take<float>(n);
  for (int k = 0; k < 2; ++k) {
    L.z[k] = c.take<float>(n);
    L.m[k] = c.take<float>(n);
    L.n[k] = c.take<float>(n);
  }
  L.zmax_raw = c.take<float>(n);
  L.zmax_wet = c.take<float>(n);
  L.arrival = c.take<int>(n);
  if (L.nl) L.dzp = c.take<float>(n);
  L.xflux = c.take<float>(2 * (size_t)nx);
  L.yflux = c.take<float>(2 * (size_t)ny);
  L.coef = c.take<float>((size_t)kNumCoef * ny);
  HIPCHK(hipMemsetAsync(L.block, 0, L.block_bytes, st));
  HIPCHK(hipMemsetAsync(L.arrival, 0xff, n * sizeof(int), st));  // -1 = not arrived
  HIPCHK(hipMemcpyAsync(L.h, H, n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(L.z[0], Z, n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(L.zmax_raw, Z, n * sizeof(float), hipMemcpyHostToDevice, st));
  L.cur = 0;
  if (id != 1) T.layers[parent].children.push_back(id);
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int tree_upload_coef(LayerRec &L, const std::vector<float> &coef) {
  if (coef.size() != (size_t)kNumCoef * L.ny) return fail(9, "coefficient table size");
  HIPCHK(hipMemcpy(L.coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

int tree_alloc_r24(LayerRec &L, const float *R2, const float *R4) {
  const size_t n = L.nxy();
  HIPCHK(hipMalloc(&L.r24, 2 * n * sizeof(float)));
  HIPCHK(hipMemcpy(L.r24, R2, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.r24 + n, R4, n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------- device views
int g_store_policy = 1;

bool streams_planes(const LayerRec &L) {
  if (g_store_policy != 1) return g_store_policy == 2;
  const long long cells = (long long)L.nx * L.ny;
  return cells > kStreamCells || cells <= kL2Cells;
}

DevLayer dev_view(const Tree &T, const LayerRec &L, bool with_max, bool advance) {
  DevLayer d{};
  d.nx = L.nx;
  d.ny = L.ny;
  d.is_root = L.id == 1;
  d.modscm = L.modscm;
  d.linchk = L.linchk;
  d.fric = L.fric;
  d.rx = L.rx;
  d.ry = L.ry;
  d.grx = L.grx;
  d.gry = L.gry;
  d.dt = L.dt;
  d.kfric = L.kfric;
  d.h = L.h;
  d.coef = L.coef;
  d.r24 = L.r24;
  d.z1 = L.z[L.cur];
  d.m1 = L.m[L.cur];
  d.n1 = L.n[L.cur];
  d.z2 = L.z[1 - L.cur];
  d.m2 = L.m[1 - L.cur];
  d.n2 = L.n[1 - L.cur];
  d.dzp = L.dzp;
  d.zmax = with_max ? L.zmax_wet : nullptr;
  d.arrival = L.arrival;
  d.ctl = T.ctl;
  d.ctl_in = T.ctl_in;
  d.advance = advance ? 1 : 0;
  d.arr_thr = T.arr_thr;
  d.zraw = (T.raw_max && with_max && advance && L.id == 1) ? L.zmax_raw : nullptr;
  d.sponge_cx = (L.sponge && L.id == 1) ? L.sponge : nullptr;
  d.sponge_cy = d.sponge_cx ? L.sponge + L.nx : nullptr;
  d.stream_out = streams_planes(L) ? 1 : 0;
  return d;
}

DevNest nest_view(const LayerRec &P, const LayerRec &C) {
  DevNest q{};
  q.pnx = P.nx;
  q.pny = P.ny;
  q.p_sph = P.sph;
  q.p_nl = P.nl;
  q.p_r24 = P.r24 != nullptr;
  q.ph = P.h;
  q.pz2 = P.z[1 - P.cur];
  q.pm1 = P.m[P.cur];
  q.pn1 = P.n[P.cur];
  q.pcoef = P.coef;
  q.pr24 = P.r24;
  q.pgrx = P.grx;
  q.pgry = P.gry;
  q.xflux = P.xflux;
  q.yflux = P.yflux;
  q.cnx = C.nx;
  q.cny = C.ny;
  q.rel_size = C.rel_size;
  q.rel_time = C.rel_time;
  q.is = C.corners[0];
  q.ie = C.corners[1];
  q.js = C.corners[2];
  q.je = C.corners[3];
  q.ch = C.h;
  q.cz1 = C.z[C.cur];
  q.cz2 = C.z[1 - C.cur];
  q.cm1 = C.m[C.cur];
  q.cn1 = C.n[C.cur];
  q.pz2_out = P.z[1 - P.cur];
  return q;
}

int tree_launch(const Tree &T, const LayerRec &L, int mode, bool with_max, bool advance,
                hipStream_t st) {
  DevLayer d = dev_view(T, L, with_max, advance);
  const int v = mode == kFused && L.variant >= 0 ? L.variant : T.variant;
  hipError_t e = launch_step(d, mode, L.nl != 0, L.sph != 0, v, st);
  if (e != hipSuccess) return fail(200 + (int)e, "step kernel launch failed: %s", hipGetErrorString(e));
  return 0;
}

// ALL_GRID (all_grids.f90:34-308) / subgrid_solver_rcsv (GPUall_grids.cu:655-699)
// The SC_OPTION = 1 coupling views of parent P and child C.
DevSc sc_view(const Tree &T, const LayerRec &P, const LayerRec &C) {
  DevSc S{};
  S.pnx = P.nx;
  S.pny = P.ny;
  S.p_nl = P.nl;
  S.p_modscm = P.modscm;
  S.p_r24 = P.r24 != nullptr;
  S.ph = P.h;
  S.pz2 = P.z[1 - P.cur];
  S.pm1 = P.m[P.cur];
  S.pn1 = P.n[P.cur];
  S.pcoef = P.coef;
  S.pr24 = P.r24;
  S.pm2 = P.m[1 - P.cur];
  S.pn2 = P.n[1 - P.cur];
  S.pz2_out = P.z[1 - P.cur];
  S.cnx = C.nx;
  S.cny = C.ny;
  S.rel_time = C.rel_time;
  S.half = (float)(C.rel_size * C.rel_size) / 2.0f;
  S.pos = C.sc_pos;
  S.cxy = C.sc_cxy;
  S.ch = C.h;
  S.cz_now = C.z[C.cur];
  S.cz2 = C.z[1 - C.cur];
  S.cm1 = C.m[C.cur];
  S.cn1 = C.n[C.cur];
  S.steps_done = T.ctl + T.ctl_in;
  for (int b = 0; b < 4; ++b)
    for (int q = 0; q < 4; ++q) S.band[b][q] = C.sc_band[b][q];
  for (int q = 0; q < 4; ++q) S.corners[q] = C.corners[q];
  S.jnz_off = C.sc_jnz;
  S.jnz_idx = C.sc_jnz + C.sc_jnz_cells + 1;
  return S;
}

int tree_enqueue_children(Tree &T, LayerRec &P, bool with_max, hipStream_t st) {
  for (int cid : P.children) {
    LayerRec &C = T.layers[cid];
    const int nhalf = (int)std::floor(C.rel_time / 2.0f) + 1;
    for (int k = 1; k <= C.rel_time; ++k) {
      hipError_t e = C.sc ? launch_sc_edges(sc_view(T, P, C), k, st)
                          : launch_jnq(nest_view(P, C), k, st);
      if (e != hipSuccess) return fail(201, "jnq launch failed: %s", hipGetErrorString(e));
      if (C.children.empty()) {
        if (int r = tree_launch(T, C, kFused, with_max, false, st)) return r;
      } else {
        if (int r = tree_launch(T, C, kMassOnly, false, false, st)) return r;
        if (int r = tree_enqueue_children(T, C, with_max, st)) return r;
        if (int r = tree_launch(T, C, kMomentOnly, with_max, false, st)) return r;
      }
      if (k == nhalf) {
        e = C.sc ? launch_sc_jnz(sc_view(T, P, C), st) : launch_jnz(nest_view(P, C), st);
        if (e != hipSuccess) return fail(202, "jnz launch failed: %s", hipGetErrorString(e));
      }
      C.cur = 1 - C.cur;  // UPDATE
    }
  }
  return 0;
}

// comcot.f90:359-439: continuity + open boundary, nested layers, momentum
// with the wet max/arrival, level rotation.
bool tree_fuses_halves(const Tree &T) {
  const LayerRec &R = T.layers[1];
  return T.fuse_halves && R.used && !R.children.empty() && !R.nl;
}

DevLayer fused_halves_view(const Tree &T, const LayerRec &R) {
  DevLayer d = dev_view(T, R, true, true);
  d.z1 = R.z[1 - R.cur];  // Z(k): continuity of step k, JNZ
  d.z2 = R.z[R.cur];      // Z(k-1), which linear momentum does not read
  return d;
}

int tree_enqueue_root_step(Tree &T, hipStream_t st, bool next) {
  LayerRec *R = T.layer(1);
  if (!R) return fail(8, "no root layer");
  if (R->children.empty()) {
    if (int r = tree_launch(T, *R, kFused, true, true, st)) return r;
  } else {
    if (!T.mass_ahead)
      if (int r = tree_launch(T, *R, kMassOnly, false, false, st)) return r;
    T.mass_ahead = false;
    if (int r = tree_enqueue_children(T, *R, true, st)) return r;
    if (next && tree_fuses_halves(T)) {
      const DevLayer d = fused_halves_view(T, *R);
      hipError_t e = launch_step(d, kMomentMass, false, R->sph != 0, T.variant, st);
      if (e != hipSuccess)
        return fail(200 + (int)e, "step kernel launch failed: %s", hipGetErrorString(e));
      T.mass_ahead = true;
    } else {
      if (int r = tree_launch(T, *R, kMomentOnly, true, true, st)) return r;
    }
  }
  R->cur = 1 - R->cur;
  T.ctl_in = 1 - T.ctl_in;
  return tree_enqueue_gauges(T, st);
}

bool tree_step2_eligible(const Tree &T) {
  const LayerRec &R = T.layers[1];
  // gauges sample every step and the sponge acts between steps, so both
  // keep one step per launch
  return T.blocking > 0 && R.used && R.children.empty() && !R.nl && T.ngauges == 0 &&
         !R.sponge;
}

// GET_TS's stencil search over X(IS..IE) (output.f90:1062-1069, 1103-1110)
static int gauge_cell(const std::vector<float> &X, int is, int ie, float x0) {
  int k = 0;
  for (int i = is; i <= ie - 1; ++i) {
    if (x0 >= X[i - 1] && x0 < X[i]) k = i;
    if (x0 < X[is - 1]) k = is;
    if (x0 >= X[ie - 1]) k = ie - 1;
  }
  return k;
}

int tree_set_gauges(Tree &T, hipStream_t st, int n, const float *x0, const float *y0,
                    const int *lid, int capacity) {
  if (n < 0 || (n > 0 && capacity < 1)) return fail(40, "gauges: bad count or capacity");
  std::vector<GaugeDev> g((size_t)n);
  for (int k = 0; k < n; ++k) {
    const LayerRec *L = T.layer(lid[k]);
    if (!L) return fail(41, "gauge %d: no layer %d", k, lid[k]);
    const auto &X = T.xs[lid[k]], &Y = T.ys[lid[k]];
    if ((int)X.size() != L->nx || (int)Y.size() != L->ny)
      return fail(42, "gauge %d: coordinates of layer %d not set", k, lid[k]);
    const int is = lid[k] == 1 ? 1 : 2;  // inner layers skip the ghost row/column
    const int ki = gauge_cell(X, is, L->nx, x0[k]), kj = gauge_cell(Y, is, L->ny, y0[k]);
    if (ki < 1 || kj < 1) return fail(43, "gauge %d: position not comparable (NaN?)", k);
    const float dxv = X[ki] - X[ki - 1], dyv = Y[kj] - Y[kj - 1];
    g[k].lid = lid[k];
    g[k].ki = ki;
    g[k].kj = kj;
    g[k].nx = L->nx;
    g[k].cx = (x0[k] - X[ki - 1]) / dxv;
    g[k].cy = (y0[k] - Y[kj - 1]) / dyv;
    g[k].tide = 0.0f;  // the layers run at mean sea level (TIDE_LEVEL = 0)
    g[k].pad = 0;
  }
  if (T.gauges) (void)hipFree(T.gauges);
  if (T.ring) (void)hipFree(T.ring);
  T.gauges = nullptr;
  T.ring = nullptr;
  T.ngauges = 0;
  if (n == 0) return 0;
  HIPCHK(hipMalloc(&T.gauges, (size_t)n * sizeof(GaugeDev)));
  HIPCHK(hipMalloc(&T.ring, (size_t)n * capacity * sizeof(float)));
  HIPCHK(hipMemcpy(T.gauges, g.data(), (size_t)n * sizeof(GaugeDev), hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(T.ring, 0, (size_t)n * capacity * sizeof(float), st));
  int done = 0;
  HIPCHK(hipStreamSynchronize(st));
  if (T.ctl) HIPCHK(hipMemcpy(&done, T.ctl + T.ctl_in, sizeof(int), hipMemcpyDeviceToHost));
  T.gauge_step0 = done;
  T.ngauges = n;
  T.gauge_cap = capacity;
  return 0;
}

int tree_enqueue_gauges(const Tree &T, hipStream_t st) {
  if (T.ngauges == 0) return 0;
  ZNow z{};
  for (int id = 1; id < kMaxLayers; ++id)
    if (T.layers[id].used) z.z[id] = T.layers[id].z[T.layers[id].cur];
  hipError_t e = launch_gauges(T.gauges, T.ngauges, z, T.ctl + T.ctl_in, T.ring, T.gauge_cap, st);
  if (e != hipSuccess) return fail(204, "gauge launch failed: %s", hipGetErrorString(e));
  return 0;
}

int tree_enqueue_root_steps2(Tree &T, hipStream_t st) {
  LayerRec *R = T.layer(1);
  if (!R) return fail(8, "no root layer");
  DevLayer d = dev_view(T, *R, true, true);
  hipError_t e = launch_step2(d, R->sph != 0, T.blocking - 1, st);
  if (e != hipSuccess) return fail(200 + (int)e, "two-step launch failed: %s", hipGetErrorString(e));
  R->cur = 1 - R->cur;
  T.ctl_in = 1 - T.ctl_in;
  return 0;
}

static long long launches_below(Tree &T, const LayerRec &P) {
  long long n = 0;
  for (int cid : P.children) {
    const LayerRec &C = T.layers[cid];
    const long long inner = C.children.empty() ? 1 : 2 + launches_below(T, C);
    // JNQ (SC: bands and edges) + step(s) per sub-step, one JNZ
    n += (long long)C.rel_time * ((C.sc ? 2 : 1) + inner) + 1;
  }
  return n;
}

long long tree_launches_per_step(Tree &T) {
  LayerRec *R = T.layer(1);
  if (!R) return 0;
  return R->children.empty() ? 1 : 2 + launches_below(T, *R);
}

bool row_factor(const float *R, const float *H, int nx, int ny, int j, bool is_x, float *F) {
  auto hpq = [&](int i) {  // PQ_DEPTH, 0-based i
    float h = H[(size_t)j * nx + i];
    if (!(h > kGx)) return 0.0f;
    if (is_x) {
      int ip1 = i + 1 >= nx - 1 ? nx - 1 : i + 1;
      return 0.5f * (h + H[(size_t)j * nx + ip1]);
    }
    int jp1 = j + 1 >= ny - 1 ? ny - 1 : j + 1;
    return 0.5f * (h + H[(size_t)jp1 * nx + i]);
  };
  int best = -1;
  float bh = 0.0f;
  for (int i = 0; i < nx; ++i) {
    float v = fmaxf(hpq(i), 0.0f);
    if (v > bh) { bh = v; best = i; }
  }
  auto check = [&](float f) {
    for (int i = 0; i < nx; ++i)
      if (!(f * fmaxf(hpq(i), 0.0f) == R[(size_t)j * nx + i])) return false;
    return true;
  };
  if (best < 0) {  // row entirely dry: R must be all zero
    *F = 0.0f;
    return check(0.0f);
  }
  float f0 = R[(size_t)j * nx + best] / bh;
  float c = f0;
  for (int k = 0; k < 4; ++k) c = nextafterf(c, -INFINITY);
  for (int k = 0; k < 9; ++k, c = nextafterf(c, INFINITY)) {
    if (check(c)) { *F = c; return true; }
  }
  return false;
}

// ------------------------------------------------------------ checkpoints
// File: "GCOMCOTCKPT1", int32 version, int32 layer count, int64 completed
// root steps; per layer an int32 record (id, parent, nx, ny, nl, laycord,
// laygov, rel_size, rel_time, corners[4], cur), the uint64 size of its
// device block and the block's bytes (every plane of the layer: depth,
// both time levels, maxima, arrival, flux caches, row coefficients).
namespace {

constexpr char kCkptMagic[12] = {'G', 'C', 'O', 'M', 'C', 'O', 'T', 'C', 'K', 'P', 'T', '1'};

struct CkptLayer {
  int32_t id, parent, nx, ny, nl, laycord, laygov, rel_size, rel_time, corners[4], cur;
  uint64_t bytes;
};

CkptLayer ckpt_record(const LayerRec &L) {
  CkptLayer c{};
  c.id = L.id;
  c.parent = L.parent;
  c.nx = L.nx;
  c.ny = L.ny;
  c.nl = L.nl;
  c.laycord = L.laycord;
  c.laygov = L.laygov;
  c.rel_size = L.rel_size;
  c.rel_time = L.rel_time;
  for (int k = 0; k < 4; ++k) c.corners[k] = L.corners[k];
  c.cur = L.cur;
  c.bytes = L.block_bytes;
  return c;
}

bool same_layer(const CkptLayer &a, const CkptLayer &b) {
  bool ok = a.id == b.id && a.parent == b.parent && a.nx == b.nx && a.ny == b.ny &&
            a.nl == b.nl && a.laycord == b.laycord && a.laygov == b.laygov &&
            a.rel_size == b.rel_size && a.rel_time == b.rel_time && a.bytes == b.bytes;
  for (int k = 0; k < 4; ++k) ok = ok && a.corners[k] == b.corners[k];
  return ok;
}

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

int tree_checkpoint_save(const Tree &T, hipStream_t st, const char *path, long long done) {
  HIPCHK(hipStreamSynchronize(st));
  File F;
  F.f = std::fopen(path, "wb");
  if (!F.f) return fail(70, "checkpoint: cannot create %s", path);
  int32_t head[2] = {1, 0};
  for (const LayerRec &L : T.layers) head[1] += L.used ? 1 : 0;
  const int64_t steps = done;
  bool ok = std::fwrite(kCkptMagic, 1, sizeof kCkptMagic, F.f) == sizeof kCkptMagic &&
            std::fwrite(head, sizeof head, 1, F.f) == 1 && std::fwrite(&steps, 8, 1, F.f) == 1;
  std::vector<char> buf;
  for (const LayerRec &L : T.layers) {
    if (!ok) break;
    if (!L.used) continue;
    const CkptLayer c = ckpt_record(L);
    buf.resize(L.block_bytes);
    HIPCHK(hipMemcpy(buf.data(), L.block, L.block_bytes, hipMemcpyDeviceToHost));
    ok = std::fwrite(&c, sizeof c, 1, F.f) == 1 &&
         std::fwrite(buf.data(), 1, buf.size(), F.f) == buf.size();
  }
  const bool closed = std::fclose(F.f) == 0;
  F.f = nullptr;
  if (!ok || !closed) return fail(71, "checkpoint: write to %s failed", path);
  return 0;
}

int tree_checkpoint_load(Tree &T, hipStream_t st, const char *path, long long *done) {
  T.mass_ahead = false;  // level 2 is no longer the next continuity
  File F;
  F.f = std::fopen(path, "rb");
  if (!F.f) return fail(72, "checkpoint: cannot open %s", path);
  char magic[sizeof kCkptMagic];
  int32_t head[2];
  int64_t steps = 0;
  if (std::fread(magic, 1, sizeof magic, F.f) != sizeof magic ||
      std::memcmp(magic, kCkptMagic, sizeof magic) != 0 || std::fread(head, sizeof head, 1, F.f) != 1 ||
      head[0] != 1 || std::fread(&steps, 8, 1, F.f) != 1 || steps < 0)
    return fail(73, "checkpoint: %s is not a checkpoint of this library", path);
  int used = 0;
  for (const LayerRec &L : T.layers) used += L.used ? 1 : 0;
  if (head[1] != used) return fail(74, "checkpoint: %d layers in the file, %d here", head[1], used);
  // every record must match a layer here before anything is overwritten
  const long start = std::ftell(F.f);
  std::vector<CkptLayer> recs;
  for (int k = 0; k < head[1]; ++k) {
    CkptLayer c;
    if (std::fread(&c, sizeof c, 1, F.f) != 1) return fail(73, "checkpoint: %s is truncated", path);
    const LayerRec *L = T.layer(c.id);
    if (!L || !same_layer(c, ckpt_record(*L)) || (c.cur != 0 && c.cur != 1))
      return fail(75, "checkpoint: layer %d of %s does not match this layer tree", c.id, path);
    if (std::fseek(F.f, (long)c.bytes, SEEK_CUR) != 0)
      return fail(73, "checkpoint: %s is truncated", path);
    recs.push_back(c);
  }
  // seeking past the end succeeds, so compare the file's size with the records'
  const long need = std::ftell(F.f);
  if (need < 0 || std::fseek(F.f, 0, SEEK_END) != 0 || std::ftell(F.f) < need)
    return fail(73, "checkpoint: %s is truncated", path);
  if (std::fseek(F.f, start, SEEK_SET) != 0) return fail(73, "checkpoint: cannot reread %s", path);
  HIPCHK(hipStreamSynchronize(st));
  std::vector<char> buf;
  for (const CkptLayer &r : recs) {
    CkptLayer c;
    buf.resize(r.bytes);
    if (std::fread(&c, sizeof c, 1, F.f) != 1 || std::fread(buf.data(), 1, buf.size(), F.f) != buf.size())
      return fail(73, "checkpoint: %s is truncated", path);
    LayerRec &L = T.layers[c.id];
    HIPCHK(hipMemcpy(L.block, buf.data(), buf.size(), hipMemcpyHostToDevice));
    L.cur = c.cur;
  }
  const int n32 = (int)steps;
  const int ctl[2] = {n32, n32};
  HIPCHK(hipMemcpy(T.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice));
  if (done) *done = steps;
  return 0;
}

// ------------------------------------------------------------ SC_OPTION 1
int tree_set_sc(Tree &T, hipStream_t st, int id, const int *pos, const float *cxy) {
  LayerRec *C = T.layer(id);
  if (!C || id == 1) return fail(81, "SC: layer %d is not a nested layer", id);
  const LayerRec &P = T.layers[C->parent];
  if (P.id != 1 || P.laycord != 0 || C->laycord != 1)
    return fail(80, "SC: layer %d must be a Cartesian child of the spherical layer 1", id);
  const int nx = C->nx, ny = C->ny;
  const size_t n = C->nxy();
  auto ki = [&](int i, int j) { return pos[(size_t)(j - 1) * nx + (i - 1)]; };
  auto kj = [&](int i, int j) { return pos[n + (size_t)(j - 1) * nx + (i - 1)]; };
  // every cell's parent cell and its +1 neighbours exist, and lie in the extent
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const int a = ki(i, j), b = kj(i, j);
      if (a < C->corners[0] || a > C->corners[1] || b < C->corners[2] || b > C->corners[3] ||
          a + 1 > P.nx || b + 1 > P.ny)
        return fail(82, "SC: layer %d cell (%d,%d) has POS (%d,%d) outside its extent", id, i, j,
                    a, b);
    }
  // NEWQ_SC's bands (all_grids.f90:1131-1271): M around the left and right
  // edges, N around the top and bottom edges
  const int c1 = C->corners[0], c2 = C->corners[1], c3 = C->corners[2], c4 = C->corners[3];
  int band[4][4] = {
      {std::min(ki(1, 1), ki(1, ny)) - 2, std::max(ki(1, 1), ki(1, ny)) + 2, c3 - 2, c4 + 2},
      {std::min(ki(nx, 1), ki(nx, ny)) - 2, std::max(ki(nx, 1), ki(nx, ny)) + 2, c3 - 2, c4 + 2},
      {c1 - 2, c2 + 2, std::min(kj(1, ny), kj(nx, ny)) - 2, std::max(kj(1, ny), kj(nx, ny)) + 2},
      {c1 - 2, c2 + 2, std::min(kj(1, 1), kj(nx, 1)) - 2, std::max(kj(1, 1), kj(nx, 1)) + 2}};
  for (int b = 0; b < 4; ++b)  // the formulas read (I+1, J) for M and (I, J+1) for N unclamped
    if (band[b][0] < 1 || band[b][2] < 1 || band[b][1] + 1 > P.nx || band[b][3] + 1 > P.ny)
      return fail(83, "SC: layer %d: NEWQ_SC band %d leaves the parent", id, b);
  // JNZ_SC's child lists per parent cell of the extent, in the Fortran loop
  // order (I outer, J inner)
  const int w = c2 - c1 + 1;
  const size_t cells = (size_t)w * (c4 - c3 + 1);
  std::vector<int> count(cells, 0), off(cells + 1, 0), idx(n);
  for (int i = 1; i <= nx; ++i)
    for (int j = 1; j <= ny; ++j) ++count[(size_t)(kj(i, j) - c3) * w + (ki(i, j) - c1)];
  for (size_t q = 0; q < cells; ++q) off[q + 1] = off[q] + count[q];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int i = 1; i <= nx; ++i)
    for (int j = 1; j <= ny; ++j)
      idx[(size_t)fill[(size_t)(kj(i, j) - c3) * w + (ki(i, j) - c1)]++] = (j - 1) * nx + (i - 1);
  std::vector<int> csr(off);
  csr.insert(csr.end(), idx.begin(), idx.end());
  if (C->sc_pos) (void)hipFree(C->sc_pos);
  if (C->sc_cxy) (void)hipFree(C->sc_cxy);
  if (C->sc_jnz) (void)hipFree(C->sc_jnz);
  C->sc_pos = nullptr;
  C->sc_cxy = nullptr;
  C->sc_jnz = nullptr;
  C->sc = 0;
  HIPCHK(hipMalloc(&C->sc_pos, 2 * n * sizeof(int)));
  HIPCHK(hipMalloc(&C->sc_cxy, 4 * n * sizeof(float)));
  HIPCHK(hipMalloc(&C->sc_jnz, csr.size() * sizeof(int)));
  HIPCHK(hipMemcpyAsync(C->sc_pos, pos, 2 * n * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(C->sc_cxy, cxy, 4 * n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(C->sc_jnz, csr.data(), csr.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b = 0; b < 4; ++b)
    for (int q = 0; q < 4; ++q) C->sc_band[b][q] = band[b][q];
  C->sc_jnz_cells = cells;
  C->sc = 1;
  return 0;
}

}  // namespace gc
