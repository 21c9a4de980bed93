// Host side of gcomcot_hip: layer registry, the reference's C ABI
// (include/gcomcot_hip.h part 1) and the whole-run device loop (part 2).
//
// Layer table: ids 1..19 as in the reference (GPUHeader.h:51), 1 = the
// outermost layer; children kept in registration order like the
// reference's sibling list (GPUcomcot.cu:157-190), which is also the order
// ALL_GRID visits them (all_grids.f90:34-35).  Each layer owns two time
// levels of Z, M, N on the device; cuda_update_layer_ swaps them.
//
// Device memory per layer: H, Z x2, M x2, N x2, raw max, wet max, arrival
// (11 planes of 4 B = 44 B/cell; +4 for the DZ plane of nonlinear
// layers), the per-row coefficient table and the interface flux cache.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "comcot_consts.h"
#include "gcomcot_hip_internal.h"
#include "kernels.h"

using namespace gc;

namespace {

constexpr int kMaxLayers = 20;

struct LayerRec {
  bool used = false;
  int id = 0, parent = -1, level = 1;
  int laycord = 0, laygov = 0, nl = 0, sph = 1;
  int rel_size = 1, rel_time = 1, corners[4] = {0, 0, 0, 0};
  int fric = 0, modscm = 0, linchk = 1;
  float fric_coef = 0, dt = 0, dx = 0, dy = 0, rx = 0, ry = 0, grx = 0, gry = 0, kfric = 0;
  int nx = 0, ny = 0;
  float *h = nullptr, *z[2] = {nullptr, nullptr}, *m[2] = {nullptr, nullptr},
        *n[2] = {nullptr, nullptr};
  float *zmax_raw = nullptr, *zmax_wet = nullptr, *coef = nullptr, *r24 = nullptr;
  float *dzp = nullptr, *xflux = nullptr, *yflux = nullptr;
  int *arrival = nullptr;
  int cur = 0;  // buffer index holding time level 1
  std::vector<int> children;
  size_t nxy() const { return (size_t)nx * ny; }
};

LayerRec g_layers[kMaxLayers];
bool g_init = false;
hipStream_t g_stream = nullptr;
int *g_ctl = nullptr;   // device int[2]: root-step counters
int g_ctl_in = 0;       // which counter holds the completed step count
int g_fatal = 1;
int g_last_code = 0;
char g_last_msg[512] = {0};
int g_variant = 0;
float g_arr_thr = 0.01f;

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

#define HIPCHK(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(100 + (int)e_, "%s at %s:%d: %s", #call, __FILE__, __LINE__,  \
                  hipGetErrorString(e_));                                        \
  } while (0)

int ensure_init() {
  if (g_init) return 0;
  HIPCHK(hipSetDevice(0));
  HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&g_ctl, 2 * sizeof(int)));
  HIPCHK(hipMemset(g_ctl, 0, 2 * sizeof(int)));
  g_ctl_in = 0;
  g_init = true;
  return 0;
}

LayerRec *layer(int id) {
  if (id < 1 || id >= kMaxLayers || !g_layers[id].used) return nullptr;
  return &g_layers[id];
}

void free_layer(LayerRec &L) {
  float *ps[] = {L.h, L.z[0], L.z[1], L.m[0], L.m[1], L.n[0], L.n[1], L.zmax_raw,
                 L.zmax_wet, L.coef, L.r24, L.dzp, L.xflux, L.yflux};
  for (float *p : ps)
    if (p) (void)hipFree(p);
  if (L.arrival) (void)hipFree(L.arrival);
  L = LayerRec();
}

int dalloc(float **p, size_t n) {
  HIPCHK(hipMalloc(p, n * sizeof(float)));
  HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(float), g_stream));
  return 0;
}

// Common part of both init entry points: table entry, device planes, H, Z.
int create_layer(int id, int parent, int level, int laycord, int laygov, int rel_size,
                 int rel_time, const int *corners, int fric_switch, float fric_coef, float dt,
                 float dx, float dy, int nx, int ny, const float *H, const float *Z) {
  if (int e = ensure_init()) return e;
  if (id < 1 || id >= kMaxLayers) return fail(1, "invalid layer id %d (1..%d)", id, kMaxLayers - 1);
  if (nx < 3 || ny < 3) return fail(2, "layer %d: grid %dx%d too small", id, nx, ny);
  if (id == 1 && parent != -1) return fail(3, "layer 1 must be the root (parent -1)");
  if (id != 1) {
    LayerRec *P = layer(parent);
    if (!P) return fail(4, "layer %d: parent %d not initialised", id, parent);
    const int is = corners[0], ie = corners[1], js = corners[2], je = corners[3];
    if (is < 4 || js < 4 || ie > P->nx - 2 || je > P->ny - 2 || ie < is || je < js)
      return fail(5, "layer %d: footprint I %d..%d J %d..%d must lie in 4..NX-2, 4..NY-2 of its parent",
                  id, is, ie, js, je);
    if (rel_size < 1 || nx != (ie - is + 1) * rel_size + 1 || ny != (je - js + 1) * rel_size + 1)
      return fail(6, "layer %d: %dx%d does not match footprint x REL_SIZE %d + 1", id, nx, ny,
                  rel_size);
    if (rel_time < 1) return fail(7, "layer %d: REL_TIME %d < 1", id, rel_time);
  }
  if (g_layers[id].used) {
    if (id == 1) {  // a new root starts a new layer tree
      for (int k = 1; k < kMaxLayers; ++k)
        if (g_layers[k].used) free_layer(g_layers[k]);
    } else {
      if (LayerRec *op = layer(g_layers[id].parent)) {
        auto &ch = op->children;
        for (size_t q = 0; q < ch.size(); ++q)
          if (ch[q] == id) { ch.erase(ch.begin() + q); break; }
      }
      free_layer(g_layers[id]);
    }
  }
  if (id == 1) g_ctl_in = 0, (void)hipMemset(g_ctl, 0, 2 * sizeof(int));
  LayerRec &L = g_layers[id];
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
  if (int e = dalloc(&L.h, n)) return e;
  for (int k = 0; k < 2; ++k) {
    if (int e = dalloc(&L.z[k], n)) return e;
    if (int e = dalloc(&L.m[k], n)) return e;
    if (int e = dalloc(&L.n[k], n)) return e;
  }
  if (int e = dalloc(&L.zmax_raw, n)) return e;
  if (int e = dalloc(&L.zmax_wet, n)) return e;
  if (L.nl)
    if (int e = dalloc(&L.dzp, n)) return e;
  if (int e = dalloc(&L.xflux, 2 * (size_t)nx)) return e;
  if (int e = dalloc(&L.yflux, 2 * (size_t)ny)) return e;
  HIPCHK(hipMalloc(&L.arrival, n * sizeof(int)));
  HIPCHK(hipMemsetAsync(L.arrival, 0xff, n * sizeof(int), g_stream));  // -1
  HIPCHK(hipMemcpyAsync(L.h, H, n * sizeof(float), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(L.z[0], Z, n * sizeof(float), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(L.zmax_raw, Z, n * sizeof(float), hipMemcpyHostToDevice, g_stream));
  L.cur = 0;
  if (id != 1) {
    LayerRec *P = layer(parent);
    P->children.push_back(id);
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  return 0;
}

int upload_coef(LayerRec &L, const std::vector<float> &coef) {
  if (int e = dalloc(&L.coef, coef.size())) return e;
  HIPCHK(hipMemcpy(L.coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------- device views
DevLayer dev_view(const LayerRec &L, bool with_max, bool advance) {
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
  d.ctl = g_ctl;
  d.ctl_in = g_ctl_in;
  d.advance = advance ? 1 : 0;
  d.arr_thr = g_arr_thr;
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
  return q;
}

int launch(const LayerRec &L, int mode, bool with_max, bool advance) {
  DevLayer d = dev_view(L, with_max, advance);
  hipError_t e = launch_step(d, mode, L.nl != 0, L.sph != 0, g_variant, g_stream);
  if (e != hipSuccess) return fail(200 + (int)e, "step kernel launch failed: %s", hipGetErrorString(e));
  return 0;
}

// Nested sub-steps of every child of P for one step of P: ALL_GRID
// (all_grids.f90:34-308) / subgrid_solver_rcsv (GPUall_grids.cu:655-699).
int enqueue_children(LayerRec &P, bool with_max) {
  for (int cid : P.children) {
    LayerRec &C = g_layers[cid];
    const int nhalf = (int)std::floor(C.rel_time / 2.0f) + 1;
    for (int k = 1; k <= C.rel_time; ++k) {
      hipError_t e = launch_jnq(nest_view(P, C), k, g_stream);
      if (e != hipSuccess) return fail(201, "jnq launch failed: %s", hipGetErrorString(e));
      if (C.children.empty()) {
        if (int r = launch(C, kFused, with_max, false)) return r;
      } else {
        if (int r = launch(C, kMassOnly, false, false)) return r;
        if (int r = enqueue_children(C, with_max)) return r;
        if (int r = launch(C, kMomentOnly, with_max, false)) return r;
      }
      if (k == nhalf) {
        DevNest q = nest_view(P, C);
        e = launch_jnz(q, P.z[1 - P.cur], g_stream);
        if (e != hipSuccess) return fail(202, "jnz launch failed: %s", hipGetErrorString(e));
      }
      C.cur = 1 - C.cur;  // UPDATE
    }
  }
  return 0;
}

// One complete root step: continuity + open boundary, nested layers,
// momentum with the wet max/arrival, level rotation (comcot.f90:359-439).
int enqueue_root_step() {
  LayerRec *R = layer(1);
  if (!R) return fail(8, "no root layer");
  if (R->children.empty()) {
    if (int r = launch(*R, kFused, true, true)) return r;
  } else {
    if (int r = launch(*R, kMassOnly, false, false)) return r;
    if (int r = enqueue_children(*R, true)) return r;
    if (int r = launch(*R, kMomentOnly, true, true)) return r;
  }
  R->cur = 1 - R->cur;
  g_ctl_in = 1 - g_ctl_in;
  return 0;
}

int sync() {
  HIPCHK(hipStreamSynchronize(g_stream));
  return 0;
}

// R2/R4 of the reference ABI: row factor F with R(I,J) == F*max(HP(I,J),0)
// for every I of the row (PARA, comcot.f90:733-744).  Returns false when no
// factor reproduces the row bit for bit.
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

}  // namespace

extern "C" {

// ------------------------------------------------------------- reference
void gcomcot_init_gpu_(void) {
  if (ensure_init()) return;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) == hipSuccess)
    printf("GPU INFORMATIONS:                      %s (%s), %d CUs, wavefront %d\n", p.name,
           p.gcnArchName, p.multiProcessorCount, p.warpSize);
}

void gcomcot_init_layer_(int *layerid, int *parent, int *level, int *rel_size, int *rel_time,
                         int *corners, float *grx, float *gry, float *R1, float *R2, float *R3,
                         float *R4, float *R5, float *R6, float *R11, float *H, float *Z,
                         int *nx, int *ny) {
  (void)grx;
  (void)gry;
  const int NX = *nx, NY = *ny;
  if (create_layer(*layerid, *parent, *level, 0, 0, *rel_size, *rel_time, corners, 1, 0.0f,
                   0.0f, 1.0f, 1.0f, NX, NY, H, Z))
    return;
  LayerRec &L = g_layers[*layerid];
  std::vector<float> coef((size_t)kNumCoef * NY, 0.0f);
  bool planes = false;
  for (int j = 0; j < NY; ++j) {
    const size_t c0 = (size_t)j * NX;  // column I=1 as in GPUcomcot.cu:62-68
    coef[kR1 * NY + j] = R1[c0];
    coef[kR11 * NY + j] = R11[c0];
    coef[kR3 * NY + j] = R3[c0];
    coef[kR5 * NY + j] = R5[c0];
    coef[kR6 * NY + j] = R6[j];
    float f2 = 0, f4 = 0;
    if (!row_factor(R2, H, NX, NY, j, true, &f2) || !row_factor(R4, H, NX, NY, j, false, &f4))
      planes = true;
    coef[kF2 * NY + j] = f2;
    coef[kF4 * NY + j] = f4;
  }
  if (upload_coef(L, coef)) return;
  if (planes) {
    if (dalloc(&L.r24, 2 * L.nxy())) return;
    if (hipMemcpy(L.r24, R2, L.nxy() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.r24 + L.nxy(), R4, L.nxy() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess) {
      fail(9, "R2/R4 upload failed");
      return;
    }
  }
}

int gcomcot_init_layer_ex_(int *layerid, int *parent, int *level, int *laycord, int *laygov,
                           int *rel_size, int *rel_time, int *corners, int *fric_switch,
                           float *fric_coef, float *dt, float *dx, float *dy, float *coef,
                           float *H, float *Z, int *nx, int *ny) {
  if (*laycord != 0 && *laycord != 1) return fail(10, "LAYCORD %d not 0/1", *laycord);
  if (!(*laygov == 0 || *laygov == 1 || *laygov == 9 || (*laycord == 1 && *laygov == 5)))
    return fail(11, "LAYGOV %d not supported (dispersive forms 2/3 are out of scope)", *laygov);
  if (*fric_switch != 0 && *fric_switch != 1)
    return fail(12, "FRIC_SWITCH %d not supported (0 constant Manning, 1 none)", *fric_switch);
  if (*dt <= 0.0f || *dx <= 0.0f || *dy <= 0.0f) return fail(13, "DT/DX/DY must be positive");
  if (int e = create_layer(*layerid, *parent, *level, *laycord, *laygov, *rel_size, *rel_time,
                           corners, *fric_switch, *fric_coef, *dt, *dx, *dy, *nx, *ny, H, Z))
    return e;
  LayerRec &L = g_layers[*layerid];
  std::vector<float> c((size_t)kNumCoef * *ny, 0.0f);
  if (*laycord == 0) {
    if (!coef) return fail(14, "spherical layer %d needs the PARA row table", *layerid);
    std::memcpy(c.data(), coef, c.size() * sizeof(float));
  }
  return upload_coef(L, c);
}

void gcomcot_show_grids_(void) {
  for (int id = 1; id < kMaxLayers; ++id) {
    const LayerRec &L = g_layers[id];
    if (!L.used) continue;
    printf("lid = %d\nparent = %d\nlevel = %d\nsize = %d x %d\n", L.id, L.parent, L.level, L.nx,
           L.ny);
  }
}

void mass_launch_(const float *Z1, float *Z2, const float *H, const int *lid) {
  (void)Z1; (void)Z2; (void)H;
  LayerRec *L = layer(*lid);
  if (!L) { fail(15, "mass_launch_: no layer %d", *lid); return; }
  if (launch(*L, kMassOnly, false, false)) return;
  sync();
}

void openbd_launch_(float *Z2) { (void)Z2; }

void momt_launch_(float *M2, float *N2, float *Z2, const int *lid) {
  (void)M2; (void)N2; (void)Z2;
  LayerRec *L = layer(*lid);
  if (!L) { fail(16, "momt_launch_: no layer %d", *lid); return; }
  if (launch(*L, kMomentOnly, false, false)) return;
  sync();
}

void all_grid_launch_(void) {
  LayerRec *R = layer(1);
  if (!R) { fail(8, "all_grid_launch_: no root layer"); return; }
  if (enqueue_children(*R, true)) return;
  sync();
}

void maxamp_launch_(const int *lid) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(17, "maxamp_launch_: no layer %d", *lid); return; }
  const bool root = L->id == 1;
  DevLayer d = dev_view(*L, true, root);
  hipError_t e = launch_max(d, L->zmax_raw, g_stream);
  if (e != hipSuccess) { fail(203, "max launch failed: %s", hipGetErrorString(e)); return; }
  if (root) g_ctl_in = 1 - g_ctl_in;
  sync();
}

void cuda_update_layer_(int *lid) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(18, "cuda_update_layer_: no layer %d", *lid); return; }
  L->cur = 1 - L->cur;
}

static void copy_plane(float *host, float *dev, size_t n, bool to_dev) {
  hipError_t e = to_dev ? hipMemcpy(dev, host, n * sizeof(float), hipMemcpyHostToDevice)
                        : hipMemcpy(host, dev, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) fail(204, "plane copy failed: %s", hipGetErrorString(e));
}

#define GC_ACCESSOR(name, expr, to_dev)                         \
  void name(float *Z, int *lid) {                               \
    LayerRec *L = layer(*lid);                                  \
    if (!L) { fail(19, #name ": no layer %d", *lid); return; } \
    if (sync()) return;                                         \
    copy_plane(Z, expr, L->nxy(), to_dev);                      \
  }
GC_ACCESSOR(cuda_getz_, L->z[1 - L->cur], false)
GC_ACCESSOR(cuda_setz_, L->z[1 - L->cur], true)
GC_ACCESSOR(cuda_getz1_, L->z[L->cur], false)
GC_ACCESSOR(cuda_setz1_, L->z[L->cur], true)
GC_ACCESSOR(cuda_getzmax_, L->zmax_raw, false)
#undef GC_ACCESSOR

#define GC_ACCESSOR2(name, lev, to_dev)                             \
  void name(float *M, float *N, int *lid) {                         \
    LayerRec *L = layer(*lid);                                      \
    if (!L) { fail(19, #name ": no layer %d", *lid); return; }     \
    if (sync()) return;                                             \
    const int b = lev ? L->cur : 1 - L->cur;                        \
    copy_plane(M, L->m[b], L->nxy(), to_dev);                       \
    copy_plane(N, L->n[b], L->nxy(), to_dev);                       \
  }
GC_ACCESSOR2(cuda_getmn_, 0, false)
GC_ACCESSOR2(cuda_setmn_, 0, true)
GC_ACCESSOR2(cuda_getmn1_, 1, false)
GC_ACCESSOR2(cuda_setmn1_, 1, true)
#undef GC_ACCESSOR2

void gcomcot_get_z_(int *lid, float *l1, float *l2, float *l3, float *l4, int *i, int *j) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(20, "gcomcot_get_z_: no layer %d", *lid); return; }
  if (*i < 1 || *i >= L->nx || *j < 1 || *j >= L->ny) {
    fail(21, "gcomcot_get_z_: (%d,%d) outside 1..%d x 1..%d", *i, *j, L->nx - 1, L->ny - 1);
    return;
  }
  if (sync()) return;
  float t[2];
  const float *z = L->z[1 - L->cur];
  if (hipMemcpy(t, z + (size_t)(*j - 1) * L->nx + (*i - 1), 2 * sizeof(float),
                hipMemcpyDeviceToHost) != hipSuccess) { fail(22, "gauge copy failed"); return; }
  *l1 = t[0];
  *l2 = t[1];
  if (hipMemcpy(t, z + (size_t)(*j) * L->nx + (*i - 1), 2 * sizeof(float),
                hipMemcpyDeviceToHost) != hipSuccess) { fail(22, "gauge copy failed"); return; }
  *l3 = t[0];
  *l4 = t[1];
}

void cuda_shutdown_(void) {
  if (!g_init) return;
  (void)hipStreamSynchronize(g_stream);
  for (int id = 0; id < kMaxLayers; ++id)
    if (g_layers[id].used) free_layer(g_layers[id]);
  (void)hipFree(g_ctl);
  g_ctl = nullptr;
  (void)hipStreamDestroy(g_stream);
  g_stream = nullptr;
  g_init = false;
}

// ------------------------------------------------------------- extended
int gcomcot_run_steps_(int *nsteps, int *use_graph) {
  if (int e = ensure_init()) return e;
  if (!layer(1)) return fail(8, "no root layer");
  int n = *nsteps;
  if (n < 0) return fail(23, "negative step count");
  if (*use_graph && n >= 2) {
    // Two root steps bring every layer's level parity back to where it
    // started (children with odd REL_TIME flip once per root step), so a
    // captured pair can be replayed.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    int r = enqueue_root_step();
    if (!r) r = enqueue_root_step();
    hipError_t ce = hipStreamEndCapture(g_stream, &graph);
    if (r) return r;
    HIPCHK(ce);
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (int k = 0; k + 1 < n; k += 2) HIPCHK(hipGraphLaunch(exec, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    n = n % 2;
  }
  for (int k = 0; k < n; ++k)
    if (int r = enqueue_root_step()) return r;
  return sync();
}

int gcomcot_time_steps_(int *nsteps, float *ms_total, float *ms_each) {
  if (int e = ensure_init()) return e;
  if (!layer(1)) return fail(8, "no root layer");
  const int n = *nsteps;
  if (n < 1) return fail(23, "step count must be >= 1");
  std::vector<hipEvent_t> ev(n + 1);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], g_stream));
  for (int k = 0; k < n; ++k) {
    if (int r = enqueue_root_step()) return r;
    HIPCHK(hipEventRecord(ev[k + 1], g_stream));
  }
  HIPCHK(hipEventSynchronize(ev[n]));
  HIPCHK(hipEventElapsedTime(ms_total, ev[0], ev[n]));
  if (ms_each)
    for (int k = 0; k < n; ++k) HIPCHK(hipEventElapsedTime(&ms_each[k], ev[k], ev[k + 1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  return 0;
}

void gcomcot_get_max_(int *lid, float *zmax_wet, int *arrival) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(24, "gcomcot_get_max_: no layer %d", *lid); return; }
  if (sync()) return;
  copy_plane(zmax_wet, L->zmax_wet, L->nxy(), false);
  if (hipMemcpy(arrival, L->arrival, L->nxy() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    fail(25, "arrival copy failed");
}

void gcomcot_set_arrival_threshold_(float *thr) { g_arr_thr = *thr; }

void gcomcot_get_flux_cache_(int *lid, float *xflux, float *yflux) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(26, "gcomcot_get_flux_cache_: no layer %d", *lid); return; }
  if (sync()) return;
  copy_plane(xflux, L->xflux, 2 * (size_t)L->nx, false);
  copy_plane(yflux, L->yflux, 2 * (size_t)L->ny, false);
}

int gcomcot_layer_dims_(int *lid, int *nx, int *ny) {
  LayerRec *L = layer(*lid);
  if (!L) return 1;
  *nx = L->nx;
  *ny = L->ny;
  return 0;
}

void gcomcot_set_error_mode_(int *fatal) { g_fatal = *fatal ? 1 : 0; }

int gcomcot_last_error_(char *buf, int *len) {
  if (buf && len && *len > 0) {
    std::strncpy(buf, g_last_msg, (size_t)*len - 1);
    buf[*len - 1] = 0;
  }
  int c = g_last_code;
  g_last_code = 0;
  g_last_msg[0] = 0;
  return c;
}

int gcomcot_set_tile_variant_(int *variant) {
  if (*variant < 0 || *variant >= num_tile_variants()) return 1;
  g_variant = *variant;
  return 0;
}

int gcomcot_tile_variants_(int *count, int *tx, int *ty) {
  *count = num_tile_variants();
  if (tx && ty)
    for (int v = 0; v < *count; ++v) tile_variant_shape(v, &tx[v], &ty[v]);
  return g_variant;
}

}  // extern "C"
