// The reference's C ABI (include/gcomcot_hip.h part 1) and the
// whole-run device loop (part 2) over one layer tree, the reference's
// global layer table (GPUcomcot.cu:5-22): ids 1..19, 1 = the outermost
// layer (GPUHeader.h:51).  Every call returns after the device work it
// issued has completed, as the reference's calls do.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "comcot_consts.h"
#include "gcomcot_hip_internal.h"
#include "hotstart.h"
#include "kernels.h"
#include "output.h"
#include "source.h"
#include "tree.h"

using namespace gc;

namespace {

Tree g_tree;
bool g_init = false;
// Faults waiting for their rupture time in the device loop.
struct PendingFault {
  float params[11];
  int fswitch;
  float t0;
};
std::vector<PendingFault> g_faults;
hipStream_t g_stream = nullptr;

int ensure_init() {
  if (g_init) return 0;
  HIPCHK(hipSetDevice(0));
  HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_init = true;
  return 0;
}

LayerRec *layer(int id) { return g_tree.layer(id); }

int sync() {
  HIPCHK(hipStreamSynchronize(g_stream));
  return 0;
}

int launch(const LayerRec &L, int mode, bool with_max, bool advance) {
  return tree_launch(g_tree, L, mode, with_max, advance, g_stream);
}

int create_layer(int id, int parent, int level, int laycord, int laygov, int rel_size,
                 int rel_time, const int *corners, int fric_switch, float fric_coef, float dt,
                 float dx, float dy, int nx, int ny, const float *H, const float *Z) {
  if (int e = tree_validate_layer(g_tree, id, parent, rel_size, rel_time, corners, nx, ny))
    return e;
  if (int e = ensure_init()) return e;
  if (id == 1 && out_active())  // a new tree: queued files are finished first
    if (int e = out_disable(g_tree)) return e;
  if (id == 1) g_faults.clear();
  return tree_create_layer(g_tree, g_stream, id, parent, level, laycord, laygov, rel_size,
                           rel_time, corners, fric_switch, fric_coef, dt, dx, dy, nx, ny, H, Z);
}

// The device half of gcomcot_detmath_eval_ and of its no-allocation
// variant: x[] (host) into d[0, n), the kernel writes d[n, 2n) and d[2n, 3n),
// both copied out; returns after the copies have completed.
int detmath_eval_staged(float *d, long long n, const float *x, float *pow_out,
                        float *exp_out) {
  const size_t bytes = (size_t)n * sizeof(float);
  hipError_t e = hipMemcpyAsync(d, x, bytes, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = launch_detmath_eval(d, n, d + n, d + 2 * n, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pow_out, d + n, bytes, hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(exp_out, d + 2 * n, bytes, hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) return fail(207, "detmath eval failed: %s", hipGetErrorString(e));
  return 0;
}

// Staging memory of gcomcot_detmath_eval_noalloc_: arguments, powers and
// exponentials of up to kNoAllocMaxArgs values (12 MiB), a static device
// array, so the entry point needs no allocation.  Each device holds its own
// copy for the life of the process, at an address hipGetSymbolAddress gives
// with that device current; host code uses that address only (the array's
// name in host code does not denote device memory).  Every call on a device
// shares its copy: calls that overlap on one device race.
constexpr long long kNoAllocMaxArgs = 1LL << 20;
__device__ __align__(256) float scratch[3 * kNoAllocMaxArgs];
constexpr int kMaxDevices = 64;
float *g_scratch[kMaxDevices] = {};

// This device's scratch address, looked up on its first use.
int noalloc_scratch(float **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(208, "device %d beyond the scratch table", dev);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(208, "scratch address of device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = static_cast<float *>(p);
  }
  *out = g_scratch[dev];
  return 0;
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
  LayerRec &L = g_tree.layers[*layerid];
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
  if (tree_upload_coef(L, coef)) return;
  if (planes) tree_alloc_r24(L, R2, R4);
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
  if (*laycord == 0 && !coef)
    return fail(14, "spherical layer %d needs the PARA row table", *layerid);
  if (int e = create_layer(*layerid, *parent, *level, *laycord, *laygov, *rel_size, *rel_time,
                           corners, *fric_switch, *fric_coef, *dt, *dx, *dy, *nx, *ny, H, Z))
    return e;
  LayerRec &L = g_tree.layers[*layerid];
  std::vector<float> c((size_t)kNumCoef * *ny, 0.0f);
  if (*laycord == 0) std::memcpy(c.data(), coef, c.size() * sizeof(float));
  return tree_upload_coef(L, c);
}

int gcomcot_init_layer_sc_(int *layerid, int *parent, int *level, int *laygov, int *rel_size,
                           int *rel_time, int *corners, int *fric_switch, float *fric_coef,
                           float *dt, float *dx, float *dy, float *H, float *Z, int *nx, int *ny,
                           int *pos, float *cxy) {
  if (!(*laygov == 0 || *laygov == 1 || *laygov == 5 || *laygov == 9))
    return fail(11, "LAYGOV %d not supported (dispersive forms 2/3 are out of scope)", *laygov);
  if (*fric_switch != 0 && *fric_switch != 1)
    return fail(12, "FRIC_SWITCH %d not supported (0 constant Manning, 1 none)", *fric_switch);
  if (*dt <= 0.0f || *dx <= 0.0f || *dy <= 0.0f) return fail(13, "DT/DX/DY must be positive");
  if (!pos || !cxy) return fail(84, "SC: layer %d needs POS and CXY", *layerid);
  if (int e = tree_validate_layer(g_tree, *layerid, *parent, *rel_size, *rel_time, corners, *nx,
                                  *ny, true))
    return e;
  if (int e = ensure_init()) return e;
  if (int e = tree_create_layer(g_tree, g_stream, *layerid, *parent, *level, 1, *laygov,
                                *rel_size, *rel_time, corners, *fric_switch, *fric_coef, *dt,
                                *dx, *dy, *nx, *ny, H, Z, true))
    return e;
  LayerRec &L = g_tree.layers[*layerid];
  std::vector<float> c((size_t)kNumCoef * *ny, 0.0f);
  int e = tree_upload_coef(L, c);
  if (!e) e = tree_set_sc(g_tree, g_stream, *layerid, pos, cxy);
  // a layer without its SC geometry must not stay in the tree: its sizes
  // are not those of a traditional nest on its corners
  if (e) tree_drop_layer(g_tree, *layerid);
  return e;
}

void gcomcot_show_grids_(void) {
  for (int id = 1; id < kMaxLayers; ++id) {
    const LayerRec &L = g_tree.layers[id];
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
  if (tree_enqueue_children(g_tree, *R, true, g_stream)) return;
  sync();
}

static void copy_plane(float *host, float *dev, size_t n, bool to_dev) {
  hipError_t e = to_dev ? hipMemcpy(dev, host, n * sizeof(float), hipMemcpyHostToDevice)
                        : hipMemcpy(host, dev, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) fail(204, "plane copy failed: %s", hipGetErrorString(e));
}

// ---- the reference's differential checks of the nest exchange
// (GPUall_grids.cu:299-394, 626-643), which its Fortran calls under
// DEBUG_CORE / DEBUG_ALL_GRID right after computing the same routine on the
// host from the state that SET_ALL gave the device (all_grids.f90:363-366,
// 416-418, 474-490, 638-650): the device runs its own JNQ/NEWQ or JNZ on
// that state, and every value that differs from the host's by 1e-4 or more
// (assert_diff with debug_option.h's TOLERANCE) is printed.
static constexpr float kDbgTolerance = 1.0e-4f;
static bool dbg_diff(float base, float eval) { return std::fabs(eval - base) >= kDbgTolerance; }

static LayerRec *dbg_child(int pid, int cid, const char *who) {
  LayerRec *P = layer(pid), *C = layer(cid);
  if (!P || !C || C->parent != pid) {
    fail(38, "%s: layer %d is not a child of layer %d", who, cid, pid);
    return nullptr;
  }
  if (C->sc) {
    fail(38, "%s: layer %d is coupled with SC_OPTION = 1", who, cid);
    return nullptr;
  }
  return C;
}

void edgeinterp_dbglaunch_(float *mO, float *nO, int *idO, float *mA, float *nA, int *idA, int *t,
                           float *yflux, float *xflux) {
  (void)mO;
  (void)nO;
  LayerRec *C = dbg_child(*idO, *idA, "edgeinterp_dbglaunch_");
  if (!C) return;
  const LayerRec *P = layer(*idO);
  if (*t < 1 || *t > C->rel_time) {
    fail(39, "edgeinterp_dbglaunch_: sub-step %d outside 1..%d", *t, C->rel_time);
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = launch_jnq(nest_view(*P, *C), *t, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { fail(209, "edgeinterp_dbglaunch_: %s", hipGetErrorString(e)); return; }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("TIME SPENT ON GPU (EDGE_INTERP->%d) %f\n", C->id, sec);
  const int nx = C->nx, ny = C->ny, pnx = P->nx, pny = P->ny;
  std::vector<float> m(C->nxy()), n(C->nxy()), xf(2 * (size_t)pnx), yf(2 * (size_t)pny);
  copy_plane(m.data(), C->m[C->cur], m.size(), false);
  copy_plane(n.data(), C->n[C->cur], n.size(), false);
  copy_plane(xf.data(), P->xflux, xf.size(), false);
  copy_plane(yf.data(), P->yflux, yf.size(), false);
  printf("==============lO:%d, lA:%d =============\n", P->id, C->id);
  if (*t == 2) {  // the flux cache NEWQ keeps at its second sub-step
    for (int i = C->corners[2] - 1; i <= C->corners[3]; ++i)
      if (dbg_diff(yf[i], yflux[i]))
        printf("(left-flux) i = %d, %f, %f, diff=%f\n", i, yf[i], yflux[i], std::fabs(yf[i] - yflux[i]));
    for (int i = C->corners[2] - 1; i <= C->corners[3]; ++i) {
      const int k = i + pny;
      if (dbg_diff(yf[k], yflux[k]))
        printf("(right-flux) i = %d, %f, %f, diff=%f\n", i, yf[k], yflux[k], std::fabs(yf[k] - yflux[k]));
    }
    for (int i = C->corners[0] - 1; i < C->corners[1]; ++i)
      if (dbg_diff(xf[i], xflux[i]))
        printf("(bottom-flux) i = %d, %f, %f, diff=%f\n", i, xf[i], xflux[i], std::fabs(xf[i] - xflux[i]));
    for (int i = C->corners[0] - 1; i <= C->corners[1]; ++i) {
      const int k = i + pnx;
      if (dbg_diff(xf[k], xflux[k]))
        printf("(top-flux) i = %d, %f, %f, diff=%f\n", i, xf[k], xflux[k], std::fabs(xf[k] - xflux[k]));
    }
  }
  for (int j = 0; j < ny; ++j) {
    const size_t k = (size_t)j * nx;
    if (dbg_diff(m[k], mA[k]))
      printf("(left) i = %d, %f, %f, diff=%f\n", j, m[k], mA[k], std::fabs(m[k] - mA[k]));
  }
  for (int j = 0; j < ny; ++j) {
    const size_t k = (size_t)j * nx + nx - 1;
    if (dbg_diff(m[k], mA[k]))
      printf("(right) i = %d, %f, %f, diff=%f\n", j, m[k], mA[k], std::fabs(m[k] - mA[k]));
  }
  for (int i = 0; i < nx; ++i)
    if (dbg_diff(n[i], nA[i]))
      printf("(bottom) i = %d, %f, %f, diff=%f\n", i, n[i], nA[i], std::fabs(n[i] - nA[i]));
  for (int i = 0; i < nx; ++i) {
    const size_t k = (size_t)(ny - 1) * nx + i;
    if (dbg_diff(n[k], nA[k]))
      printf("(top) i = %d, %f, %f, diff=%f\n", i, n[k], nA[k], std::fabs(n[k] - nA[k]));
  }
  fflush(stdout);
}

void jnz_dbglaunch_(float *Z_f, int *outid, int *inid) {
  LayerRec *C = dbg_child(*outid, *inid, "jnz_dbglaunch_");
  if (!C) return;
  LayerRec *P = layer(*outid);
  hipError_t e = launch_jnz(nest_view(*P, *C), g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { fail(209, "jnz_dbglaunch_: %s", hipGetErrorString(e)); return; }
  std::vector<float> z(P->nxy());
  copy_plane(z.data(), P->z[1 - P->cur], z.size(), false);
  printf("child corner %d %d %d %d\n", C->corners[0], C->corners[1], C->corners[2], C->corners[3]);
  for (int j = 0; j < P->ny; ++j)
    for (int i = 0; i < P->nx; ++i) {
      const size_t k = (size_t)j * P->nx + i;
      if (dbg_diff(z[k], Z_f[k])) {
        unsigned a, b;
        std::memcpy(&a, &z[k], 4);
        std::memcpy(&b, &Z_f[k], 4);
        printf("JNZ, layerid=%d(i,j)=(%d,%d), %f %f, diff=%f  %x, %x\n", C->id, i, j, z[k], Z_f[k],
               (double)std::fabs(z[k] - Z_f[k]), a, b);
      }
    }
  fflush(stdout);
}

void maxamp_launch_(const int *lid) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(17, "maxamp_launch_: no layer %d", *lid); return; }
  const bool root = L->id == 1;
  DevLayer d = dev_view(g_tree, *L, true, root);
  hipError_t e = launch_max(d, L->zmax_raw, g_stream);
  if (e != hipSuccess) { fail(203, "max launch failed: %s", hipGetErrorString(e)); return; }
  if (root) g_tree.ctl_in = 1 - g_tree.ctl_in;
  sync();
}

void cuda_update_layer_(int *lid) {
  LayerRec *L = layer(*lid);
  if (!L) { fail(18, "cuda_update_layer_: no layer %d", *lid); return; }
  L->cur = 1 - L->cur;
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
  (void)out_disable(g_tree);
  tree_free(g_tree);
  (void)hipStreamDestroy(g_stream);
  g_stream = nullptr;
  g_init = false;
}

// ------------------------------------------------------------- extended
// n root steps of the device loop, no output files
static int run_plain(int n, int use_graph) {
  if (tree_step2_eligible(g_tree) && n >= 2) {
    // a linear single layer: two root steps per launch; a graph of two
    // launches (four steps) brings the level parity back for replay
    if (use_graph && n >= 4) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
      int r = tree_enqueue_root_steps2(g_tree, g_stream);
      if (!r) r = tree_enqueue_root_steps2(g_tree, g_stream);
      hipError_t ce = hipStreamEndCapture(g_stream, &graph);
      if (r) return r;
      HIPCHK(ce);
      HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      for (int k = 0; k + 3 < n; k += 4) HIPCHK(hipGraphLaunch(exec, g_stream));
      HIPCHK(hipStreamSynchronize(g_stream));
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      n = n % 4;
    }
    for (; n >= 2; n -= 2)
      if (int r = tree_enqueue_root_steps2(g_tree, g_stream)) return r;
    for (; n > 0; --n)
      if (int r = tree_enqueue_root_step(g_tree, g_stream)) return r;
    return sync();
  }
  g_tree.mass_ahead = false;
  if (tree_fuses_halves(g_tree)) {
    // every root step but the last computes the next one's continuity in
    // its momentum launch; the first step (outside any graph) leaves the
    // tree in the state each captured pair starts and ends in
    int left = n;
    if (left > 0) {
      if (int r = tree_enqueue_root_step(g_tree, g_stream, left > 1)) return r;
      --left;
    }
    if (use_graph && left >= 3 && tree_launches_per_step(g_tree) <= kGraphMaxLaunches) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
      int r = tree_enqueue_root_step(g_tree, g_stream, true);
      if (!r) r = tree_enqueue_root_step(g_tree, g_stream, true);
      hipError_t ce = hipStreamEndCapture(g_stream, &graph);
      if (r) return r;
      HIPCHK(ce);
      HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      const int reps = (left - 1) / 2;  // at least one step stays for the end
      for (int k = 0; k < reps; ++k) HIPCHK(hipGraphLaunch(exec, g_stream));
      HIPCHK(hipStreamSynchronize(g_stream));
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      left -= 2 * reps;
    }
    for (; left > 0; --left)
      if (int r = tree_enqueue_root_step(g_tree, g_stream, left > 1)) return r;
    return sync();
  }
  if (use_graph && n >= 2 && tree_launches_per_step(g_tree) <= kGraphMaxLaunches) {
    // Two root steps bring every layer's level parity back to where it
    // started (children with odd REL_TIME flip once per root step), so a
    // captured pair can be replayed.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    int r = tree_enqueue_root_step(g_tree, g_stream);
    if (!r) r = tree_enqueue_root_step(g_tree, g_stream);
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
    if (int r = tree_enqueue_root_step(g_tree, g_stream)) return r;
  return sync();
}

// TIME of the main loop after root step k (comcot.f90:304, 323: REAL,
// one DT added per step), cached for the current layer 1
static long long g_clk_k = 0;
static float g_clk_t = 0.0f, g_clk_dt = -1.0f;

static float loop_time(long long k) {
  const float dt = g_tree.layers[1].dt;
  if (dt != g_clk_dt || k < g_clk_k) {
    g_clk_dt = dt;
    g_clk_k = 0;
    g_clk_t = 0.0f;
  }
  while (g_clk_k < k) {
    g_clk_t = g_clk_t + dt;
    ++g_clk_k;
  }
  return g_clk_t;
}

// GET_FLOOR_DEFORM's test (deform.f90:29): a fault ruptures at the call
// with TIME <= T0 < TIME + DT.  The call before the loop has TIME = 0
// (initialization.f90:476); the one of root step k, before its MASS, has
// the TIME after step k (comcot.f90:323, 341).
static bool fault_fires(const PendingFault &f, long long k) {
  const float t = loop_time(k), dt = g_tree.layers[1].dt;
  return f.t0 >= t && f.t0 < t + dt;
}

// Apply (and drop) every pending fault that fires at call k.
static int fire_faults(long long k) {
  for (size_t i = 0; i < g_faults.size();) {
    if (fault_fires(g_faults[i], k)) {
      if (int r = tree_apply_fault(g_tree, g_stream, g_faults[i].params, g_faults[i].fswitch))
        return r;
      g_faults.erase(g_faults.begin() + (long)i);
    } else {
      ++i;
    }
  }
  return 0;
}

int gcomcot_run_steps_(int *nsteps, int *use_graph) {
  int n = *nsteps;
  if (n < 0) return fail(23, "negative step count");
  if (!layer(1)) return fail(8, "no root layer");
  if (int e = ensure_init()) return e;
  if (!out_active() && g_faults.empty()) return run_plain(n, *use_graph);
  // the run is cut at every root step that writes files or before which a
  // fault ruptures; files are packed on the step stream and written while
  // the loop goes on
  if (sync()) return 1;
  int done = 0;
  HIPCHK(hipMemcpy(&done, g_tree.ctl + g_tree.ctl_in, sizeof(int), hipMemcpyDeviceToHost));
  if (done == 0) {
    if (int r = fire_faults(0)) return r;
    if (out_due(g_tree, 0))
      if (int r = out_write(g_tree, 0, g_stream)) return r;
  }
  while (n > 0) {
    if (int r = fire_faults(done + 1)) return r;
    long long m = out_active() ? out_steps_to_next(g_tree, done, n) : n;
    for (long long k = done + 2; k <= done + m && !g_faults.empty(); ++k) {
      bool any = false;
      for (const PendingFault &f : g_faults) any = any || fault_fires(f, k);
      if (any) {
        m = k - 1 - done;
        break;
      }
    }
    if (int r = run_plain((int)m, *use_graph)) return r;
    done += (int)m;
    n -= (int)m;
    if (out_due(g_tree, done))
      if (int r = out_write(g_tree, done, g_stream)) return r;
  }
  return sync();
}

int gcomcot_fault_schedule_(const float *params, int *fswitch, float *t0) {
  if (*fswitch != 0 && *fswitch != 9) return fail(60, "fault: epicentre switch %d not 0 or 9", *fswitch);
  for (int k = 0; k < 11; ++k)
    if (!std::isfinite(params[k])) return fail(61, "fault: parameter %d is not finite", k + 1);
  if (!std::isfinite(*t0)) return fail(61, "fault: rupture time is not finite");
  PendingFault f;
  std::memcpy(f.params, params, sizeof f.params);
  f.fswitch = *fswitch;
  f.t0 = *t0;
  g_faults.push_back(f);
  return 0;
}

int gcomcot_fault_pending_(void) { return (int)g_faults.size(); }

int gcomcot_output_set_(const char *dir, int *dirlen, int *iprt, int *out_opt, float *tend,
                        float *tide) {
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  return out_configure(g_tree, dir, *dirlen, *iprt, *out_opt, *tend, *tide);
}

int gcomcot_checkpoint_save_(const char *path, int *pathlen) {
  if (!layer(1)) return fail(8, "no root layer");
  if (!path || *pathlen <= 0) return fail(70, "checkpoint: empty path");
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  int done = 0;
  HIPCHK(hipMemcpy(&done, g_tree.ctl + g_tree.ctl_in, sizeof(int), hipMemcpyDeviceToHost));
  const std::string p(path, (size_t)*pathlen);
  return tree_checkpoint_save(g_tree, g_stream, p.c_str(), done);
}

int gcomcot_checkpoint_load_(const char *path, int *pathlen) {
  if (!layer(1)) return fail(8, "no root layer");
  if (!path || *pathlen <= 0) return fail(72, "checkpoint: empty path");
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  const std::string p(path, (size_t)*pathlen);
  long long done = 0;
  if (int e = tree_checkpoint_load(g_tree, g_stream, p.c_str(), &done)) return e;
  // scheduled ruptures whose call (GET_FLOOR_DEFORM's TIME <= T0 < TIME + DT)
  // lies at or before the restored step are in the restored state already
  const float next = loop_time(done) + g_tree.layers[1].dt;
  for (size_t i = 0; i < g_faults.size();)
    if (g_faults[i].t0 < next) g_faults.erase(g_faults.begin() + (long)i);
    else ++i;
  return 0;
}

int gcomcot_hotstart_write_(int *start_step, const char *dir, int *dirlen) {
  if (!layer(1)) return fail(8, "no root layer");
  if (!start_step || !dirlen || *dirlen < 0 || (*dirlen > 0 && !dir))
    return fail(95, "hot start: bad arguments");
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  return tree_hotstart_write(g_tree, g_stream, *start_step, std::string(dir ? dir : "", (size_t)*dirlen));
}

int gcomcot_hotstart_read_(int *start_type, int *start_step, const char *dir, int *dirlen,
                           int *nmissing) {
  if (!layer(1)) return fail(8, "no root layer");
  if (!start_type || !start_step || !dirlen || *dirlen < 0 || (*dirlen > 0 && !dir))
    return fail(95, "hot start: bad arguments");
  if (*start_step < 1) return fail(95, "hot start: START_STEP %d below 1", *start_step);
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  int miss = 0;
  if (int e = tree_hotstart_read(g_tree, g_stream, *start_type, *start_step,
                                 std::string(dir ? dir : "", (size_t)*dirlen), &miss))
    return e;
  // the resumed loop starts at root step K = START_STEP (comcot.f90:228-230,
  // 322), so START_STEP - 1 root steps lie behind the loaded state
  const int done = *start_step - 1;
  HIPCHK(hipMemcpy(g_tree.ctl + g_tree.ctl_in, &done, sizeof(int), hipMemcpyHostToDevice));
  const float next = loop_time(done) + g_tree.layers[1].dt;
  for (size_t i = 0; i < g_faults.size();)
    if (g_faults[i].t0 < next) g_faults.erase(g_faults.begin() + (long)i);
    else ++i;
  if (nmissing) *nmissing = miss;
  return 0;
}

int gcomcot_hotstart_file_write_(const char *path, int *pathlen, const float *plane, int *nx,
                                 int *ny, int *first) {
  if (!path || !pathlen || *pathlen <= 0 || !plane || !nx || !ny || !first)
    return fail(95, "hot start: bad arguments");
  return hotstart_write_plane(std::string(path, (size_t)*pathlen).c_str(), plane, *nx, *ny,
                              *first);
}

int gcomcot_hotstart_file_read_(const char *path, int *pathlen, float *plane, int *nx, int *ny,
                                int *first) {
  if (!path || !pathlen || *pathlen <= 0 || !plane || !nx || !ny || !first)
    return fail(95, "hot start: bad arguments");
  const std::string p(path, (size_t)*pathlen);
  const int e = hotstart_read_plane(p.c_str(), plane, *nx, *ny, *first);
  if (e == kHotstartMissing) return fail(96, "hot start: cannot open %s", p.c_str());
  return e;
}

int gcomcot_sponge_set_(int *on) {
  LayerRec *R = layer(1);
  if (!R) return fail(8, "no root layer");
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  if (R->sponge) HIPCHK(hipFree(R->sponge));
  R->sponge = nullptr;
  if (!*on) return 0;
  if ((int)g_tree.xs[1].size() != R->nx || (int)g_tree.ys[1].size() != R->ny)
    return fail(42, "sponge: coordinates of layer 1 not set");
  std::vector<float> h(R->nxy()), c((size_t)R->nx + R->ny);
  HIPCHK(hipMemcpy(h.data(), R->h, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  int laycord = R->laycord;
  gcomcot_sponge_coef_(&laycord, &R->nx, &R->ny, h.data(), g_tree.xs[1].data(),
                       g_tree.ys[1].data(), c.data(), c.data() + R->nx);
  HIPCHK(hipMalloc(&R->sponge, c.size() * sizeof(float)));
  HIPCHK(hipMemcpy(R->sponge, c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

int gcomcot_fault_deform_(const float *params, int *fswitch) {
  if (!layer(1)) return fail(8, "no root layer");
  if (int e = ensure_init()) return e;
  if (sync()) return 1;
  return tree_apply_fault(g_tree, g_stream, params, *fswitch);
}

int gcomcot_get_deform_(int *lid, float *deform) {
  LayerRec *L = layer(*lid);
  if (!L) return fail(24, "gcomcot_get_deform_: no layer %d", *lid);
  if (sync()) return 1;
  if (!L->deform) {
    std::memset(deform, 0, L->nxy() * sizeof(float));
    return 0;
  }
  HIPCHK(hipMemcpy(deform, L->deform, L->nxy() * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int gcomcot_get_depth_(int *lid, float *h) {
  LayerRec *L = layer(*lid);
  if (!L) return fail(24, "gcomcot_get_depth_: no layer %d", *lid);
  if (sync()) return 1;
  HIPCHK(hipMemcpy(h, L->h, L->nxy() * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int gcomcot_output_layer_(int *lid, int *fluxswitch) {
  return out_set_fluxswitch(*lid, *fluxswitch);
}

int gcomcot_output_flush_(long long *nfiles) {
  long long n = 0;
  if (g_init && sync()) return 1;
  const int r = out_flush(&n);
  if (nfiles) *nfiles = n;
  return r;
}

int gcomcot_time_steps_(int *nsteps, float *ms_total, float *ms_each) {
  const int n = *nsteps;
  if (n < 1) return fail(23, "step count must be >= 1");
  if (!layer(1)) return fail(8, "no root layer");
  if (int e = ensure_init()) return e;
  std::vector<hipEvent_t> ev(n + 1);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], g_stream));
  g_tree.mass_ahead = false;
  for (int k = 0; k < n; ++k) {
    if (int r = tree_enqueue_root_step(g_tree, g_stream, k + 1 < n)) return r;
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

void gcomcot_set_arrival_threshold_(float *thr) { g_tree.arr_thr = *thr; }

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
    std::strncpy(buf, last_message(), (size_t)*len - 1);
    buf[*len - 1] = 0;
  }
  return take_last_code();
}

int gcomcot_autotune_(int *reps, int *variant, int *blocking, float *ms_per_step) {
  LayerRec *R = layer(1);
  if (!R) return fail(8, "no root layer");
  if (*reps < 1) return fail(29, "autotune: reps must be >= 1");
  if (int e = ensure_init()) return e;
  // dry launches (no maximum, no counter advance) read level 1 and write
  // level 2 only, which the next real step rewrites: the state is untouched
  DevLayer d = dev_view(g_tree, *R, false, false);
  const bool fused_only = R->children.empty();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 1e30f;
  int bv = g_tree.variant, bb = 0;
  // The GPU's clocks settle over the first ~0.1 s of sustained work; the
  // candidates are timed after that, so that the first ones are not timed
  // at lower clocks than the last (dry launches of the current variant,
  // at most kSettleLaunches of them).
  {
    constexpr int kSettleLaunches = 4000;
    constexpr float kSettleMs = 100.0f;
    HIPCHK(hipEventRecord(e0, g_stream));
    float ms = 0.0f;
    for (int k = 0, batch = 8; k < kSettleLaunches && ms < kSettleMs; k += batch) {
      for (int q = 0; q < batch; ++q)
        HIPCHK(launch_step(d, kFused, R->nl != 0, R->sph != 0, g_tree.variant, g_stream));
      HIPCHK(hipEventRecord(e1, g_stream));
      HIPCHK(hipEventSynchronize(e1));
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      batch = std::min(2 * batch, 256);
    }
  }
  auto time_it = [&](auto launch, int steps_per_launch, float *out) -> int {
    HIPCHK(launch());
    HIPCHK(hipEventRecord(e0, g_stream));
    for (int k = 0; k < *reps; ++k) HIPCHK(launch());
    HIPCHK(hipEventRecord(e1, g_stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *out = ms / (*reps * steps_per_launch);
    return 0;
  };
  if (!fused_only) {
    // nested trees: each layer without nested layers of its own runs fused
    // launches; its tile variant is chosen on its own (dry launches on that
    // layer); the others run the split halves, whose tile is fixed.  The
    // reported time sums those layers' fused launches per root step, the
    // reported variant is the one of the layer with the most cell updates.
    float sum = 0.0f;
    long long most = -1;
    for (LayerRec &L : g_tree.layers) {
      if (!L.used || L.id == 1 || !L.children.empty()) continue;
      int per_root = 1;
      for (const LayerRec *q = &L; q->id != 1; q = &g_tree.layers[q->parent]) per_root *= q->rel_time;
      DevLayer dl = dev_view(g_tree, L, false, false);
      float lbest = 1e30f;
      int lv = 0;
      for (int v = 0; v < num_tile_variants(); ++v) {
        float ms;
        if (int r = time_it([&] { return launch_step(dl, kFused, L.nl != 0, L.sph != 0, v, g_stream); },
                            1, &ms))
          return r;
        if (ms < lbest) { lbest = ms; lv = v; }
      }
      L.variant = lv;
      sum += lbest * per_root;
      const long long upd = (long long)L.nx * L.ny * per_root;
      if (upd > most) { most = upd; bv = lv; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (variant) *variant = bv;
    if (blocking) *blocking = g_tree.blocking;
    if (ms_per_step) *ms_per_step = sum;
    return 0;
  }
  if (fused_only) {
    for (int v = 0; v < num_tile_variants(); ++v) {
      float ms;
      if (int r = time_it([&] { return launch_step(d, kFused, R->nl != 0, R->sph != 0, v, g_stream); },
                          1, &ms))
        return r;
      if (ms < best) { best = ms; bv = v; bb = 0; }
    }
    if (!R->nl) {
      for (int sh = 0; sh < kStep2Shapes; ++sh) {
        float ms;
        if (int r = time_it([&] { return launch_step2(d, R->sph != 0, sh, g_stream); }, 2, &ms))
          return r;
        if (ms < best) { best = ms; bb = sh + 1; }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_tree.variant = bv;
  g_tree.blocking = bb;
  if (variant) *variant = bv;
  if (blocking) *blocking = bb;
  if (ms_per_step) *ms_per_step = best;
  return 0;
}

int gcomcot_set_layer_coords_(int *lid, const float *x, const float *y) {
  LayerRec *L = layer(*lid);
  if (!L) return fail(41, "gcomcot_set_layer_coords_: no layer %d", *lid);
  g_tree.xs[*lid].assign(x, x + L->nx);
  g_tree.ys[*lid].assign(y, y + L->ny);
  return 0;
}

int gcomcot_gauges_set_(int *n, const float *x0, const float *y0, const int *lid, int *capacity) {
  if (int e = ensure_init()) return e;
  return tree_set_gauges(g_tree, g_stream, *n, x0, y0, lid, *capacity);
}

int gcomcot_gauges_read_(float *series, int *max_steps, int *nout) {
  *nout = 0;
  if (g_tree.ngauges == 0) return 0;
  if (sync()) return 1;
  int done = 0;
  HIPCHK(hipMemcpy(&done, g_tree.ctl + g_tree.ctl_in, sizeof(int), hipMemcpyDeviceToHost));
  const int cap = g_tree.gauge_cap, n = g_tree.ngauges;
  long long recorded = done - g_tree.gauge_step0;
  long long count = recorded;
  if (count > cap) count = cap;
  if (count > *max_steps) count = *max_steps;
  if (count <= 0) return 0;
  std::vector<float> ring((size_t)n * cap);
  HIPCHK(hipMemcpy(ring.data(), g_tree.ring, ring.size() * sizeof(float), hipMemcpyDeviceToHost));
  const long long first = done - count + 1;  // oldest step kept
  for (int g = 0; g < n; ++g)
    for (long long k = 0; k < count; ++k)
      series[(size_t)g * *max_steps + k] = ring[(size_t)g * cap + (size_t)((first + k - 1) % cap)];
  *nout = (int)count;
  return 0;
}

int gcomcot_set_fused_halves_(int *on) {
  g_tree.fuse_halves = *on != 0;
  return 0;
}

int gcomcot_set_store_policy_(int *mode) {
  if (*mode < 0 || *mode > 2) return fail(46, "store policy %d is not 0, 1 or 2", *mode);
  if (sync()) return 1;  // launches in flight keep the policy they were built with
  g_store_policy = *mode;
  return 0;
}

int gcomcot_set_temporal_blocking_(int *mode) {
  if (*mode < 0 || *mode > kStep2Shapes) return fail(28, "temporal blocking mode %d out of range", *mode);
  g_tree.blocking = *mode;
  return 0;
}

int gcomcot_hbm_copy_(int *mib, int *reps, float *ms_per_copy) {
  if (int e = ensure_init()) return e;
  if (*mib < 1 || *reps < 1) return fail(27, "gcomcot_hbm_copy_: sizes must be positive");
  const size_t bytes = (size_t)*mib << 20;
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  HIPCHK(hipMalloc(&b, bytes));
  HIPCHK(hipMemsetAsync(a, 0, bytes, g_stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(launch_copy(a, b, bytes, g_stream));  // warm-up
  HIPCHK(hipEventRecord(e0, g_stream));
  for (int k = 0; k < *reps; ++k) HIPCHK(launch_copy(a, b, bytes, g_stream));
  HIPCHK(hipEventRecord(e1, g_stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / *reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int gcomcot_division_check_(long long *pairs, unsigned long long *seed, long long *bad_div12,
                            long long *bad_inrange, long long *bad_sqrt, long long *tested,
                            unsigned *example_x, unsigned *example_a, unsigned *example_b,
                            unsigned *example_s) {
  if (int e = ensure_init()) return e;
  if (*pairs < 0) return fail(35, "gcomcot_division_check_: pairs must be >= 0");
  const int blocks = 4096, lanes = blocks * 256;
  const long long per = (*pairs + lanes - 1) / lanes;
  if (per > (1 << 20)) return fail(35, "gcomcot_division_check_: at most 2^20 pairs per lane");
  unsigned long long *d = nullptr;
  HIPCHK(hipMalloc(&d, 6 * (size_t)lanes * sizeof(unsigned long long)));
  hipError_t e = launch_division_check(*seed, (int)per, d, blocks, g_stream);
  std::vector<unsigned long long> h(6 * (size_t)lanes);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(205, "division check failed: %s", hipGetErrorString(e));
  long long b12 = 0, bin = 0, bsq = 0;
  *example_x = *example_a = *example_b = *example_s = 0;
  for (int g = 0; g < lanes; ++g) {
    const unsigned long long *o = h.data() + 6 * (size_t)g;
    if (o[0] && !b12) *example_x = (unsigned)o[3];
    if (o[1] && !bin) {
      *example_a = (unsigned)(o[4] >> 32);
      *example_b = (unsigned)o[4];
    }
    if (o[2] && !bsq) *example_s = (unsigned)o[5];
    b12 += (long long)o[0];
    bin += (long long)o[1];
    bsq += (long long)o[2];
  }
  *bad_div12 = b12;
  *bad_inrange = bin;
  *bad_sqrt = bsq;
  *tested = per * (long long)lanes;
  return 0;
}

int gcomcot_detmath_eval_(long long *n, const float *x, float *pow_out, float *exp_out) {
  if (*n < 0 || (*n > 0 && (!x || !pow_out || !exp_out)))
    return fail(36, "gcomcot_detmath_eval_: n must be >= 0 and the arrays given");
  if (*n == 0) return 0;
  if (int e = ensure_init()) return e;
  const size_t bytes = (size_t)*n * sizeof(float);
  float *d = nullptr;
  HIPCHK(hipMalloc(&d, 3 * bytes));
  const int rc = detmath_eval_staged(d, *n, x, pow_out, exp_out);
  (void)hipFree(d);
  return rc;
}

int gcomcot_detmath_eval_noalloc_(long long *n, const float *x, float *pow_out,
                                  float *exp_out) {
  if (!n || *n < 0 || (*n > 0 && (!x || !pow_out || !exp_out)))
    return fail(36, "gcomcot_detmath_eval_noalloc_: n must be >= 0 and the arrays given");
  if (*n > kNoAllocMaxArgs)
    return fail(37, "gcomcot_detmath_eval_noalloc_: n = %lld exceeds the %lld arguments its "
                "scratch holds", *n, kNoAllocMaxArgs);
  if (*n == 0) return 0;
  if (int e = ensure_init()) return e;
  float *d = nullptr;
  if (int e = noalloc_scratch(&d)) return e;
  return detmath_eval_staged(d, *n, x, pow_out, exp_out);
}

int gcomcot_set_tile_variant_(int *variant) {
  if (*variant < 0 || *variant >= num_tile_variants()) return 1;
  g_tree.variant = *variant;
  for (LayerRec &L : g_tree.layers) L.variant = -1;  // the autotuner's per-layer choices
  return 0;
}

int gcomcot_tile_variants_(int *count, int *tx, int *ty) {
  *count = num_tile_variants();
  if (tx && ty)
    for (int v = 0; v < *count; ++v) tile_variant_shape(v, &tx[v], &ty[v]);
  return g_tree.variant;
}

}  // extern "C"
