// Layer trees on the device: one COMCOT nest (an outermost layer and its
// nested layers), its device planes and the enqueue order of one step.
//
// The drop-in C ABI keeps one tree (the reference's global layer table,
// GPUcomcot.cu:5-22); the batch engine keeps one tree per scenario.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "kernels.h"

namespace gc {

constexpr int kMaxLayers = 20;  // layer ids 1..19 (GPUHeader.h:51)
// Launch sequences longer than this per root step are not captured into
// HIP graphs (deep sub-stepping, e.g. a REL_TIME in the thousands).
constexpr int kGraphMaxLaunches = 4096;

struct LayerRec {
  bool used = false;
  int id = 0, parent = -1, level = 1;
  int laycord = 0, laygov = 0, nl = 0, sph = 1;
  int rel_size = 1, rel_time = 1, corners[4] = {0, 0, 0, 0};
  int fric = 0, modscm = 0, linchk = 1;
  float fric_coef = 0, dt = 0, dx = 0, dy = 0, rx = 0, ry = 0, grx = 0, gry = 0, kfric = 0;
  int nx = 0, ny = 0;
  void *block = nullptr;  // one allocation holding every plane below
  size_t block_bytes = 0;
  float *h = nullptr, *z[2] = {nullptr, nullptr}, *m[2] = {nullptr, nullptr},
        *n[2] = {nullptr, nullptr};
  float *zmax_raw = nullptr, *zmax_wet = nullptr, *coef = nullptr, *r24 = nullptr;
  float *dzp = nullptr, *xflux = nullptr, *yflux = nullptr;
  int *arrival = nullptr;
  float *deform = nullptr;  // last seafloor deformation applied (on first use)
  float *sponge = nullptr;  // SPONGE_COEF: COEFX[NX] then COEFY[NY] (layer 1, when on)
  // SC_OPTION = 1 (Cartesian child of a spherical layer 1): POS, CXY, the
  // NEWQ_SC bands and JNZ_SC's child lists, on the device; corners[] is
  // then the POS extent
  int sc = 0;
  int *sc_pos = nullptr, *sc_jnz = nullptr;  // POS (2 planes); CSR offsets then indices
  float *sc_cxy = nullptr;
  int sc_band[4][4] = {};
  size_t sc_jnz_cells = 0;
  int cur = 0;  // buffer index holding time level 1
  int variant = -1;  // tile variant of this layer's fused launches (-1: the tree's)
  std::vector<int> children;
  size_t nxy() const { return (size_t)nx * ny; }
};

struct Tree {
  LayerRec layers[kMaxLayers];
  int *ctl = nullptr;  // device int[2]: root-step counters for the arrival field
  int ctl_in = 0;      // which counter holds the completed step count
  float arr_thr = 0.01f;
  // grid coordinates LAYER%X / %Y per layer (gauge stencil search), host
  std::vector<float> xs[kMaxLayers], ys[kMaxLayers];
  // gauges sampled after every root step of the device loop
  GaugeDev *gauges = nullptr;
  float *ring = nullptr;
  int ngauges = 0, gauge_cap = 0;
  long long gauge_step0 = 0;  // root steps completed when the gauges were set
  int variant = 0;     // tile variant of the step kernels
  int blocking = 1;    // two steps per launch for linear single layers: 0 off, 1+shape
  int raw_max = 0;     // device loop also keeps MAX_AMP's maximum of the outermost layer
  // a linear outermost layer with nested layers: its momentum launch also
  // computes the next root step's continuity when that step follows
  // (fuse_halves); mass_ahead: level 2 already holds that continuity (a
  // cache: continuity from level 1 gives the same bits when recomputed)
  int fuse_halves = 1;
  bool mass_ahead = false;
  LayerRec *layer(int id) {
    return (id >= 1 && id < kMaxLayers && layers[id].used) ? &layers[id] : nullptr;
  }
  size_t device_bytes() const;
};

// ---- error state shared by the whole library (reference: print + exit)
extern int g_fatal;
int fail(int code, const char *fmt, ...);
const char *last_message();
int take_last_code();

#define HIPCHK(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return ::gc::fail(100 + (int)e_, "%s at %s:%d: %s", #call, __FILE__,       \
                        __LINE__, hipGetErrorString(e_));                        \
  } while (0)

// How step launches store the new level-2 planes (DevLayer::stream_out):
// 0 through the caches, 2 streamed (nontemporal) for every layer, 1 (the
// default) streamed for layers of more than kStreamCells cells, whose one
// step (about 36 bytes per cell) moves more than three quarters of the
// 256 MB last-level cache, and for layers of at most kL2Cells cells, whose
// step fits in the 32 MB of the XCDs' L2s together: cached stores there
// leave the whole new level dirty in L2, written back at every kernel end.
// The bits are the same every way.
extern int g_store_policy;
constexpr long long kStreamCells = (192ll << 20) / 36;
constexpr long long kL2Cells = (32ll << 20) / 36;
bool streams_planes(const LayerRec &L);

// Bytes of device memory a layer of nx x ny cells takes (planes + tables).
size_t layer_bytes(int nx, int ny, bool nl, bool r24);

// Unlink nested layer id from its parent and free it (no-op for the root
// or an unused id).
void tree_drop_layer(Tree &T, int id);

// Register layer `id` of tree T: validates the nest geometry as the
// reference's SUBGRID_MATCHING does (initialization.f90:882-1131), derives
// the ALLOC scalars (comcot.f90:537-552, 602), allocates and clears the
// device planes, uploads H and Z.  A new root frees the old tree.
// The argument checks of tree_create_layer alone (no device work).
int tree_validate_layer(Tree &T, int id, int parent, int rel_size, int rel_time,
                        const int *corners, int nx, int ny, bool sc = false);
int tree_create_layer(Tree &T, hipStream_t st, int id, int parent, int level, int laycord,
                      int laygov, int rel_size, int rel_time, const int *corners,
                      int fric_switch, float fric_coef, float dt, float dx, float dy, int nx,
                      int ny, const float *H, const float *Z, bool sc = false);
// Couple child layer id to its spherical parent with SC_OPTION = 1, from
// SUBGRID_MATCHING's POS (NX*NY*2) and CXY (NX*NY*4), column-major and
// component-major as the Fortran holds them.
int tree_set_sc(Tree &T, hipStream_t st, int id, const int *pos, const float *cxy);
int tree_upload_coef(LayerRec &L, const std::vector<float> &coef);
int tree_alloc_r24(LayerRec &L, const float *R2, const float *R4);
void tree_free(Tree &T);

DevLayer dev_view(const Tree &T, const LayerRec &L, bool with_max, bool advance);
DevNest nest_view(const LayerRec &P, const LayerRec &C);
// The SC_OPTION = 1 coupling views of parent P and child C of tree T.
DevSc sc_view(const Tree &T, const LayerRec &P, const LayerRec &C);

// One layer pass (kFused / kMassOnly / kMomentOnly) on stream st.
int tree_launch(const Tree &T, const LayerRec &L, int mode, bool with_max, bool advance,
                hipStream_t st);
// Sub-steps of every child of P for one step of P (ALL_GRID order).
int tree_enqueue_children(Tree &T, LayerRec &P, bool with_max, hipStream_t st);
// One complete root step: mass+open, nests, momentum + max, rotation.
// next: another root step follows at once (its continuity may then ride in
// this step's momentum launch, tree_fuses_halves).
int tree_enqueue_root_step(Tree &T, hipStream_t st, bool next = false);
bool tree_fuses_halves(const Tree &T);
// The view of the launch that ends root step k and starts step k+1 (z1 =
// Z(k), z2 = the buffer for Z(k+1)).
DevLayer fused_halves_view(const Tree &T, const LayerRec &R);
// True when the tree is one linear layer (two steps per launch apply).
bool tree_step2_eligible(const Tree &T);
// Two root steps of such a tree in one launch (temporal blocking).
int tree_enqueue_root_steps2(Tree &T, hipStream_t st);
// Gauges: stencil cell and weights of GET_TS (output.f90:1035-1155) for
// (x0, y0) on layer lid from its coordinates; sampling launch per step.
int tree_set_gauges(Tree &T, hipStream_t st, int n, const float *x0, const float *y0,
                    const int *lid, int capacity);
int tree_enqueue_gauges(const Tree &T, hipStream_t st);
// Kernel launches one root step of the tree enqueues.
// Exact checkpoint of a tree's device state (every plane of every layer,
// the time-level parity and the root-step count) to a file, and its
// restore into a tree built with the same layers.
int tree_checkpoint_save(const Tree &T, hipStream_t st, const char *path, long long done);
int tree_checkpoint_load(Tree &T, hipStream_t st, const char *path, long long *done);

long long tree_launches_per_step(Tree &T);

// R2/R4 of the reference ABI as a row factor of max(HP,0) (PARA,
// comcot.f90:733-744); false when no factor reproduces the row exactly.
bool row_factor(const float *R, const float *H, int nx, int ny, int j, bool is_x, float *F);

}  // namespace gc
