// Scenario batch engine: many independent COMCOT nests advanced together.
//
// Each scenario is a layer tree (an outermost layer and at most one level
// of nested layers) with its own step count.  All active scenarios move
// in lockstep, one root step at a time, and every phase of the step is
// ONE launch over all of them:
//
//   fused step of every single-layer root      } persistent step kernel,
//   continuity half of every nested root       } one launch per layer kind
//   for each child slot s, sub-step k = 1..max REL_TIME:
//     JNQ/NEWQ of every child s with k <= its REL_TIME   (one launch; an
//       SC_OPTION = 1 child takes its band and edge launches instead)
//     fused step of those children                       (one per kind)
//     JNZ of the children at their half step             (one launch; JNZ_SC
//       per SC_OPTION = 1 child)
//   momentum half of every nested root (+ max / arrival)
//
// which is the per-scenario order of the drop-in path (all_grids.f90:34-
// 308, comcot.f90:359-439), so every scenario's planes are bit-identical
// to a run of that scenario alone.  Single-layer linear roots may instead
// take one launch per two root steps.  The phases are emitted in lanes of
// disjoint scenarios (single-layer scenarios; two halves of the nested
// ones) that run on three streams, forked and joined within each plan.
// The launches of a plan (two root steps, four with two-step roots) are
// captured in a HIP graph and replayed; the plan is rebuilt when the set
// of active scenarios changes (a scenario finishes, is added or removed).
//
// Step launches run (layer, tile) items either in the persistent kernel
// k_batch (8 per-XCD queues with atomic heads, cleared by a fill kernel at
// the start of each root step, inside the graph) or one workgroup per item
// (k_batch_grid).  Form and tile shape of the fused launches come from the
// autotuner below: each candidate is timed on the live batch with the
// max/arrival outputs switched off (such a "dry" launch writes only
// level-2 planes, which the next real step rewrites), and the fastest is
// kept per (layer kind, size class).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "comcot_consts.h"
#include "gcomcot_hip_internal.h"
#include "kernels.h"
#include "source.h"
#include "tree.h"

using namespace gc;

namespace {

struct Scenario {
  int sid = 0;
  Tree tree;
  long long total = 0, done = 0;
  bool finished() const { return done >= total; }
};

enum GroupType { kStepGroup = 0, kJnqGroup = 1, kJnzGroup = 2, kScEdgeGroup = 3, kScJnzGroup = 4 };
// Step-group mode beside the step modes of kernels.h: two root steps of
// single-layer linear roots in one launch (k_batch2_grid).
constexpr int kTwoStep = 4;
static_assert(kTwoStep != kFused && kTwoStep != kMassOnly && kTwoStep != kMomentOnly &&
                  kTwoStep != kMomentMass, "group modes must differ");

struct Group {
  int type = kStepGroup;
  int mode = kFused, nl = 0, sph = 1, shape = 0, t = 0;
  std::vector<BatchDesc> descs;
  std::vector<BatchItem> items;
  int qstart[9] = {0};
  std::vector<DevNest> nests;
  std::vector<DevSc> scs;  // SC_OPTION = 1 children: one launch pair each
  std::vector<int> start;  // nnest + 1
  size_t o_descs = 0, o_items = 0, o_qstart = 0, o_nests = 0, o_start = 0;
  int head = 0;            // index of this launch's 8 heads
  long long cells = 0;     // cell updates of a step group
  std::vector<std::pair<const Tree *, const LayerRec *>> recs;  // layer of each desc
  // stream lane: 0 and 2 = the two chains of nested scenarios, 1 = the
  // single-layer scenarios; the lanes share no layer, so they run side by side
  int lane = 0;
};

struct Plan {
  // root steps of the plan: 2, or 4 when single-layer linear roots take one
  // two-step launch per pair of root steps (their level parity comes back
  // after two such launches)
  std::vector<Group> step[4];
  int nsteps = 0;
  char *arena = nullptr;
  int *heads = nullptr;
  int heads_n[4] = {0, 0, 0, 0};
  int heads_off[4] = {0, 0, 0, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  long long cells_per_step = 0;
  bool owns_arena = true;  // false: the engine's reusable arena
};

struct Engine {
  bool init = false;
  hipStream_t st = nullptr;
  // lanes 1 and 2 of a plan run on st2 and st3, forked from st at the start
  // of the plan and joined back at its end (all at default priority)
  hipStream_t st2 = nullptr, st3 = nullptr;
  hipEvent_t fork = nullptr, join2 = nullptr, join3 = nullptr;
  int streams = 3;
  std::map<int, std::unique_ptr<Scenario>> scen;
  int next_sid = 1;
  float arr_thr = 0.01f;
  bool autotune = true;
  bool graphs = true;      // replay plans from HIP graphs (else enqueue each step)
  bool two_step = true;    // single-layer linear roots may take two root steps per launch
  bool fuse_halves = true; // nested linear roots: momentum + next continuity in one launch
  int step2_shape = 4;     // their tile shape (launch_step2's shapes); -1: one step per launch
  int forced_step2 = -1;   // a shape set by the caller (no tuning of it)
  int forced_shape = -1;
  std::map<long long, int> tuned;  // (kind, size class) -> shape
  std::map<long long, float> tuned_ms;  // (kind, size class) -> its time per launch
  std::map<long long, int> tuned2;  // size class of the two-step roots -> shape or -1
  Plan plan;
  bool plan_valid = false;
  int plan_policy = -1;    // the store policy its launches were built with
  char *arena = nullptr;   // descriptor/item memory reused by successive plans
  size_t arena_bytes = 0;
  long long launches = 0, graph_replays = 0, plans_built = 0;
};

Engine g_eng;

bool trace_on();

int eng_init() {
  if (g_eng.init) return 0;
  HIPCHK(hipSetDevice(0));
  HIPCHK(hipStreamCreateWithFlags(&g_eng.st, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&g_eng.st2, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&g_eng.st3, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&g_eng.fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g_eng.join2, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g_eng.join3, hipEventDisableTiming));
  g_eng.init = true;
  return 0;
}

void free_plan(Plan &P) {
  if (P.exec) (void)hipGraphExecDestroy(P.exec);
  if (P.graph) (void)hipGraphDestroy(P.graph);
  if (P.arena && P.owns_arena) (void)hipFree(P.arena);
  P = Plan();
}

void invalidate() {
  if (g_eng.plan.arena || g_eng.plan.exec) {
    (void)hipStreamSynchronize(g_eng.st);
    (void)hipStreamSynchronize(g_eng.st2);
    (void)hipStreamSynchronize(g_eng.st3);
    free_plan(g_eng.plan);
  }
  g_eng.plan_valid = false;
}

Scenario *scenario(int sid) {
  auto it = g_eng.scen.find(sid);
  return it == g_eng.scen.end() ? nullptr : it->second.get();
}

std::vector<Scenario *> active() {
  std::vector<Scenario *> a;
  for (auto &kv : g_eng.scen) {
    Scenario *s = kv.second.get();
    if (!s->finished() && s->tree.layer(1)) a.push_back(s);
  }
  return a;
}

// Tuning classes: layer kind and whether the launch alone fills the chip
// (below ~4M cells a launch leaves CUs idle and favours smaller tiles).
long long kind_key(int nl, int sph, long long cells) {
  return ((long long)nl * 2 + sph) * 2 + (cells >= (4LL << 20) ? 1 : 0);
}

int shape_for(int nl, int sph, long long cells) {
  if (g_eng.forced_shape >= 0) return g_eng.forced_shape;
  auto it = g_eng.tuned.find(kind_key(nl, sph, cells));
  return it == g_eng.tuned.end() ? batch_num_shapes() / 2 : it->second;
}

// Step group for layers of one kind and mode; tiles of tx x ty.
void add_layer_items(Group &g, const DevLayer &d, int tx, int ty) {
  BatchDesc bd{};
  bd.L = d;
  bd.tiles_x = (d.nx + tx - 1) / tx;
  bd.ntiles = bd.tiles_x * ((d.ny + ty - 1) / ty);
  const int slot = (int)g.descs.size();
  g.descs.push_back(bd);
  for (int t = 0; t < bd.ntiles; ++t) g.items.push_back(BatchItem{slot, t});
  g.cells += (long long)d.nx * d.ny;
}

void split_queues(Group &g) {
  const long long n = (long long)g.items.size();
  for (int q = 0; q <= 8; ++q) g.qstart[q] = (int)(n * q / 8);
}

struct Pending {  // layers of one launch before tiling
  int mode, nl, sph;
  std::vector<DevLayer> layers;
  std::vector<std::pair<const Tree *, const LayerRec *>> recs;
};

void flush_steps(std::vector<Pending> &pend, std::vector<Group> &out) {
  for (auto &p : pend) {
    if (p.layers.empty()) continue;
    Group g;
    g.type = kStepGroup;
    g.mode = p.mode;
    g.nl = p.nl;
    g.sph = p.sph;
    long long cells = 0;
    for (auto &d : p.layers) cells += (long long)d.nx * d.ny;
    // fused launches: the tuned form and tile; continuity/momentum halves
    // of nested roots: one workgroup per 64 x 8 tile; two-step launches:
    // the engine's two-step shape
    int tx = 64, ty = 8;
    if (p.mode == kTwoStep) {
      g.shape = g_eng.step2_shape;
      step2_shape(g.shape, &tx, &ty);
    } else {
      g.shape = p.mode == kFused ? shape_for(p.nl, p.sph, cells) : batch_num_shapes() / 2;
      batch_shape(g.shape, &tx, &ty);
    }
    for (auto &d : p.layers) add_layer_items(g, d, tx, ty);
    g.recs = p.recs;
    split_queues(g);
    out.push_back(std::move(g));
  }
  pend.clear();
}

void pend_add(std::vector<Pending> &pend, int mode, const Tree &T, const LayerRec &L,
              const DevLayer &d) {
  for (auto &p : pend)
    if (p.mode == mode && p.nl == L.nl && p.sph == L.sph) {
      p.layers.push_back(d);
      p.recs.emplace_back(&T, &L);
      return;
    }
  pend.push_back(Pending{mode, L.nl, L.sph, {d}, {{&T, &L}}});
}

// A nested linear root whose momentum launch also takes the next root
// step's continuity (see kMomentMass).
bool fuses_halves(const Tree &T) {
  const LayerRec &R = T.layers[1];
  return g_eng.fuse_halves && R.used && !R.children.empty() && !R.nl;
}

// A single-layer linear root (no sponge: it acts between steps) can take
// two root steps per launch.
bool two_step_root(const Tree &T) {
  const LayerRec &R = T.layers[1];
  return R.used && R.children.empty() && !R.nl && !R.sponge && T.ngauges == 0;
}

// Launch list of one root step of every scenario in `act`; rotates the
// host-side time levels as the launches will.  pair = -1: every layer one
// step per launch; 0 / 1: the first / second root step of a pair in which
// the two-step roots take one launch (emitted with the first, nothing with
// the second).
void emit_step(const std::vector<Scenario *> &act, std::vector<Group> &out, int pair = -1) {
  std::vector<Pending> pend;
  size_t maxsib = 0;
  for (Scenario *s : act) {
    Tree &T = s->tree;
    LayerRec &R = *T.layer(1);
    if (pair >= 0 && two_step_root(T)) {
      if (pair == 0) pend_add(pend, kTwoStep, T, R, dev_view(T, R, true, true));
    } else if (R.children.empty()) {
      pend_add(pend, kFused, T, R, dev_view(T, R, true, true));
    } else if (!T.mass_ahead) {  // else the previous momentum launch did it
      pend_add(pend, kMassOnly, T, R, dev_view(T, R, false, false));
    }
    maxsib = std::max(maxsib, R.children.size());
  }
  flush_steps(pend, out);
  for (size_t sib = 0; sib < maxsib; ++sib) {
    int maxrt = 0;
    for (Scenario *s : act) {
      LayerRec &R = *s->tree.layer(1);
      if (R.children.size() > sib)
        maxrt = std::max(maxrt, s->tree.layers[R.children[sib]].rel_time);
    }
    for (int k = 1; k <= maxrt; ++k) {
      Group jq;
      jq.type = kJnqGroup;
      jq.t = k;
      Group jz;
      jz.type = kJnzGroup;
      Group se, sz;  // the same phases of SC_OPTION = 1 children
      se.type = kScEdgeGroup;
      se.t = k;
      sz.type = kScJnzGroup;
      std::vector<LayerRec *> moved;
      for (Scenario *s : act) {
        Tree &T = s->tree;
        LayerRec &R = *T.layer(1);
        if (R.children.size() <= sib) continue;
        LayerRec &C = T.layers[R.children[sib]];
        if (k > C.rel_time) continue;
        if (C.sc) se.scs.push_back(sc_view(T, R, C));
        else jq.nests.push_back(nest_view(R, C));
        pend_add(pend, kFused, T, C, dev_view(T, C, true, false));
        moved.push_back(&C);
      }
      if (!jq.nests.empty()) out.push_back(std::move(jq));
      if (!se.scs.empty()) out.push_back(std::move(se));
      flush_steps(pend, out);
      for (Scenario *s : act) {
        Tree &T = s->tree;
        LayerRec &R = *T.layer(1);
        if (R.children.size() <= sib) continue;
        LayerRec &C = T.layers[R.children[sib]];
        if (k > C.rel_time) continue;
        const int nhalf = (int)std::floor(C.rel_time / 2.0f) + 1;
        if (k != nhalf) continue;
        if (C.sc) sz.scs.push_back(sc_view(T, R, C));
        else jz.nests.push_back(nest_view(R, C));
      }
      if (!jz.nests.empty()) out.push_back(std::move(jz));
      if (!sz.scs.empty()) out.push_back(std::move(sz));
      for (LayerRec *C : moved) C->cur = 1 - C->cur;  // UPDATE
    }
  }
  for (Scenario *s : act) {
    Tree &T = s->tree;
    LayerRec &R = *T.layer(1);
    if (R.children.empty()) continue;
    // a nested linear root's next continuity rides in its momentum launch
    // (a scenario that finishes keeps it in level 2, where nothing reads it)
    T.mass_ahead = fuses_halves(T);
    if (T.mass_ahead) pend_add(pend, kMomentMass, T, R, fused_halves_view(T, R));
    else pend_add(pend, kMomentOnly, T, R, dev_view(T, R, true, true));
  }
  flush_steps(pend, out);
  for (Scenario *s : act) {
    if (pair == 1 && two_step_root(s->tree)) continue;  // rotated with the pair's first step
    LayerRec &R = *s->tree.layer(1);
    R.cur = 1 - R.cur;
    s->tree.ctl_in = 1 - s->tree.ctl_in;
  }
}

// emit_step over the active scenarios in lanes that can run side by side:
// single-layer scenarios in lane 1, the nested ones in lanes 0 and 2 (with
// three streams; balanced by cell updates per root step, in scenario order).
void emit_lanes(const std::vector<Scenario *> &act, std::vector<Group> &out, int pair) {
  std::vector<Scenario *> single, nested[2];
  long long load[2] = {0, 0};
  for (Scenario *s : act) {
    const LayerRec &R = *s->tree.layer(1);
    if (R.children.empty()) {
      single.push_back(s);
      continue;
    }
    long long c = 0;
    for (const LayerRec &L : s->tree.layers)
      if (L.used) c += (long long)L.nx * L.ny * (L.id == 1 ? 1 : L.rel_time);
    const int q = g_eng.streams >= 3 && load[1] < load[0] ? 1 : 0;
    nested[q].push_back(s);
    load[q] += c;
  }
  auto add = [&](const std::vector<Scenario *> &v, int lane) {
    if (v.empty()) return;
    const size_t first = out.size();
    emit_step(v, out, pair);
    for (size_t k = first; k < out.size(); ++k) out[k].lane = lane;
  };
  add(nested[0], 0);
  add(single, g_eng.streams >= 2 ? 1 : 0);
  add(nested[1], 2);
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// engine_arena: the plan's descriptors go to the engine's arena, reused
// from plan to plan (the tuner's probes, the continuity prologue and the
// tail step run there too, each after the previous user has finished and
// before the next plan is built), instead of an allocation of its own.
int upload_plan(Plan &P, bool engine_arena = false) {
  size_t bytes = 0;
  int heads = 0;
  for (int r = 0; r < P.nsteps; ++r) {
    P.heads_off[r] = heads;
    for (auto &g : P.step[r]) {
      if (g.type == kStepGroup) {
        g.o_descs = bytes; bytes += al(g.descs.size() * sizeof(BatchDesc));
        g.o_items = bytes; bytes += al(g.items.size() * sizeof(BatchItem));
        g.o_qstart = bytes; bytes += al(sizeof(g.qstart));
        g.head = heads; heads += 8;
      } else if (g.type == kScEdgeGroup || g.type == kScJnzGroup) {
        // the views go to the kernels by value
      } else {
        g.start.assign(1, 0);
        for (auto &n : g.nests)
          g.start.push_back(g.start.back() +
                            (g.type == kJnqGroup ? jnq_block_count(n) : jnz_block_count(n)));
        g.o_nests = bytes; bytes += al(g.nests.size() * sizeof(DevNest));
        g.o_start = bytes; bytes += al(g.start.size() * sizeof(int));
      }
    }
    P.heads_n[r] = heads - P.heads_off[r];
  }
  const size_t o_heads = bytes;
  bytes += al((size_t)std::max(heads, 1) * sizeof(int));
  if (engine_arena) {
    if (g_eng.arena_bytes < bytes) {
      if (g_eng.arena) (void)hipFree(g_eng.arena);
      g_eng.arena = nullptr;
      g_eng.arena_bytes = 0;
      HIPCHK(hipMalloc(&g_eng.arena, bytes + bytes / 2));
      g_eng.arena_bytes = bytes + bytes / 2;
    }
    P.arena = g_eng.arena;
    P.owns_arena = false;
  } else {
    HIPCHK(hipMalloc(&P.arena, bytes));
  }
  std::vector<char> host(bytes, 0);
  for (int r = 0; r < P.nsteps; ++r)
    for (auto &g : P.step[r]) {
      if (g.type == kStepGroup) {
        std::memcpy(&host[g.o_descs], g.descs.data(), g.descs.size() * sizeof(BatchDesc));
        std::memcpy(&host[g.o_items], g.items.data(), g.items.size() * sizeof(BatchItem));
        std::memcpy(&host[g.o_qstart], g.qstart, sizeof(g.qstart));
      } else if (g.type == kScEdgeGroup || g.type == kScJnzGroup) {
        continue;
      } else {
        std::memcpy(&host[g.o_nests], g.nests.data(), g.nests.size() * sizeof(DevNest));
        std::memcpy(&host[g.o_start], g.start.data(), g.start.size() * sizeof(int));
      }
    }
  HIPCHK(hipMemcpyAsync(P.arena, host.data(), bytes, hipMemcpyHostToDevice, g_eng.st));
  HIPCHK(hipStreamSynchronize(g_eng.st));
  P.heads = reinterpret_cast<int *>(P.arena + o_heads);
  return 0;
}

BatchLaunch step_launch(const Plan &P, const Group &g) {
  BatchLaunch B;
  B.descs = reinterpret_cast<const BatchDesc *>(P.arena + g.o_descs);
  B.items = reinterpret_cast<const BatchItem *>(P.arena + g.o_items);
  B.qstart = reinterpret_cast<const int *>(P.arena + g.o_qstart);
  B.heads = P.heads + g.head;
  B.nitems = (int)g.items.size();
  return B;
}

int enqueue_group(const Plan &P, const Group &g, hipStream_t st) {
  hipError_t e = hipSuccess;
  {
    if (g.type == kStepGroup && g.mode == kTwoStep) {
      e = launch_batch2(step_launch(P, g), g.sph != 0, g.shape, st);
    } else if (g.type == kStepGroup) {
      e = launch_batch(step_launch(P, g), g.mode, g.nl != 0, g.sph != 0, g.shape, st);
    } else if (g.type == kScEdgeGroup || g.type == kScJnzGroup) {
      for (size_t q = 0; q < g.scs.size() && e == hipSuccess; ++q)
        e = g.type == kScEdgeGroup ? launch_sc_edges(g.scs[q], g.t, st)
                                   : launch_sc_jnz(g.scs[q], st);
    } else {
      const DevNest *n = reinterpret_cast<const DevNest *>(P.arena + g.o_nests);
      const int *s = reinterpret_cast<const int *>(P.arena + g.o_start);
      const int nn = (int)g.nests.size();
      e = g.type == kJnqGroup ? launch_jnq_multi(n, s, nn, g.start.back(), g.t, st)
                              : launch_jnz_multi(n, s, nn, g.start.back(), st);
    }
    if (e != hipSuccess) return fail(211, "batch launch: %s", hipGetErrorString(e));
    ++g_eng.launches;
  }
  return 0;
}

// Root steps r0..r1-1 of a plan.  The queue heads of all of them are
// cleared first (each group has heads of its own).  Each lane's groups then
// run on a stream of their own, all their steps in order: lane 0 on st,
// lanes 1 and 2 on st2 and st3, forked from st after the clearing and
// joined back at the end, so work enqueued on st afterwards (and its
// synchronisations) follows every lane.  A lane chains continuity halves,
// JNQ, children, JNZ and momentum halves of its scenarios as one stream
// would; lanes touch disjoint scenarios.
int enqueue_plan(const Plan &P, int r0, int r1) {
  hipStream_t st = g_eng.st;
  const int h0 = P.heads_off[r0], h1 = P.heads_off[r1 - 1] + P.heads_n[r1 - 1];
  hipError_t e = launch_fill_i32(P.heads + h0, h1 - h0, 0, st);
  if (e != hipSuccess) return fail(210, "batch head reset: %s", hipGetErrorString(e));
  bool used[3] = {false, false, false};
  for (int r = r0; r < r1; ++r)
    for (const auto &g : P.step[r]) used[g.lane] = true;
  hipStream_t lane_st[3] = {st, g_eng.st2, g_eng.st3};
  hipEvent_t join[3] = {nullptr, g_eng.join2, g_eng.join3};
  bool forked = false;
  for (int l = 1; l < 3; ++l) {
    if (!used[l]) continue;
    if (!forked) HIPCHK(hipEventRecord(g_eng.fork, st));
    forked = true;
    HIPCHK(hipStreamWaitEvent(lane_st[l], g_eng.fork, 0));
  }
  for (int l = 0; l < 3; ++l)
    for (int r = r0; r < r1; ++r)
      for (const auto &g : P.step[r])
        if (g.lane == l)
          if (int q = enqueue_group(P, g, lane_st[l])) return q;
  for (int l = 1; l < 3; ++l) {
    if (!used[l]) continue;
    HIPCHK(hipEventRecord(join[l], lane_st[l]));
    HIPCHK(hipStreamWaitEvent(st, join[l], 0));
  }
  return 0;
}

// Time each candidate tile shape of every fused group of one root step
// (dry launches: no max/arrival, no counter advance) and record the best.
int tune(const std::vector<Scenario *> &act) {
  std::vector<Group> probe;
  {
    // a throw-away emission to see which fused groups exist; time levels
    // are restored afterwards
    std::vector<std::vector<int>> cur;
    std::vector<int> ctl;
    std::vector<bool> ahead;
    for (Scenario *s : act) {
      std::vector<int> c;
      for (auto &L : s->tree.layers) c.push_back(L.cur);
      cur.push_back(c);
      ctl.push_back(s->tree.ctl_in);
      ahead.push_back(s->tree.mass_ahead);
    }
    emit_step(act, probe);
    for (size_t i = 0; i < act.size(); ++i) {
      for (int k = 0; k < kMaxLayers; ++k) act[i]->tree.layers[k].cur = cur[i][k];
      act[i]->tree.ctl_in = ctl[i];
      act[i]->tree.mass_ahead = ahead[i];
    }
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (auto &g : probe) {
    if (g.type != kStepGroup || g.mode != kFused) continue;
    const long long key = kind_key(g.nl, g.sph, g.cells);
    if (g_eng.tuned.count(key)) continue;
    if (trace_on())
      std::fprintf(stderr, "batch: tuning kind nl=%d sph=%d, %zu layers, %lld cells\n", g.nl,
                   g.sph, g.descs.size(), g.cells);
    int best = 0;
    float best_ms = 1e30f;
    for (int sh = 0; sh < batch_num_shapes(); ++sh) {
      int tx, ty;
      batch_shape(sh, &tx, &ty);
      Plan T;
      T.nsteps = 1;
      Group d;
      d.type = kStepGroup;
      d.mode = kFused;
      d.nl = g.nl;
      d.sph = g.sph;
      d.shape = sh;
      // views at the layers' current time levels (the probe's later
      // sub-steps saw rotated ones): read level 1, write level 2 only
      for (auto &tr : g.recs) add_layer_items(d, dev_view(*tr.first, *tr.second, false, false), tx, ty);
      split_queues(d);
      T.step[0].push_back(std::move(d));
      if (int r = upload_plan(T, true)) return r;
      float tot = 0;
      for (int rep = 0; rep < 4; ++rep) {
        if (rep == 1) HIPCHK(hipEventRecord(e0, g_eng.st));
        if (int r = enqueue_plan(T, 0, 1)) return r;
      }
      HIPCHK(hipEventRecord(e1, g_eng.st));
      HIPCHK(hipEventSynchronize(e1));
      HIPCHK(hipEventElapsedTime(&tot, e0, e1));
      free_plan(T);
      if (tot < best_ms) { best_ms = tot; best = sh; }
    }
    g_eng.tuned[key] = best;
    g_eng.tuned_ms[key] = best_ms / 3;
  }
  // Two root steps per launch for the single-layer linear roots: every
  // two-step shape is timed on them (dry, as above) and the best is kept
  // if it beats two launches of their tuned one-step shape.
  if (g_eng.two_step && g_eng.forced_step2 < 0) {
    std::vector<std::pair<const Tree *, const LayerRec *>> roots;
    long long cells = 0;
    int sph = -1;
    bool mixed = false;
    for (Scenario *s : act)
      if (two_step_root(s->tree)) {
        const LayerRec &R = *s->tree.layer(1);
        roots.emplace_back(&s->tree, &R);
        cells += R.nxy();
        if (sph >= 0 && sph != R.sph) mixed = true;
        sph = R.sph;
      }
    const long long key2 = (cells >= (4LL << 20) ? 1 : 0) + (mixed ? 4 : 2 * sph);
    if (!roots.empty() && !g_eng.tuned2.count(key2)) {
      // the one-step time of the same roots: their fused group's tuned shape
      float one = 0;
      for (auto &g : probe)
        if (g.type == kStepGroup && g.mode == kFused && g.nl == 0 && !g.recs.empty() &&
            g.recs[0].second->id == 1)  // the first flush: outermost layers only
          one += g_eng.tuned_ms.count(kind_key(0, g.sph, g.cells))
                     ? g_eng.tuned_ms[kind_key(0, g.sph, g.cells)]
                     : 0.0f;
      int best = -1;
      float best_ms = 2 * one;  // two one-step launches
      for (int sh = 0; sh < kStep2Shapes && one > 0; ++sh) {
        int tx, ty;
        step2_shape(sh, &tx, &ty);
        Plan T;
        T.nsteps = 1;
        for (int part = 0; part < 2; ++part) {  // the spherical and the Cartesian ones
          Group d;
          d.type = kStepGroup;
          d.mode = kTwoStep;
          d.nl = 0;
          d.sph = 1 - part;
          d.shape = sh;
          for (auto &tr : roots)
            if (tr.second->sph == d.sph)
              add_layer_items(d, dev_view(*tr.first, *tr.second, false, false), tx, ty);
          if (d.descs.empty()) continue;
          split_queues(d);
          T.step[0].push_back(std::move(d));
        }
        if (int r = upload_plan(T, true)) return r;
        float tot = 0;
        for (int rep = 0; rep < 4; ++rep) {
          if (rep == 1) HIPCHK(hipEventRecord(e0, g_eng.st));
          if (int r = enqueue_plan(T, 0, 1)) return r;
        }
        HIPCHK(hipEventRecord(e1, g_eng.st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&tot, e0, e1));
        free_plan(T);
        if (tot / 3 < best_ms) { best_ms = tot / 3; best = sh; }
      }
      g_eng.tuned2[key2] = best;
      if (trace_on())
        std::fprintf(stderr, "batch: two-step roots, %zu layers: one-step %.3f ms x 2, "
                     "two-step shape %d %.3f ms\n", roots.size(), one, best, best_ms);
    }
    if (!roots.empty()) g_eng.step2_shape = g_eng.tuned2[key2];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

bool any_two_step(const std::vector<Scenario *> &act) {
  if (!g_eng.two_step || g_eng.step2_shape < 0) return false;
  for (Scenario *s : act)
    if (two_step_root(s->tree)) return true;
  return false;
}

int build_plan(const std::vector<Scenario *> &act) {
  invalidate();
  g_eng.plan_policy = g_store_policy;
  if (g_eng.autotune && g_eng.forced_shape < 0)
    if (int r = tune(act)) return r;
  // nested linear roots without their next continuity in level 2 get it
  // now, outside the plan, so that every step of the plan has the same
  // launches (each takes the following step's continuity along)
  {
    Plan pre;
    pre.nsteps = 1;
    std::vector<Pending> pend;
    for (Scenario *s : act)
      if (fuses_halves(s->tree) && !s->tree.mass_ahead) {
        Tree &T = s->tree;
        pend_add(pend, kMassOnly, T, *T.layer(1), dev_view(T, *T.layer(1), false, false));
        T.mass_ahead = true;
      }
    flush_steps(pend, pre.step[0]);
    if (!pre.step[0].empty()) {
      int r = upload_plan(pre, true);
      if (!r) r = enqueue_plan(pre, 0, 1);
      if (!r) r = hipStreamSynchronize(g_eng.st) == hipSuccess ? 0 : fail(211, "batch continuity");
      free_plan(pre);
      if (r) return r;
    }
  }
  Plan &P = g_eng.plan;
  // two steps, or two pairs when two-step roots are present: either way
  // every time level is back where it started at the end of the plan
  const bool pairs = any_two_step(act);
  P.nsteps = pairs ? 4 : 2;
  for (int k = 0; k < P.nsteps; ++k) emit_lanes(act, P.step[k], pairs ? k % 2 : -1);
  if (trace_on())
    std::fprintf(stderr, "batch: plan of %d steps, %zu + %zu launches\n", P.nsteps,
                 P.step[0].size(), P.step[1].size());
  long long cells = 0;
  for (int k = 0; k < P.nsteps; ++k)
    for (auto &g : P.step[k])
      if (g.type == kStepGroup) cells += g.cells * (g.mode == kTwoStep ? 2 : 1);
  P.cells_per_step = cells / P.nsteps;
  if (int r = upload_plan(P, true)) return r;
  size_t per_step = 0;
  for (int k = 0; k < P.nsteps; ++k) per_step = std::max(per_step, P.step[k].size());
  if (!g_eng.graphs || per_step + 1 > (size_t)kGraphMaxLaunches) {
    // a very deep sub-stepping (huge REL_TIME) is enqueued step by step:
    // graphs of that many nodes are not worth building
    g_eng.plan_valid = true;
    ++g_eng.plans_built;
    return 0;
  }
  for (int r = 0; r < P.nsteps; ++r)
    for (const auto &g : P.step[r])
      if (g.type == kStepGroup && g.mode != kTwoStep)
        HIPCHK(launch_batch(step_launch(P, g), g.mode, g.nl != 0, g.sph != 0, g.shape,
                            g_eng.st, true));
  HIPCHK(hipStreamBeginCapture(g_eng.st, hipStreamCaptureModeThreadLocal));
  int r = enqueue_plan(P, 0, P.nsteps);
  hipError_t ce = hipStreamEndCapture(g_eng.st, &P.graph);
  if (r) return r;
  HIPCHK(ce);
  HIPCHK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
  g_eng.plan_valid = true;
  ++g_eng.plans_built;
  return 0;
}

bool trace_on() {
  static const int on = std::getenv("GCOMCOT_BATCH_TRACE") ? 1 : 0;
  return on != 0;
}

// Advance every active scenario by up to n root steps.
int run(long long n) {
  while (n > 0) {
    std::vector<Scenario *> act = active();
    if (act.empty()) break;
    long long m = n;
    for (Scenario *s : act) m = std::min(m, s->total - s->done);
    if (g_eng.plan_valid && g_eng.plan_policy != g_store_policy) invalidate();
    if (trace_on())
      std::fprintf(stderr, "batch: %zu active, epoch of %lld steps, plan %s\n", act.size(), m,
                   g_eng.plan_valid ? "kept" : "rebuilt");
    if (!g_eng.plan_valid)
      if (int r = build_plan(act)) return r;
    const Plan &P = g_eng.plan;
    const int ns = P.nsteps;
    for (long long k = 0; k + ns <= m; k += ns) {
      if (P.exec) {
        HIPCHK(hipGraphLaunch(P.exec, g_eng.st));
        ++g_eng.graph_replays;
      } else {
        if (int r = enqueue_plan(P, 0, ns)) return r;
      }
    }
    long long rest = m % ns;
    if (rest) {
      // the rest of the epoch from the plan's starting time levels: its
      // first pair of steps (two steps), or its first step (one); the
      // levels then differ from the captured ones, so rebuild next time
      std::vector<Group> scratch;
      if (rest >= 2) {
        if (int r = enqueue_plan(P, 0, 2)) return r;
        emit_step(act, scratch, ns == 4 ? 0 : -1);  // rotate host-side levels as they did
        emit_step(act, scratch, ns == 4 ? 1 : -1);
        rest -= 2;
      }
      if (rest == 1) {
        if (ns == 2 && scratch.empty()) {
          if (int r = enqueue_plan(P, 0, 1)) return r;
          emit_step(act, scratch);
        } else {
          // one step of every layer, two-step roots included: a plan of its own
          HIPCHK(hipStreamSynchronize(g_eng.st));
          Plan one;
          one.nsteps = 1;
          emit_step(act, one.step[0]);
          int r = upload_plan(one, true);
          if (!r) r = enqueue_plan(one, 0, 1);
          if (!r) r = hipStreamSynchronize(g_eng.st) == hipSuccess ? 0 : fail(211, "batch tail step");
          free_plan(one);
          if (r) return r;
        }
      }
      HIPCHK(hipStreamSynchronize(g_eng.st));
      invalidate();
    }
    for (Scenario *s : act) s->done += m;
    for (Scenario *s : act)
      if (s->finished()) { invalidate(); break; }
    n -= m;
  }
  HIPCHK(hipStreamSynchronize(g_eng.st));
  return 0;
}

}  // namespace

extern "C" {

int gcomcot_batch_reset_(void) {
  if (int e = eng_init()) return e;
  invalidate();
  (void)hipStreamSynchronize(g_eng.st);
  (void)hipStreamSynchronize(g_eng.st2);
  (void)hipStreamSynchronize(g_eng.st3);
  for (auto &kv : g_eng.scen) tree_free(kv.second->tree);
  g_eng.scen.clear();
  if (g_eng.arena) (void)hipFree(g_eng.arena);
  g_eng.arena = nullptr;
  g_eng.arena_bytes = 0;
  g_eng.launches = g_eng.graph_replays = g_eng.plans_built = 0;
  return 0;
}

int gcomcot_batch_new_scenario_(long long *steps, int *sid) {
  if (*steps < 0) return fail(30, "batch: negative step count");
  auto s = std::make_unique<Scenario>();
  s->sid = g_eng.next_sid++;
  s->total = *steps;
  s->tree.arr_thr = g_eng.arr_thr;
  *sid = s->sid;
  g_eng.scen[s->sid] = std::move(s);
  return 0;
}

int gcomcot_batch_add_layer_(int *sid, int *layerid, int *parent, int *laycord, int *laygov,
                             int *rel_size, int *rel_time, int *corners, int *fric_switch,
                             float *fric_coef, float *dt, float *dx, float *dy, float *coef,
                             float *H, float *Z, int *nx, int *ny) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  if (*layerid != 1 && *parent != 1)
    return fail(32, "batch: layer %d: only one level of nesting below the root", *layerid);
  if (*laycord != 0 && *laycord != 1) return fail(10, "LAYCORD %d not 0/1", *laycord);
  if (!(*laygov == 0 || *laygov == 1 || *laygov == 9 || (*laycord == 1 && *laygov == 5)))
    return fail(11, "LAYGOV %d not supported", *laygov);
  if (*fric_switch != 0 && *fric_switch != 1)
    return fail(12, "FRIC_SWITCH %d not supported", *fric_switch);
  if (*dt <= 0.0f || *dx <= 0.0f || *dy <= 0.0f) return fail(13, "DT/DX/DY must be positive");
  if (*laycord == 0 && !coef) return fail(14, "spherical layer needs the PARA row table");
  if (int e = tree_validate_layer(s->tree, *layerid, *layerid == 1 ? -1 : *parent, *rel_size,
                                  *rel_time, corners, *nx, *ny))
    return e;
  if (int e = eng_init()) return e;
  invalidate();
  const int level = *layerid == 1 ? 1 : 2;
  if (int e = tree_create_layer(s->tree, g_eng.st, *layerid, *layerid == 1 ? -1 : *parent,
                                level, *laycord, *laygov, *rel_size, *rel_time, corners,
                                *fric_switch, *fric_coef, *dt, *dx, *dy, *nx, *ny, H, Z))
    return e;
  LayerRec &L = s->tree.layers[*layerid];
  std::vector<float> c((size_t)kNumCoef * *ny, 0.0f);
  if (*laycord == 0) std::memcpy(c.data(), coef, c.size() * sizeof(float));
  return tree_upload_coef(L, c);
}

int gcomcot_batch_add_layer_sc_(int *sid, int *layerid, int *parent, int *laygov,
                                int *rel_size, int *rel_time, int *corners, int *fric_switch,
                                float *fric_coef, float *dt, float *dx, float *dy, float *H,
                                float *Z, int *nx, int *ny, int *pos, float *cxy) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  if (!(*laygov == 0 || *laygov == 1 || *laygov == 5 || *laygov == 9))
    return fail(11, "LAYGOV %d not supported", *laygov);
  if (*fric_switch != 0 && *fric_switch != 1)
    return fail(12, "FRIC_SWITCH %d not supported", *fric_switch);
  if (*dt <= 0.0f || *dx <= 0.0f || *dy <= 0.0f) return fail(13, "DT/DX/DY must be positive");
  if (!pos || !cxy) return fail(84, "SC: layer %d needs POS and CXY", *layerid);
  if (int e = tree_validate_layer(s->tree, *layerid, *parent, *rel_size, *rel_time, corners, *nx,
                                  *ny, true))
    return e;
  if (int e = eng_init()) return e;
  invalidate();
  if (int e = tree_create_layer(s->tree, g_eng.st, *layerid, *parent, 2, 1, *laygov, *rel_size,
                                *rel_time, corners, *fric_switch, *fric_coef, *dt, *dx, *dy, *nx,
                                *ny, H, Z, true))
    return e;
  std::vector<float> c((size_t)kNumCoef * *ny, 0.0f);
  int e = tree_upload_coef(s->tree.layers[*layerid], c);
  if (!e) e = tree_set_sc(s->tree, g_eng.st, *layerid, pos, cxy);
  if (e) tree_drop_layer(s->tree, *layerid);  // no half-registered SC layer
  return e;
}

int gcomcot_batch_run_(long long *nsteps, int *nfinished) {
  if (int e = eng_init()) return e;
  if (int r = run(*nsteps)) return r;
  int f = 0;
  for (auto &kv : g_eng.scen) f += kv.second->finished() ? 1 : 0;
  if (nfinished) *nfinished = f;
  return 0;
}

int gcomcot_batch_finished_(int *sids, int *cap) {
  int n = 0;
  for (auto &kv : g_eng.scen)
    if (kv.second->finished() && kv.second->tree.layer(1)) {
      if (sids && n < *cap) sids[n] = kv.first;
      ++n;
    }
  return n;
}

int gcomcot_batch_progress_(int *sid, long long *done, long long *total) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  *done = s->done;
  *total = s->total;
  return 0;
}

int gcomcot_batch_get_state_(int *sid, int *lid, float *z, float *m, float *n) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  LayerRec *L = s->tree.layer(*lid);
  if (!L) return fail(33, "batch: scenario %d has no layer %d", *sid, *lid);
  HIPCHK(hipStreamSynchronize(g_eng.st));
  const size_t b = L->nxy() * sizeof(float);
  if (z) HIPCHK(hipMemcpy(z, L->z[L->cur], b, hipMemcpyDeviceToHost));
  if (m) HIPCHK(hipMemcpy(m, L->m[L->cur], b, hipMemcpyDeviceToHost));
  if (n) HIPCHK(hipMemcpy(n, L->n[L->cur], b, hipMemcpyDeviceToHost));
  return 0;
}

int gcomcot_batch_get_max_(int *sid, int *lid, float *zmax_wet, int *arrival) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  LayerRec *L = s->tree.layer(*lid);
  if (!L) return fail(33, "batch: scenario %d has no layer %d", *sid, *lid);
  HIPCHK(hipStreamSynchronize(g_eng.st));
  if (zmax_wet)
    HIPCHK(hipMemcpy(zmax_wet, L->zmax_wet, L->nxy() * sizeof(float), hipMemcpyDeviceToHost));
  if (arrival)
    HIPCHK(hipMemcpy(arrival, L->arrival, L->nxy() * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int gcomcot_batch_checkpoint_save_(int *sid, const char *path, int *pathlen) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  if (!path || *pathlen <= 0) return fail(70, "checkpoint: empty path");
  if (int e = eng_init()) return e;
  const std::string p(path, (size_t)*pathlen);
  return tree_checkpoint_save(s->tree, g_eng.st, p.c_str(), s->done);
}

int gcomcot_batch_checkpoint_load_(int *sid, const char *path, int *pathlen) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  if (!path || *pathlen <= 0) return fail(72, "checkpoint: empty path");
  if (int e = eng_init()) return e;
  invalidate();
  const std::string p(path, (size_t)*pathlen);
  long long done = 0;
  if (int e = tree_checkpoint_load(s->tree, g_eng.st, p.c_str(), &done)) return e;
  s->done = done;
  return 0;
}

int gcomcot_batch_set_coords_(int *sid, int *lid, const float *x, const float *y) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  LayerRec *L = s->tree.layer(*lid);
  if (!L) return fail(33, "batch: scenario %d has no layer %d", *sid, *lid);
  s->tree.xs[*lid].assign(x, x + L->nx);
  s->tree.ys[*lid].assign(y, y + L->ny);
  return 0;
}

int gcomcot_batch_fault_deform_(int *sid, const float *params, int *fswitch) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  if (int e = eng_init()) return e;
  // the plan may take this scenario's next continuity as already done;
  // the deformation changes level 1, so that continuity is computed again
  invalidate();
  HIPCHK(hipStreamSynchronize(g_eng.st));
  return tree_apply_fault(s->tree, g_eng.st, params, *fswitch);
}

int gcomcot_batch_remove_(int *sid) {
  Scenario *s = scenario(*sid);
  if (!s) return fail(31, "batch: no scenario %d", *sid);
  invalidate();
  (void)hipStreamSynchronize(g_eng.st);
  tree_free(s->tree);
  g_eng.scen.erase(*sid);
  return 0;
}

long long gcomcot_batch_device_bytes_(int *sid) {
  if (*sid == 0) {
    long long b = 0;
    for (auto &kv : g_eng.scen) b += (long long)kv.second->tree.device_bytes();
    return b;
  }
  Scenario *s = scenario(*sid);
  return s ? (long long)s->tree.device_bytes() : -1;
}

long long gcomcot_layer_footprint_(int *nx, int *ny, int *nonlinear) {
  return (long long)layer_bytes(*nx, *ny, *nonlinear != 0, false);
}

void gcomcot_batch_set_arrival_threshold_(float *thr) { g_eng.arr_thr = *thr; }

int gcomcot_batch_set_tuning_(int *autotune, int *shape) {
  if (*shape >= batch_num_shapes()) return fail(34, "batch: tile shape %d out of range", *shape);
  g_eng.autotune = *autotune != 0;
  g_eng.forced_shape = *shape;
  invalidate();
  return 0;
}

int gcomcot_batch_set_graphs_(int *on) {
  g_eng.graphs = *on != 0;
  invalidate();
  return 0;
}

int gcomcot_batch_set_two_step_(int *on, int *shape) {
  if (shape && (*shape < 0 || *shape >= kStep2Shapes))
    return fail(44, "batch: two-step shape %d out of range", *shape);
  g_eng.two_step = *on != 0;
  g_eng.forced_step2 = shape ? *shape : -1;
  g_eng.step2_shape = shape ? *shape : 4;
  invalidate();
  return 0;
}

int gcomcot_batch_set_fused_halves_(int *on) {
  g_eng.fuse_halves = *on != 0;
  invalidate();
  return 0;
}

int gcomcot_batch_set_streams_(int *n) {
  if (*n < 1 || *n > 3) return fail(45, "batch: %d streams (1 to 3)", *n);
  g_eng.streams = *n;
  invalidate();
  return 0;
}

int gcomcot_batch_info_(long long *launches, long long *graph_replays, long long *plans,
                        long long *cells_per_step, int *tuned_count) {
  if (launches) *launches = g_eng.launches;
  if (graph_replays) *graph_replays = g_eng.graph_replays;
  if (plans) *plans = g_eng.plans_built;
  if (cells_per_step) *cells_per_step = g_eng.plan_valid ? g_eng.plan.cells_per_step : 0;
  if (tuned_count) *tuned_count = (int)g_eng.tuned.size();
  return 0;
}

}  // extern "C"
