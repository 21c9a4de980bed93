// CPU stand-ins for the launchers of kernels.h with checkable semantics.
//
// Every plane value of the emulated state counts steps: a continuity pass
// sets Z2 = Z1 + 1, a momentum pass M2 = M1 + 1 and N2 = N1 + 1 (checking
// that Z2 = Z1 + 1 already holds, i.e. continuity ran first), the wet max
// counts the passes that carried it and tile 0 of an advancing pass moves
// the root-step counter (mirrored into arrival[0]).  Each launch checks
// that its tiles cover every cell of every layer exactly once, that queue
// heads were cleared, and JNQ/JNZ check the sub-step they run in against
// the parent's and the child's step counts (all_grids.f90 order).  A
// violated check aborts the process.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "source.h"

using namespace gc;

namespace {

[[noreturn]] void die(const char *m) {
  std::fprintf(stderr, "emulated kernel check failed: %s\n", m);
  std::abort();
}

// kernels.hip kBatchShapes; shapes 3-5 are the same tiles in the other launch form
const int kShapes[][2] = {{64, 8}, {32, 16}, {64, 16}};

long long g_fused_tiles = 0;  // tiles of momentum + next continuity launches

void tile_op(const DevLayer &d, int mode, int t, int tiles_x, int TX, int TY,
             std::vector<unsigned char> &cov) {
  const int i0 = (t % tiles_x) * TX, j0 = (t / tiles_x) * TY;
  for (int j = j0; j < j0 + TY && j < d.ny; ++j)
    for (int i = i0; i < i0 + TX && i < d.nx; ++i) {
      const size_t c = (size_t)j * d.nx + i;
      cov[c]++;
      if (mode == kMomentMass) {  // z1 = Z(k), z2 = the buffer of Z(k-1), then Z(k+1)
        if (i == i0 && j == j0) ++g_fused_tiles;
        if (d.z1[c] != d.z2[c] + 1.0f) die("fused momentum before continuity");
        if (!d.is_root) die("fused halves of a nested layer");
        d.m2[c] = d.m1[c] + 1.0f;
        d.n2[c] = d.n1[c] + 1.0f;
        if (d.zmax) d.zmax[c] += 1.0f;
        d.z2[c] = d.z1[c] + 1.0f;  // the next root step's continuity
        continue;
      }
      if (mode != kMomentOnly) d.z2[c] = d.z1[c] + 1.0f;
      if (mode == kMomentOnly && d.z2[c] != d.z1[c] + 1.0f) die("momentum before continuity");
      if (mode != kMassOnly) {
        d.m2[c] = d.m1[c] + 1.0f;
        d.n2[c] = d.n1[c] + 1.0f;
        if (d.zmax) d.zmax[c] += 1.0f;
      }
    }
  if (mode != kMassOnly && d.zmax && d.advance && t == 0) {
    d.ctl[1 - d.ctl_in] = d.ctl[d.ctl_in] + 1;
    d.arrival[0] = d.ctl[1 - d.ctl_in];
  }
}

struct StepArg {
  DevLayer d;
  int mode;
};

void run_step(void *a) {
  const StepArg *s = (const StepArg *)a;
  const int tx = (s->d.nx + 63) / 64, ty = (s->d.ny + 7) / 8;
  std::vector<unsigned char> cov((size_t)s->d.nx * s->d.ny, 0);
  for (int t = 0; t < tx * ty; ++t) tile_op(s->d, s->mode, t, tx, 64, 8, cov);
  for (auto c : cov)
    if (c != 1) die("layer launch coverage");
}

struct BatchArg {
  BatchLaunch B;
  int mode, shape;
};

void run_batch(void *a) {
  const BatchArg *x = (const BatchArg *)a;
  const BatchLaunch &B = x->B;
  if (std::getenv("EMU_TRACE")) std::fprintf(stderr, "emu: batch launch mode %d, %d items\n", x->mode, B.nitems);
  if (B.qstart[0] != 0 || B.qstart[8] != B.nitems) die("queue bounds");
  for (int q = 0; q < 8; ++q) {
    if (B.qstart[q] > B.qstart[q + 1]) die("queue order");
    if (B.heads[q] != 0) die("queue head not cleared before the launch");
  }
  int TX = 64, TY = 8;
  if (x->mode == kFused) TX = kShapes[x->shape % 3][0], TY = kShapes[x->shape % 3][1];
  int nslot = 0;
  for (int it = 0; it < B.nitems; ++it) nslot = std::max(nslot, B.items[it].slot + 1);
  std::vector<std::vector<unsigned char>> cov(nslot);
  for (int s = 0; s < nslot; ++s) {
    const BatchDesc &bd = B.descs[s];
    if (bd.tiles_x != (bd.L.nx + TX - 1) / TX || bd.ntiles != bd.tiles_x * ((bd.L.ny + TY - 1) / TY))
      die("descriptor tiling");
    for (int r = 0; r < s; ++r)
      if (B.descs[r].L.z2 == bd.L.z2) die("layer twice in one launch");
    cov[s].assign((size_t)bd.L.nx * bd.L.ny, 0);
  }
  for (int it = 0; it < B.nitems; ++it) {
    const BatchItem w = B.items[it];
    if (w.slot < 0 || w.tile < 0 || w.tile >= B.descs[w.slot].ntiles) die("item out of range");
    tile_op(B.descs[w.slot].L, x->mode, w.tile, B.descs[w.slot].tiles_x, TX, TY, cov[w.slot]);
  }
  for (auto &c : cov)
    for (auto v : c)
      if (v != 1) die("batch launch coverage");
  for (int q = 0; q < 8; ++q) B.heads[q] = B.qstart[q + 1] - B.qstart[q] + 8;
}

int nhalf(int rt) { return (int)std::floor(rt / 2.0f) + 1; }

void check_jnq(const DevNest &N, int t) {
  if (t < 1 || t > N.rel_time) die("JNQ sub-step out of range");
  const float ps = N.pm1[0];  // parent root steps done
  for (int j = N.js; j <= N.je; ++j)
    for (int i = N.is; i <= N.ie; ++i) {
      const size_t c = (size_t)(j - 1) * N.pnx + (i - 1);
      if (N.pz2[c] != N.pm1[c] + 1.0f) die("JNQ before the parent's continuity");
    }
  const size_t cn = (size_t)N.cnx * N.cny;
  for (size_t c = 0; c < cn; ++c)
    if (N.cz1[c] != N.cm1[c] || N.cz1[c] != N.cn1[c]) {
      std::fprintf(stderr, "cell %zu of %zu (cnx %d): z1 %g m1 %g n1 %g, t=%d RT=%d parent steps %g\n", c,
                   cn, N.cnx, N.cz1[c], N.cm1[c], N.cn1[c], t, N.rel_time, ps);
      die("child level 1 inconsistent");
    }
  if (N.cz1[0] != ps * N.rel_time + (t - 1)) {
    std::fprintf(stderr, "t=%d cz1=%g pm1=%g RT=%d cnx=%d\n", t, N.cz1[0], ps, N.rel_time, N.cnx);
    die("JNQ in the wrong sub-step");
  }
}

void check_jnz(const DevNest &N) {
  if (N.pz2_out != N.pz2) die("JNZ target is not the parent's level 2");
  const size_t cn = (size_t)N.cnx * N.cny;
  for (size_t c = 0; c < cn; ++c)
    if (N.cz2[c] != N.cz1[c] + 1.0f) die("JNZ before the child's step");
  if (N.cz1[0] - N.pm1[0] * N.rel_time != nhalf(N.rel_time) - 1) die("JNZ in the wrong sub-step");
}

struct NestArg {
  DevNest N;
  int t;
};
void run_jnq(void *a) { check_jnq(((NestArg *)a)->N, ((NestArg *)a)->t); }
void run_jnz(void *a) { check_jnz(((NestArg *)a)->N); }

struct MultiArg {
  const DevNest *nests;
  const int *start;
  int nnest, nblocks, t, jnq;
};
void run_multi(void *a) {
  const MultiArg *m = (const MultiArg *)a;
  if (std::getenv("EMU_TRACE")) std::fprintf(stderr, "emu: %s of %d nests\n", m->jnq ? "jnq" : "jnz", m->nnest);
  if (m->start[0] != 0 || m->start[m->nnest] != m->nblocks) die("nest block table");
  for (int n = 0; n < m->nnest; ++n) {
    const int want = m->jnq ? jnq_block_count(m->nests[n]) : jnz_block_count(m->nests[n]);
    if (m->start[n + 1] - m->start[n] != want) die("nest block share");
    if (m->jnq) check_jnq(m->nests[n], m->t);
    else check_jnz(m->nests[n]);
  }
}

// two root steps per launch (k_batch2_grid): every plane moves by 2, the
// maximum takes both steps and the counter advances by 2
const int kStep2Tiles[][2] = {{64, 8}, {32, 16}, {64, 16}, {64, 16}, {64, 16}, {32, 16}, {64, 32}};
struct Batch2Arg {
  BatchLaunch B;
  int shape;
};
void run_batch2(void *a) {
  const Batch2Arg *x = (const Batch2Arg *)a;
  const BatchLaunch &B = x->B;
  const int TX = kStep2Tiles[x->shape][0], TY = kStep2Tiles[x->shape][1];
  int nslot = 0;
  for (int it = 0; it < B.nitems; ++it) nslot = std::max(nslot, B.items[it].slot + 1);
  std::vector<std::vector<unsigned char>> cov(nslot);
  for (int s = 0; s < nslot; ++s) {
    const BatchDesc &bd = B.descs[s];
    if (bd.tiles_x != (bd.L.nx + TX - 1) / TX || bd.ntiles != bd.tiles_x * ((bd.L.ny + TY - 1) / TY))
      die("two-step descriptor tiling");
    if (!bd.L.is_root) die("two-step launch of a nested layer");
    cov[s].assign((size_t)bd.L.nx * bd.L.ny, 0);
  }
  for (int it = 0; it < B.nitems; ++it) {
    const BatchItem w = B.items[it];
    const BatchDesc &bd = B.descs[w.slot];
    const DevLayer &d = bd.L;
    const int i0 = (w.tile % bd.tiles_x) * TX, j0 = (w.tile / bd.tiles_x) * TY;
    for (int j = j0; j < j0 + TY && j < d.ny; ++j)
      for (int i = i0; i < i0 + TX && i < d.nx; ++i) {
        const size_t c = (size_t)j * d.nx + i;
        cov[w.slot][c]++;
        d.z2[c] = d.z1[c] + 2.0f;
        d.m2[c] = d.m1[c] + 2.0f;
        d.n2[c] = d.n1[c] + 2.0f;
        if (d.zmax) d.zmax[c] += 2.0f;
      }
    if (d.zmax && d.advance && w.tile == 0) {
      d.ctl[1 - d.ctl_in] = d.ctl[d.ctl_in] + 2;
      d.arrival[0] = d.ctl[1 - d.ctl_in];
    }
  }
  for (auto &c : cov)
    for (auto v : c)
      if (v != 1) die("two-step launch coverage");
}

struct FillArg {
  int *p;
  int n, v;
};
void run_fill(void *a) {
  const FillArg *f = (const FillArg *)a;
  for (int k = 0; k < f->n; ++k) f->p[k] = f->v;
}

}  // namespace

namespace gc {

hipError_t launch_step(const DevLayer &L, int mode, bool, bool, int, hipStream_t st) {
  StepArg a{L, mode};
  emu_submit(st, run_step, &a, sizeof(a));
  return hipSuccess;
}
hipError_t launch_step2(const DevLayer &L, bool, int, hipStream_t st) {
  StepArg a{L, kFused};
  emu_submit(st, run_step, &a, sizeof(a));  // the planner tests never take this path
  return hipSuccess;
}
hipError_t launch_gauges(const GaugeDev *, int, const ZNow &, const int *, float *, int,
                         hipStream_t) {
  return hipSuccess;  // gauges belong to the single-tree device loop, not the batch planner
}
hipError_t launch_sc_edges(const DevSc &, int, hipStream_t) {
  return hipSuccess;  // SC_OPTION 1 children belong to the single-tree device loop
}
hipError_t launch_sc_jnz(const DevSc &, hipStream_t) { return hipSuccess; }
int tree_apply_fault(Tree &, hipStream_t, const float *, int) {
  return 1;  // fault sources run on the device only; the planner tests do not use them
}
int num_tile_variants() { return 1; }
void tile_variant_shape(int, int *tx, int *ty) { *tx = 64; *ty = 8; }
hipError_t launch_max(const DevLayer &, float *, hipStream_t) { return hipSuccess; }
hipError_t launch_jnq(const DevNest &N, int t, hipStream_t st) {
  NestArg a{N, t};
  emu_submit(st, run_jnq, &a, sizeof(a));
  return hipSuccess;
}
hipError_t launch_jnz(const DevNest &N, hipStream_t st) {
  NestArg a{N, 0};
  emu_submit(st, run_jnz, &a, sizeof(a));
  return hipSuccess;
}
int batch_num_shapes() { return 6; }
void batch_shape(int k, int *tx, int *ty) { *tx = kShapes[k % 3][0]; *ty = kShapes[k % 3][1]; }
hipError_t launch_batch(const BatchLaunch &B, int mode, bool, bool, int shape, hipStream_t st,
                        bool prepare) {
  if (prepare || B.nitems <= 0) return hipSuccess;
  if (shape < 0 || shape >= 6) return hipErrorInvalidValue;
  BatchArg a{B, mode, shape};
  emu_submit(st, run_batch, &a, sizeof(a));
  return hipSuccess;
}
void step2_shape(int shape, int *tx, int *ty) {
  *tx = kStep2Tiles[shape][0];
  *ty = kStep2Tiles[shape][1];
}
hipError_t launch_batch2(const BatchLaunch &B, bool, int shape, hipStream_t st) {
  if (B.nitems <= 0) return hipSuccess;
  if (shape < 0 || shape >= kStep2Shapes) return hipErrorInvalidValue;
  Batch2Arg a{B, shape};
  emu_submit(st, run_batch2, &a, sizeof(a));
  return hipSuccess;
}
int jnq_block_count(const DevNest &N) {
  return (2 * (N.je - N.js + 5) + 2 * (N.ie - N.is + 5) + 127) / 128;
}
int jnz_block_count(const DevNest &N) { return ((N.ie - N.is + 1) * (N.je - N.js + 1) + 127) / 128; }
hipError_t launch_jnq_multi(const DevNest *nests, const int *start, int nnest, int nblocks, int t,
                            hipStream_t st) {
  if (nnest <= 0) return hipSuccess;
  MultiArg a{nests, start, nnest, nblocks, t, 1};
  emu_submit(st, run_multi, &a, sizeof(a));
  return hipSuccess;
}
hipError_t launch_jnz_multi(const DevNest *nests, const int *start, int nnest, int nblocks,
                            hipStream_t st) {
  if (nnest <= 0) return hipSuccess;
  MultiArg a{nests, start, nnest, nblocks, 0, 0};
  emu_submit(st, run_multi, &a, sizeof(a));
  return hipSuccess;
}
hipError_t launch_fill_i32(int *p, int n, int v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  FillArg a{p, n, v};
  emu_submit(st, run_fill, &a, sizeof(a));
  return hipSuccess;
}

}  // namespace gc

long long emu_fused_tiles() { return g_fused_tiles; }
