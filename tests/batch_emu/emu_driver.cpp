// Drives the batch engine (batch.hip + tree.hip, built for the CPU against
// the runtime and kernel stand-ins of this folder) through random
// scenario mixes: scenarios of random sizes, nests (REL_SIZE 2-5,
// REL_TIME 1-7) and step counts, run in random chunks, with scenarios
// added while others run and finished ones removed.  Each finished
// scenario must show exactly its own number of root steps in every plane,
// REL_TIME times as many in its child, one max update per (sub-)step and
// the right root-step counter.  Usage: emu_driver SEED ROUNDS
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../include/gcomcot_hip.h"
#include "hip/hip_runtime.h"

namespace {

struct Case {
  int sid;
  long long steps;
  int pnx, pny, cnx = 0, cny = 0, rt = 0;
};

[[noreturn]] void die(const char *m, int sid) {
  std::fprintf(stderr, "scenario %d: %s\n", sid, m);
  std::exit(2);
}

Case add_case(std::mt19937 &rng) {
  auto U = [&](int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); };
  Case c;
  static const bool big = std::getenv("EMU_BIG") != nullptr;
  c.steps = big ? U(1, 6) : U(1, 40);
  c.pnx = big ? U(256, 2048) : U(12, 150);
  c.pny = big ? U(192, 1536) : U(12, 110);
  const bool nest = U(0, 9) < 4 && c.pnx >= 14 && c.pny >= 14;
  int sid = 0;
  if (gcomcot_batch_new_scenario_(&c.steps, &sid)) die("new", 0);
  c.sid = sid;
  int one = 1, minus1 = -1, zero = 0;
  int lc = U(0, 1), lg = U(0, 1) ? 1 : 0, fs = U(0, 1);
  float fc = 0.013f, dt = 1.0f, dx = 1.0f, dy = 1.0f;
  int corners[4] = {1, c.pnx, 1, c.pny};
  std::vector<float> h((size_t)c.pnx * c.pny, 100.0f), z((size_t)c.pnx * c.pny, 0.0f);
  std::vector<float> coef(10 * (size_t)c.pny, 0.0f);
  if (gcomcot_batch_add_layer_(&sid, &one, &minus1, &lc, &lg, &one, &one, corners, &fs, &fc, &dt,
                               &dx, &dy, coef.data(), h.data(), z.data(), &c.pnx, &c.pny))
    die("add root", sid);
  if (nest) {
    int R = U(2, 5), RT = U(1, 7);
    if (!big && U(0, 29) == 0) RT = 2100 + U(0, 100);  // no-graph plans (> 4096 launches)
    int fw = U(1, (c.pnx - 6) / 2), fh = U(1, (c.pny - 6) / 2);
    if (big) fw = fh = std::min(std::min(fw, fh), 1200 / R);
    int IS = U(4, c.pnx - 2 - fw + 1), JS = U(4, c.pny - 2 - fh + 1);
    int cc[4] = {IS, IS + fw - 1, JS, JS + fh - 1};
    c.cnx = fw * R + 1;
    c.cny = fh * R + 1;
    c.rt = RT;
    int two = 2, clc = U(0, 1), clg = U(0, 1);
    std::vector<float> ch((size_t)c.cnx * c.cny, 50.0f), cz((size_t)c.cnx * c.cny, 0.0f);
    std::vector<float> ccoef(10 * (size_t)c.cny, 0.0f);
    float cdt = 0.25f;
    if (gcomcot_batch_add_layer_(&sid, &two, &one, &clc, &clg, &R, &RT, cc, &fs, &fc, &cdt, &dx,
                                 &dy, ccoef.data(), ch.data(), cz.data(), &c.cnx, &c.cny))
      die("add child", sid);
  }
  (void)zero;
  return c;
}

void verify(const Case &c) {
  long long done = 0, total = 0;
  int sid = c.sid;
  if (gcomcot_batch_progress_(&sid, &done, &total) || done != c.steps || total != c.steps)
    die("progress", sid);
  for (int lid = 1; lid <= (c.cnx ? 2 : 1); ++lid) {
    const int nx = lid == 1 ? c.pnx : c.cnx, ny = lid == 1 ? c.pny : c.cny;
    const float want = (float)(lid == 1 ? c.steps : c.steps * c.rt);
    std::vector<float> z((size_t)nx * ny), m(z.size()), n(z.size()), zm(z.size());
    std::vector<int> ar(z.size());
    if (gcomcot_batch_get_state_(&sid, &lid, z.data(), m.data(), n.data())) die("state", sid);
    if (gcomcot_batch_get_max_(&sid, &lid, zm.data(), ar.data())) die("max", sid);
    for (size_t k = 0; k < z.size(); ++k) {
      if (z[k] != want || m[k] != want || n[k] != want) die("plane step count", sid);
      if (zm[k] != want) die("max update count", sid);
    }
    if (lid == 1 && ar[0] != (int)c.steps) die("root-step counter", sid);
  }
}

}  // namespace

int main(int argc, char **argv) {
  const unsigned seed = argc > 1 ? (unsigned)std::atoi(argv[1]) : 1;
  const int rounds = argc > 2 ? std::atoi(argv[2]) : 20;
  std::mt19937 rng(seed);
  auto U = [&](int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); };
  int zero = 0, on = 1;
  int shape = U(-1, 5);
  if (gcomcot_batch_reset_()) return 3;
  if (gcomcot_batch_set_tuning_(&on, &shape)) return 3;
  // two root steps per launch for linear roots: a given shape, or the tuner's choice
  int two = U(0, 3) ? 1 : 0, shape2 = U(-1, 6);
  if (std::getenv("EMU_TUNE_TWO")) {  // the tuner picks the one- and two-step shapes
    two = 1;
    shape2 = -1;
    shape = -1;
    if (gcomcot_batch_set_tuning_(&on, &shape)) return 3;
  }
  if (gcomcot_batch_set_two_step_(&two, shape2 >= 0 ? &shape2 : nullptr)) return 3;
  int fuse = std::getenv("EMU_FUSE") ? std::atoi(std::getenv("EMU_FUSE")) : 1;
  if (gcomcot_batch_set_fused_halves_(&fuse)) return 3;
  std::map<int, Case> live;
  int checked = 0;
  for (int r = 0; r < rounds; ++r) {
    const int add = U(0, 4);
    for (int k = 0; k < add; ++k) {
      Case c = add_case(rng);
      live[c.sid] = c;
    }
    long long chunk = U(1, 17);
    int nf = 0;
    if (gcomcot_batch_run_(&chunk, &nf)) return 4;
    int cap = 256;
    std::vector<int> fin(cap);
    const int nfin = gcomcot_batch_finished_(fin.data(), &cap);
    if (nfin != nf) die("finished count", 0);
    for (int k = 0; k < nfin && k < cap; ++k) {
      verify(live.at(fin[k]));
      ++checked;
      if (U(0, 2)) {  // sometimes keep finished scenarios around
        if (gcomcot_batch_remove_(&fin[k])) return 5;
        live.erase(fin[k]);
      }
    }
  }
  long long big = 1000;
  int nf = 0;
  if (gcomcot_batch_run_(&big, &nf)) return 4;
  for (auto &kv : live) {
    verify(kv.second);
    ++checked;
  }
  if (gcomcot_batch_reset_()) return 3;
  if (emu_live_bytes() != 0) {
    std::fprintf(stderr, "device memory leaked: %lld bytes\n", emu_live_bytes());
    return 6;
  }
  long long l = 0, g = 0, p = 0, cps = 0;
  int tuned = 0;
  gcomcot_batch_info_(&l, &g, &p, &cps, &tuned);
  std::printf("seed %u: %d scenarios verified, shape %d, two-step %d (shape %d), "
              "%lld cross-stream waits, fused halves %d (%lld tiles)\n", seed, checked, shape,
              two, shape2, emu_cross_stream_waits(), fuse, emu_fused_tiles());
  (void)zero;
  return 0;
}
