// Reference-format output files of the device loop: schedule, device pack,
// asynchronous copy-out and the writer thread.  File names, record layout
// and masking follow output.f90 (DATA_PRT :758-851, MAX_AMP :855-940).
#include "output.h"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace gc {
namespace {

struct FileRec {
  std::string name;
  size_t off = 0, bytes = 0;
};

enum SlotState { kFree, kCopying, kWriting };

struct Slot {
  void *dev = nullptr, *host = nullptr;
  size_t cap = 0;
  hipEvent_t packed = nullptr;
  SlotState state = kFree;
  size_t remaining = 0;  // files of the slot not yet on disk
  std::vector<FileRec> files;
};

constexpr int kSlots = 3;
// Writer threads; the files of one output go down in parallel.  A file
// name always goes to the same writer, so rewrites of one name (the final
// maximum file) land in step order.
constexpr int kWriters = 4;

struct Task {
  int slot, file;
};

struct Output {
  bool active = false;
  std::string dir;
  int iprt = 0, out_opt = 1;
  float tend = 0.0f, tide = 0.0f;
  int flux[kMaxLayers];
  bool initial_pending = false;
  bool trace = false;
  // model time after root step clk_k, accumulated in REAL as the main
  // loop does (comcot.f90:304, 323)
  long long clk_k = 0;
  float clk_t = 0.0f, clk_dt = 0.0f;
  hipStream_t copy_st = nullptr;
  Slot slots[kSlots];
  int next = 0;
  bool started = false;
  // shared with the copy callbacks and the writer threads
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Task> queue[kWriters];
  long long nfiles = 0;
  int err = 0;
  std::string errmsg;

  Output() {
    for (int &f : flux) f = 1;
    const char *e = std::getenv("GCOMCOT_OUTPUT_TRACE");
    trace = e && *e && *e != '0';
  }
};

// Never destroyed: a copy callback or a writer may still run while the
// process ends (files not flushed by then are not guaranteed).
Output &g_out = *new Output;

// Fortran Iw.w edit descriptor: zero-padded to w digits, w asterisks when
// the value does not fit.
std::string fort_int(long long v, int w) {
  char buf[32];
  if (v < 0) return std::string((size_t)w, '*');
  std::snprintf(buf, sizeof buf, "%0*lld", w, v);
  std::string s(buf);
  if ((int)s.size() > w) return std::string((size_t)w, '*');
  return s;
}

int write_record(const std::string &path, const char *data, size_t bytes, std::string *why) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    *why = "cannot open " + path;
    return 1;
  }
  // one unformatted sequential record: 4-byte length, data, 4-byte length
  const uint32_t len = (uint32_t)bytes;
  bool ok = std::fwrite(&len, 4, 1, f) == 1;
  ok = ok && std::fwrite(data, 1, bytes, f) == bytes;
  ok = ok && std::fwrite(&len, 4, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) *why = "write failed: " + path;
  return ok ? 0 : 1;
}

int writer_of(const std::string &name) {
  return (int)(std::hash<std::string>{}(name) % kWriters);
}

void writer_main(int w) {
  Output &O = g_out;
  for (;;) {
    Task t;
    {
      std::unique_lock<std::mutex> lk(O.mu);
      O.cv.wait(lk, [&] { return !O.queue[w].empty(); });
      t = O.queue[w].front();
      O.queue[w].pop_front();
    }
    // the slot's file list is not touched again until its last file is down
    Slot &S = O.slots[t.slot];
    const FileRec &f = S.files[(size_t)t.file];
    const std::string path = O.dir.empty() ? f.name : O.dir + "/" + f.name;
    std::string why;
    const auto t0 = std::chrono::steady_clock::now();
    const int bad = write_record(path, static_cast<const char *>(S.host) + f.off, f.bytes, &why);
    if (O.trace) {
      const double ms = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0).count();
      std::fprintf(stderr, "[gcomcot output] writer %d: %s %zu B %.2f ms\n", w, f.name.c_str(),
                   f.bytes + 8, ms);
    }
    {
      std::lock_guard<std::mutex> lk(O.mu);
      if (bad) {
        if (!O.err) {
          O.err = 1;
          O.errmsg = why;
        }
      } else {
        ++O.nfiles;
      }
      if (--S.remaining == 0) S.state = kFree;
    }
    O.cv.notify_all();
  }
}

// Host function on the copy stream after a slot's copy: its files go to
// the writers.  Runs on a runtime thread and makes no HIP call.
void on_copied(void *arg) {
  Output &O = g_out;
  const int s = (int)(intptr_t)arg;
  {
    std::lock_guard<std::mutex> lk(O.mu);
    Slot &S = O.slots[s];
    S.state = kWriting;
    S.remaining = S.files.size();
    for (size_t f = 0; f < S.files.size(); ++f) {
      // a queued, unstarted write of the same file is superseded: the file
      // would be overwritten by this one (the final maximum file is
      // rewritten on the last steps of a run)
      auto &q = O.queue[writer_of(S.files[f].name)];
      for (auto it = q.begin(); it != q.end(); ++it) {
        Slot &P = O.slots[it->slot];
        if (P.files[(size_t)it->file].name != S.files[f].name) continue;
        q.erase(it);
        ++O.nfiles;  // counted as written: its content is on its way
        if (--P.remaining == 0) P.state = kFree;
        break;
      }
      q.push_back({s, (int)f});
    }
    if (S.files.empty()) S.state = kFree;
  }
  O.cv.notify_all();
}

void wait_free(int s) {
  std::unique_lock<std::mutex> lk(g_out.mu);
  g_out.cv.wait(lk, [&] { return g_out.slots[s].state == kFree; });
}

float time_at(long long k) {
  Output &O = g_out;
  if (k < O.clk_k) {
    O.clk_k = 0;
    O.clk_t = 0.0f;
  }
  while (O.clk_k < k) {
    O.clk_t = O.clk_t + O.clk_dt;
    ++O.clk_k;
  }
  return O.clk_t;
}

// MAX_AMP's file test after root step k (output.f90:912-916)
bool zmax_due(long long k, float *time) {
  Output &O = g_out;
  if (!(O.out_opt == 0 || O.out_opt == 2) || k < 1) return false;
  const float t = time_at(k), dt = O.clk_dt;
  long long k1 = (long long)std::floor((t - dt) / 1800.0f);
  if (k1 <= 0) k1 = 0;
  const long long k2 = (long long)std::floor(t / 1800.0f);
  *time = t;
  return k2 > k1 || t >= O.tend - 2.0f * dt;
}

bool snap_due(long long k) {
  const Output &O = g_out;
  if (O.iprt <= 0) return false;
  return k == 0 ? O.initial_pending : k % O.iprt == 0;
}

// Give slot S at least `total` bytes of staging and pinned memory (the
// slot must be free).
int reserve(Slot &S, size_t total) {
  if (S.cap >= total) return 0;
  if (S.dev) HIPCHK(hipFree(S.dev));
  if (S.host) HIPCHK(hipHostFree(S.host));
  S.dev = S.host = nullptr;
  S.cap = 0;
  HIPCHK(hipMalloc(&S.dev, total));
  HIPCHK(hipHostMalloc(&S.host, total, hipHostMallocDefault));
  S.cap = total;
  return 0;
}

int release_slots() {
  for (int s = 0; s < kSlots; ++s) {
    Slot &S = g_out.slots[s];
    if (S.dev) (void)hipFree(S.dev);
    if (S.host) (void)hipHostFree(S.host);
    S.dev = S.host = nullptr;
    S.cap = 0;
  }
  return 0;
}

}  // namespace

int out_configure(Tree &T, const char *dir, int dirlen, int iprt, int out_opt, float tend,
                  float tide) {
  const LayerRec *R = T.layer(1);
  if (!R) return fail(8, "output: no root layer");
  if (dirlen < 0 || (dirlen > 0 && !dir)) return fail(50, "output: bad directory argument");
  if (int r = out_disable(T)) return r;
  Output &O = g_out;
  O.dir.assign(dir ? dir : "", dir ? (size_t)dirlen : 0);
  while (!O.dir.empty() && (O.dir.back() == ' ' || O.dir.back() == '\0')) O.dir.pop_back();
  O.iprt = iprt;
  O.out_opt = out_opt;
  O.tend = tend;
  O.tide = tide;
  O.initial_pending = true;
  O.clk_k = 0;
  O.clk_t = 0.0f;
  O.clk_dt = R->dt;
  {
    std::lock_guard<std::mutex> lk(O.mu);
    O.nfiles = 0;
    O.err = 0;
    O.errmsg.clear();
  }
  if (!O.copy_st) HIPCHK(hipStreamCreateWithFlags(&O.copy_st, hipStreamNonBlocking));
  for (Slot &S : O.slots)
    if (!S.packed) HIPCHK(hipEventCreateWithFlags(&S.packed, hipEventDisableTiming));
  if (!O.started) {
    for (int w = 0; w < kWriters; ++w) std::thread(writer_main, w).detach();
    O.started = true;
  }
  // the largest output of the present tree, reserved now rather than in the loop
  size_t most = 0;
  for (int id = 1; id < kMaxLayers; ++id) {
    const LayerRec *L = T.layer(id);
    if (!L) continue;
    const size_t plane = (L->nxy() * sizeof(float) + 255) & ~(size_t)255;
    const int fs = O.flux[id];
    if (iprt > 0) most += plane * ((fs <= 1 || fs == 9) ? 1 : 0) + plane * (fs == 0 ? 2 : 0);
    if (id == 1 && (out_opt == 0 || out_opt == 2)) most += plane;
  }
  for (Slot &S : O.slots)
    if (int r = reserve(S, most)) return r;
  T.raw_max = (out_opt == 0 || out_opt == 2) ? 1 : 0;
  O.active = true;
  return 0;
}

int out_set_fluxswitch(int lid, int fluxswitch) {
  if (lid < 1 || lid >= kMaxLayers) return fail(1, "invalid layer id %d (1..%d)", lid, kMaxLayers - 1);
  g_out.flux[lid] = fluxswitch;
  return 0;
}

int out_disable(Tree &T) {
  long long n = 0;
  const int r = g_out.active ? out_flush(&n) : 0;
  g_out.active = false;
  T.raw_max = 0;
  release_slots();
  std::lock_guard<std::mutex> lk(g_out.mu);
  g_out.nfiles = 0;
  g_out.err = 0;
  g_out.errmsg.clear();
  return r;
}

bool out_active() { return g_out.active; }

bool out_due(const Tree &T, long long k) {
  (void)T;
  if (!g_out.active) return false;
  float t;
  return snap_due(k) || zmax_due(k, &t);
}

long long out_steps_to_next(const Tree &T, long long done, long long n) {
  for (long long j = 1; j < n; ++j)
    if (out_due(T, done + j)) return j;
  return n;
}

int out_write(Tree &T, long long k, hipStream_t st) {
  Output &O = g_out;
  if (!O.active) return 0;
  float t = 0.0f;
  const bool snap = snap_due(k), zmax = zmax_due(k, &t);
  if (!snap && !zmax) return 0;
  if (k == 0) O.initial_pending = false;

  // the planes of this output, in staging order
  struct Plan {
    LayerRec *L;
    size_t oz, om, on, ozmax;  // byte offsets + 1 (0 = not written)
  };
  std::vector<Plan> plans;
  std::vector<FileRec> files;
  size_t total = 0;
  auto take = [&](const LayerRec &L, std::string name) {
    const size_t b = L.nxy() * sizeof(float);
    files.push_back({std::move(name), total, b});
    total += (b + 255) & ~(size_t)255;
    return files.back().off + 1;
  };
  for (int id = 1; id < kMaxLayers; ++id) {
    LayerRec *L = T.layer(id);
    if (!L) continue;
    if (L->nxy() * sizeof(float) > (size_t)INT32_MAX)
      return fail(51, "output: layer %d is too large for one record", id);
    Plan p{L, 0, 0, 0, 0};
    if (snap) {
      const int fs = O.flux[id];
      const std::string tail = fort_int(id, 2) + "_" + fort_int(k, 6) + ".dat";
      if (fs <= 1 || fs == 9) p.oz = take(*L, "z_" + tail);
      if (fs == 0) {
        p.om = take(*L, "m_" + tail);
        p.on = take(*L, "n_" + tail);
      }
    }
    if (zmax && id == 1) {
      const float tend = O.tend, dt = O.clk_dt;
      std::string name = t >= tend - 2.0f * dt
                             ? std::string("zmax_layer01.dat")
                             : "zmax_layer" + fort_int(id, 2) + "_" +
                                   fort_int(std::lround(t / 60.0f), 4) + "mins.dat";
      p.ozmax = take(*L, std::move(name));
    }
    if (p.oz || p.om || p.on || p.ozmax) plans.push_back(p);
  }
  if (plans.empty()) return 0;

  const int s = O.next;
  O.next = (O.next + 1) % kSlots;
  Slot &S = O.slots[s];
  wait_free(s);
  if (int r = reserve(S, total)) return r;
  char *base = static_cast<char *>(S.dev);
  auto at = [&](size_t off1) { return off1 ? reinterpret_cast<float *>(base + off1 - 1) : nullptr; };
  for (const Plan &p : plans) {
    const LayerRec &L = *p.L;
    OutPack q{};
    q.z = L.z[L.cur];
    q.h = L.h;
    q.m = L.m[L.cur];
    q.n = L.n[L.cur];
    q.zraw = L.zmax_raw;
    q.oz = at(p.oz);
    q.om = at(p.om);
    q.on = at(p.on);
    q.ozmax = at(p.ozmax);
    q.tide = O.tide;
    q.count = (long long)L.nxy();
    hipError_t e = launch_out_pack(q, st);
    if (e != hipSuccess) return fail(205, "output pack launch failed: %s", hipGetErrorString(e));
  }
  HIPCHK(hipEventRecord(S.packed, st));
  HIPCHK(hipStreamWaitEvent(O.copy_st, S.packed, 0));
  HIPCHK(hipMemcpyAsync(S.host, S.dev, total, hipMemcpyDeviceToHost, O.copy_st));
  {
    std::lock_guard<std::mutex> lk(O.mu);
    S.files = std::move(files);
    S.state = kCopying;
  }
  const hipError_t e = hipLaunchHostFunc(O.copy_st, on_copied, (void *)(intptr_t)s);
  if (e != hipSuccess) {  // nothing will hand the slot on: give it back
    {
      std::lock_guard<std::mutex> lk(O.mu);
      S.files.clear();
      S.state = kFree;
    }
    O.cv.notify_all();
    return fail(100 + (int)e, "output: host function launch failed: %s", hipGetErrorString(e));
  }
  return 0;
}

int out_flush(long long *nfiles) {
  Output &O = g_out;
  for (int s = 0; s < kSlots; ++s) wait_free(s);
  std::string why;
  {
    std::lock_guard<std::mutex> lk(O.mu);
    if (nfiles) *nfiles = O.nfiles;
    if (!O.err) return 0;
    why = O.errmsg;  // reported once: the next flush reports later failures only
    O.err = 0;
    O.errmsg.clear();
  }
  return fail(52, "output: %s", why.c_str());
}

}  // namespace gc
