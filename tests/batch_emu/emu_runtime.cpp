// CPU stand-in for the HIP runtime calls of tree.hip / batch.hip (see
// hip/hip_runtime.h here).  Work submitted to a stream runs at once unless
// the stream is capturing, in which case it is recorded into the graph.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "hip/hip_runtime.h"

struct Work {
  void (*fn)(void *);
  std::vector<char> arg;
};
struct EmuGraph {
  std::vector<Work> work;
  int forked = 0;  // streams that joined the capture and have not been joined back
};
struct EmuStream {
  bool capturing = false;
  bool joined = false;  // capturing because it waited on an event of a capture
  EmuGraph *graph = nullptr;
};
struct EmuEvent {
  EmuStream *src = nullptr;  // stream of the last record
  EmuGraph *graph = nullptr; // capture it was recorded in, if any
};

static std::map<void *, size_t> g_alloc;
static long long g_live = 0;

long long emu_live_bytes() { return g_live; }

void emu_submit(hipStream_t st, void (*fn)(void *), void *arg, size_t n) {
  if (st && st->capturing) {
    Work w{fn, std::vector<char>((char *)arg, (char *)arg + n)};
    st->graph->work.push_back(std::move(w));
    return;
  }
  fn(arg);
}

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->name, "emulator");
  p->multiProcessorCount = 4;
  p->warpSize = 64;
  return hipSuccess;
}
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "error"; }
hipError_t hipMalloc(void **p, size_t n) {
  *p = std::calloc(n ? n : 1, 1);
  if (!*p) return hipErrorOutOfMemory;
  // fill with a pattern so that reads of never-written memory show up
  std::memset(*p, 0x7f, n);
  g_alloc[*p] = n;
  g_live += (long long)n;
  return hipSuccess;
}
hipError_t hipFree(void *p) {
  if (!p) return hipSuccess;
  auto it = g_alloc.find(p);
  if (it == g_alloc.end()) { std::fprintf(stderr, "emu: free of unknown pointer\n"); std::abort(); }
  g_live -= (long long)it->second;
  g_alloc.erase(it);
  std::free(p);
  return hipSuccess;
}
struct SetArg { void *p; int v; size_t n; };
static void do_set(void *a) { SetArg *s = (SetArg *)a; std::memset(s->p, s->v, s->n); }
hipError_t hipMemset(void *p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st) {
  if (st && st->capturing) { std::fprintf(stderr, "emu: memset captured into a graph\n"); std::abort(); }
  SetArg a{p, v, n};
  emu_submit(st, do_set, &a, sizeof(a));
  return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st) {
  if (st && st->capturing) { std::fprintf(stderr, "emu: memcpy captured\n"); std::abort(); }
  return hipMemcpy(d, s, n, k);
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = new EmuStream(); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t s) {
  if (s && s->capturing) { std::fprintf(stderr, "emu: synchronize while capturing\n"); std::abort(); }
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  s->capturing = true;
  s->graph = new EmuGraph();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (s->graph && s->graph->forked != 0) {
    std::fprintf(stderr, "emu: capture ended with a forked stream not joined back\n");
    std::abort();
  }
  s->capturing = false;
  *g = s->graph;
  s->graph = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new EmuGraph(*g);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st) {
  for (auto &w : e->work) emu_submit(st, w.fn, w.arg.data(), w.arg.size());
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new EmuEvent(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = new EmuEvent(); return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st) {
  e->src = st;
  e->graph = st && st->capturing ? st->graph : nullptr;
  return hipSuccess;
}
static long long g_waits = 0;
long long emu_cross_stream_waits() { return g_waits; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  if (e->src && e->src != s) ++g_waits;
  if (!e->graph || e->src == s) return hipSuccess;  // outside a capture work runs in order
  if (!s->capturing) {  // fork: s joins the capture
    s->capturing = true;
    s->joined = true;
    s->graph = e->graph;
    e->graph->forked++;
  } else if (e->src->joined && e->src->graph == s->graph) {  // join back
    e->src->capturing = false;
    e->src->joined = false;
    e->src->graph = nullptr;
    s->graph->forked--;
  }
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
  static float t = 1.0f;
  t = t * 0.5f + 0.25f;  // varying, so the tuner's choice is exercised
  *ms = t;
  return hipSuccess;
}
