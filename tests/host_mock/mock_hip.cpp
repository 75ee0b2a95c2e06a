// CPU mock of the HIP runtime calls the zsmpm ABI layer makes: host memory,
// streams/events/graphs as bookkeeping objects, capture state tracked the
// way CUDA defines it (a stream joins a capture by waiting on an event
// recorded in it; the capture ends only when every joined stream rejoined).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <map>
#include <vector>
struct MStream { int id; int capture = 0; MStream* origin = nullptr; };
struct MEvent { int id; int capture = 0; MStream* rec = nullptr; };
static int g_id = 1;
static std::set<void*> g_streams, g_events, g_graphs, g_execs;
static int g_capture_seq = 0;
static std::map<int, std::set<MStream*>> g_joined;   // capture id -> streams that joined
static std::map<int, std::set<MStream*>> g_pending;  // capture id -> joined streams not yet rejoined
#define S(x) reinterpret_cast<MStream*>(x)
#define E(x) reinterpret_cast<MEvent*>(x)
static void chk_stream(hipStream_t s, const char* w) {
  if (s && !g_streams.count(s)) { fprintf(stderr, "MOCK: %s on unknown stream %p\n", w, (void*)s); abort(); }
}
static void chk_event(hipEvent_t e, const char* w) {
  if (!g_events.count(e)) { fprintf(stderr, "MOCK: %s on unknown event %p\n", w, (void*)e); abort(); }
}
extern "C" {
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) { *p = calloc(1, n ? n : 1); return hipSuccess; }
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipMemset(void* p, int v, size_t n) { memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t s) {
  chk_stream(s, "memsetAsync");
  if (s && S(s)->capture) { fprintf(stderr, "MOCK: memsetAsync captured\n"); }
  memset(p, v, n); return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* src, size_t n, hipMemcpyKind, hipStream_t s) {
  chk_stream(s, "memcpyAsync");
  if (s && S(s)->capture) { fprintf(stderr, "MOCK: memcpyAsync captured\n"); abort(); }
  memcpy(d, src, n); return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { *p = malloc(n ? n : 1); return hipSuccess; }
hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
// host-only build: the device array's host shadow stands in for its memory
hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  MStream* m = new MStream{g_id++}; *s = reinterpret_cast<hipStream_t>(m); g_streams.insert(*s); return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  chk_stream(s, "destroy"); if (S(s)->capture) { fprintf(stderr, "MOCK: destroy capturing stream %d\n", S(s)->id); abort(); }
  g_streams.erase(s); delete S(s); return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  chk_stream(s, "sync");
  if (s && S(s)->capture) { fprintf(stderr, "MOCK: sync of capturing stream %d\n", S(s)->id); return hipErrorStreamCaptureUnsupported; }
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) { MEvent* m = new MEvent{g_id++}; *e = reinterpret_cast<hipEvent_t>(m); g_events.insert(*e); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { chk_event(e, "destroy"); g_events.erase(e); delete E(e); return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  chk_event(e, "record"); chk_stream(s, "record");
  E(e)->capture = s ? S(s)->capture : 0; E(e)->rec = S(s);
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) { chk_event(e, "sync"); if (E(e)->capture) return hipErrorCapturedEvent; return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { chk_event(a, "elapsed"); chk_event(b, "elapsed"); *ms = 1.f; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  chk_event(e, "wait"); chk_stream(s, "wait");
  MStream* st = S(s);
  if (E(e)->capture) {
    if (!st) { fprintf(stderr, "MOCK: null stream waits on captured event\n"); return hipErrorStreamCaptureIsolation; }
    const int c = E(e)->capture;
    if (st->capture == 0) { st->capture = c; g_joined[c].insert(st); g_pending[c].insert(st); }
    else if (st->capture != c) { fprintf(stderr, "MOCK: cross-capture wait\n"); return hipErrorStreamCaptureIsolation; }
    // the waiting stream absorbs the recording stream's branch: rejoin
    if (E(e)->rec && E(e)->rec != st) g_pending[c].erase(E(e)->rec);
  } else if (st && st->capture && E(e)->rec && !E(e)->rec->capture) {
    // waiting on an eager event from a capturing stream: allowed (external dependency) but flag
  }
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  chk_stream(s, "begin"); if (!s) return hipErrorInvalidValue;
  if (S(s)->capture) return hipErrorIllegalState;
  S(s)->capture = ++g_capture_seq; return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  chk_stream(s, "end");
  const int c = S(s)->capture;
  if (!c) return hipErrorIllegalState;
  int unjoined = 0;
  for (MStream* m : g_pending[c]) if (m != S(s)) { fprintf(stderr, "MOCK: stream %d not rejoined at end of capture %d\n", m->id, c); ++unjoined; }
  for (MStream* m : g_joined[c]) m->capture = 0;
  S(s)->capture = 0;
  g_joined.erase(c); g_pending.erase(c);
  if (unjoined) { *g = nullptr; return hipErrorStreamCaptureUnjoined; }
  *g = reinterpret_cast<hipGraph_t>(new int(c)); g_graphs.insert(*g); return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  if (!g_graphs.count(g)) { fprintf(stderr, "MOCK: instantiate unknown graph\n"); abort(); }
  *x = reinterpret_cast<hipGraphExec_t>(new int(1)); g_execs.insert(*x); return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { if (!g_graphs.erase(g)) { fprintf(stderr, "MOCK: destroy unknown graph\n"); abort(); } delete reinterpret_cast<int*>(g); return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t x) { if (!g_execs.erase(x)) { fprintf(stderr, "MOCK: destroy unknown exec\n"); abort(); } delete reinterpret_cast<int*>(x); return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  chk_stream(s, "graph launch"); if (!g_execs.count(x)) { fprintf(stderr, "MOCK: launch unknown exec %p\n", (void*)x); abort(); } return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
static hipStream_t g_cfg_stream = nullptr;
hipError_t __hipPushCallConfiguration(dim3, dim3, size_t, hipStream_t s) { g_cfg_stream = s; return hipSuccess; }
hipError_t __hipPopCallConfiguration(dim3* g, dim3* b, size_t* m, hipStream_t* s) { *g = dim3(1); *b = dim3(1); *m = 0; *s = g_cfg_stream; return hipSuccess; }
hipError_t hipLaunchKernel(const void*, dim3, dim3, void**, size_t, hipStream_t s) { chk_stream(s, "launch"); return hipSuccess; }
void** __hipRegisterFatBinary(const void*) { static void* h[4]; return h; }
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*, void*, void*, int*) {}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterVar(void**, void*, char*, const char*, int, size_t, int, int) {}
}
