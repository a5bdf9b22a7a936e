// capi.hip -- the C ABI of libqblas_hip.so (declared in include/qblas_hip.h).
//
// Host-pointer entry points (the reference's quadblas_q* functions and the full-precision
// quadblas_*_q functions behind the C++ template API) stage their operands into per-thread device
// buffers on a per-thread stream, run the device routine and copy the outputs back before
// returning (synchronous, like the reference).  Device-pointer entry points (qblas_hip_*) only
// enqueue work on the caller's stream.
//
// There is no CPU fallback: when a HIP call fails the function records the error (thread-local,
// quadblas_hip_last_error), leaves outputs unchanged and returns.
#include <hip/hip_runtime.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/qblas_hip.h"
#include "kernels.h"

using qb::q128;

namespace {

thread_local std::string t_err;

void set_err(const char *where, hipError_t e) {
  t_err = std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
}
void set_err_msg(const std::string &m) { t_err = m; }

#define QB_CHECK(call)                 \
  do {                                 \
    hipError_t e_ = (call);            \
    if (e_ != hipSuccess) {            \
      set_err(#call, e_);              \
      return QBLAS_ERR_HIP;            \
    }                                  \
  } while (0)

std::atomic<int> g_threads{0};  // 0: host default
std::atomic<int> g_kc{126};
std::atomic<int> g_mode{-1};    // -1: not yet read from the environment
std::atomic<int> g_oz_slices{0};  // OZAKI slice count, 0: chosen per call
std::atomic<int> g_oz_method{QBLAS_OZAKI_AUTO};  // OZAKI method
std::atomic<int> g_oz_bits{0};    // CRT method: bits per scaled operand, 0: chosen per call

int default_threads() {
  // omp_get_max_threads() without an OpenMP runtime: OMP_NUM_THREADS's first entry, else the
  // cores this process may run on.
  const char *e = getenv("OMP_NUM_THREADS");
  if (e) {
    const int v = atoi(e);
    if (v > 0) return v;
  }
  cpu_set_t set;
  CPU_ZERO(&set);
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    const int c = CPU_COUNT(&set);
    if (c > 0) return c;
  }
  const long c = sysconf(_SC_NPROCESSORS_ONLN);
  return c > 0 ? (int)c : 1;
}
int cur_threads() {
  const int t = g_threads.load();
  return t > 0 ? t : default_threads();
}

int env_mode() {
  const char *e = getenv("QBLAS_MODE");
  if (!e) return QBLAS_MODE_REFERENCE;
  if (!strcmp(e, "exact")) return QBLAS_MODE_EXACT;
  if (!strcmp(e, "dd")) return QBLAS_MODE_DD;
  if (!strcmp(e, "ozaki")) return QBLAS_MODE_OZAKI;
  return QBLAS_MODE_REFERENCE;
}
// The global mode for the host-pointer entry points, per routine: OZAKI exists for gemm only,
// so dot / gemv / nrm2 then use EXACT (correctly rounded, a stronger guarantee); EXACT has no gemm,
// which then uses OZAKI (certified to 2^-104).
int cur_mode();
int vec_mode() {
  const int m = cur_mode();
  return m == QBLAS_MODE_OZAKI ? QBLAS_MODE_EXACT : m;
}
int mat_mode() {
  const int m = cur_mode();
  return m == QBLAS_MODE_EXACT ? QBLAS_MODE_OZAKI : m;
}
int cur_mode() {
  int m = g_mode.load();
  if (m < 0) {
    m = env_mode();
    int expected = -1;
    g_mode.compare_exchange_strong(expected, m);
    m = g_mode.load();
  }
  return m;
}

// ---------------------------------------------------------------------------------------------
// device workspaces: grow-only per host thread; cross-stream reuse is ordered with an event.  The
// event only orders work on the workspace's own device (a stream wait, or a host wait before
// hipFree), so it is recorded without the system-scope fence, which costs a cache writeback per
// call.
// ---------------------------------------------------------------------------------------------
struct Workspace {
  void *p = nullptr;
  size_t cap = 0;
  int dev = -1;
  hipEvent_t ev = nullptr;
  bool recorded = false;
  hipStream_t last = nullptr;
};

int ws_acquire(Workspace &w, size_t bytes, hipStream_t s, void **out) {
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  if (w.p && w.dev != dev) {
    // the thread moved to another device: drop the old buffer once its work is done
    if (w.recorded) QB_CHECK(hipEventSynchronize(w.ev));
    QB_CHECK(hipFree(w.p));
    QB_CHECK(hipEventDestroy(w.ev));
    w = Workspace();
  }
  if (!w.ev) QB_CHECK(hipEventCreateWithFlags(&w.ev, hipEventDisableTiming | hipEventDisableSystemFence));
  if (bytes > w.cap) {
    if (w.p) {
      if (w.recorded) QB_CHECK(hipEventSynchronize(w.ev));
      QB_CHECK(hipFree(w.p));
      w.p = nullptr;
      w.cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    QB_CHECK(hipMalloc(&w.p, want));
    w.cap = want;
    w.recorded = false;
    w.dev = dev;
  } else if (w.recorded && w.last != s) {
    QB_CHECK(hipStreamWaitEvent(s, w.ev, 0));
  }
  *out = w.p;
  return QBLAS_OK;
}
int ws_release(Workspace &w, hipStream_t s) {
  QB_CHECK(hipEventRecord(w.ev, s));
  w.recorded = true;
  w.last = s;
  return QBLAS_OK;
}

// Scratch of the device routines (partials, scaled x, OZAKI slices): one grow-only buffer per
// (device, stream) this host thread has called on, up to WS_STREAMS of them.  Work on one stream
// runs in order, so a stream reusing its own buffer needs no event (an event record costs a few
// microseconds of queue time per call, 2% of a memory-bound gemv).  A buffer grows after its
// stream has drained; the least recently used entry is dropped after a device synchronization.
// A buffer handed out while its stream was capturing a graph may be referenced by that graph's
// replays: it is never freed (growth and eviction retire it instead).  A capture allows neither
// synchronization nor hipMalloc: a capturing stream without a large enough buffer takes over the
// least recently used large enough buffer of another stream (which gets a new one on its next
// call); with none, the call fails.
constexpr int WS_STREAMS = 4;
struct StreamWorkspaces {
  struct Entry {
    hipStream_t s = nullptr;
    int dev = -1;
    void *p = nullptr;
    size_t cap = 0;
    uint64_t used = 0;  // 0: free slot; else the stamp of the latest acquisition
    bool captured = false;
  } e[WS_STREAMS];
  uint64_t clock = 0;
};
std::mutex g_retired_mu;
std::vector<void *> g_retired;  // buffers a captured graph may use: kept until the process ends
void ws_retire(void *p) {
  std::lock_guard<std::mutex> lk(g_retired_mu);
  g_retired.push_back(p);
}
bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return cs == hipStreamCaptureStatusActive;
}
// Where a call's status words live: valid while the workspace entry has not been handed out again
// (its stamp unchanged) -- a later call on that stream reuses, regrows or drops the buffer.
struct StatusRef {
  const uint32_t *p = nullptr;
  const StreamWorkspaces::Entry *entry = nullptr;
  uint64_t stamp = 0;
  bool valid() const { return p && entry && entry->used == stamp; }
};
int ws_acquire(StreamWorkspaces &W, size_t bytes, hipStream_t s, void **out,
               const StreamWorkspaces::Entry **used_entry = nullptr) {
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  const bool capturing = stream_capturing(s);
  StreamWorkspaces::Entry *hit = nullptr, *lru = &W.e[0];
  for (auto &x : W.e) {
    if (x.used && x.s == s && x.dev == dev) {
      hit = &x;
      break;
    }
    if (x.used < lru->used) lru = &x;
  }
  if (capturing && (!hit || bytes > hit->cap)) {
    // no allocation is allowed during a capture: the least recently used large enough buffer of
    // this device moves to the capturing stream (its former stream gets a new one next time)
    StreamWorkspaces::Entry *donor = nullptr;
    for (auto &x : W.e)
      if (&x != hit && x.used && x.dev == dev && x.p && x.cap >= bytes && (!donor || x.used < donor->used)) donor = &x;
    if (!donor) {
      set_err_msg("no workspace large enough during graph capture: make the same call once before capturing");
      return QBLAS_ERR_ARG;
    }
    if (hit) {
      if (hit->p) ws_retire(hit->p);
      *hit = StreamWorkspaces::Entry();
    }
    donor->s = s;
    hit = donor;
  }
  if (!hit) {
    hit = lru;
    if (hit->p) {
      if (hit->captured || capturing) {
        ws_retire(hit->p);  // no synchronization possible (capture) or a graph may still use it
      } else {  // drop the least recently used buffer once every stream of its device is done
        if (hit->dev != dev) QB_CHECK(hipSetDevice(hit->dev));
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipFree(hit->p);
        if (hit->dev != dev) (void)hipSetDevice(dev);
        if (e != hipSuccess) {
          set_err("workspace", e);
          return QBLAS_ERR_HIP;
        }
      }
    }
    *hit = StreamWorkspaces::Entry();
    hit->s = s;
    hit->dev = dev;
  }
  if (bytes > hit->cap) {
    if (hit->p) {
      if (hit->captured || capturing) {  // a graph may use it, or no synchronization is allowed
        ws_retire(hit->p);
      } else {
        QB_CHECK(hipStreamSynchronize(s));  // the stream's earlier calls may still read the buffer
        QB_CHECK(hipFree(hit->p));
      }
      hit->p = nullptr;
      hit->cap = 0;
      hit->captured = false;
    }
    const size_t want = bytes < 4096 ? 4096 : bytes;
    QB_CHECK(hipMalloc(&hit->p, want));
    hit->cap = want;
  }
  if (capturing) hit->captured = true;
  hit->used = ++W.clock;
  *out = hit->p;
  if (used_entry) *used_entry = hit;
  return QBLAS_OK;
}
int ws_release(StreamWorkspaces &, hipStream_t) { return QBLAS_OK; }

thread_local StreamWorkspaces t_ws;
thread_local StreamWorkspaces t_bws;                    // batched calls: prefix sums, OZAKI batch arrays
thread_local StatusRef t_exact_status;                  // status words of the last EXACT/OZAKI-form call
thread_local StatusRef t_oz_hdr;                        // header of the last OZAKI-form call
std::atomic<int> g_phase_timing{0};                     // record phase events in OZAKI calls
thread_local hipEvent_t t_oz_ev[5] = {};
thread_local bool t_oz_ev_valid = false;
thread_local int t_oz_slices = 0;                       // slice count of the last OZAKI call
thread_local int t_oz_method = 0;                       // method of the last OZAKI call
thread_local int t_oz_bits = 0;                         // CRT bits of the last OZAKI call
thread_local int t_oz_moduli = 0;                       // CRT moduli of the last OZAKI call

// the status words at `words` of workspace entry `e` (its latest acquisition) become the
// thread's EXACT status (and the OZAKI header when oz)
void set_status(const void *words, const StreamWorkspaces::Entry *e, bool oz) {
  StatusRef r;
  r.p = (const uint32_t *)words;
  r.entry = e;
  r.stamp = e ? e->used : 0;
  t_exact_status = r;
  if (oz) t_oz_hdr = r;
}

// After an EXACT- or OZAKI-form call on this thread: report a term that fell outside the
// full-range accumulator (cannot happen for binary128 inputs).  QBLAS_OK when there is no such
// call (a host-pointer entry point clears the reference first, so it sees only its own call).
int check_exact_status(hipStream_t s) {
  if (!t_exact_status.p) return QBLAS_OK;
  if (!t_exact_status.valid()) {
    t_exact_status = StatusRef();
    set_err_msg("exact status: the call's workspace was reused by a later call on its stream");
    return QBLAS_ERR_ARG;
  }
  uint32_t st = 0;
  QB_CHECK(hipMemcpyAsync(&st, t_exact_status.p, 4, hipMemcpyDeviceToHost, s));
  QB_CHECK(hipStreamSynchronize(s));
  if (st & 1u) {
    set_err_msg("exact form: a term fell outside the full-range accumulator; the output was set to NaN");
    return QBLAS_ERR_RANGE;
  }
  if (st & 2u) {
    set_err_msg("batched OZAKI form replayed from a graph: the descriptors' shapes no longer match the captured "
                "call; nothing was computed");
    return QBLAS_ERR_ARG;
  }
  return QBLAS_OK;
}
// words 2 of the same status area: items recomputed on the full-range path by that call
long long exact_wide_items(hipStream_t s) {
  if (!t_exact_status.valid()) return -1;
  uint32_t v = 0;
  if (hipMemcpyAsync(&v, t_exact_status.p + 2, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  return (long long)v;
}
thread_local Workspace t_stage[4];     // host-API operand staging
thread_local hipStream_t t_stream = nullptr;

int host_stream(hipStream_t *s) {
  if (!t_stream) QB_CHECK(hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking));
  *s = t_stream;
  return QBLAS_OK;
}

// the host gemm's second stream (C's copy in) and its two events
thread_local hipStream_t t_stream2 = nullptr;
thread_local hipEvent_t t_ev_ab = nullptr, t_ev_c = nullptr;
int host_stream2(hipStream_t *s) {
  if (!t_stream2) QB_CHECK(hipStreamCreateWithFlags(&t_stream2, hipStreamNonBlocking));
  if (!t_ev_ab) QB_CHECK(hipEventCreateWithFlags(&t_ev_ab, hipEventDisableTiming));
  if (!t_ev_c) QB_CHECK(hipEventCreateWithFlags(&t_ev_c, hipEventDisableTiming));
  *s = t_stream2;
  return QBLAS_OK;
}

// ---------------------------------------------------------------------------------------------
// host <-> device copies of the host-pointer entry points
// ---------------------------------------------------------------------------------------------
// A caller's pinned buffer (hipHostMalloc) is copied by one DMA.  A pageable one (what malloc
// gives) goes through pinned staging slots: up to stage_workers() threads (QBLAS_STAGE_WORKERS,
// 1..16, default 8; the calling thread is one of them) take interleaved chunks of stage_chunk()
// bytes (QBLAS_STAGE_CHUNK_MB, 1..64, default 8), each copying a chunk between the caller's buffer
// and one of its STAGE_SLOTS slots with memcpy while the DMA engine moves its other slot, and
// waiting on the slot's event before reusing it.  The slots (128 MiB pinned with the defaults)
// are allocated at the thread's first staged copy and kept.  Measured on MI355X: 50 GB/s host to
// device from pageable memory (56 GB/s from pinned; the runtime's own pageable path, which stages
// on one thread, 28 GB/s); 4 to 16 workers give the same rate, 8 and 16 MiB chunks beat 2 and 4
// MiB (per-copy overhead).  Spans below STAGE_DIRECT use one plain copy.
constexpr int STAGE_WORKERS = 16;  // at most
constexpr int STAGE_SLOTS = 2;
constexpr size_t STAGE_DIRECT = (size_t)2 << 20;
struct StageRing {
  bool ready = false;
  void *slot[STAGE_WORKERS][STAGE_SLOTS] = {};
  hipEvent_t ev[STAGE_WORKERS][STAGE_SLOTS] = {};
  bool pending[STAGE_WORKERS][STAGE_SLOTS] = {};  // a DMA from the slot may still be running
};
thread_local StageRing t_ring;

size_t stage_chunk() {
  static const size_t n = [] {
    const char *e = getenv("QBLAS_STAGE_CHUNK_MB");
    const int v = e ? atoi(e) : 8;
    return (size_t)(v < 1 ? 1 : (v > 64 ? 64 : v)) << 20;
  }();
  return n;
}

int stage_workers() {
  static const int n = [] {
    const char *e = getenv("QBLAS_STAGE_WORKERS");
    const int v = e ? atoi(e) : 8;
    return v < 1 ? 1 : (v > STAGE_WORKERS ? STAGE_WORKERS : v);
  }();
  return n;
}

// false when the slots cannot be had (pinned memory exhausted): the caller then copies directly
bool stage_ring() {
  if (t_ring.ready) return true;
  for (int w = 0; w < stage_workers(); ++w)
    for (int k = 0; k < STAGE_SLOTS; ++k) {
      // default event flags: completion also makes a device-to-host copy visible to the host
      if (!t_ring.slot[w][k]) {
        void *p = nullptr;
        if (hipHostMalloc(&p, stage_chunk(), hipHostMallocDefault) != hipSuccess) {
          (void)hipGetLastError();  // not an error of a later launch
          return false;
        }
        t_ring.slot[w][k] = p;
      }
      if (!t_ring.ev[w][k]) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
          (void)hipGetLastError();
          return false;
        }
        t_ring.ev[w][k] = e;
      }
    }
  t_ring.ready = true;
  return true;
}

bool host_pinned(const void *p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory may be reported as an error: not one of ours
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

// chunks w, w + nw, ... of [0, bytes) by worker w, through the calling thread's ring R (the
// workers are other threads); returns the first HIP error
hipError_t stage_chunks(StageRing &R, int w, int nw, int dev, char *d, char *h, size_t bytes, bool h2d,
                        hipStream_t s) {
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return e;
  const size_t chunk = stage_chunk();
  const size_t nch = (bytes + chunk - 1) / chunk;
  auto span = [&](size_t c, size_t *off) {
    *off = c * chunk;
    return std::min(chunk, bytes - *off);
  };
  if (h2d) {
    int i = 0;
    for (size_t c = (size_t)w; c < nch; c += (size_t)nw, ++i) {
      const int k = i % STAGE_SLOTS;
      if (R.pending[w][k] && (e = hipEventSynchronize(R.ev[w][k])) != hipSuccess) return e;
      R.pending[w][k] = false;
      size_t off;
      const size_t len = span(c, &off);
      memcpy(R.slot[w][k], h + off, len);
      if ((e = hipMemcpyAsync(d + off, R.slot[w][k], len, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
      if ((e = hipEventRecord(R.ev[w][k], s)) != hipSuccess) return e;
      R.pending[w][k] = true;
    }
    return hipSuccess;
  }
  // device to host: the DMAs of the first STAGE_SLOTS chunks go out at once; each wait is
  // followed by the host copy out of the slot and the DMA of the chunk that slot takes next
  std::vector<size_t> mine;
  for (size_t c = (size_t)w; c < nch; c += (size_t)nw) mine.push_back(c);
  auto issue = [&](size_t i) -> hipError_t {
    const int k = (int)(i % STAGE_SLOTS);
    hipError_t r;
    if (R.pending[w][k] && (r = hipEventSynchronize(R.ev[w][k])) != hipSuccess) return r;
    size_t off;
    const size_t len = span(mine[i], &off);
    if ((r = hipMemcpyAsync(R.slot[w][k], d + off, len, hipMemcpyDeviceToHost, s)) != hipSuccess) return r;
    if ((r = hipEventRecord(R.ev[w][k], s)) != hipSuccess) return r;
    R.pending[w][k] = true;
    return hipSuccess;
  };
  for (size_t i = 0; i < mine.size() && i < (size_t)STAGE_SLOTS; ++i)
    if ((e = issue(i)) != hipSuccess) return e;
  for (size_t i = 0; i < mine.size(); ++i) {
    const int k = (int)(i % STAGE_SLOTS);
    if ((e = hipEventSynchronize(R.ev[w][k])) != hipSuccess) return e;
    R.pending[w][k] = false;
    size_t off;
    const size_t len = span(mine[i], &off);
    memcpy(h + off, R.slot[w][k], len);
    if (i + STAGE_SLOTS < mine.size() && (e = issue(i + STAGE_SLOTS)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// Copy `bytes` between a caller's host buffer and device memory on stream s.  Host to device
// returns with the DMAs queued on s (later work on s is ordered after them); device to host
// returns once the data is in the host buffer.
int copy_host(void *dev_p, void *host_p, size_t bytes, bool h2d, hipStream_t s) {
  if (!bytes) return QBLAS_OK;
  if (bytes < STAGE_DIRECT || host_pinned(host_p) || !stage_ring()) {
    QB_CHECK(h2d ? hipMemcpyAsync(dev_p, host_p, bytes, hipMemcpyHostToDevice, s)
                 : hipMemcpyAsync(host_p, dev_p, bytes, hipMemcpyDeviceToHost, s));
    if (!h2d) QB_CHECK(hipStreamSynchronize(s));
    return QBLAS_OK;
  }
  StageRing &R = t_ring;
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  const size_t nch = (bytes + stage_chunk() - 1) / stage_chunk();
  const int nw = (int)std::min<size_t>(nch, (size_t)stage_workers());
  hipError_t err[STAGE_WORKERS];
  std::thread th[STAGE_WORKERS];
  bool spawned[STAGE_WORKERS] = {};
  for (int w = 1; w < nw; ++w) {
    try {
      th[w] = std::thread([&, w] { err[w] = stage_chunks(R, w, nw, dev, (char *)dev_p, (char *)host_p, bytes, h2d, s); });
      spawned[w] = true;
    } catch (...) {  // no thread to be had: this share runs on the calling thread below
    }
  }
  for (int w = 0; w < nw; ++w)
    if (!spawned[w]) err[w] = stage_chunks(R, w, nw, dev, (char *)dev_p, (char *)host_p, bytes, h2d, s);
  for (int w = 1; w < nw; ++w)
    if (spawned[w]) th[w].join();
  for (int w = 0; w < nw; ++w)
    if (err[w] != hipSuccess) {
      set_err(h2d ? "host to device copy" : "device to host copy", err[w]);
      return QBLAS_ERR_HIP;
    }
  return QBLAS_OK;
}

// A host-to-device copy staged by background threads while the calling thread goes on enqueueing
// work; once join() returns, every DMA of the copy is enqueued on its stream.
struct AsyncH2D {
  std::thread th[STAGE_WORKERS];
  hipError_t err[STAGE_WORKERS] = {};
  int nw = 0;
  hipError_t join() {
    for (auto &t : th)
      if (t.joinable()) t.join();
    for (int w = 0; w < nw; ++w)
      if (err[w] != hipSuccess) return err[w];
    return hipSuccess;
  }
  ~AsyncH2D() {
    for (auto &t : th)
      if (t.joinable()) t.join();
  }
};

int copy_h2d_async(AsyncH2D &a, void *dev_p, const void *host_p, size_t bytes, hipStream_t s) {
  if (!bytes) return QBLAS_OK;
  if (bytes < STAGE_DIRECT || host_pinned(host_p) || !stage_ring()) {
    QB_CHECK(hipMemcpyAsync(dev_p, host_p, bytes, hipMemcpyHostToDevice, s));
    return QBLAS_OK;
  }
  StageRing *R = &t_ring;
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  const size_t nch = (bytes + stage_chunk() - 1) / stage_chunk();
  const int nw = (int)std::min<size_t>(nch, (size_t)stage_workers());
  a.nw = nw;
  for (int w = 0; w < nw; ++w) {
    a.err[w] = hipSuccess;
    try {
      a.th[w] = std::thread([&a, R, w, nw, dev, dev_p, host_p, bytes, s] {
        a.err[w] = stage_chunks(*R, w, nw, dev, (char *)dev_p, (char *)host_p, bytes, true, s);
      });
    } catch (...) {  // no thread to be had: this share runs now, on the calling thread
      a.err[w] = stage_chunks(*R, w, nw, dev, (char *)dev_p, (char *)host_p, bytes, true, s);
    }
  }
  return QBLAS_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

q128 load_q(const void *p) {
  q128 v;
  memcpy(&v, p, 16);
  return v;
}

// ---------------------------------------------------------------------------------------------
// device routines
// ---------------------------------------------------------------------------------------------
int dev_dot(int mode, int64_t n, const void *x, int64_t incx, const void *y, int64_t incy, int T, void *res,
            hipStream_t s) {
  if (n < 0 || incx < 0 || incy < 0 || T < 1) {
    set_err_msg("qblas_hip_dot: negative size/stride or T < 1");
    return QBLAS_ERR_ARG;
  }
  if (!aligned16(x) || !aligned16(y) || !aligned16(res)) {
    set_err_msg("qblas_hip_dot: binary128 pointers must be 16-byte aligned");
    return QBLAS_ERR_ARG;
  }
  if (n == 0) {  // level1.hpp:83-84
    QB_CHECK(qbk::launch_zero((q128 *)res, 1, s));
    return QBLAS_OK;
  }
  if (mode == QBLAS_MODE_DD) {
    void *ws = nullptr;
    int rc = ws_acquire(t_ws, qbk::dot_dd_ws(n), s, &ws);
    if (rc) return rc;
    QB_CHECK(qbk::launch_dot_dd((const q128 *)x, incx, (const q128 *)y, incy, n, ws, (q128 *)res, s));
    return ws_release(t_ws, s);
  }
  if (mode == QBLAS_MODE_EXACT) {
    void *ws = nullptr;
    const StreamWorkspaces::Entry *we = nullptr;
    int rc = ws_acquire(t_ws, qbk::dot_exact_ws(n), s, &ws, &we);
    if (rc) return rc;
    QB_CHECK(qbk::launch_dot_exact((const q128 *)x, incx, (const q128 *)y, incy, n, ws, (q128 *)res, s));
    set_status(ws, we, false);
    return ws_release(t_ws, s);
  }
  if (mode != QBLAS_MODE_REFERENCE) {
    set_err_msg("qblas_hip_dot: mode not available");
    return QBLAS_ERR_MODE;
  }
  void *part = nullptr;
  int rc = ws_acquire(t_ws, (size_t)2 * (size_t)T * 16, s, &part);
  if (rc) return rc;
  QB_CHECK(qbk::launch_dot_ref((const q128 *)x, incx, (const q128 *)y, incy, n, T, (q128 *)part, (q128 *)res, s));
  return ws_release(t_ws, s);
}

int dev_gemv(int mode, int layout, int64_t m, int64_t n, q128 alpha, const void *A, int64_t lda, const void *x,
             int64_t incx, q128 beta, void *y, int64_t incy, hipStream_t s) {
  if (m < 0 || n < 0 || lda < 0 || incx < 0 || incy < 0) {
    set_err_msg("qblas_hip_gemv: negative size or stride");
    return QBLAS_ERR_ARG;
  }
  if (!aligned16(A) || !aligned16(x) || !aligned16(y)) {
    set_err_msg("qblas_hip_gemv: binary128 pointers must be 16-byte aligned");
    return QBLAS_ERR_ARG;
  }
  if (m == 0 || n == 0) return QBLAS_OK;  // level2.hpp:21-22,59-60
  const int colmajor = layout != 0;
  if (mode == QBLAS_MODE_DD) {
    void *ws = nullptr;
    int rc = ws_acquire(t_ws, qbk::gemv_dd_ws(colmajor, m, n), s, &ws);
    if (rc) return rc;
    QB_CHECK(qbk::launch_gemv_dd(colmajor, m, n, alpha, (const q128 *)A, lda, (const q128 *)x, incx, beta,
                                 (q128 *)y, incy, ws, s));
    return ws_release(t_ws, s);
  }
  if (mode == QBLAS_MODE_EXACT) {
    void *ws = nullptr;
    const StreamWorkspaces::Entry *we = nullptr;
    int rc = ws_acquire(t_ws, qbk::gemv_exact_ws(colmajor, m, n), s, &ws, &we);
    if (rc) return rc;
    QB_CHECK(qbk::launch_gemv_exact(colmajor, m, n, alpha, (const q128 *)A, lda, (const q128 *)x, incx, beta,
                                    (q128 *)y, incy, ws, s));
    set_status(ws, we, false);
    return ws_release(t_ws, s);
  }
  if (mode != QBLAS_MODE_REFERENCE) {
    set_err_msg("qblas_hip_gemv: mode not available");
    return QBLAS_ERR_MODE;
  }
  void *ax = nullptr;
  if (colmajor) {
    int rc = ws_acquire(t_ws, (size_t)n * 16, s, &ax);
    if (rc) return rc;
  }
  QB_CHECK(qbk::launch_gemv_ref(colmajor, m, n, alpha, (const q128 *)A, lda, (const q128 *)x, incx, beta,
                                (q128 *)y, incy, (q128 *)ax, s));
  if (colmajor) return ws_release(t_ws, s);
  return QBLAS_OK;
}

// Precision of the OZAKI form from the data: a probe computes 256 sampled entries exactly and
// measures how far below 2^(E_i + F_j) they lie; the returned value is the third lowest (so that
// the precision chosen from it certifies all but about 1% of the entries, the rest being left to
// the exact repair), INT_MAX when no sample is usable.  The probe synchronizes with the stream.
thread_local Workspace t_pws;
// The probe's workspace stays acquired (its scaling exponents serve the call); the caller releases
// it after enqueuing the call.
thread_local int *t_probe_out = nullptr;  // pinned: the probe kernel's results

// The CRT method's tables depend on the modulus count only: built once per (device, count) by this
// thread, waited for once, then read by every later call (read-only, so any stream may use them).
thread_local std::map<std::pair<int, int>, void *> t_crt_tabs;
int crt_tables(int nq, hipStream_t s, const void **out) {
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  auto it = t_crt_tabs.find({dev, nq});
  if (it == t_crt_tabs.end()) {
    void *p = nullptr;
    QB_CHECK(hipMalloc(&p, qbk::crt_tab_bytes()));
    hipError_t e = qbk::launch_crt_tables(nq, p, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipFree(p);
      set_err("crt tables", e);
      return QBLAS_ERR_HIP;
    }
    it = t_crt_tabs.emplace(std::make_pair(dev, nq), p).first;
  }
  *out = it->second;
  return QBLAS_OK;
}
int oz_probe(int colmajor, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb,
             hipStream_t s, int *out, void **pws) {
  void *ws;
  int rc = ws_acquire(t_pws, qbk::ozaki_probe_ws(m, n), s, &ws);
  if (rc) return rc;
  *pws = ws;
  constexpr int P = 256;
  // the probe writes its P results straight into pinned host memory (no copy after the kernel)
  if (!t_probe_out) QB_CHECK(hipHostMalloc((void **)&t_probe_out, P * sizeof(int), hipHostMallocDefault));
  QB_CHECK(qbk::launch_ozaki_probe(colmajor, m, n, k, (const q128 *)A, lda, (const q128 *)B, ldb, P, ws, s,
                                   t_probe_out));
  QB_CHECK(hipStreamSynchronize(s));
  int r[P];
  memcpy(r, t_probe_out, sizeof(r));
  std::sort(r, r + P);
  *out = r[2];
  return QBLAS_OK;
}

// Method and precision of an OZAKI call.  Fixed settings win (a slice count selects the slice
// method; CRT bits the CRT method); otherwise the probe's value sets both: the slice count S and
// the CRT bits t whose certification bounds pass the probe's entries.  AUTO takes the CRT method
// when t fits (<= 158 bits) and its padded GEMM work (moduli x 256-tiles) is below the slice
// method's (slice pairs x 64-tiles).
struct oz_choice {
  int method, S, t;
  void *probe_ws;  // the probe's workspace (held until the call is enqueued), or null
};
int oz_choose(int colmajor, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb,
              hipStream_t s, oz_choice *c) {
  const int method = g_oz_method.load(), fixedS = g_oz_slices.load(), fixedT = g_oz_bits.load();
  c->probe_ws = nullptr;
  c->S = fixedS;
  c->t = fixedT;
  if (method == QBLAS_OZAKI_SLICES || (method == QBLAS_OZAKI_AUTO && fixedS)) {
    c->method = QBLAS_OZAKI_SLICES;
    if (fixedS) return QBLAS_OK;
  } else if (fixedT && (method == QBLAS_OZAKI_CRT || method == QBLAS_OZAKI_AUTO)) {
    c->method = QBLAS_OZAKI_CRT;
    return QBLAS_OK;
  } else {
    c->method = method;
  }
  int r = 0;
  int rc = oz_probe(colmajor, m, n, k, A, lda, B, ldb, s, &r, &c->probe_ws);
  if (rc) return rc;
  c->S = qbk::ozaki_slices_for(k, r);
  c->t = qbk::ozaki_crt_bits_for(k, r);
  if (c->method == QBLAS_OZAKI_SLICES) return QBLAS_OK;
  if (!c->t) {  // beyond the CRT method's range
    c->method = QBLAS_OZAKI_SLICES;
    return QBLAS_OK;
  }
  if (c->method == QBLAS_OZAKI_AUTO) {
    auto up = [](int64_t v, int64_t q) { return (double)((v + q - 1) / q * q); };
    const double crt = (double)qbk::crt_moduli_for(k, c->t) * up(m, 256) * up(n, 256);
    const double sl = (double)c->S * (c->S + 1) / 2 * up(m, 64) * up(n, 64);
    c->method = crt < sl ? QBLAS_OZAKI_CRT : QBLAS_OZAKI_SLICES;
  }
  return QBLAS_OK;
}

// gate: null, or the hook that makes s wait for C's copy in (called before the first kernel that
// reads C is enqueued; the OZAKI form enqueues its probe, scaling, residues or slices and products
// first)
int dev_gemm(int mode, int layout, int64_t m, int64_t n, int64_t k, q128 alpha, const void *A, int64_t lda,
             const void *B, int64_t ldb, q128 beta, void *C, int64_t ldc, hipStream_t s,
             const qbk::c_gate *gate = nullptr) {
  if (m < 0 || n < 0 || k < 0 || lda < 0 || ldb < 0 || ldc < 0) {
    set_err_msg("qblas_hip_gemm: negative size or leading dimension");
    return QBLAS_ERR_ARG;
  }
  if (!aligned16(A) || !aligned16(B) || !aligned16(C)) {
    set_err_msg("qblas_hip_gemm: binary128 pointers must be 16-byte aligned");
    return QBLAS_ERR_ARG;
  }
  if (m == 0 || n == 0 || k == 0) return QBLAS_OK;  // level3.hpp:221-222
  const int colmajor = layout != 0;
  if (gate && mode != QBLAS_MODE_OZAKI) QB_CHECK(gate->fn(gate->arg, s));
  if (mode == QBLAS_MODE_OZAKI) {
    oz_choice ch;
    int rc0 = oz_choose(colmajor, m, n, k, A, lda, B, ldb, s, &ch);
    if (rc0) return rc0;
    // release the probe's workspace once the call's work is enqueued (its exponents are read there)
    struct probe_hold {
      void *ws;
      hipStream_t s;
      ~probe_hold() {
        if (ws) (void)ws_release(t_pws, s);
      }
    } hold{ch.probe_ws, s};
    const bool crt = ch.method == QBLAS_OZAKI_CRT;
    const size_t need = crt ? qbk::gemm_crt_ws(m, n, k, ch.t) : qbk::gemm_ozaki_ws(m, n, k, ch.S);
    if (!need) {
      set_err_msg("qblas_hip_gemm: the CRT bit count needs more moduli than the method has");
      return QBLAS_ERR_ARG;
    }
    t_oz_method = ch.method;
    t_oz_slices = crt ? 0 : ch.S;
    t_oz_bits = crt ? ch.t : 0;
    t_oz_moduli = crt ? qbk::crt_moduli_for(k, ch.t) : 0;
    void *ws;
    const StreamWorkspaces::Entry *we = nullptr;
    int rc = ws_acquire(t_ws, need, s, &ws, &we);
    if (rc) return rc;
    hipEvent_t *ev = nullptr;
    if (g_phase_timing.load()) {
      if (!t_oz_ev[0])
        for (auto &e : t_oz_ev) QB_CHECK(hipEventCreate(&e));
      ev = t_oz_ev;
    }
    t_oz_ev_valid = ev != nullptr;
    const void *tab = nullptr;
    if (crt && (rc = crt_tables(t_oz_moduli, s, &tab))) return rc;
    if (crt)
      QB_CHECK(qbk::launch_gemm_crt(colmajor, m, n, k, ch.t, alpha, (const q128 *)A, lda, (const q128 *)B, ldb, beta,
                                    (q128 *)C, ldc, ws, s, ev, ch.probe_ws, gate, tab));
    else
      QB_CHECK(qbk::launch_gemm_ozaki(colmajor, m, n, k, ch.S, alpha, (const q128 *)A, lda, (const q128 *)B, ldb,
                                      beta, (q128 *)C, ldc, ws, s, ev, ch.probe_ws, gate));
    set_status(ws, we, true);
    return ws_release(t_ws, s);
  }
  if (mode == QBLAS_MODE_DD) {
    void *ws;
    int rc = ws_acquire(t_ws, qbk::gemm_dd_ws(m, n, k), s, &ws);
    if (rc) return rc;
    QB_CHECK(qbk::launch_gemm_dd(colmajor, m, n, k, alpha, (const q128 *)A, lda, (const q128 *)B, ldb, beta,
                                 (q128 *)C, ldc, ws, s));
    return ws_release(t_ws, s);
  }
  if (mode != QBLAS_MODE_REFERENCE) {
    set_err_msg("qblas_hip_gemm: mode not available");
    return QBLAS_ERR_MODE;
  }
  // gemm_simple for m,n,k <= 64 is one chain over all of k; otherwise blocks of kc
  // (level3.hpp:224-229, :244-246).  For k <= kc both are one block.
  const int64_t kc = (m <= 64 && n <= 64 && k <= 64) ? k : (int64_t)g_kc.load();
  QB_CHECK(qbk::launch_gemm_ref(colmajor, m, n, k, kc, alpha, (const q128 *)A, lda, (const q128 *)B, ldb, beta,
                                (q128 *)C, ldc, s));
  return QBLAS_OK;
}

int dev_axpy(int64_t n, q128 alpha, const void *x, int64_t incx, void *y, int64_t incy, hipStream_t s) {
  if (n <= 0) return QBLAS_OK;  // c_interface.hpp:49-50, level1.hpp:192-193
  if (incx < 0 || incy < 0) {
    set_err_msg("qblas_hip_axpy: negative stride");
    return QBLAS_ERR_ARG;
  }
  if (!aligned16(x) || !aligned16(y)) {
    set_err_msg("qblas_hip_axpy: binary128 pointers must be 16-byte aligned");
    return QBLAS_ERR_ARG;
  }
  QB_CHECK(qbk::launch_axpy(n, alpha, (const q128 *)x, incx, (q128 *)y, incy, s));
  return QBLAS_OK;
}

// RN(sqrt(dot(x, x))) with the dot of the given mode (c_interface.hpp:34-46; cpp_classes.hpp:78-81)
int dev_nrm2(int mode, int64_t n, const void *x, int64_t incx, int T, void *res, hipStream_t s) {
  int rc = dev_dot(mode, n < 0 ? 0 : n, x, incx, x, incx, T, res, s);
  if (rc) return rc;
  QB_CHECK(qbk::launch_sqrt1((q128 *)res, s));
  return QBLAS_OK;
}

// ---------------------------------------------------------------------------------------------
// host-pointer wrappers
// ---------------------------------------------------------------------------------------------
// bytes spanned by `count` strided elements
size_t span_bytes(int64_t count, int64_t inc) { return count <= 0 ? 0 : (size_t)((count - 1) * inc + 1) * 16; }

int stage_in(int slot, const void *host, size_t bytes, hipStream_t s, void **dev) {
  int rc = ws_acquire(t_stage[slot], bytes ? bytes : 16, s, dev);
  if (rc) return rc;
  return copy_host(*dev, (void *)host, bytes, true, s);
}

int host_dot(int64_t n, const void *x, int64_t incx, const void *y, int64_t incy, q128 *out) {
  if (n <= 0) {
    *out = qb::q_zero(0);
    return QBLAS_OK;
  }
  if (incx < 0 || incy < 0) {
    set_err_msg("quadblas dot: negative stride");
    return QBLAS_ERR_ARG;
  }
  t_exact_status = StatusRef();  // check_exact_status below sees this call only
  hipStream_t s;
  int rc = host_stream(&s);
  if (rc) return rc;
  void *dx, *dy, *dr;
  if ((rc = stage_in(0, x, span_bytes(n, incx), s, &dx))) return rc;
  if ((rc = stage_in(1, y, span_bytes(n, incy), s, &dy))) return rc;
  if ((rc = stage_in(2, nullptr, 0, s, &dr))) return rc;
  if ((rc = dev_dot(vec_mode(), n, dx, incx, dy, incy, cur_threads(), dr, s))) return rc;
  q128 r;
  QB_CHECK(hipMemcpyAsync(&r, dr, 16, hipMemcpyDeviceToHost, s));
  QB_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < 3; ++i) ws_release(t_stage[i], s);
  *out = r;
  return check_exact_status(s);
}

int host_gemv(int layout, int64_t m, int64_t n, q128 alpha, const void *A, int64_t lda, const void *x,
              int64_t incx, q128 beta, void *y, int64_t incy) {
  if (m <= 0 || n <= 0) return QBLAS_OK;
  if (lda < 0 || incx < 0 || incy < 0) {
    set_err_msg("quadblas gemv: negative stride");
    return QBLAS_ERR_ARG;
  }
  t_exact_status = StatusRef();  // check_exact_status below sees this call only
  hipStream_t s;
  int rc = host_stream(&s);
  if (rc) return rc;
  const size_t abytes = layout == 0 ? span_bytes(m, lda) - 16 + (size_t)n * 16 : span_bytes(n, lda) - 16 + (size_t)m * 16;
  const size_t xbytes = span_bytes(n, incx), ybytes = span_bytes(m, incy);
  void *dA, *dx, *dy;
  if ((rc = stage_in(0, A, abytes, s, &dA))) return rc;
  if ((rc = stage_in(1, x, xbytes, s, &dx))) return rc;
  if ((rc = stage_in(2, y, ybytes, s, &dy))) return rc;
  if ((rc = dev_gemv(vec_mode(), layout, m, n, alpha, dA, lda, dx, incx, beta, dy, incy, s))) return rc;
  if ((rc = copy_host(dy, y, ybytes, false, s))) return rc;
  for (int i = 0; i < 3; ++i) ws_release(t_stage[i], s);
  return check_exact_status(s);
}

int host_gemm(int layout, int64_t m, int64_t n, int64_t k, q128 alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, q128 beta, void *C, int64_t ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return QBLAS_OK;
  if (lda < 0 || ldb < 0 || ldc < 0) {
    set_err_msg("quadblas gemm: negative leading dimension");
    return QBLAS_ERR_ARG;
  }
  t_exact_status = StatusRef();  // check_exact_status below sees this call only
  hipStream_t s;
  int rc = host_stream(&s);
  if (rc) return rc;
  size_t abytes, bbytes, cbytes;
  if (layout == 0) {
    abytes = (size_t)((m - 1) * lda + k) * 16;
    bbytes = (size_t)((k - 1) * ldb + n) * 16;
    cbytes = (size_t)((m - 1) * ldc + n) * 16;
  } else {
    abytes = (size_t)((k - 1) * lda + m) * 16;
    bbytes = (size_t)((n - 1) * ldb + k) * 16;
    cbytes = (size_t)((n - 1) * ldc + m) * 16;
  }
  void *dA, *dB, *dC;
  if ((rc = stage_in(0, A, abytes, s, &dA))) return rc;
  if ((rc = stage_in(1, B, bbytes, s, &dB))) return rc;
  // C goes in after A and B, on a second stream and by background threads, while the kernels
  // that do not read C run on s (the OZAKI form's probe, residues or slices, and products); the
  // gate makes s wait for C's copy before the first kernel that reads it
  hipStream_t s2;
  if ((rc = host_stream2(&s2))) return rc;
  if ((rc = ws_acquire(t_stage[2], cbytes, s, &dC))) return rc;
  QB_CHECK(hipEventRecord(t_ev_ab, s));
  QB_CHECK(hipStreamWaitEvent(s2, t_ev_ab, 0));
  AsyncH2D cin;
  if ((rc = copy_h2d_async(cin, dC, C, cbytes, s2))) return rc;
  struct GateArg {
    AsyncH2D *a;
    hipStream_t s2;
    bool called;
  } ga{&cin, s2, false};
  const qbk::c_gate gate{[](void *p, hipStream_t st) -> hipError_t {
                           auto *g = (GateArg *)p;
                           g->called = true;
                           hipError_t e = g->a->join();
                           if (e == hipSuccess) e = hipEventRecord(t_ev_c, g->s2);
                           if (e == hipSuccess) e = hipStreamWaitEvent(st, t_ev_c, 0);
                           return e;
                         },
                         &ga};
  if ((rc = dev_gemm(mat_mode(), layout, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, s, &gate))) return rc;
  if (!ga.called) QB_CHECK(gate.fn(gate.arg, s));  // C is copied back even if no kernel read it
  if ((rc = copy_host(dC, C, cbytes, false, s))) return rc;
  for (int i = 0; i < 3; ++i) ws_release(t_stage[i], s);
  return check_exact_status(s);
}

// ---------------------------------------------------------------------------------------------
// batched problems: grow-only buffers per (device, stream) of the calling thread (t_bws: the
// prefix sums, the EXACT status words, the OZAKI batch's arrays), so repeated calls of the same
// count launch kernels and nothing else, and a buffer used in a captured graph is never freed;
// and the tuned tile configuration per (mode, layout, count)
// ---------------------------------------------------------------------------------------------
int batch_ws(size_t bytes, hipStream_t s, void **out, const StreamWorkspaces::Entry **e = nullptr) {
  return ws_acquire(t_bws, bytes, s, out, e);
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int64_t>, int> g_tuned;

int tuned_config(int mode, int layout, int64_t count) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(std::make_tuple(mode, layout, count));
  return it == g_tuned.end() ? 0 : it->second;
}

int check_batched(const char *what, int mode, int64_t count, const void *P, bool gemm = false) {
  if (count < 0 || (count > 0 && (!P || ((uintptr_t)P % 8) != 0))) {
    set_err_msg(std::string(what) + ": bad count or descriptor pointer");
    return QBLAS_ERR_ARG;
  }
  if (mode != QBLAS_MODE_REFERENCE && mode != QBLAS_MODE_DD && !(gemm && mode == QBLAS_MODE_OZAKI)) {
    set_err_msg(std::string(what) + ": batched problems support QBLAS_MODE_REFERENCE and QBLAS_MODE_DD" +
                (gemm ? " (gemm also QBLAS_MODE_OZAKI)" : ""));
    return QBLAS_ERR_MODE;
  }
  if (gemm && mode == QBLAS_MODE_OZAKI && count > 0xffffffffll) {
    set_err_msg(std::string(what) + ": QBLAS_MODE_OZAKI takes at most 2^32 - 1 problems");
    return QBLAS_ERR_ARG;
  }
  return QBLAS_OK;
}

int batched_configs(int mode) {
  return mode == QBLAS_MODE_OZAKI ? qbk::ozaki_batched_configs() : qbk::gemm_batched_configs(mode == QBLAS_MODE_DD);
}

// Totals of batched OZAKI calls by (device, descriptor array, count, layout, slices): a call under
// graph capture cannot read its totals back, so it sizes the workspace and grids from those of the
// last eager call on the same descriptors, and the device checks at each replay that the
// descriptors still give them (otherwise nothing is computed and the status reports it).
using BoKey = std::tuple<int, uintptr_t, int64_t, int, int>;
std::mutex g_bo_mu;
std::map<BoKey, std::array<int64_t, qbk::OZ_BATCHED_TOTALS>> g_bo_plans;

// Batched OZAKI form with S slices: the per-problem counts are summed on the device and their
// totals read back (one synchronization) to size the workspace, then the grouped launches follow.
// Under graph capture the totals of the last eager call on the same descriptors are used instead.
int gemm_batched_ozaki(int layout, int64_t count, const qblas_gemm_problem *P, int S, hipStream_t s) {
  void *ws, *ws2;
  const StreamWorkspaces::Entry *we = nullptr;
  int rc;
  int dev = 0;
  QB_CHECK(hipGetDevice(&dev));
  const BoKey key(dev, (uintptr_t)P, count, layout != 0, S);
  if (stream_capturing(s)) {
    std::array<int64_t, qbk::OZ_BATCHED_TOTALS> plan;
    {
      std::lock_guard<std::mutex> lk(g_bo_mu);
      auto it = g_bo_plans.find(key);
      if (it == g_bo_plans.end()) {
        set_err_msg("qblas_hip_gemm_batched: QBLAS_MODE_OZAKI under graph capture needs one eager call on the same "
                    "descriptors (same count, layout and configuration) first");
        return QBLAS_ERR_ARG;
      }
      plan = it->second;
    }
    if ((rc = batch_ws(qbk::gemm_ozaki_batched_ws(count, plan.data(), S), s, &ws2, &we))) return rc;
    QB_CHECK(qbk::launch_gemm_ozaki_batched_scan(layout != 0, count, P, ws2, s));
    QB_CHECK(qbk::launch_gemm_ozaki_batched(layout != 0, count, P, S, plan.data(), ws2, s, true));
    set_status(ws2, we, true);
    t_oz_method = QBLAS_OZAKI_SLICES;
    t_oz_slices = S;
    t_oz_bits = 0;
    t_oz_moduli = 0;
    t_oz_ev_valid = false;
    return QBLAS_OK;
  }
  if ((rc = batch_ws(qbk::gemm_ozaki_batched_scan_ws(count), s, &ws))) return rc;
  QB_CHECK(qbk::launch_gemm_ozaki_batched_scan(layout != 0, count, P, ws, s));
  int64_t tot[qbk::OZ_BATCHED_TOTALS];
  QB_CHECK(hipMemcpyAsync(tot, (const char *)ws + qbk::gemm_ozaki_batched_totals_offset(count), sizeof(tot),
                          hipMemcpyDeviceToHost, s));
  QB_CHECK(hipStreamSynchronize(s));
  if ((rc = batch_ws(qbk::gemm_ozaki_batched_ws(count, tot, S), s, &ws2, &we))) return rc;
  if (ws2 != ws) QB_CHECK(qbk::launch_gemm_ozaki_batched_scan(layout != 0, count, P, ws2, s));  // grown: new buffer
  QB_CHECK(qbk::launch_gemm_ozaki_batched(layout != 0, count, P, S, tot, ws2, s));
  {
    std::lock_guard<std::mutex> lk(g_bo_mu);
    auto &pl = g_bo_plans[key];
    for (int f = 0; f < qbk::OZ_BATCHED_TOTALS; ++f) pl[f] = tot[f];
  }
  set_status(ws2, we, true);
  t_oz_method = QBLAS_OZAKI_SLICES;
  t_oz_slices = S;
  t_oz_bits = 0;
  t_oz_moduli = 0;
  t_oz_ev_valid = false;
  return QBLAS_OK;
}

int gemm_batched(int mode, int layout, int64_t count, const qblas_gemm_problem *P, int cfg, hipStream_t s) {
  int rc = check_batched("qblas_hip_gemm_batched", mode, count, P, true);
  if (rc || count == 0) return rc;
  if (mode == QBLAS_MODE_OZAKI) return gemm_batched_ozaki(layout, count, P, qbk::ozaki_batched_slices(cfg), s);
  void *ws;
  if ((rc = batch_ws(qbk::batched_ws(count), s, &ws))) return rc;
  QB_CHECK(qbk::launch_gemm_batched(mode == QBLAS_MODE_DD, layout != 0, count, P, cfg, (int64_t)g_kc.load(), ws, s));
  return QBLAS_OK;
}

// Replica of the first min(count, 512) problems: same m, n, k, packed leading dimensions,
// synthetic full-mantissa data; every configuration runs once untimed and three times timed.
// QBLAS_MODE_OZAKI (configurations = slice counts): the replica reads the caller's A and B, since
// the share of entries the certification sends to the exact repair depends on the data; only C
// is replaced.
int gemm_batched_tune(int mode, int layout, int64_t count, const qblas_gemm_problem *dP, hipStream_t s) {
  int rc = check_batched("qblas_hip_gemm_batched_tune", mode, count, dP, true);
  if (rc) return -rc;
  if (count == 0) return 0;
  const bool oz = mode == QBLAS_MODE_OZAKI;
  const int64_t R = count < 512 ? count : 512;
  std::vector<qblas_gemm_problem> hp(R);
  {
    hipError_t e = hipMemcpyAsync(hp.data(), dP, R * sizeof(qblas_gemm_problem), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_err("tune: reading the descriptors", e);
      return -QBLAS_ERR_HIP;
    }
  }
  int64_t elems = 0;
  for (auto &p : hp) {
    const int64_t m = p.m > 0 ? p.m : 0, n = p.n > 0 ? p.n : 0, k = p.k > 0 ? p.k : 0;
    elems += (oz ? 0 : m * k + k * n) + m * n;
  }
  char *scratch = nullptr;
  const size_t dbytes = R * sizeof(qblas_gemm_problem);
  const size_t bytes = (size_t)elems * 16 + dbytes + qbk::batched_ws(R) + 256;
  {
    const hipError_t e = hipMalloc(&scratch, bytes);
    if (e != hipSuccess) {
      set_err("tune: scratch allocation", e);
      return -QBLAS_ERR_HIP;
    }
  }
  std::vector<qblas_gemm_problem> rp(hp);
  q128 *data = (q128 *)scratch;
  int64_t off = 0;
  for (auto &p : rp) {
    const int64_t m = p.m > 0 ? p.m : 0, n = p.n > 0 ? p.n : 0, k = p.k > 0 ? p.k : 0;
    if (!oz) {
      p.A = data + off;
      p.lda = layout ? (m > 0 ? m : 1) : (k > 0 ? k : 1);
      off += m * k;
      p.B = data + off;
      p.ldb = layout ? (k > 0 ? k : 1) : (n > 0 ? n : 1);
      off += k * n;
    }
    p.C = data + off;
    p.ldc = layout ? (m > 0 ? m : 1) : (n > 0 ? n : 1);
    off += m * n;
  }
  char *dR = scratch + (size_t)elems * 16;
  char *ws = dR + ((dbytes + 255) / 256) * 256;
  int best = 0;
  float best_ms = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto fail = [&](hipError_t e, const char *where) {
    set_err(where, e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    return -QBLAS_ERR_HIP;
  };
  hipError_t e;
  if ((e = qbk::launch_fill(0, 1234, elems, 0, 0, data, s)) != hipSuccess) return fail(e, "tune fill");
  if ((e = hipMemcpyAsync(dR, rp.data(), dbytes, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e, "tune copy");
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return fail(e, "tune events");
  const int ncfg = batched_configs(mode);
  const qblas_gemm_problem *P = (const qblas_gemm_problem *)dR;
  auto launch = [&](int c) -> hipError_t {
    if (!oz) return qbk::launch_gemm_batched(mode == QBLAS_MODE_DD, layout != 0, R, P, c, (int64_t)g_kc.load(), ws, s);
    if (gemm_batched_ozaki(layout, R, P, qbk::ozaki_batched_slices(c), s) != QBLAS_OK) {
      const hipError_t le = hipGetLastError();
      return le != hipSuccess ? le : hipErrorUnknown;
    }
    return hipSuccess;
  };
  for (int c = 0; c < ncfg; ++c) {
    if ((e = launch(c)) != hipSuccess) return fail(e, "tune launch");
    if ((e = hipEventRecord(e0, s)) != hipSuccess) return fail(e, "tune record");
    for (int rep = 0; rep < 3; ++rep)
      if ((e = launch(c)) != hipSuccess) return fail(e, "tune launch");
    if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess)
      return fail(e, "tune sync");
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return fail(e, "tune time");
    if (c == 0 || ms < best_ms) {
      best = c;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (oz) {  // the replica's status and repair count are not the caller's
    t_exact_status = StatusRef();
    t_oz_hdr = StatusRef();
  }
  {
    const hipError_t e = hipFree(scratch);
    if (e != hipSuccess) {
      set_err("tune: freeing the scratch", e);
      return -QBLAS_ERR_HIP;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(mode, layout, count)] = best;
  }
  return best;
}

int host_axpy(int64_t n, q128 alpha, const void *x, int64_t incx, void *y, int64_t incy) {
  if (n <= 0) return QBLAS_OK;
  if (incx < 0 || incy < 0) {
    set_err_msg("quadblas axpy: negative stride");
    return QBLAS_ERR_ARG;
  }
  hipStream_t s;
  int rc = host_stream(&s);
  if (rc) return rc;
  const size_t xb = span_bytes(n, incx), yb = span_bytes(n, incy);
  void *dx, *dy;
  if ((rc = stage_in(0, x, xb, s, &dx))) return rc;
  if ((rc = stage_in(1, y, yb, s, &dy))) return rc;
  if ((rc = dev_axpy(n, alpha, dx, incx, dy, incy, s))) return rc;
  if ((rc = copy_host(dy, y, yb, false, s))) return rc;
  for (int i = 0; i < 2; ++i) ws_release(t_stage[i], s);
  return QBLAS_OK;
}

int host_nrm2(int64_t n, const void *x, int64_t incx, q128 *out) {
  if (n <= 0) {
    *out = qb::q_zero(0);
    return QBLAS_OK;
  }
  if (incx < 0) {
    set_err_msg("quadblas nrm2: negative stride");
    return QBLAS_ERR_ARG;
  }
  t_exact_status = StatusRef();  // check_exact_status below sees this call only
  hipStream_t s;
  int rc = host_stream(&s);
  if (rc) return rc;
  void *dx, *dr;
  if ((rc = stage_in(0, x, span_bytes(n, incx), s, &dx))) return rc;
  if ((rc = stage_in(2, nullptr, 0, s, &dr))) return rc;
  if ((rc = dev_nrm2(vec_mode(), n, dx, incx, cur_threads(), dr, s))) return rc;
  q128 r;
  QB_CHECK(hipMemcpyAsync(&r, dr, 16, hipMemcpyDeviceToHost, s));
  QB_CHECK(hipStreamSynchronize(s));
  ws_release(t_stage[0], s);
  ws_release(t_stage[2], s);
  *out = r;
  return check_exact_status(s);
}

bool is_colmajor_char(char c) { return c == 'C' || c == 'c'; }
bool is_trans_char(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }

}  // namespace

namespace qbk {
void record_error(const char *msg) { set_err_msg(msg); }
void forget_exact_status() { t_exact_status = StatusRef(); }
}  // namespace qbk

// =============================================================================================
// exported C ABI
// =============================================================================================
extern "C" {

double quadblas_qdot(int n, void *x, int incx, void *y, int incy) {
  q128 r;
  if (host_dot(n, x, incx, y, incy, &r) != QBLAS_OK) return std::nan("");
  return qb::to_double(r);
}

void quadblas_qgemv(char layout, char trans, int m, int n, double alpha, void *A, int lda, void *x, int incx,
                    double beta, void *y, int incy) {
  int lay = is_colmajor_char(layout) ? 1 : 0;  // c_interface.hpp:76
  if (is_trans_char(trans)) {                  // c_interface.hpp:79-85
    const int t = m;
    m = n;
    n = t;
    lay = 1 - lay;
  }
  host_gemv(lay, m, n, qb::from_double(alpha), A, lda, x, incx, qb::from_double(beta), y, incy);
}

void quadblas_qgemm(char layout, char transa, char transb, int m, int n, int k, double alpha, void *A, int lda,
                    void *B, int ldb, double beta, void *C, int ldc) {
  (void)transa;  // c_interface.hpp:109-112: ignored
  (void)transb;
  host_gemm(is_colmajor_char(layout) ? 1 : 0, m, n, k, qb::from_double(alpha), A, lda, B, ldb,
            qb::from_double(beta), C, ldc);
}

double quadblas_qnrm2(int n, void *x, int incx) {
  q128 r;
  if (host_nrm2(n, x, incx, &r) != QBLAS_OK) return std::nan("");
  return qb::to_double(r);
}

void quadblas_qaxpy(int n, double alpha, void *x, int incx, void *y, int incy) {
  host_axpy(n, qb::from_double(alpha), x, incx, y, incy);  // c_interface.hpp:49-58
}

void quadblas_set_num_threads(int num_threads) { g_threads.store(num_threads > 0 ? num_threads : 0); }
int quadblas_get_num_threads(void) { return cur_threads(); }

const char *quadblas_get_version(void) { return "QuadBLAS 1.0.0 - MI355X (gfx950) HIP backend"; }
int quadblas_is_aligned(const void *ptr) { return ((uintptr_t)ptr % 32u) == 0; }

// On failure the binary128 result is NaN (as quadblas_qdot's double), never a stale value.
void quadblas_dot_q(size_t n, const void *x, size_t incx, const void *y, size_t incy, void *result) {
  q128 r;
  if (host_dot((int64_t)n, x, (int64_t)incx, y, (int64_t)incy, &r) != QBLAS_OK) r = qb::q_nan();
  memcpy(result, &r, 16);
}
void quadblas_axpy_q(size_t n, const void *alpha, const void *x, size_t incx, void *y, size_t incy) {
  host_axpy((int64_t)n, load_q(alpha), x, (int64_t)incx, y, (int64_t)incy);
}
void quadblas_nrm2_q(size_t n, const void *x, size_t incx, void *result) {
  q128 r;
  if (host_nrm2((int64_t)n, x, (int64_t)incx, &r) != QBLAS_OK) r = qb::q_nan();
  memcpy(result, &r, 16);
}
void quadblas_gemv_q(int layout, size_t m, size_t n, const void *alpha, const void *A, size_t lda, const void *x,
                     size_t incx, const void *beta, void *y, size_t incy) {
  host_gemv(layout != 0, (int64_t)m, (int64_t)n, load_q(alpha), A, (int64_t)lda, x, (int64_t)incx, load_q(beta), y,
            (int64_t)incy);
}
void quadblas_gemm_q(int layout, size_t m, size_t n, size_t k, const void *alpha, const void *A, size_t lda,
                     const void *B, size_t ldb, const void *beta, void *C, size_t ldc) {
  host_gemm(layout != 0, (int64_t)m, (int64_t)n, (int64_t)k, load_q(alpha), A, (int64_t)lda, B, (int64_t)ldb,
            load_q(beta), C, (int64_t)ldc);
}

int qblas_set_mode(int mode) {
  if (mode < QBLAS_MODE_REFERENCE || mode > QBLAS_MODE_OZAKI) return QBLAS_ERR_ARG;
  g_mode.store(mode);
  return QBLAS_OK;
}
int qblas_get_mode(void) { return cur_mode(); }
int qblas_set_ozaki_slices(int slices) {
  if (slices != 0 && (slices < qbk::OZ_SMIN || slices > qbk::OZ_SMAX)) return QBLAS_ERR_ARG;
  g_oz_slices.store(slices);
  return QBLAS_OK;
}
int qblas_get_ozaki_slices(void) { return g_oz_slices.load(); }
int qblas_set_ozaki_method(int method) {
  if (method != QBLAS_OZAKI_AUTO && method != QBLAS_OZAKI_SLICES && method != QBLAS_OZAKI_CRT) return QBLAS_ERR_ARG;
  g_oz_method.store(method);
  return QBLAS_OK;
}
int qblas_get_ozaki_method(void) { return g_oz_method.load(); }
int qblas_set_ozaki_crt_bits(int bits) {
  if (bits != 0 && (bits < qbk::OZ_CRT_TMIN || bits > qbk::OZ_CRT_TMAX)) return QBLAS_ERR_ARG;
  g_oz_bits.store(bits);
  return QBLAS_OK;
}
int qblas_get_ozaki_crt_bits(void) { return g_oz_bits.load(); }
int qblas_hip_ozaki_last_method(void) { return t_oz_method; }
int qblas_hip_ozaki_last_crt_bits(void) { return t_oz_bits; }
int qblas_hip_ozaki_last_moduli(void) { return t_oz_moduli; }
long long qblas_hip_ozaki_repairs(void *stream) {
  if (!t_oz_hdr.valid()) return -1;
  uint32_t v = 0;
  if (qbk::ozaki_repair_count(t_oz_hdr.p, &v, (hipStream_t)stream) != hipSuccess) return -1;
  return (long long)v;
}
long long qblas_hip_exact_wide_items(void *stream) { return exact_wide_items((hipStream_t)stream); }
int qblas_hip_ozaki_last_slices(void) { return t_oz_slices; }
void qblas_hip_set_phase_timing(int on) { g_phase_timing.store(on ? 1 : 0); }
int qblas_hip_phase_times(float *ms) {
  if (!t_oz_ev_valid) return -1;
  if (hipEventSynchronize(t_oz_ev[4]) != hipSuccess) return -1;
  for (int i = 0; i < 4; ++i)
    if (hipEventElapsedTime(&ms[i], t_oz_ev[i], t_oz_ev[i + 1]) != hipSuccess) return -1;
  return 4;
}
void qblas_set_kc(int kc) { g_kc.store(kc > 0 ? kc : 126); }
int qblas_get_kc(void) { return g_kc.load(); }

const char *quadblas_hip_last_error(void) { return t_err.c_str(); }

int qblas_hip_exact_status(void *stream) {
  const int rc = check_exact_status((hipStream_t)stream);
  return rc == QBLAS_ERR_RANGE ? 1 : (rc == QBLAS_OK ? 0 : -1);
}
void qblas_clear_error(void) { t_err.clear(); }

int qblas_hip_dot(int mode, int64_t n, const void *d_x, int64_t incx, const void *d_y, int64_t incy, int T,
                  void *d_result, void *stream) {
  return dev_dot(mode, n, d_x, incx, d_y, incy, T, d_result, (hipStream_t)stream);
}
int qblas_hip_axpy(int64_t n, const void *alpha, const void *d_x, int64_t incx, void *d_y, int64_t incy,
                   void *stream) {
  return dev_axpy(n, load_q(alpha), d_x, incx, d_y, incy, (hipStream_t)stream);
}
int qblas_hip_nrm2(int mode, int64_t n, const void *d_x, int64_t incx, int T, void *d_result, void *stream) {
  return dev_nrm2(mode, n, d_x, incx, T, d_result, (hipStream_t)stream);
}
int qblas_hip_gemv(int mode, int layout, int64_t m, int64_t n, const void *alpha, const void *d_A, int64_t lda,
                   const void *d_x, int64_t incx, const void *beta, void *d_y, int64_t incy, void *stream) {
  return dev_gemv(mode, layout, m, n, load_q(alpha), d_A, lda, d_x, incx, load_q(beta), d_y, incy,
                  (hipStream_t)stream);
}
int qblas_hip_gemm(int mode, int layout, int64_t m, int64_t n, int64_t k, const void *alpha, const void *d_A,
                   int64_t lda, const void *d_B, int64_t ldb, const void *beta, void *d_C, int64_t ldc,
                   void *stream) {
  return dev_gemm(mode, layout, m, n, k, load_q(alpha), d_A, lda, d_B, ldb, load_q(beta), d_C, ldc,
                  (hipStream_t)stream);
}
int qblas_hip_gemm_batched(int mode, int layout, int64_t count, const qblas_gemm_problem *d_problems, void *stream) {
  return gemm_batched(mode, layout, count, d_problems, tuned_config(mode, layout, count), (hipStream_t)stream);
}
int qblas_hip_gemv_batched(int mode, int layout, int64_t count, const qblas_gemv_problem *d_problems, void *stream) {
  const int exact = mode == QBLAS_MODE_EXACT;
  int rc = check_batched("qblas_hip_gemv_batched", exact ? QBLAS_MODE_REFERENCE : mode, count, d_problems);
  if (rc || count == 0) return rc;
  void *ws;
  const StreamWorkspaces::Entry *we = nullptr;
  const hipStream_t s = (hipStream_t)stream;
  if ((rc = batch_ws(qbk::batched_ws(count), s, &ws, &we))) return rc;
  const int form = exact ? 2 : (mode == QBLAS_MODE_DD ? 1 : 0);
  QB_CHECK(qbk::launch_gemv_batched(form, layout != 0, count, d_problems, ws, s));
  if (exact) set_status((const char *)ws + qbk::batched_status_offset(count), we, false);
  return QBLAS_OK;
}
int qblas_hip_gemm_batched_tune(int mode, int layout, int64_t count, const qblas_gemm_problem *d_problems,
                                void *stream) {
  return gemm_batched_tune(mode, layout, count, d_problems, (hipStream_t)stream);
}
const char *qblas_hip_gemm_batched_config(int mode, int idx) {
  if (mode == QBLAS_MODE_OZAKI) {
    static const std::vector<std::string> names = [] {
      std::vector<std::string> v;
      for (int c = 0; c < qbk::ozaki_batched_configs(); ++c)
        v.push_back(std::to_string(qbk::ozaki_batched_slices(c)) + " slices");
      return v;
    }();
    return (idx >= 0 && idx < (int)names.size()) ? names[idx].c_str() : nullptr;
  }
  if (mode != QBLAS_MODE_REFERENCE && mode != QBLAS_MODE_DD) return nullptr;
  return qbk::gemm_batched_config_name(mode == QBLAS_MODE_DD, idx);
}
int qblas_hip_gemm_batched_select(int mode, int layout, int64_t count, int idx) {
  if ((mode != QBLAS_MODE_REFERENCE && mode != QBLAS_MODE_DD && mode != QBLAS_MODE_OZAKI) || idx < 0 ||
      idx >= batched_configs(mode))
    return QBLAS_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[std::make_tuple(mode, layout, count)] = idx;
  return QBLAS_OK;
}
int qblas_hip_fill(int kind, uint64_t seed, int64_t n, int lo_exp, int hi_exp, void *d_out, void *stream) {
  if (n < 0 || kind < 0 || kind > 4 || (kind == 1 && hi_exp <= lo_exp) || !aligned16(d_out)) {
    set_err_msg("qblas_hip_fill: bad arguments");
    return QBLAS_ERR_ARG;
  }
  if (kind == 1 && (lo_exp < -16382 || hi_exp > 16384)) {
    set_err_msg("qblas_hip_fill: exponent range outside binary128 normals");
    return QBLAS_ERR_ARG;
  }
  QB_CHECK(qbk::launch_fill(kind, seed, n, lo_exp, hi_exp, (q128 *)d_out, (hipStream_t)stream));
  return QBLAS_OK;
}

int qblas_hip_fill_2d(int kind, uint64_t seed, int64_t rows, int64_t cols, int64_t ld, int layout, int lo_exp,
                      int hi_exp, void *d_out, void *stream) {
  if (rows < 0 || cols < 0 || kind < 0 || kind > 4 || (kind == 1 && hi_exp <= lo_exp) || !aligned16(d_out) ||
      ld < (layout ? rows : cols)) {
    set_err_msg("qblas_hip_fill_2d: bad arguments");
    return QBLAS_ERR_ARG;
  }
  if (kind == 1 && (lo_exp < -16382 || hi_exp > 16384)) {
    set_err_msg("qblas_hip_fill_2d: exponent range outside binary128 normals");
    return QBLAS_ERR_ARG;
  }
  QB_CHECK(qbk::launch_fill_2d(kind, seed, rows, cols, ld, layout != 0, lo_exp, hi_exp, (q128 *)d_out,
                               (hipStream_t)stream));
  return QBLAS_OK;
}

}  // extern "C"
