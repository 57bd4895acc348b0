// server.cpp -- serving runtime: continuous batching of RGB-D frame streams.
//
// Each stream keeps its last frame on the device (colour and depth planes,
// boundary maps, and the colour and depth pyramids it got as a pair's frame
// 2), so a new frame is uploaded once and the pair (last frame, new frame)
// reuses frame 1's pyramids.  With warm starts the stream also keeps its last
// flow, and its next pair is solved as sceneFlowWithInit from it.  A worker thread forms batches from pending
// requests -- at most one pair per stream in flight, at most max_batch pairs,
// admitted while the in-flight pixels fit the HBM budget -- and runs them on
// `lanes` executor contexts with their own HIP streams, so a batch formed
// while another is on the GPU runs beside it.  Host copies happen in the
// client's threads, never in the worker's: stream_submit copies the frame to
// a device staging buffer, and a finished pair's flow is kept in a device
// result buffer until stream_poll copies it to the caller's arrays, so the
// worker only launches work and device-to-device copies.  Results do not depend on how
// requests are batched: every pair's solve is the same whatever else is in
// its batch.
//
// Every HIP API call the server makes, in any thread, runs under one mutex
// (hip_mu): the worker holds it while it enqueues a batch -- including the
// capture, instantiation and launch of the batch's HIP graph -- and the
// clients while they allocate, free or enqueue a copy.  Only waits (stream
// synchronisation, event queries) run outside it.  Host frames and results
// pass through pinned host buffers, so a client holds the mutex only for
// the time it takes to enqueue a DMA copy.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rgbdflow/rgbdflow.h"
#include "capi_internal.hpp"
#include "internal.hpp"

using namespace rgbdflow;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes an in-flight pair uses per pixel in an executor context
// (inputs, pyramids, work buffers; see capi.cpp's pair slots).
constexpr double kPairBytesPerPixel = 640.0;

struct Frame {
  double *im = nullptr, *dm = nullptr, *pim = nullptr, *pd = nullptr;
  float* hpb = nullptr;
  float* sigma = nullptr;  // fill_holes: depth stdevs of the depth as submitted
  double* dmh = nullptr;   // fill_holes: the depth with small holes colour-filled (boundary maps)
  double* raw = nullptr;   // fill_holes with warm starts: the depth as submitted (a warm pair's maps)
  bool pyr = false;  // pim / pd hold this frame's pyramids
};

struct StreamState {
  int W = 0, H = 0, min_width = 0;
  rgbdflow_camera cam{};
  size_t N = 0, rest = 0, hpb_n = 0;
  Frame f[2];
  double* flow = nullptr;  // [3][N] the stream's last result (warm starts)
  int *fill_st = nullptr, *fill_todo = nullptr, *fill_left = nullptr;  // fill_holes scratch
  int *fill_label = nullptr, *fill_cmax = nullptr;
  bool has_flow = false;
  int cur = -1;          // index of the last frame, -1 before the first
  bool busy = false;     // a request of this stream is in a batch
  int queued = 0;        // requests of this stream waiting
  bool closing = false;
  void* mem = nullptr;
  size_t bytes = 0;
};

// Device staging of one submitted frame: u8 BGR [N][3], then f32 depth [N],
// then (HPB) the boundary maps, each 256-byte aligned.
inline size_t r256(size_t n) { return (n + 255) & ~size_t(255); }
struct Staged {
  unsigned char* p = nullptr;
  size_t bytes = 0, depth_off = 0, hpb_off = 0, hpb_n = 0;
};

struct Request {
  unsigned long long ticket = 0;
  int stream = -1;
  StreamState* st = nullptr;  // set when picked (the stream stays open while busy)
  Staged in;
  Clock::time_point t_submit;
  int next = 0;  // frame slot the new frame goes to
  bool warm = false;  // solved from the stream's last flow
};

struct Result {
  bool done = false;
  double* flow = nullptr;  // [3][N] device result buffer (has_flow), freed by poll
  size_t bytes = 0;
  rgbdflow_request_info info{};
};

struct Lane {
  rgbdflow_ctx* ctx = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool warm = false;             // the batch's pairs are warm starts
  std::vector<Request> reqs;     // in dispatch order
  std::vector<int> pair_of;      // request -> pair index in the batch, -1 for a first frame
  int status = RGBDFLOW_OK;
  size_t px = 0;
  Clock::time_point t_start;
};

}  // namespace

struct rgbdflow_server {
  int device = 0;
  int max_batch = 16;
  rgbdflow_params params{};
  bool warm = false;
  bool fill = false;              // hole-fill each frame's depth on the device
  rgbdflow_params warm_params{};  // params of warm-started pairs
  unsigned long long budget = 0;
  unsigned long long state_bytes = 0;
  size_t inflight_px = 0;
  std::vector<Lane> lanes;
  std::map<int, std::unique_ptr<StreamState>> streams;
  std::deque<Request> pending;
  std::map<unsigned long long, Result> results;
  unsigned long long next_ticket = 1;
  int next_stream = 0;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  std::string err;
  hipStream_t copy_stream = nullptr;       // host<->device copies of submit and poll
  std::multimap<size_t, void*> free_bufs;  // device staging and result buffers by size (lock m)
  std::multimap<size_t, void*> free_pinned;  // pinned host buffers by size (lock m)
  std::mutex hip_mu;                       // every HIP API call except waits (see the file comment)
};

namespace {

int sfail(rgbdflow_server* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

// A device buffer of exactly `bytes` from the server's free list, else a new
// one (lock not held).
void* buf_get(rgbdflow_server* S, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->free_bufs.find(bytes);
    if (it != S->free_bufs.end()) {
      void* p = it->second;
      S->free_bufs.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  std::lock_guard<std::mutex> hk(S->hip_mu);
  hipSetDevice(S->device);
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

// A pinned host buffer of exactly `bytes` from the server's pool, else a new
// one (no lock held).
void* pinned_get(rgbdflow_server* S, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->free_pinned.find(bytes);
    if (it != S->free_pinned.end()) {
      void* p = it->second;
      S->free_pinned.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  std::lock_guard<std::mutex> hk(S->hip_mu);
  hipSetDevice(S->device);
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void pinned_put(rgbdflow_server* S, void* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(S->m);
  S->free_pinned.emplace(bytes, p);
}

void buf_put(rgbdflow_server* S, void* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(S->m);
  S->free_bufs.emplace(bytes, p);
}

// buf_put with the lock already held
void buf_put_locked(rgbdflow_server* S, void* p, size_t bytes) {
  if (p) S->free_bufs.emplace(bytes, p);
}

size_t budget_px(const rgbdflow_server* s) {
  if (!s->budget) return (size_t)-1;
  if (s->state_bytes >= s->budget) return 0;
  return (size_t)((double)(s->budget - s->state_bytes) / kPairBytesPerPixel);
}

// Uploads and converts the new frames, loads the pairs and enqueues the batch
// on the lane's stream (worker thread, no lock held; the streams involved are
// busy, so nothing else touches their state).
void dispatch(rgbdflow_server* S, Lane& ln) {
  hipStream_t hs = ctx_stream(ln.ctx);
  std::vector<PairDev> pairs;
  ln.pair_of.assign(ln.reqs.size(), -1);
  ln.status = RGBDFLOW_OK;
  for (size_t r = 0; r < ln.reqs.size(); r++) {
    Request& q = ln.reqs[r];
    StreamState& st = *q.st;
    Frame& F = st.f[q.next];
    const float* d32 = reinterpret_cast<const float*>(q.in.p + q.in.depth_off);
    launch_ingest(hs, q.in.p, d32, (int)st.N, F.im, F.dm);
    if (F.raw) hipMemcpyAsync(F.raw, F.dm, st.N * 8, hipMemcpyDeviceToDevice, hs);
    if (S->fill) {
      // the depth stdevs of the depth as submitted (runFlowTwoFrames.cpp:89-91);
      // small holes colour-filled for the boundary maps (sceneFlowMultiscale.cpp:78-83);
      // every hole filled for the rest (sceneFlowMultiscale.cpp:54,61)
      launch_depth_sigma(hs, d32, st.W, st.H, 4, F.sigma);
      hipMemcpyAsync(F.dmh, F.dm, st.N * 8, hipMemcpyDeviceToDevice, hs);
      launch_colour_fill(hs, F.dmh, F.im, st.W, st.H, 60, st.fill_label, st.fill_st, st.fill_todo, st.fill_cmax,
                         st.fill_left);
      launch_median_fill(hs, F.dm, st.W, st.H, st.fill_st, st.fill_todo, st.fill_left);
    }
    if (q.in.hpb_n) hipMemcpyAsync(F.hpb, q.in.p + q.in.hpb_off, q.in.hpb_n * 4, hipMemcpyDeviceToDevice, hs);
    F.pyr = false;
    if (st.cur >= 0) {
      const Frame& P = st.f[st.cur];
      PairDev pd{};
      pd.W = st.W;
      pd.H = st.H;
      pd.min_width = st.min_width;
      pd.cam = st.cam;
      // the boundary maps of frame 1 (made on the device unless submitted) start from
      // the colour-filled depth for a coarse-to-fine pair (sceneFlowMultiscale.cpp:78-83)
      // and from the depth as submitted for a warm start (sceneFlowWithInit.cpp:47)
      const double* hpb_dm = S->fill ? (q.warm ? P.raw : P.dmh) : nullptr;
      pd.f1 = FrameDev{P.im, P.dm, P.pyr ? P.pim : nullptr, P.pyr ? P.pd : nullptr, q.in.hpb_n ? P.hpb : nullptr,
                       S->fill ? P.sigma : nullptr, hpb_dm};
      pd.f2 = FrameDev{F.im, F.dm, nullptr, nullptr, nullptr, nullptr, nullptr};
      pd.init = q.warm ? st.flow : nullptr;
      ln.pair_of[r] = (int)pairs.size();
      pairs.push_back(pd);
    }
  }
  if (!pairs.empty()) {
    const rgbdflow_params* prm = ln.warm ? &S->warm_params : &S->params;
    int rc = batch_load_device(ln.ctx, prm, (int)pairs.size(), pairs.data());
    if (rc == RGBDFLOW_OK) rc = rgbdflow_batch_run(ln.ctx, prm, nullptr);
    ln.status = rc;
  }
  hipEventRecord(ln.done, hs);
  ln.t_start = Clock::now();
}

// Collects a finished lane: keeps each stream's new frame with its pyramids,
// downloads the flows, and publishes the results.
void finish(rgbdflow_server* S, Lane& ln) {
  hipStream_t hs = ctx_stream(ln.ctx);
  const Clock::time_point t_end = Clock::now();
  std::vector<Result> out(ln.reqs.size());
  for (size_t r = 0; r < ln.reqs.size(); r++) {
    const Request& q = ln.reqs[r];
    StreamState& st = *q.st;
    Result& res = out[r];
    res.info.stream = q.stream;
    res.info.status = ln.status;
    res.info.batch_pairs = 0;
    if (ln.pair_of[r] < 0 || ln.status != RGBDFLOW_OK) continue;
    FrameDev f2;
    const double* flow;
    size_t rest;
    if (pair_outputs(ln.ctx, ln.pair_of[r], &f2, &flow, &rest) != RGBDFLOW_OK) {
      res.info.status = RGBDFLOW_ERR_STATE;
      continue;
    }
    Frame& F = st.f[q.next];
    F.pyr = !q.warm;  // a warm start builds no pyramids
    res.bytes = 3 * st.N * 8;
    res.flow = static_cast<double*>(buf_get(S, res.bytes));
    if (!res.flow) {
      res.info.status = RGBDFLOW_ERR_HIP;
      continue;
    }
    std::lock_guard<std::mutex> hk(S->hip_mu);
    if (rest > 0 && !q.warm) {
      hipMemcpyAsync(F.pim, f2.pim, 3 * rest * 8, hipMemcpyDeviceToDevice, hs);
      hipMemcpyAsync(F.pd, f2.pd, rest * 8, hipMemcpyDeviceToDevice, hs);
    }
    if (S->warm) hipMemcpyAsync(st.flow, flow, 3 * st.N * 8, hipMemcpyDeviceToDevice, hs);
    hipMemcpyAsync(res.flow, flow, res.bytes, hipMemcpyDeviceToDevice, hs);
    res.info.has_flow = 1;
  }
  const hipError_t e = hipStreamSynchronize(hs);
  const Clock::time_point t_done = Clock::now();
  int n_pairs = 0;
  for (int p : ln.pair_of) n_pairs += p >= 0;
  std::lock_guard<std::mutex> lk(S->m);
  for (size_t r = 0; r < ln.reqs.size(); r++) {
    const Request& q = ln.reqs[r];
    StreamState& st = *q.st;
    st.cur = q.next;
    st.busy = false;
    if (S->warm && out[r].info.has_flow && e == hipSuccess) st.has_flow = true;
    Result& res = S->results[q.ticket];
    res.flow = out[r].flow;
    res.bytes = out[r].bytes;
    res.info = out[r].info;
    buf_put_locked(S, q.in.p, q.in.bytes);  // the batch has read the staged frame
    if (e != hipSuccess) res.info.status = RGBDFLOW_ERR_HIP;
    res.info.batch_pairs = n_pairs;
    res.info.queue_ms = ms_between(q.t_submit, ln.t_start);
    res.info.batch_ms = ms_between(ln.t_start, t_end);
    res.info.latency_ms = ms_between(q.t_submit, t_done);
    res.done = true;
  }
  S->inflight_px -= ln.px;
  ln.busy = false;
  ln.reqs.clear();
  S->cv_done.notify_all();
}

// Picks the requests of the next batch (lock held): in submission order, one
// per stream, while the pixels fit the budget; warm starts and coarse-to-fine
// pairs go in separate batches (their params differ), the first pending pair
// deciding which.
void pick(rgbdflow_server* S, Lane& ln) {
  const size_t cap = budget_px(S);
  size_t px = 0;
  int n_pairs = 0, kind = -1;
  for (auto it = S->pending.begin(); it != S->pending.end() && n_pairs < S->max_batch;) {
    StreamState& st = *S->streams[it->stream];
    if (st.busy) {
      ++it;
      continue;
    }
    const bool pair = st.cur >= 0;
    const int warm = (S->warm && st.has_flow) ? 1 : 0;
    if (pair && kind >= 0 && warm != kind) {
      ++it;
      continue;
    }
    if (pair) kind = warm;
    it->warm = warm != 0;
    if (pair && S->inflight_px + px + st.N > cap && (n_pairs > 0 || S->inflight_px > 0)) break;
    it->next = st.cur < 0 ? 0 : 1 - st.cur;
    it->st = &st;
    st.busy = true;
    st.queued--;
    if (pair) {
      px += st.N;
      n_pairs++;
    }
    ln.reqs.push_back(std::move(*it));
    it = S->pending.erase(it);
  }
  ln.px = px;
  ln.warm = kind == 1;
  S->inflight_px += px;
}

void worker_main(rgbdflow_server* S) {
  hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->m);
  for (;;) {
    bool idle = true, progressed = false;
    for (Lane& ln : S->lanes) {
      if (!ln.busy) continue;
      idle = false;
      hipError_t q;
      {
        std::lock_guard<std::mutex> hk(S->hip_mu);
        q = hipEventQuery(ln.done);
      }
      if (q == hipSuccess) {
        lk.unlock();
        finish(S, ln);
        lk.lock();
        progressed = true;
      }
    }
    for (Lane& ln : S->lanes) {
      if (ln.busy || S->pending.empty()) continue;
      pick(S, ln);
      if (ln.reqs.empty()) continue;
      ln.busy = true;
      idle = false;
      lk.unlock();
      {
        std::lock_guard<std::mutex> hk(S->hip_mu);
        dispatch(S, ln);
      }
      lk.lock();
      progressed = true;
    }
    if (S->stop && idle && S->pending.empty()) break;
    if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(100));
  }
}

void free_stream(StreamState& st) {
  if (st.mem) hipFree(st.mem);
  st.mem = nullptr;
}

}  // namespace

extern "C" {

int rgbdflow_server_create(rgbdflow_server** out, const rgbdflow_server_config* cfg, const rgbdflow_params* params) {
  if (!out || !cfg || !params) return RGBDFLOW_ERR_ARG;
  *out = nullptr;
  if (params->variant != RGBDFLOW_VARIANT_FULL) return RGBDFLOW_ERR_UNSUPPORTED;
  std::unique_ptr<rgbdflow_server> S(new rgbdflow_server());
  S->device = cfg->device;
  S->max_batch = cfg->max_batch > 0 ? cfg->max_batch : 16;
  S->budget = cfg->hbm_budget;
  S->params = *params;
  S->warm = cfg->warm_start != 0;
  S->fill = cfg->fill_holes != 0;
  S->warm_params = *params;
  if (cfg->warm_n_outer > 0) S->warm_params.n_outer = cfg->warm_n_outer;
  const int n_lanes = cfg->lanes > 0 ? cfg->lanes : 2;
  S->lanes.resize(n_lanes);
  auto undo = [&](int rc) {
    for (Lane& ln : S->lanes) {
      if (ln.done) hipEventDestroy(ln.done);
      if (ln.ctx) rgbdflow_destroy(ln.ctx);
    }
    if (S->copy_stream) hipStreamDestroy(S->copy_stream);
    return rc;
  };
  for (Lane& ln : S->lanes) {
    const int rc = rgbdflow_create(&ln.ctx, cfg->device, 0);
    if (rc != RGBDFLOW_OK) return undo(rc);
    if (hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) != hipSuccess) return undo(RGBDFLOW_ERR_HIP);
  }
  hipSetDevice(cfg->device);
  if (hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess) return undo(RGBDFLOW_ERR_HIP);
  rgbdflow_server* raw = S.release();
  raw->worker = std::thread(worker_main, raw);
  *out = raw;
  return RGBDFLOW_OK;
}

int rgbdflow_server_destroy(rgbdflow_server* S) {
  if (!S) return RGBDFLOW_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(S->m);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  hipSetDevice(S->device);
  for (Lane& ln : S->lanes) {
    if (ln.done) hipEventDestroy(ln.done);
    if (ln.ctx) rgbdflow_destroy(ln.ctx);
  }
  for (auto& kv : S->streams) free_stream(*kv.second);
  for (auto& kv : S->results)
    if (kv.second.flow) hipFree(kv.second.flow);
  for (const Request& q : S->pending)
    if (q.in.p) hipFree(q.in.p);
  for (auto& kv : S->free_bufs) hipFree(kv.second);
  for (auto& kv : S->free_pinned) hipHostFree(kv.second);
  if (S->copy_stream) hipStreamDestroy(S->copy_stream);
  delete S;
  return RGBDFLOW_OK;
}

const char* rgbdflow_server_last_error(const rgbdflow_server* S) { return S ? S->err.c_str() : "null server"; }

int rgbdflow_stream_open(rgbdflow_server* S, int W, int H, const rgbdflow_camera* cam, int min_width, int* stream_id) {
  if (!S) return RGBDFLOW_ERR_ARG;
  if (W < 2 || H < 2 || (long long)W * H > (1LL << 28) || !stream_id)
    return sfail(S, RGBDFLOW_ERR_ARG, "bad stream arguments");
  std::unique_ptr<StreamState> st(new StreamState());
  st->W = W;
  st->H = H;
  st->min_width = min_width > 0 ? min_width : S->params.min_width;
  st->cam = cam ? *cam : S->params.cam;
  if ((int)st->cam.xRes != W || (int)st->cam.yRes != H)
    return sfail(S, RGBDFLOW_ERR_ARG, "camera resolution differs from the stream size");
  st->N = (size_t)W * H;
  st->rest = pyramid_rest_elems(W, H, S->params.pyramid_ratio, st->min_width);
  st->hpb_n = pyramid_all_elems(W, H, S->params.pyramid_ratio, st->min_width) * 2;
  // one allocation: two frames (planes, pyramids, boundary maps) and the last flow
  const size_t frame_bytes = r256(3 * st->N * 8) + r256(st->N * 8) + r256(3 * st->rest * 8 + 8) +
                             r256(st->rest * 8 + 8) + r256(st->hpb_n * 4);
  const size_t fill_bytes =
      S->fill ? 2 * (r256(st->N * 4) + r256(st->N * 8) + (S->warm ? r256(st->N * 8) : 0)) + 3 * r256(st->N * 4) +
                    r256(2 * st->N * 4) + r256(4)
              : 0;
  const size_t bytes = 2 * frame_bytes + (S->warm ? r256(3 * st->N * 8) : 0) + fill_bytes;
  {
    std::lock_guard<std::mutex> lk(S->m);
    if (S->budget && S->state_bytes + bytes > S->budget)
      return sfail(S, RGBDFLOW_ERR_BUDGET, "stream state exceeds the HBM budget");
  }
  {
    std::lock_guard<std::mutex> hk(S->hip_mu);
    hipSetDevice(S->device);
    if (hipMalloc(&st->mem, bytes) != hipSuccess) return sfail(S, RGBDFLOW_ERR_HIP, "hipMalloc of the stream state");
  }
  st->bytes = bytes;
  char* p = static_cast<char*>(st->mem);
  auto take = [&](size_t n) {
    char* q = p;
    p += r256(n);
    return q;
  };
  for (Frame& F : st->f) {
    F.im = reinterpret_cast<double*>(take(3 * st->N * 8));
    F.dm = reinterpret_cast<double*>(take(st->N * 8));
    F.pim = reinterpret_cast<double*>(take(3 * st->rest * 8 + 8));
    F.pd = reinterpret_cast<double*>(take(st->rest * 8 + 8));
    F.hpb = reinterpret_cast<float*>(take(st->hpb_n * 4));
  }
  if (S->warm) st->flow = reinterpret_cast<double*>(take(3 * st->N * 8));
  if (S->fill) {
    for (Frame& F : st->f) {
      F.sigma = reinterpret_cast<float*>(take(st->N * 4));
      F.dmh = reinterpret_cast<double*>(take(st->N * 8));
      if (S->warm) F.raw = reinterpret_cast<double*>(take(st->N * 8));
    }
    st->fill_st = reinterpret_cast<int*>(take(st->N * 4));
    st->fill_label = reinterpret_cast<int*>(take(st->N * 4));
    st->fill_cmax = reinterpret_cast<int*>(take(st->N * 4));
    st->fill_todo = reinterpret_cast<int*>(take(2 * st->N * 4));
    st->fill_left = reinterpret_cast<int*>(take(4));
  }
  std::lock_guard<std::mutex> lk(S->m);
  S->state_bytes += st->bytes;
  *stream_id = S->next_stream++;
  S->streams[*stream_id] = std::move(st);
  return RGBDFLOW_OK;
}

int rgbdflow_stream_submit(rgbdflow_server* S, int stream_id, const unsigned char* bgr, const float* depth,
                           const float* hpb, unsigned long long* ticket) {
  if (!S) return RGBDFLOW_ERR_ARG;
  if (!bgr || !depth || !ticket) return sfail(S, RGBDFLOW_ERR_ARG, "null frame");
  Request q;
  size_t N, hn;
  {
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->streams.find(stream_id);
    if (it == S->streams.end() || it->second->closing) return sfail(S, RGBDFLOW_ERR_ARG, "no such stream");
    N = it->second->N;
    hn = it->second->hpb_n;
  }
  // the frame goes to a device staging buffer here, in the caller's thread
  Staged& in = q.in;
  in.depth_off = r256(3 * N);
  in.hpb_off = in.depth_off + r256(N * 4);
  in.hpb_n = hpb ? hn : 0;
  in.bytes = in.hpb_off + r256(in.hpb_n * 4);
  in.p = static_cast<unsigned char*>(buf_get(S, in.bytes));
  if (!in.p) return sfail(S, RGBDFLOW_ERR_HIP, "hipMalloc of a frame staging buffer");
  // the frame into a pinned buffer (host copy, no lock), then one DMA copy
  unsigned char* h = static_cast<unsigned char*>(pinned_get(S, in.bytes));
  if (!h) {
    buf_put(S, in.p, in.bytes);
    return sfail(S, RGBDFLOW_ERR_HIP, "hipHostMalloc of a frame staging buffer");
  }
  std::memcpy(h, bgr, 3 * N);
  std::memcpy(h + in.depth_off, depth, N * 4);
  if (hpb) std::memcpy(h + in.hpb_off, hpb, hn * 4);
  hipError_t e;
  {
    std::lock_guard<std::mutex> hk(S->hip_mu);
    hipSetDevice(S->device);
    e = hipMemcpyAsync(in.p, h, in.bytes, hipMemcpyHostToDevice, S->copy_stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
  pinned_put(S, h, in.bytes);
  if (e != hipSuccess) {
    buf_put(S, in.p, in.bytes);
    return sfail(S, RGBDFLOW_ERR_HIP, std::string("frame upload: ") + hipGetErrorString(e));
  }
  q.stream = stream_id;
  q.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->m);
    q.ticket = S->next_ticket++;
    *ticket = q.ticket;
    S->results[q.ticket] = Result();
    S->streams[stream_id]->queued++;
    S->pending.push_back(std::move(q));
  }
  S->cv_work.notify_all();
  return RGBDFLOW_OK;
}

int rgbdflow_stream_poll(rgbdflow_server* S, unsigned long long ticket, int wait_ms, double* u, double* v, double* w,
                         rgbdflow_request_info* info) {
  if (!S) return RGBDFLOW_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->m);
  auto it = S->results.find(ticket);
  if (it == S->results.end()) return sfail(S, RGBDFLOW_ERR_ARG, "unknown ticket");
  if (!it->second.done && wait_ms != 0) {
    auto ready = [&] { return S->results[ticket].done; };
    if (wait_ms < 0) S->cv_done.wait(lk, ready);
    else S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    it = S->results.find(ticket);
  }
  if (!it->second.done) return 0;
  const Result r = it->second;
  S->results.erase(it);
  lk.unlock();
  if (info) *info = r.info;
  int status = r.info.status;
  if (r.flow && status == RGBDFLOW_OK && r.info.has_flow) {
    // the result leaves the device here, in the caller's thread
    const size_t n = r.bytes / 24;
    double* h = static_cast<double*>(pinned_get(S, r.bytes));
    hipError_t e = h ? hipSuccess : hipErrorOutOfMemory;
    if (h) {
      {
        std::lock_guard<std::mutex> hk(S->hip_mu);
        hipSetDevice(S->device);
        e = hipMemcpyAsync(h, r.flow, r.bytes, hipMemcpyDeviceToHost, S->copy_stream);
      }
      if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
      double* out[3] = {u, v, w};
      for (int c = 0; c < 3 && e == hipSuccess; c++)
        if (out[c]) std::memcpy(out[c], h + c * n, n * 8);
      pinned_put(S, h, r.bytes);
    }
    if (e != hipSuccess) status = RGBDFLOW_ERR_HIP;
  }
  buf_put(S, r.flow, r.bytes);
  if (status != RGBDFLOW_OK) {
    std::lock_guard<std::mutex> lk2(S->m);
    return sfail(S, status, "the batch of this request failed");
  }
  return 1;
}

int rgbdflow_server_set_relax_schedule(rgbdflow_server* S, int W, int H, const rgbdflow_tile_choice* choice) {
  if (!S) return RGBDFLOW_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->m);
  // the lanes' contexts are touched only by the worker, and only while it has
  // work: wait until it has none (it cannot get any while the lock is held)
  S->cv_done.wait(lk, [&] {
    if (!S->pending.empty()) return false;
    for (const Lane& ln : S->lanes)
      if (ln.busy) return false;
    return true;
  });
  std::lock_guard<std::mutex> hk(S->hip_mu);
  for (Lane& ln : S->lanes) {
    const int rc = rgbdflow_set_relax_schedule(ln.ctx, W, H, choice);
    if (rc != RGBDFLOW_OK) return sfail(S, rc, rgbdflow_last_error(ln.ctx));
  }
  return RGBDFLOW_OK;
}

int rgbdflow_stream_close(rgbdflow_server* S, int stream_id) {
  if (!S) return RGBDFLOW_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->m);
  auto it = S->streams.find(stream_id);
  if (it == S->streams.end()) return sfail(S, RGBDFLOW_ERR_ARG, "no such stream");
  StreamState* st = it->second.get();
  st->closing = true;
  S->cv_done.wait(lk, [&] { return !st->busy && st->queued == 0; });
  S->state_bytes -= st->bytes;
  {
    std::lock_guard<std::mutex> hk(S->hip_mu);
    free_stream(*st);
  }
  S->streams.erase(it);
  return RGBDFLOW_OK;
}

}  // extern "C"
