// wfa_server.cpp — serving loop: continuous batching of alignment requests.
//
// The reference aligns one pair per call on one thread (src/aligner.rs:620,
// :707); a service built on it runs one aligner per request. Here requests
// go to a queue instead and the device runs them in shared launches:
//
// - Every request carries its own configuration (penalties, span,
//   heuristic, memory mode, scope) and sequences, copied at submission.
// - Requests wait in per-class FIFO queues. A class is one configuration,
//   one sequence format and one length bucket: a launch is planned for its
//   longest pair, so short reads never share a launch with long ones.
//   Requests that differ only in span (ends-free or extension, any free
//   ends) or only in heuristic share a class: the launch takes each pair's
//   span and heuristic from a per-pair configuration, within the class's
//   envelope (end-to-end or ends-free; no heuristic or some heuristic).
// - Long-read launches (longer than 1024 bases) may take at most half the
//   engines and half a workgroup per CU each, so short requests always find
//   an engine and room on the CUs.
// - A few engines, each with its own HIP stream, scratch and worker thread,
//   pull work. An idle worker takes the class whose oldest request has
//   waited longest among those that fit, admits that class's requests in
//   arrival order while the predicted wavefront memory of all work on the
//   device stays within the budget (the first request always fits on an
//   idle device), and runs them as one persistent launch.
// - Requests that arrive while launches run are admitted as soon as an
//   engine and memory free up; launches of different engines overlap on
//   the device, and batches grow with the load.
//
// A request's result does not depend on what it shares a launch with: every
// pair of a launch is aligned independently by the same code.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "wfa_engine.h"

using namespace wfahost;

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t tag;
  int32_t plen, tlen;
  size_t pbytes, tbytes;
  std::vector<uint8_t> seq;  // pattern bytes, then text bytes
  int64_t footprint;         // predicted device bytes while it runs
  wfa_pair_config_t cfg;     // this request's span and heuristic
  Clock::time_point t0;
};

struct Cls {
  wfa_attr_t attr;  // the envelope: per-pair spans and heuristics within it
  int32_t fmt;
  int bucket;
  bool per_pair;    // launches carry each request's span and heuristic
  std::deque<Req> q;
};

// buckets above this one hold reads longer than 1024 bases
constexpr int kLongBucket = 1;

struct Out {
  uint64_t tag;
  int32_t error, score, status;
  std::vector<uint8_t> cigar;
  double latency_us;
};

struct Worker {
  wfa_engine_t* e = nullptr;
  hipStream_t st = nullptr;
  uint8_t* h = nullptr;  // pinned staging of inputs and outputs
  size_t h_cap = 0;
  uint8_t* d = nullptr;  // device copy of the same layout
  size_t d_cap = 0;
  std::thread th;
};

// Length buckets: <= 256 bases, then one bucket per factor of four.
int bucket_of(int len) {
  int b = 0;
  for (int64_t lim = 256; len > lim && b < 12; lim *= 4) ++b;
  return b;
}

// Predicted device bytes of one request while its launch runs. The score
// of a pair of length L with up to 10% divergence is taken as
// S = 0.1 * L * max(mismatch, gap open + extension); a wavefront widens by
// at most two diagonals per gap-extension cost, so W <= 2 * S / e + 1,
// bounded by plen + tlen + 1. High-memory CIGARs keep every M row (4 bytes
// a cell, about S * W / 2 cells), medium/low memory one flag byte a cell,
// score-only and BiWFA only the live score window of their rows (O(S)).
int64_t predict_footprint(const wfa_attr_t& a, const Internal& ip, int plen, int tlen) {
  const int64_t len = std::max(plen, tlen);
  int cost = 1, e = 1;
  if (ip.metric == WFA_METRIC_GAP_AFFINE || ip.metric == WFA_METRIC_GAP_AFFINE_2P) {
    cost = std::max(ip.x, ip.o1 + ip.e1);
    e = ip.metric == WFA_METRIC_GAP_AFFINE_2P ? std::min(ip.e1, ip.e2) : ip.e1;
  } else if (ip.metric == WFA_METRIC_GAP_LINEAR) {
    cost = std::max(ip.x, ip.o1);
    e = ip.o1;
  }
  const int64_t s = std::max<int64_t>(1, (len * cost + 9) / 10);
  const int64_t w = std::min<int64_t>(2 * s / std::max(e, 1) + 1, (int64_t)plen + tlen + 1);
  const int64_t rows = score_scope(ip) * 5 + 2;
  const int64_t live = rows * w * 4;
  const int64_t seq = (int64_t)plen + tlen;
  if (a.alignment_scope == WFA_SCOPE_SCORE || a.memory_mode == WFA_MEMORY_ULTRALOW)
    return live + 2 * seq;
  const int64_t cells = s * w / 2 + 1;
  const int64_t hist = a.memory_mode == WFA_MEMORY_HIGH ? 4 * cells : cells;
  return live + hist + 3 * seq;
}

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

bool acgt_only(const uint8_t* s, int n) {
  for (int i = 0; i < n; ++i)
    if (s[i] != 'A' && s[i] != 'C' && s[i] != 'G' && s[i] != 'T') return false;
  return true;
}

// 2-bit packing, 4 bases per byte, LSB first, A=0 C=1 G=2 T=3
// (src/sequences.rs:55-74)
void pack2(const uint8_t* s, int n, uint8_t* out) {
  for (int i = 0; i < n; i += 4) {
    uint8_t b = 0;
    for (int j = 0; j < 4 && i + j < n; ++j) {
      const uint8_t c = s[i + j];
      const int v = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3;
      b |= (uint8_t)(v << (2 * j));
    }
    out[i / 4] = b;
  }
}

}  // namespace

struct wfa_server_t {
  wfa_server_config_t cfg;
  std::vector<std::unique_ptr<Worker>> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<Cls> classes;
  std::deque<Out> done;
  std::vector<Out> polled;  // results of the last poll (their CIGARs stay valid)
  int64_t in_flight = 0;    // predicted device bytes of running launches
  int long_running = 0;     // long-read launches on the device
  int64_t queued = 0, running = 0, launches = 0, launched_pairs = 0;
  bool stop = false;
};

namespace {

// Long-read requests may fill the budget only up to 7/8: the rest stays
// for short requests, whatever the long ones queue.
int64_t limit_of(const wfa_server_t* s, const Cls& c) {
  return c.bucket > kLongBucket ? s->cfg.memory_budget - s->cfg.memory_budget / 8
                                : s->cfg.memory_budget;
}

// Pick the next launch under s->mu: the fitting class whose head waited
// longest, then its requests in order while they fit.
bool pick(wfa_server_t* s, wfa_attr_t* attr_out, int* fmt_out, bool* per_pair_out,
          bool* long_out, std::vector<Req>* batch, int64_t* fp_out) {
  int best = -1;
  for (int i = 0; i < (int)s->classes.size(); ++i) {
    const Cls& c = s->classes[i];
    if (c.q.empty()) continue;
    if (c.bucket > kLongBucket && 2 * (s->long_running + 1) > std::max(2, s->cfg.num_engines))
      continue;
    const bool fits = s->in_flight == 0 || s->in_flight + c.q.front().footprint <= limit_of(s, c);
    if (!fits) continue;
    if (best < 0 || c.q.front().t0 < s->classes[best].q.front().t0) best = i;
  }
  if (best < 0) return false;
  Cls& c = s->classes[best];
  const int64_t limit = limit_of(s, c);
  int64_t fp = 0;
  batch->clear();
  while (!c.q.empty() && (int)batch->size() < s->cfg.max_batch_pairs) {
    const int64_t f = c.q.front().footprint;
    if (!batch->empty() && s->in_flight + fp + f > limit) break;
    if (batch->empty() && s->in_flight > 0 && s->in_flight + f > limit) break;
    fp += f;
    batch->push_back(std::move(c.q.front()));
    c.q.pop_front();
  }
  if (batch->empty()) return false;
  *attr_out = c.attr;
  *fmt_out = c.fmt;
  *per_pair_out = c.per_pair;
  *long_out = c.bucket > kLongBucket;
  *fp_out = fp;
  // classes come and go with the traffic (ends-free requests carry their
  // own free ends): forget emptied ones
  if (c.q.empty()) s->classes.erase(s->classes.begin() + best);
  return true;
}

// staging each worker holds from the start (host pinned and device)
constexpr size_t kStagingBytes = size_t(32) << 20;

int grow_pinned(uint8_t** p, size_t* cap, size_t need) {
  if (need <= *cap) return WFA_OK;
  const size_t n = std::max(need, *cap * 2);
  if (*p) hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipHostMalloc((void**)p, n, hipHostMallocDefault) != hipSuccess) return WFA_ERR_OUT_OF_MEMORY;
  *cap = n;
  return WFA_OK;
}

int grow_device(uint8_t** p, size_t* cap, size_t need) {
  return grow((void**)p, cap, need);
}

// One launch: stage the batch, align it on the worker's stream, read back.
// Layout (host and device alike): sequences | pattern offsets | text
// offsets | CIGAR offsets (int64) | pattern lengths | text lengths (int32)
// || scores | statuses | CIGAR lengths (int32) | CIGAR ops; the part after
// || is output.
int run_batch(Worker& w, const wfa_attr_t& attr, int fmt, bool per_pair, std::vector<Req>& reqs,
              std::vector<Out>* outs) {
  const size_t n = reqs.size();
  const bool cig = attr.alignment_scope == WFA_SCOPE_ALIGNMENT;
  size_t seq_bytes = 0, cig_bytes = 0;
  int max_len = 0;
  for (const Req& r : reqs) {
    seq_bytes += r.pbytes + r.tbytes;
    if (cig) cig_bytes += (size_t)r.plen + r.tlen;
    max_len = std::max(max_len, std::max(r.plen, r.tlen));
  }
  const size_t o_poff = align256(seq_bytes);
  const size_t o_toff = o_poff + align256(n * 8);
  const size_t o_coff = o_toff + align256(n * 8);
  const size_t o_plen = o_coff + align256(n * 8);
  const size_t o_tlen = o_plen + align256(n * 4);
  const size_t o_cfg = o_tlen + align256(n * 4);
  const size_t o_score =
      o_cfg + (per_pair ? align256(n * sizeof(wfa_pair_config_t)) : 0);  // outputs start here
  const size_t o_status = o_score + align256(n * 4);
  const size_t o_clen = o_status + align256(n * 4);
  const size_t o_cig = o_clen + align256(n * 4);
  const size_t total = o_cig + align256(std::max<size_t>(cig_bytes, 1));
  int r = grow_pinned(&w.h, &w.h_cap, total);
  if (r) return r;
  r = grow_device(&w.d, &w.d_cap, total);
  if (r) return r;
  int64_t* poff = (int64_t*)(w.h + o_poff);
  int64_t* toff = (int64_t*)(w.h + o_toff);
  int64_t* coff = (int64_t*)(w.h + o_coff);
  int32_t* plen = (int32_t*)(w.h + o_plen);
  int32_t* tlen = (int32_t*)(w.h + o_tlen);
  wfa_pair_config_t* cfg = (wfa_pair_config_t*)(w.h + o_cfg);
  size_t pos = 0, cpos = 0;
  for (size_t i = 0; i < n; ++i) {
    const Req& q = reqs[i];
    if (!q.seq.empty()) std::memcpy(w.h + pos, q.seq.data(), q.seq.size());
    poff[i] = (int64_t)pos;
    toff[i] = (int64_t)(pos + q.pbytes);
    pos += q.pbytes + q.tbytes;
    coff[i] = (int64_t)cpos;
    if (cig) cpos += (size_t)q.plen + q.tlen;
    plen[i] = q.plen;
    tlen[i] = q.tlen;
    if (per_pair) cfg[i] = q.cfg;
  }
  HIPCHK(hipMemcpyAsync(w.d, w.h, o_score, hipMemcpyHostToDevice, w.st));
  wfa_batch_t b;
  std::memset(&b, 0, sizeof(b));
  b.num_pairs = (int32_t)n;
  b.seq_format = fmt;
  b.max_length = max_len;
  b.seq_data = w.d;
  b.pattern_offset = (const int64_t*)(w.d + o_poff);
  b.text_offset = (const int64_t*)(w.d + o_toff);
  b.pattern_length = (const int32_t*)(w.d + o_plen);
  b.text_length = (const int32_t*)(w.d + o_tlen);
  b.score = (int32_t*)(w.d + o_score);
  b.status = (int32_t*)(w.d + o_status);
  if (per_pair) b.pair_config = (const wfa_pair_config_t*)(w.d + o_cfg);
  if (cig) {
    b.cigar_ops = w.d + o_cig;
    b.cigar_offset = (const int64_t*)(w.d + o_coff);
    b.cigar_length = (int32_t*)(w.d + o_clen);
  }
  r = wfa_engine_align_batch(w.e, &attr, &b, w.st);
  if (r) return r;
  const size_t out_end = cig ? o_cig + cig_bytes : o_clen;
  HIPCHK(hipMemcpyAsync(w.h + o_score, w.d + o_score, out_end - o_score, hipMemcpyDeviceToHost,
                        w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  const int32_t* score = (const int32_t*)(w.h + o_score);
  const int32_t* status = (const int32_t*)(w.h + o_status);
  const int32_t* clen = (const int32_t*)(w.h + o_clen);
  const Clock::time_point t1 = Clock::now();
  outs->resize(n);
  for (size_t i = 0; i < n; ++i) {
    Out& o = (*outs)[i];
    o.tag = reqs[i].tag;
    o.error = WFA_OK;
    o.score = score[i];
    o.status = status[i];
    o.cigar.clear();
    if (cig && clen[i] > 0) {
      const uint8_t* src = w.h + o_cig + coff[i];
      o.cigar.assign(src, src + clen[i]);
    }
    o.latency_us = std::chrono::duration<double, std::micro>(t1 - reqs[i].t0).count();
  }
  return WFA_OK;
}

void worker_main(wfa_server_t* s, Worker* w) {
  hipSetDevice(s->cfg.device);
  std::unique_lock<std::mutex> lk(s->mu);
  std::vector<Req> batch;
  std::vector<Out> outs;
  for (;;) {
    wfa_attr_t attr;
    int fmt = 0;
    int64_t fp = 0;
    bool got = false, per_pair = false, is_long = false;
    s->cv_work.wait(lk, [&] {
      return s->stop || (got = pick(s, &attr, &fmt, &per_pair, &is_long, &batch, &fp));
    });
    if (!got) break;
    if (is_long) s->long_running += 1;
    const int64_t n = (int64_t)batch.size();
    s->queued -= n;
    s->running += n;
    s->in_flight += fp;
    s->launches += 1;
    s->launched_pairs += n;
    lk.unlock();
    const int r = run_batch(*w, attr, fmt, per_pair, batch, &outs);
    if (r) {
      const Clock::time_point t1 = Clock::now();
      outs.resize(batch.size());
      for (size_t i = 0; i < batch.size(); ++i) {
        outs[i].tag = batch[i].tag;
        outs[i].error = r;
        outs[i].score = 0;
        outs[i].status = WFA_STATUS_OK;
        outs[i].cigar.clear();
        outs[i].latency_us = std::chrono::duration<double, std::micro>(t1 - batch[i].t0).count();
      }
    }
    lk.lock();
    for (Out& o : outs) s->done.push_back(std::move(o));
    outs.clear();
    batch.clear();
    s->running -= n;
    s->in_flight -= fp;
    if (is_long) s->long_running -= 1;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

void destroy(wfa_server_t* s) {
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& w : s->workers)
    if (w->th.joinable()) w->th.join();
  hipSetDevice(s->cfg.device);
  for (auto& w : s->workers) {
    if (w->st) hipStreamSynchronize(w->st);
    if (w->h) hipHostFree(w->h);
    if (w->d) hipFree(w->d);
    if (w->st) hipStreamDestroy(w->st);
    if (w->e) wfa_engine_free(w->e);
  }
  delete s;
}

}  // namespace

extern "C" {

void wfa_server_config_default(wfa_server_config_t* c) {
  if (!c) return;
  c->device = 0;
  c->num_engines = 4;
  c->max_batch_pairs = 1 << 16;
  c->memory_budget = (int64_t)64 << 30;
}

wfa_server_t* wfa_server_new(const wfa_server_config_t* cfg, int32_t* err) {
  wfa_server_config_t c;
  wfa_server_config_default(&c);
  if (cfg) c = *cfg;
  if (c.num_engines < 1 || c.num_engines > 16 || c.max_batch_pairs < 1 || c.memory_budget <= 0) {
    if (err) *err = WFA_ERR_INVALID_ARGUMENT;
    return nullptr;
  }
  auto* s = new wfa_server_t();
  s->cfg = c;
  for (int i = 0; i < c.num_engines; ++i) {
    auto w = std::make_unique<Worker>();
    int32_t e = WFA_OK;
    w->e = wfa_engine_new(c.device, &e);
    if (!w->e) {
      s->workers.push_back(std::move(w));
      destroy(s);
      if (err) *err = e;
      return nullptr;
    }
    // Admission keeps the predicted total within the budget; an engine's
    // own launches may use up to twice its even share (at most the whole
    // budget), so one large request fits in one engine
    w->e->auto_budget = false;
    w->e->scratch_budget = std::max<int64_t>(
        std::min(c.memory_budget, 2 * c.memory_budget / c.num_engines), (int64_t)1 << 30);
    // a long-read launch takes at most half a workgroup per CU on average,
    // so short requests' launches always find room on the CUs
    w->e->long_grid_cap = std::max(1, w->e->num_cus / 2);
    // (its history pool may take the whole arena: no allocation is made
    // for it, so a long launch needs fewer retry passes)
    w->e->long_pool_whole_budget = true;
    if (hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking) != hipSuccess) {
      s->workers.push_back(std::move(w));
      destroy(s);
      if (err) *err = WFA_ERR_DEVICE;
      return nullptr;
    }
    // Allocations wait for the whole device (hipMalloc / hipFree, and the
    // stream-ordered allocator as well), so while one engine's long launch
    // runs, another engine that grows a buffer would stall for as long.
    // Every engine therefore takes its scratch from one arena allocated
    // here (its launch plans fit the arena, the pairs that do not end as
    // WFA_STATUS_OOM), and the staging buffers start at a size that holds
    // a typical launch.
    void* arena = nullptr;
    const size_t arena_bytes = (size_t)w->e->scratch_budget;
    if (hipMalloc(&arena, arena_bytes) != hipSuccess) {
      (void)hipGetLastError();
      s->workers.push_back(std::move(w));
      destroy(s);
      if (err) *err = WFA_ERR_OUT_OF_MEMORY;
      return nullptr;
    }
    w->e->arena = static_cast<uint8_t*>(arena);
    w->e->arena_cap = arena_bytes;
    w->e->own_arena = true;
    if (grow_pinned(&w->h, &w->h_cap, kStagingBytes) ||
        grow_device(&w->d, &w->d_cap, kStagingBytes)) {
      s->workers.push_back(std::move(w));
      destroy(s);
      if (err) *err = WFA_ERR_OUT_OF_MEMORY;
      return nullptr;
    }
    s->workers.push_back(std::move(w));
  }
  for (auto& w : s->workers) w->th = std::thread(worker_main, s, w.get());
  if (err) *err = WFA_OK;
  return s;
}

int32_t wfa_server_submit(wfa_server_t* s, const wfa_server_request_t* rq) {
  if (!s || !rq) return WFA_ERR_INVALID_ARGUMENT;
  Internal ip;
  int r = make_internal(rq->attr.penalties, &ip);
  if (r) return r;
  r = check_supported(rq->attr);
  if (r) return r;
  if (rq->seq_format != WFA_SEQ_ASCII && rq->seq_format != WFA_SEQ_PACKED2BITS)
    return WFA_ERR_INVALID_ARGUMENT;
  if (rq->pattern_length < 0 || rq->text_length < 0 || (rq->pattern_length && !rq->pattern) ||
      (rq->text_length && !rq->text))
    return WFA_ERR_INVALID_ARGUMENT;
  Req q;
  q.tag = rq->tag;
  q.plen = rq->pattern_length;
  q.tlen = rq->text_length;
  // ASCII sequences of A/C/G/T only travel 2-bit packed (a quarter of the
  // bytes; the kernels stage them 2-bit either way)
  int fmt = rq->seq_format;
  if (fmt == WFA_SEQ_ASCII && acgt_only(rq->pattern, q.plen) && acgt_only(rq->text, q.tlen))
    fmt = WFA_SEQ_PACKED2BITS;
  const bool packed = fmt == WFA_SEQ_PACKED2BITS;
  q.pbytes = packed ? ((size_t)q.plen + 3) / 4 : (size_t)q.plen;
  q.tbytes = packed ? ((size_t)q.tlen + 3) / 4 : (size_t)q.tlen;
  q.seq.resize(q.pbytes + q.tbytes);
  if (rq->seq_format == WFA_SEQ_ASCII && packed) {
    pack2(rq->pattern, q.plen, q.seq.data());
    pack2(rq->text, q.tlen, q.seq.data() + q.pbytes);
  } else {
    if (q.pbytes) std::memcpy(q.seq.data(), rq->pattern, q.pbytes);
    if (q.tbytes) std::memcpy(q.seq.data() + q.pbytes, rq->text, q.tbytes);
  }
  q.footprint = predict_footprint(rq->attr, ip, q.plen, q.tlen);
  const int bucket = bucket_of(std::max(q.plen, q.tlen));
  // the class key: the configuration with span and heuristic reduced to
  // their envelope; the request's own values travel per pair
  wfa_attr_t key = rq->attr;
  std::memset(&q.cfg, 0, sizeof(q.cfg));
  q.cfg.span = key.span;
  q.cfg.extension = key.extension;
  q.cfg.pattern_begin_free = key.pattern_begin_free;
  q.cfg.pattern_end_free = key.pattern_end_free;
  q.cfg.text_begin_free = key.text_begin_free;
  q.cfg.text_end_free = key.text_end_free;
  q.cfg.heuristic = key.heuristic;
  for (int i = 0; i < 3; ++i) q.cfg.heuristic_p[i] = key.heuristic_p[i];
  const bool per_pair = key.memory_mode != WFA_MEMORY_ULTRALOW;
  if (per_pair) {
    key.extension = 0;
    key.pattern_begin_free = key.pattern_end_free = key.text_begin_free = key.text_end_free = 0;
    if (key.heuristic != WFA_HEURISTIC_NONE) key.heuristic = WFA_HEURISTIC_WFADAPTIVE;
    key.heuristic_p[0] = key.heuristic_p[1] = key.heuristic_p[2] = 0;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return WFA_ERR_INVALID_ARGUMENT;
    q.t0 = Clock::now();
    Cls* c = nullptr;
    for (Cls& k : s->classes)
      if (k.fmt == fmt && k.bucket == bucket &&
          std::memcmp(&k.attr, &key, sizeof(wfa_attr_t)) == 0) {
        c = &k;
        break;
      }
    if (!c) {
      s->classes.push_back(Cls{key, fmt, bucket, per_pair, {}});
      c = &s->classes.back();
    }
    c->q.push_back(std::move(q));
    s->queued += 1;
  }
  s->cv_work.notify_one();
  return WFA_OK;
}

int32_t wfa_server_poll(wfa_server_t* s, wfa_server_result_t* out, int32_t max_results,
                        int32_t timeout_us) {
  if (!s || (max_results > 0 && !out) || max_results < 0) return WFA_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return !s->done.empty() || (s->queued == 0 && s->running == 0); };
  if (timeout_us < 0)
    s->cv_done.wait(lk, ready);
  else if (timeout_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready);
  s->polled.clear();
  while (!s->done.empty() && (int32_t)s->polled.size() < max_results) {
    s->polled.push_back(std::move(s->done.front()));
    s->done.pop_front();
  }
  for (size_t i = 0; i < s->polled.size(); ++i) {
    const Out& o = s->polled[i];
    wfa_server_result_t& d = out[i];
    d.tag = o.tag;
    d.error = o.error;
    d.score = o.score;
    d.status = o.status;
    d.cigar_length = (int32_t)o.cigar.size();
    d.cigar_ops = o.cigar.empty() ? nullptr : o.cigar.data();
    d.latency_us = o.latency_us;
  }
  return (int32_t)s->polled.size();
}

int32_t wfa_server_stats(wfa_server_t* s, int64_t* queued, int64_t* running, int64_t* launches,
                         int64_t* launched_pairs) {
  if (!s) return WFA_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(s->mu);
  if (queued) *queued = s->queued;
  if (running) *running = s->running;
  if (launches) *launches = s->launches;
  if (launched_pairs) *launched_pairs = s->launched_pairs;
  return WFA_OK;
}

void wfa_server_free(wfa_server_t* s) {
  if (s) destroy(s);
}

}  // extern "C"
