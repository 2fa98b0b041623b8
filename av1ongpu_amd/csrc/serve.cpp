/*
 * Serving loop over the batch entry point: decode requests of many concurrent
 * streams are admitted against a device memory budget, planned by a pool of
 * host threads, and decoded in batches formed continuously - a frame that
 * arrives while a batch is on the device joins the next one. Frames of one
 * stream decode in submission order (each predicts from the ones before it);
 * frames of different streams share every launch of a batch.
 *
 * The reference has no serving layer; its nearest relative is the pipelined
 * ingest of dsfilter/decode_core.{hpp,cpp} (planner, worker and output
 * threads around Decoder::plan_frame / decode_reference_planned,
 * decode_core.hpp:164-292) for one stream. Here the worker is one scheduler
 * thread that owns the device and batches across streams.
 */
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "av1hip.h"

namespace {

double clock_ms() {
  using clk = std::chrono::steady_clock;
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

struct SFrame {
  uint64_t tag = 0;
  int stream = -1;
  double arrival = 0;
  std::vector<uint8_t> wire; /* to be planned, or empty */
  const uint8_t* wire_ref = nullptr; /* borrowed wire bytes to be planned */
  size_t wire_ref_len = 0;
  double plan_start = 0, planned_at = 0; /* planning of a wire frame */
  void* plan = nullptr;
  bool own_plan = true;
  bool planned = false;
  bool planning = false;
  int plan_rc = AV1H_OK;
  int reserved = -1;
  std::vector<int> release;
  bool heavy = false; /* its queue's critical path is long (a key frame) */
};

enum class StreamState { Waiting, Admitted };

struct SStream {
  int id = -1;
  int w = 0, h = 0, bd = 8, ssx = 1, ssy = 1, slots = 8;
  int64_t bytes = 0;
  av1h_ctx* ctx = nullptr;
  StreamState state = StreamState::Waiting;
  bool closing = false;
  bool busy = false; /* a frame of it is in the running batch */
  std::deque<SFrame*> q;
};

}  // namespace

struct av1h_server {
  av1h_server_config cfg;
  std::mutex mu;
  std::condition_variable cv_work, cv_plan, cv_done;
  std::map<int, SStream*> streams;
  std::deque<int> waiting; /* admission order */
  int next_id = 0;
  int64_t used = 0;
  bool stop = false;
  std::deque<SFrame*> to_plan;
  std::deque<av1h_completion> done;
  std::vector<std::thread> planners;
  std::string err;
  av1h_server_stats stats;
  int running = 0;         /* batches on the device, over every launcher */
  double busy_since = 0;   /* when `running` last left 0               */
  /* one scheduler thread per launcher context: each forms and runs its own
   * batches, so a long batch (a large key frame) does not hold back the
   * frames that arrive behind it */
  std::vector<av1h_ctx*> launchers;
  std::vector<std::thread> scheds;
};

namespace {

/* Admits waiting streams in arrival order while their state fits the budget;
 * the first stream is always admitted so an undersized budget still makes
 * progress. Called with the lock held. */
void admit(av1h_server* s) {
  while (!s->waiting.empty()) {
    SStream* st = s->streams[s->waiting.front()];
    if (s->used > 0 && s->used + st->bytes > s->cfg.hbm_budget) return;
    av1h_ctx* ctx = nullptr;
    if (av1h_create(s->cfg.device, &ctx) != AV1H_OK) {
      s->err = "stream context creation failed";
      return;
    }
    av1h_set_pool_slots(ctx, st->slots);
    if (s->cfg.nv12_pitch && st->ssx == 1 && st->ssy == 1) av1h_set_decode_nv12(ctx, s->cfg.nv12_pitch);
    if (s->cfg.hash_pictures) av1h_set_output_hash(ctx, 1);
    st->ctx = ctx;
    st->state = StreamState::Admitted;
    s->used += st->bytes;
    s->waiting.pop_front();
    ++s->stats.admitted;
    for (SFrame* f : st->q)
      if (!f->planned && !f->planning) s->to_plan.push_back(f);
    s->cv_plan.notify_all();
  }
}

/* Ends a closed stream whose frames are all done. Lock held. */
void retire(av1h_server* s, SStream* st) {
  if (!st->closing || st->busy || !st->q.empty()) return;
  if (st->state == StreamState::Admitted) {
    av1h_destroy(st->ctx);
    s->used -= st->bytes;
  } else {
    s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), st->id), s->waiting.end());
  }
  s->streams.erase(st->id);
  delete st;
  admit(s);
}

void complete(av1h_server* s, SFrame* f, int slot, int status, double start, double end, int batch,
              uint16_t* pic, int64_t samples, uint64_t hash) {
  av1h_completion c;
  std::memset(&c, 0, sizeof(c));
  c.tag = f->tag;
  c.stream = f->stream;
  c.slot = slot;
  c.status = status;
  c.batch_frames = batch;
  c.arrival_ms = f->arrival;
  c.start_ms = start;
  c.done_ms = end;
  c.picture = pic;
  c.picture_samples = samples;
  c.hash = hash;
  c.plan_start_ms = f->plan_start > 0 ? f->plan_start : f->arrival;
  c.planned_ms = f->planned_at > 0 ? f->planned_at : f->arrival;
  s->done.push_back(c);
}

void planner_main(av1h_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_plan.wait(lk, [&] { return s->stop || !s->to_plan.empty(); });
    if (s->stop) return;
    SFrame* f = s->to_plan.front();
    s->to_plan.pop_front();
    SStream* st = s->streams[f->stream];
    av1h_ctx* ctx = st->ctx;
    f->planning = true;
    f->plan_start = clock_ms();
    lk.unlock();
    void* plan = nullptr;
    const int rc = f->wire_ref ? av1h_plan_frame_wire(ctx, f->wire_ref, f->wire_ref_len, &plan)
                               : av1h_plan_frame_wire(ctx, f->wire.data(), f->wire.size(), &plan);
    const double t_planned = clock_ms();
    lk.lock();
    f->planned_at = t_planned;
    f->planning = false;
    f->planned = true;
    f->plan_rc = rc;
    f->plan = rc == AV1H_OK ? plan : nullptr;
    f->heavy = f->plan && s->cfg.heavy_levels > 0 && av1h_plan_levels(f->plan) > s->cfg.heavy_levels;
    std::vector<uint8_t>().swap(f->wire);
    f->wire_ref = nullptr;
    s->cv_work.notify_one();
  }
}

void capture(av1h_ctx* ctx, uint16_t** pic, int64_t* samples) {
  *pic = nullptr;
  *samples = 0;
  av1h_frame_info g;
  if (av1h_read_slot(ctx, -1, nullptr, &g) != AV1H_OK) return;
  const int cw = (g.width + (1 << g.ss_x) - 1) >> g.ss_x, ch = (g.height + (1 << g.ss_y) - 1) >> g.ss_y;
  const int64_t n = (int64_t)g.width * g.height + 2 * (int64_t)cw * ch;
  uint16_t* buf = new uint16_t[(size_t)n];
  av1h_planes pl;
  pl.plane[0] = buf;
  pl.plane[1] = buf + (int64_t)g.width * g.height;
  pl.plane[2] = pl.plane[1] + (int64_t)cw * ch;
  pl.stride[0] = g.width;
  pl.stride[1] = pl.stride[2] = cw;
  if (av1h_read_slot(ctx, -1, &pl, nullptr) != AV1H_OK) {
    delete[] buf;
    return;
  }
  *pic = buf;
  *samples = n;
}

void sched_main(av1h_server* s, av1h_ctx* launcher, int role) {
  /* role 0: any frame; 1: heavy frames only; 2: light frames only */
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->stop) return;
    admit(s);
    /* the head frame of every idle admitted stream, oldest arrival first */
    std::vector<SStream*> ready;
    std::vector<SStream*> failed;
    for (auto& kv : s->streams) {
      SStream* st = kv.second;
      if (st->state != StreamState::Admitted || st->busy || st->q.empty()) continue;
      SFrame* f = st->q.front();
      if (!f->planned) continue;
      if (f->plan_rc != AV1H_OK)
        failed.push_back(st);
      else if (role == 0 || (role == 1) == f->heavy)
        ready.push_back(st);
    }
    for (SStream* st : failed) {
      SFrame* f = st->q.front();
      st->q.pop_front();
      const double t = clock_ms();
      complete(s, f, -1, f->plan_rc, t, t, 0, nullptr, 0, 0);
      delete f;
      s->cv_done.notify_all();
      retire(s, st);
    }
    if (ready.empty()) {
      if (!failed.empty()) continue;
      s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    std::sort(ready.begin(), ready.end(),
              [](const SStream* a, const SStream* b) { return a->q.front()->arrival < b->q.front()->arrival; });
    /* a heavy frame's queue is bound by its own critical path; heavy
     * frames batch among themselves (heavy_batch), beside the light ones */
    const int cap = role == 1 ? std::max(s->cfg.heavy_batch, 1) : s->cfg.max_batch;
    if ((int)ready.size() > cap) ready.resize(cap);
    const int n = (int)ready.size();
    std::vector<av1h_ctx*> ctxs(n);
    std::vector<void*> plans(n);
    std::vector<int> keep(n), reserved(n), slots(n, -1);
    std::vector<SFrame*> frames(n);
    for (int i = 0; i < n; ++i) {
      SStream* st = ready[i];
      st->busy = true;
      frames[i] = st->q.front();
      ctxs[i] = st->ctx;
      plans[i] = frames[i]->plan;
      keep[i] = frames[i]->reserved >= 0;
      reserved[i] = frames[i]->reserved;
    }
    const double t0 = clock_ms();
    /* busy time is the union of the launchers' batch spans */
    if (s->running++ == 0) s->busy_since = t0;
    lk.unlock();
    const int rc = av1h_decode_planned_batch_on(launcher, ctxs.data(), plans.data(), keep.data(), reserved.data(),
                                                n, slots.data());
    const double t1 = clock_ms();
    std::vector<uint16_t*> pics(n, nullptr);
    std::vector<int64_t> samples(n, 0);
    std::vector<uint64_t> hashes(n, 0);
    if (rc == AV1H_OK && s->cfg.hash_pictures)
      for (int i = 0; i < n; ++i) av1h_last_hash(ctxs[i], &hashes[i]);
    if (rc == AV1H_OK && s->cfg.capture)
      for (int i = 0; i < n; ++i) capture(ctxs[i], &pics[i], &samples[i]);
    std::string msg = rc == AV1H_OK ? std::string() : std::string(av1h_last_error(launcher));
    for (int i = 0; i < n; ++i) {
      SFrame* f = frames[i];
      for (int r : f->release) av1h_release_reference(ctxs[i], r);
      if (f->own_plan && f->plan) av1h_discard_plan(ctxs[i], f->plan);
    }
    lk.lock();
    if (rc != AV1H_OK) s->err = msg;
    if (--s->running == 0) s->stats.busy_ms += clock_ms() - s->busy_since;
    ++s->stats.batches;
    s->stats.frames += n;
    s->stats.max_batch_frames = std::max(s->stats.max_batch_frames, n);
    for (int i = 0; i < n; ++i) {
      SStream* st = ready[i];
      SFrame* f = frames[i];
      st->q.pop_front();
      st->busy = false;
      complete(s, f, rc == AV1H_OK ? slots[i] : -1, rc, t0, t1, n, pics[i], samples[i], hashes[i]);
      delete f;
    }
    for (int i = 0; i < n; ++i) retire(s, ready[i]);
    s->cv_done.notify_all();
  }
}

int queue_frame(av1h_server* s, int stream, SFrame* f) {
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->streams.find(stream);
  if (it == s->streams.end() || it->second->closing) {
    delete f;
    return AV1H_ERR_INVALID;
  }
  SStream* st = it->second;
  f->stream = stream;
  st->q.push_back(f);
  if (!f->planned && st->state == StreamState::Admitted) {
    s->to_plan.push_back(f);
    s->cv_plan.notify_one();
  }
  s->cv_work.notify_one();
  return AV1H_OK;
}

}  // namespace

extern "C" {

double av1h_server_now_ms(void) { return clock_ms(); }

int av1h_server_create(const av1h_server_config* cfg, av1h_server** out) {
  if (!cfg || !out || cfg->max_batch < 1 || cfg->max_batch > 128 || cfg->planner_threads < 0 ||
      cfg->launchers < 0 || cfg->launchers > 8 || cfg->heavy_batch < 0 || cfg->heavy_batch > 128 ||
      (cfg->nv12_pitch & 1))
    return AV1H_ERR_INVALID;
  *out = nullptr;
  av1h_device_info d;
  const int rc = av1h_describe_device(cfg->device, &d);
  if (rc != AV1H_OK) return rc;
  av1h_server* s = new av1h_server;
  s->cfg = *cfg;
  const int nl = cfg->launchers > 0 ? cfg->launchers : 2;
  for (int i = 0; i < nl; ++i) {
    av1h_ctx* l = nullptr;
    if (av1h_create(cfg->device, &l) != AV1H_OK) {
      for (av1h_ctx* c : s->launchers) av1h_destroy(c);
      delete s;
      return AV1H_ERR_DEVICE;
    }
    /* batches of changing size and content: one captured graph per size */
    av1h_set_graphs(l, 2);
    s->launchers.push_back(l);
  }
  std::memset(&s->stats, 0, sizeof(s->stats));
  for (int i = 0; i < cfg->planner_threads; ++i) s->planners.emplace_back(planner_main, s);
  const bool split = cfg->heavy_levels > 0 && nl > 1;
  const int nh = split ? std::min(std::max(cfg->heavy_launchers, 1), nl - 1) : 0;
  for (int i = 0; i < nl; ++i) s->scheds.emplace_back(sched_main, s, s->launchers[(size_t)i], split ? (i < nh ? 1 : 2) : 0);
  *out = s;
  return AV1H_OK;
}

void av1h_server_destroy(av1h_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_plan.notify_all();
  s->cv_work.notify_all();
  for (std::thread& t : s->scheds) t.join();
  for (std::thread& t : s->planners) t.join();
  for (auto& kv : s->streams) {
    SStream* st = kv.second;
    for (SFrame* f : st->q) {
      if (f->own_plan && f->plan) av1h_discard_plan(st->ctx, f->plan);
      delete f;
    }
    if (st->ctx) av1h_destroy(st->ctx);
    delete st;
  }
  for (av1h_completion& c : s->done) delete[] c.picture;
  for (av1h_ctx* l : s->launchers) av1h_destroy(l);
  delete s;
}

int av1h_server_open_stream(av1h_server* s, int width, int height, int bit_depth, int ss_x, int ss_y,
                            int pool_slots) {
  if (!s) return AV1H_ERR_INVALID;
  const int64_t bytes =
      av1h_stream_bytes(width, height, bit_depth, ss_x, ss_y, pool_slots, ss_x == 1 && ss_y == 1 ? s->cfg.nv12_pitch : 0);
  if (bytes < 0) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  SStream* st = new SStream;
  st->id = s->next_id++;
  st->w = width;
  st->h = height;
  st->bd = bit_depth;
  st->ssx = ss_x;
  st->ssy = ss_y;
  st->slots = pool_slots;
  st->bytes = bytes;
  s->streams[st->id] = st;
  s->waiting.push_back(st->id);
  s->cv_work.notify_one();
  return st->id;
}

int av1h_server_submit_wire(av1h_server* s, int stream, const uint8_t* wire, size_t len, int reserved_slot,
                            const int32_t* release, int nrelease, double arrival_ms, uint64_t tag) {
  if (!s || !wire || nrelease < 0 || (nrelease && !release)) return AV1H_ERR_INVALID;
  if (s->cfg.planner_threads < 1) return AV1H_ERR_UNSUPPORTED;
  SFrame* f = new SFrame;
  f->wire.assign(wire, wire + len);
  f->reserved = reserved_slot;
  f->release.assign(release, release + nrelease);
  f->arrival = arrival_ms;
  f->tag = tag;
  return queue_frame(s, stream, f);
}

int av1h_server_submit_wire_ref(av1h_server* s, int stream, const uint8_t* wire, size_t len, int reserved_slot,
                                const int32_t* release, int nrelease, double arrival_ms, uint64_t tag) {
  if (!s || !wire || nrelease < 0 || (nrelease && !release)) return AV1H_ERR_INVALID;
  if (s->cfg.planner_threads < 1) return AV1H_ERR_UNSUPPORTED;
  SFrame* f = new SFrame;
  f->wire_ref = wire;
  f->wire_ref_len = len;
  f->reserved = reserved_slot;
  f->release.assign(release, release + nrelease);
  f->arrival = arrival_ms;
  f->tag = tag;
  return queue_frame(s, stream, f);
}

int av1h_server_submit_plan(av1h_server* s, int stream, void* plan, int reserved_slot, const int32_t* release,
                            int nrelease, double arrival_ms, uint64_t tag) {
  if (!s || !plan || nrelease < 0 || (nrelease && !release)) return AV1H_ERR_INVALID;
  SFrame* f = new SFrame;
  f->plan = plan;
  f->own_plan = false;
  f->planned = true;
  f->heavy = s->cfg.heavy_levels > 0 && av1h_plan_levels(plan) > s->cfg.heavy_levels;
  f->reserved = reserved_slot;
  f->release.assign(release, release + nrelease);
  f->arrival = arrival_ms;
  f->tag = tag;
  return queue_frame(s, stream, f);
}

int av1h_server_close_stream(av1h_server* s, int stream) {
  if (!s) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->streams.find(stream);
  if (it == s->streams.end()) return AV1H_ERR_INVALID;
  it->second->closing = true;
  retire(s, it->second);
  s->cv_work.notify_one();
  return AV1H_OK;
}

int av1h_server_poll(av1h_server* s, av1h_completion* out, int max, int wait_ms) {
  if (!s || !out || max < 1) return AV1H_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
  int n = 0;
  while (n < max && !s->done.empty()) {
    out[n++] = s->done.front();
    s->done.pop_front();
  }
  return n;
}

void av1h_server_free_picture(uint16_t* picture) { delete[] picture; }

int av1h_server_get_stats(av1h_server* s, av1h_server_stats* out) {
  if (!s || !out) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->budget_used = s->used;
  out->waiting = (int64_t)s->waiting.size();
  out->active = (int64_t)s->streams.size() - out->waiting;
  return AV1H_OK;
}

av1h_ctx* av1h_server_launcher(av1h_server* s, int index) {
  if (!s || index < 0 || index >= (int)s->launchers.size()) return nullptr;
  return s->launchers[index];
}

const char* av1h_server_last_error(av1h_server* s) { return s ? s->err.c_str() : "no server"; }

}  // extern "C"
