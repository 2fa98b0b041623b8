/*
 * Ingest pipeline for one stream: submitted frames are flattened and uploaded
 * by a pool of planner threads (av1h_plan_frame, which runs concurrently with
 * decodes and with itself), while one worker thread launches them in
 * submission order (av1h_decode_planned_batch_async, two frames in flight),
 * releases the slots each frame lets go, takes the results in order and
 * recycles the plans. Frame N+1.. are planned while frame N is on the
 * device, so a stream's throughput is set by the slower of the two: the
 * planners together, or the device.
 *
 * Reference: the pipelined ingest of dsfilter/decode_core.{hpp,cpp} (a
 * planner thread builds frame N's plan while the worker runs frame N-1,
 * decode_core.hpp:164-292, with kPipeDepth frames in flight). Here the planner
 * stage is several threads, since one 1080p frame's flattening takes longer
 * than its decode.
 */
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "av1hip.h"

namespace {

struct PFrame {
  int64_t seq = 0;
  av1h_frame_desc desc;
  int keep = 1, reserved = -1;
  std::vector<int32_t> release;
  void* plan = nullptr;
  int plan_rc = AV1H_OK;
  bool planned = false, planning = false;
  double plan_start_ms = 0, planned_ms = 0;
};

}  // namespace

struct av1h_pipe {
  av1h_ctx* ctx = nullptr;
  int depth = 16;
  std::mutex mu;
  std::condition_variable cv_plan, cv_work, cv_done, cv_space;
  std::deque<PFrame*> frames; /* submitted, not yet decoded, in order */
  int64_t next_seq = 0;       /* the next submission's number          */
  int64_t completed = 0;      /* frames [0, completed) are done         */
  std::vector<av1h_pipe_frame> results;
  int64_t first_failed = -1;
  std::string err;
  bool stop = false;
  std::vector<std::thread> planners;
  std::thread worker;
};

namespace {

void planner_main(av1h_pipe* p) {
  std::unique_lock<std::mutex> lk(p->mu);
  for (;;) {
    PFrame* f = nullptr;
    for (PFrame* c : p->frames)
      if (!c->planned && !c->planning) {
        f = c;
        break;
      }
    if (!f) {
      if (p->stop) return;
      p->cv_plan.wait(lk);
      continue;
    }
    f->planning = true;
    f->plan_start_ms = av1h_server_now_ms();
    lk.unlock();
    void* plan = nullptr;
    const int rc = av1h_plan_frame(p->ctx, &f->desc, &plan);
    std::string msg = rc != AV1H_OK ? av1h_last_error(p->ctx) : "";
    const double t = av1h_server_now_ms();
    lk.lock();
    f->planned_ms = t;
    f->plan = plan;
    f->plan_rc = rc;
    f->planned = true;
    f->planning = false;
    if (rc != AV1H_OK && p->err.empty()) p->err = msg;
    p->cv_work.notify_all();
  }
}

/* A frame the worker has launched (or failed) and not yet completed. */
struct Flight {
  PFrame* f = nullptr;
  bool pending = false; /* its batch is in flight on the device */
  int rc = AV1H_OK;
  std::string msg;
  int slot = -1;
  uint64_t hash = 0;
  double t_start = 0, flatten_ms = 0, upload_ms = 0, launch_ms = 0, wait_ms = 0, device_ms = 0;
};

/* the device time of the context's last finished batch */
double last_device_ms(av1h_ctx* c) {
  av1h_stats st;
  if (av1h_get_stats(c, &st) != AV1H_OK) return 0;
  return st.device_recon_ms + st.device_filter_ms;
}

/* Completes the oldest frame in flight: waits for its batch, recycles its
 * plan and records its result. Called without the pipeline's lock. */
void complete(av1h_pipe* p, std::deque<Flight>* flight) {
  Flight x = std::move(flight->front());
  flight->pop_front();
  if (x.pending) {
    const double t0 = av1h_server_now_ms();
    x.rc = av1h_batch_finish(p->ctx);
    x.wait_ms = av1h_server_now_ms() - t0;
    if (x.rc != AV1H_OK) {
      x.msg = av1h_last_error(p->ctx);
    } else {
      (void)av1h_last_hash(p->ctx, &x.hash);
      x.device_ms = last_device_ms(p->ctx);
    }
  }
  if (x.f->plan) av1h_discard_plan(p->ctx, x.f->plan);
  const double t = av1h_server_now_ms();
  std::lock_guard<std::mutex> lk(p->mu);
  av1h_pipe_frame& r = p->results[(size_t)x.f->seq];
  r.slot = x.rc == AV1H_OK ? x.slot : -1;
  r.status = x.rc;
  r.plan_start_ms = x.f->plan_start_ms;
  r.planned_ms = x.f->planned_ms;
  r.decode_start_ms = x.t_start;
  r.done_ms = t;
  r.flatten_ms = x.flatten_ms;
  r.upload_ms = x.upload_ms;
  r.hash = x.hash;
  r.launch_ms = x.launch_ms;
  r.wait_ms = x.wait_ms;
  r.device_ms = x.device_ms;
  if (x.rc != AV1H_OK) {
    if (p->first_failed < 0) p->first_failed = x.f->seq;
    if (p->err.empty()) p->err = x.msg;
  }
  p->frames.pop_front();
  p->completed = x.f->seq + 1;
  delete x.f;
  p->cv_done.notify_all();
  p->cv_space.notify_all();
  p->cv_plan.notify_all();
}

/* Launches frame f as a batch of its own without waiting for it
 * (av1h_decode_planned_batch_async), then lets go of the slots it names. A
 * launch that would wait for the device anyway (buffers that grow, stage
 * timing) first completes every frame in flight, so that results are taken
 * in order. Called without the pipeline's lock. */
void launch(av1h_pipe* p, PFrame* f, std::deque<Flight>* flight) {
  Flight x;
  x.f = f;
  x.t_start = av1h_server_now_ms();
  x.rc = f->plan_rc;
  if (f->plan) (void)av1h_plan_host_times(f->plan, &x.flatten_ms, &x.upload_ms);
  if (x.rc == AV1H_OK) {
    av1h_ctx* c = p->ctx;
    if (av1h_plan_needs_idle(c, f->plan) != 0)
      while (!flight->empty()) complete(p, flight);
    void* plan = f->plan;
    int keep = f->keep, reserved = f->reserved;
    const double t0 = av1h_server_now_ms();
    x.rc = av1h_decode_planned_batch_async(c, &c, &plan, &keep, &reserved, 1, &x.slot);
    x.launch_ms = av1h_server_now_ms() - t0;
    if (x.rc != AV1H_OK) {
      x.msg = av1h_last_error(c);
    } else if (av1h_batch_pending(c) == 0) {
      (void)av1h_last_hash(c, &x.hash); /* the launch waited for it */
      x.device_ms = last_device_ms(c);
    } else {
      x.pending = true;
    }
  } else {
    std::lock_guard<std::mutex> lk(p->mu);
    x.msg = p->err;
  }
  for (int32_t s : f->release) av1h_release_reference(p->ctx, s);
  flight->push_back(std::move(x));
}

/* Decodes the planned frames in submission order with up to two in flight
 * on the device: frame N + 1 is queued behind frame N before frame N is
 * waited for, so the device does not idle while the host takes a result and
 * launches the next frame. */
void worker_main(av1h_pipe* p) {
  std::deque<Flight> flight;
  std::unique_lock<std::mutex> lk(p->mu);
  for (;;) {
    PFrame* next = flight.size() < p->frames.size() ? p->frames[flight.size()] : nullptr;
    if (next && next->planned && flight.size() < 2) {
      lk.unlock();
      launch(p, next, &flight);
      lk.lock();
      continue;
    }
    if (!flight.empty()) {
      lk.unlock();
      complete(p, &flight);
      lk.lock();
      continue;
    }
    if (p->stop && p->frames.empty()) return;
    p->cv_work.wait(lk);
  }
}

}  // namespace

extern "C" {

int av1h_pipe_create(av1h_ctx* ctx, int planner_threads, int depth, av1h_pipe** out) {
  if (!ctx || !out || planner_threads < 1 || planner_threads > 64 || depth < 1 || depth > 256)
    return AV1H_ERR_INVALID;
  av1h_pipe* p = new av1h_pipe;
  p->ctx = ctx;
  p->depth = depth;
  for (int i = 0; i < planner_threads; ++i) p->planners.emplace_back(planner_main, p);
  p->worker = std::thread(worker_main, p);
  *out = p;
  return AV1H_OK;
}

int64_t av1h_pipe_submit(av1h_pipe* p, const av1h_frame_desc* desc, int keep, int reserved,
                         const int32_t* release, int nrelease) {
  if (!p || !desc || nrelease < 0 || (nrelease && !release)) return AV1H_ERR_INVALID;
  PFrame* f = new PFrame;
  f->desc = *desc;
  f->keep = keep;
  f->reserved = reserved;
  f->release.assign(release, release + nrelease);
  std::unique_lock<std::mutex> lk(p->mu);
  while (p->next_seq - p->completed >= p->depth) p->cv_space.wait(lk);
  f->seq = p->next_seq++;
  av1h_pipe_frame r;
  std::memset(&r, 0, sizeof(r));
  r.slot = -1;
  r.submit_ms = av1h_server_now_ms();
  p->results.push_back(r);
  p->frames.push_back(f);
  p->cv_plan.notify_one();
  return f->seq;
}

int64_t av1h_pipe_submit_wire(av1h_pipe* p, const uint8_t* wire, size_t len, int keep, int reserved,
                              const int32_t* release, int nrelease) {
  av1h_frame_desc d;
  if (!p || !wire || av1h_wire_parse(wire, len, &d) != AV1H_OK) return AV1H_ERR_INVALID;
  return av1h_pipe_submit(p, &d, keep, reserved, release, nrelease);
}

int av1h_pipe_wait(av1h_pipe* p, int64_t seq, int64_t* failed_seq) {
  if (!p) return AV1H_ERR_INVALID;
  std::unique_lock<std::mutex> lk(p->mu);
  const int64_t upto = seq < 0 ? p->next_seq : std::min(seq + 1, p->next_seq);
  while (p->completed < upto) p->cv_done.wait(lk);
  if (failed_seq) *failed_seq = p->first_failed;
  if (p->first_failed >= 0 && p->first_failed < upto) return p->results[(size_t)p->first_failed].status;
  return AV1H_OK;
}

int av1h_pipe_result(av1h_pipe* p, int64_t seq, av1h_pipe_frame* out) {
  if (!p || !out) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(p->mu);
  if (seq < 0 || seq >= p->completed) return AV1H_ERR_INVALID;
  *out = p->results[(size_t)seq];
  return AV1H_OK;
}

const char* av1h_pipe_error(av1h_pipe* p) {
  if (!p) return "no pipeline";
  std::lock_guard<std::mutex> lk(p->mu);
  return p->err.c_str();
}

void av1h_pipe_destroy(av1h_pipe* p) {
  if (!p) return;
  {
    std::unique_lock<std::mutex> lk(p->mu);
    while (p->completed < p->next_seq) p->cv_done.wait(lk);
    p->stop = true;
    p->cv_plan.notify_all();
    p->cv_work.notify_all();
  }
  for (std::thread& t : p->planners) t.join();
  p->worker.join();
  delete p;
}

}  // extern "C"
