/*
 * Ingest pipeline for one stream: submitted frames are flattened and uploaded
 * by a pool of planner threads (av1h_plan_frame, which runs concurrently with
 * decodes and with itself), while one worker thread decodes them in
 * submission order (av1h_decode_planned), releases the slots each frame lets
 * go and recycles its plan. Frame N+1.. are planned while frame N is on the
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

void worker_main(av1h_pipe* p) {
  std::unique_lock<std::mutex> lk(p->mu);
  for (;;) {
    if (p->frames.empty() || !p->frames.front()->planned) {
      if (p->stop && p->frames.empty()) return;
      p->cv_work.wait(lk);
      continue;
    }
    PFrame* f = p->frames.front();
    lk.unlock();
    const double t_start = av1h_server_now_ms();
    double flatten_ms = 0, upload_ms = 0;
    if (f->plan) (void)av1h_plan_host_times(f->plan, &flatten_ms, &upload_ms);
    int slot = -1;
    int rc = f->plan_rc;
    std::string msg;
    uint64_t hash = 0;
    if (rc == AV1H_OK) {
      rc = av1h_decode_planned(p->ctx, f->plan, f->keep, f->reserved, nullptr, nullptr, &slot);
      if (rc != AV1H_OK)
        msg = av1h_last_error(p->ctx);
      else
        (void)av1h_last_hash(p->ctx, &hash);
    }
    if (f->plan) av1h_discard_plan(p->ctx, f->plan);
    for (int32_t s : f->release) av1h_release_reference(p->ctx, s);
    const double t = av1h_server_now_ms();
    lk.lock();
    av1h_pipe_frame& r = p->results[(size_t)f->seq];
    r.slot = slot;
    r.status = rc;
    r.plan_start_ms = f->plan_start_ms;
    r.planned_ms = f->planned_ms;
    r.decode_start_ms = t_start;
    r.done_ms = t;
    r.flatten_ms = flatten_ms;
    r.upload_ms = upload_ms;
    r.hash = hash;
    if (rc != AV1H_OK) {
      if (p->first_failed < 0) p->first_failed = f->seq;
      if (p->err.empty()) p->err = msg;
    }
    p->frames.pop_front();
    p->completed = f->seq + 1;
    delete f;
    p->cv_done.notify_all();
    p->cv_space.notify_all();
    p->cv_plan.notify_all();
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
