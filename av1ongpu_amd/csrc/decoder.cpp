/*
 * libav1gpu.so: the reference's av1gpu::Decoder API over the C ABI of
 * libav1hip.so. FrameSyntax is flattened into an av1h_frame_desc (borrowing
 * the syntax's pools) and handed to the HIP backend; the decoder keeps the
 * reference's slot bookkeeping (src/core/engine.cpp:50-433): the slots a
 * decode keeps are held until released, reset, or destruction, and a
 * pre-planned frame is consumed by its decode.
 */
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "av1gpu/av1gpu.hpp"
#include "av1hip.h"
#include "syntax_desc.hpp"

namespace av1gpu {

namespace {

const char* status_text(int rc) {
  switch (rc) {
    case AV1H_OK: return "ok";
    case AV1H_ERR_INVALID: return "invalid argument";
    case AV1H_ERR_UNSUPPORTED: return "unsupported";
    case AV1H_ERR_DEVICE: return "device error";
    case AV1H_ERR_NO_DEVICE: return "no usable gfx950 device";
    case AV1H_ERR_SLOTS: return "reference pool exhausted";
    default: return "error";
  }
}

void alloc_like(Frame* out, const av1h_frame_info& info) {
  if (!out->matches(info.width, info.height, info.bit_depth, info.ss_x, info.ss_y))
    out->allocate(info.width, info.height, info.bit_depth, info.ss_x, info.ss_y);
}

class HipDecoder final : public Decoder {
 public:
  HipDecoder(av1h_ctx* ctx, const BackendOptions& options) : ctx_(ctx), options_(options) {
    av1h_device_info di;
    std::memset(&di, 0, sizeof(di));
    av1h_get_device_info(ctx_, &di);
    info_.kind = BackendKind::HIP;
    info_.name = di.name;
    info_.vendor = "AMD";
    info_.version = di.arch;
    info_.detail = std::string("HIP, ") + std::to_string(di.compute_units) + " CUs";
    info_.available = true;
    info_.global_mem = static_cast<uint64_t>(di.global_mem);
    info_.compute_units = di.compute_units;
  }

  ~HipDecoder() override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (int s : held_) av1h_release_reference(ctx_, s);
      held_.clear();
    }
    av1h_destroy(ctx_);
  }

  const DeviceInfo& device() const override { return info_; }
  const BackendOptions& options() const override { return options_; }
  const DecodeStats& stats() const override { return stats_; }

  int decode_reference(const FrameSyntax& syntax, Frame* out, std::string* error) override {
    return run(syntax, true, -1, out, true, error);
  }

  bool decode(const FrameSyntax& syntax, Frame* out, std::string* error) override {
    /* a frame nobody keeps takes no pool slot (engine.cpp:85-93) */
    return run(syntax, false, -1, out, true, error) == -1 && last_ok_;
  }

  int decode_reference_reserved(const FrameSyntax& syntax, int slot, Frame* out, std::string* error,
                                bool read_back) override {
    if (slot < 0) {
      decode(syntax, out, error);
      return -1;
    }
    return run(syntax, true, slot, out, read_back, error);
  }

  int reserve_reference() override { return av1h_reserve_reference(ctx_); }

  bool plan_frame(const FrameSyntax& syntax, void** plan, std::string* error,
                  std::atomic<bool>* cancel) override {
    if (!plan) return fail(error, "plan_frame: no output"), false;
    *plan = nullptr;
    if (cancel && cancel->load()) return fail(error, "plan_frame: cancelled"), false;
    detail::DescStorage st;
    std::string e;
    if (!detail::flatten(syntax, &st, &e)) return fail(error, e), false;
    const int rc = av1h_plan_frame(ctx_, &st.desc, plan);
    if (rc != AV1H_OK) return fail_rc(error, rc), false;
    return true;
  }

  int decode_reference_planned(void* plan, int slot, Frame* out, std::string* error,
                               bool read_back) override {
    if (!plan) return fail(error, "decode_reference_planned: no plan"), -1;
    /* the plan is consumed by its decode, as in the reference
     * (engine.cpp:180-253) */
    const int got = run_planned(plan, slot, out, read_back, error);
    av1h_discard_plan(ctx_, plan);
    return got;
  }

  void discard_frame_plan(void* plan) override {
    if (plan) av1h_discard_plan(ctx_, plan);
  }

  bool supports_preplanning() const override { return true; }

  /* every slot packs on the device (av1h_to_nv12), the latest with its film
   * grain, so a pipeline may decode with read_back false and pack the slot */
  bool packs_nv12_on_device() const override { return true; }

  /* a frame whose syntax carries coded levels is dequantised on the device
   * (inside the transforms' coefficient staging): the raster pool may be
   * left empty */
  bool dequantises_on_device() const override { return true; }

  int nv12_pack_begin(int slot, size_t pitch, std::string* error) override {
    const int rc = av1h_nv12_pack_begin(ctx_, slot, pitch);
    if (rc < 0) return fail_rc(error, rc), -1;
    return rc;
  }

  bool nv12_pack_wait(int handle, std::vector<uint8_t>* nv12, std::string* error) override {
    size_t bytes = 0;
    int rc = av1h_nv12_pack_wait(ctx_, handle, nullptr, 0, &bytes);
    if (rc != AV1H_OK) return fail_rc(error, rc), false;
    if (nv12) {
      nv12->resize(bytes);
      rc = av1h_nv12_pack_wait(ctx_, handle, nv12->data(), nv12->size(), &bytes);
      if (rc != AV1H_OK) return fail_rc(error, rc), false;
    }
    return true;
  }

  void nv12_pack_release(int handle) override { av1h_nv12_pack_release(ctx_, handle); }

  /* the output gate of a pipelined caller: stored, as the reference engine
   * does (engine.cpp:464-466); decoding itself is unaffected */
  void hold_output(bool hold) override { hold_.store(hold, std::memory_order_relaxed); }

  void release_reference(int slot) override {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = std::find(held_.begin(), held_.end(), slot);
    if (it != held_.end()) held_.erase(it);
    av1h_release_reference(ctx_, slot);
  }

  bool to_rgb(int slot, ColorMatrix matrix, ColorRange range, std::vector<uint8_t>* rgb,
              std::string* error) override {
    if (!rgb) return fail(error, "to_rgb: no output"), false;
    av1h_frame_info info;
    if (!geometry(slot, &info, error)) return false;
    rgb->assign(static_cast<size_t>(info.width) * info.height * 3, 0);
    const int rc = av1h_to_rgb(ctx_, slot, static_cast<int>(matrix), static_cast<int>(range), rgb->data(),
                               rgb->size());
    if (rc != AV1H_OK) return fail_rc(error, rc), false;
    return true;
  }

  bool to_nv12(int slot, size_t pitch, std::vector<uint8_t>* nv12, std::string* error) override {
    if (!nv12) return fail(error, "to_nv12: no output"), false;
    av1h_frame_info info;
    if (!geometry(slot, &info, error)) return false;
    const size_t bytes = av1h_nv12_size(info.width, info.height, info.ss_x, info.ss_y, pitch);
    if (bytes == 0)
      return fail(error, "NV12 output needs a 4:2:0 frame and an even pitch at least as wide"), false;
    nv12->assign(bytes, 0);
    const int rc = av1h_to_nv12(ctx_, slot, pitch, nv12->data(), nv12->size());
    if (rc != AV1H_OK) return fail_rc(error, rc), false;
    return true;
  }

  void reset() override {
    std::lock_guard<std::mutex> lk(mu_);
    held_.clear();
    av1h_reset(ctx_);
    stats_ = DecodeStats();
  }

 private:
  bool fail(std::string* error, const std::string& msg) {
    if (error) *error = msg;
    return false;
  }
  bool fail_rc(std::string* error, int rc) {
    const char* e = av1h_last_error(ctx_);
    return fail(error, e && *e ? std::string(e) : std::string(status_text(rc)));
  }

  /* the geometry a slot's picture was decoded with (a read with no planes
   * fills only `info`) */
  bool geometry(int slot, av1h_frame_info* info, std::string* error) {
    const int rc = av1h_read_slot(ctx_, slot, nullptr, info);
    if (rc != AV1H_OK) return fail_rc(error, rc);
    return true;
  }

  void take_stats() {
    av1h_stats s;
    if (av1h_get_stats(ctx_, &s) != AV1H_OK) return;
    stats_.blocks = s.blocks;
    stats_.waves = s.waves;
    stats_.kernel_launches = s.kernel_launches;
    stats_.reconstruct_ms = s.reconstruct_ms;
    stats_.filter_ms = s.filter_ms;
    stats_.readback_ms = s.readback_ms;
    stats_.total_ms = s.total_ms;
  }

  void hold(int slot) {
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(mu_);
    held_.push_back(slot);
  }

  /* decode with the picture read into `out` (when given and read_back) */
  int finish(int rc, int slot, Frame* out, const av1h_frame_info& info, std::string* error) {
    last_ok_ = rc == AV1H_OK;
    if (rc != AV1H_OK) return fail_rc(error, rc), -1;
    take_stats();
    hold(slot);
    if (out) alloc_like(out, info);
    return slot;
  }

  int run(const FrameSyntax& syntax, bool keep, int reserved, Frame* out, bool read_back, std::string* error) {
    last_ok_ = false;
    detail::DescStorage st;
    std::string e;
    if (!detail::flatten(syntax, &st, &e)) return fail(error, e), -1;
    av1h_frame_info info;
    std::memset(&info, 0, sizeof(info));
    int slot = -1;
    if (out && read_back) {
      /* the output frame is allocated for the display geometry first, so the
       * decode writes straight into its planes */
      out->allocate(syntax.width, syntax.height, syntax.bit_depth, syntax.ss_x, syntax.ss_y);
      av1h_planes pl;
      for (int p = 0; p < 3; ++p) {
        pl.plane[p] = out->plane(p);
        pl.stride[p] = out->stride[p];
      }
      const int rc = av1h_decode(ctx_, &st.desc, keep ? 1 : 0, reserved, &pl, &info, &slot);
      return finish(rc, slot, nullptr, info, error);
    }
    const int rc = av1h_decode(ctx_, &st.desc, keep ? 1 : 0, reserved, nullptr, &info, &slot);
    return finish(rc, slot, out, info, error);
  }

  int run_planned(void* plan, int slot, Frame* out, bool read_back, std::string* error) {
    last_ok_ = false;
    av1h_frame_info info;
    std::memset(&info, 0, sizeof(info));
    int got = -1;
    const int keep = slot >= 0 ? 1 : 0;
    if (out && read_back) {
      /* the plan's geometry is known once it is decoded: decode, then read
       * the picture (the last decoded one) into the sized frame */
      const int rc0 = av1h_decode_planned(ctx_, plan, keep, slot, nullptr, &info, &got);
      if (rc0 != AV1H_OK) return finish(rc0, -1, nullptr, info, error);
      alloc_like(out, info);
      av1h_planes pl;
      for (int p = 0; p < 3; ++p) {
        pl.plane[p] = out->plane(p);
        pl.stride[p] = out->stride[p];
      }
      const int rc = av1h_read_slot(ctx_, -1, &pl, nullptr);
      return finish(rc, got, nullptr, info, error);
    }
    const int rc = av1h_decode_planned(ctx_, plan, keep, slot, nullptr, &info, &got);
    return finish(rc, got, out, info, error);
  }

  av1h_ctx* ctx_;
  BackendOptions options_;
  DeviceInfo info_;
  DecodeStats stats_;
  std::mutex mu_;
  std::vector<int> held_;
  bool last_ok_ = false;
  std::atomic<bool> hold_{false};
};

}  // namespace

Decoder::~Decoder() = default;

bool Decoder::decode(const FrameSyntax& syntax, Frame* out, std::string* error) {
  const int slot = decode_reference(syntax, out, error);
  if (slot < 0) return false;
  release_reference(slot);
  return true;
}

std::unique_ptr<Decoder> Decoder::create(const BackendOptions& options, std::string* error) {
  std::string local;
  if (!error) error = &local;
  error->clear();
  if (options.kind != BackendKind::Auto && options.kind != BackendKind::HIP) {
    *error = std::string("backend '") + to_string(options.kind) +
             "' is not part of this library (it provides the HIP backend for gfx950)";
    return nullptr;
  }
  av1h_ctx* ctx = nullptr;
  const int rc = av1h_create(options.device_index, &ctx);
  if (rc != AV1H_OK || !ctx) {
    *error = std::string("HIP backend unavailable: ") + status_text(rc);
    return nullptr;
  }
  return std::unique_ptr<Decoder>(new HipDecoder(ctx, options));
}

std::unique_ptr<Decoder> create_best_decoder(const BackendOptions& options, std::string* error) {
  BackendOptions o = options;
  o.kind = BackendKind::HIP;
  return Decoder::create(o, error);
}

}  // namespace av1gpu
