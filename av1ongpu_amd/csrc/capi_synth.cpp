/*
 * C entry points of the workload generator (libav1syn.so), for the Python
 * tests and bench: seeded clips of the measured workloads as wire frames.
 */
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "av1hip.h"
#include "synth.hpp"
#include "syntax_desc.hpp"

using av1gpu::synth::Clip;
using av1gpu::synth::Config;

extern "C" {

void* av1s_config_new(const char* workload, uint64_t seed) {
  try {
    return new Config(av1gpu::synth::workload(workload ? workload : "c1", seed));
  } catch (...) {
    return nullptr;
  }
}

void av1s_config_free(void* cfg) { delete static_cast<Config*>(cfg); }

/* Overrides one field; returns 0, or -1 for an unknown key. */
int av1s_config_set(void* h, const char* key, double v) {
  Config* c = static_cast<Config*>(h);
  const std::string k(key ? key : "");
  const int i = static_cast<int>(v);
  if (k == "width") c->width = i;
  else if (k == "height") c->height = i;
  else if (k == "bit_depth") c->bit_depth = i;
  else if (k == "coded_width") c->coded_width = i;
  else if (k == "frames") c->frames = i;
  else if (k == "num_refs") c->num_refs = i;
  else if (k == "gop") c->gop = i;
  else if (k == "key_interval") c->key_interval = i;
  else if (k == "max_tx") c->max_tx = i;
  else if (k == "rect_tx") c->rect_tx = i != 0;
  else if (k == "tx_ext") c->tx_ext = i != 0;
  else if (k == "mv_gaussian") c->mv_gaussian = i != 0;
  else if (k == "mv_range") c->mv_range = i;
  else if (k == "mv_sigma") c->mv_sigma = v;
  else if (k == "all_filters") c->all_filters = i != 0;
  else if (k == "compound_frac") c->compound_frac = v;
  else if (k == "intra_frac") c->intra_frac = v;
  else if (k == "coef_zero") c->coef_zero = v;
  else if (k == "coef_scale") c->coef_scale = v;
  else if (k == "tx_skip") c->tx_skip = v;
  else if (k == "deblock_lo") c->deblock_lo = i;
  else if (k == "deblock_hi") c->deblock_hi = i;
  else if (k == "cdef") c->cdef = i != 0;
  else if (k == "lr") c->lr = i;
  else if (k == "grain") c->grain = i != 0;
  else if (k == "contract_tools") c->contract_tools = i != 0;
  else if (k == "inter_tools") c->inter_tools = v;
  else if (k == "intrabc") c->intrabc = v;
  else if (k == "avail_auto") c->avail_auto = v;
  else if (k == "residual_frac") c->residual_frac = v;
  else if (k == "lr_log2") c->lr_log2 = i;
  else if (k == "levels") c->levels = v != 0;
  else if (k == "seed") c->seed = static_cast<uint64_t>(v);
  else return -1;
  return 0;
}

struct SynClip {
  Clip clip;
  std::vector<std::vector<uint8_t>> wire;
};

void* av1s_clip_new(const void* cfg) {
  try {
    SynClip* s = new SynClip;
    s->clip = av1gpu::synth::generate(*static_cast<const Config*>(cfg));
    s->wire.resize(s->clip.frames.size());
    return s;
  } catch (...) {
    return nullptr;
  }
}

void av1s_clip_free(void* h) { delete static_cast<SynClip*>(h); }

int av1s_clip_frames(const void* h) {
  return static_cast<int>(static_cast<const SynClip*>(h)->clip.frames.size());
}

int av1s_clip_refs(const void* h, int f, int32_t* out, int cap) {
  const SynClip* s = static_cast<const SynClip*>(h);
  const std::vector<int>& r = s->clip.refs[static_cast<size_t>(f)];
  for (int i = 0; i < static_cast<int>(r.size()) && i < cap; ++i) out[i] = r[static_cast<size_t>(i)];
  return static_cast<int>(r.size());
}

int av1s_clip_last_use(const void* h, int f) {
  return static_cast<const SynClip*>(h)->clip.last_use[static_cast<size_t>(f)];
}

int av1s_clip_key(const void* h, int f) {
  return static_cast<const SynClip*>(h)->clip.key[static_cast<size_t>(f)];
}

/* Wire bytes of frame f; the first call serialises and caches them. */
int64_t av1s_clip_wire(void* h, int f, const uint8_t** out) {
  SynClip* s = static_cast<SynClip*>(h);
  std::vector<uint8_t>& w = s->wire[static_cast<size_t>(f)];
  if (w.empty()) w = av1gpu::detail::to_wire(s->clip.frames[static_cast<size_t>(f)]);
  if (out) *out = w.data();
  return static_cast<int64_t>(w.size());
}

/* Display luma pixels of frame f. */
int64_t av1s_clip_pixels(const void* h, int f) {
  const av1gpu::FrameSyntax& fs = static_cast<const SynClip*>(h)->clip.frames[static_cast<size_t>(f)];
  return static_cast<int64_t>(fs.width) * fs.height;
}

/* A wire frame with its coded levels and quantiser matrices dropped (the
 * coefficient raster left as the only residual): the form a decoder that
 * does not dequantise on the device reads. Returns the byte size written,
 * or the size needed when `out` is null or short, negative on a malformed
 * frame. */
int64_t av1s_wire_without_levels(const uint8_t* wire, int64_t len, uint8_t* out, int64_t cap) {
  av1h_frame_desc d;
  if (av1h_wire_parse(wire, static_cast<size_t>(len), &d) != AV1H_OK) return -1;
  d.num_levels = 0;
  d.levels = nullptr;
  d.num_qm = 0;
  d.qm_pool = nullptr;
  const size_t n = av1h_wire_size(&d);
  if (!out || cap < static_cast<int64_t>(n)) return static_cast<int64_t>(n);
  return static_cast<int64_t>(av1h_wire_write(&d, out));
}

/* Round trip check of the wire format: wire -> FrameSyntax -> wire must give
 * the same bytes. Returns 1 when it does. */
int av1s_wire_roundtrip(const uint8_t* wire, int64_t len) {
  av1h_frame_desc d;
  if (av1h_wire_parse(wire, static_cast<size_t>(len), &d) != AV1H_OK) return 0;
  av1gpu::FrameSyntax fs;
  av1gpu::detail::unflatten(d, &fs);
  const std::vector<uint8_t> again = av1gpu::detail::to_wire(fs);
  return again.size() == static_cast<size_t>(len) && std::memcmp(again.data(), wire, again.size()) == 0;
}

}  // extern "C"

extern "C" uint64_t av1s_fnv1a64(const uint8_t* p, int64_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (int64_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}
