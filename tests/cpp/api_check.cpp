/*
 * Drives the C++ av1gpu::Decoder API (libav1gpu.so) the way a reference
 * caller does (tools/av1gpu_decode.cpp: decode_reference, release_reference,
 * to_nv12, to_rgb), for the pytest suite:
 *
 *   api_check host
 *       host-side contracts: backend names, Frame geometry, frame_to_nv12 /
 *       frame_to_yuv16 / frame_to_yuv8 bytes, Decoder::create refusals;
 *   api_check decode <workload> <seed> <nv12 pitch> [planned|streamed] [nocoef] [key=value ...]
 *       decodes a synthetic clip on the GPU and prints one line per frame:
 *       "<frame> <slot> <yuv16 fnv> <nv12 fnv> <rgb fnv>"; with `planned`
 *       every other frame goes through plan_frame + decode_reference_planned;
 *       with `streamed` every frame is planned, decoded into a reserved slot
 *       with read_back false (the backend packs on the device) and packed
 *       through nv12_pack_begin / wait / release, the way the reference's
 *       pipelined caller does (dsfilter/decode_core.hpp), yuv16 printed as "-";
 *       with `nocoef` every frame's coefficient raster is dropped, so a clip
 *       with coded levels decodes from its levels only.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "av1gpu/av1gpu.hpp"
#include "synth.hpp"

using namespace av1gpu;

static uint64_t fnv(const std::vector<uint8_t>& v) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (uint8_t b : v) {
    h ^= b;
    h *= 0x100000001b3ull;
  }
  return h;
}

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "%s:%d: CHECK(%s)\n", __FILE__, __LINE__, #c); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

static int host() {
  CHECK(backend_kind_from_string("HIP") == BackendKind::HIP);
  CHECK(backend_kind_from_string("cpu") == BackendKind::Cpu);
  CHECK(backend_kind_from_string("nonsense") == BackendKind::Auto);
  CHECK(std::string(to_string(BackendKind::HIP)) == "hip");
  Frame f(37, 21, 10, 1, 1);
  CHECK(f.plane_width(0) == 37 && f.plane_width(1) == 19 && f.plane_height(1) == 11);
  CHECK(f.stride[0] == 64 && f.stride[1] == 32);
  CHECK(f.storage[0].size() == 64u * 24 && f.storage[1].size() == 32u * 16);
  CHECK(f.matches(37, 21, 10, 1, 1) && !f.matches(38, 21, 10, 1, 1));
  for (int p = 0; p < 3; ++p)
    for (int y = 0; y < f.plane_height(p); ++y)
      for (int x = 0; x < f.plane_width(p); ++x) f.plane(p)[y * f.stride[p] + x] = (uint16_t)((p * 300 + y * 37 + x * 11) & 1023);
  std::vector<uint8_t> nv;
  CHECK(!frame_to_nv12(f, 37, &nv));  /* odd pitch */
  CHECK(!frame_to_nv12(f, 36, &nv));  /* narrower than a chroma row */
  CHECK(frame_to_nv12(f, 40, &nv) && nv.size() == 40u * (21 + 11));
  bool ok = true;
  for (int y = 0; y < 21; ++y)
    for (int x = 0; x < 40; ++x) {
      const int want = x < 37 ? (f.plane(0)[y * f.stride[0] + x] >> 2) : 0;
      ok = ok && nv[y * 40 + x] == want;
    }
  for (int y = 0; y < 11; ++y)
    for (int x = 0; x < 40; ++x) {
      int want = 0;
      if (x < 38) want = f.plane(1 + (x & 1))[y * f.stride[1] + x / 2] >> 2;
      ok = ok && nv[(21 + y) * 40 + x] == want;
    }
  CHECK(ok);
  std::vector<uint8_t> y16, y8;
  frame_to_yuv16(f, &y16);
  frame_to_yuv8(f, &y8);
  CHECK(y16.size() == 2u * (37 * 21 + 2 * 19 * 11) && y8.size() == 37u * 21 + 2 * 19 * 11);
  CHECK(y16[0] == (f.plane(0)[0] & 0xff) && y16[1] == (f.plane(0)[0] >> 8) && y8[1] == (f.plane(0)[1] >> 2));
  Psnr same = frame_psnr(f, f);
  CHECK(same.y == 99.0 && same.u == 99.0);
  std::string err;
  BackendOptions o;
  o.kind = BackendKind::Cpu;
  CHECK(!Decoder::create(o, &err) && !err.empty());
  std::printf("host checks: %d failed\n", fails);
  return fails ? 1 : 0;
}

static int decode(int argc, char** argv) {
  const std::string wl = argv[2];
  const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
  const size_t pitch = std::strtoull(argv[4], nullptr, 10);
  bool planned = false, streamed = false, nocoef = false;
  synth::Config cfg = synth::workload(wl, seed);
  for (int i = 5; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "planned" || a == "streamed") {
      (a == "planned" ? planned : streamed) = true;
      continue;
    }
    if (a == "nocoef") {
      nocoef = true;
      continue;
    }
    const size_t eq = a.find('=');
    const std::string k = a.substr(0, eq);
    const double v = std::atof(a.c_str() + eq + 1);
    if (k == "key_interval") cfg.key_interval = (int)v;
    else if (k == "frames") cfg.frames = (int)v;
    else if (k == "cdef") cfg.cdef = v != 0;
    else if (k == "lr") cfg.lr = (int)v;
    else if (k == "deblock_lo") cfg.deblock_lo = (int)v;
    else if (k == "deblock_hi") cfg.deblock_hi = (int)v;
    else if (k == "grain") cfg.grain = v != 0;
    else if (k == "levels") cfg.levels = v != 0;
    else if (k == "residual_frac") cfg.residual_frac = v;
    else if (k == "avail_auto") cfg.avail_auto = v;
    else {
      std::fprintf(stderr, "unknown override %s\n", k.c_str());
      return 2;
    }
  }
  synth::Clip clip = synth::generate(cfg);
  std::string err;
  BackendOptions o;
  o.kind = BackendKind::HIP;
  auto dec = Decoder::create(o, &err);
  if (!dec) {
    std::fprintf(stderr, "create: %s\n", err.c_str());
    return 3;
  }
  std::map<int, int> slot_of;
  for (size_t f = 0; f < clip.frames.size(); ++f) {
    FrameSyntax s = clip.frames[f];
    if (nocoef) {
      /* a backend that dequantises on the device needs no raster pool */
      if (!dec->dequantises_on_device() || s.levels.empty()) {
        std::fprintf(stderr, "no device dequantisation or no levels\n");
        return 4;
      }
      s.coefficients.clear();
    }
    s.reference_slots.clear();
    for (int r : clip.refs[f]) s.reference_slots.push_back(slot_of[r]);
    Frame out;
    int slot;
    if (streamed) {
      if (!dec->packs_nv12_on_device()) {
        std::fprintf(stderr, "the backend does not pack on the device\n");
        return 4;
      }
      void* plan = nullptr;
      if (!dec->plan_frame(s, &plan, &err)) {
        std::fprintf(stderr, "plan_frame %zu: %s\n", f, err.c_str());
        return 4;
      }
      const int want = dec->reserve_reference();
      slot = dec->decode_reference_planned(plan, want, &out, &err, false);
      if (slot != want || out.width != s.width || out.height != s.height) {
        std::fprintf(stderr, "streamed decode %zu: slot %d, reserved %d, %dx%d (%s)\n", f, slot, want, out.width,
                     out.height, err.c_str());
        return 4;
      }
      if (f == 0 && dec->stats().waves <= 0) {
        std::fprintf(stderr, "a key frame decoded with %d waves\n", dec->stats().waves);
        return 4;
      }
      const int handle = dec->nv12_pack_begin(slot, pitch, &err);
      std::vector<uint8_t> nv, rgb;
      if (handle < 0 || !dec->nv12_pack_wait(handle, &nv, &err)) {
        std::fprintf(stderr, "streamed pack %zu: handle %d (%s)\n", f, handle, err.c_str());
        return 5;
      }
      dec->nv12_pack_release(handle);
      if (!dec->to_rgb(slot, ColorMatrix::Bt709, ColorRange::Limited, &rgb, &err)) {
        std::fprintf(stderr, "rgb %zu: %s\n", f, err.c_str());
        return 5;
      }
      std::printf("%zu %d - %016llx %016llx\n", f, slot, (unsigned long long)fnv(nv), (unsigned long long)fnv(rgb));
    } else if (planned && (f & 1)) {
      void* plan = nullptr;
      if (!dec->plan_frame(s, &plan, &err)) {
        std::fprintf(stderr, "plan_frame %zu: %s\n", f, err.c_str());
        return 4;
      }
      const int want = dec->reserve_reference();
      slot = dec->decode_reference_planned(plan, want, &out, &err, true);
      if (slot != want) {
        std::fprintf(stderr, "planned decode %zu: slot %d, reserved %d (%s)\n", f, slot, want, err.c_str());
        return 4;
      }
    } else {
      slot = dec->decode_reference(s, &out, &err);
    }
    if (slot < 0) {
      std::fprintf(stderr, "decode %zu: %s\n", f, err.c_str());
      return 4;
    }
    if (!streamed) {
      std::vector<uint8_t> y16, nv, rgb;
      frame_to_yuv16(out, &y16);
      if (!dec->to_nv12(slot, pitch, &nv, &err) ||
          !dec->to_rgb(slot, ColorMatrix::Bt709, ColorRange::Limited, &rgb, &err)) {
        std::fprintf(stderr, "pack %zu: %s\n", f, err.c_str());
        return 5;
      }
      std::printf("%zu %d %016llx %016llx %016llx\n", f, slot, (unsigned long long)fnv(y16),
                  (unsigned long long)fnv(nv), (unsigned long long)fnv(rgb));
    }
    const bool keep = clip.last_use[f] > (int)f;
    if (keep)
      slot_of[(int)f] = slot;
    else
      dec->release_reference(slot);
    for (int r : clip.refs[f])
      if (clip.last_use[r] == (int)f) {
        dec->release_reference(slot_of[r]);
        slot_of.erase(r);
      }
  }
  return 0;
}

/* A key frame with film grain whose first luma auto-regressive coefficient
 * lies outside the coded range [-128, 127] is refused with an error; the same
 * frame with the coefficient at the range's ends decodes. */
static int grain_range() {
  synth::Config cfg = synth::workload("c2_mini", 1);
  cfg.frames = 1;
  synth::Clip clip = synth::generate(cfg);
  std::string err;
  BackendOptions o;
  o.kind = BackendKind::HIP;
  auto dec = Decoder::create(o, &err);
  if (!dec) {
    std::fprintf(stderr, "create: %s\n", err.c_str());
    return 3;
  }
  for (const int v : {-129, 128, -128, 127}) {
    FrameSyntax s = clip.frames[0];
    s.reference_slots.clear();
    s.grain.ar_coeffs_y[0] = v;
    Frame out;
    err.clear();
    const int slot = dec->decode_reference(s, &out, &err);
    const bool in_range = v >= -128 && v <= 127;
    if (in_range != (slot >= 0) || (!in_range && err.find("auto-regressive") == std::string::npos)) {
      std::fprintf(stderr, "coefficient %d: slot %d, error '%s'\n", v, slot, err.c_str());
      return 4;
    }
    if (slot >= 0) dec->release_reference(slot);
  }
  std::printf("grain range ok\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "host") return host();
  if (argc >= 2 && std::string(argv[1]) == "grain_range") return grain_range();
  if (argc >= 5 && std::string(argv[1]) == "decode") return decode(argc, argv);
  std::fprintf(stderr, "usage: api_check host | grain_range | decode <workload> <seed> <nv12 pitch> [planned] [key=value...]\n");
  return 2;
}
