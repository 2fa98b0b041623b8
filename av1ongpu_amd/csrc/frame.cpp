/*
 * Frame and the host picture helpers of the av1gpu API (reference
 * src/core/factory.cpp:16-278): backend names, device enumeration, Frame
 * allocation, PSNR, YUV and NV12 packing of a host Frame. The GPU path packs
 * NV12 and RGB on the device (av1h_to_nv12 / av1h_to_rgb); these are the
 * reference's host utilities over a Frame already on the host.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "av1gpu/av1gpu.hpp"
#include "av1hip.h"

namespace av1gpu {

const char* to_string(BackendKind kind) {
  switch (kind) {
    case BackendKind::Cpu: return "cpu";
    case BackendKind::OpenCL: return "opencl";
    case BackendKind::CUDA: return "cuda";
    case BackendKind::HIP: return "hip";
    case BackendKind::Auto:
    default: return "auto";
  }
}

BackendKind backend_kind_from_string(const std::string& name) {
  std::string s;
  for (char c : name) s.push_back(static_cast<char>(c >= 'A' && c <= 'Z' ? c - 'A' + 'a' : c));
  if (s == "cpu" || s == "reference") return BackendKind::Cpu;
  if (s == "opencl" || s == "cl") return BackendKind::OpenCL;
  if (s == "cuda" || s == "nvidia") return BackendKind::CUDA;
  if (s == "hip" || s == "rocm" || s == "amd" || s == "gpu") return BackendKind::HIP;
  return BackendKind::Auto;
}

std::vector<DeviceInfo> enumerate_devices() {
  std::vector<DeviceInfo> out;
  for (int i = 0;; ++i) {
    av1h_device_info d;
    const int rc = av1h_describe_device(i, &d);
    if (rc == AV1H_ERR_NO_DEVICE || rc == AV1H_ERR_INVALID) break;
    DeviceInfo info;
    info.kind = BackendKind::HIP;
    info.name = d.name;
    info.vendor = "AMD";
    info.version = d.arch;
    info.detail = std::to_string(d.compute_units) + " CUs";
    info.available = rc == AV1H_OK;
    info.global_mem = static_cast<uint64_t>(d.global_mem);
    info.compute_units = d.compute_units;
    out.push_back(info);
  }
  return out;
}

std::string build_info() {
  return "AV1onGPU-compatible reconstruction library, MI355X build\n"
         "  HIP backend: gfx950 (reconstruction queue, loop filters, superres, restoration, film grain, "
         "NV12/RGB pack on the device)\n"
         "  CPU, OpenCL and CUDA backends: not part of this library\n";
}

int Frame::plane_width(int p) const { return p == 0 ? width : (ss_x ? (width + 1) / 2 : width); }
int Frame::plane_height(int p) const { return p == 0 ? height : (ss_y ? (height + 1) / 2 : height); }

void Frame::allocate(int w, int h, int bd, int sx, int sy) {
  width = w;
  height = h;
  bit_depth = bd;
  ss_x = sx;
  ss_y = sy;
  for (int p = 0; p < 3; ++p) {
    /* stride rounded to 32 samples, rows to 8 (factory.cpp:91-110) */
    stride[p] = (plane_width(p) + 31) & ~31;
    const int rows = (plane_height(p) + 7) & ~7;
    storage[p].assign(static_cast<size_t>(stride[p]) * rows, 0);
  }
}

bool Frame::matches(int w, int h, int bd, int sx, int sy) const {
  if (width != w || height != h || bit_depth != bd || ss_x != sx || ss_y != sy) return false;
  for (int p = 0; p < 3; ++p)
    if (storage[p].size() < static_cast<size_t>(stride[p]) * plane_height(p)) return false;
  return true;
}

void Frame::fill(uint16_t value) {
  for (auto& s : storage) std::fill(s.begin(), s.end(), value);
}

Psnr frame_psnr(const Frame& a, const Frame& b) {
  Psnr r;
  if (a.width != b.width || a.height != b.height || a.bit_depth != b.bit_depth) return r;
  const double peak = static_cast<double>((1 << a.bit_depth) - 1);
  double v[3];
  for (int p = 0; p < 3; ++p) {
    const int w = a.plane_width(p), h = a.plane_height(p);
    long double sse = 0;
    for (int y = 0; y < h; ++y) {
      const uint16_t* pa = a.plane(p) + static_cast<size_t>(y) * a.stride[p];
      const uint16_t* pb = b.plane(p) + static_cast<size_t>(y) * b.stride[p];
      for (int x = 0; x < w; ++x) {
        const double d = static_cast<double>(pa[x]) - static_cast<double>(pb[x]);
        sse += d * d;
      }
    }
    const double mse = static_cast<double>(sse) / (static_cast<double>(w) * h);
    v[p] = mse <= 1e-12 ? 99.0 : 10.0 * std::log10(peak * peak / mse);
  }
  r.y = v[0];
  r.u = v[1];
  r.v = v[2];
  return r;
}

void frame_to_yuv16(const Frame& f, std::vector<uint8_t>* out) {
  size_t n = 0;
  for (int p = 0; p < 3; ++p) n += static_cast<size_t>(f.plane_width(p)) * f.plane_height(p);
  out->resize(n * 2);
  uint8_t* d = out->data();
  for (int p = 0; p < 3; ++p)
    for (int y = 0; y < f.plane_height(p); ++y) {
      const uint16_t* row = f.plane(p) + static_cast<size_t>(y) * f.stride[p];
      for (int x = 0; x < f.plane_width(p); ++x) {
        *d++ = static_cast<uint8_t>(row[x] & 0xff);
        *d++ = static_cast<uint8_t>(row[x] >> 8);
      }
    }
}

void frame_to_yuv8(const Frame& f, std::vector<uint8_t>* out) {
  size_t n = 0;
  for (int p = 0; p < 3; ++p) n += static_cast<size_t>(f.plane_width(p)) * f.plane_height(p);
  out->resize(n);
  const int shift = f.bit_depth - 8;
  uint8_t* d = out->data();
  for (int p = 0; p < 3; ++p)
    for (int y = 0; y < f.plane_height(p); ++y) {
      const uint16_t* row = f.plane(p) + static_cast<size_t>(y) * f.stride[p];
      for (int x = 0; x < f.plane_width(p); ++x) *d++ = static_cast<uint8_t>(row[x] >> shift);
    }
}

/* NV12 of a host Frame: the luma rows, then interleaved U/V rows, samples
 * shifted down to 8 bits, padding bytes zero (kernels/av1gpu_kernels.h:3432-3472
 * for the byte layout; factory.cpp:236-278 for the size rule). */
bool frame_to_nv12(const Frame& f, size_t pitch, std::vector<uint8_t>* out) {
  const size_t bytes = av1h_nv12_size(f.width, f.height, f.ss_x, f.ss_y, pitch);
  if (bytes == 0 || !out) return false;
  out->assign(bytes, 0);
  const int shift = f.bit_depth - 8;
  for (int y = 0; y < f.height; ++y) {
    const uint16_t* src = f.plane(0) + static_cast<size_t>(y) * f.stride[0];
    uint8_t* d = out->data() + static_cast<size_t>(y) * pitch;
    for (int x = 0; x < f.width; ++x) d[x] = static_cast<uint8_t>(src[x] >> shift);
  }
  const int cw = f.plane_width(1), ch = f.plane_height(1);
  for (int y = 0; y < ch; ++y) {
    const uint16_t* u = f.plane(1) + static_cast<size_t>(y) * f.stride[1];
    const uint16_t* v = f.plane(2) + static_cast<size_t>(y) * f.stride[2];
    uint8_t* d = out->data() + static_cast<size_t>(f.height + y) * pitch;
    for (int x = 0; x < cw; ++x) {
      d[2 * x] = static_cast<uint8_t>(u[x] >> shift);
      d[2 * x + 1] = static_cast<uint8_t>(v[x] >> shift);
    }
  }
  return true;
}

}  // namespace av1gpu
