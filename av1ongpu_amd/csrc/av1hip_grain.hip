/*
 * Film grain synthesis (specification 7.18.3) on gfx950.
 *
 * The reference synthesises the grain on the host every frame and uploads a
 * full-frame noise image (src/core/grain.cpp:70-139, 6.2 MB at 1080p). Here:
 *   k_grain_template - one workgroup builds the three grain templates
 *     (82x73 luma, 44x38 chroma at 4:2:0): the Gaussian white noise with the
 *     16 bit LFSR jumped ahead per row (so every row starts in parallel), then
 *     the auto-regressive filter as a wavefront over rows (a row lags its
 *     predecessor by four columns, which is all the lag-3 filter reads), and
 *     the per stripe block offsets;
 *   k_grain_apply - the noise sample of every picture sample is derived on
 *     the fly from the templates (the block and stripe overlap blends of
 *     av1gpu_kernels.h:3218-3291 evaluated per sample) and blended in, chroma
 *     before luma inside each tile because chroma reads the luma before grain.
 * The arithmetic is the reference's (av1gpu_kernels.h:3039-3365).
 */
#define AV1T_TABLE_QUAL static __constant__ const
#include "av1hip_internal.hpp"
#include "spec_tables.h"

namespace av1hip {
namespace {

#define SPEC(off) (av1t_spec_table + (off))

constexpr int kLW = 82, kLH = 73;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int grnd(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }

__device__ __forceinline__ unsigned lfsr_step(unsigned r) {
  const unsigned bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1u;
  return (r >> 1) | (bit << 15);
}
/* apply a 16x16 GF(2) matrix given by its columns */
__device__ __forceinline__ unsigned mat_apply(const uint16_t* cols, unsigned v) {
  unsigned r = 0;
  for (int j = 0; j < 16; ++j)
    if (v & (1u << j)) r ^= cols[j];
  return r;
}

/* white noise row `y` of a w wide template (av1gpu_kernels.h:3070-3079),
 * the Gaussian sequence `gs` staged in LDS */
__device__ void noise_row(int16_t* t, int w, int y, unsigned seed, const uint16_t* jump, int shift,
                          const int16_t* gs) {
  unsigned st = seed;
  for (int k = 0; k < y; ++k) st = mat_apply(jump, st);
  for (int x = 0; x < w; ++x) {
    st = lfsr_step(st);
    t[y * w + x] = (int16_t)grnd(gs[(st >> 5) & 2047], shift);
  }
}

/* LDS accesses of one wave complete in issue order; this keeps the compiler
 * from moving a later read above an earlier write and waits for the writes. */
__device__ __forceinline__ void wave_lds_sync() { __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

/* The auto-regressive filter of one template (av1gpu_kernels.h:3087-3122) run
 * by ONE wave as a wavefront: lane l filters row 3 + l (and row 67 + l when the
 * template is taller than 67 rows) and lags the lane above by four columns,
 * which is exactly the lag-3 filter's reach into the previous row, so no two
 * waves ever synchronise. The coefficients stay in registers. */
template <int LAG>
__device__ __forceinline__ void ar_plane(int16_t* pl, int w, int h, const int* coeffs, const int16_t* luma, int ss_x, int ss_y,
                         int shift, int lo, int hi) {
  constexpr int kTaps = (2 * LAG + 1) * LAG + LAG;
  int c[kTaps + 1];
#pragma unroll
  for (int i = 0; i < kTaps; ++i) c[i] = coeffs[i];
  c[kTaps] = luma ? coeffs[kTaps] : 0; /* chroma only: the luma tap */
  const int lane = threadIdx.x & 63;
  const int rows = h - 3, cols = w - 6;
  const int steps = (cols - 1) + 4 * (rows - 1) + 1;
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = lane + 64 * p;
      const int x = s - 4 * r + 3, y = 3 + r;
      if (r < rows && x >= 3 && x < w - 3) {
        int sum = 0, pos = 0;
#pragma unroll
        for (int dr = -LAG; dr <= 0; ++dr)
#pragma unroll
          for (int dc = -LAG; dc <= LAG; ++dc) {
            if (dr == 0 && dc >= 0) break;
            sum += __mul24(pl[(y + dr) * w + x + dc], c[pos]); /* 16 x 8 bit */
            ++pos;
          }
        if (luma) {
          const int lx = ((x - 3) << ss_x) + 3, ly = ((y - 3) << ss_y) + 3;
          int acc = 0;
          for (int i = 0; i <= ss_y; ++i)
            for (int k = 0; k <= ss_x; ++k) acc += luma[(ly + i) * kLW + lx + k];
          sum += __mul24(grnd(acc, ss_x + ss_y), c[kTaps]);
        }
        pl[y * w + x] = (int16_t)clampi(pl[y * w + x] + grnd(sum, shift), lo, hi);
      }
    }
    wave_lds_sync();
  }
}

/* inlined into its callers so that the template pointers stay LDS accesses */
__device__ __forceinline__ void ar_dispatch(int lag, int16_t* pl, int w, int h, const int* coeffs, const int16_t* luma, int ss_x,
                            int ss_y, int shift, int lo, int hi) {
  switch (lag) {
    case 0: ar_plane<0>(pl, w, h, coeffs, luma, ss_x, ss_y, shift, lo, hi); break;
    case 1: ar_plane<1>(pl, w, h, coeffs, luma, ss_x, ss_y, shift, lo, hi); break;
    case 2: ar_plane<2>(pl, w, h, coeffs, luma, ss_x, ss_y, shift, lo, hi); break;
    default: ar_plane<3>(pl, w, h, coeffs, luma, ss_x, ss_y, shift, lo, hi); break;
  }
}

constexpr int kTmplThreads = 192; /* three waves: luma, cb, cr */

/* Templates and block offsets. `tmpl` receives luma (82x73) then cb and cr
 * (cw x ch each); `offs` the 8 bit offset draw of every (stripe, block).
 * Phases: white noise (one row per thread); luma filter on wave 0 while
 * waves 1-2 draw the offsets; cb on wave 1 and cr on wave 2 (both read the
 * finished luma template) while wave 0 stores luma. */
__device__ void grain_template(const DevGrain& g, const uint16_t* jm82, const uint16_t* jm44, int ss_x, int ss_y,
                               int16_t* tmpl, uint8_t* offs, int16_t* s_l, int16_t (*s_c)[kLW * kLH],
                               const int16_t* s_gauss) {
  const int tid = threadIdx.x, wave = tid >> 6;
  const int bd = g.bit_depth;
  const int lo = -(128 << (bd - 8)), hi = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const int cw = ss_x ? 44 : kLW, ch = ss_y ? 38 : kLH;
  const int chroma = g.cfl || g.num_cb > 0 || g.num_cr > 0;
  const unsigned seed = (unsigned)g.seed;
  /* white noise (av1gpu_kernels.h:3070-3079), one row per thread */
  const uint16_t* jc = ss_x ? jm44 : jm82;
  for (int r = tid; r < kLH + 2 * ch; r += kTmplThreads) {
    if (r < kLH) {
      if (g.num_y > 0)
        noise_row(s_l, kLW, r, seed, jm82, g.scale_shift, s_gauss);
      else
        for (int x = 0; x < kLW; ++x) s_l[r * kLW + x] = 0;
    } else if (chroma) {
      const int pl = (r - kLH) / ch, y = (r - kLH) - pl * ch;
      if (pl == 0 ? (g.num_cb > 0 || g.cfl) : (g.num_cr > 0 || g.cfl))
        noise_row(s_c[pl], cw, y, seed ^ (pl == 0 ? 0xb524u : 0x49d8u), jc, g.scale_shift, s_gauss);
      else
        for (int x = 0; x < cw; ++x) s_c[pl][y * cw + x] = 0;
    }
  }
  __syncthreads();
  if (wave == 0) {
    if (g.num_y > 0) ar_dispatch(g.lag, s_l, kLW, kLH, g.ar_y, nullptr, 0, 0, g.shift, lo, hi);
  } else {
    /* per stripe block offsets (av1gpu_kernels.h:3236-3247) */
    for (int st = tid - 64; st < g.stripes; st += kTmplThreads - 64) {
      unsigned r = seed;
      r ^= (unsigned)(((st * 37 + 178) & 255) << 8);
      r ^= (unsigned)((st * 173 + 105) & 255);
      for (int b = 0; b < g.blocks; ++b) {
        r = lfsr_step(r);
        offs[st * g.blocks + b] = (uint8_t)((r >> 8) & 255);
      }
    }
  }
  __syncthreads();
  if (wave == 0) {
    for (int i = tid; i < kLW * kLH; i += 64) tmpl[i] = s_l[i];
  } else if (chroma) {
    const int pl = wave - 1;
    ar_dispatch(g.lag, s_c[pl], cw, ch, pl ? g.ar_cr : g.ar_cb, g.num_y > 0 ? s_l : nullptr, ss_x, ss_y, g.shift,
                lo, hi);
    for (int i = tid & 63; i < cw * ch; i += 64) tmpl[kLW * kLH + pl * cw * ch + i] = s_c[pl][i];
  }
}

/* one workgroup per frame with film grain */
__global__ __launch_bounds__(kTmplThreads) void k_grain_template(BatchDev B) {
  __shared__ int16_t s_l[kLW * kLH];
  __shared__ int16_t s_c[2][kLW * kLH];
  __shared__ int16_t s_gauss[2048];
  {
    /* the Gaussian sequence (12 bit values), every load issued before the
     * LDS writes */
    constexpr int kPer = (2048 + kTmplThreads - 1) / kTmplThreads;
    const int* gs = SPEC(AV1T_OFF_GAUSSIAN);
    int v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + kTmplThreads * u;
      v[u] = i < 2048 ? gs[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + kTmplThreads * u;
      if (i < 2048) s_gauss[i] = (int16_t)v[u];
    }
  }
  __syncthreads();
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKGrainTemplate);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    FArgs a = filter_args(B, find_job(pre, B.njobs, v));
    grain_template(*a.plan.grain, B.jump82, B.jump44, a.lay.ss_x, a.lay.ss_y, a.grain_tmpl, a.grain_offs, s_l, s_c,
                   s_gauss);
    __syncthreads();
  }
}

struct NoiseCtx {
  const int16_t* t;   /* template of this plane */
  const uint8_t* offs;
  int tw, blocks, sx, sy, overlap, lo, hi;
  __device__ __forceinline__ int G(int s, int b, int i, int j) const {
    const int r = offs[s * blocks + b];
    const int ox = r >> 4, oy = r & 15;
    const int gx = sx ? 6 + ox : 9 + 2 * ox, gy = sy ? 6 + oy : 9 + 2 * oy;
    return t[(gy + i) * tw + gx + j];
  }
  /* the stripe buffer after the horizontal overlap blends */
  __device__ __forceinline__ int H(int s, int i, int x) const {
    const int lbw = 5 - sx, bw = 1 << lbw;
    const int b = x >> lbw, j = x & (bw - 1);
    int g = G(s, b, i, j);
    if (overlap && b > 0 && j < (sx ? 1 : 2)) {
      const int old = G(s, b - 1, i, j + bw);
      g = clampi(grnd(sx ? old * 23 + g * 22 : (j == 0 ? old * 27 + g * 17 : old * 17 + g * 27), 5), lo, hi);
    }
    return g;
  }
  /* the noise image sample after the vertical blends */
  __device__ __forceinline__ int N(int x, int y) const {
    const int lbh = 5 - sy, bh = 1 << lbh;
    const int s = y >> lbh, row = y & (bh - 1);
    int v = H(s, row, x);
    if (overlap && s > 0 && row < (sy ? 1 : 2)) {
      const int old = H(s - 1, row + bh, x);
      v = clampi(grnd(sy ? old * 23 + v * 22 : (row == 0 ? old * 27 + v * 17 : old * 17 + v * 27), 5), lo, hi);
    }
    return v;
  }
};

__device__ __forceinline__ int scale_lut(const int* lut, int index, int bd) {
  const int sh = bd - 8;
  const int x = sh > 0 ? index >> sh : index;
  if (bd == 8 || x >= 255) return lut[x > 255 ? 255 : x];
  const int rem = index - (x << sh);
  return lut[x] + grnd((lut[x + 1] - lut[x]) * rem, sh);
}

/* 64x32 luma tiles (32x16 chroma at 4:2:0), 256 threads */
__device__ void grain_tile(FArgs a, int t, int* s_lut) {
  LayRef L = a.lay;
  const DevGrain& g = *a.plan.grain;
  const int16_t* tmpl = a.grain_tmpl;
  const uint8_t* offs = a.grain_offs;
  uint16_t* frame = a.frame;
  const int bd = L.bit_depth;
  const int maxv = (1 << bd) - 1;
  const int lo = -(128 << (bd - 8)), hi = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const int ntx = (L.out_w[0] + 63) / 64;
  const int tx0 = (t % ntx) * 64, ty0 = (t / ntx) * 32;
  const int cw = L.ss_x ? 44 : kLW, ch = L.ss_y ? 38 : kLH;
  const int chroma = g.cfl || g.num_cb > 0 || g.num_cr > 0;
  const uint16_t* luma = frame + L.plane_off[0];
  const int ls = L.stride[0];
  /* the three scaling tables in LDS (indexed by each sample's value) */
  for (int i = threadIdx.x; i < 768; i += 256) s_lut[i] = g.lut[i];
  __syncthreads();
  /* Each thread's samples (8 luma, up to 8 chroma per plane) go in three
   * batches: every load first (samples, luma, noise template), then the
   * arithmetic, then the stores, so no load waits behind another sample's
   * store. */
  constexpr int kPer = 64 * 32 / 256;
  for (int pl = 1; pl < 3 && chroma; ++pl) {
    if (!g.cfl && (pl == 1 ? g.num_cb : g.num_cr) <= 0) continue;
    NoiseCtx nc{tmpl + kLW * kLH + (pl - 1) * cw * ch, offs, cw, g.blocks, L.ss_x, L.ss_y, g.overlap, lo, hi};
    const int pw = L.out_w[pl], ph = L.out_h[pl];
    const int x0 = tx0 >> L.ss_x, y0 = ty0 >> L.ss_y;
    const int tw = 64 >> L.ss_x, th = 32 >> L.ss_y;
    const int* lut = s_lut + 256 * pl;
    const int mult = pl == 1 ? g.cb_mult : g.cr_mult;
    const int lmult = pl == 1 ? g.cb_luma_mult : g.cr_luma_mult;
    const int offset = pl == 1 ? g.cb_offset : g.cr_offset;
    uint16_t* plane = frame + L.plane_off[pl];
    const int st = L.stride[pl];
    const int lw = L.out_w[0];
    const int ltw = 6 - L.ss_x;
    int orig[kPer], avg[kPer], nv[kPer];
    bool ok[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 256 * u;
      const int x = x0 + (i & (tw - 1)), y = y0 + (i >> ltw);
      ok[u] = i < tw * th && x < pw && y < ph;
      orig[u] = avg[u] = 0;
      if (ok[u]) {
        orig[u] = plane[(size_t)y * st + x];
        const int lx = x << L.ss_x, ly = y << L.ss_y;
        const int nx = lx + 1 < lw ? lx + 1 : lw - 1;
        const int l0 = luma[(size_t)ly * ls + lx];
        avg[u] = L.ss_x ? grnd(l0 + luma[(size_t)ly * ls + nx], 1) : l0;
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 256 * u;
      nv[u] = ok[u] ? nc.N(x0 + (i & (tw - 1)), y0 + (i >> ltw)) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (!ok[u]) continue;
      const int i = threadIdx.x + 256 * u;
      const int x = x0 + (i & (tw - 1)), y = y0 + (i >> ltw);
      int index;
      if (g.cfl) {
        index = avg[u];
      } else {
        const int comb = avg[u] * (lmult - 128) + orig[u] * (mult - 128);
        index = clampi((comb >> 6) + (offset - 256) * (1 << (bd - 8)), 0, maxv);
      }
      const int gv = grnd(scale_lut(lut, index, bd) * nv[u], g.scaling_shift);
      plane[(size_t)y * st + x] = (uint16_t)clampi(orig[u] + gv, g.min_value, g.max_chroma);
    }
  }
  __syncthreads();
  if (g.num_y > 0) {
    NoiseCtx nc{tmpl, offs, kLW, g.blocks, 0, 0, g.overlap, lo, hi};
    uint16_t* plane = frame + L.plane_off[0];
    int orig[kPer], nv[kPer];
    bool ok[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 256 * u;
      const int x = tx0 + (i & 63), y = ty0 + (i >> 6);
      ok[u] = x < L.out_w[0] && y < L.out_h[0];
      orig[u] = ok[u] ? plane[(size_t)y * ls + x] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 256 * u;
      nv[u] = ok[u] ? nc.N(tx0 + (i & 63), ty0 + (i >> 6)) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (!ok[u]) continue;
      const int i = threadIdx.x + 256 * u;
      const int gv = grnd(scale_lut(s_lut, orig[u], bd) * nv[u], g.scaling_shift);
      plane[(size_t)(ty0 + (i >> 6)) * ls + tx0 + (i & 63)] = (uint16_t)clampi(orig[u] + gv, g.min_value, g.max_luma);
    }
  }
}

__global__ __launch_bounds__(256) void k_grain_apply(BatchDev B) {
  __shared__ int s_lut[768];
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKGrainApply);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    grain_tile(filter_args(B, j), v - pre[j], s_lut);
    __syncthreads();
  }
}

unsigned host_step(unsigned r) {
  const unsigned bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1u;
  return (r >> 1) | (bit << 15);
}

}  // namespace

void grain_jump_matrices(uint16_t m82[16], uint16_t m44[16]) {
  for (int j = 0; j < 16; ++j) {
    unsigned a = 1u << j, b = 1u << j;
    for (int k = 0; k < 82; ++k) a = host_step(a);
    for (int k = 0; k < 44; ++k) b = host_step(b);
    m82[j] = (uint16_t)a;
    m44[j] = (uint16_t)b;
  }
}

int grain_tiles(const Layout& lay) { return (lay.out_w[0] + 63) / 64 * ((lay.out_h[0] + 31) / 32); }

hipError_t launch_grain_kernel(const BatchDev& b, int k, int total, int cap, hipStream_t s) {
  const dim3 grid(min(total, cap));
  switch (k) {
    case kKGrainTemplate: hipLaunchKernelGGL(k_grain_template, grid, dim3(kTmplThreads), 0, s, b); break;
    case kKGrainApply: hipLaunchKernelGGL(k_grain_apply, grid, dim3(256), 0, s, b); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace av1hip
