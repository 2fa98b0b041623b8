/*
 * Reconstruction kernels for gfx950: prediction + inverse transform + residual.
 *
 * Per frame, in stream order:
 *   k_mc     - motion compensation of every inter block, one thread per
 *              output sample, straight into the frame;
 *   k_itx<G> - every transform block, grouped by its larger side G: G lanes
 *              per transform. Inter residuals are added into the frame, intra
 *              residuals are kept in a residual plane;
 *   k_queue  - a persistent work queue over the intra blocks: waves take
 *              blocks in coding order, spin on per cell epoch flags until the
 *              cells the block reads are final, predict the block, add its
 *              residual and publish its own cells. This replaces the
 *              reference's host wavefront schedule (src/core/engine.cpp:
 *              695-926): the dependencies are the same cells that scheduler
 *              waits for, but they resolve on the device as soon as each
 *              neighbour is done instead of one global wave at a time.
 *
 * This is the reference CPU backend's stage order inside a wave (inter, intra,
 * transforms, residual add; src/backends/cpu/cpu_backend.cpp:213-328) with
 * the transforms of intra blocks hoisted out of the dependency chain: their
 * residual does not depend on the prediction. The arithmetic is the
 * reference's (kernels/av1gpu_kernels.h), cited per function.
 */
#define AV1T_TABLE_QUAL static __constant__ const
#include "av1hip_internal.hpp"
#include "av1hip_itx.hpp"

namespace av1hip {
namespace {

#define SPEC(off) (av1t_spec_table + (off))

constexpr int kWork = 2048;            /* int32 intra work area (edges, CFL)    */
constexpr int kSpinLimit = 1 << 22;    /* dependency wait bound (then an error) */

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int rnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ int rnd2s(int v, int n) {
  if (n <= 0) return v;
  return v >= 0 ? (v + (1 << (n - 1))) >> n : -((-v + (1 << (n - 1))) >> n);
}
__device__ __forceinline__ int rnd2s_add(int v, int n) { return (v + (1 << (n - 1)) + (v < 0 ? -1 : 0)) >> n; }
__device__ __forceinline__ int ilog2(int v) { return 31 - __builtin_clz(v); }

/* one wave per workgroup: the workgroup barrier is the wave barrier */
__device__ __forceinline__ void wsync() { __syncthreads(); }

__device__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

/* ------------------------------------------------------------------------ */
/* inverse transform of one transform block into the LDS prediction          */
/* ------------------------------------------------------------------------ */

/* row pass (av1gpu_kernels.h:869-904): lane = row */
template <int N>
__device__ void row_pass(int* tmp, int lane, int w, int h, int tx_type, int lossless, int row_shift, int bd,
                         int rect) {
  const int y = lane;
  if (y >= h) return;
  const int cs = w < 32 ? w : 32;
  const int st = w + 1;
  const int cr = (1 << (max(bd + 6, 16) - 1));
  if (y >= 32) {
#pragma unroll
    for (int x = 0; x < N; ++x) tmp[y * st + x] = 0;
    return;
  }
  int t[N];
#pragma unroll
  for (int x = 0; x < N; ++x) t[x] = x < cs ? tmp[y * st + x] : 0;
  if (rect) {
#pragma unroll
    for (int x = 0; x < N; ++x) t[x] = itx::r12((int64_t)t[x] * 2896);
  }
  if (lossless) {
    if constexpr (N == 4) itx::iwht(t, 2);
  } else {
    const int r = bd + 8;
    itx::tx1d<N>(t, itx::row_kind(tx_type), -(1 << (r - 1)), (1 << (r - 1)) - 1);
  }
#pragma unroll
  for (int x = 0; x < N; ++x) {
    const int v = lossless ? t[x] : rnd2(t[x], row_shift);
    tmp[y * st + x] = clampi(v, -cr, cr - 1);
  }
}

/* column pass (av1gpu_kernels.h:907-930): lane = column; hands each residual
 * sample to `out(x, y, value)` */
template <int M, typename Out>
__device__ void col_pass(const int* tmp, int lane, int w, int tx_type, int lossless, int bd, Out&& out) {
  const int x = lane;
  if (x >= w) return;
  const int st = w + 1;
  int t[M];
#pragma unroll
  for (int i = 0; i < M; ++i) t[i] = tmp[i * st + x];
  if (lossless) {
    if constexpr (M == 4) itx::iwht(t, 0);
  } else {
    const int r = max(bd + 6, 16);
    itx::tx1d<M>(t, itx::col_kind(tx_type), -(1 << (r - 1)), (1 << (r - 1)) - 1);
  }
  const int ud = itx::flip_ud(tx_type), lr = itx::flip_lr(tx_type);
  const int xx = lr ? w - 1 - x : x;
#pragma unroll
  for (int i = 0; i < M; ++i) out(xx, ud ? M - 1 - i : i, lossless ? t[i] : rnd2(t[i], 4));
}

/* the row / column pass of the size `l2` (log2 points), instantiated only up
 * to the class size G */
template <int G>
__device__ void row_pass_n(int l2, int* tmp, int lane, int w, int h, int tx_type, int lossless, int row_shift,
                           int bd, int rect) {
  switch (l2) {
    case 2: row_pass<4>(tmp, lane, w, h, tx_type, lossless, row_shift, bd, rect); break;
    case 3: if constexpr (G >= 8) row_pass<8>(tmp, lane, w, h, tx_type, lossless, row_shift, bd, rect); break;
    case 4: if constexpr (G >= 16) row_pass<16>(tmp, lane, w, h, tx_type, lossless, row_shift, bd, rect); break;
    case 5: if constexpr (G >= 32) row_pass<32>(tmp, lane, w, h, tx_type, lossless, row_shift, bd, rect); break;
    default: if constexpr (G >= 64) row_pass<64>(tmp, lane, w, h, tx_type, lossless, row_shift, bd, rect); break;
  }
}
template <int G, typename Out>
__device__ void col_pass_n(int l2, const int* tmp, int lane, int w, int tx_type, int lossless, int bd, Out&& out) {
  switch (l2) {
    case 2: col_pass<4>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 3: if constexpr (G >= 8) col_pass<8>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 4: if constexpr (G >= 16) col_pass<16>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 5: if constexpr (G >= 32) col_pass<32>(tmp, lane, w, tx_type, lossless, bd, out); break;
    default: if constexpr (G >= 64) col_pass<64>(tmp, lane, w, tx_type, lossless, bd, out); break;
  }
}

/* Every transform block of class G (larger side G): G lanes per transform,
 * 64 / G transforms per wave, coefficients staged in LDS between the row and
 * the column pass. The residual of an inter block is added into the frame
 * (its prediction is already there); that of an intra block goes to `res`,
 * where the dependency queue adds it after the prediction. */
template <int G>
__global__ __launch_bounds__(64) void k_itx(ReconArgs a, int cls) {
  constexpr int kPer = 64 / G;
  __shared__ int s_tmp[kPer * G * (G + 1)];
  const int grp = threadIdx.x / G, l = threadIdx.x - grp * G;
  const int idx = blockIdx.x * kPer + grp;
  const bool live = idx < a.plan.itx_n[cls];
  DevTx tx;
  if (live) {
    tx = a.plan.tx[a.plan.itx_list[cls][idx]];
  } else {
    tx.w = tx.h = 4;
    tx.flags = kTxRaw;
    tx.plane = 0;
    tx.x = tx.y = tx.coef = tx.tx_type = tx.row_shift = tx.eob = 0;
  }
  int* tmp = s_tmp + grp * G * (G + 1);
  const int w = tx.w, h = tx.h;
  const int stride = a.lay.stride[tx.plane];
  const int64_t off = a.lay.plane_off[tx.plane] + (int64_t)tx.y * stride + tx.x;
  const bool to_res = (tx.flags & kTxToResidual) != 0;
  const int maxv = a.maxv;
  auto out = [&](int xx, int yy, int v) {
    const int64_t o = off + (int64_t)yy * stride + xx;
    if (to_res)
      a.res[o] = (int16_t)clampi(v, -32768, 32767);
    else
      a.frame[o] = (uint16_t)clampi((int)a.frame[o] + v, 0, maxv);
  };
  const bool coded = live && !(tx.flags & kTxRaw);
  if (live && !coded) {
    /* raw residual (av1gpu_kernels.h:933-937) */
    const int lw = ilog2(w);
    for (int i = l; i < w * h; i += G) out(i & (w - 1), i >> lw, a.plan.residual[tx.coef + i]);
  }
  const int lw = ilog2(w), lh = ilog2(h);
  if (coded) {
    const int cs = w < 32 ? w : 32, cr = h < 32 ? h : 32;
    const int lcs = ilog2(cs), st = w + 1;
    for (int i = l; i < cs * cr; i += G) tmp[(i >> lcs) * st + (i & (cs - 1))] = a.plan.coef[tx.coef + i];
  }
  __syncthreads();
  const int lossless = (tx.flags & kTxLossless) ? 1 : 0;
  if (coded)
    row_pass_n<G>(lw, tmp, l, w, h, tx.tx_type, lossless, tx.row_shift, a.lay.bit_depth,
                  (lw - lh == 1 || lh - lw == 1) ? 1 : 0);
  __syncthreads();
  if (coded) col_pass_n<G>(lh, tmp, l, w, tx.tx_type, lossless, a.lay.bit_depth, out);
}

/* ------------------------------------------------------------------------ */
/* inter prediction                                                          */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ int filt_row(int cls, int extent) {
  if (extent > 4) return cls;
  if (cls == 0 || cls == 2) return 4;
  if (cls == 1) return 5;
  return cls;
}

struct McPos {
  int scaled;
  int bx, by, sub_x, sub_y;
  int ix, iy, fx, fy;          /* unscaled: integer offset and phase      */
  int sx0, sy0, stx, sty;      /* scaled: start and step in 1/1024         */
  __device__ __forceinline__ void col(int x, int* c, int* f) const {
    if (!scaled) {
      *c = bx + x + ix;
      *f = fx;
    } else {
      const int p = sx0 + stx * x;
      *c = p >> 10;
      *f = (p & 1023) >> 6;
    }
  }
  __device__ __forceinline__ void row(int y, int* c, int* f) const {
    if (!scaled) {
      *c = by + y + iy;
      *f = fy;
    } else {
      const int p = sy0 + sty * y;
      *c = p >> 10;
      *f = (p & 1023) >> 6;
    }
  }
};

/* the position arithmetic of av1gpu_kernels.h:1761-1763 (unscaled) and
 * 1819-1840 (scaled, 7.11.3.3) */
__device__ McPos mc_pos(int bx, int by, int mvx, int mvy, int sub_x, int sub_y, int scx, int scy) {
  McPos p;
  p.bx = bx;
  p.by = by;
  p.sub_x = sub_x;
  p.sub_y = sub_y;
  p.scaled = !(scx == (1 << 14) && scy == (1 << 14));
  const int mx = mvx * 2, my = mvy * 2;
  p.ix = mx >> (4 + sub_x);
  p.fx = (mx >> sub_x) & 15;
  p.iy = my >> (4 + sub_y);
  p.fy = (my >> sub_y) & 15;
  const int ox = (bx << 4) + ((2 * mvx) >> sub_x) + 8;
  const int oy = (by << 4) + ((2 * mvy) >> sub_y) + 8;
  p.sx0 = rnd2s_add(ox * scx - (8 << 14), 8) + 32;
  p.sy0 = rnd2s_add(oy * scy - (8 << 14), 8) + 32;
  p.stx = rnd2s_add(scx, 4);
  p.sty = rnd2s_add(scy, 4);
  return p;
}

__device__ const uint16_t* ref_plane(const ReconArgs& a, int slot, int plane) {
  const uint16_t* base = slot >= 0 ? a.pool + (size_t)slot * a.slot_size : a.frame;
  return base + a.lay.plane_off[plane];
}

/* One reference's prediction of sample (x, y) of a block: eight horizontally
 * filtered rows, rounded, then the vertical filter (av1gpu_kernels.h:
 * 1756-1857). Every thread filters its own eight rows; the reference samples
 * its neighbours share come from the cache. */
__device__ __forceinline__ int mc_sample(const uint16_t* ref, int stride, int ref_w, int ref_h, const McPos& P,
                                         int x, int y, const int* hk0, const int* vk0, int round0, int round1) {
  int cx, fx, cy, fy;
  P.col(x, &cx, &fx);
  P.row(y, &cy, &fy);
  const int* hk = hk0 + fx * 8;
  const int* vk = vk0 + fy * 8;
  int cols[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) cols[t] = clampi(cx - 3 + t, 0, ref_w - 1);
  int acc = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const uint16_t* row = ref + (size_t)clampi(cy - 3 + r, 0, ref_h - 1) * stride;
    int s = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) s += (int)row[cols[t]] * hk[t];
    acc += ((s + (1 << (round0 - 1))) >> round0) * vk[r];
  }
  return (acc + (1 << (round1 - 1))) >> round1;
}

/* ------------------------------------------------------------------------ */
/* intra prediction                                                          */
/* ------------------------------------------------------------------------ */

constexpr int kEdge = 160;   /* edge entries (w + h <= 128, upsampled small) */
constexpr int kEdgeOff = 8;

__device__ __forceinline__ int edge_strength(int w, int h, int ft, int delta) {
  const int d = delta < 0 ? -delta : delta, blk = w + h;
  if (ft == 0) {
    if (blk <= 8) return d >= 56 ? 1 : 0;
    if (blk <= 16) return d >= 40 ? 1 : 0;
    if (blk <= 24) return d >= 32 ? 3 : (d >= 16 ? 2 : (d >= 8 ? 1 : 0));
    if (blk <= 32) return d >= 32 ? 3 : (d >= 4 ? 2 : 1);
    return 3;
  }
  if (blk <= 8) return d >= 64 ? 2 : (d >= 40 ? 1 : 0);
  if (blk <= 16) return d >= 48 ? 2 : (d >= 20 ? 1 : 0);
  if (blk <= 24) return d >= 4 ? 3 : 0;
  return 3;
}
__device__ __forceinline__ int upsample_sel(int w, int h, int ft, int delta) {
  const int d = delta < 0 ? -delta : delta;
  if (d <= 0 || d >= 40) return 0;
  return ft == 0 ? (w + h <= 16) : (w + h <= 8);
}

/* 5 tap edge filter over `sz` entries (av1gpu_kernels.h:1157-1173); `buf` is
 * block relative (entry i - 1 is edge sample i) */
__device__ void edge_filter(int* buf, int* tmp, int sz, int strength) {
  const int lane = threadIdx.x;
  if (strength <= 0) return;
  for (int i = lane; i < sz; i += 64) tmp[i] = buf[i - 1];
  wsync();
  const int* k = SPEC(AV1T_OFF_EDGE_KERNEL) + 5 * (strength - 1);
  for (int i = 1 + lane; i < sz; i += 64) {
    int s = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) s += k[t] * tmp[clampi(i - 2 + t, 0, sz - 1)];
    buf[i - 1] = (s + 8) >> 4;
  }
  wsync();
}

/* 2x upsampling (av1gpu_kernels.h:1177-1191) */
__device__ void edge_upsample(int* buf, int* tmp, int n, int maxv) {
  const int lane = threadIdx.x;
  for (int i = lane - 1; i < n; i += 64) tmp[i + 2] = buf[i];
  if (lane == 0) {
    tmp[0] = buf[-1];
    tmp[n + 2] = buf[n - 1];
  }
  wsync();
  if (lane == 0) buf[-2] = tmp[0];
  for (int i = lane; i < n; i += 64) {
    const int s = -tmp[i] + 9 * tmp[i + 1] + 9 * tmp[i + 2] - tmp[i + 3];
    buf[2 * i - 1] = clampi(rnd2(s, 4), 0, maxv);
    buf[2 * i] = tmp[i + 2];
  }
  wsync();
}

/* The whole intra prediction of one block into s_pred (av1gpu_kernels.h:
 * 1578-1677 with the edges of 1108-1151 and the directional preparation of
 * 1452-1481). */
__device__ void predict_intra(const ReconArgs& a, const DevBlock& b, int* work, uint16_t* pred) {
  const int lane = threadIdx.x;
  const int w = b.w, h = b.h, lw = ilog2(w), p = b.plane;
  const int stride = a.lay.stride[p];
  const int bd = a.lay.bit_depth, maxv = a.maxv;
  const uint16_t* src = a.frame + a.lay.plane_off[p] + (size_t)b.y * stride + b.x;
  if (b.flags & kPalette) {
    for (int s = lane; s < w * h; s += 64) {
      const int y = s >> lw, x = s & (w - 1);
      pred[s] = (uint16_t)b.palette[a.plan.palette[b.palette_off + y * b.palette_stride + x]];
    }
    wsync();
    return;
  }
  int* above = work + kEdgeOff;
  int* left = work + kEdgeOff + kEdge + 2 * kEdgeOff;
  int* tmp = work + 2 * (kEdge + 2 * kEdgeOff);
  const int have_left = (b.flags & kHaveLeft) != 0, have_above = (b.flags & kHaveAbove) != 0;
  const int mid = 1 << (bd - 1);
  const int max_x = a.lay.grid_w[p] - 1, max_y = a.lay.grid_h[p] - 1;
  const int n = w + h;
  {
    const int alast = min(max_x, b.x + ((b.flags & kHaveAR) ? 2 * w : w) - 1);
    const int llast = min(max_y, b.y + ((b.flags & kHaveBL) ? 2 * h : h) - 1);
    for (int i = lane; i < n; i += 64) {
      above[i] = have_above ? (int)src[-stride + (min(b.x + i, alast) - b.x)]
                            : (have_left ? (int)src[-1] : mid - 1);
      left[i] = have_left ? (int)src[(ptrdiff_t)(min(b.y + i, llast) - b.y) * stride - 1]
                          : (have_above ? (int)src[-stride] : mid + 1);
    }
    if (lane == 0) {
      int c;
      if (have_above && have_left)
        c = src[-stride - 1];
      else if (have_above)
        c = src[-stride];
      else if (have_left)
        c = src[-1];
      else
        c = mid;
      above[-1] = c;
      left[-1] = c;
    }
  }
  wsync();
  const int mode = b.mode;
  /* DC of the edges that exist (av1gpu_kernels.h:1430-1447) */
  auto dc_value = [&]() {
    int s = 0;
    if (have_above)
      for (int k = lane; k < w; k += 64) s += above[k];
    if (have_left)
      for (int k = lane; k < h; k += 64) s += left[k];
    s = wave_sum(s);
    if (have_above && have_left) return (s + ((w + h) >> 1)) / (w + h);
    if (have_left) return (s + (h >> 1)) >> ilog2(h);
    if (have_above) return (s + (w >> 1)) >> ilog2(w);
    return 1 << (bd - 1);
  };
  if (mode == AV1H_PRED_CFL) {
    /* chroma from luma (av1gpu_kernels.h:1603-1644) */
    const int dc = dc_value();
    const int sx = a.lay.ss_x, sy = a.lay.ss_y;
    const int luma_w = (b.x + w) << sx, luma_h = (b.y + h) << sy;
    const int xlim = luma_w - (1 << sx), ylim = luma_h - (1 << sy);
    const uint16_t* luma = a.frame + a.lay.plane_off[0];
    const int ls = a.lay.stride[0];
    int part = 0;
    for (int s = lane; s < w * h; s += 64) {
      const int y = s >> lw, x = s & (w - 1);
      const int ly = min((b.y + y) << sy, ylim), lx = min((b.x + x) << sx, xlim);
      int t = 0;
      for (int dy = 0; dy <= sy; ++dy)
        for (int dx = 0; dx <= sx; ++dx) t += luma[(size_t)(ly + dy) * ls + lx + dx];
      const int v = t << (3 - sx - sy);
      tmp[s] = v;
      part += v;
    }
    const int sum = wave_sum(part);
    const int shift = ilog2(w) + ilog2(h);
    const int avg = (sum + (1 << (shift - 1))) >> shift;
    for (int s = lane; s < w * h; s += 64)
      pred[s] = (uint16_t)clampi(dc + rnd2s(b.cfl_alpha * (tmp[s] - avg), 6), 0, maxv);
    wsync();
    return;
  }
  if ((b.flags & kFilterIntra) && w >= 4 && h >= 2) {
    /* recursive filter intra (av1gpu_kernels.h:1392-1427), one anti-diagonal
     * of 4x2 sub-blocks at a time: a sub-block reads only the ones above, to
     * its left and above-left */
    const int w4 = w >> 2, h2 = h >> 1;
    const int* taps = SPEC(AV1T_OFF_FILTER_INTRA) + 56 * b.fi_mode;
    for (int d = 0; d < w4 + h2 - 1; ++d) {
      const int c_lo = max(0, d - (h2 - 1)), c_hi = min(d, w4 - 1);
      const int nsub = c_hi - c_lo + 1;
      for (int idx = lane; idx < nsub * 8; idx += 64) {
        const int c4 = c_lo + (idx >> 3), k = idx & 7, r2 = d - c4;
        int pv[7];
        for (int i = 0; i < 5; ++i) {
          if (r2 == 0)
            pv[i] = above[(c4 << 2) + i - 1];
          else if (c4 == 0 && i == 0)
            pv[i] = left[(r2 << 1) - 1];
          else
            pv[i] = pred[((r2 << 1) - 1) * w + (c4 << 2) + i - 1];
        }
        for (int i = 5; i < 7; ++i)
          pv[i] = c4 == 0 ? left[(r2 << 1) + i - 5] : (int)pred[((r2 << 1) + i - 5) * w + (c4 << 2) - 1];
        int acc = 0;
        for (int i = 0; i < 7; ++i) acc += taps[k * 7 + i] * pv[i];
        pred[((r2 << 1) + (k >> 2)) * w + (c4 << 2) + (k & 3)] = (uint16_t)clampi(rnd2(acc, 4), 0, maxv);
      }
      wsync();
    }
    return;
  }
  if (mode >= 1 && mode <= 8) {
    const int pa = SPEC(AV1T_OFF_MODE_TO_ANGLE)[mode] + 3 * b.angle;
    int ua = 0, ul = 0;
    if (b.filter_type >= 0 && pa != 90 && pa != 180) {
      if (pa > 90 && pa < 180 && w + h >= 24) {
        if (lane == 0) {
          const int c = rnd2(left[0] * 5 + above[-1] * 6 + above[0] * 5, 4);
          above[-1] = c;
          left[-1] = c;
        }
        wsync();
      }
      if (have_above)
        edge_filter(above, tmp, min(w, max_x - b.x + 1) + (pa < 90 ? h : 0) + 1,
                    edge_strength(w, h, b.filter_type, pa - 90));
      if (have_left)
        edge_filter(left, tmp, min(h, max_y - b.y + 1) + (pa > 180 ? w : 0) + 1,
                    edge_strength(w, h, b.filter_type, pa - 180));
      ua = upsample_sel(w, h, b.filter_type, pa - 90);
      if (ua) edge_upsample(above, tmp, w + (pa < 90 ? h : 0), maxv);
      ul = upsample_sel(w, h, b.filter_type, pa - 180);
      if (ul) edge_upsample(left, tmp, h + (pa > 180 ? w : 0), maxv);
    }
    /* av1gpu_kernels.h:1245-1313 */
    const int* dr = SPEC(AV1T_OFF_DR_DERIV);
    int dx = 0, dy = 0;
    if (pa < 90)
      dx = dr[pa];
    else if (pa > 90 && pa < 180)
      dx = dr[180 - pa];
    if (pa > 90 && pa < 180)
      dy = dr[pa - 90];
    else if (pa > 180)
      dy = dr[270 - pa];
    for (int s = lane; s < w * h; s += 64) {
      const int i = s >> lw, c = s & (w - 1);
      int v;
      if (pa < 90) {
        const int maxb = (w + h - 1) << ua;
        const int idx = (i + 1) * dx;
        const int base = (idx >> (6 - ua)) + (c << ua);
        const int sh = ((idx << ua) >> 1) & 31;
        v = base < maxb ? rnd2(above[base] * (32 - sh) + above[base + 1] * sh, 5) : above[maxb];
      } else if (pa > 90 && pa < 180) {
        int idx = (c << 6) - (i + 1) * dx;
        int base = idx >> (6 - ua);
        if (base >= -(1 << ua)) {
          const int sh = ((idx * (1 << ua)) >> 1) & 31;
          v = rnd2(above[base] * (32 - sh) + above[base + 1] * sh, 5);
        } else {
          idx = (i << 6) - (c + 1) * dy;
          base = idx >> (6 - ul);
          const int sh = ((idx * (1 << ul)) >> 1) & 31;
          v = rnd2(left[base] * (32 - sh) + left[base + 1] * sh, 5);
        }
      } else if (pa > 180) {
        const int idx = (c + 1) * dy;
        const int base = (idx >> (6 - ul)) + (i << ul);
        const int sh = ((idx << ul) >> 1) & 31;
        v = rnd2(left[base] * (32 - sh) + left[base + 1] * sh, 5);
      } else {
        v = pa == 90 ? above[c] : left[i];
      }
      pred[s] = (uint16_t)clampi(v, 0, maxv);
    }
    wsync();
    return;
  }
  if (mode >= AV1H_PRED_SMOOTH && mode <= AV1H_PRED_SMOOTH_H) {
    /* av1gpu_kernels.h:1333-1353 */
    const int* sw = SPEC(AV1T_OFF_SM_WEIGHTS);
    for (int s = lane; s < w * h; s += 64) {
      const int y = s >> lw, x = s & (w - 1);
      int v;
      if (mode == AV1H_PRED_SMOOTH) {
        const int wy = sw[h - 4 + y], wx = sw[w - 4 + x];
        v = rnd2(wy * above[x] + (256 - wy) * left[h - 1] + wx * left[y] + (256 - wx) * above[w - 1], 9);
      } else if (mode == AV1H_PRED_SMOOTH_V) {
        const int wy = sw[h - 4 + y];
        v = rnd2(wy * above[x] + (256 - wy) * left[h - 1], 8);
      } else {
        const int wx = sw[w - 4 + x];
        v = rnd2(wx * left[y] + (256 - wx) * above[w - 1], 8);
      }
      pred[s] = (uint16_t)clampi(v, 0, maxv);
    }
    wsync();
    return;
  }
  if (mode == AV1H_PRED_PAETH) {
    /* av1gpu_kernels.h:1364-1382 */
    for (int s = lane; s < w * h; s += 64) {
      const int y = s >> lw, x = s & (w - 1);
      const int base = above[x] + left[y] - above[-1];
      const int pl = abs(base - left[y]), pt = abs(base - above[x]), ptl = abs(base - above[-1]);
      const int v = (pl <= pt && pl <= ptl) ? left[y] : (pt <= ptl ? above[x] : above[-1]);
      pred[s] = (uint16_t)clampi(v, 0, maxv);
    }
    wsync();
    return;
  }
  const int dc = clampi(dc_value(), 0, maxv);
  for (int s = lane; s < w * h; s += 64) pred[s] = (uint16_t)dc;
  wsync();
}

/* ------------------------------------------------------------------------ */
/* block driver                                                              */
/* ------------------------------------------------------------------------ */

__device__ void store_block(const ReconArgs& a, const DevBlock& b, const uint16_t* pred) {
  const int p = b.plane, lw = ilog2(b.w);
  const int stride = a.lay.stride[p];
  uint16_t* dst = a.frame + a.lay.plane_off[p] + (size_t)b.y * stride + b.x;
  for (int s = threadIdx.x; s < b.w * b.h; s += 64) dst[(size_t)(s >> lw) * stride + (s & (b.w - 1))] = pred[s];
}

__device__ __forceinline__ int64_t cell_index(const Layout& L, int p, int x, int y) {
  return L.cell_off[p] + (int64_t)(y >> L.cell_shift[p]) * L.cells_w[p] + (x >> L.cell_shift[p]);
}

/* the cells of a block, flagged with this frame's epoch */
__device__ void mark_cells(const ReconArgs& a, const DevBlock& b, bool atomic) {
  const int p = b.plane, sh = a.lay.cell_shift[p];
  const int cw = max(b.w >> sh, 1), ch = max(b.h >> sh, 1);
  for (int i = threadIdx.x; i < cw * ch; i += 64) {
    const int cx = (b.x >> sh) + i % cw, cy = (b.y >> sh) + i / cw;
    if (cx >= a.lay.cells_w[p] || cy >= a.lay.cells_h[p]) continue;
    uint32_t* c = a.cells + a.lay.cell_off[p] + (int64_t)cy * a.lay.cells_w[p] + cx;
    if (atomic)
      __hip_atomic_store(c, a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      *c = a.epoch;
  }
}

/* Wait until every cell the block reads carries this frame's epoch: the left
 * column, the above row and the corner the reference scheduler waits for
 * (engine.cpp:811-875), and for chroma from luma the co-located luma cells
 * (engine.cpp:908-916). */
__device__ void wait_cells(const ReconArgs& a, const DevBlock& b) {
  const Layout& L = a.lay;
  const int p = b.plane, sh = L.cell_shift[p], cell = 1 << sh;
  const int gw = L.grid_w[p], gh = L.grid_h[p];
  int nl = 0, na = 0, nc = 0, ncfl = 0;
  int ly0 = b.y & ~(cell - 1), ax0 = b.x & ~(cell - 1);
  if (b.x - 1 >= 0 && b.dep_y1 >= ly0) nl = (min(b.dep_y1, gh - 1) - ly0) / cell + 1;
  if (b.y - 1 >= 0 && b.dep_x1 >= ax0) na = (min(b.dep_x1, gw - 1) - ax0) / cell + 1;
  if (nl < 0) nl = 0;
  if (na < 0) na = 0;
  if (b.dep_corner && b.x > 0 && b.y > 0) nc = 1;
  int lx0 = 0, ly0l = 0, lcw = 0;
  if (b.mode == AV1H_PRED_CFL && p > 0) {
    const int s0 = L.cell_shift[0];
    lx0 = (b.x << L.ss_x) >> s0;
    ly0l = (b.y << L.ss_y) >> s0;
    const int lx1 = min(((b.x + b.w - 1) << L.ss_x), L.grid_w[0] - 1) >> s0;
    const int ly1 = min(((b.y + b.h - 1) << L.ss_y), L.grid_h[0] - 1) >> s0;
    lcw = lx1 - lx0 + 1;
    ncfl = lcw * (ly1 - ly0l + 1);
  }
  const int total = nl + na + nc + ncfl;
  if (total == 0) return;
  for (int iter = 0;; ++iter) {
    int ok = 1;
    for (int i = threadIdx.x; i < total; i += 64) {
      int64_t idx;
      if (i < nl)
        idx = cell_index(L, p, b.x - 1, ly0 + i * cell);
      else if (i < nl + na)
        idx = cell_index(L, p, ax0 + (i - nl) * cell, b.y - 1);
      else if (i < nl + na + nc)
        idx = cell_index(L, p, b.x - 1, b.y - 1);
      else {
        const int k = i - nl - na - nc;
        idx = L.cell_off[0] + (int64_t)(ly0l + k / lcw) * L.cells_w[0] + lx0 + k % lcw;
      }
      if (__hip_atomic_load(a.cells + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.epoch) ok = 0;
    }
    if (__all(ok)) break;
    if (iter >= kSpinLimit) {
      if (threadIdx.x == 0) atomicOr(a.error, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

/* Intra block: the prediction, then the residuals of its transforms, which
 * k_itx computed before the queue started (av1gpu_kernels.h:933-937 for the
 * add), then one store. */
__device__ void reconstruct_intra(const ReconArgs& a, const DevBlock& b, int* work, uint16_t* pred) {
  predict_intra(a, b, work, pred);
  const int p = b.plane, stride = a.lay.stride[p];
  for (int t = 0; t < b.tx_count; ++t) {
    const DevTx& tx = a.plan.tx[b.tx_first + t];
    const int ox = tx.x - b.x, oy = tx.y - b.y, ltw = ilog2(tx.w);
    const int16_t* r = a.res + a.lay.plane_off[p] + (int64_t)tx.y * stride + tx.x;
    for (int s = threadIdx.x; s < tx.w * tx.h; s += 64) {
      const int yy = s >> ltw, xx = s & (tx.w - 1);
      uint16_t* q = pred + (oy + yy) * b.w + ox + xx;
      *q = (uint16_t)clampi((int)*q + r[(int64_t)yy * stride + xx], 0, a.maxv);
    }
  }
  wsync();
  store_block(a, b, pred);
}

/* Motion compensation of every inter block, one thread per output sample:
 * the samples of the inter blocks are numbered block after block, and a
 * workgroup takes 256 consecutive ones. */
constexpr int kMcChunk = kMcChunkSamples;

__global__ __launch_bounds__(kMcChunk) void k_mc(ReconArgs a) {
  __shared__ int s_filt[768];
  for (int i = threadIdx.x; i < 768; i += kMcChunk) s_filt[i] = av1t_spec_table[AV1T_OFF_SUBPEL + i];
  __syncthreads();
  const int g = blockIdx.x * kMcChunk + threadIdx.x;
  if (g >= a.plan.mc_total) return;
  int i = a.plan.mc_chunk[blockIdx.x];
  while (a.plan.mc_px0[i + 1] <= g) ++i;
  const DevBlock& b = a.plan.blocks[a.plan.inter_list[i]];
  const int s = g - a.plan.mc_px0[i];
  const int w = b.w, lw = ilog2(w), x = s & (w - 1), y = s >> lw;
  const int p = b.plane;
  const int sub_x = p > 0 ? a.lay.ss_x : 0, sub_y = p > 0 ? a.lay.ss_y : 0;
  const int stride = a.lay.stride[p];
  const int bd = a.lay.bit_depth;
  const int compound = b.compound != 0;
  const int round0 = bd >= 12 ? 5 : 3;
  int round1 = compound ? 7 : 11;
  if (bd >= 12 && !compound) round1 -= 2;
  const int* hk0 = s_filt + filt_row(b.filt, w) * 128;
  const int* vk0 = s_filt + filt_row(b.filt_v, b.h) * 128;
  const McPos P0 = mc_pos(b.x, b.y, b.mv[0][0], b.mv[0][1], sub_x, sub_y, b.scale_x, b.scale_y);
  int out = mc_sample(ref_plane(a, b.slot[0], p), stride, b.ref_w, b.ref_h, P0, x, y, hk0, vk0, round0, round1);
  if (compound) {
    /* av1gpu_kernels.h:2207-2293 */
    const McPos P1 = mc_pos(b.x, b.y, b.mv[1][0], b.mv[1][1], sub_x, sub_y, b.scale_x, b.scale_y);
    const int v1 =
        mc_sample(ref_plane(a, b.slot[1], p), stride, b.ref_w, b.ref_h, P1, x, y, hk0, vk0, round0, round1);
    const int post = 2 * 7 - (round0 + 7);
    if (b.compound == AV1H_COMPOUND_AVERAGE)
      out = (out + v1 + (1 << post)) >> (1 + post);
    else
      out = (b.w0 * out + b.w1 * v1 + (1 << (3 + post))) >> (4 + post);
  }
  a.frame[a.lay.plane_off[p] + (int64_t)(b.y + y) * stride + b.x + x] = (uint16_t)clampi(out, 0, a.maxv);
  if (s == 0) {
    /* the block's dependency cells: final once this launch and the
     * transforms after it have run, which the queue launch follows */
    const int sh = a.lay.cell_shift[p];
    const int cw = max(b.w >> sh, 1), ch = max(b.h >> sh, 1);
    for (int k = 0; k < cw * ch; ++k) {
      const int cx = (b.x >> sh) + k % cw, cy = (b.y >> sh) + k / cw;
      if (cx < a.lay.cells_w[p] && cy < a.lay.cells_h[p])
        a.cells[a.lay.cell_off[p] + (int64_t)cy * a.lay.cells_w[p] + cx] = a.epoch;
    }
  }
}

__global__ __launch_bounds__(64) void k_queue(ReconArgs a) {
  __shared__ int s_work[kWork];
  __shared__ uint16_t s_pred[64 * 64];
  for (;;) {
    int idx = 0;
    if (threadIdx.x == 0) idx = (int)atomicAdd(a.queue_counter, 1u);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= a.plan.n_queue) return;
    const int bi = __builtin_amdgcn_readfirstlane(a.plan.queue_list[idx]);
    const DevBlock& b = a.plan.blocks[bi];
    wait_cells(a, b);
    reconstruct_intra(a, b, s_work, s_pred);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    mark_cells(a, b, true);
    wsync();
  }
}

__global__ void k_zero(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

hipError_t launch_recon(const ReconArgs& a, int part, int cus, hipStream_t s, int* launches) {
  if (part == 0) {
    if (a.plan.mc_total > 0) {
      hipLaunchKernelGGL(k_mc, dim3(a.plan.mc_chunks), dim3(kMcChunk), 0, s, a);
      ++*launches;
    }
    const int per[kTxClasses] = {16, 8, 4, 2, 1};
    for (int c = 0; c < kTxClasses; ++c) {
      const int n = a.plan.itx_n[c];
      if (n <= 0) continue;
      const dim3 grid((n + per[c] - 1) / per[c]);
      switch (c) {
        case 0: hipLaunchKernelGGL(k_itx<4>, grid, dim3(64), 0, s, a, c); break;
        case 1: hipLaunchKernelGGL(k_itx<8>, grid, dim3(64), 0, s, a, c); break;
        case 2: hipLaunchKernelGGL(k_itx<16>, grid, dim3(64), 0, s, a, c); break;
        case 3: hipLaunchKernelGGL(k_itx<32>, grid, dim3(64), 0, s, a, c); break;
        default: hipLaunchKernelGGL(k_itx<64>, grid, dim3(64), 0, s, a, c); break;
      }
      ++*launches;
    }
  }
  if (part == 1 && a.plan.n_queue > 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, a.queue_counter, 1);
    const int grid = min(a.plan.n_queue, cus * 6);
    hipLaunchKernelGGL(k_queue, dim3(grid), dim3(64), 0, s, a);
    *launches += 2;
  }
  return hipGetLastError();
}

}  // namespace av1hip
