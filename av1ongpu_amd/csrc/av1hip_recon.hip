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
#include "scan_order.hpp"

namespace av1hip {
/* the diagonal scans (scan_order.hpp), in device memory */
__device__ const av1scan::DiagTables kScanTab = av1scan::make_diag();

namespace {

#define SPEC(off) (av1t_spec_table + (off))

constexpr int kWork = 2048;            /* int32 intra work area (edges, CFL)    */

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int rnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ int rnd2s(int v, int n) {
  if (n <= 0) return v;
  return v >= 0 ? (v + (1 << (n - 1))) >> n : -((-v + (1 << (n - 1))) >> n);
}
__device__ __forceinline__ int rnd2s_add(int v, int n) { return (v + (1 << (n - 1)) + (v < 0 ? -1 : 0)) >> n; }
__device__ __forceinline__ int ilog2(int v) { return 31 - __builtin_clz(v); }

/* Samples other waves wrote in this launch are exchanged at device scope:
 * the stores and loads go to the level every XCD sees (the queue publishes a
 * block's cells only after its stores completed), so no cache writeback or
 * invalidate is needed around the dependency flags. */
__device__ __forceinline__ int ld_dev(const uint16_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint16_t* p, uint16_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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
__device__ __forceinline__ void row_pass(int* tmp, int lane, int w, int h, int tx_type, int lossless, int row_shift, int bd,
                         int rect) {
  const int y = lane;
  if (y >= h) return;
  /* a row holds min(N, 32) coefficients (AV1 codes none beyond 32): the
   * other inputs are constant zeros, which folds the first stages of the
   * 64-point transform */
  constexpr int kCs = N < 32 ? N : 32;
  const int st = w + 1;
  const int cr = (1 << (max(bd + 6, 16) - 1));
  if (y >= 32) {
#pragma unroll
    for (int x = 0; x < N; ++x) tmp[y * st + x] = 0;
    return;
  }
  int t[N];
#pragma unroll
  for (int x = 0; x < N; ++x) t[x] = x < kCs ? tmp[y * st + x] : 0;
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
__device__ __forceinline__ void col_pass(const int* tmp, int lane, int w, int tx_type, int lossless, int bd, Out&& out) {
  const int x = lane;
  if (x >= w) return;
  const int st = w + 1;
  int t[M];
  /* rows from 32 on are zero after the row pass (no coefficients there) */
#pragma unroll
  for (int i = 0; i < M; ++i) t[i] = i < 32 ? tmp[i * st + x] : 0;
  if (lossless) {
    if constexpr (M == 4) itx::iwht(t, 0);
  } else {
    const int r = max(bd + 6, 16);
    itx::tx1d<M>(t, itx::col_kind(tx_type), -(1 << (r - 1)), (1 << (r - 1)) - 1);
  }
  const int ud = itx::flip_ud(tx_type), lr = itx::flip_lr(tx_type);
  const int xx = lr ? w - 1 - x : x;
  /* up to 32 rows at a time: their current samples (an inter block's
   * prediction) are all loaded before any is stored, so the adds do not wait
   * on one load after another */
  constexpr int kB = M < 32 ? M : 32;
#pragma unroll
  for (int i0 = 0; i0 < M; i0 += kB) {
    int cur[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) cur[u] = out.load(xx, ud ? M - 1 - (i0 + u) : i0 + u);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = i0 + u;
      out.store(xx, ud ? M - 1 - i : i, cur[u], lossless ? t[i] : rnd2(t[i], 4));
    }
  }
}

/* the row / column pass of the size `l2` (log2 points), instantiated only up
 * to the class size G */
template <int G>
__device__ __forceinline__ void row_pass_n(int l2, int* tmp, int lane, int w, int h, int tx_type, int lossless, int row_shift,
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
__device__ __forceinline__ void col_pass_n(int l2, const int* tmp, int lane, int w, int tx_type, int lossless, int bd, Out&& out) {
  switch (l2) {
    case 2: col_pass<4>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 3: if constexpr (G >= 8) col_pass<8>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 4: if constexpr (G >= 16) col_pass<16>(tmp, lane, w, tx_type, lossless, bd, out); break;
    case 5: if constexpr (G >= 32) col_pass<32>(tmp, lane, w, tx_type, lossless, bd, out); break;
    default: if constexpr (G >= 64) col_pass<64>(tmp, lane, w, tx_type, lossless, bd, out); break;
  }
}

/* The transform blocks of class G (larger side G): G lanes per transform,
 * 64 / G transforms per wave, coefficients staged in LDS between the row and
 * the column pass. The residual of an inter block is added into the frame
 * (its prediction is already there); that of an intra block goes to `res`,
 * where the dependency queue adds it after the prediction. */
template <int G>
__device__ __forceinline__ void itx_group(RArgs a, const int32_t* list, int n, int wg, int* s_tmp) {
  constexpr int kPer = 64 / G;
  const int grp = threadIdx.x / G, l = threadIdx.x - grp * G;
  const int idx = wg * kPer + grp;
  const bool live = idx < n;
  DevTx tx;
  if (live) {
    tx = a.plan.tx[list[idx]];
  } else {
    tx.w = tx.h = 4;
    tx.flags = kTxRaw;
    tx.plane = 0;
    tx.x = tx.y = tx.coef = tx.tx_type = tx.row_shift = tx.eob = 0;
  }
  int* tmp = s_tmp + grp * G * (G + 1);
  const int w = tx.w, h = tx.h;
  const int stride = plane_vals(a.lay.stride)[tx.plane];
  const int64_t poff = plane_vals(a.lay.plane_off)[tx.plane];
  const int64_t off = poff + (int64_t)tx.y * stride + tx.x;
  const bool to_res = (tx.flags & kTxToResidual) != 0;
  const int maxv = a.maxv;
  struct {
    RArgs a;
    int64_t off;
    int stride, maxv;
    bool to_res;
    __device__ __forceinline__ int load(int xx, int yy) const {
      return to_res ? 0 : (int)a.frame[off + (int64_t)yy * stride + xx];
    }
    __device__ __forceinline__ void store(int xx, int yy, int cur, int v) const {
      const int64_t o = off + (int64_t)yy * stride + xx;
      if (to_res)
        a.res[o] = (int16_t)clampi(v, -32768, 32767);
      else
        a.frame[o] = (uint16_t)clampi(cur + v, 0, maxv);
    }
  } out{a, off, stride, maxv, to_res};
  const bool coded = live && !(tx.flags & kTxRaw);
  const int lw = ilog2(w), lh = ilog2(h);
  if (live && !coded) {
    /* raw residual (av1gpu_kernels.h:933-937), eight samples per lane at a
     * time: residuals and current samples loaded before any store */
    const int n = w * h;
    for (int i0 = l; i0 < n; i0 += 8 * G) {
      int r[8], cur[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * G;
        r[u] = i < n ? a.plan.residual[tx.coef + i] : 0;
        cur[u] = i < n ? out.load(i & (w - 1), i >> lw) : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * G;
        if (i < n) out.store(i & (w - 1), i >> lw, cur[u], r[u]);
      }
    }
  }
  if (coded) {
    /* every coefficient a lane stages is loaded before the first LDS write:
     * one memory round trip instead of one per coefficient row */
    constexpr int kC = G < 32 ? G : 32, kIt = kC * kC / G;
    const int cs = w < 32 ? w : 32, cr = h < 32 ? h : 32;
    const int lcs = ilog2(cs), st = w + 1, n = cs * cr;
    int v[kIt];
    if (a.plan.levels) {
      /* dequantised here from the coded levels: raster position i holds
       * scan index si, coded when below the transform's count */
      const DevDq q = a.plan.dq[list[idx]];
      const int cls = av1scan::tx_class(tx.tx_type), lcw = lcs - 2, lch = ilog2(cr) - 2;
      const int count = tx.eob < n ? tx.eob : n;
      const int16_t* lv = a.plan.levels + q.level_off;
      int si[kIt];
#pragma unroll
      for (int k = 0; k < kIt; ++k) {
        const int i = l + k * G;
        si[k] = i < n ? av1scan::scan_index(kScanTab.iscan, cls, lcw, lch, i) : n;
      }
#pragma unroll
      for (int k = 0; k < kIt; ++k) v[k] = si[k] < count ? (int)lv[si[k]] : 0;
      if (q.qm_off >= 0) {
        const uint8_t* qm = a.plan.qm + q.qm_off;
#pragma unroll
        for (int k = 0; k < kIt; ++k) {
          const int i = l + k * G;
          if (v[k]) v[k] = av1scan::dequant(v[k], i ? q.ac_step : q.dc_step, qm[i], q.denom_shift, a.lay.bit_depth);
        }
      } else {
#pragma unroll
        for (int k = 0; k < kIt; ++k) {
          const int i = l + k * G;
          if (v[k]) v[k] = av1scan::dequant(v[k], i ? q.ac_step : q.dc_step, -1, q.denom_shift, a.lay.bit_depth);
        }
      }
    } else if (a.plan.coef16) {
      const int16_t* src = static_cast<const int16_t*>(a.plan.coef) + tx.coef;
#pragma unroll
      for (int k = 0; k < kIt; ++k) v[k] = l + k * G < n ? (int)src[l + k * G] : 0;
    } else {
      const int32_t* src = static_cast<const int32_t*>(a.plan.coef) + tx.coef;
#pragma unroll
      for (int k = 0; k < kIt; ++k) v[k] = l + k * G < n ? src[l + k * G] : 0;
    }
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
      const int i = l + k * G;
      if (i < n) tmp[(i >> lcs) * st + (i & (cs - 1))] = v[k];
    }
  }
  __syncthreads();
  const int lossless = (tx.flags & kTxLossless) ? 1 : 0;
  if (coded)
    row_pass_n<G>(lw, tmp, l, w, h, tx.tx_type, lossless, tx.row_shift, a.lay.bit_depth,
                  (lw - lh == 1 || lh - lw == 1) ? 1 : 0);
  __syncthreads();
  if (coded) col_pass_n<G>(lh, tmp, l, w, tx.tx_type, lossless, a.lay.bit_depth, out);
}

/* The classes [C0, C1) of a set in one launch, the larger first: two
 * launches (4-16 and 32-64) so that the small transforms, most of them, run
 * with the registers and LDS they need rather than the 64-point ones'. */
template <int C0, int C1>
__device__ __forceinline__ void itx_body(RArgs a, const int32_t* offs, int b, int* s_tmp) {
#pragma unroll
  for (int c = C1 - 1; c >= C0; --c) {
    const int o0 = offs[c], n = offs[c + 1] - o0;
    const int per = 16 >> c;
    const int wgs = (n + per - 1) / per;
    if (b < wgs) {
      const int32_t* list = a.plan.itx_list + o0;
      switch (c) {
        case 0: itx_group<4>(a, list, n, b, s_tmp); break;
        case 1: itx_group<8>(a, list, n, b, s_tmp); break;
        case 2: itx_group<16>(a, list, n, b, s_tmp); break;
        case 3: itx_group<32>(a, list, n, b, s_tmp); break;
        default: itx_group<64>(a, list, n, b, s_tmp); break;
      }
      return;
    }
    b -= wgs;
  }
}

template <int C0, int C1>
__global__ __launch_bounds__(64) void k_itx(BatchDev B, int set) {
  constexpr int kG = 4 << (C1 - 1);
  __shared__ int s_tmp[64 * (kG + 1)];
  const int kid = C0 == 0 ? (set ? kKItxIntra : kKItxInter) : (set ? kKItxIntraL : kKItxInterL);
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kid);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    RArgs a = recon_args(B, j);
    itx_body<C0, C1>(a, a.plan.itx_offs + set * (kTxClasses + 1), v - pre[j], s_tmp);
    __syncthreads();
  }
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

__device__ const uint16_t* ref_plane(RArgs a, int slot, int plane) {
  const uint16_t* base = slot >= 0 ? a.pool + (size_t)slot * a.slot_size : a.frame;
  return base + a.lay.plane_off[plane];
}

/* ------------------------------------------------------------------------ */
/* the remaining inter tools, one sample at a time (av1gpu_kernels.h:        */
/* 1880-2326): global warp, wedge and difference weighted masks, OBMC,       */
/* inter/intra, intra block copy. They are rare, so they are computed the    */
/* direct way; samples of the frame being decoded are read at device scope.  */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ int fetch_px(const uint16_t* p, bool cur, int stride, int x, int y, int x1, int y1) {
  x = clampi(x, 0, x1);
  y = clampi(y, 0, y1);
  const uint16_t* q = p + (size_t)y * stride + x;
  return cur ? ld_dev(q) : (int)*q;
}

/* one sample of one reference through the two stage filter, unscaled or
 * scaled (av1gpu_kernels.h:1756-1857) */
__device__ int mc_tap(const uint16_t* ref, bool cur, int stride, int ref_w, int ref_h, int sub_x, int sub_y,
                      int scx, int scy, int bx, int by, int ew, int eh, int mvx, int mvy, int filt, int filt_v,
                      int x, int y, int compound, int bd) {
  const int round0 = bd >= 12 ? 5 : 3;
  int round1 = compound ? 7 : 11;
  if (bd >= 12 && !compound) round1 -= 2;
  const McPos P = mc_pos(bx, by, mvx, mvy, sub_x, sub_y, scx, scy);
  int cx, fx, cy, fy;
  P.col(x, &cx, &fx);
  P.row(y, &cy, &fy);
  const int* hk = SPEC(AV1T_OFF_SUBPEL) + (filt_row(filt, ew) * 16 + fx) * 8;
  const int* vk = SPEC(AV1T_OFF_SUBPEL) + (filt_row(filt_v, eh) * 16 + fy) * 8;
  int acc = 0;
  for (int r = 0; r < 8; ++r) {
    int s = 0;
    for (int t = 0; t < 8; ++t) s += fetch_px(ref, cur, stride, cx - 3 + t, cy - 3 + r, ref_w - 1, ref_h - 1) * hk[t];
    acc += ((s + (1 << (round0 - 1))) >> round0) * vk[r];
  }
  return (acc + (1 << (round1 - 1))) >> round1;
}

__device__ __forceinline__ int warp_offs(int v) { return clampi(((v + 512) >> 10) + 64, 0, 192); }

/* one sample of a warped reference (av1gpu_kernels.h:1919-1958) */
__device__ int warp_tap(int sub_x, int sub_y, int ref_w, int ref_h, const AV1H_CMEM int* wp, const AV1H_CMEM int* abcd,
                        const uint16_t* ref, bool cur, int stride, int bx, int by, int x, int y, int bd,
                        int compound) {
  const int* wf = SPEC(AV1T_OFF_WARP);
  const int i1 = (y & 7) - 4, i2 = (x & 7) - 4;
  const int64_t srcx = (int64_t)((bx + (x >> 3) * 8 + 4) << sub_x);
  const int64_t srcy = (int64_t)((by + (y >> 3) * 8 + 4) << sub_y);
  const int64_t dstx = (int64_t)wp[2] * srcx + (int64_t)wp[3] * srcy + wp[0];
  const int64_t dsty = (int64_t)wp[4] * srcx + (int64_t)wp[5] * srcy + wp[1];
  const int64_t x4 = dstx >> sub_x, y4 = dsty >> sub_y;
  const int ix4 = (int)(x4 >> 16), sx4 = (int)(x4 & 0xffff);
  const int iy4 = (int)(y4 >> 16), sy4 = (int)(y4 & 0xffff);
  const int round0 = bd >= 12 ? 5 : 3;
  int round1 = compound ? 7 : 11;
  if (bd >= 12 && !compound) round1 -= 2;
  const int oy = warp_offs(sy4 + abcd[2] * i2 + abcd[3] * i1);
  int acc = 0;
  for (int k = 0; k < 8; ++k) {
    const int row = i1 - 3 + k;
    const int ox = warp_offs(sx4 + abcd[0] * i2 + abcd[1] * row);
    int s = 0;
    for (int t = 0; t < 8; ++t)
      s += wf[ox * 8 + t] * fetch_px(ref, cur, stride, ix4 + i2 - 3 + t, iy4 + row, ref_w - 1, ref_h - 1);
    acc += ((s + (1 << (round0 - 1))) >> round0) * wf[oy * 8 + k];
  }
  return (acc + (1 << (round1 - 1))) >> round1;
}

/* the wedge master masks (av1gpu_kernels.h:1968-2014) */
__constant__ const int kWedgeEven[8] = {1, 4, 11, 27, 46, 58, 62, 63};
__constant__ const int kWedgeOdd[8] = {1, 2, 6, 18, 37, 53, 60, 63};
__constant__ const int kWedgeVert[6] = {2, 7, 21, 43, 57, 62};

__device__ int wedge_1d(int which, int j) {
  if (which == 2) return j < 29 ? 0 : (j >= 35 ? 64 : kWedgeVert[j - 29]);
  if (j < 28) return 0;
  if (j >= 36) return 64;
  return which == 0 ? kWedgeEven[j - 28] : kWedgeOdd[j - 28];
}
__device__ int wedge_ob63(int i, int j) {
  const int shift = (i & 1) ? 15 - (i >> 1) : 16 - (i >> 1);
  return wedge_1d(i & 1, clampi(j - shift, 0, 63));
}
__device__ int wedge_master(int dir, int i, int j) {
  switch (dir) {
    case 0: return wedge_1d(2, i);
    case 1: return wedge_1d(2, j);
    case 2: return wedge_ob63(j, i);
    case 4: return 64 - wedge_ob63(i, 63 - j);
    case 5: return 64 - wedge_ob63(j, 63 - i);
    default: return wedge_ob63(i, j);
  }
}
__device__ int wedge_px(BlkRef b, int i, int k) {
  const int m = wedge_master(b.wedge_dir, clampi(b.wedge_yoff + i, 0, 63), clampi(b.wedge_xoff + k, 0, 63));
  return b.wedge_flip ? 64 - m : m;
}

/* av1gpu_kernels.h:2025-2029 */
__device__ __forceinline__ int diffwtd(int d, int sh, int sign) {
  d = (d + (1 << (sh - 1))) >> sh;
  d = clampi(38 + (d >> 4), 0, 64);
  return sign ? 64 - d : d;
}

/* av1gpu_kernels.h:2076-2100 */
__constant__ const int kObmcMask[62] = {45, 64,
                                        39, 50, 59, 64,
                                        36, 42, 48, 53, 57, 61, 64, 64,
                                        34, 37, 40, 43, 46, 49, 52, 54, 56, 58, 60, 61, 64, 64, 64, 64,
                                        33, 35, 36, 38, 40, 41, 43, 44, 45, 47, 48, 50, 51, 52, 53, 55,
                                        56, 57, 58, 59, 60, 60, 61, 62, 64, 64, 64, 64, 64, 64, 64, 64};
__device__ __forceinline__ int obmc_mask(int len, int i) {
  const int base = len == 2 ? 0 : (len == 4 ? 2 : (len == 8 ? 6 : (len == 16 ? 14 : 30)));
  return kObmcMask[base + i];
}

/* inter/intra edges and prediction (av1gpu_kernels.h:2120-2205) */
struct IiCtx {
  const uint16_t* p; /* the block's top left sample in the frame being decoded */
  int stride, x, y, w, h, max_x, max_y, have_left, have_above, bd;
  __device__ int above(int i) const {
    const int mid = 1 << (bd - 1);
    if (!have_above) return have_left ? ld_dev(p - 1) : mid - 1;
    if (x + i > max_x) i = max_x - x;
    return ld_dev(p - stride + i);
  }
  __device__ int left(int i) const {
    const int mid = 1 << (bd - 1);
    if (!have_left) return have_above ? ld_dev(p - stride) : mid + 1;
    if (y + i > max_y) i = max_y - y;
    return ld_dev(p + (ptrdiff_t)i * stride - 1);
  }
};

__device__ int ii_sample(const IiCtx& c, int mode, int xs, int ys) {
  if (mode == 1) return c.above(xs);
  if (mode == 2) return c.left(ys);
  const int* sw = SPEC(AV1T_OFF_SM_WEIGHTS);
  if (mode == 3) {
    const int wy = sw[c.h - 4 + ys], wx = sw[c.w - 4 + xs];
    return rnd2(wy * c.above(xs) + (256 - wy) * c.left(c.h - 1) + wx * c.left(ys) + (256 - wx) * c.above(c.w - 1), 9);
  }
  int sum = 0;
  if (c.have_above && c.have_left) {
    for (int k = 0; k < c.w; ++k) sum += c.above(k);
    for (int k = 0; k < c.h; ++k) sum += c.left(k);
    return (sum + ((c.w + c.h) >> 1)) / (c.w + c.h);
  }
  if (c.have_left) {
    for (int k = 0; k < c.h; ++k) sum += c.left(k);
    return (sum + (c.h >> 1)) >> ilog2(c.h);
  }
  if (c.have_above) {
    for (int k = 0; k < c.w; ++k) sum += c.above(k);
    return (sum + (c.w >> 1)) >> ilog2(c.w);
  }
  return 1 << (c.bd - 1);
}

__device__ int ii_mask(BlkRef b, int sub_x, int sub_y, int xs, int ys) {
  if (b.flags & kIIWedge) {
    const int ny = sub_y + 1, nx = sub_x + 1, n = nx * ny;
    int sum = 0;
    for (int i = 0; i < ny; ++i)
      for (int k = 0; k < nx; ++k) sum += wedge_px(b, (ys << sub_y) + i, (xs << sub_x) + k);
    return n == 1 ? sum : (n == 2 ? (sum + 1) >> 1 : (sum + 2) >> 2);
  }
  if (b.ii_mode == 0) return 32;
  const int scale = 128 / max(b.w, b.h);
  int k = 0;
  if (b.ii_mode == 1)
    k = ys * scale;
  else if (b.ii_mode == 2)
    k = xs * scale;
  else if (b.ii_mode == 3)
    k = min(xs, ys) * scale;
  return SPEC(AV1T_OFF_II_WEIGHTS)[min(k, 127)];
}

/* Sample (xs, ys) of a block using any inter tool (av1gpu_kernels.h:2207-2326).
 * References of slot -1 (intra block copy) are the frame being decoded. */
__device__ __forceinline__ int inter_sample(RArgs a, BlkRef b, int xs, int ys) {
  LayRef L = a.lay;
  const int p = b.plane, bd = L.bit_depth;
  const int sub_x = p > 0 ? L.ss_x : 0, sub_y = p > 0 ? L.ss_y : 0;
  const int stride = L.stride[p];
  const int compound = b.compound != 0;
  int v[2] = {0, 0};
  for (int r = 0; r < (compound ? 2 : 1); ++r) {
    const uint16_t* base = ref_plane(a, b.slot[r], p);
    const bool cur = b.slot[r] < 0;
    const int cmp = r ? 1 : compound;
    v[r] = (b.flags & (r ? kWarp1 : kWarp0))
               ? warp_tap(sub_x, sub_y, b.ref_w, b.ref_h, b.warp_params[r], b.warp_abcd[r], base, cur, stride,
                          b.x, b.y, xs, ys, bd, cmp)
               : mc_tap(base, cur, stride, b.ref_w, b.ref_h, sub_x, sub_y, b.scale_x, b.scale_y, b.x, b.y, b.w,
                        b.h, b.mv[r][0], b.mv[r][1], b.filt, b.filt_v, xs, ys, cmp, bd);
  }
  int out = v[0];
  if (compound) {
    const int round0 = bd >= 12 ? 5 : 3;
    const int post = 2 * 7 - (round0 + 7);
    if (b.mask_kind != 0) {
      const int sh = post + (bd - 8);
      const int ny = sub_y + 1, nx = sub_x + 1;
      int sum = 0, n = 0;
      for (int i = 0; i < ny; ++i)
        for (int k = 0; k < nx; ++k) {
          const int lx = (xs << sub_x) + k, ly = (ys << sub_y) + i;
          int m;
          if (b.mask_kind == 1) {
            m = wedge_px(b, ly, lx);
            if (b.mask_sign) m = 64 - m;
          } else if (p == 0) {
            m = diffwtd(abs(v[0] - v[1]), sh, b.mask_sign);
          } else {
            /* the luma difference at the co-located luma sample */
            int lv[2];
            for (int r = 0; r < 2; ++r) {
              const uint16_t* lb = ref_plane(a, b.slot[r], 0);
              const bool cur = b.slot[r] < 0;
              lv[r] = (b.flags & (r ? kMaskWarp1 : kMaskWarp0))
                          ? warp_tap(0, 0, L.out_w[0], L.out_h[0], b.warp_params[r], b.warp_abcd[r], lb, cur,
                                     L.stride[0], b.x << L.ss_x, b.y << L.ss_y, lx, ly, bd, 1)
                          : mc_tap(lb, cur, L.stride[0], L.out_w[0], L.out_h[0], 0, 0, b.scale_x, b.scale_y,
                                   b.x << L.ss_x, b.y << L.ss_y, b.w << sub_x, b.h << sub_y, b.mv[r][0],
                                   b.mv[r][1], b.filt, b.filt_v, lx, ly, 1, bd);
            }
            m = diffwtd(abs(lv[0] - lv[1]), sh, b.mask_sign);
          }
          sum += m;
          ++n;
        }
      const int m = n == 1 ? sum : (n == 2 ? (sum + 1) >> 1 : (sum + 2) >> 2);
      out = (m * v[0] + (64 - m) * v[1] + (1 << (5 + post))) >> (6 + post);
    } else if (b.compound == AV1H_COMPOUND_AVERAGE) {
      out = (v[0] + v[1] + (1 << post)) >> (1 + post);
    } else {
      out = (b.w0 * v[0] + b.w1 * v[1] + (1 << (3 + post))) >> (4 + post);
    }
  }
  out = clampi(out, 0, a.maxv);
  if (b.flags & kInterIntra) {
    IiCtx c;
    c.p = a.frame + L.plane_off[p] + (size_t)b.y * stride + b.x;
    c.stride = stride;
    c.x = b.x;
    c.y = b.y;
    c.w = b.w;
    c.h = b.h;
    c.max_x = L.grid_w[p] - 1;
    c.max_y = L.grid_h[p] - 1;
    c.have_left = (b.flags & kHaveLeft) != 0;
    c.have_above = (b.flags & kHaveAbove) != 0;
    c.bd = bd;
    const int m = ii_mask(b, sub_x, sub_y, xs, ys);
    out = (m * ii_sample(c, b.ii_mode, xs, ys) + (64 - m) * out + 32) >> 6;
  }
  for (int k = 0; k < b.obmc_count; ++k) {
    const DevObmc& o = a.plan.obmc[b.obmc_first + k];
    if (xs < o.rx || xs >= o.rx + o.w || ys < o.ry || ys >= o.ry + o.h) continue;
    const int cj = xs - o.rx, ci = ys - o.ry;
    const int m = o.pass == 0 ? obmc_mask(o.h, ci) : obmc_mask(o.w, cj);
    const uint16_t* base = ref_plane(a, o.slot, p);
    int pred = mc_tap(base, o.slot < 0, stride, b.ref_w, b.ref_h, sub_x, sub_y, b.scale_x, b.scale_y, b.x + o.rx,
                      b.y + o.ry, o.w, o.h, o.mvx, o.mvy, o.filt, o.filt_v, cj, ci, 0, bd);
    pred = clampi(pred, 0, a.maxv);
    out = (m * out + (64 - m) * pred + 32) >> 6;
  }
  return out;
}

/* ------------------------------------------------------------------------ */
/* intra prediction                                                          */
/* ------------------------------------------------------------------------ */

constexpr int kEdge = 160;   /* edge entries (w + h <= 128, upsampled small) */
constexpr int kEdgeOff = 8;

/* The tables intra prediction indexes per lane (smooth weights, edge
 * kernels, filter intra taps), staged in LDS by the queue kernel: one
 * contiguous range of the specification table. */
constexpr int kTabLo = AV1T_OFF_SM_WEIGHTS, kTabHi = AV1T_OFF_II_WEIGHTS;
static_assert(AV1T_OFF_EDGE_KERNEL > kTabLo && AV1T_OFF_FILTER_INTRA > kTabLo && AV1T_OFF_FILTER_INTRA < kTabHi,
              "staged table range");
#define TAB(off) (tab + ((off) - kTabLo))

/* The edge preparation of a directional block (av1gpu_kernels.h:1452-1481,
 * with the filter of 1157-1173 and the upsampling of 1177-1191), its
 * choices precomputed on the host (DevBlock::dir_a): the corner filter, the
 * 5 tap filter of each edge into fa / fl and the 2x upsampling into ga / gl,
 * out of place so that each step is one LDS pass. `above` / `left` hold
 * entries -1 (the corner) .. w + h - 1; *ea / *el get the edges to predict
 * from. */
constexpr int kPrep1 = 2 * (kEdge + 2 * kEdgeOff), kPrep2 = 4 * (kEdge + 2 * kEdgeOff);
static_assert(kPrep2 + 2 * (kEdge + 2 * kEdgeOff) <= kWork, "edge preparation area");

/* What a queued block's work after its dependency wait reads of its
 * descriptor and of the frame's layout, loaded before the wait. Read after
 * it, each field is a scalar load, and the addresses chain several of them
 * one after another (the plane's strides and offsets, then the address):
 * about 0.3 us of each block's work after its wait (2.13 -> 1.85 us on the
 * 1080p key frame) that now overlaps the wait. */
#define AV1H_GLOBAL __attribute__((address_space(1)))
struct QBlk {
  int x, y, w, h, plane, kind, flags, mode, dir_a, dir_b;
  int stride, max_x, max_y, maxv, bd;
  uint32_t epoch;
  /* global (not generic) pointers: their stores need no wait for LDS */
  AV1H_GLOBAL uint16_t* dst;     /* sample (x, y) of the block's plane in the frame */
  AV1H_GLOBAL uint64_t* row_pub; /* its last row's packed edge words (kEdgeRow)     */
  AV1H_GLOBAL uint64_t* col_pub; /* its last column's packed edge words (kEdgeCol)  */
};

__device__ __forceinline__ QBlk queued_block(RArgs a, BlkRef b) {
  LayRef L = a.lay;
  QBlk q;
  q.x = b.x;
  q.y = b.y;
  q.w = b.w;
  q.h = b.h;
  q.plane = __builtin_amdgcn_readfirstlane(b.plane);
  q.kind = b.kind;
  q.flags = b.flags;
  q.mode = b.mode;
  q.dir_a = b.dir_a;
  q.dir_b = b.dir_b;
  const int p = q.plane, sh = L.cell_shift[p];
  q.stride = L.stride[p];
  q.max_x = L.grid_w[p] - 1;
  q.max_y = L.grid_h[p] - 1;
  q.maxv = a.maxv;
  q.bd = L.bit_depth;
  q.epoch = a.epoch;
  q.dst = (AV1H_GLOBAL uint16_t*)(a.frame + L.plane_off[p] + (int64_t)q.y * q.stride + q.x);
  q.row_pub = (AV1H_GLOBAL uint64_t*)(a.edges + L.edge_row_off[p] +
                                      (int64_t)(((q.y + q.h) >> sh) - 1) * L.edge_row_words[p] + (q.x >> 1));
  q.col_pub = (AV1H_GLOBAL uint64_t*)(a.edges + L.edge_col_off[p] +
                                      (int64_t)(((q.x + q.w) >> sh) - 1) * L.edge_col_words[p] + (q.y >> 1));
  /* every field computed here, in scalar registers, before the caller's
   * wait: the compiler would otherwise sink the loads to their uses */
  asm volatile(""
               : "+s"(q.x), "+s"(q.y), "+s"(q.w), "+s"(q.h), "+s"(q.plane), "+s"(q.kind), "+s"(q.flags),
                 "+s"(q.mode), "+s"(q.dir_a), "+s"(q.dir_b), "+s"(q.stride), "+s"(q.max_x), "+s"(q.max_y),
                 "+s"(q.maxv), "+s"(q.bd), "+s"(q.epoch), "+s"(q.dst), "+s"(q.row_pub), "+s"(q.col_pub));
  return q;
}

__device__ void prepare_edges(const QBlk& b, const int* above, const int* left, int* work, const int* tab, int maxv,
                              int max_x, int max_y, const int** ea, const int** el) {
  const int lane = threadIdx.x, w = b.w, h = b.h, n = w + h;
  const int da = b.dir_a, pa = da & 511;
  const int stra = (da >> kDirStrA) & 7, strl = (da >> kDirStrL) & 7;
  int corner = above[-1];
  if (da & kDirCorner) corner = rnd2(left[0] * 5 + above[-1] * 6 + above[0] * 5, 4);
  const int sza = stra ? min(w, max_x - b.x + 1) + (pa < 90 ? h : 0) + 1 : 0;
  const int szl = strl ? min(h, max_y - b.y + 1) + (pa > 180 ? w : 0) + 1 : 0;
  int* fa = work + kPrep1 + kEdgeOff;
  int* fl = fa + kEdge + 2 * kEdgeOff;
  const int* ka = TAB(AV1T_OFF_EDGE_KERNEL) + 5 * max(stra - 1, 0);
  const int* kl = TAB(AV1T_OFF_EDGE_KERNEL) + 5 * max(strl - 1, 0);
  /* entry k of the filtered edge: the filter over entries clamp(k + q - 2,
   * -1, sz - 2) for 0 <= k <= sz - 2, the entry itself elsewhere */
  for (int j = lane - 1; j < 2 * n + 1; j += 64) {
    const bool is_a = j < n;
    const int k = is_a ? j : j - n - 1;
    const int* src = is_a ? above : left;
    const int sz = is_a ? sza : szl;
    int v;
    if (k >= 0 && k <= sz - 2) {
      const int* kk = is_a ? ka : kl;
      int acc = 0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const int m = clampi(k + q - 2, -1, sz - 2);
        acc += kk[q] * (m < 0 ? corner : src[m]);
      }
      v = (acc + 8) >> 4;
    } else {
      v = k < 0 ? corner : src[k];
    }
    (is_a ? fa : fl)[k] = v;
  }
  wsync();
  *ea = fa;
  *el = fl;
  const int ua = (da & kDirUpA) != 0, ul = (da & kDirUpL) != 0;
  if (!ua && !ul) return;
  /* upsampled entry k: -2 the first tap, odd k = 2i - 1 the 4 tap
   * interpolation, even k = 2i the sample, past 2m - 2 (up to w + h - 1)
   * the entry as it was */
  const int na = ua ? w + (pa < 90 ? h : 0) : 0, nl = ul ? h + (pa > 180 ? w : 0) : 0;
  int* ga = work + kPrep2 + kEdgeOff;
  int* gl = ga + kEdge + 2 * kEdgeOff;
  const int span = 2 * n + 1; /* entries -2 .. 2n - 2 (upsampling needs w + h <= 16) */
  for (int j = lane; j < 2 * span; j += 64) {
    const bool is_a = j < span;
    const int m = is_a ? na : nl;
    const int k = (is_a ? j : j - span) - 2;
    if (m == 0 || (k > 2 * m - 2 && k > n - 1)) continue;
    const int* f = is_a ? fa : fl;
    auto t = [&](int i) { return f[clampi(i - 2, -1, m - 1)]; };
    int v;
    if (k == -2) {
      v = t(0);
    } else if (k <= 2 * m - 2) {
      if (k & 1) {
        const int i = (k + 1) >> 1;
        v = clampi(rnd2(-t(i) + 9 * t(i + 1) + 9 * t(i + 2) - t(i + 3), 4), 0, maxv);
      } else {
        v = t((k >> 1) + 2);
      }
    } else {
      v = f[k];
    }
    (is_a ? ga : gl)[k] = v;
  }
  wsync();
  if (ua) *ea = ga;
  if (ul) *el = gl;
}

/* The whole intra prediction of one block into s_pred (av1gpu_kernels.h:
 * 1578-1677 with the edges of 1108-1151 and the directional preparation of
 * 1452-1481). */
/* Polls a packed word until it carries `epoch`, with kPollDepth loads in
 * flight spaced about a round trip / kPollDepth apart (loads return in
 * order), so a word that lands is seen that much sooner than by one load
 * at a time. */
constexpr int kPollDepth = 1;
__device__ __forceinline__ uint64_t ld_word(const uint64_t* q) {
  return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ uint64_t poll_word(const uint64_t* q, uint32_t epoch, int limit, uint32_t* error) {
  uint64_t v[kPollDepth];
#pragma unroll
  for (int k = 0; k < kPollDepth; ++k) {
    if (k) __builtin_amdgcn_s_sleep(8);
    v[k] = ld_word(q);
  }
  for (int iter = 0;; iter += kPollDepth) {
#pragma unroll
    for (int k = 0; k < kPollDepth; ++k) {
      if ((uint32_t)(v[k] >> 32) == epoch) return v[k];
      if (kPollDepth == 1) __builtin_amdgcn_s_sleep(1);
      v[k] = ld_word(q);
    }
    if (iter >= limit) {
      atomicOr(error, 1u);
      return 0;
    }
  }
}

/* The edges of a block flagged kPackedEdges: its above row (x .. ae), left
 * column (y .. le) and corner come from the packed words its queued
 * neighbours published, each 64-bit word two samples and the epoch of the
 * frame that wrote them, so the one load per word is both the dependency
 * wait and the read (no cell flag, no second round trip to the frame), and
 * where the neighbour is outside the queue, from the frame without a wait.
 * Entries past ae / le repeat the last sample: only modes that read no
 * further than their dependencies reach use them. */
__device__ void gather_packed(RArgs a, BlkRef b, int* above, int* left, int alast, int llast, int mid,
                              uint32_t* error) {
  LayRef L = a.lay;
  const int p = b.plane, sh = L.cell_shift[p];
  const int x = b.x, y = b.y, n = b.w + b.h;
  const int ha = (b.flags & kHaveAbove) != 0, hl = (b.flags & kHaveLeft) != 0;
  const int ae = min(b.dep_x1, alast) - x, le = min(b.dep_y1, llast) - y;
  const uint64_t* rowp = a.edges + L.edge_row_off[p] + (int64_t)((y >> sh) - 1) * L.edge_row_words[p];
  const uint64_t* colp = a.edges + L.edge_col_off[p] + (int64_t)((x >> sh) - 1) * L.edge_col_words[p];
  /* samples of blocks outside the queue: final before it started */
  const int stride = L.stride[p];
  const uint16_t* fr = a.frame + L.plane_off[p] + (int64_t)y * stride + x;
  const uint32_t ma = (uint32_t)b.edge_mask_a, ml = (uint32_t)b.edge_mask_l;
  /* sample i of the above row / left column (block relative) */
  auto above_at = [&](int i) -> int {
    if ((ma >> (i >> sh)) & 1) {
      const uint64_t v = poll_word(rowp + ((x + i) >> 1), a.epoch, a.spin_limit, error);
      return (int)((v >> (16 * (i & 1))) & 0xffff);
    }
    return fr[i - stride];
  };
  auto left_at = [&](int i) -> int {
    if ((ml >> (i >> sh)) & 1) {
      const uint64_t v = poll_word(colp + ((y + i) >> 1), a.epoch, a.spin_limit, error);
      return (int)((v >> (16 * (i & 1))) & 0xffff);
    }
    return fr[(int64_t)i * stride - 1];
  };
  /* one pass over the corner (e = 0), above[0 .. n) and left[0 .. n), the
   * entries past the last sample a side depends on repeating it */
  for (int e = threadIdx.x; e < 2 * n + 1; e += 64) {
    int v;
    if (e == 0) {
      if (ha && hl) {
        if (b.flags & kCornerFrame) {
          v = fr[-stride - 1];
        } else {
          const uint64_t wv = poll_word((b.flags & kCornerRow) ? rowp + (x >> 1) - 1 : colp + (y >> 1) - 1, a.epoch,
                                        a.spin_limit, error);
          v = (int)((wv >> 16) & 0xffff); /* sample x - 1 of the row, or y - 1 of the column */
        }
      } else {
        v = ha ? above_at(0) : (hl ? left_at(0) : mid);
      }
      above[-1] = v;
      left[-1] = v;
    } else if (e <= n) {
      v = ha ? above_at(min(e - 1, ae)) : (hl ? left_at(0) : mid - 1);
      above[e - 1] = v;
    } else {
      v = hl ? left_at(min(e - n - 1, le)) : (ha ? above_at(0) : mid + 1);
      left[e - n - 1] = v;
    }
  }
}

/* The dependency wait of a kPackedEdges intra block: its edges into the
 * intra work area (entries -1 .. w + h - 1 of above / left, as
 * predict_intra reads them). */
__device__ void gather_edges(RArgs a, BlkRef b, int* work, uint32_t* error) {
  const int p = b.plane, w = b.w, h = b.h;
  gather_packed(a, b, work + kEdgeOff, work + kEdgeOff + kEdge + 2 * kEdgeOff,
                min(a.lay.grid_w[p] - 1, b.x + ((b.flags & kHaveAR) ? 2 * w : w) - 1),
                min(a.lay.grid_h[p] - 1, b.y + ((b.flags & kHaveBL) ? 2 * h : h) - 1), 1 << (a.lay.bit_depth - 1),
                error);
  wsync();
}

/* A queued block's last row and column as packed words (gather_packed). */
__device__ void write_edges(const QBlk& b, const uint16_t* pred) {
  const int w = b.w, h = b.h;
  const int nr = (b.flags & kEdgeRow) ? w >> 1 : 0, nc = (b.flags & kEdgeCol) ? h >> 1 : 0;
  const uint64_t tag = (uint64_t)b.epoch << 32;
  AV1H_GLOBAL uint64_t* rowp = b.row_pub;
  AV1H_GLOBAL uint64_t* colp = b.col_pub;
  const uint16_t* last = pred + (h - 1) * w;
  for (int k = threadIdx.x; k < nr + nc; k += 64) {
    uint64_t v;
    AV1H_GLOBAL uint64_t* q;
    if (k < nr) {
      v = tag | last[2 * k] | ((uint64_t)last[2 * k + 1] << 16);
      q = rowp + k;
    } else {
      const int j = k - nr;
      v = tag | pred[2 * j * w + w - 1] | ((uint64_t)pred[(2 * j + 1) * w + w - 1] << 16);
      q = colp + j;
    }
    __hip_atomic_store(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

/* Residual samples of a block read before its wait (they do not depend on
 * the neighbours): one transform covering the block, at most 256 samples,
 * sample lane + 64 k in v[k]. */
struct EarlyRes {
  bool on;
  int16_t v[4];
};

/* Writes prediction samples f(s), s = 0 .. n - 1, clamped, into pred; with
 * an early residual (one transform over at most 256 samples, sample
 * lane + 64 k in er.v[k]) the residual is added and clamped in the same
 * pass (av1gpu_kernels.h:933-937). */
template <class F>
__device__ __forceinline__ void emit(int n, uint16_t* pred, const EarlyRes& er, int maxv, F f) {
  if (er.on) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = threadIdx.x + 64 * k;
      if (s < n) pred[s] = (uint16_t)clampi(clampi(f(s), 0, maxv) + er.v[k], 0, maxv);
    }
  } else {
    for (int s = threadIdx.x; s < n; s += 64) pred[s] = (uint16_t)clampi(f(s), 0, maxv);
  }
}

/* Returns whether the early residual is already added. */
__device__ bool predict_intra(RArgs a, BlkRef desc, const QBlk& b, int* work, uint16_t* pred, const int* tab,
                              const EarlyRes& er, uint32_t* error) {
  const int lane = threadIdx.x;
  const int w = b.w, h = b.h, lw = ilog2(w);
  const int stride = b.stride;
  const int bd = b.bd, maxv = b.maxv;
  const uint16_t* src = (const uint16_t*)b.dst;
  if (b.flags & kPalette) {
    for (int s = lane; s < w * h; s += 64) {
      const int y = s >> lw, x = s & (w - 1);
      pred[s] = (uint16_t)desc.palette[a.plan.palette[desc.palette_off + y * desc.palette_stride + x]];
    }
    wsync();
    return false;
  }
  int* above = work + kEdgeOff;
  int* left = work + kEdgeOff + kEdge + 2 * kEdgeOff;
  int* tmp = work + 2 * (kEdge + 2 * kEdgeOff);
  const int have_left = (b.flags & kHaveLeft) != 0, have_above = (b.flags & kHaveAbove) != 0;
  const int mid = 1 << (bd - 1);
  const int max_x = b.max_x, max_y = b.max_y;
  const int n = w + h;
  if (!(b.flags & kPackedEdges)) { /* packed edges: gathered by the caller (gather_edges) */
    const int alast = min(max_x, b.x + ((b.flags & kHaveAR) ? 2 * w : w) - 1);
    const int llast = min(max_y, b.y + ((b.flags & kHaveBL) ? 2 * h : h) - 1);
    for (int i = lane; i < n; i += 64) {
      above[i] = have_above ? ld_dev(src - stride + (min(b.x + i, alast) - b.x))
                            : (have_left ? ld_dev(src - 1) : mid - 1);
      left[i] = have_left ? ld_dev(src + (ptrdiff_t)(min(b.y + i, llast) - b.y) * stride - 1)
                          : (have_above ? ld_dev(src - stride) : mid + 1);
    }
    if (lane == 0) {
      int c;
      if (have_above && have_left)
        c = ld_dev(src - stride - 1);
      else if (have_above)
        c = ld_dev(src - stride);
      else if (have_left)
        c = ld_dev(src - 1);
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
        for (int dx = 0; dx <= sx; ++dx) t += ld_dev(luma + (size_t)(ly + dy) * ls + lx + dx);
      const int v = t << (3 - sx - sy);
      tmp[s] = v;
      part += v;
    }
    const int sum = wave_sum(part);
    const int shift = ilog2(w) + ilog2(h);
    const int avg = (sum + (1 << (shift - 1))) >> shift;
    emit(w * h, pred, er, maxv, [&](int s) { return dc + rnd2s(desc.cfl_alpha * (tmp[s] - avg), 6); });
    wsync();
    return true;
  }
  if ((b.flags & kFilterIntra) && w >= 4 && h >= 2) {
    /* recursive filter intra (av1gpu_kernels.h:1392-1427), one anti-diagonal
     * of 4x2 sub-blocks at a time: a sub-block reads only the ones above, to
     * its left and above-left */
    const int w4 = w >> 2, h2 = h >> 1;
    const int* taps = TAB(AV1T_OFF_FILTER_INTRA) + 56 * desc.fi_mode;
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
    return false;
  }
  if (mode >= 1 && mode <= 8) {
    const int da = b.dir_a, pa = da & 511;
    const int ua = (da & kDirUpA) ? 1 : 0, ul = (da & kDirUpL) ? 1 : 0;
    const int* ea = above;
    const int* el = left;
    if (da & kDirFilter) prepare_edges(b, above, left, work, tab, maxv, max_x, max_y, &ea, &el);
    /* av1gpu_kernels.h:1245-1313 */
    const int dx = b.dir_b & 0xffff, dy = (int)((uint32_t)b.dir_b >> 16);
    emit(w * h, pred, er, maxv, [&](int s) {
      const int i = s >> lw, c = s & (w - 1);
      int v;
      if (pa < 90) {
        const int maxb = (w + h - 1) << ua;
        const int idx = (i + 1) * dx;
        const int base = (idx >> (6 - ua)) + (c << ua);
        const int sh = ((idx << ua) >> 1) & 31;
        v = base < maxb ? rnd2(ea[base] * (32 - sh) + ea[base + 1] * sh, 5) : ea[maxb];
      } else if (pa > 90 && pa < 180) {
        int idx = (c << 6) - (i + 1) * dx;
        int base = idx >> (6 - ua);
        if (base >= -(1 << ua)) {
          const int sh = ((idx * (1 << ua)) >> 1) & 31;
          v = rnd2(ea[base] * (32 - sh) + ea[base + 1] * sh, 5);
        } else {
          idx = (i << 6) - (c + 1) * dy;
          base = idx >> (6 - ul);
          const int sh = ((idx * (1 << ul)) >> 1) & 31;
          v = rnd2(el[base] * (32 - sh) + el[base + 1] * sh, 5);
        }
      } else if (pa > 180) {
        const int idx = (c + 1) * dy;
        const int base = (idx >> (6 - ul)) + (i << ul);
        const int sh = ((idx << ul) >> 1) & 31;
        v = rnd2(el[base] * (32 - sh) + el[base + 1] * sh, 5);
      } else {
        v = pa == 90 ? ea[c] : el[i];
      }
      return v;
    });
    wsync();
    return true;
  }
  if (mode >= AV1H_PRED_SMOOTH && mode <= AV1H_PRED_SMOOTH_H) {
    /* av1gpu_kernels.h:1333-1353 */
    const int* sw = TAB(AV1T_OFF_SM_WEIGHTS);
    emit(w * h, pred, er, maxv, [&](int s) {
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
      return v;
    });
    wsync();
    return true;
  }
  if (mode == AV1H_PRED_PAETH) {
    /* av1gpu_kernels.h:1364-1382 */
    emit(w * h, pred, er, maxv, [&](int s) {
      const int y = s >> lw, x = s & (w - 1);
      const int base = above[x] + left[y] - above[-1];
      const int pl = abs(base - left[y]), pt = abs(base - above[x]), ptl = abs(base - above[-1]);
      return (pl <= pt && pl <= ptl) ? left[y] : (pt <= ptl ? above[x] : above[-1]);
    });
    wsync();
    return true;
  }
  const int dc = clampi(dc_value(), 0, maxv);
  emit(w * h, pred, er, maxv, [&](int) { return dc; });
  wsync();
  return true;
}

/* ------------------------------------------------------------------------ */
/* block driver                                                              */
/* ------------------------------------------------------------------------ */

__device__ void store_block(const QBlk& b, const uint16_t* pred) {
  const int lw = ilog2(b.w);
  const int stride = b.stride;
  AV1H_GLOBAL uint16_t* dst = b.dst;
  if (b.flags & kSkipMark) {
    /* no block of this launch reads these samples (its readers take the
     * packed edge words): ordinary stores, written back through the L2 at
     * the end of the launch like any kernel's */
    for (int s = threadIdx.x; s < b.w * b.h; s += 64) dst[(size_t)(s >> lw) * stride + (s & (b.w - 1))] = pred[s];
    return;
  }
  for (int s = threadIdx.x; s < b.w * b.h; s += 64)
    __hip_atomic_store(dst + (size_t)(s >> lw) * stride + (s & (b.w - 1)), pred[s], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t cell_index(LayRef L, int p, int x, int y) {
  return L.cell_off[p] + (int64_t)(y >> L.cell_shift[p]) * L.cells_w[p] + (x >> L.cell_shift[p]);
}

/* the cells of a block, flagged with this frame's epoch */
__device__ void mark_cells(RArgs a, BlkRef b, bool atomic) {
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
__device__ void wait_cells(RArgs a, BlkRef b, uint32_t* error) {
  LayRef L = a.lay;
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
  int nr = 0, rcw = 0;
  if (b.flags & kIntraBC) {
    /* intra block copy: every cell of the source rectangle */
    rcw = b.dep_rx1 - b.dep_rx0 + 1;
    nr = rcw * (b.dep_ry1 - b.dep_ry0 + 1);
    if (rcw <= 0 || nr < 0) nr = 0;
  }
  const int total = nl + na + nc + ncfl + nr;
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
      else if (i < nl + na + nc + ncfl) {
        const int k = i - nl - na - nc;
        idx = L.cell_off[0] + (int64_t)(ly0l + k / lcw) * L.cells_w[0] + lx0 + k % lcw;
      } else {
        const int k = i - nl - na - nc - ncfl;
        idx = L.cell_off[p] + (int64_t)(b.dep_ry0 + k / rcw) * L.cells_w[p] + b.dep_rx0 + k % rcw;
      }
      if (__hip_atomic_load(a.cells + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.epoch) ok = 0;
    }
    if (__all(ok)) break;
    if (iter >= a.spin_limit) {
      if (threadIdx.x == 0) atomicOr(error, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  /* the neighbours' samples are read with device scope loads issued after
   * the flag loads returned */
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

/* A block of the queue: the prediction, then the residuals of its transforms, which
 * k_itx computed before the queue started (av1gpu_kernels.h:933-937 for the
 * add), into `pred`; the caller stores it. */


__device__ __forceinline__ EarlyRes early_residual(RArgs a, BlkRef b) {
  EarlyRes e;
  e.on = false;
  if (b.tx_count != 1 || b.w * b.h > 256) return e;
  const AV1H_CMEM DevTx& tx = cmem(a.plan.tx)[b.tx_first];
  if (tx.x != b.x || tx.y != b.y || tx.w != b.w || tx.h != b.h) return e;
  e.on = true;
  const int p = b.plane, stride = a.lay.stride[p], lw = ilog2(b.w);
  const int16_t* r = a.res + a.lay.plane_off[p] + (int64_t)b.y * stride + b.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int s = threadIdx.x + 64 * k;
    e.v[k] = s < b.w * b.h ? r[(int64_t)(s >> lw) * stride + (s & (b.w - 1))] : 0;
  }
  return e;
}

__device__ void reconstruct_queued(RArgs a, BlkRef b, const QBlk& q, int* work, uint16_t* pred, const EarlyRes& er,
                                   const int* tab, uint32_t* error) {
  bool fused = false;
  if (q.kind == kIntra) {
    fused = predict_intra(a, b, q, work, pred, tab, er, error);
  } else {
    /* inter/intra or intra block copy: reads the frame being decoded */
    const int lw = ilog2(b.w);
    for (int s = threadIdx.x; s < b.w * b.h; s += 64) pred[s] = (uint16_t)inter_sample(a, b, s & (b.w - 1), s >> lw);
    wsync();
  }
  const int p = q.plane, stride = q.stride;
  if (er.on) {
    if (fused) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = threadIdx.x + 64 * k;
      if (s < q.w * q.h) pred[s] = (uint16_t)clampi((int)pred[s] + er.v[k], 0, q.maxv);
    }
    wsync();
    return;
  }
  for (int t = 0; t < b.tx_count; ++t) {
    const AV1H_CMEM DevTx& tx = cmem(a.plan.tx)[b.tx_first + t];
    const int ox = tx.x - b.x, oy = tx.y - b.y, ltw = ilog2(tx.w), n = tx.w * tx.h;
    const int16_t* r = a.res + a.lay.plane_off[p] + (int64_t)tx.y * stride + tx.x;
    /* eight residual loads in flight per lane before the first add */
    for (int s0 = threadIdx.x; s0 < n; s0 += 64 * 8) {
      int rv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int s = s0 + 64 * u;
        rv[u] = s < n ? r[(int64_t)(s >> ltw) * stride + (s & (tx.w - 1))] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int s = s0 + 64 * u;
        if (s < n) {
          uint16_t* q = pred + (oy + (s >> ltw)) * b.w + ox + (s & (tx.w - 1));
          *q = (uint16_t)clampi((int)*q + rv[u], 0, a.maxv);
        }
      }
    }
  }
  wsync();
}

/* Motion compensation, one 256 thread workgroup per task (at most 512
 * samples of consecutive items in coding order, an item being a block or a
 * row strip of a large one):
 *   A  the task's item descriptors (positions, reference windows, LDS
 *      offsets, all computed by the host) are copied into LDS;
 *   B  the horizontal 8 tap pass over every reference row the item's
 *      vertical taps need, the taps read from the reference and clamped to
 *      its visible area the way each tap is (av1gpu_kernels.h:1704-1711),
 *      rounded by InterRound0, into an LDS intermediate (16 bit: at most
 *      200 * 4095 >> 5); neighbouring threads read neighbouring columns, so
 *      the 8 reads of a thread come from the same cache lines;
 *   C  per output sample: the vertical pass, InterRound1, the compound blend
 *      (av1gpu_kernels.h:2207-2293), the store. */
constexpr int kMcPairs = 2 * kMcItems;

__device__ __forceinline__ int find_prefix(const int* pre, int n, int v, int lo = 0) {
  /* the last k < n with pre[k] <= v, known to be >= lo */
  int hi = n - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (pre[m] <= v)
      lo = m;
    else
      hi = m - 1;
  }
  return lo;
}

__device__ __forceinline__ McPos mc_pos_of(const McItemDev& it, const McRefDev& r, int sub_x, int sub_y) {
  McPos P;
  P.scaled = r.scaled;
  P.bx = it.x;
  P.by = it.y;
  P.sub_x = sub_x;
  P.sub_y = sub_y;
  P.ix = r.ix;
  P.iy = r.iy;
  P.fx = r.fx;
  P.fy = r.fy;
  P.sx0 = r.sx0;
  P.sy0 = r.sy0;
  P.stx = r.stx;
  P.sty = r.sty;
  return P;
}

struct McLds {
  int16_t mid[kMcMid];
  McItemDev it[kMcItems];
  int moff[kMcPairs + 1], px[kMcItems + 1];
};

/* the 8 taps of a subpel filter row: 8 int16 in one 16-byte LDS read */
__device__ __forceinline__ void mc_taps(const int16_t* f, int* t) {
  const int4 v = *reinterpret_cast<const int4*>(f);
  t[0] = (int)(int16_t)v.x;
  t[1] = v.x >> 16;
  t[2] = (int)(int16_t)v.y;
  t[3] = v.y >> 16;
  t[4] = (int)(int16_t)v.z;
  t[5] = v.z >> 16;
  t[6] = (int)(int16_t)v.w;
  t[7] = v.w >> 16;
}

__device__ void mc_body(RArgs a, int tn, const int16_t* s_filt, McLds& L_) {
  int16_t* s_mid = L_.mid;
  McItemDev* s_it = L_.it;
  int* s_moff = L_.moff;
  int* s_px = L_.px;
  const int tid = threadIdx.x;
  const AV1H_CMEM McTaskDev& task = cmem(a.plan.mc_task)[tn];
  const int nit = task.count, npairs = 2 * nit;
  {
    /* all of a thread's loads issued before its LDS writes */
    const int32_t* src = reinterpret_cast<const int32_t*>(a.plan.mc_items + task.first);
    int32_t* dst = reinterpret_cast<int32_t*>(s_it);
    const int n = nit * (int)(sizeof(McItemDev) / 4);
    constexpr int kPer = (kMcItems * (int)(sizeof(McItemDev) / 4) + 255) / 256;
    int32_t val[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) val[u] = tid + 256 * u < n ? src[tid + 256 * u] : 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u)
      if (tid + 256 * u < n) dst[tid + 256 * u] = val[u];
  }
  __syncthreads();
  if (tid < npairs) s_moff[tid] = s_it[tid >> 1].r[tid & 1].mid_off;
  if (tid < nit) s_px[tid] = s_it[tid].px0;
  if (tid == 0) {
    s_moff[npairs] = task.mtotal;
    s_px[nit] = task.samples;
  }
  __syncthreads();
  const int bd = a.lay.bit_depth;
  const int round0 = bd >= 12 ? 5 : 3;
  const PlaneVals<int> strides = plane_vals(a.lay.stride);
  const PlaneVals<int64_t> poffs = plane_vals(a.lay.plane_off);
  /* horizontal pass over every reference row an item needs, taps read from
   * the reference (clamped to it) */
  int kprev = 0;
  for (int e = tid; e < task.mtotal; e += 256) {
    const int k = kprev = find_prefix(s_moff, npairs, e, kprev);
    const McItemDev& it = s_it[k >> 1];
    const McRefDev& q = it.r[k & 1];
    const int o = e - q.mid_off;
    const int lw = ilog2(it.w);
    const int rr = o >> lw, x = o & (it.w - 1);
    const int p = it.plane;
    const int sub_x = p > 0 ? a.lay.ss_x : 0;
    const McPos P = mc_pos_of(it, q, sub_x, 0);
    int cx, fx;
    P.col(x, &cx, &fx);
    const int y = clampi(q.wy0 + rr, 0, it.ref_h - 1);
    const uint16_t* row = a.pool + (size_t)q.slot * a.slot_size + poffs[p] + (size_t)y * strides[p];
    int hk[8];
    mc_taps(s_filt + it.hk + fx * 8, hk);
    int sum = 0;
    if (cx >= 3 && cx + 4 < it.ref_w) {
#pragma unroll
      for (int t = 0; t < 8; ++t) sum += (int)row[cx - 3 + t] * hk[t];
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t) sum += (int)row[clampi(cx - 3 + t, 0, it.ref_w - 1)] * hk[t];
    }
    s_mid[e] = (int16_t)((sum + (1 << (round0 - 1))) >> round0);
  }
  __syncthreads();
  /* vertical pass and blend, one thread per output sample */
  const PlaneVals<int> cshift = plane_vals(a.lay.cell_shift), cwidth = plane_vals(a.lay.cells_w),
                       cheight = plane_vals(a.lay.cells_h);
  const PlaneVals<int64_t> coffs = plane_vals(a.lay.cell_off);
  int iprev = 0;
  for (int si = tid; si < task.samples; si += 256) {
    const int item = iprev = find_prefix(s_px, nit, si, iprev);
    const McItemDev& it = s_it[item];
    const int s = si - it.px0;
    const int lw = ilog2(it.w), x = s & (it.w - 1), y = it.y0 + (s >> lw);
    const int compound = it.compound != 0;
    int round1 = compound ? 7 : 11;
    if (bd >= 12 && !compound) round1 -= 2;
    const int sub_y = it.plane > 0 ? a.lay.ss_y : 0;
    int v[2] = {0, 0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r == 1 && !compound) break;
      const McRefDev& q = it.r[r];
      const McPos P = mc_pos_of(it, q, 0, sub_y);
      int cy, fy;
      P.row(y, &cy, &fy);
      const int16_t* m = s_mid + q.mid_off + (cy - 3 - q.wy0) * it.w + x;
      int vk[8];
      mc_taps(s_filt + it.vk + fy * 8, vk);
      int acc = 0;
#pragma unroll
      for (int t = 0; t < 8; ++t) acc += (int)m[t * it.w] * vk[t];
      v[r] = (acc + (1 << (round1 - 1))) >> round1;
    }
    int out = v[0];
    if (compound) {
      const int post = 2 * 7 - (round0 + 7);
      if (it.compound == AV1H_COMPOUND_AVERAGE)
        out = (v[0] + v[1] + (1 << post)) >> (1 + post);
      else
        out = (it.w0 * v[0] + it.w1 * v[1] + (1 << (3 + post))) >> (4 + post);
    }
    const int p = it.plane;
    const int ax = it.x + x, ay = it.y + y;
    a.frame[poffs[p] + (int64_t)ay * strides[p] + ax] = (uint16_t)clampi(out, 0, a.maxv);
    /* the block's dependency cells, each by the thread of its first sample
     * in the block: final once the inter transforms after this launch have
     * run, which the queue launch follows */
    const int sh = cshift[p], cm = (1 << sh) - 1;
    if (((ax & cm) == 0 || x == 0) && ((ay & cm) == 0 || y == 0)) {
      const int cx = ax >> sh, cy = ay >> sh;
      if (cx < cwidth[p] && cy < cheight[p]) a.cells[coffs[p] + (int64_t)cy * cwidth[p] + cx] = a.epoch;
    }
  }
}

__global__ __launch_bounds__(256) void k_mc(BatchDev B) {
  __shared__ __align__(16) int16_t s_filt[768];
  __shared__ McLds s_lds;
  for (int i = threadIdx.x; i < 768; i += 256) s_filt[i] = (int16_t)av1t_spec_table[AV1T_OFF_SUBPEL + i];
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKMc);
  const int total = pre[B.njobs];
  for (int base = 0; base < total; base += gridDim.x) {
    const int v = xcd_group(base, total);
    if (v < 0) continue;
    const int j = find_job(pre, B.njobs, v);
    mc_body(recon_args(B, j), v - pre[j], s_filt, s_lds);
    __syncthreads();
  }
}

/* Inter blocks with warp, masks or OBMC (they read only references): one
 * workgroup per block, the direct per sample computation. */
__global__ __launch_bounds__(64) void k_inter_special(BatchDev B) {
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKSpecial);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    RArgs a = recon_args(B, j);
    const int bi = cmem(a.plan.special_list)[v - pre[j]];
    BlkRef b = cmem(a.plan.blocks)[bi];
    const int p = b.plane, stride = a.lay.stride[p], lw = ilog2(b.w);
    uint16_t* dst = a.frame + a.lay.plane_off[p] + (size_t)b.y * stride + b.x;
    for (int s = threadIdx.x; s < b.w * b.h; s += 64)
      dst[(size_t)(s >> lw) * stride + (s & (b.w - 1))] = (uint16_t)inter_sample(a, b, s & (b.w - 1), s >> lw);
    mark_cells(a, b, false);
  }
}

__global__ __launch_bounds__(64) void k_queue(BatchDev B) {
  __shared__ int s_work[kWork];
  __shared__ uint16_t s_pred[64 * 64];
  __shared__ int s_tab[kTabHi - kTabLo];
  for (int i = threadIdx.x; i < kTabHi - kTabLo; i += 64) s_tab[i] = av1t_spec_table[kTabLo + i];
  __syncthreads();
  const int nqueue = kernel_prefix(B, kKernels)[0];
  for (;;) {
    int idx = 0;
    if (threadIdx.x == 0) idx = (int)atomicAdd(B.queue_counter, 1u);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= nqueue) {
      /* every workgroup takes exactly one entry past the end; the last of
       * those is the counter's last use in this launch, so it resets the
       * counter for the next one */
      if (threadIdx.x == 0 && idx == nqueue + (int)gridDim.x - 1)
        __hip_atomic_store(B.queue_counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    const uint64_t t0 = B.trace ? (uint64_t)wall_clock64() : 0;
    int job = 0, blk;
    if (B.queue) {
      const uint32_t e = cmem(B.queue)[idx];
      job = (int)(e >> kQueueJobShift);
      blk = (int)(e & ((1u << kQueueJobShift) - 1));
    } else {
      blk = cmem(recon_args(B, 0).plan.queue_list)[idx];
    }
    RArgs a = recon_args(B, __builtin_amdgcn_readfirstlane(job));
    BlkRef b = cmem(a.plan.blocks)[__builtin_amdgcn_readfirstlane(blk)];
    const EarlyRes er = early_residual(a, b);
    const QBlk q = queued_block(a, b);
    if (b.flags & kPackedEdges)
      gather_edges(a, b, s_work, B.error);
    else
      wait_cells(a, b, B.error);
    const uint64_t t1 = B.trace ? (uint64_t)wall_clock64() : 0;
    reconstruct_queued(a, b, q, s_work, s_pred, er, s_tab, B.error);
    /* the packed edges first: the blocks waiting on this one read only them */
    if (q.flags & (kEdgeRow | kEdgeCol)) write_edges(q, s_pred);
    store_block(q, s_pred);
    if (!(q.flags & kSkipMark)) {
      /* every sample store of the block has completed before its cells flip */
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      mark_cells(a, b, true);
    }
    wsync();
    if (B.trace && threadIdx.x == 0) {
      /* diagnostics: when the block was taken, found ready, published */
      uint64_t* t = B.trace + 4 * (size_t)idx;
      t[0] = t0;
      t[1] = t1;
      t[2] = (uint64_t)wall_clock64();
      t[3] = (uint64_t)blockIdx.x;
    }
  }
}


}  // namespace

int itx_groups(const FramePlanDev& p, int set, int c0, int c1) {
  int g = 0;
  for (int c = c0; c < c1; ++c) {
    const int n = p.itx_off[set][c + 1] - p.itx_off[set][c], per = 16 >> c;
    g += (n + per - 1) / per;
  }
  return g;
}

hipError_t launch_recon_kernel(const BatchDev& b, int k, int total, int cap, hipStream_t s) {
  const dim3 grid(min(total, cap));
  switch (k) {
    case kKMc: hipLaunchKernelGGL(k_mc, grid, dim3(256), 0, s, b); break;
    case kKSpecial: hipLaunchKernelGGL(k_inter_special, grid, dim3(64), 0, s, b); break;
    case kKItxInter: hipLaunchKernelGGL((k_itx<0, 3>), grid, dim3(64), 0, s, b, 0); break;
    case kKItxIntra: hipLaunchKernelGGL((k_itx<0, 3>), grid, dim3(64), 0, s, b, 1); break;
    case kKItxInterL: hipLaunchKernelGGL((k_itx<3, 5>), grid, dim3(64), 0, s, b, 0); break;
    case kKItxIntraL: hipLaunchKernelGGL((k_itx<3, 5>), grid, dim3(64), 0, s, b, 1); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_queue(const BatchDev& b, int total, int cap, hipStream_t s, int* launches) {
  if (total == 0) return hipSuccess;
  const int grid = total < 0 ? cap : min(total, cap);
  hipLaunchKernelGGL(k_queue, dim3(grid), dim3(64), 0, s, b);
  *launches += 1;
  return hipGetLastError();
}

}  // namespace av1hip
