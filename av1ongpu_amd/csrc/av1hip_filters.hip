/*
 * Post-reconstruction passes for gfx950: deblocking (two ordered passes),
 * CDEF, superres upscaling, loop restoration, and the NV12 / RGB packs. The
 * arithmetic is the reference's (kernels/av1gpu_kernels.h), cited per function.
 */
#define AV1T_TABLE_QUAL static __constant__ const
#include "av1hip_internal.hpp"
#include "spec_tables.h"

namespace av1hip {
namespace {

#define SPEC(off) (av1t_spec_table + (off))

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int rnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ int rsr64(int64_t v, int n) { return (int)((v + ((int64_t)1 << (n - 1))) >> n); }

/* ---- deblocking (av1gpu_kernels.h:2422-2572) --------------------------- */

__device__ __forceinline__ int f4c(int v, int bd) { return clampi(v, -(1 << (bd - 1)), (1 << (bd - 1)) - 1); }

/* one line across one edge: b points at p1, s is the step across the edge */
__device__ void deblock_line(uint16_t* b, int s, int level, int sharp, int fs, int chroma, int bd,
                             int maxv) {
  const int sh = bd - 8, mid = 0x80 << sh;
  const int p1 = b[0], p0 = b[s], q0 = b[2 * s], q1 = b[3 * s];
  const int flen = fs == 4 ? 4 : (chroma ? 6 : (fs == 8 ? 8 : 16));
  const int lsh = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
  int limit = level >> lsh;
  if (limit < 1) limit = 1;
  if (sharp > 0 && limit > 9 - sharp) limit = 9 - sharp;
  const int blim = (2 * (level + 2) + limit) << sh, lim = limit << sh, thr = (level >> 4) << sh;
  const int one = 1 << sh;
  const int hev = abs(p1 - p0) > thr || abs(q1 - q0) > thr;
  int bad = abs(p1 - p0) > lim || abs(q1 - q0) > lim || abs(p0 - q0) * 2 + abs(p1 - q1) / 2 > blim;
  int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
  if (flen >= 6) {
    p2 = b[-s];
    q2 = b[4 * s];
    bad = bad || abs(p2 - p1) > lim || abs(q2 - q1) > lim;
  }
  if (flen >= 8) {
    p3 = b[-2 * s];
    q3 = b[5 * s];
    bad = bad || abs(p3 - p2) > lim || abs(q3 - q2) > lim;
  }
  if (bad) return;
  int flat = 1, flat2 = 1;
  if (fs >= 8) {
    int f = abs(p1 - p0) > one || abs(q1 - q0) > one || abs(p2 - p0) > one || abs(q2 - q0) > one;
    if (flen >= 8) f = f || abs(p3 - p0) > one || abs(q3 - q0) > one;
    flat = !f;
  }
  if (fs >= 16) {
    flat2 = !(abs(b[-5 * s] - p0) > one || abs(b[8 * s] - q0) > one || abs(b[-4 * s] - p0) > one ||
              abs(b[7 * s] - q0) > one || abs(b[-3 * s] - p0) > one || abs(b[6 * s] - q0) > one);
  }
  if (fs == 4 || !flat) {
    const int ps1 = p1 - mid, ps0 = p0 - mid, qs0 = q0 - mid, qs1 = q1 - mid;
    int f = hev ? f4c(ps1 - qs1, bd) : 0;
    f = f4c(f + 3 * (qs0 - ps0), bd);
    const int f1 = f4c(f + 4, bd) >> 3, f2 = f4c(f + 3, bd) >> 3;
    b[2 * s] = (uint16_t)clampi(f4c(qs0 - f1, bd) + mid, 0, maxv);
    b[s] = (uint16_t)clampi(f4c(ps0 + f2, bd) + mid, 0, maxv);
    if (!hev) {
      const int f3 = rnd2(f1, 1);
      b[3 * s] = (uint16_t)clampi(f4c(qs1 - f3, bd) + mid, 0, maxv);
      b[0] = (uint16_t)clampi(f4c(ps1 + f3, bd) + mid, 0, maxv);
    }
    return;
  }
  const int lg = (fs == 8 || !flat2) ? 3 : 4;
  const int n = lg == 4 ? 6 : (chroma ? 2 : 3);
  const int n2 = (lg == 3 && !chroma) ? 0 : 1;
  int v[14];
  for (int i = -(n + 1); i <= n; ++i) v[i + 7] = b[(2 + i) * s];
  int res[12];
  for (int i = -n; i < n; ++i) {
    int t = 0;
    for (int k = -n; k <= n; ++k) t += v[clampi(i + k, -(n + 1), n) + 7] * (abs(k) <= n2 ? 2 : 1);
    res[i + n] = rnd2(t, lg);
  }
  for (int i = -n; i < n; ++i) b[(2 + i) * s] = (uint16_t)clampi(res[i + n], 0, maxv);
}

/* one thread per line; edges of one direction never share samples */
__global__ void k_deblock(FilterArgs a, int first, int count, int lines) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= lines) return;
  /* binary search of the edge owning line g (prefix in line0) */
  int lo = first, hi = first + count - 1;
  const int base = a.plan.deblock[first].line0;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (a.plan.deblock[m].line0 - base <= g)
      lo = m;
    else
      hi = m - 1;
  }
  const DevDeblock& e = a.plan.deblock[lo];
  if (e.level == 0) return;
  const int k = g - (e.line0 - base);
  const int p = e.plane, st = a.lay.stride[p];
  uint16_t* plane = a.frame + a.lay.plane_off[p];
  uint16_t* b;
  int s;
  if (e.vertical) {
    b = plane + (size_t)(e.y + k) * st + (e.x - 2);
    s = 1;
  } else {
    b = plane + (size_t)(e.y - 2) * st + (e.x + k);
    s = st;
  }
  deblock_line(b, s, e.level, a.sharpness, e.filter_size, p != 0, a.lay.bit_depth, a.maxv);
}

/* ---- CDEF (av1gpu_kernels.h:2601-2775) ---------------------------------- */

__device__ void cdef_direction(const uint16_t* l, int stride, int bd, int* dir_out, int* var_out) {
  const int div[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};
  int part[8][15];
  int cost[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 15; ++j) part[i][j] = 0;
  for (int i = 0; i < 8; ++i)
    for (int k = 0; k < 8; ++k) {
      const int x = (l[i * stride + k] >> (bd - 8)) - 128;
      part[0][i + k] += x;
      part[1][i + k / 2] += x;
      part[2][i] += x;
      part[3][3 + i - k / 2] += x;
      part[4][7 + i - k] += x;
      part[5][3 - i / 2 + k] += x;
      part[6][k] += x;
      part[7][i / 2 + k] += x;
    }
  for (int i = 0; i < 8; ++i) {
    cost[2] += part[2][i] * part[2][i];
    cost[6] += part[6][i] * part[6][i];
  }
  cost[2] *= div[8];
  cost[6] *= div[8];
  for (int i = 0; i < 7; ++i) {
    cost[0] += (part[0][i] * part[0][i] + part[0][14 - i] * part[0][14 - i]) * div[i + 1];
    cost[4] += (part[4][i] * part[4][i] + part[4][14 - i] * part[4][14 - i]) * div[i + 1];
  }
  cost[0] += part[0][7] * part[0][7] * div[8];
  cost[4] += part[4][7] * part[4][7] * div[8];
  for (int i = 1; i < 8; i += 2) {
    for (int k = 0; k < 5; ++k) cost[i] += part[i][3 + k] * part[i][3 + k];
    cost[i] *= div[8];
    for (int k = 0; k < 3; ++k)
      cost[i] += (part[i][k] * part[i][k] + part[i][10 - k] * part[i][10 - k]) * div[2 * k + 2];
  }
  int best = 0, d = 0;
  for (int i = 0; i < 8; ++i)
    if (cost[i] > best) {
      best = cost[i];
      d = i;
    }
  *dir_out = d;
  *var_out = (best - cost[(d + 4) & 7]) >> 10;
}

__device__ __forceinline__ int flog2(int v) { return 31 - __builtin_clz(v); }

__device__ __forceinline__ int constrain(int diff, int thr, int damping) {
  if (!thr) return 0;
  const int adj = max(0, damping - flog2(thr));
  const int mag = abs(diff);
  const int v = clampi(thr - (mag >> adj), 0, mag);
  return diff < 0 ? -v : v;
}

/* one thread per CDEF block, reading the deblocked snapshot */
__global__ void k_cdef(FilterArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.plan.n_cdef) return;
  const DevCdef c = a.plan.cdef[i];
  const int dirs[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
                             {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},  {{1, 0}, {2, 0}}, {{1, 0}, {2, -1}}};
  const int uvdir[2][2][8] = {{{0, 1, 2, 3, 4, 5, 6, 7}, {1, 2, 2, 2, 3, 4, 6, 0}},
                              {{7, 0, 2, 4, 5, 6, 6, 6}, {0, 1, 2, 3, 4, 5, 6, 7}}};
  const int pri_taps[2][2] = {{4, 2}, {3, 3}};
  const int sec_taps[2][2] = {{2, 1}, {2, 1}};
  const Layout& L = a.lay;
  const int p = c.plane, uv = p > 0;
  const int bd = L.bit_depth, cs = bd - 8;
  const int st = L.stride[p];
  const int w = uv ? (8 >> L.ss_x) : 8, h = uv ? (8 >> L.ss_y) : 8;
  const int pw = L.grid_w[p], ph = L.grid_h[p];
  const uint16_t* src = a.scratch + L.plane_off[p] + (size_t)c.y * st + c.x;
  uint16_t* dst = a.frame + L.plane_off[p] + (size_t)c.y * st + c.x;
  int pri = c.pri << cs, sec = c.sec << cs;
  const int damping = c.damping + cs - uv;
  int dir = c.dir;
  if (dir < 0) {
    int yd = 0, var = 0;
    const int lx = uv ? c.x << L.ss_x : c.x, ly = uv ? c.y << L.ss_y : c.y;
    cdef_direction(a.scratch + L.plane_off[0] + (size_t)ly * L.stride[0] + lx, L.stride[0], bd, &yd, &var);
    if (uv) {
      dir = pri == 0 ? 0 : uvdir[L.ss_x][L.ss_y][yd];
    } else {
      const int vs = (var >> 6) ? min(flog2(var >> 6), 12) : 0;
      dir = pri == 0 ? 0 : yd;
      pri = var ? ((pri * (4 + vs) + 8) >> 4) : 0;
    }
  }
  if (pri == 0 && sec == 0) {
    for (int r = 0; r < h && c.y + r < ph; ++r)
      for (int k = 0; k < w && c.x + k < pw; ++k) dst[r * st + k] = src[r * st + k];
    return;
  }
  const int ts = (pri >> cs) & 1;
  for (int r = 0; r < h; ++r)
    for (int k = 0; k < w; ++k) {
      if (c.x + k >= pw || c.y + r >= ph) continue;
      const int x = src[r * st + k];
      int sum = 0, lo = x, hi = x;
      for (int t = 0; t < 2; ++t)
        for (int sg = -1; sg <= 1; sg += 2) {
          int dy = sg * dirs[dir][t][0], dx = sg * dirs[dir][t][1];
          int px = c.x + k + dx, py = c.y + r + dy;
          if (px >= 0 && px < pw && py >= 0 && py < ph) {
            const int v = src[(r + dy) * st + k + dx];
            sum += pri_taps[ts][t] * constrain(v - x, pri, damping);
            lo = min(lo, v);
            hi = max(hi, v);
          }
          for (int o = -2; o <= 2; o += 4) {
            const int sd = (dir + o) & 7;
            dy = sg * dirs[sd][t][0];
            dx = sg * dirs[sd][t][1];
            px = c.x + k + dx;
            py = c.y + r + dy;
            if (px >= 0 && px < pw && py >= 0 && py < ph) {
              const int v = src[(r + dy) * st + k + dx];
              sum += sec_taps[ts][t] * constrain(v - x, sec, damping);
              lo = min(lo, v);
              hi = max(hi, v);
            }
          }
        }
      dst[r * st + k] = (uint16_t)clampi(clampi(x + ((8 + sum - (sum < 0)) >> 4), lo, hi), 0, a.maxv);
    }
}

/* ---- superres (av1gpu_kernels.h:2387-2402) ------------------------------ */

struct UpscalePlane {
  int step[3], initial[3];
};

__global__ void k_upscale(FilterArgs a, UpscalePlane u, const uint16_t* src, uint16_t* dst) {
  const int p = blockIdx.z;
  const Layout& L = a.lay;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= L.out_w[p] || y >= L.out_h[p]) return;
  const int* uf = SPEC(AV1T_OFF_UPSCALE);
  const uint16_t* row = src + L.plane_off[p] + (size_t)y * L.stride[p];
  const int sx = -(1 << 14) + u.initial[p] + x * u.step[p];
  const int px = sx >> 14;
  const int* k = uf + ((sx & ((1 << 14) - 1)) >> 8) * 8;
  const int maxx = L.plane_w[p] - 1;
  int sum = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) sum += row[clampi(px + t - 3, 0, maxx)] * k[t];
  dst[L.plane_off[p] + (size_t)y * L.stride[p] + x] = (uint16_t)clampi((sum + 64) >> 7, 0, a.maxv);
}

/* ---- loop restoration (av1gpu_kernels.h:2832-2982) ---------------------- */

struct LrCtx {
  const uint16_t* cdef;
  const uint16_t* curr;
  int stride, pw, ph, s0, s1;
  __device__ __forceinline__ int src(int x, int y) const {
    x = clampi(x, 0, pw - 1);
    y = clampi(y, 0, ph - 1);
    if (y < s0) return curr[(size_t)max(y, s0 - 2) * stride + x];
    if (y > s1) return curr[(size_t)min(y, s1 + 2) * stride + x];
    return cdef[(size_t)y * stride + x];
  }
};

__device__ int wiener_px(const LrCtx& c, const DevLr& j, int bd, int x, int y) {
  const int r0 = bd == 12 ? 5 : 3, r1 = bd == 12 ? 9 : 11;
  const int off = 1 << (bd + 7 - r0 - 1);
  const int lim = (1 << (bd + 1 + 7 - r0)) - 1;
  int vt[7], ht[7];
  for (int t = 0; t < 3; ++t) {
    vt[t] = vt[6 - t] = j.vf[t];
    ht[t] = ht[6 - t] = j.hf[t];
  }
  vt[3] = 128 - 2 * (j.vf[0] + j.vf[1] + j.vf[2]);
  ht[3] = 128 - 2 * (j.hf[0] + j.hf[1] + j.hf[2]);
  int sum = 0;
  for (int t = 0; t < 7; ++t) {
    int s = 0;
    for (int k = 0; k < 7; ++k) s += ht[k] * c.src(x + k - 3, y + t - 3);
    sum += vt[t] * clampi(rsr64(s, r0), -off, lim - off);
  }
  return rsr64(sum, r1);
}

__device__ void sgr_ab(const LrCtx& c, int bd, int x, int y, int d, int eps, int* ao, int* bo) {
  const int r = d >> 1, n = d * d;
  const int64_t n2e = (int64_t)n * n * eps;
  const int64_t mt = (((int64_t)1 << 20) + n2e / 2) / n2e;
  int64_t sa = 0, sb = 0;
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const int v = c.src(x + dx, y + dy);
      sa += (int64_t)v * v;
      sb += v;
    }
  const int64_t braw = sb;
  if (bd > 8) {
    sa = rsr64(sa, 2 * (bd - 8));
    sb = rsr64(sb, bd - 8);
  }
  int64_t p = sa * n - sb * sb;
  if (p < 0) p = 0;
  const int z = rsr64(p * mt, 20);
  const int a2 = z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1));
  const int inv = ((1 << 12) + n / 2) / n;
  *ao = a2;
  *bo = rsr64((int64_t)(256 - a2) * braw * inv, 12);
}

__device__ int sgr_px(const LrCtx& c, const DevLr& j, int bd, int x, int y) {
  const int u = c.src(x, y);
  const int w0 = j.xqd[0], w1 = j.xqd[1], w2 = (1 << 7) - w0 - w1;
  int flt[2] = {0, 0};
  for (int pass = 0; pass < 2; ++pass) {
    const int r = j.r[pass];
    if (r == 0) continue;
    int a = 0, b = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int wgt = pass == 0 ? (((y + dy) & 1) ? (dx == 0 ? 6 : 5) : 0) : ((dx == 0 || dy == 0) ? 4 : 3);
        if (!wgt) continue;
        int a2, b2;
        sgr_ab(c, bd, x + dx, y + dy, 2 * r + 1, j.eps[pass], &a2, &b2);
        a += wgt * a2;
        b += wgt * b2;
      }
    const int sh = (pass == 0 && (y & 1)) ? 4 : 5;
    flt[pass] = rsr64((int64_t)a * u + b, 8 + sh - 4);
  }
  const int us = u << 4;
  int64_t v = (int64_t)w1 * us;
  v += (int64_t)w0 * (j.r[0] ? flt[0] : us);
  v += (int64_t)w2 * (j.r[1] ? flt[1] : us);
  return rsr64(v, 4 + 7);
}

/* one thread per restored sample; job found by binary search on px0 */
__global__ void k_lr(FilterArgs a, const uint16_t* cdef, const uint16_t* curr) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.plan.lr_px) return;
  int lo = 0, hi = a.plan.n_lr - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (a.plan.lr[m].px0 <= g)
      lo = m;
    else
      hi = m - 1;
  }
  const DevLr& j = a.plan.lr[lo];
  const int k = g - j.px0;
  const int y = j.y + k / j.w, x = j.x + k % j.w;
  const Layout& L = a.lay;
  const int p = j.plane;
  LrCtx c;
  c.cdef = cdef + L.plane_off[p];
  c.curr = curr + L.plane_off[p];
  c.stride = L.stride[p];
  c.pw = L.out_w[p];
  c.ph = L.out_h[p];
  c.s0 = j.stripe0;
  c.s1 = j.stripe1;
  const int v = j.type == 1 ? wiener_px(c, j, L.bit_depth, x, y) : sgr_px(c, j, L.bit_depth, x, y);
  a.frame[L.plane_off[p] + (size_t)y * L.stride[p] + x] = (uint16_t)clampi(v, 0, a.maxv);
}

/* ---- NV12 and RGB packs (av1gpu_kernels.h:3379-3472) -------------------- */

/* one thread per output byte pair */
__global__ void k_nv12(Layout L, const uint16_t* src, uint8_t* out, int pitch) {
  const int x2 = (blockIdx.x * blockDim.x + threadIdx.x) * 2;
  const int row = blockIdx.y;
  if (x2 >= pitch) return;
  const int sh = L.bit_depth - 8;
  const int w = L.width, h = L.height;
  uint8_t* d = out + (size_t)row * pitch + x2;
  if (row < h) {
    const uint16_t* s = src + L.plane_off[0] + (size_t)row * L.stride[0];
    d[0] = x2 < w ? (uint8_t)(s[x2] >> sh) : 0;
    if (x2 + 1 < pitch) d[1] = x2 + 1 < w ? (uint8_t)(s[x2 + 1] >> sh) : 0;
    return;
  }
  const int cr = row - h;
  const int cw = (w + (1 << L.ss_x) - 1) >> L.ss_x;
  const int c = x2 >> 1;
  if (c < cw) {
    d[0] = (uint8_t)(src[L.plane_off[1] + (size_t)cr * L.stride[1] + c] >> sh);
    d[1] = (uint8_t)(src[L.plane_off[2] + (size_t)cr * L.stride[2] + c] >> sh);
  } else {
    d[0] = 0;
    if (x2 + 1 < pitch) d[1] = 0;
  }
}

struct ColorM {
  int m[8];
  int o[3];
};

__global__ void k_rgb(Layout L, const uint16_t* src, uint8_t* out, ColorM cm) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= L.width) return;
  const int yy = src[L.plane_off[0] + (size_t)y * L.stride[0] + x];
  const size_t co = (size_t)(y >> L.ss_y) * L.stride[1] + (x >> L.ss_x);
  const int uu = src[L.plane_off[1] + co], vv = src[L.plane_off[2] + co];
  const int r = (cm.m[0] * yy + cm.m[1] * vv + cm.o[0]) >> 16;
  const int g = (cm.m[2] * yy + cm.m[3] * uu + cm.m[4] * vv + cm.o[1]) >> 16;
  const int b = (cm.m[5] * yy + cm.m[6] * uu + cm.o[2]) >> 16;
  uint8_t* d = out + ((size_t)y * L.width + x) * 3;
  d[0] = (uint8_t)clampi(r, 0, 255);
  d[1] = (uint8_t)clampi(g, 0, 255);
  d[2] = (uint8_t)clampi(b, 0, 255);
}

}  // namespace

hipError_t launch_deblock(const FilterArgs& a, hipStream_t s, int* launches) {
  if (a.plan.n_deblock_v > 0 && a.plan.lines_v > 0) {
    hipLaunchKernelGGL(k_deblock, dim3((a.plan.lines_v + 255) / 256), dim3(256), 0, s, a, 0,
                       a.plan.n_deblock_v, a.plan.lines_v);
    ++*launches;
  }
  if (a.plan.n_deblock_h > 0 && a.plan.lines_h > 0) {
    hipLaunchKernelGGL(k_deblock, dim3((a.plan.lines_h + 255) / 256), dim3(256), 0, s, a,
                       a.plan.n_deblock_v, a.plan.n_deblock_h, a.plan.lines_h);
    ++*launches;
  }
  return hipGetLastError();
}

hipError_t launch_cdef(const FilterArgs& a, hipStream_t s, int* launches) {
  if (a.plan.n_cdef <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cdef, dim3((a.plan.n_cdef + 127) / 128), dim3(128), 0, s, a);
  ++*launches;
  return hipGetLastError();
}

void superres_params(int dw, int uw, int* step, int* initial) {
  /* av1gpu_kernels.h:2375-2383 */
  const int st = ((dw << 14) + (uw / 2)) / uw;
  const int err = uw * st - (dw << 14);
  const int sub = (-((uw - dw) << 13) + uw / 2) / uw + (1 << 7) - err / 2;
  *step = st;
  *initial = sub & ((1 << 14) - 1);
}

hipError_t launch_superres(const FilterArgs& a, const uint16_t* src, uint16_t* dst, hipStream_t s,
                           int* launches) {
  UpscalePlane u;
  for (int p = 0; p < 3; ++p) superres_params(a.lay.plane_w[p], a.lay.out_w[p], &u.step[p], &u.initial[p]);
  hipLaunchKernelGGL(k_upscale, dim3((a.lay.out_w[0] + 127) / 128, a.lay.out_h[0], 3), dim3(128), 0, s,
                     a, u, src, dst);
  ++*launches;
  return hipGetLastError();
}

hipError_t launch_lr(const FilterArgs& a, const uint16_t* cdef, const uint16_t* curr, hipStream_t s,
                     int* launches) {
  if (a.plan.lr_px <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lr, dim3((a.plan.lr_px + 255) / 256), dim3(256), 0, s, a, cdef, curr);
  ++*launches;
  return hipGetLastError();
}

hipError_t launch_nv12(const Layout& lay, const uint16_t* src, uint8_t* out, int pitch, hipStream_t s) {
  const int rows = lay.height + ((lay.height + (1 << lay.ss_y) - 1) >> lay.ss_y);
  hipLaunchKernelGGL(k_nv12, dim3((pitch / 2 + 127) / 128, rows), dim3(128), 0, s, lay, src, out, pitch);
  return hipGetLastError();
}

hipError_t launch_rgb(const Layout& lay, const uint16_t* src, uint8_t* out, const int m[8], const int o[3],
                      hipStream_t s) {
  ColorM cm;
  for (int i = 0; i < 8; ++i) cm.m[i] = m[i];
  for (int i = 0; i < 3; ++i) cm.o[i] = o[i];
  hipLaunchKernelGGL(k_rgb, dim3((lay.width + 127) / 128, lay.height), dim3(128), 0, s, lay, src, out, cm);
  return hipGetLastError();
}

}  // namespace av1hip
