/*
 * CPU oracle, pixel maths (TEST INFRASTRUCTURE ONLY - see av1o.h).
 *
 * Each function restates one function of the reference's
 * kernels/av1gpu_kernels.h (cited per function); the arithmetic, rounding and
 * clamping are the reference's, including its 64 bit products.
 */
#include <string.h>

#include "av1o_internal.h"

/* ======================================================================== */
/* inverse transform                                                        */
/* ======================================================================== */

/* av1gpu_kernels.h:472-479 */
int av1o_cos128(int angle) {
  const int* c = SPEC(AV1O_OFF_COS128);
  const int a = angle & 255;
  if (a <= 64) return c[a];
  if (a <= 128) return -c[128 - a];
  if (a <= 192) return -c[a - 128];
  return c[256 - a];
}
static int sin128(int angle) { return av1o_cos128(angle - 64); }

/* av1gpu_kernels.h:511-516 */
static int bitrev(int bits, int x) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

/* B(a, b, angle, flip) - av1gpu_kernels.h:521-545: rotation with 64 bit
 * products rounded by 12, then an exchange when flipped. At angles where cos
 * and sin have the same magnitude the reference forms the 32 bit sum first. */
static void bfly(int32_t* t, int a, int b, int angle, int flip) {
  int64_t x, y;
  if ((angle & 63) == 32) {
    const int32_t sa = (int32_t)((uint32_t)t[a] + (uint32_t)t[b]);
    const int32_t da = (int32_t)((uint32_t)t[a] - (uint32_t)t[b]);
    const int32_t v = (angle & 64) ? sa : da;
    const int32_t w = (angle & 64) ? (int32_t)(0u - (uint32_t)da) : sa;
    const int64_t c = av1o_cos128(angle);
    x = (int64_t)v * c;
    y = (int64_t)w * c;
  } else {
    const int64_t c = av1o_cos128(angle), s = sin128(angle);
    x = (int64_t)t[a] * c - (int64_t)t[b] * s;
    y = (int64_t)t[a] * s + (int64_t)t[b] * c;
  }
  const int32_t ra = rnd2_64(x, 12), rb = rnd2_64(y, 12);
  t[a] = flip ? rb : ra;
  t[b] = flip ? ra : rb;
}

/* H(a, b, flip, r) - av1gpu_kernels.h:555-568: sum and difference clamped to
 * r bits; a flipped Hadamard is H(b, a, 0, r). */
static void hada(int32_t* t, int a, int b, int flip, int r) {
  const int lo = -(1 << (r - 1)), hi = (1 << (r - 1)) - 1;
  const int32_t p = flip ? t[b] : t[a];
  const int32_t q = flip ? t[a] : t[b];
  t[flip ? b : a] = iclip(p + q, lo, hi);
  t[flip ? a : b] = iclip(p - q, lo, hi);
}

/* inverse DCT of 2^n points, n = 2..6 - av1gpu_kernels.h:571-652 (the
 * specification's butterfly network, 7.13.2.3) */
static void idct(int32_t* t, int n, int r) {
  int32_t c[64];
  const int len = 1 << n;
  for (int i = 0; i < len; ++i) c[i] = t[bitrev(n, i)];
  memcpy(t, c, sizeof(int32_t) * (size_t)len);
  if (n == 6)
    for (int i = 0; i < 16; ++i) bfly(t, 32 + i, 63 - i, 63 - 4 * bitrev(4, i), 0);
  if (n >= 5)
    for (int i = 0; i < 8; ++i) bfly(t, 16 + i, 31 - i, 6 + (bitrev(3, 7 - i) << 3), 0);
  if (n == 6)
    for (int i = 0; i < 16; ++i) hada(t, 32 + 2 * i, 33 + 2 * i, i & 1, r);
  if (n >= 4)
    for (int i = 0; i < 4; ++i) bfly(t, 8 + i, 15 - i, 12 + (bitrev(2, 3 - i) << 4), 0);
  if (n >= 5)
    for (int i = 0; i < 8; ++i) hada(t, 16 + 2 * i, 17 + 2 * i, i & 1, r);
  if (n == 6)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 2; ++j)
        bfly(t, 62 - 4 * i - j, 33 + 4 * i + j, 60 - 16 * bitrev(2, i) + 64 * j, 1);
  if (n >= 3)
    for (int i = 0; i < 2; ++i) bfly(t, 4 + i, 7 - i, 56 - 32 * i, 0);
  if (n >= 4)
    for (int i = 0; i < 4; ++i) hada(t, 8 + 2 * i, 9 + 2 * i, i & 1, r);
  if (n >= 5)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        bfly(t, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1);
  if (n == 6)
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 2; ++j) hada(t, 32 + 4 * i + j, 35 + 4 * i - j, i & 1, r);
  for (int i = 0; i < 2; ++i) bfly(t, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i);
  if (n >= 3)
    for (int i = 0; i < 2; ++i) hada(t, 4 + 2 * i, 5 + 2 * i, i, r);
  if (n >= 4)
    for (int i = 0; i < 2; ++i) bfly(t, 14 - i, 9 + i, 48 + 64 * i, 1);
  if (n >= 5)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 2; ++j) hada(t, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, r);
  if (n == 6)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 4; ++j)
        bfly(t, 61 - 8 * i - j, 34 + 8 * i + j, 56 - 32 * i + (j >> 1) * 64, 1);
  for (int i = 0; i < 2; ++i) hada(t, i, 3 - i, 0, r);
  if (n >= 3) bfly(t, 6, 5, 32, 1);
  if (n >= 4)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) hada(t, 8 + 4 * i + j, 11 + 4 * i - j, i, r);
  if (n >= 5)
    for (int i = 0; i < 4; ++i) bfly(t, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1);
  if (n == 6)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) hada(t, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, r);
  if (n >= 3)
    for (int i = 0; i < 4; ++i) hada(t, i, 7 - i, 0, r);
  if (n >= 4)
    for (int i = 0; i < 2; ++i) bfly(t, 13 - i, 10 + i, 32, 1);
  if (n >= 5)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 4; ++j) hada(t, 16 + 8 * i + j, 23 + 8 * i - j, i, r);
  if (n == 6)
    for (int i = 0; i < 8; ++i) bfly(t, 59 - i, 36 + i, i < 4 ? 48 : 112, 1);
  if (n >= 4)
    for (int i = 0; i < 8; ++i) hada(t, i, 15 - i, 0, r);
  if (n >= 5)
    for (int i = 0; i < 4; ++i) bfly(t, 27 - i, 20 + i, 32, 1);
  if (n == 6)
    for (int i = 0; i < 8; ++i) {
      hada(t, 32 + i, 47 - i, 0, r);
      hada(t, 48 + i, 63 - i, 1, r);
    }
  if (n >= 5)
    for (int i = 0; i < 16; ++i) hada(t, i, 31 - i, 0, r);
  if (n == 6)
    for (int i = 0; i < 8; ++i) bfly(t, 55 - i, 40 + i, 32, 1);
  if (n == 6)
    for (int i = 0; i < 32; ++i) hada(t, i, 63 - i, 0, r);
}

/* ADST input / output permutations - av1gpu_kernels.h:655-679 */
static void adst_in(int32_t* t, int n) {
  int32_t c[16];
  const int len = 1 << n;
  memcpy(c, t, sizeof(int32_t) * (size_t)len);
  for (int i = 0; i < len; ++i) t[i] = c[(i & 1) ? i - 1 : len - 1 - i];
}
static void adst_out(int32_t* t, int n) {
  int32_t c[16];
  const int len = 1 << n;
  memcpy(c, t, sizeof(int32_t) * (size_t)len);
  for (int i = 0; i < len; ++i) {
    const int b3 = (i >> 3) & 1, b2 = (i >> 2) & 1, b1 = (i >> 1) & 1, b0 = i & 1;
    const int idx = (((b0 ^ b1) << 3) | ((b1 ^ b2) << 2) | ((b2 ^ b3) << 1) | b3) >> (4 - n);
    t[i] = (i & 1) ? -c[idx] : c[idx];
  }
}

/* ADST4 with the SINPI constants, 64 bit - av1gpu_kernels.h:682-712 */
static void adst4(int32_t* t) {
  const int64_t k1 = 1321, k2 = 2482, k3 = 3344, k4 = 3803;
  int64_t s0 = k1 * t[0], s1 = k2 * t[0], s2 = k3 * t[1], s3 = k4 * t[2];
  const int64_t s4 = k1 * t[2], s5 = k2 * t[3], s6 = k4 * t[3];
  const int32_t b7 = (t[0] - t[2]) + t[3];
  s0 += s3;
  s1 -= s4;
  s3 = s2;
  s2 = k3 * b7;
  s0 += s5;
  s1 -= s6;
  const int64_t x0 = s0 + s3, x1 = s1 + s3, x2 = s2, x3 = s0 + s1 - s3;
  t[0] = rnd2_64(x0, 12);
  t[1] = rnd2_64(x1, 12);
  t[2] = rnd2_64(x2, 12);
  t[3] = rnd2_64(x3, 12);
}

/* av1gpu_kernels.h:714-724 */
static void adst8(int32_t* t, int r) {
  adst_in(t, 3);
  for (int i = 0; i < 4; ++i) bfly(t, 2 * i, 2 * i + 1, 60 - 16 * i, 1);
  for (int i = 0; i < 4; ++i) hada(t, i, 4 + i, 0, r);
  for (int i = 0; i < 2; ++i) bfly(t, 4 + 3 * i, 5 + i, 48 - 32 * i, 1);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) hada(t, 4 * j + i, 2 + 4 * j + i, 0, r);
  for (int i = 0; i < 2; ++i) bfly(t, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(t, 3);
}

/* av1gpu_kernels.h:726-744 */
static void adst16(int32_t* t, int r) {
  adst_in(t, 4);
  for (int i = 0; i < 8; ++i) bfly(t, 2 * i, 2 * i + 1, 62 - 8 * i, 1);
  for (int i = 0; i < 8; ++i) hada(t, i, 8 + i, 0, r);
  for (int i = 0; i < 2; ++i) {
    bfly(t, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1);
    bfly(t, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1);
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) hada(t, 8 * j + i, 4 + 8 * j + i, 0, r);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) bfly(t, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j) hada(t, 4 * j + i, 2 + 4 * j + i, 0, r);
  for (int i = 0; i < 4; ++i) bfly(t, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(t, 4);
}

/* identity scaling - av1gpu_kernels.h:759-770 */
static void iidentity(int32_t* t, int n) {
  if (n == 2) {
    for (int i = 0; i < 4; ++i) t[i] = rnd2_64((int64_t)t[i] * 5793, 12);
  } else if (n == 3) {
    for (int i = 0; i < 8; ++i) t[i] *= 2;
  } else if (n == 4) {
    for (int i = 0; i < 16; ++i) t[i] = rnd2_64((int64_t)t[i] * 11586, 12);
  } else {
    for (int i = 0; i < 32; ++i) t[i] *= 4;
  }
}

/* Walsh-Hadamard - av1gpu_kernels.h:773-790 */
static void iwht(int32_t* t, int shift) {
  int32_t a = t[0] >> shift, c = t[1] >> shift, d = t[2] >> shift, b = t[3] >> shift;
  a += c;
  d -= b;
  const int32_t e = (a - d) >> 1;
  b = e - b;
  c = e - c;
  a -= b;
  d += c;
  t[0] = a;
  t[1] = b;
  t[2] = c;
  t[3] = d;
}

/* 1D type per direction and flips - av1gpu_kernels.h:800-851 */
static int kind_rows(int tx) {
  static const int k[16] = {0, 0, 1, 1, 0, 1, 1, 1, 1, 2, 2, 0, 2, 1, 2, 1};
  return k[tx & 15];
}
static int kind_cols(int tx) {
  static const int k[16] = {0, 1, 0, 1, 1, 0, 1, 1, 1, 2, 0, 2, 1, 2, 1, 2};
  return k[tx & 15];
}
static int flips_ud(int tx) { return tx == 4 || tx == 8 || tx == 6 || tx == 14; }
static int flips_lr(int tx) { return tx == 5 || tx == 7 || tx == 6 || tx == 15; }

/* av1gpu_kernels.h:853-860 */
void av1o_tx1d(int32_t* t, int n, int kind, int r) {
  if (kind == 0) {
    idct(t, n, r);
  } else if (kind == 1) {
    if (n == 2)
      adst4(t);
    else if (n == 3)
      adst8(t, r);
    else
      adst16(t, r);
  } else {
    iidentity(t, n);
  }
}

/* row pass - av1gpu_kernels.h:869-904 */
void o_itx_row(const o_itx* j, const int32_t* coef, int32_t* tmp, int y) {
  int32_t t[64];
  const int lw = ilog2_ceil(j->w), lh = ilog2_ceil(j->h);
  const int col_r = imax(j->bitdepth + 6, 16);
  const int lo = -(1 << (col_r - 1)), hi = (1 << (col_r - 1)) - 1;
  const int cs = j->w < 32 ? j->w : 32;
  for (int x = 0; x < j->w; ++x)
    t[x] = (y < 32 && x < cs) ? coef[j->coef + (size_t)y * cs + x] : 0;
  if (iabs(lw - lh) == 1)
    for (int x = 0; x < j->w; ++x) t[x] = rnd2_64((int64_t)t[x] * 2896, 12);
  if (j->lossless)
    iwht(t, 2);
  else
    av1o_tx1d(t, lw, kind_rows(j->tx_type), j->bitdepth + 8);
  for (int x = 0; x < j->w; ++x) {
    const int v = j->lossless ? t[x] : rnd2(t[x], j->row_shift);
    tmp[j->tmp + (size_t)y * j->w + x] = iclip(v, lo, hi);
  }
}

/* column pass + add - av1gpu_kernels.h:907-930 */
void o_itx_col(const o_itx* j, const int32_t* tmp, px_t* frame, int x, int maxv) {
  int32_t t[64];
  if (j->eob == 0) return;
  const int lh = ilog2_ceil(j->h);
  for (int i = 0; i < j->h; ++i) t[i] = tmp[j->tmp + (size_t)i * j->w + x];
  if (j->lossless)
    iwht(t, 0);
  else
    av1o_tx1d(t, lh, kind_cols(j->tx_type), imax(j->bitdepth + 6, 16));
  const int ud = flips_ud(j->tx_type), lr = flips_lr(j->tx_type);
  for (int i = 0; i < j->h; ++i) {
    const int v = j->lossless ? t[i] : rnd2(t[i], 4);
    const int yy = ud ? j->h - 1 - i : i;
    const int xx = lr ? j->w - 1 - x : x;
    px_t* p = frame + j->dst + (size_t)yy * j->stride + xx;
    *p = (px_t)iclip((int)*p + v, 0, maxv);
  }
}

/* av1gpu_kernels.h:933-937 */
void o_residual_add(const o_itx* j, const int32_t* res, px_t* frame, int maxv) {
  for (int y = 0; y < j->h; ++y)
    for (int x = 0; x < j->w; ++x) {
      px_t* p = frame + j->dst + (size_t)y * j->stride + x;
      *p = (px_t)iclip((int)*p + res[j->coef + (size_t)y * j->w + x], 0, maxv);
    }
}

/* ======================================================================== */
/* intra prediction                                                         */
/* ======================================================================== */

#define EDGE_N 264
#define EDGE_OFF 4

/* AboveRow / LeftCol with the specification's fallbacks -
 * av1gpu_kernels.h:1108-1151 */
static void intra_edges(const o_intra* j, const px_t* frame, int* above, int* left) {
  const int w = j->w, h = j->h, s = j->stride;
  const px_t* p = frame + j->dst;
  const int mid = 1 << (j->bitdepth - 1);
  if (!j->have_above) {
    for (int i = 0; i < w + h; ++i) above[i] = j->have_left ? (int)p[-1] : mid - 1;
  } else {
    const int last = imin(j->max_x, j->x + (j->have_ar ? 2 * w : w) - 1);
    for (int i = 0; i < w + h; ++i) above[i] = p[-s + (imin(j->x + i, last) - j->x)];
  }
  if (!j->have_left) {
    for (int i = 0; i < w + h; ++i) left[i] = j->have_above ? (int)p[-s] : mid + 1;
  } else {
    const int last = imin(j->max_y, j->y + (j->have_bl ? 2 * h : h) - 1);
    for (int i = 0; i < w + h; ++i) left[i] = p[(ptrdiff_t)(imin(j->y + i, last) - j->y) * s - 1];
  }
  int corner;
  if (j->have_above && j->have_left)
    corner = p[-s - 1];
  else if (j->have_above)
    corner = p[-s];
  else if (j->have_left)
    corner = p[-1];
  else
    corner = mid;
  above[-1] = corner;
  left[-1] = corner;
}

/* av1gpu_kernels.h:1157-1173 */
static void edge_filter(int* buf, int sz, int strength) {
  int e[EDGE_N];
  if (strength <= 0) return;
  if (sz > EDGE_N) sz = EDGE_N;
  const int* k = SPEC(AV1O_OFF_EDGE_KERNEL) + 5 * (strength - 1);
  for (int i = 0; i < sz; ++i) e[i] = buf[i - 1];
  for (int i = 1; i < sz; ++i) {
    int s = 0;
    for (int t = 0; t < 5; ++t) s += k[t] * e[iclip(i - 2 + t, 0, sz - 1)];
    buf[i - 1] = (s + 8) >> 4;
  }
}

/* av1gpu_kernels.h:1177-1191 */
static void edge_upsample(int* buf, int n, int maxv) {
  int d[EDGE_N];
  if (n > EDGE_N - 4) n = EDGE_N - 4;
  d[0] = buf[-1];
  for (int i = -1; i < n; ++i) d[i + 2] = buf[i];
  d[n + 2] = buf[n - 1];
  buf[-2] = d[0];
  for (int i = 0; i < n; ++i) {
    const int s = -d[i] + 9 * d[i + 1] + 9 * d[i + 2] - d[i + 3];
    buf[2 * i - 1] = iclip(rnd2(s, 4), 0, maxv);
    buf[2 * i] = d[i + 2];
  }
}

/* av1gpu_kernels.h:1195-1231 */
static int edge_strength(int w, int h, int ftype, int delta) {
  const int d = iabs(delta), blk = w + h;
  if (ftype == 0) {
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

/* av1gpu_kernels.h:1233-1239 */
static int upsample_sel(int w, int h, int ftype, int delta) {
  const int d = iabs(delta);
  if (d <= 0 || d >= 40) return 0;
  return ftype == 0 ? (w + h <= 16) : (w + h <= 8);
}

static int sm_weight(int n, int i) { return SPEC(AV1O_OFF_SM_WEIGHTS)[n - 4 + i]; }

/* directional - av1gpu_kernels.h:1245-1313 */
static void intra_directional(const o_intra* j, const int* above, const int* left, int pa,
                              int ua, int ul, px_t* frame, int maxv) {
  const int w = j->w, h = j->h, s = j->stride;
  const int* dr = SPEC(AV1O_OFF_DR_DERIV);
  int dx = 0, dy = 0;
  if (pa < 90)
    dx = dr[pa];
  else if (pa > 90 && pa < 180)
    dx = dr[180 - pa];
  if (pa > 90 && pa < 180)
    dy = dr[pa - 90];
  else if (pa > 180)
    dy = dr[270 - pa];
  px_t* out = frame + j->dst;
  for (int i = 0; i < h; ++i) {
    for (int c = 0; c < w; ++c) {
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
      out[(size_t)i * s + c] = (px_t)iclip(v, 0, maxv);
    }
  }
}

/* DC - av1gpu_kernels.h:1430-1447 */
static int intra_dc(const o_intra* j, const int* above, const int* left) {
  const int w = j->w, h = j->h;
  int sum = 0;
  if (j->have_above && j->have_left) {
    for (int k = 0; k < w; ++k) sum += above[k];
    for (int k = 0; k < h; ++k) sum += left[k];
    return (sum + ((w + h) >> 1)) / (w + h);
  }
  if (j->have_left) {
    for (int k = 0; k < h; ++k) sum += left[k];
    return (sum + (h >> 1)) >> ilog2_ceil(h);
  }
  if (j->have_above) {
    for (int k = 0; k < w; ++k) sum += above[k];
    return (sum + (w >> 1)) >> ilog2_ceil(w);
  }
  return 1 << (j->bitdepth - 1);
}

/* recursive filter intra - av1gpu_kernels.h:1392-1427 */
static void intra_filter(const o_intra* j, const int* above, const int* left, px_t* frame,
                         int maxv) {
  const int s = j->stride;
  px_t* out = frame + j->dst;
  const int* taps = SPEC(AV1O_OFF_FILTER_INTRA) + 56 * j->fi_mode;
  for (int r2 = 0; r2 < (j->h >> 1); ++r2) {
    for (int c4 = 0; c4 < (j->w >> 2); ++c4) {
      int p[7];
      for (int i = 0; i < 5; ++i) {
        if (r2 == 0)
          p[i] = above[(c4 << 2) + i - 1];
        else if (c4 == 0 && i == 0)
          p[i] = left[(r2 << 1) - 1];
        else
          p[i] = out[(size_t)((r2 << 1) - 1) * s + (c4 << 2) + i - 1];
      }
      for (int i = 5; i < 7; ++i)
        p[i] = c4 == 0 ? left[(r2 << 1) + i - 5]
                       : (int)out[(size_t)((r2 << 1) + i - 5) * s + (c4 << 2) - 1];
      for (int k = 0; k < 8; ++k) {
        int acc = 0;
        for (int i = 0; i < 7; ++i) acc += taps[k * 7 + i] * p[i];
        out[(size_t)((r2 << 1) + (k >> 2)) * s + (c4 << 2) + (k & 3)] =
            (px_t)iclip(rnd2(acc, 4), 0, maxv);
      }
    }
  }
}

/* the whole prediction of one transform block - av1gpu_kernels.h:1578-1677,
 * with the directional edge preparation of 1452-1481 */
void o_intra_block(const o_intra* j, px_t* frame, const uint8_t* palette, int maxv) {
  int abuf[EDGE_N], lbuf[EDGE_N];
  int* above = abuf + EDGE_OFF;
  int* left = lbuf + EDGE_OFF;
  const int w = j->w, h = j->h, s = j->stride;
  px_t* out = frame + j->dst;

  if (j->palette) {
    const uint8_t* map = palette + j->palette_off;
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
        out[(size_t)y * s + x] = (px_t)j->palette_color[map[(size_t)y * j->palette_stride + x]];
    return;
  }
  if (j->mode == AV1H_PRED_CFL) {
    intra_edges(j, frame, above, left);
    const int dc = intra_dc(j, above, left);
    const int sx = j->ss_x, sy = j->ss_y;
    const int xlim = j->luma_w - (1 << sx), ylim = j->luma_h - (1 << sy);
    int64_t sum = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int shift = ilog2_ceil(w) + ilog2_ceil(h);
      const int avg = pass ? (int)((sum + ((int64_t)1 << (shift - 1))) >> shift) : 0;
      for (int y = 0; y < h; ++y) {
        const int ly = imin((j->y + y) << sy, ylim);
        for (int x = 0; x < w; ++x) {
          const int lx = imin((j->x + x) << sx, xlim);
          int t = 0;
          for (int dy = 0; dy <= sy; ++dy)
            for (int dx = 0; dx <= sx; ++dx)
              t += frame[j->luma + (size_t)(ly + dy) * j->luma_stride + lx + dx];
          const int v = t << (3 - sx - sy);
          if (pass == 0)
            sum += v;
          else
            out[(size_t)y * s + x] =
                (px_t)iclip(dc + rnd2s(j->cfl_alpha * (v - avg), 6), 0, maxv);
        }
      }
    }
    return;
  }
  intra_edges(j, frame, above, left);
  if (j->filter_intra && w >= 4 && h >= 2) {
    intra_filter(j, above, left, frame, maxv);
    return;
  }
  if (j->mode >= 1 && j->mode <= 8) {
    const int pa = SPEC(AV1O_OFF_MODE_TO_ANGLE)[j->mode] + 3 * j->angle;
    int ua = 0, ul = 0;
    if (j->filter_type >= 0 && pa != 90 && pa != 180) {
      if (pa > 90 && pa < 180 && w + h >= 24) {
        const int c = rnd2(left[0] * 5 + above[-1] * 6 + above[0] * 5, 4);
        above[-1] = c;
        left[-1] = c;
      }
      if (j->have_above)
        edge_filter(above, imin(w, j->max_x - j->x + 1) + (pa < 90 ? h : 0) + 1,
                    edge_strength(w, h, j->filter_type, pa - 90));
      if (j->have_left)
        edge_filter(left, imin(h, j->max_y - j->y + 1) + (pa > 180 ? w : 0) + 1,
                    edge_strength(w, h, j->filter_type, pa - 180));
      ua = upsample_sel(w, h, j->filter_type, pa - 90);
      if (ua) edge_upsample(above, w + (pa < 90 ? h : 0), maxv);
      ul = upsample_sel(w, h, j->filter_type, pa - 180);
      if (ul) edge_upsample(left, h + (pa > 180 ? w : 0), maxv);
    }
    intra_directional(j, above, left, pa, ua, ul, frame, maxv);
    return;
  }
  if (j->mode >= AV1H_PRED_SMOOTH && j->mode <= AV1H_PRED_SMOOTH_H) {
    /* av1gpu_kernels.h:1333-1353 */
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        int v;
        if (j->mode == AV1H_PRED_SMOOTH) {
          const int wy = sm_weight(h, y), wx = sm_weight(w, x);
          v = rnd2(wy * above[x] + (256 - wy) * left[h - 1] + wx * left[y] +
                       (256 - wx) * above[w - 1],
                   9);
        } else if (j->mode == AV1H_PRED_SMOOTH_V) {
          const int wy = sm_weight(h, y);
          v = rnd2(wy * above[x] + (256 - wy) * left[h - 1], 8);
        } else {
          const int wx = sm_weight(w, x);
          v = rnd2(wx * left[y] + (256 - wx) * above[w - 1], 8);
        }
        out[(size_t)y * s + x] = (px_t)iclip(v, 0, maxv);
      }
    return;
  }
  if (j->mode == AV1H_PRED_PAETH) {
    /* av1gpu_kernels.h:1364-1382 */
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const int base = above[x] + left[y] - above[-1];
        const int pl = iabs(base - left[y]), pt = iabs(base - above[x]),
                  ptl = iabs(base - above[-1]);
        const int v = (pl <= pt && pl <= ptl) ? left[y] : (pt <= ptl ? above[x] : above[-1]);
        out[(size_t)y * s + x] = (px_t)iclip(v, 0, maxv);
      }
    return;
  }
  const int dc = iclip(intra_dc(j, above, left), 0, maxv);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) out[(size_t)y * s + x] = (px_t)dc;
}

/* ======================================================================== */
/* inter prediction                                                         */
/* ======================================================================== */

/* 4-tap rows for small extents - av1gpu_kernels.h:1696-1701 */
static int filt_row(int cls, int extent) {
  if (extent > 4) return cls;
  if (cls == 0 || cls == 2) return 4;
  if (cls == 1) return 5;
  return cls;
}
static const int* subpel(int row, int phase) { return SPEC(AV1O_OFF_SUBPEL) + (row * 16 + phase) * 8; }

static int fetch(const px_t* p, int stride, int x, int y, int x1, int y1) {
  x = iclip(x, 0, x1);
  y = iclip(y, 0, y1);
  return p[(size_t)y * stride + x];
}

/* one sample of one reference: the unscaled two stage filter
 * (av1gpu_kernels.h:1756-1786) or the scaled one (1809-1857) */
static int mc_tap(const px_t* ref, int stride, int ref_w, int ref_h, int sub_x, int sub_y,
                  int scale_x, int scale_y, int bx, int by, int ew, int eh, int mvx, int mvy,
                  int filt, int filt_v, int x, int y, int compound, int bd) {
  const int round0 = bd >= 12 ? 5 : 3;
  int round1 = compound ? 7 : 11;
  if (bd >= 12 && !compound) round1 -= 2;
  const int fh = filt_row(filt, ew), fv = filt_row(filt_v, eh);
  int cx, cy, xf, yf;
  if (scale_x == (1 << 14) && scale_y == (1 << 14)) {
    const int mx = mvx * 2, my = mvy * 2;
    cx = bx + x + (mx >> (4 + sub_x));
    cy = by + y + (my >> (4 + sub_y));
    xf = (mx >> sub_x) & 15;
    yf = (my >> sub_y) & 15;
  } else {
    const int ox = (bx << 4) + ((2 * mvx) >> sub_x) + 8;
    const int oy = (by << 4) + ((2 * mvy) >> sub_y) + 8;
    const int sx = rnd2s_add(ox * scale_x - (8 << 14), 14 + 4 - 10) + 32;
    const int sy = rnd2s_add(oy * scale_y - (8 << 14), 14 + 4 - 10) + 32;
    const int px = sx + rnd2s_add(scale_x, 4) * x;
    const int py = sy + rnd2s_add(scale_y, 4) * y;
    cx = px >> 10;
    cy = py >> 10;
    xf = (px & 1023) >> 6;
    yf = (py & 1023) >> 6;
  }
  const int* hk = subpel(fh, xf);
  const int* vk = subpel(fv, yf);
  int acc = 0;
  for (int r = 0; r < 8; ++r) {
    int ra = 0;
    for (int i = 0; i < 8; ++i)
      ra += fetch(ref, stride, cx - 3 + i, cy - 3 + r, ref_w - 1, ref_h - 1) * hk[i];
    acc += ((ra + (1 << (round0 - 1))) >> round0) * vk[r];
  }
  return (acc + (1 << (round1 - 1))) >> round1;
}

static int warp_offs(int v) { return iclip(((v + 512) >> 10) + 64, 0, 192); }

/* one sample of a globally warped reference - av1gpu_kernels.h:1919-1958 */
static int warp_tap(int sub_x, int sub_y, int ref_w, int ref_h, const int* wp, const int* abcd,
                    const px_t* ref, int stride, int bx, int by, int x, int y, int bd,
                    int compound) {
  const int* wf = SPEC(AV1O_OFF_WARP);
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
      s += wf[ox * 8 + t] * fetch(ref, stride, ix4 + i2 - 3 + t, iy4 + row, ref_w - 1, ref_h - 1);
    acc += ((s + (1 << (round0 - 1))) >> round0) * wf[oy * 8 + k];
  }
  return (acc + (1 << (round1 - 1))) >> round1;
}

/* wedge master mask - av1gpu_kernels.h:1968-2014 */
static int wedge_1d(int which, int j) {
  static const int even[8] = {1, 4, 11, 27, 46, 58, 62, 63};
  static const int odd[8] = {1, 2, 6, 18, 37, 53, 60, 63};
  static const int vert[6] = {2, 7, 21, 43, 57, 62};
  if (which == 2) return j < 29 ? 0 : (j >= 35 ? 64 : vert[j - 29]);
  if (j < 28) return 0;
  if (j >= 36) return 64;
  return which == 0 ? even[j - 28] : odd[j - 28];
}
static int wedge_ob63(int i, int j) {
  const int shift = (i & 1) ? 15 - (i >> 1) : 16 - (i >> 1);
  return wedge_1d(i & 1, iclip(j - shift, 0, 63));
}
static int wedge_master(int dir, int i, int j) {
  switch (dir) {
    case 0: return wedge_1d(2, i);
    case 1: return wedge_1d(2, j);
    case 2: return wedge_ob63(j, i);
    case 4: return 64 - wedge_ob63(i, 63 - j);
    case 5: return 64 - wedge_ob63(j, 63 - i);
    default: return wedge_ob63(i, j);
  }
}
static int wedge_px(const o_inter* j, int i, int k) {
  const int m = wedge_master(j->wedge_dir, iclip(j->wedge_yoff + i, 0, 63),
                             iclip(j->wedge_xoff + k, 0, 63));
  return j->wedge_flip ? 64 - m : m;
}

/* wedge codebook entry - tables.cpp:616-635 */
int o_wedge_entry(int bw, int bh, int index, int* dir, int* xoff, int* yoff, int* flip) {
  if (index < 0 || index >= 16 || bw < 8 || bh < 8 || bw > 32 || bh > 32) return 0;
  const int shape = bh > bw ? 0 : (bh < bw ? 1 : 2);
  const int* e = SPEC(AV1O_OFF_WEDGE_CODEBOOK) + (shape * 16 + index) * 3;
  *dir = e[0];
  *xoff = 32 - ((e[1] * bw) >> 3);
  *yoff = 32 - ((e[2] * bh) >> 3);
  int sum = 0;
  for (int i = 0; i < bw; ++i) sum += wedge_master(e[0], *yoff, *xoff + i);
  for (int i = 1; i < bh; ++i) sum += wedge_master(e[0], *yoff + i, *xoff);
  const int n = bw + bh - 1;
  *flip = ((sum + n / 2) / n) < 32 ? 1 : 0;
  return 1;
}

/* av1gpu_kernels.h:2025-2029 */
static int diffwtd(int d, int sh, int sign) {
  d = (d + (1 << (sh - 1))) >> sh;
  d = iclip(38 + (d >> 4), 0, 64);
  return sign ? 64 - d : d;
}

/* av1gpu_kernels.h:2076-2100 */
static int obmc_mask(int len, int i) {
  static const int m2[2] = {45, 64}, m4[4] = {39, 50, 59, 64},
                   m8[8] = {36, 42, 48, 53, 57, 61, 64, 64},
                   m16[16] = {34, 37, 40, 43, 46, 49, 52, 54, 56, 58, 60, 61, 64, 64, 64, 64},
                   m32[32] = {33, 35, 36, 38, 40, 41, 43, 44, 45, 47, 48, 50, 51, 52, 53, 55,
                              56, 57, 58, 59, 60, 60, 61, 62, 64, 64, 64, 64, 64, 64, 64, 64};
  if (len == 2) return m2[i];
  if (len == 4) return m4[i];
  if (len == 8) return m8[i];
  if (len == 16) return m16[i];
  return m32[i];
}

/* inter/intra edges - av1gpu_kernels.h:2120-2171 */
static int ii_above(const o_inter* j, const px_t* cur, int i) {
  const px_t* p = cur + j->dst;
  const int mid = 1 << (j->bitdepth - 1);
  if (!j->have_above) return j->have_left ? (int)p[-1] : mid - 1;
  if (j->x + i > j->max_x) i = j->max_x - j->x;
  return p[-j->stride + i];
}
static int ii_left(const o_inter* j, const px_t* cur, int i) {
  const px_t* p = cur + j->dst;
  const int mid = 1 << (j->bitdepth - 1);
  if (!j->have_left) return j->have_above ? (int)p[-j->stride] : mid + 1;
  if (j->y + i > j->max_y) i = j->max_y - j->y;
  return p[(ptrdiff_t)i * j->stride - 1];
}
static int ii_sample(const o_inter* j, const px_t* cur, int x, int y) {
  if (j->ii_mode == 1) return ii_above(j, cur, x);
  if (j->ii_mode == 2) return ii_left(j, cur, y);
  if (j->ii_mode == 3) {
    const int wy = sm_weight(j->h, y), wx = sm_weight(j->w, x);
    return rnd2(wy * ii_above(j, cur, x) + (256 - wy) * ii_left(j, cur, j->h - 1) +
                    wx * ii_left(j, cur, y) + (256 - wx) * ii_above(j, cur, j->w - 1),
                9);
  }
  const int w = j->w, h = j->h;
  int sum = 0;
  if (j->have_above && j->have_left) {
    for (int k = 0; k < w; ++k) sum += ii_above(j, cur, k);
    for (int k = 0; k < h; ++k) sum += ii_left(j, cur, k);
    return (sum + ((w + h) >> 1)) / (w + h);
  }
  if (j->have_left) {
    for (int k = 0; k < h; ++k) sum += ii_left(j, cur, k);
    return (sum + (h >> 1)) >> ilog2_ceil(h);
  }
  if (j->have_above) {
    for (int k = 0; k < w; ++k) sum += ii_above(j, cur, k);
    return (sum + (w >> 1)) >> ilog2_ceil(w);
  }
  return 1 << (j->bitdepth - 1);
}
/* av1gpu_kernels.h:2178-2205 */
static int ii_mask(const o_inter* j, int x, int y) {
  if (j->ii_wedge) {
    const int ny = j->sub_y + 1, nx = j->sub_x + 1, n = nx * ny;
    int sum = 0;
    for (int a = 0; a < ny; ++a)
      for (int b = 0; b < nx; ++b) sum += wedge_px(j, (y << j->sub_y) + a, (x << j->sub_x) + b);
    return n == 1 ? sum : (n == 2 ? (sum + 1) >> 1 : (sum + 2) >> 2);
  }
  if (j->ii_mode == 0) return 32;
  const int scale = 128 / imax(j->w, j->h);
  int k = 0;
  if (j->ii_mode == 1)
    k = y * scale;
  else if (j->ii_mode == 2)
    k = x * scale;
  else if (j->ii_mode == 3)
    k = imin(x, y) * scale;
  return SPEC(AV1O_OFF_II_WEIGHTS)[imin(k, 127)];
}

static const px_t* ref_base(int slot, const px_t* pool, const px_t* cur) {
  return slot >= 0 ? pool : cur;
}

/* one prediction block - av1gpu_kernels.h:2207-2326 (sample by sample) */
void o_inter_block(const o_inter* j, px_t* frame, const px_t* pool, int maxv) {
  const int compound = j->compound != 0;
  const int scx = j->scale_x > 0 ? j->scale_x : (1 << 14);
  const int scy = j->scale_y > 0 ? j->scale_y : (1 << 14);
  const int bd = j->bitdepth;
  for (int y = 0; y < j->h; ++y) {
    for (int x = 0; x < j->w; ++x) {
      int v[2] = {0, 0};
      for (int r = 0; r < (compound ? 2 : 1); ++r) {
        const px_t* base = ref_base(j->slot[r], pool, frame) + j->ref[r];
        const int cmp = r ? 1 : compound;
        v[r] = j->warp[r]
                   ? warp_tap(j->sub_x, j->sub_y, j->ref_w, j->ref_h, j->warp_params[r],
                              j->warp_abcd[r], base, j->stride, j->x, j->y, x, y, bd, cmp)
                   : mc_tap(base, j->stride, j->ref_w, j->ref_h, j->sub_x, j->sub_y, scx, scy,
                            j->x, j->y, j->w, j->h, j->mv[r][0], j->mv[r][1], j->filt,
                            j->filt_v, x, y, cmp, bd);
      }
      int out = v[0];
      if (compound) {
        const int round0 = bd >= 12 ? 5 : 3;
        const int post = 2 * 7 - (round0 + 7);
        if (j->mask_kind != 0) {
          const int sh = post + (bd - 8);
          const int ny = j->sub_y + 1, nx = j->sub_x + 1;
          int sum = 0, n = 0;
          for (int a = 0; a < ny; ++a)
            for (int b = 0; b < nx; ++b) {
              const int lx = (x << j->sub_x) + b, ly = (y << j->sub_y) + a;
              int m;
              if (j->mask_kind == 1) {
                m = wedge_px(j, ly, lx);
                if (j->mask_sign) m = 64 - m;
              } else if (j->plane == 0) {
                m = diffwtd(iabs(v[0] - v[1]), sh, j->mask_sign);
              } else {
                int lv[2];
                for (int r = 0; r < 2; ++r) {
                  const px_t* base = ref_base(j->slot[r], pool, frame) + j->luma_ref[r];
                  lv[r] = j->mask_warp[r]
                              ? warp_tap(0, 0, j->luma_ref_w, j->luma_ref_h, j->warp_params[r],
                                         j->warp_abcd[r], base, j->luma_stride, j->luma_x,
                                         j->luma_y, lx, ly, bd, 1)
                              : mc_tap(base, j->luma_stride, j->luma_ref_w, j->luma_ref_h, 0, 0,
                                       j->scale_x, j->scale_y, j->luma_x, j->luma_y,
                                       j->w << j->sub_x, j->h << j->sub_y, j->mv[r][0],
                                       j->mv[r][1], j->filt, j->filt_v, lx, ly, 1, bd);
                }
                m = diffwtd(iabs(lv[0] - lv[1]), sh, j->mask_sign);
              }
              sum += m;
              ++n;
            }
          const int m = n == 1 ? sum : (n == 2 ? (sum + 1) >> 1 : (sum + 2) >> 2);
          out = (m * v[0] + (64 - m) * v[1] + (1 << (5 + post))) >> (6 + post);
        } else if (j->compound == 1) {
          out = (v[0] + v[1] + (1 << post)) >> (1 + post);
        } else {
          out = (j->w0 * v[0] + j->w1 * v[1] + (1 << (3 + post))) >> (4 + post);
        }
      }
      out = iclip(out, 0, maxv);
      if (j->interintra) {
        const int m = ii_mask(j, x, y);
        out = (m * ii_sample(j, frame, x, y) + (64 - m) * out + 32) >> 6;
      }
      for (int k = 0; k < j->nobmc; ++k) {
        const o_obmc* c = &j->obmc[k];
        if (x < c->rx || x >= c->rx + c->w || y < c->ry || y >= c->ry + c->h) continue;
        const int cj = x - c->rx, ci = y - c->ry;
        const int m = c->pass == 0 ? obmc_mask(c->h, ci) : obmc_mask(c->w, cj);
        const px_t* base = ref_base(c->slot, pool, frame) + c->ref;
        int pred = mc_tap(base, j->stride, j->ref_w, j->ref_h, j->sub_x, j->sub_y, scx, scy,
                          j->x + c->rx, j->y + c->ry, c->w, c->h, c->mvx, c->mvy, c->filt,
                          c->filt_v, cj, ci, 0, bd);
        pred = iclip(pred, 0, maxv);
        out = (m * out + (64 - m) * pred + 32) >> 6;
      }
      frame[j->dst + (size_t)y * j->stride + x] = (px_t)out;
    }
  }
}

/* ======================================================================== */
/* deblocking - av1gpu_kernels.h:2422-2572                                   */
/* ======================================================================== */

static int f4clamp(int v, int bd) { return iclip(v, -(1 << (bd - 1)), (1 << (bd - 1)) - 1); }

static void deblock_line(px_t* b, int s, const o_deblock* j, int maxv) {
  const int bd = j->bitdepth, fs = j->filter_size, sh = bd - 8, mid = 0x80 << sh;
  const int p1 = b[0], p0 = b[s], q0 = b[2 * s], q1 = b[3 * s];
  const int flen = fs == 4 ? 4 : (j->chroma ? 6 : (fs == 8 ? 8 : 16));
  const int shp = j->sharpness;
  const int lshift = shp > 4 ? 2 : (shp > 0 ? 1 : 0);
  int limit = j->level >> lshift;
  if (limit < 1) limit = 1;
  if (shp > 0 && limit > 9 - shp) limit = 9 - shp;
  const int blimit = 2 * (j->level + 2) + limit, thresh = j->level >> 4;
  const int lim = limit << sh, blim = blimit << sh, thr = thresh << sh, one = 1 << sh;
  const int hev = iabs(p1 - p0) > thr || iabs(q1 - q0) > thr;
  int bad = iabs(p1 - p0) > lim || iabs(q1 - q0) > lim ||
            iabs(p0 - q0) * 2 + iabs(p1 - q1) / 2 > blim;
  if (flen >= 6) bad = bad || iabs(b[-s] - p1) > lim || iabs(b[4 * s] - q1) > lim;
  if (flen >= 8) bad = bad || iabs(b[-2 * s] - b[-s]) > lim || iabs(b[5 * s] - b[4 * s]) > lim;
  if (bad) return;
  int flat = 1, flat2 = 1;
  if (fs >= 8) {
    int f = iabs(p1 - p0) > one || iabs(q1 - q0) > one || iabs(b[-s] - p0) > one ||
            iabs(b[4 * s] - q0) > one;
    if (flen >= 8) f = f || iabs(b[-2 * s] - p0) > one || iabs(b[5 * s] - q0) > one;
    flat = !f;
  }
  if (fs >= 16) {
    const int f = iabs(b[-5 * s] - p0) > one || iabs(b[8 * s] - q0) > one ||
                  iabs(b[-4 * s] - p0) > one || iabs(b[7 * s] - q0) > one ||
                  iabs(b[-3 * s] - p0) > one || iabs(b[6 * s] - q0) > one;
    flat2 = !f;
  }
  if (fs == 4 || !flat) {
    const int ps1 = p1 - mid, ps0 = p0 - mid, qs0 = q0 - mid, qs1 = q1 - mid;
    int f = hev ? f4clamp(ps1 - qs1, bd) : 0;
    f = f4clamp(f + 3 * (qs0 - ps0), bd);
    const int f1 = f4clamp(f + 4, bd) >> 3, f2 = f4clamp(f + 3, bd) >> 3;
    b[2 * s] = (px_t)iclip(f4clamp(qs0 - f1, bd) + mid, 0, maxv);
    b[s] = (px_t)iclip(f4clamp(ps0 + f2, bd) + mid, 0, maxv);
    if (!hev) {
      const int f3 = rnd2(f1, 1);
      b[3 * s] = (px_t)iclip(f4clamp(qs1 - f3, bd) + mid, 0, maxv);
      b[0] = (px_t)iclip(f4clamp(ps1 + f3, bd) + mid, 0, maxv);
    }
    return;
  }
  const int lg = (fs == 8 || !flat2) ? 3 : 4;
  const int n = lg == 4 ? 6 : (j->chroma ? 2 : 3);
  const int n2 = (lg == 3 && !j->chroma) ? 0 : 1;
  int res[12];
  for (int i = -n; i < n; ++i) {
    int t = 0;
    for (int k = -n; k <= n; ++k) {
      const int tap = iclip(i + k, -(n + 1), n);
      t += b[(2 + tap) * s] * (iabs(k) <= n2 ? 2 : 1);
    }
    res[i + n] = rnd2(t, lg);
  }
  for (int i = -n; i < n; ++i) b[(2 + i) * s] = (px_t)iclip(res[i + n], 0, maxv);
}

void o_deblock_job(const o_deblock* j, px_t* frame, int maxv) {
  if (j->level == 0) return;
  for (int k = 0; k < j->len; ++k)
    deblock_line(frame + j->p1 + (ptrdiff_t)k * j->along, j->cross, j, maxv);
}

/* ======================================================================== */
/* CDEF - av1gpu_kernels.h:2601-2775                                         */
/* ======================================================================== */

static void cdef_dir(const px_t* l, int stride, int bd, int* dir_out, int* var_out) {
  static const int div[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};
  int part[8][15];
  int cost[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  memset(part, 0, sizeof(part));
  for (int i = 0; i < 8; ++i)
    for (int k = 0; k < 8; ++k) {
      const int x = (l[(size_t)i * stride + k] >> (bd - 8)) - 128;
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

static int cdef_constrain(int diff, int thr, int damping) {
  if (!thr) return 0;
  const int adj = imax(0, damping - ilog2_floor(thr));
  const int mag = iabs(diff);
  const int v = iclip(thr - (mag >> adj), 0, mag);
  return diff < 0 ? -v : v;
}

void o_cdef_block(const o_cdef* j, const px_t* src, px_t* dst, int maxv) {
  static const int dirs[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}},
                                    {{0, 1}, {1, 2}},   {{1, 1}, {2, 2}},  {{1, 0}, {2, 1}},
                                    {{1, 0}, {2, 0}},   {{1, 0}, {2, -1}}};
  static const int uvdir[2][2][8] = {{{0, 1, 2, 3, 4, 5, 6, 7}, {1, 2, 2, 2, 3, 4, 6, 0}},
                                     {{7, 0, 2, 4, 5, 6, 6, 6}, {0, 1, 2, 3, 4, 5, 6, 7}}};
  static const int pri_taps[2][2] = {{4, 2}, {3, 3}};
  static const int sec_taps[2][2] = {{2, 1}, {2, 1}};
  const int cs = j->bitdepth - 8;
  const int w = j->uv ? (8 >> j->ss_x) : 8, h = j->uv ? (8 >> j->ss_y) : 8;
  int pri = j->pri << cs, sec = j->sec << cs;
  const int damping = j->damping + cs - (j->uv ? 1 : 0);
  int dir = j->dir;
  if (dir < 0) {
    int yd = 0, var = 0;
    cdef_dir(src + j->luma, j->luma_stride, j->bitdepth, &yd, &var);
    if (j->uv) {
      dir = pri == 0 ? 0 : uvdir[j->ss_x][j->ss_y][yd];
    } else {
      const int vs = (var >> 6) ? imin(ilog2_floor(var >> 6), 12) : 0;
      dir = pri == 0 ? 0 : yd;
      pri = var ? ((pri * (4 + vs) + 8) >> 4) : 0;
    }
  }
  if (pri == 0 && sec == 0) {
    for (int i = 0; i < h && j->y + i < j->plane_h; ++i)
      for (int k = 0; k < w && j->x + k < j->plane_w; ++k)
        dst[j->dst + (size_t)i * j->stride + k] = src[j->src + (size_t)i * j->stride + k];
    return;
  }
  const int ts = (pri >> cs) & 1;
  for (int i = 0; i < h; ++i)
    for (int k = 0; k < w; ++k) {
      const size_t at = j->src + (size_t)i * j->stride + k;
      const int x = src[at];
      int sum = 0, lo = x, hi = x;
      for (int t = 0; t < 2; ++t)
        for (int sg = -1; sg <= 1; sg += 2) {
          int dy = sg * dirs[dir][t][0], dx = sg * dirs[dir][t][1];
          int px = j->x + k + dx, py = j->y + i + dy;
          if (px >= 0 && px < j->plane_w && py >= 0 && py < j->plane_h) {
            const int p = src[at + (ptrdiff_t)dy * j->stride + dx];
            sum += pri_taps[ts][t] * cdef_constrain(p - x, pri, damping);
            lo = imin(lo, p);
            hi = imax(hi, p);
          }
          for (int o = -2; o <= 2; o += 4) {
            const int sd = (dir + o) & 7;
            dy = sg * dirs[sd][t][0];
            dx = sg * dirs[sd][t][1];
            px = j->x + k + dx;
            py = j->y + i + dy;
            if (px >= 0 && px < j->plane_w && py >= 0 && py < j->plane_h) {
              const int sv = src[at + (ptrdiff_t)dy * j->stride + dx];
              sum += sec_taps[ts][t] * cdef_constrain(sv - x, sec, damping);
              lo = imin(lo, sv);
              hi = imax(hi, sv);
            }
          }
        }
      if (j->x + k >= j->plane_w || j->y + i >= j->plane_h) continue;
      dst[j->dst + (size_t)i * j->stride + k] =
          (px_t)iclip(iclip(x + ((8 + sum - (sum < 0)) >> 4), lo, hi), 0, maxv);
    }
}

/* ======================================================================== */
/* superres - av1gpu_kernels.h:2375-2402                                     */
/* ======================================================================== */

void av1o_superres_params(int dw, int uw, int* step, int* initial) {
  const int s = ((dw << 14) + (uw / 2)) / uw;
  const int err = uw * s - (dw << 14);
  const int sub = (-((uw - dw) << 13) + uw / 2) / uw + (1 << 7) - err / 2;
  *step = s;
  *initial = sub & ((1 << 14) - 1);
}

void o_upscale_row(const o_upscale* j, const px_t* src, px_t* dst, int y, int maxv) {
  const int* uf = SPEC(AV1O_OFF_UPSCALE);
  const px_t* row = src + j->plane_off + (size_t)y * j->stride;
  px_t* out = dst + j->plane_off + (size_t)y * j->stride;
  for (int x = 0; x < j->out_w; ++x) {
    const int sx = -(1 << 14) + j->initial + x * j->step;
    const int p = sx >> 14;
    const int* k = uf + ((sx & ((1 << 14) - 1)) >> 8) * 8;
    int sum = 0;
    for (int t = 0; t < 8; ++t) sum += row[iclip(p + t - 3, 0, j->max_x)] * k[t];
    out[x] = (px_t)iclip((int)(((int64_t)sum + 64) >> 7), 0, maxv);
  }
}

/* ======================================================================== */
/* loop restoration - av1gpu_kernels.h:2832-2982                             */
/* ======================================================================== */

static int lr_src(const o_lr* j, const px_t* cdef, const px_t* curr, int x, int y) {
  x = iclip(x, 0, j->plane_w - 1);
  y = iclip(y, 0, j->plane_h - 1);
  if (y < j->stripe0) {
    y = imax(y, j->stripe0 - 2);
    return curr[j->origin + (size_t)y * j->stride + x];
  }
  if (y > j->stripe1) {
    y = imin(y, j->stripe1 + 2);
    return curr[j->origin + (size_t)y * j->stride + x];
  }
  return cdef[j->origin + (size_t)y * j->stride + x];
}

static int rshift_rnd64(int64_t v, int n) { return (int)((v + ((int64_t)1 << (n - 1))) >> n); }

static int wiener_px(const o_lr* j, const px_t* cdef, const px_t* curr, int x, int y) {
  const int bd = j->bitdepth;
  const int r0 = bd == 12 ? 5 : 3, r1 = bd == 12 ? 9 : 11;
  const int off = 1 << (bd + 7 - r0 - 1);
  const int lim = (1 << (bd + 1 + 7 - r0)) - 1;
  int vt[7], ht[7];
  for (int t = 0; t < 3; ++t) {
    vt[t] = vt[6 - t] = j->vf[t];
    ht[t] = ht[6 - t] = j->hf[t];
  }
  vt[3] = 128 - 2 * (j->vf[0] + j->vf[1] + j->vf[2]);
  ht[3] = 128 - 2 * (j->hf[0] + j->hf[1] + j->hf[2]);
  int sum = 0;
  for (int t = 0; t < 7; ++t) {
    int s = 0;
    for (int k = 0; k < 7; ++k) s += ht[k] * lr_src(j, cdef, curr, x + k - 3, y + t - 3);
    sum += vt[t] * iclip(rshift_rnd64(s, r0), -off, lim - off);
  }
  return rshift_rnd64(sum, r1);
}

static void sgr_ab(const o_lr* j, const px_t* cdef, const px_t* curr, int x, int y, int d,
                   int eps, int* a_out, int* b_out) {
  const int bd = j->bitdepth, r = d >> 1, n = d * d;
  const int64_t n2e = (int64_t)n * n * eps;
  const int64_t mt = (((int64_t)1 << 20) + n2e / 2) / n2e;
  int64_t sa = 0, sb = 0;
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const int c = lr_src(j, cdef, curr, x + dx, y + dy);
      sa += (int64_t)c * c;
      sb += c;
    }
  const int64_t braw = sb;
  if (bd > 8) {
    sa = rshift_rnd64(sa, 2 * (bd - 8));
    sb = rshift_rnd64(sb, bd - 8);
  }
  int64_t p = sa * n - sb * sb;
  if (p < 0) p = 0;
  const int z = rshift_rnd64(p * mt, 20);
  const int a2 = z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1));
  const int inv = ((1 << 12) + n / 2) / n;
  const int64_t b2 = (int64_t)(256 - a2) * braw * inv;
  *a_out = a2;
  *b_out = rshift_rnd64(b2, 12);
}

static int sgr_px(const o_lr* j, const px_t* cdef, const px_t* curr, int x, int y) {
  const int u = lr_src(j, cdef, curr, x, y);
  const int w0 = j->xqd[0], w1 = j->xqd[1], w2 = (1 << 7) - w0 - w1;
  int flt[2];
  for (int pass = 0; pass < 2; ++pass) {
    const int r = j->r[pass];
    if (r == 0) {
      flt[pass] = 0;
      continue;
    }
    int a = 0, b = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int wgt;
        if (pass == 0)
          wgt = ((y + dy) & 1) ? (dx == 0 ? 6 : 5) : 0;
        else
          wgt = (dx == 0 || dy == 0) ? 4 : 3;
        if (!wgt) continue;
        int a2, b2;
        sgr_ab(j, cdef, curr, x + dx, y + dy, 2 * r + 1, j->eps[pass], &a2, &b2);
        a += wgt * a2;
        b += wgt * b2;
      }
    const int sh = (pass == 0 && (y & 1)) ? 4 : 5;
    flt[pass] = rshift_rnd64((int64_t)a * u + b, 8 + sh - 4);
  }
  const int us = u << 4;
  int64_t v = (int64_t)w1 * us;
  v += (int64_t)w0 * (j->r[0] ? flt[0] : us);
  v += (int64_t)w2 * (j->r[1] ? flt[1] : us);
  return rshift_rnd64(v, 4 + 7);
}

void o_lr_job(const o_lr* j, const px_t* cdef, const px_t* curr, px_t* dst, int maxv) {
  for (int y = 0; y < j->h; ++y)
    for (int x = 0; x < j->w; ++x) {
      const int v = j->type == 1 ? wiener_px(j, cdef, curr, j->x + x, j->y + y)
                                 : sgr_px(j, cdef, curr, j->x + x, j->y + y);
      dst[j->dst + (size_t)y * j->stride + x] = (px_t)iclip(v, 0, maxv);
    }
}
