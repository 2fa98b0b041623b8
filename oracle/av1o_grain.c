/*
 * CPU oracle, film grain synthesis (TEST INFRASTRUCTURE ONLY - see av1o.h).
 *
 * Restates the host half of the reference (src/core/grain.cpp:70-197, with
 * kernels/av1gpu_kernels.h:3039-3291) and the per sample blend
 * (av1gpu_kernels.h:3297-3365).
 */
#include <string.h>

#include "av1o_internal.h"

#define LUMA_W 82
#define LUMA_H 73

static int gcenter(int bd) { return 128 << (bd - 8); }
static int gmin(int bd) { return -gcenter(bd); }
static int gmax(int bd) { return (256 << (bd - 8)) - 1 - gcenter(bd); }
static int grnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }

/* 16 bit LFSR - av1gpu_kernels.h:3057-3063 */
static int grand(unsigned* st, int bits) {
  unsigned r = *st;
  const unsigned bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1u;
  r = (r >> 1) | (bit << 15);
  *st = r;
  return (int)((r >> (16 - bits)) & ((1u << bits) - 1u));
}

/* white noise - av1gpu_kernels.h:3070-3079 */
static void fill_noise(unsigned st, int16_t* out, int w, int h, int active, int shift) {
  const int* gs = SPEC(AV1O_OFF_GAUSSIAN);
  for (int i = 0; i < w * h; ++i) {
    const int g = active ? gs[grand(&st, 11)] : 0;
    out[i] = (int16_t)grnd2(g, shift);
  }
}

/* auto-regressive filter - av1gpu_kernels.h:3087-3122 */
static void ar_filter(const av1h_grain* p, int bd, int16_t* pl, int w, int h, const int* coeffs,
                      const int16_t* luma, int ss_x, int ss_y) {
  const int lag = p->ar_coeff_lag, shift = p->ar_coeff_shift_minus_6 + 6;
  const int lo = gmin(bd), hi = gmax(bd);
  for (int y = 3; y < h; ++y)
    for (int x = 3; x < w - 3; ++x) {
      int sum = 0, pos = 0;
      for (int dr = -lag; dr <= 0; ++dr)
        for (int dc = -lag; dc <= lag; ++dc) {
          if (dr == 0 && dc == 0) {
            if (luma && p->num_y_points > 0) {
              const int lx = ((x - 3) << ss_x) + 3, ly = ((y - 3) << ss_y) + 3;
              int acc = 0;
              for (int i = 0; i <= ss_y; ++i)
                for (int k = 0; k <= ss_x; ++k) acc += luma[(ly + i) * LUMA_W + lx + k];
              sum += grnd2(acc, ss_x + ss_y) * coeffs[pos];
            }
            goto done;
          }
          sum += pl[(y + dr) * w + x + dc] * coeffs[pos];
          ++pos;
        }
    done:
      pl[y * w + x] = (int16_t)iclip(pl[y * w + x] + grnd2(sum, shift), lo, hi);
    }
}

/* scaling_lookup_initialization - av1gpu_kernels.h:3157-3188 */
void av1o_grain_scaling_lut(const av1h_grain* p, int plane, int* lut) {
  const int *val, *scl;
  int n;
  if (plane == 0 || p->chroma_scaling_from_luma) {
    n = p->num_y_points;
    val = p->point_y_value;
    scl = p->point_y_scaling;
  } else if (plane == 1) {
    n = p->num_cb_points;
    val = p->point_cb_value;
    scl = p->point_cb_scaling;
  } else {
    n = p->num_cr_points;
    val = p->point_cr_value;
    scl = p->point_cr_scaling;
  }
  if (n <= 0) {
    memset(lut, 0, 256 * sizeof(int));
    return;
  }
  for (int x = 0; x < val[0]; ++x) lut[x] = scl[0];
  for (int i = 0; i + 1 < n; ++i) {
    const int dy = scl[i + 1] - scl[i], dx = val[i + 1] - val[i];
    const int delta = dy * ((65536 + (dx >> 1)) / dx);
    for (int x = 0; x < dx; ++x) lut[val[i] + x] = scl[i] + ((x * delta + 32768) >> 16);
  }
  for (int x = val[n - 1]; x < 256; ++x) lut[x] = scl[n - 1];
}

/* noise image of one plane - av1gpu_kernels.h:3218-3291 */
static void noise_plane(const av1h_grain* p, int bd, const int16_t* grain, int plane, int width,
                        int height, int ss_x, int ss_y, int16_t* out) {
  const int sx = plane ? ss_x : 0, sy = plane ? ss_y : 0;
  const int pw = (width + sx) >> sx, ph = (height + sy) >> sy;
  const int rows = 34 >> sy, step = 32 >> sy;
  const size_t sw = (size_t)pw + 34;
  const int nstripes = (ph + step - 1) / step;
  const int src_w = plane == 0 ? LUMA_W : (ss_x ? 44 : LUMA_W);
  const int lo = gmin(bd), hi = gmax(bd);
  int16_t* scratch = (int16_t*)calloc((size_t)nstripes * rows * sw + 1, sizeof(int16_t));
  for (int s = 0; s < nstripes; ++s) {
    int16_t* st = scratch + (size_t)s * rows * sw;
    unsigned state = (unsigned)p->seed;
    state ^= (unsigned)(((s * 37 + 178) & 255) << 8);
    state ^= (unsigned)((s * 173 + 105) & 255);
    for (int x = 0; x < (width + 1) / 2; x += 16) {
      const int r = grand(&state, 8);
      const int ox = r >> 4, oy = r & 15;
      const int gx = sx ? 6 + ox : 9 + ox * 2;
      const int gy = sy ? 6 + oy : 9 + oy * 2;
      for (int i = 0; i < rows; ++i)
        for (int k = 0; k < (34 >> sx); ++k) {
          int g = grain[(gy + i) * src_w + gx + k];
          const int col = sx ? x + k : x * 2 + k;
          if (!sx) {
            if (k < 2 && p->overlap_flag && x > 0) {
              const int old = st[i * sw + col];
              g = iclip(grnd2(k == 0 ? old * 27 + g * 17 : old * 17 + g * 27, 5), lo, hi);
            }
          } else if (k == 0 && p->overlap_flag && x > 0) {
            const int old = st[i * sw + col];
            g = iclip(grnd2(old * 23 + g * 22, 5), lo, hi);
          }
          st[i * sw + col] = (int16_t)g;
        }
    }
  }
  for (int y = 0; y < ph; ++y) {
    const int si = y >> (5 - sy);
    const int row = y - (si << (5 - sy));
    for (int x = 0; x < pw; ++x) {
      int g = scratch[((size_t)si * rows + row) * sw + x];
      if (!sy) {
        if (row < 2 && si > 0 && p->overlap_flag) {
          const int old = scratch[((size_t)(si - 1) * rows + row + 32) * sw + x];
          g = iclip(grnd2(row == 0 ? old * 27 + g * 17 : old * 17 + g * 27, 5), lo, hi);
        }
      } else if (row < 1 && si > 0 && p->overlap_flag) {
        const int old = scratch[((size_t)(si - 1) * rows + row + 16) * sw + x];
        g = iclip(grnd2(old * 23 + g * 22, 5), lo, hi);
      }
      out[(size_t)y * pw + x] = (int16_t)g;
    }
  }
  free(scratch);
}

/* build_grain - grain.cpp:70-139 (with av1g_grain_generate,
 * av1gpu_kernels.h:3127-3153) */
void o_grain_build(const av1h_grain* p, int bd, int ss_x, int ss_y, int identity, int luma_w,
                   int luma_h, o_grain* out) {
  memset(out, 0, sizeof(*out));
  const int chroma = p->chroma_scaling_from_luma || p->num_cb_points > 0 || p->num_cr_points > 0;
  const int cw = ss_x ? 44 : LUMA_W, ch = ss_y ? 38 : LUMA_H;
  int16_t luma_g[LUMA_W * LUMA_H];
  int16_t* cb = (int16_t*)calloc((size_t)cw * ch, sizeof(int16_t));
  int16_t* cr = (int16_t*)calloc((size_t)cw * ch, sizeof(int16_t));
  const int shift = 12 - bd + p->grain_scale_shift;
  fill_noise((unsigned)p->seed, luma_g, LUMA_W, LUMA_H, p->num_y_points > 0, shift);
  if (p->num_y_points > 0)
    ar_filter(p, bd, luma_g, LUMA_W, LUMA_H, p->ar_coeffs_y, NULL, ss_x, ss_y);
  if (chroma) {
    fill_noise((unsigned)p->seed ^ 0xb524u, cb, cw, ch,
               p->num_cb_points > 0 || p->chroma_scaling_from_luma, shift);
    fill_noise((unsigned)p->seed ^ 0x49d8u, cr, cw, ch,
               p->num_cr_points > 0 || p->chroma_scaling_from_luma, shift);
    ar_filter(p, bd, cb, cw, ch, p->ar_coeffs_cb, luma_g, ss_x, ss_y);
    ar_filter(p, bd, cr, cw, ch, p->ar_coeffs_cr, luma_g, ss_x, ss_y);
  }
  const int maxv = (1 << bd) - 1;
  out->min_value = 0;
  out->max_luma = maxv;
  out->max_chroma = maxv;
  if (p->clip_to_restricted_range) {
    out->min_value = 16 << (bd - 8);
    out->max_luma = 235 << (bd - 8);
    out->max_chroma = identity ? out->max_luma : (240 << (bd - 8));
  }
  out->planes = chroma ? 3 : 1;
  size_t total = 0;
  for (int pl = 0; pl < out->planes; ++pl) {
    const int sx = pl ? ss_x : 0, sy = pl ? ss_y : 0;
    out->off[pl] = total;
    out->stride[pl] = (luma_w + sx) >> sx;
    total += (size_t)out->stride[pl] * ((luma_h + sy) >> sy);
  }
  out->noise = (int16_t*)calloc(total + 1, sizeof(int16_t));
  for (int pl = 0; pl < out->planes; ++pl) {
    const int16_t* src = pl == 0 ? luma_g : (pl == 1 ? cb : cr);
    noise_plane(p, bd, src, pl, luma_w, luma_h, ss_x, ss_y, out->noise + out->off[pl]);
    av1o_grain_scaling_lut(p, pl, out->lut + 256 * pl);
  }
  free(cb);
  free(cr);
}

void o_grain_free(o_grain* g) {
  free(g->noise);
  g->noise = NULL;
}

/* scale_lut - av1gpu_kernels.h:3297-3305 */
static int scale_lut(const int* lut, int index, int bd) {
  const int sh = bd - 8;
  const int x = sh > 0 ? index >> sh : index;
  if (bd == 8 || x >= 255) return lut[x > 255 ? 255 : x];
  const int rem = index - (x << sh);
  return lut[x] + grnd2((lut[x + 1] - lut[x]) * rem, sh);
}

/* the blend, chroma planes first, then luma - cpu_backend.cpp:467-486 with
 * av1g_grain_px (av1gpu_kernels.h:3310-3365) and the plane gating of
 * append_grain_jobs (grain.cpp:141-155) */
void o_grain_apply(const o_grain* g, const av1h_grain* p, int bd, int ss_x, int ss_y,
                   const o_layout* lay, px_t* frame) {
  const int maxv = (1 << bd) - 1;
  const int scaling_shift = p->grain_scaling_minus_8 + 8;
  static const int order[3] = {1, 2, 0};
  for (int oi = 0; oi < 3; ++oi) {
    const int pl = order[oi];
    if (pl == 0) {
      if (p->num_y_points <= 0) continue;
    } else if (!p->chroma_scaling_from_luma &&
               (pl == 1 ? p->num_cb_points : p->num_cr_points) <= 0) {
      continue;
    }
    if (pl >= g->planes) continue;
    const int pw = lay->out_w[pl], ph = lay->out_h[pl];
    const int sx = pl ? ss_x : 0, sy = pl ? ss_y : 0;
    const int* lut = g->lut + 256 * pl;
    const int mult = pl == 1 ? p->cb_mult : p->cr_mult;
    const int lmult = pl == 1 ? p->cb_luma_mult : p->cr_luma_mult;
    const int offset = pl == 1 ? p->cb_offset : p->cr_offset;
    const int vmax = pl == 0 ? g->max_luma : g->max_chroma;
    const int luma_w = lay->out_w[0];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ph; ++y) {
      px_t* row = frame + lay->plane_off[pl] + (size_t)y * lay->stride[pl];
      const px_t* lrow = frame + lay->plane_off[0] + (size_t)(y << sy) * lay->stride[0];
      const int16_t* nrow = g->noise + g->off[pl] + (size_t)y * g->stride[pl];
      for (int x = 0; x < pw; ++x) {
        const int orig = row[x];
        int index;
        if (pl == 0) {
          index = orig;
        } else {
          const int lx = x << sx;
          const int nx = lx + 1 < luma_w ? lx + 1 : luma_w - 1;
          int avg = lrow[lx];
          if (sx) avg = grnd2(avg + lrow[nx], 1);
          if (p->chroma_scaling_from_luma) {
            index = avg;
          } else {
            const int comb = avg * (lmult - 128) + orig * (mult - 128);
            index = iclip((comb >> 6) + (offset - 256) * (1 << (bd - 8)), 0, maxv);
          }
        }
        const int n = nrow[x];
        const int gv = grnd2(scale_lut(lut, index, bd) * n, scaling_shift);
        row[x] = (px_t)iclip(orig + gv, g->min_value, vmax);
      }
    }
    (void)sx;
  }
}
