/*
 * CPU oracle, frame driver (TEST INFRASTRUCTURE ONLY - see av1o.h).
 *
 * Restates the reference's planner (src/core/engine.cpp:519-1494: layout,
 * sample owner grid, wavefront schedule, job field derivation, filter job
 * lists) and its CPU backend (src/backends/cpu/cpu_backend.cpp:149-607: stage
 * order, snapshots, reference store before grain, slot and gather rules).
 */
#include <stdio.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "av1o_internal.h"

#define SLOTS 32

struct av1o_decoder {
  o_layout lay;
  size_t cap; /* samples per buffer */
  px_t *frame, *scratch, *lr_curr, *lr_cdef, *lr_prev_up, *pool;
  int used[SLOTS];
  o_layout slot_lay[SLOTS];
  int last_slot;
  int have_frame;
  int last_waves;
  char err[256];
};

static int g_threads = 0;
void av1o_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}
int av1o_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

const int* av1o_spec_table_ptr(int* size) {
  if (size) *size = AV1O_SPEC_TABLE_SIZE;
  return av1o_spec_table;
}

uint64_t av1o_fnv1a64(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

/* Transform_Row_Shift - tables.cpp:357-369 */
int av1o_tx_row_shift(int w, int h) {
  static const int t[19][3] = {{4, 4, 0},   {8, 8, 1},   {16, 16, 2}, {32, 32, 2}, {64, 64, 2},
                               {4, 8, 0},   {8, 4, 0},   {8, 16, 1},  {16, 8, 1},  {16, 32, 1},
                               {32, 16, 1}, {32, 64, 1}, {64, 32, 1}, {4, 16, 1},  {16, 4, 1},
                               {8, 32, 2},  {32, 8, 2},  {16, 64, 2}, {64, 16, 2}};
  for (int i = 0; i < 19; ++i)
    if (t[i][0] == w && t[i][1] == h) return t[i][2];
  return 0;
}

/* setup shear and resolve divisor - tables.cpp:426-475 */
static int64_t rnd2s64(int64_t v, int n) {
  const int64_t r = (int64_t)1 << (n - 1);
  return v >= 0 ? (v + r) >> n : -((-v + r) >> n);
}
int av1o_setup_shear(const int32_t* p, int* alpha, int* beta, int* gamma, int* delta) {
  const int* lut = SPEC(AV1O_OFF_DIV_LUT);
  const int a0 = iclip(p[2] - (1 << 16), -32768, 32767);
  const int b0 = iclip(p[3], -32768, 32767);
  const int ad = p[2] < 0 ? -p[2] : p[2];
  const int n = ilog2_floor(ad);
  const int e = ad - (1 << n);
  const int f = n > 8 ? (e + (1 << (n - 8 - 1))) >> (n - 8) : e << (8 - n);
  const int dshift = n + 14;
  const int dfac = p[2] < 0 ? -lut[f] : lut[f];
  const int64_t v = (int64_t)p[4] * ((int64_t)1 << 16);
  const int g0 = iclip((int)rnd2s64(v * dfac, dshift), -32768, 32767);
  const int64_t w = (int64_t)p[3] * p[4];
  const int d0 = iclip((int)(p[5] - rnd2s64(w * dfac, dshift) - (1 << 16)), -32768, 32767);
  *alpha = rnd2s(a0, 6) * 64;
  *beta = rnd2s(b0, 6) * 64;
  *gamma = rnd2s(g0, 6) * 64;
  *delta = rnd2s(d0, 6) * 64;
  if (4 * iabs(*alpha) + 7 * iabs(*beta) >= (1 << 16)) return 0;
  if (4 * iabs(*gamma) + 4 * iabs(*delta) >= (1 << 16)) return 0;
  return 1;
}

/* Q16 YUV -> RGB coefficients - tables.cpp:481-533 */
void av1o_color_matrix(int matrix, int range, int bd, int m[8], int o[3]) {
  const int sh = bd - 8;
  const int yo = range == 0 ? (16 << sh) : 0;
  const int co = range == 0 ? (128 << sh) : (1 << (bd - 1));
  const int half = 1 << 15;
  memset(m, 0, 8 * sizeof(int));
  if (matrix == 3) {
    m[0] = m[2] = m[5] = 65536;
    o[0] = o[1] = o[2] = 0;
    return;
  }
  if (matrix == 2) {
    m[0] = 76293; m[1] = 109969; m[2] = 76293; m[3] = -12255; m[4] = -42598;
    m[5] = 76293; m[6] = 121634;
  } else {
    m[0] = 76293; m[1] = 117504; m[2] = 76293; m[3] = -13926; m[4] = -34937;
    m[5] = 76293; m[6] = 138412;
  }
  o[0] = half - m[0] * yo - m[1] * co;
  o[1] = half - m[2] * yo - m[3] * co - m[4] * co;
  o[2] = half - m[5] * yo - m[6] * co;
}

/* make_plan_layout - factory.cpp:114-152 */
void o_make_layout(o_layout* g, int width, int height, int bd, int ss_x, int ss_y,
                   int coded_width, int align) {
  memset(g, 0, sizeof(*g));
  if (coded_width <= 0 || coded_width > width) coded_width = width;
  g->width = width;
  g->height = height;
  g->bit_depth = bd;
  g->ss_x = ss_x;
  g->ss_y = ss_y;
  g->out_w[0] = width;
  g->out_h[0] = height;
  g->out_w[1] = g->out_w[2] = (width + (1 << ss_x) - 1) >> ss_x;
  g->out_h[1] = g->out_h[2] = (height + (1 << ss_y) - 1) >> ss_y;
  g->plane_w[0] = coded_width;
  g->plane_h[0] = height;
  g->plane_w[1] = g->plane_w[2] = (coded_width + (1 << ss_x) - 1) >> ss_x;
  g->plane_h[1] = g->plane_h[2] = g->out_h[1];
  size_t size = 0;
  for (int p = 0; p < 3; ++p) {
    const int pad = p == 0 ? 64 : (64 >> (ss_x ? 1 : 0));
    g->stride[p] = (g->out_w[p] + pad + align - 1) & ~(align - 1);
    g->plane_off[p] = size;
    size += (size_t)g->stride[p] * (size_t)(g->out_h[p] + pad);
  }
  g->size = size;
}

av1o_decoder* av1o_create(void) {
  av1o_decoder* d = (av1o_decoder*)calloc(1, sizeof(av1o_decoder));
  if (d) d->last_slot = -1;
  return d;
}

void av1o_destroy(av1o_decoder* d) {
  if (!d) return;
  free(d->frame);
  free(d->scratch);
  free(d->lr_curr);
  free(d->lr_cdef);
  free(d->lr_prev_up);
  free(d->pool);
  free(d);
}

const char* av1o_last_error(const av1o_decoder* d) { return d ? d->err : "no decoder"; }

static int fail(av1o_decoder* d, int code, const char* msg) {
  snprintf(d->err, sizeof(d->err), "%s", msg);
  return code;
}

/* CpuBackend::prepare - cpu_backend.cpp:77-91: buffers only grow; the pool is
 * re-created (zeroed) when it has to grow */
static int prepare(av1o_decoder* d, const o_layout* g) {
  if (d->cap < g->size) {
    free(d->frame);
    free(d->scratch);
    free(d->lr_curr);
    free(d->lr_cdef);
    free(d->lr_prev_up);
    free(d->pool);
    d->frame = (px_t*)calloc(g->size, sizeof(px_t));
    d->scratch = (px_t*)calloc(g->size, sizeof(px_t));
    d->lr_curr = (px_t*)calloc(g->size, sizeof(px_t));
    d->lr_cdef = (px_t*)calloc(g->size, sizeof(px_t));
    d->lr_prev_up = (px_t*)calloc(g->size, sizeof(px_t));
    d->pool = (px_t*)calloc(g->size * SLOTS, sizeof(px_t));
    if (!d->frame || !d->scratch || !d->lr_curr || !d->lr_cdef || !d->lr_prev_up || !d->pool) {
      d->cap = 0;
      return 0;
    }
    d->cap = g->size;
  }
  d->lay = *g;
  return 1;
}

static int alloc_slot(av1o_decoder* d) {
  for (int i = 0; i < SLOTS; ++i)
    if (!d->used[i]) return i;
  return -1;
}

int av1o_reserve_reference(av1o_decoder* d) {
  const int s = alloc_slot(d);
  if (s < 0) return fail(d, AV1H_ERR_SLOTS, "CPU reference pool exhausted");
  d->used[s] = 1;
  return s;
}

void av1o_release_reference(av1o_decoder* d, int slot) {
  if (slot >= 0 && slot < SLOTS) d->used[slot] = 0;
}

void av1o_reset(av1o_decoder* d) {
  memset(d->used, 0, sizeof(d->used));
  d->last_slot = -1;
  d->have_frame = 0;
  if (d->frame) memset(d->frame, 0, d->cap * sizeof(px_t));
  if (d->pool) memset(d->pool, 0, d->cap * SLOTS * sizeof(px_t));
}

int av1o_last_waves(const av1o_decoder* d) { return d->last_waves; }

/* ---- planning ------------------------------------------------------------ */

typedef struct {
  int nintra, ninter, nitx, nres;
  int cintra, cinter, citx, cres;
  o_intra* intra;
  o_inter* inter;
  o_itx* itx;
  o_itx* res;
} o_wave;

static const int kModeAngle[13] = {0, 90, 180, 45, 135, 113, 157, 203, 67, 0, 0, 0, 0};

#define PUSH(arr, n, cap, T, v)                                    \
  do {                                                             \
    if ((n) == (cap)) {                                            \
      (cap) = (cap) ? 2 * (cap) : 16;                              \
      (arr) = (T*)realloc((arr), sizeof(T) * (size_t)(cap));        \
    }                                                              \
    (arr)[(n)++] = (v);                                            \
  } while (0)

typedef struct {
  int grid_w[3], grid_h[3];
  int shift[3], cw[3], ch[3];
  int* owner[3];
} o_owner;

static int owner_at(const o_owner* o, int p, int x, int y) {
  if (x < 0 || y < 0 || x >= o->grid_w[p] || y >= o->grid_h[p] || !o->owner[p]) return -1;
  return o->owner[p][(size_t)(y >> o->shift[p]) * o->cw[p] + (x >> o->shift[p])];
}

/* BlockDecoded over a rectangle - engine.cpp:779-793 */
static int decoded_before(const o_owner* o, int p, int x0, int x1, int y0, int y1, int bi) {
  const int oc = 1 << o->shift[p];
  for (int y = y0 & ~(oc - 1); y <= y1; y += oc)
    for (int x = x0 & ~(oc - 1); x <= x1; x += oc) {
      const int w = owner_at(o, p, x, y);
      if (w >= 0 && w >= bi) return 0;
    }
  return 1;
}

static int resolve_slot(const av1h_frame_header* h, int ref) {
  if (ref < 0 || ref >= h->num_reference_slots) return -1;
  return h->reference_slots[ref];
}

static int validate(av1o_decoder* d, const av1h_frame_desc* f, const o_layout* g) {
  const av1h_frame_header* h = &f->hdr;
  if (h->width <= 0 || h->height <= 0) return fail(d, AV1H_ERR_INVALID, "frame has no size");
  if (h->bit_depth < 8 || h->bit_depth > 12) return fail(d, AV1H_ERR_INVALID, "bad bit depth");
  if (h->num_reference_slots < 0 || h->num_reference_slots > AV1H_MAX_REFS)
    return fail(d, AV1H_ERR_INVALID, "bad reference slot count");
  for (int64_t i = 0; i < f->num_blocks; ++i) {
    const av1h_block* b = &f->blocks[i];
    if (b->plane < 0 || b->plane > 2 || b->x < 0 || b->y < 0 || b->w < 2 || b->h < 2 ||
        b->w > 128 || b->h > 128)
      return fail(d, AV1H_ERR_INVALID, "block out of range");
    const int gw = g->stride[b->plane], gh = g->out_h[b->plane] + (b->plane ? 32 : 64);
    if (b->x + b->w > gw || b->y + b->h > gh) return fail(d, AV1H_ERR_INVALID, "block outside the plane");
    if (b->tx_first < 0 || b->tx_count < 0 || b->tx_first + b->tx_count > f->num_transforms)
      return fail(d, AV1H_ERR_INVALID, "transform range out of bounds");
    if (b->obmc_first < 0 || b->obmc_count < 0 || b->obmc_first + b->obmc_count > f->num_obmc)
      return fail(d, AV1H_ERR_INVALID, "obmc range out of bounds");
    for (int t = 0; t < b->tx_count; ++t) {
      const av1h_transform* tx = &f->transforms[b->tx_first + t];
      if (tx->w < 4 || tx->h < 4 || tx->w > 64 || tx->h > 64 || tx->x < 0 || tx->y < 0 ||
          tx->x + tx->w > b->w || tx->y + tx->h > b->h || tx->coef_offset < 0)
        return fail(d, AV1H_ERR_INVALID, "transform out of range");
    }
    if (!b->is_intra) {
      for (int r = 0; r < 2; ++r) {
        const int s = resolve_slot(h, b->ref[r]);
        if (s >= SLOTS || (s >= 0 && !d->used[s]))
          return fail(d, AV1H_ERR_INVALID, "a block names a reference slot that holds no picture");
      }
    }
    if (b->use_palette && (b->palette_off < 0 ||
                           (int64_t)b->palette_off + (int64_t)(b->h - 1) * b->palette_stride + b->w >
                               f->num_palette))
      return fail(d, AV1H_ERR_INVALID, "palette map out of bounds");
  }
  for (int64_t i = 0; i < f->num_deblock; ++i) {
    const av1h_deblock_edge* e = &f->deblock[i];
    if (e->plane < 0 || e->plane > 2 || e->length < 0)
      return fail(d, AV1H_ERR_INVALID, "deblock edge out of range");
    /* how far the filter reads across the edge (p6..q6 for the widest) */
    const int reach = e->filter_size >= 16 ? 7 : (e->filter_size >= 8 ? 4 : 2);
    const int rows = g->out_h[e->plane] + (e->plane ? 32 : 64);
    const int across = e->vertical ? e->x : e->y, along = e->vertical ? e->y : e->x;
    const int across_lim = e->vertical ? g->stride[e->plane] : rows;
    const int along_lim = e->vertical ? rows : g->stride[e->plane];
    if (across < reach || across + reach > across_lim || along < 0 || along + e->length > along_lim)
      return fail(d, AV1H_ERR_INVALID, "deblock edge out of range");
  }
  for (int64_t i = 0; i < f->num_cdef; ++i) {
    const av1h_cdef_block* c = &f->cdef[i];
    if (c->plane < 0 || c->plane > 2 || c->x < 0 || c->y < 0)
      return fail(d, AV1H_ERR_INVALID, "cdef block out of range");
  }
  return AV1H_OK;
}

int av1o_decode(av1o_decoder* d, const av1h_frame_desc* f, int keep, int reserved,
                const av1h_planes* out, av1h_frame_info* info, int* slot_out) {
  const av1h_frame_header* hdr = &f->hdr;
  if (slot_out) *slot_out = -1;
  d->err[0] = 0;
  o_layout g;
  o_make_layout(&g, hdr->width, hdr->height, hdr->bit_depth, hdr->ss_x, hdr->ss_y,
                hdr->coded_width > 0 ? hdr->coded_width : hdr->width, 32);
  if (hdr->width <= 0 || hdr->height <= 0) return fail(d, AV1H_ERR_INVALID, "frame has no size");
  if (!prepare(d, &g)) return fail(d, AV1H_ERR_MEMORY, "out of memory");
  int rc = validate(d, f, &g);
  if (rc != AV1H_OK) return rc;
  for (int i = 0; i < 7; ++i) {
    const int rw = hdr->reference_width[i] > 0 ? hdr->reference_width[i] : hdr->width;
    const int rh = hdr->reference_height[i] > 0 ? hdr->reference_height[i] : hdr->height;
    if (rw != hdr->width || rh != hdr->height)
      return fail(d, AV1H_ERR_UNSUPPORTED,
                  "reference scaling between pictures of different sizes is not supported");
  }
  const int bd = hdr->bit_depth, ss_x = hdr->ss_x, ss_y = hdr->ss_y;
  const int maxv = (1 << bd) - 1;
  const int fw = g.plane_w[0], fh = g.plane_h[0];
  const int upscaled = g.out_w[0] != g.plane_w[0];
  px_t* frame = d->frame;
  const int nb = (int)f->num_blocks;

  /* pools padded the way build_frame_plan pads them (engine.cpp:595-660) */
  size_t cend = 0, rend = 0;
  for (int i = 0; i < nb; ++i) {
    const av1h_block* b = &f->blocks[i];
    for (int t = 0; t < b->tx_count; ++t) {
      const av1h_transform* tx = &f->transforms[b->tx_first + t];
      const size_t e = (size_t)tx->coef_offset + (size_t)tx->w * tx->h;
      if (b->residual_provided) {
        if (e > rend) rend = e;
      } else if (e > cend) {
        cend = e;
      }
    }
  }
  size_t csz = cend + 64, rsz = rend + 64;
  if ((size_t)f->num_coefficients > csz) csz = (size_t)f->num_coefficients;
  if ((size_t)f->num_residual > rsz) rsz = (size_t)f->num_residual;
  int32_t* coef = (int32_t*)calloc(csz, sizeof(int32_t));
  int32_t* resid = (int32_t*)calloc(rsz, sizeof(int32_t));
  if (f->num_coefficients) memcpy(coef, f->coefficients, (size_t)f->num_coefficients * 4);
  if (f->num_residual) memcpy(resid, f->residual, (size_t)f->num_residual * 4);

  /* ---- owner grid and wavefront (engine.cpp:695-926) ------------------- */
  o_owner own;
  memset(&own, 0, sizeof(own));
  const int cell[3] = {4, ss_x ? 2 : 4, ss_y ? 2 : 4};
  const int csh[3] = {2, ss_x ? 1 : 2, ss_y ? 1 : 2};
  int need[3] = {0, 0, 0};
  for (int p = 0; p < 3; ++p) {
    const int sx = p && ss_x ? 1 : 0, sy = p && ss_y ? 1 : 0;
    own.grid_w[p] = ((fw + 7) & ~7) >> sx;
    own.grid_h[p] = ((fh + 7) & ~7) >> sy;
    own.shift[p] = csh[p];
    own.cw[p] = (own.grid_w[p] + cell[p] - 1) / cell[p];
    own.ch[p] = (own.grid_h[p] + cell[p] - 1) / cell[p];
  }
  for (int i = 0; i < nb; ++i) {
    const av1h_block* b = &f->blocks[i];
    if (b->mode == AV1H_PRED_CFL) need[0] = 1;
    if (b->is_intra || b->intrabc || b->interintra) need[b->plane] = 1;
  }
  for (int p = 0; p < 3; ++p)
    if (need[p]) {
      const size_t n = (size_t)own.cw[p] * own.ch[p];
      own.owner[p] = (int*)malloc(n * sizeof(int));
      for (size_t k = 0; k < n; ++k) own.owner[p][k] = -1;
    }
  for (int i = 0; i < nb; ++i) {
    const av1h_block* b = &f->blocks[i];
    const int p = b->plane;
    if (!own.owner[p]) continue;
    const int oc = cell[p];
    const int cx1 = imin((b->x + b->w + oc - 1) / oc, own.cw[p]);
    const int cy1 = imin((b->y + b->h + oc - 1) / oc, own.ch[p]);
    for (int cy = imax(b->y, 0) / oc; cy < cy1; ++cy)
      for (int cx = imax(b->x, 0) / oc; cx < cx1; ++cx) own.owner[p][(size_t)cy * own.cw[p] + cx] = i;
  }
  int* wave = (int*)calloc((size_t)nb + 1, sizeof(int));
  char* ar = (char*)calloc((size_t)nb + 1, 1);
  char* bl = (char*)calloc((size_t)nb + 1, 1);
  int max_wave = 0;
  for (int bi = 0; bi < nb; ++bi) {
    const av1h_block* b = &f->blocks[bi];
    const int p = b->plane;
    const int oc = cell[p];
    int dep = -1;
    if (b->is_intra || b->interintra) {
      if (b->is_intra) {
        ar[bi] = (char)(b->have_above_right >= 0
                            ? b->have_above_right != 0
                            : (b->y > 0 && decoded_before(&own, p, b->x + b->w, b->x + 2 * b->w - 1,
                                                          b->y - 1, b->y - 1, bi)));
        bl[bi] = (char)(b->have_below_left >= 0
                            ? b->have_below_left != 0
                            : (b->x > 0 && decoded_before(&own, p, b->x - 1, b->x - 1, b->y + b->h,
                                                          b->y + 2 * b->h - 1, bi)));
      }
      int ly1 = b->y + b->h - 1, ax1 = b->x + b->w - 1;
      if (b->is_intra && !b->use_palette) {
        if (b->mode >= 1 && b->mode <= 8) {
          const int pa = kModeAngle[b->mode] + 3 * b->angle_delta;
          if (pa < 90 && ar[bi]) ax1 = b->x + 2 * b->w - 1;
          if (pa > 180 && bl[bi]) ly1 = b->y + 2 * b->h - 1;
        }
      } else if (b->is_intra) {
        ax1 = b->x - 1;
        ly1 = b->y - 1;
      }
      for (int y = b->y & ~(oc - 1); y <= ly1; y += oc) {
        const int o = owner_at(&own, p, b->x - 1, y);
        if (o >= 0) dep = imax(dep, wave[o]);
      }
      for (int x = b->x & ~(oc - 1); x <= ax1; x += oc) {
        const int o = owner_at(&own, p, x, b->y - 1);
        if (o >= 0) dep = imax(dep, wave[o]);
      }
      if (ax1 >= b->x || ly1 >= b->y) {
        const int o = owner_at(&own, p, b->x - 1, b->y - 1);
        if (o >= 0) dep = imax(dep, wave[o]);
      }
    }
    if (b->intrabc) {
      const int msx = 3 + (p > 0 && ss_x ? 1 : 0), msy = 3 + (p > 0 && ss_y ? 1 : 0);
      const int sx = b->x + (b->mv[0][0] >> msx), sy = b->y + (b->mv[0][1] >> msy);
      const int x1 = imin(sx + b->w, own.grid_w[p]), y1 = imin(sy + b->h, own.grid_h[p]);
      for (int y = imax(sy, 0) & ~(oc - 1); y < y1; y += oc)
        for (int x = imax(sx, 0) & ~(oc - 1); x < x1; x += oc) {
          const int o = owner_at(&own, p, x, y);
          if (o >= 0) dep = imax(dep, wave[o]);
        }
    }
    if (b->mode == AV1H_PRED_CFL) {
      for (int y = 0; y < b->h; ++y)
        for (int x = 0; x < b->w; ++x) {
          const int o = owner_at(&own, 0, (b->x + x) << ss_x, (b->y + y) << ss_y);
          if (o >= 0) dep = imax(dep, wave[o]);
        }
    }
    if (dep + 1 > wave[bi]) wave[bi] = dep + 1;
    if (wave[bi] > max_wave) max_wave = wave[bi];
  }
  d->last_waves = max_wave + 1;

  /* ---- reference scaling (engine.cpp:936-947) --------------------------- */
  int scale_x = 1 << 14, scale_y = 1 << 14;
  if (upscaled) {
    scale_x = ((hdr->width << 14) + (fw / 2)) / fw;
    scale_y = ((hdr->height << 14) + (fh / 2)) / fh;
  }

  /* ---- jobs (engine.cpp:983-1224) --------------------------------------- */
  o_wave* waves = (o_wave*)calloc((size_t)max_wave + 1, sizeof(o_wave));
  size_t tmp_need = 0;
  rc = AV1H_OK;
  for (int bi = 0; bi < nb && rc == AV1H_OK; ++bi) {
    const av1h_block* b = &f->blocks[bi];
    o_wave* wv = &waves[wave[bi]];
    const int p = b->plane;
    const int stride = g.stride[p];
    const size_t dst = g.plane_off[p] + (size_t)b->y * stride + b->x;
    const int max_x = (((fw + 7) & ~7) >> (p ? ss_x : 0)) - 1;
    const int max_y = (((fh + 7) & ~7) >> (p ? ss_y : 0)) - 1;
    if (b->is_intra) {
      o_intra j;
      memset(&j, 0, sizeof(j));
      j.dst = dst;
      j.stride = stride;
      j.x = b->x;
      j.y = b->y;
      j.w = b->w;
      j.h = b->h;
      j.plane = p;
      j.mode = b->mode;
      j.angle = b->angle_delta;
      j.bitdepth = b->bit_depth;
      j.have_left = b->have_left && b->x > 0;
      j.have_above = b->have_above && b->y > 0;
      j.have_ar = b->have_above_right >= 0 ? b->have_above_right != 0 : ar[bi] != 0;
      j.have_bl = b->have_below_left >= 0 ? b->have_below_left != 0 : bl[bi] != 0;
      j.filter_type = hdr->enable_intra_edge_filter ? b->filter_type : -1;
      j.filter_intra = b->use_filter_intra && p == 0;
      j.fi_mode = b->filter_intra_mode;
      j.cfl_alpha = b->cfl_alpha;
      j.luma = g.plane_off[0];
      j.luma_stride = g.stride[0];
      j.luma_w = (b->x + b->w) << ss_x;
      j.luma_h = (b->y + b->h) << ss_y;
      j.ss_x = ss_x;
      j.ss_y = ss_y;
      j.max_x = max_x;
      j.max_y = max_y;
      if (b->use_palette) {
        j.palette = 1;
        for (int k = 0; k < 8; ++k) j.palette_color[k] = b->palette[k];
        j.palette_off = b->palette_off;
        j.palette_stride = b->palette_stride;
      }
      PUSH(wv->intra, wv->nintra, wv->cintra, o_intra, j);
    } else {
      o_inter j;
      memset(&j, 0, sizeof(j));
      j.dst = dst;
      j.stride = stride;
      j.plane = p;
      j.w = b->w;
      j.h = b->h;
      j.x = b->x;
      j.y = b->y;
      j.ref_w = g.out_w[p];
      j.ref_h = g.out_h[p];
      j.scale_x = scale_x;
      j.scale_y = scale_y;
      if (b->intrabc) {
        const int sx = p > 0 && ss_x ? 1 : 0, sy = p > 0 && ss_y ? 1 : 0;
        j.ref_w = ((fw + 3) & ~3) >> sx;
        j.ref_h = ((fh + 3) & ~3) >> sy;
        j.scale_x = j.scale_y = 1 << 14;
      }
      for (int r = 0; r < 2; ++r) {
        j.mv[r][0] = b->mv[r][0];
        j.mv[r][1] = b->mv[r][1];
        j.slot[r] = resolve_slot(hdr, b->ref[r]);
        j.ref[r] = g.plane_off[p] + (j.slot[r] >= 0 ? (size_t)j.slot[r] * g.size : 0);
      }
      j.filt = b->filter;
      j.filt_v = b->filter_v;
      j.sub_x = ss_x && p > 0 ? 1 : 0;
      j.sub_y = ss_y && p > 0 ? 1 : 0;
      j.compound = b->compound;
      j.w0 = b->blend[0];
      j.w1 = b->blend[1];
      j.bitdepth = b->bit_depth;
      if (b->compound == AV1H_COMPOUND_WEDGE || b->compound == AV1H_COMPOUND_DIFFWTD) {
        const int lw = b->w << (p > 0 ? ss_x : 0), lh = b->h << (p > 0 ? ss_y : 0);
        j.mask_kind = b->compound == AV1H_COMPOUND_WEDGE ? 1 : 2;
        j.mask_sign = b->mask_sign;
        if (j.mask_kind == 1 &&
            !o_wedge_entry(lw, lh, b->wedge_index, &j.wedge_dir, &j.wedge_xoff, &j.wedge_yoff,
                           &j.wedge_flip)) {
          rc = fail(d, AV1H_ERR_INVALID, "a wedge on a block size that has none");
          break;
        }
        if (j.mask_kind == 2 && p > 0) {
          for (int r = 0; r < 2; ++r)
            j.luma_ref[r] = g.plane_off[0] + (j.slot[r] >= 0 ? (size_t)j.slot[r] * g.size : 0);
          j.luma_stride = g.stride[0];
          j.luma_x = b->x << ss_x;
          j.luma_y = b->y << ss_y;
          j.luma_ref_w = g.out_w[0];
          j.luma_ref_h = g.out_h[0];
        }
      }
      j.interintra = b->interintra ? 1 : 0;
      j.ii_mode = b->interintra_mode;
      j.have_left = b->have_left && b->x > 0;
      j.have_above = b->have_above && b->y > 0;
      j.max_x = max_x;
      j.max_y = max_y;
      if (b->interintra && b->interintra_wedge) {
        const int lw = b->w << (p > 0 ? ss_x : 0), lh = b->h << (p > 0 ? ss_y : 0);
        if (!o_wedge_entry(lw, lh, b->wedge_index, &j.wedge_dir, &j.wedge_xoff, &j.wedge_yoff,
                           &j.wedge_flip)) {
          rc = fail(d, AV1H_ERR_INVALID, "an inter/intra wedge on a block size that has none");
          break;
        }
        j.ii_wedge = 1;
      }
      for (int k = 0; k < b->obmc_count && j.nobmc < 8; ++k) {
        const av1h_obmc* c = &f->obmc[b->obmc_first + k];
        o_obmc* o = &j.obmc[j.nobmc++];
        o->pass = c->pass;
        o->rx = c->x;
        o->ry = c->y;
        o->w = c->w;
        o->h = c->h;
        o->mvx = c->mv_x;
        o->mvy = c->mv_y;
        o->filt = c->filter;
        o->filt_v = c->filter_v;
        o->slot = resolve_slot(hdr, c->ref);
        o->ref = g.plane_off[p] + (o->slot >= 0 ? (size_t)o->slot * g.size : 0);
      }
      for (int r = 0; r < 2; ++r) {
        if (!b->warp[r] && !b->mask_warp[r]) continue;
        j.warp[r] = b->warp[r] ? 1 : 0;
        j.mask_warp[r] = b->mask_warp[r] ? 1 : 0;
        int a, be, ga, de;
        if (!av1o_setup_shear(b->warp_params[r], &a, &be, &ga, &de)) continue;
        for (int k = 0; k < 6; ++k) j.warp_params[r][k] = b->warp_params[r][k];
        j.warp_abcd[r][0] = a;
        j.warp_abcd[r][1] = be;
        j.warp_abcd[r][2] = ga;
        j.warp_abcd[r][3] = de;
      }
      PUSH(wv->inter, wv->ninter, wv->cinter, o_inter, j);
    }
    for (int t = 0; t < b->tx_count; ++t) {
      const av1h_transform* tx = &f->transforms[b->tx_first + t];
      o_itx j;
      memset(&j, 0, sizeof(j));
      j.coef = (size_t)tx->coef_offset;
      j.dst = dst + (size_t)tx->y * stride + tx->x;
      j.stride = stride;
      j.w = tx->w;
      j.h = tx->h;
      j.tx_type = tx->tx_type;
      j.eob = tx->eob;
      j.lossless = b->lossless ? 1 : 0;
      j.row_shift = b->lossless ? 0 : av1o_tx_row_shift(tx->w, tx->h);
      j.bitdepth = b->bit_depth;
      if (j.eob <= 0) continue;
      if (b->residual_provided)
        PUSH(wv->res, wv->nres, wv->cres, o_itx, j);
      else
        PUSH(wv->itx, wv->nitx, wv->citx, o_itx, j);
    }
  }
  (void)tmp_need;
  free(ar);
  free(bl);
  free(wave);
  for (int p = 0; p < 3; ++p) free(own.owner[p]);
  if (rc != AV1H_OK) {
    for (int w = 0; w <= max_wave; ++w) {
      free(waves[w].intra);
      free(waves[w].inter);
      free(waves[w].itx);
      free(waves[w].res);
    }
    free(waves);
    free(coef);
    free(resid);
    return rc;
  }

  /* ---- reconstruction (cpu_backend.cpp:203-336) ------------------------- */
  const uint8_t* palette = f->palette_map;
  for (int w = 0; w <= max_wave; ++w) {
    o_wave* wv = &waves[w];
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < wv->ninter; ++i) o_inter_block(&wv->inter[i], frame, d->pool, maxv);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < wv->nintra; ++i) o_intra_block(&wv->intra[i], frame, palette, maxv);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < wv->nitx; ++i) {
      int32_t tmp[64 * 64];
      o_itx j = wv->itx[i];
      j.tmp = 0;
      for (int y = 0; y < j.h; ++y) o_itx_row(&j, coef, tmp, y);
      for (int x = 0; x < j.w; ++x) o_itx_col(&j, tmp, frame, x, maxv);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < wv->nres; ++i) o_residual_add(&wv->res[i], resid, frame, maxv);
    free(wv->intra);
    free(wv->inter);
    free(wv->itx);
    free(wv->res);
  }
  free(waves);
  free(coef);
  free(resid);

  /* ---- deblocking (engine.cpp:1265-1297, cpu_backend.cpp:345-356) ------- */
  for (int pass = 0; pass < 2; ++pass) {
    const int vertical = pass == 0;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < f->num_deblock; ++i) {
      const av1h_deblock_edge* e = &f->deblock[i];
      if ((e->vertical != 0) != vertical) continue;
      const int p = e->plane;
      o_deblock j;
      j.p1 = g.plane_off[p] + (vertical ? (size_t)e->y * g.stride[p] + (e->x - 2)
                                        : (size_t)(e->y - 2) * g.stride[p] + e->x);
      j.cross = vertical ? 1 : g.stride[p];
      j.along = vertical ? g.stride[p] : 1;
      j.len = e->length;
      j.level = e->level;
      j.sharpness = hdr->sharpness;
      j.filter_size = e->filter_size;
      j.chroma = p != 0;
      j.bitdepth = bd;
      o_deblock_job(&j, frame, maxv);
    }
  }

  int uses_lr = 0;
  for (int p = 0; p < 3; ++p)
    if (hdr->restoration_type[p] != AV1H_RESTORE_NONE) uses_lr = 1;
  /* the planner emits no LR job for a plane without units, and the snapshot
   * is only taken when some job exists; mirror that by building jobs first */
  o_lr* lr = NULL;
  int nlr = 0, clr = 0;
  for (int p = 0; p < 3 && uses_lr; ++p) {
    if (hdr->restoration_type[p] == AV1H_RESTORE_NONE) continue;
    const int sy = p > 0 ? ss_y : 0;
    const int usz = 1 << hdr->restoration_unit_log2[p];
    const int ucols = imax(hdr->restoration_unit_cols[p], 1);
    const int pw = g.out_w[p], ph = g.out_h[p];
    const int fh2 = (hdr->height + (sy ? 1 << (sy - 1) : 0)) >> sy;
    const int urows = imax((fh2 + (usz >> 1)) / usz, 1);
    for (int y = 0; y < ph;) {
      const int stripe = ((y << sy) + 8) / 64;
      const int s0 = (-8 + stripe * 64) >> sy;
      const int s1 = s0 + (64 >> sy) - 1;
      const int y1 = imin(ph, s1 + 1);
      const int urow = imin(urows - 1, (((y << sy) + 8) >> sy) / usz);
      for (int uc = 0; uc < ucols; ++uc) {
        const size_t idx = (size_t)urow * ucols + uc;
        if ((int64_t)idx >= f->num_restoration[p]) continue;
        const av1h_lr_unit* u = &f->restoration[p][idx];
        if (u->type == 0) continue;
        const int ux0 = uc * usz;
        const int ux1 = uc + 1 == ucols ? pw : imin(ux0 + usz, pw);
        if (ux1 <= ux0) continue;
        o_lr j;
        memset(&j, 0, sizeof(j));
        j.origin = g.plane_off[p];
        j.dst = g.plane_off[p] + (size_t)y * g.stride[p] + ux0;
        j.stride = g.stride[p];
        j.x = ux0;
        j.y = y;
        j.w = ux1 - ux0;
        j.h = y1 - y;
        j.plane_w = pw;
        j.plane_h = ph;
        j.stripe0 = s0;
        j.stripe1 = s1;
        /* engine.cpp:1435-1448: the unit's type is a RestorationType, so
         * kRestoreWiener (2) is Wiener and any other non-zero value the self
         * guided filter */
        j.type = u->type == AV1H_RESTORE_WIENER ? 1 : 2;
        if (u->type == AV1H_RESTORE_WIENER) {
          for (int k = 0; k < 3; ++k) {
            j.vf[k] = u->wiener[0][k];
            j.hf[k] = u->wiener[1][k];
          }
        } else {
          const int* sp = SPEC(AV1O_OFF_SGR_PARAMS) + 4 * (u->sgr_set & 15);
          for (int k = 0; k < 2; ++k) {
            j.r[k] = sp[2 * k];
            j.eps[k] = sp[2 * k + 1];
            j.xqd[k] = u->sgr_xqd[k];
          }
        }
        j.bitdepth = bd;
        PUSH(lr, nlr, clr, o_lr, j);
      }
      y = y1;
    }
  }
  uses_lr = nlr > 0;
  if (uses_lr) memcpy(d->lr_curr, frame, g.size * sizeof(px_t));

  /* ---- CDEF (engine.cpp:1299-1347, cpu_backend.cpp:364-380) -------------- */
  if (f->num_cdef > 0) {
    memcpy(d->scratch, frame, g.size * sizeof(px_t));
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < f->num_cdef; ++i) {
      const av1h_cdef_block* c = &f->cdef[i];
      const int p = c->plane;
      const int ch = p > 0;
      o_cdef j;
      j.x = c->x;
      j.y = c->y;
      j.src = j.dst = g.plane_off[p] + (size_t)c->y * g.stride[p] + c->x;
      j.stride = g.stride[p];
      j.luma = g.plane_off[0] + (size_t)(ch ? c->y << ss_y : c->y) * g.stride[0] +
               (ch ? c->x << ss_x : c->x);
      j.luma_stride = g.stride[0];
      j.pri = c->pri_strength;
      j.sec = c->sec_strength;
      j.damping = c->damping;
      j.dir = c->dir;
      j.uv = ch;
      j.ss_x = ss_x;
      j.ss_y = ss_y;
      j.bitdepth = bd;
      j.plane_w = own.grid_w[p];
      j.plane_h = own.grid_h[p];
      o_cdef_block(&j, d->scratch, frame, maxv);
    }
  }

  /* ---- superres (engine.cpp:1353-1380, cpu_backend.cpp:386-411) ---------- */
  if (upscaled) {
    o_upscale up[3];
    for (int p = 0; p < 3; ++p) {
      up[p].plane_off = g.plane_off[p];
      up[p].stride = g.stride[p];
      up[p].rows = g.out_h[p];
      up[p].out_w = g.out_w[p];
      up[p].max_x = g.plane_w[p] - 1;
      av1o_superres_params(g.plane_w[p], g.out_w[p], &up[p].step, &up[p].initial);
      up[p].bitdepth = bd;
    }
    memcpy(d->scratch, frame, g.size * sizeof(px_t));
    for (int p = 0; p < 3; ++p) {
#pragma omp parallel for schedule(static)
      for (int y = 0; y < up[p].rows; ++y) o_upscale_row(&up[p], d->scratch, frame, y, maxv);
    }
    if (uses_lr) {
      for (int p = 0; p < 3; ++p) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < up[p].rows; ++y) o_upscale_row(&up[p], d->lr_curr, d->lr_prev_up, y, maxv);
      }
    }
  }

  /* ---- loop restoration (cpu_backend.cpp:419-432) ------------------------- */
  if (uses_lr) {
    memcpy(d->lr_cdef, frame, g.size * sizeof(px_t));
    const px_t* curr = upscaled ? d->lr_prev_up : d->lr_curr;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nlr; ++i) o_lr_job(&lr[i], d->lr_cdef, curr, frame, maxv);
  }
  free(lr);

  /* ---- reference store before grain (cpu_backend.cpp:439-465) ------------- */
  int slot = -1;
  if (keep) {
    if (reserved >= 0) {
      if (reserved >= SLOTS) return fail(d, AV1H_ERR_INVALID, "invalid reference slot");
      slot = reserved;
    } else {
      slot = alloc_slot(d);
      if (slot < 0) return fail(d, AV1H_ERR_SLOTS, "CPU reference pool exhausted");
    }
    d->used[slot] = 1;
    d->slot_lay[slot] = g;
    memcpy(d->pool + (size_t)slot * g.size, frame, g.size * sizeof(px_t));
  }

  /* ---- film grain (engine.cpp:1467-1479, cpu_backend.cpp:467-486) -------- */
  if (hdr->grain.apply) {
    o_grain gr;
    o_grain_build(&hdr->grain, bd, ss_x, ss_y, hdr->color_matrix == 3, g.out_w[0], g.out_h[0], &gr);
    o_grain_apply(&gr, &hdr->grain, bd, ss_x, ss_y, &g, frame);
    o_grain_free(&gr);
  }

  /* ---- read back (cpu_backend.cpp:522-535) -------------------------------- */
  if (out) {
    for (int p = 0; p < 3; ++p)
      for (int y = 0; y < g.out_h[p]; ++y)
        memcpy(out->plane[p] + (size_t)y * out->stride[p], frame + g.plane_off[p] + (size_t)y * g.stride[p],
               (size_t)g.out_w[p] * sizeof(px_t));
  }
  if (info) {
    info->width = hdr->width;
    info->height = hdr->height;
    info->bit_depth = bd;
    info->ss_x = ss_x;
    info->ss_y = ss_y;
  }
  d->last_slot = slot;
  d->have_frame = 1;
  if (slot_out) *slot_out = slot;
  return AV1H_OK;
}

int av1o_decode_wire(av1o_decoder* d, const uint8_t* wire, size_t len, int keep, int reserved,
                     const av1h_planes* out, av1h_frame_info* info, int* slot_out) {
  av1h_frame_desc f;
  if (av1h_wire_parse(wire, len, &f) != AV1H_OK)
    return fail(d, AV1H_ERR_INVALID, "not a wire frame");
  return av1o_decode(d, &f, keep, reserved, out, info, slot_out);
}

/* gather's slot rule (cpu_backend.cpp:557-587): the latest slot (or -1) is the
 * post-grain frame, any older one its pre-grain reference. The display
 * picture is gathered (the decoded output's geometry). */
static const px_t* gather(av1o_decoder* d, int slot, o_layout* lay) {
  if (slot >= 0) {
    if (slot >= SLOTS || !d->used[slot]) {
      fail(d, AV1H_ERR_INVALID, "invalid reference slot");
      return NULL;
    }
    if (slot != d->last_slot) {
      *lay = d->slot_lay[slot];
      return d->pool + (size_t)slot * d->lay.size;
    }
  }
  if (!d->have_frame) {
    fail(d, AV1H_ERR_INVALID, "no frame to convert");
    return NULL;
  }
  *lay = d->lay;
  return d->frame;
}

int av1o_read_slot(av1o_decoder* d, int slot, const av1h_planes* out, av1h_frame_info* info) {
  o_layout g;
  const px_t* s = gather(d, slot, &g);
  if (!s) return AV1H_ERR_INVALID;
  for (int p = 0; p < 3; ++p)
    for (int y = 0; y < g.out_h[p]; ++y)
      memcpy(out->plane[p] + (size_t)y * out->stride[p], s + g.plane_off[p] + (size_t)y * g.stride[p],
             (size_t)g.out_w[p] * sizeof(px_t));
  if (info) {
    info->width = g.width;
    info->height = g.height;
    info->bit_depth = g.bit_depth;
    info->ss_x = g.ss_x;
    info->ss_y = g.ss_y;
  }
  return AV1H_OK;
}

/* frame_to_nv12 / av1g_nv12_rows - factory.cpp:236-278, av1gpu_kernels.h:3432-3472 */
void o_nv12_pack(const px_t* yp, int ys, const px_t* up, const px_t* vp, int cs, int w, int h,
                 int ss_x, int ss_y, int bd, size_t pitch, uint8_t* out) {
  const int sh = bd - 8;
  const int ch = (h + (1 << ss_y) - 1) >> ss_y, cw = (w + (1 << ss_x) - 1) >> ss_x;
  memset(out, 0, pitch * (size_t)(h + ch));
  for (int r = 0; r < h; ++r)
    for (int x = 0; x < w; ++x) out[(size_t)r * pitch + x] = (uint8_t)(yp[(size_t)r * ys + x] >> sh);
  for (int r = 0; r < ch; ++r)
    for (int x = 0; x < cw; ++x) {
      out[(size_t)(h + r) * pitch + 2 * x] = (uint8_t)(up[(size_t)r * cs + x] >> sh);
      out[(size_t)(h + r) * pitch + 2 * x + 1] = (uint8_t)(vp[(size_t)r * cs + x] >> sh);
    }
}

size_t av1h_nv12_size_oracle(int width, int height, int ss_x, int ss_y, size_t pitch);
size_t av1h_nv12_size_oracle(int width, int height, int ss_x, int ss_y, size_t pitch) {
  if (width <= 0 || height <= 0 || ss_x != 1 || ss_y != 1) return 0;
  const size_t crow = 2 * (((size_t)width + 1) / 2);
  if ((pitch & 1) || pitch < crow) return 0;
  return pitch * ((size_t)height + ((size_t)height + 1) / 2);
}

int av1o_to_nv12(av1o_decoder* d, int slot, size_t pitch, uint8_t* out, size_t out_bytes) {
  o_layout g;
  const px_t* s = gather(d, slot, &g);
  if (!s) return AV1H_ERR_INVALID;
  const size_t need = av1h_nv12_size_oracle(g.width, g.height, g.ss_x, g.ss_y, pitch);
  if (need == 0)
    return fail(d, AV1H_ERR_INVALID,
                "NV12 output needs a 4:2:0 frame and an even pitch at least as wide");
  if (out_bytes < need) return fail(d, AV1H_ERR_INVALID, "output buffer too small");
  o_nv12_pack(s + g.plane_off[0], g.stride[0], s + g.plane_off[1], s + g.plane_off[2], g.stride[1],
              g.width, g.height, g.ss_x, g.ss_y, g.bit_depth, pitch, out);
  return AV1H_OK;
}

/* frame_to_rgb - tables.cpp:535-561 */
int av1o_to_rgb(av1o_decoder* d, int slot, int matrix, int range, uint8_t* out, size_t out_bytes) {
  o_layout g;
  const px_t* s = gather(d, slot, &g);
  if (!s) return AV1H_ERR_INVALID;
  if (out_bytes < (size_t)g.width * g.height * 3) return fail(d, AV1H_ERR_INVALID, "output buffer too small");
  int m[8], o[3];
  av1o_color_matrix(matrix, range, g.bit_depth, m, o);
  for (int y = 0; y < g.height; ++y) {
    const px_t* yr = s + g.plane_off[0] + (size_t)y * g.stride[0];
    const px_t* ur = s + g.plane_off[1] + (size_t)(y >> g.ss_y) * g.stride[1];
    const px_t* vr = s + g.plane_off[2] + (size_t)(y >> g.ss_y) * g.stride[2];
    uint8_t* row = out + (size_t)y * g.width * 3;
    for (int x = 0; x < g.width; ++x) {
      const int yy = yr[x], uu = ur[x >> g.ss_x], vv = vr[x >> g.ss_x];
      const int r = (m[0] * yy + m[1] * vv + o[0]) >> 16;
      const int gg = (m[2] * yy + m[3] * uu + m[4] * vv + o[1]) >> 16;
      const int b = (m[5] * yy + m[6] * uu + o[2]) >> 16;
      row[3 * x] = (uint8_t)iclip(r, 0, 255);
      row[3 * x + 1] = (uint8_t)iclip(gg, 0, 255);
      row[3 * x + 2] = (uint8_t)iclip(b, 0, 255);
    }
  }
  return AV1H_OK;
}

/* itx of one block, for the known-answer tests */
void av1o_itx_add(const int32_t* coef, int w, int h, int tx_type, int lossless, int bd,
                  uint16_t* dst, int stride) {
  int32_t tmp[64 * 64];
  o_itx j;
  memset(&j, 0, sizeof(j));
  j.stride = stride;
  j.w = w;
  j.h = h;
  j.tx_type = tx_type;
  j.eob = 1;
  j.lossless = lossless;
  j.row_shift = lossless ? 0 : av1o_tx_row_shift(w, h);
  j.bitdepth = bd;
  for (int y = 0; y < h; ++y) o_itx_row(&j, coef, tmp, y);
  for (int x = 0; x < w; ++x) o_itx_col(&j, tmp, dst, x, (1 << bd) - 1);
}

/* ---- dequantisation of coded levels ---------------------------------------
 * Restates the reference's av1g_dequant_one (kernels/av1gpu_kernels.h:
 * 1024-1083): the raster of a transform is min(32, w) x min(32, h), its
 * levels come in scan order (a diagonal scan for 2D types, row by row for
 * V_* types, column by column for H_*: av1gpu_kernels.h:1036-1056 with the
 * scans of src/frontend/coef_tables.hpp), each level times its step (the DC
 * step at position 0, weighted by the quantiser matrix entry when there is
 * one), the magnitude cut to 24 bits, divided by the denominator, clipped to
 * 7 + bit depth bits (1065-1081). */

/* raster position of scan index i of a w x h raster (sides 4..32) of class
 * cls (0 2D, 1 horizontal, 2 vertical) */
static int scan_position(int cls, int w, int h, int i) {
  if (cls == 2) return i;
  if (cls == 1) return (i % h) * w + i / h;
  /* the diagonal scan: anti-diagonals in order; on a square raster the
   * even ones run from the left column up to the top row, the odd ones back;
   * on a wider raster every one runs upwards, on a taller one downwards */
  int n = 0;
  for (int d = 0; d < w + h - 1; ++d) {
    const int r0 = d - (w - 1) > 0 ? d - (w - 1) : 0, r1 = d < h - 1 ? d : h - 1;
    const int len = r1 - r0 + 1;
    if (i < n + len) {
      const int k = i - n;
      const int up = w == h ? (d % 2 == 0) : (w > h);
      const int r = up ? r1 - k : r0 + k;
      return r * w + (d - r);
    }
    n += len;
  }
  return -1;
}

int64_t av1o_dequant_frame(const av1h_frame_desc* f, int32_t* raster) {
  if (f->num_levels <= 0) return -1;
  for (int64_t i = 0; i < f->num_coefficients; ++i) raster[i] = 0;
  int64_t n = 0;
  for (int64_t bi = 0; bi < f->num_blocks; ++bi) {
    const av1h_block* b = &f->blocks[bi];
    if (b->residual_provided) continue;
    for (int k = 0; k < b->tx_count; ++k) {
      const av1h_transform* t = &f->transforms[b->tx_first + k];
      if (t->eob <= 0) continue;
      const int w = t->w < 32 ? t->w : 32, h = t->h < 32 ? t->h : 32, area = w * h;
      const int type = t->tx_type;
      const int cls = (type == 10 || type == 12 || type == 14) ? 2 : (type == 11 || type == 13 || type == 15) ? 1 : 0;
      const int count = t->eob < area ? t->eob : area;
      for (int i = 0; i < count; ++i) {
        const int level = f->levels[t->level_offset + i];
        if (!level) continue;
        const int pos = scan_position(cls, w, h, i);
        int step = pos == 0 ? t->dc_step : t->ac_step;
        if (t->qm_off >= 0) step = (step * f->qm_pool[t->qm_off + pos] + 16) >> 5;
        const long long product = (long long)level * step;
        long long mag = product < 0 ? -product : product;
        mag &= 0xFFFFFF;
        const long long scaled = mag / t->denominator;
        long long v = product < 0 ? -scaled : scaled;
        const long long lim = 1LL << (7 + b->bit_depth);
        if (v > lim - 1) v = lim - 1;
        if (v < -lim) v = -lim;
        raster[t->coef_offset + pos] = (int32_t)v;
      }
      ++n;
    }
  }
  return n;
}

int64_t av1o_dequant_wire(const uint8_t* wire, size_t len, int32_t* raster, int64_t cap) {
  av1h_frame_desc d;
  if (av1h_wire_parse(wire, len, &d) != AV1H_OK || cap < d.num_coefficients) return -2;
  return av1o_dequant_frame(&d, raster);
}
