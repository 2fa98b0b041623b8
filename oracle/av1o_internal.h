/*
 * Shared helpers and job descriptors of the CPU oracle (test infrastructure).
 * The jobs carry what the reference's job structs carry
 * (kernels/av1gpu_kernels.h:101-371), derived the way
 * src/core/engine.cpp:983-1462 derives them; the layout and names are this
 * file's own.
 */
#ifndef AV1O_INTERNAL_H
#define AV1O_INTERNAL_H

#include <stdint.h>
#include <stdlib.h>

#include "av1o.h"
#include "av1o_spec_tables.h"

typedef uint16_t px_t;

#define SPEC(off) (av1o_spec_table + (off))

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }
/* Round2 for n >= 1 (av1gpu_kernels.h:484-487 with the n <= 0 guard) */
static inline int rnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
static inline int rnd2_64(int64_t v, int n) {
  return n <= 0 ? (int)v : (int)((v + ((int64_t)1 << (n - 1))) >> n);
}
/* Round2Signed on the magnitude (av1gpu_kernels.h:489-493) */
static inline int rnd2s(int v, int n) {
  if (n <= 0) return v;
  return v >= 0 ? (v + (1 << (n - 1))) >> n : -((-v + (1 << (n - 1))) >> n);
}
/* the signed rounding MC positions use (av1gpu_kernels.h:81-84) */
static inline int rnd2s_add(int v, int n) {
  const int s = v < 0 ? -1 : 0;
  return (v + (1 << (n - 1)) + s) >> n;
}
static inline int ilog2_ceil(int v) {
  int n = 0;
  while ((1 << n) < v) ++n;
  return n;
}
static inline int ilog2_floor(int v) {
  int n = 0;
  while (v > 1) {
    v >>= 1;
    ++n;
  }
  return n;
}

/* ---- layout (reference make_plan_layout, factory.cpp:114-152) ----------- */
typedef struct {
  int width, height, bit_depth, ss_x, ss_y; /* display geometry       */
  int plane_w[3], plane_h[3];               /* coded visible samples  */
  int out_w[3], out_h[3];                   /* display samples        */
  int stride[3];
  size_t plane_off[3];
  size_t size;
} o_layout;

void o_make_layout(o_layout* g, int width, int height, int bit_depth, int ss_x, int ss_y,
                   int coded_width, int align);

/* ---- jobs ---------------------------------------------------------------- */
typedef struct {
  size_t dst;        /* offset of the block's first sample           */
  int stride;
  int x, y, w, h, plane;
  int mode, angle, bitdepth;
  int have_left, have_above, have_ar, have_bl;
  int filter_type;   /* -1: edge filter off                           */
  int filter_intra, fi_mode;
  int cfl_alpha;
  size_t luma;       /* start of the luma plane                       */
  int luma_stride, luma_w, luma_h;
  int ss_x, ss_y;
  int max_x, max_y;
  int palette, palette_color[8], palette_off, palette_stride;
} o_intra;

typedef struct {
  int pass, rx, ry, w, h, mvx, mvy, filt, filt_v, slot;
  size_t ref; /* absolute offset of the plane in the pool or current frame */
} o_obmc;

typedef struct {
  size_t dst;
  int stride, plane, w, h;
  int x, y;            /* patch origin in the plane                        */
  int ref_w, ref_h;    /* visible reference size                           */
  int slot[2];         /* pool slots, -1: the frame being decoded          */
  size_t ref[2];       /* absolute offsets of the reference planes         */
  int mv[2][2];
  int sub_x, sub_y;
  int scale_x, scale_y;
  int filt, filt_v;
  int compound, w0, w1, bitdepth;
  int mask_kind, mask_sign, wedge_dir, wedge_xoff, wedge_yoff, wedge_flip;
  size_t luma_ref[2];
  int luma_stride, luma_x, luma_y, luma_ref_w, luma_ref_h;
  int nobmc;
  o_obmc obmc[8];
  int warp[2], warp_params[2][6], warp_abcd[2][4], mask_warp[2];
  int interintra, ii_mode, ii_wedge;
  int have_left, have_above, max_x, max_y;
} o_inter;

typedef struct {
  size_t coef, tmp, dst;
  int stride, w, h, tx_type, eob, row_shift, lossless, bitdepth;
} o_itx;

typedef struct {
  size_t p1; /* sample p1 of the first group */
  int cross, along, len, level, sharpness, filter_size, chroma, bitdepth;
} o_deblock;

typedef struct {
  int x, y;
  size_t src, dst, luma;
  int stride, luma_stride;
  int pri, sec, damping, dir, uv, ss_x, ss_y, bitdepth, plane_w, plane_h;
} o_cdef;

typedef struct {
  size_t plane_off; /* same offset in source and target buffers */
  int stride, rows, out_w, max_x, step, initial, bitdepth;
} o_upscale;

typedef struct {
  size_t dst, origin; /* dst: sample (x, y); origin: sample (0, 0) of the plane */
  int stride, x, y, w, h, plane_w, plane_h, stripe0, stripe1, type;
  int vf[3], hf[3], r[2], eps[2], xqd[2], bitdepth;
} o_lr;

/* ---- maths (av1o_dsp.c) -------------------------------------------------- */
void o_itx_row(const o_itx* j, const int32_t* coef, int32_t* tmp, int y);
void o_itx_col(const o_itx* j, const int32_t* tmp, px_t* frame, int x, int maxv);
void o_residual_add(const o_itx* j, const int32_t* res, px_t* frame, int maxv);
void o_intra_block(const o_intra* j, px_t* frame, const uint8_t* palette, int maxv);
void o_inter_block(const o_inter* j, px_t* frame, const px_t* pool, int maxv);
void o_deblock_job(const o_deblock* j, px_t* frame, int maxv);
void o_cdef_block(const o_cdef* j, const px_t* src, px_t* dst, int maxv);
void o_upscale_row(const o_upscale* j, const px_t* src, px_t* dst, int y, int maxv);
void o_lr_job(const o_lr* j, const px_t* cdef, const px_t* curr, px_t* dst, int maxv);
int o_wedge_entry(int bw, int bh, int index, int* dir, int* xoff, int* yoff, int* flip);

/* ---- grain (av1o_grain.c) ------------------------------------------------ */
typedef struct {
  int16_t* noise; /* three planes back to back */
  size_t off[3];
  int stride[3];
  int lut[768];
  int min_value, max_luma, max_chroma;
  int planes;
} o_grain;
void o_grain_build(const av1h_grain* g, int bit_depth, int ss_x, int ss_y, int identity,
                   int luma_w, int luma_h, o_grain* out);
void o_grain_free(o_grain* g);
void o_grain_apply(const o_grain* g, const av1h_grain* p, int bit_depth, int ss_x, int ss_y,
                   const o_layout* lay, px_t* frame);

/* ---- packing (av1o_frame.c) ----------------------------------------------- */
void o_nv12_pack(const px_t* y, int ys, const px_t* u, const px_t* v, int cs, int w, int h,
                 int ss_x, int ss_y, int bitdepth, size_t pitch, uint8_t* out);

#endif
