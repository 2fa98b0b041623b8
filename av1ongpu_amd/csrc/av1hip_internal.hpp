/*
 * Internal definitions shared by the host side and the gfx950 kernels of
 * libav1hip.so: the device frame layout, the compact per frame work
 * descriptors the host flattens FrameSyntax into, and the kernel launchers.
 */
#ifndef AV1HIP_INTERNAL_HPP
#define AV1HIP_INTERNAL_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "av1hip.h"

namespace av1hip {

constexpr int kSlots = 32;          /* reference pictures per context          */
constexpr int kWavesPerGroup = 2;   /* recon kernels: independent waves per WG  */

/* ---- frame layout --------------------------------------------------------
 * One flat buffer of 16 bit samples, the three planes back to back. The stride
 * follows the display width plus a border of 64 luma / 32 chroma samples, and
 * is rounded up to 128 samples (256 bytes) so rows start on cache lines. */
struct Layout {
  int width, height, bit_depth, ss_x, ss_y;
  int plane_w[3], plane_h[3]; /* coded visible samples            */
  int out_w[3], out_h[3];     /* display samples                  */
  int grid_w[3], grid_h[3];   /* coded size rounded up to 8 luma  */
  int stride[3];
  int64_t plane_off[3];
  int64_t size;               /* samples                          */
  int cell_shift[3];          /* dependency cell size per plane   */
  int cells_w[3], cells_h[3];
  int64_t cell_off[3];
  int64_t cells;
};

void make_layout(Layout* g, int width, int height, int bit_depth, int ss_x, int ss_y,
                 int coded_width);

/* ---- per frame descriptors --------------------------------------------------- */

enum BlockKind { kInter = 0, kIntra = 1 };

/* flag bits of DevBlock::flags */
enum {
  kHaveLeft = 1, kHaveAbove = 2, kHaveAR = 4, kHaveBL = 8, kFilterIntra = 16,
  kPalette = 32, kLossless = 64, kResidualRaw = 128, kInterIntra = 256,
  kIIWedge = 512, kWarp0 = 1024, kWarp1 = 2048, kMaskWarp0 = 4096, kMaskWarp1 = 8192,
  kIntraBC = 16384
};

struct DevBlock {
  int32_t x, y, w, h;
  int32_t plane, kind, flags, mode;
  int32_t angle, filter_type, fi_mode, cfl_alpha;
  int32_t tx_first, tx_count, obmc_first, obmc_count;
  int32_t slot[2];       /* pool slots, -1: the frame being decoded              */
  int32_t mv[2][2];
  int32_t filt, filt_v, compound, w0, w1;
  int32_t ref_w, ref_h;  /* visible reference size, for clamping                */
  int32_t scale_x, scale_y;
  int32_t mask_kind, mask_sign, wedge_dir, wedge_xoff, wedge_yoff, wedge_flip;
  int32_t ii_mode;
  int32_t palette_off, palette_stride;
  int32_t palette[8];
  int32_t warp_params[2][6];
  int32_t warp_abcd[2][4];
  /* the cells this block must wait for (the reference scheduler's read set):
   * left column rows [dep_y0, dep_y1], above row columns [dep_x0, dep_x1],
   * the corner when dep_corner; dep_x1 < x and dep_y1 < y disable a side */
  int32_t dep_x1, dep_y1, dep_corner;
  int32_t dep_rx0, dep_ry0, dep_rx1, dep_ry1; /* intra block copy: source cells */
};

/* flag bits of DevTx::flags */
enum { kTxLossless = 1, kTxRaw = 2, kTxToResidual = 4 };

struct DevTx {
  int32_t x, y, w, h;   /* in the plane                      */
  int32_t coef;         /* raster offset in the coefficient (or residual) pool */
  int32_t tx_type;
  int32_t row_shift;
  int32_t eob;
  int32_t plane;
  int32_t flags;        /* kTx*: lossless (WHT), raw residual, intra block  */
};

/* transform classes: the larger side, 4 8 16 32 64 */
constexpr int kTxClasses = 5;
/* Motion compensation tasks: at most kMcSamples samples and kMcItems items
 * (blocks, or row strips of large blocks) whose reference windows fit
 * kMcWin samples and whose horizontal pass fits kMcMid values. The host
 * computes each item's positions and windows (the position arithmetic of the
 * reference, av1gpu_kernels.h:1761-1763 and 1819-1840) and its place in the
 * task's LDS, so a workgroup reads one descriptor array and nothing else. */
constexpr int kMcSamples = 256;
constexpr int kMcItems = 16;
constexpr int kMcWin = 6144;
constexpr int kMcMid = 3072;

struct McRefDev {
  int32_t slot;
  int32_t scaled, ix, iy, fx, fy, sx0, sy0, stx, sty; /* positions */
  int32_t wx0, wy0, ww, wh;                           /* reference window */
  int32_t win_off, mid_off;                           /* in the task's LDS */
};

struct McItemDev {
  int32_t x, y, w, h, plane, y0, rows, px0; /* block, strip, first sample in the task */
  int32_t compound, w0, w1, hk, vk, ref_w, ref_h, nref;
  McRefDev r[2];
};

struct McTaskDev {
  int32_t first, count, wtotal, mtotal, samples, pad[3];
};

constexpr int kLrTileCols = 64, kLrTileRows = 16; /* loop restoration tile per workgroup */

struct DevObmc {
  int32_t pass, rx, ry, w, h, mvx, mvy, filt, filt_v, slot;
};

/* one line of one deblocking edge: the four samples p1..q1 plus the taps the
 * filter may read, at `p1` + k * cross */
struct DevDeblock {
  int32_t plane, x, y;   /* q0 of the first line                */
  int32_t vertical, len, level, filter_size;
  int32_t line0;         /* prefix: first global line index   */
};

struct DevCdef {
  int32_t plane, x, y, pri, sec, damping, dir, pad;
};

struct DevLr {
  int32_t plane, x, y, w, h;
  int32_t stripe0, stripe1, type;
  int32_t vf[3], hf[3];
  int32_t r[2], eps[2], xqd[2];
  int32_t px0;           /* prefix: first output sample index */
};

/* grain parameters the kernels read */
struct DevGrain {
  int32_t seed, bit_depth, num_y, num_cb, num_cr, cfl;
  int32_t lag, shift, scale_shift, scaling_shift, overlap;
  int32_t cb_mult, cb_luma_mult, cb_offset, cr_mult, cr_luma_mult, cr_offset;
  int32_t min_value, max_luma, max_chroma;
  int32_t ar_y[24], ar_cb[25], ar_cr[25];
  int32_t lut[768];
  int32_t stripes, blocks; /* offsets table size                 */
};

/* The frame plan as the kernels see it: pointers into one device buffer. */
struct FramePlanDev {
  const DevBlock* blocks;
  const DevTx* tx;
  const DevObmc* obmc;
  const int32_t* inter_list;
  const int32_t* queue_list;
  const int32_t* special_list; /* inter blocks with warp, masks or OBMC */
  const int32_t* coef;
  const int32_t* residual;
  const uint8_t* palette;
  const DevDeblock* deblock;  /* vertical edges first, then horizontal */
  const DevCdef* cdef;
  const DevLr* lr;
  const DevGrain* grain;
  /* motion compensation: item descriptors in coding order, grouped into tasks */
  const McItemDev* mc_items;
  const McTaskDev* mc_task;
  int32_t mc_tasks;
  /* transform blocks: set 0 those of inter blocks, set 1 those of intra
   * blocks; inside a set by class (larger side 4 8 16 32 64), inside a class
   * by size; set s class c is itx_list[itx_off[s][c] .. itx_off[s][c + 1]).
   * itx_offs is the same table in device memory. */
  const int32_t* itx_list;
  const int32_t* itx_offs;
  int32_t itx_off[2][kTxClasses + 1];
  int32_t n_special;
  int32_t n_inter, n_queue, n_deblock_v, n_deblock_h, lines_v, lines_h, n_cdef, n_lr, lr_px;
  /* deblocking: the first edge of every 256 line chunk, vertical pass chunks
   * then horizontal */
  const int32_t* db_chunk;
  /* CDEF: luma blocks first; the luma 8x8 blocks (x8 | y8 << 16) whose
   * direction the listed blocks need */
  const int32_t* cdef_dir_list;
  int32_t n_cdef_luma, n_cdef_dir;
  /* loop restoration tiles: (job, x offset | y offset << 16) */
  const int2* lr_tiles;
  int32_t n_lr_tiles;
};

/* launchers (one per translation unit, all on the context's stream) */
struct ReconArgs {
  FramePlanDev plan;
  Layout lay;
  uint16_t* frame;
  const uint16_t* pool;
  int64_t slot_size;
  uint32_t* cells;      /* epoch per dependency cell          */
  uint32_t epoch;
  uint32_t* queue_counter;
  uint32_t* error;      /* bit 0: a dependency wait timed out  */
  int16_t* res;         /* residual of intra blocks, frame layout */
  uint64_t* trace;      /* null, or 4 timestamps per queue entry (diagnostics) */
  int maxv;
  int edge_filter;      /* enable_intra_edge_filter           */
};

/* motion compensation of every inter block into the frame */
hipError_t launch_mc(const ReconArgs& a, int cus, hipStream_t s, int* launches);
/* the inter blocks with warp, masks or OBMC, after motion compensation */
hipError_t launch_special(const ReconArgs& a, hipStream_t s, int* launches);
/* the transforms of one set: 0 inter blocks (residual added into the frame,
 * after motion compensation), 1 intra blocks (residual into `res`) */
hipError_t launch_itx(const ReconArgs& a, int set, hipStream_t s, int* launches);
/* the dependency queue over the intra blocks */
hipError_t launch_queue(const ReconArgs& a, int cus, hipStream_t s, int* launches);
/* workgroups a transform set launches (0: nothing to do) */
int itx_groups(const FramePlanDev& p, int set);

struct FilterArgs {
  FramePlanDev plan;
  Layout lay;
  uint16_t* frame;
  uint16_t* scratch;
  uint16_t* lr_curr;
  uint16_t* lr_cdef;
  uint16_t* lr_prev_up;
  int32_t* cdef_dir;    /* (direction, variance) per luma 8x8 block */
  int cdef_dir_w;       /* luma 8x8 blocks per row                  */
  int sharpness;
  int maxv;
};

hipError_t launch_deblock(const FilterArgs& a, hipStream_t s, int* launches);
hipError_t launch_cdef(const FilterArgs& a, hipStream_t s, int* launches);
/* upscales src into dst, and src2 into dst2 when src2 is not null, in one launch */
hipError_t launch_superres(const FilterArgs& a, const uint16_t* src, uint16_t* dst, const uint16_t* src2,
                           uint16_t* dst2, hipStream_t s, int* launches);
hipError_t launch_lr(const FilterArgs& a, const uint16_t* cdef, const uint16_t* curr, hipStream_t s,
                     int* launches);
/* the three grain templates and the block offsets (one workgroup; depends only
 * on the grain parameters), then the blend into the picture */
hipError_t launch_grain_template(const Layout& lay, const DevGrain* g, int16_t* templates, uint8_t* offsets,
                                 hipStream_t s, int* launches);
hipError_t launch_grain_apply(const Layout& lay, const DevGrain* g, uint16_t* frame, const int16_t* templates,
                              const uint8_t* offsets, hipStream_t s, int* launches);
hipError_t launch_nv12(const Layout& lay, const uint16_t* src, uint8_t* out, int pitch, hipStream_t s);
hipError_t launch_rgb(const Layout& lay, const uint16_t* src, uint8_t* out, const int m[8],
                      const int o[3], hipStream_t s);

/* film grain: the LFSR jump matrices the template kernel uses */
void grain_jump_matrices(uint16_t m82[16], uint16_t m44[16]);

}  // namespace av1hip

#endif
