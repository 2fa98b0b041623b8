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
  /* packed edges: per plane, the last sample row of every cell row
   * (edge_row_words 64-bit words each) and the last sample column of every
   * cell column (edge_col_words each); a word holds two samples and the
   * epoch of the frame that wrote them */
  int edge_row_words[3], edge_col_words[3];
  int64_t edge_row_off[3], edge_col_off[3];
  int64_t edge_words;
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
  kIntraBC = 16384,
  /* queue protocol: kPackedEdges - every edge sample the block reads comes
   * from a queued block's packed edge words or from the frame where no
   * queued block wrote it (no cell wait); kCornerRow / kCornerFrame - its
   * corner sample is in the packed row / the frame, else the packed column;
   * kEdgeRow / kEdgeCol - the block publishes its last row / column as
   * packed words; kSkipMark - no block waits on its cells */
  kPackedEdges = 1 << 15, kCornerRow = 1 << 16, kEdgeRow = 1 << 17, kEdgeCol = 1 << 18, kSkipMark = 1 << 19,
  kCornerFrame = 1 << 20
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
  /* directional intra: dir_a the prediction angle (bits 0-8) and the edge
   * preparation (kDir*), dir_b the two Dr_Intra_Derivative steps (dx in
   * bits 0-15, dy in 16-31), all fixed by the syntax */
  int32_t dir_a, dir_b;
  /* kPackedEdges: the source of each cell of the above / left span (bit i:
   * 1 a packed word of a queued block, 0 the frame, final before the queue) */
  int32_t edge_mask_a, edge_mask_l;
};

enum { kDirFilter = 1 << 9, kDirCorner = 1 << 10, kDirUpA = 1 << 11, kDirUpL = 1 << 12 };
constexpr int kDirStrA = 13, kDirStrL = 16; /* 3 bit edge filter strengths */

/* flag bits of DevTx::flags */
enum { kTxLossless = 1, kTxRaw = 2, kTxToResidual = 4 };

/* one transform block in 16 bytes */
struct DevTx {
  uint16_t x, y;        /* in the plane                      */
  int32_t coef;         /* raster offset in the coefficient (or residual) pool */
  uint16_t eob;         /* end of block, at most 65535 (the device reads min(eob, w * h)) */
  uint8_t w, h;
  uint8_t tx_type, plane;
  uint8_t flags;        /* kTx*: lossless (WHT), raw residual, intra block  */
  int8_t row_shift;
};
static_assert(sizeof(DevTx) == 16, "transform record layout");

/* the dequantiser of a transform whose frame carries coded levels (one per
 * DevTx, indexed alike): where its levels start, its quantiser matrix (-1:
 * none), its DC and AC steps and its denominator as a shift */
struct DevDq {
  int32_t level_off, qm_off;
  uint16_t dc_step, ac_step;
  int32_t denom_shift;
};


/* transform classes: the larger side, 4 8 16 32 64 */
constexpr int kTxClasses = 5;
/* Motion compensation tasks: at most kMcSamples samples and kMcItems items
 * (blocks, or row strips of large blocks) whose horizontal pass fits kMcMid
 * values. The host computes each item's positions and reference windows (the
 * position arithmetic of the reference, av1gpu_kernels.h:1761-1763 and
 * 1819-1840) and its place in the task's LDS, so a workgroup reads one
 * descriptor array and the reference rows and nothing else. */
constexpr int kMcSamples = 512;
constexpr int kMcItems = 32;
constexpr int kMcMid = 4096;

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

/* one deblocking edge in 16 bytes: its first global line index (prefix),
 * q0 of its first line, its length in lines, level, filter size, plane
 * (bits 0-1) and direction (bit 2: vertical) */
struct DevDeblock {
  int32_t line0;
  uint16_t x, y;
  uint16_t len;
  uint8_t level, filter_size;
  uint8_t plane_vertical;
  uint8_t pad[3];
};
static_assert(sizeof(DevDeblock) == 16, "deblocking record layout");

/* one CDEF block in 8 bytes: its position in the plane, strengths, plane
 * (bits 0-1) and damping (above), direction (-1: searched on the device) */
struct DevCdef {
  uint16_t x, y;
  uint8_t pri, sec;
  uint8_t plane_damping;
  int8_t dir;
};
static_assert(sizeof(DevCdef) == 8, "CDEF record layout");

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
  /* the dequantised coefficients: int16 when every value of the frame fits
   * (coef16), else int32; the lists above index `blocks`, which holds the
   * descriptors of queued and special blocks only */
  const void* coef;
  int32_t coef16;
  /* a frame with coded levels: the transforms dequantise while they stage
   * their coefficients (levels, matrices, one DevDq per transform); coef is
   * then unused. Null otherwise. */
  const int16_t* levels;
  const uint8_t* qm;
  const DevDq* dq;
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

/* ---- one decode: the frames of a batch ---------------------------------- */

/* What the reconstruction kernels read of one frame. */
/* Polls a queued block's dependency wait makes before it gives up and the
 * decode reports a time-out: about a second of polling. */
constexpr int kSpinLimitDefault = 1 << 22;

struct ReconArgs {
  FramePlanDev plan;
  Layout lay;
  uint16_t* frame;
  const uint16_t* pool;
  int64_t slot_size;
  uint32_t* cells;      /* epoch per dependency cell          */
  uint64_t* edges;      /* packed edge words (Layout::edge_*)  */
  uint32_t epoch;
  int16_t* res;         /* residual of intra blocks, frame layout */
  int maxv;
  int edge_filter;      /* enable_intra_edge_filter           */
  int spin_limit;       /* polls a dependency wait makes before it reports a time-out */
};

/* What the filter kernels read of one frame. */
struct FilterArgs {
  FramePlanDev plan;
  Layout lay;
  uint16_t* frame;
  uint16_t* scratch;
  uint16_t* lr_curr;
  uint16_t* lr_cdef;
  uint16_t* lr_prev_up;
  const uint16_t* cdef_src; /* the deblocked picture CDEF reads: lr_curr or scratch */
  uint16_t* cdef_dst;   /* where CDEF writes: the frame, or scratch (superres reads it) */
  uint16_t* cdef_dst2;  /* a second destination of CDEF's output (lr_cdef), or null */
  uint16_t* snap2;      /* a second destination of the deblocked snapshot, or null */
  uint16_t* up_dst2;    /* a second destination of the upscaled frame (lr_cdef), or null */
  int32_t* cdef_dir;    /* (direction, variance) per luma 8x8 block */
  int cdef_dir_w;       /* luma 8x8 blocks per row                  */
  int sharpness;
  int maxv;
  int32_t up_step[3], up_initial[3]; /* superres per plane          */
  int upscaled, uses_lr, has_grain;
  int16_t* grain_tmpl;
  uint8_t* grain_offs;
  uint16_t* store;      /* reference slot to store into, or null  */
  uint8_t* nv12;        /* NV12 output on the device, or null     */
  int32_t nv12_pitch;
  uint64_t* hash;       /* picture checksum accumulator, or null  */
};

struct FrameJob {
  ReconArgs r;
  FilterArgs f;
};

/* Every kernel runs the work of all frames of a batch in one launch: its
 * workgroups walk "virtual workgroups", numbered frame after frame, and find
 * their frame in a prefix table (kernel k, frame j: prefix[k * (jobs + 1) + j]). */
enum KernelId {
  kKMc, kKSpecial, kKItxInter, kKItxIntra, kKItxInterL, kKItxIntraL, kKDeblockV, kKDeblockH, kKCdefDir,
  kKCdefLuma, kKCdefChroma,
  kKUpscale, kKLr, kKGrainTemplate, kKGrainApply, kKNv12, kKHash,
  kKCopyLrCurr, kKCopyCdefSrc, kKCopyUpSrc, kKCopyStore,
  kKernels
};

/* The frame table, the prefix table and the plan descriptors are written
 * before a launch and only read while kernels run: read through the constant
 * address space, the reads the compiler sees are uniform become scalar loads
 * that stay valid across the kernels' stores. */
#define AV1H_CMEM __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const AV1H_CMEM T* cmem(const T* p) {
  return (const AV1H_CMEM T*)p;
}

/* the frame a virtual workgroup belongs to: the last j < n with pre[j] <= v */
/* This workgroup's virtual workgroup in the sweep of the grid that starts
 * at `base` (sweeps of gridDim.x until `total`), or -1: the sweep's r
 * virtual workgroups are cut into 8 consecutive runs, run x for the
 * workgroups on XCD x (dispatched round robin: workgroup b on XCD b % 8),
 * so neighbouring blocks, which read overlapping reference and picture
 * data, share one XCD's L2, and a sweep shorter than the grid still spreads
 * over all XCDs. */
__device__ __forceinline__ int xcd_group(int base, int total) {
  const int g = (int)gridDim.x, b = (int)blockIdx.x;
  if (g & 7) return base + b < total ? base + b : -1;
  const int r = min(g, total - base), c = (r + 7) >> 3, j = b >> 3, v = (b & 7) * c + j;
  return j < c && v < r ? base + v : -1;
}

__device__ __forceinline__ int find_job(const AV1H_CMEM int32_t* pre, int n, int v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (pre[m] <= v)
      lo = m;
    else
      hi = m - 1;
  }
  return lo;
}

/* Queue entries of a batch: frame << kQueueJobShift | block. A batch of one
 * frame hands out the frame's own queue list. */
constexpr int kQueueJobShift = 24;
constexpr int kMaxBatch = 128;

/* CDEF: groups of 256 / (block samples) blocks a workgroup takes per step
 * (a thread's samples of up to kCdefMaxBatch blocks computed together) */
constexpr int kCdefMaxBatch = 4;
/* CDEF direction search: luma 8x8 blocks per workgroup (one per lane) */
constexpr int kCdefDirPerGroup = 256;

/* samples one checksum workgroup reads */
constexpr int kHashChunk = 4096;

/* The picture checksum: the display samples of Y, U, V in raster order, the
 * i-th weighted by mix64(i), summed modulo 2^64 (mix64: the SplitMix64
 * finaliser of i + 1). Order independent to compute, position dependent in
 * value: any change of a sample or of where it is changes it. */
__host__ __device__ inline uint64_t mix64(uint64_t i) {
  uint64_t z = (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* bytes one copy workgroup moves */
constexpr int kCopyChunk = 16384;

struct BatchDev {
  const FrameJob* jobs;
  const int32_t* prefix;
  int32_t njobs;
  const uint32_t* queue;     /* intra blocks of all frames in hand-out order, or null */
  uint32_t* queue_counter;
  uint32_t* error;           /* bit 0: a dependency wait timed out   */
  uint64_t* trace;           /* null, or 4 timestamps per queue entry */
  uint16_t jump82[16], jump44[16]; /* grain LFSR jump matrices        */
};

typedef const AV1H_CMEM ReconArgs& RArgs;
typedef const AV1H_CMEM FilterArgs& FArgs;
typedef const AV1H_CMEM Layout& LayRef;
typedef const AV1H_CMEM FramePlanDev& PlanRef;
typedef const AV1H_CMEM DevBlock& BlkRef;

/* kernel k's row of the prefix table */
__device__ __forceinline__ const AV1H_CMEM int32_t* kernel_prefix(const BatchDev& B, int k) {
  return cmem(B.prefix) + k * (B.njobs + 1);
}
__device__ __forceinline__ RArgs recon_args(const BatchDev& B, int j) { return cmem(B.jobs)[j].r; }

/* One layout field of the three planes in scalar registers. Indexed by a
 * plane that differs between lanes, a field in constant memory becomes a
 * vector load on the dependent path of every sample; read once here, the
 * planes' values are selected instead. */
template <typename T>
struct PlaneVals {
  T v0, v1, v2;
  __device__ __forceinline__ T operator[](int p) const { return p == 0 ? v0 : (p == 1 ? v1 : v2); }
};
__device__ __forceinline__ int32_t sreg(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t sreg(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ PlaneVals<T> plane_vals(const AV1H_CMEM T* f) {
  return {sreg(f[0]), sreg(f[1]), sreg(f[2])};
}
__device__ __forceinline__ FArgs filter_args(const BatchDev& B, int j) { return cmem(B.jobs)[j].f; }

/* Kernel k over every frame of the batch, `total` virtual workgroups; `cap`
 * caps the grid, whose workgroups then loop. By file: */
hipError_t launch_recon_kernel(const BatchDev& b, int k, int total, int cap, hipStream_t s);
hipError_t launch_filter_kernel(const BatchDev& b, int k, int total, int cap, int cdef_batch, hipStream_t s);
hipError_t launch_grain_kernel(const BatchDev& b, int k, int total, int cap, hipStream_t s);

/* Launch shapes the autotuner chooses: persistent workgroups per compute
 * unit of every kernel (the queue's at kKernels), and CDEF blocks per
 * thread. */
constexpr int kTuneQueue = kKernels, kTuneCdefBatch = kKernels + 1, kTuneParams = kKernels + 2;
struct Tuning {
  int32_t v[kTuneParams];
};
/* the dependency queue over the intra blocks of the batch; the entry count
 * is read on the device (the word after the prefix table), so a captured
 * launch serves any batch of the same size. `total` < 0: the grid for any
 * count. */
hipError_t launch_queue(const BatchDev& b, int total, int cap, hipStream_t s, int* launches);
/* the grids every kernel takes in a captured graph (its cap) */
constexpr int kAnyTotal = 1 << 30;
/* superres: step and initial subpixel position of a plane */
void superres_params(int dw, int uw, int* step, int* initial);
/* upscaling workgroups of one frame (1 or 2 pictures) */
int upscale_tiles(const Layout& lay, int pictures);
/* grain blend workgroups of one frame */
int grain_tiles(const Layout& lay);
/* clears n 64 bit words (a kernel, so captured launches carry no memset) */
hipError_t launch_zero64(uint64_t* p, int n, hipStream_t s);
/* The batch table from its pinned staging buffer ([byte count][..][table at
 * kTableHead]) into device memory; the error word and checksums back into
 * the pinned read-back buffer. */
constexpr size_t kTableHead = 256;
hipError_t launch_fetch_table(const void* pin, void* dev, int frames, hipStream_t s);
hipError_t launch_readback(uint32_t* err, const uint64_t* hashes, int n, uint32_t* out_err, uint64_t* out_hash,
                           hipStream_t s);
/* picture conversions outside a decode */
hipError_t launch_nv12(const Layout& lay, const uint16_t* src, uint8_t* out, int pitch, hipStream_t s);
hipError_t launch_rgb(const Layout& lay, const uint16_t* src, uint8_t* out, const int m[8],
                      const int o[3], hipStream_t s);
/* transform workgroups of classes [c0, c1) of one set of a frame (0 inter,
 * 1 intra) */
int itx_groups(const FramePlanDev& p, int set, int c0, int c1);

/* film grain: the LFSR jump matrices the template kernel uses */
void grain_jump_matrices(uint16_t m82[16], uint16_t m44[16]);

}  // namespace av1hip

#endif
