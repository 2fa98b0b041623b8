/*
 * libav1hip.so host side: decoder contexts, frame plans and the C ABI of
 * include/av1hip.h.
 *
 * A context owns, in HBM: the frame being decoded, the snapshots the filter
 * stages read (the CDEF source, the pre-CDEF picture and its upscaled copy,
 * the CDEF output), the 32 slot reference pool, and one epoch word per
 * dependency cell. A frame plan is the syntax flattened into compact
 * descriptors in one device buffer (the reference's FramePlan,
 * src/core/backend.hpp:140-206, without the host wave schedule: dependencies
 * resolve on the device). Decoding runs the stage order of the reference CPU
 * backend (src/backends/cpu/cpu_backend.cpp:149-553) on the context's stream.
 */
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <deque>
#include <functional>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

#define AV1T_TABLE_QUAL static const
#include "av1hip_internal.hpp"
#include "scan_order.hpp"
#include "spec_tables.h"

using namespace av1hip;

namespace av1hip {

void make_layout(Layout* g, int width, int height, int bd, int ss_x, int ss_y, int coded_width) {
  std::memset(g, 0, sizeof(*g));
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
  int64_t size = 0, cells = 0, edges = 0;
  for (int p = 0; p < 3; ++p) {
    const int sx = p ? ss_x : 0, sy = p ? ss_y : 0;
    g->grid_w[p] = ((coded_width + 7) & ~7) >> sx;
    g->grid_h[p] = ((height + 7) & ~7) >> sy;
    const int pad = p == 0 ? 64 : (64 >> (ss_x ? 1 : 0));
    g->stride[p] = (g->out_w[p] + pad + 127) & ~127;
    g->plane_off[p] = size;
    size += (int64_t)g->stride[p] * (g->out_h[p] + pad);
    /* dependency cells: 4 luma samples, 2 for a subsampled chroma axis (the
     * reference's owner grid, engine.cpp:739-750) */
    g->cell_shift[p] = p == 0 ? 2 : (ss_x ? 1 : 2);
    g->cells_w[p] = (g->grid_w[p] + (1 << g->cell_shift[p]) - 1) >> g->cell_shift[p];
    g->cells_h[p] = (g->grid_h[p] + (1 << g->cell_shift[p]) - 1) >> g->cell_shift[p];
    g->cell_off[p] = cells;
    cells += (int64_t)g->cells_w[p] * g->cells_h[p];
    g->edge_row_words[p] = (g->cells_w[p] << g->cell_shift[p]) >> 1;
    g->edge_col_words[p] = (g->cells_h[p] << g->cell_shift[p]) >> 1;
    g->edge_row_off[p] = edges;
    edges += (int64_t)g->cells_h[p] * g->edge_row_words[p];
    g->edge_col_off[p] = edges;
    edges += (int64_t)g->cells_w[p] * g->edge_col_words[p];
  }
  g->size = size;
  g->cells = cells;
  g->edge_words = edges;
}

}  // namespace av1hip

namespace {

double now_ms() {
  using clk = std::chrono::steady_clock;
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Plan {
  Layout lay;
  av1h_frame_header hdr;
  FramePlanDev dev;
  DevGrain grain;
  void* dbuf = nullptr;
  size_t dbytes = 0;
  bool uses_lr = false, upscaled = false, has_grain = false;
  int blocks = 0;
  double flatten_ms = 0, upload_ms = 0; /* host time of the plan's making */
  std::vector<int> slots; /* pool slots the frame reads */
  int64_t alg_bytes[AV1H_STAGES] = {}; /* compulsory HBM bytes per stage */
  /* the queue list in hand-out order, and where each superblock diagonal key
   * starts in it (key_off[k] .. key_off[k + 1]), for merging the queues of a
   * batch */
  std::vector<int32_t> queue_host, key_off;
  /* superblocks (64x64 luma): features, the superblock of every queue entry
   * (static order), the static order itself, intra block copy edges
   * (source superblock, superblock) */
  int sb_w = 0, sb_h = 0;
  std::vector<float> sb_feat;
  std::vector<int32_t> queue_sb, queue_sb_static, queue_static, key_off_static;

  /* the queued blocks in coding order (index qi): block, predecessors
   * (q_dep[q_dep_off[qi] .. q_dep_off[qi + 1])), and the qi of every entry
   * of the static order */
  std::vector<int32_t> q_block, q_dep_off, q_dep, queue_qi, q_area, q_level;
  /* queued blocks, packed-edge readers, edge publishers, blocks no wait
   * reads the cells of (av1h_plan_queue_counts) */
  int32_t q_counts[4] = {0, 0, 0, 0};
};

struct Ctx;
struct Plan;
/* a launched batch, until its completion is taken */
struct JobSpec {
  Ctx* c;
  Plan* plan;
  int keep, reserved;
  const av1h_planes* out;
  av1h_frame_info* info;
  int* slot_out;
  int slot;
};
struct BatchTicket {
  std::vector<JobSpec> js;
  std::vector<char> nv12; /* per frame: NV12 packed */
  int launches = 0;
  bool any_hash = false, live = false, graph = false;
  double t0 = 0, t_sub = 0;
};

/* one entry of the streamed NV12 pack ring: its own device output buffer,
 * the pinned buffer the copy lands in, and the event that ends the copy */
struct PackEntry {
  uint8_t* dev = nullptr;
  size_t dev_cap = 0;
  void* pin = nullptr;
  size_t pin_cap = 0;
  hipEvent_t done = nullptr;
  size_t bytes = 0;
  bool busy = false;
};
constexpr int kPackRing = 3;

/* a pinned staging buffer and the stream its uploads run on */
struct Stager {
  void* pin = nullptr;
  size_t cap = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop;
  int cus = 0;
  std::string err;
  std::mutex mu;
  Layout lay;
  bool have_layout = false;
  size_t cap = 0; /* samples per buffer */
  uint16_t *frame = nullptr, *scratch = nullptr, *lr_curr = nullptr, *lr_cdef = nullptr, *lr_prev_up = nullptr;
  uint16_t* pool = nullptr;
  uint32_t* cells = nullptr;
  int64_t cells_cap = 0;
  uint64_t* edges = nullptr; /* packed edge words, cleared with the cells */
  int64_t edges_cap = 0;
  uint32_t epoch = 0;
  uint32_t* counters = nullptr; /* [0] queue counter, [1] error word */
  int32_t* cdef_dir = nullptr;   /* CDEF direction map                */
  bool trace_on = false;         /* queue timestamps (diagnostics)    */
  int spin_limit = kSpinLimitDefault; /* polls of a dependency wait before a time-out */
  uint64_t* trace = nullptr;
  size_t trace_cap = 0;
  int trace_n = 0;
  const void* trace_plan = nullptr; /* the plan of a traced decode of one frame */
  size_t cdef_dir_cap = 0;
  int16_t* grain_tmpl = nullptr;
  uint8_t* grain_offs = nullptr;
  size_t grain_offs_cap = 0;
  uint8_t* pack = nullptr;
  size_t pack_cap = 0;
  void* host_pin = nullptr;
  size_t host_pin_cap = 0;
  int pool_slots = kSlots;       /* reference slots the pool holds   */
  int pool_alloc = 0;            /* slots of the allocated pool      */
  bool used[kSlots] = {};
  bool has_pic[kSlots] = {};     /* the slot holds a decoded picture */
  Layout slot_lay[kSlots];
  int last_slot = -1;
  bool last_grain = false;       /* the most recent frame had film grain */
  bool have_frame = false;
  Plan scratch_plan;
  av1h_stats stats;
  hipEvent_t ev[4];
  /* grain templates depend only on the grain parameters: they are built on a
   * second stream while the frame reconstructs */
  hipStream_t side = nullptr;
  hipEvent_t ev_side_go = nullptr, ev_side_done = nullptr, ev_itx_done = nullptr;
  bool timing = false;
  hipEvent_t sev[2 * AV1H_STAGES] = {};
  /* the frame table of a batch this context leads: jobs, prefix table,
   * merged queue */
  bool hash_on = false;          /* every decode also checksums its picture */
  bool hash_valid = false;
  uint64_t last_hash = 0;
  uint64_t* hash_dev = nullptr;  /* checksum accumulators of a batch it leads */
  size_t decode_nv12_pitch = 0; /* > 0: every decode also packs NV12 into `pack` */
  bool nv12_current = false;     /* `pack` holds the NV12 of the latest frame       */
  size_t nv12_pitch_done = 0;
  /* two sets, so a batch can be prepared and launched while the one before
   * it runs: the table (device and pinned), the pinned error word and
   * checksums read back after it, its events (start, queue end, end, done) */
  void* batch_dev[2] = {};
  void* batch_pin[2] = {};
  size_t batch_cap[2] = {};
  void* rb_pin[2] = {};
  hipEvent_t bev[2][4] = {};
  BatchTicket ticket[2];
  std::deque<int> pending; /* launched batches this context leads, oldest first */
  int next_buf = 0;
  /* launched batches holding a frame of this context, and their lead: the
   * context's pictures are final only once they are finished */
  int inflight = 0;
  Ctx* inflight_lead = nullptr;
  bool dirty = false;            /* work queued on `stream` nobody waited for */
  /* captured launch sequences of batches this context leads, by frame count
   * (valid while batch_dev stays where it is) */
  bool graphs = true;
  bool graph_sets = true; /* graphs keyed by the set of kernels with work too */
  Tuning tune;                   /* launch shapes of the batches it leads */
  std::vector<std::pair<uint64_t, hipGraphExec_t>> graph_exec; /* (kernel set << 8 | frames, graph) */
  bool stage_hit[AV1H_STAGES] = {};
  double stage_ms[AV1H_STAGES] = {};
  int64_t stage_n[AV1H_STAGES] = {};
  int64_t stage_bytes[AV1H_STAGES] = {};
  PackEntry ring[kPackRing];     /* streamed NV12 packs                   */
  /* pre-planning (any thread): stagers, upload streams and device buffers
   * for reuse */
  std::mutex plan_mu;
  size_t stage_hint = 0; /* the largest plan staged so far */
  std::vector<Stager*> stagers;
  std::vector<hipStream_t> uploads; /* copy-only streams, one per planning thread at a time */
  std::vector<std::pair<size_t, void*>> plan_bufs;
  size_t plan_buf_bytes = 0;
};

int set_err(std::string* e, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  *e = buf;
  return code;
}

int set_err(Ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}

#define HIP_OK(c, call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_err((c), AV1H_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_));         \
  } while (0)

int tx_row_shift(int w, int h) {
  /* Transform_Row_Shift (reference src/core/tables.cpp:357-369) by
   * [log2(w) - 2][log2(h) - 2]; -1: not an AV1 transform size */
  static constexpr int8_t t[5][5] = {{0, 0, 1, -1, -1}, {0, 1, 1, 2, -1}, {1, 1, 2, 1, 2},
                                     {-1, 2, 1, 2, 1},  {-1, -1, 2, 1, 2}};
  if (w < 4 || w > 64 || h < 4 || h > 64 || (w & (w - 1)) || (h & (h - 1))) return -1;
  return t[__builtin_ctz((unsigned)w) - 2][__builtin_ctz((unsigned)h) - 2];
}

bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

int kind_ok(int kind, int n) { return kind == 0 ? n <= 64 : (kind == 1 ? n <= 16 : n <= 32); }
int row_kind(int t) { return (0x66295450u >> (2 * t)) & 3; }
int col_kind(int t) { return (0x99895144u >> (2 * t)) & 3; }

const int kModeAngle[13] = {0, 90, 180, 45, 135, 113, 157, 203, 67, 0, 0, 0, 0};

/* intra edge filter strength and upsampling choice (av1gpu_kernels.h:
 * 1452-1481; the specification's 7.11.2.9 and 7.11.2.10) */
int edge_strength(int w, int h, int ft, int delta) {
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
int upsample_sel(int w, int h, int ft, int delta) {
  const int d = delta < 0 ? -delta : delta;
  if (d <= 0 || d >= 40) return 0;
  return ft == 0 ? (w + h <= 16) : (w + h <= 8);
}

/* DevBlock::dir_a / dir_b of a directional block with prediction angle pa:
 * the edge preparation av1gpu_kernels.h:1452-1481 performs for it (the
 * corner filter, both edges' filter strengths and upsampling) and the
 * derivative steps of 1245-1313; the block's flags and filter_type are set */
void directional_params(DevBlock* d, int pa) {
  d->dir_a = pa;
  const int w = d->w, h = d->h, ft = d->filter_type;
  if (ft >= 0 && pa != 90 && pa != 180) {
    d->dir_a |= kDirFilter;
    if (pa > 90 && pa < 180 && w + h >= 24) d->dir_a |= kDirCorner;
    if (d->flags & kHaveAbove) d->dir_a |= edge_strength(w, h, ft, pa - 90) << kDirStrA;
    if (d->flags & kHaveLeft) d->dir_a |= edge_strength(w, h, ft, pa - 180) << kDirStrL;
    if (upsample_sel(w, h, ft, pa - 90)) d->dir_a |= kDirUpA;
    if (upsample_sel(w, h, ft, pa - 180)) d->dir_a |= kDirUpL;
  }
  const int* dr = av1t_spec_table + AV1T_OFF_DR_DERIV;
  int dx = 0, dy = 0;
  if (pa < 90)
    dx = dr[pa];
  else if (pa > 90 && pa < 180)
    dx = dr[180 - pa];
  if (pa > 90 && pa < 180)
    dy = dr[pa - 90];
  else if (pa > 180)
    dy = dr[270 - pa];
  d->dir_b = dx | (dy << 16);
}

/* scaling_lookup_initialization (reference av1gpu_kernels.h:3157-3188) */
void scaling_lut(const av1h_grain& p, int plane, int* lut) {
  const int *val, *scl;
  int n;
  if (plane == 0 || p.chroma_scaling_from_luma) {
    n = p.num_y_points;
    val = p.point_y_value;
    scl = p.point_y_scaling;
  } else if (plane == 1) {
    n = p.num_cb_points;
    val = p.point_cb_value;
    scl = p.point_cb_scaling;
  } else {
    n = p.num_cr_points;
    val = p.point_cr_value;
    scl = p.point_cr_scaling;
  }
  if (n <= 0) {
    std::memset(lut, 0, 256 * sizeof(int));
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

/* wedge master masks and the codebook entry of a block (tables.cpp:616-635;
 * the masks of av1gpu_kernels.h:1968-2014) */
int wedge_1d(int which, int j) {
  static const int even[8] = {1, 4, 11, 27, 46, 58, 62, 63};
  static const int odd[8] = {1, 2, 6, 18, 37, 53, 60, 63};
  static const int vert[6] = {2, 7, 21, 43, 57, 62};
  if (which == 2) return j < 29 ? 0 : (j >= 35 ? 64 : vert[j - 29]);
  if (j < 28) return 0;
  if (j >= 36) return 64;
  return which == 0 ? even[j - 28] : odd[j - 28];
}
int wedge_ob63(int i, int j) {
  const int shift = (i & 1) ? 15 - (i >> 1) : 16 - (i >> 1);
  return wedge_1d(i & 1, std::min(std::max(j - shift, 0), 63));
}
int wedge_master(int dir, int i, int j) {
  switch (dir) {
    case 0: return wedge_1d(2, i);
    case 1: return wedge_1d(2, j);
    case 2: return wedge_ob63(j, i);
    case 4: return 64 - wedge_ob63(i, 63 - j);
    case 5: return 64 - wedge_ob63(j, 63 - i);
    default: return wedge_ob63(i, j);
  }
}
bool wedge_entry(int bw, int bh, int index, int* dir, int* xoff, int* yoff, int* flip) {
  if (index < 0 || index >= 16 || bw < 8 || bh < 8 || bw > 32 || bh > 32) return false;
  const int shape = bh > bw ? 0 : (bh < bw ? 1 : 2);
  const int* e = av1t_spec_table + AV1T_OFF_WEDGE_CODEBOOK + (shape * 16 + index) * 3;
  *dir = e[0];
  *xoff = 32 - ((e[1] * bw) >> 3);
  *yoff = 32 - ((e[2] * bh) >> 3);
  int sum = 0;
  for (int i = 0; i < bw; ++i) sum += wedge_master(e[0], *yoff, *xoff + i);
  for (int i = 1; i < bh; ++i) sum += wedge_master(e[0], *yoff + i, *xoff);
  const int n = bw + bh - 1;
  *flip = ((sum + n / 2) / n) < 32 ? 1 : 0;
  return true;
}

/* warp shear parameters and validity (tables.cpp:444-475) */
int64_t round2s64(int64_t v, int n) {
  if (n <= 0) return v;
  const int64_t h = (int64_t)1 << (n - 1);
  return v >= 0 ? (v + h) >> n : -((-v + h) >> n);
}
int round2s(int v, int n) { return v >= 0 ? (v + (1 << (n - 1))) >> n : -((-v + (1 << (n - 1))) >> n); }
bool setup_shear(const int32_t* p, int* alpha, int* beta, int* gamma, int* delta) {
  const int* lut = av1t_spec_table + AV1T_OFF_DIV_LUT;
  const int a0 = std::min(32767, std::max(-32768, p[2] - (1 << 16)));
  const int b0 = std::min(32767, std::max(-32768, p[3]));
  const int ad = p[2] < 0 ? -p[2] : p[2];
  int n = 0;
  for (int v = ad; v > 1; v >>= 1) ++n;
  const int e = ad - (1 << n);
  const int f = n > 8 ? (e + (1 << (n - 8 - 1))) >> (n - 8) : e << (8 - n);
  const int dshift = n + 14;
  const int dfac = p[2] < 0 ? -lut[f] : lut[f];
  const int64_t v = (int64_t)p[4] * ((int64_t)1 << 16);
  const int g0 = (int)std::min<int64_t>(32767, std::max<int64_t>(-32768, round2s64(v * dfac, dshift)));
  const int64_t w = (int64_t)p[3] * p[4];
  const int d0 = (int)std::min<int64_t>(32767, std::max<int64_t>(-32768, p[5] - round2s64(w * dfac, dshift) - (1 << 16)));
  *alpha = round2s(a0, 6) * 64;
  *beta = round2s(b0, 6) * 64;
  *gamma = round2s(g0, 6) * 64;
  *delta = round2s(d0, 6) * 64;
  if (4 * std::abs(*alpha) + 7 * std::abs(*beta) >= (1 << 16)) return false;
  if (4 * std::abs(*gamma) + 4 * std::abs(*delta) >= (1 << 16)) return false;
  return true;
}

/* motion compensation positions on the host: the arithmetic of
 * av1gpu_kernels.h:1761-1763 (unscaled) and 1819-1840 (scaled, 7.11.3.3),
 * which the kernel then evaluates from these fields */
int rnd2s_add(int v, int n) { return (v + (1 << (n - 1)) + (v < 0 ? -1 : 0)) >> n; }
void mc_position(int bx, int by, int mvx, int mvy, int sub_x, int sub_y, int scx, int scy, McRefDev* q) {
  q->scaled = !(scx == (1 << 14) && scy == (1 << 14));
  const int mx = mvx * 2, my = mvy * 2;
  q->ix = mx >> (4 + sub_x);
  q->fx = (mx >> sub_x) & 15;
  q->iy = my >> (4 + sub_y);
  q->fy = (my >> sub_y) & 15;
  const int ox = (bx << 4) + ((2 * mvx) >> sub_x) + 8;
  const int oy = (by << 4) + ((2 * mvy) >> sub_y) + 8;
  q->sx0 = rnd2s_add(ox * scx - (8 << 14), 8) + 32;
  q->sy0 = rnd2s_add(oy * scy - (8 << 14), 8) + 32;
  q->stx = rnd2s_add(scx, 4);
  q->sty = rnd2s_add(scy, 4);
}
int mc_col(int bx, const McRefDev& q, int x) { return q.scaled ? (q.sx0 + q.stx * x) >> 10 : bx + x + q.ix; }
int mc_row(int by, const McRefDev& q, int y) { return q.scaled ? (q.sy0 + q.sty * y) >> 10 : by + y + q.iy; }
/* the 4 tap rows for extents of 4 or less (av1gpu_kernels.h:1696-1701) */
int mc_filter_row(int cls, int extent) {
  if (extent > 4) return cls;
  if (cls == 0 || cls == 2) return 4;
  if (cls == 1) return 5;
  return cls;
}

/* ---- flattening ----------------------------------------------------------- */

struct Host;
/* the arrays of a frame plan's device buffer */
enum {
  kPartBlocks, kPartTx, kPartInter, kPartQueue, kPartSpecial, kPartObmc, kPartCoef, kPartResidual, kPartPalette,
  kPartDeblock, kPartCdef, kPartLr, kPartMcItems, kPartMcTasks, kPartItx, kPartItxOffs, kPartDbChunk, kPartCdefDir,
  kPartLrTiles, kPartGrain, kPartLevels, kPartQm, kPartDq, kParts
};
void sb_begin(const Layout& L, Plan* plan, Host* H);
void sb_block(const Layout& L, const DevBlock& d, const DevTx* tx, Plan* plan, Host* H);
void sb_queued(const Layout& L, const DevBlock& d, int qi, Plan* plan, Host* H);
void sb_finish(const Host& H, const av1h_frame_header& hdr, const Layout& L, Plan* plan);

/* An allocator whose value-initialising construct (emplace_back() with no
 * arguments) leaves a trivially constructible element uninitialised: the
 * planner writes each block descriptor's fields itself, and only the ones
 * its readers use. */
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) {}
  template <typename U>
  void construct(U* p) {
    static_assert(std::is_trivially_default_constructible<U>::value, "default-initialised element");
    ::new (static_cast<void*>(p)) U;
  }
  template <typename U, typename... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};
using BlockVec = std::vector<DevBlock, NoInitAlloc<DevBlock>>;

struct Host {
  std::vector<DevDq> dq; /* a frame with coded levels: one per transform */
  BlockVec blocks;
  BlockVec dev_blocks; /* the ones the device reads */
  bool dev_all = false;             /* ... which are all of them: `blocks` goes */
  std::vector<DevTx> tx;
  std::vector<int32_t> inter_list, queue_list, special_list;
  std::vector<DevDeblock> s_deblock_h; /* horizontal deblocking edges, before they are appended */
  std::vector<DevObmc> obmc;
  std::vector<int32_t> itx_all;
  std::vector<McItemDev> mc_items;
  std::vector<McTaskDev> mc_task;
  std::vector<int32_t> db_chunk, cdef_dir_list;
  std::vector<int32_t> lr_tiles; /* pairs */
  std::vector<DevDeblock> deblock;
  std::vector<DevCdef> cdef;
  std::vector<DevLr> lr;
  DevGrain grain;
  int32_t itx_off[2][kTxClasses + 1];
  /* the device buffer's arrays: source, bytes, zero bytes after them, offset */
  struct Part {
    const void* src;
    size_t bytes, fill, off;
  };
  std::vector<Part> parts;
  int part_of[kParts];
  size_t total = 0;
  /* scratch of one flattening: queued block per cell, queued cells, device
   * index of each block, luma 8x8 blocks whose CDEF direction is needed */
  std::vector<int32_t> s_owner[3];
  std::vector<char> s_queued[3];
  std::vector<int32_t> s_dev_of;
  std::vector<int32_t> s_sb_inter, s_sb_sub; /* inter blocks, those with subpel motion, per superblock */
  std::vector<int32_t> s_sb_q;               /* superblock of each queued block              */
  std::vector<uint8_t> s_need;
};

/* BlockDecoded availability for blocks whose syntax asks to derive it
 * (reference engine.cpp:779-832): every in-grid cell of the rectangle belongs
 * to a block earlier in coding order. */
struct Owner {
  const Layout* L;
  std::vector<int> grid[3];
  int shift[3], cw[3], ch[3];
  void build(const Layout& lay, const av1h_frame_desc& f) {
    L = &lay;
    for (int p = 0; p < 3; ++p) {
      shift[p] = lay.cell_shift[p];
      cw[p] = lay.cells_w[p];
      ch[p] = lay.cells_h[p];
      grid[p].assign((size_t)cw[p] * ch[p], -1);
    }
    for (int64_t i = 0; i < f.num_blocks; ++i) {
      const av1h_block& b = f.blocks[i];
      const int p = b.plane, oc = 1 << shift[p];
      const int cx1 = std::min((b.x + b.w + oc - 1) / oc, cw[p]);
      const int cy1 = std::min((b.y + b.h + oc - 1) / oc, ch[p]);
      for (int cy = std::max(b.y, 0) / oc; cy < cy1; ++cy)
        for (int cx = std::max(b.x, 0) / oc; cx < cx1; ++cx) grid[p][(size_t)cy * cw[p] + cx] = (int)i;
    }
  }
  int at(int p, int x, int y) const {
    if (x < 0 || y < 0 || x >= L->grid_w[p] || y >= L->grid_h[p]) return -1;
    return grid[p][(size_t)(y >> shift[p]) * cw[p] + (x >> shift[p])];
  }
  bool decoded(int p, int x0, int x1, int y0, int y1, int bi) const {
    const int oc = 1 << shift[p];
    for (int y = y0 & ~(oc - 1); y <= y1; y += oc)
      for (int x = x0 & ~(oc - 1); x <= x1; x += oc) {
        const int o = at(p, x, y);
        if (o >= 0 && o >= bi) return false;
      }
    return true;
  }
};

/* The superblock features the cost model reads (their list is in the
 * header, av1h_plan_sb_features) and the superblock of every queue entry:
 * sb_begin before the blocks, sb_block for each block as it is flattened
 * (its descriptor and transforms still in cache), sb_finish once the queue
 * is known. */
int sb_of(const Layout& L, int sw, int sh, int plane, int x, int y) {
  const int sx = plane ? L.ss_x : 0, sy = plane ? L.ss_y : 0;
  const int bx = std::min(std::max((x << sx) >> 6, 0), sw - 1), by = std::min(std::max((y << sy) >> 6, 0), sh - 1);
  return by * sw + bx;
}

void sb_begin(const Layout& L, Plan* plan, Host* H) {
  const int sw = (L.grid_w[0] + 63) >> 6, sh = (L.grid_h[0] + 63) >> 6;
  plan->sb_w = sw;
  plan->sb_h = sh;
  const size_t nsb = (size_t)sw * sh;
  plan->sb_feat.assign(nsb * AV1H_CM_FEATURES, 0.f);
  H->s_sb_inter.assign(nsb, 0);
  H->s_sb_sub.assign(nsb, 0);
}

void sb_block(const Layout& L, const DevBlock& d, const DevTx* tx, Plan* plan, Host* H) {
  const int sb = sb_of(L, plan->sb_w, plan->sb_h, d.plane, d.x, d.y);
  float* f = &plan->sb_feat[(size_t)sb * AV1H_CM_FEATURES];
  const float area = (float)d.w * d.h / 4096.f;
  f[0] += 1.f;
  for (int t = 0; t < d.tx_count; ++t) {
    const int side = std::max(tx[t].w, tx[t].h);
    const int cls = side <= 4 ? 0 : side <= 8 ? 1 : side <= 16 ? 2 : side <= 32 ? 3 : 4;
    f[1 + cls] += 1.f;
    f[6] += (float)tx[t].eob / 64.f;
    f[16] = std::max(f[16], (float)side / 64.f);
  }
  if (d.kind == kIntra) {
    f[7] += 1.f;
    if (d.mode >= AV1H_PRED_V && d.mode <= AV1H_PRED_D67) f[11] += 1.f;
    if (d.mode >= AV1H_PRED_SMOOTH && d.mode <= AV1H_PRED_PAETH) f[12] += 1.f;
    if (d.mode == AV1H_PRED_CFL || (d.flags & (kPalette | kFilterIntra))) f[13] += 1.f;
  } else {
    f[8] += 1.f;
    f[15] += area;
    ++H->s_sb_inter[(size_t)sb];
    if (d.compound) f[9] += 1.f;
    bool sub = false;
    for (int r = 0; r < (d.compound ? 2 : 1); ++r) sub = sub || (d.mv[r][0] & 7) || (d.mv[r][1] & 7);
    if (sub) ++H->s_sb_sub[(size_t)sb];
    if (d.flags & (kInterIntra | kWarp0 | kWarp1 | kMaskWarp0 | kMaskWarp1) || d.obmc_count > 0 || d.mask_kind)
      f[20] += 1.f;
  }
  if (d.flags & (kLossless | kResidualRaw)) f[17] += 1.f;
  if (d.flags & kIntraBC) f[18] += 1.f;
  if (d.plane) f[19] += 1.f;
}

/* a queued block (queue graph order qi) */
void sb_queued(const Layout& L, const DevBlock& d, int qi, Plan* plan, Host* H) {
  const int sb = sb_of(L, plan->sb_w, plan->sb_h, d.plane, d.x, d.y);
  H->s_sb_q[(size_t)qi] = sb;
  float* f = &plan->sb_feat[(size_t)sb * AV1H_CM_FEATURES];
  f[14] += (float)d.w * d.h / 4096.f;
  f[22] += 1.f;
}

void sb_finish(const Host& H, const av1h_frame_header& hdr, const Layout& L, Plan* plan) {
  const int sw = plan->sb_w, sh = plan->sb_h, nsb = sw * sh;
  std::vector<float>& F = plan->sb_feat;
  /* queue entry k is queued block queue_qi[k] */
  plan->queue_sb.resize(plan->queue_host.size());
  for (size_t k = 0; k < plan->queue_host.size(); ++k) plan->queue_sb[k] = H.s_sb_q[(size_t)plan->queue_qi[k]];
  plan->queue_sb_static = plan->queue_sb;
  for (int sb = 0; sb < nsb; ++sb) {
    float* f = &F[(size_t)sb * AV1H_CM_FEATURES];
    const int ni = H.s_sb_inter[(size_t)sb];
    f[10] = ni ? (float)H.s_sb_sub[(size_t)sb] / (float)ni : 0.f;
    f[21] = hdr.enable_intra_edge_filter ? 1.f : 0.f;
    f[23] = hdr.bit_depth > 8 ? 1.f : 0.f;
  }
}

int resolve_slot(const av1h_frame_header& h, int ref) {
  if (ref < 0 || ref >= h.num_reference_slots) return -1;
  return h.reference_slots[ref];
}

/* Whether intra block d can take its edges without a cell wait: every
 * sample of its above row (x .. the last it depends on) is either the last
 * row of a queued block decoded before it (read as a packed word) or
 * belongs to a block outside the queue (final before the queue starts,
 * read from the frame); the same for its left column with last columns,
 * and its corner either way. Then the queued owners are flagged to publish
 * those rows and columns, and d gets its per-cell sources (edge_mask_a /
 * edge_mask_l, bit i for the i-th cell of the span: 1 packed, 0 frame) and
 * its corner's. `owner` maps the plane's cells to the queued blocks decoded
 * before d (index into queue_list), -1 elsewhere; `queued` flags the cells
 * of every queued block of the plane. */
bool packed_edges(const Layout& L, DevBlock& d, const std::vector<int32_t>& owner, const std::vector<char>& queued,
                  BlockVec& blocks, const std::vector<int32_t>& queue_list) {
  if (d.kind != kIntra || (d.flags & (kPalette | kIntraBC)) || (d.mode == AV1H_PRED_CFL && d.plane > 0)) return false;
  const int p = d.plane, sh = L.cell_shift[p], cell = 1 << sh;
  const bool ha = (d.flags & kHaveAbove) != 0, hl = (d.flags & kHaveLeft) != 0;
  const int alast = std::min(L.grid_w[p] - 1, d.x + ((d.flags & kHaveAR) ? 2 * d.w : d.w) - 1);
  const int llast = std::min(L.grid_h[p] - 1, d.y + ((d.flags & kHaveBL) ? 2 * d.h : d.h) - 1);
  const int ae = std::min(d.dep_x1, alast), le = std::min(d.dep_y1, llast);
  if ((ha && ae < d.x) || (hl && le < d.y)) return false;
  if ((ha && ((ae - d.x) >> sh) >= 32) || (hl && ((le - d.y) >> sh) >= 32)) return false;
  /* 1: a queued block decoded before d, 0: outside the queue, -1: neither */
  auto src = [&](int px, int py, DevBlock** o) -> int {
    const size_t c = (size_t)(py >> sh) * L.cells_w[p] + (px >> sh);
    *o = nullptr;
    if (owner[c] >= 0) {
      *o = &blocks[(size_t)queue_list[(size_t)owner[c]]];
      return 1;
    }
    return queued[c] ? -1 : 0;
  };
  uint32_t ma = 0, ml = 0;
  DevBlock* o;
  if (ha)
    for (int px = d.x; px <= ae; px += cell) {
      const int k = src(px, d.y - 1, &o);
      if (k < 0 || (k && o->y + o->h != d.y)) return false;
      if (k) ma |= 1u << ((px - d.x) >> sh);
    }
  if (hl)
    for (int py = d.y; py <= le; py += cell) {
      const int k = src(d.x - 1, py, &o);
      if (k < 0 || (k && o->x + o->w != d.x)) return false;
      if (k) ml |= 1u << ((py - d.y) >> sh);
    }
  int corner = 0; /* 1: packed row, 2: packed column, 3: frame */
  if (ha && hl) {
    const int k = src(d.x - 1, d.y - 1, &o);
    if (k < 0) return false;
    corner = !k ? 3 : o->y + o->h == d.y ? 1 : (o->x + o->w == d.x ? 2 : 0);
    if (!corner) return false;
  }
  if (ha)
    for (int px = d.x; px <= ae; px += cell)
      if (src(px, d.y - 1, &o) > 0) o->flags |= kEdgeRow;
  if (hl)
    for (int py = d.y; py <= le; py += cell)
      if (src(d.x - 1, py, &o) > 0) o->flags |= kEdgeCol;
  if (corner == 1 || corner == 2) {
    src(d.x - 1, d.y - 1, &o);
    o->flags |= corner == 1 ? kEdgeRow : kEdgeCol;
  }
  if (corner == 1) d.flags |= kCornerRow;
  if (corner == 3) d.flags |= kCornerFrame;
  d.edge_mask_a = (int32_t)ma;
  d.edge_mask_l = (int32_t)ml;
  return true;
}

int flatten(const av1h_frame_desc& f, Plan* plan, Host* H, std::string* err) {
  const av1h_frame_header& hdr = f.hdr;
  if (hdr.width <= 0 || hdr.height <= 0 || hdr.width > 16384 || hdr.height > 16384)
    return set_err(err, AV1H_ERR_INVALID, "frame has no size");
  if (hdr.bit_depth != 8 && hdr.bit_depth != 10 && hdr.bit_depth != 12)
    return set_err(err, AV1H_ERR_INVALID, "bit depth %d", hdr.bit_depth);
  if (hdr.ss_x < 0 || hdr.ss_x > 1 || hdr.ss_y < 0 || hdr.ss_y > 1 || (hdr.ss_y && !hdr.ss_x))
    return set_err(err, AV1H_ERR_UNSUPPORTED, "chroma subsampling %d,%d", hdr.ss_x, hdr.ss_y);
  if (hdr.num_reference_slots < 0 || hdr.num_reference_slots > AV1H_MAX_REFS)
    return set_err(err, AV1H_ERR_INVALID, "bad reference slot count");
  for (int i = 0; i < 7; ++i) {
    const int rw = hdr.reference_width[i] > 0 ? hdr.reference_width[i] : hdr.width;
    const int rh = hdr.reference_height[i] > 0 ? hdr.reference_height[i] : hdr.height;
    if (rw != hdr.width || rh != hdr.height)
      return set_err(err, AV1H_ERR_UNSUPPORTED,
                     "reference scaling between pictures of different sizes is not supported");
  }
  Layout& L = plan->lay;
  make_layout(&L, hdr.width, hdr.height, hdr.bit_depth, hdr.ss_x, hdr.ss_y,
              hdr.coded_width > 0 ? hdr.coded_width : hdr.width);
  plan->hdr = hdr;
  plan->upscaled = L.out_w[0] != L.plane_w[0];
  const int fw = L.plane_w[0], fh = L.plane_h[0];
  int scale_x = 1 << 14, scale_y = 1 << 14;
  if (plan->upscaled) {
    scale_x = ((hdr.width << 14) + fw / 2) / fw;
    scale_y = ((hdr.height << 14) + fh / 2) / fh;
    if (scale_x > 2 << 14 || scale_y > 2 << 14)
      return set_err(err, AV1H_ERR_UNSUPPORTED, "superres beyond 2:1");
  }
  const int nb = (int)f.num_blocks;
  /* coded levels: the device dequantises them; the raster pool is not read */
  const bool levels = f.num_levels > 0;
  if (levels && !f.levels) return set_err(err, AV1H_ERR_INVALID, "levels without an array");
  H->dq.clear();
  if (levels) H->dq.reserve((size_t)f.num_transforms);
  H->blocks.clear();
  H->tx.clear();
  H->inter_list.clear();
  H->queue_list.clear();
  H->special_list.clear();
  H->obmc.clear();
  H->blocks.reserve(nb);
  H->tx.reserve((size_t)f.num_transforms);
  plan->slots.clear();
  bool derive = false;
  for (int i = 0; i < nb; ++i)
    if (f.blocks[i].is_intra && (f.blocks[i].have_above_right < 0 || f.blocks[i].have_below_left < 0))
      derive = true;
  Owner own;
  if (derive) own.build(L, f);
  sb_begin(L, plan, H);
  size_t cend = 0, rend = 0;
  for (int i = 0; i < nb; ++i) {
    const av1h_block& b = f.blocks[i];
    const int p = b.plane;
    if (p < 0 || p > 2) return set_err(err, AV1H_ERR_INVALID, "block %d: plane %d", i, p);
    if (!pow2_in(b.w, 4, 64) || !pow2_in(b.h, 4, 64) || b.x < 0 || b.y < 0 || (b.x & 3) || (b.y & 3) ||
        b.x + b.w > L.grid_w[p] || b.y + b.h > L.grid_h[p])
      return set_err(err, AV1H_ERR_UNSUPPORTED, "block %d: %dx%d at %d,%d outside the coded grid", i, b.w,
                     b.h, b.x, b.y);
    if (b.bit_depth != hdr.bit_depth)
      return set_err(err, AV1H_ERR_UNSUPPORTED, "block %d: bit depth differs from the frame", i);
    if (b.tx_first < 0 || b.tx_count < 0 || b.tx_first + (int64_t)b.tx_count > f.num_transforms)
      return set_err(err, AV1H_ERR_INVALID, "block %d: transform range", i);
    /* built in place (room for every block is reserved). A plain inter
     * block (its prediction done by the motion compensation tasks, which read
     * its descriptor up to the mask kind) has only that head cleared; every
     * other block's descriptor reaches the device and is cleared whole. */
    DevBlock& d = H->blocks.emplace_back();
    const bool plain_inter = !b.is_intra && !b.intrabc && !b.interintra && b.compound != AV1H_COMPOUND_WEDGE &&
                             b.compound != AV1H_COMPOUND_DIFFWTD && b.obmc_count <= 0 && !b.warp[0] && !b.warp[1] &&
                             !b.mask_warp[0] && !b.mask_warp[1];
    static_assert(offsetof(DevBlock, mask_kind) + sizeof(int32_t) <= 128, "the head of a plain inter block");
    std::memset(&d, 0, plain_inter ? 128 : sizeof(d));
    d.x = b.x;
    d.y = b.y;
    d.w = b.w;
    d.h = b.h;
    d.plane = p;
    d.slot[0] = d.slot[1] = -1;
    if (b.have_left && b.x > 0) d.flags |= kHaveLeft;
    if (b.have_above && b.y > 0) d.flags |= kHaveAbove;
    if (b.lossless) d.flags |= kLossless;
    if (b.residual_provided) d.flags |= kResidualRaw;
    if (b.lossless && !b.residual_provided)
      for (int t = 0; t < b.tx_count; ++t) {
        const av1h_transform& tx = f.transforms[b.tx_first + t];
        if (tx.w != 4 || tx.h != 4) return set_err(err, AV1H_ERR_UNSUPPORTED, "lossless transform %dx%d", tx.w, tx.h);
      }
    if (b.is_intra) {
      d.kind = kIntra;
      d.mode = b.mode;
      if (b.mode < 0 || b.mode > 13) return set_err(err, AV1H_ERR_INVALID, "block %d: intra mode %d", i, b.mode);
      if (b.mode == AV1H_PRED_CFL && p == 0)
        return set_err(err, AV1H_ERR_UNSUPPORTED, "chroma from luma on the luma plane");
      if (b.mode == AV1H_PRED_CFL && (b.w > 32 || b.h > 32))
        return set_err(err, AV1H_ERR_INVALID, "block %d: chroma from luma on a %dx%d block", i, b.w, b.h);
      d.angle = b.angle_delta;
      d.filter_type = hdr.enable_intra_edge_filter ? b.filter_type : -1;
      d.fi_mode = b.filter_intra_mode;
      d.cfl_alpha = b.cfl_alpha;
      const int ar = b.have_above_right >= 0
                         ? b.have_above_right != 0
                         : (b.y > 0 && own.decoded(p, b.x + b.w, b.x + 2 * b.w - 1, b.y - 1, b.y - 1, i));
      const int bl = b.have_below_left >= 0
                         ? b.have_below_left != 0
                         : (b.x > 0 && own.decoded(p, b.x - 1, b.x - 1, b.y + b.h, b.y + 2 * b.h - 1, i));
      if (ar) d.flags |= kHaveAR;
      if (bl) d.flags |= kHaveBL;
      if (b.use_filter_intra && p == 0) {
        d.flags |= kFilterIntra;
        if (b.filter_intra_mode < 0 || b.filter_intra_mode > 4)
          return set_err(err, AV1H_ERR_INVALID, "filter intra mode %d", b.filter_intra_mode);
      }
      /* the read set the reference scheduler waits for (engine.cpp:811-875) */
      int ly1 = b.y + b.h - 1, ax1 = b.x + b.w - 1;
      if (b.use_palette) {
        d.flags |= kPalette;
        if (b.palette_off < 0 || b.palette_stride < b.w ||
            (int64_t)b.palette_off + (int64_t)(b.h - 1) * b.palette_stride + b.w > f.num_palette)
          return set_err(err, AV1H_ERR_INVALID, "block %d: palette map out of bounds", i);
        for (int k = 0; k < 8; ++k) d.palette[k] = b.palette[k];
        d.palette_off = b.palette_off;
        d.palette_stride = b.palette_stride;
        ax1 = b.x - 1;
        ly1 = b.y - 1;
      } else if (b.mode >= 1 && b.mode <= 8) {
        const int pa = kModeAngle[b.mode] + 3 * b.angle_delta;
        if (pa < 90 && ar) ax1 = b.x + 2 * b.w - 1;
        if (pa > 180 && bl) ly1 = b.y + 2 * b.h - 1;
        directional_params(&d, pa);
      }
      d.dep_x1 = ax1;
      d.dep_y1 = ly1;
      d.dep_corner = (ax1 >= b.x || ly1 >= b.y) ? 1 : 0;
      d.dep_rx1 = d.dep_rx0 - 1;
      H->queue_list.push_back(i);
    } else {
      d.kind = kInter;
      if (b.compound < 0 || b.compound > 4) return set_err(err, AV1H_ERR_INVALID, "compound %d", b.compound);
      const int nref = b.compound ? 2 : 1;
      const int sx = p > 0 ? hdr.ss_x : 0, sy = p > 0 ? hdr.ss_y : 0;
      for (int r = 0; r < nref; ++r) {
        const int s = b.intrabc ? -1 : resolve_slot(hdr, b.ref[r]);
        if (s < 0 && !b.intrabc)
          return set_err(err, AV1H_ERR_UNSUPPORTED, "block %d: inter prediction without a reference", i);
        if (s >= kSlots) return set_err(err, AV1H_ERR_INVALID, "block %d: reference slot %d", i, s);
        d.slot[r] = s;
        if (s >= 0 && std::find(plan->slots.begin(), plan->slots.end(), s) == plan->slots.end())
          plan->slots.push_back(s);
      }
      for (int r = 0; r < 2; ++r) {
        d.mv[r][0] = b.mv[r][0];
        d.mv[r][1] = b.mv[r][1];
      }
      if (b.filter < 0 || b.filter > 3 || b.filter_v < 0 || b.filter_v > 3)
        return set_err(err, AV1H_ERR_INVALID, "block %d: filter", i);
      d.filt = b.filter;
      d.filt_v = b.filter_v;
      d.compound = b.compound;
      d.w0 = b.blend[0];
      d.w1 = b.blend[1];
      d.ref_w = L.out_w[p];
      d.ref_h = L.out_h[p];
      d.scale_x = scale_x;
      d.scale_y = scale_y;
      bool special = false;
      if (b.intrabc) {
        /* the frame being decoded, 4-rounded coded size, unscaled
         * (engine.cpp job fields; av1gpu_kernels.h intrabc) */
        if (b.compound) return set_err(err, AV1H_ERR_INVALID, "block %d: compound intra block copy", i);
        d.flags |= kIntraBC;
        d.ref_w = ((fw + 3) & ~3) >> sx;
        d.ref_h = ((fh + 3) & ~3) >> sy;
        d.scale_x = d.scale_y = 1 << 14;
        const int msx = 3 + sx, msy = 3 + sy;
        const int x0 = b.x + (b.mv[0][0] >> msx), y0 = b.y + (b.mv[0][1] >> msy);
        const int x1 = std::min(x0 + b.w, L.grid_w[p]) - 1, y1 = std::min(y0 + b.h, L.grid_h[p]) - 1;
        const int sh = L.cell_shift[p];
        d.dep_rx0 = std::max(x0, 0) >> sh;
        d.dep_ry0 = std::max(y0, 0) >> sh;
        d.dep_rx1 = x1 < 0 ? -1 : x1 >> sh;
        d.dep_ry1 = y1 < 0 ? -1 : y1 >> sh;
        if (x1 < std::max(x0, 0) || y1 < std::max(y0, 0)) d.dep_rx1 = d.dep_rx0 - 1; /* nothing to wait for */
      }
      if (b.compound == AV1H_COMPOUND_WEDGE || b.compound == AV1H_COMPOUND_DIFFWTD) {
        d.mask_kind = b.compound == AV1H_COMPOUND_WEDGE ? 1 : 2;
        d.mask_sign = b.mask_sign;
        if (d.mask_kind == 1 &&
            !wedge_entry(b.w << sx, b.h << sy, b.wedge_index, &d.wedge_dir, &d.wedge_xoff, &d.wedge_yoff,
                         &d.wedge_flip))
          return set_err(err, AV1H_ERR_INVALID, "a wedge on a block size that has none");
        special = true;
      }
      if (b.interintra) {
        d.flags |= kInterIntra;
        d.ii_mode = b.interintra_mode;
        if (b.interintra_mode < 0 || b.interintra_mode > 3)
          return set_err(err, AV1H_ERR_INVALID, "block %d: inter/intra mode", i);
        if (b.interintra_wedge) {
          if (!wedge_entry(b.w << sx, b.h << sy, b.wedge_index, &d.wedge_dir, &d.wedge_xoff, &d.wedge_yoff,
                           &d.wedge_flip))
            return set_err(err, AV1H_ERR_INVALID, "an inter/intra wedge on a block size that has none");
          d.flags |= kIIWedge;
        }
      }
      if (b.obmc_count > 0) {
        if (b.obmc_first < 0 || b.obmc_first + (int64_t)b.obmc_count > f.num_obmc)
          return set_err(err, AV1H_ERR_INVALID, "block %d: OBMC range", i);
        d.obmc_first = (int)H->obmc.size();
        for (int k = 0; k < b.obmc_count && k < 8; ++k) {
          const av1h_obmc& o = f.obmc[b.obmc_first + k];
          const int s = resolve_slot(hdr, o.ref);
          if (s < 0 || s >= kSlots) return set_err(err, AV1H_ERR_UNSUPPORTED, "block %d: OBMC without a reference", i);
          if (o.w != 2 && o.w != 4 && o.w != 8 && o.w != 16 && o.w != 32 && o.pass != 0)
            return set_err(err, AV1H_ERR_INVALID, "block %d: OBMC width %d", i, o.w);
          if (o.h != 2 && o.h != 4 && o.h != 8 && o.h != 16 && o.h != 32 && o.pass == 0)
            return set_err(err, AV1H_ERR_INVALID, "block %d: OBMC height %d", i, o.h);
          if (o.filter < 0 || o.filter > 3 || o.filter_v < 0 || o.filter_v > 3)
            return set_err(err, AV1H_ERR_INVALID, "block %d: OBMC filter", i);
          if (std::find(plan->slots.begin(), plan->slots.end(), s) == plan->slots.end()) plan->slots.push_back(s);
          H->obmc.push_back(DevObmc{o.pass, o.x, o.y, o.w, o.h, o.mv_x, o.mv_y, o.filter, o.filter_v, s});
        }
        d.obmc_count = (int)H->obmc.size() - d.obmc_first;
        special = true;
      }
      for (int r = 0; r < 2; ++r) {
        if (!b.warp[r] && !b.mask_warp[r]) continue;
        if (b.warp[r]) d.flags |= r ? kWarp1 : kWarp0;
        if (b.mask_warp[r]) d.flags |= r ? kMaskWarp1 : kMaskWarp0;
        special = true;
        int al, be, ga, de;
        /* parameters kept only when the shear is valid (engine.cpp job fields) */
        if (!setup_shear(b.warp_params[r], &al, &be, &ga, &de)) continue;
        for (int k = 0; k < 6; ++k) d.warp_params[r][k] = b.warp_params[r][k];
        d.warp_abcd[r][0] = al;
        d.warp_abcd[r][1] = be;
        d.warp_abcd[r][2] = ga;
        d.warp_abcd[r][3] = de;
      }
      if (b.intrabc || b.interintra) {
        /* reads the frame being decoded: the dependency queue. Inter/intra
         * waits for its left column, above row and corner (engine.cpp:
         * 811-875); intra block copy for its source rectangle */
        if (b.interintra) {
          d.flags |= (b.have_left && b.x > 0 ? kHaveLeft : 0) | (b.have_above && b.y > 0 ? kHaveAbove : 0);
          d.dep_x1 = b.x + b.w - 1;
          d.dep_y1 = b.y + b.h - 1;
          d.dep_corner = 1;
        } else {
          d.dep_x1 = b.x - 1;
          d.dep_y1 = b.y - 1;
          d.dep_corner = 0;
        }
        if (!b.intrabc) d.dep_rx1 = d.dep_rx0 - 1;
        H->queue_list.push_back(i);
      } else if (special) {
        H->special_list.push_back(i);
      } else {
        H->inter_list.push_back(i);
      }
    }
    d.tx_first = (int)H->tx.size();
    for (int t = 0; t < b.tx_count; ++t) {
      const av1h_transform& tx = f.transforms[b.tx_first + t];
      if (tx.eob <= 0) continue;
      if (!pow2_in(tx.w, 4, 64) || !pow2_in(tx.h, 4, 64) || tx.x < 0 || tx.y < 0 || tx.x + tx.w > b.w ||
          tx.y + tx.h > b.h || tx.coef_offset < 0 || tx.tx_type < 0 || tx.tx_type > 15)
        return set_err(err, AV1H_ERR_INVALID, "block %d: transform %d out of range", i, t);
      const int rs = b.lossless ? 0 : tx_row_shift(tx.w, tx.h);
      if (rs < 0) return set_err(err, AV1H_ERR_UNSUPPORTED, "transform size %dx%d", tx.w, tx.h);
      if (!b.lossless && !b.residual_provided &&
          (!kind_ok(row_kind(tx.tx_type), tx.w) || !kind_ok(col_kind(tx.tx_type), tx.h)))
        return set_err(err, AV1H_ERR_UNSUPPORTED, "transform type %d is not defined for %dx%d", tx.tx_type, tx.w, tx.h);
      DevTx dt;
      dt.x = (uint16_t)(b.x + tx.x);
      dt.y = (uint16_t)(b.y + tx.y);
      dt.w = (uint8_t)tx.w;
      dt.h = (uint8_t)tx.h;
      dt.coef = tx.coef_offset;
      dt.tx_type = (uint8_t)tx.tx_type;
      dt.row_shift = (int8_t)rs;
      dt.eob = (uint16_t)std::min(tx.eob, 65535);
      dt.plane = (uint8_t)p;
      dt.flags = (uint8_t)((b.lossless ? kTxLossless : 0) | (b.residual_provided ? kTxRaw : 0) |
                           ((b.is_intra || b.intrabc || b.interintra) ? kTxToResidual : 0));
      const size_t end = (size_t)tx.coef_offset + (size_t)tx.w * tx.h;
      if (b.residual_provided)
        rend = std::max(rend, end);
      else if (!levels)
        cend = std::max(cend, end);
      if (levels) {
        DevDq q{0, -1, 0, 0, 0};
        if (!b.residual_provided) {
          const int64_t area = (int64_t)std::min(tx.w, 32) * std::min(tx.h, 32);
          const int64_t count = std::min<int64_t>(tx.eob, area);
          const int den = tx.denominator;
          if (tx.level_offset < 0 || tx.level_offset + count > f.num_levels)
            return set_err(err, AV1H_ERR_INVALID, "block %d: transform %d levels out of range", i, t);
          if (tx.dc_step < 1 || tx.dc_step > 65535 || tx.ac_step < 1 || tx.ac_step > 65535 ||
              (den != 1 && den != 2 && den != 4))
            return set_err(err, AV1H_ERR_INVALID, "block %d: transform %d dequantiser", i, t);
          if (av1scan::tx_class(tx.tx_type) != 0 && (tx.w > 16 || tx.h > 16))
            return set_err(err, AV1H_ERR_UNSUPPORTED,
                           "block %d: transform type %d on a %dx%d transform with coded levels: AV1's transform "
                           "sets have the 1D types only up to 16x16", i, tx.tx_type, tx.w, tx.h);
          if (tx.qm_off >= 0 && (!f.qm_pool || tx.qm_off + area > f.num_qm))
            return set_err(err, AV1H_ERR_INVALID, "block %d: transform %d quantiser matrix out of range", i, t);
          q = DevDq{tx.level_offset, tx.qm_off >= 0 ? tx.qm_off : -1, (uint16_t)tx.dc_step, (uint16_t)tx.ac_step,
                    den == 1 ? 0 : (den == 2 ? 1 : 2)};
        }
        H->dq.push_back(q);
      }
      H->tx.push_back(dt);
    }
    d.tx_count = (int)H->tx.size() - d.tx_first;
    sb_block(L, d, H->tx.data() + d.tx_first, plan, H);
  }
  /* pools padded like the reference planner's (engine.cpp:595-660): the
   * caller's arrays are staged as they are, zeros after them */
  const size_t coef_n = std::max((size_t)f.num_coefficients, cend + 64);
  const size_t res_n = std::max((size_t)f.num_residual, rend + 64);

  /* The order the queue hands blocks out. The queued blocks form a graph:
   * a block depends on the queued blocks that own the cells it waits for
   * (left column, above row, corner, co-located luma for chroma from luma,
   * the source of an intra block copy), all earlier in coding order. The
   * static order is by dependency level (a block's longest chain of queued
   * predecessors), coding order inside a level: every block a block waits
   * for is handed out before it, so the queue always progresses, and the
   * blocks a level holds can all run at once. */
  {
    const int nq = (int)H->queue_list.size();
    std::vector<int32_t>* owner = H->s_owner;
    for (int p = 0; p < 3; ++p) owner[p].assign((size_t)L.cells_w[p] * L.cells_h[p], -1);
    plan->q_dep_off.assign(1, 0);
    plan->q_dep.clear();
    plan->q_area.assign((size_t)nq, 0);
    H->s_sb_q.resize((size_t)nq);
    std::vector<int32_t> level((size_t)nq, 0);
    std::vector<int32_t> deps;
    /* packed edges need every queued block on whole cells, so that each
     * packed word has one writer */
    bool packed = true;
    for (int qi = 0; qi < nq; ++qi) {
      const DevBlock& d = H->blocks[(size_t)H->queue_list[(size_t)qi]];
      if ((d.x | d.y | d.w | d.h) & ((1 << L.cell_shift[d.plane]) - 1)) packed = false;
    }
    std::vector<char> need_mark((size_t)nq, 0);
    std::vector<char>* queued = H->s_queued;
    for (int p = 0; p < 3; ++p) queued[p].assign((size_t)L.cells_w[p] * L.cells_h[p], 0);
    for (int qi = 0; qi < nq; ++qi) {
      const DevBlock& d = H->blocks[(size_t)H->queue_list[(size_t)qi]];
      const int p = d.plane, sh = L.cell_shift[p];
      for (int cy = d.y >> sh; cy < std::min((d.y + d.h + (1 << sh) - 1) >> sh, L.cells_h[p]); ++cy)
        for (int cx = d.x >> sh; cx < std::min((d.x + d.w + (1 << sh) - 1) >> sh, L.cells_w[p]); ++cx)
          queued[p][(size_t)cy * L.cells_w[p] + cx] = 1;
    }
    for (int qi = 0; qi < nq; ++qi) {
      DevBlock& d = H->blocks[(size_t)H->queue_list[(size_t)qi]];
      const int p = d.plane, sh = L.cell_shift[p], cell = 1 << sh;
      const int gw = L.grid_w[p], gh = L.grid_h[p];
      deps.clear();
      auto dep = [&](int pl, int cx, int cy) {
        if (cx < 0 || cy < 0 || cx >= L.cells_w[pl] || cy >= L.cells_h[pl]) return;
        const int o = owner[pl][(size_t)cy * L.cells_w[pl] + cx];
        if (o >= 0 && (deps.empty() || deps.back() != o)) deps.push_back(o);
      };
      const int ly0 = d.y & ~(cell - 1), ax0 = d.x & ~(cell - 1);
      if (d.x > 0 && d.dep_y1 >= ly0)
        for (int y = ly0; y <= std::min(d.dep_y1, gh - 1); y += cell) dep(p, (d.x - 1) >> sh, y >> sh);
      if (d.y > 0 && d.dep_x1 >= ax0)
        for (int x = ax0; x <= std::min(d.dep_x1, gw - 1); x += cell) dep(p, x >> sh, (d.y - 1) >> sh);
      if (d.dep_corner && d.x > 0 && d.y > 0) dep(p, (d.x - 1) >> sh, (d.y - 1) >> sh);
      if (d.kind == kIntra && d.mode == AV1H_PRED_CFL && p > 0) {
        const int s0 = L.cell_shift[0];
        const int lx0 = (d.x << L.ss_x) >> s0, ly0l = (d.y << L.ss_y) >> s0;
        const int lx1 = std::min((d.x + d.w - 1) << L.ss_x, L.grid_w[0] - 1) >> s0;
        const int ly1 = std::min((d.y + d.h - 1) << L.ss_y, L.grid_h[0] - 1) >> s0;
        for (int cy = ly0l; cy <= ly1; ++cy)
          for (int cx = lx0; cx <= lx1; ++cx) dep(0, cx, cy);
      }
      if ((d.flags & kIntraBC) && d.dep_rx1 >= d.dep_rx0 && d.dep_ry1 >= d.dep_ry0)
        for (int cy = d.dep_ry0; cy <= d.dep_ry1; ++cy)
          for (int cx = d.dep_rx0; cx <= d.dep_rx1; ++cx) dep(p, cx, cy);
      /* a handful of entries: insertion sort, duplicates dropped */
      {
        size_t m = 0;
        for (size_t i = 0; i < deps.size(); ++i) {
          const int32_t v = deps[i];
          size_t j = m;
          while (j > 0 && deps[j - 1] > v) --j;
          if (j > 0 && deps[j - 1] == v) continue;
          for (size_t k = m; k > j; --k) deps[k] = deps[k - 1];
          deps[j] = v;
          ++m;
        }
        deps.resize(m);
      }
      if (packed && packed_edges(L, d, owner[p], queued[p], H->blocks, H->queue_list)) {
        d.flags |= kPackedEdges;
      } else {
        for (int32_t o : deps) need_mark[(size_t)o] = 1;
      }
      int lv = 0;
      for (int32_t o : deps) lv = std::max(lv, level[(size_t)o] + 1);
      level[(size_t)qi] = lv;
      plan->q_area[(size_t)qi] = d.w * d.h;
      sb_queued(L, d, qi, plan, H);
      plan->q_dep.insert(plan->q_dep.end(), deps.begin(), deps.end());
      plan->q_dep_off.push_back((int32_t)plan->q_dep.size());
      /* the block's own cells */
      const int cw = std::max(d.w >> sh, 1), ch = std::max(d.h >> sh, 1);
      for (int cy = d.y >> sh; cy < std::min((d.y >> sh) + ch, L.cells_h[p]); ++cy)
        for (int cx = d.x >> sh; cx < std::min((d.x >> sh) + cw, L.cells_w[p]); ++cx)
          owner[p][(size_t)cy * L.cells_w[p] + cx] = qi;
    }
    for (int k = 0; k < 4; ++k) plan->q_counts[k] = 0;
    plan->q_counts[0] = nq;
    for (int qi = 0; qi < nq; ++qi) {
      DevBlock& d = H->blocks[(size_t)H->queue_list[(size_t)qi]];
      if (!need_mark[(size_t)qi]) d.flags |= kSkipMark;
      plan->q_counts[1] += (d.flags & kPackedEdges) != 0;
      plan->q_counts[2] += (d.flags & (kEdgeRow | kEdgeCol)) != 0;
      plan->q_counts[3] += (d.flags & kSkipMark) != 0;
    }
    plan->q_block = H->queue_list; /* coding order */
    plan->q_level = level;
    /* by level, coding order inside a level: a counting sort */
    std::vector<int32_t> ord((size_t)nq);
    {
      int top = 0;
      for (int k = 0; k < nq; ++k) top = std::max(top, level[(size_t)k]);
      std::vector<int32_t> start((size_t)top + 2, 0);
      for (int k = 0; k < nq; ++k) ++start[(size_t)level[(size_t)k] + 1];
      for (int v = 0; v <= top; ++v) start[(size_t)v + 1] += start[(size_t)v];
      for (int k = 0; k < nq; ++k) ord[(size_t)start[(size_t)level[(size_t)k]]++] = k;
    }
    std::vector<int32_t> q((size_t)nq);
    for (int k = 0; k < nq; ++k) q[(size_t)k] = H->queue_list[(size_t)ord[(size_t)k]];
    H->queue_list.swap(q);
    plan->queue_host = H->queue_list;
    plan->queue_qi = ord;
    plan->key_off.clear();
    for (int k = 0; k < nq; ++k) {
      const int kk = level[(size_t)ord[(size_t)k]];
      while ((int)plan->key_off.size() <= kk) plan->key_off.push_back((int32_t)k);
    }
    plan->key_off.push_back((int32_t)nq);
    plan->queue_static = plan->queue_host;
    plan->key_off_static = plan->key_off;
  }
  sb_finish(*H, hdr, L, plan);

  /* motion compensation tasks: items (whole blocks, or row strips of blocks
   * over kMcSamples samples) in coding order, with their positions, windows
   * and LDS offsets, packed into tasks within the kernel's budgets */
  int64_t inter_bytes = 0; /* compulsory bytes of the inter blocks: prediction written, references read */
  H->mc_items.clear();
  H->mc_items.reserve(H->inter_list.size() + H->inter_list.size() / 4);
  H->mc_task.clear();
  {
    McTaskDev cur{};
    auto close_task = [&]() {
      if (cur.count > 0) H->mc_task.push_back(cur);
      cur = McTaskDev{};
      cur.first = (int32_t)H->mc_items.size();
    };
    close_task();
    for (int32_t bi : H->inter_list) {
      const DevBlock& d = H->blocks[bi];
      inter_bytes += (int64_t)d.w * d.h * (2 + 2 * (d.compound ? 2 : 1));
      const int sx = d.plane ? hdr.ss_x : 0, sy = d.plane ? hdr.ss_y : 0;
      const int per = d.w * d.h <= kMcSamples ? d.h : kMcSamples / d.w;
      for (int y0 = 0; y0 < d.h; y0 += per) {
        /* built in place; a task that closes before it starts at it */
        McItemDev& it = H->mc_items.emplace_back();
        std::memset(&it, 0, sizeof(it));
        it.x = d.x;
        it.y = d.y;
        it.w = d.w;
        it.h = d.h;
        it.plane = d.plane;
        it.y0 = y0;
        it.rows = std::min(per, d.h - y0);
        it.compound = d.compound;
        it.w0 = d.w0;
        it.w1 = d.w1;
        it.hk = mc_filter_row(d.filt, d.w) * 128;
        it.vk = mc_filter_row(d.filt_v, d.h) * 128;
        it.ref_w = d.ref_w;
        it.ref_h = d.ref_h;
        it.nref = d.compound ? 2 : 1;
        int msz = 0;
        for (int r = 0; r < it.nref; ++r) {
          McRefDev& q = it.r[r];
          q.slot = d.slot[r];
          mc_position(d.x, d.y, d.mv[r][0], d.mv[r][1], sx, sy, d.scale_x, d.scale_y, &q);
          const int c0 = mc_col(d.x, q, 0), c1 = mc_col(d.x, q, d.w - 1);
          const int c2 = mc_row(d.y, q, y0), c3 = mc_row(d.y, q, y0 + it.rows - 1);
          q.wx0 = c0 - 3;
          q.wy0 = c2 - 3;
          q.ww = c1 - c0 + 8;
          q.wh = c3 - c2 + 8;
          msz += q.wh * d.w;
        }
        const int sb = it.rows * d.w;
        if (cur.count + 1 > kMcItems || cur.samples + sb > kMcSamples || cur.mtotal + msz > kMcMid) {
          close_task();
          cur.first = (int32_t)H->mc_items.size() - 1;
        }
        if (msz > kMcMid) return set_err(err, AV1H_ERR_UNSUPPORTED, "motion compensation window");
        it.px0 = cur.samples;
        for (int r = 0; r < 2; ++r) {
          McRefDev& q = it.r[r];
          q.win_off = cur.wtotal;
          q.mid_off = cur.mtotal;
          if (r < it.nref) {
            cur.wtotal += q.ww * q.wh;
            cur.mtotal += q.wh * d.w;
          }
        }
        cur.samples += sb;
        ++cur.count;
      }
    }
    close_task();
  }
  const int32_t mc_tasks = (int32_t)H->mc_task.size();
  /* transforms: inter set then intra set, each by class (larger side), each
   * class by size (width, then height) so a wave's transforms mostly share
   * one; coding order between equals. A counting sort over the 2 x 25 (set,
   * width, height) keys, which are in that order already. */
  H->itx_all.assign(H->tx.size(), 0);
  int32_t itx_off[2][kTxClasses + 1];
  {
    constexpr int kKeys = 2 * 25;
    /* class c holds the 2c + 1 sizes whose larger side is 4 << c, from key
     * c * c on: widths below 4 << c (height 4 << c) first, then width 4 << c
     * by height */
    auto key = [](const DevTx& t) {
      const int lw = __builtin_ctz((unsigned)t.w) - 2, lh = __builtin_ctz((unsigned)t.h) - 2;
      const int cls = std::max(lw, lh);
      const int set = (t.flags & kTxToResidual) ? 1 : 0;
      return set * 25 + cls * cls + (lw < cls ? lw : cls + lh);
    };
    int32_t cnt[kKeys + 1] = {};
    for (const DevTx& t : H->tx) ++cnt[key(t) + 1];
    for (int k = 0; k < kKeys; ++k) cnt[k + 1] += cnt[k];
    for (int set = 0; set < 2; ++set)
      for (int c = 0; c <= kTxClasses; ++c) itx_off[set][c] = cnt[set * 25 + c * c];
    for (int t = 0; t < (int)H->tx.size(); ++t) H->itx_all[(size_t)cnt[key(H->tx[(size_t)t])]++] = t;
  }

  /* deblocking: vertical edges then horizontal, with a line prefix each;
   * one pass over the caller's edges, the horizontal ones gathered apart
   * and appended */
  H->deblock.clear();
  H->deblock.reserve((size_t)f.num_deblock);
  std::vector<DevDeblock>& horiz = H->s_deblock_h;
  horiz.clear();
  int lines[2] = {0, 0}, nedge[2] = {0, 0};
  {
    for (int64_t i = 0; i < f.num_deblock; ++i) {
      const av1h_deblock_edge& e = f.deblock[i];
      const int pass = e.vertical != 0 ? 0 : 1;
      if (e.plane < 0 || e.plane > 2 || e.length < 0 || e.level < 0 || e.level > 255 || e.filter_size < 0 ||
          e.filter_size > 255)
        return set_err(err, AV1H_ERR_INVALID, "deblock edge %lld", (long long)i);
      const int reach = e.filter_size >= 16 ? 7 : (e.filter_size >= 8 ? 4 : 2);
      const int rows = L.out_h[e.plane] + (e.plane ? 32 : 64);
      const int across = e.vertical ? e.x : e.y, along = e.vertical ? e.y : e.x;
      const int across_lim = e.vertical ? L.stride[e.plane] : rows;
      const int along_lim = e.vertical ? rows : L.stride[e.plane];
      if (across < reach || across + reach > across_lim || along < 0 || along + e.length > along_lim)
        return set_err(err, AV1H_ERR_INVALID, "deblock edge %lld out of range", (long long)i);
      if (e.length == 0) continue;
      DevDeblock d{};
      d.plane_vertical = (uint8_t)(e.plane | (e.vertical ? 4 : 0));
      d.x = (uint16_t)e.x;
      d.y = (uint16_t)e.y;
      d.len = (uint16_t)e.length;
      d.level = (uint8_t)e.level;
      d.filter_size = (uint8_t)e.filter_size;
      d.line0 = lines[pass];
      lines[pass] += e.length;
      ++nedge[pass];
      (pass ? horiz : H->deblock).push_back(d);
    }
    H->deblock.insert(H->deblock.end(), horiz.begin(), horiz.end());
  }
  /* the first edge of every 256 line chunk of each pass */
  H->db_chunk.clear();
  {
    size_t e0 = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const size_t n = (size_t)nedge[pass];
      for (int64_t k = 0, e = 0; k < (lines[pass] + 255) / 256; ++k) {
        while ((size_t)e + 1 < n && H->deblock[e0 + e + 1].line0 <= k * 256) ++e;
        H->db_chunk.push_back((int32_t)e);
      }
      e0 += n;
    }
  }
  /* CDEF: luma blocks first (one launch per block size), coding order
   * inside each; and the luma 8x8 blocks whose direction is searched */
  int32_t n_cdef_luma = 0;
  {
    const int cbw = 8 >> hdr.ss_x, cbh = 8 >> hdr.ss_y;
    for (int64_t i = 0; i < f.num_cdef; ++i) {
      const av1h_cdef_block& cb = f.cdef[i];
      if (cb.plane < 0 || cb.plane > 2 || cb.x < 0 || cb.y < 0 || cb.x >= L.grid_w[cb.plane] ||
          cb.y >= L.grid_h[cb.plane] || cb.damping < 0 || cb.damping > 15 || cb.dir > 7 ||
          cb.pri_strength < 0 || cb.pri_strength > 255 || cb.sec_strength < 0 || cb.sec_strength > 255)
        return set_err(err, AV1H_ERR_INVALID, "cdef block %lld", (long long)i);
      const int bw = cb.plane ? cbw : 8, bh = cb.plane ? cbh : 8;
      if (cb.x % bw || cb.y % bh)
        return set_err(err, AV1H_ERR_UNSUPPORTED, "cdef block %lld is not on the %dx%d grid", (long long)i, bw, bh);
      n_cdef_luma += cb.plane == 0;
    }
    H->cdef.resize((size_t)f.num_cdef);
    size_t nl = 0, nc = (size_t)n_cdef_luma;
    for (int64_t i = 0; i < f.num_cdef; ++i) {
      const av1h_cdef_block& cb = f.cdef[i];
      H->cdef[cb.plane == 0 ? nl++ : nc++] =
          DevCdef{(uint16_t)cb.x, (uint16_t)cb.y, (uint8_t)cb.pri_strength, (uint8_t)cb.sec_strength,
                  (uint8_t)(cb.plane | cb.damping << 2), (int8_t)(cb.dir < 0 ? -1 : cb.dir)};
    }
  }
  H->cdef_dir_list.clear();
  {
    const int bw8 = L.grid_w[0] >> 3, bh8 = L.grid_h[0] >> 3;
    std::vector<uint8_t>& need = H->s_need;
    need.assign((size_t)bw8 * bh8, 0);
    for (const DevCdef& d : H->cdef) {
      if (d.dir >= 0) continue;
      const int pl = d.plane_damping & 3;
      const int lx = pl ? d.x << hdr.ss_x : d.x, ly = pl ? d.y << hdr.ss_y : d.y;
      const int bx = lx >> 3, by = ly >> 3;
      if (bx < bw8 && by < bh8 && !need[(size_t)by * bw8 + bx]) {
        need[(size_t)by * bw8 + bx] = 1;
        H->cdef_dir_list.push_back(bx | (by << 16));
      }
    }
  }
  /* loop restoration jobs: one per unit per stripe (engine.cpp:1382-1462) */
  H->lr.clear();
  int lr_px = 0;
  for (int p = 0; p < 3; ++p) {
    if (hdr.restoration_type[p] == AV1H_RESTORE_NONE) continue;
    const int sy = p > 0 ? hdr.ss_y : 0;
    if (hdr.restoration_unit_log2[p] < 0 || hdr.restoration_unit_log2[p] > 10)
      return set_err(err, AV1H_ERR_INVALID, "restoration unit size");
    const int usz = 1 << hdr.restoration_unit_log2[p];
    const int ucols = std::max(hdr.restoration_unit_cols[p], 1);
    const int pw = L.out_w[p], ph = L.out_h[p];
    const int fh2 = (hdr.height + (sy ? 1 << (sy - 1) : 0)) >> sy;
    const int urows = std::max((fh2 + (usz >> 1)) / usz, 1);
    for (int y = 0; y < ph;) {
      const int stripe = ((y << sy) + 8) / 64;
      const int s0 = (-8 + stripe * 64) >> sy;
      const int s1 = s0 + (64 >> sy) - 1;
      const int y1 = std::min(ph, s1 + 1);
      const int urow = std::min(urows - 1, (((y << sy) + 8) >> sy) / usz);
      for (int uc = 0; uc < ucols; ++uc) {
        const int64_t idx = (int64_t)urow * ucols + uc;
        if (idx >= f.num_restoration[p]) continue;
        const av1h_lr_unit& u = f.restoration[p][idx];
        if (u.type == AV1H_RESTORE_NONE) continue;
        const int ux0 = uc * usz;
        const int ux1 = uc + 1 == ucols ? pw : std::min(ux0 + usz, pw);
        if (ux1 <= ux0) continue;
        DevLr j;
        std::memset(&j, 0, sizeof(j));
        j.plane = p;
        j.x = ux0;
        j.y = y;
        j.w = ux1 - ux0;
        j.h = y1 - y;
        j.stripe0 = s0;
        j.stripe1 = s1;
        j.type = u.type == AV1H_RESTORE_WIENER ? 1 : 2;
        if (j.type == 1) {
          for (int k = 0; k < 3; ++k) {
            j.vf[k] = u.wiener[0][k];
            j.hf[k] = u.wiener[1][k];
          }
        } else {
          const int* sp = av1t_spec_table + AV1T_OFF_SGR_PARAMS + 4 * (u.sgr_set & 15);
          for (int k = 0; k < 2; ++k) {
            j.r[k] = sp[2 * k];
            j.eps[k] = sp[2 * k + 1];
            j.xqd[k] = u.sgr_xqd[k];
          }
        }
        j.px0 = lr_px;
        lr_px += j.w * j.h;
        H->lr.push_back(j);
      }
      y = y1;
    }
  }
  plan->uses_lr = !H->lr.empty();
  H->lr_tiles.clear();
  for (int jn = 0; jn < (int)H->lr.size(); ++jn)
    for (int ty = 0; ty < H->lr[jn].h; ty += kLrTileRows)
      for (int tx = 0; tx < H->lr[jn].w; tx += kLrTileCols) {
        H->lr_tiles.push_back(jn);
        H->lr_tiles.push_back(tx | (ty << 16));
      }
  /* film grain parameters */
  std::memset(&H->grain, 0, sizeof(H->grain));
  plan->has_grain = hdr.grain.apply != 0;
  if (plan->has_grain) {
    const av1h_grain& g = hdr.grain;
    DevGrain& d = H->grain;
    if (g.seed < 0 || g.seed > 65535) return set_err(err, AV1H_ERR_INVALID, "grain seed");
    if (g.ar_coeff_lag < 0 || g.ar_coeff_lag > 3) return set_err(err, AV1H_ERR_INVALID, "grain lag");
    if (g.num_y_points < 0 || g.num_y_points > 16 || g.num_cb_points < 0 || g.num_cb_points > 16 ||
        g.num_cr_points < 0 || g.num_cr_points > 16)
      return set_err(err, AV1H_ERR_INVALID, "grain points");
    /* the auto-regressive coefficients are coded as 8 bit values minus 128
     * (src/frontend/headers.cpp:716-725); the template filter multiplies
     * them with 24 bit products */
    for (int k = 0; k < 25; ++k)
      if ((k < 24 && (g.ar_coeffs_y[k] < -128 || g.ar_coeffs_y[k] > 127)) || g.ar_coeffs_cb[k] < -128 ||
          g.ar_coeffs_cb[k] > 127 || g.ar_coeffs_cr[k] < -128 || g.ar_coeffs_cr[k] > 127)
        return set_err(err, AV1H_ERR_INVALID, "grain auto-regressive coefficient out of range");
    d.seed = g.seed;
    d.bit_depth = hdr.bit_depth;
    d.num_y = g.num_y_points;
    d.num_cb = g.num_cb_points;
    d.num_cr = g.num_cr_points;
    d.cfl = g.chroma_scaling_from_luma;
    d.lag = g.ar_coeff_lag;
    d.shift = g.ar_coeff_shift_minus_6 + 6;
    d.scale_shift = 12 - hdr.bit_depth + g.grain_scale_shift;
    d.scaling_shift = g.grain_scaling_minus_8 + 8;
    d.overlap = g.overlap_flag;
    d.cb_mult = g.cb_mult;
    d.cb_luma_mult = g.cb_luma_mult;
    d.cb_offset = g.cb_offset;
    d.cr_mult = g.cr_mult;
    d.cr_luma_mult = g.cr_luma_mult;
    d.cr_offset = g.cr_offset;
    const int maxv = (1 << hdr.bit_depth) - 1;
    d.min_value = 0;
    d.max_luma = d.max_chroma = maxv;
    if (g.clip_to_restricted_range) {
      d.min_value = 16 << (hdr.bit_depth - 8);
      d.max_luma = 235 << (hdr.bit_depth - 8);
      d.max_chroma = hdr.color_matrix == 3 ? d.max_luma : (240 << (hdr.bit_depth - 8));
    }
    for (int k = 0; k < 24; ++k) d.ar_y[k] = g.ar_coeffs_y[k];
    for (int k = 0; k < 25; ++k) {
      d.ar_cb[k] = g.ar_coeffs_cb[k];
      d.ar_cr[k] = g.ar_coeffs_cr[k];
    }
    const bool chroma = g.chroma_scaling_from_luma || g.num_cb_points > 0 || g.num_cr_points > 0;
    for (int pl = 0; pl < 3; ++pl) {
      if (pl > 0 && !chroma)
        std::memset(d.lut + 256 * pl, 0, 256 * sizeof(int));
      else
        scaling_lut(g, pl, d.lut + 256 * pl);
    }
    const int sy = hdr.ss_y;
    const int ph = (L.out_h[0] + sy) >> sy;
    d.stripes = std::max((L.out_h[0] + 31) / 32, (ph + (32 >> sy) - 1) / (32 >> sy));
    d.blocks = ((L.out_w[0] + 1) / 2 + 15) / 16;
  }
  plan->blocks = nb;
  plan->grain = H->grain;

  /* compulsory HBM bytes of each stage at 16 bit samples: what a stage must
   * read and write at least, independent of how the kernels tile it */
  for (int64_t& v : plan->alg_bytes) v = 0;
  {
    /* part 0: inter samples written once and read from each reference, the
     * coefficients of every transform (int16 each); the queue: per queued
     * sample its residual read and the picture sample written */
    plan->alg_bytes[AV1H_STAGE_INTER] = inter_bytes;
    for (const DevTx& t : H->tx)
      if (!(t.flags & kTxToResidual)) plan->alg_bytes[AV1H_STAGE_INTER] += 2 * (int64_t)t.w * t.h;
    for (const int32_t a : plan->q_area) plan->alg_bytes[AV1H_STAGE_QUEUE] += 4 * (int64_t)a;
    bool has_edges[3] = {false, false, false};
    for (const DevDeblock& d : H->deblock) {
      has_edges[d.plane_vertical & 3] = true;
      if (has_edges[0] && has_edges[1] && has_edges[2]) break;
    }
    int64_t disp = 0;
    for (int p = 0; p < 3; ++p) {
      if (has_edges[p]) plan->alg_bytes[AV1H_STAGE_DEBLOCK] += 4 * (int64_t)L.plane_w[p] * L.plane_h[p];
      if (plan->upscaled)
        plan->alg_bytes[AV1H_STAGE_SUPERRES] +=
            (plan->uses_lr ? 2 : 1) * 2 * (int64_t)(L.plane_w[p] + L.out_w[p]) * L.out_h[p];
      disp += (int64_t)L.out_w[p] * L.out_h[p];
    }
    plan->alg_bytes[AV1H_STAGE_CDEF] = 4 * 64 * (int64_t)n_cdef_luma +
                                       4 * (int64_t)(8 >> hdr.ss_x) * (8 >> hdr.ss_y) *
                                           (int64_t)(H->cdef.size() - (size_t)n_cdef_luma);
    for (const DevLr& d : H->lr) plan->alg_bytes[AV1H_STAGE_LR] += 4 * (int64_t)d.w * d.h + 8 * (int64_t)d.w;
    plan->alg_bytes[AV1H_STAGE_STORE] = 4 * disp;
    if (plan->has_grain) {
      const DevGrain& g = H->grain;
      const bool chroma = g.cfl || g.num_cb > 0 || g.num_cr > 0;
      plan->alg_bytes[AV1H_STAGE_GRAIN] =
          (g.num_y > 0 ? 4 * (int64_t)L.out_w[0] * L.out_h[0] : 0) +
          (chroma ? 8 * (int64_t)L.out_w[1] * L.out_h[1] : 0);
    }
  }

  /* The device reads the descriptors of queued and special blocks only
   * (through the queue and the special list): those are uploaded, in coding
   * order, and every list that names blocks for the device is renumbered. */
  {
    std::vector<int32_t>& dev_of = H->s_dev_of;
    dev_of.assign(H->blocks.size(), -1);
    size_t ndev = 0;
    for (int32_t b : H->queue_list) ndev += dev_of[(size_t)b] < 0 ? (dev_of[(size_t)b] = 0, 1) : 0;
    for (int32_t b : H->special_list) ndev += dev_of[(size_t)b] < 0 ? (dev_of[(size_t)b] = 0, 1) : 0;
    H->dev_blocks.clear();
    /* every block read by the device (an intra frame): the numbering stays */
    H->dev_all = ndev == H->blocks.size();
    if (!H->dev_all) {
      for (size_t b = 0; b < H->blocks.size(); ++b)
        if (dev_of[b] == 0) {
          dev_of[b] = (int32_t)H->dev_blocks.size();
          H->dev_blocks.push_back(H->blocks[b]);
        }
      for (std::vector<int32_t>* v : {&H->queue_list, &H->special_list, &plan->queue_host, &plan->queue_static,
                                      &plan->q_block})
        for (int32_t& b : *v) b = dev_of[(size_t)b];
    }
  }

  /* one device buffer: the arrays back to back, 256 byte aligned; the
   * coefficients last, room for int32 kept, int16 uploaded when they fit */
  H->parts.clear();
  auto add = [&](int k, const void* src, size_t bytes, size_t fill = 0) {
    H->part_of[k] = (int)H->parts.size();
    H->parts.push_back(Host::Part{src, bytes, fill, 0});
  };
  const BlockVec& dblk = H->dev_all ? H->blocks : H->dev_blocks;
  add(kPartBlocks, dblk.data(), dblk.size() * sizeof(DevBlock));
  add(kPartTx, H->tx.data(), H->tx.size() * sizeof(DevTx));
  add(kPartInter, H->inter_list.data(), H->inter_list.size() * 4);
  add(kPartQueue, H->queue_list.data(), H->queue_list.size() * 4);
  add(kPartSpecial, H->special_list.data(), H->special_list.size() * 4);
  add(kPartObmc, H->obmc.data(), H->obmc.size() * sizeof(DevObmc));
  add(kPartResidual, f.residual, (size_t)f.num_residual * 4, (res_n - (size_t)f.num_residual) * 4);
  add(kPartPalette, f.palette_map, (size_t)f.num_palette, 8);
  add(kPartDeblock, H->deblock.data(), H->deblock.size() * sizeof(DevDeblock));
  add(kPartCdef, H->cdef.data(), H->cdef.size() * sizeof(DevCdef));
  add(kPartLr, H->lr.data(), H->lr.size() * sizeof(DevLr));
  add(kPartMcItems, H->mc_items.data(), H->mc_items.size() * sizeof(McItemDev));
  add(kPartMcTasks, H->mc_task.data(), H->mc_task.size() * sizeof(McTaskDev));
  add(kPartItx, H->itx_all.data(), H->itx_all.size() * 4);
  std::memcpy(H->itx_off, itx_off, sizeof(itx_off));
  add(kPartItxOffs, H->itx_off, sizeof(itx_off));
  add(kPartDbChunk, H->db_chunk.data(), H->db_chunk.size() * 4);
  add(kPartCdefDir, H->cdef_dir_list.data(), H->cdef_dir_list.size() * 4);
  add(kPartLrTiles, H->lr_tiles.data(), H->lr_tiles.size() * 4);
  add(kPartGrain, &H->grain, sizeof(DevGrain));
  add(kPartLevels, levels ? f.levels : nullptr, levels ? (size_t)f.num_levels * 2 : 0);
  add(kPartQm, levels ? f.qm_pool : nullptr, levels ? (size_t)f.num_qm : 0);
  add(kPartDq, H->dq.data(), H->dq.size() * sizeof(DevDq));
  if (levels)
    add(kPartCoef, nullptr, 0);
  else
    add(kPartCoef, f.coefficients, (size_t)f.num_coefficients * 4, (coef_n - (size_t)f.num_coefficients) * 4);
  size_t total = 0;
  for (Host::Part& pt : H->parts) {
    pt.off = total;
    total += align256(pt.bytes + pt.fill);
  }
  H->total = total;
  FramePlanDev& P = plan->dev;
  std::memcpy(P.itx_off, itx_off, sizeof(itx_off));
  P.n_special = (int32_t)H->special_list.size();
  P.mc_tasks = mc_tasks;
  P.n_cdef_luma = n_cdef_luma;
  P.n_cdef_dir = (int32_t)H->cdef_dir_list.size();
  P.n_lr_tiles = (int32_t)(H->lr_tiles.size() / 2);
  P.n_inter = (int)H->inter_list.size();
  P.n_queue = (int)H->queue_list.size();
  P.n_deblock_v = nedge[0];
  P.n_deblock_h = nedge[1];
  P.lines_v = lines[0];
  P.lines_h = lines[1];
  P.n_cdef = (int)H->cdef.size();
  P.n_lr = (int)H->lr.size();
  P.lr_px = lr_px;
  return AV1H_OK;
}

/* Writes the arrays of a flattened frame into `dst` (room for Host::total
 * bytes: the staging buffer of its upload), zeros after the pools that are
 * padded. The coefficients, last, go as int16 when every value fits (*coef16
 * set), else as int32. Returns the bytes to upload. */
size_t stage_parts(const Host& H, uint8_t* dst, int32_t* coef16, const std::function<void(size_t)>& staged) {
  const Host::Part* coef = &H.parts[(size_t)H.part_of[kPartCoef]];
  for (const Host::Part& pt : H.parts) {
    if (&pt == coef) continue;
    if (pt.bytes) std::memcpy(dst + pt.off, pt.src, pt.bytes);
    if (pt.fill) std::memset(dst + pt.off + pt.bytes, 0, pt.fill);
    if (staged) staged(pt.off + pt.bytes + pt.fill);
  }
  if (staged) staged(coef->off);
  const size_t n = coef->bytes / 4, pad = coef->fill / 4;
  const int32_t* src = static_cast<const int32_t*>(coef->src);
  int16_t* d16 = reinterpret_cast<int16_t*>(dst + coef->off);
  uint32_t wide = 0;
  for (size_t i = 0; i < n; ++i) {
    const int32_t v = src[i];
    wide |= (uint32_t)(v + 32768) > 65535u;
    d16[i] = (int16_t)v;
  }
  if (!wide) {
    std::memset(d16 + n, 0, pad * 2);
    *coef16 = 1;
    return coef->off + (n + pad) * 2;
  }
  if (n) std::memcpy(dst + coef->off, src, n * 4);
  std::memset(dst + coef->off + n * 4, 0, pad * 4);
  *coef16 = 0;
  return coef->off + (n + pad) * 4;
}

/* Points the plan's device descriptor at its arrays in `plan->dbuf`. */
void bind_dev(Plan* plan, const Host& H) {
  uint8_t* db = static_cast<uint8_t*>(plan->dbuf);
  auto at = [&](int k) { return db + H.parts[(size_t)H.part_of[k]].off; };
  FramePlanDev& P = plan->dev;
  P.blocks = reinterpret_cast<const DevBlock*>(at(kPartBlocks));
  P.tx = reinterpret_cast<const DevTx*>(at(kPartTx));
  P.obmc = reinterpret_cast<const DevObmc*>(at(kPartObmc));
  P.special_list = reinterpret_cast<const int32_t*>(at(kPartSpecial));
  P.inter_list = reinterpret_cast<const int32_t*>(at(kPartInter));
  P.queue_list = reinterpret_cast<const int32_t*>(at(kPartQueue));
  P.coef = at(kPartCoef);
  const bool lv = H.dq.size() > 0;
  P.levels = lv ? reinterpret_cast<const int16_t*>(at(kPartLevels)) : nullptr;
  P.qm = lv ? at(kPartQm) : nullptr;
  P.dq = lv ? reinterpret_cast<const DevDq*>(at(kPartDq)) : nullptr;
  P.residual = reinterpret_cast<const int32_t*>(at(kPartResidual));
  P.palette = at(kPartPalette);
  P.deblock = reinterpret_cast<const DevDeblock*>(at(kPartDeblock));
  P.cdef = reinterpret_cast<const DevCdef*>(at(kPartCdef));
  P.lr = reinterpret_cast<const DevLr*>(at(kPartLr));
  P.grain = reinterpret_cast<const DevGrain*>(at(kPartGrain));
  P.itx_list = reinterpret_cast<const int32_t*>(at(kPartItx));
  P.itx_offs = reinterpret_cast<const int32_t*>(at(kPartItxOffs));
  P.mc_items = reinterpret_cast<const McItemDev*>(at(kPartMcItems));
  P.mc_task = reinterpret_cast<const McTaskDev*>(at(kPartMcTasks));
  P.db_chunk = reinterpret_cast<const int32_t*>(at(kPartDbChunk));
  P.cdef_dir_list = reinterpret_cast<const int32_t*>(at(kPartCdefDir));
  P.lr_tiles = reinterpret_cast<const int2*>(at(kPartLrTiles));
}

/* The synchronous path's upload: the scratch plan's device buffer (grown when
 * it must), staged through the context's pinned buffer so the copy is a DMA,
 * on the decode stream. */
int upload(Ctx* c, Plan* plan, const Host& H) {
  const size_t n = H.total;
  if (plan->dbytes < n) {
    if (plan->dbuf) (void)hipFree(plan->dbuf);
    plan->dbuf = nullptr;
    plan->dbytes = 0;
    HIP_OK(c, hipMalloc(&plan->dbuf, n));
    plan->dbytes = n;
  }
  bind_dev(plan, H);
  if (c->host_pin_cap < n) {
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (c->host_pin) (void)hipHostFree(c->host_pin);
    c->host_pin = nullptr;
    c->host_pin_cap = 0;
    HIP_OK(c, hipHostMalloc(&c->host_pin, n, hipHostMallocDefault));
    c->host_pin_cap = n;
  }
  HIP_OK(c, hipStreamSynchronize(c->stream));
  const size_t up = stage_parts(H, static_cast<uint8_t*>(c->host_pin), &plan->dev.coef16, nullptr);
  HIP_OK(c, hipMemcpyAsync(plan->dbuf, c->host_pin, up, hipMemcpyHostToDevice, c->stream));
  c->dirty = true;
  return AV1H_OK;
}

/* ---- pre-planning from several threads ------------------------------------
 * A plan owns a device buffer from the context's recycling pool and is
 * uploaded through a stager (a pinned buffer and a stream of its own), so
 * any number of threads plan at once, without the context's lock and without
 * waiting for the decode stream. */
constexpr size_t kPlanPoolBytes = (size_t)2 << 30; /* device bytes a context keeps for reuse */
constexpr size_t kPlanGrain = (size_t)1 << 20;

int take_buffer(Ctx* c, size_t n, void** out, size_t* got, std::string* err) {
  {
    std::lock_guard<std::mutex> lk(c->plan_mu);
    size_t best = c->plan_bufs.size();
    for (size_t i = 0; i < c->plan_bufs.size(); ++i)
      if (c->plan_bufs[i].first >= n && c->plan_bufs[i].first <= 2 * n + kPlanGrain &&
          (best == c->plan_bufs.size() || c->plan_bufs[i].first < c->plan_bufs[best].first))
        best = i;
    if (best < c->plan_bufs.size()) {
      *got = c->plan_bufs[best].first;
      *out = c->plan_bufs[best].second;
      c->plan_buf_bytes -= *got;
      c->plan_bufs.erase(c->plan_bufs.begin() + (std::ptrdiff_t)best);
      return AV1H_OK;
    }
  }
  const size_t sz = (n + kPlanGrain - 1) / kPlanGrain * kPlanGrain;
  const hipError_t e = hipMalloc(out, sz);
  if (e != hipSuccess) {
    *out = nullptr;
    return set_err(err, AV1H_ERR_DEVICE, "hipMalloc of a frame plan (%zu bytes): %s", sz, hipGetErrorString(e));
  }
  *got = sz;
  return AV1H_OK;
}

/* back to the pool, or freed when the pool is full */
void give_buffer(Ctx* c, void* p, size_t n) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(c->plan_mu);
    if (c->plan_buf_bytes + n <= kPlanPoolBytes) {
      c->plan_bufs.emplace_back(n, p);
      c->plan_buf_bytes += n;
      return;
    }
  }
  (void)hipFree(p);
}

/* The plan's upload: staged into a pinned buffer, copied on a stream that
 * carries nothing but such copies, and waited for by the planning thread, so
 * the plan is complete when av1h_plan_frame returns and no decode waits for
 * it on the device. An event recorded behind an upload instead (the decode's
 * wait for it) is a barrier in a hardware queue until the copy ends, and
 * the decode work that shares that queue (a process gets few) waited behind
 * it: 25 us per 1080p frame with eight threads planning. A stream per
 * planning thread keeps their copies apart. */
constexpr size_t kUploadChunk = (size_t)2 << 20;

int upload_plan(Ctx* c, Plan* plan, const Host& H, std::string* err) {
  const size_t n = H.total;
  int rc = take_buffer(c, n, &plan->dbuf, &plan->dbytes, err);
  if (rc != AV1H_OK) return rc;
  bind_dev(plan, H);
  Stager* st = nullptr;
  hipStream_t us = nullptr;
  size_t hint = n;
  {
    std::lock_guard<std::mutex> lk(c->plan_mu);
    /* the smallest free stager that holds the plan, else the largest (grown
     * below: growing a pinned buffer, free then allocate, costs
     * milliseconds), else a new one */
    size_t pick = c->stagers.size();
    for (size_t i = 0; i < c->stagers.size(); ++i) {
      const size_t ci = c->stagers[i]->cap;
      if (pick == c->stagers.size()) {
        pick = i;
        continue;
      }
      const size_t cp = c->stagers[pick]->cap;
      if ((ci >= n && (cp < n || ci < cp)) || (ci < n && cp < n && ci > cp)) pick = i;
    }
    if (pick < c->stagers.size()) {
      st = c->stagers[pick];
      c->stagers.erase(c->stagers.begin() + (std::ptrdiff_t)pick);
    }
    c->stage_hint = std::max(c->stage_hint, n);
    hint = c->stage_hint;
    if (!c->uploads.empty()) {
      us = c->uploads.back();
      c->uploads.pop_back();
    }
  }
  hipError_t e = hipSuccess;
  if (!us) e = hipStreamCreateWithFlags(&us, hipStreamNonBlocking);
  if (!st) st = new Stager;
  if (e == hipSuccess && st->cap < n) {
    if (st->pin) (void)hipHostFree(st->pin);
    st->pin = nullptr;
    st->cap = 0;
    /* sized for the largest plan staged so far, so that the stagers stop
     * growing once every kind of frame of the stream has passed */
    const size_t want = std::max(n, hint) + std::max(n, hint) / 4;
    const size_t sz = (want + 4 * kPlanGrain - 1) / (4 * kPlanGrain) * (4 * kPlanGrain);
    e = hipHostMalloc(&st->pin, sz, hipHostMallocDefault);
    if (e == hipSuccess) st->cap = sz;
  }
  if (e == hipSuccess) {
    /* copied in chunks as they are staged: the copy of the earlier arrays
     * overlaps the staging of the later ones (the coefficients' int16
     * conversion last) */
    uint8_t* pin = static_cast<uint8_t*>(st->pin);
    uint8_t* dev = static_cast<uint8_t*>(plan->dbuf);
    size_t issued = 0;
    auto staged = [&](size_t end) {
      if (e == hipSuccess && end >= issued + kUploadChunk) {
        e = hipMemcpyAsync(dev + issued, pin + issued, end - issued, hipMemcpyHostToDevice, us);
        issued = end;
      }
    };
    const size_t up = stage_parts(H, pin, &plan->dev.coef16, staged);
    if (e == hipSuccess && up > issued)
      e = hipMemcpyAsync(dev + issued, pin + issued, up - issued, hipMemcpyHostToDevice, us);
    /* waited for on failure too: the stager and the device buffer go back
     * to their pools only once no copy reads or writes them */
    const hipError_t es = hipStreamSynchronize(us);
    if (e == hipSuccess) e = es;
  }
  {
    std::lock_guard<std::mutex> lk(c->plan_mu);
    c->stagers.push_back(st);
    if (us) c->uploads.push_back(us);
  }
  if (e != hipSuccess) return set_err(err, AV1H_ERR_DEVICE, "frame plan upload: %s", hipGetErrorString(e));
  return AV1H_OK;
}

void free_plan(Plan* p) {
  if (p->dbuf) (void)hipFree(p->dbuf);
  p->dbuf = nullptr;
  p->dbytes = 0;
}

/* grows the per geometry buffers; the pool is re-created when it must grow
 * (as the reference CPU backend does, cpu_backend.cpp:77-91) */
int prepare(Ctx* c, const Layout& L) {
  if (c->cap < (size_t)L.size) {
    for (uint16_t** b : {&c->frame, &c->scratch, &c->lr_curr, &c->lr_cdef, &c->lr_prev_up, &c->pool})
      if (*b) {
        (void)hipFree(*b);
        *b = nullptr;
      }
    c->cap = 0;
    const size_t bytes = (size_t)L.size * 2;
    HIP_OK(c, hipMalloc(&c->frame, bytes));
    HIP_OK(c, hipMalloc(&c->scratch, bytes));
    HIP_OK(c, hipMalloc(&c->lr_curr, bytes));
    HIP_OK(c, hipMalloc(&c->lr_cdef, bytes));
    HIP_OK(c, hipMalloc(&c->lr_prev_up, bytes));
    HIP_OK(c, hipMemsetAsync(c->frame, 0, bytes, c->stream));
    c->dirty = true;
    c->cap = (size_t)L.size;
    c->pool_alloc = 0;
    std::memset(c->used, 0, sizeof(c->used));
    std::memset(c->has_pic, 0, sizeof(c->has_pic));
    c->last_slot = -1;
  }
  if (c->pool_alloc != c->pool_slots) {
    if (c->pool) (void)hipFree(c->pool);
    c->pool = nullptr;
    c->pool_alloc = 0;
    const size_t bytes = c->cap * 2 * (size_t)c->pool_slots;
    HIP_OK(c, hipMalloc(&c->pool, bytes));
    HIP_OK(c, hipMemsetAsync(c->pool, 0, bytes, c->stream));
    c->pool_alloc = c->pool_slots;
    c->dirty = true;
    std::memset(c->used, 0, sizeof(c->used));
    std::memset(c->has_pic, 0, sizeof(c->has_pic));
    c->last_slot = -1;
  }
  const size_t ndir = (size_t)(L.grid_w[0] >> 3) * (L.grid_h[0] >> 3) * 2;
  if (c->cdef_dir_cap < ndir) {
    if (c->cdef_dir) (void)hipFree(c->cdef_dir);
    c->cdef_dir = nullptr;
    c->cdef_dir_cap = 0;
    HIP_OK(c, hipMalloc(&c->cdef_dir, ndir * 4));
    c->cdef_dir_cap = ndir;
  }
  if (c->cells_cap < L.cells || c->edges_cap < L.edge_words) {
    /* epochs restart at 1: every cell and edge word starts below */
    if (c->cells) (void)hipFree(c->cells);
    if (c->edges) (void)hipFree(c->edges);
    c->cells = nullptr;
    c->edges = nullptr;
    c->cells_cap = c->edges_cap = 0;
    HIP_OK(c, hipMalloc(&c->cells, (size_t)L.cells * 4));
    HIP_OK(c, hipMalloc(&c->edges, (size_t)L.edge_words * 8));
    HIP_OK(c, hipMemsetAsync(c->cells, 0, (size_t)L.cells * 4, c->stream));
    HIP_OK(c, hipMemsetAsync(c->edges, 0, (size_t)L.edge_words * 8, c->stream));
    c->dirty = true;
    c->cells_cap = L.cells;
    c->edges_cap = L.edge_words;
    c->epoch = 0;
  }
  c->lay = L;
  c->have_layout = true;
  return AV1H_OK;
}

void drop_graphs(Ctx* c) {
  for (auto& g : c->graph_exec) (void)hipGraphExecDestroy(g.second);
  c->graph_exec.clear();
}

int alloc_slot(Ctx* c) {
  for (int i = 0; i < c->pool_slots; ++i)
    if (!c->used[i]) return i;
  return -1;
}

int ensure_pack(Ctx* c, size_t bytes) {
  if (c->pack_cap < bytes) {
    /* frames in flight may pack into the old buffer: they end first */
    if (c->pack && c->inflight) HIP_OK(c, hipStreamSynchronize(c->stream));
    if (c->pack) (void)hipFree(c->pack);
    c->pack = nullptr;
    c->pack_cap = 0;
    HIP_OK(c, hipMalloc(&c->pack, bytes));
    c->pack_cap = bytes;
  }
  return AV1H_OK;
}

/* ---- batches --------------------------------------------------------------
 * Every decode runs as a batch: the frames of one or more contexts (one per
 * stream), each kernel launched once for all of them. */

int batch_finish(Ctx* lead);

/* whether a decode of `plan` on `c` must replace a buffer of the context
 * (what prepare and the grain and NV12 buffers grow) */
bool must_grow(const Ctx* c, const Plan* plan) {
  const Layout& L = plan->lay;
  const size_t ndir = (size_t)(L.grid_w[0] >> 3) * (L.grid_h[0] >> 3) * 2;
  if (c->cap < (size_t)L.size || c->pool_alloc != c->pool_slots || c->cdef_dir_cap < ndir ||
      c->cells_cap < L.cells || c->edges_cap < L.edge_words)
    return true;
  if (plan->has_grain && c->grain_offs_cap < (size_t)plan->grain.stripes * plan->grain.blocks + 64) return true;
  if (c->decode_nv12_pitch > 0) {
    const size_t need = av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, c->decode_nv12_pitch);
    if (need > 0 && c->pack_cap < need) return true;
  }
  return false;
}

int copy_groups(const Layout& L) { return (int)((L.size * 2 + kCopyChunk - 1) / kCopyChunk); }

void default_tuning_impl(Tuning* t) {
  for (int k = 0; k < kTuneParams; ++k) t->v[k] = 16;
  t->v[kKMc] = 8;
  t->v[kKUpscale] = 8;
  t->v[kKLr] = 8;
  t->v[kKHash] = 8;
  t->v[kKGrainTemplate] = 1;
  for (int k : {kKCopyLrCurr, kKCopyCdefSrc, kKCopyUpSrc, kKCopyStore}) t->v[k] = 8;
  t->v[kTuneQueue] = 6;
  t->v[kTuneCdefBatch] = 1;
}

/* virtual workgroups of kernel k for one frame */
int kernel_groups(const FrameJob& fj, int k, int cdef_batch) {
  const FramePlanDev& p = fj.r.plan;
  const FilterArgs& f = fj.f;
  const Layout& L = fj.r.lay;
  switch (k) {
    case kKMc: return p.mc_tasks;
    case kKSpecial: return p.n_special;
    case kKItxInter: return itx_groups(p, 0, 0, 3);
    case kKItxIntra: return itx_groups(p, 1, 0, 3);
    case kKItxInterL: return itx_groups(p, 0, 3, kTxClasses);
    case kKItxIntraL: return itx_groups(p, 1, 3, kTxClasses);
    case kKDeblockV: return p.n_deblock_v > 0 ? (p.lines_v + 255) / 256 : 0;
    case kKDeblockH: return p.n_deblock_h > 0 ? (p.lines_h + 255) / 256 : 0;
    case kKCdefDir: return p.n_cdef > 0 ? (p.n_cdef_dir + kCdefDirPerGroup - 1) / kCdefDirPerGroup : 0;
    case kKCdefLuma: return p.n_cdef > 0 ? (p.n_cdef_luma + 4 * cdef_batch - 1) / (4 * cdef_batch) : 0;
    case kKCdefChroma: {
      const int nc = p.n_cdef - p.n_cdef_luma;
      const int per = cdef_batch * 256 / ((8 >> L.ss_x) * (8 >> L.ss_y));
      return nc > 0 ? (nc + per - 1) / per : 0;
    }
    case kKUpscale: return f.upscaled ? upscale_tiles(L, f.uses_lr ? 2 : 1) : 0;
    case kKLr: return f.uses_lr ? p.n_lr_tiles : 0;
    case kKGrainTemplate: return f.has_grain ? 1 : 0;
    case kKGrainApply: return f.has_grain ? grain_tiles(L) : 0;
    case kKNv12: return f.nv12 ? (L.height + ((L.height + 1) >> 1)) * ((f.nv12_pitch + 1023) / 1024) : 0;
    case kKHash: {
      const int64_t n = (int64_t)L.out_w[0] * L.out_h[0] + 2 * (int64_t)L.out_w[1] * L.out_h[1];
      return f.hash ? (int)((n + kHashChunk - 1) / kHashChunk) : 0;
    }
    case kKCopyLrCurr: return f.uses_lr || (f.upscaled && p.n_cdef > 0) ? copy_groups(L) : 0;
    case kKCopyCdefSrc: return p.n_cdef > 0 && !f.uses_lr && !f.upscaled ? copy_groups(L) : 0;
    case kKCopyUpSrc: return f.upscaled && !f.uses_lr && p.n_cdef == 0 ? copy_groups(L) : 0;
    case kKCopyStore: return f.store ? copy_groups(L) : 0;
    default: return 0;
  }
}

/* The queues of a batch merged into one hand-out order. Inside a frame the
 * order stays its own (superblock diagonals, a topological order); across
 * frames a diagonal is ranked by the diagonals still after it in its frame,
 * so the frame with the longest chain of diagonals starts first and all
 * frames finish their last diagonals together. */
void merge_queues(const std::vector<JobSpec>& js, uint32_t* out) {
  int span = 0;
  for (const JobSpec& j : js) span = std::max(span, (int)j.plan->key_off.size() - 1);
  size_t n = 0;
  for (int lvl = 0; lvl < span; ++lvl)
    for (size_t i = 0; i < js.size(); ++i) {
      const std::vector<int32_t>& ko = js[i].plan->key_off;
      const int keys = (int)ko.size() - 1;
      const int k = lvl - (span - keys);
      if (k < 0) continue;
      const int32_t* q = js[i].plan->queue_host.data();
      for (int32_t e = ko[k]; e < ko[k + 1]; ++e) out[n++] = (uint32_t)i << kQueueJobShift | (uint32_t)q[e];
    }
}

/* Prepares a batch and queues its launches, its read backs and its
 * completion event on the lead's stream; the slots it keeps are taken at
 * once, so the next batch may already name them. At most two batches are in
 * flight per lead (double-buffered tables): a third waits for the oldest. */
int batch_launch(Ctx* lead, std::vector<JobSpec>& js, double t0) {
  const int n = (int)js.size();
  for (JobSpec& j : js) {
    j.slot = -1;
    if (j.slot_out) *j.slot_out = -1;
  }
  if (n < 1 || n > kMaxBatch) return set_err(lead, AV1H_ERR_INVALID, "a batch holds 1 to %d frames", kMaxBatch);
  for (int i = 0; i < n; ++i) {
    if (js[i].c->device != lead->device)
      return set_err(lead, AV1H_ERR_INVALID, "the contexts of a batch must share one device");
    for (int k = 0; k < i; ++k)
      if (js[k].c == js[i].c) return set_err(lead, AV1H_ERR_INVALID, "a context appears twice in one batch");
  }
  for (JobSpec& j : js)
    if (j.c->inflight && j.c->inflight_lead != lead)
      return set_err(lead, AV1H_ERR_INVALID, "a frame of context %d's stream is in flight in another launcher's batch",
                     (int)(&j - js.data()));
  /* buffers that must grow are replaced only with nothing in flight */
  bool grow = false;
  for (JobSpec& j : js) grow = grow || must_grow(j.c, j.plan);
  while (grow && !lead->pending.empty()) {
    const int rc = batch_finish(lead);
    if (rc != AV1H_OK) return rc;
  }
  if (lead->pending.size() >= 2) {
    const int rc = batch_finish(lead);
    if (rc != AV1H_OK) return rc;
  }
  const int b = lead->next_buf;
  /* per frame buffers, slots, epochs */
  int64_t nq = 0;
  for (JobSpec& j : js) {
    Ctx* c = j.c;
    Plan* plan = j.plan;
    const Layout& L = plan->lay;
    int rc = prepare(c, L);
    if (rc != AV1H_OK) return c == lead ? rc : set_err(lead, rc, "%s", c->err.c_str());
    for (int sl : plan->slots)
      if (!c->used[sl])
        return set_err(lead, AV1H_ERR_INVALID, "a block names reference slot %d, which holds no picture", sl);
    if (j.keep) {
      if (j.reserved >= 0) {
        if (j.reserved >= c->pool_slots) return set_err(lead, AV1H_ERR_INVALID, "invalid reference slot");
        j.slot = j.reserved;
      } else {
        j.slot = alloc_slot(c);
        if (j.slot < 0) return set_err(lead, AV1H_ERR_SLOTS, "reference pool exhausted");
      }
      if (j.slot_out) *j.slot_out = j.slot;
    }
    if (plan->has_grain) {
      const size_t need = (size_t)plan->grain.stripes * plan->grain.blocks + 64;
      if (c->grain_offs_cap < need) {
        HIP_OK(lead, hipStreamSynchronize(c->stream));
        if (c->grain_offs) (void)hipFree(c->grain_offs);
        c->grain_offs = nullptr;
        HIP_OK(lead, hipMalloc(&c->grain_offs, need));
        c->grain_offs_cap = need;
      }
    }
    c->nv12_current = false;
    if (c->decode_nv12_pitch > 0 &&
        av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, c->decode_nv12_pitch) > 0) {
      rc = ensure_pack(c, av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, c->decode_nv12_pitch));
      if (rc != AV1H_OK) return c == lead ? rc : set_err(lead, rc, "%s", c->err.c_str());
    }
    if ((int64_t)plan->queue_host.size() >= (1 << kQueueJobShift))
      return set_err(lead, AV1H_ERR_INVALID, "too many intra blocks in one frame");
    nq += plan->dev.n_queue;
  }
  hipStream_t s = lead->stream;
  /* the frame table */
  std::vector<FrameJob> jobs(n);
  bool any_hash = false;
  for (int i = 0; i < n; ++i) {
    Ctx* c = js[i].c;
    Plan* plan = js[i].plan;
    const Layout& L = plan->lay;
    ++c->epoch;
    ReconArgs& ra = jobs[i].r;
    ra.plan = plan->dev;
    ra.lay = L;
    ra.frame = c->frame;
    ra.pool = c->pool;
    ra.slot_size = L.size;
    ra.cells = c->cells;
    ra.edges = c->edges;
    ra.epoch = c->epoch;
    ra.res = reinterpret_cast<int16_t*>(c->scratch); /* free until the CDEF snapshot */
    ra.maxv = (1 << L.bit_depth) - 1;
    ra.edge_filter = plan->hdr.enable_intra_edge_filter;
    ra.spin_limit = c->spin_limit;
    FilterArgs& fa = jobs[i].f;
    fa.plan = plan->dev;
    fa.lay = L;
    fa.frame = c->frame;
    fa.scratch = c->scratch;
    fa.lr_curr = c->lr_curr;
    fa.lr_cdef = c->lr_cdef;
    fa.lr_prev_up = c->lr_prev_up;
    /* The pictures between the filters, with no copy that a kernel's second
     * store can replace: the deblocked snapshot (kKCopyLrCurr) goes to
     * lr_curr, and also to scratch when superres follows (CDEF writes its
     * output there, superres reads it) or to lr_cdef when restoration reads
     * the CDEF output at the coded size (CDEF writes there too); the
     * upscaled frame also goes to lr_cdef when restoration follows. Without
     * restoration or superres, CDEF reads a snapshot in scratch
     * (kKCopyCdefSrc) and writes the frame; superres without CDEF and
     * restoration reads a copy of the frame in scratch (kKCopyUpSrc). */
    const bool snap = plan->uses_lr || plan->upscaled;
    fa.cdef_src = snap ? c->lr_curr : c->scratch;
    fa.cdef_dst = plan->upscaled ? c->scratch : c->frame;
    fa.cdef_dst2 = plan->uses_lr && !plan->upscaled ? c->lr_cdef : nullptr;
    fa.snap2 = plan->upscaled ? c->scratch : (plan->uses_lr ? c->lr_cdef : nullptr);
    fa.up_dst2 = plan->uses_lr ? c->lr_cdef : nullptr;
    fa.cdef_dir = c->cdef_dir;
    fa.cdef_dir_w = L.grid_w[0] >> 3;
    fa.sharpness = plan->hdr.sharpness;
    fa.maxv = ra.maxv;
    for (int p = 0; p < 3; ++p) superres_params(L.plane_w[p], L.out_w[p], &fa.up_step[p], &fa.up_initial[p]);
    fa.upscaled = plan->upscaled;
    fa.uses_lr = plan->uses_lr;
    fa.has_grain = plan->has_grain;
    fa.grain_tmpl = c->grain_tmpl;
    fa.grain_offs = c->grain_offs;
    fa.store = js[i].slot >= 0 ? c->pool + (size_t)js[i].slot * L.size : nullptr;
    const bool nv12 = c->decode_nv12_pitch > 0 &&
                      av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, c->decode_nv12_pitch) > 0;
    fa.nv12 = nv12 ? c->pack : nullptr;
    fa.nv12_pitch = nv12 ? (int32_t)c->decode_nv12_pitch : 0;
    fa.hash = c->hash_on ? lead->hash_dev + i : nullptr;
    any_hash = any_hash || c->hash_on;
  }
  std::vector<int32_t> prefix((size_t)kKernels * (n + 1) + 1, 0);
  for (int k = 0; k < kKernels; ++k)
    for (int i = 0; i < n; ++i)
      prefix[(size_t)k * (n + 1) + i + 1] =
          prefix[(size_t)k * (n + 1) + i] + kernel_groups(jobs[i], k, lead->tune.v[kTuneCdefBatch]);
  prefix[(size_t)kKernels * (n + 1)] = (int32_t)nq; /* the queue's entry count */
  auto total_of = [&](int k) { return prefix[(size_t)k * (n + 1) + n]; };
  /* upload: [jobs][prefix][merged queue] */
  const size_t o_prefix = align256(sizeof(FrameJob) * n);
  const size_t o_queue = o_prefix + align256(prefix.size() * 4);
  const size_t bytes_tab = o_queue + (n > 1 ? (size_t)nq * 4 : 0);
  if (lead->batch_cap[b] < bytes_tab) {
    /* the other set may be in flight: wait for it too */
    HIP_OK(lead, hipStreamSynchronize(s));
    drop_graphs(lead);
    if (lead->batch_dev[b]) (void)hipFree(lead->batch_dev[b]);
    if (lead->batch_pin[b]) (void)hipHostFree(lead->batch_pin[b]);
    lead->batch_dev[b] = lead->batch_pin[b] = nullptr;
    lead->batch_cap[b] = 0;
    const size_t cap = std::max(bytes_tab + bytes_tab / 4, (size_t)65536);
    HIP_OK(lead, hipMalloc(&lead->batch_dev[b], cap));
    HIP_OK(lead, hipHostMalloc(&lead->batch_pin[b], kTableHead + cap, hipHostMallocDefault));
    lead->batch_cap[b] = cap;
  }
  for (int k = 0; k < 4; ++k)
    if (!lead->bev[b][k])
      HIP_OK(lead, hipEventCreateWithFlags(&lead->bev[b][k], hipEventDefault));
  if (!lead->rb_pin[b]) HIP_OK(lead, hipHostMalloc(&lead->rb_pin[b], 64 + kMaxBatch * 8, hipHostMallocDefault));
  uint32_t* rb_err = static_cast<uint32_t*>(lead->rb_pin[b]);
  uint64_t* rb_hash = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(lead->rb_pin[b]) + 64);
  /* [byte count][..][the table at kTableHead]: one frame's table, a few KB,
   * is fetched by the batch's first launch; a larger one (several frames'
   * jobs, a merged queue of up to MBs) is copied ahead of the batch, which
   * moves it in larger transfers than the fetch's system-scope 8 byte loads
   * (in the fetch: 90 us per round of the 16-stream mix) */
  *static_cast<uint64_t*>(lead->batch_pin[b]) = bytes_tab;
  uint8_t* pin = static_cast<uint8_t*>(lead->batch_pin[b]) + kTableHead;
  std::memcpy(pin, jobs.data(), sizeof(FrameJob) * n);
  std::memcpy(pin + o_prefix, prefix.data(), prefix.size() * 4);
  if (n > 1) merge_queues(js, reinterpret_cast<uint32_t*>(pin + o_queue));
  /* everything queued on the other contexts' streams (buffer clears) comes
   * first; the plans' uploads ended before av1h_plan_frame returned */
  for (JobSpec& j : js)
    if (j.c != lead && j.c->dirty) {
      HIP_OK(lead, hipEventRecord(j.c->ev[3], j.c->stream));
      HIP_OK(lead, hipStreamWaitEvent(s, j.c->ev[3], 0));
      j.c->dirty = false;
    }
  uint8_t* tab = static_cast<uint8_t*>(lead->batch_dev[b]);
  if (n > 1) HIP_OK(lead, hipMemcpyAsync(tab, pin, bytes_tab, hipMemcpyHostToDevice, s));
  BatchDev B;
  B.jobs = reinterpret_cast<const FrameJob*>(tab);
  B.prefix = reinterpret_cast<const int32_t*>(tab + o_prefix);
  B.njobs = n;
  B.queue = n > 1 ? reinterpret_cast<const uint32_t*>(tab + o_queue) : nullptr;
  B.queue_counter = lead->counters;
  B.error = lead->counters + 1;
  B.trace = nullptr;
  grain_jump_matrices(B.jump82, B.jump44);
  if (lead->trace_on && nq > 0) {
    const size_t need = (size_t)nq * 4;
    if (lead->trace_cap < need) {
      HIP_OK(lead, hipStreamSynchronize(s));
      if (lead->trace) (void)hipFree(lead->trace);
      lead->trace = nullptr;
      lead->trace_cap = 0;
      HIP_OK(lead, hipMalloc(&lead->trace, need * 8));
      lead->trace_cap = need;
    }
    B.trace = lead->trace;
  }
  lead->trace_n = lead->trace_on ? (int)nq : 0;
  lead->trace_plan = lead->trace_on && n == 1 ? js[0].plan : nullptr;

  int launches = 0;
  /* Captured launches: every kernel with work in the batch at its full grid,
   * counts read on the device, so one graph per frame count and set of
   * kernels with work serves every batch of that shape (a kernel without
   * work would still dispatch its whole grid). Stage timing and the queue
   * trace take the eager path. */
  const bool use_graph = lead->graphs && !lead->timing && !lead->trace_on;
  uint64_t live = 0; /* bit k: kernel k has work; bit kKernels: the queue */
  for (int k = 0; k < kKernels; ++k)
    if (total_of(k) > 0) live |= 1ull << k;
  if (nq > 0) live |= 1ull << kKernels;
  if (any_hash) live |= 1ull << kKHash;
  if (use_graph && !lead->graph_sets) live = ~0ull >> (63 - kKernels); /* every kernel, the queue */
  static_assert(kKernels < 64, "kernel set mask");
  auto has = [&](int k) { return ((live >> k) & 1) != 0; };
  auto run = [&](int k, hipStream_t st) -> hipError_t {
    /* CDEF's luma launch also filters the chroma blocks */
    if (k == kKCdefChroma) return hipSuccess;
    const bool work = k == kKCdefLuma ? has(kKCdefLuma) || has(kKCdefChroma) : has(k);
    const int total = use_graph ? (work ? kAnyTotal : 0)
                                : total_of(k) + (k == kKCdefLuma ? total_of(kKCdefChroma) : 0);
    if (total <= 0) return hipSuccess;
    ++launches;
    const int cap = lead->cus * lead->tune.v[k];
    switch (k) {
      case kKMc:
      case kKSpecial:
      case kKItxInter:
      case kKItxIntra:
      case kKItxInterL:
      case kKItxIntraL: return launch_recon_kernel(B, k, total, cap, st);
      case kKGrainTemplate:
      case kKGrainApply: return launch_grain_kernel(B, k, total, cap, st);
      default: return launch_filter_kernel(B, k, total, cap, lead->tune.v[kTuneCdefBatch], st);
    }
  };
  auto stage = [&](int st, std::initializer_list<int> ks) -> hipError_t {
    bool any = false;
    for (int k : ks) any = any || has(k);
    if (!any) return hipSuccess;
    if (lead->timing) (void)hipEventRecord(lead->sev[2 * st], s);
    for (int k : ks) {
      const hipError_t e = run(k, s);
      if (e != hipSuccess) return e;
    }
    if (lead->timing) {
      (void)hipEventRecord(lead->sev[2 * st + 1], s);
      lead->stage_hit[st] = true;
    }
    return hipSuccess;
  };
  /* The second stream takes what does not depend on the prediction: the
   * transforms of intra blocks (the queue waits for them) and the grain
   * templates (the grain blend waits for them). */
  const bool side_itx = has(kKItxIntra) || has(kKItxIntraL);
  const bool side_grain = has(kKGrainTemplate);
  auto sequence = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
#define SEQ(call)        \
  do {                   \
    e = (call);          \
    if (e != hipSuccess) return e; \
  } while (0)
    if (n == 1) SEQ(launch_fetch_table(lead->batch_pin[b], tab, n, s));
    if (side_itx || side_grain) {
      SEQ(hipEventRecord(lead->ev_side_go, s));
      SEQ(hipStreamWaitEvent(lead->side, lead->ev_side_go, 0));
    }
    if (side_itx) {
      SEQ(run(kKItxIntraL, lead->side));
      SEQ(run(kKItxIntra, lead->side));
      SEQ(hipEventRecord(lead->ev_itx_done, lead->side));
    }
    if (side_grain) {
      SEQ(run(kKGrainTemplate, lead->side));
      SEQ(hipEventRecord(lead->ev_side_done, lead->side));
    }
    if (!use_graph) SEQ(hipEventRecord(lead->bev[b][0], s));
    SEQ(stage(AV1H_STAGE_INTER, {kKMc, kKSpecial, kKItxInterL, kKItxInter}));
    if (side_itx) SEQ(hipStreamWaitEvent(s, lead->ev_itx_done, 0));
    if (has(kKernels)) {
      if (lead->timing) (void)hipEventRecord(lead->sev[2 * AV1H_STAGE_QUEUE], s);
      SEQ(launch_queue(B, use_graph ? -1 : (int)nq, lead->cus * lead->tune.v[kTuneQueue], s, &launches));
      if (lead->timing) {
        (void)hipEventRecord(lead->sev[2 * AV1H_STAGE_QUEUE + 1], s);
        lead->stage_hit[AV1H_STAGE_QUEUE] = true;
      }
    }
    if (!use_graph) SEQ(hipEventRecord(lead->bev[b][1], s));
    SEQ(stage(AV1H_STAGE_DEBLOCK, {kKDeblockV, kKDeblockH}));
    SEQ(run(kKCopyLrCurr, s));
    SEQ(run(kKCopyCdefSrc, s));
    SEQ(stage(AV1H_STAGE_CDEF, {kKCdefDir, kKCdefLuma, kKCdefChroma}));
    SEQ(run(kKCopyUpSrc, s));
    SEQ(stage(AV1H_STAGE_SUPERRES, {kKUpscale}));
    SEQ(stage(AV1H_STAGE_LR, {kKLr}));
    SEQ(stage(AV1H_STAGE_STORE, {kKCopyStore}));
    if (side_grain) SEQ(hipStreamWaitEvent(s, lead->ev_side_done, 0));
    SEQ(stage(AV1H_STAGE_GRAIN, {kKGrainApply}));
    SEQ(run(kKNv12, s));
    if (has(kKHash)) {
      SEQ(launch_zero64(lead->hash_dev, n, s));
      SEQ(run(kKHash, s));
    }
    SEQ(launch_readback(lead->counters + 1, has(kKHash) ? lead->hash_dev : nullptr, n, rb_err, rb_hash, s));
#undef SEQ
    return e;
  };
  for (bool& h : lead->stage_hit) h = false;
  if (use_graph) {
    hipGraphExec_t exec = nullptr;
    const uint64_t key = (uint64_t)b << 62 | live << 8 | (uint64_t)n; /* n <= 128 */
    for (auto& g : lead->graph_exec)
      if (g.first == key) exec = g.second;
    if (!exec) {
      HIP_OK(lead, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const hipError_t e = sequence();
      hipGraph_t graph = nullptr;
      const hipError_t e2 = hipStreamEndCapture(s, &graph);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_err(lead, AV1H_ERR_DEVICE, "capturing the batch launches: %s",
                       hipGetErrorString(e != hipSuccess ? e : e2));
      }
      const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e3 != hipSuccess)
        return set_err(lead, AV1H_ERR_DEVICE, "instantiating the batch graph: %s", hipGetErrorString(e3));
      if (lead->graph_exec.size() >= 64) {
        (void)hipGraphExecDestroy(lead->graph_exec.front().second);
        lead->graph_exec.erase(lead->graph_exec.begin());
      }
      lead->graph_exec.emplace_back(key, exec);
    }
    HIP_OK(lead, hipEventRecord(lead->bev[b][0], s));
    HIP_OK(lead, hipGraphLaunch(exec, s));
  } else {
    HIP_OK(lead, sequence());
    HIP_OK(lead, hipEventRecord(lead->bev[b][2], s));
  }
  const double t_sub = now_ms();
  /* read back */
  for (JobSpec& j : js) {
    if (!j.out) continue;
    Ctx* c = j.c;
    const size_t host_need = (size_t)j.plan->lay.size * 2;
    if (c->host_pin_cap < host_need) {
      HIP_OK(lead, hipStreamSynchronize(s));
      HIP_OK(lead, hipStreamSynchronize(c->stream));
      if (c->host_pin) (void)hipHostFree(c->host_pin);
      c->host_pin = nullptr;
      c->host_pin_cap = 0;
      HIP_OK(lead, hipHostMalloc(&c->host_pin, host_need, hipHostMallocDefault));
      c->host_pin_cap = host_need;
    }
    HIP_OK(lead, hipMemcpyAsync(c->host_pin, c->frame, host_need, hipMemcpyDeviceToHost, s));
  }
  HIP_OK(lead, hipEventRecord(lead->bev[b][3], s));
  /* the kept slots are taken now; their pictures count once the batch is done */
  for (JobSpec& j : js)
    if (j.slot >= 0) {
      j.c->used[j.slot] = true;
      j.c->has_pic[j.slot] = false;
      if (j.c->last_slot == j.slot) j.c->last_slot = -1;
    }
  for (JobSpec& j : js) {
    ++j.c->inflight;
    j.c->inflight_lead = lead;
  }
  BatchTicket& tk = lead->ticket[b];
  tk.js = js;
  tk.nv12.assign((size_t)n, 0);
  for (int i = 0; i < n; ++i) tk.nv12[(size_t)i] = jobs[(size_t)i].f.nv12 != nullptr;
  tk.launches = launches;
  tk.any_hash = any_hash;
  tk.graph = use_graph;
  tk.t0 = t0;
  tk.t_sub = t_sub;
  tk.live = true;
  lead->pending.push_back(b);
  lead->next_buf = 1 - b;
  return AV1H_OK;
}

/* Waits for the oldest launched batch of this lead and takes its results:
 * the error word, checksums, stage times, slot bookkeeping, read backs. */
int batch_finish(Ctx* lead) {
  if (lead->pending.empty()) return AV1H_OK;
  const int b = lead->pending.front();
  lead->pending.pop_front();
  BatchTicket& tk = lead->ticket[b];
  tk.live = false;
  std::vector<JobSpec>& js = tk.js;
  for (JobSpec& j : js)
    if (--j.c->inflight == 0) j.c->inflight_lead = nullptr;
  hipStream_t s = lead->stream;
  const uint32_t* err_pin = static_cast<const uint32_t*>(lead->rb_pin[b]);
  const uint64_t* hashes = reinterpret_cast<const uint64_t*>(static_cast<uint8_t*>(lead->rb_pin[b]) + 64);
  HIP_OK(lead, hipEventSynchronize(lead->bev[b][3]));
  const uint32_t err_word = *err_pin;
  const double t_done = now_ms();
  if (err_word) {
    HIP_OK(lead, hipMemsetAsync(lead->counters + 1, 0, 4, s));
    HIP_OK(lead, hipStreamSynchronize(s));
    for (JobSpec& j : js) {
      j.c->last_slot = -1;
      j.c->have_frame = false;
      if (j.slot >= 0) {
        /* a slot this batch took is given back; a reserved one stays the
         * caller's, holding no picture */
        if (j.reserved < 0) j.c->used[j.slot] = false;
        j.c->has_pic[j.slot] = false;
      }
    }
    return set_err(lead, AV1H_ERR_INVALID,
                   "a block waited too long for a neighbour: the availability flags name blocks "
                   "that come later in coding order");
  }
  const int launches = tk.launches;
  const double t0 = tk.t0, t_sub = tk.t_sub;
  float recon = 0, filt = 0;
  if (tk.graph) {
    /* one captured launch: its whole time counts as reconstruction */
    (void)hipEventElapsedTime(&recon, lead->bev[b][0], lead->bev[b][3]);
  } else {
    (void)hipEventElapsedTime(&recon, lead->bev[b][0], lead->bev[b][1]);
    (void)hipEventElapsedTime(&filt, lead->bev[b][1], lead->bev[b][2]);
  }
  if (lead->timing)
    for (int st = 0; st < AV1H_STAGES; ++st)
      if (lead->stage_hit[st]) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, lead->sev[2 * st], lead->sev[2 * st + 1]) == hipSuccess) {
          lead->stage_ms[st] += ms;
          ++lead->stage_n[st];
          for (JobSpec& j : js) lead->stage_bytes[st] += j.plan->alg_bytes[st];
        }
      }
  for (JobSpec& j : js) {
    Ctx* c = j.c;
    const Layout& L = j.plan->lay;
    c->have_frame = true;
    if (j.slot >= 0) {
      c->has_pic[j.slot] = true;
      c->slot_lay[j.slot] = L;
    }
    c->last_slot = j.slot;
    c->last_grain = j.plan->has_grain;
    c->nv12_current = tk.nv12[(size_t)(&j - js.data())] != 0;
    c->hash_valid = c->hash_on;
    if (c->hash_on) c->last_hash = hashes[&j - js.data()];
    c->nv12_pitch_done = c->decode_nv12_pitch;
    if (j.out) {
      const uint16_t* h = static_cast<const uint16_t*>(c->host_pin);
      for (int p = 0; p < 3; ++p)
        for (int y = 0; y < L.out_h[p]; ++y)
          std::memcpy(j.out->plane[p] + (size_t)y * j.out->stride[p], h + L.plane_off[p] + (size_t)y * L.stride[p],
                      (size_t)L.out_w[p] * 2);
    }
    if (j.info) {
      j.info->width = L.width;
      j.info->height = L.height;
      j.info->bit_depth = L.bit_depth;
      j.info->ss_x = L.ss_x;
      j.info->ss_y = L.ss_y;
    }
    c->stats.blocks = j.plan->blocks;
    /* the queue's dependency levels: the counterpart of the reference
     * scheduler's wave count */
    c->stats.waves = (int32_t)j.plan->key_off_static.size() - 1;
    c->stats.kernel_launches = launches;
    c->stats.device_recon_ms = recon;
    c->stats.device_filter_ms = filt;
    c->stats.reconstruct_ms = recon;
    c->stats.filter_ms = filt;
    c->stats.readback_ms = t_done - t_sub;
    c->stats.total_ms = t_done - t0;
    if (j.slot_out) *j.slot_out = j.slot;
  }
  return AV1H_OK;
}

/* A batch launched and waited for (with any launched before it). */
int run_batch(Ctx* lead, std::vector<JobSpec>& js, double t0) {
  int rc = batch_launch(lead, js, t0);
  while (rc == AV1H_OK && !lead->pending.empty()) rc = batch_finish(lead);
  return rc;
}

int run_plan(Ctx* c, Plan* plan, int keep, int reserved, const av1h_planes* out, av1h_frame_info* info,
             int* slot_out, double t0) {
  std::vector<JobSpec> js(1);
  js[0] = JobSpec{c, plan, keep, reserved, out, info, slot_out, -1};
  return run_batch(c, js, t0);
}

/* the picture a slot names (the reference's gather rule,
 * cpu_backend.cpp:557-587) */
const uint16_t* gather(Ctx* c, int slot, Layout* L) {
  /* frames in flight on another launcher's stream: nothing this context
   * queues is ordered behind them */
  if (c->inflight && c->inflight_lead != c) {
    set_err(c, AV1H_ERR_INVALID, "a batch launched by another context holds this context's frames: finish it first");
    return nullptr;
  }
  if (slot >= 0) {
    if (slot >= c->pool_slots || !c->used[slot] || !c->has_pic[slot]) {
      set_err(c, AV1H_ERR_INVALID, c->inflight && slot < c->pool_slots && c->used[slot]
                                       ? "the frame of this slot is still in flight: finish it first"
                                       : "invalid reference slot");
      return nullptr;
    }
    /* an older slot: its picture is final and no frame in flight writes it
     * (a frame in flight took its own slot at launch), so a pack or copy
     * queued on this context's stream behind the frames in flight reads it
     * as it is. So is the most
     * recent one's when its frame had no film grain (the slot then holds
     * the output picture too) */
    if (slot != c->last_slot || (c->inflight && !c->last_grain)) {
      *L = c->slot_lay[slot];
      return c->pool + (size_t)slot * c->lay.size;
    }
  }
  if (c->inflight) {
    set_err(c, AV1H_ERR_INVALID, "a launched batch holds this context's frame: finish it first (av1h_batch_finish)");
    return nullptr;
  }
  if (!c->have_frame) {
    set_err(c, AV1H_ERR_INVALID, "no frame to convert");
    return nullptr;
  }
  *L = c->lay;
  return c->frame;
}

}  // namespace

/* ============================================================== C ABI === */

struct av1h_ctx {
  Ctx c;
};

extern "C" {

int av1h_create(int device_index, av1h_ctx** out) {
  if (!out) return AV1H_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device_index || device_index < 0) return AV1H_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) return AV1H_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return AV1H_ERR_NO_DEVICE;
  av1h_ctx* h = new av1h_ctx;
  Ctx* c = &h->c;
  c->device = device_index;
  c->prop = prop;
  c->cus = prop.multiProcessorCount;
  default_tuning_impl(&c->tune);
  std::memset(&c->stats, 0, sizeof(c->stats));
  if (hipSetDevice(device_index) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->counters, 64) != hipSuccess || hipMemset(c->counters, 0, 64) != hipSuccess ||
      hipMalloc(&c->hash_dev, kMaxBatch * 8) != hipSuccess ||
      hipMalloc(&c->grain_tmpl, 3 * 82 * 73 * 2) != hipSuccess) {
    delete h;
    return AV1H_ERR_DEVICE;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete h;
      return AV1H_ERR_DEVICE;
    }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_side_go, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_side_done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_itx_done, hipEventDisableTiming) != hipSuccess) {
    delete h;
    return AV1H_ERR_DEVICE;
  }
  *out = h;
  return AV1H_OK;
}

void av1h_destroy(av1h_ctx* h) {
  if (!h) return;
  Ctx* c = &h->c;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  if (c->hash_dev) (void)hipFree(c->hash_dev);
  for (void* p : {(void*)c->frame, (void*)c->scratch, (void*)c->lr_curr, (void*)c->lr_cdef, (void*)c->lr_prev_up,
                  (void*)c->pool, (void*)c->cells, (void*)c->edges, (void*)c->counters, (void*)c->grain_tmpl, (void*)c->grain_offs,
                  (void*)c->pack, (void*)c->cdef_dir, (void*)c->trace})
    if (p) (void)hipFree(p);
  if (c->host_pin) (void)hipHostFree(c->host_pin);
  while (!c->pending.empty()) (void)batch_finish(c);
  for (int b = 0; b < 2; ++b) {
    if (c->batch_dev[b]) (void)hipFree(c->batch_dev[b]);
    if (c->batch_pin[b]) (void)hipHostFree(c->batch_pin[b]);
    if (c->rb_pin[b]) (void)hipHostFree(c->rb_pin[b]);
    for (hipEvent_t e : c->bev[b])
      if (e) (void)hipEventDestroy(e);
  }
  free_plan(&c->scratch_plan);
  for (PackEntry& r : c->ring) {
    if (r.done) (void)hipEventSynchronize(r.done);
    if (r.done) (void)hipEventDestroy(r.done);
    if (r.dev) (void)hipFree(r.dev);
    if (r.pin) (void)hipHostFree(r.pin);
  }
  for (auto& b : c->plan_bufs) (void)hipFree(b.second);
  c->plan_bufs.clear();
  for (Stager* st : c->stagers) {
    if (st->pin) (void)hipHostFree(st->pin);
    delete st;
  }
  c->stagers.clear();
  for (hipStream_t u : c->uploads) (void)hipStreamDestroy(u);
  c->uploads.clear();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->sev)
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->ev_side_go) (void)hipEventDestroy(c->ev_side_go);
  if (c->ev_side_done) (void)hipEventDestroy(c->ev_side_done);
  if (c->ev_itx_done) (void)hipEventDestroy(c->ev_itx_done);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete h;
}

const char* av1h_last_error(const av1h_ctx* h) { return h ? h->c.err.c_str() : "no context"; }

namespace {
void describe(const hipDeviceProp_t& p, av1h_device_info* out) {
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->l2_bytes = p.l2CacheSize;
  out->global_mem = (int64_t)p.totalGlobalMem;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_width = p.memoryBusWidth;
}
}  // namespace

int av1h_get_device_info(av1h_ctx* h, av1h_device_info* out) {
  if (!h || !out) return AV1H_ERR_INVALID;
  describe(h->c.prop, out);
  return AV1H_OK;
}

int av1h_describe_device(int device_index, av1h_device_info* out) {
  if (!out) return AV1H_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device_index < 0 || device_index >= n) return AV1H_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) return AV1H_ERR_NO_DEVICE;
  describe(prop, out);
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? AV1H_OK : AV1H_ERR_UNSUPPORTED;
}

namespace {
/* Planning state, whose arrays keep their capacity from frame to frame,
 * pooled across the planning threads: a frame takes the pooled state sized
 * for the fewest blocks that still holds its own, else the largest, so a
 * key frame finds the arrays an earlier key frame grew (and whose pages it
 * touched) on whichever thread planned it. Grown on first touch, a 1080p
 * key frame's arrays cost about 3.5 ms more planning. */
std::mutex s_host_mu;
std::vector<Host*> s_host_free;

Host* take_host(int64_t blocks) {
  std::lock_guard<std::mutex> lk(s_host_mu);
  if (s_host_free.empty()) return new Host;
  size_t best = 0;
  for (size_t i = 1; i < s_host_free.size(); ++i) {
    const int64_t cb = (int64_t)s_host_free[best]->blocks.capacity(), ci = (int64_t)s_host_free[i]->blocks.capacity();
    const bool fits_i = ci >= blocks, fits_b = cb >= blocks;
    if ((fits_i && (!fits_b || ci < cb)) || (!fits_i && !fits_b && ci > cb)) best = i;
  }
  Host* h = s_host_free[best];
  s_host_free.erase(s_host_free.begin() + (ptrdiff_t)best);
  return h;
}

void give_host(Host* h) {
  std::lock_guard<std::mutex> lk(s_host_mu);
  s_host_free.push_back(h);
}
}  // namespace

int av1h_decode(av1h_ctx* h, const av1h_frame_desc* f, int keep, int reserved, const av1h_planes* out,
                av1h_frame_info* info, int* slot_out) {
  if (!h || !f) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  const double t0 = now_ms();
  Host* host = take_host(f->num_blocks);
  int rc = flatten(*f, &c->scratch_plan, host, &c->err);
  if (rc == AV1H_OK) rc = upload(c, &c->scratch_plan, *host);
  give_host(host);
  if (rc != AV1H_OK) return rc;
  return run_plan(c, &c->scratch_plan, keep, reserved, out, info, slot_out, t0);
}

int av1h_decode_wire(av1h_ctx* h, const uint8_t* wire, size_t len, int keep, int reserved, const av1h_planes* out,
                     av1h_frame_info* info, int* slot_out) {
  av1h_frame_desc d;
  if (!h) return AV1H_ERR_INVALID;
  if (av1h_wire_parse(wire, len, &d) != AV1H_OK) {
    std::lock_guard<std::mutex> lk(h->c.mu);
    h->c.err = "not a wire frame";
    return AV1H_ERR_INVALID;
  }
  return av1h_decode(h, &d, keep, reserved, out, info, slot_out);
}

/* Pre-planning (reference Decoder::plan_frame / decode_reference_planned,
 * av1gpu.hpp:670-712): the syntax is flattened and uploaded into device
 * memory owned by the plan, so the decode itself starts from HBM. Safe to
 * call from several threads at once and while the context decodes: it takes
 * the context's lock only to report an error. */
int av1h_plan_frame(av1h_ctx* h, const av1h_frame_desc* f, void** plan_out) {
  if (!h || !f || !plan_out) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  (void)hipSetDevice(c->device);
  Plan* p = new Plan;
  Host* host = take_host(f->num_blocks);
  std::string err;
  const double t0 = now_ms();
  int rc = flatten(*f, p, host, &err);
  const double t1 = now_ms();
  if (rc == AV1H_OK) rc = upload_plan(c, p, *host, &err);
  give_host(host);
  p->flatten_ms = t1 - t0;
  p->upload_ms = now_ms() - t1;
  if (rc != AV1H_OK) {
    give_buffer(c, p->dbuf, p->dbytes);
    delete p;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err = err;
    return rc;
  }
  *plan_out = p;
  return AV1H_OK;
}

int av1h_plan_frame_wire(av1h_ctx* h, const uint8_t* wire, size_t len, void** plan_out) {
  av1h_frame_desc d;
  if (!h) return AV1H_ERR_INVALID;
  if (av1h_wire_parse(wire, len, &d) != AV1H_OK) {
    std::lock_guard<std::mutex> lk(h->c.mu);
    h->c.err = "not a wire frame";
    return AV1H_ERR_INVALID;
  }
  return av1h_plan_frame(h, &d, plan_out);
}

int av1h_decode_planned(av1h_ctx* h, void* plan, int keep, int reserved, const av1h_planes* out,
                        av1h_frame_info* info, int* slot_out) {
  if (!h || !plan) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  return run_plan(c, static_cast<Plan*>(plan), keep, reserved, out, info, slot_out, now_ms());
}

void av1h_discard_plan(av1h_ctx* h, void* plan) {
  if (!plan) return;
  Plan* p = static_cast<Plan*>(plan);
  if (h) (void)hipSetDevice(h->c.device);
  if (h) {
    give_buffer(&h->c, p->dbuf, p->dbytes);
    p->dbuf = nullptr;
    p->dbytes = 0;
  }
  free_plan(p);
  delete p;
}

int av1h_reserve_reference(av1h_ctx* h) {
  if (!h) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->c.mu);
  const int s = alloc_slot(&h->c);
  if (s < 0) return set_err(&h->c, AV1H_ERR_SLOTS, "reference pool exhausted");
  h->c.used[s] = true;
  return s;
}

void av1h_release_reference(av1h_ctx* h, int slot) {
  if (!h) return;
  std::lock_guard<std::mutex> lk(h->c.mu);
  if (slot >= 0 && slot < h->c.pool_slots) h->c.used[slot] = h->c.has_pic[slot] = false;
}

size_t av1h_nv12_size(int width, int height, int ss_x, int ss_y, size_t pitch) {
  /* reference nv12_buffer_size, src/core/factory.cpp:236-243 */
  if (width <= 0 || height <= 0 || ss_x != 1 || ss_y != 1) return 0;
  const size_t crow = 2 * (((size_t)width + 1) / 2);
  if ((pitch & 1) || pitch < crow) return 0;
  return pitch * ((size_t)height + ((size_t)height + 1) / 2);
}

int av1h_to_nv12(av1h_ctx* h, int slot, size_t pitch, uint8_t* out, size_t out_bytes) {
  if (!h) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Layout L;
  const uint16_t* src = gather(c, slot, &L);
  if (!src) return AV1H_ERR_INVALID;
  const size_t need = av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, pitch);
  if (need == 0)
    return set_err(c, AV1H_ERR_INVALID, "NV12 output needs a 4:2:0 frame and an even pitch at least as wide");
  if (out && out_bytes < need) return set_err(c, AV1H_ERR_INVALID, "output buffer too small");
  int rc = ensure_pack(c, need);
  if (rc != AV1H_OK) return rc;
  /* the decode packed this picture already */
  const bool packed = c->nv12_current && pitch == c->nv12_pitch_done && src == c->frame;
  if (!packed) {
    c->nv12_current = false;
    HIP_OK(c, launch_nv12(L, src, c->pack, (int)pitch, c->stream));
  }
  if (out) HIP_OK(c, hipMemcpyAsync(out, c->pack, need, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return AV1H_OK;
}

const uint8_t* av1h_nv12_device_buffer(const av1h_ctx* h) { return h ? h->c.pack : nullptr; }

/* Streamed NV12 pack (reference Decoder::nv12_pack_begin / wait / release,
 * av1gpu.hpp:748-773, the CUDA backend's 3-entry ring, cuda_backend.cpp:
 * 1807-1998). The pack kernel and the copy into the entry's pinned buffer go
 * on the decode stream, after the decode that made the picture and before any
 * later one that reuses the slot; the decoding thread returns at once and the
 * thread that shows the frame waits for the copy. */
int av1h_nv12_pack_begin(av1h_ctx* h, int slot, size_t pitch) {
  if (!h) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Layout L;
  const uint16_t* src = gather(c, slot, &L);
  if (!src) return AV1H_ERR_INVALID;
  const size_t need = av1h_nv12_size(L.width, L.height, L.ss_x, L.ss_y, pitch);
  if (need == 0)
    return set_err(c, AV1H_ERR_INVALID, "NV12 output needs a 4:2:0 frame and an even pitch at least as wide");
  PackEntry* r = nullptr;
  for (PackEntry& e : c->ring)
    if (!e.busy) {
      r = &e;
      break;
    }
  if (!r) return set_err(c, AV1H_ERR_UNSUPPORTED, "every streamed NV12 pack is in flight: pack synchronously");
  if (!r->done) HIP_OK(c, hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
  if (r->pin_cap < need) {
    if (r->pin) (void)hipHostFree(r->pin);
    r->pin = nullptr;
    r->pin_cap = 0;
    HIP_OK(c, hipHostMalloc(&r->pin, need, hipHostMallocDefault));
    r->pin_cap = need;
  }
  /* the decode packed this picture at this pitch already: copy that */
  const bool packed = c->nv12_current && pitch == c->nv12_pitch_done && src == c->frame;
  const uint8_t* from = c->pack;
  if (!packed) {
    if (r->dev_cap < need) {
      if (r->dev) (void)hipFree(r->dev);
      r->dev = nullptr;
      r->dev_cap = 0;
      HIP_OK(c, hipMalloc(&r->dev, need));
      r->dev_cap = need;
    }
    HIP_OK(c, launch_nv12(L, src, r->dev, (int)pitch, c->stream));
    from = r->dev;
  }
  HIP_OK(c, hipMemcpyAsync(r->pin, from, need, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipEventRecord(r->done, c->stream));
  c->dirty = true;
  r->bytes = need;
  r->busy = true;
  return (int)(r - c->ring);
}

int av1h_nv12_pack_wait(av1h_ctx* h, int handle, uint8_t* out, size_t out_bytes, size_t* bytes) {
  if (!h || handle < 0 || handle >= kPackRing) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  PackEntry* r = &c->ring[handle];
  hipEvent_t ev;
  size_t n;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (!r->busy) return set_err(c, AV1H_ERR_INVALID, "the streamed NV12 pack is not in flight");
    ev = r->done;
    n = r->bytes;
  }
  (void)hipSetDevice(c->device);
  if (hipEventSynchronize(ev) != hipSuccess) {
    std::lock_guard<std::mutex> lk(c->mu);
    return set_err(c, AV1H_ERR_DEVICE, "the streamed NV12 pack failed");
  }
  if (bytes) *bytes = n;
  if (out) {
    if (out_bytes < n) {
      std::lock_guard<std::mutex> lk(c->mu);
      return set_err(c, AV1H_ERR_INVALID, "output buffer too small");
    }
    std::memcpy(out, r->pin, n);
  }
  return AV1H_OK;
}

void av1h_nv12_pack_release(av1h_ctx* h, int handle) {
  if (!h || handle < 0 || handle >= kPackRing) return;
  Ctx* c = &h->c;
  PackEntry* r = &c->ring[handle];
  hipEvent_t ev = nullptr;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (!r->busy) return;
    ev = r->done;
  }
  (void)hipSetDevice(c->device);
  (void)hipEventSynchronize(ev);
  std::lock_guard<std::mutex> lk(c->mu);
  r->busy = false;
}

int av1h_plan_host_times(const void* plan, double* flatten_ms, double* upload_ms) {
  if (!plan) return AV1H_ERR_INVALID;
  const Plan* p = static_cast<const Plan*>(plan);
  if (flatten_ms) *flatten_ms = p->flatten_ms;
  if (upload_ms) *upload_ms = p->upload_ms;
  return AV1H_OK;
}

int av1h_plan_levels(const void* plan) {
  if (!plan) return AV1H_ERR_INVALID;
  const Plan* p = static_cast<const Plan*>(plan);
  return p->key_off_static.empty() ? 0 : (int)p->key_off_static.size() - 1;
}

int av1h_plan_queue_counts(const void* plan, int32_t* counts) {
  if (!plan || !counts) return AV1H_ERR_INVALID;
  const Plan* p = static_cast<const Plan*>(plan);
  for (int k = 0; k < 4; ++k) counts[k] = p->q_counts[k];
  return AV1H_OK;
}

int av1h_plan_sb_features(const void* plan, float* features) {
  if (!plan) return AV1H_ERR_INVALID;
  const Plan* p = static_cast<const Plan*>(plan);
  if (features && !p->sb_feat.empty()) std::memcpy(features, p->sb_feat.data(), p->sb_feat.size() * sizeof(float));
  return p->sb_w * p->sb_h;
}

int av1h_plan_sb_times(av1h_ctx* h, const void* plan, float* us) {
  if (!h || !plan || !us) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  const Plan* p = static_cast<const Plan*>(plan);
  const int nsb = p->sb_w * p->sb_h;
  for (int i = 0; i < nsb; ++i) us[i] = 0.f;
  if (c->trace_plan != plan || c->trace_n != (int)p->queue_host.size())
    return set_err(c, AV1H_ERR_INVALID, "the last decode on this context was not a traced decode of this plan alone");
  if (c->trace_n == 0) return nsb;
  (void)hipSetDevice(c->device);
  std::vector<uint64_t> t((size_t)c->trace_n * 4);
  HIP_OK(c, hipMemcpy(t.data(), c->trace, t.size() * 8, hipMemcpyDeviceToHost));
  int khz = 100000;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
  for (int k = 0; k < c->trace_n; ++k) {
    const double ticks = (double)(t[4 * (size_t)k + 2] - t[4 * (size_t)k + 1]);
    us[p->queue_sb[(size_t)k]] += (float)(ticks * 1000.0 / khz);
  }
  return nsb;
}

namespace {
/* Queue orders beyond the static one. Every queued block costs its
 * superblock's cost shared by sample count (its sample count when no costs
 * are given), and its remaining path is its cost plus the longest remaining
 * path of what waits for it.
 *  - AV1H_ORDER_LEVEL: by dependency level, the longest remaining path first
 *    inside a level, coding order between equals;
 *  - AV1H_ORDER_CODING: coding order (the row order of the superblocks);
 *  - AV1H_ORDER_PATH: the longest remaining path first (list scheduling by
 *    critical path), coding order between equals.
 * Each is a topological order of the block graph (a block's remaining path
 * exceeds that of every block waiting for it), so a queue handed out in it
 * always progresses. */
bool schedule_order(Plan* p, int order, const float* cost) {
  const int nq = (int)p->q_block.size();
  const int nsb = p->sb_w * p->sb_h;
  std::vector<double> area((size_t)nq, 0.0), sb_area((size_t)nsb, 0.0);
  std::vector<int32_t> sb_of((size_t)nq, 0);
  for (size_t k = 0; k < p->queue_static.size(); ++k) sb_of[(size_t)p->queue_qi[k]] = p->queue_sb_static[k];
  for (int qi = 0; qi < nq; ++qi) {
    const double a = (double)p->q_area[(size_t)qi];
    area[(size_t)qi] = a;
    sb_area[(size_t)sb_of[(size_t)qi]] += a;
  }
  std::vector<double> path((size_t)nq, 0.0), succ((size_t)nq, 0.0);
  for (int qi = nq - 1; qi >= 0; --qi) {
    const int sb = sb_of[(size_t)qi];
    double c = area[(size_t)qi] / 64.0;
    if (cost) c = sb_area[(size_t)sb] > 0 ? (double)cost[sb] * area[(size_t)qi] / sb_area[(size_t)sb] : 0.0;
    path[(size_t)qi] = std::max(c, 0.01) + succ[(size_t)qi];
    for (int32_t e = p->q_dep_off[(size_t)qi]; e < p->q_dep_off[(size_t)qi + 1]; ++e) {
      const int32_t o = p->q_dep[(size_t)e];
      if (o >= qi) return false; /* a predecessor later in coding order */
      succ[(size_t)o] = std::max(succ[(size_t)o], path[(size_t)qi]);
    }
  }
  std::vector<int32_t> ord((size_t)nq);
  for (int k = 0; k < nq; ++k) ord[(size_t)k] = k;
  if (order == AV1H_ORDER_LEVEL)
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
      const int32_t la = p->q_level[(size_t)a], lb = p->q_level[(size_t)b];
      return la != lb ? la < lb : path[(size_t)a] > path[(size_t)b];
    });
  else if (order == AV1H_ORDER_PATH)
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return path[(size_t)a] > path[(size_t)b]; });
  /* the keys a batch merges its frames' queues by (key_off: where each key
   * starts; keys non-decreasing along a queue): the level; for the path
   * order the remaining path in quarter units counted down from the frame's
   * longest, so that merged frames line up by remaining path; for coding
   * order runs of 64 entries */
  const double top = nq ? path[(size_t)ord[0]] : 0.0;
  p->queue_host.resize((size_t)nq);
  p->queue_sb.resize((size_t)nq);
  p->key_off.clear();
  for (int k = 0; k < nq; ++k) {
    const int32_t qi = ord[(size_t)k];
    p->queue_host[(size_t)k] = p->q_block[(size_t)qi];
    p->queue_sb[(size_t)k] = sb_of[(size_t)qi];
    const int key = order == AV1H_ORDER_LEVEL ? p->q_level[(size_t)qi]
                    : order == AV1H_ORDER_PATH ? (int)((top - path[(size_t)qi]) * 4.0)
                                               : k >> 6;
    while ((int)p->key_off.size() <= key) p->key_off.push_back((int32_t)k);
  }
  p->key_off.push_back((int32_t)nq);
  return true;
}

int upload_queue(Ctx* c, Plan* p) {
  if (p->queue_host.empty()) return AV1H_OK;
  HIP_OK(c, hipMemcpy(const_cast<int32_t*>(p->dev.queue_list), p->queue_host.data(), p->queue_host.size() * 4,
                      hipMemcpyHostToDevice));
  return AV1H_OK;
}
}  // namespace

int av1h_plan_schedule(av1h_ctx* h, void* plan, const float* sb_cost_us) {
  if (!h || !plan) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Plan* p = static_cast<Plan*>(plan);
  if (!sb_cost_us) {
    p->queue_host = p->queue_static;
    p->key_off = p->key_off_static;
    p->queue_sb = p->queue_sb_static;
  } else if (!schedule_order(p, AV1H_ORDER_LEVEL, sb_cost_us)) {
    return set_err(c, AV1H_ERR_INVALID, "the superblock dependencies of this plan are cyclic");
  }
  return upload_queue(c, p);
}

int av1h_plan_schedule_order(av1h_ctx* h, void* plan, int order, const float* sb_cost_us) {
  if (!h || !plan || order < AV1H_ORDER_LEVEL || order > AV1H_ORDER_PATH) return AV1H_ERR_INVALID;
  if (order == AV1H_ORDER_LEVEL && !sb_cost_us) return av1h_plan_schedule(h, plan, nullptr);
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Plan* p = static_cast<Plan*>(plan);
  if (!schedule_order(p, order, sb_cost_us))
    return set_err(c, AV1H_ERR_INVALID, "the superblock dependencies of this plan are cyclic");
  return upload_queue(c, p);
}

int av1h_plan_schedule_model(av1h_ctx* h, void* plan, av1h_costmodel* model) {
  if (!h || !plan || !model) return AV1H_ERR_INVALID;
  Plan* p = static_cast<Plan*>(plan);
  const int n = p->sb_w * p->sb_h;
  std::vector<float> cost((size_t)n);
  const int rc = av1h_cm_predict(model, p->sb_feat.data(), n, cost.data());
  if (rc != AV1H_OK) return set_err(&h->c, rc, "cost model inference failed");
  return av1h_plan_schedule(h, plan, cost.data());
}

int av1h_set_pool_slots(av1h_ctx* h, int slots) {
  if (!h || slots < 1 || slots > kSlots) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  for (int i = 0; i < kSlots; ++i)
    if (c->used[i]) return set_err(c, AV1H_ERR_INVALID, "the pool can be resized only while no slot is in use");
  c->pool_slots = slots;
  return AV1H_OK;
}

int64_t av1h_stream_bytes(int width, int height, int bit_depth, int ss_x, int ss_y, int pool_slots,
                          size_t nv12_pitch) {
  if (width <= 0 || height <= 0 || pool_slots < 1 || pool_slots > kSlots) return -1;
  Layout L;
  make_layout(&L, width, height, bit_depth, ss_x, ss_y, width);
  const int64_t pic = L.size * 2;
  int64_t b = pic * (5 + pool_slots);                                  /* work pictures + pool */
  b += L.cells * 4 + L.edge_words * 8;                                  /* cells, packed edges */
  b += (int64_t)(L.grid_w[0] >> 3) * (L.grid_h[0] >> 3) * 8;             /* CDEF directions */
  b += 3 * 82 * 73 * 2 + 65536;                                          /* grain templates, offsets */
  if (nv12_pitch) b += (int64_t)av1h_nv12_size(width, height, ss_x, ss_y, nv12_pitch);
  return b;
}

int av1h_set_output_hash(av1h_ctx* h, int enable) {
  if (!h) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->c.mu);
  h->c.hash_on = enable != 0;
  return AV1H_OK;
}

int av1h_last_hash(av1h_ctx* h, uint64_t* out) {
  if (!h || !out) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->c.mu);
  if (!h->c.hash_valid) return set_err(&h->c, AV1H_ERR_INVALID, "the last decode computed no checksum");
  *out = h->c.last_hash;
  return AV1H_OK;
}

int av1h_tune_params(void) { return kTuneParams; }

const char* av1h_tune_name(int i) {
  static const char* names[kTuneParams] = {
      "mc", "inter_special", "itx_inter", "itx_intra", "itx_inter_large", "itx_intra_large", "deblock_v",
      "deblock_h", "cdef_dir", "cdef_luma",
      "cdef_chroma", "upscale", "lr", "grain_template", "grain_apply", "nv12", "hash", "copy_lr_curr",
      "copy_cdef_src", "copy_upscale_src", "copy_store", "queue", "cdef_blocks_per_thread"};
  return i >= 0 && i < kTuneParams ? names[i] : nullptr;
}

int av1h_tune_get(const av1h_ctx* h, int32_t* params, int n) {
  if (!h || !params || n < kTuneParams) return AV1H_ERR_INVALID;
  for (int i = 0; i < kTuneParams; ++i) params[i] = h->c.tune.v[i];
  return kTuneParams;
}

int av1h_tune_set(av1h_ctx* h, const int32_t* params, int n) {
  if (!h || !params || n != kTuneParams) return AV1H_ERR_INVALID;
  for (int i = 0; i < kTuneParams; ++i) {
    const int lim = i == kTuneCdefBatch ? kCdefMaxBatch : 64;
    if (params[i] < 1 || params[i] > lim) return AV1H_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(h->c.mu);
  (void)hipSetDevice(h->c.device);
  for (int i = 0; i < kTuneParams; ++i) h->c.tune.v[i] = params[i];
  drop_graphs(&h->c);
  return AV1H_OK;
}

int av1h_set_graphs(av1h_ctx* h, int mode) {
  if (!h || mode < 0 || mode > 2) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->c.mu);
  const bool changed = (h->c.graphs != (mode != 0)) || (mode != 0 && h->c.graph_sets != (mode == 1));
  h->c.graphs = mode != 0;
  h->c.graph_sets = mode == 1;
  if (changed) {
    (void)hipSetDevice(h->c.device);
    drop_graphs(&h->c);
  }
  return AV1H_OK;
}

int av1h_set_decode_nv12(av1h_ctx* h, size_t pitch) {
  if (!h || (pitch & 1)) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->c.mu);
  h->c.decode_nv12_pitch = pitch;
  h->c.nv12_current = false;
  return AV1H_OK;
}

int av1h_to_rgb(av1h_ctx* h, int slot, int matrix, int range, uint8_t* out, size_t out_bytes) {
  if (!h || !out) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Layout L;
  const uint16_t* src = gather(c, slot, &L);
  if (!src) return AV1H_ERR_INVALID;
  const size_t need = (size_t)L.width * L.height * 3;
  if (out_bytes < need) return set_err(c, AV1H_ERR_INVALID, "output buffer too small");
  /* Q16 coefficients (reference src/core/tables.cpp:481-533) */
  int m[8] = {0}, o[3] = {0};
  const int sh = L.bit_depth - 8;
  const int yo = range == 0 ? (16 << sh) : 0;
  const int co = range == 0 ? (128 << sh) : (1 << (L.bit_depth - 1));
  if (matrix == 3) {
    m[0] = m[2] = m[5] = 65536;
  } else {
    if (matrix == 2) {
      const int v[7] = {76293, 109969, 76293, -12255, -42598, 76293, 121634};
      std::memcpy(m, v, sizeof(v));
    } else {
      const int v[7] = {76293, 117504, 76293, -13926, -34937, 76293, 138412};
      std::memcpy(m, v, sizeof(v));
    }
    o[0] = (1 << 15) - m[0] * yo - m[1] * co;
    o[1] = (1 << 15) - m[2] * yo - m[3] * co - m[4] * co;
    o[2] = (1 << 15) - m[5] * yo - m[6] * co;
  }
  int rc = ensure_pack(c, need);
  if (rc != AV1H_OK) return rc;
  HIP_OK(c, launch_rgb(L, src, c->pack, m, o, c->stream));
  HIP_OK(c, hipMemcpyAsync(out, c->pack, need, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return AV1H_OK;
}

int av1h_read_slot(av1h_ctx* h, int slot, const av1h_planes* out, av1h_frame_info* info) {
  if (!h) return AV1H_ERR_INVALID;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Layout L;
  const uint16_t* src = gather(c, slot, &L);
  if (!src) return AV1H_ERR_INVALID;
  if (out) {
    for (int p = 0; p < 3; ++p)
      HIP_OK(c, hipMemcpy2DAsync(out->plane[p], (size_t)out->stride[p] * 2, src + L.plane_off[p],
                                 (size_t)L.stride[p] * 2, (size_t)L.out_w[p] * 2, L.out_h[p],
                                 hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
  }
  if (info) {
    info->width = L.width;
    info->height = L.height;
    info->bit_depth = L.bit_depth;
    info->ss_x = L.ss_x;
    info->ss_y = L.ss_y;
  }
  return AV1H_OK;
}

void av1h_reset(av1h_ctx* h) {
  if (!h) return;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  while (!c->pending.empty()) (void)batch_finish(c);
  std::memset(c->used, 0, sizeof(c->used));
  std::memset(c->has_pic, 0, sizeof(c->has_pic));
  c->last_slot = -1;
  c->have_frame = false;
  if (c->frame) (void)hipMemsetAsync(c->frame, 0, c->cap * 2, c->stream);
  if (c->pool) (void)hipMemsetAsync(c->pool, 0, c->cap * 2 * c->pool_alloc, c->stream);
  (void)hipStreamSynchronize(c->stream);
  /* streamed packs are complete after the sync; their handles end here */
  for (PackEntry& r : c->ring) r.busy = false;
  std::memset(&c->stats, 0, sizeof(c->stats));
}

int av1h_get_stats(const av1h_ctx* h, av1h_stats* out) {
  if (!h || !out) return AV1H_ERR_INVALID;
  *out = h->c.stats;
  return AV1H_OK;
}

void av1h_set_stage_timing(av1h_ctx* h, int enable) {
  if (!h) return;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  if (enable && !c->sev[0])
    for (auto& e : c->sev)
      if (hipEventCreate(&e) != hipSuccess) {
        set_err(c, AV1H_ERR_DEVICE, "hipEventCreate failed");
        return;
      }
  c->timing = enable != 0 && c->sev[2 * AV1H_STAGES - 1] != nullptr;
  for (int st = 0; st < AV1H_STAGES; ++st) {
    c->stage_ms[st] = 0;
    c->stage_n[st] = 0;
    c->stage_bytes[st] = 0;
  }
}

int av1h_get_stage_times(const av1h_ctx* h, double ms[AV1H_STAGES], int64_t count[AV1H_STAGES],
                         int64_t bytes[AV1H_STAGES]) {
  if (!h) return AV1H_ERR_INVALID;
  for (int st = 0; st < AV1H_STAGES; ++st) {
    if (ms) ms[st] = h->c.stage_ms[st];
    if (count) count[st] = h->c.stage_n[st];
    if (bytes) bytes[st] = h->c.stage_bytes[st];
  }
  return AV1H_OK;
}

void av1h_set_queue_trace(av1h_ctx* h, int enable) {
  if (h) h->c.trace_on = enable != 0;
}

void av1h_set_wait_limit(av1h_ctx* h, int polls) {
  if (h) h->c.spin_limit = polls > 0 ? polls : kSpinLimitDefault;
}

int64_t av1h_get_queue_trace(av1h_ctx* h, uint64_t* out, int64_t cap) {
  if (!h) return -1;
  Ctx* c = &h->c;
  std::lock_guard<std::mutex> lk(c->mu);
  const int64_t n = c->trace_on && c->trace ? c->trace_n : 0;
  if (out && cap >= 4 * n && n > 0) {
    (void)hipSetDevice(c->device);
    HIP_OK(c, hipMemcpy(out, c->trace, (size_t)n * 32, hipMemcpyDeviceToHost));
  }
  return n;
}

}  // extern "C"

namespace {
/* locks the contexts of a batch in address order */
struct BatchLock {
  std::vector<std::unique_lock<std::mutex>> locks;
  explicit BatchLock(std::vector<Ctx*> cs) {
    std::sort(cs.begin(), cs.end());
    cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
    for (Ctx* c : cs) locks.emplace_back(c->mu);
  }
};

int check_batch_args(av1h_ctx* const* ctxs, int count) {
  if (!ctxs || count < 1 || count > kMaxBatch) return AV1H_ERR_INVALID;
  for (int i = 0; i < count; ++i)
    if (!ctxs[i]) return AV1H_ERR_INVALID;
  return AV1H_OK;
}
}  // namespace

extern "C" {

int av1h_decode_batch(av1h_ctx* const* ctxs, const av1h_frame_desc* const* frames, int count, int* slots_out) {
  int rc = check_batch_args(ctxs, count);
  if (rc != AV1H_OK || !frames) return AV1H_ERR_INVALID;
  std::vector<Ctx*> cs(count);
  for (int i = 0; i < count; ++i) cs[i] = &ctxs[i]->c;
  BatchLock lk(cs);
  Ctx* lead = cs[0];
  (void)hipSetDevice(lead->device);
  const double t0 = now_ms();
  std::vector<JobSpec> js(count);
  for (int i = 0; i < count; ++i) {
    if (!frames[i]) return set_err(lead, AV1H_ERR_INVALID, "frame %d of the batch is null", i);
    for (int k = 0; k < i; ++k)
      if (cs[k] == cs[i]) return set_err(lead, AV1H_ERR_INVALID, "a context appears twice in one batch");
    Host host;
    rc = flatten(*frames[i], &cs[i]->scratch_plan, &host, &cs[i]->err);
    if (rc == AV1H_OK) rc = upload(cs[i], &cs[i]->scratch_plan, host);
    if (rc != AV1H_OK) return cs[i] == lead ? rc : set_err(lead, rc, "frame %d: %s", i, cs[i]->err.c_str());
    js[i] = JobSpec{cs[i], &cs[i]->scratch_plan, 1, -1, nullptr, nullptr, slots_out ? &slots_out[i] : nullptr, -1};
  }
  return run_batch(lead, js, t0);
}

int av1h_decode_planned_batch(av1h_ctx* const* ctxs, void* const* plans, const int* keep, const int* reserved,
                              int count, int* slots_out) {
  if (!ctxs || count < 1) return AV1H_ERR_INVALID;
  return av1h_decode_planned_batch_on(ctxs[0], ctxs, plans, keep, reserved, count, slots_out);
}

int av1h_decode_planned_batch_on(av1h_ctx* launcher, av1h_ctx* const* ctxs, void* const* plans, const int* keep,
                                 const int* reserved, int count, int* slots_out) {
  int rc = check_batch_args(ctxs, count);
  if (rc != AV1H_OK || !plans || !launcher) return AV1H_ERR_INVALID;
  std::vector<Ctx*> cs(count);
  for (int i = 0; i < count; ++i) cs[i] = &ctxs[i]->c;
  std::vector<Ctx*> locked = cs;
  locked.push_back(&launcher->c);
  BatchLock lk(locked);
  Ctx* lead = &launcher->c;
  (void)hipSetDevice(lead->device);
  std::vector<JobSpec> js(count);
  for (int i = 0; i < count; ++i) {
    if (!plans[i]) return set_err(lead, AV1H_ERR_INVALID, "plan %d of the batch is null", i);
    js[i] = JobSpec{cs[i], static_cast<Plan*>(plans[i]), keep ? keep[i] : 1, reserved ? reserved[i] : -1, nullptr,
                    nullptr, slots_out ? &slots_out[i] : nullptr, -1};
  }
  return run_batch(lead, js, now_ms());
}

/* The batch launched without waiting for it (av1h_decode_planned_batch_on
 * split in two): the slots are known at once, the results come with
 * av1h_batch_finish. Stage timing and the queue trace take the waiting
 * path. */
int av1h_decode_planned_batch_async(av1h_ctx* launcher, av1h_ctx* const* ctxs, void* const* plans, const int* keep,
                                    const int* reserved, int count, int* slots_out) {
  int rc = check_batch_args(ctxs, count);
  if (rc != AV1H_OK || !plans || !launcher) return AV1H_ERR_INVALID;
  std::vector<Ctx*> cs(count);
  for (int i = 0; i < count; ++i) cs[i] = &ctxs[i]->c;
  std::vector<Ctx*> locked = cs;
  locked.push_back(&launcher->c);
  BatchLock lk(locked);
  Ctx* lead = &launcher->c;
  (void)hipSetDevice(lead->device);
  std::vector<JobSpec> js(count);
  for (int i = 0; i < count; ++i) {
    if (!plans[i]) return set_err(lead, AV1H_ERR_INVALID, "plan %d of the batch is null", i);
    js[i] = JobSpec{cs[i], static_cast<Plan*>(plans[i]), keep ? keep[i] : 1, reserved ? reserved[i] : -1, nullptr,
                    nullptr, slots_out ? &slots_out[i] : nullptr, -1};
  }
  rc = batch_launch(lead, js, now_ms());
  if (rc == AV1H_OK && (lead->timing || lead->trace_on))
    while (rc == AV1H_OK && !lead->pending.empty()) rc = batch_finish(lead);
  for (JobSpec& j : lead->ticket[0].js) j.slot_out = nullptr;
  for (JobSpec& j : lead->ticket[1].js) j.slot_out = nullptr;
  return rc;
}

int av1h_batch_finish(av1h_ctx* launcher) {
  if (!launcher) return AV1H_ERR_INVALID;
  Ctx* lead = &launcher->c;
  std::vector<Ctx*> locked{lead};
  {
    std::lock_guard<std::mutex> lk(lead->mu);
    if (lead->pending.empty()) return AV1H_OK;
    for (JobSpec& j : lead->ticket[lead->pending.front()].js) locked.push_back(j.c);
  }
  BatchLock lk(locked);
  (void)hipSetDevice(lead->device);
  return batch_finish(lead);
}

int av1h_batch_pending(av1h_ctx* launcher) {
  if (!launcher) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(launcher->c.mu);
  return (int)launcher->c.pending.size();
}

int av1h_plan_needs_idle(av1h_ctx* ctx, void* plan) {
  if (!ctx || !plan) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->c.mu);
  /* stage timing and the queue trace take the waiting path */
  const Ctx* c = &ctx->c;
  return c->timing || c->trace_on || must_grow(c, static_cast<const Plan*>(plan)) ? 1 : 0;
}

}  // extern "C"
