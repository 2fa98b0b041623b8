/*
 * Seeded synthetic FrameSyntax generator (see synth.hpp for the contract).
 *
 * Every random draw comes from one splitmix64 stream per frame, so a clip is a
 * pure function of (config, seed) on every platform.
 */
#include "synth.hpp"
#include "scan_order.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace av1gpu {
namespace synth {
namespace {

/* the diagonal scans (scan_order.hpp) */
constexpr av1scan::DiagTables kDiag = av1scan::make_diag();

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  /* inclusive range */
  int range(int lo, int hi) {
    return lo + static_cast<int>(next() % static_cast<uint64_t>(hi - lo + 1));
  }
  bool chance(double p) { return uni() < p; }
  double gauss() {
    double u1 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

/* Per plane bookkeeping on a 4 sample cell grid (every block and transform of
 * this generator is a multiple of 4 samples in every plane). */
struct PlaneState {
  int w = 0, h = 0;         /* grid size in samples (coded width rounded up to 8) */
  int cw = 0, ch = 0;       /* cells                                             */
  std::vector<uint8_t> decoded;
  std::vector<uint8_t> smooth;
  std::vector<int16_t> txw, txh, tx_x0, tx_y0;
  std::vector<uint8_t> level;
  void init(int pw, int ph) {
    w = pw;
    h = ph;
    cw = (pw + 3) / 4;
    ch = (ph + 3) / 4;
    const size_t n = static_cast<size_t>(cw) * ch;
    decoded.assign(n, 0);
    smooth.assign(n, 0);
    txw.assign(n, 0);
    txh.assign(n, 0);
    tx_x0.assign(n, 0);
    tx_y0.assign(n, 0);
    level.assign(n, 0);
  }
  size_t idx(int x, int y) const { return static_cast<size_t>(y / 4) * cw + (x / 4); }
  /* every in-grid cell of the rectangle was emitted before (cells outside the
   * grid count as decoded, the reference's owner -1) */
  bool rect_decoded(int x0, int x1, int y0, int y1) const {
    for (int y = std::max(y0, 0) & ~3; y <= y1 && y < h; y += 4)
      for (int x = std::max(x0, 0) & ~3; x <= x1 && x < w; x += 4)
        if (!decoded[idx(x, y)]) return false;
    return true;
  }
  void mark(int x, int y, int bw, int bh, bool is_smooth, int tw, int th, int lvl) {
    for (int yy = y; yy < y + bh; yy += 4)
      for (int xx = x; xx < x + bw; xx += 4) {
        if (xx >= w || yy >= h) continue;
        const size_t i = idx(xx, yy);
        decoded[i] = 1;
        smooth[i] = is_smooth ? 1 : 0;
        level[i] = static_cast<uint8_t>(lvl);
        (void)tw;
        (void)th;
      }
  }
  void mark_tx(int x, int y, int tw, int th) {
    for (int yy = y; yy < y + th; yy += 4)
      for (int xx = x; xx < x + tw; xx += 4) {
        if (xx >= w || yy >= h) continue;
        const size_t i = idx(xx, yy);
        txw[i] = static_cast<int16_t>(tw);
        txh[i] = static_cast<int16_t>(th);
        tx_x0[i] = static_cast<int16_t>(x);
        tx_y0[i] = static_cast<int16_t>(y);
      }
  }
};

bool is_smooth_mode(int mode) { return mode >= 9 && mode <= 11; }

/* which 1D transform a direction of a Tx_Type uses: 0 DCT, 1 ADST, 2 identity */
int row_kind(int t) {
  static const int k[16] = {0, 0, 1, 1, 0, 1, 1, 1, 1, 2, 2, 0, 2, 1, 2, 1};
  return k[t];
}
int col_kind(int t) {
  static const int k[16] = {0, 1, 0, 1, 1, 0, 1, 1, 1, 2, 0, 2, 1, 2, 1, 2};
  return k[t];
}
bool kind_fits(int kind, int n) {
  if (kind == 0) return n <= 64;
  if (kind == 1) return n <= 16;
  return n <= 32;
}

class FrameBuilder {
 public:
  FrameBuilder(const Config& c, int f, const std::vector<int>& refs, bool key)
      : cfg_(c), rng_(c.seed * 0x100000001B3ull + static_cast<uint64_t>(f) * 7919u + 17u),
        key_(key), nrefs_(static_cast<int>(refs.size())) {
    fs_.width = c.width;
    fs_.height = c.height;
    fs_.coded_width = c.coded_width > 0 && c.coded_width < c.width ? c.coded_width : 0;
    fs_.bit_depth = c.bit_depth;
    fs_.ss_x = 1;
    fs_.ss_y = 1;
    fs_.base_q_idx = rng_.range(20, 200);
    fs_.sharpness = rng_.range(0, 7);
    fs_.enable_intra_edge_filter = c.contract_tools ? rng_.chance(0.8) : true;
    fs_.reference_slots = refs;
    fs_.color.matrix = ColorMatrix::Bt709;
    fs_.color.range = ColorRange::Limited;
    coded_w_ = fs_.coded_width > 0 ? fs_.coded_width : c.width;
    grid_w_ = (coded_w_ + 7) & ~7;
    grid_h_ = (c.height + 7) & ~7;
    planes_[0].init(grid_w_, grid_h_);
    planes_[1].init(grid_w_ >> 1, grid_h_ >> 1);
    planes_[2].init(grid_w_ >> 1, grid_h_ >> 1);
  }

  FrameSyntax build() {
    for (int sy = 0; sy < grid_h_; sy += 64)
      for (int sx = 0; sx < grid_w_; sx += 64) partition(sx, sy, 64);
    build_deblock();
    if (cfg_.cdef) build_cdef();
    if (cfg_.lr) build_lr();
    if (cfg_.grain) build_grain();
    /* the pool the itx jobs read is sized to the largest w*h raster end plus
     * slack; keep a zero tail so every raster is in bounds */
    fs_.coefficients.resize(fs_.coefficients.size() + 64, 0);
    if (!fs_.residual.empty()) fs_.residual.resize(fs_.residual.size() + 64, 0);
    return std::move(fs_);
  }

 private:
  struct InterParams {
    int ref[2] = {0, -1};
    int mv[2][2] = {{0, 0}, {0, 0}};
    int filt = 0, filt_v = 0;
    int compound = 0;
    int blend[2] = {8, 8};
    int wedge_index = 0, mask_sign = 0;
    bool interintra = false, interintra_wedge = false;
    int interintra_mode = 0;
    bool warp = false;
    int32_t warp_params[6] = {0, 0, 0, 0, 0, 0};
    bool intrabc = false;
    std::vector<ObmcCandidate> obmc; /* luma geometry */
  };
  const Config& cfg_;
  Rng rng_;
  bool key_;
  int nrefs_;
  FrameSyntax fs_;
  int coded_w_ = 0, grid_w_ = 0, grid_h_ = 0;
  bool steps_ready_ = false; /* levels: the frame's steps and matrices are drawn */
  int dc_step_ = 0, ac_step_ = 0;
  PlaneState planes_[3];

  void partition(int x, int y, int size) {
    if (x >= grid_w_ || y >= grid_h_) return;
    const bool crosses = x + size > grid_w_ || y + size > grid_h_;
    if (size == 8) {
      coding_block(x, y, 8, 8);
      return;
    }
    int p;
    if (crosses) {
      p = 3;
    } else {
      const double u = rng_.uni();
      p = u < 0.3 ? 0 : (u < 0.4 ? 1 : (u < 0.5 ? 2 : 3));
    }
    const int half = size / 2;
    switch (p) {
      case 0: coding_block(x, y, size, size); break;
      case 1:
        coding_block(x, y, size, half);
        coding_block(x, y + half, size, half);
        break;
      case 2:
        coding_block(x, y, half, size);
        coding_block(x + half, y, half, size);
        break;
      default:
        partition(x, y, half);
        partition(x + half, y, half);
        partition(x, y + half, half);
        partition(x + half, y + half, half);
        break;
    }
  }

  int pick_type(int tw, int th) {
    if (!cfg_.tx_ext) {
      if (tw <= 16 && th <= 16) return rng_.range(0, 3);
      return 0;
    }
    int cand[16];
    int n = 0;
    for (int t = 0; t < 16; ++t)
      if (kind_fits(row_kind(t), tw) && kind_fits(col_kind(t), th)) {
        /* with coded levels, the types AV1's transform sets allow: the 1D
         * types (V_*, H_*, scanned row or column wise) only up to 16x16 */
        if (cfg_.levels && t >= 10 && (tw > 16 || th > 16)) continue;
        cand[n++] = t;
      }
    return cand[rng_.range(0, n - 1)];
  }

  int coef_value() {
    /* Laplacian(scale) via the inverse CDF, rounded, clamped to int16 */
    double u = rng_.uni() - 0.5;
    const double sign = u < 0 ? -1.0 : 1.0;
    double a = 1.0 - 2.0 * std::fabs(u);
    if (a < 1e-300) a = 1e-300;
    double v = -cfg_.coef_scale * sign * std::log(a);
    long r = std::lround(v);
    if (r == 0) r = sign < 0 ? -1 : 1;
    if (r > 32767) r = 32767;
    if (r < -32768) r = -32768;
    return static_cast<int>(r);
  }

  TransformBlock make_tx(int rx, int ry, int tw, int th, bool lossless, bool raw = false) {
    TransformBlock t;
    t.x = rx;
    t.y = ry;
    t.w = tw;
    t.h = th;
    t.tx_type = static_cast<TransformType>(lossless ? 0 : pick_type(tw, th));
    t.qindex = fs_.base_q_idx;
    t.denominator = 1;
    if (rng_.chance(cfg_.tx_skip)) {
      t.eob = 0;
      t.coef_offset = 0;
      return t;
    }
    if (raw) {
      /* raw residual samples (residual_provided): the whole tw x th raster,
       * half of it zero, the rest reaching past the sample range on both
       * sides so the add's clip is exercised */
      const int lim = (1 << cfg_.bit_depth) + 64;
      t.coef_offset = static_cast<int>(fs_.residual.size());
      int nnz = 0;
      for (int i = 0; i < tw * th; ++i) {
        const int v = rng_.chance(0.5) ? 0 : rng_.range(-lim, lim);
        nnz += v != 0;
        fs_.residual.push_back(v);
      }
      t.eob = std::max(nnz, 1);
      return t;
    }
    const int cw = std::min(32, tw), ch = std::min(32, th);
    t.coef_offset = static_cast<int>(fs_.coefficients.size());
    if (cfg_.levels) return levels_tx(t, tw, th, cw, ch, lossless);
    int nnz = 0;
    for (int i = 0; i < cw * ch; ++i) {
      int v = 0;
      if (!rng_.chance(cfg_.coef_zero)) {
        v = coef_value();
        if (lossless) v = std::max(-2000, std::min(2000, v));
      }
      if (v) ++nnz;
      fs_.coefficients.push_back(v);
    }
    t.eob = nnz;
    return t;
  }

  /* The coded levels of a transform, in scan order, and the dequantiser a
   * frontend would derive for it (the frame's steps, the transform's
   * denominator, a quantiser matrix for some); the coefficient raster is
   * their dequantisation, so the frame carries both, as a parsed stream
   * does. */
  TransformBlock levels_tx(TransformBlock t, int tw, int th, int cw, int ch, bool lossless) {
    if (!steps_ready_) {
      /* the frame's quantiser and matrices, from a stream of their own */
      Rng q(cfg_.seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(fs_.base_q_idx) * 131u + 7u);
      const int sh = cfg_.bit_depth - 8;
      dc_step_ = q.range(24, 80) << sh;
      ac_step_ = q.range(32, 96) << sh;
      for (int m = 0; m < 4; ++m)
        for (int i = 0; i < 1024; ++i) fs_.qm_pool.push_back(static_cast<uint8_t>(q.range(20, 64)));
      steps_ready_ = true;
    }
    const int area = cw * ch, lw = __builtin_ctz(static_cast<unsigned>(cw)) - 2,
              lh = __builtin_ctz(static_cast<unsigned>(ch)) - 2;
    const int cls = av1scan::tx_class(static_cast<int>(t.tx_type));
    t.dc_step = lossless ? 4 : dc_step_;
    t.ac_step = lossless ? 4 : ac_step_;
    t.denominator = 1 << ((tw * th > 256) + (tw * th > 1024));
    t.qm_off = !lossless && rng_.chance(0.2) ? 1024 * rng_.range(0, 3) : -1;
    std::vector<int> lv(static_cast<size_t>(area), 0);
    int eob = 0;
    for (int i = 0; i < area; ++i) {
      if (rng_.chance(cfg_.coef_zero)) continue;
      const int step = i == 0 ? t.dc_step : t.ac_step;
      long v = std::lround(static_cast<double>(coef_value()) / step);
      if (v == 0) v = rng_.chance(0.5) ? -1 : 1;
      const long cap = lossless ? 500 : 32767;
      lv[static_cast<size_t>(i)] = static_cast<int>(std::max(-cap, std::min(cap, v)));
      eob = i + 1;
    }
    t.eob = eob;
    t.level_offset = static_cast<int>(fs_.levels.size());
    for (int i = 0; i < eob; ++i) fs_.levels.push_back(static_cast<int16_t>(lv[static_cast<size_t>(i)]));
    fs_.coefficients.resize(fs_.coefficients.size() + static_cast<size_t>(area), 0);
    int32_t* raster = fs_.coefficients.data() + t.coef_offset;
    for (int i = 0; i < eob; ++i) {
      const int pos = av1scan::scan_pos(kDiag.scan, cls, lw, lh, i);
      const int qm = t.qm_off >= 0 ? fs_.qm_pool[static_cast<size_t>(t.qm_off + pos)] : -1;
      raster[pos] = av1scan::dequant(lv[static_cast<size_t>(i)], pos == 0 ? t.dc_step : t.ac_step, qm,
                                     __builtin_ctz(static_cast<unsigned>(t.denominator)), cfg_.bit_depth);
    }
    return t;
  }

  /* tiling of a w x h region with transforms of tw x th, raster order */
  void tx_size_for(int plane, int w, int h, bool intra, int* tw, int* th) {
    const int cap = std::min(cfg_.max_tx, plane ? 32 : 64);
    if (plane) {
      /* the chroma transform is the largest that fits (4:2:0, capped at 32) */
      if (cfg_.rect_tx && (w == 2 * h || h == 2 * w) && w <= cap && h <= cap) {
        *tw = w;
        *th = h;
        return;
      }
      const int s = std::min(std::min(w, h), cap);
      *tw = *th = s;
      return;
    }
    if (cfg_.rect_tx && (w == 2 * h || h == 2 * w) && w <= cap && h <= cap && rng_.chance(0.35)) {
      *tw = w;
      *th = h;
      return;
    }
    int sizes[5];
    int n = 0;
    for (int s = 4; s <= std::min(std::min(w, h), cap); s *= 2) sizes[n++] = s;
    const int s = sizes[rng_.range(0, n - 1)];
    (void)intra;
    *tw = *th = s;
  }

  void coding_block(int x, int y, int w, int h) {
    const bool intra = key_ || nrefs_ == 0 || rng_.chance(cfg_.intra_frac);
    const int level = rng_.range(cfg_.deblock_lo, cfg_.deblock_hi);
    const bool lossless = cfg_.contract_tools && intra && w == 8 && h == 8 && rng_.chance(0.05);
    const bool raw = cfg_.residual_frac > 0 && !lossless && rng_.chance(cfg_.residual_frac);
    if (intra && key_ && cfg_.intrabc > 0 && y >= 64 && rng_.chance(cfg_.intrabc)) {
      intrabc_block(x, y, w, h, level, raw);
    } else if (intra) {
      const int ymode = rng_.range(0, 12);
      const int yangle = (ymode >= 1 && ymode <= 8) ? rng_.range(-3, 3) : 0;
      intra_block(0, x, y, w, h, ymode, yangle, level, lossless, 0, raw);
      int uvmode = rng_.range(0, 12);
      if (cfg_.contract_tools && rng_.chance(0.2)) uvmode = 13;
      const int uvangle = (uvmode >= 1 && uvmode <= 8) ? rng_.range(-3, 3) : 0;
      const int cfl_alpha = rng_.range(-16, 16);
      for (int p = 1; p < 3; ++p)
        intra_block(p, x >> 1, y >> 1, w >> 1, h >> 1, uvmode, uvangle, level, lossless,
                    cfl_alpha, raw);
    } else {
      InterParams ip = draw_inter();
      if (cfg_.inter_tools > 0 && w >= 8 && h >= 8 && w <= 32 && h <= 32 && rng_.chance(cfg_.inter_tools))
        draw_tool(&ip, w, h);
      for (int p = 0; p < 3; ++p) {
        const int s = p ? 1 : 0;
        inter_block(p, x >> s, y >> s, w >> s, h >> s, ip, level, raw);
      }
    }
  }

  /* One of the contract inter tools for a coding block (luma w x h). */
  void draw_tool(InterParams* ip, int w, int h) {
    const int tool = rng_.range(0, 4);
    auto second_ref = [&]() {
      if (nrefs_ >= 2) {
        int r1 = rng_.range(0, nrefs_ - 2);
        if (r1 >= ip->ref[0]) ++r1;
        ip->ref[1] = r1;
      } else {
        ip->ref[1] = ip->ref[0];
      }
    };
    if (tool == 0 || tool == 1) {
      /* wedge or difference weighted compound */
      ip->compound = tool == 0 ? 3 : 4;
      second_ref();
      ip->wedge_index = rng_.range(0, 15);
      ip->mask_sign = rng_.range(0, 1);
    } else if (tool == 2) {
      /* inter/intra: single reference */
      ip->compound = 0;
      ip->ref[1] = -1;
      ip->interintra = true;
      ip->interintra_mode = rng_.range(0, 3);
      ip->interintra_wedge = rng_.chance(0.5);
      ip->wedge_index = rng_.range(0, 15);
    } else if (tool == 3) {
      /* OBMC: single reference; the above neighbours in 8 wide strips of
       * half the block height, the left ones in 8 tall strips of half its width */
      ip->compound = 0;
      ip->ref[1] = -1;
      for (int rx = 0; rx < w; rx += 8) ip->obmc.push_back(obmc_candidate(0, rx, 0, 8, h / 2));
      for (int ry = 0; ry < h; ry += 8) ip->obmc.push_back(obmc_candidate(1, 0, ry, w / 2, 8));
    } else {
      /* global warp of the first reference, close to the identity */
      ip->warp = true;
      int32_t* p = ip->warp_params;
      p[0] = rng_.range(-8 << 16, 8 << 16);
      p[1] = rng_.range(-8 << 16, 8 << 16);
      p[2] = (1 << 16) + rng_.range(-3000, 3000);
      p[3] = rng_.range(-2000, 2000);
      p[4] = rng_.range(-2000, 2000);
      p[5] = (1 << 16) + rng_.range(-3000, 3000);
      if (ip->compound > 2) ip->compound = 1;
    }
  }

  ObmcCandidate obmc_candidate(int pass, int x, int y, int w, int h) {
    ObmcCandidate c;
    c.pass = pass;
    c.x = x;
    c.y = y;
    c.w = w;
    c.h = h;
    c.mv[0] = draw_mv_component();
    c.mv[1] = draw_mv_component();
    if (cfg_.all_filters) {
      c.filter = static_cast<FilterType>(rng_.range(0, 2));
      c.filter_v = static_cast<FilterType>(rng_.range(0, 2));
    }
    c.ref = rng_.range(0, nrefs_ - 1);
    return c;
  }

  /* Intra block copy from the superblock row above (key frames): every plane
   * copies the co-located block 64 luma rows up, shifted by a whole number of
   * chroma samples sideways. */
  void intrabc_block(int x, int y, int w, int h, int level, bool raw) {
    const int span = grid_w_ - w;
    const int sx = (rng_.range(0, span) / 16) * 16;
    InterParams ip;
    ip.ref[0] = -1;
    ip.ref[1] = -1;
    ip.mv[0][0] = (sx - x) * 8;
    ip.mv[0][1] = -64 * 8;
    ip.intrabc = true;
    for (int p = 0; p < 3; ++p) {
      const int s = p ? 1 : 0;
      inter_block(p, x >> s, y >> s, w >> s, h >> s, ip, level, raw);
    }
  }

  void intra_block(int plane, int x, int y, int w, int h, int mode, int angle, int level,
                   bool lossless, int cfl_alpha, bool raw) {
    PlaneState& ps = planes_[plane];
    int tw, th;
    tx_size_for(plane, w, h, true, &tw, &th);
    if (lossless) tw = th = 4;
    /* filterType (the specification's get_filter_type): a smooth neighbour above
     * or to the left of the coding block */
    int filter_type = 0;
    if (y > 0 && ps.smooth[ps.idx(x, y - 1)]) filter_type = 1;
    if (x > 0 && ps.smooth[ps.idx(x - 1, y)]) filter_type = 1;
    const bool use_palette = cfg_.contract_tools && plane == 0 && w <= 32 && h <= 32 &&
                             rng_.chance(0.05);
    const bool use_fi = cfg_.contract_tools && plane == 0 && !use_palette && w <= 32 &&
                        h <= 32 && rng_.chance(0.15);
    int palette[8] = {0};
    int palette_size = 0;
    int palette_base = 0;
    if (use_palette) {
      palette_size = rng_.range(2, 8);
      const int maxv = (1 << cfg_.bit_depth) - 1;
      for (int i = 0; i < palette_size; ++i) palette[i] = rng_.range(0, maxv);
      palette_base = static_cast<int>(fs_.palette_map.size());
      for (int i = 0; i < w * h; ++i)
        fs_.palette_map.push_back(static_cast<uint8_t>(rng_.range(0, palette_size - 1)));
    }
    const int fi_mode = rng_.range(0, 4);
    for (int ty = 0; ty < h; ty += th) {
      for (int tx = 0; tx < w; tx += tw) {
        BlockSyntax b;
        b.x = x + tx;
        b.y = y + ty;
        b.w = tw;
        b.h = th;
        b.plane = plane;
        b.bit_depth = cfg_.bit_depth;
        b.is_intra = true;
        b.mode = static_cast<PredictionMode>(mode);
        b.angle_delta = angle;
        b.have_left = b.x > 0;
        b.have_above = b.y > 0;
        b.have_above_right =
            (b.y > 0 && ps.rect_decoded(b.x + b.w, b.x + 2 * b.w - 1, b.y - 1, b.y - 1)) ? 1 : 0;
        b.have_below_left =
            (b.x > 0 && ps.rect_decoded(b.x - 1, b.x - 1, b.y + b.h, b.y + 2 * b.h - 1)) ? 1 : 0;
        /* left to the decoder: -1 asks it to derive the flag from coding order */
        if (cfg_.avail_auto > 0 && rng_.chance(cfg_.avail_auto)) b.have_above_right = -1;
        if (cfg_.avail_auto > 0 && rng_.chance(cfg_.avail_auto)) b.have_below_left = -1;
        b.residual_provided = raw;
        b.filter_type = filter_type;
        b.lossless = lossless;
        b.cfl_alpha = cfl_alpha;
        if (use_fi) {
          b.use_filter_intra = true;
          b.filter_intra_mode = static_cast<FilterIntraMode>(fi_mode);
        }
        if (use_palette) {
          b.use_palette = true;
          b.palette_size = palette_size;
          for (int i = 0; i < 8; ++i) b.palette[i] = palette[i];
          b.palette_off = palette_base + ty * w + tx;
          b.palette_stride = w;
        }
        b.transforms.push_back(make_tx(0, 0, tw, th, lossless, raw));
        fs_.blocks.push_back(std::move(b));
        ps.mark(x + tx, y + ty, tw, th, is_smooth_mode(mode), tw, th, level);
        ps.mark_tx(x + tx, y + ty, tw, th);
      }
    }
  }



  int draw_mv_component() {
    if (cfg_.mv_gaussian) {
      long v = std::lround(rng_.gauss() * cfg_.mv_sigma);
      return static_cast<int>(std::max(-2048L, std::min(2048L, v)));
    }
    return rng_.range(-cfg_.mv_range, cfg_.mv_range);
  }

  InterParams draw_inter() {
    InterParams ip;
    ip.ref[0] = rng_.range(0, nrefs_ - 1);
    for (int i = 0; i < 2; ++i) {
      ip.mv[i][0] = draw_mv_component();
      ip.mv[i][1] = draw_mv_component();
    }
    if (cfg_.all_filters) {
      ip.filt = rng_.range(0, 2);
      ip.filt_v = rng_.range(0, 2);
    }
    if (cfg_.compound_frac > 0 && rng_.chance(cfg_.compound_frac)) {
      ip.compound = rng_.chance(0.5) ? 1 : 2;
      if (nrefs_ >= 2) {
        int r1 = rng_.range(0, nrefs_ - 2);
        if (r1 >= ip.ref[0]) ++r1;
        ip.ref[1] = r1;
      } else {
        ip.ref[1] = ip.ref[0];
      }
      if (ip.compound == 2) {
        static const int pairs[4][2] = {{9, 7}, {11, 5}, {12, 4}, {13, 3}};
        const int k = rng_.range(0, 3);
        const bool swap = rng_.chance(0.5);
        ip.blend[0] = pairs[k][swap ? 1 : 0];
        ip.blend[1] = pairs[k][swap ? 0 : 1];
      }
    }
    return ip;
  }

  void inter_block(int plane, int x, int y, int w, int h, const InterParams& ip, int level,
                   bool raw) {
    PlaneState& ps = planes_[plane];
    int tw, th;
    tx_size_for(plane, w, h, false, &tw, &th);
    BlockSyntax b;
    b.x = x;
    b.y = y;
    b.w = w;
    b.h = h;
    b.plane = plane;
    b.bit_depth = cfg_.bit_depth;
    b.is_intra = false;
    b.have_left = x > 0;
    b.have_above = y > 0;
    b.ref[0] = ip.ref[0];
    b.ref[1] = ip.ref[1];
    for (int i = 0; i < 2; ++i) {
      b.mv[i][0] = ip.mv[i][0];
      b.mv[i][1] = ip.mv[i][1];
    }
    b.filter = static_cast<FilterType>(ip.filt);
    b.filter_v = static_cast<FilterType>(ip.filt_v);
    b.compound = static_cast<CompoundType>(ip.compound);
    b.blend[0] = ip.blend[0];
    b.blend[1] = ip.blend[1];
    b.wedge_index = ip.wedge_index;
    b.mask_sign = ip.mask_sign;
    b.interintra = ip.interintra;
    b.interintra_mode = ip.interintra_mode;
    b.interintra_wedge = ip.interintra_wedge;
    b.intrabc = ip.intrabc;
    b.residual_provided = raw;
    if (ip.warp) {
      b.warp[0] = true;
      for (int k = 0; k < 6; ++k) b.warp_params[0][k] = ip.warp_params[k];
    }
    const int sh = plane ? 1 : 0;
    for (const ObmcCandidate& c : ip.obmc) {
      ObmcCandidate o = c;
      o.x >>= sh;
      o.y >>= sh;
      o.w >>= sh;
      o.h >>= sh;
      b.obmc.push_back(o);
    }
    for (int ty = 0; ty < h; ty += th)
      for (int tx = 0; tx < w; tx += tw) {
        b.transforms.push_back(make_tx(tx, ty, tw, th, false, raw));
        ps.mark_tx(x + tx, y + ty, tw, th);
      }
    fs_.blocks.push_back(std::move(b));
    ps.mark(x, y, w, h, false, tw, th, level);
  }

  /* Deblocking edges on the transform grid. filterSize is the smaller of the
   * two transform sides across the edge, at most 16 for luma and 8 for chroma
   * (which the filter turns into its 6 tap variant). Runs of 4 sample segments
   * with the same size and level are merged into one edge. */
  void build_deblock() {
    for (int plane = 0; plane < 3; ++plane) {
      const PlaneState& ps = planes_[plane];
      const int cap = plane ? 8 : 16;
      for (int dir = 0; dir < 2; ++dir) {
        const bool vertical = dir == 0;
        const int outer = vertical ? ps.cw : ps.ch; /* edge position, in cells  */
        const int inner = vertical ? ps.ch : ps.cw; /* along the edge, in cells */
        for (int o = 1; o < outer; ++o) {
          int run_start = -1, run_fs = 0, run_lvl = 0;
          auto flush = [&](int end) {
            if (run_start < 0) return;
            DeblockEdge e;
            e.plane = plane;
            e.vertical = vertical;
            e.x = vertical ? o * 4 : run_start * 4;
            e.y = vertical ? run_start * 4 : o * 4;
            e.length = (end - run_start) * 4;
            e.level = run_lvl;
            e.filter_size = run_fs;
            fs_.deblock.push_back(e);
            run_start = -1;
          };
          for (int i = 0; i < inner; ++i) {
            const int qx = vertical ? o * 4 : i * 4;
            const int qy = vertical ? i * 4 : o * 4;
            const int px = vertical ? qx - 4 : qx;
            const int py = vertical ? qy : qy - 4;
            const size_t q = ps.idx(qx, qy), p = ps.idx(px, py);
            const bool edge = vertical ? (ps.tx_x0[q] == qx) : (ps.tx_y0[q] == qy);
            if (!edge) {
              flush(i);
              continue;
            }
            int fsz = vertical ? std::min(ps.txw[q], ps.txw[p]) : std::min(ps.txh[q], ps.txh[p]);
            fsz = std::min(fsz, cap);
            const int lvl = ps.level[q] ? ps.level[q] : ps.level[p];
            if (run_start >= 0 && (fsz != run_fs || lvl != run_lvl || i - run_start >= 16))
              flush(i);
            if (run_start < 0) {
              run_start = i;
              run_fs = fsz;
              run_lvl = lvl;
            }
          }
          flush(inner);
        }
      }
    }
  }

  void build_cdef() {
    const int damping = rng_.range(3, 6);
    const int sb_cols = (grid_w_ + 63) / 64;
    const int sb_rows = (grid_h_ + 63) / 64;
    struct Strength {
      int skip, ypri, ysec, uvpri, uvsec;
    };
    std::vector<Strength> sb(static_cast<size_t>(sb_cols) * sb_rows);
    static const int secs[4] = {0, 1, 2, 4};
    for (Strength& s : sb) {
      s.skip = rng_.chance(0.1);
      s.ypri = rng_.range(0, 15);
      s.ysec = secs[rng_.range(0, 3)];
      s.uvpri = rng_.range(0, 15);
      s.uvsec = secs[rng_.range(0, 3)];
    }
    for (int y = 0; y < grid_h_; y += 8)
      for (int x = 0; x < grid_w_; x += 8) {
        const Strength& s = sb[static_cast<size_t>(y / 64) * sb_cols + x / 64];
        if (s.skip) continue;
        CdefBlock c;
        c.plane = 0;
        c.x = x;
        c.y = y;
        c.pri_strength = s.ypri;
        c.sec_strength = s.ysec;
        c.damping = damping;
        c.dir = -1;
        fs_.cdef.push_back(c);
        for (int p = 1; p < 3; ++p) {
          CdefBlock u = c;
          u.plane = p;
          u.x = x >> 1;
          u.y = y >> 1;
          u.pri_strength = s.uvpri;
          u.sec_strength = s.uvsec;
          fs_.cdef.push_back(u);
        }
      }
  }

  static int count_units(int unit, int size) { return std::max((size + (unit >> 1)) / unit, 1); }

  void build_lr() {
    for (int plane = 0; plane < 3; ++plane) {
      const int s = plane ? 1 : 0;
      const int l0 = cfg_.lr_log2 > 0 ? cfg_.lr_log2 : 6;
      const int log2 = plane ? std::max(5, l0 - 1) : l0;
      const int unit = 1 << log2;
      const int pw = (cfg_.width + s) >> s;
      const int ph = (cfg_.height + s) >> s;
      const int cols = count_units(unit, pw), rows = count_units(unit, ph);
      fs_.restoration_type[plane] = cfg_.lr == 1 ? kRestoreWiener : kRestoreSwitchable;
      fs_.restoration_unit_log2[plane] = log2;
      fs_.restoration_unit_cols[plane] = cols;
      fs_.restoration[plane].resize(static_cast<size_t>(cols) * rows);
      for (RestorationUnit& u : fs_.restoration[plane]) {
        int type;
        if (cfg_.lr == 1) {
          type = rng_.chance(0.85) ? 1 : 0;
        } else {
          const double r = rng_.uni();
          type = r < 0.4 ? 1 : (r < 0.85 ? 2 : 0);
        }
        /* unit types use the RestorationType values, as the reference planner
         * reads them (kRestoreWiener, kRestoreSgrproj) */
        u.type = type == 1 ? kRestoreWiener : (type == 2 ? kRestoreSgrproj : kRestoreNone);
        if (type == 1) {
          for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < 3; ++i) u.wiener[pass][i] = rng_.range(kWienerTapsMin[i], kWienerTapsMax[i]);
        } else if (type == 2) {
          u.sgr_set = rng_.range(0, 15);
          u.sgr_xqd[0] = rng_.range(kSgrXqdMin[0], kSgrXqdMax[0]);
          u.sgr_xqd[1] = rng_.range(kSgrXqdMin[1], kSgrXqdMax[1]);
          if (kSgrParams[u.sgr_set][0] == 0) u.sgr_xqd[0] = 0;
          if (kSgrParams[u.sgr_set][2] == 0)
            u.sgr_xqd[1] = std::max(kSgrXqdMin[1], std::min(kSgrXqdMax[1], 128 - u.sgr_xqd[0]));
        }
      }
    }
  }

  void points(int n, int* value, int* scaling) {
    /* strictly increasing values over 0..255 */
    int v = rng_.range(0, 32);
    for (int i = 0; i < n; ++i) {
      value[i] = v;
      scaling[i] = rng_.range(0, 255);
      v += rng_.range(1, std::max(1, (255 - v) / std::max(1, n - i)));
      if (v > 255) v = 255;
      if (i + 1 < n && v <= value[i]) v = value[i] + 1;
    }
  }

  void build_grain() {
    FilmGrainParams& g = fs_.grain;
    g.apply = true;
    g.seed = rng_.range(0, 65535);
    g.num_y_points = rng_.range(4, 10);
    points(g.num_y_points, g.point_y_value, g.point_y_scaling);
    g.chroma_scaling_from_luma = rng_.chance(0.25);
    if (!g.chroma_scaling_from_luma) {
      g.num_cb_points = rng_.range(1, 6);
      g.num_cr_points = rng_.range(1, 6);
      points(g.num_cb_points, g.point_cb_value, g.point_cb_scaling);
      points(g.num_cr_points, g.point_cr_value, g.point_cr_scaling);
    }
    g.grain_scaling_minus_8 = rng_.range(0, 3);
    g.ar_coeff_lag = 3;
    for (int i = 0; i < 24; ++i) g.ar_coeffs_y[i] = rng_.range(-24, 24);
    for (int i = 0; i < 25; ++i) {
      g.ar_coeffs_cb[i] = rng_.range(-24, 24);
      g.ar_coeffs_cr[i] = rng_.range(-24, 24);
    }
    g.ar_coeff_shift_minus_6 = rng_.range(0, 3);
    g.grain_scale_shift = rng_.range(0, 1);
    g.cb_mult = rng_.range(0, 255);
    g.cb_luma_mult = rng_.range(0, 255);
    g.cb_offset = rng_.range(0, 511);
    g.cr_mult = rng_.range(0, 255);
    g.cr_luma_mult = rng_.range(0, 255);
    g.cr_offset = rng_.range(0, 511);
    g.overlap_flag = rng_.chance(0.5);
    g.clip_to_restricted_range = rng_.chance(0.5);
  }
};

/* Decode order and references. One reference: each frame reads the one before.
 * Three: mini-GOPs of `gop` frames coded as a dyadic pyramid (anchor first,
 * then the midpoints), every frame reading up to three already decoded frames
 * nearest to it in display order. */
void plan_order(const Config& c, std::vector<std::vector<int>>* refs, std::vector<int>* key) {
  const int n = c.frames;
  refs->assign(n, {});
  key->assign(n, 0);
  if (n <= 0) return;
  if (c.num_refs <= 1) {
    for (int f = 0; f < n; ++f) {
      const bool k = f == 0 || (c.key_interval > 0 && f % c.key_interval == 0);
      (*key)[f] = k;
      if (!k) (*refs)[f] = {f - 1};
    }
    return;
  }
  /* display order positions in decode order */
  std::vector<int> order;
  order.push_back(0);
  int base = 0;
  while (base + 1 < n) {
    const int end = std::min(base + c.gop, n - 1);
    order.push_back(end);
    std::vector<std::pair<int, int>> spans = {{base, end}};
    while (!spans.empty()) {
      std::vector<std::pair<int, int>> next;
      for (auto& s : spans) {
        if (s.second - s.first < 2) continue;
        const int mid = (s.first + s.second) / 2;
        order.push_back(mid);
        next.push_back({s.first, mid});
        next.push_back({mid, s.second});
      }
      spans.swap(next);
    }
    base = end;
  }
  /* decode index of each display position */
  std::vector<int> dec_of(n, -1);
  for (int d = 0; d < static_cast<int>(order.size()); ++d) dec_of[order[d]] = d;
  for (int d = 0; d < static_cast<int>(order.size()); ++d) {
    const int disp = order[d];
    if (d == 0) {
      (*key)[d] = 1;
      continue;
    }
    std::vector<std::pair<int, int>> cand; /* (distance, decode index) */
    for (int e = 0; e < d; ++e) cand.push_back({std::abs(order[e] - disp), e});
    std::sort(cand.begin(), cand.end());
    for (int i = 0; i < std::min(c.num_refs, static_cast<int>(cand.size())); ++i)
      (*refs)[d].push_back(cand[i].second);
  }
}

}  // namespace

Config workload(const std::string& name, uint64_t seed) {
  Config c;
  c.seed = seed;
  if (name == "c1" || name == "c1_mini") {
    c.width = 1280;
    c.height = 720;
    c.bit_depth = 8;
    c.frames = 30;
    c.max_tx = 32;
    c.lr = 1;
    if (name == "c1_mini") {
      c.width = 160;
      c.height = 96;
      c.frames = 4;
    }
  } else if (name == "c2" || name == "c2_mini") {
    c.width = 1920;
    c.height = 1080;
    c.coded_width = 1440;
    c.bit_depth = 10;
    c.frames = 60;
    c.num_refs = 3;
    c.gop = 16;
    c.max_tx = 64;
    c.rect_tx = true;
    c.tx_ext = true;
    c.mv_gaussian = true;
    c.mv_sigma = 192.0;
    c.all_filters = true;
    c.compound_frac = 0.5;
    c.lr = 2;
    c.grain = true;
    if (name == "c2_mini") {
      c.width = 256;
      c.height = 144;
      c.coded_width = 192;
      c.frames = 5;
      c.gop = 4;
      c.mv_sigma = 48.0;
    }
  } else if (name == "tools_mini") {
    /* every inter tool of the contract, intra block copy, 10 bit */
    c.width = 192;
    c.height = 144;
    c.bit_depth = 10;
    c.frames = 4;
    c.num_refs = 3;
    c.gop = 4;
    c.max_tx = 32;
    c.rect_tx = true;
    c.tx_ext = true;
    c.all_filters = true;
    c.mv_gaussian = true;
    c.mv_sigma = 48.0;
    c.compound_frac = 0.3;
    c.intra_frac = 0.1;
    c.inter_tools = 0.6;
    c.intrabc = 0.25;
    c.lr = 2;
    c.grain = true;
    c.contract_tools = true;
  } else if (name == "contract_mini") {
    c.width = 136;
    c.height = 72;
    c.bit_depth = 8;
    c.frames = 3;
    c.max_tx = 64;
    c.rect_tx = true;
    c.tx_ext = true;
    c.intra_frac = 0.3;
    c.lr = 2;
    c.contract_tools = true;
  } else {
    throw std::invalid_argument("unknown workload " + name);
  }
  return c;
}

FrameSyntax generate_frame(const Config& config, int f, const std::vector<int>& refs, bool key) {
  FrameBuilder b(config, f, refs, key);
  return b.build();
}

Clip generate(const Config& config) {
  Clip clip;
  clip.config = config;
  std::vector<std::vector<int>> refs;
  std::vector<int> key;
  plan_order(config, &refs, &key);
  clip.refs = refs;
  clip.key = key;
  clip.last_use.assign(config.frames, -1);
  for (int f = 0; f < config.frames; ++f)
    for (int r : refs[f]) clip.last_use[r] = std::max(clip.last_use[r], f);
  clip.frames.reserve(config.frames);
  for (int f = 0; f < config.frames; ++f)
    clip.frames.push_back(generate_frame(config, f, refs[f], key[f] != 0));
  return clip;
}

}  // namespace synth
}  // namespace av1gpu
