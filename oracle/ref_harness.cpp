/*
 * Reference harness (TEST INFRASTRUCTURE ONLY, built into oracle/_ref/ and
 * never shipped): runs the reference's own CPU backend
 * (src/backends/cpu/cpu_backend.cpp, compiled from /root/reference in place,
 * with its kernels/av1gpu_kernels.h, src/core/tables.cpp, grain.cpp and
 * factory.cpp) on wire frames, to produce golden outputs.
 *
 * The reference's planner (src/core/engine.cpp) cannot be compiled here (it
 * calls Win32 timing functions outside _WIN32 guards), so this file builds the
 * reference's internal::FramePlan itself, transcribing build_frame_plan
 * (engine.cpp:519-1494) field for field. `serial` puts every block in a wave of
 * its own, in coding order - the specification's sequential decode - which the
 * golden script uses to check that the wavefront schedule changes nothing.
 */
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "av1hip.h"
#include "core/backend.hpp"
#include "frontend/coef_tables.hpp"

using namespace av1gpu;
using namespace av1gpu::internal;

namespace {

void syntax_from_desc(const av1h_frame_desc& d, FrameSyntax* s) {
  const av1h_frame_header& h = d.hdr;
  s->width = h.width;
  s->coded_width = h.coded_width;
  s->height = h.height;
  s->bit_depth = h.bit_depth;
  s->ss_x = h.ss_x;
  s->ss_y = h.ss_y;
  s->base_q_idx = h.base_q_idx;
  s->sharpness = h.sharpness;
  s->enable_intra_edge_filter = h.enable_intra_edge_filter != 0;
  s->reference_slots.assign(h.reference_slots, h.reference_slots + h.num_reference_slots);
  for (int i = 0; i < 7; ++i) {
    s->reference_width[i] = h.reference_width[i];
    s->reference_height[i] = h.reference_height[i];
  }
  s->color.matrix = static_cast<ColorMatrix>(h.color_matrix);
  s->color.range = static_cast<ColorRange>(h.color_range);
  for (int p = 0; p < 3; ++p) {
    s->restoration_type[p] = h.restoration_type[p];
    s->restoration_unit_log2[p] = h.restoration_unit_log2[p];
    s->restoration_unit_cols[p] = h.restoration_unit_cols[p];
  }
  const av1h_grain& g = h.grain;
  FilmGrainParams& o = s->grain;
  o.apply = g.apply != 0;
  o.seed = g.seed;
  o.num_y_points = g.num_y_points;
  o.chroma_scaling_from_luma = g.chroma_scaling_from_luma != 0;
  o.num_cb_points = g.num_cb_points;
  o.num_cr_points = g.num_cr_points;
  for (int i = 0; i < 16; ++i) {
    o.point_y_value[i] = g.point_y_value[i];
    o.point_y_scaling[i] = g.point_y_scaling[i];
    o.point_cb_value[i] = g.point_cb_value[i];
    o.point_cb_scaling[i] = g.point_cb_scaling[i];
    o.point_cr_value[i] = g.point_cr_value[i];
    o.point_cr_scaling[i] = g.point_cr_scaling[i];
  }
  o.grain_scaling_minus_8 = g.grain_scaling_minus_8;
  o.ar_coeff_lag = g.ar_coeff_lag;
  for (int i = 0; i < 24; ++i) o.ar_coeffs_y[i] = g.ar_coeffs_y[i];
  for (int i = 0; i < 25; ++i) {
    o.ar_coeffs_cb[i] = g.ar_coeffs_cb[i];
    o.ar_coeffs_cr[i] = g.ar_coeffs_cr[i];
  }
  o.ar_coeff_shift_minus_6 = g.ar_coeff_shift_minus_6;
  o.grain_scale_shift = g.grain_scale_shift;
  o.cb_mult = g.cb_mult;
  o.cb_luma_mult = g.cb_luma_mult;
  o.cb_offset = g.cb_offset;
  o.cr_mult = g.cr_mult;
  o.cr_luma_mult = g.cr_luma_mult;
  o.cr_offset = g.cr_offset;
  o.overlap_flag = g.overlap_flag != 0;
  o.clip_to_restricted_range = g.clip_to_restricted_range != 0;
  s->blocks.resize(static_cast<size_t>(d.num_blocks));
  for (int64_t i = 0; i < d.num_blocks; ++i) {
    const av1h_block& b = d.blocks[i];
    BlockSyntax& q = s->blocks[static_cast<size_t>(i)];
    q.x = b.x;
    q.y = b.y;
    q.w = b.w;
    q.h = b.h;
    q.plane = b.plane;
    q.bit_depth = b.bit_depth;
    q.is_intra = b.is_intra != 0;
    q.intrabc = b.intrabc != 0;
    q.mode = static_cast<PredictionMode>(b.mode);
    q.have_left = b.have_left != 0;
    q.have_above = b.have_above != 0;
    q.have_above_right = b.have_above_right;
    q.have_below_left = b.have_below_left;
    q.angle_delta = b.angle_delta;
    q.filter_type = b.filter_type;
    q.use_filter_intra = b.use_filter_intra != 0;
    q.filter_intra_mode = static_cast<FilterIntraMode>(b.filter_intra_mode);
    q.cfl_alpha = b.cfl_alpha;
    q.lossless = b.lossless != 0;
    q.use_palette = b.use_palette != 0;
    q.palette_size = b.palette_size;
    for (int k = 0; k < 8; ++k) q.palette[k] = b.palette[k];
    q.palette_off = b.palette_off;
    q.palette_stride = b.palette_stride;
    for (int r = 0; r < 2; ++r) {
      q.ref[r] = b.ref[r];
      q.mv[r][0] = b.mv[r][0];
      q.mv[r][1] = b.mv[r][1];
      q.blend[r] = b.blend[r];
      q.warp[r] = b.warp[r] != 0;
      q.mask_warp[r] = b.mask_warp[r] != 0;
      for (int k = 0; k < 6; ++k) q.warp_params[r][k] = b.warp_params[r][k];
    }
    q.filter = static_cast<FilterType>(b.filter);
    q.filter_v = static_cast<FilterType>(b.filter_v);
    q.motion_mode = b.motion_mode;
    q.compound = static_cast<CompoundType>(b.compound);
    q.wedge_index = b.wedge_index;
    q.mask_sign = b.mask_sign;
    q.interintra = b.interintra != 0;
    q.interintra_mode = b.interintra_mode;
    q.interintra_wedge = b.interintra_wedge != 0;
    q.residual_provided = b.residual_provided != 0;
    for (int t = 0; t < b.tx_count; ++t) {
      const av1h_transform& r = d.transforms[b.tx_first + t];
      TransformBlock tb;
      tb.x = r.x;
      tb.y = r.y;
      tb.w = r.w;
      tb.h = r.h;
      tb.coef_offset = r.coef_offset;
      tb.eob = r.eob;
      tb.tx_type = static_cast<TransformType>(r.tx_type);
      tb.level_offset = r.level_offset;
      tb.qindex = r.qindex;
      tb.dc_step = r.dc_step;
      tb.ac_step = r.ac_step;
      tb.denominator = r.denominator;
      tb.qm_off = r.qm_off;
      q.transforms.push_back(tb);
    }
    for (int k = 0; k < b.obmc_count; ++k) {
      const av1h_obmc& r = d.obmc[b.obmc_first + k];
      ObmcCandidate c;
      c.pass = r.pass;
      c.x = r.x;
      c.y = r.y;
      c.w = r.w;
      c.h = r.h;
      c.mv[0] = r.mv_x;
      c.mv[1] = r.mv_y;
      c.filter = static_cast<FilterType>(r.filter);
      c.filter_v = static_cast<FilterType>(r.filter_v);
      c.ref = r.ref;
      q.obmc.push_back(c);
    }
  }
  s->coefficients.assign(d.coefficients, d.coefficients + d.num_coefficients);
  s->residual.assign(d.residual, d.residual + d.num_residual);
  s->palette_map.assign(d.palette_map, d.palette_map + d.num_palette);
  for (int64_t i = 0; i < d.num_deblock; ++i) {
    const av1h_deblock_edge& e = d.deblock[i];
    DeblockEdge o2;
    o2.plane = e.plane;
    o2.x = e.x;
    o2.y = e.y;
    o2.vertical = e.vertical != 0;
    o2.length = e.length;
    o2.level = e.level;
    o2.filter_size = e.filter_size;
    s->deblock.push_back(o2);
  }
  for (int64_t i = 0; i < d.num_cdef; ++i) {
    const av1h_cdef_block& c = d.cdef[i];
    CdefBlock o3;
    o3.plane = c.plane;
    o3.x = c.x;
    o3.y = c.y;
    o3.pri_strength = c.pri_strength;
    o3.sec_strength = c.sec_strength;
    o3.damping = c.damping;
    o3.dir = c.dir;
    s->cdef.push_back(o3);
  }
  for (int p = 0; p < 3; ++p)
    for (int64_t i = 0; i < d.num_restoration[p]; ++i) {
      const av1h_lr_unit& r = d.restoration[p][i];
      RestorationUnit u;
      u.type = r.type;
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k) u.wiener[a][k] = r.wiener[a][k];
      u.sgr_set = r.sgr_set;
      u.sgr_xqd[0] = r.sgr_xqd[0];
      u.sgr_xqd[1] = r.sgr_xqd[1];
      s->restoration[p].push_back(u);
    }
}

int mode_angle(PredictionMode m) {
  static const int a[] = {0, 90, 180, 45, 135, 113, 157, 203, 67, 0, 0, 0, 0};
  const int i = static_cast<int>(m);
  return (i < 0 || i >= 13) ? 0 : a[i];
}

int slot_of(const FrameSyntax& s, int ref) {
  if (ref < 0 || ref >= static_cast<int>(s.reference_slots.size())) return -1;
  return s.reference_slots[ref];
}

/* build_frame_plan (engine.cpp:519-1494), transcribed for this harness */
bool build_plan(const FrameSyntax& syntax, const PlanLayout& g, int capacity, bool serial,
                FramePlan* plan, std::string* err) {
  internal::FrameGeometry geom;
  geom.width = syntax.width;
  geom.height = syntax.height;
  geom.bit_depth = syntax.bit_depth;
  geom.ss_x = syntax.ss_x;
  geom.ss_y = syntax.ss_y;
  plan->layout = g;
  plan->maxv = (1 << geom.bit_depth) - 1;
  const int frame_w = g.plane_w[0], frame_h = g.plane_h[0];
  auto at = [&](int p, int x, int y) { return static_cast<int>(g.plane_off[p] + y * g.stride[p] + x); };

  size_t coef_size = 0, residual_size = 0;
  for (const BlockSyntax& b : syntax.blocks)
    for (const TransformBlock& tx : b.transforms) {
      const size_t end = tx.coef_offset + static_cast<size_t>(tx.w) * tx.h;
      if (b.residual_provided)
        residual_size = std::max(residual_size, end);
      else
        coef_size = std::max(coef_size, end);
    }
  plan->coefficients = syntax.coefficients;
  plan->residual = syntax.residual;
  plan->coef_pool_size = coef_size + 64;
  plan->residual_pool_size = residual_size + 64;
  plan->coefficients.resize(std::max(plan->coefficients.size(), plan->coef_pool_size), 0);
  plan->residual.resize(std::max(plan->residual.size(), plan->residual_pool_size), 0);
  plan->palette_map.assign(syntax.palette_map.begin(), syntax.palette_map.end());
  plan->palette_map.resize(plan->palette_map.size() + 8, 0);
  for (int i = 0; i < 2; ++i) {
    plan->ref_slot[i] = -1;
    if (i < static_cast<int>(syntax.reference_slots.size())) {
      const int s = syntax.reference_slots[i];
      if (s >= 0 && s < capacity) plan->ref_slot[i] = s;
    }
  }

  int grid_w[3], grid_h[3];
  for (int p = 0; p < 3; ++p) {
    const int sx = p > 0 && geom.ss_x ? 1 : 0, sy = p > 0 && geom.ss_y ? 1 : 0;
    grid_w[p] = ((frame_w + 7) & ~7) >> sx;
    grid_h[p] = ((frame_h + 7) & ~7) >> sy;
  }
  const int oc_[3] = {4, geom.ss_x ? 2 : 4, geom.ss_y ? 2 : 4};
  const int osh[3] = {2, geom.ss_x ? 1 : 2, geom.ss_y ? 1 : 2};
  int cw[3], ch[3];
  for (int p = 0; p < 3; ++p) {
    cw[p] = (grid_w[p] + oc_[p] - 1) / oc_[p];
    ch[p] = (grid_h[p] + oc_[p] - 1) / oc_[p];
  }
  std::vector<std::vector<int>> owner(3);
  bool need[3] = {false, false, false};
  for (const BlockSyntax& b : syntax.blocks) {
    if (b.mode == PredictionMode::ChromaFromLuma) need[0] = true;
    if (b.is_intra || b.intrabc || b.interintra) need[b.plane] = true;
  }
  for (int p = 0; p < 3; ++p)
    if (need[p]) owner[p].assign(static_cast<size_t>(cw[p]) * ch[p], -1);
  for (size_t bi = 0; bi < syntax.blocks.size(); ++bi) {
    const BlockSyntax& b = syntax.blocks[bi];
    const int p = b.plane;
    if (!need[p]) continue;
    const int oc = oc_[p];
    const int cx1 = std::min((b.x + b.w + oc - 1) / oc, cw[p]);
    const int cy1 = std::min((b.y + b.h + oc - 1) / oc, ch[p]);
    for (int cy = std::max(b.y, 0) / oc; cy < cy1; ++cy)
      for (int cx = std::max(b.x, 0) / oc; cx < cx1; ++cx)
        owner[p][static_cast<size_t>(cy) * cw[p] + cx] = static_cast<int>(bi);
  }
  auto sample_owner = [&](int p, int x, int y) {
    if (x < 0 || y < 0 || x >= grid_w[p] || y >= grid_h[p]) return -1;
    if (owner[p].empty()) return -1;
    return owner[p][static_cast<size_t>(y >> osh[p]) * cw[p] + (x >> osh[p])];
  };
  auto decoded = [&](int p, int x0, int x1, int y0, int y1, size_t bi) {
    const int oc = 1 << osh[p];
    for (int y = y0 & ~(oc - 1); y <= y1; y += oc)
      for (int x = x0 & ~(oc - 1); x <= x1; x += oc) {
        const int o = sample_owner(p, x, y);
        if (o >= 0 && static_cast<size_t>(o) >= bi) return false;
      }
    return true;
  };
  const size_t nb = syntax.blocks.size();
  std::vector<int> wave(nb, 0);
  std::vector<char> ar(nb, 0), bl(nb, 0);
  int max_wave = 0;
  for (size_t bi = 0; bi < nb; ++bi) {
    const BlockSyntax& b = syntax.blocks[bi];
    const int p = b.plane;
    int dep = -1;
    if (b.is_intra || b.interintra) {
      if (b.is_intra) {
        ar[bi] = static_cast<char>(b.have_above_right >= 0
                                       ? (b.have_above_right != 0)
                                       : (b.y > 0 && decoded(p, b.x + b.w, b.x + 2 * b.w - 1,
                                                             b.y - 1, b.y - 1, bi)));
        bl[bi] = static_cast<char>(b.have_below_left >= 0
                                       ? (b.have_below_left != 0)
                                       : (b.x > 0 && decoded(p, b.x - 1, b.x - 1, b.y + b.h,
                                                             b.y + 2 * b.h - 1, bi)));
      }
      int left_y1 = b.y + b.h - 1, above_x1 = b.x + b.w - 1;
      if (b.is_intra && !b.use_palette) {
        const int mode = static_cast<int>(b.mode);
        if (mode >= 1 && mode <= 8) {
          const int pangle = mode_angle(b.mode) + 3 * b.angle_delta;
          if (pangle < 90 && ar[bi]) above_x1 = b.x + 2 * b.w - 1;
          if (pangle > 180 && bl[bi]) left_y1 = b.y + 2 * b.h - 1;
        }
      } else if (b.is_intra) {
        above_x1 = b.x - 1;
        left_y1 = b.y - 1;
      }
      const int oc = 1 << osh[p];
      for (int y = b.y & ~(oc - 1); y <= left_y1; y += oc) {
        const int o = sample_owner(p, b.x - 1, y);
        if (o >= 0) dep = std::max(dep, wave[static_cast<size_t>(o)]);
      }
      for (int x = b.x & ~(oc - 1); x <= above_x1; x += oc) {
        const int o = sample_owner(p, x, b.y - 1);
        if (o >= 0) dep = std::max(dep, wave[static_cast<size_t>(o)]);
      }
      if (above_x1 >= b.x || left_y1 >= b.y) {
        const int o = sample_owner(p, b.x - 1, b.y - 1);
        if (o >= 0) dep = std::max(dep, wave[static_cast<size_t>(o)]);
      }
    }
    if (b.intrabc) {
      const int msx = 3 + (p > 0 && geom.ss_x ? 1 : 0), msy = 3 + (p > 0 && geom.ss_y ? 1 : 0);
      const int sx = b.x + (b.mv[0][0] >> msx), sy = b.y + (b.mv[0][1] >> msy);
      const int x1 = std::min(sx + b.w, grid_w[p]), y1 = std::min(sy + b.h, grid_h[p]);
      const int oc = 1 << osh[p];
      for (int y = std::max(sy, 0) & ~(oc - 1); y < y1; y += oc)
        for (int x = std::max(sx, 0) & ~(oc - 1); x < x1; x += oc) {
          const int o = sample_owner(p, x, y);
          if (o >= 0) dep = std::max(dep, wave[static_cast<size_t>(o)]);
        }
    }
    if (b.mode == PredictionMode::ChromaFromLuma)
      for (int y = 0; y < b.h; ++y)
        for (int x = 0; x < b.w; ++x) {
          const int o = sample_owner(0, (b.x + x) << geom.ss_x, (b.y + y) << geom.ss_y);
          if (o >= 0) dep = std::max(dep, wave[static_cast<size_t>(o)]);
        }
    wave[bi] = std::max(wave[bi], dep + 1);
    max_wave = std::max(max_wave, wave[bi]);
  }
  if (serial) {
    for (size_t bi = 0; bi < nb; ++bi) wave[bi] = static_cast<int>(bi);
    max_wave = nb ? static_cast<int>(nb) - 1 : 0;
  }
  plan->waves.assign(static_cast<size_t>(max_wave) + 1, WaveJobs());

  int scale_x = AV1G_SCALE_ONE, scale_y = AV1G_SCALE_ONE;
  for (int i = 0; i < 7; ++i) {
    const int rw = syntax.reference_width[i] > 0 ? syntax.reference_width[i] : syntax.width;
    const int rh = syntax.reference_height[i] > 0 ? syntax.reference_height[i] : syntax.height;
    if (rw != syntax.width || rh != syntax.height) {
      *err = "reference scaling between pictures of different sizes is not supported";
      return false;
    }
  }
  if (g.upscaled()) {
    scale_x = ((syntax.width << AV1G_REF_SCALE_SHIFT) + (frame_w / 2)) / frame_w;
    scale_y = ((syntax.height << AV1G_REF_SCALE_SHIFT) + (frame_h / 2)) / frame_h;
  }

  size_t tmp_cursor = 0;
  for (size_t bi = 0; bi < nb; ++bi) {
    const BlockSyntax& b = syntax.blocks[bi];
    WaveJobs& wj = plan->waves[static_cast<size_t>(wave[bi])];
    const int plane = b.plane;
    const int stride = g.stride[plane];
    const int dst_off = at(plane, b.x, b.y);
    const int max_x = (((frame_w + 7) & ~7) >> (plane ? geom.ss_x : 0)) - 1;
    const int max_y = (((frame_h + 7) & ~7) >> (plane ? geom.ss_y : 0)) - 1;
    if (b.is_intra) {
      Av1gIntraJob job;
      std::memset(&job, 0, sizeof(job));
      job.dst_off = dst_off;
      job.dst_stride = stride;
      job.w = b.w;
      job.h = b.h;
      job.mode = static_cast<int>(b.mode);
      job.bitdepth = b.bit_depth;
      job.x = b.x;
      job.y = b.y;
      job.have_left = b.have_left && b.x > 0;
      job.have_above = b.have_above && b.y > 0;
      job.have_above_right = b.have_above_right >= 0 ? b.have_above_right != 0 : ar[bi] != 0;
      job.have_below_left = b.have_below_left >= 0 ? b.have_below_left != 0 : bl[bi] != 0;
      job.filter_type = syntax.enable_intra_edge_filter ? b.filter_type : -1;
      job.filter_intra = b.use_filter_intra && plane == 0;
      job.filter_intra_mode = static_cast<int>(b.filter_intra_mode);
      job.angle = b.angle_delta;
      job.cfl_alpha = b.cfl_alpha;
      job.luma_off = static_cast<int>(g.plane_off[0]);
      job.luma_stride = g.stride[0];
      job.luma_w = (b.x + b.w) << geom.ss_x;
      job.luma_h = (b.y + b.h) << geom.ss_y;
      job.ss_x = geom.ss_x;
      job.ss_y = geom.ss_y;
      job.max_x = max_x;
      job.max_y = max_y;
      if (b.use_palette) {
        job.palette = 1;
        job.palette_size = b.palette_size;
        for (int i = 0; i < 8; ++i) job.palette_color[i] = b.palette[i];
        job.palette_off = b.palette_off;
        job.palette_stride = b.palette_stride;
      }
      wj.intra.push_back(job);
    } else {
      Av1gInterJob job;
      std::memset(&job, 0, sizeof(job));
      job.dst_off = dst_off;
      job.dst_stride = stride;
      job.plane = plane;
      job.w = b.w;
      job.h = b.h;
      job.ref_x = b.x;
      job.ref_y = b.y;
      job.ref_w = g.out_w[plane];
      job.ref_h = g.out_h[plane];
      job.scale_x = scale_x;
      job.scale_y = scale_y;
      if (b.intrabc) {
        const int sx = plane > 0 && geom.ss_x ? 1 : 0, sy = plane > 0 && geom.ss_y ? 1 : 0;
        job.ref_w = ((frame_w + 3) & ~3) >> sx;
        job.ref_h = ((frame_h + 3) & ~3) >> sy;
        job.scale_x = AV1G_SCALE_ONE;
        job.scale_y = AV1G_SCALE_ONE;
      }
      job.border = 0;
      job.mv0_x = b.mv[0][0];
      job.mv0_y = b.mv[0][1];
      job.mv1_x = b.mv[1][0];
      job.mv1_y = b.mv[1][1];
      job.filt = static_cast<int>(b.filter);
      job.filt_v = static_cast<int>(b.filter_v);
      job.sub_x = geom.ss_x && plane > 0 ? 1 : 0;
      job.sub_y = geom.ss_y && plane > 0 ? 1 : 0;
      job.compound = static_cast<int>(b.compound);
      job.w0 = b.blend[0];
      job.w1 = b.blend[1];
      job.bitdepth = b.bit_depth;
      job.mask_kind = AV1G_MASK_NONE;
      if (b.compound == CompoundType::Wedge || b.compound == CompoundType::DiffWtd) {
        const int lw = b.w << (plane > 0 ? geom.ss_x : 0), lh = b.h << (plane > 0 ? geom.ss_y : 0);
        job.mask_kind = b.compound == CompoundType::Wedge ? AV1G_MASK_WEDGE : AV1G_MASK_DIFFWTD;
        job.mask_sign = b.mask_sign;
        if (job.mask_kind == AV1G_MASK_WEDGE) {
          int dir = 0, xoff = 0, yoff = 0, flip = 0;
          if (!tables::wedge_entry(lw, lh, b.wedge_index, &dir, &xoff, &yoff, &flip)) {
            *err = "a wedge on a block size that has none";
            return false;
          }
          job.wedge_dir = dir;
          job.wedge_xoff = xoff;
          job.wedge_yoff = yoff;
          job.wedge_flip = flip;
        }
        if (job.mask_kind == AV1G_MASK_DIFFWTD && plane > 0) {
          job.luma_ref0_off = static_cast<int>(g.plane_off[0]);
          job.luma_ref1_off = static_cast<int>(g.plane_off[0]);
          job.luma_stride = g.stride[0];
          job.luma_x = b.x << geom.ss_x;
          job.luma_y = b.y << geom.ss_y;
          job.luma_ref_w = g.out_w[0];
          job.luma_ref_h = g.out_h[0];
        }
      }
      job.interintra = b.interintra ? 1 : 0;
      job.interintra_mode = b.interintra_mode;
      job.have_left = b.have_left && b.x > 0;
      job.have_above = b.have_above && b.y > 0;
      job.max_x = max_x;
      job.max_y = max_y;
      if (b.interintra && b.interintra_wedge) {
        const int lw = b.w << (plane > 0 ? geom.ss_x : 0), lh = b.h << (plane > 0 ? geom.ss_y : 0);
        int dir = 0, xoff = 0, yoff = 0, flip = 0;
        if (!tables::wedge_entry(lw, lh, b.wedge_index, &dir, &xoff, &yoff, &flip)) {
          *err = "an inter/intra wedge on a block size that has none";
          return false;
        }
        job.interintra_wedge = 1;
        job.wedge_dir = dir;
        job.wedge_xoff = xoff;
        job.wedge_yoff = yoff;
        job.wedge_flip = flip;
      }
      job.ref0_slot = slot_of(syntax, b.ref[0]);
      job.ref1_slot = slot_of(syntax, b.ref[1]);
      job.ref0_off = static_cast<int>(g.plane_off[plane]);
      job.ref1_off = static_cast<int>(g.plane_off[plane]);
      job.ref0_stride = stride;
      job.ref1_stride = stride;
      job.obmc_count = 0;
      for (const ObmcCandidate& c : b.obmc) {
        if (job.obmc_count >= 8) break;
        Av1gObmcCandidate& o = job.obmc[job.obmc_count];
        std::memset(&o, 0, sizeof(o));
        o.pass = c.pass;
        o.rx = c.x;
        o.ry = c.y;
        o.w = c.w;
        o.h = c.h;
        o.mv_x = c.mv[0];
        o.mv_y = c.mv[1];
        o.filt = static_cast<int>(c.filter);
        o.filt_v = static_cast<int>(c.filter_v);
        o.ref_slot = slot_of(syntax, c.ref);
        o.ref_off = static_cast<int>(g.plane_off[plane]);
        ++job.obmc_count;
      }
      for (int i = 0; i < 2; ++i) {
        job.warp[i] = 0;
        job.mask_warp[i] = 0;
        if (!b.warp[i] && !b.mask_warp[i]) continue;
        if (b.warp[i]) job.warp[i] = 1;
        if (b.mask_warp[i]) job.mask_warp[i] = 1;
        int alpha = 0, beta = 0, gamma = 0, delta = 0;
        if (!tables::gm_setup_shear(b.warp_params[i], &alpha, &beta, &gamma, &delta)) continue;
        for (int j = 0; j < 6; ++j) job.warp_params[i][j] = b.warp_params[i][j];
        job.warp_abcd[i][0] = alpha;
        job.warp_abcd[i][1] = beta;
        job.warp_abcd[i][2] = gamma;
        job.warp_abcd[i][3] = delta;
      }
      wj.inter.push_back(job);
    }
    for (const TransformBlock& tx : b.transforms) {
      Av1gItxJob itx;
      std::memset(&itx, 0, sizeof(itx));
      itx.coef_off = tx.coef_offset;
      itx.tmp_off = static_cast<int>(tmp_cursor);
      tmp_cursor += static_cast<size_t>(tx.w) * static_cast<size_t>(tx.h);
      itx.dst_off = dst_off + tx.y * stride + tx.x;
      itx.dst_stride = stride;
      itx.w = tx.w;
      itx.h = tx.h;
      itx.tx_type = static_cast<int>(tx.tx_type);
      itx.eob = tx.eob;
      itx.lossless = b.lossless ? 1 : 0;
      itx.row_shift = b.lossless ? 0 : tables::tx_row_shift(tx.w, tx.h);
      itx.bitdepth = b.bit_depth;
      itx.plane = b.plane;
      if (itx.eob <= 0) continue;
      if (b.residual_provided)
        wj.residual.push_back(itx);
      else
        wj.itx.push_back(itx);
    }
  }
  plan->tmp_size = tmp_cursor;

  for (const DeblockEdge& e : syntax.deblock) {
    Av1gDeblockJob job;
    std::memset(&job, 0, sizeof(job));
    const int plane = e.plane;
    job.p_off = e.vertical ? at(plane, e.x - 2, e.y) : at(plane, e.x, e.y - 2);
    job.stride = e.vertical ? 1 : g.stride[plane];
    job.step = e.vertical ? g.stride[plane] : 1;
    job.len = e.length;
    job.level = e.level;
    job.sharpness = syntax.sharpness;
    job.filter_size = e.filter_size;
    job.is_chroma = plane != 0;
    job.bitdepth = geom.bit_depth;
    (e.vertical ? plan->deblock_v : plan->deblock_h).push_back(job);
  }
  for (const CdefBlock& c : syntax.cdef) {
    Av1gCdefJob job;
    std::memset(&job, 0, sizeof(job));
    const int plane = c.plane;
    const bool chroma = plane > 0;
    job.x = c.x;
    job.y = c.y;
    job.src_off = at(plane, c.x, c.y);
    job.src_stride = g.stride[plane];
    job.dst_off = job.src_off;
    job.dst_stride = g.stride[plane];
    job.luma_off = at(0, chroma ? c.x << geom.ss_x : c.x, chroma ? c.y << geom.ss_y : c.y);
    job.luma_stride = g.stride[0];
    job.pri_strength = c.pri_strength;
    job.sec_strength = c.sec_strength;
    job.damping = c.damping;
    job.dir = c.dir;
    job.uv = chroma ? 1 : 0;
    job.ss_x = geom.ss_x;
    job.ss_y = geom.ss_y;
    job.bitdepth = geom.bit_depth;
    job.plane_w = grid_w[plane];
    job.plane_h = grid_h[plane];
    plan->cdef.push_back(job);
  }
  if (g.upscaled()) {
    for (int plane = 0; plane < 3; ++plane) {
      Av1gUpscaleJob job;
      std::memset(&job, 0, sizeof(job));
      job.src_off = static_cast<int>(g.plane_off[plane]);
      job.dst_off = static_cast<int>(g.plane_off[plane]);
      job.src_stride = g.stride[plane];
      job.dst_stride = g.stride[plane];
      job.y0 = 0;
      job.rows = g.out_h[plane];
      job.upscaled_w = g.out_w[plane];
      job.max_x = g.plane_w[plane] - 1;
      av1g_superres_params(g.plane_w[plane], g.out_w[plane], &job.step_x, &job.initial_subpel_x);
      job.bitdepth = geom.bit_depth;
      plan->upscale.push_back(job);
    }
  }
  for (int plane = 0; plane < 3; ++plane) {
    if (syntax.restoration_type[plane] == kRestoreNone) continue;
    const int sub_y = plane > 0 ? geom.ss_y : 0;
    const int unit_size = 1 << syntax.restoration_unit_log2[plane];
    const int unit_cols = std::max(syntax.restoration_unit_cols[plane], 1);
    const int plane_w = g.out_w[plane], plane_h = g.out_h[plane];
    const int fh = (geom.height + (sub_y ? 1 << (sub_y - 1) : 0)) >> sub_y;
    const int unit_rows = std::max((fh + (unit_size >> 1)) / unit_size, 1);
    for (int y = 0; y < plane_h;) {
      const int stripe = ((y << sub_y) + 8) / 64;
      const int stripe_start = (-8 + stripe * 64) >> sub_y;
      const int stripe_end = stripe_start + (64 >> sub_y) - 1;
      const int y1 = std::min(plane_h, stripe_end + 1);
      const int unit_row = std::min(unit_rows - 1, (((y << sub_y) + 8) >> sub_y) / unit_size);
      for (int uc = 0; uc < unit_cols; ++uc) {
        const size_t index = static_cast<size_t>(unit_row) * unit_cols + uc;
        if (index >= syntax.restoration[plane].size()) continue;
        const RestorationUnit& unit = syntax.restoration[plane][index];
        if (unit.type == kRestoreNone) continue;
        const int ux0 = uc * unit_size;
        const int ux1 = uc + 1 == unit_cols ? plane_w : std::min(ux0 + unit_size, plane_w);
        if (ux1 <= ux0) continue;
        Av1gLrJob job;
        std::memset(&job, 0, sizeof(job));
        job.plane = plane;
        job.dst_off = at(plane, ux0, y);
        job.dst_stride = g.stride[plane];
        job.cdef_off = at(plane, 0, 0);
        job.cdef_stride = g.stride[plane];
        job.curr_off = at(plane, 0, 0);
        job.curr_stride = g.stride[plane];
        job.x = ux0;
        job.y = y;
        job.w = ux1 - ux0;
        job.h = y1 - y;
        job.plane_w = plane_w;
        job.plane_h = plane_h;
        job.stripe_start_y = stripe_start;
        job.stripe_end_y = stripe_end;
        job.type = unit.type == kRestoreWiener ? AV1G_LR_WIENER : AV1G_LR_SGRPROJ;
        if (unit.type == kRestoreWiener) {
          for (int i = 0; i < 3; ++i) {
            job.vfilter[i] = unit.wiener[0][i];
            job.hfilter[i] = unit.wiener[1][i];
          }
        } else {
          for (int i = 0; i < 2; ++i) {
            job.sgr_radius[i] = kSgrParams[unit.sgr_set & 15][i * 2];
            job.sgr_eps[i] = kSgrParams[unit.sgr_set & 15][i * 2 + 1];
            job.sgr_xqd[i] = unit.sgr_xqd[i];
          }
        }
        job.bitdepth = geom.bit_depth;
        plan->restoration.push_back(job);
      }
      y = y1;
    }
  }
  if (syntax.grain.apply) {
    const bool identity = syntax.color.matrix == ColorMatrix::Identity;
    build_grain(syntax.grain, geom.bit_depth, geom.ss_x, geom.ss_y, identity, g.out_w[0],
                g.out_h[0], &plan->grain_noise);
    for (int plane = 0; plane < 3; ++plane)
      append_grain_jobs(plan->grain_noise, plane, g.out_w[plane], g.out_h[plane],
                        static_cast<int>(g.plane_off[plane]), g.stride[plane],
                        static_cast<int>(g.plane_off[0]), g.stride[0], g.out_w[0], geom.ss_x,
                        geom.ss_y, &plan->grain_chroma, &plan->grain_luma);
  }
  plan->blocks = static_cast<int>(nb);
  return true;
}

struct RefDecoder {
  std::unique_ptr<Backend> backend;
  int waves = 0;
  Frame last;
};

void set_err(char* err, int n, const std::string& s) {
  if (err && n > 0) {
    std::strncpy(err, s.c_str(), static_cast<size_t>(n) - 1);
    err[n - 1] = 0;
  }
}

}  // namespace

extern "C" {

void* av1r_create(void) {
  BackendOptions opts;
  opts.kind = BackendKind::Cpu;
  std::string error;
  RefDecoder* d = new RefDecoder;
  d->backend = create_cpu_backend(opts, &error);
  return d;
}

void av1r_destroy(void* h) { delete static_cast<RefDecoder*>(h); }

int av1r_decode_wire(void* h, const uint8_t* wire, size_t len, int keep, int reserved, int serial,
                     const av1h_planes* out, int* slot_out, char* err, int errlen) {
  RefDecoder* d = static_cast<RefDecoder*>(h);
  av1h_frame_desc desc;
  if (av1h_wire_parse(wire, len, &desc) != AV1H_OK) {
    set_err(err, errlen, "not a wire frame");
    return -1;
  }
  FrameSyntax syntax;
  syntax_from_desc(desc, &syntax);
  internal::FrameGeometry geom;
  geom.width = syntax.width;
  geom.height = syntax.height;
  geom.bit_depth = syntax.bit_depth;
  geom.ss_x = syntax.ss_x;
  geom.ss_y = syntax.ss_y;
  const PlanLayout layout =
      d->backend->prepare(geom, syntax.coded_width > 0 ? syntax.coded_width : syntax.width);
  FramePlan plan;
  std::string error;
  if (!build_plan(syntax, layout, d->backend->reference_capacity(), serial != 0, &plan, &error)) {
    set_err(err, errlen, error);
    return -1;
  }
  d->waves = static_cast<int>(plan.waves.size());
  int slot = -1;
  DecodeStats stats;
  bool ok;
  if (keep && reserved >= 0)
    ok = d->backend->decode_reserved(plan, reserved, &d->last, &slot, &stats, &error);
  else
    ok = d->backend->decode(plan, keep != 0, &d->last, &slot, &stats, &error);
  if (!ok) {
    set_err(err, errlen, error);
    return -1;
  }
  if (out)
    for (int p = 0; p < 3; ++p)
      for (int y = 0; y < d->last.plane_height(p); ++y)
        std::memcpy(out->plane[p] + static_cast<size_t>(y) * out->stride[p],
                    d->last.plane(p) + static_cast<size_t>(y) * d->last.stride[p],
                    static_cast<size_t>(d->last.plane_width(p)) * 2);
  if (slot_out) *slot_out = slot;
  return 0;
}

int av1r_reserve(void* h) { return static_cast<RefDecoder*>(h)->backend->reserve_reference(); }
void av1r_release(void* h, int slot) { static_cast<RefDecoder*>(h)->backend->release_reference(slot); }
int av1r_waves(void* h) { return static_cast<RefDecoder*>(h)->waves; }

/* the reference's NV12 and RGB packing of the last decoded picture
 * (factory.cpp:247-278, tables.cpp:535-561) */
int64_t av1r_last_nv12(void* h, size_t pitch, uint8_t* out, int64_t cap) {
  RefDecoder* d = static_cast<RefDecoder*>(h);
  std::vector<uint8_t> v;
  if (!frame_to_nv12(d->last, pitch, &v)) return -1;
  if (out && cap >= static_cast<int64_t>(v.size())) std::memcpy(out, v.data(), v.size());
  return static_cast<int64_t>(v.size());
}

int64_t av1r_last_rgb(void* h, int matrix, int range, uint8_t* out, int64_t cap) {
  RefDecoder* d = static_cast<RefDecoder*>(h);
  std::vector<uint8_t> v;
  tables::frame_to_rgb(d->last, static_cast<ColorMatrix>(matrix), static_cast<ColorRange>(range), &v);
  if (out && cap >= static_cast<int64_t>(v.size())) std::memcpy(out, v.data(), v.size());
  return static_cast<int64_t>(v.size());
}

/* the backend's own to_nv12 of a slot (its gather rule) */
int64_t av1r_slot_nv12(void* h, int slot, size_t pitch, uint8_t* out, int64_t cap) {
  RefDecoder* d = static_cast<RefDecoder*>(h);
  std::vector<uint8_t> v;
  std::string error;
  if (!d->backend->to_nv12(slot, pitch, &v, &error, nullptr)) return -1;
  if (out && cap >= static_cast<int64_t>(v.size())) std::memcpy(out, v.data(), v.size());
  return static_cast<int64_t>(v.size());
}

/* the backend's own to_rgb of a slot (its gather rule) */
int64_t av1r_slot_rgb(void* h, int slot, int matrix, int range, uint8_t* out, int64_t cap) {
  RefDecoder* d = static_cast<RefDecoder*>(h);
  std::vector<uint8_t> v;
  std::string error;
  if (!d->backend->to_rgb(slot, static_cast<ColorMatrix>(matrix), static_cast<ColorRange>(range), &v, &error))
    return -1;
  if (out && cap >= static_cast<int64_t>(v.size())) std::memcpy(out, v.data(), v.size());
  return static_cast<int64_t>(v.size());
}

/* The reference's own dequantisation (av1g_dequant_one, kernels/
 * av1gpu_kernels.h:1024-1083, with the scan tables of src/frontend/
 * coef_tables.hpp) of every coded transform of a wire frame that carries
 * levels, compared with the frame's coefficient raster: returns the count of
 * raster entries that differ (0: the raster is the levels' dequantisation),
 * -1 when the frame has no levels, -2 for a malformed frame; *coded gets the
 * number of transforms dequantised. */
int64_t av1r_dequant_check(const uint8_t* wire, size_t len, int64_t* coded) {
  av1h_frame_desc d;
  if (av1h_wire_parse(wire, len, &d) != AV1H_OK) return -2;
  if (d.num_levels <= 0) return -1;
  std::vector<int> pool(static_cast<size_t>(d.num_coefficients), 0);
  int64_t n = 0;
  for (int64_t bi = 0; bi < d.num_blocks; ++bi) {
    const av1h_block& b = d.blocks[bi];
    if (b.residual_provided) continue;
    for (int k = 0; k < b.tx_count; ++k) {
      const av1h_transform& t = d.transforms[b.tx_first + k];
      if (t.eob <= 0) continue;
      Av1gItxJob job;
      std::memset(&job, 0, sizeof(job));
      job.coef_off = t.coef_offset;
      job.w = t.w;
      job.h = t.h;
      job.tx_type = t.tx_type;
      job.eob = t.eob;
      job.bitdepth = b.bit_depth;
      job.level_offset = t.level_offset;
      job.plane = b.plane;
      job.dc_step = t.dc_step;
      job.ac_step = t.ac_step;
      job.denominator = t.denominator;
      job.qm_off = t.qm_off;
      av1g_dequant_one(&job, d.levels, d.num_qm > 0 ? d.qm_pool : nullptr, av1gpu::frontend::coef_tables::kDefaultScan,
                       av1gpu::frontend::coef_tables::kMrowScan, av1gpu::frontend::coef_tables::kMcolScan,
                       pool.data());
      ++n;
    }
  }
  if (coded) *coded = n;
  int64_t diff = 0;
  for (size_t i = 0; i < pool.size(); ++i) diff += pool[i] != d.coefficients[i];
  return diff;
}

}  // extern "C"
