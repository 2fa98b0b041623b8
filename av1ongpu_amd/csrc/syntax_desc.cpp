#include "syntax_desc.hpp"

#include <cstring>

namespace av1gpu {
namespace detail {

static void header_from(const FrameSyntax& s, av1h_frame_header* h) {
  std::memset(h, 0, sizeof(*h));
  h->width = s.width;
  h->coded_width = s.coded_width;
  h->height = s.height;
  h->bit_depth = s.bit_depth;
  h->ss_x = s.ss_x;
  h->ss_y = s.ss_y;
  h->base_q_idx = s.base_q_idx;
  h->sharpness = s.sharpness;
  h->enable_intra_edge_filter = s.enable_intra_edge_filter ? 1 : 0;
  h->num_reference_slots = static_cast<int32_t>(s.reference_slots.size());
  for (size_t i = 0; i < s.reference_slots.size() && i < AV1H_MAX_REFS; ++i)
    h->reference_slots[i] = s.reference_slots[i];
  for (int i = 0; i < 7; ++i) {
    h->reference_width[i] = s.reference_width[i];
    h->reference_height[i] = s.reference_height[i];
  }
  h->color_matrix = static_cast<int32_t>(s.color.matrix);
  h->color_range = static_cast<int32_t>(s.color.range);
  for (int p = 0; p < 3; ++p) {
    h->restoration_type[p] = s.restoration_type[p];
    h->restoration_unit_log2[p] = s.restoration_unit_log2[p];
    h->restoration_unit_cols[p] = s.restoration_unit_cols[p];
  }
  const FilmGrainParams& g = s.grain;
  av1h_grain& o = h->grain;
  o.apply = g.apply;
  o.seed = g.seed;
  o.num_y_points = g.num_y_points;
  o.chroma_scaling_from_luma = g.chroma_scaling_from_luma;
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
  o.overlap_flag = g.overlap_flag;
  o.clip_to_restricted_range = g.clip_to_restricted_range;
}

bool flatten(const FrameSyntax& s, DescStorage* out, std::string* error) {
  if (s.reference_slots.size() > AV1H_MAX_REFS) {
    if (error) *error = "a frame names more than 8 reference slots";
    return false;
  }
  out->blocks.clear();
  out->transforms.clear();
  out->obmc.clear();
  out->blocks.reserve(s.blocks.size());
  size_t ntx = 0;
  for (const BlockSyntax& b : s.blocks) ntx += b.transforms.size();
  out->transforms.reserve(ntx);
  for (const BlockSyntax& b : s.blocks) {
    av1h_block o;
    std::memset(&o, 0, sizeof(o));
    o.x = b.x;
    o.y = b.y;
    o.w = b.w;
    o.h = b.h;
    o.plane = b.plane;
    o.bit_depth = b.bit_depth;
    o.is_intra = b.is_intra;
    o.intrabc = b.intrabc;
    o.mode = static_cast<int32_t>(b.mode);
    o.have_left = b.have_left;
    o.have_above = b.have_above;
    o.have_above_right = b.have_above_right;
    o.have_below_left = b.have_below_left;
    o.angle_delta = b.angle_delta;
    o.filter_type = b.filter_type;
    o.use_filter_intra = b.use_filter_intra;
    o.filter_intra_mode = static_cast<int32_t>(b.filter_intra_mode);
    o.cfl_alpha = b.cfl_alpha;
    o.lossless = b.lossless;
    o.use_palette = b.use_palette;
    o.palette_size = b.palette_size;
    for (int i = 0; i < 8; ++i) o.palette[i] = b.palette[i];
    o.palette_off = b.palette_off;
    o.palette_stride = b.palette_stride;
    for (int i = 0; i < 2; ++i) {
      o.ref[i] = b.ref[i];
      o.mv[i][0] = b.mv[i][0];
      o.mv[i][1] = b.mv[i][1];
      o.blend[i] = b.blend[i];
      o.warp[i] = b.warp[i];
      o.mask_warp[i] = b.mask_warp[i];
      for (int k = 0; k < 6; ++k) o.warp_params[i][k] = b.warp_params[i][k];
    }
    o.filter = static_cast<int32_t>(b.filter);
    o.filter_v = static_cast<int32_t>(b.filter_v);
    o.motion_mode = b.motion_mode;
    o.compound = static_cast<int32_t>(b.compound);
    o.wedge_index = b.wedge_index;
    o.mask_sign = b.mask_sign;
    o.interintra = b.interintra;
    o.interintra_mode = b.interintra_mode;
    o.interintra_wedge = b.interintra_wedge;
    o.residual_provided = b.residual_provided;
    o.tx_first = static_cast<int32_t>(out->transforms.size());
    o.tx_count = static_cast<int32_t>(b.transforms.size());
    o.obmc_first = static_cast<int32_t>(out->obmc.size());
    o.obmc_count = static_cast<int32_t>(b.obmc.size());
    for (const TransformBlock& t : b.transforms) {
      av1h_transform r;
      r.x = t.x;
      r.y = t.y;
      r.w = t.w;
      r.h = t.h;
      r.coef_offset = t.coef_offset;
      r.eob = t.eob;
      r.tx_type = static_cast<int32_t>(t.tx_type);
      r.level_offset = t.level_offset;
      r.qindex = t.qindex;
      r.dc_step = t.dc_step;
      r.ac_step = t.ac_step;
      r.denominator = t.denominator;
      r.qm_off = t.qm_off;
      out->transforms.push_back(r);
    }
    for (const ObmcCandidate& c : b.obmc) {
      av1h_obmc r;
      r.pass = c.pass;
      r.x = c.x;
      r.y = c.y;
      r.w = c.w;
      r.h = c.h;
      r.mv_x = c.mv[0];
      r.mv_y = c.mv[1];
      r.filter = static_cast<int32_t>(c.filter);
      r.filter_v = static_cast<int32_t>(c.filter_v);
      r.ref = c.ref;
      out->obmc.push_back(r);
    }
    out->blocks.push_back(o);
  }
  out->deblock.resize(s.deblock.size());
  for (size_t i = 0; i < s.deblock.size(); ++i) {
    const DeblockEdge& e = s.deblock[i];
    out->deblock[i] = {e.plane, e.x, e.y, e.vertical ? 1 : 0, e.length, e.level, e.filter_size};
  }
  out->cdef.resize(s.cdef.size());
  for (size_t i = 0; i < s.cdef.size(); ++i) {
    const CdefBlock& c = s.cdef[i];
    out->cdef[i] = {c.plane, c.x, c.y, c.pri_strength, c.sec_strength, c.damping, c.dir};
  }
  for (int p = 0; p < 3; ++p) {
    out->lr[p].resize(s.restoration[p].size());
    for (size_t i = 0; i < s.restoration[p].size(); ++i) {
      const RestorationUnit& u = s.restoration[p][i];
      av1h_lr_unit& r = out->lr[p][i];
      r.type = u.type;
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k) r.wiener[a][k] = u.wiener[a][k];
      r.sgr_set = u.sgr_set;
      r.sgr_xqd[0] = u.sgr_xqd[0];
      r.sgr_xqd[1] = u.sgr_xqd[1];
    }
  }
  av1h_frame_desc& d = out->desc;
  std::memset(&d, 0, sizeof(d));
  header_from(s, &d.hdr);
  d.blocks = out->blocks.data();
  d.num_blocks = static_cast<int64_t>(out->blocks.size());
  d.transforms = out->transforms.data();
  d.num_transforms = static_cast<int64_t>(out->transforms.size());
  d.obmc = out->obmc.data();
  d.num_obmc = static_cast<int64_t>(out->obmc.size());
  d.coefficients = s.coefficients.data();
  d.num_coefficients = static_cast<int64_t>(s.coefficients.size());
  d.residual = s.residual.data();
  d.num_residual = static_cast<int64_t>(s.residual.size());
  d.levels = s.levels.data();
  d.num_levels = static_cast<int64_t>(s.levels.size());
  d.qm_pool = s.qm_pool.data();
  d.num_qm = static_cast<int64_t>(s.qm_pool.size());
  d.palette_map = s.palette_map.data();
  d.num_palette = static_cast<int64_t>(s.palette_map.size());
  d.deblock = out->deblock.data();
  d.num_deblock = static_cast<int64_t>(out->deblock.size());
  d.cdef = out->cdef.data();
  d.num_cdef = static_cast<int64_t>(out->cdef.size());
  for (int p = 0; p < 3; ++p) {
    d.restoration[p] = out->lr[p].data();
    d.num_restoration[p] = static_cast<int64_t>(out->lr[p].size());
  }
  return true;
}

void unflatten(const av1h_frame_desc& d, FrameSyntax* s) {
  const av1h_frame_header& h = d.hdr;
  *s = FrameSyntax();
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
    BlockSyntax& o2 = s->blocks[static_cast<size_t>(i)];
    o2.x = b.x;
    o2.y = b.y;
    o2.w = b.w;
    o2.h = b.h;
    o2.plane = b.plane;
    o2.bit_depth = b.bit_depth;
    o2.is_intra = b.is_intra != 0;
    o2.intrabc = b.intrabc != 0;
    o2.mode = static_cast<PredictionMode>(b.mode);
    o2.have_left = b.have_left != 0;
    o2.have_above = b.have_above != 0;
    o2.have_above_right = b.have_above_right;
    o2.have_below_left = b.have_below_left;
    o2.angle_delta = b.angle_delta;
    o2.filter_type = b.filter_type;
    o2.use_filter_intra = b.use_filter_intra != 0;
    o2.filter_intra_mode = static_cast<FilterIntraMode>(b.filter_intra_mode);
    o2.cfl_alpha = b.cfl_alpha;
    o2.lossless = b.lossless != 0;
    o2.use_palette = b.use_palette != 0;
    o2.palette_size = b.palette_size;
    for (int k = 0; k < 8; ++k) o2.palette[k] = b.palette[k];
    o2.palette_off = b.palette_off;
    o2.palette_stride = b.palette_stride;
    for (int r = 0; r < 2; ++r) {
      o2.ref[r] = b.ref[r];
      o2.mv[r][0] = b.mv[r][0];
      o2.mv[r][1] = b.mv[r][1];
      o2.blend[r] = b.blend[r];
      o2.warp[r] = b.warp[r] != 0;
      o2.mask_warp[r] = b.mask_warp[r] != 0;
      for (int k = 0; k < 6; ++k) o2.warp_params[r][k] = b.warp_params[r][k];
    }
    o2.filter = static_cast<FilterType>(b.filter);
    o2.filter_v = static_cast<FilterType>(b.filter_v);
    o2.motion_mode = b.motion_mode;
    o2.compound = static_cast<CompoundType>(b.compound);
    o2.wedge_index = b.wedge_index;
    o2.mask_sign = b.mask_sign;
    o2.interintra = b.interintra != 0;
    o2.interintra_mode = b.interintra_mode;
    o2.interintra_wedge = b.interintra_wedge != 0;
    o2.residual_provided = b.residual_provided != 0;
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
      o2.transforms.push_back(tb);
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
      o2.obmc.push_back(c);
    }
  }
  s->coefficients.assign(d.coefficients, d.coefficients + d.num_coefficients);
  s->residual.assign(d.residual, d.residual + d.num_residual);
  s->levels.assign(d.levels, d.levels + d.num_levels);
  s->qm_pool.assign(d.qm_pool, d.qm_pool + d.num_qm);
  s->palette_map.assign(d.palette_map, d.palette_map + d.num_palette);
  s->deblock.resize(static_cast<size_t>(d.num_deblock));
  for (int64_t i = 0; i < d.num_deblock; ++i) {
    const av1h_deblock_edge& e = d.deblock[i];
    DeblockEdge& o3 = s->deblock[static_cast<size_t>(i)];
    o3.plane = e.plane;
    o3.x = e.x;
    o3.y = e.y;
    o3.vertical = e.vertical != 0;
    o3.length = e.length;
    o3.level = e.level;
    o3.filter_size = e.filter_size;
  }
  s->cdef.resize(static_cast<size_t>(d.num_cdef));
  for (int64_t i = 0; i < d.num_cdef; ++i) {
    const av1h_cdef_block& c = d.cdef[i];
    CdefBlock& o4 = s->cdef[static_cast<size_t>(i)];
    o4.plane = c.plane;
    o4.x = c.x;
    o4.y = c.y;
    o4.pri_strength = c.pri_strength;
    o4.sec_strength = c.sec_strength;
    o4.damping = c.damping;
    o4.dir = c.dir;
  }
  for (int p = 0; p < 3; ++p) {
    s->restoration[p].resize(static_cast<size_t>(d.num_restoration[p]));
    for (int64_t i = 0; i < d.num_restoration[p]; ++i) {
      const av1h_lr_unit& r = d.restoration[p][i];
      RestorationUnit& u = s->restoration[p][static_cast<size_t>(i)];
      u.type = r.type;
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k) u.wiener[a][k] = r.wiener[a][k];
      u.sgr_set = r.sgr_set;
      u.sgr_xqd[0] = r.sgr_xqd[0];
      u.sgr_xqd[1] = r.sgr_xqd[1];
    }
  }
}

std::vector<uint8_t> to_wire(const FrameSyntax& s) {
  DescStorage st;
  std::string err;
  std::vector<uint8_t> out;
  if (!flatten(s, &st, &err)) return out;
  out.resize(av1h_wire_size(&st.desc));
  av1h_wire_write(&st.desc, out.data());
  return out;
}

}  // namespace detail
}  // namespace av1gpu
