/*
 * AV1 coefficient scan orders and dequantisation of coded levels, for the
 * host (the workload generator, plan validation) and the device (the
 * transform kernel dequantises while it stages a transform's coefficients).
 *
 * A transform's coded levels come in scan order; position p of its raster
 * (min(32, w) x min(32, h), row major) holds scan index iscan(p). The scan
 * depends on the transform class (reference kernels/av1gpu_kernels.h:
 * 1008-1021): the vertical 1D types (V_DCT, V_ADST, V_FLIPADST) scan the
 * raster row by row, the horizontal ones (H_*) column by column, every other
 * type diagonally: anti-diagonal after anti-diagonal, alternating direction
 * on square rasters (starting rightwards along the top row), bottom-left to
 * top-right on rasters wider than tall and top-right to bottom-left on ones
 * taller than wide. These are the normative tables of the AV1 specification
 * (Default_Scan_*, Mrow_Scan_*, Mcol_Scan_*); here they are generated.
 */
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define AV1S_FN __host__ __device__ inline
#else
#define AV1S_FN inline
#endif

namespace av1scan {

/* Transform classes (Tx_Type -> class): 0 2D, 1 horizontal, 2 vertical.
 * Tx_Type numbering: V_DCT 10, H_DCT 11, V_ADST 12, H_ADST 13,
 * V_FLIPADST 14, H_FLIPADST 15. */
AV1S_FN int tx_class(int tx_type) {
  if (tx_type == 10 || tx_type == 12 || tx_type == 14) return 2;
  if (tx_type == 11 || tx_type == 13 || tx_type == 15) return 1;
  return 0;
}

/* the diagonal tables of every raster size (4..32 per side), concatenated:
 * size (lw, lh) (log2 of side / 4) starts at offset(lw, lh) */
constexpr int kDiagEntries = 3600;
constexpr int diag_offset(int lw, int lh) {
  int o = 0;
  for (int k = 0; k < lw * 4 + lh; ++k) o += 16 << ((k >> 2) + (k & 3));
  return o;
}

struct DiagTables {
  uint16_t scan[kDiagEntries];  /* scan index -> raster position */
  uint16_t iscan[kDiagEntries]; /* raster position -> scan index */
};

constexpr DiagTables make_diag() {
  DiagTables t{};
  for (int lw = 0; lw < 4; ++lw)
    for (int lh = 0; lh < 4; ++lh) {
      const int w = 4 << lw, h = 4 << lh, o = diag_offset(lw, lh);
      int n = 0;
      for (int d = 0; d < w + h - 1; ++d) {
        /* rows r of anti-diagonal d, r ascending: top-right to bottom-left */
        const int r0 = d - (w - 1) > 0 ? d - (w - 1) : 0, r1 = d < h - 1 ? d : h - 1;
        const bool up = w == h ? (d & 1) == 0 : w > h;
        for (int k = 0; k <= r1 - r0; ++k) {
          const int r = up ? r1 - k : r0 + k;
          const int p = r * w + (d - r);
          t.scan[o + n] = (uint16_t)p;
          t.iscan[o + p] = (uint16_t)n;
          ++n;
        }
      }
    }
  return t;
}

/* scan index of raster position p of a w x h raster (sides 4..32) of class
 * `cls`; `diag_iscan` is DiagTables::iscan */
AV1S_FN int scan_index(const uint16_t* diag_iscan, int cls, int lw, int lh, int p) {
  const int w = 4 << lw, h = 4 << lh;
  if (cls == 2) return p;                            /* row by row */
  if (cls == 1) return (p & (w - 1)) * h + (p >> (lw + 2)); /* column by column */
  return diag_iscan[diag_offset(lw, lh) + p];
}

/* raster position of scan index i */
AV1S_FN int scan_pos(const uint16_t* diag_scan, int cls, int lw, int lh, int i) {
  const int w = 4 << lw, h = 4 << lh;
  if (cls == 2) return i;
  if (cls == 1) return (i % h) * w + i / h;
  return diag_scan[diag_offset(lw, lh) + i];
}

/* One coefficient from its coded level (reference dequantisation,
 * kernels/av1gpu_kernels.h:1065-1081): the step (the DC step at position 0),
 * weighted by the quantiser matrix entry when there is one (qm >= 0), times
 * the level, its magnitude cut to 24 bits and divided by the transform's
 * denominator (1 << denom_shift), clipped to the range a `bit_depth`
 * coefficient holds. The magnitude is formed in 32 bits: only the low 24
 * bits of |level| x step are kept, and those are the same in 32-bit
 * wrapping arithmetic as in the 64-bit product (steps are positive). */
AV1S_FN int32_t dequant(int level, int step, int qm, int denom_shift, int bit_depth) {
  if (level == 0) return 0;
  if (qm >= 0) step = (step * qm + 16) >> 5;
  const uint32_t mag = ((uint32_t)(level < 0 ? -level : level) * (uint32_t)step) & 0xFFFFFFu;
  int32_t v = (int32_t)(mag >> denom_shift);
  if (level < 0) v = -v;
  const int32_t lim = 1 << (7 + bit_depth);
  return v > lim - 1 ? lim - 1 : (v < -lim ? -lim : v);
}

}  // namespace av1scan
