/*
 * Seeded synthetic FrameSyntax generator for the measured workloads.
 *
 * The reference decodes syntax its own closed-loop encoder writes
 * (src/core/synthetic.cpp:169-510), which is far narrower than the workloads
 * this backend is measured on (no partitions, one transform size, DC intra
 * only). This generator draws syntax with the statistics each workload names:
 * recursive partitions, every intra mode class, subpel and compound motion,
 * every transform size and type family, deblocking, CDEF, loop restoration,
 * superres and film grain - always spec consistent, so the decode is
 * independent of the order blocks of one dependency level run in:
 *   - blocks tile [0, ceil8(W)) x [0, ceil8(H)) of every plane exactly;
 *   - an intra coding block is emitted as one BlockSyntax per transform block,
 *     in coding order, with have_above_right / have_below_left set from what
 *     has been emitted before it (the specification's BlockDecoded);
 *   - deblocking filter sizes come from the transform sizes either side of an
 *     edge, so edges of one direction never touch each other's samples.
 */
#ifndef AV1GPU_SYNTH_HPP
#define AV1GPU_SYNTH_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "av1gpu/av1gpu.hpp"

namespace av1gpu {
namespace synth {

struct Config {
  int width = 1280;
  int height = 720;
  int bit_depth = 8;
  int coded_width = 0;       /* superres: the width the tiles code (0 = none)     */
  int frames = 30;
  int num_refs = 1;          /* 1: the previous frame; 3: hierarchical mini-GOPs  */
  int gop = 8;               /* mini-GOP length when num_refs > 1                 */
  int key_interval = 0;      /* 0: only the first frame is a key frame            */
  int max_tx = 32;           /* largest transform side (32 or 64)                 */
  bool rect_tx = false;      /* 2:1 transforms on 2:1 blocks                      */
  bool tx_ext = false;       /* identity, flip-ADST and the 1D transform types    */
  bool mv_gaussian = false;  /* vectors ~ N(0, mv_sigma) instead of U[-r, r]      */
  int mv_range = 128;        /* 1/8 luma samples                                  */
  double mv_sigma = 192.0;
  bool all_filters = false;  /* regular / smooth / sharp per direction            */
  double compound_frac = 0.0;
  double intra_frac = 0.05;  /* intra coding blocks inside inter frames           */
  double coef_zero = 0.85;
  double coef_scale = 40.0;
  double tx_skip = 0.2;      /* transform blocks coded with no residual at all    */
  int deblock_lo = 8;
  int deblock_hi = 32;
  bool cdef = true;
  int lr = 1;                /* 0 none, 1 Wiener, 2 switchable Wiener / self guided */
  bool grain = false;
  double inter_tools = 0.0;  /* inter coding blocks 8x8..32x32 using wedge / difference
                                weighted compound, inter/intra, OBMC or global warp */
  double intrabc = 0.0;      /* key frame coding blocks (below the first superblock
                                row) copied from the superblock row above */
  bool contract_tools = false; /* CFL, filter intra, palette, lossless, OBMC, masks,
                                  inter-intra, global warp, edge filter off        */
  double avail_auto = 0.0;   /* intra blocks whose have_above_right / have_below_left
                                are -1 (the decoder derives them from coding order) */
  double residual_frac = 0.0; /* coding blocks carrying raw residual samples
                                 (residual_provided) instead of coefficients       */
  int lr_log2 = 0;           /* luma restoration unit log2 (0: 64); chroma one less,
                                at least 32                                       */
  bool levels = false;       /* coded levels in scan order with their dequantiser
                                (steps, denominator, quantiser matrices); the raster
                                is their dequantisation                            */
  uint64_t seed = 1;
};

/* The named workloads: 1 = 720p 8-bit single reference, 2 = 1080p 10-bit full
 * tool, "mini" variants are the same statistics at fixture size. */
Config workload(const std::string& name, uint64_t seed);

struct Clip {
  Config config;
  std::vector<FrameSyntax> frames;        /* decode order                          */
  std::vector<std::vector<int>> refs;     /* frame indices each frame predicts from */
  std::vector<int> last_use;              /* last frame index that reads frame i   */
  std::vector<int> key;                   /* 1 for key frames                      */
};

/* Generates a whole clip. reference_slots of frame f holds the frame indices
 * of refs[f]; a driver replaces them with the slots the decoder returned. */
Clip generate(const Config& config);

/* One frame only (frame index `f` of a clip with this config). */
FrameSyntax generate_frame(const Config& config, int f, const std::vector<int>& refs, bool key);

}  // namespace synth
}  // namespace av1gpu

#endif
