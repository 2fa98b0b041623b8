/*
 * FrameSyntax <-> av1h_frame_desc: the host flattening the C++ Decoder API does
 * before it calls the C ABI, and its inverse for replaying wire frames.
 */
#ifndef AV1GPU_SYNTAX_DESC_HPP
#define AV1GPU_SYNTAX_DESC_HPP

#include <string>
#include <vector>

#include "av1gpu/av1gpu.hpp"
#include "av1hip.h"

namespace av1gpu {
namespace detail {

/* Owns the flat records; the coefficient, residual, level, QM and palette
 * pools are borrowed from the FrameSyntax, which must outlive the desc. */
struct DescStorage {
  std::vector<av1h_block> blocks;
  std::vector<av1h_transform> transforms;
  std::vector<av1h_obmc> obmc;
  std::vector<av1h_deblock_edge> deblock;
  std::vector<av1h_cdef_block> cdef;
  std::vector<av1h_lr_unit> lr[3];
  av1h_frame_desc desc;
};

bool flatten(const FrameSyntax& s, DescStorage* out, std::string* error);
void unflatten(const av1h_frame_desc& d, FrameSyntax* out);
std::vector<uint8_t> to_wire(const FrameSyntax& s);

}  // namespace detail
}  // namespace av1gpu

#endif
