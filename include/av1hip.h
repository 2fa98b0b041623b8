/*
 * av1hip.h - the C ABI of the MI355X (gfx950) AV1 reconstruction backend.
 *
 * This is the drop-in boundary. A caller hands over the decoded *syntax* of one
 * frame - prediction blocks, transform blocks, dequantised coefficients,
 * deblocking edges, CDEF blocks, loop restoration units, film grain
 * parameters - as flat arrays of plain records, and the library reconstructs
 * the picture on the GPU and keeps it resident in HBM as a reference slot.
 *
 * The records mirror the reference's public syntax types one field for one
 * field (include/av1gpu/av1gpu.hpp:187-537 in the reference), with the nested
 * std::vectors of BlockSyntax (transforms, OBMC strips) turned into index
 * ranges into frame-wide arrays, so any FFI (ctypes, cgo, JNI, N-API) can fill
 * them without C++.
 *
 * The entry points mirror av1gpu::Decoder (reference
 * include/av1gpu/av1gpu.hpp:615-787): decode_reference / decode /
 * decode_reference_reserved -> av1h_decode, release_reference ->
 * av1h_release_reference, reserve_reference -> av1h_reserve_reference, to_nv12
 * -> av1h_to_nv12, to_rgb -> av1h_to_rgb, reset -> av1h_reset, stats ->
 * av1h_get_stats. Every call returns an int status (AV1H_OK or a negative
 * code) and leaves a message in av1h_last_error(); nothing throws.
 *
 * All integers are little-endian two's complement; every struct is made of
 * int32_t fields only, so its layout is the same for every compiler.
 */
#ifndef AV1HIP_H
#define AV1HIP_H

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

#define AV1H_OK 0
#define AV1H_ERR_INVALID -1      /* bad argument or malformed syntax          */
#define AV1H_ERR_UNSUPPORTED -2  /* a tool this library does not reconstruct  */
#define AV1H_ERR_NO_DEVICE -3    /* no usable gfx950 device                    */
#define AV1H_ERR_DEVICE -4       /* a HIP call failed                          */
#define AV1H_ERR_SLOTS -5        /* the reference pool is exhausted            */
#define AV1H_ERR_MEMORY -6       /* an allocation (host or HBM) failed         */

/* ------------------------------------------------------------------ syntax */
/* Numbering follows the reference enums (av1gpu.hpp:117-183), which are the
 * specification's own numbers. */

enum {
  AV1H_PRED_DC = 0, AV1H_PRED_V = 1, AV1H_PRED_H = 2, AV1H_PRED_D45 = 3,
  AV1H_PRED_D135 = 4, AV1H_PRED_D113 = 5, AV1H_PRED_D157 = 6, AV1H_PRED_D203 = 7,
  AV1H_PRED_D67 = 8, AV1H_PRED_SMOOTH = 9, AV1H_PRED_SMOOTH_V = 10,
  AV1H_PRED_SMOOTH_H = 11, AV1H_PRED_PAETH = 12, AV1H_PRED_CFL = 13
};
enum {
  AV1H_FILT_REGULAR = 0, AV1H_FILT_SMOOTH = 1, AV1H_FILT_SHARP = 2, AV1H_FILT_BILINEAR = 3
};
enum {
  AV1H_COMPOUND_NONE = 0, AV1H_COMPOUND_AVERAGE = 1, AV1H_COMPOUND_DISTANCE = 2,
  AV1H_COMPOUND_WEDGE = 3, AV1H_COMPOUND_DIFFWTD = 4
};
enum { AV1H_RESTORE_NONE = 0, AV1H_RESTORE_SWITCHABLE = 1, AV1H_RESTORE_WIENER = 2,
       AV1H_RESTORE_SGRPROJ = 3 };

/* TransformBlock (reference av1gpu.hpp:187-208). */
typedef struct av1h_transform {
  int32_t x, y, w, h;        /* position inside the prediction block, size      */
  int32_t coef_offset;       /* raster start in the coefficient (or residual) pool */
  int32_t eob;               /* 0: the residual is zero                          */
  int32_t tx_type;           /* 0..15, Tx_Type                                   */
  int32_t level_offset, qindex, dc_step, ac_step, denominator, qm_off;
} av1h_transform;

/* ObmcCandidate (reference av1gpu.hpp:215-225). */
typedef struct av1h_obmc {
  int32_t pass, x, y, w, h;
  int32_t mv_x, mv_y;
  int32_t filter, filter_v;
  int32_t ref;
} av1h_obmc;

/* BlockSyntax (reference av1gpu.hpp:229-342); `transforms` and `obmc` are the
 * ranges [tx_first, tx_first + tx_count) and [obmc_first, obmc_first +
 * obmc_count) of the frame's transform and OBMC arrays. */
typedef struct av1h_block {
  int32_t x, y, w, h, plane, bit_depth;
  int32_t is_intra, intrabc;
  int32_t mode, have_left, have_above, have_above_right, have_below_left;
  int32_t angle_delta, filter_type, use_filter_intra, filter_intra_mode, cfl_alpha;
  int32_t lossless;
  int32_t use_palette, palette_size, palette[8], palette_off, palette_stride;
  int32_t ref[2];
  int32_t mv[2][2];
  int32_t filter, filter_v, motion_mode, compound, blend[2], wedge_index, mask_sign;
  int32_t interintra, interintra_mode, interintra_wedge;
  int32_t warp[2];
  int32_t warp_params[2][6];
  int32_t mask_warp[2];
  int32_t residual_provided;
  int32_t tx_first, tx_count, obmc_first, obmc_count;
} av1h_block;

/* DeblockEdge (reference av1gpu.hpp:347-358). */
typedef struct av1h_deblock_edge {
  int32_t plane, x, y, vertical, length, level, filter_size;
} av1h_deblock_edge;

/* CdefBlock (reference av1gpu.hpp:361-369). */
typedef struct av1h_cdef_block {
  int32_t plane, x, y, pri_strength, sec_strength, damping, dir;
} av1h_cdef_block;

/* RestorationUnit (reference av1gpu.hpp:418-426). `type` is a RestorationType
 * value, the way the reference planner reads it (engine.cpp:1425-1448):
 * AV1H_RESTORE_NONE skips the unit, AV1H_RESTORE_WIENER is the Wiener filter and
 * any other value the self guided filter. */
typedef struct av1h_lr_unit {
  int32_t type;
  int32_t wiener[2][3];
  int32_t sgr_set;
  int32_t sgr_xqd[2];
} av1h_lr_unit;

/* FilmGrainParams (reference av1gpu.hpp:434-468). */
typedef struct av1h_grain {
  int32_t apply, seed;
  int32_t num_y_points, point_y_value[16], point_y_scaling[16];
  int32_t chroma_scaling_from_luma;
  int32_t num_cb_points, point_cb_value[16], point_cb_scaling[16];
  int32_t num_cr_points, point_cr_value[16], point_cr_scaling[16];
  int32_t grain_scaling_minus_8, ar_coeff_lag;
  int32_t ar_coeffs_y[24], ar_coeffs_cb[25], ar_coeffs_cr[25];
  int32_t ar_coeff_shift_minus_6, grain_scale_shift;
  int32_t cb_mult, cb_luma_mult, cb_offset, cr_mult, cr_luma_mult, cr_offset;
  int32_t overlap_flag, clip_to_restricted_range;
} av1h_grain;

#define AV1H_MAX_REFS 8

/* The scalar part of FrameSyntax (reference av1gpu.hpp:476-537). */
typedef struct av1h_frame_header {
  int32_t width, coded_width, height, bit_depth, ss_x, ss_y;
  int32_t base_q_idx, sharpness, enable_intra_edge_filter;
  int32_t num_reference_slots;
  int32_t reference_slots[AV1H_MAX_REFS];
  int32_t reference_width[7], reference_height[7];
  int32_t color_matrix, color_range;
  int32_t restoration_type[3], restoration_unit_log2[3], restoration_unit_cols[3];
  av1h_grain grain;
} av1h_frame_header;

/* One frame's syntax: the header plus borrowed pointers to the pools. The
 * library copies what it needs before av1h_decode returns. */
typedef struct av1h_frame_desc {
  av1h_frame_header hdr;
  const av1h_block* blocks;
  const av1h_transform* transforms;
  const av1h_obmc* obmc;
  const int32_t* coefficients;
  const int32_t* residual;
  const int16_t* levels;
  const uint8_t* qm_pool;
  const uint8_t* palette_map;
  const av1h_deblock_edge* deblock;
  const av1h_cdef_block* cdef;
  const av1h_lr_unit* restoration[3];
  int64_t num_blocks, num_transforms, num_obmc, num_coefficients, num_residual;
  int64_t num_levels, num_qm, num_palette, num_deblock, num_cdef;
  int64_t num_restoration[3];
} av1h_frame_desc;

/* ------------------------------------------------------------- wire format */
/*
 * A frame's syntax as one byte string, for fixtures and replay: the magic
 * "AV1WSYN1", a uint32 version (1), a uint32 byte size of av1h_frame_header, the
 * header, then thirteen int64 counts (blocks, transforms, obmc, coefficients,
 * residual, levels, qm, palette, deblock, cdef, lr0, lr1, lr2) and the thirteen
 * arrays in that order, each starting on an 8 byte boundary. Parsing only
 * points into the buffer; it copies nothing.
 */
#define AV1H_WIRE_MAGIC "AV1WSYN1"
#define AV1H_WIRE_VERSION 1u

static inline size_t av1h__pad8(size_t v) { return (v + 7u) & ~(size_t)7u; }

static inline size_t av1h__wire_elem(int i) {
  static const size_t sizes[13] = {sizeof(av1h_block), sizeof(av1h_transform),
                                   sizeof(av1h_obmc), 4, 4, 2, 1, 1,
                                   sizeof(av1h_deblock_edge), sizeof(av1h_cdef_block),
                                   sizeof(av1h_lr_unit), sizeof(av1h_lr_unit),
                                   sizeof(av1h_lr_unit)};
  return sizes[i];
}

static inline void av1h__wire_counts(const av1h_frame_desc* d, int64_t n[13]) {
  n[0] = d->num_blocks; n[1] = d->num_transforms; n[2] = d->num_obmc;
  n[3] = d->num_coefficients; n[4] = d->num_residual; n[5] = d->num_levels;
  n[6] = d->num_qm; n[7] = d->num_palette; n[8] = d->num_deblock; n[9] = d->num_cdef;
  n[10] = d->num_restoration[0]; n[11] = d->num_restoration[1];
  n[12] = d->num_restoration[2];
}

static inline size_t av1h_wire_size(const av1h_frame_desc* d) {
  int64_t n[13];
  size_t size = av1h__pad8(16 + sizeof(av1h_frame_header)) + 13 * 8;
  int i;
  av1h__wire_counts(d, n);
  for (i = 0; i < 13; ++i) size += av1h__pad8((size_t)n[i] * av1h__wire_elem(i));
  return size;
}

/* Writes the wire form of `d` into `out` (av1h_wire_size bytes); returns the
 * number of bytes written. */
static inline size_t av1h_wire_write(const av1h_frame_desc* d, uint8_t* out) {
  const void* arrays[13];
  int64_t n[13];
  const uint32_t version = AV1H_WIRE_VERSION, hsize = (uint32_t)sizeof(av1h_frame_header);
  size_t pos = 0;
  int i;
  av1h__wire_counts(d, n);
  arrays[0] = d->blocks; arrays[1] = d->transforms; arrays[2] = d->obmc;
  arrays[3] = d->coefficients; arrays[4] = d->residual; arrays[5] = d->levels;
  arrays[6] = d->qm_pool; arrays[7] = d->palette_map; arrays[8] = d->deblock;
  arrays[9] = d->cdef; arrays[10] = d->restoration[0]; arrays[11] = d->restoration[1];
  arrays[12] = d->restoration[2];
  memcpy(out, AV1H_WIRE_MAGIC, 8);
  memcpy(out + 8, &version, 4);
  memcpy(out + 12, &hsize, 4);
  memcpy(out + 16, &d->hdr, sizeof(av1h_frame_header));
  pos = av1h__pad8(16 + sizeof(av1h_frame_header));
  for (i = 0; i < 13; ++i) {
    memcpy(out + pos, &n[i], 8);
    pos += 8;
  }
  for (i = 0; i < 13; ++i) {
    const size_t bytes = (size_t)n[i] * av1h__wire_elem(i);
    if (bytes) memcpy(out + pos, arrays[i], bytes);
    if (av1h__pad8(bytes) > bytes) memset(out + pos + bytes, 0, av1h__pad8(bytes) - bytes);
    pos += av1h__pad8(bytes);
  }
  return pos;
}

/* Points `d` into `buf` (which must stay alive and be 8 byte aligned). Returns
 * AV1H_OK, or AV1H_ERR_INVALID when the buffer is not a complete version 1
 * wire frame. */
static inline int av1h_wire_parse(const uint8_t* buf, size_t len, av1h_frame_desc* d) {
  uint32_t version = 0, hsize = 0;
  int64_t n[13];
  const void* arrays[13];
  size_t pos;
  int i;
  if (!buf || !d || len < 16 || memcmp(buf, AV1H_WIRE_MAGIC, 8) != 0) return AV1H_ERR_INVALID;
  memcpy(&version, buf + 8, 4);
  memcpy(&hsize, buf + 12, 4);
  if (version != AV1H_WIRE_VERSION || hsize != sizeof(av1h_frame_header)) return AV1H_ERR_INVALID;
  pos = av1h__pad8(16 + sizeof(av1h_frame_header));
  if (len < pos + 13 * 8) return AV1H_ERR_INVALID;
  memset(d, 0, sizeof(*d));
  memcpy(&d->hdr, buf + 16, sizeof(av1h_frame_header));
  for (i = 0; i < 13; ++i) {
    memcpy(&n[i], buf + pos, 8);
    if (n[i] < 0 || n[i] > ((int64_t)1 << 40)) return AV1H_ERR_INVALID;
    pos += 8;
  }
  for (i = 0; i < 13; ++i) {
    const size_t bytes = (size_t)n[i] * av1h__wire_elem(i);
    if (len < pos + bytes) return AV1H_ERR_INVALID;
    arrays[i] = bytes ? (const void*)(buf + pos) : (const void*)0;
    pos += av1h__pad8(bytes);
  }
  d->blocks = (const av1h_block*)arrays[0];
  d->transforms = (const av1h_transform*)arrays[1];
  d->obmc = (const av1h_obmc*)arrays[2];
  d->coefficients = (const int32_t*)arrays[3];
  d->residual = (const int32_t*)arrays[4];
  d->levels = (const int16_t*)arrays[5];
  d->qm_pool = (const uint8_t*)arrays[6];
  d->palette_map = (const uint8_t*)arrays[7];
  d->deblock = (const av1h_deblock_edge*)arrays[8];
  d->cdef = (const av1h_cdef_block*)arrays[9];
  d->restoration[0] = (const av1h_lr_unit*)arrays[10];
  d->restoration[1] = (const av1h_lr_unit*)arrays[11];
  d->restoration[2] = (const av1h_lr_unit*)arrays[12];
  d->num_blocks = n[0]; d->num_transforms = n[1]; d->num_obmc = n[2];
  d->num_coefficients = n[3]; d->num_residual = n[4]; d->num_levels = n[5];
  d->num_qm = n[6]; d->num_palette = n[7]; d->num_deblock = n[8]; d->num_cdef = n[9];
  d->num_restoration[0] = n[10]; d->num_restoration[1] = n[11];
  d->num_restoration[2] = n[12];
  return AV1H_OK;
}

/* Byte offset of hdr.reference_slots inside a wire buffer: a driver that
 * replays a clip patches the slots a frame names there, because slot numbers
 * are handed out by the decoder at run time. */
#define AV1H_WIRE_REFSLOTS_OFFSET (16 + offsetof(av1h_frame_header, reference_slots))
#define AV1H_WIRE_NUMREFS_OFFSET (16 + offsetof(av1h_frame_header, num_reference_slots))

/* ---------------------------------------------------------------- decoding */

typedef struct av1h_ctx av1h_ctx;

/* DeviceInfo (reference av1gpu.hpp:556-566). */
typedef struct av1h_device_info {
  char name[128];
  char arch[64];
  int32_t compute_units;
  int32_t l2_bytes;
  int64_t global_mem;
  int32_t clock_khz;
  int32_t mem_clock_khz;
  int32_t mem_bus_width;
} av1h_device_info;

/* DecodeStats (reference av1gpu.hpp:544-552), plus the device time of the two
 * halves measured with HIP events on the decode stream. */
typedef struct av1h_stats {
  int32_t blocks, waves, kernel_launches;
  double reconstruct_ms, filter_ms, readback_ms, total_ms;
  double device_recon_ms, device_filter_ms;
} av1h_stats;

/* Geometry of a decoded picture (what Frame::allocate needs). */
typedef struct av1h_frame_info {
  int32_t width, height, bit_depth, ss_x, ss_y;
} av1h_frame_info;

/* The output planes of a read back: three planes of 16 bit samples, row pitch
 * in samples. Rows of the display picture are written, nothing else. */
typedef struct av1h_planes {
  uint16_t* plane[3];
  int32_t stride[3];
} av1h_planes;

/* Creates a decoder on HIP device `device_index` (its reference pool, streams
 * and work buffers). Replaces Decoder::create with BackendKind::HIP
 * (reference src/core/engine.cpp:1509-1558). */
int av1h_create(int device_index, av1h_ctx** out);
void av1h_destroy(av1h_ctx* ctx);
const char* av1h_last_error(const av1h_ctx* ctx);
int av1h_get_device_info(av1h_ctx* ctx, av1h_device_info* out);
/* Describes HIP device `device_index` without creating a context (for
 * enumerate_devices, reference src/core/factory.cpp:40-54): AV1H_OK for a
 * gfx950 device, AV1H_ERR_UNSUPPORTED (info filled) for another HIP device,
 * AV1H_ERR_NO_DEVICE past the last one. */
int av1h_describe_device(int device_index, av1h_device_info* out);

/* Decodes one frame (reference Decoder::decode_reference, av1gpu.hpp:635;
 * Decoder::decode, 733; Decoder::decode_reference_reserved, 645).
 *   keep_reference 0: decode and keep no slot (`*slot_out` = -1);
 *   keep_reference 1: store the pre-grain picture in a reference slot - the
 *                     reserved one when `reserved_slot` >= 0, else the lowest
 *                     free one - and return it in `*slot_out`.
 * `out` (nullable) receives the post-grain display picture; a null `out`
 * leaves it on the device. `info` (nullable) receives the geometry. The call
 * returns once the picture is complete on the device. */
int av1h_decode(av1h_ctx* ctx, const av1h_frame_desc* frame, int keep_reference,
                int reserved_slot, const av1h_planes* out, av1h_frame_info* info,
                int* slot_out);

/* av1h_decode of a wire frame (see the wire format above). */
int av1h_decode_wire(av1h_ctx* ctx, const uint8_t* wire, size_t len, int keep_reference,
                     int reserved_slot, const av1h_planes* out, av1h_frame_info* info,
                     int* slot_out);

/* Pre-planning (reference Decoder::plan_frame, av1gpu.hpp:670-689, and
 * Decoder::decode_reference_planned, 700-709): av1h_plan_frame flattens the
 * syntax and uploads it into device memory owned by the returned plan, so a
 * pipelined caller can prepare frame N+1 while frame N decodes and the decode
 * itself starts from HBM. A plan is consumed by nothing: decode it (any number
 * of times) with av1h_decode_planned, and free it with av1h_discard_plan. The
 * reference slots a plan names are resolved when it is decoded. */
int av1h_plan_frame(av1h_ctx* ctx, const av1h_frame_desc* frame, void** plan_out);
int av1h_plan_frame_wire(av1h_ctx* ctx, const uint8_t* wire, size_t len, void** plan_out);
int av1h_decode_planned(av1h_ctx* ctx, void* plan, int keep_reference, int reserved_slot,
                        const av1h_planes* out, av1h_frame_info* info, int* slot_out);
void av1h_discard_plan(av1h_ctx* ctx, void* plan);

/* Sizes the context's reference pool to `slots` (1..32, default 32) slots;
 * allowed while no slot is in use, the pool is re-allocated at the next
 * decode. Callers that admit streams against a memory budget use it with
 * av1h_stream_bytes. No counterpart in the reference (a fixed 32 slot pool,
 * cuda_backend.cpp:242-279). */
int av1h_set_pool_slots(av1h_ctx* ctx, int slots);

/* Device bytes a context holds for pictures of this geometry with a pool of
 * `pool_slots` slots (and decode-time NV12 at `nv12_pitch`, 0 for none),
 * syntax plans excluded; -1 for invalid arguments. */
int64_t av1h_stream_bytes(int width, int height, int bit_depth, int ss_x, int ss_y, int pool_slots,
                          size_t nv12_pitch);

/* Reserves a reference slot (reference Decoder::reserve_reference,
 * av1gpu.hpp:661). Returns the slot or a negative status. */
int av1h_reserve_reference(av1h_ctx* ctx);

/* Releases a slot (reference Decoder::release_reference, av1gpu.hpp:736).
 * Unknown and already released slots are ignored. */
void av1h_release_reference(av1h_ctx* ctx, int slot);

/* Bytes an NV12 picture of this geometry and pitch takes, or 0 when it cannot
 * be packed (not 4:2:0, odd pitch, pitch narrower than a chroma row): the rule
 * of the reference's nv12_buffer_size (src/core/factory.cpp:236-243). */
size_t av1h_nv12_size(int width, int height, int ss_x, int ss_y, size_t pitch);

/* Packs a slot's picture into NV12 on the device and copies it to `out`
 * (reference Decoder::to_nv12, av1gpu.hpp:745). The most recently decoded
 * slot packs the post-grain picture, an older slot its pre-grain reference
 * (reference src/backends/cpu/cpu_backend.cpp:557-587). `slot` -1 packs the
 * last decoded picture. */
int av1h_to_nv12(av1h_ctx* ctx, int slot, size_t pitch, uint8_t* out, size_t out_bytes);
/* With `out` null, av1h_to_nv12 packs on the device only: the picture stays in
 * the context's pack buffer, returned here (valid until the next pack on this
 * context), for consumers that take the device picture (no counterpart in the
 * reference, whose to_nv12 always copies to the host). */
const uint8_t* av1h_nv12_device_buffer(const av1h_ctx* ctx);
/* Streamed NV12 pack (replaces Decoder::nv12_pack_begin / nv12_pack_wait /
 * nv12_pack_release, reference include/av1gpu/av1gpu.hpp:748-773): begin
 * queues the pack of `slot` (the slot rule of av1h_to_nv12) and its copy into
 * page-locked memory on the decode stream and returns a handle (0-2) at once,
 * or a negative status (AV1H_ERR_UNSUPPORTED: every handle is in flight, pack
 * synchronously instead). wait blocks until the copy is done, reports the
 * byte count and copies the bytes when `out` is given (out_bytes at least
 * that); the handle stays in flight until release, which waits for the copy
 * first. */
int av1h_nv12_pack_begin(av1h_ctx* ctx, int slot, size_t pitch);
int av1h_nv12_pack_wait(av1h_ctx* ctx, int handle, uint8_t* out, size_t out_bytes, size_t* bytes);
void av1h_nv12_pack_release(av1h_ctx* ctx, int handle);
/* With `pitch` > 0 (even), every decode on this context also packs the
 * finished picture into NV12 at that pitch on the device, in the same launch
 * sequence (4:2:0 frames only); av1h_to_nv12 of that picture at that pitch
 * then only copies. 0 turns it off. No counterpart in the reference, whose
 * decoder packs only when asked (Decoder::to_nv12, av1gpu.hpp:745). */
int av1h_set_decode_nv12(av1h_ctx* ctx, size_t pitch);

/* Captured launches. Mode 1 (the default): the first batch of each frame
 * count and set of kernels with work this context leads records its launch
 * sequence into a HIP graph, each of those kernels at its full persistent
 * grid reading its work counts from the batch's device tables; later
 * batches of that shape replay it with one launch. Mode 2: one graph per
 * frame count with every kernel launched (kernels without work exit at
 * once), for callers whose batches change shape often (the serving loop's
 * launchers): fewer captures, each replay dispatching every kernel. Mode 0:
 * every decode launches eagerly. Decodes with stage timing or the queue
 * trace on launch eagerly. With captured launches the stats'
 * device_recon_ms holds the whole batch and device_filter_ms 0. */
int av1h_set_graphs(av1h_ctx* ctx, int mode);

/* Launch shapes of the batches this context leads, the parameters an
 * autotuner searches (av1ongpu_amd/autotune.py does): for every kernel the
 * persistent workgroups per compute unit of its grid, the queue's, and the
 * sets of CDEF blocks a workgroup stages together (1..4); CDEF's chroma
 * blocks run in the luma launch, so "cdef_chroma" has no effect.
 * av1h_tune_params gives
 * their count and av1h_tune_name their names; setting them drops the
 * captured launch sequences. No counterpart in the reference (fixed launch
 * shapes, av1_cuda.cu). */
int av1h_tune_params(void);
const char* av1h_tune_name(int index);
int av1h_tune_get(const av1h_ctx* ctx, int32_t* params, int n);
int av1h_tune_set(av1h_ctx* ctx, const int32_t* params, int n);

/* With `enable`, every decode on this context also computes a 64 bit
 * checksum of its output picture on the device (the display samples of Y,
 * U, V in raster order, the i-th multiplied by SplitMix64's finaliser of
 * i + 1, summed modulo 2^64), read back with the decode's completion:
 * callers that keep pictures on the device compare outputs without copying
 * them. No counterpart in the reference. */
int av1h_set_output_hash(av1h_ctx* ctx, int enable);
int av1h_last_hash(av1h_ctx* ctx, uint64_t* out);

/* YUV -> RGB24 of a slot (reference Decoder::to_rgb, av1gpu.hpp:739): `matrix`
 * 0 BT.601, 1 BT.709, 2 BT.2020, 3 identity; `range` 0 limited, 1 full. */
int av1h_to_rgb(av1h_ctx* ctx, int slot, int matrix, int range, uint8_t* out,
                size_t out_bytes);

/* Copies a slot's picture (same slot rule as av1h_to_nv12) to the host; a
 * null `out` returns only the geometry in `info`. */
int av1h_read_slot(av1h_ctx* ctx, int slot, const av1h_planes* out, av1h_frame_info* info);

/* Drops every slot and clears the pool (reference Decoder::reset). */
void av1h_reset(av1h_ctx* ctx);

int av1h_get_stats(const av1h_ctx* ctx, av1h_stats* out);

/* Stage timing: with timing enabled, every decode records HIP events on the
 * decode stream between its stages and accumulates the device time and the
 * count of each stage, kernels only, the snapshot copies between stages
 * excluded (the reference's env-gated timers,
 * src/backends/cuda/cuda_backend.cpp:1048-1056, as an API). */
enum {
  AV1H_STAGE_INTER = 0,   /* motion compensation and every inverse transform */
  AV1H_STAGE_QUEUE = 1,   /* intra prediction: the dependency queue          */
  AV1H_STAGE_DEBLOCK = 2, /* both deblocking passes                          */
  AV1H_STAGE_CDEF = 3,    /* CDEF                                            */
  AV1H_STAGE_SUPERRES = 4,/* upscaling (the frame and the pre-CDEF copy)     */
  AV1H_STAGE_LR = 5,      /* loop restoration                                */
  AV1H_STAGE_STORE = 6,   /* reference store                                 */
  AV1H_STAGE_GRAIN = 7,   /* template synthesis + blend                      */
  AV1H_STAGES = 8
};
void av1h_set_stage_timing(av1h_ctx* ctx, int enable);
/* Accumulated device milliseconds, decodes, and compulsory HBM bytes per stage
 * since timing was enabled (a stage a frame skips is not counted for it).
 * Compulsory bytes, at 16 bit samples: motion compensation and transforms 2 B
 * written per inter sample, 2 B per reference read, 2 B of coefficients per
 * transform sample; intra prediction 2 B written per sample; deblocking, CDEF, grain and the reference store 4 B (one read, one
 * write) per sample they cover; restoration 4 B per sample plus the two
 * boundary rows above and below each stripe; upscaling 2 B per coded and 2 B
 * per upscaled sample, for each picture it upscales. Any argument may be
 * null. */
int av1h_get_stage_times(const av1h_ctx* ctx, double ms[AV1H_STAGES], int64_t count[AV1H_STAGES],
                         int64_t bytes[AV1H_STAGES]);

/* Diagnostics of the dependency queue: with tracing on, each decode records
 * for every intra block (in queue order) four u64 values: the device clock
 * (wall_clock64, 100 MHz on gfx950) when a wave took the block, when its
 * neighbours were final, when it was published, and the workgroup that ran it.
 * av1h_get_queue_trace copies the last decode's records (4 per block) and
 * returns the block count, or that count with nothing copied when `cap` is
 * too small. No counterpart in the reference (its schedule is on the host). */
void av1h_set_queue_trace(av1h_ctx* ctx, int enable);
int64_t av1h_get_queue_trace(av1h_ctx* ctx, uint64_t* out, int64_t cap);

/* Diagnostics: how many polls a queued block's dependency wait makes before
 * it gives up, for the frames of this context decoded from now on (0 or
 * less: the default, 4,194,304, about a second). A wait that gives up fails
 * the decode with AV1H_ERR_INVALID and the context decodes on. A small limit
 * forces that path from the host for tests of it. No counterpart in the
 * reference (its schedule is on the host). */
void av1h_set_wait_limit(av1h_ctx* ctx, int polls);

/* --------------------------------------------------------------- batching */
/*
 * Multi-stream batching: frames of several independent streams (each with its
 * own context) decoded together, one launch per stage group for the whole
 * batch. `frames[i]` is decoded by `ctxs[i]` exactly as av1h_decode would
 * (keep_reference 1, no read back); slots land in `slots_out[i]`. All contexts
 * must live on the same device and be distinct (at most 128). Each stream's
 * result is bit identical to decoding it alone: every kernel runs the work of
 * all frames of the batch in one launch, and the dependency queue hands out
 * the intra blocks of all frames from one merged order.
 */
int av1h_decode_batch(av1h_ctx* const* ctxs, const av1h_frame_desc* const* frames, int count,
                      int* slots_out);

/* The batch of planned frames: `plans[i]` (from av1h_plan_frame on
 * `ctxs[i]`) decoded by `ctxs[i]` as av1h_decode_planned would with
 * keep_reference `keep[i]` and reserved slot `reserved[i]` (null arrays: keep
 * 1, reserved -1). Contexts must be distinct; the batch's launches go to the
 * stream of ctxs[0], errors to av1h_last_error(ctxs[0]). */
int av1h_decode_planned_batch(av1h_ctx* const* ctxs, void* const* plans, const int* keep,
                              const int* reserved, int count, int* slots_out);
/* The same with the launches on `launcher`'s stream (its captured launch
 * sequences, its batch tables), which need not be one of the batch's
 * contexts: a caller that forms batches from changing sets of streams keeps
 * one launcher, whose captured sequences then serve them all (one per batch
 * size with av1h_set_graphs mode 2, as the serving loop's launchers use). */
int av1h_decode_planned_batch_on(av1h_ctx* launcher, av1h_ctx* const* ctxs, void* const* plans,
                                 const int* keep, const int* reserved, int count, int* slots_out);
/* The same batch launched without waiting for it (no counterpart in the
 * reference): returns once the batch is queued on the launcher's stream,
 * with its slots in slots_out; av1h_batch_finish(launcher) waits for the
 * oldest launched batch and returns its status. At most two batches per
 * launcher are in flight (a third launch first finishes the oldest), so the
 * host prepares batch r + 1 while the device runs batch r. Until its batch
 * is finished, only the same launcher may put a context's next frame in a
 * batch; the context's finished pictures can be read and packed (to_nv12,
 * to_rgb, read_slot, the NV12 ring) while its frames are in flight only
 * when it launched them itself (they are then queued behind them on its
 * stream): an older slot, or the latest when that frame had no film grain. */
int av1h_decode_planned_batch_async(av1h_ctx* launcher, av1h_ctx* const* ctxs, void* const* plans, const int* keep,
                                    const int* reserved, int count, int* slots_out);
int av1h_batch_finish(av1h_ctx* launcher);
/* Launched batches not yet finished. */
int av1h_batch_pending(av1h_ctx* launcher);
/* 1 when a launch of `plan` on ctx would wait for every batch in flight on
 * it: a device buffer of ctx has to grow for the plan's picture (the first
 * frame, a larger resolution, a wider NV12 pack), or stage timing or the
 * queue trace is on (they take the waiting path); 0 when not. A caller that
 * finishes its own batches in order (the ingest pipeline) finishes them
 * before such a launch. */
int av1h_plan_needs_idle(av1h_ctx* ctx, void* plan);

/* ------------------------------------------------------ superblock costs */
/*
 * The queue's schedule. By default (the static schedule) the dependency queue
 * hands intra blocks out by superblock anti-diagonal. A schedule from costs
 * orders superblocks by their longest remaining path through the superblock
 * dependency graph (left, above, above-left, above-right and intra block copy
 * sources), each superblock weighted by its predicted work: blocks of the
 * superblock with the longest path go first, coding order inside one. Every
 * dependency still comes earlier than what waits for it, so the decode is
 * correct under either schedule and the pixels are the same.
 * No counterpart in the reference (engine.cpp:1080-1302 orders by wave).
 */
enum { AV1H_CM_FEATURES = 24, AV1H_CM_PARAMS = 5825 };

/* Superblocks (64x64 luma) of a plan and their 24 syntax features each:
 * blocks; transforms by larger side 4/8/16/32/64; nonzero coefficients / 64;
 * intra, inter, compound blocks; fraction of inter blocks with subpel
 * motion; directional, smooth/Paeth, CFL/palette/filter intra blocks; queued
 * and inter samples / 4096; largest transform side / 64; lossless or raw
 * residual blocks; intra block copies; chroma blocks; warp/mask/OBMC/
 * inter-intra blocks; edge filter on; queued blocks; bit depth above 8.
 * Returns the superblock count; `features` (count x 24) may be null. */
int av1h_plan_sb_features(const void* plan, float* features);
/* Host time of the plan's making: flattening and upload (ms). */
int av1h_plan_host_times(const void* plan, double* flatten_ms, double* upload_ms);
/* Dependency levels of the plan's queue (its critical path in blocks; 0
 * when nothing is queued): what bounds the frame's queue time. */
int av1h_plan_levels(const void* plan);
/* counts[0..3]: queued blocks; blocks that read their edges from packed
 * words (one load per word is both wait and read); blocks that publish
 * packed edge words; blocks whose cells no other block waits on. */
int av1h_plan_queue_counts(const void* plan, int32_t* counts);
/* After a decode of `plan` alone with the queue trace on: the queue work
 * (sample time after the wait, per block, summed) of every superblock in
 * microseconds. Returns the superblock count. */
int av1h_plan_sb_times(av1h_ctx* ctx, const void* plan, float* us);
/* Orders the plan's queue by `sb_cost_us` (one per superblock; null restores
 * the static schedule). */
int av1h_plan_schedule(av1h_ctx* ctx, void* plan, const float* sb_cost_us);
/* The queue in another order (no counterpart in the reference, whose
 * scheduler hands blocks out by wave): AV1H_ORDER_LEVEL by dependency level
 * (with costs: the longest remaining path first inside a level; without:
 * the static order), AV1H_ORDER_CODING in coding order, superblock row after
 * row (the static row-order schedule), AV1H_ORDER_PATH the longest remaining
 * path first, paths summed from `sb_cost_us` (per superblock; null: each
 * block costs its sample count). A batch merges its frames' queues by key,
 * the frames' last keys aligned: the level, the remaining path in quarter
 * units, or runs of 64 entries in coding order. Pixels do not depend on the
 * order. */
enum { AV1H_ORDER_LEVEL = 0, AV1H_ORDER_CODING = 1, AV1H_ORDER_PATH = 2 };
int av1h_plan_schedule_order(av1h_ctx* ctx, void* plan, int order, const float* sb_cost_us);

/* The cost model (a 24-64-64-1 FP32 perceptron trained and run by HIP
 * kernels on the device). */
typedef struct av1h_costmodel av1h_costmodel;
int av1h_cm_create(int device, uint64_t seed, av1h_costmodel** out);
void av1h_cm_destroy(av1h_costmodel* model);
/* W1[64][24], b1[64], W2[64][64], b2[64], W3[64], b3 */
int av1h_cm_set_params(av1h_costmodel* model, const float* params);
int av1h_cm_get_params(av1h_costmodel* model, float* params);
/* feature standardisation and the time scale the network predicts in */
int av1h_cm_set_norm(av1h_costmodel* model, const float* mean, const float* dev, float y_mean, float y_dev);
int av1h_cm_get_norm(av1h_costmodel* model, float* mean, float* dev, float* y_mean, float* y_dev);
int av1h_cm_predict(av1h_costmodel* model, const float* features, int64_t n, float* out_us);
/* av1h_cm_predict that allocates nothing (no counterpart in the reference:
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or a serving loop that forbids it). The features and
 * predictions are staged in a static device array of AV1H_CM_NOALLOC_MAX
 * superblocks (25 floats each, 400 KiB), whose address each device looks up
 * when a model is created on it; more superblocks than that are refused with
 * AV1H_ERR_UNSUPPORTED. Calls that overlap on one device share that array
 * and race. */
#define AV1H_CM_NOALLOC_MAX 4096
int av1h_cm_predict_noalloc(av1h_costmodel* model, const float* features, int64_t n, float* out_us);
/* Gradient of the mean squared error (standardised time) over the n samples,
 * and that loss: what training steps on. */
int av1h_cm_gradients(av1h_costmodel* model, const float* features, const float* y_us, int64_t n, float* grad,
                      float* loss);
/* Adam, batches of 64 samples in a seeded shuffle per epoch; `set_norm`
 * first sets the standardisation from these samples. */
int av1h_cm_train(av1h_costmodel* model, const float* features, const float* y_us, int64_t n, int epochs,
                  float lr, uint64_t seed, int set_norm, float* loss_per_epoch);
/* Predicts the plan's superblock costs and schedules its queue by them. */
int av1h_plan_schedule_model(av1h_ctx* ctx, void* plan, av1h_costmodel* model);

/* -------------------------------------------------------- ingest pipeline */
/*
 * One stream's frames through planner threads and an in-order decode worker
 * (reference: the planner and worker threads of dsfilter/decode_core.hpp:
 * 164-292 around Decoder::plan_frame / decode_reference_planned, kPipeDepth
 * frames in flight; here `planner_threads` planners). Each submitted frame is
 * flattened and uploaded by av1h_plan_frame on some planner while earlier
 * frames decode; the worker decodes in submission order with
 * av1h_decode_planned(keep, reserved) (no read back: the picture stays in its
 * slot, and in the NV12 buffer when the context packs NV12 with every
 * decode), then releases the slots the frame lists. At most `depth` frames
 * are in flight: submit blocks until one completes.
 */
typedef struct av1h_pipe av1h_pipe;
int av1h_pipe_create(av1h_ctx* ctx, int planner_threads, int depth, av1h_pipe** out);
/* The frame's arrays (desc's pointers, or the wire bytes) are borrowed until
 * the frame completes (av1h_pipe_wait past its number). Returns the frame's
 * sequence number (0, 1, ...) or a negative error. */
int64_t av1h_pipe_submit(av1h_pipe* pipe, const av1h_frame_desc* desc, int keep, int reserved,
                         const int32_t* release, int nrelease);
int64_t av1h_pipe_submit_wire(av1h_pipe* pipe, const uint8_t* wire, size_t len, int keep, int reserved,
                              const int32_t* release, int nrelease);
/* Waits until frames 0..seq are complete (seq < 0: every frame submitted).
 * Returns AV1H_OK or the status of the first failed frame among them, whose
 * number goes to *failed_seq (-1 when none failed). */
int av1h_pipe_wait(av1h_pipe* pipe, int64_t seq, int64_t* failed_seq);
/* A completed frame: its slot (-1 when not kept) and status; its times on the
 * av1h_server_now_ms clock: submitted, taken by a planner, planned (flattened
 * and uploaded), taken by the worker, complete on the device; how long its
 * flattening and its upload took; and, when the context checksums every
 * decoded picture (av1h_set_output_hash), the picture's checksum (else 0). */
typedef struct av1h_pipe_frame {
  int32_t slot, status;
  double submit_ms, plan_start_ms, planned_ms, decode_start_ms, done_ms;
  double flatten_ms, upload_ms;
  uint64_t hash;
  /* the worker's host time inside the frame's launch call and blocked
   * waiting for its result; the device time of its launches (first to last
   * event of its batch) */
  double launch_ms, wait_ms, device_ms;
} av1h_pipe_frame;
int av1h_pipe_result(av1h_pipe* pipe, int64_t seq, av1h_pipe_frame* out);
/* The first failure's message ("" when none). */
const char* av1h_pipe_error(av1h_pipe* pipe);
/* Waits for every submitted frame, then ends the threads. */
void av1h_pipe_destroy(av1h_pipe* pipe);

/* ---------------------------------------------------------------- serving */
/*
 * A serving loop over av1h_decode_planned_batch (no counterpart in the
 * reference; its nearest relative is the single stream pipelined ingest of
 * dsfilter/decode_core.hpp:164-292). Streams open with their geometry and
 * pool size and are admitted in order while the device memory of admitted
 * streams (av1h_stream_bytes) fits `hbm_budget`; the first stream is always
 * admitted. Frames are submitted as wire bytes (planned by `planner_threads`
 * host threads once the stream is admitted) or as plans; frames of a stream
 * decode in submission order. `launchers` scheduler threads, each with its
 * own launcher context and stream, form batches from the oldest planned
 * head frames of the idle streams, up to `max_batch` frames, so batches run
 * on the device at the same time and a long one (a large key frame) does not
 * hold back the frames arriving behind it. A closed stream gives its memory
 * back after its last frame.
 */
typedef struct av1h_server av1h_server;

typedef struct av1h_server_config {
  int32_t device;
  int32_t max_batch;        /* frames per batch, 1..128                        */
  int32_t planner_threads;  /* host threads flattening submitted wire frames   */
  int32_t capture;          /* 1: every completion carries a host copy         */
  int32_t hash_pictures;    /* 1: every completion carries a picture checksum  */
  int32_t launchers;        /* batches in flight at once (0: 2)                */
  int32_t heavy_levels;     /* > 0 and launchers > 1: frames whose queue has
                               more dependency levels than this (key frames)
                               run on the first heavy_launchers launchers,
                               the others on the rest                       */
  int32_t heavy_launchers;  /* launchers for heavy frames (0: 1)              */
  int32_t heavy_batch;      /* heavy frames per batch (0: 1); their queues
                               merge, so a batch of them takes about as long
                               as its longest critical path                 */
  int64_t hbm_budget;       /* bytes of admitted stream state                  */
  size_t nv12_pitch;        /* > 0: NV12 packed on the device with each frame   */
} av1h_server_config;

typedef struct av1h_completion {
  uint64_t tag;             /* the submitter's                                  */
  int32_t stream, slot;     /* slot the frame was kept in, or -1                */
  int32_t status;           /* AV1H_OK or the error                             */
  int32_t batch_frames;     /* frames of the batch it ran in                    */
  double arrival_ms, start_ms, done_ms; /* av1h_server_now_ms clock            */
  uint16_t* picture;        /* capture: Y, U, V display planes back to back     */
  int64_t picture_samples;  /* free with av1h_server_free_picture               */
  uint64_t hash;            /* checksum of the picture (hash_pictures)          */
  double plan_start_ms, planned_ms; /* planning of a wire frame (else arrival)  */
} av1h_completion;

typedef struct av1h_server_stats {
  int64_t frames, batches, admitted, waiting, active, budget_used;
  double busy_ms;           /* wall time with at least one batch decoding       */
  int32_t max_batch_frames;
} av1h_server_stats;

double av1h_server_now_ms(void);
int av1h_server_create(const av1h_server_config* cfg, av1h_server** out);
/* Stops the threads; frames not yet decoded are dropped. */
void av1h_server_destroy(av1h_server* server);
/* Returns the stream id (>= 0) or a negative status. */
int av1h_server_open_stream(av1h_server* server, int width, int height, int bit_depth, int ss_x, int ss_y,
                            int pool_slots);
/* `reserved_slot`: the slot to keep the frame in (-1: not kept); `release`:
 * slots given back once the frame is decoded (as av1h_release_reference). */
int av1h_server_submit_wire(av1h_server* server, int stream, const uint8_t* wire, size_t len,
                            int reserved_slot, const int32_t* release, int nrelease, double arrival_ms,
                            uint64_t tag);
/* The same with the wire bytes borrowed, not copied: they must stay valid
 * until the frame's completion is polled. */
int av1h_server_submit_wire_ref(av1h_server* server, int stream, const uint8_t* wire, size_t len,
                                int reserved_slot, const int32_t* release, int nrelease, double arrival_ms,
                                uint64_t tag);
/* A plan from av1h_plan_frame on any context of the server's device; the
 * caller keeps it and may submit it again (plans are read only). */
int av1h_server_submit_plan(av1h_server* server, int stream, void* plan, int reserved_slot,
                            const int32_t* release, int nrelease, double arrival_ms, uint64_t tag);
int av1h_server_close_stream(av1h_server* server, int stream);
/* Up to `max` completions, waiting up to `wait_ms` for the first. */
int av1h_server_poll(av1h_server* server, av1h_completion* out, int max, int wait_ms);
void av1h_server_free_picture(uint16_t* picture);
int av1h_server_get_stats(av1h_server* server, av1h_server_stats* out);
/* Launcher context `index` (0 .. launchers-1, the heavy-frame launchers
 * first), owned by the server: its launch shapes can be set with
 * av1h_tune_set before frames are submitted. NULL for a bad index. */
av1h_ctx* av1h_server_launcher(av1h_server* server, int index);
const char* av1h_server_last_error(av1h_server* server);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* AV1HIP_H */
