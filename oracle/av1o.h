/*
 * av1o.h - the CPU oracle: a plain C restatement of the reference's AV1
 * reconstruction path (trprkkk/AV1onGPU: kernels/av1gpu_kernels.h,
 * src/core/engine.cpp:519-1494, src/core/grain.cpp, src/core/factory.cpp,
 * src/core/tables.cpp and src/backends/cpu/cpu_backend.cpp).
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is the checker the GPU path is compared
 * against (tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * call it); the product library never links, loads or calls it.
 *
 * It consumes the same av1h_frame_desc records the C ABI takes (include/
 * av1hip.h), and mirrors the Decoder contract: a 32 slot reference pool that
 * stores the picture before film grain, and an output that carries it.
 */
#ifndef AV1O_H
#define AV1O_H

#include <stddef.h>
#include <stdint.h>

#include "av1hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct av1o_decoder av1o_decoder;

av1o_decoder* av1o_create(void);
void av1o_destroy(av1o_decoder* dec);
const char* av1o_last_error(const av1o_decoder* dec);

/* Decode one frame (reference CpuBackend::decode_impl, cpu_backend.cpp:149-553,
 * fed by build_frame_plan, engine.cpp:519-1494). keep_reference / reserved_slot
 * / out follow av1h_decode. Returns AV1H_OK or a negative status. */
int av1o_decode(av1o_decoder* dec, const av1h_frame_desc* frame, int keep_reference,
                int reserved_slot, const av1h_planes* out, av1h_frame_info* info,
                int* slot_out);
/* The same, from a wire buffer. */
int av1o_decode_wire(av1o_decoder* dec, const uint8_t* wire, size_t len, int keep_reference,
                     int reserved_slot, const av1h_planes* out, av1h_frame_info* info,
                     int* slot_out);

int av1o_reserve_reference(av1o_decoder* dec);
void av1o_release_reference(av1o_decoder* dec, int slot);
int av1o_to_nv12(av1o_decoder* dec, int slot, size_t pitch, uint8_t* out, size_t out_bytes);
int av1o_to_rgb(av1o_decoder* dec, int slot, int matrix, int range, uint8_t* out,
                size_t out_bytes);
int av1o_read_slot(av1o_decoder* dec, int slot, const av1h_planes* out, av1h_frame_info* info);
void av1o_reset(av1o_decoder* dec);
/* waves of the last decode (the wavefront schedule length, reference
 * DecodeStats::waves) */
int av1o_last_waves(const av1o_decoder* dec);
/* OpenMP threads the oracle uses (0 = the OpenMP default) */
void av1o_set_threads(int n);
int av1o_get_threads(void);

/* ---- single functions, for known-answer tests ------------------------- */

/* cos128 / sin128 (av1gpu_kernels.h:472-479) */
int av1o_cos128(int angle);
/* 1D inverse transform of 2^n points in place: kind 0 DCT, 1 ADST, 2 identity
 * (av1gpu_kernels.h:571-770) */
void av1o_tx1d(int32_t* t, int n, int kind, int r);
/* Full 2D inverse transform + add of one block into a w x h sample block of
 * stride `stride` (av1gpu_kernels.h:869-930). */
void av1o_itx_add(const int32_t* coef, int w, int h, int tx_type, int lossless, int bitdepth,
                  uint16_t* dst, int stride);
/* Transform_Row_Shift (tables.cpp:357-369) */
int av1o_tx_row_shift(int w, int h);
/* Scaling LUT (av1gpu_kernels.h:3157-3188): plane 0..2, 256 entries. */
void av1o_grain_scaling_lut(const av1h_grain* g, int plane, int* lut);
/* superres step/initial phase (av1gpu_kernels.h:2375-2383) */
void av1o_superres_params(int downscaled_w, int upscaled_w, int* step, int* initial);
/* setup shear (tables.cpp:444-475); returns warpValid */
int av1o_setup_shear(const int32_t* params, int* alpha, int* beta, int* gamma, int* delta);
/* Q16 colour matrix (tables.cpp:481-533) */
void av1o_color_matrix(int matrix, int range, int bit_depth, int m[8], int o[3]);
/* the filter table the predictions index: subpel, then warp, then upscale */
const int* av1o_spec_table_ptr(int* size);
/* The coefficient raster of every coded transform of a frame that carries
 * coded levels, dequantised on the host (into `raster`, num_coefficients
 * entries, zeros elsewhere); returns the transforms dequantised, or -1 when
 * the frame has no levels. */
int64_t av1o_dequant_frame(const av1h_frame_desc* frame, int32_t* raster);
/* The same for a wire frame; `cap` entries of room (at least the frame's
 * num_coefficients, else -2). */
int64_t av1o_dequant_wire(const uint8_t* wire, size_t len, int32_t* raster, int64_t cap);
/* FNV-1a 64 over bytes (the reference's --hash-check, tools/av1gpu_decode.cpp:133-139) */
uint64_t av1o_fnv1a64(const uint8_t* p, size_t n);

#ifdef __cplusplus
}
#endif

#endif
