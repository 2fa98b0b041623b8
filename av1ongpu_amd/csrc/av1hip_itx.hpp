/*
 * Inverse transforms on VALU for gfx950: the specification's integer
 * butterfly networks (the arithmetic of the reference's kernels/av1gpu_kernels.h:
 * 472-930), written as fully unrolled templates so one lane keeps a whole 1D
 * transform (up to 64 points) in VGPRs. Products are 64 bit (v_mad_i64_i32) and
 * rounded back by 12, exactly like the reference, so any int32 input gives the
 * reference's result. No MFMA: the per stage rounding and clamping is not a
 * dense contraction.
 */
#ifndef AV1HIP_ITX_HPP
#define AV1HIP_ITX_HPP

#include "spec_tables.h"

namespace av1hip {
namespace itx {

__device__ __forceinline__ constexpr int cos128(int angle) {
  return (angle & 255) <= 64    ? av1t_cos128[angle & 255]
         : (angle & 255) <= 128 ? -av1t_cos128[128 - (angle & 255)]
         : (angle & 255) <= 192 ? -av1t_cos128[(angle & 255) - 128]
                                : av1t_cos128[256 - (angle & 255)];
}
__device__ __forceinline__ constexpr int brev(int bits, int x) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
__device__ __forceinline__ int r12(int64_t v) { return (int)((v + 2048) >> 12); }

/* B(a, b, angle, flip) (av1gpu_kernels.h:521-545). The products are formed
 * in 64 bits as the reference does; when both inputs lie in [-2^17, 2^17),
 * as they do for every conforming stream up to 10 bits, every product is
 * below 2^29 in magnitude and each sum below 2^30, so 24-bit multiplies
 * (full rate, exact for such inputs) and 32-bit sums give the same bits.
 * Used up to 16 points: the 32- and 64-point transforms measured faster
 * without the test. */
template <int N>
__device__ __forceinline__ void bf(int (&t)[N], int a, int b, int angle, int flip) {
  const int ta = t[a], tb = t[b];
  if (N <= 16 &&
      __builtin_expect((((unsigned)ta + (1u << 17)) | ((unsigned)tb + (1u << 17))) < (1u << 18), 1)) {
    const int c = cos128(angle);
    int x, y;
    if ((angle & 63) == 32) {
      const int s = ta + tb, d = ta - tb;
      x = __mul24((angle & 64) ? s : d, c);
      y = __mul24((angle & 64) ? -d : s, c);
    } else {
      const int sn = cos128(angle - 64);
      x = __mul24(ta, c) - __mul24(tb, sn);
      y = __mul24(ta, sn) + __mul24(tb, c);
    }
    const int ra = (x + 2048) >> 12, rb = (y + 2048) >> 12;
    t[a] = flip ? rb : ra;
    t[b] = flip ? ra : rb;
    return;
  }
  int64_t x, y;
  const int64_t c = cos128(angle);
  if ((angle & 63) == 32) {
    const int s = (int)((unsigned)t[a] + (unsigned)t[b]);
    const int d = (int)((unsigned)t[a] - (unsigned)t[b]);
    x = (int64_t)((angle & 64) ? s : d) * c;
    y = (int64_t)((angle & 64) ? (int)(0u - (unsigned)d) : s) * c;
  } else {
    const int64_t sn = cos128(angle - 64);
    x = (int64_t)t[a] * c - (int64_t)t[b] * sn;
    y = (int64_t)t[a] * sn + (int64_t)t[b] * c;
  }
  const int ra = r12(x), rb = r12(y);
  t[a] = flip ? rb : ra;
  t[b] = flip ? ra : rb;
}

/* H(a, b, flip, r) (av1gpu_kernels.h:555-568) */
template <int N>
__device__ __forceinline__ void hd(int (&t)[N], int a, int b, int flip, int lo, int hi) {
  const int p = flip ? t[b] : t[a];
  const int q = flip ? t[a] : t[b];
  const int s = min(max(p + q, lo), hi);
  const int d = min(max(p - q, lo), hi);
  if (flip) {
    t[b] = s;
    t[a] = d;
  } else {
    t[a] = s;
    t[b] = d;
  }
}

/* inverse DCT of N = 2^n points (av1gpu_kernels.h:571-652) */
template <int N>
__device__ __forceinline__ void idct(int (&t)[N], int lo, int hi) {
  constexpr int n = N == 4 ? 2 : N == 8 ? 3 : N == 16 ? 4 : N == 32 ? 5 : 6;
  int c[N];
#pragma unroll
  for (int i = 0; i < N; ++i) c[i] = t[brev(n, i)];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = c[i];
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 16; ++i) bf(t, 32 + i, 63 - i, 63 - 4 * brev(4, i), 0);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 8; ++i) bf(t, 16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 16; ++i) hd(t, 32 + 2 * i, 33 + 2 * i, i & 1, lo, hi);
  }
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i) bf(t, 8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 8; ++i) hd(t, 16 + 2 * i, 17 + 2 * i, i & 1, lo, hi);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) bf(t, 62 - 4 * i - j, 33 + 4 * i + j, 60 - 16 * brev(2, i) + 64 * j, 1);
  }
  if (n >= 3) {
#pragma unroll
    for (int i = 0; i < 2; ++i) bf(t, 4 + i, 7 - i, 56 - 32 * i, 0);
  }
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i) hd(t, 8 + 2 * i, 9 + 2 * i, i & 1, lo, hi);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) bf(t, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) hd(t, 32 + 4 * i + j, 35 + 4 * i - j, i & 1, lo, hi);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) bf(t, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i);
  if (n >= 3) {
#pragma unroll
    for (int i = 0; i < 2; ++i) hd(t, 4 + 2 * i, 5 + 2 * i, i, lo, hi);
  }
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; ++i) bf(t, 14 - i, 9 + i, 48 + 64 * i, 1);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) hd(t, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, lo, hi);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) bf(t, 61 - 8 * i - j, 34 + 8 * i + j, 56 - 32 * i + (j >> 1) * 64, 1);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) hd(t, i, 3 - i, 0, lo, hi);
  if (n >= 3) bf(t, 6, 5, 32, 1);
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) hd(t, 8 + 4 * i + j, 11 + 4 * i - j, i, lo, hi);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; ++i) bf(t, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) hd(t, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, lo, hi);
  }
  if (n >= 3) {
#pragma unroll
    for (int i = 0; i < 4; ++i) hd(t, i, 7 - i, 0, lo, hi);
  }
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; ++i) bf(t, 13 - i, 10 + i, 32, 1);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) hd(t, 16 + 8 * i + j, 23 + 8 * i - j, i, lo, hi);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; ++i) bf(t, 59 - i, 36 + i, i < 4 ? 48 : 112, 1);
  }
  if (n >= 4) {
#pragma unroll
    for (int i = 0; i < 8; ++i) hd(t, i, 15 - i, 0, lo, hi);
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; ++i) bf(t, 27 - i, 20 + i, 32, 1);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      hd(t, 32 + i, 47 - i, 0, lo, hi);
      hd(t, 48 + i, 63 - i, 1, lo, hi);
    }
  }
  if (n >= 5) {
#pragma unroll
    for (int i = 0; i < 16; ++i) hd(t, i, 31 - i, 0, lo, hi);
  }
  if (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; ++i) bf(t, 55 - i, 40 + i, 32, 1);
#pragma unroll
    for (int i = 0; i < 32; ++i) hd(t, i, 63 - i, 0, lo, hi);
  }
}

/* ADST permutations (av1gpu_kernels.h:655-679) */
template <int N>
__device__ __forceinline__ void adst_in(int (&t)[N]) {
  int c[N];
#pragma unroll
  for (int i = 0; i < N; ++i) c[i] = t[i];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = c[(i & 1) ? i - 1 : N - 1 - i];
}
template <int N>
__device__ __forceinline__ void adst_out(int (&t)[N]) {
  constexpr int n = N == 8 ? 3 : 4;
  int c[N];
#pragma unroll
  for (int i = 0; i < N; ++i) c[i] = t[i];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int b3 = (i >> 3) & 1, b2 = (i >> 2) & 1, b1 = (i >> 1) & 1, b0 = i & 1;
    const int idx = (((b0 ^ b1) << 3) | ((b1 ^ b2) << 2) | ((b2 ^ b3) << 1) | b3) >> (4 - n);
    t[i] = (i & 1) ? -c[idx] : c[idx];
  }
}

/* ADST4 (av1gpu_kernels.h:682-712) */
__device__ __forceinline__ void adst4(int (&t)[4]) {
  const int64_t k1 = 1321, k2 = 2482, k3 = 3344, k4 = 3803;
  int64_t s0 = k1 * t[0], s1 = k2 * t[0], s2 = k3 * t[1], s3 = k4 * t[2];
  const int64_t s4 = k1 * t[2], s5 = k2 * t[3], s6 = k4 * t[3];
  const int b7 = (t[0] - t[2]) + t[3];
  s0 += s3;
  s1 -= s4;
  s3 = s2;
  s2 = k3 * b7;
  s0 += s5;
  s1 -= s6;
  t[0] = r12(s0 + s3);
  t[1] = r12(s1 + s3);
  t[2] = r12(s2);
  t[3] = r12(s0 + s1 - s3);
}
/* ADST8 / ADST16 (av1gpu_kernels.h:714-744) */
__device__ __forceinline__ void adst8(int (&t)[8], int lo, int hi) {
  adst_in(t);
#pragma unroll
  for (int i = 0; i < 4; ++i) bf(t, 2 * i, 2 * i + 1, 60 - 16 * i, 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) hd(t, i, 4 + i, 0, lo, hi);
#pragma unroll
  for (int i = 0; i < 2; ++i) bf(t, 4 + 3 * i, 5 + i, 48 - 32 * i, 1);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) hd(t, 4 * j + i, 2 + 4 * j + i, 0, lo, hi);
#pragma unroll
  for (int i = 0; i < 2; ++i) bf(t, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(t);
}
__device__ __forceinline__ void adst16(int (&t)[16], int lo, int hi) {
  adst_in(t);
#pragma unroll
  for (int i = 0; i < 8; ++i) bf(t, 2 * i, 2 * i + 1, 62 - 8 * i, 1);
#pragma unroll
  for (int i = 0; i < 8; ++i) hd(t, i, 8 + i, 0, lo, hi);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    bf(t, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1);
    bf(t, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) hd(t, 8 * j + i, 4 + 8 * j + i, 0, lo, hi);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) bf(t, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) hd(t, 4 * j + i, 2 + 4 * j + i, 0, lo, hi);
#pragma unroll
  for (int i = 0; i < 4; ++i) bf(t, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(t);
}

/* identity (av1gpu_kernels.h:759-770) */
template <int N>
__device__ __forceinline__ void iidt(int (&t)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (N == 4)
      t[i] = r12((int64_t)t[i] * 5793);
    else if (N == 8)
      t[i] = t[i] * 2;
    else if (N == 16)
      t[i] = r12((int64_t)t[i] * 11586);
    else if (i < 32)
      t[i] = t[i] * 4;
  }
}

/* Walsh-Hadamard (av1gpu_kernels.h:773-790) */
__device__ __forceinline__ void iwht(int (&t)[4], int shift) {
  int a = t[0] >> shift, c = t[1] >> shift, d = t[2] >> shift, b = t[3] >> shift;
  a += c;
  d -= b;
  const int e = (a - d) >> 1;
  b = e - b;
  c = e - c;
  a -= b;
  d += c;
  t[0] = a;
  t[1] = b;
  t[2] = c;
  t[3] = d;
}

/* kind: 0 DCT, 1 ADST, 2 identity (kind tables av1gpu_kernels.h:800-840) */
template <int N>
__device__ __forceinline__ void tx1d(int (&t)[N], int kind, int lo, int hi) {
  if (kind == 0) {
    idct<N>(t, lo, hi);
  } else if (kind == 1) {
    if constexpr (N == 4) adst4(t);
    if constexpr (N == 8) adst8(t, lo, hi);
    if constexpr (N == 16) adst16(t, lo, hi);
  } else {
    iidt<N>(t);
  }
}

/* two bits per Tx_Type, type 0 in the low bits */
__device__ __forceinline__ int row_kind(int tx) { return (0x66295450u >> (2 * tx)) & 3; }
__device__ __forceinline__ int col_kind(int tx) { return (0x99895144u >> (2 * tx)) & 3; }
__device__ __forceinline__ int flip_ud(int tx) { return tx == 4 || tx == 8 || tx == 6 || tx == 14; }
__device__ __forceinline__ int flip_lr(int tx) { return tx == 5 || tx == 7 || tx == 6 || tx == 15; }

}  // namespace itx
}  // namespace av1hip

#endif
