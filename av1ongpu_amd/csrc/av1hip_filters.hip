/*
 * Post-reconstruction passes for gfx950: deblocking (two ordered passes),
 * CDEF, superres upscaling, loop restoration, and the NV12 / RGB packs. The
 * arithmetic is the reference's (kernels/av1gpu_kernels.h), cited per function.
 */
#define AV1T_TABLE_QUAL static __constant__ const
#include "av1hip_internal.hpp"
#include "spec_tables.h"

#include <climits>

namespace av1hip {
namespace {

#define SPEC(off) (av1t_spec_table + (off))

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int rnd2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ int rsr64(int64_t v, int n) { return (int)((v + ((int64_t)1 << (n - 1))) >> n); }

/* ---- deblocking (av1gpu_kernels.h:2422-2572) --------------------------- */

__device__ __forceinline__ int f4c(int v, int bd) { return clampi(v, -(1 << (bd - 1)), (1 << (bd - 1)) - 1); }

/* The flat filters: outputs p(N-1)..q(N-1) from taps p(N)..q(N), the centre
 * N2 taps weighted twice, all indices fixed at compile time
 * (av1gpu_kernels.h:2540-2568). v[j] holds the sample at b[(j - 5) s]. */
template <int N, int N2, int LG>
__device__ __forceinline__ void wide_filter(const int* v, uint16_t* b, int s, int maxv) {
  int res[2 * N];
#pragma unroll
  for (int i = -N; i < N; ++i) {
    int t = 0;
#pragma unroll
    for (int k = -N; k <= N; ++k) {
      const int idx = i + k < -(N + 1) ? -(N + 1) : (i + k > N ? N : i + k);
      t += v[idx + 7] * ((k <= N2 && k >= -N2) ? 2 : 1);
    }
    res[i + N] = rnd2(t, LG);
  }
#pragma unroll
  for (int i = -N; i < N; ++i) b[(2 + i) * s] = (uint16_t)clampi(res[i + N], 0, maxv);
}

/* one line across one edge: b points at p1, s is the step across the edge.
 * Every sample the line's filter size can use is loaded first, together, so
 * the decisions do not wait on one load after another. */
__device__ void deblock_line(uint16_t* b, int s, int level, int sharp, int fs, int chroma, int bd,
                             int maxv) {
  const int sh = bd - 8, mid = 0x80 << sh;
  const int flen = fs == 4 ? 4 : (chroma ? 6 : (fs == 8 ? 8 : 16));
  /* v[j] = b[(j - 5) s]: p6 .. p1 at 0 .. 5, p0 6, q0 7, q1 .. q6 at 8 .. 13 */
  const int jlo = fs >= 16 ? 0 : (flen >= 8 ? 3 : (flen >= 6 ? 4 : 5));
  const int jhi = fs >= 16 ? 13 : (flen >= 8 ? 10 : (flen >= 6 ? 9 : 8));
  int v[14];
#pragma unroll
  for (int j = 0; j < 14; ++j) v[j] = (j >= jlo && j <= jhi) ? (int)b[(j - 5) * s] : 0;
  const int p1 = v[5], p0 = v[6], q0 = v[7], q1 = v[8];
  const int lsh = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
  int limit = level >> lsh;
  if (limit < 1) limit = 1;
  if (sharp > 0 && limit > 9 - sharp) limit = 9 - sharp;
  const int blim = (2 * (level + 2) + limit) << sh, lim = limit << sh, thr = (level >> 4) << sh;
  const int one = 1 << sh;
  const int hev = abs(p1 - p0) > thr || abs(q1 - q0) > thr;
  int bad = abs(p1 - p0) > lim || abs(q1 - q0) > lim || abs(p0 - q0) * 2 + abs(p1 - q1) / 2 > blim;
  const int p2 = v[4], q2 = v[9], p3 = v[3], q3 = v[10];
  if (flen >= 6) bad = bad || abs(p2 - p1) > lim || abs(q2 - q1) > lim;
  if (flen >= 8) bad = bad || abs(p3 - p2) > lim || abs(q3 - q2) > lim;
  if (bad) return;
  int flat = 1, flat2 = 1;
  if (fs >= 8) {
    int f = abs(p1 - p0) > one || abs(q1 - q0) > one || abs(p2 - p0) > one || abs(q2 - q0) > one;
    if (flen >= 8) f = f || abs(p3 - p0) > one || abs(q3 - q0) > one;
    flat = !f;
  }
  if (fs >= 16) {
    flat2 = !(abs(v[0] - p0) > one || abs(v[13] - q0) > one || abs(v[1] - p0) > one || abs(v[12] - q0) > one ||
              abs(v[2] - p0) > one || abs(v[11] - q0) > one);
  }
  if (fs == 4 || !flat) {
    const int ps1 = p1 - mid, ps0 = p0 - mid, qs0 = q0 - mid, qs1 = q1 - mid;
    int f = hev ? f4c(ps1 - qs1, bd) : 0;
    f = f4c(f + 3 * (qs0 - ps0), bd);
    const int f1 = f4c(f + 4, bd) >> 3, f2 = f4c(f + 3, bd) >> 3;
    b[2 * s] = (uint16_t)clampi(f4c(qs0 - f1, bd) + mid, 0, maxv);
    b[s] = (uint16_t)clampi(f4c(ps0 + f2, bd) + mid, 0, maxv);
    if (!hev) {
      const int f3 = rnd2(f1, 1);
      b[3 * s] = (uint16_t)clampi(f4c(qs1 - f3, bd) + mid, 0, maxv);
      b[0] = (uint16_t)clampi(f4c(ps1 + f3, bd) + mid, 0, maxv);
    }
    return;
  }
  if (fs >= 16 && flat2)
    wide_filter<6, 1, 4>(v, b, s, maxv);
  else if (chroma)
    wide_filter<2, 1, 3>(v, b, s, maxv);
  else
    wide_filter<3, 0, 3>(v, b, s, maxv);
}

/* One thread per line; edges of one direction never share samples. The edge
 * owning line g is found from the first edge of the thread's 256 line chunk
 * (a table the host builds) and a short forward scan. */
__device__ void deblock_chunk(FArgs a, int first, int count, int lines, const int32_t* chunk,
                              int ci, int* s_line0) {
  /* the chunk's lines lie in the 256 edges from its first one (an edge
   * covers at least one line): their first lines go to LDS, and every
   * thread finds its edge by binary search there */
  const DevDeblock* edges = a.plan.deblock + first;
  const int e0 = chunk[ci], ne = min(count - e0, 256);
  if ((int)threadIdx.x < ne) s_line0[threadIdx.x] = edges[e0 + threadIdx.x].line0;
  __syncthreads();
  const int g = ci * 256 + threadIdx.x;
  if (g >= lines) return;
  int lo = 0, hi = ne - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (s_line0[m] <= g)
      lo = m;
    else
      hi = m - 1;
  }
  const DevDeblock& e = edges[e0 + lo];
  if (e.level == 0) return;
  const int k = g - e.line0;
  const int p = e.plane_vertical & 3, st = a.lay.stride[p];
  uint16_t* plane = a.frame + a.lay.plane_off[p];
  uint16_t* b;
  int s;
  if (e.plane_vertical & 4) {
    b = plane + (size_t)(e.y + k) * st + (e.x - 2);
    s = 1;
  } else {
    b = plane + (size_t)(e.y - 2) * st + (e.x + k);
    s = st;
  }
  deblock_line(b, s, e.level, a.sharpness, e.filter_size, p != 0, a.lay.bit_depth, a.maxv);
}

/* the vertical edges (pass 0) or the horizontal edges (pass 1) of every frame */
__global__ __launch_bounds__(256) void k_deblock(BatchDev B, int pass) {
  __shared__ int s_line0[256];
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, (pass ? kKDeblockH : kKDeblockV));
  const int total = pre[B.njobs];
  for (int base = 0; base < total; base += gridDim.x) {
    const int v = xcd_group(base, total);
    if (v < 0) continue;
    const int j = find_job(pre, B.njobs, v);
    FArgs a = filter_args(B, j);
    PlanRef P = a.plan;
    if (pass)
      deblock_chunk(a, P.n_deblock_v, P.n_deblock_h, P.lines_h, P.db_chunk + (P.lines_v + 255) / 256, v - pre[j],
                    s_line0);
    else
      deblock_chunk(a, 0, P.n_deblock_v, P.lines_v, P.db_chunk, v - pre[j], s_line0);
    __syncthreads();
  }
}

/* ---- CDEF (av1gpu_kernels.h:2601-2775) ---------------------------------- */

__device__ __forceinline__ int flog2(int v) { return 31 - __builtin_clz(v); }

__device__ __forceinline__ int constrain(int diff, int thr, int damping) {
  if (!thr) return 0;
  const int adj = max(0, damping - flog2(thr));
  const int mag = abs(diff);
  const int v = clampi(thr - (mag >> adj), 0, mag);
  return diff < 0 ? -v : v;
}

/* A sample at byte offset o of a picture whose base is the same for the
 * whole wave: one scalar base plus a 32-bit vector offset, no per-lane
 * 64-bit address arithmetic (a picture buffer is far below 4 GB) */
__device__ __forceinline__ int tap16(const char* base, unsigned o) {
  return *reinterpret_cast<const uint16_t*>(base + o);
}

/* constrain() of tap v around the centre x, the shift of its threshold
 * precomputed: |v - x| is one sum-of-absolute-differences instruction
 * (samples are at most 16 bit, the upper halves zero) */
__device__ __forceinline__ int constrain_vx(int v, int x, int thr, int adj) {
  const int mag = (int)__builtin_amdgcn_sad_u16((unsigned)v, (unsigned)x, 0u);
  const int c = max(0, min(thr - (mag >> adj), mag));
  return v < x ? -c : c;
}

__constant__ const int kCdefDirs[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}},
                                             {{0, 1}, {1, 2}},   {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},
                                             {{1, 0}, {2, 0}},   {{1, 0}, {2, -1}}};
__constant__ const int kCdefUvDir[2][2][8] = {{{0, 1, 2, 3, 4, 5, 6, 7}, {1, 2, 2, 2, 3, 4, 6, 0}},
                                              {{7, 0, 2, 4, 5, 6, 6, 6}, {0, 1, 2, 3, 4, 5, 6, 7}}};

/* Direction search of luma 8x8 blocks (av1gpu_kernels.h:2601-2660), one
 * block per lane: a lane reads its block's eight rows (one 16-byte load
 * each: rows start on 8-sample boundaries of 128-sample aligned strides),
 * keeps the eight directions' line sums in registers (every index a
 * compile-time constant), weighs them and picks the first direction of
 * largest cost. No LDS and no cross-lane traffic: consecutive lanes take
 * consecutive listed blocks, so a wave's row loads cover whole 128-byte
 * spans of a block row. Writes (direction, variance) for each block into
 * `dirmap`. */
__device__ __forceinline__ void cdef_dir_block(FArgs a, int bi) {
  const int pos = a.plan.cdef_dir_list[bi];
  const int bx = pos & 0xffff, by = pos >> 16;
  const int sh = a.lay.bit_depth - 8;
  const uint16_t* src = a.cdef_src + a.lay.plane_off[0] + (size_t)(by * 8) * a.lay.stride[0] + bx * 8;
  int p0[15] = {}, p1[11] = {}, p2[8] = {}, p3[11] = {}, p4[15] = {}, p5[11] = {}, p6[8] = {}, p7[11] = {};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint4 r = *reinterpret_cast<const uint4*>(src + (size_t)i * a.lay.stride[0]);
    const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int x = (int)(((w[k >> 1] >> (16 * (k & 1))) & 0xffffu) >> sh) - 128;
      p0[i + k] += x;
      p1[i + k / 2] += x;
      p2[i] += x;
      p3[3 + i - k / 2] += x;
      p4[7 + i - k] += x;
      p5[3 - i / 2 + k] += x;
      p6[k] += x;
      p7[i / 2 + k] += x;
    }
  }
  /* 840 / n for a line of n samples */
  constexpr int div[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};
  auto full = [&](const int* q) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) c += q[t] * q[t];
    return c * div[8];
  };
  auto diag = [&](const int* q) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < 7; ++t) c += (q[t] * q[t] + q[14 - t] * q[14 - t]) * div[t + 1];
    return c + q[7] * q[7] * div[8];
  };
  auto half = [&](const int* q) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) c += q[3 + t] * q[3 + t];
    c *= div[8];
#pragma unroll
    for (int t = 0; t < 3; ++t) c += (q[t] * q[t] + q[10 - t] * q[10 - t]) * div[2 * t + 2];
    return c;
  };
  const int cost[8] = {diag(p0), half(p1), full(p2), half(p3), diag(p4), half(p5), full(p6), half(p7)};
  /* the first direction with the largest cost (a strict > scan from 0); a
   * block whose costs are all zero keeps direction 0 */
  int best = 0, dd = 0;
#pragma unroll
  for (int d = 0; d < 8; ++d)
    if (cost[d] > best) {
      best = cost[d];
      dd = d;
    }
  int opp = cost[0];
#pragma unroll
  for (int d = 1; d < 8; ++d)
    if (d == ((dd + 4) & 7)) opp = cost[d];
  int* m = a.cdef_dir + 2 * (by * a.cdef_dir_w + bx);
  m[0] = dd;
  m[1] = (best - opp) >> 10;
}

__global__ __launch_bounds__(256) void k_cdef_dir(BatchDev B) {
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKCdefDir);
  const int total = pre[B.njobs];
  for (int base = 0; base < total; base += gridDim.x) {
    const int v = xcd_group(base, total);
    if (v < 0) continue;
    const int j = find_job(pre, B.njobs, v);
    FArgs a = filter_args(B, j);
    const int bi = (v - pre[j]) * kCdefDirPerGroup + (int)threadIdx.x;
    if (bi < a.plan.n_cdef_dir) cdef_dir_block(a, bi);
  }
}

/* CDEF of the listed blocks (av1gpu_kernels.h:2662-2775). A wave takes
 * 64 / (W x H) blocks of one plane kind (W x H samples each: 8x8 luma, the
 * chroma size by subsampling) and first stages each block with a 2-sample
 * border from the deblocked snapshot into LDS, one 32-bit load per sample
 * pair (rows start on even samples of 128-sample aligned strides), marking
 * samples outside the plane; then one lane per sample reads its twelve taps
 * from LDS. Samples of unlisted blocks are already final in the frame. */
constexpr uint16_t kCdefOutside = 0xffff; /* no sample reaches this value */
constexpr int kCdefTileWave = 256;         /* LDS samples per wave         */

/* the tap of direction table entry e = (dy, dx) as an offset in a tile of
 * row pitch tw */
__device__ __forceinline__ int cdef_tap(const int* e, int tw) { return e[0] * tw + e[1]; }

/* A block's CDEF record and, when its direction is searched, its entry of
 * the direction map: loaded before its tile is staged, so that the two
 * loads' latencies overlap. A luma wave holds one block: its values are
 * wave-uniform and come through the scalar cache (the map was written by an
 * earlier kernel); chroma waves hold several blocks. */
struct CdefBlockIn {
  int plane, x, y, pri, sec, damping, dir, yd, var;
};

template <bool kLuma>
__device__ __forceinline__ CdefBlockIn cdef_block_in(FArgs a, int first, int bi) {
  CdefBlockIn q;
  if (kLuma) {
    bi = __builtin_amdgcn_readfirstlane(bi);
    const AV1H_CMEM DevCdef& cu = cmem(a.plan.cdef)[first + bi];
    q = CdefBlockIn{cu.plane_damping & 3, cu.x, cu.y, cu.pri, cu.sec, cu.plane_damping >> 2, cu.dir, 0, 0};
    if (q.dir < 0) {
      const AV1H_CMEM int* m = cmem(a.cdef_dir) + 2 * ((q.y >> 3) * a.cdef_dir_w + (q.x >> 3));
      q.yd = m[0];
      q.var = m[1];
    }
  } else {
    const DevCdef c = a.plan.cdef[first + bi];
    q = CdefBlockIn{c.plane_damping & 3, c.x, c.y, c.pri, c.sec, c.plane_damping >> 2, c.dir, 0, 0};
    if (q.dir < 0) {
      LayRef L = a.lay;
      const int lx = c.x << L.ss_x, ly = c.y << L.ss_y;
      q.yd = a.cdef_dir[2 * ((ly >> 3) * a.cdef_dir_w + (lx >> 3))];
    }
  }
  return q;
}

/* One sample (r, k) of block `c`, its tile at `t` (row pitch TW, the sample
 * at t[(r + 2) * TW + k + 2]). False when it keeps the snapshot's value.
 * Luma reads the direction table from constant memory (its direction is
 * wave-uniform), chroma from LDS. */
template <int W, int H, bool kLuma>
__device__ __forceinline__ bool cdef_tile_sample(FArgs a, const CdefBlockIn& c, int r, int k, const uint16_t* t,
                                                 const int (*sdirs)[2][2], int64_t* off, int* val) {
  constexpr int TW = W + 4;
  const int (*dirs)[2][2] = kLuma ? kCdefDirs : sdirs;
  LayRef L = a.lay;
  const int p = kLuma ? 0 : c.plane;
  const int cs = L.bit_depth - 8;
  const int pw = L.grid_w[p], ph = L.grid_h[p];
  if (c.x + k >= pw || c.y + r >= ph) return false;
  int pri = c.pri << cs;
  const int sec = c.sec << cs;
  const int damping = c.damping + cs - (kLuma ? 0 : 1);
  int dir = c.dir;
  if (dir < 0) {
    if (!kLuma) {
      dir = pri == 0 ? 0 : kCdefUvDir[L.ss_x][L.ss_y][c.yd];
    } else {
      const int vs = (c.var >> 6) ? min(flog2(c.var >> 6), 12) : 0;
      dir = pri == 0 ? 0 : c.yd;
      pri = c.var ? ((pri * (4 + vs) + 8) >> 4) : 0;
    }
  }
  if (pri == 0 && sec == 0) return false; /* the destination already holds the snapshot's value */
  const int adjp = pri ? max(0, damping - flog2(pri)) : 0, adjs = sec ? max(0, damping - flog2(sec)) : 0;
  const int ts = (pri >> cs) & 1;
  const int pt0 = ts ? 3 : 4, pt1 = ts ? 3 : 2;
  const int d1 = (dir + 2) & 7, d2 = (dir + 6) & 7;
  const uint16_t* s0 = t + (r + 2) * TW + k + 2;
  const int x = s0[0];
  int sum = 0, lo = x, hi = x;
  /* every tap inside the plane: no marks to test */
  const bool inner = c.x >= 2 && c.y >= 2 && c.x + W + 2 <= pw && c.y + H + 2 <= ph;
#pragma unroll
  for (int tt = 0; tt < 2; ++tt) {
    const int op = cdef_tap(dirs[dir][tt], TW), o1 = cdef_tap(dirs[d1][tt], TW), o2 = cdef_tap(dirs[d2][tt], TW);
    const int wp = tt ? pt1 : pt0, ws = tt ? 1 : 2;
#pragma unroll
    for (int sg = -1; sg <= 1; sg += 2) {
      const int vp = s0[sg * op], v1 = s0[sg * o1], v2 = s0[sg * o2];
      if (inner) {
        if (pri) sum += __mul24(wp, constrain_vx(vp, x, pri, adjp));
        if (sec) sum += __mul24(ws, constrain_vx(v1, x, sec, adjs) + constrain_vx(v2, x, sec, adjs));
        lo = min(lo, min(vp, min(v1, v2)));
        hi = max(hi, max(vp, max(v1, v2)));
      } else {
        if (vp != kCdefOutside) {
          if (pri) sum += __mul24(wp, constrain_vx(vp, x, pri, adjp));
          lo = min(lo, vp);
          hi = max(hi, vp);
        }
        if (v1 != kCdefOutside) {
          if (sec) sum += __mul24(ws, constrain_vx(v1, x, sec, adjs));
          lo = min(lo, v1);
          hi = max(hi, v1);
        }
        if (v2 != kCdefOutside) {
          if (sec) sum += __mul24(ws, constrain_vx(v2, x, sec, adjs));
          lo = min(lo, v2);
          hi = max(hi, v2);
        }
      }
    }
  }
  *off = L.plane_off[p] + (int64_t)(c.y + r) * L.stride[p] + c.x + k;
  *val = clampi(clampi(x + ((8 + sum - (sum < 0)) >> 4), lo, hi), 0, a.maxv);
  return true;
}

/* `batch` (1 to kSets) sets of 256 / (W x H) blocks per workgroup: every
 * lane loads its block record (and direction) of each set, the wave stages
 * all its sets' tiles, and only then filters them, so that a wave has the
 * loads of all its sets in flight at once. kSets bounds the registers the
 * sets take (one set keeps the kernel at 8 waves per SIMD). */
template <int W, int H, bool kLuma, int kSets>
__device__ __forceinline__ void cdef_tiles(FArgs a, int first, int count, int grp, int batch, uint16_t* tile,
                                           const int (*dirs)[2][2]) {
  constexpr int S = W * H, TW = W + 4, TH = H + 4, kWave = 64 / S, kPer = 256 / S;
  constexpr int kPairs = TW / 2, kDw = kPairs * TH; /* 32-bit words per block tile */
  static_assert(kWave * TW * TH <= kCdefTileWave, "tiles of a wave exceed its LDS");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint16_t* wt = tile + wave * kCdefTileWave * kSets;
  LayRef L = a.lay;
  const int sb = lane / S, sm = lane - sb * S, r = sm / W, k = sm - r * W;
  CdefBlockIn q[kSets];
  bool live[kSets];
#pragma unroll
  for (int u = 0; u < kSets; ++u) {
    const int b0 = (grp * batch + u) * kPer + wave * kWave;
    live[u] = u < batch && b0 + sb < count;
    q[u] = CdefBlockIn{};
    if (live[u]) q[u] = cdef_block_in<kLuma>(a, first, b0 + sb);
  }
#pragma unroll
  for (int u = 0; u < kSets; ++u) {
    if (u >= batch) break;
    const int b0 = (grp * batch + u) * kPer + wave * kWave;
    uint16_t* ut = wt + u * kCdefTileWave;
    for (int e = lane; e < kWave * kDw; e += 64) {
      const int eb = e / kDw, rem = e - eb * kDw, row = rem / kPairs, cp = rem - row * kPairs;
      uint32_t v = 0xffffffffu;
      if (b0 + eb < count) {
        /* a luma wave stages its own block; a chroma lane may stage another */
        int p = 0, cx = q[u].x, cy = q[u].y;
        if (!kLuma) {
          const DevCdef& c = a.plan.cdef[first + b0 + eb];
          p = c.plane_damping & 3;
          cx = c.x;
          cy = c.y;
        }
        const int px = cx - 2 + 2 * cp, py = cy - 2 + row;
        /* grid widths are even: a pair is inside or outside as a whole */
        if (px >= 0 && px < L.grid_w[p] && py >= 0 && py < L.grid_h[p])
          v = *reinterpret_cast<const uint32_t*>(a.cdef_src + L.plane_off[p] + (int64_t)py * L.stride[p] + px);
      }
      *reinterpret_cast<uint32_t*>(ut + eb * TW * TH + row * TW + 2 * cp) = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kSets; ++u) {
    int64_t off = 0;
    int val = 0;
    if (live[u] &&
        cdef_tile_sample<W, H, kLuma>(a, q[u], r, k, wt + u * kCdefTileWave + sb * TW * TH, dirs, &off, &val)) {
      a.cdef_dst[off] = (uint16_t)val;
      if (a.cdef_dst2) a.cdef_dst2[off] = (uint16_t)val;
    }
  }
  __syncthreads();
}

/* The luma blocks (plane 0) and then the chroma blocks (planes 1 and 2) of
 * every frame in one launch: the luma groups of every frame come first, then
 * the chroma groups (both read the deblocked snapshot and the direction
 * map, and write different planes, so nothing orders them). `batch` sets
 * per workgroup, at most kSets (a kernel per bound: the registers of more
 * sets would lower the occupancy of one). */
template <int kSets>
__global__ __launch_bounds__(256) void k_cdef(BatchDev B, int batch) {
  /* the direction table in LDS: lanes index it by their own direction */
  __shared__ int s_dirs[8][2][2];
  __shared__ __align__(16) uint16_t s_tile[4 * kCdefTileWave * kSets];
  if (threadIdx.x < 32) (&s_dirs[0][0][0])[threadIdx.x] = (&kCdefDirs[0][0][0])[threadIdx.x];
  __syncthreads();
  const AV1H_CMEM int32_t* pre_l = kernel_prefix(B, kKCdefLuma);
  const AV1H_CMEM int32_t* pre_c = kernel_prefix(B, kKCdefChroma);
  const int tl = pre_l[B.njobs], total = tl + pre_c[B.njobs];
  for (int base = 0; base < total; base += gridDim.x) {
    const int v = xcd_group(base, total);
    if (v < 0) continue;
    if (v < tl) {
      const int j = find_job(pre_l, B.njobs, v);
      FArgs a = filter_args(B, j);
      cdef_tiles<8, 8, true, kSets>(a, 0, a.plan.n_cdef_luma, v - pre_l[j], batch, s_tile, s_dirs);
      continue;
    }
    const int vc = v - tl;
    const int j = find_job(pre_c, B.njobs, vc);
    FArgs a = filter_args(B, j);
    const int nl = a.plan.n_cdef_luma, nc = a.plan.n_cdef - nl, grp = vc - pre_c[j];
    switch (a.lay.ss_x * 2 + a.lay.ss_y) {
      case 3: cdef_tiles<4, 4, false, kSets>(a, nl, nc, grp, batch, s_tile, s_dirs); break;
      case 2: cdef_tiles<4, 8, false, kSets>(a, nl, nc, grp, batch, s_tile, s_dirs); break;
      case 1: cdef_tiles<8, 4, false, kSets>(a, nl, nc, grp, batch, s_tile, s_dirs); break;
      default: cdef_tiles<8, 8, false, kSets>(a, nl, nc, grp, batch, s_tile, s_dirs); break;
    }
  }
}

/* ---- superres (av1gpu_kernels.h:2387-2402) ------------------------------ */

/* 256 x 4 output samples of one plane of one picture per workgroup: the
 * source span of the 256 columns (at most 264 samples when upscaling) is
 * staged in LDS per row, clamped to the coded width as each tap is
 * (av1gpu_kernels.h:2387-2402). Picture 0 is the frame (from the CDEF
 * output in `scratch`; also stored to lr_cdef when loop restoration reads
 * it), picture 1 the deblocked picture loop restoration reads (lr_curr into
 * lr_prev_up). */
constexpr int kUpW = 256, kUpRows = 8, kUpSpan = 272;

/* output tiles of a plane of ow x oh display samples */
__host__ __device__ __forceinline__ int up_tiles(int ow, int oh) {
  return (ow + kUpW - 1) / kUpW * ((oh + kUpRows - 1) / kUpRows);
}

__device__ void upscale_tile(FArgs a, int t, const int* s_k, uint16_t (*s_row)[kUpSpan]) {
  LayRef L = a.lay;
  const int t0 = up_tiles(L.out_w[0], L.out_h[0]), t1 = up_tiles(L.out_w[1], L.out_h[1]);
  const int tp = t0 + t1 + up_tiles(L.out_w[2], L.out_h[2]);
  const int pic = t >= tp;
  int rem = t - pic * tp, p = 0;
  if (rem >= t0) {
    rem -= t0;
    p = 1;
    if (rem >= t1) {
      rem -= t1;
      p = 2;
    }
  }
  const int ow = L.out_w[p], oh = L.out_h[p];
  const int tx = (ow + kUpW - 1) / kUpW;
  const int by = rem / tx, bx = rem - by * tx;
  const uint16_t* src = pic ? a.lr_curr : a.scratch;
  uint16_t* dst = pic ? a.lr_prev_up : a.frame;
  const int x0 = bx * kUpW, y0 = by * kUpRows;
  const int step = a.up_step[p], init = a.up_initial[p];
  const int maxx = L.plane_w[p] - 1;
  const int stride = L.stride[p];
  const int64_t poff = L.plane_off[p];
  const int xl = min(x0 + kUpW, ow) - 1;
  const int ws = ((-(1 << 14) + init + x0 * step) >> 14) - 3;
  const int wn = ((-(1 << 14) + init + xl * step) >> 14) + 4 - ws + 1;
  const int rows = min(kUpRows, oh - y0);
  {
    /* all of a thread's loads issued before its LDS writes */
    constexpr int kPer = (kUpSpan + 255) / 256;
    uint16_t val[kUpRows][kPer];
#pragma unroll
    for (int r = 0; r < kUpRows; ++r)
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int c = threadIdx.x + 256 * u;
        val[r][u] = r < rows && c < wn ? src[poff + (size_t)(y0 + r) * stride + clampi(ws + c, 0, maxx)] : 0;
      }
#pragma unroll
    for (int r = 0; r < kUpRows; ++r)
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int c = threadIdx.x + 256 * u;
        if (r < rows && c < wn) s_row[r][c] = val[r][u];
      }
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x > xl) return;
  const int sx = -(1 << 14) + init + x * step;
  const int o = (sx >> 14) - 3 - ws;
  const int* k = s_k + ((sx & ((1 << 14) - 1)) >> 8) * 8;
  uint16_t* dst2 = pic ? nullptr : a.up_dst2;
  for (int r = 0; r < rows; ++r) {
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) sum += s_row[r][o + q] * k[q];
    const uint16_t v = (uint16_t)clampi((sum + 64) >> 7, 0, a.maxv);
    const size_t at = poff + (size_t)(y0 + r) * stride + x;
    dst[at] = v;
    if (dst2) dst2[at] = v;
  }
}

__global__ __launch_bounds__(256) void k_upscale(BatchDev B) {
  __shared__ int s_k[512];
  __shared__ uint16_t s_row[kUpRows][kUpSpan];
  for (int i = threadIdx.x; i < 512; i += 256) s_k[i] = av1t_spec_table[AV1T_OFF_UPSCALE + i];
  __syncthreads();
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKUpscale);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    upscale_tile(filter_args(B, j), v - pre[j], s_k, s_row);
    __syncthreads();
  }
}

/* ---- loop restoration (av1gpu_kernels.h:2832-2982) ---------------------- */

/* Row / column of element i of a w-wide 2D range walked by 256 threads
 * (i = threadIdx.x, then + 256 each step), without a division per step. */
struct Walk256 {
  int r, q, dr, dq, w;
  __device__ __forceinline__ explicit Walk256(int width) : w(width) {
    r = (int)threadIdx.x / w;
    q = (int)threadIdx.x - r * w;
    dr = 256 / w;
    dq = 256 - dr * w;
  }
  __device__ __forceinline__ void next() {
    q += dq;
    r += dr;
    if (q >= w) {
      q -= w;
      ++r;
    }
  }
};

/* a2 of the self-guided filter for z = 0 .. 255 (the specification's
 * x_by_xplus1 mapping, av1gpu_kernels.h:2898): z >= 255 gives 256 */
__device__ __forceinline__ int sgr_a2(int z) { return z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1); }

struct LrCtx {
  const uint16_t* cdef;
  const uint16_t* curr;
  int stride, pw, ph, s0, s1;
  __device__ __forceinline__ int src(int x, int y) const {
    x = clampi(x, 0, pw - 1);
    y = clampi(y, 0, ph - 1);
    if (y < s0) return curr[(size_t)max(y, s0 - 2) * stride + x];
    if (y > s1) return curr[(size_t)min(y, s1 + 2) * stride + x];
    return cdef[(size_t)y * stride + x];
  }
};

/* Loop restoration of one tile (at most kLrTileW x kLrTileH samples of one
 * restoration job, i.e. one unit inside one 64 row stripe) per workgroup:
 * the source window with its 3 sample halo is staged in LDS through the
 * stripe rule (LrCtx::src), then
 *   Wiener - the horizontal 7 tap pass into an LDS intermediate, rounded and
 *            clamped, then the vertical pass (av1gpu_kernels.h:2832-2870);
 *   SGR    - the box sums, A and B of both passes at every position the 3x3
 *            neighbourhoods of the tile touch, into LDS, then the weighted
 *            neighbourhood sums and the projection (av1gpu_kernels.h:2872-2982).
 * Each intermediate is computed once instead of once per output sample. */
constexpr int kLrTileW = kLrTileCols, kLrTileH = kLrTileRows;
constexpr int kLrSrcW = kLrTileW + 6, kLrSrcH = kLrTileH + 6;
constexpr int kLrAbW = kLrTileW + 2, kLrAbH = kLrTileH + 2;

__device__ void lr_tile(FArgs a, int t, int* s_src, int* s_mid, const int* s_a2) {
  const uint16_t* cdef = a.lr_cdef;
  const uint16_t* curr = a.upscaled ? a.lr_prev_up : a.lr_curr;
  const int2 tile = a.plan.lr_tiles[t];
  const DevLr& j = a.plan.lr[tile.x];
  const int x0 = j.x + (tile.y & 0xffff), y0 = j.y + (tile.y >> 16);
  const int tw = min(kLrTileW, j.x + j.w - x0), th = min(kLrTileH, j.y + j.h - y0);
  LayRef L = a.lay;
  const int p = j.plane, bd = L.bit_depth;
  LrCtx c;
  c.cdef = cdef + L.plane_off[p];
  c.curr = curr + L.plane_off[p];
  c.stride = L.stride[p];
  c.pw = L.out_w[p];
  c.ph = L.out_h[p];
  c.s0 = j.stripe0;
  c.s1 = j.stripe1;
  {
    /* all of a thread's loads issued before its LDS writes */
    constexpr int kStage = (kLrSrcW * kLrSrcH + 255) / 256;
    int val[kStage], at[kStage];
    Walk256 it(tw + 6);
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      at[u] = it.r < th + 6 ? it.r * kLrSrcW + it.q : -1;
      val[u] = at[u] >= 0 ? c.src(x0 - 3 + it.q, y0 - 3 + it.r) : 0;
      it.next();
    }
#pragma unroll
    for (int u = 0; u < kStage; ++u)
      if (at[u] >= 0) s_src[at[u]] = val[u];
  }
  __syncthreads();
  uint16_t* dst = a.frame + L.plane_off[p] + (size_t)y0 * c.stride + x0;
  if (j.type == 1) {
    const int r0 = bd == 12 ? 5 : 3, r1 = bd == 12 ? 9 : 11;
    const int off = 1 << (bd + 7 - r0 - 1);
    const int lim = (1 << (bd + 1 + 7 - r0)) - 1;
    int ht[7], vt[7];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      vt[t] = vt[6 - t] = j.vf[t];
      ht[t] = ht[6 - t] = j.hf[t];
    }
    vt[3] = 128 - 2 * (j.vf[0] + j.vf[1] + j.vf[2]);
    ht[3] = 128 - 2 * (j.hf[0] + j.hf[1] + j.hf[2]);
    for (Walk256 it(tw); it.r < th + 6; it.next()) {
      const int r = it.r, q = it.q;
      const int* row = s_src + r * kLrSrcW + q;
      int sum = 0;
#pragma unroll
      for (int k = 0; k < 7; ++k) sum += ht[k] * row[k];
      s_mid[r * kLrTileW + q] = clampi(rsr64(sum, r0), -off, lim - off);
    }
    __syncthreads();
    for (Walk256 it(tw); it.r < th; it.next()) {
      const int r = it.r, q = it.q;
      int sum = 0;
#pragma unroll
      for (int t = 0; t < 7; ++t) sum += vt[t] * s_mid[(r + t) * kLrTileW + q];
      dst[(size_t)r * c.stride + q] = (uint16_t)clampi(rsr64(sum, r1), 0, a.maxv);
    }
    return;
  }
  /* self guided: A and B of each pass on the (th + 2) x (tw + 2) positions
   * around the tile */
  for (int pass = 0; pass < 2; ++pass) {
    const int rr = j.r[pass];
    if (rr == 0) continue;
    const int d = 2 * rr + 1, n = d * d;
    const int64_t n2e = (int64_t)n * n * j.eps[pass];
    const int64_t mt = (((int64_t)1 << 20) + n2e / 2) / n2e;
    const int inv = ((1 << 12) + n / 2) / n;
    int* A = s_mid + (2 * pass) * kLrAbH * kLrAbW;
    int* B = A + kLrAbH * kLrAbW;
    for (Walk256 it(tw + 2); it.r < th + 2; it.next()) {
      const int r = it.r, q = it.q;
      /* position (x0 - 1 + q, y0 - 1 + r); its window starts at src row
       * r + 2 - rr, column q + 2 - rr of the staged window */
      /* at most 25 samples of 12 bits: the sums fit 32 bits */
      int sa32 = 0, sb32 = 0;
      for (int dy = 0; dy < d; ++dy) {
        const int* row = s_src + (r + 2 - rr + dy) * kLrSrcW + q + 2 - rr;
        for (int dx = 0; dx < d; ++dx) {
          const int v = row[dx];
          sa32 += v * v;
          sb32 += v;
        }
      }
      int64_t sa = sa32, sb = sb32;
      const int64_t braw = sb;
      if (bd > 8) {
        sa = rsr64(sa, 2 * (bd - 8));
        sb = rsr64(sb, bd - 8);
      }
      int64_t pp = sa * n - sb * sb;
      if (pp < 0) pp = 0;
      const int z = rsr64(pp * mt, 20);
      const int a2 = z >= 255 ? 256 : s_a2[z];
      A[r * kLrAbW + q] = a2;
      B[r * kLrAbW + q] = rsr64((int64_t)(256 - a2) * braw * inv, 12);
    }
  }
  __syncthreads();
  const int w0 = j.xqd[0], w1 = j.xqd[1], w2 = (1 << 7) - w0 - w1;
  for (Walk256 it(tw); it.r < th; it.next()) {
    const int r = it.r, q = it.q;
    const int y = y0 + r;
    const int u = s_src[(r + 3) * kLrSrcW + q + 3];
    int flt[2] = {0, 0};
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      if (j.r[pass] == 0) continue;
      const int* A = s_mid + (2 * pass) * kLrAbH * kLrAbW;
      const int* B = A + kLrAbH * kLrAbW;
      int sa = 0, sb = 0;
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int wgt =
              pass == 0 ? (((y + dy) & 1) ? (dx == 0 ? 6 : 5) : 0) : ((dx == 0 || dy == 0) ? 4 : 3);
          if (!wgt) continue;
          const int o = (r + 1 + dy) * kLrAbW + q + 1 + dx;
          sa += wgt * A[o];
          sb += wgt * B[o];
        }
      const int sh = (pass == 0 && (y & 1)) ? 4 : 5;
      flt[pass] = rsr64((int64_t)sa * u + sb, 8 + sh - 4);
    }
    const int us = u << 4;
    int64_t v = (int64_t)w1 * us;
    v += (int64_t)w0 * (j.r[0] ? flt[0] : us);
    v += (int64_t)w2 * (j.r[1] ? flt[1] : us);
    dst[(size_t)r * c.stride + q] = (uint16_t)clampi(rsr64(v, 4 + 7), 0, a.maxv);
  }
}

__global__ __launch_bounds__(256) void k_lr(BatchDev B) {
  __shared__ int s_src[kLrSrcH * kLrSrcW];
  __shared__ int s_mid[4 * kLrAbH * kLrAbW]; /* Wiener rows, or A0 B0 A1 B1 */
  __shared__ int s_a2[256];
  s_a2[threadIdx.x] = sgr_a2(threadIdx.x);
  __syncthreads();
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKLr);
  const int total = pre[B.njobs];
  for (int base = 0; base < total; base += gridDim.x) {
    const int v = xcd_group(base, total);
    if (v < 0) continue;
    const int j = find_job(pre, B.njobs, v);
    lr_tile(filter_args(B, j), v - pre[j], s_src, s_mid, s_a2);
    __syncthreads();
  }
}

/* ---- whole picture copies (the snapshots the filters read, the reference
 * slot store) ---------------------------------------------------------------- */

__device__ __forceinline__ void copy_kind(FArgs a, int k, const uint4** src, uint4** dst) {
  *src = reinterpret_cast<const uint4*>(a.frame);
  switch (k) {
    case kKCopyLrCurr: *dst = reinterpret_cast<uint4*>(a.lr_curr); break;
    case kKCopyCdefSrc:
    case kKCopyUpSrc: *dst = reinterpret_cast<uint4*>(a.scratch); break;
    default: *dst = reinterpret_cast<uint4*>(a.store); break;
  }
}

/* kCopyChunk bytes per workgroup, 16 bytes per thread and step */
__global__ __launch_bounds__(256) void k_copy(BatchDev B, int k) {
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, k);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    FArgs a = filter_args(B, j);
    const uint4* src;
    uint4* dst;
    copy_kind(a, k, &src, &dst);
    /* the deblocked snapshot may have a second destination */
    uint4* dst2 = k == kKCopyLrCurr ? reinterpret_cast<uint4*>(a.snap2) : nullptr;
    const int64_t n = a.lay.size * 2 / 16;
    const int64_t base = (int64_t)(v - pre[j]) * (kCopyChunk / 16);
#pragma unroll
    for (int i = 0; i < kCopyChunk / 16 / 256; ++i) {
      const int64_t e = base + i * 256 + threadIdx.x;
      if (e < n) {
        const uint4 x = src[e];
        dst[e] = x;
        if (dst2) dst2[e] = x;
      }
    }
  }
}

/* ---- NV12 and RGB packs (av1gpu_kernels.h:3379-3472) -------------------- */

/* one thread per output byte pair */
__global__ void k_nv12(Layout L, const uint16_t* src, uint8_t* out, int pitch) {
  const int x2 = (blockIdx.x * blockDim.x + threadIdx.x) * 2;
  const int row = blockIdx.y;
  if (x2 >= pitch) return;
  const int sh = L.bit_depth - 8;
  const int w = L.width, h = L.height;
  uint8_t* d = out + (size_t)row * pitch + x2;
  if (row < h) {
    const uint16_t* s = src + L.plane_off[0] + (size_t)row * L.stride[0];
    d[0] = x2 < w ? (uint8_t)(s[x2] >> sh) : 0;
    if (x2 + 1 < pitch) d[1] = x2 + 1 < w ? (uint8_t)(s[x2 + 1] >> sh) : 0;
    return;
  }
  const int cr = row - h;
  const int cw = (w + (1 << L.ss_x) - 1) >> L.ss_x;
  const int c = x2 >> 1;
  if (c < cw) {
    d[0] = (uint8_t)(src[L.plane_off[1] + (size_t)cr * L.stride[1] + c] >> sh);
    d[1] = (uint8_t)(src[L.plane_off[2] + (size_t)cr * L.stride[2] + c] >> sh);
  } else {
    d[0] = 0;
    if (x2 + 1 < pitch) d[1] = 0;
  }
}

/* NV12 of each frame of a batch that asks for it, after film grain: 1024
 * output bytes of one row per workgroup, 4 per thread */
__device__ void nv12_chunk(FArgs a, int t) {
  LayRef L = a.lay;
  const int pitch = a.nv12_pitch;
  const int cpr = (pitch + 1023) >> 10;
  const int row = t / cpr;
  const int x0 = (t - row * cpr) * 1024 + threadIdx.x * 4;
  if (x0 >= pitch) return;
  const int sh = L.bit_depth - 8, w = L.width, h = L.height;
  uint32_t word = 0;
  if (row < h) {
    const uint16_t* src = a.frame + L.plane_off[0] + (size_t)row * L.stride[0];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (x0 + k < w) word |= (uint32_t)(src[x0 + k] >> sh) << (8 * k);
  } else {
    const int cr = row - h, cw = (w + (1 << L.ss_x) - 1) >> L.ss_x;
    const uint16_t* u = a.frame + L.plane_off[1] + (size_t)cr * L.stride[1];
    const uint16_t* v = a.frame + L.plane_off[2] + (size_t)cr * L.stride[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = (x0 + k) >> 1;
      if (c < cw) word |= (uint32_t)(((x0 + k) & 1 ? v[c] : u[c]) >> sh) << (8 * k);
    }
  }
  uint8_t* d = a.nv12 + (size_t)row * pitch + x0;
  if ((pitch & 3) == 0 && x0 + 4 <= pitch) {
    *reinterpret_cast<uint32_t*>(d) = word;
  } else {
    for (int k = 0; k < 4 && x0 + k < pitch; ++k) d[k] = (uint8_t)(word >> (8 * k));
  }
}

__global__ __launch_bounds__(256) void k_nv12_batch(BatchDev B) {
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKNv12);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    nv12_chunk(filter_args(B, j), v - pre[j]);
  }
}

/* Picture checksums (mix64 weights, see the header): kHashChunk samples per
 * workgroup, 16 per thread, a wave sum added to the frame's accumulator. */
__global__ __launch_bounds__(256) void k_hash(BatchDev B) {
  const AV1H_CMEM int32_t* pre = kernel_prefix(B, kKHash);
  const int total = pre[B.njobs];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int j = find_job(pre, B.njobs, v);
    FArgs a = filter_args(B, j);
    LayRef L = a.lay;
    const int64_t n0 = (int64_t)L.out_w[0] * L.out_h[0], n1 = (int64_t)L.out_w[1] * L.out_h[1];
    const int64_t n = n0 + 2 * n1;
    uint64_t acc = 0;
    const int64_t base = (int64_t)(v - pre[j]) * kHashChunk;
#pragma unroll 4
    for (int k = 0; k < kHashChunk / 256; ++k) {
      const int64_t i = base + k * 256 + threadIdx.x;
      if (i < n) {
        int p = 0;
        int64_t r = i;
        if (r >= n0) {
          r -= n0;
          p = 1;
          if (r >= n1) {
            r -= n1;
            p = 2;
          }
        }
        const int w = L.out_w[p];
        const int y = (int)(r / w), x = (int)(r - (int64_t)y * w);
        acc += (uint64_t)a.frame[L.plane_off[p] + (int64_t)y * L.stride[p] + x] * mix64((uint64_t)i);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)a.hash, (unsigned long long)acc);
  }
}

__global__ void k_zero64(uint64_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

/* The batch table (jobs, kernel prefixes, merged queue) copied from its
 * pinned staging buffer into device memory as the batch's first launch: the
 * byte count is the staging buffer's first word and the table follows at
 * kTableHead, so one captured launch serves every table size, and no copy
 * command of its own sits between two batches on the stream. The host
 * rewrites the staging buffer between the batches that use it: the loads
 * are system scope, past the device's caches. */
__global__ void k_fetch_table(const uint64_t* src, uint64_t* dst) {
  const uint64_t words = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >> 3;
  const uint64_t* from = src + kTableHead / 8;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += 4 * stride) {
    uint64_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v[k] = i + k * stride < words ? __hip_atomic_load(from + i + k * stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
                                    : 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + k * stride < words) dst[i + k * stride] = v[k];
  }
}

/* The batch's results written into its pinned read-back buffer as its last
 * launch, instead of copies queued behind the batch: the error word (then
 * cleared for the next batch on the stream) and, with `hashes`, the n
 * pictures' checksums; system-scope stores, which the host reads once the
 * batch's completion event has fired. */
__global__ void k_readback(uint32_t* err, const uint64_t* hashes, int n, uint32_t* out_err, uint64_t* out_hash) {
  if (threadIdx.x == 0) {
    __hip_atomic_store(out_err, *err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    *err = 0;
  }
  if (hashes)
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      __hip_atomic_store(out_hash + i, hashes[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct ColorM {
  int m[8];
  int o[3];
};

__global__ void k_rgb(Layout L, const uint16_t* src, uint8_t* out, ColorM cm) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= L.width) return;
  const int yy = src[L.plane_off[0] + (size_t)y * L.stride[0] + x];
  const size_t co = (size_t)(y >> L.ss_y) * L.stride[1] + (x >> L.ss_x);
  const int uu = src[L.plane_off[1] + co], vv = src[L.plane_off[2] + co];
  const int r = (cm.m[0] * yy + cm.m[1] * vv + cm.o[0]) >> 16;
  const int g = (cm.m[2] * yy + cm.m[3] * uu + cm.m[4] * vv + cm.o[1]) >> 16;
  const int b = (cm.m[5] * yy + cm.m[6] * uu + cm.o[2]) >> 16;
  uint8_t* d = out + ((size_t)y * L.width + x) * 3;
  d[0] = (uint8_t)clampi(r, 0, 255);
  d[1] = (uint8_t)clampi(g, 0, 255);
  d[2] = (uint8_t)clampi(b, 0, 255);
}

}  // namespace

void superres_params(int dw, int uw, int* step, int* initial) {
  /* av1gpu_kernels.h:2375-2383 */
  const int st = ((dw << 14) + (uw / 2)) / uw;
  const int err = uw * st - (dw << 14);
  const int sub = (-((uw - dw) << 13) + uw / 2) / uw + (1 << 7) - err / 2;
  *step = st;
  *initial = sub & ((1 << 14) - 1);
}

int upscale_tiles(const Layout& L, int pictures) {
  return (up_tiles(L.out_w[0], L.out_h[0]) + up_tiles(L.out_w[1], L.out_h[1]) + up_tiles(L.out_w[2], L.out_h[2])) *
         pictures;
}

hipError_t launch_filter_kernel(const BatchDev& b, int k, int total, int cap, int cdef_batch, hipStream_t s) {
  const dim3 blk(256), grid(min(total, cap));
  switch (k) {
    case kKDeblockV:
    case kKDeblockH: hipLaunchKernelGGL(k_deblock, grid, blk, 0, s, b, k == kKDeblockH ? 1 : 0); break;
    case kKCdefDir: hipLaunchKernelGGL(k_cdef_dir, grid, blk, 0, s, b); break;
    case kKCdefLuma:
      /* luma and chroma in one launch (kKCdefLuma's): `total` counts both */
      if (cdef_batch <= 1)
        hipLaunchKernelGGL(k_cdef<1>, grid, blk, 0, s, b, 1);
      else if (cdef_batch == 2)
        hipLaunchKernelGGL(k_cdef<2>, grid, blk, 0, s, b, 2);
      else
        hipLaunchKernelGGL(k_cdef<kCdefMaxBatch>, grid, blk, 0, s, b, cdef_batch);
      break;
    case kKUpscale: hipLaunchKernelGGL(k_upscale, grid, blk, 0, s, b); break;
    case kKLr: hipLaunchKernelGGL(k_lr, grid, blk, 0, s, b); break;
    case kKNv12: hipLaunchKernelGGL(k_nv12_batch, grid, blk, 0, s, b); break;
    case kKHash: hipLaunchKernelGGL(k_hash, grid, blk, 0, s, b); break;
    case kKCopyLrCurr:
    case kKCopyCdefSrc:
    case kKCopyUpSrc:
    case kKCopyStore: hipLaunchKernelGGL(k_copy, grid, blk, 0, s, b, k); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_zero64(uint64_t* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_fetch_table(const void* pin, void* dev, int frames, hipStream_t s) {
  /* a table the launch fetches is one frame's, a few KB */
  (void)frames;
  hipLaunchKernelGGL(k_fetch_table, dim3(1), dim3(256), 0, s, static_cast<const uint64_t*>(pin),
                     static_cast<uint64_t*>(dev));
  return hipGetLastError();
}

hipError_t launch_readback(uint32_t* err, const uint64_t* hashes, int n, uint32_t* out_err, uint64_t* out_hash,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, s, err, hashes, n, out_err, out_hash);
  return hipGetLastError();
}

hipError_t launch_nv12(const Layout& lay, const uint16_t* src, uint8_t* out, int pitch, hipStream_t s) {
  const int rows = lay.height + ((lay.height + (1 << lay.ss_y) - 1) >> lay.ss_y);
  hipLaunchKernelGGL(k_nv12, dim3((pitch / 2 + 127) / 128, rows), dim3(128), 0, s, lay, src, out, pitch);
  return hipGetLastError();
}

hipError_t launch_rgb(const Layout& lay, const uint16_t* src, uint8_t* out, const int m[8], const int o[3],
                      hipStream_t s) {
  ColorM cm;
  for (int i = 0; i < 8; ++i) cm.m[i] = m[i];
  for (int i = 0; i < 3; ++i) cm.o[i] = o[i];
  hipLaunchKernelGGL(k_rgb, dim3((lay.width + 127) / 128, lay.height), dim3(128), 0, s, lay, src, out, cm);
  return hipGetLastError();
}

}  // namespace av1hip
