// normalize.hip -- local contrast normalization on gfx950.
//
// Reference semantics: src/3D/NormalizeFilter.cu:53-654.
//   E = box_mean(min_filter(I)), D = box_mean(max_filter(I)), both with clamp-to-edge
//   borders and the same radius for the min/max window and the box;
//   norm = I >= min_intensity ? (I - E) / (D - E) : 0.
//
// MI355X design:
//   * Erosion and dilation run together: each voxel is packed into one 32-bit word
//     (v, 65535 - v) so a single v_pk_min_u16 advances both filters (max(v) =
//     65535 - min(65535 - v)).
//   * Each 1-D min/max pass is van Herk / Gil-Werman (3 packed mins per output for any
//     radius) on an LDS tile; tiles are x-contiguous so every HBM access coalesces.
//   * Min/max over a clamped box is separable and exact in any order, so the y and z
//     passes run first (for radius 10 in one z walk, k_minmax_zy_walk; otherwise z then
//     y) and the x pass is fused with the x pass of the box filter (sums of integers,
//     exact in any order, so a sliding sum is used).  The y and z box passes sum FP32
//     values and keep the reference's sequential order (c-r .. c+r); the z pass is fused
//     with the normalization.
//   * For the radii in use, the z box pass (and the separate y / z min/max passes) are
//     column walkers: one thread per line keeps a compile-time-sized register ring, reads
//     every input once and needs no LDS.  Other radii use the LDS-tile kernels.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;
// slices of raw input in flight per thread in the fused y / z min/max walk (a divisor of
// the z window 2 RZ + 1)
#ifndef HDOG_ZY_PF
#define HDOG_ZY_PF 3
#endif
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pkmin(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2, a),
                                                                __builtin_bit_cast(us2, b)));
}

// min of w[3 .. 2R] (the window middle shared by four window minima) in four independent
// chains: consecutive dependent packed 16-bit ops need a wait state between them on
// gfx950, interleaved chains do not (min is exact, so the order is free)
template <int R>
__device__ __forceinline__ uint32_t mid_min(const uint32_t* w) {
  if constexpr (2 * R - 2 < 4) {  // w[3 .. 2R]: a short chain
    uint32_t m = w[3];
#pragma unroll
    for (int t = 4; t <= 2 * R; ++t) m = pkmin(m, w[t]);
    return m;
  } else {
    uint32_t m[4] = {w[3], w[4], w[5], w[6]};
#pragma unroll
    for (int t = 7; t <= 2 * R; ++t) m[(t - 3) & 3] = pkmin(m[(t - 3) & 3], w[t]);
    return pkmin(pkmin(m[0], m[1]), pkmin(m[2], m[3]));
  }
}
__device__ __forceinline__ uint32_t pack(uint32_t v) { return v | ((65535u - v) << 16); }

// van Herk prefix (P) / suffix (in place over A) minima over blocks of w along one LDS
// line whose consecutive elements are `stride` words apart.
__device__ __forceinline__ void vh_block(uint32_t* A, uint32_t* P, int p0, int p1, int stride) {
  uint32_t m = A[p0 * stride];
  P[p0 * stride] = m;
#pragma unroll 4
  for (int p = p0 + 1; p < p1; ++p) {
    m = pkmin(m, A[p * stride]);
    P[p * stride] = m;
  }
  m = A[(p1 - 1) * stride];
#pragma unroll 4
  for (int p = p1 - 2; p >= p0; --p) {
    m = pkmin(m, A[p * stride]);
    A[p * stride] = m;
  }
}

// 1-D packed min/max along y (sa = W) or z (sa = W*H).  Work-group tile: XW consecutive
// x positions times C consecutive outputs along the axis, at one coordinate of the
// remaining axis (stride so).  LDS holds the clamped input span (C + 2r) x XW twice:
// A (input, then suffix minima in place) and P (prefix minima).
template <int XW, bool FROM_U16>
__global__ __launch_bounds__(NT) void k_minmax_line(const void* __restrict__ src,
                                                    uint32_t* __restrict__ dst, int W, int L,
                                                    i64 sa, i64 so, int r, int C, int nxt,
                                                    int nct) {
  extern __shared__ uint32_t smem[];
  const int span = C + 2 * r;
  uint32_t* A = smem;
  uint32_t* P = smem + span * XW;
  unsigned b = blockIdx.x;
  const int xt = b % nxt; b /= nxt;
  const int ct = b % nct; b /= nct;
  const i64 base = (i64)b * so;
  const int x0 = xt * XW, c0 = ct * C;
  stage_tile<8, NT, uint32_t>(
      span * XW,
      [&](int e) -> uint32_t {
        const int p = e / XW, x = x0 + e % XW;
        if (x >= W) return 0u;
        const i64 idx = base + (i64)clampi(c0 - r + p, 0, L - 1) * sa + x;
        return FROM_U16 ? pack(((const uint16_t*)src)[idx]) : ((const uint32_t*)src)[idx];
      },
      [&](int e, uint32_t v) { A[e] = v; });
  __syncthreads();
  const int w = 2 * r + 1;
  const int nb = (span + w - 1) / w;
  for (int t = threadIdx.x; t < nb * XW; t += NT) {
    const int blk = t / XW, xl = t % XW;
    const int p0 = blk * w;
    vh_block(A + xl, P + xl, p0, min(p0 + w, span), XW);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c < L && x < W)
      dst[base + (i64)c * sa + x] = pkmin(A[i * XW + xl], P[(i + 2 * r) * XW + xl]);
  }
}

// x pass of min/max (packed) fused with the x pass of the box filter.  One wave per row
// (4 rows per work-group), C = 64 * RUN outputs per row tile.  The min/max result is
// needed on [c0-r, c0+C+r) clamped to [0, W) because the box pass clamps to the edge of
// the *filtered* row.  Phase 1 stages the clamped packed input of the row in LDS; phase 2
// gives each lane 4 consecutive window minima (the window middle, shared by all four,
// plus short prefix / suffix runs: about (2r + 12) / 4 packed mins per output); phase 3
// gives each lane RUN consecutive box sums by sliding (integers: exact in any order).
// Output: float2 (E_x, D_x) = (sum(min) * k, sum(max) * k).  R > 0 fixes the radius at
// compile time (r must equal it); R = 0 takes r at run time.
template <int R, int RUN>
__global__ __launch_bounds__(NT) void k_minmax_x_box(const uint32_t* __restrict__ src,
                                                     float2* __restrict__ dst, int W, i64 nrows,
                                                     int r_rt, int nct, float k) {
  static_assert(RUN == 4, "four outputs per lane");
  constexpr int C = 64 * RUN;
  const int r = R > 0 ? R : r_rt;
  extern __shared__ __align__(16) uint32_t smem_x[];
  const int spanmax = C + 4 * r + 4;  // a multiple of 4: both rows stay 16-byte aligned
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t* A = smem_x + wave * 2 * spanmax;  // clamped packed input of the row
  uint32_t* B = A + spanmax;                  // window minima
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = tile % nct;
  const i64 row = (i64)(tile / nct) * (NT / 64) + wave;
  const bool live = row < nrows;  // wave-uniform
  const int c0 = ct * C;
  const int qa = max(0, c0 - r), qb = min(W, c0 + C + r);
  const int nq = qb - qa, span = nq + 2 * r;
  const uint32_t* srow = src + (live ? row : 0) * W;
  if (live)
    for (int p = lane; p < span; p += 64) A[p] = srow[clampi(qa - r + p, 0, W - 1)];
  __syncthreads();
  // window minima B[j] = min(A[j .. j + 2r]), j in [0, nq), four per lane: the window
  // middle A[j0+3 .. j0+2r] is shared by all four, plus short prefix / suffix runs
  for (int j0 = lane * 4; live && j0 < nq; j0 += 256) {
    uint32_t o[4];
    if constexpr (R >= 2) {  // compile-time radius: 16-byte LDS reads, conflict-free
      constexpr int NV = (2 * R + 4) / 4;
      uint32_t w[4 * NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const uint4 q = *reinterpret_cast<const uint4*>(A + j0 + 4 * v);
        w[4 * v] = q.x; w[4 * v + 1] = q.y; w[4 * v + 2] = q.z; w[4 * v + 3] = q.w;
      }
      const uint32_t mid = mid_min<R>(w);
      const uint32_t l2 = w[2], l1 = pkmin(w[1], l2), l0 = pkmin(w[0], l1);
      const uint32_t r1 = w[2 * R + 1], r2 = pkmin(r1, w[2 * R + 2]), r3 = pkmin(r2, w[2 * R + 3]);
      o[0] = pkmin(mid, l0);
      o[1] = pkmin(pkmin(mid, l1), r1);
      o[2] = pkmin(pkmin(mid, l2), r2);
      o[3] = pkmin(mid, r3);
    } else if (2 * r + 1 >= 4) {
      uint32_t mid = A[j0 + 3];
      for (int t = 4; t <= 2 * r; ++t) mid = pkmin(mid, A[j0 + t]);
      const uint32_t l2 = A[j0 + 2], l1 = pkmin(A[j0 + 1], l2), l0 = pkmin(A[j0], l1);
      const uint32_t r1 = A[j0 + 2 * r + 1], r2 = pkmin(r1, A[j0 + 2 * r + 2]),
                     r3 = pkmin(r2, A[j0 + 2 * r + 3]);
      o[0] = pkmin(mid, l0);
      o[1] = pkmin(pkmin(mid, l1), r1);
      o[2] = pkmin(pkmin(mid, l2), r2);
      o[3] = pkmin(mid, r3);
    } else {  // r = 1
#pragma unroll
      for (int q = 0; q < 4; ++q) o[q] = pkmin(pkmin(A[j0 + q], A[j0 + q + 1]), A[j0 + q + 2]);
    }
    *reinterpret_cast<uint4*>(B + j0) = make_uint4(o[0], o[1], o[2], o[3]);
  }
  __syncthreads();
  const int xs = c0 + lane * RUN;
  if (!live || xs >= W) return;
  const bool interior = c0 - r >= 0 && c0 + C + r <= W;  // no clamping in this tile
  const uint32_t full = (uint32_t)(2 * r + 1) * 65535u;
  float2 out[RUN];
  if (R > 0 && interior) {
    // the window minima of x - r .. x + 3 + r are B[4 lane .. 4 lane + 2r + 3] (qa = c0 - r)
    constexpr int RR = R > 0 ? R : 1;
    constexpr int NV = (2 * RR + 4 + 3) / 4;
    uint32_t w[4 * NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const uint4 q = *reinterpret_cast<const uint4*>(B + lane * 4 + 4 * v);
      w[4 * v] = q.x; w[4 * v + 1] = q.y; w[4 * v + 2] = q.z; w[4 * v + 3] = q.w;
    }
    uint32_t se = 0, sdn = 0;  // sum of min, sum of (65535 - max)
#pragma unroll
    for (int t = 0; t <= 2 * RR; ++t) {
      se += w[t] & 0xFFFFu;
      sdn += w[t] >> 16;
    }
#pragma unroll
    for (int q = 0; q < RUN; ++q) {
      if (q > 0) {
        const uint32_t vin = w[2 * RR + q], vout = w[q - 1];
        se += (vin & 0xFFFFu) - (vout & 0xFFFFu);
        sdn += (vin >> 16) - (vout >> 16);
      }
      out[q] = make_float2((float)se * k, (float)(full - sdn) * k);
    }
  } else {
    const uint32_t* m = B - qa;  // m[x]: window minimum at x, x in [qa, qb)
    auto at = [&](int x) { return m[clampi(x, 0, W - 1)]; };
    uint32_t se = 0, sdn = 0;
    for (int j = -r; j <= r; ++j) {
      const uint32_t v = at(xs + j);
      se += v & 0xFFFFu;
      sdn += v >> 16;
    }
#pragma unroll
    for (int q = 0; q < RUN; ++q) {
      if (q > 0) {
        const int x = xs + q;
        const uint32_t vin = at(x + r), vout = at(x - r - 1);
        se += (vin & 0xFFFFu) - (vout & 0xFFFFu);
        sdn += (vin >> 16) - (vout >> 16);
      }
      out[q] = make_float2((float)se * k, (float)(full - sdn) * k);
    }
  }
  float2* drow = dst + row * W;
  // 16-byte stores (a wave writes 1 KB in two instructions) where the row start allows
  if (xs + RUN <= W && ((row * W) & 1) == 0) {
#pragma unroll
    for (int q = 0; q < RUN; q += 2)
      *reinterpret_cast<float4*>(drow + xs + q) =
          make_float4(out[q].x, out[q].y, out[q + 1].x, out[q + 1].y);
  } else {
#pragma unroll
    for (int q = 0; q < RUN; ++q)
      if (xs + q < W) drow[xs + q] = out[q];
  }
}

// With NORM the z pass also normalizes against img.
template <int XW, bool NORM, typename TS>
__global__ __launch_bounds__(NT) void k_box_line(const float2* __restrict__ src,
                                                 float2* __restrict__ dst, TS* __restrict__ norm,
                                                 const uint16_t* __restrict__ img, float min_int,
                                                 int W, int L, i64 sa, i64 so, int r, int C,
                                                 int nxt, int nct, float k) {
  extern __shared__ float2 smem2[];
  const int span = C + 2 * r;
  unsigned b = blockIdx.x;
  const int xt = b % nxt; b /= nxt;
  const int ct = b % nct; b /= nct;
  const i64 base = (i64)b * so;
  const int x0 = xt * XW, c0 = ct * C;
  stage_tile<8, NT, float2>(
      span * XW,
      [&](int e) -> float2 {
        const int p = e / XW, x = x0 + e % XW;
        if (x >= W) return make_float2(0.0f, 0.0f);
        return src[base + (i64)clampi(c0 - r + p, 0, L - 1) * sa + x];
      },
      [&](int e, float2 v) { smem2[e] = v; });
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c >= L || x >= W) continue;
    float ve = 0.0f, vd = 0.0f;
#pragma unroll 4
    for (int j = 0; j <= 2 * r; ++j) {
      const float2 t = smem2[(i + j) * XW + xl];
      ve += t.x;
      vd += t.y;
    }
    ve *= k;
    vd *= k;
    const i64 idx = base + (i64)c * sa + x;
    if (NORM) {
      const float I = (float)img[idx];
      norm[idx] = cvt<TS>((I >= min_int) ? (I - ve) / (vd - ve) : 0.0f);
    } else {
      dst[idx] = make_float2(ve, vd);
    }
  }
}

// ---- column walkers (compile-time radius) ------------------------------------------
// One thread owns one line along the axis (x-consecutive threads -> coalesced rows) and
// walks it front to back, so every input is read once and no LDS is needed.
//
// Min / max: van Herk blocks of w = 2R+1 over the clamp-extended line g[k] = f[c(k-R)].
// With S = suffix minima of the previous block and P = prefix minima of the current
// one, output c = (b-1)w + j is S[0] for j = 0 and min(S[j], P[j-1]) otherwise.
template <int R, bool FROM_U16>
__global__ __launch_bounds__(NT) void k_minmax_walk(const void* __restrict__ src,
                                                    uint32_t* __restrict__ dst, int W, int L,
                                                    i64 sa, i64 so, int nxb) {
  constexpr int WB = 2 * R + 1;
  const int x = (blockIdx.x % nxb) * NT + threadIdx.x;
  if (x >= W) return;
  const i64 base = (i64)(blockIdx.x / nxb) * so + x;
  auto load = [&](int k) -> uint32_t {
    const i64 idx = base + (i64)clampi(k - R, 0, L - 1) * sa;
    return FROM_U16 ? pack(((const uint16_t*)src)[idx]) : ((const uint32_t*)src)[idx];
  };
  uint32_t S[WB], cur[WB];
#pragma unroll
  for (int j = 0; j < WB; ++j) cur[j] = load(j);
  S[WB - 1] = cur[WB - 1];
#pragma unroll
  for (int j = WB - 2; j >= 0; --j) S[j] = pkmin(cur[j], S[j + 1]);
  const int nb = (L + WB - 1) / WB;
  for (int b = 1; b <= nb; ++b) {
#pragma unroll
    for (int j = 0; j < WB; ++j) cur[j] = load(b * WB + j);
    const int c0 = (b - 1) * WB;
    uint32_t p = cur[0];
#pragma unroll
    for (int j = 0; j < WB; ++j) {
      if (j > 0) {
        if (c0 + j < L) st_nt(dst + base + (i64)(c0 + j) * sa, pkmin(S[j], p));
        p = pkmin(p, cur[j]);
      } else if (c0 < L) {
        st_nt(dst + base + (i64)c0 * sa, S[0]);
      }
    }
    S[WB - 1] = cur[WB - 1];
#pragma unroll
    for (int j = WB - 2; j >= 0; --j) S[j] = pkmin(cur[j], S[j + 1]);
  }
}

// Min / max along y and z in one pass: a work-group of 16 waves owns a 64-column x 32-row
// tile of every plane and walks z.  Per slice:
//   1. the packed rows y0-RY .. y0+31+RY of the slice (clamped to the plane; raw u16
//      loaded PF slices ahead) enter an LDS buffer (two buffers, alternating: the next
//      slice is stored while this one is read, one barrier per slice);
//   2. each thread (lane = column, wave = row pair) forms the y window minima of its two
//      rows from the 2RY+2 values of its column: a shared window middle plus one end value
//      each;
//   3. these enter the thread's z walk: van Herk blocks of WB = 2RZ+1 slices over the
//      clamp-extended line g[k] = slice(clamp(k - RZ)), held in place in a register array:
//      slot j holds the suffix minimum S[j] of the previous block until slice j of the
//      current block takes it, and a running prefix minimum P covers the current block,
//      so output c = min(S[j+1], P) leaves when g[c + 2RZ] arrives; at a block's end its
//      suffix minima replace it in place and give the output c = k - 2RZ as S[0].
// Min is exact, so y then z equals the z then y of the unfused passes bit for bit.  The
// z-min/maxed volume between them (8 B/voxel of traffic) is never stored.
template <int RZ, int RY, int PF>
__global__ __launch_bounds__(1024) void k_minmax_zy_walk(const uint16_t* __restrict__ img,
                                                        uint32_t* __restrict__ dst, int W, int H,
                                                        int D, int nyt) {
  constexpr int TY = 32, KR = 2, NWV = TY / KR;  // rows per tile, rows per thread, waves
  constexpr int SR = TY + 2 * RY;                 // staged rows per slice
  constexpr int NLD = (SR + NWV - 1) / NWV;       // row loads per lane and slice
  constexpr int SRP = NLD * NWV;  // rows of an LDS buffer: every wave loads and stores NLD
                                  // rows (rows past SR repeat a clamped row, unused)
  constexpr int WB = 2 * RZ + 1;
  constexpr int NW = 2 * RY + KR;                 // window values per thread
  static_assert(WB % PF == 0, "queue slots must be static in the unrolled block");
  static_assert(RY >= 2, "shared window middle");
  __shared__ uint32_t S[2][SRP * 64];
  const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
  const int yt = (int)(t % (unsigned)nyt), xt = (int)(t / (unsigned)nyt);
  const int lane = (int)(threadIdx.x & 63);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform
  const int x = xt * 64 + lane, y0 = yt * TY;
  const int xc = min(x, W - 1);
  const size_t WH = (size_t)W * (size_t)H;
  // staged row wv + NWV j of every slice: its offset in the slice (rows clamped); no
  // load or LDS store is under a condition
  uint32_t off[NLD];
#pragma unroll
  for (int j = 0; j < NLD; ++j)
    off[j] = (uint32_t)clampi(y0 - RY + wv + NWV * j, 0, H - 1) * (uint32_t)W + (uint32_t)xc;
  // loads and stores go through a buffer descriptor of the slice (scalar) with 32-bit
  // lane offsets: with 64-bit lane addresses the compiler kept one per unrolled step
  uint32_t q[PF][NLD];  // raw values of slice k wait in slot k % PF
  auto fetch = [&](int k, int slot) {
    const __amdgpu_buffer_rsrc_t r =
        buf_rsrc(img + (size_t)clampi(k - RZ, 0, D - 1) * WH, (uint32_t)(WH * 2));
#pragma unroll
    for (int j = 0; j < NLD; ++j)
      q[slot][j] = __builtin_amdgcn_raw_buffer_load_b16(r, (int)(off[j] * 2u), 0, 0);
  };
  // slice k's packed rows into buffer k & 1, the loads of slice k + PF into the freed slot.
  // Step k stores slice k + 1 and reads slice k's rows, then one barrier: every wave's
  // stores of slice k + 1 are then done before step k + 1 reads them, and its reads of
  // slice k before step k + 1 stores slice k + 2 over them (one wait per slice, not two).
  auto stage = [&](int k, int slot) {
    uint32_t* Sb = S[k & 1];
#pragma unroll
    for (int j = 0; j < NLD; ++j)
      Sb[(wv + NWV * j) * 64 + lane] = pack(q[slot][j]);
    fetch(k + PF, slot);
  };
  const int ty = KR * wv;  // first output row of this thread in the tile
  // y window minima of rows ty, ty + 1: staged rows ty .. ty + 2RY + 1 of column lane
  auto ymin = [&](const uint32_t* Sb, uint32_t o[KR]) {
    uint32_t w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = Sb[(ty + i) * 64 + lane];
    // middle w[1 .. 2RY], shared by both windows, in four interleaved chains
    uint32_t m[4] = {w[1], w[2], w[3], w[4]};
#pragma unroll
    for (int i = 5; i <= 2 * RY; ++i) m[(i - 1) & 3] = pkmin(m[(i - 1) & 3], w[i]);
    const uint32_t mid = pkmin(pkmin(m[0], m[1]), pkmin(m[2], m[3]));
    o[0] = pkmin(mid, w[0]);
    o[1] = pkmin(mid, w[2 * RY + 1]);
  };
  // byte offsets of the thread's outputs in a slice; outside the plane past the slice
  // descriptor's record count, so those stores are dropped
  uint32_t ro[KR];
#pragma unroll
  for (int i = 0; i < KR; ++i)
    ro[i] = x < W && y0 + ty + i < H ? ((uint32_t)(y0 + ty + i) * (uint32_t)W + (uint32_t)x) * 4u
                                     : 0xFFFFFFF0u;
  auto put = [&](int c, const uint32_t v[KR]) {
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(dst + (size_t)c * WH, (uint32_t)(WH * 4));
#pragma unroll
    for (int i = 0; i < KR; ++i)  // streaming (nontemporal) stores
      __builtin_amdgcn_raw_buffer_store_b32(v[i], r, (int)ro[i], 0, kCachePolicyNT);
  };
  uint32_t zs[KR][WB], P[KR];
  auto suffix = [&]() {
#pragma unroll
    for (int i = 0; i < KR; ++i)
#pragma unroll
      for (int j = WB - 2; j >= 0; --j) zs[i][j] = pkmin(zs[i][j], zs[i][j + 1]);
  };
#pragma unroll
  for (int k = 0; k < PF; ++k) fetch(k, k);
  stage(0, 0);
  __syncthreads();
  // block 0 (slices k = 0 .. WB-1) only fills the slots; its suffix minimum S[0] is
  // output 0
#pragma unroll
  for (int j = 0; j < WB; ++j) {
    stage(j + 1, (j + 1) % PF);
    uint32_t v[KR];
    ymin(S[j & 1], v);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KR; ++i) zs[i][j] = v[i];
  }
  suffix();
  {
    uint32_t v[KR];
#pragma unroll
    for (int i = 0; i < KR; ++i) v[i] = zs[i][0];
    put(0, v);
  }
  // output c leaves when slice k = c + 2RZ has arrived; the last is c = D - 1.  The walk's
  // end leaves the kernel (uniform over the work-group: every wave has the same k).
  const int kend = D + 2 * RZ;
  for (int k0 = WB;; k0 += WB) {
#pragma unroll
    for (int j = 0; j < WB; ++j) {
      const int k = k0 + j;
      if (k >= kend) return;
      stage(k + 1, (j + 1) % PF);  // (k0 + j) % PF == j % PF
      uint32_t v[KR], o[KR];
      ymin(S[k & 1], v);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < KR; ++i) {
        P[i] = j == 0 ? v[i] : pkmin(P[i], v[i]);
        if (j < WB - 1) o[i] = pkmin(zs[i][j + 1], P[i]);
        zs[i][j] = v[i];
      }
      if (j < WB - 1) {
        put(k - 2 * RZ, o);
      } else {
        suffix();
#pragma unroll
        for (int i = 0; i < KR; ++i) o[i] = zs[i][0];
        put(k - 2 * RZ, o);
      }
    }
  }
}

// Box mean: a register ring of the last 2R+1 inputs (clamp), unrolled by the ring period
// so every index is static; each output sums its window in the reference order.  Ring
// refills are loaded PF steps before they enter the ring (PF divides the ring period, so
// every queue slot is static too), and the walk leaves the kernel at the line's end
// instead of skipping the remaining unrolled steps, so no merged control path makes the
// compiler wait for the whole queue.
template <int R, bool NORM, typename TS>
__global__ __launch_bounds__(NT) void k_box_walk(const float2* __restrict__ src,
                                                 float2* __restrict__ dst, TS* __restrict__ norm,
                                                 const uint16_t* __restrict__ img, float min_int,
                                                 int W, int L, i64 sa, i64 so, int nxb, float k) {
  constexpr int WB = 2 * R + 1;
  constexpr int PF = WB % 3 == 0 ? 3 : WB % 5 == 0 ? 5 : WB % 7 == 0 ? 7 : 1;
  const int x = (blockIdx.x % nxb) * NT + threadIdx.x;
  if (x >= W) return;
  const i64 base = (i64)(blockIdx.x / nxb) * so + x;
  auto at = [&](int p) { return src[base + (i64)clampi(p, 0, L - 1) * sa]; };
  float2 ring[WB];  // position p lives in ring[(p + R) % WB]
#pragma unroll
  for (int q = 0; q < WB; ++q) ring[q] = at(q - R);
  float2 pf[PF];  // the refill of step c (position c + R + 1) waits in slot c % PF
#pragma unroll
  for (int j = 0; j < PF; ++j) pf[j] = at(R + 1 + j);
  for (int c0 = 0;; c0 += WB) {
    // with the normalization fused, the raw intensities of this block's WB outputs are
    // loaded up front: loading each one at its output made every output wait for all
    // outstanding loads, the ring refills included
    float iv[WB];
    if (NORM) {
#pragma unroll
      for (int q = 0; q < WB; ++q)
        iv[q] = (float)img[base + (i64)min(c0 + q, L - 1) * sa];
    }
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      const int c = c0 + q;
      if (c >= L) return;
      float ve = 0.0f, vd = 0.0f;
#pragma unroll
      for (int j = 0; j <= 2 * R; ++j) {  // position c - R + j -> ring[(q + j) % WB]
        ve += ring[(q + j) % WB].x;
        vd += ring[(q + j) % WB].y;
      }
      ve *= k;
      vd *= k;
      const i64 idx = base + (i64)c * sa;
      if (NORM) {
        const float I = iv[q];
        st_nt(norm + idx, (I >= min_int) ? (I - ve) / (vd - ve) : 0.0f);
      } else {
        st_nt(dst + idx, make_float2(ve, vd));
      }
      // slot of position c - R is reused for position c + R + 1
      ring[q] = pf[q % PF];
      pf[q % PF] = at(c + R + 1 + PF);
    }
  }
}

// min/max along x, box along x and box along y fused.  A one-wave work-group walks the
// rows of one plane of the y-min/maxed (packed) volume for a strip of CW = 128 columns,
// two per lane.  Per row:
//   1. the clamped packed row segment x0-2R .. x0+CW-1+2R enters an LDS line A;
//   2. window minima B[b] = min(A[b .. b+2R]) (the min/max at x = x0-R+b) are formed four
//      per task (a shared window middle plus short prefix / suffix runs); on a strip at
//      the plane's edge, positions outside [0, W) then take the value at the edge, since
//      the box pass clamps to the edge of the filtered row;
//   3. each lane sums B over its two windows (integers, exact in any order) -> (E_x, D_x)
//      of its columns; B keeps the min channel and the 65535 - max channel in two 16-bit
//      arrays, so one v_dot2 with (1, 1) adds two window positions;
//   4. these enter a register ring of the last 2R+1 rows (rows clamped), and each output
//      row sums its window in the reference order, as the y pass walker does.
// The (E_x, D_x) volume of the unfused passes is never stored.  Rows are fetched PF rows
// ahead of their use; the plane's end leaves the kernel (see the DoG row walker).
template <int R>
__global__ __launch_bounds__(64) void k_norm_xy_walk(const uint32_t* __restrict__ src,
                                                     float4* __restrict__ dst, int W, int H,
                                                     int nxs, int nys, int rows, float k) {
  constexpr int WB = 2 * R + 1;
  constexpr int CW = 128;                    // output columns per wave
  constexpr int HX = (2 * R + 3) / 4 * 4;    // line halo: 16-byte aligned task reads
  constexpr int NB = CW + 2 * R;             // window minima per row
  constexpr int NTASK = (NB + 3) / 4;        // four minima per task
  constexpr int LW = 64 * ((4 * NTASK + 2 * R + 4 + 63) / 64);  // line, whole lane loads
  constexpr int NLD = LW / 64;               // loads per lane per row
  constexpr int PF = WB % 3 == 0 ? 3 : WB % 5 == 0 ? 5 : WB % 7 == 0 ? 7 : 1;
  static_assert(HX == 2 * R, "task windows assume a halo of exactly 2R");
  static_assert(4 * NTASK + 2 * R + 4 <= LW, "task reads stay inside the line");
  __shared__ __align__(16) uint32_t A[LW];
  // B[b] = window minimum b: its min channel in Bm, its 65535 - max channel in Bx, as
  // 16-bit entries, positions 2i and 2i + 1 in word i
  __shared__ __align__(16) uint32_t Bm[2 * NTASK];
  __shared__ __align__(16) uint32_t Bx[2 * NTASK];
  // work-group = one strip of one plane over the rows [y0, y1) of one of the nys segments
  // of the plane (rows: a multiple of WB)
  const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
  const int y0 = (int)(t % nys) * rows, y1 = min(H, y0 + rows);
  const unsigned tile = t / nys;
  const int l = threadIdx.x;
  const int x0 = (int)(tile % nxs) * CW;
  const uint32_t* plane = src + (i64)(tile / nxs) * H * W;
  int xl[NLD];  // clamped column of each line position this lane loads
#pragma unroll
  for (int j = 0; j < NLD; ++j) xl[j] = clampi(x0 - HX + 64 * j + l, 0, W - 1);
  uint32_t qv[PF][NLD];  // row y is held in slot y % PF
  auto fetch = [&](int y, int s) {
    const uint32_t* row = plane + (i64)min(y, H - 1) * W;
#pragma unroll
    for (int j = 0; j < NLD; ++j) qv[s][j] = row[xl[j]];
  };
  // row y waits in queue slot y % PF (y0 is a multiple of WB, so of PF)
  if (y0 == 0) {
#pragma unroll
    for (int s = 0; s < PF; ++s) fetch(s, s);
  } else {
#pragma unroll
    for (int s = 0; s < PF; ++s) fetch(y0 - R + s, ((s - R) % PF + PF) % PF);
  }
  const bool edge = x0 - R < 0 || x0 + CW + R > W;  // wave-uniform
  const int blo = R - x0, bhi = W - 1 - x0 + R;     // B index of x = 0 and x = W - 1
  const uint32_t full = (uint32_t)WB * 65535u;
  const us2 lo1 = {1, 0}, hi1 = {0, 1}, one = {1, 1};
  // (E_x, D_x) of the lane's two columns for row y (slot s); fetches row y + PF into s
  auto xrow = [&](int y, int s) -> float4 {
#pragma unroll
    for (int j = 0; j < NLD; ++j) A[64 * j + l] = qv[s][j];
    fetch(y + PF, s);
    wave_lds_sync();
    if (l < NTASK) {
      constexpr int NV = (2 * R + 4) / 4;
      uint32_t w[4 * NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const uint4 q = *reinterpret_cast<const uint4*>(A + 4 * l + 4 * v);
        w[4 * v] = q.x; w[4 * v + 1] = q.y; w[4 * v + 2] = q.z; w[4 * v + 3] = q.w;
      }
      const uint32_t mid = mid_min<R>(w);
      const uint32_t l2 = w[2], l1 = pkmin(w[1], l2), l0 = pkmin(w[0], l1);
      const uint32_t r1 = w[2 * R + 1], r2 = pkmin(r1, w[2 * R + 2]), r3 = pkmin(r2, w[2 * R + 3]);
      const uint32_t b0 = pkmin(mid, l0), b1 = pkmin(pkmin(mid, l1), r1),
                     b2 = pkmin(pkmin(mid, l2), r2), b3 = pkmin(mid, r3);
      // v_perm_b32: perm(b1, b0, 0x05040100) = lo16(b0) | lo16(b1) << 16,
      //             perm(b1, b0, 0x07060302) = hi16(b0) | hi16(b1) << 16
      *reinterpret_cast<uint2*>(Bm + 2 * l) =
          make_uint2(__builtin_amdgcn_perm(b1, b0, 0x05040100u),
                     __builtin_amdgcn_perm(b3, b2, 0x05040100u));
      *reinterpret_cast<uint2*>(Bx + 2 * l) =
          make_uint2(__builtin_amdgcn_perm(b1, b0, 0x07060302u),
                     __builtin_amdgcn_perm(b3, b2, 0x07060302u));
    }
    wave_lds_sync();
    if (edge) {  // positions outside the plane take the edge value
      constexpr int NF = (4 * NTASK + 63) / 64;
      uint16_t* bm = reinterpret_cast<uint16_t*>(Bm);
      uint16_t* bx = reinterpret_cast<uint16_t*>(Bx);
      uint16_t fm[NF], fx[NF];
#pragma unroll
      for (int j = 0; j < NF; ++j) {
        const int b = 64 * j + l;
        const int c = clampi(b, blo, bhi);
        fm[j] = b < 4 * NTASK ? bm[c] : (uint16_t)0;
        fx[j] = b < 4 * NTASK ? bx[c] : (uint16_t)0;
      }
      wave_lds_sync();
#pragma unroll
      for (int j = 0; j < NF; ++j)
        if (64 * j + l < 4 * NTASK) {
          bm[64 * j + l] = fm[j];
          bx[64 * j + l] = fx[j];
        }
      wave_lds_sync();
    }
    // column x0 + 2l sums positions 2l .. 2l + 2R: words l .. l + R - 1 and the low half
    // of word l + R; column x0 + 2l + 1 sums 2l + 1 .. 2l + 2R + 1: the high half of word
    // l and words l + 1 .. l + R
    uint32_t m[R + 1], x[R + 1];
#pragma unroll
    for (int t = 0; t <= R; ++t) {
      m[t] = Bm[l + t];
      x[t] = Bx[l + t];
    }
    uint32_t sm = 0, sx = 0;
#pragma unroll
    for (int t = 1; t < R; ++t) {
      sm = __builtin_amdgcn_udot2(__builtin_bit_cast(us2, m[t]), one, sm, false);
      sx = __builtin_amdgcn_udot2(__builtin_bit_cast(us2, x[t]), one, sx, false);
    }
    const us2 m0 = __builtin_bit_cast(us2, m[0]), mr = __builtin_bit_cast(us2, m[R]);
    const us2 x0w = __builtin_bit_cast(us2, x[0]), xr = __builtin_bit_cast(us2, x[R]);
    const uint32_t se = __builtin_amdgcn_udot2(m0, one, __builtin_amdgcn_udot2(mr, lo1, sm, false), false);
    const uint32_t se1 = __builtin_amdgcn_udot2(m0, hi1, __builtin_amdgcn_udot2(mr, one, sm, false), false);
    const uint32_t sdn = __builtin_amdgcn_udot2(x0w, one, __builtin_amdgcn_udot2(xr, lo1, sx, false), false);
    const uint32_t sdn1 = __builtin_amdgcn_udot2(x0w, hi1, __builtin_amdgcn_udot2(xr, one, sx, false), false);
    float4 o = make_float4((float)se * k, (float)(full - sdn) * k, (float)se1 * k,
                           (float)(full - sdn1) * k);
    // pin the row's outputs here (see the DoG row walker)
    asm volatile("" : "+v"(o.x), "+v"(o.y), "+v"(o.z), "+v"(o.w));
    return o;
  };
  const int xc = x0 + 2 * l;
  const bool st2 = xc + 1 < W && (W & 1) == 0, st1 = xc < W;
  float4* dq = dst + ((tile / nxs) * (i64)H * W + xc) / 2;  // valid when st2
  float2* dp = reinterpret_cast<float2*>(dst) + (tile / nxs) * (i64)H * W + xc;
  float4 ring[WB];  // row p lives in ring[(p + R) % WB]; rows clamp to [0, H)
  if (y0 == 0) {
    ring[R] = xrow(0, 0);
#pragma unroll
    for (int q = 0; q < R; ++q) ring[q] = ring[R];
#pragma unroll
    for (int q = R + 1; q < WB; ++q) ring[q] = xrow(min(q - R, H - 1), (q - R) % PF);
  } else {  // a later segment starts from the R rows above it
#pragma unroll
    for (int q = 0; q < WB; ++q) ring[q] = xrow(y0 - R + q, ((q - R) % PF + PF) % PF);
  }
  for (int c0 = y0;; c0 += WB) {
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      const int c = c0 + q;
      if (c >= y1) return;
      float e0 = 0.0f, d0 = 0.0f, e1 = 0.0f, d1 = 0.0f;
#pragma unroll
      for (int j = 0; j <= 2 * R; ++j) {  // row c - R + j -> ring[(q + j) % WB]
        const float4 t = ring[(q + j) % WB];
        e0 += t.x;
        d0 += t.y;
        e1 += t.z;
        d1 += t.w;
      }
      const float4 o = make_float4(e0 * k, d0 * k, e1 * k, d1 * k);
      if (st2) {
        st_nt(dq + (i64)c * W / 2, o);
      } else if (st1) {
        dp[(i64)c * W] = make_float2(o.x, o.y);
        if (xc + 1 < W) dp[(i64)c * W + 1] = make_float2(o.z, o.w);
      }
      ring[q] = xrow(min(c + R + 1, H - 1), (q + R + 1) % PF);  // replaces row c - R
    }
  }
}

template <typename TS>
__global__ __launch_bounds__(NT) void k_copy_norm(const uint16_t* __restrict__ img,
                                                  TS* __restrict__ norm, float min_int, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const float v = (float)img[i];
    norm[i] = cvt<TS>((v >= min_int) ? v : 0.0f);
  }
}

struct LineCfg { int XW, C; };
LineCfg line_cfg(int r, int L, size_t elem_bytes, int arrays) {
  LineCfg c{64, 64};
  if (r > 24) c.XW = 32;
  if ((size_t)arrays * (c.C + 2 * r) * c.XW * elem_bytes > 60 * 1024) c.XW = 16;
  if (L < c.C) c.C = L;
  return c;
}

template <typename F>
hipError_t dispatch_xw(int XW, F&& f) {
  if (XW == 64) return f(std::integral_constant<int, 64>{});
  if (XW == 32) return f(std::integral_constant<int, 32>{});
  return f(std::integral_constant<int, 16>{});
}

template <int R>
hipError_t minmax_walk(const void* src, bool from_u16, uint32_t* dst, int W, int L, i64 sa, i64 so,
                       i64 n_other, hipStream_t s) {
  const int nxb = (W + NT - 1) / NT;
  const i64 nblk = (i64)nxb * n_other;
  if (from_u16)
    hipLaunchKernelGGL((k_minmax_walk<R, true>), dim3((unsigned)nblk), dim3(NT), 0, s, src, dst, W,
                       L, sa, so, nxb);
  else
    hipLaunchKernelGGL((k_minmax_walk<R, false>), dim3((unsigned)nblk), dim3(NT), 0, s, src, dst,
                       W, L, sa, so, nxb);
  return hipGetLastError();
}

hipError_t minmax_line(const void* src, bool from_u16, uint32_t* dst, int W, int L, i64 sa,
                       i64 so, i64 n_other, int r, hipStream_t s) {
  switch (r) {  // radii with a column-walker kernel
    case 1: return minmax_walk<1>(src, from_u16, dst, W, L, sa, so, n_other, s);
    case 2: return minmax_walk<2>(src, from_u16, dst, W, L, sa, so, n_other, s);
    case 3: return minmax_walk<3>(src, from_u16, dst, W, L, sa, so, n_other, s);
    case 5: return minmax_walk<5>(src, from_u16, dst, W, L, sa, so, n_other, s);
    case 10: return minmax_walk<10>(src, from_u16, dst, W, L, sa, so, n_other, s);
    default: break;
  }
  const LineCfg c = line_cfg(r, L, 4, 2);
  const int nxt = (W + c.XW - 1) / c.XW, nct = (L + c.C - 1) / c.C;
  const i64 nblk = (i64)nxt * nct * n_other;
  const size_t lds = 2 * (size_t)(c.C + 2 * r) * c.XW * 4;
  return dispatch_xw(c.XW, [&](auto xw) {
    constexpr int XW = decltype(xw)::value;
    if (from_u16)
      hipLaunchKernelGGL((k_minmax_line<XW, true>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                         dst, W, L, sa, so, r, c.C, nxt, nct);
    else
      hipLaunchKernelGGL((k_minmax_line<XW, false>), dim3((unsigned)nblk), dim3(NT), lds, s,
                         src, dst, W, L, sa, so, r, c.C, nxt, nct);
    return hipGetLastError();
  });
}

template <int R, bool NORM, typename TS>
hipError_t box_line_fast(const float2* src, float2* dst, TS* norm, const uint16_t* img,
                         float min_int, int W, int L, i64 sa, i64 so, i64 n_other, float k,
                         hipStream_t s) {
  const int nxb = (W + NT - 1) / NT;
  const i64 nblk = (i64)nxb * n_other;
  hipLaunchKernelGGL((k_box_walk<R, NORM, TS>), dim3((unsigned)nblk), dim3(NT), 0, s, src, dst,
                     norm, img, min_int, W, L, sa, so, nxb, k);
  return hipGetLastError();
}

template <bool NORM, typename TS>
hipError_t box_line(const float2* src, float2* dst, TS* norm, const uint16_t* img, float min_int,
                    int W, int L, i64 sa, i64 so, i64 n_other, int r, hipStream_t s) {
  const float k = 1.0f / (float)(2 * r + 1);
  switch (r) {  // radii with a register-window kernel
    case 1: return box_line_fast<1, NORM, TS>(src, dst, norm, img, min_int, W, L, sa, so, n_other, k, s);
    case 2: return box_line_fast<2, NORM, TS>(src, dst, norm, img, min_int, W, L, sa, so, n_other, k, s);
    case 3: return box_line_fast<3, NORM, TS>(src, dst, norm, img, min_int, W, L, sa, so, n_other, k, s);
    case 5: return box_line_fast<5, NORM, TS>(src, dst, norm, img, min_int, W, L, sa, so, n_other, k, s);
    case 10: return box_line_fast<10, NORM, TS>(src, dst, norm, img, min_int, W, L, sa, so, n_other, k, s);
    default: break;
  }
  const LineCfg c = line_cfg(r, L, 8, 1);
  const int nxt = (W + c.XW - 1) / c.XW, nct = (L + c.C - 1) / c.C;
  const i64 nblk = (i64)nxt * nct * n_other;
  const size_t lds = (size_t)(c.C + 2 * r) * c.XW * 8;
  return dispatch_xw(c.XW, [&](auto xw) {
    constexpr int XW = decltype(xw)::value;
    hipLaunchKernelGGL((k_box_line<XW, NORM, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, norm, img, min_int, W, L, sa, so, r, c.C, nxt, nct, k);
    return hipGetLastError();
  });
}

__global__ __launch_bounds__(NT) void k_split_ed(const float2* __restrict__ ed,
                                                 float* __restrict__ ero, float* __restrict__ dil,
                                                 i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const float2 v = ed[i];
    ero[i] = v.x;
    dil[i] = v.y;
  }
}

}  // namespace

hipError_t launch_box_pair(const float2* src, float2* dst, int W, int L, i64 sa, i64 so,
                           i64 n_other, int r, hipStream_t s) {
  return box_line<false, float>(src, dst, nullptr, nullptr, 0.0f, W, L, sa, so, n_other, r, s);
}

template <typename TS>
hipError_t launch_normalize(const uint16_t* img, TS* norm, void* tmp1, void* tmp2,
                            float min_intensity, int W, int H, int D, int rxy, int rz,
                            hipStream_t s, float* ero, float* dil, int segments) {
  const i64 WH = (i64)W * H, N = WH * D;
  if (rxy == 0 || rz == 0) {
    const i64 nb = std::min<i64>((N + NT - 1) / NT, 1 << 20);
    hipLaunchKernelGGL((k_copy_norm<TS>), dim3((unsigned)nb), dim3(NT), 0, s, img, norm,
                       min_intensity, N);
    mark("normalize/copy", s);
    return hipGetLastError();
  }
  hipError_t e;
  uint32_t* p1 = (uint32_t*)tmp1;
  uint32_t* p2 = (uint32_t*)tmp2;
  if (rxy == 10 && rz == 10 && WH * 4 < 0xFFFFFFF0ll) {
    // min/max along y and z in one z walk over 64 x 32 tiles (slices addressed by buffer
    // descriptors with 32-bit byte offsets)
    const int nxt = (W + 63) / 64, nyt = (H + 31) / 32;
    hipLaunchKernelGGL((k_minmax_zy_walk<10, 10, HDOG_ZY_PF>), dim3((unsigned)((i64)nxt * nyt)),
                       dim3(1024), 0, s, img, p2, W, H, D, nyt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("normalize/minmax_zy", s);
  } else {
    // min/max along z (lines over x, y fixed per work-group), then along y
    if ((e = minmax_line(img, true, p1, W, D, WH, W, H, rz, s)) != hipSuccess) return e;
    mark("normalize/minmax_z", s);
    if ((e = minmax_line(p1, false, p2, W, H, W, WH, D, rxy, s)) != hipSuccess) return e;
    mark("normalize/minmax_y", s);
  }
  if (rxy == 10 && !(ero && dil)) {
    // min/max along x, box along x and box along y in one row walker -> (E, D) in tmp1,
    // then box along z fused with normalization
    constexpr int CW = 128;
    const int nxs = (W + CW - 1) / CW;
    int nys, rows;
    walk_segments(H, 21, &nys, &rows, segments);
    const i64 nblk = (i64)nxs * D * nys;
    hipLaunchKernelGGL((k_norm_xy_walk<10>), dim3((unsigned)nblk), dim3(64), 0, s, p2,
                       (float4*)tmp1, W, H, nxs, nys, rows, 1.0f / 21.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("normalize/minmax_x_box_xy", s);
    e = box_line<true, TS>((const float2*)tmp1, nullptr, norm, img, min_intensity, W, D, WH, W, H,
                           rz, s);
    mark("normalize/box_z_norm", s);
    return e;
  }
  // min/max along x fused with the box x pass -> float2 (E, D) in tmp1
  {
    constexpr int RUN = 4, C = 64 * RUN, ROWS = NT / 64;
    const int nct = (W + C - 1) / C;
    const i64 nrows = (i64)H * D;
    const i64 nblk = ((nrows + ROWS - 1) / ROWS) * nct;
    const size_t lds = (size_t)ROWS * 2 * (C + 4 * rxy + 4) * 4;
    const float k = 1.0f / (float)(2 * rxy + 1);
    if (rxy == 10)
      hipLaunchKernelGGL((k_minmax_x_box<10, RUN>), dim3((unsigned)nblk), dim3(NT), lds, s, p2,
                         (float2*)tmp1, W, nrows, rxy, nct, k);
    else
      hipLaunchKernelGGL((k_minmax_x_box<0, RUN>), dim3((unsigned)nblk), dim3(NT), lds, s, p2,
                         (float2*)tmp1, W, nrows, rxy, nct, k);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("normalize/minmax_x_box_x", s);
  }
  // box along y, then along z fused with normalization
  if ((e = box_line<false, TS>((const float2*)tmp1, (float2*)tmp2, nullptr, nullptr, 0.0f, W, H,
                               W, WH, D, rxy, s)) != hipSuccess)
    return e;
  mark("normalize/box_y", s);
  if (ero && dil) {  // E and D on request: the same z pass without normalization, then split
    if ((e = box_line<false, TS>((const float2*)tmp2, (float2*)tmp1, nullptr, nullptr, 0.0f, W, D,
                                 WH, W, H, rz, s)) != hipSuccess)
      return e;
    const i64 nb = std::min<i64>((N + NT - 1) / NT, 1 << 20);
    hipLaunchKernelGGL(k_split_ed, dim3((unsigned)nb), dim3(NT), 0, s, (const float2*)tmp1, ero,
                       dil, N);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  e = box_line<true, TS>((const float2*)tmp2, nullptr, norm, img, min_intensity, W, D, WH, W, H,
                         rz, s);
  mark("normalize/box_z_norm", s);
  return e;
}

template hipError_t launch_normalize<float>(const uint16_t*, float*, void*, void*, float, int,
                                            int, int, int, int, hipStream_t, float*, float*, int);
template hipError_t launch_normalize<__half>(const uint16_t*, __half*, void*, void*, float, int,
                                             int, int, int, int, hipStream_t, float*, float*, int);
template hipError_t launch_normalize<__hip_bfloat16>(const uint16_t*, __hip_bfloat16*, void*,
                                                     void*, float, int, int, int, int, int,
                                                     hipStream_t, float*, float*, int);

}  // namespace hdog
