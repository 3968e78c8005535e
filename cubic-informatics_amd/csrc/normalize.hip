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
//   * Min/max over a clamped box is separable and exact in any order, so the passes run
//     z, y, x and the x pass is fused with the x pass of the box filter (integer-valued
//     sums, exact in any order).  The y and z box passes sum FP32 values and keep the
//     reference's sequential order (c-r .. c+r) for bit-exactness; the z pass is fused
//     with the normalization.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pkmin(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2, a),
                                                                __builtin_bit_cast(us2, b)));
}
__device__ __forceinline__ uint32_t pack(uint32_t v) { return v | ((65535u - v) << 16); }

// 1-D packed min/max along y (sa = W) or z (sa = W*H).  Work-group tile: XW consecutive
// x positions times C consecutive outputs along the axis, at one coordinate `o` of the
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
  for (int e = threadIdx.x; e < span * XW; e += NT) {
    const int p = e / XW, xl = e % XW, x = x0 + xl;
    uint32_t v = 0;
    if (x < W) {
      const i64 idx = base + (i64)clampi(c0 - r + p, 0, L - 1) * sa + x;
      v = FROM_U16 ? pack(((const uint16_t*)src)[idx]) : ((const uint32_t*)src)[idx];
    }
    A[e] = v;
  }
  __syncthreads();
  const int w = 2 * r + 1;
  const int nb = (span + w - 1) / w;
  for (int t = threadIdx.x; t < nb * XW; t += NT) {
    const int blk = t / XW, xl = t % XW;
    const int p0 = blk * w, p1 = min(p0 + w, span);
    uint32_t m = A[p0 * XW + xl];
    P[p0 * XW + xl] = m;
    for (int p = p0 + 1; p < p1; ++p) { m = pkmin(m, A[p * XW + xl]); P[p * XW + xl] = m; }
    m = A[(p1 - 1) * XW + xl];
    for (int p = p1 - 2; p >= p0; --p) { m = pkmin(m, A[p * XW + xl]); A[p * XW + xl] = m; }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c < L && x < W)
      dst[base + (i64)c * sa + x] = pkmin(A[i * XW + xl], P[(i + 2 * r) * XW + xl]);
  }
}

// x pass of min/max (packed) fused with the x pass of the box filter.  Tile: ROWS rows
// times C outputs.  The min/max result is needed on [c0-r, c0+C+r) clamped to [0, W)
// because the box pass clamps to the edge of the *filtered* row.
// Output: float2 (E_x, D_x) = (sum_x(min) * k, sum_x(max) * k); sums of integers.
template <int ROWS>
__global__ __launch_bounds__(NT) void k_minmax_x_box(const uint32_t* __restrict__ src,
                                                     float2* __restrict__ dst, int W, i64 nrows,
                                                     int r, int C, int nct, float k) {
  extern __shared__ uint32_t smem[];
  const int spanmax = C + 4 * r;
  uint32_t* A = smem;
  uint32_t* P = smem + ROWS * spanmax;
  const int ct = blockIdx.x % nct;
  const i64 row0 = (i64)(blockIdx.x / nct) * ROWS;
  const int c0 = ct * C;
  const int qa = max(0, c0 - r), qb = min(W, c0 + C + r);
  const int nq = qb - qa;
  const int span = nq + 2 * r;
  for (int e = threadIdx.x; e < ROWS * span; e += NT) {
    const int row = e / span, p = e % span;
    const i64 grow = row0 + row;
    uint32_t v = 0;
    if (grow < nrows) v = src[grow * W + clampi(qa - r + p, 0, W - 1)];
    A[row * spanmax + p] = v;
  }
  __syncthreads();
  const int w = 2 * r + 1;
  const int nb = (span + w - 1) / w;
  for (int t = threadIdx.x; t < ROWS * nb; t += NT) {
    const int row = t / nb, blk = t % nb;
    uint32_t* a = A + row * spanmax;
    uint32_t* pp = P + row * spanmax;
    const int p0 = blk * w, p1 = min(p0 + w, span);
    uint32_t m = a[p0];
    pp[p0] = m;
    for (int p = p0 + 1; p < p1; ++p) { m = pkmin(m, a[p]); pp[p] = m; }
    m = a[p1 - 1];
    for (int p = p1 - 2; p >= p0; --p) { m = pkmin(m, a[p]); a[p] = m; }
  }
  __syncthreads();
  // filtered value at position qa + j, stored in place over the suffix array
  for (int e = threadIdx.x; e < ROWS * nq; e += NT) {
    const int row = e / nq, j = e % nq;
    A[row * spanmax + j] = pkmin(A[row * spanmax + j], P[row * spanmax + j + 2 * r]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ROWS * C; e += NT) {
    const int row = e / C, i = e % C, x = c0 + i;
    const i64 grow = row0 + row;
    if (x >= W || grow >= nrows) continue;
    const uint32_t* m = A + row * spanmax - qa;
    float se = 0.0f, sd = 0.0f;
    for (int j = -r; j <= r; ++j) {
      const uint32_t v = m[clampi(x + j, 0, W - 1)];
      se += (float)(v & 0xFFFFu);
      sd += (float)(65535u - (v >> 16));
    }
    dst[grow * W + x] = make_float2(se * k, sd * k);
  }
}

// Box pass on float2 (E, D) along y or z with clamp-to-edge and the sequential FP32 sum
// order of the reference.  With NORM the z pass also normalizes against img.
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
  for (int e = threadIdx.x; e < span * XW; e += NT) {
    const int p = e / XW, xl = e % XW, x = x0 + xl;
    float2 v = make_float2(0.0f, 0.0f);
    if (x < W) v = src[base + (i64)clampi(c0 - r + p, 0, L - 1) * sa + x];
    smem2[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c >= L || x >= W) continue;
    float ve = 0.0f, vd = 0.0f;
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

hipError_t minmax_line(const void* src, bool from_u16, uint32_t* dst, int W, int L, i64 sa,
                       i64 so, i64 n_other, int r, hipStream_t s) {
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

template <bool NORM, typename TS>
hipError_t box_line(const float2* src, float2* dst, TS* norm, const uint16_t* img, float min_int,
                    int W, int L, i64 sa, i64 so, i64 n_other, int r, hipStream_t s) {
  const LineCfg c = line_cfg(r, L, 8, 1);
  const int nxt = (W + c.XW - 1) / c.XW, nct = (L + c.C - 1) / c.C;
  const i64 nblk = (i64)nxt * nct * n_other;
  const size_t lds = (size_t)(c.C + 2 * r) * c.XW * 8;
  const float k = 1.0f / (float)(2 * r + 1);
  return dispatch_xw(c.XW, [&](auto xw) {
    constexpr int XW = decltype(xw)::value;
    hipLaunchKernelGGL((k_box_line<XW, NORM, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, norm, img, min_int, W, L, sa, so, r, c.C, nxt, nct, k);
    return hipGetLastError();
  });
}

}  // namespace

template <typename TS>
hipError_t launch_normalize(const uint16_t* img, TS* norm, void* tmp1, void* tmp2,
                            float min_intensity, int W, int H, int D, int rxy, int rz,
                            hipStream_t s) {
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
  // min/max along z (lines over x, y fixed per work-group), then along y
  if ((e = minmax_line(img, true, p1, W, D, WH, W, H, rz, s)) != hipSuccess) return e;
  mark("normalize/minmax_z", s);
  if ((e = minmax_line(p1, false, p2, W, H, W, WH, D, rxy, s)) != hipSuccess) return e;
  mark("normalize/minmax_y", s);
  // min/max along x fused with the box x pass -> float2 (E, D) in tmp1
  {
    constexpr int ROWS = 4;
    const int C = 256;
    const int nct = (W + C - 1) / C;
    const i64 nrows = (i64)H * D;
    const i64 nblk = ((nrows + ROWS - 1) / ROWS) * nct;
    const size_t lds = 2 * (size_t)ROWS * (C + 4 * rxy) * 4;
    const float k = 1.0f / (float)(2 * rxy + 1);
    hipLaunchKernelGGL((k_minmax_x_box<ROWS>), dim3((unsigned)nblk), dim3(NT), lds, s, p2,
                       (float2*)tmp1, W, nrows, rxy, C, nct, k);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("normalize/minmax_x_box_x", s);
  }
  // box along y, then along z fused with normalization
  if ((e = box_line<false, TS>((const float2*)tmp1, (float2*)tmp2, nullptr, nullptr, 0.0f, W, H,
                               W, WH, D, rxy, s)) != hipSuccess)
    return e;
  mark("normalize/box_y", s);
  e = box_line<true, TS>((const float2*)tmp2, nullptr, norm, img, min_intensity, W, D, WH, W, H,
                         rz, s);
  mark("normalize/box_z_norm", s);
  return e;
}

template hipError_t launch_normalize<float>(const uint16_t*, float*, void*, void*, float, int,
                                            int, int, int, int, hipStream_t);
template hipError_t launch_normalize<__half>(const uint16_t*, __half*, void*, void*, float, int,
                                             int, int, int, int, hipStream_t);
template hipError_t launch_normalize<__hip_bfloat16>(const uint16_t*, __hip_bfloat16*, void*,
                                                     void*, float, int, int, int, int, int,
                                                     hipStream_t);

}  // namespace hdog
