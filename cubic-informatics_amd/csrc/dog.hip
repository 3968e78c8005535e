// dog.hip -- separable anisotropic difference of Gaussians on gfx950.
//
// Reference semantics: src/3D/DoG3DFilter.cu:70-437.  Each Gaussian is x -> y (xy
// weights) -> z (z weights) with zero padding; every output is the FMA chain
// sum = fmaf(w[r-j], s[c+j], sum), j = -r..r, from 0.  dog = (G1 - G2) * normalizer.
//
// MI355X design: both Gaussians are evaluated from one read of each tile.  The x pass
// reads `norm` once and writes (G1_x, G2_x) as float2, the y pass works on the pair, and
// the z pass forms the difference directly, so `norm` is read once and the two
// intermediate volumes of the reference become one interleaved stream.  Intermediates
// stay FP32 whatever the storage type of norm / dog.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

template <int ROWS, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_x(const TS* __restrict__ src, float2* __restrict__ dst,
                                                const float* __restrict__ w1,
                                                const float* __restrict__ w2, int r1, int r2,
                                                int W, i64 nrows, int C, int nct) {
  extern __shared__ float smf[];
  const int R = max(r1, r2);
  const int span = C + 2 * R;
  const int ct = blockIdx.x % nct;
  const i64 row0 = (i64)(blockIdx.x / nct) * ROWS;
  const int c0 = ct * C;
  for (int e = threadIdx.x; e < ROWS * span; e += NT) {
    const int row = e / span, p = e % span;
    const i64 grow = row0 + row;
    const int x = c0 - R + p;
    float v = 0.0f;
    if (grow < nrows && x >= 0 && x < W) v = ld(src, grow * W + x);
    smf[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ROWS * C; e += NT) {
    const int row = e / C, i = e % C, x = c0 + i;
    const i64 grow = row0 + row;
    if (x >= W || grow >= nrows) continue;
    const float* t = smf + row * span + i + R;
    float s1 = 0.0f, s2 = 0.0f;
    for (int j = -r1; j <= r1; ++j) s1 = fmaf(w1[r1 - j], t[j], s1);
    for (int j = -r2; j <= r2; ++j) s2 = fmaf(w2[r2 - j], t[j], s2);
    dst[grow * W + x] = make_float2(s1, s2);
  }
}

// y or z pass on the (G1, G2) pair with zero padding; LAST forms the scaled difference.
template <int XW, bool LAST, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_line(const float2* __restrict__ src,
                                                   float2* __restrict__ dst, TS* __restrict__ dog,
                                                   const float* __restrict__ w1,
                                                   const float* __restrict__ w2, int r1, int r2,
                                                   float nz, int W, int L, i64 sa, i64 so, int C,
                                                   int nxt, int nct) {
  extern __shared__ float2 smf2[];
  const int R = max(r1, r2);
  const int span = C + 2 * R;
  unsigned b = blockIdx.x;
  const int xt = b % nxt; b /= nxt;
  const int ct = b % nct; b /= nct;
  const i64 base = (i64)b * so;
  const int x0 = xt * XW, c0 = ct * C;
  for (int e = threadIdx.x; e < span * XW; e += NT) {
    const int p = e / XW, xl = e % XW, x = x0 + xl, c = c0 - R + p;
    float2 v = make_float2(0.0f, 0.0f);
    if (x < W && c >= 0 && c < L) v = src[base + (i64)c * sa + x];
    smf2[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c >= L || x >= W) continue;
    const float2* t = smf2 + (i + R) * XW + xl;
    float s1 = 0.0f, s2 = 0.0f;
    for (int j = -r1; j <= r1; ++j) s1 = fmaf(w1[r1 - j], t[j * XW].x, s1);
    for (int j = -r2; j <= r2; ++j) s2 = fmaf(w2[r2 - j], t[j * XW].y, s2);
    const i64 idx = base + (i64)c * sa + x;
    if (LAST)
      dog[idx] = cvt<TS>((s1 - s2) * nz);
    else
      dst[idx] = make_float2(s1, s2);
  }
}

template <bool LAST, typename TS>
hipError_t gauss_line(const float2* src, float2* dst, TS* dog, const float* w1, const float* w2,
                      int r1, int r2, float nz, int W, int L, i64 sa, i64 so, i64 n_other,
                      hipStream_t s) {
  const int R = std::max(r1, r2);
  int XW = R > 24 ? 32 : 64;
  int C = 64;
  if (L < C) C = L;
  const int nxt = (W + XW - 1) / XW, nct = (L + C - 1) / C;
  const i64 nblk = (i64)nxt * nct * n_other;
  const size_t lds = (size_t)(C + 2 * R) * XW * 8;
  if (XW == 64)
    hipLaunchKernelGGL((k_gauss_line<64, LAST, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, dog, w1, w2, r1, r2, nz, W, L, sa, so, C, nxt, nct);
  else
    hipLaunchKernelGGL((k_gauss_line<32, LAST, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, dog, w1, w2, r1, r2, nz, W, L, sa, so, C, nxt, nct);
  return hipGetLastError();
}

}  // namespace

template <typename TS>
hipError_t launch_dog(const TS* norm, TS* dog, void* tmp1, void* tmp2, const float* weights,
                      const int r[4], float normalizer, int W, int H, int D, hipStream_t s) {
  const i64 WH = (i64)W * H;
  const float* wxy1 = weights;
  const float* wz1 = weights + kMaxTaps;
  const float* wxy2 = weights + 2 * kMaxTaps;
  const float* wz2 = weights + 3 * kMaxTaps;
  hipError_t e;
  {
    constexpr int ROWS = 4;
    const int C = 256;
    const int nct = (W + C - 1) / C;
    const i64 nrows = (i64)H * D;
    const i64 nblk = ((nrows + ROWS - 1) / ROWS) * nct;
    const size_t lds = (size_t)ROWS * (C + 2 * std::max(r[0], r[2])) * 4;
    hipLaunchKernelGGL((k_gauss_x<ROWS, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, norm,
                       (float2*)tmp1, wxy1, wxy2, r[0], r[2], W, nrows, C, nct);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("dog/gauss_x", s);
  }
  if ((e = gauss_line<false, TS>((const float2*)tmp1, (float2*)tmp2, nullptr, wxy1, wxy2, r[0],
                                 r[2], 0.0f, W, H, W, WH, D, s)) != hipSuccess)
    return e;
  mark("dog/gauss_y", s);
  e = gauss_line<true, TS>((const float2*)tmp2, nullptr, dog, wz1, wz2, r[1], r[3], normalizer, W,
                           D, WH, W, H, s);
  mark("dog/gauss_z_dog", s);
  return e;
}

template hipError_t launch_dog<float>(const float*, float*, void*, void*, const float*,
                                      const int*, float, int, int, int, hipStream_t);
template hipError_t launch_dog<__half>(const __half*, __half*, void*, void*, const float*,
                                       const int*, float, int, int, int, hipStream_t);
template hipError_t launch_dog<__hip_bfloat16>(const __hip_bfloat16*, __hip_bfloat16*, void*,
                                               void*, const float*, const int*, float, int, int,
                                               int, hipStream_t);

}  // namespace hdog
