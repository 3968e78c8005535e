// hessian.hip -- Hessian elements and the negative-definite candidate mask, one kernel.
//
// Reference semantics: src/3D/Hessian3DFilter_element.cu:22-388.  Nine central-difference
// passes g_a(f)[c] = ((0 + f[c+1]) - f[c-1]) * 0.5f, each zero-padded at the volume
// border: Hxx = g_x(g_x d), Hxy = g_y(g_x d), Hxz = g_z(g_x d), Hyy = g_y(g_y d),
// Hyz = g_z(g_y d), Hzz = g_z(g_z d).  Candidate: xx < 0, xx*yy - xy*xy > 0 and the 3x3
// determinant < 0, evaluated in C parse order without contraction (the library is
// built with -ffp-contract=off).
//
// MI355X design: one pass over `dog` instead of nine.  A work-group stages a
// (64+4) x (4+4) x (4+4) tile of dog in LDS, forms the first derivatives Hx, Hy, Hz on
// the halo positions the second pass needs (zero outside the volume, as the reference's
// second pass pads its own input), then writes the six elements (SoA, pitch N) and the
// mask with coalesced stores.
//
// Reduced-precision storage (FP16 / BF16): the elements are computed in FP32 from the
// stored DoG and rounded only on the store; the candidate test uses the FP32 values, and
// when a leading minor lies within 1e-3 of the sum of the magnitudes of its terms, the
// minors are re-evaluated in FP64 from the same FP32 elements.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int TX = 64, TY = 4, TZ = 4;
constexpr int NT = 256;

__device__ __forceinline__ float g2(float hi, float lo) {
  float s = 0.0f;
  s += hi;
  s -= lo;
  s *= 0.5f;
  return s;
}

__device__ __forceinline__ uint8_t negdef32(float xx, float xy, float xz, float yy, float yz,
                                            float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

__device__ __forceinline__ uint8_t negdef_band(float xx, float xy, float xz, float yy, float yz,
                                               float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  const float s2 = fabsf(xx * yy) + fabsf(xy * xy);
  const float s3 = fabsf(xx * yy * zz) + fabsf(2 * xy * yz * xz) + fabsf(xx * yz * yz) +
                   fabsf(yy * xz * xz) + fabsf(zz * xy * xy);
  if (fabsf(m2) < 1e-3f * s2 || fabsf(m3) < 1e-3f * s3) {
    const double a = xx, b = xy, c = xz, d = yy, e = yz, f = zz;
    const double n2 = a * d - b * b;
    const double n3 = a * d * f + 2 * b * e * c - a * e * e - d * c * c - f * b * b;
    return (a < 0 && n2 > 0 && n3 < 0) ? 1 : 0;
  }
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

template <typename TS, bool BAND>
__global__ __launch_bounds__(NT) void k_hessian(const TS* __restrict__ dog, TS* __restrict__ hess,
                                                uint8_t* __restrict__ pd, int W, int H, int D,
                                                int nbx, int nby) {
  __shared__ float sD[TZ + 4][TY + 4][TX + 4];
  __shared__ float sHx[TZ + 2][TY + 2][TX + 2];
  __shared__ float sHy[TZ + 2][TY + 2][TX];
  __shared__ float sHz[TZ + 2][TY][TX];
  unsigned b = blockIdx.x;
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x0 = bx * TX, y0 = by * TY, z0 = (int)b * TZ;
  const i64 WH = (i64)W * H;
  const i64 N = WH * D;

  constexpr int ND = (TZ + 4) * (TY + 4) * (TX + 4);
  for (int e = threadIdx.x; e < ND; e += NT) {
    const int lx = e % (TX + 4), t = e / (TX + 4), ly = t % (TY + 4), lz = t / (TY + 4);
    const int gx = x0 - 2 + lx, gy = y0 - 2 + ly, gz = z0 - 2 + lz;
    float v = 0.0f;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
      v = ld(dog, gz * WH + (i64)gy * W + gx);
    sD[lz][ly][lx] = v;
  }
  __syncthreads();
  // first derivatives on the halo positions; zero outside the volume
  constexpr int NHX = (TZ + 2) * (TY + 2) * (TX + 2);
  for (int e = threadIdx.x; e < NHX; e += NT) {
    const int lx = e % (TX + 2), t = e / (TX + 2), ly = t % (TY + 2), lz = t / (TY + 2);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
    float v = 0.0f;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
      v = g2(sD[lz + 1][ly + 1][lx + 2], sD[lz + 1][ly + 1][lx]);
    sHx[lz][ly][lx] = v;
  }
  constexpr int NHY = (TZ + 2) * (TY + 2) * TX;
  for (int e = threadIdx.x; e < NHY; e += NT) {
    const int lx = e % TX, t = e / TX, ly = t % (TY + 2), lz = t / (TY + 2);
    const int gx = x0 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
    float v = 0.0f;
    if (gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
      v = g2(sD[lz + 1][ly + 2][lx + 2], sD[lz + 1][ly][lx + 2]);
    sHy[lz][ly][lx] = v;
  }
  constexpr int NHZ = (TZ + 2) * TY * TX;
  for (int e = threadIdx.x; e < NHZ; e += NT) {
    const int lx = e % TX, t = e / TX, ly = t % TY, lz = t / TY;
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 - 1 + lz;
    float v = 0.0f;
    if (gx < W && gy < H && gz >= 0 && gz < D)
      v = g2(sD[lz + 2][ly + 2][lx + 2], sD[lz][ly + 2][lx + 2]);
    sHz[lz][ly][lx] = v;
  }
  __syncthreads();
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gx = x0 + tx, gy = y0 + ty;
  if (gx >= W || gy >= H) return;
#pragma unroll
  for (int tz = 0; tz < TZ; ++tz) {
    const int gz = z0 + tz;
    if (gz >= D) break;
    const float xx = g2(sHx[tz + 1][ty + 1][tx + 2], sHx[tz + 1][ty + 1][tx]);
    const float xy = g2(sHx[tz + 1][ty + 2][tx + 1], sHx[tz + 1][ty][tx + 1]);
    const float xz = g2(sHx[tz + 2][ty + 1][tx + 1], sHx[tz][ty + 1][tx + 1]);
    const float yy = g2(sHy[tz + 1][ty + 2][tx], sHy[tz + 1][ty][tx]);
    const float yz = g2(sHy[tz + 2][ty + 1][tx], sHy[tz][ty + 1][tx]);
    const float zz = g2(sHz[tz + 2][ty][tx], sHz[tz][ty][tx]);
    const i64 idx = gz * WH + (i64)gy * W + gx;
    hess[idx] = cvt<TS>(xx);
    hess[N + idx] = cvt<TS>(xy);
    hess[2 * N + idx] = cvt<TS>(xz);
    hess[3 * N + idx] = cvt<TS>(yy);
    hess[4 * N + idx] = cvt<TS>(yz);
    hess[5 * N + idx] = cvt<TS>(zz);
    pd[idx] = BAND ? negdef_band(xx, xy, xz, yy, yz, zz) : negdef32(xx, xy, xz, yy, yz, zz);
  }
}

}  // namespace

template <typename TS>
hipError_t launch_hessian(const TS* dog, TS* hess, uint8_t* pd, int W, int H, int D,
                          hipStream_t s) {
  const int nbx = (W + TX - 1) / TX, nby = (H + TY - 1) / TY, nbz = (D + TZ - 1) / TZ;
  const i64 nblk = (i64)nbx * nby * nbz;
  constexpr bool band = !std::is_same<TS, float>::value;
  hipLaunchKernelGGL((k_hessian<TS, band>), dim3((unsigned)nblk), dim3(NT), 0, s, dog, hess, pd,
                     W, H, D, nbx, nby);
  mark("hessian/hessian_pd", s);
  return hipGetLastError();
}

template hipError_t launch_hessian<float>(const float*, float*, uint8_t*, int, int, int,
                                          hipStream_t);
template hipError_t launch_hessian<__half>(const __half*, __half*, uint8_t*, int, int, int,
                                           hipStream_t);
template hipError_t launch_hessian<__hip_bfloat16>(const __hip_bfloat16*, __hip_bfloat16*,
                                                   uint8_t*, int, int, int, hipStream_t);

}  // namespace hdog
