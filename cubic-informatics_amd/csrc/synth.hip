// synth.hip -- seeded synthetic stacks rendered on the device (benchmark input).
//
// Same model as hdog/synth.py: counter-based normal noise (splitmix64 of the voxel
// index, Box-Muller in FP64), background mean linear in z, clipped at 0, plus Gaussian
// blobs evaluated within 4 sigma of their centre, then rint and clip to uint16.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ u64 splitmix64(u64 z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit53(u64 h) {
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(NT) void k_noise(float* __restrict__ out, i64 N, i64 WH, int D,
                                              u64 key, double mean0, double mean1, double sd) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const double u1 = unit53(splitmix64(((u64)i * 2ull) ^ key));
    const double u2 = unit53(splitmix64(((u64)i * 2ull + 1ull) ^ key));
    const double n = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    const i64 z = i / WH;
    const double mean = mean0 + (mean1 - mean0) * (double)z / (double)(D > 1 ? D - 1 : 1);
    out[i] = (float)fmax(mean + sd * n, 0.0);
  }
}

__global__ __launch_bounds__(NT) void k_blobs(float* __restrict__ out, int W, int H, int D,
                                              const double* __restrict__ blobs) {
  const double* b = blobs + 7 * (i64)blockIdx.x;
  const double cz = b[0], cy = b[1], cx = b[2], sz = b[3], sy = b[4], sx = b[5], A = b[6];
  const int z0 = max(0, (int)floor(cz - 4 * sz)), z1 = min(D, (int)ceil(cz + 4 * sz) + 1);
  const int y0 = max(0, (int)floor(cy - 4 * sy)), y1 = min(H, (int)ceil(cy + 4 * sy) + 1);
  const int x0 = max(0, (int)floor(cx - 4 * sx)), x1 = min(W, (int)ceil(cx + 4 * sx) + 1);
  const int nx = x1 - x0, ny = y1 - y0, nz = z1 - z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const int tot = nx * ny * nz;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int x = x0 + e % nx, t = e / nx, y = y0 + t % ny, z = z0 + t / ny;
    const double g = exp(-(z - cz) * (z - cz) / (2 * sz * sz)) *
                     exp(-(y - cy) * (y - cy) / (2 * sy * sy)) *
                     exp(-(x - cx) * (x - cx) / (2 * sx * sx));
    unsafeAtomicAdd(out + ((i64)z * H + y) * W + x, (float)(A * g));
  }
}

__global__ __launch_bounds__(NT) void k_quantize(const float* __restrict__ v,
                                                 uint16_t* __restrict__ img, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT)
    img[i] = (uint16_t)fminf(fmaxf(rintf(v[i]), 0.0f), 65535.0f);
}

}  // namespace

hipError_t launch_synth(uint16_t* img, float* scratch, int W, int H, int D, u64 seed, double mean0,
                        double mean1, double sd, const double* d_blobs, i64 n_blobs,
                        hipStream_t s) {
  const i64 WH = (i64)W * H, N = WH * D;
  const u64 key = seed * 0x632BE59BD9B4E019ull;
  const i64 nb = std::min<i64>((N + NT - 1) / NT, 1 << 20);
  hipLaunchKernelGGL(k_noise, dim3((unsigned)nb), dim3(NT), 0, s, scratch, N, WH, D, key, mean0,
                     mean1, sd);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_blobs > 0) {
    hipLaunchKernelGGL(k_blobs, dim3((unsigned)n_blobs), dim3(NT), 0, s, scratch, W, H, D, d_blobs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_quantize, dim3((unsigned)nb), dim3(NT), 0, s, scratch, img, N);
  return hipGetLastError();
}

}  // namespace hdog
