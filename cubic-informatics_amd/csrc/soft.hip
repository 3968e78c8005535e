// soft.hip -- kernels of the differentiable blob response (no counterpart in the
// reference, which only thresholds the Hessian).
//
//   dog  = s * (G_s1 v - G_s2 v),  s = (sx1 * sy1 * sz1)^gamma, anisotropic sigmas, fixed radii
//   H    = Hessian of dog (the same zero-padded central differences as the detector)
//   M    = candidate mask (all leading minors of H negative), held constant
//   u    = -det(H) / tau,  loss = mean over M of BCE(sigmoid(u), target)
//
// Backward pass (hand-written adjoints):
//   dL/du   = (sigmoid(u) - t) / |M|                    (pointwise, FP64)
//   dL/dH_k = dL/du * (-1/tau) * d det / d H_k           (cofactors)
//   dL/ddog = sum_k A_k^T dL/dH_k, with A_k the pair of first differences producing H_k; the
//             zero-padded central difference g satisfies g^T = -g, so A_k^T applies the same
//             two differences in reverse order:  g_x(P_x) + g_y(P_y) + g_z(P_z) with
//             P_x = g_x Gxx + g_y Gxy + g_z Gxz, P_y = g_y Gyy + g_z Gyz, P_z = g_z Gzz.
//   dL/dv   = s * (G_s1 - G_s2) dL/ddog (a symmetric zero-padded Gaussian is self-adjoint)
//   dL/dsigma: one separable pass with the weight derivative along one axis, dotted with
//             dL/ddog, plus the normalizer term (gamma / sigma_a1) <dL/ddog, dog>.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ double det3(double xx, double xy, double xz, double yy, double yz,
                                       double zz) {
  return xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
}

__device__ __forceinline__ double softplus(double u) {
  return u > 0 ? u + log1p(exp(-u)) : log1p(exp(u));
}

// sums[0] += sum over M of softplus(u) - t*u ; sums[1] += |M|
__global__ __launch_bounds__(NT) void k_soft_loss(const float* __restrict__ hess,
                                                  const uint8_t* __restrict__ mask,
                                                  const uint8_t* __restrict__ target, i64 N,
                                                  double inv_tau, double* sums) {
  double ls = 0.0, cnt = 0.0;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    if (!mask[i]) continue;
    const double d = det3(hess[i], hess[N + i], hess[2 * N + i], hess[3 * N + i],
                          hess[4 * N + i], hess[5 * N + i]);
    const double u = -d * inv_tau;
    ls += softplus(u) - (target[i] ? u : 0.0);
    cnt += 1.0;
  }
  ls = wave_sum(ls);
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(sums, ls);
    unsafeAtomicAdd(sums + 1, cnt);
  }
}

// hess (6N) is overwritten by dL/dH; sums[0] += dL/dtau contributions
__global__ __launch_bounds__(NT) void k_soft_grad_h(float* __restrict__ hess,
                                                    const uint8_t* __restrict__ mask,
                                                    const uint8_t* __restrict__ target, i64 N,
                                                    double inv_tau, double inv_count,
                                                    double* sums) {
  double gt = 0.0;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    if (!mask[i]) {
#pragma unroll
      for (int k = 0; k < 6; ++k) hess[k * N + i] = 0.0f;
      continue;
    }
    const double xx = hess[i], xy = hess[N + i], xz = hess[2 * N + i], yy = hess[3 * N + i],
                 yz = hess[4 * N + i], zz = hess[5 * N + i];
    const double d = det3(xx, xy, xz, yy, yz, zz);
    const double u = -d * inv_tau;
    const double p = 1.0 / (1.0 + exp(-u));
    const double gu = (p - (target[i] ? 1.0 : 0.0)) * inv_count;  // dL/du
    gt += gu * d * inv_tau * inv_tau;                               // du/dtau = det / tau^2
    const double gd = -gu * inv_tau;                                // dL/ddet
    hess[i] = (float)(gd * (yy * zz - yz * yz));
    hess[N + i] = (float)(gd * 2.0 * (xz * yz - zz * xy));
    hess[2 * N + i] = (float)(gd * 2.0 * (xy * yz - yy * xz));
    hess[3 * N + i] = (float)(gd * (xx * zz - xz * xz));
    hess[4 * N + i] = (float)(gd * 2.0 * (xy * xz - xx * yz));
    hess[5 * N + i] = (float)(gd * (xx * yy - xy * xy));
  }
  gt = wave_sum(gt);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(sums, gt);
}

// zero-padded central difference of field f along axis `a` at (x, y, z)
__device__ __forceinline__ float cdiff(const float* f, int x, int y, int z, int a, int W, int H,
                                       int D, i64 WH) {
  int xp = x, xm = x, yp = y, ym = y, zp = z, zm = z;
  bool okp, okm;
  if (a == 0) { xp = x + 1; xm = x - 1; okp = xp < W; okm = xm >= 0; }
  else if (a == 1) { yp = y + 1; ym = y - 1; okp = yp < H; okm = ym >= 0; }
  else { zp = z + 1; zm = z - 1; okp = zp < D; okm = zm >= 0; }
  const float hi = okp ? f[(i64)zp * WH + (i64)yp * W + xp] : 0.0f;
  const float lo = okm ? f[(i64)zm * WH + (i64)ym * W + xm] : 0.0f;
  return (hi - lo) * 0.5f;
}

__global__ __launch_bounds__(NT) void k_hadj1(const float* __restrict__ g, float* __restrict__ P,
                                              int W, int H, int D) {
  const i64 WH = (i64)W * H, N = WH * D;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / WH);
    const float* gxx = g; const float* gxy = g + N; const float* gxz = g + 2 * N;
    const float* gyy = g + 3 * N; const float* gyz = g + 4 * N; const float* gzz = g + 5 * N;
    P[i] = cdiff(gxx, x, y, z, 0, W, H, D, WH) + cdiff(gxy, x, y, z, 1, W, H, D, WH) +
           cdiff(gxz, x, y, z, 2, W, H, D, WH);
    P[N + i] = cdiff(gyy, x, y, z, 1, W, H, D, WH) + cdiff(gyz, x, y, z, 2, W, H, D, WH);
    P[2 * N + i] = cdiff(gzz, x, y, z, 2, W, H, D, WH);
  }
}

__global__ __launch_bounds__(NT) void k_hadj2(const float* __restrict__ P, float* __restrict__ out,
                                              int W, int H, int D) {
  const i64 WH = (i64)W * H, N = WH * D;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / WH);
    out[i] = cdiff(P, x, y, z, 0, W, H, D, WH) + cdiff(P + N, x, y, z, 1, W, H, D, WH) +
             cdiff(P + 2 * N, x, y, z, 2, W, H, D, WH);
  }
}

// out[0] += <a, pair.x>, out[1] += <a, pair.y>
__global__ __launch_bounds__(NT) void k_dot_pair(const float* __restrict__ a,
                                                 const float2* __restrict__ pair, i64 N,
                                                 double* out) {
  double s0 = 0.0, s1 = 0.0;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const double v = a[i];
    const float2 p = pair[i];
    s0 += v * p.x;
    s1 += v * p.y;
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(out, s0);
    unsafeAtomicAdd(out + 1, s1);
  }
}

__global__ __launch_bounds__(NT) void k_dot(const float* __restrict__ a,
                                            const float* __restrict__ b, i64 N, double* out) {
  double s = 0.0;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT)
    s += (double)a[i] * b[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(out, s);
}

__global__ void k_zero_f64(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

unsigned grid_for(i64 N) { return (unsigned)std::min<i64>((N + NT - 1) / NT, 1 << 16); }

}  // namespace

hipError_t launch_soft_loss(const float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                            double inv_tau, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, sums, 2);
  hipLaunchKernelGGL(k_soft_loss, dim3(grid_for(N)), dim3(NT), 0, s, hess, mask, target, N, inv_tau,
                     sums);
  return hipGetLastError();
}

hipError_t launch_soft_grad_h(float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                              double inv_tau, double inv_count, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, sums, 1);
  hipLaunchKernelGGL(k_soft_grad_h, dim3(grid_for(N)), dim3(NT), 0, s, hess, mask, target, N,
                     inv_tau, inv_count, sums);
  return hipGetLastError();
}

hipError_t launch_hessian_adjoint(const float* gh, float* P, float* out, int W, int H, int D,
                                  hipStream_t s) {
  const i64 N = (i64)W * H * D;
  hipLaunchKernelGGL(k_hadj1, dim3(grid_for(N)), dim3(NT), 0, s, gh, P, W, H, D);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_hadj2, dim3(grid_for(N)), dim3(NT), 0, s, P, out, W, H, D);
  return hipGetLastError();
}

hipError_t launch_dot_pair(const float* a, const float2* pair, i64 N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, out, 2);
  hipLaunchKernelGGL(k_dot_pair, dim3(grid_for(N)), dim3(NT), 0, s, a, pair, N, out);
  return hipGetLastError();
}

hipError_t launch_dot(const float* a, const float* b, i64 N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, out, 1);
  hipLaunchKernelGGL(k_dot, dim3(grid_for(N)), dim3(NT), 0, s, a, b, N, out);
  return hipGetLastError();
}

}  // namespace hdog
