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
//
// Local-contrast normalization in front of the DoG (src/3D/NormalizeFilter.cu:53-654 on a
// float volume v): E = box(min3(v)), D = box(max3(v)) with clamp-to-edge windows of
// radius (rxy, rxy, rz) for both the min / max and the box, n = v >= m ? (v - E)/(D - E)
// : 0.  Its adjoint, given dL/dn:
//   dL/dv  = [v >= m] dL/dn / (D - E)                          (direct term)
//          + route(argmin, box^T(dL/dE)) + route(argmax, box^T(dL/dD))
//   dL/dE  = [v >= m] dL/dn (v - D) / (D - E)^2,  dL/dD = -[v >= m] dL/dn (v - E) / (D - E)^2
// box^T is the transpose of the clamped box mean (the clamped border positions collect
// the weight of every window tap that clamps onto them); the min filter is taken by its
// subgradient: each output's gradient goes to the voxel that attains the window minimum
// (the lowest linear index among equal minima), likewise for the max.  The argmin / argmax
// of the 3-D window are found by three separable passes over (value, index) pairs, which
// is exact because the order "smaller value, then smaller index" is total.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Adds the work-group's sums of a[k] to out[k] (K <= 2) with one FP64 atomic each: the
// reductions run on a grid of at most kReduceGroups work-groups, so a handful of atomics
// meet at one address (one atomic per wave over a grid sized for the volume serialized
// hundreds of thousands of them there).
constexpr int kReduceGroups = 1024;
template <int K>
__device__ __forceinline__ void group_add(double (&a)[K], double* out) {
  __shared__ double part[NT / 64][K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = wave_sum(a[k]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[threadIdx.x >> 6][k] = a[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t += part[w][threadIdx.x];
    unsafeAtomicAdd(out + threadIdx.x, t);
  }
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
  double a[2] = {ls, cnt};
  group_add<2>(a, sums);
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
  double a[1] = {gt};
  group_add<1>(a, sums);
}

// Line kernels run one element per thread on a grid of nxb x-blocks per row (y + H z):
// x, y, z come from the thread and the block with 32-bit arithmetic, not from divisions
// of a 64-bit linear index.
struct Pos {
  int x, y, z;
  i64 i;  // linear index
};
__device__ __forceinline__ bool pos_of(int W, int H, unsigned nxb, Pos& p) {
  const unsigned row = blockIdx.x / nxb;
  p.x = (int)(blockIdx.x - row * nxb) * NT + (int)threadIdx.x;
  p.y = (int)(row % (unsigned)H);
  p.z = (int)(row / (unsigned)H);
  p.i = (i64)row * W + p.x;
  return p.x < W;
}
unsigned xblocks(int W) { return (unsigned)((W + NT - 1) / NT); }
unsigned grid_rows(int W, int H, int D) { return xblocks(W) * (unsigned)H * (unsigned)D; }

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
                                              int W, int H, int D, unsigned nxb) {
  const i64 WH = (i64)W * H, N = WH * D;
  Pos q;
  if (!pos_of(W, H, nxb, q)) return;
  {
    const i64 i = q.i;
    const int x = q.x, y = q.y, z = q.z;
    const float* gxx = g; const float* gxy = g + N; const float* gxz = g + 2 * N;
    const float* gyy = g + 3 * N; const float* gyz = g + 4 * N; const float* gzz = g + 5 * N;
    P[i] = cdiff(gxx, x, y, z, 0, W, H, D, WH) + cdiff(gxy, x, y, z, 1, W, H, D, WH) +
           cdiff(gxz, x, y, z, 2, W, H, D, WH);
    P[N + i] = cdiff(gyy, x, y, z, 1, W, H, D, WH) + cdiff(gyz, x, y, z, 2, W, H, D, WH);
    P[2 * N + i] = cdiff(gzz, x, y, z, 2, W, H, D, WH);
  }
}

__global__ __launch_bounds__(NT) void k_hadj2(const float* __restrict__ P, float* __restrict__ out,
                                              int W, int H, int D, unsigned nxb) {
  const i64 WH = (i64)W * H, N = WH * D;
  Pos q;
  if (!pos_of(W, H, nxb, q)) return;
  {
    const i64 i = q.i;
    const int x = q.x, y = q.y, z = q.z;
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
  double t[2] = {s0, s1};
  group_add<2>(t, out);
}

__global__ __launch_bounds__(NT) void k_dot(const float* __restrict__ a,
                                            const float* __restrict__ b, i64 N, double* out) {
  double s = 0.0;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT)
    s += (double)a[i] * b[i];
  double t[1] = {s};
  group_add<1>(t, out);
}

__global__ __launch_bounds__(NT) void k_pack_pair(const float* __restrict__ a,
                                                  const float* __restrict__ b,
                                                  float2* __restrict__ out, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT)
    out[i] = make_float2(a[i], b[i]);
}

__global__ void k_zero_f64(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

unsigned grid_for(i64 N) { return (unsigned)std::min<i64>((N + NT - 1) / NT, 1 << 16); }
unsigned grid_reduce(i64 N) {
  return (unsigned)std::max<i64>(1, std::min<i64>((N + NT - 1) / NT, kReduceGroups));
}

struct Line {  // the line through a voxel along one axis
  i64 base, stride;
  int c, L;
};
__device__ __forceinline__ Line line_at(const Pos& p, int axis, int W, int H, int D) {
  Line l;
  if (axis == 0) { l.c = p.x; l.L = W; l.stride = 1; }
  else if (axis == 1) { l.c = p.y; l.L = H; l.stride = W; }
  else { l.c = p.z; l.L = D; l.stride = (i64)W * H; }
  l.base = p.i - (i64)l.c * l.stride;
  return l;
}

// one separable pass of the windowed (value, index) minimum and maximum, clamp-to-edge;
// first pass: indices are the voxels' own (imn == NULL)
__global__ __launch_bounds__(NT) void k_soft_argminmax(const float* __restrict__ vmn,
                                                       const uint32_t* __restrict__ imn,
                                                       const float* __restrict__ vmx,
                                                       const uint32_t* __restrict__ imx,
                                                       float* __restrict__ omn, uint32_t* __restrict__ oimn,
                                                       float* __restrict__ omx, uint32_t* __restrict__ oimx,
                                                       int W, int H, int D, int axis, int r,
                                                       unsigned nxb) {
  Pos q;
  if (!pos_of(W, H, nxb, q)) return;
  const i64 i = q.i;
  {
    const Line l = line_at(q, axis, W, H, D);
    float bmn = INFINITY, bmx = -INFINITY;
    uint32_t jmn = 0xFFFFFFFFu, jmx = 0xFFFFFFFFu;
    for (int t = -r; t <= r; ++t) {
      const i64 p = l.base + (i64)clampi(l.c + t, 0, l.L - 1) * l.stride;
      const float a = vmn[p], b = vmx[p];
      const uint32_t ja = imn ? imn[p] : (uint32_t)p, jb = imx ? imx[p] : (uint32_t)p;
      if (a < bmn || (a == bmn && ja < jmn)) { bmn = a; jmn = ja; }
      if (b > bmx || (b == bmx && jb < jmx)) { bmx = b; jmx = jb; }
    }
    omn[i] = bmn; oimn[i] = jmn;
    omx[i] = bmx; oimx[i] = jmx;
  }
}

// The x passes with the work-group's row segment x0 - r .. x0 + NT - 1 + r staged in LDS
// (clamped to the row): the same taps, in the same order, as the line kernels, each read
// from LDS instead of global memory.  For radii up to kXR; others use the line kernels.
constexpr int kXR = 64;

// stages the clamped row segment of the work-group; returns the row's linear base, x0
template <typename T>
__device__ __forceinline__ i64 stage_row(const T* __restrict__ in, T* sv, int W, int r,
                                         unsigned nxb, int& x0) {
  const unsigned row = blockIdx.x / nxb;
  x0 = (int)(blockIdx.x - row * nxb) * NT;
  const i64 base = (i64)row * W;
  for (int e = (int)threadIdx.x; e < NT + 2 * r; e += NT) sv[e] = in[base + clampi(x0 - r + e, 0, W - 1)];
  __syncthreads();
  return base;
}

// first (x) pass of the windowed (value, index) minimum and maximum: value v[p], index p
__global__ __launch_bounds__(NT) void k_soft_argminmax_x(const float* __restrict__ v,
                                                         float* __restrict__ omn,
                                                         uint32_t* __restrict__ oimn,
                                                         float* __restrict__ omx,
                                                         uint32_t* __restrict__ oimx, int W,
                                                         int r, unsigned nxb) {
  __shared__ float sv[NT + 2 * kXR];
  int x0;
  const i64 base = stage_row(v, sv, W, r, nxb, x0);
  const int x = x0 + (int)threadIdx.x;
  if (x >= W) return;
  float bmn = INFINITY, bmx = -INFINITY;
  uint32_t jmn = 0xFFFFFFFFu, jmx = 0xFFFFFFFFu;
  for (int t = -r; t <= r; ++t) {
    const int c = clampi(x + t, 0, W - 1);
    const float a = sv[x + t - x0 + r];  // the value at c
    const uint32_t j = (uint32_t)(base + c);
    if (a < bmn || (a == bmn && j < jmn)) { bmn = a; jmn = j; }
    if (a > bmx || (a == bmx && j < jmx)) { bmx = a; jmx = j; }
  }
  omn[base + x] = bmn; oimn[base + x] = jmn;
  omx[base + x] = bmx; oimx[base + x] = jmx;
}

// clamped box mean of the (E, D) pair along x (FP32 sums in window order)
__global__ __launch_bounds__(NT) void k_soft_box_x(const float2* __restrict__ in,
                                                   float2* __restrict__ out, int W, int r,
                                                   unsigned nxb) {
  __shared__ float2 sv[NT + 2 * kXR];
  const float k = 1.0f / (float)(2 * r + 1);
  int x0;
  const i64 base = stage_row(in, sv, W, r, nxb, x0);
  const int x = x0 + (int)threadIdx.x;
  if (x >= W) return;
  float se = 0.0f, sd = 0.0f;
  for (int t = -r; t <= r; ++t) {
    const float2 v = sv[x + t - x0 + r];
    se += v.x;
    sd += v.y;
  }
  out[base + x] = make_float2(se * k, sd * k);
}

// transpose of the clamped box mean along x (k_soft_box_adj's sums, from LDS)
__global__ __launch_bounds__(NT) void k_soft_box_adj_x(const float2* __restrict__ in,
                                                       float2* __restrict__ out, int W, int r,
                                                       unsigned nxb) {
  __shared__ float2 sv[NT + 2 * kXR];
  const float k = 1.0f / (float)(2 * r + 1);
  int x0;
  const i64 base = stage_row(in, sv, W, r, nxb, x0);
  const int p = x0 + (int)threadIdx.x, L = W;
  if (p >= W) return;
  const int o = r - x0;  // LDS slot of position c: c + o
  float se = 0.0f, sd = 0.0f;
  for (int c = max(0, p - r); c <= min(L - 1, p + r); ++c) {
    const float2 v = sv[c + o];
    se += v.x;
    sd += v.y;
  }
  if (p == 0)
    for (int c = 0; c <= min(L - 1, r - 1); ++c) {
      const float2 v = sv[c + o];
      se += v.x * (float)(r - c);
      sd += v.y * (float)(r - c);
    }
  if (p == L - 1)
    for (int c = max(0, L - r); c <= L - 1; ++c) {
      const float2 v = sv[c + o];
      se += v.x * (float)(c + r - L + 1);
      sd += v.y * (float)(c + r - L + 1);
    }
  out[base + p] = make_float2(se * k, sd * k);
}

// The same pass along y or z (second and third passes, indices given) as a column walker
// for radius R: one thread per line (x-consecutive threads), van Herk blocks of 2R+1 over
// the clamp-extended line g[k] = f[clamp(k - R)] held in place in registers (slot j keeps
// the previous block's suffix extreme until the current block's j-th element takes it; a
// running prefix extreme covers the current block), so every element is read once instead
// of 2R+1 times.  Output c leaves when g[c + 2R] has arrived.  (value, index) pairs under
// "smaller value, then smaller index" (min) and "larger value, then smaller index" (max)
// form total orders, so the blocks give the window extremes exactly.
struct Ext {
  float vn, vx;       // min value, max value
  uint32_t jn, jx;    // their indices
};
// (by value: a select between two elements' fields must not become a select between
// their addresses, which keeps the register arrays in scratch memory)
__device__ __forceinline__ Ext ext_join(Ext a, Ext b) {
  const float an = a.vn, bn = b.vn, ax = a.vx, bx = b.vx;
  const uint32_t ajn = a.jn, bjn = b.jn, ajx = a.jx, bjx = b.jx;
  const bool mn = an < bn || (an == bn && ajn < bjn);
  const bool mx = ax > bx || (ax == bx && ajx < bjx);
  Ext o;
  o.vn = mn ? an : bn;
  o.jn = mn ? ajn : bjn;
  o.vx = mx ? ax : bx;
  o.jx = mx ? ajx : bjx;
  return o;
}
template <int R>
__global__ __launch_bounds__(NT) void k_soft_argminmax_walk(
    const float* __restrict__ vmn, const uint32_t* __restrict__ imn,
    const float* __restrict__ vmx, const uint32_t* __restrict__ imx, float* __restrict__ omn,
    uint32_t* __restrict__ oimn, float* __restrict__ omx, uint32_t* __restrict__ oimx, int W,
    int L, i64 sa, i64 so, unsigned nxb) {
  constexpr int WB = 2 * R + 1;
  constexpr int PF = WB % 3 == 0 ? 3 : 1;  // loads queued PF steps ahead (static slots)
  const int x = (int)(blockIdx.x % nxb) * NT + (int)threadIdx.x;
  if (x >= W) return;
  const i64 base = (i64)(blockIdx.x / nxb) * so + x;
  auto load = [&](int k) -> Ext {
    const i64 p = base + (i64)clampi(k - R, 0, L - 1) * sa;
    Ext e;
    e.vn = vmn[p];
    e.jn = imn[p];
    e.vx = vmx[p];
    e.jx = imx[p];
    return e;
  };
  auto put = [&](int c, const Ext& e) {
    const i64 p = base + (i64)c * sa;
    omn[p] = e.vn;
    oimn[p] = e.jn;
    omx[p] = e.vx;
    oimx[p] = e.jx;
  };
  Ext slot[WB], q[PF], P;
#pragma unroll
  for (int k = 0; k < PF; ++k) q[k] = load(k);
  // block 0 (k = 0 .. WB-1) fills the slots; its suffix extreme S[0] is output 0
#pragma unroll
  for (int j = 0; j < WB; ++j) {
    slot[j] = q[j % PF];
    q[j % PF] = load(j + PF);
  }
#pragma unroll
  for (int j = WB - 2; j >= 0; --j) slot[j] = ext_join(slot[j], slot[j + 1]);
  put(0, slot[0]);
  const int kend = L + 2 * R;
  for (int k0 = WB;; k0 += WB) {
#pragma unroll
    for (int j = 0; j < WB; ++j) {
      const int k = k0 + j;
      if (k >= kend) return;
      const Ext v = q[j % PF];  // (k0 + j) % PF == j % PF
      q[j % PF] = load(k + PF);
      P = j == 0 ? v : ext_join(P, v);
      if (j < WB - 1) {
        put(k - 2 * R, ext_join(slot[j + 1], P));
        slot[j] = v;
      } else {
        slot[j] = v;
#pragma unroll
        for (int t = WB - 2; t >= 0; --t) slot[t] = ext_join(slot[t], slot[t + 1]);
        put(k - 2 * R, slot[0]);
      }
    }
  }
}

// clamped box mean of the (E, D) pair along one axis (FP32 sums in window order)
__global__ __launch_bounds__(NT) void k_soft_box(const float2* __restrict__ in,
                                                 float2* __restrict__ out, int W, int H, int D,
                                                 int axis, int r, unsigned nxb) {
  const float k = 1.0f / (float)(2 * r + 1);
  Pos q;
  if (!pos_of(W, H, nxb, q)) return;
  const i64 i = q.i;
  {
    const Line l = line_at(q, axis, W, H, D);
    float se = 0.0f, sd = 0.0f;
    for (int t = -r; t <= r; ++t) {
      const float2 v = in[l.base + (i64)clampi(l.c + t, 0, l.L - 1) * l.stride];
      se += v.x;
      sd += v.y;
    }
    out[i] = make_float2(se * k, sd * k);
  }
}

// transpose of the clamped box mean along one axis: position p receives every window tap
// that lands on it, border positions also the taps that clamp onto them
__global__ __launch_bounds__(NT) void k_soft_box_adj(const float2* __restrict__ in,
                                                     float2* __restrict__ out, int W, int H,
                                                     int D, int axis, int r, unsigned nxb) {
  const float k = 1.0f / (float)(2 * r + 1);
  Pos q;
  if (!pos_of(W, H, nxb, q)) return;
  const i64 i = q.i;
  {
    const Line l = line_at(q, axis, W, H, D);
    const int p = l.c, L = l.L;
    float se = 0.0f, sd = 0.0f;
    for (int c = max(0, p - r); c <= min(L - 1, p + r); ++c) {
      const float2 v = in[l.base + (i64)c * l.stride];
      se += v.x;
      sd += v.y;
    }
    if (p == 0)  // taps c - r + j < 0 of the windows of c < r clamp onto 0
      for (int c = 0; c <= min(L - 1, r - 1); ++c) {
        const float2 v = in[l.base + (i64)c * l.stride];
        se += v.x * (float)(r - c);
        sd += v.y * (float)(r - c);
      }
    if (p == L - 1)  // taps c - r + j > L - 1 of the windows of c > L - 1 - r clamp onto L - 1
      for (int c = max(0, L - r); c <= L - 1; ++c) {
        const float2 v = in[l.base + (i64)c * l.stride];
        se += v.x * (float)(c + r - L + 1);
        sd += v.y * (float)(c + r - L + 1);
      }
    out[i] = make_float2(se * k, sd * k);
  }
}

// The transpose along y or z for radius R as a column walker: a register ring of the
// 2R+1 inputs around each position (zero outside the line), unrolled by the ring period;
// the two border positions add the taps that clamp onto them from direct loads.
template <int R>
__global__ __launch_bounds__(NT) void k_soft_box_adj_walk(const float2* __restrict__ in,
                                                          float2* __restrict__ out, int W, int L,
                                                          i64 sa, i64 so, unsigned nxb) {
  constexpr int WB = 2 * R + 1;
  const float k = 1.0f / (float)WB;
  const int x = (int)(blockIdx.x % nxb) * NT + (int)threadIdx.x;
  if (x >= W) return;
  const i64 base = (i64)(blockIdx.x / nxb) * so + x;
  auto at = [&](int c) -> float2 {
    return (c >= 0 && c < L) ? in[base + (i64)c * sa] : make_float2(0.0f, 0.0f);
  };
  float2 ring[WB];  // position p lives in ring[(p + R) % WB]
#pragma unroll
  for (int q = 0; q < WB; ++q) ring[q] = at(q - R);
  for (int c0 = 0;; c0 += WB) {
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      const int c = c0 + q;
      if (c >= L) return;
      float se = 0.0f, sd = 0.0f;
#pragma unroll
      for (int j = 0; j < WB; ++j) {  // positions c - R .. c + R
        se += ring[(q + j) % WB].x;
        sd += ring[(q + j) % WB].y;
      }
      if (c == 0)  // taps t - R + j < 0 of the windows of t < R clamp onto 0
        for (int t = 0; t <= min(L - 1, R - 1); ++t) {
          const float2 v = in[base + (i64)t * sa];
          se += v.x * (float)(R - t);
          sd += v.y * (float)(R - t);
        }
      if (c == L - 1)  // taps beyond L - 1 of the windows of t > L - 1 - R clamp onto L - 1
        for (int t = max(0, L - R); t <= L - 1; ++t) {
          const float2 v = in[base + (i64)t * sa];
          se += v.x * (float)(t + R - L + 1);
          sd += v.y * (float)(t + R - L + 1);
        }
      out[base + (i64)c * sa] = make_float2(se * k, sd * k);
      ring[q] = at(c + R + 1);  // replaces position c - R
    }
  }
}

__global__ __launch_bounds__(NT) void k_soft_norm(const float* __restrict__ v,
                                                  const float2* __restrict__ ed,
                                                  float* __restrict__ norm, float min_int, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const float I = v[i];
    const float2 e = ed[i];
    norm[i] = I >= min_int ? (I - e.x) / (e.y - e.x) : 0.0f;
  }
}

// direct term into gv; (dL/dE, dL/dD) into g_ed
__global__ __launch_bounds__(NT) void k_soft_norm_adj(const float* __restrict__ v,
                                                      const float2* __restrict__ ed,
                                                      const float* __restrict__ gn,
                                                      float* __restrict__ gv,
                                                      float2* __restrict__ g_ed, float min_int,
                                                      i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const float I = v[i];
    const float2 e = ed[i];
    const float g = I >= min_int ? gn[i] : 0.0f;
    const float inv = 1.0f / (e.y - e.x);
    gv[i] = g * inv;
    g_ed[i] = make_float2(g * (I - e.y) * inv * inv, -g * (I - e.x) * inv * inv);
  }
}

// Sum of v over the run of equal consecutive keys (lanes of one wave) that starts at each
// run's first lane, where *head is set; every lane of the wave must be active.  A segmented
// suffix sum by lane shifts: after the step of offset o each lane holds the sum over
// [lane, lane + 2o) within its run.
__device__ __forceinline__ float run_sum(uint32_t key, float v, bool* head) {
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t prev = __shfl_up(key, 1, 64);
  *head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(*head);
  const unsigned long long after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
  const int tail = after ? __builtin_ctzll(after) - 1 : 63;  // last lane of the run
  for (int o = 1; o < 64; o <<= 1) {
    const float w = __shfl_down(v, o, 64);
    if (lane + o <= tail) v += w;
  }
  return v;
}

// subgradient routing: gv[argmin(i)] += dL/dmin(i), gv[argmax(i)] += dL/dmax(i).  Windows
// of neighbouring voxels mostly share their extreme, so a wave first sums its lanes' values
// per run of equal targets and adds each run with one atomic (one atomic per voxel made the
// contended targets serialize).
__global__ __launch_bounds__(NT) void k_soft_route(const float2* __restrict__ g,
                                                   const uint32_t* __restrict__ amn,
                                                   const uint32_t* __restrict__ amx,
                                                   float* __restrict__ gv, i64 N) {
  constexpr uint32_t kNone = 0xFFFFFFFFu;  // lanes past the volume (N < 2^32 - 1)
  const int lane = (int)(threadIdx.x & 63);
  for (i64 b = (i64)blockIdx.x * NT + (threadIdx.x - lane); b < N; b += (i64)gridDim.x * NT) {
    const i64 i = b + lane;  // the whole wave iterates together
    const bool ok = i < N;
    const float2 q = ok ? g[i] : make_float2(0.0f, 0.0f);
    const uint32_t kmn = ok ? amn[i] : kNone, kmx = ok ? amx[i] : kNone;
    bool hmn, hmx;
    const float smn = run_sum(kmn, q.x, &hmn);
    const float smx = run_sum(kmx, q.y, &hmx);
    if (hmn && kmn != kNone && smn != 0.0f) atomicAdd(gv + kmn, smn);
    if (hmx && kmx != kNone && smx != 0.0f) atomicAdd(gv + kmx, smx);
  }
}

}  // namespace

// norm, and for the backward pass ed (E, D pairs) and arg (argmin N, then argmax N);
// scratch: 4 N words + 4 N words
hipError_t launch_soft_normalize(const float* v, float* norm, float2* ed, uint32_t* arg,
                                 void* scratch, float min_int, int W, int H, int D, int rxy,
                                 int rz, hipStream_t s) {
  const i64 N = (i64)W * H * D;
  float* a = (float*)scratch;  // ping-pong (value, index) min / max planes
  float* fmn[2] = {a, a + 4 * N};
  const unsigned g = grid_for(N);
  const int rad[3] = {rxy, rxy, rz};
  for (int axis = 0; axis < 3; ++axis) {
    float* src = fmn[(axis + 1) & 1];
    float* dst = fmn[axis & 1];
    const float* vmn = axis ? src : v;
    const float* vmx = axis ? src + 2 * N : v;
    const uint32_t* imn = axis ? (const uint32_t*)(src + N) : nullptr;
    const uint32_t* imx = axis ? (const uint32_t*)(src + 3 * N) : nullptr;
    if (axis > 0 && rad[axis] == 10) {  // y or z: column walker
      const i64 WH = (i64)W * H;
      const int L = axis == 1 ? H : D;
      const i64 sa = axis == 1 ? W : WH, so = axis == 1 ? WH : W;
      const i64 n_other = axis == 1 ? D : H;
      hipLaunchKernelGGL((k_soft_argminmax_walk<10>), dim3((unsigned)(xblocks(W) * n_other)),
                         dim3(NT), 0, s, vmn, imn, vmx, imx, dst, (uint32_t*)(dst + N), dst + 2 * N,
                         (uint32_t*)(dst + 3 * N), W, L, sa, so, xblocks(W));
    } else if (axis == 0 && rad[0] <= kXR) {  // x, the first pass: indices are positions
      hipLaunchKernelGGL(k_soft_argminmax_x, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, v, dst,
                         (uint32_t*)(dst + N), dst + 2 * N, (uint32_t*)(dst + 3 * N), W, rad[0],
                         xblocks(W));
    } else {
      hipLaunchKernelGGL(k_soft_argminmax, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, vmn, imn,
                         vmx, imx, dst, (uint32_t*)(dst + N), dst + 2 * N,
                         (uint32_t*)(dst + 3 * N), W, H, D, axis, rad[axis], xblocks(W));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const float* res = fmn[0];  // axis 2 wrote plane set 0
  hipError_t e = hipMemcpyAsync(arg, res + N, N * 4, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(arg + N, res + 3 * N, N * 4, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  // (min, max) pairs -> box along x, y, z -> (E, D)
  float2* p0 = (float2*)fmn[1];
  float2* p1 = ed;
  hipLaunchKernelGGL(k_pack_pair, dim3(g), dim3(NT), 0, s, res, res + 2 * N, p0, N);
  float2* bufs[2] = {p0, p1};
  const i64 WH = (i64)W * H;
  for (int axis = 0; axis < 3; ++axis) {
    // x: p0 -> ed, y: ed -> p0, z: p0 -> ed; y and z by the normalization's box passes
    if (axis == 0 && rad[0] <= kXR) {
      hipLaunchKernelGGL(k_soft_box_x, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, bufs[0],
                         bufs[1], W, rad[0], xblocks(W));
    } else if (axis == 0) {
      hipLaunchKernelGGL(k_soft_box, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, bufs[0], bufs[1],
                         W, H, D, axis, rad[axis], xblocks(W));
    } else {
      e = launch_box_pair(bufs[axis & 1], bufs[(axis + 1) & 1], W, axis == 1 ? H : D,
                          axis == 1 ? W : WH, axis == 1 ? WH : W, axis == 1 ? D : H, rad[axis], s);
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(k_soft_norm, dim3(g), dim3(NT), 0, s, v, ed, norm, min_int, N);
  return hipGetLastError();
}

hipError_t launch_soft_normalize_adjoint(const float* v, const float2* ed, const uint32_t* arg,
                                         const float* gn, float* gv, void* scratch,
                                         float min_int, int W, int H, int D, int rxy, int rz,
                                         hipStream_t s) {
  const i64 N = (i64)W * H * D;
  float2* q0 = (float2*)scratch;
  float2* q1 = q0 + N;
  const unsigned g = grid_for(N);
  const int rad[3] = {rxy, rxy, rz};
  hipLaunchKernelGGL(k_soft_norm_adj, dim3(g), dim3(NT), 0, s, v, ed, gn, gv, q0, min_int, N);
  // the box was applied x, y, z: its transpose runs z, y, x
  float2* bufs[2] = {q0, q1};
  const i64 WH = (i64)W * H;
  for (int k = 0; k < 3; ++k) {
    const int axis = 2 - k;
    if (axis > 0 && rad[axis] == 10) {  // z, y: column walker
      const i64 n_other = axis == 1 ? D : H;
      hipLaunchKernelGGL((k_soft_box_adj_walk<10>), dim3((unsigned)(xblocks(W) * n_other)),
                         dim3(NT), 0, s, bufs[k & 1], bufs[(k + 1) & 1], W, axis == 1 ? H : D,
                         axis == 1 ? (i64)W : WH, axis == 1 ? WH : (i64)W, xblocks(W));
    } else if (axis == 0 && rad[0] <= kXR) {
      hipLaunchKernelGGL(k_soft_box_adj_x, dim3(grid_rows(W, H, D)), dim3(NT), 0, s,
                         bufs[k & 1], bufs[(k + 1) & 1], W, rad[0], xblocks(W));
    } else {
      hipLaunchKernelGGL(k_soft_box_adj, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, bufs[k & 1],
                         bufs[(k + 1) & 1], W, H, D, axis, rad[axis], xblocks(W));
    }
  }
  hipLaunchKernelGGL(k_soft_route, dim3(g), dim3(NT), 0, s, bufs[1], arg, arg + N, gv, N);
  return hipGetLastError();
}

hipError_t launch_soft_loss(const float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                            double inv_tau, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, sums, 2);
  hipLaunchKernelGGL(k_soft_loss, dim3(grid_reduce(N)), dim3(NT), 0, s, hess, mask, target, N,
                     inv_tau, sums);
  return hipGetLastError();
}

hipError_t launch_soft_grad_h(float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                              double inv_tau, double inv_count, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, sums, 1);
  hipLaunchKernelGGL(k_soft_grad_h, dim3(grid_reduce(N)), dim3(NT), 0, s, hess, mask, target, N,
                     inv_tau, inv_count, sums);
  return hipGetLastError();
}

hipError_t launch_hessian_adjoint(const float* gh, float* P, float* out, int W, int H, int D,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_hadj1, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, gh, P, W, H, D,
                     xblocks(W));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_hadj2, dim3(grid_rows(W, H, D)), dim3(NT), 0, s, P, out, W, H, D,
                     xblocks(W));
  return hipGetLastError();
}

hipError_t launch_dot_pair(const float* a, const float2* pair, i64 N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, out, 2);
  hipLaunchKernelGGL(k_dot_pair, dim3(grid_reduce(N)), dim3(NT), 0, s, a, pair, N, out);
  return hipGetLastError();
}

hipError_t launch_dot(const float* a, const float* b, i64 N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, out, 1);
  hipLaunchKernelGGL(k_dot, dim3(grid_reduce(N)), dim3(NT), 0, s, a, b, N, out);
  return hipGetLastError();
}

}  // namespace hdog
