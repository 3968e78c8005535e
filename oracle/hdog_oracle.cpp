// hdog_oracle.cpp -- CPU restatement of the reference HDoG candidate-detection path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
// smoke check and the CPU-baseline leg of the benchmark call it to decide what
// the right answer is.  No product code path links, loads or calls it.
//
// Every function restates the per-voxel semantics of one reference stage under
// the declared FP32 arithmetic policy below.  Reference files are cited as
// path:line relative to the lsb-riken/CUBIC-informatics checkout.
//
// Declared FP32 arithmetic policy (the HIP kernels follow the same rules, which
// is what "bit-exact" is measured against):
//   P1  box sums: val = 0; val += s[c-r+j] for j = 0..2r in that order, then
//       val * fl(1.0f/(2r+1))                   (src/3D/NormalizeFilter.cu:69,100-103)
//   P2  Gaussian: sum = 0; sum = fmaf(w[r-j], s[c+j], sum) for j = -r..r
//                                                 (src/3D/DoG3DFilter.cu:119-126)
//   P3  gradients: ((0 + s[c+1]) - s[c-1]) * 0.5f (src/3D/Hessian3DFilter_element.cu:67-70)
//   P4  leading minors with no contraction, C parse order
//                                                 (src/3D/Hessian3DFilter_element.cu:317-321)
//   P5  IEEE correctly rounded FP32 divide         (src/3D/NormalizeFilter.cu:585)
//   P6  eigen-features in FP64 with libm acos/cos  (src/3D/Eigenvalues.cu:32-94)
//   P7  region sums: integers exact (u64); Hessian and intensity sums accumulated in
//       FP64 in linear-index order and rounded once to FP32.  The reference sums in
//       FP32 in CUB's order, which no test can reproduce (src/3D/RegionalFeatures.cu:316-327).
//
// Compile with -ffp-contract=off so that no multiply-add is fused except the explicit
// fmaf of P2.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

typedef int64_t i64;
typedef uint64_t u64;

static inline i64 clampi(i64 v, i64 lo, i64 hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Line iteration helper: a volume (W,H,D) x-fastest; an "axis pass" processes all lines
// along one axis independently.
struct Geo {
  i64 W, H, D, N;
  Geo(int w, int h, int d) : W(w), H(h), D(d), N((i64)w * h * d) {}
  i64 len(int axis) const { return axis == 0 ? W : (axis == 1 ? H : D); }
  i64 stride(int axis) const { return axis == 0 ? 1 : (axis == 1 ? W : W * H); }
  i64 nlines(int axis) const { return N / len(axis); }
  // base offset of line k along axis
  i64 base(int axis, i64 k) const {
    if (axis == 0) return k * W;                 // k = z*H + y
    if (axis == 1) { i64 z = k / W, x = k % W; return z * W * H + x; }  // k = z*W + x
    return k;                                    // k = y*W + x
  }
};

extern "C" {

// ---------------------------------------------------------------------------
// a1: separable min (erosion) / max (dilation) filter, clamp-to-edge borders.
// src/3D/NormalizeFilter.cu:291-530 (kernels), :549-569 (x -> y -> z sequence).
// Min/max over a clamped box is separable and exact, so the pass order is free.
void oracle_minmax3d(const uint16_t* in, uint16_t* out, int W, int H, int D,
                     int rxy, int rz, int is_min) {
  Geo g(W, H, D);
  std::vector<uint16_t> a(in, in + g.N), b(g.N);
  for (int axis = 0; axis < 3; ++axis) {
    const int r = axis == 2 ? rz : rxy;
    const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
    for (i64 k = 0; k < nl; ++k) {
      const i64 b0 = g.base(axis, k);
      for (i64 c = 0; c < L; ++c) {
        uint16_t v = a[b0 + clampi(c - r, 0, L - 1) * st];
        for (int j = 1; j <= 2 * r; ++j) {
          uint16_t s = a[b0 + clampi(c - r + j, 0, L - 1) * st];
          v = is_min ? std::min(v, s) : std::max(v, s);
        }
        b[b0 + c * st] = v;
      }
    }
    a.swap(b);
  }
  std::memcpy(out, a.data(), g.N * sizeof(uint16_t));
}

// a2: separable box mean, clamp-to-edge, x (on u16 values) -> y -> z, policy P1.
// src/3D/NormalizeFilter.cu:53-286, :537-546.
void oracle_box3d(const uint16_t* in, float* out, int W, int H, int D, int rxy, int rz) {
  Geo g(W, H, D);
  std::vector<float> a(g.N), b(g.N);
  for (i64 i = 0; i < g.N; ++i) a[i] = (float)in[i];
  for (int axis = 0; axis < 3; ++axis) {
    const int r = axis == 2 ? rz : rxy;
    const float k = 1.0f / (float)(2 * r + 1);
    const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
    for (i64 l = 0; l < nl; ++l) {
      const i64 b0 = g.base(axis, l);
      for (i64 c = 0; c < L; ++c) {
        float val = 0.0f;
        for (int j = 0; j <= 2 * r; ++j) val += a[b0 + clampi(c - r + j, 0, L - 1) * st];
        b[b0 + c * st] = val * k;
      }
    }
    a.swap(b);
  }
  std::memcpy(out, a.data(), g.N * sizeof(float));
}

// a3: local contrast normalization.  src/3D/NormalizeFilter.cu:571-654.
// erosion_l / dilation_l may be NULL; when given they receive E and D.
void oracle_normalize3d(const uint16_t* img, float* norm, float min_intensity,
                        int W, int H, int D, int rxy, int rz,
                        float* erosion_l, float* dilation_l) {
  Geo g(W, H, D);
  if (rxy == 0 || rz == 0) {  // Copy3DKernel, :588-601 and :612-623
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < g.N; ++i) {
      const float v = (float)img[i];
      norm[i] = (v >= min_intensity) ? v : 0.0f;
    }
    return;
  }
  std::vector<uint16_t> t(g.N);
  std::vector<float> E(g.N), Dl(g.N);
  oracle_minmax3d(img, t.data(), W, H, D, rxy, rz, 1);
  oracle_box3d(t.data(), E.data(), W, H, D, rxy, rz);
  oracle_minmax3d(img, t.data(), W, H, D, rxy, rz, 0);
  oracle_box3d(t.data(), Dl.data(), W, H, D, rxy, rz);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < g.N; ++i) {
    const float v = (float)img[i];
    norm[i] = (v >= min_intensity) ? (v - E[i]) / (Dl[i] - E[i]) : 0.0f;
  }
  if (erosion_l) std::memcpy(erosion_l, E.data(), g.N * sizeof(float));
  if (dilation_l) std::memcpy(dilation_l, Dl.data(), g.N * sizeof(float));
}

// ---------------------------------------------------------------------------
// a4: Gaussian weights.  src/3D/DoG3DFilter.cu:372-392 with TRUNCATE (:15) made a
// parameter.  radius = (int)(truncate*sigma + 0.5) evaluated in double; weights are
// float(exp(-0.5*i*i/(double)fl(sigma^2))) normalised by a float running sum that
// starts at 1 and adds 2p per tap.  w must hold 2r+1 floats.  Returns r.
int oracle_gaussian_radius(float sigma, double truncate) {
  return (int)(truncate * (double)sigma + 0.5);
}
int oracle_gaussian_weights(float sigma, double truncate, float* w) {
  const int lw = oracle_gaussian_radius(sigma, truncate);
  float s2 = sigma * sigma;
  float sum = 1.0f;
  w[lw] = 1.0f;
  for (int i = 1; i <= lw; ++i) {
    float p = (float)std::exp(-0.5 * i * i / (double)s2);
    w[lw - i] = p;
    w[lw + i] = p;
    sum += p * 2;
  }
  for (int i = 0; i < 2 * lw + 1; ++i) w[i] /= sum;
  return lw;
}

// src/3D/DoG3DFilter.cu:432 -- powf(sigma_xy1, 2*gamma) * powf(sigma_z1, gamma) in FP32.
float oracle_dog_normalizer(float sigma_xy1, float sigma_z1, float gamma) {
  return powf(sigma_xy1, gamma * 2) * powf(sigma_z1, gamma);
}

// a5: separable Gaussian x -> y (xy weights) -> z (z weights), zero padding, policy P2.
// src/3D/DoG3DFilter.cu:70-343, :417-429.
void oracle_gaussian3d(const float* in, float* out, int W, int H, int D,
                       const float* wxy, int rxy, const float* wz, int rz) {
  Geo g(W, H, D);
  std::vector<float> a(in, in + g.N), b(g.N);
  for (int axis = 0; axis < 3; ++axis) {
    const int r = axis == 2 ? rz : rxy;
    const float* w = axis == 2 ? wz : wxy;
    const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
    for (i64 l = 0; l < nl; ++l) {
      const i64 b0 = g.base(axis, l);
      for (i64 c = 0; c < L; ++c) {
        float sum = 0.0f;
        for (int j = -r; j <= r; ++j) {
          const i64 p = c + j;
          const float s = (p >= 0 && p < L) ? a[b0 + p * st] : 0.0f;
          sum = fmaf(w[r - j], s, sum);
        }
        b[b0 + c * st] = sum;
      }
    }
    a.swap(b);
  }
  std::memcpy(out, a.data(), g.N * sizeof(float));
}

// a6: DoG = (G_sigma1 - G_sigma2) * normalizer.  src/3D/DoG3DFilter.cu:346-366, :431-437.
// Returns the z radius of sigma2 (what initGaussian3DKernel returns, :394-415).
int oracle_dog3d(const float* norm, float* dog, int W, int H, int D,
                 float sxy1, float sz1, float sxy2, float sz2, double truncate, float gamma) {
  Geo g(W, H, D);
  std::vector<float> w0(2 * oracle_gaussian_radius(sxy1, truncate) + 1),
      w1(2 * oracle_gaussian_radius(sz1, truncate) + 1),
      w2(2 * oracle_gaussian_radius(sxy2, truncate) + 1),
      w3(2 * oracle_gaussian_radius(sz2, truncate) + 1);
  const int r0 = oracle_gaussian_weights(sxy1, truncate, w0.data());
  const int r1 = oracle_gaussian_weights(sz1, truncate, w1.data());
  const int r2 = oracle_gaussian_weights(sxy2, truncate, w2.data());
  const int r3 = oracle_gaussian_weights(sz2, truncate, w3.data());
  std::vector<float> g2(g.N);
  oracle_gaussian3d(norm, g2.data(), W, H, D, w2.data(), r2, w3.data(), r3);
  oracle_gaussian3d(norm, dog, W, H, D, w0.data(), r0, w1.data(), r1);
  const float nz = oracle_dog_normalizer(sxy1, sz1, gamma);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < g.N; ++i) dog[i] = (dog[i] - g2[i]) * nz;
  return r3;
}

// ---------------------------------------------------------------------------
// a7: central difference along one axis with zero padding, policy P3.
// src/3D/Hessian3DFilter_element.cu:22-288.
static void gradient_axis(const float* in, float* out, const Geo& g, int axis) {
  const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
  for (i64 l = 0; l < nl; ++l) {
    const i64 b0 = g.base(axis, l);
    for (i64 c = 0; c < L; ++c) {
      const float hi = (c + 1 < L) ? in[b0 + (c + 1) * st] : 0.0f;
      const float lo = (c - 1 >= 0) ? in[b0 + (c - 1) * st] : 0.0f;
      float sum = 0.0f;
      sum += hi;
      sum -= lo;
      sum *= 0.5f;
      out[b0 + c * st] = sum;
    }
  }
}

// a8: Sylvester test for negative definiteness, policy P4.
// src/3D/Hessian3DFilter_element.cu:309-321.
uint8_t oracle_negdef(float xx, float xy, float xz, float yy, float yz, float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

// a7+a8: Hessian elements (SoA [xx,xy,xz,yy,yz,zz] at pitch N) and candidate mask.
// Pass sequence of src/3D/Hessian3DFilter_element.cu:373-385.
void oracle_hessian3d(const float* dog, float* hess, uint8_t* pd, int W, int H, int D) {
  Geo g(W, H, D);
  const i64 N = g.N;
  float *Hxx = hess, *Hxy = hess + N, *Hxz = hess + 2 * N, *Hyy = hess + 3 * N,
        *Hyz = hess + 4 * N, *Hzz = hess + 5 * N;
  std::vector<float> t(N);
  gradient_axis(dog, t.data(), g, 0);
  gradient_axis(t.data(), Hxx, g, 0);
  gradient_axis(t.data(), Hxy, g, 1);
  gradient_axis(t.data(), Hxz, g, 2);
  gradient_axis(dog, t.data(), g, 1);
  gradient_axis(t.data(), Hyy, g, 1);
  gradient_axis(t.data(), Hyz, g, 2);
  gradient_axis(dog, t.data(), g, 2);
  gradient_axis(t.data(), Hzz, g, 2);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < N; ++i)
    pd[i] = oracle_negdef(Hxx[i], Hxy[i], Hxz[i], Hyy[i], Hyz[i], Hzz[i]);
}

// ---------------------------------------------------------------------------
// a9: connected components.  src/3D/CCL3D.cu:17-117.  The reference iterates to the
// fixpoint where every foreground voxel holds the minimum linear index of its
// component; adjacency is the 27 linear offsets dz*W*H + dy*W + dx kept when the
// target lies in [0, N) (so x and y borders wrap into the next row / slice), and the
// voxel at index 0 is never foreground (label = index * mask).  This restatement
// reaches the same fixpoint with a sequential union-find that always links the
// larger root under the smaller one.
// Writes the distinct positive offsets (at most 13) into offs and returns how many.
int oracle_positive_offsets(int W, int H, i64* offs) {
  std::vector<i64> v;
  const i64 WH = (i64)W * H;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const i64 o = dz * WH + dy * (i64)W + dx;
        if (o > 0) v.push_back(o);
      }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  for (size_t i = 0; i < v.size(); ++i) offs[i] = v[i];
  return (int)v.size();
}

static i64 uf_find(std::vector<i64>& p, i64 x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

void oracle_ccl3d(const uint8_t* pd, uint32_t* labels, int W, int H, int D) {
  Geo g(W, H, D);
  const i64 N = g.N;
  i64 offs[13];
  const int no = oracle_positive_offsets(W, H, offs);
  std::vector<i64> p(N);
  for (i64 i = 0; i < N; ++i) p[i] = i;
  for (i64 i = 1; i < N; ++i) {
    if (!pd[i]) continue;
    for (int k = 0; k < no; ++k) {
      const i64 j = i + offs[k];
      if (j >= N || !pd[j]) continue;
      i64 a = uf_find(p, i), b = uf_find(p, j);
      if (a == b) continue;
      if (a < b) std::swap(a, b);
      p[a] = b;
    }
  }
  labels[0] = 0;
  for (i64 i = 1; i < N; ++i) labels[i] = pd[i] ? (uint32_t)uf_find(p, i) : 0u;
}

// ---------------------------------------------------------------------------
// a11: eigen-features of one summed Hessian (Eberly's closed form, FP64).
// src/3D/Eigenvalues.cu:32-94.  out2[0] = structureness, out2[1] = blobness.
void oracle_eigen_features(const float* h6, float* out2) {
  const double xx = h6[0], xy = h6[1], xz = h6[2], yy = h6[3], yz = h6[4], zz = h6[5];
  const double max0 = std::fmax(std::fabs(xx), std::fabs(xy));
  const double max1 = std::fmax(std::fabs(xz), std::fabs(yy));
  const double max2 = std::fmax(std::fabs(yz), std::fabs(zz));
  const double m = std::fmax(std::fmax(max0, max1), max2);
  if (m < 1e-30) { out2[0] = 0.0f; out2[1] = 0.0f; return; }
  const double norm = xy * xy + xz * xz + yz * yz;
  double e0, e1, e2;
  if (norm > 1e-30) {
    const double t = (xx + yy + zz) / 3.0;
    const double b00 = xx - t, b11 = yy - t, b22 = zz - t;
    const double den = std::sqrt((b00 * b00 + b11 * b11 + b22 * b22 + norm * 2.0) / 6.0);
    const double c00 = b11 * b22 - yz * yz;
    const double c01 = xy * b22 - yz * xz;
    const double c02 = xy * yz - b11 * xz;
    double half = (b00 * c00 - xy * c01 + xz * c02) / (den * den * den) * 0.5;
    half = std::fmin(std::fmax(half, -1.0), 1.0);
    const double ang = std::acos(half) / 3.0;
    const double beta2 = std::cos(ang) * 2.0;
    const double beta0 = std::cos(ang + 2.094395102393195) * 2.0;
    const double beta1 = -(beta0 + beta2);
    e0 = t + den * beta0;
    e1 = t + den * beta1;
    e2 = t + den * beta2;
  } else {
    e0 = xx; e1 = yy; e2 = zz;
  }
  out2[0] = (float)(e0 * e0 + e1 * e1 + e2 * e2);
  out2[1] = (float)(e2 * e2 / (e0 * e1));
}

// a10, a12, a13: per-region reductions.  src/3D/RegionalFeatures.cu:65-88 (max),
// :114-137 (sum, "ave mode"), :139-255 (size RLE into u16, centroid = float sum /
// u16 size), :294-329 (Hessian sums), src/3D/Eigenvalues.cu (a11).
// Regions are the distinct labels in ascending order, region 0 = background (label 0);
// returns R (including region 0) or -1 if R exceeds cap.  Arrays have cap entries;
// entry 0 is left zero (the reference computes it but never emits it).
// Extra integer outputs that have no float rounding: count (u64), coordinate sums,
// raw-intensity sum and raw-intensity max (src/3D/RegionalFeatures.cu:39-63).
i64 oracle_regions(const uint32_t* labels, const float* hess, const float* norm,
                   const uint16_t* img, int W, int H, int D, int is_ave_mode, i64 cap,
                   uint32_t* out_label, uint16_t* out_size16, u64* out_count,
                   u64* out_sumxyz /*3*cap*/, float* out_centroid /*3*cap*/,
                   float* out_hsum /*6*cap*/, float* out_eig /*2*cap*/,
                   float* out_norm /*cap: max or sum*/, u64* out_isum, uint16_t* out_imax) {
  Geo g(W, H, D);
  const i64 N = g.N;
  // region numbering: ascending label order (the CUB sort order)
  std::vector<i64> rid(N, -1);
  i64 R = 1;
  for (i64 i = 1; i < N; ++i)
    if (labels[i] == (uint32_t)i) rid[i] = R++;
  if (R > cap) return -1;
  std::vector<double> hs(6 * R, 0.0), ns(R, 0.0);
  std::vector<float> nmax(R, -INFINITY);
  std::vector<u64> cnt(R, 0), sx(3 * R, 0), isum(R, 0);
  std::vector<uint16_t> imax(R, 0);
  for (i64 i = 0; i < N; ++i) {
    const uint32_t l = labels[i];
    if (!l) continue;
    const i64 r = rid[l];
    const i64 x = i % W, y = (i / W) % H, z = i / ((i64)W * H);
    cnt[r] += 1;
    sx[3 * r] += x; sx[3 * r + 1] += y; sx[3 * r + 2] += z;
    for (int k = 0; k < 6; ++k) hs[6 * r + k] += (double)hess[k * N + i];
    const float v = norm[i];
    if (v > nmax[r]) nmax[r] = v;   // MaxOp (b > a) ? b : a, RegionalFeatures.cu:13-20
    ns[r] += (double)v;
    isum[r] += img[i];
    imax[r] = std::max(imax[r], img[i]);
  }
  std::memset(out_label, 0, cap * sizeof(uint32_t));
  std::memset(out_size16, 0, cap * sizeof(uint16_t));
  std::memset(out_count, 0, cap * sizeof(u64));
  std::memset(out_sumxyz, 0, 3 * cap * sizeof(u64));
  std::memset(out_centroid, 0, 3 * cap * sizeof(float));
  std::memset(out_hsum, 0, 6 * cap * sizeof(float));
  std::memset(out_eig, 0, 2 * cap * sizeof(float));
  std::memset(out_norm, 0, cap * sizeof(float));
  std::memset(out_isum, 0, cap * sizeof(u64));
  std::memset(out_imax, 0, cap * sizeof(uint16_t));
  for (i64 i = 1; i < N; ++i) {
    if (rid[i] < 0) continue;
    const i64 r = rid[i];
    out_label[r] = (uint32_t)i;
    const uint16_t s16 = (uint16_t)cnt[r];  // u16 RLE count wraps, RegionalFeatures.cu:148-150
    out_size16[r] = s16;
    out_count[r] = cnt[r];
    for (int a = 0; a < 3; ++a) {
      out_sumxyz[3 * r + a] = sx[3 * r + a];
      out_centroid[3 * r + a] = (float)sx[3 * r + a] / (float)s16;  // DivideKernel :192-199
    }
    float h6[6];
    for (int k = 0; k < 6; ++k) h6[k] = (float)hs[6 * r + k];
    for (int k = 0; k < 6; ++k) out_hsum[6 * r + k] = h6[k];
    oracle_eigen_features(h6, out_eig + 2 * r);
    out_norm[r] = is_ave_mode ? (float)ns[r] : nmax[r];
    out_isum[r] = isum[r];
    out_imax[r] = imax[r];
  }
  return R;
}

// ---------------------------------------------------------------------------
// Whole substack in the order of src/3D/HDoG3D_main.cpp:171-205.  Intermediates are
// returned to the caller (norm N, dog N, hess 6N, pd N, labels N) so the step
// outputs can all be compared.  Returns R as oracle_regions does.
i64 oracle_substack(const uint16_t* img, int W, int H, int D,
                    float min_intensity, int rxy, int rz,
                    float sxy1, float sz1, float sxy2, float sz2, double truncate, float gamma,
                    int is_ave_mode,
                    float* norm, float* dog, float* hess, uint8_t* pd, uint32_t* labels,
                    i64 cap, uint32_t* out_label, uint16_t* out_size16, u64* out_count,
                    u64* out_sumxyz, float* out_centroid, float* out_hsum, float* out_eig,
                    float* out_norm, u64* out_isum, uint16_t* out_imax) {
  oracle_normalize3d(img, norm, min_intensity, W, H, D, rxy, rz, nullptr, nullptr);
  oracle_dog3d(norm, dog, W, H, D, sxy1, sz1, sxy2, sz2, truncate, gamma);
  oracle_hessian3d(dog, hess, pd, W, H, D);
  oracle_ccl3d(pd, labels, W, H, D);
  return oracle_regions(labels, hess, norm, img, W, H, D, is_ave_mode, cap, out_label,
                        out_size16, out_count, out_sumxyz, out_centroid, out_hsum, out_eig,
                        out_norm, out_isum, out_imax);
}

// ---------------------------------------------------------------------------
// FP64 restatement of the filter stages: the reference for the precision forms (FP16 /
// BF16 storage), against which their per-stage error and candidate agreement are stated.
// The same stage semantics as above in double throughout: box means of the (exact,
// integer) min / max filters with double sums and 1/(2r+1) in double; the divide in
// double; Gaussian weights exp(-0.5 i^2 / sigma^2) normalised in double (the recipe of
// src/3D/DoG3DFilter.cu:372-392 without its float roundings); the separable Gaussian as
// a double fma chain with zero padding; the DoG normalizer pow(sxy1, 2 gamma) *
// pow(sz1, gamma) in double; central differences and leading minors in double.  Sigmas
// and radii are those of the float parameters, so the supports match the FP32 path.

static void box3d_f64(const uint16_t* in, double* out, int W, int H, int D, int rxy, int rz) {
  Geo g(W, H, D);
  std::vector<double> a(g.N), b(g.N);
  for (i64 i = 0; i < g.N; ++i) a[i] = (double)in[i];
  for (int axis = 0; axis < 3; ++axis) {
    const int r = axis == 2 ? rz : rxy;
    const double k = 1.0 / (double)(2 * r + 1);
    const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
    for (i64 l = 0; l < nl; ++l) {
      const i64 b0 = g.base(axis, l);
      for (i64 c = 0; c < L; ++c) {
        double val = 0.0;
        for (int j = 0; j <= 2 * r; ++j) val += a[b0 + clampi(c - r + j, 0, L - 1) * st];
        b[b0 + c * st] = val * k;
      }
    }
    a.swap(b);
  }
  std::memcpy(out, a.data(), g.N * sizeof(double));
}

void oracle64_normalize3d(const uint16_t* img, double* norm, double min_intensity, int W, int H,
                          int D, int rxy, int rz) {
  Geo g(W, H, D);
  if (rxy == 0 || rz == 0) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < g.N; ++i) {
      const double v = (double)img[i];
      norm[i] = (v >= min_intensity) ? v : 0.0;
    }
    return;
  }
  std::vector<uint16_t> t(g.N);
  std::vector<double> E(g.N), Dl(g.N);
  oracle_minmax3d(img, t.data(), W, H, D, rxy, rz, 1);
  box3d_f64(t.data(), E.data(), W, H, D, rxy, rz);
  oracle_minmax3d(img, t.data(), W, H, D, rxy, rz, 0);
  box3d_f64(t.data(), Dl.data(), W, H, D, rxy, rz);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < g.N; ++i) {
    const double v = (double)img[i];
    norm[i] = (v >= min_intensity) ? (v - E[i]) / (Dl[i] - E[i]) : 0.0;
  }
}

int oracle64_gaussian_weights(float sigma, double truncate, double* w) {
  const int lw = oracle_gaussian_radius(sigma, truncate);
  const double s2 = (double)sigma * (double)sigma;
  double sum = 1.0;
  w[lw] = 1.0;
  for (int i = 1; i <= lw; ++i) {
    const double p = std::exp(-0.5 * i * i / s2);
    w[lw - i] = p;
    w[lw + i] = p;
    sum += p * 2;
  }
  for (int i = 0; i < 2 * lw + 1; ++i) w[i] /= sum;
  return lw;
}

void oracle64_gaussian3d(const double* in, double* out, int W, int H, int D, const double* wxy,
                         int rxy, const double* wz, int rz) {
  Geo g(W, H, D);
  std::vector<double> a(in, in + g.N), b(g.N);
  for (int axis = 0; axis < 3; ++axis) {
    const int r = axis == 2 ? rz : rxy;
    const double* w = axis == 2 ? wz : wxy;
    const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
    for (i64 l = 0; l < nl; ++l) {
      const i64 b0 = g.base(axis, l);
      for (i64 c = 0; c < L; ++c) {
        double sum = 0.0;
        for (int j = -r; j <= r; ++j) {
          const i64 p = c + j;
          const double v = (p >= 0 && p < L) ? a[b0 + p * st] : 0.0;
          sum = std::fma(w[r - j], v, sum);
        }
        b[b0 + c * st] = sum;
      }
    }
    a.swap(b);
  }
  std::memcpy(out, a.data(), g.N * sizeof(double));
}

int oracle64_dog3d(const double* norm, double* dog, int W, int H, int D, float sxy1, float sz1,
                   float sxy2, float sz2, double truncate, float gamma) {
  Geo g(W, H, D);
  std::vector<double> w0(2 * oracle_gaussian_radius(sxy1, truncate) + 1),
      w1(2 * oracle_gaussian_radius(sz1, truncate) + 1),
      w2(2 * oracle_gaussian_radius(sxy2, truncate) + 1),
      w3(2 * oracle_gaussian_radius(sz2, truncate) + 1);
  const int r0 = oracle64_gaussian_weights(sxy1, truncate, w0.data());
  const int r1 = oracle64_gaussian_weights(sz1, truncate, w1.data());
  const int r2 = oracle64_gaussian_weights(sxy2, truncate, w2.data());
  const int r3 = oracle64_gaussian_weights(sz2, truncate, w3.data());
  std::vector<double> g2(g.N);
  oracle64_gaussian3d(norm, g2.data(), W, H, D, w2.data(), r2, w3.data(), r3);
  oracle64_gaussian3d(norm, dog, W, H, D, w0.data(), r0, w1.data(), r1);
  const double nz = std::pow((double)sxy1, (double)gamma * 2) * std::pow((double)sz1, (double)gamma);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < g.N; ++i) dog[i] = (dog[i] - g2[i]) * nz;
  return r3;
}

static void gradient_axis_f64(const double* in, double* out, const Geo& g, int axis) {
  const i64 L = g.len(axis), st = g.stride(axis), nl = g.nlines(axis);
#pragma omp parallel for schedule(static)
  for (i64 l = 0; l < nl; ++l) {
    const i64 b0 = g.base(axis, l);
    for (i64 c = 0; c < L; ++c) {
      const double hi = (c + 1 < L) ? in[b0 + (c + 1) * st] : 0.0;
      const double lo = (c - 1 >= 0) ? in[b0 + (c - 1) * st] : 0.0;
      out[b0 + c * st] = (hi - lo) * 0.5;
    }
  }
}

uint8_t oracle64_negdef(double xx, double xy, double xz, double yy, double yz, double zz) {
  const double m2 = xx * yy - xy * xy;
  const double m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

void oracle64_hessian3d(const double* dog, double* hess, uint8_t* pd, int W, int H, int D) {
  Geo g(W, H, D);
  const i64 N = g.N;
  double *Hxx = hess, *Hxy = hess + N, *Hxz = hess + 2 * N, *Hyy = hess + 3 * N,
         *Hyz = hess + 4 * N, *Hzz = hess + 5 * N;
  std::vector<double> t(N);
  gradient_axis_f64(dog, t.data(), g, 0);
  gradient_axis_f64(t.data(), Hxx, g, 0);
  gradient_axis_f64(t.data(), Hxy, g, 1);
  gradient_axis_f64(t.data(), Hxz, g, 2);
  gradient_axis_f64(dog, t.data(), g, 1);
  gradient_axis_f64(t.data(), Hyy, g, 1);
  gradient_axis_f64(t.data(), Hyz, g, 2);
  gradient_axis_f64(dog, t.data(), g, 2);
  gradient_axis_f64(t.data(), Hzz, g, 2);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < N; ++i)
    pd[i] = oracle64_negdef(Hxx[i], Hxy[i], Hxz[i], Hyy[i], Hyz[i], Hzz[i]);
}

}  // extern "C"
