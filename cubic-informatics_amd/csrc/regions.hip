// regions.hip -- per-region features without sorting.
//
// Reference semantics (src/3D/RegionalFeatures.cu:65-329, src/3D/Eigenvalues.cu:18-112):
// regions are the distinct labels in ascending order (region 0 = background), each with
// the six Hessian sums, max (or, in ave mode, sum) of normalized intensity, the voxel
// count as u16 (wrapping modulo 65536), the centroid sum(coord) / (float)size and the
// eigen-features of the summed Hessian in FP64.  The reference gets there with CUB radix
// sorts of (label, value) pairs and reduce-by-key, about 1 KB of traffic per voxel.
//
// MI355X design: the labels already are the component minima, so a region's id is the
// rank of its root among all roots in linear order -- a block-count / scan / assign
// pass that preserves the reference's ascending order.  Values are then reduced straight
// into one 128-byte accumulator line per region with hardware atomics: u64 for counts
// and coordinate / intensity sums (exact), FP64 for Hessian and intensity-float sums,
// ordered-integer atomicMax for the float max.  A finalize kernel turns the sums into
// the reference's output arrays.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;
constexpr int ITEMS = 16;
constexpr i64 BLK = (i64)NT * ITEMS;  // voxels per numbering block

__device__ __forceinline__ int block_sum_int(int v, int* sh) {
  // wave64 reduction, then across the 4 waves
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0) t = sh[0] + sh[1] + sh[2] + sh[3];
  return t;
}

__global__ __launch_bounds__(NT) void k_count_roots(const uint32_t* __restrict__ L, i64 N,
                                                    uint32_t* __restrict__ cnt) {
  __shared__ int sh[4];
  const i64 base = (i64)blockIdx.x * BLK;
  int c = 0;
  for (int k = 0; k < ITEMS; ++k) {
    const i64 i = base + (i64)k * NT + threadIdx.x;
    if (i > 0 && i < N && L[i] == (uint32_t)i) ++c;
  }
  const int t = block_sum_int(c, sh);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)t;
}

// exclusive scan of cnt[0..nb) in place by one work-group of 1024 threads; *d_R = 1 + total
__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* cnt, i64 nb, uint32_t* d_R) {
  __shared__ unsigned long long part[1024];
  const i64 per = (nb + 1023) / 1024;
  const i64 lo = threadIdx.x * per, hi = std::min<i64>(lo + per, nb);
  unsigned long long s = 0;
  for (i64 i = lo; i < hi; ++i) s += cnt[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    unsigned long long v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (i64 i = lo; i < hi; ++i) {
    const uint32_t c = cnt[i];
    cnt[i] = (uint32_t)run;
    run += c;
  }
  if (threadIdx.x == 1023) *d_R = (uint32_t)(1 + part[1023]);
}

// rid[root] = 1 + rank of the root in linear order; acc[rid].root = root
__global__ __launch_bounds__(NT) void k_assign_ids(const uint32_t* __restrict__ L, i64 N,
                                                   const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ rid,
                                                   RegionAcc* __restrict__ acc, i64 cap) {
  __shared__ uint32_t wsum[4];
  const i64 base = (i64)blockIdx.x * BLK;
  uint32_t run = off[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < ITEMS; ++k) {
    const i64 i = base + (i64)k * NT + threadIdx.x;
    const bool isroot = i > 0 && i < N && L[i] == (uint32_t)i;
    const unsigned long long m = __ballot(isroot);
    const uint32_t pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < 4; ++q) {
      woff += q < wave ? wsum[q] : 0;
      tot += wsum[q];
    }
    if (isroot) {
      const uint32_t r = 1 + run + woff + pre;
      rid[i] = r;
      if ((i64)r < cap) acc[r].root = (uint32_t)i;
    }
    run += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(NT) void k_zero_acc(RegionAcc* acc, i64 R, const uint32_t* d_R,
                                                  i64 cap) {
  if (d_R) R = std::min<i64>((i64)*d_R, cap);
  const i64 nwords = R * (i64)(sizeof(RegionAcc) / 8);
  unsigned long long* p = (unsigned long long*)acc;
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < nwords; i += (i64)gridDim.x * NT) p[i] = 0;
}

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdecode(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// A work-group takes a brick of 64 x BY x BZ voxels (one row of 64 x per wave and z
// step).  All loads of the brick are issued up front; every candidate voxel then folds
// itself into a per-work-group hash table in LDS keyed by region id (LDS atomics, FP64
// for the Hessian sums), and each occupied slot finally issues one set of global
// atomics, so a region pays one set per brick it touches instead of one per voxel.  A
// voxel that finds the table full falls back to global atomics directly (128 slots: at
// most 128 separate 26-connected regions fit in a 64 x 4 x 4 brick).
constexpr int BY = 4, BZ = 4, HSB = 7, HS = 1 << HSB;
constexpr uint32_t EMPTY = 0xFFFFFFFFu;

struct RunSums {
  uint32_t cnt, sx, sy, sz, isum, imax, nkey;
  double h[6], ns;
};

__device__ __forceinline__ void flush_global(RegionAcc* a, const RunSums& v, int is_ave,
                                             bool has_img) {
  atomicAdd(&a->cnt, v.cnt);
  atomicAdd(&a->sx, (unsigned long long)v.sx);
  atomicAdd(&a->sy, (unsigned long long)v.sy);
  atomicAdd(&a->sz, (unsigned long long)v.sz);
#pragma unroll
  for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&a->h[k], v.h[k]);
  if (is_ave)
    unsafeAtomicAdd(&a->nsum, v.ns);
  else
    atomicMax(&a->nmax_key, v.nkey);
  if (has_img) {
    atomicAdd(&a->isum, (unsigned long long)v.isum);
    atomicMax(&a->imax, v.imax);
  }
}

template <typename TH, typename TN>
__global__ __launch_bounds__(NT) void k_accum(const uint32_t* __restrict__ L,
                                              const uint32_t* __restrict__ rid,
                                              const TH* __restrict__ hess,
                                              const TN* __restrict__ norm,
                                              const uint16_t* __restrict__ img, RegionAcc* acc,
                                              uint32_t cap, int W, int H, int D, i64 N, int nbx,
                                              int nby, int is_ave) {
  __shared__ uint32_t key[HS];
  __shared__ uint32_t tcnt[HS], tsx[HS], tsy[HS], tsz[HS], tis[HS], tim[HS], tnk[HS];
  __shared__ double th[6][HS], tns[HS];
  for (int t = threadIdx.x; t < HS; t += NT) {
    key[t] = EMPTY;
    tcnt[t] = tsx[t] = tsy[t] = tsz[t] = tis[t] = tim[t] = tnk[t] = 0u;
#pragma unroll
    for (int k = 0; k < 6; ++k) th[k][t] = 0.0;
    tns[t] = 0.0;
  }
  unsigned b = blockIdx.x;
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x = bx * 64 + (threadIdx.x & 63);
  const int y = by * BY + (threadIdx.x >> 6);
  const int zb = (int)b * BZ;
  const i64 WH = (i64)W * H;
  const bool has_img = img != nullptr;
  // loads of all BZ rows issued before any use: two dependent rounds in total
  uint32_t l[BZ], r[BZ];
#pragma unroll
  for (int zz = 0; zz < BZ; ++zz) {
    const int z = zb + zz;
    l[zz] = (x < W && y < H && z < D) ? L[(i64)z * WH + (i64)y * W + x] : 0u;
  }
  float hv[BZ][6], nv[BZ];
  uint32_t iv[BZ];
#pragma unroll
  for (int zz = 0; zz < BZ; ++zz) {
    const i64 i = (i64)(zb + zz) * WH + (i64)y * W + x;
    const bool on = l[zz] != 0u;
    r[zz] = on ? rid[l[zz]] : 0u;
#pragma unroll
    for (int k = 0; k < 6; ++k) hv[zz][k] = on ? ld(hess, (i64)k * N + i) : 0.0f;
    nv[zz] = on ? ld(norm, i) : 0.0f;
    iv[zz] = (on && has_img) ? (uint32_t)img[i] : 0u;
  }
  __syncthreads();  // table initialised
#pragma unroll
  for (int zz = 0; zz < BZ; ++zz) {
    if (!l[zz] || r[zz] >= cap) continue;
    const uint32_t rz = r[zz];
    uint32_t slot = (rz * 2654435761u) >> (32 - HSB);
    int probes = 0;
    bool placed = false;
    while (probes < HS) {
      const uint32_t k0 = atomicCAS(&key[slot], EMPTY, rz);
      if (k0 == EMPTY || k0 == rz) { placed = true; break; }
      slot = (slot + 1) & (HS - 1);
      ++probes;
    }
    if (!placed) {  // not reachable for this brick size (see above); kept for safety
      RunSums v;
      v.cnt = 1u; v.sx = (uint32_t)x; v.sy = (uint32_t)y; v.sz = (uint32_t)(zb + zz);
      v.isum = iv[zz]; v.imax = iv[zz];
      v.nkey = is_ave ? 0u : fkey(nv[zz]);
      v.ns = is_ave ? (double)nv[zz] : 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) v.h[k] = (double)hv[zz][k];
      flush_global(acc + rz, v, is_ave, has_img);
      continue;
    }
    atomicAdd(&tcnt[slot], 1u);
    atomicAdd(&tsx[slot], (uint32_t)x);
    atomicAdd(&tsy[slot], (uint32_t)y);
    atomicAdd(&tsz[slot], (uint32_t)(zb + zz));
#pragma unroll
    for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&th[k][slot], (double)hv[zz][k]);
    if (is_ave)
      unsafeAtomicAdd(&tns[slot], (double)nv[zz]);
    else
      atomicMax(&tnk[slot], fkey(nv[zz]));
    if (has_img) {
      atomicAdd(&tis[slot], iv[zz]);
      atomicMax(&tim[slot], iv[zz]);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < HS; t += NT) {
    const uint32_t rr = key[t];
    if (rr == EMPTY) continue;
    RunSums v;
    v.cnt = tcnt[t]; v.sx = tsx[t]; v.sy = tsy[t]; v.sz = tsz[t];
    v.isum = tis[t]; v.imax = tim[t]; v.nkey = tnk[t]; v.ns = tns[t];
#pragma unroll
    for (int k = 0; k < 6; ++k) v.h[k] = th[k][t];
    flush_global(acc + rr, v, is_ave, has_img);
  }
}

// src/3D/Eigenvalues.cu:32-94 (Eberly's closed form, FP64)
__device__ void eigen_features(const float h6[6], float* st, float* bl) {
  const double xx = h6[0], xy = h6[1], xz = h6[2], yy = h6[3], yz = h6[4], zz = h6[5];
  const double m = fmax(fmax(fmax(fabs(xx), fabs(xy)), fmax(fabs(xz), fabs(yy))),
                        fmax(fabs(yz), fabs(zz)));
  if (m < 1e-30) { *st = 0.0f; *bl = 0.0f; return; }
  const double nrm = xy * xy + xz * xz + yz * yz;
  double e0, e1, e2;
  if (nrm > 1e-30) {
    const double t = (xx + yy + zz) / 3.0;
    const double b00 = xx - t, b11 = yy - t, b22 = zz - t;
    const double den = sqrt((b00 * b00 + b11 * b11 + b22 * b22 + nrm * 2.0) / 6.0);
    const double c00 = b11 * b22 - yz * yz;
    const double c01 = xy * b22 - yz * xz;
    const double c02 = xy * yz - b11 * xz;
    double half = (b00 * c00 - xy * c01 + xz * c02) / (den * den * den) * 0.5;
    half = fmin(fmax(half, -1.0), 1.0);
    const double ang = acos(half) / 3.0;
    const double beta2 = cos(ang) * 2.0;
    const double beta0 = cos(ang + 2.094395102393195) * 2.0;
    const double beta1 = -(beta0 + beta2);
    e0 = t + den * beta0;
    e1 = t + den * beta1;
    e2 = t + den * beta2;
  } else {
    e0 = xx; e1 = yy; e2 = zz;
  }
  *st = (float)(e0 * e0 + e1 * e1 + e2 * e2);
  *bl = (float)(e2 * e2 / (e0 * e1));
}

__global__ __launch_bounds__(NT) void k_finalize(const RegionAcc* __restrict__ acc, i64 R,
                                                 int is_ave, RegionOut o, i64 pitch,
                                                 const TileOut* __restrict__ tile) {
  i64 base = 0, first = 0;
  if (tile) {
    if (tile->overflow) return;
    R = (i64)tile->count + 1;
    base = (i64)tile->offset - 1;  // region r >= 1 goes to slot offset + r - 1
    first = 1;
  }
  for (i64 r = (i64)blockIdx.x * NT + threadIdx.x + first; r < R; r += (i64)gridDim.x * NT) {
    if (r == 0) {
      if (o.label) o.label[0] = 0;
      if (o.size) o.size[0] = 0;
      if (o.voxels) o.voxels[0] = 0;
      if (o.cx) { o.cx[0] = 0; o.cy[0] = 0; o.cz[0] = 0; }
      if (o.st) { o.st[0] = 0; o.bl[0] = 0; }
      if (o.nrm) o.nrm[0] = 0;
      if (o.hess) for (int k = 0; k < 6; ++k) o.hess[k * pitch] = 0;
      if (o.isum) o.isum[0] = 0;
      if (o.imax) o.imax[0] = 0;
      continue;
    }
    const RegionAcc& a = acc[r];
    const i64 q = base + r;
    const uint16_t s16 = (uint16_t)a.cnt;
    const float fs = (float)s16;
    if (o.label) o.label[q] = a.root;
    if (o.size) o.size[q] = s16;
    if (o.voxels) o.voxels[q] = a.cnt;
    if (o.cx) {
      o.cx[q] = (float)a.sx / fs;
      o.cy[q] = (float)a.sy / fs;
      o.cz[q] = (float)a.sz / fs;
    }
    float h6[6];
    for (int k = 0; k < 6; ++k) h6[k] = (float)a.h[k];
    if (o.hess) for (int k = 0; k < 6; ++k) o.hess[k * pitch + q] = h6[k];
    if (o.st) eigen_features(h6, o.st + q, o.bl + q);
    if (o.nrm) o.nrm[q] = is_ave ? (float)a.nsum : fdecode(a.nmax_key);
    if (o.isum) o.isum[q] = a.isum;
    if (o.imax) o.imax[q] = (uint16_t)a.imax;
  }
}

__global__ void k_tile_reserve(const uint32_t* __restrict__ d_R, i64 acc_cap, i64 arena_cap,
                               u64* running, TileOut* t) {
  if (threadIdx.x != 0) return;
  const i64 R = (i64)*d_R;
  const u64 off = *running;
  const bool over = R > acc_cap || (i64)off + (R - 1) > arena_cap;
  t->offset = off;
  t->count = over ? 0 : (u64)(R - 1);
  t->overflow = over ? 1u : 0u;
  t->found = (uint32_t)(R - 1);
  if (!over) *running = off + (u64)(R - 1);
}

__global__ __launch_bounds__(NT) void k_fill_u64(u64* p, i64 n, u64 v) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < n; i += (i64)gridDim.x * NT) p[i] = v;
}

}  // namespace

hipError_t launch_region_ids(const uint32_t* labels, uint32_t* rid, uint32_t* block_cnt,
                             uint32_t* d_R, i64 N, hipStream_t s, i64* nblocks_out) {
  const i64 nb = (N + BLK - 1) / BLK;
  *nblocks_out = nb;
  hipLaunchKernelGGL(k_count_roots, dim3((unsigned)nb), dim3(NT), 0, s, labels, N, block_cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark("regions/count_roots", s);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, block_cnt, nb, d_R);
  mark("regions/scan", s);
  return hipGetLastError();
}

hipError_t launch_region_assign(const uint32_t* labels, const uint32_t* block_off, uint32_t* rid,
                                RegionAcc* acc, i64 R, i64 cap, const uint32_t* d_R, i64 N,
                                hipStream_t s) {
  const i64 words = (d_R ? cap : std::min(R, cap)) * (i64)(sizeof(RegionAcc) / 8);
  const i64 zb = std::min<i64>((words + NT - 1) / NT, 1 << 16);
  mark(nullptr, s);  // the host waited for the region count: start a new interval
  hipLaunchKernelGGL(k_zero_acc, dim3((unsigned)std::max<i64>(zb, 1)), dim3(NT), 0, s, acc, R, d_R,
                     cap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark("regions/zero", s);
  const i64 nb = (N + BLK - 1) / BLK;
  hipLaunchKernelGGL(k_assign_ids, dim3((unsigned)nb), dim3(NT), 0, s, labels, N, block_off, rid,
                     acc, cap);
  mark("regions/assign_ids", s);
  return hipGetLastError();
}

template <typename TH, typename TN>
hipError_t launch_region_accum(const uint32_t* labels, const uint32_t* rid, const TH* hess,
                               const TN* norm, const uint16_t* img, RegionAcc* acc, i64 cap,
                               int W, int H, int D, int is_ave, hipStream_t s) {
  const int nbx = (W + 63) / 64, nby = (H + BY - 1) / BY, nbz = (D + BZ - 1) / BZ;
  const i64 nblk = (i64)nbx * nby * nbz;
  const i64 N = (i64)W * H * D;
  hipLaunchKernelGGL((k_accum<TH, TN>), dim3((unsigned)nblk), dim3(NT), 0, s, labels, rid, hess,
                     norm, img, acc, (uint32_t)std::min<i64>(cap, 0xFFFFFFFFll), W, H, D, N, nbx,
                     nby, is_ave);
  mark("regions/accum", s);
  return hipGetLastError();
}

#define HDOG_INST_ACC(TH, TN)                                                                  \
  template hipError_t launch_region_accum<TH, TN>(const uint32_t*, const uint32_t*, const TH*, \
                                                  const TN*, const uint16_t*, RegionAcc*, i64,  \
                                                  int, int, int, int, hipStream_t);
HDOG_INST_ACC(float, float)
HDOG_INST_ACC(__half, __half)
HDOG_INST_ACC(__hip_bfloat16, __hip_bfloat16)
#undef HDOG_INST_ACC

hipError_t launch_region_finalize(const RegionAcc* acc, i64 R, int is_ave, RegionOut o, i64 pitch,
                                  hipStream_t s, const TileOut* t) {
  const i64 nb = std::max<i64>(1, std::min<i64>((R + NT - 1) / NT, 1 << 16));
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)nb), dim3(NT), 0, s, acc, R, is_ave, o, pitch, t);
  mark("regions/finalize", s);
  return hipGetLastError();
}

hipError_t launch_tile_reserve(const uint32_t* d_R, i64 acc_cap, i64 arena_cap, u64* running,
                               TileOut* t, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_reserve, dim3(1), dim3(64), 0, s, d_R, acc_cap, arena_cap, running, t);
  return hipGetLastError();
}

hipError_t launch_fill_u64(u64* p, i64 n, u64 value, hipStream_t s) {
  const i64 nb = std::max<i64>(1, std::min<i64>((n + NT - 1) / NT, 1 << 16));
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)nb), dim3(NT), 0, s, p, n, value);
  return hipGetLastError();
}

}  // namespace hdog
