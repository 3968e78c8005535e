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
// pass that preserves the reference's ascending order.  Each work-group folds its voxels
// into an LDS table of partial sums (packed integer sums, FP64 Hessian and
// intensity-float sums, ordered-integer max for the float max); each table entry then
// ends in the region's 128-byte accumulator line, stored when the work-group holds the
// whole region, else added with hardware atomics (u64 counts and coordinate / intensity
// sums, exact; FP64 sums).  A finalize kernel turns the sums into the reference's output
// arrays.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;
constexpr int ITEMS = 16;
constexpr i64 BLK = (i64)NT * ITEMS;  // voxels per numbering block
static_assert(BLK == kRootBlock, "numbering block");

__device__ __forceinline__ int wave_sum_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;  // lane 0 holds the sum
}

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

// Per block of BLK voxels: the number of roots (L[i] == i, i > 0) and their positions as
// a bitmask (masks[64 b + 4 k + wave] bit lane <-> voxel b*BLK + k*NT + 64 wave + lane,
// so bit order is linear order).
__global__ __launch_bounds__(NT) void k_count_roots(const uint32_t* __restrict__ L, i64 N,
                                                    uint32_t* __restrict__ cnt,
                                                    unsigned long long* __restrict__ masks) {
  __shared__ int sh[4];
  const i64 base = (i64)blockIdx.x * BLK;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t v[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const i64 i = base + (i64)k * NT + threadIdx.x;
    v[k] = i < N ? L[i] : 0u;
  }
  int c = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const i64 i = base + (i64)k * NT + threadIdx.x;
    const bool root = i > 0 && i < N && v[k] == (uint32_t)i;
    const unsigned long long m = __ballot(root);
    if (lane == 0) masks[(size_t)blockIdx.x * (BLK / 64) + k * (NT / 64) + wave] = m;
    c += root ? 1 : 0;
  }
  const int t = block_sum_int(c, sh);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)t;
}

// Exclusive scan of cnt[0..nb) in place, *d_R = 1 + total (totals stay below 2^32 as
// N < 2^32), in three short kernels: the sums of chunks of 4096 counts (one work-group of
// 1024 threads, 4 consecutive counts each), the exclusive scan of the <= 256 chunk sums
// (one work-group), and every chunk's scan from its offset.
constexpr int SCAN_V = 4, SCAN_CH = 1024 * SCAN_V;

// inclusive scan of x over the 1024 threads of the work-group; *total = the sum
__device__ __forceinline__ uint32_t block_scan_1024(uint32_t x, uint32_t* wsum,
                                                    uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const uint32_t y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < 16) wsum[lane] = w;
  }
  __syncthreads();
  *total = wsum[15];
  return x + (wave ? wsum[wave - 1] : 0u);
}

__global__ __launch_bounds__(1024) void k_chunk_sums(const uint32_t* __restrict__ cnt, i64 nb,
                                                     uint32_t* __restrict__ csum) {
  __shared__ uint32_t wsum[16];
  const i64 i0 = (i64)blockIdx.x * SCAN_CH + (i64)threadIdx.x * SCAN_V;
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j) v += (i0 + j < nb) ? cnt[i0 + j] : 0u;
  uint32_t total;
  (void)block_scan_1024(v, wsum, &total);
  if (threadIdx.x == 0) csum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_chunk_offsets(uint32_t* csum, int nch,
                                                        uint32_t* d_R) {
  __shared__ uint32_t wsum[16];
  const uint32_t v = (int)threadIdx.x < nch ? csum[threadIdx.x] : 0u;
  uint32_t total;
  const uint32_t incl = block_scan_1024(v, wsum, &total);
  if ((int)threadIdx.x < nch) csum[threadIdx.x] = incl - v;
  if (threadIdx.x == 0) *d_R = 1u + total;
}

__global__ __launch_bounds__(1024) void k_chunk_scan(uint32_t* cnt, i64 nb,
                                                     const uint32_t* __restrict__ coff) {
  __shared__ uint32_t wsum[16];
  const i64 i0 = (i64)blockIdx.x * SCAN_CH + (i64)threadIdx.x * SCAN_V;
  uint32_t cur[SCAN_V], sum = 0;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j) {
    cur[j] = (i0 + j < nb) ? cnt[i0 + j] : 0u;
    sum += cur[j];
  }
  uint32_t total;
  uint32_t run = coff[blockIdx.x] + block_scan_1024(sum, wsum, &total) - sum;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j) {
    if (i0 + j < nb) cnt[i0 + j] = run;
    run += cur[j];
  }
}

// rid[root] = 1 + rank of the root in linear order; acc[rid] = zero sums with .root =
// root (a whole 128-byte line per region, which also clears the accumulator; block 0
// clears acc[0], the background's).  One wave per block of BLK voxels: lane m takes root
// mask m of the block, a wave scan of the mask popcounts gives each mask's rank offset,
// and only the (few) roots are visited.
__global__ __launch_bounds__(64) void k_assign_ids(const unsigned long long* __restrict__ masks,
                                                   const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ rid,
                                                   RegionAcc* __restrict__ acc, i64 cap) {
  static_assert(BLK / 64 == 64, "one mask per lane");
  const int lane = threadIdx.x;
  unsigned long long m = masks[(size_t)blockIdx.x * 64 + lane];
  const uint32_t n = (uint32_t)__popcll(m);
  uint32_t incl = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  uint32_t r = 1 + off[blockIdx.x] + (incl - n);
  const i64 base = (i64)blockIdx.x * BLK + 64 * lane;
  if (blockIdx.x == 0 && lane == 0 && cap > 0) acc[0] = RegionAcc{};
  while (m) {
    const int bit = __builtin_ctzll(m);
    m &= m - 1;
    const i64 i = base + bit;
    rid[i] = r;
    if ((i64)r < cap) {
      RegionAcc a{};
      a.root = (uint32_t)i;
      acc[r] = a;
    }
    ++r;
  }
}

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdecode(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

constexpr uint32_t EMPTY = 0xFFFFFFFFu;

// Partial sums of one region, then one set of global atomics per region and work-group.
struct RunSums {
  uint32_t cnt, imax, nkey;
  unsigned long long sx, sy, sz, isum;
  double h[6], ns;
};

__device__ __forceinline__ void flush_global(RegionAcc* a, const RunSums& v, int is_ave,
                                             bool has_img) {
  atomicAdd(&a->cnt, v.cnt);
  atomicAdd(&a->sx, v.sx);
  atomicAdd(&a->sy, v.sy);
  atomicAdd(&a->sz, v.sz);
#pragma unroll
  for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&a->h[k], v.h[k]);
  if (is_ave)
    unsafeAtomicAdd(&a->nsum, v.ns);
  else
    atomicMax(&a->nmax_key, v.nkey);
  if (has_img) {
    atomicAdd(&a->isum, v.isum);
    atomicMax(&a->imax, v.imax);
  }
}

// A region whose voxels all lie in one work-group's tiles: its sums are final after the
// work-group's fold and are stored, not added (the accumulator is zeroed beforehand).
__device__ __forceinline__ void store_final(RegionAcc* a, const RunSums& v) {
  a->sx = v.sx;
  a->sy = v.sy;
  a->sz = v.sz;
  a->isum = v.isum;
#pragma unroll
  for (int k = 0; k < 6; ++k) a->h[k] = v.h[k];
  a->nsum = v.ns;
  a->cnt = v.cnt;
  a->nmax_key = v.nkey;
  a->imax = v.imax;
}

// Per-work-group hash table in LDS keyed by region id (open addressing, LDS atomics, FP64
// for the Hessian and normalized sums).  Each occupied slot ends as one set of global
// atomics, so a region pays one set per work-group that touches it instead of one per
// voxel; a voxel that finds the table full goes straight to the global atomics.  A slot
// whose region the caller marks `owned` (no other work-group holds any of its voxels) is
// stored instead, unless some voxel of the work-group found the table full.
// Coordinates are summed relative to the work-group's origin (ox, oy, oz), with the count
// and the x and y sums packed in one 64-bit word (count: bits 0-15, x: 16-39, y: 40-63)
// and the z sums with the raw intensity sums in another (z: bits 0-23, intensity: 24-63):
// two LDS atomics instead of five.  For the < 2^16 voxels of a work-group whose offsets
// stay below 256 in x, y and z the fields cannot carry into each other.
template <int S>
struct LdsTable {
  uint32_t key[S];
  unsigned long long cxy[S], zis[S];
  uint32_t im[S], nk[S], own[S];
  double h[6][S], ns[S];
  uint32_t spilled, ox, oy, oz;

  __device__ void init(uint32_t x0, uint32_t y0, uint32_t z0) {
    if (threadIdx.x == 0) {
      spilled = 0u;
      ox = x0;
      oy = y0;
      oz = z0;
    }
    for (int t = threadIdx.x; t < S; t += blockDim.x) {
      key[t] = EMPTY;
      cxy[t] = zis[t] = 0ull;
      im[t] = nk[t] = own[t] = 0u;
#pragma unroll
      for (int k = 0; k < 6; ++k) h[k][t] = 0.0;
      ns[t] = 0.0;
    }
  }

  // sums of some voxels of one region: count, coordinate offsets from the origin, raw
  // intensity sum and max, max key (or sum) of the normalized value, Hessian sums
  struct Part {
    uint32_t cnt, dx, dy, dz, isum, imax, nkey;
    double ns, h[6];
  };

  __device__ void add(RegionAcc* acc, uint32_t r, uint32_t x, uint32_t y, uint32_t z,
                      const float* hv, float nv, uint32_t iv, int is_ave, bool has_img,
                      bool owned = false) {
    Part q;
    q.cnt = 1u;
    q.dx = x - ox;
    q.dy = y - oy;
    q.dz = z - oz;
    q.isum = q.imax = iv;
    q.nkey = is_ave ? 0u : fkey(nv);
    q.ns = is_ave ? (double)nv : 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) q.h[k] = (double)hv[k];
    add_part(acc, r, q, is_ave, has_img, owned);
  }

  __device__ void add_part(RegionAcc* acc, uint32_t r, const Part& q, int is_ave, bool has_img,
                           bool owned) {
    uint32_t slot = (r * 2654435761u) >> (32 - __builtin_ctz(S));
    bool placed = false;
    for (int probes = 0; probes < S; ++probes) {
      // a key, once set, never changes: a plain read finds the region's slot, and only an
      // empty slot needs the compare-and-swap
      uint32_t k0 = __atomic_load_n(&key[slot], __ATOMIC_RELAXED);
      if (k0 == EMPTY) k0 = atomicCAS(&key[slot], EMPTY, r);
      if (k0 == EMPTY || k0 == r) { placed = true; break; }
      slot = (slot + 1) & (S - 1);
    }
    if (!placed) {
      RunSums v;
      v.cnt = q.cnt;
      v.sx = q.dx + (unsigned long long)q.cnt * ox;
      v.sy = q.dy + (unsigned long long)q.cnt * oy;
      v.sz = q.dz + (unsigned long long)q.cnt * oz;
      v.isum = q.isum;
      v.imax = q.imax;
      v.nkey = q.nkey;
      v.ns = q.ns;
#pragma unroll
      for (int k = 0; k < 6; ++k) v.h[k] = q.h[k];
      flush_global(acc + r, v, is_ave, has_img);
      spilled = 1u;
      return;
    }
    if (owned) own[slot] = 1u;  // the same value from every voxel of the region
    atomicAdd(&cxy[slot], (unsigned long long)q.cnt | ((unsigned long long)q.dx << 16) |
                              ((unsigned long long)q.dy << 40));
    atomicAdd(&zis[slot], (unsigned long long)q.dz | ((unsigned long long)q.isum << 24));
#pragma unroll
    for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&h[k][slot], q.h[k]);
    if (is_ave)
      unsafeAtomicAdd(&ns[slot], q.ns);
    else
      atomicMax(&nk[slot], q.nkey);
    if (has_img) atomicMax(&im[slot], q.imax);
  }

  __device__ void flush(RegionAcc* acc, int is_ave, bool has_img) {
    for (int t = threadIdx.x; t < S; t += blockDim.x) {
      const uint32_t rr = key[t];
      if (rr == EMPTY) continue;
      RunSums v;
      const unsigned long long p = cxy[t];
      v.cnt = (uint32_t)(p & 0xFFFFull);
      v.sx = ((p >> 16) & 0xFFFFFFull) + (unsigned long long)v.cnt * ox;
      v.sy = (p >> 40) + (unsigned long long)v.cnt * oy;
      const unsigned long long q = zis[t];
      v.sz = (q & 0xFFFFFFull) + (unsigned long long)v.cnt * oz;
      v.isum = q >> 24;
      v.imax = im[t]; v.nkey = nk[t]; v.ns = ns[t];
#pragma unroll
      for (int k = 0; k < 6; ++k) v.h[k] = h[k][t];
      if (own[t] && !spilled)
        store_final(acc + rr, v);
      else
        flush_global(acc + rr, v, is_ave, has_img);
    }
  }
};

// Region sums from the materialized Hessian (the per-stage interface).  A work-group takes
// a brick of 64 x BY x BZ voxels (one row of 64 x per wave and z step); all loads of the
// brick are issued up front, then every candidate voxel folds itself into the table
// (128 slots: at most 128 separate 26-connected regions fit in a 64 x 4 x 4 brick).
constexpr int BY = 4, BZ = 4;

template <typename TH, typename TN>
__global__ __launch_bounds__(NT) void k_accum(const uint32_t* __restrict__ L,
                                              const uint32_t* __restrict__ rid,
                                              const TH* __restrict__ hess,
                                              const TN* __restrict__ norm,
                                              const uint16_t* __restrict__ img, RegionAcc* acc,
                                              uint32_t cap, int W, int H, int D, i64 N, int nbx,
                                              int nby, int is_ave) {
  __shared__ LdsTable<128> tab;
  unsigned b = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x = bx * 64 + (threadIdx.x & 63);
  const int y = by * BY + (threadIdx.x >> 6);
  const int zb = (int)b * BZ;
  tab.init((uint32_t)(bx * 64), (uint32_t)(by * BY), (uint32_t)zb);
  const i64 WH = (i64)W * H;
  const bool has_img = img != nullptr;
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
  for (int zz = 0; zz < BZ; ++zz)
    if (l[zz] && r[zz] < cap)
      tab.add(acc, r[zz], (uint32_t)x, (uint32_t)y, (uint32_t)(zb + zz), hv[zz], nv[zz], iv[zz],
              is_ave, has_img);
  __syncthreads();
  tab.flush(acc, is_ave, has_img);
}

// Region sums from the candidate records of the Hessian kernel: one work-group per
// kAccGY x kAccGZ Hessian tiles folds their records (voxel index + six elements, SoA) into
// the table, gathering the label, normalized and raw intensity of each candidate voxel; no
// pass over the non-candidate voxels.
static_assert(kAccGY * kAccGZ * kCandTile < (1 << 16) && kHessTX <= 256 &&
                  kAccGY * kHessTY <= 256 && kAccGZ * kHessTZ <= 256,
              "packed table sums: count below 2^16, offsets below 256");
// v of lane i + 1 (lane 63 keeps its own) and of lane i - 1 (lane 0 keeps its own); one
// DPP move each.  The whole wave must be active.
__device__ __forceinline__ uint32_t lane_next(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t lane_prev(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x138, 0xf, 0xf, false);
}

template <typename TN>
__global__ __launch_bounds__(NT) void k_accum_cand(const float* __restrict__ cand,
                                                   const uint32_t* __restrict__ cand_cnt,
                                                   const uint32_t* __restrict__ rlab,
                                                   const unsigned long long* __restrict__ multi,
                                                   const uint32_t* __restrict__ rid,
                                                   const TN* __restrict__ norm,
                                                   const uint16_t* __restrict__ img,
                                                   RegionAcc* acc, uint32_t cap, uint32_t W,
                                                   uint32_t WH, int is_ave, unsigned nbx,
                                                   unsigned nby, unsigned nbz) {
  // work-group p takes the tiles (bx, kAccGY gy + a, kAccGZ gz + b) of the Hessian tiling;
  // a region within them is final here, others pay one set of global atomics
  constexpr int TPW = kAccGY * kAccGZ;
  const int lane = (int)(threadIdx.x & 63);
  const unsigned p = xcd_tile(blockIdx.x, gridDim.x);
  const unsigned ngy = (nby + kAccGY - 1) / kAccGY;
  const unsigned bx = p % nbx, gy = (p / nbx) % ngy, gz = p / nbx / ngy;
  uint32_t n[TPW];
  unsigned tiles[TPW], tby[TPW], tbz[TPW];
  uint32_t any = 0;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const unsigned by = gy * kAccGY + q % kAccGY, bz = gz * kAccGZ + q / kAccGY;
    tiles[q] = (bz * nby + by) * nbx + bx;
    tby[q] = by;
    tbz[q] = bz;
    n[q] = (by < nby && bz < nbz) ? cand_cnt[tiles[q]] : 0u;
    any |= n[q];
  }
  if (any == 0) return;  // uniform over the work-group
  __shared__ LdsTable<256> tab;
  tab.init(bx * kHessTX, gy * kAccGY * kHessTY, gz * kAccGZ * kHessTZ);
  const bool has_img = img != nullptr;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const float* t = cand + (size_t)tiles[q] * kCandTile * 7;
    const uint32_t* rl = rlab + (size_t)tiles[q] * kCandTile;  // each record's root label
    // tile origin: records hold the voxel's position in the tile
    const uint32_t tx0 = bx * kHessTX, ty0 = tby[q] * kHessTY, tz0 = tbz[q] * kHessTZ;
    // the whole wave walks the records together (the group sums below move values across
    // lanes); lanes past the tile's records load nothing and add nothing
    for (uint32_t b = threadIdx.x - lane; b < n[q]; b += NT) {
      const uint32_t i = b + lane;
      const bool ok = i < n[q];
      uint32_t r = 0u, x = 0u, y = 0u, z = 0u, iv = 0u;
      float hv[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, nv = 0.0f;
      bool owned = false;
      if (ok) {
        const uint32_t v = __float_as_uint(t[i]);
        x = tx0 + (v % kHessTX);
        y = ty0 + (v / kHessTX) % kHessTY;
        z = tz0 + v / (kHessTX * kHessTY);
        const uint32_t idx = z * WH + y * W + x;
#pragma unroll
        for (int k = 0; k < 6; ++k) hv[k] = t[(size_t)kCandTile * (1 + k) + i];
        const uint32_t root = rl[i];
        r = rid[root];
        // the labeling marks every region that reaches beyond one work-group's tiles
        owned = !((multi[root >> 6] >> (root & 63)) & 1ull);
        nv = ld(norm, (i64)idx);
        iv = has_img ? (uint32_t)img[idx] : 0u;
      }
      // Consecutive records are mostly x-neighbours in one row of one region.  Runs of
      // lanes with the same region are cut into groups of at most four; the first lane of
      // a group adds the group's sums, so the table sees one set of atomics per group.
      // (the shifts run in every lane, outside any condition: a lane whose source lane is
      // inactive would get its own value back)
      const uint32_t r_prev = lane_prev(r);
      const uint32_t ok_prev = lane_prev(ok ? 1u : 0u);
      const bool run_head = (lane == 0) | !ok | (ok_prev == 0u) | (r_prev != r);
      const unsigned long long heads = __ballot(run_head);
      const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
      const int start = 63 - __builtin_clzll(heads & upto);  // first lane of the run
      const unsigned long long later = heads & ~upto;
      const int next = later ? __builtin_ctzll(later) : 64;  // first lane of the next run
      const int last = min(next - 1, lane + 3);  // last lane of a group that starts here
      typename LdsTable<256>::Part g;
      g.cnt = 1u;
      g.dx = x - tab.ox;
      g.dy = y - tab.oy;
      g.dz = z - tab.oz;
      g.isum = g.imax = iv;
      g.nkey = is_ave ? 0u : fkey(nv);
      g.ns = is_ave ? (double)nv : 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) g.h[k] = (double)hv[k];
      // the values of lanes +1, +2, +3 arrive by one-lane shifts; those past `last` are
      // left out (a shift keeps lane 63's own value, which is never inside its group)
      uint32_t sdx = g.dx, sdy = g.dy, sdz = g.dz, siv = iv, snk = g.nkey;
      float snv = nv, sh[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) sh[k] = hv[k];
#pragma unroll
      for (int o = 1; o <= 3; ++o) {
        sdx = lane_next(sdx);
        sdy = lane_next(sdy);
        sdz = lane_next(sdz);
        siv = lane_next(siv);
        if (is_ave)
          snv = __uint_as_float(lane_next(__float_as_uint(snv)));
        else
          snk = lane_next(snk);
#pragma unroll
        for (int k = 0; k < 6; ++k) sh[k] = __uint_as_float(lane_next(__float_as_uint(sh[k])));
        if (lane + o <= last) {
          g.cnt += 1u;
          g.dx += sdx;
          g.dy += sdy;
          g.dz += sdz;
          g.isum += siv;
          g.imax = max(g.imax, siv);
          if (is_ave)
            g.ns += (double)snv;
          else
            g.nkey = max(g.nkey, snk);
#pragma unroll
          for (int k = 0; k < 6; ++k) g.h[k] += (double)sh[k];
        }
      }
      const bool head = ((lane - start) & 3) == 0;
      if (ok && head && r < cap) tab.add_part(acc, r, g, is_ave, has_img, owned);
    }
  }
  __syncthreads();
  tab.flush(acc, is_ave, has_img);
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

// The arena counter is advanced by compare-and-swap: stacks of a batch may reserve their
// ranges concurrently (the branches of a batch graph).
__global__ void k_tile_reserve(const uint32_t* __restrict__ d_R, i64 acc_cap, i64 arena_cap,
                               u64* running, TileOut* t) {
  if (threadIdx.x != 0) return;
  const i64 R = (i64)*d_R;
  u64 off = __hip_atomic_load(running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool over;
  for (;;) {
    over = R > acc_cap || (i64)off + (R - 1) > arena_cap;
    if (over) break;
    const u64 seen = atomicCAS((unsigned long long*)running, (unsigned long long)off,
                               (unsigned long long)(off + (u64)(R - 1)));
    if (seen == off) break;
    off = seen;  // another stack reserved first: retry from its end
  }
  t->offset = off;
  t->count = over ? 0 : (u64)(R - 1);
  t->overflow = over ? 1u : 0u;
  t->found = (uint32_t)(R - 1);
}

__global__ __launch_bounds__(NT) void k_fill_u64(u64* p, i64 n, u64 v) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < n; i += (i64)gridDim.x * NT) p[i] = v;
}

__global__ __launch_bounds__(NT) void k_fill_u32(uint32_t* p, i64 n, uint32_t v) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < n; i += (i64)gridDim.x * NT) p[i] = v;
}

// roots per 4096-voxel block from the root bit mask (one wave per block, a word per lane)
__global__ __launch_bounds__(NT) void k_mask_counts(const unsigned long long* __restrict__ masks,
                                                    i64 nb, uint32_t* __restrict__ cnt) {
  const i64 blk = (i64)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (blk >= nb) return;  // whole waves
  const int c = __popcll(masks[blk * (kRootBlock / 64) + (threadIdx.x & 63)]);
  const int t = wave_sum_int(c);
  if ((threadIdx.x & 63) == 0) cnt[blk] = (uint32_t)t;
}

}  // namespace

// block_cnt holds block_count_words(N) words: the nb counts, then the chunk sums
hipError_t launch_scan_counts(uint32_t* block_cnt, i64 nb, uint32_t* d_R, hipStream_t s) {
  const int nch = (int)((nb + SCAN_CH - 1) / SCAN_CH);
  if (nch > 1024) return hipErrorInvalidValue;  // N < 2^32: at most 256 chunks
  uint32_t* csum = block_cnt + scan_chunk_offset(nb);
  hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)nch), dim3(1024), 0, s, block_cnt, nb, csum);
  hipLaunchKernelGGL(k_chunk_offsets, dim3(1), dim3(1024), 0, s, csum, nch, d_R);
  hipLaunchKernelGGL(k_chunk_scan, dim3((unsigned)nch), dim3(1024), 0, s, block_cnt, nb,
                     (const uint32_t*)csum);
  mark("regions/scan", s);
  return hipGetLastError();
}

hipError_t launch_region_ids(const uint32_t* labels, uint32_t* block_cnt,
                             unsigned long long* masks, uint32_t* d_R, i64 N, hipStream_t s,
                             i64* nblocks_out) {
  const i64 nb = (N + BLK - 1) / BLK;
  *nblocks_out = nb;
  hipLaunchKernelGGL(k_count_roots, dim3((unsigned)nb), dim3(NT), 0, s, labels, N, block_cnt,
                     masks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark("regions/count_roots", s);
  return launch_scan_counts(block_cnt, nb, d_R, s);
}

hipError_t launch_region_assign(const unsigned long long* masks, const uint32_t* block_off,
                                uint32_t* rid, RegionAcc* acc, i64 cap, i64 N, hipStream_t s) {
  mark(nullptr, s);  // the host waited for the region count: start a new interval
  const i64 nb = (N + BLK - 1) / BLK;
  hipLaunchKernelGGL(k_assign_ids, dim3((unsigned)nb), dim3(64), 0, s, masks, block_off, rid, acc,
                     cap);
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

template <typename TN>
hipError_t launch_region_accum_cand(const float* cand, const uint32_t* cand_cnt, i64 ntiles,
                                    const uint32_t* rlab, const unsigned long long* multi,
                                    const uint32_t* rid, const TN* norm, const uint16_t* img,
                                    RegionAcc* acc, i64 cap, int W, int H, int is_ave,
                                    hipStream_t s) {
  const unsigned nbx = (unsigned)((W + kHessTX - 1) / kHessTX),
                 nby = (unsigned)((H + kHessTY - 1) / kHessTY);
  const unsigned nbz = (unsigned)(ntiles / ((i64)nbx * nby));
  const i64 ngroups =
      (i64)nbx * ((nby + kAccGY - 1) / kAccGY) * ((nbz + kAccGZ - 1) / kAccGZ);
  hipLaunchKernelGGL((k_accum_cand<TN>), dim3((unsigned)ngroups), dim3(NT), 0, s, cand, cand_cnt,
                     rlab, multi, rid, norm, img, acc, (uint32_t)std::min<i64>(cap, 0xFFFFFFFFll),
                     (uint32_t)W, (uint32_t)((i64)W * H), is_ave, nbx, nby, nbz);
  mark("regions/accum_candidates", s);
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
#define HDOG_INST_ACCC(TN)                                                                  \
  template hipError_t launch_region_accum_cand<TN>(                                         \
      const float*, const uint32_t*, i64, const uint32_t*, const unsigned long long*,         \
      const uint32_t*, const TN*, const uint16_t*, RegionAcc*, i64, int, int, int, hipStream_t);
HDOG_INST_ACCC(float)
HDOG_INST_ACCC(__half)
HDOG_INST_ACCC(__hip_bfloat16)
#undef HDOG_INST_ACCC

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

hipError_t launch_fill_u32(uint32_t* p, i64 n, uint32_t value, hipStream_t s) {
  const i64 nb = std::max<i64>(1, std::min<i64>((n + NT - 1) / NT, 1 << 16));
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nb), dim3(NT), 0, s, p, n, value);
  return hipGetLastError();
}

hipError_t launch_region_count_masks(const unsigned long long* masks, uint32_t* block_cnt,
                                     uint32_t* d_R, i64 N, hipStream_t s, i64* nblocks_out) {
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  if (nblocks_out) *nblocks_out = nb;
  hipLaunchKernelGGL(k_mask_counts, dim3((unsigned)((nb + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                     s, masks, nb, block_cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark("regions/count_root_bits", s);
  return launch_scan_counts(block_cnt, nb, d_R, s);
}

hipError_t launch_fill_u64(u64* p, i64 n, u64 value, hipStream_t s) {
  const i64 nb = std::max<i64>(1, std::min<i64>((n + NT - 1) / NT, 1 << 16));
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)nb), dim3(NT), 0, s, p, n, value);
  return hipGetLastError();
}

}  // namespace hdog
