// ccl.hip -- 3-D connected-component labelling on gfx950.
//
// Reference semantics: src/3D/CCL3D.cu:17-117 iterates scan / pointer-jump passes,
// polling a device flag from the host, until every foreground voxel holds the minimum
// linear index of its component.  Adjacency is the set of linear offsets
// dz*W*H + dy*W + dx kept when the target is inside [0, N), so rows and slices wrap;
// voxel 0 is never foreground (label = index * mask).
//
// MI355X design: no host round trips.  Union-find over the same adjacency in three
// kernels: init (label = own index), union (each foreground voxel links with its
// foreground neighbours at the distinct positive offsets; roots are linked larger-under-
// smaller with atomicMin, so the root of every tree is the component minimum), and
// flatten (every voxel points at its root).  Each edge is visited once.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

struct Offsets {
  i64 o[13];
  int n;
};

// Parents only ever decrease, so every chain ends at a root.  Plain (cacheable) loads are
// enough: a stale parent is still an ancestor, two voxels that ever shared an ancestor are
// in the same set, and every link goes through an atomicMin whose returned value is
// current, so each retry of unite strictly lowers `a` and the loop ends.  Path halving
// (x's parent replaced by its grandparent) keeps the trees shallow; the plain store only
// ever replaces a parent by one of its ancestors, so no union is lost.
// Only the union kernel may halve paths: in the flatten kernel a halving store could
// overwrite another voxel's final root with an intermediate ancestor.
__device__ __forceinline__ uint32_t find_root(uint32_t* L, uint32_t x) {
  uint32_t p = L[x];
  while (p != x) {
    const uint32_t gp = L[p];
    if (gp != p) L[x] = gp;
    x = p;
    p = gp;
  }
  return x;
}

// links the sets of a and b; returns the root of the union at the time of the link
__device__ uint32_t unite(uint32_t* L, uint32_t a, uint32_t b) {
  while (true) {
    a = find_root(L, a);
    b = find_root(L, b);
    if (a == b) return a;
    if (a < b) { const uint32_t t = a; a = b; b = t; }
    const uint32_t old = atomicMin(L + a, b);
    if (old == a) return b;
    a = old;  // a had been linked meanwhile: continue with its new parent
  }
}

__global__ __launch_bounds__(NT) void k_ccl_init(const uint8_t* __restrict__ pd,
                                                 uint32_t* __restrict__ L, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT)
    L[i] = (pd[i] && i > 0) ? (uint32_t)i : 0u;
}

__global__ __launch_bounds__(NT) void k_ccl_union(const uint8_t* __restrict__ pd, uint32_t* L,
                                                  i64 N, Offsets offs) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    if (i == 0 || !pd[i]) continue;
    uint32_t root = (uint32_t)i;  // a representative of i's set, refreshed by each link
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      if (k < offs.n) {
        const i64 j = i + offs.o[k];
        if (j < N && pd[j]) {
          // a neighbour whose parent already is our representative is in our set
          if (L[j] != root) root = unite(L, root, (uint32_t)j);
        }
      }
    }
  }
}

__global__ __launch_bounds__(NT) void k_ccl_flatten(uint32_t* L, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const uint32_t l = L[i];
    if (l && l != (uint32_t)i) L[i] = find_root_readonly(L, l);
  }
}

// ---- tiled form (width >= 2 and height >= 2) ----------------------------------------
// Kernel 1 labels each 32 x 8 x 8 tile on its own with a union-find in LDS (links of
// voxel pairs whose neighbour lies inside the tile and inside the volume, i.e. without
// wrap); every foreground voxel then points at the global index of its tile-local root,
// which is the tile-local minimum because local order follows global order.  Kernel 2
// links, with the global union-find, only the pairs that leave the tile or wrap.  For
// width, height >= 2 the 13 forward neighbours (dz = 1; dz = 0, dy = 1; dz = dy = 0,
// dx = 1) are exactly the positive linear offsets of the reference's adjacency.
constexpr int CX = 32, CY = 8, CZ = 8, CV = CX * CY * CZ, NTC = 256, CPT = CV / NTC;
static_assert(CV % NTC == 0, "tile shape");

__device__ __forceinline__ uint32_t lds_find(const uint32_t* lab, uint32_t x) {
  uint32_t p = lab[x];
  while (p != x) {
    x = p;
    p = lab[x];
  }
  return x;
}

// union phase only: path halving keeps the LDS trees shallow (same argument as
// find_root: a parent is only ever replaced by one of its ancestors)
__device__ __forceinline__ uint32_t lds_find_halve(uint32_t* lab, uint32_t x) {
  uint32_t p = lab[x];
  while (p != x) {
    const uint32_t gp = lab[p];
    if (gp != p) lab[x] = gp;
    x = p;
    p = gp;
  }
  return x;
}

__device__ __forceinline__ uint32_t lds_unite(uint32_t* lab, uint32_t a, uint32_t b) {
  while (true) {
    a = lds_find_halve(lab, a);
    b = lds_find_halve(lab, b);
    if (a == b) return a;
    if (a < b) { const uint32_t t = a; a = b; b = t; }
    const uint32_t old = atomicMin(lab + a, b);
    if (old == a) return b;
    a = old;
  }
}

// The same union-find on 16-bit parents (tile positions below 2^16): half the LDS, so
// twice the work-groups per CU.  LDS has no 16-bit atomic minimum; the link is a 32-bit
// compare-and-swap on the word holding the entry, which fails and retries when either
// half changed.  Path-halving stores stay plain 16-bit stores, as above.
__device__ __forceinline__ uint32_t lds16_find(const uint16_t* lab, uint32_t x) {
  uint32_t p = lab[x];
  while (p != x) {
    x = p;
    p = lab[x];
  }
  return x;
}
__device__ __forceinline__ uint32_t lds16_find_halve(uint16_t* lab, uint32_t x) {
  uint32_t p = lab[x];
  while (p != x) {
    const uint32_t gp = lab[p];
    if (gp != p) lab[x] = (uint16_t)gp;
    x = p;
    p = gp;
  }
  return x;
}
// lab[a] = min(lab[a], b); returns the old lab[a]
__device__ __forceinline__ uint32_t lds16_min(uint16_t* lab, uint32_t a, uint32_t b) {
  uint32_t* w = reinterpret_cast<uint32_t*>(lab + (a & ~1u));
  const int sh = (int)(a & 1u) * 16;
  uint32_t cur = *w;
  while (true) {
    const uint32_t old = (cur >> sh) & 0xFFFFu;
    if (old <= b) return old;
    const uint32_t prev = atomicCAS(w, cur, (cur & ~(0xFFFFu << sh)) | (b << sh));
    if (prev == cur) return old;
    cur = prev;
  }
}
__device__ __forceinline__ uint32_t lds16_unite(uint16_t* lab, uint32_t a, uint32_t b) {
  while (true) {
    a = lds16_find_halve(lab, a);
    b = lds16_find_halve(lab, b);
    if (a == b) return a;
    if (a < b) { const uint32_t t = a; a = b; b = t; }
    const uint32_t old = lds16_min(lab, a, b);
    if (old == a) return b;
    a = old;
  }
}

__device__ __forceinline__ void fwd_offset(int k, int& dx, int& dy, int& dz) {
  if (k < 9) { dz = 1; dy = k / 3 - 1; dx = k % 3 - 1; }
  else if (k < 12) { dz = 0; dy = 1; dx = k - 10; }
  else { dz = 0; dy = 0; dx = 1; }
}

// Redundant links.  Every 26-adjacent candidate pair is linked by its lower voxel (in
// the tile's local pass when both lie in the tile, else in the merge), so a voxel v can
// skip a foreground forward neighbour b that is adjacent to another forward neighbour w
// already known to be in v's set: b and w are both above v, so the pair (w, b) is linked
// by a voxel above v, and by induction from the top voxel down it ends in one set.  Two
// forward offsets are adjacent when they differ by at most 1 on every axis (the linear
// offset between the two voxels is then one of the reference's offsets, wrapped or not).
// (0, 0, +1) is adjacent to all other twelve, so it is visited first: with it present, one
// link replaces up to thirteen.
constexpr int fdx(int k) { return k < 9 ? k % 3 - 1 : k < 12 ? k - 10 : 1; }
constexpr int fdy(int k) { return k < 9 ? k / 3 - 1 : k < 12 ? 1 : 0; }
constexpr int fdz(int k) { return k < 9 ? 1 : 0; }
constexpr uint32_t fwd_adj(int q) {  // bit p: forward offsets p != q, q adjacent
  uint32_t m = 0;
  for (int p = 0; p < 13; ++p) {
    const int ax = fdx(p) - fdx(q), ay = fdy(p) - fdy(q), az = fdz(p) - fdz(q);
    if (p != q && ax >= -1 && ax <= 1 && ay >= -1 && ay <= 1 && az >= -1 && az <= 1)
      m |= 1u << p;
  }
  return m;
}
constexpr int fwd_order(int j) {
  return j == 0 ? 4 : j == 1 ? 10 : j <= 5 ? j - 2 : j <= 9 ? j - 1 : j == 10 ? 9 : j;
}
static_assert(fwd_adj(4) == 0x1FFFu - (1u << 4), "(0, 0, +1) neighbours all others");
static_assert(fwd_order(2) == 0 && fwd_order(5) == 3 && fwd_order(6) == 5 && fwd_order(9) == 8 &&
                  fwd_order(10) == 9 && fwd_order(11) == 11 && fwd_order(12) == 12,
              "every forward offset once");

// Kernel 1.  With 256 threads on a 32 x 8 x 8 tile, thread t owns the column (t % 32,
// t / 32) of the tile and k = 0..7 walks its z; every index is 32-bit (N < 2^32).  The
// foreground voxels are compacted into an LDS list (wave ballots), so the union phase runs
// one dense pass over the few foreground voxels instead of eight divergent passes over all.
static_assert(NTC == CX * CY && CPT == CZ, "k_ccl_local maps one thread to one z column");

__global__ __launch_bounds__(NTC) void k_ccl_local(const uint8_t* __restrict__ pd,
                                                  uint32_t* __restrict__ L, int W, int H, int D,
                                                  int nbx, int nby) {
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  __shared__ uint32_t lab[CV];
  __shared__ uint16_t list[CV];
  __shared__ uint32_t nfg;
  if (threadIdx.x == 0) nfg = 0;
  unsigned b = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x0 = bx * CX, y0 = by * CY, z0 = (int)b * CZ;
  const uint32_t WH = (uint32_t)W * (uint32_t)H;
  const int lx = threadIdx.x & (CX - 1), ly = threadIdx.x / CX;
  const int gx = x0 + lx, gy = y0 + ly;
  const bool col = gx < W && gy < H;
  const uint32_t g0 = (uint32_t)z0 * WH + (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
  const int lane = threadIdx.x & 63;
  uint8_t fg[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const uint32_t g = g0 + (uint32_t)k * WH;
    fg[k] = (col && z0 + k < D && g > 0) ? pd[g] : (uint8_t)0;
  }
  __syncthreads();  // nfg initialised
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int v = threadIdx.x + k * NTC;
    lab[v] = fg[k] ? (uint32_t)v : NONE;
    const unsigned long long bal = __ballot(fg[k] != 0);
    if (bal) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&nfg, (uint32_t)__popcll(bal));
      base = __shfl(base, 0, 64);
      if (fg[k]) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)v;
    }
  }
  __syncthreads();
  const int n = (int)nfg;
#pragma unroll 1
  for (int i = threadIdx.x; i < n; i += NTC) {
    const int v = list[i];
    const int vx = v & (CX - 1), vy = (v / CX) & (CY - 1), vz = v / (CX * CY);
    uint32_t root = (uint32_t)v;  // a representative of v's set, refreshed by each link
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      int dx, dy, dz;
      fwd_offset(q, dx, dy, dz);
      const int nx = vx + dx, ny = vy + dy, nz = vz + dz;
      if (nx < 0 || nx >= CX || ny < 0 || ny >= CY || nz >= CZ) continue;
      const int u = v + dx + CX * (dy + CY * dz);
      const uint32_t lu = lab[u];
      // outside-volume and background positions hold NONE; a parent equal to our
      // representative means u is already in our set
      if (lu != NONE && lu != root) root = lds_unite(lab, root, (uint32_t)u);
    }
  }
  __syncthreads();
  if (!col) return;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    if (z0 + k >= D) break;
    uint32_t out = 0u;
    if (fg[k]) {
      const uint32_t r = lds_find(lab, (uint32_t)(threadIdx.x + k * NTC));
      out = (uint32_t)(z0 + (int)(r / (CX * CY))) * WH +
            (uint32_t)(y0 + (int)((r / CX) & (CY - 1))) * (uint32_t)W + (uint32_t)x0 +
            (r & (CX - 1));
    }
    L[g0 + (uint32_t)k * WH] = out;
  }
}

// Kernel 2 visits only the voxels that can have a neighbour outside their tile or the
// volume: with the tile clipped to the volume as ex x ey x ez, its top z-face (lz = ez-1)
// and the rim of every lower plane (lx in {0, ex-1} or ly in {0, ey-1}); every other
// voxel's 13 forward neighbours lie inside the clipped tile.  Those voxels are numbered
// densely so that each wave carries only rim work.
__device__ __forceinline__ void rim_voxel(int t, int ex, int ey, int ez, int& lx, int& ly,
                                          int& lz) {
  const int top = ex * ey;
  if (t < top) {
    lz = ez - 1;
    ly = t / ex;
    lx = t - ly * ex;
    return;
  }
  t -= top;
  const int rows = (ey >= 2 ? 2 : 1) * ex, nc = ex >= 2 ? 2 : 1;
  const int ring = rows + (ey >= 2 ? ey - 2 : 0) * nc;
  lz = t / ring;
  const int r = t - lz * ring;
  if (r < rows) {
    ly = r < ex ? 0 : ey - 1;
    lx = r < ex ? r : r - ex;
  } else {
    const int r2 = r - rows;
    ly = 1 + r2 / nc;
    lx = (r2 % nc) ? ex - 1 : 0;
  }
}

__device__ __forceinline__ int rim_count(int ex, int ey, int ez) {
  const int rows = (ey >= 2 ? 2 : 1) * ex, nc = ex >= 2 ? 2 : 1;
  return ex * ey + (ez - 1) * (rows + (ey >= 2 ? ey - 2 : 0) * nc);
}

constexpr int NTM = 256;
constexpr int KR = 4;  // rim voxels per thread: a 32 x 8 x 8 tile has at most 788 <= KR * NTM
static_assert(32 * 8 + 7 * (2 * 32 + 6 * 2) <= KR * 256, "rim of a full tile fits one pass");

// Kernel 2.  Background voxels carry label 0 after kernel 1, so labels double as the
// mask.  Each distinct neighbour label is linked with the voxel's own label by the global
// union-find (every voxel of a tile-local component carries the same label, so few links
// remain).  Labels read here may already have been linked by other work-groups; any value
// read is an ancestor, so every link still joins the right sets.
// Round 1: the labels of up to KR rim voxels per thread (in flight together); the
// foreground ones (label != 0) are compacted into an LDS list.  Round 2: one dense pass
// over that list loads the labels of each voxel's neighbours outside the clipped tile and
// links every distinct one with the voxel's own label.
template <bool FULL>
__device__ __forceinline__ void merge_tile(uint32_t* L, int W, uint32_t WH, uint32_t N, int x0,
                                           int y0, int z0, int ex, int ey, int ez,
                                           uint16_t* list_t, uint32_t* list_own,
                                           uint32_t* n_list) {
  if (FULL) { ex = CX; ey = CY; ez = CZ; }  // compile-time shape: the rim math folds
  const int F = rim_count(ex, ey, ez);
  const int lane = threadIdx.x & 63;
  uint32_t own[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int t = threadIdx.x + k * NTM;
    own[k] = 0u;
    if (t < F) {
      int lx, ly, lz;
      rim_voxel(t, ex, ey, ez, lx, ly, lz);
      own[k] = L[(uint32_t)(z0 + lz) * WH + (uint32_t)(y0 + ly) * (uint32_t)W +
                 (uint32_t)(x0 + lx)];  // 0: background (voxel 0 is always background)
    }
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const unsigned long long bal = __ballot(own[k] != 0u);
    if (bal) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(n_list, (uint32_t)__popcll(bal));
      base = __shfl(base, 0, 64);
      if (own[k]) {
        const uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        list_t[slot] = (uint16_t)(threadIdx.x + k * NTM);
        list_own[slot] = own[k];
      }
    }
  }
  __syncthreads();
  const int n = (int)*n_list;
#pragma unroll 1
  for (int i = threadIdx.x; i < n; i += NTM) {
    int lx, ly, lz;
    rim_voxel(list_t[i], ex, ey, ez, lx, ly, lz);
    const uint32_t me = list_own[i];
    const uint32_t g = (uint32_t)(z0 + lz) * WH + (uint32_t)(y0 + ly) * (uint32_t)W +
                       (uint32_t)(x0 + lx);
    uint32_t lj[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      int dx, dy, dz;
      fwd_offset(q, dx, dy, dz);
      // forward linear offset (positive for W, H >= 2); wraps like the reference
      const i64 off64 = (i64)dz * WH + (i64)dy * W + dx;
      const uint32_t off = (uint32_t)off64;
      const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
      const bool local = nx >= 0 && nx < ex && ny >= 0 && ny < ey && nz < ez;
      lj[q] = (!local && off64 < (i64)N && g < N - off) ? L[g + off] : 0u;
    }
    uint32_t seen1 = 0u, seen2 = 0u;
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      const uint32_t l = lj[q];
      if (!l || l == me || l == seen1 || l == seen2) continue;
      unite(L, me, l);
      seen2 = seen1;
      seen1 = l;
    }
  }
}

__global__ __launch_bounds__(NTM) void k_ccl_merge(uint32_t* L, int W, int H, int D, int nbx,
                                                   int nby) {
  __shared__ uint16_t list_t[KR * NTM];
  __shared__ uint32_t list_own[KR * NTM];
  __shared__ uint32_t n_list;
  if (threadIdx.x == 0) n_list = 0;
  __syncthreads();
  unsigned b = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x0 = bx * CX, y0 = by * CY, z0 = (int)b * CZ;
  const int ex = min(CX, W - x0), ey = min(CY, H - y0), ez = min(CZ, D - z0);
  const uint32_t WH = (uint32_t)W * (uint32_t)H, N = WH * (uint32_t)D;
  if (ex == CX && ey == CY && ez == CZ)
    merge_tile<true>(L, W, WH, N, x0, y0, z0, ex, ey, ez, list_t, list_own, &n_list);
  else
    merge_tile<false>(L, W, WH, N, x0, y0, z0, ex, ey, ez, list_t, list_own, &n_list);
}

// ---- candidate-record form (the whole-path call) ------------------------------------
// The Hessian pass leaves, per 64 x 8 x 16 tile, the positions of its candidate voxels
// (the foreground of the CCL) in the tile.  Labeling then touches only candidates: P (a
// volume-sized parent array) is written and read at candidate voxels only, and the
// per-record labels go to rlab (slot-aligned with the records).
//   kernel A: per tile, the tile's candidate bit mask (one word per row, stored tile-major
//             for kernel B) and a union-find in LDS over the tile's candidates (links
//             inside the clipped tile, no wrap); rlab = the global index of v's tile-local
//             root, the tile-local minimum, also in P where P is read;
//   kernel B: per tile, the candidates on the tile's rim link each candidate neighbour
//             outside the clipped tile (by linear offset, wrapping like the reference) with
//             the global union-find on P;
//   kernel C: every record's root (the component minimum), root bits for the region
//             numbering, and with `dense` the final label written back into P.
// 256 threads and 16-bit LDS parents (17 KB of LDS per tile): eight tiles share a CU.
constexpr int RTX = kHessTX, RTY = kHessTY, RTZ = kHessTZ, RTV = RTX * RTY * RTZ, NTR = 256;
constexpr int RROWS = RTY * RTZ;  // mask words per tile
static_assert(RTV <= 65536, "tile positions fit the 16-bit parents");
// Each thread keeps the positions of its first KREC records (slots t + k NTR) in registers,
// loaded together: a tile's record list is then read once, with all loads in flight,
// instead of one dependent load per record and phase.  Tiles with more than KREC * NTR
// candidates take their remaining slots with a plain loop.
constexpr int KREC = 2048 / NTR;
struct TileRecs {
  uint32_t v[KREC];
  __device__ __forceinline__ void load(const float* rec, uint32_t n) {
#pragma unroll
    for (int k = 0; k < KREC; ++k) {
      const uint32_t i = threadIdx.x + k * NTR;
      v[k] = i < n ? __float_as_uint(rec[i]) : 0u;
    }
  }
  // f(slot, position) for every record of this thread
  template <typename F>
  __device__ __forceinline__ void each(const float* rec, uint32_t n, F&& f) const {
#pragma unroll
    for (int k = 0; k < KREC; ++k) {
      const uint32_t i = threadIdx.x + k * NTR;
      if (i < n) f(i, v[k]);
    }
#pragma unroll 1
    for (uint32_t i = threadIdx.x + KREC * NTR; i < n; i += NTR) f(i, __float_as_uint(rec[i]));
  }
  // the same visit with f at one call site (a rolled loop; the held positions rotate
  // down one register per record), for bodies too large to copy once per record
  template <typename F>
  __device__ __forceinline__ void each_rolled(const float* rec, uint32_t n, F&& f) const {
    uint32_t w[KREC];
#pragma unroll
    for (int k = 0; k < KREC; ++k) w[k] = v[k];
#pragma unroll 1
    for (uint32_t i = threadIdx.x; i < n; i += NTR) {
      f(i, i < (uint32_t)(KREC * NTR) ? w[0] : __float_as_uint(rec[i]));
#pragma unroll
      for (int k = 0; k + 1 < KREC; ++k) w[k] = w[k + 1];
    }
  }
};
static_assert(RTX == 64 && RTV == kCandTile, "a tile row is one mask word");

struct CandTile {
  int x0, y0, z0, ex, ey, ez;
};
__device__ __forceinline__ CandTile cand_tile(unsigned tile, int W, int H, int D, int nbx,
                                              int nby) {
  CandTile t;
  const unsigned r = tile / nbx;
  t.x0 = (int)(tile % nbx) * RTX;
  t.y0 = (int)(r % nby) * RTY;
  t.z0 = (int)(r / nby) * RTZ;
  t.ex = min(RTX, W - t.x0);
  t.ey = min(RTY, H - t.y0);
  t.ez = min(RTZ, D - t.z0);
  return t;
}
// mask word of voxel (x, y, z) in the tile-major layout; its bit is x % 64
__device__ __forceinline__ size_t mask_word(uint32_t x, uint32_t y, uint32_t z, int nbx,
                                            int nby) {
  const size_t tile = ((size_t)(z / RTZ) * nby + y / RTY) * nbx + x / RTX;
  return tile * RROWS + (z % RTZ) * RTY + y % RTY;
}

__global__ __launch_bounds__(NTR) void k_cand_local(const float* __restrict__ cand,
                                                   const uint32_t* __restrict__ cand_cnt,
                                                   unsigned long long* __restrict__ fgbits,
                                                   uint32_t* __restrict__ P,
                                                   uint32_t* __restrict__ rlab, int W, int H,
                                                   int D, int nbx, int nby, int links) {
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t n = cand_cnt[tile];
  __shared__ unsigned long long bits[RROWS];  // the tile's mask rows (word = ly + RTY lz)
  __shared__ uint16_t lab[RTV];              // parents, valid at candidate positions
  const float* rec = cand + (size_t)tile * kCandTile * 7;
  TileRecs R;
  R.load(rec, n);
  if (threadIdx.x < RROWS) bits[threadIdx.x] = 0ull;
  __syncthreads();
  R.each(rec, n, [&](uint32_t, uint32_t v) {
    lab[v] = v;
    atomicOr(bits + (v >> 6), 1ull << (v & 63));
  });
  __syncthreads();
  if (threadIdx.x < RROWS) fgbits[(size_t)tile * RROWS + threadIdx.x] = bits[threadIdx.x];
  if (n == 0) return;  // uniform over the work-group
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  if (links)
    R.each_rolled(rec, n, [&](uint32_t, uint32_t pv) __attribute__((always_inline)) {
      const int v = (int)pv;
      const int vx = v & (RTX - 1), vy = (v / RTX) & (RTY - 1), vz = v / (RTX * RTY);
      uint32_t fg = 0u;  // bit q: forward neighbour q is a candidate inside the clipped tile
#pragma unroll
      for (int q = 0; q < 13; ++q) {
        const int nx = vx + fdx(q), ny = vy + fdy(q), nz = vz + fdz(q);
        if (nx < 0 || nx >= T.ex || ny < 0 || ny >= T.ey || nz >= T.ez) continue;
        if ((bits[ny + RTY * nz] >> nx) & 1ull) fg |= 1u << q;
      }
      uint32_t root = (uint32_t)v;  // a representative of v's set, refreshed by each link
      uint32_t known = 0u;          // neighbours already in v's set
#pragma unroll
      for (int j = 0; j < 13; ++j) {
        const int q = fwd_order(j);
        if (!((fg >> q) & 1u)) continue;
        const bool skip = (fwd_adj(q) & known) != 0u;
        known |= 1u << q;
        if (skip) continue;
        const int u = (vx + fdx(q)) + RTX * ((vy + fdy(q)) + RTY * (vz + fdz(q)));
        if (lab[u] != root) root = lds16_unite(lab, root, (uint32_t)u);
      }
    });
  __syncthreads();
  auto global = [&](uint32_t v) {  // linear index of tile position v
    return (uint32_t)(T.z0 + (int)(v / (RTX * RTY))) * WH +
           (uint32_t)(T.y0 + (int)((v / RTX) & (RTY - 1))) * Wu + (uint32_t)T.x0 +
           (v & (RTX - 1));
  };
  // P is read at tile-local roots (the find chains) and at candidates another tile's
  // forward neighbours reach (the bottom face and the x / y faces of the clipped tile);
  // interior entries stay unwritten (kernel C writes every entry when `dense`)
  R.each(rec, n, [&](uint32_t i, uint32_t v) {
    const uint32_t r = lds16_find(lab, v);
    const uint32_t label = global(r);
    const int lx = (int)v & (RTX - 1), ly = (int)(v / RTX) & (RTY - 1), lz = (int)(v / (RTX * RTY));
    if (r == v || lz == 0 || lx == 0 || lx == T.ex - 1 || ly == 0 || ly == T.ey - 1)
      P[global(v)] = label;
    rlab[(size_t)tile * kCandTile + i] = label;
  });
}

__global__ __launch_bounds__(NTR) void k_cand_merge(const float* __restrict__ cand,
                                                   const uint32_t* __restrict__ cand_cnt,
                                                   const unsigned long long* __restrict__ fgbits,
                                                   uint32_t* P, const uint32_t* __restrict__ rlab,
                                                   int W, int H, int D, int nbx, int nby,
                                                   int generic, Offsets offs) {
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t n = cand_cnt[tile];
  if (n == 0) return;
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  const float* rec = cand + (size_t)tile * kCandTile * 7;
  if (generic) {  // width or height 1: every candidate, the distinct positive offsets
    const i64 N = (i64)WH * D;
#pragma unroll 1
    for (uint32_t i = threadIdx.x; i < n; i += NTR) {
      const uint32_t v = __float_as_uint(rec[i]);
      const uint32_t idx = (uint32_t)(T.z0 + (int)(v / (RTX * RTY))) * WH +
                           (uint32_t)(T.y0 + (int)((v / RTX) & (RTY - 1))) * Wu +
                           (uint32_t)T.x0 + (v & (RTX - 1));
      for (int k = 0; k < offs.n; ++k) {
        const i64 u = (i64)idx + offs.o[k];
        if (u >= N) break;  // ascending offsets
        const uint32_t uu = (uint32_t)u, uz = uu / WH, rem = uu - uz * WH, uy = rem / Wu,
                       ux = rem - uy * Wu;
        if ((fgbits[mask_word(ux, uy, uz, nbx, nby)] >> (ux & 63)) & 1ull) unite(P, idx, uu);
      }
    }
    return;
  }
  // rim candidates (those with a forward neighbour outside the clipped tile: the top
  // face and the ring of every lower plane) are compacted into an LDS list first, so the
  // neighbour work below runs on full waves
  constexpr int RIM_MAX = RTX * RTY + (RTZ - 1) * (2 * RTX + 2 * (RTY - 2));
  __shared__ uint16_t rim_v[RIM_MAX];
  __shared__ uint32_t rim_me[RIM_MAX];
  __shared__ uint32_t n_rim;
  __shared__ unsigned long long tb[RROWS];  // the tile's own mask rows
  if (threadIdx.x == 0) n_rim = 0;
  if (threadIdx.x < RROWS) tb[threadIdx.x] = fgbits[(size_t)tile * RROWS + threadIdx.x];
  TileRecs R;
  R.load(rec, n);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  auto push = [&](bool on, uint32_t v, uint32_t me) {  // wave-uniform call
    const unsigned long long bal = __ballot(on);
    if (!bal) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&n_rim, (uint32_t)__popcll(bal));
    base = __shfl(base, 0, 64);
    if (on) {
      const uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
      rim_v[slot] = (uint16_t)v;
      rim_me[slot] = me;
    }
  };
  auto is_rim = [&](uint32_t v) {
    const int lx = (int)v & (RTX - 1), ly = (int)(v / RTX) & (RTY - 1), lz = (int)(v / (RTX * RTY));
    return lz == T.ez - 1 || lx == 0 || lx == T.ex - 1 || ly == 0 || ly == T.ey - 1;
  };
#pragma unroll
  for (int k = 0; k < KREC; ++k) {
    const uint32_t i = threadIdx.x + k * NTR;
    const bool on = i < n && is_rim(R.v[k]);
    push(on, R.v[k], on ? rlab[(size_t)tile * kCandTile + i] : 0u);
  }
  for (uint32_t i0 = KREC * NTR; i0 < n; i0 += NTR) {  // uniform trip count
    const uint32_t i = i0 + threadIdx.x;
    const uint32_t v = i < n ? __float_as_uint(rec[i]) : 0u;
    const bool on = i < n && is_rim(v);
    push(on, v, on ? rlab[(size_t)tile * kCandTile + i] : 0u);
  }
  __syncthreads();
  const uint32_t nr = n_rim;
#pragma unroll 1
  for (uint32_t j = threadIdx.x; j < nr; j += NTR) {
    const int v = rim_v[j];
    const int lx = v & (RTX - 1), ly = (v / RTX) & (RTY - 1), lz = v / (RTX * RTY);
    const int x = T.x0 + lx, y = T.y0 + ly, z = T.z0 + lz;
    const uint32_t me = rim_me[j];
    // candidate neighbours inside the clipped tile (linked to v by kernel A) and outside
    uint32_t in_fg = 0u, out_fg = 0u, uidx[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      const int dx = fdx(q), dy = fdy(q), dz = fdz(q);
      uidx[q] = 0u;
      const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
      if (nx >= 0 && nx < T.ex && ny >= 0 && ny < T.ey && nz < T.ez) {
        if ((tb[ny + RTY * nz] >> nx) & 1ull) in_fg |= 1u << q;
        continue;
      }
      // the neighbour at linear offset dz WH + dy W + dx (wrapping across x and y like the
      // reference; one carry per axis for width, height >= 2); it exists below N
      int ux = x + dx, uy = y + dy, uz = z + dz;
      if (ux < 0) { ux += W; --uy; } else if (ux >= W) { ux -= W; ++uy; }
      if (uy < 0) { uy += H; --uz; } else if (uy >= H) { uy -= H; ++uz; }
      if (uz >= D) continue;
      if ((fgbits[mask_word((uint32_t)ux, (uint32_t)uy, (uint32_t)uz, nbx, nby)] >> (ux & 63)) &
          1ull) {
        out_fg |= 1u << q;
        uidx[q] = (uint32_t)uz * WH + (uint32_t)uy * Wu + (uint32_t)ux;
      }
    }
    // outside neighbours not adjacent to one already in v's set (those inside included)
    uint32_t known = in_fg, link = 0u;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      const int q = fwd_order(k);
      if (!((out_fg >> q) & 1u)) continue;
      if (!(fwd_adj(q) & known)) link |= 1u << q;
      known |= 1u << q;
    }
    uint32_t lj[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) lj[q] = ((link >> q) & 1u) ? P[uidx[q]] : 0u;
    uint32_t seen1 = 0u, seen2 = 0u;
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      const uint32_t l = lj[q];
      if (!l || l == me || l == seen1 || l == seen2) continue;
      unite(P, me, l);
      seen2 = seen1;
      seen1 = l;
    }
  }
}

__global__ __launch_bounds__(NTR) void k_cand_flatten(const float* __restrict__ cand,
                                                     const uint32_t* __restrict__ cand_cnt,
                                                     uint32_t* P, uint32_t* __restrict__ rlab,
                                                     unsigned long long* __restrict__ roots,
                                                     unsigned long long* __restrict__ multi,
                                                     int dense, int W, int H, int D, int nbx,
                                                     int nby) {
  const unsigned tile = blockIdx.x;
  const uint32_t n = cand_cnt[tile];
  if (n == 0) return;
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  const float* rec = cand + (size_t)tile * kCandTile * 7;
  uint32_t* rl = rlab + (size_t)tile * kCandTile;
  TileRecs R;
  R.load(rec, n);
  uint32_t l0[KREC];  // the records' tile-local root labels, loaded together
#pragma unroll
  for (int k = 0; k < KREC; ++k) {
    const uint32_t i = threadIdx.x + k * NTR;
    l0[k] = i < n ? rl[i] : 0u;
  }
  auto one = [&](uint32_t i, uint32_t v, uint32_t l) {
    const uint32_t idx = (uint32_t)(T.z0 + (int)(v / (RTX * RTY))) * WH +
                         (uint32_t)(T.y0 + (int)((v / RTX) & (RTY - 1))) * Wu + (uint32_t)T.x0 +
                         (v & (RTX - 1));
    // any value of P met on the way is an ancestor, also where another thread has already
    // written a root
    const uint32_t r = find_root_readonly(P, l);
    rl[i] = r;
    if (dense && P[idx] != r) P[idx] = r;
    if (r == idx) {
      atomicOr(roots + (r >> 6), 1ull << (r & 63));
    } else if (l == idx) {
      // the root of a tile-local component other than the region's first: the region
      // reaches beyond the region pass's work-group of r when this component lies outside it
      const uint32_t rz = r / WH, rr = r - rz * WH, ry = rr / Wu, rx = rr - ry * Wu;
      const int ty = T.y0 / RTY, tz = T.z0 / RTZ;
      if (rx / RTX != (uint32_t)T.x0 / RTX || (ry / RTY) / kAccGY != (uint32_t)(ty / kAccGY) ||
          (rz / RTZ) / kAccGZ != (uint32_t)(tz / kAccGZ))
        atomicOr(multi + (r >> 6), 1ull << (r & 63));
    }
  };
#pragma unroll
  for (int k = 0; k < KREC; ++k) {
    const uint32_t i = threadIdx.x + k * NTR;
    if (i < n) one(i, R.v[k], l0[k]);
  }
#pragma unroll 1
  for (uint32_t i = threadIdx.x + KREC * NTR; i < n; i += NTR)
    one(i, __float_as_uint(rec[i]), rl[i]);
}

// The distinct positive linear offsets dz W H + dy W + dx of the 27-neighbourhood
// (src/3D/CCL3D.cu:43-48 keeps every offset whose target lies in [0, N)).
Offsets positive_offsets(int W, int H) {
  const i64 WH = (i64)W * H;
  Offsets offs{};
  i64 v[27];
  int n = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const i64 o = dz * WH + dy * (i64)W + dx;
        if (o <= 0) continue;
        bool dup = false;
        for (int k = 0; k < n; ++k) dup |= (v[k] == o);
        if (!dup) v[n++] = o;
      }
  std::sort(v, v + n);
  for (int k = 0; k < n; ++k) offs.o[k] = v[k];
  offs.n = n;
  return offs;
}

}  // namespace

hipError_t launch_ccl(const uint8_t* pd, uint32_t* labels, int W, int H, int D, hipStream_t s,
                      bool flatten) {
  const i64 WH = (i64)W * H, N = WH * D;
  const Offsets offs = positive_offsets(W, H);
  const i64 nb = std::min<i64>((N + NT - 1) / NT, 1 << 20);
  hipError_t e;
  if (W >= 2 && H >= 2) {
    const int nbx = (W + CX - 1) / CX, nby = (H + CY - 1) / CY, nbz = (D + CZ - 1) / CZ;
    const i64 nt = (i64)nbx * nby * nbz;
    hipLaunchKernelGGL(k_ccl_local, dim3((unsigned)nt), dim3(NTC), 0, s, pd, labels, W, H, D, nbx,
                       nby);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("ccl/local", s);
    hipLaunchKernelGGL(k_ccl_merge, dim3((unsigned)nt), dim3(NTM), 0, s, labels, W, H, D, nbx, nby);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("ccl/merge", s);
  } else {
    hipLaunchKernelGGL(k_ccl_init, dim3((unsigned)nb), dim3(NT), 0, s, pd, labels, N);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("ccl/init", s);
    hipLaunchKernelGGL(k_ccl_union, dim3((unsigned)nb), dim3(NT), 0, s, pd, labels, N, offs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    mark("ccl/union", s);
  }
  if (!flatten) return hipSuccess;
  hipLaunchKernelGGL(k_ccl_flatten, dim3((unsigned)nb), dim3(NT), 0, s, labels, N);
  mark("ccl/flatten", s);
  return hipGetLastError();
}

hipError_t launch_ccl_candidates(const float* cand, const uint32_t* cand_cnt, i64 ntiles,
                                 unsigned long long* fgbits, uint32_t* P, uint32_t* rlab,
                                 unsigned long long* roots, unsigned long long* multi,
                                 i64 root_words, bool dense, int W, int H, int D,
                                 hipStream_t s) {
  const int nbx = (W + RTX - 1) / RTX, nby = (H + RTY - 1) / RTY;
  hipError_t e;
  if (dense) {  // background of the materialized label volume
    if ((e = launch_fill_u32(P, (i64)W * H * D, 0u, s)) != hipSuccess) return e;
  }
  // the 13 forward offsets with one carry per axis are the positive offsets only for
  // width, height >= 2; otherwise every link goes through the generic merge
  const int generic = (W < 2 || H < 2) ? 1 : 0;
  hipLaunchKernelGGL(k_cand_local, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt, fgbits,
                     P, rlab, W, H, D, nbx, nby, 1 - generic);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark("ccl/cand_local", s);
  hipLaunchKernelGGL(k_cand_merge, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt,
                     (const unsigned long long*)fgbits, P, rlab, W, H, D, nbx, nby, generic,
                     positive_offsets(W, H));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark("ccl/cand_merge", s);
  if ((e = launch_fill_u64((u64*)roots, root_words, 0ull, s)) != hipSuccess) return e;
  if ((e = launch_fill_u64((u64*)multi, root_words, 0ull, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_cand_flatten, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt, P,
                     rlab, roots, multi, dense ? 1 : 0, W, H, D, nbx, nby);
  mark("ccl/cand_flatten", s);
  return hipGetLastError();
}

}  // namespace hdog
