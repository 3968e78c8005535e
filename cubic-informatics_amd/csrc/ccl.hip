// ccl.hip -- 3-D connected-component labelling on gfx950.
//
// Reference semantics: src/3D/CCL3D.cu:17-117 iterates scan / pointer-jump passes,
// polling a device flag from the host, until every foreground voxel holds the minimum
// linear index of its component.  Adjacency is the set of linear offsets
// dz*W*H + dy*W + dx kept when the target is inside [0, N), so rows and slices wrap;
// voxel 0 is never foreground (label = index * mask).
//
// MI355X design: no host round trips, and only foreground (candidate) voxels are touched.
// The labeling reads the candidate positions per 64 x 8 x 16 tile that the Hessian pass
// records (or that launch_mask_records extracts from a byte mask) and runs a union-find in
// three kernels: per tile in LDS (links inside the tile), per tile over its rim (links
// that leave the tile or wrap, in a volume-sized parent array touched only at candidates),
// then every candidate's root.  Roots are linked larger-under-smaller, so the root of every
// tree is the component minimum; each adjacent pair is linked once, by its lower voxel.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

struct Offsets {
  i64 o[13];
  int n;
};

// Every access to the parent array P while other work-groups update it (the merge and the
// flatten) is an atomic at agent scope: a relaxed load or store here, the minimum in
// unite.  Plain accesses beside the atomics would be a data race, and on gfx950 they are
// issued without the device-scope coherence bit, so they may be served by, or left in, a
// cache that is not kept coherent with the other XCDs' for the rest of the kernel; the
// relaxed atomic loads and stores carry that bit (sc1) and cost one coherent access each.
__device__ __forceinline__ uint32_t ld_parent(const uint32_t* L, uint32_t x) {
  return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_parent(uint32_t* L, uint32_t x, uint32_t v) {
  __hip_atomic_store(L + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Parents only ever decrease, so every chain ends at a root.  A parent read while others
// link is at worst out of date, and then still an ancestor: two voxels that ever shared an
// ancestor are in the same set, and every link goes through an atomicMin whose returned
// value is current, so each retry of unite strictly lowers `a` and the loop ends.  Path
// halving (x's parent replaced by its grandparent) keeps the trees shallow; the store only
// ever replaces a parent by one of its ancestors, so no union is lost (a store that lands
// after another thread's minimum on the same non-root entry is repaired by that thread's
// retry, which unites the old parent with its new one).
// Only the merge may halve paths: in the flatten kernel a halving store could overwrite
// another voxel's final root with an intermediate ancestor.
__device__ __forceinline__ uint32_t find_root(uint32_t* L, uint32_t x) {
  uint32_t p = ld_parent(L, x);
  while (p != x) {
    const uint32_t gp = ld_parent(L, p);
    if (gp != p) st_parent(L, x, gp);
    x = p;
    p = gp;
  }
  return x;
}
// the root of x while other threads only replace parents by roots (the flatten)
__device__ __forceinline__ uint32_t find_root_flat(const uint32_t* L, uint32_t x) {
  uint32_t p = ld_parent(L, x);
  while (p != x) {
    x = p;
    p = ld_parent(L, x);
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

// The tile-local union-find, in LDS on 16-bit parents (tile positions below 2^16): half
// the LDS of 32-bit parents, so twice the work-groups per CU.  LDS has no 16-bit atomic
// minimum; the link is a 32-bit compare-and-swap on the word holding the entry, which
// fails and retries when either half changed.  Path-halving stores are 16-bit stores (a
// parent is only ever replaced by one of its ancestors, as above).  While the unions run,
// every access to the array is a relaxed work-group-scope atomic (one ds_read_u16 /
// ds_write_b16 each), so the compiler keeps each load and store where it is written.
__device__ __forceinline__ uint32_t lab_ld(const uint16_t* lab, uint32_t x) {
  return __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds16_find(const uint16_t* lab, uint32_t x) {
  uint32_t p = lab[x];  // after the unions (read-only)
  while (p != x) {
    x = p;
    p = lab[x];
  }
  return x;
}
__device__ __forceinline__ uint32_t lds16_find_halve(uint16_t* lab, uint32_t x) {
  uint32_t p = lab_ld(lab, x);
  while (p != x) {
    const uint32_t gp = lab_ld(lab, p);
    if (gp != p)
      __hip_atomic_store(lab + x, (uint16_t)gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = p;
    p = gp;
  }
  return x;
}
// lab[a] = min(lab[a], b); returns the old lab[a]
__device__ __forceinline__ uint32_t lds16_min(uint16_t* lab, uint32_t a, uint32_t b) {
  uint32_t* w = reinterpret_cast<uint32_t*>(lab + (a & ~1u));
  const int sh = (int)(a & 1u) * 16;
  uint32_t cur = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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

// ---- labeling of candidate records ---------------------------------------------------
// The Hessian pass (or launch_mask_records, for a given mask) leaves, per 64 x 8 x 16
// tile, the positions of its candidate voxels (the foreground of the CCL) in the tile;
// every index is 32-bit (N < 2^32).  Labeling then touches only candidates: P (a
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
//             numbering, the roots of regions that reach beyond one region-pass
//             work-group, and with `dense` the final label written back into P.
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
                                                   int D, int nbx, int nby, int links,
                                                   int rec_words) {
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t n = cand_cnt[tile];
  __shared__ unsigned long long bits[RROWS];  // the tile's mask rows (word = ly + RTY lz)
  // parents, valid at candidate positions; linked by 32-bit compare-and-swap on the word
  // holding two of them, so the array starts on a 4-byte boundary
  __shared__ alignas(4) uint16_t lab[RTV];
  const float* rec = cand + (size_t)tile * kCandTile * rec_words;
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
      // bit q: forward neighbour q is a candidate inside the clipped tile (branch-free).
      // The 13 neighbours lie in five mask rows: x-1, x, x+1 of rows (y+1, z), (y-1, z+1),
      // (y, z+1), (y+1, z+1) in three bits each, and x+1 of (y, z).  Mask bits outside the
      // clipped tile are 0 (no candidates there), so only rows outside it need a test; a
      // column outside the 64-bit row shifts in as 0.
      const bool yl = vy > 0, yr = vy + 1 < T.ey, zr = vz + 1 < T.ez;
      auto tri = [&](int row, bool ok) -> uint32_t {  // bits of x-1, x, x+1 (0 outside)
        const unsigned long long w = bits[ok ? row : 0];
        const uint32_t t = vx ? (uint32_t)(w >> (vx - 1)) & 7u : (uint32_t)(w << 1) & 7u;
        return ok ? t : 0u;
      };
      const int r0 = vy + RTY * vz;
      const uint32_t fg = tri(r0 - 1 + RTY, yl && zr) | tri(r0 + RTY, zr) << 3 |
                          tri(r0 + 1 + RTY, yr && zr) << 6 | tri(r0 + 1, yr) << 9 |
                          ((uint32_t)(bits[r0] >> vx >> 1) & 1u) << 12;
      static_assert(fdx(0) == -1 && fdy(0) == -1 && fdz(0) == 1 && fdx(8) == 1 &&
                        fdy(8) == 1 && fdx(9) == -1 && fdy(9) == 1 && fdz(9) == 0 &&
                        fdx(12) == 1 && fdy(12) == 0 && fdz(12) == 0,
                    "bit q of the five rows is forward neighbour q");
      // the links: neighbours not adjacent to one already known to be in v's set
      uint32_t known = 0u, link = 0u;
#pragma unroll
      for (int j = 0; j < 13; ++j) {
        const int q = fwd_order(j);
        const uint32_t b = (fg >> q) & 1u;
        link |= ((fwd_adj(q) & known) == 0u ? b : 0u) << q;
        known |= b << q;
      }
      uint32_t root = (uint32_t)v;  // a representative of v's set, refreshed by each link
      while (link) {  // one copy of the union code for every link
        const int q = __builtin_ctz(link);
        link &= link - 1u;
        // tile-local offset of forward neighbour q: dz 512 + dy 64 + dx
        const int off = q < 9 ? RTX * RTY + RTX * (q / 3 - 1) + (q % 3 - 1)
                              : (q < 12 ? RTX + (q - 10) : 1);
        const uint32_t u = (uint32_t)(v + off);
        if (lab_ld(lab, u) != root) root = lds16_unite(lab, root, u);
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
                                                   int generic, Offsets offs, int rec_words) {
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t n = cand_cnt[tile];
  if (n == 0) return;
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  const float* rec = cand + (size_t)tile * kCandTile * rec_words;
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
  // the mask words of the whole tiling (gridDim.x tiles)
  const __amdgpu_buffer_rsrc_t fg_rsrc =
      buf_rsrc(fgbits, (uint32_t)gridDim.x * (uint32_t)(RROWS * sizeof(unsigned long long)));
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
    base = __builtin_amdgcn_readfirstlane(base);  // lane 0 (the whole wave is here)
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
    // candidate neighbours inside the clipped tile (linked to v by kernel A) and outside;
    // no branches: outside words come through a buffer descriptor (an offset past its end,
    // for a neighbour that does not exist or lies inside, reads 0), so all are in flight
    uint32_t in_fg = 0u, uidx[13];
    unsigned long long ow[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      const int dx = fdx(q), dy = fdy(q), dz = fdz(q);
      const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
      const bool inside = nx >= 0 && nx < T.ex && ny >= 0 && ny < T.ey && nz < T.ez;
      const int row = inside ? ny + RTY * nz : 0;
      in_fg |= (inside ? (uint32_t)(tb[row] >> (nx & (RTX - 1))) & 1u : 0u) << q;
      // the neighbour at linear offset dz WH + dy W + dx (wrapping across x and y like the
      // reference; one carry per axis for width, height >= 2); it exists below N
      int ux = x + dx, uy = y + dy, uz = z + dz;
      if (dx < 0) { const bool c = ux < 0; ux = c ? ux + W : ux; uy = c ? uy - 1 : uy; }
      if (dx > 0) { const bool c = ux >= W; ux = c ? ux - W : ux; uy = c ? uy + 1 : uy; }
      if (dy < 0 || dx < 0) { const bool c = uy < 0; uy = c ? uy + H : uy; uz = c ? uz - 1 : uz; }
      if (dy > 0 || dx > 0) { const bool c = uy >= H; uy = c ? uy - H : uy; uz = c ? uz + 1 : uz; }
      const bool exists = !inside && uz < D;
      const uint32_t word =
          (uint32_t)mask_word((uint32_t)ux, (uint32_t)uy, (uint32_t)uz, nbx, nby);
      ow[q] = __builtin_bit_cast(unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(
                                     fg_rsrc, exists ? (int)(word * 8u) : -16, 0, 0)) >>
              (ux & 63);
      uidx[q] = exists ? (uint32_t)uz * WH + (uint32_t)uy * Wu + (uint32_t)ux : 0u;
    }
    uint32_t out_fg = 0u;
#pragma unroll
    for (int q = 0; q < 13; ++q) out_fg |= ((uint32_t)ow[q] & 1u) << q;
    // outside neighbours not adjacent to one already in v's set (those inside included)
    uint32_t known = in_fg, link = 0u;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      const int q = fwd_order(k);
      const uint32_t b = (out_fg >> q) & 1u;
      link |= ((fwd_adj(q) & known) == 0u ? b : 0u) << q;
      known |= b << q;
    }
    uint32_t lj[13];  // the linked neighbours' labels, loaded together (P[0] for the rest)
#pragma unroll
    for (int q = 0; q < 13; ++q) {
      const bool on = (link >> q) & 1u;
      const uint32_t l = ld_parent(P, on ? uidx[q] : 0u);
      lj[q] = on ? l : 0u;
    }
    uint32_t seen1 = 0u, seen2 = 0u;
    while (link) {  // one copy of the union code for every link
      const int q = __builtin_ctz(link);
      link &= link - 1u;
      uint32_t l = 0u;
#pragma unroll
      for (int k = 0; k < 13; ++k) l = k == q ? lj[k] : l;
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
                                                     int nby, int rec_words) {
  const unsigned tile = blockIdx.x;
  const uint32_t n = cand_cnt[tile];
  if (n == 0) return;
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  const float* rec = cand + (size_t)tile * kCandTile * rec_words;
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
    const uint32_t r = find_root_flat(P, l);
    rl[i] = r;
    if (dense && ld_parent(P, idx) != r) st_parent(P, idx, r);
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

// Candidate positions of a byte mask, in the record layout of the Hessian pass (per
// Hessian tile a count and the tile positions, one reservation per wave and tile row, in
// lane order): the labeling of a given mask then runs the same kernels as the whole path.
// Voxel 0 is background whatever its mask value (its label is 0).
__global__ __launch_bounds__(256) void k_mask_records(const uint8_t* __restrict__ pd,
                                                      float* __restrict__ pos,
                                                      uint32_t* __restrict__ cand_cnt, int W,
                                                      int H, int D, int nbx, int nby) {
  const unsigned tile = blockIdx.x;
  const CandTile T = cand_tile(tile, W, H, D, nbx, nby);
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t WH = (uint32_t)W * (uint32_t)H;
  float* out = pos + (size_t)tile * kCandTile;
  for (int row = wave; row < RROWS; row += 4) {  // uniform over the wave
    const int ly = row % RTY, lz = row / RTY;
    bool on = false;
    if (lane < T.ex && ly < T.ey && lz < T.ez) {
      const uint32_t idx = (uint32_t)(T.z0 + lz) * WH + (uint32_t)(T.y0 + ly) * (uint32_t)W +
                           (uint32_t)(T.x0 + lane);
      on = idx != 0 && pd[idx] != 0;
    }
    const unsigned long long bal = __ballot(on);
    if (!bal) continue;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&s_cnt, (uint32_t)__popcll(bal));
    base = __builtin_amdgcn_readfirstlane(base);  // lane 0 (the whole wave is here)
    if (on)
      out[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] =
          __uint_as_float((uint32_t)(lane + RTX * (ly + RTY * lz)));
  }
  __syncthreads();
  if (threadIdx.x == 0) cand_cnt[tile] = s_cnt;
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

hipError_t launch_mask_records(const uint8_t* pd, float* pos, uint32_t* cand_cnt, int W, int H,
                               int D, hipStream_t s) {
  const int nbx = (W + RTX - 1) / RTX, nby = (H + RTY - 1) / RTY;
  hipLaunchKernelGGL(k_mask_records, dim3((unsigned)hessian_tiles(W, H, D)), dim3(256), 0, s, pd,
                     pos, cand_cnt, W, H, D, nbx, nby);
  mark("ccl/mask_records", s);
  return hipGetLastError();
}

hipError_t launch_ccl_candidates(const float* cand, const uint32_t* cand_cnt, i64 ntiles,
                                 unsigned long long* fgbits, uint32_t* P, uint32_t* rlab,
                                 unsigned long long* roots, unsigned long long* multi,
                                 i64 root_words, bool dense, int W, int H, int D, hipStream_t s,
                                 int rec_words) {
  const int nbx = (W + RTX - 1) / RTX, nby = (H + RTY - 1) / RTY;
  hipError_t e;
  if (dense) {  // background of the materialized label volume
    if ((e = launch_fill_u32(P, (i64)W * H * D, 0u, s)) != hipSuccess) return e;
  }
  // the 13 forward offsets with one carry per axis are the positive offsets only for
  // width, height >= 2; otherwise every link goes through the generic merge
  const int generic = (W < 2 || H < 2) ? 1 : 0;
  hipLaunchKernelGGL(k_cand_local, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt, fgbits,
                     P, rlab, W, H, D, nbx, nby, 1 - generic, rec_words);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark("ccl/cand_local", s);
  hipLaunchKernelGGL(k_cand_merge, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt,
                     (const unsigned long long*)fgbits, P, rlab, W, H, D, nbx, nby, generic,
                     positive_offsets(W, H), rec_words);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark("ccl/cand_merge", s);
  if ((e = launch_fill_u64((u64*)roots, root_words, 0ull, s)) != hipSuccess) return e;
  if ((e = launch_fill_u64((u64*)multi, root_words, 0ull, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_cand_flatten, dim3((unsigned)ntiles), dim3(NTR), 0, s, cand, cand_cnt, P,
                     rlab, roots, multi, dense ? 1 : 0, W, H, D, nbx, nby, rec_words);
  mark("ccl/cand_flatten", s);
  return hipGetLastError();
}

}  // namespace hdog
