// hdog_internal.hpp -- shared declarations of the HIP kernels' launchers and the
// context object behind include/hdog.h.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <string>

typedef int64_t i64;
typedef uint64_t u64;

namespace hdog {

constexpr int kMaxGaussRadius = 40;            // 2r+1 <= 81 taps per axis
constexpr int kMaxTaps = 2 * kMaxGaussRadius + 1;
constexpr int kMaxNormRadius = 200;
constexpr int kRootBlock = 4096;  // voxels per root-numbering block (64 mask words)
// per-block root counts of a volume of N voxels, followed (at a 64-word boundary) by the
// scan's chunk sums, one per 4096 blocks
inline i64 scan_chunk_offset(i64 nb) { return (nb + 63) / 64 * 64; }
inline i64 block_count_words(i64 N) {
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  return scan_chunk_offset(nb) + (nb + 4095) / 4096;
}
// word 3 of a raw buffer descriptor on gfx9-family GPUs (32-bit data format; no swizzle)
constexpr int kBufferWord3 = 0x00020000;
// cache-policy operand of a buffer store: nontemporal (streaming), as a nontemporal global
// store is compiled on gfx950
constexpr int kCachePolicyNT = 2;
// Hessian tiles: 64 x 8 outputs per work-group over 16 slices.  Candidate records are kept
// per tile in a block of kCandTile slots x 7 words (voxel index, then the six elements).
// The tile height and the region pass's tile group are build-time choices (-DHDOG_HESS_TY,
// -DHDOG_ACC_GY / GZ; static_asserts guard the limits); 8 and 2 x 2 measured fastest.
#ifndef HDOG_HESS_TY
#define HDOG_HESS_TY 8
#endif
constexpr int kHessTX = 64, kHessTY = HDOG_HESS_TY, kHessTZ = 16;
constexpr int kCandTile = kHessTX * kHessTY * kHessTZ;
// reduced-precision forms: default FP64 re-evaluation band of the leading minors (relative
// to the sum of the magnitudes of their terms)
constexpr float kDefaultMinorBand = 1e-5f;
inline i64 hessian_tiles(int W, int H, int D) {
  return (i64)((W + kHessTX - 1) / kHessTX) * ((H + kHessTY - 1) / kHessTY) *
         ((D + kHessTZ - 1) / kHessTZ);
}
// words of the candidate bit mask: one per (z, y) row of each Hessian tile, tile-major
inline i64 fg_words(int W, int H, int D) { return hessian_tiles(W, H, D) * kHessTY * kHessTZ; }
// Region sums from candidate records: one work-group folds kAccGY x kAccGZ Hessian tiles
// (y, z); a region inside one such group is summed by that work-group alone.
#ifndef HDOG_ACC_GY
#define HDOG_ACC_GY 2
#endif
#ifndef HDOG_ACC_GZ
#define HDOG_ACC_GZ 2
#endif
constexpr int kAccGY = HDOG_ACC_GY, kAccGZ = HDOG_ACC_GZ;

// ---- storage types of float intermediates --------------------------------------
template <typename T> __device__ __forceinline__ float ld(const T* p, i64 i);
template <> __device__ __forceinline__ float ld<float>(const float* p, i64 i) { return p[i]; }
template <> __device__ __forceinline__ float ld<__half>(const __half* p, i64 i) {
  return __half2float(p[i]);
}
template <> __device__ __forceinline__ float ld<__hip_bfloat16>(const __hip_bfloat16* p, i64 i) {
  return __bfloat162float(p[i]);
}
template <typename T> __device__ __forceinline__ T cvt(float v);
template <> __device__ __forceinline__ float cvt<float>(float v) { return v; }
template <> __device__ __forceinline__ __half cvt<__half>(float v) { return __float2half_rn(v); }
template <> __device__ __forceinline__ __hip_bfloat16 cvt<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}

// LDS stores before, LDS loads after, in a work-group of one wave: the LDS instructions of
// all waves execute in one total order, the order in which each wave issues them, so a
// load issued after a store sees it without a wait or a barrier instruction; the fence
// only keeps the compiler from moving the accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Raw buffer loads of one stored element (the bits, converted by from_bits): a load whose
// byte offset lies at or beyond the descriptor's record count returns zero.
template <typename T> struct StoredBits { typedef uint16_t type; };
template <> struct StoredBits<float> { typedef uint32_t type; };
template <typename T>
__device__ __forceinline__ typename StoredBits<T>::type buf_load(__amdgpu_buffer_rsrc_t r,
                                                                 uint32_t off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
  else
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
}
template <typename T> __device__ __forceinline__ float from_bits(typename StoredBits<T>::type b);
template <> __device__ __forceinline__ float from_bits<float>(uint32_t b) {
  return __uint_as_float(b);
}
template <> __device__ __forceinline__ float from_bits<__half>(uint16_t b) {
  return __half2float(__builtin_bit_cast(__half, b));
}
template <> __device__ __forceinline__ float from_bits<__hip_bfloat16>(uint16_t b) {
  return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, b));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufferWord3);
}

__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// ---- streaming (nontemporal) stores of volumes written once, read by a later kernel ----
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(uint32_t* p, uint32_t v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ void st_nt(float4* p, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<f4*>(p));
}
template <typename T> __device__ __forceinline__ void st_nt(T* p, float v) { *p = cvt<T>(v); }

// ---- tile staging -----------------------------------------------------------------
// Copies n elements global -> LDS with B independent loads in flight per lane before the
// first LDS store (a plain strided loop waits for each load in turn).
template <int B, int NTH, typename T, typename LD, typename ST>
__device__ __forceinline__ void stage_tile(int n, LD&& ldf, ST&& stf) {
  for (int base = (int)threadIdx.x; base < n; base += B * NTH) {
    T v[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int e = base + k * NTH;
      if (e < n) v[k] = ldf(e);
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int e = base + k * NTH;
      if (e < n) stf(e, v[k]);
    }
  }
}

// ---- per-region accumulator (one 128-byte line per region) -----------------------
struct alignas(128) RegionAcc {
  unsigned long long sx, sy, sz, isum;
  double h[6];
  double nsum;
  unsigned int cnt, nmax_key, imax, root;
  unsigned int pad[2];
};

// ---- optional per-kernel event marks (profiling mode of the C-ABI) ----------------
// A launcher calls mark(name, s) after each kernel it enqueues; with profiling on the
// context points tl_prof at its event pool for the duration of the call.  A mark with
// name nullptr only starts a new interval (after a host synchronization).
struct Prof {
  static constexpr int kMax = 64;
  hipEvent_t ev[kMax];
  const char* name[kMax];
  int n = 0;
};
extern thread_local Prof* tl_prof;
inline void mark(const char* name, hipStream_t s) {
  Prof* p = tl_prof;
  if (p && p->n < Prof::kMax) {
    (void)hipEventRecord(p->ev[p->n], s);
    p->name[p->n++] = name;
  }
}

// Work-group b of n runs on XCD b % 8 (the dispatcher deals work-groups to the 8 XCDs in
// turn, each XCD with its own L2).  Returns the logical tile for b such that every XCD
// walks one contiguous range of tiles in order: tiles that share cache lines (halo rows,
// 128-byte lines split between narrow tiles) then meet in the same L2.  A bijection of
// [0, n) for any n.
// Row walkers (one wave per plane strip) split tall planes into row segments, each
// starting with the 2R rows of context its window needs: a C1 plane strip as one wave left
// the last round of waves part-empty (2.5 and 1.4 rounds of resident waves for the two
// walkers).  `want` segments (a tuned launch shape), or by default two for planes of at
// least 16 ring periods; segment length: a multiple of the ring period wb.
inline void walk_segments(int H, int wb, int* nys, int* rows, int want = 0) {
  const int n = want > 0 ? want : (H >= 16 * wb ? 2 : 1);
  const int per = (H + n - 1) / n;
  *rows = (per + wb - 1) / wb * wb;
  *nys = (H + *rows - 1) / *rows;
}

__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned n) {
  constexpr unsigned X = 8;
  const unsigned q = n / X, r = n % X, j = b % X;
  return j * q + min(j, r) + b / X;
}

// One zero-padded central difference ((0 + hi) - lo) * 0.5 in the reference's order; the
// Hessian kernel and the region pass that recomputes the Hessian share it.
__device__ __forceinline__ float cdiff_half(float hi, float lo) {
  float s = 0.0f;
  s += hi;
  s -= lo;
  s *= 0.5f;
  return s;
}

// v rounded to the storage type T and widened back
template <typename T> __device__ __forceinline__ float round_to(float v) {
  const T t = cvt<T>(v);
  return ld(&t, 0);
}

// ---- launchers (return hipSuccess or the launch error) ----------------------------
// normalize stage; tmp1/tmp2 each hold >= 8*N bytes.  ero / dil (optional): the box means
// of the min and max filters, E and D, as separate float volumes.
// segments: row segments of the row walker (0: walk_segments' default).
template <typename TS>
hipError_t launch_normalize(const uint16_t* img, TS* norm, void* tmp1, void* tmp2,
                            float min_intensity, int W, int H, int D, int rxy, int rz,
                            hipStream_t s, float* ero = nullptr, float* dil = nullptr,
                            int segments = 0);
// Clamped box mean of float2 pairs along y (sa = W, so = W*H, n_other = D) or z (sa = W*H,
// so = W, n_other = H): the sequential sums c-r .. c+r times 1/(2r+1) of the normalization's
// y and z passes.
hipError_t launch_box_pair(const float2* src, float2* dst, int W, int L, i64 sa, i64 so,
                           i64 n_other, int r, hipStream_t s);
// DoG stage; weights: device array [4][kMaxTaps] (xy1, z1, xy2, z2); tmp1/tmp2 >= 8*N bytes;
// segments as for launch_normalize
template <typename TN, typename TD>
hipError_t launch_dog(const TN* norm, TD* dog, void* tmp1, void* tmp2, const float* weights,
                      const int r[4], float normalizer, int W, int H, int D, hipStream_t s,
                      int segments = 0);
hipError_t launch_gauss_pair(const float* src, void* tmp1, void* tmp2, const float* const w[6],
                             const int r[6], float nz, float* dog, float2* pair_out, int W, int H,
                             int D, hipStream_t s);
template <typename TD, typename TH>
// hess and pd may be NULL (not written).  cand / cand_cnt (optional): per Hessian tile
// (kHessTX x kHessTY x kHessTZ, tiles x-fastest), the number of candidate voxels and their
// records: the voxel's position in the tile, lx + kHessTX (ly + kHessTY lz), then the six
// elements rounded to TH, as 7 SoA words of kCandTile slots; cand holds hessian_tiles() *
// kCandTile * 7 words.  Voxel 0 never has a record (its label is 0 whatever its mask).
// band: the FP64 re-evaluation band of the leading minors when TD or TH is a 2-byte type.
hipError_t launch_hessian(const TD* dog, TH* hess, uint8_t* pd, int W, int H, int D,
                          hipStream_t s, float* cand = nullptr, uint32_t* cand_cnt = nullptr,
                          float band = kDefaultMinorBand);
// Exclusive scan of the nb per-block counts in place (block_cnt: block_count_words() words),
// *d_R = 1 + their total.
hipError_t launch_scan_counts(uint32_t* block_cnt, i64 nb, uint32_t* d_R, hipStream_t s);
// Region numbering of a label volume: root counts per block of 4096 voxels and their
// positions (masks: 64 words per block), then *d_R = 1 + number of roots.
hipError_t launch_region_ids(const uint32_t* labels, uint32_t* block_cnt,
                             unsigned long long* masks, uint32_t* d_R, i64 N, hipStream_t s,
                             i64* nblocks_out);
// Region ids from the root bits and the scanned block counts: rid at every root, and a
// cleared accumulator line (with the root) for every id below cap; ids >= cap are skipped
// (graph-captured batches, where the host never sees the count, rely on that).
hipError_t launch_region_assign(const unsigned long long* masks, const uint32_t* block_off,
                                uint32_t* rid, RegionAcc* acc, i64 cap, i64 N, hipStream_t s);
template <typename TH, typename TN>
hipError_t launch_region_accum(const uint32_t* labels, const uint32_t* rid, const TH* hess,
                               const TN* norm, const uint16_t* img, RegionAcc* acc, i64 cap,
                               int W, int H, int D, int is_ave, hipStream_t s);
// Region sums from the candidate records of launch_hessian (ntiles = hessian_tiles()).
// rlab and multi come from launch_ccl_candidates: each record's root, and the roots of
// regions reaching beyond one work-group's tiles (all others are summed by one work-group
// and stored without atomics).
template <typename TN>
hipError_t launch_region_accum_cand(const float* cand, const uint32_t* cand_cnt, i64 ntiles,
                                    const uint32_t* rlab, const unsigned long long* multi,
                                    const uint32_t* rid, const TN* norm, const uint16_t* img,
                                    RegionAcc* acc, i64 cap, int W, int H, int is_ave,
                                    hipStream_t s);
struct RegionOut {
  uint32_t* label; uint16_t* size; u64* voxels; float* cx; float* cy; float* cz;
  float* st; float* bl; float* nrm; float* hess; u64* isum; uint16_t* imax;
};
// Placement of one stack's regions in a batch's region arena (device memory): regions
// 1..count of the stack occupy arena slots offset .. offset + count - 1.
struct TileOut {
  u64 offset, count;
  uint32_t overflow, found;  // found: regions of the stack, also when they did not fit
};
// With t == nullptr: region r goes to slot r, r < R.  With t: regions 1..t->count go to
// slots t->offset + r - 1 (nothing when t->overflow); R then only sizes the grid.
hipError_t launch_region_finalize(const RegionAcc* acc, i64 R, int is_ave, RegionOut o,
                                  i64 pitch, hipStream_t s, const TileOut* t = nullptr);
// t->offset = *running, t->count = *d_R - 1, *running += t->count; overflow (count 0) when
// *d_R > acc_cap or the arena of arena_cap slots would overflow
hipError_t launch_tile_reserve(const uint32_t* d_R, i64 acc_cap, i64 arena_cap, u64* running,
                               TileOut* t, hipStream_t s);
hipError_t launch_fill_u64(u64* p, i64 n, u64 value, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, i64 n, uint32_t value, hipStream_t s);
// Labeling of the candidates the Hessian pass recorded (cand / cand_cnt, see
// launch_hessian): fgbits (fg_words() words) receives the candidate bit mask; P is a
// volume-sized parent array written and read at candidate voxels only; rlab (slot-aligned
// with the records) receives each record's label, the component's minimum linear index;
// roots (root_words 64-bit words, cleared here) receives the root bit of every component,
// in linear order, and multi (same size, cleared here) the root bit of every component
// with voxels outside the kAccGY x kAccGZ tile group of its root.  dense: P becomes the
// full label volume (0 for the background).
// rec_words: 32-bit words per record slot block (7: the Hessian pass's records; 1: the
// positions of launch_mask_records).
hipError_t launch_ccl_candidates(const float* cand, const uint32_t* cand_cnt, i64 ntiles,
                                 unsigned long long* fgbits, uint32_t* P, uint32_t* rlab,
                                 unsigned long long* roots, unsigned long long* multi,
                                 i64 root_words, bool dense, int W, int H, int D, hipStream_t s,
                                 int rec_words = 7);
// The candidate positions of a byte mask (pd != 0, voxel 0 excluded) in the Hessian pass's
// record layout with one word per slot (pos: hessian_tiles() * kCandTile words).
hipError_t launch_mask_records(const uint8_t* pd, float* pos, uint32_t* cand_cnt, int W, int H,
                               int D, hipStream_t s);
// Roots per numbering block from a root bit mask in linear order, then the exclusive scan
// of the counts; *d_R = 1 + number of roots.
hipError_t launch_region_count_masks(const unsigned long long* masks, uint32_t* block_cnt,
                                     uint32_t* d_R, i64 N, hipStream_t s, i64* nblocks_out);
// differentiable blob response (soft.hip)
hipError_t launch_soft_loss(const float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                            double inv_tau, double* sums, hipStream_t s);
hipError_t launch_soft_grad_h(float* hess, const uint8_t* mask, const uint8_t* target, i64 N,
                              double inv_tau, double inv_count, double* sums, hipStream_t s);
hipError_t launch_hessian_adjoint(const float* gh, float* P, float* out, int W, int H, int D,
                                  hipStream_t s);
// Differentiable form, local-contrast normalization of a float volume (clamped min / max
// and box of radius (rxy, rxy, rz)): norm, the (E, D) pairs and the window argmin / argmax
// (N each) for the backward pass; scratch 8 N words.
hipError_t launch_soft_normalize(const float* v, float* norm, float2* ed, uint32_t* arg,
                                 void* scratch, float min_int, int W, int H, int D, int rxy,
                                 int rz, hipStream_t s);
// Its adjoint: gv = dL/dv from gn = dL/dnorm (min / max by their subgradient); scratch 4 N
// words.
hipError_t launch_soft_normalize_adjoint(const float* v, const float2* ed, const uint32_t* arg,
                                         const float* gn, float* gv, void* scratch,
                                         float min_int, int W, int H, int D, int rxy, int rz,
                                         hipStream_t s);
hipError_t launch_dot_pair(const float* a, const float2* pair, i64 N, double* out, hipStream_t s);
hipError_t launch_dot(const float* a, const float* b, i64 N, double* out, hipStream_t s);
hipError_t launch_synth(uint16_t* img, float* scratch, int W, int H, int D, u64 seed,
                        double mean0, double mean1, double sd, const double* d_blobs,
                        i64 n_blobs, hipStream_t s);

}  // namespace hdog
