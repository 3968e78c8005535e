// hessian.hip -- Hessian elements and the negative-definite candidate mask, one kernel.
//
// Reference semantics: src/3D/Hessian3DFilter_element.cu:22-388.  Nine central-difference
// passes g_a(f)[c] = ((0 + f[c+1]) - f[c-1]) * 0.5f, each zero-padded at the volume
// border: Hxx = g_x(g_x d), Hxy = g_y(g_x d), Hxz = g_z(g_x d), Hyy = g_y(g_y d),
// Hyz = g_z(g_y d), Hzz = g_z(g_z d).  Candidate: xx < 0, xx*yy - xy*xy > 0 and the 3x3
// determinant < 0, evaluated in C parse order without contraction (the library is
// built with -ffp-contract=off).
//
// MI355X design: one pass over `dog` instead of nine.  When the elements or the mask are
// asked for, a work-group streams a 64 x 8 column of slices through LDS rings (k_hessian),
// forms the first derivatives Hx, Hy, Hz on the halo positions the second pass needs (zero
// outside the volume, as the reference's second pass pads its own input), then writes the
// six elements (SoA, pitch N) and the mask with coalesced stores.  The whole-path step
// asks for neither: k_hessian_rec keeps everything in registers (a lane per column, DPP
// shifts for the x neighbours) and writes only the candidates' records.
//
// Reduced-precision storage (FP16 / BF16): the elements are computed in FP32 from the
// stored DoG and rounded only on the store; the candidate test uses the FP32 values, and
// when a leading minor lies within a band (1e-5 by default) of the sum of the magnitudes
// of its terms, the minors are re-evaluated in FP64 from the same FP32 elements.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int TX = kHessTX, TY = kHessTY, ZC = kHessTZ;
constexpr int NT = TX * TY / 2;  // two output rows per thread

__device__ __forceinline__ float g2(float hi, float lo) { return cdiff_half(hi, lo); }

__device__ __forceinline__ uint8_t negdef32(float xx, float xy, float xz, float yy, float yz,
                                            float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

// Reduced-precision forms: a leading minor within `band` of the sum of its terms'
// magnitudes has its sign re-decided in FP64 from the same FP32 elements.  The FP32
// evaluation error of these minors is at most about 6 * 2^-24 (3.6e-7) of that sum (two
// or three rounded products per term, four rounded additions), so the default band of
// 1e-5 covers every sign the FP32 rounding can flip, with a margin of about 28.  The FP32
// minors and the term magnitudes share their products with negdef32's evaluation (same
// order, same roundings).
__device__ __forceinline__ uint8_t negdef_band(float xx, float xy, float xz, float yy, float yz,
                                               float zz, float band) {
  const float p1 = xx * yy, p2 = xy * xy;
  const float m2 = p1 - p2;
  const float t1 = p1 * zz, t2 = 2 * xy * yz * xz, t3 = xx * yz * yz, t4 = yy * xz * xz,
              t5 = zz * xy * xy;
  const float m3 = t1 + t2 - t3 - t4 - t5;
  const float s2 = fabsf(p1) + fabsf(p2);
  const float s3 = fabsf(t1) + fabsf(t2) + fabsf(t3) + fabsf(t4) + fabsf(t5);
  if (fabsf(m2) < band * s2 || fabsf(m3) < band * s3) {
    const double a = xx, b = xy, c = xz, d = yy, e = yz, f = zz;
    const double n2 = a * d - b * b;
    const double n3 = a * d * f + 2 * b * e * c - a * e * e - d * c * c - f * b * b;
    return (a < 0 && n2 > 0 && n3 < 0) ? 1 : 0;
  }
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

// z-streaming form.  A work-group owns a TX x TY column of outputs over ZC consecutive
// slices z0..z1-1 and walks s = z0 .. z1+1 with three-slot LDS rings:
//   step s: Hx(s-1), Hy(s-1) from dog(s-1) (xy halo 1), Hz(s-1) from dog(s) - dog(s-2)
//           ->  outputs of slice z = s-2, which need Hx / Hy at z-1..z+1 and Hz at
//           z-1, z+1  ->  dog(s+1) (xy halo 2) into the slot of dog(s-2), and the loads of
//           dog(s+2) issued into registers.
// Every dog value is read once per column (plus the xy halo, served by L2) instead of
// (TZ+4)/TZ times; first derivatives outside the volume are zero as in the reference's
// second pass.  Each thread's positions in the slices do not change with z, so their
// LDS and global offsets and their bounds tests are computed once, before the walk; ring
// slots rotate instead of being recomputed; global indices are 32-bit (N < 2^32).
// (the second bound keeps the register allocation within five waves per SIMD, the most
// that the LDS rings allow)
template <typename TD, typename TS, bool BAND>
__global__ __launch_bounds__(NT, 5) void k_hessian(const TD* __restrict__ dog, TS* __restrict__ hess,
                                                uint8_t* __restrict__ pd, int W, int H, int D,
                                                int nbx, int nby, float* __restrict__ cand,
                                                uint32_t* __restrict__ cand_cnt, float band) {
  constexpr int DX = TX + 4, DY = TY + 4;        // dog slice with halo 2
  constexpr int XX = TX + 2, XY = TY + 2;        // Hx slice with halo 1 in x and y
  constexpr int YX = TX, YY = TY + 2;            // Hy slice with halo 1 in y
  constexpr int NPF = (DY * DX + NT - 1) / NT;   // dog elements per thread and slice
  constexpr int NHX = (XY * XX + NT - 1) / NT, NHY = (YY * YX + NT - 1) / NT;
  constexpr int NHZ = TY * TX / NT, NOUT = TY / (NT / TX);
  __shared__ float sD[3][DY * DX];
  __shared__ float sHx[3][XY * XX];
  __shared__ float sHy[3][YY * YX];
  __shared__ float sHz[3][TY * TX];
  __shared__ uint32_t s_cnt;  // candidate records written by this work-group
  if (threadIdx.x == 0) s_cnt = 0;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  float* rec = cand ? cand + (size_t)tile * kCandTile * 7 : nullptr;
  const int lane = threadIdx.x & 63;
  unsigned b = tile;
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int x0 = bx * TX, y0 = by * TY;
  const int z0 = (int)b * ZC, z1 = min(z0 + ZC, D);
  const uint32_t Wu = (uint32_t)W, WH = Wu * (uint32_t)H;
  auto inplane = [&](int gx, int gy) { return gx >= 0 && gx < W && gy >= 0 && gy < H; };

  // slice-invariant addressing of this thread: the byte offset in the slice; positions
  // outside the plane read past the slice descriptor's record count
  uint32_t fo[NPF];
#pragma unroll
  for (int k = 0; k < NPF; ++k) {
    const int e = threadIdx.x + k * NT, lx = e % DX, ly = e / DX;
    const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
    const bool fv = e < DY * DX && inplane(gx, gy);
    const uint32_t i = (uint32_t)gy * Wu + (uint32_t)gx;
    fo[k] = fv ? i * (uint32_t)sizeof(TD) : 0xFFFFFFF0u;
  }
  int hxi[NHX], hyi[NHY], hzi[NHZ];
  bool hxe[NHX], hxv[NHX], hye[NHY], hyv[NHY], hzv[NHZ];
#pragma unroll
  for (int j = 0; j < NHX; ++j) {
    const int e = threadIdx.x + j * NT, lx = e % XX, ly = e / XX;
    hxe[j] = e < XY * XX;
    hxv[j] = hxe[j] && inplane(x0 - 1 + lx, y0 - 1 + ly);
    hxi[j] = hxe[j] ? (ly + 1) * DX + lx : 0;  // Hx = g2(d[hxi + 2], d[hxi])
  }
#pragma unroll
  for (int j = 0; j < NHY; ++j) {
    const int e = threadIdx.x + j * NT, lx = e % YX, ly = e / YX;
    hye[j] = e < YY * YX;
    hyv[j] = hye[j] && inplane(x0 + lx, y0 - 1 + ly);
    hyi[j] = hye[j] ? ly * DX + lx + 2 : 0;  // Hy = g2(d[hyi + 2 DX], d[hyi])
  }
#pragma unroll
  for (int j = 0; j < NHZ; ++j) {
    const int e = threadIdx.x + j * NT, lx = e % TX, ly = e / TX;
    hzv[j] = inplane(x0 + lx, y0 + ly);
    hzi[j] = (ly + 2) * DX + lx + 2;  // Hz = g2(dn[hzi], dp[hzi])
  }
  const int tx = threadIdx.x % TX, ty0 = threadIdx.x / TX;  // outputs (tx, ty0 + k*(NT/TX))
  bool ov[NOUT];
  uint32_t po[NOUT];
#pragma unroll
  for (int k = 0; k < NOUT; ++k) {
    const int gx = x0 + tx, gy = y0 + ty0 + k * (NT / TX);
    ov[k] = gx < W && gy < H;
    po[k] = ov[k] ? (uint32_t)gy * Wu + (uint32_t)gx : 0u;
  }

  // dog slices go global -> registers one step ahead -> LDS ring, as raw buffer loads
  // through a descriptor of the slice (a position outside the plane has an offset past
  // the record count, a slice outside the volume has no records, so both read zero) with
  // no branch around them; 16-bit storage is converted when the slice is written to LDS
  // (converting at the load made every load wait where it was issued).
  // a queue of three register slices: dog(sl) waits in the slot of its LDS ring slot, so
  // it is loaded three steps before it enters LDS
  typename StoredBits<TD>::type pf[3][NPF];
  const uint32_t slice_bytes = WH * (uint32_t)sizeof(TD);
  auto fetch = [&](int sl, auto q_c) {
    constexpr int q = decltype(q_c)::value;
    // slices past the volume, and past z1 + 1 (the last one this column reads), have no
    // records: their loads return zero without touching memory
    const bool zin = sl >= 0 && sl < D && sl <= z1 + 1;
    const __amdgpu_buffer_rsrc_t r =
        buf_rsrc(dog + (size_t)(zin ? sl : 0) * WH, zin ? slice_bytes : 0u);
#pragma unroll
    for (int k = 0; k < NPF; ++k) pf[q][k] = buf_load<TD>(r, fo[k]);
  };
  auto put = [&](float* d, auto q_c) {
    constexpr int q = decltype(q_c)::value;
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < DY * DX) d[e] = from_bits<TD>(pf[q][k]);
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  // Both rings rotate with period 3, so the walk is unrolled by three steps whose slots
  // are compile-time constants: every LDS access is then a per-thread offset plus an
  // immediate.  Slots at a step: dog(s-2), dog(s-1), dog(s) in DP, DC, DN; H(s-3), H(s-2)
  // in DP, DC and H(s-1) written into DN.
  fetch(z0 - 2, I0());
  fetch(z0 - 1, I1());
  fetch(z0, I2());
  put(sD[0], I0());
  put(sD[1], I1());
  put(sD[2], I2());
  fetch(z0 + 1, I0());
  fetch(z0 + 2, I1());
  fetch(z0 + 3, I2());
  __syncthreads();
  auto step = [&](int s, auto dp_c, auto dc_c, auto dn_c) {
    constexpr int dp = decltype(dp_c)::value, dc = decltype(dc_c)::value,
                  dn = decltype(dn_c)::value;
    // 1. first derivatives of slice s-1
    {
      const bool zin = s - 1 >= 0 && s - 1 < D;
      const float* Dc = sD[dc];
      const float* Dn = sD[dn];
      const float* Dp = sD[dp];
      float* hx = sHx[dn];
      float* hy = sHy[dn];
      float* hz = sHz[dn];
      // every operand is read first (all offsets are inside the slot), then the zero
      // padding is a select: a read under the position's condition would wait for its
      // own LDS round trip before the next one is issued
      float xa[NHX], xb[NHX], ya[NHY], yb[NHY], za[NHZ], zb[NHZ];
#pragma unroll
      for (int j = 0; j < NHX; ++j) {
        xa[j] = Dc[hxi[j] + 2];
        xb[j] = Dc[hxi[j]];
      }
#pragma unroll
      for (int j = 0; j < NHY; ++j) {
        ya[j] = Dc[hyi[j] + 2 * DX];
        yb[j] = Dc[hyi[j]];
      }
#pragma unroll
      for (int j = 0; j < NHZ; ++j) {
        za[j] = Dn[hzi[j]];
        zb[j] = Dp[hzi[j]];
      }
#pragma unroll
      for (int j = 0; j < NHX; ++j) {
        const float v = g2(xa[j], xb[j]);
        if (hxe[j]) hx[threadIdx.x + j * NT] = (zin && hxv[j]) ? v : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < NHY; ++j) {
        const float v = g2(ya[j], yb[j]);
        if (hye[j]) hy[threadIdx.x + j * NT] = (zin && hyv[j]) ? v : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < NHZ; ++j) {
        const float v = g2(za[j], zb[j]);
        hz[threadIdx.x + j * NT] = (zin && hzv[j]) ? v : 0.0f;
      }
    }
    __syncthreads();
    // 2. outputs of slice z = s - 2
    const int z = s - 2;
    if (z >= z0 && z < z1) {
      const float* hxm = sHx[dp];  // Hx(z-1)
      const float* hx0 = sHx[dc];  // Hx(z)
      const float* hxp = sHx[dn];  // Hx(z+1)
      const float* hym = sHy[dp];
      const float* hy0 = sHy[dc];
      const float* hyp = sHy[dn];
      const float* hzm = sHz[dp];
      const float* hzp = sHz[dn];
      const uint32_t zb = (uint32_t)z * WH;
      // the elements of every output row first (positions outside the plane read the
      // zero-padded slots and are dropped below), so that all their reads are in flight
      // together
      float el[NOUT][6];
#pragma unroll
      for (int k = 0; k < NOUT; ++k) {
        const int ty = ty0 + k * (NT / TX);
        const int cx = (ty + 1) * XX + tx + 1;  // Hx centre
        const int cy = (ty + 1) * YX + tx;      // Hy centre
        const int cz = ty * TX + tx;            // Hz centre
        el[k][0] = g2(hx0[cx + 1], hx0[cx - 1]);
        el[k][1] = g2(hx0[cx + XX], hx0[cx - XX]);
        el[k][2] = g2(hxp[cx], hxm[cx]);
        el[k][3] = g2(hy0[cy + YX], hy0[cy - YX]);
        el[k][4] = g2(hyp[cy], hym[cy]);
        el[k][5] = g2(hzp[cz], hzm[cz]);
      }
#pragma unroll
      for (int k = 0; k < NOUT; ++k) {
        const int ty = ty0 + k * (NT / TX);
        bool is_cand = false;
        float h6[6];
        uint32_t cidx = 0;
        if (ov[k]) {
          const float xx = el[k][0], xy = el[k][1], xz = el[k][2], yy = el[k][3],
                      yz = el[k][4], zz = el[k][5];
          const uint32_t idx = zb + po[k];
          if (hess) {  // NULL: mask only (the region pass reads the candidate records)
            const i64 N = (i64)WH * D;
            hess[idx] = cvt<TS>(xx);
            hess[N + idx] = cvt<TS>(xy);
            hess[2 * N + idx] = cvt<TS>(xz);
            hess[3 * N + idx] = cvt<TS>(yy);
            hess[4 * N + idx] = cvt<TS>(yz);
            hess[5 * N + idx] = cvt<TS>(zz);
          }
          const uint8_t m =
              BAND ? negdef_band(xx, xy, xz, yy, yz, zz, band) : negdef32(xx, xy, xz, yy, yz, zz);
          if (pd) pd[idx] = m;
          // voxel 0 is background whatever its mask (its label is 0): no record
          is_cand = m != 0 && idx != 0 && rec != nullptr;
          if (is_cand) {
            h6[0] = round_to<TS>(xx); h6[1] = round_to<TS>(xy); h6[2] = round_to<TS>(xz);
            h6[3] = round_to<TS>(yy); h6[4] = round_to<TS>(yz); h6[5] = round_to<TS>(zz);
            cidx = (uint32_t)(tx + TX * (ty + TY * (z - z0)));  // position in the tile
          }
        }
        // candidate records: one LDS reservation per wave (a wave is one row of 64 x),
        // consecutive slots for its candidates in lane order
        const unsigned long long bal = __ballot(is_cand);
        if (bal) {
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(&s_cnt, (uint32_t)__popcll(bal));
          base = __builtin_amdgcn_readfirstlane(base);  // lane 0 (the whole wave is here)
          if (is_cand) {
            const uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            rec[slot] = __uint_as_float(cidx);
#pragma unroll
            for (int q = 0; q < 6; ++q) rec[(size_t)kCandTile * (1 + q) + slot] = h6[q];
          }
        }
      }
    }
    // 3. dog(s+1) into the slot of dog(s-2) (last read above); its register slot takes
    // the loads of dog(s+4)
    put(sD[dp], decltype(dp_c)());
    fetch(s + 4, decltype(dp_c)());
    __syncthreads();
  };
  for (int s = z0; s <= z1 + 1; s += 3) {  // the step count is uniform over the work-group
    step(s, I0(), I1(), I2());
    if (s + 1 > z1 + 1) break;
    step(s + 1, I1(), I2(), I0());
    if (s + 2 > z1 + 1) break;
    step(s + 2, I2(), I0(), I1());
  }
  if (cand_cnt && threadIdx.x == 0) cand_cnt[tile] = s_cnt;
}

// ---- records-only form -----------------------------------------------------------------
// When neither the element volumes nor the byte mask are wanted (the whole-path step), the
// elements only decide the candidate test and fill the records.  The sign of a zero
// element reaches neither: the tests compare with zero (a -0 and a +0 compare alike) and
// the region sums add records into accumulators that start at +0.  So every difference is
// taken as (hi - lo) * 0.5f, without the leading 0 + of the per-stage form, which only
// turns a -0 into +0; every non-zero value is the same, bit for bit.
//
// Layout: a wave owns 64 x-adjacent columns (one lane each) and kRecTY consecutive output
// rows; a work-group is the kHessTY / kRecTY waves of one Hessian tile row block, walking
// kRecNZ Hessian tiles in z.  Everything a lane needs in y and z stays in its registers:
// dog rows y-2 .. y+kRecTY+1 of each slice arrive from global memory (one row descriptor
// each, so that rows and slices outside the volume, and columns outside the plane, read
// zero), Hy and the y second differences are register arithmetic, and the z differences
// walk three-slice register rings.  The x neighbours come from the adjacent lanes by DPP
// wave shifts.  The columns beyond the wave come from one more load per row (`halo`),
// whose lanes 0, 1, 62 and 63 hold columns x0-1, x0-2, x0+65 and x0+64: a shift keeps
// the destination's own value in the lane without a source, so copies of the halo are
// the destinations of the left and right shifts, and a pair swap of it brings x0-2 /
// x0+65 to lanes 0 / 63 for the first differences just outside the wave, Hx(x0-1) and
// Hx(x0+64).  No LDS beyond the record counters, no barrier inside the walk.  (Two halo
// loads per row, one per shift, save the copies but cost the registers of a fourth wave
// per SIMD, which measured slower.)
// A work-group spans kRecTilesY Hessian tiles in y: the dog rows above and below each
// wave's rows are then mostly read by the same work-group at the same time (its y halo is
// 4 rows for 8 kRecTilesY output rows); with one tile per work-group the halo rows were
// read again from HBM by the work-group of the neighbouring tile, which runs at another z.
#ifndef HDOG_REC_TILES_Y
#define HDOG_REC_TILES_Y 2
#endif
constexpr int kRecTY = 4;                     // output rows per lane
constexpr int kRecTilesY = HDOG_REC_TILES_Y;  // Hessian tiles per work-group, along y
constexpr int kRecWaves = kRecTilesY * kHessTY / kRecTY;  // waves per work-group
constexpr int kRecNZ = 4;                     // Hessian tiles per work-group, along z
static_assert(kHessTX == 64 && kHessTY % kRecTY == 0, "one wave spans a tile row");

// v of lane i - 1 (lane 0 keeps old) and of lane i + 1 (lane 63 keeps old); one DPP move
// each, written over old.  The whole wave must be active.
__device__ __forceinline__ float from_left(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
// lanes 2k and 2k + 1 exchange v
__device__ __forceinline__ float swap_pairs(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf, 0xf, false));
}

template <typename TD, typename TS, bool BAND>
__global__ __launch_bounds__(64 * kRecWaves) void k_hessian_rec(
    const TD* __restrict__ dog, int W, int H, int D, int nbx, int nby, int nbyg,
    float* __restrict__ cand,
    uint32_t* __restrict__ cand_cnt, float band) {
  constexpr int TY = kRecTY, NR = TY + 4, NH = TY + 2;  // dog rows, first-difference rows
  __shared__ uint32_t s_cnt[kRecTilesY][kRecNZ];
  if (threadIdx.x < kRecTilesY * kRecNZ) s_cnt[threadIdx.x / kRecNZ][threadIdx.x % kRecNZ] = 0u;
  const unsigned g = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = (int)(g % (unsigned)nbx), byg = (int)(g / (unsigned)nbx % (unsigned)nbyg);
  const int gz = (int)(g / (unsigned)nbx / (unsigned)nbyg);
  const int lane = threadIdx.x & 63;
  // first output row of the wave in the tile (wave-uniform: the row descriptors stay scalar)
  const int wy0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * TY;
  const int tt = wy0 / kHessTY, ty0 = wy0 % kHessTY;  // the wave's tile, its row in the tile
  const int by = byg * kRecTilesY + tt;  // past nby: every row is below the plane
  const int x0 = bx * 64, x = x0 + lane, y = by * kHessTY + ty0;
  const int z0 = gz * kRecNZ * kHessTZ, z1 = min(z0 + kRecNZ * kHessTZ, D);
  constexpr uint32_t SZ = sizeof(TD);
  // byte offsets in a row: a negative column wraps past the record count
  const uint32_t vo = (uint32_t)x * SZ;
  const int hc = lane == 1 ? x0 - 2 : lane == 62 ? x0 + 65 : lane == 63 ? x0 + 64 : x0 - 1;
  const uint32_t ho = (uint32_t)hc * SZ;
  const bool xin = x < W;
  const bool hin = lane == 0 ? x0 > 0 : (lane == 63 ? x0 + 64 < W : true);
  const float hsign = lane == 0 ? -0.5f : 0.5f;  // Hx(x0-1) = -(d(x0-2) - d(x0)) / 2
  const uint32_t WH = (uint32_t)W * (uint32_t)H;

  using Bits = typename StoredBits<TD>::type;
  Bits qd[3][NR], qh[3][NH];  // slices in flight: dog rows, halo columns
  float Hx[3][NH], Hy[3][NH], Hz[3][TY], Hxx[3][TY];
  const uint32_t row_bytes = (uint32_t)W * SZ;
  // CLIP: some of the wave's dog rows lie outside the plane
  auto fetch = [&](int s, auto q_c, auto clip_c) {
    constexpr int q = decltype(q_c)::value;
    constexpr bool CLIP = decltype(clip_c)::value;
    // slices past the walk's last one are loaded too, with no records: every step issues
    // the same loads
    const uint32_t nrec = s >= 0 && s < D && s <= z1 + 1 ? row_bytes : 0u;
    // rows and slices outside the volume get no records (their addresses are not used)
    const char* rows = (const char*)dog + ((int64_t)s * WH + (int64_t)(y - 2) * W) * SZ;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t n = !CLIP || (unsigned)(y - 2 + r) < (unsigned)H ? nrec : 0u;
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(rows + r * row_bytes, n);
      qd[q][r] = buf_load<TD>(rs, vo);
      if (r >= 1 && r <= NH) {
        qh[q][r - 1] = buf_load<TD>(rs, ho);
      }
    }
  };
  __syncthreads();  // counters cleared

  auto step = [&](int s, auto a_c, auto edge_c, auto clip_c) {
    constexpr int A = decltype(a_c)::value;  // ring slot of slice s (and of s - 3)
    constexpr int B = (A + 2) % 3;           // slot of s - 1 (and of s + 2)
    constexpr int C = (A + 1) % 3;           // slot of s - 2 (and of s + 1)
    constexpr bool EDGE = decltype(edge_c)::value;  // a column of the wave or halo is outside
    constexpr bool CLIP = decltype(clip_c)::value;  // a row of the wave or halo is outside
    float d[NR], h[NH];
#pragma unroll
    for (int r = 0; r < NR; ++r) d[r] = from_bits<TD>(qd[A][r]);
#pragma unroll
    for (int j = 0; j < NH; ++j) h[j] = from_bits<TD>(qh[A][j]);
    // outputs of slice z = s - 2
    const int z = s - 2;
    if (z >= z0 && z < z1) {
      const int tz = (z - z0) / kHessTZ;  // Hessian tile of the slice within the walk
      const unsigned tile = (unsigned)(((gz * kRecNZ + tz) * nby + by) * nbx + bx);
      const uint32_t lz = (uint32_t)(z - (z0 + tz * kHessTZ));
      const __amdgpu_buffer_rsrc_t rr =
          buf_rsrc(cand + (size_t)tile * kCandTile * 7, (uint32_t)kCandTile * 7u * 4u);
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        const float xx = Hxx[C][r];
        const float xy = (Hx[C][r + 2] - Hx[C][r]) * 0.5f;
        const float xz = (Hx[B][r + 1] - Hx[A][r + 1]) * 0.5f;
        const float yy = (Hy[C][r + 2] - Hy[C][r]) * 0.5f;
        const float yz = (Hy[B][r + 1] - Hy[A][r + 1]) * 0.5f;
        const float zz = (Hz[B][r] - Hz[A][r]) * 0.5f;
        bool m;
        if (BAND) {
          m = negdef_band(xx, xy, xz, yy, yz, zz, band) != 0;
        } else {  // negdef32's minors, without branches
          const float m2 = xx * yy - xy * xy;
          const float m3 =
              xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
          m = (xx < 0.0f) & (m2 > 0.0f) & (m3 < 0.0f);
        }
        // voxel 0 is background whatever its mask (its label is 0): no record (only an
        // EDGE and CLIP wave holds it)
        const bool is_c = m && (!EDGE || xin) && (!CLIP || y + r < H) &&
                          (!(EDGE && CLIP) || (x | (y + r) | z) != 0);
        // one LDS reservation per wave and row, consecutive slots in lane order (ds_append
        // would do both in one instruction, but as compiled for this target it adds to the
        // counter at LDS address 0 whatever the address asked for)
        const unsigned long long act = __ballot(is_c);
        uint32_t base = 0;
        if (act && lane == 0) base = atomicAdd(&s_cnt[tt][tz], (uint32_t)__popcll(act));
        base = __builtin_amdgcn_readfirstlane(base);
        if (is_c) {
          const uint32_t slot = base + __builtin_amdgcn_mbcnt_hi(
                                           (uint32_t)(act >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
          const uint32_t cidx = (uint32_t)lane + 64u * ((uint32_t)(ty0 + r) + kHessTY * lz);
          const int vb = (int)(slot * 4u);
          const float h6[6] = {round_to<TS>(xx), round_to<TS>(xy), round_to<TS>(xz),
                               round_to<TS>(yy), round_to<TS>(yz), round_to<TS>(zz)};
          __builtin_amdgcn_raw_buffer_store_b32(cidx, rr, vb, 0, 0);
#pragma unroll
          for (int q = 0; q < 6; ++q)
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(h6[q]), rr, vb,
                                                  (1 + q) * kCandTile * 4, 0);
        }
      }
    }
    // first differences of slice s, and its x second difference
    if (s >= z0 - 1 && s <= z1) {
      float o[TY];  // lane 0: Hx(x0-1), lane 63: Hx(x0+64)
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        o[r] = (swap_pairs(h[r + 1]) - d[r + 2]) * hsign;
        if (EDGE && !hin) o[r] = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < NH; ++j) {
        const float c = d[j + 1];
        Hx[A][j] = (from_right(h[j], c) - from_left(h[j], c)) * 0.5f;
        if (EDGE && !xin) Hx[A][j] = 0.0f;
        Hy[A][j] = (d[j + 2] - d[j]) * 0.5f;
      }
      if (CLIP) {
#pragma unroll
        for (int j = 0; j < NH; ++j)
          if (y - 1 + j < 0 || y - 1 + j >= H) Hy[A][j] = 0.0f;
      }
      if (s >= z0 && s < z1) {
#pragma unroll
        for (int r = 0; r < TY; ++r) {
          const float c = Hx[A][r + 1];
          Hxx[A][r] = (from_right(o[r], c) - from_left(o[r], c)) * 0.5f;
        }
      }
    }
    // Hz(s), zero outside the volume, from dog(s + 1) (in flight since the last step) and
    // dog(s - 1), whose registers the next loads take over
    if (s >= z0 - 1 && s <= z1) {
#pragma unroll
      for (int r = 0; r < TY; ++r)
        Hz[A][r] = (from_bits<TD>(qd[C][r + 2]) - from_bits<TD>(qd[B][r + 2])) * 0.5f;
      if (s < 0 || s >= D) {
#pragma unroll
        for (int r = 0; r < TY; ++r) Hz[A][r] = 0.0f;
      }
    }
    fetch(s + 2, std::integral_constant<int, B>(), clip_c);
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  // steps z0-2 .. z1+1, rounded up to whole rotations (the steps past z1+1 load nothing
  // and compute nothing): one loop exit, so that the waits for a slice's loads never
  // also wait for the loads issued after it
  const int rotations = (z1 - z0 + 4 + 2) / 3;
  auto walk = [&](auto edge_c, auto clip_c) {
    fetch(z0 - 2, I0(), clip_c);
    fetch(z0 - 1, I1(), clip_c);
    for (int i = 0; i < rotations; ++i) {  // uniform
      const int s = z0 - 2 + 3 * i;
      step(s, I0(), edge_c, clip_c);
      step(s + 1, I1(), edge_c, clip_c);
      step(s + 2, I2(), edge_c, clip_c);
    }
  };
  // interior waves (the common case) run without the border tests
  if (x0 == 0 || x0 + 64 >= W || y - 2 < 0 || y + TY + 1 >= H)
    walk(std::true_type(), std::true_type());
  else
    walk(std::false_type(), std::false_type());
  __syncthreads();
  if (threadIdx.x < kRecTilesY * kRecNZ) {
    const int ct = (int)threadIdx.x / kRecNZ, cz = (int)threadIdx.x % kRecNZ;
    const int cby = byg * kRecTilesY + ct;
    if (cby < nby && z0 + cz * kHessTZ < D)
      cand_cnt[(((unsigned)gz * kRecNZ + cz) * nby + cby) * nbx + bx] = s_cnt[ct][cz];
  }
}

}  // namespace

template <typename TD, typename TH>
hipError_t launch_hessian(const TD* dog, TH* hess, uint8_t* pd, int W, int H, int D,
                          hipStream_t s, float* cand, uint32_t* cand_cnt, float band) {
  const int nbx = (W + TX - 1) / TX, nby = (H + TY - 1) / TY, nbz = (D + ZC - 1) / ZC;
  const i64 nblk = (i64)nbx * nby * nbz;
  // slices are read through buffer descriptors with 32-bit byte offsets
  if ((i64)W * H * (i64)sizeof(TD) >= 0xFFFFFFF0ll) return hipErrorInvalidValue;
  // reduced-precision storage anywhere on the path: FP64 re-evaluation band for the minors
  constexpr bool reduced = !std::is_same<TD, float>::value || !std::is_same<TH, float>::value;
  if (!hess && !pd && cand && cand_cnt) {  // records only
    const int nbyg = (nby + kRecTilesY - 1) / kRecTilesY;
    const i64 ngroups = (i64)nbx * nbyg * ((nbz + kRecNZ - 1) / kRecNZ);
    hipLaunchKernelGGL((k_hessian_rec<TD, TH, reduced>), dim3((unsigned)ngroups),
                       dim3(64 * kRecWaves), 0, s, dog, W, H, D, nbx, nby, nbyg, cand, cand_cnt,
                       band);
  } else {
    hipLaunchKernelGGL((k_hessian<TD, TH, reduced>), dim3((unsigned)nblk), dim3(NT), 0, s, dog,
                       hess, pd, W, H, D, nbx, nby, cand, cand_cnt, band);
  }
  mark("hessian/hessian_pd", s);
  return hipGetLastError();
}

#define HDOG_INST_HESS(TD, TH)                                                          \
  template hipError_t launch_hessian<TD, TH>(const TD*, TH*, uint8_t*, int, int, int, \
                                             hipStream_t, float*, uint32_t*, float);
HDOG_INST_HESS(float, float)
HDOG_INST_HESS(__half, __half)
HDOG_INST_HESS(__hip_bfloat16, __hip_bfloat16)
HDOG_INST_HESS(float, __half)
HDOG_INST_HESS(float, __hip_bfloat16)
#undef HDOG_INST_HESS

}  // namespace hdog
