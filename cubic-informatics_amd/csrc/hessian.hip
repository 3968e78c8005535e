// hessian.hip -- Hessian elements and the negative-definite candidate mask, one kernel.
//
// Reference semantics: src/3D/Hessian3DFilter_element.cu:22-388.  Nine central-difference
// passes g_a(f)[c] = ((0 + f[c+1]) - f[c-1]) * 0.5f, each zero-padded at the volume
// border: Hxx = g_x(g_x d), Hxy = g_y(g_x d), Hxz = g_z(g_x d), Hyy = g_y(g_y d),
// Hyz = g_z(g_y d), Hzz = g_z(g_z d).  Candidate: xx < 0, xx*yy - xy*xy > 0 and the 3x3
// determinant < 0, evaluated in C parse order without contraction (the library is
// built with -ffp-contract=off).
//
// MI355X design: one pass over `dog` instead of nine.  A work-group streams a 64 x 8
// column of slices through LDS rings (see k_hessian), forms the first derivatives Hx, Hy,
// Hz on the halo positions the second pass needs (zero outside the volume, as the
// reference's second pass pads its own input), then writes the six elements (SoA,
// pitch N) and the mask with coalesced stores.
//
// Reduced-precision storage (FP16 / BF16): the elements are computed in FP32 from the
// stored DoG and rounded only on the store; the candidate test uses the FP32 values, and
// when a leading minor lies within 1e-3 of the sum of the magnitudes of its terms, the
// minors are re-evaluated in FP64 from the same FP32 elements.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int TX = kHessTX, TY = kHessTY, ZC = kHessTZ;
constexpr int NT = 256;

__device__ __forceinline__ float g2(float hi, float lo) { return cdiff_half(hi, lo); }

__device__ __forceinline__ uint8_t negdef32(float xx, float xy, float xz, float yy, float yz,
                                            float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

__device__ __forceinline__ uint8_t negdef_band(float xx, float xy, float xz, float yy, float yz,
                                               float zz) {
  const float m2 = xx * yy - xy * xy;
  const float m3 = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
  const float s2 = fabsf(xx * yy) + fabsf(xy * xy);
  const float s3 = fabsf(xx * yy * zz) + fabsf(2 * xy * yz * xz) + fabsf(xx * yz * yz) +
                   fabsf(yy * xz * xz) + fabsf(zz * xy * xy);
  if (fabsf(m2) < 1e-3f * s2 || fabsf(m3) < 1e-3f * s3) {
    const double a = xx, b = xy, c = xz, d = yy, e = yz, f = zz;
    const double n2 = a * d - b * b;
    const double n3 = a * d * f + 2 * b * e * c - a * e * e - d * c * c - f * b * b;
    return (a < 0 && n2 > 0 && n3 < 0) ? 1 : 0;
  }
  return (xx < 0 && m2 > 0 && m3 < 0) ? 1 : 0;
}

// z-streaming form.  A work-group owns a TX x TY column of outputs over ZC consecutive
// slices and walks s = z0-2 .. z1+1 with three-slot LDS rings:
//   step s: Hx(s-1), Hy(s-1) from dog(s-1) (xy halo 1), Hz(s-1) from dog(s) - dog(s-2)
//           ->  outputs of slice z = s-2, which need Hx / Hy at z-1..z+1 and Hz at
//           z-1, z+1  ->  dog(s+1) (xy halo 2) into the ring, fetched one step earlier.
// Every dog value is read once per column (plus the xy halo, served by L2) instead of
// (TZ+4)/TZ times; first derivatives outside the volume are zero as in the reference's
// second pass.
template <typename TD, typename TS, bool BAND>
__global__ __launch_bounds__(NT) void k_hessian(const TD* __restrict__ dog, TS* __restrict__ hess,
                                                uint8_t* __restrict__ pd, int W, int H, int D,
                                                int nbx, int nby, float* __restrict__ cand,
                                                uint32_t* __restrict__ cand_cnt) {
  constexpr int DX = TX + 4, DY = TY + 4;        // dog slice with halo 2
  constexpr int XX = TX + 2, XY = TY + 2;        // Hx slice with halo 1 in x and y
  constexpr int YX = TX, YY = TY + 2;            // Hy slice with halo 1 in y
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
  const i64 WH = (i64)W * H;
  const i64 N = WH * D;
  const int tx = threadIdx.x % TX, ty0 = threadIdx.x / TX;  // outputs (tx, ty0 + k*(NT/TX))
  // dog slices go global -> registers one step ahead -> LDS ring, so the loads of slice
  // s+2 are in flight while slice s is processed
  constexpr int NPF = (DY * DX + NT - 1) / NT;
  float pf[NPF];
  auto fetch = [&](int sl) {
    const bool zin = sl >= 0 && sl < D;
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int e = threadIdx.x + k * NT;
      const int lx = e % DX, ly = e / DX;
      const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
      pf[k] = (e < DY * DX && zin && gx >= 0 && gx < W && gy >= 0 && gy < H)
                  ? ld(dog, (i64)sl * WH + (i64)gy * W + gx)
                  : 0.0f;
    }
  };
  auto put = [&](int sl) {
    float* d = sD[(sl + 3) % 3];
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < DY * DX) d[e] = pf[k];
    }
  };
  fetch(z0 - 2);
  put(z0 - 2);
  fetch(z0 - 1);
  __syncthreads();
  for (int s = z0 - 2; s <= z1 + 1; ++s) {
    // invariant: dog(s-2..s) are in LDS, pf holds dog(s+1)
    // 1. first derivatives of slice s-1 (Hz from dog(s) - dog(s-2))
    {
      const int zc = s - 1;
      const bool zin = zc >= 0 && zc < D;
      const float* dc = sD[(zc + 3) % 3];  // dog(s-1)
      const float* dn = sD[(s + 3) % 3];   // dog(s)
      const float* dp = sD[(s + 1) % 3];   // dog(s-2)
      float* hx = sHx[(zc + 3) % 3];
      float* hy = sHy[(zc + 3) % 3];
      float* hz = sHz[(zc + 3) % 3];
      for (int e = threadIdx.x; e < XY * XX; e += NT) {
        const int lx = e % XX, ly = e / XX;
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
        float v = 0.0f;
        if (zin && gx >= 0 && gx < W && gy >= 0 && gy < H)
          v = g2(dc[(ly + 1) * DX + lx + 2], dc[(ly + 1) * DX + lx]);
        hx[e] = v;
      }
      for (int e = threadIdx.x; e < YY * YX; e += NT) {
        const int lx = e % YX, ly = e / YX;
        const int gx = x0 + lx, gy = y0 - 1 + ly;
        float v = 0.0f;
        if (zin && gx < W && gy >= 0 && gy < H)
          v = g2(dc[(ly + 2) * DX + lx + 2], dc[ly * DX + lx + 2]);
        hy[e] = v;
      }
      for (int e = threadIdx.x; e < TY * TX; e += NT) {
        const int lx = e % TX, ly = e / TX;
        const int gx = x0 + lx, gy = y0 + ly;
        float v = 0.0f;
        if (zin && gx < W && gy < H)
          v = g2(dn[(ly + 2) * DX + lx + 2], dp[(ly + 2) * DX + lx + 2]);
        hz[e] = v;
      }
    }
    __syncthreads();
    // 2. outputs of slice z = s - 2; then dog(s+1) into the slot of dog(s-2) (last read
    //    above) and start fetching dog(s+2)
    const int z = s - 2;
    if (z >= z0 && z < z1) {
      const float* hxm = sHx[(z + 2) % 3];  // Hx(z-1)
      const float* hx0 = sHx[(z + 3) % 3];  // Hx(z)
      const float* hxp = sHx[(z + 4) % 3];  // Hx(z+1)
      const float* hym = sHy[(z + 2) % 3];
      const float* hy0 = sHy[(z + 3) % 3];
      const float* hyp = sHy[(z + 4) % 3];
      const float* hzm = sHz[(z + 2) % 3];
      const float* hzp = sHz[(z + 4) % 3];
      const int gx = x0 + tx;
#pragma unroll
      for (int k = 0; k < TY / (NT / TX); ++k) {
        const int ty = ty0 + k * (NT / TX);
        const int gy = y0 + ty;
        bool is_cand = false;
        float h6[6];
        uint32_t cidx = 0;
        if (gx < W && gy < H) {
          const int cx = (ty + 1) * XX + tx + 1;  // Hx centre
          const int cy = (ty + 1) * YX + tx;      // Hy centre
          const int cz = ty * TX + tx;            // Hz centre
          const float xx = g2(hx0[cx + 1], hx0[cx - 1]);
          const float xy = g2(hx0[cx + XX], hx0[cx - XX]);
          const float xz = g2(hxp[cx], hxm[cx]);
          const float yy = g2(hy0[cy + YX], hy0[cy - YX]);
          const float yz = g2(hyp[cy], hym[cy]);
          const float zz = g2(hzp[cz], hzm[cz]);
          const i64 idx = (i64)z * WH + (i64)gy * W + gx;
          if (hess) {  // NULL: mask only (the region pass recomputes the elements)
            hess[idx] = cvt<TS>(xx);
            hess[N + idx] = cvt<TS>(xy);
            hess[2 * N + idx] = cvt<TS>(xz);
            hess[3 * N + idx] = cvt<TS>(yy);
            hess[4 * N + idx] = cvt<TS>(yz);
            hess[5 * N + idx] = cvt<TS>(zz);
          }
          const uint8_t m =
              BAND ? negdef_band(xx, xy, xz, yy, yz, zz) : negdef32(xx, xy, xz, yy, yz, zz);
          pd[idx] = m;
          is_cand = m != 0 && rec != nullptr;
          if (is_cand) {
            h6[0] = round_to<TS>(xx); h6[1] = round_to<TS>(xy); h6[2] = round_to<TS>(xz);
            h6[3] = round_to<TS>(yy); h6[4] = round_to<TS>(yz); h6[5] = round_to<TS>(zz);
            cidx = (uint32_t)idx;
          }
        }
        // candidate records: one LDS reservation per wave (a wave is one row of 64 x),
        // consecutive slots for its candidates in lane order
        const unsigned long long bal = __ballot(is_cand);
        if (bal) {
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(&s_cnt, (uint32_t)__popcll(bal));
          base = __shfl(base, 0, 64);
          if (is_cand) {
            const uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            rec[slot] = __uint_as_float(cidx);
#pragma unroll
            for (int q = 0; q < 6; ++q) rec[(size_t)kCandTile * (1 + q) + slot] = h6[q];
          }
        }
      }
    }
    put(s + 1);
    fetch(s + 2);
    __syncthreads();
  }
  if (cand_cnt && threadIdx.x == 0) cand_cnt[tile] = s_cnt;
}

}  // namespace

template <typename TD, typename TH>
hipError_t launch_hessian(const TD* dog, TH* hess, uint8_t* pd, int W, int H, int D,
                          hipStream_t s, float* cand, uint32_t* cand_cnt) {
  const int nbx = (W + TX - 1) / TX, nby = (H + TY - 1) / TY, nbz = (D + ZC - 1) / ZC;
  const i64 nblk = (i64)nbx * nby * nbz;
  // reduced-precision storage anywhere on the path: FP64 re-evaluation band for the minors
  constexpr bool band = !std::is_same<TD, float>::value || !std::is_same<TH, float>::value;
  hipLaunchKernelGGL((k_hessian<TD, TH, band>), dim3((unsigned)nblk), dim3(NT), 0, s, dog, hess,
                     pd, W, H, D, nbx, nby, cand, cand_cnt);
  mark("hessian/hessian_pd", s);
  return hipGetLastError();
}

#define HDOG_INST_HESS(TD, TH)                                                          \
  template hipError_t launch_hessian<TD, TH>(const TD*, TH*, uint8_t*, int, int, int, \
                                             hipStream_t, float*, uint32_t*);
HDOG_INST_HESS(float, float)
HDOG_INST_HESS(__half, __half)
HDOG_INST_HESS(__hip_bfloat16, __hip_bfloat16)
HDOG_INST_HESS(float, __half)
HDOG_INST_HESS(float, __hip_bfloat16)
#undef HDOG_INST_HESS

}  // namespace hdog
