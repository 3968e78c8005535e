// dog.hip -- separable anisotropic difference of Gaussians on gfx950.
//
// Reference semantics: src/3D/DoG3DFilter.cu:70-437.  Each Gaussian is x -> y (xy
// weights) -> z (z weights) with zero padding; every output is the FMA chain
// sum = fmaf(w[r-j], s[c+j], sum), j = -r..r, from 0.  dog = (G1 - G2) * normalizer.
//
// MI355X design: both Gaussians are evaluated from one read of each tile.  The x pass
// reads `norm` once and writes (G1_x, G2_x) as float2, the y pass works on the pair, and
// the z pass forms the difference directly, so `norm` is read once and the two
// intermediate volumes of the reference become one interleaved stream.  Intermediates
// stay FP32 whatever the storage type of norm / dog.
//
// Radius pairs (r1, r2) that occur in practice have compile-time kernels with the weights
// in scalar registers: the x pass makes 4 outputs per thread from 128-bit LDS reads, the
// y / z passes are column walkers (one thread per line, a register ring, every input
// read once, no LDS).  Other radii use the generic LDS-tile loops.
#include "hdog_internal.hpp"

namespace hdog {
namespace {

constexpr int NT = 256;

// ------------------------------------------------------------------- generic radii
template <int ROWS, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_x(const TS* __restrict__ src, float2* __restrict__ dst,
                                                const float* __restrict__ w1,
                                                const float* __restrict__ w2, int r1, int r2,
                                                int W, i64 nrows, int C, int nct) {
  extern __shared__ float smf[];
  const int R = max(r1, r2);
  const int span = C + 2 * R;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = tile % nct;
  const i64 row0 = (i64)(tile / nct) * ROWS;
  const int c0 = ct * C;
  stage_tile<8, NT, float>(
      ROWS * span,
      [&](int e) -> float {
        const int row = e / span, p = e % span;
        const i64 grow = row0 + row;
        const int x = c0 - R + p;
        return (grow < nrows && x >= 0 && x < W) ? ld(src, grow * W + x) : 0.0f;
      },
      [&](int e, float v) { smf[e] = v; });
  __syncthreads();
  for (int e = threadIdx.x; e < ROWS * C; e += NT) {
    const int row = e / C, i = e % C, x = c0 + i;
    const i64 grow = row0 + row;
    if (x >= W || grow >= nrows) continue;
    const float* t = smf + row * span + i + R;
    float s1 = 0.0f, s2 = 0.0f;
    for (int j = -r1; j <= r1; ++j) s1 = fmaf(w1[r1 - j], t[j], s1);
    for (int j = -r2; j <= r2; ++j) s2 = fmaf(w2[r2 - j], t[j], s2);
    dst[grow * W + x] = make_float2(s1, s2);
  }
}

template <int XW, bool LAST, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_line(const float2* __restrict__ src,
                                                   float2* __restrict__ dst, TS* __restrict__ dog,
                                                   const float* __restrict__ w1,
                                                   const float* __restrict__ w2, int r1, int r2,
                                                   float nz, int W, int L, i64 sa, i64 so, int C,
                                                   int nxt, int nct) {
  extern __shared__ float2 smf2[];
  const int R = max(r1, r2);
  const int span = C + 2 * R;
  unsigned b = blockIdx.x;
  const int xt = b % nxt; b /= nxt;
  const int ct = b % nct; b /= nct;
  const i64 base = (i64)b * so;
  const int x0 = xt * XW, c0 = ct * C;
  stage_tile<8, NT, float2>(
      span * XW,
      [&](int e) -> float2 {
        const int p = e / XW, x = x0 + e % XW, c = c0 - R + p;
        return (x < W && c >= 0 && c < L) ? src[base + (i64)c * sa + x] : make_float2(0.0f, 0.0f);
      },
      [&](int e, float2 v) { smf2[e] = v; });
  __syncthreads();
  for (int e = threadIdx.x; e < C * XW; e += NT) {
    const int i = e / XW, xl = e % XW, c = c0 + i, x = x0 + xl;
    if (c >= L || x >= W) continue;
    const float2* t = smf2 + (i + R) * XW + xl;
    float s1 = 0.0f, s2 = 0.0f;
    for (int j = -r1; j <= r1; ++j) s1 = fmaf(w1[r1 - j], t[j * XW].x, s1);
    for (int j = -r2; j <= r2; ++j) s2 = fmaf(w2[r2 - j], t[j * XW].y, s2);
    const i64 idx = base + (i64)c * sa + x;
    if (LAST)
      dog[idx] = cvt<TS>((s1 - s2) * nz);
    else
      dst[idx] = make_float2(s1, s2);
  }
}

// ------------------------------------------------------------- compile-time radii
// x pass: tile ROWS rows x (64*4) outputs; thread (row, seg) makes 4 consecutive outputs.
// LDS row position p <-> x = c0 - RP + p with RP = R rounded up to 4 (16-byte aligned
// window reads).
template <int R1, int R2, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_x_r(const TS* __restrict__ src,
                                                  float2* __restrict__ dst,
                                                  const float* __restrict__ w1,
                                                  const float* __restrict__ w2, int W, i64 nrows,
                                                  int nct) {
  constexpr int ROWS = 4, RUN = 4, C = NT / ROWS * RUN;
  constexpr int R = R1 > R2 ? R1 : R2;
  constexpr int RP = (R + 3) / 4 * 4;
  constexpr int SPAN = C + 2 * RP;
  constexpr int NV = (RUN + 2 * RP) / 4;  // float4 reads per window
  __shared__ __align__(16) float T[ROWS * SPAN];
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = tile % nct;
  const i64 row0 = (i64)(tile / nct) * ROWS;
  const int c0 = ct * C;
  stage_tile<8, NT, float>(
      ROWS * SPAN,
      [&](int e) -> float {
        const int row = e / SPAN, p = e % SPAN;
        const i64 grow = row0 + row;
        const int x = c0 - RP + p;
        return (grow < nrows && x >= 0 && x < W) ? ld(src, grow * W + x) : 0.0f;
      },
      [&](int e, float v) { T[e] = v; });
  __syncthreads();
  const int row = threadIdx.x / (NT / ROWS), seg = threadIdx.x % (NT / ROWS);
  const i64 grow = row0 + row;
  const int xs = c0 + seg * RUN;
  if (grow >= nrows || xs >= W) return;
  float win[4 * NV];
  const float4* T4 = reinterpret_cast<const float4*>(T + row * SPAN + seg * RUN);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float4 q = T4[v];
    win[4 * v] = q.x; win[4 * v + 1] = q.y; win[4 * v + 2] = q.z; win[4 * v + 3] = q.w;
  }
  float2 out[RUN];
#pragma unroll
  for (int o = 0; o < RUN; ++o) {
    // window index of x = xs + o + j is o + j + RP
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int j = -R1; j <= R1; ++j) s1 = fmaf(w1[R1 - j], win[o + j + RP], s1);
#pragma unroll
    for (int j = -R2; j <= R2; ++j) s2 = fmaf(w2[R2 - j], win[o + j + RP], s2);
    out[o] = make_float2(s1, s2);
  }
  float2* drow = dst + grow * W;
  // 16-byte stores (a wave writes 1 KB in two instructions) where the row start allows
  if (xs + RUN <= W && ((grow * W) & 1) == 0) {
#pragma unroll
    for (int o = 0; o < RUN; o += 2)
      *reinterpret_cast<float4*>(drow + xs + o) =
          make_float4(out[o].x, out[o].y, out[o + 1].x, out[o + 1].y);
  } else {
#pragma unroll
    for (int o = 0; o < RUN; ++o)
      if (xs + o < W) drow[xs + o] = out[o];
  }
}

// y / z pass as a column walker: one thread per line, a register ring of the last 2R+1
// (G1, G2) pairs (zero outside the line), unrolled by the ring period so every index is
// static; every input is read once and no LDS is used.  Ring refills are loaded PF steps
// before they enter the ring (PF divides the ring period: static queue slots), and the
// walk leaves the kernel at the line's end so no merged control path makes the compiler
// wait for the whole queue.
template <int R1, int R2, bool LAST, typename TS>
__global__ __launch_bounds__(NT) void k_gauss_walk(const float2* __restrict__ src,
                                                   float2* __restrict__ dst, TS* __restrict__ dog,
                                                   const float* __restrict__ w1,
                                                   const float* __restrict__ w2, float nz, int W,
                                                   int L, i64 sa, i64 so, int nxb) {
  constexpr int R = R1 > R2 ? R1 : R2;
  constexpr int WB = 2 * R + 1;
  constexpr int PF = WB % 3 == 0 ? 3 : WB % 5 == 0 ? 5 : WB % 7 == 0 ? 7 : 1;
  const int x = (blockIdx.x % nxb) * NT + threadIdx.x;
  if (x >= W) return;
  const i64 base = (i64)(blockIdx.x / nxb) * so + x;
  // zero padding in the load itself: position p is read through a buffer descriptor of
  // the wave's 64 columns at p (a uniform base), whose record count is 0 outside the
  // line, so those loads return zeros and no value is tested (or waited for) when it
  // enters the ring
  const int lane = (int)(threadIdx.x & 63);
  const int xw = __builtin_amdgcn_readfirstlane(x - lane);  // the wave's first column
  const i64 bw = (i64)(blockIdx.x / nxb) * so + xw;
  const int nrec = min(64, W - xw) * (int)sizeof(float2);
  auto at = [&](int p) -> float2 {
    const float2* line = src + bw + (i64)p * sa;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2*>(line), 0, (p >= 0 && p < L) ? nrec : 0, kBufferWord3);
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                          r, lane * (int)sizeof(float2), 0, 0));
  };
  float2 ring[WB];  // position p lives in ring[(p + R) % WB]
#pragma unroll
  for (int q = 0; q < WB; ++q) ring[q] = at(q - R);
  float2 pf[PF];  // the refill of step c (position c + R + 1) waits in slot c % PF
#pragma unroll
  for (int j = 0; j < PF; ++j) pf[j] = at(R + 1 + j);
  for (int c0 = 0;; c0 += WB) {
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      const int c = c0 + q;
      if (c >= L) return;
      float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
      for (int j = -R1; j <= R1; ++j) s1 = fmaf(w1[R1 - j], ring[(q + j + R) % WB].x, s1);
#pragma unroll
      for (int j = -R2; j <= R2; ++j) s2 = fmaf(w2[R2 - j], ring[(q + j + R) % WB].y, s2);
      const i64 idx = base + (i64)c * sa;
      if (LAST)
        st_nt(dog + idx, (s1 - s2) * nz);
      else
        st_nt(dst + idx, make_float2(s1, s2));
      ring[q] = pf[q % PF];  // position c + R + 1 replaces position c - R
      pf[q % PF] = at(c + R + 1 + PF);
    }
  }
}

// x and y passes fused: a one-wave work-group walks the rows of one plane for a strip of
// 64 columns.  Each row's window x0-16 .. x0+79 goes through an LDS line (two lines,
// alternating; the wave's own LDS order needs no barrier); lane l forms the x outputs
// (G1_x, G2_x) of column x0+l, which enter the register ring of the last 2R+1 rows
// exactly as in the column walker, and the y outputs leave as float2.  The (G1_x, G2_x)
// volume of the unfused passes is never stored.  Rows are fetched PF rows ahead of their
// use.  The Gaussian weights are symmetric (set pairwise from one value,
// w[r - i] == w[r + i]), so tap j reads w[r + |j|]: r + 1 distinct values per Gaussian,
// held in registers for the whole walk.
template <int R1, int R2, typename TS>
__global__ __launch_bounds__(64) void k_gauss_xy_walk(const TS* __restrict__ src,
                                                      float2* __restrict__ dst,
                                                      const float* __restrict__ w1,
                                                      const float* __restrict__ w2, int W, int H,
                                                      int nxs, int nys, int rows) {
  constexpr int R = R1 > R2 ? R1 : R2;
  constexpr int WB = 2 * R + 1;
  // rows fetched ahead: a divisor of WB, so the queue slot of every row is static in
  // the loop unrolled by WB
  constexpr int PF = WB % 3 == 0 ? 3 : WB % 5 == 0 ? 5 : WB <= 13 ? WB : 1;
  constexpr int HX = 16, LW = 128;  // line positions 96 .. 127 stay unused (zero)
  static_assert(R <= HX, "x window exceeds the LDS line");
  __shared__ float line[2][LW];
  // work-group = one 64-column strip of one plane over the rows [y0, y1) of one of the nys
  // segments of the plane (rows: a multiple of WB)
  const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
  const int y0 = (int)(t % nys) * rows, y1 = min(H, y0 + rows);
  const unsigned tile = t / nys;
  const int l = threadIdx.x;
  const int x0 = (int)(tile % nxs) * 64;
  const i64 pbase = (i64)(tile / nxs) * H * W;
  const int xa = x0 - HX + l, xb = x0 + 64 - HX + l;  // line positions l and 64 + l
  const bool va = xa >= 0 && xa < W, vb = l < 2 * HX && xb < W;
  const int xac = min(max(xa, 0), W - 1), xbc = min(xb, W - 1);
  // branch-free: every lane loads an in-bounds element (row and column clamped); zero is
  // selected outside the plane when the row is consumed, so no load sits under an exec
  // mask and none is waited for before its use
  float qa[PF], qb[PF];  // row y is held in slot y % PF
  auto fetch = [&](int y, int k) {
    const i64 rb = pbase + (i64)min(y, H - 1) * W;
    qa[k] = ld(src, rb + xac);
    qb[k] = ld(src, rb + xbc);
  };
  // the distinct weights in vector registers for the whole walk: in scalar registers the
  // unrolled walk ran out of them and reloaded the weights row after row, and every such
  // scalar load shares its wait counter with the LDS reads
  float v1[R1 + 1], v2[R2 + 1];
#pragma unroll
  for (int i = 0; i <= R1; ++i) {
    v1[i] = w1[R1 + i];
    asm volatile("" : "+v"(v1[i]));
  }
#pragma unroll
  for (int i = 0; i <= R2; ++i) {
    v2[i] = w2[R2 + i];
    asm volatile("" : "+v"(v2[i]));
  }
  // row y waits in queue slot y % PF (y0 is a multiple of WB, so of PF): every slot index
  // below is a compile-time constant
  if (y0 == 0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) fetch(k, k);
  } else {
#pragma unroll
    for (int k = 0; k < PF; ++k) fetch(y0 - R + k, ((k - R) % PF + PF) % PF);
  }
  // x outputs of row y, whose slot is k; fetches row y + PF into the freed slot.  Past
  // the plane the line holds zeros, so the outputs are +0 as the y pass's zero padding
  // requires.
  auto xrow = [&](int y, int k) -> float2 {
    const float a = (va && y < H) ? qa[k] : 0.0f, b = (vb && y < H) ? qb[k] : 0.0f;
    fetch(y + PF, k);
    float* ln = line[y & 1];
    ln[l] = a;
    ln[64 + l] = b;
    wave_lds_sync();
    const float* t = ln + l + HX;
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int j = -R1; j <= R1; ++j) s1 = fmaf(v1[j < 0 ? -j : j], t[j], s1);
#pragma unroll
    for (int j = -R2; j <= R2; ++j) s2 = fmaf(v2[j < 0 ? -j : j], t[j], s2);
    // pin the row's outputs here: otherwise the compiler sinks the FMA chains towards
    // their uses and keeps many rows' LDS windows live at once
    asm volatile("" : "+v"(s1), "+v"(s2));
    return make_float2(s1, s2);
  };
  const bool out = x0 + l < W;
  float2* dcol = dst + pbase + x0 + l;
  float2 ring[WB];  // row p lives in ring[(p + R) % WB]; rows < 0 are zero
  if (y0 == 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) ring[q] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int q = R; q < WB; ++q) ring[q] = xrow(q - R, (q - R) % PF);
  } else {  // a later segment starts from the R rows above it
#pragma unroll
    for (int q = 0; q < WB; ++q) ring[q] = xrow(y0 - R + q, ((q - R) % PF + PF) % PF);
  }
  // The segment's end leaves the kernel (a wave-uniform return) instead of skipping the
  // remaining unrolled rows: no path that skips a row's loads merges back into the loop,
  // so the compiler's memory-wait counts keep the PF rows in flight.
  for (int c0 = y0;; c0 += WB) {
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      const int c = c0 + q;
      if (c >= y1) return;
      float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
      for (int j = -R1; j <= R1; ++j) s1 = fmaf(v1[j < 0 ? -j : j], ring[(q + j + R) % WB].x, s1);
#pragma unroll
      for (int j = -R2; j <= R2; ++j) s2 = fmaf(v2[j < 0 ? -j : j], ring[(q + j + R) % WB].y, s2);
      if (out) st_nt(dcol + (i64)c * W, make_float2(s1, s2));
      ring[q] = xrow(c + R + 1, (q + R + 1) % PF);  // row c + R + 1 replaces row c - R
    }
  }
}

// (r1, r2) pairs with compile-time kernels: truncation 4 and 2 of the benchmark sigmas
// (2.0 / 3.2, 1.6 / 2.56) and of the reference's example sigmas (4.0 / 6.0, 1.8 / 2.7).
#define HDOG_GAUSS_PAIRS(X) X(8, 13) X(6, 10) X(4, 6) X(3, 5) X(8, 12) X(4, 5) X(7, 11)

template <typename TS>
hipError_t gauss_x(const TS* src, float2* dst, const float* w1, const float* w2, int r1, int r2,
                   int W, i64 nrows, hipStream_t s) {
#define X(A, B)                                                                              \
  if (r1 == A && r2 == B) {                                                                  \
    constexpr int C = NT / 4 * 4;                                                            \
    const int nct = (W + C - 1) / C;                                                         \
    const i64 nblk = ((nrows + 3) / 4) * nct;                                                \
    hipLaunchKernelGGL((k_gauss_x_r<A, B, TS>), dim3((unsigned)nblk), dim3(NT), 0, s, src,   \
                       dst, w1, w2, W, nrows, nct);                                          \
    return hipGetLastError();                                                                \
  }
  HDOG_GAUSS_PAIRS(X)
#undef X
  constexpr int ROWS = 4;
  const int C = 256;
  const int nct = (W + C - 1) / C;
  const i64 nblk = ((nrows + ROWS - 1) / ROWS) * nct;
  const size_t lds = (size_t)ROWS * (C + 2 * std::max(r1, r2)) * 4;
  hipLaunchKernelGGL((k_gauss_x<ROWS, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src, dst, w1,
                     w2, r1, r2, W, nrows, C, nct);
  return hipGetLastError();
}

// Fused x + y passes for the compile-time radius pairs; false when (r1, r2) has none.
template <typename TS>
bool gauss_xy(const TS* src, float2* dst, const float* w1, const float* w2, int r1, int r2,
              int W, int H, int D, hipStream_t s, hipError_t* e, int segments) {
#define X(A, B)                                                                              \
  if (r1 == A && r2 == B) {                                                                  \
    const int nxs = (W + 63) / 64;                                                           \
    int nys, rows;                                                                           \
    walk_segments(H, 2 * (A > B ? A : B) + 1, &nys, &rows, segments);                        \
    const i64 nblk = (i64)nxs * D * nys;                                                     \
    hipLaunchKernelGGL((k_gauss_xy_walk<A, B, TS>), dim3((unsigned)nblk), dim3(64), 0, s,    \
                       src, dst, w1, w2, W, H, nxs, nys, rows);                              \
    *e = hipGetLastError();                                                                  \
    return true;                                                                             \
  }
  HDOG_GAUSS_PAIRS(X)
#undef X
  return false;
}

template <bool LAST, typename TS>
hipError_t gauss_line(const float2* src, float2* dst, TS* dog, const float* w1, const float* w2,
                      int r1, int r2, float nz, int W, int L, i64 sa, i64 so, i64 n_other,
                      hipStream_t s) {
#define X(A, B)                                                                              \
  if (r1 == A && r2 == B) {                                                                  \
    const int nxb = (W + NT - 1) / NT;                                                       \
    const i64 nblk = (i64)nxb * n_other;                                                     \
    hipLaunchKernelGGL((k_gauss_walk<A, B, LAST, TS>), dim3((unsigned)nblk), dim3(NT), 0, s, \
                       src, dst, dog, w1, w2, nz, W, L, sa, so, nxb);                        \
    return hipGetLastError();                                                                \
  }
  HDOG_GAUSS_PAIRS(X)
#undef X
  const int R = std::max(r1, r2);
  int XW = R > 24 ? 32 : 64;
  int C = 64;
  if (L < C) C = L;
  const int nxt = (W + XW - 1) / XW, nct = (L + C - 1) / C;
  const i64 nblk = (i64)nxt * nct * n_other;
  const size_t lds = (size_t)(C + 2 * R) * XW * 8;
  if (XW == 64)
    hipLaunchKernelGGL((k_gauss_line<64, LAST, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, dog, w1, w2, r1, r2, nz, W, L, sa, so, C, nxt, nct);
  else
    hipLaunchKernelGGL((k_gauss_line<32, LAST, TS>), dim3((unsigned)nblk), dim3(NT), lds, s, src,
                       dst, dog, w1, w2, r1, r2, nz, W, L, sa, so, C, nxt, nct);
  return hipGetLastError();
}

}  // namespace

template <typename TN, typename TD>
hipError_t launch_dog(const TN* norm, TD* dog, void* tmp1, void* tmp2, const float* weights,
                      const int r[4], float normalizer, int W, int H, int D, hipStream_t s,
                      int segments) {
  const i64 WH = (i64)W * H;
  const float* wxy1 = weights;
  const float* wz1 = weights + kMaxTaps;
  const float* wxy2 = weights + 2 * kMaxTaps;
  const float* wz2 = weights + 3 * kMaxTaps;
  hipError_t e = hipSuccess;
  const float2* gxy = (const float2*)tmp2;
  if (gauss_xy<TN>(norm, (float2*)tmp1, wxy1, wxy2, r[0], r[2], W, H, D, s, &e, segments)) {
    if (e != hipSuccess) return e;
    mark("dog/gauss_xy", s);
    gxy = (const float2*)tmp1;
  } else {
    if ((e = gauss_x<TN>(norm, (float2*)tmp1, wxy1, wxy2, r[0], r[2], W, (i64)H * D, s)) !=
        hipSuccess)
      return e;
    mark("dog/gauss_x", s);
    if ((e = gauss_line<false, TD>((const float2*)tmp1, (float2*)tmp2, nullptr, wxy1, wxy2, r[0],
                                   r[2], 0.0f, W, H, W, WH, D, s)) != hipSuccess)
      return e;
    mark("dog/gauss_y", s);
  }
  e = gauss_line<true, TD>(gxy, nullptr, dog, wz1, wz2, r[1], r[3], normalizer, W, D, WH, W, H,
                           s);
  mark("dog/gauss_z_dog", s);
  return e;
}

// Separable pair of Gaussians with independent weights per axis (w[0..2] = x, y, z of the
// first, w[3..5] of the second; radii r[6]).  With dog != nullptr the z pass writes
// (G1 - G2) * nz; otherwise it writes the (G1, G2) pair to pair_out.  Used by the
// differentiable form (anisotropic sigmas and their weight derivatives).
hipError_t launch_gauss_pair(const float* src, void* tmp1, void* tmp2, const float* const w[6],
                             const int r[6], float nz, float* dog, float2* pair_out, int W, int H,
                             int D, hipStream_t s) {
  const i64 WH = (i64)W * H;
  hipError_t e;
  if ((e = gauss_x<float>(src, (float2*)tmp1, w[0], w[3], r[0], r[3], W, (i64)H * D, s)) !=
      hipSuccess)
    return e;
  if ((e = gauss_line<false, float>((const float2*)tmp1, (float2*)tmp2, nullptr, w[1], w[4], r[1],
                                    r[4], 0.0f, W, H, W, WH, D, s)) != hipSuccess)
    return e;
  if (dog)
    return gauss_line<true, float>((const float2*)tmp2, nullptr, dog, w[2], w[5], r[2], r[5], nz, W,
                                   D, WH, W, H, s);
  return gauss_line<false, float>((const float2*)tmp2, pair_out, nullptr, w[2], w[5], r[2], r[5],
                                  0.0f, W, D, WH, W, H, s);
}

#define HDOG_INST_DOG(TN, TD)                                                                \
  template hipError_t launch_dog<TN, TD>(const TN*, TD*, void*, void*, const float*, const int*, \
                                         float, int, int, int, hipStream_t, int);
HDOG_INST_DOG(float, float)
HDOG_INST_DOG(__half, __half)
HDOG_INST_DOG(__hip_bfloat16, __hip_bfloat16)
HDOG_INST_DOG(__half, float)
HDOG_INST_DOG(__hip_bfloat16, float)
#undef HDOG_INST_DOG

}  // namespace hdog
