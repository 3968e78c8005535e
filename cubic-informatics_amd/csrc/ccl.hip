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

__device__ __forceinline__ uint32_t ld_agent(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// parents only ever decrease, so every chain ends at a root even under concurrent links
__device__ __forceinline__ uint32_t find_root(uint32_t* L, uint32_t x) {
  uint32_t p = ld_agent(L + x);
  while (p != x) {
    x = p;
    p = ld_agent(L + x);
  }
  return x;
}

__device__ void unite(uint32_t* L, uint32_t a, uint32_t b) {
  while (true) {
    a = find_root(L, a);
    b = find_root(L, b);
    if (a == b) return;
    if (a < b) { const uint32_t t = a; a = b; b = t; }
    const uint32_t old = atomicMin(L + a, b);
    if (old == a) return;
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
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      if (k < offs.n) {
        const i64 j = i + offs.o[k];
        if (j < N && pd[j]) unite(L, (uint32_t)i, (uint32_t)j);
      }
    }
  }
}

__global__ __launch_bounds__(NT) void k_ccl_flatten(uint32_t* L, i64 N) {
  for (i64 i = (i64)blockIdx.x * NT + threadIdx.x; i < N; i += (i64)gridDim.x * NT) {
    const uint32_t l = L[i];
    if (l && l != (uint32_t)i) L[i] = find_root(L, l);
  }
}

}  // namespace

hipError_t launch_ccl(const uint8_t* pd, uint32_t* labels, int W, int H, int D, hipStream_t s) {
  const i64 WH = (i64)W * H, N = WH * D;
  Offsets offs{};
  {
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
  }
  const i64 nb = std::min<i64>((N + NT - 1) / NT, 1 << 20);
  hipLaunchKernelGGL(k_ccl_init, dim3((unsigned)nb), dim3(NT), 0, s, pd, labels, N);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark("ccl/init", s);
  hipLaunchKernelGGL(k_ccl_union, dim3((unsigned)nb), dim3(NT), 0, s, pd, labels, N, offs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark("ccl/union", s);
  hipLaunchKernelGGL(k_ccl_flatten, dim3((unsigned)nb), dim3(NT), 0, s, labels, N);
  mark("ccl/flatten", s);
  return hipGetLastError();
}

}  // namespace hdog
