// ccl_model.cpp -- host-threaded model of the candidate labeling kernels of
// cubic-informatics_amd/csrc/ccl.hip, for checking their concurrency on the CPU.
//
// TEST INFRASTRUCTURE.  It restates, memory operation by memory operation, the three
// kernels that label the candidate voxels of each Hessian tile (the foreground of
// src/3D/CCL3D.cu:17-117: labels are component minima over the wrapped 27-offset
// adjacency, voxel 0 is background):
//   A  k_cand_local   the tile's mask rows; a union-find over the tile's candidates in a
//                     tile-private array of 16-bit parents held in 32-bit words (links
//                     by compare-and-swap of the word, path-halving 16-bit stores); each
//                     record's tile-local root label; P written at tile-local roots and at
//                     the faces that other tiles' forward neighbours reach
//   B  k_cand_merge   the tile's rim candidates; their forward neighbours outside the tile
//                     (wrapping like the reference) with the link-skipping rule; the
//                     neighbours' labels read from P before any union; unions on P by
//                     atomic minimum with path halving (the generic form for width or
//                     height 1: every candidate, every distinct positive offset)
//   C  k_cand_flatten every record's root by a read-only walk, written back into P
//                     (dense labels) and into the record's label; root bits; the bits of
//                     roots whose region leaves the region pass's work-group
// A work-group runs on a team of OS threads (its "lanes", records dealt round-robin)
// with a barrier wherever the kernel has __syncthreads; several teams run work-groups at
// once; every memory operation on shared data may be preceded by a random yield or spin,
// so that runs explore different interleavings.  Results are compared with the oracle's
// sequential union-find (oracle/hdog_oracle.cpp, oracle_ccl3d).
//
// Access modes for P in kernels B and C:
//   atomic  every load and store is a relaxed atomic (what the kernels do)
//   plain   loads and stores are plain accesses beside the atomic minimum (what the
//           kernels did before); built with -fsanitize=thread this reports the data race
//   stale   plain accesses through a private write-back cache per lane, the way a cache
//           that is not kept coherent with the others would serve them: a load returns
//           the lane's cached copy (filled on first use, dropped at random), a store
//           stays in the lane's cache until a random write-back or the end of the
//           kernel (all lanes' leftovers written back in random order); the minimum is
//           performed on memory after the lane's own dirty copy is written back
//
// usage: ccl_model MODE ITERS SEED TEAMS LANES   (MODE atomic | plain | stale)
// prints "ok <volumes> <candidates> <unions>" or the first mismatch; exit status 0 / 1.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#ifndef PLANT
#define PLANT 0  // test builds plant one mistake (1..5) to show that the model can fail
#endif

extern "C" void oracle_ccl3d(const uint8_t* pd, uint32_t* labels, int W, int H, int D);

namespace {

typedef int64_t i64;
bool g_plain = false, g_stale = false;

// ---- random interleavings ------------------------------------------------------------
thread_local uint64_t t_rng = 1;
inline uint64_t rnd() {
  t_rng ^= t_rng << 13;
  t_rng ^= t_rng >> 7;
  t_rng ^= t_rng << 17;
  return t_rng;
}
inline void chaos() {
  const uint64_t r = rnd() & 63;
  if (r < 6) std::this_thread::yield();
  else if (r < 8) for (volatile int k = 0, n = (int)(rnd() & 255); k < n; ++k) {}
}

// ---- P: the volume-sized parent array -------------------------------------------------
uint32_t* gP = nullptr;
i64 gN = 0;
[[noreturn]] void fault(const char* what, i64 x) {
  printf("FAULT %s %lld (N %lld)\n", what, (long long)x, (long long)gN);
  fflush(stdout);
  std::_Exit(1);
}
inline void in_range(uint32_t x) {
  if ((i64)x >= gN) fault("P access out of range at", x);
}
// a lane's private cache of P (mode stale): state 0 absent, 1 clean, 2 dirty
struct TeamCache {
  std::vector<uint32_t> val;
  std::vector<uint8_t> state;
  std::vector<uint32_t> dirty;  // addresses that may be dirty
  void reset(size_t n) {
    val.assign(n, 0u);
    state.assign(n, 0);
    dirty.clear();
  }
  void write_back(uint32_t x) {
    if (state[x] == 2) {
      __atomic_store_n(gP + x, val[x], __ATOMIC_RELAXED);
      state[x] = 1;
    }
  }
  void write_back_all() {
    for (uint32_t x : dirty) write_back(x);
    dirty.clear();
  }
};
thread_local TeamCache* t_cache = nullptr;
inline void cache_noise() {  // random evictions and write-backs of this lane's lines
  TeamCache& c = *t_cache;
  const uint64_t r = rnd() & 31;
  if (r == 0 && !c.dirty.empty()) {
    const size_t k = rnd() % c.dirty.size();
    c.write_back(c.dirty[k]);
    c.dirty[k] = c.dirty.back();
    c.dirty.pop_back();
  } else if (r == 1) {
    const uint32_t x = (uint32_t)(rnd() % c.state.size());
    if (c.state[x] == 1) c.state[x] = 0;
  }
}
inline uint32_t ldP(uint32_t x) {
  in_range(x);
  chaos();
  if (g_stale) {
    TeamCache& c = *t_cache;
    cache_noise();
    if (c.state[x] == 0) {
      c.val[x] = __atomic_load_n(gP + x, __ATOMIC_RELAXED);
      c.state[x] = 1;
    }
    return c.val[x];
  }
  if (g_plain) return ((volatile uint32_t*)gP)[x];
  return __atomic_load_n(gP + x, __ATOMIC_RELAXED);
}
inline void stP(uint32_t x, uint32_t v) {
  in_range(x);
  chaos();
  if (g_stale) {
    TeamCache& c = *t_cache;
    cache_noise();
    c.val[x] = v;
    if (c.state[x] != 2) c.dirty.push_back(x);
    c.state[x] = 2;
    return;
  }
  if (g_plain) ((volatile uint32_t*)gP)[x] = v;
  else __atomic_store_n(gP + x, v, __ATOMIC_RELAXED);
}
inline uint32_t minP(uint32_t x, uint32_t v) {  // atomicMin, returns the old value
  in_range(x);
  chaos();
  if (g_stale) {  // the lane's own dirty copy reaches memory first
    t_cache->write_back(x);
    t_cache->state[x] = 0;
  }
  uint32_t cur = __atomic_load_n(gP + x, __ATOMIC_RELAXED);
  while (cur > v && !__atomic_compare_exchange_n(gP + x, &cur, v, true, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED)) {
  }
  return cur;
}
// ccl.hip find_root / unite
// a walk longer than N steps has met a cycle (parents must only decrease)
uint32_t find_root(uint32_t x) {
  uint32_t p = ldP(x);
  for (i64 steps = 0; p != x; ++steps) {
    if (steps > gN || p > x) fault("parent chain not decreasing at", x);
    const uint32_t gp = ldP(p);
    if (gp != p) stP(x, gp);
    x = p;
    p = gp;
  }
  return x;
}
std::atomic<uint64_t> g_unions{0};
uint32_t unite(uint32_t a, uint32_t b) {
  g_unions.fetch_add(1, std::memory_order_relaxed);
  while (true) {
    a = find_root(a);
    b = find_root(b);
    if (a == b) return a;
    if (a < b) std::swap(a, b);
    const uint32_t old = minP(a, b);
#if PLANT == 5  // planted mistake: a lost race on the minimum is not retried
    return b;
#endif
    if (old == a) return b;
    a = old;
  }
}
uint32_t find_root_readonly(uint32_t x) {
#if PLANT == 3  // planted mistake: the flatten walk halves paths
  return find_root(x);
#endif
  uint32_t p = ldP(x);
  for (i64 steps = 0; p != x; ++steps) {
    if (steps > gN || p > x) fault("parent chain not decreasing at", x);
    x = p;
    p = ldP(x);
  }
  return x;
}

// ---- the forward offsets and the link-skipping rule (ccl.hip fdx / fwd_adj / fwd_order)
constexpr int fdx(int k) { return k < 9 ? k % 3 - 1 : k < 12 ? k - 10 : 1; }
constexpr int fdy(int k) { return k < 9 ? k / 3 - 1 : k < 12 ? 1 : 0; }
constexpr int fdz(int k) { return k < 9 ? 1 : 0; }
constexpr uint32_t fwd_adj(int q) {
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
uint32_t links_of(uint32_t fg, uint32_t known) {
  uint32_t link = 0;
  for (int j = 0; j < 13; ++j) {
    const int q = fwd_order(j);
    const uint32_t b = (fg >> q) & 1u;
    link |= ((fwd_adj(q) & known) == 0u ? b : 0u) << q;
    known |= b << q;
  }
  return link;
}

// ---- geometry -------------------------------------------------------------------------
struct Geo {
  int W, H, D, TX, TY, TZ, nbx, nby, nbz;
  uint32_t WH;
  i64 N, ntiles;
  int rows() const { return TY * TZ; }
};
struct Tile {
  int x0, y0, z0, ex, ey, ez;
};
Tile tile_of(const Geo& g, i64 t) {
  Tile T;
  T.x0 = (int)(t % g.nbx) * g.TX;
  T.y0 = (int)((t / g.nbx) % g.nby) * g.TY;
  T.z0 = (int)(t / ((i64)g.nbx * g.nby)) * g.TZ;
  T.ex = std::min(g.TX, g.W - T.x0);
  T.ey = std::min(g.TY, g.H - T.y0);
  T.ez = std::min(g.TZ, g.D - T.z0);
  return T;
}
inline void pos_xyz(const Geo& g, uint32_t v, int& lx, int& ly, int& lz) {
  lx = (int)(v % g.TX);
  ly = (int)((v / g.TX) % g.TY);
  lz = (int)(v / (g.TX * g.TY));
}
inline uint32_t global_of(const Geo& g, const Tile& T, uint32_t v) {
  int lx, ly, lz;
  pos_xyz(g, v, lx, ly, lz);
  return (uint32_t)(T.z0 + lz) * g.WH + (uint32_t)(T.y0 + ly) * (uint32_t)g.W +
         (uint32_t)(T.x0 + lx);
}
inline size_t mask_word(const Geo& g, int x, int y, int z) {
  const size_t tile = ((size_t)(z / g.TZ) * g.nby + y / g.TY) * g.nbx + x / g.TX;
  return tile * g.rows() + (size_t)(z % g.TZ) * g.TY + y % g.TY;
}

// ---- one team = one work-group --------------------------------------------------------
struct Barrier {
  std::atomic<int> count{0}, gen{0};
  int n = 1;
  void wait() {
    const int g = gen.load();
    if (count.fetch_add(1) + 1 == n) {
      count.store(0);
      gen.fetch_add(1);
    } else {
      while (gen.load() == g) std::this_thread::yield();
    }
  }
};
struct Team {
  Barrier bar;
  std::atomic<i64> tile{0};
  std::vector<std::atomic<uint32_t>> lab;  // 16-bit parents, two per word
  std::vector<std::atomic<uint64_t>> bits;
  std::vector<uint32_t> rim_v, rim_me;
  std::atomic<uint32_t> n_rim{0};
};

// LDS union-find on 16-bit parents (ccl.hip lds16_*): a 16-bit load or store is one
// indivisible operation on its half of the word
inline uint32_t lab_ld(Team& tm, uint32_t x) {
  chaos();
  return (tm.lab[x >> 1].load(std::memory_order_relaxed) >> ((x & 1u) * 16)) & 0xFFFFu;
}
inline void lab_st(Team& tm, uint32_t x, uint32_t v) {
  chaos();
  const int sh = (int)(x & 1u) * 16;
  uint32_t cur = tm.lab[x >> 1].load(std::memory_order_relaxed);
  while (!tm.lab[x >> 1].compare_exchange_weak(cur, (cur & ~(0xFFFFu << sh)) | (v << sh),
                                               std::memory_order_relaxed)) {
  }
}
uint32_t lds16_find(Team& tm, uint32_t x) {
  uint32_t p = lab_ld(tm, x);
  while (p != x) {
    x = p;
    p = lab_ld(tm, x);
  }
  return x;
}
uint32_t lds16_find_halve(Team& tm, uint32_t x) {
  uint32_t p = lab_ld(tm, x);
  while (p != x) {
    const uint32_t gp = lab_ld(tm, p);
    if (gp != p) lab_st(tm, x, gp);
    x = p;
    p = gp;
  }
  return x;
}
uint32_t lds16_min(Team& tm, uint32_t a, uint32_t b) {
  std::atomic<uint32_t>& w = tm.lab[a >> 1];
  const int sh = (int)(a & 1u) * 16;
  chaos();
  uint32_t cur = w.load(std::memory_order_relaxed);
  while (true) {
    const uint32_t old = (cur >> sh) & 0xFFFFu;
    if (old <= b) return old;
    chaos();
    if (w.compare_exchange_strong(cur, (cur & ~(0xFFFFu << sh)) | (b << sh),
                                  std::memory_order_relaxed))
      return old;
  }
}
uint32_t lds16_unite(Team& tm, uint32_t a, uint32_t b) {
  while (true) {
    a = lds16_find_halve(tm, a);
    b = lds16_find_halve(tm, b);
    if (a == b) return a;
    if (a < b) std::swap(a, b);
    const uint32_t old = lds16_min(tm, a, b);
    if (old == a) return b;
    a = old;
  }
}

// ---- the volume's data ----------------------------------------------------------------
struct Vol {
  Geo g;
  std::vector<uint8_t> pd;
  std::vector<std::vector<uint32_t>> rec;  // candidate positions per tile (any order)
  std::vector<std::vector<uint32_t>> rlab;
  std::vector<uint64_t> fg;                // tile-major mask rows
  std::vector<uint32_t> P;
  std::vector<std::atomic<uint64_t>> roots, multi;
  bool dense = true;
  int accgy = 2, accgz = 2;
};

using Kernel = void (*)(Vol&, Team&, i64 tile, int lane, int lanes);

void run_kernel(Vol& V, std::vector<Team>& teams, int lanes, Kernel k, uint64_t seed) {
  std::atomic<i64> next{0};
  std::vector<std::thread> th;
  std::vector<TeamCache> caches(teams.size() * lanes);
  for (size_t t = 0; t < teams.size(); ++t) {
    Team& tm = teams[t];
    tm.bar.n = lanes;
    for (int l = 0; l < lanes; ++l)
      th.emplace_back([&, t, l] {
        t_rng = seed * 0x9E3779B97F4A7C15ull + t * 131 + l * 7 + 1;
        TeamCache& c = caches[t * lanes + l];
        if (g_stale) c.reset((size_t)V.g.N);
        t_cache = &c;
        while (true) {
          if (l == 0) tm.tile.store(next.fetch_add(1));
          tm.bar.wait();
          const i64 tile = tm.tile.load();
          if (tile >= V.g.ntiles) break;
          k(V, tm, tile, l, lanes);
          tm.bar.wait();
        }
      });
  }
  for (auto& x : th) x.join();
  if (g_stale) {  // the end of the kernel: every lane's dirty lines, in random order
    std::vector<std::pair<size_t, uint32_t>> left;
    for (size_t c = 0; c < caches.size(); ++c)
      for (uint32_t x : caches[c].dirty) left.emplace_back(c, x);
    for (size_t k2 = left.size(); k2 > 1; --k2) std::swap(left[k2 - 1], left[rnd() % k2]);
    for (auto& e : left) caches[e.first].write_back(e.second);
  }
}

// kernel A: k_cand_local
void kernel_local(Vol& V, Team& tm, i64 tile, int lane, int lanes) {
  const Geo& g = V.g;
  const std::vector<uint32_t>& rec = V.rec[tile];
  const uint32_t n = (uint32_t)rec.size();
  for (int r = lane; r < g.rows(); r += lanes) tm.bits[r].store(0);
  tm.bar.wait();
  for (uint32_t i = lane; i < n; i += lanes) {
    lab_st(tm, rec[i], rec[i]);
    chaos();
    tm.bits[rec[i] / g.TX].fetch_or(1ull << (rec[i] % g.TX));
  }
  tm.bar.wait();
  for (int r = lane; r < g.rows(); r += lanes) V.fg[(size_t)tile * g.rows() + r] = tm.bits[r].load();
  const Tile T = tile_of(g, tile);
  const bool links = g.W >= 2 && g.H >= 2;
  if (links && n)
    for (uint32_t i = lane; i < n; i += lanes) {
      const uint32_t v = rec[i];
      int vx, vy, vz;
      pos_xyz(g, v, vx, vy, vz);
      uint32_t fg = 0;
      for (int q = 0; q < 13; ++q) {
        const int nx = vx + fdx(q), ny = vy + fdy(q), nz = vz + fdz(q);
        if (nx < 0 || nx >= T.ex || ny < 0 || ny >= T.ey || nz >= T.ez) continue;
        fg |= (uint32_t)((tm.bits[ny + g.TY * nz].load() >> nx) & 1u) << q;
      }
      uint32_t link = links_of(fg, 0u);
      uint32_t root = v;
      while (link) {
        const int q = __builtin_ctz(link);
        link &= link - 1u;
        const uint32_t u = v + (uint32_t)(fdz(q) * g.TX * g.TY + fdy(q) * g.TX + fdx(q));
        if (lab_ld(tm, u) != root) root = lds16_unite(tm, root, u);
      }
    }
  tm.bar.wait();
  for (uint32_t i = lane; i < n; i += lanes) {
    const uint32_t v = rec[i];
    const uint32_t r = lds16_find(tm, v);
    const uint32_t label = global_of(g, T, r);
    int lx, ly, lz;
    pos_xyz(g, v, lx, ly, lz);
#if PLANT == 1  // planted mistake: the high-y face is not written
    if (r == v || lz == 0 || lx == 0 || lx == T.ex - 1 || ly == 0)
#else
    if (r == v || lz == 0 || lx == 0 || lx == T.ex - 1 || ly == 0 || ly == T.ey - 1)
#endif
      stP(global_of(g, T, v), label);
    V.rlab[tile][i] = label;
  }
}

bool fg_at(const Vol& V, int x, int y, int z) {
  return (V.fg[mask_word(V.g, x, y, z)] >> (x % V.g.TX)) & 1u;
}

// kernel B: k_cand_merge
void kernel_merge(Vol& V, Team& tm, i64 tile, int lane, int lanes) {
  const Geo& g = V.g;
  const std::vector<uint32_t>& rec = V.rec[tile];
  const uint32_t n = (uint32_t)rec.size();
  if (n == 0) return;
  const Tile T = tile_of(g, tile);
  if (g.W < 2 || g.H < 2) {  // generic: every candidate, the distinct positive offsets
    std::vector<i64> offs;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const i64 o = dz * (i64)g.WH + dy * (i64)g.W + dx;
          if (o > 0) offs.push_back(o);
        }
    std::sort(offs.begin(), offs.end());
    offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
    for (uint32_t i = lane; i < n; i += lanes) {
      const uint32_t idx = global_of(g, T, rec[i]);
      for (i64 o : offs) {
        const i64 u = (i64)idx + o;
        if (u >= g.N) break;
        const int uz = (int)(u / g.WH), uy = (int)((u % g.WH) / g.W), ux = (int)(u % g.W);
        if (fg_at(V, ux, uy, uz)) unite(idx, (uint32_t)u);
      }
    }
    return;
  }
  if (lane == 0) tm.n_rim.store(0);
  tm.bar.wait();
  for (uint32_t i = lane; i < n; i += lanes) {
    int lx, ly, lz;
    pos_xyz(g, rec[i], lx, ly, lz);
    if (lz == T.ez - 1 || lx == 0 || lx == T.ex - 1 || ly == 0 || ly == T.ey - 1) {
      const uint32_t s = tm.n_rim.fetch_add(1);
      tm.rim_v[s] = rec[i];
      tm.rim_me[s] = V.rlab[tile][i];
    }
  }
  tm.bar.wait();
  const uint32_t nr = tm.n_rim.load();
  const uint64_t* tb = V.fg.data() + (size_t)tile * g.rows();
  for (uint32_t j = lane; j < nr; j += lanes) {
    const uint32_t v = tm.rim_v[j], me = tm.rim_me[j];
    int lx, ly, lz;
    pos_xyz(g, v, lx, ly, lz);
    const int x = T.x0 + lx, y = T.y0 + ly, z = T.z0 + lz;
    uint32_t in_fg = 0, out_fg = 0, uidx[13];
    for (int q = 0; q < 13; ++q) {
      const int dx = fdx(q), dy = fdy(q), dz = fdz(q);
      const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
      const bool inside = nx >= 0 && nx < T.ex && ny >= 0 && ny < T.ey && nz < T.ez;
      if (inside) in_fg |= (uint32_t)((tb[ny + g.TY * nz] >> nx) & 1u) << q;
      int ux = x + dx, uy = y + dy, uz = z + dz;
      if (dx < 0 && ux < 0) { ux += g.W; uy -= 1; }
      if (dx > 0 && ux >= g.W) { ux -= g.W; uy += 1; }
      if ((dy < 0 || dx < 0) && uy < 0) { uy += g.H; uz -= 1; }
#if PLANT == 4  // planted mistake: no carry into z from a wrapped row
      if (dy > 0 && uy >= g.H) { uy -= g.H; uz += 1; }
#else
      if ((dy > 0 || dx > 0) && uy >= g.H) { uy -= g.H; uz += 1; }
#endif
      const bool exists = !inside && uz < g.D;
      out_fg |= (uint32_t)(exists && fg_at(V, ux, uy, uz)) << q;
      uidx[q] = exists ? (uint32_t)uz * g.WH + (uint32_t)uy * (uint32_t)g.W + (uint32_t)ux : 0u;
    }
#if PLANT == 2  // planted mistake: outside neighbours count as known
    uint32_t link = links_of(out_fg, in_fg | out_fg);
#else
    uint32_t link = links_of(out_fg, in_fg);
#endif
    uint32_t lj[13];
    for (int q = 0; q < 13; ++q) {  // loaded together, before any union
      const bool on = (link >> q) & 1u;
      const uint32_t l = ldP(on ? uidx[q] : 0u);
      lj[q] = on ? l : 0u;
    }
    uint32_t seen1 = 0, seen2 = 0;
    while (link) {
      const int q = __builtin_ctz(link);
      link &= link - 1u;
      const uint32_t l = lj[q];
      if (!l || l == me || l == seen1 || l == seen2) continue;
      unite(me, l);
      seen2 = seen1;
      seen1 = l;
    }
  }
}

// kernel C: k_cand_flatten
void kernel_flatten(Vol& V, Team&, i64 tile, int lane, int lanes) {
  const Geo& g = V.g;
  const std::vector<uint32_t>& rec = V.rec[tile];
  const uint32_t n = (uint32_t)rec.size();
  const Tile T = tile_of(g, tile);
  for (uint32_t i = lane; i < n; i += lanes) {
    const uint32_t idx = global_of(g, T, rec[i]), l = V.rlab[tile][i];
    const uint32_t r = find_root_readonly(l);
    V.rlab[tile][i] = r;
    if (V.dense && ldP(idx) != r) stP(idx, r);
    if (r == idx) {
      V.roots[r >> 6].fetch_or(1ull << (r & 63));
    } else if (l == idx) {
      const uint32_t rz = r / g.WH, rr = r % g.WH, ry = rr / g.W, rx = rr % g.W;
      const int ty = T.y0 / g.TY, tz = T.z0 / g.TZ;
      if ((int)rx / g.TX != T.x0 / g.TX || ((int)ry / g.TY) / V.accgy != ty / V.accgy ||
          ((int)rz / g.TZ) / V.accgz != tz / V.accgz)
        V.multi[r >> 6].fetch_or(1ull << (r & 63));
    }
  }
}

// ---- one volume -----------------------------------------------------------------------
struct Stats {
  uint64_t cand = 0;
};

bool check_volume(int W, int H, int D, int TX, int TY, int TZ, double dens, bool dense,
                  uint64_t seed, std::vector<Team>& teams, int lanes, Stats& st) {
  Vol V;
  Geo& g = V.g;
  g.W = W; g.H = H; g.D = D; g.TX = TX; g.TY = TY; g.TZ = TZ;
  g.nbx = (W + TX - 1) / TX; g.nby = (H + TY - 1) / TY; g.nbz = (D + TZ - 1) / TZ;
  g.WH = (uint32_t)W * (uint32_t)H;
  g.N = (i64)W * H * D;
  g.ntiles = (i64)g.nbx * g.nby * g.nbz;
  V.dense = dense;
  t_rng = seed * 2654435761u + 12345;
  V.pd.resize(g.N);
  // blobs of foreground on a sparse background, so that components span tiles
  for (i64 i = 0; i < g.N; ++i) V.pd[i] = (rnd() % 1000) < (uint64_t)(dens * 1000);
  const int nblob = 1 + (int)(rnd() % 6);
  for (int b = 0; b < nblob; ++b) {
    const int cx = (int)(rnd() % W), cy = (int)(rnd() % H), cz = (int)(rnd() % D);
    const int r = 1 + (int)(rnd() % 4);
    for (int z = std::max(0, cz - r); z <= std::min(D - 1, cz + r); ++z)
      for (int y = std::max(0, cy - r); y <= std::min(H - 1, cy + r); ++y)
        for (int x = std::max(0, cx - r); x <= std::min(W - 1, cx + r); ++x)
          if ((rnd() & 7) != 0) V.pd[((i64)z * H + y) * W + x] = 1;
  }
  V.rec.assign(g.ntiles, {});
  for (int z = 0; z < D; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const i64 i = ((i64)z * H + y) * W + x;
        if (i == 0 || !V.pd[i]) continue;  // voxel 0 is background whatever its mask
        const i64 t = ((i64)(z / TZ) * g.nby + y / TY) * g.nbx + x / TX;
        V.rec[t].push_back((uint32_t)(x % TX + TX * (y % TY + TY * (z % TZ))));
      }
  for (auto& r : V.rec) {
    for (size_t k = r.size(); k > 1; --k) std::swap(r[k - 1], r[rnd() % k]);  // any order
    st.cand += r.size();
  }
  V.rlab.resize(g.ntiles);
  for (i64 t = 0; t < g.ntiles; ++t) V.rlab[t].assign(V.rec[t].size(), 0u);
  V.fg.assign((size_t)g.ntiles * g.rows(), 0ull);
  V.P.resize(g.N);
  // P is cleared only for dense labels; else it holds whatever the last run left
  for (i64 i = 0; i < g.N; ++i) V.P[i] = dense ? 0u : (uint32_t)rnd();
  gP = V.P.data();
  gN = g.N;
  const size_t words = (size_t)(g.N + 63) / 64;
  V.roots = std::vector<std::atomic<uint64_t>>(words);
  V.multi = std::vector<std::atomic<uint64_t>>(words);
  const int tv = TX * TY * TZ;
  const int rim_max = tv;
  for (Team& tm : teams) {
    tm.lab = std::vector<std::atomic<uint32_t>>((tv + 1) / 2);
    tm.bits = std::vector<std::atomic<uint64_t>>(g.rows());
    tm.rim_v.assign(rim_max, 0);
    tm.rim_me.assign(rim_max, 0);
  }
  run_kernel(V, teams, lanes, kernel_local, seed * 3 + 1);
  run_kernel(V, teams, lanes, kernel_merge, seed * 3 + 2);
  run_kernel(V, teams, lanes, kernel_flatten, seed * 3 + 3);

  std::vector<uint32_t> want(g.N);
  oracle_ccl3d(V.pd.data(), want.data(), W, H, D);
  // every record's label, and with dense labels the whole volume
  for (i64 t = 0; t < g.ntiles; ++t) {
    const Tile T = tile_of(g, t);
    for (size_t i = 0; i < V.rec[t].size(); ++i) {
      const uint32_t idx = global_of(g, T, V.rec[t][i]);
      if (V.rlab[t][i] != want[idx]) {
        printf("MISMATCH record label at %u: model %u oracle %u (W %d H %d D %d tile %dx%dx%d "
               "seed %llu dense %d)\n", idx, V.rlab[t][i], want[idx], W, H, D, TX, TY, TZ,
               (unsigned long long)seed, (int)dense);
        return false;
      }
    }
  }
  if (dense)
    for (i64 i = 0; i < g.N; ++i)
      if (V.P[i] != want[i]) {
        printf("MISMATCH dense label at %lld: model %u oracle %u (W %d H %d D %d tile %dx%dx%d "
               "seed %llu)\n", (long long)i, V.P[i], want[i], W, H, D, TX, TY, TZ,
               (unsigned long long)seed);
        return false;
      }
  // root bits exactly at component minima; region-spanning bits exactly where a component
  // has voxels in more than one region work-group cell
  std::vector<int64_t> cell(g.N, -1);
  for (i64 i = 1; i < g.N; ++i) {
    if (!want[i]) continue;
    const bool isroot = want[i] == (uint32_t)i;
    if (((V.roots[i >> 6].load() >> (i & 63)) & 1u) != (uint64_t)isroot) {
      printf("MISMATCH root bit at %lld (seed %llu)\n", (long long)i, (unsigned long long)seed);
      return false;
    }
  }
  std::vector<uint8_t> spans(g.N, 0);
  for (i64 i = 1; i < g.N; ++i) {
    if (!want[i]) continue;
    const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / g.WH);
    const int64_t c = ((int64_t)(z / TZ / V.accgz) * 1000003 + (y / TY / V.accgy)) * 1000003 +
                      x / TX;
    const uint32_t r = want[i];
    if (cell[r] < 0) cell[r] = c;
    else if (cell[r] != c) spans[r] = 1;
  }
  for (i64 i = 1; i < g.N; ++i)
    if (want[i] == (uint32_t)i &&
        ((V.multi[i >> 6].load() >> (i & 63)) & 1u) != (uint64_t)spans[i]) {
      printf("MISMATCH region-spanning bit at %lld (seed %llu)\n", (long long)i,
             (unsigned long long)seed);
      return false;
    }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: ccl_model atomic|plain|stale ITERS SEED TEAMS LANES\n");
    return 2;
  }
  g_plain = strcmp(argv[1], "plain") == 0;
  g_stale = strcmp(argv[1], "stale") == 0;
  const int iters = atoi(argv[2]);
  const uint64_t seed0 = strtoull(argv[3], nullptr, 10);
  const int nteams = atoi(argv[4]), lanes = atoi(argv[5]);
  std::vector<Team> teams(nteams);
  Stats st;
  // tile shapes: the kernels' 64 x 8 x 16 and small ones that put many tiles, faces and
  // wrapped links into tiny volumes
  const int shapes[][3] = {{64, 8, 16}, {4, 2, 2}, {8, 4, 2}, {2, 2, 4}, {16, 2, 2}, {3, 3, 3}};
  for (int it = 0; it < iters; ++it) {
    const uint64_t seed = seed0 * 100003 + it;
    t_rng = seed * 0x2545F4914F6CDD1Dull + 7;
    const int* sh = shapes[rnd() % 6];
    int W, H, D;
    if (sh[0] == 64) { W = 1 + (int)(rnd() % 150); H = 1 + (int)(rnd() % 20); D = 1 + (int)(rnd() % 36); }
    else { W = 1 + (int)(rnd() % 14); H = 1 + (int)(rnd() % 9); D = 1 + (int)(rnd() % 9); }
    if (rnd() % 8 == 0) W = 1 + (int)(rnd() % 2);  // width or height 1 or 2 (generic merge, wrap)
    if (rnd() % 8 == 0) H = 1 + (int)(rnd() % 2);
    const double dens = 0.05 + 0.6 * (double)(rnd() % 1000) / 1000.0;
    const bool dense = (rnd() & 1) != 0;
    if (!check_volume(W, H, D, sh[0], sh[1], sh[2], dens, dense, seed, teams, lanes, st)) return 1;
  }
  printf("ok %d %llu %llu\n", iters, (unsigned long long)st.cand,
         (unsigned long long)g_unions.load());
  return 0;
}
