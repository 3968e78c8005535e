// capi.hip -- the C-ABI of include/hdog.h: context, argument checks, scratch management
// and stage sequencing.  No process-global state: everything lives in hdog_ctx.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hdog_internal.hpp"
#include "hdog.h"

using namespace hdog;

namespace {
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

// One branch of a batch graph: the scratch of one stack at a time, sized for the largest
// stack of the branch.  The separable passes' pair buffers (tmp1, tmp2: 16 bytes per voxel)
// live inside the candidate-record block (28 bytes per voxel of tile-padded address space):
// they are dead before the Hessian pass writes records, and the next stack of the branch
// writes them only after the previous stack's region pass has read its records.
struct BatchLane {
  Buf norm, dog, labels, rid, bcnt, rmask, rmulti, dR, acc, cand, cand_cnt, fgbits, cand_lab;
  void* tmp1 = nullptr;
  void* tmp2 = nullptr;
  i64 maxN = 0, maxT = 0, maxF = 0, acc_cap = 0;
  hipStream_t s = nullptr;  // capture stream of the branch
  std::vector<int> stacks;  // in launch order
};

// A captured batch of resident stacks: it owns every buffer its graph points at.  Stacks
// are dealt longest first to the branch with fewer voxels so far; the branches are
// independent chains of the graph, so one stack's labeling and region pass (latency-bound)
// run beside another stack's filters (bandwidth- and issue-bound).
struct BatchState {
  std::vector<hdog_tile> tiles;
  hdog_params p{};
  std::vector<BatchLane> lanes;
  Buf weights, tout, running;
  Buf a_label, a_size, a_vox, a_cen, a_eig, a_nrm, a_hess, a_isum, a_imax;
  i64 acc_cap = 0, arena_cap = 0;  // acc_cap: the largest of the branches'
  hipStream_t cs = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  TileOut* h_tout = nullptr;  // pinned copy of the per-stack placements
  bool ran = false;
};
}  // namespace

struct hdog_ctx {
  int device = 0;
  std::string err;
  Buf tmp1, tmp2, rid, bcnt, rmask, rmulti, dR, acc, weights, blobs, sw, sums, ddog, cand,
      cand_cnt, fgbits, cand_lab;
  Buf o_label, o_size, o_vox, o_cen, o_eig, o_nrm, o_hess, o_isum, o_imax;
  i64 out_cap = 0;
  Buf img, norm, dog, hess, pd, labels;  // substack buffers of hdog_process_substack
  bool gauss_ready = false;
  float sigma[4] = {0, 0, 0, 0};
  double truncate = 0.0;
  int r[4] = {0, 0, 0, 0};
  i64 R = 0;
  uint32_t* h_R = nullptr;  // pinned
  bool prof = false;        // record an event after every kernel of run_device
  Prof pool;
  bool ev_valid = false;
  BatchState* batch = nullptr;
  // split substack form: two device image slots filled on a copy stream
  Buf up_img[2];
  hipStream_t copy_stream = nullptr;
  hipEvent_t up_done[2] = {nullptr, nullptr};   // upload into the slot complete
  hipEvent_t run_done[2] = {nullptr, nullptr};  // last run reading the slot complete
  int up_shape[2][3] = {{0, 0, 0}, {0, 0, 0}};  // (W, H, D) of each slot's upload
  int up_valid[2] = {0, 0};                      // slices before the zero fill
  // launch shapes chosen by hdog_autotune per stack shape and storage form
  struct Tuned {
    int W, H, D, precision, norm_segments, gauss_segments;
  };
  std::vector<Tuned> tuned;
};

thread_local hdog::Prof* hdog::tl_prof = nullptr;

namespace {

int fail(hdog_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}

#define HIPCHK(...)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (__VA_ARGS__);                                                       \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, HDOG_ERR_HIP, "%s failed: %s", #__VA_ARGS__, hipGetErrorString(e_)); \
  } while (0)

hipError_t ensure(Buf& b, size_t bytes) {
  if (bytes <= b.bytes) return hipSuccess;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return e;
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e == hipSuccess) b.bytes = bytes;
  return e;
}

void release(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// ---- the no-allocation labeling (hdog_ccl3d_noalloc) ----------------------------------
// One fixed scratch array per GPU, sized for volumes of at most kNoAllocVoxels voxels once
// padded to whole Hessian tiles (2^23: a 128 x 256 x 256 stack exactly).  Its layout per
// call, every part 256-byte aligned: candidate positions and record labels (4 bytes per
// padded voxel each), per-tile counts, the candidate bit mask (one word per tile row), and
// the root and region-spanning bitmaps.  The array's name is only a host-side stand-in;
// kernels get the address hipGetSymbolAddress returns on each device (noalloc_addr).
constexpr i64 kNoAllocVoxels = (i64)1 << 23;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
struct NoAllocLayout {
  size_t cand, lab, cnt, fg, roots, multi, total;
  i64 ntiles, root_words;
};
constexpr NoAllocLayout noalloc_layout(i64 ntiles, i64 N) {
  NoAllocLayout L{};
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  L.ntiles = ntiles;
  L.root_words = nb * (kRootBlock / 64);
  L.cand = 0;
  L.lab = L.cand + align256((size_t)ntiles * kCandTile * 4);
  L.cnt = L.lab + align256((size_t)ntiles * kCandTile * 4);
  L.fg = L.cnt + align256((size_t)ntiles * 4);
  L.roots = L.fg + align256((size_t)ntiles * kHessTY * kHessTZ * 8);
  L.multi = L.roots + align256((size_t)L.root_words * 8);
  L.total = L.multi + align256((size_t)L.root_words * 8);
  return L;
}
constexpr size_t kNoAllocScratchBytes =
    noalloc_layout(kNoAllocVoxels / kCandTile, kNoAllocVoxels).total;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
void* g_noalloc_addr[kMaxDevices] = {};
std::mutex g_noalloc_mu;

// the device address of `scratch` on `device` (the current device), looked up once
hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_addr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_noalloc_addr[device] = ptr;
  return hipSuccess;
}
void* noalloc_addr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return g_noalloc_addr[device];
}

// the tuned launch shape of a stack shape, or nullptr (built-in rule)
const hdog_ctx::Tuned* find_tuned(const hdog_ctx* ctx, int W, int H, int D, int precision) {
  for (const hdog_ctx::Tuned& t : ctx->tuned)
    if (t.W == W && t.H == H && t.D == D && t.precision == precision) return &t;
  return nullptr;
}
int tuned_norm(const hdog_ctx* ctx, int W, int H, int D, int precision) {
  const hdog_ctx::Tuned* t = find_tuned(ctx, W, H, D, precision);
  return t ? t->norm_segments : 0;
}
int tuned_gauss(const hdog_ctx* ctx, int W, int H, int D, int precision) {
  const hdog_ctx::Tuned* t = find_tuned(ctx, W, H, D, precision);
  return t ? t->gauss_segments : 0;
}

// the FP64 re-evaluation band of the reduced-precision forms (0 selects the default)
float minor_band(const hdog_params* p) {
  return p->minor_band > 0.0f ? p->minor_band : kDefaultMinorBand;
}

int check_shape(hdog_ctx* ctx, int W, int H, int D) {
  if (W < 1 || H < 1 || D < 1)
    return fail(ctx, HDOG_ERR_ARG, "invalid shape %dx%dx%d (width x height x depth)", W, H, D);
  const i64 N = (i64)W * H * D;
  if (N >= (i64)0xFFFFFFFFll)
    return fail(ctx, HDOG_ERR_ARG, "volume of %lld voxels exceeds the uint32 label range",
                (long long)N);
  return HDOG_OK;
}

// calcGaussianWeight, src/3D/DoG3DFilter.cu:372-392, truncation as a parameter
int gaussian_weights(float sigma, double truncate, float* w) {
  const int lw = (int)(truncate * (double)sigma + 0.5);
  if (lw > kMaxGaussRadius) return -1;
  const float s2 = sigma * sigma;
  float sum = 1.0f;
  w[lw] = 1.0f;
  for (int i = 1; i <= lw; ++i) {
    const float p = (float)std::exp(-0.5 * i * i / (double)s2);
    w[lw - i] = p;
    w[lw + i] = p;
    sum += p * 2;
  }
  for (int i = 0; i < 2 * lw + 1; ++i) w[i] /= sum;
  return lw;
}

int ensure_regions(hdog_ctx* ctx, i64 R) {
  HIPCHK(ensure(ctx->acc, (size_t)R * sizeof(RegionAcc)));
  if (R > ctx->out_cap) {
    const i64 cap = std::max<i64>(R, ctx->out_cap * 3 / 2);
    HIPCHK(ensure(ctx->o_label, cap * 4));
    HIPCHK(ensure(ctx->o_size, cap * 2));
    HIPCHK(ensure(ctx->o_vox, cap * 8));
    HIPCHK(ensure(ctx->o_cen, cap * 3 * 4));
    HIPCHK(ensure(ctx->o_eig, cap * 2 * 4));
    HIPCHK(ensure(ctx->o_nrm, cap * 4));
    HIPCHK(ensure(ctx->o_hess, cap * 6 * 4));
    HIPCHK(ensure(ctx->o_isum, cap * 8));
    HIPCHK(ensure(ctx->o_imax, cap * 2));
    ctx->out_cap = cap;
  }
  return HDOG_OK;
}

// Candidate records of the Hessian pass, when the pipeline produced them.
struct CandRecords {
  const float* rec = nullptr;
  const uint32_t* cnt = nullptr;
  const uint32_t* rlab = nullptr;  // each record's label (launch_ccl_candidates)
  const unsigned long long* multi = nullptr;  // roots of multi-component regions (same)
  i64 ntiles = 0;
};

// With cand.rec the components come from launch_ccl_candidates (their root bits already
// in ctx->rmask, each record's label in cand.rlab) and the region sums from the records
// (labels and hess unused); else from the label volume (final labels, only read) and the
// materialized Hessian.
template <typename TH, typename TN>
int regional_features_t(hdog_ctx* ctx, hipStream_t s, uint32_t* labels, const TH* hess,
                        const TN* norm, const uint16_t* img, int W, int H, int D, int is_ave,
                        int64_t* num_regions, CandRecords cand = CandRecords()) {
  const i64 N = (i64)W * H * D;
  HIPCHK(ensure(ctx->rid, (size_t)N * 4));
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  HIPCHK(ensure(ctx->bcnt, (size_t)block_count_words(N) * 4));
  HIPCHK(ensure(ctx->rmask, (size_t)nb * (kRootBlock / 8)));
  HIPCHK(ensure(ctx->dR, 64));
  i64 nbo = 0;
  unsigned long long* masks = (unsigned long long*)ctx->rmask.p;
  if (cand.rec)
    HIPCHK(launch_region_count_masks(masks, (uint32_t*)ctx->bcnt.p, (uint32_t*)ctx->dR.p, N, s,
                                     &nbo));
  else
    HIPCHK(launch_region_ids(labels, (uint32_t*)ctx->bcnt.p, masks, (uint32_t*)ctx->dR.p, N, s,
                             &nbo));
  HIPCHK(hipMemcpyAsync(ctx->h_R, ctx->dR.p, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const i64 R = *ctx->h_R;
  int rc = ensure_regions(ctx, R);
  if (rc) return rc;
  RegionAcc* acc = (RegionAcc*)ctx->acc.p;
  HIPCHK(launch_region_assign(masks, (const uint32_t*)ctx->bcnt.p, (uint32_t*)ctx->rid.p, acc, R, N,
                              s));
  if (cand.rec)
    HIPCHK(launch_region_accum_cand<TN>(cand.rec, cand.cnt, cand.ntiles, cand.rlab, cand.multi,
                                        (const uint32_t*)ctx->rid.p, norm, img, acc, R, W, H,
                                        is_ave, s));
  else
    HIPCHK(launch_region_accum<TH, TN>(labels, (const uint32_t*)ctx->rid.p, hess, norm, img, acc, R,
                                       W, H, D, is_ave, s));
  const i64 cap = ctx->out_cap;
  RegionOut o;
  o.label = (uint32_t*)ctx->o_label.p;
  o.size = (uint16_t*)ctx->o_size.p;
  o.voxels = (u64*)ctx->o_vox.p;
  o.cx = (float*)ctx->o_cen.p;
  o.cy = o.cx + cap;
  o.cz = o.cx + 2 * cap;
  o.st = (float*)ctx->o_eig.p;
  o.bl = o.st + cap;
  o.nrm = (float*)ctx->o_nrm.p;
  o.hess = (float*)ctx->o_hess.p;
  o.isum = (u64*)ctx->o_isum.p;
  o.imax = (uint16_t*)ctx->o_imax.p;
  HIPCHK(launch_region_finalize(acc, R, is_ave, o, cap, s));
  ctx->R = R;
  if (num_regions) *num_regions = R;
  return HDOG_OK;
}

// The separable passes' pair buffers (8 bytes per voxel each) of the whole-path call and
// the tuner: they live in the candidate-record block (28 bytes per voxel of tile-padded
// address space), being dead before the Hessian pass writes the first record.
hipError_t pair_buffers(hdog_ctx* ctx, int W, int H, int D, void** tmp1, void** tmp2) {
  const i64 N = (i64)W * H * D;
  const size_t off2 = ((size_t)N * 8 + 255) / 256 * 256;
  const hipError_t e = ensure(ctx->cand, std::max((size_t)hessian_tiles(W, H, D) * kCandTile * 7 * 4,
                                                  off2 + (size_t)N * 8));
  *tmp1 = ctx->cand.p;
  *tmp2 = (char*)ctx->cand.p + off2;
  return e;
}

// pd and hess may be NULL (not materialized); labels is the union-find's parent array and
// becomes the full label volume only with `dense`.
template <typename TN, typename TD, typename TH>
int run_stages(hdog_ctx* ctx, hipStream_t s, const hdog_params* p, const uint16_t* img, TN* norm,
               TD* dog, TH* hess, uint8_t* pd, uint32_t* labels, bool dense,
               int64_t* num_regions) {
  const int W = p->width, H = p->height, D = p->depth;
  const i64 N = (i64)W * H * D;
  CandRecords cand;
  cand.ntiles = hessian_tiles(W, H, D);
  void *tmp1, *tmp2;
  HIPCHK(pair_buffers(ctx, W, H, D, &tmp1, &tmp2));
  struct ProfScope {  // points the launchers' marks at this context's pool for one call
    explicit ProfScope(Prof* p) { tl_prof = p; }
    ~ProfScope() { tl_prof = nullptr; }
  } scope(ctx->prof ? &ctx->pool : nullptr);
  ctx->ev_valid = false;
  ctx->pool.n = 0;
  mark(nullptr, s);
  HIPCHK(launch_normalize<TN>(img, norm, tmp1, tmp2, p->min_intensity, W, H, D, p->radius_xy,
                              p->radius_z, s, nullptr, nullptr,
                              tuned_norm(ctx, W, H, D, p->precision)));
  const float nz = powf(ctx->sigma[0], p->gamma * 2) * powf(ctx->sigma[1], p->gamma);
  HIPCHK(launch_dog<TN, TD>(norm, dog, tmp1, tmp2, (const float*)ctx->weights.p, ctx->r, nz, W,
                            H, D, s, tuned_gauss(ctx, W, H, D, p->precision)));
  // the Hessian pass also leaves compact candidate records and the candidate bit mask;
  // labeling and the region pass work from those, not from the mask and element volumes
  // (written only when pd / hess are non-NULL)
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  HIPCHK(ensure(ctx->cand_cnt, (size_t)cand.ntiles * 4));
  HIPCHK(ensure(ctx->cand_lab, (size_t)cand.ntiles * kCandTile * 4));
  HIPCHK(ensure(ctx->fgbits, (size_t)fg_words(W, H, D) * 8));
  HIPCHK(ensure(ctx->rmask, (size_t)nb * (kRootBlock / 8)));
  HIPCHK(ensure(ctx->rmulti, (size_t)nb * (kRootBlock / 8)));
  cand.rec = (const float*)ctx->cand.p;
  cand.cnt = (const uint32_t*)ctx->cand_cnt.p;
  cand.rlab = (const uint32_t*)ctx->cand_lab.p;
  cand.multi = (const unsigned long long*)ctx->rmulti.p;
  HIPCHK(launch_hessian<TD, TH>(dog, hess, pd, W, H, D, s, (float*)ctx->cand.p,
                                (uint32_t*)ctx->cand_cnt.p, minor_band(p)));
  HIPCHK(launch_ccl_candidates(cand.rec, cand.cnt, cand.ntiles,
                               (unsigned long long*)ctx->fgbits.p, labels,
                               (uint32_t*)ctx->cand_lab.p, (unsigned long long*)ctx->rmask.p,
                               (unsigned long long*)ctx->rmulti.p, nb * (kRootBlock / 64), dense,
                               W, H, D, s));
  int rc = regional_features_t<TH, TN>(ctx, s, labels, hess, norm, img, W, H, D, p->is_ave_mode,
                                       num_regions, cand);
  if (rc) return rc;
  ctx->ev_valid = ctx->prof;
  return HDOG_OK;
}

int check_params(hdog_ctx* ctx, const hdog_params* p) {
  if (!p) return fail(ctx, HDOG_ERR_ARG, "params is NULL");
  int rc = check_shape(ctx, p->width, p->height, p->depth);
  if (rc) return rc;
  if (p->radius_xy < 0 || p->radius_z < 0 || p->radius_xy > kMaxNormRadius ||
      p->radius_z > kMaxNormRadius)
    return fail(ctx, HDOG_ERR_ARG, "normalization radii (%d, %d) outside [0, %d]", p->radius_xy,
                p->radius_z, kMaxNormRadius);
  if (p->precision < HDOG_FP32 || p->precision > HDOG_BF16_DOG32)
    return fail(ctx, HDOG_ERR_ARG, "unknown precision %d", p->precision);
  if (!(p->minor_band >= 0.0f && p->minor_band < 0.5f))
    return fail(ctx, HDOG_ERR_ARG, "minor_band %g outside [0, 0.5)", (double)p->minor_band);
  return HDOG_OK;
}

// Normalised Gaussian weights exp(-i^2 / (2 sigma^2)) / Z on |i| <= r (FP64) and their
// sigma-derivative (w_i / sigma^3) (i^2 - sum_k w_k k^2), stored as float.
void soft_weights(double sigma, int r, float* w, float* dw) {
  std::vector<double> e(2 * r + 1);
  double Z = 0.0;
  for (int i = -r; i <= r; ++i) Z += (e[i + r] = std::exp(-(double)i * i / (2.0 * sigma * sigma)));
  double m2 = 0.0;
  for (int i = -r; i <= r; ++i) m2 += e[i + r] / Z * (double)i * i;
  for (int i = -r; i <= r; ++i) {
    const double wi = e[i + r] / Z;
    w[i + r] = (float)wi;
    if (dw) dw[i + r] = (float)(wi / (sigma * sigma * sigma) * ((double)i * i - m2));
  }
}

struct SoftSetup {
  int r[6];
  const float* w[6];   // x, y, z of sigma1, then of sigma2 (device)
  const float* dw[6];  // their sigma-derivatives (device)
  double s;            // normalizer (sx1 sy1 sz1)^gamma
};

int soft_setup(hdog_ctx* ctx, const double sig1[3], const double sig2[3], const int rad1[3],
               const int rad2[3], double gamma, SoftSetup* S) {
  std::vector<float> hw(12 * kMaxTaps, 0.0f);
  for (int k = 0; k < 6; ++k) {
    const double sg = k < 3 ? sig1[k] : sig2[k - 3];
    const int r = k < 3 ? rad1[k] : rad2[k - 3];
    if (!(sg > 0.0) || r < 0 || r > kMaxGaussRadius)
      return fail(ctx, HDOG_ERR_ARG, "sigma %g / radius %d invalid (radius at most %d)", sg, r,
                  kMaxGaussRadius);
    S->r[k] = r;
    soft_weights(sg, r, hw.data() + k * kMaxTaps, hw.data() + (6 + k) * kMaxTaps);
  }
  HIPCHK(ensure(ctx->sw, hw.size() * 4));
  HIPCHK(hipMemcpy(ctx->sw.p, hw.data(), hw.size() * 4, hipMemcpyHostToDevice));
  for (int k = 0; k < 6; ++k) {
    S->w[k] = (const float*)ctx->sw.p + k * kMaxTaps;
    S->dw[k] = (const float*)ctx->sw.p + (6 + k) * kMaxTaps;
  }
  S->s = std::pow(sig1[0] * sig1[1] * sig1[2], gamma);
  return HDOG_OK;
}


// ---- batched stacks -------------------------------------------------------------

void batch_free(BatchState* B) {
  if (!B) return;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  if (B->fork) (void)hipEventDestroy(B->fork);
  for (hipEvent_t e : B->joins)
    if (e) (void)hipEventDestroy(e);  // (a null handle would leave an error for the next call)
  for (BatchLane& L : B->lanes) {
    if (L.s && L.s != B->cs) (void)hipStreamDestroy(L.s);
    for (Buf* b : {&L.norm, &L.dog, &L.labels, &L.rid, &L.bcnt, &L.rmask, &L.rmulti, &L.dR,
                   &L.acc, &L.cand, &L.cand_cnt, &L.fgbits, &L.cand_lab})
      release(*b);
  }
  if (B->cs) (void)hipStreamDestroy(B->cs);
  if (B->h_tout) (void)hipHostFree(B->h_tout);
  for (Buf* b : {&B->weights, &B->tout, &B->running, &B->a_label, &B->a_size, &B->a_vox,
                 &B->a_cen, &B->a_eig, &B->a_nrm, &B->a_hess, &B->a_isum, &B->a_imax})
    release(*b);
  delete B;
}

// Enqueues the whole pipeline of tile t on stream s with device-side region counts: no
// host synchronization, allocation or memset, so it can be captured.
template <typename TN, typename TD, typename TH>
hipError_t batch_tile(hdog_ctx* ctx, BatchState* B, BatchLane& L, int t, hipStream_t s) {
  const hdog_tile& T = B->tiles[t];
  const int W = T.width, H = T.height, D = T.depth;
  const i64 N = (i64)W * H * D;
  const hdog_params* p = &B->p;
  TN* norm = (TN*)L.norm.p;
  TD* dog = (TD*)L.dog.p;
  TH* hess = nullptr;  // not materialized: the region pass reads the candidate records
  uint32_t* labels = (uint32_t*)L.labels.p;  // parent array, candidate voxels only
  hipError_t e;
  if ((e = launch_normalize<TN>(T.image, norm, L.tmp1, L.tmp2, p->min_intensity, W, H, D,
                                p->radius_xy, p->radius_z, s, nullptr, nullptr,
                                tuned_norm(ctx, W, H, D, p->precision))) != hipSuccess)
    return e;
  const float nz = powf(ctx->sigma[0], p->gamma * 2) * powf(ctx->sigma[1], p->gamma);
  if ((e = launch_dog<TN, TD>(norm, dog, L.tmp1, L.tmp2, (const float*)B->weights.p, ctx->r, nz, W,
                              H, D, s, tuned_gauss(ctx, W, H, D, p->precision))) != hipSuccess)
    return e;
  const i64 ntiles = hessian_tiles(W, H, D);
  if ((e = launch_hessian<TD, TH>(dog, hess, nullptr, W, H, D, s, (float*)L.cand.p,
                                  (uint32_t*)L.cand_cnt.p, minor_band(&B->p))) != hipSuccess)
    return e;
  i64 nbo = 0;
  uint32_t* dR = (uint32_t*)L.dR.p;
  unsigned long long* masks = (unsigned long long*)L.rmask.p;
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  if ((e = launch_ccl_candidates((const float*)L.cand.p, (const uint32_t*)L.cand_cnt.p, ntiles,
                                 (unsigned long long*)L.fgbits.p, labels,
                                 (uint32_t*)L.cand_lab.p, masks,
                                 (unsigned long long*)L.rmulti.p, nb * (kRootBlock / 64), false,
                                 W, H, D, s)) != hipSuccess)
    return e;
  if ((e = launch_region_count_masks(masks, (uint32_t*)L.bcnt.p, dR, N, s, &nbo)) != hipSuccess)
    return e;
  TileOut* tout = (TileOut*)B->tout.p + t;
  if ((e = launch_tile_reserve(dR, L.acc_cap, B->arena_cap, (u64*)B->running.p, tout, s)) !=
      hipSuccess)
    return e;
  RegionAcc* acc = (RegionAcc*)L.acc.p;
  if ((e = launch_region_assign(masks, (const uint32_t*)L.bcnt.p, (uint32_t*)L.rid.p, acc,
                                L.acc_cap, N, s)) != hipSuccess)
    return e;
  if ((e = launch_region_accum_cand<TN>((const float*)L.cand.p, (const uint32_t*)L.cand_cnt.p,
                                        ntiles, (const uint32_t*)L.cand_lab.p,
                                        (const unsigned long long*)L.rmulti.p,
                                        (const uint32_t*)L.rid.p, norm, T.image,
                                        acc, L.acc_cap, W, H, p->is_ave_mode, s)) != hipSuccess)
    return e;
  const i64 cap = B->arena_cap;
  RegionOut o;
  o.label = (uint32_t*)B->a_label.p;
  o.size = (uint16_t*)B->a_size.p;
  o.voxels = (u64*)B->a_vox.p;
  o.cx = (float*)B->a_cen.p;
  o.cy = o.cx + cap;
  o.cz = o.cx + 2 * cap;
  o.st = (float*)B->a_eig.p;
  o.bl = o.st + cap;
  o.nrm = (float*)B->a_nrm.p;
  o.hess = (float*)B->a_hess.p;
  o.isum = (u64*)B->a_isum.p;
  o.imax = (uint16_t*)B->a_imax.p;
  return launch_region_finalize(acc, L.acc_cap, p->is_ave_mode, o, cap, s, tout);
}

// Captures on B->cs: the arena counter is cleared, the other branches fork from it through
// an event, each branch enqueues its stacks in order, and B->cs joins them at the end.
template <typename TN, typename TD, typename TH>
hipError_t batch_capture(hdog_ctx* ctx, BatchState* B) {
  hipError_t e = launch_fill_u64((u64*)B->running.p, 1, 0, B->cs);
  if (e == hipSuccess) e = hipEventRecord(B->fork, B->cs);
  for (size_t l = 1; e == hipSuccess && l < B->lanes.size(); ++l)
    e = hipStreamWaitEvent(B->lanes[l].s, B->fork, 0);
  for (size_t l = 0; l < B->lanes.size(); ++l)
    for (int t : B->lanes[l].stacks)
      if (e == hipSuccess) e = batch_tile<TN, TD, TH>(ctx, B, B->lanes[l], t, B->lanes[l].s);
  for (size_t l = 1; l < B->lanes.size(); ++l) {
    if (e == hipSuccess) e = hipEventRecord(B->joins[l], B->lanes[l].s);
    if (e == hipSuccess) e = hipStreamWaitEvent(B->cs, B->joins[l], 0);
  }
  return e;
}

hipError_t batch_capture_any(hdog_ctx* ctx, BatchState* B) {
  typedef __half h16;
  typedef __hip_bfloat16 b16;
  switch (B->p.precision) {
    case HDOG_FP16: return batch_capture<h16, h16, h16>(ctx, B);
    case HDOG_BF16: return batch_capture<b16, b16, b16>(ctx, B);
    case HDOG_FP16_DOG32: return batch_capture<h16, float, h16>(ctx, B);
    case HDOG_BF16_DOG32: return batch_capture<b16, float, b16>(ctx, B);
    default: return batch_capture<float, float, float>(ctx, B);
  }
}

// ---- autotuner of the run-time launch shapes ---------------------------------------
// Candidate row-segment counts of the two row walkers, 1..4; counts that give the same
// segmentation as a smaller one are skipped.
constexpr int kMaxSegments = 4;

// median of three event-timed runs of f (after one untimed run)
template <typename F>
hipError_t time_median(hipStream_t s, F&& f, float* ms) {
  hipEvent_t a, b;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  if ((e = hipEventCreate(&b)) != hipSuccess) { (void)hipEventDestroy(a); return e; }
  float t[3] = {0, 0, 0};
  e = f();
  for (int k = 0; e == hipSuccess && k < 3; ++k) {
    if ((e = hipEventRecord(a, s)) != hipSuccess) break;
    if ((e = f()) != hipSuccess) break;
    if ((e = hipEventRecord(b, s)) != hipSuccess) break;
    if ((e = hipEventSynchronize(b)) != hipSuccess) break;
    e = hipEventElapsedTime(&t[k], a, b);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::sort(t, t + 3);
  *ms = t[1];
  return e;
}

// Measured search: the normalization stage with every candidate segment count of its row
// walker, then the DoG stage likewise, on the caller's stack; the fastest of each.
template <typename TN, typename TD>
int autotune_t(hdog_ctx* ctx, hipStream_t s, const hdog_params* p, const uint16_t* img,
               int* best_norm, int* best_gauss, float* ms_norm, float* ms_gauss) {
  const int W = p->width, H = p->height, D = p->depth;
  const i64 N = (i64)W * H * D;
  void *tmp1, *tmp2;
  HIPCHK(pair_buffers(ctx, W, H, D, &tmp1, &tmp2));
  HIPCHK(ensure(ctx->norm, (size_t)N * sizeof(TN)));
  HIPCHK(ensure(ctx->dog, (size_t)N * sizeof(TD)));
  TN* norm = (TN*)ctx->norm.p;
  TD* dog = (TD*)ctx->dog.p;
  const float nz = powf(ctx->sigma[0], p->gamma * 2) * powf(ctx->sigma[1], p->gamma);
  const int wb_gauss = 2 * std::max(ctx->r[0], ctx->r[2]) + 1;
  struct Stage {
    int wb, *best;
    float* ms;
    bool dog;
  } stages[2] = {{21, best_norm, ms_norm, false}, {wb_gauss, best_gauss, ms_gauss, true}};
  for (const Stage& st : stages) {
    // the built-in segment count stays unless another one is faster by more than 3 %: a
    // choice within the timing noise would change nothing but the launch shape
    int def_nys, def_rows;
    walk_segments(H, st.wb, &def_nys, &def_rows);
    float def_ms = 1e30f;
    int cand_want[kMaxSegments], ncand = 0;
    float cand_ms[kMaxSegments];
    int seen[kMaxSegments + 1] = {0};
    int nseen = 0;
    for (int want = 1; want <= kMaxSegments; ++want) {
      int nys, rows;
      walk_segments(H, st.wb, &nys, &rows, want);
      bool dup = false;
      for (int k = 0; k < nseen; ++k) dup |= seen[k] == nys;
      if (dup) continue;
      seen[nseen++] = nys;
      float ms = 0.0f;
      if (st.dog)
        HIPCHK(time_median(s, [&] {
          return launch_dog<TN, TD>(norm, dog, tmp1, tmp2,
                                    (const float*)ctx->weights.p, ctx->r, nz, W, H, D, s, want);
        }, &ms));
      else
        HIPCHK(time_median(s, [&] {
          return launch_normalize<TN>(img, norm, tmp1, tmp2, p->min_intensity, W, H,
                                      D, p->radius_xy, p->radius_z, s, nullptr, nullptr, want);
        }, &ms));
      cand_want[ncand] = want;
      cand_ms[ncand++] = ms;
      if (nys == def_nys) def_ms = ms;
    }
    *st.best = 0;
    *st.ms = def_ms;
    for (int k = 0; k < ncand; ++k) {
      int nys, rows;
      walk_segments(H, st.wb, &nys, &rows, cand_want[k]);
      if (nys == def_nys) {
        if (*st.best == 0) *st.best = cand_want[k];  // the default's own entry
      } else if (cand_ms[k] < 0.97f * def_ms && cand_ms[k] < *st.ms) {
        *st.ms = cand_ms[k];
        *st.best = cand_want[k];
      }
    }
    // the default is among the candidates (1 or 2 segments); otherwise the fastest
    if (*st.best == 0) {
      *st.best = cand_want[0];
      *st.ms = cand_ms[0];
      for (int k = 1; k < ncand; ++k)
        if (cand_ms[k] < *st.ms) {
          *st.ms = cand_ms[k];
          *st.best = cand_want[k];
        }
    }
    if (!st.dog)  // the DoG stage is timed on the normalized stack
      HIPCHK(launch_normalize<TN>(img, norm, tmp1, tmp2, p->min_intensity, W, H, D,
                                  p->radius_xy, p->radius_z, s, nullptr, nullptr, *st.best));
  }
  return HDOG_OK;
}

}  // namespace

extern "C" {

int hdog_create(hdog_ctx** out, int device) {
  if (!out) return HDOG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HDOG_ERR_HIP;
  if (device < 0 || device >= n) return HDOG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return HDOG_ERR_HIP;
  // the no-allocation labeling's scratch address on this device (set-up, once per device)
  if (noalloc_init(device) != hipSuccess) return HDOG_ERR_HIP;
  hdog_ctx* c = new hdog_ctx();
  c->device = device;
  if (hipHostMalloc((void**)&c->h_R, 64, hipHostMallocDefault) != hipSuccess) {
    delete c;
    return HDOG_ERR_HIP;
  }
  *out = c;
  return HDOG_OK;
}

int hdog_release_scratch(hdog_ctx* ctx) {
  if (!ctx) return HDOG_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  for (Buf* b : {&ctx->tmp1, &ctx->tmp2, &ctx->rid, &ctx->bcnt, &ctx->rmask, &ctx->rmulti, &ctx->dR, &ctx->acc,
                 &ctx->blobs, &ctx->cand_lab, &ctx->fgbits,
                 &ctx->sw, &ctx->sums, &ctx->ddog,
                 &ctx->o_label, &ctx->o_size, &ctx->o_vox, &ctx->o_cen, &ctx->o_eig, &ctx->o_nrm,
                 &ctx->o_hess, &ctx->o_isum, &ctx->o_imax, &ctx->img, &ctx->norm, &ctx->dog,
                 &ctx->hess, &ctx->pd, &ctx->labels, &ctx->cand, &ctx->cand_cnt})
    release(*b);
  ctx->out_cap = 0;
  ctx->R = 0;
  return HDOG_OK;
}

void hdog_destroy(hdog_ctx* ctx) {
  if (!ctx) return;
  hdog_batch_release(ctx);
  (void)hipSetDevice(ctx->device);
  for (int k = 0; k < 2; ++k) {
    if (ctx->up_done[k]) (void)hipEventDestroy(ctx->up_done[k]);
    if (ctx->run_done[k]) (void)hipEventDestroy(ctx->run_done[k]);
    release(ctx->up_img[k]);
  }
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  hdog_release_scratch(ctx);
  if (ctx->prof)
    for (hipEvent_t& e : ctx->pool.ev) (void)hipEventDestroy(e);
  release(ctx->weights);
  if (ctx->h_R) (void)hipHostFree(ctx->h_R);
  delete ctx;
}

const char* hdog_last_error(const hdog_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int hdog_set_profiling(hdog_ctx* ctx, int on) {
  if (!ctx) return HDOG_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  if (on && !ctx->prof)
    for (hipEvent_t& e : ctx->pool.ev) HIPCHK(hipEventCreate(&e));
  if (!on && ctx->prof)
    for (hipEvent_t& e : ctx->pool.ev) HIPCHK(hipEventDestroy(e));
  ctx->prof = on != 0;
  ctx->ev_valid = false;
  return HDOG_OK;
}

int hdog_stage_times(hdog_ctx* ctx, double* ms, const char** names, int max_entries) {
  if (!ctx || !ctx->prof || !ctx->ev_valid || ctx->pool.n < 2) return 0;
  Prof& p = ctx->pool;
  if (hipEventSynchronize(p.ev[p.n - 1]) != hipSuccess) return 0;
  int n = 0;
  for (int k = 1; k < p.n && n < max_entries; ++k) {
    if (!p.name[k]) continue;  // interval that only restarts timing after a host wait
    float t = 0.0f;
    if (hipEventElapsedTime(&t, p.ev[k - 1], p.ev[k]) != hipSuccess) return n;
    if (ms) ms[n] = t;
    if (names) names[n] = p.name[k];
    ++n;
  }
  return n;
}

int hdog_normalize3d(hdog_ctx* ctx, void* stream, const uint16_t* d_img, float* d_norm,
                     float min_intensity, int width, int height, int depth, int radius_xy,
                     int radius_z) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_img || !d_norm) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  if (radius_xy < 0 || radius_z < 0 || radius_xy > kMaxNormRadius || radius_z > kMaxNormRadius)
    return fail(ctx, HDOG_ERR_ARG, "normalization radii (%d, %d) outside [0, %d]", radius_xy,
                radius_z, kMaxNormRadius);
  HIPCHK(hipSetDevice(ctx->device));
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 8));
  HIPCHK(ensure(ctx->tmp2, (size_t)N * 8));
  HIPCHK(launch_normalize<float>(d_img, d_norm, ctx->tmp1.p, ctx->tmp2.p, min_intensity, width,
                                 height, depth, radius_xy, radius_z, (hipStream_t)stream));
  return HDOG_OK;
}

int hdog_normalize3d_ed(hdog_ctx* ctx, void* stream, const uint16_t* d_img, float* d_norm,
                        float* d_erosion, float* d_dilation, float min_intensity, int width,
                        int height, int depth, int radius_xy, int radius_z) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_img || !d_norm || !d_erosion || !d_dilation)
    return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  if (radius_xy < 0 || radius_z < 0 || radius_xy > kMaxNormRadius || radius_z > kMaxNormRadius)
    return fail(ctx, HDOG_ERR_ARG, "normalization radii (%d, %d) outside [0, %d]", radius_xy,
                radius_z, kMaxNormRadius);
  HIPCHK(hipSetDevice(ctx->device));
  const i64 N = (i64)width * height * depth;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 8));
  HIPCHK(ensure(ctx->tmp2, (size_t)N * 8));
  if (radius_xy == 0 || radius_z == 0) {  // thresholded copy: no filters run, E = D = 0
    HIPCHK(hipMemsetAsync(d_erosion, 0, (size_t)N * 4, s));
    HIPCHK(hipMemsetAsync(d_dilation, 0, (size_t)N * 4, s));
  }
  HIPCHK(launch_normalize<float>(d_img, d_norm, ctx->tmp1.p, ctx->tmp2.p, min_intensity, width,
                                 height, depth, radius_xy, radius_z, s, d_erosion, d_dilation));
  return HDOG_OK;
}

int hdog_init_gaussian3d(hdog_ctx* ctx, float sxy1, float sz1, float sxy2, float sz2,
                         double truncate, int* r_z2) {
  if (!ctx) return HDOG_ERR_ARG;
  const float sg[4] = {sxy1, sz1, sxy2, sz2};
  std::vector<float> w(4 * kMaxTaps, 0.0f);
  int r[4];
  for (int k = 0; k < 4; ++k) {
    if (!(sg[k] > 0.0f) || !(truncate >= 0.0))
      return fail(ctx, HDOG_ERR_ARG, "sigma %g / truncate %g must be positive", sg[k], truncate);
    r[k] = gaussian_weights(sg[k], truncate, w.data() + k * kMaxTaps);
    if (r[k] < 0)
      return fail(ctx, HDOG_ERR_ARG, "Gaussian radius of sigma %g at truncate %g exceeds %d", sg[k],
                  truncate, kMaxGaussRadius);
  }
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(ensure(ctx->weights, w.size() * 4));
  HIPCHK(hipMemcpy(ctx->weights.p, w.data(), w.size() * 4, hipMemcpyHostToDevice));
  for (int k = 0; k < 4; ++k) {
    ctx->sigma[k] = sg[k];
    ctx->r[k] = r[k];
  }
  ctx->truncate = truncate;
  ctx->gauss_ready = true;
  if (r_z2) *r_z2 = r[3];
  return HDOG_OK;
}

int hdog_dog3d(hdog_ctx* ctx, void* stream, const float* d_norm, float* d_dog, int width,
               int height, int depth, float gamma) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!ctx->gauss_ready) return fail(ctx, HDOG_ERR_STATE, "hdog_init_gaussian3d not called");
  if (!d_norm || !d_dog) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 8));
  HIPCHK(ensure(ctx->tmp2, (size_t)N * 8));
  const float nz = powf(ctx->sigma[0], gamma * 2) * powf(ctx->sigma[1], gamma);
  HIPCHK(launch_dog<float, float>(d_norm, d_dog, ctx->tmp1.p, ctx->tmp2.p, (const float*)ctx->weights.p,
                           ctx->r, nz, width, height, depth, (hipStream_t)stream));
  return HDOG_OK;
}

int hdog_hessian_pd(hdog_ctx* ctx, void* stream, const float* d_dog, float* d_hessian,
                    uint8_t* d_pd, int width, int height, int depth) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_dog || !d_hessian || !d_pd) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(launch_hessian<float, float>(d_dog, d_hessian, d_pd, width, height, depth,
                                      (hipStream_t)stream));
  return HDOG_OK;
}

int hdog_ccl3d(hdog_ctx* ctx, void* stream, const uint8_t* d_pd, uint32_t* d_labels, int width,
               int height, int depth) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_pd || !d_labels) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  // the mask's candidate positions in the record layout, then the whole path's labeling
  // of candidates with the label volume written in full
  const hipStream_t s = (hipStream_t)stream;
  const i64 N = (i64)width * height * depth, ntiles = hessian_tiles(width, height, depth);
  const i64 nb = (N + kRootBlock - 1) / kRootBlock;
  HIPCHK(ensure(ctx->cand, (size_t)ntiles * kCandTile * 4));
  HIPCHK(ensure(ctx->cand_cnt, (size_t)ntiles * 4));
  HIPCHK(ensure(ctx->cand_lab, (size_t)ntiles * kCandTile * 4));
  HIPCHK(ensure(ctx->fgbits, (size_t)fg_words(width, height, depth) * 8));
  HIPCHK(ensure(ctx->rmask, (size_t)nb * (kRootBlock / 8)));
  HIPCHK(ensure(ctx->rmulti, (size_t)nb * (kRootBlock / 8)));
  HIPCHK(launch_mask_records(d_pd, (float*)ctx->cand.p, (uint32_t*)ctx->cand_cnt.p, width, height,
                             depth, s));
  HIPCHK(launch_ccl_candidates((const float*)ctx->cand.p, (const uint32_t*)ctx->cand_cnt.p, ntiles,
                               (unsigned long long*)ctx->fgbits.p, d_labels,
                               (uint32_t*)ctx->cand_lab.p, (unsigned long long*)ctx->rmask.p,
                               (unsigned long long*)ctx->rmulti.p, nb * (kRootBlock / 64), true,
                               width, height, depth, s, 1));
  return HDOG_OK;
}

int hdog_ccl3d_noalloc(hdog_ctx* ctx, void* stream, const uint8_t* d_pd, uint32_t* d_labels,
                       int width, int height, int depth) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_pd || !d_labels) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  const i64 N = (i64)width * height * depth, ntiles = hessian_tiles(width, height, depth);
  if (ntiles * kCandTile > kNoAllocVoxels)
    return fail(ctx, HDOG_ERR_ARG,
                "hdog_ccl3d_noalloc: %d x %d x %d pads to %lld voxels of Hessian tiles, more "
                "than its fixed scratch holds (%lld)", width, height, depth,
                (long long)(ntiles * kCandTile), (long long)kNoAllocVoxels);
  HIPCHK(hipSetDevice(ctx->device));
  unsigned char* base = (unsigned char*)noalloc_addr(ctx->device);
  if (!base) return fail(ctx, HDOG_ERR_STATE, "no scratch address for device %d", ctx->device);
  const NoAllocLayout L = noalloc_layout(ntiles, N);
  const hipStream_t s = (hipStream_t)stream;
  // every part is written before it is read: the counts and positions by the mask pass, the
  // mask rows and record labels by the tile-local labeling, the bitmaps and the label
  // volume cleared by fill kernels inside launch_ccl_candidates
  HIPCHK(launch_mask_records(d_pd, (float*)(base + L.cand), (uint32_t*)(base + L.cnt), width,
                             height, depth, s));
  HIPCHK(launch_ccl_candidates((const float*)(base + L.cand), (const uint32_t*)(base + L.cnt),
                               ntiles, (unsigned long long*)(base + L.fg), d_labels,
                               (uint32_t*)(base + L.lab), (unsigned long long*)(base + L.roots),
                               (unsigned long long*)(base + L.multi), L.root_words, true, width,
                               height, depth, s, 1));
  return HDOG_OK;
}

int hdog_regional_features(hdog_ctx* ctx, void* stream, const uint32_t* d_labels,
                           const float* d_hessian, const float* d_norm, const uint16_t* d_img,
                           int width, int height, int depth, int is_ave_mode,
                           int64_t* num_regions) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_labels || !d_hessian || !d_norm) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  return regional_features_t<float, float>(ctx, (hipStream_t)stream, const_cast<uint32_t*>(d_labels),
                                           d_hessian, d_norm, d_img, width, height, depth,
                                           is_ave_mode, num_regions);
}

int hdog_regions_download(hdog_ctx* ctx, void* stream, hdog_regions* out) {
  if (!ctx || !out) return HDOG_ERR_ARG;
  const i64 R = ctx->R;
  out->count = R;
  if (R > out->capacity)
    return fail(ctx, HDOG_ERR_CAPACITY, "%lld regions exceed the caller capacity %lld",
                (long long)R, (long long)out->capacity);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const i64 cap = ctx->out_cap;
  auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  };
  const float* cen = (const float*)ctx->o_cen.p;
  const float* eig = (const float*)ctx->o_eig.p;
  HIPCHK(cp(out->label, ctx->o_label.p, R * 4));
  HIPCHK(cp(out->size, ctx->o_size.p, R * 2));
  HIPCHK(cp(out->voxels, ctx->o_vox.p, R * 8));
  HIPCHK(cp(out->centroid_x, cen, R * 4));
  HIPCHK(cp(out->centroid_y, cen + cap, R * 4));
  HIPCHK(cp(out->centroid_z, cen + 2 * cap, R * 4));
  HIPCHK(cp(out->structureness, eig, R * 4));
  HIPCHK(cp(out->blobness, eig + cap, R * 4));
  HIPCHK(cp(out->max_normalized, ctx->o_nrm.p, R * 4));
  if (out->hessian)
    for (int k = 0; k < 6; ++k)
      HIPCHK(cp(out->hessian + k * out->capacity, (const float*)ctx->o_hess.p + k * cap, R * 4));
  HIPCHK(cp(out->intensity_sum, ctx->o_isum.p, R * 8));
  HIPCHK(cp(out->intensity_max, ctx->o_imax.p, R * 2));
  HIPCHK(hipStreamSynchronize(s));
  return HDOG_OK;
}

static int ensure_gauss(hdog_ctx* ctx, const hdog_params* p) {
  if (ctx->gauss_ready && ctx->sigma[0] == p->sigma_xy1 && ctx->sigma[1] == p->sigma_z1 &&
      ctx->sigma[2] == p->sigma_xy2 && ctx->sigma[3] == p->sigma_z2 && ctx->truncate == p->truncate)
    return HDOG_OK;
  return hdog_init_gaussian3d(ctx, p->sigma_xy1, p->sigma_z1, p->sigma_xy2, p->sigma_z2,
                              p->truncate, nullptr);
}

int hdog_run_device(hdog_ctx* ctx, void* stream, const hdog_params* p, const uint16_t* d_img,
                    void* d_norm, void* d_dog, void* d_hessian, uint8_t* d_pd, uint32_t* d_labels,
                    int64_t* num_regions) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (!d_img) return fail(ctx, HDOG_ERR_ARG, "NULL image pointer");
  HIPCHK(hipSetDevice(ctx->device));
  if ((rc = ensure_gauss(ctx, p))) return rc;
  const i64 N = (i64)p->width * p->height * p->depth;
  const size_t es = p->precision == HDOG_FP32 ? 4 : 2;  // norm and hessian element size
  const size_t ed = (p->precision == HDOG_FP16 || p->precision == HDOG_BF16) ? 2 : 4;  // dog
  if (!d_norm) { HIPCHK(ensure(ctx->norm, N * es)); d_norm = ctx->norm.p; }
  if (!d_dog) { HIPCHK(ensure(ctx->dog, N * ed)); d_dog = ctx->dog.p; }
  // no Hessian or mask buffer requested: not materialized (the region pass works from the
  // candidate records); no label buffer: the union-find's parent array is context scratch,
  // written and read at candidate voxels only
  const bool dense = d_labels != nullptr;
  if (!d_labels) { HIPCHK(ensure(ctx->labels, N * 4)); d_labels = (uint32_t*)ctx->labels.p; }
  hipStream_t s = (hipStream_t)stream;
  typedef __half h16;
  typedef __hip_bfloat16 b16;
  switch (p->precision) {
    case HDOG_FP16:
      return run_stages<h16, h16, h16>(ctx, s, p, d_img, (h16*)d_norm, (h16*)d_dog, (h16*)d_hessian,
                                       d_pd, d_labels, dense, num_regions);
    case HDOG_BF16:
      return run_stages<b16, b16, b16>(ctx, s, p, d_img, (b16*)d_norm, (b16*)d_dog, (b16*)d_hessian,
                                       d_pd, d_labels, dense, num_regions);
    case HDOG_FP16_DOG32:
      return run_stages<h16, float, h16>(ctx, s, p, d_img, (h16*)d_norm, (float*)d_dog,
                                         (h16*)d_hessian, d_pd, d_labels, dense, num_regions);
    case HDOG_BF16_DOG32:
      return run_stages<b16, float, b16>(ctx, s, p, d_img, (b16*)d_norm, (float*)d_dog,
                                         (b16*)d_hessian, d_pd, d_labels, dense, num_regions);
    default:
      return run_stages<float, float, float>(ctx, s, p, d_img, (float*)d_norm, (float*)d_dog,
                                             (float*)d_hessian, d_pd, d_labels, dense, num_regions);
  }
}

int hdog_autotune(hdog_ctx* ctx, void* stream, const hdog_params* p, const uint16_t* d_img,
                  const char* cache_path, int* norm_segments, int* gauss_segments) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  const int W = p->width, H = p->height, D = p->depth, prec = p->precision;
  int ns = 0, gs = 0;
  bool found = false;
  if (cache_path && *cache_path) {  // lines "width height depth precision norm gauss"
    if (FILE* f = std::fopen(cache_path, "r")) {
      int w, h, d, pr, a, b;
      while (std::fscanf(f, "%d %d %d %d %d %d", &w, &h, &d, &pr, &a, &b) == 6)
        if (w == W && h == H && d == D && pr == prec && a >= 1 && a <= kMaxSegments && b >= 1 &&
            b <= kMaxSegments) {
          ns = a;
          gs = b;
          found = true;
        }
      std::fclose(f);
    }
  }
  if (!found) {
    if (!d_img) return fail(ctx, HDOG_ERR_ARG, "no cached launch shape and no stack to time");
    HIPCHK(hipSetDevice(ctx->device));
    if ((rc = ensure_gauss(ctx, p))) return rc;
    hipStream_t s = (hipStream_t)stream;
    float tn = 0, tg = 0;
    typedef __half h16;
    typedef __hip_bfloat16 b16;
    switch (prec) {
      case HDOG_FP16: rc = autotune_t<h16, h16>(ctx, s, p, d_img, &ns, &gs, &tn, &tg); break;
      case HDOG_BF16: rc = autotune_t<b16, b16>(ctx, s, p, d_img, &ns, &gs, &tn, &tg); break;
      case HDOG_FP16_DOG32: rc = autotune_t<h16, float>(ctx, s, p, d_img, &ns, &gs, &tn, &tg); break;
      case HDOG_BF16_DOG32: rc = autotune_t<b16, float>(ctx, s, p, d_img, &ns, &gs, &tn, &tg); break;
      default: rc = autotune_t<float, float>(ctx, s, p, d_img, &ns, &gs, &tn, &tg);
    }
    if (rc) return rc;
    if (cache_path && *cache_path) {
      if (FILE* f = std::fopen(cache_path, "a")) {
        std::fprintf(f, "%d %d %d %d %d %d\n", W, H, D, prec, ns, gs);
        std::fclose(f);
      } else {
        return fail(ctx, HDOG_ERR_ARG, "cannot append to the tuning cache %s", cache_path);
      }
    }
  }
  bool have = false;
  for (hdog_ctx::Tuned& t : ctx->tuned)
    if (t.W == W && t.H == H && t.D == D && t.precision == prec) {
      t.norm_segments = ns;
      t.gauss_segments = gs;
      have = true;
    }
  if (!have) ctx->tuned.push_back({W, H, D, prec, ns, gs});
  if (norm_segments) *norm_segments = ns;
  if (gauss_segments) *gauss_segments = gs;
  return HDOG_OK;
}

int hdog_process_substack(hdog_ctx* ctx, void* stream, const hdog_params* p,
                          const uint16_t* h_img, int depth_valid, hdog_regions* out) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (!h_img || depth_valid < 1 || depth_valid > p->depth)
    return fail(ctx, HDOG_ERR_ARG, "depth_valid %d outside [1, %d]", depth_valid, p->depth);
  HIPCHK(hipSetDevice(ctx->device));
  const i64 plane = (i64)p->width * p->height;
  const i64 N = plane * p->depth;
  HIPCHK(ensure(ctx->img, N * 2));
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(ctx->img.p, h_img, plane * depth_valid * 2, hipMemcpyHostToDevice, s));
  if (depth_valid < p->depth)  // zero-fill, src/3D/HDoG3D_main.cpp:166-169
    HIPCHK(hipMemsetAsync((uint16_t*)ctx->img.p + plane * depth_valid, 0,
                          (N - plane * depth_valid) * 2, s));
  int64_t R = 0;
  rc = hdog_run_device(ctx, stream, p, (const uint16_t*)ctx->img.p, nullptr, nullptr, nullptr,
                       nullptr, nullptr, &R);
  if (rc) return rc;
  if (!out) return HDOG_OK;
  return hdog_regions_download(ctx, stream, out);
}

int hdog_synth_volume(hdog_ctx* ctx, void* stream, uint16_t* d_img, float* d_scratch, int width,
                      int height, int depth, uint64_t seed, double mean0, double mean1, double sd,
                      const double* blobs, int64_t n_blobs) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_img || !d_scratch || (n_blobs > 0 && !blobs) || n_blobs < 0)
    return fail(ctx, HDOG_ERR_ARG, "NULL pointer or negative blob count");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  if (n_blobs > 0) {
    HIPCHK(ensure(ctx->blobs, (size_t)n_blobs * 7 * 8));
    HIPCHK(hipMemcpyAsync(ctx->blobs.p, blobs, (size_t)n_blobs * 7 * 8, hipMemcpyHostToDevice, s));
  }
  HIPCHK(launch_synth(d_img, d_scratch, width, height, depth, seed, mean0, mean1, sd,
                      (const double*)ctx->blobs.p, n_blobs, s));
  HIPCHK(hipStreamSynchronize(s));
  return HDOG_OK;
}

int hdog_soft_forward(hdog_ctx* ctx, void* stream, const float* d_v, const uint8_t* d_target,
                      int width, int height, int depth, const double sigma1[3],
                      const double sigma2[3], const int radius1[3], const int radius2[3],
                      double gamma, double tau, float* d_dog, uint8_t* d_mask, double* loss,
                      int64_t* count) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_v || !d_target || !d_dog || !d_mask || !sigma1 || !sigma2 || !radius1 || !radius2 ||
      !(tau > 0.0))
    return fail(ctx, HDOG_ERR_ARG, "NULL pointer or non-positive tau");
  HIPCHK(hipSetDevice(ctx->device));
  SoftSetup S;
  if ((rc = soft_setup(ctx, sigma1, sigma2, radius1, radius2, gamma, &S))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 8));
  HIPCHK(ensure(ctx->tmp2, (size_t)N * 8));
  HIPCHK(ensure(ctx->hess, (size_t)N * 24));
  HIPCHK(ensure(ctx->sums, 64 * 8));
  HIPCHK(launch_gauss_pair(d_v, ctx->tmp1.p, ctx->tmp2.p, S.w, S.r, (float)S.s, d_dog, nullptr,
                           width, height, depth, s));
  HIPCHK(launch_hessian<float, float>(d_dog, (float*)ctx->hess.p, d_mask, width, height, depth, s));
  HIPCHK(launch_soft_loss((const float*)ctx->hess.p, d_mask, d_target, N, 1.0 / tau,
                          (double*)ctx->sums.p, s));
  double h[2];
  HIPCHK(hipMemcpyAsync(h, ctx->sums.p, 16, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (count) *count = (int64_t)h[1];
  if (loss) *loss = h[1] > 0 ? h[0] / h[1] : 0.0;
  return HDOG_OK;
}

int hdog_soft_backward(hdog_ctx* ctx, void* stream, const float* d_v, const float* d_dog,
                       const uint8_t* d_mask, const uint8_t* d_target, int width, int height,
                       int depth, const double sigma1[3], const double sigma2[3],
                       const int radius1[3], const int radius2[3], double gamma, double tau,
                       int64_t count, double grad_out, float* d_grad_v, double grad_sigma1[3],
                       double grad_sigma2[3], double* grad_tau) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (!d_v || !d_dog || !d_mask || !d_target || !d_grad_v || !(tau > 0.0))
    return fail(ctx, HDOG_ERR_ARG, "NULL pointer or non-positive tau");
  HIPCHK(hipSetDevice(ctx->device));
  SoftSetup S;
  if ((rc = soft_setup(ctx, sigma1, sigma2, radius1, radius2, gamma, &S))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 8));
  HIPCHK(ensure(ctx->tmp2, (size_t)N * 8));
  HIPCHK(ensure(ctx->hess, (size_t)N * 24));
  HIPCHK(ensure(ctx->ddog, (size_t)N * 4));
  HIPCHK(ensure(ctx->pd, (size_t)N));
  HIPCHK(ensure(ctx->sums, 64 * 8));
  double* sums = (double*)ctx->sums.p;
  float* gh = (float*)ctx->hess.p;
  float* ddog = (float*)ctx->ddog.p;
  // H is recomputed from the checkpointed DoG; the mask from the forward pass is used
  HIPCHK(launch_hessian<float, float>(d_dog, gh, (uint8_t*)ctx->pd.p, width, height, depth, s));
  const double inv_count = count > 0 ? grad_out / (double)count : 0.0;
  HIPCHK(launch_soft_grad_h(gh, d_mask, d_target, N, 1.0 / tau, inv_count, sums, s));
  // P_x, P_y, P_z (3N floats) in tmp1, grown to 12 bytes per voxel
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 12));
  HIPCHK(launch_hessian_adjoint(gh, (float*)ctx->tmp1.p, ddog, width, height, depth, s));
  // dL/dv = s (G1 - G2) dL/ddog
  HIPCHK(launch_gauss_pair(ddog, ctx->tmp1.p, ctx->tmp2.p, S.w, S.r, (float)S.s, d_grad_v, nullptr,
                           width, height, depth, s));
  double h[16];
  HIPCHK(launch_dot(ddog, d_dog, N, sums + 1, s));  // <dL/ddog, dog>
  for (int a = 0; a < 3; ++a) {
    const float* w[6] = {S.w[0], S.w[1], S.w[2], S.w[3], S.w[4], S.w[5]};
    w[a] = S.dw[a];
    w[3 + a] = S.dw[3 + a];
    HIPCHK(launch_gauss_pair(d_v, ctx->tmp1.p, ctx->tmp2.p, w, S.r, 0.0f, nullptr,
                             (float2*)ctx->tmp1.p, width, height, depth, s));
    HIPCHK(launch_dot_pair(ddog, (const float2*)ctx->tmp1.p, N, sums + 2 + 2 * a, s));
  }
  HIPCHK(hipMemcpyAsync(h, sums, 8 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (grad_tau) *grad_tau = h[0];
  for (int a = 0; a < 3; ++a) {
    const double sg1 = sigma1[a];
    if (grad_sigma1) grad_sigma1[a] = gamma / sg1 * h[1] + S.s * h[2 + 2 * a];
    if (grad_sigma2) grad_sigma2[a] = -S.s * h[3 + 2 * a];
  }
  return HDOG_OK;
}

int soft_norm_check(hdog_ctx* ctx, int width, int height, int depth, int rxy, int rz) {
  int rc = check_shape(ctx, width, height, depth);
  if (rc) return rc;
  if (rxy < 1 || rz < 1 || rxy > kMaxNormRadius || rz > kMaxNormRadius)
    return fail(ctx, HDOG_ERR_ARG, "normalization radii (%d, %d) outside [1, %d]", rxy, rz,
                kMaxNormRadius);
  return HDOG_OK;
}

int hdog_soft_normalize_forward(hdog_ctx* ctx, void* stream, const float* d_v, float* d_norm,
                                float* d_ed, uint32_t* d_arg, float min_intensity, int width,
                                int height, int depth, int radius_xy, int radius_z) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = soft_norm_check(ctx, width, height, depth, radius_xy, radius_z);
  if (rc) return rc;
  if (!d_v || !d_norm || !d_ed || !d_arg) return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 32));
  HIPCHK(launch_soft_normalize(d_v, d_norm, (float2*)d_ed, d_arg, ctx->tmp1.p, min_intensity,
                               width, height, depth, radius_xy, radius_z, (hipStream_t)stream));
  return HDOG_OK;
}

int hdog_soft_normalize_backward(hdog_ctx* ctx, void* stream, const float* d_v, const float* d_ed,
                                 const uint32_t* d_arg, const float* d_grad_norm,
                                 float* d_grad_v, float min_intensity, int width, int height,
                                 int depth, int radius_xy, int radius_z) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = soft_norm_check(ctx, width, height, depth, radius_xy, radius_z);
  if (rc) return rc;
  if (!d_v || !d_ed || !d_arg || !d_grad_norm || !d_grad_v)
    return fail(ctx, HDOG_ERR_ARG, "NULL device pointer");
  HIPCHK(hipSetDevice(ctx->device));
  const i64 N = (i64)width * height * depth;
  HIPCHK(ensure(ctx->tmp1, (size_t)N * 16));
  HIPCHK(launch_soft_normalize_adjoint(d_v, (const float2*)d_ed, d_arg, d_grad_norm, d_grad_v,
                                       ctx->tmp1.p, min_intensity, width, height, depth,
                                       radius_xy, radius_z, (hipStream_t)stream));
  return HDOG_OK;
}

int hdog_batch_release(hdog_ctx* ctx) {
  if (!ctx) return HDOG_ERR_ARG;
  if (ctx->batch) {
    (void)hipSetDevice(ctx->device);
    batch_free(ctx->batch);
    ctx->batch = nullptr;
  }
  return HDOG_OK;
}

int hdog_batch_plan(hdog_ctx* ctx, const hdog_params* p, const hdog_tile* tiles, int n_tiles,
                    int64_t max_regions_per_stack, int64_t arena_capacity) {
  if (!ctx) return HDOG_ERR_ARG;
  if (!p || !tiles || n_tiles < 1) return fail(ctx, HDOG_ERR_ARG, "NULL params / tiles or no tiles");
  if (max_regions_per_stack < 0 || arena_capacity < 0)
    return fail(ctx, HDOG_ERR_ARG, "negative capacity");
  HIPCHK(hipSetDevice(ctx->device));
  hdog_batch_release(ctx);
  std::vector<i64> Ns(n_tiles);
  i64 totN = 0;
  for (int t = 0; t < n_tiles; ++t) {
    hdog_params q = *p;
    q.width = tiles[t].width;
    q.height = tiles[t].height;
    q.depth = tiles[t].depth;
    int rc = check_params(ctx, &q);
    if (rc) return fail(ctx, rc, "stack %d: %s", t, std::string(ctx->err).c_str());
    if (!tiles[t].image) return fail(ctx, HDOG_ERR_ARG, "stack %d: NULL image", t);
    Ns[t] = (i64)q.width * q.height * q.depth;
    totN += Ns[t];
  }
  int rc = ensure_gauss(ctx, p);
  if (rc) return rc;
  const size_t es = p->precision == HDOG_FP32 ? 4 : 2;
  const size_t ed = (p->precision == HDOG_FP16 || p->precision == HDOG_BF16) ? 2 : 4;
  // longest first to the branch with fewer voxels so far
  std::vector<int> order(n_tiles);
  for (int t = 0; t < n_tiles; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return Ns[a] > Ns[b]; });
  auto deal = [&](int nl) {
    std::vector<BatchLane> L(nl);
    std::vector<i64> load(nl, 0);
    for (int t : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[l] += Ns[t];
      BatchLane& b = L[l];
      b.stacks.push_back(t);
      b.maxN = std::max(b.maxN, Ns[t]);
      b.maxT = std::max(b.maxT, hessian_tiles(tiles[t].width, tiles[t].height, tiles[t].depth));
      b.maxF = std::max(b.maxF, fg_words(tiles[t].width, tiles[t].height, tiles[t].depth));
    }
    for (BatchLane& b : L)
      b.acc_cap = std::min<i64>(max_regions_per_stack > 0 ? max_regions_per_stack + 1
                                                          : b.maxN / 64 + 2,
                                0xFFFFFFFFll);
    return L;
  };
  // the pair buffers' place in the record block (second one 256-byte aligned)
  auto tmp2_off = [](const BatchLane& b) { return ((size_t)b.maxN * 8 + 255) / 256 * 256; };
  auto lane_bytes = [&](const BatchLane& b) -> std::vector<std::pair<Buf*, size_t>> {
    BatchLane& m = const_cast<BatchLane&>(b);
    const i64 nb = (b.maxN + kRootBlock - 1) / kRootBlock;
    return {{&m.norm, (size_t)b.maxN * es}, {&m.dog, (size_t)b.maxN * ed},
            {&m.labels, (size_t)b.maxN * 4}, {&m.rid, (size_t)b.maxN * 4},
            {&m.bcnt, (size_t)block_count_words(b.maxN) * 4},
            {&m.rmask, (size_t)nb * (kRootBlock / 8)}, {&m.rmulti, (size_t)nb * (kRootBlock / 8)},
            {&m.dR, 64}, {&m.acc, (size_t)b.acc_cap * sizeof(RegionAcc)},
            {&m.cand, std::max((size_t)b.maxT * kCandTile * 7 * 4, tmp2_off(b) + (size_t)b.maxN * 8)},
            {&m.cand_cnt, (size_t)b.maxT * 4}, {&m.cand_lab, (size_t)b.maxT * kCandTile * 4},
            {&m.fgbits, (size_t)b.maxF * 8}};
  };
  BatchState* B = new BatchState;
  B->tiles.assign(tiles, tiles + n_tiles);
  B->p = *p;
  B->arena_cap = arena_capacity > 0 ? arena_capacity : totN / 128 + n_tiles;
  const i64 cap = B->arena_cap;
  // two branches when their scratch fits in 90 % of the free device memory, else one
  // the arena's region arrays: 72 bytes per region (see `need` below)
  size_t shared_bytes = (size_t)cap * 72 + ctx->weights.bytes + (size_t)n_tiles * sizeof(TileOut);
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  B->lanes = deal(n_tiles >= 2 ? 2 : 1);
  if (B->lanes.size() > 1) {
    size_t need = shared_bytes;
    for (const BatchLane& b : B->lanes)
      for (const auto& n : lane_bytes(b)) need += n.second;
    if ((double)need > 0.9 * (double)free_b) {
      // the per-stack path's working buffers (grown again by its next call; the last
      // call's region arrays stay) may hold the memory, e.g. after the tuner ran
      for (Buf* w : {&ctx->tmp1, &ctx->tmp2, &ctx->rid, &ctx->bcnt, &ctx->rmask, &ctx->rmulti,
                     &ctx->dR, &ctx->acc, &ctx->cand_lab, &ctx->fgbits, &ctx->img, &ctx->norm,
                     &ctx->dog, &ctx->hess, &ctx->pd, &ctx->labels, &ctx->cand, &ctx->cand_cnt})
        release(*w);
      (void)hipMemGetInfo(&free_b, &total_b);
      if ((double)need > 0.9 * (double)free_b) B->lanes = deal(1);
    }
  }
  auto bad = [&](hipError_t e, const char* what) {
    batch_free(B);
    return fail(ctx, HDOG_ERR_HIP, "batch plan: %s failed: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  for (BatchLane& b : B->lanes) {
    B->acc_cap = std::max(B->acc_cap, b.acc_cap);  // for messages
    for (const auto& n : lane_bytes(b))
      if ((e = ensure(*n.first, n.second)) != hipSuccess) return bad(e, "device allocation");
    b.tmp1 = b.cand.p;
    b.tmp2 = (char*)b.cand.p + tmp2_off(b);
  }
  const struct {
    Buf* b;
    size_t bytes;
  } need[] = {{&B->weights, ctx->weights.bytes},
              {&B->tout, (size_t)n_tiles * sizeof(TileOut)}, {&B->running, 64},
              {&B->a_label, (size_t)cap * 4}, {&B->a_size, (size_t)cap * 2},
              {&B->a_vox, (size_t)cap * 8}, {&B->a_cen, (size_t)cap * 12},
              {&B->a_eig, (size_t)cap * 8}, {&B->a_nrm, (size_t)cap * 4},
              {&B->a_hess, (size_t)cap * 24}, {&B->a_isum, (size_t)cap * 8},
              {&B->a_imax, (size_t)cap * 2}};
  for (const auto& n : need)
    if ((e = ensure(*n.b, n.bytes)) != hipSuccess) return bad(e, "device allocation");
  if ((e = hipHostMalloc((void**)&B->h_tout, n_tiles * sizeof(TileOut), 0)) != hipSuccess)
    return bad(e, "hipHostMalloc");
  if ((e = hipMemcpy(B->weights.p, ctx->weights.p, ctx->weights.bytes, hipMemcpyDeviceToDevice)) !=
      hipSuccess)
    return bad(e, "weight copy");
  if ((e = hipStreamCreateWithFlags(&B->cs, hipStreamNonBlocking)) != hipSuccess)
    return bad(e, "hipStreamCreate");
  if ((e = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming)) != hipSuccess)
    return bad(e, "hipEventCreate");
  B->joins.assign(B->lanes.size(), nullptr);
  B->lanes[0].s = B->cs;
  for (size_t l = 1; l < B->lanes.size(); ++l) {
    if ((e = hipStreamCreateWithFlags(&B->lanes[l].s, hipStreamNonBlocking)) != hipSuccess)
      return bad(e, "hipStreamCreate");
    if ((e = hipEventCreateWithFlags(&B->joins[l], hipEventDisableTiming)) != hipSuccess)
      return bad(e, "hipEventCreate");
  }
  if ((e = hipStreamBeginCapture(B->cs, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return bad(e, "hipStreamBeginCapture");
  const hipError_t ec = batch_capture_any(ctx, B);
  e = hipStreamEndCapture(B->cs, &B->graph);
  if (ec != hipSuccess) return bad(ec, "kernel launch during capture");
  if (e != hipSuccess) return bad(e, "hipStreamEndCapture");
  if ((e = hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0)) != hipSuccess)
    return bad(e, "hipGraphInstantiate");
  ctx->batch = B;
  return HDOG_OK;
}

int hdog_batch_run(hdog_ctx* ctx, void* stream) {
  if (!ctx) return HDOG_ERR_ARG;
  BatchState* B = ctx->batch;
  if (!B) return fail(ctx, HDOG_ERR_STATE, "hdog_batch_plan not called");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  B->ran = false;
  HIPCHK(hipGraphLaunch(B->exec, s));
  HIPCHK(hipMemcpyAsync(B->h_tout, B->tout.p, B->tiles.size() * sizeof(TileOut),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  B->ran = true;
  for (size_t t = 0; t < B->tiles.size(); ++t)
    if (B->h_tout[t].overflow)
      return fail(ctx, HDOG_ERR_CAPACITY,
                  "stack %d: %u regions do not fit (per-stack capacity %lld, arena %lld)", (int)t,
                  B->h_tout[t].found, (long long)(B->acc_cap - 1), (long long)B->arena_cap);
  return HDOG_OK;
}

int hdog_batch_tile_info(hdog_ctx* ctx, int tile, int64_t* num_regions, int* overflow) {
  if (!ctx) return HDOG_ERR_ARG;
  BatchState* B = ctx->batch;
  if (!B || !B->ran) return fail(ctx, HDOG_ERR_STATE, "no completed hdog_batch_run");
  if (tile < 0 || tile >= (int)B->tiles.size())
    return fail(ctx, HDOG_ERR_ARG, "stack %d outside [0, %d)", tile, (int)B->tiles.size());
  if (num_regions) *num_regions = B->h_tout[tile].found;
  if (overflow) *overflow = (int)B->h_tout[tile].overflow;
  return HDOG_OK;
}

int hdog_batch_download(hdog_ctx* ctx, void* stream, int tile, hdog_regions* out) {
  if (!ctx) return HDOG_ERR_ARG;
  BatchState* B = ctx->batch;
  if (!B || !B->ran) return fail(ctx, HDOG_ERR_STATE, "no completed hdog_batch_run");
  if (tile < 0 || tile >= (int)B->tiles.size() || !out)
    return fail(ctx, HDOG_ERR_ARG, "stack %d outside [0, %d) or NULL output", tile,
                (int)B->tiles.size());
  const TileOut& T = B->h_tout[tile];
  if (T.overflow)
    return fail(ctx, HDOG_ERR_CAPACITY, "stack %d overflowed its region capacity", tile);
  const i64 n = (i64)T.count, off = (i64)T.offset, R = n + 1;
  out->count = R;
  if (R > out->capacity)
    return fail(ctx, HDOG_ERR_CAPACITY, "%lld regions exceed the caller capacity %lld",
                (long long)R, (long long)out->capacity);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const i64 cap = B->arena_cap;
  // entry 0 (background) is zero, entries 1..n come from arena slots off..off+n-1
  auto cp = [&](void* dst, const void* src, size_t esz) -> hipError_t {
    if (!dst) return hipSuccess;
    std::memset(dst, 0, esz);
    if (n == 0) return hipSuccess;
    return hipMemcpyAsync((char*)dst + esz, (const char*)src + off * esz, n * esz,
                          hipMemcpyDeviceToHost, s);
  };
  const float* cen = (const float*)B->a_cen.p;
  const float* eig = (const float*)B->a_eig.p;
  HIPCHK(cp(out->label, B->a_label.p, 4));
  HIPCHK(cp(out->size, B->a_size.p, 2));
  HIPCHK(cp(out->voxels, B->a_vox.p, 8));
  HIPCHK(cp(out->centroid_x, cen, 4));
  HIPCHK(cp(out->centroid_y, cen + cap, 4));
  HIPCHK(cp(out->centroid_z, cen + 2 * cap, 4));
  HIPCHK(cp(out->structureness, eig, 4));
  HIPCHK(cp(out->blobness, eig + cap, 4));
  HIPCHK(cp(out->max_normalized, B->a_nrm.p, 4));
  if (out->hessian)
    for (int k = 0; k < 6; ++k)
      HIPCHK(cp(out->hessian + k * out->capacity, (const float*)B->a_hess.p + k * cap, 4));
  HIPCHK(cp(out->intensity_sum, B->a_isum.p, 8));
  HIPCHK(cp(out->intensity_max, B->a_imax.p, 2));
  HIPCHK(hipStreamSynchronize(s));
  return HDOG_OK;
}

int hdog_host_alloc(hdog_ctx* ctx, size_t bytes, void** p) {
  if (!ctx || !p) return HDOG_ERR_ARG;
  *p = nullptr;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
  return HDOG_OK;
}

int hdog_host_free(hdog_ctx* ctx, void* p) {
  if (!ctx) return HDOG_ERR_ARG;
  if (p) HIPCHK(hipHostFree(p));
  return HDOG_OK;
}

}  // extern "C"

// Fills device slot `slot` on the context's copy stream: `keep` slices from slice `from`
// of the other slot's upload (device to device), then n_new host slices, then zeros up to
// p->depth.  The slot is rewritten only after its previous upload and the last run
// reading it ended; the other slot's upload precedes this copy on the same stream.
static int upload_slot(hdog_ctx* ctx, const hdog_params* p, int slot, int from, int keep,
                       const uint16_t* h_img, int n_new) {
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->copy_stream) HIPCHK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  for (hipEvent_t* ev : {&ctx->up_done[slot], &ctx->run_done[slot]})
    if (!*ev) HIPCHK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  const i64 plane = (i64)p->width * p->height;
  const i64 N = plane * p->depth;
  HIPCHK(hipEventSynchronize(ctx->up_done[slot]));
  HIPCHK(hipEventSynchronize(ctx->run_done[slot]));
  HIPCHK(ensure(ctx->up_img[slot], (size_t)N * 2));
  uint16_t* d = (uint16_t*)ctx->up_img[slot].p;
  if (keep > 0)
    HIPCHK(hipMemcpyAsync(d, (const uint16_t*)ctx->up_img[1 - slot].p + plane * from,
                          (size_t)(plane * keep) * 2, hipMemcpyDeviceToDevice, ctx->copy_stream));
  if (n_new > 0)
    HIPCHK(hipMemcpyAsync(d + plane * keep, h_img, (size_t)(plane * n_new) * 2,
                          hipMemcpyHostToDevice, ctx->copy_stream));
  const int valid = keep + n_new;
  if (valid < p->depth)  // zero-fill, src/3D/HDoG3D_main.cpp:166-169
    HIPCHK(hipMemsetAsync(d + plane * valid, 0, (size_t)(N - plane * valid) * 2,
                          ctx->copy_stream));
  HIPCHK(hipEventRecord(ctx->up_done[slot], ctx->copy_stream));
  ctx->up_shape[slot][0] = p->width;
  ctx->up_shape[slot][1] = p->height;
  ctx->up_shape[slot][2] = p->depth;
  ctx->up_valid[slot] = valid;
  return HDOG_OK;
}

extern "C" {

int hdog_substack_upload(hdog_ctx* ctx, const hdog_params* p, const uint16_t* h_img,
                         int depth_valid, int slot) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (!h_img || depth_valid < 1 || depth_valid > p->depth || slot < 0 || slot > 1)
    return fail(ctx, HDOG_ERR_ARG, "depth_valid %d outside [1, %d] or slot %d not 0 / 1",
                depth_valid, p->depth, slot);
  return upload_slot(ctx, p, slot, 0, 0, h_img, depth_valid);
}

int hdog_substack_upload_overlap(hdog_ctx* ctx, const hdog_params* p, int from, int keep,
                                 const uint16_t* h_img, int n_new, int slot) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (slot < 0 || slot > 1 || from < 0 || keep < 0 || n_new < 0 || keep + n_new < 1 ||
      keep + n_new > p->depth || (n_new > 0 && !h_img))
    return fail(ctx, HDOG_ERR_ARG,
                "slot %d, from %d, keep %d, n_new %d: need slot 0 / 1 and 1 <= keep + n_new <= %d",
                slot, from, keep, n_new, p->depth);
  const int o = 1 - slot;
  const int* up = ctx->up_shape[o];
  if (keep > 0 && (!ctx->up_done[o] || up[0] != p->width || up[1] != p->height ||
                   up[2] != p->depth || from + keep > ctx->up_valid[o]))
    return fail(ctx, HDOG_ERR_STATE,
                "slot %d holds %d valid slices of %d x %d x %d, not slices %d .. %d of "
                "%d x %d x %d", o, ctx->up_valid[o], up[0], up[1], up[2], from, from + keep - 1,
                p->width, p->height, p->depth);
  return upload_slot(ctx, p, slot, from, keep, h_img, n_new);
}

int hdog_substack_run(hdog_ctx* ctx, void* stream, const hdog_params* p, int slot,
                      hdog_regions* out) {
  if (!ctx) return HDOG_ERR_ARG;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (slot < 0 || slot > 1 || !ctx->up_done[slot])
    return fail(ctx, HDOG_ERR_STATE, "slot %d holds no upload", slot);
  const int* up = ctx->up_shape[slot];
  if (up[0] != p->width || up[1] != p->height || up[2] != p->depth ||
      ctx->up_img[slot].bytes < (size_t)p->width * p->height * p->depth * 2)
    return fail(ctx, HDOG_ERR_STATE,
                "slot %d holds an upload of %d x %d x %d, not the %d x %d x %d of these params",
                slot, up[0], up[1], up[2], p->width, p->height, p->depth);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipStreamWaitEvent(s, ctx->up_done[slot], 0));
  int64_t R = 0;
  rc = hdog_run_device(ctx, stream, p, (const uint16_t*)ctx->up_img[slot].p, nullptr, nullptr,
                       nullptr, nullptr, nullptr, &R);
  HIPCHK(hipEventRecord(ctx->run_done[slot], s));
  if (rc) return rc;
  if (!out) return HDOG_OK;
  return hdog_regions_download(ctx, stream, out);
}

}  // extern "C"
