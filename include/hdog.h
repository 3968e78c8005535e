/* hdog.h -- C-ABI of the MI355X HDoG cell-candidate detection library (libhdog.so).
 *
 * Plain C types only: device pointers are caller-owned HIP allocations, the stream
 * is a hipStream_t passed as void* (NULL = default stream).  Every entry point
 * returns an hdog_status; hdog_last_error() gives the message of the last failure
 * on that context.  One context per host thread; contexts share nothing.
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * the lsb-riken/CUBIC-informatics tree).  The reference functions are C/C++ linkage
 * launch wrappers on the legacy default stream with process-global state; these take
 * an explicit context and stream instead and hold no globals.
 *
 * Volumes are x-fastest: index i = (z*height + y)*width + x, N = width*height*depth,
 * N < 2^32 (labels are uint32; the reference's int32 labels overflow at 2^31).
 */
#ifndef HDOG_H
#define HDOG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hdog_ctx hdog_ctx;

enum hdog_status {
  HDOG_OK = 0,
  HDOG_ERR_ARG = 1,      /* invalid argument (shape, radius, NULL pointer) */
  HDOG_ERR_HIP = 2,      /* HIP runtime error (allocation, launch) */
  HDOG_ERR_STATE = 3,    /* call order violated (e.g. DoG before init_gaussian3d) */
  HDOG_ERR_CAPACITY = 4  /* caller's region arrays are too small */
};

/* Storage type of the float intermediates (normalized, DoG, Hessian).  Every
 * separable pass accumulates in FP32 whatever the storage type.  The *_DOG32 forms keep
 * the DoG volume in FP32 (its second differences feed the candidate test) and store the
 * normalized volume and the Hessian elements in 2 bytes. */
enum hdog_precision {
  HDOG_FP32 = 0,
  HDOG_FP16 = 1,        /* norm, dog, hessian in FP16 */
  HDOG_BF16 = 2,        /* norm, dog, hessian in BF16 */
  HDOG_FP16_DOG32 = 3,  /* norm, hessian in FP16; dog in FP32 */
  HDOG_BF16_DOG32 = 4   /* norm, hessian in BF16; dog in FP32 */
};

/* Per-region results, structure of arrays indexed by region id r = 0..count-1 in
 * ascending label order (region 0 = background, never filled).  The caller supplies
 * `capacity` entries for every non-NULL array; NULL arrays are skipped.
 * Mirrors the arrays src/3D/HDoG3D_main.cpp:216-227 downloads. */
typedef struct hdog_regions {
  int64_t capacity;         /* in: entries available per array */
  int64_t count;            /* out: number of regions R including region 0 */
  uint32_t* label;          /* minimum linear index of the component */
  uint16_t* size;           /* voxel count modulo 65536 (u16 RLE of the reference) */
  uint64_t* voxels;         /* exact voxel count */
  float* centroid_x;        /* sum(x) / (float)size */
  float* centroid_y;
  float* centroid_z;
  float* structureness;     /* sum of squared eigenvalues of the summed Hessian */
  float* blobness;          /* e2^2 / (e0*e1) */
  float* max_normalized;    /* max of normalized intensity (sum in ave mode) */
  float* hessian;           /* 6*capacity: [xx,xy,xz,yy,yz,zz] sums, pitch capacity */
  uint64_t* intensity_sum;  /* exact sum of raw uint16 intensity */
  uint16_t* intensity_max;  /* max raw intensity (MaxIntensity, RegionalFeatures.cu:39-63) */
} hdog_regions;

/* Parameters of one substack run (HDoG_param of the JSON, src/3D/utils.h:48-93). */
typedef struct hdog_params {
  int width, height, depth;  /* substack shape; depth = HDoG_param.depth */
  int radius_xy, radius_z;   /* radius_norm.{xy_large,z_large} */
  float min_intensity;       /* HDoG_param.min_intensity */
  float sigma_xy1, sigma_z1, sigma_xy2, sigma_z2;  /* sigma_dog */
  double truncate;           /* Gaussian radius = (int)(truncate*sigma + 0.5); reference 2.0 */
  float gamma;               /* gamma_n */
  int is_ave_mode;           /* top-level is_ave_mode: sum instead of max of normalized */
  int precision;             /* enum hdog_precision */
  /* Reduced-precision forms only: a leading minor whose magnitude is below minor_band times
   * the sum of its terms' magnitudes has its sign re-decided in FP64.  0 selects 1e-5,
   * which covers every sign the FP32 evaluation can flip; the precision-form config quotes
   * 1e-3.  Ignored by HDOG_FP32. */
  float minor_band;
} hdog_params;

/* ---- context --------------------------------------------------------------- */
/* Selects `device` (hipSetDevice) and creates an empty context.  Replaces the device
 * set-up of initializeGPU, src/3D/HDoG3D_main.cpp:58-109. */
int hdog_create(hdog_ctx** ctx, int device);
void hdog_destroy(hdog_ctx* ctx);
const char* hdog_last_error(const hdog_ctx* ctx);
/* Frees scratch buffers (they grow on demand and are kept between calls). */
int hdog_release_scratch(hdog_ctx* ctx);

/* ---- stages (device pointers, stream ordered) ---------------------------- */

/* Separable min and max filter (clamp), box means of both (clamp), then
 * norm = I >= min_intensity ? (I - E)/(D - E) : 0; a thresholded copy when either
 * radius is 0.  Replaces Normalize3DFilter, src/3D/NormalizeFilter.cuh:5-12
 * (src/3D/NormalizeFilter.cu:603-654). */
int hdog_normalize3d(hdog_ctx* ctx, void* stream, const uint16_t* d_img, float* d_norm,
                     float min_intensity, int width, int height, int depth,
                     int radius_xy, int radius_z);

/* As hdog_normalize3d, and also writes the box means of the min filter (E) and of the
 * max filter (D) that the normalization divides by: the erosion_l / dilation_l volumes
 * of Normalize3DFilter (src/3D/NormalizeFilter.cu:603-654) that test/Normalize_test.cpp
 * dumps (:105-125).  With a zero radius no filter runs and E = D = 0. */
int hdog_normalize3d_ed(hdog_ctx* ctx, void* stream, const uint16_t* d_img, float* d_norm,
                        float* d_erosion, float* d_dilation, float min_intensity, int width,
                        int height, int depth, int radius_xy, int radius_z);

/* Gaussian weights for sigma1 (xy, z) and sigma2 (xy, z); *r_z2 receives the z radius
 * of sigma2, which sets the substack depth margin.  Replaces initGaussian3DKernel,
 * src/3D/DoG3DFilter.cuh:5-9 (src/3D/DoG3DFilter.cu:372-415) with the truncation
 * (reference constant 2.0, :15) as a parameter. */
int hdog_init_gaussian3d(hdog_ctx* ctx, float sigma_xy1, float sigma_z1, float sigma_xy2,
                         float sigma_z2, double truncate, int* r_z2);

/* dog = (G_sigma1 - G_sigma2) * powf(sigma_xy1, 2 gamma) * powf(sigma_z1, gamma), zero
 * padding.  Replaces DoG3DFilter, src/3D/DoG3DFilter.cuh:17-24. */
int hdog_dog3d(hdog_ctx* ctx, void* stream, const float* d_norm, float* d_dog,
               int width, int height, int depth, float gamma);

/* Hessian elements by repeated central differences (zero padding), SoA
 * [xx,xy,xz,yy,yz,zz] at pitch N, and the negative-definite candidate mask.
 * Replaces HessianPositiveDefiniteWithElement, src/3D/Hessian3DFilter_element.cuh:4-8. */
int hdog_hessian_pd(hdog_ctx* ctx, void* stream, const float* d_dog, float* d_hessian,
                    uint8_t* d_pd, int width, int height, int depth);

/* 26-neighbour connected components over linear-index offsets (x and y borders wrap
 * as in the reference); foreground voxels receive the minimum linear index of their
 * component, background 0, voxel 0 is always background.  Replaces CCL,
 * src/3D/CCL3D.cuh:5-10. */
int hdog_ccl3d(hdog_ctx* ctx, void* stream, const uint8_t* d_pd, uint32_t* d_labels,
               int width, int height, int depth);

/* hdog_ccl3d without any allocation: no hipMalloc, hipFree or allocating library call, so
 * it can be called where allocation is not allowed (under stream capture into a HIP graph,
 * in tight serving loops, in harnesses that forbid it).  It has no counterpart in the
 * reference.  Its scratch is one fixed device array per GPU, held until the process ends,
 * sized for volumes that pad to at most 2^23 voxels of 64 x 8 x 16 Hessian tiles (about
 * 70 MB); larger volumes are refused with HDOG_ERR_ARG.  Every call on a GPU shares that
 * array, so calls that overlap on one GPU (other streams, other contexts, other threads)
 * race and must be serialized by the caller.  Same results as hdog_ccl3d. */
int hdog_ccl3d_noalloc(hdog_ctx* ctx, void* stream, const uint8_t* d_pd, uint32_t* d_labels,
                       int width, int height, int depth);

/* Per-region features of a labelled substack: Hessian sums, eigen-features, max (or
 * sum) of normalized intensity, size, centroid, plus exact intensity sum / max.
 * d_img may be NULL (intensity outputs then read 0).  Results stay on the device in
 * the context; read them with hdog_regions_download.  Synchronizes `stream` once to
 * learn the region count.  Replaces HessianFeatures + Eigenvalues + MaxNormalized /
 * SumNormalized + RegionalSizeAndCentroid (src/3D/RegionalFeatures.cuh:16-102,
 * src/3D/Eigenvalues.cuh:21-32) as sequenced at src/3D/HDoG3D_main.cpp:184-214. */
int hdog_regional_features(hdog_ctx* ctx, void* stream, const uint32_t* d_labels,
                           const float* d_hessian, const float* d_norm,
                           const uint16_t* d_img, int width, int height, int depth,
                           int is_ave_mode, int64_t* num_regions);

/* Copies the last hdog_regional_features result into caller host arrays (regions
 * 0..count-1; entry 0 zero) and synchronizes `stream`.  Replaces the D2H copies of
 * src/3D/HDoG3D_main.cpp:216-227. */
int hdog_regions_download(hdog_ctx* ctx, void* stream, hdog_regions* out);

/* ---- fused substack ------------------------------------------------------- */
/* One substack end to end: copies depth_valid host slices (h_img, x-fastest) to the
 * device, zero-fills up to p->depth, runs every stage on p->depth slices and
 * downloads the region arrays into *out.  Replaces processSubstack,
 * src/3D/HDoG3D_main.cpp:134-257 (minus file I/O, which stays in the host program).
 * Gaussian weights are (re)initialised from p when they differ from the last call. */
int hdog_process_substack(hdog_ctx* ctx, void* stream, const hdog_params* p,
                          const uint16_t* h_img, int depth_valid, hdog_regions* out);

/* Split form of hdog_process_substack, to overlap the upload of the next substack with
 * the compute of this one (the reference reads, copies and computes serially,
 * src/3D/HDoG3D_main.cpp:155-170).  hdog_substack_upload copies depth_valid host slices
 * (best from hdog_host_alloc memory: then the copy is asynchronous) into device slot 0 or 1
 * on the context's own copy stream, zero-fills up to p->depth and returns at once; the
 * host buffer must stay unchanged until the next upload into the same slot returns.
 * hdog_substack_run makes `stream` wait for that slot's upload, runs every stage on it
 * and downloads the regions into *out (NULL: keep them on the device), synchronizing
 * `stream`.  Alternate the slots: upload k+1 into one while k runs from the other. */
int hdog_substack_upload(hdog_ctx* ctx, const hdog_params* p, const uint16_t* h_img,
                         int depth_valid, int slot);
int hdog_substack_run(hdog_ctx* ctx, void* stream, const hdog_params* p, int slot,
                      hdog_regions* out);

/* hdog_substack_upload for a substack that overlaps the one uploaded into the other slot:
 * its slices 0 .. keep-1 are slices from .. from+keep-1 of that upload (copied on the
 * device, which needs from + keep no more than that upload's depth_valid), slices keep ..
 * keep+n_new-1 come from h_img (n_new host slices), and the rest up to p->depth is
 * zero-filled.  Substacks advance by depth - 2*margin slices, so consecutive substacks of
 * a stack share all but that many: the reference reads and copies every slice of every
 * substack again (src/3D/HDoG3D_main.cpp:155-170); here only the new ones cross the link.
 * Same slot rules as hdog_substack_upload; HDOG_ERR_STATE when the other slot does not
 * hold enough valid slices of this shape. */
int hdog_substack_upload_overlap(hdog_ctx* ctx, const hdog_params* p, int from, int keep,
                                 const uint16_t* h_img, int n_new, int slot);

/* Page-locked host memory (hipHostMalloc): host-to-device copies from it run at the full
 * link rate and asynchronously.  No counterpart in the reference, which reads into
 * cudaHostAlloc buffers it allocates itself (src/3D/HDoG3D_main.cpp:58-109). */
int hdog_host_alloc(hdog_ctx* ctx, size_t bytes, void** p);
int hdog_host_free(hdog_ctx* ctx, void* p);

/* Same pipeline on a volume already resident on the device.  The intermediates are
 * written to caller buffers when the pointers are non-NULL (else context scratch is
 * used); storage types follow p->precision (enum hdog_precision).  With d_hessian NULL
 * the six Hessian volumes are not materialized at all: the candidate pass writes the mask
 * and a 28-byte record (voxel index, six elements) per candidate voxel, which the region
 * pass reads instead of the volumes (identical sums; about 48 of the 80 bytes per voxel
 * of FP32 traffic saved at a few percent candidates).  Region results stay on the device
 * as for hdog_regional_features. */
int hdog_run_device(hdog_ctx* ctx, void* stream, const hdog_params* p, const uint16_t* d_img,
                    void* d_norm, void* d_dog, void* d_hessian, uint8_t* d_pd,
                    uint32_t* d_labels, int64_t* num_regions);

/* Autotuner of the run-time launch shapes for one stack shape and storage form (no
 * counterpart in the reference, whose tile shapes are compile-time constants,
 * src/3D/NormalizeFilter.cu:13-47, src/3D/DoG3DFilter.cu:29-58): the number of row
 * segments each of the normalization's and the DoG's row walkers splits a plane into,
 * chosen by timing every candidate (1..4) on d_img with HIP events (median of 3).  With
 * cache_path, a line "width height depth precision norm gauss" for the shape is read from
 * that file when present (d_img may then be NULL), and a new measurement is appended.  The
 * choice is kept in the context and used by hdog_run_device and hdog_batch_plan for stacks
 * of that shape; it changes launch shapes only, never results. */
int hdog_autotune(hdog_ctx* ctx, void* stream, const hdog_params* p, const uint16_t* d_img,
                  const char* cache_path, int* norm_segments, int* gauss_segments);

/* ---- batched stacks ------------------------------------------------------------
 * Many whole stacks (tiles of one brain) resident on the device, run as ONE captured
 * hipGraph launch.  Replaces the per-stack fan-out of script/HDoG_gpu.py:67-104, which
 * starts one HDoG3D_main process per stack and runs each as depth-32 substacks.
 *
 * hdog_batch_plan deals the stacks, longest first, to two branches of the graph (one when
 * the second branch's scratch does not fit in device memory); a branch runs its stacks in
 * order on buffers sized for its largest stack, with a region accumulator of
 * max_regions_per_stack entries (0: its largest N / 64), and the branches run
 * concurrently, so one stack's labeling overlaps another stack's filters.  A region arena
 * of arena_capacity records (0: total N / 128) is shared by all stacks in the order they
 * reserve it.  The pipeline of p (sizes ignored, precision honoured) of every tile is
 * captured into one graph.  The
 * tiles' images must stay valid and unchanged in shape while the plan exists; their
 * contents may change between runs.  hdog_batch_run launches the graph on `stream`,
 * waits for it and returns HDOG_ERR_CAPACITY (naming the stack) when a stack's regions
 * did not fit; the other stacks' results are still valid.  Region ids are per stack,
 * labels are linear indices within the stack.  One plan per context; a new plan or
 * hdog_batch_release frees the previous one.  Per-stack calls on the same context do
 * not disturb a plan (it owns its buffers). */
typedef struct hdog_tile {
  const uint16_t* image;  /* device, depth x height x width, x-fastest */
  int width, height, depth;
} hdog_tile;
int hdog_batch_plan(hdog_ctx* ctx, const hdog_params* p, const hdog_tile* tiles, int n_tiles,
                    int64_t max_regions_per_stack, int64_t arena_capacity);
int hdog_batch_run(hdog_ctx* ctx, void* stream);
/* After hdog_batch_run: regions found in stack `tile` (excluding the background) and
 * whether they fit (1 = overflow, nothing stored). */
int hdog_batch_tile_info(hdog_ctx* ctx, int tile, int64_t* num_regions, int* overflow);
/* Copies stack `tile`'s regions into *out in the layout of hdog_regions_download
 * (entry 0 = background, zero). */
int hdog_batch_download(hdog_ctx* ctx, void* stream, int tile, hdog_regions* out);
int hdog_batch_release(hdog_ctx* ctx);

/* ---- kernel timing ----------------------------------------------------------- */
/* With profiling on, hdog_run_device / hdog_process_substack record a HIP event on the
 * caller's stream after every kernel they enqueue; hdog_stage_times waits for the last
 * run and writes up to max_entries (device milliseconds, "stage/kernel" name) pairs in
 * launch order; returns how many were written (0 when none are available).  The
 * reference only prints host wall-clock per substack (src/3D/HDoG3D_main.cpp:162-230). */
int hdog_set_profiling(hdog_ctx* ctx, int on);
int hdog_stage_times(hdog_ctx* ctx, double* ms, const char** names, int max_entries);

/* ---- differentiable blob response ---------------------------------------------
 * No counterpart in the reference, which only thresholds the Hessian.  For a float
 * volume v (x-fastest), anisotropic sigmas sigma{1,2}[3] = (x, y, z) with fixed radii
 * radius{1,2}[3], normalised weights exp(-i^2 / 2 sigma^2) / Z:
 *   dog = (sx1 sy1 sz1)^gamma (G_sigma1 v - G_sigma2 v)      (zero padding)
 *   H   = Hessian of dog by the detector's central differences, M = its candidate mask
 *   loss = mean over M of BCE(sigmoid(-det(H) / tau), target)   (target: uint8 0/1)
 * hdog_soft_forward writes the DoG (checkpoint for the backward pass) and M, and
 * returns the loss and |M|; it synchronizes `stream`.  hdog_soft_backward recomputes H
 * from the DoG, holds M fixed and returns grad_out * d loss / d (v, sigma1, sigma2, tau)
 * through hand-written adjoints; it synchronizes `stream`. */
int hdog_soft_forward(hdog_ctx* ctx, void* stream, const float* d_v, const uint8_t* d_target,
                      int width, int height, int depth, const double sigma1[3],
                      const double sigma2[3], const int radius1[3], const int radius2[3],
                      double gamma, double tau, float* d_dog, uint8_t* d_mask, double* loss,
                      int64_t* count);
int hdog_soft_backward(hdog_ctx* ctx, void* stream, const float* d_v, const float* d_dog,
                       const uint8_t* d_mask, const uint8_t* d_target, int width, int height,
                       int depth, const double sigma1[3], const double sigma2[3],
                       const int radius1[3], const int radius2[3], double gamma, double tau,
                       int64_t count, double grad_out, float* d_grad_v, double grad_sigma1[3],
                       double grad_sigma2[3], double* grad_tau);

/* The local-contrast normalization in front of the differentiable blob response, on a
 * float volume: E = box(min3(v)), D = box(max3(v)) (clamp-to-edge windows and box of radius
 * (radius_xy, radius_xy, radius_z)), norm = v >= min_intensity ? (v - E)/(D - E) : 0
 * (src/3D/NormalizeFilter.cu:53-654 with a float input).  The forward pass writes norm and,
 * for the backward pass, d_ed (2N floats: (E, D) pairs) and d_arg (2N: each voxel's window
 * argmin, then argmax, the lowest linear index among equal extremes).  The backward pass
 * writes d_grad_v = dL/dv from d_grad_norm = dL/dnorm: the min and max filters by their
 * subgradient (gradient routed to the argmin / argmax), the box means by their clamped
 * transposes.  No counterpart in the reference. */
int hdog_soft_normalize_forward(hdog_ctx* ctx, void* stream, const float* d_v, float* d_norm,
                                float* d_ed, uint32_t* d_arg, float min_intensity, int width,
                                int height, int depth, int radius_xy, int radius_z);
int hdog_soft_normalize_backward(hdog_ctx* ctx, void* stream, const float* d_v, const float* d_ed,
                                 const uint32_t* d_arg, const float* d_grad_norm,
                                 float* d_grad_v, float min_intensity, int width, int height,
                                 int depth, int radius_xy, int radius_z);

/* ---- synthetic input (benchmark / tests) ---------------------------------- */
/* Renders noise + Gaussian blobs into d_img (depth x height x width uint16):
 * background mean0 + (mean1 - mean0) * z / (depth-1), sd `sd` (counter-based normal
 * variates keyed by seed), clipped at 0, plus n_blobs blobs given as rows
 * (cz, cy, cx, sz, sy, sx, A) of the host table `blobs` (double), each evaluated within
 * 4 sigma of its centre; then rint and clip to [0, 65535].  d_scratch must hold N
 * floats. */
int hdog_synth_volume(hdog_ctx* ctx, void* stream, uint16_t* d_img, float* d_scratch,
                      int width, int height, int depth, uint64_t seed, double mean0,
                      double mean1, double sd, const double* blobs, int64_t n_blobs);

#ifdef __cplusplus
}
#endif
#endif /* HDOG_H */
