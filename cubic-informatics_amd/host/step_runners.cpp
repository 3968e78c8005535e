// step_runners.cpp -- the reference's step-runner programs (test/CMakeLists.txt:3-48) on
// libhdog.so: each runs a prefix of the HDoG pipeline on one substack and dumps its
// intermediates as raw little-endian files, with the reference's CLIs and file names so
// that script/HDoG_intermediate.py drives them unchanged.
//
// One binary; the program name (argv[0]) selects the runner:
//   Normalize_test PARAM [ERO DIL NORM | NORM]   test/Normalize_test.cpp:68-128
//   DoG_test PARAM [NORM DOG | DOG]               test/DoG_test.cpp:73-138
//   Hessian_test PARAM [DOG PD [HESSIAN]]         test/Hessian_test.cpp:80-160
//   CCL_test PARAM                                test/CCL_test.cpp:82-142
//   Regional_test PARAM                           test/Regional_test.cpp:135-238
//   MaxNormalized_test PARAM [OUT]                test/MaxNormalized_test.cpp:102-173
//   AverageNormalized_test PARAM [OUT]            test/AverageNormalized_test.cpp:104-174
// (or `hdog_step_runner NAME PARAM ...`).  The parameter file is the HDoG3D_main one
// with exactly one stack of exactly HDoG_param.depth slices; outputs are written in the
// working directory unless named.  Exit codes: 0 ok, 2 usage / parameter / file error,
// 1 device error.  Differences from the reference: every runner uses devID (CCL_test
// always used device 0), labels are written as the uint32 labels of the library (the
// same bytes as the reference's int32 for volumes below 2^31 voxels), and the Gaussian
// truncation / precision keys of HDoG3D_main are honoured.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "hdog.h"
#include "hdog_params.hpp"

namespace {

struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

[[noreturn]] void die(int code, const std::string& msg) {
  std::cout << msg << std::endl;
  std::exit(code);
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) die(1, std::string(what) + ": " + hipGetErrorString(e));
}

void check(hdog_ctx* ctx, int rc, const char* what) {
  if (rc != HDOG_OK) die(rc == HDOG_ERR_HIP ? 1 : 2, std::string(what) + ": " + hdog_last_error(ctx));
}

void* dalloc(Dev& d, size_t bytes) {
  check_hip(hipMalloc(&d.p, bytes ? bytes : 1), "hipMalloc");
  return d.p;
}

template <typename T>
std::vector<T> download(const void* d, size_t n) {
  std::vector<T> h(n);
  check_hip(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return h;
}

// saveImage, src/3D/utils.h:175-185
template <typename T>
void save(const T* data, size_t n, const std::string& fname) {
  FILE* f = std::fopen(fname.c_str(), "wb");
  if (!f) die(2, "Unable to open " + fname);
  std::fwrite(data, sizeof(T), n, f);
  std::fclose(f);
}

struct Run {
  HostParams p;
  hdog_ctx* ctx = nullptr;
  size_t N = 0;
  Dev img, norm, dog, hess, pd, labels;
};

void print_dog_params(const HostParams& p) {
  std::cout << "sigma_xy1: " << p.sigma_xy1 << "\t"
            << "sigma_z1: " << p.sigma_z1 << "\t"
            << "sigma_xy2: " << p.sigma_xy2 << "\t"
            << "sigma_z2: " << p.sigma_z2 << std::endl;
}

void print_norm_params(const HostParams& p) {
  std::cout << "radius_large_xy: " << p.radius_xy << "\t"
            << "radius_large_z: " << p.radius_z << std::endl;
}

// common prologue of every runner: parameters, one stack of `depth` slices read to the
// host, then the device context and the upload
void setup(Run& r, const char* param_file, bool dog_params) {
  std::string why;
  if (!load_params(param_file, r.p, why)) die(2, std::string(param_file) + ": " + why);
  if (r.p.src_paths.size() != 1) die(2, "A single stack should be provided!");
  if ((int)r.p.src_paths[0].size() != r.p.depth)
    die(2, "depth(=" + std::to_string(r.p.src_paths[0].size()) + ") should be " +
               std::to_string(r.p.depth));
  const size_t plane = (size_t)r.p.width * r.p.height;
  r.N = plane * r.p.depth;
  std::vector<uint16_t> h(r.N);
  for (int z = 0; z < r.p.depth; ++z)
    if (!load_slice(r.p.src_paths[0][z], h.data() + plane * z, plane, why)) die(2, why);
  if (hdog_create(&r.ctx, r.p.devID) != HDOG_OK) die(1, "no HIP device " + std::to_string(r.p.devID));
  if (dog_params) {
    check(r.ctx, hdog_init_gaussian3d(r.ctx, r.p.sigma_xy1, r.p.sigma_z1, r.p.sigma_xy2,
                                      r.p.sigma_z2, r.p.truncate, nullptr),
          "initGaussian3DKernel");
    print_dog_params(r.p);
  }
  print_norm_params(r.p);
  dalloc(r.img, r.N * 2);
  check_hip(hipMemcpy(r.img.p, h.data(), r.N * 2, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void run_normalize(Run& r) {
  dalloc(r.norm, r.N * 4);
  check(r.ctx, hdog_normalize3d(r.ctx, nullptr, (const uint16_t*)r.img.p, (float*)r.norm.p,
                                (float)r.p.min_intensity, r.p.width, r.p.height, r.p.depth,
                                r.p.radius_xy, r.p.radius_z),
        "Normalize3DFilter");
}

void run_dog(Run& r) {
  run_normalize(r);
  dalloc(r.dog, r.N * 4);
  check(r.ctx, hdog_dog3d(r.ctx, nullptr, (const float*)r.norm.p, (float*)r.dog.p, r.p.width,
                          r.p.height, r.p.depth, r.p.gamma_n),
        "DoG3DFilter");
}

void run_hessian(Run& r) {
  run_dog(r);
  dalloc(r.hess, r.N * 24);
  dalloc(r.pd, r.N);
  check(r.ctx, hdog_hessian_pd(r.ctx, nullptr, (const float*)r.dog.p, (float*)r.hess.p,
                               (uint8_t*)r.pd.p, r.p.width, r.p.height, r.p.depth),
        "HessianPositiveDefiniteWithElement");
}

void run_ccl(Run& r) {
  run_hessian(r);
  dalloc(r.labels, r.N * 4);
  check(r.ctx, hdog_ccl3d(r.ctx, nullptr, (const uint8_t*)r.pd.p, (uint32_t*)r.labels.p,
                          r.p.width, r.p.height, r.p.depth),
        "CCL");
}

struct Regions {
  int64_t R = 0;
  std::vector<uint32_t> label;
  std::vector<uint16_t> size;
  std::vector<float> cx, cy, cz, st, bl, nrm;
};

Regions run_regions(Run& r, bool ave) {
  run_ccl(r);
  Regions g;
  check(r.ctx, hdog_regional_features(r.ctx, nullptr, (const uint32_t*)r.labels.p,
                                      (const float*)r.hess.p, (const float*)r.norm.p,
                                      (const uint16_t*)r.img.p, r.p.width, r.p.height, r.p.depth,
                                      ave ? 1 : 0, &g.R),
        "RegionalFeatures");
  std::cout << "num_regions: " << g.R << std::endl;
  const size_t R = (size_t)g.R;
  for (auto* v : {&g.cx, &g.cy, &g.cz, &g.st, &g.bl, &g.nrm}) v->resize(R);
  g.label.resize(R);
  g.size.resize(R);
  hdog_regions o{};
  o.capacity = g.R;
  o.label = g.label.data();
  o.size = g.size.data();
  o.centroid_x = g.cx.data();
  o.centroid_y = g.cy.data();
  o.centroid_z = g.cz.data();
  o.structureness = g.st.data();
  o.blobness = g.bl.data();
  o.max_normalized = g.nrm.data();
  check(r.ctx, hdog_regions_download(r.ctx, nullptr, &o), "download");
  return g;
}

// region arrays without region 0 (the background), as the reference saves them
template <typename T>
void save_regions(const std::vector<T>& v, const std::string& fname) {
  save(v.data() + (v.empty() ? 0 : 1), v.empty() ? 0 : v.size() - 1, fname);
}

int normalize_test(int argc, char** argv) {
  Run r;
  setup(r, argv[1], false);
  dalloc(r.norm, r.N * 4);
  Dev ero, dil;
  if (r.p.verbose) {
    dalloc(ero, r.N * 4);
    dalloc(dil, r.N * 4);
    check(r.ctx, hdog_normalize3d_ed(r.ctx, nullptr, (const uint16_t*)r.img.p, (float*)r.norm.p,
                                     (float*)ero.p, (float*)dil.p, (float)r.p.min_intensity,
                                     r.p.width, r.p.height, r.p.depth, r.p.radius_xy,
                                     r.p.radius_z),
          "Normalize3DFilter");
  } else {
    check(r.ctx, hdog_normalize3d(r.ctx, nullptr, (const uint16_t*)r.img.p, (float*)r.norm.p,
                                  (float)r.p.min_intensity, r.p.width, r.p.height, r.p.depth,
                                  r.p.radius_xy, r.p.radius_z),
          "Normalize3DFilter");
  }
  const auto norm = download<float>(r.norm.p, r.N);
  if (r.p.verbose) {
    const auto e = download<float>(ero.p, r.N), d = download<float>(dil.p, r.N);
    const bool named = argc == 5;
    save(e.data(), r.N, named ? argv[2] : "./erosion_l.bin");
    save(d.data(), r.N, named ? argv[3] : "./dilation_l.bin");
    save(norm.data(), r.N, named ? argv[4] : (argc == 3 ? argv[2] : "./normalized.bin"));
  } else {
    save(norm.data(), r.N, argc == 3 ? argv[2] : "./normalized.bin");
  }
  hdog_destroy(r.ctx);
  return 0;
}

int dog_test(int argc, char** argv) {
  Run r;
  setup(r, argv[1], true);
  run_dog(r);
  const auto dog = download<float>(r.dog.p, r.N);
  if (argc == 3) {
    save(dog.data(), r.N, argv[2]);
  } else if (argc == 4) {
    save(download<float>(r.norm.p, r.N).data(), r.N, argv[2]);
    save(dog.data(), r.N, argv[3]);
  } else {
    save(dog.data(), r.N, "./dog.bin");
    if (r.p.verbose) save(download<float>(r.norm.p, r.N).data(), r.N, "./normalized.bin");
  }
  hdog_destroy(r.ctx);
  return 0;
}

int hessian_test(int argc, char** argv) {
  Run r;
  setup(r, argv[1], true);
  run_hessian(r);
  const auto dog = download<float>(r.dog.p, r.N);
  const auto pd = download<uint8_t>(r.pd.p, r.N);
  if (r.p.verbose) {
    const auto h = download<float>(r.hess.p, r.N * 6);
    if (argc == 5 || argc == 4) {
      save(dog.data(), r.N, argv[2]);
      save(pd.data(), r.N, argv[3]);
      save(h.data(), r.N * 6, argc == 5 ? argv[4] : "hessian.bin");
    }
  } else if (argc == 4) {
    save(dog.data(), r.N, argv[2]);
    save(pd.data(), r.N, argv[3]);
  } else {
    save(dog.data(), r.N, "dog.bin");
    save(pd.data(), r.N, "hessian_pd.bin");
  }
  hdog_destroy(r.ctx);
  return 0;
}

int ccl_test(int, char** argv) {
  Run r;
  setup(r, argv[1], true);
  run_ccl(r);
  save(download<uint32_t>(r.labels.p, r.N).data(), r.N, "labels.bin");
  save(download<uint8_t>(r.pd.p, r.N).data(), r.N, "hessian_pd.bin");
  hdog_destroy(r.ctx);
  return 0;
}

int regional_test(int, char** argv) {
  Run r;
  setup(r, argv[1], true);
  const Regions g = run_regions(r, false);
  std::cout << "Saving..." << std::endl;
  save(download<uint32_t>(r.labels.p, r.N).data(), r.N, "labels.bin");
  save(download<float>(r.norm.p, r.N).data(), r.N, "normalized.bin");
  save_regions(g.label, "labels_region.bin");
  save_regions(g.size, "size_region.bin");
  save_regions(g.nrm, "max_normalized_region.bin");
  save_regions(g.st, "structureness_region.bin");
  save_regions(g.bl, "blobness_region.bin");
  save_regions(g.cx, "centroid_x.bin");
  save_regions(g.cy, "centroid_y.bin");
  save_regions(g.cz, "centroid_z.bin");
  hdog_destroy(r.ctx);
  return 0;
}

int max_normalized_test(int argc, char** argv) {
  Run r;
  setup(r, argv[1], true);
  const Regions g = run_regions(r, false);
  save_regions(g.nrm, argc == 3 ? argv[2] : "max_normalized_region.bin");
  hdog_destroy(r.ctx);
  return 0;
}

// AverageNormalized, src/3D/RegionalFeatures.cu:257-291: the FP32 sum of normalized
// intensity divided by the u16 region size
int average_normalized_test(int argc, char** argv) {
  Run r;
  setup(r, argv[1], true);
  Regions g = run_regions(r, true);
  for (size_t i = 0; i < g.nrm.size(); ++i) g.nrm[i] = g.nrm[i] / (float)g.size[i];
  save_regions(g.nrm, argc == 3 ? argv[2] : "average_normalized_region.bin");
  hdog_destroy(r.ctx);
  return 0;
}

struct Runner {
  const char* name;
  int (*fn)(int, char**);
};
const Runner kRunners[] = {{"Normalize_test", normalize_test},
                           {"DoG_test", dog_test},
                           {"Hessian_test", hessian_test},
                           {"CCL_test", ccl_test},
                           {"Regional_test", regional_test},
                           {"MaxNormalized_test", max_normalized_test},
                           {"AverageNormalized_test", average_normalized_test}};

}  // namespace

int main(int argc, char** argv) {
  std::string name = argv[0];
  const size_t slash = name.find_last_of('/');
  if (slash != std::string::npos) name = name.substr(slash + 1);
  if (name == "hdog_step_runner" && argc >= 2) {  // hdog_step_runner NAME PARAM ...
    name = argv[1];
    ++argv;
    --argc;
  }
  for (const Runner& r : kRunners)
    if (name == r.name) {
      if (argc < 2) die(2, "usage: " + name + " PARAM_FILE [outputs...]");
      return r.fn(argc, argv);
    }
  std::string names;
  for (const Runner& r : kRunners) names += std::string(" ") + r.name;
  die(2, "unknown step runner '" + name + "'; one of:" + names);
}
