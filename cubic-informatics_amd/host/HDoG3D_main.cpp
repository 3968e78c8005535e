// HDoG3D_main.cpp -- drop-in replacement of the reference's `HDoG3D_main PARAM_FILE`
// (src/3D/HDoG3D_main.cpp:260-319), driving the gfx950 kernels through libhdog.so.
//
// Same CLI, parameter JSON, slice files and 28-byte candidate records, so
// script/HDoG_gpu.py runs it unchanged (one process per GPU, device from devID).
// Per stack: the result file is removed, then each overlapping z-substack is processed
// and its candidates appended.  Substacks advance by depth - 2*margin with
// margin = r_z(sigma2) + extra_depth_margin; the last one is zero-filled to `depth`.
//
// Differences from the reference, all additive:
//   * the next substack's slices are read on a loader thread while the GPU works on the
//     current one (the reference reads serially, src/3D/HDoG3D_main.cpp:155-157);
//   * a substack step depth - 2*margin <= 0 is rejected (the reference loops forever);
//   * a missing parameter key exits with code 2 and a message (the reference aborts);
//   * optional HDoG_param.truncate / HDoG_param.precision keys.
// Exit codes: 0 ok, 2 usage / parameter / file error, 1 device error.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "hdog.h"
#include "hdog_params.hpp"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

struct RegionHost {
  std::vector<float> cx, cy, cz, st, bl, mx;
  std::vector<uint16_t> sz;
  hdog_regions view(int64_t cap) {
    if ((int64_t)cx.size() < cap) {
      for (auto* v : {&cx, &cy, &cz, &st, &bl, &mx}) v->resize(cap);
      sz.resize(cap);
    }
    hdog_regions r{};
    r.capacity = (int64_t)cx.size();
    r.centroid_x = cx.data();
    r.centroid_y = cy.data();
    r.centroid_z = cz.data();
    r.structureness = st.data();
    r.blobness = bl.data();
    r.max_normalized = mx.data();
    r.size = sz.data();
    return r;
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cout << "Usage:" << std::endl << argv[0] << " PARAM_FILE" << std::endl;
    return 2;
  }
  HostParams p;
  std::string why;
  if (!load_params(argv[1], p, why)) {
    std::cout << why << std::endl;
    return 2;
  }
  std::cout << "min_intensity_skip:" << p.min_intensity_skip << std::endl;
  std::cout << "min_intensity:" << p.min_intensity << std::endl;
  std::cout << "min_size:" << p.min_size << std::endl;
  std::cout << "number of stacks: " << p.src_paths.size() << std::endl;

  auto t0 = Clock::now();
  hdog_ctx* ctx = nullptr;
  if (hdog_create(&ctx, p.devID) != HDOG_OK) {
    std::cout << "Unable to initialise HIP device " << p.devID << std::endl;
    return 1;
  }
  int r_z2 = 0;
  if (hdog_init_gaussian3d(ctx, p.sigma_xy1, p.sigma_z1, p.sigma_xy2, p.sigma_z2, p.truncate,
                           &r_z2) != HDOG_OK) {
    std::cout << hdog_last_error(ctx) << std::endl;
    hdog_destroy(ctx);
    return 2;
  }
  p.depth_margin = r_z2 + p.extra_depth_margin;
  std::cout << "Initialization took " << ms_since(t0) << "msec" << std::endl;
  std::cout << "sigma_xy1: " << p.sigma_xy1 << "\tsigma_z1: " << p.sigma_z1
            << "\tsigma_xy2: " << p.sigma_xy2 << "\tsigma_z2: " << p.sigma_z2 << std::endl;
  std::cout << "radius_large_xy: " << p.radius_xy << "\tradius_large_z: " << p.radius_z
            << std::endl;
  std::cout << "depth_margin: " << p.depth_margin << std::endl;
  const int step = p.depth - 2 * p.depth_margin;
  if (step <= 0) {
    std::cout << "HDoG_param.depth (" << p.depth << ") must exceed 2*depth_margin ("
              << 2 * p.depth_margin << ")" << std::endl;
    hdog_destroy(ctx);
    return 2;
  }

  hdog_params hp{};
  hp.width = p.width;
  hp.height = p.height;
  hp.depth = p.depth;
  hp.radius_xy = p.radius_xy;
  hp.radius_z = p.radius_z;
  hp.min_intensity = (float)p.min_intensity;
  hp.sigma_xy1 = p.sigma_xy1;
  hp.sigma_z1 = p.sigma_z1;
  hp.sigma_xy2 = p.sigma_xy2;
  hp.sigma_z2 = p.sigma_z2;
  hp.truncate = p.truncate;
  hp.gamma = p.gamma_n;
  hp.is_ave_mode = p.is_ave_mode ? 1 : 0;
  hp.precision = precision_code(p.precision);

  const size_t plane = (size_t)p.width * p.height;
  std::vector<uint16_t> buf[2] = {std::vector<uint16_t>(plane * p.depth),
                                  std::vector<uint16_t>(plane * p.depth)};
  RegionHost reg;
  int rc = 0;

  for (size_t is = 0; is < p.src_paths.size() && rc == 0; ++is) {
    const auto& src = p.src_paths[is];
    const int L = (int)src.size();
    if (std::remove(p.dst_paths[is].c_str()) == 0)
      std::cout << "removed previous result " << is << std::endl;
    else
      std::cout << "newly create result " << is << std::endl;
    std::cout << "length: " << L << std::endl;

    std::vector<int> z0s;
    for (int z0 = 0; z0 < L - p.depth_margin; z0 += step) z0s.push_back(z0);

    // loader: fills buf[slot] with the slices of substack k; reports failure in `err`
    auto load = [&](size_t k, int slot, std::string* err, double* took) {
      auto t = Clock::now();
      const int z0 = z0s[k];
      const int d = std::min(p.depth, L - z0);
      for (int i = 0; i < d; ++i)
        if (!load_slice(src[z0 + i], buf[slot].data() + plane * i, plane, *err)) return;
      *took = ms_since(t);
    };
    std::string lerr;
    double ltook = 0;
    if (!z0s.empty()) load(0, 0, &lerr, &ltook);
    for (size_t k = 0; k < z0s.size(); ++k) {
      const int slot = (int)(k & 1);
      if (!lerr.empty()) {
        std::cout << lerr << std::endl;
        rc = 2;
        break;
      }
      const int z0 = z0s[k];
      const int d = std::min(p.depth, L - z0);
      std::cout << "----------" << std::endl;
      std::cout << "processSubstack(z0=" << z0 << ", depth=" << d << ")" << std::endl;
      std::cout << "\tfirst_src_path: " << src[z0] << std::endl;
      std::cout << "\tlast_src_path: " << src[z0 + d - 1] << std::endl;
      std::cout << "loadImage took " << ltook << "msec" << std::endl;
      // prefetch the next substack while this one runs on the GPU
      std::string nerr;
      double ntook = 0;
      std::thread pre;
      if (k + 1 < z0s.size()) pre = std::thread(load, k + 1, 1 - slot, &nerr, &ntook);

      auto t = Clock::now();
      // one extra capacity probe: process, then size the host arrays to the region count
      hdog_regions probe{};
      int e = hdog_process_substack(ctx, nullptr, &hp, buf[slot].data(), d, &probe);
      if (e == HDOG_ERR_CAPACITY) {
        hdog_regions v = reg.view(probe.count);
        e = hdog_regions_download(ctx, nullptr, &v);
        probe.count = v.count;
      } else if (e == HDOG_OK) {
        hdog_regions v = reg.view(probe.count);
        e = hdog_regions_download(ctx, nullptr, &v);
      }
      if (e != HDOG_OK) {
        std::cout << "HDoG failed: " << hdog_last_error(ctx) << std::endl;
        rc = e == HDOG_ERR_ARG ? 2 : 1;
        if (pre.joinable()) pre.join();
        break;
      }
      const int64_t R = probe.count;
      std::cout << "num_regions: " << R << std::endl;
      std::cout << "GPU memcpy & kernel execution took " << ms_since(t) << "msec" << std::endl;
      t = Clock::now();
      if (save_features(p.dst_paths[is], R, reg.cx.data(), reg.cy.data(), reg.cz.data(),
                        reg.st.data(), reg.bl.data(), reg.mx.data(), reg.sz.data(), d, z0, p) < 0) {
        std::cout << "Unable to open " << p.dst_paths[is] << std::endl;
        rc = 2;
      }
      if (d == p.depth)
        std::cout << "[save] middle of the stack : " << z0 + p.depth_margin << " - "
                  << z0 + d - p.depth_margin - 1 << " (" << d - 2 * p.depth_margin << ")"
                  << std::endl;
      else
        std::cout << "[save] end of the stack : " << z0 + p.depth_margin << " - " << z0 + d - 1
                  << " (" << d - p.depth_margin << ")" << std::endl;
      std::cout << "saveImage took " << ms_since(t) << "msec" << std::endl;
      if (pre.joinable()) pre.join();
      lerr = nerr;
      ltook = ntook;
    }
  }
  hdog_destroy(ctx);
  return rc;
}
