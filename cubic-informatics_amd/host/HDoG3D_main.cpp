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
//   * slices are read into page-locked buffers by several threads, one substack ahead,
//     and each substack's host-to-device copy runs on a copy stream while the previous
//     substack's kernels run (the reference reads, copies and computes serially,
//     src/3D/HDoG3D_main.cpp:155-170); the substacks of all stacks form one pipeline, so
//     the next stack's first substack is read while the last one of a stack runs, and the
//     records are written by a writer thread while the next substack runs;
//   * consecutive substacks of a stack share all but depth - 2*margin slices: those are
//     copied on the device from the previous substack's upload, and only the new slices
//     are read from files and copied over the link;
//   * the region arrays arrive in page-locked buffers (a direct copy from the device
//     rather than one staged through the runtime's own page-locked memory);
//   * a substack step depth - 2*margin <= 0 is rejected (the reference loops forever);
//   * a missing parameter key exits with code 2 and a message (the reference aborts);
//   * optional HDoG_param.truncate / HDoG_param.precision keys.
// Exit codes: 0 ok, 2 usage / parameter / file error, 1 device error.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
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

// The region arrays of one substack, in one page-locked block that grows when a substack
// has more regions than it holds (freed when the last owner lets go of it).
struct RegionHost {
  std::shared_ptr<void> block;
  int64_t cap = 0;
  float *cx = nullptr, *cy = nullptr, *cz = nullptr, *st = nullptr, *bl = nullptr,
        *mx = nullptr;
  uint16_t* sz = nullptr;
  bool reserve(hdog_ctx* ctx, int64_t need) {
    if (need <= cap) return true;
    const int64_t n = need + need / 4 + 1024;  // the next substacks' counts vary a little
    const size_t fb = ((size_t)n * sizeof(float) + 255) / 256 * 256;
    void* p = nullptr;
    if (hdog_host_alloc(ctx, fb * 6 + (size_t)n * sizeof(uint16_t), &p) != HDOG_OK) return false;
    block = std::shared_ptr<void>(p, [ctx](void* q) { hdog_host_free(ctx, q); });
    char* c = static_cast<char*>(p);
    float** f[6] = {&cx, &cy, &cz, &st, &bl, &mx};
    for (int k = 0; k < 6; ++k) *f[k] = reinterpret_cast<float*>(c + k * fb);
    sz = reinterpret_cast<uint16_t*>(c + 6 * fb);
    cap = n;
    return true;
  }
  hdog_regions view() const {
    hdog_regions r{};
    r.capacity = cap;
    r.centroid_x = cx;
    r.centroid_y = cy;
    r.centroid_z = cz;
    r.structureness = st;
    r.blobness = bl;
    r.max_normalized = mx;
    r.size = sz;
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
  // every substack of every stack, in order: one pipeline over the whole parameter file
  // keep: leading slices shared with the previous substack of the same stack (they are
  // copied from its device slot, from its slice `from`), so only d - keep are read
  struct Item {
    size_t is;
    int z0, d, from, keep;
  };
  std::vector<Item> items;
  for (size_t is = 0; is < p.src_paths.size(); ++is) {
    const int L = (int)p.src_paths[is].size();
    for (int z0 = 0; z0 < L - p.depth_margin; z0 += step) {
      Item it{is, z0, std::min(p.depth, L - z0), 0, 0};
      if (!items.empty() && items.back().is == is) {
        const Item& pr = items.back();
        it.from = z0 - pr.z0;
        it.keep = std::max(0, std::min(pr.d - it.from, it.d));
      }
      items.push_back(it);
    }
  }
  // three page-locked substack buffers: one being read from disk, one being copied to the
  // device, one whose copy the running substack consumed
  uint16_t* buf[3] = {nullptr, nullptr, nullptr};
  for (auto& b : buf)
    if (hdog_host_alloc(ctx, plane * p.depth * sizeof(uint16_t), (void**)&b) != HDOG_OK) {
      std::cout << "Unable to allocate host buffers: " << hdog_last_error(ctx) << std::endl;
      for (auto& q : buf) hdog_host_free(ctx, q);
      hdog_destroy(ctx);
      return 1;
    }
  const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  int rc = 0;

  // writer: announces each stack (its result file removed) and appends the records of
  // each substack, in order, on its own thread, so the next substack's kernels run
  // meanwhile.  At most two substacks' regions wait for it.
  struct Job {
    size_t is = 0;
    int z0 = 0, d = 0;
    int64_t R = -1;  // -1: announce stack `is`
    RegionHost reg;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> queue;
  std::vector<RegionHost> spare;  // region arrays returned by the writer
  bool closing = false, write_failed = false;
  std::thread writer([&] {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return closing || !queue.empty(); });
        if (queue.empty()) return;
        j = std::move(queue.front());
        queue.pop_front();
      }
      std::ostringstream out;
      if (j.R < 0) {
        if (std::remove(p.dst_paths[j.is].c_str()) == 0)
          out << "removed previous result " << j.is << "\n";
        else
          out << "newly create result " << j.is << "\n";
        out << "length: " << p.src_paths[j.is].size() << "\n";
      } else {
        auto t = Clock::now();
        bool ok = save_features(p.dst_paths[j.is], j.R, j.reg.cx, j.reg.cy, j.reg.cz,
                                j.reg.st, j.reg.bl, j.reg.mx, j.reg.sz, j.d, j.z0, p) >= 0;
        if (!ok) out << "Unable to open " << p.dst_paths[j.is] << "\n";
        if (j.d == p.depth)
          out << "[save] middle of the stack : " << j.z0 + p.depth_margin << " - "
              << j.z0 + j.d - p.depth_margin - 1 << " (" << j.d - 2 * p.depth_margin << ")\n";
        else
          out << "[save] end of the stack : " << j.z0 + p.depth_margin << " - " << j.z0 + j.d - 1
              << " (" << j.d - p.depth_margin << ")\n";
        out << "saveImage took " << ms_since(t) << "msec\n";
        std::lock_guard<std::mutex> lk(mu);
        if (!ok) write_failed = true;
        spare.push_back(std::move(j.reg));
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        std::cout << out.str() << std::flush;
      }
      cv.notify_all();
    }
  });
  auto post = [&](Job&& j) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] {
      size_t saves = 0;
      for (const Job& q : queue) saves += q.R >= 0;
      return saves < 2;
    });
    queue.push_back(std::move(j));
    cv.notify_all();
  };
  size_t announced = 0;  // stacks announced to the writer so far
  auto announce_upto = [&](size_t is) {
    for (; announced <= is && announced < p.src_paths.size(); ++announced) {
      Job j;
      j.is = announced;
      post(std::move(j));
    }
  };

  // loader: reads the slices of item k into buf[k % 3] on `nthreads` threads
  struct Load {
    std::string err;
    double took = 0;
    std::thread th;
  };
  auto load = [&](size_t k, Load* ld) {
    auto t = Clock::now();
    const Item& it = items[k];
    const auto& src = p.src_paths[it.is];
    uint16_t* dst = buf[k % 3];
    std::vector<std::string> errs(nthreads);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w)
      pool.emplace_back([&, w] {
        for (int i = (int)w; i < it.d - it.keep; i += (int)nthreads)
          if (!load_slice(src[it.z0 + it.keep + i], dst + plane * i, plane, errs[w])) return;
      });
    for (auto& th : pool) th.join();
    for (auto& e : errs)
      if (!e.empty() && ld->err.empty()) ld->err = e;
    ld->took = ms_since(t);
  };
  auto upload = [&](size_t k) {
    const Item& it = items[k];
    if (it.keep > 0)  // the previous substack is in the other slot
      return hdog_substack_upload_overlap(ctx, &hp, it.from, it.keep, buf[k % 3],
                                          it.d - it.keep, (int)(k & 1));
    return hdog_substack_upload(ctx, &hp, buf[k % 3], it.d, (int)(k & 1));
  };
  auto say = [&](const std::string& s) {
    std::lock_guard<std::mutex> lk(mu);
    std::cout << s << std::flush;
  };
  const size_t nz = items.size();
  Load cur, next;
  if (nz > 0) {
    load(0, &cur);
    if (cur.err.empty() && upload(0) != HDOG_OK) cur.err = hdog_last_error(ctx);
  }
  if (nz > 1) next.th = std::thread(load, 1, &next);
  for (size_t k = 0; k < nz; ++k) {
    const Item& it = items[k];
    announce_upto(it.is);
    if (!cur.err.empty()) {
      say(cur.err + "\n");
      rc = 2;
      break;
    }
    const auto& src = p.src_paths[it.is];
    {
      std::ostringstream out;
      out << "----------\n"
          << "processSubstack(z0=" << it.z0 << ", depth=" << it.d << ")\n"
          << "\tfirst_src_path: " << src[it.z0] << "\n"
          << "\tlast_src_path: " << src[it.z0 + it.d - 1] << "\n"
          << "loadImage took " << cur.took << "msec\n";
      say(out.str());
    }
    // item k+1: its slices are read (joined here) and its copy starts before k runs, so
    // the copy overlaps k's kernels; item k+2 is read meanwhile
    Load after;
    if (k + 1 < nz) {
      next.th.join();
      if (next.err.empty() && upload(k + 1) != HDOG_OK) next.err = hdog_last_error(ctx);
      if (k + 2 < nz) after.th = std::thread(load, k + 2, &after);
    }
    auto t = Clock::now();
    Job j;
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!spare.empty()) {
        j.reg = std::move(spare.back());
        spare.pop_back();
      }
    }
    // the first download is a capacity probe that returns the region count
    hdog_regions probe{};
    int e = hdog_substack_run(ctx, nullptr, &hp, (int)(k & 1), &probe);
    if (e == HDOG_ERR_CAPACITY || e == HDOG_OK) {
      if (j.reg.reserve(ctx, probe.count)) {
        hdog_regions v = j.reg.view();
        e = hdog_regions_download(ctx, nullptr, &v);
      } else {
        e = HDOG_ERR_HIP;
      }
    }
    if (e != HDOG_OK) {
      say(std::string("HDoG failed: ") + hdog_last_error(ctx) + "\n");
      rc = e == HDOG_ERR_ARG ? 2 : 1;
      if (after.th.joinable()) after.th.join();
      break;
    }
    {
      std::ostringstream out;
      out << "num_regions: " << probe.count << "\n"
          << "GPU memcpy & kernel execution took " << ms_since(t) << "msec\n";
      say(out.str());
    }
    j.is = it.is;
    j.z0 = it.z0;
    j.d = it.d;
    j.R = probe.count;
    post(std::move(j));
    cur.err = std::move(next.err);
    cur.took = next.took;
    next.err = std::move(after.err);
    next.took = after.took;
    next.th = std::move(after.th);
    {
      std::lock_guard<std::mutex> lk(mu);
      if (write_failed) rc = 2;
    }
    if (rc) break;
  }
  if (next.th.joinable()) next.th.join();
  if (rc == 0) announce_upto(p.src_paths.size());  // stacks too short for any substack
  {
    std::lock_guard<std::mutex> lk(mu);
    closing = true;
  }
  cv.notify_all();
  writer.join();
  if (write_failed && rc == 0) rc = 2;
  spare.clear();  // page-locked region arrays, before the context goes
  for (auto& b : buf) hdog_host_free(ctx, b);
  hdog_destroy(ctx);
  return rc;
}
