// hdog_params.hpp -- the HDoG parameter-file schema, slice loader and candidate-record
// writer of the host program.
//
// Schema and defaults: src/3D/utils.h:48-157 of the reference.  Required keys are the
// ones the reference reads with at(); optional keys are `verbose`,
// HDoG_param.cub_tmp_size_factor (accepted, unused: there is no CUB here),
// HDoG_param.min_intensity_skip (read, unused as in the reference) and top-level
// is_ave_mode.  New optional keys, ignored by the reference's reader:
// HDoG_param.truncate (Gaussian truncation, default 2.0 = the reference's constant) and
// HDoG_param.precision ("fp32" default, "fp16", "bf16").
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "hdog.h"
#include "json.hpp"

struct HostParams {
  int devID = 0;
  bool verbose = false;
  int width = 2560, height = 2160, depth = 32;
  int left_margin = 100, right_margin = 100, top_margin = 100, bottom_margin = 100;
  int depth_margin = 0;  // r_z(sigma2) + extra_depth_margin, set after the Gaussian init
  int extra_depth_margin = 3;
  float scale_x = 1.0f, scale_y = 1.0f, scale_z = 1.0f;
  float gamma_n = 1.0f;
  float sigma_xy1 = 0, sigma_z1 = 0, sigma_xy2 = 0, sigma_z2 = 0;
  int radius_xy = 0, radius_z = 0;
  unsigned short min_intensity_skip = 1000;
  unsigned short min_intensity = 1000;
  bool is_ave_mode = false;
  unsigned short min_size = 0;
  double truncate = 2.0;
  std::string precision = "fp32";
  std::vector<std::vector<std::string>> src_paths;
  std::vector<std::string> dst_paths;
};

// enum hdog_precision of an HDoG_param.precision string, -1 when unknown
inline int precision_code(const std::string& s) {
  if (s == "fp32") return HDOG_FP32;
  if (s == "fp16") return HDOG_FP16;
  if (s == "bf16") return HDOG_BF16;
  if (s == "fp16_dog32") return HDOG_FP16_DOG32;
  if (s == "bf16_dog32") return HDOG_BF16_DOG32;
  return -1;
}

// Exit code 2 for every parameter-file problem (the reference exits 2 when the file
// cannot be opened, src/3D/utils.h:96-100, and aborts on a missing key).
inline bool load_params(const std::string& fname, HostParams& p, std::string& why) {
  std::ifstream f(fname);
  if (!f.is_open()) { why = "Unable to open file."; return false; }
  std::stringstream ss;
  ss << f.rdbuf();
  try {
    const hjson::Value j = hjson::parse(ss.str());
    p.devID = j.at("devID").integer();
    if (j.has("verbose")) p.verbose = j.at("verbose").boolean();
    const auto& im = j.at("input_image_info");
    p.width = im.at("width").integer();
    p.height = im.at("height").integer();
    p.left_margin = im.at("left_margin").integer();
    p.right_margin = im.at("right_margin").integer();
    p.top_margin = im.at("top_margin").integer();
    p.bottom_margin = im.at("bottom_margin").integer();
    const auto& co = j.at("coordinate_info");
    p.scale_x = (float)co.at("scale_x").number();
    p.scale_y = (float)co.at("scale_y").number();
    p.scale_z = (float)co.at("scale_z").number();
    const auto& hp = j.at("HDoG_param");
    p.depth = hp.at("depth").integer();
    p.extra_depth_margin = hp.at("extra_depth_margin").integer();
    if (hp.has("min_intensity_skip"))
      p.min_intensity_skip = (unsigned short)hp.at("min_intensity_skip").integer();
    const auto& rn = hp.at("radius_norm");
    p.radius_xy = rn.at("xy_large").integer();
    p.radius_z = rn.at("z_large").integer();
    p.min_intensity = (unsigned short)hp.at("min_intensity").integer();
    p.gamma_n = (float)hp.at("gamma_n").number();
    const auto& sd = hp.at("sigma_dog");
    p.sigma_xy1 = (float)sd.at("xy_small").number();
    p.sigma_z1 = (float)sd.at("z_small").number();
    p.sigma_xy2 = (float)sd.at("xy_large").number();
    p.sigma_z2 = (float)sd.at("z_large").number();
    p.min_size = (unsigned short)hp.at("min_size").integer();
    if (hp.has("truncate")) p.truncate = hp.at("truncate").number();
    if (hp.has("precision")) p.precision = hp.at("precision").string();
    if (j.has("is_ave_mode")) p.is_ave_mode = j.at("is_ave_mode").boolean();
    for (const auto& st : j.at("stacks").array()) {
      std::vector<std::string> paths;
      for (const auto& s : st.at("src_paths").array()) paths.push_back(s.string());
      p.src_paths.push_back(paths);
      p.dst_paths.push_back(st.at("dst_path").string());
    }
  } catch (const hjson::JsonError& e) {
    why = e.what();
    return false;
  }
  if (p.width < 1 || p.height < 1 || p.depth < 1) { why = "non-positive image shape"; return false; }
  if (precision_code(p.precision) < 0) {
    why = "HDoG_param.precision must be fp32, fp16, bf16, fp16_dog32 or bf16_dog32";
    return false;
  }
  return true;
}

// src/3D/utils.h:160-173: one fread of W*H uint16; only a read of zero items is an error.
inline bool load_slice(const std::string& fname, uint16_t* dst, size_t n, std::string& why) {
  FILE* f = std::fopen(fname.c_str(), "rb");
  if (!f) { why = "Unable to open " + fname; return false; }
  const size_t got = std::fread(dst, sizeof(uint16_t), n, f);
  std::fclose(f);
  if (got == 0) { why = "Unable to read " + fname; return false; }
  return true;
}

// 28-byte record consumed downstream (src/3D/utils.h:188-197; script/HalfBrainCells.py dt_local)
#pragma pack(push, 1)
struct FeatureVector {
  float centroid_x, centroid_y, centroid_z;
  float structureness, blobness, max_normalized;
  unsigned short size, padding;
};
#pragma pack(pop)
static_assert(sizeof(FeatureVector) == 28, "FeatureVector must stay 28 bytes (6 f4 + 2 u2)");

// saveFeatureVector, src/3D/utils.h:204-250: skip region 0, drop regions whose centroid
// lies in a margin (depth = slices actually loaded in this substack), smaller than
// min_size, or with max_normalized <= 0; append the rest.  Returns the saved count or -1.
inline long save_features(const std::string& fname, int64_t R, const float* cx, const float* cy,
                          const float* cz, const float* st, const float* bl, const float* mx,
                          const uint16_t* sz, int depth, int z0, const HostParams& p) {
  FILE* f = std::fopen(fname.c_str(), "ab");
  if (!f) return -1;
  std::vector<FeatureVector> out;
  const float min_maxnorm = 0.0f;
  for (int64_t i = 1; i < R; ++i) {
    if (cz[i] < p.depth_margin || cz[i] >= depth - p.depth_margin) continue;
    if (cx[i] < p.left_margin || cx[i] >= p.width - p.right_margin) continue;
    if (cy[i] < p.top_margin || cy[i] >= p.height - p.bottom_margin) continue;
    if (sz[i] < p.min_size || mx[i] <= min_maxnorm) continue;
    out.push_back(FeatureVector{cx[i], cy[i], cz[i] + z0, st[i], bl[i], mx[i], sz[i], 0});
  }
  if (!out.empty()) std::fwrite(out.data(), sizeof(FeatureVector), out.size(), f);
  std::fclose(f);
  std::cout << "saved count:" << out.size() << "*" << sizeof(FeatureVector) << "="
            << out.size() * sizeof(FeatureVector) << std::endl;
  return (long)out.size();
}
