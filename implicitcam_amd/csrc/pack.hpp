// Host packer: the part (include/icam_part.h) as the record tables the kernels read.
#pragma once
#include <string>
#include <vector>

#include "../../include/icam_part.h"
#include "layout.hpp"

namespace icam {

struct Packed {
  Layout L{};
  std::vector<double> F;   // float64 records (the float32 copy is made from these)
  std::vector<uint32_t> U; // index records
  std::vector<float> pbox; // 6 per primitive (lo, hi), rounded outward
  std::vector<uint32_t> pcode;
  // the primitives' hierarchy (Layout::bvh_*): boxes, the sorted primitives, and per node
  // whether a boundary point lies in it
  std::vector<float> bvh_box;
  std::vector<uint32_t> bvh_prim;
  std::vector<uint8_t> bvh_holds;
  double origin[3] = {0, 0, 0};
  size_t nodes = 0;        // sub-patches and curve pieces made
  std::string error;
};

// Exact::gpu_pack(origin, need) (crates/icam-core/src/solid3/gpu.rs:259-701): false
// with `error` set when the part cannot be packed.
bool pack_part(const icam_part_desc* part, double need, Packed& out);

// The float32 copy of the float64 records.
std::vector<float> to_f32(const std::vector<double>& f);

}  // namespace icam
