// Device work plan of a batch of pn J-scheme H*v instances.
//
// The vectors of all instances are concatenated (instance i occupies
// [vbase_i, vbase_i + N_i)).  Two kernels make one H*v:
//   1. pn_partial: one workgroup per (item, column tile); an item is one
//      (block L, a-side J') pair.  It forms Y = sum_R r V_R Bhat^T for each
//      a-density and accumulates A Y into a private partial tile (no atomics).
//   2. reduce: one workgroup per (block L, column tile); sums the item
//      partials of L in a fixed order, adds H_a V_L + V_L H_b^T, writes W_L.
#pragma once
#include <cstdint>
#include <vector>

#include "nsx_model.hpp"

namespace nsx {

constexpr int kThreads = 256;
constexpr int kMaxE = 8;                 // outputs per thread in one tile
constexpr int kTileElems = kThreads * kMaxE;
constexpr int kMaxLdsDoubles = 8192;     // Y tile capacity (64 KiB)

struct DItem {       // one (L, Ja') partial, one column tile
  int64_t part_off;  // partial tile base (full da x db layout of the item)
  int32_t da, dar, db, j0, nc;
  int32_t a_begin, a_end;
  int32_t pad;
};
struct DATerm {
  int64_t a_off;     // dense A (da x dar, column-major)
  int32_t r_begin, r_end;
};
struct DRTerm {
  int64_t v_off;     // V_R (dar x dbr, column-major) in the batched vector
  int64_t b_off;     // Bhat (db x dbr, column-major)
  double r;
  int32_t dbr, pad;
};
struct DBlock {      // one output block L, one column tile
  int64_t v_off;     // W_L / V_L base in the batched vector
  int64_t ha_off, hb_off;  // -1 when the side has no matrix for this J
  int32_t da, db, j0, nc;
  int32_t p_begin, p_end;  // range in part_list of the item partials of L
};

struct HostPlan {
  std::vector<DItem> items;
  std::vector<DATerm> aterms;
  std::vector<DRTerm> rterms;
  std::vector<DBlock> blocks;
  std::vector<int64_t> part_list;   // partial base of every item of each block, in item order
  std::vector<double> Ad, Bh, H;
  std::vector<int64_t> vbase;       // per instance
  int64_t N = 0, part_size = 0;
  int max_lds = 0;                  // doubles of LDS the pn kernel needs
  double flops_ref = 0.0;           // algorithmic work of one batched H*v
  double flops_like = 0.0;          // of which the like-particle (reduce kernel) part
  int64_t bytes_alg = 0;            // algorithmic HBM bytes of one batched H*v (FP64)
};

// Appends one instance to a plan.
void plan_add(HostPlan& p, const Model& m);

}  // namespace nsx
