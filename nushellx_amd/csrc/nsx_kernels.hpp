// Device-side plan and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nsx_plan.hpp"

namespace nsx {

// Per-sample element strides of a linear ensemble (vectors, X, Bhat^T store,
// H store, large-block X, untransposed Bhat store) and its sample count n; all
// zero for an ordinary plan.
struct EnsStride {
  int64_t v = 0, x = 0, b = 0, h = 0, xg = 0, bn = 0;  // bn: the untransposed Bhat copy
  int32_t n = 0, pad = 0;
};

// per-XCD ranges of a task list, passed to a kernel by value
struct XcdSplit {
  int32_t s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

struct DevicePlan {
  DItem2* items2 = nullptr;
  DPhase* phases = nullptr;
  DPair* pairs = nullptr;
  int n_items2 = 0, lds_y = 1, lds_buf = 1;
  int max_units = 1;       // staged kernel: largest units per item
  int staged_nt = 128;     // staged kernel threads per workgroup
  int staged_mode = 0;     // 0: copy at phase start, 1: direct LDS loads one phase ahead
  int atomic = 0;          // 1: pn kernels add into W with FP64 atomics (no partials)
  size_t lds2_bytes = 0;
  DItem* items = nullptr;
  DATerm* aterms = nullptr;
  DRTerm* rterms = nullptr;
  DBlock* blocks = nullptr;
  int64_t* part_list = nullptr;
  double *Ad = nullptr, *Bh = nullptr, *H = nullptr, *part = nullptr;
  int n_items = 0, n_blocks = 0;
  int max_lds = 1;           // direct kernel Y tile (elements)
  int64_t N = 0;
  // stacked two-stage path (replaces the staged kernel when `stacked`)
  int stacked = 0, stk_nt = 256;
  SChunk1* s_chunk1 = nullptr;
  SUnit1* s_unit1 = nullptr;
  SChunk2* s_chunk2 = nullptr;
  STerm2* s_term2 = nullptr;
  SXRef* s_xref = nullptr;
  int n_chunk1 = 0, n_chunk2 = 0, s_lds1 = 1, s_lds2 = 1;
  void* X = nullptr;         // x_size elements of the accumulation type (FP64, or FP32 in reduced precision)
  // matrix-core stacked path (small blocks; uses s_term2, s_xref, X of the
  // stacked path, X transposed per product)
  int mstk = 0, mx_ks = 1;
  MTask1* m_task1 = nullptr;
  MUnit1* m_unit1 = nullptr;
  MTask2* m_task2 = nullptr;
  int n_mtask1 = 0, n_mtask2 = 0;
  // fused matrix-core path (small blocks; k_reduce stores the like-particle
  // part first, k_mxf adds the pn part)
  int mxf = 0, mxf_slots = kMfSlots, mxf_like = 0;
  double* Bn = nullptr;      // untransposed Bhat blocks (FP64)
  MFTask* f_mtask = nullptr;
  MFStep* f_mstep = nullptr;
  MFU* f_mu = nullptr;
  int n_mftask = 0;
  XcdSplit mxf_xcd;          // the tasks' XCD ranges (HostPlan::mxf_xcd)
  // deterministic fused path: per-task partial tiles and their W tiles
  int mxf_det = 0, n_dtile = 0;
  MFDTile* f_dtile = nullptr;
  int32_t* f_dpart = nullptr;
  double* f_dpsum = nullptr;  // n_mftask * mxf_slots * 256 partials
  // large-block GEMM path (FP64, atomic mode)
  GTask1* g_task1 = nullptr;
  GTile1* g_tile1 = nullptr;
  GTile2* g_tile2 = nullptr;
  GTile3* g_tile3 = nullptr;
  GTerm2* g_term2 = nullptr;
  SXRef* g_xref = nullptr;
  int n_gtile1 = 0, n_gtile2 = 0, n_gtile3 = 0;
  int big_rows1 = 64;        // rows of the stage-1 tiles (64, or 128 in BF16 plans)
  void* XG = nullptr;        // xg_size elements of the accumulation type (FP64, or FP32 in reduced precision)
  // fused column path (small blocks; deterministic): k_fuse writes column
  // partials, k_reduce sums them with the like-particle part
  int fused = 0, f_u = 2;    // f_u: output units per thread of the k_fuse instance
  FTask* f_task = nullptr;
  FBlk* f_blk = nullptr;
  FRound* f_round = nullptr;
  FSeg* f_seg = nullptr;
  FX* f_x = nullptr;
  FCons* f_cons = nullptr;
  FRef* f_ref = nullptr;
  DBlock* f_blocks = nullptr;
  int64_t* f_part_list = nullptr;
  double* f_part = nullptr;
  int n_ftask = 0, n_fblocks = 0, f_lds = 1;
  uint64_t gen = 0;          // changes at every upload (captured graphs hold plan pointers)
  // truncated plans: the window mask of the pn input and the masked FP64 copy
  uint8_t* vmask = nullptr;
  double* Vm = nullptr;
  // linear ensemble: the stacked and large-block kernels run ens_s samples
  // (grid y) at strides es
  int ens_s = 1;
  EnsStride es;
};

// Uploads a host plan; returns "" or an error message.
std::string upload_plan(const HostPlan& hp, DevicePlan& d, int precision);
void free_plan(DevicePlan& d);

// Operand storage of the BF16 hi/lo split: x ~ hi + lo with hi = bf16(x) and
// lo = bf16(x - hi), both round-to-nearest-even; hi in the upper 16 bits.
struct Bf16x2 {
  uint32_t bits;
};
__host__ __device__ inline uint32_t f32_bits(float x) {
  union { float f; uint32_t u; } c;
  c.f = x;
  return c.u;
}
__host__ __device__ inline float bits_f32(uint32_t u) {
  union { float f; uint32_t u; } c;
  c.u = u;
  return c.f;
}
// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint32_t bf16_rne(float x) {
  const uint32_t u = f32_bits(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__host__ __device__ inline Bf16x2 bf16x2_split(float x) {
  const uint32_t hi = bf16_rne(x);
  const float rest = x - bits_f32(hi << 16);
  return Bf16x2{(hi << 16) | bf16_rne(rest)};
}

// BF16 panels of the large-block tiles: a matrix of nx rows and inner (K)
// dimension K in 16-byte units of four K values, the hi halves of the four
// pairs first (two dwords), then the lo halves, stored step-major: unit
// (row x, 16-deep step s, quad qq) at dst + 4 (s nx + x) + qq holds K values
// 16s + 4qq .. 16s + 4qq + 3 of row x (K padded with zeros to whole steps;
// ldq = 4 ceil(K / 16) units per row).  A K step of 64 consecutive rows is
// 4 KB of contiguous memory.  A source block is column-major (rows x cols);
// kcols = 1: K runs along its columns (x = row), kcols = 0: K runs along its
// rows (x = column).
struct PanelBlk {
  int64_t src, dst;            // element offset in the Bf16x2 store, unit offset in the panel store
  int32_t rows, cols, kcols, ldq;
};
struct PanelChunk {            // 256 consecutive units of one block
  int32_t blk, u0;
};

// Reduced-precision copies of the operators (FP32 or BF16 hi/lo) and the
// converted input vector; structure arrays are shared with the FP64 plan.
struct LowpPlan {
  int precision = 0;
  void* Ad = nullptr;
  void* Bh = nullptr;
  void* H = nullptr;
  void* vin = nullptr;
  void* vin_m = nullptr;  // the masked input of the pn part (truncated plans)
  void* Bn = nullptr;     // the fused path's untransposed Bhat blocks
  // BF16 plans with large blocks: the tiles' operands as panels.  Bp (Bhat^T
  // per stage-1 task, K = its rows) and Ap (A per stage-2 term, K = its
  // columns) are made once; Vt (V_R per stage-1 task, K = its columns) is
  // made from the pn input by every H*v.  g_task1p / g_term2p are the plan's
  // records with v_off / b_off / a_off as unit offsets into these stores.
  uint4 *Vt = nullptr, *Bp = nullptr, *Ap = nullptr;
  GTask1* g_task1p = nullptr;
  GTerm2* g_term2p = nullptr;
  PanelBlk* vt_blk = nullptr;
  PanelChunk* vt_chunk = nullptr;
  int n_vt_chunk = 0;
  // the panel tiles' XCD grouping (xcd_group; 0: the hardware's round-robin order)
  int xcd_grp1 = 16, xcd_grp2 = 16;
};
// panels = false (linear ensembles, whose operators are made per sample on
// the device) keeps the large-block tiles on the Bf16x2 stores
std::string build_lowp(const HostPlan& hp, int precision, LowpPlan& lp, bool panels = true);
void free_lowp(LowpPlan& lp);

// prec: 0 FP64, 1 FP32, 2 BF16 hi/lo.  launch_hv = convert + (pn, reduce) in mode order.
// w_zero: the caller guarantees W == 0 on entry (the paths that accumulate
// every term into W then skip clearing it).
hipError_t launch_hv(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                     hipStream_t s, bool w_zero = false);
hipError_t launch_convert(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, hipStream_t s);
// panel units of Bf16x2 blocks (256 units per chunk)
hipError_t launch_panels(const PanelChunk* chunks, int n_chunks, const PanelBlk* blks, const void* src, uint4* dst,
                         hipStream_t s);
// n FP64 values -> FP32 or BF16 hi/lo (prec 1, 2), as the plan's operands are stored
hipError_t convert_to(int prec, const double* x, void* y, int64_t n, hipStream_t s);
hipError_t launch_pn(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                     hipStream_t s);
hipError_t launch_reduce(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                         hipStream_t s, bool w_zero = false);
int num_staged();
const void* k_pn_staged_ptr(int i);
const void* k_pn_partial_ptr(int prec, bool atomic);
const void* k_stk_ptr(int stage, int prec, int nt);
// k_fuse instance of a precision with u output units per thread (2, 4 or 8; 256 threads)
const void* k_fuse_ptr(int prec, int u);
constexpr int kFuseThreads = 256;

}  // namespace nsx
