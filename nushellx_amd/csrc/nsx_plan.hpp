// Device work plan of a batch of pn J-scheme H*v instances.
//
// The vectors of all instances are concatenated (instance i occupies
// [vbase_i, vbase_i + N_i)).  An item is one (block L, a-side J') pair of the
// pn part, W_L += sum_k A_k sum_R r V_R Bhat^T.  Items whose blocks all have
// sides up to kSmallMax take the stacked two-stage path (or the staged kernel);
// larger items take the large-block matrix-core tiles, or, when a block is too
// large for their K-step masks, the direct kernel (pn_partial: per (item,
// column tile), partials or atomics).  The like-particle part H_a V_L +
// V_L H_b^T is a reduce pass per block (inside the stacked stage-1 launch on
// that path), or matrix-core tiles for large blocks.
#pragma once
#include <array>
#include <cstdint>
#include <unordered_map>
#include <string>
#include <vector>

#include "nsx_model.hpp"

namespace nsx {

constexpr int kThreads = 256;
constexpr int kMaxE = 8;                 // outputs per thread in one tile
constexpr int kTileElems = kThreads * kMaxE;
constexpr int kMaxLdsDoubles = 8192;     // Y tile capacity (64 KiB)
constexpr int kSmallMax = 40;            // block sides up to this use the staged kernel
constexpr int kPrefetch = (2 * kSmallMax * kSmallMax + kThreads - 1) / kThreads;
constexpr int kColBlk = 4;               // output columns per thread in the staged kernel
constexpr int kMaxU = (kSmallMax * ((kSmallMax + kColBlk - 1) / kColBlk) + kThreads - 1) / kThreads;
constexpr int kMaxPhases = 6;            // phases per staged workgroup (long items are split)
constexpr int kFatCap = 5120;            // doubles one phase stages in LDS (V run + Bhat run + A)
constexpr int kFatPF = kFatCap / kThreads;
constexpr int kStkMeta = 64;             // stacked path: pairs per stage-1 chunk
constexpr int kStkTerms = 32;            // stacked path: a-terms per stage-2 chunk
constexpr int kStkXref = 128;            // stacked path: X references per stage-2 chunk

struct DItem {       // one (L, Ja') partial, one column tile
  int64_t part_off;  // partial tile base (full da x db layout of the item)
  int64_t w_off;     // output block in the batched vector (atomic accumulation)
  int32_t da, dar, db, j0, nc;
  int32_t a_begin, a_end;
  int32_t pad;
};
struct DATerm {
  int64_t a_off;     // dense A (da x dar, column-major)
  int32_t r_begin, r_end;
};
// Staged kernel: an item is a list of phases.  A phase stages three contiguous
// runs in LDS -- the V_R blocks (Ja', Jb') of consecutive Jb' (adjacent in the
// vector), the Bhat^T blocks of the same Jb' (laid out adjacently per
// (Jb, lam, sa, sa') group) and, in the last phase of an a-term, A -- and then
// forms Y += sum_pairs r V_R Bhat^T and, with A, W += A Y.
struct DPhase {
  int64_t v_off, b_off, a_off;   // run starts: vector, Bhat^T store, A store (a_off < 0: no A)
  int32_t nv, nb, na;            // doubles per run
  int32_t pair_begin, pair_end;
  int32_t pad;
};
struct DPair {
  double r;
  int32_t v_lds, b_lds;          // offsets in the phase's LDS buffer
  int32_t dbr, pad;
};
struct DItem2 {
  int64_t part_off;  // private partial (deterministic accumulation)
  int64_t w_off;     // output block in the batched vector (atomic accumulation)
  int32_t da, dar, db;
  int32_t ph_begin, ph_end;
  int32_t pad;
};
struct DRTerm {
  int64_t v_off;     // V_R (dar x dbr, column-major) in the batched vector
  int64_t b_off;     // Bhat^T (dbr x db, column-major)
  double r;
  int32_t dbr, pad;
};
struct DBlock {      // one output block L, one column tile
  int64_t v_off;     // W_L / V_L base in the batched vector
  int64_t ha_off, hb_off;  // -1 when the side has no matrix for this J
  int32_t da, db, j0, nc;
  int32_t p_begin, p_end;  // range in part_list of the item partials of L
};

// Stacked two-stage pn path (small blocks):
//   stage 1: X_u = V_R Bhat^T for every distinct (V block, Bhat^T block) pair u,
//            computed once and shared by all left blocks that use it; a
//            workgroup takes a chunk of pairs that share one Bhat^T block;
//   stage 2: W_L += sum_k A_k (sum_u r_u X_u); a workgroup takes a chunk of
//            the a-terms of one left block L, accumulates W_L in registers and
//            adds it into W once.
struct SChunk1 {
  int64_t b_off;            // Bhat^T (dbr x db, column-major) in Bh
  int32_t dbr, db;
  int32_t u_begin, u_end;   // SUnit1 range
  int32_t work;             // sum over units of dar * ceil(db / 4)
  int32_t lds;              // doubles staged: dbr*db + sum dar*dbr
};
struct SUnit1 {
  int64_t v_off;            // V_R (dar x dbr, column-major) in the batched vector
  int64_t x_off;            // X (dar x db, column-major) in the X buffer
  int32_t dar, lds;         // rows; offset of V_R in the chunk's LDS buffer
  int32_t work0, pad;       // first work unit of this pair in the chunk
};
struct SChunk2 {
  int64_t w_off;            // W_L (da x db, column-major) in the batched vector
  int32_t da, db;
  int32_t t_begin, t_end;   // STerm2 range (a-terms of this left block)
  int32_t lds;              // doubles staged: sum over terms of da*dar + dar*db
  int32_t x_begin, x_end;   // SXRef range of those terms (read with the terms, not after them)
  int32_t pad;
};
struct STerm2 {
  int64_t a_off;            // A (da x dar, column-major) in Ad
  int32_t dar, lds;         // inner size; offset of A in the chunk's LDS (Y follows)
  int32_t x_begin, x_end;   // SXRef range
};
struct SXRef {
  int64_t x_off;
  double r;
};

// Matrix-core stacked path (small blocks): the two stacked stages as 16 x 16
// matrix-core tiles held in registers, operands loaded straight from memory
// (no LDS), X stored transposed (X_u^T, db x dar column-major) so that both
// stages read and write 16 consecutive elements per row of lanes.
//   stage 1: one wavefront per (Bhat^T block, 16-column tile j0, run of
//            products sharing that block): Bhat^T's K steps stay in registers,
//            and every 16-row tile of each V_R is multiplied against them;
//   stage 2: one workgroup per (left block L, 16 x 16 tile (i0, j0) of W_L,
//            run of a-terms): its four wavefronts take every fourth term,
//            Y_k = sum_u r_u X_u is formed in registers as the matrix-core
//            operand, the partial tiles are summed in LDS and added into W.
// MUnit1 (one 16-row tile of a product) and STerm2 (a_off, dar, x_begin,
// x_end) / SXRef describe the products and a-terms.
constexpr int kMxTile = 16;     // output tile side (16 x 16 x 4 matrix-core steps)
struct MTask1 {
  int64_t b_off;            // Bhat^T (dbr x db, column-major) in Bh
  int32_t dbr, db, j0;      // inner size, columns, first column of the tile
  int32_t u_begin, u_end;   // MUnit1 range
  int32_t pad;
};
struct MUnit1 {             // one 16-row tile of one product X_u = V_R Bhat^T
  int64_t v_off;            // V_R rows i0.. (leading dimension ld = dar) in the batched vector
  int64_t x_off;            // X_u^T rows i0.. (db x dar column-major: X_u^T + db * i0)
  int32_t ld, rows;         // dar; rows of this tile (<= 16)
};
struct MTask2 {
  int64_t w_off;            // W_L (da x db, column-major) in the batched vector
  int32_t da, db, i0, j0;   // block, first row and column of the tile
  int32_t t_begin, t_end;   // STerm2 range
};

// Fused matrix-core path (small blocks; stacked == 4): X never leaves the
// registers.  A slab (instance, Ja', Jb, lam, sa, sa') has products
// X_u = V_(Ja',Jb'_u) Bhat_(Jb,Jb'_u)^T and consumers (the left blocks (Ja, Jb)
// whose a-term A(Ja, Ja') reads them, each with its r_u).  A task is one
// 16-row tile of Ja' (a0), one 16-column tile of Jb (j0) and up to kMfSlots
// output slots (consumer Ja, 16-row tile i0 of W_(Ja,Jb)) of one (instance,
// Ja', Jb) group; its steps are the group's slabs.  Per step a wavefront forms
// every X_u tile on the matrix cores, accumulates Y_s += r_su X_u per slot in
// registers (the 16 x 16 x 4 FP64 result layout of X is the operand layout of
// the next product's K steps), then W_s^T += Y_s^T A_s^T; the slots' tiles stay
// in registers across the steps.  The four wavefronts of a workgroup take every
// fourth step of one task; their tiles are summed in LDS and added into W.
constexpr int kMfSlots = 4;
struct MFTask {
  int64_t w_off[kMfSlots];  // W_(Ja,Jb) + i0 + da * j0 (slot's tile origin); < 0: unused slot
  int32_t da[kMfSlots];     // leading dimension of the slot's W block and A
  int32_t rows_i[kMfSlots]; // valid rows of the slot's tile (Ja states)
  int32_t ld_v, rows_a;     // d(Ja') (V's leading dimension); valid Ja' rows of the tile
  int32_t db, cols_j;       // d(Jb) (Bhat^T columns); valid Jb columns of the tile
  int32_t p[5];             // wavefront w takes steps [p[w], p[w + 1])
  int32_t pad;
};
struct MFStep {
  int64_t a_off[kMfSlots];  // A(Ja, Ja') + i0 + da * a0 per slot; < 0: slot not in this slab
  int32_t u_begin, u_end;   // MFU range
};
struct MFU {                // one product of a slab
  int64_t v_off;            // V_(Ja',Jb') + a0 (leading dimension ld_v)
  int64_t b_off;            // Bhat(Jb, Jb') + j0 in the Bn store (db x dbr column-major)
  int32_t dbr, pad;         // d(Jb')
  double r[kMfSlots];       // r_su (0 for slots not in this slab)
};
// Deterministic variant: every task stores its summed slot tiles as partials
// (16 x 16, element row + 16 col, partial index task * slots + slot) instead
// of adding them into W, and one workgroup per W tile sums the tile's
// partials (listed in task order) in a fixed order and adds the sum onto W
// (cleared or holding the stored like-particle part: one addend onto a stored
// value).
struct MFDTile {
  int64_t w_off;            // W_(Ja,Jb) + i0 + da * j0
  int32_t da, rows, cols;   // leading dimension; valid rows and columns of the tile
  int32_t c_begin, c_end;   // range of partial indices in f_dpart
  int32_t pad;
};

// Fused small-block path (owner-computes, no atomics, deterministic):
// one workgroup owns the output column Jb of one instance (the blocks
// L = (Ja, Jb), concatenated into a column partial) for a range of the column's
// slabs.  A slab is one (Ja', lam, sa, sa') (separable: (Ja', lam, n, parity)):
// its products X_R = V_R Bhat_R^T over the right blocks R = (Ja', Jb'), and
// its consumers, the a-terms of the column's blocks that read them.  Slabs
// are taken in rounds that fit the LDS buffer,
//   stage V_R, Bhat_R^T, A_k  ->  X_R = V_R Bhat_R^T  ->  Y_k = sum r X_R  ->  W_L += A_k Y_k,
// so X and Y never leave LDS; the column's outputs stay in registers across
// rounds and the task writes its column partial once.  k_reduce then sums a
// block's partials in task order and adds H_a V_L + V_L H_b^T.
constexpr int kFMaxBlk = 16;     // blocks of one column
constexpr int kFMaxX = 64;       // products of one round
constexpr int kFMaxC = 64;       // consumers (a-terms) of one round
constexpr int kFMaxRef = 256;    // X references of one round
constexpr int kFMaxSeg = 128;    // staging segments of one round
constexpr int kFMaxRounds = 24;  // rounds of one task
constexpr int kFMaxU = 8;        // output units (row, 4 columns) per thread
constexpr int kFMaxLds = 32767;  // LDS element offsets fit 16 bits
struct FTask {
  int64_t part_off;         // column partial: the column's blocks, each da x db column-major
  int32_t r_begin, r_end;   // FRound range
  int32_t b_begin, b_end;   // FBlk range (the column's blocks, Ja ascending)
  int32_t db, units;        // column width; output units = sum over blocks of da * ceil(db / 4)
};
struct FBlk {
  int32_t wc_off, da, unit0, pad;  // offset in the column partial, rows, first output unit
};
struct FRound {
  int32_t s_begin, s_end;   // FSeg range (staging)
  int32_t x_begin, x_end;   // FX range
  int32_t c_begin, c_end;   // FCons range, sorted by column block
  int32_t f_begin, f_end;   // FRef range
  int32_t xwork, ywork;     // X units (row, 4 columns); Y elements
  int32_t wwork, wp_lds;    // W items (consumer, row, 4 columns); their partial sums (4 each) in LDS
  int32_t stage;            // elements staged (sum of the segment lengths)
  int32_t cb[kFMaxBlk + 1]; // consumers of the task's block b: [cb[b], cb[b + 1]) (round-relative)
};
// The round tables hold LDS offsets and sizes in 16 bits (the plan checks they fit).
struct FSeg {
  int64_t src;              // element offset in V (kind 0), Bh (1) or Ad (2)
  int16_t lds, pre, kind, pad;  // LDS offset; first element in the round's staging order
};
struct FX {
  int16_t v_lds, b_lds, x_lds;  // V_R (dar x dbr), Bhat^T (dbr x db), X (dar x db) in LDS
  int16_t dar, dbr, work0;      // first X unit
};
struct FCons {
  int16_t a_lds, y_lds;     // A (da x dar), Y (dar x db) in LDS
  int16_t da, dar;
  int16_t f_begin, f_end;   // round-relative FRef range
  int32_t y0;               // first Y element of the round
  int32_t w0;               // first W item of the round: items are (row, 4 columns) of A Y
  int32_t pad[3];
};
struct FRef {
  double r;
  int32_t x_lds, pad;
};

// Large-block path (blocks with a side above kSmallMax, in every precision of
// an atomic-mode plan): the same two stages as tiled GEMMs on the matrix
// cores, 64 x 64 output tiles.
//   G1: X_u = V_R Bhat^T for every distinct (V block, Bhat block) pair;
//   G2: W_L += sum_k A_k Y_k, Y_k = sum_u r_u X_u, tiles of W_L over chunks of
//       its a-terms (split-K), added into W with FP64 atomics.
// The densities are block-sparse (partition pairs), so every tile carries the
// set of 16-deep K steps in which its operand is not identically zero, and
// tiles whose operands vanish are not listed.
constexpr int kBigTile = 64;     // output tile side
constexpr int kBigStep = 16;     // K step of the tiles
constexpr int kBigMaxK = 2048;   // largest inner dimension the K-step masks cover (128 steps)
constexpr int kBigKChunk = 128;  // active K steps one stage-2 tile covers (split-K)
// X of the large-block path: columns padded to a multiple of kBigStep rows and
// 128-byte aligned, so that a K step of a column is one cache line
constexpr int big_ldx(int m) { return (m + kBigStep - 1) / kBigStep * kBigStep; }
struct GTask1 {
  int64_t v_off, b_off, x_off;   // V_R (m x k), Bhat^T (k x n), X (m x n, leading dimension big_ldx(m))
  int32_t m, k, n, pad;
};
struct GTile1 {
  uint64_t kmask[2];             // K steps where this tile's Bhat^T columns are not all zero
  int32_t task, i0, j0, pad;
};
struct GTerm2 {
  uint64_t kmask[2];             // K steps where A's rows of this tile are not all zero
  int64_t a_off;                 // A (da x dar, column-major)
  int32_t dar, x_begin, x_end, pad;
};
struct GTile2 {
  int64_t w_off;                 // W_L (da x db, column-major)
  int32_t da, db, i0, j0;
  int32_t t_begin, t_end;        // GTerm2 range (this row tile's terms of one K chunk)
};
// G0: the like-particle part of a large block, W_L = H_a V_L + V_L H_b^T, one
// 64 x 64 tile each (stored, so it initialises W_L before G2 adds into it)
struct GTile3 {
  int64_t v_off, ha_off, hb_off; // ha_off / hb_off < 0: no matrix on that side
  int32_t da, db, i0, j0;
};

// Tunable plan shape (chosen by the autotuner, see nsx_tune.cpp).
struct PlanConfig {
  int fat_cap = kFatCap;     // doubles one staged phase may load (<= kFatCap)
  int max_phases = kMaxPhases;  // phases per staged workgroup
  int stacked = 0;           // 1: small blocks take the two-stage stacked path
                             // (3: on the matrix cores; 2: the fused path)
  int stk_cap = 2048;        // stacked path: doubles one chunk may stage in LDS
  int stk_work = 512;        // stacked path: work units per chunk
  int big_min = kSmallMax;   // items with a side above this take the direct / large-block paths
  int big_rows1 = kBigTile;  // rows of the stage-1 large-block tiles: 64, or 128 (BF16 plans)
  // stacked == 2: the fused path, stk_cap = LDS elements of a round,
  // stk_work = flops one task (workgroup) takes at most, about, and
  // f_rounds = rounds one task takes at most
  int f_rounds = 4;
  // stage-2 large-block tiles take about this many active K steps each (split-K
  // over several workgroups adding into one W tile)
  int big_kchunk = kBigKChunk;
  int big_rows1_req = 0;     // requested stage-1 rows (0: by precision)
  // bitwise reproducible H*v: the fused path for the small blocks and one
  // stage-2 workgroup per large W tile (its single atomic add lands on the
  // stored like-particle value, so the order of additions cannot vary)
  int deterministic = 0;
  // stacked == 3, the matrix-core stacked path: products per stage-1
  // wavefront and a-terms per stage-2 workgroup
  int mx_units1 = 4;
  int mx_terms2 = 16;
  // stacked == 4, the fused matrix-core path: steps (slabs) per wavefront, at
  // most (0: a workgroup takes all steps of its task)
  int mxf_steps = 2;
  // output slots per fused task: 4 (kMfSlots), or 2 (fewer registers, more
  // products formed twice)
  int mxf_slots = kMfSlots;
  // and the like-particle part of the small blocks folded in (two more slabs
  // per block; W cleared first), where the plan allows it
  int mxf_like = 1;
};

// Linear ensembles: every sample's operators are linear in its parameters, so
// one template plan serves all samples.  A folded Bhat^T sub-block of the store
// is sum_p c_p val[p] Bsrc_p over the recipe's pairs (val: the sample's
// .oph value, zeroed below 1e-10 as at load); a like-particle H block is a copy
// of the sample's values.
struct EnsFold {   // one Bhat^T sub-block (dbr x db, column-major) of the Bh store
  int64_t dst;     // store offset
  int32_t n;       // elements
  int32_t p_begin, p_end;
  int32_t pad;
};
struct EnsPair {   // one b-density of a fold
  int64_t src;     // offset into the transposed dense densities
  double c;        // renorm of the template
  int32_t val;     // column of the sample's values
  int32_t pad;
};
struct EnsHBlk {   // one H block of the store and where its values come from
  int64_t dst, n;
  int32_t t, side, j, pad;  // template, side (0 = a, 1 = b), 2J / 2
};

struct HostPlan {
  PlanConfig cfg;
  int max_units = 1;                 // staged kernel: largest (row, column block) count of an item
  std::vector<DItem2> items2;        // staged (small-block) items
  std::vector<DPhase> phases;
  std::vector<DPair> pairs;
  int lds_y = 1;                     // staged kernel Y region (doubles)
  int lds_buf = 1;                   // staged kernel operand buffer (doubles): largest phase
  std::vector<DItem> items;          // direct kernel (large-block) tiles
  std::vector<DATerm> aterms;
  std::vector<DRTerm> rterms;
  std::vector<DBlock> blocks;
  std::vector<int64_t> part_list;   // partial base of every item of each block, in item order
  std::vector<double> Ad, Bh, H;
  std::vector<int64_t> vbase;       // per instance
  // N-hbar-omega windows: the pn part reads its input masked to the window
  // states (NuTRL's single_productm gathers lcoef states and adds only window
  // ones); 1 for every state of an untruncated instance.  Empty when no
  // instance is truncated.  The like-particle part reads the input unmasked.
  std::vector<uint8_t> vmask;
  // identical operator matrices are stored once (content hash -> offsets)
  std::unordered_map<uint64_t, std::vector<int64_t>> dedup_a, dedup_b, dedup_h;
  // (offset, length) of every block stored in Ad, Bh, H, in insertion order
  std::vector<std::pair<int64_t, int64_t>> a_blocks, b_blocks, h_blocks;
  // per instance: store offset of each adense / bhat slot of its model
  std::vector<std::vector<int64_t>> inst_a_off, inst_b_off;
  // stacked path (filled by build_stacked)
  std::vector<SChunk1> s_chunk1;
  std::vector<SUnit1> s_unit1;
  std::vector<SChunk2> s_chunk2;
  std::vector<STerm2> s_term2;
  std::vector<SXRef> s_xref;
  int64_t x_size = 0;          // X buffer elements
  int s_lds1 = 1, s_lds2 = 1;  // largest chunk LDS (doubles)
  // matrix-core stacked path (filled by build_mstk; shares the arrays above)
  std::vector<MTask1> m_task1;
  std::vector<MUnit1> m_unit1;
  std::vector<MTask2> m_task2;
  int mx_ks = 1;               // largest inner size in K steps of 4
  // fused matrix-core path (filled by build_mxf)
  std::vector<MFTask> f_mtask;
  std::vector<MFStep> f_mstep;
  std::vector<MFU> f_mu;
  bool mxf_det = false;          // deterministic fused path: tasks store partials (cfg.deterministic)
  std::vector<MFDTile> f_dtile;  // W tiles with partials, in order of their offsets
  std::vector<int32_t> f_dpart;  // their partial indices, in task order
  int mxf_slots = kMfSlots;      // slots the plan's tasks use (2 or 4)
  // the Bhat blocks the fused path reads, untransposed (db x dbr column-major),
  // so that a row of lanes reads 16 consecutive elements (and, with the
  // like-particle part folded in, identities and H_b blocks)
  std::vector<double> Bn;
  bool mxf_like = false;        // the fused path carries the like-particle part
  // f_mtask in 8 contiguous ranges of about equal cost, one per XCD (range x
  // = [mxf_xcd[x], mxf_xcd[x + 1])), each with its longest tasks first
  std::array<int32_t, 9> mxf_xcd{};
  // Bn blocks that are transposed copies of Bh blocks (linear ensembles make
  // them per sample on the device): {Bh offset, Bn offset, dbr, db}
  std::vector<std::array<int64_t, 4>> bn_src;
  // large-block path (filled by build_big)
  std::vector<GTask1> g_task1;
  std::vector<GTile1> g_tile1;
  std::vector<GTile2> g_tile2;
  std::vector<GTile3> g_tile3;
  std::vector<GTerm2> g_term2;
  std::vector<SXRef> g_xref;
  int64_t xg_size = 0;         // large-block X buffer (doubles)
  // fused path (filled by build_fused)
  std::vector<FTask> f_task;
  std::vector<FBlk> f_blk;
  std::vector<FRound> f_round;
  std::vector<FSeg> f_seg;
  std::vector<FX> f_x;
  std::vector<FCons> f_cons;
  std::vector<FRef> f_ref;
  std::vector<DBlock> f_blocks;      // p.blocks with the fused partials
  std::vector<int64_t> f_part_list;
  int64_t f_part_size = 0;           // doubles of the column partials
  int f_lds = 1, f_units = 1;        // largest round (LDS elements); largest task's output units
  // linear ensembles (filled by plan_add when ens is set): Bh and H blocks are
  // interned by their source (files and keys) instead of their values
  bool ens = false;
  int32_t ens_t = 0;                   // template being added
  int32_t ens_opcol = 0;               // first value column of its .oph section
  std::unordered_map<std::string, int64_t> ens_key;
  std::vector<EnsFold> ens_fold;
  std::vector<EnsPair> ens_pair;
  std::vector<double> ens_bsrc;        // dense b densities in the Bhat^T layout
  std::vector<EnsHBlk> ens_h;
  std::vector<EnsHBlk> ens_h_alias;    // later templates' blocks that share a stored one
  int64_t N = 0, part_size = 0;
  int max_lds = 0;                  // doubles of LDS the pn kernel needs
  double flops_ref = 0.0;           // algorithmic work of one batched H*v
  double flops_like = 0.0;          // of which the like-particle (reduce kernel) part
  int64_t bytes_alg = 0;            // algorithmic HBM bytes of one batched H*v (FP64)
};

// Appends one instance to a plan.
void plan_add(HostPlan& p, const Model& m);
// Appends a plan built on its own (same PlanConfig) to p: offsets are moved and
// q's operator blocks are interned into p's stores.  plan_add of instances
// into separate plans (in parallel) followed by plan_merge in order gives the
// same H*v as plan_add of all of them into one plan.
void plan_merge(HostPlan& p, const HostPlan& q);
// Fills the stacked-path arrays of p for the small-block items of `models`
// (the instances planned into p, in order).  Clears p.cfg.stacked (the staged
// kernel then runs) if an a-term has more references than a chunk can hold.
void build_stacked(HostPlan& p, const std::vector<const Model*>& models);
// Fills the matrix-core stacked path's arrays (m_unit1, s_term2, s_xref,
// m_task1, m_task2, x_size) for the small-block items of `models`.
void build_mstk(HostPlan& p, const std::vector<const Model*>& models);
// Fills the fused matrix-core path's arrays (f_mtask, f_mstep, f_mu, mx_ks)
// for the small-block items of `models`; false (arrays empty) if a slab does
// not have the shape the path assumes.
bool build_mxf(HostPlan& p, const std::vector<const Model*>& models);
// Fills the fused-path arrays of p for the small-block items of `models`.
// Returns false (and leaves them empty) if a column exceeds the kernel's
// limits (blocks per column, output units) or the plan has items that only
// the direct kernel takes.
bool build_fused(HostPlan& p, const std::vector<const Model*>& models, int threads);
// Fills the large-block GEMM arrays of p for the items with a side above
// p.cfg.big_min (kSmallMax unless a test lowers it); leaves them empty, so the
// direct kernel takes those items, when a block exceeds kBigMaxK.
void build_big(HostPlan& p, const std::vector<const Model*>& models);

}  // namespace nsx
