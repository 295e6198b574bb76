// Host setup of one pn J-scheme H*v instance (one nucleus, one total J).
//
// Mirrors what NuTRL's Lnz_TRD builds before its Lanczos loop
// (src/NuTRL/NuCoreX.f90:807-1367): the (2Ja,2Jb) block basis, the Racah
// recoupling scalars (:1071-1127) and the drdbh term tables (:1428-1577).
// Instead of per-thread bitmask walks the terms are regrouped for the GPU:
//
//   W_L = H_a(Ja) V_L + V_L H_b(Jb)^T
//       + sum_{Ja'} sum_{k=(Ja,Ja',sa,sa',lam)} A_k  sum_{Jb'} r(Ja,Jb,Ja',Jb',lam)
//                                                   V_(Ja',Jb') Bhat(Jb,Jb',lam,sa,sa')^T
//
// with Bhat = sum_{sb,sb'} opmeV(sa,sa',sb,sb',lam) B(Jb,Jb',sb,sb',lam) folded once
// (the b-density list of a drdbh entry depends only on (Jb,Jb',lam,sa,sa')).
//
// The separable variant of NuTRLZ (src/NuTRLZ/NuCoreXZ.f90:1495-1887) has the same
// shape with k = (Ja,Ja',lam,parity,n) running over the eigen-terms of a .dop file:
// A_k = sum_lines e_n renorm u_n(sa,sa') A(sa,sa') and Bhat = sum_lines u_n(sb,sb') B(sb,sb').
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "nsx_io.hpp"

namespace nsx {

struct RTerm {
  int R;        // right block index
  int bhat;     // index into Model::bhat
  double r;     // Racah factor r(Ja,Jb,Ja',Jb',lam)
};
struct ATerm {
  int adens;    // index into Model::adense
  std::vector<RTerm> rt;
};
struct Item {   // one (L, Ja') work item
  int L;
  int ajr;
  std::vector<ATerm> at;
};

struct Model {
  NusInput nus;
  std::string intfile;                     // interaction name as given (NuVec prints it)
  std::string opfile_a, opfile_b;          // side file names as given
  SideFiles A, B;
  int J2 = 0, basd = 0, no_level = 5;
  double renorm = 1.0;
  int64_t N = 0;
  bool azero = false, bzero = false;
  bool separable = false;                  // pn part from a .dop file (NuTRLZ)
  std::vector<int> bas_a, bas_b, da, db;   // per block: 2Ja, 2Jb, d(Ja), d(Jb)
  std::vector<int64_t> posnv;
  std::vector<uint8_t> lcoef;              // parity mask of the basis states
  // N-hbar-omega window (min_MSon/max_MSon of the .nus): trunc marks the states
  // of the requested parity inside it, bas_active the blocks with any partition
  // pair inside it; inactive blocks get no H*v output, as in the reference
  bool truncate = false;
  std::vector<uint8_t> trunc, bas_active;
  std::vector<Item> items;
  // dense a-densities d(Jl) x d(Jr), column-major, and their keys
  std::vector<std::vector<double>> adense;
  std::vector<std::array<int, 2>> adense_J;  // (2Jl, 2Jr)
  // folded b-operators d(Jb) x d(Jb'), column-major
  std::vector<std::vector<double>> bhat;
  std::vector<std::array<int, 2>> bhat_J;    // (2Jb, 2Jb')
  std::vector<std::array<int, 5>> bhat_key;  // (2Jb, 2Jb', 2lam, sa, sa') or (2Jb, 2Jb', 2lam, n, parity)
  int64_t nnz_a = 0, nnz_b = 0;              // operator values the reference stores per side
  // linear-ensemble templates: the .oph values as read (file order), and per
  // bhat slot its fold recipe (b-density index, .oph record index); the pn
  // structure then keeps every .oph key present in the file, zero or not, so
  // that any interaction with the same keys fits it
  bool keep_zero_ops = false;
  std::vector<double> oph_me;
  std::vector<std::vector<std::pair<int, int>>> bhat_src;
  // algorithmic (reference sparse) work of one H*v, and term counts
  double flops_ref = 0.0;               // all terms
  double flops_like = 0.0;              // H_a V + V H_b^T part of flops_ref
  int64_t n_aterms = 0, n_rterms = 0, n_bterms = 0;
};

// Reads every file of one NuTRL run and builds the model.  Returns "" on success.
// With a non-empty `dopfile` the pn part comes from <dopfile>.dop as in NuTRLZ.
std::string build_model(const std::string& nus_prefix, const std::string& opfile_a,
                        const std::string& opfile_b, const std::string& intfile, Model& m,
                        const std::string& dopfile = "", bool keep_zero_ops = false);

// Restatement of the reference's float-bisection key search
// (src/NuCommonX/NuCommonX.f90:147-183); 1-based index or 0.
int tfind_ref(const std::vector<std::array<uint8_t, 8>>& keys, const uint8_t* key);

// Dense d(Jl) x d(Jr) column-major matrix of one stored density of a side.
std::vector<double> densify(const SideFiles& s, const Density& dd, int J2l, int J2r);

double sixj(int a, int b, int c, int d, int e, int f);
double racah_w(int j1, int j2, int j3, int j4, int J, int K);

}  // namespace nsx
