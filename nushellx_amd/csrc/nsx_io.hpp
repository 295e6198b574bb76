// Readers for the NuShellX operator files consumed by the pn J-scheme H*v.
//
// File set and name resolution follow program NuTRL
// (src/NuTRL/NuCoreX.f90:1634-1670): <nus>.nus, <opa>.tri/.trd, <opb>.tri/.trd,
// <int>.oph/.spe, and per side and J the .mat/.sgt pair named
// <side><file_code(J,T,nmp/2+36)><xtra> (src/NuCommon/NuModules.F90:22-35,553-597).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nsx {

struct NusInput {                     // src/NuCommon/NuCommon.f90:2546-2736
  std::string nucleus, desc[3];       // the name and the three description lines
  int no_cI = 0, no_cN = 0, tparity = 0, no_shells = 0, np_shells = 0;
  std::vector<int> sj2, sp;           // per shell: 2j, parity
  std::vector<int> sn, sl;            // per shell: radial n and l (oscillator quanta 2n + l)
  int min_cJ2 = 0, max_cJ2 = 0, min_cT2 = 0, max_cT2 = 0, nmp = 0;
  int min_MSon = 0, max_MSon = 999, del_cJ2 = 2, del_cT2 = 2;
};

struct Density {                      // one .trd key: (2Jl, 2Jr, shl, shr, 2lam)
  uint8_t key[8];
  int64_t pos = 0;                    // offset into SideFiles::vals
  int count = 0;
  std::vector<uint32_t> mask;         // [nsp columns][words]: bit npl of column npr
};

struct SideFiles {
  int nsp = 0, minJ2 = 0, maxJ2 = 0, delJ2 = 2, words = 1, max_lm = 0;
  bool proton = false, neutron = false;
  float xtag = 0.f;
  std::vector<int> nJT;               // [(J2/2) * nsp + p]
  std::vector<int> parity;            // per partition
  std::vector<int> nq;                // per partition: sum over shells of (2n + l) * occupation
  std::vector<int16_t> occ;           // [p * no_shells + shell]: occupation of each shell (all shells)
  std::vector<Density> dens;          // sorted by key
  std::vector<double> vals;           // density values (file REAL*4 or REAL*8, widened)
  // like-particle Hamiltonian per J: dense d(J) x d(J), symmetric, empty if absent
  std::vector<std::vector<double>> H;
  int nJ() const { return maxJ2 / 2 + 1; }
  int n(int J2, int p) const { return nJT[(J2 / 2) * nsp + p]; }
  int d(int J2) const;
  int offset(int J2, int p) const;    // first row of partition p inside the J basis
};

struct OphEntry { uint8_t key[8]; double me; };

// Separable (Dufour-Zuker) pn operator, <name>.dop as NuDiop writes it
// (src/NuDiop/NuDiop.f90:170-280) and NuTRLZ reads it (src/NuTRLZ/NuCoreXZ.f90:1637-1660):
// per (2lam, parity) section, eigen-terms e_n with components (proton pair,
// neutron pair, amplitude), so that the pn multipole interaction is
// V(pp', nn'; lam) = sum_n e_n u_n(pp') u_n(nn').
struct DopLine { int s[4]; double me; };        // shells p, p', n, n' (as written), amplitude
struct DopTerm { double e = 0.0; std::vector<DopLine> lines; };  // e rounded to REAL*4
struct DopSection { int lm = 0, par = 0; std::vector<DopTerm> terms; };

// Returns an empty string on success, else the error message.
std::string read_nus(const std::string& prefix, NusInput& out);
std::string read_side(const std::string& opfile, const NusInput& nus, SideFiles& out);
std::string read_mats(const std::string& opfile, const NusInput& nus, int side_nmp, char xtra,
                      SideFiles& s);
std::string read_oph(const std::string& intfile, std::vector<OphEntry>& out, int& max_lmo);
// Reads <dopfile>.dop with Fortran list-directed rules: every READ starts on a
// new line, may continue onto the next lines, and skips the rest of its last line.
std::string read_dop(const std::string& dopfile, int& max_lmo, std::vector<DopSection>& out);
std::string read_spe(const std::string& intfile, double& renorm, int& no_level);

std::string file_code(int i, int j, int k);

}  // namespace nsx
