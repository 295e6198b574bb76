#include "nsx_io.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace nsx {

static const char kAnCode[] =
    "0123456789_$cdefghijklmnopqrstuvwxyz@ABCDEFGHIJKLMNMPQRSTUVWXYZ?a%^&=b-+#";
static const int kMaxCode = 72;

std::string file_code(int i, int j, int k) {
  auto f = [](int x) { return x < 0 ? x + kMaxCode : x; };
  std::string s;
  s += kAnCode[f(i)];
  s += kAnCode[f(j)];
  s += kAnCode[f(k)];
  return s;
}

// Fortran unformatted sequential file held in memory.
class Records {
 public:
  bool open(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    data_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    pos_ = 0;
    return true;
  }
  // next record payload; false at end of file or on a corrupt marker
  bool next(const uint8_t*& p, int& len) {
    if (pos_ + 4 > data_.size()) return false;
    int32_t n;
    std::memcpy(&n, data_.data() + pos_, 4);
    if (n < 0 || pos_ + 8 + size_t(n) > data_.size()) return false;
    int32_t n2;
    std::memcpy(&n2, data_.data() + pos_ + 4 + n, 4);
    if (n2 != n) return false;
    p = reinterpret_cast<const uint8_t*>(data_.data()) + pos_ + 4;
    len = n;
    pos_ += 8 + size_t(n);
    return true;
  }

 private:
  std::vector<char> data_;
  size_t pos_ = 0;
};

static int32_t i32(const uint8_t* p, int k) {
  int32_t v;
  std::memcpy(&v, p + 4 * k, 4);
  return v;
}

int SideFiles::d(int J2) const {
  int s = 0;
  for (int p = 0; p < nsp; ++p) s += n(J2, p);
  return s;
}
int SideFiles::offset(int J2, int p) const {
  int s = 0;
  for (int q = 0; q < p; ++q) s += n(J2, q);
  return s;
}

std::string read_nus(const std::string& prefix, NusInput& n) {
  std::ifstream f(prefix + ".nus");
  if (!f) return "cannot open " + prefix + ".nus";
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line[0] == '!') continue;
    lines.push_back(line);
  }
  size_t i = 4;
  auto need = [&](size_t k) { return lines.size() > k; };
  if (!need(10)) return prefix + ".nus: too short";
  n.nucleus = lines[0];
  for (int k = 0; k < 3; ++k) n.desc[k] = lines[1 + k];
  std::istringstream(lines[i++]) >> n.no_cI >> n.no_cN;
  std::istringstream(lines[i++]) >> n.tparity;
  std::istringstream(lines[i++]) >> n.no_shells;
  std::string nuc = lines[i++];
  n.np_shells = 0;
  size_t b = nuc.find_first_not_of(' ');
  for (int k = 0; k < n.no_shells && b + k < nuc.size() && nuc[b + k] == 'p'; ++k) n.np_shells++;
  n.sj2.assign(n.no_shells, 0);
  n.sp.assign(n.no_shells, 1);
  n.sn.assign(n.no_shells, 0);
  n.sl.assign(n.no_shells, 0);
  for (int k = 0; k < n.no_shells; ++k) {
    if (!need(i)) return prefix + ".nus: shells";
    int a, nn, l, j2, t2, par;
    std::istringstream ss(lines[i++]);
    if (!(ss >> a >> nn >> l >> j2 >> t2 >> par)) return prefix + ".nus: bad shell line";
    n.sj2[k] = j2;
    n.sp[k] = par;
    n.sn[k] = nn;
    n.sl[k] = l;
  }
  if (!need(i + 2)) return prefix + ".nus: truncated";
  std::istringstream(lines[i++]) >> n.min_cJ2 >> n.max_cJ2;
  std::istringstream(lines[i++]) >> n.min_cT2 >> n.max_cT2 >> n.nmp;
  std::istringstream(lines[i++]) >> n.min_MSon >> n.max_MSon;
  i += 5;  // four shell-limit lines, output_control
  if (lines.size() > i) {
    int a = 0, c = 0;
    std::istringstream ss(lines[i]);
    if (ss >> a >> c) { n.del_cJ2 = a; n.del_cT2 = c; }
  }
  if (n.del_cJ2 < 2 || (n.del_cJ2 & 1)) n.del_cJ2 = 2;
  if (n.del_cT2 < 2 || (n.del_cT2 & 1)) n.del_cT2 = 2;
  return "";
}

std::string read_side(const std::string& opfile, const NusInput& nus, SideFiles& s) {
  Records r;
  const uint8_t* p;
  int len;
  if (!r.open(opfile + ".tri")) return "cannot open " + opfile + ".tri";
  if (!r.next(p, len) || len < 36) return opfile + ".tri: bad header";
  s.nsp = i32(p, 0);
  s.minJ2 = i32(p, 1);
  s.maxJ2 = i32(p, 2);
  s.proton = i32(p, 3) != 0;
  s.neutron = i32(p, 4) != 0;
  std::memcpy(&s.xtag, p + 20, 4);
  s.delJ2 = i32(p, 8);
  if (s.delJ2 < 2) s.delJ2 = 2;
  s.words = (s.nsp - 1) / 32 + 1;
  const int nJf = s.maxJ2 / 2 - s.minJ2 / 2 + 1;
  s.nJT.assign(size_t(s.nJ()) * s.nsp, 0);
  if (!r.next(p, len) || len < 4 * s.nsp * nJf) return opfile + ".tri: bad nJT record";
  for (int j = 0; j < nJf; ++j)
    for (int q = 0; q < s.nsp; ++q) s.nJT[size_t(s.minJ2 / 2 + j) * s.nsp + q] = i32(p, j * s.nsp + q);
  if (!r.next(p, len)) return opfile + ".tri: no partition record";
  const int ptdim = len / (2 * s.nsp);
  s.parity.assign(s.nsp, 1);
  s.nq.assign(s.nsp, 0);
  s.occ.assign(size_t(s.nsp) * nus.no_shells, 0);
  for (int q = 0; q < s.nsp; ++q) {  // cParity, src/NuCommon/NuCommon.f90:2918-2937
    int par = 1, nq = 0;
    for (int k = 0; k < nus.no_shells && k < ptdim; ++k) {
      int16_t occ;
      std::memcpy(&occ, p + 2 * (size_t(q) * ptdim + k), 2);
      s.occ[size_t(q) * nus.no_shells + k] = occ;
      for (int t = 0; t < occ; ++t) par *= nus.sp[k];
      nq += (2 * nus.sn[k] + nus.sl[k]) * occ;  // cNN, src/NuTRL/NuCoreX.f90:814-817
    }
    s.parity[q] = par;
    s.nq[q] = nq;
  }

  Records t;
  if (!t.open(opfile + ".trd")) return "cannot open " + opfile + ".trd";
  if (!t.next(p, len) || len < 12) return opfile + ".trd: bad header";
  const int trdd = i32(p, 0);
  s.max_lm = i32(p, 1);
  s.dens.assign(trdd, Density());
  if (trdd > 0) {
    if (!t.next(p, len) || len < 8 * trdd) return opfile + ".trd: bad key record";
    for (int k = 0; k < trdd; ++k) std::memcpy(s.dens[k].key, p + 8 * k, 8);
    if (!t.next(p, len) || len < 4 * trdd) return opfile + ".trd: bad count record";
    for (int k = 0; k < trdd; ++k) s.dens[k].count = i32(p, k);
    for (int k = 0; k < trdd; ++k) {
      Density& d = s.dens[k];
      d.pos = int64_t(s.vals.size());
      d.mask.assign(size_t(s.nsp) * s.words, 0u);
      if (d.count == 0) continue;
      if (!t.next(p, len)) return opfile + ".trd: truncated values";
      const int kind = len / d.count;
      for (int v = 0; v < d.count; ++v) {
        double x;
        if (kind == 8) std::memcpy(&x, p + 8 * v, 8);
        else { float y; std::memcpy(&y, p + 4 * v, 4); x = y; }
        s.vals.push_back(x);
      }
      if (!t.next(p, len) || len < 4 * s.nsp * s.words) return opfile + ".trd: truncated mask";
      std::memcpy(d.mask.data(), p, 4 * size_t(s.nsp) * s.words);
    }
  }
  return "";
}

// .mat/.sgt per J; segments expanded into dense symmetric matrices.  Value
// order follows mulmtx (src/NuCommonX/NuCommonX.f90:1009-1104): columns np
// ascending, row bits npr ascending, only same-parity segments consume values.
std::string read_mats(const std::string& opfile, const NusInput& nus, int side_nmp, char xtra,
                      SideFiles& s) {
  s.H.assign(s.nJ(), std::vector<double>());
  const int k3 = side_nmp / 2 + kMaxCode / 2;  // C and Fortran both truncate toward zero
  for (int J2 = s.minJ2; J2 <= s.maxJ2; J2 += 2) {
    for (int T2 = nus.min_cT2; T2 <= nus.max_cT2; T2 += 2) {
      const std::string name = opfile + file_code(J2 / 2, T2 / 2, k3) + xtra;
      Records sg, mt;
      if (!sg.open(name + ".sgt")) continue;
      const uint8_t* p;
      int len;
      if (!sg.next(p, len) || !sg.next(p, len) || len < 12) continue;
      const int nseg = i32(p, 0);
      float xx;
      std::memcpy(&xx, p + 8, 4);
      if (xx != s.xtag) return name + ".sgt: mat file and proj file not consistent";
      if (!mt.open(name + ".mat")) return "cannot open " + name + ".mat";
      if (!mt.next(p, len)) return name + ".mat: empty";
      const int d = s.d(J2);
      std::vector<double>& H = s.H[J2 / 2];
      H.assign(size_t(d) * d, 0.0);
      // collect segments, then consume in (column, row-bit) order
      struct Seg { int np, npr, dimpt, dimptr; std::vector<double> m; };
      std::vector<Seg> segs;
      for (int g = 0; g < nseg; ++g) {
        if (!mt.next(p, len)) return name + ".mat: truncated";
        Seg sgm;
        sgm.dimpt = i32(p, 0);
        sgm.dimptr = i32(p, 1);
        sgm.np = i32(p, 2);
        sgm.npr = i32(p, 3);
        const int n = sgm.dimpt * sgm.dimptr;
        int kind = n ? (len - 16 - 8 - 64) / n : 4;
        if (kind != 8) kind = 4;
        sgm.m.resize(n);
        for (int v = 0; v < n; ++v) {
          if (kind == 8) std::memcpy(&sgm.m[v], p + 16 + 8 * v, 8);
          else { float y; std::memcpy(&y, p + 16 + 4 * v, 4); sgm.m[v] = y; }
        }
        segs.push_back(std::move(sgm));
      }
      // concatenated value stream, walked as the reference does
      std::vector<double> stream;
      std::vector<std::vector<uint8_t>> bit(s.nsp, std::vector<uint8_t>(s.nsp, 0));
      for (auto& sgm : segs) {
        stream.insert(stream.end(), sgm.m.begin(), sgm.m.end());
        if (sgm.np < 1 || sgm.np > s.nsp || sgm.npr < 1 || sgm.npr > s.nsp)
          return name + ".mat: partition index out of range";
        bit[sgm.np - 1][sgm.npr - 1] = 1;
      }
      size_t pt = 0;
      for (int j = 0; j < s.nsp; ++j)          // column partition np
        for (int i = 0; i < s.nsp; ++i) {      // row partition npr
          if (!bit[j][i] || s.parity[i] != s.parity[j]) continue;
          const int oj = s.offset(J2, j), oi = s.offset(J2, i);
          for (int k = 0; k < s.n(J2, j); ++k)
            for (int l = 0; l < s.n(J2, i); ++l) {
              if (pt >= stream.size()) return name + ".mat: value stream too short";
              const double me = stream[pt++];
              // element H((npr,l),(np,k)); the other triangle when npr != np
              H[size_t(oj + k) * d + (oi + l)] = me;
              if (i != j) H[size_t(oi + l) * d + (oj + k)] = me;
            }
        }
      break;  // one T per J
    }
  }
  return "";
}

std::string read_oph(const std::string& intfile, std::vector<OphEntry>& out, int& max_lmo) {
  Records r;
  const uint8_t* p;
  int len;
  if (!r.open(intfile + ".oph")) return "cannot open " + intfile + ".oph";
  if (!r.next(p, len) || len < 8) return intfile + ".oph: bad header";
  const int opd = i32(p, 0);
  max_lmo = i32(p, 1);
  out.resize(opd);
  for (int k = 0; k < opd; ++k) {
    if (!r.next(p, len) || len < 16) return intfile + ".oph: truncated";
    std::memcpy(out[k].key, p, 8);
    std::memcpy(&out[k].me, p + 8, 8);
  }
  return "";
}

namespace {

// Fortran list-directed input over a text file held as lines of tokens.
class ListInput {
 public:
  bool open(const std::string& path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
      for (char& c : line)
        if (c == ',' || c == '\t' || c == '\r') c = ' ';
      std::istringstream is(line);
      std::vector<std::string> toks;
      std::string t;
      while (is >> t) toks.push_back(t);
      lines_.push_back(std::move(toks));
    }
    return true;
  }
  // one READ statement of n values; false at end of file or on a token that is not a number
  bool read(int n, double* v) {
    int got = 0;
    while (got < n) {
      if (ln_ >= lines_.size()) return false;
      const auto& toks = lines_[ln_++];
      for (size_t k = 0; k < toks.size() && got < n; ++k) {
        std::string t = toks[k];
        for (char& c : t)
          if (c == 'd' || c == 'D') c = 'e';
        char* end = nullptr;
        v[got++] = std::strtod(t.c_str(), &end);
        if (end == t.c_str() || *end != 0) return false;
      }
    }
    return true;
  }
  size_t line() const { return ln_; }

 private:
  std::vector<std::vector<std::string>> lines_;
  size_t ln_ = 0;
};

}  // namespace

std::string read_dop(const std::string& dopfile, int& max_lmo, std::vector<DopSection>& out) {
  const std::string path = dopfile + ".dop";
  ListInput in;
  if (!in.open(path)) return "cannot open " + path;
  auto bad = [&](const char* what) { return path + ": " + what + " near line " + std::to_string(in.line()); };
  double v[5];
  if (!in.read(1, v)) return bad("no header");
  max_lmo = (int)v[0];
  if (max_lmo < 0 || max_lmo > 254) return bad("bad maximum multipole");
  out.clear();
  // the reader loops over 2 * (max_lmo/2 + 1) sections (src/NuTRLZ/NuCoreXZ.f90:1638-1641)
  const int nsec = 2 * (max_lmo / 2 + 1);
  for (int s = 0; s < nsec; ++s) {
    DopSection sec;
    if (!in.read(2, v)) return bad("missing section header");
    sec.lm = (int)v[0];
    sec.par = (int)v[1];
    if (sec.lm < 0 || sec.lm > max_lmo || (sec.lm & 1) || (sec.par != 1 && sec.par != -1))
      return bad("bad section header");
    if (!in.read(1, v)) return bad("missing term count");
    const int nelm = (int)v[0];
    if (nelm < 0 || nelm > 255) return bad("bad term count");
    for (int n = 0; n < nelm; ++n) {
      DopTerm t;
      if (!in.read(1, v)) return bad("missing eigenvalue");
      t.e = (double)(float)v[0];  // REAL*4 in the reader (:1617)
      if (!in.read(1, v)) return bad("missing component count");
      const int nle = (int)v[0];
      if (nle < 0) return bad("bad component count");
      for (int c = 0; c < nle; ++c) {
        if (!in.read(5, v)) return bad("missing component");
        DopLine l;
        for (int q = 0; q < 4; ++q) {
          l.s[q] = (int)v[q];
          if (l.s[q] < 1 || l.s[q] > 255) return bad("bad shell index");
        }
        l.me = v[4];
        t.lines.push_back(l);
      }
      sec.terms.push_back(std::move(t));
    }
    out.push_back(std::move(sec));
  }
  return "";
}

std::string read_spe(const std::string& intfile, double& renorm, int& no_level) {
  std::ifstream f(intfile + ".spe");
  if (!f) return "cannot open " + intfile + ".spe";
  std::string l1, l2, l3;
  if (!std::getline(f, l1) || !std::getline(f, l2)) return intfile + ".spe: too short";
  renorm = std::atof(l2.c_str());
  no_level = 5;
  if (std::getline(f, l3)) std::istringstream(l3) >> no_level;
  return "";
}

}  // namespace nsx
