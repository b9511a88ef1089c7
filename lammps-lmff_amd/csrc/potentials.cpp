// Potential-file readers (see potentials.hpp for the reference anchors).
#include "potentials.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "lmff_hip.h"

namespace lmff {

namespace {

std::vector<std::string> split_words(const std::string &s) {
  std::vector<std::string> w;
  std::istringstream is(s);
  std::string t;
  while (is >> t) w.push_back(t);
  return w;
}

// Reads logical entries of exactly `nwords` words.  A '#' starts a comment; an entry with
// too few words continues on the following line(s), as pair_tersoff.cpp:420-450 does.
class EntryReader {
 public:
  explicit EntryReader(std::istream &is) : is_(is) {}
  // 1 = entry read, 0 = end of file, -1 = malformed entry
  int next(size_t nwords, std::vector<std::string> &out) {
    std::string line;
    for (;;) {
      if (!std::getline(is_, line)) return 0;
      strip(line);
      out = split_words(line);
      if (out.empty()) continue;
      while (out.size() < nwords) {
        std::string more;
        if (!std::getline(is_, more)) return -1;
        strip(more);
        line += " " + more;
        out = split_words(line);
      }
      return out.size() == nwords ? 1 : -1;
    }
  }

 private:
  static void strip(std::string &l) {
    auto p = l.find('#');
    if (p != std::string::npos) l.erase(p);
  }
  std::istream &is_;
};

int find_element(const ElementMap &m, const std::string &name) {
  for (size_t e = 0; e < m.elements.size(); ++e)
    if (m.elements[e] == name) return (int)e;
  return -1;
}

}  // namespace

bool map_elements(int ntypes, const char *const *names, ElementMap &out, std::string &err) {
  out.elements.clear();
  out.map.assign(ntypes + 1, -1);
  for (int t = 1; t <= ntypes; ++t) {
    if (!names[t - 1]) {
      err = "Incorrect args for pair coefficients";
      return false;
    }
    std::string nm = names[t - 1];
    if (nm == "NULL") {
      out.map[t] = -1;
      continue;
    }
    int e = find_element(out, nm);
    if (e < 0) {
      e = (int)out.elements.size();
      out.elements.push_back(nm);
    }
    out.map[t] = e;
  }
  if (out.elements.empty()) {
    err = "Incorrect args for pair coefficients";
    return false;
  }
  return true;
}

int read_tersoff(const std::string &file, TersoffTable &t, std::string &err) {
  std::ifstream is(file);
  if (!is) {
    err = "Cannot open Tersoff potential file " + file;
    return LMFF_ERR_FILE;
  }
  EntryReader rd(is);
  std::vector<std::string> w;
  t.params.clear();
  for (;;) {
    int r = rd.next(17, w);
    if (r == 0) break;
    if (r < 0) {
      err = "Incorrect format in Tersoff potential file";
      return LMFF_ERR_FILE;
    }
    int ie = find_element(t.emap, w[0]), je = find_element(t.emap, w[1]), ke = find_element(t.emap, w[2]);
    if (ie < 0 || je < 0 || ke < 0) continue;
    TersoffEntry p{};
    p.ielement = ie;
    p.jelement = je;
    p.kelement = ke;
    p.powerm = std::atof(w[3].c_str());
    p.gamma = std::atof(w[4].c_str());
    p.lam3 = std::atof(w[5].c_str());
    p.c = std::atof(w[6].c_str());
    p.d = std::atof(w[7].c_str());
    p.h = std::atof(w[8].c_str());
    p.powern = std::atof(w[9].c_str());
    p.beta = std::atof(w[10].c_str());
    p.lam2 = std::atof(w[11].c_str());
    p.bigb = std::atof(w[12].c_str());
    p.bigr = std::atof(w[13].c_str());
    p.bigd = std::atof(w[14].c_str());
    p.lam1 = std::atof(w[15].c_str());
    p.biga = std::atof(w[16].c_str());
    p.powermint = (int)p.powerm;
    if (p.c < 0.0 || p.d < 0.0 || p.powern < 0.0 || p.beta < 0.0 || p.lam2 < 0.0 || p.bigb < 0.0 ||
        p.bigr < 0.0 || p.bigd < 0.0 || p.bigd > p.bigr || p.lam1 < 0.0 || p.biga < 0.0 ||
        p.powerm - p.powermint != 0.0 || (p.powermint != 3 && p.powermint != 1) || p.gamma < 0.0) {
      err = "Illegal Tersoff parameter";
      return LMFF_ERR_PARAM;
    }
    t.params.push_back(p);
  }
  const int ne = t.nelements();
  t.elem2param.assign((size_t)ne * ne * ne, -1);
  for (int i = 0; i < ne; ++i)
    for (int j = 0; j < ne; ++j)
      for (int k = 0; k < ne; ++k) {
        int n = -1;
        for (size_t m = 0; m < t.params.size(); ++m) {
          const auto &p = t.params[m];
          if (p.ielement == i && p.jelement == j && p.kelement == k) {
            if (n >= 0) {
              err = "Potential file has duplicate entry";
              return LMFF_ERR_PARAM;
            }
            n = (int)m;
          }
        }
        if (n < 0) {
          err = "Potential file is missing an entry";
          return LMFF_ERR_PARAM;
        }
        t.elem2param[(i * ne + j) * ne + k] = n;
      }
  t.cutmax = 0.0;
  for (auto &p : t.params) {
    p.cut = p.bigr + p.bigd;
    p.cutsq = p.cut * p.cut;
    p.c1 = std::pow(2.0 * p.powern * 1.0e-16, -1.0 / p.powern);
    p.c2 = std::pow(2.0 * p.powern * 1.0e-8, -1.0 / p.powern);
    p.c3 = 1.0 / p.c2;
    p.c4 = 1.0 / p.c1;
    if (p.cut > t.cutmax) t.cutmax = p.cut;
  }
  return 0;
}

int read_ilp(const std::string &file, IlpTable &t, std::string &err) {
  std::ifstream is(file);
  if (!is) {
    err = "Cannot open ILP potential file " + file;
    return LMFF_ERR_FILE;
  }
  EntryReader rd(is);
  std::vector<std::string> w;
  t.params.clear();
  for (;;) {
    int r = rd.next(13, w);
    if (r == 0) break;
    if (r < 0) {
      err = "Insufficient format in ILP potential file";
      return LMFF_ERR_FILE;
    }
    int ie = find_element(t.emap, w[0]), je = find_element(t.emap, w[1]);
    if (ie < 0 || je < 0) continue;
    IlpEntry p{};
    p.ielement = ie;
    p.jelement = je;
    p.z0 = std::atof(w[2].c_str());
    p.alpha = std::atof(w[3].c_str());
    p.delta = std::atof(w[4].c_str());
    p.epsilon = std::atof(w[5].c_str());
    p.C = std::atof(w[6].c_str());
    p.d = std::atof(w[7].c_str());
    p.sR = std::atof(w[8].c_str());
    p.reff = std::atof(w[9].c_str());
    p.C6 = std::atof(w[10].c_str());
    p.S = std::atof(w[11].c_str());
    p.rcut = std::atof(w[12].c_str());
    // energies in meV scaled by S (ilp.cpp:354-359)
    const double meV = 1e-3 * p.S;
    p.C *= meV;
    p.C6 *= meV;
    p.epsilon *= meV;
    p.delta2inv = std::pow(p.delta, -2.0);
    p.lambda = p.alpha / p.z0;
    p.seff = p.sR * p.reff;
    t.params.push_back(p);
  }
  const int ne = t.nelements();
  t.elem2param.assign((size_t)ne * ne, -1);
  t.cutILPsq.assign((size_t)ne * ne, 0.0);
  for (int i = 0; i < ne; ++i)
    for (int j = 0; j < ne; ++j) {
      int n = -1;
      for (size_t m = 0; m < t.params.size(); ++m) {
        const auto &p = t.params[m];
        if (p.ielement == i && p.jelement == j) {
          if (n >= 0) {
            err = "Potential file has duplicate entry";
            return LMFF_ERR_PARAM;
          }
          n = (int)m;
        }
      }
      if (n < 0) {
        err = "Potential file is missing an entry";
        return LMFF_ERR_PARAM;
      }
      t.elem2param[i * ne + j] = n;
      t.cutILPsq[i * ne + j] = t.params[n].rcut * t.params[n].rcut;
    }
  return 0;
}

}  // namespace lmff
