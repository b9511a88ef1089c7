// Potential-file readers of the LMFF force path (host C++).
//
// They keep the reference's file formats and checks:
//   Tersoff: Baseline/MANYBODY/pair_tersoff.cpp:392-532 (read_file), 536-580 (setup_params)
//   ILP:     Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp:246-387 (read_file, meV scaling)
#pragma once

#include <string>
#include <vector>

namespace lmff {

struct TersoffEntry {
  int ielement, jelement, kelement, powermint;
  double powerm, gamma, lam3, c, d, h, powern, beta, lam2, bigb, bigr, bigd, lam1, biga;
  double cut, cutsq, c1, c2, c3, c4;
};

struct IlpEntry {
  int ielement, jelement;
  double z0, alpha, delta, epsilon, C, d, sR, reff, C6, S, rcut;
  double delta2inv, lambda, seff;
};

// Element mapping of "pair_coeff * * <style> <file> e1 .. e_ntypes":
// map[t] = element index of LAMMPS type t (1-based), -1 for NULL.
struct ElementMap {
  std::vector<std::string> elements;
  std::vector<int> map;  // size ntypes + 1
};

bool map_elements(int ntypes, const char *const *names, ElementMap &out, std::string &err);

struct TersoffTable {
  ElementMap emap;
  std::vector<TersoffEntry> params;
  std::vector<int> elem2param;  // [ne][ne][ne]
  double cutmax = 0.0;
  int nelements() const { return (int)emap.elements.size(); }
  int index(int i, int j, int k) const {
    const int ne = nelements();
    return elem2param[(i * ne + j) * ne + k];
  }
};

struct IlpTable {
  ElementMap emap;
  std::vector<IlpEntry> params;
  std::vector<int> elem2param;   // [ne][ne]
  std::vector<double> cutILPsq;  // [ne][ne]
  int nelements() const { return (int)emap.elements.size(); }
};

// Both return 0 on success, else an LMFF_ERR_* code with a message in err.
int read_tersoff(const std::string &file, TersoffTable &t, std::string &err);
int read_ilp(const std::string &file, IlpTable &t, std::string &err);

}  // namespace lmff
