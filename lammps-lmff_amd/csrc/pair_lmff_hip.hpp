// PairLMFFHip: host C++ adaptor with the LAMMPS Pair surface for either deck form
//   pair_style hybrid/overlay tersoff ilp/graphene/hbn <cut> [tap_flag]   (stock styles)
//   pair_style hybrid/overlay lmff ilp/graphene/hbn <cut> [tap_flag]      (LMFF fused style)
// over the C ABI of include/lmff_hip.h.
//
// Mirrors the reference interface (LMFF/pair.h:154-181): settings(narg, arg), coeff(narg, arg),
// init_style(), init_one(i, j), compute(eflag, vflag), with the same outputs (eng_vdwl,
// virial[6], eatom, vatom, pvector[2]) and error behaviour (an error aborts the command; here
// it throws lmff::PairError carrying the reference's message).  Inside LAMMPS the same class
// body sits in a `class PairLMFFHip : public Pair` (compiled against the LAMMPS headers and
// linked with the library); outside LAMMPS the caller hands in the atom arrays through
// AtomView.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "lmff_hip.h"

namespace lmff {

struct PairError : std::runtime_error {
  int code;
  PairError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct AtomView {
  int64_t nlocal = 0;
  const double *x = nullptr;       // [nlocal][3]
  double *f = nullptr;             // [nlocal][3], accumulated into (+=)
  const int32_t *type = nullptr;   // 1-based
  const int64_t *tag = nullptr;
  const int64_t *molecule = nullptr;
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {0, 0, 0};
  int periodic[3] = {1, 1, 1};
};

class PairLMFFHip {
 public:
  explicit PairLMFFHip(int ntypes, int device = 0);
  ~PairLMFFHip();
  PairLMFFHip(const PairLMFFHip &) = delete;
  PairLMFFHip &operator=(const PairLMFFHip &) = delete;

  // pair_style hybrid/overlay tersoff|lmff ilp/graphene/hbn <args>: args after "ilp/graphene/hbn"
  void settings(int narg, char **arg);
  // pair_coeff * * tersoff|lmff <file> e1..en   |   pair_coeff * * ilp/graphene/hbn <file> e1..en
  void coeff(int narg, char **arg);
  // neighbor/neigh_modify/mass state the reference reads in init
  void init_style(double skin, int every, int delay, int check, const std::vector<double> &mass);
  // hybrid cutoff of a type pair: max over sub-styles (pair_hybrid.cpp:642-685)
  double init_one(int i, int j) const;
  void set_precision(int mode);
  // Pair::compute semantics; new_list = the caller rebuilt its neighbor list this step
  void compute(int eflag, int vflag, const AtomView &atom, bool new_list);

  double eng_vdwl = 0.0;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  double pvector[2] = {0, 0};
  std::vector<double> eatom;  // [nlocal]
  std::vector<double> vatom;  // [nlocal][6]

  lmff_ctx *context() const { return ctx_; }

 private:
  void check(int rc) const;
  void apply_ilp();
  lmff_ctx *ctx_ = nullptr;
  int form_ = 0;                         // 0 not yet known, 1 tersoff, 2 lmff
  std::vector<bool> lmff_null_;          // types the lmff sub-style maps to NULL
  std::string ilp_file_;
  std::vector<std::string> ilp_elems_;
  int ntypes_;
  double cut_global_ = 16.0, tersoff_cutmax_ = 0.0, skin_ = 2.0;
  bool have_settings_ = false, have_tersoff_ = false, have_ilp_ = false;
};

}  // namespace lmff
