// PairLMFFHip: see pair_lmff_hip.hpp.
#include "pair_lmff_hip.hpp"

#include <cstdlib>
#include <cstring>

#include "potentials.hpp"

namespace lmff {

PairLMFFHip::PairLMFFHip(int ntypes, int device) : ntypes_(ntypes) {
  int rc = lmff_create(&ctx_, device);
  if (rc) throw PairError(rc, "cannot create the HIP context (no MI355X device?)");
}

PairLMFFHip::~PairLMFFHip() { lmff_destroy(ctx_); }

void PairLMFFHip::check(int rc) const {
  if (rc) throw PairError(rc, lmff_last_error(ctx_));
}

// PairILPGrapheneHBN::settings (pair_ilp_graphene_hbn.cpp:132-149): cut_global [tap_flag]
void PairLMFFHip::settings(int narg, char **arg) {
  if (narg < 1 || narg > 2) throw PairError(LMFF_ERR_ARG, "Illegal pair_style command");
  char *end = nullptr;
  cut_global_ = std::strtod(arg[0], &end);
  if (end == arg[0] || *end) throw PairError(LMFF_ERR_ARG, "Expected floating point parameter in input script");
  int tap = 1;
  if (narg == 2) tap = std::atoi(arg[1]);
  check(lmff_settings(ctx_, ntypes_, cut_global_, tap));
  have_settings_ = true;
}

// PairTersoff::coeff (pair_tersoff.cpp:293-358), PairLMFF::coeff (LMFF/pair_lmff.cpp:2432-2497)
// and PairILPGrapheneHBN::coeff (ilp.cpp:155-219).  "lmff" takes a Tersoff file like
// "tersoff" does; the two deck forms differ only in which atoms ILP sees (apply_ilp).
void PairLMFFHip::coeff(int narg, char **arg) {
  if (narg != 4 + ntypes_) throw PairError(LMFF_ERR_ARG, "Incorrect args for pair coefficients");
  if (std::strcmp(arg[0], "*") != 0 || std::strcmp(arg[1], "*") != 0)
    throw PairError(LMFF_ERR_ARG, "Incorrect args for pair coefficients");
  if (!have_settings_) throw PairError(LMFF_ERR_ARG, "Pair coeff for hybrid has invalid style");
  const std::string style = arg[2];
  const char *file = arg[3];
  std::vector<const char *> elems(arg + 4, arg + 4 + ntypes_);
  if (style == "tersoff" || style == "lmff") {
    const int form = style == "tersoff" ? 1 : 2;
    if (form_ != 0 && form_ != form)
      throw PairError(LMFF_ERR_ARG, "Pair coeff for hybrid has invalid style: " + style);
    check(lmff_coeff_tersoff(ctx_, file, elems.data()));
    TersoffTable t;
    std::string err;
    map_elements(ntypes_, elems.data(), t.emap, err);
    read_tersoff(file, t, err);
    tersoff_cutmax_ = t.cutmax;
    have_tersoff_ = true;
    form_ = form;
    lmff_null_.assign((size_t)ntypes_ + 1, false);
    if (form == 2)
      for (int t = 1; t <= ntypes_; t++) lmff_null_[t] = std::strcmp(elems[t - 1], "NULL") == 0;
    if (have_ilp_) apply_ilp();
  } else if (style == "ilp/graphene/hbn") {
    ilp_file_ = file;
    ilp_elems_.assign(elems.begin(), elems.end());
    apply_ilp();
    have_ilp_ = true;
  } else {
    throw PairError(LMFF_ERR_ARG, "Pair coeff for hybrid has invalid style: " + style);
  }
}

// In the lmff form the fused style computes ILP itself, over the neighbor list PairHybrid
// hands the lmff sub-style; that list skips every type the lmff pair_coeff maps to NULL
// (LMFF/pair_hybrid.cpp:571-619), while the ilp/graphene/hbn sub-style only loads parameters.
// So ILP leaves those types out as well.  In the stock form ILP keeps its own mapping.
void PairLMFFHip::apply_ilp() {
  std::vector<const char *> el;
  for (int t = 1; t <= ntypes_; t++) {
    const bool skip = form_ == 2 && t < (int)lmff_null_.size() && lmff_null_[t];
    el.push_back(skip ? "NULL" : ilp_elems_[t - 1].c_str());
  }
  check(lmff_coeff_ilp(ctx_, ilp_file_.c_str(), el.data()));
}

void PairLMFFHip::init_style(double skin, int every, int delay, int check_yes, const std::vector<double> &mass) {
  if (!have_tersoff_ || !have_ilp_) throw PairError(LMFF_ERR_ARG, "All pair coeffs are not set");
  if ((int)mass.size() != ntypes_) throw PairError(LMFF_ERR_ARG, "Not all per-type masses are set");
  skin_ = skin;
  check(lmff_neighbor(ctx_, skin, every, delay, check_yes));
  check(lmff_masses(ctx_, mass.data()));
}

double PairLMFFHip::init_one(int, int) const { return cut_global_ > tersoff_cutmax_ ? cut_global_ : tersoff_cutmax_; }

void PairLMFFHip::set_precision(int mode) { check(lmff_precision(ctx_, mode)); }

void PairLMFFHip::compute(int eflag, int vflag, const AtomView &a, bool new_list) {
  const int eflag_atom = (eflag / 2) & 1;
  const int vflag_atom = (vflag & 4) ? 1 : 0;
  if (eflag_atom) eatom.assign((size_t)a.nlocal, 0.0);
  if (vflag_atom) vatom.assign((size_t)a.nlocal * 6, 0.0);
  check(lmff_pair_compute(ctx_, a.nlocal, a.x, a.type, a.tag, a.molecule, a.boxlo, a.boxhi, a.periodic,
                          new_list ? 1 : 0, eflag, vflag, a.f, &eng_vdwl, virial,
                          eflag_atom ? eatom.data() : nullptr, vflag_atom ? vatom.data() : nullptr, pvector));
}

}  // namespace lmff
