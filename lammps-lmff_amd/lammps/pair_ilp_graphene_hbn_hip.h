/* LAMMPS pair style ilp/graphene/hbn/hip: the parameter-only partner of tersoff/hip and lmff/hip.

   With the "hip" suffix (lmp -sf hip) the unchanged deck line
     pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 1
   makes PairHybrid::settings create "tersoff/hip" and "ilp/graphene/hbn/hip"
   (Force::new_pair appends the suffix to every sub-style, LMFF/pair_hybrid.cpp:334,
   LMFF/force.cpp:255-272).  The MI355X library evaluates Tersoff and ILP in one pass, so the
   owning style (tersoff/hip or lmff/hip) computes both terms and this style only keeps what the
   deck gives ILP: pair_style arguments (cutoff, Tap flag) and the pair_coeff line (potential
   file, element map).  Its compute() returns without computing, as the reference's parameter
   loader for the fused LMFF style does (SWLMFF-5GCC/pair_ilp_graphene_hbn.cpp:497-503); the
   owner fills its pvector (ILP van der Waals and repulsion energies) so that
   compute pair ilp/graphene/hbn keeps working.

   Reference interface replaced: PairILPGrapheneHBN (Baseline/USER-MISC/pair_ilp_graphene_hbn.h),
   settings 132-149, coeff 155-219, init_style 393-398. */

#ifdef PAIR_CLASS

PairStyle(ilp/graphene/hbn/hip,PairILPGrapheneHBNHip)

#else

#ifndef LMP_PAIR_ILP_GRAPHENE_HBN_HIP_H
#define LMP_PAIR_ILP_GRAPHENE_HBN_HIP_H

#include <string>
#include <vector>

#include "pair.h"

namespace LAMMPS_NS {

class PairILPGrapheneHBNHip : public Pair {
 public:
  PairILPGrapheneHBNHip(class LAMMPS *);
  virtual ~PairILPGrapheneHBNHip();

  virtual void compute(int, int);
  void settings(int, char **);
  void coeff(int, char **);
  void init_style();
  double init_one(int, int);

  // read by the owning style at its init_style()
  double cut_global;
  int tap_flag;
  std::string potential_file;
  std::vector<std::string> elements;  // per atom type, "NULL" for unmapped types
  bool have_coeff;

 protected:
  void allocate();
};

}  // namespace LAMMPS_NS

#endif
#endif
