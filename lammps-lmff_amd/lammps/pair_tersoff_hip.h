/* LAMMPS pair style tersoff/hip: Tersoff + ILP/graphene/hBN on one MI355X through liblmff_hip.

   An unchanged deck selects it with the "hip" suffix (lmp -sf hip ...):
     pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 1
     pair_coeff * * tersoff BNC.tersoff C B N
     pair_coeff * * ilp/graphene/hbn BNCH.ILP C B N
   PairHybrid creates "tersoff/hip" and "ilp/graphene/hbn/hip" (LMFF/pair_hybrid.cpp:334,
   LMFF/force.cpp:255-272).  This style owns the computation: at init_style() it takes the
   ILP settings and coefficients from its ilp/graphene/hbn/hip partner and hands everything to
   the library (lmff_settings, lmff_coeff_tersoff, lmff_coeff_ilp, lmff_neighbor, lmff_masses);
   compute() is one lmff_pair_compute of both terms on the owned atoms, with LAMMPS' Pair::compute
   outputs (f +=, eng_vdwl, virial, eatom, vatom) and the partner's pvector.

   Reference interfaces replaced: PairTersoff (Baseline/MANYBODY/pair_tersoff.h, settings 284-287,
   coeff 293-358, init_style 365-377, init_one 383-388, compute 86-266) together with the compute
   of PairILPGrapheneHBN (Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp:428-675).

   Limits: one MPI rank (the library takes the whole orthogonal periodic box and handles the
   periodic images itself; ghosts are not used), newton pair on, atom attribute molecule.  The
   device is LMFF_DEVICE (default 0); LMFF_PRECISION=mixed selects the mixed-precision mode. */

#ifdef PAIR_CLASS

PairStyle(tersoff/hip,PairTersoffHip)

#else

#ifndef LMP_PAIR_TERSOFF_HIP_H
#define LMP_PAIR_TERSOFF_HIP_H

#include <cstdint>
#include <string>
#include <vector>

#include "lmff_hip.h"
#include "pair.h"

namespace LAMMPS_NS {

class PairILPGrapheneHBNHip;

class PairTersoffHip : public Pair {
 public:
  PairTersoffHip(class LAMMPS *);
  virtual ~PairTersoffHip();

  virtual void compute(int, int);
  void settings(int, char **);
  virtual void coeff(int, char **);
  void init_style();
  double init_one(int, int);

 protected:
  lmff_ctx *ctx;
  std::string tersoff_file;
  std::vector<std::string> tersoff_elements;  // per atom type, "NULL" for unmapped types
  // the lmff deck form: types the fused style maps to NULL drop out of ILP as well
  bool null_types_leave_ilp;
  PairILPGrapheneHBNHip *partner;
  double cutmax;
  std::vector<int64_t> tag64, mol64;  // tagint widened for the C ABI when tagint is 32-bit

  void allocate();
  void check(int rc);
  const int64_t *widen(const tagint *p, int n, std::vector<int64_t> &buf);
};

}  // namespace LAMMPS_NS

#endif
#endif
