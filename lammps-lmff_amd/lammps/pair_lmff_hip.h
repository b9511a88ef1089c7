/* LAMMPS pair style lmff/hip: the LMFF deck form on one MI355X through liblmff_hip.

   The decks of the reference (Cases/case_5016/in.replicate and the other cases) use the fused style:
     pair_style hybrid/overlay lmff ilp/graphene/hbn 16.0 0
     pair_coeff * * lmff C.opt.tersoff NULL C(O) C(O) C(O) C(O) C(O)
     pair_coeff * * ilp/graphene/hbn BNCH.ILP C C C C C C
   With the "hip" suffix PairHybrid creates "lmff/hip" and "ilp/graphene/hbn/hip".  This is
   tersoff/hip with the fused style's semantics for types mapped to NULL on its line: in the
   reference, PairHybrid gives the fused style a neighbor list that skips those types
   (LMFF/pair_hybrid.cpp:571-619) and the fused style computes ILP over that list while
   ilp/graphene/hbn only loads parameters, so those types drop out of ILP as well.

   Reference interface replaced: PairLMFF (LMFF/pair_lmff.h:14-16; settings
   LMFF/pair_lmff.cpp:2423-2426, coeff 2432-2500, compute 602-1613). */

#ifdef PAIR_CLASS

PairStyle(lmff/hip,PairLMFFHip)

#else

#ifndef LMP_PAIR_LMFF_HIP_H
#define LMP_PAIR_LMFF_HIP_H

#include "pair_tersoff_hip.h"

namespace LAMMPS_NS {

class PairLMFFHip : public PairTersoffHip {
 public:
  PairLMFFHip(class LAMMPS *lmp) : PairTersoffHip(lmp) { null_types_leave_ilp = true; }
};

}  // namespace LAMMPS_NS

#endif
#endif
