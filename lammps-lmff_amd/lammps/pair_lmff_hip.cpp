/* LAMMPS pair style lmff/hip: see pair_lmff_hip.h (the class is tersoff/hip with the fused
   style's NULL-type semantics; this file gives the style its own translation unit, as LAMMPS'
   build expects one .cpp per style header). */

#include "pair_lmff_hip.h"
