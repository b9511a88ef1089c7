/* LAMMPS pair style ilp/graphene/hbn/hip: see pair_ilp_graphene_hbn_hip.h. */

#include "pair_ilp_graphene_hbn_hip.h"

#include <cstring>

#include "atom.h"
#include "error.h"
#include "force.h"
#include "memory.h"

using namespace LAMMPS_NS;

PairILPGrapheneHBNHip::PairILPGrapheneHBNHip(LAMMPS *lmp) : Pair(lmp)
{
  single_enable = 0;
  restartinfo = 0;
  one_coeff = 1;
  manybody_flag = 1;
  no_virial_fdotr_compute = 1;
  // ILP van der Waals and repulsion energies, as the reference style publishes them
  nextra = 2;
  pvector = new double[nextra];
  pvector[0] = pvector[1] = 0.0;
  cut_global = 0.0;
  tap_flag = 1;
  have_coeff = false;
}

PairILPGrapheneHBNHip::~PairILPGrapheneHBNHip()
{
  delete[] pvector;
  if (allocated) {
    memory->destroy(setflag);
    memory->destroy(cutsq);
  }
}

void PairILPGrapheneHBNHip::allocate()
{
  allocated = 1;
  int n = atom->ntypes;
  memory->create(setflag, n + 1, n + 1, "pair:setflag");
  memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
  for (int i = 1; i <= n; i++)
    for (int j = i; j <= n; j++) setflag[i][j] = 0;
}

/* The owning style (tersoff/hip or lmff/hip, earlier on the hybrid/overlay line) has already
   computed the ILP term with the Tersoff term; it also wrote this style's pvector. */
void PairILPGrapheneHBNHip::compute(int eflag, int vflag)
{
  ev_init(eflag, vflag);
}

void PairILPGrapheneHBNHip::settings(int narg, char **arg)
{
  if (narg < 1 || narg > 2) error->all(FLERR, "Illegal pair_style command");
  if (strcmp(force->pair_style, "hybrid/overlay") != 0)
    error->all(FLERR, "ERROR: requires hybrid/overlay pair_style");
  cut_global = force->numeric(FLERR, arg[0]);
  if (narg == 2) tap_flag = force->inumeric(FLERR, arg[1]);
}

/* pair_coeff * * <file> e1 .. e_ntypes: kept for the owner, which reads the file through the
   library (its parser applies the reference's checks and messages). */
void PairILPGrapheneHBNHip::coeff(int narg, char **arg)
{
  if (!allocated) allocate();
  if (narg != 3 + atom->ntypes) error->all(FLERR, "Incorrect args for pair coefficients");
  if (strcmp(arg[0], "*") != 0 || strcmp(arg[1], "*") != 0)
    error->all(FLERR, "Incorrect args for pair coefficients");
  potential_file = arg[2];
  elements.assign(arg + 3, arg + narg);
  for (int i = 1; i <= atom->ntypes; i++)
    for (int j = i; j <= atom->ntypes; j++)
      setflag[i][j] = (elements[i - 1] != "NULL" && elements[j - 1] != "NULL") ? 1 : 0;
  have_coeff = true;
}

void PairILPGrapheneHBNHip::init_style()
{
  if (force->newton_pair == 0) error->all(FLERR, "Pair style ilp/graphene/hbn requires newton pair on");
  if (!atom->molecule_flag) error->all(FLERR, "Pair style ilp/graphene/hbn requires atom attribute molecule");
  // no neighbor request: the library builds its own lists on the GPU
}

double PairILPGrapheneHBNHip::init_one(int i, int j)
{
  if (setflag[i][j] == 0 && setflag[j][i] == 0) return 0.0;
  return cut_global;
}
