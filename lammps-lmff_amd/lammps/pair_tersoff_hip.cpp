/* LAMMPS pair style tersoff/hip: see pair_tersoff_hip.h. */

#include "pair_tersoff_hip.h"

#include <cstdlib>
#include <cstring>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neighbor.h"
#include "pair_ilp_graphene_hbn_hip.h"

using namespace LAMMPS_NS;

PairTersoffHip::PairTersoffHip(LAMMPS *lmp) : Pair(lmp)
{
  single_enable = 0;
  restartinfo = 0;
  one_coeff = 1;
  manybody_flag = 1;
  // the library returns the F.r virial over all periodic images itself; LAMMPS' fdotr sum over
  // owned + ghost atoms would not see the images (PairHybrid honours this, pair_hybrid.cpp:97-102)
  no_virial_fdotr_compute = 1;
  ctx = nullptr;
  null_types_leave_ilp = false;
  partner = nullptr;
  cutmax = 0.0;
  const char *dev = getenv("LMFF_DEVICE");
  if (lmff_create(&ctx, dev ? atoi(dev) : 0) != LMFF_OK)
    error->all(FLERR, "tersoff/hip: cannot open the MI355X device (lmff_create failed)");
}

PairTersoffHip::~PairTersoffHip()
{
  lmff_destroy(ctx);
  if (allocated) {
    memory->destroy(setflag);
    memory->destroy(cutsq);
  }
}

void PairTersoffHip::check(int rc)
{
  if (rc != LMFF_OK) error->all(FLERR, lmff_last_error(ctx));
}

void PairTersoffHip::allocate()
{
  allocated = 1;
  int n = atom->ntypes;
  memory->create(setflag, n + 1, n + 1, "pair:setflag");
  memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
  for (int i = 1; i <= n; i++)
    for (int j = i; j <= n; j++) setflag[i][j] = 0;
}

void PairTersoffHip::settings(int narg, char ** /*arg*/)
{
  if (narg != 0) error->all(FLERR, "Illegal pair_style command");
}

/* pair_coeff * * <file> e1 .. e_ntypes: the file is read by the library at init_style (its
   parser applies the reference's checks and messages); here only the map is kept. */
void PairTersoffHip::coeff(int narg, char **arg)
{
  if (!allocated) allocate();
  if (narg != 3 + atom->ntypes) error->all(FLERR, "Incorrect args for pair coefficients");
  if (strcmp(arg[0], "*") != 0 || strcmp(arg[1], "*") != 0)
    error->all(FLERR, "Incorrect args for pair coefficients");
  tersoff_file = arg[2];
  tersoff_elements.assign(arg + 3, arg + narg);
  int count = 0;
  for (int i = 1; i <= atom->ntypes; i++)
    for (int j = i; j <= atom->ntypes; j++) {
      setflag[i][j] = (tersoff_elements[i - 1] != "NULL" && tersoff_elements[j - 1] != "NULL") ? 1 : 0;
      count += setflag[i][j];
    }
  if (count == 0) error->all(FLERR, "Incorrect args for pair coefficients");
}

void PairTersoffHip::init_style()
{
  if (atom->tag_enable == 0) error->all(FLERR, "Pair style Tersoff requires atom IDs");
  if (force->newton_pair == 0) error->all(FLERR, "Pair style Tersoff requires newton pair on");
  if (!atom->molecule_flag) error->all(FLERR, "Pair style ilp/graphene/hbn requires atom attribute molecule");
  if (comm->nprocs != 1) error->all(FLERR, "tersoff/hip runs on one MPI rank (the library holds the whole box)");
  if (domain->triclinic) error->all(FLERR, "tersoff/hip requires an orthogonal box");
  partner = dynamic_cast<PairILPGrapheneHBNHip *>(force->pair_match("ilp/graphene/hbn", 1, 0));
  if (!partner || !partner->have_coeff)
    error->all(FLERR, "tersoff/hip computes ILP too: it needs ilp/graphene/hbn/hip with its pair_coeff "
                      "in the same hybrid/overlay");
  const int nt = atom->ntypes;
  check(lmff_settings(ctx, nt, partner->cut_global, partner->tap_flag));
  std::vector<const char *> te(nt), ie(nt);
  for (int t = 0; t < nt; t++) {
    te[t] = tersoff_elements[t].c_str();
    const bool drop = null_types_leave_ilp && tersoff_elements[t] == "NULL";
    ie[t] = drop ? "NULL" : partner->elements[t].c_str();
  }
  check(lmff_coeff_tersoff(ctx, tersoff_file.c_str(), te.data()));
  check(lmff_coeff_ilp(ctx, partner->potential_file.c_str(), ie.data()));
  check(lmff_neighbor(ctx, neighbor->skin, neighbor->every, neighbor->delay, neighbor->dist_check));
  check(lmff_masses(ctx, &atom->mass[1]));
  const char *prec = getenv("LMFF_PRECISION");
  check(lmff_precision(ctx, (prec && strcmp(prec, "mixed") == 0) ? LMFF_MIXED : LMFF_FP64));
  // PairTersoff::init_one's cutmax: the largest R + D of the mapped entries
  check(lmff_tersoff_cutmax(ctx, &cutmax));
  // no LAMMPS neighbor request: the library builds its own cell-sorted lists on the GPU
}

double PairTersoffHip::init_one(int i, int j)
{
  if (setflag[i][j] == 0) error->all(FLERR, "All pair coeffs are not set");
  return cutmax;
}

const int64_t *PairTersoffHip::widen(const tagint *p, int n, std::vector<int64_t> &buf)
{
  if (sizeof(tagint) == sizeof(int64_t)) return reinterpret_cast<const int64_t *>(p);
  buf.resize(n > 0 ? n : 1);
  for (int i = 0; i < n; i++) buf[i] = (int64_t)p[i];
  return buf.data();
}

/* One Pair::compute(eflag, vflag) of Tersoff + ILP on the owned atoms; the library reuses its
   list while LAMMPS does (neighbor->ago > 0: same atoms in the same order). */
void PairTersoffHip::compute(int eflag, int vflag)
{
  ev_init(eflag, vflag);
  const int nlocal = atom->nlocal;
  double boxlo[3] = {domain->boxlo[0], domain->boxlo[1], domain->boxlo[2]};
  double boxhi[3] = {domain->boxhi[0], domain->boxhi[1], domain->boxhi[2]};
  int per[3] = {domain->xperiodic, domain->yperiodic, domain->zperiodic};
  const int64_t *tag = widen(atom->tag, nlocal, tag64);
  const int64_t *mol = widen(atom->molecule, nlocal, mol64);
  double pv[2] = {0.0, 0.0};
  check(lmff_pair_compute(ctx, nlocal, nlocal ? &atom->x[0][0] : nullptr, atom->type, tag, mol, boxlo, boxhi, per,
                          neighbor->ago == 0 ? 1 : 0, eflag, vflag, nlocal ? &atom->f[0][0] : nullptr, &eng_vdwl,
                          virial, eflag_atom ? eatom : nullptr, (vflag_atom && nlocal) ? &vatom[0][0] : nullptr, pv));
  if (partner && partner->pvector) {
    partner->pvector[0] = pv[0];
    partner->pvector[1] = pv[1];
  }
}
