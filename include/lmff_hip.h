/*
 * lmff_hip.h -- C ABI of the MI355X LMFF force path (Tersoff + ILP/graphene/hBN).
 *
 * Plain C types only: pointers, sizes and status codes.  Every function
 * returns LMFF_OK (0) or a positive LMFF_ERR_* code; lmff_last_error() gives
 * the message.  A context is not thread-safe: drive it from one host thread.
 * Arrays are host pointers, atom-major ([n][3] = x0 y0 z0 x1 ...), in the
 * caller's atom order; the library keeps its own cell-sorted copy in HBM.
 *
 * Reference interfaces replaced (paths relative to the LAMMPS-LMFF tree):
 *   lmff_settings        PairHybrid::settings -> PairILPGrapheneHBN::settings
 *                        (Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp:132-149)
 *   lmff_coeff_tersoff   PairTersoff::coeff + read_file + setup_params
 *                        (Baseline/MANYBODY/pair_tersoff.cpp:293-358, 392-580)
 *   lmff_coeff_ilp       PairILPGrapheneHBN::coeff + read_file (pair_ilp_graphene_hbn.cpp:155-219, 246-387)
 *   lmff_neighbor        "neighbor <skin> bin" + "neigh_modify every/delay/check"
 *                        (LMFF/neighbor.cpp:87-91, 240-317, 1947-2035)
 *   lmff_pair_compute    Pair::compute(eflag, vflag) of "hybrid/overlay tersoff ilp/graphene/hbn"
 *                        (LMFF/pair_hybrid.cpp:93-210; pair_tersoff.cpp:86-266;
 *                        pair_ilp_graphene_hbn.cpp:428-675), with the SWLMFF
 *                        host->device precedent of SWLMFF/pair_lmff.cpp:37-43,376-445
 *   lmff_run             Verlet::run with fix nve / fix langevin on device
 *                        (LMFF/verlet.cpp:220-350, LMFF/fix_nve.cpp:64-143,
 *                        LMFF/fix_langevin.cpp:447-760; SWLMFF/fix_nve_sw64.c:7-161)
 */
#ifndef LMFF_HIP_H
#define LMFF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lmff_ctx lmff_ctx;

enum {
  LMFF_OK = 0,
  LMFF_ERR_ARG = 1,      /* bad argument or call order */
  LMFF_ERR_FILE = 2,     /* potential file missing or malformed */
  LMFF_ERR_PARAM = 3,    /* illegal / duplicate / missing potential entry */
  LMFF_ERR_HIP = 4,      /* HIP runtime failure */
  LMFF_ERR_NORMAL = 5,   /* >3 normal neighbours or zero normal (ilp.cpp:739,887,1030,1070) */
  LMFF_ERR_SHORT = 6,    /* too many Tersoff neighbours within cutmax */
  LMFF_ERR_NONFINITE = 7,/* non-numeric coordinates (domain.cpp:514) */
  LMFF_ERR_MEMORY = 8,   /* HBM budget exceeded */
  LMFF_ERR_BOX = 9,      /* periodic side not above twice the neighbor cutoff */
  LMFF_ERR_STATE = 10    /* no such system / not set up */
};

enum { LMFF_FP64 = 0, LMFF_MIXED = 1 };
enum { LMFF_NVE = 0, LMFF_LANGEVIN = 1 };

/* Context: owns one HIP stream on `device` and all HBM state. */
int lmff_create(lmff_ctx **ctx, int device);
void lmff_destroy(lmff_ctx *ctx);
const char *lmff_last_error(const lmff_ctx *ctx);
const char *lmff_version(void);

/* pair_style hybrid/overlay tersoff ilp/graphene/hbn <cut_global> [tap_flag]
 * ntypes = number of LAMMPS atom types (1-based types in all arrays). */
int lmff_settings(lmff_ctx *ctx, int ntypes, double ilp_cut_global, int tap_flag);
/* pair_coeff * * tersoff <file> e1 .. e_ntypes   ("NULL" leaves a type out) */
int lmff_coeff_tersoff(lmff_ctx *ctx, const char *file, const char *const *elements);
/* pair_coeff * * ilp/graphene/hbn <file> e1 .. e_ntypes */
int lmff_coeff_ilp(lmff_ctx *ctx, const char *file, const char *const *elements);
/* neighbor <skin> bin ; neigh_modify every <every> delay <delay> check <yes=1|no=0> */
int lmff_neighbor(lmff_ctx *ctx, double skin, int every, int delay, int check);
/* PairTersoff::init_one's cutoff (Baseline/MANYBODY/pair_tersoff.cpp:383-388): the largest
 * R + D of the Tersoff entries the element map uses (after lmff_coeff_tersoff). */
int lmff_tersoff_cutmax(lmff_ctx *ctx, double *cutmax);
/* LMFF_FP64 or LMFF_MIXED (FP32 pair/bond-order math, FP64 displacements and accumulation) */
int lmff_precision(lmff_ctx *ctx, int mode);
/* mass <type> <value> for all types (index type-1) */
int lmff_masses(lmff_ctx *ctx, const double *mass_per_type);

/* ---- Pair plugin entry: one force evaluation on host arrays ---------------------------
 * Semantics of Pair::compute(eflag, vflag) for hybrid/overlay tersoff ilp/graphene/hbn on
 * nlocal owned atoms of an orthogonal periodic box (no ghosts: images are handled inside).
 * eflag bit0 = global energy, bit1 = per-atom energy; vflag%4 != 0 = global virial,
 * vflag&4 = per-atom virial.  Outputs ACCUMULATE (+=) into f[n][3], eatom[n], vatom[n][6]
 * as LAMMPS pair styles do; eng_vdwl and virial[6] (xx yy zz xy xz yz) and pvector[2]
 * (ILP vdW, ILP repulsion energies) are overwritten.  new_list != 0 forces a neighbor
 * rebuild (the caller's neighbor->ago == 0); the first call always builds.  Between builds
 * the atoms must come in the same order (LAMMPS reorders them only when it reneighbours);
 * positions need not be wrapped: the library applies the wrap of its last build to them.
 * Any output pointer may be NULL. */
int lmff_pair_compute(lmff_ctx *ctx, int64_t nlocal, const double *x, const int32_t *type,
                      const int64_t *tag, const int64_t *mol, const double boxlo[3],
                      const double boxhi[3], const int periodic[3], int new_list, int eflag,
                      int vflag, double *f, double *eng_vdwl, double virial[6], double *eatom,
                      double *vatom, double pvector[2]);

/* ---- Allocation-free Pair::compute on device arrays ------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same force
 * evaluation as lmff_pair_compute, on DEVICE arrays: x[n][3], type[n] (1-based), tag[n], mol[n]
 * in, f[n][3] accumulated (+=) in the caller's order, globals[9] overwritten with {eng_vdwl,
 * virial[6] (xx yy zz xy xz yz), pvector[2]} (either output may be NULL).  Every call builds
 * its neighbor lists afresh; per-atom energy/virial (eflag bit1, vflag&4) are refused.  All
 * working arrays are carved out of one fixed device array of lmff_noalloc_scratch_bytes()
 * (256 MiB) per GPU, held until the process ends: about 50,000 atoms of a bilayer, fewer for
 * thicker stacks (their ILP rows are longer); larger inputs fail with LMFF_ERR_MEMORY.  The call
 * makes no hipMalloc/hipFree or allocating library call; it synchronises the context's stream.
 * Calls on one GPU share that scratch: calls that overlap in time on one GPU (two contexts, two
 * host threads) race.  pair_style, pair_coeff and masses must be complete (LMFF_ERR_STATE). */
int lmff_pair_compute_noalloc(lmff_ctx *ctx, int64_t nlocal, const double *x, const int32_t *type,
                              const int64_t *tag, const int64_t *mol, const double boxlo[3],
                              const double boxhi[3], const int periodic[3], int eflag, int vflag,
                              double *f, double *globals);
int64_t lmff_noalloc_scratch_bytes(void);

/* ---- Device-resident systems (standalone runs and the batched job server) ------------
 * A context holds any number of independent systems packed into one atom array; every
 * kernel launch covers all of them.  x, v are [n][3]; v may be NULL (zero). */
int lmff_system_add(lmff_ctx *ctx, int64_t natoms, const double boxlo[3], const double boxhi[3],
                    const int periodic[3], const int32_t *type, const int64_t *tag,
                    const int64_t *mol, const double *x, const double *v, int *sys_id);
int lmff_system_remove(lmff_ctx *ctx, int sys_id);
/* fix nve (style LMFF_NVE) or fix langevin T T damp seed + fix nve (LMFF_LANGEVIN) */
int lmff_system_integrator(lmff_ctx *ctx, int sys_id, int style, double dt, double t_target,
                           double damp, uint64_t seed);
int lmff_system_natoms(lmff_ctx *ctx, int sys_id, int64_t *natoms);
/* copies the system's state back in the caller's original atom order (any may be NULL) */
int lmff_system_get(lmff_ctx *ctx, int sys_id, double *x, double *v, double *f);
int lmff_system_set(lmff_ctx *ctx, int sys_id, const double *x, const double *v);
/* Image flags [n][3] in the caller's atom order (atom->image unpacked, LMFF/atom.h:57): the
 * device counts box crossings when it wraps positions at a neighbor-list build, as
 * Domain::pbc does (LMFF/domain.cpp:501); unwrapped x = x + image * box length (Domain::unmap,
 * LMFF/domain.cpp:1534).  A new system starts with image 0; set accepts -512..511. */
int lmff_system_get_image(lmff_ctx *ctx, int sys_id, int32_t *image);
int lmff_system_set_image(lmff_ctx *ctx, int sys_id, const int32_t *image);
/* per-atom energy [n] and virial [n][6] of the last evaluation with those flags */
int lmff_system_peratom(lmff_ctx *ctx, int sys_id, double *eatom, double *vatom);
/* Step counter of a system (the Langevin noise is keyed by it); the reset_timestep command
 * (LMFF/input.cpp:1865) so that a run restarted from write_data continues the stream. */
int lmff_system_set_step(lmff_ctx *ctx, int sys_id, int64_t step);
/* When enabled, the last step of every lmff_run evaluates with eflag_atom and vflag_atom set,
 * as LAMMPS does on the steps a compute pe/atom or stress/atom is invoked
 * (Integrate::ev_set, LMFF/integrate.cpp:131); lmff_system_peratom then returns that step's tallies
 * and the forces stay the integrator's (pair + thermostat). */
int lmff_peratom_on_last_step(lmff_ctx *ctx, int enable);

/* ---- Energy minimisation (min_style fire, LMFF/min_fire.cpp; minimize etol ftol maxiter
 * maxeval, LMFF/minimize.cpp) --------------------------------------------------------------
 * Every resident system is relaxed at once, each with its own FIRE state (semi-implicit Euler
 * integrator, half step back), starting from its integrator timestep; lists follow Min::init
 * (every 1, delay 0, check yes).  Velocities are zeroed at the start as MinFire::setup_style
 * does; each system's step counter advances by its iterations.  FIRE stops on etol, ftol,
 * maxiter or too many consecutive v.f <= 0 steps (it does not count force evaluations, so the
 * maxeval of the minimize command has no effect, as in the reference). */
typedef struct {
  double dmax;              /* max displacement per iteration (A), default 0.1 */
  int delaystep;            /* default 20 */
  double dtgrow, dtshrink;  /* 1.1, 0.5 */
  double alpha0, alphashrink; /* 0.25, 0.99 */
  double tmax, tmin;        /* dtmax = tmax dt, dtmin = tmin dt: 10, 0.02 */
  int halfstepback;         /* 1 */
  int delaystep_start;      /* 1 */
  int max_vdotf_negatif;    /* 2000 */
  int norm;                 /* force norm for ftol: 0 two, 1 max, 2 inf (min_modify norm) */
} lmff_fire_params;

enum {
  LMFF_STOP_MAXITER = 0, LMFF_STOP_MAXEVAL = 1, LMFF_STOP_ETOL = 2, LMFF_STOP_FTOL = 3, LMFF_STOP_DOWNHILL = 4,
  LMFF_STOP_ZEROALPHA = 5, LMFF_STOP_ZEROFORCE = 6, LMFF_STOP_ZEROQUAD = 7, LMFF_STOP_MAXVDOTF = 11
};

typedef struct {
  int stop;                 /* LMFF_STOP_* (the reference's stop_condition codes) */
  int64_t iterations, evaluations;
  double e_initial, e_next_to_last, e_final;   /* pair energy (eV) */
  double fnorm2_initial, fnorm2_final;         /* Euclidean force norm (eV/A) */
  double fnorminf_initial, fnorminf_final;     /* largest force component */
  double dt_final;
} lmff_min_stats;

/* LAMMPS' FIRE defaults (Min::Min, LMFF/min.cpp:54-71) */
void lmff_fire_defaults(lmff_fire_params *p);
/* params NULL: defaults; out[k] for the k-th resident system (in order of addition) */
int lmff_minimize(lmff_ctx *ctx, double etol, double ftol, int64_t maxiter, const lmff_fire_params *params,
                  lmff_min_stats *out, int out_cap);

/* min_style cg (the reference's default) with the quadratic line search: MinCG::iterate and
 * MinLineSearch::linemin_quadratic (LMFF/min_cg.cpp:40-190, LMFF/min_linesearch.cpp:324-503);
 * dmax is min_modify dmax (0.1), norm as in lmff_fire_params.  The context must hold exactly one
 * system.  Velocities are not touched; the step counter advances by the iterations. */
int lmff_minimize_cg(lmff_ctx *ctx, double etol, double ftol, int64_t maxiter, int64_t maxeval, double dmax,
                     int norm, lmff_min_stats *out);

/* Thermo record of one system (metal units). */
typedef struct {
  int64_t step;
  double pe;           /* eng_vdwl of the pair: Tersoff + ILP */
  double ke;
  double e_vdw;        /* pvector[0] */
  double e_rep;        /* pvector[1] */
  double e_tersoff;
  double temp;         /* 2 ke / ((3N-3) kB) */
  double virial[6];    /* pair virial, F.r convention */
} lmff_thermo;

/* Evaluate forces (and energy/virial per flags) of every resident system at the current
 * positions, rebuilding neighbor lists if positions were set since the last build. */
int lmff_evaluate(lmff_ctx *ctx, int eflag, int vflag);
int lmff_system_thermo(lmff_ctx *ctx, int sys_id, lmff_thermo *out);

/* Advance every resident system by nsteps Velocity-Verlet steps (forces from the previous
 * evaluation are reused as LAMMPS does after setup).  Energies are evaluated every
 * thermo_every steps (0 = only at the end); records go to out[] in system-major order,
 * out[s * nrec + r], nrec = nsteps / thermo_every (+1 if nsteps is not a multiple);
 * out may be NULL.  use_graph != 0 replays the steps between neighbor rebuilds as a
 * captured hipGraph. */
int lmff_run(lmff_ctx *ctx, int64_t nsteps, int64_t thermo_every, int use_graph, lmff_thermo *out,
             int64_t out_capacity);

/* Neighbor list builds and dangerous builds since the last reset (Neighbor::ncalls / ndanger,
 * LMFF/neighbor.cpp:2026,2041, reported by Finish, LMFF/finish.cpp:814-819): a dangerous build
 * is one the distance check (neigh_modify check yes) triggered on the first step it was
 * allowed to, so some atom may have moved past half the skin before the list caught up.
 * reset != 0 zeroes both after reading (Neighbor::init does so before every run). */
int lmff_neighbor_stats(lmff_ctx *ctx, int reset, int64_t *builds, int64_t *dangerous);

/* Device-side timing of the last lmff_run: total ms and per-kernel-family averages
 * (index: 0 list build, 1 normals, 2 ILP pair, 3 Tersoff bonds, 4 force gather,
 *  5 integrate, 6 reductions).  Requires lmff_profile(ctx, 1) before the run. */
int lmff_profile(lmff_ctx *ctx, int enable);
int lmff_profile_read(lmff_ctx *ctx, double *ms_total, double *ms_per_family, int64_t *launches_per_family);

/* HBM accounting: bytes currently held and the per-atom estimate used for admission. */
int lmff_memory(lmff_ctx *ctx, int64_t *bytes_in_use, double *bytes_per_atom_estimate);
/* Bytes a new system of natoms would add: every per-atom array, with list rows at the
 * context's current capacities or, if larger, what a stack of nlayers_hint sheets needs. */
int64_t lmff_system_bytes(lmff_ctx *ctx, int64_t natoms, int nlayers_hint);

/* ILP list of one atom (caller order index) as sorted (caller-order j, sx, sy, sz) rows, for
 * neighbor-set parity checks; returns the count in *n (rows up to capacity). */
int lmff_debug_ilp_list(lmff_ctx *ctx, int sys_id, int64_t atom, int32_t *rows, int capacity, int *n);
/* Counts over every resident system after the last neighbor build (it builds one if due):
 * out[0] ILP list entries, out[1] ILP ordered pairs within the ILP cutoff, out[2] Tersoff
 * neighbours within cutmax (tallied at the last force evaluation), out[3] short-list entries. */
int lmff_pair_stats(lmff_ctx *ctx, int64_t out[4]);
/* Normals (caller order): normal[n][3], count[n]. */
int lmff_debug_normals(lmff_ctx *ctx, int sys_id, double *normal, int32_t *count);

#ifdef __cplusplus
}
#endif

#endif /* LMFF_HIP_H */
