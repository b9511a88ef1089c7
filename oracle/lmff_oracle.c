/*
 * lmff_oracle.c -- TEST INFRASTRUCTURE ONLY (never part of the product path).
 *
 * A CPU FP64 restatement of the reference's LMFF force path, written in the
 * reference's own "push" form, used as the parity checker for the HIP
 * kernels and as the on-host CPU baseline.  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg load it.
 *
 * What it restates (all paths relative to the reference tree):
 *   - potential files:   Baseline/MANYBODY/pair_tersoff.cpp:392-580 (read_file,
 *                        setup_params), Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp:246-387
 *   - neighbor list:     LMFF/npair_full_bin_ghost.cpp:35-155 with ghost images
 *                        x_ghost = x_j + pbc*prd (LMFF/MOLECULE/atom_vec_full.cpp:402-412)
 *                        and the test rsq <= (cut+skin)^2
 *   - Tersoff:           Baseline/MANYBODY/pair_tersoff.cpp:86-266, 584-811 and
 *                        pair_tersoff.h:90-107
 *   - ILP:               Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp:449-1075 and
 *                        pair_ilp_graphene_hbn.h:80-117 (calc_Tap, calc_dTap)
 *   - tallies:           LMFF/pair.cpp:914 (ev_tally), 1061 (ev_tally_xyz),
 *                        1436 (v_tally3), 1555-1604 (virial_fdotr_compute);
 *                        LMFF/pair_hybrid.cpp:93-210 (global virial by F.r)
 *   - integrator:        LMFF/fix_nve.cpp:51-143, LMFF/domain.cpp:501 (pbc wrap)
 *   - the fused LMFF CPU path (orc_compute_lmff, the CPU baseline's timed path):
 *                        LMFF/pair_lmff.cpp:1121-1613 without the quirks of its recalc,
 *                        filtered short list and list-length-scaled normal
 *
 * There are no ghost atoms: every list entry carries the periodic image
 * shift s of its atom, and "x[j]" of a ghost is formed exactly as LAMMPS
 * forms it, x_j + s*prd.  Forces and per-atom tallies on an image land on
 * the owning atom (what reverse communication does).  The F.r global virial
 * over owned+ghost atoms is rebuilt as sum_i x_i F_i plus sum over image
 * contributions of (s*prd) (x) f.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms of Tap and of the
 * perfect-lattice Tersoff energy, an independent numpy brute-force ILP/Tersoff
 * energy, finite-difference forces, the strain-derivative virial, brute-force
 * neighbor sets, and NVE energy conservation.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXTYPES 8
#define MAXELEM 8
#define MAXLINE 1024
#define MY_PI2 1.57079632679489661923
#define MY_PI4 0.78539816339744830962
#define THIRD (1.0 / 3.0)

typedef struct {
  int ielement, jelement, kelement, powermint;
  double powerm, gamma, lam3, c, d, h, powern, beta, lam2, bigb, bigr, bigd, lam1, biga;
  double cut, cutsq, c1, c2, c3, c4;
} TParam;

typedef struct {
  int ielement, jelement;
  double z0, alpha, delta, epsilon, C, d, sR, reff, C6, S, rcut, delta2inv, lambda, seff;
} IParam;

typedef struct {
  int ntypes;
  /* Tersoff */
  int t_nelem;
  char t_elem[MAXELEM][64];
  int t_map[MAXTYPES + 1];
  TParam *tp;
  int ntp;
  int t_e2p[MAXELEM][MAXELEM][MAXELEM];
  double cutmax;
  /* ILP */
  int i_nelem;
  char i_elem[MAXELEM][64];
  int i_map[MAXTYPES + 1];
  IParam *ip;
  int nip;
  int i_e2p[MAXELEM][MAXELEM];
  double cutILPsq[MAXELEM][MAXELEM];
  double ilp_cut;
  int tap_flag;
  double ilp_cutsq[MAXTYPES + 1][MAXTYPES + 1];
  /* neighbor settings */
  double skin, cutneighsq;
  /* atoms */
  int n;
  double *x;  /* [n][3] */
  int *type;
  long long *tag, *mol;
  double lo[3], hi[3], prd[3];
  int per[3];
  /* full neighbor list (CSR), entry = (j, image code) */
  int *offs;
  int *ent_j;
  unsigned char *ent_s;
  long long nent, cap;
  /* scratch */
  int nthreads;
  char err[512];
} Orc;

static int count_words(const char *line) {
  int n = 0;
  const char *p = line;
  while (*p) {
    while (*p && strchr(" \t\n\r\f", *p)) p++;
    if (!*p) break;
    n++;
    while (*p && !strchr(" \t\n\r\f", *p)) p++;
  }
  return n;
}

/* reads the next logical entry with exactly nwant words; returns 1, 0 at EOF, -1 on format error.
 * Mirrors the line-concatenation rule of read_file (pair_tersoff.cpp:420-450). */
static int next_entry(FILE *fp, char *line, int nwant, char **words) {
  char *ptr;
  for (;;) {
    if (!fgets(line, MAXLINE, fp)) return 0;
    if ((ptr = strchr(line, '#'))) *ptr = '\0';
    int nwords = count_words(line);
    if (nwords == 0) continue;
    while (nwords < nwant) {
      int n = (int)strlen(line);
      if (!fgets(&line[n], MAXLINE - n, fp)) return -1;
      if ((ptr = strchr(line, '#'))) *ptr = '\0';
      nwords = count_words(line);
    }
    if (nwords != nwant) return -1;
    int k = 0;
    words[k++] = strtok(line, " \t\n\r\f");
    while ((words[k++] = strtok(NULL, " \t\n\r\f"))) continue;
    return 1;
  }
}

static int map_elements(int ntypes, const char *spec, int *map, char elem[MAXELEM][64], int *nelem) {
  /* spec: whitespace separated element names, one per atom type ("NULL" allowed) */
  char buf[1024];
  strncpy(buf, spec, sizeof(buf) - 1);
  buf[sizeof(buf) - 1] = 0;
  char *tok = strtok(buf, " \t");
  int t = 1;
  *nelem = 0;
  while (tok && t <= ntypes) {
    if (strcmp(tok, "NULL") == 0) {
      map[t] = -1;
    } else {
      int j;
      for (j = 0; j < *nelem; j++)
        if (strcmp(tok, elem[j]) == 0) break;
      map[t] = j;
      if (j == *nelem) {
        if (*nelem >= MAXELEM) return -1;
        strncpy(elem[j], tok, 63);
        (*nelem)++;
      }
    }
    tok = strtok(NULL, " \t");
    t++;
  }
  return (t == ntypes + 1 && tok == NULL) ? 0 : -1;
}

/* pair_tersoff.cpp:392-532 and 536-580 */
static int read_tersoff(Orc *o, const char *file) {
  FILE *fp = fopen(file, "r");
  if (!fp) { snprintf(o->err, sizeof o->err, "Cannot open Tersoff potential file %s", file); return -1; }
  char line[MAXLINE];
  char *words[18];
  int maxp = 0;
  o->ntp = 0;
  for (;;) {
    int r = next_entry(fp, line, 17, words);
    if (r == 0) break;
    if (r < 0) { fclose(fp); snprintf(o->err, sizeof o->err, "Incorrect format in Tersoff potential file"); return -1; }
    int ie, je, ke;
    for (ie = 0; ie < o->t_nelem; ie++) if (!strcmp(words[0], o->t_elem[ie])) break;
    if (ie == o->t_nelem) continue;
    for (je = 0; je < o->t_nelem; je++) if (!strcmp(words[1], o->t_elem[je])) break;
    if (je == o->t_nelem) continue;
    for (ke = 0; ke < o->t_nelem; ke++) if (!strcmp(words[2], o->t_elem[ke])) break;
    if (ke == o->t_nelem) continue;
    if (o->ntp == maxp) { maxp += 16; o->tp = (TParam *)realloc(o->tp, maxp * sizeof(TParam)); }
    TParam *p = &o->tp[o->ntp];
    memset(p, 0, sizeof *p);
    p->ielement = ie; p->jelement = je; p->kelement = ke;
    p->powerm = atof(words[3]); p->gamma = atof(words[4]); p->lam3 = atof(words[5]);
    p->c = atof(words[6]); p->d = atof(words[7]); p->h = atof(words[8]);
    p->powern = atof(words[9]); p->beta = atof(words[10]); p->lam2 = atof(words[11]);
    p->bigb = atof(words[12]); p->bigr = atof(words[13]); p->bigd = atof(words[14]);
    p->lam1 = atof(words[15]); p->biga = atof(words[16]);
    p->powermint = (int)p->powerm;
    if (p->c < 0.0 || p->d < 0.0 || p->powern < 0.0 || p->beta < 0.0 || p->lam2 < 0.0 ||
        p->bigb < 0.0 || p->bigr < 0.0 || p->bigd < 0.0 || p->bigd > p->bigr || p->lam1 < 0.0 ||
        p->biga < 0.0 || p->powerm - p->powermint != 0.0 ||
        (p->powermint != 3 && p->powermint != 1) || p->gamma < 0.0) {
      fclose(fp); snprintf(o->err, sizeof o->err, "Illegal Tersoff parameter"); return -1;
    }
    o->ntp++;
  }
  fclose(fp);
  for (int i = 0; i < o->t_nelem; i++)
    for (int j = 0; j < o->t_nelem; j++)
      for (int k = 0; k < o->t_nelem; k++) {
        int n = -1;
        for (int m = 0; m < o->ntp; m++)
          if (i == o->tp[m].ielement && j == o->tp[m].jelement && k == o->tp[m].kelement) {
            if (n >= 0) { snprintf(o->err, sizeof o->err, "Potential file has duplicate entry"); return -1; }
            n = m;
          }
        if (n < 0) { snprintf(o->err, sizeof o->err, "Potential file is missing an entry"); return -1; }
        o->t_e2p[i][j][k] = n;
      }
  o->cutmax = 0.0;
  for (int m = 0; m < o->ntp; m++) {
    TParam *p = &o->tp[m];
    p->cut = p->bigr + p->bigd;
    p->cutsq = p->cut * p->cut;
    p->c1 = pow(2.0 * p->powern * 1.0e-16, -1.0 / p->powern);
    p->c2 = pow(2.0 * p->powern * 1.0e-8, -1.0 / p->powern);
    p->c3 = 1.0 / p->c2;
    p->c4 = 1.0 / p->c1;
    if (p->cut > o->cutmax) o->cutmax = p->cut;
  }
  return 0;
}

/* pair_ilp_graphene_hbn.cpp:246-387 */
static int read_ilp(Orc *o, const char *file) {
  FILE *fp = fopen(file, "r");
  if (!fp) { snprintf(o->err, sizeof o->err, "Cannot open ILP potential file %s", file); return -1; }
  char line[MAXLINE];
  char *words[14];
  int maxp = 0;
  o->nip = 0;
  for (;;) {
    int r = next_entry(fp, line, 13, words);
    if (r == 0) break;
    if (r < 0) { fclose(fp); snprintf(o->err, sizeof o->err, "Insufficient format in ILP potential file"); return -1; }
    int ie, je;
    for (ie = 0; ie < o->i_nelem; ie++) if (!strcmp(words[0], o->i_elem[ie])) break;
    if (ie == o->i_nelem) continue;
    for (je = 0; je < o->i_nelem; je++) if (!strcmp(words[1], o->i_elem[je])) break;
    if (je == o->i_nelem) continue;
    if (o->nip == maxp) { maxp += 16; o->ip = (IParam *)realloc(o->ip, maxp * sizeof(IParam)); }
    IParam *p = &o->ip[o->nip];
    memset(p, 0, sizeof *p);
    p->ielement = ie; p->jelement = je;
    p->z0 = atof(words[2]); p->alpha = atof(words[3]); p->delta = atof(words[4]);
    p->epsilon = atof(words[5]); p->C = atof(words[6]); p->d = atof(words[7]);
    p->sR = atof(words[8]); p->reff = atof(words[9]); p->C6 = atof(words[10]);
    p->S = atof(words[11]); p->rcut = atof(words[12]);
    double meV = 1e-3 * p->S;
    p->C *= meV; p->C6 *= meV; p->epsilon *= meV;
    p->delta2inv = pow(p->delta, -2.0);
    p->lambda = p->alpha / p->z0;
    p->seff = p->sR * p->reff;
    o->nip++;
  }
  fclose(fp);
  for (int i = 0; i < o->i_nelem; i++)
    for (int j = 0; j < o->i_nelem; j++) {
      int n = -1;
      for (int m = 0; m < o->nip; m++)
        if (i == o->ip[m].ielement && j == o->ip[m].jelement) {
          if (n >= 0) { snprintf(o->err, sizeof o->err, "Potential file has duplicate entry"); return -1; }
          n = m;
        }
      if (n < 0) { snprintf(o->err, sizeof o->err, "Potential file is missing an entry"); return -1; }
      o->i_e2p[i][j] = n;
      o->cutILPsq[i][j] = o->ip[n].rcut * o->ip[n].rcut;
    }
  return 0;
}

/* ----------------------------------------------------------------------- */

Orc *orc_create(int ntypes, const char *tersoff_file, const char *tersoff_elems,
                const char *ilp_file, const char *ilp_elems, double ilp_cut, int tap_flag,
                double skin, char *err, int errlen) {
  Orc *o = (Orc *)calloc(1, sizeof(Orc));
  o->ntypes = ntypes;
  o->ilp_cut = ilp_cut;
  o->tap_flag = tap_flag;
  o->skin = skin;
  o->nthreads = 1;
  int rc = 0;
  if (ntypes < 1 || ntypes > MAXTYPES) { snprintf(o->err, sizeof o->err, "bad ntypes"); rc = -1; }
  if (!rc && map_elements(ntypes, tersoff_elems, o->t_map, o->t_elem, &o->t_nelem)) {
    snprintf(o->err, sizeof o->err, "Incorrect args for pair coefficients"); rc = -1;
  }
  if (!rc && map_elements(ntypes, ilp_elems, o->i_map, o->i_elem, &o->i_nelem)) {
    snprintf(o->err, sizeof o->err, "Incorrect args for pair coefficients"); rc = -1;
  }
  if (!rc) rc = read_tersoff(o, tersoff_file);
  if (!rc) rc = read_ilp(o, ilp_file);
  if (!rc) {
    /* cutsq of the ILP sub-style = cut_global^2 (init_one returns cut[i][j]) */
    for (int i = 1; i <= ntypes; i++)
      for (int j = 1; j <= ntypes; j++) o->ilp_cutsq[i][j] = ilp_cut * ilp_cut;
    /* hybrid cutoff = max over sub-styles (pair_hybrid.cpp:642-685); list cutoff (cut+skin)^2 */
    double cutforce = ilp_cut > o->cutmax ? ilp_cut : o->cutmax;
    o->cutneighsq = (cutforce + skin) * (cutforce + skin);
  }
  if (rc) {
    if (err) snprintf(err, errlen, "%s", o->err);
    free(o->tp); free(o->ip); free(o);
    return NULL;
  }
  return o;
}

void orc_destroy(Orc *o) {
  if (!o) return;
  free(o->tp); free(o->ip); free(o->x); free(o->type); free(o->tag); free(o->mol);
  free(o->offs); free(o->ent_j); free(o->ent_s);
  free(o);
}

const char *orc_error(Orc *o) { return o->err; }
double orc_cutmax(Orc *o) { return o->cutmax; }
double orc_cutneighsq(Orc *o) { return o->cutneighsq; }
void orc_set_threads(Orc *o, int nt) { o->nthreads = nt > 0 ? nt : 1; }

/* exposes parsed tables so tests can compare the product's parser with this one */
int orc_tersoff_param(Orc *o, int m, double *out /* 22 */) {
  if (m < 0 || m >= o->ntp) return -1;
  TParam *p = &o->tp[m];
  double v[22] = {p->ielement, p->jelement, p->kelement, p->powerm, p->gamma, p->lam3, p->c, p->d,
                  p->h, p->powern, p->beta, p->lam2, p->bigb, p->bigr, p->bigd, p->lam1, p->biga,
                  p->cutsq, p->c1, p->c2, p->c3, p->c4};
  memcpy(out, v, sizeof v);
  return o->ntp;
}

int orc_ilp_param(Orc *o, int m, double *out /* 16 */) {
  if (m < 0 || m >= o->nip) return -1;
  IParam *p = &o->ip[m];
  double v[16] = {p->ielement, p->jelement, p->z0, p->alpha, p->delta, p->epsilon, p->C, p->d,
                  p->sR, p->reff, p->C6, p->S, p->rcut, p->delta2inv, p->lambda, p->seff};
  memcpy(out, v, sizeof v);
  return o->nip;
}

int orc_set_atoms(Orc *o, int n, const double *x, const int *type, const long long *tag,
                  const long long *mol, const double *lo, const double *hi, const int *periodic) {
  free(o->x); free(o->type); free(o->tag); free(o->mol);
  o->n = n;
  o->x = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
  o->type = (int *)malloc(sizeof(int) * (n ? n : 1));
  o->tag = (long long *)malloc(sizeof(long long) * (n ? n : 1));
  o->mol = (long long *)malloc(sizeof(long long) * (n ? n : 1));
  memcpy(o->x, x, sizeof(double) * 3 * n);
  memcpy(o->type, type, sizeof(int) * n);
  memcpy(o->tag, tag, sizeof(long long) * n);
  memcpy(o->mol, mol, sizeof(long long) * n);
  for (int d = 0; d < 3; d++) {
    o->lo[d] = lo[d]; o->hi[d] = hi[d]; o->prd[d] = hi[d] - lo[d]; o->per[d] = periodic[d];
  }
  for (int i = 0; i < n; i++)
    if (type[i] < 1 || type[i] > o->ntypes) { snprintf(o->err, sizeof o->err, "Invalid atom type"); return -1; }
  return 0;
}

void orc_get_x(Orc *o, double *x) { memcpy(x, o->x, sizeof(double) * 3 * o->n); }
void orc_put_x(Orc *o, const double *x) { memcpy(o->x, x, sizeof(double) * 3 * o->n); }

/* Domain::pbc (LMFF/domain.cpp:501-560) for the periodic dimensions */
void orc_pbc(Orc *o) {
  for (int i = 0; i < o->n; i++)
    for (int d = 0; d < 3; d++) {
      if (!o->per[d]) continue;
      double *xi = &o->x[3 * i + d];
      if (*xi < o->lo[d]) *xi += o->prd[d];
      if (*xi >= o->hi[d]) {
        *xi -= o->prd[d];
        if (*xi < o->lo[d]) *xi = o->lo[d];
      }
    }
}

static inline int scode(int sx, int sy, int sz) { return (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1); }
static inline void sdecode(int c, int *s) { s[0] = c % 3 - 1; s[1] = (c / 3) % 3 - 1; s[2] = c / 9 - 1; }

/* image coordinate of list entry: x_j + s*prd (atom_vec_full.cpp:402-412) */
static inline void ghost_x(const Orc *o, int j, int code, double *xg) {
  int s[3];
  sdecode(code, s);
  for (int d = 0; d < 3; d++) {
    double dx = s[d] * o->prd[d];
    xg[d] = o->x[3 * j + d] + dx;
  }
}

/* Full neighbor list over owned atoms with periodic images, criterion rsq <= cutneighsq,
 * delx = xtmp - x_ghost (npair_full_bin_ghost.cpp:97-102).  Bins of half the list cutoff and
 * a +-2 bin stencil in every dimension; periodic dimensions map a stencil bin outside the grid
 * onto its periodic image and record the image shift.  Box sides must exceed twice the list
 * cutoff so that each atom appears at most once in another atom's list. */
int orc_build_list(Orc *o) {
  const int n = o->n;
  const double cutn = sqrt(o->cutneighsq);
  double glo[3], bs[3];
  int nb[3];
  for (int d = 0; d < 3; d++) {
    if (o->per[d]) {
      if (o->prd[d] <= 2.0 * cutn) { snprintf(o->err, sizeof o->err, "periodic box side must exceed twice the neighbor cutoff"); return -1; }
      glo[d] = o->lo[d];
      nb[d] = (int)(o->prd[d] / (0.5 * cutn));
      if (nb[d] < 1) nb[d] = 1;
      bs[d] = o->prd[d] / nb[d];
    } else {
      double mn = 1e300, mx = -1e300;
      for (int i = 0; i < n; i++) { double v = o->x[3 * i + d]; if (v < mn) mn = v; if (v > mx) mx = v; }
      if (n == 0) { mn = 0; mx = 0; }
      glo[d] = mn;
      nb[d] = (int)((mx - mn) / (0.5 * cutn));
      if (nb[d] < 1) nb[d] = 1;
      bs[d] = (mx - mn) / nb[d];
      if (bs[d] <= 0) bs[d] = 1.0;
    }
  }
  long long nbins = (long long)nb[0] * nb[1] * nb[2];
  int *head = (int *)malloc(sizeof(int) * nbins);
  int *next = (int *)malloc(sizeof(int) * (n ? n : 1));
  int *binof = (int *)malloc(sizeof(int) * (n ? n : 1));
  for (long long b = 0; b < nbins; b++) head[b] = -1;
  /* bins hold atoms in increasing index order */
  for (int i = n - 1; i >= 0; i--) {
    int c[3];
    for (int d = 0; d < 3; d++) {
      c[d] = (int)floor((o->x[3 * i + d] - glo[d]) / bs[d]);
      if (c[d] < 0) c[d] = 0;
      if (c[d] >= nb[d]) c[d] = nb[d] - 1;
    }
    int b = (c[2] * nb[1] + c[1]) * nb[0] + c[0];
    binof[i] = b;
    next[i] = head[b];
    head[b] = i;
  }
  free(o->offs);
  o->offs = (int *)malloc(sizeof(int) * (n + 1));
  if (o->cap == 0) {
    o->cap = 1024;
    o->ent_j = (int *)malloc(sizeof(int) * o->cap);
    o->ent_s = (unsigned char *)malloc(o->cap);
  }
  long long ne = 0;
  for (int i = 0; i < n; i++) {
    o->offs[i] = (int)ne;
    int b = binof[i];
    int c[3] = {b % nb[0], (b / nb[0]) % nb[1], b / (nb[0] * nb[1])};
    const double xt = o->x[3 * i], yt = o->x[3 * i + 1], zt = o->x[3 * i + 2];
    for (int dz = -2; dz <= 2; dz++)
      for (int dy = -2; dy <= 2; dy++)
        for (int dx = -2; dx <= 2; dx++) {
          int cc[3] = {c[0] + dx, c[1] + dy, c[2] + dz}, s[3] = {0, 0, 0}, skip = 0;
          for (int d = 0; d < 3; d++) {
            if (cc[d] < 0 || cc[d] >= nb[d]) {
              if (!o->per[d]) { skip = 1; break; }
              int w = cc[d] % nb[d];
              if (w < 0) w += nb[d];
              s[d] = (cc[d] - w) / nb[d];
              cc[d] = w;
            }
          }
          if (skip) continue;
          int code = scode(s[0], s[1], s[2]);
          int bb = (cc[2] * nb[1] + cc[1]) * nb[0] + cc[0];
          for (int j = head[bb]; j >= 0; j = next[j]) {
            if (j == i && code == 13) continue;
            double xg[3];
            ghost_x(o, j, code, xg);
            double delx = xt - xg[0];
            double dely = yt - xg[1];
            double delz = zt - xg[2];
            double rsq = delx * delx + dely * dely + delz * delz;
            if (rsq <= o->cutneighsq) {
              if (ne == o->cap) {
                o->cap *= 2;
                o->ent_j = (int *)realloc(o->ent_j, sizeof(int) * o->cap);
                o->ent_s = (unsigned char *)realloc(o->ent_s, o->cap);
              }
              o->ent_j[ne] = j;
              o->ent_s[ne] = (unsigned char)code;
              ne++;
            }
          }
        }
  }
  o->offs[n] = (int)ne;
  o->nent = ne;
  free(head); free(next); free(binof);
  return 0;
}

long long orc_list_size(Orc *o) { return o->nent; }
/* copies offsets [n+1], entries j and image shifts (sx,sy,sz) per entry */
void orc_get_list(Orc *o, int *offs, int *ent_j, int *ent_shift) {
  memcpy(offs, o->offs, sizeof(int) * (o->n + 1));
  for (long long e = 0; e < o->nent; e++) {
    ent_j[e] = o->ent_j[e];
    int s[3];
    sdecode(o->ent_s[e], s);
    ent_shift[3 * e] = s[0]; ent_shift[3 * e + 1] = s[1]; ent_shift[3 * e + 2] = s[2];
  }
}

/* ----------------------------------------------------------------------- */
/* Tersoff functions (pair_tersoff.cpp:584-811, pair_tersoff.h:90-107)       */

static inline double ters_fc(double r, const TParam *p) {
  if (r < p->bigr - p->bigd) return 1.0;
  if (r > p->bigr + p->bigd) return 0.0;
  return 0.5 * (1.0 - sin(MY_PI2 * (r - p->bigr) / p->bigd));
}
static inline double ters_fc_d(double r, const TParam *p) {
  if (r < p->bigr - p->bigd) return 0.0;
  if (r > p->bigr + p->bigd) return 0.0;
  return -(MY_PI4 / p->bigd) * cos(MY_PI2 * (r - p->bigr) / p->bigd);
}
static inline double ters_fa(double r, const TParam *p) {
  if (r > p->bigr + p->bigd) return 0.0;
  return -p->bigb * exp(-p->lam2 * r) * ters_fc(r, p);
}
static inline double ters_fa_d(double r, const TParam *p) {
  if (r > p->bigr + p->bigd) return 0.0;
  return p->bigb * exp(-p->lam2 * r) * (p->lam2 * ters_fc(r, p) - ters_fc_d(r, p));
}
static inline double ters_bij(double zeta, const TParam *p) {
  double tmp = p->beta * zeta;
  if (tmp > p->c1) return 1.0 / sqrt(tmp);
  if (tmp > p->c2) return (1.0 - pow(tmp, -p->powern) / (2.0 * p->powern)) / sqrt(tmp);
  if (tmp < p->c4) return 1.0;
  if (tmp < p->c3) return 1.0 - pow(tmp, p->powern) / (2.0 * p->powern);
  return pow(1.0 + pow(tmp, p->powern), -1.0 / (2.0 * p->powern));
}
static inline double ters_bij_d(double zeta, const TParam *p) {
  double tmp = p->beta * zeta;
  if (tmp > p->c1) return p->beta * -0.5 * pow(tmp, -1.5);
  if (tmp > p->c2)
    return p->beta * (-0.5 * pow(tmp, -1.5) *
                      (1.0 - (1.0 + 1.0 / (2.0 * p->powern)) * pow(tmp, -p->powern)));
  if (tmp < p->c4) return 0.0;
  if (tmp < p->c3) return -0.5 * p->beta * pow(tmp, p->powern - 1.0);
  double tmp_n = pow(tmp, p->powern);
  return -0.5 * pow(1.0 + tmp_n, -1.0 - (1.0 / (2.0 * p->powern))) * tmp_n / zeta;
}
static inline double ters_gijk(double costheta, const TParam *p) {
  const double c2 = p->c * p->c, d2 = p->d * p->d, hcth = p->h - costheta;
  return p->gamma * (1.0 + c2 / d2 - c2 / (d2 + hcth * hcth));
}
static inline double ters_gijk_d(double costheta, const TParam *p) {
  const double c2 = p->c * p->c, d2 = p->d * p->d, hcth = p->h - costheta;
  const double numerator = -2.0 * c2 * hcth;
  const double denominator = 1.0 / (d2 + hcth * hcth);
  return p->gamma * numerator * denominator * denominator;
}
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void ters_repulsive(const TParam *p, double rsq, double *fforce, int eflag, double *eng) {
  double r = sqrt(rsq);
  double tmp_fc = ters_fc(r, p), tmp_fc_d = ters_fc_d(r, p);
  double tmp_exp = exp(-p->lam1 * r);
  *fforce = -p->biga * tmp_exp * (tmp_fc_d - tmp_fc * p->lam1) / r;
  if (eflag) *eng = tmp_fc * p->biga * tmp_exp;
}
static double ters_zeta(const TParam *p, double rsqij, double rsqik, const double *delrij, const double *delrik) {
  double rij = sqrt(rsqij), rik = sqrt(rsqik);
  double costheta = dot3(delrij, delrik) / (rij * rik);
  double arg;
  if (p->powermint == 3) { double t = p->lam3 * (rij - rik); arg = t * t * t; }
  else arg = p->lam3 * (rij - rik);
  double ex_delr;
  if (arg > 69.0776) ex_delr = 1.e30;
  else if (arg < -69.0776) ex_delr = 0.0;
  else ex_delr = exp(arg);
  return ters_fc(rik, p) * ters_gijk(costheta, p) * ex_delr;
}
static void ters_force_zeta(const TParam *p, double rsq, double zeta_ij, double *fforce, double *prefactor,
                            int eflag, double *eng) {
  double r = sqrt(rsq);
  double fa = ters_fa(r, p), fa_d = ters_fa_d(r, p);
  double bij = ters_bij(zeta_ij, p);
  *fforce = 0.5 * bij * fa_d / r;
  *prefactor = -0.5 * fa * ters_bij_d(zeta_ij, p);
  if (eflag) *eng = 0.5 * bij * fa;
}
static void costheta_d(const double *rij_hat, double rij, const double *rik_hat, double rik,
                       double *dri, double *drj, double *drk) {
  double cos_theta = dot3(rij_hat, rik_hat);
  for (int d = 0; d < 3; d++) {
    drj[d] = (-cos_theta * rij_hat[d] + rik_hat[d]) * (1.0 / rij);
    drk[d] = (-cos_theta * rik_hat[d] + rij_hat[d]) * (1.0 / rik);
    dri[d] = -(drj[d] + drk[d]);
  }
}
static void ters_zetaterm_d(double prefactor, const double *rij_hat, double rij, const double *rik_hat, double rik,
                            double *dri, double *drj, double *drk, const TParam *p) {
  double fc = ters_fc(rik, p), dfc = ters_fc_d(rik, p), tmp;
  if (p->powermint == 3) { double t = p->lam3 * (rij - rik); tmp = t * t * t; }
  else tmp = p->lam3 * (rij - rik);
  double ex_delr;
  if (tmp > 69.0776) ex_delr = 1.e30;
  else if (tmp < -69.0776) ex_delr = 0.0;
  else ex_delr = exp(tmp);
  double ex_delr_d;
  if (p->powermint == 3) ex_delr_d = 3.0 * (p->lam3 * p->lam3 * p->lam3) * (rij - rik) * (rij - rik) * ex_delr;
  else ex_delr_d = p->lam3 * ex_delr;
  double cos_theta = dot3(rij_hat, rik_hat);
  double gijk = ters_gijk(cos_theta, p), gijk_d = ters_gijk_d(cos_theta, p);
  double dcosdri[3], dcosdrj[3], dcosdrk[3];
  costheta_d(rij_hat, rij, rik_hat, rik, dcosdri, dcosdrj, dcosdrk);
  for (int d = 0; d < 3; d++) {
    double v = -dfc * gijk * ex_delr * rik_hat[d];
    v = fc * gijk_d * ex_delr * dcosdri[d] + v;
    v = fc * gijk * ex_delr_d * rik_hat[d] + v;
    v = -fc * gijk * ex_delr_d * rij_hat[d] + v;
    dri[d] = prefactor * v;
    double w = fc * gijk_d * ex_delr * dcosdrj[d];
    w = fc * gijk * ex_delr_d * rij_hat[d] + w;
    drj[d] = prefactor * w;
    double u = dfc * gijk * ex_delr * rik_hat[d];
    u = fc * gijk_d * ex_delr * dcosdrk[d] + u;
    u = -fc * gijk * ex_delr_d * rik_hat[d] + u;
    drk[d] = prefactor * u;
  }
}
static void ters_attractive(const TParam *p, double prefactor, double rsqij, double rsqik,
                            const double *delrij, const double *delrik, double *fi, double *fj, double *fk) {
  double rij = sqrt(rsqij), rijinv = 1.0 / rij, rik = sqrt(rsqik), rikinv = 1.0 / rik;
  double rij_hat[3] = {rijinv * delrij[0], rijinv * delrij[1], rijinv * delrij[2]};
  double rik_hat[3] = {rikinv * delrik[0], rikinv * delrik[1], rikinv * delrik[2]};
  ters_zetaterm_d(prefactor, rij_hat, rij, rik_hat, rik, fi, fj, fk, p);
}

/* ----------------------------------------------------------------------- */
/* ILP functions                                                            */

static inline double calc_Tap(double r_ij, double Rcut) {
  static const double Tap_coeff[8] = {1.0, 0.0, 0.0, 0.0, -35.0, 84.0, -70.0, 20.0};
  double r = r_ij / Rcut, Tap;
  if (r >= 1.0) return 0.0;
  Tap = Tap_coeff[7] * r + Tap_coeff[6];
  Tap = Tap * r + Tap_coeff[5];
  Tap = Tap * r + Tap_coeff[4];
  Tap = Tap * r + Tap_coeff[3];
  Tap = Tap * r + Tap_coeff[2];
  Tap = Tap * r + Tap_coeff[1];
  Tap = Tap * r + Tap_coeff[0];
  return Tap;
}
static inline double calc_dTap(double r_ij, double Rcut) {
  static const double Tap_coeff[8] = {1.0, 0.0, 0.0, 0.0, -35.0, 84.0, -70.0, 20.0};
  double r = r_ij / Rcut, dTap;
  if (r >= 1.0) return 0.0;
  dTap = 7.0 * Tap_coeff[7] * r + 6.0 * Tap_coeff[6];
  dTap = dTap * r + 5.0 * Tap_coeff[5];
  dTap = dTap * r + 4.0 * Tap_coeff[4];
  dTap = dTap * r + 3.0 * Tap_coeff[3];
  dTap = dTap * r + 2.0 * Tap_coeff[2];
  dTap = dTap * r + Tap_coeff[1];
  dTap = dTap / Rcut;
  return dTap;
}

double orc_tap(double r, double rcut) { return calc_Tap(r, rcut); }
double orc_dtap(double r, double rcut) { return calc_dTap(r, rcut); }

/* per-atom normal data (calc_normal, ilp.cpp:750-1075) */
typedef struct {
  int cont;
  int nb[3];
  unsigned char ns[3];
  double normal[3];
  double dnormdri[3][3];     /* [id][ip] */
  double dnormal[3][3][3];   /* [id][ip][m] */
} Normal;

static int calc_normal_one(const double vet[3][3], int cont, Normal *nm, char *err, int errlen) {
  double pv12[3] = {0}, pv31[3] = {0}, pv23[3] = {0}, n1[3] = {0}, dni[3] = {0};
  double dnn[3][3] = {{0}}, dpvdri[3][3] = {{0}};
  double dn1[3][3][3], dpv12[3][3][3], dpv23[3][3][3], dpv31[3][3][3];
  memset(dn1, 0, sizeof dn1); memset(dpv12, 0, sizeof dpv12);
  memset(dpv23, 0, sizeof dpv23); memset(dpv31, 0, sizeof dpv31);
  memset(nm->normal, 0, sizeof nm->normal);
  memset(nm->dnormdri, 0, sizeof nm->dnormdri);
  memset(nm->dnormal, 0, sizeof nm->dnormal);
  double nn, nn2;
  if (cont <= 1) {
    nm->normal[0] = 0.0; nm->normal[1] = 0.0; nm->normal[2] = 1.0;
    return 0;
  } else if (cont == 2) {
    pv12[0] = vet[0][1] * vet[1][2] - vet[1][1] * vet[0][2];
    pv12[1] = vet[0][2] * vet[1][0] - vet[1][2] * vet[0][0];
    pv12[2] = vet[0][0] * vet[1][1] - vet[1][0] * vet[0][1];
    dpvdri[0][0] = 0.0;
    dpvdri[0][1] = vet[0][2] - vet[1][2];
    dpvdri[0][2] = vet[1][1] - vet[0][1];
    dpvdri[1][0] = vet[1][2] - vet[0][2];
    dpvdri[1][1] = 0.0;
    dpvdri[1][2] = vet[0][0] - vet[1][0];
    dpvdri[2][0] = vet[0][1] - vet[1][1];
    dpvdri[2][1] = vet[1][0] - vet[0][0];
    dpvdri[2][2] = 0.0;
    dpv12[0][0][0] = 0.0;         dpv12[0][1][0] = vet[1][2];  dpv12[0][2][0] = -vet[1][1];
    dpv12[1][0][0] = -vet[1][2];  dpv12[1][1][0] = 0.0;        dpv12[1][2][0] = vet[1][0];
    dpv12[2][0][0] = vet[1][1];   dpv12[2][1][0] = -vet[1][0]; dpv12[2][2][0] = 0.0;
    dpv12[0][0][1] = 0.0;         dpv12[0][1][1] = -vet[0][2]; dpv12[0][2][1] = vet[0][1];
    dpv12[1][0][1] = vet[0][2];   dpv12[1][1][1] = 0.0;        dpv12[1][2][1] = -vet[0][0];
    dpv12[2][0][1] = -vet[0][1];  dpv12[2][1][1] = vet[0][0];  dpv12[2][2][1] = 0.0;
    n1[0] = pv12[0]; n1[1] = pv12[1]; n1[2] = pv12[2];
    nn2 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    nn = sqrt(nn2);
    if (nn == 0) { snprintf(err, errlen, "The magnitude of the normal vector is zero"); return -1; }
    for (int d = 0; d < 3; d++) nm->normal[d] = n1[d] / nn;
    for (int ip = 0; ip < 3; ip++)
      dni[ip] = (n1[0] * dpvdri[0][ip] + n1[1] * dpvdri[1][ip] + n1[2] * dpvdri[2][ip]) / nn;
    for (int id = 0; id < 3; id++)
      for (int ip = 0; ip < 3; ip++)
        nm->dnormdri[id][ip] = dpvdri[id][ip] / nn - n1[id] * dni[ip] / nn2;
    for (int id = 0; id < 3; id++)
      for (int ip = 0; ip < 3; ip++)
        for (int m = 0; m < 3; m++) dn1[id][ip][m] = dpv12[id][ip][m];
  } else if (cont == 3) {
    pv12[0] = vet[0][1] * vet[1][2] - vet[1][1] * vet[0][2];
    pv12[1] = vet[0][2] * vet[1][0] - vet[1][2] * vet[0][0];
    pv12[2] = vet[0][0] * vet[1][1] - vet[1][0] * vet[0][1];
    dpv12[0][0][0] = 0.0;         dpv12[0][1][0] = vet[1][2];  dpv12[0][2][0] = -vet[1][1];
    dpv12[1][0][0] = -vet[1][2];  dpv12[1][1][0] = 0.0;        dpv12[1][2][0] = vet[1][0];
    dpv12[2][0][0] = vet[1][1];   dpv12[2][1][0] = -vet[1][0]; dpv12[2][2][0] = 0.0;
    dpv12[0][0][1] = 0.0;         dpv12[0][1][1] = -vet[0][2]; dpv12[0][2][1] = vet[0][1];
    dpv12[1][0][1] = vet[0][2];   dpv12[1][1][1] = 0.0;        dpv12[1][2][1] = -vet[0][0];
    dpv12[2][0][1] = -vet[0][1];  dpv12[2][1][1] = vet[0][0];  dpv12[2][2][1] = 0.0;

    pv31[0] = vet[2][1] * vet[0][2] - vet[0][1] * vet[2][2];
    pv31[1] = vet[2][2] * vet[0][0] - vet[0][2] * vet[2][0];
    pv31[2] = vet[2][0] * vet[0][1] - vet[0][0] * vet[2][1];
    dpv31[0][0][0] = 0.0;         dpv31[0][1][0] = -vet[2][2]; dpv31[0][2][0] = vet[2][1];
    dpv31[1][0][0] = vet[2][2];   dpv31[1][1][0] = 0.0;        dpv31[1][2][0] = -vet[2][0];
    dpv31[2][0][0] = -vet[2][1];  dpv31[2][1][0] = vet[2][0];  dpv31[2][2][0] = 0.0;
    dpv31[0][0][2] = 0.0;         dpv31[0][1][2] = vet[0][2];  dpv31[0][2][2] = -vet[0][1];
    dpv31[1][0][2] = -vet[0][2];  dpv31[1][1][2] = 0.0;        dpv31[1][2][2] = vet[0][0];
    dpv31[2][0][2] = vet[0][1];   dpv31[2][1][2] = -vet[0][0]; dpv31[2][2][2] = 0.0;

    pv23[0] = vet[1][1] * vet[2][2] - vet[2][1] * vet[1][2];
    pv23[1] = vet[1][2] * vet[2][0] - vet[2][2] * vet[1][0];
    pv23[2] = vet[1][0] * vet[2][1] - vet[2][0] * vet[1][1];
    dpv23[0][0][1] = 0.0;         dpv23[0][1][1] = vet[2][2];  dpv23[0][2][1] = -vet[2][1];
    dpv23[1][0][1] = -vet[2][2];  dpv23[1][1][1] = 0.0;        dpv23[1][2][1] = vet[2][0];
    dpv23[2][0][1] = vet[2][1];   dpv23[2][1][1] = -vet[2][0]; dpv23[2][2][1] = 0.0;
    dpv23[0][0][2] = 0.0;         dpv23[0][1][2] = -vet[1][2]; dpv23[0][2][2] = vet[1][1];
    dpv23[1][0][2] = vet[1][2];   dpv23[1][1][2] = 0.0;        dpv23[1][2][2] = -vet[1][0];
    dpv23[2][0][2] = -vet[1][1];  dpv23[2][1][2] = vet[1][0];  dpv23[2][2][2] = 0.0;

    for (int d = 0; d < 3; d++) n1[d] = (pv12[d] + pv31[d] + pv23[d]) / cont;
    nn2 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    nn = sqrt(nn2);
    if (nn == 0) { snprintf(err, errlen, "The magnitude of the normal vector is zero"); return -1; }
    for (int d = 0; d < 3; d++) nm->normal[d] = n1[d] / nn;
    for (int id = 0; id < 3; id++)
      for (int ip = 0; ip < 3; ip++)
        for (int m = 0; m < 3; m++)
          dn1[id][ip][m] = (dpv12[id][ip][m] + dpv23[id][ip][m] + dpv31[id][ip][m]) / cont;
  } else {
    snprintf(err, errlen, "There are too many neighbors for calculating normals");
    return -1;
  }
  for (int m = 0; m < 3; m++)
    for (int id = 0; id < 3; id++)
      dnn[id][m] = (n1[0] * dn1[0][id][m] + n1[1] * dn1[1][id][m] + n1[2] * dn1[2][id][m]) / nn;
  for (int m = 0; m < 3; m++)
    for (int id = 0; id < 3; id++)
      for (int ip = 0; ip < 3; ip++)
        nm->dnormal[id][ip][m] = dn1[id][ip][m] / nn - n1[id] * dnn[ip][m] / nn2;
  return 0;
}

/* ----------------------------------------------------------------------- */
/* Force evaluation                                                         */

typedef struct {
  double *f;        /* [n][3] */
  double *eatom;    /* [n] */
  double *vatom;    /* [n][6] */
  double fdotr_corr[6];
  double eng_ters, eng_vdw, eng_rep;
} Acc;

static inline int tag_skip(long long itag, long long jtag, const double *xi, const double *xj) {
  /* half selection of a full list by tag parity (pair_tersoff.cpp:150-158, ilp.cpp:488-496) */
  if (itag > jtag) {
    if ((itag + jtag) % 2 == 0) return 1;
  } else if (itag < jtag) {
    if ((itag + jtag) % 2 == 1) return 1;
  } else {
    if (xj[2] < xi[2]) return 1;
    if (xj[2] == xi[2] && xj[1] < xi[1]) return 1;
    if (xj[2] == xi[2] && xj[1] == xi[1] && xj[0] < xi[0]) return 1;
  }
  return 0;
}

static inline void add_f(Acc *a, const Orc *o, int j, int code, double fx, double fy, double fz, int fdotr) {
  a->f[3 * j] += fx; a->f[3 * j + 1] += fy; a->f[3 * j + 2] += fz;
  if (fdotr && code != 13) {
    int s[3];
    sdecode(code, s);
    double sh[3] = {s[0] * o->prd[0], s[1] * o->prd[1], s[2] * o->prd[2]};
    a->fdotr_corr[0] += fx * sh[0];
    a->fdotr_corr[1] += fy * sh[1];
    a->fdotr_corr[2] += fz * sh[2];
    a->fdotr_corr[3] += fy * sh[0];
    a->fdotr_corr[4] += fz * sh[0];
    a->fdotr_corr[5] += fz * sh[1];
  }
}

/* Pair::ev_tally with newton_pair on: per-atom halves */
static inline void ev_tally_atom(Acc *a, int eflag_atom, int vflag_atom, int i, int j, double evdwl,
                                 double fpair, double delx, double dely, double delz) {
  if (eflag_atom) { a->eatom[i] += 0.5 * evdwl; a->eatom[j] += 0.5 * evdwl; }
  if (vflag_atom) {
    double v[6] = {delx * delx * fpair, dely * dely * fpair, delz * delz * fpair,
                   delx * dely * fpair, delx * delz * fpair, dely * delz * fpair};
    for (int c = 0; c < 6; c++) { a->vatom[6 * i + c] += 0.5 * v[c]; a->vatom[6 * j + c] += 0.5 * v[c]; }
  }
}
/* Pair::ev_tally_xyz */
static inline void ev_tally_xyz_atom(Acc *a, int eflag_atom, int vflag_atom, int i, int j, double evdwl,
                                     double fx, double fy, double fz, double delx, double dely, double delz) {
  if (eflag_atom) { a->eatom[i] += 0.5 * evdwl; a->eatom[j] += 0.5 * evdwl; }
  if (vflag_atom) {
    double v[6] = {delx * fx, dely * fy, delz * fz, delx * fy, delx * fz, dely * fz};
    for (int c = 0; c < 6; c++) { a->vatom[6 * i + c] += 0.5 * v[c]; a->vatom[6 * j + c] += 0.5 * v[c]; }
  }
}
/* Pair::v_tally3 */
static inline void v_tally3_atom(Acc *a, int i, int j, int k, const double *fi, const double *fj,
                                 const double *drik, const double *drjk) {
  double v[6];
  v[0] = THIRD * (drik[0] * fi[0] + drjk[0] * fj[0]);
  v[1] = THIRD * (drik[1] * fi[1] + drjk[1] * fj[1]);
  v[2] = THIRD * (drik[2] * fi[2] + drjk[2] * fj[2]);
  v[3] = THIRD * (drik[0] * fi[1] + drjk[0] * fj[1]);
  v[4] = THIRD * (drik[0] * fi[2] + drjk[0] * fj[2]);
  v[5] = THIRD * (drik[1] * fi[2] + drjk[1] * fj[2]);
  for (int c = 0; c < 6; c++) {
    a->vatom[6 * i + c] += v[c]; a->vatom[6 * j + c] += v[c]; a->vatom[6 * k + c] += v[c];
  }
}

/* PairTersoff::compute for owned atoms [i0,i1) (pair_tersoff.cpp:86-266) */
static int tersoff_range(const Orc *o, Acc *a, int i0, int i1, int eflag, int eflag_atom, int vflag_atom,
                         int fdotr, char *err) {
  const double cutshortsq = o->cutmax * o->cutmax;
  int maxshort = 16;
  int *neighshort = (int *)malloc(sizeof(int) * maxshort);
  unsigned char *shortcode = (unsigned char *)malloc(maxshort);
  for (int i = i0; i < i1; i++) {
    long long itag = o->tag[i];
    int itype = o->t_map[o->type[i]];
    if (itype < 0) continue;
    const double *xi = &o->x[3 * i];
    double xtmp = xi[0], ytmp = xi[1], ztmp = xi[2];
    double fxtmp = 0, fytmp = 0, fztmp = 0;
    int numshort = 0;
    for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
      int j = o->ent_j[e], code = o->ent_s[e];
      int jt = o->t_map[o->type[j]];
      if (jt < 0) continue;  /* NULL-mapped types are skipped by hybrid's sub-style list */
      double xj[3];
      ghost_x(o, j, code, xj);
      double delx = xtmp - xj[0], dely = ytmp - xj[1], delz = ztmp - xj[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < cutshortsq) {
        neighshort[numshort] = j;
        shortcode[numshort] = (unsigned char)code;
        numshort++;
        if (numshort >= maxshort) {
          maxshort += maxshort / 2;
          neighshort = (int *)realloc(neighshort, sizeof(int) * maxshort);
          shortcode = (unsigned char *)realloc(shortcode, maxshort);
        }
      }
      if (tag_skip(itag, o->tag[j], xi, xj)) continue;
      int iparam_ij = o->t_e2p[itype][jt][jt];
      const TParam *p = &o->tp[iparam_ij];
      if (rsq >= p->cutsq) continue;
      double fpair, evdwl = 0.0;
      ters_repulsive(p, rsq, &fpair, eflag, &evdwl);
      fxtmp += delx * fpair; fytmp += dely * fpair; fztmp += delz * fpair;
      add_f(a, o, j, code, -delx * fpair, -dely * fpair, -delz * fpair, fdotr);
      if (eflag) a->eng_ters += evdwl;
      ev_tally_atom(a, eflag_atom, vflag_atom, i, j, evdwl, fpair, delx, dely, delz);
    }
    for (int jj = 0; jj < numshort; jj++) {
      int j = neighshort[jj], jcode = shortcode[jj];
      int jt = o->t_map[o->type[j]];
      int iparam_ij = o->t_e2p[itype][jt][jt];
      double xj[3];
      ghost_x(o, j, jcode, xj);
      double delr1[3] = {xj[0] - xtmp, xj[1] - ytmp, xj[2] - ztmp};
      double rsq1 = delr1[0] * delr1[0] + delr1[1] * delr1[1] + delr1[2] * delr1[2];
      if (rsq1 >= o->tp[iparam_ij].cutsq) continue;
      double fjx = 0, fjy = 0, fjz = 0, zeta_ij = 0.0;
      for (int kk = 0; kk < numshort; kk++) {
        if (jj == kk) continue;
        int k = neighshort[kk];
        int kt = o->t_map[o->type[k]];
        int iparam_ijk = o->t_e2p[itype][jt][kt];
        double xk[3];
        ghost_x(o, k, shortcode[kk], xk);
        double delr2[3] = {xk[0] - xtmp, xk[1] - ytmp, xk[2] - ztmp};
        double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 >= o->tp[iparam_ijk].cutsq) continue;
        zeta_ij += ters_zeta(&o->tp[iparam_ijk], rsq1, rsq2, delr1, delr2);
      }
      double fpair, prefactor, evdwl = 0.0;
      ters_force_zeta(&o->tp[iparam_ij], rsq1, zeta_ij, &fpair, &prefactor, eflag, &evdwl);
      fxtmp += delr1[0] * fpair; fytmp += delr1[1] * fpair; fztmp += delr1[2] * fpair;
      fjx -= delr1[0] * fpair; fjy -= delr1[1] * fpair; fjz -= delr1[2] * fpair;
      if (eflag) a->eng_ters += evdwl;
      ev_tally_atom(a, eflag_atom, vflag_atom, i, j, evdwl, -fpair, -delr1[0], -delr1[1], -delr1[2]);
      for (int kk = 0; kk < numshort; kk++) {
        if (jj == kk) continue;
        int k = neighshort[kk], kcode = shortcode[kk];
        int kt = o->t_map[o->type[k]];
        int iparam_ijk = o->t_e2p[itype][jt][kt];
        double xk[3];
        ghost_x(o, k, kcode, xk);
        double delr2[3] = {xk[0] - xtmp, xk[1] - ytmp, xk[2] - ztmp};
        double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 >= o->tp[iparam_ijk].cutsq) continue;
        double fi[3], fj[3], fk[3];
        ters_attractive(&o->tp[iparam_ijk], prefactor, rsq1, rsq2, delr1, delr2, fi, fj, fk);
        fxtmp += fi[0]; fytmp += fi[1]; fztmp += fi[2];
        fjx += fj[0]; fjy += fj[1]; fjz += fj[2];
        add_f(a, o, k, kcode, fk[0], fk[1], fk[2], fdotr);
        if (vflag_atom) v_tally3_atom(a, i, j, k, fj, fk, delr1, delr2);
      }
      add_f(a, o, j, jcode, fjx, fjy, fjz, fdotr);
    }
    add_f(a, o, i, 13, fxtmp, fytmp, fztmp, fdotr);
  }
  free(neighshort);
  free(shortcode);
  (void)err;
  return 0;
}

/* ILP_neigh + calc_normal for owned atoms (ilp.cpp:681-1075) */
static int ilp_normals(const Orc *o, Normal *nm, int i0, int i1, char *err, int errlen) {
  for (int i = i0; i < i1; i++) {
    int itype = o->i_map[o->type[i]];
    const double *xi = &o->x[3 * i];
    int n = 0;
    nm[i].cont = 0;
    if (itype >= 0) {
      for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
        int j = o->ent_j[e], code = o->ent_s[e];
        int jtype = o->i_map[o->type[j]];
        if (jtype < 0) continue;
        double xj[3];
        ghost_x(o, j, code, xj);
        double delx = xi[0] - xj[0], dely = xi[1] - xj[1], delz = xi[2] - xj[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq != 0 && rsq < o->cutILPsq[itype][jtype] && o->mol[i] == o->mol[j]) {
          if (n >= 3) {
            snprintf(err, errlen, "There are too many neighbors for some atoms, please check your configuration");
            return -1;
          }
          nm[i].nb[n] = j;
          nm[i].ns[n] = (unsigned char)code;
          n++;
        }
      }
    }
    nm[i].cont = n;
    double vet[3][3] = {{0}};
    for (int k = 0; k < n; k++) {
      double xk[3];
      ghost_x(o, nm[i].nb[k], nm[i].ns[k], xk);
      vet[k][0] = xk[0] - xi[0]; vet[k][1] = xk[1] - xi[1]; vet[k][2] = xk[2] - xi[2];
    }
    if (calc_normal_one(vet, n, &nm[i], err, errlen)) return -1;
  }
  return 0;
}

/* calc_FvdW (ilp.cpp:449-541) */
static void ilp_vdw_range(const Orc *o, Acc *a, int i0, int i1, int eflag, int eflag_atom, int vflag_atom, int fdotr) {
  for (int i = i0; i < i1; i++) {
    int it = o->type[i];
    if (o->i_map[it] < 0) continue;
    const double *xi = &o->x[3 * i];
    double xtmp = xi[0], ytmp = xi[1], ztmp = xi[2];
    for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
      int j = o->ent_j[e], code = o->ent_s[e];
      int jt = o->type[j];
      if (o->i_map[jt] < 0) continue;
      double xj[3];
      ghost_x(o, j, code, xj);
      if (tag_skip(o->tag[i], o->tag[j], xi, xj)) continue;
      double delx = xtmp - xj[0], dely = ytmp - xj[1], delz = ztmp - xj[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < o->ilp_cutsq[it][jt] && o->mol[i] != o->mol[j]) {
        const IParam *p = &o->ip[o->i_e2p[o->i_map[it]][o->i_map[jt]]];
        double r = sqrt(rsq), r2inv = 1.0 / rsq, r6inv = r2inv * r2inv * r2inv, r8inv = r6inv * r2inv;
        double Tap, dTap;
        if (o->tap_flag) {
          double Rcut = sqrt(o->ilp_cutsq[it][jt]);
          Tap = calc_Tap(r, Rcut);
          dTap = calc_dTap(r, Rcut);
        } else { Tap = 1.0; dTap = 0.0; }
        double TSvdw = 1.0 + exp(-p->d * (r / p->seff - 1.0));
        double TSvdw2inv = pow(TSvdw, -2.0);
        double Vilp = -p->C6 * r6inv / TSvdw;
        double fpair = -6.0 * p->C6 * r8inv / TSvdw + p->C6 * p->d / p->seff * (TSvdw - 1.0) * TSvdw2inv * r8inv * r;
        double fsum = fpair * Tap - Vilp * dTap / r;
        add_f(a, o, i, 13, fsum * delx, fsum * dely, fsum * delz, fdotr);
        add_f(a, o, j, code, -fsum * delx, -fsum * dely, -fsum * delz, fdotr);
        double evdwl = 0.0;
        if (eflag) { evdwl = Vilp * Tap; a->eng_vdw += evdwl; }
        ev_tally_atom(a, eflag_atom, vflag_atom, i, j, evdwl, fsum, delx, dely, delz);
      }
    }
  }
}

/* calc_FRep (ilp.cpp:547-675) */
static void ilp_rep_range(const Orc *o, const Normal *nm, Acc *a, int i0, int i1, int eflag, int eflag_atom,
                          int vflag_atom, int fdotr) {
  for (int i = i0; i < i1; i++) {
    int it = o->type[i];
    if (o->i_map[it] < 0) continue;
    const double *xi = &o->x[3 * i];
    double xtmp = xi[0], ytmp = xi[1], ztmp = xi[2];
    const Normal *ni = &nm[i];
    for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
      int j = o->ent_j[e], code = o->ent_s[e];
      int jt = o->type[j];
      if (o->i_map[jt] < 0) continue;
      double xj[3];
      ghost_x(o, j, code, xj);
      double delx = xtmp - xj[0], dely = ytmp - xj[1], delz = ztmp - xj[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < o->ilp_cutsq[it][jt] && o->mol[i] != o->mol[j]) {
        const IParam *p = &o->ip[o->i_e2p[o->i_map[it]][o->i_map[jt]]];
        double r = sqrt(rsq), Tap, dTap;
        if (o->tap_flag) {
          double Rcut = sqrt(o->ilp_cutsq[it][jt]);
          Tap = calc_Tap(r, Rcut);
          dTap = calc_dTap(r, Rcut);
        } else { Tap = 1.0; dTap = 0.0; }
        double prodnorm1 = ni->normal[0] * delx + ni->normal[1] * dely + ni->normal[2] * delz;
        double rhosq1 = rsq - prodnorm1 * prodnorm1;
        double rdsq1 = rhosq1 * p->delta2inv;
        double exp0 = exp(-p->lambda * (r - p->z0));
        double exp1 = exp(-rdsq1);
        double frho1 = exp1 * p->C;
        double Erep = 0.5 * p->epsilon + frho1;
        double Vilp = exp0 * Erep;
        double fpair = p->lambda * exp0 / r * Erep;
        double fpair1 = 2.0 * exp0 * frho1 * p->delta2inv;
        double fsum = fpair + fpair1;
        double dprodnorm1[3], fp1[3], fprod1[3];
        for (int c = 0; c < 3; c++)
          dprodnorm1[c] = ni->dnormdri[0][c] * delx + ni->dnormdri[1][c] * dely + ni->dnormdri[2][c] * delz;
        for (int c = 0; c < 3; c++) {
          fp1[c] = prodnorm1 * ni->normal[c] * fpair1;
          fprod1[c] = prodnorm1 * dprodnorm1[c] * fpair1;
        }
        double fkcx = (delx * fsum - fp1[0]) * Tap - Vilp * dTap * delx / r;
        double fkcy = (dely * fsum - fp1[1]) * Tap - Vilp * dTap * dely / r;
        double fkcz = (delz * fsum - fp1[2]) * Tap - Vilp * dTap * delz / r;
        add_f(a, o, i, 13, fkcx - fprod1[0] * Tap, fkcy - fprod1[1] * Tap, fkcz - fprod1[2] * Tap, fdotr);
        add_f(a, o, j, code, -fkcx, -fkcy, -fkcz, fdotr);
        for (int kk = 0; kk < ni->cont; kk++) {
          int k = ni->nb[kk], kcode = ni->ns[kk];
          if (k == i && kcode == 13) continue;
          double dp[3], fk[3];
          for (int c = 0; c < 3; c++)
            dp[c] = ni->dnormal[0][c][kk] * delx + ni->dnormal[1][c][kk] * dely + ni->dnormal[2][c][kk] * delz;
          for (int c = 0; c < 3; c++) fk[c] = (-prodnorm1 * dp[c] * fpair1) * Tap;
          add_f(a, o, k, kcode, fk[0], fk[1], fk[2], fdotr);
          double xk[3];
          ghost_x(o, k, kcode, xk);
          double delkj[3] = {xk[0] - xj[0], xk[1] - xj[1], xk[2] - xj[2]};
          ev_tally_xyz_atom(a, 0, vflag_atom, k, j, 0.0, fk[0], fk[1], fk[2], delkj[0], delkj[1], delkj[2]);
        }
        double evdwl = 0.0;
        if (eflag) { evdwl = Tap * Vilp; a->eng_rep += evdwl; }
        ev_tally_xyz_atom(a, eflag_atom, vflag_atom, i, j, evdwl, fkcx, fkcy, fkcz, delx, dely, delz);
      }
    }
  }
}

/*
 * Full force evaluation, Pair::compute semantics of
 * "pair_style hybrid/overlay tersoff ilp/graphene/hbn":
 *   f (+=), eng_vdwl (eng[0]), pvector (eng[1] = vdW, eng[2] = repulsive), eng[3] = Tersoff,
 *   virial[6] by F.r over owned + ghost atoms (pair_hybrid.cpp:209), eatom[n], vatom[n][6].
 * eflag: bit0 global, bit1 per atom.  vflag: vflag%4 != 0 global, vflag&4 per atom.
 * Returns 0 or -1 (message in orc_error).
 */
int orc_compute(Orc *o, int eflag, int vflag, double *f, double *eng, double *virial, double *eatom, double *vatom) {
  const int n = o->n;
  const int eflag_atom = (eflag / 2) & 1, eflag_g = eflag & 1 || eflag_atom;
  const int vflag_atom = (vflag & 4) ? 1 : 0, vflag_g = (vflag % 4) != 0;
  const int fdotr = vflag_g;
  Normal *nm = (Normal *)calloc(n ? n : 1, sizeof(Normal));
  int nt = o->nthreads;
  Acc *acc = (Acc *)calloc(nt, sizeof(Acc));
  int rc = 0;
  char errs[8][256];
  memset(errs, 0, sizeof errs);
  int fails = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : fails)
#endif
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    int i0 = (int)((long long)n * t / nt), i1 = (int)((long long)n * (t + 1) / nt);
    if (ilp_normals(o, nm, i0, i1, errs[t & 7], 256)) fails++;
  }
  if (fails) {
    for (int t = 0; t < 8; t++) if (errs[t][0]) { snprintf(o->err, sizeof o->err, "%s", errs[t]); break; }
    free(nm); free(acc);
    return -1;
  }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    Acc *a = &acc[t];
    a->f = (double *)calloc(3 * (size_t)(n ? n : 1), sizeof(double));
    if (eflag_atom) a->eatom = (double *)calloc(n ? n : 1, sizeof(double));
    if (vflag_atom) a->vatom = (double *)calloc(6 * (size_t)(n ? n : 1), sizeof(double));
    int i0 = (int)((long long)n * t / nt), i1 = (int)((long long)n * (t + 1) / nt);
    tersoff_range(o, a, i0, i1, eflag_g, eflag_atom, vflag_atom, fdotr, NULL);
    ilp_vdw_range(o, a, i0, i1, eflag_g, eflag_atom, vflag_atom, fdotr);
    ilp_rep_range(o, nm, a, i0, i1, eflag_g, eflag_atom, vflag_atom, fdotr);
  }
  /* deterministic reduction of the per-thread accumulators */
  double et = 0, ev = 0, er = 0, corr[6] = {0};
  for (int t = 0; t < nt; t++) {
    Acc *a = &acc[t];
    for (long long k = 0; k < 3LL * n; k++) f[k] += a->f[k];
    if (eflag_atom && eatom) for (int i = 0; i < n; i++) eatom[i] += a->eatom[i];
    if (vflag_atom && vatom) for (long long k = 0; k < 6LL * n; k++) vatom[k] += a->vatom[k];
    et += a->eng_ters; ev += a->eng_vdw; er += a->eng_rep;
    for (int c = 0; c < 6; c++) corr[c] += a->fdotr_corr[c];
  }
  if (eng) { eng[0] = et + ev + er; eng[1] = ev; eng[2] = er; eng[3] = et; }
  if (virial && fdotr) {
    /* virial_fdotr_compute over owned atoms with the pair forces of this call, plus images */
    double v[6] = {0};
    for (int t = 0; t < nt; t++) {
      double *ff = acc[t].f;
      for (int i = 0; i < n; i++) {
        const double *xi = &o->x[3 * i];
        v[0] += ff[3 * i] * xi[0];
        v[1] += ff[3 * i + 1] * xi[1];
        v[2] += ff[3 * i + 2] * xi[2];
        v[3] += ff[3 * i + 1] * xi[0];
        v[4] += ff[3 * i + 2] * xi[0];
        v[5] += ff[3 * i + 2] * xi[1];
      }
    }
    for (int c = 0; c < 6; c++) virial[c] = v[c] + corr[c];
  }
  for (int t = 0; t < nt; t++) { free(acc[t].f); free(acc[t].eatom); free(acc[t].vatom); }
  free(acc);
  free(nm);
  return rc;
}

/*
 * The reference's fused LMFF CPU path (LMFF/pair_lmff.cpp:1121-1613, computeLMFFTwice) restated,
 * for the CPU baseline: per owned atom one pass over its full list finds the Tersoff short
 * neighbours and the normal neighbours (1254-1295, calc_normal once), a second pass does the
 * Tersoff repulsive half pair and, for every other-molecule neighbour in the ILP cutoff, the
 * van der Waals half pair (by tag parity) and the repulsion with the atom's own normal, pushing
 * the pair force to both atoms (1320-1465); the normal-derivative forces are accumulated as
 * dprodni = sum_j -prodnorm1 fpair1 Tap del (1440-1442) and applied once per atom through
 * dnormal and dnormdri (1490-1519); then the Tersoff three-body terms (1523-1608).
 * Not reproduced (reference quirks): the recalc double count, the unbounded filtered short list,
 * the normal scaled by the full list length (calc_normal here is the stock one).
 * Returns forces and, with eflag, eng = {total, vdW, repulsion, Tersoff} (no per-atom tallies,
 * no virial: the baseline times the force path).  Same results as orc_compute up to rounding.
 */
static int lmff_fused_range(const Orc *o, Acc *a, int i0, int i1, int eflag, char *err, int errlen) {
  const double cutshortsq = o->cutmax * o->cutmax;
  int maxshort = 16;
  int *neighshort = (int *)malloc(sizeof(int) * maxshort);
  unsigned char *shortcode = (unsigned char *)malloc(maxshort);
  for (int i = i0; i < i1; i++) {
    const long long itag = o->tag[i];
    const int it = o->type[i];
    const int itype = o->t_map[it], itm = o->i_map[it];
    const double *xi = &o->x[3 * i];
    const double xtmp = xi[0], ytmp = xi[1], ztmp = xi[2];
    /* pass 1: short neighbours and normal neighbours */
    int numshort = 0, nilp = 0;
    Normal nm;
    double vet[3][3] = {{0}};
    for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
      const int j = o->ent_j[e], code = o->ent_s[e];
      double xj[3];
      ghost_x(o, j, code, xj);
      const double delx = xtmp - xj[0], dely = ytmp - xj[1], delz = ztmp - xj[2];
      const double rsq = delx * delx + dely * dely + delz * delz;
      const int jt = o->type[j];
      if (itype >= 0 && o->t_map[jt] >= 0 && rsq < cutshortsq) {
        if (numshort >= maxshort - 1) {
          maxshort += maxshort / 2;
          neighshort = (int *)realloc(neighshort, sizeof(int) * maxshort);
          shortcode = (unsigned char *)realloc(shortcode, maxshort);
        }
        neighshort[numshort] = j;
        shortcode[numshort] = (unsigned char)code;
        numshort++;
      }
      if (itm >= 0 && o->i_map[jt] >= 0 && rsq != 0 && rsq < o->cutILPsq[itm][o->i_map[jt]] && o->mol[i] == o->mol[j]) {
        if (nilp >= 3) {
          snprintf(err, errlen, "There are too many neighbors for some atoms, please check your configuration");
          free(neighshort); free(shortcode);
          return -1;
        }
        nm.nb[nilp] = j;
        nm.ns[nilp] = (unsigned char)code;
        vet[nilp][0] = -delx; vet[nilp][1] = -dely; vet[nilp][2] = -delz;
        nilp++;
      }
    }
    nm.cont = nilp;
    if (itm >= 0 && calc_normal_one(vet, nilp, &nm, err, errlen)) { free(neighshort); free(shortcode); return -1; }
    /* pass 2: Tersoff repulsive half pairs and ILP, pushed to both atoms */
    double fxtmp = 0, fytmp = 0, fztmp = 0;
    double dprodni[3] = {0, 0, 0};
    for (int e = o->offs[i]; e < o->offs[i + 1]; e++) {
      const int j = o->ent_j[e], code = o->ent_s[e];
      const int jt = o->type[j];
      double xj[3];
      ghost_x(o, j, code, xj);
      const double delx = xtmp - xj[0], dely = ytmp - xj[1], delz = ztmp - xj[2];
      const double rsq = delx * delx + dely * dely + delz * delz;
      const int owner = !tag_skip(itag, o->tag[j], xi, xj);
      const double r = sqrt(rsq), rinv = 1.0 / r;
      if (owner && itype >= 0 && o->t_map[jt] >= 0) {
        const TParam *p = &o->tp[o->t_e2p[itype][o->t_map[jt]][o->t_map[jt]]];
        if (rsq < p->cutsq) {
          double fpair, evdwl = 0.0;
          ters_repulsive(p, rsq, &fpair, eflag, &evdwl);
          fxtmp += delx * fpair; fytmp += dely * fpair; fztmp += delz * fpair;
          add_f(a, o, j, code, -delx * fpair, -dely * fpair, -delz * fpair, 0);
          if (eflag) a->eng_ters += evdwl;
        }
      }
      if (itm >= 0 && o->i_map[jt] >= 0 && rsq < o->ilp_cutsq[it][jt] && o->mol[i] != o->mol[j]) {
        const IParam *p = &o->ip[o->i_e2p[itm][o->i_map[jt]]];
        double Tap, dTap;
        if (o->tap_flag) {
          const double Rcut = sqrt(o->ilp_cutsq[it][jt]);
          Tap = calc_Tap(r, Rcut);
          dTap = calc_dTap(r, Rcut);
        } else { Tap = 1.0; dTap = 0.0; }
        double fvdw = 0.0;
        if (owner) {
          const double r2inv = rinv * rinv, r6inv = r2inv * r2inv * r2inv, r8inv = r6inv * r2inv;
          const double TSvdw = 1.0 + exp(-p->d * (r / p->seff - 1.0));
          const double TSvdw2inv = 1.0 / (TSvdw * TSvdw);
          const double vvdw = -p->C6 * r6inv / TSvdw;
          const double fp = -6.0 * p->C6 * r8inv / TSvdw + p->C6 * p->d / p->seff * (TSvdw - 1.0) * TSvdw2inv * r8inv * r;
          fvdw = fp * Tap - vvdw * dTap * rinv;
          if (eflag) a->eng_vdw += vvdw * Tap;
        }
        const double prodnorm1 = nm.normal[0] * delx + nm.normal[1] * dely + nm.normal[2] * delz;
        const double rhosq1 = rsq - prodnorm1 * prodnorm1;
        const double rdsq1 = rhosq1 * p->delta2inv;
        const double exp0 = exp(-p->lambda * (r - p->z0));
        const double exp1 = exp(-rdsq1);
        const double frho1 = exp1 * p->C;
        const double Erep = 0.5 * p->epsilon + frho1;
        const double Vilp = exp0 * Erep;
        const double fpair = p->lambda * exp0 * rinv * Erep;
        const double fpair1 = 2.0 * exp0 * frho1 * p->delta2inv;
        const double fsum = fpair + fpair1;
        const double fkcx = (delx * fsum - prodnorm1 * nm.normal[0] * fpair1) * Tap - Vilp * dTap * delx * rinv;
        const double fkcy = (dely * fsum - prodnorm1 * nm.normal[1] * fpair1) * Tap - Vilp * dTap * dely * rinv;
        const double fkcz = (delz * fsum - prodnorm1 * nm.normal[2] * fpair1) * Tap - Vilp * dTap * delz * rinv;
        dprodni[0] += -prodnorm1 * fpair1 * Tap * delx;
        dprodni[1] += -prodnorm1 * fpair1 * Tap * dely;
        dprodni[2] += -prodnorm1 * fpair1 * Tap * delz;
        const double ftx = fvdw * delx + fkcx, fty = fvdw * dely + fkcy, ftz = fvdw * delz + fkcz;
        fxtmp += ftx; fytmp += fty; fztmp += ftz;
        add_f(a, o, j, code, -ftx, -fty, -ftz, 0);
        if (eflag) a->eng_rep += Tap * Vilp;
      }
    }
    /* normal-derivative forces: on the normal neighbours and (two-neighbour normals) on i */
    if (itm >= 0) {
      for (int kk = 0; kk < nm.cont; kk++) {
        double fk[3];
        for (int c = 0; c < 3; c++)
          fk[c] = nm.dnormal[0][c][kk] * dprodni[0] + nm.dnormal[1][c][kk] * dprodni[1] + nm.dnormal[2][c][kk] * dprodni[2];
        add_f(a, o, nm.nb[kk], nm.ns[kk], fk[0], fk[1], fk[2], 0);
      }
      fxtmp += nm.dnormdri[0][0] * dprodni[0] + nm.dnormdri[1][0] * dprodni[1] + nm.dnormdri[2][0] * dprodni[2];
      fytmp += nm.dnormdri[0][1] * dprodni[0] + nm.dnormdri[1][1] * dprodni[1] + nm.dnormdri[2][1] * dprodni[2];
      fztmp += nm.dnormdri[0][2] * dprodni[0] + nm.dnormdri[1][2] * dprodni[1] + nm.dnormdri[2][2] * dprodni[2];
    }
    /* Tersoff three-body terms of i's short neighbours */
    for (int jj = 0; jj < numshort; jj++) {
      const int j = neighshort[jj], jcode = shortcode[jj];
      const int jt = o->t_map[o->type[j]];
      const int iparam_ij = o->t_e2p[itype][jt][jt];
      double xj[3];
      ghost_x(o, j, jcode, xj);
      const double delr1[3] = {xj[0] - xtmp, xj[1] - ytmp, xj[2] - ztmp};
      const double rsq1 = delr1[0] * delr1[0] + delr1[1] * delr1[1] + delr1[2] * delr1[2];
      if (rsq1 >= o->tp[iparam_ij].cutsq) continue;
      double fjx = 0, fjy = 0, fjz = 0, zeta_ij = 0.0;
      for (int kk = 0; kk < numshort; kk++) {
        if (jj == kk) continue;
        const int k = neighshort[kk];
        const int iparam_ijk = o->t_e2p[itype][jt][o->t_map[o->type[k]]];
        double xk[3];
        ghost_x(o, k, shortcode[kk], xk);
        const double delr2[3] = {xk[0] - xtmp, xk[1] - ytmp, xk[2] - ztmp};
        const double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 >= o->tp[iparam_ijk].cutsq) continue;
        zeta_ij += ters_zeta(&o->tp[iparam_ijk], rsq1, rsq2, delr1, delr2);
      }
      double fpair, prefactor, evdwl = 0.0;
      ters_force_zeta(&o->tp[iparam_ij], rsq1, zeta_ij, &fpair, &prefactor, eflag, &evdwl);
      fxtmp += delr1[0] * fpair; fytmp += delr1[1] * fpair; fztmp += delr1[2] * fpair;
      fjx -= delr1[0] * fpair; fjy -= delr1[1] * fpair; fjz -= delr1[2] * fpair;
      if (eflag) a->eng_ters += evdwl;
      for (int kk = 0; kk < numshort; kk++) {
        if (jj == kk) continue;
        const int k = neighshort[kk], kcode = shortcode[kk];
        const int iparam_ijk = o->t_e2p[itype][jt][o->t_map[o->type[k]]];
        double xk[3];
        ghost_x(o, k, kcode, xk);
        const double delr2[3] = {xk[0] - xtmp, xk[1] - ytmp, xk[2] - ztmp};
        const double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 >= o->tp[iparam_ijk].cutsq) continue;
        double fi[3], fj[3], fk[3];
        ters_attractive(&o->tp[iparam_ijk], prefactor, rsq1, rsq2, delr1, delr2, fi, fj, fk);
        fxtmp += fi[0]; fytmp += fi[1]; fztmp += fi[2];
        fjx += fj[0]; fjy += fj[1]; fjz += fj[2];
        add_f(a, o, k, kcode, fk[0], fk[1], fk[2], 0);
      }
      add_f(a, o, j, jcode, fjx, fjy, fjz, 0);
    }
    add_f(a, o, i, 13, fxtmp, fytmp, fztmp, 0);
  }
  free(neighshort);
  free(shortcode);
  return 0;
}

int orc_compute_lmff(Orc *o, int eflag, double *f, double *eng) {
  const int n = o->n;
  const int nt = o->nthreads;
  Acc *acc = (Acc *)calloc(nt, sizeof(Acc));
  char errs[8][256];
  memset(errs, 0, sizeof errs);
  int fails = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : fails)
#endif
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    Acc *a = &acc[t];
    a->f = (double *)calloc(3 * (size_t)(n ? n : 1), sizeof(double));
    int i0 = (int)((long long)n * t / nt), i1 = (int)((long long)n * (t + 1) / nt);
    if (lmff_fused_range(o, a, i0, i1, eflag & 1, errs[t & 7], 256)) fails++;
  }
  int rc = 0;
  if (fails) {
    for (int t = 0; t < 8; t++) if (errs[t][0]) { snprintf(o->err, sizeof o->err, "%s", errs[t]); break; }
    rc = -1;
  } else {
    double et = 0, ev = 0, er = 0;
    for (int t = 0; t < nt; t++) {
      for (long long k = 0; k < 3LL * n; k++) f[k] += acc[t].f[k];
      et += acc[t].eng_ters; ev += acc[t].eng_vdw; er += acc[t].eng_rep;
    }
    if (eng) { eng[0] = et + ev + er; eng[1] = ev; eng[2] = er; eng[3] = et; }
  }
  for (int t = 0; t < nt; t++) free(acc[t].f);
  free(acc);
  return rc;
}

/* Normals of owned atoms: normal[n][3], count[n], dnormal[n][3][3][3] ([id][ip][m]), dnormdri[n][3][3] */
int orc_normals(Orc *o, double *normal, int *count, double *dnormal, double *dnormdri) {
  Normal *nm = (Normal *)calloc(o->n ? o->n : 1, sizeof(Normal));
  if (ilp_normals(o, nm, 0, o->n, o->err, sizeof o->err)) { free(nm); return -1; }
  for (int i = 0; i < o->n; i++) {
    memcpy(&normal[3 * i], nm[i].normal, sizeof(double) * 3);
    count[i] = nm[i].cont;
    if (dnormal) memcpy(&dnormal[27 * i], nm[i].dnormal, sizeof(double) * 27);
    if (dnormdri) memcpy(&dnormdri[9 * i], nm[i].dnormdri, sizeof(double) * 9);
  }
  free(nm);
  return 0;
}

/*
 * Velocity-Verlet NVE (FixNVE, LMFF/fix_nve.cpp:64-143) driven as Verlet::run does
 * (LMFF/verlet.cpp:220-350): initial_integrate, rebuild decision (fixed interval
 * "neigh_modify every K delay 0 check no"), pbc + list build on rebuild steps, force,
 * final_integrate.  x, v are [n][3] in/out, mass_per_type[ntypes] (index type-1).
 * thermo[k*3 + {0,1,2}] = (step, pe, ke) every thermo_every steps (including step 0).
 */
int orc_run_nve(Orc *o, double *x, double *v, const double *mass_per_type, double dt, int nsteps,
                int rebuild_every, int thermo_every, double *thermo) {
  const int n = o->n;
  const double ftm2v = 1.0 / 1.0364269e-4, mvv2e = 1.0364269e-4;
  const double dtv = dt, dtf = 0.5 * dt * ftm2v;
  memcpy(o->x, x, sizeof(double) * 3 * n);
  double *f = (double *)calloc(3 * (size_t)(n ? n : 1), sizeof(double));
  double eng[4];
  int nth = 0;
  orc_pbc(o);
  if (orc_build_list(o)) { free(f); return -1; }
  if (orc_compute(o, 1, 0, f, eng, NULL, NULL, NULL)) { free(f); return -1; }
  int ago = 0;
  for (int step = 0; step <= nsteps; step++) {
    if (step > 0) {
      for (int i = 0; i < n; i++) {
        double dtfm = dtf / mass_per_type[o->type[i] - 1];
        for (int d = 0; d < 3; d++) {
          v[3 * i + d] += dtfm * f[3 * i + d];
          o->x[3 * i + d] += dtv * v[3 * i + d];
        }
      }
      ago++;
      if (rebuild_every > 0 && ago >= rebuild_every) {
        orc_pbc(o);
        if (orc_build_list(o)) { free(f); return -1; }
        ago = 0;
      }
      memset(f, 0, sizeof(double) * 3 * n);
      int want_e = thermo_every > 0 && (step % thermo_every == 0 || step == nsteps);
      if (orc_compute(o, want_e ? 1 : 0, 0, f, eng, NULL, NULL, NULL)) { free(f); return -1; }
      for (int i = 0; i < n; i++) {
        double dtfm = dtf / mass_per_type[o->type[i] - 1];
        for (int d = 0; d < 3; d++) v[3 * i + d] += dtfm * f[3 * i + d];
      }
    }
    if (thermo && thermo_every > 0 && (step % thermo_every == 0 || step == nsteps)) {
      double ke = 0.0;
      for (int i = 0; i < n; i++) {
        double m = mass_per_type[o->type[i] - 1];
        ke += m * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
      }
      ke *= 0.5 * mvv2e;
      thermo[3 * nth] = step; thermo[3 * nth + 1] = eng[0]; thermo[3 * nth + 2] = ke;
      nth++;
    }
  }
  memcpy(x, o->x, sizeof(double) * 3 * n);
  free(f);
  return nth;
}
