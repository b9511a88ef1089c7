// Host engine and C ABI of the LMFF force path on MI355X (see include/lmff_hip.h).
//
// One context = one HIP stream + every resident system packed into one cell-sorted atom
// array.  Per MD step (Verlet::run, LMFF/verlet.cpp:220-350):
//   initial_integrate -> [rebuild: pbc wrap, cell sort, permute, list build]
//   -> normals -> ILP pair -> Tersoff zeta -> gather (normal-derivative + Tersoff forces)
//   -> post_force/final_integrate -> [thermo reductions]
// The steps between rebuilds can be replayed from a captured hipGraph.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <functional>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "lmff_hip.h"
#include "lmff_kernels.hpp"
#include "potentials.hpp"

using namespace lmff;

// Scratch of the allocation-free entry point (lmff_pair_compute_noalloc): every per-atom array
// of one force evaluation lives here instead of in library allocations.  Each GPU the library
// runs on holds its own copy until the process ends; calls on one GPU share it (and race if they
// overlap).  Host code never passes the array's name as an address: the device address of each
// GPU's copy comes from hipGetSymbolAddress at context creation (scratch_address).
constexpr size_t LMFF_SCRATCH_BYTES = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[LMFF_SCRATCH_BYTES];

namespace {

constexpr int CHUNK = 4096;
constexpr int BS = 256;
constexpr int HALF_TILE_X = 6, HALF_TILE_Y = 3;  // bins per half-pair ILP tile

// Systems below this many atoms keep the pull ILP kernel (no tiles): a half-pair workgroup makes
// one pass per partner z bin, each with its loads, barriers and flush in a row, so a small system
// waits on that chain (about 90 us per step for a 3,000- to 76,000-atom bilayer against 11-100 us
// in the pull kernel), and a packed stream of such systems also ran slower with it.
// LMFF_HALF_MIN_ATOMS in the environment sets another bound (0: every system that fits).
int64_t half_min_atoms() {
  static const int64_t v = [] {
    const char *e = std::getenv("LMFF_HALF_MIN_ATOMS");
    return e ? std::max<int64_t>(0, std::atoll(e)) : (int64_t)100000;
  }();
  return v;
}
// window atoms the half-pair kernel keeps in LDS at most: WCAP_MAX, or LMFF_HALF_WCAP (a
// multiple of 64 from 64 up, read at every list build) to send more slots to the global sums
int half_wcap_limit() {
  const char *e = std::getenv("LMFF_HALF_WCAP");
  if (!e) return WCAP_MAX;
  return std::min(WCAP_MAX, std::max(64, std::atoi(e) / 64 * 64));
}
constexpr int MAX_DEVICES = 64;

// device address of `scratch` per GPU, looked up once per device with that device current
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

struct HipError {
  hipError_t e;
  const char *what;
};

#define HIPCHK(call)                                      \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) throw HipError{e_, #call};      \
  } while (0)

struct LmffError {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw LmffError{code, buf};
}

// Device buffer with capacity growth (contents are not preserved on growth).
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  size_t *tally = nullptr;
  void ensure(size_t need) {
    if (need <= bytes) return;
    release();
    size_t b = std::max<size_t>(need, 256);
    HIPCHK(hipMalloc(&p, b));
    bytes = b;
    if (tally) *tally += b;
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      if (tally) *tally -= bytes;
    }
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return (T *)p; }
  ~DBuf() { release(); }
};

// page-locked host memory (hipHostMalloc), grown on demand, for device-to-host copies at
// full PCIe rate
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void ensure(size_t need) {
    if (need <= bytes) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    HIPCHK(hipHostMalloc(&p, need, hipHostMallocDefault));
    bytes = need;
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

struct SysHost {
  int id;
  bool live;
  bool staged = false;
  int64_t n;
  double lo[3], hi[3];
  int per[3];
  int nc[3];
  double glo[3], bs[3];
  int cell_off, ncells;
  int row_off = 0;  // first of this system's rows of cells (fixed y and z bin)
  int tile[2] = {0, 0}, ntile[2] = {0, 0};  // half-pair ILP tiles (bins per tile, tiles)
  int tile_off = 0;                          // first tile x z-bin workgroup / window of the system
  int begin, end;
  int style = LMFF_NVE;
  double dt = 0.001, t_target = 300.0, damp = 1.0;
  unsigned long long seed = 1;
  long long step = 0;
  std::vector<int> chunk_ids;
  // sheets of a stack (open z): 0 not looked at yet, 1 its molecules are the sheets (sheet_ids,
  // sorted), -1 not a stack of at most MAX_SHEETS molecules
  int sheet_state = 0;
  std::vector<int> sheet_ids;
};

void check_box(const double lo[3], const double hi[3], const int per[3], double cutneigh);
void setup_grid(SysHost &s, const double *x, int64_t n, double cutneigh);
void grid_from_extent(SysHost &s, const double mn[3], const double mx[3], int64_t n, double cutneigh);

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t bytes_in_use = 0;
  // settings
  int ntypes = 0;
  double ilp_cut = 16.0;
  int tap_flag = 1;
  bool have_settings = false, have_tersoff = false, have_ilp = false;
  TersoffTable ters;
  IlpTable ilp;
  double skin = 2.0;
  int every = 1, delay = 0, check = 1;
  int precision = LMFF_FP64;
  std::vector<double> mass;  // per type
  bool params_dirty = true;
  // systems
  std::vector<SysHost> sys;  // in packed order
  int next_id = 0;
  int64_t n = 0;             // total atoms
  int ncells = 0;
  int nrows_cells = 0;  // rows of cells (fixed y and z bin) over all systems
  bool layout_dirty = true;  // packed arrays must be rebuilt from host staging
  bool need_rebuild = true;
  bool forces_valid = false;
  int ago = 0;
  long long global_step = 0;
  // Neighbor::ncalls and ndanger (LMFF/neighbor.cpp:2026,2041): list builds, and builds the
  // distance check asked for on the first step it was allowed to (an atom may already have
  // moved past half the skin before), since the last reset
  long long ncalls = 0, ndanger = 0;
  // the last step of every run also tallies per-atom energy and virial (for dumps of
  // compute pe/atom and stress/atom, which LAMMPS requests on the steps they are output)
  bool peratom_last = false;
  // staging for systems added since the last layout (host copies, caller order)
  struct Staged {
    int id;
    std::vector<double> x, v;
    std::vector<int> type, tag, mol, img;
  };
  std::vector<Staged> staged;
  // device state
  DBuf x, v, f, xhold, mol, tag, img, sysid, orig, key, key2, val, perm, tmp4, tmpi;
  DBuf slist, scnt, ilist, icnt, ocnt;
  DBuf nrm, dn, dndri, vet, nnb, tsh, tcnt, rtc, rtcnt, rnc, rncnt, eb, fb, wb, ft, fil, S, T, th, vatom;
  DBuf cell_start, row_mol, sort_tmp, systab, chunks, cbeg, partial, sums, rmass;
  DBuf ip_d, ip_f, tp_d, tp_f, t3, errw, ovf, dflag, fsave, farflag;
  DBuf hb_in, hb_out;  // host-buffer Pair::compute: caller-order arrays on the device
  // half-pair ILP: window slots of list entries, per-dz segments, window tables, per-system
  // largest window and mode, partner sums
  DBuf lidx, segf, segb, win, winmax, syshalf, acc, ipart;
  DBuf extent_d;  // per re-fitted system: the atoms' extent (k_extent)
  static constexpr int REGRID_EVERY = 5;
  bool regrid_due = true;
  int builds_since_regrid = 0;
  PinnedBuf regrid_host;  // the extents read back, the system table's grid fields written
  DBuf sheet_d;  // per sheet of the re-fitted stacks: z sum and count (k_sheet_z), then the ids
  PinnedBuf sheet_host;
  // full ILP rows for evaluations by the pull kernel while half systems hold half rows
  DBuf ilist_full, icnt_full, ocnt_full;
  int icap_full = 0;
  int npass = 0;               // z-bin offsets the half-pair kernel walks (0 .. npass - 1)
  int ntiles = 0;              // tile x z-bin workgroups (= windows) over every system
  std::vector<int> half_sys;   // per system: 1 when its ILP pairs run in the half-pair kernel
  bool half_any = false, pull_any = true;
  int wcap = 0;                // window atoms the half-pair kernel's LDS holds (largest half window)
  bool half_ovf = false;       // some half window is larger than wcap (the kernel's OVF variant)
  bool half_enabled = true;    // LMFF_HALF=0 in the environment keeps every system in the pull kernel
  PinnedBuf hb_host;   // and the forces on their way back
  int ntp = 0;  // Tersoff parameter records
  bool ilp_sym = false;  // every mapped ILP type pair symmetric (k_ilp's SYM variant applies)
  // every Tersoff-mapped type pair has the same repulsive parameters (A, lambda1, R, D) in the
  // (i, j, j) and (j, i, i) entries, so the tag-parity choice between them is moot
  bool ters_rep_sym = false;
  // conjugate-gradient vectors (x0, g, h), carried through rebuilds while a CG minimisation runs
  DBuf cg_x0, cg_g, cg_h;
  bool cg_active = false;
  // atom ranges whose last-step force is restored after the next set-up force evaluation
  std::vector<std::pair<int, int>> restore_f;
  int scap = 64, icap = 0;
  size_t sort_tmp_bytes = 0;
  int end_bit = 1;
  int nchunks = 0;
  // derived cutoffs
  double cutneigh = 18.0, short_cut = 4.1, short_reach = 2.1;
  // graph
  hipGraphExec_t graph = nullptr;
  int graph_len = 0;
  long long graph_gen = -1;
  long long gen = 0;
  int graph_prec = -1;
  // last evaluation flags (for per-atom outputs)
  int last_eflag = 0, last_vflag = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, std::pair<size_t, size_t>>> ev_marks;
  double prof_total = 0.0;
  double prof_ms[8] = {0};
  long long prof_cnt[8] = {0};
  // pair-plugin bookkeeping
  int pair_sys = -1;
  int64_t pair_n = -1;
  // allocation-free entry point: this device's copy of `scratch` and a small page-locked word
  // block for the status it reads back (both set up by lmff_create)
  unsigned char *scratch_dev = nullptr;
  int *note_host = nullptr;

  Ctx() {
    DBuf *all[] = {&x, &v, &f, &xhold, &mol, &tag, &img, &sysid, &orig, &key, &key2, &val, &perm, &tmp4, &tmpi,
                   &slist, &scnt, &ilist, &icnt, &ocnt, &nrm, &dn, &dndri, &vet, &nnb, &tsh, &tcnt, &rtc, &rtcnt,
                   &rnc, &rncnt, &eb, &fb, &wb, &ft,
                   &fil, &S, &T, &th, &vatom, &cell_start, &row_mol, &sort_tmp, &systab, &chunks, &cbeg, &partial, &sums,
                   &rmass, &ip_d, &ip_f, &tp_d, &tp_f, &t3, &errw, &ovf, &dflag, &fsave, &farflag,
                   &cg_x0, &cg_g, &cg_h, &hb_in, &hb_out, &lidx, &segf, &segb, &win, &winmax, &syshalf, &acc, &ipart, &ilist_full, &icnt_full, &ocnt_full, &extent_d, &sheet_d};
    for (DBuf *b : all) b->tally = &bytes_in_use;
    const char *h = std::getenv("LMFF_HALF");
    if (h && std::strcmp(h, "0") == 0) half_enabled = false;
  }
  ~Ctx() {
    if (note_host) (void)hipHostFree(note_host);
    if (graph) (void)hipGraphExecDestroy(graph);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  SysHost *find(int id) {
    for (auto &s : sys)
      if (s.id == id && s.live) return &s;
    return nullptr;
  }

  // ---------------------------------------------------------------- parameters
  template <typename R>
  void fill_params(std::vector<IlpPair<R>> &ipv, std::vector<TersP<R>> &tpv, std::vector<int> &t3v) {
    ipv.assign(MAXT * MAXT, IlpPair<R>{});
    for (int a = 1; a <= ntypes; a++)
      for (int b = 1; b <= ntypes; b++) {
        IlpPair<R> &q = ipv[(a - 1) * MAXT + (b - 1)];
        const int ea = ilp.emap.map[a], eb_ = ilp.emap.map[b];
        if (ea < 0 || eb_ < 0) continue;
        const int ne = ilp.nelements();
        const IlpEntry &p = ilp.params[ilp.elem2param[ea * ne + eb_]];
        const IlpEntry &r = ilp.params[ilp.elem2param[eb_ * ne + ea]];
        q.z0 = (R)p.z0;
        q.lambda = (R)p.lambda;
        q.half_eps = (R)(0.5 * p.epsilon);
        q.C = (R)p.C;
        q.d2inv = (R)p.delta2inv;
        q.C6 = (R)p.C6;
        q.d = (R)p.d;
        q.inv_seff = (R)(1.0 / p.seff);
        q.C6_d_seff = (R)(p.C6 * p.d / p.seff);
        q.rcutsq = ilp.cutILPsq[ea * ne + eb_];
        q.mapped = 1;
        q.vsym = (p.C6 == r.C6 && p.d == r.d && p.seff == r.seff) ? 1 : 0;
        q.esym = (p.lambda == r.lambda && p.z0 == r.z0) ? 1 : 0;
        if (!q.vsym || !q.esym) ilp_sym = false;
      }
    tpv.clear();
    for (const auto &p : ters.params) {
      TersP<R> q{};
      q.lam1 = (R)p.lam1; q.biga = (R)p.biga; q.lam2 = (R)p.lam2; q.bigb = (R)p.bigb;
      q.bigr = (R)p.bigr; q.bigd = (R)p.bigd; q.cc = (R)(p.c * p.c); q.dd = (R)(p.d * p.d);
      q.h = (R)p.h; q.gamma = (R)p.gamma; q.powern = (R)p.powern; q.beta = (R)p.beta;
      q.c1 = (R)p.c1; q.c2 = (R)p.c2; q.c3 = (R)p.c3; q.c4 = (R)p.c4; q.lam3 = (R)p.lam3;
      q.cutsq = p.cutsq;
      q.powermint = p.powermint;
      tpv.push_back(q);
    }
    if (tpv.empty()) tpv.push_back(TersP<R>{});
    t3v.assign(MAXT * MAXT * MAXT, -1);
    for (int a = 1; a <= ntypes; a++)
      for (int b = 1; b <= ntypes; b++)
        for (int c = 1; c <= ntypes; c++) {
          const int ea = ters.emap.map[a], eb_ = ters.emap.map[b], ec = ters.emap.map[c];
          if (ea < 0 || eb_ < 0 || ec < 0) continue;
          t3v[((a - 1) * MAXT + (b - 1)) * MAXT + (c - 1)] = ters.index(ea, eb_, ec);
        }
    if (sizeof(R) != 8) return;  // decided on the FP64 values (float rounding could merge them)
    ters_rep_sym = true;
    for (int a = 0; a < ntypes; a++)
      for (int b = 0; b < ntypes; b++) {
        const int pab = t3v[(a * MAXT + b) * MAXT + b], pba = t3v[(b * MAXT + a) * MAXT + a];
        if (pab < 0 || pba < 0) continue;
        const TersP<R> &u = tpv[pab], &w = tpv[pba];
        if (u.biga != w.biga || u.lam1 != w.lam1 || u.bigr != w.bigr || u.bigd != w.bigd || u.cutsq != w.cutsq)
          ters_rep_sym = false;
      }
  }

  void upload_params() {
    if (!params_dirty) return;
    if (!have_settings || !have_tersoff || !have_ilp) fail(LMFF_ERR_ARG, "pair_style/pair_coeff not complete");
    if ((int)mass.size() != ntypes) fail(LMFF_ERR_ARG, "masses not set for all atom types");
    std::vector<IlpPair<double>> ipd;
    std::vector<IlpPair<float>> ipf;
    std::vector<TersP<double>> tpd;
    std::vector<TersP<float>> tpf;
    std::vector<int> t3v;
    ilp_sym = true;
    fill_params(ipd, tpd, t3v);
    fill_params(ipf, tpf, t3v);
    ip_d.ensure(sizeof(IlpPair<double>) * ipd.size());
    ip_f.ensure(sizeof(IlpPair<float>) * ipf.size());
    tp_d.ensure(sizeof(TersP<double>) * tpd.size());
    ntp = (int)tpd.size();
    tp_f.ensure(sizeof(TersP<float>) * tpf.size());
    t3.ensure(sizeof(int) * t3v.size());
    rmass.ensure(sizeof(double) * MAXT);
    errw.ensure(sizeof(int) * 4);
    ovf.ensure(sizeof(int) * 4);
    std::vector<double> m(MAXT, 1.0);
    for (int t = 0; t < ntypes; t++) m[t] = mass[t];
    HIPCHK(hipMemcpy(ip_d.p, ipd.data(), sizeof(IlpPair<double>) * ipd.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ip_f.p, ipf.data(), sizeof(IlpPair<float>) * ipf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tp_d.p, tpd.data(), sizeof(TersP<double>) * tpd.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tp_f.p, tpf.data(), sizeof(TersP<float>) * tpf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t3.p, t3v.data(), sizeof(int) * t3v.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(rmass.p, m.data(), sizeof(double) * MAXT, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(errw.p, 0, sizeof(int) * 4));
    cutoffs(cutneigh, short_reach);
    short_cut = short_reach + skin;
    params_dirty = false;
    gen++;  // captured graphs hold the old tables' launch arguments
  }

  // neighbor cutoff (pair cutoff + skin) and the short-list reach the current settings give
  void cutoffs(double &cutn, double &reach) const {
    double rc_max = 0.0;
    for (double c : ilp.cutILPsq) rc_max = std::max(rc_max, std::sqrt(c));
    cutn = std::max(ilp_cut, ters.cutmax) + skin;
    reach = std::max(ters.cutmax, rc_max);
  }

  // Called by every setter: the device tables are refreshed as soon as the settings are complete
  // (and after every later change), so evaluations never see stale masses or coefficients.  The
  // setters refuse cutoff changes while systems are resident (they were binned for the old ones).
  void params_changed() {
    params_dirty = true;
    if (!(have_settings && have_tersoff && have_ilp && (int)mass.size() == ntypes)) return;
    upload_params();
    forces_valid = false;
  }

  // ---------------------------------------------------------------- layout
  // Re-packs the device arrays: resident systems keep their current (sorted) atom order and
  // are moved with device-to-device copies; systems added since the last layout are
  // uploaded from their host staging copies and appended.
  void relayout() {
    if (!layout_dirty) return;
    upload_params();
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<SysHost> next;
    std::vector<int> src_begin;
    std::vector<const Staged *> src_stage;
    for (auto &s : sys)
      if (s.live && !s.staged) {
        next.push_back(s);
        src_begin.push_back(s.begin);
        src_stage.push_back(nullptr);
      }
    for (auto &st : staged) {
      SysHost *s = find(st.id);
      if (!s) continue;
      next.push_back(*s);
      next.back().staged = false;
      src_begin.push_back(-1);
      src_stage.push_back(&st);
    }
    int64_t nn = 0;
    int nc = 0, nr = 0, nt = 0;
    for (auto &s : next) {
      s.begin = (int)nn;
      nn += s.n;
      s.end = (int)nn;
      s.cell_off = nc;
      nc += s.ncells;
      s.row_off = nr;
      nr += s.nc[1] * s.nc[2];
      s.tile_off = nt;
      if (s.tile[0] > 0) nt += s.ntile[0] * s.ntile[1] * s.nc[2];
    }
    if (nn >= (int64_t)IDX_MASK) fail(LMFF_ERR_MEMORY, "too many atoms (%lld) for 26-bit list indices", (long long)nn);
    const size_t m1 = std::max<size_t>((size_t)nn, 1);
    DBuf nx, nv, nmol, ntag, nimg, norig, nsys, nf;
    for (DBuf *b : {&nx, &nv, &nmol, &ntag, &nimg, &norig, &nsys, &nf}) b->tally = &bytes_in_use;
    nf.ensure(m1 * sizeof(double4));
    HIPCHK(hipMemsetAsync(nf.p, 0, m1 * sizeof(double4), stream));
    std::vector<std::pair<int, int>> keep_f;
    nx.ensure(m1 * sizeof(double4));
    nv.ensure(m1 * sizeof(double4));
    nmol.ensure(m1 * sizeof(int));
    ntag.ensure(m1 * sizeof(int));
    nimg.ensure(m1 * sizeof(int));
    norig.ensure(m1 * sizeof(int));
    nsys.ensure(m1 * sizeof(int));
    for (size_t k = 0; k < next.size(); k++) {
      const auto &s = next[k];
      const size_t m = (size_t)s.n;
      if (m == 0) continue;
      if (src_begin[k] >= 0) {
        const size_t o = (size_t)src_begin[k];
        HIPCHK(hipMemcpyAsync(nx.as<double4>() + s.begin, x.as<double4>() + o, m * sizeof(double4), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(nv.as<double4>() + s.begin, v.as<double4>() + o, m * sizeof(double4), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(nmol.as<int>() + s.begin, mol.as<int>() + o, m * sizeof(int), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(ntag.as<int>() + s.begin, tag.as<int>() + o, m * sizeof(int), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(nimg.as<int>() + s.begin, img.as<int>() + o, m * sizeof(int), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(norig.as<int>() + s.begin, orig.as<int>() + o, m * sizeof(int), hipMemcpyDeviceToDevice, stream));
        if (forces_valid) {
          // the force of the last step (pair + post_force terms such as Langevin) stays with the system
          HIPCHK(hipMemcpyAsync(nf.as<double4>() + s.begin, f.as<double4>() + o, m * sizeof(double4), hipMemcpyDeviceToDevice, stream));
          keep_f.push_back({s.begin, s.end});
        }
      } else {
        const Staged &st = *src_stage[k];
        std::vector<double4> hx(m), hv(m);
        std::vector<int> ho(m);
        for (size_t i = 0; i < m; i++) {
          hx[i] = {st.x[3 * i], st.x[3 * i + 1], st.x[3 * i + 2], (double)st.type[i]};
          hv[i] = {st.v[3 * i], st.v[3 * i + 1], st.v[3 * i + 2], 0.0};
          ho[i] = (int)i;
        }
        HIPCHK(hipMemcpy(nx.as<double4>() + s.begin, hx.data(), m * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(nv.as<double4>() + s.begin, hv.data(), m * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(nmol.as<int>() + s.begin, st.mol.data(), m * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ntag.as<int>() + s.begin, st.tag.data(), m * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(nimg.as<int>() + s.begin, st.img.data(), m * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(norig.as<int>() + s.begin, ho.data(), m * sizeof(int), hipMemcpyHostToDevice));
      }
      hipLaunchKernelGGL(k_fill_int, dim3((m + BS - 1) / BS), dim3(BS), 0, stream, (int)m, nsys.as<int>() + s.begin,
                         (int)k);
    }
    HIPCHK(hipStreamSynchronize(stream));
    adopt(x, nx);
    adopt(v, nv);
    adopt(mol, nmol);
    adopt(tag, ntag);
    adopt(img, nimg);
    adopt(orig, norig);
    adopt(sysid, nsys);
    adopt(fsave, nf);
    const bool had_new = !staged.empty();
    staged.clear();
    sys = next;
    n = nn;
    ncells = nc;
    nrows_cells = nr;
    ntiles = nt;
    alloc_atoms((size_t)n);
    win.ensure(sizeof(int) * (size_t)std::max(ntiles, 1) * WSTR);
    winmax.ensure(sizeof(int) * std::max<size_t>(sys.size(), 1));
    syshalf.ensure(sizeof(int) * std::max<size_t>(sys.size(), 1));
    HIPCHK(hipMemsetAsync(syshalf.p, 0, sizeof(int) * std::max<size_t>(sys.size(), 1), stream));
    half_sys.assign(sys.size(), 0);
    half_any = false;
    pull_any = true;
    wcap = 0;
    half_ovf = false;
    if (n) HIPCHK(hipMemcpyAsync(f.p, fsave.p, n * sizeof(double4), hipMemcpyDeviceToDevice, stream));
    restore_f = had_new ? keep_f : std::vector<std::pair<int, int>>{};
    upload_systab();
    // reduction chunks: never straddle a system
    std::vector<Chunk> ch;
    std::vector<int> cb;
    for (size_t k = 0; k < sys.size(); k++) {
      cb.push_back((int)ch.size());
      for (int b = sys[k].begin; b < sys[k].end; b += CHUNK) ch.push_back(Chunk{b, std::min(b + CHUNK, sys[k].end), (int)k, 0});
    }
    cb.push_back((int)ch.size());
    nchunks = (int)ch.size();
    chunks.ensure(sizeof(Chunk) * std::max<size_t>(ch.size(), 1));
    cbeg.ensure(sizeof(int) * cb.size());
    partial.ensure(sizeof(double2) * std::max(nchunks, 1) * 16);
    sums.ensure(sizeof(double) * std::max<size_t>(sys.size(), 1) * 16);
    if (!ch.empty()) HIPCHK(hipMemcpy(chunks.p, ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(cbeg.p, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
    size_cells();
    layout_dirty = false;
    need_rebuild = true;
    regrid_due = true;
    forces_valid = forces_valid && !had_new;
    gen++;
  }

  // cell ranges, row molecule ranges and the sort's key bits and scratch for ncells cells
  void size_cells() {
    cell_start.ensure(sizeof(int) * (size_t)(ncells + 1));
    row_mol.ensure(sizeof(int2) * (size_t)std::max(nrows_cells, 1));
    end_bit = 1;
    while ((1ll << end_bit) <= (long long)ncells) end_bit++;
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr, (int *)nullptr,
                                              (int *)nullptr, (int)std::max<int64_t>(n, 1), 0, end_bit, stream));
    sort_tmp.ensure(tb);
    sort_tmp_bytes = tb;
  }

  // Re-fits the bins of the non-periodic sides of every system of at least half_min_atoms()
  // atoms to its atoms' current extent (grid_from_extent), at the start of a list build: a
  // stack whose sheets expand or ripple would otherwise pile its outer sheets into the edge
  // bins, whose windows then outgrow the half-pair kernel's LDS.  The cell, row and tile offsets
  // of every system follow; a change in their counts invalidates captured steps.
  void regrid() {
    // sheets breathe and ripple over hundreds of steps: re-fit at every REGRID_EVERY-th build, and
    // at the first build after a layout change or positions from the host
    if (!regrid_due && ++builds_since_regrid < REGRID_EVERY) return;
    regrid_due = false;
    builds_since_regrid = 0;
    std::vector<int> which;
    for (size_t k = 0; k < sys.size(); k++) {
      const SysHost &s = sys[k];
      if (s.n > 0 && s.n >= half_min_atoms() && !(s.per[0] && s.per[1] && s.per[2])) which.push_back((int)k);
    }
    if (which.empty()) return;
    const size_t m = which.size();
    // the sheets of a stack: its molecules, when there are 2 .. MAX_SHEETS of them (looked up
    // once per system; the ILP tells layers apart by molecule, so a stack has one per sheet)
    std::vector<int> sheet_off(m, -1);
    int nsheet = 0;
    for (size_t q = 0; q < m; q++) {
      SysHost &s = sys[which[q]];
      if (s.per[2]) continue;
      if (s.sheet_state == 0) {
        auto ids = download<int>(mol, (size_t)s.begin, (size_t)s.n);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        s.sheet_state = (ids.size() >= 2 && ids.size() <= (size_t)MAX_SHEETS) ? 1 : -1;
        if (s.sheet_state == 1) s.sheet_ids = ids;
      }
      if (s.sheet_state == 1) {
        sheet_off[q] = nsheet;
        nsheet += (int)s.sheet_ids.size();
      }
    }
    // minima of every re-fitted system first, then the maxima
    const size_t eb = sizeof(unsigned long long) * 3 * m;
    extent_d.ensure(2 * eb);
    regrid_host.ensure(std::max(2 * eb, sizeof(SysDev) * sys.size()));
    HIPCHK(hipMemsetAsync(extent_d.p, 0xff, eb, stream));
    HIPCHK(hipMemsetAsync((char *)extent_d.p + eb, 0, eb, stream));
    unsigned long long *emin = extent_d.as<unsigned long long>(), *emax = emin + 3 * m;
    const size_t sb = sizeof(unsigned long long) * 2 * (size_t)nsheet;
    if (nsheet) {
      sheet_d.ensure(sb + sizeof(int) * (size_t)nsheet);
      sheet_host.ensure(sb + sizeof(int) * (size_t)nsheet);
      int *hid = (int *)((char *)sheet_host.p + sb);
      for (size_t q = 0; q < m; q++)
        if (sheet_off[q] >= 0)
          std::copy(sys[which[q]].sheet_ids.begin(), sys[which[q]].sheet_ids.end(), hid + sheet_off[q]);
      HIPCHK(hipMemsetAsync(sheet_d.p, 0, sb, stream));
      HIPCHK(hipMemcpyAsync((char *)sheet_d.p + sb, hid, sizeof(int) * (size_t)nsheet, hipMemcpyHostToDevice, stream));
    }
    for (size_t q = 0; q < m; q++) {
      const SysHost &s = sys[which[q]];
      const int blocks = (int)std::min<int64_t>(256, (s.n + BS - 1) / BS);
      hipLaunchKernelGGL(k_extent, dim3(blocks), dim3(BS), 0, stream, s.begin, s.end, (const double4 *)x.p,
                         emin + 3 * q, emax + 3 * q);
      if (sheet_off[q] >= 0)
        hipLaunchKernelGGL(k_sheet_z, dim3(blocks), dim3(BS), 0, stream, s.begin, s.end, (const double4 *)x.p,
                           (const int *)mol.p, (const int *)((char *)sheet_d.p + sb) + sheet_off[q],
                           (int)s.sheet_ids.size(), sheet_d.as<unsigned long long>() + 2 * sheet_off[q]);
    }
    HIPCHK(hipMemcpyAsync(regrid_host.p, extent_d.p, 2 * eb, hipMemcpyDeviceToHost, stream));
    if (nsheet) HIPCHK(hipMemcpyAsync(sheet_host.p, sheet_d.p, sb, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<unsigned long long> ex((const unsigned long long *)regrid_host.p,
                                       (const unsigned long long *)regrid_host.p + 6 * m);
    const unsigned long long *zs = (const unsigned long long *)sheet_host.p;
    auto value = [](unsigned long long k) {
      const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
      double v;
      std::memcpy(&v, &b, sizeof v);
      return v;
    };
    bool changed = false, counts = false;
    for (size_t q = 0; q < m; q++) {
      SysHost &s = sys[which[q]];
      double mn[3], mx[3];
      for (int d = 0; d < 3; d++) {
        mn[d] = value(ex[3 * q + d]);
        mx[d] = value(ex[3 * m + 3 * q + d]);
        if (!std::isfinite(mn[d]) || !std::isfinite(mx[d])) return;  // k_wrap_key reports the atom
      }
      const SysHost old = s;
      grid_from_extent(s, mn, mx, s.n, cutneigh);
      if (sheet_off[q] >= 0) {
        // one z bin per sheet, centred on the sheets' mean heights, when they are evenly spaced
        // (within a factor 1.5) and far enough apart for the list build's stencil (at most 8 bins
        // of reach): each sheet keeps its bin while it ripples by less than half the spacing
        const int nl = (int)s.sheet_ids.size();
        std::vector<double> zm(nl);
        bool ok = true;
        for (int r = 0; r < nl; r++) {
          const unsigned long long c = zs[2 * (sheet_off[q] + r) + 1];
          ok = ok && c > 0;
          zm[r] = c ? (double)(long long)zs[2 * (sheet_off[q] + r)] / 1048576.0 / (double)c : 0.0;
        }
        std::sort(zm.begin(), zm.end());
        double dmin = 1e300, dmax = 0.0;
        for (int r = 0; r + 1 < nl; r++) {
          dmin = std::min(dmin, zm[r + 1] - zm[r]);
          dmax = std::max(dmax, zm[r + 1] - zm[r]);
        }
        const double bsz = ok ? (zm[nl - 1] - zm[0]) / (nl - 1) : 0.0;
        if (ok && dmin > 0 && dmax <= 1.5 * dmin && bsz >= cutneigh / 8.0) {
          s.nc[2] = nl;
          s.bs[2] = bsz;
          s.glo[2] = zm[0] - 0.5 * bsz;
          s.ncells = s.nc[0] * s.nc[1] * s.nc[2];
        }
      }
      for (int d = 0; d < 3; d++) {
        changed |= s.nc[d] != old.nc[d] || s.glo[d] != old.glo[d] || s.bs[d] != old.bs[d];
        counts |= s.nc[d] != old.nc[d];
      }
      counts |= s.tile[0] != old.tile[0] || s.ntile[0] != old.ntile[0] || s.ntile[1] != old.ntile[1];
    }
    if (!changed) return;
    if (counts) {
      int nc = 0, nr = 0, nt = 0;
      for (auto &s : sys) {
        s.cell_off = nc;
        nc += s.ncells;
        s.row_off = nr;
        nr += s.nc[1] * s.nc[2];
        s.tile_off = nt;
        if (s.tile[0] > 0) nt += s.ntile[0] * s.ntile[1] * s.nc[2];
      }
      ncells = nc;
      nrows_cells = nr;
      ntiles = nt;
      size_cells();
      win.ensure(sizeof(int) * (size_t)std::max(ntiles, 1) * WSTR);
      gen++;  // launch sizes and buffers of a captured step change
    }
    // the grid fields of the device's system table (nc up to begin: bins, offsets, tiles; the
    // step counters the device advances are left alone), from page-locked staging that no
    // other copy uses before the next list build
    SysDev *img = (SysDev *)regrid_host.p;
    const size_t f0 = offsetof(SysDev, nc), f1 = offsetof(SysDev, begin);
    for (size_t k = 0; k < sys.size(); k++) {
      img[k] = sysdev_of(sys[k]);
      HIPCHK(hipMemcpyAsync((char *)systab.p + k * sizeof(SysDev) + f0, (const char *)(img + k) + f0, f1 - f0,
                            hipMemcpyHostToDevice, stream));
    }
  }

  static void adopt(DBuf &dst, DBuf &src) {
    std::swap(dst.p, src.p);
    std::swap(dst.bytes, src.bytes);
  }

  // Bytes per atom of everything alloc_atoms() sizes (term for term), plus the radix sort's
  // alternate key/value buffers and list rows of scap + icap entries.
  static double atom_bytes(int scap_, int icap_, int nlayers = 2) {
    const double d4 = 9 * 32.0;                                     // x v f xhold tmp4 nrm ft fil S
    const double i32 = 16 * 4.0;                                    // mol tag img sysid orig tmpi key key2
                                                                    // val perm scnt icnt ocnt tcnt rtcnt rncnt
    const double f64 = 8.0 * (27 + 9 + 9 + 6 + NTH + 6 + MAXSH + 3 * MAXSH * MAXSH);  // dn dndri vet T th
                                                                    // vatom eb fb
    const double u32 = 4.0 * (4 + 3 * MAXSH);                       // nnb tsh rtc rnc
    const double sort = 8.0;
    // half-pair ILP: partner sums (6 int64), per-dz segments (2 x NSEG uint16), one window slot
    // per ILP list entry (uint16), window tables (WSTR ints per 6 x 3-bin tile and z bin, about
    // 60 atoms per table in a stack of sheets with an empty z bin between some of them)
    const double half = 48.0 + 4.0 * NSEG + 2.0 * (double)icap_ + 4.0 * WSTR / 60.0 +
                        8.0 * half_pstr(true, true);  // + the atoms' own sums
    return d4 + i32 + f64 + u32 + sort + half + 4.0 * ((double)scap_ + (double)icap_);
  }
  // ILP row capacity the list build settles on for the worst atom of an AB-stacked pile of
  // nlayers sheets 3.2 A apart with at most 0.39 atoms/A^2 per sheet (graphene 0.382, hBN
  // 0.370): the atoms of the other sheets within cutneigh, grown the way build_lists() grows.
  static int ilp_row_estimate(double cutneigh, int nlayers) {
    const double rho = 0.39, gap = 3.2, pi = 3.141592653589793;
    double worst = 0.0;
    for (int p = 0; p < nlayers; p++) {
      double cnt = 0.0;
      for (int q = 0; q < nlayers; q++) {
        const double dz = gap * (p - q);
        if (q != p && dz * dz < cutneigh * cutneigh) cnt += pi * (cutneigh * cutneigh - dz * dz) * rho;
      }
      worst = std::max(worst, cnt);
    }
    return ((int)(worst * 1.15) + 64 + 63) / 64 * 64;
  }

  void alloc_atoms(size_t m) {
    m = std::max<size_t>(m, 1);
    x.ensure(m * sizeof(double4)); v.ensure(m * sizeof(double4)); f.ensure(m * sizeof(double4));
    xhold.ensure(m * sizeof(double4)); tmp4.ensure(m * sizeof(double4));
    mol.ensure(m * sizeof(int)); tag.ensure(m * sizeof(int)); img.ensure(m * sizeof(int));
    sysid.ensure(m * sizeof(int));
    orig.ensure(m * sizeof(int)); tmpi.ensure(m * sizeof(int));
    key.ensure(m * sizeof(uint32_t)); key2.ensure(m * sizeof(uint32_t));
    val.ensure(m * sizeof(int)); perm.ensure(m * sizeof(int));
    scnt.ensure(m * sizeof(int)); icnt.ensure(m * sizeof(int)); ocnt.ensure(m * sizeof(int));
    if (!farflag.p) {
      farflag.ensure(sizeof(int));
      HIPCHK(hipMemset(farflag.p, 0, sizeof(int)));
    }
    slist.ensure(m * scap * sizeof(uint32_t));
    nrm.ensure(m * sizeof(double4));
    dn.ensure(m * 27 * sizeof(double)); dndri.ensure(m * 9 * sizeof(double)); vet.ensure(m * 9 * sizeof(double));
    nnb.ensure(m * 4 * sizeof(uint32_t)); tsh.ensure(m * MAXSH * sizeof(uint32_t)); tcnt.ensure(m * sizeof(int));
    rtc.ensure(m * MAXSH * sizeof(uint32_t)); rtcnt.ensure(m * sizeof(int));
    rnc.ensure(m * MAXSH * sizeof(uint32_t)); rncnt.ensure(m * sizeof(int));
    eb.ensure(m * MAXSH * sizeof(double)); fb.ensure(m * MAXSH * MAXSH * 3 * sizeof(double));
    ft.ensure(m * sizeof(double4));
    fil.ensure(m * sizeof(double4)); S.ensure(m * sizeof(double4)); T.ensure(m * 6 * sizeof(double));
    th.ensure(m * NTH * sizeof(double)); vatom.ensure(m * 6 * sizeof(double));
    segf.ensure(m * NSEG * sizeof(uint16_t)); segb.ensure(m * NSEG * sizeof(uint16_t));
    if (acc.bytes < m * 6 * sizeof(unsigned long long)) {
      acc.ensure(m * 6 * sizeof(unsigned long long));
      // the half-pair kernel's partner sums start from zero; k_ilp_combine clears them after use
      HIPCHK(hipMemsetAsync(acc.p, 0, acc.bytes, stream));
    }
  }

  void upload_systab() {
    std::vector<SysDev> t(sys.size());
    for (size_t k = 0; k < sys.size(); k++) t[k] = sysdev_of(sys[k]);
    systab.ensure(sizeof(SysDev) * std::max<size_t>(t.size(), 1));
    if (!t.empty()) HIPCHK(hipMemcpy(systab.p, t.data(), sizeof(SysDev) * t.size(), hipMemcpyHostToDevice));
  }

  static SysDev sysdev_of(const SysHost &s) {
    SysDev d{};
    for (int a = 0; a < 3; a++) {
      d.lo[a] = s.lo[a];
      d.prd[a] = s.hi[a] - s.lo[a];
      d.per[a] = s.per[a];
      d.nc[a] = s.nc[a];
      d.glo[a] = s.glo[a];
      d.inv_bs[a] = 1.0 / s.bs[a];
    }
    d.cell_off = s.cell_off;
    d.row_off = s.row_off;
    d.tile[0] = s.tile[0];
    d.tile[1] = s.tile[1];
    d.ntile[0] = s.ntile[0];
    d.ntile[1] = s.ntile[1];
    d.tile_off = s.tile_off;
    d.begin = s.begin;
    d.end = s.end;
    d.dtv = s.dt;
    d.dtf = 0.5 * s.dt * FTM2V;
    d.style = s.style;
    d.t_target = s.t_target;
    d.damp = s.damp;
    d.seed = s.seed;
    d.step = s.step;
    return d;
  }

  void sync_steps_from_device() {
    if (sys.empty()) return;
    std::vector<SysDev> t(sys.size());
    HIPCHK(hipMemcpy(t.data(), systab.p, sizeof(SysDev) * t.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < sys.size(); k++) sys[k].step = t[k].step;
  }

  // ---------------------------------------------------------------- device pointers
  template <typename R>
  DevPtrs<R> ptrs() {
    DevPtrs<R> P{};
    P.n = (int)n;
    P.x = x.as<double4>(); P.v = v.as<double4>(); P.f = f.as<double4>();
    P.mol = mol.as<int>(); P.tag = tag.as<int>(); P.sys = sysid.as<int>();
    P.sys_tab = systab.as<SysDev>();
    P.slist = slist.as<uint32_t>(); P.scnt = scnt.as<int>(); P.scap = scap;
    P.ilist = ilist.as<uint32_t>(); P.icnt = icnt.as<int>(); P.ocnt = ocnt.as<int>(); P.icap = icap;
    P.nrm = (typename std::conditional<sizeof(R) == 8, double4, float4>::type *)nrm.p;
    P.dn = dn.as<R>(); P.dndri = dndri.as<R>(); P.vet = vet.as<double>();
    P.nnb = nnb.as<uint32_t>(); P.tsh = tsh.as<uint32_t>(); P.tcnt = tcnt.as<int>();
    P.rtc = rtc.as<uint32_t>(); P.rtcnt = rtcnt.as<int>(); P.rnc = rnc.as<uint32_t>(); P.rncnt = rncnt.as<int>();
    P.eb = eb.as<double>(); P.fb = fb.as<double>(); P.wb = wb.as<double>(); P.ft = ft.as<double4>();
    P.fil = fil.as<double4>(); P.S = S.as<double4>(); P.T = T.as<double>();
    P.th = th.as<double>(); P.vatom = vatom.as<double>();
    P.err = errw.as<int>();
    P.ovf = ovf.as<int>();
    P.farflag = farflag.as<int>();
    P.cell_start = cell_start.as<int>();
    P.lidx = lidx.as<uint16_t>();
    P.segf = segf.as<uint16_t>();
    P.segb = segb.as<uint16_t>();
    P.win = ntiles > 0 ? win.as<int>() : nullptr;
    P.winmax = winmax.as<int>();
    P.syshalf = syshalf.as<int>();
    P.acc = acc.as<unsigned long long>();
    P.ipart = ipart.as<double>();
    P.npass = npass;
    P.fx_scale = FX_SCALE;
    P.fx_magic = FX_MAGIC;
    P.nsys = (int)sys.size();
    P.half_on = 0;
    scalars(P);
    return P;
  }

  // parameter tables, cutoffs and Tap coefficients (everything of DevPtrs but the per-atom arrays)
  template <typename R>
  void scalars(DevPtrs<R> &P) {
    P.ip = (const IlpPair<R> *)(sizeof(R) == 8 ? ip_d.p : ip_f.p);
    P.tp = (const TersP<double> *)tp_d.p;
    P.ntp = ntp;
    P.t3 = t3.as<int>();
    P.cutneighsq = cutneigh * cutneigh;
    P.shortsq = short_cut * short_cut;
    P.tersoff_cutmaxsq = ters.cutmax * ters.cutmax;
    P.ilp_cutsq = ilp_cut * ilp_cut;
    P.ilp_splitsq = (ilp_cut + ilp_margin()) * (ilp_cut + ilp_margin());
    // the per-step tests reach to cutmax and rcut widened by 1e-12 in r^2 (reverse candidates)
    const double sr = short_reach * (1.0 + 1e-9) + ilp_margin();
    P.short_splitsq = sr * sr;
    P.rcut = std::sqrt(ilp_cut * ilp_cut);
    P.inv_rcut = 1.0 / P.rcut;
    P.tap_flag = tap_flag;
    P.ntypes = ntypes;
    P.ters_rep_sym = ters_rep_sym ? 1 : 0;
    P.rmass = rmass.as<double>();
    {
      const double rc = P.rcut, rc4 = rc * rc * rc * rc;
      P.tap_poly[0] = -35.0 / rc4;
      P.tap_poly[1] = 84.0 / (rc4 * rc);
      P.tap_poly[2] = -70.0 / (rc4 * rc * rc);
      P.tap_poly[3] = 20.0 / (rc4 * rc * rc * rc);
      P.tap_poly[4] = 140.0 / (rc4 * rc * rc * rc);
    }
  }

  // ILP rows keep pairs with r < rc + margin at the build in front (short lists: r < max(Tersoff
  // cutmax, ILP normal rcut) + margin); the rest is only walked once
  // some atom has moved more than margin/2 (the far flag).  The margin trades the length of the
  // front segment against how soon thermal motion raises the flag (0.3 A of displacement is
  // beyond what a 300 K atom travels in a 10-step rebuild interval); results do not depend on it.
  double ilp_margin() const { return std::min(0.6, skin); }
  void far_check() {
    const double h = 0.5 * ilp_margin();
    hipLaunchKernelGGL(k_displacement, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, (const double4 *)x.p,
                       (const double4 *)xhold.p, sysid.as<int>(), h * h, farflag.as<int>());
  }

  // ---------------------------------------------------------------- profiling helpers
  void mark_begin(int fam, size_t &slot) {
    if (!prof) return;
    if (ev_used + 2 > ev_pool.size()) {
      size_t add = std::max<size_t>(ev_pool.size(), 256);
      for (size_t k = 0; k < add; k++) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ev_pool.push_back(e);
      }
    }
    slot = ev_used;
    ev_used += 2;
    HIPCHK(hipEventRecord(ev_pool[slot], stream));
    (void)fam;
  }
  void mark_end(int fam, size_t slot) {
    if (!prof) return;
    HIPCHK(hipEventRecord(ev_pool[slot + 1], stream));
    ev_marks.push_back({fam, {slot, slot + 1}});
  }

  // ---------------------------------------------------------------- rebuild
  void rebuild() {
    if (n == 0) {
      need_rebuild = false;
      return;
    }
    size_t mk = 0;
    mark_begin(0, mk);
    regrid();
    const int nb = (int)((n + BS - 1) / BS);
    hipLaunchKernelGGL(k_wrap_key, dim3(nb), dim3(BS), 0, stream, (int)n, x.as<double4>(), img.as<int>(),
                       sysid.as<int>(), systab.as<SysDev>(), key.as<uint32_t>(), val.as<int>(), errw.as<int>());
    size_t tb = sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, tb, key.as<uint32_t>(), key2.as<uint32_t>(), val.as<int>(),
                                              perm.as<int>(), (int)n, 0, end_bit, stream));
    permute4(x);
    permute4(v);
    permute4(f);  // forces kept across a set-up rebuild stay with their atoms
    if (!restore_f.empty()) permute4(fsave);
    permutei(mol);
    permutei(tag);
    permutei(img);
    permutei(orig);
    if (cg_active) {
      permute4(cg_x0);
      permute4(cg_g);
      permute4(cg_h);
      hipLaunchKernelGGL(k_image_follow, dim3(nb), dim3(BS), 0, stream, (int)n, (const double4 *)x.p,
                         cg_x0.as<double4>(), sysid.as<int>(), systab.as<SysDev>());
    }
    hipLaunchKernelGGL(k_cell_start, dim3((ncells + 1 + BS - 1) / BS), dim3(BS), 0, stream, (int)n, ncells,
                       key2.as<uint32_t>(), cell_start.as<int>());
    hipLaunchKernelGGL(k_row_mol, dim3((nrows_cells + 3) / 4), dim3(BS), 0, stream, nrows_cells, (int)sys.size(),
                       (const SysDev *)systab.as<SysDev>(), (const int *)cell_start.as<int>(), (const int *)mol.as<int>(),
                       row_mol.as<int2>());
    if (ntiles > 0) {
      hipLaunchKernelGGL(k_fill_int, dim3(((int)sys.size() + BS - 1) / BS), dim3(BS), 0, stream, (int)sys.size(),
                         winmax.as<int>(), 0);
      hipLaunchKernelGGL(k_win_table, dim3((ntiles + 3) / 4), dim3(BS), 0, stream, ntiles, (int)sys.size(),
                         (const SysDev *)systab.as<SysDev>(), (const int *)cell_start.as<int>(),
                         std::sqrt(cutneigh * cutneigh), win.as<int>(), winmax.as<int>());
    }
    HIPCHK(hipMemcpyAsync(xhold.p, x.p, n * sizeof(double4), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, 1, farflag.as<int>(), 0);
    build_lists();
    mark_end(0, mk);
    need_rebuild = false;
    ago = 0;
    ncalls++;
  }

  void permute4(DBuf &b) {
    hipLaunchKernelGGL(k_permute<double4>, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, perm.as<int>(),
                       b.as<double4>(), tmp4.as<double4>());
    HIPCHK(hipMemcpyAsync(b.p, tmp4.p, n * sizeof(double4), hipMemcpyDeviceToDevice, stream));
  }
  void permutei(DBuf &b) {
    hipLaunchKernelGGL(k_permute<int>, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, perm.as<int>(),
                       b.as<int>(), tmpi.as<int>());
    HIPCHK(hipMemcpyAsync(b.p, tmpi.p, n * sizeof(int), hipMemcpyDeviceToDevice, stream));
  }

  // Which systems run their ILP pairs in the half-pair kernel: those with tiles, whose stencil
  // reaches at most NSEG - 2 z bins up, and whose list rows fit the 16-bit segment counts.  The
  // kernel's LDS holds the largest window up to the size that keeps two workgroups per CU
  // (half_lds_cap; LMFF_HALF_WCAP in the environment lowers that limit); slots beyond it add to
  // the global sums directly.  Decided per system from its own geometry (on the
  // device, k_half_decide, before the list build, which keeps half rows for those systems), so a
  // system's results do not depend on what else is packed with it; here the host takes the
  // decisions back and sizes the launches.
  void choose_half(const std::vector<int> &wm, const std::vector<int> &hs) {
    int wc = 0, np = 0;
    const double cut = std::sqrt(cutneigh * cutneigh);
    for (size_t k = 0; k < sys.size(); k++) {
      const SysHost &q = sys[k];
      const int rz = (int)std::ceil(cut * (1.0 / q.bs[2]));
      if (hs[k]) {
        wc = std::max(wc, wm[k]);
        // passes: z-bin offsets 0 .. rz (within the bins in a non-periodic z)
        np = std::max(np, q.per[2] ? rz + 1 : std::min(rz + 1, q.nc[2]));
      }
    }
    bool any = false, pull = false;
    for (size_t k = 0; k < sys.size(); k++) {
      if (hs[k]) any = true;
      else if (sys[k].n > 0) pull = true;
    }
    const int wmax = wc;
    if (any) half_attrs();  // here, never inside a graph capture
    wc = std::min({(wc + 63) / 64 * 64, half_wcap_limit(), half_lds_cap});
    const bool ovf = any && wmax > wc;
    if (any) {
      // the atoms' own sums at the widest layout (with the virial), so that no later evaluation
      // moves the buffer under a captured graph
      if (ipart.bytes < sizeof(double) * (size_t)n * half_pstr(true, true)) {
        ipart.ensure(sizeof(double) * (size_t)n * half_pstr(true, true));
        gen++;
      }
    }
    if (hs != half_sys || wc != wcap || any != half_any || pull != pull_any || np != npass || ovf != half_ovf) {
      half_sys = hs;
      wcap = wc;
      half_ovf = ovf;
      npass = np;
      half_any = any;
      pull_any = pull;
      gen++;  // the launches of a captured step change
    }
  }

  void build_lists() {
    for (int attempt = 0; attempt < 4; attempt++) {
      slist.ensure((size_t)n * scap * sizeof(uint32_t));
      ilist.ensure((size_t)n * std::max(icap, 1) * sizeof(uint32_t));
      if (ntiles > 0) lidx.ensure((size_t)n * std::max(icap, 1) * sizeof(uint16_t));
      HIPCHK(hipMemsetAsync(ovf.p, 0, sizeof(int) * 4, stream));
      if (ntiles > 0)
        hipLaunchKernelGGL(k_half_decide, dim3(((int)sys.size() + BS - 1) / BS), dim3(BS), 0, stream,
                           (int)sys.size(), (const SysDev *)systab.as<SysDev>(), (const int *)winmax.as<int>(),
                           std::sqrt(cutneigh * cutneigh), (half_enabled && have_ilp && icap < 0xffff) ? 1 : 0,
                           syshalf.as<int>());
      auto P = ptrs<double>();
      const int nb = (int)((n + 3) / 4);
      hipLaunchKernelGGL(k_build_lists<double>, dim3(nb), dim3(BS), 0, stream, P, (const int *)cell_start.as<int>(),
                         (const int2 *)row_mol.as<int2>());
      int o[4];
      HIPCHK(hipMemcpyAsync(o, ovf.p, sizeof o, hipMemcpyDeviceToHost, stream));
      std::vector<int> wm(sys.size(), 0), hs(sys.size(), 0);
      if (ntiles > 0 && !sys.empty()) {
        HIPCHK(hipMemcpyAsync(wm.data(), winmax.p, sizeof(int) * sys.size(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(hs.data(), syshalf.p, sizeof(int) * sys.size(), hipMemcpyDeviceToHost, stream));
      }
      HIPCHK(hipStreamSynchronize(stream));
      if (o[0] <= 0 && o[1] <= 0) choose_half(wm, hs);
      bool grow = false;
      if (o[0] > 0) {
        scap = ((int)(o[0] * 1.25) + 63) / 64 * 64;
        if (scap > 0xffff) fail(LMFF_ERR_MEMORY, "short neighbor list of %d entries per atom", o[0]);
        grow = true;
      }
      if (o[1] > 0) {
        icap = ((int)(o[1] * 1.15) + 64 + 63) / 64 * 64;
        grow = true;
      } else if (icap == 0) {
        icap = 64;  // no atom has an ILP neighbour
      }
      if (!grow) return;
      gen++;  // list pointers / capacities change: captured graphs are stale
    }
    fail(LMFF_ERR_MEMORY, "neighbor list capacity did not converge");
  }

  // ---------------------------------------------------------------- force evaluation
  template <typename R, bool LDSP>
  void launch_bonds(const DevPtrs<R> &P, int ef, int vf, int va, int nb4) {
    if (va) hipLaunchKernelGGL((k_tersoff_bonds<R, true, true, true, LDSP>), dim3(nb4), dim3(BS), 0, stream, P);
    else if (vf) hipLaunchKernelGGL((k_tersoff_bonds<R, true, true, false, LDSP>), dim3(nb4), dim3(BS), 0, stream, P);
    else if (ef) hipLaunchKernelGGL((k_tersoff_bonds<R, true, false, false, LDSP>), dim3(nb4), dim3(BS), 0, stream, P);
    else hipLaunchKernelGGL((k_tersoff_bonds<R, false, false, false, LDSP>), dim3(nb4), dim3(BS), 0, stream, P);
  }

  // integ: the Verlet loop's final integration (and the step advance) runs inside the force
  // gather; want_ke also fills the kinetic-energy column
  template <typename R, bool SYM>
  void launch_ilp(const DevPtrs<R> &P, int ef, int vf, int va, int nbw) {
    if (va) hipLaunchKernelGGL((k_ilp<R, true, true, true, SYM>), dim3(nbw), dim3(BS), 0, stream, P);
    else if (vf) hipLaunchKernelGGL((k_ilp<R, true, true, false, SYM>), dim3(nbw), dim3(BS), 0, stream, P);
    else if (ef) hipLaunchKernelGGL((k_ilp<R, true, false, false, SYM>), dim3(nbw), dim3(BS), 0, stream, P);
    else hipLaunchKernelGGL((k_ilp<R, false, false, false, SYM>), dim3(nbw), dim3(BS), 0, stream, P);
  }

  template <typename R>
  void force_impl(int ef, int vf, int va, int integ, int want_ke, int half, bool full_rows) {
    DevPtrs<R> P = ptrs<R>();
    if (full_rows) {
      P.ilist = ilist_full.as<uint32_t>();
      P.icnt = icnt_full.as<int>();
      P.ocnt = ocnt_full.as<int>();
      P.icap = icap_full;
    }
    force_kernels<R>(P, ef, vf, va, integ, want_ke, half);
  }

  // Full ILP rows of every atom for an evaluation by the pull kernel while the lists of the half
  // systems hold half rows (per-atom tallies), from the cell ranges of the last build: an atom
  // has moved less than half the skin since then, so every pair within rc now lies in the
  // stencil of the atoms' current bins (the rows are exact within rc, not within rc + skin).
  void build_full_rows() {
    if (icap_full < 2 * icap) icap_full = ((2 * icap) + 63) / 64 * 64;
    for (int attempt = 0; attempt < 4; attempt++) {
      ilist_full.ensure((size_t)n * icap_full * sizeof(uint32_t));
      icnt_full.ensure((size_t)n * sizeof(int));
      ocnt_full.ensure((size_t)n * sizeof(int));
      HIPCHK(hipMemsetAsync(ovf.p, 0, sizeof(int) * 4, stream));
      auto P = ptrs<double>();
      P.ilist = ilist_full.as<uint32_t>();
      P.icnt = icnt_full.as<int>();
      P.ocnt = ocnt_full.as<int>();
      P.icap = icap_full;
      P.ilp_only = 1;
      hipLaunchKernelGGL(k_build_lists<double>, dim3((int)((n + 3) / 4)), dim3(BS), 0, stream, P,
                         (const int *)cell_start.as<int>(), (const int2 *)row_mol.as<int2>());
      int o[4];
      HIPCHK(hipMemcpyAsync(o, ovf.p, sizeof o, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      if (o[1] <= 0) return;
      icap_full = ((int)(o[1] * 1.15) + 64 + 63) / 64 * 64;
    }
    fail(LMFF_ERR_MEMORY, "neighbor list capacity did not converge");
  }

  template <typename R, bool EF, bool VF, bool SYM>
  void launch_half_one(const DevPtrs<R> &P) {
    const size_t lds = sizeof(unsigned long long) * 6 * (size_t)wcap;
    if (half_ovf)
      hipLaunchKernelGGL((k_ilp_half<R, EF, VF, SYM, true>), dim3(ntiles), dim3(64 * HalfWaves<R>::n), lds, stream, P,
                         wcap);
    else
      hipLaunchKernelGGL((k_ilp_half<R, EF, VF, SYM, false>), dim3(ntiles), dim3(64 * HalfWaves<R>::n), lds, stream, P,
                         wcap);
  }
  template <typename R, bool SYM>
  void launch_half(const DevPtrs<R> &P, int ef, int vf) {
    if (vf) launch_half_one<R, true, true, SYM>(P);
    else if (ef) launch_half_one<R, true, false, SYM>(P);
    else launch_half_one<R, false, false, SYM>(P);
  }
  // the half-pair kernel's window sums take more than the default 64 KB of LDS per workgroup
  // (and returns the largest static LDS of the variants)
  template <typename R, bool EF, bool VF, bool SYM>
  static size_t half_attr_one() {
    size_t st = 0;
    for (const void *k : {(const void *)k_ilp_half<R, EF, VF, SYM, false>, (const void *)k_ilp_half<R, EF, VF, SYM, true>}) {
      HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(unsigned long long) * 6 * WCAP_MAX)));
      hipFuncAttributes a{};
      HIPCHK(hipFuncGetAttributes(&a, k));
      st = std::max(st, (size_t)a.sharedSizeBytes);
    }
    return st;
  }
  template <typename R>
  static size_t half_attr_prec() {
    return std::max({half_attr_one<R, false, false, false>(), half_attr_one<R, true, false, false>(),
                     half_attr_one<R, true, true, false>(), half_attr_one<R, false, false, true>(),
                     half_attr_one<R, true, false, true>(), half_attr_one<R, true, true, true>()});
  }
  bool half_attr_done = false;
  // window atoms whose sums leave room in a CU's 160 KB of LDS for two workgroups of the kernel
  // (with its static LDS): a larger window would halve the kernel's occupancy
  int half_lds_cap = 64;
  void half_attrs() {
    if (half_attr_done) return;
    const size_t st = std::max(half_attr_prec<double>(), half_attr_prec<float>());
    const size_t per_wg = (size_t)160 * 1024 / 2;
    const size_t room = per_wg > st ? per_wg - st : 0;
    half_lds_cap = std::max(64, std::min(WCAP_MAX, (int)(room / (6 * sizeof(unsigned long long))) / 64 * 64));
    half_attr_done = true;
  }

  // normals -> ILP pair -> Tersoff bonds -> gather over the arrays P points at
  // half: the systems marked half run their ILP pairs in the half-pair kernel (the others, and
  // every system in an evaluation that asks for per-atom tallies, in the pull kernel)
  template <typename R>
  void force_kernels(const DevPtrs<R> &P, int ef, int vf, int va, int integ, int want_ke, int half = 0) {
    const int n = P.n;
    const int nb1 = (int)((n + BS - 1) / BS);
    const int nbw = (int)((n + 3) / 4);
    size_t mk = 0;
    mark_begin(1, mk);
    if (vf || va) hipLaunchKernelGGL((k_normals<R, true>), dim3(nb1), dim3(BS), 0, stream, P);
    else hipLaunchKernelGGL((k_normals<R, false>), dim3(nb1), dim3(BS), 0, stream, P);
    mark_end(1, mk);
    mark_begin(2, mk);
    if (half) {
      DevPtrs<R> H = P;
      H.half_on = 1;
      if (pull_any) {
        if (ilp_sym) launch_ilp<R, true>(H, ef, vf, va, nbw);
        else launch_ilp<R, false>(H, ef, vf, va, nbw);
      }
      if (ilp_sym) launch_half<R, true>(H, ef, vf);
      else launch_half<R, false>(H, ef, vf);
      if (vf) hipLaunchKernelGGL((k_ilp_combine<R, true, true>), dim3(nb1), dim3(BS), 0, stream, H);
      else if (ef) hipLaunchKernelGGL((k_ilp_combine<R, true, false>), dim3(nb1), dim3(BS), 0, stream, H);
      else hipLaunchKernelGGL((k_ilp_combine<R, false, false>), dim3(nb1), dim3(BS), 0, stream, H);
    } else {
      if (ilp_sym) launch_ilp<R, true>(P, ef, vf, va, nbw);
      else launch_ilp<R, false>(P, ef, vf, va, nbw);
    }
    mark_end(2, mk);
    mark_begin(3, mk);
    // the bonds and gather kernels: BOND_ATOMS_PER_WAVE atoms per wave, four waves per block
    const int nb4 = (int)(((n + BOND_ATOMS_PER_WAVE - 1) / BOND_ATOMS_PER_WAVE + BS / 64 - 1) / (BS / 64));
    if (ntp <= TP_LDS) launch_bonds<R, true>(P, ef, vf, va, nb4);
    else launch_bonds<R, false>(P, ef, vf, va, nb4);
    mark_end(3, mk);
    mark_begin(4, mk);
    if (integ) {
      if (va) hipLaunchKernelGGL((k_gather<R, true, true, true>), dim3(nb4), dim3(BS), 0, stream, P, want_ke);
      else if (ef) hipLaunchKernelGGL((k_gather<R, true, false, true>), dim3(nb4), dim3(BS), 0, stream, P, want_ke);
      else hipLaunchKernelGGL((k_gather<R, false, false, true>), dim3(nb4), dim3(BS), 0, stream, P, want_ke);
    } else {
      if (va) hipLaunchKernelGGL((k_gather<R, true, true, false>), dim3(nb4), dim3(BS), 0, stream, P, 0);
      else if (ef) hipLaunchKernelGGL((k_gather<R, true, false, false>), dim3(nb4), dim3(BS), 0, stream, P, 0);
      else hipLaunchKernelGGL((k_gather<R, false, false, false>), dim3(nb4), dim3(BS), 0, stream, P, 0);
    }
    mark_end(4, mk);
  }

  void force(int eflag, int vflag, int integ = 0, int want_ke = 0) {
    if (n == 0) return;
    const int va = (vflag & 4) ? 1 : 0;
    // per-atom virial shares of the Tersoff triplets: only allocated once asked for
    if (va) wb.ensure((size_t)n * MAXSH * MAXSH * 6 * sizeof(double));
    const int vf = ((vflag % 4) != 0 || va) ? 1 : 0;
    const int ef = (eflag != 0 || vf) ? 1 : 0;
    // per-atom energy or virial: the pull kernel tallies both halves of every pair as the
    // reference's ev_tally does; the half-pair kernel gives a pair's energy to one atom
    const int half = (half_any && !va && !(eflag & 2)) ? 1 : 0;
    const bool full_rows = half_any && !half;
    if (full_rows) build_full_rows();
    if (precision == LMFF_MIXED) force_impl<float>(ef, vf, va, integ, want_ke, half, full_rows);
    else force_impl<double>(ef, vf, va, integ, want_ke, half, full_rows);
    last_eflag = ef ? (eflag | 1) : 0;
    last_vflag = va ? 4 | 2 : (vf ? 2 : 0);
    forces_valid = true;
  }

  // retry_ok: an overflow of the half-pair kernel's fixed-point partner sums (and nothing
  // else) moves this context to the pull kernel for good (full rows at the next build, which it
  // requests) and returns true, so that the caller repeats its evaluation; otherwise every error
  // bit ends the call with its message.
  bool check_errors(bool retry_ok = false) {
    int e[4];
    HIPCHK(hipMemcpyAsync(e, errw.p, sizeof e, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (e[0] == 0) return false;
    HIPCHK(hipMemset(errw.p, 0, sizeof(int) * 4));
    if (retry_ok && e[0] == EB_FIXED && half_enabled) {
      half_enabled = false;
      need_rebuild = true;
      return true;
    }
    if (e[0] & EB_NONFINITE) fail(LMFF_ERR_NONFINITE, "Non-numeric atom coords - simulation unstable");
    if (e[0] & EB_NORMAL_COUNT)
      fail(LMFF_ERR_NORMAL, "There are too many neighbors for some atoms, please check your configuration");
    if (e[0] & EB_NORMAL_ZERO) fail(LMFF_ERR_NORMAL, "The magnitude of the normal vector is zero");
    if (e[0] & EB_SHORT) fail(LMFF_ERR_SHORT, "more than %d Tersoff neighbours within cutmax", MAXSH);
    if (e[0] & EB_GRID) fail(LMFF_ERR_STATE, "bin grid too coarse for the neighbor stencil");
    if (e[0] & EB_FIXED)
      fail(LMFF_ERR_NONFINITE,
           "ILP pair terms beyond the fixed-point range of the half-pair kernel (atoms of different molecules "
           "closer than about 1.7 A); LMFF_HALF=0 in the environment evaluates such configurations");
    return false;
  }

  // ---------------------------------------------------------------- thermo
  // columns: 0 e_vdw 1 e_rep 2 e_tersoff 3-8 virial 9 ke
  // per-system sums of the first ncols thermo columns (the last nmax of them: maxima)
  void reduce(int ncols, double *dst_dev, int nmax = 0) {
    if (nchunks == 0) {
      HIPCHK(hipMemsetAsync(dst_dev, 0, sizeof(double) * ncols * std::max<size_t>(sys.size(), 1), stream));
      return;
    }
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nchunks), dim3(BS), 0, stream, (const double *)th.p, (size_t)n, ncols,
                       chunks.as<Chunk>(), partial.as<double2>(), nmax);
    hipLaunchKernelGGL(k_reduce_systems, dim3((unsigned)sys.size()), dim3(64), 0, stream, (const double2 *)partial.p,
                       ncols, cbeg.as<int>(), (int)sys.size(), dst_dev, nmax);
  }

  // ---------------------------------------------------------------- CG minimisation
  // MinCG::iterate (LMFF/min_cg.cpp:40-190) with linemin_quadratic and alpha_step
  // (LMFF/min_linesearch.cpp:324-503, 837-868); one system per context (the line search's
  // evaluation sequence is the system's own).  Lists as Min::init: every 1, delay 0, check yes.
  void minimize_cg(double etol, double ftol, int64_t maxiter, int64_t maxeval, double dmax, int norm,
                   lmff_min_stats &st) {
    relayout();
    if (sys.size() != 1) fail(LMFF_ERR_STATE, "min_style cg needs exactly one system in the context");
    st = lmff_min_stats{};
    if (n == 0) return;
    const double EMACH = 1.0e-8, EPS_QUAD = 1.0e-28, ALPHA_MAX = 1.0, ALPHA_REDUCE = 0.5, BACKTRACK_SLOPE = 0.4,
                 QUADRATIC_TOL = 0.1, eps_energy = 1.0e-8;
    const int64_t nlimit = std::min<int64_t>(2147483647, 3 * (int64_t)n);
    cg_x0.ensure(n * sizeof(double4));
    cg_g.ensure(n * sizeof(double4));
    cg_h.ensure(n * sizeof(double4));
    const int sv_every = every, sv_delay = delay, sv_check = check;
    every = 1;
    delay = 0;
    check = 1;
    cg_active = true;
    struct Restore {
      Ctx &c;
      int e, d, k;
      ~Restore() { c.every = e; c.delay = d; c.check = k; c.cg_active = false; }
    } restore{*this, sv_every, sv_delay, sv_check};
    DBuf out;
    out.ensure(sizeof(double) * 16);
    double hst[9], hd[2];
    const int nb = (int)((n + BS - 1) / BS);
    auto eval = [&]() {  // energy, then f.f, f.h, f.g, max|h_c|, max|f|^2, max|f_c| of the new forces
      ago++;
      if (rebuild_due()) rebuild();
      for (int attempt = 0;; attempt++) {
        if (attempt > 0) rebuild();  // full rows for the pull kernel
        force(1, 0);
        reduce(3, out.as<double>());
        hipLaunchKernelGGL(k_cg_stats, dim3(nb), dim3(BS), 0, stream, (int)n, (const double4 *)f.p,
                           (const double4 *)cg_g.p, (const double4 *)cg_h.p, th.as<double>());
        reduce(6, out.as<double>() + 3, 3);
        HIPCHK(hipMemcpyAsync(hst, out.p, sizeof hst, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (!check_errors(attempt == 0)) break;
      }
      return hst[0] + hst[1] + hst[2];
    };
    int64_t neval = 0;
    const double hm = 0.5 * ilp_margin();
    auto alpha_step = [&](double alpha) {
      hipLaunchKernelGGL(k_cg_alpha_step, dim3(nb), dim3(BS), 0, stream, (int)n, x.as<double4>(),
                         (const double4 *)cg_x0.p, (const double4 *)cg_h.p, alpha, (const double4 *)xhold.p, hm * hm,
                         farflag.as<int>());
      neval++;
      return eval();
    };
    auto fnorm = [&]() { return norm == 1 ? hst[7] : norm == 2 ? hst[8] : hst[3]; };
    // Min::setup: lists and forces of the start (not counted as an evaluation)
    if (need_rebuild) rebuild();
    ago = -1;  // eval() counts this evaluation's step
    double ecurrent = eval();
    hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(BS), 0, stream, (int)n, (const double4 *)f.p, cg_g.as<double4>(),
                       cg_h.as<double4>(), 0.0, 0, th.as<double>());
    double gg = hst[3];
    double fdoth = hst[3], hmax = hst[8];  // h = f
    st.e_initial = st.e_next_to_last = ecurrent;
    st.fnorm2_initial = std::sqrt(hst[3]);
    st.fnorminf_initial = hst[8];
    st.stop = LMFF_STOP_MAXITER;
    // quadratic line search along h from the current point; returns 0 or a stop code
    auto linemin = [&](double eoriginal) -> int {
      const double fdothall = fdoth;
      if (fdothall <= 0.0) return LMFF_STOP_DOWNHILL;
      const double alphamax = std::min(ALPHA_MAX, dmax / hmax);
      if (hmax == 0.0) return LMFF_STOP_ZEROFORCE;
      HIPCHK(hipMemcpyAsync(cg_x0.p, x.p, n * sizeof(double4), hipMemcpyDeviceToDevice, stream));
      double alpha = alphamax, fhprev = fdothall, engprev = eoriginal, alphaprev = 0.0;
      while (true) {
        ecurrent = alpha_step(alpha);
        const double fh = hst[4];
        const double delfh = fh - fhprev;
        if (std::fabs(fh) < EPS_QUAD || std::fabs(delfh) < EPS_QUAD) {
          ecurrent = alpha_step(0.0);
          return LMFF_STOP_ZEROQUAD;
        }
        const double relerr = std::fabs(1.0 - (0.5 * (alpha - alphaprev) * (fh + fhprev) + ecurrent) / engprev);
        const double alpha0 = alpha - (alpha - alphaprev) * fh / delfh;
        if (relerr <= QUADRATIC_TOL && alpha0 > 0.0 && alpha0 < alphamax) {
          ecurrent = alpha_step(alpha0);
          if (ecurrent - eoriginal < EMACH) return 0;
        }
        const double de_ideal = -BACKTRACK_SLOPE * alpha * fdothall;
        const double de = ecurrent - eoriginal;
        if (de <= de_ideal) return 0;
        fhprev = fh;
        engprev = ecurrent;
        alphaprev = alpha;
        alpha *= ALPHA_REDUCE;
        if (alpha <= 0.0 || de_ideal >= -EMACH) {
          ecurrent = alpha_step(0.0);
          return LMFF_STOP_ZEROALPHA;
        }
      }
    };
    int64_t niter = 0;
    double eprevious = ecurrent;
    for (int64_t iter = 0; iter < maxiter; iter++) {
      niter++;
      eprevious = ecurrent;
      const int rc = linemin(ecurrent);
      if (rc) {
        st.stop = rc;
        break;
      }
      if (neval >= maxeval) {
        st.stop = LMFF_STOP_MAXEVAL;
        break;
      }
      if (std::fabs(ecurrent - eprevious) < etol * 0.5 * (std::fabs(ecurrent) + std::fabs(eprevious) + eps_energy)) {
        st.stop = LMFF_STOP_ETOL;
        break;
      }
      const double dot0 = hst[3], dot1 = hst[5];
      if (ftol > 0.0 && fnorm() < ftol * ftol) {
        st.stop = LMFF_STOP_FTOL;
        break;
      }
      double beta = std::max(0.0, (dot0 - dot1) / gg);
      if ((niter + 1) % nlimit == 0) beta = 0.0;
      gg = dot0;
      hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(BS), 0, stream, (int)n, (const double4 *)f.p, cg_g.as<double4>(),
                         cg_h.as<double4>(), beta, 1, th.as<double>());
      reduce(2, out.as<double>() + 12, 1);
      HIPCHK(hipMemcpyAsync(hd, out.as<double>() + 12, sizeof hd, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      if (hd[0] <= 0.0) {
        hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(BS), 0, stream, (int)n, (const double4 *)f.p, cg_g.as<double4>(),
                           cg_h.as<double4>(), 0.0, 2, th.as<double>());
        fdoth = dot0;   // h = g = f
        hmax = hst[8];
      } else {
        fdoth = hd[0];
        hmax = hd[1];
      }
    }
    st.iterations = niter;
    st.evaluations = neval;
    st.e_next_to_last = eprevious;
    st.e_final = ecurrent;
    st.fnorm2_final = std::sqrt(hst[3]);
    st.fnorminf_final = hst[8];
    sync_steps_from_device();
    sys[0].step += niter;
    upload_systab();
    gen++;
    forces_valid = true;
  }

  // ---------------------------------------------------------------- FIRE minimisation
  // MinFire::iterate (LMFF/min_fire.cpp:104-460) with the semi-implicit Euler integrator, run
  // for every resident system at once, each with its own FIRE state; neighbor lists follow
  // Min::init (every 1, delay 0, check yes: LMFF/min.cpp:190-196).
  void minimize(double etol, double ftol, int64_t maxiter, const lmff_fire_params &fp, std::vector<lmff_min_stats> &st) {
    relayout();
    const int nsys = (int)sys.size();
    st.assign(nsys, lmff_min_stats{});
    if (n == 0 || nsys == 0) return;
    const double ftm2v = 1.0 / MVV2E, eps_energy = 1.0e-8;
    struct S {
      double dt, dtmax, dtmin, alpha, dtv_prev = 0, ecur = 0, eprev = 0;
      long long ntimestep, last_negative, start;
      int negatif = 0, done = 0;
    };
    std::vector<S> fs(nsys);
    for (int k = 0; k < nsys; k++) {
      fs[k].dt = sys[k].dt;
      fs[k].dtmax = fp.tmax * sys[k].dt;
      fs[k].dtmin = fp.tmin * sys[k].dt;
      fs[k].alpha = fp.alpha0;
      fs[k].ntimestep = fs[k].last_negative = fs[k].start = sys[k].step;
    }
    const int sv_every = every, sv_delay = delay, sv_check = check;
    every = 1;
    delay = 0;
    check = 1;
    DBuf out, steps;
    out.ensure(sizeof(double) * 16 * nsys);
    steps.ensure(sizeof(FireStep) * nsys);
    std::vector<double> h(16 * (size_t)nsys);
    // energies (columns 0-2 of the thermo block) and the FIRE statistics of the current state
    auto measure = [&]() {
      for (int attempt = 0;; attempt++) {
        if (attempt > 0) rebuild();  // full rows for the pull kernel
        force(1, 0);
        reduce(3, out.as<double>());
        hipLaunchKernelGGL(k_fire_stats, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n,
                           (const double4 *)v.p, (const double4 *)f.p, th.as<double>());
        reduce(6, out.as<double>() + 3 * (size_t)nsys, 3);
        HIPCHK(hipMemcpyAsync(h.data(), out.p, sizeof(double) * 9 * nsys, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (!check_errors(attempt == 0)) break;
      }
    };
    auto energy = [&](int k) { return h[3 * k] + h[3 * k + 1] + h[3 * k + 2]; };
    auto stat = [&](int k, int c) { return h[3 * (size_t)nsys + 6 * k + c]; };
    auto fnorm = [&](int k) {  // the quantity compared with ftol^2 (Min::fnorm_sqr / _max / _inf)
      return fp.norm == 1 ? stat(k, 4) : fp.norm == 2 ? stat(k, 5) : stat(k, 2);
    };
    // Min setup: velocities zeroed (MinFire::setup_style), lists and forces of the start
    hipLaunchKernelGGL(k_zero_v, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, v.as<double4>());
    if (need_rebuild) rebuild();
    measure();
    for (int k = 0; k < nsys; k++) {
      fs[k].ecur = energy(k);
      st[k].e_initial = st[k].e_next_to_last = st[k].e_final = fs[k].ecur;
      st[k].fnorm2_initial = std::sqrt(stat(k, 2));
      st[k].fnorminf_initial = stat(k, 5);
      st[k].stop = LMFF_STOP_MAXITER;
    }
    std::vector<FireStep> q(nsys);
    for (int64_t iter = 0; iter < maxiter; iter++) {
      bool any = false;
      for (int k = 0; k < nsys; k++) {
        S &a = fs[k];
        FireStep &c = q[k];
        c = FireStep{};
        if (a.done) continue;
        const long long ts = ++a.ntimestep;
        st[k].iterations++;
        const double vdotf = stat(k, 0);
        if (vdotf > 0.0) {
          a.negatif = 0;
          const double vdotv = stat(k, 1), fdotf = stat(k, 2);
          c.scale1 = 1.0 - a.alpha;
          c.scale2 = fdotf <= 1e-20 ? 0.0 : a.alpha * std::sqrt(vdotv / fdotf);
          c.mix = 1;
          if (ts - a.last_negative > fp.delaystep) {
            a.dt = std::min(a.dt * fp.dtgrow, a.dtmax);
            a.alpha *= fp.alphashrink;
          }
        } else {
          a.last_negative = ts;
          const bool delayflag = !(ts - a.start < fp.delaystep && fp.delaystep_start);
          if (delayflag) {
            a.alpha = fp.alpha0;
            if (a.dt * fp.dtshrink >= a.dtmin) a.dt *= fp.dtshrink;
          }
          a.negatif++;
          if (fp.max_vdotf_negatif > 0 && a.negatif > fp.max_vdotf_negatif) {
            a.done = 1;
            st[k].stop = LMFF_STOP_MAXVDOTF;
            continue;
          }
          c.zero = 1;
          c.back = fp.halfstepback ? 0.5 * a.dtv_prev : 0.0;
        }
        // no atom moves further than dmax (velocities after the zeroing above)
        const double vmax = c.zero ? 0.0 : stat(k, 3);
        double dtv = a.dt;
        if (dtv * vmax > fp.dmax) dtv = fp.dmax / vmax;
        a.dtv_prev = dtv;
        c.dtv = dtv;
        c.dtf = dtv * ftm2v;
        c.active = 1;
        any = true;
      }
      if (!any) break;
      HIPCHK(hipMemcpyAsync(steps.p, q.data(), sizeof(FireStep) * nsys, hipMemcpyHostToDevice, stream));
      const double hm = 0.5 * ilp_margin();
      hipLaunchKernelGGL(k_fire_step, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, x.as<double4>(),
                         v.as<double4>(), (const double4 *)f.p, sysid.as<int>(), steps.as<FireStep>(),
                         rmass.as<double>(), (const double4 *)xhold.p, hm * hm, farflag.as<int>());
      ago++;
      if (rebuild_due()) rebuild();
      measure();
      for (int k = 0; k < nsys; k++) {
        S &a = fs[k];
        if (a.done || !q[k].active) continue;
        st[k].evaluations++;
        a.eprev = a.ecur;
        a.ecur = energy(k);
        st[k].e_next_to_last = a.eprev;
        st[k].e_final = a.ecur;
        if (etol > 0.0 && a.ntimestep - a.last_negative > fp.delaystep &&
            std::fabs(a.ecur - a.eprev) < etol * 0.5 * (std::fabs(a.ecur) + std::fabs(a.eprev) + eps_energy)) {
          a.done = 1;
          st[k].stop = LMFF_STOP_ETOL;
        } else if (ftol > 0.0 && fnorm(k) < ftol * ftol) {
          a.done = 1;
          st[k].stop = LMFF_STOP_FTOL;
        }
      }
    }
    for (int k = 0; k < nsys; k++) {
      st[k].fnorm2_final = std::sqrt(stat(k, 2));
      st[k].fnorminf_final = stat(k, 5);
      st[k].dt_final = fs[k].dt;
      sys[k].step = fs[k].ntimestep;
    }
    every = sv_every;
    delay = sv_delay;
    check = sv_check;
    upload_systab();
    gen++;
    forces_valid = true;
  }

  void zero_ke() { HIPCHK(hipMemsetAsync(th.as<double>() + 9 * (size_t)n, 0, n * sizeof(double), stream)); }

  // ---------------------------------------------------------------- stepping
  // advance: this step's final integration runs fused into the force gather, so the step
  // counter moves here instead of in k_advance_step
  void initial_integrate(bool advance) {
    const double h = 0.5 * ilp_margin();
    hipLaunchKernelGGL(k_initial_integrate, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, x.as<double4>(),
                       v.as<double4>(), (const double4 *)f.p, sysid.as<int>(), systab.as<SysDev>(),
                       rmass.as<double>(), (const double4 *)xhold.p, h * h, farflag.as<int>(),
                       advance ? (int)sys.size() : 0);
  }
  void final_integrate(int want_ke) {
    hipLaunchKernelGGL(k_final_integrate, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n,
                       (const double4 *)x.p, v.as<double4>(), f.as<double4>(), sysid.as<int>(), tag.as<int>(),
                       systab.as<SysDev>(), rmass.as<double>(), th.as<double>() + 9 * (size_t)n, want_ke);
    hipLaunchKernelGGL(k_advance_step, dim3(((int)sys.size() + 63) / 64), dim3(64), 0, stream, (int)sys.size(),
                       systab.as<SysDev>());
  }
  // forces, then post_force + final_integrate and the step advance, fused into the force
  // gather (with no atoms only the step advance remains)
  void force_and_final(int eflag, int vflag, int want_ke) {
    if (n == 0) {
      final_integrate(want_ke);
      return;
    }
    force(eflag, vflag, 1, want_ke);
  }
  void plain_step_body() {
    size_t mk = 0;
    mark_begin(5, mk);
    initial_integrate(n > 0);
    mark_end(5, mk);
    force_and_final(0, 0, 0);
  }

  bool rebuild_due() {
    // Neighbor::decide (neighbor.cpp:1947-1962) with ago already incremented
    if (!(ago >= delay && ago % every == 0)) return false;
    if (!check) return true;
    DBuf &flag = dflag;
    flag.ensure(sizeof(int));
    HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), stream));
    const double trig = 0.5 * skin;
    hipLaunchKernelGGL(k_displacement, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, (const double4 *)x.p,
                       (const double4 *)xhold.p, sysid.as<int>(), trig * trig, flag.as<int>());
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (h && ago == std::max(every, delay)) ndanger++;
    return h != 0;
  }

  void capture_graph(int len) {
    if (graph) {
      (void)hipGraphExecDestroy(graph);
      graph = nullptr;
    }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < len; k++) plain_step_body();
    HIPCHK(hipStreamEndCapture(stream, &g));
    HIPCHK(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    graph_len = len;
    graph_gen = gen;
    graph_prec = precision;
  }

  void evaluate(int eflag, int vflag) {
    for (int attempt = 0;; attempt++) {
      relayout();
      if (need_rebuild) rebuild();
      force(eflag, vflag);
      if (eflag || vflag) {
        if (eflag || (vflag % 4)) reduce(9, sums.as<double>());
      }
      if (!check_errors(attempt == 0)) break;  // a repeat runs the pull kernel
    }
  }

  void run(int64_t nsteps, int64_t thermo_every, int use_graph, lmff_thermo *out, int64_t out_cap) {
    relayout();
    if (n == 0) return;
    // Verlet set-up: lists if due; forces only where no valid force of the last step exists
    // (systems that kept theirs through a re-pack get it back bit for bit)
    if (need_rebuild) rebuild();
    if (!forces_valid) {
      force(0, 0);
      // FixLangevin::setup: drag and noise on the set-up force (systems that kept their last-step
      // force get it back below, unchanged)
      hipLaunchKernelGGL(k_langevin_setup, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, (int)n, (const double4 *)x.p,
                         (const double4 *)v.p, f.as<double4>(), (const int *)sysid.p, (const int *)tag.p,
                         (const SysDev *)systab.p, (const double *)rmass.p);
      for (auto &r : restore_f)
        HIPCHK(hipMemcpyAsync(f.as<double4>() + r.first, fsave.as<double4>() + r.first,
                              (size_t)(r.second - r.first) * sizeof(double4), hipMemcpyDeviceToDevice, stream));
    }
    restore_f.clear();
    const int nsys = (int)sys.size();
    int64_t nrec = 0;
    std::vector<int64_t> rec_steps;
    for (int64_t s = 1; s <= nsteps; s++)
      if ((thermo_every > 0 && s % thermo_every == 0) || s == nsteps) rec_steps.push_back(s);
    DBuf recs;
    recs.ensure(sizeof(double) * 16 * std::max<size_t>(rec_steps.size(), 1) * std::max(nsys, 1));
    std::vector<long long> base_step(nsys);
    for (int k = 0; k < nsys; k++) base_step[k] = sys[k].step;
    if (prof) {
      ev_used = 0;
      ev_marks.clear();
      for (auto &m : prof_ms) m = 0;
      for (auto &c : prof_cnt) c = 0;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    HIPCHK(hipEventRecord(t0, stream));
    int64_t step = 0;
    size_t ri = 0;
    while (step < nsteps) {
      const int64_t s = step + 1;
      const bool thermo = ri < rec_steps.size() && rec_steps[ri] == s;
      // can the next block of steps be replayed from a graph?  (no rebuild, no thermo)
      if (use_graph && !prof && !check) {
        // steps s .. s+len-1 are plain when ago+1 .. ago+len are not rebuild steps
        int len = 0;
        while (s + len <= nsteps) {
          const int ago_k = ago + len + 1;
          const bool rb = (ago_k >= delay && ago_k % every == 0);
          const bool th_k = ri < rec_steps.size() && rec_steps[ri] == s + len;
          if (rb || th_k) break;
          len++;
        }
        if (len >= 2) {
          if (!graph || graph_gen != gen || graph_prec != precision) capture_graph(1);
          for (int k = 0; k < len; k++) HIPCHK(hipGraphLaunch(graph, stream));
          ago += len;
          step += len;
          global_step += len;
          continue;
        }
      }
      size_t mk = 0;
      mark_begin(5, mk);
      initial_integrate(n > 0);
      mark_end(5, mk);
      ago++;
      if (rebuild_due()) rebuild();
      if (thermo) {
        if (peratom_last && s == nsteps) force_and_final(3, 6, 1);
        else force_and_final(1, 2, 1);
        mark_begin(6, mk);
        reduce(10, recs.as<double>() + (size_t)ri * nsys * 16);
        mark_end(6, mk);
        ri++;
      } else {
        force_and_final(0, 0, 0);
      }
      step++;
      global_step++;
    }
    HIPCHK(hipEventRecord(t1, stream));
    HIPCHK(hipStreamSynchronize(stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    prof_total = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (prof) {
      for (auto &m : ev_marks) {
        float e = 0;
        HIPCHK(hipEventElapsedTime(&e, ev_pool[m.second.first], ev_pool[m.second.second]));
        prof_ms[m.first] += e;
        prof_cnt[m.first]++;
      }
    }
    check_errors();
    sync_steps_from_device();
    // thermo records (reduce() wrote ncols=10 values per system at stride 10; copy out)
    nrec = (int64_t)rec_steps.size();
    if (out && nrec) {
      std::vector<double> h((size_t)nrec * nsys * 16);
      HIPCHK(hipMemcpy(h.data(), recs.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
      for (int k = 0; k < nsys; k++)
        for (int64_t r = 0; r < nrec; r++) {
          if ((int64_t)k * nrec + r >= out_cap) continue;
          const double *c = &h[(size_t)r * nsys * 16 + (size_t)k * 10];
          lmff_thermo &o = out[(size_t)k * nrec + r];
          o.step = base_step[k] + rec_steps[r];
          o.e_vdw = c[0];
          o.e_rep = c[1];
          o.e_tersoff = c[2];
          o.pe = c[0] + c[1] + c[2];
          for (int q = 0; q < 6; q++) o.virial[q] = c[3 + q];
          o.ke = c[9];
          const double dof = 3.0 * (double)sys[k].n - 3.0;
          o.temp = dof > 0 ? 2.0 * o.ke / (dof * BOLTZ) : 0.0;
        }
    }
  }

  // ---------------------------------------------------------------- allocation-free Pair::compute
  // One Pair::compute(eflag, vflag) on device arrays with every per-atom array carved out of this
  // GPU's `scratch`: no hipMalloc/hipFree, no allocating library call.  Each call wraps, sorts and
  // lists the atoms afresh (nothing is kept between calls; what a kernel needs zeroed is zeroed
  // here).  The ILP rows get whatever scratch is left after the other arrays.
  struct Carve {
    unsigned char *base;
    size_t cap, used;
    template <typename T>
    T *take(size_t count, const char *what) {
      const size_t off = (used + 255) & ~(size_t)255;
      const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
      if (off + bytes > cap)
        fail(LMFF_ERR_MEMORY, "the %zu MiB scratch of the allocation-free entry point is too small (%s)",
             cap >> 20, what);
      used = off + bytes;
      return (T *)(base + off);
    }
    size_t left() const { return cap - std::min(cap, (used + 255) & ~(size_t)255); }
  };

  void raise_device_errors(int e) {
    if (e & EB_NONFINITE) fail(LMFF_ERR_NONFINITE, "Non-numeric atom coords - simulation unstable");
    if (e & EB_NORMAL_COUNT)
      fail(LMFF_ERR_NORMAL, "There are too many neighbors for some atoms, please check your configuration");
    if (e & EB_NORMAL_ZERO) fail(LMFF_ERR_NORMAL, "The magnitude of the normal vector is zero");
    if (e & EB_SHORT) fail(LMFF_ERR_SHORT, "more than %d Tersoff neighbours within cutmax", MAXSH);
    if (e & EB_GRID) fail(LMFF_ERR_STATE, "bin grid too coarse for the neighbor stencil");
  }

  void pair_compute_noalloc(int64_t nlocal, const double *x_d, const int32_t *type_d, const int64_t *tag_d,
                            const int64_t *mol_d, const double lo[3], const double hi[3], const int per[3], int eflag,
                            int vflag, double *f_d, double *glob_d) {
    if (!scratch_dev) fail(LMFF_ERR_STATE, "no scratch on this device");
    if (params_dirty) fail(LMFF_ERR_STATE, "pair_style, pair_coeff and masses must be complete first");
    if ((eflag & 2) || (vflag & 4))
      fail(LMFF_ERR_ARG, "per-atom energy or virial: use lmff_pair_compute (this entry point keeps no per-atom tallies)");
    if (nlocal < 0 || nlocal >= (int64_t)IDX_MASK) fail(LMFF_ERR_ARG, "bad atom count %lld", (long long)nlocal);
    if (nlocal > 0 && (!x_d || !type_d || !tag_d || !mol_d)) fail(LMFF_ERR_ARG, "bad atom arrays");
    check_box(lo, hi, per, cutneigh);
    const int m = (int)nlocal;
    // bins from the box alone (non-periodic sides: the box bounds; atoms beyond them fall into
    // the edge bins, which reach to infinity)
    SysHost g{};
    for (int d = 0; d < 3; d++) {
      g.lo[d] = lo[d];
      g.hi[d] = hi[d];
      g.per[d] = per[d] ? 1 : 0;
    }
    setup_grid(g, nullptr, 0, cutneigh);
    SysDev sd{};
    for (int d = 0; d < 3; d++) {
      sd.lo[d] = g.lo[d];
      sd.prd[d] = g.hi[d] - g.lo[d];
      sd.per[d] = g.per[d];
      sd.nc[d] = g.nc[d];
      sd.glo[d] = g.glo[d];
      sd.inv_bs[d] = 1.0 / g.bs[d];
    }
    sd.begin = 0;
    sd.end = m;
    const int nch = std::max(1, (m + CHUNK - 1) / CHUNK);
    int ebit = 1;
    while ((1ll << ebit) <= (long long)g.ncells) ebit++;
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr, (int *)nullptr,
                                              (int *)nullptr, std::max(m, 1), 0, ebit, stream));
    const size_t M = (size_t)std::max(m, 1);
    const int scap_nf = std::max(scap, 64);
    Carve cv{scratch_dev, LMFF_SCRATCH_BYTES, 0};
    SysDev *systab_d = cv.take<SysDev>(1, "system table");
    Chunk *chunks_d = cv.take<Chunk>(nch, "chunks");
    int *cbeg_d = cv.take<int>(2, "chunks");
    double2 *partial_d = cv.take<double2>((size_t)nch * 9, "partial sums");
    double *sums_d = cv.take<double>(16, "sums");
    int *word_d = cv.take<int>(16, "status words");  // [0,4) errors, [4,8) list overflow, [8] far flag
    double4 *xa = cv.take<double4>(M, "positions"), *xb = cv.take<double4>(M, "positions");
    double4 *fa = cv.take<double4>(M, "forces");
    int *taga = cv.take<int>(M, "tags"), *tagb = cv.take<int>(M, "tags");
    int *mola = cv.take<int>(M, "molecules"), *molb = cv.take<int>(M, "molecules");
    int *imga = cv.take<int>(M, "images"), *origa = cv.take<int>(M, "order"), *origb = cv.take<int>(M, "order");
    int *sys_d = cv.take<int>(M, "system ids");
    uint32_t *key_d = cv.take<uint32_t>(M, "keys"), *key2_d = cv.take<uint32_t>(M, "keys");
    int *val_d = cv.take<int>(M, "sort values"), *perm_d = cv.take<int>(M, "sort values");
    int *cstart_d = cv.take<int>((size_t)g.ncells + 1, "cell ranges");
    int2 *rmol_d = cv.take<int2>((size_t)g.nc[1] * g.nc[2], "row molecule ranges");
    void *sorttmp_d = cv.take<unsigned char>(tb, "sort");
    DevPtrs<double> P{};
    P.n = m;
    P.x = xb;
    P.v = xa;  // no kernel of a plain evaluation touches velocities
    P.f = fa;
    P.mol = molb;
    P.tag = tagb;
    P.sys = sys_d;
    P.sys_tab = systab_d;
    P.scnt = cv.take<int>(M, "list counts");
    P.icnt = cv.take<int>(M, "list counts");
    P.ocnt = cv.take<int>(M, "list counts");
    P.scap = scap_nf;
    P.slist = cv.take<uint32_t>(M * scap_nf, "short lists");
    P.nrm = cv.take<double4>(M, "normals");
    P.dn = cv.take<double>(M * 27, "normal derivatives");
    P.dndri = cv.take<double>(M * 9, "normal derivatives");
    P.vet = cv.take<double>(M * 9, "normal neighbours");
    P.nnb = cv.take<uint32_t>(M * 4, "normal neighbours");
    P.tsh = cv.take<uint32_t>(M * MAXSH, "Tersoff neighbours");
    P.tcnt = cv.take<int>(M, "Tersoff neighbours");
    P.rtc = cv.take<uint32_t>(M * MAXSH, "reverse candidates");
    P.rtcnt = cv.take<int>(M, "reverse candidates");
    P.rnc = cv.take<uint32_t>(M * MAXSH, "reverse candidates");
    P.rncnt = cv.take<int>(M, "reverse candidates");
    P.eb = cv.take<double>(M * MAXSH, "bond energies");
    P.fb = cv.take<double>(M * MAXSH * MAXSH * 3, "bond forces");
    P.ft = cv.take<double4>(M, "Tersoff forces");
    P.fil = cv.take<double4>(M, "ILP forces");
    P.S = cv.take<double4>(M, "normal weights");
    P.th = cv.take<double>(M * NTH, "thermo columns");
    P.wb = nullptr;
    P.T = nullptr;
    P.vatom = nullptr;
    const size_t row = cv.left() / (4 * M);
    const int icap_nf = (int)std::min<size_t>(row / 64 * 64, (size_t)1 << 16);
    if (icap_nf < 64)
      fail(LMFF_ERR_MEMORY, "%lld atoms do not fit the %zu MiB scratch of the allocation-free entry point",
           (long long)nlocal, LMFF_SCRATCH_BYTES >> 20);
    P.icap = icap_nf;
    P.ilist = cv.take<uint32_t>(M * icap_nf, "ILP lists");
    P.err = word_d;
    P.ovf = word_d + 4;
    P.farflag = word_d + 8;
    scalars(P);
    const bool prof_was = prof;
    prof = false;  // no timing events: they would be created on demand
    struct Restore {
      bool &p;
      bool v;
      ~Restore() { p = v; }
    } restore{prof, prof_was};
    const int nb = (m + BS - 1) / BS;
    hipLaunchKernelGGL(k_store_sysdev, dim3(1), dim3(64), 0, stream, sd, systab_d);
    hipLaunchKernelGGL(k_one_system_chunks, dim3((nch + BS - 1) / BS), dim3(BS), 0, stream, m, CHUNK, chunks_d, nch,
                       cbeg_d);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, 16, word_d, 0);
    if (m > 0) {
      hipLaunchKernelGGL(k_fill_int, dim3(nb), dim3(BS), 0, stream, m, sys_d, 0);
      hipLaunchKernelGGL(k_hb_load, dim3(nb), dim3(BS), 0, stream, m, x_d, type_d, (const long long *)tag_d,
                         (const long long *)mol_d, xa, xb, taga, mola, origa, imga);
      hipLaunchKernelGGL(k_check_types, dim3(nb), dim3(BS), 0, stream, m, (const double4 *)xa, ntypes,
                         (const long long *)tag_d, (const long long *)mol_d, word_d + 9);
      // the list build and the force kernels index parameter tables by type: check first
      HIPCHK(hipMemcpyAsync(note_host + 9, word_d + 9, sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      if (note_host[9]) fail(LMFF_ERR_ARG, "atom types must be 1..%d, IDs 1..2^31-1, molecule IDs 32-bit", ntypes);
      hipLaunchKernelGGL(k_wrap_key, dim3(nb), dim3(BS), 0, stream, m, xa, imga, (const int *)sys_d,
                         (const SysDev *)systab_d, key_d, val_d, word_d);
      size_t tb2 = tb;
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(sorttmp_d, tb2, key_d, key2_d, val_d, perm_d, m, 0, ebit, stream));
      hipLaunchKernelGGL(k_permute<double4>, dim3(nb), dim3(BS), 0, stream, m, (const int *)perm_d,
                         (const double4 *)xa, xb);
      hipLaunchKernelGGL(k_permute<int>, dim3(nb), dim3(BS), 0, stream, m, (const int *)perm_d, (const int *)taga, tagb);
      hipLaunchKernelGGL(k_permute<int>, dim3(nb), dim3(BS), 0, stream, m, (const int *)perm_d, (const int *)mola, molb);
      hipLaunchKernelGGL(k_permute<int>, dim3(nb), dim3(BS), 0, stream, m, (const int *)perm_d, (const int *)origa,
                         origb);
      hipLaunchKernelGGL(k_cell_start, dim3((g.ncells + 1 + BS - 1) / BS), dim3(BS), 0, stream, m, g.ncells,
                         (const uint32_t *)key2_d, cstart_d);
      hipLaunchKernelGGL(k_row_mol, dim3((g.nc[1] * g.nc[2] + 3) / 4), dim3(BS), 0, stream, g.nc[1] * g.nc[2], 1,
                         (const SysDev *)systab_d, (const int *)cstart_d, (const int *)molb, rmol_d);
      hipLaunchKernelGGL(k_build_lists<double>, dim3((m + 3) / 4), dim3(BS), 0, stream, P, (const int *)cstart_d,
                         (const int2 *)rmol_d);
      // lists must be whole before any force is added into the caller's array
      HIPCHK(hipMemcpyAsync(note_host, word_d, 16 * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      raise_device_errors(note_host[0]);
      if (note_host[4] > 0)
        fail(LMFF_ERR_MEMORY, "short neighbor lists need %d entries per atom (the allocation-free entry point holds %d)",
             note_host[4], scap_nf);
      if (note_host[5] > 0)
        fail(LMFF_ERR_MEMORY, "ILP lists need %d entries per atom; %lld atoms leave room for %d in the %zu MiB scratch",
             note_host[5], (long long)nlocal, icap_nf, LMFF_SCRATCH_BYTES >> 20);
      const int vf = (vflag % 4) != 0 ? 1 : 0;
      const int ef = (eflag != 0 || vf) ? 1 : 0;
      if (precision == LMFF_MIXED) {
        DevPtrs<float> Q{};
        std::memcpy((void *)&Q, (const void *)&P, sizeof(Q));  // same layout, R-typed arrays reinterpreted
        Q.nrm = (float4 *)P.nrm;
        Q.dn = (float *)P.dn;
        Q.dndri = (float *)P.dndri;
        Q.wb = nullptr;
        scalars(Q);
        force_kernels<float>(Q, ef, vf, 0, 0, 0);
      } else {
        force_kernels<double>(P, ef, vf, 0, 0, 0);
      }
      if (f_d) hipLaunchKernelGGL(k_add_f_caller, dim3(nb), dim3(BS), 0, stream, m, (const int *)origb,
                                  (const double4 *)fa, f_d);
      if ((eflag & 1) || vf) {
        hipLaunchKernelGGL(k_reduce_chunks, dim3(nch), dim3(BS), 0, stream, (const double *)P.th, (size_t)m, 9,
                           (const Chunk *)chunks_d, partial_d, 0);
        hipLaunchKernelGGL(k_reduce_systems, dim3(1), dim3(64), 0, stream, (const double2 *)partial_d, 9,
                           (const int *)cbeg_d, 1, sums_d, 0);
      }
    }
    if (glob_d)
      hipLaunchKernelGGL(k_pair_globals, dim3(1), dim3(64), 0, stream, (const double *)sums_d, m > 0 ? eflag : 0,
                         m > 0 ? vflag : 0, glob_d);
    HIPCHK(hipMemcpyAsync(note_host, word_d, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    raise_device_errors(note_host[0]);
  }

  // ---------------------------------------------------------------- downloads
  template <typename T>
  std::vector<T> download(const DBuf &b, size_t off, size_t count) {
    std::vector<T> h(count);
    if (count) HIPCHK(hipMemcpy(h.data(), (const T *)b.p + off, count * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

int guard(lmff_ctx *c, const std::function<void()> &fn);

}  // namespace

struct lmff_ctx {
  Ctx c;
};

namespace {

int guard(lmff_ctx *ctx, const std::function<void()> &fn) {
  if (!ctx) return LMFF_ERR_ARG;
  try {
    fn();
    ctx->c.err.clear();
    return LMFF_OK;
  } catch (const LmffError &e) {
    ctx->c.err = e.msg;
    return e.code;
  } catch (const HipError &e) {
    ctx->c.err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return LMFF_ERR_HIP;
  } catch (const std::bad_alloc &) {
    ctx->c.err = "host allocation failed";
    return LMFF_ERR_MEMORY;
  }
}

void check_box(const double lo[3], const double hi[3], const int per[3], double cutneigh) {
  for (int d = 0; d < 3; d++) {
    if (!(hi[d] > lo[d])) fail(LMFF_ERR_BOX, "box dimension %d is empty", d);
    if (per[d] && hi[d] - lo[d] <= 2.0 * cutneigh)
      fail(LMFF_ERR_BOX, "periodic box side %d (%.3f) must exceed twice the neighbor cutoff (%.3f)", d, hi[d] - lo[d],
           cutneigh);
  }
}

// Bins: at least a sixth of the list cutoff along x and a third along y and z (so the list
// build's stencil has at most 7 x 7 rows; the factor keeps ceil(cut / bin) at 3 whatever the
// rounding).  A non-periodic side spans the atoms' extent, or the box bounds when x is null
// (atoms beyond the bins fall into the edge bins, which reach to infinity).
void setup_grid(SysHost &s, const double *x, int64_t n, double cutneigh) {
  double mn[3], mx[3];
  for (int d = 0; d < 3; d++) {
    mn[d] = 1e300;
    mx[d] = -1e300;
    if (!x) {
      mn[d] = s.lo[d];
      mx[d] = s.hi[d];
    } else {
      for (int64_t i = 0; i < n; i++) {
        mn[d] = std::min(mn[d], x[3 * i + d]);
        mx[d] = std::max(mx[d], x[3 * i + d]);
      }
      if (n == 0) mn[d] = mx[d] = 0.0;
    }
  }
  grid_from_extent(s, mn, mx, n, cutneigh);
}

// The grid of setup_grid from the extent [mn, mx] of the atoms along the non-periodic sides
// (periodic sides take the box).  Also what the list builds of a running system re-fit its
// non-periodic sides with, so that the edge bins of a stack that expands or ripples do not
// gather the atoms of two sheets.
void grid_from_extent(SysHost &s, const double mn_[3], const double mx_[3], int64_t n, double cutneigh) {
  // bins of at least a sixth of the list cutoff in x and z and a third in y (the list build's
  // stencil then has at most 7 rows in y and 13 in a periodic z); a non-periodic z (the normal
  // of a stack) splits the atoms' extent into bins of at most a sixth instead, so that sheets
  // 3.35 A apart fall into bins of their own (at most 15 rows: the extent over such bins)
  const double target[3] = {cutneigh / 6.0, cutneigh / 3.0 * (1.0 + 1e-6), cutneigh / 6.0 * (1.0 + 1e-6)};
  s.ncells = 1;
  for (int d = 0; d < 3; d++) {
    const double half = target[d];
    if (s.per[d]) {
      const double L = s.hi[d] - s.lo[d];
      s.nc[d] = std::max(1, (int)(L / half));
      s.glo[d] = s.lo[d];
      s.bs[d] = L / s.nc[d];
    } else {
      const double mn = mn_[d], mx = mx_[d];
      const double ext = mx - mn;
      s.glo[d] = mn;
      if (d == 2 && ext >= half) {
        s.nc[d] = (int)std::ceil(ext / half);  // bins in [half / 2, half]
        s.bs[d] = ext / s.nc[d];
      } else {
        s.nc[d] = std::max(1, (int)(ext / half));
        s.bs[d] = ext > 0 ? ext / s.nc[d] : half;
        if (s.bs[d] < half) s.bs[d] = half;
      }
    }
  }
  const int64_t cells = (int64_t)s.nc[0] * s.nc[1] * s.nc[2];
  if (cells > (int64_t)1 << 28) fail(LMFF_ERR_BOX, "box needs %lld neighbor bins (limit 2^28)", (long long)cells);
  s.ncells = (int)cells;
  // half-pair ILP tiles: 6 x 3 bins (about 18 x 18 A); the window around a tile (the stencil's
  // reach on each side, rx <= 7 bins in x and ry <= 3 rows in y) must fit its table
  const int rx = (int)std::ceil((cutneigh + 2e-6) / s.bs[0]) + 1, ry = (int)std::ceil(cutneigh / s.bs[1]) + 1;
  s.tile[0] = HALF_TILE_X;
  s.tile[1] = HALF_TILE_Y;
  if ((s.tile[0] + 2 * rx) * (s.tile[1] + 2 * ry) >= WSTR || n < half_min_atoms()) s.tile[0] = s.tile[1] = 0;
  s.ntile[0] = s.tile[0] ? (s.nc[0] + s.tile[0] - 1) / s.tile[0] : 0;
  s.ntile[1] = s.tile[1] ? (s.nc[1] + s.tile[1] - 1) / s.tile[1] : 0;
}

}  // namespace

// ======================================================================== C ABI

extern "C" {

const char *lmff_version(void) { return "lmff-mi355x 0.1 (gfx950)"; }

int lmff_create(lmff_ctx **out, int device) {
  if (!out) return LMFF_ERR_ARG;
  *out = nullptr;
  lmff_ctx *ctx = new lmff_ctx();
  int rc = guard(ctx, [&] {
    if (device < 0 || device >= MAX_DEVICES) fail(LMFF_ERR_ARG, "device %d out of range", device);
    ctx->c.device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    {
      // this device's copy of the allocation-free entry point's scratch (with the device current)
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
      }
      ctx->c.scratch_dev = (unsigned char *)g_scratch_ptr[device];
    }
    HIPCHK(hipHostMalloc((void **)&ctx->c.note_host, 16 * sizeof(int), hipHostMallocDefault));
  });
  if (rc) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return LMFF_OK;
}

void lmff_destroy(lmff_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  delete ctx;
}

const char *lmff_last_error(const lmff_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int lmff_settings(lmff_ctx *ctx, int ntypes, double ilp_cut_global, int tap_flag) {
  return guard(ctx, [&] {
    if (ntypes < 1 || ntypes > MAXT) fail(LMFF_ERR_ARG, "ntypes must be 1..%d", MAXT);
    if (!(ilp_cut_global > 0)) fail(LMFF_ERR_ARG, "Illegal pair_style command");
    auto &c = ctx->c;
    if (!c.sys.empty() && ntypes != c.ntypes) fail(LMFF_ERR_STATE, "cannot change ntypes with resident systems");
    if (!c.sys.empty() && ilp_cut_global != c.ilp_cut)
      fail(LMFF_ERR_STATE, "cannot change the ILP cutoff with resident systems");
    c.ntypes = ntypes;
    c.ilp_cut = ilp_cut_global;
    c.tap_flag = tap_flag ? 1 : 0;
    c.have_settings = true;
    c.params_changed();
  });
}

int lmff_coeff_tersoff(lmff_ctx *ctx, const char *file, const char *const *elements) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!c.have_settings) fail(LMFF_ERR_ARG, "lmff_settings must precede pair_coeff");
    if (!file || !elements) fail(LMFF_ERR_ARG, "Incorrect args for pair coefficients");
    TersoffTable t;
    std::string err;
    if (!map_elements(c.ntypes, elements, t.emap, err)) fail(LMFF_ERR_ARG, "%s", err.c_str());
    int rc = read_tersoff(file, t, err);
    if (rc) fail(rc, "%s", err.c_str());
    if (!c.sys.empty() && c.have_tersoff && t.cutmax != c.ters.cutmax)
      fail(LMFF_ERR_STATE, "cannot change the Tersoff cutoff with resident systems");
    c.ters = std::move(t);
    c.have_tersoff = true;
    c.params_changed();
  });
}

int lmff_coeff_ilp(lmff_ctx *ctx, const char *file, const char *const *elements) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!c.have_settings) fail(LMFF_ERR_ARG, "lmff_settings must precede pair_coeff");
    if (!file || !elements) fail(LMFF_ERR_ARG, "Incorrect args for pair coefficients");
    IlpTable t;
    std::string err;
    if (!map_elements(c.ntypes, elements, t.emap, err)) fail(LMFF_ERR_ARG, "%s", err.c_str());
    int rc = read_ilp(file, t, err);
    if (rc) fail(rc, "%s", err.c_str());
    if (!c.sys.empty() && c.have_ilp) {
      double r0 = 0.0, r1 = 0.0;
      for (double q : c.ilp.cutILPsq) r0 = std::max(r0, q);
      for (double q : t.cutILPsq) r1 = std::max(r1, q);
      if (r0 != r1) fail(LMFF_ERR_STATE, "cannot change the ILP normal cutoff with resident systems");
    }
    c.ilp = std::move(t);
    c.have_ilp = true;
    c.params_changed();
  });
}

int lmff_neighbor(lmff_ctx *ctx, double skin, int every, int delay, int check) {
  return guard(ctx, [&] {
    if (skin < 0 || every < 1 || delay < 0) fail(LMFF_ERR_ARG, "Illegal neighbor command");
    auto &c = ctx->c;
    if (!c.sys.empty() && skin != c.skin) fail(LMFF_ERR_STATE, "cannot change skin with resident systems");
    c.skin = skin;
    c.every = every;
    c.delay = delay;
    c.check = check ? 1 : 0;
    c.params_changed();
  });
}

int lmff_tersoff_cutmax(lmff_ctx *ctx, double *cutmax) {
  return guard(ctx, [&] {
    if (!cutmax) fail(LMFF_ERR_ARG, "null output");
    if (!ctx->c.have_tersoff) fail(LMFF_ERR_STATE, "pair_coeff for tersoff has not been given");
    *cutmax = ctx->c.ters.cutmax;
  });
}

int lmff_precision(lmff_ctx *ctx, int mode) {
  return guard(ctx, [&] {
    if (mode != LMFF_FP64 && mode != LMFF_MIXED) fail(LMFF_ERR_ARG, "precision must be LMFF_FP64 or LMFF_MIXED");
    ctx->c.precision = mode;
    ctx->c.forces_valid = false;
  });
}

int lmff_masses(lmff_ctx *ctx, const double *m) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!c.have_settings || !m) fail(LMFF_ERR_ARG, "masses need lmff_settings first");
    for (int t = 0; t < c.ntypes; t++)
      if (!(m[t] > 0)) fail(LMFF_ERR_ARG, "Invalid mass value");
    c.mass.assign(m, m + c.ntypes);
    c.params_changed();
  });
}

int lmff_system_add(lmff_ctx *ctx, int64_t natoms, const double boxlo[3], const double boxhi[3], const int periodic[3],
                    const int32_t *type, const int64_t *tag, const int64_t *mol, const double *x, const double *v,
                    int *sys_id) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    c.upload_params();
    if (natoms < 0 || (natoms > 0 && (!type || !tag || !mol || !x))) fail(LMFF_ERR_ARG, "bad system arrays");
    if (!boxlo || !boxhi || !periodic) fail(LMFF_ERR_ARG, "bad box");
    check_box(boxlo, boxhi, periodic, c.cutneigh);
    SysHost s{};
    s.id = c.next_id++;
    s.live = true;
    s.staged = true;
    s.n = natoms;
    for (int d = 0; d < 3; d++) {
      s.lo[d] = boxlo[d];
      s.hi[d] = boxhi[d];
      s.per[d] = periodic[d] ? 1 : 0;
    }
    Ctx::Staged st;
    st.id = s.id;
    st.x.assign(x, x + 3 * natoms);
    st.v.assign(3 * natoms, 0.0);
    if (v) st.v.assign(v, v + 3 * natoms);
    st.type.resize(natoms);
    st.tag.resize(natoms);
    st.mol.resize(natoms);
    st.img.assign(natoms, IMG_ZERO);
    for (int64_t i = 0; i < natoms; i++) {
      if (type[i] < 1 || type[i] > c.ntypes) fail(LMFF_ERR_ARG, "Invalid atom type %d", (int)type[i]);
      if (tag[i] <= 0 || tag[i] > INT32_MAX) fail(LMFF_ERR_ARG, "atom IDs must be in 1..2^31-1");
      if (mol[i] < INT32_MIN || mol[i] > INT32_MAX) fail(LMFF_ERR_ARG, "molecule IDs must fit in 32 bits");
      for (int d = 0; d < 3; d++)
        if (!std::isfinite(x[3 * i + d])) fail(LMFF_ERR_NONFINITE, "Non-numeric atom coords");
      st.type[i] = type[i];
      st.tag[i] = (int)tag[i];
      st.mol[i] = (int)mol[i];
    }
    setup_grid(s, x, natoms, c.cutneigh);
    c.sys.push_back(s);
    c.staged.push_back(std::move(st));
    c.layout_dirty = true;
    if (sys_id) *sys_id = s.id;
  });
}

int lmff_system_remove(lmff_ctx *ctx, int sys_id) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    // marked only: the next re-pack (relayout) drops every removed system in one pass and keeps
    // the others' resident data; until then the system tables keep their order
    s->live = false;
    c.layout_dirty = true;
    if (c.pair_sys == sys_id) c.pair_sys = -1;
  });
}

int lmff_system_integrator(lmff_ctx *ctx, int sys_id, int style, double dt, double t_target, double damp,
                           uint64_t seed) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    if (style != LMFF_NVE && style != LMFF_LANGEVIN) fail(LMFF_ERR_ARG, "unknown integrator style");
    if (!(dt > 0)) fail(LMFF_ERR_ARG, "timestep must be positive");
    if (style == LMFF_LANGEVIN && (!(damp > 0) || t_target < 0)) fail(LMFF_ERR_ARG, "Illegal fix langevin command");
    s->style = style;
    s->dt = dt;
    s->t_target = t_target;
    s->damp = damp;
    s->seed = seed;
    if (!c.layout_dirty) {
      c.sync_steps_from_device();
      c.upload_systab();
    }
    c.gen++;
  });
}

void lmff_fire_defaults(lmff_fire_params *p) {
  if (!p) return;
  // Min::Min (LMFF/min.cpp:54-71)
  p->dmax = 0.1;
  p->delaystep = 20;
  p->dtgrow = 1.1;
  p->dtshrink = 0.5;
  p->alpha0 = 0.25;
  p->alphashrink = 0.99;
  p->tmax = 10.0;
  p->tmin = 0.02;
  p->halfstepback = 1;
  p->delaystep_start = 1;
  p->max_vdotf_negatif = 2000;
  p->norm = 0;
}

int lmff_minimize(lmff_ctx *ctx, double etol, double ftol, int64_t maxiter, const lmff_fire_params *params,
                  lmff_min_stats *out, int out_cap) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (etol < 0.0 || ftol < 0.0) fail(LMFF_ERR_ARG, "Illegal minimize command");
    if (maxiter < 0) fail(LMFF_ERR_ARG, "Illegal minimize command");
    lmff_fire_params p;
    lmff_fire_defaults(&p);
    if (params) p = *params;
    if (p.tmax < p.tmin) fail(LMFF_ERR_ARG, "tmax has to be larger than tmin");
    if (p.dtgrow < 1.0) fail(LMFF_ERR_ARG, "dtgrow has to be larger than 1.0");
    if (p.dtshrink > 1.0) fail(LMFF_ERR_ARG, "dtshrink has to be smaller than 1.0");
    if (p.norm < 0 || p.norm > 2) fail(LMFF_ERR_ARG, "Illegal min_modify command");
    std::vector<lmff_min_stats> st;
    c.minimize(etol, ftol, maxiter, p, st);
    if (out)
      for (size_t k = 0; k < st.size() && (int)k < out_cap; k++) out[k] = st[k];
  });
}

int lmff_minimize_cg(lmff_ctx *ctx, double etol, double ftol, int64_t maxiter, int64_t maxeval, double dmax, int norm,
                     lmff_min_stats *out) {
  return guard(ctx, [&] {
    if (etol < 0.0 || ftol < 0.0 || maxiter < 0 || maxeval < 0) fail(LMFF_ERR_ARG, "Illegal minimize command");
    if (!(dmax > 0.0)) fail(LMFF_ERR_ARG, "Illegal min_modify command");
    if (norm < 0 || norm > 2) fail(LMFF_ERR_ARG, "Illegal min_modify command");
    lmff_min_stats st;
    ctx->c.minimize_cg(etol, ftol, maxiter, maxeval, dmax, norm, st);
    if (out) *out = st;
  });
}

int lmff_peratom_on_last_step(lmff_ctx *ctx, int enable) {
  return guard(ctx, [&] { ctx->c.peratom_last = enable != 0; });
}

int lmff_system_set_step(lmff_ctx *ctx, int sys_id, int64_t step) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    if (step < 0) fail(LMFF_ERR_ARG, "Timestep must be >= 0");
    if (!c.layout_dirty) c.sync_steps_from_device();
    s->step = step;
    if (!c.layout_dirty) c.upload_systab();
    c.gen++;
  });
}

int lmff_system_natoms(lmff_ctx *ctx, int sys_id, int64_t *natoms) {
  return guard(ctx, [&] {
    SysHost *s = ctx->c.find(sys_id);
    if (!s || !natoms) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    *natoms = s->n;
  });
}

int lmff_system_get(lmff_ctx *ctx, int sys_id, double *xo, double *vo, double *fo) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    c.relayout();
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    auto put = [&](const DBuf &b, double *o) {
      if (!o) return;
      auto h = c.download<double4>(b, s->begin, m);
      for (size_t q = 0; q < m; q++) {
        const size_t i = (size_t)org[q];
        o[3 * i] = h[q].x;
        o[3 * i + 1] = h[q].y;
        o[3 * i + 2] = h[q].z;
      }
    };
    put(c.x, xo);
    put(c.v, vo);
    put(c.f, fo);
  });
}

int lmff_system_set(lmff_ctx *ctx, int sys_id, const double *xi, const double *vi) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    c.relayout();
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    auto upd = [&](DBuf &b, const double *in) {
      if (!in) return;
      auto h = c.download<double4>(b, s->begin, m);
      for (size_t q = 0; q < m; q++) {
        const size_t i = (size_t)org[q];
        h[q].x = in[3 * i];
        h[q].y = in[3 * i + 1];
        h[q].z = in[3 * i + 2];
      }
      HIPCHK(hipMemcpy(b.as<double4>() + s->begin, h.data(), m * sizeof(double4), hipMemcpyHostToDevice));
    };
    upd(c.x, xi);
    upd(c.v, vi);
    if (xi) {
      c.need_rebuild = true;
      c.regrid_due = true;
      c.forces_valid = false;
    }
  });
}

int lmff_system_get_image(lmff_ctx *ctx, int sys_id, int32_t *image) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!image) fail(LMFF_ERR_ARG, "bad image array");
    c.relayout();
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    auto h = c.download<int>(c.img, s->begin, m);
    for (size_t q = 0; q < m; q++)
      for (int d = 0; d < 3; d++)
        image[3 * (size_t)org[q] + d] = (int32_t)(((uint32_t)h[q] >> (IMG_BITS * d)) & IMG_MASK) - IMG_MAX;
  });
}

int lmff_system_set_image(lmff_ctx *ctx, int sys_id, const int32_t *image) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!image) fail(LMFF_ERR_ARG, "bad image array");
    c.relayout();
    SysHost *s = c.find(sys_id);
    if (!s) fail(LMFF_ERR_STATE, "no system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    std::vector<int> h(m);
    for (size_t q = 0; q < m; q++) {
      uint32_t packed = 0;
      for (int d = 0; d < 3; d++) {
        const int32_t v = image[3 * (size_t)org[q] + d];
        if (v < -IMG_MAX || v >= IMG_MAX) fail(LMFF_ERR_ARG, "image flag %d out of range", (int)v);
        packed |= (uint32_t)(v + IMG_MAX) << (IMG_BITS * d);
      }
      h[q] = (int)packed;
    }
    if (m) HIPCHK(hipMemcpy(c.img.as<int>() + s->begin, h.data(), m * sizeof(int), hipMemcpyHostToDevice));
  });
}

int lmff_system_peratom(lmff_ctx *ctx, int sys_id, double *eatom, double *vatom) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s || c.layout_dirty) fail(LMFF_ERR_STATE, "no evaluated system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    if (eatom) {
      if (!(c.last_eflag & 2) && !(c.last_eflag & 1)) fail(LMFF_ERR_STATE, "last evaluation had no energy");
      auto e0 = c.download<double>(c.th, 0 * (size_t)c.n + s->begin, m);
      auto e1 = c.download<double>(c.th, 1 * (size_t)c.n + s->begin, m);
      auto e2 = c.download<double>(c.th, 2 * (size_t)c.n + s->begin, m);
      for (size_t q = 0; q < m; q++) eatom[org[q]] = e0[q] + e1[q] + e2[q];
    }
    if (vatom) {
      if (!(c.last_vflag & 4)) fail(LMFF_ERR_STATE, "last evaluation had no per-atom virial");
      for (int k = 0; k < 6; k++) {
        auto col = c.download<double>(c.vatom, (size_t)k * c.n + s->begin, m);
        for (size_t q = 0; q < m; q++) vatom[6 * (size_t)org[q] + k] = col[q];
      }
    }
  });
}

int lmff_evaluate(lmff_ctx *ctx, int eflag, int vflag) {
  return guard(ctx, [&] { ctx->c.evaluate(eflag, vflag); });
}

int lmff_system_thermo(lmff_ctx *ctx, int sys_id, lmff_thermo *out) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!out) fail(LMFF_ERR_ARG, "null output");
    int k = -1;
    for (size_t q = 0; q < c.sys.size(); q++)
      if (c.sys[q].id == sys_id && c.sys[q].live) k = (int)q;
    if (k < 0 || c.layout_dirty) fail(LMFF_ERR_STATE, "no evaluated system %d", sys_id);
    // ke of the current velocities
    const SysHost &s = c.sys[k];
    const size_t m = (size_t)(s.end - s.begin);
    auto hv = c.download<double4>(c.v, s.begin, m);
    auto hx = c.download<double4>(c.x, s.begin, m);
    double ke = 0;
    for (size_t q = 0; q < m; q++) {
      const double mass = c.mass[(int)hx[q].w - 1];
      ke += mass * (hv[q].x * hv[q].x + hv[q].y * hv[q].y + hv[q].z * hv[q].z);
    }
    ke *= 0.5 * MVV2E;
    double h[9];
    HIPCHK(hipMemcpy(h, c.sums.as<double>() + (size_t)k * 9, sizeof h, hipMemcpyDeviceToHost));
    out->step = s.step;
    out->e_vdw = h[0];
    out->e_rep = h[1];
    out->e_tersoff = h[2];
    out->pe = h[0] + h[1] + h[2];
    for (int q = 0; q < 6; q++) out->virial[q] = h[3 + q];
    out->ke = ke;
    const double dof = 3.0 * (double)s.n - 3.0;
    out->temp = dof > 0 ? 2.0 * ke / (dof * BOLTZ) : 0.0;
  });
}

int lmff_run(lmff_ctx *ctx, int64_t nsteps, int64_t thermo_every, int use_graph, lmff_thermo *out,
             int64_t out_capacity) {
  return guard(ctx, [&] {
    if (nsteps < 0) fail(LMFF_ERR_ARG, "Illegal run command");
    ctx->c.run(nsteps, thermo_every, use_graph, out, out_capacity);
  });
}

int lmff_pair_compute(lmff_ctx *ctx, int64_t nlocal, const double *x, const int32_t *type, const int64_t *tag,
                      const int64_t *mol, const double boxlo[3], const double boxhi[3], const int periodic[3],
                      int new_list, int eflag, int vflag, double *f, double *eng_vdwl, double virial[6],
                      double *eatom, double *vatom, double pvector[2]) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (nlocal < 0 || (nlocal > 0 && !x)) fail(LMFF_ERR_ARG, "bad atom arrays");
    SysHost *ps = c.pair_sys >= 0 ? c.find(c.pair_sys) : nullptr;
    const bool fresh = new_list || !ps || c.pair_n != nlocal;
    const int nb = (int)((nlocal + BS - 1) / BS);
    // a fresh list for the same number of atoms in the same box reloads the system where it
    // is (no re-pack of the context); the next evaluation sorts and lists it
    bool in_place = false;
    if (fresh && ps && !c.layout_dirty && c.pair_n == nlocal && nlocal > 0 && type && tag && mol) {
      in_place = true;
      for (int d = 0; d < 3; d++)
        if (ps->lo[d] != boxlo[d] || ps->hi[d] != boxhi[d] || ps->per[d] != (periodic[d] ? 1 : 0)) in_place = false;
      if (in_place) {
        SysHost g = *ps;
        setup_grid(g, x, nlocal, c.cutneigh);
        for (int d = 0; d < 3; d++)
          if (g.nc[d] != ps->nc[d]) in_place = false;
        if (in_place) {
          for (int64_t i = 0; i < nlocal; i++) {
            if (type[i] < 1 || type[i] > c.ntypes) fail(LMFF_ERR_ARG, "Invalid atom type %d", (int)type[i]);
            if (tag[i] <= 0 || tag[i] > INT32_MAX) fail(LMFF_ERR_ARG, "atom IDs must be in 1..2^31-1");
            if (mol[i] < INT32_MIN || mol[i] > INT32_MAX) fail(LMFF_ERR_ARG, "molecule IDs must fit in 32 bits");
            for (int d = 0; d < 3; d++)
              if (!std::isfinite(x[3 * i + d])) fail(LMFF_ERR_NONFINITE, "Non-numeric atom coords");
          }
          c.sync_steps_from_device();  // other systems' step counters live on the device
          for (int d = 0; d < 3; d++) {
            ps->glo[d] = g.glo[d];
            ps->bs[d] = g.bs[d];
          }
          c.upload_systab();
          c.gen++;
          const size_t m = (size_t)nlocal;
          c.hb_in.ensure(m * (3 * sizeof(double) + 2 * sizeof(int64_t) + sizeof(int32_t)));
          double *dx = c.hb_in.as<double>();
          int64_t *dtag = (int64_t *)(dx + 3 * m), *dmol = dtag + m;
          int32_t *dtype = (int32_t *)(dmol + m);
          HIPCHK(hipMemcpyAsync(dx, x, m * 3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
          HIPCHK(hipMemcpyAsync(dtag, tag, m * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
          HIPCHK(hipMemcpyAsync(dmol, mol, m * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
          HIPCHK(hipMemcpyAsync(dtype, type, m * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
          const int b = ps->begin;
          hipLaunchKernelGGL(k_hb_load, dim3(nb), dim3(BS), 0, c.stream, (int)m, dx, dtype, (const long long *)dtag,
                             (const long long *)dmol, c.x.as<double4>() + b, c.v.as<double4>() + b, c.tag.as<int>() + b,
                             c.mol.as<int>() + b, c.orig.as<int>() + b, c.img.as<int>() + b);
          c.need_rebuild = true;
          c.regrid_due = true;
          c.forces_valid = false;
        }
      }
    }
    if (fresh && !in_place) {
      if (ps) {
        ps->live = false;
        c.layout_dirty = true;
      }
      int id = -1;
      int rc = lmff_system_add(ctx, nlocal, boxlo, boxhi, periodic, type, tag, mol, x, nullptr, &id);
      if (rc) throw LmffError{rc, c.err};
      c.pair_sys = id;
      c.pair_n = nlocal;
      c.relayout();
    } else if (!fresh && nlocal > 0) {
      // positions only: caller order -> the system's sorted slots, on the device
      const size_t m = (size_t)nlocal;
      c.hb_in.ensure(m * 3 * sizeof(double));
      HIPCHK(hipMemcpyAsync(c.hb_in.p, x, m * 3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_hb_scatter_x, dim3(nb), dim3(BS), 0, c.stream, (int)m, c.orig.as<int>() + ps->begin,
                         c.img.as<int>() + ps->begin, ps->hi[0] - ps->lo[0], ps->hi[1] - ps->lo[1],
                         ps->hi[2] - ps->lo[2], c.hb_in.as<double>(), c.x.as<double4>() + ps->begin);
      c.far_check();
    }
    c.evaluate(eflag, vflag);  // ends with the error check, a synchronisation
    SysHost *s = c.find(c.pair_sys);
    int k = 0;
    for (size_t q = 0; q < c.sys.size(); q++)
      if (c.sys[q].id == c.pair_sys) k = (int)q;
    const size_t m = (size_t)(s->end - s->begin);
    const size_t total = (f && m > 0) ? 3 * m : 0;
    // page-locked landing area: the forces, then the nine global sums
    c.hb_host.ensure((total + 16) * sizeof(double));
    double *hpin = (double *)c.hb_host.p;
    double *h9 = hpin + total;
    const bool want_sums = eflag || (vflag % 4);
    if (want_sums)
      HIPCHK(hipMemcpyAsync(h9, c.sums.as<double>() + (size_t)k * 9, 9 * sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
    if (total) {
      // forces gathered into caller order on the device and brought down in chunks; the host adds
      // each chunk into the caller's array while the next one is in flight
      c.hb_out.ensure(total * sizeof(double));
      hipLaunchKernelGGL(k_hb_gather_f, dim3((int)((m + BS - 1) / BS)), dim3(BS), 0, c.stream, (int)m,
                         c.orig.as<int>() + s->begin, c.f.as<double4>() + s->begin, c.hb_out.as<double>());
      const int nchunk = total > ((size_t)1 << 18) ? 4 : 1;
      std::vector<hipEvent_t> done(nchunk, nullptr);
      for (int q = 0; q < nchunk; q++) {
        const size_t b = total * q / nchunk, e = total * (q + 1) / nchunk;
        HIPCHK(hipMemcpyAsync(hpin + b, c.hb_out.as<double>() + b, (e - b) * sizeof(double), hipMemcpyDeviceToHost,
                              c.stream));
        HIPCHK(hipEventCreateWithFlags(&done[q], hipEventDisableTiming));
        HIPCHK(hipEventRecord(done[q], c.stream));
      }
      for (int q = 0; q < nchunk; q++) {
        const size_t b = total * q / nchunk, e = total * (q + 1) / nchunk;
        HIPCHK(hipEventSynchronize(done[q]));
        for (size_t t = b; t < e; t++) f[t] += hpin[t];
      }
      for (auto ev : done) (void)hipEventDestroy(ev);
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    std::vector<int> org;
    if ((eatom && (eflag & 2)) || (vatom && (vflag & 4))) org = c.download<int>(c.orig, s->begin, m);
    if (!want_sums)
      for (int q = 0; q < 9; q++) h9[q] = 0.0;
    if (eng_vdwl) *eng_vdwl = (eflag & 1) ? h9[0] + h9[1] + h9[2] : 0.0;
    if (pvector) {
      pvector[0] = (eflag & 1) ? h9[0] : 0.0;
      pvector[1] = (eflag & 1) ? h9[1] : 0.0;
    }
    if (virial && (vflag % 4))
      for (int q = 0; q < 6; q++) virial[q] = h9[3 + q];
    if (eatom && (eflag & 2)) {
      auto e0 = c.download<double>(c.th, 0 * (size_t)c.n + s->begin, m);
      auto e1 = c.download<double>(c.th, 1 * (size_t)c.n + s->begin, m);
      auto e2 = c.download<double>(c.th, 2 * (size_t)c.n + s->begin, m);
      for (size_t q = 0; q < m; q++) eatom[org[q]] += e0[q] + e1[q] + e2[q];
    }
    if (vatom && (vflag & 4)) {
      for (int kk = 0; kk < 6; kk++) {
        auto col = c.download<double>(c.vatom, (size_t)kk * c.n + s->begin, m);
        for (size_t q = 0; q < m; q++) vatom[6 * (size_t)org[q] + kk] += col[q];
      }
    }
  });
}

int lmff_pair_compute_noalloc(lmff_ctx *ctx, int64_t nlocal, const double *x, const int32_t *type, const int64_t *tag,
                              const int64_t *mol, const double boxlo[3], const double boxhi[3], const int periodic[3],
                              int eflag, int vflag, double *f, double *globals) {
  return guard(ctx, [&] {
    if (!boxlo || !boxhi || !periodic) fail(LMFF_ERR_ARG, "box arguments missing");
    HIPCHK(hipSetDevice(ctx->c.device));
    ctx->c.pair_compute_noalloc(nlocal, x, type, tag, mol, boxlo, boxhi, periodic, eflag, vflag, f, globals);
  });
}

int64_t lmff_noalloc_scratch_bytes(void) { return (int64_t)LMFF_SCRATCH_BYTES; }

int lmff_neighbor_stats(lmff_ctx *ctx, int reset, int64_t *builds, int64_t *dangerous) {
  return guard(ctx, [&] {
    if (builds) *builds = ctx->c.ncalls;
    if (dangerous) *dangerous = ctx->c.ndanger;
    if (reset) ctx->c.ncalls = ctx->c.ndanger = 0;
  });
}

int lmff_profile(lmff_ctx *ctx, int enable) {
  return guard(ctx, [&] { ctx->c.prof = enable != 0; });
}

int lmff_profile_read(lmff_ctx *ctx, double *ms_total, double *ms_per_family, int64_t *launches) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (ms_total) *ms_total = c.prof_total;
    for (int k = 0; k < 7; k++) {
      if (ms_per_family) ms_per_family[k] = c.prof_cnt[k] ? c.prof_ms[k] / c.prof_cnt[k] : 0.0;
      if (launches) launches[k] = c.prof_cnt[k];
    }
  });
}

int lmff_memory(lmff_ctx *ctx, int64_t *bytes, double *per_atom) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (bytes) *bytes = (int64_t)c.bytes_in_use;
    if (per_atom) *per_atom = Ctx::atom_bytes(c.scap, std::max(c.icap, 64));
  });
}

int64_t lmff_system_bytes(lmff_ctx *ctx, int64_t natoms, int nlayers_hint) {
  if (!ctx) return -1;
  auto &c = ctx->c;
  // List rows are shared by every packed system, so a new system pays the context's current
  // capacities, or what its own stack needs if that is more.
  const int icap = std::max(c.icap, Ctx::ilp_row_estimate(c.cutneigh, std::max(nlayers_hint, 2)));
  return (int64_t)std::ceil((double)natoms * Ctx::atom_bytes(std::max(c.scap, 64), icap, std::max(nlayers_hint, 2)));
}

int lmff_debug_ilp_list(lmff_ctx *ctx, int sys_id, int64_t atom, int32_t *rows, int capacity, int *nout) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s || c.layout_dirty) fail(LMFF_ERR_STATE, "no evaluated system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    int64_t q = -1;
    for (size_t p = 0; p < m; p++)
      if (org[p] == atom) q = (int64_t)p;
    if (q < 0) fail(LMFF_ERR_ARG, "atom out of range");
    const size_t gq = (size_t)(s->begin + q);
    int nfront = 0, nback = 0;
    HIPCHK(hipMemcpy(&nfront, c.icnt.as<int>() + gq, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&nback, c.ocnt.as<int>() + gq, sizeof(int), hipMemcpyDeviceToHost));
    const int cnt = nfront + nback;
    auto ent = c.download<uint32_t>(c.ilist, gq * (size_t)c.icap, (size_t)nfront);
    auto back = c.download<uint32_t>(c.ilist, gq * (size_t)c.icap + (size_t)(c.icap - nback), (size_t)nback);
    ent.insert(ent.end(), back.begin(), back.end());
    auto orgall = c.download<int>(c.orig, 0, (size_t)c.n);
    std::vector<std::array<int32_t, 4>> r;
    for (int k = 0; k < cnt; k++) {
      const uint32_t e = ent[k];
      const int j = (int)(e & IDX_MASK);
      r.push_back({orgall[j], (int32_t)((e >> IDX_BITS) & 3) - 1, (int32_t)((e >> (IDX_BITS + 2)) & 3) - 1,
                   (int32_t)((e >> (IDX_BITS + 4)) & 3) - 1});
    }
    std::sort(r.begin(), r.end());
    for (int k = 0; k < cnt && k < capacity; k++)
      for (int d = 0; d < 4; d++) rows[4 * k + d] = r[k][d];
    *nout = cnt;
  });
}

int lmff_pair_stats(lmff_ctx *ctx, int64_t out[4]) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    if (!out) fail(LMFF_ERR_ARG, "null output");
    c.relayout();
    if (c.need_rebuild) c.rebuild();
    DBuf cnt;
    cnt.ensure(4 * sizeof(unsigned long long));
    HIPCHK(hipMemsetAsync(cnt.p, 0, 4 * sizeof(unsigned long long), c.stream));
    if (c.n) {
      auto P = c.ptrs<double>();
      hipLaunchKernelGGL(k_pair_stats<double>, dim3((c.n + BS - 1) / BS), dim3(BS), 0, c.stream, P,
                         cnt.as<unsigned long long>());
    }
    unsigned long long h[4];
    HIPCHK(hipMemcpyAsync(h, cnt.p, sizeof h, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 4; k++) out[k] = (int64_t)h[k];
  });
}

int lmff_debug_normals(lmff_ctx *ctx, int sys_id, double *normal, int32_t *count) {
  return guard(ctx, [&] {
    auto &c = ctx->c;
    SysHost *s = c.find(sys_id);
    if (!s || c.layout_dirty) fail(LMFF_ERR_STATE, "no evaluated system %d", sys_id);
    HIPCHK(hipStreamSynchronize(c.stream));
    const size_t m = (size_t)(s->end - s->begin);
    auto org = c.download<int>(c.orig, s->begin, m);
    auto nb = c.download<uint32_t>(c.nnb, (size_t)s->begin * 4, m * 4);
    std::vector<double> nv(m * 3);
    if (c.precision == LMFF_MIXED) {
      auto h = c.download<float4>(c.nrm, s->begin, m);
      for (size_t q = 0; q < m; q++) { nv[3 * q] = h[q].x; nv[3 * q + 1] = h[q].y; nv[3 * q + 2] = h[q].z; }
    } else {
      auto h = c.download<double4>(c.nrm, s->begin, m);
      for (size_t q = 0; q < m; q++) { nv[3 * q] = h[q].x; nv[3 * q + 1] = h[q].y; nv[3 * q + 2] = h[q].z; }
    }
    for (size_t q = 0; q < m; q++) {
      const size_t i = (size_t)org[q];
      if (normal) for (int d = 0; d < 3; d++) normal[3 * i + d] = nv[3 * q + d];
      if (count) count[i] = (int32_t)nb[4 * q];
    }
  });
}

}  // extern "C"
