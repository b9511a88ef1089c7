// HIP kernels of the LMFF force path for CDNA4 (gfx950, wave64).
//
// Layout in HBM (all per-atom arrays are in cell-sorted order, one packed array for every
// resident system):
//   x, v, f        double4 per atom (x.w holds the atom type as an exact double)
//   mol, tag, sys  int32 per atom; orig = index of the atom in its system's caller order
//   slist          [scap][n] uint32 (transposed): every atom within max(Tersoff cutmax,
//                  ILP rcut)+skin
//   ilist          [n][icap] uint32: other-molecule atoms within ILP cut + skin (the
//                  layer-filtered ILP list)
//   list entries   bits 0-25 atom index, bits 26-27/28-29/30-31 image shift+1 in x/y/z
//   nrm, dn, dndri normal (R4; w = |n1| before normalisation), d n / d x_k [3 nbrs][3 comps]
//                  [3 coords] (27 R, virial runs only), d n / d x_i (9 R) -- R = double (FP64
//                  mode) or float (mixed mode)
//   nnb            [n][4] uint32: count + up to 3 normal neighbours
//   tsh            [n][MAXSH] uint32: Tersoff neighbours within cutmax
//   fz, pref, eb   [n][MAXSH] double: per-bond force_zeta scalar, prefactor, energy
//   fil, S         double4: ILP pair force and S_i = sum_j w_ij del_ij
//   th             [NTH][n] double columns: e_vdw, e_rep, e_tersoff, virial share (6), ke
//
// Forces are gathered ("pulled") by the atom they act on: no kernel writes another atom's
// force, so results are deterministic and independent of launch order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace lmff {

constexpr int MAXT = 8;        // LAMMPS atom types supported on device
constexpr int MAXSH = 8;       // Tersoff neighbours within cutmax per atom
constexpr int IDX_BITS = 26;
constexpr uint32_t IDX_MASK = (1u << IDX_BITS) - 1u;
constexpr int NTH = 10;        // thermo columns
// image flags, as LAMMPS' default 32-bit imageint (LMFF/lmptype.h:98-101)
constexpr int IMG_BITS = 10;
constexpr uint32_t IMG_MASK = (1u << IMG_BITS) - 1u;
constexpr int IMG_MAX = 1 << (IMG_BITS - 1);
constexpr int IMG_ZERO = IMG_MAX | (IMG_MAX << IMG_BITS) | (IMG_MAX << (2 * IMG_BITS));
constexpr double FTM2V = 1.0 / 1.0364269e-4;  // LMFF/update.cpp:170
constexpr double MVV2E = 1.0364269e-4;
constexpr double BOLTZ = 8.617343e-5;
constexpr double MY_PI2 = 1.57079632679489661923;
constexpr double MY_PI4 = 0.78539816339744830962;

enum ErrBits : int {
  EB_NORMAL_COUNT = 1,
  EB_NORMAL_ZERO = 2,
  EB_SHORT = 4,
  EB_NONFINITE = 8,
  EB_GRID = 16,       // bin grid coarser than the list build's stencil allows
  EB_FIXED = 32,      // an ILP pair term beyond the range of the fixed-point partner sums
};

// Half-pair ILP: per ILP list row, the entries with z-bin offset dz < 0 come first, then dz = 0,
// 1, 2, ... (the build's stencil order); seg[q] counts the entries with dz <= q - 1 (so the
// entries of dz = d >= 0 are [seg[d], seg[d + 1])), for the front and back segments separately.
constexpr int NSEG = 16;
// window tables: per window (tile, partner z bin) the exclusive prefix of atoms over its bins
// (row-major over (tile[1] + 2 ry) rows of (tile[0] + 2 rx) bins), the total at [bins]
constexpr int WSTR = 256;
// window atoms the half-pair kernel can hold in LDS (six int64 sums each: 72 KB); the host keeps
// the windows it launches with below what leaves room for two workgroups per CU (about 1,400
// atoms with the kernel's static LDS), and the partner sums of slots beyond that (windows of
// rippled or crowded sheets) go to the global sums with device atomics instead
constexpr int WCAP_MAX = 1536;
// wavefronts per half-pair workgroup and the waves per SIMD the kernel is compiled for: two
// workgroups fit a CU's LDS, so 8 waves give 4 per SIMD (128 VGPRs; 10-wave workgroups for FP32,
// at 96 VGPRs, measured a third slower)
template <typename R> struct HalfWaves {
  static constexpr int n = 8;
  static constexpr int per_simd = 4;
};
// partner sums in fixed point: units of 2^-44 eV/A (forces) and of the normal weight S, added as
// exact 64-bit integers so that the order of the adds does not matter.  A term is converted by
// one fma with 1.5 * 2^52 (its rounded integer lands in the low mantissa bits) and an integer
// subtract; |term| < 2^7 keeps it exact to the unit (the sums hold 2^19).
constexpr double FX_SCALE = 17592186044416.0;     // 2^44
constexpr double FX_INV = 1.0 / 17592186044416.0;
constexpr double FX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
constexpr unsigned long long FX_MAGIC_BITS = 0x4338000000000000ull;
// a pair's partner terms are bounded by (rc + 2) max(|C|, |Tap a_ij|, |Tap a_ji|) < 2^7 when the
// three coefficients stay below this (rc <= 18 A)
constexpr double FX_COEF_MAX = 6.0;

struct SysDev {
  double lo[3], prd[3];
  int per[3];
  int nc[3];
  double glo[3], inv_bs[3];
  int cell_off;
  int row_off;        // first row of cells (fixed y and z bin) of this system
  // tiles of the half-pair ILP kernel: tile[0] x tile[1] bins in x and y, ntile[] of them; one
  // workgroup (and one window table) per tile and z bin, numbered from tile_off: (z ntile[1] +
  // ty) ntile[0] + tx.  tile[0] == 0: no tiles (the system only takes the pull kernel).
  int tile[2], ntile[2];
  int tile_off;
  int begin, end;
  double dtv, dtf;
  int style;          // 0 = NVE, 1 = Langevin
  double t_target, damp;
  unsigned long long seed;
  long long step;
};

template <typename R>
struct IlpPair {
  R z0, lambda, half_eps, C, d2inv;  // repulsion (normal of the first atom)
  R C6, d, inv_seff, C6_d_seff;      // van der Waals
  double rcutsq;                     // normal-neighbour cutoff^2 (cutILPsq)
  int mapped;                        // both types mapped to ILP elements
  int vsym;                          // vdW parameters equal for (ti,tj) and (tj,ti)
  int esym;                          // lambda, z0 equal for (ti,tj) and (tj,ti)
  int pad;
};

template <typename R>
struct TersP {
  R lam1, biga, lam2, bigb, bigr, bigd, cc, dd, h, gamma, powern, beta, c1, c2, c3, c4, lam3;
  double cutsq;
  int powermint, pad;
};

template <typename R>
struct DevPtrs {
  int n;
  double4 *x, *v, *f;
  int *mol, *tag, *sys;
  const SysDev *sys_tab;
  // lists
  uint32_t *slist; int *scnt; int scap;  // scnt: front count | back count << 16
  // ILP rows: a front segment [0, icnt) and a back segment [icap - ocnt, icap)
  uint32_t *ilist; int *icnt, *ocnt; int icap;
  // per step
  typename std::conditional<sizeof(R) == 8, double4, float4>::type *nrm;
  R *dn, *dndri;
  double *vet;        // [n][9] double, positions of normal neighbours relative to the atom
  uint32_t *nnb;      // [n][4]
  uint32_t *tsh; int *tcnt;
  uint32_t *rtc; int *rtcnt;  // [n][MAXSH] atoms that may hold i in their Tersoff list
  uint32_t *rnc; int *rncnt;  // [n][MAXSH] atoms that may hold i as a normal neighbour
  double *eb;         // [n][MAXSH] bond energy 1/2 b_ij f_A
  double *fb;         // per atom a block of MAXSH*MAXSH*3: the Tersoff force of bond (i, jj) on i's
                      // t-th neighbour at (t*ns + jj)*3, ns = tcnt[i] (see bond_slot)
  double *wb;         // the same with 6 per-atom virial shares per entry (per-atom virial only)
  double4 *ft;        // [n] Tersoff force of i's own bonds and repulsive halves
  double4 *fil, *S;
  double *T;          // [n][6] (per-atom virial only)
  double *th;         // [NTH][n]
  double *vatom;      // [6][n]
  // parameters
  const IlpPair<R> *ip;   // [MAXT][MAXT], types 0-based
  const TersP<double> *tp;  // [ntp]; Tersoff math is FP64 in both precision modes
  int ntp;
  const int *t3;          // [MAXT][MAXT][MAXT] -> tp index or -1
  int *err;
  int *ovf;               // [0] max short count, [1] max ILP count
  double cutneighsq, shortsq, tersoff_cutmaxsq, ilp_cutsq, inv_rcut, rcut;
  double ilp_splitsq;     // ILP rows hold entries with r^2 < ilp_splitsq (at build time) first
  double short_splitsq;   // short lists likewise hold entries with r^2 < short_splitsq first
  const int *farflag;     // set once an atom has moved more than half the split margin since the build
  const int *cell_start;  // [cells + 1] first sorted atom of each cell (of the last list build)
  int ilp_only;           // list build: full ILP rows only (no short lists, segments or window slots)
  int tap_flag;
  int ntypes;             // atom types 1..ntypes
  int ters_rep_sym;       // repulsive Tersoff parameters equal in both orders of every type pair
  // Tap(r) = 1 + r^4 (c4 + r (c5 + r (c6 + r c7))) and dTap/dr = d7 r^3 (r - rcut)^3:
  // {c4, c5, c6, c7, d7}
  double tap_poly[5];
  // final integration fused into the force gather (the Verlet loop's plain and thermo steps)
  const double *rmass;    // [MAXT] mass per type
  // half-pair ILP (systems with syshalf[s] != 0; all pointers null where no system has tiles)
  uint16_t *lidx;         // [n][icap] window slot of each ILP entry with dz >= 0 (parallel to ilist)
  uint16_t *segf, *segb;  // [n][NSEG] entries per z-bin offset, front and back segments
  int *win;               // [windows][WSTR] window tables
  int *winmax;            // [nsys] largest window of each system (atoms)
  const int *syshalf;     // [nsys] 1: the system's ILP pairs run in the half-pair kernel
  unsigned long long *acc;  // [6][n] partner sums (fx fy fz Sx Sy Sz) in fixed point
  double *ipart;          // [n][half_pstr] each atom's own sums, added pass by pass
  int npass;              // passes of the half-pair kernel: z-bin offsets 0 .. npass - 1
  double fx_scale, fx_magic;  // FX_SCALE and FX_MAGIC (launch arguments, see fx_of)
  int nsys;
  int half_on;            // this evaluation runs the half-pair kernel for the half systems
};

// ---------------------------------------------------------------------------------------
// math helpers

template <typename R> struct MathOps;
template <> struct MathOps<double> {
  // e^a for finite a: Cody-Waite reduction a = k ln2 + r, |r| <= ln2/2, with ln2 split into
  // a double and its remainder (0x1.abc9e3b39803fp-56 = ln2 - hi, from 60-digit arithmetic),
  // then a degree-11 polynomial for e^r fitted by least squares on the relative error at
  // Chebyshev nodes of [-ln2/2, ln2/2], and 2^k applied by ldexp (which over- and underflows
  // like exp).  Measured against 50-digit arithmetic on 6e4 arguments in [-700, 50], [-40, 0]
  // and [-2, 2]: at most 0.99 ulp.  (A degree-10 fit, at 2.9 ulp, measured no faster in the ILP
  // kernel.)  No special-case branches: five instructions fewer than the library exp, 3 %
  // off the FP64 ILP pair kernel.  The Tersoff kernels call the library exp (the inline
  // polynomial raises the register pressure of the three-wave bonds kernel).
  static __device__ __forceinline__ double ex(double a) {
    const double k = rint(a * 0x1.71547652b82fep+0);
    double r = fma(-k, 0x1.62e42fefa39efp-1, a);
    r = fma(-k, 0x1.abc9e3b39803fp-56, r);
    double p = 0x1.ad5126a516463p-26;
    p = fma(p, r, 0x1.28b2f03db05a6p-22);
    p = fma(p, r, 0x1.71df5344a07b1p-19);
    p = fma(p, r, 0x1.a01992f900b49p-16);
    p = fma(p, r, 0x1.a01a010963f3fp-13);
    p = fma(p, r, 0x1.6c16c18770609p-10);
    p = fma(p, r, 0x1.11111111325cap-7);
    p = fma(p, r, 0x1.555555554f3cbp-5);
    p = fma(p, r, 0x1.5555555555499p-3);
    p = fma(p, r, 0x1.0000000000010p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
  }
  static __device__ __forceinline__ double sq(double a) { return sqrt(a); }
  static __device__ __forceinline__ double inv(double a) { return 1.0 / a; }
  // v_rsq_f64 / v_rcp_f64 seeds refined by two Newton steps (relative error at the last
  // bits of a double, not correctly rounded)
  static __device__ __forceinline__ double rsqrt(double a) {
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    y = y * fma(-h, y * y, 1.5);
    y = y * fma(-h, y * y, 1.5);
    return y;
  }
  static __device__ __forceinline__ double rcp(double a) {
    double y = __builtin_amdgcn_rcp(a);
    double e = fma(-a, y, 1.0);
    y = fma(y, e, y);
    e = fma(-a, y, 1.0);
    y = fma(y, e, y);
    return y;
  }
  static __device__ __forceinline__ double sn(double a) { return sin(a); }
  static __device__ __forceinline__ double cs(double a) { return cos(a); }
  static __device__ __forceinline__ double pw(double a, double b) { return pow(a, b); }
};
template <> struct MathOps<float> {
  static __device__ __forceinline__ float ex(float a) { return __expf(a); }
  static __device__ __forceinline__ float sq(float a) { return __fsqrt_rn(a); }
  static __device__ __forceinline__ float inv(float a) { return __builtin_amdgcn_rcpf(a); }
  static __device__ __forceinline__ float rsqrt(float a) { return __builtin_amdgcn_rsqf(a); }
  static __device__ __forceinline__ float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
  static __device__ __forceinline__ float sn(float a) { return sinf(a); }
  static __device__ __forceinline__ float cs(float a) { return cosf(a); }
  static __device__ __forceinline__ float pw(float a, float b) { return powf(a, b); }
};

__device__ __forceinline__ int e_idx(uint32_t e) { return (int)(e & IDX_MASK); }
__device__ __forceinline__ int e_s(uint32_t e, int d) { return (int)((e >> (IDX_BITS + 2 * d)) & 3u) - 1; }
__device__ __forceinline__ uint32_t e_make(int j, int sx, int sy, int sz) {
  return (uint32_t)j | ((uint32_t)(sx + 1) << IDX_BITS) | ((uint32_t)(sy + 1) << (IDX_BITS + 2)) |
         ((uint32_t)(sz + 1) << (IDX_BITS + 4));
}

// ghost coordinate of a list entry as LAMMPS forms it: x_j + pbc*prd (atom_vec_full.cpp:402-412)
__device__ __forceinline__ void ghost(const double4 &xj, uint32_t e, const double *prd, double &gx, double &gy,
                                      double &gz) {
  gx = __dadd_rn(xj.x, (double)e_s(e, 0) * prd[0]);
  gy = __dadd_rn(xj.y, (double)e_s(e, 1) * prd[1]);
  gz = __dadd_rn(xj.z, (double)e_s(e, 2) * prd[2]);
}

// r^2 = dx*dx + dy*dy + dz*dz as the reference computes it (three rounded products, two
// rounded sums, left to right; x86-64 code without FMA): every cutoff decision uses this, so
// a pair within an ulp of a cutoff falls on the same side as in the reference.  The pair
// arithmetic after the decision may contract freely.
__device__ __forceinline__ double rsq_ref(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  const double a = dx * dx, b = dy * dy, c = dz * dz;
  return (a + b) + c;
}

// Copies a small read-only table into LDS with every thread of the block (call before any
// thread leaves; a __syncthreads must follow).  sizeof(T) is a multiple of 4.
template <typename T>
__device__ __forceinline__ void stage_table(T *dst, const T *src, int count) {
  const int words = count * (int)(sizeof(T) / 4);
  for (int k = threadIdx.x; k < words; k += blockDim.x) ((int *)dst)[k] = ((const int *)src)[k];
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8).  Atoms are
// cell-sorted, so handing each XCD one contiguous range of workgroups keeps the neighbour data
// a workgroup gathers in the L2 of the XCD that already holds it.  Bijective on [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}

// Workgroups dealt to the XCDs in chunks: within each run of 8 K workgroups, XCD x takes K
// consecutive ones, so the chunks follow the launch order (lower z bins of a stack first, the
// short workgroups of the top bins last) while each XCD still works on neighbouring tiles.
// Bijective on [0, gridDim.x) (a last partial chunk keeps the launch order).
__device__ __forceinline__ int xcd_chunk_block(int K) {
  const int b = blockIdx.x;
  const int base = b / (8 * K) * (8 * K);
  if ((int)gridDim.x - base < 8 * K) return b;
  return base + (b % 8) * K + (b - base) / 8;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Wave sums of eight values at once: each of the first three exchanges halves the values a lane
// carries (it keeps one half and adds its partner's copy of that half), the last three finish
// within groups of eight lanes.  Lane l returns the sum of v[l >> 3]; ten exchanges instead of
// the 48 of eight separate butterflies.
__device__ __forceinline__ double wsum8(const double v[8], int lane) {
  const bool u32 = (lane & 32) != 0, u16 = (lane & 16) != 0, u8 = (lane & 8) != 0;
  double a[4], h[2];
#pragma unroll
  for (int k = 0; k < 4; k++) a[k] = (u32 ? v[4 + k] : v[k]) + __shfl_xor(u32 ? v[k] : v[4 + k], 32, 64);
#pragma unroll
  for (int k = 0; k < 2; k++) h[k] = (u16 ? a[2 + k] : a[k]) + __shfl_xor(u16 ? a[k] : a[2 + k], 16, 64);
  double c = (u8 ? h[1] : h[0]) + __shfl_xor(u8 ? h[0] : h[1], 8, 64);
  c += __shfl_xor(c, 4, 64);
  c += __shfl_xor(c, 2, 64);
  c += __shfl_xor(c, 1, 64);
  return c;
}

// half selection of a full list by tag parity (pair_tersoff.cpp:150-158, ilp.cpp:488-496):
// true when the pair (i,j) is computed by j in the reference
__device__ __forceinline__ bool tag_skip(int itag, int jtag, double xi, double yi, double zi, double xj, double yj,
                                         double zj) {
  if (itag > jtag) return ((itag + jtag) % 2) == 0;
  if (itag < jtag) return ((itag + jtag) % 2) == 1;
  if (zj < zi) return true;
  if (zj == zi && yj < yi) return true;
  if (zj == zi && yj == yi && xj < xi) return true;
  return false;
}

// ---------------------------------------------------------------------------------------
// Rebuild: wrap (Domain::pbc, domain.cpp:501-560), cell key, sort, permute, cell ranges

// order-preserving integer image of a double (a larger value gives a larger integer)
__device__ __forceinline__ unsigned long long ord_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Per sheet (molecule) of atoms [begin, end): the z sum in fixed point (z 2^20, exact and
// independent of the order of the adds) and the atom count, into acc[2 r], acc[2 r + 1] for the
// sheet whose molecule id is ids[r] (ids sorted, nl <= MAX_SHEETS; the caller clears acc).
constexpr int MAX_SHEETS = 64;
__global__ __launch_bounds__(256) void k_sheet_z(int begin, int end, const double4 *x, const int *mol,
                                                 const int *ids, int nl, unsigned long long *acc) {
  __shared__ unsigned long long s_acc[2 * MAX_SHEETS];
  for (int k = threadIdx.x; k < 2 * nl; k += blockDim.x) s_acc[k] = 0ull;
  __syncthreads();
  for (int p = begin + blockIdx.x * blockDim.x + threadIdx.x; p < end; p += gridDim.x * blockDim.x) {
    const int m = mol[p];
    int lo = 0, hi = nl - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ids[mid] < m) lo = mid + 1;
      else hi = mid;
    }
    if (ids[lo] != m) continue;  // not one of the sheets (the host checked there is none)
    atomicAdd(&s_acc[2 * lo], (unsigned long long)llrint(x[p].z * 1048576.0));
    atomicAdd(&s_acc[2 * lo + 1], 1ull);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 2 * nl; k += blockDim.x)
    if (s_acc[k] != 0ull) atomicAdd(acc + k, s_acc[k]);
}

// The extent of atoms [begin, end) per dimension as ord_key images: emin[d], emax[d] (the caller
// sets them to ~0 and 0 first).  256 threads per block, grid-stride.
__global__ __launch_bounds__(256) void k_extent(int begin, int end, const double4 *x, unsigned long long *emin,
                                                unsigned long long *emax) {
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int p = begin + blockIdx.x * blockDim.x + threadIdx.x; p < end; p += gridDim.x * blockDim.x) {
    const double4 q = x[p];
    mn[0] = fmin(mn[0], q.x); mn[1] = fmin(mn[1], q.y); mn[2] = fmin(mn[2], q.z);
    mx[0] = fmax(mx[0], q.x); mx[1] = fmax(mx[1], q.y); mx[2] = fmax(mx[2], q.z);
  }
  __shared__ double s_m[4][6];
#pragma unroll
  for (int d = 0; d < 3; d++)
    for (int o = 32; o > 0; o >>= 1) {
      mn[d] = fmin(mn[d], __shfl_xor(mn[d], o, 64));
      mx[d] = fmax(mx[d], __shfl_xor(mx[d], o, 64));
    }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; d++) {
      s_m[wv][d] = mn[d];
      s_m[wv][3 + d] = mx[d];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    double v = s_m[0][c];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = c < 3 ? fmin(v, s_m[w][c]) : fmax(v, s_m[w][c]);
    if (v > -1e300 && v < 1e300) {  // a block with atoms
      if (c < 3) atomicMin(emin + c, ord_key(v));
      else atomicMax(emax + (c - 3), ord_key(v));
    }
  }
}

// Domain::pbc (LMFF/domain.cpp:501): wraps into the periodic box and counts the crossing in
// the atom's image flags, packed as LAMMPS' 32-bit imageint (10 bits per dimension, offset
// IMG_MAX; x in bits 0-9, y in 10-19, z in 20-29)
__global__ void k_wrap_key(int n, double4 *x, int *img, const int *sys, const SysDev *st, uint32_t *key, int *val,
                           int *err) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SysDev &s = st[sys[p]];
  double4 xp = x[p];
  double c[3] = {xp.x, xp.y, xp.z};
  if (!isfinite(c[0]) || !isfinite(c[1]) || !isfinite(c[2])) atomicOr(err, EB_NONFINITE);
  int cell[3];
  uint32_t im = (uint32_t)img[p];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (s.per[d]) {
      const int sh = IMG_BITS * d;
      uint32_t idim = (im >> sh) & IMG_MASK;
      double hi = s.lo[d] + s.prd[d];
      if (c[d] < s.lo[d]) {
        c[d] += s.prd[d];
        idim = (idim - 1) & IMG_MASK;
      }
      if (c[d] >= hi) {
        c[d] -= s.prd[d];
        c[d] = c[d] > s.lo[d] ? c[d] : s.lo[d];
        idim = (idim + 1) & IMG_MASK;
      }
      im = (im & ~(IMG_MASK << sh)) | (idim << sh);
    }
    double f = floor((c[d] - s.glo[d]) * s.inv_bs[d]);
    int ci = isfinite(f) ? (int)fmax(fmin(f, (double)(s.nc[d] - 1)), 0.0) : 0;
    cell[d] = ci;
  }
  xp.x = c[0]; xp.y = c[1]; xp.z = c[2];
  x[p] = xp;
  img[p] = (int)im;
  key[p] = (uint32_t)(s.cell_off + (cell[2] * s.nc[1] + cell[1]) * s.nc[0] + cell[0]);
  val[p] = p;
}

__global__ void k_fill_int(int n, int *dst, int value) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = value;
}

template <typename T>
__global__ void k_permute(int n, const int *perm, const T *src, T *dst) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[perm[q]];
}

// Host-buffer Pair::compute: orig[q] is the caller's index of the q-th atom of the system in
// the device's cell-sorted order.  Positions in caller order are scattered into place (type in
// w kept), forces gathered back into caller order, and a fresh set of atoms is loaded in caller
// order (the next list build sorts it).  The image flags of a loaded system count the wraps of
// its last list build, so a caller position that the build moved by a box length is shifted
// the same way here and stays consistent with the list's image shifts (a LAMMPS caller wraps
// before it asks for a new list and never needs it).
__global__ void k_hb_scatter_x(int m, const int *orig, const int *img, double prd0, double prd1, double prd2,
                               const double *xin, double4 *x) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t i = (size_t)orig[q];
  const uint32_t im = (uint32_t)img[q];
  const int ix = (int)(im & IMG_MASK) - IMG_MAX, iy = (int)((im >> IMG_BITS) & IMG_MASK) - IMG_MAX,
            iz = (int)(im >> (2 * IMG_BITS)) - IMG_MAX;
  double4 p = x[q];
  p.x = xin[3 * i] - ix * prd0;
  p.y = xin[3 * i + 1] - iy * prd1;
  p.z = xin[3 * i + 2] - iz * prd2;
  x[q] = p;
}

__global__ void k_hb_gather_f(int m, const int *orig, const double4 *f, double *fout) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t i = (size_t)orig[q];
  const double4 a = f[q];
  fout[3 * i] = a.x;
  fout[3 * i + 1] = a.y;
  fout[3 * i + 2] = a.z;
}

__global__ void k_hb_load(int m, const double *xin, const int *type, const long long *tag, const long long *mol,
                          double4 *x, double4 *v, int *tagd, int *mold, int *orig, int *img) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  x[q] = {xin[3 * (size_t)q], xin[3 * (size_t)q + 1], xin[3 * (size_t)q + 2], (double)type[q]};
  v[q] = {0.0, 0.0, 0.0, 0.0};
  tagd[q] = (int)tag[q];
  mold[q] = (int)mol[q];
  orig[q] = q;
  img[q] = IMG_ZERO;
}

// Forces of the system added into the caller's device array in caller order (f += semantics
// of Pair::compute), for the entry point that takes device arrays.
__global__ void k_add_f_caller(int m, const int *orig, const double4 *f, double *fout) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t i = (size_t)orig[q];
  const double4 a = f[q];
  fout[3 * i] += a.x;
  fout[3 * i + 1] += a.y;
  fout[3 * i + 2] += a.z;
}

// Atom types (in x.w), tags and molecule IDs the device layout can hold: flag != 0 otherwise.
__global__ void k_check_types(int m, const double4 *x, int ntypes, const long long *tag, const long long *mol,
                              int *flag) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const double t = x[q].w;
  const long long tg = tag[q], mo = mol[q];
  if (!(t >= 1.0 && t <= (double)ntypes) || tg <= 0 || tg > 2147483647LL || mo < -2147483648LL || mo > 2147483647LL)
    atomicOr(flag, 1);
}

// A one-system table written by a kernel from its launch argument (no host-to-device copy).
__global__ void k_store_sysdev(SysDev s, SysDev *dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = s;
}

// Pair::compute's global outputs from the per-system column sums (e_vdw, e_rep, e_tersoff,
// virial[6]): out = {eng_vdwl, virial[6], pvector[2]}, zero where the flags ask for nothing.
__global__ void k_pair_globals(const double *sums, int eflag, int vflag, double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const bool e = (eflag & 1) != 0, v = (vflag % 4) != 0;
  out[0] = e ? sums[0] + sums[1] + sums[2] : 0.0;
  for (int c = 0; c < 6; c++) out[1 + c] = v ? sums[3 + c] : 0.0;
  out[7] = e ? sums[0] : 0.0;
  out[8] = e ? sums[1] : 0.0;
}

// cell_start[c] = first sorted position with key >= c, for c in [0, ncells]: one thread per cell,
// a binary search over the sorted keys (a run of empty cells costs nothing extra)
__global__ void k_cell_start(int n, int ncells, const uint32_t *key, int *start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int)key[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  start[c] = lo;
}

// Molecule range {min, max} of the atoms of each row of cells (one system, fixed y and z bin;
// {INT_MAX, INT_MIN} when the row is empty), one wave per row.
__global__ void k_row_mol(int nrows, int nsys, const SysDev *st, const int *cell_start, const int *mol, int2 *rm) {
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  // the last system whose rows start at or before r (every system has at least one row)
  int lo_s = 0, hi_s = nsys - 1;
  while (lo_s < hi_s) {
    const int mid = (lo_s + hi_s + 1) >> 1;
    if (st[mid].row_off <= r) lo_s = mid;
    else hi_s = mid - 1;
  }
  const SysDev &s = st[lo_s];
  const int base = s.cell_off + (r - s.row_off) * s.nc[0];
  const int a0 = cell_start[base], a1 = cell_start[base + s.nc[0]];
  int lo = 0x7fffffff, hi = (int)0x80000000;
  for (int q = a0 + lane; q < a1; q += 64) {
    const int m = mol[q];
    lo = min(lo, m);
    hi = max(hi, m);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  if (lane == 0) rm[r] = make_int2(lo, hi);
}

// unwrapped bin u of a periodic dimension of nc bins -> its bin in [0, nc) and the image shift
// (u - bin) / nc; u is within a few bins of the grid, so one add or subtract does unless the
// dimension has very few bins (no integer division on the common path)
__device__ __forceinline__ int wrap_bin(int u, int nc, int &shift) {
  int w = u, sh = 0;
  if (w < 0) { w += nc; sh = -1; }
  else if (w >= nc) { w -= nc; sh = 1; }
  if (w < 0 || w >= nc) {
    w = ((u % nc) + nc) % nc;
    sh = (u - w) / nc;
  }
  shift = sh;
  return w;
}

// Reach of the list build's stencil around an atom's bin: rows in y, bins in x.  The build and
// the half-pair window tables use these same expressions, so every stencil run of an atom lies
// inside the window of its tile.
__device__ __forceinline__ int stencil_ry(const SysDev &s, double cut) { return (int)ceil(cut * s.inv_bs[1]); }
__device__ __forceinline__ int stencil_rx(const SysDev &s, double cut) { return (int)ceil((cut + 2e-6) * s.inv_bs[0]); }

// the last system whose numbered items (rows, tiles) start at or before r
__device__ __forceinline__ int sys_of_tile(const SysDev *st, int nsys, int r) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (st[mid].tile_off <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Window tables of the half-pair ILP kernel, one wavefront per window (system, partner z bin,
// tile).  A window is the rectangle of bins of one z bin that holds every stencil run of every
// atom of the tile (rx bins and ry rows beyond the tile on each side, unwrapped: a periodic
// bin reached twice appears twice, as two images).  win[w][b] = atoms in the window's bins
// before bin b (row-major), win[w][bins] = its size; winmax[s] = the largest window of s.
__global__ void k_win_table(int nwin, int nsys, const SysDev *st, const int *cell_start, double cut, int *win,
                            int *winmax) {
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nwin) return;
  const int si = sys_of_tile(st, nsys, w);
  const SysDev &s = st[si];
  if (s.tile[0] <= 0) return;
  const int loc = w - s.tile_off;
  const int tx = loc % s.ntile[0], ty = (loc / s.ntile[0]) % s.ntile[1], zb = loc / (s.ntile[0] * s.ntile[1]);
  const int rx = stencil_rx(s, cut), ry = stencil_ry(s, cut);
  const int wxb = s.tile[0] + 2 * rx, wyb = s.tile[1] + 2 * ry, nb = wxb * wyb;
  const int x0 = tx * s.tile[0] - rx, y0 = ty * s.tile[1] - ry;
  int cnt[4], sum = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int b = 4 * lane + q;
    cnt[q] = 0;
    if (b < nb) {
      const int k = b / wxb, u = b - k * wxb;
      int uy = y0 + k, ux = x0 + u, sh;
      bool ok = true;
      if (s.per[1]) uy = wrap_bin(uy, s.nc[1], sh);
      else ok = uy >= 0 && uy < s.nc[1];
      if (s.per[0]) ux = wrap_bin(ux, s.nc[0], sh);
      else ok = ok && ux >= 0 && ux < s.nc[0];
      if (ok) {
        const int c = s.cell_off + (zb * s.nc[1] + uy) * s.nc[0] + ux;
        cnt[q] = cell_start[c + 1] - cell_start[c];
      }
    }
    sum += cnt[q];
  }
  int scan = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int a = __shfl_up(scan, o, 64);
    if (lane >= o) scan += a;
  }
  int run = scan - sum;
  int *row = win + (size_t)w * WSTR;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int b = 4 * lane + q;
    if (b < nb) row[b] = run;
    run += cnt[q];
  }
  const int total = __shfl(scan, 63, 64);
  if (lane == 0) {
    row[nb] = total;
    atomicMax(winmax + si, total);
  }
}

// Which systems run their ILP pairs in the half-pair kernel (one thread per system), decided
// before the list build from each system's own geometry: tiles, atoms in some window and fewer
// than 0xffff in every one, the stencil within NSEG - 2 z bins (a window larger than the
// kernel's LDS holds is allowed: the slots beyond it add to the global sums directly).  allow = 0 keeps every system in the pull
// kernel.
__global__ void k_half_decide(int nsys, const SysDev *st, const int *winmax, double cut, int allow, int *syshalf) {
  const int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsys) return;
  const SysDev &s = st[si];
  const int rz = (int)ceil(cut * s.inv_bs[2]);
  // window slots are 16-bit in the lists (0xffff: none)
  syshalf[si] = (allow && s.tile[0] > 0 && s.end > s.begin && winmax[si] > 0 && winmax[si] < 0xffff &&
                 rz + 1 <= NSEG - 2) ? 1 : 0;
}

// ---------------------------------------------------------------------------------------
// Neighbor build (npair_full_bin_ghost.cpp:35-155 semantics: every image of every atom with
// rsq <= (cut+skin)^2, evaluated in the reference's operation order with no fused
// multiply-adds, 97-102).  One wavefront per atom.  Bins are at least a third of the list
// cutoff in y (at most 7 rows) and at least a sixth in x and z (in a non-periodic z the atoms'
// extent is split into bins of at most a sixth, so that the sheets of a stack fall into bins of
// their own).  Lane l takes stencil rows l and l + 64 of the (dy, dz) rows the grid holds (at
// most 7 x 15): it drops a row whose nearest point is beyond the cutoff, else keeps the x bins
// the cutoff sphere reaches on it, only those within the short-list reach
// when the whole row of bins holds only atoms of i's own molecule (no entry of the ILP list
// can come from it), and keeps the rest as at most three runs of consecutive sorted atoms (one per periodic image in x).
// The wave lays the runs end to end in LDS and then tests the candidates 64 at a time across
// run boundaries, so no lane idles at the end of a short row.  Bins in a non-periodic
// dimension hold every atom beyond the grid in the edge bin, so edge rows and bins reach to
// infinity.

constexpr int RUNS_PER_ROW = 3;
constexpr int ROWS_PER_LANE = 2;
constexpr int RUN_SLOTS = 64 * RUNS_PER_ROW * ROWS_PER_LANE;

// nearest distance from coordinate p to the extent of unwrapped bin u of one dimension
__device__ __forceinline__ double bin_gap(double p, int u, int nc, bool per, double glo, double bs) {
  const double lo = (!per && u == 0) ? -1e300 : glo + (double)u * bs;
  const double hi = (!per && u == nc - 1) ? 1e300 : glo + (double)(u + 1) * bs;
  return fmax(0.0, fmax(lo - p, p - hi));
}

template <typename R>
__global__ __launch_bounds__(256) void k_build_lists(DevPtrs<R> P, const int *cell_start, const int2 *row_mol) {
  // run k of the wave: its first sorted atom with the image shift of its bins in the shift bits
  // of a list entry (so a candidate's entry is that word plus its position in the run), and the
  // offset of its first candidate in the wave's candidate sequence
  __shared__ uint32_t run_word[4][RUN_SLOTS];
  __shared__ int run_off[4][RUN_SLOTS + 1];
  // half-pair tables: the window slot of a run's first atom (runs with dz > 0), the run's z-bin
  // offset dz, and the wave's ILP entry counts per dz (front, back)
  __shared__ uint32_t run_loc[4][RUN_SLOTS];  // window slot | (dz + 32) << 16
  __shared__ int seg_cnt[4][2][64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = __builtin_amdgcn_readfirstlane(xcd_block() * (blockDim.x >> 6) + wv);
  const int lane = threadIdx.x & 63;
  if (i >= P.n) return;
  const SysDev &s = P.sys_tab[P.sys[i]];
  const double4 xi = P.x[i];
  const int mi = P.mol[i];
  const int ti = (int)xi.w - 1;
  // bit tj: the ILP acts between types ti and tj (wave-uniform)
  unsigned ilp_mask = 0;
  for (int t = 0; t < MAXT; t++) ilp_mask |= (P.ip[ti * MAXT + t].mapped ? 1u : 0u) << t;
  const double pos[3] = {xi.x, xi.y, xi.z};
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double f = floor((pos[d] - s.glo[d]) * s.inv_bs[d]);
    c[d] = (int)fmax(fmin(f, (double)(s.nc[d] - 1)), 0.0);
  }
  // geometric slack for the bin arithmetic (coordinates are at most ~1e4 A, so their
  // rounding is ~1e-12 A); pruning with it never drops an atom the test below accepts
  const double slack = 1e-6;
  const double cut = sqrt(P.cutneighsq);
  const double bsy = 1.0 / s.inv_bs[1], bsz = 1.0 / s.inv_bs[2];
  const int ry = (int)ceil(cut * s.inv_bs[1]), rz = (int)ceil(cut * s.inv_bs[2]);
  // the stencil rows the grid holds: in a non-periodic dimension only the bins that exist
  const int dylo = s.per[1] ? -ry : max(-ry, -c[1]), dyhi = s.per[1] ? ry : min(ry, s.nc[1] - 1 - c[1]);
  const int dzlo = s.per[2] ? -rz : max(-rz, -c[2]), dzhi = s.per[2] ? rz : min(rz, s.nc[2] - 1 - c[2]);
  const int wy = dyhi - dylo + 1, nrows = wy * (dzhi - dzlo + 1);
  if (nrows > 64 * ROWS_PER_LANE) {
    if (lane == 0) atomicOr(P.err, EB_GRID);
    return;
  }
  // the tile of i and the origin of its windows (unwrapped bins), for a system in the half-pair
  // kernel (the pull kernel reads neither window slots nor segments)
  const bool tiled = P.win != nullptr && s.tile[0] > 0 && !P.ilp_only && P.syshalf[P.sys[i]] != 0;
  // half rows: a system in the half-pair kernel keeps only the ILP entries with dz > 0, and with
  // dz = 0 those of a higher index (the partner in a lower z bin, or of the lower index, computes
  // the rest); the rows below i's z bin are scanned only for the short list
  const bool halfrows = tiled;
  const int rxs = stencil_rx(s, cut);
  int win0 = 0, wx0 = 0, wy0 = 0, wxb = 1;
  if (tiled) {
    const int tx = c[0] / s.tile[0], ty = c[1] / s.tile[1];
    win0 = s.tile_off + ty * s.ntile[0] + tx;  // + z bin x ntile[0] ntile[1]
    wx0 = tx * s.tile[0] - rxs;
    wy0 = ty * s.tile[1] - ry;
    wxb = s.tile[0] + 2 * rxs;
    seg_cnt[wv][0][lane] = 0;
    seg_cnt[wv][1][lane] = 0;
  }
  // ---- lane l: stencil rows l, l + 64 -> up to three runs each, laid end to end ----
  int total = 0, nruns = 0;  // candidates and runs laid out so far (wave-uniform)
  const float inv_wy = 1.0f / (float)wy;
  for (int it = 0; it < nrows; it += 64) {
    const int row = it + lane;
    int nrun = 0, rb[RUNS_PER_ROW], rc_[RUNS_PER_ROW], rl[RUNS_PER_ROW];
    uint32_t rs[RUNS_PER_ROW];
    int rdz = 0;
    if (row < nrows) {
      // row = qz * wy + qy (row < 128, wy <= 13: the float quotient is exact after truncation)
      const int qz = (int)(((float)row + 0.5f) * inv_wy), qy = row - qz * wy;
      int uy = c[1] + dylo + qy, uz = c[2] + dzlo + qz;
      const int uy_unw = uy;
      rdz = dzlo + qz;
      const double gy = fmax(0.0, bin_gap(pos[1], uy, s.nc[1], s.per[1], s.glo[1], bsy) - slack);
      const double gz = fmax(0.0, bin_gap(pos[2], uz, s.nc[2], s.per[2], s.glo[2], bsz) - slack);
      const double d2 = gy * gy + gz * gz;
      if (d2 <= P.cutneighsq) {
        int sy = 0, sz = 0;
        if (s.per[1]) uy = wrap_bin(uy, s.nc[1], sy);
        if (s.per[2]) uz = wrap_bin(uz, s.nc[2], sz);
        const double w = sqrt(P.cutneighsq - d2) + slack;
        // bins in reach along x; in a periodic dimension at most nc + 2 of them (w < L/2), so
        // at most three images; in a non-periodic one the edge bins hold what lies beyond
        const int rx = rxs;
        int xlo = (int)fmax(floor((pos[0] - w - s.glo[0]) * s.inv_bs[0]), (double)(c[0] - rx));
        int xhi = (int)fmin(floor((pos[0] + w - s.glo[0]) * s.inv_bs[0]), (double)(c[0] + rx));
        const int2 rm = row_mol[s.row_off + uz * s.nc[1] + uy];
        if (rm.x > rm.y) {
          xhi = xlo - 1;  // no atoms in this row of bins
        } else if ((rm.x == mi && rm.y == mi) || (halfrows && rdz < 0)) {
          // only atoms of i's own molecule (or, in half rows, a row below i's z bin): just the
          // short lists can take entries from this row
          if (d2 <= P.shortsq) {
            const double ws = sqrt(P.shortsq - d2) + slack;
            xlo = max(xlo, (int)floor((pos[0] - ws - s.glo[0]) * s.inv_bs[0]));
            xhi = min(xhi, (int)floor((pos[0] + ws - s.glo[0]) * s.inv_bs[0]));
          } else {
            xhi = xlo - 1;
          }
        }
        if (!s.per[0] && xlo <= xhi) {
          xlo = min(max(xlo, 0), s.nc[0] - 1);
          xhi = min(max(xhi, 0), s.nc[0] - 1);
        }
        const int rowbase = s.cell_off + (uz * s.nc[1] + uy) * s.nc[0];
        // window slots of this row (dz >= 0: the half-pair kernel computes the pair from i's side
        // -- for dz = 0 only when j > i -- and its partner terms land in the window of (i's tile,
        // uz))
        const int *wrow = nullptr;
        if (tiled && rdz >= 0)
          wrow = P.win + (size_t)(win0 + uz * s.ntile[0] * s.ntile[1]) * WSTR + (uy_unw - wy0) * wxb;
        int u = xlo;
        while (u <= xhi && nrun < RUNS_PER_ROW) {
          int sx = 0, wx = u, uend = xhi;
          if (s.per[0]) {
            wx = wrap_bin(u, s.nc[0], sx);
            uend = min(xhi, u + (s.nc[0] - 1 - wx));
          }
          const int a0 = cell_start[rowbase + wx], a1 = cell_start[rowbase + wx + (uend - u) + 1];
          rb[nrun] = a0;
          rc_[nrun] = a1 - a0;
          rl[nrun] = wrow ? wrow[u - wx0] : 0;
          rs[nrun] = (uint32_t)((sx + 1) | ((sy + 1) << 2) | ((sz + 1) << 4)) << IDX_BITS;
          nrun++;
          u = uend + 1;
        }
      }
    }
    // wave prefix sums of candidate and run counts place this pass's runs after the earlier ones
    int mycnt = 0, mynr = 0;
    for (int k = 0; k < nrun; k++)
      if (rc_[k] > 0) { mycnt += rc_[k]; mynr++; }
    int scnt_ = mycnt, snr = mynr;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int a = __shfl_up(scnt_, o, 64), b = __shfl_up(snr, o, 64);
      if (lane >= o) { scnt_ += a; snr += b; }
    }
    int off = total + scnt_ - mycnt, slot = nruns + snr - mynr;
    for (int k = 0; k < nrun; k++) {
      if (rc_[k] <= 0) continue;
      run_word[wv][slot] = (uint32_t)rb[k] | rs[k];
      run_off[wv][slot] = off;
      run_loc[wv][slot] = ((uint32_t)rl[k] & 0xffffu) | ((uint32_t)(rdz + 32) << 16);
      off += rc_[k];
      slot++;
    }
    // wave-uniform, so the candidate loop below is a scalar loop
    total += __builtin_amdgcn_readfirstlane(__shfl(scnt_, 63, 64));
    nruns += __builtin_amdgcn_readfirstlane(__shfl(snr, 63, 64));
  }
  if (lane == 0) run_off[wv][nruns] = total;
  // the table is read by other lanes of this wave only
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  // ILP rows are partitioned: entries within the split radius at build time are appended from
  // the front, the rest from the back.  The per-step pair kernel walks the front segment and
  // then the back one in chunks of 64, so the chunks past the split radius have (almost) no
  // pair in range and skip the pair arithmetic as a whole.
  int ns = 0, nsb = 0, ni = 0, no = 0;
  uint32_t *irow = P.ilist + (size_t)i * P.icap;
  // box lengths in registers (the list stores below could alias the system table as far as
  // the compiler knows, which would reload them for every chunk)
  const double prd0 = s.prd[0], prd1 = s.prd[1], prd2 = s.prd[2];
  // every type i can meet has the ILP with ti: the per-candidate type test is moot
  const bool every_tj = (ilp_mask | ~((1u << P.ntypes) - 1u)) == 0xffffffffu;
  int r = 0;
  // entry counts of the current z-bin offset (wave-uniform)
  int cur_d = -64, cur_f = 0, cur_b = 0;
  for (int v0 = 0; v0 < total; v0 += 64) {
    // lanes past the end repeat the last candidate and take no entry, so the candidate test
    // below runs without a divergent branch
    const bool live = v0 + lane < total;
    const int v = live ? v0 + lane : total - 1;
    while (run_off[wv][r + 1] <= v) r++;
    const uint32_t ent = run_word[wv][r] + (uint32_t)(v - run_off[wv][r]);
    const int j = e_idx(ent);
    const double4 xj = P.x[j];
    const int mj = P.mol[j];  // issued with the position, not after the distance test
    const double gx = __dadd_rn(xj.x, (double)e_s(ent, 0) * prd0);
    const double gy = __dadd_rn(xj.y, (double)e_s(ent, 1) * prd1);
    const double gz = __dadd_rn(xj.z, (double)e_s(ent, 2) * prd2);
    const double delx = __dsub_rn(xi.x, gx), dely = __dsub_rn(xi.y, gy), delz = __dsub_rn(xi.z, gz);
    const double rsq = rsq_ref(delx, dely, delz);
    const bool take = live && !((ent >> IDX_BITS) == 0x15u && j == i);
    bool type_ok = true;
    if (!every_tj) type_ok = ((ilp_mask >> ((int)xj.w - 1)) & 1u) != 0;
    uint32_t rlw = 0;
    int dzv = 0;
    if (tiled) {
      rlw = run_loc[wv][r];
      dzv = (int)(rlw >> 16) - 32;
    }
    // half rows: only partners in z bins at or above i's, and within i's own z bin only those
    // of a higher index (the pair is the lower index's)
    const bool ilp = take && (rsq <= P.cutneighsq) && (mj != mi) && type_ok &&
                     !(halfrows && (dzv < 0 || (dzv == 0 && j < i)));
    const bool near_i = rsq < P.ilp_splitsq;
    const bool in_i = ilp && near_i, in_o = ilp && !near_i;
    const bool sht = take && (rsq <= P.shortsq) && !P.ilp_only;
    const bool near_s = rsq < P.short_splitsq;
    const bool in_s = sht && near_s, in_sb = sht && !near_s;
    const unsigned long long bi = __builtin_amdgcn_ballot_w64(in_i), bo = __builtin_amdgcn_ballot_w64(in_o),
                             bs = __builtin_amdgcn_ballot_w64(in_s), bsb = __builtin_amdgcn_ballot_w64(in_sb);
    // set bits of each mask below this lane (v_mbcnt)
    auto below = [](unsigned long long m) {
      return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    };
    const int pi = below(bi), po = below(bo), ps = below(bs), psb = below(bsb);
    if (in_i && ni + pi < P.icap) irow[ni + pi] = ent;
    if (in_o && no + po < P.icap) irow[P.icap - 1 - (no + po)] = ent;
    if (tiled) {
      // the entry's window slot (dz > 0) and the per-dz entry counts; dz does not decrease along
      // the candidate sequence, so a chunk spans one or two values
      const uint16_t loc = dzv >= 0 ? (uint16_t)(rlw + (uint32_t)(v - run_off[wv][r])) : (uint16_t)0xffffu;
      uint16_t *lrow = P.lidx + (size_t)i * P.icap;
      if (in_i && ni + pi < P.icap) lrow[ni + pi] = loc;
      if (in_o && no + po < P.icap) lrow[P.icap - 1 - (no + po)] = loc;
      const int d_lo = __builtin_amdgcn_readfirstlane(dzv);
      const int d_hi = __builtin_amdgcn_readlane(dzv, 63);
      if (d_lo == cur_d && d_hi == cur_d) {
        cur_f += (int)__popcll(bi);
        cur_b += (int)__popcll(bo);
      } else {
        for (int d = d_lo; d <= d_hi; d++) {
          const unsigned long long m = __builtin_amdgcn_ballot_w64(dzv == d);
          if (d != cur_d) {
            // dz never decreases, so the finished value's counts are final
            if (cur_d > -64 && lane == 0) {
              seg_cnt[wv][0][cur_d + 32] = cur_f;
              seg_cnt[wv][1][cur_d + 32] = cur_b;
            }
            cur_d = d;
            cur_f = 0;
            cur_b = 0;
          }
          cur_f += (int)__popcll(bi & m);
          cur_b += (int)__popcll(bo & m);
        }
      }
    }
    // short list is stored transposed ([slot][atom]) so the per-step thread-per-atom
    // readers coalesce; it is split like the ILP rows (front from slot 0, back from scap-1)
    if (in_s && ns + ps < P.scap) P.slist[(size_t)(ns + ps) * P.n + i] = ent;
    if (in_sb && nsb + psb < P.scap) P.slist[(size_t)(P.scap - 1 - (nsb + psb)) * P.n + i] = ent;
    ni += __popcll(bi);
    no += __popcll(bo);
    ns += __popcll(bs);
    nsb += __popcll(bsb);
  }
  if (tiled) {
    if (cur_d > -64 && lane == 0) {
      seg_cnt[wv][0][cur_d + 32] = cur_f;
      seg_cnt[wv][1][cur_d + 32] = cur_b;
    }
    // seg[q] = entries with dz <= q - 1: lane 31 + q after an inclusive scan over dz = lane - 32
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int cf = seg_cnt[wv][0][lane], cb = seg_cnt[wv][1][lane];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int a = __shfl_up(cf, o, 64), b = __shfl_up(cb, o, 64);
      if (lane >= o) { cf += a; cb += b; }
    }
    const int q = lane - 31;
    if (q >= 0 && q < NSEG) {
      P.segf[(size_t)i * NSEG + q] = (uint16_t)min(cf, 0xffff);
      P.segb[(size_t)i * NSEG + q] = (uint16_t)min(cb, 0xffff);
    }
  }
  if (lane == 0) {
    const int front = ni < P.icap ? ni : P.icap;
    P.icnt[i] = front;
    P.ocnt[i] = no < P.icap - front ? no : P.icap - front;
    ni += no;
    const int sfront = ns < P.scap ? ns : P.scap;
    const int sback = nsb < P.scap - sfront ? nsb : P.scap - sfront;
    if (!P.ilp_only) P.scnt[i] = sfront | (sback << 16);
    ns += nsb;
    // only an overflowing row reports its size (a per-wave atomic on one address would
    // serialize every wave of the launch)
    if (ni > P.icap) atomicMax(&P.ovf[1], ni);
    if (ns > P.scap) atomicMax(&P.ovf[0], ns);
  }
}

// slot of the k-th entry of an ILP row: front segment first, then the back segment
// (back_off = icap - ocnt - icnt)
__device__ __forceinline__ int ilp_slot(int k, int nfront, int back_off) { return k < nfront ? k : k + back_off; }

// ---------------------------------------------------------------------------------------
// Per step 1: Tersoff neighbours within cutmax, ILP normal neighbours and the normal with its
// derivatives (ILP_neigh + calc_normal, pair_ilp_graphene_hbn.cpp:681-1075).  One thread/atom.

// The unnormalised normal n1 of an atom from its normal-neighbour vectors vet[k] = x_k - x_i
// (calc_normal, ilp.cpp:835-1067): nn = 2, n1 = v1 x v2 (it depends on x_i); nn = 3,
// n1 = (v1 x v2 + v3 x v1 + v2 x v3) / 3 (it does not).  Also |n1|^2 and |n1|, both set to 1
// when n1 = 0 (the return value is then true).
template <typename R>
__device__ __forceinline__ bool normal_raw(const R vet[3][3], int nn, R n1[3], R &nn2, R &len) {
  using M = MathOps<R>;
  if (nn == 2) {
    n1[0] = vet[0][1] * vet[1][2] - vet[1][1] * vet[0][2];
    n1[1] = vet[0][2] * vet[1][0] - vet[1][2] * vet[0][0];
    n1[2] = vet[0][0] * vet[1][1] - vet[1][0] * vet[0][1];
  } else {
    R pv12[3], pv31[3], pv23[3];
    pv12[0] = vet[0][1] * vet[1][2] - vet[1][1] * vet[0][2];
    pv12[1] = vet[0][2] * vet[1][0] - vet[1][2] * vet[0][0];
    pv12[2] = vet[0][0] * vet[1][1] - vet[1][0] * vet[0][1];
    pv31[0] = vet[2][1] * vet[0][2] - vet[0][1] * vet[2][2];
    pv31[1] = vet[2][2] * vet[0][0] - vet[0][2] * vet[2][0];
    pv31[2] = vet[2][0] * vet[0][1] - vet[0][0] * vet[2][1];
    pv23[0] = vet[1][1] * vet[2][2] - vet[2][1] * vet[1][2];
    pv23[1] = vet[1][2] * vet[2][0] - vet[2][2] * vet[1][0];
    pv23[2] = vet[1][0] * vet[2][1] - vet[2][0] * vet[1][1];
    for (int a = 0; a < 3; a++) n1[a] = (pv12[a] + pv31[a] + pv23[a]) / (R)3;
  }
  nn2 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
  len = M::sq(nn2);
  if (len == (R)0) {
    len = 1;
    nn2 = 1;
    return true;
  }
  return false;
}

// d n1 / d x_m for neighbour slot m is the cross-product derivative [w]x =
// [[0, w2, -w1], [-w2, 0, w0], [w1, -w0, 0]] with w = (v_{m+1} - v_{m+2}) / 3 (cyclic) for three
// neighbours and w = v2 (m = 0), -v1 (m = 1) or 0 (m = 2) for two.  Reads only the rows of vet
// other than m.
template <typename R>
__device__ __forceinline__ void normal_dslot_w(const R vet[3][3], int nn, int m, R w[3]) {
  if (nn == 2) {
    for (int d = 0; d < 3; d++) w[d] = m == 0 ? vet[1][d] : m == 1 ? -vet[0][d] : (R)0;
  } else {
    const R third = (R)1 / (R)3;
    const int p = m == 0 ? 1 : m == 1 ? 2 : 0, q = m == 0 ? 2 : m == 1 ? 0 : 1;
    for (int d = 0; d < 3; d++) w[d] = (vet[p][d] - vet[q][d]) * third;
  }
}

// d n / d x_m of the normal n = n1 / |n1| for neighbour slot m (calc_normal's dnormal):
// D[a][b] = d n[a] / d x_m[b] = dn1[a][b] / |n1| - n1[a] dnn[b] / |n1|^2 with
// dnn[b] = n1 . dn1[:, b] / |n1| and dn1 = [w]x.  Stored by the normals kernel for the virial
// tallies only; the force gather uses the contracted form -(d n / d x_m)^T S = -w x P / |n1|,
// P = S - (n . S) n.
template <typename R>
__device__ __forceinline__ void normal_dslot(const R vet[3][3], int nn, int m, const R n1[3], R nn2, R len,
                                             R D[3][3]) {
  R w[3];
  normal_dslot_w<R>(vet, nn, m, w);
  const R dn1[3][3] = {{0, w[2], -w[1]}, {-w[2], 0, w[0]}, {w[1], -w[0], 0}};
  for (int b = 0; b < 3; b++) {
    const R dnn = (n1[0] * dn1[0][b] + n1[1] * dn1[1][b] + n1[2] * dn1[2][b]) / len;
    for (int a = 0; a < 3; a++) D[a][b] = dn1[a][b] / len - n1[a] * dnn / nn2;
  }
}

// short-list entries fetched together per thread in the normals kernel
#ifndef LMFF_NORMALS_BATCH
#define LMFF_NORMALS_BATCH 8
#endif

template <typename R, bool WANT_VET>
__global__ __launch_bounds__(256) void k_normals(DevPtrs<R> P) {
  // per type pair (ti, tj): Tersoff-mapped, ILP-mapped, cutILPsq and the widened reverse cutoff,
  // built once per block in LDS (read per short-list entry)
  struct NPair {
    double rcutsq, rc2w;
    int tmapped, imapped;
  };
  __shared__ NPair s_np[MAXT * MAXT];
  const double widen = 1.0 + 1e-12;
  if (threadIdx.x < MAXT * MAXT) {
    const int a = threadIdx.x / MAXT, b = threadIdx.x % MAXT;
    const IlpPair<R> &q = P.ip[threadIdx.x];
    s_np[threadIdx.x] = {q.rcutsq, fmax(q.rcutsq, P.ip[b * MAXT + a].rcutsq) * widen,
                         P.t3[(a * MAXT + b) * MAXT + b] >= 0 ? 1 : 0, q.mapped};
  }
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SysDev &s = P.sys_tab[P.sys[i]];
  const double4 xi = P.x[i];
  const int ti = (int)xi.w - 1;
  const int mi = P.mol[i];
  // the back segment (r beyond the split at the build) can hold a neighbour within reach
  // only after some atom has moved more than half the split margin since the build
  const int sc = P.scnt[i];
  const int sfront = sc & 0xffff, cnt = sfront + (*P.farflag ? (sc >> 16) : 0);
  const int sback_off = P.scap - 1 + sfront;  // slot of entry k >= sfront: sback_off - k
  const uint32_t *row = P.slist + i;  // transposed: entry k at row[k * n]
  const size_t sstride = (size_t)P.n;
  int ns = 0, nn = 0, nrt = 0, nrn = 0;
  uint32_t nb0 = 0, nb1 = 0, nb2 = 0;
  double vetd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  bool too_many = false;
  // reverse candidates: every atom j that may list i as a Tersoff neighbour or as a normal
  // neighbour (j's own test runs from j's side, so the cutoffs are widened by 1e-12 to form a
  // superset whatever the last-bit rounding of j's r^2)
  uint32_t *rt_row = P.rtc + (size_t)i * MAXSH;
  uint32_t *rn_row = P.rnc + (size_t)i * MAXSH;
  // entries are fetched in batches of NB so that NB independent gathers are in flight; every
  // index below is a compile-time constant after unrolling (no scratch memory)
  constexpr int NB = LMFF_NORMALS_BATCH;
  for (int k0 = 0; k0 < cnt; k0 += NB) {
    uint32_t eb_[NB];
    double4 xb_[NB];
    int mb_[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int k = k0 + b;
      eb_[b] = (k < cnt) ? row[(size_t)(k < sfront ? k : sback_off - k) * sstride] : 0u;
    }
#pragma unroll
    for (int b = 0; b < NB; b++)
      if (k0 + b < cnt) {
        xb_[b] = P.x[e_idx(eb_[b])];
        mb_[b] = P.mol[e_idx(eb_[b])];
      }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if (k0 + b >= cnt) break;
      const uint32_t e = eb_[b];
      const int j = e_idx(e);
      const double4 xj = xb_[b];
      double gx, gy, gz;
      ghost(xj, e, s.prd, gx, gy, gz);
      const double delx = xi.x - gx, dely = xi.y - gy, delz = xi.z - gz;
      const double rsq = rsq_ref(delx, dely, delz);
      const int tj = (int)xj.w - 1;
      const NPair &pp = s_np[ti * MAXT + tj];
      const bool tmapped = pp.tmapped != 0;
      if (rsq < P.tersoff_cutmaxsq && tmapped) {
        if (ns < MAXSH) P.tsh[(size_t)i * MAXSH + ns] = e;
        ns++;
      }
      if (tmapped && rsq < P.tersoff_cutmaxsq * widen) {
        if (nrt < MAXSH) rt_row[nrt] = (uint32_t)j;
        nrt++;
      }
      if (pp.imapped && mb_[b] == mi && rsq != 0.0) {
        if (rsq < pp.rc2w) {
          if (nrn < MAXSH) rn_row[nrn] = (uint32_t)j;
          nrn++;
        }
        if (rsq < pp.rcutsq) {
          const double v0 = gx - xi.x, v1 = gy - xi.y, v2 = gz - xi.z;
          // fixed-register slots (selects, not an indexed array)
          if (nn == 0) { nb0 = e; vetd[0][0] = v0; vetd[0][1] = v1; vetd[0][2] = v2; }
          else if (nn == 1) { nb1 = e; vetd[1][0] = v0; vetd[1][1] = v1; vetd[1][2] = v2; }
          else if (nn == 2) { nb2 = e; vetd[2][0] = v0; vetd[2][1] = v1; vetd[2][2] = v2; }
          else too_many = true;
          nn++;
        }
      }
    }
  }
  if (ns > MAXSH || nrt > MAXSH || nrn > MAXSH) {
    atomicOr(P.err, EB_SHORT);
    ns = ns < MAXSH ? ns : MAXSH;
    nrt = nrt < MAXSH ? nrt : MAXSH;
    nrn = nrn < MAXSH ? nrn : MAXSH;
  }
  P.tcnt[i] = ns;
  P.rtcnt[i] = nrt;
  P.rncnt[i] = nrn;
  if (too_many) {
    atomicOr(P.err, EB_NORMAL_COUNT);
    nn = 3;
  }
  uint32_t *nrow = P.nnb + (size_t)i * 4;
  nrow[0] = (uint32_t)nn;
  nrow[1] = nb0;
  nrow[2] = nb1;
  nrow[3] = nb2;
  if (WANT_VET) {
    double *vo = P.vet + (size_t)i * 9;
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 3; d++) vo[k * 3 + d] = vetd[k][d];
  }
  R vet[3][3];
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) vet[k][d] = (R)vetd[k][d];
  R *dno = P.dn + (size_t)i * 27;
  if (nn <= 1 || !s_np[ti * MAXT + ti].imapped) {
    P.nrm[i] = {0, 0, 1, 0};
    if (WANT_VET)
      for (int q = 0; q < 27; q++) dno[q] = 0;
    nrow[0] = (uint32_t)(nn <= 1 ? nn : 0);
    return;
  }
  R n1[3], nn2, nrm_len;
  if (normal_raw<R>(vet, nn, n1, nn2, nrm_len)) atomicOr(P.err, EB_NORMAL_ZERO);
  if (nn == 2) {
    // d n / d x_i of the two-neighbour normal n1 = v1 x v2 (ilp.cpp:835-927)
    R dpvdri[3][3];
    dpvdri[0][0] = 0; dpvdri[0][1] = vet[0][2] - vet[1][2]; dpvdri[0][2] = vet[1][1] - vet[0][1];
    dpvdri[1][0] = vet[1][2] - vet[0][2]; dpvdri[1][1] = 0; dpvdri[1][2] = vet[0][0] - vet[1][0];
    dpvdri[2][0] = vet[0][1] - vet[1][1]; dpvdri[2][1] = vet[1][0] - vet[0][0]; dpvdri[2][2] = 0;
    R dni[3];
    for (int b = 0; b < 3; b++) dni[b] = (n1[0] * dpvdri[0][b] + n1[1] * dpvdri[1][b] + n1[2] * dpvdri[2][b]) / nrm_len;
    R *dro = P.dndri + (size_t)i * 9;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) dro[a * 3 + b] = dpvdri[a][b] / nrm_len - n1[a] * dni[b] / nn2;
  }
  const R rl = (R)1 / nrm_len;
  P.nrm[i] = {n1[0] * rl, n1[1] * rl, n1[2] * rl, nrm_len};
  // d n / d x_k of every neighbour slot, for the virial tallies of the ILP kernel; the force
  // gather evaluates the one slot it needs itself, from positions, so plain force steps
  // neither store nor load these 27 values per atom
  if (WANT_VET)
    for (int m = 0; m < 3; m++) {
      R D[3][3];
      normal_dslot<R>(vet, nn, m, n1, nn2, nrm_len, D);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) dno[m * 9 + a * 3 + b] = D[a][b];
    }
}

// ---------------------------------------------------------------------------------------
// Per step 2: ILP pair kernel (calc_FvdW + calc_FRep, ilp.cpp:449-675), one wavefront per
// atom i.  For each j in the layer-filtered list with r < rc the wave evaluates
//   vdW(i,j), repulsion (i,j) with n_i and repulsion (j,i) with n_j,
// and keeps only what acts on i: f_i, S_i = sum_j w_ij del_ij (w = prodnorm*fpair1*Tap,
// the weight of the normal-derivative forces on i's normal neighbours), per-atom energy halves
// and the virial.  Lanes own list entries; the wave reduces with butterfly shuffles.

template <typename R>
struct Tap {
  R t, dt;
};

// Tap(x) = 20x^7 - 70x^6 + 84x^5 - 35x^4 + 1 with x = r/rcut (pair_ilp_graphene_hbn.h:80-117)
// written in r, 1 + r^4 (c4 + r (c5 + r (c6 + r c7))) with c_k = a_k / rcut^k, and its
// derivative 140 x^3 (x - 1)^3 / rcut = d7 r^3 (r - rcut)^3 with d7 = 140 / rcut^7 (the
// reference's expanded dTap polynomial, factored).  The coefficients are launch arguments,
// so they stay in scalar registers.
template <typename R>
__device__ __forceinline__ Tap<R> calc_tap(R r, const double *poly, double rcut, int tap_flag) {
  Tap<R> o;
  if (!tap_flag) {
    o.t = 1;
    o.dt = 0;
    return o;
  }
  const R r2 = r * r;
  R p = (R)poly[3] * r + (R)poly[2];
  p = p * r + (R)poly[1];
  p = p * r + (R)poly[0];
  o.t = (r2 * r2) * p + (R)1;
  const R y = r - (R)rcut;
  o.dt = (R)poly[4] * (r2 * r) * (y * y * y);
  return o;
}

// The ILP cutoff test of a displacement (x_i - x_j), returning r^2 in the pair precision: FP64
// in the reference's operation order; in mixed precision on the FP32 r^2 of the FP64
// displacement (a pair within rounding of rc contributes Tap(rc) = dTap(rc) = 0 either way).
// Both walks of a row use this one test, so they agree on every entry.
template <typename R>
__device__ __forceinline__ bool ilp_cut_test(double dxd, double dyd, double dzd, double cutsq, R &rsq) {
  if (sizeof(R) == 8) {
    const double r2 = rsq_ref(dxd, dyd, dzd);
    rsq = (R)r2;
    return r2 < cutsq;
  } else {
#pragma clang fp contract(off)
    const R dx = (R)dxd, dy = (R)dyd, dz = (R)dzd;
    const R r2 = (dx * dx + dy * dy) + dz * dz;
    rsq = r2;
    return r2 < (R)cutsq;
  }
}

// minimum resident waves per SIMD requested for the plain (force-only) ILP pair kernel: with
// the prefetched chunk in registers four waves fit without spilling (five spill and lose 20 %);
// the tally variants keep their registers
#ifndef LMFF_ILP_MIN_WAVES
#define LMFF_ILP_MIN_WAVES 4
#endif

// SYM: every mapped type pair has the same vdW and exp(-lambda (r - z0)) parameters in both
// orders (true of both shipped parameter files), so neither per-lane choice below is needed
template <typename R, bool EF, bool VF, bool VA, bool SYM>
__global__ __launch_bounds__(256, (EF || VF || VA) ? 1 : LMFF_ILP_MIN_WAVES) void k_ilp(DevPtrs<R> P) {
  using M = MathOps<R>;
  // one wavefront per atom: the atom index is wave-uniform (scalar registers)
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = __builtin_amdgcn_readfirstlane(xcd_block() * (blockDim.x >> 6) + wv);
  const int lane = threadIdx.x & 63;
  // the pair-parameter table (MAXT x MAXT entries) is read per lane and per entry: from LDS
  __shared__ IlpPair<R> s_ip[MAXT * MAXT];
  // per wave: back-segment entries within rc waiting for a full batch
  __shared__ uint32_t s_q[4][128];
  stage_table(s_ip, P.ip, MAXT * MAXT);
  __syncthreads();
  if (i >= P.n) return;
  // atoms of systems the half-pair kernel takes in this evaluation
  if (P.half_on && P.syshalf[P.sys[i]]) return;
  const SysDev &s = P.sys_tab[P.sys[i]];
  const double prd[3] = {s.prd[0], s.prd[1], s.prd[2]};
  const double4 xi = P.x[i];
  const int ti = (int)xi.w - 1;
  const auto ni4 = P.nrm[i];
  const R nix = ni4.x, niy = ni4.y, niz = ni4.z;
  // the back segment (r > rc + margin at the build) can hold a pair within rc only after some
  // atom has moved more than margin/2 since the build
  const int nfront = P.icnt[i], nback = *P.farflag ? P.ocnt[i] : 0;
  const uint32_t *row = P.ilist + (size_t)i * P.icap;
  const int itag = P.tag[i];
  double fx = 0, fy = 0, fz = 0, sx = 0, sy = 0, sz = 0, ev = 0, er = 0;
  double w[6] = {0, 0, 0, 0, 0, 0}, va[6] = {0, 0, 0, 0, 0, 0}, tt[6] = {0, 0, 0, 0, 0, 0};
  // mixed precision outside the tallies: each lane sums its pairs' force and S terms in FP32
  // (a lane sees one pair in 64 of the row: at most a few tens) and adds that sum to the FP64
  // accumulators once, after the loop; the wave reduction and everything after it stay FP64
  constexpr bool LANE32 = sizeof(R) == 4 && !VF && !VA;
  R fxl = 0, fyl = 0, fzl = 0, sxl = 0, syl = 0, szl = 0;
  // the pair terms of list entry e (j = its atom, xj its position) for this lane
  auto pair_terms = [&](const uint32_t e, const double4 xj) {
      const int j = e_idx(e);
      double gx, gy, gz;
      // x_j + 0*prd == x_j exactly, so the image arithmetic is skipped when no lane needs it
      if (__ballot((e >> IDX_BITS) != 0x15u) == 0) {
        gx = xj.x;
        gy = xj.y;
        gz = xj.z;
      } else {
        ghost(xj, e, prd, gx, gy, gz);
      }
      const double dxd = xi.x - gx, dyd = xi.y - gy, dzd = xi.z - gz;
      const R dx = (R)dxd, dy = (R)dyd, dz = (R)dzd;
      R rsq;
      if (!ilp_cut_test<R>(dxd, dyd, dzd, P.ilp_cutsq, rsq)) return;
      const int tj = (int)xj.w - 1;
      const IlpPair<R> &pij = s_ip[ti * MAXT + tj];
      const IlpPair<R> &pji = s_ip[tj * MAXT + ti];
      const R rinv = M::rsqrt(rsq);
      const R r = rsq * rinv;
      const Tap<R> tp = calc_tap<R>(r, P.tap_poly, P.rcut, P.tap_flag);
      // ---- van der Waals, one half per ordered pair ----
      const IlpPair<R> *pv = &pij;
      if (!SYM && !pij.vsym && tag_skip(itag, P.tag[j], xi.x, xi.y, xi.z, gx, gy, gz)) pv = &pji;
      // E = -C6 / (r^6 (1 + e)), e = exp(-d (r/seff - 1)) (ilp.cpp:510-530); fpv = -dE/dr / r
      const R r2inv = rinv * rinv;
      const R edamp = M::ex(-pv->d * (r * pv->inv_seff - (R)1));
      const R TSinv = M::rcp((R)1 + edamp);
      const R w6 = r2inv * r2inv * r2inv * TSinv;
      const R Vvdw = -pv->C6 * w6;
      const R fpv = r2inv * w6 * (pv->C6_d_seff * r * edamp * TSinv - (R)6 * pv->C6);
      // ---- repulsion (i,j) with n_i ----
      const R exp0 = M::ex(-pij.lambda * (r - pij.z0));
      const R pn1 = nix * dx + niy * dy + niz * dz;
      const R rd1 = (rsq - pn1 * pn1) * pij.d2inv;
      const R frho1 = M::ex(-rd1) * pij.C;
      const R V1 = exp0 * (pij.half_eps + frho1);
      const R fq1 = (R)2 * exp0 * frho1 * pij.d2inv;
      const R a1 = pn1 * fq1;
      // ---- repulsion (j,i) with n_j; del_ji = -del ----
      const auto nj4 = P.nrm[j];
      const R exp0b = (SYM || pij.esym) ? exp0 : M::ex(-pji.lambda * (r - pji.z0));
      const R pn2 = -(nj4.x * dx + nj4.y * dy + nj4.z * dz);
      const R rd2 = (rsq - pn2 * pn2) * pji.d2inv;
      const R frho2 = M::ex(-rd2) * pji.C;
      const R V2 = exp0b * (pji.half_eps + frho2);
      const R fq2 = (R)2 * exp0b * frho2 * pji.d2inv;
      const R a2 = pn2 * fq2;
      // Force on i: fsum_vdw del + fkc_ij - fkc_ji (ilp.cpp:525-534, 640-648) with
      //   fsum_vdw = fpv Tap - V_vdw dTap / r,  fkc_ij = (del fs1 - a1 n_i) Tap - V1 dTap del / r,
      //   fkc_ji = (-del fs2 - a2 n_j) Tap + V2 dTap del / r,  fs = lambda V / r + fq,
      // collected as  C del + Tap (a2 n_j - a1 n_i)  (one scalar per pair instead of three
      // vectors).  The weight of the normal-derivative forces is w_ij = a1 Tap.
      const R ta1 = a1 * tp.t, ta2 = a2 * tp.t;
      const R Cd = tp.t * (fpv + rinv * (pij.lambda * V1 + pji.lambda * V2) + fq1 + fq2) -
                   tp.dt * rinv * (Vvdw + V1 + V2);
      if (LANE32) {
        fxl += Cd * dx + (ta2 * nj4.x - ta1 * nix);
        fyl += Cd * dy + (ta2 * nj4.y - ta1 * niy);
        fzl += Cd * dz + (ta2 * nj4.z - ta1 * niz);
        sxl += ta1 * dx;
        syl += ta1 * dy;
        szl += ta1 * dz;
      } else {
        fx += (double)(Cd * dx + (ta2 * nj4.x - ta1 * nix));
        fy += (double)(Cd * dy + (ta2 * nj4.y - ta1 * niy));
        fz += (double)(Cd * dz + (ta2 * nj4.z - ta1 * niz));
        sx += (double)(ta1 * dx);
        sy += (double)(ta1 * dy);
        sz += (double)(ta1 * dz);
      }
      if (EF) {
        ev += 0.5 * (double)(Vvdw * tp.t);
        er += 0.5 * (double)(tp.t * V1) + 0.5 * (double)(tp.t * V2);
      }
      if (VF || VA) {
        // the separate force terms, for the tallies
        const R fsv = fpv * tp.t - Vvdw * tp.dt * rinv;
        const R fs1 = pij.lambda * rinv * V1 + fq1, b1 = V1 * tp.dt * rinv;
        const R fs2 = pji.lambda * rinv * V2 + fq2, b2 = V2 * tp.dt * rinv;
        const R fkx = (dx * fs1 - a1 * nix) * tp.t - b1 * dx;
        const R fky = (dy * fs1 - a1 * niy) * tp.t - b1 * dy;
        const R fkz = (dz * fs1 - a1 * niz) * tp.t - b1 * dz;
        const R gkx = (-dx * fs2 - a2 * nj4.x) * tp.t + b2 * dx;
        const R gky = (-dy * fs2 - a2 * nj4.y) * tp.t + b2 * dy;
        const R gkz = (-dz * fs2 - a2 * nj4.z) * tp.t + b2 * dz;
        const R wgt = ta1;
        // vdW: half of del (x) del * fsum per ordered pair; repulsion (i,j): del (x) fkc
        const double hv = 0.5 * (double)fsv;
        const double dd[3] = {dxd, dyd, dzd};
        const double fk[3] = {(double)fkx, (double)fky, (double)fkz};
        const double gk[3] = {(double)gkx, (double)gky, (double)gkz};
        if (VF) {
          w[0] += hv * dd[0] * dd[0] + dd[0] * fk[0];
          w[1] += hv * dd[1] * dd[1] + dd[1] * fk[1];
          w[2] += hv * dd[2] * dd[2] + dd[2] * fk[2];
          w[3] += hv * dd[0] * dd[1] + dd[0] * fk[1];
          w[4] += hv * dd[0] * dd[2] + dd[0] * fk[2];
          w[5] += hv * dd[1] * dd[2] + dd[1] * fk[2];
        }
        if (VA) {
          // ev_tally (vdW) half, ev_tally_xyz(i,j) half for (i,j), and the j-side half of (j,i):
          // 0.5 * del_ji (x) fkc_ji = 0.5 * (-del) (x) (-gk)... fkc_ji = gk, del_ji = -del
          va[0] += hv * dd[0] * dd[0] + 0.5 * (dd[0] * fk[0] - dd[0] * gk[0]);
          va[1] += hv * dd[1] * dd[1] + 0.5 * (dd[1] * fk[1] - dd[1] * gk[1]);
          va[2] += hv * dd[2] * dd[2] + 0.5 * (dd[2] * fk[2] - dd[2] * gk[2]);
          va[3] += hv * dd[0] * dd[1] + 0.5 * (dd[0] * fk[1] - dd[0] * gk[1]);
          va[4] += hv * dd[0] * dd[2] + 0.5 * (dd[0] * fk[2] - dd[0] * gk[2]);
          va[5] += hv * dd[1] * dd[2] + 0.5 * (dd[1] * fk[2] - dd[1] * gk[2]);
          const double wd = (double)wgt;
          tt[0] += wd * dd[0] * dd[0];
          tt[1] += wd * dd[1] * dd[1];
          tt[2] += wd * dd[2] * dd[2];
          tt[3] += wd * dd[0] * dd[1];
          tt[4] += wd * dd[0] * dd[2];
          tt[5] += wd * dd[1] * dd[2];
          // k-terms of pair (j,i): half of (x_k - x_i) (x) fk lands on i (ev_tally_xyz(k, i)),
          // fk = -w_ji * D_k(j)^T del_ji, x_k - x_i = vet_k(j) + del_ji
          const uint32_t *nj = P.nnb + (size_t)j * 4;
          const int cj = (int)nj[0];
          const double wji = (double)(a2 * tp.t);
          for (int kk = 0; kk < cj; kk++) {
            const R *D = P.dn + (size_t)j * 27 + kk * 9;
            const double *vt = P.vet + (size_t)j * 9 + kk * 3;
            double fkk[3];
            for (int c = 0; c < 3; c++)
              fkk[c] = -wji * (-(double)D[0 * 3 + c] * dd[0] - (double)D[1 * 3 + c] * dd[1] - (double)D[2 * 3 + c] * dd[2]);
            const double rk[3] = {vt[0] - dd[0], vt[1] - dd[1], vt[2] - dd[2]};
            va[0] += 0.5 * rk[0] * fkk[0];
            va[1] += 0.5 * rk[1] * fkk[1];
            va[2] += 0.5 * rk[2] * fkk[2];
            va[3] += 0.5 * rk[0] * fkk[1];
            va[4] += 0.5 * rk[0] * fkk[2];
            va[5] += 0.5 * rk[1] * fkk[2];
          }
        }
      }
  };
  // front segment (r < rc + margin at the build; most of it within rc): software pipeline, the
  // entry two chunks ahead and the position one chunk ahead are in flight while a chunk
  // computes (15 % off the FP32 kernel, 2 % off the FP64 one)
  uint32_t e_n1 = 0, e_n2 = 0;
  double4 x_n1 = {0, 0, 0, 0};
  if (lane < nfront) e_n1 = row[lane];
  if (lane + 64 < nfront) e_n2 = row[lane + 64];
  if (lane < nfront) x_n1 = P.x[e_idx(e_n1)];
  for (int k = lane; k < nfront; k += 64) {
    const uint32_t e = e_n1;
    const double4 xj = x_n1;
    e_n1 = e_n2;
    // unconditional (past the row's end a lane reloads an entry it holds or atom 0, in bounds):
    // no branch, so the waits inside the pair terms can leave these loads in flight
    x_n1 = P.x[e_idx(e_n1)];
    e_n2 = row[k + 128 < nfront ? k + 128 : nfront - 1];
    pair_terms(e, xj);
  }
  // back segment (walked once the far flag is up): only a few of its entries have come within
  // rc, but nearly every chunk of 64 holds one, so the entries are tested first and the pair
  // terms run on full batches of 64 entries that passed, queued in LDS
  if (nback > 0) {
    const uint32_t *brow = row + (P.icap - nback);
    int qn = 0;  // queued entries (wave-uniform, below 64 between chunks)
    // pair terms of the first min(qn, 64) queued entries; the rest moves to the queue's front
    auto drain = [&]() {
      // the queue is written and read by lanes of this wave only
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const int take = qn < 64 ? qn : 64;
      const uint32_t qe = s_q[wv][lane];
      const uint32_t rest = lane + 64 < qn ? s_q[wv][lane + 64] : 0u;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (lane + 64 < qn) s_q[wv][lane] = rest;
      qn -= take;
      if (lane < take) pair_terms(qe, P.x[e_idx(qe)]);
    };
    for (int k0 = 0; k0 < nback; k0 += 64) {
      const int k = k0 + lane;
      bool in = false;
      uint32_t e = 0;
      if (k < nback) {
        e = brow[k];
        const double4 xj = P.x[e_idx(e)];
        double gx, gy, gz;
        ghost(xj, e, prd, gx, gy, gz);
        R r2;
        in = ilp_cut_test<R>(xi.x - gx, xi.y - gy, xi.z - gz, P.ilp_cutsq, r2);
      }
      const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
      const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      if (in) s_q[wv][qn + below] = e;
      qn += (int)__popcll(m);
      if (qn >= 64) drain();
    }
    if (qn > 0) drain();
  }
  if (LANE32) {
    fx += (double)fxl; fy += (double)fyl; fz += (double)fzl;
    sx += (double)sxl; sy += (double)syl; sz += (double)szl;
  }
  fx = wsum(fx); fy = wsum(fy); fz = wsum(fz);
  sx = wsum(sx); sy = wsum(sy); sz = wsum(sz);
  if (EF) { ev = wsum(ev); er = wsum(er); }
  if (VF) for (int c = 0; c < 6; c++) w[c] = wsum(w[c]);
  if (VA) for (int c = 0; c < 6; c++) { va[c] = wsum(va[c]); tt[c] = wsum(tt[c]); }
  if (lane != 0) return;
  P.fil[i] = {fx, fy, fz, 0.0};
  P.S[i] = {sx, sy, sz, 0.0};
  const size_t n = (size_t)P.n;
  if (EF) { P.th[0 * n + i] = ev; P.th[1 * n + i] = er; }
  if (VF) {
    // virial share of the normal-derivative forces of i's normal: sum_k vet_k (x) F_k,
    // F_k = -D_k^T S_i  (equals the F.r virial of those forces by translation invariance)
    const uint32_t *nrow = P.nnb + (size_t)i * 4;
    const int ci = (int)nrow[0];
    for (int kk = 0; kk < ci; kk++) {
      const R *D = P.dn + (size_t)i * 27 + kk * 9;
      const double *vt = P.vet + (size_t)i * 9 + kk * 3;
      double Fk[3];
      for (int c = 0; c < 3; c++) Fk[c] = -((double)D[0 * 3 + c] * sx + (double)D[1 * 3 + c] * sy + (double)D[2 * 3 + c] * sz);
      w[0] += vt[0] * Fk[0];
      w[1] += vt[1] * Fk[1];
      w[2] += vt[2] * Fk[2];
      w[3] += vt[0] * Fk[1];
      w[4] += vt[0] * Fk[2];
      w[5] += vt[1] * Fk[2];
    }
    for (int c = 0; c < 6; c++) P.th[(3 + c) * n + i] = w[c];
  }
  if (VA) {
    for (int c = 0; c < 6; c++) {
      P.vatom[c * n + i] = va[c];
      P.T[(size_t)i * 6 + c] = tt[c];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Half-pair ILP (calc_FvdW + calc_FRep, ilp.cpp:449-675, each unordered pair once).  The pull
// kernel above evaluates every pair from both of its atoms; here one workgroup takes the atoms
// of one tile (tile[0] x tile[1] bins) of one z bin and evaluates each pair once, from the atom
// whose z bin is lower (from the lower atom index within one z bin).  The terms acting on i go
// into FP64 lane sums as in the pull kernel; the partner's terms (-F_ij and -Tap a_ji del) are
// added in exact 64-bit fixed point to LDS sums over a window (the bins of the partner z bin
// around the tile, one slot per atom image, slot from the list build), and each window goes out
// to the per-atom sums acc[] with integer atomics when the pass over its z bin ends.  Integer
// sums do not depend on the order of the adds, so results are reproducible bit for bit and do
// not depend on the other systems of a launch.  k_ilp_combine then adds acc into fil and S.

// scale and magic are passed at run time (equal to FX_SCALE and FX_MAGIC), so that the compiler
// keeps them in registers across the pair loop: one fma and an add to the high word per value
// (as compile-time constants the magic was rebuilt by two moves before every fma)
__device__ __forceinline__ unsigned long long fx_of(double v, double scale, double magic) {
  return (unsigned long long)__double_as_longlong(fma(v, scale, magic)) - FX_MAGIC_BITS;
}

// values of an atom's own sums: fx fy fz Sx Sy Sz, then e_vdw e_rep, then the virial
__host__ __device__ constexpr int half_pstr(bool EF, bool VF) { return 6 + (EF || VF ? 2 : 0) + (VF ? 6 : 0); }

// OVF: some window holds more atoms than wcap (the LDS); without it every slot is below wcap
template <typename R, bool EF, bool VF, bool SYM, bool OVF>
__global__ __launch_bounds__(64 * HalfWaves<R>::n, HalfWaves<R>::per_simd) void k_ilp_half(DevPtrs<R> P, int wcap) {
  using M = MathOps<R>;
  constexpr int PSTR = half_pstr(EF, VF);
  extern __shared__ unsigned long long s_acc[];  // [6][wcap]: fx fy fz Sx Sy Sz of the window's atoms
  __shared__ IlpPair<R> s_ip[MAXT * MAXT];
  constexpr int HALF_WAVES = HalfWaves<R>::n;
  __shared__ uint32_t s_q[HALF_WAVES][128];
  __shared__ int s_bin[WSTR];
  __shared__ int s_batom[WSTR];  // first atom of each window bin (bins that hold atoms)
  __shared__ int s_row[4][2];
  __shared__ int s_next;  // the next tile atom of the pass to hand out
  stage_table(s_ip, P.ip, MAXT * MAXT);
  for (int p = threadIdx.x; p < 6 * wcap; p += blockDim.x) s_acc[p] = 0;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = xcd_chunk_block(16);
  const int si = sys_of_tile(P.sys_tab, P.nsys, b);
  const SysDev &s = P.sys_tab[si];
  const int ntt = s.ntile[0] * s.ntile[1];
  const int loc = b - s.tile_off;
  const bool mine = s.tile[0] > 0 && P.syshalf[si] && loc >= 0 && loc < ntt * s.nc[2];
  const int tx = mine ? loc % s.ntile[0] : 0, ty = mine ? (loc / s.ntile[0]) % s.ntile[1] : 0;
  const int za = mine ? loc / ntt : 0;
  const int x0 = tx * s.tile[0], x1 = min(x0 + s.tile[0], s.nc[0]);
  if (threadIdx.x < 4) {
    int a0 = 0, cnt = 0;
    const int yy = ty * s.tile[1] + (int)threadIdx.x;
    if (mine && (int)threadIdx.x < s.tile[1] && yy < s.nc[1]) {
      const int base = s.cell_off + (za * s.nc[1] + yy) * s.nc[0];
      a0 = P.cell_start[base + x0];
      cnt = P.cell_start[base + x1] - a0;
    }
    s_row[threadIdx.x][0] = a0;
    s_row[threadIdx.x][1] = cnt;
  }
  __syncthreads();
  const int n0 = s_row[0][1], n1 = n0 + s_row[1][1], n2 = n1 + s_row[2][1], ntile_atoms = n2 + s_row[3][1];
  if (ntile_atoms == 0) return;  // the whole workgroup (nothing here or not a half system)
  auto tile_atom = [&](int k) {
    return k < n0 ? s_row[0][0] + k : k < n1 ? s_row[1][0] + (k - n0) : k < n2 ? s_row[2][0] + (k - n1)
                                                                                 : s_row[3][0] + (k - n2);
  };
  const double prd[3] = {s.prd[0], s.prd[1], s.prd[2]};
  const int far = *P.farflag;
  const double cut = sqrt(P.cutneighsq);
  const int rx = stencil_rx(s, cut), ry = stencil_ry(s, cut);
  const int wxb = s.tile[0] + 2 * rx, nbins = wxb * (s.tile[1] + 2 * ry);
  const int rz = (int)ceil(cut * s.inv_bs[2]);
  const int dzmax = min(rz, NSEG - 2);
  const size_t n = (size_t)P.n;
  const int icap = P.icap;
  constexpr bool LANE32 = sizeof(R) == 4 && !VF;
  R cmax = 0;  // largest |C|, |Tap a_ij|, |Tap a_ji| of this lane (fixed-point range check)
  const double fxs = P.fx_scale, fxm = P.fx_magic;
  // an atom's own sums go to its slot of ipart: the first pass (dz = 0) stores them (zeros when
  // it has no entries there), every later pass with entries for the atom adds its sums to them.
  // One workgroup makes all passes of its tile's atoms, and the barrier at the end of a pass
  // orders the passes, so the sums are added in pass order whichever wave takes the atom.
  auto put0 = [&](int i, int dz) {
    if (dz == 0 && lane < PSTR) P.ipart[(size_t)i * PSTR + lane] = 0.0;
  };
  // the sums of eight values (wsum8), lanes 0, 8, ..: value c + off at slot c + off
  auto put8 = [&](int i, int dz, double v, int off) {
    const int c = off + (lane >> 3);
    if ((lane & 7) == 0 && c < PSTR) {
      double *q = P.ipart + (size_t)i * PSTR + c;
      *q = dz == 0 ? v : *q + v;
    }
  };
  for (int dz = 0; dz < P.npass; dz++) {
    int zb = za + dz;
    bool live = dz <= dzmax;
    if (zb >= s.nc[2]) {
      if (s.per[2]) zb -= s.nc[2];
      else live = false;
    }
    const int w = live ? s.tile_off + zb * ntt + ty * s.ntile[0] + tx : 0;
    // does any atom of the tile have entries at this dz?
    int any = 0;
    if (threadIdx.x == 0) s_next = HALF_WAVES;
    if (live) {
      const int *wt = P.win + (size_t)w * WSTR;
      for (int q = threadIdx.x; q <= nbins; q += blockDim.x) {
        const int w0 = wt[q];
        s_bin[q] = w0;
        if (q < nbins && wt[q + 1] > w0) {
          // a bin that holds atoms lies in the grid (k_win_table counts no other)
          const int kr = q / wxb, u = q - kr * wxb;
          int uy = ty * s.tile[1] - ry + kr, ux = tx * s.tile[0] - rx + u, sh;
          if (s.per[1]) uy = wrap_bin(uy, s.nc[1], sh);
          if (s.per[0]) ux = wrap_bin(ux, s.nc[0], sh);
          s_batom[q] = P.cell_start[s.cell_off + (zb * s.nc[1] + uy) * s.nc[0] + ux];
        }
      }
      for (int k = wv; k < ntile_atoms; k += HALF_WAVES) {
        const int i = tile_atom(k);
        const uint16_t *sf = P.segf + (size_t)i * NSEG, *sb = P.segb + (size_t)i * NSEG;
        any |= (sf[dz + 1] > sf[dz]) || (far && sb[dz + 1] > sb[dz]);
      }
    }
    if (!__syncthreads_or(any)) {
      for (int k = wv; k < ntile_atoms; k += HALF_WAVES) put0(tile_atom(k), dz);
      continue;
    }
    // a window larger than the LDS holds (OVF kernels only): the slots from wcap on add to acc[]
    // directly
    const bool wovf = OVF && __builtin_amdgcn_readfirstlane(s_bin[nbins]) > wcap;
    // the waves take the tile's atoms one at a time from a shared counter, so that a wave with
    // long rows does not hold the others at the barrier that ends the pass (which wave takes an
    // atom does not change any result: an atom's own sums come from its row alone, and the
    // partner sums are integers)
    auto next_atom = [&]() {
      int t = 0;
      if (lane == 0) t = atomicAdd(&s_next, 1);
      return __builtin_amdgcn_readfirstlane(t);
    };
    for (int k = wv; k < ntile_atoms; k = next_atom()) {
      const int i = __builtin_amdgcn_readfirstlane(tile_atom(k));
      const uint16_t *sf = P.segf + (size_t)i * NSEG, *sb = P.segb + (size_t)i * NSEG;
      const int f0 = sf[dz], f1 = sf[dz + 1];
      const int b0 = far ? (int)sb[dz] : 0, b1 = far ? (int)sb[dz + 1] : 0;
      if (f1 <= f0 && b1 <= b0) {
        put0(i, dz);
        continue;
      }
      const double4 xi = P.x[i];
      const int ti = (int)xi.w - 1;
      const auto ni4 = P.nrm[i];
      const R nix = ni4.x, niy = ni4.y, niz = ni4.z;
      const int itag = P.tag[i];
      const uint32_t *row = P.ilist + (size_t)i * icap;
      const uint16_t *lrow = P.lidx + (size_t)i * icap;
      double fx = 0, fy = 0, fz = 0, sx = 0, sy = 0, sz = 0, ev = 0, er = 0;
      double wv6[6] = {0, 0, 0, 0, 0, 0};
      R fxl = 0, fyl = 0, fzl = 0, sxl = 0, syl = 0, szl = 0;
      using N4 = typename std::conditional<sizeof(R) == 8, double4, float4>::type;
      // mode: where j's terms go -- 0 the window's LDS sums (slot < wcap), 1 the global sums
      // (slot >= wcap), 2 decided per entry
      auto terms = [&](auto mode, const uint32_t e, const uint32_t slot, const double4 xj, const N4 nj4) {
        constexpr int MODE = decltype(mode)::value;
        const int j = e_idx(e);  // within i's own z bin (dz = 0) the rows hold only j > i
        double gx, gy, gz;
        if (__ballot((e >> IDX_BITS) != 0x15u) == 0) {
          gx = xj.x;
          gy = xj.y;
          gz = xj.z;
        } else {
          ghost(xj, e, prd, gx, gy, gz);
        }
        const double dxd = xi.x - gx, dyd = xi.y - gy, dzd = xi.z - gz;
        const R dx = (R)dxd, dy = (R)dyd, dzz = (R)dzd;
        R rsq;
        if (!ilp_cut_test<R>(dxd, dyd, dzd, P.ilp_cutsq, rsq)) return;
        const int tj = (int)xj.w - 1;
        const IlpPair<R> &pij = s_ip[ti * MAXT + tj];
        const IlpPair<R> &pji = s_ip[tj * MAXT + ti];
        const R rinv = M::rsqrt(rsq);
        const R r = rsq * rinv;
        const Tap<R> tp = calc_tap<R>(r, P.tap_poly, P.rcut, P.tap_flag);
        // van der Waals with the parameters of the atom the reference gives the pair to
        const IlpPair<R> *pv = &pij;
        if (!SYM && !pij.vsym && tag_skip(itag, P.tag[j], xi.x, xi.y, xi.z, gx, gy, gz)) pv = &pji;
        const R r2inv = rinv * rinv;
        const R edamp = M::ex(-pv->d * (r * pv->inv_seff - (R)1));
        const R TSinv = M::rcp((R)1 + edamp);
        const R w6 = r2inv * r2inv * r2inv * TSinv;
        const R Vvdw = -pv->C6 * w6;
        const R fpv = r2inv * w6 * (pv->C6_d_seff * r * edamp * TSinv - (R)6 * pv->C6);
        // repulsion (i,j) with n_i and (j,i) with n_j
        const R exp0 = M::ex(-pij.lambda * (r - pij.z0));
        const R pn1 = nix * dx + niy * dy + niz * dzz;
        const R rd1 = (rsq - pn1 * pn1) * pij.d2inv;
        const R frho1 = M::ex(-rd1) * pij.C;
        const R V1 = exp0 * (pij.half_eps + frho1);
        const R fq1 = (R)2 * exp0 * frho1 * pij.d2inv;
        const R a1 = pn1 * fq1;
        const R exp0b = (SYM || pij.esym) ? exp0 : M::ex(-pji.lambda * (r - pji.z0));
        const R pn2 = -(nj4.x * dx + nj4.y * dy + nj4.z * dzz);
        const R rd2 = (rsq - pn2 * pn2) * pji.d2inv;
        const R frho2 = M::ex(-rd2) * pji.C;
        const R V2 = exp0b * (pji.half_eps + frho2);
        const R fq2 = (R)2 * exp0b * frho2 * pji.d2inv;
        const R a2 = pn2 * fq2;
        const R ta1 = a1 * tp.t, ta2 = a2 * tp.t;
        const R Cd = tp.t * (fpv + rinv * (pij.lambda * V1 + pji.lambda * V2) + fq1 + fq2) -
                     tp.dt * rinv * (Vvdw + V1 + V2);
        // force of the pair on i: C del + Tap (a_ji n_j - a_ij n_i); on j its negative
        const R Fx = Cd * dx + (ta2 * nj4.x - ta1 * nix);
        const R Fy = Cd * dy + (ta2 * nj4.y - ta1 * niy);
        const R Fz = Cd * dzz + (ta2 * nj4.z - ta1 * niz);
        if (LANE32) {
          fxl += Fx; fyl += Fy; fzl += Fz;
          sxl += ta1 * dx; syl += ta1 * dy; szl += ta1 * dzz;
        } else {
          fx += (double)Fx; fy += (double)Fy; fz += (double)Fz;
          sx += (double)(ta1 * dx); sy += (double)(ta1 * dy); sz += (double)(ta1 * dzz);
        }
        // j's terms: -F and S_j += Tap a_ji del_ji = -Tap a_ji del
        auto lds_add = [&]() {
          unsigned long long *pa = s_acc + slot;
          __hip_atomic_fetch_add(pa, fx_of(-(double)Fx, fxs, fxm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(pa + wcap, fx_of(-(double)Fy, fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(pa + 2 * wcap, fx_of(-(double)Fz, fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(pa + 3 * wcap, fx_of(-(double)(ta2 * dx), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(pa + 4 * wcap, fx_of(-(double)(ta2 * dy), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(pa + 5 * wcap, fx_of(-(double)(ta2 * dzz), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        // a slot beyond what the LDS holds (a window larger than wcap): straight to acc[]
        auto global_add = [&]() {
          unsigned long long *ga = P.acc + j;
          __hip_atomic_fetch_add(ga, fx_of(-(double)Fx, fxs, fxm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(ga + n, fx_of(-(double)Fy, fxs, fxm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(ga + 2 * n, fx_of(-(double)Fz, fxs, fxm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(ga + 3 * n, fx_of(-(double)(ta2 * dx), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(ga + 4 * n, fx_of(-(double)(ta2 * dy), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(ga + 5 * n, fx_of(-(double)(ta2 * dzz), fxs, fxm), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        };
        if constexpr (MODE == 0) {
          lds_add();
        } else if constexpr (MODE == 1) {
          global_add();
        } else {
          if (slot < (uint32_t)wcap) lds_add();
          else global_add();
        }
        cmax = fmax(cmax, fmax(fabs(Cd), fmax(fabs(ta1), fabs(ta2))));
        if (EF) {
          ev += (double)(Vvdw * tp.t);
          er += (double)(tp.t * V1) + (double)(tp.t * V2);
        }
        if (VF) {
          // del (x) F_ij: the pair's share of the F.r virial (both atoms' halves)
          wv6[0] += dxd * (double)Fx;
          wv6[1] += dyd * (double)Fy;
          wv6[2] += dzd * (double)Fz;
          wv6[3] += dxd * (double)Fy;
          wv6[4] += dxd * (double)Fz;
          wv6[5] += dyd * (double)Fz;
        }
      };
      if constexpr (!OVF) {
        // front entries of this dz, pipelined as in the pull kernel
        const int cnt = f1 - f0;
        const uint32_t *fr = row + f0;
        const uint16_t *fl = lrow + f0;
        uint32_t e_n1 = 0, e_n2 = 0;
        uint32_t l_n1 = 0, l_n2 = 0;
        double4 x_n1 = {0, 0, 0, 0};
        N4 n_n1 = {0, 0, 0, 0};
        if (lane < cnt) { e_n1 = fr[lane]; l_n1 = fl[lane]; }
        if (lane + 64 < cnt) { e_n2 = fr[lane + 64]; l_n2 = fl[lane + 64]; }
        // FP32: the partner's normal is fetched a chunk ahead with its position (FP64: in the
        // chunk, which measured faster at the FP64 kernel's register budget)
        constexpr bool NPRE = sizeof(R) == 4;
        if (lane < cnt) {
          x_n1 = P.x[e_idx(e_n1)];
          if (NPRE) n_n1 = P.nrm[e_idx(e_n1)];
        }
        for (int k2 = lane; k2 < cnt; k2 += 64) {
          const uint32_t e = e_n1, l = l_n1;
          const double4 xj = x_n1;
          // FP64: this chunk's normal is loaded before the prefetches, and the prefetches are
          // unconditional (past the row's end a lane reloads an entry it already holds or atom
          // 0, both in bounds), so that waiting for the normal leaves them in flight (vector
          // loads complete in issue order, and a wait after a branch counts every load)
          const N4 nj = NPRE ? n_n1 : P.nrm[e_idx(e)];
          e_n1 = e_n2;
          l_n1 = l_n2;
          x_n1 = P.x[e_idx(e_n1)];
          if (NPRE) n_n1 = P.nrm[e_idx(e_n1)];
          {
            const int kn = k2 + 128 < cnt ? k2 + 128 : cnt - 1;
            e_n2 = fr[kn];
            l_n2 = fl[kn];
          }
          terms(std::integral_constant<int, 0>{}, e, l, xj, nj);
        }
      } else {
        // the same loop over entries [k0, k1) of the front segment, their partner sums going to
        // the LDS window (mode 0) or to acc[] (mode 1)
        auto front = [&](auto mode, const int k0, const int k1) {
          const int cnt = k1 - k0;
          if (cnt <= 0) return;
          const uint32_t *fr = row + k0;
          const uint16_t *fl = lrow + k0;
          uint32_t e_n1 = 0, e_n2 = 0;
          uint32_t l_n1 = 0, l_n2 = 0;
          double4 x_n1 = {0, 0, 0, 0};
          N4 n_n1 = {0, 0, 0, 0};
          if (lane < cnt) { e_n1 = fr[lane]; l_n1 = fl[lane]; }
          if (lane + 64 < cnt) { e_n2 = fr[lane + 64]; l_n2 = fl[lane + 64]; }
          // FP32: the partner's normal is fetched a chunk ahead with its position (FP64: in the
          // chunk, which measured faster at the FP64 kernel's register budget)
          constexpr bool NPRE = sizeof(R) == 4;
          if (lane < cnt) {
            x_n1 = P.x[e_idx(e_n1)];
            if (NPRE) n_n1 = P.nrm[e_idx(e_n1)];
          }
          for (int k2 = lane; k2 < cnt; k2 += 64) {
            const uint32_t e = e_n1, l = l_n1;
            const double4 xj = x_n1;
            // FP64: this chunk's normal is loaded before the prefetches, and the prefetches are
            // unconditional (past the row's end a lane reloads an entry it already holds or atom
            // 0, both in bounds), so that waiting for the normal leaves them in flight (vector
            // loads complete in issue order, and a wait after a branch counts every load)
            const N4 nj = NPRE ? n_n1 : P.nrm[e_idx(e)];
            e_n1 = e_n2;
            l_n1 = l_n2;
            x_n1 = P.x[e_idx(e_n1)];
            if (NPRE) n_n1 = P.nrm[e_idx(e_n1)];
            {
              const int kn = k2 + 128 < cnt ? k2 + 128 : cnt - 1;
              e_n2 = fr[kn];
              l_n2 = fl[kn];
            }
            terms(mode, e, l, xj, nj);
          }
        };
        if (!wovf) {
          front(std::integral_constant<int, 0>{}, f0, f1);
        } else {
          // the entries of one dz come in ascending window slots (the build scans the window's
          // bins in its row-major order), so those at or beyond wcap are the segment's tail
          int lo = f0, hi = f1;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)lrow[mid] < wcap) lo = mid + 1;
            else hi = mid;
          }
          front(std::integral_constant<int, 0>{}, f0, lo);
          front(std::integral_constant<int, 1>{}, lo, f1);
        }
      }
      // back entries of this dz (appended from slot icap - 1 downwards): tested first, the ones
      // within rc queued and run in batches of 64
      if (b1 > b0) {
        int qn = 0;
        auto drain = [&]() {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          const int take = qn < 64 ? qn : 64;
          const uint32_t qs = s_q[wv][lane];
          const uint32_t rest = lane + 64 < qn ? s_q[wv][lane + 64] : 0u;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          if (lane + 64 < qn) s_q[wv][lane] = rest;
          qn -= take;
          if (lane < take) {
            const uint32_t e = row[qs];
            terms(std::integral_constant<int, OVF ? 2 : 0>{}, e, lrow[qs], P.x[e_idx(e)], P.nrm[e_idx(e)]);
          }
        };
        for (int k0 = b0; k0 < b1; k0 += 64) {
          const int k2 = k0 + lane;
          bool in = false;
          uint32_t slot = 0;
          if (k2 < b1) {
            slot = (uint32_t)(icap - 1 - k2);
            const uint32_t e = row[slot];
            const double4 xj = P.x[e_idx(e)];
            double gx, gy, gz;
            ghost(xj, e, prd, gx, gy, gz);
            R r2;
            in = ilp_cut_test<R>(xi.x - gx, xi.y - gy, xi.z - gz, P.ilp_cutsq, r2);
          }
          const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
          const int below =
              (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
          if (in) s_q[wv][qn + below] = slot;
          qn += (int)__popcll(m);
          if (qn >= 64) drain();
        }
        if (qn > 0) drain();
      }
      if (LANE32) {
        fx += (double)fxl; fy += (double)fyl; fz += (double)fzl;
        sx += (double)sxl; sy += (double)syl; sz += (double)szl;
      }
      {
        const double v8[8] = {fx, fy, fz, sx, sy, sz, ev, er};
        put8(i, dz, wsum8(v8, lane), 0);
      }
      if (VF) {
        const double w8[8] = {wv6[0], wv6[1], wv6[2], wv6[3], wv6[4], wv6[5], 0.0, 0.0};
        put8(i, dz, wsum8(w8, lane), 8);
      }
    }
    // the window's partner sums go out to acc[] (integer adds: any order gives the same sums)
    __syncthreads();
    const int wsize = OVF ? min(s_bin[nbins], wcap) : s_bin[nbins];
    for (int p = threadIdx.x; p < wsize; p += blockDim.x) {
      unsigned long long vq[6];
      bool nz = false;
#pragma unroll
      for (int c = 0; c < 6; c++) {
        vq[c] = s_acc[c * wcap + p];
        nz |= vq[c] != 0ull;
        s_acc[c * wcap + p] = 0ull;
      }
      if (!nz) continue;
      // the bin of slot p: the last bin whose first slot is at or before p (it holds p)
      int lo = 0, hi = nbins - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_bin[mid] <= p) lo = mid;
        else hi = mid - 1;
      }
      const int atom = s_batom[lo] + (p - s_bin[lo]);
#pragma unroll
      for (int c = 0; c < 6; c++)
        if (vq[c] != 0ull)
          __hip_atomic_fetch_add(P.acc + (size_t)c * n + atom, vq[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
  // fixed-point range: |term| < 2^7 needs the coefficients below FX_COEF_MAX
  double cm = (double)cmax;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cm = fmax(cm, __shfl_xor(cm, o, 64));
  if (lane == 0 && !(cm < FX_COEF_MAX)) atomicOr(P.err, EB_FIXED);
}

// The ILP force, S and tallies of every atom of a half system: its pass sums in pass order plus
// the partner sums (which are then cleared for the next evaluation); with the virial, S is now
// whole, so the virial of the normal-derivative forces of the atom's normal is added here (as the
// pull kernel's tail adds it).
template <typename R, bool EF, bool VF>
__global__ void k_ilp_combine(DevPtrs<R> P) {
  constexpr int PSTR = half_pstr(EF, VF);
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= P.n) return;
  if (!P.syshalf[P.sys[a]]) return;
  const size_t n = (size_t)P.n;
  double v[PSTR];
#pragma unroll
  for (int c = 0; c < PSTR; c++) v[c] = P.ipart[(size_t)a * PSTR + c];
  long long q[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    q[c] = (long long)P.acc[c * n + a];
    P.acc[c * n + a] = 0ull;
  }
  const double4 f = {v[0] + (double)q[0] * FX_INV, v[1] + (double)q[1] * FX_INV, v[2] + (double)q[2] * FX_INV, 0.0};
  const double4 S4 = {v[3] + (double)q[3] * FX_INV, v[4] + (double)q[4] * FX_INV, v[5] + (double)q[5] * FX_INV, 0.0};
  P.fil[a] = f;
  P.S[a] = S4;
  if (EF || VF) {
    P.th[0 * n + a] = v[6];
    P.th[1 * n + a] = v[7];
  }
  if (VF) {
    const uint32_t *nrow = P.nnb + (size_t)a * 4;
    const int ci = (int)nrow[0];
    double w[6] = {v[8], v[9], v[10], v[11], v[12], v[13]};
    for (int kk = 0; kk < ci; kk++) {
      const R *D = P.dn + (size_t)a * 27 + kk * 9;
      const double *vt = P.vet + (size_t)a * 9 + kk * 3;
      double Fk[3];
      for (int c = 0; c < 3; c++)
        Fk[c] = -((double)D[0 * 3 + c] * S4.x + (double)D[1 * 3 + c] * S4.y + (double)D[2 * 3 + c] * S4.z);
      w[0] += vt[0] * Fk[0];
      w[1] += vt[1] * Fk[1];
      w[2] += vt[2] * Fk[2];
      w[3] += vt[0] * Fk[1];
      w[4] += vt[0] * Fk[2];
      w[5] += vt[1] * Fk[2];
    }
    for (int c = 0; c < 6; c++) P.th[(3 + c) * n + a] = w[c];
  }
}

// ---------------------------------------------------------------------------------------
// Tersoff functions (pair_tersoff.cpp:584-811, pair_tersoff.h:90-107)

template <typename R>
__device__ __forceinline__ R t_fc(R r, const TersP<R> &p) {
  if (r < p.bigr - p.bigd) return (R)1;
  if (r > p.bigr + p.bigd) return (R)0;
  return (R)0.5 * ((R)1 - MathOps<R>::sn((R)MY_PI2 * (r - p.bigr) / p.bigd));
}
template <typename R>
__device__ __forceinline__ R t_fc_d(R r, const TersP<R> &p) {
  if (r < p.bigr - p.bigd) return (R)0;
  if (r > p.bigr + p.bigd) return (R)0;
  return -((R)MY_PI4 / p.bigd) * MathOps<R>::cs((R)MY_PI2 * (r - p.bigr) / p.bigd);
}
template <typename R>
__device__ __forceinline__ R t_gijk(R costheta, const TersP<R> &p) {
  const R hcth = p.h - costheta;
  return p.gamma * ((R)1 + p.cc / p.dd - p.cc / (p.dd + hcth * hcth));
}
template <typename R>
__device__ __forceinline__ R t_gijk_d(R costheta, const TersP<R> &p) {
  const R hcth = p.h - costheta;
  const R numerator = (R)-2 * p.cc * hcth;
  const R denominator = (R)1 / (p.dd + hcth * hcth);
  return p.gamma * numerator * denominator * denominator;
}
template <typename R>
__device__ __forceinline__ R t_exdelr(R rij, R rik, const TersP<R> &p, R &ex_delr_d) {
  // lam3 = 0 (every entry of C.opt.tersoff and BNC.tersoff): the reference's exp(0) = 1 and
  // a zero derivative, without evaluating them
  if (p.lam3 == (R)0) {
    ex_delr_d = 0;
    return 1;
  }
  R arg;
  if (p.powermint == 3) {
    const R t = p.lam3 * (rij - rik);
    arg = t * t * t;
  } else {
    arg = p.lam3 * (rij - rik);
  }
  R ex;
  if (arg > (R)69.0776) ex = (R)1.e30;
  else if (arg < (R)-69.0776) ex = (R)0;
  else ex = exp(arg);
  if (p.powermint == 3) ex_delr_d = (R)3 * (p.lam3 * p.lam3 * p.lam3) * (rij - rik) * (rij - rik) * ex;
  else ex_delr_d = p.lam3 * ex;
  return ex;
}
template <typename R>
__device__ __forceinline__ R t_bij(R zeta, const TersP<R> &p) {
  using M = MathOps<R>;
  const R tmp = p.beta * zeta;
  if (tmp > p.c1) return (R)1 / M::sq(tmp);
  if (tmp > p.c2) return ((R)1 - M::pw(tmp, -p.powern) / ((R)2 * p.powern)) / M::sq(tmp);
  if (tmp < p.c4) return (R)1;
  if (tmp < p.c3) return (R)1 - M::pw(tmp, p.powern) / ((R)2 * p.powern);
  return M::pw((R)1 + M::pw(tmp, p.powern), (R)-1 / ((R)2 * p.powern));
}
template <typename R>
__device__ __forceinline__ R t_bij_d(R zeta, const TersP<R> &p) {
  using M = MathOps<R>;
  const R tmp = p.beta * zeta;
  if (tmp > p.c1) return p.beta * (R)-0.5 * M::pw(tmp, (R)-1.5);
  if (tmp > p.c2)
    return p.beta * ((R)-0.5 * M::pw(tmp, (R)-1.5) * ((R)1 - ((R)1 + (R)1 / ((R)2 * p.powern)) * M::pw(tmp, -p.powern)));
  if (tmp < p.c4) return (R)0;
  if (tmp < p.c3) return (R)-0.5 * p.beta * M::pw(tmp, p.powern - (R)1);
  const R tmp_n = M::pw(tmp, p.powern);
  return (R)-0.5 * M::pw((R)1 + tmp_n, (R)-1 - ((R)1 / ((R)2 * p.powern))) * tmp_n / zeta;
}

// b_ij and d b_ij / d zeta together (ters_bij + ters_bij_d, the same branches): in the
// common branch both share (1 + (beta zeta)^n), and pow(1 + t, -1 - 1/2n) is taken as
// pow(1 + t, -1/2n) / (1 + t)
template <typename R>
__device__ __forceinline__ void t_bij_both(R zeta, const TersP<R> &p, R &bij, R &bij_d) {
  using M = MathOps<R>;
  const R tmp = p.beta * zeta;
  if (tmp > p.c1 || tmp > p.c2 || tmp < p.c4 || tmp < p.c3) {
    bij = t_bij(zeta, p);
    bij_d = t_bij_d(zeta, p);
    return;
  }
  const R tmp_n = M::pw(tmp, p.powern);
  const R base = (R)1 + tmp_n;
  bij = M::pw(base, (R)-1 / ((R)2 * p.powern));
  bij_d = (R)-0.5 * (bij / base) * tmp_n / zeta;
}

// attractive(): forces on i, j, k of the zeta term of bond (i,j) through neighbour k
// (pair_tersoff.cpp:642-661 -> ters_zetaterm_d 738-795, costheta_d 797-811)
template <typename R>
__device__ __forceinline__ void t_attractive(const TersP<R> &p, R prefactor, R rij, R rik, const R *delrij,
                                             const R *delrik, R *fi, R *fj, R *fk) {
  const R rijinv = (R)1 / rij;
  const R rikinv = (R)1 / rik;
  R rij_hat[3], rik_hat[3];
  for (int d = 0; d < 3; d++) {
    rij_hat[d] = rijinv * delrij[d];
    rik_hat[d] = rikinv * delrik[d];
  }
  const R fc = t_fc(rik, p), dfc = t_fc_d(rik, p);
  R ex_delr_d;
  const R ex_delr = t_exdelr(rij, rik, p, ex_delr_d);
  const R cos_theta = rij_hat[0] * rik_hat[0] + rij_hat[1] * rik_hat[1] + rij_hat[2] * rik_hat[2];
  const R gijk = t_gijk(cos_theta, p), gijk_d = t_gijk_d(cos_theta, p);
  R dcosdri[3], dcosdrj[3], dcosdrk[3];
  for (int d = 0; d < 3; d++) {
    dcosdrj[d] = (-cos_theta * rij_hat[d] + rik_hat[d]) * ((R)1 / rij);
    dcosdrk[d] = (-cos_theta * rik_hat[d] + rij_hat[d]) * ((R)1 / rik);
    dcosdri[d] = -(dcosdrj[d] + dcosdrk[d]);
  }
  for (int d = 0; d < 3; d++) {
    R v = -dfc * gijk * ex_delr * rik_hat[d];
    v = fc * gijk_d * ex_delr * dcosdri[d] + v;
    v = fc * gijk * ex_delr_d * rik_hat[d] + v;
    v = -fc * gijk * ex_delr_d * rij_hat[d] + v;
    fi[d] = prefactor * v;
    R u = fc * gijk_d * ex_delr * dcosdrj[d];
    u = fc * gijk * ex_delr_d * rij_hat[d] + u;
    fj[d] = prefactor * u;
    R q = dfc * gijk * ex_delr * rik_hat[d];
    q = fc * gijk_d * ex_delr * dcosdrk[d] + q;
    q = -fc * gijk * ex_delr_d * rik_hat[d] + q;
    fk[d] = prefactor * q;
  }
}

// Entry of bond (i, jj)'s force on i's t-th Tersoff neighbour in fb (times 3) and wb (times
// 6): target-major with the atom's own bond count as the stride, so that all of one atom's
// entries are contiguous and the gather reads one run of ns entries per (i, target).
__device__ __forceinline__ size_t bond_slot(int i, int t, int jj, int ns) {
  return (size_t)i * (MAXSH * MAXSH) + (size_t)(t * ns + jj);
}

// displacement x_j(image) - x_i of a list entry, its square and j's type (the Tersoff kernels
// recompute these where needed instead of keeping per-neighbour arrays, which would live in
// scratch memory because their indices are not compile-time constants)
__device__ __forceinline__ void nb_del(const double4 *x, uint32_t e, const double *prd, const double4 &xi, double d[3],
                                       double &rsq, int &t) {
  const double4 xj = x[e_idx(e)];
  double gx, gy, gz;
  ghost(xj, e, prd, gx, gy, gz);
  d[0] = gx - xi.x;
  d[1] = gy - xi.y;
  d[2] = gz - xi.z;
  rsq = rsq_ref(d[0], d[1], d[2]);
  t = (int)xj.w - 1;
}

// Per step 3: the Tersoff bonds of i, one lane per bond (three lanes per atom, the bond count of
// an atom of a graphene or hBN sheet; lane q takes bonds q, q+3, ...; a wave holds 21 atoms and
// its last lane idles).  For bond (i, j): zeta_ij and b_ij (zeta 599, force_zeta 621-640), the
// three-body forces of every triplet (i, j, k) (attractive 642, ters_zetaterm_d 738-795), and
// the repulsive pair with i's half (584-597; pair_tersoff.cpp:141-262 for the loop structure).
// Every triplet is evaluated once.  What acts on i is summed over the atom's lanes into ft[i]; the
// force on j and on every k is stored per bond in fb[i][jj][kk] (kk == jj: the force on j),
// where the gather kernel picks it up.  The global virial of the bond is kept at i; per-atom
// virial shares (v_tally3 thirds, ev_tally halves) are stored in wb for the gather.
// With LDSP the type-triple index and the parameter records (at most TP_LDS of them) are
// staged in LDS, which shortens the chains of dependent loads that bound this kernel.  Three
// resident waves per SIMD (a few bytes of spill) measured 10 % faster than two.
constexpr int TP_LDS = 64;
// lanes per atom in the bonds kernel and the atoms a wave holds
constexpr int BOND_LANES = 3;
constexpr int BOND_ATOMS_PER_WAVE = 64 / BOND_LANES;

template <typename R, bool EF, bool VF, bool VA, bool LDSP>
__global__ __launch_bounds__(256, 3) void k_tersoff_bonds(DevPtrs<R> P) {
  using RT = double;
  using M = MathOps<RT>;
  __shared__ int s_t3[LDSP ? MAXT * MAXT * MAXT : 1];
  __shared__ TersP<double> s_tp[LDSP ? TP_LDS : 1];
  if (LDSP) {
    stage_table(s_t3, P.t3, MAXT * MAXT * MAXT);
    stage_table(s_tp, P.tp, P.ntp);
    __syncthreads();
  }
  const int *t3 = LDSP ? s_t3 : P.t3;
  const TersP<double> *tp = LDSP ? s_tp : P.tp;
  const int wl = threadIdx.x & 63;
  const int la = wl / BOND_LANES;               // atom within the wave
  const int lane = wl - BOND_LANES * la;        // lane within the atom
  const int i = ((xcd_block() * blockDim.x + threadIdx.x) >> 6) * BOND_ATOMS_PER_WAVE + la;
  if (la >= BOND_ATOMS_PER_WAVE || i >= P.n) return;  // all of an atom's lanes leave together
  const SysDev &s = P.sys_tab[P.sys[i]];
  const double prd[3] = {s.prd[0], s.prd[1], s.prd[2]};
  const double4 xi = P.x[i];
  const int ti = (int)xi.w - 1;
  const int ns = P.tcnt[i];
  const int itag = P.tag[i];
  const uint32_t *row = P.tsh + (size_t)i * MAXSH;
  const size_t n = (size_t)P.n;
  // the atom's lanes stage i's neighbours once (x_k - x_i, r^2, r, type per slot); both loops of
  // every bond read them from here instead of gathering and recomputing them per triplet
  __shared__ double s_nb[4 * BOND_ATOMS_PER_WAVE][MAXSH][6];
  double(*nbr)[6] = s_nb[(threadIdx.x >> 6) * BOND_ATOMS_PER_WAVE + la];
  for (int kk = lane; kk < ns; kk += BOND_LANES) {
    double d[3], rsq;
    int t;
    nb_del(P.x, row[kk], prd, xi, d, rsq, t);
    nbr[kk][0] = d[0];
    nbr[kk][1] = d[1];
    nbr[kk][2] = d[2];
    nbr[kk][3] = rsq;
    nbr[kk][4] = M::sq(rsq);
    nbr[kk][5] = (double)t;
  }
  // read by the other lanes of this atom only (one wave)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  double4 F = {0.0, 0.0, 0.0, 0.0};
  double et = 0.0;
  double vg[6] = {0, 0, 0, 0, 0, 0};  // global virial of i's bonds (VF)
  double vo[6] = {0, 0, 0, 0, 0, 0};  // i's own per-atom share (VA)
  const double third = 1.0 / 3.0;
  for (int jj = lane; jj < ns; jj += BOND_LANES) {
    const uint32_t ej = row[jj];
    const size_t slot = (size_t)i * MAXSH + jj;
    const double dj[3] = {nbr[jj][0], nbr[jj][1], nbr[jj][2]};
    const double rsqj = nbr[jj][3], rj = nbr[jj][4];
    const int tj = (int)nbr[jj][5];
    // ---- repulsion, i's half; the parameter is the one of whichever atom computes the pair
    //      in the reference (tag parity), so both halves agree ----
    {
      // the coordinates only decide between two images of the same atom; with equal
      // parameters in both orders the choice is moot and the tag is not loaded
      bool skip = false;
      if (!P.ters_rep_sym) {
        const int jtag = P.tag[e_idx(ej)];
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (jtag == itag) ghost(P.x[e_idx(ej)], ej, prd, gx, gy, gz);
        skip = tag_skip(itag, jtag, xi.x, xi.y, xi.z, gx, gy, gz);
      }
      const int pr = skip ? t3[(tj * MAXT + ti) * MAXT + ti] : t3[(ti * MAXT + tj) * MAXT + tj];
      const TersP<double> &p = tp[pr];
      if (rsqj < p.cutsq) {
        const double r = rj;
        const double tmp_fc = t_fc(r, p), tmp_fc_d = t_fc_d(r, p);
        const double tmp_exp = exp(-p.lam1 * r);
        const double fpair = -p.biga * tmp_exp * (tmp_fc_d - tmp_fc * p.lam1) / r;
        // dj is x_j - x_i; the reference's del is x_i - x_j
        F.x -= dj[0] * fpair;
        F.y -= dj[1] * fpair;
        F.z -= dj[2] * fpair;
        if (EF) et += 0.5 * (tmp_fc * p.biga * tmp_exp);
        if (VF || VA) {
          const double hf = 0.5 * fpair;
          const double w[6] = {hf * dj[0] * dj[0], hf * dj[1] * dj[1], hf * dj[2] * dj[2],
                               hf * dj[0] * dj[1], hf * dj[0] * dj[2], hf * dj[1] * dj[2]};
          for (int c = 0; c < 6; c++) { vg[c] += w[c]; vo[c] += w[c]; }
        }
      }
    }
    // ---- bond (i, j) ----
    const TersP<RT> &p = tp[t3[(ti * MAXT + tj) * MAXT + tj]];
    if (rsqj >= p.cutsq) {
      for (int kk = 0; kk < ns; kk++) {
        double *fv = P.fb + bond_slot(i, kk, jj, ns) * 3;
        fv[0] = 0.0; fv[1] = 0.0; fv[2] = 0.0;
        if (VA)
          for (int c = 0; c < 6; c++) P.wb[bond_slot(i, kk, jj, ns) * 6 + c] = 0.0;
      }
      if (EF) P.eb[slot] = 0.0;
      continue;
    }
    RT zeta = 0;
    const RT rij = rj;
    for (int kk = 0; kk < ns; kk++) {
      if (kk == jj) continue;
      const double dk[3] = {nbr[kk][0], nbr[kk][1], nbr[kk][2]};
      const double rsqk = nbr[kk][3];
      const int tk = (int)nbr[kk][5];
      const TersP<RT> &q = tp[t3[(ti * MAXT + tj) * MAXT + tk]];
      if (rsqk >= q.cutsq) continue;
      const RT rik = nbr[kk][4];
      const RT costheta = (dj[0] * dk[0] + dj[1] * dk[1] + dj[2] * dk[2]) / (rij * rik);
      RT dd;
      const RT ex = t_exdelr(rij, rik, q, dd);
      zeta += t_fc(rik, q) * t_gijk(costheta, q) * ex;
    }
    const RT r = rij;
    RT fa, fa_d;
    if (r > p.bigr + p.bigd) {
      fa = 0;
      fa_d = 0;
    } else {
      const RT e2 = exp(-p.lam2 * r);
      fa = -p.bigb * e2 * t_fc(r, p);
      fa_d = p.bigb * e2 * (p.lam2 * t_fc(r, p) - t_fc_d(r, p));
    }
    RT bij, bij_d;
    t_bij_both(zeta, p, bij, bij_d);
    const double fzv = (double)((RT)0.5 * bij * fa_d / r);
    const double pref = (double)((RT)-0.5 * fa * bij_d);
    if (EF) {
      const double eb = (double)((RT)0.5 * bij * fa);
      P.eb[slot] = eb;
      et += 0.5 * eb;
    }
    // force_zeta: +del*fpair on i, -del*fpair on j (ev_tally halves of the virial)
    F.x += dj[0] * fzv;
    F.y += dj[1] * fzv;
    F.z += dj[2] * fzv;
    double fjx = -dj[0] * fzv, fjy = -dj[1] * fzv, fjz = -dj[2] * fzv;
    double vj[6] = {0, 0, 0, 0, 0, 0};  // j's per-atom share (VA)
    if (VF || VA) {
      const double hf = -0.5 * fzv;
      const double w[6] = {hf * dj[0] * dj[0], hf * dj[1] * dj[1], hf * dj[2] * dj[2],
                           hf * dj[0] * dj[1], hf * dj[0] * dj[2], hf * dj[1] * dj[2]};
      for (int c = 0; c < 6; c++) { vg[c] += 2.0 * w[c]; vo[c] += w[c]; vj[c] = w[c]; }
    }
    for (int kk = 0; kk < ns; kk++) {
      if (kk == jj) continue;
      const double dk[3] = {nbr[kk][0], nbr[kk][1], nbr[kk][2]};
      const double rsqk = nbr[kk][3];
      const int tk = (int)nbr[kk][5];
      const TersP<double> &q = tp[t3[(ti * MAXT + tj) * MAXT + tk]];
      double *fv = P.fb + bond_slot(i, kk, jj, ns) * 3;
      if (rsqk >= q.cutsq) {
        fv[0] = 0.0; fv[1] = 0.0; fv[2] = 0.0;
        if (VA)
          for (int c = 0; c < 6; c++) P.wb[bond_slot(i, kk, jj, ns) * 6 + c] = 0.0;
        continue;
      }
      double fi[3], fj[3], fk[3];
      t_attractive(q, pref, rj, nbr[kk][4], dj, dk, fi, fj, fk);
      F.x += fi[0];
      F.y += fi[1];
      F.z += fi[2];
      fjx += fj[0];
      fjy += fj[1];
      fjz += fj[2];
      fv[0] = fk[0]; fv[1] = fk[1]; fv[2] = fk[2];
      if (VF || VA) {
        // v_tally3: delr1 (x) fj + delr2 (x) fk, a third to each of i, j, k
        const double w[6] = {dj[0] * fj[0] + dk[0] * fk[0], dj[1] * fj[1] + dk[1] * fk[1],
                             dj[2] * fj[2] + dk[2] * fk[2], dj[0] * fj[1] + dk[0] * fk[1],
                             dj[0] * fj[2] + dk[0] * fk[2], dj[1] * fj[2] + dk[1] * fk[2]};
        for (int c = 0; c < 6; c++) {
          vg[c] += w[c];
          const double t = third * w[c];
          vo[c] += t;
          vj[c] += t;
          if (VA) P.wb[bond_slot(i, kk, jj, ns) * 6 + c] = t;
        }
      }
    }
    double *fjv = P.fb + bond_slot(i, jj, jj, ns) * 3;
    fjv[0] = fjx; fjv[1] = fjy; fjv[2] = fjz;
    if (VA)
      for (int c = 0; c < 6; c++) P.wb[bond_slot(i, jj, jj, ns) * 6 + c] = vj[c];
  }
  // sum over the atom's lanes in lane order (the same in every lane)
  const int l0 = BOND_LANES * la;
  auto lsum = [&](double v) { return (__shfl(v, l0, 64) + __shfl(v, l0 + 1, 64)) + __shfl(v, l0 + 2, 64); };
  F.x = lsum(F.x);
  F.y = lsum(F.y);
  F.z = lsum(F.z);
  if (EF) et = lsum(et);
  if (VF)
    for (int c = 0; c < 6; c++) vg[c] = lsum(vg[c]);
  if (VA)
    for (int c = 0; c < 6; c++) vo[c] = lsum(vo[c]);
  if (lane != 0) return;
  F.w = 0.0;
  P.ft[i] = F;
  if (EF) P.th[2 * n + i] = et;
  if (VF)
    for (int c = 0; c < 6; c++) P.th[(3 + c) * n + i] += vg[c];
  if (VA)
    for (int c = 0; c < 6; c++) P.vatom[c * n + i] += vo[c];
}

// Per step 4: the force on atom a, by BOND_LANES lanes (21 atoms per wave, as in the bonds
// kernel).  Lane 0 adds the ILP pair sum and a's own Tersoff bonds (both from the kernels
// before) and a's own d n_a / d x_a term.  The atom's lanes then walk one list of items, lane q
// taking items q, q+3, ...: first a's reverse Tersoff
// partners i (what i's bonds left for a: fb[i][slot of a][jj] over jj, a as j when jj is a's
// own slot, as k otherwise), then a's reverse normal candidates m (the normal-derivative force
// on a as a normal neighbour of m).  Every item is a short chain of dependent loads; spreading
// them over the lanes keeps several chains in flight per atom.  The lanes are summed in lane
// order.  With INT (the Verlet loop), lane 0 then applies post_force
// (Langevin) and final_integrate to a with the same arithmetic as k_final_integrate, which
// saves a launch and a re-read of f and v per step.  In that loop k_initial_integrate has
// already advanced the step counter, so the noise is keyed by step - 1, the step
// k_final_integrate sees before k_advance_step.
__device__ __forceinline__ void philox4x32(uint32_t ctr[4], uint32_t key[2]);
// Langevin noise: Philox4x32-10 with counter {step low, step high, atom tag, word}, key = the
// system's 64-bit seed; the per-step draws and the run set-up draw use different words
constexpr uint32_t PHILOX_STEP_WORD = 0x4c4d4646u;   // "LMFF"
constexpr uint32_t PHILOX_SETUP_WORD = 0x53455455u;  // "SETU"

template <typename R, bool EF, bool VA, bool INT>
__device__ __forceinline__ void gather_atom(const DevPtrs<R> &P, int a, int lane, int want_ke) {
  const size_t n = (size_t)P.n;
  double4 F = {0.0, 0.0, 0.0, 0.0};
  double et = 0.0;
  double va[6] = {0, 0, 0, 0, 0, 0};   // Tersoff shares of i's bonds (per-atom only)
  double vk[6] = {0, 0, 0, 0, 0, 0};   // ILP k-term halves landing on a (per-atom only)
  const int rt = P.rtcnt[a], rn = P.rncnt[a];
  const double *prd = P.sys_tab[P.sys[a]].prd;  // a's reverse candidates share a's system
  if (lane == 0) {
    const double4 f0 = P.fil[a], f1 = P.ft[a];
    F = {f0.x + f1.x, f0.y + f1.y, f0.z + f1.z, 0.0};
    // ---- ILP: d n_a / d x_a term (2-neighbour normals only) ----
    const uint32_t *na = P.nnb + (size_t)a * 4;
    if (na[0] == 2) {
      const double4 Sa = P.S[a];
      const R *Dri = P.dndri + (size_t)a * 9;
      for (int c = 0; c < 3; c++) {
        const double v = (double)Dri[0 * 3 + c] * Sa.x + (double)Dri[1 * 3 + c] * Sa.y + (double)Dri[2 * 3 + c] * Sa.z;
        (&F.x)[c] -= v;
      }
    }
  }
  for (int q = lane; q < rt + rn; q += BOND_LANES) {
    if (q < rt) {
      // ---- Tersoff: what the bonds of i left for a ----
      const int i = (int)P.rtc[(size_t)a * MAXSH + q];
      const int ni = P.tcnt[i];
      const uint32_t *irow = P.tsh + (size_t)i * MAXSH;
      int slot_a = -1;
      for (int b = 0; b < ni; b++)
        if (e_idx(irow[b]) == a) { slot_a = b; break; }
      if (slot_a < 0) continue;
      const double *fv = P.fb + bond_slot(i, slot_a, 0, ni) * 3;
      for (int jj = 0; jj < ni; jj++) {
        F.x += fv[3 * jj + 0];
        F.y += fv[3 * jj + 1];
        F.z += fv[3 * jj + 2];
        if (VA)
          for (int c = 0; c < 6; c++) va[c] += P.wb[bond_slot(i, slot_a, jj, ni) * 6 + c];
      }
      if (EF) et += 0.5 * P.eb[(size_t)i * MAXSH + slot_a];
    } else {
      // ---- ILP: a as a normal neighbour of m ----
      const int m = (int)P.rnc[(size_t)a * MAXSH + (q - rt)];
      const uint4 nm = *(const uint4 *)(P.nnb + (size_t)m * 4);
      const int cm = (int)nm.x;
      if (cm < 2) continue;  // the normal is z and does not move with its neighbour
      const int kk = e_idx(nm.y) == a ? 0 : e_idx(nm.z) == a ? 1 : (cm == 3 && e_idx(nm.w) == a) ? 2 : -1;
      if (kk < 0) continue;
      {
        // F_a = -(d n_m / d x_a)^T S_m = -w x P / |n1_m| with P = S_m - (n_m . S_m) n_m and w
        // the vector of normal_dslot_w from m's other normal neighbours (nrm.w holds |n1_m|)
        const auto n4 = P.nrm[m];
        const double4 Sm = P.S[m];
        double w0, w1, w2;
        if (cm == 3) {
          // w = (x_p - x_q) / 3 over m's two other neighbours (x_m cancels)
          const uint32_t ep = kk == 0 ? nm.z : kk == 1 ? nm.w : nm.y;
          const uint32_t eq = kk == 0 ? nm.w : kk == 1 ? nm.y : nm.z;
          double px_, py_, pz_, qx_, qy_, qz_;
          ghost(P.x[e_idx(ep)], ep, prd, px_, py_, pz_);
          ghost(P.x[e_idx(eq)], eq, prd, qx_, qy_, qz_);
          const R third = (R)1 / (R)3;
          w0 = (double)((R)(px_ - qx_) * third);
          w1 = (double)((R)(py_ - qy_) * third);
          w2 = (double)((R)(pz_ - qz_) * third);
        } else {
          // w = v_other (kk = 0) or -v_other (kk = 1), v = x_other - x_m
          const uint32_t eo = kk == 0 ? nm.z : nm.y;
          const double4 xm = P.x[m];
          double ox, oy, oz;
          ghost(P.x[e_idx(eo)], eo, prd, ox, oy, oz);
          const double sg = kk == 0 ? 1.0 : -1.0;
          w0 = sg * (double)(R)(ox - xm.x);
          w1 = sg * (double)(R)(oy - xm.y);
          w2 = sg * (double)(R)(oz - xm.z);
        }
        const double nx = n4.x, ny = n4.y, nz = n4.z;
        const double ns = nx * Sm.x + ny * Sm.y + nz * Sm.z;
        const double px = Sm.x - ns * nx, py = Sm.y - ns * ny, pz = Sm.z - ns * nz;
        const double rl = 1.0 / (double)n4.w;
        double Fk[3];
        Fk[0] = -(w1 * pz - w2 * py) * rl;
        Fk[1] = -(w2 * px - w0 * pz) * rl;
        Fk[2] = -(w0 * py - w1 * px) * rl;
        F.x += Fk[0];
        F.y += Fk[1];
        F.z += Fk[2];
        if (VA) {
          // 0.5 * (vet_k(m) (x) F_k - T_m D_k), D_k as the normals kernel stored it
          const R *D = P.dn + (size_t)m * 27 + kk * 9;
          const double *vt = P.vet + (size_t)m * 9 + kk * 3;
          const double *Tm = P.T + (size_t)m * 6;
          const double Tf[3][3] = {{Tm[0], Tm[3], Tm[4]}, {Tm[3], Tm[1], Tm[5]}, {Tm[4], Tm[5], Tm[2]}};
          double TD[3][3];
          for (int r0 = 0; r0 < 3; r0++)
            for (int c = 0; c < 3; c++)
              TD[r0][c] = Tf[r0][0] * (double)D[0 * 3 + c] + Tf[r0][1] * (double)D[1 * 3 + c] + Tf[r0][2] * (double)D[2 * 3 + c];
          vk[0] += 0.5 * (vt[0] * Fk[0] - TD[0][0]);
          vk[1] += 0.5 * (vt[1] * Fk[1] - TD[1][1]);
          vk[2] += 0.5 * (vt[2] * Fk[2] - TD[2][2]);
          vk[3] += 0.5 * (vt[0] * Fk[1] - TD[0][1]);
          vk[4] += 0.5 * (vt[0] * Fk[2] - TD[0][2]);
          vk[5] += 0.5 * (vt[1] * Fk[2] - TD[1][2]);
        }
      }
    }
  }
  // sum over the atom's lanes in lane order (the same in every lane)
  const int l0 = ((int)(threadIdx.x & 63) / BOND_LANES) * BOND_LANES;
  auto lsum = [&](double v) { return (__shfl(v, l0, 64) + __shfl(v, l0 + 1, 64)) + __shfl(v, l0 + 2, 64); };
  F.x = lsum(F.x);
  F.y = lsum(F.y);
  F.z = lsum(F.z);
  if (EF) et = lsum(et);
  if (VA)
    for (int c = 0; c < 6; c++) {
      va[c] = lsum(va[c]);
      vk[c] = lsum(vk[c]);
    }
  if (lane != 0) return;
  F.w = 0.0;
  if (INT) {
    // post_force (Langevin) + final_integrate, as k_final_integrate does them
    const SysDev &s = P.sys_tab[P.sys[a]];
    const double m = P.rmass[(int)P.x[a].w - 1];
    double4 vp = P.v[a];
    if (s.style == 1) {
      const double gamma1 = -m / s.damp / FTM2V;
      const double gamma2 = sqrt(m) * sqrt(24.0 * BOLTZ / s.damp / s.dtv / MVV2E) / FTM2V * sqrt(s.t_target);
      const long long step = s.step - 1;
      uint32_t ctr[4] = {(uint32_t)step, (uint32_t)((unsigned long long)step >> 32), (uint32_t)P.tag[a], PHILOX_STEP_WORD};
      uint32_t key[2] = {(uint32_t)s.seed, (uint32_t)(s.seed >> 32)};
      philox4x32(ctr, key);
      const double u0 = (ctr[0] + 0.5) * 2.3283064365386963e-10 - 0.5;
      const double u1 = (ctr[1] + 0.5) * 2.3283064365386963e-10 - 0.5;
      const double u2 = (ctr[2] + 0.5) * 2.3283064365386963e-10 - 0.5;
      F.x += gamma1 * vp.x + gamma2 * u0;
      F.y += gamma1 * vp.y + gamma2 * u1;
      F.z += gamma1 * vp.z + gamma2 * u2;
    }
    const double dtfm = s.dtf / m;
    vp.x += dtfm * F.x;
    vp.y += dtfm * F.y;
    vp.z += dtfm * F.z;
    P.v[a] = vp;
    if (want_ke) P.th[9 * n + a] = 0.5 * MVV2E * m * (vp.x * vp.x + vp.y * vp.y + vp.z * vp.z);
  }
  P.f[a] = F;
  if (EF) P.th[2 * n + a] += et;
  if (VA)
    for (int c = 0; c < 6; c++) P.vatom[c * n + a] += va[c] + vk[c];
}

template <typename R, bool EF, bool VA, bool INT>
__global__ __launch_bounds__(256) void k_gather(DevPtrs<R> P, int want_ke) {
  const int wl = threadIdx.x & 63;
  const int la = wl / BOND_LANES;
  const int a = ((xcd_block() * blockDim.x + threadIdx.x) >> 6) * BOND_ATOMS_PER_WAVE + la;
  // all of an atom's lanes or none of them
  if (la < BOND_ATOMS_PER_WAVE && a < P.n) gather_atom<R, EF, VA, INT>(P, a, wl - BOND_LANES * la, want_ke);
}

// ---------------------------------------------------------------------------------------
// Integration (FixNVE, LMFF/fix_nve.cpp:64-143) and Langevin (FixLangevin::post_force,
// LMFF/fix_langevin.cpp:771-900 with the default uniform noise)

__device__ __forceinline__ void philox4x32(uint32_t ctr[4], uint32_t key[2]) {
  // Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ ctr[1] ^ key[0];
    const uint32_t n2 = hi0 ^ ctr[3] ^ key[1];
    ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
    key[0] += 0x9E3779B9u;
    key[1] += 0xBB67AE85u;
  }
}

// initial_integrate; also raises the far flag (sticky until the next build) when an atom is now
// more than half the ILP split margin away from its position at the last build: only then can a
// pair from the back segment of an ILP row be within the cutoff
// advance_nsys > 0: block 0 also advances every system's step counter (the fused
// force-gather integration of this step then keys its noise by step - 1); nothing in this
// kernel reads the counter
__global__ void k_initial_integrate(int n, double4 *x, double4 *v, const double4 *f, const int *sys,
                                    SysDev *st, const double *rmass_type, const double4 *xhold,
                                    double far_sq, int *farflag, int advance_nsys) {
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < advance_nsys; k += blockDim.x) st[k].step++;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SysDev &s = st[sys[p]];
  double4 xp = x[p], vp = v[p];
  const double4 fp = f[p];
  const double dtfm = s.dtf / rmass_type[(int)xp.w - 1];
  vp.x += dtfm * fp.x;
  vp.y += dtfm * fp.y;
  vp.z += dtfm * fp.z;
  xp.x += s.dtv * vp.x;
  xp.y += s.dtv * vp.y;
  xp.z += s.dtv * vp.z;
  x[p] = xp;
  v[p] = vp;
  const double4 h = xhold[p];
  const double dx = xp.x - h.x, dy = xp.y - h.y, dz = xp.z - h.z;
  const unsigned long long far = __ballot(dx * dx + dy * dy + dz * dz > far_sq);
  // one atomic per wave, none once the flag is up
  if (far && (threadIdx.x & 63) == __ffsll((long long)far) - 1 && *(volatile int *)farflag == 0) atomicOr(farflag, 1);
}

// post_force (Langevin) + final_integrate; ke column on thermo steps
__global__ void k_final_integrate(int n, const double4 *x, double4 *v, double4 *f, const int *sys, const int *tag,
                                  SysDev *st, const double *rmass_type, double *ke, int want_ke) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SysDev &s = st[sys[p]];
  const double4 xp = x[p];
  double4 vp = v[p], fp = f[p];
  const double m = rmass_type[(int)xp.w - 1];
  if (s.style == 1) {
    const double gamma1 = -m / s.damp / FTM2V;
    const double gamma2 = sqrt(m) * sqrt(24.0 * BOLTZ / s.damp / s.dtv / MVV2E) / FTM2V * sqrt(s.t_target);
    uint32_t ctr[4] = {(uint32_t)s.step, (uint32_t)((unsigned long long)s.step >> 32), (uint32_t)tag[p], PHILOX_STEP_WORD};
    uint32_t key[2] = {(uint32_t)s.seed, (uint32_t)(s.seed >> 32)};
    philox4x32(ctr, key);
    const double u0 = (ctr[0] + 0.5) * 2.3283064365386963e-10 - 0.5;
    const double u1 = (ctr[1] + 0.5) * 2.3283064365386963e-10 - 0.5;
    const double u2 = (ctr[2] + 0.5) * 2.3283064365386963e-10 - 0.5;
    fp.x += gamma1 * vp.x + gamma2 * u0;
    fp.y += gamma1 * vp.y + gamma2 * u1;
    fp.z += gamma1 * vp.z + gamma2 * u2;
    f[p] = fp;
  }
  const double dtfm = s.dtf / m;
  vp.x += dtfm * fp.x;
  vp.y += dtfm * fp.y;
  vp.z += dtfm * fp.z;
  v[p] = vp;
  if (want_ke) ke[p] = 0.5 * MVV2E * m * (vp.x * vp.x + vp.y * vp.y + vp.z * vp.z);
}

// FixLangevin::setup applies post_force to the set-up force (LMFF/fix_langevin.cpp:376-377), so
// the first half-kick of a run already carries drag and noise.  The set-up draw has its own
// counter word, keyed by the step counter at set-up and the atom tag.
__global__ void k_langevin_setup(int n, const double4 *x, const double4 *v, double4 *f, const int *sys,
                                 const int *tag, const SysDev *st, const double *rmass_type) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SysDev &s = st[sys[p]];
  if (s.style != 1) return;
  const double m = rmass_type[(int)x[p].w - 1];
  const double4 vp = v[p];
  double4 fp = f[p];
  const double gamma1 = -m / s.damp / FTM2V;
  const double gamma2 = sqrt(m) * sqrt(24.0 * BOLTZ / s.damp / s.dtv / MVV2E) / FTM2V * sqrt(s.t_target);
  uint32_t ctr[4] = {(uint32_t)s.step, (uint32_t)((unsigned long long)s.step >> 32), (uint32_t)tag[p],
                     PHILOX_SETUP_WORD};
  uint32_t key[2] = {(uint32_t)s.seed, (uint32_t)(s.seed >> 32)};
  philox4x32(ctr, key);
  const double u0 = (ctr[0] + 0.5) * 2.3283064365386963e-10 - 0.5;
  const double u1 = (ctr[1] + 0.5) * 2.3283064365386963e-10 - 0.5;
  const double u2 = (ctr[2] + 0.5) * 2.3283064365386963e-10 - 0.5;
  fp.x += gamma1 * vp.x + gamma2 * u0;
  fp.y += gamma1 * vp.y + gamma2 * u1;
  fp.z += gamma1 * vp.z + gamma2 * u2;
  f[p] = fp;
}

__global__ void k_advance_step(int nsys, SysDev *st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) st[s].step++;
}

// ---------------------------------------------------------------------------------------
// Deterministic per-system reductions (Neumaier-compensated FP64): stage 1 over fixed chunks
// that never straddle a system, stage 2 one block per system in chunk order.

struct Chunk {
  int begin, end, sys, pad;
};

// chunks of a single system of n atoms: [k chunk, min((k+1) chunk, n)), system 0
__global__ void k_one_system_chunks(int n, int chunk, Chunk *chunks, int nchunks, int *cbeg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nchunks) chunks[k] = Chunk{k * chunk, min((k + 1) * chunk, n), 0, 0};
  if (k == 0) {
    cbeg[0] = 0;
    cbeg[1] = nchunks;
  }
}

__device__ __forceinline__ void neumaier(double &sum, double &comp, double v) {
  const double t = sum + v;
  if (fabs(sum) >= fabs(v)) comp += (sum - t) + v;
  else comp += (v - t) + sum;
  sum = t;
}

// columns c >= ncols - nmax are reduced by max (non-negative values) instead of summed
__global__ __launch_bounds__(256) void k_reduce_chunks(const double *cols, size_t stride, int ncols,
                                                       const Chunk *chunks, double2 *partial, int nmax) {
  __shared__ double ss[256], sc[256];
  const Chunk ch = chunks[blockIdx.x];
  for (int c = 0; c < ncols; c++) {
    if (c >= ncols - nmax) {
      double m = 0;
      for (int p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) m = fmax(m, cols[c * stride + p]);
      ss[threadIdx.x] = m;
      __syncthreads();
      for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) ss[threadIdx.x] = fmax(ss[threadIdx.x], ss[threadIdx.x + o]);
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[(size_t)blockIdx.x * ncols + c] = make_double2(ss[0], 0.0);
      __syncthreads();
      continue;
    }
    double sum = 0, comp = 0;
    for (int p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) neumaier(sum, comp, cols[c * stride + p]);
    ss[threadIdx.x] = sum;
    sc[threadIdx.x] = comp;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        double s2 = ss[threadIdx.x], c2 = sc[threadIdx.x] + sc[threadIdx.x + o];
        neumaier(s2, c2, ss[threadIdx.x + o]);
        ss[threadIdx.x] = s2;
        sc[threadIdx.x] = c2;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * ncols + c] = make_double2(ss[0], sc[0]);
    __syncthreads();
  }
}

// out[s * ncols + c] = sum over the chunks of system s (chunk ranges [cbeg[s], cbeg[s+1]))
__global__ void k_reduce_systems(const double2 *partial, int ncols, const int *cbeg, int nsys, double *out, int nmax) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
    if (c >= ncols - nmax) {
      double m = 0;
      for (int k = cbeg[s]; k < cbeg[s + 1]; k++) m = fmax(m, partial[(size_t)k * ncols + c].x);
      out[(size_t)s * ncols + c] = m;
      continue;
    }
    double sum = 0, comp = 0;
    for (int k = cbeg[s]; k < cbeg[s + 1]; k++) {
      const double2 v = partial[(size_t)k * ncols + c];
      neumaier(sum, comp, v.x);
      comp += v.y;
    }
    out[(size_t)s * ncols + c] = sum + comp;
  }
}

// ---------------------------------------------------------------------------------------
// FIRE minimisation (MinFire::iterate, LMFF/min_fire.cpp:104-460, semi-implicit Euler
// integrator with the half step back).  The scalar decisions are made per system on the host
// from the statistics below; the per-atom update is one kernel.

struct FireStep {
  double dtv, dtf;         // position and velocity step (dtf = dtv * ftm2v)
  double scale1, scale2;   // velocity mixing v = scale1 v + scale2 f (when mix)
  double back;             // x -= back * v before v is zeroed (half step back)
  int mix, zero, active, pad;
};

__global__ void k_zero_v(int n, double4 *v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) v[p] = {0.0, 0.0, 0.0, 0.0};
}

// per atom: v.f, v.v, f.f (summed per system) and max|v_c|, |f|^2, max|f_c| (maxed per system)
__global__ void k_fire_stats(int n, const double4 *v, const double4 *f, double *cols) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double4 a = v[p], b = f[p];
  const double ff = b.x * b.x + b.y * b.y + b.z * b.z;
  cols[0 * (size_t)n + p] = a.x * b.x + a.y * b.y + a.z * b.z;
  cols[1 * (size_t)n + p] = a.x * a.x + a.y * a.y + a.z * a.z;
  cols[2 * (size_t)n + p] = ff;
  cols[3 * (size_t)n + p] = fmax(fmax(fabs(a.x), fabs(a.y)), fabs(a.z));
  cols[4 * (size_t)n + p] = ff;
  cols[5 * (size_t)n + p] = fmax(fmax(fabs(b.x), fabs(b.y)), fabs(b.z));
}

__global__ void k_fire_step(int n, double4 *x, double4 *v, const double4 *f, const int *sys, const FireStep *fs,
                            const double *rmass_type, const double4 *xhold, double far_sq, int *farflag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const FireStep q = fs[sys[p]];
  double4 xp = x[p], vp = v[p];
  if (q.active) {
    const double4 fp = f[p];
    if (q.zero) {
      xp.x -= q.back * vp.x;
      xp.y -= q.back * vp.y;
      xp.z -= q.back * vp.z;
      vp.x = vp.y = vp.z = 0.0;
    }
    const double dtfm = q.dtf / rmass_type[(int)xp.w - 1];
    vp.x += dtfm * fp.x;
    vp.y += dtfm * fp.y;
    vp.z += dtfm * fp.z;
    if (q.mix) {
      vp.x = q.scale1 * vp.x + q.scale2 * fp.x;
      vp.y = q.scale1 * vp.y + q.scale2 * fp.y;
      vp.z = q.scale1 * vp.z + q.scale2 * fp.z;
    }
    xp.x += q.dtv * vp.x;
    xp.y += q.dtv * vp.y;
    xp.z += q.dtv * vp.z;
    x[p] = xp;
    v[p] = vp;
  }
  const double4 h = xhold[p];
  const double dx = xp.x - h.x, dy = xp.y - h.y, dz = xp.z - h.z;
  if (dx * dx + dy * dy + dz * dz > far_sq && *(volatile int *)farflag == 0) atomicOr(farflag, 1);
}

// ---------------------------------------------------------------------------------------
// Conjugate gradients with the quadratic line search (MinCG::iterate, LMFF/min_cg.cpp:40-190;
// MinLineSearch::linemin_quadratic and alpha_step, LMFF/min_linesearch.cpp:324-503,837-868).
// Per-atom vectors: x0 (start of the line search), g (last gradient), h (search direction).

// per atom: f.f, f.h, f.g (summed) and max|h_c|, |f|^2, max|f_c| (maxed)
__global__ void k_cg_stats(int n, const double4 *f, const double4 *g, const double4 *h, double *cols) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double4 a = f[p], b = g[p], c = h[p];
  const double ff = a.x * a.x + a.y * a.y + a.z * a.z;
  cols[0 * (size_t)n + p] = ff;
  cols[1 * (size_t)n + p] = a.x * c.x + a.y * c.y + a.z * c.z;
  cols[2 * (size_t)n + p] = a.x * b.x + a.y * b.y + a.z * b.z;
  cols[3 * (size_t)n + p] = fmax(fmax(fabs(c.x), fabs(c.y)), fabs(c.z));
  cols[4 * (size_t)n + p] = ff;
  cols[5 * (size_t)n + p] = fmax(fmax(fabs(a.x), fabs(a.y)), fabs(a.z));
}

// mode 0: h = g = f;  mode 1: g = f, h = g + beta h, cols <- g.h and max|h_c|;  mode 2: h = g
__global__ void k_cg_dir(int n, const double4 *f, double4 *g, double4 *h, double beta, int mode, double *cols) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (mode == 2) {
    h[p] = g[p];
    return;
  }
  const double4 a = f[p];
  double4 c = mode == 0 ? a : h[p];
  if (mode == 1) {
    c.x = a.x + beta * c.x;
    c.y = a.y + beta * c.y;
    c.z = a.z + beta * c.z;
    cols[0 * (size_t)n + p] = a.x * c.x + a.y * c.y + a.z * c.z;
    cols[1 * (size_t)n + p] = fmax(fmax(fabs(c.x), fabs(c.y)), fabs(c.z));
  }
  g[p] = a;
  h[p] = c;
}

// x = x0 + alpha h (x = x0 for alpha <= 0), keeping x.w (type); raises the far flag
__global__ void k_cg_alpha_step(int n, double4 *x, const double4 *x0, const double4 *h, double alpha,
                                const double4 *xhold, double far_sq, int *farflag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double4 xp = x0[p];
  xp.w = x[p].w;
  if (alpha > 0.0) {
    const double4 c = h[p];
    xp.x += alpha * c.x;
    xp.y += alpha * c.y;
    xp.z += alpha * c.z;
  }
  x[p] = xp;
  const double4 q = xhold[p];
  const double dx = xp.x - q.x, dy = xp.y - q.y, dz = xp.z - q.z;
  if (dx * dx + dy * dy + dz * dz > far_sq && *(volatile int *)farflag == 0) atomicOr(farflag, 1);
}

// after a rebuild wrapped x into the box: x0 follows by the same periodic image
// (FixMinimize::reset_coords, LMFF/fix_minimize.cpp)
__global__ void k_image_follow(int n, const double4 *x, double4 *x0, const int *sys, const SysDev *st) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SysDev &s = st[sys[p]];
  const double4 a = x[p];
  double4 b = x0[p];
  double *bb = &b.x;
  const double aa[3] = {a.x, a.y, a.z};
  for (int d = 0; d < 3; d++)
    if (s.per[d]) bb[d] += s.prd[d] * rint((aa[d] - bb[d]) / s.prd[d]);
  x0[p] = b;
}

// max squared displacement since the last build, per system (neigh_modify check yes)
__global__ void k_displacement(int n, const double4 *x, const double4 *xhold, const int *sys, double trigger_sq,
                               int *flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double dx = x[p].x - xhold[p].x, dy = x[p].y - xhold[p].y, dz = x[p].z - xhold[p].z;
  if (dx * dx + dy * dy + dz * dz > trigger_sq) atomicOr(flag, 1);
  (void)sys;
}

// counts: [0] ILP list entries, [1] ILP ordered pairs within rc, [2] Tersoff neighbours within
// cutmax, [3] short-list entries
template <typename R>
__global__ void k_pair_stats(DevPtrs<R> P, unsigned long long *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const SysDev &s = P.sys_tab[P.sys[i]];
  const double4 xi = P.x[i];
  const int nfront = P.icnt[i], nback = P.ocnt[i];
  const int cnt = nfront + nback, back_off = P.icap - nback - nfront;
  const uint32_t *row = P.ilist + (size_t)i * P.icap;
  // half rows hold each pair with dz > 0 once (counted for both of its ordered pairs) and the
  // pairs within one z bin from both atoms
  const bool half = P.syshalf && P.syshalf[P.sys[i]] && !P.ilp_only;
  const int f1 = half ? P.segf[(size_t)i * NSEG + 1] : 0, b1 = half ? P.segb[(size_t)i * NSEG + 1] : 0;
  unsigned long long np = 0, ne = 0;
  for (int k = 0; k < cnt; k++) {
    const uint32_t e = row[ilp_slot(k, nfront, back_off)];
    // back entries lie in memory in reverse order of their append index
    const unsigned w = (!half || (k < nfront ? k < f1 : (nback - 1 - (k - nfront)) < b1)) ? 1u : 2u;
    double gx, gy, gz;
    ghost(P.x[e_idx(e)], e, s.prd, gx, gy, gz);
    const double dx = xi.x - gx, dy = xi.y - gy, dz = xi.z - gz;
    ne += w;
    if (dx * dx + dy * dy + dz * dz < P.ilp_cutsq) np += w;
  }
  atomicAdd(&out[0], ne);
  atomicAdd(&out[1], np);
  atomicAdd(&out[2], (unsigned long long)P.tcnt[i]);
  atomicAdd(&out[3], (unsigned long long)((P.scnt[i] & 0xffff) + (P.scnt[i] >> 16)));
}

}  // namespace lmff
