// lmff_run: runs the subset of a LAMMPS input deck that drives the LMFF force path on one
// MI355X, so an unchanged deck of the reference's form
//   units metal / atom_style full / boundary / variable (index, equal) / read_data / mass /
//   replicate / pair_style hybrid/overlay tersoff|lmff ilp/graphene/hbn / pair_coeff /
//   neighbor / neigh_modify / timestep / velocity create (dist, mom, rot) | scale /
//   fix nve | fix langevin / thermo / run
// executes with the force field, the neighbor build and the integrator on the GPU.
// Output follows LAMMPS' log format (thermo table, "Loop time", "Performance:" as in
// LMFF/finish.cpp:124,162).  Also: compute pe/atom and stress/atom, dump custom (text,
// DumpCustom::header_item, LMFF/dump_custom.cpp:422-440; atoms sorted by id), dump atom
// (DumpAtom, LMFF/dump_atom.cpp; scaled or, with dump_modify scale no, plain), write_data
// (WriteData, LMFF/write_data.cpp:236-420, atom_style full with image flags and velocities),
// write_restart / read_restart / restart N file [file2] (a binary file of this runner's own
// layout) and reset_timestep.  Image flags count box crossings as Domain::pbc does.  Extra
// command: "lmff_precision fp64|mixed".
//
// usage: lmff_run -in deck [-var NAME VALUE ...]
#include <algorithm>
#include <array>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <cctype>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "lmff_hip.h"
#include "pair_lmff_hip.hpp"

namespace {

constexpr double BOLTZ = 8.617343e-5, MVV2E = 1.0364269e-4, NKTV2P = 1.6021765e6;

struct Dump {
  std::string id, file;
  std::string style = "custom";  // custom | atom
  bool scale = true;             // atom style: scaled coordinates (dump_modify scale)
  int64_t every = 1;
  std::vector<std::string> fields;
  std::string fmt_float = "%g";
  FILE *fp = nullptr;
  int64_t last = -1;  // a step is written once (a run's set-up repeats the previous run's end)
};

struct Data {
  int64_t n = 0;
  int ntypes = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  std::vector<double> x, v;
  std::vector<int> img;  // image flags (3 per atom) from the Atoms section, 0 when absent
  std::vector<int32_t> type;
  std::vector<int64_t> tag, mol;
  std::map<int, double> mass;
};

[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "ERROR: %s\n", m.c_str());
  std::exit(1);
}

std::vector<std::string> words(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> w;
  std::string t;
  while (is >> t) w.push_back(t);
  return w;
}

// read_data for atom_style full (LMFF/read_data.cpp:116): header, Masses, Atoms, Velocities
Data read_data(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("Cannot open file " + path);
  Data d;
  std::string line, section;
  std::getline(in, line);  // title
  std::map<int64_t, size_t> index;
  std::vector<std::array<double, 3>> vel;
  std::vector<int64_t> vel_id;
  while (std::getline(in, line)) {
    auto h = line.find('#');
    std::string body = h == std::string::npos ? line : line.substr(0, h);
    auto w = words(body);
    if (w.empty()) continue;
    if (w.size() >= 2 && w[1] == "atoms" && section.empty()) { d.n = std::stoll(w[0]); continue; }
    if (w.size() >= 3 && w[1] == "atom" && w[2] == "types") { d.ntypes = std::stoi(w[0]); continue; }
    if (w.size() == 4 && w[2].size() == 3 && w[2].substr(1) == "lo") {
      int k = w[2][0] - 'x';
      d.lo[k] = std::stod(w[0]);
      d.hi[k] = std::stod(w[1]);
      continue;
    }
    if (w[0] == "Masses" || w[0] == "Atoms" || w[0] == "Velocities" || w[0] == "Bonds" || w[0] == "Pair") {
      section = w[0];
      continue;
    }
    if (section == "Masses") d.mass[std::stoi(w[0])] = std::stod(w[1]);
    else if (section == "Atoms") {
      if (w.size() < 7) die("Incorrect atom format in data file");
      index[std::stoll(w[0])] = d.tag.size();
      d.tag.push_back(std::stoll(w[0]));
      d.mol.push_back(std::stoll(w[1]));
      d.type.push_back(std::stoi(w[2]));
      for (int k = 0; k < 3; k++) d.x.push_back(std::stod(w[4 + k]));
      for (int k = 0; k < 3; k++) d.img.push_back(w.size() >= 10 ? std::stoi(w[7 + k]) : 0);
    } else if (section == "Velocities") {
      vel_id.push_back(std::stoll(w[0]));
      vel.push_back({std::stod(w[1]), std::stod(w[2]), std::stod(w[3])});
    }
  }
  if ((int64_t)d.tag.size() != d.n) die("Did not assign all atoms correctly");
  d.v.assign(3 * d.n, 0.0);
  for (size_t k = 0; k < vel.size(); k++) {
    auto it = index.find(vel_id[k]);
    if (it == index.end()) die("Invalid atom ID in Velocities section of data file");
    for (int c = 0; c < 3; c++) d.v[3 * it->second + c] = vel[k][c];
  }
  return d;
}

// replicate nx ny nz (Replicate::command, LMFF/replicate.cpp): atoms are unwrapped with their
// image flags (137), the box grows to n times each side (212-214), copies are made in ix, iy, iz
// loop order (384-386), shifted by whole old box lengths, remapped into the new box (533-544),
// and renumbered with id + (iz*ny*nx + iy*nx + ix)*maxtag and, for molecule ids > 0,
// mol + (iz*ny*nx + iy*nx + ix)*maxmol (556-571; maxtag/maxmol at 83-100).
void replicate(Data &d, int nx, int ny, int nz, const int periodic[3]) {
  if (nx < 1 || ny < 1 || nz < 1) die("Illegal replicate command");
  const int nrep = nx * ny * nz;
  const double prd[3] = {d.hi[0] - d.lo[0], d.hi[1] - d.lo[1], d.hi[2] - d.lo[2]};
  int64_t maxtag = 0, maxmol = 0;
  for (int64_t i = 0; i < d.n; i++) {
    maxtag = std::max(maxtag, d.tag[i]);
    maxmol = std::max(maxmol, d.mol[i]);
  }
  if (maxtag + (int64_t)(nrep - 1) * d.n >= ((int64_t)1 << 53)) die("Replicated system atom IDs are too big");
  Data r = d;
  r.x.clear(); r.v.clear(); r.img.clear(); r.type.clear(); r.tag.clear(); r.mol.clear();
  const int nrep_k[3] = {nx, ny, nz};
  for (int k = 0; k < 3; k++) r.hi[k] = d.lo[k] + nrep_k[k] * prd[k];
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        const int64_t rep = (int64_t)iz * ny * nx + (int64_t)iy * nx + ix;
        const int shift[3] = {ix, iy, iz};
        for (int64_t i = 0; i < d.n; i++) {
          for (int k = 0; k < 3; k++) {
            const int im0 = d.img.size() == 3 * (size_t)d.n ? d.img[3 * i + k] : 0;
            double xk = d.x[3 * i + k] + im0 * prd[k] + shift[k] * prd[k];
            int im = 0;
            if (periodic[k]) {
              const double L = r.hi[k] - r.lo[k];
              while (xk < r.lo[k]) { xk += L; im--; }
              while (xk >= r.hi[k]) { xk -= L; im++; }
            }
            r.x.push_back(xk);
            r.img.push_back(im);
            r.v.push_back(d.v[3 * i + k]);
          }
          r.type.push_back(d.type[i]);
          r.tag.push_back(d.tag[i] + rep * maxtag);
          r.mol.push_back(d.mol[i] > 0 ? d.mol[i] + rep * maxmol : d.mol[i]);
        }
      }
  r.n = d.n * nrep;
  d = std::move(r);
}

// Equal-style variable formulas (Variable::evaluate, LMFF/variable.cpp:1185): numbers, PI,
// v_name, + - * / % ^ with LAMMPS' precedence (variable.cpp:108-116: unary minus binds
// tighter than ^, and ^ groups left to right), parentheses and the math functions sqrt exp
// ln log abs sin cos tan asin acos atan atan2 ceil floor round.
struct Formula {
  const std::string &s;
  std::function<double(const std::string &)> var;
  size_t p = 0;
  void skip() { while (p < s.size() && std::isspace((unsigned char)s[p])) p++; }
  [[noreturn]] void bad() { die("Invalid syntax in variable formula: " + s); }
  double add() {
    double a = mul();
    for (skip(); p < s.size() && (s[p] == '+' || s[p] == '-'); skip()) {
      const char op = s[p++];
      const double b = mul();
      a = op == '+' ? a + b : a - b;
    }
    return a;
  }
  double mul() {
    double a = pw();
    for (skip(); p < s.size() && (s[p] == '*' || s[p] == '/' || s[p] == '%'); skip()) {
      const char op = s[p++];
      const double b = pw();
      if (op == '*') a *= b;
      else if (b == 0.0) die("Divide by 0 in variable formula");
      else a = op == '/' ? a / b : std::fmod(a, b);
    }
    return a;
  }
  double pw() {
    double a = unary();
    for (skip(); p < s.size() && s[p] == '^'; skip()) {
      p++;
      const double b = unary();
      if (b == 0.0) a = 1.0;
      else if (a == 0.0 && b < 0.0) die("Invalid power expression in variable formula");
      else a = std::pow(a, b);
    }
    return a;
  }
  double unary() {
    skip();
    if (p < s.size() && s[p] == '-') { p++; return -unary(); }
    if (p < s.size() && s[p] == '+') { p++; return unary(); }
    return primary();
  }
  double primary() {
    skip();
    if (p >= s.size()) bad();
    if (s[p] == '(') {
      p++;
      const double a = add();
      skip();
      if (p >= s.size() || s[p] != ')') bad();
      p++;
      return a;
    }
    if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
      size_t q = p;
      while (q < s.size() && (std::isdigit((unsigned char)s[q]) || s[q] == '.')) q++;
      if (q < s.size() && (s[q] == 'e' || s[q] == 'E')) {
        q++;
        if (q < s.size() && (s[q] == '+' || s[q] == '-')) q++;
        while (q < s.size() && std::isdigit((unsigned char)s[q])) q++;
      }
      const double a = std::stod(s.substr(p, q - p));
      p = q;
      return a;
    }
    size_t q = p;
    while (q < s.size() && (std::isalnum((unsigned char)s[q]) || s[q] == '_')) q++;
    if (q == p) bad();
    const std::string word = s.substr(p, q - p);
    p = q;
    if (word.compare(0, 2, "v_") == 0) return var(word.substr(2));
    if (word == "PI") return 3.14159265358979323846;
    skip();
    if (p >= s.size() || s[p] != '(') die("Invalid thermo keyword '" + word + "' in variable formula");
    p++;
    const double a = add();
    skip();
    double b = 0.0;
    const bool two = p < s.size() && s[p] == ',';
    if (two) { p++; b = add(); skip(); }
    if (p >= s.size() || s[p] != ')') bad();
    p++;
    if (two && word == "atan2") return std::atan2(a, b);
    if (two) die("Invalid math function in variable formula: " + word);
    if (word == "sqrt") { if (a < 0.0) die("Sqrt of negative value in variable formula"); return std::sqrt(a); }
    if (word == "exp") return std::exp(a);
    if (word == "ln" || word == "log") {
      if (a <= 0.0) die("Log of zero/negative value in variable formula");
      return word == "ln" ? std::log(a) : std::log10(a);
    }
    if (word == "abs") return std::fabs(a);
    if (word == "sin") return std::sin(a);
    if (word == "cos") return std::cos(a);
    if (word == "tan") return std::tan(a);
    if (word == "asin") return std::asin(a);
    if (word == "acos") return std::acos(a);
    if (word == "atan") return std::atan(a);
    if (word == "ceil") return std::ceil(a);
    if (word == "floor") return std::floor(a);
    if (word == "round") return std::round(a);
    die("Invalid math function in variable formula: " + word);
  }
  double eval() {
    const double a = add();
    skip();
    if (p != s.size()) bad();
    return a;
  }
};

// write_data: header (WriteData::header), Masses, Atoms # full (AtomVecFull::write_data: wrapped
// positions and image flags), Velocities (AtomVec::write_vel)
void write_data(const std::string &path, const Data &d, int64_t step, const std::vector<double> &x,
                const std::vector<double> &v, const std::vector<int> &img) {
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) die("Cannot open data file " + path);
  std::fprintf(fp, "LAMMPS data file via write_data, lmff_run, timestep = %lld\n\n", (long long)step);
  std::fprintf(fp, "%lld atoms\n%d atom types\n\n", (long long)d.n, d.ntypes);
  const char *ax[3] = {"x", "y", "z"};
  for (int k = 0; k < 3; k++) std::fprintf(fp, "%-1.16e %-1.16e %slo %shi\n", d.lo[k], d.hi[k], ax[k], ax[k]);
  std::fprintf(fp, "\nMasses\n\n");
  for (int t = 1; t <= d.ntypes; t++) std::fprintf(fp, "%d %g\n", t, d.mass.count(t) ? d.mass.at(t) : 0.0);
  std::vector<int64_t> order(d.n);
  for (int64_t i = 0; i < d.n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return d.tag[a] < d.tag[b]; });
  std::fprintf(fp, "\nAtoms # full\n\n");
  for (int64_t i : order)
    std::fprintf(fp, "%lld %lld %d %-1.16e %-1.16e %-1.16e %-1.16e %d %d %d\n", (long long)d.tag[i],
                 (long long)d.mol[i], d.type[i], 0.0, x[3 * i], x[3 * i + 1], x[3 * i + 2], img[3 * i], img[3 * i + 1],
                 img[3 * i + 2]);
  std::fprintf(fp, "\nVelocities\n\n");
  for (int64_t i : order)
    std::fprintf(fp, "%lld %-1.16e %-1.16e %-1.16e\n", (long long)d.tag[i], v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  std::fclose(fp);
}

// Binary restart file of the device-resident state (the role of WriteRestart::write /
// ReadRestart::command, LMFF/write_restart.cpp:238, LMFF/read_restart.cpp:76; not LAMMPS' own
// binary layout): the timestep, dt, boundary, box, per-type masses and, for every atom in id
// order, id, molecule, type, image flags, wrapped position and velocity.  Little-endian, as
// written by this host; pair settings are not stored (tersoff and ilp/graphene/hbn keep none in
// LAMMPS restarts either, so a deck repeats pair_style and pair_coeff after read_restart).
constexpr char RESTART_MAGIC[8] = {'L', 'M', 'F', 'F', 'R', 'S', 'T', '1'};
constexpr char RESTART_END[8] = {'L', 'M', 'F', 'F', 'E', 'N', 'D', '1'};

struct Restart {
  Data data;
  int64_t step = 0;
  double dt = 0.001;
  int periodic[3] = {1, 1, 1};
  std::string bstr[3];
};

void write_restart(const std::string &path, const Data &d, int64_t step, double dt, const int periodic[3],
                   const std::string bstr[3], const std::vector<double> &x, const std::vector<double> &v,
                   const std::vector<int> &img) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) die("Cannot open restart file " + path);
  auto put = [&](const void *p, size_t nb) {
    if (std::fwrite(p, 1, nb, fp) != nb) die("Cannot write restart file " + path);
  };
  const int32_t version = 1;
  put(RESTART_MAGIC, 8);
  put(&version, 4);
  put(&step, 8);
  put(&dt, 8);
  for (int k = 0; k < 3; k++) {
    const int32_t pk = periodic[k];
    char b[4] = {0, 0, 0, 0};
    std::strncpy(b, bstr[k].c_str(), 3);
    put(&pk, 4);
    put(b, 4);
  }
  put(d.lo, sizeof d.lo);
  put(d.hi, sizeof d.hi);
  const int32_t nt = d.ntypes;
  put(&nt, 4);
  for (int t = 1; t <= d.ntypes; t++) {
    const double m = d.mass.count(t) ? d.mass.at(t) : std::nan("");
    put(&m, 8);
  }
  put(&d.n, 8);
  std::vector<int64_t> order(d.n);
  for (int64_t i = 0; i < d.n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return d.tag[a] < d.tag[b]; });
  for (int64_t i : order) {
    const int32_t ty = d.type[i];
    const int32_t im[3] = {img[3 * i], img[3 * i + 1], img[3 * i + 2]};
    put(&d.tag[i], 8);
    put(&d.mol[i], 8);
    put(&ty, 4);
    put(im, 12);
    put(&x[3 * i], 24);
    put(&v[3 * i], 24);
  }
  put(RESTART_END, 8);
  if (std::fclose(fp) != 0) die("Cannot write restart file " + path);
}

Restart read_restart(const std::string &path) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) die("Cannot open restart file " + path);
  auto get = [&](void *p, size_t nb) {
    if (std::fread(p, 1, nb, fp) != nb) die("Unexpected end of restart file " + path);
  };
  char magic[8];
  get(magic, 8);
  if (std::memcmp(magic, RESTART_MAGIC, 8) != 0) die("Invalid restart file " + path + " (not written by lmff_run)");
  int32_t version = 0;
  get(&version, 4);
  if (version != 1) die("Restart file " + path + " has an unsupported version");
  Restart r;
  get(&r.step, 8);
  get(&r.dt, 8);
  for (int k = 0; k < 3; k++) {
    int32_t pk = 0;
    char b[4];
    get(&pk, 4);
    get(b, 4);
    b[3] = 0;
    r.periodic[k] = pk ? 1 : 0;
    r.bstr[k] = b;
  }
  Data &d = r.data;
  get(d.lo, sizeof d.lo);
  get(d.hi, sizeof d.hi);
  int32_t nt = 0;
  get(&nt, 4);
  if (nt < 1 || nt > 64) die("Invalid number of atom types in restart file " + path);
  d.ntypes = nt;
  for (int t = 1; t <= nt; t++) {
    double m = 0;
    get(&m, 8);
    if (!std::isnan(m)) d.mass[t] = m;
  }
  get(&d.n, 8);
  if (d.n < 0) die("Invalid atom count in restart file " + path);
  d.x.resize(3 * d.n);
  d.v.resize(3 * d.n);
  d.img.resize(3 * d.n);
  d.type.resize(d.n);
  d.tag.resize(d.n);
  d.mol.resize(d.n);
  for (int64_t i = 0; i < d.n; i++) {
    int32_t ty = 0, im[3];
    get(&d.tag[i], 8);
    get(&d.mol[i], 8);
    get(&ty, 4);
    get(im, 12);
    get(&d.x[3 * i], 24);
    get(&d.v[3 * i], 24);
    if (ty < 1 || ty > nt) die("Invalid atom type in restart file " + path);
    d.type[i] = ty;
    for (int k = 0; k < 3; k++) d.img[3 * i + k] = im[k];
  }
  get(magic, 8);
  if (std::memcmp(magic, RESTART_END, 8) != 0) die("Restart file " + path + " is corrupted");
  std::fclose(fp);
  return r;
}

// one snapshot of dump custom: columns id mol type x y z vx vy vz fx fy fz and c_ID / c_ID[k] /
// c_ID[*] of compute pe/atom (eV) and stress/atom (compute_stress_atom.cpp:254-340: minus
// (m v_a v_b + per-atom virial) in pressure*volume units, order xx yy zz xy xz yz)
void write_dump(Dump &dp, const Data &d, int64_t step, const std::string &boundstr, const std::vector<double> &x,
                const std::vector<double> &v, const std::vector<double> &f, const std::vector<int> &img,
                const std::vector<double> &eatom, const std::vector<double> &vatom,
                const std::map<std::string, std::string> &computes) {
  if (!dp.fp) {
    dp.fp = std::fopen(dp.file.c_str(), "w");
    if (!dp.fp) die("Cannot open dump file " + dp.file);
  }
  std::vector<int64_t> order(d.n);
  for (int64_t i = 0; i < d.n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return d.tag[a] < d.tag[b]; });
  // dump atom (DumpAtom, LMFF/dump_atom.cpp:210-247,504-513): id type xs ys zs (scaled by the
  // box, "%d %d %g %g %g") or id type x y z with dump_modify scale no
  if (dp.style == "atom") {
    FILE *fp = dp.fp;
    std::fprintf(fp, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%lld\n", (long long)step, (long long)d.n);
    std::fprintf(fp, "ITEM: BOX BOUNDS %s\n", boundstr.c_str());
    for (int k = 0; k < 3; k++) std::fprintf(fp, "%-1.16e %-1.16e\n", d.lo[k], d.hi[k]);
    std::fprintf(fp, "ITEM: ATOMS %s\n", dp.scale ? "id type xs ys zs" : "id type x y z");
    double inv[3];
    for (int k = 0; k < 3; k++) inv[k] = 1.0 / (d.hi[k] - d.lo[k]);
    for (int64_t i : order) {
      double c[3];
      for (int k = 0; k < 3; k++) c[k] = dp.scale ? (x[3 * i + k] - d.lo[k]) * inv[k] : x[3 * i + k];
      std::fprintf(fp, "%lld %d %g %g %g\n", (long long)d.tag[i], d.type[i], c[0], c[1], c[2]);
    }
    std::fflush(fp);
    return;
  }
  // expand c_ID[*] into its six columns
  std::vector<std::string> cols;
  for (const auto &c : dp.fields) {
    if (c.size() > 5 && c.compare(c.size() - 3, 3, "[*]") == 0) {
      for (int k = 1; k <= 6; k++) cols.push_back(c.substr(0, c.size() - 3) + "[" + std::to_string(k) + "]");
    } else {
      cols.push_back(c);
    }
  }
  FILE *fp = dp.fp;
  std::fprintf(fp, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%lld\n", (long long)step, (long long)d.n);
  std::fprintf(fp, "ITEM: BOX BOUNDS %s\n", boundstr.c_str());
  for (int k = 0; k < 3; k++) std::fprintf(fp, "%-1.16e %-1.16e\n", d.lo[k], d.hi[k]);
  std::fprintf(fp, "ITEM: ATOMS");
  for (const auto &c : cols) std::fprintf(fp, " %s", c.c_str());
  std::fprintf(fp, "\n");
  for (int64_t i : order) {
    bool first = true;
    for (const auto &c : cols) {
      if (!first) std::fputc(' ', fp);
      first = false;
      if (c == "id") { std::fprintf(fp, "%lld", (long long)d.tag[i]); continue; }
      if (c == "mol") { std::fprintf(fp, "%lld", (long long)d.mol[i]); continue; }
      if (c == "type") { std::fprintf(fp, "%d", d.type[i]); continue; }
      if (c == "ix" || c == "iy" || c == "iz") { std::fprintf(fp, "%d", img[3 * i + (c[1] - 'x')]); continue; }
      double val = 0.0;
      if (c == "x" || c == "y" || c == "z") val = x[3 * i + (c[0] - 'x')];
      else if (c == "xu" || c == "yu" || c == "zu") {
        // unwrapped (DumpCustom::pack_xu, Domain::unmap): x + image * box length
        const int k = c[0] - 'x';
        val = x[3 * i + k] + img[3 * i + k] * (d.hi[k] - d.lo[k]);
      }
      else if (c == "vx" || c == "vy" || c == "vz") val = v[3 * i + (c[1] - 'x')];
      else if (c == "fx" || c == "fy" || c == "fz") val = f[3 * i + (c[1] - 'x')];
      else {
        const size_t br = c.find('[');
        const std::string id = c.substr(2, br == std::string::npos ? std::string::npos : br - 2);
        const std::string &style = computes.at(id);
        if (style == "pe/atom") {
          val = eatom[i];
        } else {
          const int k = br == std::string::npos ? 1 : std::stoi(c.substr(br + 1));
          static const int a[6] = {0, 1, 2, 0, 0, 1}, b[6] = {0, 1, 2, 1, 2, 2};
          const double m = d.mass.at(d.type[i]);
          val = -NKTV2P * (MVV2E * m * v[3 * i + a[k - 1]] * v[3 * i + b[k - 1]] + vatom[6 * i + (k - 1)]);
        }
      }
      std::fprintf(fp, dp.fmt_float.c_str(), val);
    }
    std::fputc('\n', fp);
  }
  std::fflush(fp);
}

}  // namespace

int run_deck(int argc, char **argv) {
  std::string deck;
  std::map<std::string, std::string> vars;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if ((a == "-in" || a == "-i") && i + 1 < argc) deck = argv[++i];
    else if ((a == "-var" || a == "-v") && i + 2 < argc) { vars[argv[i + 1]] = argv[i + 2]; i += 2; }
  }
  if (deck.empty()) die("usage: lmff_run -in deck [-var NAME VALUE]");
  std::ifstream in(deck);
  if (!in) die("Cannot open input script " + deck);

  Data data;
  bool have_data = false;
  int periodic[3] = {1, 1, 1};
  std::vector<std::string> pair_args;
  std::vector<std::vector<std::string>> coeffs;
  double skin = 2.0, dt = 0.001;
  int every = 1, delay = 10, check = 1, thermo_every = 0, precision = LMFF_FP64;
  double t_target = 0, damp = 1;
  uint64_t lseed = 1;
  std::map<std::string, std::string> fixes;  // fix ID -> "nve" | "langevin"
  std::unique_ptr<lmff::PairLMFFHip> pair;
  int sys = -1;
  int64_t step = 0;
  std::map<std::string, std::string> computes;  // id -> pe/atom | stress/atom
  std::vector<Dump> dumps;
  std::string bstr[3] = {"p", "p", "p"};
  int64_t restart_every = 0;  // restart N file [file2] (Output::restart): written at multiples of N
  std::vector<std::string> restart_files;
  int restart_toggle = 0;
  std::string min_style = "cg";
  lmff_fire_params fire;
  lmff_fire_defaults(&fire);

  std::map<std::string, std::string> formulas;  // equal-style variables
  int depth = 0;
  std::function<double(const std::string &)> value = [&](const std::string &name) -> double {
    if (vars.count(name)) return std::stod(vars[name]);
    if (!formulas.count(name)) die("Invalid variable name in variable formula: " + name);
    if (++depth > 64) die("Variable formula refers to itself: " + name);
    const double a = Formula{formulas[name], value}.eval();
    depth--;
    return a;
  };
  // ${name}: an index-style value as written, an equal-style one as its formula's value in
  // "%.15g" (Variable::retrieve, LMFF/variable.cpp:871)
  auto subst = [&](std::string s) {
    for (size_t p; (p = s.find("${")) != std::string::npos;) {
      size_t q = s.find('}', p);
      if (q == std::string::npos) die("Invalid variable name in " + s);
      std::string name = s.substr(p + 2, q - p - 2);
      std::string val;
      if (vars.count(name)) {
        val = vars[name];
      } else if (formulas.count(name)) {
        char b[64];
        std::snprintf(b, sizeof b, "%.15g", value(name));
        val = b;
      } else {
        die("Substitution for illegal variable " + name);
      }
      s = s.substr(0, p) + val + s.substr(q + 1);
    }
    return s;
  };
  auto setup_pair = [&]() {
    if (pair) return;
    if (!have_data) die("Pair style requires read_data first");
    pair.reset(new lmff::PairLMFFHip(data.ntypes));
    std::vector<char *> a;
    for (auto &s : pair_args) a.push_back(const_cast<char *>(s.c_str()));
    pair->settings((int)a.size(), a.data());
    for (auto &c : coeffs) {
      std::vector<char *> b;
      for (auto &s : c) b.push_back(const_cast<char *>(s.c_str()));
      pair->coeff((int)b.size(), b.data());
    }
    std::vector<double> m(data.ntypes);
    for (int t = 1; t <= data.ntypes; t++) {
      if (!data.mass.count(t)) die("Not all per-type masses are set");
      m[t - 1] = data.mass[t];
    }
    pair->init_style(skin, every, delay, check, m);
    pair->set_precision(precision);
    lmff_ctx *ctx = pair->context();
    if (lmff_system_add(ctx, data.n, data.lo, data.hi, periodic, data.type.data(), data.tag.data(), data.mol.data(),
                        data.x.data(), data.v.data(), &sys))
      die(lmff_last_error(ctx));
    if (step > 0 && lmff_system_set_step(ctx, sys, step)) die(lmff_last_error(ctx));
    if (data.img.size() == 3 * (size_t)data.n && data.n > 0 && lmff_system_set_image(ctx, sys, data.img.data()))
      die(lmff_last_error(ctx));
  };
  // the current positions, velocities and image flags: from the device once a system is
  // resident, from the data read otherwise
  auto current = [&](std::vector<double> &x, std::vector<double> &v, std::vector<int> &im) {
    x = data.x;
    v = data.v;
    im = data.img;
    im.resize(3 * data.n, 0);
    if (!pair) return;
    lmff_ctx *ctx = pair->context();
    if (lmff_system_get(ctx, sys, x.data(), v.data(), nullptr)) die(lmff_last_error(ctx));
    if (data.n > 0 && lmff_system_get_image(ctx, sys, im.data())) die(lmff_last_error(ctx));
  };
  auto restart_now = [&](const std::string &file) {
    std::vector<double> x, v;
    std::vector<int> im;
    current(x, v, im);
    std::string f = file;
    const size_t star = f.find('*');
    if (star != std::string::npos) f = f.substr(0, star) + std::to_string(step) + f.substr(star + 1);
    write_restart(f, data, step, dt, periodic, bstr, x, v, im);
  };

  std::string line;
  while (std::getline(in, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line.erase(h);
    line = subst(line);
    auto w = words(line);
    if (w.empty()) continue;
    const std::string &cmd = w[0];
    if (cmd == "units") {
      if (w.size() < 2 || w[1] != "metal") die("only units metal is supported");
    } else if (cmd == "atom_style") {
      if (w.size() < 2 || w[1] != "full") die("only atom_style full is supported");
    } else if (cmd == "dimension") {
      if (w.size() < 2 || w[1] != "3") die("only dimension 3 is supported");
    } else if (cmd == "boundary") {
      for (int k = 0; k < 3 && k + 1 < (int)w.size(); k++) {
        periodic[k] = w[k + 1] == "p" ? 1 : 0;
        bstr[k] = w[k + 1];
      }
    } else if (cmd == "variable") {
      // index: the first definition (or -var on the command line) wins; equal: redefinable
      if (w.size() < 4) die("Illegal variable command");
      if (w[2] == "equal") {
        if (vars.count(w[1])) die("Cannot redefine variable as a different style");
        std::string f;
        for (size_t k = 3; k < w.size(); k++) f += w[k];
        if (f.size() >= 2 && (f.front() == '"' || f.front() == '\'') && f.back() == f.front())
          f = f.substr(1, f.size() - 2);
        formulas[w[1]] = f;
      } else if (w[2] == "index") {
        if (!vars.count(w[1])) vars[w[1]] = w[3];
      } else {
        die("lmff_run supports variable styles index and equal");
      }
    } else if (cmd == "print") {
      // print "text" (Input::print): the substituted text, quotes removed
      std::string t = line.substr(line.find("print") + 5);
      const size_t a = t.find_first_not_of(" \t"), b = t.find_last_not_of(" \t\r");
      t = a == std::string::npos ? "" : t.substr(a, b - a + 1);
      if (t.size() >= 2 && (t.front() == '"' || t.front() == '\'') && t.back() == t.front()) t = t.substr(1, t.size() - 2);
      std::printf("%s\n", t.c_str());
    } else if (cmd == "replicate") {
      if (!have_data) die("Replicate command before simulation box is defined");
      if (sys >= 0) die("lmff_run supports replicate only before the first run or minimize");
      if (w.size() < 4) die("Illegal replicate command");
      replicate(data, std::stoi(w[1]), std::stoi(w[2]), std::stoi(w[3]), periodic);
      std::printf("  %lld atoms after replicate\n", (long long)data.n);
    } else if (cmd == "read_data") {
      data = read_data(w.at(1));
      have_data = true;
      std::printf("  %lld atoms\n", (long long)data.n);
    } else if (cmd == "mass") {
      data.mass[std::stoi(w.at(1))] = std::stod(w.at(2));
      if (pair) {
        // after the first run or minimize the library holds the masses: forward the change
        std::vector<double> m(data.ntypes);
        for (int t = 1; t <= data.ntypes; t++) m[t - 1] = data.mass.count(t) ? data.mass[t] : 0.0;
        if (lmff_masses(pair->context(), m.data())) die(lmff_last_error(pair->context()));
      }
    } else if (cmd == "pair_style") {
      if (w.size() < 5 || w[1] != "hybrid/overlay" || (w[2] != "tersoff" && w[2] != "lmff") ||
          w[3] != "ilp/graphene/hbn" || w.size() > 6)
        die("lmff_run supports pair_style hybrid/overlay tersoff|lmff ilp/graphene/hbn <cut> [tap]");
      if (pair) die("lmff_run cannot change pair_style after the first run or minimize");
      pair_args.assign(w.begin() + 4, w.end());
    } else if (cmd == "pair_coeff") {
      coeffs.push_back(std::vector<std::string>(w.begin() + 1, w.end()));
      if (pair) {
        std::vector<char *> b;
        for (auto &q : coeffs.back()) b.push_back(const_cast<char *>(q.c_str()));
        pair->coeff((int)b.size(), b.data());
      }
    } else if (cmd == "neighbor") {
      skin = std::stod(w.at(1));
      if (pair && lmff_neighbor(pair->context(), skin, every, delay, check)) die(lmff_last_error(pair->context()));
    } else if (cmd == "neigh_modify") {
      for (size_t k = 1; k + 1 < w.size(); k++) {
        if (w[k] == "every") every = std::stoi(w[k + 1]);
        else if (w[k] == "delay") delay = std::stoi(w[k + 1]);
        else if (w[k] == "check") check = w[k + 1] == "yes";
      }
      if (pair && lmff_neighbor(pair->context(), skin, every, delay, check)) die(lmff_last_error(pair->context()));
    } else if (cmd == "timestep") {
      dt = std::stod(w.at(1));
    } else if (cmd == "lmff_precision") {
      precision = w.at(1) == "mixed" ? LMFF_MIXED : LMFF_FP64;
      if (pair) pair->set_precision(precision);
    } else if (cmd == "velocity") {
      // velocity all create T seed [dist uniform|gaussian] [mom yes|no] [rot yes|no] and
      // velocity all scale T (Velocity::create / scale, LMFF/velocity.cpp:162-380, 579):
      // draws scaled by 1/sqrt(m), net momentum and angular momentum about the centre of
      // mass removed as asked (zero_momentum / zero_rotation, 755-830, on unwrapped
      // positions), then rescaled to T over 3N-3 degrees of freedom.  The stream is this
      // runner's own (mt19937_64), not LAMMPS' RanPark.
      if (w.size() < 4 || w[1] != "all" || (w[2] != "create" && w[2] != "scale"))
        die("lmff_run supports velocity all create T seed [keywords] and velocity all scale T");
      if (!have_data) die("Velocity command before simulation box is defined");
      if (data.n == 0) die("Cannot zero momentum of no atoms");
      if (pair) current(data.x, data.v, data.img);  // a resident system: its present state
      const double T = std::stod(w[3]);
      bool gaussian = false, mom = true, rot = false;
      const size_t kw0 = w[2] == "create" ? 5 : 4;
      if (w[2] == "create" && w.size() < 5) die("Illegal velocity command");
      for (size_t k = kw0; k < w.size(); k += 2) {
        if (k + 1 >= w.size()) die("Illegal velocity command");
        const std::string &key = w[k], &val = w[k + 1];
        if (key == "dist" && (val == "uniform" || val == "gaussian")) gaussian = val == "gaussian";
        else if (key == "mom" && (val == "yes" || val == "no")) mom = val == "yes";
        else if (key == "rot" && (val == "yes" || val == "no")) rot = val == "yes";
        else if (key == "loop" || key == "sum" || key == "units" || key == "bias" || key == "temp") continue;
        else die("Illegal velocity command: " + key + " " + val);
      }
      auto mass_of = [&](int64_t i) { return data.mass.at(data.type[i]); };
      if (w[2] == "create") {
        std::mt19937_64 rng(std::stoull(w[4]));
        std::normal_distribution<double> g(0.0, 1.0);
        std::uniform_real_distribution<double> u(-0.5, 0.5);
        for (int64_t i = 0; i < data.n; i++) {
          const double f = 1.0 / std::sqrt(mass_of(i));
          for (int c = 0; c < 3; c++) data.v[3 * i + c] = (gaussian ? g(rng) : u(rng)) * f;
        }
        double mt = 0, p[3] = {0, 0, 0};
        for (int64_t i = 0; i < data.n; i++) {
          mt += mass_of(i);
          for (int c = 0; c < 3; c++) p[c] += mass_of(i) * data.v[3 * i + c];
        }
        if (mom)
          for (int64_t i = 0; i < data.n; i++)
            for (int c = 0; c < 3; c++) data.v[3 * i + c] -= p[c] / mt;
        if (rot) {
          // omega = I^-1 L about the centre of mass of the unwrapped positions
          std::vector<double> xu(3 * data.n);
          for (int64_t i = 0; i < data.n; i++)
            for (int c = 0; c < 3; c++)
              xu[3 * i + c] = data.x[3 * i + c] +
                              (data.img.size() == 3 * (size_t)data.n ? data.img[3 * i + c] : 0) * (data.hi[c] - data.lo[c]);
          double xcm[3] = {0, 0, 0}, L[3] = {0, 0, 0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
          for (int64_t i = 0; i < data.n; i++)
            for (int c = 0; c < 3; c++) xcm[c] += mass_of(i) * xu[3 * i + c] / mt;
          for (int64_t i = 0; i < data.n; i++) {
            const double m = mass_of(i);
            const double d[3] = {xu[3 * i] - xcm[0], xu[3 * i + 1] - xcm[1], xu[3 * i + 2] - xcm[2]};
            const double *vv = &data.v[3 * i];
            L[0] += m * (d[1] * vv[2] - d[2] * vv[1]);
            L[1] += m * (d[2] * vv[0] - d[0] * vv[2]);
            L[2] += m * (d[0] * vv[1] - d[1] * vv[0]);
            I[0][0] += m * (d[1] * d[1] + d[2] * d[2]);
            I[1][1] += m * (d[0] * d[0] + d[2] * d[2]);
            I[2][2] += m * (d[0] * d[0] + d[1] * d[1]);
            I[0][1] -= m * d[0] * d[1];
            I[1][2] -= m * d[1] * d[2];
            I[0][2] -= m * d[0] * d[2];
          }
          I[1][0] = I[0][1];
          I[2][1] = I[1][2];
          I[2][0] = I[0][2];
          const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) -
                             I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                             I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
          if (det > 1e-6) {
            double inv[3][3];
            inv[0][0] = (I[1][1] * I[2][2] - I[1][2] * I[2][1]) / det;
            inv[0][1] = -(I[0][1] * I[2][2] - I[0][2] * I[2][1]) / det;
            inv[0][2] = (I[0][1] * I[1][2] - I[0][2] * I[1][1]) / det;
            inv[1][0] = -(I[1][0] * I[2][2] - I[1][2] * I[2][0]) / det;
            inv[1][1] = (I[0][0] * I[2][2] - I[0][2] * I[2][0]) / det;
            inv[1][2] = -(I[0][0] * I[1][2] - I[0][2] * I[1][0]) / det;
            inv[2][0] = (I[1][0] * I[2][1] - I[1][1] * I[2][0]) / det;
            inv[2][1] = -(I[0][0] * I[2][1] - I[0][1] * I[2][0]) / det;
            inv[2][2] = (I[0][0] * I[1][1] - I[0][1] * I[1][0]) / det;
            double om[3];
            for (int r = 0; r < 3; r++) om[r] = inv[r][0] * L[0] + inv[r][1] * L[1] + inv[r][2] * L[2];
            for (int64_t i = 0; i < data.n; i++) {
              const double d[3] = {xu[3 * i] - xcm[0], xu[3 * i + 1] - xcm[1], xu[3 * i + 2] - xcm[2]};
              data.v[3 * i] -= om[1] * d[2] - om[2] * d[1];
              data.v[3 * i + 1] -= om[2] * d[0] - om[0] * d[2];
              data.v[3 * i + 2] -= om[0] * d[1] - om[1] * d[0];
            }
          }
        }
      }
      double ke = 0;
      for (int64_t i = 0; i < data.n; i++)
        for (int c = 0; c < 3; c++) ke += 0.5 * MVV2E * mass_of(i) * data.v[3 * i + c] * data.v[3 * i + c];
      const double t_now = data.n > 1 ? 2 * ke / ((3.0 * data.n - 3) * BOLTZ) : 0.0;
      const double sc = t_now > 0 ? std::sqrt(T / t_now) : 0.0;
      for (auto &q : data.v) q *= sc;
      if (pair && lmff_system_set(pair->context(), sys, nullptr, data.v.data())) die(lmff_last_error(pair->context()));
    } else if (cmd == "fix") {
      // fix ID all nve | fix ID all langevin T T damp seed [keyword no ...] (FixNVE,
      // FixLangevin: LMFF/fix_nve.cpp:30-48, LMFF/fix_langevin.cpp:58-150).  The device
      // integrates every atom, so other groups, other fix styles and the keywords that change the
      // dynamics are refused rather than skipped.
      if (w.size() < 4) die("Illegal fix command");
      const std::string &id = w[1], &grp = w[2], &fstyle = w[3];
      if (fixes.count(id)) die("lmff_run cannot replace fix " + id + "; unfix it first");
      if (fstyle != "nve" && fstyle != "langevin")
        die("fix style " + fstyle + " is not supported by lmff_run (only fix nve and fix langevin)");
      if (grp != "all") die("lmff_run integrates group all only: fix " + id + " " + grp + " " + fstyle);
      for (const auto &f : fixes)
        if (f.second == fstyle) die("lmff_run supports one fix " + fstyle + " (fix " + f.first + " exists)");
      if (fstyle == "nve") {
        if (w.size() != 4) die("Illegal fix nve command");
      } else {
        if (w.size() < 8 || (w.size() - 8) % 2) die("Illegal fix langevin command");
        if (std::stod(w[4]) != std::stod(w[5])) die("lmff_run supports fix langevin with Tstart = Tstop");
        for (size_t k = 8; k < w.size(); k += 2) {
          const std::string &kw = w[k], &val = w[k + 1];
          const bool known = kw == "zero" || kw == "tally" || kw == "gjf" || kw == "omega" || kw == "angmom";
          if (!known) die("Illegal fix langevin command: " + kw);
          if (val != "no") die("fix langevin " + kw + " " + val + " is not supported by lmff_run");
        }
        t_target = std::stod(w[4]);
        damp = std::stod(w[6]);
        lseed = std::stoull(w[7]);
      }
      fixes[id] = fstyle;
    } else if (cmd == "unfix") {
      if (w.size() != 2) die("Illegal unfix command");
      if (!fixes.erase(w[1])) die("Could not find fix ID " + w[1] + " to delete");
    } else if (cmd == "thermo") {
      thermo_every = std::stoi(w.at(1));
    } else if (cmd == "run") {
      const int64_t nsteps = std::stoll(w.at(1));
      setup_pair();
      lmff_ctx *ctx = pair->context();
      bool has_nve = false, has_langevin = false;
      for (const auto &f : fixes) {
        has_nve = has_nve || f.second == "nve";
        has_langevin = has_langevin || f.second == "langevin";
      }
      // without fix nve LAMMPS moves no atom; the device loop always integrates, so such runs
      // are refused (run 0 only evaluates)
      if (nsteps > 0 && !has_nve) die("lmff_run needs fix ID all nve for a run of more than 0 steps");
      const int style = has_langevin ? LMFF_LANGEVIN : LMFF_NVE;
      if (lmff_system_integrator(ctx, sys, style, dt, t_target, damp, lseed)) die(lmff_last_error(ctx));
      if (lmff_neighbor_stats(ctx, 1, nullptr, nullptr)) die(lmff_last_error(ctx));  // Neighbor::init
      bool want_peratom = false;
      for (const auto &dp : dumps)
        for (const auto &c : dp.fields)
          if (c.rfind("c_", 0) == 0) want_peratom = true;
      std::string boundstr;
      for (int k = 0; k < 3; k++) boundstr += (k ? " " : "") + (bstr[k].size() == 1 ? bstr[k] + bstr[k] : bstr[k]);
      std::vector<double> hx(3 * data.n), hv(3 * data.n), hf(3 * data.n), he, hw;
      std::vector<int> him(3 * data.n, 0);
      auto snapshot = [&](int64_t at) {
        bool any = false;
        for (const auto &dp : dumps) any = any || (at % dp.every == 0 && dp.last != at);
        if (!any) return;
        if (lmff_system_get(ctx, sys, hx.data(), hv.data(), hf.data())) die(lmff_last_error(ctx));
        if (data.n > 0 && lmff_system_get_image(ctx, sys, him.data())) die(lmff_last_error(ctx));
        if (want_peratom) {
          he.assign(data.n, 0.0);
          hw.assign(6 * data.n, 0.0);
          if (lmff_system_peratom(ctx, sys, he.data(), hw.data())) die(lmff_last_error(ctx));
        }
        for (auto &dp : dumps)
          if (at % dp.every == 0 && dp.last != at) {
            write_dump(dp, data, at, boundstr, hx, hv, hf, him, he, hw, computes);
            dp.last = at;
          }
      };
      if (lmff_evaluate(ctx, want_peratom ? 3 : 1, want_peratom ? 6 : 2)) die(lmff_last_error(ctx));
      lmff_thermo t0;
      lmff_system_thermo(ctx, sys, &t0);
      double vol = 1;
      for (int k = 0; k < 3; k++) vol *= data.hi[k] - data.lo[k];
      auto press = [&](const lmff_thermo &t) {
        return ((3.0 * data.n - 3) * BOLTZ * t.temp + t.virial[0] + t.virial[1] + t.virial[2]) / 3.0 / vol * NKTV2P;
      };
      std::printf("Step Temp E_pair TotEng Press\n");
      std::printf("%8lld %12.6g %16.10g %16.10g %12.6g\n", (long long)step, t0.temp, t0.pe, t0.pe + t0.ke, press(t0));
      snapshot(step);
      if (lmff_peratom_on_last_step(ctx, want_peratom ? 1 : 0)) die(lmff_last_error(ctx));
      // the run is cut at thermo and dump steps; each piece is one lmff_run (device-resident)
      const int64_t te = thermo_every > 0 ? thermo_every : nsteps;
      const int64_t end = step + nsteps;
      double secs = 0;
      int64_t at = step;
      while (at < end) {
        int64_t next = std::min(end, (at / te + 1) * te);
        for (const auto &dp : dumps) next = std::min(next, (at / dp.every + 1) * dp.every);
        if (restart_every > 0) next = std::min(next, (at / restart_every + 1) * restart_every);
        const int64_t len = next - at;
        lmff_thermo t;
        auto c0 = std::chrono::steady_clock::now();
        if (lmff_run(ctx, len, len, 1, &t, 1)) die(lmff_last_error(ctx));
        secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
        at = next;
        if (at % te == 0 || at == end)
          std::printf("%8lld %12.6g %16.10g %16.10g %12.6g\n", (long long)at, t.temp, t.pe, t.pe + t.ke, press(t));
        snapshot(at);
        if (restart_every > 0 && at % restart_every == 0) {
          step = at;
          restart_now(restart_files[restart_files.size() == 2 ? restart_toggle : 0]);
          restart_toggle ^= 1;
        }
      }
      if (lmff_peratom_on_last_step(ctx, 0)) die(lmff_last_error(ctx));
      step = end;
      std::printf("Loop time of %g on 1 procs for %lld steps with %lld atoms\n\n", secs, (long long)nsteps,
                  (long long)data.n);
      const double tps = secs > 0 ? nsteps / secs : 0;
      std::printf("Performance: %.3f ns/day, %.3f hours/ns, %.3f timesteps/s\n", tps * dt * 86400.0 / 1000.0,
                  tps > 0 ? 1.0 / (tps * dt * 3.6) : 0.0, tps);
      std::printf("Performance: %.4g atom-steps/s\n", tps * data.n);
      int64_t nbuild = 0, ndanger = 0;
      if (lmff_neighbor_stats(ctx, 0, &nbuild, &ndanger)) die(lmff_last_error(ctx));
      // Finish's neighbor lines (LMFF/finish.cpp:814-819); builds counted on the device
      std::printf("Neighbor list builds = %lld\n", (long long)nbuild);
      if (check) std::printf("Dangerous builds = %lld\n", (long long)ndanger);
      else std::printf("Dangerous builds not checked\n");
    } else if (cmd == "compute") {
      if (w.size() >= 4 && (w[3] == "pe/atom" || w[3] == "stress/atom")) computes[w[1]] = w[3];
      else std::printf("WARNING: compute %s ignored by lmff_run\n", w.size() > 3 ? w[3].c_str() : "?");
    } else if (cmd == "dump") {
      if (w.size() < 6 || (w[3] != "custom" && w[3] != "atom"))
        die("lmff_run supports dump ID all custom N file fields... and dump ID all atom N file");
      if (w[2] != "all") die("lmff_run dumps group all only");
      Dump dp;
      dp.id = w[1];
      dp.style = w[3];
      dp.every = std::stoll(w[4]);
      if (dp.every <= 0) die("Illegal dump command");
      dp.file = w[5];
      if (dp.style == "atom" && w.size() > 6) die("Illegal dump atom command");
      for (size_t k = 6; k < w.size(); k++) {
        const std::string &c = w[k];
        static const char *plain[] = {"id", "mol", "type", "x", "y", "z", "xu", "yu", "zu", "ix", "iy", "iz",
                                      "vx", "vy", "vz", "fx", "fy", "fz"};
        bool ok = std::find_if(std::begin(plain), std::end(plain), [&](const char *p) { return c == p; }) != std::end(plain);
        if (!ok && c.rfind("c_", 0) == 0) {
          const size_t br = c.find('[');
          const std::string id = c.substr(2, br == std::string::npos ? std::string::npos : br - 2);
          if (!computes.count(id)) die("Could not find dump custom compute ID " + id);
          ok = computes[id] == "pe/atom" ? br == std::string::npos : true;
        }
        if (!ok) die("Invalid attribute " + c + " in dump custom command");
        dp.fields.push_back(c);
      }
      dumps.push_back(dp);
    } else if (cmd == "dump_modify") {
      // atoms are always written sorted by id: "sort id" / "sort 1" (the id column) and
      // "sort off" (any order) are all satisfied by that
      bool found = false;
      for (auto &dp : dumps) {
        if (dp.id != w.at(1)) continue;
        found = true;
        for (size_t k = 2; k < w.size(); k++) {
          const std::string &key = w[k];
          const std::string val = k + 1 < w.size() ? w[k + 1] : "";
          if (key == "format" && val == "float" && k + 2 < w.size()) { dp.fmt_float = w[k + 2]; k += 2; }
          else if (key == "scale") { dp.scale = val == "yes"; k++; }
          else if (key == "sort" && (val == "off" || val == "id" || val == "1")) k++;
          else if (key == "first" && val == "no") k++;
          else die("lmff_run does not support dump_modify " + key + " " + val);
        }
      }
      if (!found) die("Could not find dump_modify ID " + w.at(1));
    } else if (cmd == "undump") {
      for (auto it = dumps.begin(); it != dumps.end(); ++it)
        if (it->id == w.at(1)) {
          if (it->fp) std::fclose(it->fp);
          dumps.erase(it);
          break;
        }
    } else if (cmd == "reset_timestep") {
      step = std::stoll(w.at(1));
      if (step < 0) die("Timestep must be >= 0");
      if (pair && lmff_system_set_step(pair->context(), sys, step)) die(lmff_last_error(pair->context()));
    } else if (cmd == "write_data") {
      if (!have_data) die("Write_data command before simulation box is defined");
      std::vector<double> x, v;
      std::vector<int> im;
      current(x, v, im);
      write_data(w.at(1), data, step, x, v, im);
    } else if (cmd == "write_restart") {
      if (!have_data) die("Write_restart command before simulation box is defined");
      restart_now(w.at(1));
    } else if (cmd == "restart") {
      restart_every = std::stoll(w.at(1));
      if (restart_every < 0) die("Illegal restart command");
      restart_files.assign(w.begin() + 2, w.end());
      if (restart_every > 0 && (restart_files.empty() || restart_files.size() > 2)) die("Illegal restart command");
      if (restart_files.size() == 2 && restart_files[0].find('*') != std::string::npos) die("Illegal restart command");
      restart_toggle = 0;
    } else if (cmd == "read_restart") {
      if (have_data) die("Cannot read_restart after simulation box is defined");
      Restart r = read_restart(w.at(1));
      data = std::move(r.data);
      step = r.step;
      dt = r.dt;
      for (int k = 0; k < 3; k++) {
        periodic[k] = r.periodic[k];
        bstr[k] = r.bstr[k];
      }
      have_data = true;
      std::printf("  %lld atoms\n", (long long)data.n);
    } else if (cmd == "min_style") {
      min_style = w.at(1);
    } else if (cmd == "min_modify") {
      for (size_t k = 1; k + 1 < w.size(); k += 2) {
        const std::string &key = w[k], &val = w[k + 1];
        if (key == "dmax") fire.dmax = std::stod(val);
        else if (key == "delaystep") fire.delaystep = std::stoi(val);
        else if (key == "dtgrow") fire.dtgrow = std::stod(val);
        else if (key == "dtshrink") fire.dtshrink = std::stod(val);
        else if (key == "alpha0") fire.alpha0 = std::stod(val);
        else if (key == "alphashrink") fire.alphashrink = std::stod(val);
        else if (key == "tmax") fire.tmax = std::stod(val);
        else if (key == "tmin") fire.tmin = std::stod(val);
        else if (key == "halfstepback") fire.halfstepback = val == "yes";
        else if (key == "initialdelay") fire.delaystep_start = val == "yes";
        else if (key == "vdfmax") fire.max_vdotf_negatif = std::stoi(val);
        else if (key == "norm") fire.norm = val == "max" ? 1 : val == "inf" ? 2 : 0;
        else std::printf("WARNING: min_modify %s ignored by lmff_run\n", key.c_str());
      }
    } else if (cmd == "minimize") {
      // minimize etol ftol maxiter maxeval (LMFF/minimize.cpp) on the device: min_style cg (the
      // default, quadratic line search) or fire
      if (w.size() < 5) die("Illegal minimize command");
      if (min_style != "fire" && min_style != "cg")
        std::printf("WARNING: min_style %s is not available in lmff_run; minimizing with cg\n", min_style.c_str());
      setup_pair();
      lmff_ctx *ctx = pair->context();
      // fix langevin has no minimizer hook: the minimizers see the conservative force only
      if (lmff_system_integrator(ctx, sys, LMFF_NVE, dt, t_target, damp, lseed)) die(lmff_last_error(ctx));
      lmff_min_stats ms;
      auto c0 = std::chrono::steady_clock::now();
      const int rc = min_style == "fire"
                         ? lmff_minimize(ctx, std::stod(w[1]), std::stod(w[2]), std::stoll(w[3]), &fire, &ms, 1)
                         : lmff_minimize_cg(ctx, std::stod(w[1]), std::stod(w[2]), std::stoll(w[3]), std::stoll(w[4]),
                                            fire.dmax, fire.norm, &ms);
      if (rc) die(lmff_last_error(ctx));
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
      static const char *stops[] = {"max iterations", "max force evaluations", "energy tolerance", "force tolerance",
                                    "search direction is not downhill", "linesearch alpha is zero", "forces are zero",
                                    "quadratic factors are zero", "trust region too small", "HFTN minimizer error",
                                    "walltime limit reached", "max iterations with v.f negative"};
      step += ms.iterations;
      std::printf("Loop time of %g on 1 procs for %lld steps with %lld atoms\n\n", secs, (long long)ms.iterations,
                  (long long)data.n);
      std::printf("Minimization stats:\n  Stopping criterion = %s\n", stops[ms.stop]);
      std::printf("  Energy initial, next-to-last, final = \n    %18.15g %18.15g %18.15g\n", ms.e_initial,
                  ms.e_next_to_last, ms.e_final);
      std::printf("  Force two-norm initial, final = %g %g\n", ms.fnorm2_initial, ms.fnorm2_final);
      std::printf("  Force max component initial, final = %g %g\n", ms.fnorminf_initial, ms.fnorminf_final);
      std::printf("  Iterations, force evaluations = %lld %lld\n\n", (long long)ms.iterations, (long long)ms.evaluations);
    } else if (cmd == "group" || cmd == "thermo_style" || cmd == "fix_modify" || cmd == "thermo_modify") {
      std::printf("WARNING: %s ignored by lmff_run\n", cmd.c_str());
    } else {
      die("Unknown command: " + line);
    }
  }
  for (auto &dp : dumps)
    if (dp.fp) std::fclose(dp.fp);
  return 0;
}

int main(int argc, char **argv) {
  // library errors from the Pair adaptor and malformed numbers end the run as LAMMPS'
  // error->all does: the message and a non-zero exit
  try {
    return run_deck(argc, argv);
  } catch (const lmff::PairError &e) {
    die(e.what());
  } catch (const std::exception &e) {
    die(std::string("lmff_run: ") + e.what());
  }
}
