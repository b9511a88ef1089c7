// lmff_run: runs the subset of a LAMMPS input deck that drives the LMFF force path on one
// MI355X, so an unchanged deck of the reference's form
//   units metal / atom_style full / boundary / read_data / mass / pair_style hybrid/overlay
//   tersoff ilp/graphene/hbn / pair_coeff / neighbor / neigh_modify / timestep / velocity /
//   fix nve | fix langevin / thermo / run
// executes with the force field, the neighbor build and the integrator on the GPU.
// Output follows LAMMPS' log format (thermo table, "Loop time", "Performance:" as in
// LMFF/finish.cpp:124,162).  Extra command: "lmff_precision fp64|mixed".
//
// usage: lmff_run -in deck [-var NAME VALUE ...]
#include <array>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
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

struct Data {
  int64_t n = 0;
  int ntypes = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  std::vector<double> x, v;
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

}  // namespace

int main(int argc, char **argv) {
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
  int style = LMFF_NVE;
  double t_target = 0, damp = 1;
  uint64_t lseed = 1;
  bool have_fix = false;
  std::unique_ptr<lmff::PairLMFFHip> pair;
  int sys = -1;
  int64_t step = 0;

  auto subst = [&](std::string s) {
    for (size_t p; (p = s.find("${")) != std::string::npos;) {
      size_t q = s.find('}', p);
      if (q == std::string::npos) die("Invalid variable name in " + s);
      std::string name = s.substr(p + 2, q - p - 2);
      if (!vars.count(name)) die("Substitution for illegal variable " + name);
      s = s.substr(0, p) + vars[name] + s.substr(q + 1);
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
      for (int k = 0; k < 3 && k + 1 < (int)w.size(); k++) periodic[k] = w[k + 1] == "p" ? 1 : 0;
    } else if (cmd == "variable") {
      if (w.size() >= 4 && !vars.count(w[1])) vars[w[1]] = w[3];
    } else if (cmd == "read_data") {
      data = read_data(w.at(1));
      have_data = true;
      std::printf("  %lld atoms\n", (long long)data.n);
    } else if (cmd == "mass") {
      data.mass[std::stoi(w.at(1))] = std::stod(w.at(2));
    } else if (cmd == "pair_style") {
      if (w.size() < 4 || w[1] != "hybrid/overlay" || w[2] != "tersoff" || w[3] != "ilp/graphene/hbn")
        die("lmff_run supports pair_style hybrid/overlay tersoff ilp/graphene/hbn <cut> [tap]");
      pair_args.assign(w.begin() + 4, w.end());
    } else if (cmd == "pair_coeff") {
      coeffs.push_back(std::vector<std::string>(w.begin() + 1, w.end()));
    } else if (cmd == "neighbor") {
      skin = std::stod(w.at(1));
    } else if (cmd == "neigh_modify") {
      for (size_t k = 1; k + 1 < w.size(); k++) {
        if (w[k] == "every") every = std::stoi(w[k + 1]);
        else if (w[k] == "delay") delay = std::stoi(w[k + 1]);
        else if (w[k] == "check") check = w[k + 1] == "yes";
      }
    } else if (cmd == "timestep") {
      dt = std::stod(w.at(1));
    } else if (cmd == "lmff_precision") {
      precision = w.at(1) == "mixed" ? LMFF_MIXED : LMFF_FP64;
    } else if (cmd == "velocity") {
      // velocity all create T seed: Gaussian, zero net momentum, scaled to T (own RNG stream)
      if (w.size() < 5 || w[2] != "create") die("lmff_run supports velocity all create T seed");
      if (!have_data) die("Velocity command before simulation box is defined");
      const double T = std::stod(w[3]);
      std::mt19937_64 rng(std::stoull(w[4]));
      std::normal_distribution<double> g(0.0, 1.0);
      double p[3] = {0, 0, 0}, mt = 0, ke = 0;
      for (int64_t i = 0; i < data.n; i++) {
        const double m = data.mass.at(data.type[i]);
        for (int c = 0; c < 3; c++) {
          data.v[3 * i + c] = g(rng) * std::sqrt(BOLTZ * T / (m * MVV2E));
          p[c] += m * data.v[3 * i + c];
        }
        mt += m;
      }
      for (int64_t i = 0; i < data.n; i++) {
        const double m = data.mass.at(data.type[i]);
        for (int c = 0; c < 3; c++) {
          data.v[3 * i + c] -= p[c] / mt;
          ke += 0.5 * MVV2E * m * data.v[3 * i + c] * data.v[3 * i + c];
        }
      }
      const double t_now = 2 * ke / ((3.0 * data.n - 3) * BOLTZ);
      const double sc = t_now > 0 ? std::sqrt(T / t_now) : 0.0;
      for (auto &q : data.v) q *= sc;
    } else if (cmd == "fix") {
      if (w.size() >= 4 && w[3] == "nve") {
        have_fix = true;
      } else if (w.size() >= 8 && w[3] == "langevin") {
        style = LMFF_LANGEVIN;
        t_target = std::stod(w[4]);
        damp = std::stod(w[6]);
        lseed = std::stoull(w[7]);
      } else {
        std::printf("WARNING: fix %s ignored by lmff_run\n", w.size() > 3 ? w[3].c_str() : "?");
      }
    } else if (cmd == "thermo") {
      thermo_every = std::stoi(w.at(1));
    } else if (cmd == "run") {
      const int64_t nsteps = std::stoll(w.at(1));
      setup_pair();
      lmff_ctx *ctx = pair->context();
      if (!have_fix) std::printf("WARNING: No fixes defined, atoms won't move\n");
      if (lmff_system_integrator(ctx, sys, style, dt, t_target, damp, lseed)) die(lmff_last_error(ctx));
      if (lmff_evaluate(ctx, 1, 2)) die(lmff_last_error(ctx));
      lmff_thermo t0;
      lmff_system_thermo(ctx, sys, &t0);
      double vol = 1;
      for (int k = 0; k < 3; k++) vol *= data.hi[k] - data.lo[k];
      auto press = [&](const lmff_thermo &t) {
        return ((3.0 * data.n - 3) * BOLTZ * t.temp + t.virial[0] + t.virial[1] + t.virial[2]) / 3.0 / vol * NKTV2P;
      };
      std::printf("Step Temp E_pair TotEng Press\n");
      std::printf("%8lld %12.6g %16.10g %16.10g %12.6g\n", (long long)step, t0.temp, t0.pe, t0.pe + t0.ke, press(t0));
      const int64_t te = thermo_every > 0 ? thermo_every : nsteps;
      const int64_t nrec = nsteps / te + (nsteps % te ? 1 : 0);
      std::vector<lmff_thermo> rec((size_t)std::max<int64_t>(nrec, 1));
      auto c0 = std::chrono::steady_clock::now();
      if (lmff_run(ctx, nsteps, te, 1, rec.data(), nrec)) die(lmff_last_error(ctx));
      double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
      for (int64_t r = 0; r < nrec; r++) {
        const auto &t = rec[r];
        std::printf("%8lld %12.6g %16.10g %16.10g %12.6g\n", (long long)(step + (r + 1) * te > step + nsteps ? step + nsteps : step + (r + 1) * te),
                    t.temp, t.pe, t.pe + t.ke, press(t));
      }
      step += nsteps;
      std::printf("Loop time of %g on 1 procs for %lld steps with %lld atoms\n\n", secs, (long long)nsteps,
                  (long long)data.n);
      const double tps = secs > 0 ? nsteps / secs : 0;
      std::printf("Performance: %.3f ns/day, %.3f hours/ns, %.3f timesteps/s\n", tps * dt * 86400.0 / 1000.0,
                  tps > 0 ? 1.0 / (tps * dt * 3.6) : 0.0, tps);
      std::printf("Performance: %.4g atom-steps/s\n", tps * data.n);
    } else if (cmd == "group" || cmd == "compute" || cmd == "thermo_style" || cmd == "dump" || cmd == "dump_modify" ||
               cmd == "replicate" || cmd == "reset_timestep" || cmd == "minimize" || cmd == "unfix") {
      std::printf("WARNING: %s ignored by lmff_run\n", cmd.c_str());
    } else {
      die("Unknown command: " + line);
    }
  }
  return 0;
}
