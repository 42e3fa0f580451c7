// potentials -- per-atom quasi-potential export on the MI355X engine.
// Same options, stage messages and output files as the reference tool
// (potentials.cpp:30-207): for every input state, V of |f|^2 at each atom's
// cell, written to <output-dir>/potential-<basename> as "%.12le" of the
// real part, one atom per line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "clat.h"
#include "cli_common.hpp"

namespace {

const char* kUsage =
    "Coulombo v2.0 (c) 2023 [Computer Physics Communications] Rozanski & Zielinski:\n"
    "Exploiting underlying crystal lattice for efficient computation of Coulomb matrix elements in "
    "multi-million atoms nanostructures\n"
    "(MI355X engine)\n"
    "\nUSAGE:\n"
    " potentials [FLAGS/OPTIONS] data files ...\n"
    "\nOPTIONS:\n"
    "  --atoms=PATH           path to *.3d file with atoms' positions\n"
    "  --dielectric=VALUE     dielectric constant, default: 1\n"
    "  --onsite=ENERGY        energy for on-site contribution, default: 0 (eV)\n"
    "  --orbitals=N           number of (spin-)orbitals per atom, default: 20\n"
    "  --output-dir=DIR       directory for output files, default: current\n"
    "  --skip-lines=N         number of lines to be skipped on top of each LCAO file, default: 0\n"
    "  --threads-per-node=N   number of host threads for reading input files, default: all (max 16)\n"
    "  --tf-lattice=VALUE     lattice constant (A) for Thomas-Fermi-Resta model,\n"
    "                           default: model not applied\n"
    "  --precision=P          fp64 (default), fp32x (complex64 passes, FP64 last pass) or fp32\n"
    "\n";

int parse_positive(const cli::Options& o, const char* name, int def, const char* err) {
  std::string v;
  if (!o.has(name, v)) return def;
  const int n = std::atoi(v.c_str());
  if (n <= 0) cli::failf("%s", err);
  return n;
}

void run(const cli::Options& o) {
  cli::Timer timer;
  std::string value;
  const int threads = parse_positive(o, "threads-per-node", cli::parse_threads(), "invalid value for threads-per-node");

  timer.start("Reading atom positions");
  const int n_orb = parse_positive(o, "orbitals", 20, "invalid value for orbitals");
  const int skip = parse_positive(o, "skip-lines", 0, "invalid value for skip-lines");
  std::vector<double> xyz = cli::read_atoms(o.get("atoms"), threads);
  const int64_t n_atoms = (int64_t)xyz.size() / 3;
  int raw[3], pad[3];
  double origin[3], step[3];
  cli::check(clat_lattice_grid(xyz.data(), n_atoms, raw, pad, origin, step));
  std::vector<int32_t> cells((size_t)n_atoms);
  cli::check(clat_lattice_cells(xyz.data(), n_atoms, pad, origin, step, cells.data()));
  double dielectric = 1.0;
  if (o.has("dielectric", value)) dielectric = std::atof(value.c_str());
  std::string out_dir;
  if (o.has("output-dir", out_dir) && !out_dir.empty()) out_dir += '/';
  int precision = CLAT_PREC_FP64;
  if (o.has("precision", value)) {
    if (value == "fp64") precision = CLAT_PREC_FP64;
    else if (value == "fp32") precision = CLAT_PREC_FP32;
    else if (value == "fp32x") precision = CLAT_PREC_FP32X;
    else cli::failf("invalid value for precision: %s", value.c_str());
  }

  timer.start("Reading wavefunctions");
  const auto& args = o.args();
  if ((int)args.size() >= std::numeric_limits<short>::max()) cli::failf("too many input files");
  std::vector<std::vector<double>> states = cli::read_lcao_all(args, skip, n_atoms, n_orb, threads);
  std::vector<std::string> names;
  for (const std::string& a : args) names.push_back(cli::basename_of(a));

  timer.start("Initializing calculator");
  double onsite = 0.0;
  if (o.has("onsite", value)) onsite = std::atof(value.c_str());
  clat_engine* eng = nullptr;
  cli::check(clat_create(pad[0], pad[1], pad[2], precision, CLAT_SCHED_SHARED, &eng));
  struct Guard {
    clat_engine* e;
    ~Guard() { clat_destroy(e); }
  } guard{eng};
  if (o.has("tf-lattice", value)) {
    const double a = std::atof(value.c_str());
    if (!std::isgreater(a, 0.0)) cli::failf("--tf-lattice parameter must be positive");
    if (!std::isgreater(dielectric, 1.0)) cli::failf("dielectric constant must be >1 to use Thomas-Fermi model");
    cli::check(clat_set_interaction_thomas_fermi(eng, step, onsite, dielectric, a));
  } else {
    cli::check(clat_set_interaction_simple(eng, step, onsite, dielectric));
  }
  cli::check(clat_set_atoms(eng, cells.data(), n_atoms, n_orb));
  for (auto& s : states) {
    cli::check(clat_add_state(eng, s.data()));
    std::vector<double>().swap(s);
  }

  timer.start("Computing all quasi-potentials");
  std::vector<double> v(2 * (size_t)std::max<int64_t>(1, n_atoms));
  for (size_t k = 0; k < names.size(); ++k) {
    cli::check(clat_compute_potential_at_atoms(eng, (int)k + 1, v.data()));
    const std::string path = out_dir + "potential-" + names[k];
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) cli::failf("cannot open file %s for writing", path.c_str());
    for (int64_t a = 0; a < n_atoms; ++a) std::fprintf(f, "%.12le\n", v[2 * a]);
    std::fclose(f);
  }
}

}  // namespace

int main(int argc, char** argv) {
  try {
    cli::Options o;
    o.value("atoms").value("dielectric").value("onsite").value("orbitals").value("output-dir");
    o.value("skip-lines").value("threads-per-node").value("tf-lattice").value("precision");
    o.parse(argc, argv);
    if (o.args().empty()) {
      std::fputs(kUsage, stderr);
      std::exit(EXIT_FAILURE);
    }
    run(o);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "ERROR: %s\n", ex.what());
    std::exit(EXIT_FAILURE);
  }
  return 0;
}
