// coulombo -- drop-in command-line driver of the MI355X engine.
// Same options, input files, stage messages and output files as the
// reference driver (coulombo.cpp:30-315); the FFT path runs on the GPU
// through the C ABI of include/clat.h.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "clat.h"
#include "cli_common.hpp"
#include "host_logic.hpp"

namespace {

const char* kUsage =
    "Coulombo v2.0 (c) 2023 [Computer Physics Communications] Rozanski & Zielinski:\n"
    "Exploiting underlying crystal lattice for efficient computation of Coulomb matrix elements in "
    "multi-million atoms nanostructures\n"
    "(MI355X engine)\n"
    "\nUSAGE:\n"
    " coulombo [FLAGS/OPTIONS] data files ...\n"
    "\nOPTIONS:\n"
    "  --atoms=PATH           path to *.3d file with atoms' positions\n"
    "  --dielectric=VALUE     dielectric constant, default: 1\n"
    "  --integrals=LIST       comma-separated list of integrals to be computed\n"
    "                           (eg. \"eeee,hhhh,ehhe,eheh\"),\n"
    "                           default: all integrals are computed\n"
    "  --onsite=ENERGY        energy for on-site contribution, default: 0 (eV)\n"
    "  --orbitals=N           number of (spin-)orbitals per atom, default: 20\n"
    "  --output-dir=DIR       directory for output files, default: current\n"
    "  --skip-lines=N         number of lines to be skipped on top of each LCAO file, default: 0\n"
    "  --threads-per-node=N   number of host threads for reading input files, default: all (max 16)\n"
    "  --tf-lattice=VALUE     lattice constant (A) for Thomas-Fermi-Resta model,\n"
    "                           default: model not applied\n"
    "  --precision=P          fp64 (default); fp32x: complex64 passes, FP64 last pass and sums\n"
    "                           (each integral within 1e-5 of itself vs fp64, or within 3e-10 of\n"
    "                           the largest integral if that is larger); fp32: complex64\n"
    "                           throughout (within 1e-5 of max(|I|, 1e-3 of the largest))\n"
    "  --schedule=S           shared (default): the 8 rounds share x/y passes; rounds: 8 full rounds\n"
    "\nFLAGS:\n"
    "  --reference-stale-values  reuse the previous overlap when only the left product changes,\n"
    "                            exactly as the reference main loop does\n"
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
  std::string integrals = "****";
  o.has("integrals", integrals);
  std::string out_dir;
  if (o.has("output-dir", out_dir) && !out_dir.empty()) out_dir += '/';
  int precision = CLAT_PREC_FP64, schedule = CLAT_SCHED_SHARED;
  if (o.has("precision", value)) {
    if (value == "fp64") precision = CLAT_PREC_FP64;
    else if (value == "fp32") precision = CLAT_PREC_FP32;
    else if (value == "fp32x") precision = CLAT_PREC_FP32X;
    else cli::failf("invalid value for precision: %s", value.c_str());
  }
  if (o.has("schedule", value)) {
    if (value == "shared") schedule = CLAT_SCHED_SHARED;
    else if (value == "rounds") schedule = CLAT_SCHED_ROUNDS;
    else cli::failf("invalid value for schedule: %s", value.c_str());
  }

  timer.start("Reading wavefunctions");
  const auto& args = o.args();
  const int n_inputs = (int)args.size();
  if (n_inputs >= std::numeric_limits<short>::max()) cli::failf("too many input files");
  // name checks in argument order (coulombo.cpp:96-117); the files before the
  // first bad name are read first, so that their own errors come first
  int n_h = 0, n_e = 0;
  std::string name_error;
  size_t n_ok = args.size();
  for (size_t i = 0; i < args.size() && name_error.empty(); ++i) {
    const std::string base = cli::basename_of(args[i]);
    if (!base.empty() && base[0] == 'h') {
      if (n_e) name_error = "hole states must appear before electron states";
      else ++n_h;
    } else if (!base.empty() && base[0] == 'e') {
      ++n_e;
    } else {
      name_error = "invalid state file name: " + args[i];
    }
    if (!name_error.empty()) n_ok = i;
  }
  std::vector<std::vector<double>> states = cli::read_lcao_all(
      std::vector<std::string>(args.begin(), args.begin() + (long)n_ok), skip, n_atoms, n_orb, threads);
  if (!name_error.empty()) cli::failf("%s", name_error.c_str());

  timer.start("Preparing plan");
  std::vector<int32_t> quads;
  try {
    clat::host::Pattern pattern(integrals, n_h);
    quads = clat::host::enumerate_integrals(pattern, n_inputs);
  } catch (const clat::host::Error& e) {
    throw cli::Fail(e.what());
  }
  const int64_t n_int = (int64_t)quads.size() / 4;

  timer.start("Initializing calculator");
  double onsite = 0.0;
  if (o.has("onsite", value)) onsite = std::atof(value.c_str());
  clat_engine* eng = nullptr;
  cli::check(clat_create(pad[0], pad[1], pad[2], precision, schedule, &eng));
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

  timer.start("Computing requested integrals");
  std::vector<double> vals(2 * (size_t)std::max<int64_t>(1, n_int));
  const int flags = o.flag_set("reference-stale-values") ? CLAT_FLAG_STALE_VALUES : 0;
  if (n_int > 0) cli::check(clat_compute_integrals(eng, quads.data(), n_int, flags, vals.data()));

  timer.start("Exporting results");
  // coulombo.cpp:213-262
  std::map<std::array<int, 4>, std::array<double, 2>> values;
  for (int64_t i = 0; i < n_int; ++i)
    values[{quads[4 * i], quads[4 * i + 1], quads[4 * i + 2], quads[4 * i + 3]}] = {vals[2 * i], vals[2 * i + 1]};
  char core[5];
  core[4] = 0;
  for (int ti = 0; ti < 2; ++ti)
    for (int tj = 0; tj < 2; ++tj)
      for (int tk = 0; tk < 2; ++tk)
        for (int tl = 0; tl < 2; ++tl) {
          FILE* file = nullptr;
          core[0] = ti ? 'e' : 'h';
          core[1] = tj ? 'e' : 'h';
          core[2] = tk ? 'e' : 'h';
          core[3] = tl ? 'e' : 'h';
          const int Ni = ti ? n_e : n_h, Nj = tj ? n_e : n_h, Nk = tk ? n_e : n_h, Nl = tl ? n_e : n_h;
          for (int ni = 1; ni <= Ni; ++ni)
            for (int nj = 1; nj <= Nj; ++nj)
              for (int nk = 1; nk <= Nk; ++nk)
                for (int nl = 1; nl <= Nl; ++nl) {
                  std::array<int, 4> spec = {ti ? n_h + ni : n_h + 1 - ni, tj ? n_h + nj : n_h + 1 - nj,
                                             tk ? n_h + nk : n_h + 1 - nk, tl ? n_h + nl : n_h + 1 - nl};
                  auto it = values.find(spec);
                  if (it == values.end()) continue;
                  if (!file) {
                    const std::string path = out_dir + core + ".txt";
                    file = std::fopen(path.c_str(), "w");
                    if (!file) cli::failf("cannot open file %s for writing", path.c_str());
                  }
                  std::fprintf(file, "%2d %2d %2d %2d   %17.14f %17.14f\n", ni, nj, nk, nl, it->second[0],
                               it->second[1]);
                }
          if (file) std::fclose(file);
        }
}

}  // namespace

int main(int argc, char** argv) {
  try {
    cli::Options o;
    o.value("atoms").value("dielectric").value("integrals").value("onsite").value("orbitals");
    o.value("output-dir").value("skip-lines").value("threads-per-node").value("tf-lattice");
    o.value("precision").value("schedule").flag("reference-stale-values");
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
