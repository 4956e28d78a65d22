// benchmark.cpp — the reference's Benchmark mode (Benchmark::Run,
// benchmark.cc:12-52) and its Cavitation project (Cavitation::Run,
// cavitation.cc:238-287) on one GPU through libmdx.
//
//   mdx_benchmark <input.cfg> [--procs P] [--threads T]
//
// Mode=Cavitation: the same initial conditions, ThermalizeLoop steps under
// the heat bath (HeatbathType, AimedTemperature, HeatbathTau), the box and
// positions scaled by ChangeScale (MDManager::ChangeScale), then TotalLoop
// + 1 NVE steps with the BubbleHist density grid (GridSize) analysed and
// written to conf%04d.dat every ObserveLoop steps, and the reference's
// "# Thermalize" / "# Observe" lines (time, temperature, potential energy,
// pressure, total energy per particle, seconds).  It runs through the C++
// mirror of the reference interface (mdacp_amd/host/mdacp_gpu.h).
//
// The input file has the reference's format (Parameter: Key=Value lines,
// '#' comments) and keys: Mode=Benchmark, Density (0.7), UnitLength or
// SystemSize, TimeStep (0.001), ThermalizeLoop (150), TotalLoop (1000),
// ObserveLoop (100), InitialVelocity (1.0).  P x T is the reference run's
// MPI processes x OpenMP threads (default 1 x 1); it fixes the domain units
// whose initial conditions are reproduced (ExtractConfigurationMaker and
// SetInitialVelocity per unit), all of which then advance as one periodic
// system.  Output follows the reference: the system information lines,
// one "#observe" line (time, temperature, pressure, energy per particle)
// every ObserveLoop steps of the timed loop, and "# N = ... [SEC] ...
// [MUPS]" with MUPS = TotalLoop N / seconds / 1e6, the observations inside
// the timed loop as the reference has them.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <string>

#include "../../include/mdx.h"
#include "confmaker.h"
#include "mdacp_gpu.h"

namespace {

// Parameter::LoadFromFile (parameter.cc): Key=Value, '#' starts a comment
std::map<std::string, std::string> LoadParameters(const char *path) {
  std::map<std::string, std::string> m;
  std::ifstream in(path);
  if (!in) {
    std::fprintf(stderr, "cannot open %s\n", path);
    std::exit(2);
  }
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      const size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    m[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return m;
}

double GetDouble(const std::map<std::string, std::string> &m, const char *k, double def) {
  auto it = m.find(k);
  return it == m.end() ? def : std::atof(it->second.c_str());
}
int GetInt(const std::map<std::string, std::string> &m, const char *k, int def) {
  auto it = m.find(k);
  return it == m.end() ? def : std::atoi(it->second.c_str());
}

void Check(mdx_system *s, int rc, const char *what) {
  if (rc) {
    std::fprintf(stderr, "%s: %s (status %d)\n", what, mdx_last_error(s), rc);
    std::exit(1);
  }
}

std::string GetString(const std::map<std::string, std::string> &m, const char *k, const char *def) {
  auto it = m.find(k);
  return it == m.end() ? std::string(def) : it->second;
}

// Cavitation::Run (cavitation.cc:238-287) for the units of P x T as one
// periodic system, through the mirror's MDUnit and BubbleHist
int RunCavitation(const std::map<std::string, std::string> &prm, int procs, int threads,
                  double unit_length) {
  const double density = GetDouble(prm, "Density", 0.7);
  const double v0 = GetDouble(prm, "InitialVelocity", 1.0);
  const double dt = GetDouble(prm, "TimeStep", 0.001);
  const int t_loop = GetInt(prm, "ThermalizeLoop", 150);
  const int loop = GetInt(prm, "TotalLoop", 1000);
  const int observe_loop = GetInt(prm, "ObserveLoop", 100);
  const double change_scale = GetDouble(prm, "ChangeScale", 1.05);
  const double grid_size = GetDouble(prm, "GridSize", 3.0);
  const mdx::BenchmarkState st = mdx::MakeBenchmarkState(procs, threads, unit_length, density, v0);
  const int n = static_cast<int>(st.q.size() / 3);
  try {
    // SimulationInfo (simulationinfo.cc:5-30) with the heat bath keys
    mdx::SimulationInfo sinfo(st.L[0], st.L[1], st.L[2], dt);
    sinfo.AimedTemperature = GetDouble(prm, "AimedTemperature", 1.0);
    sinfo.HeatbathTau = GetDouble(prm, "HeatbathTau", 0.1);
    sinfo.HeatbathGamma = GetDouble(prm, "HeatbathGamma", 0.1);
    sinfo.HeatbathType =
        GetString(prm, "HeatbathType", "NoseHoover") == "Langevin" ? mdx::HT_LANGEVIN : mdx::HT_NOSEHOOVER;
    mdx::MDUnit unit(&sinfo);
    for (int i = 0; i < n; i++) {
      double x[mdx::D] = {st.q[3 * i], st.q[3 * i + 1], st.q[3 * i + 2], 0.0};
      double v[mdx::D] = {st.p[3 * i], st.p[3 * i + 1], st.p[3 * i + 2], 0.0};
      unit.AddParticle(x, v, i);
    }
    std::cout << "# Mode = Cavitation" << std::endl;
    std::cout << "# " << procs << " MPI Process(es), " << threads << " OpenMP Thread(s), Total "
              << procs * threads << " Unit(s) on one GPU" << std::endl;
    std::cout << "# Number of Particles = " << n << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    auto seconds = [&]() {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    // MDManager::Temperature / PotentialEnergy / Pressure / TotalEnergy
    // (mdmanager.cc:531-570), per particle
    auto line = [&](const char *tag) {
      const double T = unit.Temperature();
      const double pe = unit.PotentialEnergy();
      std::cout << unit.GetVariables()->SimulationTime << " " << T << " " << pe << " "
                << unit.Pressure() << " " << (1.5 * T + pe) << " " << seconds() << " # " << tag
                << std::endl;
    };
    sinfo.ControlTemperature = true;   // MDManager::SetControlTemperature
    for (int i = 0; i < t_loop; i++) {
      unit.Calculate();
      if (i % observe_loop == 0) line("Thermalize");
    }
    // MDManager::ChangeScale (mdmanager.cc:572-584)
    unit.MakePairList();
    for (int d = 0; d < 3; d++) sinfo.L[d] *= change_scale;
    unit.ChangeScale(change_scale);
    unit.MakePairList();
    std::cout << "# Change Scale " << change_scale << std::endl;
    sinfo.ControlTemperature = false;
    mdx::BubbleHist bhist(&unit, grid_size);
    for (int i = 0; i <= loop; i++) {
      unit.Calculate();
      if (i % observe_loop == 0) {
        bhist.Analyse(&unit);
        line("Observe");
      }
    }
  } catch (const mdx::Error &e) {
    std::fprintf(stderr, "mdx error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.cfg [--procs P] [--threads T]\n", argv[0]);
    return 2;
  }
  int procs = 1, threads = 1;
  for (int a = 2; a + 1 < argc; a += 2) {
    const std::string k = argv[a];
    if (k == "--procs") procs = std::atoi(argv[a + 1]);
    else if (k == "--threads") threads = std::atoi(argv[a + 1]);
  }
  const auto prm = LoadParameters(argv[1]);
  auto mode = prm.find("Mode");
  if (mode == prm.end() || (mode->second != "Benchmark" && mode->second != "Cavitation")) {
    std::fprintf(stderr, "only Mode=Benchmark and Mode=Cavitation are supported\n");
    return 2;
  }
  const mdx::UnitGrid grid(procs, threads);
  double unit_length = GetDouble(prm, "UnitLength", 0.0);
  if (prm.count("SystemSize")) unit_length = GetDouble(prm, "SystemSize", 0.0) / grid.size[0];
  if (!(unit_length > 0.0) || grid.size[0] != grid.size[1] || grid.size[0] != grid.size[2]) {
    std::fprintf(stderr, "need UnitLength, or SystemSize on a cubic unit grid\n");
    return 2;
  }
  if (mode->second == "Cavitation") return RunCavitation(prm, procs, threads, unit_length);
  const double density = GetDouble(prm, "Density", 0.7);
  const double v0 = GetDouble(prm, "InitialVelocity", 1.0);
  const double dt = GetDouble(prm, "TimeStep", 0.001);
  const int t_loop = GetInt(prm, "ThermalizeLoop", 150);
  const int loop = GetInt(prm, "TotalLoop", 1000);
  const int observe_loop = GetInt(prm, "ObserveLoop", 100);

  // initial conditions of every unit (confmaker.cc:174-255, variables.cc:87-113)
  const mdx::BenchmarkState st = mdx::MakeBenchmarkState(procs, threads, unit_length, density, v0);
  const int n = static_cast<int>(st.q.size() / 3);
  mdx_params p;
  mdx_default_params(&p);
  for (int d = 0; d < 3; d++) p.L[d] = st.L[d];
  p.dt = dt;
  mdx_system *sys = nullptr;
  Check(nullptr, mdx_create(&p, &sys), "mdx_create");
  Check(sys, mdx_set_state(sys, n, st.q.data(), st.p.data()), "mdx_set_state");

  // MDManager::ShowSystemInformation (simulationinfo.cc:54-71)
  const double V = st.L[0] * st.L[1] * st.L[2];
  std::cout << "# Mode = Benchmark" << std::endl;
  std::cout << "# " << procs << " MPI Process(es), " << threads << " OpenMP Thread(s), Total "
            << procs * threads << " Unit(s) on one GPU" << std::endl;
  std::cout << "# Number of Particles = " << n << std::endl;
  std::cout << "# System Size = (" << st.L[0] << "," << st.L[1] << "," << st.L[2] << ")" << std::endl;
  std::cout << "# Density = " << static_cast<double>(n) / V << std::endl;
  std::cout << "# Cutoff Length  = " << p.cutoff << std::endl;
  std::cout << "# TimeStep = " << dt << std::endl;
  std::cout << "# ControlTemperature = no" << std::endl;
  std::cout << "# IsPeriodic = yes" << std::endl;

  Check(sys, mdx_make_pairlist(sys), "mdx_make_pairlist");
  double s_time = 0.0;
  auto advance = [&](int k) {
    if (k <= 0) return;
    Check(sys, mdx_calculate(sys, k, nullptr), "mdx_calculate");
    for (int i = 0; i < k; i++) s_time += dt;   // MDManager::Calculate, mdmanager.cc:184
  };
  advance(t_loop);
  const auto t0 = std::chrono::steady_clock::now();
  // the reference observes after the steps i = 0, OBSERVE, 2 OBSERVE, ...
  // of the timed loop (benchmark.cc:32-41), ghosts as Calculate leaves them
  int done = 0;
  for (int i = 0; i < loop; i += observe_loop) {
    advance(i + 1 - done);
    done = i + 1;
    double ke = 0, pe = 0, vir = 0;
    Check(sys, mdx_observe(sys, &ke, &pe, &vir), "mdx_observe");
    const double T = ke / n / 1.5;                              // mdmanager.cc:531-533
    const double P = (T - vir / n) * n / V;                     // mdmanager.cc:563-570
    std::cout << s_time << " " << T << " " << P << " " << (ke / n + pe / n) << " #observe"
              << std::endl;
  }
  advance(loop - done);
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double mups = static_cast<double>(loop) * static_cast<double>(n) / sec / 1.0e6;
  std::cout << "# N = " << n << " " << sec << " [SEC] " << mups << " [MUPS]" << std::endl;
  mdx_destroy(sys);
  return 0;
}
