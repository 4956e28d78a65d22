// Host code under AddressSanitizer and UndefinedBehaviorSanitizer, run on a
// machine without a GPU: the C ABI's argument checks and every path that
// returns before touching a device, the Benchmark-mode initial conditions
// (confmaker.cpp), the C restatement of the reference path (oracle/
// mdoracle.c: build, steps, rebuilds, heat baths, exports) and the mirror's
// host-only StopWatch.  Built by tests/cpp/sanitize.mk.
//
// usage: sanitize_driver capi
//          argument checks; prints "checks <n>" and "failures <k>"
//        sanitize_driver conf <procs> <threads> <unit_length> <density> <v0> <prefix>
//          MakeBenchmarkState and the C entry; writes <prefix>_q.f64, _p.f64
//        sanitize_driver oracle <state.bin> <steps>
//          the oracle over <steps> steps, then 20 Nose-Hoover and 20
//          Langevin steps; prints the observables and counts
//        sanitize_driver stopwatch <dir>
//          StopWatch without a device; writes <dir>/sw00000.dat
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "../../mdacp_amd/host/confmaker.h"
#include "../../mdacp_amd/host/mdacp_gpu.h"
extern "C" {
#include "../../oracle/mdoracle.h"
}

namespace {

int checks = 0, failures = 0;

void Expect(bool ok, const char *what) {
  checks++;
  if (!ok) {
    failures++;
    std::fprintf(stderr, "FAILED: %s\n", what);
  }
}

int Capi() {
  mdx_params p;
  mdx_default_params(&p);
  Expect(p.cutoff == 3.0 && p.buffer_length > 0.0 && p.dt > 0.0, "default params");
  mdx_system *s = reinterpret_cast<mdx_system *>(0x1);
  Expect(mdx_create(nullptr, &s) == MDX_ERR_ARG, "create(NULL params)");
  Expect(mdx_create(&p, nullptr) == MDX_ERR_ARG, "create(NULL out)");
  for (int d = 0; d < 3; d++) p.L[d] = 20.0;
  p.L[1] = 6.0;   // below two search lengths
  Expect(mdx_create(&p, &s) == MDX_ERR_BOX && s == nullptr, "create(small box)");
  Expect(std::strlen(mdx_last_error(nullptr)) > 0, "last_error after a failed create");
  p.L[1] = 20.0;
  p.dt = -1.0;
  Expect(mdx_create(&p, &s) == MDX_ERR_ARG, "create(dt < 0)");
  p.dt = 0.001;
  p.L[0] = NAN;
  Expect(mdx_create(&p, &s) == MDX_ERR_BOX, "create(NaN box)");
  p.L[0] = 20.0;

  // every entry point with a NULL system
  double d3[3] = {0, 0, 0}, x = 0.0;
  int i1 = 0, i2 = 0;
  int32_t ids[4] = {0, 0, 0, 0};
  int64_t c64 = 0;
  uint8_t bytes[8];
  Expect(mdx_destroy(nullptr) == MDX_OK, "destroy(NULL)");
  Expect(mdx_set_state(nullptr, 1, d3, d3) == MDX_ERR_ARG, "set_state(NULL)");
  Expect(mdx_get_state(nullptr, d3, d3) == MDX_ERR_ARG, "get_state(NULL)");
  Expect(mdx_particle_number(nullptr, &i1, &i2) == MDX_ERR_ARG, "particle_number(NULL)");
  Expect(mdx_get_images(nullptr, d3, d3, ids, ids) == MDX_ERR_ARG, "get_images(NULL)");
  Expect(mdx_synchronize(nullptr) == MDX_ERR_ARG, "synchronize(NULL)");
  Expect(mdx_make_pairlist(nullptr) == MDX_ERR_ARG, "make_pairlist(NULL)");
  Expect(mdx_is_pairlist_expired(nullptr, &i1) == MDX_ERR_ARG, "is_pairlist_expired(NULL)");
  Expect(mdx_update_position_half(nullptr) == MDX_ERR_ARG, "update_position_half(NULL)");
  Expect(mdx_refresh_images(nullptr) == MDX_ERR_ARG, "refresh_images(NULL)");
  Expect(mdx_calculate_force(nullptr) == MDX_ERR_ARG, "calculate_force(NULL)");
  Expect(mdx_calculate(nullptr, 1, &i1) == MDX_ERR_ARG, "calculate(NULL)");
  mdx_heatbath hb{};
  Expect(mdx_set_heatbath(nullptr, &hb) == MDX_ERR_ARG, "set_heatbath(NULL)");
  Expect(mdx_get_heatbath(nullptr, &hb) == MDX_ERR_ARG, "get_heatbath(NULL)");
  Expect(mdx_heatbath_zeta(nullptr, 1.0) == MDX_ERR_ARG, "heatbath_zeta(NULL)");
  Expect(mdx_heatbath_momenta(nullptr) == MDX_ERR_ARG, "heatbath_momenta(NULL)");
  Expect(mdx_langevin(nullptr) == MDX_ERR_ARG, "langevin(NULL)");
  Expect(mdx_observe(nullptr, &x, &x, &x) == MDX_ERR_ARG, "observe(NULL)");
  Expect(mdx_forces(nullptr, d3) == MDX_ERR_ARG, "forces(NULL)");
  Expect(mdx_pair_number(nullptr, &c64) == MDX_ERR_ARG, "pair_number(NULL)");
  Expect(mdx_pairs_canonical(nullptr, ids, 1, &c64) == MDX_ERR_ARG, "pairs_canonical(NULL)");
  Expect(mdx_pairs_canonical_noalloc(nullptr, ids, 1, &c64) == MDX_ERR_ARG, "pairs_canonical_noalloc(NULL)");
  Expect(mdx_pairs_csr(nullptr, ids, ids, ids, 1, &c64) == MDX_ERR_ARG, "pairs_csr(NULL)");
  Expect(mdx_adjust_periodic_boundary(nullptr) == MDX_ERR_ARG, "adjust_periodic_boundary(NULL)");
  Expect(mdx_scale_positions(nullptr, 1.1) == MDX_ERR_ARG, "scale_positions(NULL)");
  Expect(mdx_set_box(nullptr, d3) == MDX_ERR_ARG, "set_box(NULL)");
  int32_t dims[3];
  Expect(mdx_density_grid(nullptr, 3.0, dims, bytes, 8) == MDX_ERR_ARG, "density_grid(NULL)");
  Expect(mdx_get_cell_order(nullptr, ids) == MDX_ERR_ARG, "get_cell_order(NULL)");
  double ms = 0.0;
  int64_t n64 = 0;
  Expect(mdx_timing(nullptr, &ms, &ms, &ms, &n64, &n64) == MDX_ERR_ARG, "timing(NULL)");
  Expect(mdx_timing_reset(nullptr, 1) == MDX_ERR_ARG, "timing_reset(NULL)");
  Expect(mdx_kick_profile(nullptr, &ms, &ms) == MDX_ERR_ARG, "kick_profile(NULL)");

  // the ensemble's checks
  mdx_ensemble *e = reinterpret_cast<mdx_ensemble *>(0x1);
  int n2[2] = {100, 0};
  double L6[6] = {20, 20, 20, 20, 20, 20};
  Expect(mdx_ensemble_create(&p, 0, n2, L6, &e) == MDX_ERR_ARG, "ensemble_create(0 replicas)");
  Expect(mdx_ensemble_create(&p, 2, n2, L6, &e) == MDX_ERR_ARG, "ensemble_create(empty replica)");
  n2[1] = 50;
  L6[4] = 5.0;
  Expect(mdx_ensemble_create(&p, 2, n2, L6, &e) == MDX_ERR_BOX && e == nullptr,
         "ensemble_create(small box)");
  Expect(std::strlen(mdx_ensemble_last_error(nullptr)) > 0, "ensemble_last_error");
  Expect(mdx_ensemble_destroy(nullptr) == MDX_OK, "ensemble_destroy(NULL)");
  Expect(mdx_ensemble_size(nullptr, &i1, &i2) == MDX_ERR_ARG, "ensemble_size(NULL)");
  Expect(mdx_ensemble_calculate(nullptr, 1, &i1) == MDX_ERR_ARG, "ensemble_calculate(NULL)");
  Expect(mdx_ensemble_set_state(nullptr, d3, d3) == MDX_ERR_ARG, "ensemble_set_state(NULL)");
  Expect(mdx_ensemble_get_state(nullptr, d3, d3) == MDX_ERR_ARG, "ensemble_get_state(NULL)");
  Expect(mdx_ensemble_status(nullptr, &i1, &i2) == MDX_ERR_ARG, "ensemble_status(NULL)");
  Expect(mdx_ensemble_observe(nullptr, &x, &x, &x) == MDX_ERR_ARG, "ensemble_observe(NULL)");
  mdx_ensemble_config cfg{};
  Expect(mdx_ensemble_get_config(nullptr, &cfg) == MDX_ERR_ARG, "ensemble_get_config(NULL)");
  Expect(mdx_ensemble_set_config(nullptr, 128, 1) == MDX_ERR_ARG, "ensemble_set_config(NULL)");
  Expect(mdx_ensemble_autotune(nullptr, 4, &cfg) == MDX_ERR_ARG, "ensemble_autotune(NULL)");

  // a valid request on a machine without a device fails cleanly
  if (std::getenv("SANITIZE_NO_DEVICE")) {
    Expect(mdx_create(&p, &s) == MDX_ERR_HIP && s == nullptr, "create without a device");
  }
  std::printf("checks %d\nfailures %d\n", checks, failures);
  return failures ? 1 : 0;
}

bool WriteF64(const std::string &path, const std::vector<double> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(v.data(), 8, v.size(), f) == v.size();
  std::fclose(f);
  return ok;
}

int Conf(int argc, char **argv) {
  if (argc < 8) return 2;
  const int procs = std::atoi(argv[2]), threads = std::atoi(argv[3]);
  const double ul = std::atof(argv[4]), rho = std::atof(argv[5]), v0 = std::atof(argv[6]);
  const mdx::BenchmarkState st = mdx::MakeBenchmarkState(procs, threads, ul, rho, v0);
  const long n = static_cast<long>(st.q.size() / 3);
  // the C entry: sizing call, a short buffer, then the whole state
  long m = -1;
  double L3[3];
  Expect(mdacp_benchmark_state(procs, threads, ul, rho, v0, nullptr, nullptr, 0, &m, L3) == 0 &&
             m == n,
         "C entry sizing call");
  std::vector<double> q(3 * (n / 2)), p(3 * (n / 2));
  Expect(mdacp_benchmark_state(procs, threads, ul, rho, v0, q.data(), p.data(), n / 2, &m, L3) == 0 &&
             m == n && std::memcmp(q.data(), st.q.data(), 8 * q.size()) == 0,
         "C entry, short buffer");
  q.resize(3 * n);
  p.resize(3 * n);
  Expect(mdacp_benchmark_state(procs, threads, ul, rho, v0, q.data(), p.data(), n, &m, L3) == 0 &&
             std::memcmp(p.data(), st.p.data(), 8 * p.size()) == 0 && L3[0] == st.L[0],
         "C entry, whole state");
  Expect(mdacp_benchmark_state(0, threads, ul, rho, v0, nullptr, nullptr, 0, &m, L3) == -1,
         "C entry, zero processes");
  const std::string pre = argv[7];
  if (!WriteF64(pre + "_q.f64", st.q) || !WriteF64(pre + "_p.f64", st.p)) return 2;
  std::printf("n %ld\nL %.17g %.17g %.17g\nfailures %d\n", n, st.L[0], st.L[1], st.L[2], failures);
  return failures ? 1 : 0;
}

struct State {
  int64_t n = 0;
  double L[3] = {0, 0, 0}, dt = 0.001;
  std::vector<double> q, p;
};

bool Read(const char *path, State &s) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  bool ok = std::fread(&s.n, 8, 1, f) == 1 && std::fread(s.L, 8, 3, f) == 3 &&
            std::fread(&s.dt, 8, 1, f) == 1;
  s.q.resize(3 * s.n);
  s.p.resize(3 * s.n);
  ok = ok && std::fread(s.q.data(), 8, 3 * s.n, f) == static_cast<size_t>(3 * s.n) &&
       std::fread(s.p.data(), 8, 3 * s.n, f) == static_cast<size_t>(3 * s.n);
  std::fclose(f);
  return ok;
}

int Oracle(int argc, char **argv) {
  if (argc < 4) return 2;
  State st;
  if (!Read(argv[2], st)) return 2;
  const int steps = std::atoi(argv[3]);
  orc_params prm{};
  for (int d = 0; d < 3; d++) prm.L[d] = st.L[d];
  prm.dt = st.dt;
  prm.cutoff = 3.0;
  prm.buffer_length = 0.3;
  prm.check_list_length = 500;
  const int n = static_cast<int>(st.n);
  orc_sys *s = orc_create(&prm, 2 * n + 4096);
  if (!s) return 3;
  int rc = orc_set_state(s, n, st.q.data(), st.p.data(), nullptr);
  if (!rc) rc = orc_make_pairlist(s);
  const long pairs0 = orc_pair_number(s);
  const long canon = orc_pairs_canonical(s, nullptr, 0);
  std::vector<int32_t> tri(3 * (canon > 0 ? canon : 1));
  Expect(orc_pairs_canonical(s, tri.data(), canon) == canon, "oracle canonical export");
  std::vector<double> f(3 * n);
  if (!rc) rc = orc_forces(s, f.data());
  int rebuilds = 0;
  if (!rc) rc = orc_calculate(s, steps, &rebuilds);
  orc_refresh_images(s);
  const double ke = orc_kinetic_energy(s), pe = orc_potential_energy(s), vir = orc_virial(s);
  orc_heatbath hb{};
  hb.type = 1;
  hb.aimed_temperature = 1.2;
  hb.tau = 0.1;
  orc_set_heatbath(s, &hb);
  int r2 = 0;
  if (!rc) rc = orc_calculate(s, 20, &r2);
  orc_get_heatbath(s, &hb);
  const double zeta = hb.zeta;
  hb.type = 2;
  hb.gamma = 0.1;
  hb.seed = 7;
  orc_set_heatbath(s, &hb);
  if (!rc) rc = orc_calculate(s, 20, &r2);
  std::vector<double> q(3 * n), p(3 * n);
  if (!rc) rc = orc_get_state(s, q.data(), p.data());
  double qsum = 0.0;
  for (double v : q) qsum += v;
  orc_destroy(s);
  std::printf("status %d\npairs %ld\ncanonical %ld\nrebuilds %d\nke %.17g\npe %.17g\nvirial %.17g\n"
              "zeta %.17g\nqsum %.17g\nf0 %.17g %.17g %.17g\nfailures %d\n",
              rc, pairs0, canon, rebuilds, ke, pe, vir, zeta, qsum, f[0], f[1], f[2], failures);
  return rc || failures ? 1 : 0;
}

int Stopwatch(int argc, char **argv) {
  if (argc < 3) return 2;
  const std::string base = std::string(argv[2]) + "/sw";
  {
    mdx::StopWatch sw(0, base.c_str());
    for (int k = 0; k < 3; k++) {
      sw.Start();
      volatile double acc = 0.0;
      for (int i = 0; i < 1000; i++) acc = acc + std::sqrt(static_cast<double>(i));
      sw.Stop();
    }
    Expect(sw.GetData().size() == 3 && sw.GetBackData() >= 0.0, "stopwatch laps");
  }   // saved on destruction (rank 0)
  std::printf("failures %d\n", failures);
  return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s capi|conf|oracle|stopwatch ...\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  if (mode == "capi") return Capi();
  if (mode == "conf") return Conf(argc, argv);
  if (mode == "oracle") return Oracle(argc, argv);
  if (mode == "stopwatch") return Stopwatch(argc, argv);
  return 2;
}
