// mdacp_gpu.cpp — see mdacp_gpu.h.
#include "mdacp_gpu.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>

namespace mdx {
namespace {

void check(mdx_system *s, int rc) {
  if (rc != MDX_OK) throw Error(rc, mdx_last_error(s));
}

}  // namespace

SimulationInfo::SimulationInfo(double lx, double ly, double lz, double time_step) {
  L[X] = lx;
  L[Y] = ly;
  L[Z] = lz;
  L[3] = 0.0;
  TimeStep = time_step;
  BufferLength = 0.3;                         // simulationinfo.cc:6
  SearchLength = CUTOFF_LENGTH + BufferLength;  // simulationinfo.cc:7
  CheckListLength = 500;                      // simulationinfo.cc:29
}

MDRect::MDRect(const double s2[D], const double e2[D]) {
  for (int i = 0; i < D; i++) {
    s[i] = s2[i];
    e[i] = e2[i];
  }
}

// ------------------------------------------------------------ Variables
Variables::Variables(SimulationInfo *si) : sinfo(si) {
  // variables.cc:21-27
  const double s2 = 1.0 / (CUTOFF_LENGTH * CUTOFF_LENGTH);
  const double s6 = s2 * s2 * s2, s8 = s6 * s2, s12 = s6 * s6, s14 = s12 * s2;
  C2 = 6.0 * s14 - 3.0 * s8;
  C0 = -s12 + s6 - C2 / s2;
}

Variables::~Variables() {
  if (sys) mdx_destroy(sys);
}

void Variables::Grow(int need) {
  if (need <= capacity) return;
  int cap = capacity ? capacity : 1024;
  while (cap < need) cap *= 2;
  qbuf.resize(static_cast<size_t>(cap) * D, 0.0);
  pbuf.resize(static_cast<size_t>(cap) * D, 0.0);
  typebuf.resize(cap, 0);
  type = typebuf.data();
  q = reinterpret_cast<double(*)[D]>(qbuf.data());
  p = reinterpret_cast<double(*)[D]>(pbuf.data());
  capacity = cap;
}

int Variables::GetTotalParticleNumber() const {
  if (!sys || host_dirty || images_dropped) return particle_number;
  int pn = 0, tn = 0;
  mdx_particle_number(sys, &pn, &tn);
  return tn;
}

void Variables::SetParticleNumber(int pn) {
  if (pn < 0) throw Error(MDX_ERR_ARG, "negative particle number");
  SyncToHost();
  Grow(pn);
  for (int i = particle_number; i < pn; i++) type[i] = 0;
  particle_number = pn;
  host_dirty = true;
}

void Variables::SetTotalParticleNumber(int tn) {
  if (tn == particle_number) {
    images_dropped = true;   // until the next pairlist build
    return;
  }
  if (tn != GetTotalParticleNumber())
    throw Error(MDX_ERR_ARG, "the images are made on the device: the total is the particle "
                             "number (images dropped) or the device's own total");
}

void Variables::AdjustPeriodicBoundary(SimulationInfo *si) {
  if (sys && !host_dirty) {
    check(sys, mdx_adjust_periodic_boundary(sys));
    return;
  }
  // on the newer host values, the rule of variables.cc:64-84
  for (int i = 0; i < particle_number; i++) {
    for (int d = 0; d < 3; d++) {
      if (q[i][d] < 0.0) q[i][d] += si->L[d];
      else if (q[i][d] > si->L[d]) q[i][d] -= si->L[d];
      if (q[i][d] < 0.0 || q[i][d] > si->L[d] || q[i][d] != q[i][d])
        throw Error(MDX_ERR_INVALID_POS, "Invalid position");
    }
  }
  host_dirty = true;
}

void Variables::ChangeScale(double alpha) {
  if (sys && !host_dirty) {
    check(sys, mdx_scale_positions(sys, alpha));
    return;
  }
  for (int i = 0; i < particle_number; i++)
    for (int d = 0; d < 3; d++) q[i][d] *= alpha;
  host_dirty = true;
}

void Variables::AddParticle(double x[D], double v[D], int t) {
  Grow(particle_number + 1);
  for (int d = 0; d < 3; d++) {
    q[particle_number][d] = x[d];
    p[particle_number][d] = v[d];
  }
  q[particle_number][3] = p[particle_number][3] = 0.0;
  type[particle_number] = t;
  particle_number++;
  host_dirty = true;
}

void Variables::SetInitialVelocity(const double V0, const int seed) {
  SyncToHost();
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  for (int i = 0; i < particle_number; i++) {
    const double z = ud(mt) * 2.0 - 1.0;
    const double phi = ud(mt) * 2.0 * M_PI;
    p[i][X] = V0 * std::sqrt(1 - z * z) * std::cos(phi);
    p[i][Y] = V0 * std::sqrt(1 - z * z) * std::sin(phi);
    p[i][Z] = V0 * z;
  }
  double av[3] = {0, 0, 0};
  for (int i = 0; i < particle_number; i++)
    for (int d = 0; d < 3; d++) av[d] += p[i][d];
  for (int d = 0; d < 3; d++) av[d] /= static_cast<double>(particle_number);
  for (int i = 0; i < particle_number; i++)
    for (int d = 0; d < 3; d++) p[i][d] -= av[d];
  host_dirty = true;
}

void Variables::SyncToDevice() {
  if (particle_number == 0) return;
  if (!sys) {
    mdx_params prm;
    mdx_default_params(&prm);
    for (int d = 0; d < 3; d++) prm.L[d] = sinfo->L[d];
    prm.dt = sinfo->TimeStep;
    prm.buffer_length = sinfo->BufferLength;
    prm.cutoff = sinfo->SearchLength - sinfo->BufferLength;
    prm.check_list_length = sinfo->CheckListLength;
    prm.device = sinfo->Device;
    int rc = mdx_create(&prm, &sys);
    if (rc != MDX_OK) throw Error(rc, mdx_last_error(nullptr));
  }
  std::vector<double> q3(3 * static_cast<size_t>(particle_number)), p3(q3.size());
  for (int i = 0; i < particle_number; i++)
    for (int d = 0; d < 3; d++) {
      q3[3 * i + d] = q[i][d];
      p3[3 * i + d] = p[i][d];
    }
  check(sys, mdx_set_state(sys, particle_number, q3.data(), p3.data()));
  // the device numbers particles by their position in this upload
  host_ids.assign(type, type + particle_number);
  host_dirty = false;
  images_dropped = false;
}

void Variables::SyncToHost() {
  if (!sys || host_dirty) return;
  std::vector<double> q3(3 * static_cast<size_t>(particle_number)), p3(q3.size());
  check(sys, mdx_get_state(sys, q3.data(), p3.data()));
  for (int i = 0; i < particle_number; i++) {
    for (int d = 0; d < 3; d++) {
      q[i][d] = q3[3 * i + d];
      p[i][d] = p3[3 * i + d];
    }
    type[i] = host_ids[i];
  }
  // the images (the reference's ghosts), while a pairlist exists
  if (images_dropped) return;
  int pn = 0, tn = 0;
  check(sys, mdx_particle_number(sys, &pn, &tn));
  const int g = tn - pn;
  if (g <= 0) return;
  std::vector<double> gq(3 * static_cast<size_t>(g)), gp(gq.size());
  std::vector<int32_t> gid(g);
  if (mdx_get_images(sys, gq.data(), gp.data(), gid.data(), nullptr) != MDX_OK) return;
  Grow(tn);
  for (int k = 0; k < g; k++) {
    for (int d = 0; d < 3; d++) {
      q[pn + k][d] = gq[3 * k + d];
      p[pn + k][d] = gp[3 * k + d];
    }
    q[pn + k][3] = p[pn + k][3] = 0.0;
    type[pn + k] = host_ids[gid[k]];
  }
}

void Variables::SaveConfiguration(std::ostream &fs) {
  SyncToHost();
  auto row = [&fs](const double *a, const double *b) {
    fs << a[X];
    fs << " " << a[Y];
    fs << " " << a[Z];
    fs << " " << b[X];
    fs << " " << b[Y];
    fs << " " << b[Z];
    fs << "\n";
  };
  for (int i = 0; i < particle_number; i++) row(q[i], p[i]);
  if (!sys) return;
  int pn = 0, tn = 0;
  check(sys, mdx_particle_number(sys, &pn, &tn));
  const int g = tn - pn;
  if (g <= 0) return;
  std::vector<double> gq(3 * static_cast<size_t>(g));
  if (mdx_get_images(sys, gq.data(), nullptr, nullptr, nullptr) != MDX_OK) return;  // no list yet
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < g; k++) row(&gq[3 * k], zero);
}

mdx_system *Variables::Device() {
  if (host_dirty || !sys) SyncToDevice();
  if (!sys) throw Error(MDX_ERR_STATE, "no particles");
  return sys;
}

// ------------------------------------------------------------ MeshList
MeshList::MeshList(SimulationInfo *sinfo, MDRect &r) {
  (void)sinfo;
  for (int d = 0; d < 3; d++) box[d] = r.GetWidth(d);
}

void MeshList::ChangeScale(SimulationInfo *sinfo, MDRect &myrect) {
  (void)sinfo;
  for (int d = 0; d < 3; d++) box[d] = myrect.GetWidth(d);
  box_pending = true;
  fetched = false;
  if (bound && bound->sys && !bound->host_dirty) ApplyBox(bound);
}

void MeshList::ApplyBox(Variables *vars) {
  if (!box_pending) return;
  mdx_system *s = vars->Device();
  check(s, mdx_set_box(s, box));
  box_pending = false;
}

void MeshList::MakeList(Variables *vars, SimulationInfo *sinfo, MDRect &myrect) {
  (void)sinfo;
  (void)myrect;
  bound = vars;
  ApplyBox(vars);
  mdx_system *s = vars->Device();
  check(s, mdx_make_pairlist(s));
  vars->images_dropped = false;
  fetched = false;
  number_of_constructions++;
}

void MeshList::Sort(Variables *vars, SimulationInfo *sinfo, MDRect &myrect) {
  (void)myrect;
  if (!sinfo->SortParticle) return;
  if (number_of_constructions % sort_interval != 0) return;
  if (!vars->sys || vars->host_dirty) return;   // nothing binned on the device yet
  const int pn = vars->GetParticleNumber();
  std::vector<int32_t> order(pn);
  if (mdx_get_cell_order(vars->sys, order.data()) != MDX_OK) return;   // no list yet
  vars->SyncToHost();
  std::vector<double> q2(static_cast<size_t>(pn) * D), p2(q2.size());
  std::vector<int> t2(pn);
  for (int i = 0; i < pn; i++) {
    const int j = order[i];
    for (int d = 0; d < D; d++) {
      q2[static_cast<size_t>(i) * D + d] = vars->q[j][d];
      p2[static_cast<size_t>(i) * D + d] = vars->p[j][d];
    }
    t2[i] = vars->type[j];
  }
  for (int i = 0; i < pn; i++) {
    for (int d = 0; d < D; d++) {
      vars->q[i][d] = q2[static_cast<size_t>(i) * D + d];
      vars->p[i][d] = p2[static_cast<size_t>(i) * D + d];
    }
    vars->type[i] = t2[i];
  }
  vars->host_dirty = true;   // uploaded in the new order by the next device call
  std::cout << "# Sorted!" << std::endl;
}

void MeshList::Fetch() {
  if (fetched) return;
  if (!bound) throw Error(MDX_ERR_STATE, "no pairlist: MakeList first");
  mdx_system *s = bound->Device();
  int pn = 0, tn = 0;
  check(s, mdx_particle_number(s, &pn, &tn));
  key_pointer.assign(tn, 0);
  number_of_partners.assign(tn, 0);
  int64_t c = 0;
  check(s, mdx_pairs_csr(s, key_pointer.data(), number_of_partners.data(), nullptr, 0, &c));
  sorted_list.assign(static_cast<size_t>(c), 0);
  check(s, mdx_pairs_csr(s, nullptr, nullptr, sorted_list.data(), c, &c));
  key_partner_pairs.resize(2 * static_cast<size_t>(c));
  for (int i = 0; i < tn; i++)
    for (int k = key_pointer[i]; k < key_pointer[i] + number_of_partners[i]; k++) {
      key_partner_pairs[2 * static_cast<size_t>(k)] = i;
      key_partner_pairs[2 * static_cast<size_t>(k) + 1] = sorted_list[k];
    }
  fetched = true;
}

int (*MeshList::GetKeyPartnerPairs())[2] {
  Fetch();
  return reinterpret_cast<int(*)[2]>(key_partner_pairs.data());
}
int MeshList::GetPairNumber() {
  Fetch();
  return static_cast<int>(sorted_list.size());
}
int MeshList::GetPartnerNumber(int i) {
  Fetch();
  return number_of_partners.at(i);
}
int MeshList::GetKeyPointer(int i) {
  Fetch();
  return key_pointer.at(i);
}
int *MeshList::GetSortedList() {
  Fetch();
  return sorted_list.data();
}
int *MeshList::GetKeyPointerP() {
  Fetch();
  return key_pointer.data();
}
int *MeshList::GetNumberOfPartners() {
  Fetch();
  return number_of_partners.data();
}

void MeshList::ShowPairs() {
  Fetch();
  for (size_t k = 0; k < sorted_list.size(); k++)
    std::printf("(%05d,%05d)\n", key_partner_pairs[2 * k], key_partner_pairs[2 * k + 1]);
}

void MeshList::ShowSortedList(Variables *vars) {
  Fetch();
  const int tn = vars->GetTotalParticleNumber();
  for (int i = 0; i < tn && i < static_cast<int>(key_pointer.size()); i++)
    for (int k = 0; k < number_of_partners[i]; k++)
      std::printf("(%05d,%05d)\n", i, sorted_list[key_pointer[i] + k]);
}

void MeshList::RefreshImages(Variables *vars) {
  mdx_system *s = vars->Device();
  check(s, mdx_refresh_images(s));
}

long MeshList::GetCanonicalPairNumber(Variables *vars) {
  mdx_system *s = vars->Device();
  int64_t c = 0;
  check(s, mdx_pair_number(s, &c));
  return static_cast<long>(c);
}

std::vector<int> MeshList::GetPairsCanonical(Variables *vars) {
  mdx_system *s = vars->Device();
  int64_t c = 0;
  check(s, mdx_pairs_canonical(s, nullptr, 0, &c));
  std::vector<int32_t> t(3 * static_cast<size_t>(c));
  check(s, mdx_pairs_canonical(s, t.data(), c, &c));
  std::vector<int> out(t.size());
  for (size_t k = 0; k < t.size(); k += 3) {   // device ids -> the caller's type[] ids
    out[k] = vars->host_ids[t[k]];
    out[k + 1] = vars->host_ids[t[k + 1]];
    out[k + 2] = t[k + 2];
  }
  return out;
}

// ------------------------------------------------------------ PairList
bool PairList::IsPairListExpired(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  (void)mesh;
  (void)sinfo;
  if (!initialized) return true;
  int expired = 0;
  mdx_system *s = vars->Device();
  check(s, mdx_is_pairlist_expired(s, &expired));
  return expired != 0;
}

// ------------------------------------------------------------ ForceCalculator
namespace ForceCalculator {
void CalculateForce(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  (void)mesh;
  (void)sinfo;
  mdx_system *s = vars->Device();
  check(s, mdx_calculate_force(s));
}

void UpdatePositionHalf(Variables *vars, SimulationInfo *sinfo) {
  (void)sinfo;
  mdx_system *s = vars->Device();
  check(s, mdx_update_position_half(s));
}

// the heat bath of sinfo with zeta, handed to the device
static void PushHeatbath(Variables *vars, SimulationInfo *sinfo, int type) {
  mdx_system *s = vars->Device();
  mdx_heatbath hb{};
  hb.type = type;
  hb.aimed_temperature = sinfo->AimedTemperature;
  hb.tau = sinfo->HeatbathTau;
  hb.gamma = sinfo->HeatbathGamma;
  hb.zeta = vars->Zeta;
  hb.seed = sinfo->LangevinSeed;
  check(s, mdx_set_heatbath(s, &hb));
}

// fcalculator.cc:979-985, on the host as the reference keeps Zeta
void HeatbathZeta(Variables *vars, double current_temperature, SimulationInfo *sinfo) {
  const double dt2 = sinfo->TimeStep * 0.5;
  const double tau = sinfo->HeatbathTau;
  const double t1 = (current_temperature - sinfo->AimedTemperature) / (tau * tau);
  vars->Zeta += t1 * dt2;
}

// fcalculator.cc:987-999
void HeatbathMomenta(Variables *vars, SimulationInfo *sinfo) {
  PushHeatbath(vars, sinfo, MDX_HEATBATH_NOSE_HOOVER);
  mdx_system *s = vars->Device();
  check(s, mdx_heatbath_momenta(s));
}

// fcalculator.cc:1001-1017 (the noise: mdx_langevin)
void Langevin(Variables *vars, SimulationInfo *sinfo) {
  PushHeatbath(vars, sinfo, MDX_HEATBATH_LANGEVIN);
  mdx_system *s = vars->Device();
  check(s, mdx_langevin(s));
}

std::vector<double> Forces(Variables *vars, MeshList *mesh) {
  (void)mesh;
  mdx_system *s = vars->Device();
  const int n = vars->GetParticleNumber();
  std::vector<double> f(3 * static_cast<size_t>(n));
  check(s, mdx_forces(s, f.data()));
  return f;
}
}  // namespace ForceCalculator

// ------------------------------------------------------------ observers
double KineticEnergyObserver::Observe(Variables *vars, MeshList *) {
  mdx_system *s = vars->Device();
  double ke = 0;
  check(s, mdx_observe(s, &ke, nullptr, nullptr));
  return ke;
}

double PotentialEnergyObserver::Observe(Variables *vars, MeshList *) {
  mdx_system *s = vars->Device();
  double pe = 0;
  check(s, mdx_observe(s, nullptr, &pe, nullptr));
  return pe;
}

double VirialObserver::Observe(Variables *vars, MeshList *) {
  mdx_system *s = vars->Device();
  double vir = 0;
  check(s, mdx_observe(s, nullptr, nullptr, &vir));
  return vir;
}

// ------------------------------------------------------------ MDUnit
MDUnit::MDUnit(SimulationInfo *si) : sinfo(si) {
  vars = new Variables(si);
  plist = new PairList();
  double s[D] = {0, 0, 0, 0}, e[D] = {si->L[X], si->L[Y], si->L[Z], 0};
  myrect = MDRect(s, e);
  mesh = new MeshList(si, myrect);
}

MDUnit::MDUnit(int id_, SimulationInfo *si, ParaInfo *pi) : sinfo(si), id(id_) {
  int size[D], pos[D];
  pi->GetGridSize(size);
  pi->GetGridPosition(id_, pos);
  double s[D] = {0, 0, 0, 0}, e[D] = {0, 0, 0, 0};
  for (int d = 0; d < 3; d++) {   // mdunit.cc:21-26
    const double ul = si->L[d] / static_cast<double>(size[d]);
    s[d] = ul * static_cast<double>(pos[d]);
    e[d] = s[d] + ul;
  }
  if (size[X] != 1 || size[Y] != 1 || size[Z] != 1)
    throw Error(MDX_ERR_ARG, "a GPU system is one periodic domain: the grid must be one unit");
  myrect = MDRect(s, e);
  vars = new Variables(si);
  plist = new PairList();
  mesh = new MeshList(si, myrect);
}

MDUnit::~MDUnit() {
  // the timers write their files (rank 0), as the reference's static ones
  // do at exit
  delete swAll;
  delete swForce;
  delete swComm;
  delete swPair;
  delete mesh;
  delete plist;
  delete vars;
}

void MDUnit::AddParticle(double x[D], double v[D], int type) {
  if (myrect.IsInside(x)) vars->AddParticle(x, v, type);   // mdunit.cc:61-65
}

void MDUnit::AddParticle(double x[D], int type) {
  double v[D] = {0.0, 0.0, 0.0, 0.0};
  AddParticle(x, v, type);
}

void MDUnit::MakePairList() {   // MDUnit::MakePairList (mdunit.cc:211-215)
  mesh->Sort(vars, sinfo, myrect);
  plist->Init(vars, sinfo);
  mesh->MakeList(vars, sinfo, myrect);
}

void MDUnit::ChangeScale(double alpha) {   // mdunit.cc:217-222
  myrect.ChangeScale(alpha);
  vars->ChangeScale(alpha);
  mesh->ChangeScale(sinfo, myrect);
}

// MDManager::Calculate (mdmanager.cc:154-186) with CalculateNoseHoover
// (:255-287) and CalculateLangevin (:289-315) for one domain
// ------------------------------------------------------------ StopWatch
namespace {
double WallSeconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void StopWatch::Start() {
  if (sys) check(sys, mdx_synchronize(sys));
  start = WallSeconds();
}

void StopWatch::Stop() {
  if (sys) check(sys, mdx_synchronize(sys));
  data.push_back(WallSeconds() - start);
}

void StopWatch::SaveToFile() const {
  char filename[256];
  std::snprintf(filename, sizeof filename, "%s%05d.dat", base.c_str(), id);
  std::ofstream ofs(filename);
  for (double d : data) ofs << d << "\n";
}

void MDUnit::EnableStopWatches(int rank) {
  if (swAll) return;
  swAll = new StopWatch(rank, "all");
  swForce = new StopWatch(rank, "force");
  swComm = new StopWatch(rank, "comm");
  swPair = new StopWatch(rank, "pair");
}

void MDUnit::Calculate() {
  // the device system exists once particles are loaded; the timers wait on it
  mdx_system *dev = swAll ? vars->Device() : nullptr;
  for (StopWatch *w : {swAll, swForce, swComm, swPair})
    if (w) w->SetSystem(dev);
  if (swAll) swAll->Start();
  if (IsPairListExpired()) {
    if (swPair) swPair->Start();
    MakePairList();
    if (swPair) swPair->Stop();
  }
  ForceCalculator::UpdatePositionHalf(vars, sinfo);
  if (swComm) swComm->Start();
  mesh->RefreshImages(vars);
  if (swComm) swComm->Stop();
  if (swForce && !sinfo->ControlTemperature) swForce->Start();
  if (sinfo->ControlTemperature && sinfo->HeatbathType == HT_NOSEHOOVER) {
    ForceCalculator::HeatbathZeta(vars, Temperature(), sinfo);
    ForceCalculator::HeatbathMomenta(vars, sinfo);
    ForceCalculator::CalculateForce(vars, mesh, sinfo);
    ForceCalculator::HeatbathMomenta(vars, sinfo);
    ForceCalculator::HeatbathZeta(vars, Temperature(), sinfo);
    ForceCalculator::UpdatePositionHalf(vars, sinfo);
  } else if (sinfo->ControlTemperature) {
    ForceCalculator::CalculateForce(vars, mesh, sinfo);
    ForceCalculator::Langevin(vars, sinfo);
  } else {
    ForceCalculator::CalculateForce(vars, mesh, sinfo);
    ForceCalculator::UpdatePositionHalf(vars, sinfo);
  }
  if (swForce && !sinfo->ControlTemperature) swForce->Stop();
  vars->SimulationTime += sinfo->TimeStep;
  if (swAll) swAll->Stop();
}

void MDUnit::SaveAsCdview(std::ostream &ofs) {
  vars->SyncToHost();
  for (int i = 0; i < vars->GetParticleNumber(); i++) {
    ofs << "0 ";
    ofs << vars->type[i] << " ";
    ofs << vars->q[i][X] << " ";
    ofs << vars->q[i][Y] << " ";
    ofs << vars->q[i][Z] << " ";
    ofs << vars->p[i][X] << " ";
    ofs << vars->p[i][Y] << " ";
    ofs << vars->p[i][Z] << "\n";
  }
}

void MDUnit::SaveConfiguration(int id) {
  char filename[256];
  std::snprintf(filename, sizeof filename, "conf%03d.cnf", id);
  std::ofstream fs(filename);
  vars->SaveConfiguration(fs);
}

void SaveAsCdview(MDUnit &unit, const char *filename) {
  std::ofstream ofs(filename);
  unit.SaveAsCdview(ofs);
}

void SaveAsCdviewSequential(MDUnit &unit) {
  static int index = 0;
  char filename[256];
  std::snprintf(filename, sizeof filename, "conf%04d.cd", index);
  index++;
  SaveAsCdview(unit, filename);
}

// ------------------------------------------------------------ BubbleHist
BubbleHist::BubbleHist(MDUnit *unit, double gsize) : gsize_asked(gsize) {
  MDRect *r = unit->GetRect();   // cavitation.cc:41-57
  for (int d = 0; d < 3; d++) grid[d] = static_cast<int>(r->GetWidth(d) / gsize + 0.5);
  grid_size = r->GetWidth(X) / static_cast<double>(grid[0]);
  data.assign(static_cast<size_t>(grid[0]) * grid[1] * grid[2], 0);
  std::cout << "# Grid Information ------" << std::endl;
  std::cout << "# grid_size = " << grid_size << std::endl;
  std::cout << "# grid: " << grid[0] << "," << grid[1] << "," << grid[2] << std::endl;
  std::cout << "# grid_count = " << data.size() << std::endl;
  std::cout << "# ------" << std::endl;
}

void BubbleHist::Analyse(MDUnit *unit) {   // cavitation.cc:142-166
  unit->MakePairList();
  mdx_system *s = unit->GetVariables()->Device();
  int32_t dims[3];
  check(s, mdx_density_grid(s, gsize_asked, dims, data.data(), static_cast<int64_t>(data.size())));
  SaveDensity();
  analyse_count++;
}

void BubbleHist::SaveDensity() const {   // cavitation.cc:85-90
  char filename[256];
  std::snprintf(filename, sizeof filename, "conf%04d.dat", analyse_count);
  std::ofstream ofs(filename, std::ios::binary);
  ofs.write(reinterpret_cast<const char *>(data.data()), static_cast<std::streamsize>(data.size()));
}

double MDUnit::Temperature() {
  KineticEnergyObserver obs;
  return obs.Observe(vars, mesh) / vars->GetParticleNumber() / 1.5;
}

double MDUnit::PotentialEnergy() {
  PotentialEnergyObserver obs(sinfo);
  return obs.Observe(vars, mesh) / vars->GetParticleNumber();
}

double MDUnit::Pressure() {
  VirialObserver obs;
  const double pn = vars->GetParticleNumber();
  const double phi = obs.Observe(vars, mesh) / pn;
  const double T = Temperature();
  const double V = sinfo->L[X] * sinfo->L[Y] * sinfo->L[Z];
  return (T - phi) * pn / V;
}

}  // namespace mdx
