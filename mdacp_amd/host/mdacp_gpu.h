// mdacp_gpu.h — the reference's hot-path C++ interface (the calls MDUnit
// makes, mdunit.h:59-102) implemented over libmdx (include/mdx.h).
//
// Names, arguments and call order follow the reference:
//   Variables          variables.h:12-49
//   SimulationInfo     simulationinfo.h:11-34 (the fields the path reads)
//   MDRect             mdrect.h:9-45
//   MeshList           meshlist.h:13-102
//   PairList           pairlist.h:10-31
//   ForceCalculator    fcalculator.h:10-44
//   *Observer          observer.h:8-38
//   Executor, MDUnit   mdunit.h:22-102
//   ParaInfo           parainfo.h:9-34 (one domain per GPU system)
// so that MDManager::Calculate (mdmanager.cc:154-186) reads unchanged:
//   if (plist->IsPairListExpired(vars, mesh, sinfo)) { plist->Init(...); mesh->MakeList(...); }
//   ForceCalculator::UpdatePositionHalf(vars, sinfo);
//   <image refresh>;  ForceCalculator::CalculateForce(vars, mesh, sinfo);
//   ForceCalculator::UpdatePositionHalf(vars, sinfo);
//
// What differs, because the state lives in HBM:
//   * q / p / type are host mirrors.  Variables::SyncToHost() refreshes them
//     from the device; after writing them on the host call
//     Variables::SyncToDevice().  AddParticle and SetInitialVelocity do this
//     bookkeeping themselves.
//   * Periodic images are made and refreshed on the device: MeshList::MakeList
//     builds them (the reference's MDManager::MakePairList halo exchange) and
//     MeshList::RefreshImages replaces MDManager::SendBorderParticles.
//   * Fatal conditions throw mdx::Error (status code + message) where the
//     reference calls exit(1) or assert.
#ifndef MDACP_GPU_H
#define MDACP_GPU_H

#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mdx.h"

namespace mdx {

constexpr int D = 4;  // mdconfig.h:9, lane 3 unused
constexpr int X = 0, Y = 1, Z = 2;
constexpr double CUTOFF_LENGTH = 3.0;  // mdconfig.h:31

struct Error : std::runtime_error {
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
  int code;
};

enum { HT_NOSEHOOVER, HT_LANGEVIN };   // heat bath types (mdconfig.h)

// the SimulationInfo fields on the path (simulationinfo.cc:5-51)
struct SimulationInfo {
  SimulationInfo(double lx, double ly, double lz, double time_step = 0.001);
  double L[D];
  double TimeStep;
  double SearchLength;   // CUTOFF_LENGTH + BufferLength
  double BufferLength;   // 0.3
  int CheckListLength;   // 500
  // heat bath (simulationinfo.cc:11-23)
  bool ControlTemperature = false;
  int HeatbathType = HT_NOSEHOOVER;
  double AimedTemperature = 1.0;
  double HeatbathTau = 0.1;
  double HeatbathGamma = 0.1;
  unsigned long long LangevinSeed = 0;   // noise stream (mdx_heatbath.seed)
  bool SortParticle = false;   // simulationinfo.cc:16 (MeshList::Sort)
  int Device = 0;
};

class MDRect {  // mdrect.h:9-45, one domain [0, L)
 public:
  MDRect() = default;
  MDRect(const double s2[D], const double e2[D]);
  double s[D] = {0, 0, 0, 0}, e[D] = {0, 0, 0, 0};
  double GetWidth(int dir) const { return e[dir] - s[dir]; }
  void ChangeScale(double alpha) {   // mdrect.cc:54-60
    for (int i = 0; i < D; i++) {
      s[i] *= alpha;
      e[i] *= alpha;
    }
  }
  bool IsInside(const double q[D]) const {
    return s[X] <= q[X] && q[X] < e[X] && s[Y] <= q[Y] && q[Y] < e[Y] && s[Z] <= q[Z] && q[Z] < e[Z];
  }
};

class Variables {
 public:
  explicit Variables(SimulationInfo *sinfo);
  ~Variables();
  Variables(const Variables &) = delete;
  Variables &operator=(const Variables &) = delete;

  // host mirrors (variables.h:19-28); see SyncToHost / SyncToDevice.
  // [0, particle_number) are the real particles; after SyncToHost, while a
  // pairlist exists, [particle_number, total) hold the periodic images (the
  // reference's ghosts): position at the last refresh, the source's momentum
  // and type.
  int *type = nullptr;
  double (*q)[D] = nullptr;
  double (*p)[D] = nullptr;
  double Zeta = 0.0;
  double SimulationTime = 0.0;

  double GetC0() const { return C0; }
  double GetC2() const { return C2; }
  int GetParticleNumber() const { return particle_number; }
  int GetTotalParticleNumber() const;
  // variables.h:39-40.  SetParticleNumber keeps the first pn particles (or
  // extends to pn with the host values held); SetTotalParticleNumber(pn)
  // drops the images until the next pairlist build, as MDManager::
  // MakePairList does before the halo exchange (any other total is an
  // error: images are made on the device).
  void SetParticleNumber(int pn);
  void SetTotalParticleNumber(int tn);
  // variables.cc:64-84 on the device, or on the host values if they are newer
  void AdjustPeriodicBoundary(SimulationInfo *sinfo);
  // variables.cc:138-145: q *= alpha for the real particles
  void ChangeScale(double alpha);
  void AddParticle(double x[D], double v[D], int t = 0);   // variables.cc:36-47
  void SetInitialVelocity(double V0, int seed);             // variables.cc:87-113
  // variables.cc:49-61: "q p" per line for the reals, then the ghosts
  // (here the periodic images, mdx_get_images), default stream format.
  // The reference's border exchange copies ghost positions only
  // (mdunit.cc:181-208), so its file shows whatever its ghost momentum
  // slots last held (0 in fresh storage); they are written as 0 here.
  void SaveConfiguration(std::ostream &fs);
  void SyncToHost();
  void SyncToDevice();
  mdx_system *Device();   // uploads pending host changes first

 private:
  void Grow(int need);
  SimulationInfo *sinfo;
  mdx_system *sys = nullptr;
  std::vector<double> qbuf, pbuf;
  std::vector<int> typebuf;
  bool images_dropped = false;
  std::vector<int> host_ids;   // type[] of each particle at the last upload
  int particle_number = 0;
  int capacity = 0;
  bool host_dirty = false;
  double C0, C2;
  friend class MeshList;
};

class MeshList {
 public:
  MeshList(SimulationInfo *sinfo, MDRect &r);
  // MeshList::ChangeScale (meshlist.cc:41-65): the cell grid of the new
  // rectangle (here the whole box of the GPU system, mdx_set_box)
  void ChangeScale(SimulationInfo *sinfo, MDRect &myrect);
  // MeshList::MakeList (meshlist.cc:68-104), including the periodic images
  // the reference builds in MDManager::MakePairList (mdmanager.cc:388-417)
  void MakeList(Variables *vars, SimulationInfo *sinfo, MDRect &myrect);
  // MeshList::Sort (meshlist.cc:140-173): with SortParticle, every
  // sort_interval-th construction, the particles are put in the order the
  // device keeps them (cell order of the last build).  type[] moves with
  // q and p here (the reference leaves it behind, an oversight kept out).
  void Sort(Variables *vars, SimulationInfo *sinfo, MDRect &myrect);
  // MDManager::SendBorderParticles (mdmanager.cc:420-429)
  void RefreshImages(Variables *vars);

  // The CSR half list of the last MakeList and its getters
  // (meshlist.h:66-88), fetched from the device on first use: indices in
  // Variables' host arrays, reals [0, pn) and images [pn, tn); key = the
  // smaller index of each pair; a pair across the box appears once from
  // each side with an image as partner (observers weight those by 1/2,
  // observer.cc:36-38).  Valid until the next MakeList.
  enum { KEY = 0, PARTNER = 1 };
  int (*GetKeyPartnerPairs())[2];
  int GetPairNumber();
  int GetPartnerNumber(int i);
  int GetKeyPointer(int i);
  int *GetSortedList();
  int *GetKeyPointerP();
  int *GetNumberOfPartners();
  void ShowPairs();                      // meshlist.cc:597-603
  void ShowSortedList(Variables *vars);  // meshlist.cc:605-615

  long GetCanonicalPairNumber(Variables *vars);   // distinct pairs, images resolved
  // canonical pair set, (id_lo, id_hi, image code) triples
  std::vector<int> GetPairsCanonical(Variables *vars);
  int GetNumberOfConstructions() const { return number_of_constructions; }
  void ClearNumberOfConstructions() { number_of_constructions = 0; }

 private:
  void Fetch();
  void ApplyBox(Variables *vars);
  int number_of_constructions = 0;
  int sort_interval = 10;   // meshlist.cc:14
  double box[3];
  bool box_pending = false;
  Variables *bound = nullptr;
  bool fetched = false;
  std::vector<int> key_pointer, number_of_partners, sorted_list, key_partner_pairs;
};

class PairList {  // pairlist.h:10-31
 public:
  void Init(Variables *vars, SimulationInfo *sinfo) { (void)vars; (void)sinfo; initialized = true; }
  bool IsPairListExpired(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);

 private:
  bool initialized = false;
};

namespace ForceCalculator {  // fcalculator.h:10-44
void CalculateForce(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void UpdatePositionHalf(Variables *vars, SimulationInfo *sinfo);
// heat bath (fcalculator.cc:979-1017); Variables::Zeta is the host value
void HeatbathZeta(Variables *vars, double current_temperature, SimulationInfo *sinfo);
void HeatbathMomenta(Variables *vars, SimulationInfo *sinfo);
void Langevin(Variables *vars, SimulationInfo *sinfo);
// forces at dt = 1 with p untouched, [pn][3] by particle id (test hook)
std::vector<double> Forces(Variables *vars, MeshList *mesh);
}  // namespace ForceCalculator

class IntegerObserver {  // observer.h:8-11
 public:
  virtual ~IntegerObserver() = default;
  virtual int Observe(Variables *vars, MeshList *mesh) = 0;
};
class ParticleNumberObserver : public IntegerObserver {  // observer.h:17-20
 public:
  int Observe(Variables *vars, MeshList *mesh = nullptr) override {
    (void)mesh;
    return vars->GetParticleNumber();
  }
};
class DoubleObserver {  // observer.h:13-16
 public:
  virtual ~DoubleObserver() = default;
  virtual double Observe(Variables *vars, MeshList *mesh) = 0;
};
class KineticEnergyObserver : public DoubleObserver {
 public:
  double Observe(Variables *vars, MeshList *mesh) override;
};
class PotentialEnergyObserver : public DoubleObserver {
 public:
  explicit PotentialEnergyObserver(SimulationInfo *si = nullptr) { (void)si; }
  double Observe(Variables *vars, MeshList *mesh) override;
};
class VirialObserver : public DoubleObserver {
 public:
  double Observe(Variables *vars, MeshList *mesh) override;
};

// StopWatch (stopwatch.h:84-130): one duration in seconds per Start/Stop,
// written one per line to <basename><rank as %05d>.dat when it is
// destroyed (rank 0 only, as the reference).  With a system given, Start
// and Stop first wait for its queued work, as the reference's GPU timer
// waits on its event.
class StopWatch {
 public:
  StopWatch(int rank, const char *basename, mdx_system *sys = nullptr)
      : id(rank), base(basename), sys(sys) {}
  ~StopWatch() {
    if (id == 0) SaveToFile();
  }
  StopWatch(const StopWatch &) = delete;
  StopWatch &operator=(const StopWatch &) = delete;
  void SetSystem(mdx_system *s) { sys = s; }
  void Start();
  void Stop();
  void SaveToFile() const;
  double GetBackData() const { return data.back(); }
  const std::vector<double> &GetData() const { return data; }

 private:
  int id;
  std::string base;
  mdx_system *sys;
  double start = 0.0;
  std::vector<double> data;
};

// The process/thread grid of the reference (parainfo.h:9-34) reduced to
// what MDUnit reads: grid size and a unit's grid position (x fastest, as
// ParaInfo::SetGrid numbers the units).  A GPU system is one periodic
// domain, so MDUnit accepts a grid of one unit only.
class ParaInfo {
 public:
  explicit ParaInfo(int gx = 1, int gy = 1, int gz = 1) : size{gx, gy, gz} {}
  void GetGridSize(int grid_size[D]) const {
    for (int d = 0; d < 3; d++) grid_size[d] = size[d];
    grid_size[3] = 1;
  }
  void GetGridPosition(int id, int grid_position[D]) const {
    grid_position[X] = id % size[X];
    grid_position[Y] = (id / size[X]) % size[Y];
    grid_position[Z] = id / size[X] / size[Y];
    grid_position[3] = 0;
  }
  bool IsValid() const { return size[X] > 0 && size[Y] > 0 && size[Z] > 0; }

 private:
  int size[3];
};

class MDUnit;
class Executor {  // mdunit.h:22-25
 public:
  virtual ~Executor() = default;
  virtual void Execute(MDUnit *mdu) { (void)mdu; }
};

// One periodic domain driven in the order of MDManager::Calculate
// (mdmanager.cc:154-186) and MDManager::MakePairList (:388-417).
class MDUnit {
 public:
  MDUnit(SimulationInfo *si);
  // mdunit.cc:8-33: the unit's rectangle from its grid position; the GPU
  // system covers the whole box, so the grid must be one unit
  MDUnit(int id, SimulationInfo *si, ParaInfo *pi);
  ~MDUnit();
  int GetID() const { return id; }
  MDRect *GetRect() { return &myrect; }
  Variables *GetVariables() { return vars; }
  MeshList *GetMeshList() { return mesh; }
  double *GetSystemSize() { return sinfo->L; }
  double GetTimeStep() { return sinfo->TimeStep; }
  void AddParticle(double x[D], double v[D], int type = 1);
  void AddParticle(double x[D], int type = 1);   // at rest (mdunit.cc:52-58)
  void SetInitialVelocity(double v0) { vars->SetInitialVelocity(v0, id); }
  int GetParticleNumber() { return vars->GetParticleNumber(); }
  int GetTotalParticleNumber() { return vars->GetTotalParticleNumber(); }
  void SetTotalParticleNumber(int n) { vars->SetTotalParticleNumber(n); }
  void CalculateForce() { ForceCalculator::CalculateForce(vars, mesh, sinfo); }
  void UpdatePositionHalf() { ForceCalculator::UpdatePositionHalf(vars, sinfo); }
  void HeatbathZeta(double t) { ForceCalculator::HeatbathZeta(vars, t, sinfo); }
  void HeatbathMomenta() { ForceCalculator::HeatbathMomenta(vars, sinfo); }
  void Langevin() { ForceCalculator::Langevin(vars, sinfo); }
  void AdjustPeriodicBoundary() { vars->AdjustPeriodicBoundary(sinfo); }
  int IntegerDouble(IntegerObserver *obs) { return obs->Observe(vars, mesh); }
  void Execute(Executor *ex) { ex->Execute(this); }
  void ShowPairs() { mesh->ShowPairs(); }
  // mdunit.cc:217-222: rectangle, positions and cell grid scaled by alpha
  // (MDManager::ChangeScale scales SimulationInfo::L and rebuilds around it)
  void ChangeScale(double alpha);
  void MakePairList();
  bool IsPairListExpired() { return plist->IsPairListExpired(vars, mesh, sinfo); }
  void Calculate();  // one step, NVE or under the heat bath of SimulationInfo
  // the timers MDManager::Calculate keeps (mdmanager.cc:155-185): all,
  // force, comm (image refresh), pair (rebuilds); their files are written
  // when the unit is destroyed
  void EnableStopWatches(int rank = 0);
  double ObserveDouble(DoubleObserver *obs) { return obs->Observe(vars, mesh); }
  // mdunit.cc:67-82: "0 type q p" per real, the CDView particle format
  void SaveAsCdview(std::ostream &ofs);
  // mdunit.cc:85-91: conf<id>.cnf (id as %03d) in the working directory
  void SaveConfiguration(int id = 0);
  double Temperature();        // KE/N/1.5 (mdmanager.cc:531-533)
  double PotentialEnergy();    // per particle (mdmanager.cc:551-555)
  double Pressure();           // (T - vir/N) N/V (mdmanager.cc:563-570)

 private:
  SimulationInfo *sinfo;
  Variables *vars;
  MeshList *mesh;
  PairList *plist;
  MDRect myrect;
  int id = 0;
  StopWatch *swAll = nullptr, *swForce = nullptr, *swComm = nullptr, *swPair = nullptr;
};

// BubbleHist of the Cavitation project (cavitation.cc:13-167) for one
// domain: the particle count of each box of a grid of about gsize per side
// (mdx_density_grid), kept as unsigned char; Analyse rebuilds the pairlist
// first as the reference's does, then SaveDensity writes the counts, box
// index ix + iy gx + iz gx gy, to conf%04d.dat (numbered by analysis).  The
// constructor prints the reference's grid information lines.
class BubbleHist {
 public:
  BubbleHist(MDUnit *unit, double gsize);
  void Analyse(MDUnit *unit);
  void SaveDensity() const;
  const std::vector<unsigned char> &GetData() const { return data; }
  int GetGridSize(int d) const { return grid[d]; }
  int GetAnalyseCount() const { return analyse_count; }

 private:
  double gsize_asked, grid_size;
  int grid[3];
  int analyse_count = 0;
  std::vector<unsigned char> data;
};

// MDManager::SaveAsCdview / SaveAsCdviewSequential (mdmanager.cc:118-143):
// the CDView file of a unit, and the numbered conf%04d.cd series
void SaveAsCdview(MDUnit &unit, const char *filename);
void SaveAsCdviewSequential(MDUnit &unit);

}  // namespace mdx
#endif
