/*
 * mdx.h — C ABI of libmdx, the MI355X (gfx950) short-range Lennard-Jones
 * step of MDACP: image construction, cell binning, pairlist build, force
 * kick with virial / potential energy, half drift and pairlist expiry, all
 * resident in HBM of one GPU.
 *
 * The reference has no C ABI for this path; its operator interface is the
 * C++ triple that MDUnit calls (mdunit.h:59-102).  Each entry point below
 * names the reference call it replaces.  The C++ mirror of that interface
 * (mdacp_amd/host/mdacp_gpu.h: Variables, MeshList, ForceCalculator,
 * PairList, observers, MDUnit) is built on these functions.
 *
 * Conventions
 *   - Every function returns an int status: 0 = MDX_OK, negative = error.
 *     Nothing calls exit(); mdx_last_error() gives the message.  The fatal
 *     exits of the reference map onto status codes (see enum mdx_status).
 *   - Host arrays are caller-owned, [n][3] doubles in particle-id order
 *     (id = position in the array given to mdx_set_state).  On the device
 *     particles are kept in cell order; ids travel with them.
 *   - All device work is ordered on the handle's HIP stream.  A handle is
 *     not thread-safe; independent handles are.
 *   - Units: sigma = epsilon = m = k_B = 1 (hard-wired in the reference,
 *     variables.cc:16-33, mdconfig.h:31).
 */
#ifndef MDX_H
#define MDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdx_system mdx_system;

typedef struct mdx_params {
  double L[3];            /* SimulationInfo::L (simulationinfo.cc:31-49) */
  double dt;              /* SimulationInfo::TimeStep (simulationinfo.cc:9) */
  double cutoff;          /* CUTOFF_LENGTH (mdconfig.h:31), 3.0 */
  double buffer_length;   /* SimulationInfo::BufferLength (simulationinfo.cc:6), 0.3 */
  int check_list_length;  /* SimulationInfo::CheckListLength (simulationinfo.cc:29), 500 */
  int capacity;           /* max particles; 0 = sized by the first mdx_set_state */
  int device;             /* HIP device ordinal */
  void *stream;           /* hipStream_t to order work on; NULL = own stream */
} mdx_params;

enum mdx_status {
  MDX_OK = 0,
  MDX_ERR_ARG = -1,          /* bad argument */
  MDX_ERR_CAPACITY = -2,     /* exceeds capacity (variables.cc:46, meshlist.cc:582) */
  MDX_ERR_INVALID_POS = -3,  /* "Invalid position" (variables.cc:77-81) or non-finite */
  MDX_ERR_TOO_FAST = -4,     /* "Too fast particles" (pairlist.cc:59-63) */
  MDX_ERR_BOX = -5,          /* box edge < 2 search lengths: images would alias */
  MDX_ERR_HIP = -6,          /* HIP runtime error */
  MDX_ERR_STATE = -7         /* call out of order (e.g. force before pairlist) */
};

/* Fills the reference defaults (dt 0.001, cutoff 3.0, buffer 0.3, 500). */
void mdx_default_params(mdx_params *prm);

/* MDUnit::MDUnit (mdunit.cc:8-39): Variables + MeshList + PairList of one
 * periodic domain, here one GPU-resident system. */
int mdx_create(const mdx_params *prm, mdx_system **out);
/* MDUnit::~MDUnit (mdunit.cc:41-49) */
int mdx_destroy(mdx_system *sys);
/* message of the last failure on sys (or of mdx_create when sys is NULL) */
const char *mdx_last_error(const mdx_system *sys);

/* Variables::AddParticle for all n particles at once (variables.cc:36-47);
 * resets the pairlist (PairList ctor, pairlist.h:25). */
int mdx_set_state(mdx_system *sys, int n, const double *q3, const double *p3);
/* Variables::q / p read back, ordered by particle id; either may be NULL. */
int mdx_get_state(mdx_system *sys, double *q3, double *p3);
/* Variables::GetParticleNumber / GetTotalParticleNumber (variables.h:25-26) */
int mdx_particle_number(const mdx_system *sys, int *pn, int *tn);
/* The periodic images, as the reference keeps its ghosts in Variables
 * beyond particle_number (q, p, type, variables.h:19-28; a ghost carries
 * its source's momentum, mdunit.cc:181-191): position at the last refresh,
 * the source's momentum, the source's id and the image shift code
 * (sx+1) + 3(sy+1) + 9(sz+1).  Each array holds tn - pn entries
 * (mdx_particle_number), in the library's image order; any may be NULL.
 * Valid once a pairlist exists. */
int mdx_get_images(mdx_system *sys, double *q3, double *p3, int32_t *id, int32_t *code);

/* Waits until the work queued on sys has finished (the reference's GPU
 * StopWatch waits on its CUDA event before it reads the time,
 * stopwatch.h:48-82).  Every other call here that returns data waits by
 * itself; this one is for timing. */
int mdx_synchronize(mdx_system *sys);

/* MDManager::MakePairList (mdmanager.cc:388-417): wrap into the box
 * (variables.cc:64-84), build periodic images on the device (replacing
 * mdmanager.cc:395-401), bin by counting sort (meshlist.cc:176-228), build
 * the neighbour list (meshlist.cc:68-104, 231-415) and snapshot positions
 * (PairList::Init, pairlist.cc:11-23). */
int mdx_make_pairlist(mdx_system *sys);
/* PairList::IsPairListExpired (pairlist.cc:26-121) */
int mdx_is_pairlist_expired(mdx_system *sys, int *expired);
/* ForceCalculator::UpdatePositionHalf (fcalculator.cc:18-30) */
int mdx_update_position_half(mdx_system *sys);
/* MDManager::SendBorderParticles (mdmanager.cc:420-429): image refresh */
int mdx_refresh_images(mdx_system *sys);
/* ForceCalculator::CalculateForce (fcalculator.cc:33-50): p += F dt */
int mdx_calculate_force(mdx_system *sys);
/* MDManager::Calculate (mdmanager.cc:154-186), nsteps times: NVE, or
 * under the heat bath set by mdx_set_heatbath (CalculateNoseHoover /
 * CalculateLangevin, mdmanager.cc:255-315).  *rebuilds (may be NULL)
 * receives the number of pairlist constructions. */
int mdx_calculate(mdx_system *sys, int nsteps, int *rebuilds);

/* Heat bath: SimulationInfo::ControlTemperature / HeatbathType /
 * AimedTemperature / HeatbathTau / HeatbathGamma (simulationinfo.cc:11-23)
 * and Variables::Zeta (variables.h:29).
 *   MDX_HEATBATH_NONE         NVE (ControlTemperature = no)
 *   MDX_HEATBATH_NOSE_HOOVER  zeta += (T - T0) / tau^2 dt/2 from the
 *                             temperature KE/N/1.5, p *= exp(-zeta dt/2),
 *                             before and after the kick (mdmanager.cc:255-287)
 *   MDX_HEATBATH_LANGEVIN     after the kick p += (-gamma p + r) dt with
 *                             r ~ N(0, 2 gamma T0 / dt) on all four lanes
 *                             (the padding lane too, which then enters KE),
 *                             and no trailing half drift (mdmanager.cc:289-315)
 * The reference draws Langevin noise from a per-thread std::mt19937
 * (fcalculator.cc:1001-1017), which fixes no particle-to-number map; here
 * the noise of particle `id` in the k-th Langevin kick since mdx_set_state
 * (k from 0) is a pure function of (seed, id, k): Philox4x32-10 blocks with counter
 * (id, k lo, k hi, 0 / 1) and key (seed lo, seed hi), eight words -> four
 * 53-bit uniforms in (0, 1] -> two Box-Muller pairs.  The same stream
 * whatever the binning or the GPU.  Zeta is kept on the device;
 * mdx_get_heatbath reads it back. */
enum {
  MDX_HEATBATH_NONE = 0,
  MDX_HEATBATH_NOSE_HOOVER = 1,
  MDX_HEATBATH_LANGEVIN = 2
};
typedef struct mdx_heatbath {
  int type;                  /* MDX_HEATBATH_* */
  double aimed_temperature;  /* AimedTemperature, 1.0 */
  double tau;                /* HeatbathTau, 0.1 (Nose-Hoover, > 0) */
  double gamma;              /* HeatbathGamma, 0.1 (Langevin, >= 0) */
  double zeta;               /* Variables::Zeta (Nose-Hoover state) */
  uint64_t seed;             /* Langevin noise stream */
} mdx_heatbath;
int mdx_set_heatbath(mdx_system *sys, const mdx_heatbath *hb);
int mdx_get_heatbath(mdx_system *sys, mdx_heatbath *hb);
/* ForceCalculator::HeatbathZeta (fcalculator.cc:979-985): zeta += (t -
 * T0) / tau^2 dt/2 for the given current temperature t */
int mdx_heatbath_zeta(mdx_system *sys, double current_temperature);
/* ForceCalculator::HeatbathMomenta (fcalculator.cc:987-999): p *= exp(-zeta dt/2) */
int mdx_heatbath_momenta(mdx_system *sys);
/* ForceCalculator::Langevin (fcalculator.cc:1001-1017): the next Langevin
 * kick (its noise counter advances by one) */
int mdx_langevin(mdx_system *sys);

/* KineticEnergyObserver / PotentialEnergyObserver / VirialObserver
 * (observer.cc:5-79): sums over the system (not per particle); KE over all
 * four momentum lanes as the reference's (the fourth is zero unless the
 * Langevin bath ran); virial is sum r V'(r) / 3.  Any pointer may be NULL. */
int mdx_observe(mdx_system *sys, double *ke, double *pe, double *virial);
/* Forces F_i = sum_j f_ij of the current list, [n][3] by id; p untouched
 * (the reference's kick at dt = 1 on p = 0, fcalculator.cc:182-223). */
int mdx_forces(mdx_system *sys, double *f3);
/* number of distinct pairs within the search length (each pair once,
 * images resolved), i.e. the canonical pair set's size */
int mdx_pair_number(mdx_system *sys, int64_t *count);
/* canonical pair set: sorted (id_lo, id_hi, image code) int32 triples,
 * code = (sx+1) + 3(sy+1) + 9(sz+1) for the partner image q_hi + s.L.
 * Writes min(cap, count) triples; *count receives the full count. */
int mdx_pairs_canonical(mdx_system *sys, int32_t *out, int64_t cap, int64_t *count);

/* MeshList::MakeList's CSR half list and its getters (meshlist.cc:68-104,
 * 556-583; meshlist.h:72-88: GetKeyPartnerPairs, GetPairNumber,
 * GetPartnerNumber, GetKeyPointer, GetSortedList, GetKeyPointerP,
 * GetNumberOfPartners) in the reference's index space: real particles by id
 * in [0, pn), periodic image k as pn + k in mdx_get_images order (the
 * reference's ghosts beyond particle_number).  Each pair has key = its
 * smaller index: a pair of reals within the search length once, a real and
 * an image from the real's side, so a pair across the box appears once from
 * each side with an image as partner, as the reference's real-ghost pairs do
 * (its observers weight those by 1/2, observer.cc:36-38); image-image pairs
 * none.  number_of_partners and key_pointer ([tn] each, may be NULL) are
 * the per-key counts and their exclusive prefix sum (images: 0 partners);
 * sorted_list receives the partners key by key, min(cap, count) of them,
 * and *count the total, MeshList::GetPairNumber. */
int mdx_pairs_csr(mdx_system *sys, int32_t *key_pointer, int32_t *number_of_partners,
                  int32_t *sorted_list, int64_t cap, int64_t *count);
/* Variables::AdjustPeriodicBoundary (variables.cc:64-84): q < 0 -> q + L,
 * q > L -> q - L per axis; MDX_ERR_INVALID_POS ("Invalid position") if a
 * coordinate is still outside.  The pairlist is void afterwards (the
 * reference wraps only inside MDManager::MakePairList). */
int mdx_adjust_periodic_boundary(mdx_system *sys);
/* Variables::ChangeScale (variables.cc:138-145): q *= alpha for the real
 * particles, momenta unchanged; the pairlist is void afterwards. */
int mdx_scale_positions(mdx_system *sys, double alpha);
/* MeshList::ChangeScale (meshlist.cc:41-65) with SimulationInfo::L: a new
 * box L3[3], its cell grid (floor(L / search length) cells per axis) and
 * arrays; MDX_ERR_BOX if an edge is below 2 search lengths.  The pairlist is
 * void afterwards.  MDUnit::ChangeScale (mdunit.cc:218-222) is
 * mdx_scale_positions(alpha) then mdx_set_box(alpha L). */
int mdx_set_box(mdx_system *sys, const double *L3);
/* The density grid of the Cavitation project's BubbleHist (cavitation.cc:
 * 38-80, 117-141, one domain at the origin): g_d = (int)(L_d / grid_size +
 * 0.5) boxes per axis, one box width w = L_x / g_x for every axis, each
 * real particle counted in box (trunc(x / w), trunc(y / w), trunc(z / w))
 * at index ix + iy g_x + iz g_x g_y, the count kept as the reference's
 * unsigned char (modulo 256).  dims[3] receives g; data (may be NULL: only
 * the dimensions) receives the g_x g_y g_z bytes that BubbleHist::
 * SaveDensity writes (cavitation.cc:85-90).  Positions are taken as they
 * stand (the reference's Analyse first calls MakePairList, which wraps
 * them); a particle outside the grid is MDX_ERR_INVALID_POS (the
 * reference's assert). */
int mdx_density_grid(mdx_system *sys, double grid_size, int32_t *dims, uint8_t *data, int64_t cap);
/* The ids of the real particles in the order the device keeps them (cell
 * order of the last pairlist build, by x inside a cell where the reference
 * keeps index order), pn entries: the permutation MeshList::Sort applies
 * to Variables (meshlist.cc:140-173). */
int mdx_get_cell_order(mdx_system *sys, int32_t *ids);

/* Allocation-free variant of mdx_pairs_canonical.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as mdx_pairs_canonical (the
 * canonical triples, sorted; min(cap, count) of them written; *count the
 * full count; out = NULL or cap = 0 only counts), but the triples are
 * staged and sorted in a fixed device array of MDX_NOALLOC_PAIRS pairs
 * (12 MiB, held by every GPU the library runs on until the process ends)
 * instead of a buffer allocated per call, and nothing is allocated on the
 * device or the host.  A pair set larger than MDX_NOALLOC_PAIRS returns
 * MDX_ERR_CAPACITY (with *count set).  Every system on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two streams, two
 * threads) race; calls ordered one after another do not. */
#define MDX_NOALLOC_PAIRS (1 << 20)
int mdx_pairs_canonical_noalloc(mdx_system *sys, int32_t *out, int64_t cap, int64_t *count);

/* Differentiable form (no counterpart in the reference, whose epsilon and
 * sigma are hard-wired to 1, variables.cc:16-33): V(r) = eps v(r / sigma)
 * with v the reference's shifted LJ (variables.cc:21-27) cut at 3 sigma.
 * From the loaded state, nsteps drift-kick-drift steps; over the last nobs
 * of them (observed after each step)
 *   loss = (mean P - p_target)^2 + (mean PE/N - pe_target)^2,
 *   P = (2 KE / 3 - W) / V  (MDManager::Pressure, mdmanager.cc:563-570),
 *   W = sum r V'(r) / 3     (VirialObserver, observer.cc:50-79).
 * The pairlist is rebuilt every checkpoint_every steps (and must stay valid
 * over each interval: MDX_ERR_STATE otherwise).  With grad != 0 the
 * adjoint (Hessian-vector products of the forces, dF/deps, dF/dsigma)
 * runs backwards over the checkpoints, recomputing each interval, and
 * fills dloss_deps, dloss_dsigma and, if dloss_dq0 is not NULL, dL/dq0 as
 * [n][3] by id.  Leaves the system at the final state without a list. */
typedef struct mdx_lj_result {
  double loss;
  double pressure_mean;
  double pe_mean;         /* per particle */
  double dloss_deps;
  double dloss_dsigma;
  double forward_ms;      /* wall time of the forward pass */
  double backward_ms;     /* wall time of the reverse sweep (0 without grad) */
} mdx_lj_result;

int mdx_lj_trajectory(mdx_system *sys, double eps, double sigma, int nsteps, int nobs,
                      int checkpoint_every, double p_target, double pe_target, int grad,
                      mdx_lj_result *out, double *dloss_dq0);

/* ------------------------------------------------------------------
 * Replica ensemble (no counterpart in the reference, which runs one
 * system per MPI job): nrep independent periodic systems of uneven sizes,
 * densities and temperatures advanced together on one GPU.  Each replica
 * runs exactly the step of mdx_calculate on its own (binning, images,
 * pairlist expiry, kick arithmetic), so its trajectory is the one a
 * separate mdx_system would produce; one captured HIP graph per step
 * advances all of them, and a replica whose pairlist expired is rebuilt
 * inside that graph while the others keep stepping.
 *
 * Particles of all replicas are concatenated in replica order (replica r
 * owns ids [sum_{s<r} n_s, sum_{s<=r} n_s)); L3 is [nrep][3].  prm gives
 * dt, cutoff, buffer_length, check_list_length (<= 1024), device and
 * stream; prm->L and prm->capacity are not used. */
typedef struct mdx_ensemble mdx_ensemble;

typedef struct mdx_ensemble_config {
  int tile;            /* reals aimed at per tile at first: 512 or 256 (a tile holds at most 512) */
  int grid_mult;       /* persistent step grid = grid_mult x resident workgroups */
  int step_grid;       /* workgroups of the step kernel */
  int stage_cap;       /* particles one tile may stage in LDS */
  double ms_per_step;  /* device time per step of the last mdx_ensemble_calculate */
} mdx_ensemble_config;

int mdx_ensemble_create(const mdx_params *prm, int nrep, const int *n, const double *L3,
                        mdx_ensemble **out);
int mdx_ensemble_destroy(mdx_ensemble *e);
const char *mdx_ensemble_last_error(const mdx_ensemble *e);
int mdx_ensemble_size(const mdx_ensemble *e, int *nrep, int *ntotal);
/* loads every replica (q3, p3 are [ntotal][3]) and builds the first
 * pairlists (the construction the first step of mdx_calculate makes) */
int mdx_ensemble_set_state(mdx_ensemble *e, const double *q3, const double *p3);
int mdx_ensemble_get_state(mdx_ensemble *e, double *q3, double *p3);
/* nsteps NVE steps of every running replica; rebuilds (may be NULL,
 * [nrep]) receives each replica's pairlist constructions since
 * mdx_ensemble_set_state.  A replica that fails (too fast, invalid
 * position, capacity) stops; the others go on; the first failure is
 * returned and mdx_ensemble_status lists all.  A replica left short of its
 * steps after the capacity regrowth rounds also returns MDX_ERR_CAPACITY
 * (never MDX_OK with steps missing). */
int mdx_ensemble_calculate(mdx_ensemble *e, int nsteps, int *rebuilds);
int mdx_ensemble_status(mdx_ensemble *e, int *status, int *steps);
/* per replica KE, PE, virial ([nrep] each, any may be NULL), as
 * mdx_observe of each replica alone */
int mdx_ensemble_observe(mdx_ensemble *e, double *ke, double *pe, double *vir);
int mdx_ensemble_get_config(const mdx_ensemble *e, mdx_ensemble_config *cfg);
/* tile (512 or 256) changes only before the first step after mdx_ensemble_set_state */
int mdx_ensemble_set_config(mdx_ensemble *e, int tile, int grid_mult);
/* times trial_steps steps of every tile target and grid multiple from the
 * loaded state (after one untimed trial per target), keeps the fastest;
 * only before the first step; the loaded state is left in place, and
 * results do not depend on the choice */
int mdx_ensemble_autotune(mdx_ensemble *e, int trial_steps, mdx_ensemble_config *best);

/* cumulative device time (ms) of the pairlist build, force kick and
 * remaining step kernels, and counts, measured with HIP events since the
 * last reset; any pointer may be NULL. */
int mdx_timing(mdx_system *sys, double *build_ms, double *force_ms, double *other_ms,
               int64_t *builds, int64_t *steps);
int mdx_timing_reset(mdx_system *sys, int enable);
/* average device time (ms) of one force-kick launch, and algorithmic HBM
 * bytes of that launch (positions, momenta, neighbour indices) */
int mdx_kick_profile(mdx_system *sys, double *avg_ms, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
