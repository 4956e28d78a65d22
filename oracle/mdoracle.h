/*
 * mdoracle — CPU restatement of MDACP's short-range LJ step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libmdx) never calls it and has no CPU fallback.
 *
 * Every function restates one piece of the reference's algorithm and cites
 * the reference file:line it follows (paths relative to the reference tree).
 * Layout follows the reference: q, p are AoS double[n][4] with lane 3 unused
 * (variables.h:19-28); indices [0, pn) are real particles, [pn, tn) periodic
 * images ("ghosts") appended in the reference's six-direction order.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   - pair search, forces, PE, virial: against the compiled reference
 *     (oracle/_ref, golden fixtures in tests/golden/), against closed forms
 *     of the perfect FCC lattice, and against O(N^2) minimum-image brute force;
 *   - drift, step order and pairlist expiry: against the reference's
 *     1000-step trajectory fixture (observables and final state);
 *   - image construction: every pair found through an image is re-found by
 *     the minimum-image brute force;
 *   - Nose-Hoover heat bath: against the reference's thermostatted trajectory
 *     fixture, and zeta against its closed-form update on a frozen system;
 *   - Langevin heat bath: its noise (Philox4x32-10 + Box-Muller) against the
 *     published Philox known-answer vectors and normal statistics, its
 *     momentum update against the closed form p(1 - gamma dt) + r dt.
 */
#ifndef MDORACLE_H
#define MDORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double L[3];            /* box lengths (simulationinfo.cc:31-49) */
  double dt;              /* TimeStep (simulationinfo.cc:9) */
  double cutoff;          /* CUTOFF_LENGTH = 3.0 (mdconfig.h:31) */
  double buffer_length;   /* BufferLength = 0.3 (simulationinfo.cc:6) */
  int check_list_length;  /* CheckListLength = 500 (simulationinfo.cc:29) */
} orc_params;

typedef struct orc_sys orc_sys;

/* Heat bath (SimulationInfo::ControlTemperature, HeatbathType,
 * AimedTemperature, HeatbathTau, HeatbathGamma, simulationinfo.cc:11-23;
 * Variables::Zeta, variables.h:29).  type 0 = no control (NVE), 1 =
 * Nose-Hoover, 2 = Langevin.  The reference draws Langevin noise from a
 * per-thread std::mt19937 (fcalculator.cc:1002-1017), which fixes no
 * particle-to-number map; here the noise of particle `tag` in the k-th
 * Langevin kick since orc_set_state (k from 0) is
 * orc_langevin_normals(seed, tag, k). */
typedef struct {
  int type;
  double aimed_temperature;
  double tau;
  double gamma;
  double zeta;
  uint64_t seed;
} orc_heatbath;

/* status codes: 0 ok, negative = fatal condition the reference exit()s on */
enum {
  ORC_OK = 0,
  ORC_ERR_CAPACITY = -1,      /* particle/pair capacity (variables.cc:46, meshlist.cc:582) */
  ORC_ERR_INVALID_POS = -2,   /* "Invalid position" (variables.cc:77-81) */
  ORC_ERR_TOO_FAST = -3,      /* "Too fast particles" (pairlist.cc:59-63) */
  ORC_ERR_INVALID_INDEX = -4, /* "Invalid index" (meshlist.cc:202-207) */
  ORC_ERR_BOX = -5            /* box narrower than two search lengths */
};

orc_sys *orc_create(const orc_params *prm, int capacity);
void orc_destroy(orc_sys *s);

/* load n particles, q3/p3 are [n][3]; tag[i] identifies the particle (the
 * reference keeps it in type[], variables.cc:36-47).  tag may be NULL. */
int orc_set_state(orc_sys *s, int n, const double *q3, const double *p3, const int *tag);
/* copy state out, ordered by tag (tag must be a permutation of 0..n-1) */
int orc_get_state(orc_sys *s, double *q3, double *p3);
int orc_particle_number(const orc_sys *s);
int orc_total_particle_number(const orc_sys *s);

/* MDManager::MakePairList (mdmanager.cc:388-417) for one periodic domain */
int orc_make_pairlist(orc_sys *s);
/* MDManager::IsPairListExpired -> PairList::IsPairListExpired (pairlist.cc:26-121) */
int orc_is_pairlist_expired(orc_sys *s, int *expired);
/* ForceCalculator::UpdatePositionHalf (fcalculator.cc:18-30) */
void orc_update_position_half(orc_sys *s);
/* MDManager::SendBorderParticles for one domain (mdmanager.cc:420-429) */
void orc_refresh_images(orc_sys *s);
/* ForceCalculator::CalculateForce, half list with reaction (fcalculator.cc:182-223) */
void orc_calculate_force(orc_sys *s);
/* MDManager::Calculate (mdmanager.cc:154-186): NVE, or the heat bath set
 * by orc_set_heatbath (CalculateNoseHoover / CalculateLangevin,
 * mdmanager.cc:255-315) */
int orc_calculate(orc_sys *s, int nsteps, int *rebuilds);

void orc_set_heatbath(orc_sys *s, const orc_heatbath *hb);
void orc_get_heatbath(const orc_sys *s, orc_heatbath *hb);
/* ForceCalculator::HeatbathZeta / HeatbathMomenta / Langevin
 * (fcalculator.cc:979-1017) */
void orc_heatbath_zeta(orc_sys *s, double current_temperature);
void orc_heatbath_momenta(orc_sys *s);
void orc_langevin(orc_sys *s);
/* MDManager::Temperature (mdmanager.cc:531-533): KE per particle / 1.5 */
double orc_temperature(orc_sys *s);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): the 128-bit block of counter ctr under key. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* four standard normals: two Philox blocks with counter (tag, step lo,
 * step hi, 0 / 1) and key (seed lo, seed hi), eight words -> four 53-bit
 * uniforms u in (0, 1] -> two Box-Muller pairs
 * sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2) */
void orc_langevin_normals(uint64_t seed, uint32_t tag, uint64_t step, double out[4]);

/* observers (observer.cc:5-79); sums, not per-particle */
double orc_kinetic_energy(orc_sys *s);
double orc_potential_energy(orc_sys *s);
double orc_virial(orc_sys *s);

/* forces F = dp/dt of the current pairlist, written [n][3] by tag, with p
 * untouched (the reference's kick evaluated at dt = 1 on p = 0) */
int orc_forces(orc_sys *s, double *f3);

/* number of (key, partner) entries, as MeshList::GetPairNumber */
long orc_pair_number(const orc_sys *s);
/* canonical pair set: (tag_lo, tag_hi, image code) triples, deduplicated,
 * sorted; code = (sx+1) + 3(sy+1) + 9(sz+1), the partner image being
 * q_hi + (sx Lx, sy Ly, sz Lz).  Returns the count; out may be NULL. */
long orc_pairs_canonical(orc_sys *s, int32_t *out, long cap);

/* O(N^2) minimum-image brute force (semantics of meshlist.cc:107-123 and
 * fcalculator.cc:53-78 on an unwrapped periodic box) */
long orc_bruteforce_pairs(const orc_params *prm, int n, const double *q3,
                          double r2max, int32_t *out, long cap);
void orc_bruteforce_forces(const orc_params *prm, int n, const double *q3,
                           double *f3, double *pe, double *vir);

/* the quadratic-shift constants (variables.cc:21-27) */
void orc_constants(double cutoff, double *c0, double *c2);

#ifdef __cplusplus
}
#endif
#endif
