/*
 * mdoracle.c — CPU restatement of MDACP's short-range LJ step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * mdoracle.h for the pinning of each function.  Built with
 * -ffp-contract=off so every operation below rounds exactly as written;
 * where the reference's AVX2 build uses fused multiply-adds for the search
 * distance (meshlist.cc:358-362) this file writes fma() explicitly.
 *
 * Citations are file:line in the reference tree.
 */
#include "mdoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { D_LEFT, D_RIGHT, D_BACK, D_FORWARD, D_DOWN, D_UP, MAX_DIR }; /* mdconfig.h:41 */

struct orc_sys {
  orc_params prm;
  double C0, C2;
  int cap;               /* capacity of q/p incl. images */
  int pn, tn;            /* real / real+image counts (variables.h:14-15) */
  double (*q)[4], (*p)[4];
  int *tag;              /* the reference's type[] used as a particle id */
  int *img_src;          /* for index >= pn: the real particle it images */
  int *img_code;         /* image shift code, see mdoracle.h */
  int *border[MAX_DIR];  /* MDUnit::border_particles (mdunit.h:35) */
  int nborder[MAX_DIR];
  /* mesh (meshlist.h:15-35) */
  int mx, my, mz, nmesh;
  double msize[3];
  int *particle_position, *mesh_index, *mesh_index2, *mesh_particle_number, *sortbuf;
  /* pairs (meshlist.h:30-47) */
  long npairs, pair_cap;
  int (*pairs)[2];
  int *key_pointer, *number_of_partners, *key_pointer2, *sorted_list;
  /* PairList (pairlist.h:10-31) */
  int initialized;
  double buffer_length;
  double (*qb_old)[4];
  double *disp;
  int *cand;
  /* heat bath (simulationinfo.cc:11-23, variables.h:29) */
  orc_heatbath hb;
  uint64_t step;         /* Langevin kicks since orc_set_state (noise counter) */
};

/* variables.cc:21-27 */
void orc_constants(double cutoff, double *c0, double *c2) {
  const double s2 = 1.0 / (cutoff * cutoff);
  const double s6 = s2 * s2 * s2;
  const double s8 = s6 * s2;
  const double s12 = s6 * s6;
  const double s14 = s12 * s2;
  const double C2 = 6.0 * s14 - 3.0 * s8;
  *c2 = C2;
  *c0 = -s12 + s6 - C2 / s2;
}

static double search_length(const orc_params *prm) { /* simulationinfo.cc:7 */
  return prm->cutoff + prm->buffer_length;
}

orc_sys *orc_create(const orc_params *prm, int capacity) {
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  if (!s) return NULL;
  s->prm = *prm;
  orc_constants(prm->cutoff, &s->C0, &s->C2);
  s->cap = capacity;
  s->q = calloc((size_t)capacity, sizeof(double[4]));
  s->p = calloc((size_t)capacity, sizeof(double[4]));
  s->qb_old = calloc((size_t)capacity, sizeof(double[4]));
  s->disp = calloc((size_t)capacity, sizeof(double));
  s->cand = calloc((size_t)capacity, sizeof(int));
  s->tag = calloc((size_t)capacity, sizeof(int));
  s->img_src = calloc((size_t)capacity, sizeof(int));
  s->img_code = calloc((size_t)capacity, sizeof(int));
  for (int d = 0; d < MAX_DIR; d++) s->border[d] = calloc((size_t)capacity, sizeof(int));
  s->particle_position = calloc((size_t)capacity, sizeof(int));
  s->sortbuf = calloc((size_t)capacity, sizeof(int));
  s->key_pointer = calloc((size_t)capacity + 1, sizeof(int));
  s->key_pointer2 = calloc((size_t)capacity, sizeof(int));
  s->number_of_partners = calloc((size_t)capacity, sizeof(int));
  s->pair_cap = 0;
  s->pairs = NULL;
  s->sorted_list = NULL;

  /* MeshList::ChangeScale (meshlist.cc:41-65) with the domain [0, L) */
  const double SL = search_length(prm);
  int ms[3];
  for (int d = 0; d < 3; d++) {
    ms[d] = (int)(prm->L[d] / SL);
    s->msize[d] = prm->L[d] / (double)ms[d];
  }
  s->mx = (int)(prm->L[0] / s->msize[0]) + 2;
  s->my = (int)(prm->L[1] / s->msize[1]) + 2;
  s->mz = (int)(prm->L[2] / s->msize[2]) + 2;
  s->nmesh = s->mx * s->my * s->mz;
  s->mesh_index = calloc((size_t)s->nmesh, sizeof(int));
  s->mesh_index2 = calloc((size_t)s->nmesh, sizeof(int));
  s->mesh_particle_number = calloc((size_t)s->nmesh, sizeof(int));
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->q); free(s->p); free(s->qb_old); free(s->disp); free(s->cand);
  free(s->tag); free(s->img_src); free(s->img_code);
  for (int d = 0; d < MAX_DIR; d++) free(s->border[d]);
  free(s->particle_position); free(s->sortbuf);
  free(s->key_pointer); free(s->key_pointer2); free(s->number_of_partners);
  free(s->mesh_index); free(s->mesh_index2); free(s->mesh_particle_number);
  free(s->pairs); free(s->sorted_list);
  free(s);
}

int orc_set_state(orc_sys *s, int n, const double *q3, const double *p3, const int *tag) {
  if (n > s->cap) return ORC_ERR_CAPACITY;
  for (int i = 0; i < n; i++) { /* Variables::AddParticle (variables.cc:36-47) */
    for (int d = 0; d < 3; d++) {
      s->q[i][d] = q3[3 * i + d];
      s->p[i][d] = p3[3 * i + d];
    }
    s->q[i][3] = 0.0;
    s->p[i][3] = 0.0;
    s->tag[i] = tag ? tag[i] : i;
  }
  s->pn = s->tn = n;
  s->step = 0;
  s->initialized = 0; /* PairList ctor (pairlist.h:25) */
  s->npairs = 0;
  return ORC_OK;
}

int orc_get_state(orc_sys *s, double *q3, double *p3) {
  for (int i = 0; i < s->pn; i++) {
    const int t = s->tag[i];
    if (t < 0 || t >= s->pn) return ORC_ERR_INVALID_INDEX;
    for (int d = 0; d < 3; d++) {
      if (q3) q3[3 * t + d] = s->q[i][d];
      if (p3) p3[3 * t + d] = s->p[i][d];
    }
  }
  return ORC_OK;
}

int orc_particle_number(const orc_sys *s) { return s->pn; }
int orc_total_particle_number(const orc_sys *s) { return s->tn; }
long orc_pair_number(const orc_sys *s) { return s->npairs; }

/* Variables::AdjustPeriodicBoundary (variables.cc:64-84) */
static int adjust_periodic_boundary(orc_sys *s) {
  for (int i = 0; i < s->pn; i++) {
    for (int d = 0; d < 3; d++) {
      if (s->q[i][d] < 0.0) {
        s->q[i][d] += s->prm.L[d];
      } else if (s->q[i][d] > s->prm.L[d]) {
        s->q[i][d] -= s->prm.L[d];
      }
    }
  }
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 3; d++)
      if (s->q[i][d] < 0.0 || s->q[i][d] > s->prm.L[d] || s->q[i][d] != s->q[i][d])
        return ORC_ERR_INVALID_POS;
  return ORC_OK;
}

/* MDRect::IsInsideEdge (mdrect.cc:24-47) with s = 0, e = L */
static int is_inside_edge(const orc_sys *s, int dir, const double *q) {
  const double SL = search_length(&s->prm);
  const double *L = s->prm.L;
  switch (dir) {
  case D_LEFT:    return q[0] < 0.0 + SL && q[0] > 0.0;
  case D_RIGHT:   return q[0] > L[0] - SL && q[0] < L[0];
  case D_BACK:    return q[1] < 0.0 + SL && q[1] > 0.0;
  case D_FORWARD: return q[1] > L[1] - SL && q[1] < L[1];
  case D_DOWN:    return q[2] < 0.0 + SL && q[2] > 0.0;
  case D_UP:      return q[2] > L[2] - SL && q[2] < L[2];
  }
  return 0;
}

/* the diff[] of MDUnit::MakeBufferForBorderParticles (mdunit.cc:157-179),
 * a single domain being its own neighbour in every direction */
static void dir_shift(int dir, int sh[3]) {
  sh[0] = sh[1] = sh[2] = 0;
  switch (dir) {
  case D_LEFT: sh[0] = 1; break;
  case D_RIGHT: sh[0] = -1; break;
  case D_BACK: sh[1] = 1; break;
  case D_FORWARD: sh[1] = -1; break;
  case D_DOWN: sh[2] = 1; break;
  case D_UP: sh[2] = -1; break;
  }
}

static int code_of(const int sh[3]) { return (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1); }
static void shift_of(int code, int sh[3]) {
  sh[0] = code % 3 - 1;
  sh[1] = (code / 3) % 3 - 1;
  sh[2] = code / 9 - 1;
}

/* append the images of border_particles[dir]
 * (MDUnit::MakeBufferForBorderParticles + ReceiveBorderParticles,
 *  mdunit.cc:153-208) */
static int send_border_sub(orc_sys *s, int dir) {
  int sh[3];
  dir_shift(dir, sh);
  double diff[3];
  for (int d = 0; d < 3; d++) diff[d] = sh[d] == 0 ? 0.0 : (sh[d] > 0 ? s->prm.L[d] : -s->prm.L[d]);
  int index = s->tn;
  if (index + s->nborder[dir] > s->cap) return ORC_ERR_CAPACITY;
  for (int k = 0; k < s->nborder[dir]; k++) {
    const int j = s->border[dir][k];
    s->q[index][0] = s->q[j][0] + diff[0];
    s->q[index][1] = s->q[j][1] + diff[1];
    s->q[index][2] = s->q[j][2] + diff[2];
    s->q[index][3] = 0.0;
    s->p[index][0] = s->p[index][1] = s->p[index][2] = s->p[index][3] = 0.0;
    const int src = j < s->pn ? j : s->img_src[j];
    int shj[3] = {0, 0, 0};
    if (j >= s->pn) shift_of(s->img_code[j], shj);
    int shg[3] = {shj[0] + sh[0], shj[1] + sh[1], shj[2] + sh[2]};
    s->img_src[index] = src;
    s->img_code[index] = code_of(shg);
    index++;
  }
  s->tn = index;
  return ORC_OK;
}

/* MDManager::SendBorderParticles for one domain (mdmanager.cc:420-429) */
void orc_refresh_images(orc_sys *s) {
  s->tn = s->pn;
  for (int dir = 0; dir < MAX_DIR; dir++) send_border_sub(s, dir);
}

/* MeshList::MakeMesh (meshlist.cc:176-228) */
static int make_mesh(orc_sys *s) {
  const int pn = s->tn;
  const double imx = 1.0 / s->msize[0];
  const double imy = 1.0 / s->msize[1];
  const double imz = 1.0 / s->msize[2];
  for (int i = 0; i < s->nmesh; i++) s->mesh_particle_number[i] = 0;
  for (int i = 0; i < pn; i++) {
    int ix = (int)((s->q[i][0] - 0.0) * imx) + 1;
    int iy = (int)((s->q[i][1] - 0.0) * imy) + 1;
    int iz = (int)((s->q[i][2] - 0.0) * imz) + 1;
    if (ix < 0) ix = s->mx - 1;
    else if (ix >= s->mx) ix = 0;
    if (iy < 0) iy = s->my - 1;
    else if (iy >= s->my) iy = 0;
    if (iz < 0) iz = s->mz - 1;
    else if (iz >= s->mz) iz = 0;
    const int index = s->mx * s->my * iz + s->mx * iy + ix;
    if (index >= s->nmesh || index < 0) return ORC_ERR_INVALID_INDEX;
    s->particle_position[i] = index;
    s->mesh_particle_number[index]++;
  }
  s->mesh_index[0] = 0;
  int sum = 0;
  for (int i = 0; i < s->nmesh - 1; i++) {
    sum += s->mesh_particle_number[i];
    s->mesh_index[i + 1] = sum;
  }
  for (int i = 0; i < s->nmesh; i++) s->mesh_index2[i] = 0;
  for (int i = 0; i < pn; i++) {
    const int index = s->particle_position[i];
    const int j = s->mesh_index[index] + s->mesh_index2[index];
    s->sortbuf[j] = i;
    s->mesh_index2[index]++;
  }
  return ORC_OK;
}

/* MeshList::RegisterPair (meshlist.cc:556-583), half list */
static int register_pair(orc_sys *s, int index1, int index2) {
  if (s->npairs >= s->pair_cap) {
    long nc = s->pair_cap ? s->pair_cap * 2 : 1 << 16;
    int (*np_)[2] = realloc(s->pairs, (size_t)nc * sizeof(int[2]));
    if (!np_) return ORC_ERR_CAPACITY;
    s->pairs = np_;
    s->pair_cap = nc;
  }
  const int i1 = index1 < index2 ? index1 : index2;
  const int i2 = index1 < index2 ? index2 : index1;
  s->pairs[s->npairs][0] = i1;
  s->pairs[s->npairs][1] = i2;
  s->npairs++;
  return ORC_OK;
}

/* MeshList::AppendList (meshlist.cc:231-240) */
static int append_list(const orc_sys *s, int ix, int iy, int iz, int *v, int nv) {
  if (ix < 0 || ix >= s->mx) return nv;
  if (iy < 0 || iy >= s->my) return nv;
  if (iz < 0 || iz >= s->mz) return nv;
  const int index = s->mx * s->my * iz + s->mx * iy + ix;
  const int in = s->mesh_particle_number[index];
  const int mi = s->mesh_index[index];
  for (int k = 0; k < in; k++) v[nv++] = s->sortbuf[mi + k];
  return nv;
}

/* MeshList::SearchMesh (meshlist.cc:243-292); the distance is the AVX2
 * build's fused form fma(dx,dx, fma(dy,dy, dz*dz)) of meshlist.cc:358-360
 * and the acceptance r2 <= SL^2 of meshlist.cc:362 */
static int search_mesh(orc_sys *s, int index, int *v) {
  int ix = index % s->mx;
  int iy = (index / s->mx) % s->my;
  int iz = index / s->mx / s->my;
  int nv = 0;
  nv = append_list(s, ix, iy, iz, v, nv);
  nv = append_list(s, ix + 1, iy, iz, v, nv);
  nv = append_list(s, ix - 1, iy + 1, iz, v, nv);
  nv = append_list(s, ix, iy + 1, iz, v, nv);
  nv = append_list(s, ix + 1, iy + 1, iz, v, nv);
  nv = append_list(s, ix - 1, iy, iz + 1, v, nv);
  nv = append_list(s, ix, iy, iz + 1, v, nv);
  nv = append_list(s, ix + 1, iy, iz + 1, v, nv);
  nv = append_list(s, ix - 1, iy - 1, iz + 1, v, nv);
  nv = append_list(s, ix, iy - 1, iz + 1, v, nv);
  nv = append_list(s, ix + 1, iy - 1, iz + 1, v, nv);
  nv = append_list(s, ix - 1, iy + 1, iz + 1, v, nv);
  nv = append_list(s, ix, iy + 1, iz + 1, v, nv);
  nv = append_list(s, ix + 1, iy + 1, iz + 1, v, nv);

  const double SL = search_length(&s->prm);
  const double S2 = SL * SL;
  const int pn = s->pn;
  const int in = s->mesh_particle_number[index];
  for (int i = 0; i < in; i++) {
    const int i1 = v[i];
    const double x1 = s->q[i1][0], y1 = s->q[i1][1], z1 = s->q[i1][2];
    for (int j = i + 1; j < nv; j++) {
      const int i2 = v[j];
      if (i1 >= pn && i2 >= pn) continue;
      const double dx = s->q[i2][0] - x1;
      const double dy = s->q[i2][1] - y1;
      const double dz = s->q[i2][2] - z1;
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      if (!(r2 <= S2)) continue;
      int rc = register_pair(s, i1, i2);
      if (rc) return rc;
    }
  }
  return ORC_OK;
}

/* MeshList::MakeList (meshlist.cc:68-104) */
static int make_list(orc_sys *s) {
  s->npairs = 0;
  const int pn = s->tn;
  for (int i = 0; i < pn; i++) s->number_of_partners[i] = 0;
  int rc = make_mesh(s);
  if (rc) return rc;
  int *v = malloc((size_t)s->tn * sizeof(int)); /* candidate buffer of one stencil */
  if (!v) return ORC_ERR_CAPACITY;
  for (int c = 0; c < s->nmesh; c++) {
    rc = search_mesh(s, c, v);
    if (rc) { free(v); return rc; }
  }
  free(v);
  for (long k = 0; k < s->npairs; k++) s->number_of_partners[s->pairs[k][0]]++;
  int pos = 0;
  s->key_pointer[0] = 0;
  for (int i = 0; i < pn - 1; i++) {
    pos += s->number_of_partners[i];
    s->key_pointer[i + 1] = pos;
  }
  for (int i = 0; i < pn; i++) s->key_pointer2[i] = 0;
  free(s->sorted_list);
  s->sorted_list = malloc((size_t)(s->npairs + 8) * sizeof(int));
  if (!s->sorted_list) return ORC_ERR_CAPACITY;
  for (long k = 0; k < s->npairs; k++) {
    const int i = s->pairs[k][0];
    const int j = s->pairs[k][1];
    const int index = s->key_pointer[i] + s->key_pointer2[i];
    s->sorted_list[index] = j;
    s->key_pointer2[i]++;
  }
  return ORC_OK;
}

/* PairList::Init (pairlist.cc:11-23) */
static void pairlist_init(orc_sys *s) {
  s->initialized = 1;
  s->buffer_length = s->prm.buffer_length;
  for (int i = 0; i < s->tn; i++) {
    s->qb_old[i][0] = s->q[i][0];
    s->qb_old[i][1] = s->q[i][1];
    s->qb_old[i][2] = s->q[i][2];
  }
}

/* MDManager::MakePairList (mdmanager.cc:388-417), one domain: migration
 * within a single domain only reorders particles (mdunit.cc:94-138), which
 * this restatement skips because tags carry identity */
int orc_make_pairlist(orc_sys *s) {
  const double SL = search_length(&s->prm);
  for (int d = 0; d < 3; d++)
    if (s->prm.L[d] < 2.0 * SL) return ORC_ERR_BOX;
  int rc = adjust_periodic_boundary(s);
  if (rc) return rc;
  s->tn = s->pn; /* mdmanager.cc:390-394 */
  for (int dir = 0; dir < MAX_DIR; dir++) {
    /* MDUnit::FindBorderParticles (mdunit.cc:141-150) */
    int nb = 0;
    for (int i = 0; i < s->tn; i++)
      if (is_inside_edge(s, dir, s->q[i])) s->border[dir][nb++] = i;
    s->nborder[dir] = nb;
    rc = send_border_sub(s, dir);
    if (rc) return rc;
  }
  pairlist_init(s);   /* MDUnit::MakePairList (mdunit.cc:211-215) */
  return make_list(s);
}

/* PairList::UpdatePairListValidity (pairlist.cc:48-65) */
static int update_validity(orc_sys *s) {
  const int pn = s->pn;
  double max_velocity = s->p[0][0] * s->p[0][0] + s->p[0][1] * s->p[0][1] + s->p[0][2] * s->p[0][2];
  for (int i = 1; i < pn; i++) {
    const double v = s->p[i][0] * s->p[i][0] + s->p[i][1] * s->p[i][1] + s->p[i][2] * s->p[i][2];
    if (v > max_velocity) max_velocity = v;
  }
  max_velocity = sqrt(max_velocity);
  if (max_velocity * 2.0 * s->prm.dt > s->prm.buffer_length) return ORC_ERR_TOO_FAST;
  s->buffer_length = s->buffer_length - max_velocity * 2.0 * s->prm.dt;
  return ORC_OK;
}

/* PairList::CheckByDisplacement (pairlist.cc:68-121) */
static int check_by_displacement(orc_sys *s) {
  const int pn = s->tn;
  double dr2_max = 0.0;
  for (int i = 0; i < pn; i++) {
    const double dx = s->qb_old[i][0] - s->q[i][0];
    const double dy = s->qb_old[i][1] - s->q[i][1];
    const double dz = s->qb_old[i][2] - s->q[i][2];
    const double dr2 = dx * dx + dy * dy + dz * dz;
    s->disp[i] = dr2;
    if (dr2_max < dr2) dr2_max = dr2;
  }
  const int Nl = pn > s->prm.check_list_length ? s->prm.check_list_length : pn;
  const double c_distance = s->prm.buffer_length - sqrt(dr2_max);
  const double c_d2 = c_distance * c_distance;
  int l = 0;
  for (int i = 0; i < pn; i++) {
    if (s->disp[i] > c_d2) {
      s->cand[l++] = i;
      if (l >= Nl) return 1;
    }
  }
  const double SL = search_length(&s->prm);
  const double SL2 = SL * SL;
  const double CL2 = s->prm.cutoff * s->prm.cutoff;
  for (int i = 0; i < l - 1; i++) {
    for (int j = i + 1; j < l; j++) {
      const int i1 = s->cand[i], i2 = s->cand[j];
      double dx = s->qb_old[i1][0] - s->qb_old[i2][0];
      double dy = s->qb_old[i1][1] - s->qb_old[i2][1];
      double dz = s->qb_old[i1][2] - s->qb_old[i2][2];
      double dr2 = dx * dx + dy * dy + dz * dz;
      if (dr2 < SL2) continue;
      dx = s->q[i1][0] - s->q[i2][0];
      dy = s->q[i1][1] - s->q[i2][1];
      dz = s->q[i1][2] - s->q[i2][2];
      dr2 = dx * dx + dy * dy + dz * dz;
      if (dr2 < CL2) return 1;
    }
  }
  return 0;
}

/* PairList::IsPairListExpired (pairlist.cc:26-45) */
int orc_is_pairlist_expired(orc_sys *s, int *expired) {
  if (!s->initialized) {
    *expired = 1;
    return ORC_OK;
  }
  if (s->buffer_length > 0.0) {
    int rc = update_validity(s);
    if (rc) return rc;
  }
  *expired = 0;
  if (s->buffer_length < 0.0) *expired = check_by_displacement(s);
  return ORC_OK;
}

/* ForceCalculator::UpdatePositionHalf (fcalculator.cc:18-30) */
void orc_update_position_half(orc_sys *s) {
  const double dt2 = s->prm.dt * 0.5;
  for (int i = 0; i < s->pn; i++) {
    s->q[i][0] += s->p[i][0] * dt2;
    s->q[i][1] += s->p[i][1] * dt2;
    s->q[i][2] += s->p[i][2] * dt2;
  }
}

/* ForceCalculator::CalculateForceSorted (fcalculator.cc:182-223): the
 * arithmetic of the default kernels (CalculateForceNext :229-290,
 * CalculateForceAVX2 :294-466) in plain list order */
static void kick(orc_sys *s, double dt) {
  const double CL2 = s->prm.cutoff * s->prm.cutoff;
  const double C2 = s->C2;
  const int pn = s->pn;
  for (int i = 0; i < pn; i++) {
    const double qx_key = s->q[i][0], qy_key = s->q[i][1], qz_key = s->q[i][2];
    const int np = s->number_of_partners[i];
    const int kp = s->key_pointer[i];
    double pfx = 0, pfy = 0, pfz = 0;
    for (int k = 0; k < np; k++) {
      const int j = s->sorted_list[kp + k];
      const double dx = s->q[j][0] - qx_key;
      const double dy = s->q[j][1] - qy_key;
      const double dz = s->q[j][2] - qz_key;
      const double r2 = (dx * dx + dy * dy + dz * dz);
      const double r6 = r2 * r2 * r2;
      double df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0) * dt;
      if (r2 > CL2) df = 0.0;
      pfx += df * dx;
      pfy += df * dy;
      pfz += df * dz;
      s->p[j][0] -= df * dx;
      s->p[j][1] -= df * dy;
      s->p[j][2] -= df * dz;
    }
    s->p[i][0] += pfx;
    s->p[i][1] += pfy;
    s->p[i][2] += pfz;
  }
}

void orc_calculate_force(orc_sys *s) { kick(s, s->prm.dt); }

int orc_forces(orc_sys *s, double *f3) {
  double (*save)[4] = malloc((size_t)s->tn * sizeof(double[4]));
  if (!save) return ORC_ERR_CAPACITY;
  memcpy(save, s->p, (size_t)s->tn * sizeof(double[4]));
  for (int i = 0; i < s->tn; i++) s->p[i][0] = s->p[i][1] = s->p[i][2] = 0.0;
  kick(s, 1.0);
  for (int i = 0; i < s->pn; i++) {
    const int t = s->tag[i];
    f3[3 * t + 0] = s->p[i][0];
    f3[3 * t + 1] = s->p[i][1];
    f3[3 * t + 2] = s->p[i][2];
  }
  memcpy(s->p, save, (size_t)s->tn * sizeof(double[4]));
  free(save);
  return ORC_OK;
}

/* MDManager::Calculate (mdmanager.cc:154-186).  NVE: CalculateForce with
 * its trailing half drift (mdmanager.cc:227-252).  Nose-Hoover
 * (CalculateNoseHoover, :255-287): zeta from the temperature, momenta
 * scaled by exp(-zeta dt/2) around the kick, zeta again, half drift.
 * Langevin (CalculateLangevin, :289-315): kick, then the noise step, and no
 * trailing half drift. */
int orc_calculate(orc_sys *s, int nsteps, int *rebuilds) {
  int nreb = 0;
  for (int step = 0; step < nsteps; step++) {
    int expired = 0;
    int rc = orc_is_pairlist_expired(s, &expired);
    if (rc) return rc;
    if (expired) {
      rc = orc_make_pairlist(s);
      if (rc) return rc;
      nreb++;
    }
    orc_update_position_half(s);
    orc_refresh_images(s);
    if (s->hb.type == 1) {
      orc_heatbath_zeta(s, orc_temperature(s));
      orc_heatbath_momenta(s);
      orc_calculate_force(s);
      orc_heatbath_momenta(s);
      orc_heatbath_zeta(s, orc_temperature(s));
      orc_update_position_half(s);
    } else if (s->hb.type == 2) {
      orc_calculate_force(s);
      orc_langevin(s);
    } else {
      orc_calculate_force(s);
      orc_update_position_half(s);
    }
  }
  if (rebuilds) *rebuilds = nreb;
  return ORC_OK;
}

void orc_set_heatbath(orc_sys *s, const orc_heatbath *hb) { s->hb = *hb; }
void orc_get_heatbath(const orc_sys *s, orc_heatbath *hb) { *hb = s->hb; }

/* ForceCalculator::HeatbathZeta (fcalculator.cc:979-985) */
void orc_heatbath_zeta(orc_sys *s, double current_temperature) {
  const double dt2 = s->prm.dt * 0.5;
  const double tau = s->hb.tau;
  const double t1 = (current_temperature - s->hb.aimed_temperature) / (tau * tau);
  s->hb.zeta += t1 * dt2;
}

/* ForceCalculator::HeatbathMomenta (fcalculator.cc:987-999), all four lanes */
void orc_heatbath_momenta(orc_sys *s) {
  const double dt2 = s->prm.dt * 0.5;
  const double exp1 = exp(-dt2 * s->hb.zeta);
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 4; d++) s->p[i][d] *= exp1;
}

/* ForceCalculator::Langevin (fcalculator.cc:1001-1017): all four lanes, the
 * padding lane included, with r ~ N(0, sqrt(2 gamma T / dt)) */
void orc_langevin(orc_sys *s) {
  const double dt = s->prm.dt;
  const double hb_gamma = s->hb.gamma;
  const double T = s->hb.aimed_temperature;
  const double hb_D = sqrt(2.0 * hb_gamma * T / dt);
  for (int i = 0; i < s->pn; i++) {
    double nrm[4];
    orc_langevin_normals(s->hb.seed, (uint32_t)s->tag[i], s->step, nrm);
    for (int d = 0; d < 4; d++) {
      const double r = hb_D * nrm[d];
      s->p[i][d] += (-hb_gamma * s->p[i][d] + r) * dt;
    }
  }
  s->step++;
}

/* MDManager::Temperature (mdmanager.cc:531-533) over KineticEnergy
 * (mdmanager.cc:543-548) */
double orc_temperature(orc_sys *s) {
  return orc_kinetic_energy(s) / (double)s->pn / 1.5;
}

/* Philox4x32-10: ten rounds of two 32x32->64 multiplies and xors, the key
 * bumped by the Weyl constants between rounds (Salmon et al., SC'11) */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t a = (uint64_t)0xD2511F53u * c0, b = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0, n1 = (uint32_t)b;
    const uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1, n3 = (uint32_t)a;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double uniform53(uint32_t a, uint32_t b) {   /* (0, 1] */
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)m + 1.0) * (1.0 / 9007199254740992.0);
}

void orc_langevin_normals(uint64_t seed, uint32_t tag, uint64_t step, double out[4]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[8];
  for (int h = 0; h < 2; h++) {
    const uint32_t ctr[4] = {tag, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)h};
    orc_philox4x32(ctr, key, w + 4 * h);
  }
  const double two_pi = 6.283185307179586;
  for (int k = 0; k < 2; k++) {
    const double u1 = uniform53(w[4 * k], w[4 * k + 1]);
    const double u2 = uniform53(w[4 * k + 2], w[4 * k + 3]);
    const double rr = sqrt(-2.0 * log(u1));
    out[2 * k] = rr * cos(two_pi * u2);
    out[2 * k + 1] = rr * sin(two_pi * u2);
  }
}

/* KineticEnergyObserver::Observe (observer.cc:5-15) */
double orc_kinetic_energy(orc_sys *s) {
  double e = 0;
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 4; d++) e += 0.5 * s->p[i][d] * s->p[i][d];
  return e;
}

/* PotentialEnergyObserver::Observe (observer.cc:18-47), half list */
double orc_potential_energy(orc_sys *s) {
  const double CL2 = s->prm.cutoff * s->prm.cutoff;
  const double C2 = s->C2, C0 = s->C0;
  const int pn = s->pn;
  double energy = 0.0;
  for (long k = 0; k < s->npairs; k++) {
    const int i = s->pairs[k][0], j = s->pairs[k][1];
    const double dx = s->q[i][0] - s->q[j][0];
    const double dy = s->q[i][1] - s->q[j][1];
    const double dz = s->q[i][2] - s->q[j][2];
    const double r2 = (dx * dx + dy * dy + dz * dz);
    if (r2 > CL2) continue;
    double e = 4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / (r2 * r2 * r2) + C2 * r2 + C0);
    if (i >= pn || j >= pn) e *= 0.5;
    energy += e;
  }
  return energy;
}

/* VirialObserver::Observe (observer.cc:50-79), half list; returns sum/3 */
double orc_virial(orc_sys *s) {
  const double CL2 = s->prm.cutoff * s->prm.cutoff;
  const double C_2 = s->C2;
  const int pn = s->pn;
  double phi = 0.0;
  for (long k = 0; k < s->npairs; k++) {
    const int i = s->pairs[k][0], j = s->pairs[k][1];
    const double dx = s->q[j][0] - s->q[i][0];
    const double dy = s->q[j][1] - s->q[i][1];
    const double dz = s->q[j][2] - s->q[i][2];
    const double r2 = (dx * dx + dy * dy + dz * dz);
    if (r2 > CL2) continue;
    const double r6 = r2 * r2 * r2;
    double df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C_2 * 8.0) * r2;
    if (i >= pn || j >= pn) df *= 0.5;
    phi += df;
  }
  return phi / 3.0;
}

static int cmp_triple(const void *a, const void *b) {
  const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
  for (int k = 0; k < 3; k++) {
    if (x[k] < y[k]) return -1;
    if (x[k] > y[k]) return 1;
  }
  return 0;
}

static long sort_unique(int32_t *t, long n) {
  qsort(t, (size_t)n, sizeof(int32_t[3]), cmp_triple);
  long m = 0;
  for (long k = 0; k < n; k++) {
    if (m > 0 && cmp_triple(t + 3 * (m - 1), t + 3 * k) == 0) continue;
    if (m != k) memcpy(t + 3 * m, t + 3 * k, sizeof(int32_t[3]));
    m++;
  }
  return m;
}

long orc_pairs_canonical(orc_sys *s, int32_t *out, long cap) {
  int32_t *t = malloc((size_t)(s->npairs + 1) * sizeof(int32_t[3]));
  if (!t) return ORC_ERR_CAPACITY;
  for (long k = 0; k < s->npairs; k++) {
    int a = s->pairs[k][0], b = s->pairs[k][1];
    int ta = a < s->pn ? s->tag[a] : s->tag[s->img_src[a]];
    int ca = a < s->pn ? 13 : s->img_code[a];
    int tb = b < s->pn ? s->tag[b] : s->tag[s->img_src[b]];
    int cb = b < s->pn ? 13 : s->img_code[b];
    /* relative image of b seen from a: shift(b) - shift(a) */
    int sa[3], sb[3], rel[3];
    shift_of(ca, sa);
    shift_of(cb, sb);
    for (int d = 0; d < 3; d++) rel[d] = sb[d] - sa[d];
    if (ta < tb) {
      t[3 * k + 0] = ta; t[3 * k + 1] = tb; t[3 * k + 2] = code_of(rel);
    } else {
      int neg[3] = {-rel[0], -rel[1], -rel[2]};
      t[3 * k + 0] = tb; t[3 * k + 1] = ta; t[3 * k + 2] = code_of(neg);
    }
  }
  long m = sort_unique(t, s->npairs);
  if (out) {
    if (m > cap) { free(t); return ORC_ERR_CAPACITY; }
    memcpy(out, t, (size_t)m * sizeof(int32_t[3]));
  }
  free(t);
  return m;
}

/* the image of q_j nearest to q_i along one axis, formed as the reference
 * forms ghosts: (q_j + s L) - q_i with s in {-1, 0, 1} */
static double min_image(double qi, double qj, double L, int *s_out) {
  double best = qj - qi;
  int bs = 0;
  const double a = (qj + L) - qi, b = (qj - L) - qi;
  if (fabs(a) < fabs(best)) { best = a; bs = 1; }
  if (fabs(b) < fabs(best)) { best = b; bs = -1; }
  *s_out = bs;
  return best;
}

long orc_bruteforce_pairs(const orc_params *prm, int n, const double *q3, double r2max,
                          int32_t *out, long cap) {
  long m = 0;
  for (int i = 0; i < n; i++) {
    for (int j = i + 1; j < n; j++) {
      int sh[3];
      const double dx = min_image(q3[3 * i + 0], q3[3 * j + 0], prm->L[0], &sh[0]);
      const double dy = min_image(q3[3 * i + 1], q3[3 * j + 1], prm->L[1], &sh[1]);
      const double dz = min_image(q3[3 * i + 2], q3[3 * j + 2], prm->L[2], &sh[2]);
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      if (!(r2 <= r2max)) continue;
      if (out) {
        if (m >= cap) return ORC_ERR_CAPACITY;
        out[3 * m + 0] = i;
        out[3 * m + 1] = j;
        out[3 * m + 2] = code_of(sh);
      }
      m++;
    }
  }
  return m;
}

/* CalculateForceBruteforce (fcalculator.cc:53-78) at dt = 1, with the
 * PE and virial terms of observer.cc:18-79 */
void orc_bruteforce_forces(const orc_params *prm, int n, const double *q3, double *f3,
                           double *pe, double *vir) {
  double C0, C2;
  orc_constants(prm->cutoff, &C0, &C2);
  const double CL2 = prm->cutoff * prm->cutoff;
  double e = 0, phi = 0;
  for (int i = 0; i < 3 * n; i++) f3[i] = 0.0;
  for (int i = 0; i < n; i++) {
    for (int j = i + 1; j < n; j++) {
      int sh[3];
      const double dx = min_image(q3[3 * i + 0], q3[3 * j + 0], prm->L[0], &sh[0]);
      const double dy = min_image(q3[3 * i + 1], q3[3 * j + 1], prm->L[1], &sh[1]);
      const double dz = min_image(q3[3 * i + 2], q3[3 * j + 2], prm->L[2], &sh[2]);
      const double r2 = (dx * dx + dy * dy + dz * dz);
      if (r2 > CL2) continue;
      const double r6 = r2 * r2 * r2;
      const double df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0);
      f3[3 * i + 0] += df * dx;
      f3[3 * i + 1] += df * dy;
      f3[3 * i + 2] += df * dz;
      f3[3 * j + 0] -= df * dx;
      f3[3 * j + 1] -= df * dy;
      f3[3 * j + 2] -= df * dz;
      e += 4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / (r2 * r2 * r2) + C2 * r2 + C0);
      phi += df * r2;
    }
  }
  if (pe) *pe = e;
  if (vir) *vir = phi / 3.0;
}
