/*
 * ref_cpu.c -- the oracle: a CPU restatement of dynLattice's explicit
 * dynamics path (rod element, rod materials, block-lumped mass, gyro,
 * LeapFrog with store, rod-rod contact, commit energies).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py.  The product library never links it.
 *
 * Operation order follows the reference line by line (matrix route: the
 * Crisfield-Jelenic interpolation through logMat/expVec exactly as
 * Line3D.cpp and helpers.cpp evaluate it, material calls as in
 * ElasticRodMaterial.cpp / ElastoPlasticRodMaterial.cpp).  Repeated helper
 * calls on identical inputs are evaluated once and reused: they are
 * deterministic, so the values are the same bits.
 *
 * Third-party predicates the reference takes from JEM/JIVE (not vendored)
 * are fixed here as named choices:
 *   REF_IS_TINY           jem::isTiny(x)  -> |x| < DBL_MIN
 *   ref_contains          StdLine::containsLocalPoint(u) -> -1 <= u <= 1
 *   shape functions       StdLine2: N = (0.5(1-u), 0.5(1+u))
 *   Gauss points          Gauss1: u=0, w=2;  Gauss2: u=-+1/sqrt(3), w=1
 *   yield gradient        analytic, zeroed where |s_i| < DBL_EPSILON
 *   consistent solve      exact per-node 6x6 solve of the reduced system
 * Parity at these points is therefore "unpinned" against the real JEM/JIVE
 * build, which cannot be built here.
 */
#include "ref_cpu.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY_H (DBL_EPSILON * 1e6) /* helpers.h:55 */
#ifdef REF_OMP
#define REF_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define REF_PARALLEL_FOR
#endif
#define REF_IS_TINY(x) (fabs(x) < DBL_MIN)

/* Planted mistakes: the tests that pin the contact forces and the
 * gyroscopic term build this file once per mistake k with -DREF_PLANT=k
 * (into a temporary directory) and show that their check fails on it.
 * Every build of the oracle itself has REF_PLANT 0.
 *   1 STS force sign flipped        2 STS shape weights of A swapped
 *   3 NTS penalty doubled           4 NTS master weights swapped
 *   5 gyroscopic term sign flipped  6 gyroscopic term once per node
 *                                     instead of once per rod model */
#ifndef REF_PLANT
#define REF_PLANT 0
#endif

/* ------------------------------------------------------------------ */
/* small dense helpers (row-major 3x3)                                 */
/* ------------------------------------------------------------------ */
static void eye3(double R[9]) {
  memset(R, 0, 9 * sizeof(double));
  R[0] = R[4] = R[8] = 1.0;
}
static void skew3(const double v[3], double M[9]) { /* helpers.cpp:242-255 */
  M[0] = 0.0;   M[1] = -1 * v[2]; M[2] = 1 * v[1];
  M[3] = 1 * v[2]; M[4] = 0.0;   M[5] = -1 * v[0];
  M[6] = -1 * v[1]; M[7] = 1 * v[0]; M[8] = 0.0;
}
static void unskew3(const double M[9], double v[3]) { /* helpers.cpp:257-281 */
  v[0] = M[7] - M[5];
  v[1] = M[2] - M[6];
  v[2] = M[3] - M[1];
  v[0] /= 2.; v[1] /= 2.; v[2] /= 2.;
}
static void mm3(const double A[9], const double B[9], double C[9]) {
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = A[3 * i] * B[j];
      s += A[3 * i + 1] * B[3 + j];
      s += A[3 * i + 2] * B[6 + j];
      T[3 * i + j] = s;
    }
  memcpy(C, T, sizeof T);
}
static void mtm3(const double A[9], const double B[9], double C[9]) { /* A^T B */
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = A[i] * B[j];
      s += A[3 + i] * B[3 + j];
      s += A[6 + i] * B[6 + j];
      T[3 * i + j] = s;
    }
  memcpy(C, T, sizeof T);
}
static void mmt3(const double A[9], const double B[9], double C[9]) { /* A B^T */
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = A[3 * i] * B[3 * j];
      s += A[3 * i + 1] * B[3 * j + 1];
      s += A[3 * i + 2] * B[3 * j + 2];
      T[3 * i + j] = s;
    }
  memcpy(C, T, sizeof T);
}
static void mv3(const double A[9], const double x[3], double y[3]) {
  double t[3];
  for (int i = 0; i < 3; i++) {
    double s = A[3 * i] * x[0];
    s += A[3 * i + 1] * x[1];
    s += A[3 * i + 2] * x[2];
    t[i] = s;
  }
  y[0] = t[0]; y[1] = t[1]; y[2] = t[2];
}
static void mtv3(const double A[9], const double x[3], double y[3]) { /* A^T x */
  double t[3];
  for (int i = 0; i < 3; i++) {
    double s = A[i] * x[0];
    s += A[3 + i] * x[1];
    s += A[6 + i] * x[2];
    t[i] = s;
  }
  y[0] = t[0]; y[1] = t[1]; y[2] = t[2];
}
static double dot3(const double a[3], const double b[3]) {
  double s = a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}
static double norm3(const double a[3]) { return sqrt(dot3(a, a)); }

/* ------------------------------------------------------------------ */
/* SO(3) maps: helpers.cpp                                             */
/* ------------------------------------------------------------------ */
void ref_expvec(const double psi[3], double R[9]) { /* helpers.cpp:139-157 */
  eye3(R);
  const double theta = norm3(psi);
  if (REF_IS_TINY(theta)) {
    double K[9];
    skew3(psi, K);
    for (int i = 0; i < 9; i++) R[i] += K[i];
  } else {
    double k[3] = {psi[0] / theta, psi[1] / theta, psi[2] / theta};
    double K[9], KK[9];
    skew3(k, K);
    const double s = sin(theta);
    for (int i = 0; i < 9; i++) R[i] += s * K[i];
    mm3(K, K, KK);
    const double c1 = 1 - cos(theta);
    for (int i = 0; i < 9; i++) R[i] += c1 * KK[i];
  }
}

void ref_expvecp(const double psi[3], const double psiP[3], double out[9]) { /* helpers.cpp:159-197 */
  const double theta = norm3(psi);
  if (theta < TINY_H) {
    skew3(psiP, out);
    return;
  }
  const double thetaP = dot3(psi, psiP) / theta;
  double k[3], kP[3], K[9], KP[9], T[9];
  for (int i = 0; i < 3; i++) {
    k[i] = psi[i] / theta;
    kP[i] = (psiP[i] * theta - psi[i] * thetaP) / theta / theta;
  }
  skew3(k, K);
  skew3(kP, KP);
  for (int i = 0; i < 9; i++) out[i] = 0;
  const double ct = cos(theta), st = sin(theta);
  for (int i = 0; i < 9; i++) out[i] += ct * thetaP * K[i];
  for (int i = 0; i < 9; i++) out[i] += st * KP[i];
  mm3(K, K, T);
  for (int i = 0; i < 9; i++) out[i] += st * thetaP * T[i];
  mm3(K, KP, T);
  for (int i = 0; i < 9; i++) out[i] += (1 - ct) * T[i];
  mm3(KP, K, T);
  for (int i = 0; i < 9; i++) out[i] += (1 - ct) * T[i];
}

void ref_logmat(const double R[9], double rv[3]) { /* helpers.cpp:99-111 */
  double tr = R[0];
  tr += R[4];
  tr += R[8];
  const double tr_R = tr <= 3. ? tr : 3.;
  const double theta = acos((tr_R - 1.) / 2.);
  double M[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[3 * i + j] = R[3 * i + j] - R[3 * j + i];
  unskew3(M, rv);
  if (REF_IS_TINY(theta)) {
    const double h = 1 / 2.;
    rv[0] *= h; rv[1] *= h; rv[2] *= h;
  } else {
    const double f = theta / (2. * sin(theta));
    rv[0] *= f; rv[1] *= f; rv[2] *= f;
  }
}

/* ------------------------------------------------------------------ */
/* model                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t n_el;
  int32_t *elems;   /* global element indices in group order */
  int32_t n_nd;
  int32_t *nodes;   /* sorted unique global node indices (getNodeIndices) */
  double *lambdaN;  /* [9*n_nd] LambdaN_ per rod node (SpecialCosseratRodModel.cpp:598) */
  double box[6];    /* scratch: rod AABB */
} strut_t;

struct ref_model {
  int64_t N, E;
  int32_t R;
  double *X0;        /* [3N] */
  int32_t *en;       /* [2E] */
  int32_t *es;       /* [E] */
  strut_t *st;
  int32_t *el_local; /* [E] index of element within its strut */
  /* sections / materials per strut */
  double *K;         /* [6R] diag material stiffness (calcMaterialStiff_) */
  double *Mm;        /* [6R] diag material mass (calcMaterialMass_) */
  double *area, *rho;
  /* plastic */
  ref_plastic pl;
  double *pw;        /* [6R] yield weights */
  double *Hk;        /* [6R] kinematic hardening diag */
  /* material state per element (ipCount = 1) */
  double *strain0;   /* [6E] matStrain0_ */
  double *curr_strain, *old_strain; /* [6E] */
  double *epot;      /* [E] energyPot_ */
  double *curr_ep, *old_ep;         /* [6E] */
  double *curr_al, *old_al;         /* [7E] iso, kin... */
  double *dflow;     /* [E] currDeltaFlow_ */
  double *ediss;     /* [E] energyDiss_ */
  double *ehard;     /* [E] */
  uint8_t *active;   /* [E] plastic correction taken in last call */
  /* contact */
  ref_contact_opts co;
  double *crad;      /* [R] */
  int64_t n_black;
  int64_t *black;    /* sorted keys (min<<32 | max) of blacklisted pairs */
  int64_t n_act, cap_act;
  int32_t *actA, *actB;
  int use_grid;
  /* bcs */
  uint8_t *cmask;
  double *cval, *fext;
  /* boundary programs: dl_set_bcs-style base values scaled per step */
  int n_prog;
  ref_prog prog[REF_MAX_PROG];
  int32_t *dof_prog;  /* [6N] or NULL */
  double *cval0, *fext0;
  double clock_t;
  int64_t clock_k;
  /* updateWhen of the explicit modules (ExplicitModule.cpp:119-126,
   * 256-266): the mass is assembled at the start of a step only when the
   * condition holds or no valid mass exists (after INIT) */
  char *mu_cond;      /* NULL: default, every step */
  int mu_valid;
  double *mu_blocks;  /* [36N] the mass in use */
  /* point masses (MassModel) */
  double *pm_mass, *pm_inertia;
  /* ImpactBC groups */
  int imp_n, imp_ready;
  int32_t *imp_nodes, *imp_ptr, *imp_dof;
  double *imp_w;
  /* options */
  int so3;
  double inertia_correct;
  int err;
  /* all-cores variant: strut-local node rows and the node -> rows map */
  int64_t *st_off, *nd_ptr, *nd_row;
  double *loc_blk, *loc_f, *loc_g;
};

int ref_last_error(const ref_model *m) { return m->err; }

static void impact_constraints(ref_model *m, const double *fint);

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}
static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

static int rod_local_node(const strut_t *s, int32_t g) {
  int lo = 0, hi = s->n_nd - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (s->nodes[mid] == g) return mid;
    if (s->nodes[mid] < g) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

ref_model *ref_create(const ref_mesh *mesh, const ref_sections *sec,
                      const ref_plastic *pl, const ref_contact_opts *co,
                      const ref_options *opt, const double *disp0) {
  (void)disp0;
  ref_model *m = (ref_model *)calloc(1, sizeof(ref_model));
  m->N = mesh->n_nodes;
  m->E = mesh->n_elems;
  m->R = mesh->n_struts;
  const int64_t N = m->N, E = m->E;
  const int32_t R = m->R;
  m->X0 = (double *)malloc(sizeof(double) * 3 * N);
  memcpy(m->X0, mesh->coords, sizeof(double) * 3 * N);
  m->en = (int32_t *)malloc(sizeof(int32_t) * 2 * E);
  memcpy(m->en, mesh->elem_nodes, sizeof(int32_t) * 2 * E);
  m->es = (int32_t *)malloc(sizeof(int32_t) * E);
  memcpy(m->es, mesh->elem_strut, sizeof(int32_t) * E);
  m->el_local = (int32_t *)malloc(sizeof(int32_t) * E);

  /* struts: element groups "beam_<s+1>" (LatticeModel.cpp:61-74) */
  m->st = (strut_t *)calloc(R, sizeof(strut_t));
  int32_t *cnt = (int32_t *)calloc(R, sizeof(int32_t));
  for (int64_t e = 0; e < E; e++) {
    if (m->es[e] < 0 || m->es[e] >= R) { m->err = REF_ERR_INPUT; free(cnt); return m; }
    cnt[m->es[e]]++;
  }
  for (int32_t s = 0; s < R; s++) {
    m->st[s].elems = (int32_t *)malloc(sizeof(int32_t) * (cnt[s] ? cnt[s] : 1));
    m->st[s].n_el = 0;
  }
  for (int64_t e = 0; e < E; e++) {
    strut_t *s = &m->st[m->es[e]];
    m->el_local[e] = s->n_el;
    s->elems[s->n_el++] = (int32_t)e;
  }
  for (int32_t s = 0; s < R; s++) {
    strut_t *S = &m->st[s];
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * 2 * (S->n_el ? S->n_el : 1));
    for (int i = 0; i < S->n_el; i++) {
      tmp[2 * i] = m->en[2 * S->elems[i]];
      tmp[2 * i + 1] = m->en[2 * S->elems[i] + 1];
    }
    qsort(tmp, 2 * S->n_el, sizeof(int32_t), cmp_i32);
    int n = 0;
    for (int i = 0; i < 2 * S->n_el; i++)
      if (i == 0 || tmp[i] != tmp[i - 1]) tmp[n++] = tmp[i];
    S->n_nd = n;
    S->nodes = tmp;
    S->lambdaN = (double *)calloc(9 * (n ? n : 1), sizeof(double));
  }
  free(cnt);

  /* materials (ElasticRodMaterial.cpp:265-287) */
  m->K = (double *)malloc(sizeof(double) * 6 * R);
  m->Mm = (double *)malloc(sizeof(double) * 6 * R);
  m->area = (double *)malloc(sizeof(double) * R);
  m->rho = (double *)malloc(sizeof(double) * R);
  m->inertia_correct = sec->inertia_correct;
  for (int32_t s = 0; s < R; s++) {
    const double G = sec->shear_modulus[s], Y = sec->young[s], k = sec->shear_factor[s];
    const double A = sec->area[s], I1 = sec->I1[s], I2 = sec->I2[s], J = sec->J[s], rho = sec->density[s];
    double *K = m->K + 6 * s, *M = m->Mm + 6 * s;
    K[0] = G * k * A;
    K[1] = G * k * A;
    K[2] = Y * A;
    K[3] = Y * I1;
    K[4] = Y * I2;
    K[5] = G * J;
    M[0] = rho * A; M[1] = rho * A; M[2] = rho * A;
    M[3] = rho * I1; M[4] = rho * I2; M[5] = rho * J;
    if (sec->inertia_correct != 1.0)
      for (int i = 3; i < 6; i++) M[i] *= sec->inertia_correct;
    m->area[s] = A;
    m->rho[s] = rho;
  }

  /* plastic */
  memset(&m->pl, 0, sizeof m->pl);
  m->pl.model = REF_ELASTIC;
  if (pl) {
    m->pl = *pl;
    if (pl->model == REF_ELASTOPLASTIC) {
      m->pw = (double *)malloc(sizeof(double) * 6 * R);
      memcpy(m->pw, pl->weights, sizeof(double) * 6 * R);
      m->Hk = (double *)malloc(sizeof(double) * 6 * R);
      for (int32_t s = 0; s < R; s++)
        for (int i = 0; i < 6; i++)
          m->Hk[6 * s + i] = pl->kin_const[i] + pl->kin_ratio * m->K[6 * s + i];
    }
    m->pl.weights = NULL;
  }

  m->strain0 = (double *)calloc(6 * E, sizeof(double));
  m->curr_strain = (double *)calloc(6 * E, sizeof(double));
  m->old_strain = (double *)calloc(6 * E, sizeof(double));
  m->epot = (double *)calloc(E, sizeof(double));
  m->curr_ep = (double *)calloc(6 * E, sizeof(double));
  m->old_ep = (double *)calloc(6 * E, sizeof(double));
  m->curr_al = (double *)calloc(7 * E, sizeof(double));
  m->old_al = (double *)calloc(7 * E, sizeof(double));
  m->dflow = (double *)calloc(E, sizeof(double));
  m->ediss = (double *)calloc(E, sizeof(double));
  m->ehard = (double *)calloc(E, sizeof(double));
  m->active = (uint8_t *)calloc(E, 1);

  memset(&m->co, 0, sizeof m->co);
  if (co && co->enabled) {
    m->co = *co;
    m->crad = (double *)malloc(sizeof(double) * R);
    memcpy(m->crad, co->radius, sizeof(double) * R);
    m->co.radius = NULL;
  }
  m->cap_act = 1024;
  m->actA = (int32_t *)malloc(sizeof(int32_t) * m->cap_act);
  m->actB = (int32_t *)malloc(sizeof(int32_t) * m->cap_act);

  m->cmask = (uint8_t *)calloc(6 * N, 1);
  m->cval = (double *)calloc(6 * N, sizeof(double));
  m->fext = (double *)calloc(6 * N, sizeof(double));
  m->so3 = opt ? opt->so3 : 1;
  return m;
}

void ref_destroy(ref_model *m) {
  if (!m) return;
  for (int32_t s = 0; s < m->R; s++) {
    free(m->st[s].elems);
    free(m->st[s].nodes);
    free(m->st[s].lambdaN);
  }
  free(m->st); free(m->X0); free(m->en); free(m->es); free(m->el_local);
  free(m->K); free(m->Mm); free(m->area); free(m->rho); free(m->pw); free(m->Hk);
  free(m->strain0); free(m->curr_strain); free(m->old_strain); free(m->epot);
  free(m->curr_ep); free(m->old_ep); free(m->curr_al); free(m->old_al);
  free(m->dflow); free(m->ediss); free(m->ehard); free(m->active);
  free(m->crad); free(m->black); free(m->actA); free(m->actB);
  free(m->mu_cond); free(m->mu_blocks);
  free(m->cmask); free(m->cval); free(m->fext);
  for (int q = 0; q < m->n_prog; q++) free(m->prog[q].expr);
  free(m->dof_prog); free(m->cval0); free(m->fext0);
  free(m->pm_mass); free(m->pm_inertia);
  free(m->imp_nodes); free(m->imp_ptr); free(m->imp_dof); free(m->imp_w);
  free(m->st_off); free(m->nd_ptr); free(m->nd_row);
  free(m->loc_blk); free(m->loc_f); free(m->loc_g);
  free(m);
}

int ref_set_bcs(ref_model *m, const uint8_t *cmask, const double *cval,
                const double *fext) {
  const int64_t n = 6 * m->N;
  if (cmask) memcpy(m->cmask, cmask, n); else memset(m->cmask, 0, n);
  if (cval) memcpy(m->cval, cval, n * sizeof(double)); else memset(m->cval, 0, n * sizeof(double));
  if (fext) memcpy(m->fext, fext, n * sizeof(double)); else memset(m->fext, 0, n * sizeof(double));
  if (m->cval0) {
    memcpy(m->cval0, m->cval, n * sizeof(double));
    memcpy(m->fext0, m->fext, n * sizeof(double));
  }
  return REF_OK;
}

/* ------------------------------------------------------------------ */
/* Boundary programs                                                   */
/* ------------------------------------------------------------------ */
/* Expression evaluation by recursive descent directly on the string (JEM
 * function syntax subset of LoadScale scaleFunc strings): numbers, t, i,
 * PI, + - * / % ^ (unary minus looser than ^), comparisons, && ||, !,
 * if(c,a,b) and the one- and two-argument functions listed below. */
typedef struct {
  const char *s;
  size_t p;
  double t, i;
  int err;
} ref_ex;

static double ex_or(ref_ex *e);
static void ex_ws(ref_ex *e) {
  while (e->s[e->p] == ' ' || e->s[e->p] == '\t' || e->s[e->p] == '\n') e->p++;
}
static int ex_eat(ref_ex *e, const char *tok) {
  ex_ws(e);
  size_t n = strlen(tok);
  if (strncmp(e->s + e->p, tok, n) == 0) {
    e->p += n;
    return 1;
  }
  return 0;
}
static double ex_unary(ref_ex *e);
static double ex_primary(ref_ex *e) {
  ex_ws(e);
  char c = e->s[e->p];
  if ((c >= '0' && c <= '9') || c == '.') {
    char *end;
    double v = strtod(e->s + e->p, &end);
    if (end == e->s + e->p) { e->err = 1; return 0.; }
    e->p = (size_t)(end - e->s);
    return v;
  }
  if (c == '(') {
    e->p++;
    double v = ex_or(e);
    if (!ex_eat(e, ")")) e->err = 1;
    return v;
  }
  if ((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_') {
    char id[16];
    int n = 0;
    while (((e->s[e->p] >= 'a' && e->s[e->p] <= 'z') || (e->s[e->p] >= 'A' && e->s[e->p] <= 'Z') ||
            (e->s[e->p] >= '0' && e->s[e->p] <= '9') || e->s[e->p] == '_') && n < 15)
      id[n++] = e->s[e->p++];
    id[n] = 0;
    if (!strcmp(id, "t")) return e->t;
    if (!strcmp(id, "i")) return e->i;
    if (!strcmp(id, "PI")) return 3.14159265358979323846;
    double a[3] = {0., 0., 0.};
    int na = (!strcmp(id, "if")) ? 3 : (!strcmp(id, "min") || !strcmp(id, "max") || !strcmp(id, "pow")) ? 2 : 1;
    if (!ex_eat(e, "(")) { e->err = 1; return 0.; }
    for (int k = 0; k < na; k++) {
      if (k > 0 && !ex_eat(e, ",")) { e->err = 1; return 0.; }
      a[k] = ex_or(e);
    }
    if (!ex_eat(e, ")")) { e->err = 1; return 0.; }
    if (!strcmp(id, "if")) return a[0] != 0. ? a[1] : a[2];
    if (!strcmp(id, "min")) return a[0] < a[1] ? a[0] : a[1];
    if (!strcmp(id, "max")) return a[0] > a[1] ? a[0] : a[1];
    if (!strcmp(id, "pow")) return pow(a[0], a[1]);
    if (!strcmp(id, "sin")) return sin(a[0]);
    if (!strcmp(id, "cos")) return cos(a[0]);
    if (!strcmp(id, "tan")) return tan(a[0]);
    if (!strcmp(id, "asin")) return asin(a[0]);
    if (!strcmp(id, "acos")) return acos(a[0]);
    if (!strcmp(id, "atan")) return atan(a[0]);
    if (!strcmp(id, "exp")) return exp(a[0]);
    if (!strcmp(id, "log")) return log(a[0]);
    if (!strcmp(id, "sqrt")) return sqrt(a[0]);
    if (!strcmp(id, "abs")) return fabs(a[0]);
    if (!strcmp(id, "floor")) return floor(a[0]);
    if (!strcmp(id, "ceil")) return ceil(a[0]);
    e->err = 1;
    return 0.;
  }
  e->err = 1;
  return 0.;
}
static double ex_power(ref_ex *e) {
  double x = ex_primary(e);
  ex_ws(e);
  if (e->s[e->p] == '^') {
    e->p++;
    return pow(x, ex_unary(e));
  }
  return x;
}
static double ex_unary(ref_ex *e) {
  ex_ws(e);
  if (e->s[e->p] == '-') { e->p++; return -ex_unary(e); }
  if (e->s[e->p] == '+') { e->p++; return ex_unary(e); }
  if (e->s[e->p] == '!' && e->s[e->p + 1] != '=') { e->p++; return ex_unary(e) == 0. ? 1. : 0.; }
  return ex_power(e);
}
static double ex_mul(ref_ex *e) {
  double x = ex_unary(e);
  for (;;) {
    ex_ws(e);
    char c = e->s[e->p];
    if (c == '*') { e->p++; x = x * ex_unary(e); }
    else if (c == '/') { e->p++; x = x / ex_unary(e); }
    else if (c == '%') { e->p++; x = fmod(x, ex_unary(e)); }
    else return x;
  }
}
static double ex_add(ref_ex *e) {
  double x = ex_mul(e);
  for (;;) {
    ex_ws(e);
    char c = e->s[e->p];
    if (c == '+') { e->p++; x = x + ex_mul(e); }
    else if (c == '-') { e->p++; x = x - ex_mul(e); }
    else return x;
  }
}
static double ex_cmp(ref_ex *e) {
  double x = ex_add(e);
  if (ex_eat(e, "<=")) return x <= ex_add(e);
  if (ex_eat(e, ">=")) return x >= ex_add(e);
  if (ex_eat(e, "==")) return x == ex_add(e);
  if (ex_eat(e, "!=")) return x != ex_add(e);
  if (ex_eat(e, "<")) return x < ex_add(e);
  if (ex_eat(e, ">")) return x > ex_add(e);
  return x;
}
static double ex_and(ref_ex *e) {
  double x = ex_cmp(e);
  while (ex_eat(e, "&&")) {
    double y = ex_cmp(e);
    x = (x != 0.) && (y != 0.);
  }
  return x;
}
static double ex_or(ref_ex *e) {
  double x = ex_and(e);
  while (ex_eat(e, "||")) {
    double y = ex_and(e);
    x = (x != 0.) || (y != 0.);
  }
  return x;
}

int ref_expr_eval(const char *expr, double t, double i, double *value) {
  ref_ex e = {expr, 0, t, i, 0};
  double v = ex_or(&e);
  ex_ws(&e);
  if (e.err || e.s[e.p]) return REF_ERR_INPUT;
  *value = v;
  return REF_OK;
}

/* STEP accumulator as the reference iterates it: NeumannModel keeps
 * loadScale0 and adds loadIncr at every advance, committing the sum
 * (NeumannModel.cpp:347-368); DirichletModel starts from initDisp - dispIncr
 * (DirichletModel.cpp:165-172) and does the same: scale_k = acc_k + b,
 * acc_{k+1} = scale_k, acc_0 = a + b (c - 1) */
static void prog_reset_acc(ref_model *m) {
  for (int q = 0; q < m->n_prog; q++) {
    ref_prog *g = &m->prog[q];
    g->acc = g->a + g->b * ((double)m->clock_k + g->c - 1.);
  }
}

int ref_set_programs(ref_model *m, int n, const int32_t *kind, const double *abc,
                     const char *const *exprs, const int32_t *dof_prog) {
  if (n < 0 || n > REF_MAX_PROG) return REF_ERR_INPUT;
  const int64_t nd = 6 * m->N;
  for (int q = 0; q < m->n_prog; q++) free(m->prog[q].expr);
  m->n_prog = 0;
  free(m->dof_prog);
  m->dof_prog = NULL;
  if (n == 0) {
    if (m->cval0) {
      memcpy(m->cval, m->cval0, nd * sizeof(double));
      memcpy(m->fext, m->fext0, nd * sizeof(double));
    }
    return REF_OK;
  }
  for (int q = 0; q < n; q++) {
    ref_prog *g = &m->prog[q];
    g->kind = kind[q];
    g->a = abc[3 * q]; g->b = abc[3 * q + 1]; g->c = abc[3 * q + 2];
    g->expr = NULL;
    if (g->kind == REF_PROG_EXPR) {
      double v;
      if (!exprs || !exprs[q] || ref_expr_eval(exprs[q], 0., 0., &v)) { m->n_prog = q; return REF_ERR_INPUT; }
      const size_t len = strlen(exprs[q]) + 1;
      g->expr = (char *)malloc(len);
      memcpy(g->expr, exprs[q], len);
    }
  }
  m->n_prog = n;
  m->dof_prog = (int32_t *)malloc(sizeof(int32_t) * nd);
  memcpy(m->dof_prog, dof_prog, sizeof(int32_t) * nd);
  if (!m->cval0) {
    m->cval0 = (double *)malloc(nd * sizeof(double));
    m->fext0 = (double *)malloc(nd * sizeof(double));
    memcpy(m->cval0, m->cval, nd * sizeof(double));
    memcpy(m->fext0, m->fext, nd * sizeof(double));
  }
  prog_reset_acc(m);
  return REF_OK;
}

int ref_set_time(ref_model *m, double t, int64_t k) {
  m->clock_t = t;
  m->clock_k = k;
  prog_reset_acc(m);
  return REF_OK;
}

int ref_get_time(const ref_model *m, double *t, int64_t *k) {
  if (t) *t = m->clock_t;
  if (k) *k = m->clock_k;
  return REF_OK;
}

/* boundary values of the step that ends at time t_end */
static int prog_apply(ref_model *m, double t_end, double *scale) {
  for (int q = 0; q < m->n_prog; q++) {
    const ref_prog *g = &m->prog[q];
    if (g->kind == REF_PROG_CONST) scale[q] = g->a;
    else if (g->kind == REF_PROG_STEP) scale[q] = g->acc + g->b;
    else if (ref_expr_eval(g->expr, t_end, (double)m->clock_k, &scale[q])) return REF_ERR_INPUT;
  }
  const int64_t nd = 6 * m->N;
  for (int64_t i = 0; i < nd; i++) {
    const int32_t q = m->dof_prog[i];
    if (q < 0) continue;
    if (m->cmask[i]) m->cval[i] = m->cval0[i] * scale[q];
    else m->fext[i] = m->fext0[i] * scale[q];
  }
  return REF_OK;
}

/* ExplicitModule::advance (ExplicitModule.cpp:256-266): valid_ = valid_ &&
 * !updCond(t, i); the mass is re-assembled (updateMass_) when not valid, at
 * the step's start state; the clock is the one before the step's time
 * advance.  Without a condition the mass is assembled every step. */
int ref_set_mass_update(ref_model *m, const char *cond) {
  free(m->mu_cond);
  m->mu_cond = NULL;
  m->mu_valid = 0;
  if (!cond) return REF_OK;
  double v;
  if (ref_expr_eval(cond, 0., 0., &v)) return REF_ERR_INPUT;
  m->mu_cond = (char *)malloc(strlen(cond) + 1);
  strcpy(m->mu_cond, cond);
  if (!m->mu_blocks) m->mu_blocks = (double *)malloc(sizeof(double) * 36 * m->N);
  return REF_OK;
}

static int mass_in_use(ref_model *m, const double *disp, double *blocks) {
  if (!m->mu_cond) return ref_mass_blocks(m, disp, blocks, !m->so3);
  double c = 1.;
  if (m->mu_valid && ref_expr_eval(m->mu_cond, m->clock_t, (double)m->clock_k, &c))
    return REF_ERR_INPUT;
  if (!m->mu_valid || c != 0.) {
    ref_mass_blocks(m, disp, m->mu_blocks, !m->so3);
    m->mu_valid = 1;
  }
  memcpy(blocks, m->mu_blocks, sizeof(double) * 36 * m->N);
  return REF_OK;
}

/* the mass of the last update ([36N] blocks), for the tests */
int ref_mass_in_use(ref_model *m, double *blocks) {
  if (!m->mu_cond || !m->mu_valid) return REF_ERR_INPUT;
  memcpy(blocks, m->mu_blocks, sizeof(double) * 36 * m->N);
  return REF_OK;
}

/* the step is taken (committed): the clock and the accumulators move */
static void prog_commit(ref_model *m, double dt, const double *scale) {
  m->clock_t += dt;
  m->clock_k += 1;
  for (int q = 0; q < m->n_prog; q++)
    if (m->prog[q].kind == REF_PROG_STEP) m->prog[q].acc = scale[q];
}

int ref_set_point_masses(ref_model *m, const double *mass, const double *inertia) {
  const int64_t N = m->N;
  free(m->pm_mass); free(m->pm_inertia);
  m->pm_mass = m->pm_inertia = NULL;
  if (mass) { m->pm_mass = (double *)malloc(N * sizeof(double)); memcpy(m->pm_mass, mass, N * sizeof(double)); }
  if (inertia) { m->pm_inertia = (double *)malloc(N * sizeof(double)); memcpy(m->pm_inertia, inertia, N * sizeof(double)); }
  return REF_OK;
}

/* MassModel's GET_MATRIX2 (MassModel.cpp:92-108): diagonal additions */
static void add_point_masses(const ref_model *m, double *blocks) {
  for (int64_t n = 0; n < m->N; n++) {
    double *B = blocks + 36 * n;
    if (m->pm_mass)
      for (int i = 0; i < 3; i++) B[7 * i] += m->pm_mass[n];
    if (m->pm_inertia)
      for (int i = 3; i < 6; i++) B[7 * i] += m->pm_inertia[n];
  }
}

/* ------------------------------------------------------------------ */
/* Line3D (2-node)                                                     */
/* ------------------------------------------------------------------ */
/* getShapeGradients (Line3D.cpp:102-115) with Gauss1: g[2], w */
static void shape_grads_g1(const double c0[3], const double c1[3], double g[2], double *w) {
  double d[3] = {c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]};
  const double L = norm3(d) + 0.;
  /* ParametricLine: J = dN0/du * 0 + dN1/du * L  (local coords c1 = [0, L]) */
  const double J = -0.5 * 0. + 0.5 * L;
  g[0] = -0.5 / J;
  g[1] = 0.5 / J;
  *w = 2. * J;
}

/* getRefRot_ + getNodeRotVecs_ (Line3D.cpp:249-273) */
static void node_rot_vecs(const double Rn0[9], const double Rn1[9], double Lr[9],
                          double psi0[3], double psi1[3]) {
  double T[9], phi[3], E[9];
  mtm3(Rn0, Rn1, T);
  ref_logmat(T, phi);
  phi[0] = 0.5 * phi[0]; phi[1] = 0.5 * phi[1]; phi[2] = 0.5 * phi[2];
  ref_expvec(phi, E);
  mm3(Rn0, E, Lr);
  mtm3(Lr, Rn0, T);
  ref_logmat(T, psi0);
  mtm3(Lr, Rn1, T);
  ref_logmat(T, psi1);
}

/* ------------------------------------------------------------------ */
/* materials                                                           */
/* ------------------------------------------------------------------ */
static void elastic_stress(const double *K, const double *strain, double *stress) {
  for (int i = 0; i < 6; i++) stress[i] = K[i] * strain[i];
}

/* power-sum yield family: value and analytic gradient over (s, h0, hk[6])
 *   f = sum_i |w_i a_i|^p_i - Y (1 + ci q0),  a_i = s_i + ck_i q_i
 * p_i = 1 and p_i = 2 are evaluated without pow(). */
static double pterm(double x, double p) { /* |x|^p */
  const double ax = fabs(x);
  if (p == 1.0) return ax;
  if (p == 2.0) return ax * ax;
  return pow(ax, p);
}
static double pderiv(double x, double p) { /* d|x|^p / dx without the sign */
  const double ax = fabs(x);
  if (p == 1.0) return 1.0;
  if (p == 2.0) return 2.0 * ax;
  return p * pow(ax, p - 1.0);
}
static double yield_value(const ref_model *m, const double *w, const double s[6],
                          const double q[7]) {
  double f = 0.;
  for (int i = 0; i < 6; i++) {
    const double a = s[i] + (m->pl.has_kin ? m->pl.kin_coupling[i] * q[1 + i] : 0.);
    f += pterm(w[i] * a, m->pl.exponent[i]);
  }
  f -= m->pl.yield_level * (1. + (m->pl.has_iso ? m->pl.iso_coupling * q[0] : 0.));
  return f;
}
static double sgn(double x) { return (x > 0) - (x < 0); }
static void yield_grad(const ref_model *m, const double *w, const double s[6],
                       const double q[7], double gs[6], double gq[7]) {
  for (int i = 0; i < 6; i++) {
    const double a = s[i] + (m->pl.has_kin ? m->pl.kin_coupling[i] * q[1 + i] : 0.);
    const double d = pderiv(w[i] * a, m->pl.exponent[i]) * w[i] * sgn(a);
    gs[i] = d;
    gq[1 + i] = m->pl.has_kin ? d * m->pl.kin_coupling[i] : 0.;
    /* numeric-derivative path: zero where |args_i| < eps (ElastoPlasticRodMaterial.cpp:249-253) */
    if (fabs(s[i]) < DBL_EPSILON) gs[i] = 0.;
  }
  gq[0] = m->pl.has_iso ? -m->pl.yield_level * m->pl.iso_coupling : 0.;
}

/* Material::getStress(stress, strain, ielem, ip, inelastic) */
static int material_stress(ref_model *m, int64_t e, const double strain[6],
                           double stress[6], int inelastic) {
  const int32_t s = m->es[e];
  const double *K = m->K + 6 * s;
  if (m->pl.model == REF_ELASTIC) { /* ElasticRodMaterial.cpp:363-367 */
    memcpy(m->curr_strain + 6 * e, strain, 6 * sizeof(double));
    elastic_stress(K, strain, stress);
    return REF_OK;
  }
  /* ElastoPlasticRodMaterial.cpp:198-269 (cutting plane) */
  const double *w = m->pw + 6 * s;
  const double *Hk = m->Hk + 6 * s;
  double ep[6], al[7], q[7], gs[6], gq[7], el[6];
  memcpy(ep, m->old_ep + 6 * e, sizeof ep);
  memcpy(al, m->old_al + 7 * e, sizeof al);
  double dflow = 0.;
  int liter = 0;
  int took = 0;
  for (;;) {
    for (int i = 0; i < 6; i++) el[i] = strain[i] - ep[i];
    elastic_stress(K, el, stress);
    /* getHardVals: q = -H alpha, H = blockdiag(H_iso, diag(Hk)) */
    q[0] = -1. * (m->pl.has_iso ? m->pl.iso_modulus * al[0] : 0.);
    for (int i = 0; i < 6; i++) q[1 + i] = -1. * (m->pl.has_kin ? Hk[i] * al[1 + i] : 0.);
    if (!inelastic) break;
    const double f = yield_value(m, w, stress, q);
    if (!(liter < m->pl.max_iter)) { m->err = REF_ERR_PLASTIC_ITER; return REF_ERR_PLASTIC_ITER; }
    if (f < m->pl.precision) break;
    yield_grad(m, w, stress, q, gs, gq);
    double den = 0.;
    {
      double a = gs[0] * (K[0] * gs[0]);
      for (int i = 1; i < 6; i++) a += gs[i] * (K[i] * gs[i]);
      double b = 0.;
      int first = 1;
      if (m->pl.has_iso) { b = gq[0] * (m->pl.iso_modulus * gq[0]); first = 0; }
      if (m->pl.has_kin)
        for (int i = 0; i < 6; i++) {
          const double t = gq[1 + i] * (Hk[i] * gq[1 + i]);
          if (first) { b = t; first = 0; } else b += t;
        }
      den = a + b;
    }
    const double ddl = f / den;
    for (int i = 0; i < 6; i++) ep[i] += ddl * gs[i];
    if (m->pl.has_iso) al[0] += ddl * gq[0];
    if (m->pl.has_kin)
      for (int i = 0; i < 6; i++) al[1 + i] += ddl * gq[1 + i];
    dflow += ddl;
    liter++;
    took = 1;
  }
  memcpy(m->curr_strain + 6 * e, strain, 6 * sizeof(double));
  memcpy(m->curr_ep + 6 * e, ep, sizeof ep);
  memcpy(m->curr_al + 7 * e, al, sizeof al);
  m->dflow[e] = dflow;
  m->active[e] = (uint8_t)took;
  return REF_OK;
}

/* applyDeform (ElasticRodMaterial.cpp:377-389; ElastoPlasticRodMaterial.cpp:271-307) */
static void apply_deform(ref_model *m) {
  const int64_t E = m->E;
  if (m->pl.model == REF_ELASTIC) {
    for (int64_t e = 0; e < E; e++) {
      const double *K = m->K + 6 * m->es[e];
      const double *c = m->curr_strain + 6 * e;
      double s[6];
      elastic_stress(K, c, s);
      double d = c[0] * s[0];
      for (int i = 1; i < 6; i++) d += c[i] * s[i];
      m->epot[e] = 0.5 * d;
    }
    memcpy(m->old_strain, m->curr_strain, 6 * E * sizeof(double));
    return;
  }
  for (int64_t e = 0; e < E; e++) {
    const int32_t s = m->es[e];
    const double *K = m->K + 6 * s;
    const double *Hk = m->Hk + 6 * s;
    double oel[6], cel[6], dep[6], os[6], cs[6];
    for (int i = 0; i < 6; i++) {
      oel[i] = m->old_strain[6 * e + i] - m->old_ep[6 * e + i];
      cel[i] = m->curr_strain[6 * e + i] - m->curr_ep[6 * e + i];
      dep[i] = m->curr_ep[6 * e + i] - m->old_ep[6 * e + i];
    }
    elastic_stress(K, oel, os);
    elastic_stress(K, cel, cs);
    {
      const double *a = m->curr_al + 7 * e;
      double h = 0.;
      int first = 1;
      if (m->pl.has_iso) { h = a[0] * (m->pl.iso_modulus * a[0]); first = 0; }
      if (m->pl.has_kin)
        for (int i = 0; i < 6; i++) {
          const double t = a[1 + i] * (Hk[i] * a[1 + i]);
          if (first) { h = t; first = 0; } else h += t;
        }
      m->ehard[e] = 0.5 * h;
    }
    double d = cel[0] * cs[0];
    for (int i = 1; i < 6; i++) d += cel[i] * cs[i];
    m->epot[e] = 0.5 * d;
    if (fabs(m->dflow[e]) > DBL_EPSILON) {
      double t = ((os[0] + cs[0]) / 2.) * dep[0];
      for (int i = 1; i < 6; i++) t += ((os[i] + cs[i]) / 2.) * dep[i];
      m->ediss[e] += t;
    }
  }
  memcpy(m->old_al, m->curr_al, 7 * E * sizeof(double));
  memcpy(m->old_ep, m->curr_ep, 6 * E * sizeof(double));
  memcpy(m->old_strain, m->curr_strain, 6 * E * sizeof(double));
  memset(m->dflow, 0, E * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* rod element                                                         */
/* ------------------------------------------------------------------ */
/* getDisplacments_ (SpecialCosseratRodModel.cpp:773-798): X0, u, Lambda */
static void elem_gather(const ref_model *m, int64_t e, const double *disp,
                        double c[2][3], double u[2][3], double L[2][9]) {
  const strut_t *S = &m->st[m->es[e]];
  for (int k = 0; k < 2; k++) {
    const int32_t n = m->en[2 * e + k];
    for (int i = 0; i < 3; i++) {
      c[k][i] = m->X0[3 * n + i];
      u[k][i] = disp ? disp[6 * n + i] : 0.;
    }
    double th[3] = {disp ? disp[6 * n + 3] : 0., disp ? disp[6 * n + 4] : 0.,
                    disp ? disp[6 * n + 5] : 0.};
    double Ex[9];
    ref_expvec(th, Ex);
    const int ln = rod_local_node(S, n);
    mm3(Ex, S->lambdaN + 9 * ln, L[k]);
  }
}

/* getStrains_ (SpecialCosseratRodModel.cpp:694-744), material frame, plus the
 * interpolated rotation Lambda_ip and phi' used by getXi / getPi. */
static void elem_strains(const double c[2][3], const double u[2][3],
                         const double L[2][9], const double *strain0,
                         double strain[6], double Lip[9], double phiP[3],
                         double g[2], double *w) {
  shape_grads_g1(c[0], c[1], g, w);
  for (int i = 0; i < 3; i++) {
    double s = (c[0][i] + u[0][i]) * g[0];
    s += (c[1][i] + u[1][i]) * g[1];
    phiP[i] = s;
  }
  double Lr[9], psi0[3], psi1[3];
  node_rot_vecs(L[0], L[1], Lr, psi0, psi1);
  /* getRotations: ip_psi = node_psi * N,  N = (0.5, 0.5) at Gauss1 */
  double ippsi[3], ippsiP[3], Ex[9], LP[9], T[9];
  for (int i = 0; i < 3; i++) {
    double s = psi0[i] * 0.5;
    s += psi1[i] * 0.5;
    ippsi[i] = s;
    double t = psi0[i] * g[0];
    t += psi1[i] * g[1];
    ippsiP[i] = t;
  }
  ref_expvec(ippsi, Ex);
  mm3(Lr, Ex, Lip);
  /* getRotationGradients */
  ref_expvecp(ippsi, ippsiP, T);
  mm3(Lr, T, LP);
  mtv3(Lip, phiP, strain);
  mtm3(Lip, LP, T);
  unskew3(T, strain + 3);
  if (strain0)
    for (int i = 0; i < 6; i++) strain[i] -= strain0[i];
}

int ref_elem_fint(ref_model *m, int64_t e, const double *disp, double *f, int inelastic) {
  double c[2][3], u[2][3], L[2][9], strain[6], Lip[9], phiP[3], g[2], w, sm[6], ss[6];
  elem_gather(m, e, disp, c, u, L);
  elem_strains(c, u, L, m->strain0 + 6 * e, strain, Lip, phiP, g, &w);
  int rc = material_stress(m, e, strain, sm, inelastic);
  if (rc) return rc;
  /* spatial: PI * stress (SpecialCosseratRodModel.cpp:764-770) */
  mv3(Lip, sm, ss);
  mv3(Lip, sm + 3, ss + 3);
  /* XI (Line3D.cpp:141-162) and fint[Idofs] += w * XI * stress (:945) */
  double SK[9];
  skew3(phiP, SK);
  for (int k = 0; k < 2; k++) {
    const double N = 0.5;
    double r[6];
    for (int i = 0; i < 3; i++) r[i] = g[k] * ss[i];
    for (int i = 0; i < 3; i++) {
      double t = (-1. * N * SK[3 * i + 0]) * ss[0];
      t += (-1. * N * SK[3 * i + 1]) * ss[1];
      t += (-1. * N * SK[3 * i + 2]) * ss[2];
      t += g[k] * ss[3 + i];
      r[3 + i] = t;
    }
    for (int i = 0; i < 6; i++) f[6 * k + i] = w * r[i];
  }
  return REF_OK;
}

/* ------------------------------------------------------------------ */
/* INIT: initRotation_ + initStrain_ (SpecialCosseratRodModel.cpp:536-658) */
/* ------------------------------------------------------------------ */
static void init_rotation(ref_model *m, strut_t *S) {
  const int n = S->n_nd;
  double *dirs = (double *)calloc(3 * (n ? n : 1), sizeof(double));
  for (int i = 0; i < S->n_el; i++) {
    const int64_t e = S->elems[i];
    const int32_t a = m->en[2 * e], b = m->en[2 * e + 1];
    const int la = rod_local_node(S, a), lb = rod_local_node(S, b);
    for (int d = 0; d < 3; d++) {
      const double t = m->X0[3 * b + d] - m->X0[3 * a + d];
      dirs[3 * la + d] += t;
      dirs[3 * lb + d] += t;
    }
  }
  for (int k = 0; k < n; k++) {
    double *d = dirs + 3 * k;
    const double nn = norm3(d);
    d[0] = d[0] / nn; d[1] = d[1] / nn; d[2] = d[2] / nn;
    /* v = e3^T skew(d) = e3 x d; c = d . e3 */
    double v[3] = {-1 * d[1], 1 * d[0], 0.0};
    const double cc = d[0] * 0. + d[1] * 0. + d[2] * 1.;
    double *Rm = S->lambdaN + 9 * k;
    eye3(Rm);
    if (fabs(cc + 1.0) > DBL_EPSILON) {
      double Sv[9], SS[9];
      skew3(v, Sv);
      mm3(Sv, Sv, SS);
      const double f = 1 / (1 + cc);
      for (int i = 0; i < 9; i++) Rm[i] += Sv[i] + f * SS[i];
    } else {
      for (int i = 0; i < 9; i++) Rm[i] *= -1.;
    }
  }
  free(dirs);
}

/* ------------------------------------------------------------------ */
/* contact                                                             */
/* ------------------------------------------------------------------ */
static int ref_contains(double u) { return u >= -1.0 && u <= 1.0; }

static void global_point(double u, const double p0[3], const double p1[3], double x[3]) {
  const double N0 = 0.5 * (1. - u), N1 = 0.5 * (1. + u);
  for (int i = 0; i < 3; i++) {
    double s = p0[i] * N0;
    s += p1[i] * N1;
    x[i] = s;
  }
}

static void closest_points(const double pA[2][3], const double pB[2][3], double *uA, double *uB) {
  double bA[3], bB[3], tA[3], tB[3], d[3], vA[3], vB[3];
  for (int i = 0; i < 3; i++) {
    bA[i] = pA[1][i] + pA[0][i];
    bB[i] = pB[1][i] + pB[0][i];
    tA[i] = pA[1][i] - pA[0][i];
    tB[i] = pB[1][i] - pB[0][i];
    d[i] = bB[i] - bA[i];
  }
  const double tBtA = dot3(tB, tA), tBtB = dot3(tB, tB), tAtA = dot3(tA, tA);
  for (int i = 0; i < 3; i++) {
    vA[i] = tB[i] * tBtA - tA[i] * tBtB;
    vB[i] = tA[i] * tBtA - tB[i] * tAtA;
  }
  const double den = tBtB * tAtA - tBtA * tBtA;
  *uA = -1. * dot3(d, vA) / den;
  *uB = +1. * dot3(d, vB) / den;
}

static double closest_point(const double s[3], const double pM[2][3]) {
  double bM[3], tM[3];
  for (int i = 0; i < 3; i++) {
    bM[i] = pM[1][i] + pM[0][i];
    tM[i] = pM[1][i] - pM[0][i];
  }
  return (2 * dot3(s, tM) - dot3(tM, bM)) / dot3(tM, tM);
}

/* computeSTS_ force part (RodContactModel.cpp:1031-1152); f may be NULL */
static int contact_sts(const double pA[2][3], const double pB[2][3], double uA, double uB,
                       double R, double kap, double f[12]) {
  double xA[3], xB[3], d[3];
  global_point(uA, pA[0], pA[1], xA);
  global_point(uB, pB[0], pB[1], xB);
  for (int i = 0; i < 3; i++) d[i] = xB[i] - xA[i];
  const double dist = norm3(d);
  if (REF_IS_TINY(dist)) return 0;
  if (dist > R) return 0;
  if (!f) return 1;
  double n[3] = {d[0] / dist, d[1] / dist, d[2] / dist};
#if REF_PLANT == 2
  const double NA[2] = {0.5 * (1. + uA), 0.5 * (1. - uA)};
#else
  const double NA[2] = {0.5 * (1. - uA), 0.5 * (1. + uA)};
#endif
  const double NB[2] = {0.5 * (1. - uB), 0.5 * (1. + uB)};
#if REF_PLANT == 1
  const double sc = kap * (R - dist);
#else
  const double sc = kap * (dist - R);
#endif
  for (int i = 0; i < 3; i++) {
    f[0 + i] += sc * ((-1. * NA[0]) * n[i]);
    f[3 + i] += sc * ((-1. * NA[1]) * n[i]);
    f[6 + i] += sc * (NB[0] * n[i]);
    f[9 + i] += sc * (NB[1] * n[i]);
  }
  return 1;
}

/* computeNTS_ force part (RodContactModel.cpp:1157-1208); f[9] may be NULL */
static int contact_nts(const double s[3], const double pM[2][3], double uM, double R,
                       double kap, double f[9]) {
  double xM[3], d[3];
  global_point(uM, pM[0], pM[1], xM);
  for (int i = 0; i < 3; i++) d[i] = xM[i] - s[i];
  const double dist = norm3(d);
  if (REF_IS_TINY(dist)) return 0;
  if (dist > R) return 0;
  if (!f) return 1;
  double n[3];
  for (int i = 0; i < 3; i++) n[i] = (s[i] - xM[i]) / dist;
#if REF_PLANT == 3
  const double sc = 2. * kap * (dist - R);
#else
  const double sc = kap * (dist - R);
#endif
#if REF_PLANT == 4
  const double a = -.5 * (1. + uM), b = -.5 * (1. - uM);
#else
  const double a = -.5 * (1. - uM), b = -.5 * (1. + uM);
#endif
  for (int i = 0; i < 3; i++) {
    f[i] += sc * n[i];
    f[3 + i] += sc * (a * n[i]);
    f[6 + i] += sc * (b * n[i]);
  }
  return 1;
}

static void node_pos(const ref_model *m, int32_t n, const double *disp, double x[3]) {
  for (int i = 0; i < 3; i++) x[i] = m->X0[3 * n + i] + (disp ? disp[6 * n + i] : 0.);
}

static int strut_has_node(const ref_model *m, int32_t s, int32_t n) {
  return rod_local_node(&m->st[s], n) >= 0;
}

/* element AABB inflated by the strut's contact radius */
static void elem_box(const ref_model *m, int64_t e, const double *disp, double box[6]) {
  double a[3], b[3];
  node_pos(m, m->en[2 * e], disp, a);
  node_pos(m, m->en[2 * e + 1], disp, b);
  const double r = m->crad[m->es[e]];
  for (int i = 0; i < 3; i++) {
    box[i] = (a[i] < b[i] ? a[i] : b[i]) - r;
    box[3 + i] = (a[i] > b[i] ? a[i] : b[i]) + r;
  }
}
static int box_overlap(const double A[6], const double B[6]) {
  int ok = 1;
  for (int i = 0; i < 3; i++) ok &= A[3 + i] >= B[i] && B[3 + i] >= A[i];
  return ok;
}
/* element A may meet strut B: neither of its nodes belongs to strut B (RodContactModel.cpp:393,413) */
static int elem_free_of(const ref_model *m, int64_t e, int32_t s) {
  return !(strut_has_node(m, s, m->en[2 * e]) || strut_has_node(m, s, m->en[2 * e + 1]));
}

typedef struct {
  int64_t n, cap;
  int32_t *a, *b;
} pairbuf;
static void pb_push(pairbuf *p, int32_t a, int32_t b) {
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 1024;
    p->a = (int32_t *)realloc(p->a, sizeof(int32_t) * p->cap);
    p->b = (int32_t *)realloc(p->b, sizeof(int32_t) * p->cap);
  }
  p->a[p->n] = a;
  p->b[p->n] = b;
  p->n++;
}

/* findPossibleBeams_ + findPossibleElements_ (RodContactModel.cpp:245-446), O(R^2) */
static void broadphase_brute(ref_model *m, const double *disp, pairbuf *out) {
  const int32_t R = m->R;
  for (int32_t s = 0; s < R; s++) {
    strut_t *S = &m->st[s];
    const double r = m->crad[s];
    for (int i = 0; i < 3; i++) { S->box[i] = INFINITY; S->box[3 + i] = -INFINITY; }
    for (int k = 0; k < S->n_nd; k++) {
      double x[3];
      node_pos(m, S->nodes[k], disp, x);
      for (int i = 0; i < 3; i++) {
        if (x[i] < S->box[i]) S->box[i] = x[i];
        if (x[i] > S->box[3 + i]) S->box[3 + i] = x[i];
      }
    }
    for (int i = 0; i < 3; i++) { S->box[i] -= r; S->box[3 + i] += r; }
  }
  for (int32_t i = 0; i < R; i++)
    for (int32_t j = R - 1; j > i; j--) {
      if (!box_overlap(m->st[i].box, m->st[j].box)) continue;
      const strut_t *A = &m->st[i], *B = &m->st[j];
      for (int ka = 0; ka < A->n_el; ka++) {
        const int64_t ea = A->elems[ka];
        if (!elem_free_of(m, ea, j)) continue;
        double ba[6];
        elem_box(m, ea, disp, ba);
        for (int kb = 0; kb < B->n_el; kb++) {
          const int64_t eb = B->elems[kb];
          if (!elem_free_of(m, eb, i)) continue;
          double bb[6];
          elem_box(m, eb, disp, bb);
          if (box_overlap(ba, bb)) pb_push(out, (int32_t)ea, (int32_t)eb);
        }
      }
    }
}

/* Uniform-grid broadphase producing the same candidate set (order differs):
 * cell = element AABB min corner / h, h >= largest AABB extent. */
static void broadphase_grid(ref_model *m, const double *disp, pairbuf *out) {
  const int64_t E = m->E;
  double *boxes = (double *)malloc(sizeof(double) * 6 * E);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, ext = 0.;
  for (int64_t e = 0; e < E; e++) {
    double *b = boxes + 6 * e;
    elem_box(m, e, disp, b);
    for (int i = 0; i < 3; i++) {
      if (b[i] < lo[i]) lo[i] = b[i];
      if (b[3 + i] - b[i] > ext) ext = b[3 + i] - b[i];
    }
  }
  const double h = ext * 1.0001 + 1e-300;
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t e = 0; e < E; e++)
    for (int i = 0; i < 3; i++)
      if (boxes[6 * e + i] > hi[i]) hi[i] = boxes[6 * e + i];
  int64_t dim[3];
  for (int i = 0; i < 3; i++) dim[i] = (int64_t)floor((hi[i] - lo[i]) / h) + 1;
  const int64_t ncell = dim[0] * dim[1] * dim[2];
  int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * E);
  int64_t *start = (int64_t *)calloc(ncell + 1, sizeof(int64_t));
  int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * E);
  for (int64_t e = 0; e < E; e++) {
    int64_t c[3];
    for (int i = 0; i < 3; i++) {
      c[i] = (int64_t)floor((boxes[6 * e + i] - lo[i]) / h);
      if (c[i] < 0) c[i] = 0;
      if (c[i] >= dim[i]) c[i] = dim[i] - 1;
    }
    cell[e] = (c[2] * dim[1] + c[1]) * dim[0] + c[0];
    start[cell[e] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ncell);
  memcpy(fill, start, sizeof(int64_t) * ncell);
  for (int64_t e = 0; e < E; e++) sorted[fill[cell[e]]++] = (int32_t)e;
#ifdef REF_OMP
  /* all-cores variant: count per element, then fill at the prefix offsets,
   * which keeps the serial candidate order */
  int64_t *cnt = (int64_t *)calloc(E + 1, sizeof(int64_t));
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int64_t a = 0; a < E; a++) cnt[a + 1] += cnt[a];
      out->a = (int32_t *)malloc(sizeof(int32_t) * (cnt[E] ? cnt[E] : 1));
      out->b = (int32_t *)malloc(sizeof(int32_t) * (cnt[E] ? cnt[E] : 1));
      out->n = out->cap = cnt[E];
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t a = 0; a < E; a++) {
      int64_t k_out = pass ? cnt[a] : 0;
#else
  for (int64_t a = 0; a < E; a++) {
#endif
    const int64_t ca = cell[a];
    const int64_t cx = ca % dim[0], cy = (ca / dim[0]) % dim[1], cz = ca / (dim[0] * dim[1]);
    const int32_t sa = m->es[a];
    for (int64_t dz = -1; dz <= 1; dz++)
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dx = -1; dx <= 1; dx++) {
          const int64_t x = cx + dx, y = cy + dy, z = cz + dz;
          if (x < 0 || y < 0 || z < 0 || x >= dim[0] || y >= dim[1] || z >= dim[2]) continue;
          const int64_t c = (z * dim[1] + y) * dim[0] + x;
          for (int64_t k = start[c]; k < start[c + 1]; k++) {
            const int32_t b = sorted[k];
            const int32_t sb = m->es[b];
            if (sb <= sa) continue; /* A belongs to the lower strut index */
            if (!elem_free_of(m, a, sb) || !elem_free_of(m, b, sa)) continue;
            if (box_overlap(boxes + 6 * a, boxes + 6 * b)) {
#ifdef REF_OMP
              if (pass) { out->a[k_out] = (int32_t)a; out->b[k_out] = b; k_out++; }
              else cnt[a + 1]++;
#else
              pb_push(out, (int32_t)a, b);
#endif
            }
          }
        }
  }
#ifdef REF_OMP
  }
  free(cnt);
#endif
  free(boxes); free(cell); free(start); free(sorted); free(fill);
}

static void broadphase(ref_model *m, const double *disp, pairbuf *out) {
  out->n = 0;
  if (m->use_grid) broadphase_grid(m, disp, out);
  else broadphase_brute(m, disp, out);
}

static int is_blacklisted(const ref_model *m, int32_t a, int32_t b) {
  const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
  const int64_t key = (lo << 32) | hi;
  int64_t l = 0, h = m->n_black - 1;
  while (l <= h) {
    const int64_t mid = (l + h) >> 1;
    if (m->black[mid] == key) return 1;
    if (m->black[mid] < key) l = mid + 1; else h = mid - 1;
  }
  return 0;
}

/* per-candidate narrowphase shared by computeContacts_ and computeBlacklist_.
 * Returns 1 if the contact is closed; when f != NULL fills f[12] (A0,A1,B0,B1
 * translational) with the contribution laid out as in computeContacts_. */
static int narrowphase(const ref_model *m, int32_t ea, int32_t eb, const double *disp,
                       double f[12]) {
  const int32_t nA[2] = {m->en[2 * ea], m->en[2 * ea + 1]};
  const int32_t nB[2] = {m->en[2 * eb], m->en[2 * eb + 1]};
  if (nA[0] == nB[0] || nA[0] == nB[1] || nA[1] == nB[0] || nA[1] == nB[1]) return 0;
  double pA[2][3], pB[2][3];
  for (int k = 0; k < 2; k++) {
    node_pos(m, nA[k], disp, pA[k]);
    node_pos(m, nB[k], disp, pB[k]);
  }
  const double R = m->crad[m->es[ea]] + m->crad[m->es[eb]];
  const double kS = m->co.penalty_sts, kN = m->co.penalty_nts;
  double uA, uB;
  closest_points(pA, pB, &uA, &uB);
  if (f) memset(f, 0, 12 * sizeof(double));
  if (ref_contains(uA) && ref_contains(uB)) {
    return contact_sts(pA, pB, uA, uB, R, kS, f);
  } else if (!ref_contains(uA) && ref_contains(uB)) {
    const int iA = (uA < -1.) ? 0 : 1;
    uB = closest_point(pB[iA], pA); /* RodContactModel.cpp:575: node of B projected on A */
    if (!ref_contains(uB)) return 0;
    double g[9] = {0};
    if (!contact_nts(pA[iA], pB, uB, R, kN, f ? g : NULL)) return 0;
    if (f) {
      for (int i = 0; i < 3; i++) {
        f[3 * iA + i] = g[i];
        f[6 + i] = g[3 + i];
        f[9 + i] = g[6 + i];
      }
    }
    return 1;
  } else if (!ref_contains(uB) && ref_contains(uA)) {
    const int iB = (uB < -1.) ? 0 : 1;
    uA = closest_point(pB[iB], pA);
    if (!ref_contains(uA)) return 0;
    double g[9] = {0};
    if (!contact_nts(pB[iB], pA, uA, R, kN, f ? g : NULL)) return 0;
    if (f) {
      for (int i = 0; i < 3; i++) {
        f[6 + 3 * iB + i] = g[i];
        f[0 + i] = g[3 + i];
        f[3 + i] = g[6 + i];
      }
    }
    return 1;
  } else {
    int closed = 0;
    for (int iA = 0; iA < 2; iA++) {
      const double u = closest_point(pA[iA], pB);
      if (!ref_contains(u)) continue;
      double g[9] = {0};
      if (contact_nts(pA[iA], pB, u, R, kN, f ? g : NULL)) {
        if (f)
          for (int i = 0; i < 3; i++) {
            f[3 * iA + i] += g[i];
            f[6 + i] += g[3 + i];
            f[9 + i] += g[6 + i];
          }
        closed = 1;
      }
    }
    for (int iB = 0; iB < 2; iB++) {
      const double u = closest_point(pB[iB], pA);
      if (!ref_contains(u)) continue;
      double g[9] = {0};
      if (contact_nts(pB[iB], pA, u, R, kN, f ? g : NULL)) {
        if (f)
          for (int i = 0; i < 3; i++) {
            f[0 + i] += g[3 + i];
            f[3 + i] += g[6 + i];
            f[6 + 3 * iB + i] += g[i];
          }
        closed = 1;
      }
    }
    return closed;
  }
}

static void compute_blacklist(ref_model *m, const double *disp) { /* RodContactModel.cpp:743-936 */
  pairbuf pb = {0};
  broadphase(m, disp, &pb);
  free(m->black);
  m->black = (int64_t *)malloc(sizeof(int64_t) * (pb.n ? pb.n : 1));
  m->n_black = 0;
  for (int64_t i = 0; i < pb.n; i++)
    if (narrowphase(m, pb.a[i], pb.b[i], disp, NULL)) {
      const int64_t a = pb.a[i], b = pb.b[i];
      const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
      m->black[m->n_black++] = (lo << 32) | hi;
    }
  qsort(m->black, m->n_black, sizeof(int64_t), cmp_i64);
  free(pb.a); free(pb.b);
}

static void act_push(ref_model *m, int32_t a, int32_t b) {
  if (m->n_act == m->cap_act) {
    m->cap_act *= 2;
    m->actA = (int32_t *)realloc(m->actA, sizeof(int32_t) * m->cap_act);
    m->actB = (int32_t *)realloc(m->actB, sizeof(int32_t) * m->cap_act);
  }
  m->actA[m->n_act] = a;
  m->actB[m->n_act] = b;
  m->n_act++;
}

int64_t ref_contact_force(ref_model *m, const double *disp, double *fint) {
  m->n_act = 0;
  if (!m->co.enabled) return 0;
  pairbuf pb = {0};
  broadphase(m, disp, &pb);
#ifdef REF_OMP
  /* all-cores variant: pair forces in parallel, added in candidate order */
  double *pf = (double *)malloc(sizeof(double) * 12 * (pb.n ? pb.n : 1));
  uint8_t *closed = (uint8_t *)malloc(pb.n ? pb.n : 1);
  #pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < pb.n; i++)
    closed[i] = !is_blacklisted(m, pb.a[i], pb.b[i]) &&
                narrowphase(m, pb.a[i], pb.b[i], disp, pf + 12 * i);
#endif
  for (int64_t i = 0; i < pb.n; i++) {
    const int32_t ea = pb.a[i], eb = pb.b[i];
#ifdef REF_OMP
    if (!closed[i]) continue;
    const double *f = pf + 12 * i;
#else
    if (is_blacklisted(m, ea, eb)) continue;
    double f[12];
    if (!narrowphase(m, ea, eb, disp, f)) continue;
#endif
    const int32_t nd[4] = {m->en[2 * ea], m->en[2 * ea + 1], m->en[2 * eb], m->en[2 * eb + 1]};
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) fint[6 * nd[k] + d] += f[3 * k + d];
    act_push(m, ea, eb);
  }
#ifdef REF_OMP
  free(pf); free(closed);
#endif
  free(pb.a); free(pb.b);
  return m->n_act;
}

int64_t ref_contact_candidates(ref_model *m, const double *disp, int32_t *a, int32_t *b, int64_t cap) {
  if (!m->co.enabled) return 0;
  pairbuf pb = {0};
  broadphase(m, disp, &pb);
  for (int64_t i = 0; i < pb.n && i < cap; i++) { a[i] = pb.a[i]; b[i] = pb.b[i]; }
  const int64_t n = pb.n;
  free(pb.a); free(pb.b);
  return n;
}

int64_t ref_contact_active(ref_model *m, int32_t *a, int32_t *b, int64_t cap) {
  for (int64_t i = 0; i < m->n_act && i < cap; i++) { a[i] = m->actA[i]; b[i] = m->actB[i]; }
  return m->n_act;
}

int64_t ref_blacklist(ref_model *m, int32_t *a, int32_t *b, int64_t cap) {
  for (int64_t i = 0; i < m->n_black && i < cap; i++) {
    a[i] = (int32_t)(m->black[i] >> 32);
    b[i] = (int32_t)(m->black[i] & 0xffffffff);
  }
  return m->n_black;
}

void ref_set_grid(ref_model *m, int use_grid) { m->use_grid = use_grid; }

int ref_init(ref_model *m, const double *disp) {
  m->mu_valid = 0; /* ExplicitModule::init ends with updateMass_ */
  for (int32_t s = 0; s < m->R; s++) init_rotation(m, &m->st[s]);
  /* initStrain_: strains at null displacement, Lambda = LambdaN */
  for (int64_t e = 0; e < m->E; e++) {
    double c[2][3], u[2][3], L[2][9], Lip[9], phiP[3], g[2], w;
    elem_gather(m, e, NULL, c, u, L);
    elem_strains(c, u, L, NULL, m->strain0 + 6 * e, Lip, phiP, g, &w);
  }
  if (m->co.enabled) compute_blacklist(m, disp);
  m->imp_ready = 0; /* no output yet: ImpactBC dofs start free */
  return m->err;
}

int ref_frames(ref_model *m, double *lambda_n, double *strain0) {
  for (int64_t e = 0; e < m->E; e++) {
    const strut_t *S = &m->st[m->es[e]];
    for (int k = 0; k < 2; k++) {
      const int ln = rod_local_node(S, m->en[2 * e + k]);
      if (lambda_n) memcpy(lambda_n + 18 * e + 9 * k, S->lambdaN + 9 * ln, 9 * sizeof(double));
    }
  }
  if (strain0) memcpy(strain0, m->strain0, 6 * m->E * sizeof(double));
  return REF_OK;
}

/* ------------------------------------------------------------------ */
/* mass (assembleM_, SpecialCosseratRodModel.cpp:969-1031)             */
/* ------------------------------------------------------------------ */
/* assembleM_ contributions of one element (SpecialCosseratRodModel.cpp:
 * 969-1031): SI[in][jn] is the 6x6 block that row node in receives for
 * column node jn, in the order the reference adds them */
static void elem_mass_si(const ref_model *m, int32_t s, int64_t e, const double *disp,
                         double SI[2][2][36]) {
  const double xi[2] = {-1. / sqrt(3.), 1. / sqrt(3.)};
  const double *Mm = m->Mm + 6 * s;
  double c[2][3], u[2][3], L[2][9];
  elem_gather(m, e, disp, c, u, L);
  /* Gauss2 rotations (getRotations with massShape) */
  double Lr[9], psi0[3], psi1[3], Lip[2][9];
  node_rot_vecs(L[0], L[1], Lr, psi0, psi1);
  double sh[2][2]; /* sh[node][ip] */
  for (int ip = 0; ip < 2; ip++) {
    sh[0][ip] = 0.5 * (1. - xi[ip]);
    sh[1][ip] = 0.5 * (1. + xi[ip]);
    double ps[3], Ex[9];
    for (int i = 0; i < 3; i++) {
      double t = psi0[i] * sh[0][ip];
      t += psi1[i] * sh[1][ip];
      ps[i] = t;
    }
    ref_expvec(ps, Ex);
    mm3(Lr, Ex, Lip[ip]);
  }
  /* getIntegrationWeights: w_ip = 1 * |J|, J = 0.5 L */
  double d[3] = {c[1][0] - c[0][0], c[1][1] - c[0][1], c[1][2] - c[0][2]};
  const double Lc = norm3(d) + 0.;
  const double Jc = -0.5 * 0. + 0.5 * Lc;
  const double wts[2] = {1. * Jc, 1. * Jc};
  const double l = (wts[0] + wts[1]) / (2 - 1);
  /* material masses */
  const double Mt[9] = {Mm[0], 0, 0, 0, Mm[1], 0, 0, 0, Mm[2]};
  double Mr[9] = {Mm[3] * l, 0, 0, 0, Mm[4] * l, 0, 0, 0, Mm[5] * l};
  const double add = m->area[s] * m->rho[s] * pow(l, 3) / 12.;
  Mr[0] += add;
  Mr[4] += add;
  for (int in = 0; in < 2; in++) {
    for (int jn = 0; jn < 2; jn++) {
      double *S = SI[in][jn];
      memset(S, 0, 36 * sizeof(double));
      for (int ip = 0; ip < 2; ip++) {
        double T[9], T2[9];
        mm3(Lip[ip], Mt, T);
        mmt3(T, Lip[ip], T2);
        const double sc = wts[ip] * sh[in][ip] * sh[jn][ip];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) S[6 * i + j] += sc * T2[3 * i + j];
      }
      if (in == jn) {
        double T[9], T2[9];
        mm3(L[in], Mr, T);
        mmt3(T, L[in], T2);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) S[6 * (3 + i) + 3 + j] += T2[3 * i + j];
        if (in == 0 || in == 1)
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S[6 * (3 + i) + 3 + j] /= 2.;
      }
    }
  }
}

static void add_block(double *B, const double *S, int lumped) {
  if (lumped) {
    for (int i = 0; i < 6; i++) B[7 * i] += S[7 * i];
  } else {
    for (int i = 0; i < 36; i++) B[i] += S[i];
  }
}

#ifndef REF_OMP
int ref_mass_blocks(ref_model *m, const double *disp, double *blocks, int lumped) {
  const int64_t N = m->N;
  memset(blocks, 0, sizeof(double) * 36 * N);
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      double SI[2][2][36];
      elem_mass_si(m, s, e, disp, SI);
      for (int in = 0; in < 2; in++)
        for (int jn = 0; jn < 2; jn++) add_block(blocks + 36 * m->en[2 * e + in], SI[in][jn], lumped);
    }
  }
  add_point_masses(m, blocks);
  return REF_OK;
}
#endif

/* ------------------------------------------------------------------ */
/* GET_INT_VECTOR                                                      */
/* ------------------------------------------------------------------ */
#ifdef REF_OMP
/* All-cores variant (libref_cpu_omp.so, the CPU baseline on every host
 * core): struts run in parallel, each accumulating its elements in order
 * into strut-local node rows; every node then sums its struts' rows in strut
 * order.  Deterministic for any thread count; equal to the serial path up
 * to the association of the sums at nodes shared by several struts. */
static int omp_prepare(ref_model *m) {
  if (m->nd_ptr) return REF_OK;
  const int64_t N = m->N;
  m->st_off = (int64_t *)malloc(sizeof(int64_t) * (m->R + 1));
  m->st_off[0] = 0;
  for (int32_t s = 0; s < m->R; s++) m->st_off[s + 1] = m->st_off[s] + m->st[s].n_nd;
  m->nd_ptr = (int64_t *)calloc(N + 1, sizeof(int64_t));
  for (int32_t s = 0; s < m->R; s++)
    for (int k = 0; k < m->st[s].n_nd; k++) m->nd_ptr[m->st[s].nodes[k] + 1]++;
  for (int64_t n = 0; n < N; n++) m->nd_ptr[n + 1] += m->nd_ptr[n];
  m->nd_row = (int64_t *)malloc(sizeof(int64_t) * (m->nd_ptr[N] ? m->nd_ptr[N] : 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (N ? N : 1));
  memcpy(fill, m->nd_ptr, sizeof(int64_t) * N);
  for (int32_t s = 0; s < m->R; s++)
    for (int k = 0; k < m->st[s].n_nd; k++)
      m->nd_row[fill[m->st[s].nodes[k]]++] = m->st_off[s] + k;
  free(fill);
  const int64_t rows = m->st_off[m->R];
  m->loc_blk = (double *)malloc(sizeof(double) * 36 * (rows ? rows : 1));
  m->loc_f = (double *)malloc(sizeof(double) * 6 * (rows ? rows : 1));
  m->loc_g = (double *)malloc(sizeof(double) * 6 * (rows ? rows : 1));
  return m->loc_blk && m->loc_f && m->loc_g ? REF_OK : REF_ERR_INPUT;
}

int ref_mass_blocks(ref_model *m, const double *disp, double *blocks, int lumped) {
  if (omp_prepare(m)) return REF_ERR_INPUT;
  #pragma omp parallel for schedule(dynamic, 16)
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    double *loc = m->loc_blk + 36 * m->st_off[s];
    memset(loc, 0, sizeof(double) * 36 * S->n_nd);
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      double SI[2][2][36];
      elem_mass_si(m, s, e, disp, SI);
      for (int in = 0; in < 2; in++)
        for (int jn = 0; jn < 2; jn++)
          add_block(loc + 36 * rod_local_node(S, m->en[2 * e + in]), SI[in][jn], lumped);
    }
  }
  #pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < m->N; n++) {
    double *B = blocks + 36 * n;
    memset(B, 0, sizeof(double) * 36);
    for (int64_t t = m->nd_ptr[n]; t < m->nd_ptr[n + 1]; t++)
      for (int i = 0; i < 36; i++) B[i] += m->loc_blk[36 * m->nd_row[t] + i];
  }
  add_point_masses(m, blocks);
  return REF_OK;
}

static int int_vector(ref_model *m, const double *disp, const double *velo, double *fint,
                      double *fgyro, const double *blocks, int inelastic) {
  if (omp_prepare(m)) return REF_ERR_INPUT;
  int rc_all = REF_OK;
  #pragma omp parallel for schedule(dynamic, 16)
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    double *lf = m->loc_f + 6 * m->st_off[s], *lg = m->loc_g + 6 * m->st_off[s];
    memset(lf, 0, sizeof(double) * 6 * S->n_nd);
    memset(lg, 0, sizeof(double) * 6 * S->n_nd);
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      double f[12];
      const int rc = ref_elem_fint(m, e, disp, f, inelastic);
      if (rc) {
        #pragma omp critical
        rc_all = rc;
        break;
      }
      for (int k = 0; k < 2; k++) {
        double *r = lf + 6 * rod_local_node(S, m->en[2 * e + k]);
        for (int i = 0; i < 6; i++) r[i] += f[6 * k + i];
      }
    }
    if (fgyro) /* assembleGyro_: per rod, its nodes */
      for (int k = 0; k < S->n_nd; k++) {
        const int32_t n = S->nodes[k];
        const double *B = blocks + 36 * n;
        const double *v = velo + 6 * n;
        double t[3], Sw[9], g[3];
        for (int i = 0; i < 3; i++) {
          double a = B[6 * (3 + i)] * v[0];
          for (int j = 1; j < 6; j++) a += B[6 * (3 + i) + j] * v[j];
          t[i] = a;
        }
        skew3(v + 3, Sw);
        mv3(Sw, t, g);
        for (int i = 0; i < 3; i++) lg[6 * k + 3 + i] += g[i];
      }
  }
  if (rc_all) return rc_all;
  #pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < m->N; n++)
    for (int64_t t = m->nd_ptr[n]; t < m->nd_ptr[n + 1]; t++) {
      const int64_t r = m->nd_row[t];
      for (int i = 0; i < 6; i++) fint[6 * n + i] += m->loc_f[6 * r + i];
      if (fgyro)
        for (int i = 0; i < 6; i++) fgyro[6 * n + i] += m->loc_g[6 * r + i];
    }
  return REF_OK;
}
#else
static int int_vector(ref_model *m, const double *disp, const double *velo, double *fint,
                      double *fgyro, const double *blocks, int inelastic) {
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      double f[12];
      const int rc = ref_elem_fint(m, e, disp, f, inelastic);
      if (rc) return rc;
      for (int k = 0; k < 2; k++) {
        const int32_t n = m->en[2 * e + k];
        for (int i = 0; i < 6; i++) fint[6 * n + i] += f[6 * k + i];
      }
    }
    if (fgyro && REF_PLANT != 6) { /* assembleGyro_ (SpecialCosseratRodModel.cpp:951-967): per rod, its nodes */
      for (int k = 0; k < S->n_nd; k++) {
        const int32_t n = S->nodes[k];
        const double *B = blocks + 36 * n;
        const double *v = velo + 6 * n;
        double t[3], Sw[9], g[3];
        for (int i = 0; i < 3; i++) {
          double a = B[6 * (3 + i)] * v[0];
          for (int j = 1; j < 6; j++) a += B[6 * (3 + i) + j] * v[j];
          t[i] = a;
        }
        skew3(v + 3, Sw);
        mv3(Sw, t, g);
        for (int i = 0; i < 3; i++) fgyro[6 * n + 3 + i] += (REF_PLANT == 5 ? -1. : 1.) * g[i];
      }
    }
  }
#if REF_PLANT == 6
  if (fgyro)
    for (int64_t n = 0; n < m->N; n++) {
      const double *B = blocks + 36 * n, *v = velo + 6 * n;
      double t[3], Sw[9], g[3];
      for (int i = 0; i < 3; i++) {
        double a = B[6 * (3 + i)] * v[0];
        for (int j = 1; j < 6; j++) a += B[6 * (3 + i) + j] * v[j];
        t[i] = a;
      }
      skew3(v + 3, Sw);
      mv3(Sw, t, g);
      for (int i = 0; i < 3; i++) fgyro[6 * n + 3 + i] += g[i];
    }
#endif
  return REF_OK;
}
#endif

int ref_get_int_vector(ref_model *m, const double *disp, const double *velo, double *fint,
                       int inelastic, int with_gyro, int with_contact) {
  double *blocks = NULL, *fg = NULL;
  if (with_gyro) {
    blocks = (double *)malloc(sizeof(double) * 36 * m->N);
    ref_mass_blocks(m, disp, blocks, 0);
    fg = (double *)calloc(6 * m->N, sizeof(double));
  }
  int rc = int_vector(m, disp, velo, fint, fg, blocks, inelastic);
  if (!rc && with_contact) ref_contact_force(m, disp, fint);
  if (fg)
    for (int64_t i = 0; i < 6 * m->N; i++) fint[i] += fg[i];
  free(blocks);
  free(fg);
  return rc;
}

/* solve the 6x6 block system with constrained dofs (reduced system, Gauss
 * elimination with partial pivoting) */
static void solve_block(const double *B, const double *f, const uint8_t *cm, const double *cv,
                        double *a) {
  int fr[6], nf = 0;
  for (int i = 0; i < 6; i++) {
    if (cm[i]) a[i] = cv[i];
    else fr[nf++] = i;
  }
  if (!nf) return;
  double A[6][7];
  for (int r = 0; r < nf; r++) {
    double rhs = f[fr[r]];
    for (int j = 0; j < 6; j++)
      if (cm[j]) rhs -= B[6 * fr[r] + j] * cv[j];
    for (int c = 0; c < nf; c++) A[r][c] = B[6 * fr[r] + fr[c]];
    A[r][nf] = rhs;
  }
  for (int k = 0; k < nf; k++) {
    int p = k;
    for (int r = k + 1; r < nf; r++)
      if (fabs(A[r][k]) > fabs(A[p][k])) p = r;
    if (p != k)
      for (int c = 0; c <= nf; c++) { double t = A[k][c]; A[k][c] = A[p][c]; A[p][c] = t; }
    for (int r = k + 1; r < nf; r++) {
      const double l = A[r][k] / A[k][k];
      for (int c = k; c <= nf; c++) A[r][c] -= l * A[k][c];
    }
  }
  double x[6];
  for (int k = nf - 1; k >= 0; k--) {
    double s = A[k][nf];
    for (int c = k + 1; c < nf; c++) s -= A[k][c] * x[c];
    x[k] = s / A[k][k];
  }
  for (int r = 0; r < nf; r++) a[fr[r]] = x[r];
}

int ref_commit(ref_model *m, const double *disp, double *e_pot, double *e_diss) {
  (void)disp;
  apply_deform(m);
  /* getPotentialEnergy_ / getDissipatedEnergy_: sum_e sum_I w N_I energy_e */
  double ep = 0., ed = 0.;
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    double sp = 0., sd = 0.;
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      const int32_t a = m->en[2 * e], b = m->en[2 * e + 1];
      double g[2], w;
      shape_grads_g1(m->X0 + 3 * a, m->X0 + 3 * b, g, &w);
      for (int k = 0; k < 2; k++) {
        sp += w * 0.5 * m->epot[e];
        sd += w * 0.5 * (m->pl.model == REF_ELASTOPLASTIC ? m->ediss[e] : 0.);
      }
    }
    ed += sd;
    ep += sp;
  }
  if (e_pot) *e_pot = ep;
  if (e_diss) *e_diss = ed;
  return REF_OK;
}

/* GroupOutputModule::run (GroupOutputModule.cpp:100-116,162) after a step:
 * LOAD_CASE "output", the model's GET_INT_VECTOR at the current state, the
 * group's dof summed.  The rod elements use the elastic predictor from the
 * committed plastic state (SpecialCosseratRodModel.cpp:762 passes
 * inelastic = false to getStress, ElastoPlasticRodMaterial.cpp:198-269 then
 * skips the return map); the contact model searches the contacts again at
 * this state (RodContactModel.cpp:153-166, the update condition defaults to
 * true, :84-88) and adds their forces.  Only elements with a node in the
 * group reach the group's dofs, so only those are evaluated; the material
 * state is not written (the reference overwrites its current state here,
 * but the next step's evaluation overwrites it again before anything reads
 * it). */
static int output_reaction(ref_model *m, const double *disp, const int32_t *group,
                           int64_t group_n, int32_t dof, double *r) {
  const int64_t N = m->N;
  double *f = (double *)calloc(6 * N, sizeof(double));
  char *in = (char *)calloc(N, 1);
  for (int64_t k = 0; k < group_n; k++) in[group[k]] = 1;
  for (int32_t s = 0; s < m->R; s++) {
    const strut_t *S = &m->st[s];
    for (int ie = 0; ie < S->n_el; ie++) {
      const int64_t e = S->elems[ie];
      const int32_t a = m->en[2 * e], b = m->en[2 * e + 1];
      if (!in[a] && !in[b]) continue;
      double c[2][3], u[2][3], L[2][9], strain[6], Lip[9], phiP[3], g[2], w, sm[6], ss[6];
      elem_gather(m, e, disp, c, u, L);
      elem_strains(c, u, L, m->strain0 + 6 * e, strain, Lip, phiP, g, &w);
      const double *K = m->K + 6 * s;
      if (m->pl.model == REF_ELASTIC) {
        elastic_stress(K, strain, sm);
      } else {
        double el[6];
        for (int i = 0; i < 6; i++) el[i] = strain[i] - m->old_ep[6 * e + i];
        elastic_stress(K, el, sm);
      }
      mv3(Lip, sm, ss);
      mv3(Lip, sm + 3, ss + 3);
      double SK[9];
      skew3(phiP, SK);
      for (int k = 0; k < 2; k++) {
        double rr[6];
        for (int i = 0; i < 3; i++) rr[i] = g[k] * ss[i];
        for (int i = 0; i < 3; i++) {
          double t = (-1. * 0.5 * SK[3 * i + 0]) * ss[0];
          t += (-1. * 0.5 * SK[3 * i + 1]) * ss[1];
          t += (-1. * 0.5 * SK[3 * i + 2]) * ss[2];
          t += g[k] * ss[3 + i];
          rr[3 + i] = t;
        }
        const int32_t n = m->en[2 * e + k];
        for (int i = 0; i < 6; i++) f[6 * n + i] += w * rr[i];
      }
    }
  }
  if (m->co.enabled) ref_contact_force(m, disp, f);
  double sum = 0.;
  for (int64_t k = 0; k < group_n; k++) sum += f[6 * (int64_t)group[k] + dof];
  *r = sum;
  free(f);
  free(in);
  return REF_OK;
}

int ref_output_reaction(ref_model *m, const double *disp, const int32_t *group, int64_t group_n,
                        int32_t dof, double *r) {
  return output_reaction(m, disp, group, group_n, dof, r);
}

/* ExplicitModule::advance/commit + LeapFrogModule::solve with the intended
 * store (LeapFrogModule.cpp:44-84; ExplicitModule.cpp:259-419).  react[st]:
 * the GroupOutput sum after step st (output_reaction at the new state). */
int ref_leapfrog(ref_model *m, double *disp, double *velo, double dt, int64_t nsteps,
                 const int32_t *group, int64_t group_n, int32_t group_dof, double *react,
                 int commit_energy, double *energies, double *unused, int64_t *n_contacts) {
  (void)unused;
  const int64_t N = m->N, nd = 6 * N;
  double *blocks = (double *)malloc(sizeof(double) * 36 * N);
  double *fint = (double *)malloc(sizeof(double) * nd);
  double *fg = (double *)malloc(sizeof(double) * nd);
  double *acc = (double *)malloc(sizeof(double) * nd);
  int rc = REF_OK;
  double scale[REF_MAX_PROG];
  for (int64_t st = 0; st < nsteps && !rc; st++) {
    /* advance: boundary values at the new time; updCond default true ->
     * updateMass_ at the current state */
    if (m->n_prog && (rc = prog_apply(m, m->clock_t + dt, scale))) break;
    if ((rc = mass_in_use(m, disp, blocks))) break;
    if (!m->so3) {
      for (int64_t n = 0; n < N; n++)
        for (int i = 0; i < 6; i++)
          if (!(blocks[36 * n + 7 * i] > DBL_MIN)) { rc = REF_ERR_ZERO_MASS; m->err = rc; }
      if (rc) break;
    }
    /* getForce */
    memset(fint, 0, sizeof(double) * nd);
    memset(fg, 0, sizeof(double) * nd);
    rc = int_vector(m, disp, velo, fint, m->so3 ? fg : NULL, blocks, 1);
    if (rc) break;
    int64_t nc = 0;
    if (m->co.enabled) nc = ref_contact_force(m, disp, fint);
    if (n_contacts) n_contacts[st] = nc;
    if (m->imp_n) impact_constraints(m, fint);
    /* fres = fext - fint; getAcce */
    REF_PARALLEL_FOR
    for (int64_t n = 0; n < N; n++) {
      double fr[6];
      for (int i = 0; i < 6; i++) fr[i] = m->fext[6 * n + i] - (fint[6 * n + i] + fg[6 * n + i]);
      if (m->so3) {
        solve_block(blocks + 36 * n, fr, m->cmask + 6 * n, m->cval + 6 * n, acc + 6 * n);
      } else {
        for (int i = 0; i < 6; i++) {
          acc[6 * n + i] = (1 / blocks[36 * n + 7 * i]) * fr[i];
          if (m->cmask[6 * n + i]) acc[6 * n + i] = m->cval[6 * n + i];
        }
      }
    }
    /* v_new = v + dt a;  u_new = u + dt v_new (+ SO(3) composition) */
    REF_PARALLEL_FOR
    for (int64_t i = 0; i < nd; i++) velo[i] = velo[i] + dt * acc[i];
    REF_PARALLEL_FOR
    for (int64_t n = 0; n < N; n++) {
      double *u = disp + 6 * n, *v = velo + 6 * n;
      double th_old[3] = {u[3], u[4], u[5]};
      for (int i = 0; i < 6; i++) u[i] = u[i] + dt * v[i];
      if (m->so3) {
        double dr[3] = {dt * v[3], dt * v[4], dt * v[5]};
        double Ro[9], Vu[9], Rn[9];
        ref_expvec(th_old, Ro);
        ref_expvec(dr, Vu);
        mm3(Vu, Ro, Rn);
        ref_logmat(Rn, u + 3);
      }
    }
    /* commit (error = 0 -> accepted; dt pinned) */
    if (commit_energy && energies) {
      double ep, ed;
      ref_commit(m, disp, &ep, &ed);
      double ek = 0.;
      for (int64_t n = 0; n < N; n++) {
        const double *B = blocks + 36 * n, *v = velo + 6 * n;
        double t = 0.;
        for (int i = 0; i < 6; i++) {
          double a = B[6 * i] * v[0];
          for (int j = 1; j < 6; j++) a += B[6 * i + j] * v[j];
          t += v[i] * a;
        }
        ek += 0.5 * t;
      }
      energies[3 * st] = ep;
      energies[3 * st + 1] = ed;
      energies[3 * st + 2] = ek;
    } else {
      apply_deform(m);
    }
    if (m->n_prog || m->mu_cond) prog_commit(m, dt, scale);
    m->imp_ready = 1;
    /* output modules after the solver's commit */
    if (react) output_reaction(m, disp, group, group_n, group_dof, react + st);
  }
  free(blocks); free(fint); free(fg); free(acc);
  return rc;
}

/* SO(3)-aware state update y_new = y_old (+) dy (ExplicitModule.cpp:360-388):
 * translations add; rotation vectors compose as log(exp(dr) exp(r_old)) */
static void update_vec(const ref_model *m, const double *y_old, const double *dy,
                       double *y_new, int rot) {
  const int64_t N = m->N;
  REF_PARALLEL_FOR
  for (int64_t n = 0; n < N; n++) {
    for (int i = 0; i < 6; i++) y_new[6 * n + i] = y_old[6 * n + i] + dy[6 * n + i];
    if (rot && m->so3) {
      double Ro[9], Vu[9], Rn[9];
      ref_expvec(y_old + 6 * n + 3, Ro);
      ref_expvec(dy + 6 * n + 3, Vu);
      mm3(Vu, Ro, Rn);
      ref_logmat(Rn, y_new + 6 * n + 3);
    }
  }
}

/* getQuality (ExplicitModule.cpp:490-499): RMS of y_pre - y_cor over all
 * dofs, translational dofs divided by the length scale */
static double quality(const ref_model *m, const double *y_pre, const double *y_cor,
                      double len_scale) {
  const int64_t nd = 6 * m->N;
  double s = 0.;
  for (int64_t i = 0; i < nd; i++) {
    double d = y_pre[i] - y_cor[i];
    if (i % 6 < 3 || !m->so3) d /= len_scale;
    s += d * d;
  }
  return sqrt(s) / sqrt((double)nd);
}

/* ImpactBC (ImpactModel.cpp:74-104): group g's dof gets the prescribed
 * acceleration -R_g / w_g, R_g the group's internal-force sum at the start
 * state (the reference reads it from GroupOutput's last output); before the
 * first output (m->imp_ready == 0) the dof is free. */
static void impact_constraints(ref_model *m, const double *fint) {
  for (int g = 0; g < m->imp_n; g++) {
    const int32_t dof = m->imp_dof[g];
    double R = 0.;
    for (int32_t j = m->imp_ptr[g]; j < m->imp_ptr[g + 1]; j++)
      R += fint[6 * (int64_t)m->imp_nodes[j] + dof];
    const double a = -R / m->imp_w[g];
    for (int32_t j = m->imp_ptr[g]; j < m->imp_ptr[g + 1]; j++) {
      const int64_t i = 6 * (int64_t)m->imp_nodes[j] + dof;
      m->cmask[i] = (uint8_t)(m->imp_ready ? 1 : 0);
      m->cval[i] = a;
    }
  }
}

int ref_set_impact_groups(ref_model *m, int n, const int32_t *nodes, const int32_t *ptr,
                          const int32_t *dofs, const double *weights) {
  free(m->imp_nodes); free(m->imp_ptr); free(m->imp_dof); free(m->imp_w);
  m->imp_nodes = m->imp_ptr = m->imp_dof = NULL;
  m->imp_w = NULL;
  m->imp_n = n;
  m->imp_ready = 0;
  if (n <= 0) { m->imp_n = 0; return REF_OK; }
  const int32_t nn = ptr[n];
  m->imp_ptr = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  m->imp_nodes = (int32_t *)malloc(sizeof(int32_t) * (nn ? nn : 1));
  m->imp_dof = (int32_t *)malloc(sizeof(int32_t) * n);
  m->imp_w = (double *)malloc(sizeof(double) * n);
  memcpy(m->imp_ptr, ptr, sizeof(int32_t) * (n + 1));
  memcpy(m->imp_nodes, nodes, sizeof(int32_t) * nn);
  memcpy(m->imp_dof, dofs, sizeof(int32_t) * n);
  memcpy(m->imp_w, weights, sizeof(double) * n);
  return REF_OK;
}

/* internal forces at (disp, velo) with the mass blocks of the step's start
 * state (getForce / updForce): fint (+ contact), gyro into fg */
static int milne_forces(ref_model *m, const double *disp, const double *velo,
                        const double *blocks, double *fint, double *fg, int64_t *nc) {
  const int64_t nd = 6 * m->N;
  memset(fint, 0, sizeof(double) * nd);
  memset(fg, 0, sizeof(double) * nd);
  int rc = int_vector(m, disp, velo, fint, m->so3 ? fg : NULL, blocks, 1);
  if (rc) return rc;
  const int64_t c = m->co.enabled ? ref_contact_force(m, disp, fint) : 0;
  if (nc) *nc = c;
  return REF_OK;
}

/* a = M^-1 (fext - fint - fgyro) with the constraints (getAcce) */
static void milne_solve(ref_model *m, const double *blocks, const double *fint, const double *fg,
                        double *acc) {
  const int64_t N = m->N;
  REF_PARALLEL_FOR
  for (int64_t n = 0; n < N; n++) {
    double fr[6];
    for (int i = 0; i < 6; i++) fr[i] = m->fext[6 * n + i] - (fint[6 * n + i] + fg[6 * n + i]);
    if (m->so3) {
      solve_block(blocks + 36 * n, fr, m->cmask + 6 * n, m->cval + 6 * n, acc + 6 * n);
    } else {
      for (int i = 0; i < 6; i++) {
        acc[6 * n + i] = (1 / blocks[36 * n + 7 * i]) * fr[i];
        if (m->cmask[6 * n + i]) acc[6 * n + i] = m->cval[6 * n + i];
      }
    }
  }
}

/* MilneDevice attempts (MilneDeviceModule.cpp:41-112) with the adaptive step
 * of ExplicitModule::commit (ExplicitModule.cpp:290-335), order 1:
 *   a_cur = a(u, v);  v_pre = v + dt a_cur;  u_pre = u (+) dt v
 *   a_pre = a(u_pre, v_pre) (mass of the start state);
 *   v_new = v + dt a_pre;  u_new = u (+) dt v_pre
 *   err = (quality(u_pre, u_new) + dt quality(v_pre, v_new)) / 2
 *   dt_opt = dt (prec / err)^(1/2); accepted if err <= prec or dt <= dt_min,
 *   then dt = max(min(s dt_opt, incr dt, dt_max), decr dt, dt_min), the
 *   material state of the last evaluation is committed and the state moves
 *   to (u_new, v_new); otherwise dt = max(s dt_opt, decr dt, dt_min).
 * Per attempt: out[6*k] = dt used, err, accepted, the GroupOutput sum after
 * the attempt (output_reaction at the accepted state, or at the start
 * state a rejected attempt leaves in place), contact pairs at the start
 * state, new dt. */
int ref_milne(ref_model *m, double *disp, double *velo, const ref_milne_opts *o,
              int64_t n_attempts, const int32_t *group, int64_t group_n, int32_t group_dof,
              double *out) {
  const int64_t N = m->N, nd = 6 * N;
  double *blocks = (double *)malloc(sizeof(double) * 36 * N);
  double *fint = (double *)malloc(sizeof(double) * nd);
  double *fg = (double *)malloc(sizeof(double) * nd);
  double *acc = (double *)malloc(sizeof(double) * nd);
  double *dy = (double *)malloc(sizeof(double) * nd);
  double *u_pre = (double *)malloc(sizeof(double) * nd);
  double *v_pre = (double *)malloc(sizeof(double) * nd);
  double *u_new = (double *)malloc(sizeof(double) * nd);
  double *v_new = (double *)malloc(sizeof(double) * nd);
  double dt = o->dt;
  int rc = REF_OK;
  double scale[REF_MAX_PROG];
  for (int64_t k = 0; k < n_attempts && !rc; k++) {
    /* advance: boundary values at the attempt's end time */
    if (m->n_prog && (rc = prog_apply(m, m->clock_t + dt, scale))) break;
    if ((rc = mass_in_use(m, disp, blocks))) break;
    int64_t nc = 0;
    if ((rc = milne_forces(m, disp, velo, blocks, fint, fg, &nc))) break;
    if (m->imp_n) impact_constraints(m, fint);
    milne_solve(m, blocks, fint, fg, acc);
    /* predictor (ABupdate: dy = dt f) */
    for (int64_t i = 0; i < nd; i++) dy[i] = dt * acc[i];
    update_vec(m, velo, dy, v_pre, 0);
    for (int64_t i = 0; i < nd; i++) dy[i] = dt * velo[i];
    update_vec(m, disp, dy, u_pre, 1);
    if ((rc = milne_forces(m, u_pre, v_pre, blocks, fint, fg, NULL))) break;
    milne_solve(m, blocks, fint, fg, acc);
    /* corrector (AMupdate_: dy = dt f_pre) */
    for (int64_t i = 0; i < nd; i++) dy[i] = dt * acc[i];
    update_vec(m, velo, dy, v_new, 0);
    for (int64_t i = 0; i < nd; i++) dy[i] = dt * v_pre[i];
    update_vec(m, disp, dy, u_new, 1);
    double err = 0.;
    err += quality(m, u_pre, u_new, o->len_scale);
    err += quality(m, v_pre, v_new, o->len_scale) * dt;
    err = 0.5 * err;
    /* ExplicitModule::commit, order 1 */
    const double dt_opt = dt * pow(o->precision / err, 1. / (1. + 1.));
    const int accept = (err <= o->precision || dt <= o->dt_min);
    double dt_next;
    if (accept) {
      double t = o->safety * dt_opt;
      if (o->incr * dt < t) t = o->incr * dt;
      if (o->dt_max < t) t = o->dt_max;
      if (t < o->decr * dt) t = o->decr * dt;
      if (t < o->dt_min) t = o->dt_min;
      dt_next = t;
      apply_deform(m);
      memcpy(disp, u_new, sizeof(double) * nd);
      memcpy(velo, v_new, sizeof(double) * nd);
      if (m->n_prog || m->mu_cond) prog_commit(m, dt, scale);
      m->imp_ready = 1;
    } else {
      double t = o->safety * dt_opt;
      if (t < o->decr * dt) t = o->decr * dt;
      if (t < o->dt_min) t = o->dt_min;
      dt_next = t;
    }
    double r = 0.;
    if (group_n > 0) output_reaction(m, disp, group, group_n, group_dof, &r);
    if (out) {
      out[6 * k] = dt;
      out[6 * k + 1] = err;
      out[6 * k + 2] = accept;
      out[6 * k + 3] = r;
      out[6 * k + 4] = (double)nc;
      out[6 * k + 5] = dt_next;
    }
    dt = dt_next;
  }
  free(blocks); free(fint); free(fg); free(acc); free(dy);
  free(u_pre); free(v_pre); free(u_new); free(v_new);
  return rc;
}

/* GET_TABLE at the state disp (SpecialCosseratRodModel.cpp:221-262,
 * 426-531): kind 0 strain (PI x material), 1 mat_strain, 2 stress (PI x
 * material), 3 mat_stress, the material stress with the default load case
 * (return map from the committed state; getStresses_'s loadCase = "").
 * out [E][6] (may be NULL); epot [E] (may be NULL): w 1/2 (eps - eps_p) .
 * sigma of that evaluation (ElastoPlasticRodMaterial.cpp:291-292). */
int ref_elem_table(ref_model *m, const double *disp, int kind, double *out, double *epot) {
  for (int64_t e = 0; e < m->E; e++) {
    double c[2][3], u[2][3], L[2][9], strain[6], Lip[9], phiP[3], g[2], w, sm[6];
    elem_gather(m, e, disp, c, u, L);
    elem_strains(c, u, L, m->strain0 + 6 * e, strain, Lip, phiP, g, &w);
    int rc = material_stress(m, e, strain, sm, 1);
    if (rc) return rc;
    if (epot) {
      const double *ep = (m->pl.model == REF_ELASTIC) ? NULL : m->curr_ep + 6 * e;
      double t = 0.;
      for (int i = 0; i < 6; i++) t += (strain[i] - (ep ? ep[i] : 0.)) * sm[i];
      epot[e] = w * (0.5 * t);
    }
    if (!out) continue;
    const double *v = (kind <= 1) ? strain : sm;
    double *o = out + 6 * e;
    if (kind == 1 || kind == 3) {
      for (int i = 0; i < 6; i++) o[i] = v[i];
    } else {
      mv3(Lip, v, o);
      mv3(Lip, v + 3, o + 3);
    }
  }
  return REF_OK;
}

/* LatticeModel::getKineticEnergy (LatticeModel.cpp:195-222): per node
 * 1/2 v . M v with the solver's mass blocks at disp (lumped in the additive
 * rotation mode) */
int ref_kinetic_energy(ref_model *m, const double *disp, const double *velo, double *ekin) {
  const int64_t N = m->N;
  double *blocks = (double *)malloc(sizeof(double) * 36 * N);
  ref_mass_blocks(m, disp, blocks, !m->so3);
  for (int64_t n = 0; n < N; n++) {
    const double *B = blocks + 36 * n, *v = velo + 6 * n;
    double t = 0.;
    for (int i = 0; i < 6; i++) {
      double a = 0.;
      for (int j = 0; j < 6; j++) a += B[6 * i + j] * v[j];
      t += v[i] * a;
    }
    ekin[n] = 0.5 * t;
  }
  free(blocks);
  return REF_OK;
}

int ref_plastic_state(ref_model *m, double *eps_p, double *alpha, double *dflow, uint8_t *active) {
  const int64_t E = m->E;
  if (eps_p) memcpy(eps_p, m->curr_ep, 6 * E * sizeof(double));
  if (alpha) memcpy(alpha, m->curr_al, 7 * E * sizeof(double));
  if (dflow) memcpy(dflow, m->dflow, E * sizeof(double));
  if (active) memcpy(active, m->active, E);
  return REF_OK;
}

/* sets the committed (and current) plastic strain [E][6] and hardening
 * parameters [E][7], e.g. to continue from a state another run reached */
int ref_set_plastic_state(ref_model *m, const double *eps_p, const double *alpha) {
  const int64_t E = m->E;
  if (eps_p) {
    memcpy(m->old_ep, eps_p, 6 * E * sizeof(double));
    memcpy(m->curr_ep, eps_p, 6 * E * sizeof(double));
  }
  if (alpha) {
    memcpy(m->old_al, alpha, 7 * E * sizeof(double));
    memcpy(m->curr_al, alpha, 7 * E * sizeof(double));
  }
  return REF_OK;
}

int ref_curr_strains(ref_model *m, double *strain) {
  memcpy(strain, m->curr_strain, 6 * m->E * sizeof(double));
  return REF_OK;
}
