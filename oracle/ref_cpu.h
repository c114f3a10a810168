/*
 * ref_cpu.h -- CPU restatement ("the oracle") of dynLattice's explicit-dynamics
 * hot path: Simo-Reissner rod element, elastic / elastoplastic rod material,
 * block-lumped mass, gyroscopic term, LeapFrog step (with store), rod-rod
 * penalty contact and the commit energies.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is used by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg as the timed CPU baseline.
 *
 * Every routine follows the reference's operation order (matrix route, the
 * same repeated helper calls); the reference file:line it restates is given
 * at each definition in ref_cpu.c.
 */
#ifndef DL_REF_CPU_H
#define DL_REF_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* material models (MaterialFactory names "ElasticRod" / "ElastoPlasticRod") */
#define REF_ELASTIC 0
#define REF_ELASTOPLASTIC 1

typedef struct ref_model ref_model;

typedef struct {
  int64_t n_nodes;
  int64_t n_elems;
  int32_t n_struts;
  const double *coords;      /* [3*n_nodes] node-major x,y,z */
  const int32_t *elem_nodes; /* [2*n_elems] */
  const int32_t *elem_strut; /* [n_elems] rod group ("beam_<s+1>") of each element */
} ref_mesh;

/* Per-strut section + elastic constants (ElasticRodMaterial::configure). */
typedef struct {
  const double *area;          /* [n_struts] A   */
  const double *I1;            /* [n_struts] area moment about material x */
  const double *I2;            /* [n_struts] area moment about material y */
  const double *J;             /* [n_struts] polar moment */
  const double *shear_factor;  /* [n_struts] kappa */
  const double *young;         /* [n_struts] E */
  const double *shear_modulus; /* [n_struts] G */
  const double *density;       /* [n_struts] rho */
  double inertia_correct;      /* multiplies rotational material mass (ElasticRodMaterial.cpp:197-200); 1 = off */
} ref_sections;

/* Elastoplastic parameters: the power-sum yield family
 *   f = sum_i |w_i (s_i + ck_i q_i)|^p_i - Y (1 + ci q_0),   q = -H alpha
 * p_i = 1 is the sum-abs shape of tests/plastic/test1.pro:22-28,
 * test2a.pro:22, test2b.pro:22; p_i ~ 2 the shape of test3.pro:22-28. */
typedef struct {
  int32_t model;           /* REF_ELASTIC / REF_ELASTOPLASTIC */
  const double *weights;   /* [6*n_struts] w_i per strut */
  double yield_level;      /* Y */
  double kin_coupling[6];  /* ck_i */
  double iso_coupling;     /* ci */
  int32_t has_iso;         /* isotropicCoefficient given */
  double iso_modulus;      /* H_iso */
  int32_t has_kin;         /* kinematicTensor given (diagonal) */
  double kin_const[6];     /* H_kin,i = kin_const[i] + kin_ratio * C_ii */
  double kin_ratio;
  double precision;        /* 1e-5 (ElastoPlasticRodMaterial.cpp:43) */
  int32_t max_iter;        /* 20 (ElastoPlasticRodMaterial.cpp:42) */
  double exponent[6];      /* p_i */
} ref_plastic;

typedef struct {
  int32_t enabled;
  double penalty_sts;
  double penalty_nts;
  const double *radius; /* [n_struts] contact radius per strut (reference: one scalar) */
} ref_contact_opts;

typedef struct {
  int32_t so3;        /* dofs_SO3 set -> consistent block mass, gyro, Lie-group update */
} ref_options;

/* error codes */
#define REF_OK 0
#define REF_ERR_INPUT -1
#define REF_ERR_ZERO_MASS -2
#define REF_ERR_PLASTIC_ITER -3

ref_model *ref_create(const ref_mesh *mesh, const ref_sections *sec,
                      const ref_plastic *pl, const ref_contact_opts *co,
                      const ref_options *opt, const double *disp0);
void ref_destroy(ref_model *m);
int ref_last_error(const ref_model *m);

/* INIT: initRotation_, initStrain_ per strut; contact blacklist at disp. */
int ref_init(ref_model *m, const double *disp);

/* GET_INT_VECTOR: fint += internal force (elements [+ gyro] [+ contact]).
 * inelastic: load case is not "output".  velo may be NULL when with_gyro==0. */
int ref_get_int_vector(ref_model *m, const double *disp, const double *velo,
                       double *fint, int inelastic, int with_gyro,
                       int with_contact);

/* element-only internal force of one element, f[12] = (node a 6, node b 6) */
int ref_elem_fint(ref_model *m, int64_t ielem, const double *disp, double *f,
                  int inelastic);

/* GET_MATRIX2: 6x6 mass block per node (row-major), blocks[36*n_nodes].
 * Diagonal-only when lumped != 0 (DiagMatrixObject path). */
int ref_mass_blocks(ref_model *m, const double *disp, double *blocks, int lumped);

/* Constraints: cmask[6N] (1 = constrained dof), cval[6N] prescribed
 * acceleration; fext[6N] external force (constant). */
int ref_set_bcs(ref_model *m, const uint8_t *cmask, const double *cval,
                const double *fext);

/* GroupOutput's reaction (GroupOutputModule.cpp:100-116,162) at the state
 * disp: the sum over group nodes of fint[6*node+dof] with the output load
 * case (elastic predictor from the committed plastic state, contact forces
 * of a fresh search at disp). */
int ref_output_reaction(ref_model *m, const double *disp, const int32_t *group, int64_t group_n,
                        int32_t dof, double *r);

/* LeapFrog step with store (ExplicitModule + LeapFrogModule), nsteps times at
 * fixed dt.  disp/velo updated in place.  Reaction group: after each step,
 * GroupOutput's sum (ref_output_reaction at the new state) is written to
 * react[step] (may be NULL).  If
 * commit_energy != 0, energies[3*step + (0,1,2)] = (E_pot, E_diss, E_kin)
 * after each step's COMMIT. n_contacts[step] = active contact pairs. */
int ref_leapfrog(ref_model *m, double *disp, double *velo, double dt,
                 int64_t nsteps, const int32_t *group, int64_t group_n,
                 int32_t group_dof, double *react, int commit_energy,
                 double *energies, double *unused, int64_t *n_contacts);

/* 1: broadphase by uniform grid (same candidate set, different order);
 * 0: faithful O(R^2) rod/element AABB loops (default). */
void ref_set_grid(ref_model *m, int use_grid);

/* COMMIT energies (SpecialCosseratRodModel.cpp:367-388) for the current
 * committed state: applyDeform, then potential and dissipated energy passes. */
int ref_commit(ref_model *m, const double *disp, double *e_pot, double *e_diss);

/* Boundary programs: every constrained dof's prescribed acceleration and
 * every free dof's external force with dof_prog[i] = q >= 0 becomes (its
 * ref_set_bcs value) x scale_q, evaluated before every step / attempt at the
 * step's end time (clock + dt) after ExplicitModule::advance
 * (ExplicitModule.cpp:259-274): CONST a; STEP as NeumannModel / Dirichlet
 * iterate it (scale_k = a + b (k + c) by repeated addition); EXPR a LoadScale
 * scaleFunc of t and i.  The clock (t, k) moves with each taken step. */
#define REF_MAX_PROG 64
#define REF_PROG_CONST 0
#define REF_PROG_STEP 1
#define REF_PROG_EXPR 2
typedef struct {
  int kind;
  double a, b, c;
  char *expr;
  double acc;
} ref_prog;
int ref_set_programs(ref_model *m, int n, const int32_t *kind, const double *abc,
                     const char *const *exprs, const int32_t *dof_prog);
int ref_set_time(ref_model *m, double t, int64_t k);
/* updateWhen of the explicit modules: cond(t, i) or NULL (every step) */
int ref_set_mass_update(ref_model *m, const char *cond);
int ref_mass_in_use(ref_model *m, double *blocks);
int ref_get_time(const ref_model *m, double *t, int64_t *k);
int ref_expr_eval(const char *expr, double t, double i, double *value);
/* GET_TABLE at disp: kind 0 strain, 1 mat_strain, 2 stress, 3 mat_stress
 * into out [E][6]; element potential energy into epot [E] (either NULL) */
int ref_elem_table(ref_model *m, const double *disp, int kind, double *out, double *epot);
/* per-node kinetic energy 1/2 v . M v (mass blocks at disp) */
int ref_kinetic_energy(ref_model *m, const double *disp, const double *velo, double *ekin);
/* ImpactBC groups: nodes[ptr[g] .. ptr[g+1]) get on dofs[g] the prescribed
 * acceleration -R_g / weights[g], R_g their internal-force sum on that dof
 * at the step's start state; free until the first taken step after INIT */
int ref_set_impact_groups(ref_model *m, int n, const int32_t *nodes, const int32_t *ptr,
                          const int32_t *dofs, const double *weights);
/* MassModel point masses: per node, added to the translational and the
 * rotational diagonal of the mass blocks (either may be NULL) */
int ref_set_point_masses(ref_model *m, const double *mass, const double *inertia);

/* MilneDevice predictor-corrector with the adaptive step of
 * ExplicitModule::commit (MilneDeviceModule.cpp:41-112,
 * ExplicitModule.cpp:290-335); see ref_cpu.c */
typedef struct {
  double dt;         /* first step size (deltaTime) */
  double dt_min, dt_max;
  double precision;  /* prec_ */
  double safety;     /* stepSaftey, 0.9 */
  double incr, decr; /* increaseFactor 1.2, decreaseFactor 0.8 */
  double len_scale;  /* lengthScale, 1e-3 */
} ref_milne_opts;
int ref_milne(ref_model *m, double *disp, double *velo, const ref_milne_opts *o,
              int64_t n_attempts, const int32_t *group, int64_t group_n, int32_t group_dof,
              double *out);

/* Contact broadphase, faithful O(R^2) (RodContactModel.cpp:245-446).
 * Writes candidate pairs (A,B), returns the count (or, if more than cap,
 * the count without writing beyond cap). */
int64_t ref_contact_candidates(ref_model *m, const double *disp, int32_t *a,
                               int32_t *b, int64_t cap);
/* Active contact pairs of the last force evaluation (contactsA_/B_). */
int64_t ref_contact_active(ref_model *m, int32_t *a, int32_t *b, int64_t cap);
/* Contact force only (fint += contact), returns the active pair count. */
int64_t ref_contact_force(ref_model *m, const double *disp, double *fint);
int64_t ref_blacklist(ref_model *m, int32_t *a, int32_t *b, int64_t cap);

/* Plastic state read-out: per element eps_p[6], alpha[7] (iso, kin...),
 * delta_flow, and the 'plastic active' flag of the last material call. */
int ref_set_plastic_state(ref_model *m, const double *eps_p, const double *alpha);
int ref_plastic_state(ref_model *m, double *eps_p, double *alpha,
                      double *dflow, uint8_t *active);
/* Material strains of the last call (currStrains_), [6*n_elems]. */
int ref_curr_strains(ref_model *m, double *strain);

/* Per-(element,end) reference frame LambdaN (row-major 3x3), [18*n_elems],
 * and initial material strain matStrain0_, [6*n_elems]. */
int ref_frames(ref_model *m, double *lambda_n, double *strain0);

/* SO(3) helpers exposed for tests (helpers.cpp). */
void ref_expvec(const double psi[3], double R[9]);
void ref_logmat(const double R[9], double rv[3]);
void ref_expvecp(const double psi[3], const double psiP[3], double out[9]);

#ifdef __cplusplus
}
#endif
#endif
