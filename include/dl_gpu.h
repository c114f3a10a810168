/*
 * dl_gpu.h -- C ABI of the MI355X (gfx950) explicit-dynamics path for
 * geometrically exact rod lattices (dynLattice hot path).
 *
 * The ABI mirrors the reference's JIVE model action surface
 *   bool Model::takeAction(const String& action, const Properties& params,
 *                          const Properties& globdat)
 * (src/models/SpecialCosseratRodModel.h:164-166, src/models/RodContactModel.h:112-114,
 *  dispatched by src/models/LatticeModel.cpp:115-120) and the explicit
 * time-marching module's step (src/modules/ExplicitModule.cpp:259-459,
 * src/modules/LeapFrogModule.cpp:44-84).
 *
 * Conventions
 *  - Every entry point returns DL_OK (0) or a negative DL_ERR_* code; no C++
 *    exception crosses the ABI.  dl_last_error() gives a message.
 *  - Host arrays are caller-owned and copied in/out.  Device state (SoA,
 *    FP64) is owned by the context and stays resident in HBM between calls.
 *  - DOF vectors use the reference's layout: node-major, 6 DOFs per node in
 *    the order dx, dy, dz, rx, ry, rz (SpecialCosseratRodModel.cpp:131-137).
 *  - A context belongs to one HIP stream and is not thread-safe; separate
 *    contexts are independent.
 */
#ifndef DL_GPU_H
#define DL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DL_OK 0
#define DL_ERR_INPUT -1        /* IllegalInputException / JEM_PRECHECK on input */
#define DL_ERR_ZERO_MASS -2    /* ArithmeticException, ExplicitModule.cpp:452-454 */
#define DL_ERR_PLASTIC_ITER -3 /* JEM_PRECHECK2(liter < maxIter_), ElastoPlasticRodMaterial.cpp:234 */
#define DL_ERR_HIP -4          /* HIP runtime error */
#define DL_ERR_CAPACITY -5     /* contact pair buffer too small */
#define DL_ERR_STATE -6        /* call out of order (e.g. step before INIT) */

/* material models: MaterialFactory "ElasticRod" / "ElastoPlasticRod"
 * (ElasticRodMaterial.cpp:420-424, ElastoPlasticRodMaterial.cpp:388-392) */
#define DL_ELASTIC 0
#define DL_ELASTOPLASTIC 1

/* actions: jive::model::Actions names handled by the rod/contact models */
#define DL_INIT 0            /* Actions::INIT: frames, initial strain, contact blacklist */
#define DL_GET_INT_VECTOR 1  /* Actions::GET_INT_VECTOR (+= into params.int_vector) */
#define DL_GET_MATRIX2 2     /* Actions::GET_MATRIX2: block-lumped mass (6x6 per node) */
#define DL_COMMIT 3          /* Actions::COMMIT: material applyDeform + energies */
#define DL_CANCEL 4          /* Actions::CANCEL: material rejectDeform */

/* load cases (globdat LOAD_CASE) */
#define DL_LOAD_CASE_DEFAULT 0 /* inelastic material update */
#define DL_LOAD_CASE_OUTPUT 1  /* "output": elastic predictor only (SpecialCosseratRodModel.cpp:762) */

/* dl_options.flags */
#define DL_TRACK_DISSIPATION 1u /* per-step plastic dissipation history */
#define DL_FRAMES_PER_ELEMENT 2u /* the element kernel reads both end frames of
                                  * every element instead of taking an inner
                                  * element's end-b frame from its neighbour
                                  * (same values; a layout the sweep
                                  * autotuner tries) */
#define DL_JOINTS_IN_MEMORY 4u   /* every shared node goes through the slot
                                  * buffer and the joint kernel, also those
                                  * whose chains all sit in one wavefront
                                  * (same bits; a layout the sweep autotuner
                                  * tries, faster for some instance mixes) */

typedef struct dl_ctx dl_ctx;

/* Mesh: nodes, 2-node line elements, and the element group ("beam_<s+1>")
 * of every element.  The elements of one strut must be contiguous and form a
 * chain in order (element k = (p_k, p_k+1)), which is what one GMSH curve
 * meshed with Line2 elements is (GMSHInputModule.cpp:308-415). */
typedef struct {
  int64_t n_nodes;
  int64_t n_elems;
  int32_t n_struts;
  const double *coords;      /* [3*n_nodes] */
  const int32_t *elem_nodes; /* [2*n_elems] */
  const int32_t *elem_strut; /* [n_elems] */
} dl_mesh;

/* Per-strut section and elastic constants, i.e. the ElasticRodMaterial of
 * each rod model (ElasticRodMaterial.cpp:65-228). */
typedef struct {
  const double *area;          /* A */
  const double *I1;            /* area moment, material x */
  const double *I2;            /* area moment, material y */
  const double *J;             /* polar moment */
  const double *shear_factor;  /* kappa */
  const double *young;         /* E */
  const double *shear_modulus; /* G */
  const double *density;       /* rho */
  double inertia_correct;      /* ElasticRodMaterial.cpp:197-200; 1 = off */
} dl_sections;

/* Elastoplastic rod material (ElastoPlasticRodMaterial.cpp:55-147) with the
 * yield condition restricted to the power-sum family
 *   f = sum_i |w_i (s_i + ck_i q_i)|^p_i - Y (1 + ci q_0),  q = -H alpha,
 *   H = blockdiag(H_iso, diag(kin_const + kin_ratio * C)),
 * p_i = 1: the sum-abs conditions of tests/plastic/test1,2a,2b.pro;
 * p_i ~ 2: the shape of tests/plastic/test3.pro. */
typedef struct {
  int32_t model;          /* DL_ELASTIC / DL_ELASTOPLASTIC */
  const double *weights;  /* [6*n_struts] w_i */
  double yield_level;     /* Y */
  double kin_coupling[6]; /* ck_i */
  double iso_coupling;    /* ci */
  int32_t has_iso;        /* "isotropicCoefficient" given */
  double iso_modulus;     /* H_iso */
  int32_t has_kin;        /* "kinematicTensor" given (diagonal) */
  double kin_const[6];
  double kin_ratio;
  double precision;       /* default 1e-5 */
  int32_t max_iter;       /* default 20 */
  double exponent[6];     /* p_i (>= 1) */
} dl_material;

/* Rod-rod penalty contact (RodContactModel.cpp:60-98).  radius is per strut;
 * the pair threshold is r_A + r_B, equal to the reference's 2 r when all
 * radii are equal. */
typedef struct {
  int32_t enabled;
  double penalty_sts;
  double penalty_nts;
  const double *radius; /* [n_struts] */
  int64_t max_pairs;    /* step candidate-list capacity (partners summed over the
                           elements; max_pairs / n_elems per element, pairs
                           blacklisted at INIT included); 0 = 48 * n_elems;
                           overflow is DL_ERR_CAPACITY */
  double skin;          /* candidate-list skin distance of the step: 0 = half the
                           largest radius, < 0 = full broadphase every step */
} dl_contact;

typedef struct {
  int32_t so3;      /* dofs_SO3 given: consistent block mass, gyro, Lie-group update */
  int32_t device;   /* HIP device ordinal */
  void *stream;     /* hipStream_t to run on; NULL = the context creates its own */
  uint32_t flags;   /* DL_TRACK_DISSIPATION */
} dl_options;

/* Action parameters (ActionParams::INT_VECTOR / MATRIX2, LOAD_CASE). */
typedef struct {
  double *int_vector;  /* [6*n_nodes] host, accumulated with += (GET_INT_VECTOR) */
  double *matrix2;     /* [36*n_nodes] host 6x6 row-major blocks, += (GET_MATRIX2) */
  int32_t load_case;   /* DL_LOAD_CASE_* */
  int32_t with_gyro;   /* MATRIX2 passed with GET_INT_VECTOR: add gyroscopic term */
  int32_t lumped;      /* GET_MATRIX2 into a diagonal (LUMPED) builder */
  double e_pot;        /* COMMIT: potentialEnergy */
  double e_diss;       /* COMMIT: dissipatedEnergy */
} dl_params;

/* Output of dl_explicit_step: values after the last step. */
typedef struct {
  double reaction;      /* GroupOutput sum of the reaction group after the last step */
  double e_diss;        /* accumulated plastic dissipation (DL_TRACK_DISSIPATION) */
  int64_t n_contacts;   /* active contact pairs in the last step */
  int64_t n_candidates; /* broadphase candidates in the last step */
  int64_t steps_done;
  int64_t n_rebuilds;   /* contact candidate-list rebuilds during the call */
} dl_step_out;

/* Replaces: model tree construction by factories (ModelFactory "Lattice" ->
 * N x "specialCosseratRod" + "RodContact", LatticeModel.cpp:61-96;
 * SpecialCosseratRodModel.cpp:100-186) and ExplicitModule::init mass set-up
 * (ExplicitModule.cpp:58-196). */
int dl_create(const dl_mesh *mesh, const dl_sections *sec, const dl_material *mat,
              const dl_contact *contact, const dl_options *opt, dl_ctx **out);
int dl_destroy(dl_ctx *ctx);

/* Replaces: Model::takeAction(action, params, globdat) for the rod lattice
 * (SpecialCosseratRodModel.cpp:196-400; RodContactModel.cpp:107-240;
 * LatticeModel.cpp:110-192). */
int dl_take_action(dl_ctx *ctx, int32_t action, dl_params *params);

/* Replaces: StateVector::get/store(disp|velo, STATE0|STATE1, dofs, globdat)
 * (used at SpecialCosseratRodModel.cpp:349, LeapFrogModule.cpp:62-64). */
int dl_set_state(dl_ctx *ctx, const double *disp, const double *velo, int64_t ndof);
int dl_get_state(dl_ctx *ctx, double *disp, double *velo, int64_t ndof);

/* Replaces: Constraints from GET_CONSTRAINTS (DirichletModel.cpp:326-366,
 * ImpactModel.cpp:74-104) and f_ext from GET_EXT_VECTOR
 * (NeumannModel.cpp:291-352): cmask[6N] 1 = constrained, cval[6N] the
 * prescribed acceleration, fext[6N] constant external force. */
int dl_set_bcs(dl_ctx *ctx, const uint8_t *cmask, const double *cval, const double *fext,
               int64_t ndof);

/* Replaces: GroupOutputModule's resp sum over a node group's DOF
 * (GroupOutputModule.cpp:100-116,162), recorded once per step: after the
 * step, with LOAD_CASE "output", the internal force at the new state -- the
 * rod elements' elastic predictor from the committed plastic state
 * (SpecialCosseratRodModel.cpp:762) plus the contact forces of a contact
 * search at that state (RodContactModel.cpp:153-166) -- summed over the
 * group's nodes on dof.  The sum of step k is evaluated right after the
 * contact pass of step k + 1 (the same state); the last step's by a closing
 * contact pass and evaluation at the end of the call. */
int dl_set_reaction_group(dl_ctx *ctx, const int32_t *nodes, int64_t n, int32_t dof);

/* Replaces: ImpactModel (ImpactBC, ImpactModel.cpp:74-104): every node of
 * group g gets on dof dofs[g] the prescribed acceleration -R_g / weights[g],
 * R_g the group's internal-force sum on that dof at the step's start state
 * (the "resp" the reference's GroupOutput computed after the previous step;
 * that is the output load case, which for an element yielding in the step
 * differs from the step's own evaluation used here -- identical for the
 * elastic material).  Until the first step (LeapFrog) or accepted attempt (MilneDevice) after
 * INIT there is no such output and the dof stays free, as in the reference.
 * nodes[group_ptr[g] .. group_ptr[g+1]); a node belongs to one group at
 * most, and the dof must not be constrained by dl_set_bcs.  n_groups = 0
 * removes them.  Not with per-node time steps or the adjoint. */
int dl_set_impact_groups(dl_ctx *ctx, const int32_t *nodes, const int32_t *group_ptr,
                         int32_t n_groups, const int32_t *dofs, const double *weights);

/* Several reaction groups at once (e.g. one per instance of a batched design
 * sweep): nodes[group_ptr[g] .. group_ptr[g+1]) form group g; every step
 * records one sum per group (dl_get_group_history).  Generalises the
 * GroupOutputModule resp sum (GroupOutputModule.cpp:121-169) to n_groups. */
int dl_set_reaction_groups(dl_ctx *ctx, const int32_t *nodes, const int32_t *group_ptr,
                           int32_t n_groups, int32_t dof);

/* Per-node time step for independent instances packed into one context (a
 * batched sweep: every instance keeps its own stable dt, ExplicitModule's
 * deltaTime per model).  NULL restores the scalar dt of dl_explicit_step. */
int dl_set_node_dt(dl_ctx *ctx, const double *dt, int64_t n_nodes);

/* Replay the step sequence from a captured hipGraph of steps_per_graph steps
 * (0 = direct launches).  The graph is re-captured when dt or any boundary /
 * group data changes. */
int dl_set_graph(dl_ctx *ctx, int32_t steps_per_graph);

/* Replaces: ControlModule -> ExplicitModule::advance / LeapFrogModule::solve /
 * ExplicitModule::commit, nsteps times with fixed dt (maxDtime = minDtime =
 * deltaTime), with the intended store of u, v (LeapFrogModule.cpp:44-84,
 * ExplicitModule.cpp:259-419).  Runs asynchronously on the context's stream;
 * out (may be NULL) is filled after a stream synchronisation.  Errors the
 * kernels raise (DL_ERR_PLASTIC_ITER, DL_ERR_CAPACITY) are kept in a device
 * flag and returned by the first call that synchronises and reads it: this
 * call when out is not NULL, else a later dl_explicit_step with out,
 * dl_milne_step, dl_sync or an action.  After DL_ERR_CAPACITY the candidate
 * lists stay marked stale, so every later step rebuilds them (and raises the
 * error again while they still overflow). */
int dl_explicit_step(dl_ctx *ctx, int64_t nsteps, double dt, dl_step_out *out);

/* Allocation-free variant of dl_explicit_step, for callers that cannot
 * allocate device memory (code that runs under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same steps and the same histories as dl_explicit_step, with
 * direct launches (no internal graph); the step counter and the history
 * rows live in a fixed per-GPU scratch array of DL_NOALLOC_SCRATCH_DOUBLES
 * doubles (2 MiB, held by every GPU the library runs on until the process
 * ends) instead of the context's history buffers, so no hipMalloc, hipFree
 * or allocating library call happens.  nsteps * (n_groups + 3) + 1 must not
 * exceed DL_NOALLOC_SCRATCH_DOUBLES (else DL_ERR_INPUT).  The rows are copied
 * out before returning (after a stream synchronisation): reaction
 * [nsteps][max(n_groups, 1)], e_diss [nsteps], n_contacts [nsteps], each may
 * be NULL.  dl_get_history keeps describing the last dl_explicit_step call.
 * All contexts on one GPU share its scratch array: calls that overlap on one
 * GPU (other threads, other streams) race and must be serialised by the
 * caller. */
#define DL_NOALLOC_SCRATCH_DOUBLES (1 << 18)
int dl_explicit_step_noalloc(dl_ctx *ctx, int64_t nsteps, double dt, double *reaction,
                             double *e_diss, int64_t *n_contacts);

/* Per-step history of the last dl_explicit_step call: GroupOutput reaction
 * after the step, dissipation after the step, active contact pairs of the
 * step's force evaluation (its start state); n = number of steps recorded. */
int dl_get_history(dl_ctx *ctx, double *reaction, double *e_diss, int64_t *n_contacts,
                   int64_t cap, int64_t *n);
/* All reaction groups of the last dl_explicit_step call, [steps][n_groups]. */
int dl_get_group_history(dl_ctx *ctx, double *reaction, int64_t cap_steps, int64_t *n);

/* Replaces: RodContactModel contactsA_/contactsB_ (RodContactModel.cpp:732-733):
 * active pairs of the last force evaluation, sorted by (a, b). */
int dl_contact_pairs(dl_ctx *ctx, int32_t *a, int32_t *b, int64_t cap, int64_t *n);
/* Broadphase candidates of the last force evaluation, sorted by (a, b). */
int dl_contact_candidates(dl_ctx *ctx, int32_t *a, int32_t *b, int64_t cap, int64_t *n);
/* The step's candidate lists: partners listed, summed over the elements
 * (compare with dl_contact.max_pairs, the capacity), and list builds since
 * the context was created.  Either pointer may be NULL; zeros without
 * lists. */
int dl_contact_list_stats(dl_ctx *ctx, int64_t *entries, int64_t *builds);

/* Replaces: the plastic-strain table of ElastoPlasticRodMaterial::getTable
 * (ElastoPlasticRodMaterial.cpp:318-340): per element the plastic-correction
 * flag of the last material update (mask[n_elems]), plastic strain [6E] and
 * hardening parameters [7E] (iso, kin...) of the committed state. */
int dl_plastic_active(dl_ctx *ctx, uint8_t *mask, int64_t n);
int dl_plastic_state(dl_ctx *ctx, double *eps_p, double *alpha, int64_t n_elems);
/* Sets the committed plastic state (the layout of dl_plastic_state), e.g. to
 * continue a run from a saved state together with dl_set_state; the
 * counterpart of the material's committed history (ElastoPlasticRodMaterial
 * oldPlastStrains_ / oldHardParams_).  Either array may be NULL. */
int dl_set_plastic_state(dl_ctx *ctx, const double *eps_p, const double *alpha, int64_t n_elems);

/* Device timing of the kernels of the last dl_explicit_step (or
 * dl_milne_step) call, measured with HIP events on the context's stream:
 * total milliseconds of the named phase ("step", "contact" -- of which
 * "refresh" is the candidate-list rebuild chain --, "react" (the reaction
 * groups' entries), "strut", "joint"; for MilneDevice both force
 * evaluations of an attempt, and "node": predictor, corrector, step control
 * and commit) and the number of steps (attempts). */
int dl_phase_time(dl_ctx *ctx, const char *phase, double *ms, int64_t *count);
int dl_set_timing(dl_ctx *ctx, int32_t enabled);

/* Replaces the per-strut section data (and, with contact, the per-strut
 * contact radii; either may be NULL) of an existing context: stiffness,
 * hardening and lumped mass are rebuilt as dl_create builds them, i.e. the
 * ElasticRodMaterial configuration (ElasticRodMaterial.cpp:65-228) and
 * ExplicitModule's mass set-up (ExplicitModule.cpp:58-196).  The contact
 * blacklist keeps the radii of the last INIT. */
int dl_set_sections(dl_ctx *ctx, const dl_sections *sec, const double *contact_radius);

/* Replaces the yield weights of the elastoplastic material ([n_struts][6],
 * the table dl_material.weights gave at create: the per-rod-model yield
 * function's coefficients, ElastoPlasticRodMaterial.cpp:131-141), e.g. after
 * a change of the strut radii or yield stresses. */
int dl_set_yield_weights(dl_ctx *ctx, const double *weights);

/* Gradients of the per-strut section data (each array [n_struts] or NULL to
 * skip it) and, with the elastoplastic material, of the yield weights
 * ([n_struts][6] like dl_material.weights, or NULL; zero for the elastic
 * material). */
typedef struct {
  double *area, *I1, *I2, *J, *shear_factor, *young, *shear_modulus, *density;
  double *contact_radius;
  double *yield_weight;
} dl_section_grads;

/* Adjoint design gradients (no counterpart in the reference, which has no
 * sensitivities).  dl_adjoint_forward runs nsteps explicit steps like
 * dl_explicit_step (same histories; the GroupOutput reactions
 * reaction[nsteps][n_groups] returned when not NULL) and keeps a checkpoint
 * of the state every checkpoint_every steps.  dl_adjoint_backward then takes the adjoint of an objective with
 * respect to that reaction history, reaction_bar[nsteps][n_groups] (e.g.
 * dt * v_plate for the absorbed energy, a one-hot row for the peak), and
 * returns the gradient of the objective with respect to the section data by
 * a reverse sweep: every segment is recomputed from its checkpoint, then
 * reversed through hand-written adjoint kernels of the node update, the
 * element internal force and the contact pair forces; the last row's weight
 * enters through the internal force's adjoint at the final state.
 * Elastoplastic material (ElastoPlasticRodMaterial.cpp:198-269): the
 * committed plastic state is checkpointed and taped with the node state,
 * the cutting-plane return map is reversed through every correction it
 * took (the derivative of the algorithm as executed: exact wherever the
 * number of corrections does not change), and the reaction rows are
 * reversed through the output load case (the elastic predictor) that
 * GroupOutput records.
 * SO(3) rotation update only.  The forward run's final state (and plastic
 * state) is restored afterwards; backward may be called again with another
 * reaction_bar until the state, boundary data or sections change. */
int dl_adjoint_forward(dl_ctx *ctx, int64_t nsteps, double dt, int32_t checkpoint_every,
                       double *reaction);
int dl_adjoint_backward(dl_ctx *ctx, const double *reaction_bar, dl_section_grads *grads);

/* MilneDevice integrator settings: the solver module's configuration
 * (defaults ExplicitModule.cpp:37-46, lengthScale :76, the rest :128-149;
 * MilneDeviceModule.cpp:28-31 sets order 1).  dt_min and dt_max default to
 * dt / 100 and dt * 100 there. */
typedef struct {
  double dt;        /* step size of the first attempt (deltaTime) */
  double dt_min;    /* minDtime: attempts at or below it are always accepted */
  double dt_max;    /* maxDtime */
  double precision; /* precision: error bound for acceptance */
  double safety;    /* stepSaftey, 0.9 */
  double incr;      /* increaseFactor, 1.2 */
  double decr;      /* decreaseFactor, 0.8 */
  double len_scale; /* lengthScale: translational dofs are divided by it in the error, 1e-3 */
} dl_milne_opts;

/* Output of dl_milne_step, filled after a stream synchronisation. */
typedef struct {
  int64_t attempts; /* attempts made */
  int64_t accepted; /* of which accepted */
  double time;      /* sum of the accepted step sizes */
  double dt_next;   /* step size for the next attempt (pass it as opts.dt to continue) */
} dl_milne_out;

/* Replaces: ControlModule -> ExplicitModule::advance / MilneDeviceModule::solve /
 * ExplicitModule::commit|cancel with adaptive step size, n_attempts times
 * (MilneDeviceModule.cpp:41-112, ExplicitModule.cpp:259-335).  Each attempt:
 * forces at the start state and the predictor, forces at the predicted state
 * and the corrector (mass of the start state), the error estimate, then
 * accept (the material commits its predicted-state update and the state
 * moves to the corrected one) or reject, and the next step size -- all on the
 * device, without a host round trip (dl_set_graph replays attempts from a
 * captured graph).  Histories as dl_explicit_step, one row per attempt: the
 * GroupOutput reaction at the state the attempt leaves (the accepted state,
 * or the start state a rejected attempt restores) and the contact pairs at
 * the attempt's start state.  attempts (may
 * be NULL): host [n_attempts][4] = dt, error, accepted (0/1), next dt.
 * Elastic and elastoplastic material, contact, both rotation updates; not
 * with per-node time steps (dl_set_node_dt). */
int dl_milne_step(dl_ctx *ctx, int64_t n_attempts, const dl_milne_opts *opts, double *attempts,
                  dl_milne_out *out);

/* Time programs of the boundary data.  Every prescribed acceleration
 * (constrained dof) and external force (free dof) of dl_set_bcs may follow a
 * program: its value becomes (the dl_set_bcs value) x (the program's scale).
 * Replaces the time dependence the reference's boundary models give the
 * solver through GET_CONSTRAINTS / GET_EXT_VECTOR after
 * ExplicitModule::advance (ExplicitModule.cpp:259-274, 397-419):
 *   DL_PROG_STEP, scale = a + b (k + c): NeumannModel initLoad/loadIncr
 *     (a = initLoad, b = loadIncr, c = 1; NeumannModel.cpp:118-147,347-368)
 *     and DirichletModel initDisp/dispIncr (c = 0; DirichletModel.cpp:163-172,
 *     290-306,400-406);
 *   DL_PROG_EXPR, scale = f(t, i): JIVE's LoadScale scaleFunc (e.g.
 *     tests/transient/test1.pro, test3.pro), f in JEM function syntax:
 *     numbers, t, i, PI, + - * / % ^, comparisons, && ||, !, if(c,a,b),
 *     sin cos tan asin acos atan exp log sqrt abs floor ceil min max pow;
 *   DL_PROG_CONST, scale = a.
 * t is the time at the end of the step being taken (clock time + dt) and
 * k (i) the number of steps taken on the clock; both are device-resident:
 * every dl_explicit_step step and every accepted dl_milne_step attempt
 * advances them.  Scales are evaluated on the device before every step /
 * attempt (no host round trip; graph replay included).  dof_prog[6N]: the
 * program of each dof or -1.  n_progs = 0 removes all programs.  Not with
 * per-node time steps or the adjoint. */
#define DL_PROG_CONST 0
#define DL_PROG_STEP 1
#define DL_PROG_EXPR 2
#define DL_MAX_PROGRAMS 64
typedef struct {
  int32_t kind;
  double a, b, c;
  const char *expr; /* DL_PROG_EXPR */
} dl_program;
int dl_set_programs(dl_ctx *ctx, const dl_program *progs, int32_t n_progs, const int32_t *dof_prog,
                    int64_t ndof);
/* The programs' clock: time t and steps taken. */
int dl_set_time(dl_ctx *ctx, double t, int64_t step);

/* updateWhen of the explicit solver modules (ExplicitModule.cpp:119-126,
 * 256-266): the block-lumped mass (and with it the rotational inertia of the
 * solve and of the gyroscopic term) is re-assembled at the current rotations
 * only at the start of the steps (MilneDevice: attempts) where cond(t, i)
 * holds on the context's clock (dl_set_time; it moves by every step, and by
 * every accepted attempt), and at the first step after INIT or
 * dl_set_state; in between the last assembled mass is used.  cond is a
 * program expression of t and i (dl_set_programs' syntax); NULL or "" gives
 * the default, an update at every step.  GET_MATRIX2 still assembles at the
 * current state (the reference's action does).  A schedule and per-node time
 * steps (dl_set_node_dt) exclude each other: DL_ERR_STATE. */
int dl_set_mass_update(dl_ctx *ctx, const char *cond);
int dl_get_time(dl_ctx *ctx, double *t, int64_t *step);
/* Evaluates a program expression on the host with the compiler and
 * evaluator the device uses (to check an expression before a run). */
int dl_program_eval(const char *expr, double t, double i, double *value);

/* Replaces: MassModel's GET_MATRIX2 contribution (MassModel.cpp:92-108):
 * per node, a point mass added to the three translational dofs and a
 * rotational inertia added to the three rotational dofs (either array may
 * be NULL; NULL for both removes them).  The reference adds a group's
 * totalMass / group size to each listed dof; a group listing all
 * translational and/or all rotational dofs maps onto these arrays exactly
 * (an isotropic addition commutes with the node rotation). */
int dl_set_point_masses(dl_ctx *ctx, const double *mass, const double *inertia, int64_t n_nodes);

/* Replaces: Actions::GET_TABLE of the rod models (SpecialCosseratRodModel.cpp:
 * 221-262, 426-531) and LatticeModel's kinetic energy and mass tables
 * (LatticeModel.cpp:195-275), evaluated at the current state:
 *   "strain", "mat_strain", "stress", "mat_stress"  [n_elems][6]
 *       (gamma_x,y,z, kappa_x,y,z / n_x,y,z, m_x,y,z; spatial = PI x material;
 *       stress of the elastoplastic material: the return map from the
 *       committed state, which stays unchanged);
 *   "potentialEnergy", "kineticEnergy", "mass"  [n_nodes][1]
 *       (element potential energy w 1/2 eps_e . sigma shared equally by the
 *       element's two nodes; 1/2 v . M v with the block-lumped mass; the
 *       translational mass).
 * n_rows and n_cols must match the table. */
int dl_get_table(dl_ctx *ctx, const char *name, double *out, int64_t n_rows, int32_t n_cols);
/* Totals at the current state: kinetic and potential energy as above, the
 * dissipated energy accumulated so far (DL_TRACK_DISSIPATION steps and
 * COMMIT actions) and the total mass (LatticeModel.cpp:160-190 sets
 * kineticEnergy and mass, SpecialCosseratRodModel.cpp:367-388
 * potentialEnergy and dissipatedEnergy).  Any pointer may be NULL. */
int dl_get_energies(dl_ctx *ctx, double *e_kin, double *e_pot, double *e_diss, double *mass);

/* Wait for the context's stream; returns (and clears) a pending device
 * error of the kernels run since the last check. */
int dl_sync(dl_ctx *ctx);

/* Message of the last failed call on this thread. */
const char *dl_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
