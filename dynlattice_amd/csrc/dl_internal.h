// dl_internal.h -- device data layout and launcher prototypes shared by the
// host API (dl_api.cpp) and the kernels (dl_kernels.hip, dl_contact.hip).
//
// HBM layout: structure-of-arrays, FP64, component-major ([c][n]) so that
// consecutive threads touch consecutive addresses.
#ifndef DL_INTERNAL_H
#define DL_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dl_device.h"
#include "dl_program.h"

// Nodes, elements, struts and the shared-node gather map.
struct DevData {
  int64_t N, E;
  int32_t R;
  // nodes ([c][N])
  const double *X0;  // 3
  double *u;         // 3
  double *rot;       // so3: quaternion 4 (w,x,y,z); additive: rotation vector 3
  double *v;         // 6
  const double *mt;  // translational lumped mass
  const double *ndt; // per-node time step (independent batched instances) or null
  const double *J0;  // 6: rotational inertia at theta = 0 (xx yy zz xy xz yz)
  // updateWhen: the spatial inertia R J0 R^T of the last mass update, packed
  // like J0, which the node updates use instead of turning J0 by the
  // current rotation (null: the mass follows the rotation at every step)
  const double *Js;
  const double *gm;  // gyro multiplicity (number of struts holding the node)
  const int32_t *bc; // -1 or record index
  // boundary records
  const uint32_t *bc_mask; // bit i: dof i constrained
  const double *bc_a;      // [6][nbc] prescribed acceleration
  const double *bc_f;      // [6][nbc] external force
  const int32_t *bc_rslot; // reaction-group node slot or -1 (the adjoint's objective weight)
  const int32_t *bc_imp;   // impact group of the record or -1 (null: no impact groups)
  int32_t nbc;
  int32_t react_dof;
  // struts
  const double *C;      // [6][R] material stiffness diag
  // element -> lane packing: chains of <= 64 elements in consecutive lanes
  int64_t NL;             // lanes (multiple of the block size)
  const int32_t *lane_e;  // [NL] element of the lane or -1
  int32_t lane_ident;     // lane L holds element L for L < E (no padding): lane_e not read
  const int32_t *en0;     // [E] node a
  const int32_t *en1;     // [E] node b
  const int32_t *es;      // [E] strut
  uint8_t *eflag;         // [E] bit 0: first element of its chain, bit 1: last,
                          //     bit 2: end frames -I (strut along -e3; elem_mirror),
                          //     bit 3: an end at a reaction-group node (rslot_e),
                          //     bit 4: committed plastic state possibly nonzero
                          //     (set by every writer of pst; where it is clear the
                          //     record is zero and is not read)
  const int32_t *sa;      // [E] slot of node a if shared, else -1
  const int32_t *sb;      // [E] slot of node b (chain end) if shared, else -1
  // elements ([c][E])
  const double *qN;     // 8: LambdaN at end a (4), end b (4)
  // frames at end b in k_strut when qN_share: an inner element's end-b
  // frame is the next lane's end-a frame (the strut's frame at their common
  // node); only chain ends read theirs, from a compact array
  int32_t qN_share;
  int32_t nqe;             // chain ends
  const double *qNe;       // [4][chain ends] end-b frame of each chain's last element
  const int32_t *qNe_wave; // [NL / 64] chain ends in the wavefronts before
  double *eps0;         // 6
  // plastic
  const double *pw;     // [6][R] yield weights
  const double *Hk;     // [6][R] kinematic hardening diag
  // element-major records: an element's plastic state is one 96-byte
  // record, read only where eflag bit 4 is set and written only where the
  // return map corrected it, so yielding elements scattered over the mesh
  // touch whole lines instead of a part of one line per component
  double *pst;          // [E][12] committed plastic strain (6), kinematic
                        //         hardening parameters alpha_1..6 (6)
  double *aiso;         // [E] committed isotropic hardening parameter alpha_0
  double *pst_cur;      // [E][12] trial state (action path, Milne)
  double *aiso_cur;     // [E]
  double *dflow;        // [E]
  uint8_t *active;      // [E]
  double *eo;           // [E][6] committed total strain (dissipation)
  double *ecur;         // [E][6] last material strain (action path)
  // shared nodes
  int32_t NJ;
  const int32_t *jnode; // [NJ]
  const int32_t *jptr;  // [NJ+1]
  const int32_t *jslot; // [jptr[NJ]]
  int32_t nslot;
  double *sbuf;         // [6][nslot]
  // outputs
  double *fout;         // [6][N] (action path)
  // contact: per-element translational contact force on (node a, node b)
  const unsigned long long *cbits; // [(E+63)/64] bit e: element e has a contact force row
  const double *fcs;    // [6][E] (a: 0..2, b: 3..5), row e valid where its bit is set
  // GroupOutput reaction: entries (element << 1 | end) at the reaction
  // groups' nodes, group by group; rbuf holds each entry's share of the
  // group dof of the output-load-case internal force
  const int32_t *rent;  // [nrent]
  int64_t nrent;
  double *rbuf;         // [nrent]
  // entry of the element's end a ([0][E]) and end b ([1][E]) or -1: the
  // step's k_strut writes the rows of the elements flagged in eflag bit 3
  const int32_t *rslot_e;
  double *dbuf;         // [R] per-strut dissipation increment
  int32_t *err;         // device error flag
  // adjoint tape row of the current step ([19][N]: u 3, rot 4, v 6, f 6) or null
  double *tape;
  // contact candidate lists: a node whose displacement has moved by more
  // than ctrig (any component) since the lists' build sets *cgate
  const double *uref;
  int32_t *cgate;
  double ctrig;
};

struct StepParams {
  double dt;
  int32_t inelastic;
  int32_t track_diss;
};

// strut / joint kernel modes: the explicit step (node update), force only
// (nodal force to fout), and the MilneDevice predictor / corrector fused
// into the force evaluation
enum { MODE_STEP = 0, MODE_FORCE = 1, MODE_MPRED = 2, MODE_MCORR = 3 };

// Reverse sweep of the explicit step (SO(3) mode, elastic or elastoplastic
// material).  Adjoints ("bar") of the state after the step being reversed
// are turned into those of the state before it; parameter adjoints
// accumulate.
struct AdjDev {
  const double *tape;   // [19][N] start state of the step: u 3, rot 4, v 6, f 6
  double *ubar;         // [3][N]
  double *qbar;         // [4][N]
  double *vbar;         // [6][N]
  double *fbar;         // [6][N] adjoint of the nodal internal force of the step
  double *ebuf;         // [14][E] per element end: ubar_a 3, qbar_a 4, ubar_b 3, qbar_b 4
  double *Cbar;         // [6][E] stiffness adjoint, summed over steps
  double *mtbar;        // [N]
  double *J0bar;        // [9][N] full 3x3
  double *crbar;        // [E] contact-distance adjoint, summed over steps
  const int32_t *nptr;  // [N+1] node -> element ends
  const int32_t *nlist; // (element << 1) | end
  const int32_t *slot_group; // reaction slot -> group
  const double *rbar;   // [groups] objective adjoint of this step's reactions
  // elastoplastic material (null for the elastic one)
  const double *ptape;  // committed plastic state at the step's start: the
                        // [E][12] records of DevData::pst, then [E] alpha_0
  double *pbar;         // [13][E] adjoint of the committed plastic state
                        // (eps_p 6, alpha 7): of the step's end state on
                        // entry to the element kernel, of its start on exit
  double *Wbar;         // [6][E] yield-weight adjoint, summed over steps
  double *rnode;        // [N] objective weight on the node's reaction dof
  double *fstep_r;      // [N] the step force's adjoint on that dof (without it)
  int32_t seed;         // closing evaluation: the output load case only
};

// boundary-data time programs: scales evaluated on the device from the
// clock (t, k), written into the boundary records' values
struct ProgDev {
  int32_t nprog;
  const int32_t *kind;   // [nprog]
  const double *abc;     // [nprog][3]
  const int32_t *cptr;   // [nprog+1] expression code ranges
  const ProgOp *code;
  int32_t nent;          // program-driven record values
  const int32_t *ent_rec;
  const int32_t *ent_dof;  // dof 0..5, + 8 when constrained (acceleration)
  const int32_t *ent_prog;
  const double *ent_base;  // dl_set_bcs value
  double *bc_a;          // record arrays written ([6][nbc])
  double *bc_f;
  int32_t nbc;
  double *clock_t;       // [1] time
  int64_t *clock_k;      // [1] steps taken
  int32_t clock_on;      // the clock moves without programs too (updateWhen)
};

// ImpactBC groups (ImpactModel.cpp:74-104): the group's dof gets the
// prescribed acceleration -R / w, R the group's internal-force sum at the
// step's start state (the reaction the reference's GroupOutput computed
// after the previous step); before the first output the dof is free
struct ImpDev {
  int32_t ngroups;
  const int32_t *ptr;    // [ngroups+1]
  const int32_t *nodes;
  const int32_t *dof;    // [ngroups]
  const double *w;       // [ngroups]
  double *R;             // [ngroups]
  int32_t *ready;        // [1]
  double *bc_a;          // record accelerations written ([6][nbc])
  int32_t nbc;
};

// element-major plastic records (DevData::pst, eo, ecur)
#define PST_FLAG 16u // eflag bit 4
__device__ __forceinline__ void pst_load(const double *pst, const double *aiso, int64_t e, bool iso,
                                         double ep[6], double al[7]) {
  const double2 *p = (const double2 *)(pst + 12 * e);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double2 a = p[k], b = p[3 + k];
    ep[2 * k] = a.x;
    ep[2 * k + 1] = a.y;
    al[1 + 2 * k] = b.x;
    al[2 + 2 * k] = b.y;
  }
  al[0] = iso ? aiso[e] : 0.0;
}
__device__ __forceinline__ void pst_zero(double ep[6], double al[7]) {
#pragma unroll
  for (int i = 0; i < 6; i++) ep[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 7; i++) al[i] = 0.0;
}
__device__ __forceinline__ void pst_store(double *pst, double *aiso, int64_t e, bool iso,
                                          const double ep[6], const double al[7]) {
  double2 *p = (double2 *)(pst + 12 * e);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    p[k] = make_double2(ep[2 * k], ep[2 * k + 1]);
    p[3 + k] = make_double2(al[1 + 2 * k], al[2 + 2 * k]);
  }
  if (iso) aiso[e] = al[0];
}
__device__ __forceinline__ void rec6_load(const double *a, int64_t e, double v[6]) {
  const double2 *p = (const double2 *)(a + 6 * e);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double2 x = p[k];
    v[2 * k] = x.x;
    v[2 * k + 1] = x.y;
  }
}
__device__ __forceinline__ void rec6_store(double *a, int64_t e, const double v[6]) {
  double2 *p = (double2 *)(a + 6 * e);
#pragma unroll
  for (int k = 0; k < 3; k++) p[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// end-of-step sums of an explicit step (k_step_end1, or one more block of
// the step's k_joint launch): the reaction groups' rows and the dissipation
struct StepEnd {
  const double *rbuf;
  const int32_t *rgptr;
  int32_t ngroups;
  int32_t nsums;        // ngroups (+ 1 with the dissipation sum)
  const double *dbuf;
  int64_t ndb;
  double *diss_total;
  double *hist_react;
  double *hist_diss;
  int32_t *step;
  int32_t tail;         // reaction row only, the counter stays
  // blocks: 1 (up to 8 sums, two at a time) or one per sum; with several,
  // the last block to finish (counted in *done) moves the step counter
  int32_t nblocks;
  int32_t *done;
};

// MilneDevice attempts: predicted then corrected state, error partials and
// the step-size control state, all on the device
struct MilneDev {
  double *u0;     // the context's state (start of the attempt, then corrected)
  double *r0;
  double *v0;
  double *u1;     // [3][N] predicted displacement, then the start state's
  double *r1;     // [4][N] rotation (quaternion; additive mode: 3)
  double *v1;     // [6][N]
  double *part;   // [2][nblk] per-block sums of squared differences (u, v)
  int32_t nblk;   // partial sums per quantity
  int32_t jofs;   // joint kernel's first slot (fused corrector), else 0
  double ls;      // length scale of the error measure
  double *dt;     // [1] step size of the current attempt
  double *tacc;   // [1] sum of the accepted step sizes
  int32_t *acc;   // [1] last attempt accepted
  int64_t *nacc;  // [1] accepted attempts
  double *log;    // [attempts][4]: dt, error, accepted, next dt
  double *clock_t;   // program clock advanced on accept, or null
  int64_t *clock_k;
  int32_t *imp_ready; // impact groups: set on accept, or null
  const int32_t *imp_dof; // impact group -> dof
};

struct MilneCtl {
  double dt_min, dt_max, prec, safety, incr, decr, len_scale;
  int64_t ndof;
};

// launchers (dl_kernels.hip)
void launch_init_strain(const DevData &d, hipStream_t s);
// m: the MilneDevice buffers of the fused modes (null otherwise)
void launch_strut_pp(const DevData &d, const StepParams &sp, const PlasticParams &pp, int mode,
                     bool so3, int mat, bool contact, hipStream_t s, const MilneDev *m = nullptr);
void launch_joint(const DevData &d, const StepParams &sp, int mode, bool so3, hipStream_t s,
                  const MilneDev *m = nullptr, const StepEnd *se = nullptr);
void launch_force_extras(const DevData &d, bool so3, bool gyro, hipStream_t s);
// step end: the reaction groups' sums of rbuf into history row *step - 1
// (the reaction of the previous step, evaluated at this step's start state;
// none for *step == 0), the dissipation increment into row *step, then
// *step += 1.  tail: only the reaction row *step - 1 (the call's closing
// evaluation), the counter stays.
void launch_step_end(const DevData &d, const int32_t *rgptr, int32_t ngroups, bool diss,
                     int64_t ndb, double *diss_total, double *hist_react, double *hist_diss,
                     int32_t *step, hipStream_t s, bool tail = false);
// the sums of launch_step_end as blocks to hand to launch_joint (done: a
// device counter, zero between steps, for more than 8 sums)
StepEnd make_step_end(const DevData &d, const int32_t *rgptr, int32_t ngroups, bool diss,
                      int64_t ndb, double *diss_total, double *hist_react, double *hist_diss,
                      int32_t *step, int32_t *done, bool tail = false);
// GroupOutput reaction entries at the current state (after the contact pass
// of this state): rbuf[k] = the dof share of entry k
void launch_react(const DevData &d, bool so3, bool plastic, bool contact, hipStream_t s);
void launch_set_i32(int32_t *p, int32_t v, hipStream_t s);
// updateWhen (ExplicitModule.cpp:256-266): at the start of a step or
// attempt, Js = R J0 R^T at the current rotations where the condition
// program holds at the clock (t, k) or *force is set (then cleared)
struct MassUpd {
  const ProgOp *code;
  int32_t nops;
  const double *clock_t;
  const int64_t *clock_k;
  int32_t *force;
  double *Js;
};
void launch_mass_update(const DevData &d, const MassUpd &mu, bool so3, hipStream_t s);
// p[0 .. n) = 0 by a kernel (no memset node when a caller captures the stream)
void launch_zero_f64(double *p, int64_t n, hipStream_t s);
void launch_commit_pl(const DevData &d, bool plastic, double *epot, double *ediss, hipStream_t s);
void launch_mass_blocks(const DevData &d, bool so3, bool lumped, double *blocks, hipStream_t s);
void launch_state_io(const DevData &d, bool so3, bool to_dof, double *disp, double *velo,
                     hipStream_t s);
// adjoint (dl_adjoint.hip): [gather of the element ends +] node update, element force
void launch_adj_node(const DevData &d, const AdjDev &a, double dt, bool gather, hipStream_t s);
void launch_adj_elem(const DevData &d, const AdjDev &a, hipStream_t s);
// elastoplastic element force (return map and output load case), reversed;
// expk: the yield family's exponent kind (return_map's template argument)
void launch_adj_elem_pl(const DevData &d, const AdjDev &a, const PlasticParams &pp, int expk,
                        hipStream_t s);
// objective weight of the call's closing reaction row: fbar = rbar at the
// reaction groups' dofs, zero elsewhere
void launch_adj_seed(const DevData &d, const AdjDev &a, hipStream_t s);
// MilneDevice (dl_milne.hip): predictor and corrector node passes over the
// internal force in d.fout, step-size control, and the accept/reject commit
void launch_milne_pred(const DevData &d, const MilneDev &m, bool so3, hipStream_t s);
void launch_milne_corr(const DevData &d, const MilneDev &m, bool so3, hipStream_t s);
void launch_milne_ctl(const MilneDev &m, const MilneCtl &mc, const int32_t *step, hipStream_t s);
void launch_milne_restore(const DevData &d, const MilneDev &m, bool so3, hipStream_t s);
void launch_milne_commit_pl(const DevData &d, const MilneDev &m, hipStream_t s);
void launch_set_f64(double *p, double v, hipStream_t s);
// observables (dl_tables.hip): element table kind 0 strain, 1 mat_strain,
// 2 stress, 3 mat_stress into out [E][6] (may be null), element potential
// energy into epot [E] (may be null); node kinetic energy [N]
void launch_elem_table(const DevData &d, const PlasticParams &pp, bool so3, int mat, int kind,
                       double *out, double *epot, hipStream_t s);
void launch_node_ekin(const DevData &d, bool so3, double *ekin, hipStream_t s);
// boundary programs (dl_kernels.hip): values at clock time + dt (dtp: device
// dt, else dt); clock advance by dt
void launch_bc_program(const ProgDev &P, double dt, const double *dtp, hipStream_t s);
void launch_clock_advance(const ProgDev &P, double dt, hipStream_t s);
// ImpactBC (dl_kernels.hip): after the joint kernel of a step, the deferred
// impact nodes (their force in fout) are summed per group and updated
void launch_impact_step(const DevData &d, const StepParams &sp, const ImpDev &I, bool so3,
                        hipStream_t s);
// group sums of fout into I.R and the records' accelerations (MilneDevice)
void launch_impact_sum(const ImpDev &I, const DevData &d, hipStream_t s);

#endif
