// dl_ctx.h -- the context behind the opaque dl_ctx handle of the C ABI,
// shared by the host files (dl_api.cpp: set-up, actions, the explicit step;
// dl_adjoint_api.cpp: checkpointed forward and adjoint).
#ifndef DL_CTX_H
#define DL_CTX_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dl_gpu.h"
#include "dl_contact.h"
#include "dl_internal.h"

// records the message for dl_last_error and returns code
int fail(int code, const std::string &msg);

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t _e = (x);                                                                 \
    if (_e != hipSuccess)                                                                \
      return fail(DL_ERR_HIP, std::string("HIP: ") + hipGetErrorString(_e) + " at " #x); \
  } while (0)

// boundary data kept on the host for re-uploads (constraints, loads, reaction group)
struct BcHost {
  std::vector<uint8_t> cmask;
  std::vector<double> cval, fext;
  std::vector<int32_t> react;     // reaction nodes, concatenated by group
  // ImpactBC groups: nodes concatenated by group, per group dof and weight
  std::vector<int32_t> imp_nodes, imp_ptr, imp_dof;
  std::vector<double> imp_w;
};

// host copy of the mesh (set-up tables are rebuilt from it when the
// sections change, and the adjoint maps its results back through it)
struct MeshHost {
  int64_t N = 0, E = 0;
  int32_t R = 0;
  std::vector<double> X;        // [N][3]
  std::vector<int32_t> en;      // [E][2]
  std::vector<int32_t> es;      // [E]
  std::vector<int64_t> eptr;    // [R+1] elements of strut s: [eptr[s], eptr[s+1])
  std::vector<int32_t> cnode;   // [E+R] chain nodes of strut s: [eptr[s]+s, eptr[s+1]+s]
};

// section data kept on the host (per strut)
struct SecHost {
  std::vector<double> area, I1, I2, J, kappa, young, shear, rho;
  double inertia_correct = 1.0;
};

// adjoint: checkpoints of the last checkpointed forward run and the buffers
// of the reverse sweep
struct AdjState {
  bool valid = false;
  int64_t nsteps = 0;
  int32_t every = 0;
  double dt = 0.0;
  int64_t ckpt_slots = 0;
  double *ckpt = nullptr;      // [slots][13][N]: u 3, rot 4, v 6
  double *final_state = nullptr; // [13][N]
  int64_t tape_rows = 0;
  double *tape = nullptr;      // [rows][19][N]
  bool ready = false;          // reverse-sweep buffers allocated
  double *ubar = nullptr, *qbar = nullptr, *vbar = nullptr, *fbar = nullptr;
  double *ebuf = nullptr, *Cbar = nullptr, *mtbar = nullptr, *J0bar = nullptr, *crbar = nullptr;
  int32_t *nptr = nullptr, *nlist = nullptr, *slot_group = nullptr;
  double *rbar = nullptr;
  int64_t rbar_cap = 0;
  // elastoplastic material: committed plastic state (the [E][12] records,
  // then [E] alpha_0) at every checkpoint, at the end of the forward run and
  // at the start of every step of the segment being reversed
  int64_t ckpt_pl_slots = 0;
  double *ckpt_pl = nullptr;   // [slots][13 E]
  double *final_pl = nullptr;  // [13 E]
  uint8_t *final_active = nullptr; // [E] plastic active flags after the forward run
  int64_t ptape_rows = 0;
  double *ptape = nullptr;     // [rows][13 E]
  double *pbar = nullptr, *Wbar = nullptr, *rnode = nullptr, *fstep_r = nullptr;
  // contact candidate lists: the steps of the forward run that built them
  // (every checkpoint builds), so that the reverse sweep rebuilds them only
  // there, from the taped state of the build
  int64_t blog_cap = 0;
  uint8_t *blog = nullptr;     // [nsteps] device
  std::vector<uint8_t> blog_h; // [nsteps] host copy
};

// boundary-data time programs (host copy)
struct ProgState {
  std::vector<int32_t> kind, cptr, dof_prog;
  std::vector<double> abc;
  std::vector<ProgOp> code;
};

// MilneDevice attempts: device buffers and the captured attempt graph
struct MilneState {
  bool ready = false;
  MilneDev m{};
  int64_t part_cap = 0;
  int64_t log_cap = 0;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_gen = -1;
  MilneCtl graph_ctl{};
};

struct dl_ctx {
  BcHost bch;
  std::vector<double> ediss_acc_h; // per-element dissipation of the action path (energyDiss_)
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool so3 = true;
  bool plastic = false;
  int mat_kind = 0; // 0 elastic; 1/2/3 plastic with yield exponents all 1 / all 2 / general
  bool contact = false;
  bool initialized = false;
  bool track_diss = false;
  int64_t n_mirrored = 0; // elements with -I end frames (struts along -e3)
  int64_t n_wave_joints = 0; // shared nodes summed within one wavefront (k_strut)
  uint32_t flags = 0;
  PlasticParams pp{};
  DevData d{};
  std::vector<void *> allocs;
  // host copies needed by actions
  std::vector<double> w_e;        // Gauss weight (L0) per element
  std::vector<int32_t> elem_perm; // identity (elements are stored in input order)
  // reaction groups: entries [rgptr[g], rgptr[g+1]) of rbuf belong to group
  // g; the group's nodes are slots [rn_ptr_h[g], rn_ptr_h[g+1]) of the
  // boundary records' reaction slots
  int64_t n_react = 0;
  int32_t n_groups = 0;
  int32_t *rgptr = nullptr;
  std::vector<int32_t> rn_ptr_h;
  int32_t *step_ctr = nullptr;   // device step counter (history row)
  int32_t *end_done = nullptr;   // blocks of the end-of-step sums finished (zero between steps)
  double *ndt = nullptr;          // per-node dt
  // captured step graphs
  int32_t graph_steps = 0;
  hipGraphExec_t gexec = nullptr;
  int64_t gen = 0, graph_gen = -1;
  double graph_dt = 0.0;
  // boundary records (host) for re-upload
  std::vector<int32_t> bc_host;
  // device buffers for history
  double *hist_react = nullptr, *hist_diss = nullptr;
  int64_t *hist_nc = nullptr, *hist_ncand = nullptr;
  int64_t hist_cap = 0, hist_n = 0;
  double *diss_total = nullptr;
  int32_t strut_blocks = 0;
  // action path energy accumulators
  double *ediss_acc = nullptr, *epot_e = nullptr, *ediss_inc = nullptr;
  // device address of this GPU's scratch array (dl_explicit_step_noalloc)
  double *scratch = nullptr;
  // staging buffer of the DOF-layout conversions (GET_MATRIX2, set/get_state)
  double *stage = nullptr;
  size_t stage_cap = 0;
  // timing
  bool timing = false;
  double t_step = 0, t_contact = 0, t_refresh = 0, t_react = 0, t_strut = 0, t_joint = 0, t_node = 0;
  int64_t t_count = 0;
  // contact
  ContactState cs{};
  // host copies of the mesh and sections, and the material
  MeshHost mesh;
  SecHost sec;
  dl_material mat{};
  std::vector<double> pl_weights; // [R][6] yield weights (elastoplastic)
  std::vector<double> crad;       // [R] contact radius
  std::vector<int32_t> rgptr_h;   // host copy of the reaction group pointer
  // adjoint
  AdjState adj;
  // MilneDevice
  MilneState milne;
  // boundary programs, point masses (per node, added to mt and J0's diagonal)
  ProgState prog;
  ProgDev pd{};
  // updateWhen of the explicit modules (dl_set_mass_update)
  MassUpd mu{};
  ImpDev imp{};
  std::vector<double> pm_mass, pm_inertia;
};

template <class T>
static inline int dalloc(dl_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
  c->allocs.push_back((void *)*p);
  return DL_OK;
}

// frees a buffer from dalloc (no-op for null)
void dfree(dl_ctx *c, void *p);

template <class T>
static inline int upload(dl_ctx *c, T **p, const std::vector<T> &h) {
  int rc = dalloc(c, p, h.size());
  if (rc) return rc;
  if (!h.empty()) HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return DL_OK;
}

// one explicit step (contact, reaction entries, strut chains, joints, step end)
int launch_one_step(dl_ctx *c, const StepParams &sp, hipEvent_t *ev);
// the call's closing GroupOutput evaluation: contact search and reaction
// entries at the state the call ends in, summed into the last history row
int launch_reaction_tail(dl_ctx *c);
// zeroes the per-step histories for nsteps rows and rewinds the step counter
int prepare_history(dl_ctx *c, int64_t nsteps);
// nsteps further steps (captured graph when configured), continuing the step counter
int advance_steps(dl_ctx *c, int64_t nsteps, const StepParams &sp);
int check_err(dl_ctx *c);
// rebuilds the device list of program-driven boundary values (dl_program.cpp)
int prog_entries(dl_ctx *c);

#endif
