// dl_api.cpp -- host side of the C ABI (include/dl_gpu.h).
//
// dl_create plays the role of the reference's model-tree construction and
// INIT-time set-up: strut chains, per-(strut,node) reference frames
// (initRotation_, SpecialCosseratRodModel.cpp:574-658), material tables
// (ElasticRodMaterial.cpp:265-287), block-lumped mass (assembleM_,
// SpecialCosseratRodModel.cpp:969-1031, split into a constant per-node part:
// J_node(theta) = R(theta) J0 R(theta)^T), and the shared-node gather map.
// Everything is uploaded once; the step runs on resident device state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dl_gpu.h"
#include "dl_contact.h"
#include "dl_ctx.h"
#include "dl_internal.h"
#include "dl_setup.h"

static thread_local std::string g_err;

// ------------------------------------------------------------------
// scratch of the allocation-free step (dl_explicit_step_noalloc): the step
// counter and the per-step history rows live in this array instead of the
// context's history buffers.  Each GPU holds its own copy; its device
// address is looked up once per device (dl_create, with the device
// current) and only that address is used from the host.
// ------------------------------------------------------------------
static __device__ __align__(256) double scratch[DL_NOALLOC_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_dev; // per device ordinal

static int scratch_address(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DL_ERR_HIP;
    }
    g_scratch_dev[device] = (double *)p;
  }
  *out = g_scratch_dev[device];
  return DL_OK;
}

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

extern "C" const char *dl_last_error(void) { return g_err.c_str(); }

void dfree(dl_ctx *c, void *p) {
  if (!p) return;
  for (size_t i = 0; i < c->allocs.size(); i++)
    if (c->allocs[i] == p) {
      c->allocs[i] = c->allocs.back();
      c->allocs.pop_back();
      hipFree(p);
      return;
    }
}

// ------------------------------------------------------------------
// section-dependent tables: stiffness diag C = (kGA, kGA, EA, EI1, EI2, GJ)
// (ElasticRodMaterial.cpp:265-287), kinematic hardening Hk, lumped mass
// ------------------------------------------------------------------
static int store_sections(dl_ctx *c, const dl_sections *sec) {
  const int32_t R = c->mesh.R;
  SecHost h;
  h.area.assign(sec->area, sec->area + R);
  h.I1.assign(sec->I1, sec->I1 + R);
  h.I2.assign(sec->I2, sec->I2 + R);
  h.J.assign(sec->J, sec->J + R);
  h.kappa.assign(sec->shear_factor, sec->shear_factor + R);
  h.young.assign(sec->young, sec->young + R);
  h.shear.assign(sec->shear_modulus, sec->shear_modulus + R);
  h.rho.assign(sec->density, sec->density + R);
  h.inertia_correct = sec->inertia_correct;
  for (int32_t s = 0; s < R; s++)
    if (!(h.area[s] > 0) || !(h.young[s] > 0) || !(h.shear[s] > 0) || !(h.I1[s] > 0) ||
        !(h.I2[s] > 0) || !(h.J[s] > 0) || h.rho[s] < 0)
      return fail(DL_ERR_INPUT, "invalid section/material data for strut " + std::to_string(s));
  c->sec = std::move(h);
  return DL_OK;
}

// builds C, Hk (elastoplastic), mt, J0 from the stored sections; allocates
// the device tables when first, else overwrites them in place
static int section_tables(dl_ctx *c, bool first) {
  const MeshHost &m = c->mesh;
  const SecHost &sc = c->sec;
  const int64_t N = m.N;
  const int32_t R = m.R;
  std::vector<double> Cd(6 * (size_t)R), mt(N, 0.0), J0(6 * N, 0.0), pw, Hk;
  if (c->plastic) {
    pw.resize(6 * (size_t)R);
    Hk.resize(6 * (size_t)R);
  }
  for (int32_t s = 0; s < R; s++) {
    const double G = sc.shear[s], Y = sc.young[s], kap = sc.kappa[s], A = sc.area[s];
    Cd[0 * R + s] = G * kap * A;
    Cd[1 * R + s] = G * kap * A;
    Cd[2 * R + s] = Y * A;
    Cd[3 * R + s] = Y * sc.I1[s];
    Cd[4 * R + s] = Y * sc.I2[s];
    Cd[5 * R + s] = G * sc.J[s];
    if (c->plastic)
      for (int i = 0; i < 6; i++) {
        pw[i * R + s] = c->pl_weights[6 * s + i];
        Hk[i * R + s] = c->mat.kin_const[i] + c->mat.kin_ratio * Cd[i * R + s];
      }
  }
  walk_elements(m, [&](int32_t s, int64_t, int32_t a, int32_t b, const double *La, const double *Lb) {
    add_elem_mass(sc, s, elem_jac(m.X.data(), a, b), a, b, La, Lb, N, mt.data(), J0.data());
  });
  // point masses (MassModel.cpp:92-108): isotropic additions to the
  // translational mass and to the rotational diagonal (R (J0 + m I) R^T =
  // R J0 R^T + m I, so the addition can sit in the reference inertia)
  if (!c->pm_mass.empty())
    for (int64_t n = 0; n < N; n++) mt[n] += c->pm_mass[n];
  if (!c->pm_inertia.empty())
    for (int64_t n = 0; n < N; n++)
      for (int i = 0; i < 3; i++) J0[i * N + n] += c->pm_inertia[n];
  // mass checks (ExplicitModule.cpp:452-454): positive mass and inertia
  for (int64_t n = 0; n < N; n++) {
    const double a = J0[n], b = J0[N + n], cc = J0[2 * N + n], xy = J0[3 * N + n],
                 xz = J0[4 * N + n], yz = J0[5 * N + n];
    const double det = a * (b * cc - yz * yz) - xy * (xy * cc - yz * xz) + xz * (xy * yz - b * xz);
    if (!(mt[n] > 2.2250738585072014e-308) || !(a > 0) || !(a * b - xy * xy > 0) || !(det > 0))
      return fail(DL_ERR_ZERO_MASS, "Zero (or negative) masses cannot be inversed! (node " +
                                        std::to_string(n) + ")");
  }
  DevData &d = c->d;
  if (first) {
    double *p;
    int rc;
    if ((rc = upload(c, &p, mt))) return rc;
    d.mt = p;
    if ((rc = upload(c, &p, J0))) return rc;
    d.J0 = p;
    if ((rc = upload(c, &p, Cd))) return rc;
    d.C = p;
    if (c->plastic) {
      if ((rc = upload(c, &p, pw))) return rc;
      d.pw = p;
      if ((rc = upload(c, &p, Hk))) return rc;
      d.Hk = p;
    }
    return DL_OK;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy((void *)d.mt, mt.data(), mt.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.J0, J0.data(), J0.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.C, Cd.data(), Cd.size() * sizeof(double), hipMemcpyHostToDevice));
  if (c->plastic)
    HIPCHK(hipMemcpy((void *)d.Hk, Hk.data(), Hk.size() * sizeof(double), hipMemcpyHostToDevice));
  return DL_OK;
}

// ------------------------------------------------------------------
// dl_create
// ------------------------------------------------------------------
extern "C" int dl_create(const dl_mesh *mesh, const dl_sections *sec, const dl_material *mat,
                         const dl_contact *contact, const dl_options *opt, dl_ctx **out) {
  if (!out) return fail(DL_ERR_INPUT, "out is NULL");
  *out = nullptr;
  if (!mesh || !sec) return fail(DL_ERR_INPUT, "mesh and sections are required");
  const int64_t N = mesh->n_nodes, E = mesh->n_elems;
  const int32_t R = mesh->n_struts;
  if (N <= 0 || E <= 0 || R <= 0) return fail(DL_ERR_INPUT, "empty mesh");
  if (N >= INT32_MAX || E + R >= INT32_MAX) return fail(DL_ERR_INPUT, "mesh too large for int32 indices");
  const int32_t *en = mesh->elem_nodes, *es = mesh->elem_strut;
  for (int64_t e = 0; e < E; e++) {
    if (en[2 * e] < 0 || en[2 * e] >= N || en[2 * e + 1] < 0 || en[2 * e + 1] >= N)
      return fail(DL_ERR_INPUT, "element node index out of range at element " + std::to_string(e));
    if (en[2 * e] == en[2 * e + 1]) return fail(DL_ERR_INPUT, "degenerate element " + std::to_string(e));
    if (es[e] < 0 || es[e] >= R) return fail(DL_ERR_INPUT, "element strut out of range");
    if (e > 0 && es[e] < es[e - 1])
      return fail(DL_ERR_INPUT, "elements of a strut must be contiguous (elem_strut non-decreasing)");
    if (e > 0 && es[e] == es[e - 1] && en[2 * e] != en[2 * e - 1])
      return fail(DL_ERR_INPUT, "strut elements must form a chain at element " + std::to_string(e));
  }
  std::vector<int64_t> eptr(R + 1, 0);
  for (int64_t e = 0; e < E; e++) eptr[es[e] + 1]++;
  for (int32_t s = 0; s < R; s++) {
    if (eptr[s + 1] == 0) return fail(DL_ERR_INPUT, "strut without elements: " + std::to_string(s));
    eptr[s + 1] += eptr[s];
  }
  const bool plastic = mat && mat->model == DL_ELASTOPLASTIC;
  if (plastic && !mat->weights) return fail(DL_ERR_INPUT, "plastic material needs yield weights");
  if (plastic)
    for (int i = 0; i < 6; i++)
      if (!(mat->exponent[i] >= 1.0)) return fail(DL_ERR_INPUT, "yield exponents must be >= 1");
  const bool so3 = opt ? opt->so3 != 0 : true;

  dl_ctx *c = new dl_ctx();
  c->so3 = so3;
  c->plastic = plastic;
  c->flags = opt ? opt->flags : 0u;
  c->track_diss = plastic && (c->flags & DL_TRACK_DISSIPATION);
  c->device = opt ? opt->device : 0;
  int rc = DL_OK;
#define TRY(x)                 \
  do {                         \
    rc = (x);                  \
    if (rc) {                  \
      dl_destroy(c);           \
      return rc;               \
    }                          \
  } while (0)
  {
    hipError_t he = hipSetDevice(c->device);
    if (he != hipSuccess) {
      delete c;
      return fail(DL_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(he));
    }
  }
  {
    const int rc0 = scratch_address(c->device, &c->scratch);
    if (rc0) {
      delete c;
      return rc0;
    }
  }
  if (opt && opt->stream) {
    c->stream = (hipStream_t)opt->stream;
  } else {
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
      delete c;
      return fail(DL_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    c->own_stream = true;
  }

  // chains
  std::vector<int32_t> cnode(E + R);
  for (int32_t s = 0; s < R; s++) {
    const int64_t c0 = eptr[s] + s;
    cnode[c0] = en[2 * eptr[s]];
    for (int64_t e = eptr[s]; e < eptr[s + 1]; e++) cnode[c0 + (e - eptr[s]) + 1] = en[2 * e + 1];
  }
  // gyro multiplicity: number of distinct struts holding the node
  std::vector<int32_t> nstr(N, 0), last(N, -1);
  for (int32_t s = 0; s < R; s++)
    for (int64_t p = eptr[s] + s; p <= eptr[s + 1] + s; p++) {
      const int32_t n = cnode[p];
      if (last[n] != s) { nstr[n]++; last[n] = s; }
    }
  for (int64_t n = 0; n < N; n++)
    if (nstr[n] == 0) {
      dl_destroy(c);
      return fail(DL_ERR_ZERO_MASS, "node " + std::to_string(n) + " belongs to no element (zero mass)");
    }
  // chains: struts cut into pieces of at most 64 elements (one wavefront);
  // a cut node then occurs twice and becomes a shared node
  std::vector<int64_t> gptr;
  for (int32_t s = 0; s < R; s++)
    for (int64_t b = eptr[s]; b < eptr[s + 1]; b += 64) gptr.push_back(b);
  gptr.push_back(E);
  const int64_t NG = (int64_t)gptr.size() - 1;
  auto gnode = [&](int64_t g, int64_t k) -> int32_t { // chain position k of chain g
    return k == 0 ? en[2 * gptr[g]] : en[2 * (gptr[g] + k - 1) + 1];
  };
  std::vector<int32_t> occ(N, 0);
  for (int64_t g = 0; g < NG; g++)
    for (int64_t k = 0; k <= gptr[g + 1] - gptr[g]; k++) occ[gnode(g, k)]++;
  // wavefront packing: a chain never straddles two 64-lane wavefronts
  std::vector<int32_t> lane_e, elane(E, -1);
  for (int64_t g = 0; g < NG; g++) {
    const int64_t len = gptr[g + 1] - gptr[g];
    if ((int64_t)(lane_e.size() % 64) + len > 64)
      while (lane_e.size() % 64) lane_e.push_back(-1);
    for (int64_t k = 0; k < len; k++) {
      elane[gptr[g] + k] = (int32_t)lane_e.size();
      lane_e.push_back((int32_t)(gptr[g] + k));
    }
  }
  while (lane_e.size() % 256) lane_e.push_back(-1);
  // In-wavefront joints: a shared node all of whose chain positions are
  // node a of an element (no chain end) in one wavefront -- a BCC cell
  // centre's eight struts -- is summed by shuffles in k_strut, in lane order
  // (= chain order, the order of its slots), and updated by its lowest lane.
  // Its elements carry sa = -2 - (that lane within the wavefront); it gets no
  // slots and no joint-kernel thread.
  std::vector<int32_t> jwave(N, -2), jlow(N, 64);  // -2 unseen, -1 not eligible
  if (!(c->flags & DL_JOINTS_IN_MEMORY))
    for (int64_t g = 0; g < NG; g++) {
      const int64_t len = gptr[g + 1] - gptr[g];
      for (int64_t k = 0; k <= len; k++) {
        const int32_t n = gnode(g, k);
        if (occ[n] < 2 || jwave[n] == -1) continue;
        if (k == len) { jwave[n] = -1; continue; } // a chain end
        const int32_t L = elane[gptr[g] + k];
        if (jwave[n] == -2) jwave[n] = L / 64;
        else if (jwave[n] != L / 64) { jwave[n] = -1; continue; }
        jlow[n] = std::min(jlow[n], L % 64);
      }
    }
  auto in_wave = [&](int32_t n) { return occ[n] > 1 && jwave[n] >= 0; };
  // slots for the other shared nodes in (chain, position) order, i.e. strut order
  std::vector<int32_t> sa(E, -1), sb(E, -1), jidx(N, -1), jnode;
  std::vector<uint8_t> eflag(E, 0);
  std::vector<std::pair<int32_t, int32_t>> slot_node; // (slot, node)
  int32_t nslot = 0;
  for (int64_t g = 0; g < NG; g++) {
    const int64_t len = gptr[g + 1] - gptr[g];
    for (int64_t k = 0; k <= len; k++) {
      const int32_t n = gnode(g, k);
      int32_t sl = -1;
      if (in_wave(n)) {
        sl = -2 - jlow[n];
      } else if (occ[n] > 1) {
        sl = nslot++;
        slot_node.push_back({sl, n});
      }
      if (k < len) sa[gptr[g] + k] = sl;
      else sb[gptr[g] + len - 1] = sl;
    }
    eflag[gptr[g]] |= 1u;
    eflag[gptr[g] + len - 1] |= 2u;
  }
  for (int64_t n = 0; n < N; n++)
    if (occ[n] > 1 && !in_wave(n)) { jidx[n] = (int32_t)jnode.size(); jnode.push_back((int32_t)n); }
  const int32_t NJ = (int32_t)jnode.size();
  // renumber the slots so that the contributions to one shared node are
  // contiguous (in chain order): the joint kernel then reads
  // sbuf[jptr[j] .. jptr[j+1]) without indirection
  std::vector<int32_t> jptr(NJ + 1, 0), jslot(nslot);
  for (auto &sn : slot_node) jptr[jidx[sn.second] + 1]++;
  for (int32_t j = 0; j < NJ; j++) jptr[j + 1] += jptr[j];
  {
    std::vector<int32_t> fill(jptr.begin(), jptr.end() - 1), remap(nslot);
    for (auto &sn : slot_node) remap[sn.first] = fill[jidx[sn.second]]++;
    for (int64_t e = 0; e < E; e++) {
      if (sa[e] >= 0) sa[e] = remap[sa[e]];
      if (sb[e] >= 0) sb[e] = remap[sb[e]];
    }
    c->n_wave_joints = 0;
    for (int64_t n = 0; n < N; n++) c->n_wave_joints += in_wave((int32_t)n) ? 1 : 0;
    for (int32_t t = 0; t < nslot; t++) jslot[t] = t;
  }
  // host copies of the mesh, sections and material (tables are rebuilt from
  // them when the sections change; the adjoint maps back through them)
  {
    MeshHost &m = c->mesh;
    m.N = N; m.E = E; m.R = R;
    m.X.assign(mesh->coords, mesh->coords + 3 * N);
    m.en.assign(en, en + 2 * E);
    m.es.assign(es, es + E);
    m.eptr = eptr;
    m.cnode = cnode;
  }
  if (plastic) {
    c->mat = *mat;
    c->pl_weights.assign(mat->weights, mat->weights + 6 * (size_t)R);
    c->mat.weights = nullptr;
  }
  TRY(store_sections(c, sec));
  // per (element, end) frame quaternions and Gauss weights.  The -I frame
  // of a strut along -e3 (SpecialCosseratRodModel.cpp:651-654) is no
  // rotation: such elements keep identity quaternions and flag bit 2; an
  // element with one -I end and one rotation end (a curved strut passing
  // through -e3) has a reflection as its relative rotation, which no
  // quaternion route can follow
  std::vector<double> qN(8 * E), wE(E);
  auto minus_eye = [](const double *L) {
    for (int i = 0; i < 9; i++)
      if (L[i] != ((i % 4 == 0) ? -1.0 : 0.0)) return false;
    return true;
  };
  int64_t mixed = -1;
  walk_elements(c->mesh, [&](int32_t, int64_t e, int32_t a, int32_t b, const double *La,
                             const double *Lb) {
    const bool ma = minus_eye(La), mb = minus_eye(Lb);
    if (ma != mb && mixed < 0) mixed = e;
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const Quat qa = mat2quat(ma ? I : La), qb = mat2quat(mb ? I : Lb);
    if (ma && mb) {
      eflag[e] |= 4u;
      c->n_mirrored++;
    }
    qN[0 * E + e] = qa.w; qN[1 * E + e] = qa.x; qN[2 * E + e] = qa.y; qN[3 * E + e] = qa.z;
    qN[4 * E + e] = qb.w; qN[5 * E + e] = qb.x; qN[6 * E + e] = qb.y; qN[7 * E + e] = qb.z;
    wE[e] = 2. * elem_jac(c->mesh.X.data(), a, b);
  });
  c->w_e = wE;
  if (mixed >= 0) {
    dl_destroy(c);
    return fail(DL_ERR_INPUT, "element " + std::to_string(mixed) +
                                  " joins a -I frame (tangent along -e3) and a rotation frame");
  }

  // upload
  DevData &d = c->d;
  d.N = N;
  d.E = E;
  d.R = R;
  {
    std::vector<double> X0(3 * N);
    for (int64_t n = 0; n < N; n++)
      for (int k = 0; k < 3; k++) X0[k * N + n] = mesh->coords[3 * n + k];
    double *p;
    TRY(upload(c, &p, X0));
    d.X0 = p;
  }
  {
    std::vector<double> z3(3 * N, 0.0), z6(6 * N, 0.0);
    TRY(upload(c, &d.u, z3));
    TRY(upload(c, &d.v, z6));
    std::vector<double> r(so3 ? 4 * N : 3 * N, 0.0);
    if (so3)
      for (int64_t n = 0; n < N; n++) r[n] = 1.0;
    TRY(upload(c, &d.rot, r));
  }
  {
    double *p;
    std::vector<double> g(N);
    for (int64_t n = 0; n < N; n++) g[n] = (double)nstr[n];
    TRY(upload(c, &p, g)); d.gm = p;
    TRY(upload(c, &p, qN)); d.qN = p;
    TRY(dalloc(c, &d.eps0, 6 * E));
    TRY(section_tables(c, true));
  }
  {
    std::vector<int32_t> bc(N, -1);
    int32_t *pi;
    TRY(upload(c, &pi, bc)); d.bc = pi;
    c->bc_host = bc;
    d.nbc = 0;
    d.bc_mask = nullptr; d.bc_a = nullptr; d.bc_f = nullptr; d.bc_rslot = nullptr;
  }
  {
    int32_t *pi;
    std::vector<int32_t> e0(E), e1(E), esv(es, es + E);
    for (int64_t e = 0; e < E; e++) {
      e0[e] = en[2 * e];
      e1[e] = en[2 * e + 1];
    }
    TRY(upload(c, &pi, e0)); d.en0 = pi;
    TRY(upload(c, &pi, e1)); d.en1 = pi;
    TRY(upload(c, &pi, esv)); d.es = pi;
    TRY(upload(c, &pi, sa)); d.sa = pi;
    TRY(upload(c, &pi, sb)); d.sb = pi;
    TRY(upload(c, &pi, lane_e)); d.lane_e = pi;
    d.NL = (int64_t)lane_e.size();
    d.lane_ident = 1;
    for (int64_t L = 0; L < d.NL && d.lane_ident; L++)
      d.lane_ident = lane_e[L] == (L < E ? (int32_t)L : -1);
    uint8_t *pu;
    TRY(upload(c, &pu, eflag)); d.eflag = pu;
    // end-b frames for k_strut: by shuffle from the next lane where it holds
    // the same frame bit for bit, else the per-element array
    {
      std::vector<double> qe[4];
      std::vector<int32_t> wave((size_t)(d.NL / 64), 0);
      bool share = !(c->flags & DL_FRAMES_PER_ELEMENT);
      int32_t ends = 0;
      for (int64_t L = 0; L < d.NL && share; L++) {
        if (L % 64 == 0) wave[L / 64] = ends;
        const int32_t e = lane_e[L];
        if (e < 0) continue;
        if (eflag[e] & 2u) {
          for (int k = 0; k < 4; k++) qe[k].push_back(qN[(4 + k) * E + e]);
          ends++;
          continue;
        }
        const int32_t e1 = (L + 1 < d.NL && (L + 1) % 64) ? lane_e[L + 1] : -1;
        for (int k = 0; k < 4 && share; k++)
          share = e1 >= 0 && std::memcmp(&qN[(4 + k) * E + e], &qN[k * E + e1], sizeof(double)) == 0;
      }
      d.qN_share = share ? 1 : 0;
      d.nqe = share ? ends : 0;
      d.qNe = nullptr;
      d.qNe_wave = nullptr;
      if (share) {
        std::vector<double> flat;
        for (int k = 0; k < 4; k++) flat.insert(flat.end(), qe[k].begin(), qe[k].end());
        double *p;
        TRY(upload(c, &p, flat)); d.qNe = p;
        TRY(upload(c, &pi, wave)); d.qNe_wave = pi;
      }
    }
    d.NJ = NJ;
    d.nslot = nslot;
    TRY(upload(c, &pi, jnode)); d.jnode = pi;
    TRY(upload(c, &pi, jptr)); d.jptr = pi;
    TRY(upload(c, &pi, jslot)); d.jslot = pi;
    TRY(dalloc(c, &d.sbuf, 6 * (size_t)std::max<int32_t>(nslot, 1)));
  }
  // material state
  {
    std::vector<double> z6e(6 * E, 0.0), z7e(7 * E, 0.0);
    TRY(upload(c, &d.ecur, z6e));
    TRY(upload(c, &d.eo, z6e));
    if (plastic) {
      std::vector<double> z12e(12 * E, 0.0), z1e(E, 0.0);
      TRY(upload(c, &d.pst, z12e));
      TRY(upload(c, &d.pst_cur, z12e));
      TRY(upload(c, &d.aiso, z1e));
      TRY(upload(c, &d.aiso_cur, z1e));
      std::vector<double> zE(E, 0.0);
      TRY(upload(c, &d.dflow, zE));
      std::vector<uint8_t> zb(E, 0);
      TRY(upload(c, &d.active, zb));
      PlasticParams &P = c->pp;
      P.yield_level = mat->yield_level;
      for (int i = 0; i < 6; i++) P.kin_coupling[i] = mat->kin_coupling[i];
      P.iso_coupling = mat->iso_coupling;
      P.iso_modulus = mat->iso_modulus;
      P.precision = mat->precision;
      P.has_iso = mat->has_iso;
      P.has_kin = mat->has_kin;
      P.max_iter = mat->max_iter;
      for (int i = 0; i < 6; i++) P.exponent[i] = mat->exponent[i];
      for (int i = 0; i < 6; i++) P.kin_const[i] = mat->kin_const[i];
      P.kin_ratio = mat->kin_ratio;
      bool all1 = true, all2 = true;
      for (int i = 0; i < 6; i++) {
        all1 = all1 && mat->exponent[i] == 1.0;
        all2 = all2 && mat->exponent[i] == 2.0;
      }
      c->mat_kind = all1 ? 1 : (all2 ? 2 : 3);
    }
  }
  TRY(dalloc(c, &d.fout, 6 * N));
  TRY(dalloc(c, &d.err, 1));
  HIPCHK(hipMemset(d.err, 0, sizeof(int32_t)));
  TRY(dalloc(c, &c->end_done, 1));
  HIPCHK(hipMemset(c->end_done, 0, sizeof(int32_t)));
  c->strut_blocks = (int32_t)(d.NL / 256);
  TRY(dalloc(c, &d.dbuf, (size_t)c->strut_blocks));
  TRY(dalloc(c, &c->diss_total, 1));
  HIPCHK(hipMemset(c->diss_total, 0, sizeof(double)));
  d.rbuf = nullptr;
  d.rslot_e = nullptr;
  d.react_dof = 2;
  // contact
  if (contact && contact->enabled) {
    c->contact = true;
    // the allocator outlives dl_create (pair lists grow later): capture the context by value
    TRY(contact_create(
        c->cs, d, mesh, contact,
        [c](void **p, size_t bytes) {
          hipError_t he = hipMalloc(p, bytes ? bytes : 1);
          if (he != hipSuccess) return fail(DL_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(he));
          c->allocs.push_back(*p);
          return (int)DL_OK;
        },
        [c](void *p) { dfree(c, p); }));
  }
#undef TRY
  *out = c;
  return DL_OK;
}

extern "C" int dl_destroy(dl_ctx *c) {
  if (!c) return DL_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->milne.gexec) hipGraphExecDestroy(c->milne.gexec);
  for (void *p : c->allocs) hipFree(p);
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  delete c;
  return DL_OK;
}

extern "C" int dl_set_point_masses(dl_ctx *c, const double *mass, const double *inertia,
                                   int64_t n_nodes) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if ((mass || inertia) && n_nodes != c->d.N) return fail(DL_ERR_INPUT, "n_nodes mismatch");
  for (const double *a : {mass, inertia})
    if (a)
      for (int64_t n = 0; n < n_nodes; n++)
        if (!(a[n] >= 0.0) || !std::isfinite(a[n]))
          return fail(DL_ERR_INPUT, "point masses must be finite and >= 0");
  hipSetDevice(c->device);
  c->adj.valid = false;
  std::vector<double> pm(mass ? mass : (const double *)nullptr, mass ? mass + n_nodes : nullptr);
  std::vector<double> pi(inertia ? inertia : (const double *)nullptr,
                         inertia ? inertia + n_nodes : nullptr);
  c->pm_mass.swap(pm);
  c->pm_inertia.swap(pi);
  return section_tables(c, false);
}

extern "C" int dl_set_sections(dl_ctx *c, const dl_sections *sec, const double *contact_radius) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!sec && !contact_radius) return fail(DL_ERR_INPUT, "nothing to set");
  if (contact_radius && !c->contact) return fail(DL_ERR_INPUT, "contact radii given without contact");
  hipSetDevice(c->device);
  c->adj.valid = false;
  if (sec) {
    SecHost keep = c->sec;
    int rc = store_sections(c, sec);
    if (rc) return rc;
    if ((rc = section_tables(c, false))) {
      c->sec = std::move(keep);
      return rc;
    }
  }
  if (contact_radius) {
    int rc = contact_set_radius(c->cs, contact_radius, c->stream);
    if (rc) return rc;
    contact_invalidate(c->cs, c->stream);
  }
  return DL_OK;
}

extern "C" int dl_set_yield_weights(dl_ctx *c, const double *weights) {
  if (!c || !weights) return fail(DL_ERR_INPUT, "NULL argument");
  if (!c->plastic) return fail(DL_ERR_STATE, "yield weights need the elastoplastic material");
  const int32_t R = c->mesh.R;
  for (int64_t k = 0; k < 6 * (int64_t)R; k++)
    if (!(weights[k] >= 0.0) || !std::isfinite(weights[k]))
      return fail(DL_ERR_INPUT, "yield weights must be finite and >= 0");
  hipSetDevice(c->device);
  c->adj.valid = false;
  c->pl_weights.assign(weights, weights + 6 * (size_t)R);
  std::vector<double> pw(6 * (size_t)R);
  for (int32_t s = 0; s < R; s++)
    for (int i = 0; i < 6; i++) pw[i * (size_t)R + s] = weights[6 * s + i];
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy((void *)c->d.pw, pw.data(), pw.size() * sizeof(double), hipMemcpyHostToDevice));
  return DL_OK;
}

extern "C" int dl_sync(dl_ctx *c) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  hipSetDevice(c->device);
  return check_err(c);
}

// reads and clears the device error flag (after a stream synchronisation);
// a capacity overflow also marks the candidate lists stale, so the next
// step rebuilds them (and reports the overflow again if it persists)
int check_err(dl_ctx *c) {
  int32_t e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->d.err, sizeof e, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (e == DL_ERR_PLASTIC_ITER) {
    HIPCHK(hipMemsetAsync(c->d.err, 0, sizeof(int32_t), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return fail(DL_ERR_PLASTIC_ITER, "Too many iterations in plasticity loop");
  }
  if (e == DL_ERR_CAPACITY) {
    HIPCHK(hipMemsetAsync(c->d.err, 0, sizeof(int32_t), c->stream));
    if (c->contact) contact_invalidate(c->cs, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    return fail(DL_ERR_CAPACITY, "contact candidate lists exceed their capacity "
                                 "(dl_contact.max_pairs); the step results are invalid");
  }
  if (e) return fail(e, "device error");
  return DL_OK;
}

// the context's staging buffer for DOF-layout conversions (grown, never
// shrunk): GET_MATRIX2, set_state and get_state allocate nothing per call
static int stage_buffer(dl_ctx *c, size_t n_doubles, double **p) {
  if (c->stage_cap < n_doubles) {
    dfree(c, c->stage);
    c->stage = nullptr;
    c->stage_cap = 0;
    int rc = dalloc(c, &c->stage, n_doubles);
    if (rc) return rc;
    c->stage_cap = n_doubles;
  }
  *p = c->stage;
  return DL_OK;
}

// ------------------------------------------------------------------
// state and boundary data
// ------------------------------------------------------------------
extern "C" int dl_set_state(dl_ctx *c, const double *disp, const double *velo, int64_t ndof) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (ndof != 6 * c->d.N) return fail(DL_ERR_INPUT, "ndof must be 6 * n_nodes");
  c->adj.valid = false;
  hipSetDevice(c->device);
  double *tmp = nullptr;
  int rc = stage_buffer(c, 2 * ndof, &tmp);
  if (rc) return rc;
  hipError_t e1 = hipSuccess, e2 = hipSuccess;
  if (disp) e1 = hipMemcpyAsync(tmp, disp, ndof * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (velo) e2 = hipMemcpyAsync(tmp + ndof, velo, ndof * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e1 == hipSuccess && e2 == hipSuccess)
    launch_state_io(c->d, c->so3, false, disp ? tmp : nullptr, velo ? tmp + ndof : nullptr, c->stream);
  if (c->contact) contact_invalidate(c->cs, c->stream);
  // updateWhen: the next step assembles the mass at the new state
  if (c->mu.force) launch_set_i32(c->mu.force, 1, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return fail(DL_ERR_HIP, "dl_set_state copy failed");
  return DL_OK;
}

extern "C" int dl_get_state(dl_ctx *c, double *disp, double *velo, int64_t ndof) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (ndof != 6 * c->d.N) return fail(DL_ERR_INPUT, "ndof must be 6 * n_nodes");
  hipSetDevice(c->device);
  double *tmp = nullptr;
  int rc = stage_buffer(c, 2 * ndof, &tmp);
  if (rc) return rc;
  launch_state_io(c->d, c->so3, true, disp ? tmp : nullptr, velo ? tmp + ndof : nullptr, c->stream);
  hipError_t e1 = hipSuccess, e2 = hipSuccess;
  if (disp) e1 = hipMemcpyAsync(disp, tmp, ndof * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (velo) e2 = hipMemcpyAsync(velo, tmp + ndof, ndof * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return fail(DL_ERR_HIP, "dl_get_state copy failed");
  return DL_OK;
}

// (re)build the boundary record table from host masks; keeps reaction slots
static int upload_bcs(dl_ctx *c, const uint8_t *cmask, const double *cval, const double *fext,
                      const std::vector<int32_t> &react_nodes) {
  c->adj.valid = false;
  const int64_t N = c->d.N;
  std::vector<int32_t> bc(N, -1), rs;
  std::vector<uint32_t> mask;
  std::vector<double> a, f;
  std::vector<char> inreact(N, 0);
  for (int32_t n : react_nodes) inreact[n] = 1;
  const BcHost &hb = c->bch;
  std::vector<int32_t> impg(N, -1);
  for (size_t g = 0; g + 1 < hb.imp_ptr.size(); g++)
    for (int32_t j = hb.imp_ptr[g]; j < hb.imp_ptr[g + 1]; j++) impg[hb.imp_nodes[j]] = (int32_t)g;
  int32_t nbc = 0;
  for (int64_t n = 0; n < N; n++) {
    bool any = inreact[n] || impg[n] >= 0;
    for (int i = 0; i < 6 && !any; i++)
      any = (cmask && cmask[6 * n + i]) || (fext && fext[6 * n + i] != 0.0) ||
            (cval && cmask && cmask[6 * n + i] && cval[6 * n + i] != 0.0);
    if (any) bc[n] = nbc++;
  }
  mask.assign(std::max(nbc, 1), 0u);
  a.assign(6 * (size_t)std::max(nbc, 1), 0.0);
  f.assign(6 * (size_t)std::max(nbc, 1), 0.0);
  rs.assign(std::max(nbc, 1), -1);
  for (int64_t n = 0; n < N; n++) {
    const int32_t b = bc[n];
    if (b < 0) continue;
    uint32_t m = 0;
    for (int i = 0; i < 6; i++) {
      if (cmask && cmask[6 * n + i]) {
        m |= 1u << i;
        a[(size_t)i * nbc + b] = cval ? cval[6 * n + i] : 0.0;
      }
      f[(size_t)i * nbc + b] = fext ? fext[6 * n + i] : 0.0;
    }
    mask[b] = m;
  }
  for (size_t k = 0; k < react_nodes.size(); k++) rs[bc[react_nodes[k]]] = (int32_t)k;
  std::vector<int32_t> bimp;
  if (!hb.imp_ptr.empty()) {
    bimp.assign(std::max(nbc, 1), -1);
    for (int64_t n = 0; n < N; n++)
      if (impg[n] >= 0) {
        bimp[bc[n]] = impg[n];
        mask[bc[n]] |= 1u << hb.imp_dof[impg[n]];
      }
  }
  c->gen++;
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy((void *)c->d.bc, bc.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
  c->bc_host = bc;
  uint32_t *pm;
  double *pa, *pf;
  int32_t *pr;
  int rc;
  for (const void *q : {(const void *)c->d.bc_mask, (const void *)c->d.bc_a, (const void *)c->d.bc_f,
                        (const void *)c->d.bc_rslot, (const void *)c->d.bc_imp})
    dfree(c, (void *)q);
  c->d.bc_mask = nullptr;
  c->d.bc_a = c->d.bc_f = nullptr;
  c->d.bc_rslot = c->d.bc_imp = nullptr;
  if ((rc = upload(c, &pm, mask))) return rc;
  if ((rc = upload(c, &pa, a))) return rc;
  if ((rc = upload(c, &pf, f))) return rc;
  if ((rc = upload(c, &pr, rs))) return rc;
  c->d.bc_mask = pm;
  c->d.bc_a = pa;
  c->d.bc_f = pf;
  c->d.bc_rslot = pr;
  c->d.nbc = nbc;
  c->d.bc_imp = nullptr;
  if (!bimp.empty()) {
    int32_t *pb;
    if ((rc = upload(c, &pb, bimp))) return rc;
    c->d.bc_imp = pb;
  }
  c->imp.bc_a = pa;
  c->imp.nbc = nbc;
  return prog_entries(c);
}

static BcHost &bch(dl_ctx *c) { return c->bch; }

extern "C" int dl_set_bcs(dl_ctx *c, const uint8_t *cmask, const double *cval, const double *fext,
                          int64_t ndof) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (ndof != 6 * c->d.N) return fail(DL_ERR_INPUT, "ndof must be 6 * n_nodes");
  BcHost &h = bch(c);
  h.cmask.assign(cmask ? cmask : (const uint8_t *)nullptr, cmask ? cmask + ndof : nullptr);
  h.cval.assign(cval ? cval : (const double *)nullptr, cval ? cval + ndof : nullptr);
  h.fext.assign(fext ? fext : (const double *)nullptr, fext ? fext + ndof : nullptr);
  return upload_bcs(c, h.cmask.empty() ? nullptr : h.cmask.data(), h.cval.empty() ? nullptr : h.cval.data(),
                    h.fext.empty() ? nullptr : h.fext.data(), h.react);
}

extern "C" int dl_set_reaction_groups(dl_ctx *c, const int32_t *nodes, const int32_t *group_ptr,
                                      int32_t n_groups, int32_t dof) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (dof < 0 || dof > 5) return fail(DL_ERR_INPUT, "dof must be in [0, 5]");
  if (n_groups < 0 || (n_groups > 0 && (!group_ptr || group_ptr[0] != 0)))
    return fail(DL_ERR_INPUT, "group_ptr must start at 0");
  for (int32_t g = 0; g < n_groups; g++)
    if (group_ptr[g + 1] < group_ptr[g]) return fail(DL_ERR_INPUT, "group_ptr must be non-decreasing");
  const int64_t n = n_groups ? group_ptr[n_groups] : 0;
  BcHost &h = bch(c);
  h.react.assign(nodes, nodes + n);
  for (int32_t x : h.react)
    if (x < 0 || x >= c->d.N) return fail(DL_ERR_INPUT, "reaction node out of range");
  std::vector<int32_t> srt = h.react;
  std::sort(srt.begin(), srt.end());
  if (std::adjacent_find(srt.begin(), srt.end()) != srt.end())
    return fail(DL_ERR_INPUT, "a node may belong to one reaction group only");
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->d.react_dof = dof;
  c->n_react = n;
  c->n_groups = n_groups;
  std::vector<int32_t> gp(group_ptr, group_ptr + n_groups + 1);
  if (n_groups == 0) gp.assign(1, 0);
  c->rn_ptr_h = gp;
  // GroupOutput entries: every (element, end) at a node of group g, the
  // group's nodes in their order, a node's elements in id order
  const MeshHost &m = c->mesh;
  std::vector<int32_t> nptr(m.N + 1, 0), nlist(2 * m.E);
  for (int64_t e = 0; e < m.E; e++) {
    nptr[m.en[2 * e] + 1]++;
    nptr[m.en[2 * e + 1] + 1]++;
  }
  for (int64_t k = 0; k < m.N; k++) nptr[k + 1] += nptr[k];
  {
    std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t e = 0; e < m.E; e++)
      for (int end = 0; end < 2; end++) nlist[fill[m.en[2 * e + end]]++] = (int32_t)(2 * e + end);
  }
  std::vector<int32_t> rent, egp(n_groups + 1, 0);
  for (int32_t g = 0; g < n_groups; g++) {
    for (int32_t j = gp[g]; j < gp[g + 1]; j++) {
      const int32_t x = h.react[j];
      for (int32_t t = nptr[x]; t < nptr[x + 1]; t++) rent.push_back(nlist[t]);
    }
    if (rent.size() >= (size_t)INT32_MAX) return fail(DL_ERR_INPUT, "reaction groups too large");
    egp[g + 1] = (int32_t)rent.size();
  }
  if (n_groups == 0) egp.assign(1, 0);
  dfree(c, c->d.rbuf);
  dfree(c, (void *)c->d.rent);
  dfree(c, (void *)c->d.rslot_e);
  dfree(c, c->rgptr);
  c->d.rbuf = nullptr;
  c->d.rent = nullptr;
  c->d.rslot_e = nullptr;
  c->rgptr = nullptr;
  c->d.nrent = (int64_t)rent.size();
  int rc;
  int32_t *pe;
  if ((rc = upload(c, &pe, rent))) return rc;
  c->d.rent = pe;
  // the step writes the entries from k_strut: per element the entry of each
  // end (or -1), and flag bit 3 on the elements that have one
  {
    // the device flags (bit 4 is kept: the kernels set it)
    std::vector<uint8_t> fl(m.E);
    HIPCHK(hipMemcpy(fl.data(), c->d.eflag, m.E, hipMemcpyDeviceToHost));
    for (auto &x : fl) x &= (uint8_t)~8u;
    if (!rent.empty()) {
      std::vector<int32_t> rs(2 * (size_t)m.E, -1);
      for (size_t k = 0; k < rent.size(); k++) {
        const int32_t e = rent[k] >> 1, end = rent[k] & 1;
        rs[(size_t)end * m.E + e] = (int32_t)k;
        fl[e] |= 8u;
      }
      int32_t *pr;
      if ((rc = upload(c, &pr, rs))) return rc;
      c->d.rslot_e = pr;
    }
    HIPCHK(hipMemcpy(c->d.eflag, fl.data(), fl.size(), hipMemcpyHostToDevice));
  }
  double *rb;
  if ((rc = dalloc(c, &rb, (size_t)std::max<int64_t>(c->d.nrent, 1)))) return rc;
  c->d.rbuf = rb;
  HIPCHK(hipMemset(rb, 0, sizeof(double) * std::max<int64_t>(c->d.nrent, 1)));
  if ((rc = upload(c, &c->rgptr, egp))) return rc;
  c->rgptr_h = egp;
  c->hist_cap = 0; // history rows are n_groups wide: reallocate on the next step call
  c->gen++;
  return upload_bcs(c, h.cmask.empty() ? nullptr : h.cmask.data(), h.cval.empty() ? nullptr : h.cval.data(),
                    h.fext.empty() ? nullptr : h.fext.data(), h.react);
}

extern "C" int dl_set_impact_groups(dl_ctx *c, const int32_t *nodes, const int32_t *group_ptr,
                                    int32_t n_groups, const int32_t *dofs, const double *weights) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (n_groups < 0 || (n_groups > 0 && (!group_ptr || !nodes || !dofs || !weights || group_ptr[0] != 0)))
    return fail(DL_ERR_INPUT, "impact groups need nodes, group_ptr (from 0), dofs, weights");
  const int64_t N = c->d.N;
  BcHost &h = bch(c);
  std::vector<int32_t> seen(N, 0);
  for (int32_t g = 0; g < n_groups; g++) {
    if (group_ptr[g + 1] < group_ptr[g]) return fail(DL_ERR_INPUT, "group_ptr must be non-decreasing");
    if (dofs[g] < 0 || dofs[g] > 5) return fail(DL_ERR_INPUT, "impact dof must be in [0, 5]");
    if (!(weights[g] != 0.0) || !std::isfinite(weights[g]))
      return fail(DL_ERR_INPUT, "impact weights must be finite and nonzero");
    for (int32_t j = group_ptr[g]; j < group_ptr[g + 1]; j++) {
      const int32_t n = nodes[j];
      if (n < 0 || n >= N) return fail(DL_ERR_INPUT, "impact node out of range");
      if (seen[n]++) return fail(DL_ERR_INPUT, "a node may belong to one impact group only");
      if (!h.cmask.empty() && h.cmask[6 * (int64_t)n + dofs[g]])
        return fail(DL_ERR_INPUT, "impact dof already constrained by dl_set_bcs");
    }
  }
  if (n_groups > 0 && c->d.ndt) return fail(DL_ERR_STATE, "impact groups with per-node time steps");
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t nn = n_groups ? group_ptr[n_groups] : 0;
  h.imp_nodes.assign(nodes, nodes + nn);
  h.imp_ptr.clear();
  h.imp_dof.clear();
  h.imp_w.clear();
  if (n_groups > 0) {
    h.imp_ptr.assign(group_ptr, group_ptr + n_groups + 1);
    h.imp_dof.assign(dofs, dofs + n_groups);
    h.imp_w.assign(weights, weights + n_groups);
  }
  ImpDev &I = c->imp;
  for (void *q : {(void *)I.ptr, (void *)I.nodes, (void *)I.dof, (void *)I.w, (void *)I.R})
    dfree(c, q);
  I.ptr = I.nodes = I.dof = nullptr;
  I.w = nullptr;
  I.R = nullptr;
  I.ngroups = n_groups;
  int rc;
  if (n_groups > 0) {
    int32_t *pi;
    double *pd;
    if ((rc = upload(c, &pi, h.imp_ptr))) return rc;
    I.ptr = pi;
    if ((rc = upload(c, &pi, h.imp_nodes))) return rc;
    I.nodes = pi;
    if ((rc = upload(c, &pi, h.imp_dof))) return rc;
    I.dof = pi;
    if ((rc = upload(c, &pd, h.imp_w))) return rc;
    I.w = pd;
    if ((rc = dalloc(c, &I.R, n_groups))) return rc;
    if (!I.ready && (rc = dalloc(c, &I.ready, 1))) return rc;
    HIPCHK(hipMemset(I.ready, 0, sizeof(int32_t)));
  }
  c->adj.valid = false;
  return upload_bcs(c, h.cmask.empty() ? nullptr : h.cmask.data(), h.cval.empty() ? nullptr : h.cval.data(),
                    h.fext.empty() ? nullptr : h.fext.data(), h.react);
}

extern "C" int dl_set_reaction_group(dl_ctx *c, const int32_t *nodes, int64_t n, int32_t dof) {
  if (n < 0 || n >= INT32_MAX) return fail(DL_ERR_INPUT, "bad reaction group size");
  const int32_t gp[2] = {0, (int32_t)n};
  return dl_set_reaction_groups(c, nodes, gp, 1, dof);
}

extern "C" int dl_set_node_dt(dl_ctx *c, const double *dt, int64_t n_nodes) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (dt && c->pd.nprog > 0) return fail(DL_ERR_STATE, "per-node time steps with time programs");
  if (dt && c->imp.ngroups > 0) return fail(DL_ERR_STATE, "per-node time steps with impact groups");
  if (dt && c->mu.code) return fail(DL_ERR_STATE, "per-node time steps with a mass update schedule");
  c->adj.valid = false;
  if (!dt) {
    c->d.ndt = nullptr;
    c->gen++;
    return DL_OK;
  }
  if (n_nodes != c->d.N) return fail(DL_ERR_INPUT, "n_nodes mismatch");
  for (int64_t i = 0; i < n_nodes; i++)
    if (!(dt[i] > 0.0)) return fail(DL_ERR_INPUT, "node time steps must be > 0");
  if (!c->ndt) {
    int rc = dalloc(c, &c->ndt, n_nodes);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(c->ndt, dt, n_nodes * sizeof(double), hipMemcpyHostToDevice));
  c->d.ndt = c->ndt;
  c->gen++;
  return DL_OK;
}

extern "C" int dl_set_graph(dl_ctx *c, int32_t steps_per_graph) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (steps_per_graph < 0) return fail(DL_ERR_INPUT, "steps_per_graph must be >= 0");
  c->graph_steps = steps_per_graph;
  c->gen++;
  return DL_OK;
}

// ------------------------------------------------------------------
// actions
// ------------------------------------------------------------------
extern "C" int dl_take_action(dl_ctx *c, int32_t action, dl_params *p) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  c->adj.valid = false;
  hipSetDevice(c->device);
  const int64_t N = c->d.N, E = c->d.E;
  switch (action) {
  case DL_INIT: {
    // the blacklist's pointer and count live in the captured step graphs:
    // re-capture after every INIT
    c->gen++;
    launch_init_strain(c->d, c->stream);
    // ExplicitModule::init ends with updateMass_: the next step assembles
    if (c->mu.force) launch_set_i32(c->mu.force, 1, c->stream);
    // no output has run yet: ImpactBC dofs start free
    if (c->imp.ready) launch_set_i32(c->imp.ready, 0, c->stream);
    if (c->contact) contact_invalidate(c->cs, c->stream);
    HIPCHK(hipGetLastError());
    if (c->contact) {
      int rc = contact_init_blacklist(c->cs, c->d, c->stream);
      if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = true;
    return DL_OK;
  }
  case DL_GET_INT_VECTOR: {
    if (!c->initialized) return fail(DL_ERR_STATE, "GET_INT_VECTOR before INIT");
    if (!p || !p->int_vector) return fail(DL_ERR_INPUT, "params.int_vector is required");
    StepParams sp{0.0, p->load_case != DL_LOAD_CASE_OUTPUT, 0};
    if (c->contact) {
      int rc = contact_force(c->cs, c->d, c->stream);
      if (rc) return rc;
    }
    launch_strut_pp(c->d, sp, c->pp, MODE_FORCE, c->so3, c->mat_kind, c->contact, c->stream);
    launch_joint(c->d, sp, MODE_FORCE, c->so3, c->stream);
    launch_force_extras(c->d, c->so3, p->with_gyro != 0, c->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> h(6 * N);
    HIPCHK(hipMemcpyAsync(h.data(), c->d.fout, 6 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = check_err(c);
    if (rc) return rc;
    for (int64_t n = 0; n < N; n++)
      for (int i = 0; i < 6; i++) p->int_vector[6 * n + i] += h[i * N + n];
    return DL_OK;
  }
  case DL_GET_MATRIX2: {
    if (!p || !p->matrix2) return fail(DL_ERR_INPUT, "params.matrix2 is required");
    double *blk = nullptr;
    int rc = stage_buffer(c, 36 * N, &blk);
    if (rc) return rc;
    launch_mass_blocks(c->d, c->so3, p->lumped != 0, blk, c->stream);
    std::vector<double> h(36 * N);
    hipError_t e1 = hipMemcpyAsync(h.data(), blk, 36 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(DL_ERR_HIP, "GET_MATRIX2 copy failed");
    for (int64_t i = 0; i < 36 * N; i++) p->matrix2[i] += h[i];
    return DL_OK;
  }
  case DL_COMMIT: {
    if (!c->epot_e) {
      int rc;
      if ((rc = dalloc(c, &c->epot_e, E))) return rc;
      if ((rc = dalloc(c, &c->ediss_inc, E))) return rc;
      c->ediss_acc_h.assign(E, 0.0);
    }
    launch_commit_pl(c->d, c->plastic, c->epot_e, c->ediss_inc, c->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> ep(E), ed(E);
    HIPCHK(hipMemcpyAsync(ep.data(), c->epot_e, E * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ed.data(), c->ediss_inc, E * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> &A = c->ediss_acc_h;
    // E_pot / E_diss = sum_e sum_I w N_I energy_e  (SpecialCosseratRodModel.cpp:1074-1197)
    double sp = 0.0, sd = 0.0;
    for (int64_t e = 0; e < E; e++) {
      A[e] += ed[e];
      sp += c->w_e[e] * 0.5 * ep[e];
      sp += c->w_e[e] * 0.5 * ep[e];
      sd += c->w_e[e] * 0.5 * A[e];
      sd += c->w_e[e] * 0.5 * A[e];
    }
    if (p) {
      p->e_pot = sp;
      p->e_diss = sd;
    }
    return DL_OK;
  }
  case DL_CANCEL: {
    if (c->plastic) {
      HIPCHK(hipMemcpyAsync(c->d.pst_cur, c->d.pst, 12 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->d.aiso_cur, c->d.aiso, E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemsetAsync(c->d.dflow, 0, E * sizeof(double), c->stream));
    }
    HIPCHK(hipMemcpyAsync(c->d.ecur, c->d.eo, 6 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DL_OK;
  }
  default:
    return fail(DL_ERR_INPUT, "unknown action " + std::to_string(action));
  }
}

// ------------------------------------------------------------------
// explicit step
// ------------------------------------------------------------------
extern "C" int dl_set_timing(dl_ctx *c, int32_t enabled) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  c->timing = enabled != 0;
  return DL_OK;
}

extern "C" int dl_phase_time(dl_ctx *c, const char *phase, double *ms, int64_t *count) {
  if (!c || !phase) return fail(DL_ERR_INPUT, "NULL argument");
  const std::string ph(phase);
  double v;
  if (ph == "step") v = c->t_step;
  else if (ph == "contact") v = c->t_contact;
  else if (ph == "react") v = c->t_react;
  else if (ph == "refresh") v = c->t_refresh;
  else if (ph == "strut") v = c->t_strut;
  else if (ph == "joint") v = c->t_joint;
  else if (ph == "node") v = c->t_node;
  else return fail(DL_ERR_INPUT, "unknown phase " + ph);
  if (ms) *ms = v;
  if (count) *count = c->t_count;
  return DL_OK;
}

// events per timed step: start, after the candidate-list refresh, after
// contact, after the reaction entries, after the strut kernel, after the
// joint (and impact) kernels
#define STEP_EVENTS 6

// one explicit step: kernels only (capturable into a hipGraph)
int launch_one_step(dl_ctx *c, const StepParams &sp, hipEvent_t *ev) {
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  // updateWhen: the mass of the step (ExplicitModule::advance)
  launch_mass_update(c->d, c->mu, c->so3, c->stream);
  // boundary values at the end time of the step (after advance)
  launch_bc_program(c->pd, sp.dt, nullptr, c->stream);
  if (c->contact) {
    int rc = contact_step(c->cs, c->d, c->stream, c->hist_nc, c->hist_ncand, c->step_ctr,
                          ev ? ev[1] : nullptr);
    if (rc) return rc;
  }
  if (ev && !(c->contact && c->cs.c.skin > 0.0)) HIPCHK(hipEventRecord(ev[1], c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
  // the previous step's GroupOutput row, at this step's start state: written
  // by k_strut from the trial stress (see there)
  if (ev) HIPCHK(hipEventRecord(ev[3], c->stream));
  launch_strut_pp(c->d, sp, c->pp, MODE_STEP, c->so3, c->mat_kind, c->contact, c->stream);
  if (ev) HIPCHK(hipEventRecord(ev[4], c->stream));
  // the end-of-step sums ride in the joint kernel's launch (one block more,
  // or one per sum beyond 8) unless impact groups update nodes in between or
  // there are no shared nodes
  const bool fuse = c->d.NJ > 0 && c->imp.ngroups == 0;
  if (fuse) {
    const StepEnd se = make_step_end(c->d, c->rgptr, c->n_groups, c->track_diss, c->strut_blocks,
                                     c->diss_total, c->hist_react, c->hist_diss, c->step_ctr,
                                     c->end_done);
    launch_joint(c->d, sp, MODE_STEP, c->so3, c->stream, nullptr, &se);
    if (ev) HIPCHK(hipEventRecord(ev[5], c->stream));
  } else {
    launch_joint(c->d, sp, MODE_STEP, c->so3, c->stream);
    launch_impact_step(c->d, sp, c->imp, c->so3, c->stream);
    if (ev) HIPCHK(hipEventRecord(ev[5], c->stream));
    launch_step_end(c->d, c->rgptr, c->n_groups, c->track_diss, c->strut_blocks, c->diss_total,
                    c->hist_react, c->hist_diss, c->step_ctr, c->stream);
  }
  launch_clock_advance(c->pd, sp.dt, c->stream);
  return DL_OK;
}

int launch_reaction_tail(dl_ctx *c) {
  if (c->d.nrent == 0) return DL_OK;
  if (c->contact) {
    int rc = contact_step(c->cs, c->d, c->stream, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  launch_react(c->d, c->so3, c->plastic, c->contact, c->stream);
  launch_step_end(c->d, c->rgptr, c->n_groups, false, 0, nullptr, c->hist_react, c->hist_diss,
                  c->step_ctr, c->stream, true);
  return DL_OK;
}

int prepare_history(dl_ctx *c, int64_t nsteps) {
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  if (nsteps > c->hist_cap || !c->step_ctr) {
    int rc;
    const int64_t cap = std::max<int64_t>(nsteps, std::max<int64_t>(c->hist_cap, 16));
    dfree(c, c->hist_react);
    dfree(c, c->hist_diss);
    dfree(c, c->hist_nc);
    dfree(c, c->hist_ncand);
    if ((rc = dalloc(c, &c->hist_react, cap * ng))) return rc;
    if ((rc = dalloc(c, &c->hist_diss, cap))) return rc;
    if ((rc = dalloc(c, &c->hist_nc, cap))) return rc;
    if ((rc = dalloc(c, &c->hist_ncand, cap))) return rc;
    if (!c->step_ctr && (rc = dalloc(c, &c->step_ctr, 1))) return rc;
    c->hist_cap = cap;
    c->gen++;
  }
  HIPCHK(hipMemsetAsync(c->hist_react, 0, nsteps * ng * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(c->hist_diss, 0, nsteps * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(c->hist_nc, 0, nsteps * sizeof(int64_t), c->stream));
  HIPCHK(hipMemsetAsync(c->hist_ncand, 0, nsteps * sizeof(int64_t), c->stream));
  launch_set_i32(c->step_ctr, 0, c->stream);
  return DL_OK;
}

int advance_steps(dl_ctx *c, int64_t nsteps, const StepParams &sp) {
  int64_t done = 0;
  const int64_t K = c->graph_steps;
  const double dt = sp.dt;
  if (K > 0 && nsteps >= K) {
    if (!c->gexec || c->graph_gen != c->gen || c->graph_dt != dt) {
      if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
      }
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = DL_OK;
      for (int64_t k = 0; k < K && !rc; k++) rc = launch_one_step(c, sp, nullptr);
      hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (rc) return rc;
      if (ec != hipSuccess) return fail(DL_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
      ec = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (ec != hipSuccess) return fail(DL_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec));
      c->graph_gen = c->gen;
      c->graph_dt = dt;
    }
    for (; done + K <= nsteps; done += K) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  }
  for (; done < nsteps; done++) {
    int rc = launch_one_step(c, sp, nullptr);
    if (rc) return rc;
  }
  return DL_OK;
}

extern "C" int dl_explicit_step(dl_ctx *c, int64_t nsteps, double dt, dl_step_out *out) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->initialized) return fail(DL_ERR_STATE, "explicit step before INIT");
  if (nsteps < 0 || !(dt > 0.0)) return fail(DL_ERR_INPUT, "nsteps >= 0 and dt > 0 required");
  hipSetDevice(c->device);
  c->adj.valid = false;
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  {
    int rc = prepare_history(c, nsteps);
    if (rc) return rc;
  }
  StepParams sp{dt, 1, c->track_diss ? 1 : 0};
  int64_t builds0 = 0;
  if (out && c->contact) {
    int rc = contact_rebuilds(c->cs, c->stream, &builds0);
    if (rc) return rc;
  }
  if (c->timing) {
    const int K = STEP_EVENTS;
    std::vector<hipEvent_t> ev(K * (size_t)nsteps + 1);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    for (int64_t st = 0; st < nsteps; st++) {
      int rc = launch_one_step(c, sp, &ev[K * st]);
      if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ev[K * nsteps], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->t_step = c->t_contact = c->t_refresh = c->t_react = c->t_strut = c->t_joint = c->t_node = 0.0;
    for (int64_t st = 0; st < nsteps; st++) {
      float t[STEP_EVENTS], all;
      for (int i = 0; i < K; i++) HIPCHK(hipEventElapsedTime(&t[i], ev[K * st + i], ev[K * st + i + 1]));
      HIPCHK(hipEventElapsedTime(&all, ev[K * st], ev[K * st + K]));
      c->t_refresh += t[0];
      c->t_contact += t[0] + t[1];
      c->t_react += t[2];
      c->t_strut += t[3];
      c->t_joint += t[4];
      c->t_step += all;
    }
    c->t_count = nsteps;
    for (auto &e : ev) hipEventDestroy(e);
  } else {
    int rc = advance_steps(c, nsteps, sp);
    if (rc) return rc;
  }
  if (nsteps > 0) {
    int rc = launch_reaction_tail(c);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  c->hist_n = nsteps;
  if (out) {
    int rc = check_err(c);
    if (rc) return rc;
    out->steps_done = nsteps;
    out->n_rebuilds = 0;
    if (c->contact) {
      int64_t builds1 = 0;
      if ((rc = contact_rebuilds(c->cs, c->stream, &builds1))) return rc;
      out->n_rebuilds = builds1 - builds0;
    }
    out->reaction = 0.0;
    out->e_diss = 0.0;
    out->n_contacts = 0;
    out->n_candidates = 0;
    if (nsteps > 0) {
      HIPCHK(hipMemcpy(&out->reaction, c->hist_react + (nsteps - 1) * ng, sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(&out->n_contacts, c->hist_nc + nsteps - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(&out->n_candidates, c->hist_ncand + nsteps - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
    }
    HIPCHK(hipMemcpy(&out->e_diss, c->diss_total, sizeof(double), hipMemcpyDeviceToHost));
  }
  return DL_OK;
}

extern "C" int dl_explicit_step_noalloc(dl_ctx *c, int64_t nsteps, double dt, double *reaction,
                                        double *e_diss, int64_t *n_contacts) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->initialized) return fail(DL_ERR_STATE, "explicit step before INIT");
  if (nsteps < 0 || !(dt > 0.0)) return fail(DL_ERR_INPUT, "nsteps >= 0 and dt > 0 required");
  if (!c->scratch) return fail(DL_ERR_STATE, "no scratch address for this device");
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  // layout: [0] step counter, then reaction rows [nsteps][ng], dissipation
  // [nsteps], active pairs [nsteps] and candidates [nsteps] (int64)
  const int64_t need = 1 + nsteps * (ng + 3);
  if (need > DL_NOALLOC_SCRATCH_DOUBLES)
    return fail(DL_ERR_INPUT, "dl_explicit_step_noalloc: nsteps * (n_groups + 3) + 1 = " +
                                  std::to_string(need) + " exceeds the scratch of " +
                                  std::to_string(DL_NOALLOC_SCRATCH_DOUBLES) + " doubles");
  hipSetDevice(c->device);
  c->adj.valid = false;
  double *base = c->scratch;
  int32_t *ctr = (int32_t *)base;
  double *hr = base + 1, *hd = hr + nsteps * ng;
  int64_t *hn = (int64_t *)(hd + nsteps), *hc = hn + nsteps;
  // the array keeps the last call's rows: clear what this call reads back
  launch_zero_f64(base, need, c->stream);
  double *keep_r = c->hist_react, *keep_d = c->hist_diss;
  int64_t *keep_n = c->hist_nc, *keep_c = c->hist_ncand;
  int32_t *keep_s = c->step_ctr;
  c->hist_react = hr;
  c->hist_diss = hd;
  c->hist_nc = hn;
  c->hist_ncand = hc;
  c->step_ctr = ctr;
  StepParams sp{dt, 1, c->track_diss ? 1 : 0};
  int rc = DL_OK;
  for (int64_t st = 0; st < nsteps && !rc; st++) rc = launch_one_step(c, sp, nullptr);
  if (!rc && nsteps > 0) rc = launch_reaction_tail(c);
  c->hist_react = keep_r;
  c->hist_diss = keep_d;
  c->hist_nc = keep_n;
  c->hist_ncand = keep_c;
  c->step_ctr = keep_s;
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  if ((rc = check_err(c))) return rc;
  if (nsteps > 0) {
    if (reaction) HIPCHK(hipMemcpy(reaction, hr, nsteps * ng * sizeof(double), hipMemcpyDeviceToHost));
    if (e_diss) {
      if (c->track_diss) HIPCHK(hipMemcpy(e_diss, hd, nsteps * sizeof(double), hipMemcpyDeviceToHost));
      else std::fill(e_diss, e_diss + nsteps, 0.0);
    }
    if (n_contacts) HIPCHK(hipMemcpy(n_contacts, hn, nsteps * sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  return DL_OK;
}

extern "C" int dl_get_group_history(dl_ctx *c, double *reaction, int64_t cap_steps, int64_t *n) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  const int64_t k = std::min(cap_steps, c->hist_n);
  if (n) *n = c->hist_n;
  if (k > 0 && reaction)
    HIPCHK(hipMemcpy(reaction, c->hist_react, k * ng * sizeof(double), hipMemcpyDeviceToHost));
  return DL_OK;
}

extern "C" int dl_get_history(dl_ctx *c, double *reaction, double *e_diss, int64_t *n_contacts,
                              int64_t cap, int64_t *n) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t k = std::min(cap, c->hist_n);
  if (n) *n = c->hist_n;
  if (k <= 0) return DL_OK;
  if (reaction) {
    const int64_t ng = std::max<int32_t>(c->n_groups, 1);
    std::vector<double> h(k * ng);
    HIPCHK(hipMemcpy(h.data(), c->hist_react, k * ng * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < k; i++) reaction[i] = h[i * ng];
  }
  if (e_diss) {
    if (c->track_diss) HIPCHK(hipMemcpy(e_diss, c->hist_diss, k * sizeof(double), hipMemcpyDeviceToHost));
    else std::fill(e_diss, e_diss + k, 0.0);
  }
  if (n_contacts) HIPCHK(hipMemcpy(n_contacts, c->hist_nc, k * sizeof(int64_t), hipMemcpyDeviceToHost));
  return DL_OK;
}

extern "C" int dl_plastic_active(dl_ctx *c, uint8_t *mask, int64_t n) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (n != c->d.E) return fail(DL_ERR_INPUT, "n must be n_elems");
  if (!c->plastic) {
    std::fill(mask, mask + n, 0);
    return DL_OK;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(mask, c->d.active, n, hipMemcpyDeviceToHost));
  return DL_OK;
}

extern "C" int dl_plastic_state(dl_ctx *c, double *eps_p, double *alpha, int64_t n_elems) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  const int64_t E = c->d.E;
  if (n_elems != E) return fail(DL_ERR_INPUT, "n_elems mismatch");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->plastic) {
    if (eps_p) std::fill(eps_p, eps_p + 6 * E, 0.0);
    if (alpha) std::fill(alpha, alpha + 7 * E, 0.0);
    return DL_OK;
  }
  // element records: ep (6), alpha_1..6 (6); alpha_0 apart.  A record
  // whose eflag bit 4 is clear was never written and is zero.
  std::vector<double> h(12 * E), a0(E);
  HIPCHK(hipMemcpy(h.data(), c->d.pst, 12 * E * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(a0.data(), c->d.aiso, E * sizeof(double), hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < E; e++) {
    if (eps_p)
      for (int i = 0; i < 6; i++) eps_p[6 * e + i] = h[12 * e + i];
    if (alpha) {
      alpha[7 * e] = a0[e];
      for (int i = 0; i < 6; i++) alpha[7 * e + 1 + i] = h[12 * e + 6 + i];
    }
  }
  return DL_OK;
}

extern "C" int dl_set_plastic_state(dl_ctx *c, const double *eps_p, const double *alpha,
                                    int64_t n_elems) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  const int64_t E = c->d.E;
  if (n_elems != E) return fail(DL_ERR_INPUT, "n_elems mismatch");
  if (!c->plastic) return fail(DL_ERR_STATE, "elastic material has no plastic state");
  hipSetDevice(c->device);
  c->adj.valid = false;
  HIPCHK(hipStreamSynchronize(c->stream));
  // merge into the device records (a NULL part stays), then flag the
  // elements whose record is nonzero (eflag bit 4) and clear the others
  std::vector<double> h(12 * E), a0(E);
  std::vector<uint8_t> fl(E);
  HIPCHK(hipMemcpy(h.data(), c->d.pst, 12 * E * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(a0.data(), c->d.aiso, E * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(fl.data(), c->d.eflag, E, hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < E; e++) {
    if (eps_p)
      for (int i = 0; i < 6; i++) h[12 * e + i] = eps_p[6 * e + i];
    if (alpha) {
      a0[e] = alpha[7 * e];
      for (int i = 0; i < 6; i++) h[12 * e + 6 + i] = alpha[7 * e + 1 + i];
    }
    bool nz = a0[e] != 0.0;
    for (int i = 0; i < 12; i++) nz = nz || h[12 * e + i] != 0.0;
    fl[e] = nz ? (uint8_t)(fl[e] | 16u) : (uint8_t)(fl[e] & ~16u);
  }
  HIPCHK(hipMemcpy(c->d.pst, h.data(), 12 * E * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d.pst_cur, h.data(), 12 * E * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d.aiso, a0.data(), E * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d.aiso_cur, a0.data(), E * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d.eflag, fl.data(), E, hipMemcpyHostToDevice));
  return DL_OK;
}

extern "C" int dl_contact_pairs(dl_ctx *c, int32_t *a, int32_t *b, int64_t cap, int64_t *n) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->contact) {
    if (n) *n = 0;
    return DL_OK;
  }
  return contact_read_pairs(c->cs, c->stream, true, a, b, cap, n);
}

extern "C" int dl_contact_list_stats(dl_ctx *c, int64_t *entries, int64_t *builds) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (entries) *entries = 0;
  if (builds) *builds = 0;
  if (!c->contact) return DL_OK;
  hipSetDevice(c->device);
  int rc;
  if (entries && (rc = contact_list_entries(c->cs, c->stream, entries))) return rc;
  if (builds && (rc = contact_rebuilds(c->cs, c->stream, builds))) return rc;
  return DL_OK;
}

extern "C" int dl_contact_candidates(dl_ctx *c, int32_t *a, int32_t *b, int64_t cap, int64_t *n) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->contact) {
    if (n) *n = 0;
    return DL_OK;
  }
  return contact_read_pairs(c->cs, c->stream, false, a, b, cap, n);
}

int dl_set_error(int code, const char *msg) { return fail(code, msg); }
