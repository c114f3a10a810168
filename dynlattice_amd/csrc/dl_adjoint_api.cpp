// dl_adjoint_api.cpp -- checkpointed forward run and reverse sweep behind
// dl_adjoint_forward / dl_adjoint_backward (include/dl_gpu.h).
//
// Forward: the explicit step as usual (histories included), with the state
// (u, rotation, v) copied to a checkpoint every `every` steps.
// Backward: segments in reverse order; each segment is recomputed from its
// checkpoint with a tape (the start state and internal force of every node
// at every step of the segment), then reversed step by step:
//   k_adj_node (gather of the previous reversal's element ends + node VJP)
//   -> k_adj_elem -> [contact list refresh + pair VJPs].
// Parameter adjoints (per-element stiffness, per-node mass and inertia,
// per-element contact distance, per-element yield weights) accumulate over
// all steps and are mapped to the per-strut section data on the host,
// through the same set-up functions that built the device tables from the
// sections.  With the elastoplastic material the committed plastic state is
// checkpointed and taped too, and its adjoint travels back step by step
// beside the node state's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "dl_ctx.h"
#include "dl_setup.h"

// copies (u, rotation, v) between the context state and a [13][N] buffer
static int state_copy(dl_ctx *c, double *buf, bool save) {
  const int64_t N = c->d.N;
  const size_t b = sizeof(double);
  double *parts[3] = {c->d.u, c->d.rot, c->d.v};
  const int64_t off[3] = {0, 3 * N, 7 * N}, len[3] = {3 * N, 4 * N, 6 * N};
  for (int i = 0; i < 3; i++) {
    if (save)
      HIPCHK(hipMemcpyAsync(buf + off[i], parts[i], len[i] * b, hipMemcpyDeviceToDevice, c->stream));
    else
      HIPCHK(hipMemcpyAsync(parts[i], buf + off[i], len[i] * b, hipMemcpyDeviceToDevice, c->stream));
  }
  return DL_OK;
}

// copies the committed plastic state (the [E][12] records, then [E] alpha_0)
// between the context and a [13 E] buffer
static int plastic_copy(dl_ctx *c, double *buf, bool save) {
  const int64_t E = c->d.E;
  const size_t b = sizeof(double);
  if (save) {
    HIPCHK(hipMemcpyAsync(buf, c->d.pst, 12 * E * b, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(buf + 12 * E, c->d.aiso, E * b, hipMemcpyDeviceToDevice, c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(c->d.pst, buf, 12 * E * b, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.aiso, buf + 12 * E, E * b, hipMemcpyDeviceToDevice, c->stream));
  }
  return DL_OK;
}

// return_map's exponent kind for the context's yield family
static int plastic_expk(const dl_ctx *c) { return c->mat_kind == 3 ? 0 : c->mat_kind; }

static int adjoint_supported(dl_ctx *c) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->initialized) return fail(DL_ERR_STATE, "adjoint before INIT");
  if (!c->so3) return fail(DL_ERR_STATE, "the adjoint supports the SO(3) rotation update only");
  if (c->pd.nprog > 0) return fail(DL_ERR_STATE, "the adjoint does not support time programs");
  if (c->d.Js)
    return fail(DL_ERR_STATE, "the adjoint does not support an updateWhen mass schedule");
  if (c->imp.ngroups > 0) return fail(DL_ERR_STATE, "the adjoint does not support impact groups");
  if (c->n_mirrored > 0)
    return fail(DL_ERR_STATE, "the adjoint does not support -I strut frames (struts along -e3)");
  return DL_OK;
}

extern "C" int dl_adjoint_forward(dl_ctx *c, int64_t nsteps, double dt, int32_t every,
                                  double *reaction) {
  int rc = adjoint_supported(c);
  if (rc) return rc;
  if (nsteps < 1 || every < 1 || !(dt > 0.0))
    return fail(DL_ERR_INPUT, "nsteps >= 1, checkpoint_every >= 1 and dt > 0 required");
  hipSetDevice(c->device);
  AdjState &A = c->adj;
  A.valid = false;
  const int64_t N = c->d.N;
  const int64_t nseg = (nsteps + every - 1) / every;
  if (nseg > A.ckpt_slots) {
    dfree(c, A.ckpt);
    A.ckpt = nullptr;
    A.ckpt_slots = 0;
    if ((rc = dalloc(c, &A.ckpt, (size_t)nseg * 13 * N))) return rc;
    A.ckpt_slots = nseg;
  }
  if (!A.final_state && (rc = dalloc(c, &A.final_state, 13 * (size_t)N))) return rc;
  const int64_t E = c->d.E;
  if (c->plastic) {
    if (nseg > A.ckpt_pl_slots) {
      dfree(c, A.ckpt_pl);
      A.ckpt_pl = nullptr;
      A.ckpt_pl_slots = 0;
      if ((rc = dalloc(c, &A.ckpt_pl, (size_t)nseg * 13 * E))) return rc;
      A.ckpt_pl_slots = nseg;
    }
    if (!A.final_pl && (rc = dalloc(c, &A.final_pl, 13 * (size_t)E))) return rc;
    if (!A.final_active) {
      double *p = nullptr;
      if ((rc = dalloc(c, &p, (E + 7) / 8))) return rc;
      A.final_active = (uint8_t *)p;
    }
  }
  if ((rc = prepare_history(c, nsteps))) return rc;
  const bool lists = c->contact && c->cs.c.skin > 0.0;
  if (lists) {
    // log the steps that build the candidate lists (the kernels take the
    // log's pointer by value: captured step graphs are re-captured)
    if (A.blog_cap < nsteps) {
      dfree(c, A.blog);
      A.blog = nullptr;
      A.blog_cap = 0;
      double *p = nullptr;
      if ((rc = dalloc(c, &p, (nsteps + 7) / 8))) return rc;
      A.blog = (uint8_t *)p;
      A.blog_cap = nsteps;
    }
    HIPCHK(hipMemsetAsync(A.blog, 0, nsteps, c->stream));
    c->cs.c.blog = A.blog;
    c->cs.c.bstep = c->step_ctr;
    c->gen++;
  }
  const StepParams sp{dt, 1, 0};
  for (int64_t seg = 0; seg < nseg; seg++) {
    if ((rc = state_copy(c, A.ckpt + seg * 13 * N, true))) return rc;
    if (c->plastic && (rc = plastic_copy(c, A.ckpt_pl + seg * 13 * E, true))) return rc;
    // every segment starts with lists built at its checkpoint, as its
    // recomputation does
    if (lists) contact_invalidate(c->cs, c->stream);
    rc = advance_steps(c, std::min<int64_t>(every, nsteps - seg * every), sp);
    if (rc) {
      if (lists) {
        c->cs.c.blog = nullptr;
        c->gen++;
      }
      return rc;
    }
  }
  if (lists) {
    c->cs.c.blog = nullptr;
    c->gen++;
    A.blog_h.resize(nsteps);
    HIPCHK(hipMemcpyAsync(A.blog_h.data(), A.blog, nsteps, hipMemcpyDeviceToHost, c->stream));
  }
  if ((rc = launch_reaction_tail(c))) return rc;
  if ((rc = state_copy(c, A.final_state, true))) return rc;
  if (c->plastic) {
    if ((rc = plastic_copy(c, A.final_pl, true))) return rc;
    HIPCHK(hipMemcpyAsync(A.final_active, c->d.active, E, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipGetLastError());
  c->hist_n = nsteps;
  if ((rc = check_err(c))) return rc;
  A.valid = true;
  A.nsteps = nsteps;
  A.every = every;
  A.dt = dt;
  if (reaction) {
    const int64_t ng = std::max<int32_t>(c->n_groups, 1);
    HIPCHK(hipMemcpy(reaction, c->hist_react, nsteps * ng * sizeof(double), hipMemcpyDeviceToHost));
  }
  return DL_OK;
}

// reverse-sweep buffers and the node -> element-end map (once per context)
static int adjoint_buffers(dl_ctx *c) {
  AdjState &A = c->adj;
  const MeshHost &m = c->mesh;
  const int64_t N = m.N, E = m.E;
  int rc;
  if (!A.ready) {
    if ((rc = dalloc(c, &A.ubar, 3 * N))) return rc;
    if ((rc = dalloc(c, &A.qbar, 4 * N))) return rc;
    if ((rc = dalloc(c, &A.vbar, 6 * N))) return rc;
    if ((rc = dalloc(c, &A.fbar, 6 * N))) return rc;
    if ((rc = dalloc(c, &A.ebuf, 14 * E))) return rc;
    if ((rc = dalloc(c, &A.Cbar, 6 * E))) return rc;
    if ((rc = dalloc(c, &A.mtbar, N))) return rc;
    if ((rc = dalloc(c, &A.J0bar, 9 * N))) return rc;
    if ((rc = dalloc(c, &A.crbar, E))) return rc;
    std::vector<int32_t> nptr(N + 1, 0), nlist(2 * E);
    for (int64_t e = 0; e < E; e++) {
      nptr[m.en[2 * e] + 1]++;
      nptr[m.en[2 * e + 1] + 1]++;
    }
    for (int64_t n = 0; n < N; n++) nptr[n + 1] += nptr[n];
    std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t e = 0; e < E; e++)
      for (int end = 0; end < 2; end++) nlist[fill[m.en[2 * e + end]]++] = (int32_t)(2 * e + end);
    if ((rc = upload(c, &A.nptr, nptr))) return rc;
    if ((rc = upload(c, &A.nlist, nlist))) return rc;
    A.ready = true;
  }
  if (c->plastic && !A.pbar) {
    if ((rc = dalloc(c, &A.pbar, 13 * E))) return rc;
    if ((rc = dalloc(c, &A.Wbar, 6 * E))) return rc;
    if ((rc = dalloc(c, &A.rnode, N))) return rc;
    if ((rc = dalloc(c, &A.fstep_r, N))) return rc;
  }
  if (c->plastic && A.ptape_rows < A.every) {
    dfree(c, A.ptape);
    A.ptape = nullptr;
    A.ptape_rows = 0;
    if ((rc = dalloc(c, &A.ptape, (size_t)A.every * 13 * E))) return rc;
    A.ptape_rows = A.every;
  }
  // reaction slot -> group (the groups may have changed since the last call)
  std::vector<int32_t> sg(std::max<int64_t>(c->n_react, 1), 0);
  for (int32_t g = 0; g < c->n_groups; g++)
    for (int32_t k = c->rn_ptr_h[g]; k < c->rn_ptr_h[g + 1]; k++) sg[k] = g;
  dfree(c, A.slot_group);
  A.slot_group = nullptr;
  if ((rc = upload(c, &A.slot_group, sg))) return rc;
  if (A.tape_rows < A.every) {
    dfree(c, A.tape);
    A.tape = nullptr;
    A.tape_rows = 0;
    if ((rc = dalloc(c, &A.tape, (size_t)A.every * 19 * N))) return rc;
    A.tape_rows = A.every;
  }
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  if (A.rbar_cap < A.nsteps * ng) {
    dfree(c, A.rbar);
    A.rbar = nullptr;
    A.rbar_cap = 0;
    if ((rc = dalloc(c, &A.rbar, A.nsteps * ng))) return rc;
    A.rbar_cap = A.nsteps * ng;
  }
  return DL_OK;
}

// per-strut section gradients from the accumulated device adjoints
static int compose_gradients(dl_ctx *c, dl_section_grads *g) {
  const MeshHost &m = c->mesh;
  const SecHost &sc = c->sec;
  const AdjState &A = c->adj;
  const int64_t N = m.N, E = m.E;
  const int32_t R = m.R;
  std::vector<double> Cb(6 * E), mtb(N), J0b(9 * N), crb(E), Wb;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->plastic) {
    Wb.resize(6 * E);
    HIPCHK(hipMemcpy(Wb.data(), A.Wbar, Wb.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  HIPCHK(hipMemcpy(Cb.data(), A.Cbar, Cb.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(mtb.data(), A.mtbar, mtb.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(J0b.data(), A.J0bar, J0b.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(crb.data(), A.crbar, crb.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> gA(R, 0.0), gI1(R, 0.0), gI2(R, 0.0), gJ(R, 0.0), gk(R, 0.0), gY(R, 0.0),
      gG(R, 0.0), grho(R, 0.0), gcr(R, 0.0);
  // stiffness C = (G k A, G k A, Y A, Y I1, Y I2, G J)
  for (int32_t s = 0; s < R; s++) {
    double cb[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t e = m.eptr[s]; e < m.eptr[s + 1]; e++) {
      for (int i = 0; i < 6; i++) cb[i] += Cb[i * E + e];
      gcr[s] += crb[e];
    }
    const double G = sc.shear[s], Y = sc.young[s], k = sc.kappa[s], Ar = sc.area[s];
    const double shr = cb[0] + cb[1];
    gA[s] += G * k * shr + Y * cb[2];
    gI1[s] += Y * cb[3];
    gI2[s] += Y * cb[4];
    gJ[s] += G * cb[5];
    gY[s] += Ar * cb[2] + sc.I1[s] * cb[3] + sc.I2[s] * cb[4];
    gG[s] += k * Ar * shr + sc.J[s] * cb[5];
    gk[s] += G * Ar * shr;
  }
  // lumped mass (the reverse of add_elem_mass): mt += c_in rho A;
  // J0 += packed(L diag(Mr) L^T / 2), Mr = (c rho I1 l + rho A l^3/12,
  // c rho I2 l + rho A l^3/12, c rho J l)
  const double ic = sc.inertia_correct;
  double sh[2][2];
  mass_shapes(sh);
  walk_elements(m, [&](int32_t s, int64_t, int32_t a, int32_t b, const double *La, const double *Lb) {
    const double Jc = elem_jac(m.X.data(), a, b);
    const double wts = 1. * Jc;
    const double l = (wts + wts) / (2 - 1);
    const int32_t nodes2[2] = {a, b};
    double rhoA_bar = 0.0, Mrb[3] = {0.0, 0.0, 0.0};
    for (int in = 0; in < 2; in++) {
      double cm = 0.0;
      for (int jn = 0; jn < 2; jn++)
        for (int ip = 0; ip < 2; ip++) cm += wts * sh[in][ip] * sh[jn][ip];
      const int32_t n = nodes2[in];
      rhoA_bar += cm * mtb[n];
      const double *L = in == 0 ? La : Lb;
      double Jb[3][3];
      for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) Jb[r][q] = J0b[(3 * r + q) * N + n];
      // adjoints of the packed entries xx, yy, zz, xy, xz, yz and their (i, j)
      const double pb[6] = {Jb[0][0], Jb[1][1], Jb[2][2], Jb[0][1] + Jb[1][0], Jb[0][2] + Jb[2][0],
                            Jb[1][2] + Jb[2][1]};
      const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
      for (int t = 0; t < 6; t++)
        for (int k = 0; k < 3; k++) Mrb[k] += pb[t] * L[3 * pi[t] + k] * L[3 * pj[t] + k] / 2.;
    }
    const double rho = sc.rho[s], Ar = sc.area[s];
    const double l3 = std::pow(l, 3) / 12.;
    gI1[s] += rho * ic * l * Mrb[0];
    gI2[s] += rho * ic * l * Mrb[1];
    gJ[s] += rho * ic * l * Mrb[2];
    grho[s] += ic * l * (sc.I1[s] * Mrb[0] + sc.I2[s] * Mrb[1] + sc.J[s] * Mrb[2]);
    gA[s] += rho * (l3 * (Mrb[0] + Mrb[1]) + rhoA_bar);
    grho[s] += Ar * (l3 * (Mrb[0] + Mrb[1]) + rhoA_bar);
  });
  auto put = [&](double *dst, const std::vector<double> &src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  put(g->area, gA);
  put(g->I1, gI1);
  put(g->I2, gI2);
  put(g->J, gJ);
  put(g->shear_factor, gk);
  put(g->young, gY);
  put(g->shear_modulus, gG);
  put(g->density, grho);
  put(g->contact_radius, gcr);
  if (g->yield_weight) {
    // yield weights per strut, [R][6] like the material's table
    for (int32_t s = 0; s < R; s++)
      for (int i = 0; i < 6; i++) {
        double t = 0.0;
        if (c->plastic)
          for (int64_t e = m.eptr[s]; e < m.eptr[s + 1]; e++) t += Wb[i * E + e];
        g->yield_weight[6 * s + i] = t;
      }
  }
  return DL_OK;
}

extern "C" int dl_adjoint_backward(dl_ctx *c, const double *reaction_bar, dl_section_grads *grads) {
  int rc = adjoint_supported(c);
  if (rc) return rc;
  if (!reaction_bar || !grads) return fail(DL_ERR_INPUT, "reaction_bar and grads are required");
  AdjState &A = c->adj;
  if (!A.valid)
    return fail(DL_ERR_STATE, "no checkpointed forward run to reverse (dl_adjoint_forward first; "
                              "stepping, state, boundary or section changes discard it)");
  hipSetDevice(c->device);
  if ((rc = adjoint_buffers(c))) return rc;
  const int64_t N = c->d.N, E = c->d.E;
  const int64_t ng = std::max<int32_t>(c->n_groups, 1);
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(A.rbar, reaction_bar, A.nsteps * ng * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(A.ubar, 0, 3 * N * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.qbar, 0, 4 * N * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.vbar, 0, 6 * N * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.Cbar, 0, 6 * E * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.mtbar, 0, N * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.J0bar, 0, 9 * N * sizeof(double), s));
  HIPCHK(hipMemsetAsync(A.crbar, 0, E * sizeof(double), s));
  AdjDev a{};
  a.ubar = A.ubar; a.qbar = A.qbar; a.vbar = A.vbar; a.fbar = A.fbar; a.ebuf = A.ebuf;
  a.Cbar = A.Cbar; a.mtbar = A.mtbar; a.J0bar = A.J0bar; a.crbar = A.crbar;
  a.nptr = A.nptr; a.nlist = A.nlist; a.slot_group = A.slot_group;
  const bool pl = c->plastic;
  if (pl) {
    HIPCHK(hipMemsetAsync(A.Wbar, 0, 6 * E * sizeof(double), s));
    a.pbar = A.pbar; a.Wbar = A.Wbar; a.rnode = A.rnode; a.fstep_r = A.fstep_r;
  }
  // the element force's reverse for the context's material
  auto adj_elem = [&](const AdjDev &aa) {
    if (pl) launch_adj_elem_pl(c->d, aa, c->pp, plastic_expk(c), s);
    else launch_adj_elem(c->d, aa, s);
  };
  const StepParams sp{A.dt, 1, 0};
  const int64_t nseg = (A.nsteps + A.every - 1) / A.every;
  // Row t of the reaction history is the GroupOutput sum at the state after
  // step t, i.e. the internal force step t + 1 starts from; the last row is
  // the closing evaluation at the final state.  Its weight enters first:
  // the internal force's VJP at the final state (elements and contact) into
  // the element-end buffers, which the last step's node kernel gathers.
  a.rbar = A.rbar + (A.nsteps - 1) * ng;
  a.tape = A.final_state; // rows 0..6: u, rotation (the tape's layout)
  launch_adj_seed(c->d, a, s);
  if (pl) {
    // the closing row is the output load case at the final state: elastic
    // predictor from the final committed plastic state, which the seed
    // evaluation also gives its first adjoint (a.pbar)
    a.seed = 1;
    a.ptape = A.final_pl;
  }
  adj_elem(a);
  a.seed = 0;
  if (c->contact) {
    contact_invalidate(c->cs, s);
    DevData dn = c->d;
    dn.u = A.final_state;
    if ((rc = contact_adjoint(c->cs, dn, a, s))) return rc;
  }
  // candidate lists: the forward run's build log says at which steps they
  // were built (always at a segment's first step); the recomputation builds
  // them at the same steps from the same states, and the reverse sweep
  // rebuilds, from the taped state, the lists each step was evaluated with
  const bool lists = c->contact && c->cs.c.skin > 0.0 &&
                     (int64_t)A.blog_h.size() >= A.nsteps;
  for (int64_t seg = nseg - 1; seg >= 0; seg--) {
    const int64_t t0 = seg * A.every, len = std::min<int64_t>(A.every, A.nsteps - t0);
    if ((rc = state_copy(c, A.ckpt + seg * 13 * N, false))) return rc;
    if (pl && (rc = plastic_copy(c, A.ckpt_pl + seg * 13 * E, false))) return rc;
    if (c->contact && !lists) contact_invalidate(c->cs, s);
    // the build in effect at each step of the segment
    std::vector<int64_t> built(len, 0);
    if (lists)
      for (int64_t j = 1; j < len; j++) built[j] = A.blog_h[t0 + j] ? j : built[j - 1];
    // recompute the segment, taping every step's start state and forces
    // (and the committed plastic state the step starts from)
    for (int64_t j = 0; j < len; j++) {
      DevData dj = c->d;
      dj.tape = A.tape + j * 19 * N;
      if (pl && (rc = plastic_copy(c, A.ptape + j * 13 * E, true))) return rc;
      const int refresh = !lists ? REFRESH_GATED : (built[j] == j ? REFRESH_FORCE : REFRESH_SKIP);
      if (c->contact && (rc = contact_step(c->cs, dj, s, nullptr, nullptr, nullptr, nullptr,
                                           refresh)))
        return rc;
      launch_strut_pp(dj, sp, c->pp, MODE_STEP, true, c->mat_kind, c->contact, s);
      launch_joint(dj, sp, MODE_STEP, true, s);
    }
    int64_t cur = built[len - 1]; // the lists the recomputation left
    // reverse it; the node kernel first gathers the element ends of the
    // step reversed before (before the last step: the closing evaluation)
    for (int64_t j = len - 1; j >= 0; j--) {
      a.tape = A.tape + j * 19 * N;
      // the start-state force of step t is row t - 1's reaction
      a.rbar = (t0 + j >= 1) ? A.rbar + (t0 + j - 1) * ng : nullptr;
      if (pl) a.ptape = A.ptape + j * 13 * E;
      launch_adj_node(c->d, a, A.dt, true, s);
      adj_elem(a);
      if (c->contact) {
        DevData dn = c->d;
        dn.u = (double *)a.tape; // rows 0..2 of the tape: the step's start displacement
        if (lists && built[j] != cur) {
          DevData db = c->d;
          db.u = A.tape + built[j] * 19 * N;
          contact_rebuild(c->cs, db, s);
          cur = built[j];
        }
        if ((rc = contact_adjoint(c->cs, dn, a, s, lists ? REFRESH_SKIP : REFRESH_GATED)))
          return rc;
      }
    }
    HIPCHK(hipGetLastError());
  }
  if ((rc = state_copy(c, A.final_state, false))) return rc;
  if (pl) {
    if ((rc = plastic_copy(c, A.final_pl, false))) return rc;
    HIPCHK(hipMemcpyAsync(c->d.active, A.final_active, E, hipMemcpyDeviceToDevice, s));
  }
  if (c->contact) contact_invalidate(c->cs, s);
  if ((rc = check_err(c))) return rc;
  return compose_gradients(c, grads);
}
