// dl_tables_api.cpp -- dl_get_table / dl_get_energies (include/dl_gpu.h):
// element and node tables and energy totals at the current state, from the
// kernels in dl_tables.hip.  Element energies reach the nodes on the host in
// element order (half to each end: N_I = 1/2 at the one Gauss point), and
// totals are sums in index order, so repeated calls give the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "dl_ctx.h"

namespace {

// scratch device buffer of n doubles, freed on scope exit
struct DevBuf {
  double *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

int elem_values(dl_ctx *c, int kind, std::vector<double> *table, std::vector<double> *epot) {
  const int64_t E = c->d.E;
  DevBuf t, e;
  if (table) HIPCHK(hipMalloc((void **)&t.p, 6 * (size_t)std::max<int64_t>(E, 1) * sizeof(double)));
  if (epot) HIPCHK(hipMalloc((void **)&e.p, (size_t)std::max<int64_t>(E, 1) * sizeof(double)));
  launch_elem_table(c->d, c->pp, c->so3, c->mat_kind, kind, t.p, e.p, c->stream);
  HIPCHK(hipGetLastError());
  if (table) {
    table->resize(6 * E);
    HIPCHK(hipMemcpyAsync(table->data(), t.p, 6 * E * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  }
  if (epot) {
    epot->resize(E);
    HIPCHK(hipMemcpyAsync(epot->data(), e.p, E * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_err(c);
}

int node_ekin(dl_ctx *c, std::vector<double> &ek) {
  const int64_t N = c->d.N;
  DevBuf b;
  HIPCHK(hipMalloc((void **)&b.p, (size_t)std::max<int64_t>(N, 1) * sizeof(double)));
  launch_node_ekin(c->d, c->so3, b.p, c->stream);
  HIPCHK(hipGetLastError());
  ek.resize(N);
  HIPCHK(hipMemcpyAsync(ek.data(), b.p, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DL_OK;
}

int node_mass(dl_ctx *c, std::vector<double> &m) {
  const int64_t N = c->d.N;
  m.resize(N);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(m.data(), c->d.mt, N * sizeof(double), hipMemcpyDeviceToHost));
  return DL_OK;
}

} // namespace

extern "C" int dl_get_table(dl_ctx *c, const char *name, double *out, int64_t n_rows,
                            int32_t n_cols) {
  if (!c || !name || !out) return fail(DL_ERR_INPUT, "NULL argument");
  hipSetDevice(c->device);
  const std::string nm(name);
  const int64_t E = c->d.E, N = c->d.N;
  static const char *elem_tables[4] = {"strain", "mat_strain", "stress", "mat_stress"};
  for (int k = 0; k < 4; k++)
    if (nm == elem_tables[k]) {
      if (n_rows != E || n_cols != 6) return fail(DL_ERR_INPUT, nm + " is an [n_elems][6] table");
      std::vector<double> t;
      int rc = elem_values(c, k, &t, nullptr);
      if (rc) return rc;
      std::memcpy(out, t.data(), t.size() * sizeof(double));
      return DL_OK;
    }
  if (nm != "potentialEnergy" && nm != "kineticEnergy" && nm != "mass")
    return fail(DL_ERR_INPUT, "unknown table " + nm);
  if (n_rows != N || n_cols != 1) return fail(DL_ERR_INPUT, nm + " is an [n_nodes][1] table");
  std::vector<double> v;
  int rc;
  if (nm == "potentialEnergy") {
    std::vector<double> ep;
    if ((rc = elem_values(c, 0, nullptr, &ep))) return rc;
    v.assign(N, 0.0);
    const std::vector<int32_t> &en = c->mesh.en;
    for (int64_t e = 0; e < E; e++) {
      v[en[2 * e]] += 0.5 * ep[e];
      v[en[2 * e + 1]] += 0.5 * ep[e];
    }
  } else if (nm == "kineticEnergy") {
    if ((rc = node_ekin(c, v))) return rc;
  } else {
    if ((rc = node_mass(c, v))) return rc;
  }
  std::memcpy(out, v.data(), N * sizeof(double));
  return DL_OK;
}

extern "C" int dl_get_energies(dl_ctx *c, double *e_kin, double *e_pot, double *e_diss,
                               double *mass) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  hipSetDevice(c->device);
  int rc;
  if (e_kin) {
    std::vector<double> v;
    if ((rc = node_ekin(c, v))) return rc;
    double s = 0.0;
    for (double x : v) s += x;
    *e_kin = s;
  }
  if (e_pot) {
    std::vector<double> ep;
    if ((rc = elem_values(c, 0, nullptr, &ep))) return rc;
    double s = 0.0;
    for (double x : ep) s += x;
    *e_pot = s;
  }
  if (e_diss) {
    double s = 0.0;
    if (c->track_diss) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(&s, c->diss_total, sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int64_t e = 0; e < (int64_t)c->ediss_acc_h.size(); e++) s += c->w_e[e] * c->ediss_acc_h[e];
    *e_diss = s;
  }
  if (mass) {
    std::vector<double> m;
    if ((rc = node_mass(c, m))) return rc;
    double s = 0.0;
    for (double x : m) s += x;
    *mass = s;
  }
  return DL_OK;
}
