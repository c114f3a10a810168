// dl_program.cpp -- boundary-data time programs behind dl_set_programs /
// dl_set_time / dl_get_time (include/dl_gpu.h), and point masses
// (dl_set_point_masses).
//
// The reference drives its boundary data through model actions evaluated by
// the solver module every step: DirichletModel (INCREMENT: initDisp +
// k dispIncr, DirichletModel.cpp:163-172,290-306,400-406), NeumannModel
// (initLoad + (k+1) loadIncr, NeumannModel.cpp:118-147,347-368, or the
// parent's scale factor), and JIVE's LoadScale model (scaleFunc(t) applied
// as SCALE_FACTOR to its child), all read after ExplicitModule::advance has
// moved the time to t + dt (ExplicitModule.cpp:259-274, 397-419).  In
// dynamic runs the Dirichlet value is the prescribed acceleration.  Here
// every such value is (its dl_set_bcs factor) x (the scale of its program);
// the scales are computed on the device by one small kernel before every
// step / attempt, from a device-resident clock (t, k), so a captured graph
// replays them correctly.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dl_ctx.h"

static_assert(DL_MAX_PROGRAMS == DL_PROG_MAXPROGS, "program count limits differ");

namespace {

// recursive-descent compiler: expression string -> postfix program
struct Compiler {
  const char *s;
  size_t p = 0;
  std::vector<ProgOp> out;
  int depth = 0, maxdepth = 0;
  std::string err;

  explicit Compiler(const char *src) : s(src) {}

  void emit(int32_t op, double v = 0.0) {
    ProgOp o;
    o.op = op;
    o.pad = 0;
    o.v = v;
    out.push_back(o);
    if (op == PO_CONST || op == PO_T || op == PO_I) depth += 1;
    else if (op == PO_IF) depth -= 2;
    else if ((op >= PO_ADD && op <= PO_POW) || (op >= PO_LT && op <= PO_OR) || op == PO_MIN ||
             op == PO_MAX)
      depth -= 1;
    if (depth > maxdepth) maxdepth = depth;
  }
  void ws() {
    while (s[p] && std::isspace((unsigned char)s[p])) p++;
  }
  bool eat(const char *tok) {
    ws();
    const size_t n = std::strlen(tok);
    if (std::strncmp(s + p, tok, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  bool fail(const std::string &m) {
    if (err.empty()) err = m + " at position " + std::to_string(p);
    return false;
  }
  bool expr() { return orx(); }
  bool orx() {
    if (!andx()) return false;
    while (eat("||")) {
      if (!andx()) return false;
      emit(PO_OR);
    }
    return true;
  }
  bool andx() {
    if (!cmp()) return false;
    while (eat("&&")) {
      if (!cmp()) return false;
      emit(PO_AND);
    }
    return true;
  }
  bool cmp() {
    if (!add()) return false;
    static const struct { const char *t; int op; } ops[] = {
        {"<=", PO_LE}, {">=", PO_GE}, {"==", PO_EQ}, {"!=", PO_NE}, {"<", PO_LT}, {">", PO_GT}};
    for (const auto &o : ops)
      if (eat(o.t)) {
        if (!add()) return false;
        emit(o.op);
        return true;
      }
    return true;
  }
  bool add() {
    if (!mul()) return false;
    for (;;) {
      ws();
      if (s[p] == '+') { p++; if (!mul()) return false; emit(PO_ADD); }
      else if (s[p] == '-') { p++; if (!mul()) return false; emit(PO_SUB); }
      else return true;
    }
  }
  bool mul() {
    if (!unary()) return false;
    for (;;) {
      ws();
      if (s[p] == '*') { p++; if (!unary()) return false; emit(PO_MUL); }
      else if (s[p] == '/') { p++; if (!unary()) return false; emit(PO_DIV); }
      else if (s[p] == '%') { p++; if (!unary()) return false; emit(PO_MOD); }
      else return true;
    }
  }
  // unary minus binds looser than ^: -2^2 = -4, 2^-1 = 0.5
  bool unary() {
    ws();
    if (s[p] == '-') { p++; if (!unary()) return false; emit(PO_NEG); return true; }
    if (s[p] == '+') { p++; return unary(); }
    if (s[p] == '!' && s[p + 1] != '=') { p++; if (!unary()) return false; emit(PO_NOT); return true; }
    return power();
  }
  bool power() {
    if (!primary()) return false;
    ws();
    if (s[p] == '^') {
      p++;
      if (!unary()) return false;
      emit(PO_POW);
    }
    return true;
  }
  bool args(int n) {
    if (!eat("(")) return fail("'(' expected");
    for (int i = 0; i < n; i++) {
      if (i > 0 && !eat(",")) return fail("',' expected");
      if (!expr()) return false;
    }
    if (!eat(")")) return fail("')' expected");
    return true;
  }
  bool primary() {
    ws();
    const char c = s[p];
    if (std::isdigit((unsigned char)c) || c == '.') {
      char *end = nullptr;
      const double v = std::strtod(s + p, &end);
      if (end == s + p) return fail("number expected");
      p = (size_t)(end - s);
      emit(PO_CONST, v);
      return true;
    }
    if (c == '(') {
      p++;
      if (!expr()) return false;
      if (!eat(")")) return fail("')' expected");
      return true;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t q = p;
      while (std::isalnum((unsigned char)s[q]) || s[q] == '_') q++;
      const std::string id(s + p, q - p);
      p = q;
      if (id == "t") { emit(PO_T); return true; }
      if (id == "i") { emit(PO_I); return true; }
      if (id == "PI") { emit(PO_CONST, 3.14159265358979323846); return true; }
      if (id == "if") { if (!args(3)) return false; emit(PO_IF); return true; }
      static const struct { const char *n; int op; int nargs; } fns[] = {
          {"sin", PO_SIN, 1},   {"cos", PO_COS, 1},     {"tan", PO_TAN, 1},   {"asin", PO_ASIN, 1},
          {"acos", PO_ACOS, 1}, {"atan", PO_ATAN, 1},   {"exp", PO_EXP, 1},   {"log", PO_LOG, 1},
          {"sqrt", PO_SQRT, 1}, {"abs", PO_ABS, 1},     {"floor", PO_FLOOR, 1}, {"ceil", PO_CEIL, 1},
          {"min", PO_MIN, 2},   {"max", PO_MAX, 2},     {"pow", PO_POW, 2}};
      for (const auto &f : fns)
        if (id == f.n) {
          if (!args(f.nargs)) return false;
          emit(f.op);
          return true;
        }
      return fail("unknown name '" + id + "'");
    }
    return fail("unexpected character");
  }
  bool run() {
    if (!expr()) return false;
    ws();
    if (s[p]) return fail("trailing characters");
    if (maxdepth > DL_PROG_STACK) return fail("expression too deep");
    if (out.size() > DL_PROG_MAXOPS) return fail("expression too long");
    return true;
  }
};

} // namespace

// compiles one expression; message in err on failure
bool prog_compile(const char *src, std::vector<ProgOp> &code, std::string &err) {
  Compiler cp(src);
  if (!cp.run()) {
    err = cp.err;
    return false;
  }
  code = cp.out;
  return true;
}

// host-side evaluation of an expression string through the same compiler
// and evaluator as the device (for tests); returns 0 or -1 on a syntax error
extern "C" int dl_program_eval(const char *expr, double t, double k, double *value) {
  std::vector<ProgOp> code;
  std::string err;
  if (!expr || !prog_compile(expr, code, err)) return dl_set_error(DL_ERR_INPUT, err.c_str());
  if (value) *value = prog_eval_expr(code.data(), (int32_t)code.size(), t, k);
  return DL_OK;
}

// (re)builds the device list of program-driven boundary values from the
// current boundary records
int prog_entries(dl_ctx *c) {
  ProgState &S = c->prog;
  if (S.kind.empty()) {
    c->pd.nprog = 0;
    c->pd.nent = 0;
    return DL_OK;
  }
  const BcHost &h = c->bch;
  const int64_t N = c->d.N;
  std::vector<int32_t> rec, dof, pr;
  std::vector<double> base;
  for (int64_t n = 0; n < N; n++) {
    const int32_t b = c->bc_host.empty() ? -1 : c->bc_host[n];
    for (int i = 0; i < 6; i++) {
      const int32_t q = S.dof_prog[6 * n + i];
      if (q < 0 || b < 0) continue;
      const bool cons = !h.cmask.empty() && h.cmask[6 * n + i];
      const double v = cons ? (h.cval.empty() ? 0.0 : h.cval[6 * n + i])
                            : (h.fext.empty() ? 0.0 : h.fext[6 * n + i]);
      rec.push_back(b);
      dof.push_back(i | (cons ? 8 : 0));
      pr.push_back(q);
      base.push_back(v);
    }
  }
  ProgDev &P = c->pd;
  int rc;
  for (void *q : {(void *)P.ent_rec, (void *)P.ent_dof, (void *)P.ent_prog, (void *)P.ent_base})
    dfree(c, q);
  P.ent_rec = P.ent_dof = P.ent_prog = nullptr;
  P.ent_base = nullptr;
  int32_t *pi;
  double *pdv;
  if ((rc = upload(c, &pi, rec))) return rc;
  P.ent_rec = pi;
  if ((rc = upload(c, &pi, dof))) return rc;
  P.ent_dof = pi;
  if ((rc = upload(c, &pi, pr))) return rc;
  P.ent_prog = pi;
  if ((rc = upload(c, &pdv, base))) return rc;
  P.ent_base = pdv;
  P.nent = (int32_t)rec.size();
  P.nprog = (int32_t)S.kind.size();
  P.bc_a = (double *)c->d.bc_a;
  P.bc_f = (double *)c->d.bc_f;
  P.nbc = c->d.nbc;
  c->gen++;
  return DL_OK;
}

extern "C" int dl_set_programs(dl_ctx *c, const dl_program *progs, int32_t n_progs,
                               const int32_t *dof_prog, int64_t ndof) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (n_progs < 0 || n_progs > DL_MAX_PROGRAMS)
    return fail(DL_ERR_INPUT, "n_progs must be in [0, " + std::to_string(DL_MAX_PROGRAMS) + "]");
  if (n_progs > 0 && (!progs || !dof_prog || ndof != 6 * c->d.N))
    return fail(DL_ERR_INPUT, "programs need dof_prog[6 * n_nodes]");
  if (n_progs > 0 && c->d.ndt)
    return fail(DL_ERR_STATE, "time programs with per-node time steps");
  ProgState S;
  S.cptr.push_back(0);
  for (int32_t j = 0; j < n_progs; j++) {
    const dl_program &g = progs[j];
    if (g.kind != DL_PROG_CONST && g.kind != DL_PROG_STEP && g.kind != DL_PROG_EXPR)
      return fail(DL_ERR_INPUT, "program " + std::to_string(j) + ": unknown kind");
    S.kind.push_back(g.kind);
    S.abc.push_back(g.a);
    S.abc.push_back(g.b);
    S.abc.push_back(g.c);
    if (g.kind == DL_PROG_EXPR) {
      std::vector<ProgOp> code;
      std::string err;
      if (!g.expr || !prog_compile(g.expr, code, err))
        return fail(DL_ERR_INPUT, "program " + std::to_string(j) + ": " +
                                      (g.expr ? err : std::string("NULL expression")));
      S.code.insert(S.code.end(), code.begin(), code.end());
    }
    S.cptr.push_back((int32_t)S.code.size());
  }
  if (n_progs > 0) {
    S.dof_prog.assign(dof_prog, dof_prog + ndof);
    for (int32_t q : S.dof_prog)
      if (q < -1 || q >= n_progs) return fail(DL_ERR_INPUT, "dof_prog entry out of range");
  }
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->adj.valid = false;
  c->prog = std::move(S);
  ProgDev &P = c->pd;
  int rc;
  if (!P.clock_t) {
    if ((rc = dalloc(c, &P.clock_t, 1))) return rc;
    if ((rc = dalloc(c, &P.clock_k, 1))) return rc;
    HIPCHK(hipMemset(P.clock_t, 0, sizeof(double)));
    HIPCHK(hipMemset(P.clock_k, 0, sizeof(int64_t)));
  }
  for (void *q : {(void *)P.kind, (void *)P.cptr, (void *)P.abc, (void *)P.code}) dfree(c, q);
  P.kind = P.cptr = nullptr;
  P.abc = nullptr;
  P.code = nullptr;
  if (n_progs > 0) {
    int32_t *pi;
    double *pdv;
    ProgOp *po;
    if ((rc = upload(c, &pi, c->prog.kind))) return rc;
    P.kind = pi;
    if ((rc = upload(c, &pi, c->prog.cptr))) return rc;
    P.cptr = pi;
    if ((rc = upload(c, &pdv, c->prog.abc))) return rc;
    P.abc = pdv;
    std::vector<ProgOp> code = c->prog.code;
    if (code.empty()) code.resize(1);
    if ((rc = upload(c, &po, code))) return rc;
    P.code = po;
  }
  return prog_entries(c);
}

extern "C" int dl_set_mass_update(dl_ctx *c, const char *cond) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  // the inertia pointer and the condition live in captured step graphs
  c->gen++;
  c->adj.valid = false;
  MassUpd &M = c->mu;
  for (void *q : {(void *)M.code, (void *)M.force, (void *)M.Js}) dfree(c, q);
  M = MassUpd{};
  c->d.Js = nullptr;
  c->pd.clock_on = 0;
  if (!cond || !*cond) return DL_OK;
  // the schedule runs on the context's clock, which per-node time steps do
  // not define
  if (c->d.ndt) return fail(DL_ERR_STATE, "a mass update schedule with per-node time steps");
  std::vector<ProgOp> code;
  std::string err;
  if (!prog_compile(cond, code, err)) return fail(DL_ERR_INPUT, "updateWhen: " + err);
  ProgDev &P = c->pd;
  int rc;
  if (!P.clock_t) {
    if ((rc = dalloc(c, &P.clock_t, 1))) return rc;
    if ((rc = dalloc(c, &P.clock_k, 1))) return rc;
    HIPCHK(hipMemset(P.clock_t, 0, sizeof(double)));
    HIPCHK(hipMemset(P.clock_k, 0, sizeof(int64_t)));
  }
  ProgOp *po;
  if ((rc = upload(c, &po, code))) return rc;
  M.code = po;
  M.nops = (int32_t)code.size();
  int32_t *f;
  if ((rc = upload(c, &f, std::vector<int32_t>{1}))) return rc; // first step: assemble
  M.force = f;
  double *js;
  if ((rc = dalloc(c, &js, 6 * c->d.N))) return rc;
  M.Js = js;
  M.clock_t = P.clock_t;
  M.clock_k = P.clock_k;
  c->d.Js = js;
  P.clock_on = 1;
  return DL_OK;
}

extern "C" int dl_set_time(dl_ctx *c, double t, int64_t step) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!std::isfinite(t) || step < 0) return fail(DL_ERR_INPUT, "finite t and step >= 0 required");
  hipSetDevice(c->device);
  ProgDev &P = c->pd;
  int rc;
  if (!P.clock_t) {
    if ((rc = dalloc(c, &P.clock_t, 1))) return rc;
    if ((rc = dalloc(c, &P.clock_k, 1))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(P.clock_t, &t, sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.clock_k, &step, sizeof(int64_t), hipMemcpyHostToDevice));
  c->adj.valid = false;
  return DL_OK;
}

extern "C" int dl_get_time(dl_ctx *c, double *t, int64_t *step) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  double tv = 0.0;
  int64_t kv = 0;
  if (c->pd.clock_t) {
    HIPCHK(hipMemcpy(&tv, c->pd.clock_t, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&kv, c->pd.clock_k, sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  if (t) *t = tv;
  if (step) *step = kv;
  return DL_OK;
}
