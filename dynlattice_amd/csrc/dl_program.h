// dl_program.h -- time programs of the boundary data: scale factors of
// prescribed accelerations and external forces, evaluated on the device
// before every step / attempt.
//
// A program is one of
//   CONST  s = a
//   STEP   s = a + b (k + c)        (k: steps taken on the program clock)
//   EXPR   s = f(t, k)              compiled from an expression string
// The expression language is the subset of JEM's function syntax used by
// the reference's LoadScale scaleFunc strings (e.g. tests/transient/test1.pro:
// "if (t<15, 6/15 * (1 - cos(2*PI/15 * t)), 0)"): numbers, t, i (= k), PI,
// + - * / % ^, unary - + !, < <= > >= == !=, && ||, parentheses, and
// if(c,a,b) sin cos tan asin acos atan exp log sqrt abs floor ceil
// min(a,b) max(a,b) pow(a,b).  It is compiled on the host to a postfix
// program that this header evaluates with a fixed-size stack.
#ifndef DL_PROGRAM_H
#define DL_PROGRAM_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#ifndef DL_HD
#define DL_HD __host__ __device__ __forceinline__
#endif

enum {
  PO_CONST = 0, PO_T, PO_I,
  PO_ADD, PO_SUB, PO_MUL, PO_DIV, PO_MOD, PO_POW,
  PO_NEG, PO_NOT,
  PO_LT, PO_LE, PO_GT, PO_GE, PO_EQ, PO_NE, PO_AND, PO_OR,
  PO_IF,
  PO_SIN, PO_COS, PO_TAN, PO_ASIN, PO_ACOS, PO_ATAN, PO_EXP, PO_LOG, PO_SQRT, PO_ABS,
  PO_FLOOR, PO_CEIL, PO_MIN, PO_MAX,
  PO_COUNT
};

// one postfix instruction: op, and the value of PO_CONST
struct ProgOp {
  int32_t op;
  int32_t pad;
  double v;
};

#define DL_PROG_STACK 16 // deepest stack the compiler accepts
#define DL_PROG_MAXOPS 256
#define DL_PROG_MAXPROGS 64 // programs per context (the C ABI's DL_MAX_PROGRAMS)

enum { PK_CONST = 0, PK_STEP = 1, PK_EXPR = 2 };

// evaluates a compiled expression (the compiler guarantees a well-formed
// program of stack depth <= DL_PROG_STACK)
DL_HD double prog_eval_expr(const ProgOp *code, int32_t n, double t, double k) {
  double st[DL_PROG_STACK];
  int sp = 0;
  for (int32_t j = 0; j < n; j++) {
    const int32_t op = code[j].op;
    if (op == PO_CONST) { st[sp++] = code[j].v; continue; }
    if (op == PO_T) { st[sp++] = t; continue; }
    if (op == PO_I) { st[sp++] = k; continue; }
    if (op == PO_IF) {
      const double b = st[--sp], a = st[--sp], c = st[--sp];
      st[sp++] = (c != 0.0) ? a : b;
      continue;
    }
    if (op >= PO_ADD && op <= PO_POW) {
      const double y = st[--sp], x = st[--sp];
      double r;
      switch (op) {
      case PO_ADD: r = x + y; break;
      case PO_SUB: r = x - y; break;
      case PO_MUL: r = x * y; break;
      case PO_DIV: r = x / y; break;
      case PO_MOD: r = fmod(x, y); break;
      default: r = pow(x, y); break;
      }
      st[sp++] = r;
      continue;
    }
    if ((op >= PO_LT && op <= PO_OR) || op == PO_MIN || op == PO_MAX) {
      const double y = st[--sp], x = st[--sp];
      double r;
      switch (op) {
      case PO_LT: r = x < y; break;
      case PO_LE: r = x <= y; break;
      case PO_GT: r = x > y; break;
      case PO_GE: r = x >= y; break;
      case PO_EQ: r = x == y; break;
      case PO_NE: r = x != y; break;
      case PO_AND: r = (x != 0.0) && (y != 0.0); break;
      case PO_OR: r = (x != 0.0) || (y != 0.0); break;
      case PO_MIN: r = x < y ? x : y; break;
      default: r = x > y ? x : y; break;
      }
      st[sp++] = r;
      continue;
    }
    // unary
    const double x = st[--sp];
    double r;
    switch (op) {
    case PO_NEG: r = -x; break;
    case PO_NOT: r = (x == 0.0) ? 1.0 : 0.0; break;
    case PO_SIN: r = sin(x); break;
    case PO_COS: r = cos(x); break;
    case PO_TAN: r = tan(x); break;
    case PO_ASIN: r = asin(x); break;
    case PO_ACOS: r = acos(x); break;
    case PO_ATAN: r = atan(x); break;
    case PO_EXP: r = exp(x); break;
    case PO_LOG: r = log(x); break;
    case PO_SQRT: r = sqrt(x); break;
    case PO_ABS: r = fabs(x); break;
    case PO_FLOOR: r = floor(x); break;
    default: r = ceil(x); break;
    }
    st[sp++] = r;
  }
  return st[0];
}

// scale factor of program p at time t after k steps
DL_HD double prog_scale(int32_t kind, const double abc[3], const ProgOp *code, int32_t n, double t,
                        double k) {
  if (kind == PK_CONST) return abc[0];
  if (kind == PK_STEP) return abc[0] + abc[1] * (k + abc[2]);
  return prog_eval_expr(code, n, t, k);
}

#endif
