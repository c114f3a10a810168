// dl_contact.hip -- rod-rod penalty contact on gfx950.
//
// Per force evaluation:
//   kc_boxes     element AABBs (nodes +- own contact radius), largest extent
//                (integer atomic max of the bit patterns)
//   kc_cells     cell size h = max(h0, 1.001 * largest extent); integer cell of
//                each AABB's min corner (clamped to the grid), per-cell counts
//                with integer atomics
//   scan         exclusive scan of the counts: chunk-local scans and the
//                chunk offsets (2 launches); a cell's start is their sum
//   kc_scatter   element ids into their cells; kc_place writes each element's
//                record at its stable (id-ordered) position in the cell
//   kc_narrow    per element X: every Y in the 27 neighbour cells that the
//                reference would pair with X (different strut, no node of X in
//                Y's strut and vice versa, AABB overlap) -> closest points ->
//                STS / NTS branch -> penalty force on X's own two nodes.
// Two AABBs that overlap have min corners within h of each other, so their
// cells differ by at most one in every direction (clamping is 1-Lipschitz in
// the cell index, so it keeps that property).  The candidate set therefore
// equals the reference's all-pairs set (RodContactModel.cpp:316-350,388-440).
// Each pair is evaluated once from each side; each side accumulates only its
// own nodes' forces, in a fixed order, so no floating-point atomics are used
// and the result is bitwise deterministic.  The strut kernel adds the
// per-element force rows (flagged by one bit per element) to its nodal sums.
//
// The explicit step does not run this per step: it keeps per-element
// candidate lists built by the same broadphase with boxes inflated by a skin
// (kc_vbuild) and rebuilds them only when a node has moved by
// nearly the skin (kc_vmark evaluates each listed pair once for both its
// elements, kc_vsum sums each element's pairs in list order; see the list
// section below).
// The action path (GET_INT_VECTOR, pair listings, blacklist) and the adjoint
// without lists use the cell scan directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "dl_adjoint.h"
#include "dl_contact.h"
#include "dl_contact_math.h"

#define CB 256

enum { LIST_CAND = 0, LIST_ACTIVE = 1, LIST_CLOSED = 2 };

// ------------------------------------------------------------------
// A gated kernel does nothing unless *gate != 0 (null gate: always runs):
// the candidate-list rebuild is launched every step and runs only when the
// list has expired, which keeps the step sequence fixed for graph capture.
#define GATED(gate) \
  if ((gate) && *(gate) == 0) return

// The broadphase and list-build kernels loop over their range with a grid
// stride and are launched with at most GS_CAP blocks: full-chip occupancy
// when they work, and a cheap dispatch when they are gated off.
#define GS_CAP 2048
#define GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < (n); i += (int64_t)gridDim.x * CB)

// AABB of an element inflated by r (the element's contact radius, plus the
// list skin when building candidate lists), in one fixed operation order
__device__ __forceinline__ void elem_box(const DevData &d, int64_t N, int32_t a, int32_t b,
                                         double r, double lo[3], double hi[3]) {
  for (int i = 0; i < 3; i++) {
    const double xa = d.X0[i * N + a] + d.u[i * N + a];
    const double xb = d.X0[i * N + b] + d.u[i * N + b];
    lo[i] = (xa < xb ? xa : xb) - r;
    hi[i] = (xa > xb ? xa : xb) + r;
  }
}

__global__ __launch_bounds__(CB) void kc_boxes(ContactDev c, DevData d, double infl,
                                               const int32_t *gate) {
  GATED(gate);
  double ext = 0.0;
  GRID_STRIDE(e, c.E) {
    const int64_t N = c.N, E = c.E;
    double lo[3], hi[3];
    elem_box(d, N, c.en0[e], c.en1[e], c.crad[c.es[e]] + infl, lo, hi);
    for (int i = 0; i < 3; i++) {
      c.box[i * E + e] = lo[i];
      c.box[(3 + i) * E + e] = hi[i];
      ext = fmax(ext, hi[i] - lo[i]);
    }
  }
  __shared__ double red[CB];
  red[threadIdx.x] = ext;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  // largest extent over the blocks: the bit patterns of non-negative doubles
  // order like the values, so an integer max is exact and order-independent
  if (threadIdx.x == 0) atomicMax(c.extbits, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ __launch_bounds__(CB) void kc_fill_i32(int32_t *p, int64_t n, int32_t v,
                                                  const int32_t *gate) {
  GATED(gate);
  const int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ int32_t cell_coord(double x, double o, double h, int32_t dim) {
  double v = floor((x - o) / h);
  v = fmin(fmax(v, 0.0), (double)(dim - 1));
  return (int32_t)v;
}

// cell of every element (clamped cell of its box's min corner) and its rank
// among the cell's elements; ccount is all zero on entry (kc_scatter leaves
// it so)
__global__ __launch_bounds__(CB) void kc_cells(ContactDev c, const int32_t *gate) {
  GATED(gate);
  // cell size h = max(h0, 1.001 * largest extent)
  const double h = fmax(c.h0, 1.001 * __longlong_as_double((long long)*c.extbits));
  const int64_t E = c.E;
  GRID_STRIDE(e, E) {
    const int32_t cx = cell_coord(c.box[e], c.org[0], h, c.dim[0]);
    const int32_t cy = cell_coord(c.box[E + e], c.org[1], h, c.dim[1]);
    const int32_t cz = cell_coord(c.box[2 * E + e], c.org[2], h, c.dim[2]);
    const int32_t lin = (cz * c.dim[1] + cy) * c.dim[0] + cx;
    c.cell[e] = lin;
    c.rank[e] = atomicAdd(&c.ccount[lin], 1);
  }
}

// exclusive scan of n int32 values (chunks of 1024 per block)
__global__ __launch_bounds__(CB) void kc_scan1(const int32_t *in, int64_t n, int32_t *out,
                                               int32_t *bsum, const int32_t *gate) {
  GATED(gate);
  __shared__ int32_t s[CB];
  const int64_t nchunk = (n + 1023) / 1024;
  for (int64_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const int64_t base = ch * 1024 + 4 * threadIdx.x;
    int32_t v[4];
    int32_t t = 0;
    for (int k = 0; k < 4; k++) {
      v[k] = (base + k < n) ? in[base + k] : 0;
      t += v[k];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < CB; off <<= 1) {
      const int32_t x = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += x;
      __syncthreads();
    }
    int32_t run = s[threadIdx.x] - t;
    for (int k = 0; k < 4; k++) {
      if (base + k < n) out[base + k] = run;
      run += v[k];
    }
    if (threadIdx.x == CB - 1) bsum[ch] = s[CB - 1];
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void kc_scan2(int32_t *bsum, int64_t nb, int32_t *total,
                                                 const int32_t *gate) {
  GATED(gate);
  __shared__ int32_t s[1024];
  int32_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
    const int64_t i = b0 + threadIdx.x;
    const int32_t v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t x = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    const int32_t tot = s[1023];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(CB) void kc_scan3(int32_t *out, int64_t n, const int32_t *bsum,
                                               const int32_t *gate) {
  GATED(gate);
  GRID_STRIDE(i, n) out[i] += bsum[i / 1024];
}

// first record of cell k in cell order (k = ncell: the number of records)
__device__ __forceinline__ int32_t cell_start(const ContactDev &c, int64_t k) {
  return k == c.ncell ? c.cstart[k] : c.cstart[k] + c.cbsum[k >> 10];
}

// cell-ordered element record: AABB, element, strut, cell, joint bits of
// its nodes (ContactDev::ejoint; one 64-byte line)
struct __align__(64) CRec {
  double lo[3], hi[3];
  int32_t e, s, cell, jn;
};

// element ids into their cells (order within a cell set by the atomics)
// element ids into their cells; the counts are consumed by now and are
// reset to zero for the next broadphase
__global__ __launch_bounds__(CB) void kc_scatter(ContactDev c, const int32_t *gate) {
  GATED(gate);
  // the extent maximum has been used (kc_cells): reset it for the next
  // broadphase
  if (blockIdx.x == 0 && threadIdx.x == 0) *c.extbits = 0ull;
  GRID_STRIDE(e, c.E) {
    const int32_t cell = c.cell[e];
    c.tmp[cell_start(c, cell) + c.rank[e]] = (int32_t)e;
    c.ccount[cell] = 0;
  }
}

// stable placement: an element's slot in its cell is the number of elements
// of the same cell with a smaller id, so every cell lists its elements in id
// order whatever order the atomics gave; the 64-byte record goes there
__device__ __forceinline__ void place_one(const ContactDev &c, int64_t e) {
  const int64_t E = c.E;
  const int32_t cell = c.cell[e];
  const int32_t b = cell_start(c, cell), n = cell_start(c, cell + 1) - b;
  int32_t pos = b;
  for (int32_t j = 0; j < n; j++) pos += (c.tmp[b + j] < (int32_t)e) ? 1 : 0;
  CRec r;
  for (int i = 0; i < 3; i++) {
    r.lo[i] = c.box[i * E + e];
    r.hi[i] = c.box[(3 + i) * E + e];
  }
  r.e = (int32_t)e;
  r.s = c.es[e];
  r.cell = cell;
  r.jn = c.ejoint[e];
  ((CRec *)c.rec)[pos] = r;
  // compact prefilter record: bounds rounded outwards to float, strut id
  float4 lo, hi;
  lo.x = __double2float_rd(r.lo[0]); lo.y = __double2float_rd(r.lo[1]);
  lo.z = __double2float_rd(r.lo[2]); lo.w = __int_as_float(r.s);
  hi.x = __double2float_ru(r.hi[0]); hi.y = __double2float_ru(r.hi[1]);
  hi.z = __double2float_ru(r.hi[2]); hi.w = __int_as_float(r.e);
  ((float4 *)c.frec)[2 * pos] = lo;
  ((float4 *)c.frec)[2 * pos + 1] = hi;
}

// zero: a counter the list build that follows appends to (or null)
__global__ __launch_bounds__(CB) void kc_place(ContactDev c, const int32_t *gate, int32_t *zero) {
  GATED(gate);
  if (zero && blockIdx.x == 0 && threadIdx.x == 0) zero[0] = zero[1] = 0;
  GRID_STRIDE(e, c.E) place_one(c, e);
}

__device__ __forceinline__ bool node_in_strut(const ContactDev &c, int32_t n, int32_t s) {
  for (int32_t k = c.nsptr[n]; k < c.nsptr[n + 1]; k++)
    if (c.nslist[k] == s) return true;
  return false;
}

__device__ __forceinline__ bool blacklisted(const ContactDev &c, int32_t a, int32_t b) {
  if (c.nblack == 0) return false;
  const unsigned long long lo = (unsigned long long)(a < b ? a : b);
  const unsigned long long hi = (unsigned long long)(a < b ? b : a);
  const unsigned long long key = (lo << 32) | hi;
  int64_t l = 0, r = c.nblack - 1;
  while (l <= r) {
    const int64_t m = (l + r) >> 1;
    const unsigned long long v = c.black[m];
    if (v == key) return true;
    if (v < key) l = m + 1; else r = m - 1;
  }
  return false;
}

__device__ __forceinline__ void elem_pos(const DevData &d, int32_t n0, int32_t n1, double p[2][3]) {
  const int64_t N = d.N;
  for (int i = 0; i < 3; i++) {
    p[0][i] = d.X0[i * N + n0] + d.u[i * N + n0];
    p[1][i] = d.X0[i * N + n1] + d.u[i * N + n1];
  }
}

// Candidate scan of the element X of cell-order record k: every Y in the 27
// neighbour cells with a different strut, overlapping AABBs (float prefilter
// with outward-rounded bounds, then the exact double test) and no node of X
// in Y's strut nor of Y in X's strut.  fn(Y, sY, ya, yb, XisA) is called per
// candidate in cell order, three contiguous cells per row.
template <class Fn>
__device__ __forceinline__ void scan_candidates(const ContactDev &c, int64_t k, const CRec &x,
                                                Fn fn) {
  const CRec *rec = (const CRec *)c.rec;
  const int32_t X = x.e, sX = x.s;
  const int32_t xa = c.en0[X], xb = c.en1[X];
  const int32_t cx = x.cell % c.dim[0];
  const int32_t cy = (x.cell / c.dim[0]) % c.dim[1];
  const int32_t cz = x.cell / (c.dim[0] * c.dim[1]);
  const float4 xfl = ((const float4 *)c.frec)[2 * k], xfh = ((const float4 *)c.frec)[2 * k + 1];
  // X's side of the strut filter in registers, loaded once: the struts at
  // X's joint nodes and the joint nodes of X's strut (the lists of a BCC
  // joint and strut fit; longer ones are walked in memory).  A node of one
  // strut only is in no other strut, and Y's strut is never X's here, so
  // only joint nodes need the test.
  constexpr int KS = 8, KJ = 4;
  int32_t nla, nlb, nlj;
  int4 la0, la1, lb0, lb1, lj0;
  {
    const int32_t a0 = (x.jn & 1) ? c.nsptr[xa] : 0, a1 = (x.jn & 1) ? c.nsptr[xa + 1] : 0;
    const int32_t b0 = (x.jn & 2) ? c.nsptr[xb] : 0, b1 = (x.jn & 2) ? c.nsptr[xb + 1] : 0;
    const int32_t j0 = c.sjptr[sX], j1 = c.sjptr[sX + 1];
    nla = a1 - a0;
    nlb = b1 - b0;
    nlj = j1 - j0;
    auto ld = [&](const int32_t *l, int32_t b, int32_t n, int q) { return q < n ? l[b + q] : -1; };
    la0 = make_int4(ld(c.nslist, a0, nla, 0), ld(c.nslist, a0, nla, 1), ld(c.nslist, a0, nla, 2),
                    ld(c.nslist, a0, nla, 3));
    la1 = make_int4(ld(c.nslist, a0, nla, 4), ld(c.nslist, a0, nla, 5), ld(c.nslist, a0, nla, 6),
                    ld(c.nslist, a0, nla, 7));
    lb0 = make_int4(ld(c.nslist, b0, nlb, 0), ld(c.nslist, b0, nlb, 1), ld(c.nslist, b0, nlb, 2),
                    ld(c.nslist, b0, nlb, 3));
    lb1 = make_int4(ld(c.nslist, b0, nlb, 4), ld(c.nslist, b0, nlb, 5), ld(c.nslist, b0, nlb, 6),
                    ld(c.nslist, b0, nlb, 7));
    lj0 = make_int4(ld(c.sjlist, j0, nlj, 0), ld(c.sjlist, j0, nlj, 1), ld(c.sjlist, j0, nlj, 2),
                    ld(c.sjlist, j0, nlj, 3));
  }
  auto eq4 = [](const int4 a, int32_t v) { return (a.x == v) | (a.y == v) | (a.z == v) | (a.w == v); };
  for (int dz = -1; dz <= 1; dz++) {
    const int32_t z = cz + dz;
    if (z < 0 || z >= c.dim[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int32_t y = cy + dy;
      if (y < 0 || y >= c.dim[1]) continue;
      // the three cells x-1..x+1 of a row are contiguous in cell order
      const int32_t x0 = cx > 0 ? cx - 1 : 0;
      const int32_t x1 = cx + 1 < c.dim[0] ? cx + 1 : c.dim[0] - 1;
      const int32_t row = (z * c.dim[1] + y) * c.dim[0];
      const int32_t j1 = cell_start(c, row + x1 + 1);
      const float4 *fr = (const float4 *)c.frec;
      for (int32_t j0 = cell_start(c, row + x0); j0 < j1; j0 += 4) {
        // four independent prefilter loads in flight
        float4 flo[4], fhi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int32_t jj = (j0 + q < j1) ? j0 + q : j1 - 1;
          flo[q] = fr[2 * jj];
          fhi[q] = fr[2 * jj + 1];
        }
        unsigned hit = 0u;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const bool ok = (j0 + q < j1) && (__float_as_int(flo[q].w) != sX) &&
                          (xfh.x >= flo[q].x) && (fhi[q].x >= xfl.x) &&
                          (xfh.y >= flo[q].y) && (fhi[q].y >= xfl.y) &&
                          (xfh.z >= flo[q].z) && (fhi[q].z >= xfl.z);
          hit |= ok ? (1u << q) : 0u;
        }
        while (hit) {
          const int q = __builtin_ctz(hit);
          hit &= hit - 1u;
          // Y and its strut from the prefilter record, so that Y's nodes
          // load together with its full record
          int32_t Y = __float_as_int(fhi[0].w), sY = __float_as_int(flo[0].w);
#pragma unroll
          for (int p = 1; p < 4; p++) { // select, not a register-array index
            Y = q == p ? __float_as_int(fhi[p].w) : Y;
            sY = q == p ? __float_as_int(flo[p].w) : sY;
          }
          const int32_t ya = c.en0[Y], yb = c.en1[Y];
          const CRec yr = rec[j0 + q];
          bool ov = true;
          for (int i = 0; i < 3; i++) ov = ov && (x.hi[i] >= yr.lo[i]) && (yr.hi[i] >= x.lo[i]);
          if (!ov) continue;
          // the strut filter
          bool ex = false;
          if (x.jn & 1) ex = nla > KS ? node_in_strut(c, xa, sY) : (eq4(la0, sY) | eq4(la1, sY));
          if (!ex && (x.jn & 2))
            ex = nlb > KS ? node_in_strut(c, xb, sY) : (eq4(lb0, sY) | eq4(lb1, sY));
          if (!ex && (yr.jn & 1)) ex = nlj > KJ ? node_in_strut(c, ya, sX) : eq4(lj0, ya);
          if (!ex && (yr.jn & 2)) ex = nlj > KJ ? node_in_strut(c, yb, sX) : eq4(lj0, yb);
          if (ex) continue;
          fn(Y, sY, ya, yb, sX < sY);
        }
      }
    }
  }
}

// One thread per element, in cell order (thread k owns the element of
// record k).  MODE 0: contact forces on the element's own nodes (+ counts of
// active pairs and candidates, each pair counted on its A side).  MODE 1:
// count the listed pairs of `kind` (A side) into lcount[k].  MODE 2: emit
// them at loff[k].
template <int MODE>
__global__ __launch_bounds__(CB) void kc_narrow(ContactDev c, DevData d, int kind,
                                                unsigned long long *cnt_act,
                                                unsigned long long *cnt_cand,
                                                const int32_t *step) {
  const int64_t k = (int64_t)blockIdx.x * CB + threadIdx.x;
  unsigned long long nact = 0, ncand = 0;
  if (k < c.E) {
    const CRec x = ((const CRec *)c.rec)[k];
    const int32_t X = x.e, sX = x.s;
    double pX[2][3];
    elem_pos(d, c.en0[X], c.en1[X], pX);
    double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0};
    bool any = false;
    int32_t nlist = 0;
    const int32_t off = (MODE == 2) ? c.loff[k] : 0;
    scan_candidates(c, k, x, [&](int32_t Y, int32_t sY, int32_t ya, int32_t yb, bool XisA) {
      if (XisA) ncand++;
      if (MODE != 0 && kind == LIST_CAND) {
        if (XisA) {
          if (MODE == 2) { c.la[off + nlist] = X; c.lb[off + nlist] = Y; }
          nlist++;
        }
        return;
      }
      if (MODE != 0 && !XisA) return;
      if (kind != LIST_CLOSED && blacklisted(c, XisA ? X : Y, XisA ? Y : X)) return;
      double pY[2][3];
      elem_pos(d, ya, yb, pY);
      const double R = c.crad[sX] + c.crad[sY];
      if (MODE == 0) {
        // this pair's contribution summed from zero, then added
        double g0[3] = {0.0, 0.0, 0.0}, g1[3] = {0.0, 0.0, 0.0};
        const bool closed = XisA ? cm_narrowphase<true, 0>(pX, pY, R, c.kS, c.kN, g0, g1)
                                 : cm_narrowphase<true, 1>(pY, pX, R, c.kS, c.kN, g0, g1);
        if (closed) {
          any = true;
          for (int i = 0; i < 3; i++) {
            fa[i] += g0[i];
            fb[i] += g1[i];
          }
          if (XisA) nact++;
        }
      } else {
        double g0[3], g1[3];
        if (cm_narrowphase<false, 0>(pX, pY, R, c.kS, c.kN, g0, g1)) {
          if (MODE == 2) { c.la[off + nlist] = X; c.lb[off + nlist] = Y; }
          nlist++;
        }
      }
    });
    if (MODE == 0 && any) {
      // the element's force row and its bit (cleared by kc_clear_bits)
      const int64_t E = c.E;
      for (int i = 0; i < 3; i++) {
        c.fcs[i * E + X] = fa[i];
        c.fcs[(3 + i) * E + X] = fb[i];
      }
      atomicOr(c.cbits + (X >> 6), 1ull << (X & 63));
    }
    if (MODE == 1) c.lcount[k] = nlist;
  }
  if (MODE == 0) {
    __shared__ unsigned long long ra[CB], rc[CB];
    ra[threadIdx.x] = nact;
    rc[threadIdx.x] = ncand;
    __syncthreads();
    for (int w = CB / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        ra[threadIdx.x] += ra[threadIdx.x + w];
        rc[threadIdx.x] += rc[threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0 && cnt_act) {
      // history row of the current step (device counter: graph-replay safe)
      const int64_t st = *step;
      if (ra[0]) atomicAdd(cnt_act + st, ra[0]);
      if (rc[0]) atomicAdd(cnt_cand + st, rc[0]);
    }
  }
}

// Adjoint of the contact forces of one step, one thread per element X in
// cell order: for every closed pair, the pair VJP from X's side (adjoints of
// both elements' nodal forces) adds to X's element-end displacement adjoints
// and to X's contact-distance adjoint.  Each element is owned by one thread,
// so the additions are plain and ordered.
__global__ __launch_bounds__(CB) void kc_adj_narrow(ContactDev c, DevData d, AdjDev a) {
  const int64_t k = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (k >= c.E) return;
  const int64_t N = c.N, E = c.E;
  const CRec x = ((const CRec *)c.rec)[k];
  const int32_t X = x.e, sX = x.s;
  const int32_t xa = c.en0[X], xb = c.en1[X];
  double pX[2][3], fbX[6];
  elem_pos(d, xa, xb, pX);
  for (int i = 0; i < 3; i++) {
    fbX[i] = a.fbar[i * N + xa];
    fbX[3 + i] = a.fbar[i * N + xb];
  }
  double g[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool any = false;
  scan_candidates(c, k, x, [&](int32_t Y, int32_t sY, int32_t ya, int32_t yb, bool XisA) {
    if (blacklisted(c, XisA ? X : Y, XisA ? Y : X)) return;
    double pY[2][3], fbY[6], o[7];
    elem_pos(d, ya, yb, pY);
    for (int i = 0; i < 3; i++) {
      fbY[i] = a.fbar[i * N + ya];
      fbY[3 + i] = a.fbar[i * N + yb];
    }
    if (contact_pair_vjp(pX, pY, XisA, c.crad[sX] + c.crad[sY], c.kS, c.kN, fbX, fbY, o)) {
      any = true;
      for (int i = 0; i < 7; i++) g[i] += o[i];
    }
  });
  if (!any) return;
  for (int i = 0; i < 3; i++) {
    a.ebuf[i * E + X] += g[i];
    a.ebuf[(7 + i) * E + X] += g[3 + i];
  }
  a.crbar[X] += g[6];
}

// ------------------------------------------------------------------
// candidate lists of the step (skin lists)
//
// Built from the broadphase with boxes inflated by the skin s: every
// partner Y of X whose inflated box overlaps X's (plus the static filters:
// other strut, no node in the other strut; pairs blacklisted at INIT are
// kept but marked, so that they count as candidates as in the cell scan and
// are never evaluated).  While
// no node has moved by more than the trigger distance (< s) since the build,
// two boxes that overlap now overlapped inflated by s at the build, so the
// list holds every current candidate; the step then tests only the listed
// pairs with the exact boxes, i.e. it finds exactly the broadphase's
// candidate set.  Lists are sorted by partner id, so the force on X is
// summed in an order that does not depend on the cell layout or on when the
// lists were built.
// ------------------------------------------------------------------
// A blacklisted partner stays in the list, stored as ~Y (negative): it is
// counted as a candidate like in the cell scan but never evaluated.
__device__ __forceinline__ int32_t list_partner(int32_t v) { return v < 0 ? ~v : v; }

__global__ __launch_bounds__(CB) void kc_vbuild(ContactDev c, DevData d, const int32_t *gate) {
  GATED(gate);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(c.nbuild, 1ull);
    if (c.blog) c.blog[*c.bstep] = 1;
  }
  GRID_STRIDE(k, c.E) {
    const CRec x = ((const CRec *)c.rec)[k];
    const int32_t X = x.e;
    int32_t n = 0;
    int32_t *lst = c.vlist + (int64_t)X * c.vk;
    const int32_t cap = c.vk;
    scan_candidates(c, k, x, [&](int32_t Y, int32_t, int32_t, int32_t, bool XisA) {
      const bool bl = blacklisted(c, XisA ? X : Y, XisA ? Y : X);
      if (n < cap) {
        // insertion by partner id
        int32_t i = n;
        while (i > 0 && list_partner(lst[i - 1]) > Y) {
          lst[i] = lst[i - 1];
          i--;
        }
        lst[i] = bl ? ~Y : Y;
      }
      n++;
    });
    if (n > cap) *c.err = DL_ERR_CAPACITY;
    c.vcount[X] = n < cap ? n : cap;
    if (n > 0) {
      const int32_t m = n < cap ? n : cap;
      c.vbase[X] = atomicAdd(c.nentries, m); // the element's fpair rows
      // the pair records of kc_vmark, one per listed pair on its A side (the
      // element of the lower strut id; the order of the records is
      // irrelevant): the entry's bit, B, both struts and both elements' nodes
      const int32_t sX = x.s;
      int32_t na = 0;
      for (int32_t t = 0; t < m; t++) na += sX < c.es[list_partner(lst[t])] ? 1 : 0;
      if (na > 0) {
        int64_t q = atomicAdd(c.nentries + 1, na);
        const int32_t xa = c.en0[X], xb = c.en1[X];
        for (int32_t t = 0; t < m; t++) {
          const int32_t code = lst[t], Y = list_partner(code), sY = c.es[Y];
          if (!(sX < sY)) continue;
          const int32_t bit = (int32_t)((int64_t)X * c.vkb + t);
          int4 *r = (int4 *)(c.vrec + 8 * q++);
          r[0] = make_int4(code < 0 ? ~bit : bit, Y, sX, sY);
          r[1] = make_int4(xa, xb, c.en0[Y], c.en1[Y]);
        }
      }
    }
  }
  // the lists are built: remember the displacement they were built at (no
  // kernel here reads uref; the node updates that follow compare against it)
  GRID_STRIDE(n, c.N) {
    for (int i = 0; i < 3; i++) c.uref[i * c.N + n] = d.u[i * c.N + n];
  }
}

// ------------------------------------------------------------------
// The step's pair pass from the lists, in two kernels.
//   kc_vmark  parallel over listed pairs (the records of kc_vbuild, one per
//             pair, each pair listed on both its elements): the exact boxes
//             of A and B, the candidate and active counts, and for every
//             closed pair one evaluation that gives its force on A's two
//             nodes and on B's, stored in the fpair rows of both entries,
//             with both entries' bits in vmark and both elements' bits in
//             cnew.  A thread takes one pair at a time, so the work follows
//             the number of pairs and not the longest list, and a light load
//             spreads over as many wavefronts as it has pairs.
//   kc_vsum   per element with a closed pair (4 elements per lane, the
//             wavefront owning 256 elements = 4 mask words): the stored
//             forces of its closed pairs summed in list order, i.e. by
//             partner id -- the operations and the order of one serial pass
//             over the list, so the force does not depend on the cell layout,
//             on when the lists were built or on the split into two kernels.  It
//             writes the element's force row, stores the wavefront's contact
//             mask words whole (k_strut reads them) and clears the pair and
//             element bits it used, so both arrays are zero between steps.
// ------------------------------------------------------------------
// the force of a closed entry on its element's nodes, stored in the entry's
// fpair row, and the entry's bit and its element's bit
__device__ __forceinline__ void vstore(const ContactDev &c, int32_t bit, const double g0[3],
                                       const double g1[3]) {
  const int32_t X = bit / c.vkb, t = bit - X * c.vkb;
  double2 *fp = (double2 *)(c.fpair + 6 * ((int64_t)c.vbase[X] + t));
  fp[0] = make_double2(g0[0], g0[1]);
  fp[1] = make_double2(g0[2], g1[0]);
  fp[2] = make_double2(g1[1], g1[2]);
  atomicOr(&c.vmark[bit >> 5], 1u << (bit & 31));
  atomicOr(&c.cnew[X >> 6], 1ull << (X & 63));
}

// the pair's force on the nodes of both elements (one evaluation), stored
// in the fpair rows of both entries; each row is summed by kc_vsum in its
// element's list order
__device__ __forceinline__ bool vpair_force(const ContactDev &c, const DevData &d, int64_t q) {
  const int4 r0 = ((const int4 *)c.vrec)[2 * q], r1 = ((const int4 *)c.vrec)[2 * q + 1];
  const int32_t bitA = r0.x, B = r0.y; // bitA >= 0: not blacklisted
  double pA[2][3], pB[2][3];
  double g0[3] = {0.0, 0.0, 0.0}, g1[3] = {0.0, 0.0, 0.0};
  double h0[3] = {0.0, 0.0, 0.0}, h1[3] = {0.0, 0.0, 0.0};
  elem_pos(d, r1.x, r1.y, pA);
  elem_pos(d, r1.z, r1.w, pB);
  const double R = c.crad[r0.z] + c.crad[r0.w];
  if (!cm_narrowphase<true, 2>(pA, pB, R, c.kS, c.kN, g0, g1, h0, h1)) return false;
  vstore(c, bitA, g0, g1);
  // B's entry of the pair: A in B's list, which is sorted by partner id.
  // The candidate test is symmetric, so B's list holds A unless an overflow
  // truncated it (reported by the list build).
  const int32_t A = bitA / c.vkb;
  const int32_t *lb = c.vlist + (int64_t)B * c.vk;
  const int32_t mb = c.vcount[B];
  int32_t lo = 0, hi = mb;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (list_partner(lb[mid]) < A) lo = mid + 1;
    else hi = mid;
  }
  if (lo < mb && list_partner(lb[lo]) == A) vstore(c, (int32_t)((int64_t)B * c.vkb + lo), h0, h1);
  return true;
}

// A block takes 256 records at a time: the box test of each, then the
// pairs that need the contact evaluation (boxes overlap, not blacklisted)
// join a queue in LDS, which is evaluated 256 at a time, so that the
// evaluation runs on full wavefronts although only a fraction of the listed
// pairs overlap.  512 blocks (two wavefronts per SIMD) loop over the
// records: on the 40 %-strain crush state (1.9 M records) 512 blocks ran the
// step 8 % faster than 2,048 and 2-5 % faster than 1,024 or 256, and it
// matched 2,048 on the light load at 20 % strain.
#define VM_GRID 512
__global__ __launch_bounds__(CB) void kc_vmark(ContactDev c, DevData d,
                                               unsigned long long *cnt_act,
                                               unsigned long long *cnt_cand,
                                               const int32_t *step) {
  __shared__ int32_t qq[2 * CB];
  __shared__ int32_t qn;
  const int lane = threadIdx.x & 63;
  const int64_t N = c.N;
  const int64_t ne = c.nentries[1];
  unsigned long long nact = 0, ncand = 0;
  int32_t n = 0; // queue length, the same in every thread
  if (threadIdx.x == 0) qn = 0;
  __syncthreads();
  for (int64_t q0 = (int64_t)blockIdx.x * CB;; q0 += (int64_t)gridDim.x * CB) {
    const bool more = q0 < ne; // the same in every thread
    if (more) {
      const int64_t q = q0 + threadIdx.x;
      bool ev = false;
      if (q < ne) {
        const int4 r0 = ((const int4 *)c.vrec)[2 * q], r1 = ((const int4 *)c.vrec)[2 * q + 1];
        const double rA = c.crad[r0.z], rB = c.crad[r0.w];
        double alo[3], ahi[3], blo[3], bhi[3];
        elem_box(d, N, r1.x, r1.y, rA, alo, ahi);
        elem_box(d, N, r1.z, r1.w, rB, blo, bhi);
        bool ov = true;
        for (int i = 0; i < 3; i++) ov = ov && (ahi[i] >= blo[i]) && (bhi[i] >= alo[i]);
        if (ov) ncand++;
        ev = ov && r0.x >= 0;
      }
      // append to the queue (n < CB here, so at most 2 CB - 1 entries)
      const unsigned long long bal = __ballot(ev);
      int32_t wb = 0;
      if (lane == 0 && bal) wb = atomicAdd(&qn, __popcll(bal));
      wb = __shfl(wb, 0, 64);
      if (ev) qq[n + wb + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)q;
      __syncthreads();
      n += qn;
      __syncthreads();
      if (threadIdx.x == 0) qn = 0;
    }
    // a full wave set of pairs, or the rest after the last records
    if (n >= CB || (!more && n > 0)) {
      const int32_t k = n < CB ? n : CB;
      if (threadIdx.x < k && vpair_force(c, d, qq[threadIdx.x])) nact++;
      __syncthreads();
      if (threadIdx.x < n - k) qq[threadIdx.x] = qq[k + threadIdx.x];
      n -= k;
    }
    __syncthreads();
    if (!more) break;
  }
  for (int o = 32; o > 0; o >>= 1) {
    nact += __shfl_xor(nact, o, 64);
    ncand += __shfl_xor(ncand, o, 64);
  }
  if (lane == 0 && cnt_act) {
    const int64_t st = *step;
    if (nact) atomicAdd(cnt_act + st, nact);
    if (ncand) atomicAdd(cnt_cand + st, ncand);
  }
}

// the force row of X from its closed listed pairs, in list order: the pairs
// kc_vmark marked, whose forces it stored per entry
__device__ __forceinline__ void vsum_elem(const ContactDev &c, const DevData &d, int64_t X) {
  const int64_t E = c.E;
  double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0};
  const double2 *fp = (const double2 *)(c.fpair + 6 * (int64_t)c.vbase[X]);
  uint32_t *mk = c.vmark + X * (c.vkb / 32);
  const int32_t t1 = c.vcount[X];
  for (int32_t w0 = 0; w0 < t1; w0 += 32) {
    uint32_t m = mk[w0 >> 5];
    if (!m) continue;
    mk[w0 >> 5] = 0u;
    while (m) {
      const int t = w0 + __ffs(m) - 1;
      m &= m - 1u;
      const double2 a = fp[3 * t], b = fp[3 * t + 1], e = fp[3 * t + 2];
      const double g0[3] = {a.x, a.y, b.x}, g1[3] = {b.y, e.x, e.y};
      for (int i = 0; i < 3; i++) {
        fa[i] += g0[i];
        fb[i] += g1[i];
      }
    }
  }
  for (int i = 0; i < 3; i++) {
    c.fcs[i * E + X] = fa[i];
    c.fcs[(3 + i) * E + X] = fb[i];
  }
}

// elements per lane: 4 ran the 40 %-strain crush state 1 % faster than 8
// and 16, and 32 slower; the same at 20 %
#define VN_PER 4
__global__ __launch_bounds__(CB) void kc_vsum(ContactDev c, DevData d) {
  __shared__ int32_t qe[CB / 64][64 * VN_PER];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t T = (int64_t)blockIdx.x * CB + threadIdx.x;
  // the lists are current (list_refresh ran before): clear the rebuild flag;
  // the step's node updates may set it again.  Lists truncated by an overflow
  // stay stale, so every later step rebuilds them and reports the overflow
  // again instead of running on the truncated lists.
  if (T == 0 && *c.err != DL_ERR_CAPACITY) *c.gate = 0;
  const int64_t xw = (T - lane) * VN_PER; // the wavefront's first element
  const int64_t nw = (c.E + 63) / 64;
  // the wavefront's VN_PER mask words: lane k < VN_PER holds word k
  unsigned long long word = 0ull;
  const int64_t wi = (xw >> 6) + lane;
  if (lane < VN_PER && wi < nw) word = c.cnew[wi];
  // queue of the elements with a closed pair, in element order
  const int cnt = lane < VN_PER ? __popcll(word) : 0;
  int pre = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(pre, o, 64);
    if (lane >= o) pre += v;
  }
  const int total = __shfl(pre, 63, 64);
  pre -= cnt;
  for (unsigned long long m = word; m; m &= m - 1ull)
    qe[wv][pre++] = (int32_t)((wi << 6) + __ffsll((long long)m) - 1);
  __syncthreads();
  for (int i = lane; i < total; i += 64) vsum_elem(c, d, qe[wv][i]);
  if (lane < VN_PER && wi < nw) {
    c.cbits[wi] = word;
    c.cnew[wi] = 0ull;
  }
}

// clears the contact bit mask before a cell-scan evaluation sets its bits
__global__ __launch_bounds__(CB) void kc_clear_bits(unsigned long long *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Adjoint of the contact forces from the lists (see kc_adj_narrow), one
// thread per element in id order.
__global__ __launch_bounds__(CB) void kc_adj_vnarrow(ContactDev c, DevData d, AdjDev a) {
  const int64_t X = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (X == 0 && *c.err != DL_ERR_CAPACITY) *c.gate = 0; // lists current (see kc_vsum)
  if (X >= c.E) return;
  const int64_t N = c.N, E = c.E;
  const int32_t sX = c.es[X], xa = c.en0[X], xb = c.en1[X];
  const double rX = c.crad[sX];
  double pX[2][3], xlo[3], xhi[3], fbX[6];
  elem_pos(d, xa, xb, pX);
  elem_box(d, N, xa, xb, rX, xlo, xhi);
  for (int i = 0; i < 3; i++) {
    fbX[i] = a.fbar[i * N + xa];
    fbX[3 + i] = a.fbar[i * N + xb];
  }
  double g[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool any = false;
  const int32_t *lst = c.vlist + X * c.vk;
  const int32_t t1 = c.vcount[X];
  for (int32_t t = 0; t < t1; t++) {
    const int32_t Y = lst[t];
    if (Y < 0) continue; // blacklisted at INIT
    const int32_t sY = c.es[Y], ya = c.en0[Y], yb = c.en1[Y];
    const double rY = c.crad[sY];
    double ylo[3], yhi[3];
    elem_box(d, N, ya, yb, rY, ylo, yhi);
    bool ov = true;
    for (int i = 0; i < 3; i++) ov = ov && (xhi[i] >= ylo[i]) && (yhi[i] >= xlo[i]);
    if (!ov) continue;
    double pY[2][3], fbY[6], o[7];
    elem_pos(d, ya, yb, pY);
    for (int i = 0; i < 3; i++) {
      fbY[i] = a.fbar[i * N + ya];
      fbY[3 + i] = a.fbar[i * N + yb];
    }
    if (contact_pair_vjp(pX, pY, sX < sY, rX + rY, c.kS, c.kN, fbX, fbY, o)) {
      any = true;
      for (int i = 0; i < 7; i++) g[i] += o[i];
    }
  }
  if (!any) return;
  for (int i = 0; i < 3; i++) {
    a.ebuf[i * E + X] += g[i];
    a.ebuf[(7 + i) * E + X] += g[3 + i];
  }
  a.crbar[X] += g[6];
}

// ------------------------------------------------------------------
// host side
// ------------------------------------------------------------------
static unsigned nb(int64_t n, int b = CB) { return (unsigned)((n + b - 1) / b); }

#define CHK(x)                                                                          \
  do {                                                                                  \
    hipError_t _e = (x);                                                                \
    if (_e != hipSuccess)                                                               \
      return dl_set_error(DL_ERR_HIP, (std::string("HIP: ") + hipGetErrorString(_e)).c_str()); \
  } while (0)

template <class T>
static int calloc_dev(ContactState &cs, T **p, size_t n) {
  return cs.alloc((void **)p, sizeof(T) * (n ? n : 1));
}

template <class T>
static int up(ContactState &cs, T **p, const std::vector<T> &h) {
  int rc = calloc_dev(cs, p, h.size());
  if (rc) return rc;
  if (!h.empty()) CHK(hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return DL_OK;
}

int contact_create(ContactState &cs, DevData &d, const dl_mesh *mesh, const dl_contact *co,
                   const DevAlloc &alloc, const DevFree &release) {
  cs.alloc = alloc;
  cs.release = release;
  ContactDev &c = cs.c;
  const int64_t E = mesh->n_elems, N = mesh->n_nodes;
  const int32_t R = mesh->n_struts;
  c.E = E;
  c.N = N;
  c.R = R;
  if (!co->radius) return dl_set_error(DL_ERR_INPUT, "contact radius array is required");
  if (!(co->penalty_sts >= 0.0) || !(co->penalty_nts >= 0.0))
    return dl_set_error(DL_ERR_INPUT, "contact penalties must be >= 0");
  std::vector<double> crad(co->radius, co->radius + R);
  for (double r : crad)
    if (!(r > 0.0)) return dl_set_error(DL_ERR_INPUT, "contact radius must be > 0");
  c.kS = co->penalty_sts;
  c.kN = co->penalty_nts;
  const int32_t *en = mesh->elem_nodes, *esp = mesh->elem_strut;
  std::vector<int32_t> es(esp, esp + E), e0(E), e1(E);
  for (int64_t e = 0; e < E; e++) {
    e0[e] = en[2 * e];
    e1[e] = en[2 * e + 1];
  }
  // node -> struts (sorted, unique)
  std::vector<int64_t> key(2 * E);
  for (int64_t e = 0; e < E; e++) {
    key[2 * e] = ((int64_t)e0[e] << 32) | (uint32_t)es[e];
    key[2 * e + 1] = ((int64_t)e1[e] << 32) | (uint32_t)es[e];
  }
  std::sort(key.begin(), key.end());
  key.erase(std::unique(key.begin(), key.end()), key.end());
  std::vector<int32_t> nsptr(N + 1, 0), nslist(key.size());
  for (size_t k = 0; k < key.size(); k++) {
    nsptr[(key[k] >> 32) + 1]++;
    nslist[k] = (int32_t)(key[k] & 0xffffffff);
  }
  for (int64_t n = 0; n < N; n++) nsptr[n + 1] += nsptr[n];
  // grid from the rest configuration with a margin
  const double *X = mesh->coords;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t n = 0; n < N; n++)
    for (int i = 0; i < 3; i++) {
      lo[i] = std::min(lo[i], X[3 * n + i]);
      hi[i] = std::max(hi[i], X[3 * n + i]);
    }
  double ext = 0.0, rmax = 0.0;
  for (int64_t e = 0; e < E; e++) {
    const double r = crad[es[e]];
    rmax = std::max(rmax, r);
    for (int i = 0; i < 3; i++)
      ext = std::max(ext, std::fabs(X[3 * e1[e] + i] - X[3 * e0[e] + i]) + 2 * r);
  }
  double h0 = 1.001 * ext;
  double span = 0.0;
  for (int i = 0; i < 3; i++) span = std::max(span, hi[i] - lo[i]);
  const double margin = 0.1 * span + 2 * h0 + 2 * rmax;
  int64_t dims[3], ncell;
  for (;;) {
    ncell = 1;
    for (int i = 0; i < 3; i++) {
      dims[i] = (int64_t)std::ceil((hi[i] - lo[i] + 2 * margin) / h0) + 1;
      ncell *= dims[i];
    }
    if (ncell <= 4 * E + 4096) break;
    h0 *= 1.25;
  }
  for (int i = 0; i < 3; i++) {
    c.org[i] = lo[i] - margin;
    c.dim[i] = (int32_t)dims[i];
  }
  c.ncell = ncell;
  c.h0 = h0;
  int rc;
  int32_t *pi;
  double *pd;
  if ((rc = up(cs, &pi, es))) return rc; c.es = pi;
  if ((rc = up(cs, &pi, e0))) return rc; c.en0 = pi;
  if ((rc = up(cs, &pi, e1))) return rc; c.en1 = pi;
  if ((rc = up(cs, &pd, crad))) return rc; c.crad = pd;
  if ((rc = up(cs, &pi, nsptr))) return rc; c.nsptr = pi;
  if ((rc = up(cs, &pi, nslist))) return rc; c.nslist = pi;
  {
    std::vector<uint8_t> ej(E);
    for (int64_t e = 0; e < E; e++)
      ej[e] = (uint8_t)((nsptr[e0[e] + 1] - nsptr[e0[e]] > 1 ? 1 : 0) |
                        (nsptr[e1[e] + 1] - nsptr[e1[e]] > 1 ? 2 : 0));
    uint8_t *pu;
    if ((rc = up(cs, &pu, ej))) return rc;
    c.ejoint = pu;
    // strut -> joint nodes, in node order
    std::vector<int32_t> sjptr((size_t)R + 1, 0), sjlist;
    for (size_t k = 0; k < key.size(); k++) {
      const int64_t n = key[k] >> 32;
      if (nsptr[n + 1] - nsptr[n] > 1) sjptr[(key[k] & 0xffffffff) + 1]++;
    }
    for (int32_t s = 0; s < R; s++) sjptr[s + 1] += sjptr[s];
    sjlist.resize(sjptr[R]);
    std::vector<int32_t> fill(sjptr.begin(), sjptr.end() - 1);
    for (size_t k = 0; k < key.size(); k++) {
      const int64_t n = key[k] >> 32;
      if (nsptr[n + 1] - nsptr[n] > 1) sjlist[fill[key[k] & 0xffffffff]++] = (int32_t)n;
    }
    if ((rc = up(cs, &pi, sjptr))) return rc; c.sjptr = pi;
    if ((rc = up(cs, &pi, sjlist))) return rc; c.sjlist = pi;
  }
  cs.nblk_e = (int32_t)nb(E);
  cs.nscan_blk = (std::max<int64_t>(ncell, E) + 1023) / 1024 + 1;
  if ((rc = calloc_dev(cs, &c.extbits, 1))) return rc;
  CHK(hipMemset(c.extbits, 0, sizeof(unsigned long long)));
  if ((rc = calloc_dev(cs, &c.box, 6 * E))) return rc;
  if ((rc = calloc_dev(cs, &c.cell, E))) return rc;
  if ((rc = calloc_dev(cs, &c.rank, E))) return rc;
  if ((rc = calloc_dev(cs, &c.ccount, ncell))) return rc;
  CHK(hipMemset(c.ccount, 0, sizeof(int32_t) * ncell)); // kc_cells expects zero counts
  if ((rc = calloc_dev(cs, &c.cstart, ncell + 1))) return rc;
  if ((rc = calloc_dev(cs, &c.cbsum, (ncell + 1023) / 1024 + 1))) return rc;
  if ((rc = calloc_dev(cs, &c.bsum, cs.nscan_blk + 1))) return rc;
  if ((rc = calloc_dev(cs, (double **)&c.rec, 8 * (size_t)E))) return rc;
  if ((rc = calloc_dev(cs, &c.tmp, E))) return rc;
  if ((rc = calloc_dev(cs, (double **)&c.frec, 4 * (size_t)E))) return rc;
  if ((rc = calloc_dev(cs, &c.cbits, (E + 63) / 64))) return rc;
  if ((rc = calloc_dev(cs, &c.fcs, 6 * E))) return rc;
  if ((rc = calloc_dev(cs, &c.lcount, E))) return rc;
  if ((rc = calloc_dev(cs, &c.loff, E + 1))) return rc;
  c.black = nullptr;
  c.nblack = 0;
  c.la = c.lb = nullptr;
  c.lcap = 0;
  // no contact before the first evaluation
  CHK(hipMemset(c.cbits, 0, sizeof(unsigned long long) * ((E + 63) / 64)));
  d.cbits = c.cbits;
  d.fcs = c.fcs;
  // candidate lists of the step: skin 0 = half the largest radius, < 0 = off
  c.skin = co->skin < 0.0 ? 0.0 : (co->skin > 0.0 ? co->skin : 0.5 * rmax);
  c.err = d.err;
  d.uref = nullptr;
  d.cgate = nullptr;
  d.ctrig = 0.0;
  if (c.skin > 0.0) {
    // list capacity per element: max_pairs / E partners, default 48
    const int64_t vk = co->max_pairs > 0 ? std::max<int64_t>(1, co->max_pairs / E) : 48;
    if (vk > INT32_MAX) return dl_set_error(DL_ERR_INPUT, "max_pairs too large");
    c.vk = (int32_t)vk;
    if ((rc = calloc_dev(cs, &c.gate, 1))) return rc;
    if ((rc = calloc_dev(cs, &c.nbuild, 1))) return rc;
    CHK(hipMemset(c.nbuild, 0, sizeof(unsigned long long)));
    if ((rc = calloc_dev(cs, &c.uref, 3 * N))) return rc;
    if ((rc = calloc_dev(cs, &c.vcount, E))) return rc;
    if ((rc = calloc_dev(cs, &c.vlist, (size_t)E * c.vk))) return rc;
    c.vkb = (c.vk + 31) & ~31;
    if ((uint64_t)E * (uint64_t)c.vkb > (uint64_t)INT32_MAX)
      return dl_set_error(DL_ERR_INPUT,
                          "n_elems * (max_pairs / n_elems rounded up to 32) must stay below 2^31");
    if ((rc = calloc_dev(cs, &c.vrec, 8 * (size_t)E * c.vk))) return rc;
    if ((rc = calloc_dev(cs, &c.vbase, E))) return rc;
    if ((rc = calloc_dev(cs, &c.fpair, 6 * (size_t)E * c.vk))) return rc;
    if ((rc = calloc_dev(cs, &c.nentries, 2))) return rc;
    if ((rc = calloc_dev(cs, &c.vmark, (size_t)E * (c.vkb / 32)))) return rc;
    if ((rc = calloc_dev(cs, &c.cnew, (E + 63) / 64))) return rc;
    CHK(hipMemset(c.nentries, 0, 2 * sizeof(int32_t)));
    CHK(hipMemset(c.vmark, 0, sizeof(uint32_t) * (size_t)E * (c.vkb / 32)));
    CHK(hipMemset(c.cnew, 0, sizeof(unsigned long long) * ((E + 63) / 64)));
    CHK(hipMemset(c.vcount, 0, sizeof(int32_t) * E));
    const int32_t one = 1; // lists are built at the first step
    CHK(hipMemcpy(c.gate, &one, sizeof one, hipMemcpyHostToDevice));
    d.uref = c.uref;
    d.cgate = c.gate;
    // rebuild once a node has moved by 0.95 s: the margin covers rounding
    // of the box bounds, far below 5 % of s
    d.ctrig = 0.95 * c.skin;
  }
  return DL_OK;
}

// grid of a grid-stride kernel over n items
static unsigned gs(int64_t n) { return std::min<unsigned>(nb(n), GS_CAP); }

static void scan_i32(ContactState &cs, const int32_t *in, int64_t n, int32_t *out, hipStream_t s,
                     const int32_t *gate = nullptr) {
  const int64_t nchunk = (n + 1023) / 1024;
  ContactDev &c = cs.c;
  const unsigned g1 = (unsigned)std::min<int64_t>(nchunk, GS_CAP);
  hipLaunchKernelGGL(kc_scan1, dim3(g1), dim3(CB), 0, s, in, n, out, c.bsum, gate);
  hipLaunchKernelGGL(kc_scan2, dim3(1), dim3(1024), 0, s, c.bsum, nchunk, out + n, gate);
  hipLaunchKernelGGL(kc_scan3, dim3(gs(n)), dim3(CB), 0, s, out, n, (const int32_t *)c.bsum, gate);
}

// broadphase up to sorted cell lists, boxes inflated by infl beyond the
// contact radii; gated (see GATED) when gate is not null
static void clear_bits(const ContactDev &c, hipStream_t s) {
  const int64_t n = (c.E + 63) / 64;
  hipLaunchKernelGGL(kc_clear_bits, dim3((unsigned)((n + CB - 1) / CB)), dim3(CB), 0, s, c.cbits, n);
}

static void broadphase(ContactState &cs, const DevData &d, hipStream_t s, double infl = 0.0,
                       const int32_t *gate = nullptr, int32_t *zero = nullptr) {
  ContactDev &c = cs.c;
  const unsigned ge = gs(c.E);
  hipLaunchKernelGGL(kc_boxes, dim3(ge), dim3(CB), 0, s, c, d, infl, gate);
  hipLaunchKernelGGL(kc_cells, dim3(ge), dim3(CB), 0, s, c, gate);
  // chunk-local scan and the chunk offsets; readers add the two (cell_start)
  const int64_t nchunk = (c.ncell + 1023) / 1024;
  hipLaunchKernelGGL(kc_scan1, dim3((unsigned)std::min<int64_t>(nchunk, GS_CAP)), dim3(CB), 0, s,
                     (const int32_t *)c.ccount, c.ncell, c.cstart, c.cbsum, gate);
  hipLaunchKernelGGL(kc_scan2, dim3(1), dim3(1024), 0, s, c.cbsum, nchunk, c.cstart + c.ncell, gate);
  hipLaunchKernelGGL(kc_scatter, dim3(ge), dim3(CB), 0, s, c, gate);
  hipLaunchKernelGGL(kc_place, dim3(ge), dim3(CB), 0, s, c, gate, zero);
}

static int list_pairs(ContactState &cs, const DevData &d, hipStream_t s, int kind,
                      std::vector<int32_t> &A, std::vector<int32_t> &B) {
  ContactDev &c = cs.c;
  hipLaunchKernelGGL(kc_narrow<1>, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, kind, nullptr, nullptr,
                     nullptr);
  scan_i32(cs, c.lcount, c.E, c.loff, s);
  int32_t total = 0;
  CHK(hipMemcpyAsync(&total, c.loff + c.E, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CHK(hipStreamSynchronize(s));
  if (total > c.lcap) {
    int rc;
    const int64_t cap = std::max<int64_t>(total, 1024);
    cs.release(c.la);
    cs.release(c.lb);
    c.la = c.lb = nullptr;
    c.lcap = 0;
    if ((rc = calloc_dev(cs, &c.la, cap))) return rc;
    if ((rc = calloc_dev(cs, &c.lb, cap))) return rc;
    c.lcap = cap;
  }
  A.assign(total, 0);
  B.assign(total, 0);
  if (total > 0) {
    hipLaunchKernelGGL(kc_narrow<2>, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, kind, nullptr, nullptr,
                       nullptr);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(A.data(), c.la, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s));
    CHK(hipMemcpyAsync(B.data(), c.lb, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
  }
  // canonical order: sorted by (a, b)
  std::vector<int64_t> k(total);
  for (int32_t i = 0; i < total; i++) k[i] = ((int64_t)A[i] << 32) | (uint32_t)B[i];
  std::sort(k.begin(), k.end());
  for (int32_t i = 0; i < total; i++) {
    A[i] = (int32_t)(k[i] >> 32);
    B[i] = (int32_t)(k[i] & 0xffffffff);
  }
  return DL_OK;
}

int contact_init_blacklist(ContactState &cs, const DevData &d, hipStream_t s) {
  ContactDev &c = cs.c;
  c.nblack = 0;
  broadphase(cs, d, s);
  std::vector<int32_t> A, B;
  int rc = list_pairs(cs, d, s, LIST_CLOSED, A, B);
  if (rc) return rc;
  std::vector<unsigned long long> keys(A.size());
  for (size_t i = 0; i < A.size(); i++) {
    const unsigned long long lo = (unsigned long long)std::min(A[i], B[i]);
    const unsigned long long hi = (unsigned long long)std::max(A[i], B[i]);
    keys[i] = (lo << 32) | hi;
  }
  std::sort(keys.begin(), keys.end());
  // one buffer for the life of the context, grown when an INIT finds more
  // pairs than it holds
  if ((int64_t)keys.size() > cs.black_cap || !c.black) {
    cs.release((void *)c.black);
    c.black = nullptr;
    cs.black_cap = 0;
    unsigned long long *pk;
    if ((rc = calloc_dev(cs, &pk, std::max<size_t>(keys.size(), 64)))) return rc;
    c.black = pk;
    cs.black_cap = (int64_t)std::max<size_t>(keys.size(), 64);
  }
  if (!keys.empty())
    CHK(hipMemcpy((void *)c.black, keys.data(), sizeof(unsigned long long) * keys.size(),
                  hipMemcpyHostToDevice));
  c.nblack = (int64_t)keys.size();
  cs.last_a.clear();
  cs.last_b.clear();
  return DL_OK;
}

int contact_force(ContactState &cs, const DevData &d, hipStream_t s) {
  ContactDev &c = cs.c;
  broadphase(cs, d, s);
  clear_bits(c, s);
  hipLaunchKernelGGL(kc_narrow<0>, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, (int)LIST_ACTIVE,
                     nullptr, nullptr, nullptr);
  CHK(hipGetLastError());
  int rc = list_pairs(cs, d, s, LIST_ACTIVE, cs.last_a, cs.last_b);
  if (rc) return rc;
  return list_pairs(cs, d, s, LIST_CAND, cs.cand_a, cs.cand_b);
}

// rebuilds the lists when stale (gated), as the step does
static void list_refresh(ContactState &cs, const DevData &d, hipStream_t s) {
  ContactDev &c = cs.c;
  const int32_t *g = c.gate;
  broadphase(cs, d, s, c.skin, g, c.nentries);
  hipLaunchKernelGGL(kc_vbuild, dim3(gs(std::max(c.E, c.N))), dim3(CB), 0, s, c, d, g);
}

int contact_step(ContactState &cs, const DevData &d, hipStream_t s, int64_t *n_active,
                 int64_t *n_cand, const int32_t *step, hipEvent_t after_refresh, int refresh) {
  ContactDev &c = cs.c;
  if (!(c.skin > 0.0)) {
    // no lists: full broadphase and cell scan every step
    broadphase(cs, d, s);
    clear_bits(c, s);
    hipLaunchKernelGGL(kc_narrow<0>, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, (int)LIST_ACTIVE,
                       (unsigned long long *)n_active, (unsigned long long *)n_cand, step);
    CHK(hipGetLastError());
    return DL_OK;
  }
  // list rebuild, gated on the trigger (no-ops while the lists are valid)
  if (refresh == REFRESH_FORCE) contact_invalidate(cs, s);
  if (refresh != REFRESH_SKIP) list_refresh(cs, d, s);
  if (after_refresh) CHK(hipEventRecord(after_refresh, s));
  const int64_t nthr = (c.E + VN_PER - 1) / VN_PER;
  const unsigned gm = (unsigned)std::min<int64_t>(VM_GRID, (c.E + CB - 1) / CB);
  hipLaunchKernelGGL(kc_vmark, dim3(gm), dim3(CB), 0, s, c, d, (unsigned long long *)n_active,
                     (unsigned long long *)n_cand, step);
  hipLaunchKernelGGL(kc_vsum, dim3((unsigned)((nthr + CB - 1) / CB)), dim3(CB), 0, s, c, d);
  CHK(hipGetLastError());
  return DL_OK;
}

void contact_rebuild(ContactState &cs, const DevData &d, hipStream_t s) {
  if (!(cs.c.skin > 0.0)) return;
  contact_invalidate(cs, s);
  list_refresh(cs, d, s);
}

void contact_invalidate(ContactState &cs, hipStream_t s) {
  if (cs.c.skin > 0.0) hipLaunchKernelGGL(kc_fill_i32, dim3(1), dim3(CB), 0, s, cs.c.gate, 1, 1,
                                          (const int32_t *)nullptr);
}

int contact_list_entries(ContactState &cs, hipStream_t s, int64_t *n) {
  *n = 0;
  if (!(cs.c.skin > 0.0)) return DL_OK;
  std::vector<int32_t> h(cs.c.E);
  CHK(hipStreamSynchronize(s));
  CHK(hipMemcpy(h.data(), cs.c.vcount, sizeof(int32_t) * cs.c.E, hipMemcpyDeviceToHost));
  int64_t t = 0;
  for (int32_t v : h) t += v;
  *n = t;
  return DL_OK;
}

int contact_rebuilds(ContactState &cs, hipStream_t s, int64_t *n) {
  *n = 0;
  if (!(cs.c.skin > 0.0)) return DL_OK;
  unsigned long long v = 0;
  CHK(hipMemcpyAsync(&v, cs.c.nbuild, sizeof v, hipMemcpyDeviceToHost, s));
  CHK(hipStreamSynchronize(s));
  *n = (int64_t)v;
  return DL_OK;
}

int contact_read_pairs(ContactState &cs, hipStream_t s, bool active, int32_t *a, int32_t *b,
                       int64_t cap, int64_t *n) {
  CHK(hipStreamSynchronize(s));
  const std::vector<int32_t> &A = active ? cs.last_a : cs.cand_a;
  const std::vector<int32_t> &B = active ? cs.last_b : cs.cand_b;
  if (n) *n = (int64_t)A.size();
  const int64_t k = std::min<int64_t>(cap, (int64_t)A.size());
  for (int64_t i = 0; i < k; i++) {
    if (a) a[i] = A[i];
    if (b) b[i] = B[i];
  }
  return DL_OK;
}

int contact_adjoint(ContactState &cs, const DevData &d, const AdjDev &a, hipStream_t s,
                    int refresh) {
  ContactDev &c = cs.c;
  if (!(c.skin > 0.0)) {
    broadphase(cs, d, s);
    hipLaunchKernelGGL(kc_adj_narrow, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, a);
  } else {
    if (refresh == REFRESH_FORCE) contact_invalidate(cs, s);
    if (refresh != REFRESH_SKIP) list_refresh(cs, d, s);
    hipLaunchKernelGGL(kc_adj_vnarrow, dim3(cs.nblk_e), dim3(CB), 0, s, c, d, a);
  }
  CHK(hipGetLastError());
  return DL_OK;
}

int contact_set_radius(ContactState &cs, const double *r, hipStream_t s) {
  const int32_t R = cs.c.R;
  for (int32_t i = 0; i < R; i++)
    if (!(r[i] > 0.0)) return dl_set_error(DL_ERR_INPUT, "contact radius must be > 0");
  CHK(hipStreamSynchronize(s));
  CHK(hipMemcpy((void *)cs.c.crad, r, sizeof(double) * R, hipMemcpyHostToDevice));
  return DL_OK;
}
