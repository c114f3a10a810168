// dl_contact.h -- rod-rod penalty contact on the device: uniform-grid
// broadphase (integer cell keys, counting sort, per-cell id order) and the
// closest-point narrowphase of RodContactModel.cpp:245-1208.
#ifndef DL_CONTACT_H
#define DL_CONTACT_H

#include <functional>
#include <vector>

#include "../../include/dl_gpu.h"
#include "dl_internal.h"

// Device-side contact data (passed by value to the contact kernels).
struct ContactDev {
  int64_t E, N;
  int32_t R;
  const int32_t *es;    // [E] strut of element
  const int32_t *en0;   // [E] element node a
  const int32_t *en1;   // [E] element node b
  const double *crad;   // [R] contact radius
  const int32_t *nsptr; // [N+1] node -> struts CSR
  const int32_t *nslist;
  const uint8_t *ejoint; // [E] bit 0: node a belongs to more than one strut, bit 1: node b
  const int32_t *sjptr; // [R+1] strut -> its joint nodes (nodes of more than one strut) CSR
  const int32_t *sjlist;
  double kS, kN;
  // grid
  double org[3];
  int32_t dim[3];
  int64_t ncell;
  double h0;
  unsigned long long *extbits; // [1] largest box extent (double bits) of the current evaluation
  double *box;          // [6][E]
  int32_t *cell;        // [E]
  int32_t *rank;        // [E]
  int32_t *ccount;      // [ncell]
  int32_t *cstart;      // [ncell+1] chunk-local exclusive scan of the counts (1024 cells a
                        // chunk), the total at [ncell]; cell k starts at
                        // cstart[k] + cbsum[k / 1024] (cell_start)
  int32_t *cbsum;       // [ncell / 1024 + 1] exclusive scan of the chunk totals
  int32_t *bsum;        // scan block sums
  void *rec;            // [E] cell-ordered 64-byte records (box, element, strut, cell)
  int32_t *tmp;         // [E] element ids in cell order, unsorted within a cell
  void *frec;           // [E] 32-byte float prefilter records (outward-rounded box, strut, element)
  // blacklist (sorted keys lo<<32|hi)
  const unsigned long long *black;
  int64_t nblack;
  // per-element contact forces
  unsigned long long *cbits; // [(E+63)/64] bit X: element X has a closed pair this evaluation
  double *fcs;          // [6][E] force rows by element
  // listing
  int32_t *lcount;      // [E]
  int32_t *loff;        // [E+1]
  int32_t *la, *lb;     // [lcap]
  int64_t lcap;
  // per-element candidate lists of the step (skin lists): partners whose
  // boxes inflated by the skin overlapped at the last build, sorted by id;
  // rebuilt when a node has moved by more than the trigger distance
  double skin;          // > 0: lists on
  int32_t *gate;        // [1] 1 = rebuild the lists at the next step
  double *uref;         // [3][N] displacement at the last build
  int32_t *vcount;      // [E] partners listed
  int32_t *vlist;       // [E][vk] partner ids
  int32_t vk;           // list capacity per element
  // the step's pair pass: a 32-byte record per listed pair, on its A side
  // (the element of the lower strut id): A's entry bit A * vkb + t (~bit
  // when blacklisted at INIT), B, both struts and both elements' nodes; a
  // bit per closed list entry ([E][vkb] bits, vkb = vk rounded up to 32, so
  // an element owns whole words) and a bit per element with a closed pair
  int32_t *vrec;        // [E * vk][8] (bitA, B, sA, sB, a0, a1, b0, b1)
  int32_t *vbase;       // [E] an element's first entry in fpair
  double *fpair;        // [E * vk][6] force of a closed entry on its element's nodes a, b
  int32_t *nentries;    // [2] list entries, pair records
  int32_t vkb;
  uint32_t *vmark;      // [E * vkb / 32]
  unsigned long long *cnew; // [(E+63)/64]
  int32_t *err;         // device error flag (capacity)
  unsigned long long *nbuild; // [1] list builds so far
  // while set (the adjoint's checkpointed forward run): blog[*bstep] = 1 at
  // every list build, the step's entry of a build log
  uint8_t *blog;
  const int32_t *bstep;
};

struct ContactState {
  ContactDev c{};
  int32_t nblk_e = 0;
  int64_t nscan_blk = 0;
  std::vector<int32_t> last_a, last_b;     // active pairs of the last listed evaluation
  std::vector<int32_t> cand_a, cand_b;     // candidates of the last listed evaluation
  int64_t black_cap = 0;                   // blacklist buffer capacity (keys)
  std::function<int(void **, size_t)> alloc;
  std::function<void(void *)> release;     // frees a buffer from alloc
};

using DevAlloc = std::function<int(void **, size_t)>;
using DevFree = std::function<void(void *)>;

int contact_create(ContactState &cs, DevData &d, const dl_mesh *mesh, const dl_contact *co,
                   const DevAlloc &alloc, const DevFree &release);
// pairs in contact at the current state, stored in one device buffer that
// is reused (and grown) by every INIT; the caller re-captures any step graph
// since the kernels take the blacklist's pointer and count by value
int contact_init_blacklist(ContactState &cs, const DevData &d, hipStream_t s);
// evaluation at the current state for the action path (also lists pairs)
int contact_force(ContactState &cs, const DevData &d, hipStream_t s);
// evaluation inside the explicit step: per-element forces + counts added to
// n_active[*step], n_cand[*step] (kernels only: safe inside a captured graph)
// refresh: REFRESH_GATED rebuilds the lists when the trigger has fired;
// REFRESH_FORCE rebuilds them; REFRESH_SKIP launches no rebuild stage (the
// caller knows the lists hold: the adjoint's recorded build schedule)
enum { REFRESH_GATED = 0, REFRESH_FORCE = 1, REFRESH_SKIP = 2 };
int contact_step(ContactState &cs, const DevData &d, hipStream_t s, int64_t *n_active,
                 int64_t *n_cand, const int32_t *step, hipEvent_t after_refresh = nullptr,
                 int refresh = REFRESH_GATED);
// candidate-list entries summed over the elements (after a stream sync)
int contact_list_entries(ContactState &cs, hipStream_t s, int64_t *n);
// adjoint of the contact forces at the state d (the taped start state of a
// step): broadphase, then per closed pair the VJP into a.ebuf and a.crbar
int contact_adjoint(ContactState &cs, const DevData &d, const AdjDev &a, hipStream_t s,
                    int refresh = REFRESH_GATED);
// marks the candidate lists stale (the state was set from outside the step)
void contact_invalidate(ContactState &cs, hipStream_t s);
// rebuilds the candidate lists at the state d (kernels only)
void contact_rebuild(ContactState &cs, const DevData &d, hipStream_t s);
// candidate-list builds since the context was created (after a stream sync)
int contact_rebuilds(ContactState &cs, hipStream_t s, int64_t *n);
// replaces the per-strut contact radii (host array of n_struts values)
int contact_set_radius(ContactState &cs, const double *r, hipStream_t s);
int contact_read_pairs(ContactState &cs, hipStream_t s, bool active, int32_t *a, int32_t *b,
                       int64_t cap, int64_t *n);
int dl_set_error(int code, const char *msg);

#endif
