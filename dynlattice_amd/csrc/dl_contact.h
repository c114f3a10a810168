// dl_contact.h -- rod-rod penalty contact on the device (uniform-grid
// broadphase + closest-point narrowphase), RodContactModel.cpp:245-1208.
#ifndef DL_CONTACT_H
#define DL_CONTACT_H

#include <functional>

#include "../../include/dl_gpu.h"
#include "dl_internal.h"

struct ContactState {
  double *fc = nullptr; // [3][N] per-node contact force of the last evaluation
};

using DevAlloc = std::function<int(void **, size_t)>;

int contact_create(ContactState &cs, DevData &d, const dl_mesh *mesh, const dl_contact *co,
                   const DevAlloc &alloc);
int contact_init_blacklist(ContactState &cs, const DevData &d, hipStream_t s);
int contact_force(ContactState &cs, const DevData &d, hipStream_t s);
int contact_step(ContactState &cs, const DevData &d, hipStream_t s, int64_t *n_active,
                 int64_t *n_cand);
int contact_read_pairs(ContactState &cs, hipStream_t s, bool active, int32_t *a, int32_t *b,
                       int64_t cap, int64_t *n);
int dl_set_error(int code, const char *msg);

#endif
