// dl_contact.hip -- rod-rod contact (placeholder until the grid broadphase lands).
#include "dl_contact.h"

int contact_create(ContactState &, DevData &, const dl_mesh *, const dl_contact *, const DevAlloc &) {
  return dl_set_error(DL_ERR_INPUT, "contact is not available in this build");
}
int contact_init_blacklist(ContactState &, const DevData &, hipStream_t) { return DL_OK; }
int contact_force(ContactState &, const DevData &, hipStream_t) { return DL_OK; }
int contact_step(ContactState &, const DevData &, hipStream_t, int64_t *, int64_t *) { return DL_OK; }
int contact_read_pairs(ContactState &, hipStream_t, bool, int32_t *, int32_t *, int64_t, int64_t *n) {
  if (n) *n = 0;
  return DL_OK;
}
