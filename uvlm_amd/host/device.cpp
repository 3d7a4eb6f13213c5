// device.cpp -- per-thread C-ABI context of the host API.
#include "device.h"

#include <cstdio>
#include <cstdlib>

namespace UVLM {
namespace gpu {

namespace {
struct Holder {
  uvlm_ctx* ctx = nullptr;
  ~Holder() {
    if (ctx) uvlm_destroy(ctx);
  }
};
thread_local Holder tl_holder;
thread_local unsigned tl_flags = UVLM_PREC_F64_EXACT;
}  // namespace

uvlm_ctx* Context() {
  if (!tl_holder.ctx) {
    int dev = 0;
    if (const char* e = std::getenv("UVLM_DEVICE")) dev = std::atoi(e);
    uvlm_ctx* c = nullptr;
    const int rc = uvlm_create(dev, &c);
    if (rc != UVLM_OK) {
      std::fprintf(stderr,
                   "UVLM: no usable GPU (uvlm_create(device=%d) returned %d); "
                   "the UVLM hot path runs only on the GPU\n",
                   dev, rc);
      std::abort();
    }
    tl_holder.ctx = c;
  }
  return tl_holder.ctx;
}

void Check(int rc, const char* what, uvlm_ctx* ctx) {
  if (rc == UVLM_OK) return;
  std::fprintf(stderr, "UVLM: %s failed (%d): %s\n", what, rc,
               uvlm_last_error(ctx ? ctx : tl_holder.ctx));
  std::abort();
}

unsigned DefaultFlags() { return tl_flags; }
void SetDefaultFlags(unsigned flags) { tl_flags = flags; }

}  // namespace gpu
}  // namespace UVLM
