// device.h -- the host API's handle on the gfx950 library (include/uvlm_gpu.h).
//
// Every host-API call that evaluates Biot-Savart sums, assembles the
// influence matrix or solves the circulation system goes through the C-ABI
// on a per-thread context.  Errors abort with the message of the C-ABI, as
// the reference's CHECK / LOG(FATAL) do (no exceptions, no silent CPU path).
#pragma once

#include <string>

#include "../../include/uvlm_gpu.h"

namespace UVLM {
namespace gpu {

// Context of the calling thread on device $UVLM_DEVICE (default 0), created
// on first use.
uvlm_ctx* Context();

// Abort with `what` and the C-ABI's last error when rc != UVLM_OK.
void Check(int rc, const char* what, uvlm_ctx* ctx = nullptr);

// Precision flags used by the host API's stateless calls (UVLM_PREC_*).
unsigned DefaultFlags();
void SetDefaultFlags(unsigned flags);

}  // namespace gpu
}  // namespace UVLM
