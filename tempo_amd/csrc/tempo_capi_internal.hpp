// Library-internal entry points shared between the C ABI's source files.
#pragma once

#include <stddef.h>

#include "tempo_hip.h"

namespace tempo {

// One step of tempo_hip_run with its per-point scratch carved from the
// caller's device block [base, base + bytes) on the context's device (which
// must be current): nothing is allocated.  TEMPO_HIP_E_ARG when the batch
// needs more than `bytes` (the message says how much).
int run_with_scratch(tempo_hip_ctx* x, const tempo_hip_fields* f, const tempo_hip_diags* d, const tempo_hip_dims* dm,
                     double dt, int itimestep, int precision, void* stream, void* base, size_t bytes);

// Records msg as the calling thread's last error (tempo_hip_last_error) and
// returns code.
int set_error(int code, const char* msg);

}  // namespace tempo
