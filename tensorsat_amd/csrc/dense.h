// Routing of dense Boolean steps to the int8 MFMA kernel (dense.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tensorsat_amd.h"

namespace tsd {

// True when the step is a K-innermost two-operand product with at least 2^6 contracted
// entries and at least 4 lhs-only and 4 rhs-only out bits, operands 16-byte aligned, and
// the MFMA route is enabled (ts_dense_config).
bool dense_eligible(const ts_step& st, const void* lhs, const void* rhs);

int launch_dense(const uint8_t* lhs, const uint8_t* rhs, uint8_t* out, const ts_step& st, int mode,
                 hipStream_t s);

}  // namespace tsd
