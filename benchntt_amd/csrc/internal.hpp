// internal.hpp -- host helpers shared by the C ABI translation units
// (hidden symbols; the exported entry points are those of include/bntt.h).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bntt.h"
#include "u128.hpp"

// Record a failure for bntt_last_error() and return `code`.
int bntt_set_error(int code, const char *msg);
int bntt_set_hip_error(hipError_t e, const char *what);
// Check the current device is gfx950 (once per device).
int bntt_ensure_init();
// Modulus descriptor, including the check that selects the fast reduction.
bntt::Mod bntt_make_mod(const bntt_u128 *m, const bntt_u128 *mu);
