// Host-side driver of the device engine (engine.hip).
#pragma once
#include <cstdint>

#include "engine/handle.hpp"

namespace lblg {
struct DeviceFactor;
// upload values from the handle's host copy and factor
void factor(Handle& h);
// re-factor from the values already resident on the device
void factor_device(Handle& h);
// solve nrhs right-hand sides (column-major, leading dimensions ldb/ldx) with
// the refinement the settings ask for; pointers are host or device memory
void solve(Handle& h, int nrhs, const double* B, int64_t ldb, double* X, int64_t ldx, bool device_ptrs,
           double* bwd_out);
// fill the schedule knobs still unset (-1) with the environment's or built-in defaults
void resolve_tuning(Handle& h);
// time candidate device schedules (h.tune) on the handle's matrix and keep the
// fastest; the handle is left factored with it.  factor_ms / solve_ms: the
// chosen schedule's times; tried: number of schedules timed
void autotune(Handle& h, int nrhs, int reps, double& factor_ms, double& solve_ms, int& tried);
}  // namespace lblg
