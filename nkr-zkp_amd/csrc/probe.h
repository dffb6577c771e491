// Measured integer-VALU rates of the device (probe.hip).
#pragma once

#include "field.h"
#include "runtime.h"

namespace nkr {

// v_mad_u64_u32 per second over the whole device, and this library's Fr
// Montgomery products per second (two independent chains per lane).
void valu_probe(int device, double *mad_per_s, double *fr_mul_per_s);

}  // namespace nkr
