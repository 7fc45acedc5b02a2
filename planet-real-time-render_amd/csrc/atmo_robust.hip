// atmo_robust.hip -- the robust-precision variant of the atmosphere
// kernels (atmo::robust): the stages of atmo_stages.h compiled with the
// float64 islands, Kahan-compensated optical depth and expm1 segment
// integral of ATMO_ROBUST (see atmo_device.h).  The C-ABI in
// atmo_kernels.hip launches them for handles set to
// ATMO_PRECISION_ROBUST.
#define ATMO_ROBUST 1
// Three waves per SIMD for the robust composite: 167 VGPRs and no scratch,
// 4% faster at 1920x1080 than four waves, which spill 116 B per lane.
#ifndef ATMO_COMPOSITE_WAVES
#define ATMO_COMPOSITE_WAVES 3
#endif
#include "atmo_stages.h"
#include "atmo_launch.h"
