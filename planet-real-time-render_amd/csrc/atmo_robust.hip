// atmo_robust.hip -- the robust-precision variant of the atmosphere
// kernels (atmo::robust): the stages of atmo_stages.h compiled with the
// float64 islands, Kahan-compensated optical depth and cancellation-free
// segment integral of ATMO_ROBUST (see atmo_device.h).  The C-ABI in
// atmo_kernels.hip launches them for handles set to
// ATMO_PRECISION_ROBUST.
#define ATMO_ROBUST 1
// Four waves per SIMD for the robust composite, as for the reference one:
// 128 VGPRs with a 48 B per lane spill, 337 us at 1920x1080 against 367 us
// at three waves (151 VGPRs, no scratch).
#ifndef ATMO_COMPOSITE_WAVES
#define ATMO_COMPOSITE_WAVES 4
#endif
#include "atmo_stages.h"
#include "atmo_launch.h"
