// atmo_launch.h -- launchers of the atmosphere stages in both precision
// variants (defined in atmo_stages.h, compiled into atmo_kernels.hip for
// atmo::ref and atmo_robust.hip for atmo::robust).  Include after
// atmo_device.h.
#pragma once

#include <hip/hip_runtime.h>

#include "atmo_device.h"
#include "atmo_handle.h"

#define ATMO_DECLARE_LAUNCHERS                                                                   \
    void launch_transmittance(const Consts& k, float bottom, float* tlut, hipStream_t s, int g);  \
    void launch_multi_scattering(const Consts& k, float bottom, f3 irradiance, float ar,         \
                                 const float* tlut, float* mslut, hipStream_t s);                \
    void launch_sky_view(const Consts& k, float bottom, float cam_alt, float sun_zc,             \
                         f3 irradiance, float ar, const float* tlut, const float* mslut,         \
                         float* svlut, hipStream_t s, int g);                                    \
    void launch_aerial(const Consts& k, int width, int height, float bottom, float cam_alt,      \
                       f3 sun, f3 irradiance, float ar, f3 right, f3 up, f3 forward, float thf,  \
                       const float* tlut, const float* mslut, float* ap, hipStream_t s);         \
    void launch_prepare(const Consts& k, long npix, const float* position_view,                 \
                        const int32_t* surface_id, float bottom, float cam_alt, f3 sun,          \
                        float ar, f3 right, f3 up, f3 forward, const float* tlut,                \
                        const float* mslut, float* sun_t, float* sky_l, hipStream_t s);          \
    void launch_composite(const Consts& k, const CompositeArgs& a, hipStream_t s);              \
    void launch_composite_list(const Consts& k, const CompositeArgs& a, hipStream_t s);         \
    void launch_composite_queue(const Consts& k, const CompositeArgs& a, hipStream_t s);

// Each namespace is reopened with the inline-ness this translation unit
// gave it in atmo_device.h.
#if ATMO_ROBUST
#define ATMO_REF_NAMESPACE namespace ref
#define ATMO_ROBUST_NAMESPACE inline namespace robust
#else
#define ATMO_REF_NAMESPACE inline namespace ref
#define ATMO_ROBUST_NAMESPACE namespace robust
#endif

namespace atmo {
ATMO_REF_NAMESPACE {
ATMO_DECLARE_LAUNCHERS
// Reference precision only: the batched LUT builds.
// `stages`: bits 1, 2, 4, 8 select the T, MS, SV and AP launches.
void launch_batch(const Consts& k, const BatchState* states, unsigned count, int width, int height,
                  float* const luts[4], const long strides[4], hipStream_t s, int g_t, int g_sv,
                  int stages = 15);
}  // namespace ref
ATMO_ROBUST_NAMESPACE {
ATMO_DECLARE_LAUNCHERS
}  // namespace robust
}  // namespace atmo

#undef ATMO_DECLARE_LAUNCHERS
#undef ATMO_REF_NAMESPACE
#undef ATMO_ROBUST_NAMESPACE
