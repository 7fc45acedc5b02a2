// atmo_handle.h -- the library handle and host-side helpers shared by the
// translation units of libatmo.so (not part of the public C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/atmo.h"
#include "atmo_device.h"

constexpr uint32_t kAtmoMagic = 0x41544d4fu;  // "ATMO"

struct AtmoHandle {
    uint32_t magic;
    AtmoParams params;
    atmo::Consts k;
    int width;
    int height;
    int device;
    float* tlut;
    float* mslut;
    float* svlut;
    float* aplut;
    float* sun_t;
    float* sky_l;
    bool borrowed[6];  // buffer supplied by the caller through atmo_bind_buffer
    // Pixels deferred by the COMPOSITE passes (refinement, surrogate
    // fallback): [0], [1] = two counters, then width x height pixel indices.
    // A call counts in [parity]; its last pass zeroes [1 - parity] for the
    // next.  The same allocation holds a second queue of that layout right
    // after it (refine_queue()): the surrogate fallback's pixels that need
    // the terminator refinement.
    int* queue;
    int queue_parity;
    bool queue_reset;  // a composite launch failed: clear both counters first
    int precision;     // ATMO_PRECISION_REFERENCE or ATMO_PRECISION_ROBUST
    int group_t;       // lanes per transmittance texel (1, 2, 4, 8, 16), autotuned
    int group_sv;      // lanes per sky-view texel (1, 2, 4, 8, 16), autotuned
    AtmoLutCache cache;  // the host policy of atmo_update / atmo_prepare_frame
    char error[512];
};

// Arguments of the composite kernels (renderer.py:1827-1849).
struct CompositeArgs {
    const float* scene;
    const float* space;
    const int32_t* sid;
    const float* pos;
    float* hdr;
    const float* tlut;
    const float* mslut;
    const float* svlut;
    const float* aplut;
    int width, height;
    float bottom, cam_alt, sky_alt;
    atmo::f3 sun, irradiance, disk;
    float ar;
    atmo::f3 right, up, forward;
    float thf, c0, c1;
    int view;
    int* queue;  // COMPOSITE view: deferred pixels, layout of AtmoHandle::queue
    int parity;  // the counter of this call: queue[parity]
    long queue_capacity;  // index slots after the two counters
    // a second queue of the same layout and parity for the pixels the
    // coarse pass over `queue` leaves to the refinement, or null
    int* refine = nullptr;
    // the queue the list pass (k_composite_list) works through: `queue` by
    // default, `refine` after the coarse pass over `queue`
    int* list = nullptr;
    // frame constants (composite_frame_constants)
    float aspect;        // width / height
    atmo::f3 fwd_t;      // forward / thf
    float pixel_filter;  // atan2(thf * 1.41421356 / height, 1), the sun disk's pixel filter
    float sky_horizon;   // sky_horizon_zenith at sky_alt
};

// Fills the frame constants of `a` from its other fields, with the float32
// operations of the reference's per-pixel expressions (renderer.py:1850-1859,
// 482-490, 1930-1936) evaluated once on the host (atmo_kernels.hip).
void composite_frame_constants(const atmo::Consts& k, CompositeArgs& a);

// Float32 kernel arguments of one state's LUT builds, derived on the host
// exactly as the reference's update / prepare_frame round them
// (renderer.py:2082-2187).
struct BatchState {
    atmo::Consts k;
    float bottom;          // planet radius
    atmo::f3 irradiance;   // solar irradiance
    float ar_update;       // float32(radians(alpha)), update()'s MS / SV argument
    float sky_alt;         // float32(max(r - R, 0)): sky-view altitude key
    float sun_zc;          // float32(clip(normalised sun.y)): sky-view sun key
    float cam_alt;         // max(r - R, 0): aerial-perspective altitude
    atmo::f3 sun;          // sun_local as given (aerial perspective)
    float ar;              // radians(alpha) (aerial perspective)
    atmo::f3 right, up, forward;
    float thf;
};

// Launch the full surface branch of the COMPOSITE view for the pixels
// deferred to a.list (a.queue when null), in the handle's precision variant
// (atmo_kernels.hip).
int32_t atmo_launch_composite_list(AtmoHandle* h, const CompositeArgs& a, hipStream_t stream);

// Launch the coarse-only surface branch for every pixel of a.queue (the
// surrogate composite's fallback): finished pixels are written, the rest are
// appended to a.refine for atmo_launch_composite_list (atmo_kernels.hip).
int32_t atmo_launch_composite_queue(AtmoHandle* h, const CompositeArgs& a, hipStream_t stream);

// Clears both counters of a deferred-pixel queue with a one-wave kernel
// (no memset node, so the launch can sit in a captured graph).
void atmo_launch_clear_queue(int* queue, hipStream_t stream);

// Prepares the handle's queues for one COMPOSITE-view call: clears the
// counters if an earlier call failed, and fills a.queue, a.refine, a.parity
// and a.queue_capacity (atmo_kernels.hip).
int32_t atmo_begin_queue(AtmoHandle* h, CompositeArgs& a, hipStream_t stream);

// The handle's second queue (AtmoHandle::queue).
inline int* refine_queue(const AtmoHandle* h) {
    return h->queue + 2 + (long)h->width * h->height;
}

inline atmo::f3 mk3(const float* a) { return atmo::f3{a[0], a[1], a[2]}; }

inline bool atmo_handle_ok(const AtmoHandle* h) { return h != nullptr && h->magic == kAtmoMagic; }

// Records a message (plain, or printf-formatted with arguments) on the handle.
inline void atmo_set_error(AtmoHandle* h, const char* message) {
    if (h) snprintf(h->error, sizeof h->error, "%s", message);
}

template <typename A, typename... Args>
inline void atmo_set_error(AtmoHandle* h, const char* fmt, A a, Args... args) {
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wformat-security"
#pragma clang diagnostic ignored "-Wformat-nonliteral"
    if (h) snprintf(h->error, sizeof h->error, fmt, a, args...);
#pragma clang diagnostic pop
}

// Status of the last kernel launch, recorded on the handle on failure.
inline int32_t atmo_launch_status(AtmoHandle* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        atmo_set_error(h, "%s: %s", what, hipGetErrorString(e));
        return ATMO_EHIP;
    }
    return ATMO_OK;
}
