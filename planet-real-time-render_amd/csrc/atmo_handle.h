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
    char error[512];
};

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
