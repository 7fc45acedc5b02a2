// atmo_kernels.hip -- the C-ABI of libatmo.so (include/atmo.h) and the
// reference-precision kernels (atmo_stages.h compiled for atmo::ref).  Each
// entry point validates its arguments and launches the stage of the
// handle's precision variant: atmo::ref here, atmo::robust from
// atmo_robust.hip.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>

#include "../../include/atmo.h"
#include "atmo_stages.h"
#include "atmo_launch.h"

using namespace atmo;

// Calls the launcher `fn` of the handle's precision variant.
#define ATMO_STAGE(h, fn, ...)                                              \
    do {                                                                    \
        if ((h)->precision == ATMO_PRECISION_ROBUST)                        \
            atmo::robust::fn(__VA_ARGS__);                                  \
        else                                                                \
            atmo::ref::fn(__VA_ARGS__);                                     \
    } while (0)

namespace {

thread_local char g_create_error[512];

void set_error(AtmoHandle* h, const char* fmt, const char* a = "", const char* b = "") {
    char* dst = h ? h->error : g_create_error;
    snprintf(dst, 512, fmt, a, b);
}


// AtmosphereConfig.__post_init__ (config.py:61-128), same messages.
int32_t validate(const AtmoParams* p, char* msg, size_t cap) {
#define FAIL(...)                          \
    do {                                   \
        if (msg) snprintf(msg, cap, __VA_ARGS__); \
        return ATMO_EINVAL;                \
    } while (0)
    struct {
        const char* name;
        const double* v;
    } groups[] = {
        {"rayleigh_scattering_per_m", p->rayleigh_scattering_per_m},
        {"mie_scattering_per_m", p->mie_scattering_per_m},
        {"mie_extinction_per_m", p->mie_extinction_per_m},
        {"absorption_extinction_per_m", p->absorption_extinction_per_m},
        {"ground_albedo", p->ground_albedo},
    };
    for (auto& g : groups)
        for (int c = 0; c < 3; ++c)
            if (!(isfinite(g.v[c]) && g.v[c] >= 0.0))
                FAIL("atmosphere.%s must contain three non-negative values", g.name);
    for (int c = 0; c < 3; ++c)
        if (p->ground_albedo[c] > 1.0) FAIL("atmosphere.ground_albedo must remain in [0, 1]");
    struct {
        const char* name;
        double v;
    } positive[] = {
        {"top_altitude_m", p->top_altitude_m},
        {"rayleigh_scale_height_m", p->rayleigh_scale_height_m},
        {"mie_scale_height_m", p->mie_scale_height_m},
        {"absorption_half_width_m", p->absorption_half_width_m},
    };
    for (auto& q : positive)
        if (!isfinite(q.v) || q.v <= 0.0) FAIL("atmosphere.%s must be positive and finite", q.name);
    if (!(0.0 <= p->absorption_peak_altitude_m && p->absorption_peak_altitude_m <= p->top_altitude_m))
        FAIL("atmosphere absorption peak must lie inside the shell");
    if (!(-0.99 < p->mie_phase_g && p->mie_phase_g < 0.99))
        FAIL("atmosphere.mie_phase_g must be in (-0.99, 0.99)");
    for (int c = 0; c < 3; ++c)
        if (p->mie_scattering_per_m[c] > p->mie_extinction_per_m[c])
            FAIL("Mie scattering cannot exceed Mie extinction");
    struct {
        int v, lo, hi;
        const char* text;
    } ranges[] = {
        {p->transmittance_lut_width, 16, 2048, "transmittance LUT width must be in 16..2048"},
        {p->transmittance_lut_height, 16, 1024, "transmittance LUT height must be in 16..1024"},
        {p->sky_view_lut_width, 16, 2048, "sky-view LUT width must be in 16..2048"},
        {p->sky_view_lut_height, 16, 2048, "sky-view LUT height must be in 16..2048"},
        {p->multi_scattering_lut_width, 8, 256, "multi-scattering LUT width must be in 8..256"},
        {p->multi_scattering_lut_height, 8, 256, "multi-scattering LUT height must be in 8..256"},
        {p->aerial_lut_width, 4, 256, "aerial LUT width must be in 4..256"},
        {p->aerial_lut_height, 4, 256, "aerial LUT height must be in 4..256"},
        {p->aerial_lut_depth, 8, 256, "aerial LUT depth must be in 8..256"},
        {p->transmittance_steps, 4, 256, "transmittance_steps must be in 4..256"},
        {p->multi_scattering_directions, 4, 128, "multi_scattering_directions must be in 4..128"},
        {p->multi_scattering_steps, 4, 128, "multi_scattering_steps must be in 4..128"},
        {p->sky_view_steps, 4, 256, "sky_view_steps must be in 4..256"},
        {p->aerial_steps_per_slice, 1, 16, "aerial_steps_per_slice must be in 1..16"},
        {p->aerial_horizon_raymarch_steps, 4, 64, "aerial_horizon_raymarch_steps must be in 4..64"},
        {p->aerial_terminator_substeps, 2, 16, "aerial_terminator_substeps must be in 2..16"},
    };
    for (auto& r : ranges)
        if (r.v < r.lo || r.v > r.hi) FAIL("%s", r.text);
#undef FAIL
    return ATMO_OK;
}

Consts make_consts(const AtmoParams& p) {
    Consts k;
    const double pi = acos(-1.0);
    const double g = p.mie_phase_g;
    k.pi = (float)pi;
    k.pi16 = (float)(16.0 * pi);
    k.large = 1.0e30f;
    k.top = (float)p.top_altitude_m;
    for (int c = 0; c < 3; ++c) {
        k.beta_r[c] = (float)p.rayleigh_scattering_per_m[c];
        k.beta_ms[c] = (float)p.mie_scattering_per_m[c];
        k.beta_me[c] = (float)p.mie_extinction_per_m[c];
        k.beta_a[c] = (float)p.absorption_extinction_per_m[c];
        k.albedo[c] = (float)p.ground_albedo[c];
    }
    k.h_r = (float)p.rayleigh_scale_height_m;
    k.h_m = (float)p.mie_scale_height_m;
    k.peak = (float)p.absorption_peak_altitude_m;
    k.half_width = (float)p.absorption_half_width_m;
    k.mie_a = (float)(1.0 + g * g);
    k.mie_b = (float)(2.0 * g);
    k.mie_c = (float)(3.0 * (1.0 - g * g));
    k.mie_d = (float)(8.0 * pi * (2.0 + g * g));
    // correctly rounded float reciprocals of the divisors of quot()
    k.r_h_r = 1.0f / k.h_r;
    k.r_h_m = 1.0f / k.h_m;
    k.r_half_width = 1.0f / k.half_width;
    k.r_top = 1.0f / k.top;
    k.tw = p.transmittance_lut_width;
    k.th = p.transmittance_lut_height;
    k.mw = p.multi_scattering_lut_width;
    k.mh = p.multi_scattering_lut_height;
    k.sw = p.sky_view_lut_width;
    k.sh = p.sky_view_lut_height;
    k.tw_f = (float)k.tw;
    k.th_f = (float)k.th;
    k.mw_f = (float)k.mw;
    k.mh_f = (float)k.mh;
    k.sw_f = (float)k.sw;
    k.sh_f = (float)k.sh;
    k.tw1_f = (float)(k.tw - 1);
    k.th1_f = (float)(k.th - 1);
    k.mw1_f = (float)(k.mw - 1);
    k.mh1_f = (float)(k.mh - 1);
    k.sw1_f = (float)(k.sw - 1);
    k.sh1_f = (float)(k.sh - 1);
    k.aw = p.aerial_lut_width;
    k.ah = p.aerial_lut_height;
    k.ad = p.aerial_lut_depth;
    k.t_steps = p.transmittance_steps;
    k.ms_dirs = p.multi_scattering_directions;
    k.ms_steps = p.multi_scattering_steps;
    k.sv_steps = p.sky_view_steps;
    k.ap_spp = p.aerial_steps_per_slice;
    k.hz_steps = p.aerial_horizon_raymarch_steps;
    k.term_sub = p.aerial_terminator_substeps;
    return k;
}

f3 host3(const float* a) { return f3{a[0], a[1], a[2]}; }

}  // namespace

// ------------------------------------------------------------ C-ABI

namespace {

unsigned blocks_for(long threads, int block) { return (unsigned)((threads + block - 1) / block); }

float** buffer_slot(AtmoHandle* h, int which) {
    switch (which) {
        case ATMO_BUF_TRANSMITTANCE: return &h->tlut;
        case ATMO_BUF_MULTI_SCATTERING: return &h->mslut;
        case ATMO_BUF_SKY_VIEW: return &h->svlut;
        case ATMO_BUF_AERIAL_SCATTERING: return &h->aplut;
        case ATMO_BUF_SURFACE_SUN_T: return &h->sun_t;
        default: return &h->sky_l;
    }
}

void buffer_shape(const AtmoHandle* h, int which, int64_t s[4]) {
    const Consts& k = h->k;
    s[0] = s[1] = s[2] = s[3] = 0;
    switch (which) {
        case ATMO_BUF_TRANSMITTANCE: s[0] = k.tw; s[1] = k.th; s[2] = 3; break;
        case ATMO_BUF_MULTI_SCATTERING: s[0] = k.mw; s[1] = k.mh; s[2] = 3; break;
        case ATMO_BUF_SKY_VIEW: s[0] = k.sw; s[1] = k.sh; s[2] = 3; break;
        case ATMO_BUF_AERIAL_SCATTERING: s[0] = k.aw; s[1] = k.ah; s[2] = k.ad; s[3] = 3; break;
        default: s[0] = h->width; s[1] = h->height; s[2] = 3; break;
    }
}

size_t buffer_floats(const AtmoHandle* h, int which) {
    int64_t s[4];
    buffer_shape(h, which, s);
    size_t n = 1;
    for (int d = 0; d < 4; ++d)
        if (s[d] > 0) n *= (size_t)s[d];
    return n;
}

void free_buffers(AtmoHandle* h) {
    for (int b = 0; b < 6; ++b) {
        float** slot = buffer_slot(h, b);
        if (*slot && !h->borrowed[b]) (void)hipFree(*slot);
        *slot = nullptr;
    }
    if (h->queue) (void)hipFree(h->queue);
    h->queue = nullptr;
}

}  // namespace

extern "C" {

int32_t atmo_default_params(AtmoParams* out) {
    if (!out) return ATMO_EINVAL;
    // atmosphere/config.py:33-59
    AtmoParams p = {};
    p.top_altitude_m = 100000.0;
    const double rayleigh[3] = {5.802e-6, 13.558e-6, 33.100e-6};
    const double absorption[3] = {0.650e-6, 1.881e-6, 0.085e-6};
    for (int c = 0; c < 3; ++c) {
        p.rayleigh_scattering_per_m[c] = rayleigh[c];
        p.mie_scattering_per_m[c] = 3.996e-6;
        p.mie_extinction_per_m[c] = 4.440e-6;
        p.absorption_extinction_per_m[c] = absorption[c];
        p.ground_albedo[c] = 0.10;
    }
    p.rayleigh_scale_height_m = 8000.0;
    p.mie_scale_height_m = 1200.0;
    p.mie_phase_g = 0.80;
    p.absorption_peak_altitude_m = 25000.0;
    p.absorption_half_width_m = 15000.0;
    p.transmittance_lut_width = 256;
    p.transmittance_lut_height = 64;
    p.multi_scattering_lut_width = 32;
    p.multi_scattering_lut_height = 32;
    p.sky_view_lut_width = 192;
    p.sky_view_lut_height = 256;
    p.aerial_lut_width = 32;
    p.aerial_lut_height = 18;
    p.aerial_lut_depth = 32;
    p.transmittance_steps = 40;
    p.multi_scattering_directions = 16;
    p.multi_scattering_steps = 12;
    p.sky_view_steps = 24;
    p.aerial_steps_per_slice = 2;
    p.aerial_horizon_raymarch_steps = 12;
    p.aerial_terminator_substeps = 8;
    *out = p;
    return ATMO_OK;
}

int32_t atmo_validate_params(const AtmoParams* params, char* message, int32_t capacity) {
    if (!params) {
        if (message && capacity > 0) snprintf(message, (size_t)capacity, "params is NULL");
        return ATMO_EINVAL;
    }
    return validate(params, message, capacity > 0 ? (size_t)capacity : 0);
}

int32_t atmo_create(const AtmoParams* params, int32_t width, int32_t height, AtmoHandle** out) {
    if (!out || !params) {
        set_error(nullptr, "atmo_create: NULL argument");
        return ATMO_EINVAL;
    }
    *out = nullptr;
    if (validate(params, g_create_error, sizeof g_create_error) != ATMO_OK) return ATMO_EINVAL;
    if (width <= 0 || height <= 0) {
        set_error(nullptr, "width and height must be positive");
        return ATMO_EINVAL;
    }
    if ((long)width * (long)height > (long)ATMO_MAX_PIXELS) {
        char msg[160];
        snprintf(msg, sizeof msg, "atmo_create: %d x %d pixels exceed the limit of %ld pixels",
                 (int)width, (int)height, (long)ATMO_MAX_PIXELS);
        set_error(nullptr, msg);
        return ATMO_EINVAL;
    }
    AtmoHandle* h = new (std::nothrow) AtmoHandle();
    if (!h) {
        set_error(nullptr, "atmo_create: host allocation failed");
        return ATMO_ENOMEM;
    }
    h->magic = kAtmoMagic;
    h->params = *params;
    h->k = make_consts(*params);
    h->width = width;
    h->height = height;
    h->group_t = 8;
    h->group_sv = 8;
    h->error[0] = 0;
    for (int b = 0; b < 6; ++b) {
        *buffer_slot(h, b) = nullptr;
        h->borrowed[b] = false;
    }
    h->queue = nullptr;
    h->queue_parity = 0;
    h->queue_reset = false;
    atmo_lut_cache_init(&h->cache);
    if (hipGetDevice(&h->device) != hipSuccess) {
        set_error(nullptr, "atmo_create: no HIP device");
        delete h;
        return ATMO_EHIP;
    }
    for (int b = 0; b < 6; ++b) {
        size_t bytes = buffer_floats(h, b) * sizeof(float);
        if (hipMalloc((void**)buffer_slot(h, b), bytes) != hipSuccess ||
            hipMemset(*buffer_slot(h, b), 0, bytes) != hipSuccess) {
            set_error(nullptr, "atmo_create: device allocation failed");
            free_buffers(h);
            delete h;
            return ATMO_ENOMEM;
        }
    }
    h->queue = nullptr;
    // the deferred-pixel queue and the refinement queue after it
    if (hipMalloc((void**)&h->queue, 2 * ((size_t)width * (size_t)height + 2) * sizeof(int)) !=
            hipSuccess ||
        hipMemset(h->queue, 0, 2 * sizeof(int)) != hipSuccess ||
        hipMemset(refine_queue(h), 0, 2 * sizeof(int)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        set_error(nullptr, "atmo_create: device allocation failed");
        free_buffers(h);
        delete h;
        return ATMO_ENOMEM;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error(nullptr, "atmo_create: device synchronize failed");
        free_buffers(h);
        delete h;
        return ATMO_EHIP;
    }
    *out = h;
    return ATMO_OK;
}

int32_t atmo_destroy(AtmoHandle* h) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    free_buffers(h);
    h->magic = 0;
    delete h;
    return ATMO_OK;
}

const char* atmo_last_error(const AtmoHandle* h) {
    if (!h) return g_create_error;
    if (!atmo_handle_ok(h)) return "invalid atmosphere handle";
    return h->error;
}

int32_t atmo_get_buffer(AtmoHandle* h, int32_t which, void** dptr, int64_t shape[4]) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!dptr || !shape) {
        set_error(h, "atmo_get_buffer: NULL argument");
        return ATMO_EINVAL;
    }
    if (which < 0 || which > 5) {
        set_error(h, "atmo_get_buffer: unknown buffer id");
        return ATMO_EINVAL;
    }
    buffer_shape(h, which, shape);
    *dptr = *buffer_slot(h, which);
    return ATMO_OK;
}

int32_t atmo_bind_buffer(AtmoHandle* h, int32_t which, void* dptr, int64_t bytes) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (which < 0 || which > 5) {
        set_error(h, "atmo_bind_buffer: unknown buffer id");
        return ATMO_EINVAL;
    }
    if (!dptr) {
        set_error(h, "atmo_bind_buffer: NULL device pointer");
        return ATMO_EINVAL;
    }
    if (bytes != (int64_t)(buffer_floats(h, which) * sizeof(float))) {
        set_error(h, "atmo_bind_buffer: buffer size does not match the field shape");
        return ATMO_ESHAPE;
    }
    float** slot = buffer_slot(h, which);
    if (*slot && !h->borrowed[which]) {
        if (hipDeviceSynchronize() != hipSuccess || hipFree(*slot) != hipSuccess) {
            set_error(h, "atmo_bind_buffer: failed to release the owned buffer");
            return ATMO_EHIP;
        }
    }
    *slot = (float*)dptr;
    h->borrowed[which] = true;
    return ATMO_OK;
}

int32_t atmo_set_precision(AtmoHandle* h, int32_t precision) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (precision != ATMO_PRECISION_REFERENCE && precision != ATMO_PRECISION_ROBUST) {
        set_error(h, "atmo_set_precision: precision must be ATMO_PRECISION_REFERENCE or "
                     "ATMO_PRECISION_ROBUST");
        return ATMO_EINVAL;
    }
    h->precision = precision;
    return ATMO_OK;
}

int32_t atmo_build_transmittance(AtmoHandle* h, float bottom, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    ATMO_STAGE(h, launch_transmittance, h->k, bottom, h->tlut, (hipStream_t)stream, h->group_t);
    return atmo_launch_status(h, "atmo_build_transmittance");
}

int32_t atmo_build_multi_scattering(AtmoHandle* h, float bottom, const float irradiance[3],
                                    float ar, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!irradiance) {
        set_error(h, "atmo_build_multi_scattering: NULL irradiance");
        return ATMO_EINVAL;
    }
    ATMO_STAGE(h, launch_multi_scattering, h->k, bottom, host3(irradiance), ar, h->tlut, h->mslut,
               (hipStream_t)stream);
    return atmo_launch_status(h, "atmo_build_multi_scattering");
}

int32_t atmo_build_sky_view(AtmoHandle* h, float bottom, float cam_alt, float sun_zc,
                            const float irradiance[3], float ar, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!irradiance) {
        set_error(h, "atmo_build_sky_view: NULL irradiance");
        return ATMO_EINVAL;
    }
    ATMO_STAGE(h, launch_sky_view, h->k, bottom, cam_alt, sun_zc, host3(irradiance), ar, h->tlut,
               h->mslut, h->svlut, (hipStream_t)stream, h->group_sv);
    return atmo_launch_status(h, "atmo_build_sky_view");
}

int32_t atmo_build_aerial_perspective(AtmoHandle* h, float bottom, float cam_alt,
                                      const float sun[3], const float irradiance[3], float ar,
                                      const float right[3], const float up[3],
                                      const float forward[3], float thf, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!sun || !irradiance || !right || !up || !forward) {
        set_error(h, "atmo_build_aerial_perspective: NULL vector argument");
        return ATMO_EINVAL;
    }
    ATMO_STAGE(h, launch_aerial, h->k, h->width, h->height, bottom, cam_alt, host3(sun),
               host3(irradiance), ar, host3(right), host3(up), host3(forward), thf, h->tlut,
               h->mslut, h->aplut, (hipStream_t)stream);
    return atmo_launch_status(h, "atmo_build_aerial_perspective");
}

int32_t atmo_prepare_surface_lighting(AtmoHandle* h, const float* position_view,
                                      const int32_t* surface_id, float bottom, float cam_alt,
                                      const float sun[3], float ar, const float right[3],
                                      const float up[3], const float forward[3], void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!position_view || !surface_id || !sun || !right || !up || !forward) {
        set_error(h, "atmo_prepare_surface_lighting: NULL argument");
        return ATMO_EINVAL;
    }
    const long npix = (long)h->width * h->height;
    ATMO_STAGE(h, launch_prepare, h->k, npix, position_view, surface_id, bottom, cam_alt,
               host3(sun), ar, host3(right), host3(up), host3(forward), h->tlut, h->mslut,
               h->sun_t, h->sky_l, (hipStream_t)stream);
    return atmo_launch_status(h, "atmo_prepare_surface_lighting");
}

}  // extern "C"

void composite_frame_constants(const atmo::Consts& k, CompositeArgs& a) {
    // camera_ray (renderer.py:1850-1859): width / height and forward / thf
    a.aspect = (float)a.width / (float)a.height;
    a.fwd_t = atmo::f3{a.forward.x / a.thf, a.forward.y / a.thf, a.forward.z / a.thf};
    // the sun disk's one-pixel filter (renderer.py:1930-1936)
    a.pixel_filter = (float)atan2((double)(a.thf * 1.41421356f / (float)a.height), 1.0);
    // sky_horizon_zenith at the sky-view altitude (renderer.py:482-490)
    const float alt = fmaxf(a.sky_alt, 0.0f);
    const float radius = a.bottom + alt;
    const float rho = sqrtf(alt * (2.0f * a.bottom + alt));
    const float tc = rho / fmaxf(radius, 1.0f);
    a.sky_horizon = k.pi - (float)acos((double)fminf(1.0f, fmaxf(0.0f, tc)));
}


int32_t atmo_launch_composite_list(AtmoHandle* h, const CompositeArgs& a, hipStream_t stream) {
    ATMO_STAGE(h, launch_composite_list, h->k, a, stream);
    return atmo_launch_status(h, "k_composite_list");
}

int32_t atmo_launch_composite_queue(AtmoHandle* h, const CompositeArgs& a, hipStream_t stream) {
    ATMO_STAGE(h, launch_composite_queue, h->k, a, stream);
    return atmo_launch_status(h, "k_composite_queue_coarse");
}

namespace {

__global__ void k_clear_queue(int* queue) {
    if (threadIdx.x < 2) queue[threadIdx.x] = 0;
}

}  // namespace

void atmo_launch_clear_queue(int* queue, hipStream_t stream) {
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(64), 0, stream, queue);
}

int32_t atmo_begin_queue(AtmoHandle* h, CompositeArgs& a, hipStream_t stream) {
    if (h->queue_reset) {
        atmo_launch_clear_queue(h->queue, stream);
        atmo_launch_clear_queue(refine_queue(h), stream);
        if (atmo_launch_status(h, "atmo_composite: clearing the pixel queue") != ATMO_OK)
            return ATMO_EHIP;
        h->queue_parity = 0;
        h->queue_reset = false;
    }
    a.queue = h->queue;
    a.refine = refine_queue(h);
    a.parity = h->queue_parity;
    a.queue_capacity = (long)h->width * h->height;
    h->queue_parity ^= 1;
    return ATMO_OK;
}

extern "C" {

int32_t atmo_composite(AtmoHandle* h, const float* scene_hdr, const float* space_hdr,
                       const int32_t* surface_id, const float* position_view, float* hdr,
                       float bottom, float cam_alt, float sky_alt, const float sun[3],
                       const float irradiance[3], const float disk[3], float ar,
                       const float right[3], const float up[3], const float forward[3], float thf,
                       float c0, float c1, int32_t view, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!scene_hdr || !space_hdr || !surface_id || !position_view || !hdr || !sun ||
        !irradiance || !disk || !right || !up || !forward) {
        set_error(h, "atmo_composite: NULL argument");
        return ATMO_EINVAL;
    }
    if (view < ATMO_VIEW_COMPOSITE || view > ATMO_VIEW_AERIAL_BLEND_WEIGHT) {
        set_error(h, "atmo_composite: diagnostic_view must be in 0..10");
        return ATMO_EINVAL;
    }
    CompositeArgs a;
    a.scene = scene_hdr;
    a.space = space_hdr;
    a.sid = surface_id;
    a.pos = position_view;
    a.hdr = hdr;
    a.tlut = h->tlut;
    a.mslut = h->mslut;
    a.svlut = h->svlut;
    a.aplut = h->aplut;
    a.width = h->width;
    a.height = h->height;
    a.bottom = bottom;
    a.cam_alt = cam_alt;
    a.sky_alt = sky_alt;
    a.sun = host3(sun);
    a.irradiance = host3(irradiance);
    a.disk = host3(disk);
    a.ar = ar;
    a.right = host3(right);
    a.up = host3(up);
    a.forward = host3(forward);
    a.thf = thf;
    a.c0 = c0;
    a.c1 = c1;
    a.view = view;
    a.queue = h->queue;
    a.parity = h->queue_parity;
    a.queue_capacity = (long)h->width * h->height;
    composite_frame_constants(h->k, a);
    if (view != ATMO_VIEW_COMPOSITE) {
        ATMO_STAGE(h, launch_composite, h->k, a, (hipStream_t)stream);
        return atmo_launch_status(h, "atmo_composite");
    }
    if (atmo_begin_queue(h, a, (hipStream_t)stream) != ATMO_OK) return ATMO_EHIP;
    ATMO_STAGE(h, launch_composite, h->k, a, (hipStream_t)stream);
    if (atmo_launch_status(h, "atmo_composite") != ATMO_OK) {
        h->queue_reset = true;
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

int32_t atmo_composite_queued(AtmoHandle* h, int64_t* count, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!count) {
        set_error(h, "atmo_composite_queued: NULL count");
        return ATMO_EINVAL;
    }
    // atmo_begin_queue hands a call the current parity and then flips it,
    // so the last call's counter is the other one; that call's list pass
    // cleared only the counter of the call after it
    int n = 0;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(&n, h->queue + (h->queue_parity ^ 1), sizeof(int), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error(h, "atmo_composite_queued: reading the queue counter failed");
        return ATMO_EHIP;
    }
    const long capacity = (long)h->width * h->height;
    *count = n < 0 ? 0 : (n > capacity ? capacity : n);
    return ATMO_OK;
}

}  // extern "C"

// ------------------------------------------- the allocation-free composite

namespace {

// The deferred-pixel queue of atmo_composite_noalloc: two counters, then
// up to ATMO_NOALLOC_MAX_PIXELS pixel indices (32 MiB).  Every GPU the
// library runs on holds its own copy from the load of the code object
// until the process ends; calls on one GPU share it.
__device__ __align__(256) int scratch[ATMO_NOALLOC_MAX_PIXELS + 2];

constexpr int kMaxDevices = 64;
int* g_scratch[kMaxDevices];  // scratch's device address, per device
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device, looked up once per
// device with that device current.  NULL with the message in g_create_error
// on failure.
int* scratch_on_current_device() {
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
        set_error(nullptr, "atmo_composite_noalloc: no usable HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(nullptr, "atmo_composite_noalloc: hipGetSymbolAddress(scratch): %s",
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[device] = static_cast<int*>(ptr);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" {

int32_t atmo_composite_noalloc(const AtmoParams* params, const float* transmittance_lut,
                               const float* multi_scattering_lut, const float* sky_view_lut,
                               const float* aerial_lut, int32_t width, int32_t height,
                               const float* scene_hdr, const float* space_hdr,
                               const int32_t* surface_id, const float* position_view, float* hdr,
                               float bottom, float cam_alt, float sky_alt, const float sun[3],
                               const float irradiance[3], const float disk[3], float ar,
                               const float right[3], const float up[3], const float forward[3],
                               float thf, float c0, float c1, int32_t view, void* stream) {
    if (!params || !transmittance_lut || !multi_scattering_lut || !sky_view_lut || !aerial_lut ||
        !scene_hdr || !space_hdr || !surface_id || !position_view || !hdr || !sun ||
        !irradiance || !disk || !right || !up || !forward) {
        set_error(nullptr, "atmo_composite_noalloc: NULL argument");
        return ATMO_EINVAL;
    }
    if (validate(params, g_create_error, sizeof g_create_error) != ATMO_OK) return ATMO_EINVAL;
    if (view < ATMO_VIEW_COMPOSITE || view > ATMO_VIEW_AERIAL_BLEND_WEIGHT) {
        set_error(nullptr, "atmo_composite_noalloc: diagnostic_view must be in 0..10");
        return ATMO_EINVAL;
    }
    if (width <= 0 || height <= 0) {
        set_error(nullptr, "atmo_composite_noalloc: width and height must be positive");
        return ATMO_EINVAL;
    }
    if ((long)width * (long)height > (long)ATMO_NOALLOC_MAX_PIXELS) {
        char msg[160];
        snprintf(msg, sizeof msg,
                 "atmo_composite_noalloc: %d x %d pixels exceed the limit of %d pixels",
                 (int)width, (int)height, (int)ATMO_NOALLOC_MAX_PIXELS);
        set_error(nullptr, "%s", msg);
        return ATMO_ESHAPE;
    }
    int* queue = scratch_on_current_device();
    if (!queue) return ATMO_EHIP;
    const Consts k = make_consts(*params);
    CompositeArgs a;
    a.scene = scene_hdr;
    a.space = space_hdr;
    a.sid = surface_id;
    a.pos = position_view;
    a.hdr = hdr;
    a.tlut = transmittance_lut;
    a.mslut = multi_scattering_lut;
    a.svlut = sky_view_lut;
    a.aplut = aerial_lut;
    a.width = width;
    a.height = height;
    a.bottom = bottom;
    a.cam_alt = cam_alt;
    a.sky_alt = sky_alt;
    a.sun = host3(sun);
    a.irradiance = host3(irradiance);
    a.disk = host3(disk);
    a.ar = ar;
    a.right = host3(right);
    a.up = host3(up);
    a.forward = host3(forward);
    a.thf = thf;
    a.c0 = c0;
    a.c1 = c1;
    a.view = view;
    a.queue = queue;
    a.parity = 0;
    a.queue_capacity = (long)width * height;
    composite_frame_constants(k, a);
    const hipStream_t s = (hipStream_t)stream;
    if (view == ATMO_VIEW_COMPOSITE) atmo_launch_clear_queue(queue, s);
    atmo::ref::launch_composite(k, a, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(nullptr, "atmo_composite_noalloc: %s", hipGetErrorString(e));
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

}  // extern "C"

// ------------------------------------------------------------ autotuner

namespace {

constexpr int kGroups[5] = {1, 2, 4, 8, 16};

// Mean duration in ms of `launch` on `stream` over `reps` runs after one
// warm-up run; a negative value if a HIP call failed.
template <typename F>
float time_launch(hipStream_t stream, int reps, F&& launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0f;
    launch();
    (void)hipEventRecord(e0, stream);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, stream);
    float ms = -1.0f;
    if (hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
        ms /= (float)reps;
    else
        ms = -1.0f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

}  // namespace

extern "C" {

int32_t atmo_autotune(AtmoHandle* h, float bottom, const float irradiance[3], float ar,
                      float cam_alt, float sun_zc, void* stream, int32_t chosen[2]) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!irradiance) {
        set_error(h, "atmo_autotune: NULL irradiance");
        return ATMO_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const f3 e = host3(irradiance);
    float best = 1.0e30f;
    int gt = h->group_t;
    for (int g : kGroups) {
        float ms = time_launch(s, 5, [&] { ATMO_STAGE(h, launch_transmittance, h->k, bottom, h->tlut, s, g); });
        if (ms < 0.0f) return atmo_launch_status(h, "atmo_autotune");
        if (ms < best) best = ms, gt = g;
    }
    h->group_t = gt;
    ATMO_STAGE(h, launch_transmittance, h->k, bottom, h->tlut, s, gt);
    ATMO_STAGE(h, launch_multi_scattering, h->k, bottom, e, ar, h->tlut, h->mslut, s);
    best = 1.0e30f;
    int gs = h->group_sv;
    for (int g : kGroups) {
        float ms = time_launch(s, 5, [&] {
            ATMO_STAGE(h, launch_sky_view, h->k, bottom, cam_alt, sun_zc, e, ar, h->tlut, h->mslut,
                       h->svlut, s, g);
        });
        if (ms < 0.0f) return atmo_launch_status(h, "atmo_autotune");
        if (ms < best) best = ms, gs = g;
    }
    h->group_sv = gs;
    // the LUTs now hold this call's arguments: no cached key still applies
    h->cache.has_planet_key = h->cache.has_multi_scattering_key = h->cache.has_sky_key = 0;
    // leave the sky-view LUT as the chosen variant builds it
    ATMO_STAGE(h, launch_sky_view, h->k, bottom, cam_alt, sun_zc, e, ar, h->tlut, h->mslut,
               h->svlut, s, gs);
    if (chosen) {
        chosen[0] = gt;
        chosen[1] = gs;
    }
    return atmo_launch_status(h, "atmo_autotune");
}

int32_t atmo_get_tuning(AtmoHandle* h, int32_t tuning[2]) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!tuning) return ATMO_EINVAL;
    tuning[0] = h->group_t;
    tuning[1] = h->group_sv;
    return ATMO_OK;
}

int32_t atmo_set_tuning(AtmoHandle* h, const int32_t tuning[2]) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!tuning || !atmo::ref::valid_group(tuning[0]) || !atmo::ref::valid_group(tuning[1])) {
        set_error(h, "atmo_set_tuning: lanes per texel must be 1, 2, 4, 8 or 16");
        return ATMO_EINVAL;
    }
    h->group_t = tuning[0];
    h->group_sv = tuning[1];
    // the next update rebuilds every LUT with the new shapes
    h->cache.has_planet_key = h->cache.has_multi_scattering_key = h->cache.has_sky_key = 0;
    return ATMO_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batches

struct AtmoBatch {
    uint32_t magic;
    int count;
    int width;
    int height;
    Consts k0;           // LUT dimensions shared by every state
    int group_t;         // lanes per transmittance texel (autotuned)
    int group_sv;        // lanes per sky-view texel (autotuned)
    BatchState* states;  // device state table [count]
    float* luts[4];      // T, MS, SV, AP arrays [count * per-state floats]
    bool borrowed[4];    // array supplied by the caller
    long strides[4];     // per-state floats of each LUT
    hipStream_t capture;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool captured;
    char error[512];
};

namespace {

constexpr uint32_t kBatchMagic = 0x42415443u;  // "BATC"

bool batch_ok(const AtmoBatch* b) { return b != nullptr && b->magic == kBatchMagic; }

void batch_free(AtmoBatch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    if (b->states) (void)hipFree(b->states);
    for (int l = 0; l < 4; ++l)
        if (b->luts[l] && !b->borrowed[l]) (void)hipFree(b->luts[l]);
}

bool same_dims(const Consts& a, const Consts& b) {
    return a.tw == b.tw && a.th == b.th && a.mw == b.mw && a.mh == b.mh && a.sw == b.sw &&
           a.sh == b.sh && a.aw == b.aw && a.ah == b.ah && a.ad == b.ad;
}

// The reference's host-side rounding of one state's kernel arguments:
// update (renderer.py:2082-2137) and prepare_frame (renderer.py:2153-2184).
BatchState derive_state(const AtmoBatchState& s) {
    BatchState b;
    const double R = s.planet_radius_m, r = s.camera_radius_m;
    const double deg_to_rad = acos(-1.0) / 180.0;
    b.k = make_consts(s.params);
    b.bottom = (float)R;
    b.irradiance = f3{(float)s.solar_irradiance[0], (float)s.solar_irradiance[1],
                      (float)s.solar_irradiance[2]};
    b.ar_update = (float)(s.sun_angular_radius_deg * deg_to_rad);
    const double alt = r - R > 0.0 ? r - R : 0.0;
    b.sky_alt = (float)alt;
    const double* v = s.sun_local;
    double len = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    double y = v[1] / (len > 1.0e-12 ? len : 1.0e-12);
    b.sun_zc = (float)(y < -1.0 ? -1.0 : (y > 1.0 ? 1.0 : y));
    b.cam_alt = (float)alt;
    b.sun = f3{(float)v[0], (float)v[1], (float)v[2]};
    b.ar = (float)(s.sun_angular_radius_deg * deg_to_rad);
    b.right = f3{(float)s.right[0], (float)s.right[1], (float)s.right[2]};
    b.up = f3{(float)s.view_up[0], (float)s.view_up[1], (float)s.view_up[2]};
    b.forward = f3{(float)s.forward[0], (float)s.forward[1], (float)s.forward[2]};
    b.thf = (float)s.tangent_half_fov;
    return b;
}

void batch_error(AtmoBatch* b, const char* msg) { snprintf(b->error, sizeof b->error, "%s", msg); }

}  // namespace

extern "C" {

int32_t atmo_batch_create(const AtmoBatchState* states, int32_t count, int32_t width, int32_t height,
                          float* const luts[4], AtmoBatch** out) {
    if (!out || !states || count <= 0 || count > 65535 || width <= 0 || height <= 0) {
        set_error(nullptr, "atmo_batch_create: bad arguments");
        return ATMO_EINVAL;
    }
    *out = nullptr;
    BatchState* host = new (std::nothrow) BatchState[count];
    if (!host) return ATMO_ENOMEM;
    Consts k0 = make_consts(states[0].params);
    for (int i = 0; i < count; ++i) {
        char msg[400];
        if (validate(&states[i].params, msg, sizeof msg) != ATMO_OK) {
            char full[512];
            snprintf(full, sizeof full, "state %d: %s", i, msg);
            set_error(nullptr, "%s", full);
            delete[] host;
            return ATMO_EINVAL;
        }
        const AtmoBatchState& s = states[i];
        if (!(isfinite(s.planet_radius_m) && s.planet_radius_m > 1.0 && isfinite(s.camera_radius_m) &&
              isfinite(s.tangent_half_fov) && s.tangent_half_fov > 0.0)) {
            char full[512];
            snprintf(full, sizeof full, "state %d: planet radius, camera radius and tangent_half_fov "
                     "must be finite and positive", i);
            set_error(nullptr, "%s", full);
            delete[] host;
            return ATMO_EINVAL;
        }
        host[i] = derive_state(s);
        if (!same_dims(host[i].k, k0)) {
            set_error(nullptr, "atmo_batch_create: every state must have the same LUT dimensions");
            delete[] host;
            return ATMO_ESHAPE;
        }
    }
    AtmoBatch* b = new (std::nothrow) AtmoBatch();
    if (!b) {
        delete[] host;
        return ATMO_ENOMEM;
    }
    memset(b, 0, sizeof *b);
    b->magic = kBatchMagic;
    b->group_t = 8;
    b->group_sv = 8;
    b->count = count;
    b->width = width;
    b->height = height;
    b->k0 = k0;
    b->strides[0] = (long)k0.tw * k0.th * 3;
    b->strides[1] = (long)k0.mw * k0.mh * 3;
    b->strides[2] = (long)k0.sw * k0.sh * 3;
    b->strides[3] = (long)k0.aw * k0.ah * k0.ad * 3;
    bool ok = hipMalloc((void**)&b->states, sizeof(BatchState) * count) == hipSuccess &&
              hipMemcpy(b->states, host, sizeof(BatchState) * count, hipMemcpyHostToDevice) == hipSuccess;
    for (int l = 0; l < 4 && ok; ++l) {
        if (luts && luts[l]) {
            b->luts[l] = luts[l];
            b->borrowed[l] = true;
        } else {
            ok = hipMalloc((void**)&b->luts[l], sizeof(float) * b->strides[l] * count) == hipSuccess;
        }
    }
    delete[] host;
    if (!ok) {
        set_error(nullptr, "atmo_batch_create: device allocation failed");
        batch_free(b);
        delete b;
        return ATMO_ENOMEM;
    }
    *out = b;
    return ATMO_OK;
}

int32_t atmo_batch_destroy(AtmoBatch* b) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    (void)hipDeviceSynchronize();
    batch_free(b);
    b->magic = 0;
    delete b;
    return ATMO_OK;
}

const char* atmo_batch_last_error(const AtmoBatch* b) {
    if (!b) return g_create_error;
    return batch_ok(b) ? b->error : "invalid batch handle";
}

int32_t atmo_batch_run(AtmoBatch* b, void* stream) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    if (!b->captured) {
        const Consts& k = b->k0;
        const unsigned n = (unsigned)b->count;
        if ((!b->capture && hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess) ||
            hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            batch_error(b, "atmo_batch_run: cannot start graph capture");
            return ATMO_EHIP;
        }
        hipStream_t cs = b->capture;
        atmo::ref::launch_batch(k, b->states, n, b->width, b->height, b->luts, b->strides, cs,
                                b->group_t, b->group_sv);
        hipError_t launch = hipGetLastError();
        hipError_t end = hipStreamEndCapture(cs, &b->graph);
        if (launch != hipSuccess || end != hipSuccess ||
            hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            batch_error(b, "atmo_batch_run: graph capture failed");
            return ATMO_EHIP;
        }
        b->captured = true;
    }
    if (hipGraphLaunch(b->exec, (hipStream_t)stream) != hipSuccess) {
        batch_error(b, "atmo_batch_run: graph launch failed");
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

int32_t atmo_batch_run_stages(AtmoBatch* b, int32_t stages, void* stream) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    if (stages <= 0 || stages > 15) {
        batch_error(b, "atmo_batch_run_stages: stages must be a non-empty subset of 1|2|4|8");
        return ATMO_EINVAL;
    }
    atmo::ref::launch_batch(b->k0, b->states, (unsigned)b->count, b->width, b->height, b->luts,
                            b->strides, (hipStream_t)stream, b->group_t, b->group_sv, stages);
    if (hipGetLastError() != hipSuccess) {
        batch_error(b, "atmo_batch_run_stages: launch failed");
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

int32_t atmo_batch_autotune(AtmoBatch* b, void* stream, int32_t chosen[2]) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    hipStream_t s = (hipStream_t)stream;
    // The two shapes are independent: time the transmittance stage alone
    // for each G_T, then (with T and MS built) the sky-view stage alone for
    // each G_SV.
    int bt = b->group_t, bs = b->group_sv;
    const auto run = [&](int gt, int gs, int stages) {
        atmo::ref::launch_batch(b->k0, b->states, (unsigned)b->count, b->width, b->height, b->luts,
                                b->strides, s, gt, gs, stages);
    };
    float best = 1.0e30f;
    for (int gt : kGroups) {
        float ms = time_launch(s, 3, [&] { run(gt, bs, 1); });
        if (ms < 0.0f) {
            batch_error(b, "atmo_batch_autotune: launch failed");
            return ATMO_EHIP;
        }
        if (ms < best) best = ms, bt = gt;
    }
    run(bt, bs, 3);
    best = 1.0e30f;
    for (int gs : kGroups) {
        float ms = time_launch(s, 3, [&] { run(bt, gs, 4); });
        if (ms < 0.0f) {
            batch_error(b, "atmo_batch_autotune: launch failed");
            return ATMO_EHIP;
        }
        if (ms < best) best = ms, bs = gs;
    }
    if (bt != b->group_t || bs != b->group_sv) {
        // the captured graph holds the old shapes: capture again on the next run
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->captured = false;
        b->group_t = bt;
        b->group_sv = bs;
    }
    if (chosen) {
        chosen[0] = bt;
        chosen[1] = bs;
    }
    return ATMO_OK;
}

int32_t atmo_batch_get_buffer(AtmoBatch* b, int32_t which, void** dptr, int64_t shape[5]) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    if (!dptr || !shape || which < 0 || which > 3) {
        batch_error(b, "atmo_batch_get_buffer: bad arguments");
        return ATMO_EINVAL;
    }
    const Consts& k = b->k0;
    int64_t s[5] = {b->count, 0, 0, 0, 0};
    switch (which) {
        case 0: s[1] = k.tw; s[2] = k.th; s[3] = 3; break;
        case 1: s[1] = k.mw; s[2] = k.mh; s[3] = 3; break;
        case 2: s[1] = k.sw; s[2] = k.sh; s[3] = 3; break;
        default: s[1] = k.aw; s[2] = k.ah; s[3] = k.ad; s[4] = 3; break;
    }
    for (int d = 0; d < 5; ++d) shape[d] = s[d];
    *dptr = b->luts[which];
    return ATMO_OK;
}

}  // extern "C"
