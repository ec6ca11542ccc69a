// host_ref.cpp - host-side functions of the reference's film and colour
// modules that sit on the step's boundary: FilmSettings defaults, the
// environment toggles, pigment/CMF tables, weave and substrate init, Deposit
// geometry and PaintMix packing.  Pure host C++ (FP32 as in the Rust source).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "../../include/canvaz_film.h"
#include "czmath.h"
#include "host_ref.h"

namespace czh {
/* Host-side sin/cos/tan: evaluated in double and rounded once to float, so
 * that the result does not depend on which libm code path the CPU's feature
 * set selects (the reference's f32 calls are within one ULP of these). */
static inline float hsinf(float x) { return (float)sin((double)x); }
static inline float hcosf(float x) { return (float)cos((double)x); }
static inline float htanf(float x) { return (float)tan((double)x); }

static const float PI_F = 3.14159265358979323846f;

float to_radians(float d) { return d * (PI_F / 180.0f); }
float powi4(float x) {
    float a = x * x;
    return a * a;
}

// Deposit::profile (src/film/mod.rs:152-185)
float height_at(const cz_deposit& d, float px, float py) {
    float dx = d.p1[0] - d.p0[0], dy = d.p1[1] - d.p0[1];
    float l = sqrtf(dx * dx + dy * dy);
    float dir0 = 1.0f, dir1 = 0.0f;
    if (l > 0.0f) {
        dir0 = dx / l;
        dir1 = dy / l;
    } else {
        l = 0.0f;
    }
    bool cap_start = (d.kind & CZ_DEPOSIT_CAP_START) != 0;
    bool cap_end = (d.kind & CZ_DEPOSIT_CAP_END) != 0;
    float q0x = px - d.p0[0], q0y = py - d.p0[1];
    float q1x = px - d.p1[0], q1y = py - d.p1[1];
    if (l <= 0.0f) {
        float x = sqrtf(q0x * q0x + q0y * q0y) / d.r0;
        return (x < 1.0f) ? d.h0 * (1.0f - x * x) : 0.0f;
    }
    if (!cap_start && q0x * d.n0[0] + q0y * d.n0[1] < 0.0f) return 0.0f;
    if (!cap_end && q1x * d.n1[0] + q1y * d.n1[1] >= 0.0f) return 0.0f;
    float u = q0x * dir0 + q0y * dir1;
    if (u < -d.r0 || u > l + d.r1) return 0.0f;
    float dist = fabsf(q0x * dir1 - q0y * dir0);
    if (cap_start && u < 0.0f) dist = sqrtf(q0x * q0x + q0y * q0y);
    if (cap_end && u > l) dist = sqrtf(q1x * q1x + q1y * q1y);
    float t = std::min(std::max(u / l, 0.0f), 1.0f);
    float rt = d.r0 + (d.r1 - d.r0) * t;
    float x = dist / rt;
    return (x < 1.0f) ? d.h0 * (rt / d.r0) * (1.0f - x * x) : 0.0f;
}

// Deposit::sampled_volume_in (src/film/mod.rs:117-131)
float sampled_volume_in(const cz_deposit& d, float lattice, float lo0, float lo1, float hi0, float hi1) {
    float rm = std::max(d.r0, d.r1) + lattice;
    int x0 = (int)floorf(std::max(std::min(d.p0[0], d.p1[0]) - rm, lo0) / lattice);
    int x1 = (int)ceilf(std::min(std::max(d.p0[0], d.p1[0]) + rm, hi0) / lattice);
    int y0 = (int)floorf(std::max(std::min(d.p0[1], d.p1[1]) - rm, lo1) / lattice);
    int y1 = (int)ceilf(std::min(std::max(d.p0[1], d.p1[1]) + rm, hi1) / lattice);
    float sum = 0.0f;
    for (int i = x0; i < x1; i++)
        for (int j = y0; j < y1; j++) sum += height_at(d, ((float)i + 0.5f) * lattice, ((float)j + 0.5f) * lattice);
    return sum * lattice * lattice;
}

float sampled_volume(const cz_deposit& d, float lattice) {
    return sampled_volume_in(d, lattice, -INFINITY, -INFINITY, INFINITY, INFINITY);
}

// Deposit::inside_fraction (src/film/mod.rs:133-145)
float inside_fraction(const cz_deposit& d, float w, float h) {
    float rb = std::max(d.r0, d.r1);
    bool full = std::min(d.p0[0], d.p1[0]) - rb >= 0.0f && std::min(d.p0[1], d.p1[1]) - rb >= 0.0f &&
                std::max(d.p0[0], d.p1[0]) + rb <= w && std::max(d.p0[1], d.p1[1]) + rb <= h;
    if (full) return 1.0f;
    float lattice = 0.25f * std::max(std::min(d.r0, d.r1), 1e-4f);
    float all = sampled_volume(d, lattice);
    if (all <= 0.0f) return 0.0f;
    float v = sampled_volume_in(d, lattice, 0.0f, 0.0f, w, h) / all;
    return std::min(std::max(v, 0.0f), 1.0f);
}

float max_height(const cz_deposit& d) { return d.h0 * std::max(d.r0, d.r1) / d.r0; }

}  // namespace czh

using namespace czh;

extern "C" {

void cz_settings_default(cz_settings* s) {
    memset(s, 0, sizeof(*s));
    s->width_m = 1.2f;
    s->height_m = 0.8f;
    s->rho = 1000.0f;
    s->sigma = 0.035f;
    s->n_index = 1.0f;
    s->m_cap = 1.0e-8f;
    s->contact_angle_deg = 45.0f;
    s->hysteresis_deg = 15.0f;
    s->h_star = 2.0e-5f;
    s->raw_canvas = 1;
    s->weave_um = 25.0f;
    s->cure_rate = 1.0f / 600.0f;
    s->tau_cure = 300.0f;
    s->lnmu_cure = 5.0f;
    s->v_cap = 1.0f;
    s->tilt_deg = 0.0f;
    s->tilt_dir_deg = -90.0f;
    s->gravity = 9.81f;
    s->medium_mu = 0.08f;
    s->debug_view = 0;
    s->h_min = 1.5e-5f;
    s->m_cap_hydro = 1.0e-7f;
    s->max_substeps = 64;
    s->min_substeps = 4;
    s->implicit_iters = 40;
    s->implicit_hydro = 1;
    s->slump = 0;
    s->capillary = 1;
    s->capillary_implicit = 0;
}

void cz_overrides_default(cz_overrides* o) {
    memset(o, 0, sizeof(*o));
    o->capcond = 80.0f;
    o->fcap = 2.5f;
    o->kshape = 1.0f;
    o->kline = 0.0f;
}

static bool env_set(const char* k) { return getenv(k) != nullptr; }
static bool env_float(const char* k, float* out) {
    const char* v = getenv(k);
    if (!v) return false;
    char* end = nullptr;
    float f = strtof(v, &end);
    if (end == v) return false;
    *out = f;
    return true;
}

// The env reads of update_params (src/film/mod.rs:870-915).
void cz_overrides_from_env(cz_overrides* o) {
    cz_overrides_default(o);
    o->no_implicit = env_set("PAINT_NO_IMPLICIT");
    o->no_cap = env_set("PAINT_NO_CAP");
    env_float("PAINT_CAPCOND", &o->capcond);
    env_float("PAINT_FCAP", &o->fcap);
    o->no_hold = env_set("PAINT_NO_HOLD");
    env_float("PAINT_KSHAPE", &o->kshape);
    o->no_transl = env_set("PAINT_NO_TRANSL");
    o->no_shape = env_set("PAINT_NO_SHAPE");
    env_float("PAINT_KLINE", &o->kline);
    o->has_hyst = env_float("PAINT_HYST", &o->hyst_deg);
}

// pigment_table (src/color/mod.rs:243-258)
void cz_pigment_table(const float* refl, const float* scatter, const float* tint, const float* viscosity,
                      const float* yield_stress, float* t) {
    memset(t, 0, CZ_NPIG * CZ_PIG_STRIDE * sizeof(float));
    for (int i = 0; i < CZ_NPIG; i++) {
        for (int b = 0; b < CZ_NBAND; b++) {
            float r = std::min(std::max(refl[i * CZ_NBAND + b], 0.01f), 0.98f);
            float ks = (1.0f - r) * (1.0f - r) / (2.0f * r);
            float s = scatter[i];
            t[i * CZ_PIG_STRIDE + b] = ks * s;
            t[i * CZ_PIG_STRIDE + CZ_NBAND + b] = s;
        }
        t[i * CZ_PIG_STRIDE + 32] = viscosity[i];
        t[i * CZ_PIG_STRIDE + 33] = yield_stress[i];
        t[i * CZ_PIG_STRIDE + 34] = tint[i];
    }
}

// default_pigments (src/color/mod.rs:173-224): piecewise-linear reflectance
// samples evaluated at the band centres 410..710 nm, clamped to [0.01, 0.98].
void cz_default_pigments(float* refl, float* scatter, float* tint, float* viscosity, float* yield_stress) {
    struct Pt { float l, r; };
    struct Pg { const char* name; float s, mu, ty; int n; Pt pts[8]; };
    static const Pg pigs[12] = {
        {"Titanium White", 3.0e4f, 40.0f, 40.0f, 4, {{400, 0.86f}, {450, 0.92f}, {550, 0.94f}, {700, 0.93f}}},
        {"Ivory Black", 6.0e3f, 30.0f, 30.0f, 3, {{400, 0.045f}, {550, 0.045f}, {700, 0.055f}}},
        {"Cadmium Yellow", 2.2e4f, 35.0f, 35.0f, 7, {{400, 0.05f}, {460, 0.06f}, {490, 0.15f}, {510, 0.45f}, {530, 0.75f}, {560, 0.86f}, {700, 0.88f}}},
        {"Cadmium Red", 2.0e4f, 35.0f, 35.0f, 6, {{400, 0.05f}, {560, 0.06f}, {590, 0.20f}, {610, 0.55f}, {630, 0.78f}, {700, 0.82f}}},
        {"Alizarin Crimson", 9.0e2f, 25.0f, 15.0f, 8, {{400, 0.12f}, {430, 0.14f}, {470, 0.07f}, {540, 0.04f}, {580, 0.06f}, {610, 0.22f}, {640, 0.48f}, {700, 0.60f}}},
        {"Ultramarine Blue", 4.0e3f, 30.0f, 20.0f, 6, {{400, 0.25f}, {440, 0.36f}, {470, 0.32f}, {520, 0.12f}, {580, 0.05f}, {700, 0.06f}}},
        {"Phthalo Blue", 1.0e3f, 28.0f, 15.0f, 6, {{400, 0.12f}, {440, 0.22f}, {470, 0.26f}, {510, 0.14f}, {560, 0.04f}, {700, 0.03f}}},
        {"Phthalo Green", 1.0e3f, 28.0f, 15.0f, 7, {{400, 0.04f}, {470, 0.09f}, {500, 0.27f}, {520, 0.30f}, {550, 0.17f}, {600, 0.04f}, {700, 0.03f}}},
        {"Yellow Ochre", 1.5e4f, 38.0f, 40.0f, 6, {{400, 0.07f}, {470, 0.10f}, {520, 0.22f}, {570, 0.42f}, {620, 0.52f}, {700, 0.56f}}},
        {"Burnt Umber", 4.0e3f, 38.0f, 40.0f, 4, {{400, 0.04f}, {500, 0.05f}, {600, 0.10f}, {700, 0.17f}}},
        {"Viridian", 2.0e3f, 30.0f, 20.0f, 6, {{400, 0.06f}, {460, 0.12f}, {500, 0.26f}, {530, 0.22f}, {580, 0.07f}, {700, 0.05f}}},
        {"Cadmium Orange", 2.0e4f, 35.0f, 35.0f, 6, {{400, 0.05f}, {520, 0.06f}, {550, 0.15f}, {580, 0.50f}, {610, 0.80f}, {700, 0.85f}}},
    };
    for (int i = 0; i < 12; i++) {
        const Pg& p = pigs[i];
        for (int b = 0; b < CZ_NBAND; b++) {
            float lam = 410.0f + 20.0f * (float)b;
            float v = p.pts[p.n - 1].r;
            if (lam <= p.pts[0].l) {
                v = p.pts[0].r;
            } else {
                for (int k = 0; k + 1 < p.n; k++) {
                    if (lam >= p.pts[k].l && lam <= p.pts[k + 1].l) {
                        float t = (lam - p.pts[k].l) / (p.pts[k + 1].l - p.pts[k].l);
                        v = p.pts[k].r + t * (p.pts[k + 1].r - p.pts[k].r);
                        break;
                    }
                }
            }
            refl[i * CZ_NBAND + b] = std::min(std::max(v, 0.01f), 0.98f);
        }
        scatter[i] = p.s;
        viscosity[i] = p.mu;
        yield_stress[i] = p.ty;
        std::string nm(p.name);
        float tn = 1.0f;  // default_tint (src/color/mod.rs:217-224)
        if (nm == "Phthalo Blue" || nm == "Phthalo Green") tn = 25.0f;
        else if (nm == "Alizarin Crimson") tn = 4.0f;
        else if (nm == "Ultramarine Blue" || nm == "Viridian" || nm == "Ivory Black") tn = 1.5f;
        tint[i] = tn;
    }
}

// cmf_table (src/color/mod.rs:110-171): CIE 1931 x D65, band pairs, sum(y) = 1.
void cz_cmf_table(float* t) {
    static const float CMF[32][3] = {
        {0.0143f, 0.0004f, 0.0679f}, {0.0435f, 0.0012f, 0.2074f}, {0.1344f, 0.0040f, 0.6456f}, {0.2839f, 0.0116f, 1.3856f},
        {0.3483f, 0.0230f, 1.7471f}, {0.3362f, 0.0380f, 1.7721f}, {0.2908f, 0.0600f, 1.6692f}, {0.1954f, 0.0910f, 1.2876f},
        {0.0956f, 0.1390f, 0.8130f}, {0.0320f, 0.2080f, 0.4652f}, {0.0049f, 0.3230f, 0.2720f}, {0.0093f, 0.5030f, 0.1582f},
        {0.0633f, 0.7100f, 0.0782f}, {0.1655f, 0.8620f, 0.0422f}, {0.2904f, 0.9540f, 0.0203f}, {0.4334f, 0.9950f, 0.0087f},
        {0.5945f, 0.9950f, 0.0039f}, {0.7621f, 0.9520f, 0.0021f}, {0.9163f, 0.8700f, 0.0017f}, {1.0263f, 0.7570f, 0.0011f},
        {1.0622f, 0.6310f, 0.0008f}, {1.0026f, 0.5030f, 0.0003f}, {0.8544f, 0.3810f, 0.0002f}, {0.6424f, 0.2650f, 0.0000f},
        {0.4479f, 0.1750f, 0.0f},    {0.2835f, 0.1070f, 0.0f},    {0.1649f, 0.0610f, 0.0f},    {0.0874f, 0.0320f, 0.0f},
        {0.0468f, 0.0170f, 0.0f},    {0.0227f, 0.0082f, 0.0f},    {0.0114f, 0.0041f, 0.0f},    {0.0058f, 0.0021f, 0.0f}};
    static const float D65[32] = {82.8f, 91.5f, 93.4f, 86.7f, 104.9f, 117.0f, 117.8f, 114.9f, 115.9f, 108.8f, 109.4f,
                                  107.8f, 104.8f, 107.7f, 104.4f, 104.0f, 100.0f, 96.3f, 95.8f, 88.7f, 90.0f, 89.6f,
                                  87.7f, 83.3f, 83.7f, 80.0f, 80.2f, 82.3f, 78.3f, 69.7f, 71.6f, 74.3f};
    float ysum = 0.0f;
    for (int b = 0; b < CZ_NBAND; b++) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < 2; s++) {
            int k = 2 * b + s;
            float w = D65[k];
            acc[0] += CMF[k][0] * w;
            acc[1] += CMF[k][1] * w;
            acc[2] += CMF[k][2] * w;
        }
        t[b * 3] = acc[0];
        t[b * 3 + 1] = acc[1];
        t[b * 3 + 2] = acc[2];
        ysum += acc[1];
    }
    for (int i = 0; i < CZ_NBAND * 3; i++) t[i] /= ysum;
}

// weave (src/film/mod.rs:212-233)
void cz_weave(uint32_t w, uint32_t h, float dx, float amp, float* v) {
    memset(v, 0, (size_t)w * h * sizeof(float));
    if (amp <= 0.0f) return;
    float period = 1.4e-3f;
    float k = (2.0f * PI_F) / period;
    for (uint32_t j = 0; j < h; j++) {
        float y = ((float)j + 0.5f) * dx;
        float sy = hsinf(k * y);
        for (uint32_t i = 0; i < w; i++) {
            float x = ((float)i + 0.5f) * dx;
            uint32_t hsh = (i * 73856093u) ^ (j * 19349663u);
            hsh ^= hsh >> 13;
            hsh = hsh * 0x5bd1e995u;
            hsh ^= hsh >> 15;
            float r = (float)(hsh & 0xffffu) / 65535.0f * 2.0f - 1.0f;
            float s = hsinf(k * x) * sy;
            v[(size_t)j * w + i] = amp * (1.0f + 0.6f * s + 0.4f * r);
        }
    }
}

// The canvas colour (src/color/mod.rs:27-100).  f32::powf and f32::exp
// are evaluated in double and rounded once.
namespace {
constexpr float VIEW_GAIN[3] = {1.94f, 1.95f, 2.17f};
float powf_d(float x, float y) { return (float)pow((double)x, (double)y); }
float expf_d(float x) { return (float)exp((double)x); }
// sRGB transfer function and its inverse
float srgb_dec(float c) { return c <= 0.04045f ? c / 12.92f : powf_d((c + 0.055f) / 1.055f, 2.4f); }
float srgb_enc(float c) { return c <= 0.0031308f ? c * 12.92f : 1.055f * powf_d(c, 1.0f / 2.4f) - 0.055f; }
float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
// the display tone curve's inverse: x with x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14) = y
float untone(float y) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    y = clampf(y, 0.0f, 0.995f);
    const float qa = a - y * c;
    const float qb = b - y * d;
    return (-qb + sqrtf(qb * qb + 4.0f * qa * y * e)) / (2.0f * qa);
}
// color::canvas_albedo (src/color/mod.rs:29-44): the albedo that shows as `shown`
void canvas_albedo(const float shown[3], float out[3]) {
    for (int i = 0; i < 3; i++) out[i] = srgb_enc(clampf(untone(srgb_dec(shown[i])) / VIEW_GAIN[i], 0.005f, 0.9f));
}
float det3(const float a[3][3]) {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}
}  // namespace

// color::reflectance_of (src/color/mod.rs:46-100): three smooth band bases
// (red, green, blue) mixed to hit the colour's XYZ through the CMF table, the
// clamped result corrected six times
void cz_canvas_reflectance(const float rgb[3], const float* cmf, float out[16]) {
    float albedo[3];
    canvas_albedo(rgb, albedo);
    float lin[3];
    for (int i = 0; i < 3; i++) lin[i] = srgb_dec(albedo[i]);
    const float target[3] = {0.4124f * lin[0] + 0.3576f * lin[1] + 0.1805f * lin[2],
                             0.2126f * lin[0] + 0.7152f * lin[1] + 0.0722f * lin[2],
                             0.0193f * lin[0] + 0.1192f * lin[1] + 0.9505f * lin[2]};
    float basis[16][3];
    for (int b = 0; b < 16; b++) {
        const float l = 410.0f + 20.0f * (float)b;
        const float blue = 1.0f / (1.0f + expf_d((l - 490.0f) / 18.0f));
        const float red = 1.0f / (1.0f + expf_d(-(l - 585.0f) / 18.0f));
        const float green = 1.0f - blue - red;
        basis[b][0] = red;
        basis[b][1] = green > 0.0f ? green : 0.0f;
        basis[b][2] = blue;
    }
    float m[3][3] = {};
    for (int b = 0; b < 16; b++)
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) m[i][c] = m[i][c] + basis[b][c] * cmf[b * 3 + i];
    const float d = det3(m);
    float want[3] = {target[0], target[1], target[2]};
    for (int it = 0; it < 6; it++) {
        float k[3];
        for (int c = 0; c < 3; c++) {
            float a[3][3];
            memcpy(a, m, sizeof(a));
            for (int i = 0; i < 3; i++) a[i][c] = want[i];
            k[c] = fabsf(d) > 1e-12f ? det3(a) / d : 0.0f;
        }
        for (int b = 0; b < 16; b++)
            out[b] = clampf(basis[b][0] * k[0] + basis[b][1] * k[1] + basis[b][2] * k[2], 0.005f, 0.97f);
        float got[3] = {0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 16; b++)
            for (int i = 0; i < 3; i++) got[i] = got[i] + out[b] * cmf[b * 3 + i];
        for (int i = 0; i < 3; i++) want[i] = want[i] + (target[i] - got[i]);
    }
}

void cz_substrate_words(const cz_settings* s, const float* cmf, uint32_t out[8]) {
    float bands[16];
    if (s->has_canvas_rgb) {
        cz_canvas_reflectance(s->canvas_rgb, cmf, bands);
    } else {
        for (int b = 0; b < 16; b++) {
            float t = (float)b / 15.0f;
            bands[b] = s->raw_canvas ? 0.50f + 0.16f * t : 0.84f + 0.04f * t;
        }
    }
    for (int k = 0; k < 8; k++) out[k] = cz_f2h(bands[2 * k]) | (cz_f2h(bands[2 * k + 1]) << 16);
}

// substrate_refl without a canvas colour (src/film/mod.rs:188-210)
void cz_substrate_refl(uint64_t n, uint32_t raw, uint32_t* v) {
    float bands[16];
    for (int b = 0; b < 16; b++) {
        float t = (float)b / 15.0f;
        bands[b] = raw ? 0.50f + 0.16f * t : 0.84f + 0.04f * t;
    }
    uint32_t packed[8];
    for (int k = 0; k < 8; k++) packed[k] = cz_f2h(bands[2 * k]) | (cz_f2h(bands[2 * k + 1]) << 16);
    for (uint64_t c = 0; c < n; c++) memcpy(&v[c * 8], packed, sizeof(packed));
}

// Deposit::ribbon (src/film/mod.rs:63-80)
void cz_deposit_ribbon(const float p0[2], const float p1[2], const float n0[2], const float n1[2], float r0, float r1,
                       uint32_t caps, float volume, const uint32_t pig[8], float dx, cz_deposit* out) {
    float min_r = 0.8f * dx;
    float a = std::max(r0, min_r), b = std::max(r1, min_r);
    cz_deposit d;
    memset(&d, 0, sizeof(d));
    d.p0[0] = p0[0]; d.p0[1] = p0[1]; d.p1[0] = p1[0]; d.p1[1] = p1[1];
    d.n0[0] = n0[0]; d.n0[1] = n0[1]; d.n1[0] = n1[0]; d.n1[1] = n1[1];
    d.r0 = a; d.r1 = b; d.h0 = a; d.vol = volume; d.kind = caps;
    memcpy(d.pig, pig, 8 * sizeof(uint32_t));
    float lattice = 0.25f * dx;
    float sampled = sampled_volume(d, lattice);
    float rm = std::max(a, b);
    float disc = PI_F * rm * rm * rm / 2.0f;
    if (sampled < 0.25f * disc) {
        float mid0 = 0.5f * (p0[0] + p1[0]), mid1 = 0.5f * (p0[1] + p1[1]);
        d.p0[0] = mid0; d.p0[1] = mid1; d.p1[0] = mid0; d.p1[1] = mid1;
        d.r0 = rm; d.r1 = rm; d.h0 = rm;
        d.h0 = volume / sampled_volume(d, lattice) * rm;
        *out = d;
        return;
    }
    d.h0 = volume / sampled * a;
    *out = d;
}

static void segment_dir(const float* p0, const float* p1, float* d, float* l) {
    float dx = p1[0] - p0[0], dy = p1[1] - p0[1];
    float len = sqrtf(dx * dx + dy * dy);
    if (len > 0.0f) {
        d[0] = dx / len; d[1] = dy / len; *l = len;
    } else {
        d[0] = 1.0f; d[1] = 0.0f; *l = 0.0f;
    }
}

// Deposit::capsule / stroke (src/film/mod.rs:82-93)
void cz_deposit_capsule(const float p0[2], const float p1[2], float r, float volume, const uint32_t pig[8], float dx,
                        cz_deposit* out) {
    float d[2], l;
    segment_dir(p0, p1, d, &l);
    cz_deposit_ribbon(p0, p1, d, d, r, r, CZ_DEPOSIT_CAP_START | CZ_DEPOSIT_CAP_END, volume, pig, dx, out);
}

void cz_deposit_stroke(const float p0[2], const float p1[2], float r, float volume, const uint32_t pig[8], float dx,
                       cz_deposit* out) {
    float d[2], l;
    segment_dir(p0, p1, d, &l);
    if (l < std::max(r, 0.8f * dx)) {
        cz_deposit_capsule(p0, p1, r, volume, pig, dx, out);
        return;
    }
    cz_deposit_ribbon(p0, p1, d, d, r, r, 0, volume, pig, dx, out);
}

void cz_deposit_laid_by(cz_deposit* d, uint32_t method) { d->kind |= method << CZ_DEPOSIT_METHOD_SHIFT; }

float cz_deposit_sampled_volume(const cz_deposit* d, float lattice) { return sampled_volume(*d, lattice); }
float cz_deposit_height_at(const cz_deposit* d, float x, float y) { return height_at(*d, x, y); }
float cz_deposit_inside_fraction(const cz_deposit* d, float w, float h) { return inside_fraction(*d, w, h); }

// gravity_vec (src/film/mod.rs:861-867)
void cz_gravity_vec_host(const cz_settings* s, float out[3]) {
    float tilt = to_radians(s->tilt_deg);
    float dir = to_radians(s->tilt_dir_deg);
    float gxy = s->gravity * hsinf(tilt);
    out[0] = gxy * hcosf(dir);
    out[1] = gxy * hsinf(dir);
    out[2] = fmaxf(s->gravity * hcosf(tilt), 0.0f);
}

// update_params (src/film/mod.rs:869-917): the uniform block of one frame.
void cz_update_params_host(const cz_settings* sp, const cz_overrides* ov, uint32_t w, uint32_t h, float frame_dt,
                           uint32_t substeps, const cz_host_state* hs, cz_params* P, uint32_t* implicit_steps) {
    const cz_settings& s = *sp;
    float g[3];
    cz_gravity_vec_host(sp, g);
    float dx = s.width_m / (float)w;
    float theta = to_radians(s.contact_angle_deg);
    float pi_a = 3.0f * s.sigma * (1.0f - hcosf(theta)) / s.h_star;
    float dt = frame_dt / (float)substeps;
    float m_cap_exp = 0.5f * dx * dx * dx * dx / (32.0f * s.sigma * dt);
    float m_cap = fmaxf(fminf(s.m_cap, m_cap_exp), 1e-14f);
    float m_exp = 0.1f * dx * dx / (dt * s.rho * fmaxf(g[2], 0.1f));
    P->size[0] = w; P->size[1] = h;
    P->tiles[0] = (w + 63u) / 64u; P->tiles[1] = (h + 63u) / 64u;
    P->dx = dx; P->inv_dx = 1.0f / dx;
    P->dt = dt;
    P->frame_dt = frame_dt;
    P->g[0] = g[0]; P->g[1] = g[1]; P->g[2] = g[2]; P->g[3] = s.gravity;
    P->phys[0] = s.rho; P->phys[1] = s.sigma; P->phys[2] = s.n_index;
    P->phys[3] = (s.capillary && !ov->no_cap) ? m_cap : 0.0f;
    float lc = sqrtf(s.sigma / (s.rho * 9.81f));
    P->ein[0] = fmaxf(htanf(theta), 0.05f); P->ein[1] = lc * hsinf(0.5f * theta); P->ein[2] = pi_a; P->ein[3] = s.h_min;
    P->cure[0] = s.cure_rate; P->cure[1] = s.tau_cure; P->cure[2] = s.lnmu_cure; P->cure[3] = s.h_star;
    P->misc[0] = 2.0e-6f; P->misc[1] = 0.12f; P->misc[2] = s.medium_mu; P->misc[3] = fminf(m_exp, s.m_cap_hydro);
    P->flags[0] = s.debug_view; P->flags[1] = s.slump; P->flags[2] = s.n_per_cell ? CZ_FLAG_CELL_N : 0u; P->flags[3] = 0;
    float lc_z = sqrtf(s.sigma / (s.rho * fmaxf(g[2], 0.1f)));
    float cap_explicit = fminf(m_exp, s.m_cap_hydro);
    float h_big = fmaxf(fmaxf(hs->h_max, hs->h_max_pending), 1e-4f);
    float m_phys = h_big * h_big * h_big / (3.0f * fmaxf(hs->mu_min, 1e-3f)) - cap_explicit;
    float k_max = 40.0f;
    float kappa_frame = frame_dt * s.rho * fmaxf(g[2], 0.1f) / (dx * dx);
    uint32_t n_i = 0;
    if (m_phys > 0.0f) n_i = std::min((uint32_t)fmaxf(ceilf(kappa_frame * m_phys / k_max), 1.0f), 4u);
    float dt_i = frame_dt / (float)std::max(n_i, 1u);
    float m_imp_max = k_max * dx * dx / (dt_i * s.rho * fmaxf(g[2], 0.1f));
    *implicit_steps = n_i;
    float m_phys_cap = h_big * h_big * h_big / (3.0f * fmaxf(hs->mu_min, 1e-3f));
    float m_cap_impl = (ov->capcond / 28.0f) * powi4(dx) / (dt_i * s.sigma);
    float m_c = s.capillary_implicit ? fminf(m_phys_cap, m_cap_impl) : 0.0f;
    float cc = dt_i * s.sigma * m_c / powi4(dx);
    float kc = m_c * s.sigma / (dx * dx * dx);
    P->front3[0] = cc; P->front3[1] = kc; P->front3[2] = m_c; P->front3[3] = ov->fcap;
    P->front[0] = ov->no_hold ? 0.0f : 2.0f * lc_z * hsinf(0.5f * theta);
    P->front[1] = dx / lc_z; P->front[2] = m_imp_max; P->front[3] = dt_i;
    uint32_t flags = (ov->no_transl ? 1u : 0u) | ((ov->no_shape ? 1u : 0u) << 1);
    float hyst = ov->has_hyst ? ov->hyst_deg : s.hysteresis_deg;
    P->front2[0] = to_radians(hyst); P->front2[1] = ov->kshape; P->front2[2] = ov->kline;
    P->front2[3] = (float)flags;
}

// choose_substeps (src/film/mod.rs:919-933)
uint32_t cz_choose_substeps_host(const cz_settings* sp, uint32_t w, float frame_dt, const cz_host_state* hs) {
    const cz_settings& s = *sp;
    float g[3];
    cz_gravity_vec_host(sp, g);
    float dx = s.width_m / (float)w;
    float gxy = sqrtf(g[0] * g[0] + g[1] * g[1]);
    float h_max = fmaxf(fmaxf(hs->h_max, hs->h_max_pending), 1e-5f);
    float mu = fmaxf(hs->mu_min, 1e-3f);
    float speed = fminf(s.rho * gxy * h_max * h_max / mu, s.v_cap);
    float dt_cfl = (speed > 1e-6f) ? 0.5f * dx / speed : frame_dt;
    float mob = h_max * h_max * h_max / (3.0f * mu);
    uint32_t n_hyd = (uint32_t)fmaxf(ceilf(frame_dt * s.rho * g[2] * fminf(mob, s.m_cap_hydro) / (0.2f * dx * dx)), 1.0f);
    float dx4 = powi4(dx);
    uint32_t n_cap = (uint32_t)fmaxf(ceilf(frame_dt * 32.0f * s.sigma * fminf(mob, s.m_cap) / dx4), 1.0f);
    uint32_t n_cfl = (uint32_t)fmaxf(ceilf(frame_dt / dt_cfl), 1.0f);
    uint32_t nn = std::max(std::max(n_cfl, n_hyd), n_cap);
    return std::min(std::max(nn, s.min_substeps), s.max_substeps);
}

// PaintMix::normalized (src/color/mod.rs:358-371) and packed (348-356)
void cz_paint_pack(const float* frac16, int normalize, uint32_t out[8]) {
    float f[16];
    memcpy(f, frac16, sizeof(f));
    if (normalize) {
        float s = 0.0f;
        for (int i = 0; i < 13; i++) s += f[i];
        if (s > 0.0f)
            for (int i = 0; i < 13; i++) f[i] /= s;
    }
    for (int k = 0; k < 8; k++) out[k] = cz_f2h(f[2 * k]) | (cz_f2h(f[2 * k + 1]) << 16);
}

}  // extern "C"
