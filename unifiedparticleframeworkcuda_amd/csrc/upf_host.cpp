// upf_host.cpp -- host-side set-up of the particle pipeline: parameter
// derivation, kernel look-up tables, the PCISPH density error factor and the
// seeded scene generators.  Nothing here runs per step; it replaces the
// reference's UnifiedConstants / WeightingKernel / Scene set-up code, which
// also runs on the host before the GPU path starts.
//
// Float expressions keep the reference's evaluation order; build with
// -ffp-contract=off.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "upf.h"
#include "upf_internal.h"

namespace upf {

// ---------------------------------------------------------------- Morton
uint32_t spread10(uint32_t v) {  // bit i of v -> bit 3i (the z_table_ of src/cpu/zIndex.cpp:10-24)
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) { return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2); }

// float -> unsigned short of the scaled coordinate; false if outside [0, 1024)
static bool grid_coord(float v, uint32_t& c) {
    if (!(v >= 0.0f) || v >= 1024.0f) { c = v >= 1024.0f ? 1023u : 0u; return false; }
    c = (uint32_t)(unsigned short)v;
    return true;
}

bool host_key(const upf_params& p, float x, float y, float z, uint32_t& key) {
    uint32_t c[3];
    bool ok = grid_coord(p.scales[0] * x, c[0]);
    ok &= grid_coord(p.scales[1] * y, c[1]);
    ok &= grid_coord(p.scales[2] * z, c[2]);
    key = morton3(c[0], c[1], c[2]);
    return ok;
}

// ------------------------------------------------------------- kernels
// WeightingKernel (src/cpu/WeightingKernel.cpp:75-263), float arithmetic with
// the reference's double literals.
static float m4(float dist, float sr) {
    float s = dist / sr;
    float factor = 2.546479089470325472f / (sr * sr * sr);
    if (dist < 0.0f || dist >= sr) return 0.0f;
    float r;
    if (s < 0.5f) r = (float)(1.0f - 6.0 * s * s + 6.0f * s * s * s);
    else { float t = 1.0f - s; r = (float)(2.0 * t * t * t); }
    return factor * r;
}
float pressure_grad(float dist, float sr) {
    if (dist == 0) return 0.0f;
    if (dist > sr) return 0.0f;
    float c = -45.f / ((float(3.1415926535897932384626) * sr * sr * sr * sr * sr * sr));
    return c / dist * (sr - dist) * (sr - dist);
}
static float visc_lap(float dist, float sr) {
    if (dist > sr) return 0.0f;
    float c = 45.f / ((float(3.1415926535897932384626) * sr * sr * sr * sr * sr * sr));
    return c * (sr - dist);
}
static float tension(float dist, float sr) {
    const float c = 32.0f / ((float(3.1415926535897932384626) * sr * sr * sr * sr * sr * sr * sr * sr * sr));
    if (dist > 0.5 * sr && dist <= sr) return c * (sr - dist) * (sr - dist) * (sr - dist) * dist * dist * dist;
    if (dist > 0 && dist <= 0.5 * sr)
        return c * (2.0f * (sr - dist) * (sr - dist) * (sr - dist) * dist * dist * dist - sr * sr * sr * sr * sr * sr / 64.0f);
    return 0.0f;
}
static float adhesion(float dist, float sr) {
    const float c = 0.007f / (powf(sr, 3.25f));
    if (dist > 0.5 * sr && dist <= sr) {
        const float b = -4.0f * dist * dist / sr + 6.0f * dist - 2.0f * sr;
        return c * powf(b, 0.25f);
    }
    return 0.0f;
}

void build_luts(float sr, float* out) {
    for (int i = 0; i < 300; ++i) {
        float dist = sr * i / 300;
        out[i] = m4(dist, sr);
        out[301 + i] = pressure_grad(dist, sr);
        out[602 + i] = visc_lap(dist, sr);
        out[903 + i] = tension(dist, sr);
        out[1204 + i] = adhesion(dist, sr);
    }
    for (int t = 0; t < 5; ++t) out[t * 301 + 300] = 0.0f;
}

// smallest float T with sqrtf(T) >= h (sqrtf is correctly rounded and monotone,
// so `sqrtf(d2) < h` <=> `d2 < T`)
float sqrt_threshold(float h) {
    float t = h * h;
    while (sqrtf(t) >= h) t = nextafterf(t, 0.0f);
    while (!(sqrtf(t) >= h)) t = nextafterf(t, INFINITY);
    return t;
}

int check_params(const upf_params* p) {
    if (!p) return UPF_ERR_ARG;
    // the z-index grid: block_size = 2^block_shift cells of the 1024-wide
    // grid per axis (ZIndex::SetBlockSize, src/cpu/zIndex.cpp:27-40), so a
    // cell code key >> 3 block_shift indexes a grid_dim^3 table
    const uint32_t b = p->block_size;
    if (p->block_shift < 1 || p->block_shift > 9 || b != (1u << p->block_shift)) return UPF_ERR_ARG;
    if (p->grid_dim != 1024u / b) return UPF_ERR_ARG;
    // the list builds take candidates from the 3x3x3 cells around a particle's
    // cell: the query box [x - h, x + h] must not reach a fourth cell, so a
    // cell is wider than the scaled support radius by more than rounding
    for (int a = 0; a < 3; ++a)
        if (!((float)b > p->scales[a] * p->support_radius * 1.0001f)) return UPF_ERR_ARG;
    if (p->vbox_min[0] != 0.0f || p->vbox_min[1] != 0.0f || p->vbox_min[2] != 0.0f) return UPF_ERR_ARG;
    if (p->max_neighbors <= 0 || p->max_neighbors > 4096) return UPF_ERR_ARG;
    if (p->add_wall_weight != 0 && p->add_wall_weight != 1) return UPF_ERR_ARG;
    if (p->add_wall_weight)
        for (int a = 0; a < 3; ++a)
            if (!(p->container_min[a] < p->container_max[a])) return UPF_ERR_ARG;
    return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

// UnifiedConstants ctor + setParticleSpacing + computeGasConstAndTimeStep
// (src/sceneConfigure/UnifiedConstants.cpp:7-191,241-306)
int upf_params_init(float spacing, upf_params* p) {
    if (!p || !(spacing > 0.0f)) return UPF_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->rest_density = 1000.0f;
    p->gravity = 9.81f;
    p->gamma = 7.0f;
    p->spacing = spacing;
    p->particle_radius = spacing * 0.5f;
    p->support_radius = (float)(spacing * 2.0 * 1.00001);
    p->force_distance = (float)(0.25 * p->support_radius);
    p->initial_mass = spacing * spacing * spacing * p->rest_density;
    p->visc_const = 1.0f;
    p->max_density_error_allowed = 1.0f;
    p->dt = 0.0012f;                       // deltaT_pcisph
    p->gas_const_wcsph = 2000.0f;
    p->gas_const_sph = 1000.0f;  // setSimulationMethod (UnifiedConstants.cpp:260,280)
    {
        const float f = p->dt / 0.0002f;
        p->density_error_factor_parameter = (float)(0.05 * f * f);
    }
    p->spring_coefficient = 300.0f;
    p->spring_coefficient_boundary = 300.0f;
    p->damping_coefficient = 0.5f;
    p->surface_tension_gamma = 0.3f;
    p->surface_adhesion_beta = 0.0f;
    p->max_boundary_force = 1.0f * 64 * spacing * spacing;
    p->min_loops = 3;
    p->max_loops = 50;
    p->terminal_speed = 20.0f;
    p->boundary_pressure_mode = UPF_BOUNDARY_CPU_REF;
    p->renormalize_quaternion = 0;
    p->rigid_predicted_current = 0;
    p->kernel_self = m4(0.0f, p->support_radius);
    p->support_radius_sq_limit = sqrt_threshold(p->support_radius);
    p->max_neighbors = 160;
    return UPF_OK;
}

int upf_params_set_domain(upf_params* p, const float vmax[3], const float cmin[3], const float cmax[3]) {
    if (!p || !vmax || !cmin || !cmax) return UPF_ERR_ARG;
    float bs = 0.0f;
    for (int a = 0; a < 3; ++a) {
        if (!(vmax[a] > 0.0f)) return UPF_ERR_ARG;
        p->vbox_min[a] = 0.0f;
        p->vbox_max[a] = vmax[a];
        p->scales[a] = 1024.0f / (vmax[a] - 0.0f);
        p->collision_min[a] = cmin[a];
        p->collision_max[a] = cmax[a];
        // setBoundingBoxValues: collisionBox = realBoxContainer -/+ particleRadius (:374-375)
        p->container_min[a] = cmin[a] - p->particle_radius;
        p->container_max[a] = cmax[a] + p->particle_radius;
        const float s = p->scales[a] * p->support_radius;
        bs = a == 0 ? s : (bs >= s ? bs : s);
    }
    // ZIndex::SetBlockSize takes an int: the scaled support radius is truncated
    const int size = (int)bs;
    uint32_t b = 1;
    while ((int)b < size) b *= 2;
    p->block_size = b;
    uint32_t sh = 0;
    while ((1u << sh) < b) ++sh;
    p->block_shift = sh;
    p->grid_dim = 1024u / b;
    // a cell must hold the support radius with room for rounding (a query box
    // then spans at most 3 cells per axis)
    if (b < 2 || !((float)b > bs * 1.0001f)) return UPF_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (!(cmin[a] >= 0.0f) || !(cmax[a] * p->scales[a] < 1024.0f) || !(cmin[a] < cmax[a])) return UPF_ERR_ARG;
    return UPF_OK;
}

int upf_build_luts(const upf_params* p, float* out) {
    if (!p || !out) return UPF_ERR_ARG;
    build_luts(p->support_radius, out);
    return UPF_OK;
}

// ComputeDensityErrorFactor (src/cpu/UnifiedPhysics.cpp:5525-5689).  The
// prototype block of CreatePreParticlesISPH (:5577-5606) is sorted by its
// z-index keys; its neighbour lists are found by exhaustive search in that
// order, which yields the same ascending lists as the reference's box query.
int upf_compute_density_error_factor(upf_params* p, int* max_neighbors_out) {
    if (!p) return UPF_ERR_ARG;
    const float spacing = p->spacing;
    if (spacing == 0.0125f) return UPF_ERR_ARG;  // the reference reads uninitialised bounds for this spacing
    const float scale = (float)(spacing / 0.0125);
    float mn[3] = {-0.1f * scale, 0.0f * scale, -0.1f * scale};
    float mx[3] = {0.1f * scale, 0.2f * scale, 0.1f * scale};
    int num[3];
    for (int a = 0; a < 3; ++a) num[a] = (int)((mx[a] - mn[a]) / spacing);
    struct Q { float x, y, z; uint32_t key; };
    std::vector<Q> q;
    q.reserve((size_t)num[0] * num[1] * num[2]);
    for (int iy = 0; iy < num[1]; ++iy) {
        const float y = 0.0f + iy * spacing;
        for (int ix = 0; ix < num[0]; ++ix) {
            const float x = 0.0f + ix * spacing;
            for (int iz = 0; iz < num[2]; ++iz) {
                const float z = 0.0f + iz * spacing;
                Q e{x, y, z, 0};
                if (!host_key(*p, x, y, z, e.key)) return UPF_ERR_RANGE;
                q.push_back(e);
            }
        }
    }
    std::stable_sort(q.begin(), q.end(), [](const Q& a, const Q& b) { return a.key < b.key; });
    const int n = (int)q.size();
    const float h = p->support_radius;
    std::vector<float> sx(n, 0.0f), sy(n, 0.0f), sz(n, 0.0f), sd(n, 0.0f);
    std::vector<int> cnt(n, 0);
    auto dist = [&](int i, int j) {
        const float dx = q[i].x - q[j].x, dy = q[i].y - q[j].y, dz = q[i].z - q[j].z;
        return sqrtf(dx * dx + dy * dy + dz * dz);
    };
    // ComputeGradWValues (:5546-5575): each pair is visited from both sides
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const float d = dist(i, j);
            if (!(d < h)) continue;
            ++cnt[i];
            const float g = pressure_grad(d, h);
            const float gx = (q[i].x - q[j].x) * g, gy = (q[i].y - q[j].y) * g, gz = (q[i].z - q[j].z) * g;
            sx[i] = sx[i] + gx; sy[i] = sy[i] + gy; sz[i] = sz[i] + gz;
            sx[j] = sx[j] - gx; sy[j] = sy[j] - gy; sz[j] = sz[j] - gz;
            const float gg = gx * gx + gy * gy + gz * gz;
            sd[i] = sd[i] + gg;
            const float nx = gx * -1.0f, ny = gy * -1.0f, nz = gz * -1.0f;
            sd[j] = sd[j] + (nx * nx + ny * ny + nz * nz);
        }
    }
    // ComputeFactor (:5638-5689)
    int best = 0, maxn = 0;
    for (int i = 0; i < n; ++i)
        if (cnt[i] > maxn) { maxn = cnt[i]; best = i; }
    const float f0 = p->dt / 0.0002f;
    p->density_error_factor_parameter = (float)(0.05 * f0 * f0);
    const float restVol = p->initial_mass / p->rest_density;
    const float pre = restVol * restVol * p->dt * p->dt;
    const float ax = sx[best] * -1.0f, ay = sy[best] * -1.0f, az = sz[best] * -1.0f;
    const float gw = (ax * sx[best] + ay * sy[best] + az * sz[best]) - sd[best];
    const float divisor = pre * gw;
    if (max_neighbors_out) *max_neighbors_out = maxn;
    if (divisor == 0) return UPF_ERR_DELTA;
    float f = (float)(-1.0 / divisor);
    f *= p->density_error_factor_parameter;
    p->density_error_factor = f;
    return UPF_OK;
}

}  // extern "C"
