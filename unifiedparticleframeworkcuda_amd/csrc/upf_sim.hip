// upf_sim.hip -- the per-step particle pipeline on gfx950 and its C ABI.
//
// One step (PCISPH, static boundaries), all on one HIP stream:
//   radix sort of z-index keys -> gather of the particle state into key order
//   -> dense cell table (first/end slot per occupied cell, by boundary
//   detection, no atomics) -> neighbour lists + external forces + first
//   prediction (one pass over the query-box candidates) -> K x [predicted
//   density & pressure with the convergence test on the device; correction
//   pressure force fused with the next prediction] -> integration + new keys.
//
// Float arithmetic follows the reference CPU path (src/cpu/UnifiedPhysics.cpp)
// operation by operation, neighbours are visited in ascending sorted order
// like the reference's neighbour lists, and the file is compiled with
// -ffp-contract=off and correctly rounded division / sqrt, so results match
// the CPU restatement bit for bit where the reference's arithmetic is IEEE.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "upf.h"
#include "upf_device.h"
#include "upf_internal.h"
#include "upf_kernels.h"
#include "upf_sort.h"

namespace upf {

__device__ inline float lut_at(const float* lut, float dist, float factor) {
    // kernel*Lut(dist): index = dist * (1/h) * kLutSize (src/cpu/WeightingKernel.cpp:122-127)
    const int idx = (int)(dist * factor * 300.0f);
    return idx >= 300 ? 0.0f : lut[idx];
}

// DistanceToWall (src/cpu/UnifiedPhysics.cpp:5926-5988) = DistanceToWallDevice
// (src/gpu/particles_kernel.cuh:612-650): distance to the nearest face of the
// container box (0 outside it), at most the box's longest side.
__device__ inline float wall_distance(const upf_params& P, float x, float y, float z) {
    const float* lo = P.container_min;
    const float* hi = P.container_max;
    float r = fmaxf(hi[0] - lo[0], fmaxf(hi[1] - lo[1], hi[2] - lo[2]));
    r = fminf(r, fmaxf(y - lo[1], 0.0f));
    r = fminf(r, fmaxf(hi[1] - y, 0.0f));
    r = fminf(r, fmaxf(x - lo[0], 0.0f));
    r = fminf(r, fmaxf(hi[0] - x, 0.0f));
    r = fminf(r, fmaxf(z - lo[2], 0.0f));
    r = fminf(r, fmaxf(hi[2] - z, 0.0f));
    return r;
}

// WallWeight (src/cpu/UnifiedPhysics.cpp:6032-6060) / WallWeightDevice
// (src/gpu/particles_kernel.cuh:652-693): m times the kernel sum over ten
// virtual wall particles, two layers below the wall (d1, d2) and one and two
// spacings aside (both sides).  The table index is the CPU path's
// (int)(d * (1/h) * 300) when CPU_FORM, else the GPU path's (d * 300) * (1/h).
template <bool CPU_FORM>
__device__ inline float wall_weight(const upf_params& P, float dist, float inv_h, const float* lw) {
    if (dist >= P.support_radius - P.particle_radius) return 0.0f;
    const float s = P.spacing, s2 = s * 2.0f;
    const float d1 = dist + P.particle_radius, d2 = d1 + s;
    const float d3 = sqrtf(d1 * d1 + s * s), d4 = sqrtf(d1 * d1 + s2 * s2);
    const float d5 = sqrtf(d2 * d2 + s * s), d6 = sqrtf(d2 * d2 + s2 * s2);
    auto k = [&](float d) {
        const int idx = CPU_FORM ? (int)(d * inv_h * 300.0f) : (int)(d * 300.0f * inv_h);
        return idx >= 300 ? 0.0f : lw[idx];
    };
    return P.initial_mass * (k(d1) + k(d2) + 2.0f * (k(d3) + k(d4) + k(d5) + k(d6)));
}

__device__ inline int ptype(const float4& x) { return __float_as_int(x.w); }

// The wave's lanes (among the active ones) whose predicate holds.  The
// builtin takes the predicate as a condition, so a compare result is the mask
// itself (no select to an integer and compare back).
__device__ __forceinline__ uint64_t wballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ inline void set_err(Ctl* ctl, int code) { atomicOr(&ctl->err, 1u << code); }

struct StepArgs {
    upf_params P;        // physics shared by all scenes
    float lut_factor;    // (float)(1.0 / h)
    float inv_mass_f;    // 1.0f / m (vmmlib operator/(float))
    float inv_mass_d;    // (float)(1.0 / m) (static const float invMass, :3714)
    float rest_vol;      // m / rho0
    float dens_sq;       // rho0 * rho0
    int n;
    int stride;          // column stride of the neighbour lists
    int nbr_ld;          // row pitch of the slot-major packed lists (max_neighbors rounded up to 4)
    Dom dom0;            // the domain when there is one scene
    const Dom* doms;     // per-scene domains (batched form), else null
    const int* scene_of; // scene of every slot (batched form), else null
    SceneCtl* sctl;      // per-scene loop control
    uint32_t* max_slots; // per-scene max slots (x 32 uints of padding)
    int n_scenes;
    int slots_per_scene;
    int code_lists;      // the correction loop reads delta-coded lists (the external-force pass writes them)
    int xcd;             // remap blocks to XCDs by contiguous ranges (grids are multiples of 8)
    // batched form: the scenes' slots in chunks of 64 (a wave; chunks never
    // straddle scenes), and per PCISPH iteration parity the list of the
    // chunks whose scene runs that iteration (act_n[p] < 0: all chunks, in order)
    int n_chunks;
    const int* chunk_begin;  // first slot of chunk c
    const int* chunk_end;    // one past its last slot
    const int* sc_chunk0;    // first chunk of scene s
    const int* sc_nchunks;   // its number of chunks
    int* act;                // [2][n_chunks] chunk lists
    int* act_n;              // [2] list lengths
};

__device__ inline int scene_at(const StepArgs& A, int i) { return A.scene_of ? A.scene_of[i] : 0; }
__device__ inline Dom dom_at(const StepArgs& A, int i) { return A.scene_of ? A.doms[A.scene_of[i]] : A.dom0; }
__device__ inline int gstride() { return gridDim.x * blockDim.x; }

// Batched form, iteration t: true when the block's slots [i0, i0 + blockDim.x)
// all belong to one scene whose loop stopped before t (a scene's slots are
// contiguous), so that late iterations pass over converged scenes without
// reading their slots.  A scene's count of iterations only grows, and the
// loop test of iteration t leaves it above t if it was, so a read racing with
// that test decides the same way.
__device__ inline bool block_idle(const StepArgs& A, int i0, int t) {
    if (!A.scene_of || i0 >= A.n) return false;
    const int s0 = A.scene_of[i0];
    const int s1 = A.scene_of[min(i0 + (int)blockDim.x, A.n) - 1];
    return s0 == s1 && t >= A.sctl[s0].want;
}

// Workgroups go to the 8 XCDs round-robin, each XCD with its own L2.  With the
// grid a multiple of 8, remap gives every XCD one contiguous eighth of the
// block range, so that blocks of neighbouring slots -- which gather the same
// neighbour data -- share an L2.
__device__ inline int xblk(int remap) {
    if (!remap) return blockIdx.x;
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// The same blocks in the opposite order: each XCD walks its eighth from the
// top down.  A pass that follows one walked upwards starts on the slots whose
// lists and records the previous pass left in that XCD's L2.
__device__ inline int xblk_down(int remap) {
    if (!remap) return gridDim.x - 1 - blockIdx.x;
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (per - 1 - (blockIdx.x >> 3));
}

// ------------------------------------------------------------------ keys
__global__ void k_keys(StepArgs A, const float4* __restrict__ pos, uint32_t* __restrict__ keys, Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float4 x = pos[i];
    const Dom d = dom_at(A, i);
    bool ok;
    keys[i] = dkey(x.x, x.y, x.z, d.scales, ok);
    if (!ok) set_err(ctl, UPF_ERR_RANGE);
}

// --------------------------------------------------------------- gather
// CopySortedParticleValuesKernel* (src/gpu/particles_kernel.cu:544-654):
// the persistent particle state follows the sort permutation.  Pressures of
// the finished loop move one generation back (p_last -> p_prev).
__global__ void k_gather(int n, const uint32_t* __restrict__ perm, Slots src, Slots dst, const float* __restrict__ p_work,
                         const float* __restrict__ rho_work, int rigid) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int o = (int)perm[s];
    dst.pos[s] = src.pos[o];
    dst.vel[s] = src.vel[o];
    dst.id[s] = src.id[o];
    dst.p_prev[s] = src.p_last[o];
    dst.p_last[s] = p_work[o];
    dst.rho_last[s] = rho_work[o];
    dst.psi[s] = src.psi[o];
    if (rigid) {
        dst.x0[s] = src.x0[o];
        dst.body[s] = src.body[o];
        dst.rel[s] = src.rel[o];
    }
}

// ------------------------------------------------------------ cell table
__global__ void k_fill_cells(uint2* __restrict__ cells, uint32_t ncells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncells) cells[c] = make_uint2(0u, 0u);
}

// first / end slot of every occupied cell by boundary detection over the
// sorted keys (BlockGeneration, src/gpu/particles.cu:904-964, without atomics)
__global__ void k_cell_bounds(StepArgs A, const uint32_t* __restrict__ keys, uint2* __restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const Dom d = dom_at(A, i);
    const uint32_t shift3 = 3u * d.block_shift;
    const uint32_t c = keys[i] >> shift3;
    uint2* T = cells + d.cell_off;
    if (i == d.slot_off || (keys[i - 1] >> shift3) != c) T[c].x = (uint32_t)i;
    if (i == d.slot_off + d.n - 1 || (keys[i + 1] >> shift3) != c) T[c].y = (uint32_t)(i + 1);
}

// Visit the candidates of the reference's query box around x in ascending
// slot order; f(j) is called for every j in the covered cells.
template <class F>
__device__ __forceinline__ void for_each_candidate(const float4& x, float h, const Dom& d, const uint2* __restrict__ cells,
                                                   F f) {
    BoxCells bc;
    bc.init(x.x, x.y, x.z, h, d.scales, d.block_shift);
    const int cnt = bc.count();
    const uint2* T = cells + d.cell_off;
    for (int e = 0; e < cnt; ++e) {
        uint32_t c[3];
        if (!bc.cell(e, c)) continue;
        if (c[0] >= d.grid_dim || c[1] >= d.grid_dim || c[2] >= d.grid_dim) continue;
        const uint2 r = T[dmorton3(c[0], c[1], c[2])];
        for (uint32_t j = r.x; j < r.y; ++j) f((int)j);
    }
}

// --------------------------------------------------- boundary volume (Psi)
// CalculateWeightedVolume (src/cpu/UnifiedPhysics.cpp:2703-2746): Psi_b = 1 / sum_b W
__global__ void k_psi(StepArgs A, const float4* __restrict__ pos, const uint2* __restrict__ cells,
                      const float* __restrict__ luts, float* __restrict__ psi, Ctl* ctl) {
    __shared__ float lw[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) lw[t] = luts[t];
    __syncthreads();
    const upf_params& P = A.P;
    const float lut_factor = A.lut_factor;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const Dom d = dom_at(A, i);
    const float4 xi = pos[i];
    const int ti = ptype(xi);
    if (ti != UPF_FROZEN && ti != UPF_RIGID) return;
    const float T = P.support_radius_sq_limit;
    // Akinci et al. 2012 eq. 4 sums over the boundary neighbourhood including the
    // particle itself (W(0) = the table's entry 0); the CPU path excludes it
    float inv = P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI ? lw[0] : 0.0f;
    for_each_candidate(xi, P.support_radius, d, cells, [&](int j) __attribute__((always_inline)) {
        if (j == i) return;
        const float4 xj = pos[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 < T)) return;
        const int tj = ptype(xj);
        if (tj == UPF_FROZEN || tj == UPF_RIGID) inv += lut_at(lw, sqrtf(d2), lut_factor);
    });
    // CPU path: 1 / sum W (:2735); Akinci mode: the GPU path's factor 0.7
    // (src/gpu/particles_kernel.cu:5927-5946) on the self-inclusive sum
    if (inv != 0.0f) psi[i] = P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI ? (float)(0.7 / (double)inv) : (float)(1.0 / (double)inv);
    else set_err(ctl, UPF_ERR_ISOLATED);
}


// BoxBoundaryForce (src/cpu/UnifiedPhysics.cpp:3794-3842): the optional wall
// force of the static drivers (addBoundaryForce), twice as strong at the floor
__device__ inline void box_force_static(const upf_params& P, const Dom& d, const float4& x, float& fx, float& fy,
                                        float& fz) {
    const float fd = P.force_distance;
    const float inv = (float)(1.0 / (double)P.force_distance);
    const float F = P.max_boundary_force;
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    if (x.y < d.cmin[1] + fd) by = by + ((d.cmin[1] + fd - x.y) * inv * 2.0f) * F;
    else if (x.y > d.cmax[1] - fd) by = by + ((x.y - (d.cmax[1] - fd)) * inv) * F * -1.0f;
    if (x.x < d.cmin[0] + fd) bx = bx + ((d.cmin[0] + fd - x.x) * inv) * F;
    if (x.x > d.cmax[0] - fd) bx = bx + ((x.x - (d.cmax[0] - fd)) * inv) * F * -1.0f;
    if (x.z < d.cmin[2] + fd) bz = bz + ((d.cmin[2] + fd - x.z) * inv) * F;
    if (x.z > d.cmax[2] - fd) bz = bz + ((x.z - (d.cmax[2] - fd)) * inv) * F * -1.0f;
    fx = fx + bx;
    fy = fy + by;
    fz = fz + bz;
}

// Predicted position + CollisionHandling (src/cpu/UnifiedPhysics.cpp:5426-5454, :5739-5757)
__device__ inline float4 predict(const float4& x, const float4& v, const float4& F, const float4& Fp,
                                 const StepArgs& A, const Dom& d, Ctl* ctl) {
    const upf_params& P = A.P;
    float ax = F.x + Fp.x, ay = F.y + Fp.y, az = F.z + Fp.z;
    ax = ax * P.dt; ay = ay * P.dt; az = az * P.dt;
    ax = ax * A.inv_mass_f; ay = ay * A.inv_mass_f; az = az * A.inv_mass_f;
    const float vx = v.x + ax, vy = v.y + ay, vz = v.z + az;
    float px = x.x + vx * P.dt, py = x.y + vy * P.dt, pz = x.z + vz * P.dt;
    px = px < d.cmin[0] ? d.cmin[0] : px;
    px = px > d.cmax[0] ? d.cmax[0] : px;
    py = py < d.cmin[1] ? d.cmin[1] : py;
    py = py > d.cmax[1] ? d.cmax[1] : py;
    pz = pz < d.cmin[2] ? d.cmin[2] : pz;
    pz = pz > d.cmax[2] ? d.cmax[2] : pz;
    if (!(px == px) || !(py == py) || !(pz == pz)) set_err(ctl, UPF_ERR_NAN);
    return make_float4(px, py, pz, x.w);
}

// ------------------------------------------------------- neighbour lists
// Entries pack the neighbour slot and the kernel-table bin of the pair's
// distance at the positions of the step: (slot << 9) | bin.  The bin is the
// index of kernel*Lut(dist) (src/cpu/WeightingKernel.cpp:122-127), 300 when the
// table returns 0; the passes that work on the step's positions read their
// kernel values from it without recomputing the distance.
constexpr int NB_SLOT_SHIFT = 9;
constexpr uint32_t NB_BIN_MASK = 511u;

__device__ inline uint32_t compact10(uint32_t v) {  // inverse of dspread10
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0x030000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

__device__ inline uint32_t lut_bin(float dist, float factor) {
    const int idx = (int)(dist * factor * 300.0f);
    return idx >= 300 ? 300u : (uint32_t)idx;
}

// lut_bin(sqrt(d2)) with the correctly rounded square root, without
// computing that root.  The bin is a non-decreasing step function of d2, so
// it is fixed by its thresholds thr[k] = min{ d2 : bin(d2) >= k } (k = 1 ..
// 300; thr[0] = 0, thr[301] = +inf), which the host finds by bisection over
// the float bit patterns.  The hardware square root (within 1 ulp) gives a
// bin b0 at most one away from the true one, and the two thresholds around
// b0 (th[b0] = (thr[b0], thr[b0 + 1])) settle it.  upf_create checks this
// equality on the device for every float below 2 thr[300].
constexpr int LUT_TH_OFF = 1508;  // float offset of th[] after the 5 tables (8-byte aligned)
constexpr int LUT_TOTAL = LUT_TH_OFF + 2 * LUT_N;
__device__ __forceinline__ uint32_t bin_of_d2(float d2, float factor, const float2* th) {
    const uint32_t b0 = lut_bin(__builtin_amdgcn_sqrtf(d2), factor);
    const float2 t = th[b0];
    return b0 - (d2 < t.x ? 1u : 0u) + (d2 >= t.y ? 1u : 0u);
}

// Counts the floats d2 in [0, limit) whose bin_of_d2 differs from
// lut_bin(sqrt_rn(d2)).
__global__ void k_check_bins(float factor, const float* __restrict__ luts, uint32_t limit_bits,
                             unsigned long long* bad) {
    __shared__ float2 th[LUT_N];
    for (int k = threadIdx.x; k < LUT_N; k += blockDim.x)
        th[k] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[k];
    __syncthreads();
    unsigned long long n_bad = 0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < limit_bits; u += gridDim.x * blockDim.x) {
        const float d2 = __uint_as_float(u);
        n_bad += bin_of_d2(d2, factor, th) != lut_bin(sqrtf(d2), factor) ? 1ull : 0ull;
    }
    if (n_bad) atomicAdd(bad, n_bad);
}

// Delta coding of a list for the passes that re-read it every iteration:
// the list is ascending, so consecutive slots mostly differ by less than
// 2^16; larger jumps (between stencil cells far apart in Morton order) go to
// a per-particle escape column.  Eight 16-bit increments share one 16-byte
// load, so a reader fetches eight entries per memory round trip.
struct ListEncoder {
    uint32_t* first;
    uint4* rows;
    uint32_t* e;
    size_t st;
    uint32_t prev = 0;
    int esc = 0;
    uint64_t lo = 0, hi = 0;  // the row being filled: a 128-bit shift register of 16-bit words
    int nw = 0, row = 0;
    __device__ ListEncoder(const Work& W, int i, size_t stride)
        : first(W.nbf + i), rows(W.nbw + i), e(W.nbe + i), st(stride) {}
    __device__ void push(uint32_t w) {
        lo = (lo >> 16) | (hi << 48);
        hi = (hi >> 16) | ((uint64_t)w << 48);
        if (++nw == 8) {
            rows[(size_t)row * st] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
            ++row;
            nw = 0;
        }
    }
    __device__ void put(int k, uint32_t j, Ctl* ctl) {
        uint32_t w = 0;
        if (k == 0) {
            *first = j;
        } else {
            const uint32_t d = j - prev;
            if (d <= 65535u) {
                w = d;
            } else {
                if (esc < NB_ESC) e[(size_t)esc * st] = d;
                else set_err(ctl, UPF_ERR_NBR_OVERFLOW);
                ++esc;
            }
        }
        prev = j;
        push(w);
    }
    __device__ void finish() {
        while (nw != 0) push(0u);
    }
};

struct ListDecoder {
    const uint4* rows;
    const uint32_t* e;
    size_t st;
    uint32_t j;
    int esc = 0;
    __device__ ListDecoder(const Work& W, int i, size_t stride)
        : rows(W.nbw + i), e(W.nbe + i), st(stride), j(W.nbf[i]) {}
    // the 16-byte word of entries k0 .. k0 + 7 (k0 a multiple of 8)
    __device__ void load(int k0, uint64_t& lo, uint64_t& hi) const {
        const uint4 r = rows[(size_t)(k0 >> 3) * st];
        lo = (uint64_t)r.x | ((uint64_t)r.y << 32);
        hi = (uint64_t)r.z | ((uint64_t)r.w << 32);
    }
    // slot of entry k0 + u (entries decoded in order)
    __device__ int next(int k0, int u, uint64_t lo, uint64_t hi) {
        if (k0 + u > 0) {
            const uint32_t wd = (uint32_t)(((u < 4 ? lo : hi) >> (16 * (u & 3))) & 0xFFFFu);
            // more escapes than NB_ESC cannot occur (one per gap between the
            // <= 27 stencil cells); the encoder flags it, the clamp keeps the
            // read inside the column
            j += wd ? wd : e[(size_t)min(esc++, NB_ESC - 1) * st];
        }
        return (int)j;
    }
};

// Visits the list of slot i in order: load(j) for G neighbours is issued
// together, then use(j, entry, value) runs for them in list order.  CODED:
// the delta-coded lists (entry = 0; eight entries per list load), else the
// packed 32-bit lists (entry = (slot << 9) | bin).
// base[j] addressed with a 32-bit byte offset (slots < 2^23, records <= 16
// bytes), which lets the load take its base from scalar registers
template <class T>
__device__ __forceinline__ T at_u32(const T* __restrict__ base, int j) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint32_t)j * (uint32_t)sizeof(T));
}

template <int CODED, int G, class Load, class Use>
__device__ __forceinline__ void for_each_listed(const Work& W, int i, size_t st, int cnt, Load load, Use use) {
    using V = decltype(load(0));
    if (CODED == 2) {
        // union indices: eight 16-bit entries per row, each the neighbour's
        // place in the chunk's union (load(u) reads it from LDS)
        const uint4* rows = W.nbu + i;
        int k0 = 0;
        for (; k0 + 8 <= cnt; k0 += 8) {
            const uint4 r = rows[(size_t)(k0 >> 3) * st];
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
            int u[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) u[q] = (int)((w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu);
#pragma unroll
            for (int h = 0; h < 8; h += G) {
                V v[G];
#pragma unroll
                for (int q = 0; q < G; ++q) v[q] = load(u[h + q]);
#pragma unroll
                for (int q = 0; q < G; ++q) use(u[h + q], 0u, v[q]);
            }
        }
        if (k0 < cnt) {
            const uint4 r = rows[(size_t)(k0 >> 3) * st];
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
            for (int q = 0; k0 + q < cnt; ++q) {
                const int u = (int)((w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu);
                use(u, 0u, load(u));
            }
        }
    } else if (CODED) {
        ListDecoder dec(W, i, st);
        int k0 = 0;
        // full rows: the next row's load is issued before this row's gathers
        uint64_t lo = 0, hi = 0;
        if (cnt > 0) dec.load(0, lo, hi);
        for (; k0 + 8 <= cnt; k0 += 8) {
            uint64_t nlo = 0, nhi = 0;
            if (k0 + 8 < cnt) dec.load(k0 + 8, nlo, nhi);
            int j[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) j[u] = dec.next(k0, u, lo, hi);
#pragma unroll
            for (int h = 0; h < 8; h += G) {
                V v[G];
#pragma unroll
                for (int u = 0; u < G; ++u) v[u] = load(j[h + u]);
#pragma unroll
                for (int u = 0; u < G; ++u) use(j[h + u], 0u, v[u]);
            }
            lo = nlo;
            hi = nhi;
        }
        // last, partial row (1 .. 7 entries): its gathers are issued together
        // as well, so that the row costs one round trip, not one per entry
        if (k0 < cnt) {
            const int rem = cnt - k0;
            int j[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) j[u] = u < rem ? dec.next(k0, u, lo, hi) : i;
#pragma unroll
            for (int h = 0; h < 8; h += G) {
                V v[G];
#pragma unroll
                for (int u = 0; u < G; ++u)
                    if (h + u < rem) v[u] = load(j[h + u]);
#pragma unroll
                for (int u = 0; u < G; ++u)
                    if (h + u < rem) use(j[h + u], 0u, v[u]);
            }
        }
    } else {
        const uint32_t* L = W.nbr + i;
        int k0 = 0;
        for (; k0 + G <= cnt; k0 += G) {
            uint32_t e[G];
            V v[G];
#pragma unroll
            for (int u = 0; u < G; ++u) e[u] = L[(size_t)(k0 + u) * st];
#pragma unroll
            for (int u = 0; u < G; ++u) v[u] = load((int)(e[u] >> NB_SLOT_SHIFT));
#pragma unroll
            for (int u = 0; u < G; ++u) use((int)(e[u] >> NB_SLOT_SHIFT), e[u], v[u]);
        }
        // the last cnt - k0 < G entries, their gathers issued together
        if (k0 < cnt) {
            const int rem = cnt - k0;
            uint32_t e[G];
            V v[G];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (u < rem) e[u] = L[(size_t)(k0 + u) * st];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (u < rem) v[u] = load((int)(e[u] >> NB_SLOT_SHIFT));
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (u < rem) use((int)(e[u] >> NB_SLOT_SHIFT), e[u], v[u]);
        }
    }
}

// Union of the neighbour lists of each chunk of UNION_CHUNK consecutive slots
// (one block per chunk), for the correction loop: the distinct slots go into
// an LDS hash table (linear probing, at most UNION_TAB probes), are sorted
// (bitonic, in LDS) and numbered in slot order -- so that the lanes of a wave,
// consecutive slots, read nearby LDS words of the union and few bank
// conflicts -- and every list is rewritten as those numbers (nbu), in list
// order.  A chunk whose union exceeds UNION_CAP keeps un = -1 and its passes
// read the delta-coded lists.  Reads the packed lists of the step (nbr); the
// sums keep their order, so results do not change.
__global__ __launch_bounds__(UNION_CHUNK) void k_union(StepArgs A, Work W) {
    __shared__ uint32_t tab[UNION_TAB];
    __shared__ uint32_t occ[UNION_TAB / 32], pre[UNION_TAB / 32];
    __shared__ uint32_t srt[UNION_CAP];
    __shared__ uint16_t rank_at[UNION_TAB];
    __shared__ int s_bad, s_total;
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    const int c = blockIdx.x, tid = threadIdx.x;
    const int i = c * UNION_CHUNK + tid;
    const size_t st = (size_t)A.stride;
    const int cnt = i < A.n ? min(W.ncount[i], A.P.max_neighbors) : 0;
    const uint32_t* L = W.nbr + i;
    for (int k = tid; k < UNION_TAB; k += UNION_CHUNK) tab[k] = EMPTY;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    auto home = [](uint32_t slot) { return (slot * 2654435761u) >> (32 - 11); };  // UNION_TAB = 2^11
    for (int k = 0; k < cnt; ++k) {
        const uint32_t slot = L[(size_t)k * st] >> NB_SLOT_SHIFT;
        uint32_t h = home(slot);
        bool done = false;
        for (int probe = 0; probe < UNION_TAB && !done; ++probe) {
            const uint32_t old = atomicCAS(&tab[h], EMPTY, slot);
            done = old == EMPTY || old == slot;
            h = (h + 1) & (UNION_TAB - 1);
        }
        if (!done) s_bad = 1;  // table full: more than UNION_TAB distinct slots
    }
    __syncthreads();
    // rank of the occupied table positions: one word of the occupancy map per
    // lane of the first wave, then a prefix over the 64 words
    if (tid < UNION_TAB / 32) {
        uint32_t w = 0;
        for (int b = 0; b < 32; ++b) w |= (tab[tid * 32 + b] != EMPTY ? 1u : 0u) << b;
        occ[tid] = w;
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t pc = tid < UNION_TAB / 32 ? (uint32_t)__popc(occ[tid]) : 0u;
        uint32_t incl = pc;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
            if (tid >= o) incl += y;
        }
        if (tid < UNION_TAB / 32) pre[tid] = incl - pc;
        if (tid == 63) s_total = (int)incl;
    }
    __syncthreads();
    const int U = s_total;
    if (s_bad || U > UNION_CAP) {
        if (tid == 0) W.un[c] = -1;
        return;
    }
    // compact, then sort ascending (padding sorts last)
    for (int k = tid; k < UNION_CAP; k += UNION_CHUNK) srt[k] = EMPTY;
    __syncthreads();
    for (int k = tid; k < UNION_TAB; k += UNION_CHUNK)
        if (tab[k] != EMPTY) srt[pre[k >> 5] + __popc(occ[k >> 5] & ((1u << (k & 31)) - 1u))] = tab[k];
    __syncthreads();
    for (int size = 2; size <= UNION_CAP; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = tid; q < UNION_CAP / 2; q += UNION_CHUNK) {
                const int a = 2 * q - (q & (stride - 1)), b = a + stride;  // a has bit `stride` clear
                const bool up = (a & size) == 0;
                const uint32_t x = srt[a], y = srt[b];
                if ((x > y) == up) {
                    srt[a] = y;
                    srt[b] = x;
                }
            }
            __syncthreads();
        }
    // number of each slot = its place in the sorted union
    for (int q = tid; q < U; q += UNION_CHUNK) {
        const uint32_t slot = srt[q];
        uint32_t h = home(slot);
        for (int probe = 0; probe < UNION_TAB && tab[h] != slot; ++probe) h = (h + 1) & (UNION_TAB - 1);
        rank_at[h] = (uint16_t)q;
        W.uslots[(size_t)c * UNION_CAP + q] = slot;
    }
    __syncthreads();
    // the list as union numbers, eight per row
    uint4* rows = W.nbu + i;
    uint32_t wd[4] = {0u, 0u, 0u, 0u};
    for (int k = 0; k < cnt; ++k) {
        const uint32_t slot = L[(size_t)k * st] >> NB_SLOT_SHIFT;
        uint32_t h = home(slot);
        for (int probe = 0; probe < UNION_TAB && tab[h] != slot; ++probe) h = (h + 1) & (UNION_TAB - 1);
        const uint32_t u = rank_at[h];
        const int q = k & 7;
        wd[q >> 1] |= u << (16 * (q & 1));
        if (q == 7 || k + 1 == cnt) {
            rows[(size_t)(k >> 3) * st] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
            wd[0] = wd[1] = wd[2] = wd[3] = 0u;
        }
    }
    if (tid == 0) W.un[c] = U;
}

// GetNeighbors (src/cpu/UnifiedPhysics.cpp:2242-2295) for every slot of type
// OWNER (liquid; rigid for the coupling step's rigid lists), one wave per 64
// consecutive slots.  For each cell met in its slots the wave lists the
// occupied cells of the 3x3x3 stencil in Morton order (so that the candidates
// come in ascending slot order) and their candidates in an LDS table (slot,
// cell rank), then tests the candidates 64 at
// a time against each particle of the cell: the candidate's cell must lie in
// the particle's query box (the reference's BoxQuery; decided once per cell
// and particle for the occupied stencil cells; SKIP: a round of 64 candidates
// none of whose cells lies in the box is passed over) and its distance must be
// below h.  Ballots keep each list in ascending order (BIN: entries carry
// the pair's table bin, else only the slot).  ROUNDS x 64
// candidates are held in registers per chunk.  The lists of the wave's 64
// slots are staged in LDS (row k, one column per slot, padded against bank
// conflicts) and leave in full 256-byte rows of the column-major layout;
// entries past the first STAGE of a list are written directly (STAGE 0: all,
// which keeps the LDS free for more waves).  SLOT_MAJOR (STAGE 0): each list
// in its own row of A.nbr_ld words instead, where a round's entries are
// consecutive words (for a reader in the same kernel).
constexpr int CAND_CAP = 512;  // stencil candidates held in the LDS candidate table
template <int ROUNDS, int STAGE, int BIN, int SKIP, int KEEP_RIGID = 0, int OWNER = UPF_LIQUID,
          bool SLOT_MAJOR = false, int SPLIT = 1>
__device__ __forceinline__ void build_wave(const StepArgs& A, const float4* __restrict__ pos,
                                           const uint32_t* __restrict__ keys, const uint2* __restrict__ cells,
                                           uint32_t* __restrict__ nbr, int* __restrict__ ncount, Ctl* ctl,
                                           uint32_t* __restrict__ cand) {
    constexpr int NB_CHUNK = 64 * ROUNDS;
    // cell coordinates and query boxes are packed in 10-bit fields (x, y, z)
    __shared__ uint32_t r_start[32], r_pref[33], r_cell[32];
    __shared__ uint32_t stage[(STAGE > 0 ? STAGE : 1) * 65];
    const upf_params& P = A.P;
    const float lut_factor = A.lut_factor;
    const int n = A.n;
    const int stride = A.stride;
    const int lane = threadIdx.x;
    // SPLIT waves share a window of 64 slots, each building the lists of the
    // owners in its 64 / SPLIT lanes (more, shorter waves where owners are
    // few and dense, as the rigid particles of the coupling step)
    const int vb = xblk(A.xcd);
    const int base = (vb / SPLIT) * 64;
    const int part = vb % SPLIT;
    if (base >= n) return;
    const int wend = min(base + 64, n);
    const float T = P.support_radius_sq_limit;
    const float h = P.support_radius;
    const int cap = P.max_neighbors;
    // the wave's 64 slots, one per lane: position, key, scene and the query
    // box of the slot's particle; the serial loop below reads them from the
    // owning lane into scalars
    const int me = base + lane;
    const bool mine_valid = me < wend;
    const float4 my_x = mine_valid ? pos[me] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t my_key = mine_valid ? keys[me] : 0xFFFFFFFFu;
    const int my_scene = mine_valid ? scene_at(A, me) : -1;
    // owners: the slots whose lists this build writes (liquid, or the rigid
    // slots of the coupling step)
    const bool my_owner = mine_valid && ptype(my_x) == OWNER && (SPLIT == 1 || lane / (64 / SPLIT) == part);
    uint32_t my_blo = 0u, my_bhi = 0u;
    if (my_owner) {
        // BoxQuery of the particle in cell coordinates, clipped to the grid
        const Dom dm = A.scene_of ? A.doms[my_scene] : A.dom0;
        const float xv[3] = {my_x.x, my_x.y, my_x.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float lo_f = dm.scales[a] * (xv[a] - h);
            float hi_f = dm.scales[a] * (xv[a] + h);
            lo_f = lo_f >= 0.0f ? lo_f : 0.0f;
            hi_f = hi_f <= 1024.0f ? hi_f : 1024.0f;
            const uint32_t l = ((uint32_t)(unsigned short)lo_f) >> dm.block_shift;
            uint32_t u = ((uint32_t)(unsigned short)hi_f) >> dm.block_shift;
            u = u < dm.grid_dim ? u : dm.grid_dim - 1;  // cells lie below G <= 512
            my_blo |= l << (10 * a);
            my_bhi |= u << (10 * a);
        }
    }
    // the key whose cell centres a slot's stencil: its sort key, or for a
    // rigid owner -- keyed at its position before the last rigid update -- the
    // key of its current position, whose cell's stencil holds its query box
    uint32_t my_seg = my_key;
    if constexpr (OWNER == UPF_RIGID) {
        if (my_owner) {
            const Dom dm = A.scene_of ? A.doms[my_scene] : A.dom0;
            bool ok;
            my_seg = dkey(my_x.x, my_x.y, my_x.z, dm.scales, ok);
        }
    }
    const uint64_t owners = wballot(my_owner);
    int my_cnt = 0;  // this lane's slot: list length so far
    int s = base;
    while (s < wend) {
        const int sl = s - base;
        // the cell's values are the same on every lane: read from lane sl into
        // scalar registers, so that the cell's set-up runs on the scalar unit
        const int sc = __builtin_amdgcn_readlane(my_scene, sl);
        const Dom d = A.scene_of ? A.doms[sc] : A.dom0;  // the cell's scene
        const uint32_t shift3 = 3u * d.block_shift;
        const uint32_t G = d.grid_dim;
        const uint2* T3 = cells + d.cell_off;
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)my_seg, sl) >> shift3;
        // end of the cell within the wave: first later slot of another cell or scene
        const uint64_t other = wballot(mine_valid && me > s && ((my_seg >> shift3) != c || my_scene != sc));
        const int pe = other ? base + __ffsll((long long)other) - 1 : wend;
        // the cell's owner slots (bits sl .. pe - base - 1)
        const uint64_t upto = (pe - base) >= 64 ? ~0ull : ((1ull << (pe - base)) - 1ull);
        const uint64_t cell_owners = owners & upto & ~((1ull << sl) - 1ull);
        if (cell_owners == 0ull) {  // no list to build in this cell (e.g. walls)
            s = pe;
            continue;
        }
        const uint32_t cc[3] = {compact10(c), compact10(c >> 1), compact10(c >> 2)};
        // stencil cells in Morton order: lane e names enumeration index e
        BoxCells bc;
        {
            // a degenerate "box" whose cells are the clipped 3x3x3 stencil
            int nn[3];
            uint32_t lo[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = cc[a] > 0 ? cc[a] - 1 : 0;
                const uint32_t hi = cc[a] + 1 < G ? cc[a] + 1 : G - 1;
                nn[a] = (int)(hi - lo[a]) + 1;
            }
            bc.init_cells(lo, nn);
        }
        uint32_t rs = 0, rl = 0, rcell = 0;
        bool valid = false;
        if (lane < bc.count()) {
            uint32_t q[3];
            if (bc.cell(lane, q)) {
                const uint2 r = T3[dmorton3(q[0], q[1], q[2])];
                valid = r.y > r.x;
                rs = r.x;
                rl = r.y - r.x;
                rcell = q[0] | (q[1] << 10) | (q[2] << 20);
            }
        }
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint64_t vm = wballot(valid);
        const int nr = __popcll(vm);
        const int rank = __popcll(vm & lt);
        // prefix of the occupied cells' lengths in rank (= lane) order
        uint32_t incl = valid ? rl : 0u;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (valid) { r_start[rank] = rs; r_cell[rank] = rcell; r_pref[rank + 1] = incl; }
        if (lane == 0) r_pref[0] = 0;
        __syncthreads();
        const int M = (int)r_pref[nr];
        // candidate table (cand, CAND_CAP entries of LDS): slot << 5 | cell rank
        // of every candidate, written by the lanes of the occupied cells, so
        // that each round reads its candidates at once; a stencil with more
        // candidates finds their cells by a walk over the prefix sums
        const bool table = M <= CAND_CAP;
        if (table) {
            const uint32_t len = valid ? rl : 0u;
            const int beg = (int)(incl - len);  // r_pref[rank] of a valid lane
            for (uint32_t t = 0; wballot(t < len) != 0ull; ++t)
                if (t < len) cand[beg + (int)t] = ((rs + t) << 5) | (uint32_t)rank;
            __syncthreads();
        }
        // Which occupied stencil cells (bit k: rank k) lie in the query box of
        // this lane's particle -- the box test of every candidate of the cell,
        // done once per cell: lane k < nr holds rank k's cell; per axis, the
        // ranks one below, at and one above the cell's own coordinate, and the
        // lane's box (blo .. bhi per field) keeps those of each class it meets.
        uint32_t in_box;
        {
            const uint32_t qk = lane < nr ? r_cell[lane] : 0u;
            in_box = (1u << nr) - 1u;  // nr <= 27
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint32_t qa = (qk >> (10 * a)) & 0x3FFu;
                const uint32_t below = (uint32_t)wballot(lane < nr && qa < cc[a]);
                const uint32_t above = (uint32_t)wballot(lane < nr && qa > cc[a]);
                const uint32_t at = (uint32_t)wballot(lane < nr && qa == cc[a]);
                const uint32_t l = (my_blo >> (10 * a)) & 0x3FFu, u = (my_bhi >> (10 * a)) & 0x3FFu;
                // stencil coordinates are cc - 1, cc, cc + 1 (cc - 1 only when cc > 0)
                if (!(cc[a] > 0 && l <= cc[a] - 1 && cc[a] - 1 <= u)) in_box &= ~below;
                if (!(l <= cc[a] && cc[a] <= u)) in_box &= ~at;
                if (!(l <= cc[a] + 1 && cc[a] + 1 <= u)) in_box &= ~above;
            }
        }
        for (int c0 = 0; c0 < M; c0 += NB_CHUNK) {
            // this lane's candidates of the chunk; a lane past the end holds
            // an infinitely distant candidate in no cell, which fails both tests
            float4 cx[ROUNDS];
            int cj[ROUNDS];
            uint32_t cbit[ROUNDS];  // 1 << rank of the candidate's cell
            uint32_t rbits[ROUNDS]; // SKIP: ranks met by the round's candidates (uniform)
            int k = 0;  // range of the lane's candidate; rises with the round
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int ci = c0 + r * 64 + lane;
                cj[r] = -1;
                cbit[r] = 0u;
                cx[r] = make_float4(__int_as_float(0x7f800000), 0.f, 0.f, 0.f);
                if (ci < M) {
                    if (table) {
                        const uint32_t v = cand[ci];
                        cj[r] = (int)(v >> 5);
                        k = (int)(v & 31u);
                    } else {
                        while (k + 1 < nr && (int)r_pref[k + 1] <= ci) ++k;
                        cj[r] = (int)r_start[k] + (ci - (int)r_pref[k]);
                    }
                    cbit[r] = 1u << k;
                    cx[r] = pos[cj[r]];
                }
                if constexpr (SKIP != 0) {
                    // the round's ranks run from its first lane's to its last lane's
                    const int last = min(63, M - 1 - (c0 + r * 64));
                    const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)cbit[r], 0);
                    const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)cbit[r], last < 0 ? 0 : last);
                    rbits[r] = last < 0 ? 0u : (b1 << 1) - b0;
                }
            }
            const int rounds = min(ROUNDS, (M - c0 + 63) >> 6);
            uint64_t pm = cell_owners;
            while (pm) {
                const int pl = __builtin_ctzll(pm);
                pm &= pm - 1ull;
                const int p = base + pl;
                const float px = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_x.x), pl));
                const float py = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_x.y), pl));
                const float pz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_x.z), pl));
                const uint32_t boxm = (uint32_t)__builtin_amdgcn_readlane((int)in_box, pl);
                int cnt = __builtin_amdgcn_readlane(my_cnt, pl);
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    if (r >= rounds) break;
                    if constexpr (SKIP != 0)
                        if ((boxm & rbits[r]) == 0u) continue;  // no candidate of the round in the box
                    const bool inbox = (boxm & cbit[r]) != 0u;
                    const float dx = px - cx[r].x, dy = py - cx[r].y, dz = pz - cx[r].z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const bool near = d2 < T, other = cj[r] != p;
                    const bool pass = inbox & near & other;
                    // the AND of three compare masks (each ballot is its compare)
                    const uint64_t m = wballot(inbox) & wballot(near) & wballot(other);
                    if (m == 0ull) continue;
                    if (pass) {
                        // place in the list: cnt + the passing lanes below this one
                        const int slot = (int)__builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)cnt));
                        const uint32_t below = (uint32_t)(slot - cnt);
                        // BIN == 0: the table bin is left to the readers (the
                        // external-force pass computes it from the same distance)
                        const uint32_t e = ((uint32_t)cj[r] << NB_SLOT_SHIFT) | (BIN ? lut_bin(sqrtf(d2), lut_factor) : 0u);
                        if constexpr (STAGE > 0) {
                            if (slot < STAGE) stage[slot * 65 + pl] = e;
                            else if (slot < cap) nbr[(size_t)slot * stride + p] = e;
                        } else if constexpr (SLOT_MAJOR) {
                            // the particle's own row (a uniform base): the
                            // round's entries are consecutive words
                            uint32_t* row = nbr + (size_t)p * A.nbr_ld;
                            if (slot < cap)
                                *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(row) + (uint32_t)slot * 4u) = e;
                        } else {
                            // row cnt of the list column, then `below` rows on
                            // (slots < 2^23: the offset fits 24-bit multiplies)
                            if (slot < cap) (nbr + (size_t)cnt * stride + p)[__umul24(below, (uint32_t)stride)] = e;
                        }
                    }
                    cnt += __popcll(m);
                }
                if (lane == pl) my_cnt = cnt;
            }
        }
        __syncthreads();  // r_* are rewritten for the next cell
        s = pe;
    }
    // KEEP_RIGID: rigid slots keep the counts of their own lists; a rigid-owner
    // build writes only those
    if (OWNER == UPF_LIQUID ? mine_valid && !(KEEP_RIGID && ptype(my_x) == UPF_RIGID) : my_owner) {
        ncount[me] = my_cnt < cap ? my_cnt : cap;
        if (my_cnt > cap) set_err(ctl, UPF_ERR_NBR_OVERFLOW);
    }
    if (STAGE > 0) {
        __syncthreads();
        const int mine = min(min(my_cnt, cap), STAGE);
        int rows = mine;
        for (int o = 32; o > 0; o >>= 1) rows = max(rows, __shfl_xor(rows, o, 64));
        for (int k = 0; k < rows; ++k)
            if (k < mine) nbr[(size_t)k * stride + me] = stage[k * 65 + lane];
    }
}

template <int ROUNDS, int STAGE, int BIN, int SKIP = 0>
__global__ __launch_bounds__(64) void k_nbr_build(StepArgs A, const float4* __restrict__ pos,
                                                  const uint32_t* __restrict__ keys, const uint2* __restrict__ cells,
                                                  uint32_t* __restrict__ nbr, int* __restrict__ ncount, Ctl* ctl) {
    __shared__ uint32_t cand[CAND_CAP];
    build_wave<ROUNDS, STAGE, BIN, SKIP>(A, pos, keys, cells, nbr, ncount, ctl, cand);
}

// The lists of the rigid slots for the coupling step (GetNeighbors of a rigid
// particle: every particle type, entries with their table bins), by the same
// wave-cooperative build; the other slots' counts are left as they are.
// SLOT_MAJOR: the lists in slot-major rows (as the fused coupling kernel
// writes and reads the liquid ones).
// The rigid particles of a body are packed densely in few windows of 64
// slots: four waves share each window (RIGID_SPLIT), so that many more waves
// run the build at once.
constexpr int RIGID_SPLIT = 4;
template <int ROUNDS, bool SLOT_MAJOR>
__global__ __launch_bounds__(64) void k_nbr_build_rigid(StepArgs A, const float4* __restrict__ pos,
                                                        const uint32_t* __restrict__ keys,
                                                        const uint2* __restrict__ cells, uint32_t* __restrict__ nbr,
                                                        int* __restrict__ ncount, Ctl* ctl) {
    __shared__ uint32_t cand[CAND_CAP];
    build_wave<ROUNDS, 0, 1, 0, 0, UPF_RIGID, SLOT_MAJOR, RIGID_SPLIT>(A, pos, keys, cells, nbr, ncount, ctl, cand);
}

// GetNeighbors again, tile form: one wave per 64 consecutive slots of one
// scene.  The occupied cells of the box around the slots' cells (grown by one
// cell) are ranked into Morton order -- the slot order -- in LDS; their
// particles are staged in LDS CH at a time, and every lane tests all of them
// against its own particle (query box of the reference's BoxQuery, distance
// below h), appending to its list in ascending slot order.  No per-particle
// serialisation; a slot range whose box exceeds TB_CELLS cells is halved
// until it fits.
constexpr int TB_CELLS = 128;

__device__ inline uint32_t wave_min_u(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ inline uint32_t wave_max_u(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

template <int CH>
__global__ __launch_bounds__(64) void k_nbr_build_tile(StepArgs A, const float4* __restrict__ pos,
                                                       const uint32_t* __restrict__ keys,
                                                       const uint2* __restrict__ cells, uint32_t* __restrict__ nbr,
                                                       int* __restrict__ ncount, Ctl* ctl) {
    __shared__ float4 cpos[CH];
    __shared__ uint32_t cpk[CH], cslot[CH];
    __shared__ uint32_t o_start[TB_CELLS], o_len[TB_CELLS], o_cell[TB_CELLS];
    __shared__ uint32_t s_start[TB_CELLS], s_cell[TB_CELLS], s_pref[TB_CELLS + 1];
    const upf_params& P = A.P;
    const int lane = threadIdx.x;
    const int base = xblk(A.xcd) * 64;
    if (base >= A.n) return;
    const int wend = min(base + 64, A.n);
    const int p = base + lane;
    const bool valid = p < wend;
    const float4 xp = valid ? pos[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool liquid = valid && ptype(xp) == UPF_LIQUID;
    const int my_scene = valid ? scene_at(A, p) : -1;
    const float T = P.support_radius_sq_limit;
    const float h = P.support_radius;
    const int cap = P.max_neighbors;
    const size_t stride = (size_t)A.stride;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int cnt = 0;
    int s = base;
    while (s < wend) {
        const int sc = scene_at(A, s);
        const Dom d = dom_at(A, s);
        const uint32_t shift3 = 3u * d.block_shift, G = d.grid_dim;
        const uint2* T3 = cells + d.cell_off;
        const uint64_t other = wballot(valid && p >= s && my_scene != sc);
        int se = other ? base + __ffsll((long long)other) - 1 : wend;
        // this lane's query box in cell coordinates (its own scene's grid)
        uint32_t blo[3] = {0u, 0u, 0u}, bhi[3] = {0u, 0u, 0u};
        if (liquid && my_scene == sc) {
            const float xv[3] = {xp.x, xp.y, xp.z};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float lo_f = d.scales[a] * (xv[a] - h);
                float hi_f = d.scales[a] * (xv[a] + h);
                lo_f = lo_f >= 0.0f ? lo_f : 0.0f;
                hi_f = hi_f <= 1024.0f ? hi_f : 1024.0f;
                blo[a] = ((uint32_t)(unsigned short)lo_f) >> d.block_shift;
                bhi[a] = ((uint32_t)(unsigned short)hi_f) >> d.block_shift;
            }
        }
        for (;;) {  // shrink [s, se) until its cell box fits
            const bool in_seg = valid && p >= s && p < se;
            const uint32_t c = in_seg ? keys[p] >> shift3 : 0u;
            const uint32_t cx = compact10(c), cy = compact10(c >> 1), cz = compact10(c >> 2);
            uint32_t lo[3], ext[3];
            const uint32_t cv[3] = {cx, cy, cz};
            uint32_t nbox = 1;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint32_t mn = wave_min_u(in_seg ? cv[a] : 0xFFFFFFFFu);
                const uint32_t mx = wave_max_u(in_seg ? cv[a] : 0u);
                lo[a] = mn > 0 ? mn - 1 : 0;
                const uint32_t hi = mx + 1 < G ? mx + 1 : G - 1;
                ext[a] = hi - lo[a] + 1;
                nbox *= ext[a];
            }
            if (nbox > (uint32_t)TB_CELLS) {
                if (se - s > 1) {
                    se = s + (se - s) / 2;
                    continue;
                }
                set_err(ctl, UPF_ERR_NBR_OVERFLOW);  // one cell's box holds <= 27 cells: cannot happen
                break;
            }
            // occupied cells of the box (unsorted), two per lane
            int n_occ = 0;
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const uint32_t q = (uint32_t)(lane + 64 * r);
                bool occ = false;
                uint2 rg = make_uint2(0u, 0u);
                uint32_t pk = 0u;
                if (q < nbox) {
                    const uint32_t qx = lo[0] + q % ext[0], qy = lo[1] + (q / ext[0]) % ext[1],
                                   qz = lo[2] + q / (ext[0] * ext[1]);
                    rg = T3[dmorton3(qx, qy, qz)];
                    occ = rg.y > rg.x;
                    pk = qx | (qy << 10) | (qz << 20);
                }
                const uint64_t m = wballot(occ);
                if (occ) {
                    const int k = n_occ + __popcll(m & lt);
                    o_start[k] = rg.x;
                    o_len[k] = rg.y - rg.x;
                    o_cell[k] = pk;
                }
                n_occ += __popcll(m);
            }
            __syncthreads();
            // rank by first slot = Morton order of the cells
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int e = lane + 64 * r;
                if (e < n_occ) {
                    const uint32_t st0 = o_start[e];
                    int rank = 0;
                    for (int f = 0; f < n_occ; ++f) rank += o_start[f] < st0;
                    s_start[rank] = st0;
                    s_cell[rank] = o_cell[e];
                    s_pref[rank + 1] = o_len[e];
                }
            }
            __syncthreads();
            // prefix of the cell lengths (two per lane, wave scan)
            {
                const uint32_t a0 = 2 * lane < n_occ ? s_pref[2 * lane + 1] : 0u;
                const uint32_t a1 = 2 * lane + 1 < n_occ ? s_pref[2 * lane + 2] : 0u;
                uint32_t incl = a0 + a1;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
                    if (lane >= o) incl += y;
                }
                const uint32_t excl = incl - (a0 + a1);
                __syncthreads();
                if (2 * lane < n_occ) s_pref[2 * lane + 1] = excl + a0;
                if (2 * lane + 1 < n_occ) s_pref[2 * lane + 2] = excl + a0 + a1;
                if (lane == 0) s_pref[0] = 0u;
            }
            __syncthreads();
            const int M = (int)s_pref[n_occ];
            for (int c0 = 0; c0 < M; c0 += CH) {
                const int chunk = min(CH, M - c0);
                for (int idx = lane; idx < chunk; idx += 64) {
                    const uint32_t g = (uint32_t)(c0 + idx);
                    int k0 = 0, k1 = n_occ;  // last cell with s_pref[k] <= g
                    while (k1 - k0 > 1) {
                        const int mid = (k0 + k1) >> 1;
                        if (s_pref[mid] <= g) k0 = mid; else k1 = mid;
                    }
                    const uint32_t slot = s_start[k0] + (g - s_pref[k0]);
                    cslot[idx] = slot;
                    cpk[idx] = s_cell[k0];
                    cpos[idx] = pos[slot];
                }
                __syncthreads();
                if (liquid && p >= s && p < se) {
                    for (int idx = 0; idx < chunk; ++idx) {
                        const uint32_t qc = cpk[idx];
                        const uint32_t qx = qc & 1023u, qy = (qc >> 10) & 1023u, qz = qc >> 20;
                        if (qx < blo[0] || qx > bhi[0] || qy < blo[1] || qy > bhi[1] || qz < blo[2] || qz > bhi[2])
                            continue;
                        const uint32_t j = cslot[idx];
                        if ((int)j == p) continue;
                        const float4 xj = cpos[idx];
                        const float dx = xp.x - xj.x, dy = xp.y - xj.y, dz = xp.z - xj.z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 < T) {
                            if (cnt < cap)
                                nbr[(size_t)cnt * stride + p] = (j << NB_SLOT_SHIFT) | lut_bin(sqrtf(d2), A.lut_factor);
                            ++cnt;
                        }
                    }
                }
                __syncthreads();
            }
            break;
        }
        s = se;
    }
    if (valid) {
        ncount[p] = cnt < cap ? cnt : cap;
        if (cnt > cap) set_err(ctl, UPF_ERR_NBR_OVERFLOW);
    }
}

// CalculateExternalForcesStaticBoundariesPCISPH (:5198-5250) over the lists +
// the first PredictPositionAndVelocity of the loop; every slot resets its
// loop pressure.
// w of a particle's pressure-force record at the start of a step.  Liquid:
// p / rho0^2 = +0 until the density pass writes it (p >= 0, so never -0).
// Other particles have p = 0; with Akinci boundaries they carry their mass
// rho0 psi as -(rho0 psi), whose sign bit marks them (psi >= 0).
template <int MODE>
__device__ __forceinline__ float pressure_record_w(int type, float psi, float rho0) {
    if (MODE == UPF_BOUNDARY_AKINCI && type != UPF_LIQUID) return -(rho0 * psi);
    // a liquid particle's w is p / rho0^2 >= +0 (+0 until the first density
    // pass); any other particle's has its sign bit set, which is how the
    // pressure-force pass tells them apart (-0 adds as +0 to a pressure term)
    return type == UPF_LIQUID ? 0.0f : -0.0f;
}

// The predicted-position record of a non-liquid particle: with Akinci
// boundaries its w carries the particle's mass in the density sums, negated,
// -(rho0 psi) (not the type bits of a liquid particle), so that the density
// pass reads it with the position instead of gathering psi.
template <int MODE>
__device__ __forceinline__ float4 boundary_xs(float4 x, float psi, float rho0) {
    if (MODE == UPF_BOUNDARY_AKINCI) x.w = -(rho0 * psi);
    return x;
}

// The external-force work of slot i (one thread), shared by the
// external-force pass and the fused list build: per-step records of the slot,
// and for a liquid slot the forces over its list, the first predicted
// position and the delta-coded copy of the list.  lvl, lpg, th: the
// viscosity and pressure-gradient tables and the bin thresholds in LDS.
template <int MODE, bool SLOT_MAJOR = false>
__device__ __forceinline__ void ext_static_slot(const StepArgs& A, const Slots& S, const Work& W, int i,
                                                const float* lvl, const float* lpg, const float2* th, Ctl* ctl) {
    const upf_params& P = A.P;
    const float4 xi = S.pos[i];
    const int ti = ptype(xi);
    W.p[i] = 0.0f;
    W.xpd[i] = make_float4(xi.x, xi.y, xi.z, pressure_record_w<MODE>(ti, S.psi[i], P.rest_density));
    W.fp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    W.rho[i] = S.rho_last[i];
    if (ti != UPF_LIQUID) {
        W.xs[i] = boundary_xs<MODE>((ti == UPF_RIGID && !P.rigid_predicted_current) ? S.x0[i] : xi, S.psi[i],
                                    P.rest_density);
        W.force[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4 vi = S.vel[i];
    const float h = P.support_radius;
    const float rv = A.rest_vol, visc = P.visc_const;
    const float own_p = MODE == UPF_BOUNDARY_CPU_REF ? S.p_prev[i] : S.p_last[i];
    const float inv = 1.0f / S.rho_last[i];
    const float pg = own_p * inv * inv;
    const float own_psi = S.psi[i];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int cnt = W.ncount[i];
    // the packed list: column-major, or (SLOT_MAJOR) the slot's own row
    const uint32_t* L = SLOT_MAJOR ? W.nbr + (size_t)i * A.nbr_ld : W.nbr + i;
    const bool bin_free = A.code_lists != 0;  // the lists of this step carry no bins
    const size_t st = (size_t)A.stride;  // column stride (the coded lists; column-major packed lists)
    auto pair = [&](uint32_t e, const float4& xj, const float4& vj) __attribute__((always_inline)) {
        const int tj = ptype(xj);
        if (tj == UPF_LIQUID) {
            uint32_t bin = e & NB_BIN_MASK;
            if (bin_free) {  // the list build's expression on the same positions
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                bin = bin_of_d2(dx * dx + dy * dy + dz * dz, A.lut_factor, th);
            }
            const float kv = lvl[bin];
            const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
            fx = fx - vx * rv * rv * visc * kv;
            fy = fy - vy * rv * rv * visc * kv;
            fz = fz - vz * rv * rv * visc * kv;
        } else if (tj == UPF_FROZEN) {
            // CalculateBoundaryFluidPressureForceHost (:4240-4250).  The CPU path
            // passes the integer bin as a distance to the table look-up, which
            // always yields the table's +0 (entry 0 or past the end).
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            float g, wv;
            if (MODE == UPF_BOUNDARY_CPU_REF) {
                g = 0.0f;
                wv = own_psi;
            } else {
                const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                uint32_t dl = (uint32_t)(dist * 300.0f / h);
                dl = dl > 300u ? 300u : dl;
                g = lpg[dl];
                wv = S.psi[(int)(e >> NB_SLOT_SHIFT)];
            }
            const float s1 = (((dx * g) * pg) * P.initial_mass) * P.rest_density;
            const float s2 = (((dy * g) * pg) * P.initial_mass) * P.rest_density;
            const float s3 = (((dz * g) * pg) * P.initial_mass) * P.rest_density;
            fx = fx + (s1 * wv) * -1.0f;
            fy = fy + (s2 * wv) * -1.0f;
            fz = fz + (s3 * wv) * -1.0f;
        }
    };
    ListEncoder enc(W, i, st);
    const bool code = A.code_lists != 0;
    // groups of four entries; the next group's entries are loaded before this
    // group's gathers, and the last, partial group's gathers go out together
    // (a slot-major row holds its groups as aligned 16-byte words; entries
    // past the list's end are loaded but never used)
    auto group = [&](int k0, uint32_t* out) __attribute__((always_inline)) {
        if constexpr (SLOT_MAJOR) {
            uint4 q = make_uint4(0u, 0u, 0u, 0u);
            if (k0 < cnt) q = *reinterpret_cast<const uint4*>(L + k0);
            out[0] = q.x;
            out[1] = q.y;
            out[2] = q.z;
            out[3] = q.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) out[u] = k0 + u < cnt ? L[(size_t)(k0 + u) * st] : 0u;
        }
    };
    uint32_t e[4];
    group(0, e);
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
        uint32_t en[4];
        group(k + 4, en);
        float4 xj[4], vj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = (int)(e[u] >> NB_SLOT_SHIFT);
            if (code) enc.put(k + u, (uint32_t)j, ctl);
            xj[u] = S.pos[j];
            vj[u] = S.vel[j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) pair(e[u], xj[u], vj[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = en[u];
    }
    if (k < cnt) {
        const int rem = cnt - k;
        float4 xj[4], vj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u < rem) {
                const int j = (int)(e[u] >> NB_SLOT_SHIFT);
                if (code) enc.put(k + u, (uint32_t)j, ctl);
                xj[u] = S.pos[j];
                vj[u] = S.vel[j];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < rem) pair(e[u], xj[u], vj[u]);
    }
    if (code) enc.finish();
    fy = fy - P.initial_mass * P.gravity;
    const Dom d = dom_at(A, i);
    if (P.add_boundary_force) box_force_static(P, d, xi, fx, fy, fz);
    const float4 F = make_float4(fx, fy, fz, 0.0f);
    W.force[i] = F;
    W.xs[i] = predict(xi, vi, F, make_float4(0.f, 0.f, 0.f, 0.f), A, d, ctl);
}

template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_ext_static(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                      Ctl* ctl) {
    __shared__ float lvl[LUT_N], lpg[LUT_N];
    __shared__ float2 th[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) {
        lvl[t] = luts[2 * LUT_N + t];
        lpg[t] = luts[LUT_N + t];
        th[t] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[t];
    }
    __syncthreads();
    for (int i = xblk(A.xcd) * blockDim.x + threadIdx.x; i < A.n; i += gstride())
        ext_static_slot<MODE>(A, S, W, i, lvl, lpg, th, ctl);
}

// The list build and the external-force pass in one kernel: a wave builds
// the lists of its 64 slots (k_nbr_build, packed entries without bins), then
// each lane runs the external-force work of its own slot over the list the
// wave just wrote.  Nothing crosses waves between the two parts, so the
// memory-bound force work of some waves overlaps the VALU-bound list work of
// others, and the lists are read back while they are still in the cache.
template <int ROUNDS, int SKIP, int MODE, bool SLOT_MAJOR>
__global__ __launch_bounds__(64) void k_nbr_build_ext(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                      Ctl* ctl) {
    // the build's candidate table, then the tables of the force pass
    __shared__ float4 tabs[LUT_N];
    static_assert(4 * LUT_N >= CAND_CAP, "candidate table");
    build_wave<ROUNDS, 0, 0, SKIP, 0, UPF_LIQUID, SLOT_MAJOR>(A, S.pos, W.keys_sorted, W.cells, W.nbr, W.ncount, ctl,
                                                             reinterpret_cast<uint32_t*>(tabs));
    // the block is the wave: its lists (written by all its lanes) are complete
    // and visible to every lane after the barrier, and so are the tables
    __syncthreads();
    float* lvl = reinterpret_cast<float*>(tabs);
    float* lpg = lvl + LUT_N;
    float2* th = reinterpret_cast<float2*>(lvl + 2 * LUT_N);
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) {
        lvl[t] = luts[2 * LUT_N + t];
        lpg[t] = luts[LUT_N + t];
        th[t] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[t];
    }
    __syncthreads();
    const int i = xblk(A.xcd) * 64 + (int)threadIdx.x;
    if (i < A.n) ext_static_slot<MODE, SLOT_MAJOR>(A, S, W, i, lvl, lpg, th, ctl);
}

// ------------------------------------------------------ PCISPH iteration
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Loads the union of chunk c (records rec_src[slot] of its slots) into LDS
// and returns its size, or -1 if the chunk reads the delta-coded lists.  All
// threads of the block call it (it synchronises); blockDim.x == UNION_CHUNK.
__device__ inline int stage_union(const StepArgs& A, const Work& W, const float4* __restrict__ rec_src,
                                  float4* urec, int c) {
    const int un = c * UNION_CHUNK < A.n ? W.un[c] : -1;
    __syncthreads();  // the previous chunk's readers are done
    const uint32_t* us = W.uslots + (size_t)c * UNION_CAP;
    for (int u = threadIdx.x; u < un; u += blockDim.x) urec[u] = at_u32(rec_src, (int)us[u]);
    __syncthreads();
    return un;
}

// ComputePredictedDensityAndPressure (:5456-5491) + the max reduction of the
// predicted density per scene: maxima fold into a few padded slots per scene
// (atomics spread over cache lines).  Persistent grid-stride kernel; iteration
// t of a scene runs only while t < its want, and the kernel returns at once
// when no scene runs iteration t.
template <int MODE, int CODED, int G>
__global__ __launch_bounds__(BLOCK) void k_density(int t, StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                   Ctl* ctl) {
    static_assert(!(MODE == UPF_BOUNDARY_AKINCI && CODED == 2), "the Akinci density reads boundary volumes by slot");
    if (t >= __atomic_load_n(&ctl->want_any, __ATOMIC_RELAXED)) return;
    __shared__ float lw[LUT_N];
    __shared__ float2 th[LUT_N];
    __shared__ float wmax[BLOCK / 64];  // blockDim.x <= BLOCK
    __shared__ float4 urec[CODED == 2 ? UNION_CAP : 1];  // the chunk's union: predicted positions
    // batched: empty the next iteration's chunk list, which the force pass
    // of this iteration fills (a different list from the one read here)
    if (A.act && blockIdx.x == 0 && threadIdx.x == 0) A.act_n[(t + 1) & 1] = 0;
    const upf_params& P = A.P;
    const int lane = threadIdx.x & 63;
    const int sps = A.slots_per_scene;
    // the predicted density of slot i (this iteration's pressure, the
    // force pass's record), or 0 with sc = -1 when the slot has none
    auto density_of = [&](int i, int un, float& mine, int& sc) __attribute__((always_inline)) {
        mine = 0.0f;
        sc = scene_at(A, i);
        // batched: a converged scene's slot stops before its record is
        // read (one scene runs iteration t: the kernel's entry test decided it)
        const bool runs = !A.scene_of || t < A.sctl[sc].want;
        const float4 xi = runs ? W.xs[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (!(runs && ptype(xi) == UPF_LIQUID)) {
            sc = -1;
            return;
        }
        const float m = P.initial_mass;
        const float T = P.support_radius_sq_limit;
        float rho = P.kernel_self * m;
        if (P.add_wall_weight)  // the GPU path's order: walls before neighbours
            rho = rho + wall_weight<false>(P, wall_distance(P, xi.x, xi.y, xi.z), A.lut_factor, lw);
        const int cnt = W.ncount[i];
        const size_t st = (size_t)A.stride;
        // G neighbours per round: their loads are issued together,
        // the sum still runs in list order
        auto use = [&](int j, uint32_t, const float4& xj) __attribute__((always_inline)) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if constexpr (MODE == UPF_BOUNDARY_AKINCI) {
                if (d2 < T) {
                    const float w = lw[bin_of_d2(d2, A.lut_factor, th)];
                    // a boundary neighbour's record carries -(rho0 psi)
                    const float mj = ptype(xj) == UPF_LIQUID ? m : -xj.w;
                    rho = rho + w * mj;
                }
            } else {
                // without a branch: a pair at d2 >= T adds w = +0, and
                // rho + (+0) * m is rho exactly (rho > 0)
                const float wt = lw[bin_of_d2(d2, A.lut_factor, th)];
                const float w = d2 < T ? wt : 0.0f;
                rho = rho + w * m;
            }
        };
        auto from_global = [&](int j) { return at_u32(W.xs, j); };
        if constexpr (CODED == 2) {
            if (un >= 0) for_each_listed<2, G>(W, i, st, cnt, [&](int u) { return urec[u]; }, use);
            else for_each_listed<1, G>(W, i, st, cnt, from_global, use);
        } else {
            for_each_listed<CODED, G>(W, i, st, cnt, from_global, use);
        }
        const float err = fmax_ref(rho - P.rest_density, 0.0f);
        const float delta = A.scene_of ? A.doms[sc].delta : P.density_error_factor;
        const float p = W.p[i] + fmax_ref(err * delta, 0.0f);
        W.p[i] = p;
        // the force pass gathers current position and p / rho0^2 of a
        // neighbour in one 16-byte load; the position part was written
        // at the start of the step
        const float pd = p / (P.rest_density * P.rest_density);
        reinterpret_cast<float*>(W.xpd)[4 * (size_t)i + 3] = pd;
        W.rho[i] = rho;
        mine = rho;
    };
    if constexpr (CODED != 2) {
        if (A.act) {
            // batched: one wave per chunk of this iteration's list
            int cnt = A.act_n[t & 1];
            const bool all = cnt < 0;
            if (all) cnt = -cnt;
            // consecutive blocks go to different XCDs: a short list (late in
            // the loop, few scenes left) spreads over the whole chip
            const int wpb = blockDim.x >> 6;
            const int w0 = (int)blockIdx.x * wpb;
            if (w0 >= cnt) return;  // nothing for this block: no tables
            for (int k = threadIdx.x; k < LUT_N; k += blockDim.x) {
                lw[k] = k < 300 ? luts[k] : 0.0f;  // bin 300 (dist >= h) weighs 0, as in lut_at
                th[k] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[k];
            }
            __syncthreads();
            const int* list = A.act + (size_t)(t & 1) * A.n_chunks;
            for (int q = w0 + (int)(threadIdx.x >> 6); q < cnt; q += (int)gridDim.x * wpb) {
                const int c = all ? q : list[q];
                const int i = A.chunk_begin[c] + lane;
                float mine = 0.0f;
                int sc = -1;
                if (i < A.chunk_end[c]) density_of(i, -1, mine, sc);
                // a chunk holds one scene: one maximum per wave
                const uint64_t act = wballot(sc >= 0);
                if (act == 0ull) continue;
                const int first = __shfl(sc, __ffsll((long long)act) - 1, 64);
                const float mx = wave_max(mine);
                if (lane == 0) atomicMax(&A.max_slots[(first * sps + q % sps) * 32], __float_as_uint(mx));
            }
            return;
        }
    }
    for (int k = threadIdx.x; k < LUT_N; k += blockDim.x) {
        lw[k] = k < 300 ? luts[k] : 0.0f;  // bin 300 (dist >= h) weighs 0, as in lut_at
        th[k] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[k];
    }
    __syncthreads();
    float run = 0.0f;  // single scene: running max over the thread's slots
    // top down (the list build and the force pass walk upwards), so that the
    // previous kernel's last slots come first, while they are in L2
    const int n_round = (A.n + gstride() - 1) / gstride();
    for (int r = n_round - 1; r >= 0; --r) {
        const int i = r * gstride() + xblk_down(A.xcd) * blockDim.x + threadIdx.x;
        if (block_idle(A, i - (int)threadIdx.x, t)) continue;
        int un = -1;
        if constexpr (CODED == 2) un = stage_union(A, W, W.xs, urec, (i - (int)threadIdx.x) / UNION_CHUNK);
        float mine = 0.0f;
        int sc = -1;
        if (i < A.n) density_of(i, un, mine, sc);
        if (A.scene_of == nullptr) {
            run = fmaxf(run, mine);
            continue;
        }
        // several scenes: reduce per wave when the wave holds one scene
        const uint64_t act = wballot(sc >= 0);
        if (act == 0ull) continue;
        const int first = __shfl(sc, __ffsll((long long)act) - 1, 64);
        const bool uniform = wballot(sc >= 0 && sc != first) == 0ull;
        if (uniform) {
            const float m = wave_max(mine);
            if (lane == 0)
                atomicMax(&A.max_slots[(first * sps + ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % sps)) * 32],
                          __float_as_uint(m));
        } else if (sc >= 0) {
            atomicMax(&A.max_slots[(sc * sps + lane % sps) * 32], __float_as_uint(mine));
        }
    }
    if (A.scene_of == nullptr) {
        run = wave_max(run);
        if (lane == 0) wmax[threadIdx.x >> 6] = run;
        __syncthreads();
        if (threadIdx.x == 0) {
            float b = wmax[0];
            for (int k = 1; k < (int)(blockDim.x >> 6); ++k) b = fmaxf(b, wmax[k]);
            if (b > 0.0f) atomicMax(&A.max_slots[(blockIdx.x % sps) * 32], __float_as_uint(b));
        }
    }
}

// The loop test of PredictionCorrectionStep (:5376-5398) on the device for
// scene sc after iteration t's densities (one wave): max over the scene's
// slots (which it clears for the next iteration), densityErrorInPercent, the
// too-large latch and the number of iterations the scene's loop runs; that
// number also raises ctl->want_any.  Every scene's count only grows during a
// step, so want_any needs no reset between iterations.
__device__ inline void loop_test(const StepArgs& A, int sc, int t, int lane, Ctl* ctl) {
    const upf_params& P = A.P;
    SceneCtl& C = A.sctl[sc];
    if (t >= C.want) return;
    uint32_t b = 0u;
    if (lane < A.slots_per_scene) {
        uint32_t* slot = &A.max_slots[(sc * A.slots_per_scene + lane) * 32];
        b = *slot;
        *slot = 0u;
    }
    for (int o = 32; o > 0; o >>= 1) b = max(b, (uint32_t)__shfl_xor((int)b, o, 64));
    int want = 0;
    if (lane == 0) {
        const float mx = __uint_as_float(b);
        const float pct = fmax_ref(0.1f * mx - 100.0f, 0.0f);
        int too_large = C.too_large;
        if (pct < P.max_density_error_allowed) too_large = 0;
        const int it = t + 1;
        const bool next = it < P.min_loops || (too_large && it < P.max_loops);
        want = next ? it + 1 : it;
        C.too_large = too_large;
        C.last_max = mx;
        C.want = want;
        atomicMax(&ctl->want_any, want);
    }
    // batched: a scene that runs iteration t + 1 puts its chunks on that
    // iteration's list (emptied by the density pass of iteration t)
    if (A.act) {
        want = __shfl(want, 0, 64);
        if (want > t + 1) {
            const int nch = A.sc_nchunks[sc], c0 = A.sc_chunk0[sc];
            int base = 0;
            if (lane == 0) base = atomicAdd(&A.act_n[(t + 1) & 1], nch);
            base = __shfl(base, 0, 64);
            int* list = A.act + (size_t)((t + 1) & 1) * A.n_chunks;
            for (int k = lane; k < nch; k += 64) list[base + k] = c0 + k;
        }
    }
}

// Start of a step's loop: every scene runs iteration 0 (the CPU loop starts
// with density_error_too_large_ = true) unless its loop limits are zero.
__global__ void k_loop_init(StepArgs A, Ctl* ctl) {
    const upf_params& P = A.P;
    const int w = (0 < P.min_loops || 0 < P.max_loops) ? 1 : 0;
    for (int sc = blockIdx.x * blockDim.x + threadIdx.x; sc < A.n_scenes; sc += gstride()) {
        A.sctl[sc].want = w;
        A.sctl[sc].too_large = 1;
        A.sctl[sc].last_max = 0.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->want_any = w;
        if (A.act) A.act_n[0] = -A.n_chunks;  // iteration 0: every chunk
    }
}

// Per-step record of every scene's iteration count and last max density
// (one block), then the step counter advances.
__global__ void k_record(StepArgs A, Ctl* ctl, int* __restrict__ k_out, float* __restrict__ r_out, int cap_steps) {
    const uint32_t step = ctl->step;
    if ((int)step < cap_steps) {
        for (int sc = threadIdx.x; sc < A.n_scenes; sc += blockDim.x) {
            k_out[(size_t)step * A.n_scenes + sc] = A.sctl[sc].want;
            r_out[(size_t)step * A.n_scenes + sc] = A.sctl[sc].last_max;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) ctl->step = step + 1u;
}

// Loop state for the host: the control words and scene 0's loop state are
// written to pinned host memory, then a sequence number the host waits for.
struct CtlMirror {
    Ctl ctl;
    SceneCtl scene0;
    uint32_t seq;
    uint32_t pad[3];
};

__global__ void k_publish(const Ctl* __restrict__ ctl, const SceneCtl* __restrict__ sctl, CtlMirror* m, uint32_t seq) {
    if (threadIdx.x != 0) return;
    m->ctl = *ctl;
    m->scene0 = *sctl;
    __threadfence_system();
    *(volatile uint32_t*)&m->seq = seq;
}

// ComputeCorrectivePressureForce (:5493-5522), then the prediction of the
// next iteration for the same particle; the loop test of the iteration runs
// first (the force pass does not depend on it).
struct PosVel {
    float4 x, v;
};

template <int MODE, int CODED, int G>
__global__ __launch_bounds__(BLOCK) void k_force(int t, StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                 Ctl* ctl) {
    if (t >= __atomic_load_n(&ctl->want_any, __ATOMIC_RELAXED)) return;
    // the loop test of iteration t, one wave per scene: it only ever raises a
    // scene's count past t, so the particle loop below reads either value
    {
        const int wpb = blockDim.x >> 6;
        const int waves = gridDim.x * wpb;
        for (int sc = blockIdx.x * wpb + (threadIdx.x >> 6); sc < A.n_scenes; sc += waves)
            loop_test(A, sc, t, threadIdx.x & 63, ctl);
    }
    __shared__ float lpg[LUT_N];
    __shared__ float2 th[LUT_N];
    __shared__ float4 urec[CODED == 2 ? UNION_CAP : 1];  // the chunk's union: (position, p / rho0^2)
    const upf_params& P = A.P;
    // the correction pressure force of slot i and its next prediction
    auto force_of = [&](int i, int un) __attribute__((always_inline)) {
        // batched: a converged scene's slot stops before its record is read (one
        // scene runs iteration t: the kernel's entry test decided it)
        if (A.scene_of && t >= A.sctl[A.scene_of[i]].want) return;
        // the slot's own record: current position and p / rho0^2 -- the density
        // pass's p / (rho0 rho0), the same division as W.p[i] / dens_sq -- with
        // the sign bit set for a non-liquid particle
        const float4 xr = W.xpd[i];
        if (__builtin_signbit(xr.w)) return;
        const float4 xi = make_float4(xr.x, xr.y, xr.z, __int_as_float(UPF_LIQUID));
        const float m = P.initial_mass;
        const float pi_d = xr.w;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        const int cnt = W.ncount[i];
        const size_t st = (size_t)A.stride;
        // every listed pair is closer than h at these (current) positions, so the
        // reference's `dist < h` test holds; the table bin is the one the list
        // build computed from the same positions (carried by the packed entry,
        // recomputed by the same expression for the coded lists)
        // (one 16-byte record per neighbour: current position and p / rho0^2,
        // or for an Akinci boundary neighbour its mass rho0 psi, negated)
        auto use = [&](int j, uint32_t e, const float4& xj) __attribute__((always_inline)) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const uint32_t bin = CODED ? bin_of_d2(dx * dx + dy * dy + dz * dz, A.lut_factor, th) : (e & NB_BIN_MASK);
            const float g = lpg[bin];
            // grad = p_i / densSq + p_j / (rho0 * rho0); the neighbour term is
            // precomputed per particle by the density pass (same division)
            float pd = xj.w, mj = m;
            if constexpr (MODE == UPF_BOUNDARY_AKINCI) {
                if (__builtin_signbit(xj.w)) {  // boundary: p_j = 0, mass rho0 psi_j
                    mj = -xj.w;
                    pd = 0.0f;
                }
            }
            const float grad = pi_d + pd;
            fx = fx - (((dx * g) * grad) * m) * mj;
            fy = fy - (((dy * g) * grad) * m) * mj;
            fz = fz - (((dz * g) * grad) * m) * mj;
        };
        auto from_global = [&](int j) { return at_u32(W.xpd, j); };
        if constexpr (CODED == 2) {
            if (un >= 0)
                for_each_listed<2, G>(W, i, st, cnt, [&](int u) { return urec[u]; }, use);
            else
                for_each_listed<1, G>(W, i, st, cnt, from_global, use);
        } else {
            for_each_listed<CODED, G>(W, i, st, cnt, from_global, use);
        }
        const float4 Fp = make_float4(fx, fy, fz, 0.0f);
        // only the last iteration's pressure force is read (by the integration):
        // it is written when this iteration may be the scene's last -- the scene's
        // loop count is t + 1 until this iteration's test (at the start of this
        // kernel) raises it, so a slot that reads it before the test writes too
        const int want = __atomic_load_n(&A.sctl[A.scene_of ? A.scene_of[i] : 0].want, __ATOMIC_RELAXED);
        if (t + 1 >= want) W.fp[i] = Fp;
        W.xs[i] = predict(xi, S.vel[i], W.force[i], Fp, A, dom_at(A, i), ctl);
    };
    auto load_tables = [&]() {
        for (int k = threadIdx.x; k < LUT_N; k += blockDim.x) {
            lpg[k] = luts[LUT_N + k];
            th[k] = reinterpret_cast<const float2*>(luts + LUT_TH_OFF)[k];
        }
        __syncthreads();
    };
    if constexpr (CODED != 2) {
        if (A.act) {
            // batched: one wave per chunk of this iteration's list
            int cnt = A.act_n[t & 1];
            const bool all = cnt < 0;
            if (all) cnt = -cnt;
            // consecutive blocks go to different XCDs: a short list (late in
            // the loop, few scenes left) spreads over the whole chip
            const int wpb = blockDim.x >> 6;
            const int w0 = (int)blockIdx.x * wpb;
            if (w0 >= cnt) return;  // nothing for this block: no tables
            load_tables();
            const int* list = A.act + (size_t)(t & 1) * A.n_chunks;
            for (int q = w0 + (int)(threadIdx.x >> 6); q < cnt; q += (int)gridDim.x * wpb) {
                const int c = all ? q : list[q];
                const int i = A.chunk_begin[c] + (int)(threadIdx.x & 63);
                if (i < A.chunk_end[c]) force_of(i, -1);
            }
            return;
        }
    }
    load_tables();
    const int n_round = (A.n + gstride() - 1) / gstride();
    for (int r = 0; r < n_round; ++r) {
        const int i = r * gstride() + xblk(A.xcd) * blockDim.x + threadIdx.x;
        if (block_idle(A, i - (int)threadIdx.x, t)) continue;
        int un = -1;
        if constexpr (CODED == 2) un = stage_union(A, W, W.xpd, urec, (i - (int)threadIdx.x) / UNION_CHUNK);
        if (i < A.n) force_of(i, un);
    }
}

// ------------------------------------------------------------ integration
// UpdateLiquidParticlePosVelSphCPU (:3705-3742) + BoundaryHandlingBoxPerParticleOMP
// (:3940-3970) / BoundaryHandlingPureFluidOMP (:3972-3998): symplectic Euler,
// box clamp with damping -0.5, new z-index key of every particle.
__global__ void k_integrate(StepArgs A, Slots S, Work W, int pcisph, int pure_fluid_clamp, int has_bodies, Ctl* ctl) {
    const upf_params& P = A.P;
    for (int i = xblk(A.xcd) * blockDim.x + threadIdx.x; i < A.n; i += gstride()) {
    float4 x = S.pos[i];
    const int ti = ptype(x);
    const Dom d = dom_at(A, i);
    if (ti == UPF_FROZEN) {
        bool ok;
        W.keys[i] = dkey(x.x, x.y, x.z, d.scales, ok);
        if (!ok) set_err(ctl, UPF_ERR_RANGE);
        continue;
    }
    float4 v = S.vel[i];
    if (ti == UPF_LIQUID) {
        float4 F = W.force[i];
        if (pcisph) {
            const float4 Fp = W.fp[i];
            F.x = F.x + Fp.x; F.y = F.y + Fp.y; F.z = F.z + Fp.z;
            W.force[i] = F;
        }
        v.x = v.x + (F.x * A.inv_mass_d) * P.dt;
        v.y = v.y + (F.y * A.inv_mass_d) * P.dt;
        v.z = v.z + (F.z * A.inv_mass_d) * P.dt;
        x.x = x.x + v.x * P.dt;
        x.y = x.y + v.y * P.dt;
        x.z = x.z + v.z * P.dt;
    }
    // static driver: every non-frozen particle without a parent body;
    // coupling driver: liquid particles only
    const bool clamp = pure_fluid_clamp ? (ti == UPF_LIQUID) : (!has_bodies || S.body[i] < 0);
    if (clamp) {
        const float damping = -0.5f;
        if (x.y < d.cmin[1]) { x.y = d.cmin[1]; v.y *= damping; }
        if (x.y > d.cmax[1]) { x.y = d.cmax[1]; v.y *= damping; }
        if (x.x < d.cmin[0]) { x.x = d.cmin[0]; v.x *= damping; }
        if (x.x > d.cmax[0]) { x.x = d.cmax[0]; v.x *= damping; }
        if (x.z < d.cmin[2]) { x.z = d.cmin[2]; v.z *= damping; }
        if (x.z > d.cmax[2]) { x.z = d.cmax[2]; v.z *= damping; }
    }
    if (ti == UPF_LIQUID || clamp) {
        S.pos[i] = x;
        S.vel[i] = v;
    }
    bool ok;
    W.keys[i] = dkey(x.x, x.y, x.z, d.scales, ok);
    if (!ok) set_err(ctl, UPF_ERR_RANGE);
    }
}

// ------------------------------------------------------------------- WCSPH
// CalculateDensityPressure (:2339-2406), the wall weight when
// P.add_wall_weight, then the pressure of physicsType 'w' (TAIT,
// CalculatePressureWCSPH :2491-2501) or 'o' (CalculatePressureSPH
// :2479-2489), over the lists
// The density work of slot i (one thread) over its packed list (entries
// carry the table bin): the density pass and the fused list build share it.
template <bool TAIT>
__device__ __forceinline__ void density_wcsph_slot(const StepArgs& A, const Slots& S, const Work& W, int i,
                                                   const float* lw) {
    const upf_params& P = A.P;
    const float4 xi = S.pos[i];
    if (ptype(xi) != UPF_LIQUID) {
        W.rho[i] = 0.0f;
        W.pres[i] = 0.0f;
        W.force[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        W.xpd[i] = make_float4(xi.x, xi.y, xi.z, -1.0f);  // w < 0: not a fluid neighbour
        return;
    }
    const float m = P.initial_mass;
    float rho = P.kernel_self * m;
    const int cnt = W.ncount[i];
    for (int k = 0; k < cnt; ++k) rho = rho + lw[W.nbr[(size_t)k * A.stride + i] & NB_BIN_MASK] * m;
    if (P.add_wall_weight) rho = rho + wall_weight<true>(P, wall_distance(P, xi.x, xi.y, xi.z), A.lut_factor, lw);
    float pres;
    if constexpr (TAIT) {
        const float gamma = 7.0f;
        const float b = P.rest_density * P.gas_const_wcsph / gamma;
        const float frac = rho / P.rest_density;
        const float pg = powf(frac, gamma);
        pres = fmax_ref(0.0f, b * (pg - 1.0f));
    } else {  // gamma = 1: p = k (rho - rho0) written as the reference writes it
        const float b = P.rest_density * P.gas_const_sph / 1.0f;
        const float frac = rho / P.rest_density;
        pres = fmax_ref(0.0f, b * (frac - 1.0f));
    }
    W.rho[i] = rho;
    W.pres[i] = pres;
    // what the force pass needs of this particle as a neighbour, in two
    // 16-byte records: (position, m / rho) and (velocity, p / rho^2) -- the
    // same expressions it would evaluate per pair
    const float4 vi = S.vel[i];
    W.xpd[i] = make_float4(xi.x, xi.y, xi.z, m / rho);
    W.vpg[i] = make_float4(vi.x, vi.y, vi.z, pres / (rho * rho));
}

template <bool TAIT>
__global__ __launch_bounds__(BLOCK) void k_density_wcsph(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                         Ctl* ctl) {
    __shared__ float lw[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) lw[t] = luts[t];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) density_wcsph_slot<TAIT>(A, S, W, i, lw);
}

// The list build (entries with bins) and the density pass in one kernel:
// each lane sums the density of its own slot over the list its wave just
// wrote (as k_nbr_build_ext does for the PCISPH external forces).
template <int ROUNDS, bool TAIT>
__global__ __launch_bounds__(64) void k_nbr_build_wcsph(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                        Ctl* ctl) {
    // the build's candidate table, then the density table
    __shared__ float lw[LUT_N > CAND_CAP ? LUT_N : CAND_CAP];
    build_wave<ROUNDS, 0, 1, 0>(A, S.pos, W.keys_sorted, W.cells, W.nbr, W.ncount, ctl,
                                reinterpret_cast<uint32_t*>(lw));
    __syncthreads();  // the block is the wave: its lists are complete
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) lw[t] = luts[t];
    __syncthreads();
    const int i = xblk(A.xcd) * 64 + (int)threadIdx.x;
    if (i < A.n) density_wcsph_slot<TAIT>(A, S, W, i, lw);
}

// CalculateLiquidParticleForces (:3009-3083) + gravity (:2822-2824)
__global__ __launch_bounds__(BLOCK) void k_force_wcsph(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                       Ctl* ctl) {
    __shared__ float lvl[LUT_N], lpg[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) {
        lvl[t] = luts[2 * LUT_N + t];
        lpg[t] = luts[LUT_N + t];
    }
    __syncthreads();
    const upf_params& P = A.P;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float4 xi = S.pos[i];
    if (ptype(xi) != UPF_LIQUID) return;
    const float m = P.initial_mass;
    const float rho_i = W.rho[i], p_i = W.pres[i];
    const float invD = (float)(1.0 / (double)rho_i);
    const float pVol = m * invD;
    const float4 vi = S.vel[i];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int cnt = W.ncount[i];
    // two 16-byte records per neighbour, four neighbours' loads in flight
    for_each_listed<0, 4>(
        W, i, (size_t)A.stride, cnt, [&](int j) { return PosVel{at_u32(W.xpd, j), at_u32(W.vpg, j)}; },
        [&](int, uint32_t e, const PosVel& q) __attribute__((always_inline)) {
            const float4& xj = q.x;
            if (xj.w < 0.0f) return;  // not a fluid neighbour
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float kv = lvl[e & NB_BIN_MASK];
            const float nVol = xj.w;  // m / rho_j
            const float g = lpg[e & NB_BIN_MASK];
            const float grad = p_i * invD * invD + q.v.w;  // + p_j / rho_j^2
            fx = fx - (((dx * g) * grad) * m) * m;
            fy = fy - (((dy * g) * grad) * m) * m;
            fz = fz - (((dz * g) * grad) * m) * m;
            const float vx = vi.x - q.v.x, vy = vi.y - q.v.y, vz = vi.z - q.v.z;
            fx = fx - vx * pVol * nVol * P.visc_const * kv;
            fy = fy - vy * pVol * nVol * P.visc_const * kv;
            fz = fz - vz * pVol * nVol * P.visc_const * kv;
        });
    if (P.add_boundary_force) box_force_static(P, dom_at(A, i), xi, fx, fy, fz);
    fy = fy - P.initial_mass * P.gravity;
    W.force[i] = make_float4(fx, fy, fz, 0.0f);
}

// ------------------------------------------------------ rigid coupling
// Neighbour lists of rigid particles.  Their keys are those of the position
// before the last rigid update (the reference recomputes keys before it moves
// the rigid particles, src/cpu/UnifiedPhysics.cpp:5007-5022), so their query box
// need not sit on their key's cell stencil: each lane walks its own box.
__global__ __launch_bounds__(BLOCK) void k_nbr_build_box(StepArgs A, const float4* __restrict__ pos,
                                                         const uint2* __restrict__ cells, int owner_type,
                                                         uint32_t* __restrict__ nbr, int* __restrict__ ncount,
                                                         Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float4 xi = pos[i];
    if (ptype(xi) != owner_type) return;
    const upf_params& P = A.P;
    const float lut_factor = A.lut_factor;
    const int stride = A.stride;
    const float T = P.support_radius_sq_limit;
    const int cap = P.max_neighbors;
    int cnt = 0;
    for_each_candidate(xi, P.support_radius, dom_at(A, i), cells, [&](int j) __attribute__((always_inline)) {
        if (j == i) return;
        const float4 xj = pos[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 < T)) return;
        if (cnt < cap) nbr[(size_t)cnt * stride + i] = ((uint32_t)j << NB_SLOT_SHIFT) | lut_bin(sqrtf(d2), lut_factor);
        ++cnt;
    });
    ncount[i] = cnt < cap ? cnt : cap;
    if (cnt > cap) set_err(ctl, UPF_ERR_NBR_OVERFLOW);
}

// Boundary volumes of the rigid particles only (walls keep theirs).
// (entry k of slot i at nbr[k * kstride + i * istride]: column-major or slot-major lists)
__global__ void k_psi_rigid(int n, const float4* __restrict__ pos, const uint32_t* __restrict__ nbr, int kstride,
                            int istride, const int* __restrict__ ncount, upf_params P, const float* __restrict__ luts,
                            float* __restrict__ psi, Ctl* ctl) {
    __shared__ float lw[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) lw[t] = luts[t];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (ptype(pos[i]) != UPF_RIGID) return;
    float inv = P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI ? lw[0] : 0.0f;
    const int cnt = ncount[i];
    auto add = [&](uint32_t e) __attribute__((always_inline)) {
        const int tj = ptype(pos[e >> NB_SLOT_SHIFT]);
        if (tj == UPF_FROZEN || tj == UPF_RIGID) inv += lw[e & NB_BIN_MASK];
    };
    if (kstride == 1) {
        // a slot-major row (its pitch a multiple of four): 16-byte groups
        const uint32_t* L = nbr + (size_t)i * istride;
        for (int k = 0; k < cnt; k += 4) {
            const uint4 q = *reinterpret_cast<const uint4*>(L + k);
            const uint32_t e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (k + u < cnt) add(e[u]);
        }
    } else {
        for (int k = 0; k < cnt; ++k) add(nbr[(size_t)k * kstride + (size_t)i * istride]);
    }
    if (inv != 0.0f) psi[i] = P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI ? (float)(0.7 / (double)inv) : (float)(1.0 / (double)inv);
    else set_err(ctl, UPF_ERR_ISOLATED);
}

// CalculateBoundaryFluidPressureForceHost (:4240-4250) for a liquid particle
// at xl (density rho, pressure pr, volume wv) and a boundary particle at xb
__device__ inline void boundary_pressure(int mode, float dist, float h, const float4& xl, const float4& xb, float rho,
                                         float pr, float wv, const float* lpg, const upf_params& P, float& ox,
                                         float& oy, float& oz) {
    uint32_t dl = (uint32_t)(dist * 300.0f / h);
    dl = dl > 300u ? 300u : dl;
    const float g = mode == UPF_BOUNDARY_CPU_REF ? 0.0f : lpg[dl];  // the CPU path's look-up always yields +0
    const float inv = 1.0f / rho;
    const float pg = pr * inv * inv;
    const float dx = xl.x - xb.x, dy = xl.y - xb.y, dz = xl.z - xb.z;
    ox = (((((dx * g) * pg) * P.initial_mass) * P.rest_density) * wv) * -1.0f;
    oy = (((((dy * g) * pg) * P.initial_mass) * P.rest_density) * wv) * -1.0f;
    oz = (((((dz * g) * pg) * P.initial_mass) * P.rest_density) * wv) * -1.0f;
}

// CalculateExternalForcesVersatilCouplingPCISPH (:5252-5365) over the lists
// for liquid and rigid particles, + the first prediction of the loop; every
// slot resets its loop pressure.  Pressures and densities of neighbours are
// the ones the previous step's loop ended with.
// The coupling force work of slot i (one thread): shared by the coupling
// pass and the fused list build.  lvl, lpg, lst, lad: the viscosity,
// pressure-gradient, tension and adhesion tables in LDS.
template <int MODE, bool SLOT_MAJOR = false>
__device__ __forceinline__ void ext_coupling_slot(const StepArgs& A, const Slots& S, const Work& W, int i,
                                                  const float* lvl, const float* lpg, const float* lst,
                                                  const float* lad, Ctl* ctl) {
    const upf_params& P = A.P;
    const Dom d = dom_at(A, i);
    const float4 xi = S.pos[i];
    const int ti = ptype(xi);
    W.p[i] = 0.0f;
    W.xpd[i] = make_float4(xi.x, xi.y, xi.z, pressure_record_w<MODE>(ti, S.psi[i], P.rest_density));
    W.fp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    W.rho[i] = S.rho_last[i];
    if (ti != UPF_LIQUID && ti != UPF_RIGID) {
        W.xs[i] = boundary_xs<MODE>(xi, S.psi[i], P.rest_density);
        W.force[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4 vi = S.vel[i];
    const float h = P.support_radius;
    const float m = P.initial_mass;
    const float rv = A.rest_vol, visc = P.visc_const;
    const float quarter = h * 0.25f;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int cnt = W.ncount[i];
    // the liquid particle's own pressure / density / volume (CPU path:
    // previous_correction_pressure_ and its own weighted_volume_)
    const float own_p = MODE == UPF_BOUNDARY_CPU_REF ? S.p_prev[i] : S.p_last[i];
    const float own_rho = S.rho_last[i];
    const float own_psi = S.psi[i];
    ListEncoder enc(W, i, (size_t)A.stride);
    const bool code = A.code_lists != 0;
    // one neighbour (entry e, record xj)
    auto pair = [&](uint32_t e, const float4& xj) __attribute__((always_inline)) {
        const int j = (int)(e >> NB_SLOT_SHIFT);
        const int tj = ptype(xj);
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        uint32_t dl = (uint32_t)(dist * 300.0f / h);
        dl = dl > 300u ? 300u : dl;
        if (ti == UPF_LIQUID) {
            if (tj == UPF_LIQUID) {
                const float kv = lvl[e & NB_BIN_MASK];
                const float4 vj = S.vel[j];
                const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
                fx = fx - vx * rv * rv * visc * kv;
                fy = fy - vy * rv * rv * visc * kv;
                fz = fz - vz * rv * rv * visc * kv;
                if (dist >= quarter) {  // CalculateSurfaceTensionForcePCISPHHost (:1990-1998)
                    const float sc = lst[dl];
                    const float id = 1.0f / dist;
                    fx = fx - dx * P.surface_tension_gamma * m * m * sc * id;
                    fy = fy - dy * P.surface_tension_gamma * m * m * sc * id;
                    fz = fz - dz * P.surface_tension_gamma * m * m * sc * id;
                }
            } else if (tj == UPF_FROZEN || tj == UPF_RIGID) {
                float ox, oy, oz;
                const float wv = MODE == UPF_BOUNDARY_CPU_REF ? own_psi : S.psi[j];
                boundary_pressure(MODE, dist, h, xi, xj, own_rho, own_p, wv, lpg, P, ox, oy, oz);
                fx = fx + ox; fy = fy + oy; fz = fz + oz;
                if (dist >= quarter) {  // CalculateSurfaceCohesionForcePCISPHHost (:2000-2008)
                    const float sc = lad[dl];
                    const float id = 1.0f / dist;
                    const float pj = S.psi[j];
                    fx = fx - dx * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                    fy = fy - dy * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                    fz = fz - dz * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                }
            }
        } else {  // rigid
            if (tj == UPF_RIGID) {
                const float overlap = 2.0f * P.particle_radius - dist;
                if (overlap > 0.00001f) {
                    // CalculateSpringForceHost + CalculateDampingForceHost (:2010-2030), r_ij = x_j - x_i
                    const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                    const float sx = -1.0f * P.spring_coefficient * overlap * rx / dist;
                    const float sy = -1.0f * P.spring_coefficient * overlap * ry / dist;
                    const float sz = -1.0f * P.spring_coefficient * overlap * rz / dist;
                    const float4 vj = S.vel[j];
                    const float cx = P.damping_coefficient * (vj.x - vi.x);
                    const float cy = P.damping_coefficient * (vj.y - vi.y);
                    const float cz = P.damping_coefficient * (vj.z - vi.z);
                    fx = fx + (sx + cx);
                    fy = fy + (sy + cy);
                    fz = fz + (sz + cz);
                }
            } else if (tj == UPF_LIQUID) {
                // the liquid's pressure force on this boundary particle, reversed (:5316-5321)
                float ox, oy, oz;
                const float wv = MODE == UPF_BOUNDARY_CPU_REF ? S.psi[j] : S.psi[i];
                boundary_pressure(MODE, dist, h, xj, xi, S.rho_last[j], S.p_last[j], wv, lpg, P, ox, oy, oz);
                fx = fx - ox; fy = fy - oy; fz = fz - oz;
                if (dist >= quarter) {
                    const float sc = lad[dl];
                    const float id = 1.0f / dist;
                    const float pj = S.psi[j];
                    const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                    fx = fx + rx * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                    fy = fy + ry * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                    fz = fz + rz * P.surface_adhesion_beta * m * P.rest_density * pj * sc * id;
                }
            }
        }
    };
    // list entries in groups of four, the next group's loaded ahead; the next
    // neighbour's record is loaded before this neighbour's pair runs.  The
    // packed list: column-major, or (SLOT_MAJOR) the slot's own row, whose
    // groups are aligned 16-byte words (entries past the end are not used)
    const size_t st = (size_t)A.stride;
    const uint32_t* L = SLOT_MAJOR ? W.nbr + (size_t)i * A.nbr_ld : W.nbr + i;
    auto group = [&](int k0, uint32_t* out) __attribute__((always_inline)) {
        if constexpr (SLOT_MAJOR) {
            uint4 q = make_uint4(0u, 0u, 0u, 0u);
            if (k0 < cnt) q = *reinterpret_cast<const uint4*>(L + k0);
            out[0] = q.x;
            out[1] = q.y;
            out[2] = q.z;
            out[3] = q.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) out[u] = k0 + u < cnt ? L[(size_t)(k0 + u) * st] : 0u;
        }
    };
    uint32_t e[4];
    group(0, e);
    float4 xn = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cnt > 0) xn = S.pos[e[0] >> NB_SLOT_SHIFT];
    for (int k0 = 0; k0 < cnt; k0 += 4) {
        uint32_t en[4];
        group(k0 + 4, en);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u;
            if (k < cnt) {
                const float4 xj = xn;
                if (k + 1 < cnt) xn = S.pos[(u < 3 ? e[u + 1] : en[0]) >> NB_SLOT_SHIFT];
                if (code) enc.put(k, e[u] >> NB_SLOT_SHIFT, ctl);
                pair(e[u], xj);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = en[u];
    }
    if (code) enc.finish();
    if (ti == UPF_LIQUID) {
        fy = fy - P.initial_mass * P.gravity;
        // CalculateBoundaryForcePerLiquidParticleHost (:2032-2069)
        const float inv = 1.0f / P.force_distance;
        const float fd = P.force_distance, fmax = P.max_boundary_force;
        const float c[3] = {xi.x, xi.y, xi.z};
        float b[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float mn = d.cmin[a], mx = d.cmax[a];
            if (c[a] < mn + fd) {
                const float v = (float)((double)((mn + fd - c[a]) * inv) * 2.0 * (double)fmax);
#pragma unroll
                for (int q = 0; q < 3; ++q) b[q] = b[q] + (q == a ? 1.0f : 0.0f) * v;
            }
            if (c[a] > mx - fd) {
                const float v = (float)((double)((c[a] + fd - mx) * inv) * 2.0 * (double)fmax);
#pragma unroll
                for (int q = 0; q < 3; ++q) b[q] = b[q] + (q == a ? -1.0f : 0.0f) * v;
            }
        }
        fx = fx + b[0]; fy = fy + b[1]; fz = fz + b[2];
        const float4 F = make_float4(fx, fy, fz, 0.0f);
        W.force[i] = F;
        W.xs[i] = predict(xi, vi, F, make_float4(0.f, 0.f, 0.f, 0.f), A, d, ctl);
    } else {
        // CalculateBoundaryForcePerRigidParticleHost (:2071-2132)
        const float r = P.particle_radius, kb = P.spring_coefficient_boundary, cd = P.damping_coefficient;
        const float c[3] = {xi.x, xi.y, xi.z};
        float b[3] = {0.0f, 0.0f, 0.0f};
        bool hit = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float mn = d.cmin[a], mx = d.cmax[a];
            if (c[a] - r < mn) {
                hit = true;
                const float v = mn + r - c[a];
#pragma unroll
                for (int q = 0; q < 3; ++q) b[q] = b[q] + ((q == a ? 1.0f : 0.0f) * v) * kb;
            }
            if (c[a] + r > mx) {
                hit = true;
                const float v = c[a] + r - mx;
#pragma unroll
                for (int q = 0; q < 3; ++q) b[q] = b[q] + ((q == a ? -1.0f : 0.0f) * v) * kb;
            }
        }
        if (hit) { b[0] -= cd * vi.x; b[1] -= cd * vi.y; b[2] -= cd * vi.z; }
        fx = fx + b[0]; fy = fy + b[1]; fz = fz + b[2];
        W.force[i] = make_float4(fx, fy, fz, 0.0f);
        W.xs[i] = boundary_xs<MODE>(P.rigid_predicted_current ? xi : S.x0[i], S.psi[i], P.rest_density);
    }
}

template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_ext_coupling(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                        Ctl* ctl) {
    __shared__ float lvl[LUT_N], lpg[LUT_N], lst[LUT_N], lad[LUT_N];
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) {
        lpg[t] = luts[LUT_N + t];
        lvl[t] = luts[2 * LUT_N + t];
        lst[t] = luts[3 * LUT_N + t];
        lad[t] = luts[4 * LUT_N + t];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) ext_coupling_slot<MODE>(A, S, W, i, lvl, lpg, lst, lad, ctl);
}

// The liquid list build (entries with bins) and the coupling pass in one
// kernel: the rigid particles' lists and boundary volumes come first (their
// counts are kept), then a wave builds the lists of its liquid slots and each
// lane runs the coupling work of its own slot.
template <int ROUNDS, int MODE, bool SLOT_MAJOR>
__global__ __launch_bounds__(64) void k_nbr_build_coupling(StepArgs A, Slots S, Work W, const float* __restrict__ luts,
                                                           Ctl* ctl) {
    // the build's candidate table, then the tables of the coupling pass
    __shared__ float tabs[4 * LUT_N];
    static_assert(4 * LUT_N >= CAND_CAP, "candidate table");
    build_wave<ROUNDS, 0, 1, 0, 1, UPF_LIQUID, SLOT_MAJOR>(A, S.pos, W.keys_sorted, W.cells, W.nbr, W.ncount, ctl,
                                                          reinterpret_cast<uint32_t*>(tabs));
    __syncthreads();  // the block is the wave: its lists are complete
    float* lvl = tabs;
    float* lpg = tabs + LUT_N;
    float* lst = tabs + 2 * LUT_N;
    float* lad = tabs + 3 * LUT_N;
    for (int t = threadIdx.x; t < LUT_N; t += blockDim.x) {
        lpg[t] = luts[LUT_N + t];
        lvl[t] = luts[2 * LUT_N + t];
        lst[t] = luts[3 * LUT_N + t];
        lad[t] = luts[4 * LUT_N + t];
    }
    __syncthreads();
    const int i = xblk(A.xcd) * 64 + (int)threadIdx.x;
    if (i < A.n) ext_coupling_slot<MODE, SLOT_MAJOR>(A, S, W, i, lvl, lpg, lst, lad, ctl);
}

// key of a slot for grouping rigid particles by body (non-rigid last)
__global__ void k_body_keys(int n, const int* __restrict__ body, const float4* __restrict__ pos,
                            uint32_t* __restrict__ keys, uint32_t none) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = body[i];
    keys[i] = (b >= 0 && ptype(pos[i]) == UPF_RIGID) ? (uint32_t)b : none;
}

__global__ void k_body_bounds(int n, const uint32_t* __restrict__ keys, int nb, int* __restrict__ start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (k >= (uint32_t)nb) {
        if (i == 0 || keys[i - 1] < (uint32_t)nb) start[nb] = i;  // end of the rigid block
        return;
    }
    if (i == 0 || keys[i - 1] != k) start[k] = i;
    if (i == n - 1) start[nb] = n;
}

__device__ inline void mat3mul(const float* A, const float* B, float* C) {  // Matrix3x3::operator*
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = A[r * 3 + 0] * B[0 * 3 + c] + A[r * 3 + 1] * B[1 * 3 + c] + A[r * 3 + 2] * B[2 * 3 + c];
}

// The force, torque and weight terms of every body member of
// CalculateRigidBodyForceTorqueMomenta (:3620-3703), one thread per member
// position k of the members grouped by body (ascending slots within a body):
// terms[c * ld + k] for f.x f.y f.z t.x t.y t.z gy.
__global__ void k_body_terms(int n, int nb, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ members,
                             const upf_body* __restrict__ bodies, const float4* __restrict__ pos,
                             const float4* __restrict__ force, const float4* __restrict__ rel, upf_params P,
                             float* __restrict__ terms, int ld) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t b = keys[k];
    if (b >= (uint32_t)nb) return;
    const int i = (int)members[k];
    const float4 f = force[i];
    const float4 x = pos[i];
    const float r0 = x.x - bodies[b].pos[0], r1 = x.y - bodies[b].pos[1], r2 = x.z - bodies[b].pos[2];
    terms[0 * (size_t)ld + k] = f.x;
    terms[1 * (size_t)ld + k] = f.y;
    terms[2 * (size_t)ld + k] = f.z;
    terms[3 * (size_t)ld + k] = r1 * f.z - r2 * f.y;
    terms[4 * (size_t)ld + k] = r2 * f.x - r0 * f.z;
    terms[5 * (size_t)ld + k] = r0 * f.y - r1 * f.x;
    terms[6 * (size_t)ld + k] = -rel[i].w * P.gravity;
}

// The sums of CalculateRigidBodyForceTorqueMomenta with RigidBody::UpdateMomenta,
// PerformStep, UpdateRotationMatrix (src/cpu/RigidBody.cpp:55-160): one wave
// per body.  The wave loads the terms of 64 members at a time (coalesced,
// RB_DEPTH batches in flight) into LDS; six lanes then run the six sums --
// F.x, F.y, F.z, T.x, T.y, T.z -- each in member order with the reference's
// additions (F.x + f.x + 0, F.y + f.y + gy, F.z + f.z + 0, T + t), so every
// sum has the serial loop's rounding.  The kernel takes as long as the
// largest body's F.y chain (two dependent additions per member).
constexpr int RB_CHUNK = 64;
constexpr int RB_DEPTH = 4;
__global__ __launch_bounds__(RB_CHUNK) void k_rigid_bodies(int nb, upf_body* __restrict__ bodies,
                                                           const int* __restrict__ start,
                                                           const float* __restrict__ terms, int ld, upf_params P) {
    const int b = blockIdx.x;
    if (b >= nb) return;
    __shared__ float tt[8][RB_CHUNK];  // the seven term rows and a row of +0
    __shared__ float sums[6];
    __shared__ int range[2];
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int s0 = start[b];
        int s1 = nb;
        for (int k = b + 1; k <= nb; ++k) {
            if (start[k] >= 0) { s1 = start[k]; break; }
        }
        range[0] = s0;
        range[1] = s0 >= 0 ? s1 : s0;
    }
    __syncthreads();
    const int s0 = range[0], s1 = range[1];
    // member k0 + tid's seven terms (a clamped index past the body's end;
    // those terms are replaced by +0 before they are added)
    auto load = [&](int k0, float* out) __attribute__((always_inline)) {
        const int k = min(k0 + tid, s1 - 1);
#pragma unroll
        for (int c = 0; c < 7; ++c) out[c] = terms[(size_t)c * ld + k];
    };
    float r[RB_DEPTH][7];
    tt[7][tid] = 0.0f;
#pragma unroll
    for (int d = 0; d < RB_DEPTH; ++d)
        if (s0 + d * RB_CHUNK < s1) load(s0 + d * RB_CHUNK, r[d]);
    float acc = 0.0f;  // lane a < 6: sum a
    const int gy_row = tid == 1 ? 6 : 7;  // row 7: zeros
    for (int kb = s0; kb < s1; kb += RB_DEPTH * RB_CHUNK) {
#pragma unroll
        for (int d = 0; d < RB_DEPTH; ++d) {
            const int k0 = kb + d * RB_CHUNK;
            if (k0 >= s1) break;
            // past the body's last member a term is +0, which leaves a sum
            // unchanged (a sum from +0 is never -0)
            const bool in = k0 + tid < s1;
#pragma unroll
            for (int c = 0; c < 7; ++c) tt[c][tid] = in ? r[d][c] : 0.0f;
            if (k0 + RB_DEPTH * RB_CHUNK < s1) load(k0 + RB_DEPTH * RB_CHUNK, r[d]);
            __syncthreads();
            if (tid < 6) {
                // every lane adds a second term per member: gy for F.y, +0
                // for the others (a no-op on a sum that is never -0), which
                // keeps the six dependency chains free of selects
                const int m = min(RB_CHUNK, s1 - k0);
                for (int u0 = 0; u0 < m; u0 += 16) {
                    float v[16], g[16];
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        v[q] = tt[tid][u0 + q];
                        g[q] = tt[gy_row][u0 + q];
                    }
#pragma unroll
                    for (int q = 0; q < 16; ++q) acc = (acc + v[q]) + g[q];
                }
            }
            __syncthreads();
        }
    }
    if (tid < 6) sums[tid] = acc;
    __syncthreads();
    if (tid != 0) return;
    upf_body B = bodies[b];
    const float F[3] = {sums[0], sums[1], sums[2]}, Tq[3] = {sums[3], sums[4], sums[5]};
    const float dt = P.dt;
    const float terminal = B.mass * P.terminal_speed;
#pragma unroll
    for (int a = 0; a < 3; ++a) B.linear_momentum[a] = B.linear_momentum[a] + F[a] * dt;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (B.linear_momentum[a] > 0.0f) B.linear_momentum[a] = fmin_ref(B.linear_momentum[a], terminal);
        else B.linear_momentum[a] = fmax_ref(B.linear_momentum[a], -terminal);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) B.angular_momentum[a] = B.angular_momentum[a] + Tq[a] * dt;
    float Rt[9], tmp[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Rt[r * 3 + c] = B.rotation[c * 3 + r];
    mat3mul(B.rotation, B.inv_inertia_local, tmp);
    mat3mul(tmp, Rt, B.inv_inertia_world);
    {
        const float inv = 1.0f / B.mass;
#pragma unroll
        for (int a = 0; a < 3; ++a) B.vel[a] = B.linear_momentum[a] * inv;
#pragma unroll
        for (int a = 0; a < 3; ++a) B.pos[a] = B.pos[a] + B.vel[a] * dt;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        B.angular_velocity[r] = B.inv_inertia_world[r * 3 + 0] * B.angular_momentum[0] +
                                B.inv_inertia_world[r * 3 + 1] * B.angular_momentum[1] +
                                B.inv_inertia_world[r * 3 + 2] * B.angular_momentum[2];
    const float* w = B.angular_velocity;
    const float sp = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (sp > 0.0f) {
        const float ll = (float)(1.0 / (double)sp);
        const float ax = w[0] * ll, ay = w[1] * ll, az = w[2] * ll;
        const float ang = sp * dt;
        const float ds = cosf(ang / 2.0f);
        const float sn = sinf(ang / 2.0f);
        const float dvx = ax * sn, dvy = ay * sn, dvz = az * sn;
        const float s = B.quaternion[0], vx = B.quaternion[1], vy = B.quaternion[2], vz = B.quaternion[3];
        B.quaternion[0] = s * ds - vx * dvx - vy * dvy - vz * dvz;
        B.quaternion[1] = ds * vx + s * dvx + dvy * vz - dvz * vy;
        B.quaternion[2] = ds * vy + s * dvy + dvz * vx - dvx * vz;
        B.quaternion[3] = ds * vz + s * dvz + dvx * vy - dvy * vx;
    }
    if (P.renormalize_quaternion) {
        const float* q = B.quaternion;
        const float l = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        if (l > 0.0f) {
            const float il = 1.0f / l;
#pragma unroll
            for (int a = 0; a < 4; ++a) B.quaternion[a] *= il;
        }
    }
    const float qw = B.quaternion[0], qx = B.quaternion[1], qy = B.quaternion[2], qz = B.quaternion[3];
    const float xx = qx * qx, yy = qy * qy, zz = qz * qz, xy = qx * qy, xz = qx * qz, yz = qy * qz, wx = qw * qx,
                wy = qw * qy, wz = qw * qz;
    B.rotation[0] = 1.0f - 2.0f * (yy + zz); B.rotation[1] = 2.0f * (xy - wz); B.rotation[2] = 2.0f * (xz + wy);
    B.rotation[3] = 2.0f * (xy + wz); B.rotation[4] = 1.0f - 2.0f * (xx + zz); B.rotation[5] = 2.0f * (yz - wx);
    B.rotation[6] = 2.0f * (xz - wy); B.rotation[7] = 2.0f * (yz + wx); B.rotation[8] = 1.0f - 2.0f * (xx + yy);
#pragma unroll
    for (int a = 0; a < 3; ++a) { B.force[a] = F[a]; B.torque[a] = Tq[a]; }
    bodies[b] = B;
}

// SynRigidParticlesPerRigidBody (:3744-3792): x = x_cm + R r0, v = v_cm + w x (R r0)
__global__ void k_rigid_sync(int n, Slots S, const upf_body* __restrict__ bodies) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = S.pos[i];
    if (ptype(x) != UPF_RIGID) return;
    const int b = S.body[i];
    if (b < 0) return;
    const upf_body& B = bodies[b];
    const float4 r0 = S.rel[i];
    float r[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) r[q] = B.rotation[q * 3 + 0] * r0.x + B.rotation[q * 3 + 1] * r0.y + B.rotation[q * 3 + 2] * r0.z;
    const float* w = B.angular_velocity;
    const float c0 = w[1] * r[2] - w[2] * r[1], c1 = w[2] * r[0] - w[0] * r[2], c2 = w[0] * r[1] - w[1] * r[0];
    S.vel[i] = make_float4(B.vel[0] + c0, B.vel[1] + c1, B.vel[2] + c2, 0.0f);
    x.x = B.pos[0] + r[0];
    x.y = B.pos[1] + r[1];
    x.z = B.pos[2] + r[2];
    S.pos[i] = x;
}

// ------------------------------------------------------ batched sort
// Second, stable pass of the batched sort: the scene of each key-sorted slot.
__global__ void k_scene_keys(int n, const uint32_t* __restrict__ perm1, const int* __restrict__ scene_of,
                             uint32_t* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[s] = (uint32_t)scene_of[perm1[s]];
}

__global__ void k_gather_keys(int n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[s] = keys[perm[s]];
}

// -------------------------------------------------------- id-order output
__global__ void k_scatter_by_id(int n, const int* __restrict__ id, const float4* __restrict__ a4, float4* __restrict__ o4,
                                const float* __restrict__ a1, float* __restrict__ o1) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int d = id[s];
    if (a4) o4[d] = a4[s];
    if (a1) o1[d] = a1[s];
}

}  // namespace upf

// ======================================================================
// host side
// ======================================================================
using namespace upf;

#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return UPF_ERR_HIP;  \
    } while (0)

enum ProfClass { PC_SORT = 0, PC_GATHER, PC_CELLS, PC_PSI, PC_NBR_BUILD, PC_NBR_EXT, PC_DENSITY, PC_FORCE,
                 PC_INTEGRATE, PC_WCSPH_DENSITY, PC_WCSPH_FORCE, PC_RIGID, PC_UNION, PC_NBR_BUILD_EXT,
                 PC_NBR_BUILD_DENSITY, PC_COUNT };
static const char* kProfNames[PC_COUNT] = {"radix_sort", "gather", "cell_table", "boundary_volume",
                                          "neighbour_lists", "external_force_predict", "predicted_density",
                                          "pressure_force_predict", "integrate", "wcsph_density", "wcsph_force",
                                          "rigid_bodies", "neighbour_union", "neighbour_lists_external_force",
                                          "neighbour_lists_wcsph_density"};

struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, int>> pending;  // (class, event index of start; stop = start + 1)
    int used = 0;
    double ms[PC_COUNT] = {};
    long long cnt[PC_COUNT] = {};
    void begin(int cls, hipStream_t st) {
        if (!on) return;
        if (used + 2 > (int)pool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                pool.push_back(e);
            }
        }
        pending.push_back(std::make_pair(cls, used));
        (void)hipEventRecord(pool[used], st);
        used += 2;
    }
    void end(hipStream_t st) {
        if (!on || pending.empty()) return;
        (void)hipEventRecord(pool[pending.back().second + 1], st);
    }
    // start and stop events for a launch that records them itself
    // (hipExtLaunchKernel: the dispatch's own start and end, no marker
    // packets around it); false when profiling is off
    bool span(int cls, hipEvent_t& e0, hipEvent_t& e1) {
        if (!on) return false;
        if (used + 2 > (int)pool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return false;
                pool.push_back(e);
            }
        }
        pending.push_back(std::make_pair(cls, used));
        e0 = pool[used];
        e1 = pool[used + 1];
        used += 2;
        return true;
    }
    // Folds the recorded spans into the totals once the last of them has
    // completed (stream order: then all have), and frees the events for
    // reuse; otherwise leaves them for a later call.  A span is never read
    // before its stop event reports completion.
    void collect() {
        if (pending.empty() || hipEventQuery(pool[pending.back().second + 1]) != hipSuccess) return;
        for (auto& pr : pending) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, pool[pr.second], pool[pr.second + 1]) == hipSuccess) {
                ms[pr.first] += t;
                cnt[pr.first] += 1;
            }
        }
        pending.clear();
        used = 0;
    }
    ~Prof() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct upf_sim {
    Prof prof;
    upf_params P;
    int cap = 0;
    int n = 0;
    int cur = 0;            // which Slots buffer holds the current state
    bool uploaded = false;
    bool psi_ready = false;
    bool rigid = false;
    int n_bodies = 0;
    int last_method = UPF_PCISPH_STATIC;
    bool union_ready = false;  // this step's union lists are built (static PCISPH)
    hipStream_t st = nullptr;
    Slots S[2];
    Work W;
    Ctl* ctl = nullptr;
    Ctl* h_ctl = nullptr;   // pinned mirror
    SceneCtl* sctl = nullptr;
    SceneCtl* h_sctl = nullptr;  // pinned mirror of scene 0
    CtlMirror* mirror = nullptr; // pinned, written by k_publish
    CtlMirror* mirror_dev = nullptr;  // its device address
    uint32_t mirror_seq = 0;
    // scenes (a plain simulation is one scene with no per-slot arrays)
    int n_scenes = 1;
    int sps = MAX_SLOTS;          // max-reduction slots per scene
    Dom* d_doms = nullptr;
    int* d_scene_of = nullptr;
    // batched form: 64-slot chunks of the scenes and the per-iteration lists
    int n_chunks = 0;
    int* d_chunk_begin = nullptr;
    int* d_chunk_end = nullptr;
    int* d_sc_chunk0 = nullptr;
    int* d_sc_nchunks = nullptr;
    int* d_act = nullptr;
    int* d_act_n = nullptr;
    uint32_t* b_keys1 = nullptr;  // batched sort: key-sorted keys, their permutation, scene keys
    uint32_t* b_perm1 = nullptr;
    uint32_t* b_sk = nullptr;
    uint32_t* b_sk2 = nullptr;
    int* d_kstat = nullptr;       // per-step per-scene iteration counts and max densities
    float* d_rstat = nullptr;
    int stat_cap = 0;
    int steps_done = 0;
    upf_tuning tune = {256, 8192, 4, 0, 4, 1, 8, 0, 4, 1, 0, 1};
    bool static_lists = false;  // the lists being built serve the static PCISPH step only
    float* luts = nullptr;  // 5 * 301
    void* sort_tmp = nullptr;
    uint32_t ncells = 0;
    std::vector<int> k_hist;
    std::vector<uint32_t> key0;  // z-index of every particle's uploaded position, by upload id
    std::vector<float> rho_hist;
    std::vector<void*> allocs;
    // rigid bodies
    upf_body* d_bodies = nullptr;
    int* d_body_start = nullptr;       // first member slot of each body (n_bodies + 1, -1 when empty)
    uint32_t* d_body_keys = nullptr;   // scratch: body key per slot
    uint32_t* d_body_keys2 = nullptr;
    uint32_t* d_body_members = nullptr;  // slots grouped by body, ascending within a body
    float* d_body_terms = nullptr;       // 7 rows of cap: the members' force, torque and weight terms

    StepArgs args() const {
        StepArgs A;
        A.P = P;
        A.lut_factor = (float)(1.0 / (double)P.support_radius);
        A.inv_mass_f = 1.0f / P.initial_mass;
        A.inv_mass_d = (float)(1.0 / (double)P.initial_mass);
        A.rest_vol = P.initial_mass / P.rest_density;
        A.dens_sq = P.rest_density * P.rest_density;
        A.n = n;
        A.stride = cap;
        A.nbr_ld = (P.max_neighbors + 3) & ~3;
        A.dom0.scales[0] = P.scales[0];
        A.dom0.scales[1] = P.scales[1];
        A.dom0.scales[2] = P.scales[2];
        for (int a = 0; a < 3; ++a) {
            A.dom0.cmin[a] = P.collision_min[a];
            A.dom0.cmax[a] = P.collision_max[a];
        }
        A.dom0.block_shift = P.block_shift;
        A.dom0.grid_dim = P.grid_dim;
        A.dom0.cell_off = 0;
        A.dom0.slot_off = 0;
        A.dom0.n = n;
        A.dom0.delta = P.density_error_factor;
        A.doms = n_scenes > 1 ? d_doms : nullptr;
        A.scene_of = n_scenes > 1 ? d_scene_of : nullptr;
        A.sctl = sctl;
        A.max_slots = W.max_slots;
        A.n_scenes = n_scenes;
        A.slots_per_scene = sps;
        A.code_lists = tune.loop_list;
        A.xcd = tune.xcd_remap;
        A.n_chunks = n_chunks;
        A.chunk_begin = d_chunk_begin;
        A.chunk_end = d_chunk_end;
        A.sc_chunk0 = d_sc_chunk0;
        A.sc_nchunks = d_sc_nchunks;
        A.act = n_scenes > 1 ? d_act : nullptr;
        A.act_n = d_act_n;
        return A;
    }
    template <class T> hipError_t alloc(T** p, size_t count) {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) allocs.push_back((void*)*p);
        return e;
    }
};

static inline dim3 grid_for(int n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }
// block and grid of the persistent (grid-stride) kernels, from the tuning
static inline dim3 block_loop(const upf_sim* s) { return dim3((unsigned)s->tune.loop_block); }
static inline dim3 grid_persist(const upf_sim* s, int n) {
    const int b = s->tune.loop_block;
    const int g = std::max(1, std::min((n + b - 1) / b, s->tune.grid_cap));
    return dim3((unsigned)((g + 7) / 8 * 8));  // a multiple of 8 (XCD remap)
}

static int alloc_slots(upf_sim* s, Slots& S, int cap) {
    HIPCHK(s->alloc(&S.pos, cap));
    HIPCHK(s->alloc(&S.vel, cap));
    HIPCHK(s->alloc(&S.id, cap));
    HIPCHK(s->alloc(&S.p_last, cap));
    HIPCHK(s->alloc(&S.p_prev, cap));
    HIPCHK(s->alloc(&S.rho_last, cap));
    HIPCHK(s->alloc(&S.psi, cap));
    HIPCHK(s->alloc(&S.x0, cap));
    HIPCHK(s->alloc(&S.body, cap));
    HIPCHK(s->alloc(&S.rel, cap));
    return UPF_OK;
}

// sort + gather + cell table (the first part of every step)
static int sort_and_bin(upf_sim* s) {
    const int n = s->n;
    Work& W = s->W;
    s->prof.begin(PC_SORT, s->st);
    if (s->n_scenes == 1) {
        HIPCHK(radix_sort(W.keys, nullptr, W.keys_sorted, W.perm, n, 30, s->sort_tmp, s->st));
    } else {
        // stable sort by (scene, key): keys first, then the scene as the most
        // significant digit; scenes keep their slot ranges
        HIPCHK(radix_sort(W.keys, nullptr, s->b_keys1, s->b_perm1, n, 30, s->sort_tmp, s->st));
        hipLaunchKernelGGL(k_scene_keys, grid_for(n), dim3(BLOCK), 0, s->st, n, s->b_perm1, s->d_scene_of, s->b_sk);
        const int sbits = s->n_scenes <= 256 ? 8 : 16;
        HIPCHK(radix_sort(s->b_sk, s->b_perm1, s->b_sk2, W.perm, n, sbits, s->sort_tmp, s->st));
        hipLaunchKernelGGL(k_gather_keys, grid_for(n), dim3(BLOCK), 0, s->st, n, W.perm, W.keys, W.keys_sorted);
    }
    s->prof.end(s->st);
    const int nxt = 1 - s->cur;
    s->prof.begin(PC_GATHER, s->st);
    hipLaunchKernelGGL(k_gather, grid_for(n), dim3(BLOCK), 0, s->st, n, W.perm, s->S[s->cur], s->S[nxt], W.p, W.rho,
                       s->rigid ? 1 : 0);
    s->prof.end(s->st);
    s->cur = nxt;
    s->prof.begin(PC_CELLS, s->st);
    hipLaunchKernelGGL(k_fill_cells, dim3((s->ncells + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s->st, W.cells, s->ncells);
    s->prof.end(s->st);
    s->prof.begin(PC_CELLS, s->st);
    hipLaunchKernelGGL(k_cell_bounds, grid_for(n), dim3(BLOCK), 0, s->st, s->args(), W.keys_sorted, W.cells);
    s->prof.end(s->st);
    HIPCHK(hipGetLastError());
    return UPF_OK;
}

// The loop state after all work enqueued so far: a one-thread kernel writes
// it to pinned host memory with a sequence number the host spins on (a
// blocking stream synchronisation wakes the host up later than the GPU
// finishes, and the GPU idles meanwhile).
static int read_ctl(upf_sim* s) {
    const uint32_t seq = ++s->mirror_seq;
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s->st, s->ctl, s->sctl, s->mirror_dev, seq);
    HIPCHK(hipGetLastError());
    const volatile uint32_t* flag = &s->mirror->seq;
    for (uint64_t spins = 1; *flag != seq; ++spins) {
        if ((spins & 4095u) == 0) {
            const hipError_t e = hipStreamQuery(s->st);
            if (e == hipSuccess) {
                if (*flag == seq) break;
                return UPF_ERR_HIP;  // the stream drained without the mirror update
            }
            if (e != hipErrorNotReady) return UPF_ERR_HIP;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    *s->h_ctl = s->mirror->ctl;
    *s->h_sctl = s->mirror->scene0;
    s->prof.collect();
    return UPF_OK;
}

static int err_code(uint32_t bits) {
    const int order[] = {UPF_ERR_NAN, UPF_ERR_RANGE, UPF_ERR_NBR_OVERFLOW, UPF_ERR_ISOLATED};
    for (int c : order)
        if (bits & (1u << c)) return c;
    return UPF_OK;
}

template <int MODE>
static void enqueue_iteration(upf_sim* s, const StepArgs& A, int t) {
    const dim3 g = grid_persist(s, s->n), b = block_loop(s);
    Slots& S = s->S[s->cur];
    const int lw = s->tune.loop_width, ll = s->tune.loop_list;
    // the iteration passes are timed by their own dispatches when profiling
    // (their device time is the figure the roofline is quoted on)
#define UPF_PASS(KERNEL, CLS)                                                                               \
    do {                                                                                                   \
        hipEvent_t e0, e1;                                                                                 \
        if (s->prof.span(CLS, e0, e1))                                                                     \
            hipExtLaunchKernelGGL(KERNEL, g, b, 0, s->st, e0, e1, 0, t, A, S, s->W, s->luts, s->ctl);      \
        else                                                                                               \
            hipLaunchKernelGGL(KERNEL, g, b, 0, s->st, t, A, S, s->W, s->luts, s->ctl);                    \
    } while (0)
    // union lists (the step built them): both passes read their neighbours'
    // records from LDS; the Akinci density needs boundary volumes by slot and
    // keeps the delta-coded lists
    constexpr int DU = MODE == UPF_BOUNDARY_AKINCI ? 1 : 2;
    if (ll == 2 && s->union_ready) {
        if (lw >= 8) UPF_PASS((k_density<MODE, DU, 8>), PC_DENSITY);
        else UPF_PASS((k_density<MODE, DU, 4>), PC_DENSITY);
        if (lw >= 8) UPF_PASS((k_force<MODE, 2, 8>), PC_FORCE);
        else UPF_PASS((k_force<MODE, 2, 4>), PC_FORCE);
        return;
    }
    if (ll && lw >= 8) UPF_PASS((k_density<MODE, 1, 8>), PC_DENSITY);
    else if (ll) UPF_PASS((k_density<MODE, 1, 4>), PC_DENSITY);
    else if (lw >= 8) UPF_PASS((k_density<MODE, 0, 8>), PC_DENSITY);
    else UPF_PASS((k_density<MODE, 0, 4>), PC_DENSITY);
    if (ll && lw >= 8) UPF_PASS((k_force<MODE, 1, 8>), PC_FORCE);
    else if (ll) UPF_PASS((k_force<MODE, 1, 4>), PC_FORCE);
    else if (lw >= 8) UPF_PASS((k_force<MODE, 0, 8>), PC_FORCE);
    else UPF_PASS((k_force<MODE, 0, 4>), PC_FORCE);
#undef UPF_PASS
}

// PredictionCorrectionStep (:5367-5424).  The loop test runs on the device;
// the host enqueues iterations ahead (as many as the previous step needed)
// and reads the decision back once the enqueued ones are done.  Iterations
// beyond the device's decision return at their first instruction.
template <int MODE>
static int pcisph_loop(upf_sim* s, const StepArgs& A) {
    const int maxl = std::max(s->P.max_loops, 1);
    const int lo = std::min(std::max(s->P.min_loops, 1), maxl);
    // an iteration past the decision costs two launches that return at once,
    // a read-back round trip several of them: enqueue as many as the previous
    // step needed (at most spec_first past the minimum), then `speculate`
    // more per round trip
    const int prev = s->k_hist.empty() ? lo : s->k_hist.back();
    const int spec = std::max(1, s->tune.speculate);
    const int ahead = std::min(std::min(std::max(lo, prev), lo + std::max(0, s->tune.spec_first)), maxl);
    hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(256), 0, s->st, A, s->ctl);
    int t = 0;
    for (; t < ahead; ++t) enqueue_iteration<MODE>(s, A, t);
    for (;;) {
        int rc = read_ctl(s);
        if (rc) return rc;
        const int want = s->h_ctl->want_any;  // final once want <= enqueued iterations
        if (want <= t || (s->h_ctl->err & (1u << UPF_ERR_NAN))) break;
        const int stop = std::min(t + spec, maxl);
        for (; t < stop; ++t) enqueue_iteration<MODE>(s, A, t);
    }
    s->k_hist.push_back(s->h_ctl->want_any);
    s->rho_hist.push_back(s->n_scenes == 1 ? s->h_sctl->last_max : 0.0f);
    return UPF_OK;
}

// The whole correction loop without a host decision: all max_loops
// iterations are enqueued (for graph capture); those at or beyond every
// scene's decision return at their first instruction.
template <int MODE>
static void pcisph_loop_device(upf_sim* s, const StepArgs& A) {
    const int maxl = std::max(s->P.max_loops, 1);
    hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(256), 0, s->st, A, s->ctl);
    for (int t = 0; t < maxl; ++t) enqueue_iteration<MODE>(s, A, t);
}

static void build_lists(upf_sim* s, const StepArgs& A) {
    Slots& S = s->S[s->cur];
    s->prof.begin(PC_NBR_BUILD, s->st);
    const dim3 g((unsigned)(((s->n + 63) / 64 + 7) / 8 * 8));  // a multiple of 8 (XCD remap)
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, g, dim3(64), 0, s->st, A, S.pos, s->W.keys_sorted, s->W.cells, s->W.nbr, s->W.ncount,
                           s->ctl);
    };
    const int r = s->tune.nbr_rounds;
    // the correction loop of the delta-coded lists never reads the bins
    const bool bins = !(A.code_lists && s->static_lists);
    if (s->tune.nbr_tile == 256) {
        launch(k_nbr_build_tile<256>);
    } else if (s->tune.nbr_tile == 512) {
        launch(k_nbr_build_tile<512>);
    } else if (s->tune.nbr_stage == 0) {
        if (!bins && s->tune.nbr_skip) {
            if (r == 2) launch(k_nbr_build<2, 0, 0, 1>);
            else if (r == 8) launch(k_nbr_build<8, 0, 0, 1>);
            else launch(k_nbr_build<4, 0, 0, 1>);
        } else if (!bins) {
            if (r == 2) launch(k_nbr_build<2, 0, 0>);
            else if (r == 8) launch(k_nbr_build<8, 0, 0>);
            else launch(k_nbr_build<4, 0, 0>);
        } else {
            if (r == 2) launch(k_nbr_build<2, 0, 1>);
            else if (r == 8) launch(k_nbr_build<8, 0, 1>);
            else launch(k_nbr_build<4, 0, 1>);
        }
    } else if (s->tune.nbr_stage == 32) {
        if (r == 2) launch(k_nbr_build<2, 32, 1>);
        else if (r == 8) launch(k_nbr_build<8, 32, 1>);
        else launch(k_nbr_build<4, 32, 1>);
    } else {
        if (r == 2) launch(k_nbr_build<2, 64, 1>);
        else if (r == 8) launch(k_nbr_build<8, 64, 1>);
        else launch(k_nbr_build<4, 64, 1>);
    }
    s->prof.end(s->st);
}

// The parts of a static-boundary step: the head (sort, cell table, boundary
// volumes on the first step, lists, external forces and first prediction)
// and the tail (integration, per-scene record).  The correction loop runs
// between them; its iterations only touch the slot buffer the head sorted
// into, so a batch captures head, loop chunks and tail as separate graphs.
static int step_static_head(upf_sim* s) {
    int rc = sort_and_bin(s);
    if (rc) return rc;
    const StepArgs A = s->args();
    Slots& S = s->S[s->cur];
    const dim3 g = grid_for(s->n);
    const dim3 gp = grid_persist(s, s->n), bl = block_loop(s);
    const bool akinci = s->P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI;
    if (!s->psi_ready) {
        s->prof.begin(PC_PSI, s->st);
        hipLaunchKernelGGL(k_psi, g, dim3(BLOCK), 0, s->st, A, S.pos, s->W.cells, s->luts, S.psi, s->ctl);
        s->prof.end(s->st);
        s->psi_ready = true;
    }
    // the fused list build + external forces needs the cell-by-cell build
    // without LDS staging and the delta-coded loop lists (no bins)
    if (s->tune.nbr_fuse && s->tune.nbr_stage == 0 && s->tune.nbr_tile == 0 && A.code_lists) {
        s->prof.begin(PC_NBR_BUILD_EXT, s->st);
        const dim3 gb((unsigned)(((s->n + 63) / 64 + 7) / 8 * 8));  // a multiple of 8 (XCD remap)
        auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, s->st, A, S, s->W, s->luts, s->ctl); };
        const int r = s->tune.nbr_rounds, k = s->tune.nbr_skip;
        // slot-major packed lists when only this kernel reads them (the delta-
        // coded loop lists); the chunk unions read the column-major ones
        auto go = [&](auto sm) {
            constexpr bool SM = decltype(sm)::value;
            if (akinci) {
                if (r == 2) k ? launch(k_nbr_build_ext<2, 1, UPF_BOUNDARY_AKINCI, SM>) : launch(k_nbr_build_ext<2, 0, UPF_BOUNDARY_AKINCI, SM>);
                else if (r == 8) k ? launch(k_nbr_build_ext<8, 1, UPF_BOUNDARY_AKINCI, SM>) : launch(k_nbr_build_ext<8, 0, UPF_BOUNDARY_AKINCI, SM>);
                else k ? launch(k_nbr_build_ext<4, 1, UPF_BOUNDARY_AKINCI, SM>) : launch(k_nbr_build_ext<4, 0, UPF_BOUNDARY_AKINCI, SM>);
            } else {
                if (r == 2) k ? launch(k_nbr_build_ext<2, 1, UPF_BOUNDARY_CPU_REF, SM>) : launch(k_nbr_build_ext<2, 0, UPF_BOUNDARY_CPU_REF, SM>);
                else if (r == 8) k ? launch(k_nbr_build_ext<8, 1, UPF_BOUNDARY_CPU_REF, SM>) : launch(k_nbr_build_ext<8, 0, UPF_BOUNDARY_CPU_REF, SM>);
                else k ? launch(k_nbr_build_ext<4, 1, UPF_BOUNDARY_CPU_REF, SM>) : launch(k_nbr_build_ext<4, 0, UPF_BOUNDARY_CPU_REF, SM>);
            }
        };
        if (s->tune.loop_list == 1) go(std::true_type{});
        else go(std::false_type{});
        s->prof.end(s->st);
    } else {
        s->static_lists = true;
        build_lists(s, A);
        s->static_lists = false;
        s->prof.begin(PC_NBR_EXT, s->st);
        if (akinci)
            hipLaunchKernelGGL(k_ext_static<UPF_BOUNDARY_AKINCI>, gp, bl, 0, s->st, A, S, s->W, s->luts, s->ctl);
        else
            hipLaunchKernelGGL(k_ext_static<UPF_BOUNDARY_CPU_REF>, gp, bl, 0, s->st, A, S, s->W, s->luts, s->ctl);
        s->prof.end(s->st);
    }
    s->union_ready = s->tune.loop_list == 2;
    if (s->union_ready) {
        s->prof.begin(PC_UNION, s->st);
        hipLaunchKernelGGL(k_union, dim3((unsigned)((s->n + UNION_CHUNK - 1) / UNION_CHUNK)), dim3(UNION_CHUNK), 0,
                           s->st, A, s->W);
        s->prof.end(s->st);
    }
    HIPCHK(hipGetLastError());
    return UPF_OK;
}

// correction iterations [t0, t1) of the step whose head ran (no loop init)
static void step_static_iterations(upf_sim* s, int t0, int t1) {
    const StepArgs A = s->args();
    const bool akinci = s->P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI;
    for (int t = t0; t < t1; ++t) {
        if (akinci) enqueue_iteration<UPF_BOUNDARY_AKINCI>(s, A, t);
        else enqueue_iteration<UPF_BOUNDARY_CPU_REF>(s, A, t);
    }
}

static int step_static_tail(upf_sim* s) {
    const StepArgs A = s->args();
    Slots& S = s->S[s->cur];
    s->prof.begin(PC_INTEGRATE, s->st);
    hipLaunchKernelGGL(k_integrate, grid_persist(s, s->n), block_loop(s), 0, s->st, A, S, s->W, 1, 0,
                       s->rigid ? 1 : 0, s->ctl);
    s->prof.end(s->st);
    if (s->d_kstat)
        hipLaunchKernelGGL(k_record, dim3(1), dim3(256), 0, s->st, A, s->ctl, s->d_kstat, s->d_rstat, s->stat_cap);
    HIPCHK(hipGetLastError());
    return UPF_OK;
}

// CpuParticlePhysicsStaticBoundariesPCISPH (src/cpu/UnifiedPhysics.cpp:4220-4238).
// device_loop: the loop decision stays on the device and every iteration up
// to max_loops is enqueued; otherwise the host reads the decision back.
static int step_pcisph_static(upf_sim* s, bool device_loop) {
    int rc = step_static_head(s);
    if (rc) return rc;
    const StepArgs A = s->args();
    const bool akinci = s->P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI;
    if (device_loop) {
        if (akinci) pcisph_loop_device<UPF_BOUNDARY_AKINCI>(s, A);
        else pcisph_loop_device<UPF_BOUNDARY_CPU_REF>(s, A);
    } else {
        rc = akinci ? pcisph_loop<UPF_BOUNDARY_AKINCI>(s, A) : pcisph_loop<UPF_BOUNDARY_CPU_REF>(s, A);
        if (rc) return rc;
    }
    return step_static_tail(s);
}

// CpuParticlePhysicsVersatileCouplingPCISPH (src/cpu/UnifiedPhysics.cpp:4991-5024)
static int step_pcisph_rigid(upf_sim* s) {
    int rc = sort_and_bin(s);
    if (rc) return rc;
    s->union_ready = false;  // the coupling loop reads the delta-coded lists
    const StepArgs A = s->args();
    Slots& S = s->S[s->cur];
    const dim3 g = grid_for(s->n);
    const bool akinci = s->P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI;
    // fused: rigid lists and volumes first, then the liquid lists and the
    // coupling pass as one kernel (the liquid build keeps the rigid counts)
    const bool fuse = s->tune.nbr_fuse && s->tune.nbr_stage == 0 && s->tune.nbr_tile == 0;
    if (!fuse) build_lists(s, A);  // liquid owners
    s->prof.begin(PC_NBR_BUILD, s->st);
    if (s->tune.nbr_tile == 0) {
        // RIGID_SPLIT waves per window of 64 slots, a multiple of 8 (XCD remap)
        const dim3 gb((unsigned)(((s->n + 63) / 64 * RIGID_SPLIT + 7) / 8 * 8));
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, gb, dim3(64), 0, s->st, A, S.pos, s->W.keys_sorted, s->W.cells, s->W.nbr,
                               s->W.ncount, s->ctl);
        };
        const int r = s->tune.nbr_rounds;
        // slot-major when the fused coupling kernel reads them
        if (fuse) {
            if (r == 2) launch(k_nbr_build_rigid<2, true>);
            else if (r == 8) launch(k_nbr_build_rigid<8, true>);
            else launch(k_nbr_build_rigid<4, true>);
        } else {
            if (r == 2) launch(k_nbr_build_rigid<2, false>);
            else if (r == 8) launch(k_nbr_build_rigid<8, false>);
            else launch(k_nbr_build_rigid<4, false>);
        }
    } else {
        hipLaunchKernelGGL(k_nbr_build_box, g, dim3(BLOCK), 0, s->st, A, S.pos, s->W.cells, (int)UPF_RIGID, s->W.nbr,
                           s->W.ncount, s->ctl);
    }
    s->prof.end(s->st);
    if (akinci) {
        s->prof.begin(PC_PSI, s->st);
        if (!s->psi_ready) {  // walls: once (they do not move)
            hipLaunchKernelGGL(k_psi, g, dim3(BLOCK), 0, s->st, A, S.pos, s->W.cells, s->luts, S.psi, s->ctl);
            s->psi_ready = true;
        }
        hipLaunchKernelGGL(k_psi_rigid, g, dim3(BLOCK), 0, s->st, s->n, S.pos, s->W.nbr, fuse ? 1 : A.stride,
                           fuse ? A.nbr_ld : 1, s->W.ncount, s->P,
                           s->luts, S.psi, s->ctl);
        s->prof.end(s->st);
    }
    if (fuse) {
        s->prof.begin(PC_NBR_BUILD_EXT, s->st);
        const dim3 gb((unsigned)(((s->n + 63) / 64 + 7) / 8 * 8));  // a multiple of 8 (XCD remap)
        auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, s->st, A, S, s->W, s->luts, s->ctl); };
        const int r = s->tune.nbr_rounds;
        if (akinci) {
            if (r == 2) launch(k_nbr_build_coupling<2, UPF_BOUNDARY_AKINCI, true>);
            else if (r == 8) launch(k_nbr_build_coupling<8, UPF_BOUNDARY_AKINCI, true>);
            else launch(k_nbr_build_coupling<4, UPF_BOUNDARY_AKINCI, true>);
        } else {
            if (r == 2) launch(k_nbr_build_coupling<2, UPF_BOUNDARY_CPU_REF, true>);
            else if (r == 8) launch(k_nbr_build_coupling<8, UPF_BOUNDARY_CPU_REF, true>);
            else launch(k_nbr_build_coupling<4, UPF_BOUNDARY_CPU_REF, true>);
        }
        s->prof.end(s->st);
    } else {
        s->prof.begin(PC_NBR_EXT, s->st);
        if (akinci)
            hipLaunchKernelGGL(k_ext_coupling<UPF_BOUNDARY_AKINCI>, g, dim3(BLOCK), 0, s->st, A, S, s->W, s->luts, s->ctl);
        else
            hipLaunchKernelGGL(k_ext_coupling<UPF_BOUNDARY_CPU_REF>, g, dim3(BLOCK), 0, s->st, A, S, s->W, s->luts, s->ctl);
        s->prof.end(s->st);
    }
    rc = s->P.boundary_pressure_mode == UPF_BOUNDARY_AKINCI ? pcisph_loop<UPF_BOUNDARY_AKINCI>(s, A)
                                                            : pcisph_loop<UPF_BOUNDARY_CPU_REF>(s, A);
    if (rc) return rc;
    s->prof.begin(PC_INTEGRATE, s->st);
    hipLaunchKernelGGL(k_integrate, grid_persist(s, s->n), block_loop(s), 0, s->st, A, S, s->W, 1, 1, 1, s->ctl);
    s->prof.end(s->st);
    // rigid bodies: members grouped by body in slot order, one thread per body
    s->prof.begin(PC_RIGID, s->st);
    // body keys below 255: one 8-bit pass of the stable sort (non-rigid: 255)
    const int body_bits = s->n_bodies < 255 ? 8 : 16;
    hipLaunchKernelGGL(k_body_keys, g, dim3(BLOCK), 0, s->st, s->n, S.body, S.pos, s->d_body_keys,
                       body_bits == 8 ? 0xFFu : 0xFFFFu);
    HIPCHK(radix_sort(s->d_body_keys, nullptr, s->d_body_keys2, s->d_body_members, s->n, body_bits, s->sort_tmp, s->st));
    HIPCHK(hipMemsetAsync(s->d_body_start, 0xFF, (size_t)(s->n_bodies + 1) * sizeof(int), s->st));
    hipLaunchKernelGGL(k_body_bounds, g, dim3(BLOCK), 0, s->st, s->n, s->d_body_keys2, s->n_bodies, s->d_body_start);
    hipLaunchKernelGGL(k_body_terms, g, dim3(BLOCK), 0, s->st, s->n, s->n_bodies, s->d_body_keys2, s->d_body_members,
                       s->d_bodies, S.pos, s->W.force, S.rel, s->P, s->d_body_terms, s->cap);
    hipLaunchKernelGGL(k_rigid_bodies, dim3(s->n_bodies), dim3(RB_CHUNK), 0, s->st, s->n_bodies, s->d_bodies,
                       s->d_body_start, s->d_body_terms, s->cap, s->P);
    hipLaunchKernelGGL(k_rigid_sync, g, dim3(BLOCK), 0, s->st, s->n, S, s->d_bodies);
    s->prof.end(s->st);
    HIPCHK(hipGetLastError());
    return UPF_OK;
}

// The methods whose pressure is the state pressure W.pres (no correction loop).
static inline bool sph_family(int method) { return method == UPF_WCSPH_STATIC || method == UPF_SPH_STATIC; }

static int step_wcsph_static(upf_sim* s, bool tait) {
    int rc = sort_and_bin(s);
    if (rc) return rc;
    const StepArgs A = s->args();
    Slots& S = s->S[s->cur];
    const dim3 g = grid_for(s->n);
    if (s->tune.nbr_fuse && s->tune.nbr_stage == 0 && s->tune.nbr_tile == 0) {
        // the list build (with bins) and the density pass as one kernel
        s->prof.begin(PC_NBR_BUILD_DENSITY, s->st);
        const dim3 gb((unsigned)(((s->n + 63) / 64 + 7) / 8 * 8));  // a multiple of 8 (XCD remap)
        auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, s->st, A, S, s->W, s->luts, s->ctl); };
        const int r = s->tune.nbr_rounds;
        if (tait) {
            if (r == 2) launch(k_nbr_build_wcsph<2, true>);
            else if (r == 8) launch(k_nbr_build_wcsph<8, true>);
            else launch(k_nbr_build_wcsph<4, true>);
        } else {
            if (r == 2) launch(k_nbr_build_wcsph<2, false>);
            else if (r == 8) launch(k_nbr_build_wcsph<8, false>);
            else launch(k_nbr_build_wcsph<4, false>);
        }
        s->prof.end(s->st);
    } else {
        build_lists(s, A);
        s->prof.begin(PC_WCSPH_DENSITY, s->st);
        if (tait) hipLaunchKernelGGL(k_density_wcsph<true>, g, dim3(BLOCK), 0, s->st, A, S, s->W, s->luts, s->ctl);
        else hipLaunchKernelGGL(k_density_wcsph<false>, g, dim3(BLOCK), 0, s->st, A, S, s->W, s->luts, s->ctl);
        s->prof.end(s->st);
    }
    s->prof.begin(PC_WCSPH_FORCE, s->st);
    hipLaunchKernelGGL(k_force_wcsph, g, dim3(BLOCK), 0, s->st, A, S, s->W, s->luts, s->ctl);
    s->prof.end(s->st);
    s->prof.begin(PC_INTEGRATE, s->st);
    hipLaunchKernelGGL(k_integrate, grid_persist(s, s->n), block_loop(s), 0, s->st, A, S, s->W, 0, 0,
                       s->rigid ? 1 : 0, s->ctl);
    s->prof.end(s->st);
    HIPCHK(hipGetLastError());
    rc = read_ctl(s);
    if (rc) return rc;
    s->k_hist.push_back(0);
    s->rho_hist.push_back(0.0f);
    return UPF_OK;
}

extern "C" {

const char* upf_version(void) { return "upf-mi355x 0.1 (gfx950)"; }

// Device buffers of a simulation of `capacity` slots in `n_scenes` scenes
// whose cell tables hold `ncells` entries together.
// Thresholds of the table bin as a function of the squared distance:
// thr[k] = the smallest float d2 >= 0 with lut_bin(sqrt_rn(d2)) >= k, found by
// bisection over the (ordered) bit patterns of the non-negative floats; the
// host's sqrtf is the correctly rounded IEEE root.  out[2b] = thr[b],
// out[2b + 1] = thr[b + 1] for b = 0 .. 300, with thr[0] = 0, thr[301] = +inf.
static void build_bin_thresholds(float factor, float* out) {
    auto bin = [&](uint32_t u) {
        float d2;
        std::memcpy(&d2, &u, 4);
        const float v = std::sqrt(d2) * factor * 300.0f;
        if (!(v < 2147483648.0f)) return 300;  // the device conversion saturates
        const int idx = (int)v;
        return idx >= 300 ? 300 : idx;
    };
    std::vector<float> thr(302);
    thr[0] = 0.0f;
    const uint32_t inf_bits = 0x7f800000u;
    for (int k = 1; k <= 300; ++k) {
        uint32_t lo = 0, hi = inf_bits;  // bin(lo) < k <= bin(hi)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (bin(mid) >= k) hi = mid; else lo = mid;
        }
        std::memcpy(&thr[k], &hi, 4);
    }
    thr[301] = std::numeric_limits<float>::infinity();
    for (int b = 0; b <= 300; ++b) {
        out[2 * b] = thr[b];
        out[2 * b + 1] = thr[b + 1];
    }
}

static int sim_alloc(upf_sim* s, int capacity, uint32_t ncells, int n_scenes) {
    const upf_params* p = &s->P;
    s->cap = capacity;
    s->n_scenes = n_scenes;
    s->sps = n_scenes == 1 ? MAX_SLOTS : 16;
    s->ncells = ncells;
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) return UPF_ERR_HIP;
    if (alloc_slots(s, s->S[0], capacity) || alloc_slots(s, s->S[1], capacity)) return UPF_ERR_HIP;
    Work& W = s->W;
    const size_t nslots = (size_t)n_scenes * s->sps * 32;
    HIPCHK(s->alloc(&W.keys, capacity));
    HIPCHK(s->alloc(&W.keys_sorted, capacity));
    HIPCHK(s->alloc(&W.perm, capacity));
    HIPCHK(s->alloc(&W.force, capacity));
    HIPCHK(s->alloc(&W.fp, capacity));
    HIPCHK(s->alloc(&W.xs, capacity));
    HIPCHK(s->alloc(&W.p, capacity));
    HIPCHK(s->alloc(&W.rho, capacity));
    HIPCHK(s->alloc(&W.pres, capacity));
    HIPCHK(s->alloc(&W.xpd, capacity));
    HIPCHK(s->alloc(&W.vpg, capacity));
    HIPCHK(s->alloc(&W.max_slots, nslots));
    HIPCHK(s->alloc(&W.ncount, capacity));
    // column-major (k * capacity + slot) or, for the fused build, slot-major
    // (slot * pitch + k, the pitch rounded up to four entries)
    HIPCHK(s->alloc(&W.nbr, (size_t)capacity * ((p->max_neighbors + 3) & ~3)));
    HIPCHK(s->alloc(&W.nbw, (size_t)capacity * ((p->max_neighbors + 7) / 8)));
    HIPCHK(s->alloc(&W.nbf, capacity));
    HIPCHK(s->alloc(&W.nbe, (size_t)capacity * NB_ESC));
    HIPCHK(s->alloc(&W.nbu, (size_t)capacity * ((p->max_neighbors + 7) / 8)));
    const size_t nchunks = ((size_t)capacity + UNION_CHUNK - 1) / UNION_CHUNK;
    HIPCHK(s->alloc(&W.uslots, nchunks * UNION_CAP));
    HIPCHK(s->alloc(&W.un, nchunks));
    HIPCHK(s->alloc(&W.cells, ncells));
    HIPCHK(s->alloc(&s->luts, LUT_TOTAL));
    HIPCHK(s->alloc(&s->ctl, 1));
    HIPCHK(s->alloc(&s->sctl, n_scenes));
    void* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, radix_sort_temp_bytes(capacity)));
    s->sort_tmp = tmp;
    s->allocs.push_back(tmp);
    HIPCHK(hipHostMalloc((void**)&s->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&s->h_sctl, sizeof(SceneCtl), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&s->mirror, sizeof(CtlMirror), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(s->mirror, 0, sizeof(CtlMirror));
    HIPCHK(hipHostGetDevicePointer((void**)&s->mirror_dev, s->mirror, 0));
    std::vector<float> L(LUT_TOTAL, 0.0f);
    build_luts(p->support_radius, L.data());
    const float factor = (float)(1.0 / (double)p->support_radius);  // StepArgs::lut_factor
    build_bin_thresholds(factor, L.data() + LUT_TH_OFF);
    HIPCHK(hipMemcpy(s->luts, L.data(), L.size() * sizeof(float), hipMemcpyHostToDevice));
    {
        // the sqrt-free bins equal the correctly rounded ones for every float
        // below 2 thr[300] (above it both are 300)
        unsigned long long* bad = nullptr;
        HIPCHK(s->alloc(&bad, 1));
        HIPCHK(hipMemset(bad, 0, sizeof(unsigned long long)));
        const float top = 2.0f * L[LUT_TH_OFF + 2 * 300];
        uint32_t top_bits;
        std::memcpy(&top_bits, &top, 4);
        hipLaunchKernelGGL(k_check_bins, dim3(4096), dim3(256), 0, s->st, factor, s->luts, top_bits, bad);
        unsigned long long h_bad = 1;
        HIPCHK(hipMemcpyAsync(&h_bad, bad, sizeof(h_bad), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        if (h_bad != 0) {
            std::fprintf(stderr, "upf: hardware square root bins disagree with the correctly rounded ones (%llu floats)\n", h_bad);
            return UPF_ERR_HIP;
        }
    }
    HIPCHK(hipMemset(s->ctl, 0, sizeof(Ctl)));
    HIPCHK(hipMemset(s->sctl, 0, (size_t)n_scenes * sizeof(SceneCtl)));
    HIPCHK(hipMemset(W.max_slots, 0, nslots * sizeof(uint32_t)));
    if (n_scenes > 1) {
        HIPCHK(s->alloc(&s->d_doms, n_scenes));
        HIPCHK(s->alloc(&s->d_scene_of, capacity));
        HIPCHK(s->alloc(&s->b_keys1, capacity));
        HIPCHK(s->alloc(&s->b_perm1, capacity));
        HIPCHK(s->alloc(&s->b_sk, capacity));
        HIPCHK(s->alloc(&s->b_sk2, capacity));
    }
    return UPF_OK;
}

// per-step statistics on the device (batched form and graph replay)
static int stats_alloc(upf_sim* s, int cap_steps) {
    s->stat_cap = cap_steps;
    HIPCHK(s->alloc(&s->d_kstat, (size_t)cap_steps * s->n_scenes));
    HIPCHK(s->alloc(&s->d_rstat, (size_t)cap_steps * s->n_scenes));
    return UPF_OK;
}

int upf_create(const upf_params* p, int capacity, upf_sim** out) {
    // list entries hold the neighbour slot in 23 bits
    if (!out || capacity <= 0 || capacity >= (1 << (32 - NB_SLOT_SHIFT))) return UPF_ERR_ARG;
    int rc = upf::check_params(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPF_ERR_NO_DEVICE;
    upf_sim* s = new upf_sim();
    s->P = *p;
    rc = sim_alloc(s, capacity, p->grid_dim * p->grid_dim * p->grid_dim, 1);
    if (rc != UPF_OK) {
        upf_destroy(s);
        return rc;
    }
    *out = s;
    return UPF_OK;
}

int upf_destroy(upf_sim* s) {
    if (!s) return UPF_ERR_ARG;
    if (s->st) (void)hipStreamSynchronize(s->st);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->d_bodies) (void)hipFree(s->d_bodies);
    if (s->d_body_start) (void)hipFree(s->d_body_start);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    if (s->h_sctl) (void)hipHostFree(s->h_sctl);
    if (s->mirror) (void)hipHostFree(s->mirror);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
    return UPF_OK;
}

int upf_set_params(upf_sim* s, const upf_params* p) {
    if (!s || !p) return UPF_ERR_ARG;
    int rc = upf::check_params(p);
    if (rc) return rc;
    if (p->grid_dim != s->P.grid_dim || p->block_size != s->P.block_size || p->block_shift != s->P.block_shift ||
        p->max_neighbors != s->P.max_neighbors || p->support_radius != s->P.support_radius)
        return UPF_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (p->scales[a] != s->P.scales[a]) return UPF_ERR_ARG;
    s->P = *p;
    return UPF_OK;
}

}  // extern "C"

// Uploads n particles; scene_of / scene_params give every particle's scene and
// the scenes' domains (null: one scene with the simulation's params).  p0 /
// rho0, if given, are the correction pressure and predicted density of the
// step before (a restart); else 0 and the rest density.
static int upload_common(upf_sim* s, int n, const float* pos4, const float* vel4, const int* type,
                         const int* scene_of, const upf_params* scene_params, const float* p0 = nullptr,
                         const float* rho0 = nullptr) {
    if (!s || n <= 0 || n > s->cap || !pos4 || !type) return UPF_ERR_ARG;
    std::vector<float> pos((size_t)4 * n), vel((size_t)4 * n, 0.0f), x0((size_t)4 * n);
    std::vector<int> id(n), body(n, -1);
    std::vector<uint32_t> key0(n);
    std::vector<float> rho(n, s->P.rest_density), zero(n, 0.0f), psi(n, s->P.initial_mass / s->P.rest_density);
    for (int i = 0; i < n; ++i) {
        const upf_params& Q = scene_of ? scene_params[scene_of[i]] : s->P;
        const int t = type[i];
        if (t != UPF_LIQUID && t != UPF_FROZEN && t != UPF_RIGID) return UPF_ERR_ARG;
        for (int a = 0; a < 3; ++a) {
            pos[4 * i + a] = pos4[4 * i + a];
            x0[4 * i + a] = pos4[4 * i + a];
            if (vel4) vel[4 * i + a] = vel4[4 * i + a];
        }
        int tb = t;
        float w;
        std::memcpy(&w, &tb, 4);
        pos[4 * i + 3] = w;
        x0[4 * i + 3] = w;
        vel[4 * i + 3] = 0.0f;
        id[i] = i;
        uint32_t k;
        if (!host_key(Q, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], k)) return UPF_ERR_RANGE;
        key0[i] = k;
        // the query box of every particle must stay inside the grid (beyond it the
        // reference's cell ranges overflow, src/cpu/zIndex.cpp:66)
        for (int a = 0; a < 3; ++a)
            if (!(Q.scales[a] * (pos4[4 * i + a] + Q.support_radius) < 1024.0f)) return UPF_ERR_RANGE;
    }
    Slots& S = s->S[0];
    s->cur = 0;
    s->n = n;
    const size_t b4 = (size_t)n * sizeof(float4), b1 = (size_t)n * sizeof(float);
    HIPCHK(hipMemcpyAsync(S.pos, pos.data(), b4, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.vel, vel.data(), b4, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.x0, x0.data(), b4, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.id, id.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.body, body.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(S.rel, 0, b4, s->st));
    HIPCHK(hipMemcpyAsync(S.p_last, zero.data(), b1, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.p_prev, zero.data(), b1, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.rho_last, rho.data(), b1, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.psi, psi.data(), b1, hipMemcpyHostToDevice, s->st));
    // loop arrays of the "previous step" (gathered into p_last / rho_last by the first sort)
    HIPCHK(hipMemcpyAsync(s->W.p, p0 ? p0 : zero.data(), b1, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->W.rho, rho0 ? rho0 : rho.data(), b1, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(s->ctl, 0, sizeof(Ctl), s->st));
    if (scene_of) HIPCHK(hipMemcpyAsync(s->d_scene_of, scene_of, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s->st));
    hipLaunchKernelGGL(k_keys, grid_for(n), dim3(BLOCK), 0, s->st, s->args(), S.pos, s->W.keys, s->ctl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
    s->uploaded = true;
    s->key0.swap(key0);
    s->psi_ready = false;
    s->k_hist.clear();
    s->rho_hist.clear();
    s->steps_done = 0;
    return UPF_OK;
}

extern "C" {

int upf_upload_particles(upf_sim* s, int n, const float* pos4, const float* vel4, const int* type) {
    if (s && s->n_scenes != 1) return UPF_ERR_STATE;
    return upload_common(s, n, pos4, vel4, type, nullptr, nullptr);
}


int upf_step(upf_sim* s, int nsteps, int method) {
    if (!s || nsteps < 0) return UPF_ERR_ARG;
    if (!s->uploaded) return UPF_ERR_STATE;
    if (method == UPF_PCISPH_RIGID && !s->rigid) return UPF_ERR_STATE;
    if (method == UPF_PCISPH_RIGID && s->P.add_wall_weight) return UPF_ERR_ARG;
    for (int k = 0; k < nsteps; ++k) {
        int rc;
        if (method == UPF_PCISPH_STATIC) rc = step_pcisph_static(s, false);
        else if (method == UPF_WCSPH_STATIC) rc = step_wcsph_static(s, true);
        else if (method == UPF_SPH_STATIC) rc = step_wcsph_static(s, false);
        else if (method == UPF_PCISPH_RIGID) rc = step_pcisph_rigid(s);
        else return UPF_ERR_ARG;
        s->last_method = method;
        if (rc) return rc;
        const int e = err_code(s->h_ctl->err);
        if (e) return e;
    }
    if (nsteps > 0 && (method == UPF_PCISPH_STATIC || method == UPF_PCISPH_RIGID)) {
        // the loop's read-back comes before integration: errors that the last
        // step's integration raised (a position leaving the z-index grid)
        // are read once more here, so that this call reports them
        const int rc = read_ctl(s);
        if (rc) return rc;
        return err_code(s->h_ctl->err);
    }
    HIPCHK(hipStreamSynchronize(s->st));
    s->prof.collect();
    return UPF_OK;
}

int upf_set_tuning(upf_sim* s, const upf_tuning* t) {
    if (!s || !t) return UPF_ERR_ARG;
    if (t->loop_block != 64 && t->loop_block != 128 && t->loop_block != 256) return UPF_ERR_ARG;
    if (t->grid_cap < 1 || (t->nbr_rounds != 2 && t->nbr_rounds != 4 && t->nbr_rounds != 8) ||
        (t->nbr_stage != 0 && t->nbr_stage != 32 && t->nbr_stage != 64) || t->speculate < 1 || t->speculate > 64 ||
        (t->loop_list < 0 || t->loop_list > 2) || (t->loop_list == 2 && t->loop_block != UNION_CHUNK) ||
        (t->loop_width != 4 && t->loop_width != 8) ||
        (t->nbr_tile != 0 && t->nbr_tile != 256 && t->nbr_tile != 512) || t->spec_first < 0 || t->spec_first > 64 ||
        (t->xcd_remap != 0 && t->xcd_remap != 1) || (t->nbr_skip != 0 && t->nbr_skip != 1) ||
        (t->nbr_fuse != 0 && t->nbr_fuse != 1))
        return UPF_ERR_ARG;
    s->tune = *t;
    return UPF_OK;
}

int upf_get_tuning(const upf_sim* s, upf_tuning* t) {
    if (!s || !t) return UPF_ERR_ARG;
    *t = s->tune;
    return UPF_OK;
}

int upf_profile_enable(upf_sim* s, int on) {
    if (!s) return UPF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(s->st));
    s->prof.collect();
    s->prof.on = on != 0;
    return UPF_OK;
}
int upf_profile_reset(upf_sim* s) {
    if (!s) return UPF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(s->st));
    s->prof.collect();
    for (int c = 0; c < PC_COUNT; ++c) { s->prof.ms[c] = 0.0; s->prof.cnt[c] = 0; }
    return UPF_OK;
}
int upf_profile_classes(void) { return PC_COUNT; }
const char* upf_profile_name(int cls) { return cls >= 0 && cls < PC_COUNT ? kProfNames[cls] : ""; }
int upf_profile_read(upf_sim* s, int cls, double* total_ms, long long* launches) {
    if (!s || cls < 0 || cls >= PC_COUNT) return UPF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(s->st));
    s->prof.collect();
    if (total_ms) *total_ms = s->prof.ms[cls];
    if (launches) *launches = s->prof.cnt[cls];
    return UPF_OK;
}

int upf_sort_and_bin(upf_sim* s) {
    if (!s) return UPF_ERR_ARG;
    if (!s->uploaded) return UPF_ERR_STATE;
    int rc = sort_and_bin(s);
    if (rc) return rc;
    // the state is now in key order: its keys are the sorted ones
    HIPCHK(hipMemcpyAsync(s->W.keys, s->W.keys_sorted, (size_t)s->n * 4, hipMemcpyDeviceToDevice, s->st));
    // loop arrays follow the new order too (identity gather for the next sort)
    Slots& S = s->S[s->cur];
    HIPCHK(hipMemcpyAsync(s->W.p, S.p_last, (size_t)s->n * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(S.p_last, S.p_prev, (size_t)s->n * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->W.rho, S.rho_last, (size_t)s->n * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return UPF_OK;
}

int upf_download_neighbors(upf_sim* s, int* counts, int32_t* flat, int cap) {
    if (!s || !counts) return -UPF_ERR_ARG;
    if (!s->uploaded || s->n_scenes != 1) return -UPF_ERR_STATE;
    int rc = upf_sort_and_bin(s);
    if (rc) return -rc;
    build_lists(s, s->args());
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->st) != hipSuccess) return -UPF_ERR_HIP;
    const int n = s->n, mx = s->P.max_neighbors;
    std::vector<int> cnt(n);
    std::vector<uint32_t> L((size_t)mx * s->cap);
    if (hipMemcpy(cnt.data(), s->W.ncount, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) return -UPF_ERR_HIP;
    if (hipMemcpy(L.data(), s->W.nbr, L.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -UPF_ERR_HIP;
    int w = 0;
    for (int i = 0; i < n; ++i) {
        counts[i] = cnt[i];
        for (int k = 0; k < cnt[i]; ++k, ++w)
            if (flat && w < cap) flat[w] = (int32_t)(L[(size_t)k * s->cap + i] >> NB_SLOT_SHIFT);
    }
    Ctl h;
    if (hipMemcpy(&h, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost) != hipSuccess) return -UPF_ERR_HIP;
    const int e = err_code(h.err);
    return e ? -e : w;
}

int upf_download_sorted(upf_sim* s, uint32_t* keys, int32_t* ids, uint32_t* cells, int* n_cells) {
    if (!s) return UPF_ERR_ARG;
    const int n = s->n;
    if (keys) HIPCHK(hipMemcpy(keys, s->W.keys_sorted, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (ids) HIPCHK(hipMemcpy(ids, s->S[s->cur].id, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (n_cells) {
        std::vector<uint2> t(s->ncells);
        HIPCHK(hipMemcpy(t.data(), s->W.cells, (size_t)s->ncells * sizeof(uint2), hipMemcpyDeviceToHost));
        std::vector<uint32_t> codes;
        for (uint32_t c = 0; c < s->ncells; ++c)
            if (t[c].y > t[c].x) codes.push_back(c);
        std::sort(codes.begin(), codes.end(), [&](uint32_t a, uint32_t b) { return t[a].x < t[b].x; });
        *n_cells = (int)codes.size();
        if (cells)
            for (size_t k = 0; k < codes.size(); ++k) {
                cells[3 * k] = codes[k];
                cells[3 * k + 1] = t[codes[k]].x;
                cells[3 * k + 2] = t[codes[k]].y;
            }
    }
    return UPF_OK;
}

int upf_download(upf_sim* s, float* pos4, float* vel4, float* rho, float* p, float* force4, float* psi, int* type) {
    if (!s) return UPF_ERR_ARG;
    if (!s->uploaded) return UPF_ERR_STATE;
    const int n = s->n;
    Slots& S = s->S[s->cur];
    HIPCHK(hipStreamSynchronize(s->st));
    std::vector<int> id(n);
    HIPCHK(hipMemcpy(id.data(), S.id, (size_t)n * 4, hipMemcpyDeviceToHost));
    std::vector<float> b4((size_t)4 * n), b1(n);
    auto get4 = [&](const float4* src, float* dst, bool is_pos) -> int {
        if (!dst) return UPF_OK;
        HIPCHK(hipMemcpy(b4.data(), src, (size_t)n * 16, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) {
            const int d = id[k];
            for (int a = 0; a < 3; ++a) dst[4 * d + a] = b4[4 * k + a];
            dst[4 * d + 3] = 0.0f;
            if (is_pos && type) {
                int t;
                std::memcpy(&t, &b4[4 * k + 3], 4);
                type[d] = t;
            }
        }
        return UPF_OK;
    };
    auto get1 = [&](const float* src, float* dst) -> int {
        if (!dst) return UPF_OK;
        HIPCHK(hipMemcpy(b1.data(), src, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) dst[id[k]] = b1[k];
        return UPF_OK;
    };
    // after a step the slot arrays of the loop (W.*) are in the current slot order
    std::vector<float> tmp4;
    if (pos4 || type) {
        if (!pos4) { tmp4.resize((size_t)4 * n); pos4 = tmp4.data(); }
        if (get4(S.pos, pos4, true)) return UPF_ERR_HIP;
    }
    if (get4(S.vel, vel4, false)) return UPF_ERR_HIP;
    if (get4(s->W.force, force4, false)) return UPF_ERR_HIP;
    if (get1(s->W.rho, rho)) return UPF_ERR_HIP;
    if (get1(sph_family(s->last_method) ? s->W.pres : s->W.p, p)) return UPF_ERR_HIP;
    if (get1(S.psi, psi)) return UPF_ERR_HIP;
    return UPF_OK;
}

int upf_upload_bodies(upf_sim* s, int n_bodies, const upf_body* bodies, const int* body_id, const float* rel4) {
    if (!s || n_bodies <= 0 || n_bodies >= 0xFFFF || !bodies || !body_id || !rel4) return UPF_ERR_ARG;
    if (!s->uploaded) return UPF_ERR_STATE;
    const int n = s->n;
    // per-particle body data in the current slot order (slots are still in
    // upload order unless a step ran, so scatter through the ids), checked on
    // the host before anything on the device changes
    std::vector<int> id(n), bid(n);
    std::vector<float> rel((size_t)4 * n), psi(n);
    Slots& S = s->S[s->cur];
    HIPCHK(hipMemcpy(id.data(), S.id, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(psi.data(), S.psi, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
        const int u = id[k];
        bid[k] = body_id[u];
        if (bid[k] < -1 || bid[k] >= n_bodies) return UPF_ERR_ARG;
        for (int a = 0; a < 4; ++a) rel[4 * k + a] = rel4[4 * u + a];
        // a rigid particle's weighted_volume_ = its mass / rest density (UnifiedParticle.cpp:24)
        if (bid[k] >= 0) psi[k] = rel4[4 * u + 3] / s->P.rest_density;
    }
    // the previous bodies are dropped first: after a failure below the
    // simulation has none (a rigid step returns UPF_ERR_STATE), never a body
    // count that disagrees with the device arrays
    s->rigid = false;
    s->n_bodies = 0;
    if (s->d_bodies) (void)hipFree(s->d_bodies);
    if (s->d_body_start) (void)hipFree(s->d_body_start);
    s->d_bodies = nullptr;
    s->d_body_start = nullptr;
    upf_body* d_b = nullptr;
    int* d_start = nullptr;
    bool ok = hipMalloc((void**)&d_b, (size_t)n_bodies * sizeof(upf_body)) == hipSuccess &&
              hipMalloc((void**)&d_start, (size_t)(n_bodies + 1) * sizeof(int)) == hipSuccess;
    if (ok && !s->d_body_keys)
        ok = s->alloc(&s->d_body_keys, s->cap) == hipSuccess && s->alloc(&s->d_body_keys2, s->cap) == hipSuccess &&
             s->alloc(&s->d_body_members, s->cap) == hipSuccess &&
             s->alloc(&s->d_body_terms, (size_t)7 * s->cap) == hipSuccess;
    ok = ok && hipMemcpy(d_b, bodies, (size_t)n_bodies * sizeof(upf_body), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(S.body, bid.data(), (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(S.rel, rel.data(), (size_t)n * 16, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(S.psi, psi.data(), (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        if (d_b) (void)hipFree(d_b);
        if (d_start) (void)hipFree(d_start);
        return UPF_ERR_HIP;
    }
    s->d_bodies = d_b;
    s->d_body_start = d_start;
    s->n_bodies = n_bodies;
    s->rigid = true;
    return UPF_OK;
}

int upf_upload_restart(upf_sim* s, int n, const float* pos4, const float* vel4, const int* type,
                       const float* corr_pressure, const float* predicted_density) {
    if (s && s->n_scenes != 1) return UPF_ERR_STATE;
    if (!corr_pressure || !predicted_density) return UPF_ERR_ARG;
    return upload_common(s, n, pos4, vel4, type, nullptr, nullptr, corr_pressure, predicted_density);
}

int upf_save_text(upf_sim* s, const char* path, int which) {
    if (!s || !path || which < UPF_TEXT_LIQUID || which > UPF_TEXT_RESTART) return UPF_ERR_ARG;
    if (!s->uploaded) return UPF_ERR_STATE;
    const int n = s->n;
    const Slots& S = s->S[s->cur];
    HIPCHK(hipStreamSynchronize(s->st));
    // slot order: the arrays as the last sort (and the step after it) left them
    std::vector<float> pos((size_t)4 * n), vel, p, rho;
    HIPCHK(hipMemcpy(pos.data(), S.pos, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (which == UPF_TEXT_RESTART) {
        vel.resize((size_t)4 * n);
        p.resize(n);
        rho.resize(n);
        HIPCHK(hipMemcpy(vel.data(), S.vel, (size_t)n * 16, hipMemcpyDeviceToHost));
        const float* src_p = sph_family(s->last_method) ? s->W.pres : s->W.p;
        HIPCHK(hipMemcpy(p.data(), src_p, (size_t)n * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rho.data(), s->W.rho, (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    std::vector<int> type(n);
    for (int i = 0; i < n; ++i) std::memcpy(&type[i], &pos[4 * (size_t)i + 3], 4);
    return write_text(path, which, n, pos.data(), type.data(), vel.data(), p.data(), rho.data());
}

int upf_dump_cells(upf_sim* s, const char* particles_path, const char* blocks_path) {
    if (!s || !particles_path || !blocks_path) return UPF_ERR_ARG;
    if (!s->uploaded || s->n_scenes != 1) return UPF_ERR_STATE;
    int rc = upf_sort_and_bin(s);
    if (rc) return rc;
    const int n = s->n;
    const Slots& S = s->S[s->cur];
    std::vector<float> rec((size_t)4 * n);
    std::vector<int> id(n);
    std::vector<uint2> cells(s->ncells);
    HIPCHK(hipMemcpy(rec.data(), S.pos, (size_t)n * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(id.data(), S.id, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cells.data(), s->W.cells, (size_t)s->ncells * sizeof(uint2), hipMemcpyDeviceToHost));
    // w: the uploaded particle's z-index as a float (the reference uploads
    // particles_h[i].index into w and never rewrites it)
    for (int k = 0; k < n; ++k) rec[4 * (size_t)k + 3] = (float)s->key0[id[k]];
    // occupied cells as (row-major block index, first slot, count), by block index
    const uint32_t dim = s->P.grid_dim;
    std::vector<uint32_t> blocks;
    for (uint32_t c = 0; c < s->ncells; ++c) {
        if (cells[c].y <= cells[c].x) continue;
        uint32_t b[3] = {0, 0, 0};
        for (int bit = 0; bit < 10; ++bit)
            for (int a = 0; a < 3; ++a) b[a] |= ((c >> (3 * bit + a)) & 1u) << bit;
        blocks.push_back((dim * dim) * b[2] + dim * b[1] + b[0]);
        blocks.push_back(cells[c].x);
        blocks.push_back(cells[c].y - cells[c].x);
    }
    return write_cell_dump(particles_path, blocks_path, n, rec.data(), blocks);
}

int upf_download_bodies(upf_sim* s, upf_body* out) {
    if (!s || !out) return UPF_ERR_ARG;
    if (!s->d_bodies) return UPF_ERR_STATE;
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(out, s->d_bodies, (size_t)s->n_bodies * sizeof(upf_body), hipMemcpyDeviceToHost));
    return UPF_OK;
}

int upf_get_step_stats(upf_sim* s, int max_steps, int* iterations, float* max_rho) {
    if (!s) return UPF_ERR_ARG;
    const int m = (int)s->k_hist.size();
    for (int i = 0; i < m && i < max_steps; ++i) {
        if (iterations) iterations[i] = s->k_hist[i];
        if (max_rho) max_rho[i] = s->rho_hist[i];
    }
    return m;
}

int upf_keys(const upf_params* p, int n, const float* pos4, uint32_t* keys) {
    if (!p || n <= 0 || !pos4 || !keys) return UPF_ERR_ARG;
    float4* d_pos = nullptr;
    uint32_t* d_keys = nullptr;
    Ctl* ctl = nullptr;
    int rc = UPF_ERR_HIP;
    do {
        if (hipMalloc((void**)&d_pos, (size_t)n * 16) != hipSuccess) break;
        if (hipMalloc((void**)&d_keys, (size_t)n * 4) != hipSuccess) break;
        if (hipMalloc((void**)&ctl, sizeof(Ctl)) != hipSuccess) break;
        if (hipMemset(ctl, 0, sizeof(Ctl)) != hipSuccess) break;
        if (hipMemcpy(d_pos, pos4, (size_t)n * 16, hipMemcpyHostToDevice) != hipSuccess) break;
        upf_sim tmp;
        tmp.P = *p;
        tmp.n = n;
        hipLaunchKernelGGL(k_keys, grid_for(n), dim3(BLOCK), 0, 0, tmp.args(), d_pos, d_keys, ctl);
        if (hipGetLastError() != hipSuccess) break;
        if (hipMemcpy(keys, d_keys, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
        Ctl h;
        if (hipMemcpy(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost) != hipSuccess) break;
        rc = err_code(h.err);
    } while (0);
    (void)hipFree(d_pos);
    (void)hipFree(d_keys);
    (void)hipFree(ctl);
    return rc;
}

int upf_radix_sort(int n, const uint32_t* keys_in, uint32_t* keys_out, int32_t* perm_out) {
    if (n <= 0 || !keys_in || !keys_out || !perm_out) return UPF_ERR_ARG;
    uint32_t *dk = nullptr, *dko = nullptr, *dv = nullptr;
    void* tmp = nullptr;
    int rc = UPF_ERR_HIP;
    do {
        if (hipMalloc((void**)&dk, (size_t)n * 4) != hipSuccess) break;
        if (hipMalloc((void**)&dko, (size_t)n * 4) != hipSuccess) break;
        if (hipMalloc((void**)&dv, (size_t)n * 4) != hipSuccess) break;
        if (hipMalloc(&tmp, radix_sort_temp_bytes(n)) != hipSuccess) break;
        if (hipMemcpy(dk, keys_in, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) break;
        if (radix_sort(dk, nullptr, dko, dv, n, 32, tmp, 0) != hipSuccess) break;
        if (hipMemcpy(keys_out, dko, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
        if (hipMemcpy(perm_out, dv, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
        rc = UPF_OK;
    } while (0);
    (void)hipFree(dk);
    (void)hipFree(dko);
    (void)hipFree(dv);
    (void)hipFree(tmp);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------ batched form
struct upf_batch {
    upf_sim* s = nullptr;
    std::vector<upf_params> sp;  // per-scene params (domains)
    std::vector<int> counts;
    std::vector<int> scene_of;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // one step, by the slot buffer it starts from
    // split form, by the slot buffer the step starts from: the head with the
    // first LOOP_CHUNK iterations, LOOP_CHUNK further iterations each, the tail
    hipGraphExec_t head[2] = {nullptr, nullptr};
    std::vector<hipGraphExec_t> chunk[2];
    hipGraphExec_t tail[2] = {nullptr, nullptr};
    int last_k = 0;                                 // split form: iterations of the previous step
    upf_tuning captured{};                          // launch shapes baked into the graphs
    upf_params captured_p{};                        // physics baked into the graphs (StepArgs, loop bounds)

    void drop_graphs() {
        for (auto& g : gexec)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        for (int c = 0; c < 2; ++c) {
            if (head[c]) { (void)hipGraphExecDestroy(head[c]); head[c] = nullptr; }
            if (tail[c]) { (void)hipGraphExecDestroy(tail[c]); tail[c] = nullptr; }
            for (auto& g : chunk[c])
                if (g) (void)hipGraphExecDestroy(g);
            chunk[c].clear();
        }
    }
};

constexpr int LOOP_CHUNK = 4;  // correction iterations per graph of the split form

// Captures one static-boundary PCISPH step starting from slot buffer `cur`
// (the whole correction loop with the decision on the device).
static int capture_step(upf_sim* s, int cur, hipGraphExec_t* out) {
    const bool prof = s->prof.on;
    s->prof.on = false;
    s->cur = cur;
    hipGraph_t g = nullptr;
    int rc = UPF_OK;
    if (hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = UPF_ERR_HIP;
    if (!rc) {
        rc = step_pcisph_static(s, true);
        const hipError_t e = hipStreamEndCapture(s->st, &g);
        if (!rc && e != hipSuccess) rc = UPF_ERR_HIP;
    }
    s->prof.on = prof;
    if (!rc && hipGraphInstantiate(out, g, nullptr, nullptr, 0) != hipSuccess) rc = UPF_ERR_HIP;
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

template <class F>
static int capture_graph(upf_sim* s, hipGraphExec_t* out, F enqueue) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return UPF_ERR_HIP;
    int rc = enqueue();
    const hipError_t e = hipStreamEndCapture(s->st, &g);
    if (!rc && e != hipSuccess) rc = UPF_ERR_HIP;
    if (!rc && hipGraphInstantiate(out, g, nullptr, nullptr, 0) != hipSuccess) rc = UPF_ERR_HIP;
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

// The split form's graphs for both starting slot buffers.
static int capture_split(upf_batch* b) {
    upf_sim* s = b->s;
    const bool prof = s->prof.on;
    s->prof.on = false;
    const int cur0 = s->cur;
    const int maxl = std::max(s->P.max_loops, 1);
    int rc = UPF_OK;
    for (int c = 0; c < 2 && !rc; ++c) {
        s->cur = c;
        rc = capture_graph(s, &b->head[c], [&] {
            const int r = step_static_head(s);  // sorts into buffer 1 - c
            if (r) return r;
            hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(256), 0, s->st, s->args(), s->ctl);
            step_static_iterations(s, 0, std::min(LOOP_CHUNK, maxl));
            return (int)UPF_OK;
        });
        for (int t0 = LOOP_CHUNK; !rc && t0 < maxl; t0 += LOOP_CHUNK) {
            hipGraphExec_t ge = nullptr;
            rc = capture_graph(s, &ge, [&] {
                step_static_iterations(s, t0, std::min(t0 + LOOP_CHUNK, maxl));
                return (int)UPF_OK;
            });
            if (ge) b->chunk[c].push_back(ge);
        }
        if (!rc) rc = capture_graph(s, &b->tail[c], [&] { return step_static_tail(s); });
    }
    s->cur = cur0;
    s->prof.on = prof;
    return rc;
}

static bool same_physics(const upf_params& a, const upf_params& b) {
    upf_params x = b;
    for (int k = 0; k < 3; ++k) {
        x.vbox_min[k] = a.vbox_min[k];
        x.vbox_max[k] = a.vbox_max[k];
        x.scales[k] = a.scales[k];
        x.collision_min[k] = a.collision_min[k];
        x.collision_max[k] = a.collision_max[k];
        x.container_min[k] = a.container_min[k];
        x.container_max[k] = a.container_max[k];
    }
    x.block_size = a.block_size;
    x.block_shift = a.block_shift;
    x.grid_dim = a.grid_dim;
    x.density_error_factor = a.density_error_factor;
    return std::memcmp(&x, &a, sizeof(upf_params)) == 0;
}

extern "C" {

int upf_batch_create(int n_scenes, const upf_params* scene_params, const int* counts, int max_record_steps,
                     upf_batch** out) {
    if (!out || n_scenes < 1 || n_scenes > 65535 || !scene_params || !counts || max_record_steps < 0)
        return UPF_ERR_ARG;
    long long total = 0;
    uint64_t ncells = 0;
    for (int k = 0; k < n_scenes; ++k) {
        if (counts[k] <= 0) return UPF_ERR_ARG;
        int rc = upf::check_params(&scene_params[k]);
        if (!rc && scene_params[k].add_wall_weight) rc = UPF_ERR_ARG;  // single-scene steps only
        if (rc) return rc;
        if (!same_physics(scene_params[0], scene_params[k])) return UPF_ERR_ARG;
        total += counts[k];
        const uint64_t g = scene_params[k].grid_dim;
        ncells += g * g * g;
    }
    if (total >= (1 << (32 - NB_SLOT_SHIFT)) || ncells >= (1ull << 32)) return UPF_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPF_ERR_NO_DEVICE;
    upf_batch* b = new upf_batch();
    b->sp.assign(scene_params, scene_params + n_scenes);
    b->counts.assign(counts, counts + n_scenes);
    b->s = new upf_sim();
    upf_sim* s = b->s;
    s->P = scene_params[0];
    int rc = sim_alloc(s, (int)total, (uint32_t)ncells, n_scenes);
    if (!rc) rc = stats_alloc(s, std::max(max_record_steps, 1));
    if (!rc) {
        std::vector<Dom> doms(n_scenes);
        uint32_t coff = 0;
        int soff = 0;
        for (int k = 0; k < n_scenes; ++k) {
            const upf_params& Q = scene_params[k];
            Dom& d = doms[k];
            for (int a = 0; a < 3; ++a) {
                d.scales[a] = Q.scales[a];
                d.cmin[a] = Q.collision_min[a];
                d.cmax[a] = Q.collision_max[a];
            }
            d.block_shift = Q.block_shift;
            d.grid_dim = Q.grid_dim;
            d.cell_off = coff;
            d.slot_off = soff;
            d.n = counts[k];
            d.delta = Q.density_error_factor;
            coff += Q.grid_dim * Q.grid_dim * Q.grid_dim;
            b->scene_of.insert(b->scene_of.end(), (size_t)counts[k], k);
            soff += counts[k];
        }
        // one scene: its domain travels in the kernel arguments
        if (n_scenes > 1 &&
            hipMemcpy(s->d_doms, doms.data(), doms.size() * sizeof(Dom), hipMemcpyHostToDevice) != hipSuccess)
            rc = UPF_ERR_HIP;
        // the scenes' 64-slot chunks (a chunk never straddles two scenes) for
        // the correction loop's per-iteration lists of running scenes
        if (!rc && n_scenes > 1) {
            std::vector<int> cb, ce, c0(n_scenes), nc(n_scenes);
            int off = 0;
            for (int k = 0; k < n_scenes; ++k) {
                c0[k] = (int)cb.size();
                for (int b0 = 0; b0 < counts[k]; b0 += 64) {
                    cb.push_back(off + b0);
                    ce.push_back(off + std::min(b0 + 64, counts[k]));
                }
                nc[k] = (int)cb.size() - c0[k];
                off += counts[k];
            }
            s->n_chunks = (int)cb.size();
            if (s->alloc(&s->d_chunk_begin, cb.size()) != hipSuccess || s->alloc(&s->d_chunk_end, ce.size()) != hipSuccess ||
                s->alloc(&s->d_sc_chunk0, (size_t)n_scenes) != hipSuccess ||
                s->alloc(&s->d_sc_nchunks, (size_t)n_scenes) != hipSuccess ||
                s->alloc(&s->d_act, 2 * cb.size()) != hipSuccess || s->alloc(&s->d_act_n, 2) != hipSuccess ||
                hipMemcpy(s->d_chunk_begin, cb.data(), cb.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s->d_chunk_end, ce.data(), ce.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s->d_sc_chunk0, c0.data(), c0.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s->d_sc_nchunks, nc.data(), nc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(s->d_act_n, 0, 2 * sizeof(int)) != hipSuccess)
                rc = UPF_ERR_HIP;
        }
    }
    if (rc) {
        upf_batch_destroy(b);
        return rc;
    }
    *out = b;
    return UPF_OK;
}

int upf_batch_destroy(upf_batch* b) {
    if (!b) return UPF_ERR_ARG;
    b->drop_graphs();
    if (b->s) upf_destroy(b->s);
    delete b;
    return UPF_OK;
}

int upf_batch_upload(upf_batch* b, const float* pos4, const float* vel4, const int* type) {
    if (!b) return UPF_ERR_ARG;
    const bool many = b->s->n_scenes > 1;
    return upload_common(b->s, (int)b->scene_of.size(), pos4, vel4, type, many ? b->scene_of.data() : nullptr,
                         many ? b->sp.data() : nullptr);
}

int upf_batch_step(upf_batch* b, int nsteps, int use_graph) {
    if (!b || nsteps < 0) return UPF_ERR_ARG;
    upf_sim* s = b->s;
    if (!s->uploaded) return UPF_ERR_STATE;
    int k = 0;
    // the first step also computes the boundary volumes: eager
    for (; k < nsteps && (!use_graph || !s->psi_ready); ++k) {
        int rc = step_pcisph_static(s, false);
        if (rc) return rc;
        ++s->steps_done;
        const int e = err_code(s->h_ctl->err);
        if (e) return e;
    }
    // the graphs bake in the launch shapes and the physics (every kernel's
    // StepArgs, the number of loop iterations): recapture when upf_set_tuning
    // or upf_set_params (through upf_batch_sim) changed either
    if ((b->gexec[0] || b->head[0]) && (std::memcmp(&b->captured, &s->tune, sizeof(upf_tuning)) != 0 ||
                                        std::memcmp(&b->captured_p, &s->P, sizeof(upf_params)) != 0))
        b->drop_graphs();
    if (use_graph == 2) {
        // split form: head graph, then chunk graphs while the device's loop
        // decision (read back) asks for more iterations, then the tail graph
        if (k < nsteps && !b->head[0]) {
            b->captured = s->tune;
            b->captured_p = s->P;
            const int rc = capture_split(b);
            if (rc) {
                b->drop_graphs();
                return rc;
            }
        }
        const int maxl = std::max(s->P.max_loops, 1);
        for (; k < nsteps; ++k) {
            const int c = s->cur;
            HIPCHK(hipGraphLaunch(b->head[c], s->st));
            int t = std::min(LOOP_CHUNK, maxl);
            size_t q = 0;
            // iterations the previous step needed go ahead of the first read-back,
            // at most spec_first past the head's, as in the host loop
            const int ahead = std::min(std::min(b->last_k, t + s->tune.spec_first), maxl);
            while (t < ahead && q < b->chunk[c].size()) {
                HIPCHK(hipGraphLaunch(b->chunk[c][q++], s->st));
                t = std::min(t + LOOP_CHUNK, maxl);
            }
            for (;;) {
                const int rc = read_ctl(s);
                if (rc) return rc;
                if (s->h_ctl->want_any <= t || q >= b->chunk[c].size() || (s->h_ctl->err & (1u << UPF_ERR_NAN)))
                    break;
                HIPCHK(hipGraphLaunch(b->chunk[c][q++], s->st));
                t = std::min(t + LOOP_CHUNK, maxl);
            }
            b->last_k = s->h_ctl->want_any;
            HIPCHK(hipGraphLaunch(b->tail[c], s->st));
            s->cur = 1 - c;
            ++s->steps_done;
            const int e = err_code(s->h_ctl->err);
            if (e) return e;
        }
    } else {
        if (k < nsteps && !b->gexec[0]) {
            b->captured = s->tune;
            b->captured_p = s->P;
            const int cur = s->cur;
            int rc = capture_step(s, 0, &b->gexec[0]);
            if (!rc) rc = capture_step(s, 1, &b->gexec[1]);
            s->cur = cur;
            if (rc) return rc;
        }
        for (; k < nsteps; ++k) {
            HIPCHK(hipGraphLaunch(b->gexec[s->cur], s->st));
            s->cur = 1 - s->cur;
            ++s->steps_done;
        }
    }
    int rc = read_ctl(s);
    if (rc) return rc;
    return err_code(s->h_ctl->err);
}

int upf_batch_download(upf_batch* b, float* pos4, float* vel4, float* rho) {
    if (!b) return UPF_ERR_ARG;
    return upf_download(b->s, pos4, vel4, rho, nullptr, nullptr, nullptr, nullptr);
}

// iterations / max_rho: [step][scene], up to max_steps steps; returns the
// number of steps recorded
int upf_batch_get_stats(upf_batch* b, int max_steps, int* iterations, float* max_rho) {
    if (!b || max_steps < 0) return -UPF_ERR_ARG;
    upf_sim* s = b->s;
    if (read_ctl(s)) return -UPF_ERR_HIP;
    const int m = std::min((int)s->h_ctl->step, s->stat_cap);
    const int c = std::min(m, max_steps);
    const size_t cnt = (size_t)c * s->n_scenes;
    if (iterations && cnt && hipMemcpy(iterations, s->d_kstat, cnt * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return -UPF_ERR_HIP;
    if (max_rho && cnt && hipMemcpy(max_rho, s->d_rstat, cnt * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return -UPF_ERR_HIP;
    return m;
}

int upf_batch_sim(upf_batch* b, upf_sim** sim) {
    if (!b || !sim) return UPF_ERR_ARG;
    *sim = b->s;
    return UPF_OK;
}

// Host-side mirror of the query-box enumeration, for tests: the cells (Morton
// codes of cells inside the grid) visited for a particle at (x, y, z), in
// visiting order.  Returns their number.
int upf_debug_box_cells(const upf_params* p, float x, float y, float z, uint32_t* out, int cap) {
    if (!p) return -1;
    BoxCells bc;
    bc.init(x, y, z, *p);
    int m = 0;
    for (int e = 0; e < bc.count(); ++e) {
        uint32_t c[3];
        if (!bc.cell(e, c)) continue;
        if (c[0] >= p->grid_dim || c[1] >= p->grid_dim || c[2] >= p->grid_dim) continue;
        if (m < cap) out[m] = dmorton3(c[0], c[1], c[2]);
        ++m;
    }
    return m;
}

}  // extern "C"

