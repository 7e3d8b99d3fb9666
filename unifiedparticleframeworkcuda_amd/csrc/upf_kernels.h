// upf_kernels.h -- device state layout and kernel declarations of the step.
#ifndef UPF_KERNELS_H_
#define UPF_KERNELS_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "upf.h"

namespace upf {

constexpr int LUT_N = 301;
constexpr int BLOCK = 256;
constexpr int NB_ESC = 27;     // escapes per delta-coded list at most (one per gap between stencil cells)
// neighbour unions of the correction loop: per chunk of UNION_CHUNK
// consecutive slots, the distinct neighbour slots of its lists (at most
// UNION_CAP, else the chunk reads through the delta-coded lists), found
// through an LDS hash table of UNION_TAB entries
constexpr int UNION_CHUNK = 256;
constexpr int UNION_CAP = 1024;
constexpr int UNION_TAB = 2048;

// Loop control of the PCISPH correction loop, kept on the device.
constexpr int MAX_SLOTS = 64;  // the max reduction spreads its atomics over 64 cache lines

struct Ctl {              // whole simulation
    int want_any;        // largest number of iterations any scene's loop must run
    uint32_t err;        // bit (1 << UPF_ERR_*) set on failures
    uint32_t step;       // steps recorded in the per-scene statistics
    uint32_t pad;
};

struct SceneCtl {         // one scene's PCISPH loop
    int want;            // number of iterations the loop must run (grows as the loop decides)
    int too_large;       // density_error_too_large_ latch (src/cpu/UnifiedPhysics.cpp:5373,5397-5398)
    float last_max;      // max predicted density of the last completed iteration
    int pad;
};

// The domain of one scene: its z-index grid, collision box, cell table, slot
// range and PCISPH density error factor.  A plain simulation is one scene.
struct Dom {
    float scales[3];
    float cmin[3], cmax[3];
    uint32_t block_shift, grid_dim;
    uint32_t cell_off;   // first entry of the scene's cell table
    int slot_off, n;     // the scene's slot range
    float delta;         // density error factor (computed on the scene's own prototype block)
};

// Per-particle state of one sorted slot range (scene).  Arrays indexed by slot.
struct Slots {
    float4* pos;        // xyz, w = particle type (int bits)
    float4* vel;
    int* id;            // upload index
    float* p_last;      // correction pressure at the end of the previous step's loop
    float* p_prev;      // the one before (previous_correction_pressure_ of the CPU path)
    float* rho_last;    // predicted density of the previous step's last iteration
    float* psi;         // boundary volume (weighted_volume_)
    float4* x0;         // initial position (predicted position of rigid particles in the CPU path)
    int* body;          // rigid body index or -1
    float4* rel;        // initial offset from the body's centre of mass
};

struct Work {           // per-step arrays, slot order
    uint32_t* keys;     // keys of the slot order before sorting (written by integration)
    uint32_t* keys_sorted;
    uint32_t* perm;     // sorted slot -> previous slot
    float4* force;      // external force, after integration force + pressure force
    float4* fp;         // correction pressure force
    float4* xs;         // predicted position
    float* p;           // correction pressure of the running loop
    float4* xpd;        // (current position, p / (rho0 * rho0) -- or -(rho0 psi) for an Akinci
                        // boundary particle): one gather per pressure-force pair; WCSPH:
                        // (position, m / rho), w < 0 for a non-fluid particle
    float4* vpg;        // WCSPH: (velocity, p / rho^2) of a particle, the force pass's second record
    uint32_t* max_slots; // MAX_SLOTS x 32 uints: per-slot max of the predicted density (bits)
    float* rho;         // predicted density (PCISPH) / density (WCSPH)
    float* pres;        // WCSPH pressure
    uint32_t* nbr;      // neighbour lists, column-major: nbr[k * stride + slot]
    int* ncount;
    // the same lists for the correction loop, delta-coded: first neighbour
    // slot in nbf, then 16-bit slot increments, eight per 16-byte word of
    // nbw (row r = entries 8r..8r+7, column-major); increment 0 = escape:
    // the increment is the particle's next entry of nbe
    uint32_t* nbf;
    uint4* nbw;
    uint32_t* nbe;
    // the same lists as indices into their chunk's union (16-bit, eight per
    // 16-byte word of nbu, column-major like nbw); uslots[c * UNION_CAP + u]
    // is the slot of union entry u of chunk c, un[c] the union's size (-1:
    // the chunk uses the delta-coded lists)
    uint4* nbu;
    uint32_t* uslots;
    int* un;
    uint2* cells;       // dense cell table indexed by cell Morton code: [first, end)
};

}  // namespace upf

#endif
