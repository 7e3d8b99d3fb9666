// upf_device.h -- device-side helpers shared by the HIP translation units.
#ifndef UPF_DEVICE_H_
#define UPF_DEVICE_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "upf.h"

namespace upf {

// ------------------------------------------------------------- Morton codes
__host__ __device__ inline uint32_t dspread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint32_t dmorton3(uint32_t x, uint32_t y, uint32_t z) {
    return dspread10(x) | (dspread10(y) << 1) | (dspread10(z) << 2);
}

// unsigned short conversion of a scaled coordinate (the reference builds
// its grid coordinate with vmmlib's float -> unsigned short converting
// constructor, src/vmmlib/vector3.h:300-304); `ok` drops to false outside
// the 1024-entry z_table.
__host__ __device__ inline uint32_t dgrid(float v, bool& ok) {
    if (!(v >= 0.0f) || v >= 1024.0f) { ok = false; return v >= 1024.0f ? 1023u : 0u; }
    return (uint32_t)(unsigned short)v;
}

// CalcIndex(scales * position) (src/cpu/zIndex.h:40-45, UnifiedPhysics.cpp:3967)
__host__ __device__ inline uint32_t dkey(float x, float y, float z, const float* scales, bool& ok) {
    ok = true;
    const uint32_t cx = dgrid(scales[0] * x, ok);
    const uint32_t cy = dgrid(scales[1] * y, ok);
    const uint32_t cz = dgrid(scales[2] * z, ok);
    return dmorton3(cx, cy, cz);
}

__host__ __device__ inline int dmsb(uint32_t v) {  // index of the highest set bit, v != 0
#ifdef __HIP_DEVICE_COMPILE__
    return 31 - __clz(v);
#else
    return 31 - __builtin_clz(v);
#endif
}

// Cells of the query box of GetNeighbors (src/cpu/UnifiedPhysics.cpp:2246-2256,
// ZIndex::BoxQuery src/cpu/zIndex.cpp:43-74), visited in ascending Morton order
// of the cell so that the particles of the cells come out in ascending sorted
// order -- the order of the reference's neighbour lists.
//
// Per axis the box covers n = 1..3 consecutive cell coordinates.  Two cells
// compare in Morton order by the highest interleaved bit in which they differ;
// among n consecutive coordinates that bit can only be one of <= 2 "split"
// levels per axis.  Ranking the <= 6 splits by (3 * level + axis) and counting
// through them as a binary number enumerates the cells in Morton order.
// Every array index below is a compile-time constant after unrolling, so the
// state lives in registers.
struct BoxCells {
    uint32_t lo[3];
    int n[3];
    int hcase[3];   // n == 3: 0 -> the high split separates {v0}|{v1,v2}; 1 -> {v0,v1}|{v2}
    int shA[3];     // shift of the axis' first split bit in the enumeration index
    int shB[3];     // shift of its second split bit (n == 3 only)
    int nd;         // number of splits

    __host__ __device__ void init(float x, float y, float z, const upf_params& P) {
        init(x, y, z, P.support_radius, P.scales, P.block_shift);
    }

    __host__ __device__ void init(float x, float y, float z, float h, const float* scales, uint32_t block_shift) {
        const float c[3] = {x, y, z};
        uint32_t l3[3];
        int n3[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float lo_f = scales[a] * (c[a] - h);
            float hi_f = scales[a] * (c[a] + h);
            lo_f = lo_f >= 0.0f ? lo_f : 0.0f;       // MAX(min_query_box, 0)
            hi_f = hi_f <= 1024.0f ? hi_f : 1024.0f; // MIN(max_query_box, GRID_RESOLUTION)
            const uint32_t l = ((uint32_t)(unsigned short)lo_f) >> block_shift;
            const uint32_t u = ((uint32_t)(unsigned short)hi_f) >> block_shift;
            l3[a] = l;
            n3[a] = (int)(u - l) + 1;
        }
        init_cells(l3, n3);
    }

    // the cells lo[a] .. lo[a] + n[a] - 1 per axis (n[a] in 1..3)
    __host__ __device__ void init_cells(const uint32_t l3[3], const int n3[3]) {
        int prA[3], prB[3];
        bool vA[3], vB[3];
        nd = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t l = l3[a];
            int cnt = n3[a];
            cnt = cnt > 3 ? 3 : (cnt < 1 ? 1 : cnt);
            lo[a] = l;
            n[a] = cnt;
            const int m01 = dmsb(l ^ (l + 1)), m12 = dmsb((l + 1) ^ (l + 2));
            hcase[a] = m01 > m12 ? 0 : 1;
            vA[a] = cnt >= 2;
            vB[a] = cnt == 3;
            prA[a] = 3 * (cnt == 2 ? m01 : (m01 > m12 ? m01 : m12)) + a;
            prB[a] = 3 * (m01 > m12 ? m12 : m01) + a;
            nd += (int)vA[a] + (int)vB[a];
        }
        // rank of a split = number of valid splits with a higher priority
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int rA = 0, rB = 0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                rA += (vA[b] && prA[b] > prA[a]) + (vB[b] && prB[b] > prA[a]);
                rB += (vA[b] && prA[b] > prB[a]) + (vB[b] && prB[b] > prB[a]);
            }
            shA[a] = nd - 1 - rA;
            shB[a] = nd - 1 - rB;
        }
    }

    __host__ __device__ int count() const { return 1 << nd; }

    // cell of enumeration index e (0 .. count()-1); false for indices that
    // do not name a cell (duplicates of the n == 3 singleton group)
    __host__ __device__ bool cell(int e, uint32_t out[3]) const {
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int off = 0;
            if (n[a] == 2) {
                off = (e >> shA[a]) & 1;
            } else if (n[a] == 3) {
                const int hb = (e >> shA[a]) & 1;
                const int lb = (e >> shB[a]) & 1;
                if (hcase[a] == 0) {
                    ok = ok && !(hb == 0 && lb);
                    off = hb ? 1 + lb : 0;
                } else {
                    ok = ok && !(hb == 1 && lb);
                    off = hb ? 2 : lb;
                }
            }
            out[a] = lo[a] + (uint32_t)off;
        }
        return ok;
    }
};

// MAX / MIN of src/sceneConfigure/global.h:4-5 (keep their NaN behaviour)
__host__ __device__ inline float fmax_ref(float a, float b) { return a >= b ? a : b; }
__host__ __device__ inline float fmin_ref(float a, float b) { return a <= b ? a : b; }

}  // namespace upf

#endif
