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
// levels per axis.  Sorting the <= 6 splits by (3 * level + axis) and counting
// through them as a binary number enumerates the cells in Morton order.
struct BoxCells {
    uint32_t lo[3];
    int n[3];
    int nd;                 // number of splits
    int dax[6], dkind[6];   // axis, kind (0: n==2 split, 1: high split, 2: low split)
    int hcase[3];           // for n==3: 0 -> high split separates {v0}|{v1,v2}; 1 -> {v0,v1}|{v2}
    int pos_of[3][3];       // bit position (0 = most significant) of an axis' split kinds

    __host__ __device__ void init(float x, float y, float z, const upf_params& P) {
        const float h = P.support_radius;
        const float c[3] = {x, y, z};
        int prio[6];
        nd = 0;
        for (int a = 0; a < 3; ++a) {
            float lo_f = P.scales[a] * (c[a] - h);
            float hi_f = P.scales[a] * (c[a] + h);
            lo_f = lo_f >= 0.0f ? lo_f : 0.0f;       // MAX(min_query_box, 0)
            hi_f = hi_f <= 1024.0f ? hi_f : 1024.0f; // MIN(max_query_box, GRID_RESOLUTION)
            const uint32_t l = ((uint32_t)(unsigned short)lo_f) >> P.block_shift;
            const uint32_t u = ((uint32_t)(unsigned short)hi_f) >> P.block_shift;
            lo[a] = l;
            n[a] = (int)(u - l) + 1;
            hcase[a] = 0;
            if (n[a] == 2) {
                dax[nd] = a; dkind[nd] = 0; prio[nd] = 3 * dmsb(l ^ (l + 1)) + a; ++nd;
            } else if (n[a] >= 3) {
                n[a] = 3;
                const int m01 = dmsb(l ^ (l + 1)), m12 = dmsb((l + 1) ^ (l + 2));
                hcase[a] = m01 > m12 ? 0 : 1;
                const int hi = m01 > m12 ? m01 : m12, low = m01 > m12 ? m12 : m01;
                dax[nd] = a; dkind[nd] = 1; prio[nd] = 3 * hi + a; ++nd;
                dax[nd] = a; dkind[nd] = 2; prio[nd] = 3 * low + a; ++nd;
            }
        }
        // sort splits by priority, descending (insertion sort, <= 6 entries)
        for (int i = 1; i < nd; ++i) {
            for (int j = i; j > 0 && prio[j] > prio[j - 1]; --j) {
                int t = prio[j]; prio[j] = prio[j - 1]; prio[j - 1] = t;
                t = dax[j]; dax[j] = dax[j - 1]; dax[j - 1] = t;
                t = dkind[j]; dkind[j] = dkind[j - 1]; dkind[j - 1] = t;
            }
        }
        for (int a = 0; a < 3; ++a) pos_of[a][0] = pos_of[a][1] = pos_of[a][2] = -1;
        for (int k = 0; k < nd; ++k) pos_of[dax[k]][dkind[k]] = k;
    }

    __host__ __device__ int count() const { return 1 << nd; }

    // cell of enumeration index e (0 .. count()-1); false for indices that
    // do not name a cell (duplicates of the n == 3 singleton group)
    __host__ __device__ bool cell(int e, uint32_t out[3]) const {
        for (int a = 0; a < 3; ++a) {
            int off = 0;
            if (n[a] == 2) {
                off = (e >> (nd - 1 - pos_of[a][0])) & 1;
            } else if (n[a] == 3) {
                const int hb = (e >> (nd - 1 - pos_of[a][1])) & 1;
                const int lb = (e >> (nd - 1 - pos_of[a][2])) & 1;
                if (hcase[a] == 0) {
                    if (hb == 0) { if (lb) return false; off = 0; }
                    else off = 1 + lb;
                } else {
                    if (hb == 1) { if (lb) return false; off = 2; }
                    else off = lb;
                }
            }
            out[a] = lo[a] + (uint32_t)off;
        }
        return true;
    }
};

// MAX / MIN of src/sceneConfigure/global.h:4-5 (keep their NaN behaviour)
__host__ __device__ inline float fmax_ref(float a, float b) { return a >= b ? a : b; }
__host__ __device__ inline float fmin_ref(float a, float b) { return a <= b ? a : b; }

}  // namespace upf

#endif
