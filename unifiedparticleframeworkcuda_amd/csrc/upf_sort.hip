// upf_sort.hip -- stable LSD radix sort of (32-bit key, 32-bit value) pairs
// for gfx950 wave64.  Replaces the reference's RadixSort
// (src/gpu/particles.cu:31-60, kernels RadixSum / RadixPrefixSum /
// RadixAddOffsetsAndShuffle in src/gpu/radixsort_kernel.cu:64,216,339) with a
// design of its own:
//   * 8-bit digits, one pass per byte of the key range that is in use;
//   * tile = 256 threads x IPT items (4, 8 or 16: the largest that still
//     gives at least 512 tiles, so that long sorts write longer runs per
//     digit and every CU still gets tiles); each wave ranks 64 consecutive items per
//     round with 8 ballots (lanes with equal digits form a peer mask), so the
//     rank of an item is the number of equal digits before it -> stable;
//   * per-tile digit counts (digit-major) are scanned by one workgroup per
//     digit; the scatter orders its tile by digit in LDS first, so that the
//     items of one digit leave as one contiguous (coalesced) run at digit
//     base + tile offset.
#include <hip/hip_runtime.h>

#include "upf_sort.h"

namespace upf {

constexpr int RS_THREADS = 256;
constexpr int RS_MIN_IPT = 4;  // the smallest tile (most tiles) of any sort
constexpr int RS_WAVES = RS_THREADS / 64;

template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void rs_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                      uint32_t* __restrict__ hist, int ntiles) {
    constexpr int RS_TILE = RS_THREADS * RS_IPT;
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * RS_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// one workgroup per digit: exclusive scan of the digit's row of tile counts,
// total of the row into totals[digit]
__global__ __launch_bounds__(RS_THREADS) void rs_scan(uint32_t* __restrict__ hist, int ntiles,
                                                      uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[RS_THREADS];
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    const int per = (ntiles + RS_THREADS - 1) / RS_THREADS;
    const int beg = threadIdx.x * per;
    const int end = min(beg + per, ntiles);
    uint32_t s = 0;
    for (int i = beg; i < end; ++i) s += row[i];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 partial sums
    for (int off = 1; off < RS_THREADS; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int i = beg; i < end; ++i) {
        const uint32_t c = row[i];
        row[i] = run;
        run += c;
    }
    if (threadIdx.x == RS_THREADS - 1) totals[blockIdx.x] = part[RS_THREADS - 1];
}

template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void rs_scatter(const uint32_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin,
                                                         uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                         int n, int shift, const uint32_t* __restrict__ hist,
                                                         int ntiles, const uint32_t* __restrict__ totals) {
    constexpr int RS_TILE = RS_THREADS * RS_IPT;
    __shared__ uint32_t digit_base[256];  // global start of each digit, then of this tile's run of it
    __shared__ uint32_t local_start[256]; // start of each digit within the tile
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t skey[RS_TILE], sval[RS_TILE];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    // exclusive scan of the digit totals
    digit_base[t] = totals[t];
#pragma unroll
    for (int k = 0; k < RS_WAVES; ++k) wcnt[k][t] = 0;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = t >= off ? digit_base[t - off] : 0u;
        __syncthreads();
        digit_base[t] += v;
        __syncthreads();
    }
    const uint32_t my_incl = digit_base[t];
    __syncthreads();
    digit_base[t] = my_incl - totals[t];
    __syncthreads();

    uint32_t key[RS_IPT], val[RS_IPT], rank[RS_IPT];
    const int tile0 = blockIdx.x * RS_TILE;
    const int base = tile0 + w * 64 * RS_IPT;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        const bool valid = idx < n;
        key[r] = valid ? kin[idx] : 0u;
        val[r] = valid ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        const uint32_t cur = wcnt[w][d];
        rank[r] = cur + before;
        const bool leader = valid && ((peers & lt) == 0ull);
        // all lanes of the wave have read wcnt before the leaders write it
        __builtin_amdgcn_wave_barrier();
        if (leader) wcnt[w][d] = cur + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, and the tile's count
    uint32_t tile_cnt = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; ++k) {
        const uint32_t c = wcnt[k][t];
        wcnt[k][t] = tile_cnt;
        tile_cnt += c;
    }
    // exclusive scan of the tile's digit counts: where each digit starts in the tile
    local_start[t] = tile_cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = t >= off ? local_start[t - off] : 0u;
        __syncthreads();
        local_start[t] += v;
        __syncthreads();
    }
    const uint32_t ls = local_start[t] - tile_cnt;
    __syncthreads();
    local_start[t] = ls;
    // destination of the tile's run of digit t, minus the run's local start
    digit_base[t] = digit_base[t] + hist[t * ntiles + blockIdx.x] - ls;
    __syncthreads();
    // the tile in digit order (stable: wave, round, lane order within a digit)
#pragma unroll
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & 255u;
            const uint32_t lp = local_start[d] + wcnt[w][d] + rank[r];
            skey[lp] = key[r];
            sval[lp] = val[r];
        }
    }
    __syncthreads();
    const int cnt = min(RS_TILE, n - tile0);
    for (int q = t; q < cnt; q += RS_THREADS) {
        const uint32_t k = skey[q];
        const uint32_t dst = digit_base[(k >> shift) & 255u] + (uint32_t)q;
        kout[dst] = k;
        vout[dst] = sval[q];
    }
}

// items per thread of a sort of n keys: the largest of 16, 8, 4 that leaves
// at least 512 tiles
static int sort_ipt(int n) {
    if (n >= 512 * RS_THREADS * 16) return 16;
    if (n >= 512 * RS_THREADS * 8) return 8;
    return RS_MIN_IPT;
}

size_t radix_sort_temp_bytes(int n) {
    const int tile = RS_THREADS * sort_ipt(n);
    const int ntiles = (n + tile - 1) / tile;
    return ((size_t)256 * ntiles + 256) * sizeof(uint32_t) + 2 * (size_t)n * sizeof(uint32_t);
}

// Sorts keys[0..n) with values; vals == nullptr means values = index.  The
// result lands in keys_out / vals_out.  `bits` = number of key bits in use.
hipError_t radix_sort(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out, int n,
                      int bits, void* temp, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int ipt = sort_ipt(n);
    const int ntiles = (n + RS_THREADS * ipt - 1) / (RS_THREADS * ipt);
    uint32_t* hist = (uint32_t*)temp;
    uint32_t* totals = hist + (size_t)256 * ntiles;
    uint32_t* kbuf = totals + 256;
    uint32_t* vbuf = kbuf + n;
    const int passes = (bits + 7) / 8;
    // ping-pong so that the last pass writes keys_out / vals_out
    const uint32_t* kin = keys;
    const uint32_t* vin = vals;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        // passes alternate between (kbuf, vbuf) and (keys_out, vals_out)
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* ko = to_out ? keys_out : kbuf;
        uint32_t* vo = to_out ? vals_out : vbuf;
        (void)last;
        auto pass = [&](auto hist_kernel, auto scatter_kernel) {
            hipLaunchKernelGGL(hist_kernel, dim3(ntiles), dim3(RS_THREADS), 0, st, kin, n, 8 * p, hist, ntiles);
            hipLaunchKernelGGL(rs_scan, dim3(256), dim3(RS_THREADS), 0, st, hist, ntiles, totals);
            hipLaunchKernelGGL(scatter_kernel, dim3(ntiles), dim3(RS_THREADS), 0, st, kin, vin, ko, vo, n, 8 * p, hist,
                               ntiles, totals);
        };
        if (ipt == 16) pass(rs_hist<16>, rs_scatter<16>);
        else if (ipt == 8) pass(rs_hist<8>, rs_scatter<8>);
        else pass(rs_hist<4>, rs_scatter<4>);
        kin = ko;
        vin = vo;
    }
    return hipGetLastError();
}

}  // namespace upf
