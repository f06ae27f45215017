// Wavefront helpers shared by the kernels that work one bin cell per
// wavefront: the neighbor-row build and the LJ tile force kernel.
#pragma once

#include "ctx.hpp"

namespace mdd {

// Blocks are dispatched round-robin over the 8 XCDs; map them so that each
// XCD works on one contiguous range of cells, whose neighbourhoods share
// that XCD's L2.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + i;
}

// lanes above / below lane `rel` of a 64-lane chunk
__device__ inline uint64_t lanes_above(int rel) {
    if (rel < 0) return ~0ull;
    if (rel >= 63) return 0ull;
    return ~0ull << (rel + 1);
}
__device__ inline uint64_t lanes_below(int rel) {
    if (rel <= 0) return 0ull;
    if (rel >= 64) return ~0ull;
    return (1ull << rel) - 1ull;
}

__device__ inline double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(b), l));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(b >> 32), l));
    return __longlong_as_double(static_cast<long long>((uint64_t(hi) << 32) | lo));
}

// Runs of equal keys inside a wavefront: the lane of the run head (h) and
// of the run tail (t). Invalid lanes form runs of their own.
__device__ inline void wave_runs(uint32_t key, bool valid, int& h, int& t) {
    const int lane = threadIdx.x & 63;
    const uint32_t kp = __shfl_up(key, 1, 64);
    const int vp = __shfl_up(int(valid), 1, 64);
    const uint32_t kn = __shfl_down(key, 1, 64);
    const int vn = __shfl_down(int(valid), 1, 64);
    const bool head = lane == 0 || !valid || !vp || kp != key;
    const bool tail = lane == 63 || !valid || !vn || kn != key;
    int hv = head ? lane : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int x = __shfl_up(hv, d, 64);
        if (lane >= d && x > hv) hv = x;
    }
    int tv = tail ? lane : 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int x = __shfl_down(tv, d, 64);
        if (lane + d < 64 && x < tv) tv = x;
    }
    h = hv;
    t = tv;
}

__device__ inline uint32_t lanes_before(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// The candidate list of a bin cell: the cell's stencil cells form runs of
// consecutive cells, i.e. contiguous bin ranges, concatenated in bin order.
// Lane r < nruns holds run r: its flat offset excl and rbase = bin start -
// excl, so bin index = flat index + rbase (mod 2^32).
struct CellRuns {
    uint32_t excl, rbase;  // this lane's run
    uint32_t total;        // candidates in the list
    uint32_t base0;        // rbase of the run holding the cell itself
    int nruns;
};

// rf / rl: the run table of this cell's system; cells outside
// [cmin, cmax] are clamped away
__device__ inline CellRuns cell_runs(int64_t C, int64_t cmin, int64_t cmax,
                                     const int* __restrict__ rf, const int* __restrict__ rl,
                                     int nruns, int run0, const uint32_t* __restrict__ bin_start) {
    const int lane = int(threadIdx.x & 63);
    uint32_t rcnt = 0, rb = 0;
    if (lane < nruns) {
        int64_t lo = C + rf[lane], hi = C + rl[lane];
        if (lo < cmin) lo = cmin;
        if (hi > cmax) hi = cmax;
        if (lo <= hi) {
            rb = bin_start[lo];
            rcnt = bin_start[hi + 1] - rb;
        }
    }
    uint32_t incl = rcnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    CellRuns R;
    R.excl = incl - rcnt;
    R.rbase = rb - R.excl;
    R.total = __builtin_amdgcn_readlane(incl, nruns - 1);
    R.base0 = __builtin_amdgcn_readlane(R.rbase, run0);
    R.nruns = nruns;
    return R;
}

// bin index of flat candidate f of the list (f per lane)
__device__ inline uint32_t flat_to_bin(const CellRuns& R, uint32_t f) {
    uint32_t base = 0;
    for (int r = 0; r < R.nruns; r++) {
        const uint32_t pr = __builtin_amdgcn_readlane(R.excl, r);
        const uint32_t br = __builtin_amdgcn_readlane(R.rbase, r);
        if (f >= pr) base = br;
    }
    return f + base;
}

// The tile force kernel works a cell with kTileWaves wavefronts; wave w
// takes the cell's particles [w mw, (w + 1) mw), mw = ceil(M / kTileWaves),
// and G lanes share a particle, G = 64 / mw capped at 16 (1 for mw >= 64);
// a wave with more than 64 / G particles takes them in passes.
constexpr uint32_t kTileWaves = 1;

__device__ __host__ inline int tile_lanes(uint32_t M) {
    const uint32_t mw = (M + kTileWaves - 1u) / kTileWaves;
    if (mw >= 64u) return 1;
    const uint32_t g = 64u / (mw ? mw : 1u);
    return g > 16u ? 16 : int(g);
}

// Tile-row layout of a cell holding particles [s0, s0 + M): entry k of the
// cell's particle m sits at s0 * cap + (k / G) * M * G + m * G + k % G, so
// the G lanes of a particle read G contiguous entries and the M particles
// of a cell are contiguous. Rows of up to maxnb entries stay inside the
// cell's region [s0 cap, (s0 + M) cap) when cap >= maxnb + 15 (G <= 16).
__device__ __host__ inline uint32_t tile_addr(uint32_t s0, uint32_t cap, uint32_t M, uint32_t G,
                                              uint32_t m, uint32_t k) {
    return s0 * cap + (k / G) * M * G + m * G + k % G;
}

}  // namespace mdd
