// Device-wide exclusive scan and stable counting sort (cell binning).
//
// The counting sort reproduces the order of the reference's binning
// (crates/mddem_core/src/neighbor.rs:857-889): within a cell, elements stay
// in index order. Elements are first placed with atomics (aggregated over
// runs of equal keys inside a wavefront, since the input is mostly sorted
// already), then every element computes its index rank inside its cell in
// parallel and is written to its final slot, so the result is deterministic.
#include "cells.hpp"

namespace mdd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, double v, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f32(float* __restrict__ p, float v, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// inclusive scan across one 64-lane wavefront
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive scan across a 256-thread block; returns the block total in *total
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* lds4, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; k++) off += lds4[k];
    uint32_t tot = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* __restrict__ in,
                                                              int64_t n, uint32_t* sums) {
    __shared__ uint32_t lds4[4];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_excl_scan(s, lds4, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Second pass: each block first sums the block totals before it (a few
// thousand at most, read from L2), so no separate scan of the totals runs.
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out,
                                                             int64_t n,
                                                             const uint32_t* __restrict__ sums,
                                                             int64_t nb) {
    __shared__ uint32_t lds4[4];
    uint32_t before = 0;
    for (int64_t j = threadIdx.x; j < int64_t(blockIdx.x); j += kScanThreads) before += sums[j];
    uint32_t off;
    block_excl_scan(before, lds4, &off);
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, lds4, &tot) + off;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = off + tot;
}

// Exclusive scan of in[0..n) by one workgroup, for arrays up to
// kOneBlockScan elements: out[0..n] (out[n] = total). zero_in: in[i] is
// cleared once read (a counting sort's count array becomes its cursor).
// add/out2: out2[i] = out[i] + add[i] as well (bin starts: particles plus
// images of each cell).
constexpr int kOneThreads = 1024;
constexpr int kOneItems = 8;
constexpr int kOneTile = kOneThreads * kOneItems;
constexpr int64_t kOneBlockScan = int64_t(1) << 17;

// Tiles of kOneTile elements pass through LDS: read and written coalesced,
// scanned as kOneItems contiguous elements per thread. The next tile's
// loads are issued before the current tile is scanned, so one global
// latency per tile overlaps the LDS work instead of adding to it.
__global__ void __launch_bounds__(kOneThreads) k_scan_one(uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, int64_t n,
                                                          int zero_in,
                                                          const uint32_t* __restrict__ add,
                                                          uint32_t* __restrict__ out2) {
    __shared__ uint4 tile4[kOneTile / 4];
    __shared__ uint32_t wsum[kOneThreads / 64];
    uint32_t* tile = reinterpret_cast<uint32_t*>(tile4);
    const int tid = int(threadIdx.x), lane = tid & 63, w = tid >> 6;
    uint32_t v[kOneItems], ad[kOneItems];
    auto fetch = [&](int64_t base) {
#pragma unroll
        for (int k = 0; k < kOneItems; k++) {
            const int64_t i = base + int64_t(k) * kOneThreads + tid;
            v[k] = i < n ? in[i] : 0u;
            ad[k] = out2 && i < n ? add[i] : 0u;
        }
    };
    fetch(0);
    uint32_t carry = 0;
    for (int64_t base = 0; base < n; base += kOneTile) {
        uint32_t adc[kOneItems];
#pragma unroll
        for (int k = 0; k < kOneItems; k++) {
            const int64_t i = base + int64_t(k) * kOneThreads + tid;
            tile[k * kOneThreads + tid] = v[k];
            if (zero_in && i < n) in[i] = 0u;
            adc[k] = ad[k];
        }
        if (base + kOneTile < n) fetch(base + kOneTile);
        __syncthreads();
        const uint4 a = tile4[2 * tid], b = tile4[2 * tid + 1];
        const uint32_t s = (a.x + a.y + a.z + a.w) + (b.x + b.y + b.z + b.w);
        const uint32_t inc = wave_incl_scan(s);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t off = carry, tot = 0;
#pragma unroll
        for (int k = 0; k < kOneThreads / 64; k++) {
            const uint32_t x = wsum[k];
            if (k < w) off += x;
            tot += x;
        }
        uint4 ra, rb;
        ra.x = off + inc - s;
        ra.y = ra.x + a.x;
        ra.z = ra.y + a.y;
        ra.w = ra.z + a.z;
        rb.x = ra.w + a.w;
        rb.y = rb.x + b.x;
        rb.z = rb.y + b.y;
        rb.w = rb.z + b.z;
        tile4[2 * tid] = ra;
        tile4[2 * tid + 1] = rb;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kOneItems; k++) {
            const int64_t i = base + int64_t(k) * kOneThreads + tid;
            if (i < n) {
                const uint32_t r = tile[k * kOneThreads + tid];
                out[i] = r;
                if (out2) out2[i] = r + adc[k];
            }
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[n] = carry;
        if (out2) out2[n] = carry + add[n];
    }
}

// out[i] = a[i] + b[i], i < n
__global__ void k_add_u32(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                          uint32_t* __restrict__ out, int64_t n) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

// count per key; n taken from n_dev when given (device-side element count)
__global__ void k_count(const uint32_t* __restrict__ key, int64_t n,
                        const uint32_t* __restrict__ n_dev, uint32_t* cnt) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int64_t nn = n_dev ? int64_t(*n_dev) : n;
    const bool valid = i < nn;
    const uint32_t k = valid ? key[i] : 0u;
    int h, t;
    wave_runs(k, valid, h, t);
    const int lane = threadIdx.x & 63;
    if (valid && lane == t) atomicAdd(&cnt[k], uint32_t(t - h + 1));
}

__global__ void k_scatter(const uint32_t* __restrict__ key, int64_t n,
                          const uint32_t* __restrict__ n_dev,
                          const uint32_t* __restrict__ start, uint32_t* cursor,
                          uint32_t* __restrict__ order) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int64_t nn = n_dev ? int64_t(*n_dev) : n;
    const bool valid = i < nn;
    const uint32_t k = valid ? key[i] : 0u;
    int h, t;
    wave_runs(k, valid, h, t);
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (valid && lane == t) base = start[k] + atomicAdd(&cursor[k], uint32_t(t - h + 1));
    base = __shfl(base, t, 64);
    if (valid) order[base + uint32_t(lane - h)] = uint32_t(i);
}

// final slot of every element: its cell's start plus the number of elements
// of the same cell with a smaller index (index order inside each cell)
__global__ void k_rank(const uint32_t* __restrict__ key, int64_t n,
                       const uint32_t* __restrict__ n_dev, const uint32_t* __restrict__ start,
                       const uint32_t* __restrict__ placed, uint32_t* __restrict__ order,
                       uint32_t* __restrict__ slot_of) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int64_t nn = n_dev ? int64_t(*n_dev) : n;
    if (s >= nn) return;
    const uint32_t v = placed[s];
    const uint32_t c = key[v];
    const uint32_t b = start[c], e = start[c + 1];
    uint32_t r = 0;
    for (uint32_t t = b; t < e; t++) r += placed[t] < v ? 1u : 0u;
    order[b + r] = v;
    if (slot_of) slot_of[v] = b + r;
}

int fill_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return int(b > 0 ? b : 1);
}

}  // namespace

void fill_u32(Ctx& c, uint32_t* p, uint32_t v, int64_t n) {
    if (n <= 0) return;
    launch(c, k_fill_u32, fill_blocks(n), 256, p, v, n);
    MDD_HIP(hipGetLastError());
}
void fill_f64(Ctx& c, double* p, double v, int64_t n) {
    if (n <= 0) return;
    launch(c, k_fill_f64, fill_blocks(n), 256, p, v, n);
    MDD_HIP(hipGetLastError());
}
void fill_f32(Ctx& c, float* p, float v, int64_t n) {
    if (n <= 0) return;
    launch(c, k_fill_f32, fill_blocks(n), 256, p, v, n);
    MDD_HIP(hipGetLastError());
}

// out[0..n] = exclusive scan of in[0..n), out[n] = total
void exclusive_scan(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n) {
    if (n <= kOneBlockScan) {
        launch(c, k_scan_one, 1, kOneThreads, const_cast<uint32_t*>(in), out, n, 0,
               static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr));
        return;
    }
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb < 1) nb = 1;
    c.scan_tmp.resize(size_t(nb + 2));
    launch(c, k_scan_reduce, int(nb), kScanThreads, in, n, c.scan_tmp.p);
    launch(c, k_scan_apply, int(nb), kScanThreads, in, out, n, c.scan_tmp.p, nb);
    MDD_HIP(hipGetLastError());
}

// Stable counting sort of indices 0..n-1 by key in [0, ncells). The element
// count is n, or *n_dev (<= n) when n_dev is given. start: ncells+1;
// order: n (slot -> index); slot_of: n (index -> slot) or null; cnt: ncells
// scratch, already zero when cnt_zero, or already holding the per-key counts
// when counted (the kernel that wrote the keys counted them); placed: n
// scratch. With add given, out2[c] = start[c] + add[c] for c <= ncells as well.
void counting_sort(Ctx& c, const uint32_t* key, int64_t n, const uint32_t* n_dev, int64_t ncells,
                   uint32_t* cnt, uint32_t* start, uint32_t* placed, uint32_t* order,
                   uint32_t* slot_of, bool cnt_zero, const uint32_t* add, uint32_t* out2,
                   bool counted) {
    if (!counted) {
        if (!cnt_zero) fill_u32(c, cnt, 0u, ncells);
        if (n > 0) launch(c, k_count, blocks_for(n, 256), 256, key, n, n_dev, cnt);
    }
    if (ncells <= kOneBlockScan) {
        // scan and cursor reset in one pass
        launch(c, k_scan_one, 1, kOneThreads, cnt, start, ncells, 1, add, out2);
    } else {
        exclusive_scan(c, cnt, start, ncells);
        fill_u32(c, cnt, 0u, ncells);
        if (add) launch(c, k_add_u32, blocks_for(ncells + 1, 256), 256, start, add, out2, ncells + 1);
    }
    if (n > 0) {
        launch(c, k_scatter, blocks_for(n, 256), 256, key, n, n_dev, start, cnt, placed);
        launch(c, k_rank, blocks_for(n, 256), 256, key, n, n_dev, start, placed, order, slot_of);
    }
    MDD_HIP(hipGetLastError());
}

}  // namespace mdd
