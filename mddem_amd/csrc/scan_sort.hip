// Device-wide exclusive scan and stable counting sort (cell binning).
//
// The counting sort reproduces the order of the reference's binning
// (crates/mddem_core/src/neighbor.rs:857-889): within a cell, elements stay
// in index order. Placement uses atomics for speed, and a per-cell insertion
// sort then restores index order, so the result is deterministic.
#include "ctx.hpp"

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

// single block: exclusive scan of nb block sums in place; sums[nb] = total
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(uint32_t* sums, int64_t nb) {
    __shared__ uint32_t lds4[4];
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += kScanThreads) {
        int64_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, lds4, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out,
                                                             int64_t n,
                                                             const uint32_t* __restrict__ sums,
                                                             int64_t nb) {
    __shared__ uint32_t lds4[4];
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
    uint32_t run = block_excl_scan(s, lds4, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = sums[nb];
}

__global__ void k_count(const uint32_t* __restrict__ key, int64_t n, uint32_t* cnt) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < n) atomicAdd(&cnt[key[i]], 1u);
}

__global__ void k_scatter(const uint32_t* __restrict__ key, int64_t n,
                          const uint32_t* __restrict__ start, uint32_t* cursor,
                          uint32_t* __restrict__ order) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < n) {
        uint32_t c = key[i];
        uint32_t p = atomicAdd(&cursor[c], 1u);
        order[start[c] + p] = uint32_t(i);
    }
}

// restore index order inside each cell (insertion sort; cells hold few items)
__global__ void k_segsort(const uint32_t* __restrict__ start, int64_t ncells,
                          uint32_t* __restrict__ order) {
    int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (c >= ncells) return;
    uint32_t b = start[c], e = start[c + 1];
    for (uint32_t i = b + 1; i < e; i++) {
        uint32_t v = order[i];
        uint32_t j = i;
        while (j > b && order[j - 1] > v) {
            order[j] = order[j - 1];
            j--;
        }
        order[j] = v;
    }
}

__global__ void k_invert(const uint32_t* __restrict__ order, int64_t n,
                         uint32_t* __restrict__ slot_of) {
    int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s < n) slot_of[order[s]] = uint32_t(s);
}

int fill_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return int(b > 0 ? b : 1);
}

}  // namespace

void fill_u32(Ctx& c, uint32_t* p, uint32_t v, int64_t n) {
    if (n <= 0) return;
    k_fill_u32<<<fill_blocks(n), 256, 0, c.stream>>>(p, v, n);
    MDD_HIP(hipGetLastError());
}
void fill_f64(Ctx& c, double* p, double v, int64_t n) {
    if (n <= 0) return;
    k_fill_f64<<<fill_blocks(n), 256, 0, c.stream>>>(p, v, n);
    MDD_HIP(hipGetLastError());
}
void fill_f32(Ctx& c, float* p, float v, int64_t n) {
    if (n <= 0) return;
    k_fill_f32<<<fill_blocks(n), 256, 0, c.stream>>>(p, v, n);
    MDD_HIP(hipGetLastError());
}

// out[0..n] = exclusive scan of in[0..n), out[n] = total
void exclusive_scan(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n) {
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb < 1) nb = 1;
    c.scan_tmp.resize(size_t(nb + 1));
    k_scan_reduce<<<int(nb), kScanThreads, 0, c.stream>>>(in, n, c.scan_tmp.p);
    k_scan_sums<<<1, kScanThreads, 0, c.stream>>>(c.scan_tmp.p, nb);
    k_scan_apply<<<int(nb), kScanThreads, 0, c.stream>>>(in, out, n, c.scan_tmp.p, nb);
    MDD_HIP(hipGetLastError());
}

// Stable counting sort of indices 0..n-1 by key in [0, ncells).
// start: ncells+1, order: n (slot -> index), slot_of: n (index -> slot) or null.
// cnt: ncells scratch.
void counting_sort(Ctx& c, const uint32_t* key, int64_t n, int64_t ncells, uint32_t* cnt,
                   uint32_t* start, uint32_t* order, uint32_t* slot_of) {
    fill_u32(c, cnt, 0u, ncells);
    if (n > 0) k_count<<<blocks_for(n, 256), 256, 0, c.stream>>>(key, n, cnt);
    exclusive_scan(c, cnt, start, ncells);
    fill_u32(c, cnt, 0u, ncells);
    if (n > 0) {
        k_scatter<<<blocks_for(n, 256), 256, 0, c.stream>>>(key, n, start, cnt, order);
        k_segsort<<<blocks_for(ncells, 256), 256, 0, c.stream>>>(start, ncells, order);
        if (slot_of) k_invert<<<blocks_for(n, 256), 256, 0, c.stream>>>(order, n, slot_of);
    }
    MDD_HIP(hipGetLastError());
}

}  // namespace mdd
