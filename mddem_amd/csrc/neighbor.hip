// Neighbor rebuild for the gfx950 particle path.
//
// One rebuild reproduces, on the device, the reference's FullRebuild phases
// for a single process:
//   pbc (crates/mddem_core/src/domain.rs:413-459)
//   sort_atoms_by_bin (crates/mddem_core/src/neighbor.rs:686-794), stable
//   borders self-sends (crates/mddem_core/src/comm.rs:449-570) as periodic images
//   bin_neighbor_list (crates/mddem_core/src/neighbor.rs:810-1080)
// on the reference's bin grid, so the reference half list (same owners,
// same FP64 fma distance test) comes out bit for bit. Instead of storing
// only the half list, every particle gets a full row: the pairs it owns
// plus the pairs other particles own with it, each entry flagged, so the
// force kernels gather without atomics. Rows are built one wavefront per bin
// cell (k_build_cells): a candidate after the particle in bin order (a
// forward cell, or later in the same cell) is a pair the particle owns; a
// particle before it owns the pair instead, and for two particles the
// distance test is symmetric bit for bit. Pairs owned through a particle's
// periodic images are collected by the images themselves and appended to
// the particle's row. Passing candidates are compacted with a wavefront
// ballot, in a fixed order, so rows are deterministic.
//
// A batched ensemble shares one cell space: system s owns cells
// [cell_off, cell_off + ncells) and its own stencil runs and cutoffs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cells.hpp"

namespace mdd {

namespace {

__device__ inline uint32_t cell_of(const GridP& g, double x, double y, double z) {
    int cx = int(floor((x - g.origin[0]) * g.inv_bs[0]));
    int cy = int(floor((y - g.origin[1]) * g.inv_bs[1]));
    int cz = int(floor((z - g.origin[2]) * g.inv_bs[2]));
    cx = cx < 0 ? 0 : (cx > g.n[0] - 1 ? g.n[0] - 1 : cx);
    cy = cy < 0 ? 0 : (cy > g.n[1] - 1 ? g.n[1] - 1 : cy);
    cz = cz < 0 ? 0 : (cz > g.n[2] - 1 ? g.n[2] - 1 : cz);
    return uint32_t(cx * g.n[1] * g.n[2] + cy * g.n[2] + cz);
}

// Start of a rebuild: the step sequence resumes (abort cleared), the build
// status is reset and the sort's per-cell counts are zeroed.
__global__ void k_rebuild_begin(Status* st, uint32_t* __restrict__ cnt, int64_t ncells) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < ncells;
         i += int64_t(gridDim.x) * blockDim.x)
        cnt[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->abort_step = INT_MAX;
        st->overflow = 0u;
        st->max_row = 0;
        st->max_half = 0;
        st->ghost_total = 0u;
    }
}

// pbc wrap (wrap_periodic, domain.rs:396-407) + bin key of every particle,
// counted per cell for the sort (cnt zeroed by k_rebuild_begin; equal keys
// of a wavefront are counted by one atomic). Batched ensembles pass the
// per-system table: each particle then uses its own system's box and grid,
// and its key is offset into the shared cell space.
__global__ void k_wrap_key(int64_t n, double4* __restrict__ pos, int4* __restrict__ image,
                           BoxP box, GridP g, uint32_t* __restrict__ key, Status* st,
                           const SysP* __restrict__ sysp, const uint32_t* __restrict__ psys,
                           uint32_t* __restrict__ cnt) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const bool valid = i < n;
    uint32_t kc = 0u;
    if (valid) {
        int64_t coff = 0;
        if (sysp) {
            const SysP& S = sysp[psys[i]];
            box = S.box;
            g = S.grid;
            coff = S.cell_off;
        }
        double4 p = pos[i];
        int4 im = image[i];
        double x[3] = {p.x, p.y, p.z};
        int m[3] = {im.x, im.y, im.z};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (box.periodic[d]) {
                double high = box.lo[d] + box.L[d];
                if (x[d] < box.lo[d]) {
                    x[d] += box.L[d];
                    m[d] += -1;
                } else if (x[d] >= high) {
                    x[d] -= box.L[d];
                    m[d] += 1;
                }
            } else if (x[d] < box.lo[d] || x[d] >= box.hi[d]) {
                atomicOr(&st->overflow, 4u);  // left a fixed boundary
            }
        }
        p.x = x[0];
        p.y = x[1];
        p.z = x[2];
        pos[i] = p;
        image[i] = make_int4(m[0], m[1], m[2], 0);
        kc = uint32_t(coff + cell_of(g, x[0], x[1], x[2]));
        key[i] = kc;
    }
    int h, t;
    wave_runs(kc, valid, h, t);
    if (valid && int(threadIdx.x & 63) == t) atomicAdd(&cnt[kc], uint32_t(t - h + 1));
}

struct PermArgs {
    int64_t n, npad;
    const uint32_t* order;
    const double4 *pos, *vel, *omega, *quat;
    double4 *pos2, *vel2, *omega2, *quat2, *build_pos;
    const uint32_t *tag, *type, *key;
    uint32_t *tag2, *type2, *key2;
    const int4* image;
    int4* image2;
    const double *mass, *cutr;
    double *mass2, *cutr2;
    const double *f64, *t64;
    double *f64b, *t64b;
    const float *f32, *t32;
    float *f32b, *t32b;
    int4* xq;   // LJ mixed: the fixed-point copy of the permuted positions, or null
    int4* bxq;  // with xq: the same values as the fixed-point build positions
    double qs[3];
    int dem;
};

// apply the sort permutation to every per-particle array (Atom::apply_permutation,
// atom.rs:383-394) and record the build positions (save_build_positions)
__global__ void k_permute(PermArgs a) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (k >= a.n) return;
    uint32_t o = a.order[k];
    double4 p = a.pos[o];
    a.pos2[k] = p;
    a.build_pos[k] = p;
    if (a.xq) {
        const int4 q = make_int4(quantize_axis(p.x, a.qs[0]), quantize_axis(p.y, a.qs[1]),
                                 quantize_axis(p.z, a.qs[2]), 0);
        a.xq[k] = q;
        a.bxq[k] = q;
    }
    a.vel2[k] = a.vel[o];
    a.tag2[k] = a.tag[o];
    a.type2[k] = a.type[o];
    a.key2[k] = a.key[o];
    a.image2[k] = a.image[o];
    a.mass2[k] = a.mass[o];
    a.cutr2[k] = a.cutr[o];
    if (a.f64) {
        for (int d = 0; d < 3; d++) a.f64b[d * a.npad + k] = a.f64[d * a.npad + o];
    }
    if (a.f32) {
        for (int d = 0; d < 3; d++) a.f32b[d * a.npad + k] = a.f32[d * a.npad + o];
    }
    if (a.dem) {
        a.omega2[k] = a.omega[o];
        a.quat2[k] = a.quat[o];
        if (a.t64)
            for (int d = 0; d < 3; d++) a.t64b[d * a.npad + k] = a.t64[d * a.npad + o];
        if (a.t32)
            for (int d = 0; d < 3; d++) a.t32b[d * a.npad + k] = a.t32[d * a.npad + o];
    }
}

// image flags of one particle: which faces it is within the ghost cutoff of
// (pack_border_atoms, comm.rs:256-284; swap 0 near the low face gets +L,
// swap 1 near the high face gets -L)
__device__ inline void image_flags(const BoxP& box, double gc, const double x[3], int lowf[3],
                                   int highf[3]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lowf[d] = box.periodic[d] && (x[d] < box.lo[d] + gc);
        highf[d] = box.periodic[d] && (x[d] >= box.hi[d] - gc);
    }
}

// Start of one build attempt (a retry repeats it with larger capacities):
// status reset and the image sort's per-cell counts zeroed, fused with the
// image count of every particle.
__global__ void k_ghost_count(int64_t n, const double4* __restrict__ pos, BoxP box, double gc,
                              uint32_t* __restrict__ gcnt, const SysP* __restrict__ sysp,
                              const uint32_t* __restrict__ psys, Status* st,
                              uint32_t* __restrict__ cnt, int64_t ncells) {
    for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < ncells;
         c += int64_t(gridDim.x) * blockDim.x)
        cnt[c] = 0u;
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i == 0) {
        st->abort_step = INT_MAX;
        st->overflow &= 4u;  // keep the fixed-boundary exit seen by the wrap
        st->max_row = 0;
        st->max_half = 0;
        st->ghost_total = 0u;
    }
    if (i >= n) return;
    if (sysp) {
        const SysP& S = sysp[psys[i]];
        box = S.box;
        gc = S.gc;
    }
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    int lf[3], hf[3];
    image_flags(box, gc, x, lf, hf);
    gcnt[i] = uint32_t((1 + lf[0] + hf[0]) * (1 + lf[1] + hf[1]) * (1 + lf[2] + hf[2]) - 1);
}

// periodic images of every particle (x shifted by exactly +-L per axis),
// counted per bin cell for the image sort (cnt zeroed by k_ghost_count);
// when they do not all fit, none is written and the live count is 0
__global__ void k_ghost_emit(int64_t n, const double4* __restrict__ pos, BoxP box, double gc,
                             GridP grid, const uint32_t* __restrict__ goff, int64_t ghost_cap,
                             uint32_t* __restrict__ groot, uint8_t* __restrict__ gcode,
                             double4* __restrict__ gpos, uint32_t* __restrict__ gkey,
                             uint32_t* ng_live, Status* st, const SysP* __restrict__ sysp,
                             const uint32_t* __restrict__ psys, uint32_t* __restrict__ cnt) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const uint32_t total = goff[n];
    const bool fits = int64_t(total) <= ghost_cap;
    if (i == 0) {
        st->ghost_total = total;
        *ng_live = fits ? total : 0u;
        if (!fits) atomicOr(&st->overflow, 1u);
    }
    if (!fits) return;
    // every lane stays to the end: the per-cell counts are added once per
    // run of equal cells among the wavefront's images of one shift (the
    // particles are in cell order, so such images are too)
    const bool live = i < n;
    if (!live) i = 0;
    int64_t coff = 0;
    if (sysp) {
        const SysP& S = sysp[psys[i]];
        box = S.box;
        gc = S.gc;
        grid = S.grid;
        coff = S.cell_off;
    }
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    int lf[3], hf[3];
    image_flags(box, gc, x, lf, hf);
    uint32_t g = live ? goff[i] : 0u;
    const int sv[3] = {0, 1, -1};
    for (int a = 0; a < 3; a++) {
        const int sx = sv[a];
        const bool ha = live && !((sx == 1 && !lf[0]) || (sx == -1 && !hf[0]));
        if (__ballot(ha) == 0ull) continue;
        for (int b = 0; b < 3; b++) {
            const int sy = sv[b];
            const bool hb = ha && !((sy == 1 && !lf[1]) || (sy == -1 && !hf[1]));
            if (__ballot(hb) == 0ull) continue;
            for (int cz = 0; cz < 3; cz++) {
                const int sz = sv[cz];
                if (sx == 0 && sy == 0 && sz == 0) continue;
                const bool has = hb && !((sz == 1 && !lf[2]) || (sz == -1 && !hf[2]));
                if (__ballot(has) == 0ull) continue;
                uint32_t key = 0u;
                if (has) {
                    const double gx = x[0] + image_shift(sx, box.L[0]);
                    const double gy = x[1] + image_shift(sy, box.L[1]);
                    const double gz = x[2] + image_shift(sz, box.L[2]);
                    key = uint32_t(coff + cell_of(grid, gx, gy, gz));
                    groot[g] = uint32_t(i);
                    gcode[g] = uint8_t(code_of(sx, sy, sz));
                    gpos[g] = make_double4(gx, gy, gz, 0.0);
                    gkey[g] = key;
                    g++;
                }
                // the image sort's per-cell count
                int h, t;
                wave_runs(key, has, h, t);
                if (has && int(threadIdx.x & 63) == t) atomicAdd(&cnt[key], uint32_t(t - h + 1));
            }
        }
    }
}

// combined bin starts: particles of a cell first (already sorted), then images

// binned copies of position / neighbor skin radius and the encoded entry
__global__ void k_bin_fill(int64_t n, int64_t ghost_cap, const uint32_t* __restrict__ ng_live,
                           const uint32_t* __restrict__ key, const uint32_t* __restrict__ lstart,
                           const uint32_t* __restrict__ gstart,
                           const uint32_t* __restrict__ bstart,
                           const uint32_t* __restrict__ gorder, const uint32_t* __restrict__ gkey,
                           const double4* __restrict__ pos, const double* __restrict__ cutr,
                           const double4* __restrict__ gpos, const uint32_t* __restrict__ groot,
                           const uint8_t* __restrict__ gcode, double4* __restrict__ bin_pos,
                           uint32_t* __restrict__ bin_entry, uint32_t* __restrict__ bin_gidx,
                           uint32_t* __restrict__ slot_local, uint32_t* __restrict__ bin_cell) {
    const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (t < n) {
        const uint32_t c = key[t];
        const uint32_t s = bstart[c] + uint32_t(t - lstart[c]);
        const double4 p = pos[t];
        bin_pos[s] = make_double4(p.x, p.y, p.z, cutr[t]);
        bin_entry[s] = enc_entry(uint32_t(t), kCodeNone, false);
        slot_local[t] = s;
        if (bin_cell) bin_cell[s] = c;
        return;
    }
    const int64_t gs = t - n;
    if (gs >= int64_t(*ng_live) || gs >= ghost_cap) return;
    const uint32_t g = gorder[gs];
    const uint32_t c = gkey[g];
    const uint32_t s = bstart[c] + (lstart[c + 1] - lstart[c]) + (uint32_t(gs) - gstart[c]);
    const double4 q = gpos[g];
    const uint32_t r = groot[g];
    bin_pos[s] = make_double4(q.x, q.y, q.z, cutr[r]);
    bin_entry[s] = enc_entry(r, gcode[g], false);
    bin_gidx[s] = g;
    if (bin_cell) bin_cell[s] = c;
}

constexpr int kCellWaves = 4;  // cells (one per wavefront) per 256-thread block
constexpr int kNch = 10;       // register-resident 64-candidate chunks per window
constexpr int kMaxRuns = 64;   // contiguous bin ranges of one cell's stencil

struct CellArgs {
    int64_t ncells, npad;
    int maxnb, lg;
    const double4* __restrict__ bin_pos;
    const uint32_t* __restrict__ bin_entry;
    const uint32_t* __restrict__ bin_gidx;
    const uint32_t* __restrict__ bin_start;
    const int* __restrict__ run_first;
    const int* __restrict__ run_last;
    int nruns, run0;
    double cutoff_sq, sf2;
    uint32_t* __restrict__ nbr;
    uint32_t* __restrict__ nnb;
    uint32_t* __restrict__ nhalf;
    uint32_t* __restrict__ gnbr;
    uint32_t* __restrict__ gcnt;
    Status* st;
    const SysP* __restrict__ sysp;
    const int64_t* __restrict__ sys_cell_off;
    int nsys;
    // tile rows (LJ): 16-bit candidate indices, laid out per cell by tile_addr
    uint16_t* __restrict__ crow;
    const uint32_t* __restrict__ lstart;  // first particle of each cell
    int tile_cap;
};

// Pair cutoff of the cell's system: uniform (r2 < cut2) or per pair
// (r2 < (s_i + s_j)^2 sf2), as neighbor.rs:938 and 1003-1004 choose it.
struct CellCut {
    double cut2, sf2;
    int uni;
};

template <bool UNIFORM>
__device__ inline bool within(const CellCut& cc, double r2, double si, double sj) {
    if (UNIFORM) return r2 < cc.cut2;
    if (cc.uni) return r2 < cc.cut2;
    const double s = si + sj;
    return r2 < s * s * cc.sf2;
}

// Candidate chunks of one window, held in registers: lane l of chunk c is
// flat candidate win + 64 c + l of the cell's stencil ranges (bin order).
template <bool UNIFORM>
struct Window {
    double x[kNch], y[kNch], z[kNch], w[kNch];
    uint32_t e[kNch];
    // lanes a particle of the cell may pair with: every particle, and every
    // periodic image that is not in a backward cell (those pairs belong to
    // the image's particle, found through its own images)
    uint64_t okp[kNch];
};

// Pair tests of one particle of the cell (flat index fs relative to the
// window) against the window. Candidates before it in bin order own the
// pair with it (entry flagged), candidates after it are pairs it owns.
//
// TILE: the row is written as 16-bit indices into the cell's candidate list
// (flat index win + 64 c + lane) at tile_addr(s0, cap, M, Gt, m, k), for the
// tile force kernel; tg = (s0, M, Gt, m, 65536 / Gt rounded up).
struct TileMember {
    uint32_t s0, M, G, m, ginv;
};

template <bool UNIFORM, bool TILE>
__device__ inline void test_particle(const CellArgs& a, const CellCut& cc,
                                     const Window<UNIFORM>& W, int nch, int fs, const double4& xi,
                                     uint32_t q, uint32_t& cnt, uint32_t& nrev, uint32_t win,
                                     const TileMember& tm) {
    const int lane = int(threadIdx.x & 63);
    const uint32_t G = 1u << a.lg;
#pragma unroll
    for (int c = 0; c < kNch; c++) {
        if (c < nch) {
            const int f = c * 64 + lane;
            const double dx = W.x[c] - xi.x, dy = W.y[c] - xi.y, dz = W.z[c] - xi.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            const bool in = within<UNIFORM>(cc, r2, xi.w, W.w[c]);
            const uint64_t pass = __ballot(in) & __ballot(f != fs) & W.okp[c];
            if (pass != 0ull) {
                const uint64_t before = __ballot(f < fs);
                const uint32_t k = cnt + lanes_before(pass);
                if (__builtin_amdgcn_inverse_ballot_w64(pass) && k < uint32_t(a.maxnb)) {
                    if (TILE) {
                        const uint32_t kq = (k * tm.ginv) >> 16;  // k / G (k < 2^15)
                        const uint32_t at = tm.s0 * uint32_t(a.tile_cap) + kq * (tm.M * tm.G) +
                                            tm.m * tm.G + (k - kq * tm.G);
                        a.crow[at] = uint16_t(win + uint32_t(f));
                    } else {
                        // row_addr with 24-bit multiplies (npad < 2^24 is checked by the host)
                        const uint32_t at = ((__umul24(k >> a.lg, uint32_t(a.npad)) + q) << a.lg) +
                                            (k & (G - 1u));
                        a.nbr[at] = __builtin_amdgcn_inverse_ballot_w64(before)
                                        ? (W.e[c] | kRevBit)
                                        : W.e[c];
                    }
                }
                cnt += uint32_t(__popcll(pass));
                nrev += uint32_t(__popcll(pass & before));
            }
        }
    }
}

// Pair tests of one periodic image of the cell: the particles before it in
// bin order (backward cells and this cell) own the pair with it; they are
// collected, flagged, into the image's scratch row.
template <bool UNIFORM>
__device__ inline void test_image(const CellArgs& a, const CellCut& cc,
                                  const Window<UNIFORM>& W, int nch, int f_end,
                                  const double4& xi, uint32_t code_i, uint32_t* __restrict__ out,
                                  uint32_t& cnt) {
    const int lane = int(threadIdx.x & 63);
    const uint32_t tagbits = (code_i << kIdxBits) | kRevBit;
#pragma unroll
    for (int c = 0; c < kNch; c++) {
        const uint64_t cm =
            c < nch ? (__ballot(entry_code(W.e[c]) == kCodeNone) & lanes_below(f_end - c * 64))
                    : 0ull;
        if (cm != 0ull) {
            const double dx = W.x[c] - xi.x, dy = W.y[c] - xi.y, dz = W.z[c] - xi.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            const bool in = within<UNIFORM>(cc, r2, xi.w, W.w[c]);
            const uint64_t pass = __ballot(in) & cm;
            if (pass != 0ull) {
                const uint32_t k = cnt + lanes_before(pass);
                if (__builtin_amdgcn_inverse_ballot_w64(pass) && k < uint32_t(a.maxnb))
                    out[k] = entry_idx(W.e[c]) | tagbits;
                cnt += uint32_t(__popcll(pass));
            }
        }
    }
}

// Neighbor rows, one wavefront per bin cell. The cell's stencil cells form
// a few runs of consecutive cells, i.e. contiguous bin ranges; concatenated
// they are the cell's candidate list, whose order is bin order. Candidates
// are loaded 64 per chunk into registers once, and every member of the cell
// (particles, then periodic images) is tested against them with uniform
// control flow; passing lanes are compacted with a ballot, so each row lists
// its entries in bin order.
//
// In a batched ensemble (BATCH) the cell space is shared: the wave finds its
// cell's system and uses that system's cell range, stencil runs and cutoff.
template <bool UNIFORM, bool BATCH, bool TILE>
__global__ void __launch_bounds__(256) k_build_cells(CellArgs a) {
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = int(threadIdx.x & 63);
    const int64_t C = int64_t(xcd_block(blockIdx.x, gridDim.x)) * kCellWaves + w;
    if (C >= a.ncells) return;
    const uint32_t i0 = a.bin_start[C], i1 = a.bin_start[C + 1];
    if (i0 == i1) return;
    int64_t cmin = 0, cmax = a.ncells - 1;
    int run_off = 0, nruns = a.nruns, run0 = a.run0;
    CellCut cc{a.cutoff_sq, a.sf2, UNIFORM ? 1 : 0};
    if (BATCH) {
        // the system whose cell range holds C
        int lo = 0, hi = a.nsys - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.sys_cell_off[mid] <= C)
                lo = mid;
            else
                hi = mid - 1;
        }
        const SysP& S = a.sysp[lo];
        cmin = S.cell_off;
        cmax = S.cell_off + S.grid.ncells - 1;
        run_off = S.run_off;
        nruns = S.nruns;
        run0 = S.run0;
        cc.cut2 = S.cutoff_sq;
        cc.sf2 = S.sf2;
        cc.uni = S.uniform;
    }
    // stencil runs in lanes: bin range [rb, rb + rcnt), flat offset excl
    uint32_t rcnt = 0, rb = 0;
    if (lane < nruns) {
        int64_t lo = C + a.run_first[run_off + lane], hi = C + a.run_last[run_off + lane];
        if (lo < cmin) lo = cmin;
        if (hi > cmax) hi = cmax;
        if (lo <= hi) {
            rb = a.bin_start[lo];
            rcnt = a.bin_start[hi + 1] - rb;
        }
    }
    uint32_t incl = rcnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    const uint32_t excl = incl - rcnt;
    const uint32_t rbase = rb - excl;  // bin index = flat index + rbase (mod 2^32)
    const uint32_t total = __builtin_amdgcn_readlane(incl, nruns - 1);
    const uint32_t base0 = __builtin_amdgcn_readlane(rbase, run0);
    // flat indices of this cell's first slot and of the slot after it
    const int f_i0 = int(i0 - base0), f_i1 = int(i1 - base0);
    const uint32_t ni = i1 - i0;
    TileMember tm{0u, 0u, 1u, 0u, 65536u};
    if (TILE) {
        tm.s0 = a.lstart[C];
        tm.M = a.lstart[C + 1] - tm.s0;
        tm.G = uint32_t(tile_lanes(tm.M));
        tm.ginv = (65536u + tm.G - 1u) / tm.G;
        if (total > 65535u) atomicOr(&a.st->overflow, 8u);  // tile index range
    }
    for (uint32_t ib = 0; ib < ni; ib += 64) {
        const uint32_t nib = ni - ib < 64u ? ni - ib : 64u;
        uint32_t cntv = 0, revv = 0;  // lane l: member ib + l
        // the members of this batch, lane l holding member ib + l
        double4 mx = make_double4(0.0, 0.0, 0.0, 0.0);
        uint32_t me = 0u, mg = 0u;
        if (uint32_t(lane) < nib) {
            const uint32_t slot = i0 + ib + uint32_t(lane);
            mx = a.bin_pos[slot];
            me = a.bin_entry[slot];
            if (entry_code(me) != kCodeNone) mg = a.bin_gidx[slot];
        }
        // members ib .. ib + nloc - 1 are particles (they precede the images)
        const uint32_t nloc =
            uint32_t(__popcll(__ballot(uint32_t(lane) < nib && entry_code(me) == kCodeNone)));
        for (uint32_t win = 0; win < total; win += kNch * 64) {
            Window<UNIFORM> W;
            uint32_t base[kNch];
#pragma unroll
            for (int c = 0; c < kNch; c++) base[c] = 0;
            for (int r = 0; r < nruns; r++) {
                const uint32_t pr = __builtin_amdgcn_readlane(excl, r);
                const uint32_t br = __builtin_amdgcn_readlane(rbase, r);
#pragma unroll
                for (int c = 0; c < kNch; c++)
                    if (win + uint32_t(c * 64 + lane) >= pr) base[c] = br;
            }
            // every chunk's loads are issued before any is used (lanes past
            // the list read slot 0 and are cleared), so the window costs one
            // memory latency rather than one per chunk
            double4 qv[kNch];
            uint32_t ev[kNch];
#pragma unroll
            for (int c = 0; c < kNch; c++) {
                const uint32_t f = win + uint32_t(c * 64 + lane);
                const uint32_t src = f < total ? f + base[c] : 0u;
                qv[c] = a.bin_pos[src];
                ev[c] = a.bin_entry[src];
            }
#pragma unroll
            for (int c = 0; c < kNch; c++) {
                const uint32_t f = win + uint32_t(c * 64 + lane);
                const bool valid = f < total;
                const double4 q = valid ? qv[c] : make_double4(0.0, 0.0, 0.0, 0.0);
                const uint32_t e = valid ? ev[c] : 0u;
                W.x[c] = q.x;
                W.y[c] = q.y;
                W.z[c] = q.z;
                W.w[c] = q.w;
                W.e[c] = e;
                // invalid lanes hold entry 0, whose image code is not kCodeNone
                const bool particle = valid && entry_code(e) == kCodeNone;
                W.okp[c] = __ballot(particle || (valid && int(f) >= f_i0));
            }
            const uint32_t left = (total - win + 63u) / 64u;
            const int nch = left < uint32_t(kNch) ? int(left) : kNch;
            const int shift = int(win);
            for (uint32_t ii = 0; ii < nloc; ii++) {
                const uint32_t ei = uint32_t(__builtin_amdgcn_readlane(int(me), int(ii)));
                double4 xi;
                xi.x = readlane_f64(mx.x, int(ii));
                xi.y = readlane_f64(mx.y, int(ii));
                xi.z = readlane_f64(mx.z, int(ii));
                xi.w = UNIFORM ? 0.0 : readlane_f64(mx.w, int(ii));
                uint32_t cnt = __builtin_amdgcn_readlane(cntv, int(ii));
                uint32_t nrev = __builtin_amdgcn_readlane(revv, int(ii));
                const int fs = f_i0 + int(ib + ii) - shift;
                if (TILE) tm.m = ib + ii;
                test_particle<UNIFORM, TILE>(a, cc, W, nch, fs, xi, entry_idx(ei), cnt, nrev, win,
                                             tm);
                if (uint32_t(lane) == ii) {
                    cntv = cnt;
                    revv = nrev;
                }
            }
            for (uint32_t ii = nloc; ii < nib; ii++) {
                const uint32_t ei = uint32_t(__builtin_amdgcn_readlane(int(me), int(ii)));
                double4 xi;
                xi.x = readlane_f64(mx.x, int(ii));
                xi.y = readlane_f64(mx.y, int(ii));
                xi.z = readlane_f64(mx.z, int(ii));
                xi.w = UNIFORM ? 0.0 : readlane_f64(mx.w, int(ii));
                uint32_t cnt = __builtin_amdgcn_readlane(cntv, int(ii));
                const uint32_t g = uint32_t(__builtin_amdgcn_readlane(int(mg), int(ii)));
                test_image<UNIFORM>(a, cc, W, nch, f_i1 - shift, xi, uint32_t(entry_code(ei)),
                                    a.gnbr + size_t(g) * uint32_t(a.maxnb), cnt);
                if (uint32_t(lane) == ii) cntv = cnt;
            }
        }
        if (uint32_t(lane) < nib) {
            if (entry_code(me) == kCodeNone) {
                a.nnb[entry_idx(me)] = cntv;
                a.nhalf[entry_idx(me)] = cntv - revv;
            } else {
                a.gcnt[mg] = cntv;
            }
        }
    }
}

// The same rows for sparse grids (most cells empty or nearly so, where a
// wavefront per cell idles): one lane per member slot, walking its cell's
// stencil runs in bin order. Candidate order, the ownership flags and the
// counts are those of k_build_cells, so the rows are identical.
template <bool UNIFORM>
__global__ void __launch_bounds__(256) k_build_sparse(CellArgs a,
                                                      const uint32_t* __restrict__ bin_cell,
                                                      const uint32_t* __restrict__ ng_live,
                                                      int64_t n) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= n + int64_t(*ng_live)) return;
    const int64_t C = int64_t(bin_cell[s]);
    const uint32_t i0 = a.bin_start[C], i1 = a.bin_start[C + 1];
    const CellCut cc{a.cutoff_sq, a.sf2, UNIFORM ? 1 : 0};
    // flat index of the cell's first slot: the runs before run0, then the
    // cells of run0 before C
    uint32_t f_i0 = 0;
    for (int r = 0; r < a.nruns; r++) {
        int64_t lo = C + a.run_first[r], hi = C + a.run_last[r];
        if (lo < 0) lo = 0;
        if (hi > a.ncells - 1) hi = a.ncells - 1;
        if (lo > hi) continue;
        if (r == a.run0) {
            f_i0 += i0 - a.bin_start[lo];
            break;
        }
        f_i0 += a.bin_start[hi + 1] - a.bin_start[lo];
    }
    const uint32_t e = a.bin_entry[s];
    const double4 xi = a.bin_pos[s];
    const bool image = entry_code(e) != kCodeNone;
    const uint32_t fs = f_i0 + (uint32_t(s) - i0), f_end = f_i0 + (i1 - i0);
    const uint32_t q = entry_idx(e);
    const uint32_t G = 1u << a.lg;
    uint32_t* out = nullptr;
    uint32_t tagbits = 0u;
    if (image) {
        out = a.gnbr + size_t(a.bin_gidx[s]) * uint32_t(a.maxnb);
        tagbits = (uint32_t(entry_code(e)) << kIdxBits) | kRevBit;
    }
    uint32_t cnt = 0, nrev = 0, flat = 0;
    for (int r = 0; r < a.nruns; r++) {
        int64_t lo = C + a.run_first[r], hi = C + a.run_last[r];
        if (lo < 0) lo = 0;
        if (hi > a.ncells - 1) hi = a.ncells - 1;
        if (lo > hi) continue;
        const uint32_t b0 = a.bin_start[lo], b1 = a.bin_start[hi + 1];
        for (uint32_t t = b0; t < b1; t++) {
            const uint32_t f = flat + (t - b0);
            const uint32_t et = a.bin_entry[t];
            const bool particle = entry_code(et) == kCodeNone;
            if (image ? (!particle || f >= f_end) : (f == fs || !(particle || f >= f_i0)))
                continue;
            const double4 pt = a.bin_pos[t];
            const double dx = pt.x - xi.x, dy = pt.y - xi.y, dz = pt.z - xi.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (!within<UNIFORM>(cc, r2, xi.w, pt.w)) continue;
            const uint32_t k = cnt++;
            if (image) {
                if (k < uint32_t(a.maxnb)) out[k] = entry_idx(et) | tagbits;
            } else {
                const bool before = f < fs;
                if (before) nrev++;
                if (k < uint32_t(a.maxnb))
                    a.nbr[((size_t(k >> a.lg) * size_t(a.npad) + q) << a.lg) + (k & (G - 1u))] =
                        before ? (et | kRevBit) : et;
            }
        }
        flat += b1 - b0;
    }
    if (image) {
        a.gcnt[a.bin_gidx[s]] = cnt;
    } else {
        a.nnb[q] = cnt;
        a.nhalf[q] = cnt - nrev;
    }
}

// Row length of every particle: its own cell pass plus the pairs owned
// through each of its periodic images. Row maxima go to per-block slots
// (blk[2 b], blk[2 b + 1]); k_rows_gate reduces them, so no two blocks contend
// for one address.
__global__ void __launch_bounds__(256) k_row_totals(int64_t n, const uint32_t* __restrict__ goff,
                                                    const uint32_t* __restrict__ ng_live,
                                                    const uint32_t* __restrict__ gcnt,
                                                    const uint32_t* __restrict__ nmain,
                                                    const uint32_t* __restrict__ nhalf,
                                                    uint32_t* __restrict__ nnb,
                                                    uint32_t* __restrict__ blk,
                                                    uint32_t* __restrict__ nx) {
    __shared__ uint32_t lds[2][4];
    const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    uint32_t cnt = 0, nh = 0;
    if (q < n) {
        cnt = nmain[q];
        nh = nhalf[q];
        uint32_t xs = 0;
        if (*ng_live != 0u)
            for (uint32_t g = goff[q]; g < goff[q + 1]; g++) xs += gcnt[g];
        cnt += xs;
        nnb[q] = cnt;
        if (nx) nx[q] = xs;  // tile rows: image-owned pairs kept apart
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(cnt, o, 64), y = __shfl_xor(nh, o, 64);
        cnt = x > cnt ? x : cnt;
        nh = y > nh ? y : nh;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[0][w] = cnt;
        lds[1][w] = nh;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t mc = 0, mh = 0;
        for (int k = 0; k < int(blockDim.x >> 6); k++) {
            mc = lds[0][k] > mc ? lds[0][k] : mc;
            mh = lds[1][k] > mh ? lds[1][k] : mh;
        }
        blk[2 * blockIdx.x] = mc;
        blk[2 * blockIdx.x + 1] = mh;
    }
}


// Copy each image's scratch row to the end of its particle's row: after the
// particle's own pass and the rows of the particle's earlier images.
constexpr int kCopyLanes = 16;
__global__ void __launch_bounds__(256) k_copy_image_rows(int64_t ghost_cap, int64_t npad,
                                                         int maxnb, int lg,
                                                         const uint32_t* __restrict__ ng_live,
                                                         const uint32_t* __restrict__ groot,
                                                         const uint32_t* __restrict__ goff,
                                                         const uint32_t* __restrict__ gcnt,
                                                         const uint32_t* __restrict__ nmain,
                                                         const uint32_t* __restrict__ gnbr,
                                                         uint32_t* __restrict__ nbr,
                                                         uint32_t* __restrict__ xnbr) {
    const int64_t gt = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int64_t g = gt / kCopyLanes;
    const uint32_t t = uint32_t(gt % kCopyLanes);
    if (g >= ghost_cap || g >= int64_t(*ng_live)) return;
    const uint32_t p = groot[g];
    // tile rows (xnbr given): a separate row per particle, p * maxnb + k
    uint32_t off = xnbr ? 0u : nmain[p];
    for (uint32_t h = goff[p]; h < uint32_t(g); h++) off += gcnt[h];
    uint32_t c = gcnt[g];
    if (c > uint32_t(maxnb)) c = uint32_t(maxnb);
    for (uint32_t k = t; k < c; k += kCopyLanes) {
        const uint32_t kk = off + k;
        if (kk >= uint32_t(maxnb)) continue;
        const uint32_t e = gnbr[size_t(g) * uint32_t(maxnb) + k];
        if (xnbr)
            xnbr[size_t(p) * uint32_t(maxnb) + kk] = e;
        else
            nbr[row_addr(p, int(kk), npad, lg)] = e;
    }
}

// The particles whose row is not empty (the order is free: each particle's
// force is computed by one thread alone). A workgroup lists kListPer x 256
// particles into LDS and reserves their slots with one atomic: on the 4M
// sphere bed a per-wavefront atomic on the one counter (62,500 of them,
// serialized at the memory side) took 0.7 ms.
constexpr int kListPer = 16;

__global__ void __launch_bounds__(256) k_list_rows(int64_t n, const uint32_t* __restrict__ nnb,
                                                   uint32_t* __restrict__ listed,
                                                   uint32_t* __restrict__ count) {
    __shared__ uint32_t buf[256 * kListPer];
    __shared__ uint32_t wcount[4];
    __shared__ uint32_t base_s;
    const int w = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63u);
    const int64_t b0 = int64_t(blockIdx.x) * (256 * kListPer);
    uint32_t* mine = buf + w * 64 * kListPer;
    uint32_t local = 0u;  // this wavefront's listed count (uniform)
    for (int p = 0; p < kListPer; p++) {
        const int64_t i = b0 + (int64_t(w) * kListPer + p) * 64 + lane;
        const bool has = i < n && nnb[i] != 0u;
        const uint64_t m = __ballot(has);
        if (has) mine[local + lanes_before(m)] = uint32_t(i);
        local += uint32_t(__popcll(m));
    }
    if (lane == 0) wcount[w] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        base_s = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    uint32_t off = base_s;
    for (int k = 0; k < w; k++) off += wcount[k];
    for (uint32_t t = uint32_t(lane); t < local; t += 64u) listed[off + t] = mine[t];
}

// The DEM pair pass's particle order: a band of kTileCols bin columns
// along y, x after x, band after band. Each (band, x) segment is the
// particles of cells [x, y0 .. y1 - 1, all z], one contiguous range of the
// bin-sorted particles; in flight on one XCD they and their neighbors span
// about 13 x 18 columns instead of three whole x slabs.
constexpr int kTileCols = 16;

__global__ void k_tile_count(GridP g, const uint32_t* __restrict__ start,
                             uint32_t* __restrict__ cnt, int64_t nseg) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= nseg) return;
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t band = s / nx, x = s - band * nx;
    const int64_t y0 = band * kTileCols, y1 = y0 + kTileCols < ny ? y0 + kTileCols : ny;
    cnt[s] = start[(x * ny + y1) * nz] - start[(x * ny + y0) * nz];
}

// one workgroup per segment
__global__ void __launch_bounds__(256) k_tile_emit(GridP g, const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ list) {
    const int64_t s = blockIdx.x;
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t band = s / nx, x = s - band * nx;
    const int64_t y0 = band * kTileCols;
    const uint32_t lo = start[(x * ny + y0) * nz];
    const uint32_t o = off[s], len = off[s + 1] - o;
    for (uint32_t t = threadIdx.x; t < len; t += 256u) list[o + t] = lo + t;
}

// the listed count into the status, where the host reads it
__global__ void k_note_listed(const uint32_t* __restrict__ count, Status* st) {
    st->listed = *count;
}

// carry shear history across a rebuild: a new slot takes the old slot of the
// same particle whose partner has the same tag (the reference keys history
// by partner tag, tangential.rs:54-133, and the spring is stored in the
// lower tag's frame, so it transfers unchanged). Only slots that were in
// contact at the last force pass carry history (the reference prunes the
// rest, contact.rs:586-589); their contact bits move with them. With
// by_owner set (the reference's behaviour) the pair must also sit in the
// same particle's half-list row as before: the reference keeps a contact's
// history in its owner's list only (contact.rs:358-364, 575-582), so when a
// rebuild moves the pair into the partner's row the partner finds no entry
// and the spring restarts from zero, and the old owner prunes its entry.
__global__ void k_remap_history(int64_t n, int64_t npad, int maxnb, int lg,
                                const uint32_t* __restrict__ nbr,
                                const uint32_t* __restrict__ nnb,
                                const uint32_t* __restrict__ tag,
                                const uint32_t* __restrict__ order, int64_t npad_old,
                                int maxnb_old, int lg_old, const uint32_t* __restrict__ nbr_old,
                                const uint32_t* __restrict__ nnb_old,
                                const uint32_t* __restrict__ tag_old,
                                const double* __restrict__ hist_old,
                                const uint64_t* __restrict__ cmask_old,
                                double* __restrict__ hist, uint64_t* __restrict__ cmask,
                                int have_old, int nslot, int by_owner) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int64_t cs = int64_t(maxnb) * npad;
    const int64_t cso = int64_t(maxnb_old) * npad_old;
    int cnt = int(nnb[i]);
    if (cnt > maxnb) cnt = maxnb;
    const uint32_t oi = order[i];
    uint64_t om = have_old ? cmask_old[oi] : 0ull;
    int ocnt = have_old ? int(nnb_old[oi]) : 0;
    if (ocnt > maxnb_old) ocnt = maxnb_old;
    uint64_t nm = 0ull;
    if (om) {
        for (int k = 0; k < cnt; k++) {
            const int64_t s = row_addr(i, k, npad, lg);
            const uint32_t e = nbr[s];
            const uint32_t tp = tag[entry_idx(e)];
            for (int ko = 0; ko < ocnt; ko++) {
                if (!((om >> ko) & 1ull)) continue;
                const int64_t so = row_addr(oi, ko, npad_old, lg_old);
                const uint32_t eo = nbr_old[so];
                if (tag_old[entry_idx(eo)] == tp) {
                    // by owner: only the old slot on the same side of the
                    // pair (a pair the reference lists twice, through two
                    // periodic images, has one slot on each side)
                    if (by_owner && entry_rev(eo) != entry_rev(e)) continue;
                    for (int q = 0; q < nslot; q++) hist[q * cs + s] = hist_old[q * cso + so];
                    nm |= 1ull << k;
                    break;
                }
            }
        }
    }
    cmask[i] = nm;
}

// sort_atoms_by_bin's periodic sort (neighbor.rs:686-794) rebuilds only when
// the stable sort would move a particle, i.e. when the bin keys of the
// current positions are not non-decreasing in particle order: then the
// step asks for a rebuild as a drift does
__global__ void k_sort_check(int64_t n, const double4* __restrict__ pos, GridP g,
                             const SysP* __restrict__ sysp, const uint32_t* __restrict__ psys,
                             Status* st, int step) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i + 1 >= n) return;
    auto key = [&](int64_t k) {
        const double4 p = pos[k];
        if (sysp) {
            const SysP& S = sysp[psys[k]];
            return uint64_t(S.cell_off) + cell_of(S.grid, p.x, p.y, p.z);
        }
        return uint64_t(cell_of(g, p.x, p.y, p.z));
    };
    if (key(i) > key(i + 1)) atomicMin(&st->abort_step, step);
}

// ---- shear-history remap by sorted pair keys ------------------------------
// The old list's contacts become one record per pair (from its owner's
// slot): the pair as particle indices of the NEW order, lo < hi, with the
// new index of the old owner and the old history slot. Records are bucketed
// by lo and sorted by hi inside each bucket, so (lo, hi) -- the 64-bit key
// lo << 32 | hi -- is sorted over the whole table; every new row slot then
// finds its pair's record by a binary search of bucket lo.
struct HistRec {
    uint32_t hi, owner, src, pad;
};
constexpr uint32_t kNoIndex = 0xFFFFFFFFu;

// new index of every old particle (kNoIndex for a removed one)
__global__ void k_hk_invert(int64_t n, const uint32_t* __restrict__ order,
                            uint32_t* __restrict__ new_of_old) {
    const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (k < n) new_of_old[order[k]] = uint32_t(k);
}

// owned contact slots of every old particle
__global__ void k_hk_count(int64_t n_old, int64_t npad_old, int maxnb_old, int lg_old,
                           const uint32_t* __restrict__ nbr_old,
                           const uint32_t* __restrict__ nnb_old,
                           const uint64_t* __restrict__ cmask_old, uint32_t* __restrict__ cnt) {
    const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (p >= n_old) return;
    uint64_t m = cmask_old[p];
    int len = int(nnb_old[p]);
    if (len > maxnb_old) len = maxnb_old;
    uint32_t c = 0;
    while (m) {
        const int k = __builtin_ctzll(m);
        m &= m - 1ull;
        if (k >= len) break;
        if (!entry_rev(nbr_old[row_addr(p, k, npad_old, lg_old)])) c++;
    }
    cnt[p] = c;
}

// the records, at their old particle's offset; records of pairs with a
// removed particle are marked with lo = kNoIndex. More records than the
// capacity: status bit 16, the total in Status::hist_records.
__global__ void k_hk_emit(int64_t n_old, int64_t npad_old, int maxnb_old, int lg_old,
                          const uint32_t* __restrict__ nbr_old,
                          const uint32_t* __restrict__ nnb_old,
                          const uint64_t* __restrict__ cmask_old,
                          const uint32_t* __restrict__ off, const uint32_t* __restrict__ new_of_old,
                          uint32_t cap, uint32_t* __restrict__ lo_out,
                          HistRec* __restrict__ rec_out, Status* st) {
    const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (p == 0) {
        const uint32_t total = off[n_old];
        st->hist_records = total;
        if (total > cap) atomicOr(&st->overflow, 16u);
    }
    if (p >= n_old) return;
    uint64_t m = cmask_old[p];
    int len = int(nnb_old[p]);
    if (len > maxnb_old) len = maxnb_old;
    uint32_t r = off[p];
    const uint32_t ni = new_of_old[p];
    while (m) {
        const int k = __builtin_ctzll(m);
        m &= m - 1ull;
        if (k >= len) break;
        const int64_t so = row_addr(p, k, npad_old, lg_old);
        const uint32_t e = nbr_old[so];
        if (entry_rev(e)) continue;
        if (r < cap) {
            const uint32_t nj = new_of_old[entry_idx(e)];
            const bool gone = ni == kNoIndex || nj == kNoIndex;
            lo_out[r] = gone ? kNoIndex : (ni < nj ? ni : nj);
            HistRec h;
            h.hi = ni < nj ? nj : ni;
            h.owner = ni;
            h.src = uint32_t(so);
            h.pad = 0u;
            rec_out[r] = h;
        }
        r++;
    }
}

// records per bucket lo
__global__ void k_hk_bucket_count(const uint32_t* __restrict__ total_dev, uint32_t cap,
                                  const uint32_t* __restrict__ lo, uint32_t* __restrict__ cnt) {
    const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    uint32_t tot = *total_dev;
    if (tot > cap) tot = cap;
    if (r >= int64_t(tot)) return;
    const uint32_t b = lo[r];
    if (b != kNoIndex) atomicAdd(&cnt[b], 1u);
}

// records into their buckets (any order inside a bucket: the sort follows)
__global__ void k_hk_bucket_fill(const uint32_t* __restrict__ total_dev, uint32_t cap,
                                 const uint32_t* __restrict__ lo,
                                 const HistRec* __restrict__ rec,
                                 const uint32_t* __restrict__ bstart, uint32_t* __restrict__ cursor,
                                 HistRec* __restrict__ out) {
    const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    uint32_t tot = *total_dev;
    if (tot > cap) tot = cap;
    if (r >= int64_t(tot)) return;
    const uint32_t b = lo[r];
    if (b == kNoIndex) return;
    out[bstart[b] + atomicAdd(&cursor[b], 1u)] = rec[r];
}

// each bucket sorted by hi, then by old slot (insertion sort: a bucket
// holds the few contacts of one particle). The old slot orders the records
// of a pair listed twice (through two periodic images) whatever order the
// fill's atomics left them in.
__device__ inline bool rec_before(const HistRec& a, const HistRec& b) {
    return a.hi < b.hi || (a.hi == b.hi && a.src < b.src);
}

__global__ void k_hk_bucket_sort(int64_t n, const uint32_t* __restrict__ bstart,
                                 HistRec* __restrict__ rec) {
    const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (b >= n) return;
    const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
    for (uint32_t a = s0 + 1; a < s1; a++) {
        const HistRec x = rec[a];
        uint32_t t = a;
        while (t > s0 && rec_before(x, rec[t - 1])) {
            rec[t] = rec[t - 1];
            t--;
        }
        rec[t] = x;
    }
}

// every new row slot looks its pair (lo, hi) up: a binary search over hi in
// bucket lo. A slot takes the record's history when the record exists and,
// with by_owner (the reference's per-owner store), when the pair still has
// the same owner; its contact bit is set.
__global__ void k_remap_sorted(int64_t n, int64_t npad, int maxnb, int lg,
                               const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ nnb,
                               const uint32_t* __restrict__ bstart,
                               const HistRec* __restrict__ rec, int64_t cso,
                               const double* __restrict__ hist_old, double* __restrict__ hist,
                               uint64_t* __restrict__ cmask, int nslot, int by_owner) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int64_t cs = int64_t(maxnb) * npad;
    int cnt = int(nnb[i]);
    if (cnt > maxnb) cnt = maxnb;
    uint64_t nm = 0ull;
    for (int k = 0; k < cnt; k++) {
        const int64_t s = row_addr(i, k, npad, lg);
        const uint32_t e = nbr[s];
        const uint32_t j = entry_idx(e);
        const uint32_t lo = uint32_t(i) < j ? uint32_t(i) : j;
        const uint32_t hi = uint32_t(i) < j ? j : uint32_t(i);
        uint32_t a = bstart[lo], b = bstart[lo + 1];
        while (a < b) {  // first record with rec.hi >= hi
            const uint32_t mid = (a + b) >> 1;
            if (rec[mid].hi < hi)
                a = mid + 1;
            else
                b = mid;
        }
        // the first record of the pair (in old-slot order) whose owner is
        // still the owner, with by_owner; else the first record
        const uint32_t owner = entry_rev(e) ? j : uint32_t(i);
        const uint32_t end = bstart[lo + 1];
        for (; a < end && rec[a].hi == hi; a++)
            if (!by_owner || rec[a].owner == owner) break;
        if (a >= end || rec[a].hi != hi) continue;
        const HistRec h = rec[a];
        for (int q = 0; q < nslot; q++) hist[q * cs + s] = hist_old[q * cso + h.src];
        nm |= 1ull << k;
    }
    cmask[i] = nm;
}

}  // namespace

// neighbor_setup + compute_bin_grid (neighbor.rs:480-541, 290-371) on the host
void neighbor_setup(Ctx& c);

namespace {

// End of a build attempt: row maxima from the per-block slots into the
// status, and a build that did not fit stops the step sequence at `step`.
__global__ void __launch_bounds__(256) k_rows_gate(int nblk, const uint32_t* __restrict__ blk,
                                                   int maxnb, Status* st, int step) {
    __shared__ uint32_t lds[2][4];
    uint32_t mc = 0, mh = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        mc = blk[2 * b] > mc ? blk[2 * b] : mc;
        mh = blk[2 * b + 1] > mh ? blk[2 * b + 1] : mh;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(mc, o, 64), y = __shfl_xor(mh, o, 64);
        mc = x > mc ? x : mc;
        mh = y > mh ? y : mh;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[0][w] = mc;
        lds[1][w] = mh;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < int(blockDim.x >> 6); k++) {
            mc = lds[0][k] > mc ? lds[0][k] : mc;
            mh = lds[1][k] > mh ? lds[1][k] : mh;
        }
        st->max_row = int(mc);
        st->max_half = int(mh);
        uint32_t ov = st->overflow;
        if (mc > uint32_t(maxnb)) {
            ov |= 2u;
            st->overflow = ov;
        }
        // only the bits that make the host repeat the build (0-4) stop the step
        if ((ov & 31u) != 0u && step < st->abort_step) st->abort_step = step;
    }
}


void swap_alternates(Ctx& c);

// Gather every per-particle array through `order` (n entries) into the
// alternates, record the build positions, then swap: the arrays are in the
// new order and the alternates keep the previous one (tag2: the previous
// tags, the source of the shear-history remap).
void permute_all(Ctx& c, const uint32_t* order, int64_t n) {
    const bool dem = c.model == 1;
    const bool mixed = c.precision == 1;
    PermArgs a;
    std::memset(&a, 0, sizeof(a));  // padding too: the bytes are a launch signature
    a.n = n;
    a.npad = c.npad;
    a.order = order;
    a.pos = c.pos.p;
    a.vel = c.vel.p;
    a.omega = c.omega.p;
    a.quat = c.quat.p;
    a.pos2 = c.pos2.p;
    a.vel2 = c.vel2.p;
    a.omega2 = c.omega2.p;
    a.quat2 = c.quat2.p;
    a.build_pos = c.build_pos.p;
    a.tag = c.tag.p;
    a.type = c.type.p;
    a.key = c.key.p;
    a.tag2 = c.tag2.p;
    a.type2 = c.type2.p;
    a.key2 = c.key2.p;
    a.image = c.image.p;
    a.image2 = c.image2.p;
    a.mass = c.mass.p;
    a.cutr = c.cutr.p;
    a.mass2 = c.mass2.p;
    a.cutr2 = c.cutr2.p;
    if (mixed) {
        a.f32 = c.f32.p;
        a.f32b = c.f32b.p;
        a.t32 = dem ? c.t32.p : nullptr;
        a.t32b = dem ? c.t32b.p : nullptr;
    } else {
        a.f64 = c.f64.p;
        a.f64b = c.f64b.p;
        a.t64 = dem ? c.t64.p : nullptr;
        a.t64b = dem ? c.t64b.p : nullptr;
    }
    a.dem = dem;
    // the LJ mixed pass's fixed-point positions come with the permutation
    const bool xq = c.xq_usable();
    if (xq) {
        c.xq.resize(size_t(c.npad));
        c.bxq.resize(size_t(c.npad));
        a.xq = c.xq.p;
        a.bxq = c.bxq.p;
        for (int d = 0; d < 3; d++) a.qs[d] = 4294967296.0 / c.box.L[d];
    }
    launch(c, k_permute, blocks_for(n, 256), 256, a);
    MDD_HIP(hipGetLastError());
    swap_alternates(c);
    c.xq_valid = xq;
    c.bxq_valid = xq;
}

// exchange every per-particle array with its alternate
void swap_alternates(Ctx& c) {
    const bool dem = c.model == 1;
    const bool mixed = c.precision == 1;
    c.pos.swap(c.pos2);
    c.vel.swap(c.vel2);
    c.key.swap(c.key2);
    c.tag.swap(c.tag2);  // tag2 keeps the pre-sort tags for the history remap
    c.type.swap(c.type2);
    c.image.swap(c.image2);
    c.mass.swap(c.mass2);
    c.cutr.swap(c.cutr2);
    if (mixed) {
        c.f32.swap(c.f32b);
        if (dem) c.t32.swap(c.t32b);
    } else {
        c.f64.swap(c.f64b);
        if (dem) c.t64.swap(c.t64b);
    }
    if (dem) {
        c.omega.swap(c.omega2);
        c.quat.swap(c.quat2);
    }
}

// pbc, spatial sort of the particles, and the previous list kept as the
// source of the shear-history remap
void rebuild_prepare(Ctx& c) {
    if (!c.setup_done) neighbor_setup(c);
    c.xq_valid = false;  // the wrap moves particles
    const int64_t n = c.n;
    const int64_t ncells = c.grid.ncells;
    const bool dem = c.model == 1;
    const bool mixed = c.precision == 1;
    c.key.resize(n);
    c.key2.resize(n);
    c.order.resize(n);
    c.cnt.resize(ncells);
    c.start.resize(ncells + 1);
    c.bin_start.resize(ncells + 1);
    c.gstart.resize(ncells + 1);
    c.gcnt.resize(n);
    c.goff.resize(n + 1);
    c.placed.resize(std::max<int64_t>(n, c.ghost_cap) + 1);
    // scan scratch sized once for every scan of this rebuild, so no buffer is
    // reallocated while kernels that use it are queued
    c.scan_tmp.resize(size_t((std::max<int64_t>(ncells, n) + 2047) / 2048 + 2));

    launch(c, k_rebuild_begin, blocks_for(ncells, 256) < 512 ? blocks_for(ncells, 256) : 512, 256,
           c.status.p, c.cnt.p, ncells);
    // pbc + keys
    launch(c, k_wrap_key, blocks_for(n, 256), 256, n, c.pos.p, c.image.p, c.box, c.grid,
                                                          c.key.p, c.status.p, c.sysp_or_null(),
                                                          c.psys.p, c.cnt.p);
    // sort_atoms_by_bin: stable counting sort of particles by bin
    counting_sort(c, c.key.p, n, nullptr, ncells, c.cnt.p, c.start.p, c.placed.p, c.order.p,
                  nullptr, true, nullptr, nullptr, true);
    // permute every particle array into the alternates, then swap
    permute_all(c, c.order.p, n);
    // the previous list becomes the remap source
    c.remap_have_old = c.list_valid && dem;
    c.n_list_old = n;
    c.nbr.swap(c.nbr_old);
    c.nnb.swap(c.nnb_old);
    c.hist.swap(c.hist_old);
    c.cmask.swap(c.cmask_old);
    c.npad_old = c.npad;
    c.maxnb_old = c.maxnb;
    c.lg_old = c.list_lg;
}

// one build of images, bins and rows with the current capacities; the
// status says afterwards whether everything fitted
void rebuild_attempt(Ctx& c) {
    const int64_t n = c.n;
    const int64_t ncells = c.grid.ncells;
    const bool dem = c.model == 1;
    if (c.ghost_cap < 1024) c.ghost_cap = 1024;
    c.ghost_root.resize(c.ghost_cap);
    c.ghost_code.resize(c.ghost_cap);
    c.ghost_pos.resize(c.ghost_cap);
    c.ghost_key.resize(c.ghost_cap + 1);  // last word: live image count
    c.ghost_order.resize(c.ghost_cap);
    c.placed.resize(std::max<int64_t>(n, c.ghost_cap) + 1);
    const int64_t nall_max = n + c.ghost_cap;
    c.bin_pos.resize(nall_max);
    c.bin_entry.resize(nall_max);
    c.bin_gidx.resize(nall_max);
    c.slot_of.resize(n);
    // sparse grids (fewer particles than cells): the per-member build kernel
    const bool sparse = c.nsys == 0 && !c.want_tile() &&
                        (c.build_mode == 2 || (c.build_mode == 0 && n < ncells));
    if (sparse) c.bin_cell.resize(nall_max);
    const int G = 1 << c.lg;
    if (c.maxnb < G) c.maxnb = G;
    c.maxnb = (c.maxnb + G - 1) / G * G;
    if (int64_t(c.maxnb) * c.npad >= (int64_t(1) << 31) || c.npad >= (int64_t(1) << 24))
        throw ApiError(-5, "neighbor list too large for 32-bit row indexing");
    c.nbr.resize(size_t(c.maxnb) * c.npad);
    // LJ: rows as 16-bit indices into each cell's candidate list (tile force kernel)
    const bool tile = c.want_tile();
    c.tile_cap = (c.maxnb + 15 + 7) / 8 * 8;  // >= maxnb + 15: rows of any G <= 16 fit
    if (tile) {
        c.crow.resize(size_t(c.tile_cap) * size_t(n));
        c.xnbr.resize(size_t(c.maxnb) * size_t(n));
        c.nx.resize(size_t(n));
    }
    c.gnbr.resize(size_t(c.ghost_cap) * c.maxnb);
    c.gcnt_img.resize(c.ghost_cap);
    c.nnb.resize(c.npad);
    c.nmain.resize(c.npad);
    c.nhalf.resize(c.npad);
    if (dem) {
        if (c.maxnb > 64)
            throw ApiError(-5, "DEM neighbor rows above 64 entries are not supported");
        c.hist.resize(size_t(c.hist_slots) * c.maxnb * c.npad);
        c.cmask.resize(c.npad);
    }
    uint32_t* ng_live = c.ghost_key.p + c.ghost_cap;

    // borders: periodic images and their bins
    launch(c, k_ghost_count, blocks_for(n, 256), 256, n, c.pos.p, c.box, c.ghost_cutoff,
                                                             c.gcnt.p, c.sysp_or_null(), c.psys.p,
                                                             c.status.p, c.cnt.p, ncells);
    exclusive_scan(c, c.gcnt.p, c.goff.p, n);
    launch(c, k_ghost_emit, blocks_for(n, 256), 256,
        n, c.pos.p, c.box, c.ghost_cutoff, c.grid, c.goff.p, c.ghost_cap, c.ghost_root.p,
        c.ghost_code.p, c.ghost_pos.p, c.ghost_key.p, ng_live, c.status.p, c.sysp_or_null(),
        c.psys.p, c.cnt.p);
    // images binned after the particles of each cell (bin_neighbor_list
    // steps 1-2b: counting sort in index order, images have the larger indices)
    // bin_start = particle start + image start of every cell, from the same scan
    counting_sort(c, c.ghost_key.p, c.ghost_cap, ng_live, ncells, c.cnt.p, c.gstart.p, c.placed.p,
                  c.ghost_order.p, nullptr, true, c.start.p, c.bin_start.p, true);
    launch(c, k_bin_fill, blocks_for(nall_max, 256), 256,
        n, c.ghost_cap, ng_live, c.key.p, c.start.p, c.gstart.p, c.bin_start.p, c.ghost_order.p,
        c.ghost_key.p, c.pos.p, c.cutr.p, c.ghost_pos.p, c.ghost_root.p, c.ghost_code.p,
        c.bin_pos.p, c.bin_entry.p, c.bin_gidx.p, c.slot_of.p, sparse ? c.bin_cell.p : nullptr);
    // rows: one wavefront per cell (or one lane per member on sparse grids),
    // then image-owned pairs appended
    CellArgs ca;
    std::memset(&ca, 0, sizeof(ca));
    ca.ncells = ncells;
    ca.npad = c.npad;
    ca.maxnb = c.maxnb;
    ca.lg = c.lg;
    ca.bin_pos = c.bin_pos.p;
    ca.bin_entry = c.bin_entry.p;
    ca.bin_gidx = c.bin_gidx.p;
    ca.bin_start = c.bin_start.p;
    ca.run_first = c.run_first.p;
    ca.run_last = c.run_last.p;
    ca.nruns = c.nruns;
    ca.run0 = c.run0;
    ca.cutoff_sq = c.cutoff_sq;
    ca.sf2 = c.sf2;
    ca.nbr = c.nbr.p;
    ca.nnb = c.nmain.p;
    ca.nhalf = c.nhalf.p;
    ca.gnbr = c.gnbr.p;
    ca.gcnt = c.gcnt_img.p;
    ca.sysp = c.sysp_or_null();
    ca.sys_cell_off = c.sys_cell_off.p;
    ca.nsys = c.nsys;
    ca.st = c.status.p;
    ca.crow = tile ? c.crow.p : nullptr;
    ca.lstart = c.start.p;
    ca.tile_cap = c.tile_cap;
    const int cblocks = blocks_for(ncells, kCellWaves);
    if (sparse && c.uniform)
        launch(c, k_build_sparse<true>, blocks_for(nall_max, 256), 256, ca,
               static_cast<const uint32_t*>(c.bin_cell.p), static_cast<const uint32_t*>(ng_live),
               n);
    else if (sparse)
        launch(c, k_build_sparse<false>, blocks_for(nall_max, 256), 256, ca,
               static_cast<const uint32_t*>(c.bin_cell.p), static_cast<const uint32_t*>(ng_live),
               n);
    else if (c.nsys > 0)
        launch(c, k_build_cells<false, true, false>, cblocks, 256, ca);
    else if (tile && c.uniform)
        launch(c, k_build_cells<true, false, true>, cblocks, 256, ca);
    else if (tile)
        launch(c, k_build_cells<false, false, true>, cblocks, 256, ca);
    else if (c.uniform)
        launch(c, k_build_cells<true, false, false>, cblocks, 256, ca);
    else
        launch(c, k_build_cells<false, false, false>, cblocks, 256, ca);
    const int tblocks = blocks_for(n, 256);
    c.row_blk.resize(size_t(2) * tblocks);
    launch(c, k_row_totals, tblocks, 256, n, c.goff.p, ng_live, c.gcnt_img.p, c.nmain.p,
                                                c.nhalf.p, c.nnb.p, c.row_blk.p,
                                                tile ? c.nx.p : nullptr);
    c.row_nblk = tblocks;
    launch(c, k_copy_image_rows, blocks_for(c.ghost_cap * kCopyLanes, 256), 256,
        c.ghost_cap, c.npad, c.maxnb, c.lg, ng_live, c.ghost_root.p, c.goff.p, c.gcnt_img.p,
        c.nmain.p, c.gnbr.p, c.nbr.p, tile ? c.xnbr.p : nullptr);
    MDD_HIP(hipGetLastError());
    c.list_tile = tile;
    if (dem && c.remap_sorted && c.remap_have_old) {
        // the old contacts as records sorted by pair key, then every new
        // row slot looks its pair up
        const int64_t n_old = c.n_list_old;
        if (c.hk_cap < 1024) c.hk_cap = std::max<int64_t>(1024, 4 * n_old);
        c.hk_inv.resize(size_t(n_old));
        c.hk_cnt.resize(size_t(n_old) + 1);
        c.hk_off.resize(size_t(n_old) + 1);
        c.hk_lo.resize(size_t(c.hk_cap));
        c.hk_rec.resize(size_t(c.hk_cap));
        c.hk_rec2.resize(size_t(c.hk_cap));
        c.hk_bcnt.resize(size_t(n) + 1);
        c.hk_bstart.resize(size_t(n) + 1);
        HistRec* rec = reinterpret_cast<HistRec*>(c.hk_rec.p);
        HistRec* rec2 = reinterpret_cast<HistRec*>(c.hk_rec2.p);
        const uint32_t cap = uint32_t(c.hk_cap);
        const uint32_t* total = c.hk_off.p + n_old;
        fill_u32(c, c.hk_inv.p, kNoIndex, n_old);
        launch(c, k_hk_invert, blocks_for(n, 256), 256, n, static_cast<const uint32_t*>(c.order.p),
               c.hk_inv.p);
        launch(c, k_hk_count, blocks_for(n_old, 256), 256, n_old, c.npad_old, c.maxnb_old,
               c.lg_old, static_cast<const uint32_t*>(c.nbr_old.p),
               static_cast<const uint32_t*>(c.nnb_old.p),
               static_cast<const uint64_t*>(c.cmask_old.p), c.hk_cnt.p);
        exclusive_scan(c, c.hk_cnt.p, c.hk_off.p, n_old);
        launch(c, k_hk_emit, blocks_for(n_old, 256), 256, n_old, c.npad_old, c.maxnb_old,
               c.lg_old, static_cast<const uint32_t*>(c.nbr_old.p),
               static_cast<const uint32_t*>(c.nnb_old.p),
               static_cast<const uint64_t*>(c.cmask_old.p),
               static_cast<const uint32_t*>(c.hk_off.p), static_cast<const uint32_t*>(c.hk_inv.p),
               cap, c.hk_lo.p, rec, c.status.p);
        fill_u32(c, c.hk_bcnt.p, 0u, n + 1);
        launch(c, k_hk_bucket_count, blocks_for(c.hk_cap, 256), 256, total, cap,
               static_cast<const uint32_t*>(c.hk_lo.p), c.hk_bcnt.p);
        exclusive_scan(c, c.hk_bcnt.p, c.hk_bstart.p, n);
        fill_u32(c, c.hk_bcnt.p, 0u, n);
        launch(c, k_hk_bucket_fill, blocks_for(c.hk_cap, 256), 256, total, cap,
               static_cast<const uint32_t*>(c.hk_lo.p), static_cast<const HistRec*>(rec),
               static_cast<const uint32_t*>(c.hk_bstart.p), c.hk_bcnt.p, rec2);
        launch(c, k_hk_bucket_sort, blocks_for(n, 256), 256, n,
               static_cast<const uint32_t*>(c.hk_bstart.p), rec2);
        launch(c, k_remap_sorted, blocks_for(n, 256), 256, n, c.npad, c.maxnb, c.lg,
               static_cast<const uint32_t*>(c.nbr.p), static_cast<const uint32_t*>(c.nnb.p),
               static_cast<const uint32_t*>(c.hk_bstart.p), static_cast<const HistRec*>(rec2),
               int64_t(c.maxnb_old) * c.npad_old, static_cast<const double*>(c.hist_old.p),
               c.hist.p, c.cmask.p, c.hist_slots, c.hist_by_owner ? 1 : 0);
    } else if (dem) {
        launch(c, k_remap_history, blocks_for(n, 256), 256,
            n, c.npad, c.maxnb, c.lg, c.nbr.p, c.nnb.p, c.tag.p, c.order.p, c.npad_old,
            c.maxnb_old, c.lg_old, c.nbr_old.p, c.nnb_old.p, c.tag2.p, c.hist_old.p,
            c.cmask_old.p, c.hist.p, c.cmask.p, c.remap_have_old ? 1 : 0, c.hist_slots,
            c.hist_by_owner ? 1 : 0);
    }
    if (dem) {
        c.listed.resize(size_t(c.npad));
        c.listed_n.resize(1);
        fill_u32(c, c.listed_n.p, 0u, 1);
        launch(c, k_list_rows, blocks_for(n, 256 * kListPer), 256, n, c.nnb.p, c.listed.p,
               c.listed_n.p);
        launch(c, k_note_listed, 1, 1, static_cast<const uint32_t*>(c.listed_n.p), c.status.p);
        MDD_HIP(hipGetLastError());
    }
    c.tile_ok = false;
    if (dem && c.nsys == 0 && c.tile_order && n > 0) {
        const int64_t nseg = int64_t(c.grid.n[0]) * ((c.grid.n[1] + kTileCols - 1) / kTileCols);
        c.tile_list.resize(size_t(n));
        c.tile_cnt.resize(size_t(nseg) + 1);
        c.tile_off.resize(size_t(nseg) + 1);
        launch(c, k_tile_count, blocks_for(nseg, 256), 256, c.grid,
               static_cast<const uint32_t*>(c.start.p), c.tile_cnt.p, nseg);
        exclusive_scan(c, c.tile_cnt.p, c.tile_off.p, nseg);
        launch(c, k_tile_emit, int(nseg), 256, c.grid, static_cast<const uint32_t*>(c.start.p),
               static_cast<const uint32_t*>(c.tile_off.p), c.tile_list.p);
        MDD_HIP(hipGetLastError());
        c.tile_ok = true;
    }
    c.listed_ok = dem;
    c.list_valid = true;
    c.list_injected = false;
    c.list_lg = c.lg;
    c.npairs = -1;  // computed on request
}

}  // namespace

// Capacities for the next attempt after a build that did not fit; false when
// the status reports a complete build.
bool rebuild_grow(Ctx& c, const Status& s) {
    bool again = false;
    if (s.overflow & 1u) {
        c.ghost_cap = int64_t(s.ghost_total) + int64_t(s.ghost_total) / 4 + 1024;
        again = true;
    }
    if (s.overflow & 2u) {
        c.maxnb = s.max_row + s.max_row / 8 + 8;
        if (c.model == 1 && c.maxnb > 64) c.maxnb = s.max_row <= 64 ? 64 : s.max_row;
        again = true;
    }
    if (s.overflow & 8u) {
        c.tile_blocked = true;  // a candidate list beyond 16-bit indices
        again = true;
    }
    if (s.overflow & 16u) {
        c.hk_cap = int64_t(s.hist_records) + int64_t(s.hist_records) / 4 + 1024;
        again = true;
    }
    return again;
}

namespace {
__global__ void k_clear_exit(Status* st) { st->overflow &= ~4u; }
}  // namespace

// pbc's removal of the particles outside a non-periodic face (domain.rs:
// 413-459), after a rebuild's wrap and sort saw them (status bit 4). The
// reference walks the particles from the last index down and swap_removes
// each one outside (the last particle takes its slot), then sorts by bin
// (sort_atoms_by_bin, stable here); that order is rebuilt from the sort's
// permutation on the host, and the survivors are gathered from the
// pre-sort arrays in it. The history remap then reads the pre-sort rows
// through the new order as after any sort. A rare event: done
// synchronously, the build then repeats on the remaining particles.
void remove_outside(Ctx& c) {
    if (c.nsys > 0)
        throw ApiError(-5, "a particle of a batched system left a non-periodic boundary");
    MDD_HIP(hipStreamSynchronize(c.stream));
    const int64_t n = c.n;
    const int64_t ncells = c.grid.ncells;
    const size_t un = static_cast<size_t>(n);
    std::vector<double4> pos(un);
    std::vector<uint32_t> key(un), order(un);
    d2h(c, pos.data(), c.pos.p, sizeof(double4) * size_t(n));
    d2h(c, key.data(), c.key.p, sizeof(uint32_t) * size_t(n));
    d2h(c, order.data(), c.order.p, sizeof(uint32_t) * size_t(n));
    std::vector<uint8_t> out(un, 0);
    std::vector<uint32_t> key_pre(un);
    for (int64_t k = 0; k < n; k++) {
        const double x[3] = {pos[k].x, pos[k].y, pos[k].z};
        bool o = false;
        for (int d = 0; d < 3; d++)
            if (!c.box.periodic[d] && (x[d] < c.box.lo[d] || x[d] >= c.box.hi[d])) o = true;
        out[order[k]] = o ? 1 : 0;
        key_pre[order[k]] = key[k];
    }
    // swap_remove from the last index down (the particle at the end has
    // already been looked at and stays)
    std::vector<uint32_t> cur(un);
    for (int64_t i = 0; i < n; i++) cur[i] = uint32_t(i);
    int64_t len = n;
    for (int64_t i = n - 1; i >= 0; i--)
        if (out[cur[i]]) cur[i] = cur[--len];
    if (len == 0) throw ApiError(-1, "every particle left the box through a fixed boundary");
    // stable counting sort of the survivors by bin
    std::vector<uint32_t> start(static_cast<size_t>(ncells) + 1, 0u);
    for (int64_t j = 0; j < len; j++) start[key_pre[cur[j]] + 1]++;
    for (int64_t q = 0; q < ncells; q++) start[q + 1] += start[q];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    std::vector<uint32_t> fin(static_cast<size_t>(len));
    for (int64_t j = 0; j < len; j++) fin[fill[key_pre[cur[j]]]++] = cur[j];
    // the pre-sort arrays back in place, then the survivors gathered
    swap_alternates(c);
    h2d(c, c.order.p, fin.data(), sizeof(uint32_t) * size_t(len));
    h2d(c, c.start.p, start.data(), sizeof(uint32_t) * (size_t(ncells) + 1));
    permute_all(c, c.order.p, len);
    k_clear_exit<<<1, 1, 0, c.stream>>>(c.status.p);
    MDD_HIP(hipGetLastError());
    MDD_HIP(hipStreamSynchronize(c.stream));
    c.n = len;
    c.removed += n - len;
}

// full rebuild, checked on the host before it returns
void rebuild(Ctx& c) {
    rebuild_prepare(c);
    bool fitted = false;
    for (int attempt = 0; attempt < 6 && !fitted; attempt++) {
        rebuild_attempt(c);
        k_rows_gate<<<1, 256, 0, c.stream>>>(c.row_nblk, c.row_blk.p, c.maxnb, c.status.p,
                                             INT_MAX);
        MDD_HIP(hipMemcpyAsync(c.status_host, c.status.p, sizeof(Status), hipMemcpyDeviceToHost,
                               c.stream));
        MDD_HIP(hipStreamSynchronize(c.stream));
        if (c.status_host->overflow & 4u) {
            remove_outside(c);
            attempt--;
            continue;
        }
        fitted = !rebuild_grow(c, *c.status_host);
    }
    if (!fitted) throw ApiError(-4, "neighbor rebuild did not fit after retries");
    c.nghost = c.status_host->ghost_total;
    c.rebuilds++;
}

// Rebuild inside a step sequence without waiting for the device: if the
// build does not fit, the step's later kernels are gated off (as after a
// rebuild request) and the host repeats the build with larger capacities,
// starting from the same sorted state (retry = true), when it next reads
// the status.
namespace {

// Issue the recorded launches: replay the captured graph whose signature is
// the same, or capture them as a new graph (at most kGraphs kept; the two
// alternating buffer sets of successive rebuilds give two signatures).
constexpr size_t kGraphs = 4;

void issue_recorded(Ctx& c) {
    for (size_t k = 0; k < c.graphs.size(); k++) {
        if (c.graphs[k].sig == c.rec.sig) {
            MDD_HIP(hipGraphLaunch(c.graphs[k].exec, c.stream));
            c.graph_hits++;
            std::rotate(c.graphs.begin() + long(k), c.graphs.begin() + long(k) + 1,
                        c.graphs.end());
            return;
        }
    }
    // capture; if the runtime refuses it, the launches go out one by one and
    // this context stops recording (the results are the same either way)
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    bool captured = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (captured) {
        try {
            for (auto& f : c.rec.calls) f(c.stream);
        } catch (const HipError&) {
            captured = false;
        }
        const hipError_t ee = hipStreamEndCapture(c.stream, &g);
        captured = captured && ee == hipSuccess && g != nullptr;
        if (captured) captured = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();  // a refused capture leaves its error behind
    if (!captured) {
        c.graphs_on = false;
        for (auto& f : c.rec.calls) f(c.stream);
        return;
    }
    MDD_HIP(hipGraphLaunch(ex, c.stream));
    c.graph_captures++;
    if (c.graphs.size() >= kGraphs) {
        (void)hipGraphExecDestroy(c.graphs.front().exec);
        c.graphs.erase(c.graphs.begin());
    }
    c.graphs.push_back(GraphEntry{c.rec.sig, ex});
}

void end_recording(Ctx& c) {
    c.rec.on = false;
    deferred_frees() = nullptr;
}

}  // namespace

void release_graphs(Ctx& c) {
    for (auto& g : c.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c.graphs.clear();
}

// One rebuild. A first attempt (not a retry) is recorded and issued as a
// graph: its ~30 launches reach the device at once instead of one host
// launch at a time.
void launch_sort_check(Ctx& c, int step) {
    if (c.n < 2) return;
    launch(c, k_sort_check, blocks_for(c.n, 256), 256, c.n, static_cast<const double4*>(c.pos.p),
           c.grid, c.sysp_or_null(), static_cast<const uint32_t*>(c.psys.p), c.status.p, step);
    MDD_HIP(hipGetLastError());
}

void rebuild_async(Ctx& c, int step, bool retry) {
    const bool timed = c.prof_every > 0 && c.rb_pending < int64_t(c.rb_ev.size() / 3);
    if (timed) {
        MDD_HIP(hipEventRecord(c.rb_ev[3 * c.rb_pending + 1], c.stream));
        c.rb_has_gap[size_t(c.rb_pending)] = c.rb_chunk_mark ? 1 : 0;
        if (c.rb_chunk_mark) std::swap(c.rb_ev[3 * c.rb_pending], c.rb_mark);
    }
    c.rb_chunk_mark = false;
    const bool record = !retry && c.graphs_on && c.setup_done;
    if (record) {
        c.rec.sig.clear();
        c.rec.calls.clear();
        c.rec.frees.clear();
        c.rec.on = true;
        deferred_frees() = &c.rec.frees;
    }
    try {
        if (!retry) {
            rebuild_prepare(c);
            c.rebuilds++;
        }
        rebuild_attempt(c);
    } catch (...) {
        if (record) end_recording(c);
        throw;
    }
    if (record) {
        end_recording(c);
        issue_recorded(c);
        c.rec.calls.clear();
        if (!c.rec.frees.empty()) {
            // buffers replaced during the rebuild: free them once the
            // launches that still name them have run
            MDD_HIP(hipStreamSynchronize(c.stream));
            for (void* p : c.rec.frees) (void)hipFree(p);
            c.rec.frees.clear();
        }
    }
    k_rows_gate<<<1, 256, 0, c.stream>>>(c.row_nblk, c.row_blk.p, c.maxnb, c.status.p, step);
    MDD_HIP(hipGetLastError());
    if (timed) MDD_HIP(hipEventRecord(c.rb_ev[3 * c.rb_pending++ + 2], c.stream));
}

namespace {

// neighbor_setup + compute_bin_grid for one box and radius range
struct Geometry {
    GridP grid;
    double max_cutoff = 0, ghost_cutoff = 0, bin_size[3] = {1, 1, 1};
    std::vector<int> full, run_first, run_last;
    int run0 = -1, nfwd = 0;
    bool uniform = true;
    double cutoff_sq = 0, sf2 = 1, min_skin = 0;
};

Geometry make_geometry(const BoxP& box, double mn, double mx, double skin_fraction,
                       double bin_size) {
    // neighbor_setup (neighbor.rs:480-541), single process
    Geometry g;
    const double max_cutoff = mx > 0.0 ? 2.0 * mx * skin_fraction : bin_size;
    const double buffer = mn < 1.7976931348623157e308 * 0.5 ? (skin_fraction - 1.0) * mn : 0.0;
    double ghost_cut = max_cutoff + 2.0 * buffer;
    if (ghost_cut < 0.0) ghost_cut = 0.0;
    g.ghost_cutoff = ghost_cut;
    g.max_cutoff = max_cutoff;
    const double bin_min = bin_size < max_cutoff ? max_cutoff : bin_size;
    // compute_bin_grid (neighbor.rs:290-371), comm_size = 1
    int nint[3], sl[3];
    for (int d = 0; d < 3; d++) {
        double q = std::floor(box.L[d] / bin_min);
        if (q < 1.0) q = 1.0;
        nint[d] = int(q);
        g.bin_size[d] = box.L[d] / double(nint[d]);
    }
    const double gcut = g.ghost_cutoff > max_cutoff ? g.ghost_cutoff : max_cutoff;
    for (int d = 0; d < 3; d++) {
        const int range = int(std::ceil(max_cutoff / g.bin_size[d]));
        sl[d] = int(std::ceil(gcut / g.bin_size[d])) + range;
        g.grid.n[d] = nint[d] + 2 * sl[d];
        g.grid.origin[d] = box.lo[d] - g.bin_size[d] * double(sl[d]);
        g.grid.inv_bs[d] = 1.0 / g.bin_size[d];
    }
    g.grid.ncells = int64_t(g.grid.n[0]) * g.grid.n[1] * g.grid.n[2];
    if (g.grid.ncells > (int64_t(1) << 31))
        throw ApiError(-5, "bin grid too large; raise bin_size");
    const double cutoff_sq = max_cutoff * max_cutoff;
    const int ny = g.grid.n[1], nz = g.grid.n[2];
    for (int dx = -sl[0]; dx <= sl[0]; dx++)
        for (int dy = -sl[1]; dy <= sl[1]; dy++)
            for (int dz = -sl[2]; dz <= sl[2]; dz++) {
                const int ax = std::abs(dx) - 1, ay = std::abs(dy) - 1, az = std::abs(dz) - 1;
                const double mdx = double(ax > 0 ? ax : 0) * g.bin_size[0];
                const double mdy = double(ay > 0 ? ay : 0) * g.bin_size[1];
                const double mdz = double(az > 0 ? az : 0) * g.bin_size[2];
                if (mdx * mdx + mdy * mdy + mdz * mdz < cutoff_sq) {
                    const int off = dx * ny * nz + dy * nz + dz;
                    g.full.push_back(off);
                    if (off > 0) g.nfwd++;
                }
            }
    g.grid.nfwd = g.nfwd;
    // the full stencil as runs of consecutive cell offsets (contiguous bins)
    std::sort(g.full.begin(), g.full.end());
    for (size_t k = 0; k < g.full.size(); k++) {
        if (k == 0 || g.full[k] != g.full[k - 1] + 1) {
            g.run_first.push_back(g.full[k]);
            g.run_last.push_back(g.full[k]);
        } else {
            g.run_last.back() = g.full[k];
        }
        if (g.full[k] == 0) g.run0 = int(g.run_first.size()) - 1;
    }
    if (g.run_first.size() > size_t(kMaxRuns) || g.run0 < 0)
        throw ApiError(-5, "bin stencil has too many cell runs; raise bin_size");
    // save_build_positions cutoff cache (neighbor.rs:545-561)
    g.min_skin = mn;
    if (std::fabs(mx - mn) < 1e-15) {
        const double cutoff = 2.0 * mn * skin_fraction;
        g.uniform = true;
        g.cutoff_sq = cutoff * cutoff;
    } else {
        g.uniform = false;
    }
    g.sf2 = skin_fraction * skin_fraction;
    return g;
}

// expected periodic images per particle: prod over periodic axes of (1 + 2 gc / L) - 1
double images_per_particle(const BoxP& box, double gc) {
    double per = 1.0;
    for (int d = 0; d < 3; d++)
        if (box.periodic[d]) per *= 1.0 + std::min(1.0, 2.0 * gc / box.L[d]);
    return per - 1.0;
}

void upload_runs(Ctx& c, const std::vector<int>& rf, const std::vector<int>& rl) {
    c.run_first.resize(rf.size());
    c.run_last.resize(rl.size());
    h2d(c, c.run_first.p, rf.data(), sizeof(int) * rf.size());
    h2d(c, c.run_last.p, rl.data(), sizeof(int) * rl.size());
}

// one geometry per system, cells and runs concatenated
void ensemble_setup(Ctx& c) {
    std::vector<int> rf, rl;
    std::vector<int64_t> cell_off(size_t(c.nsys) + 1, 0);
    double images = 0.0;
    for (int s = 0; s < c.nsys; s++) {
        SysP& S = c.sys_h[s];
        const Geometry g =
            make_geometry(S.box, c.sys_rmin[s], c.sys_rmax[s], c.skin_fraction, c.bin_min_size);
        S.grid = g.grid;
        S.cell_off = cell_off[s];
        S.gc = g.ghost_cutoff;
        S.cutoff_sq = g.cutoff_sq;
        S.sf2 = g.sf2;
        const double skin = (c.skin_fraction - 1.0) * g.min_skin;
        S.thr2 = skin * skin;
        S.uniform = g.uniform ? 1 : 0;
        S.run_off = int(rf.size());
        S.nruns = int(g.run_first.size());
        S.run0 = g.run0;
        rf.insert(rf.end(), g.run_first.begin(), g.run_first.end());
        rl.insert(rl.end(), g.run_last.begin(), g.run_last.end());
        cell_off[s + 1] = cell_off[s] + g.grid.ncells;
        images += double(c.sys_first[s + 1] - c.sys_first[s]) * images_per_particle(S.box, S.gc);
    }
    if (cell_off[c.nsys] >= (int64_t(1) << 32))
        throw ApiError(-5, "ensemble bin grids too large; raise bin_size");
    c.grid.ncells = cell_off[c.nsys];
    upload_runs(c, rf, rl);
    c.sysp.resize(c.sys_h.size());
    h2d(c, c.sysp.p, c.sys_h.data(), sizeof(SysP) * c.sys_h.size());
    c.sys_cell_off.resize(cell_off.size());
    h2d(c, c.sys_cell_off.p, cell_off.data(), sizeof(int64_t) * cell_off.size());
    if (c.ghost_cap == 0) c.ghost_cap = int64_t(images * 1.25) + 1024;
    c.setup_done = true;
}

}  // namespace

void neighbor_setup(Ctx& c) {
    if (c.n <= 0) throw ApiError(-1, "neighbor_setup needs particles");
    if (c.nsys > 0) {
        ensemble_setup(c);
        return;
    }
    std::vector<double> cr(c.n);
    d2h(c, cr.data(), c.cutr.p, sizeof(double) * c.n);
    double mx = 0.0, mn = 1.7976931348623157e308;
    for (double v : cr) {
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    const Geometry g = make_geometry(c.box, mn, mx, c.skin_fraction, c.bin_min_size);
    c.grid = g.grid;
    c.max_cutoff = g.max_cutoff;
    c.ghost_cutoff = g.ghost_cutoff;
    for (int d = 0; d < 3; d++) c.bin_size[d] = g.bin_size[d];
    c.full = g.full;
    c.nruns = int(g.run_first.size());
    c.run0 = g.run0;
    upload_runs(c, g.run_first, g.run_last);
    c.min_skin = g.min_skin;
    c.uniform = g.uniform;
    c.cutoff_sq = g.cutoff_sq;
    c.sf2 = g.sf2;
    if (c.ghost_cap == 0)
        c.ghost_cap = int64_t(double(c.n) * images_per_particle(c.box, c.ghost_cutoff) * 1.25) + 1024;
    c.setup_done = true;
}

}  // namespace mdd
