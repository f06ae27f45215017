// Neighbor rebuild for the gfx950 particle path.
//
// One rebuild reproduces, on the device, the reference's FullRebuild phases
// for a single process:
//   pbc (crates/mddem_core/src/domain.rs:413-459)
//   sort_atoms_by_bin (crates/mddem_core/src/neighbor.rs:686-794), stable
//   borders self-sends (crates/mddem_core/src/comm.rs:449-570) as periodic images
//   bin_neighbor_list (crates/mddem_core/src/neighbor.rs:810-1080)
// and then derives, for every particle, a full row: the pairs it owns in the
// reference half list (same bin grid, same candidate order, same FP64 fma
// distance test) followed by the pairs that other particles own with it. A
// particle finds the latter by scanning its backward stencil cells and the
// cells around each of its periodic images, applying the owner's distance
// test bit for bit, so no atomics or transposition are needed and every row
// comes out in a deterministic order.
#include "ctx.hpp"

namespace mdd {

void counting_sort_dev(Ctx& c, const uint32_t* key, const uint32_t* n_dev, int64_t n_max,
                       int64_t ncells, uint32_t* cnt, uint32_t* start, uint32_t* order,
                       uint32_t* slot_of);

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

__global__ void k_status_reset_build(Status* st) {
    st->overflow = 0u;
    st->max_row = 0;
    st->max_half = 0;
    st->ghost_total = 0u;
}

// pbc wrap (wrap_periodic, domain.rs:396-407) + bin key of every particle
__global__ void k_wrap_key(int64_t n, double4* __restrict__ pos, int4* __restrict__ image,
                           BoxP box, GridP g, uint32_t* __restrict__ key, Status* st) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
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
    key[i] = cell_of(g, x[0], x[1], x[2]);
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

__global__ void k_ghost_count(int64_t n, const double4* __restrict__ pos, BoxP box, double gc,
                              uint32_t* __restrict__ gcnt) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    int lf[3], hf[3];
    image_flags(box, gc, x, lf, hf);
    gcnt[i] = uint32_t((1 + lf[0] + hf[0]) * (1 + lf[1] + hf[1]) * (1 + lf[2] + hf[2]) - 1);
}

__global__ void k_ghost_emit(int64_t n, const double4* __restrict__ pos, BoxP box, double gc,
                             const uint32_t* __restrict__ goff, int64_t ghost_cap,
                             uint32_t* __restrict__ groot, uint8_t* __restrict__ gcode,
                             double4* __restrict__ gpos, Status* st) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    uint32_t total = goff[n];
    if (i == 0) st->ghost_total = total;
    if (int64_t(total) > ghost_cap) {
        if (i == 0) atomicOr(&st->overflow, 1u);
        return;
    }
    if (i >= n) return;
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    int lf[3], hf[3];
    image_flags(box, gc, x, lf, hf);
    uint32_t g = goff[i];
    int sxs[3] = {0, 1, -1};
    for (int a = 0; a < 3; a++) {
        int sx = sxs[a];
        if ((sx == 1 && !lf[0]) || (sx == -1 && !hf[0])) continue;
        for (int b = 0; b < 3; b++) {
            int sy = sxs[b];
            if ((sy == 1 && !lf[1]) || (sy == -1 && !hf[1])) continue;
            for (int cidx = 0; cidx < 3; cidx++) {
                int sz = sxs[cidx];
                if ((sz == 1 && !lf[2]) || (sz == -1 && !hf[2])) continue;
                if (sx == 0 && sy == 0 && sz == 0) continue;
                int code = code_of(sx, sy, sz);
                groot[g] = uint32_t(i);
                gcode[g] = uint8_t(code);
                gpos[g] = make_double4(x[0] + image_shift(sx, box.L[0]),
                                       x[1] + image_shift(sy, box.L[1]),
                                       x[2] + image_shift(sz, box.L[2]), 0.0);
                g++;
            }
        }
    }
}

// bin keys of particles (already sorted) followed by their periodic images
__global__ void k_all_keys(int64_t n, int64_t ghost_cap, const uint32_t* __restrict__ key_local,
                           const double4* __restrict__ gpos, const uint32_t* __restrict__ goff,
                           GridP g, uint32_t* __restrict__ all_key, uint32_t* n_all) {
    int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    int64_t ng = goff[n];
    if (ng > ghost_cap) ng = ghost_cap;
    if (t == 0) *n_all = uint32_t(n + ng);
    if (t < n) {
        all_key[t] = key_local[t];
    } else if (t < n + ng) {
        double4 q = gpos[t - n];
        all_key[t] = cell_of(g, q.x, q.y, q.z);
    }
}

// binned copies of position / neighbor skin radius and the pre-encoded entry
__global__ void k_bin_fill(int64_t n, const uint32_t* __restrict__ n_all,
                           const uint32_t* __restrict__ order, const double4* __restrict__ pos,
                           const double* __restrict__ cutr, const double4* __restrict__ gpos,
                           const uint32_t* __restrict__ groot,
                           const uint8_t* __restrict__ gcode, double4* __restrict__ bin_pos,
                           uint32_t* __restrict__ bin_entry, uint32_t* __restrict__ slot_local) {
    int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= int64_t(*n_all)) return;
    uint32_t a = order[s];
    if (a < n) {
        double4 p = pos[a];
        bin_pos[s] = make_double4(p.x, p.y, p.z, cutr[a]);
        bin_entry[s] = enc_entry(a, kCodeNone, false);
        slot_local[a] = uint32_t(s);
    } else {
        uint32_t gi = a - uint32_t(n);
        double4 q = gpos[gi];
        uint32_t r = groot[gi];
        bin_pos[s] = make_double4(q.x, q.y, q.z, cutr[r]);
        bin_entry[s] = enc_entry(r, gcode[gi], false);
    }
}

struct ListArgs {
    int64_t n, npad;
    int maxnb;
    const double4* bin_pos;
    const uint32_t* bin_entry;
    const uint32_t* bin_start;
    const uint32_t* slot_local;
    const uint32_t* key;
    const int* fwd;
    int nfwd;
    BoxP box;
    GridP grid;
    double gc;
    int uniform;
    double cutoff_sq, sf2;
    uint32_t* nbr;
    uint32_t* nnb;
    uint32_t* nhalf;
    Status* st;
};

// reference pair test, evaluated from the owner's side: r2 of
// (partner image - owner) with the fma chain of neighbor.rs:955-958 and
// either the uniform cutoff or (s_i + s_j)^2 sf^2 (neighbor.rs:1003-1004)
__device__ inline bool pair_in(const ListArgs& a, double dx, double dy, double dz,
                               double s_owner, double s_partner) {
    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    if (a.uniform) return r2 < a.cutoff_sq;
    double sum = s_owner + s_partner;
    return r2 < sum * sum * a.sf2;
}

__global__ void __launch_bounds__(256) k_build_list(ListArgs a) {
    int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (p >= a.n) return;
    const uint32_t slot = a.slot_local[p];
    const uint32_t c = a.key[p];
    const double4 xp = a.bin_pos[slot];
    const double sp = xp.w;
    int cnt = 0;
    auto push = [&](uint32_t e) {
        if (cnt < a.maxnb) a.nbr[int64_t(cnt) * a.npad + p] = e;
        cnt++;
    };
    // (1) pairs this particle owns: rest of its own cell, then forward cells
    {
        const uint32_t end = a.bin_start[c + 1];
        for (uint32_t m = slot + 1; m < end; m++) {
            double4 q = a.bin_pos[m];
            if (pair_in(a, q.x - xp.x, q.y - xp.y, q.z - xp.z, sp, q.w)) push(a.bin_entry[m]);
        }
        for (int k = 0; k < a.nfwd; k++) {
            const int64_t cc = int64_t(c) + a.fwd[k];
            if (cc < 0 || cc >= a.grid.ncells) continue;  // only off-grid (clamped) cells
            const uint32_t b = a.bin_start[cc], e = a.bin_start[cc + 1];
            for (uint32_t m = b; m < e; m++) {
                double4 q = a.bin_pos[m];
                if (pair_in(a, q.x - xp.x, q.y - xp.y, q.z - xp.z, sp, q.w))
                    push(a.bin_entry[m]);
            }
        }
    }
    const int nh = cnt;
    // (2) pairs owned by other particles with this particle itself: earlier
    //     slots of its cell and particles (not images) of backward cells
    {
        for (uint32_t m = a.bin_start[c]; m < slot; m++) {
            double4 q = a.bin_pos[m];
            if (pair_in(a, xp.x - q.x, xp.y - q.y, xp.z - q.z, q.w, sp))
                push(a.bin_entry[m] | kRevBit);
        }
        for (int k = 0; k < a.nfwd; k++) {
            const int64_t cc = int64_t(c) - a.fwd[k];
            if (cc < 0 || cc >= a.grid.ncells) continue;
            const uint32_t b = a.bin_start[cc], e = a.bin_start[cc + 1];
            for (uint32_t m = b; m < e; m++) {
                uint32_t ent = a.bin_entry[m];
                if (entry_code(ent) != kCodeNone) continue;
                double4 q = a.bin_pos[m];
                if (pair_in(a, xp.x - q.x, xp.y - q.y, xp.z - q.z, q.w, sp))
                    push(ent | kRevBit);
            }
        }
    }
    // (3) pairs owned by particles near one of this particle's periodic images
    {
        double x[3] = {xp.x, xp.y, xp.z};
        int lf[3], hf[3];
        image_flags(a.box, a.gc, x, lf, hf);
        if (lf[0] | hf[0] | lf[1] | hf[1] | lf[2] | hf[2]) {
            const int sv[3] = {0, 1, -1};
            for (int ia = 0; ia < 3; ia++) {
                int sx = sv[ia];
                if ((sx == 1 && !lf[0]) || (sx == -1 && !hf[0])) continue;
                for (int ib = 0; ib < 3; ib++) {
                    int sy = sv[ib];
                    if ((sy == 1 && !lf[1]) || (sy == -1 && !hf[1])) continue;
                    for (int ic = 0; ic < 3; ic++) {
                        int sz = sv[ic];
                        if ((sz == 1 && !lf[2]) || (sz == -1 && !hf[2])) continue;
                        if (sx == 0 && sy == 0 && sz == 0) continue;
                        const int code = code_of(sx, sy, sz);
                        const double gx = x[0] + image_shift(sx, a.box.L[0]);
                        const double gy = x[1] + image_shift(sy, a.box.L[1]);
                        const double gz = x[2] + image_shift(sz, a.box.L[2]);
                        const uint32_t cg = cell_of(a.grid, gx, gy, gz);
                        for (int k = -1; k < a.nfwd; k++) {
                            const int64_t cc = k < 0 ? int64_t(cg) : int64_t(cg) - a.fwd[k];
                            if (cc < 0 || cc >= a.grid.ncells) continue;
                            const uint32_t b = a.bin_start[cc], e = a.bin_start[cc + 1];
                            for (uint32_t m = b; m < e; m++) {
                                uint32_t ent = a.bin_entry[m];
                                if (entry_code(ent) != kCodeNone) continue;
                                double4 q = a.bin_pos[m];
                                if (pair_in(a, gx - q.x, gy - q.y, gz - q.z, q.w, sp))
                                    push(enc_entry(entry_idx(ent), code, true));
                            }
                        }
                    }
                }
            }
        }
    }
    a.nnb[p] = uint32_t(cnt);
    a.nhalf[p] = uint32_t(nh);
    atomicMax(&a.st->max_row, cnt);
    atomicMax(&a.st->max_half, nh);
    if (cnt > a.maxnb) atomicOr(&a.st->overflow, 2u);
}

// carry shear history across a rebuild: a new slot takes the old slot of the
// same particle whose partner has the same tag (the reference keys history
// by partner tag, tangential.rs:54-133, and the spring is stored in the
// lower tag's frame, so it transfers unchanged). Only slots that were in
// contact at the last force pass carry history (the reference prunes the
// rest, contact.rs:586-589); their contact bits move with them.
__global__ void k_remap_history(int64_t n, int64_t npad, int maxnb,
                                const uint32_t* __restrict__ nbr,
                                const uint32_t* __restrict__ nnb,
                                const uint32_t* __restrict__ tag,
                                const uint32_t* __restrict__ order, int64_t npad_old,
                                int maxnb_old, const uint32_t* __restrict__ nbr_old,
                                const uint32_t* __restrict__ nnb_old,
                                const uint32_t* __restrict__ tag_old,
                                const double* __restrict__ hist_old,
                                const uint64_t* __restrict__ cmask_old,
                                double* __restrict__ hist, uint64_t* __restrict__ cmask,
                                int have_old) {
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
            const uint32_t tp = tag[entry_idx(nbr[int64_t(k) * npad + i])];
            for (int ko = 0; ko < ocnt; ko++) {
                if (!((om >> ko) & 1ull)) continue;
                const uint32_t eo = nbr_old[int64_t(ko) * npad_old + oi];
                if (tag_old[entry_idx(eo)] == tp) {
                    const int64_t so = int64_t(ko) * npad_old + oi;
                    const int64_t s = int64_t(k) * npad + i;
                    hist[s] = hist_old[so];
                    hist[cs + s] = hist_old[cso + so];
                    hist[2 * cs + s] = hist_old[2 * cso + so];
                    nm |= 1ull << k;
                    break;
                }
            }
        }
    }
    cmask[i] = nm;
}

__global__ void k_count_dev(const uint32_t* __restrict__ key, const uint32_t* __restrict__ n_dev,
                            uint32_t* cnt) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < int64_t(*n_dev)) atomicAdd(&cnt[key[i]], 1u);
}

__global__ void k_scatter_dev(const uint32_t* __restrict__ key,
                              const uint32_t* __restrict__ n_dev,
                              const uint32_t* __restrict__ start, uint32_t* cursor,
                              uint32_t* __restrict__ order) {
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < int64_t(*n_dev)) {
        uint32_t c = key[i];
        uint32_t p = atomicAdd(&cursor[c], 1u);
        order[start[c] + p] = uint32_t(i);
    }
}

__global__ void k_segsort2(const uint32_t* __restrict__ start, int64_t ncells,
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

}  // namespace

void counting_sort_dev(Ctx& c, const uint32_t* key, const uint32_t* n_dev, int64_t n_max,
                       int64_t ncells, uint32_t* cnt, uint32_t* start, uint32_t* order,
                       uint32_t* slot_of) {
    (void)slot_of;
    fill_u32(c, cnt, 0u, ncells);
    if (n_max > 0) k_count_dev<<<blocks_for(n_max, 256), 256, 0, c.stream>>>(key, n_dev, cnt);
    exclusive_scan(c, cnt, start, ncells);
    fill_u32(c, cnt, 0u, ncells);
    if (n_max > 0) {
        k_scatter_dev<<<blocks_for(n_max, 256), 256, 0, c.stream>>>(key, n_dev, start, cnt,
                                                                   order);
        k_segsort2<<<blocks_for(ncells, 256), 256, 0, c.stream>>>(start, ncells, order);
    }
    MDD_HIP(hipGetLastError());
}

// neighbor_setup + compute_bin_grid (neighbor.rs:480-541, 290-371) on the host
void neighbor_setup(Ctx& c);

void rebuild(Ctx& c) {
    if (!c.setup_done) neighbor_setup(c);
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
    c.gcnt.resize(n);
    c.goff.resize(n + 1);

    k_status_reset_build<<<1, 1, 0, c.stream>>>(c.status.p);
    // pbc + keys
    k_wrap_key<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, c.pos.p, c.image.p, c.box, c.grid,
                                                          c.key.p, c.status.p);
    // sort_atoms_by_bin: stable counting sort of particles by bin
    counting_sort(c, c.key.p, n, ncells, c.cnt.p, c.start.p, c.order.p, nullptr);
    // permute every particle array into the alternates, then swap
    {
        PermArgs a{};
        a.n = n;
        a.npad = c.npad;
        a.order = c.order.p;
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
        k_permute<<<blocks_for(n, 256), 256, 0, c.stream>>>(a);
        MDD_HIP(hipGetLastError());
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
    // the previous list becomes the remap source
    const bool have_old = c.list_valid && dem;
    c.nbr.swap(c.nbr_old);
    c.nnb.swap(c.nnb_old);
    c.hist.swap(c.hist_old);
    c.cmask.swap(c.cmask_old);
    c.npad_old = c.npad;
    c.maxnb_old = c.maxnb;

    for (int attempt = 0; attempt < 6; attempt++) {
        if (c.ghost_cap < 1024) c.ghost_cap = 1024;
        c.ghost_root.resize(c.ghost_cap);
        c.ghost_code.resize(c.ghost_cap);
        c.ghost_pos.resize(c.ghost_cap);
        const int64_t nall_max = n + c.ghost_cap;
        c.all_key.resize(nall_max + 1);  // last word holds the live count
        c.bin_order.resize(nall_max);
        c.bin_pos.resize(nall_max);
        c.bin_entry.resize(nall_max);
        c.slot_of.resize(n);
        if (c.maxnb < 8) c.maxnb = 8;
        c.nbr.resize(size_t(c.maxnb) * c.npad);
        c.nnb.resize(c.npad);
        c.nhalf.resize(c.npad);
        if (dem) {
            if (c.maxnb > 64)
                throw ApiError(-5, "DEM neighbor rows above 64 entries are not supported");
            c.hist.resize(size_t(3) * c.maxnb * c.npad);
            c.cmask.resize(c.npad);
        }
        uint32_t* n_all = c.all_key.p + nall_max;

        k_status_reset_build<<<1, 1, 0, c.stream>>>(c.status.p);
        // borders: periodic images
        k_ghost_count<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, c.pos.p, c.box, c.ghost_cutoff,
                                                                 c.gcnt.p);
        exclusive_scan(c, c.gcnt.p, c.goff.p, n);
        k_ghost_emit<<<blocks_for(n, 256), 256, 0, c.stream>>>(
            n, c.pos.p, c.box, c.ghost_cutoff, c.goff.p, c.ghost_cap, c.ghost_root.p,
            c.ghost_code.p, c.ghost_pos.p, c.status.p);
        // bin particles and images (bin_neighbor_list steps 1-2b)
        k_all_keys<<<blocks_for(nall_max, 256), 256, 0, c.stream>>>(
            n, c.ghost_cap, c.key.p, c.ghost_pos.p, c.goff.p, c.grid, c.all_key.p, n_all);
        counting_sort_dev(c, c.all_key.p, n_all, nall_max, ncells, c.cnt.p, c.bin_start.p,
                          c.bin_order.p, nullptr);
        k_bin_fill<<<blocks_for(nall_max, 256), 256, 0, c.stream>>>(
            n, n_all, c.bin_order.p, c.pos.p, c.cutr.p, c.ghost_pos.p, c.ghost_root.p,
            c.ghost_code.p, c.bin_pos.p, c.bin_entry.p, c.slot_of.p);
        // full rows (bin_neighbor_list step 3 plus the mirrored entries)
        ListArgs la{};
        la.n = n;
        la.npad = c.npad;
        la.maxnb = c.maxnb;
        la.bin_pos = c.bin_pos.p;
        la.bin_entry = c.bin_entry.p;
        la.bin_start = c.bin_start.p;
        la.slot_local = c.slot_of.p;
        la.key = c.key.p;
        la.fwd = c.stencil.p;
        la.nfwd = int(c.fwd.size());
        la.box = c.box;
        la.grid = c.grid;
        la.gc = c.ghost_cutoff;
        la.uniform = c.uniform ? 1 : 0;
        la.cutoff_sq = c.cutoff_sq;
        la.sf2 = c.sf2;
        la.nbr = c.nbr.p;
        la.nnb = c.nnb.p;
        la.nhalf = c.nhalf.p;
        la.st = c.status.p;
        k_build_list<<<blocks_for(n, 256), 256, 0, c.stream>>>(la);
        MDD_HIP(hipGetLastError());
        if (dem) {
            k_remap_history<<<blocks_for(n, 256), 256, 0, c.stream>>>(
                n, c.npad, c.maxnb, c.nbr.p, c.nnb.p, c.tag.p, c.order.p, c.npad_old,
                c.maxnb_old, c.nbr_old.p, c.nnb_old.p, c.tag2.p, c.hist_old.p, c.cmask_old.p,
                c.hist.p, c.cmask.p, have_old ? 1 : 0);
            MDD_HIP(hipGetLastError());
        }
        MDD_HIP(hipMemcpyAsync(c.status_host, c.status.p, sizeof(Status), hipMemcpyDeviceToHost,
                               c.stream));
        MDD_HIP(hipStreamSynchronize(c.stream));
        const Status& s = *c.status_host;
        if (s.overflow & 4u)
            throw ApiError(-5, "a particle left a non-periodic boundary (removal is not supported)");
        if (s.overflow & 1u) {
            c.ghost_cap = int64_t(s.ghost_total) + int64_t(s.ghost_total) / 4 + 1024;
            continue;
        }
        if (s.overflow & 2u) {
            c.maxnb = s.max_row + s.max_row / 8 + 8;
            if (dem && c.maxnb > 64) c.maxnb = s.max_row <= 64 ? 64 : s.max_row;
            continue;
        }
        c.nghost = s.ghost_total;
        break;
    }
    if (c.status_host->overflow) throw ApiError(-4, "neighbor rebuild did not fit after retries");
    c.list_valid = true;
    c.list_injected = false;
    c.npairs = -1;  // computed on request
    c.rebuilds++;
}

void neighbor_setup(Ctx& c) {
    // neighbor_setup (neighbor.rs:480-541), single process
    if (c.n <= 0) throw ApiError(-1, "neighbor_setup needs particles");
    std::vector<double> cr(c.n);
    MDD_HIP(hipMemcpy(cr.data(), c.cutr.p, sizeof(double) * c.n, hipMemcpyDeviceToHost));
    double mx = 0.0, mn = 1.7976931348623157e308;
    for (double v : cr) {
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    double max_cutoff = mx > 0.0 ? 2.0 * mx * c.skin_fraction : c.bin_min_size;
    double buffer = mn < 1.7976931348623157e308 * 0.5 ? (c.skin_fraction - 1.0) * mn : 0.0;
    double ghost_cut = max_cutoff + 2.0 * buffer;
    if (ghost_cut < 0.0) ghost_cut = 0.0;
    c.ghost_cutoff = ghost_cut;
    c.max_cutoff = max_cutoff;
    double bin_min = c.bin_min_size < max_cutoff ? max_cutoff : c.bin_min_size;
    c.bin_min_size = bin_min;
    // compute_bin_grid (neighbor.rs:290-371), comm_size = 1
    int nint[3], sl[3];
    for (int d = 0; d < 3; d++) {
        double q = std::floor(c.box.L[d] / bin_min);
        if (q < 1.0) q = 1.0;
        nint[d] = int(q);
        c.bin_size[d] = c.box.L[d] / double(nint[d]);
    }
    double gcut = c.ghost_cutoff > max_cutoff ? c.ghost_cutoff : max_cutoff;
    for (int d = 0; d < 3; d++) {
        int range = int(std::ceil(max_cutoff / c.bin_size[d]));
        sl[d] = int(std::ceil(gcut / c.bin_size[d])) + range;
        c.grid.n[d] = nint[d] + 2 * sl[d];
        c.grid.origin[d] = c.box.lo[d] - c.bin_size[d] * double(sl[d]);
        c.grid.inv_bs[d] = 1.0 / c.bin_size[d];
    }
    c.grid.ncells = int64_t(c.grid.n[0]) * c.grid.n[1] * c.grid.n[2];
    if (c.grid.ncells > (int64_t(1) << 31))
        throw ApiError(-5, "bin grid too large; raise bin_size");
    double cutoff_sq = max_cutoff * max_cutoff;
    c.fwd.clear();
    const int ny = c.grid.n[1], nz = c.grid.n[2];
    for (int dx = -sl[0]; dx <= sl[0]; dx++)
        for (int dy = -sl[1]; dy <= sl[1]; dy++)
            for (int dz = -sl[2]; dz <= sl[2]; dz++) {
                int ax = std::abs(dx) - 1, ay = std::abs(dy) - 1, az = std::abs(dz) - 1;
                double mdx = double(ax > 0 ? ax : 0) * c.bin_size[0];
                double mdy = double(ay > 0 ? ay : 0) * c.bin_size[1];
                double mdz = double(az > 0 ? az : 0) * c.bin_size[2];
                if (mdx * mdx + mdy * mdy + mdz * mdz < cutoff_sq) {
                    int off = dx * ny * nz + dy * nz + dz;
                    if (off > 0) c.fwd.push_back(off);
                }
            }
    c.grid.nfwd = int(c.fwd.size());
    c.stencil.resize(c.fwd.size());
    MDD_HIP(hipMemcpy(c.stencil.p, c.fwd.data(), sizeof(int) * c.fwd.size(),
                      hipMemcpyHostToDevice));
    // save_build_positions cutoff cache (neighbor.rs:545-561)
    c.min_skin = mn;
    if (std::fabs(mx - mn) < 1e-15) {
        double cutoff = 2.0 * mn * c.skin_fraction;
        c.uniform = true;
        c.cutoff_sq = cutoff * cutoff;
    } else {
        c.uniform = false;
    }
    c.sf2 = c.skin_fraction * c.skin_fraction;
    for (int code = 0; code < 27; code++) {
        int s[3] = {code / 9 - 1, (code / 3) % 3 - 1, code % 3 - 1};
        for (int d = 0; d < 3; d++)
            c.shifts.v[code][d] = s[d] == 0 ? 0.0 : (s[d] > 0 ? c.box.L[d] : -c.box.L[d]);
    }
    if (c.ghost_cap == 0) c.ghost_cap = c.n / 2 + 1024;
    c.setup_done = true;
}

}  // namespace mdd
