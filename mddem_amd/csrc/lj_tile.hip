// LJ pair forces on cell tiles, with the step's final integration fused in.
//
// One wavefront per bin cell. The cell's candidate list (its stencil cells'
// particles and periodic images, in bin order) is loaded once per step into
// LDS: current positions, images shifted by exactly +-L as the reference's
// ghost copies are (comm.rs:280). Each particle of the cell then reads its
// row of 16-bit indices into that list, so a pair costs LDS reads instead of
// a gather through L1. G lanes share a particle (tile_lanes), 64 / G
// particles per pass. A lane's first kJ row entries are loaded before the
// tile, so the two gathers are in flight together; entries past the row
// point at a far-away slot, so the pair loop runs in branch-free groups.
//
// Every tile entry evaluates d = y - x_i with y the candidate's (shifted)
// position. For a pair the particle owns that is the reference's owner-frame
// separation; for a pair the candidate owns it is the exact negation
// (fl(a - b) = -fl(b - a)), so F_i = -fpair d in both cases (lj_force,
// md_lj/src/lib.rs:488-633). Pairs owned through the particle's own periodic
// images are kept in a separate row with global indices and the shift
// applied to the particle: d = x_j - (x_i + S).
//
// Fused post: the particle's final Verlet / Nose-Hoover half step
// (mddem_verlet/src/lib.rs:135-154, md_thermostat/src/lib.rs:313-353) runs
// right after its force is known, and each cell leaves its kinetic-energy
// sum for the thermostat.
#include "cells.hpp"

namespace mdd {

namespace {

constexpr uint32_t kTileThreads = 64u * kTileWaves;
constexpr int kTile = 768;  // candidates staged in LDS (18 KB); longer lists read HBM
constexpr int kJ = 32;      // row entries per lane held in registers per batch
constexpr int kSlots = kTile / 64;  // tile slots per lane
constexpr int kMaxTileRuns = 32;    // stencil runs held in the LDS run table
static_assert(kTile % kTileThreads == 0, "tile slots per thread");

struct TileArgs {
    int64_t ncells, npad;
    const double4* pos;
    double4* vel;
    const double* mass;
    const uint32_t* bin_start;
    const uint32_t* bin_entry;
    const int* run_first;
    const int* run_last;
    int nruns, run0;
    const uint32_t* lstart;
    const uint16_t* crow;
    const uint32_t* ncr;
    int cap, maxnb;
    const uint32_t* xnbr;
    const uint32_t* nx;
    double L[3];
    double lj1, lj2, cut2, e12, e6;
    double* f64;
    float* f32;
    double* pe;
    double* vir;  // energy passes: per-particle half share of sum d (x) F, 6 x npad
    Status* st;
    int step;
    int vv, nh, q;  // fused post: Verlet half kick, thermostat, scale slot
    double dt;
    double* partial;  // per-cell sum of m v^2
};

template <typename R>
__device__ inline R rcp_t(R x);
template <>
__device__ inline double rcp_t<double>(double x) {
    return 1.0 / x;
}
template <>
__device__ inline float rcp_t<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}

// energy passes: pair energy and virial (xx yy zz xy xz yz) of one particle
struct EAcc {
    double pe = 0.0;
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
};

template <typename R, bool ENERGY>
__device__ inline void lj_pair(const TileArgs& a, double dx, double dy, double dz, R& fx, R& fy,
                               R& fz, EAcc& pe) {
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const bool in = r2 < a.cut2;
    const R r2inv = rcp_t<R>(R(r2));
    const R r6inv = r2inv * r2inv * r2inv;
    const R fpair = r2inv * r6inv * fma(R(a.lj1), r6inv, -R(a.lj2));
    const R s = in ? -fpair : R(0);
    fx = fma(s, R(dx), fx);
    fy = fma(s, R(dy), fy);
    fz = fma(s, R(dz), fz);
    if (ENERGY && in) {
        const double r6 = double(r6inv);
        pe.pe += 0.5 * (r6 * (a.e12 * r6 - a.e6));
        // virial (md_lj/src/lib.rs:555-561), half on each side of the full rows
        const double sd = double(s);
        pe.w[0] += 0.5 * (dx * (sd * dx));
        pe.w[1] += 0.5 * (dy * (sd * dy));
        pe.w[2] += 0.5 * (dz * (sd * dz));
        pe.w[3] += 0.5 * (dx * (sd * dy));
        pe.w[4] += 0.5 * (dx * (sd * dz));
        pe.w[5] += 0.5 * (dy * (sd * dz));
    }
}

// candidate `f` of the cell's list, read from HBM (lists beyond the tile)
__device__ inline void candidate_global(const TileArgs& a, const CellRuns& Rn, uint32_t f,
                                        double& yx, double& yy, double& yz) {
    const uint32_t e = a.bin_entry[flat_to_bin(Rn, f)];
    const int code = entry_code(e);
    const double4 p = a.pos[entry_idx(e)];
    double sx, sy, sz;
    code_shifts(code, a.L[0], a.L[1], a.L[2], sx, sy, sz);
    yx = p.x + sx;
    yy = p.y + sy;
    yz = p.z + sz;
}

struct TileLds {
    // slot kTile is a candidate far away from every particle: rows are padded
    // with it, so a padded pair is out of range and adds exactly zero
    double x[kTile + 1], y[kTile + 1], z[kTile + 1];
    uint32_t run_excl[kMaxTileRuns + 1], run_base[kMaxTileRuns];
};

// the lane's row entries jb .. jb + kJ - 1 (entry k = j G + t, j = jb + u);
// entries past the row point at the far-away slot
__device__ inline void load_batch(const uint16_t* row, uint32_t stride, uint32_t jb, uint32_t G,
                                  uint32_t t, uint32_t cnt, uint32_t (&idx)[kJ]) {
#pragma unroll
    for (int u = 0; u < kJ; u++) {
        const uint32_t j = jb + uint32_t(u);
        idx[u] = j * G + t < cnt ? uint32_t(row[size_t(j) * stride]) : uint32_t(kTile);
    }
}

// pairs of the lane's row from the tile, in groups of 4 (their LDS reads go
// out together); a batch of kJ entries is preloaded, further ones load here
template <typename R, bool ENERGY>
__device__ inline void row_pairs_tile(const TileArgs& a, const TileLds& sh, const uint16_t* row,
                                      uint32_t stride, uint32_t G, uint32_t t, uint32_t cnt,
                                      const double4& xi, uint32_t (&idx)[kJ], R& fx, R& fy, R& fz,
                                      EAcc& pe) {
    for (uint32_t jb = 0;;) {
#pragma unroll
        for (int u0 = 0; u0 < kJ; u0 += 4) {
            if (!__any((jb + uint32_t(u0)) * G + t < cnt)) break;
#pragma unroll
            for (int u = u0; u < u0 + 4; u++) {
                const uint32_t id = idx[u];
                lj_pair<R, ENERGY>(a, sh.x[id] - xi.x, sh.y[id] - xi.y, sh.z[id] - xi.z, fx, fy,
                                   fz, pe);
            }
        }
        jb += uint32_t(kJ);
        if (!__any(jb * G + t < cnt)) break;
        load_batch(row, stride, jb, G, t, cnt, idx);
    }
}

// the same with every candidate read from HBM (a list longer than the tile)
template <typename R, bool ENERGY>
__device__ inline void row_pairs_global(const TileArgs& a, const CellRuns& Rn,
                                        const uint16_t* row, uint32_t stride, uint32_t G,
                                        uint32_t t, uint32_t cnt, const double4& xi, R& fx,
                                        R& fy, R& fz, EAcc& pe) {
    uint32_t j = 0;
    for (uint32_t k = t; k < cnt; k += G, j++) {
        double yx, yy, yz;
        candidate_global(a, Rn, uint32_t(row[size_t(j) * stride]), yx, yy, yz);
        lj_pair<R, ENERGY>(a, yx - xi.x, yy - xi.y, yz - xi.z, fx, fy, fz, pe);
    }
}

// The cell's candidate positions into LDS. Lane l fills slots l + 64 j; the
// run holding each slot is found with a cursor over the run table (slots
// increase with j), then all bin-entry loads and all position gathers go
// out together.
__device__ inline void fill_tile(const TileArgs& a, TileLds& sh, const CellRuns& Rn, int lane) {
    if (lane < Rn.nruns) {
        sh.run_excl[lane] = Rn.excl;
        sh.run_base[lane] = Rn.rbase;
    }
    if (lane == 0) {
        sh.run_excl[Rn.nruns] = Rn.total;
        sh.x[kTile] = 1e30;
        sh.y[kTile] = 1e30;
        sh.z[kTile] = 1e30;
    }
    __syncthreads();
    uint32_t bin[kSlots];
    int r = 0;
#pragma unroll
    for (int j = 0; j < kSlots; j++) {
        const uint32_t f = uint32_t(lane + 64 * j);
        bin[j] = 0u;
        if (f < Rn.total) {
            while (f >= sh.run_excl[r + 1]) r++;
            bin[j] = f + sh.run_base[r];
        }
    }
    uint32_t e[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; j++)
        e[j] = uint32_t(lane + 64 * j) < Rn.total ? a.bin_entry[bin[j]] : 0u;
#pragma unroll
    for (int j = 0; j < kSlots; j++) {
        const uint32_t f = uint32_t(lane + 64 * j);
        if (f < Rn.total) {
            const int code = entry_code(e[j]);
            const double4 p = a.pos[entry_idx(e[j])];
            double sx, sy, sz;
            code_shifts(code, a.L[0], a.L[1], a.L[2], sx, sy, sz);
            sh.x[f] = p.x + sx;
            sh.y[f] = p.y + sy;
            sh.z[f] = p.z + sz;
        }
    }
    __syncthreads();
}

template <typename R, bool ENERGY, bool FUSE>
__global__ void __launch_bounds__(kTileThreads) k_lj_tile(TileArgs a) {
    __shared__ TileLds sh;
    const int lane = int(threadIdx.x & 63u);
    const int64_t C = int64_t(xcd_block(blockIdx.x, gridDim.x));
    if (a.step >= 0 && a.st->abort_step <= a.step) return;
    const uint32_t s0 = a.lstart[C], M = a.lstart[C + 1] - s0;
    if (M == 0) {
        if (FUSE && a.nh && lane == 0) a.partial[C] = 0.0;
        return;
    }
    const CellRuns Rn =
        cell_runs(C, 0, a.ncells - 1, a.run_first, a.run_last, a.nruns, a.run0, a.bin_start);
    const bool tiled = Rn.total <= uint32_t(kTile) && Rn.nruns <= kMaxTileRuns;
    const uint32_t G = uint32_t(tile_lanes(M));
    const uint32_t ginv = (65536u + G - 1u) / G;
    const uint32_t sub = (uint32_t(lane) * ginv) >> 16;  // lane / G
    const uint32_t t = uint32_t(lane) - sub * G;
    const uint32_t per = 64u / G;
    const uint32_t stride = M * G;
    const uint16_t* rows = a.crow + size_t(s0) * uint32_t(a.cap) + t;
    // first pass: this lane's particle, count and first row batch (issued
    // before the tile's gathers, so both are in flight together)
    uint32_t mb = 0;
    uint32_t m = sub;
    bool act = sub < per && m < M;
    uint32_t q = s0 + m, cnt = 0u;
    double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
    uint32_t idx[kJ];
    if (act) {
        cnt = min(a.ncr[q], uint32_t(a.maxnb));
        xi = a.pos[q];
    }
    if (tiled) {
        load_batch(rows + m * G, stride, 0u, G, t, cnt, idx);
        fill_tile(a, sh, Rn, lane);
    }
    double ke = 0.0;
    for (;;) {
        R fx = R(0), fy = R(0), fz = R(0);
        EAcc pe;
        if (act) {
            const uint16_t* row = rows + m * G;
            if (tiled)
                row_pairs_tile<R, ENERGY>(a, sh, row, stride, G, t, cnt, xi, idx, fx, fy, fz, pe);
            else
                row_pairs_global<R, ENERGY>(a, Rn, row, stride, G, t, cnt, xi, fx, fy, fz, pe);
            uint32_t nxq = a.nx[q];
            if (nxq > uint32_t(a.maxnb)) nxq = uint32_t(a.maxnb);
            for (uint32_t k = t; k < nxq; k += G) {
                const uint32_t e = a.xnbr[size_t(q) * uint32_t(a.maxnb) + k];
                const int code = entry_code(e);
                const double4 pj = a.pos[entry_idx(e)];
                double sx, sy, sz;
                code_shifts(code, a.L[0], a.L[1], a.L[2], sx, sy, sz);
                lj_pair<R, ENERGY>(a, pj.x - (xi.x + sx), pj.y - (xi.y + sy), pj.z - (xi.z + sz),
                                   fx, fy, fz, pe);
            }
        }
        // the particle's G lanes, summed by a fixed pairwise tree
#pragma unroll
        for (uint32_t off = 1; off < 16u; off <<= 1) {
            if (off < G) {
                const R ox = __shfl(fx, lane + int(off), 64);
                const R oy = __shfl(fy, lane + int(off), 64);
                const R oz = __shfl(fz, lane + int(off), 64);
                EAcc op;
                if (ENERGY) {
                    op.pe = __shfl(pe.pe, lane + int(off), 64);
#pragma unroll
                    for (int k = 0; k < 6; k++) op.w[k] = __shfl(pe.w[k], lane + int(off), 64);
                }
                if ((t & (2u * off - 1u)) == 0u && t + off < G) {
                    fx += ox;
                    fy += oy;
                    fz += oz;
                    if (ENERGY) {
                        pe.pe += op.pe;
#pragma unroll
                        for (int k = 0; k < 6; k++) pe.w[k] += op.w[k];
                    }
                }
            }
        }
        if (act && t == 0) {
            if constexpr (sizeof(R) == 8) {
                a.f64[q] = fx;
                a.f64[a.npad + q] = fy;
                a.f64[2 * a.npad + q] = fz;
            } else {
                a.f32[q] = fx;
                a.f32[a.npad + q] = fy;
                a.f32[2 * a.npad + q] = fz;
            }
            if (ENERGY) {
                a.pe[q] = pe.pe;
#pragma unroll
                for (int k = 0; k < 6; k++) a.vir[k * a.npad + q] = pe.w[k];
            }
            if (FUSE) {
                double4 v = a.vel[q];
                const double h = 0.5 * a.dt * v.w;
                const double Fx = double(fx), Fy = double(fy), Fz = double(fz);
                if (a.vv) {
                    v.x += h * Fx;
                    v.y += h * Fy;
                    v.z += h * Fz;
                }
                if (a.nh) {
                    const double scale = a.st->nh_scale[a.q];
                    v.x += h * Fx;
                    v.y += h * Fy;
                    v.z += h * Fz;
                    v.x *= scale;
                    v.y *= scale;
                    v.z *= scale;
                    ke += a.mass[q] * (v.x * v.x + v.y * v.y + v.z * v.z);
                }
                a.vel[q] = v;
            }
        }
        mb += per;
        if (mb >= M) break;
        // further passes (cells of more than 64 particles)
        m = mb + sub;
        act = sub < per && m < M;
        q = s0 + m;
        cnt = 0u;
        if (act) {
            cnt = min(a.ncr[q], uint32_t(a.maxnb));
            xi = a.pos[q];
        }
        if (tiled) load_batch(rows + m * G, stride, 0u, G, t, cnt, idx);
    }
    if (FUSE && a.nh) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
        if (lane == 0) a.partial[C] = ke;
    }
}

// sum of `count` partials in a fixed order -> out[0]
__global__ void __launch_bounds__(1024) k_ke_reduce(const double* __restrict__ partial,
                                                    int64_t count, double* __restrict__ out,
                                                    const Status* st, int step) {
    __shared__ double lds[16];
    if (step >= 0 && st->abort_step <= step) return;
    const int64_t tid = threadIdx.x;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t b0 = 0; b0 < count; b0 += 8 * 1024) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t b = b0 + u * 1024 + tid;
            if (b < count) s[u] += partial[b];
        }
    }
    double v = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double r = 0.0;
        for (int k = 0; k < 16; k++) r += lds[k];
        out[0] = r;
    }
}

// Tile rows back to full 32-bit rows (row_addr layout), for the pair
// downloads: entry = the candidate's bin entry, flagged when the candidate
// precedes the particle in bin order; then the image-owned pairs.
__global__ void __launch_bounds__(64) k_expand_tile(TileArgs a, uint32_t* __restrict__ nbr, int lg) {
    const int lane = int(threadIdx.x);
    const int64_t C = int64_t(blockIdx.x);
    const uint32_t s0 = a.lstart[C], M = a.lstart[C + 1] - s0;
    if (M == 0) return;
    const CellRuns Rn =
        cell_runs(C, 0, a.ncells - 1, a.run_first, a.run_last, a.nruns, a.run0, a.bin_start);
    const uint32_t i0 = a.bin_start[C];
    const uint32_t G = uint32_t(tile_lanes(M));
    for (uint32_t m = 0; m < M; m++) {
        const uint32_t q = s0 + m;
        const uint32_t fs = i0 + m - Rn.base0;
        uint32_t cnt = a.ncr[q];
        if (cnt > uint32_t(a.maxnb)) cnt = uint32_t(a.maxnb);
        for (uint32_t k = uint32_t(lane); k < cnt; k += 64u) {
            const uint32_t idx = a.crow[tile_addr(s0, uint32_t(a.cap), M, G, m, k)];
            const uint32_t e = a.bin_entry[flat_to_bin(Rn, idx)];
            nbr[row_addr(q, int(k), a.npad, lg)] = idx < fs ? (e | kRevBit) : e;
        }
        const uint32_t nxq = a.nx[q];
        for (uint32_t k = uint32_t(lane); k < nxq; k += 64u)
            if (cnt + k < uint32_t(a.maxnb))
                nbr[row_addr(q, int(cnt + k), a.npad, lg)] =
                    a.xnbr[size_t(q) * uint32_t(a.maxnb) + k];
    }
}

template <typename R, bool E, bool F>
void launch_t(Ctx& c, const TileArgs& a) {
    k_lj_tile<R, E, F><<<int(c.grid.ncells), kTileThreads, 0, c.stream>>>(a);
}

}  // namespace

namespace {
TileArgs tile_args(Ctx& c, int step) {
    TileArgs a{};
    a.ncells = c.grid.ncells;
    a.npad = c.npad;
    a.pos = c.pos.p;
    a.vel = c.vel.p;
    a.mass = c.mass.p;
    a.bin_start = c.bin_start.p;
    a.bin_entry = c.bin_entry.p;
    a.run_first = c.run_first.p;
    a.run_last = c.run_last.p;
    a.nruns = c.nruns;
    a.run0 = c.run0;
    a.lstart = c.start.p;
    a.crow = c.crow.p;
    a.ncr = c.nmain.p;
    a.cap = c.tile_cap;
    a.maxnb = c.maxnb;
    a.xnbr = c.xnbr.p;
    a.nx = c.nx.p;
    for (int d = 0; d < 3; d++) a.L[d] = c.box.L[d];
    a.lj1 = c.lj1;
    a.lj2 = c.lj2;
    a.cut2 = c.lj_cut2;
    a.e12 = c.lj_e12;
    a.e6 = c.lj_e6;
    a.f64 = c.f64.p;
    a.f32 = c.f32.p;
    a.pe = c.pe.p;
    a.vir = c.vir.p;
    a.st = c.status.p;
    a.step = step;
    a.vv = c.vv_on;
    a.nh = c.nh_on;
    a.q = step & 1;
    a.dt = c.dt;
    a.partial = c.partial.p;
    return a;
}
}  // namespace

// tile force pass; fuse: also the step's final integration (scheduled steps)
void launch_lj_tile(Ctx& c, int step, bool energy, bool fuse) {
    if (fuse && c.nh_on) c.partial.resize(size_t(c.grid.ncells));
    const TileArgs a = tile_args(c, step);
    const bool mixed = c.precision == 1;
    if (energy) {
        if (mixed)
            launch_t<float, true, false>(c, a);
        else
            launch_t<double, true, false>(c, a);
    } else if (fuse) {
        if (mixed)
            launch_t<float, false, true>(c, a);
        else
            launch_t<double, false, true>(c, a);
    } else {
        if (mixed)
            launch_t<float, false, false>(c, a);
        else
            launch_t<double, false, false>(c, a);
    }
    MDD_HIP(hipGetLastError());
}

void expand_tile_rows(Ctx& c) {
    const TileArgs a = tile_args(c, -1);
    k_expand_tile<<<int(c.grid.ncells), 64, 0, c.stream>>>(a, c.nbr.p, c.list_lg);
    MDD_HIP(hipGetLastError());
}

// thermostat kinetic energy: sum the first `count` partials into ke_total
void launch_ke_reduce(Ctx& c, int step, int64_t count) {
    c.ke_total.resize(1);
    k_ke_reduce<<<1, 1024, 0, c.stream>>>(c.partial.p, count, c.ke_total.p, c.status.p, step);
    MDD_HIP(hipGetLastError());
}

}  // namespace mdd
