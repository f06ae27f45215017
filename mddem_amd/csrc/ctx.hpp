// Internal types of the gfx950 particle path: device buffers, the context
// behind the C ABI, and the encoding of neighbor-list entries.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace mdd {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* expr, const char* file, int line)
        : std::runtime_error(std::string(hipGetErrorString(c)) + " in " + expr + " (" + file +
                             ":" + std::to_string(line) + ")"),
          code(c) {}
};

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

#define MDD_HIP(expr)                                                   \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) throw ::mdd::HipError(e_, #expr, __FILE__, __LINE__); \
    } while (0)

// ---- neighbor entry encoding -------------------------------------------------
// One 32-bit word per (particle, partner) entry of the full list:
//   bits  0..25  partner particle index (root particle for periodic images)
//   bits 26..30  image code = (sx+1)*9 + (sy+1)*3 + (sz+1); 13 = no shift.
//                The shift applies to the non-owner of the pair.
//   bit  31      set when this particle is NOT the owner of the pair, i.e.
//                the pair sits in the partner's half-list row.
constexpr uint32_t kIdxBits = 26;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
constexpr uint32_t kRevBit = 1u << 31;
constexpr int kCodeNone = 13;
constexpr int64_t kMaxParticles = int64_t(1) << kIdxBits;
// kinetic-energy partials of a thermostatted post kernel (one per block)
constexpr int kMaxKePartials = 1024;

__host__ __device__ inline uint32_t enc_entry(uint32_t idx, int code, bool rev) {
    return idx | (uint32_t(code) << kIdxBits) | (rev ? kRevBit : 0u);
}
__host__ __device__ inline uint32_t entry_idx(uint32_t e) { return e & kIdxMask; }
__host__ __device__ inline int entry_code(uint32_t e) { return int((e >> kIdxBits) & 31u); }
__host__ __device__ inline bool entry_rev(uint32_t e) { return (e & kRevBit) != 0u; }
__host__ __device__ inline int code_of(int sx, int sy, int sz) {
    return (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1);
}
// periodic displacement of one axis: exactly +L, -L or 0.0 (comm.rs:280 packs
// pos + periodic_offset * size with periodic_offset = +-1.0)
__host__ __device__ inline double image_shift(int s, double L) {
    return s == 0 ? 0.0 : (s > 0 ? L : -L);
}
// Neighbor-row layout: entry k of particle p. G consecutive entries of a
// row are contiguous (G lanes of one particle read them together), and
// entry group k / G of consecutive particles is contiguous across particles.
__host__ __device__ inline int64_t row_addr(int64_t p, int k, int64_t npad, int lg) {
    return ((int64_t(k >> lg) * npad + p) << lg) + (k & ((1 << lg) - 1));
}
__host__ __device__ inline int code_sx(int code) { return code / 9 - 1; }
__host__ __device__ inline int code_sy(int code) { return (code / 3) % 3 - 1; }
__host__ __device__ inline int code_sz(int code) { return code % 3 - 1; }
// The three digits of an image code with the divisions by 9 and by 3 done
// as multiply-shifts (exact for codes below 27, checked below), and the
// shifts as s * L: exactly +L, -L or +0.0, the values image_shift gives.
__host__ __device__ constexpr int code_q9(int code) { return (code * 57) >> 9; }
__host__ __device__ constexpr int code_q3(int code) { return (code * 11) >> 5; }
__host__ __device__ inline void code_shifts(int code, double Lx, double Ly, double Lz,
                                            double& sx, double& sy, double& sz) {
    const int q9 = code_q9(code), q3 = code_q3(code);
    sx = double(q9 - 1) * Lx;
    sy = double(q3 - 3 * q9 - 1) * Ly;
    sz = double(code - 3 * q3 - 1) * Lz;
}
constexpr bool code_digits_exact() {
    for (int c = 0; c < 27; c++)
        if (code_q9(c) != c / 9 || code_q3(c) != c / 3) return false;
    return true;
}
static_assert(code_digits_exact(), "multiply-shift digits of image codes");

// Periodic image displacement per image code: exactly +L, -L or 0.0 per axis,
// so x + shift reproduces the reference ghost coordinate fl(x +- L).
struct Shift27 {
    double v[27][3];
};

struct BoxP {
    double lo[3], hi[3], L[3];
    int periodic[3];
};

// A position in fixed point: x * 2^32 / L per axis, rounded to the nearest
// integer and kept modulo 2^32 (qs[d] = 2^32 / L[d]).
__host__ __device__ inline int quantize_axis(double x, double qs) {
    return int(uint32_t(uint64_t(llrint(x * qs))));
}

struct GridP {
    double origin[3], inv_bs[3];
    int n[3];
    int nfwd;
    int64_t ncells;
};

// One system of a batched ensemble: its box, bin grid (cells
// cell_off .. cell_off + grid.ncells - 1 of the shared cell space), stencil
// runs (run_off .. run_off + nruns - 1 of the shared run tables), pair
// cutoff, rebuild threshold and material constants. Particles of system s
// occupy one contiguous index range.
struct SysP {
    BoxP box;
    GridP grid;
    int64_t cell_off;
    double gc;            // ghost (image) cutoff
    double cutoff_sq;     // uniform pair cutoff squared
    double sf2;           // skin_fraction^2 (per-pair cutoff)
    double thr2;          // rebuild threshold: ((skin_fraction - 1) * min radius)^2
    int uniform;
    int run_off, nruns, run0;
    double beta, mu, eeff, geff;  // single-material pair constants
};

// One wall (dem_wall/src/lib.rs:302-453). kind 0: plane through p with unit
// normal n, active inside [blo, bhi]; 1: cylinder about axis through (c0, c1)
// of radius `radius`, axial range [lo, hi]; 2: sphere about c of radius
// `radius`; 3: region surface. inside: particles live inside the curved surface. Plane motion
// (WallMotion, lib.rs:262-293): 0 static, 1 constant velocity (v), 2
// oscillate (a0 amplitude, a1 frequency), 3 servo (a0 target force, a1 max
// velocity, a2 gain). p is the configured (origin) point; the current point
// and velocity live on the device (WallsDyn).
constexpr int kMaxRegionPrims = 8;
struct Wall {
    double p[3], n[3], blo[3], bhi[3];
    int mat;
    int kind, active, motion;
    double v[3], a0, a1, a2;
    double c[3], radius, lo, hi;
    int axis, inside;
    // kind 3, region surface (WallRegion, lib.rs:435-447): the region tree's
    // leaf primitives in depth-first order, 12 parameters each (block: min,
    // max; sphere: center, radius; cylinder: center[2], radius, axis, lo, hi;
    // cone: center[2], axis, rad_lo, rad_hi, lo, hi; plane: point, normal)
    int nprim;
    int ptype[kMaxRegionPrims];
    double prm[kMaxRegionPrims][12];
};
constexpr int kMaxWalls = 12;
// scalar contact force sums (force_accumulator) are kept in fixed point:
// integer sums do not depend on the order the atomics land in
constexpr double kWallAccScale = 4294967296.0;  // 2^32 per newton
struct WallDyn {
    double p[3], v[3];
    unsigned long long acc;
    unsigned long long pad;
};
struct WallsDyn {
    double time;  // Walls::time
    double pad;
    WallDyn w[kMaxWalls];
};
constexpr int kMaxMat = 16;
constexpr int kNumExtras = 10;  // optional per-material contact properties

struct WallsP {
    int count;
    Wall w[kMaxWalls];
};

// wall_move (dem_wall/src/lib.rs:766-815) for every plane wall, then the
// force sums are cleared for the coming force pass (wall_zero_force_
// accumulators, lib.rs:821-834). One thread runs it.
__host__ __device__ inline void walls_advance(const WallsP& W, WallsDyn& D, double dt,
                                              bool move = true, bool zero = true) {
    const double PI = 3.141592653589793;
    const double time = D.time;
    for (int k = 0; k < W.count; k++) {
        const Wall& w = W.w[k];
        WallDyn& d = D.w[k];
        if (move && w.kind == 0 && w.active) {
            if (w.motion == 1) {
                for (int q = 0; q < 3; q++) d.p[q] += d.v[q] * dt;
            } else if (w.motion == 2) {
                const double phase = 2.0 * PI * w.a1 * (time + dt);
                const double disp = w.a0 * sin(phase);
                for (int q = 0; q < 3; q++) d.p[q] = w.p[q] + disp * w.n[q];
                const double vel_mag = w.a0 * 2.0 * PI * w.a1 * cos(phase);
                for (int q = 0; q < 3; q++) d.v[q] = vel_mag * w.n[q];
            } else if (w.motion == 3) {
                const double measured = double(static_cast<long long>(d.acc)) / kWallAccScale;
                double vel_mag = w.a2 * (w.a0 - measured);
                if (vel_mag < -w.a1) vel_mag = -w.a1;
                if (vel_mag > w.a1) vel_mag = w.a1;
                for (int q = 0; q < 3; q++) d.v[q] = vel_mag * w.n[q];
                for (int q = 0; q < 3; q++) d.p[q] += d.v[q] * dt;
            }
        }
        if (zero) d.acc = 0ull;
    }
    if (move) D.time = time + dt;
}

// Material pair tables (beta, friction, E*, G*) in row-major nmat x nmat.
struct MatP {
    int n;
    double beta[kMaxMat * kMaxMat];
    double mu[kMaxMat * kMaxMat];
    double eeff[kMaxMat * kMaxMat];
    double geff[kMaxMat * kMaxMat];
};

// Device-resident run status, read by the host at chunk boundaries.
struct Status {
    int abort_step;          // first step whose drift needs a rebuild; INT_MAX none
    int error;               // 0 or an MDD_ERR_* code raised on the device
    unsigned overflow;       // bit 0: image capacity, bit 1: neighbor row capacity,
                             // bit 2: particle outside a non-periodic face,
                             // bit 3: candidate list beyond 16-bit tile indices
                             // bit 4: more old contacts than history records
    int max_row;             // longest full-list row at the last build
    int max_half;            // longest half-list row at the last build
    unsigned ghost_total;    // periodic images at the last build
    int overlap_count;       // large-overlap pairs in the current step
    int nh_kind[2];          // per step parity: 1 = nh_p is p_xi after a post
                             // half-update (or the initial p_xi), 0 = after a pre
    unsigned hist_records;   // contacts of the old list at the last history remap
    unsigned listed;         // DEM: particles with a nonempty row at the last build
    double ke;               // kinetic energy at the last thermostat finish
    double p_xi;             // Nose-Hoover momentum at the last thermostat finish
    double nh_p[2];          // p_xi by step parity
    double nh_scale[2];      // velocity scale by step parity
};

// MaterialTable::build_pair_tables (crates/dem_atom/src/lib.rs:389-515) for
// one pair of materials: beta, friction, E*, G*
inline void pair_constants(double Ei, double nui, double ei, double mui, double Ej, double nuj,
                           double ej, double muj, double out[4]) {
    const double PI = 3.141592653589793;
    const double e_ij = __builtin_sqrt(ei * ej);
    const double log_e = __builtin_log(e_ij);
    out[0] = -log_e / __builtin_sqrt(PI * PI + log_e * log_e);
    out[1] = __builtin_sqrt(mui * muj);
    out[2] = 1.0 / ((1.0 - nui * nui) / Ei + (1.0 - nuj * nuj) / Ej);
    out[3] = 1.0 / (2.0 * (2.0 - nui) * (1.0 + nui) / Ei + 2.0 * (2.0 - nuj) * (1.0 + nuj) / Ej);
}

// Device pointers whose hipFree waits until recorded launches are issued
// (set by the recording thread, see LaunchLog).
inline std::vector<void*>*& deferred_frees() {
    static thread_local std::vector<void*>* list = nullptr;
    return list;
}

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            // while launches are being recorded, a buffer they name stays
            // allocated until they have been issued
            if (deferred_frees())
                deferred_frees()->push_back(p);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    void resize(size_t count) {
        if (count <= n && p) return;
        release();
        size_t c = count ? count : 1;
        MDD_HIP(hipMalloc(&p, c * sizeof(T)));
        n = c;
    }
    void swap(DBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

// Launch recording. While `on`, launch() does not launch: it appends the
// kernel, its grid, block and argument bytes to `sig` and a closure that
// issues the same launch to `calls`. A launch sequence whose signature
// equals that of a captured graph is replayed as that graph; two sequences
// with equal signatures launch the same kernels on the same arguments.
struct LaunchLog {
    bool on = false;
    std::vector<uint8_t> sig;
    std::vector<std::function<void(hipStream_t)>> calls;
    std::vector<void*> frees;  // buffers released while recording
};

struct GraphEntry {
    std::vector<uint8_t> sig;
    hipGraphExec_t exec = nullptr;
};

struct Ctx {
    int device = 0, model = 0, precision = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // ---- configuration
    BoxP box{};
    double dt = 1.0;
    double skin_fraction = 1.0, bin_min_size = 1.0;
    int every = 0, check = 1, sort_every = 1000, newton = 1;
    int box_exit_rebuild = 1;
    MatP mat{};
    double lj1 = 0, lj2 = 0, lj_cut2 = 0, lj_e12 = 0, lj_e6 = 0;
    double lj_eps = 1.0, lj_sigma = 1.0;  // for the tail corrections
    int nh_on = 0;
    double nh_T = 0, nh_tau = 1, nh_ndof = 0, nh_Q = 1;
    bool nh_ready = false;
    int vv_on = 0, rot_on = 0;
    int gravity_on = 0;
    double g[3] = {0, 0, 0};
    WallsP walls{};
    DBuf<WallsP> walls_dev;   // device copy of the wall table (integrators)
    DBuf<WallsDyn> wdyn;      // current wall points, velocities, force sums
    bool walls_moving = false;

    // ---- batched ensemble (nsys > 0): particles of system s are
    // [sys_first[s], sys_first[s + 1]); per-particle system index psys
    int nsys = 0;
    std::vector<SysP> sys_h;
    std::vector<int64_t> sys_first;
    std::vector<double> sys_rmin, sys_rmax;  // neighbor radius range per system
    DBuf<SysP> sysp;
    DBuf<uint32_t> psys;
    DBuf<int64_t> sys_cell_off;  // nsys + 1 cell offsets (last = total cells)
    struct Staged {
        double lo[3], hi[3];
        double youngs, poisson, restitution, friction;
        std::vector<uint32_t> tag;
        std::vector<double> pos, vel, omega, radius, cutr, mass;
    };
    std::vector<Staged> staged;
    bool systems_dirty = false;
    const SysP* sysp_or_null() const { return nsys > 0 ? sysp.p : nullptr; }

    // ---- neighbor grid (neighbor_setup)
    bool setup_done = false;
    double max_cutoff = 0, ghost_cutoff = 0, bin_size[3] = {1, 1, 1};
    GridP grid{};
    bool uniform = true;
    double cutoff_sq = 0, sf2 = 1, min_skin = 0;
    Shift27 shifts{};

    // ---- sizes
    int64_t n = 0, cap = 0, npad = 0;
    int64_t ghost_cap = 0, nghost = 0;
    int maxnb = 0;          // row capacity of the full list (multiple of 1 << lg)
    int lg = 2;             // log2 of lanes per particle in the pair-force kernels
    int list_lg = 2;        // the row layout the current list was built with
    int lg_old = 2;         // the row layout of the previous list
    int64_t npairs = 0;
    bool list_valid = false;
    bool list_injected = false;
    // DEM: the particles with a nonempty row at the last rebuild and their
    // count (the force pass at one lane per particle is split by them)
    DBuf<uint32_t> listed, listed_n;
    bool listed_ok = false;
    // DEM force pass at one lane per particle: -1 split by row length when
    // fewer than half the particles have neighbors (Status::listed of the
    // last build the host has read), 0 one kernel, 1 always split
    int dem_split = -1;
    // DEM pair pass in tile order (single systems): the particles listed
    // tile by tile, a tile being kTileCols bin columns along y at one x,
    // tiles x-major inside each band of columns; rebuilt with the list
    DBuf<uint32_t> tile_list, tile_cnt, tile_off;
    bool tile_order = true;  // mdd_set_dem_tiles
    bool tile_ok = false;
    // DEM pair pass at one lane per particle with the pairs in contact
    // compacted per wavefront (k_hertz_wave; mdd_set_dem_wave)
    bool dem_wave = true;

    // ---- particle arrays (current order)
    DBuf<double4> pos, vel, omega, quat;  // pos.w radius, vel.w 1/m, omega.w 1/I
    DBuf<uint32_t> tag, type;
    DBuf<int4> image;
    DBuf<double> mass, cutr;
    DBuf<double> f64, t64;  // 3*npad SoA force / torque (FP64 mode)
    DBuf<float> f32, t32;   // 3*npad SoA force / torque (mixed mode)
    DBuf<double4> build_pos;
    // alternates for the rebuild permutation
    DBuf<double4> pos2, vel2, omega2, quat2;
    DBuf<uint32_t> tag2, type2;
    DBuf<int4> image2;
    DBuf<double> mass2, cutr2;
    DBuf<double> f64b, t64b;
    DBuf<float> f32b, t32b;

    // ---- rebuild scratch
    DBuf<uint32_t> key, key2, order, slot_of, cnt, start, scan_tmp, gcnt, goff;
    DBuf<uint32_t> ghost_root;
    DBuf<uint8_t> ghost_code;
    DBuf<double4> ghost_pos;
    DBuf<uint32_t> bin_start, bin_entry, gstart, placed, ghost_key, ghost_order;
    DBuf<double4> bin_pos;
    std::vector<int> full;   // all offsets whose cells can hold a pair, ascending
    DBuf<int> run_first, run_last;  // the full stencil as runs of consecutive offsets
    int nruns = 0, run0 = 0;        // run0: the run holding offset 0
    DBuf<uint32_t> bin_gidx;        // bin slot of a periodic image -> image index
    DBuf<uint32_t> gnbr, gcnt_img;  // per image: particles owning a pair with it
    DBuf<uint32_t> nmain;           // row length from the particle's own cell pass
    // LJ tile rows: crow (16-bit candidate indices, tile_addr layout, count
    // nmain) and xnbr (pairs owned through periodic images, p * maxnb + k,
    // count nx)
    // tile_enabled: the caller's choice (mdd_set_force_lanes); tile_blocked:
    // a candidate list outgrew 16-bit indices
    bool tile_enabled = false, tile_blocked = false, list_tile = false;
    bool want_tile() const { return model == 0 && nsys == 0 && tile_enabled && !tile_blocked; }
    // LJ mixed row pass on fixed-point positions: each coordinate in units
    // of L / 2^32 of its axis, so that the 32-bit difference of two
    // particles wraps to the minimum image (every axis at least twice the
    // list cutoff: the listed image is the nearest). The step kernels keep
    // xq current while xq_valid and the rebuild's permutation rewrites it;
    // an upload or a box change clears it and the next pair pass refills it.
    DBuf<int4> xq;
    bool lj_fixed = true;  // mdd_set_lj_positions
    bool xq_valid = false;
    // the same fixed-point form of the build positions, written by the
    // permutation beside xq: the first half step's displacement check reads
    // 16 B instead of the 32 B double4 and goes back to the double4 only
    // near the threshold. Valid from such a permutation until an upload
    // rewrites the build positions.
    DBuf<int4> bxq;
    bool bxq_valid = false;
    bool xq_usable() const {
        if (!(model == 0 && precision == 1 && nsys == 0 && lj_fixed && !list_tile)) return false;
        for (int d = 0; d < 3; d++)
            if (!(box.L[d] > 2.0 * max_cutoff)) return false;
        return true;
    }
    bool last_force_fused = false;  // the last scheduled force pass also integrated
    int64_t fused_parts = 0;        // its kinetic-energy partials (cell tiles)
    // the mixed LJ row pass with the thermostatted final half fused in
    // (mdd_set_lj_post_fused); fused_rows: the last pass was that one, and
    // left the post kernel's partials
    bool lj_post_fused = true;
    bool fused_rows = false;
    // the final half step of step `step` is deferred into the next step's
    // first pass (runs without a thermostat)
    bool post_pending = false;
    // neighbor build kernel: 0 chosen by occupancy, 1 one wavefront per cell,
    // 2 one lane per member slot (sparse grids)
    int build_mode = 0;
    DBuf<uint32_t> bin_cell;  // cell of every bin slot (sparse build)
    int tile_cap = 0;
    DBuf<uint16_t> crow;
    DBuf<uint32_t> xnbr, nx;
    DBuf<double> ke_total;          // NH kinetic energy sum for the next pre kernel
    DBuf<uint32_t> row_blk;         // per-block row maxima of the last build
    DBuf<double> vir;               // energy passes: per-particle virial, 6 x npad
    // DEM contact virial: computed in the force pass of every step whose
    // number is a multiple of virial_every (0: never); vir then holds the
    // pass of step vir_step
    int virial_every = 0;
    bool vir_valid = false;
    int64_t vir_step = -1;

    // ---- full neighbor list: entry k of particle i at row_addr(i, k, npad, lg)
    DBuf<uint32_t> nbr, nbr_old;
    DBuf<uint32_t> nnb, nnb_old, nhalf;
    int64_t npad_old = 0;
    bool remap_have_old = false;
    int maxnb_old = 0;
    // DEM shear history per list slot, canonical sign (lower tag's frame)
    DBuf<double> hist, hist_old;  // hist_slots * maxnb * npad, SoA by component
    // bit k of cmask[i]: slot k of particle i was in contact at the last force pass
    DBuf<uint64_t> cmask, cmask_old;
    DBuf<double> mat_dev;         // beta | mu | E* | G*, each nmat x nmat
    // optional contact models: pair tables of the optional material
    // properties (rolling / twisting friction, cohesion, surface energy, Hooke
    // and SDS stiffness and damping), each nmat x nmat, and the model choices
    DBuf<double> mext_dev;
    bool ext_on = false;
    int cm_hooke = 0, cm_dmt = 0, cm_roll_sds = 0, cm_twist_sds = 0;
    double mext_host[kNumExtras * kMaxMat * kMaxMat] = {};
    // history values per contact: 3 (tangential spring) or, with the optional
    // models, 7 (spring, rolling displacement, twist angle)
    int hist_slots = 3;
    // 1: a contact's history survives a rebuild only if the same particle
    // still owns the pair (the reference's per-owner store); 0: it survives
    // whichever side owns the pair (keyed by the pair alone)
    int hist_by_owner = 1;
    // history remap: by sorted pair keys (default) or by a search of each
    // particle's old row
    bool remap_sorted = true;
    int64_t n_list_old = 0;   // particles of the list the remap reads
    int64_t hk_cap = 0;       // capacity of the contact records
    DBuf<uint32_t> hk_inv, hk_cnt, hk_off, hk_lo, hk_bcnt, hk_bstart;
    DBuf<uint4> hk_rec, hk_rec2;
    DBuf<double> pe;              // per-particle LJ energy (thermo only)

    // ---- run state
    DBuf<Status> status;
    Status* status_host = nullptr;  // pinned mirror
    DBuf<double> partial;           // per-block kinetic-energy partial sums
    DBuf<double> partial2;          // per-block thermo partial sums
    int64_t step = 0, rebuilds = 0;
    int64_t removed = 0;  // particles removed at non-periodic faces (pbc)
    int64_t sort_counter = 0;
    int chunk = 8;
    // chunk sizing: step of the last rebuild and the smoothed number of steps
    // between rebuilds (0 until two rebuilds were seen)
    int64_t last_rebuild = 0;
    double rb_interval = 0.0;
    int retries = 0;  // consecutive builds that did not fit
    int64_t last_resume = -1;  // step of the last rebuild inside a run and
    int repeats = 0;           // how many times in a row it was that step

    // ---- rebuild launch sequences captured as graphs
    LaunchLog rec;
    int row_nblk = 0;  // per-block row-maxima slots of the last build
    // device time of whole rebuilds (first launch to the gate) and of the
    // idle gap before them (end of the step chunk that asked for the rebuild
    // to its first launch), when profiling: events 3k, 3k+1, 3k+2
    std::vector<hipEvent_t> rb_ev;
    std::vector<char> rb_has_gap;
    hipEvent_t rb_mark = nullptr;  // end of the last step chunk
    bool rb_chunk_mark = false;
    int64_t rb_pending = 0, rb_count = 0, rb_gap_count = 0;
    double rb_ms_sum = 0, rb_gap_ms_sum = 0;
    std::vector<GraphEntry> graphs;  // most recent last
    bool graphs_on = true;
    int64_t graph_hits = 0, graph_captures = 0;

    // ---- profiling of the pair-force kernel
    int prof_every = 0;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms_sum = 0;
    int64_t prof_count = 0;
    int64_t force_launches = 0;
    int64_t prof_ctr = 0;
    int64_t prof_pending = 0;
    std::vector<int> prof_step;  // step of each pending timed launch
};

// ---- launchers implemented in the .hip files
// scan_sort.hip
void exclusive_scan(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n);
void counting_sort(Ctx& c, const uint32_t* key, int64_t n, const uint32_t* n_dev, int64_t ncells,
                   uint32_t* cnt, uint32_t* start, uint32_t* placed, uint32_t* order,
                   uint32_t* slot_of, bool cnt_zero = false, const uint32_t* add = nullptr,
                   uint32_t* out2 = nullptr, bool counted = false);
void fill_u32(Ctx& c, uint32_t* p, uint32_t v, int64_t n);
void fill_f64(Ctx& c, double* p, double v, int64_t n);
void fill_f32(Ctx& c, float* p, float v, int64_t n);
// neighbor.hip
void rebuild(Ctx& c);
void rebuild_async(Ctx& c, int step, bool retry);
// drop the particles a rebuild found outside a non-periodic face (pbc)
void remove_outside(Ctx& c);
// the periodic sort's test at `step`: a rebuild request unless the
// particles are still in bin order
void launch_sort_check(Ctx& c, int step);
void release_graphs(Ctx& c);
bool rebuild_grow(Ctx& c, const Status& s);
// forces.hip
void launch_pair_forces(Ctx& c, int step, bool energy);
void launch_walls_gravity(Ctx& c, bool walls, bool gravity);
// standalone wall_move / wall_zero_force_accumulators
void launch_wall_systems(Ctx& c, bool move, bool zero);
// lj_tile.hip
void launch_lj_tile(Ctx& c, int step, bool energy, bool fuse);
void launch_ke_reduce(Ctx& c, int step, int64_t count);
void expand_tile_rows(Ctx& c);
void upload_materials(Ctx& c);
// switch between the plain Hertz-Mindlin kernel and the optional-model kernel;
// a change of history width drops the stored history and rebuilds the list
inline void set_contact_path(Ctx& c, bool on) {
    const int slots = on ? 7 : 3;
    c.ext_on = on;
    if (slots != c.hist_slots) {
        c.hist_slots = slots;
        c.list_valid = false;
    }
}
// integrate.hip
void launch_pre(Ctx& c, int step, bool check_rebuild);
void launch_post_pre(Ctx& c, int step, bool check_rebuild);
void launch_post(Ctx& c, int step);
void launch_initial_integration(Ctx& c, int step, bool check_rebuild);
void launch_final_integration(Ctx& c);
void launch_initial_rotation(Ctx& c);
void launch_final_rotation(Ctx& c);
void launch_ke_partials(Ctx& c);
void launch_nh_finish(Ctx& c, int parity);
void launch_nh_seed(Ctx& c, int parity);
void launch_nh_pre_standalone(Ctx& c);
void launch_nh_post_standalone(Ctx& c);
// totals of sum m v^2, sum m v (3), sum m, sum m |v - vbar|^2, half-list
// pairs and contact bits: through a context buffer, or through the static
// scratch of the context's device (allocates nothing; at most
// thermo_static_max_particles() particles)
void thermo_reduce(Ctx& c, double out[8]);
void thermo_reduce_static(Ctx& c, double out[8]);
void init_thermo_static_scratch(int device);
int64_t thermo_static_max_particles();

inline int blocks_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return int(b > 0 ? b : 1);
}

// Copies and fills on the context's own stream, finished before returning.
// The stream is non-blocking, so nothing here may rely on the legacy
// stream's implicit ordering (and other threads' legacy-stream work never
// waits on this stream, e.g. while it records a graph).
inline void h2d(Ctx& c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    MDD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
    MDD_HIP(hipStreamSynchronize(c.stream));
}
inline void d2h(Ctx& c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    MDD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    MDD_HIP(hipStreamSynchronize(c.stream));
}
inline void dzero(Ctx& c, void* dst, size_t bytes) {
    if (!bytes) return;
    MDD_HIP(hipMemsetAsync(dst, 0, bytes, c.stream));
    MDD_HIP(hipStreamSynchronize(c.stream));
}

template <typename V>
inline void sig_put(std::vector<uint8_t>& s, const V& v) {
    const uint8_t* b = reinterpret_cast<const uint8_t*>(&v);
    s.insert(s.end(), b, b + sizeof(V));
}

// Launch kernel k on the context's stream, or record the launch (c.rec.on).
template <typename... P, typename... A>
inline void launch(Ctx& c, void (*k)(P...), dim3 grid, dim3 block, A&&... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    if (!c.rec.on) {
        k<<<grid, block, 0, c.stream>>>(std::forward<A>(a)...);
        return;
    }
    auto args = std::make_shared<std::tuple<std::decay_t<P>...>>(
        static_cast<std::decay_t<P>>(std::forward<A>(a))...);
    sig_put(c.rec.sig, reinterpret_cast<uintptr_t>(k));
    const uint32_t dims[6] = {grid.x, grid.y, grid.z, block.x, block.y, block.z};
    sig_put(c.rec.sig, dims);
    std::apply([&](const auto&... v) { (sig_put(c.rec.sig, v), ...); }, *args);
    c.rec.calls.push_back([k, grid, block, args](hipStream_t s) {
        void* argv[sizeof...(P) + 1] = {};
        size_t i = 0;
        std::apply([&](auto&... v) { ((argv[i++] = static_cast<void*>(&v)), ...); }, *args);
        MDD_HIP(hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, argv, 0, s));
    });
}

}  // namespace mdd
