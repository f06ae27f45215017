// Pair-force kernels of the gfx950 particle path: Lennard-Jones 12-6 and
// fused Hertz-Mindlin contact with tangential shear history, plus plane walls
// and gravity.
//
// A group of G = 2^LG lanes owns one particle and gathers over its full
// neighbor row (lane t takes entries t, t+G, ...), so forces accumulate in
// registers, are combined by a fixed shuffle tree and written once: no
// atomics, no reverse (image -> particle) pass, and bitwise run-to-run
// determinism. Every pair is evaluated in its owner's frame (the frame the
// reference half list evaluates it in), so the two particles of a pair see
// bitwise-equal, opposite pair forces.
//
// Precision: R = double reproduces the reference arithmetic (operation order
// and fused multiply-adds as written there). R = float ("mixed") keeps the
// FP64 pair geometry (image shift, separation, cutoff test, overlap, normal,
// relative velocity) and the FP64 shear spring, and does the force algebra
// and per-particle accumulation in FP32.
#include <hip/hip_ext.h>

#include "cells.hpp"
#include "ctx.hpp"

// row entries per lane gathered together by the LJ row kernel
#ifndef MDD_LJ_BATCH
#define MDD_LJ_BATCH 3
#endif

namespace mdd {

namespace {

// dem_atom/src/lib.rs:63
constexpr double kSqrt53 = 0.9128709291752768;
// dem_granular/src/lib.rs:93-108
constexpr double kTangEps = 1e-30;
constexpr double kLargeOverlap = 0.90;
constexpr int kMaxOverlapWarnings = 500;
// optional pair tables in ForceArgs::mext, each nm x nm (the C API's order)
constexpr int kXRolling = 0, kXTwisting = 1, kXCohesion = 2, kXSurface = 3, kXKn = 4, kXKt = 5,
              kXRollK = 6, kXRollG = 7, kXTwistK = 8, kXTwistG = 9;

struct ForceArgs {
    int64_t n, npad;
    int maxnb;
    const double4* pos;
    const double4* vel;
    const double4* omega;
    const uint32_t* tag;
    const uint32_t* type;
    const double* mass;
    const uint32_t* nbr;
    const uint32_t* nnb;
    const uint32_t* gcnt;  // periodic images per particle at the last build, or null
    double L[3];
    double lj1, lj2, cut2, e12, e6;
    const double* mat;
    int nm;
    const double* mext;  // optional pair tables (kX* order), each nm x nm, or null
    int hooke, dmt, roll_sds, twist_sds;
    double dt;
    double* hist;
    uint64_t* cmask;
    int64_t hist_cs;
    const WallsP* walls;  // wall table (device copy)
    WallsDyn* wdyn;       // current wall points / velocities and force sums
    int has_regions;      // region walls present: applied in their own pass
    int regions_only;     // that pass: region walls (then gravity) only
    int use_walls, use_gravity, use_pairs;
    double g[3];
    double* f64;
    float* f32;
    double* t64;
    float* t32;
    double* pe;
    double* vir;  // energy passes: per-particle half share of sum d (x) F, 6 x npad
                  // (LJ); DEM virial passes: the owned pairs' full share
    int dem_virial;  // DEM: this pass writes vir
    Status* st;
    int step;
    const SysP* sysp;     // batched ensemble: per-system box and materials, or null
    const uint32_t* psys;
    const int4* xq;            // LJ mixed: fixed-point positions, or null
    float qf[3];               // their unit L / 2^32 per axis, FP32 and FP64
    double qd[3];
    const uint32_t* listed;    // DEM: particles with a nonempty row (last rebuild)
    const uint32_t* listed_n;  // their count
};

__device__ inline bool gated(const ForceArgs& a) {
    return a.step >= 0 && a.st->abort_step <= a.step;
}

template <typename R>
__device__ inline R rcp(R x);
template <>
__device__ inline double rcp<double>(double x) {
    return 1.0 / x;
}
template <>
__device__ inline float rcp<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}

template <typename R>
__device__ inline R sqrt_r(R x);
template <>
__device__ inline double sqrt_r<double>(double x) {
    return sqrt(x);
}
template <>
__device__ inline float sqrt_r<float>(float x) {
    return __builtin_sqrtf(x);
}

template <int G, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int G>
__device__ inline uint64_t group_or(uint64_t v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor(uint32_t(v), o, 64);
        const uint32_t hi = __shfl_xor(uint32_t(v >> 32), o, 64);
        v |= (uint64_t(hi) << 32) | lo;
    }
    return v;
}

// VirialStress::add_pair (mddem_core/src/virial.rs:59-67): d = x_j - x_i,
// f = force on i from j; xx yy zz xy xz yz
__device__ inline void virial_add(double w[6], double dx, double dy, double dz, double fx,
                                  double fy, double fz) {
    w[0] += dx * fx;
    w[1] += dy * fy;
    w[2] += dz * fz;
    w[3] += dx * fy;
    w[4] += dx * fz;
    w[5] += dy * fz;
}

// partner (j) and own (i) periodic shift of a row entry, 64 B per entry
struct alignas(16) ShiftPair {
    double j[3], i[3], pad[2];
};

// Mixed LJ row of particle i (lane t of G) on fixed-point positions: one
// 16-byte gather per neighbor; the 32-bit coordinate differences are the
// minimum-image separation in units of L / 2^32, so image entries need no
// shift. FP32 pair arithmetic; a pair within 1e-6 of the cutoff (the FP32
// r^2 is good to ~3e-7) is decided on the FP64 r^2 of the same differences.
template <int LG>
__device__ inline void lj_row_fixed_force(const ForceArgs& a, int64_t i, int t, float& fx_out,
                                          float& fy_out, float& fz_out) {
    constexpr int G = 1 << LG;
    constexpr int kB = MDD_LJ_BATCH;
    int cnt = int(a.nnb[i]);
    if (cnt > a.maxnb) cnt = a.maxnb;
    const uint32_t* row = a.nbr + (i << LG) + t;
    const int64_t stride = a.npad << LG;
    const int4 qi = a.xq[i];
    const float sx = a.qf[0], sy = a.qf[1], sz = a.qf[2];
    const float cut2 = float(a.cut2), lj1 = float(a.lj1), lj2 = float(a.lj2);
    const float cut2_eps = 1e-6f * cut2;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    uint32_t en[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
        const int k = t + u * G;
        en[u] = k < cnt ? row[int64_t(k >> LG) * stride] : 0u;
    }
    for (int kb = t; kb < cnt; kb += kB * G) {
        int4 qj[kB];
#pragma unroll
        for (int u = 0; u < kB; u++) qj[u] = a.xq[entry_idx(en[u])];
#pragma unroll
        for (int u = 0; u < kB; u++) {
            const int k = kb + (kB + u) * G;
            en[u] = k < cnt ? row[int64_t(k >> LG) * stride] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kB; u++) {
            const int ix = int(uint32_t(qj[u].x) - uint32_t(qi.x));
            const int iy = int(uint32_t(qj[u].y) - uint32_t(qi.y));
            const int iz = int(uint32_t(qj[u].z) - uint32_t(qi.z));
            const float dx = float(ix) * sx, dy = float(iy) * sy, dz = float(iz) * sz;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const bool valid = kb + u * G < cnt;
            // the sign of the rounded difference is the sign of r2 - cut2
            const float dr = r2 - cut2;
            bool in = valid && dr < 0.0f;
            if (__builtin_expect(fabsf(dr) <= cut2_eps, 0)) {
                const double ex = double(ix) * a.qd[0], ey = double(iy) * a.qd[1],
                             ez = double(iz) * a.qd[2];
                in = valid && fma(ex, ex, fma(ey, ey, ez * ez)) < a.cut2;
            }
            const float r2inv = __builtin_amdgcn_rcpf(r2);
            const float r6inv = r2inv * r2inv * r2inv;
            const float fpair = r2inv * r6inv * fmaf(lj1, r6inv, -lj2);
            const float sc = in ? -fpair : 0.0f;
            fx = fmaf(sc, dx, fx);
            fy = fmaf(sc, dy, fy);
            fz = fmaf(sc, dz, fz);
        }
    }
    fx_out = group_sum<G>(fx);
    fy_out = group_sum<G>(fy);
    fz_out = group_sum<G>(fz);
}

template <int LG>
__device__ inline void lj_row_fixed(const ForceArgs& a, int64_t i, int t) {
    float fx, fy, fz;
    lj_row_fixed_force<LG>(a, i, t, fx, fy, fz);
    if (t != 0) return;
    a.f32[i] = fx;
    a.f32[a.npad + i] = fy;
    a.f32[2 * a.npad + i] = fz;
}

// The thermostatted final half step (k_post of integrate.hip) fused into
// the mixed LJ row pass: each particle's kick with its new force and this
// step's thermostat rescale right after the force, and the kinetic-energy
// partial of each 256 particles summed exactly as k_post sums it (one
// particle per thread, xor butterfly per wavefront, the four wavefront sums
// in order), so the next first-half kernel reads the same partials bit for
// bit. One workgroup per 256 particles, 1 << LG lanes each.
struct PostArgs {
    double4* vel;
    const double* mass;
    double* partial;
    int q;  // thermostat slot of this step
    double dt;
};

template <int LG>
__global__ void __launch_bounds__(256 << LG) k_lj_post(ForceArgs a, PostArgs p) {
    constexpr int G = 1 << LG;
    __shared__ double ke_lds[256];
    __shared__ double lds4[4];
    if (gated(a)) return;
    const int64_t pb = xcd_block(blockIdx.x, gridDim.x);
    const int local = int(threadIdx.x >> LG);
    const int t = int(threadIdx.x & (G - 1));
    const int64_t i = pb * 256 + local;
    double ke = 0.0;
    if (i < a.n) {
        float fx, fy, fz;
        lj_row_fixed_force<LG>(a, i, t, fx, fy, fz);
        if (t == 0) {
            a.f32[i] = fx;
            a.f32[a.npad + i] = fy;
            a.f32[2 * a.npad + i] = fz;
            double4 v = p.vel[i];
            const double h = 0.5 * p.dt * v.w;
            v.x += h * double(fx);
            v.y += h * double(fy);
            v.z += h * double(fz);
            const double scale = a.st->nh_scale[p.q];
            v.x *= scale;
            v.y *= scale;
            v.z *= scale;
            p.vel[i] = v;
            ke = p.mass[i] * (v.x * v.x + v.y * v.y + v.z * v.z);
        }
    }
    if (t == 0) ke_lds[local] = ke;
    __syncthreads();
    if (threadIdx.x < 256u) {
        double v = ke_lds[threadIdx.x];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63u) == 0u) lds4[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) p.partial[pb] = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
}

// lj_force (md_lj/src/lib.rs:488-633): for each neighbor inside the cutoff,
// fpair = r^-2 r^-6 (lj1 r^-6 - lj2), owner gets fma(-fpair, d, f),
// partner fma(fpair, d, f), d = partner image - owner.
//
// Seen from particle i, with the periodic shift S applied to the non-owner:
//   i owns the pair:  d  = (x_j + S) - x_i,  F_i = -fpair d
//   j owns the pair:  d' = (x_i + S) - x_j,  F_i = +fpair d'
// and since fl(a - b) = -fl(b - a), d'' = x_j - (x_i + S) = -d' exactly, so
// F_i = -fpair d'' in both cases: d = (x_j + S_j) - (x_i + S_i) with S on
// the non-owner. Without an image both shifts vanish and d = x_j - x_i.
// Pairs outside the cutoff contribute fma(0, d, f) = f.
template <typename R, bool ENERGY, int LG>
__global__ void __launch_bounds__(256) k_lj(ForceArgs a) {
    constexpr int G = 1 << LG;
    // XCD-aware block order: each XCD's L2 holds the positions of one
    // contiguous (spatially sorted) range of particles and its neighbors
    const int64_t gt = int64_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const int64_t i = gt >> LG;
    const int t = int(gt & (G - 1));
    // The periodic shifts of a row entry, by its top six bits (image code
    // and owner flag): the shift goes on the non-owner, so entry bits
    // (code, rev) give the partner's shift j and this particle's shift i,
    // one of them s = code_shifts(code) and the other +0 (exact values).
    __shared__ ShiftPair shp[64];
    if (threadIdx.x < 64u) {
        const int code = int(threadIdx.x & 31u);
        const bool rev = threadIdx.x >= 32u;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (code < 27) code_shifts(code, a.L[0], a.L[1], a.L[2], sx, sy, sz);
        ShiftPair q;
        q.j[0] = rev ? 0.0 : sx;
        q.j[1] = rev ? 0.0 : sy;
        q.j[2] = rev ? 0.0 : sz;
        q.i[0] = rev ? sx : 0.0;
        q.i[1] = rev ? sy : 0.0;
        q.i[2] = rev ? sz : 0.0;
        q.pad[0] = q.pad[1] = 0.0;
        shp[threadIdx.x] = q;
    }
    __syncthreads();
    if (i >= a.n) return;
    if (gated(a)) return;
    if constexpr (sizeof(R) == 4 && !ENERGY) {
        if (a.xq) {
            lj_row_fixed<LG>(a, i, t);
            return;
        }
    }
    const double4 xi = a.pos[i];
    int cnt = int(a.nnb[i]);
    if (cnt > a.maxnb) cnt = a.maxnb;
    R fx = R(0), fy = R(0), fz = R(0);
    double pe = 0.0;
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // virial (xx yy zz xy xz yz)
    const R lj1 = R(a.lj1), lj2 = R(a.lj2);
    const uint32_t* row = a.nbr + (i << LG) + t;
    const int64_t stride = a.npad << LG;
    // a row holds periodic-image entries only if its particle had images at
    // the last build; wavefronts without any take the shift-free loop
    const bool img = a.gcnt == nullptr || a.gcnt[i] != 0u;
    auto pair = [&](bool valid, double dx, double dy, double dz) {
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const bool in = valid && r2 < a.cut2;
        const R r2inv = rcp<R>(R(r2));
        const R r6inv = r2inv * r2inv * r2inv;
        const R fpair = r2inv * r6inv * fma(lj1, r6inv, -lj2);
        const R s = in ? -fpair : R(0);
        fx = fma(s, R(dx), fx);
        fy = fma(s, R(dy), fy);
        fz = fma(s, R(dz), fz);
        if (ENERGY && in) {
            const double r6 = double(r6inv);
            pe += 0.5 * (r6 * (a.e12 * r6 - a.e6));
            // virial (md_lj/src/lib.rs:555-561): d (x) F per pair, half on
            // each side of the full rows (both sides see the same d (x) F)
            const double sd = double(s);
            w[0] += 0.5 * (dx * (sd * dx));
            w[1] += 0.5 * (dy * (sd * dy));
            w[2] += 0.5 * (dz * (sd * dz));
            w[3] += 0.5 * (dx * (sd * dy));
            w[4] += 0.5 * (dx * (sd * dz));
            w[5] += 0.5 * (dy * (sd * dz));
        }
    };
    // Batches of kB entries per lane: the batch's kB gathers are issued
    // together, and the next batch's row entries are loaded while they are
    // in flight, so a wave waits on memory once per batch. The gathers are
    // unconditional (an entry past the row is 0 and reads particle 0; its
    // pair is masked out of the sum): a gather under a branch makes the
    // compiler wait for it inside the branch, one gather at a time.
    constexpr int kB = MDD_LJ_BATCH;
    uint32_t en[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
        const int k = t + u * G;
        en[u] = k < cnt ? row[int64_t(k >> LG) * stride] : 0u;
    }
    const bool any_img = __ballot(img) != 0ull;
    for (int kb = t; kb < cnt; kb += kB * G) {
        uint32_t e[kB];
        double px[kB], py[kB], pz[kB];
#pragma unroll
        for (int u = 0; u < kB; u++) {
            e[u] = en[u];
            const double4 xj = a.pos[entry_idx(e[u])];
            px[u] = xj.x;
            py[u] = xj.y;
            pz[u] = xj.z;
        }
#pragma unroll
        for (int u = 0; u < kB; u++) {
            const int k = kb + (kB + u) * G;
            en[u] = k < cnt ? row[int64_t(k >> LG) * stride] : 0u;
        }
        if (!any_img) {
#pragma unroll
            for (int u = 0; u < kB; u++)
                pair(kb + u * G < cnt, px[u] - xi.x, py[u] - xi.y, pz[u] - xi.z);
        } else {
#pragma unroll
            for (int u = 0; u < kB; u++) {
                const ShiftPair& q = shp[e[u] >> kIdxBits];
                pair(kb + u * G < cnt, (px[u] + q.j[0]) - (xi.x + q.i[0]),
                     (py[u] + q.j[1]) - (xi.y + q.i[1]), (pz[u] + q.j[2]) - (xi.z + q.i[2]));
            }
        }
    }
    fx = group_sum<G>(fx);
    fy = group_sum<G>(fy);
    fz = group_sum<G>(fz);
    if (ENERGY) {
        pe = group_sum<G>(pe);
#pragma unroll
        for (int k = 0; k < 6; k++) w[k] = group_sum<G>(w[k]);
    }
    if (t != 0) return;
    if constexpr (sizeof(R) == 8) {
        a.f64[i] = fx;
        a.f64[a.npad + i] = fy;
        a.f64[2 * a.npad + i] = fz;
    } else {
        a.f32[i] = fx;
        a.f32[a.npad + i] = fy;
        a.f32[2 * a.npad + i] = fz;
    }
    if (ENERGY) {
        a.pe[i] = pe;
#pragma unroll
        for (int k = 0; k < 6; k++) a.vir[k * a.npad + i] = w[k];
    }
}

template <typename R>
__device__ inline void store_force(const ForceArgs& a, int64_t i, R fx, R fy, R fz, R tx, R ty,
                                   R tz) {
    if constexpr (sizeof(R) == 8) {
        a.f64[i] = fx;
        a.f64[a.npad + i] = fy;
        a.f64[2 * a.npad + i] = fz;
        a.t64[i] = tx;
        a.t64[a.npad + i] = ty;
        a.t64[2 * a.npad + i] = tz;
    } else {
        a.f32[i] = fx;
        a.f32[a.npad + i] = fy;
        a.f32[2 * a.npad + i] = fz;
        a.t32[i] = tx;
        a.t32[a.npad + i] = ty;
        a.t32[2 * a.npad + i] = tz;
    }
}

// wall_contact_force (dem_wall/src/lib.rs:848-1241) for one particle: plane
// walls (JKR / DMT adhesion, SJKR cohesion, twisting friction, moving
// walls), then cylinder and sphere walls (Hertz, SJKR cohesion). FP64 pair
// arithmetic; each wall's scalar force joins its fixed-point sum.
__device__ inline double curved_wall_force(const ForceArgs& a, int k, double radius,
                                           double mass, const double4& v, double delta,
                                           double nx, double ny, double nz) {
    const double PI = 3.141592653589793;
    const int nn = a.nm * a.nm;
    const double e_eff = a.mat[2 * nn + k];
    const double sdr = sqrt(delta * radius);
    const double k_n = 4.0 / 3.0 * e_eff * sdr;
    const double s_n = 2.0 * e_eff * sdr;
    const double v_n = v.x * nx + v.y * ny + v.z * nz;
    const double beta = a.mat[k];
    const double cohesion = a.mext ? a.mext[kXCohesion * nn + k] : 0.0;
    const double f_diss = 2.0 * beta * kSqrt53 * sqrt(s_n * mass) * v_n;
    if (cohesion > 0.0) return k_n * delta - f_diss - cohesion * PI * delta * radius;
    const double f = k_n * delta - f_diss;
    return f > 0.0 ? f : 0.0;
}

// The wall table and the walls' current points and velocities do not change
// while a force kernel runs (the host and the integrator kernels write them).
// Read through the constant address space, their wavefront-uniform fields
// come in scalar loads even though the kernel stores forces and wall sums;
// through generic pointers every field was a dependent vector load per wall.
typedef const __attribute__((address_space(4))) WallsP CWallsP;
typedef const __attribute__((address_space(4))) Wall CWall;
typedef const __attribute__((address_space(4))) WallDyn CWallDyn;
__device__ inline CWallsP& wall_table(const ForceArgs& a) { return *(CWallsP*)a.walls; }
__device__ inline CWallDyn& wall_state(const ForceArgs& a, int w) {
    return *(CWallDyn*)&a.wdyn->w[w];
}

// Region::closest_point_on_surface (mddem_core/src/region.rs:260-772):
// signed distance (positive outside), surface point, outward normal
struct Surf {
    double pt[3], nrm[3], d;
};

__device__ inline double clamp_d(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// (axial, offsets from the centre) for an axis, and back
__device__ inline void axis_split(const double* x, const double* c, int ax, double& a,
                                  double& d0, double& d1) {
    a = x[ax];
    d0 = x[ax == 0 ? 1 : 0] - c[0];
    d1 = x[ax == 2 ? 1 : 2] - c[1];
}
__device__ inline void axis_join(double a, double r0, double r1, const double* c, int ax,
                                 double* out) {
    out[ax] = a;
    out[ax == 0 ? 1 : 0] = c[0] + r0;
    out[ax == 2 ? 1 : 2] = c[1] + r1;
}
__device__ inline void axis_join_n(double a, double n0, double n1, int ax, double* out) {
    out[ax] = a;
    out[ax == 0 ? 1 : 0] = n0;
    out[ax == 2 ? 1 : 2] = n1;
}

__device__ inline Surf surf_from(const double* pt, const double* nrm, double d) {
    Surf r;
    for (int q = 0; q < 3; q++) {
        r.pt[q] = pt[q];
        r.nrm[q] = nrm[q];
    }
    r.d = d;
    return r;
}

// a point outside: the candidate's point, the normal along x - point
__device__ inline Surf surf_outside(const double* x, const double* pt, const double* nrm) {
    const double dx = x[0] - pt[0], dy = x[1] - pt[1], dz = x[2] - pt[2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < 1e-30) return surf_from(pt, nrm, 0.0);
    const double n[3] = {dx / dist, dy / dist, dz / dist};
    return surf_from(pt, n, dist);
}

__device__ inline double norm_diff(const double* a, const double* b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ Surf surf_block(const double* x, const double* mn, const double* mx) {
    const bool in = x[0] >= mn[0] && x[0] <= mx[0] && x[1] >= mn[1] && x[1] <= mx[1] &&
                    x[2] >= mn[2] && x[2] <= mx[2];
    if (!in) {
        const double cl[3] = {clamp_d(x[0], mn[0], mx[0]), clamp_d(x[1], mn[1], mx[1]),
                              clamp_d(x[2], mn[2], mx[2])};
        const double up[3] = {0.0, 0.0, 1.0};
        return surf_outside(x, cl, up);
    }
    double best = x[0] - mn[0];
    int bi = 0;
    for (int k = 1; k < 6; k++) {
        const int axk = k >> 1;
        const double dk = (k & 1) ? mx[axk] - x[axk] : x[axk] - mn[axk];
        if (dk < best) {
            best = dk;
            bi = k;
        }
    }
    const int ax = bi >> 1;
    double pt[3] = {x[0], x[1], x[2]}, n[3] = {0.0, 0.0, 0.0};
    pt[ax] = (bi & 1) ? mx[ax] : mn[ax];
    n[ax] = (bi & 1) ? 1.0 : -1.0;
    return surf_from(pt, n, -best);
}

__device__ Surf surf_sphere(const double* x, const double* c, double rad) {
    const double dx = x[0] - c[0], dy = x[1] - c[1], dz = x[2] - c[2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < 1e-30) {
        const double pt[3] = {c[0] + rad, c[1], c[2]}, n[3] = {1.0, 0.0, 0.0};
        return surf_from(pt, n, -rad);
    }
    const double inv = 1.0 / dist;
    const double n[3] = {dx * inv, dy * inv, dz * inv};
    const double pt[3] = {c[0] + n[0] * rad, c[1] + n[1] * rad, c[2] + n[2] * rad};
    return surf_from(pt, n, dist - rad);
}

__device__ Surf surf_cylinder(const double* x, const double* c, double rad, int ax, double lo,
                              double hi) {
    double a, d0, d1;
    axis_split(x, c, ax, a, d0, d1);
    const double rd = sqrt(d0 * d0 + d1 * d1);
    const double ac = clamp_d(a, lo, hi);
    const bool a_in = a >= lo && a <= hi;
    double cpt[3], cn[3], cdist;
    if (rd < 1e-30) {
        axis_join(ac, rad, 0.0, c, ax, cpt);
        axis_join_n(0.0, 1.0, 0.0, ax, cn);
        const double da = a < lo ? lo - a : a - hi;
        cdist = a_in ? -rad : sqrt(rad * rad + da * da);
    } else {
        const double inv = 1.0 / rd;
        axis_join(ac, d0 * inv * rad, d1 * inv * rad, c, ax, cpt);
        axis_join_n(0.0, d0 * inv, d1 * inv, ax, cn);
        if (a_in) {
            cdist = rd - rad;
        } else {
            const double da = a < lo ? lo - a : a - hi;
            const double dr = rd - rad;
            cdist = sqrt(dr * dr + da * da) * ((dr < 0.0 && da == 0.0) ? -1.0 : 1.0);
        }
    }
    const double cap_r = rd <= rad ? rd : rad;
    double c0 = 0.0, c1 = 0.0;
    if (!(rd < 1e-30)) {
        const double inv = 1.0 / rd;
        c0 = d0 * inv * cap_r;
        c1 = d1 * inv * cap_r;
    }
    double lpt[3], ln[3], hpt[3], hn[3];
    axis_join(lo, c0, c1, c, ax, lpt);
    axis_join_n(-1.0, 0.0, 0.0, ax, ln);
    axis_join(hi, c0, c1, c, ax, hpt);
    axis_join_n(1.0, 0.0, 0.0, ax, hn);
    const bool in = a_in && rd <= rad;
    if (in) {
        const double dc = fabs(rad - rd), dl = a - lo, dh = hi - a;
        if (dc <= dl && dc <= dh) return surf_from(cpt, cn, rd - rad);
        if (dl <= dh) return surf_from(lpt, ln, -(a - lo));
        return surf_from(hpt, hn, -(hi - a));
    }
    const double el = norm_diff(x, lpt), eh = norm_diff(x, hpt);
    const double lo_d = a < lo ? (a <= lo ? lo - a : el) : el;
    const double hi_d = a > hi ? (a >= hi ? a - hi : eh) : eh;
    const double k0 = fabs(cdist);
    int best = 0;
    double bk = k0;
    if (lo_d < bk) {
        best = 1;
        bk = lo_d;
    }
    if (hi_d < bk) best = 2;
    if (best == 0) return surf_outside(x, cpt, cn);
    if (best == 1) return surf_outside(x, lpt, ln);
    return surf_outside(x, hpt, hn);
}

__device__ Surf surf_cone(const double* x, const double* c, int ax, double rlo, double rhi,
                          double lo, double hi) {
    double a, d0, d1;
    axis_split(x, c, ax, a, d0, d1);
    const double rd = sqrt(d0 * d0 + d1 * d1);
    const double h = hi - lo;
    double n0 = 1.0, n1 = 0.0;
    if (!(rd < 1e-30)) {
        n0 = d0 / rd;
        n1 = d1 / rd;
    }
    const bool a_in = a >= lo && a <= hi;
    const double r_at =
        fabs(h) < 1e-30 ? (rlo + rhi) * 0.5 : rlo + clamp_d((a - lo) / h, 0.0, 1.0) * (rhi - rlo);
    const bool in = a_in && rd <= r_at;
    // lateral surface: nearest point of the segment (lo, rlo)-(hi, rhi) in
    // (axial, radial) coordinates
    const double sa = hi - lo, sr = rhi - rlo;
    const double s2 = sa * sa + sr * sr;
    const double t = s2 < 1e-30 ? 0.5 : clamp_d(((a - lo) * sa + (rd - rlo) * sr) / s2, 0.0, 1.0);
    const double pa = lo + t * sa, pr = rlo + t * sr;
    const double da = a - pa, dr = rd - pr;
    const double d2 = sqrt(da * da + dr * dr);
    double cpt[3], cn[3];
    axis_join(pa, n0 * pr, n1 * pr, c, ax, cpt);
    double na = -sr, nr = sa;
    const double nl = sqrt(na * na + nr * nr);
    if (nl > 1e-30) {
        na /= nl;
        nr /= nl;
    } else {
        na = 0.0;
        nr = 1.0;
    }
    axis_join_n(na, nr * n0, nr * n1, ax, cn);
    const double rr = rd > 0.0 ? rd : 0.0;
    const double lcr = rlo < rr ? rlo : rr, hcr = rhi < rr ? rhi : rr;
    double lpt[3], ln[3], hpt[3], hn[3];
    axis_join(lo, n0 * lcr, n1 * lcr, c, ax, lpt);
    axis_join_n(-1.0, 0.0, 0.0, ax, ln);
    axis_join(hi, n0 * hcr, n1 * hcr, c, ax, hpt);
    axis_join_n(1.0, 0.0, 0.0, ax, hn);
    if (in) {
        const double dl = a - lo, dh = hi - a;
        if (d2 <= dl && d2 <= dh) return surf_from(cpt, cn, -d2);
        if (dl <= dh) return surf_from(lpt, ln, -dl);
        return surf_from(hpt, hn, -dh);
    }
    const double el = norm_diff(x, lpt), eh = norm_diff(x, hpt);
    int best = 0;
    double bk = d2;
    if (el < bk) {
        best = 1;
        bk = el;
    }
    if (eh < bk) best = 2;
    if (best == 0) return surf_outside(x, cpt, cn);
    if (best == 1) return surf_outside(x, lpt, ln);
    return surf_outside(x, hpt, hn);
}

__device__ Surf surf_plane(const double* x, const double* p0, const double* nv) {
    const double mag = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    double n[3] = {0.0, 0.0, 1.0};
    if (mag > 1e-30)
        for (int q = 0; q < 3; q++) n[q] = nv[q] / mag;
    const double d = (x[0] - p0[0]) * n[0] + (x[1] - p0[1]) * n[1] + (x[2] - p0[2]) * n[2];
    const double pt[3] = {x[0] - d * n[0], x[1] - d * n[1], x[2] - d * n[2]};
    return surf_from(pt, n, d);
}

__device__ Surf surf_prim(int type, const double* q, const double* x) {
    switch (type) {
        case 0: return surf_block(x, q, q + 3);
        case 1: return surf_sphere(x, q, q[3]);
        case 2: return surf_cylinder(x, q, q[2], int(q[3]), q[4], q[5]);
        case 3: return surf_cone(x, q, int(q[2]), q[3], q[4], q[5], q[6]);
        default: return surf_plane(x, q, q + 3);
    }
}

// nearest leaf surface, the first of equals (Union / Intersect, region.rs:282-309)
__device__ Surf surf_region(CWall& W, const double* x) {
    Surf best;
    for (int k = 0; k < W.nprim; k++) {
        double q[12];
        for (int j = 0; j < 12; j++) q[j] = W.prm[k][j];
        const Surf r = surf_prim(W.ptype[k], q, x);
        if (k == 0 || fabs(r.d) < fabs(best.d)) best = r;
    }
    return best;
}

// a wall's force sum: the lanes of a wavefront in contact with it are summed
// in fixed point by the scalar unit and the lowest active lane adds the total
// (one atomic per wavefront, where thousands of spheres may rest on one wall)
__device__ inline void wall_acc(const ForceArgs& a, int w, bool on, double f) {
    const long long v = on ? __double2ll_rn(f * kWallAccScale) : 0ll;
    uint64_t m = __ballot(v != 0ll);
    if (m == 0ull) return;
    const uint64_t act = __ballot(true);
    long long sum = 0;
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1ull;
        const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), l));
        const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(uint64_t(v) >> 32)), l));
        sum += static_cast<long long>((uint64_t(hi) << 32) | lo);
    }
    if (int(threadIdx.x & 63u) == __builtin_ctzll(act))
        atomicAdd(&a.wdyn->w[w].acc, static_cast<unsigned long long>(sum));
}

template <typename R>
__device__ inline void wall_forces(const ForceArgs& a, const double4& p, const double4* vp,
                                   const double4* om, double mass, uint32_t ti, R& fx, R& fy,
                                   R& fz, R& tx, R& ty, R& tz) {
    // the velocity is read only by particles touching a wall
    const double PI = 3.141592653589793;
    const int nm = a.nm, nn = nm * nm;
    const double radius = p.w;
    CWallsP& WS = wall_table(a);
    for (int w = 0; w < WS.count; w++) {
        CWall& W = WS.w[w];
        if (W.kind != 0 || !W.active) continue;
        double f_net = 0.0;
        bool on = false;
        do {
            if (!(p.x >= W.blo[0] && p.x <= W.bhi[0] && p.y >= W.blo[1] && p.y <= W.bhi[1] &&
                  p.z >= W.blo[2] && p.z <= W.bhi[2]))
                break;
            CWallDyn& D = wall_state(a, w);
            const double dx = p.x - D.p[0], dy = p.y - D.p[1], dz = p.z - D.p[2];
            const double distance = dx * W.n[0] + dy * W.n[1] + dz * W.n[2];
            if (distance <= 0.0) break;
            const int k = int(ti) * nm + W.mat;
            const double e_eff = a.mat[2 * nn + k];
            const double surface = a.mext ? a.mext[kXSurface * nn + k] : 0.0;
            double d_po = 0.0;
            if (surface > 0.0 && !a.dmt)
                d_po = cbrt(PI * PI * surface * surface * radius / (4.0 * e_eff * e_eff));
            double delta = radius - distance;
            if (0.5 * radius < delta) delta = 0.5 * radius;
            if (delta <= 0.0 && surface <= 0.0) break;
            if (delta < -d_po) break;
            const bool adhesion_only = surface > 0.0 && !a.dmt && delta <= 0.0;
            double s_n = 0.0, k_n = 0.0;
            if (delta > 0.0) {
                const double sdr = sqrt(delta * radius);
                s_n = 2.0 * e_eff * sdr;
                k_n = 4.0 / 3.0 * e_eff * sdr;
            }
            const double4 v = *vp;
            const double v_n =
                (v.x - D.v[0]) * W.n[0] + (v.y - D.v[1]) * W.n[1] + (v.z - D.v[2]) * W.n[2];
            const double beta = a.mat[k];
            const double cohesion = a.mext ? a.mext[kXCohesion * nn + k] : 0.0;
            const double f_diss = 2.0 * beta * kSqrt53 * sqrt(s_n * mass) * v_n;
            on = true;
            if (surface > 0.0 && a.dmt) {
                f_net = k_n * delta - f_diss - 2.0 * PI * surface * radius;
            } else if (surface > 0.0) {
                const double f_adh = 1.5 * PI * surface * radius;
                f_net = adhesion_only ? -f_adh : k_n * delta - f_diss - f_adh;
            } else if (cohesion > 0.0) {
                f_net = k_n * delta - f_diss - cohesion * PI * delta * radius;
            } else {
                f_net = k_n * delta - f_diss;
                if (!(f_net > 0.0)) f_net = 0.0;
            }
            fx += R(f_net * W.n[0]);
            fy += R(f_net * W.n[1]);
            fz += R(f_net * W.n[2]);
            const double mu_tw = a.mext ? a.mext[kXTwisting * nn + k] : 0.0;
            if (delta > 0.0 && mu_tw > 0.0 && om) {
                const double twist = om->x * W.n[0] + om->y * W.n[1] + om->z * W.n[2];
                if (fabs(twist) > 1e-30) {
                    const double tau = (twist > 0.0 ? -1.0 : 1.0) * (mu_tw * fabs(f_net) * radius);
                    tx += R(tau * W.n[0]);
                    ty += R(tau * W.n[1]);
                    tz += R(tau * W.n[2]);
                }
            }
        } while (false);
        wall_acc(a, w, on, f_net);
    }
    for (int w = 0; w < WS.count; w++) {
        CWall& W = WS.w[w];
        if (W.kind != 1 || !W.active) continue;
        double f_net = 0.0;
        bool on = false;
        do {
            double axial, d0, d1;
            if (W.axis == 0) {
                axial = p.x;
                d0 = p.y - W.c[0];
                d1 = p.z - W.c[1];
            } else if (W.axis == 1) {
                axial = p.y;
                d0 = p.x - W.c[0];
                d1 = p.z - W.c[1];
            } else {
                axial = p.z;
                d0 = p.x - W.c[0];
                d1 = p.y - W.c[1];
            }
            if (axial < W.lo || axial > W.hi) break;
            const double radial = sqrt(d0 * d0 + d1 * d1);
            if (radial < 1e-30) break;
            const double inv_r = 1.0 / radial;
            const double delta0 =
                W.inside ? radius - (W.radius - radial) : radius - (radial - W.radius);
            const double n0 = W.inside ? -d0 * inv_r : d0 * inv_r;
            const double n1 = W.inside ? -d1 * inv_r : d1 * inv_r;
            const double nx = W.axis == 0 ? 0.0 : n0;
            const double ny = W.axis == 0 ? n0 : (W.axis == 1 ? 0.0 : n1);
            const double nz = W.axis == 2 ? 0.0 : n1;
            if (delta0 <= 0.0) break;
            const double delta = 0.5 * radius < delta0 ? 0.5 * radius : delta0;
            on = true;
            f_net =
                curved_wall_force(a, int(ti) * nm + W.mat, radius, mass, *vp, delta, nx, ny, nz);
            fx += R(f_net * nx);
            fy += R(f_net * ny);
            fz += R(f_net * nz);
        } while (false);
        wall_acc(a, w, on, f_net);
    }
    for (int w = 0; w < WS.count; w++) {
        CWall& W = WS.w[w];
        if (W.kind != 2 || !W.active) continue;
        double f_net = 0.0;
        bool on = false;
        do {
            const double dx = p.x - W.c[0], dy = p.y - W.c[1], dz = p.z - W.c[2];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < 1e-30) break;
            const double inv = 1.0 / dist;
            const double sg = W.inside ? -1.0 : 1.0;
            const double nx = sg * dx * inv, ny = sg * dy * inv, nz = sg * dz * inv;
            const double delta0 = W.inside ? radius - (W.radius - dist) : radius - (dist - W.radius);
            if (delta0 <= 0.0) break;
            const double delta = 0.5 * radius < delta0 ? 0.5 * radius : delta0;
            on = true;
            f_net =
                curved_wall_force(a, int(ti) * nm + W.mat, radius, mass, *vp, delta, nx, ny, nz);
            fx += R(f_net * nx);
            fy += R(f_net * ny);
            fz += R(f_net * nz);
        } while (false);
        wall_acc(a, w, on, f_net);
    }
}

// region-surface walls (dem_wall/src/lib.rs:1170-1231), the last wall type a
// particle sees; kept out of the pair kernels (a separate pass after them)
template <typename R>
__device__ void region_wall_forces(const ForceArgs& a, const double4& p, const double4* vp,
                                   double mass, uint32_t ti, R& fx, R& fy, R& fz) {
    const int nm = a.nm;
    const double radius = p.w;
    CWallsP& WS = wall_table(a);
    for (int w = 0; w < WS.count; w++) {
        CWall& W = WS.w[w];
        if (W.kind != 3 || !W.active) continue;
        double f_net = 0.0;
        bool on = false;
        do {
            const double x[3] = {p.x, p.y, p.z};
            const Surf sr = surf_region(W, x);
            const double gap = W.inside ? -sr.d : sr.d;
            const double delta0 = radius - gap;
            if (delta0 <= 0.0) break;
            const double delta = 0.5 * radius < delta0 ? 0.5 * radius : delta0;
            const double sg = W.inside ? -1.0 : 1.0;
            const double nx = sg * sr.nrm[0], ny = sg * sr.nrm[1], nz = sg * sr.nrm[2];
            on = true;
            f_net =
                curved_wall_force(a, int(ti) * nm + W.mat, radius, mass, *vp, delta, nx, ny, nz);
            fx += R(f_net * nx);
            fy += R(f_net * ny);
            fz += R(f_net * nz);
        } while (false);
        wall_acc(a, w, on, f_net);
    }
}

// One row entry's pair geometry in the owner frame (the particle that
// lists the pair without the rev flag): d = partner image - owner, r1 the
// owner's radius. Returns whether the pair is in contact (delta > 0), with
// the large-overlap warning raised for owned pairs (contact.rs:147-217).
struct PairGeom {
    double dx, dy, dz, distance, delta, r1, r2;
};

__device__ inline bool pair_geom(const ForceArgs& a, const double4& pi, const double4& pj,
                                 uint32_t e, double Lx, double Ly, double Lz, PairGeom& g) {
    const bool rev = entry_rev(e);
    double shx, shy, shz;
    code_shifts(entry_code(e), Lx, Ly, Lz, shx, shy, shz);
    if (rev) {
        g.dx = (pi.x + shx) - pj.x;
        g.dy = (pi.y + shy) - pj.y;
        g.dz = (pi.z + shz) - pj.z;
        g.r1 = pj.w;
        g.r2 = pi.w;
    } else {
        g.dx = (pj.x + shx) - pi.x;
        g.dy = (pj.y + shy) - pi.y;
        g.dz = (pj.z + shz) - pi.z;
        g.r1 = pi.w;
        g.r2 = pj.w;
    }
    const double dist_sq = g.dx * g.dx + g.dy * g.dy + g.dz * g.dz;
    const double sum_r = g.r1 + g.r2;
    if (dist_sq >= sum_r * sum_r) return false;
    g.distance = sqrt(dist_sq);
    if (g.distance == 0.0) return false;
    const double r_min = g.r1 < g.r2 ? g.r1 : g.r2;
    g.delta = sum_r - g.distance;
    if (0.5 * r_min < g.delta) g.delta = 0.5 * r_min;
    if (g.delta > 0.0 && !rev && g.distance / sum_r < kLargeOverlap) {
        int c = atomicAdd(&a.st->overlap_count, 1) + 1;
        if (c > kMaxOverlapWarnings) {
            atomicCAS(&a.st->error, 0, -3);
            if (a.step >= 0) atomicMin(&a.st->abort_step, a.step);
        }
    }
    return g.delta > 0.0;
}

// Hertz-Mindlin force of one pair in contact (contact.rs:218-424 with
// surface energy, cohesion and rolling / twisting friction all zero), seen
// from particle i of the row (entry e, rev: i is the partner): the normal
// and tangential forces on the owner, the torque on i, and the shear spring
// of the row slot, read when the slot held a contact (had) and written back.
template <typename R>
struct PairOut {
    R fnx, fny, fnz, ftx, fty, ftz, tqx, tqy, tqz;
};

template <typename R>
__device__ inline void contact_pair(const ForceArgs& a, const SysP* S, int64_t i, uint32_t ti,
                                    uint32_t e, const PairGeom& g, int64_t slot, bool had,
                                    PairOut<R>& o) {
    const uint32_t j = entry_idx(e);
    const bool rev = entry_rev(e);
    const int nm = a.nm;
    const double dx = g.dx, dy = g.dy, dz = g.dz, distance = g.distance, delta = g.delta;
    const double r1 = g.r1, r2 = g.r2, sum_r = r1 + r2;
    const double4 vi = a.vel[i], vj = a.vel[j];
    const double4 wi = a.omega[i], wj = a.omega[j];
    const uint32_t tagi = a.tag[i], tagj = a.tag[j];
    const uint32_t tj = a.type[j];
    // owner (O) / partner (P) quantities
    const double4& vO = rev ? vj : vi;
    const double4& vP = rev ? vi : vj;
    const double4& wO = rev ? wj : wi;
    const double4& wP = rev ? wi : wj;
    const uint32_t tagO = rev ? tagj : tagi;
    const uint32_t tagP = rev ? tagi : tagj;
    const int mO = int(rev ? tj : ti), mP = int(rev ? ti : tj);
    const int mij = mO * nm + mP;
    const R r_eff = R(r1) * R(r2) / R(sum_r);
    const R e_eff = R(S ? S->eeff : a.mat[2 * nm * nm + mij]);
    const R g_eff = R(S ? S->geff : a.mat[3 * nm * nm + mij]);
    const R beta = R(S ? S->beta : a.mat[mij]);
    const R mu = R(S ? S->mu : a.mat[nm * nm + mij]);
    const double inv_dist = 1.0 / distance;
    const double nx = dx * inv_dist, ny = dy * inv_dist, nz = dz * inv_dist;
    const R m_r = R(1.0) / (R(vO.w) + R(vP.w));
    const R sdr = sqrt_r<R>(R(delta) * r_eff);
    const R s_n = R(2.0) * e_eff * sdr;
    const R k_n = R(4.0 / 3.0) * e_eff * sdr;
    const R k_t = R(8.0) * g_eff * sdr;
    // contact-point velocities v + w x (+-r n), FP64
    const double r1nx = r1 * nx, r1ny = r1 * ny, r1nz = r1 * nz;
    const double vcix = vO.x + (wO.y * r1nz - wO.z * r1ny);
    const double vciy = vO.y + (wO.z * r1nx - wO.x * r1nz);
    const double vciz = vO.z + (wO.x * r1ny - wO.y * r1nx);
    const double r2nx = r2 * nx, r2ny = r2 * ny, r2nz = r2 * nz;
    const double vcjx = vP.x + (-wP.y * r2nz + wP.z * r2ny);
    const double vcjy = vP.y + (-wP.z * r2nx + wP.x * r2nz);
    const double vcjz = vP.z + (-wP.x * r2ny + wP.y * r2nx);
    const double vrx = vcjx - vcix, vry = vcjy - vciy, vrz = vcjz - vciz;
    const double v_n = vrx * nx + vry * ny + vrz * nz;
    const R f_diss_n = R(2.0) * beta * R(kSqrt53) * sqrt_r<R>(s_n * m_r) * R(v_n);
    R f_n_mag = k_n * R(delta) - f_diss_n;
    if (!(f_n_mag > R(0))) f_n_mag = R(0);
    o.fnx = f_n_mag * R(nx);
    o.fny = f_n_mag * R(ny);
    o.fnz = f_n_mag * R(nz);
    const double vtx = vrx - v_n * nx, vty = vry - v_n * ny, vtz = vrz - v_n * nz;
    // shear spring in the lower tag's frame (tangential.rs:23-26)
    const double sign = tagO < tagP ? 1.0 : -1.0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (had) {
        sx = sign * a.hist[slot];
        sy = sign * a.hist[a.hist_cs + slot];
        sz = sign * a.hist[2 * a.hist_cs + slot];
    }
    const double s_dot_n = sx * nx + sy * ny + sz * nz;
    sx -= s_dot_n * nx;
    sy -= s_dot_n * ny;
    sz -= s_dot_n * nz;
    sx += vtx * a.dt;
    sy += vty * a.dt;
    sz += vtz * a.dt;
    const R s_mag = R(sqrt(sx * sx + sy * sy + sz * sz));
    const R f_t_spring_mag = k_t * s_mag;
    const R f_t_max = mu * fabs(f_n_mag);
    if (f_t_spring_mag > f_t_max && f_t_spring_mag > R(kTangEps)) {
        const R scale = f_t_max / f_t_spring_mag;
        sx *= double(scale);
        sy *= double(scale);
        sz *= double(scale);
    }
    const R gamma_t = R(2.0) * R(kSqrt53) * beta * sqrt_r<R>(k_t * m_r);
    R ftx = k_t * R(sx) - gamma_t * R(vtx);
    R fty = k_t * R(sy) - gamma_t * R(vty);
    R ftz = k_t * R(sz) - gamma_t * R(vtz);
    const R f_t_mag = sqrt_r<R>(ftx * ftx + fty * fty + ftz * ftz);
    if (f_t_mag > f_t_max && f_t_mag > R(kTangEps)) {
        const R scale = f_t_max / f_t_mag;
        ftx *= scale;
        fty *= scale;
        ftz *= scale;
    }
    o.ftx = ftx;
    o.fty = fty;
    o.ftz = ftz;
    // torque on i: (r2 n) x F_t as the partner, (r1 n) x F_t as the owner
    if (rev) {
        o.tqx = R(r2ny) * ftz - R(r2nz) * fty;
        o.tqy = R(r2nz) * ftx - R(r2nx) * ftz;
        o.tqz = R(r2nx) * fty - R(r2ny) * ftx;
    } else {
        o.tqx = R(r1ny) * ftz - R(r1nz) * fty;
        o.tqy = R(r1nz) * ftx - R(r1nx) * ftz;
        o.tqz = R(r1nx) * fty - R(r1ny) * ftx;
    }
    a.hist[slot] = sign * sx;
    a.hist[a.hist_cs + slot] = sign * sy;
    a.hist[2 * a.hist_cs + slot] = sign * sz;
}

// adds one pair's forces to particle i's sums: the partner gets +F_n - F_t,
// the owner -F_n + F_t
template <typename R>
__device__ inline void add_pair(bool rev, const PairOut<R>& o, R& fx, R& fy, R& fz, R& tx, R& ty,
                                R& tz) {
    if (rev) {
        fx += o.fnx;
        fy += o.fny;
        fz += o.fnz;
        fx -= o.ftx;
        fy -= o.fty;
        fz -= o.ftz;
    } else {
        fx -= o.fnx;
        fy -= o.fny;
        fz -= o.fnz;
        fx += o.ftx;
        fy += o.fty;
        fz += o.ftz;
    }
    tx += o.tqx;
    ty += o.tqy;
    tz += o.tqz;
}

// the particle's box lengths: the context's, or its system's in a batched
// ensemble (with the system's single-material constants)
__device__ inline const SysP* particle_box(const ForceArgs& a, int64_t i, double& Lx, double& Ly,
                                          double& Lz) {
    Lx = a.L[0];
    Ly = a.L[1];
    Lz = a.L[2];
    const SysP* S = a.sysp ? &a.sysp[a.psys[i]] : nullptr;
    if (S) {
        Lx = S->box.L[0];
        Ly = S->box.L[1];
        Lz = S->box.L[2];
    }
    return S;
}

// walls, gravity and the store of one particle's sums (lane t == 0)
template <typename R>
__device__ inline void finish_particle(const ForceArgs& a, int64_t i, const double4& pi, uint32_t ti,
                                       R fx, R fy, R fz, R tx, R ty, R tz) {
    if (a.use_walls)
        wall_forces<R>(a, pi, &a.vel[i], a.omega ? &a.omega[i] : nullptr, a.mass[i], ti, fx, fy, fz,
                       tx, ty, tz);
    if (a.use_gravity) {
        const double m = a.mass[i];
        fx += R(m * a.g[0]);
        fy += R(m * a.g[1]);
        fz += R(m * a.g[2]);
    }
    store_force<R>(a, i, fx, fy, fz, tx, ty, tz);
}

// hertz_mindlin_contact_force (dem_granular/src/contact.rs:103-615) for
// one particle (lane t of its G lanes): its row in order, each pair in
// contact through contact_pair. VIR: also the contact virial
// (contact.rs:559-569): each particle sums d (x) F over the pairs it owns,
// d = partner image - owner, F = -F_n + F_t the force on the owner, into
// vir[q * npad + i] (xx yy zz xy xz yz).
template <typename R, int LG, bool VIR>
__device__ inline void hertz_particle(const ForceArgs& a, int64_t i, int t) {
    constexpr int G = 1 << LG;
    const double4 pi = a.pos[i];
    R fx = R(0), fy = R(0), fz = R(0), tx = R(0), ty = R(0), tz = R(0);
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
    const uint32_t ti = a.type[i];
    double Lx, Ly, Lz;
    const SysP* S = particle_box(a, i, Lx, Ly, Lz);
    if (a.use_pairs) {
        int cnt = int(a.nnb[i]);
        if (cnt > a.maxnb) cnt = a.maxnb;
        // contact bits only ever cover row slots: an empty row has none, so
        // a particle without neighbors reads and writes no contact state
        uint64_t mask_in = 0ull;
        if (cnt > 0) mask_in = a.cmask[i];
        uint64_t mask_out = 0ull;
        // the next neighbor's position and the entry after it are loaded
        // while this one is computed (entries past the row read entry and
        // particle 0 and are not used); the partner's velocity and spin are
        // read only for pairs in contact, so they are not fetched ahead
        uint32_t e_n = t < cnt ? a.nbr[row_addr(i, t, a.npad, LG)] : 0u;
        uint32_t e_nn = t + G < cnt ? a.nbr[row_addr(i, t + G, a.npad, LG)] : 0u;
        double4 pj_n = a.pos[entry_idx(e_n)];
        for (int k = t; k < cnt; k += G) {
            const uint32_t e = e_n;
            const double4 pj = pj_n;
            e_n = e_nn;
            pj_n = a.pos[entry_idx(e_n)];
            e_nn = k + 2 * G < cnt ? a.nbr[row_addr(i, k + 2 * G, a.npad, LG)] : 0u;
            PairGeom g;
            if (!pair_geom(a, pi, pj, e, Lx, Ly, Lz, g)) continue;
            PairOut<R> o;
            contact_pair<R>(a, S, i, ti, e, g, row_addr(i, k, a.npad, LG), (mask_in >> k) & 1ull,
                            o);
            if constexpr (VIR) {
                if (!entry_rev(e)) {
                    const double vfx = double(-o.fnx + o.ftx), vfy = double(-o.fny + o.fty),
                                 vfz = double(-o.fnz + o.ftz);
                    w0 += g.dx * vfx;
                    w1 += g.dy * vfy;
                    w2 += g.dz * vfz;
                    w3 += g.dx * vfy;
                    w4 += g.dx * vfz;
                    w5 += g.dy * vfz;
                }
            }
            add_pair<R>(entry_rev(e), o, fx, fy, fz, tx, ty, tz);
            mask_out |= 1ull << k;
        }
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        tx = group_sum<G>(tx);
        ty = group_sum<G>(ty);
        tz = group_sum<G>(tz);
        mask_out = group_or<G>(mask_out);
        if (t == 0 && mask_out != mask_in) a.cmask[i] = mask_out;
        if constexpr (VIR) {
            w0 = group_sum<G>(w0);
            w1 = group_sum<G>(w1);
            w2 = group_sum<G>(w2);
            w3 = group_sum<G>(w3);
            w4 = group_sum<G>(w4);
            w5 = group_sum<G>(w5);
        }
    }
    if (t != 0) return;
    if constexpr (VIR) {
        const int64_t np = a.npad;
        a.vir[i] = w0;
        a.vir[np + i] = w1;
        a.vir[2 * np + i] = w2;
        a.vir[3 * np + i] = w3;
        a.vir[4 * np + i] = w4;
        a.vir[5 * np + i] = w5;
    }
    finish_particle<R>(a, i, pi, ti, fx, fy, fz, tx, ty, tz);
}

template <typename R, int LG, bool VIR>
__global__ void __launch_bounds__(256) k_hertz(ForceArgs a) {
    constexpr int G = 1 << LG;
    // XCD-aware block order: each XCD's L2 holds the positions of one
    // contiguous (spatially sorted) range of particles and its neighbors
    const int64_t gt = int64_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const int64_t i = gt >> LG;
    const int t = int(gt & (G - 1));
    if (i >= a.n) return;
    if (gated(a)) return;
    hertz_particle<R, LG, VIR>(a, i, t);
}

// The same pass over the particles in tile order (order[k], k < n): each
// XCD works through a band of bin columns, so the partners it gathers stay
// in its L2; each particle's arithmetic is that of k_hertz.
template <typename R, int LG, bool VIR>
__global__ void __launch_bounds__(256) k_hertz_tiles(ForceArgs a, const uint32_t* __restrict__ order) {
    constexpr int G = 1 << LG;
    const int64_t gt = int64_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const int64_t k = gt >> LG;
    const int t = int(gt & (G - 1));
    if (k >= a.n) return;
    if (gated(a)) return;
    hertz_particle<R, LG, VIR>(a, int64_t(order[k]), t);
}

// The same pass with the pairs in contact compacted per wavefront (one lane
// per particle, no virial). In hertz_particle a wavefront runs the contact
// arithmetic whenever any lane's pair at that row position is in contact,
// with every other lane idle (its pair is apart, or its row has ended).
// Here the lanes walk their rows together testing the geometry only, and
// queue each pair in contact (geometry, entry, row position, lane) in LDS;
// each time 64 are queued, every lane runs the contact arithmetic of one of
// them, and then each particle adds the results of its own queued pairs in
// row order. A pair's arithmetic (contact_pair) and the order of each
// particle's sums are those of hertz_particle, so forces, torques and
// springs are the same bit for bit.
constexpr int kWaveQ = 64;

struct WaveQueue {
    // per queued pair: dx dy dz distance delta r1 r2; the contact
    // arithmetic overwrites slots 0..63 with its 9 results (the last two in
    // res), while slots 64..127 hold the pairs queued after the 64th
    double g[7][2 * kWaveQ];
    double res[2][kWaveQ];
    uint32_t e[2 * kWaveQ];
    uint32_t lk[2 * kWaveQ];  // lane | row position << 8 | spring held << 16
};

// orders the wavefront's LDS accesses on either side (lanes read what other
// lanes of the same wavefront wrote)
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// contact arithmetic of the queued pairs 0..nq-1 (nq <= 64), then each lane
// adds those it queued (bits of own) to its sums
template <typename R>
__device__ inline void wave_contacts(const ForceArgs& a, WaveQueue& q, int nq, uint32_t my_i,
                                     uint64_t own, R& fx, R& fy, R& fz, R& tx, R& ty, R& tz) {
    const int lane = int(threadIdx.x & 63);
    wave_lds_sync();
    const uint32_t lk = q.lk[lane];
    const uint32_t io = uint32_t(__shfl(int(my_i), int(lk & 63u), 64));
    if (lane < nq) {
        const uint32_t e = q.e[lane];
        PairGeom g;
        g.dx = q.g[0][lane];
        g.dy = q.g[1][lane];
        g.dz = q.g[2][lane];
        g.distance = q.g[3][lane];
        g.delta = q.g[4][lane];
        g.r1 = q.g[5][lane];
        g.r2 = q.g[6][lane];
        const SysP* S = a.sysp ? &a.sysp[a.psys[io]] : nullptr;
        PairOut<R> o;
        contact_pair<R>(a, S, io, a.type[io], e, g, row_addr(io, int((lk >> 8) & 255u), a.npad, 0),
                        ((lk >> 16) & 1u) != 0u, o);
        q.g[0][lane] = double(o.fnx);
        q.g[1][lane] = double(o.fny);
        q.g[2][lane] = double(o.fnz);
        q.g[3][lane] = double(o.ftx);
        q.g[4][lane] = double(o.fty);
        q.g[5][lane] = double(o.ftz);
        q.g[6][lane] = double(o.tqx);
        q.res[0][lane] = double(o.tqy);
        q.res[1][lane] = double(o.tqz);
    }
    wave_lds_sync();
    while (own != 0ull) {
        const int p = __builtin_ctzll(own);
        own &= own - 1ull;
        PairOut<R> o;
        o.fnx = R(q.g[0][p]);
        o.fny = R(q.g[1][p]);
        o.fnz = R(q.g[2][p]);
        o.ftx = R(q.g[3][p]);
        o.fty = R(q.g[4][p]);
        o.ftz = R(q.g[5][p]);
        o.tqx = R(q.g[6][p]);
        o.tqy = R(q.res[0][p]);
        o.tqz = R(q.res[1][p]);
        add_pair<R>(entry_rev(q.e[p]), o, fx, fy, fz, tx, ty, tz);
    }
    wave_lds_sync();
}

// order: tile order (particle order[k] for k < n), or null for XCD-aware
// block order over the particles
template <typename R>
__global__ void __launch_bounds__(256) k_hertz_wave(ForceArgs a, const uint32_t* __restrict__ order) {
    __shared__ WaveQueue queues[4];
    WaveQueue& q = queues[threadIdx.x >> 6];
    const int lane = int(threadIdx.x & 63);
    if (gated(a)) return;
    const int64_t k = int64_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const bool live = k < a.n;
    const int64_t i = live ? (order ? int64_t(order[k]) : k) : 0;
    const double4 pi = a.pos[i];
    R fx = R(0), fy = R(0), fz = R(0), tx = R(0), ty = R(0), tz = R(0);
    const uint32_t ti = a.type[i];
    double Lx, Ly, Lz;
    particle_box(a, i, Lx, Ly, Lz);
    if (a.use_pairs) {
        int cnt = live ? int(a.nnb[i]) : 0;
        if (cnt > a.maxnb) cnt = a.maxnb;
        uint64_t mask_in = 0ull;
        if (cnt > 0) mask_in = a.cmask[i];
        uint64_t mask_out = 0ull;
        // the partner positions of the next row steps are loaded ahead: two
        // steps in FP32 (160 VGPRs, still 3 wavefronts per SIMD), one in
        // FP64, whose registers allow only one at 3 wavefronts
        constexpr bool kTwoAhead = std::is_same<R, float>::value;
        uint32_t e_n = 0 < cnt ? a.nbr[row_addr(i, 0, a.npad, 0)] : 0u;
        uint32_t e_nn = 1 < cnt ? a.nbr[row_addr(i, 1, a.npad, 0)] : 0u;
        uint32_t e_n3 = kTwoAhead && 2 < cnt ? a.nbr[row_addr(i, 2, a.npad, 0)] : 0u;
        double4 pj_n = a.pos[entry_idx(e_n)];
        double4 pj_nn = kTwoAhead ? a.pos[entry_idx(e_nn)] : pj_n;
        // queued pairs: nq (wavefront-uniform); this lane's among them at
        // the bits of own_lo (slots 0..63) and own_hi (slots 64..127)
        int nq = 0;
        uint64_t own_lo = 0ull, own_hi = 0ull;
        for (int t = 0; __ballot(t < cnt) != 0ull; t++) {
            bool in = false;
            uint32_t e = 0u;
            PairGeom g;
            if (t < cnt) {
                e = e_n;
                const double4 pj = pj_n;
                e_n = e_nn;
                if constexpr (kTwoAhead) {
                    pj_n = pj_nn;
                    e_nn = e_n3;
                    pj_nn = a.pos[entry_idx(e_nn)];
                    e_n3 = t + 3 < cnt ? a.nbr[row_addr(i, t + 3, a.npad, 0)] : 0u;
                } else {
                    pj_n = a.pos[entry_idx(e_n)];
                    e_nn = t + 2 < cnt ? a.nbr[row_addr(i, t + 2, a.npad, 0)] : 0u;
                }
                in = pair_geom(a, pi, pj, e, Lx, Ly, Lz, g);
            }
            const uint64_t b = __ballot(in);
            if (b == 0ull) continue;
            if (in) {
                const int p = nq + int(lanes_before(b));
                q.g[0][p] = g.dx;
                q.g[1][p] = g.dy;
                q.g[2][p] = g.dz;
                q.g[3][p] = g.distance;
                q.g[4][p] = g.delta;
                q.g[5][p] = g.r1;
                q.g[6][p] = g.r2;
                q.e[p] = e;
                q.lk[p] = uint32_t(lane) | (uint32_t(t) << 8) | (uint32_t((mask_in >> t) & 1ull) << 16);
                if (p < kWaveQ)
                    own_lo |= 1ull << p;
                else
                    own_hi |= 1ull << (p - kWaveQ);
                mask_out |= 1ull << t;
            }
            nq += __popcll(b);
            if (nq >= kWaveQ) {
                wave_contacts<R>(a, q, kWaveQ, uint32_t(i), own_lo, fx, fy, fz, tx, ty, tz);
                // the pairs queued past the 64th move to the front
                const int rest = nq - kWaveQ;
                if (lane < rest) {
#pragma unroll
                    for (int c = 0; c < 7; c++) q.g[c][lane] = q.g[c][kWaveQ + lane];
                    q.e[lane] = q.e[kWaveQ + lane];
                    q.lk[lane] = q.lk[kWaveQ + lane];
                }
                wave_lds_sync();
                own_lo = own_hi;
                own_hi = 0ull;
                nq = rest;
            }
        }
        if (nq > 0) wave_contacts<R>(a, q, nq, uint32_t(i), own_lo, fx, fy, fz, tx, ty, tz);
        if (live && mask_out != mask_in) a.cmask[i] = mask_out;
    }
    if (live) finish_particle<R>(a, i, pi, ti, fx, fy, fz, tx, ty, tz);
}

// The same pass split by row length (one lane per particle): the particles
// whose row is not empty, listed at the last rebuild, run the pair kernel
// above (grid-stride over the list); every other particle has no pair
// state to read or write and only meets the walls and gravity, in a kernel
// light enough to keep many wavefronts in flight. Each particle's
// arithmetic is that of k_hertz, so the forces are the same bit for bit.
template <typename R>
__global__ void __launch_bounds__(256) k_hertz_listed(ForceArgs a) {
    if (gated(a)) return;
    const uint32_t nl = *a.listed_n;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += stride)
        hertz_particle<R, 0, false>(a, int64_t(a.listed[k]), 0);
}

template <typename R>
__global__ void __launch_bounds__(256) k_hertz_unlisted(ForceArgs a) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= a.n) return;
    if (gated(a)) return;
    // every load is issued before the row-length test: one memory latency
    // per wavefront (a listed particle's loads are wasted)
    const uint32_t nn = a.nnb[i];
    const double4 pi = a.pos[i];
    const uint32_t ti = a.type[i];
    const double mi = a.mass[i];
    // (an empty statement that takes the loaded values keeps the compiler
    // from sinking their loads below the test)
    __asm__ volatile("" ::"v"(pi.x), "v"(pi.y), "v"(pi.z), "v"(pi.w), "v"(ti), "v"(mi));
    if (nn != 0u) return;  // a listed particle
    R fx = R(0), fy = R(0), fz = R(0), tx = R(0), ty = R(0), tz = R(0);
    if (a.use_walls)
        wall_forces<R>(a, pi, &a.vel[i], a.omega ? &a.omega[i] : nullptr, mi, ti, fx, fy, fz, tx,
                       ty, tz);
    if (a.use_gravity) {
        const double m = mi;
        fx += R(m * a.g[0]);
        fy += R(m * a.g[1]);
        fz += R(m * a.g[2]);
    }
    store_force<R>(a, i, fx, fy, fz, tx, ty, tz);
}

// The contact model family on the same rows: Hertz-Mindlin with JKR / DMT
// adhesion or SJKR cohesion (contact.rs:103-615) or, with the Hooke model,
// hooke_contact_force (contact.rs:622-990); constant or SDS rolling resistance
// and twisting friction (contact.rs:427-557) with their history slots (spring
// 0-2, rolling 3-5, twist 6, canonical sign). Every pair quantity is FP64;
// only the sums are kept in R. Pairs are evaluated in the owner frame as in
// k_hertz, so both rows of a pair carry the same history.
template <typename R, int LG>
__global__ void __launch_bounds__(256) k_contact_ext(ForceArgs a) {
    constexpr int G = 1 << LG;
    // XCD-aware block order: each XCD's L2 holds the positions of one
    // contiguous (spatially sorted) range of particles and its neighbors
    const int64_t gt = int64_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const int64_t i = gt >> LG;
    const int t = int(gt & (G - 1));
    if (i >= a.n) return;
    if (gated(a)) return;
    const double PI = 3.141592653589793;
    const double4 pi = a.pos[i];
    const double4 vi = a.vel[i];
    R fx = R(0), fy = R(0), fz = R(0), tx = R(0), ty = R(0), tz = R(0);
    double vw[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // contact virial (owned pairs)
    const uint32_t ti = a.type[i];
    const double Lx = a.L[0], Ly = a.L[1], Lz = a.L[2];
    if (a.use_pairs) {
        const double4 wi = a.omega[i];
        const uint32_t tagi = a.tag[i];
        int cnt = int(a.nnb[i]);
        if (cnt > a.maxnb) cnt = a.maxnb;
        const uint64_t mask_in = a.cmask[i];
        uint64_t mask_out = 0ull;
        const int nm = a.nm, nn = nm * nm;
        const double dt = a.dt;
        const int64_t cs = a.hist_cs;
        for (int k = t; k < cnt; k += G) {
            const int64_t slot = row_addr(i, k, a.npad, LG);
            const uint32_t e = a.nbr[slot];
            const uint32_t j = entry_idx(e);
            const bool rev = entry_rev(e);
            const double4 pj = a.pos[j];
            double shx, shy, shz;
            code_shifts(entry_code(e), Lx, Ly, Lz, shx, shy, shz);
            double dx, dy, dz, r1, r2;
            if (rev) {
                dx = (pi.x + shx) - pj.x;
                dy = (pi.y + shy) - pj.y;
                dz = (pi.z + shz) - pj.z;
                r1 = pj.w;
                r2 = pi.w;
            } else {
                dx = (pj.x + shx) - pi.x;
                dy = (pj.y + shy) - pi.y;
                dz = (pj.z + shz) - pi.z;
                r1 = pi.w;
                r2 = pj.w;
            }
            const double dist_sq = dx * dx + dy * dy + dz * dz;
            const double sum_r = r1 + r2;
            const uint32_t tj = a.type[j];
            const int mO = int(rev ? tj : ti), mP = int(rev ? ti : tj);
            const int mij = mO * nm + mP;
            const double r_eff = (r1 * r2) / sum_r;
            const double e_eff = a.mat[2 * nn + mij];
            const double surface = a.hooke ? 0.0 : a.mext[kXSurface * nn + mij];
            double d_po = 0.0;  // JKR pull-off distance (contact.rs:163-171)
            if (surface > 0.0 && !a.dmt)
                d_po = cbrt(PI * PI * surface * surface * r_eff / (4.0 * e_eff * e_eff));
            const double reach = sum_r + d_po;
            if (dist_sq >= reach * reach) continue;
            const double distance = sqrt(dist_sq);
            if (distance == 0.0) continue;
            const double r_min = r1 < r2 ? r1 : r2;
            double delta = sum_r - distance;
            if (0.5 * r_min < delta) delta = 0.5 * r_min;
            if (a.hooke && delta <= 0.0) continue;
            if (delta > 0.0 && !rev && distance / sum_r < kLargeOverlap) {
                int c = atomicAdd(&a.st->overlap_count, 1) + 1;
                if (c > kMaxOverlapWarnings) {
                    atomicCAS(&a.st->error, 0, -3);
                    if (a.step >= 0) atomicMin(&a.st->abort_step, a.step);
                }
            }
            if (delta <= 0.0 && surface <= 0.0) continue;
            const double4 vj = a.vel[j];
            const double4 wj = a.omega[j];
            const uint32_t tagj = a.tag[j];
            const double4& vO = rev ? vj : vi;
            const double4& vP = rev ? vi : vj;
            const double4& wO = rev ? wj : wi;
            const double4& wP = rev ? wi : wj;
            const uint32_t tagO = rev ? tagj : tagi;
            const uint32_t tagP = rev ? tagi : tagj;
            const double inv_dist = 1.0 / distance;
            const double nx = dx * inv_dist, ny = dy * inv_dist, nz = dz * inv_dist;
            const double m_r = 1.0 / (vO.w + vP.w);
            const double beta = a.mat[mij];
            const double mu = a.mat[nn + mij];
            const double cohesion = a.mext[kXCohesion * nn + mij];
            const bool adhesion_only = surface > 0.0 && !a.dmt && delta <= 0.0;
            double s_n = 0.0, k_n = 0.0, k_t = 0.0, gamma_n = 0.0;
            if (a.hooke) {
                k_n = a.mext[kXKn * nn + mij];
                k_t = a.mext[kXKt * nn + mij];
                gamma_n = 2.0 * beta * sqrt(k_n * m_r);
            } else if (delta > 0.0) {
                const double sdr = sqrt(delta * r_eff);
                s_n = 2.0 * e_eff * sdr;
                k_n = 4.0 / 3.0 * e_eff * sdr;
                k_t = 8.0 * a.mat[3 * nn + mij] * sdr;
            }
            const double r1nx = r1 * nx, r1ny = r1 * ny, r1nz = r1 * nz;
            const double vcix = vO.x + (wO.y * r1nz - wO.z * r1ny);
            const double vciy = vO.y + (wO.z * r1nx - wO.x * r1nz);
            const double vciz = vO.z + (wO.x * r1ny - wO.y * r1nx);
            const double r2nx = r2 * nx, r2ny = r2 * ny, r2nz = r2 * nz;
            const double vcjx = vP.x + (-wP.y * r2nz + wP.z * r2ny);
            const double vcjy = vP.y + (-wP.z * r2nx + wP.x * r2nz);
            const double vcjz = vP.z + (-wP.x * r2ny + wP.y * r2nx);
            const double vrx = vcjx - vcix, vry = vcjy - vciy, vrz = vcjz - vciz;
            const double v_n = vrx * nx + vry * ny + vrz * nz;
            double f_n;
            if (a.hooke) {
                f_n = k_n * delta - gamma_n * v_n;
                if (cohesion > 0.0)
                    f_n -= cohesion * PI * delta * r_eff;
                else if (!(f_n > 0.0))
                    f_n = 0.0;
            } else {
                const double f_diss = 2.0 * beta * kSqrt53 * sqrt(s_n * m_r) * v_n;
                if (surface > 0.0 && a.dmt) {
                    f_n = k_n * delta - f_diss - 2.0 * PI * surface * r_eff;
                } else if (surface > 0.0) {
                    const double f_adh = 1.5 * PI * surface * r_eff;
                    f_n = adhesion_only ? -f_adh : k_n * delta - f_diss - f_adh;
                } else if (cohesion > 0.0) {
                    f_n = k_n * delta - f_diss - cohesion * PI * delta * r_eff;
                } else {
                    f_n = k_n * delta - f_diss;
                    if (!(f_n > 0.0)) f_n = 0.0;
                }
            }
            // owner: -F_n, partner: +F_n
            const double sn = rev ? 1.0 : -1.0;
            fx += R(sn * (f_n * nx));
            fy += R(sn * (f_n * ny));
            fz += R(sn * (f_n * nz));
            if (adhesion_only) {  // no tangential part, no history
                if (a.dem_virial && !rev)
                    virial_add(vw, dx, dy, dz, -(f_n * nx), -(f_n * ny), -(f_n * nz));
                continue;
            }
            const double vtx = vrx - v_n * nx, vty = vry - v_n * ny, vtz = vrz - v_n * nz;
            const double sign = tagO < tagP ? 1.0 : -1.0;
            double h[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if ((mask_in >> k) & 1ull) {
#pragma unroll
                for (int q = 0; q < 7; q++) h[q] = sign * a.hist[q * cs + slot];
            }
            double sx = h[0], sy = h[1], sz = h[2];
            const double s_dot_n = sx * nx + sy * ny + sz * nz;
            sx -= s_dot_n * nx;
            sy -= s_dot_n * ny;
            sz -= s_dot_n * nz;
            sx += vtx * dt;
            sy += vty * dt;
            sz += vtz * dt;
            const double s_mag = sqrt(sx * sx + sy * sy + sz * sz);
            const double f_t_spring = k_t * s_mag;
            const double f_t_max = mu * fabs(f_n);
            if (f_t_spring > f_t_max && f_t_spring > kTangEps) {
                const double sc = f_t_max / f_t_spring;
                sx *= sc;
                sy *= sc;
                sz *= sc;
            }
            const double gamma_t = 2.0 * kSqrt53 * beta * sqrt(k_t * m_r);
            double ftx = k_t * sx - gamma_t * vtx;
            double fty = k_t * sy - gamma_t * vty;
            double ftz = k_t * sz - gamma_t * vtz;
            const double f_t_mag = sqrt(ftx * ftx + fty * fty + ftz * ftz);
            if (f_t_mag > f_t_max && f_t_mag > kTangEps) {
                const double sc = f_t_max / f_t_mag;
                ftx *= sc;
                fty *= sc;
                ftz *= sc;
            }
            if (a.dem_virial && !rev)
                virial_add(vw, dx, dy, dz, -(f_n * nx) + ftx, -(f_n * ny) + fty,
                           -(f_n * nz) + ftz);
            // owner: +F_t, torque (r1 n) x F_t; partner: -F_t, torque (r2 n) x F_t
            const double rnx = rev ? r2nx : r1nx, rny = rev ? r2ny : r1ny, rnz = rev ? r2nz : r1nz;
            fx += R(-sn * ftx);
            fy += R(-sn * fty);
            fz += R(-sn * ftz);
            double ttx = rny * ftz - rnz * fty;
            double tty = rnz * ftx - rnx * ftz;
            double ttz = rnx * fty - rny * ftx;
            // rolling resistance and twisting friction (contact.rs:427-557):
            // +tau on the owner, -tau on the partner
            const double orx = wO.x - wP.x, ory = wO.y - wP.y, orz = wO.z - wP.z;
            const double or_n = orx * nx + ory * ny + orz * nz;
            const double rlx = orx - or_n * nx, rly = ory - or_n * ny, rlz = orz - or_n * nz;
            double rdx = h[3], rdy = h[4], rdz = h[5], tw = h[6];
            const double mu_r = a.mext[kXRolling * nn + mij];
            const double mu_tw = a.mext[kXTwisting * nn + mij];
            const double st = rev ? -1.0 : 1.0;
            if (mu_r > 0.0) {
                const double roll_mag = sqrt(rlx * rlx + rly * rly + rlz * rlz);
                const double tau_max = mu_r * fabs(f_n) * r_eff;
                if (a.roll_sds) {
                    const double k_roll = a.mext[kXRollK * nn + mij];
                    const double g_roll = a.mext[kXRollG * nn + mij];
                    const double rd_n = rdx * nx + rdy * ny + rdz * nz;
                    rdx -= rd_n * nx;
                    rdy -= rd_n * ny;
                    rdz -= rd_n * nz;
                    rdx += rlx * dt;
                    rdy += rly * dt;
                    rdz += rlz * dt;
                    double trx = -k_roll * rdx - g_roll * rlx;
                    double try_ = -k_roll * rdy - g_roll * rly;
                    double trz = -k_roll * rdz - g_roll * rlz;
                    const double tr_mag = sqrt(trx * trx + try_ * try_ + trz * trz);
                    if (tr_mag > tau_max && tr_mag > kTangEps) {
                        const double sc = tau_max / tr_mag;
                        trx *= sc;
                        try_ *= sc;
                        trz *= sc;
                        if (k_roll > kTangEps) {
                            rdx = (trx + g_roll * rlx) / (-k_roll);
                            rdy = (try_ + g_roll * rly) / (-k_roll);
                            rdz = (trz + g_roll * rlz) / (-k_roll);
                        }
                    }
                    ttx += st * trx;
                    tty += st * try_;
                    ttz += st * trz;
                } else if (roll_mag > 1e-30) {
                    const double inv_roll = tau_max / roll_mag;
                    ttx += st * (-inv_roll * rlx);
                    tty += st * (-inv_roll * rly);
                    ttz += st * (-inv_roll * rlz);
                }
            }
            if (mu_tw > 0.0) {
                const double tau_max = mu_tw * fabs(f_n) * r_eff;
                double tau = 0.0;
                if (a.twist_sds) {
                    const double k_tw = a.mext[kXTwistK * nn + mij];
                    const double g_tw = a.mext[kXTwistG * nn + mij];
                    tw += or_n * dt;
                    tau = -k_tw * tw - g_tw * or_n;
                    if (fabs(tau) > tau_max) {
                        tau = (signbit(tau) ? -1.0 : 1.0) * tau_max;
                        if (k_tw > kTangEps) tw = (tau + g_tw * or_n) / (-k_tw);
                    }
                } else if (fabs(or_n) > 1e-30) {
                    tau = (or_n > 0.0 ? -1.0 : 1.0) * tau_max;
                }
                ttx += st * (tau * nx);
                tty += st * (tau * ny);
                ttz += st * (tau * nz);
            }
            tx += R(ttx);
            ty += R(tty);
            tz += R(ttz);
            a.hist[slot] = sign * sx;
            a.hist[cs + slot] = sign * sy;
            a.hist[2 * cs + slot] = sign * sz;
            a.hist[3 * cs + slot] = sign * rdx;
            a.hist[4 * cs + slot] = sign * rdy;
            a.hist[5 * cs + slot] = sign * rdz;
            a.hist[6 * cs + slot] = sign * tw;
            mask_out |= 1ull << k;
        }
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        tx = group_sum<G>(tx);
        ty = group_sum<G>(ty);
        tz = group_sum<G>(tz);
        mask_out = group_or<G>(mask_out);
        if (t == 0) a.cmask[i] = mask_out;
        if (a.dem_virial) {
#pragma unroll
            for (int q = 0; q < 6; q++) vw[q] = group_sum<G>(vw[q]);
            if (t == 0)
                for (int q = 0; q < 6; q++) a.vir[q * a.npad + i] = vw[q];
        }
    }
    if (t != 0) return;
    if (a.use_walls)
        wall_forces<R>(a, pi, &a.vel[i], a.omega ? &a.omega[i] : nullptr, a.mass[i], ti, fx, fy, fz, tx,
                       ty, tz);
    if (a.use_gravity) {
        const double m = a.mass[i];
        fx += R(m * a.g[0]);
        fy += R(m * a.g[1]);
        fz += R(m * a.g[2]);
    }
    store_force<R>(a, i, fx, fy, fz, tx, ty, tz);
}

// standalone wall / gravity system: adds to the current forces
template <typename R>
__global__ void k_walls_gravity(ForceArgs a) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= a.n) return;
    if (gated(a)) return;
    R* f = nullptr;
    if constexpr (sizeof(R) == 8)
        f = a.f64;
    else
        f = a.f32;
    R fx = f[i], fy = f[a.npad + i], fz = f[2 * a.npad + i];
    R* t = nullptr;
    if constexpr (sizeof(R) == 8)
        t = a.t64;
    else
        t = a.t32;
    R tx = t ? t[i] : R(0), ty = t ? t[a.npad + i] : R(0), tz = t ? t[2 * a.npad + i] : R(0);
    if (a.use_walls) {
        if (!a.regions_only)
            wall_forces<R>(a, a.pos[i], &a.vel[i], a.omega ? &a.omega[i] : nullptr, a.mass[i],
                           a.type[i], fx, fy, fz, tx, ty, tz);
        if (a.has_regions) region_wall_forces<R>(a, a.pos[i], &a.vel[i], a.mass[i], a.type[i], fx, fy, fz);
    }
    if (t) {
        t[i] = tx;
        t[a.npad + i] = ty;
        t[2 * a.npad + i] = tz;
    }
    if (a.use_gravity) {
        const double m = a.mass[i];
        fx += R(m * a.g[0]);
        fy += R(m * a.g[1]);
        fz += R(m * a.g[2]);
    }
    f[i] = fx;
    f[a.npad + i] = fy;
    f[2 * a.npad + i] = fz;
}

__global__ void k_reset_overlap(Status* st) { st->overlap_count = 0; }

ForceArgs make_args(Ctx& c, int step) {
    ForceArgs a{};
    a.n = c.n;
    a.npad = c.npad;
    a.maxnb = c.maxnb;
    a.pos = c.pos.p;
    a.vel = c.vel.p;
    a.omega = c.omega.p;
    a.tag = c.tag.p;
    a.type = c.type.p;
    a.mass = c.mass.p;
    a.nbr = c.nbr.p;
    a.nnb = c.nnb.p;
    a.gcnt = c.list_injected ? nullptr : c.gcnt.p;
    for (int d = 0; d < 3; d++) a.L[d] = c.box.L[d];
    a.lj1 = c.lj1;
    a.lj2 = c.lj2;
    a.cut2 = c.lj_cut2;
    a.e12 = c.lj_e12;
    a.e6 = c.lj_e6;
    a.mat = c.mat_dev.p;
    a.nm = c.mat.n;
    a.mext = c.ext_on ? c.mext_dev.p : nullptr;
    a.hooke = c.cm_hooke;
    a.dmt = c.cm_dmt;
    a.roll_sds = c.cm_roll_sds;
    a.twist_sds = c.cm_twist_sds;
    a.dt = c.dt;
    a.hist = c.hist.p;
    a.cmask = c.cmask.p;
    a.hist_cs = int64_t(c.maxnb) * c.npad;
    a.walls = c.walls_dev.p;
    a.wdyn = c.wdyn.p;
    for (int k = 0; k < c.walls.count; k++) a.has_regions |= c.walls.w[k].kind == 3 ? 1 : 0;
    a.g[0] = c.g[0];
    a.g[1] = c.g[1];
    a.g[2] = c.g[2];
    a.f64 = c.f64.p;
    a.f32 = c.f32.p;
    a.t64 = c.t64.p;
    a.t32 = c.t32.p;
    a.pe = c.pe.p;
    a.vir = c.vir.p;
    a.st = c.status.p;
    a.step = step;
    a.sysp = c.sysp_or_null();
    a.psys = c.psys.p;
    return a;
}

// A pair-force launch of nb blocks of 256 threads. Timing events, when
// given, are attached to the dispatch itself (hipExtLaunchKernelGGL), so
// they time the kernel alone, as a kernel trace does.
template <typename K>
void launch_force_kernel(Ctx& c, K kern, int nb, const ForceArgs& a, hipEvent_t e0,
                         hipEvent_t e1) {
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, c.stream, e0, e1, 0, a);
    else
        kern<<<nb, 256, 0, c.stream>>>(a);
}
template <typename K>
void launch_force_kernel(Ctx& c, K kern, int nb, const ForceArgs& a, hipEvent_t e0,
                         hipEvent_t e1, const uint32_t* order) {
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, c.stream, e0, e1, 0, a, order);
    else
        kern<<<nb, 256, 0, c.stream>>>(a, order);
}

// fixed-point positions of every particle (x * qs per axis, modulo 2^32)
__global__ void k_fill_xq(int64_t n, const double4* __restrict__ pos, int4* __restrict__ xq,
                          double qx, double qy, double qz) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    xq[i] = make_int4(quantize_axis(p.x, qx), quantize_axis(p.y, qy), quantize_axis(p.z, qz), 0);
}

// the fixed-point positions of the mixed LJ pass (filled when stale) and
// their units, into the pass's arguments
void lj_fixed_setup(Ctx& c, ForceArgs& a) {
    const double two32 = 4294967296.0;
    if (!c.xq_valid) {
        c.xq.resize(size_t(c.npad));
        k_fill_xq<<<blocks_for(c.n, 256), 256, 0, c.stream>>>(
            c.n, static_cast<const double4*>(c.pos.p), c.xq.p, two32 / c.box.L[0],
            two32 / c.box.L[1], two32 / c.box.L[2]);
        MDD_HIP(hipGetLastError());
        c.xq_valid = true;
    }
    a.xq = c.xq.p;
    for (int d = 0; d < 3; d++) {
        a.qd[d] = c.box.L[d] / two32;
        a.qf[d] = float(a.qd[d]);
    }
}

template <typename R, bool E>
void launch_lj(Ctx& c, const ForceArgs& a0, hipEvent_t e0, hipEvent_t e1) {
    const int64_t threads = c.n << c.list_lg;
    const int nb = blocks_for(threads, 256);
    ForceArgs a = a0;
    if (sizeof(R) == 4 && !E && c.xq_usable()) lj_fixed_setup(c, a);
    switch (c.list_lg) {
        case 0: launch_force_kernel(c, k_lj<R, E, 0>, nb, a, e0, e1); break;
        case 1: launch_force_kernel(c, k_lj<R, E, 1>, nb, a, e0, e1); break;
        case 3: launch_force_kernel(c, k_lj<R, E, 3>, nb, a, e0, e1); break;
        default: launch_force_kernel(c, k_lj<R, E, 2>, nb, a, e0, e1); break;
    }
}

// The mixed LJ row pass with the thermostatted final half step fused in
// (k_lj_post), for a scheduled step of a single Nose-Hoover system whose
// post kernel would use one block per 256 particles: returns false (and
// launches nothing) otherwise.
bool launch_lj_post(Ctx& c, const ForceArgs& a0, hipEvent_t e0, hipEvent_t e1, int step) {
    const int64_t nb = blocks_for(c.n, 256);
    if (!c.lj_post_fused || step < 0 || !c.nh_on || c.vv_on || c.rot_on || c.nsys != 0 ||
        c.list_lg > 2 || nb > kMaxKePartials || !c.xq_usable() || c.list_tile)
        return false;
    ForceArgs a = a0;
    lj_fixed_setup(c, a);
    c.partial.resize(kMaxKePartials);
    PostArgs p;
    p.vel = c.vel.p;
    p.mass = c.mass.p;
    p.partial = c.partial.p;
    p.q = step & 1;
    p.dt = c.dt;
    const dim3 block(256u << c.list_lg);
    auto go = [&](auto kern) {
        if (e0)
            hipExtLaunchKernelGGL(kern, dim3(int(nb)), block, 0, c.stream, e0, e1, 0, a, p);
        else
            kern<<<dim3(int(nb)), block, 0, c.stream>>>(a, p);
    };
    switch (c.list_lg) {
        case 0: go(k_lj_post<0>); break;
        case 1: go(k_lj_post<1>); break;
        default: go(k_lj_post<2>); break;
    }
    MDD_HIP(hipGetLastError());
    return true;
}

template <typename R, bool VIR>
void launch_hertz(Ctx& c, const ForceArgs& a, hipEvent_t e0, hipEvent_t e1) {
    const int64_t threads = c.n << c.list_lg;
    const int nb = blocks_for(threads, 256);
    if (c.ext_on) {
        switch (c.list_lg) {
            case 0: launch_force_kernel(c, k_contact_ext<R, 0>, nb, a, e0, e1); break;
            case 1: launch_force_kernel(c, k_contact_ext<R, 1>, nb, a, e0, e1); break;
            case 3: launch_force_kernel(c, k_contact_ext<R, 3>, nb, a, e0, e1); break;
            default: launch_force_kernel(c, k_contact_ext<R, 2>, nb, a, e0, e1); break;
        }
        return;
    }
    const bool tiles = c.tile_order && c.tile_ok && c.nsys == 0;
    if (c.dem_wave && c.list_lg == 0 && !VIR) {
        launch_force_kernel(c, k_hertz_wave<R>, nb, a, e0, e1,
                            tiles ? static_cast<const uint32_t*>(c.tile_list.p) : nullptr);
        return;
    }
    if (tiles) {
        const uint32_t* ord = c.tile_list.p;
        switch (c.list_lg) {
            case 0: launch_force_kernel(c, k_hertz_tiles<R, 0, VIR>, nb, a, e0, e1, ord); break;
            case 1: launch_force_kernel(c, k_hertz_tiles<R, 1, VIR>, nb, a, e0, e1, ord); break;
            case 3: launch_force_kernel(c, k_hertz_tiles<R, 3, VIR>, nb, a, e0, e1, ord); break;
            default: launch_force_kernel(c, k_hertz_tiles<R, 2, VIR>, nb, a, e0, e1, ord); break;
        }
        return;
    }
    switch (c.list_lg) {
        case 0: launch_force_kernel(c, k_hertz<R, 0, VIR>, nb, a, e0, e1); break;
        case 1: launch_force_kernel(c, k_hertz<R, 1, VIR>, nb, a, e0, e1); break;
        case 3: launch_force_kernel(c, k_hertz<R, 3, VIR>, nb, a, e0, e1); break;
        default: launch_force_kernel(c, k_hertz<R, 2, VIR>, nb, a, e0, e1); break;
    }
}

}  // namespace

void upload_materials(Ctx& c) {
    const int nm = c.mat.n;
    std::vector<double> h(size_t(4) * nm * nm);
    for (int k = 0; k < nm * nm; k++) {
        h[k] = c.mat.beta[k];
        h[nm * nm + k] = c.mat.mu[k];
        h[2 * nm * nm + k] = c.mat.eeff[k];
        h[3 * nm * nm + k] = c.mat.geff[k];
    }
    c.mat_dev.resize(h.size());
    h2d(c, c.mat_dev.p, h.data(), sizeof(double) * h.size());
}

// one pair-force pass (Force phase). step < 0: ungated standalone call.
void launch_pair_forces(Ctx& c, int step, bool energy) {
    ForceArgs a = make_args(c, step);
    const bool mixed = c.precision == 1;
    if (step < 0) {
        k_reset_overlap<<<1, 1, 0, c.stream>>>(c.status.p);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool prof = c.prof_every > 0 && step >= 0 && (c.prof_ctr++ % c.prof_every) == 0 &&
                      c.prof_pending < int64_t(c.ev_pool.size() / 2);
    // the pass is one kernel except for cell tiles (staging) and region
    // walls (a second pass): the events go on that kernel, else around all
    // DEM at one lane per particle: the pass is split by row length
    // (single systems: in batched ensembles the grid-stride pair kernel lost
    // to the one-kernel pass, 1.64e10 vs 1.72e10 particle-steps/s)
    // DEM contact virial on this pass (VirialStress active: every step
    // number that is a multiple of virial_every; standalone calls always)
    const bool dem_vir = c.model == 1 && c.nsys == 0 && c.virial_every > 0 &&
                         (step < 0 || step % c.virial_every == 0);
    if (dem_vir) {
        c.vir.resize(size_t(6) * c.npad);
        a.vir = c.vir.p;
        a.dem_virial = 1;
        c.vir_step = step < 0 ? c.step : step;
        c.vir_valid = true;
    }
    // the split pass pays off when most particles have no neighbors (a
    // dilute cloud): the light kernel then covers most of them. In a dense
    // packing every particle is listed and the one-kernel pass (XCD-aware
    // block order, one thread per particle) is the faster one.
    const bool sparse_rows = c.dem_split == 1 ||
                             (c.dem_split < 0 && 2 * int64_t(c.status_host->listed) < c.n);
    const bool split = c.model == 1 && c.listed_ok && c.list_valid && !c.list_injected &&
                       c.list_lg == 0 && !c.ext_on && c.nsys == 0 && !dem_vir && sparse_rows;
    const bool one_kernel = !(c.model == 0 && c.list_tile) && !split &&
                            !(c.model == 1 && step >= 0 && c.walls.count > 0 && a.has_regions);
    if (prof) {
        e0 = c.ev_pool[2 * c.prof_pending];
        e1 = c.ev_pool[2 * c.prof_pending + 1];
        if (!one_kernel) MDD_HIP(hipEventRecord(e0, c.stream));
    }
    hipEvent_t k0 = one_kernel ? e0 : nullptr, k1 = one_kernel ? e1 : nullptr;
    c.last_force_fused = false;
    c.fused_rows = false;
    if (c.model == 0 && c.list_tile) {
        if (energy) {
            c.pe.resize(c.npad);
            c.vir.resize(size_t(6) * c.npad);
        }
        const bool fuse = !energy && step >= 0;
        launch_lj_tile(c, step, energy, fuse);
        c.last_force_fused = fuse;
        c.fused_parts = c.grid.ncells;
    } else if (c.model == 0) {
        if (energy) {
            c.pe.resize(c.npad);
            c.vir.resize(size_t(6) * c.npad);
            a.pe = c.pe.p;
            a.vir = c.vir.p;
            if (mixed)
                launch_lj<float, true>(c, a, k0, k1);
            else
                launch_lj<double, true>(c, a, k0, k1);
        } else if (mixed && launch_lj_post(c, a, k0, k1, step)) {
            c.last_force_fused = true;
            c.fused_rows = true;
        } else {
            if (mixed)
                launch_lj<float, false>(c, a, k0, k1);
            else
                launch_lj<double, false>(c, a, k0, k1);
        }
    } else {
        a.use_pairs = 1;
        a.use_walls = step >= 0 && c.walls.count > 0;
        a.use_gravity = step >= 0 && c.gravity_on;
        // region walls come last among the walls and before gravity: their
        // own pass follows the pair pass and adds gravity after them
        const bool regions = a.use_walls && a.has_regions;
        if (regions) a.use_gravity = 0;
        if (split) {
            a.listed = c.listed.p;
            a.listed_n = c.listed_n.p;
            const int nb = blocks_for(c.n, 256);
            const int nbl = nb < 2048 ? nb : 2048;
            if (mixed) {
                k_hertz_listed<float><<<nbl, 256, 0, c.stream>>>(a);
                k_hertz_unlisted<float><<<nb, 256, 0, c.stream>>>(a);
            } else {
                k_hertz_listed<double><<<nbl, 256, 0, c.stream>>>(a);
                k_hertz_unlisted<double><<<nb, 256, 0, c.stream>>>(a);
            }
        } else if (dem_vir) {
            if (mixed)
                launch_hertz<float, true>(c, a, k0, k1);
            else
                launch_hertz<double, true>(c, a, k0, k1);
        } else if (mixed) {
            launch_hertz<float, false>(c, a, k0, k1);
        } else {
            launch_hertz<double, false>(c, a, k0, k1);
        }
        if (regions) {
            ForceArgs b = a;
            b.regions_only = 1;
            b.use_gravity = c.gravity_on ? 1 : 0;
            const int nb = blocks_for(c.n, 256);
            if (mixed)
                k_walls_gravity<float><<<nb, 256, 0, c.stream>>>(b);
            else
                k_walls_gravity<double><<<nb, 256, 0, c.stream>>>(b);
        }
    }
    MDD_HIP(hipGetLastError());
    c.force_launches++;
    if (prof) {
        if (!one_kernel) MDD_HIP(hipEventRecord(e1, c.stream));
        c.prof_step[size_t(c.prof_pending)] = step;
        c.prof_pending++;
    }
}

namespace {
__global__ void k_wall_systems(const WallsP* w, WallsDyn* d, double dt, int move, int zero) {
    if (threadIdx.x == 0) walls_advance(*w, *d, dt, move != 0, zero != 0);
}
}  // namespace

void launch_wall_systems(Ctx& c, bool move, bool zero) {
    if (c.walls.count == 0 || !c.wdyn.p) return;
    k_wall_systems<<<1, 64, 0, c.stream>>>(c.walls_dev.p, c.wdyn.p, c.dt, move ? 1 : 0,
                                           zero ? 1 : 0);
    MDD_HIP(hipGetLastError());
}

void launch_walls_gravity(Ctx& c, bool walls, bool gravity) {
    ForceArgs a = make_args(c, -1);
    a.use_walls = walls && c.walls.count > 0;
    a.use_gravity = gravity && c.gravity_on;
    const int nb = blocks_for(c.n, 256);
    if (c.precision == 1)
        k_walls_gravity<float><<<nb, 256, 0, c.stream>>>(a);
    else
        k_walls_gravity<double><<<nb, 256, 0, c.stream>>>(a);
    MDD_HIP(hipGetLastError());
}

}  // namespace mdd
