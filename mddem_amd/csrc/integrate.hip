// Time integration kernels of the gfx950 particle path: Velocity Verlet
// (crates/mddem_verlet/src/lib.rs:97-154), quaternion rotation
// (crates/dem_granular/src/rotational.rs:59-100) and the Nose-Hoover fused
// half steps (crates/md_thermostat/src/lib.rs:254-353).
//
// The pre kernel of a step ends with the reference's rebuild decision
// (decide_rebuild, crates/mddem_core/src/neighbor.rs:639-674): a particle
// that moved more than the skin budget since the last build, or (when
// enabled) left the periodic box, records the step in Status::abort_step.
// Every later kernel of the step sequence then returns at once, the host
// rebuilds the neighbor list and resumes with the force pass of that step.
//
// Nose-Hoover without a grid-wide barrier: the post kernel of step s leaves
// per-block kinetic-energy partials (a fixed number of blocks, grid-stride),
// and every block of the pre kernel of step s+1 sums all of them in the same
// fixed order, so each block derives the bitwise-identical thermostat update
// by itself. The reference computes that same KE twice (end of step s and
// start of step s+1, md_thermostat/src/lib.rs:262-265 and 345-352) and
// advances p_xi by two equal terms; both are applied here. The thermostat
// momentum lives in two slots indexed by step parity: the pre kernel of step
// s reads slot (s-1)&1 and its block 0 writes slot s&1, so no block reads a
// value another block of the same launch is writing.
#include <mutex>
#include <string>

#include "ctx.hpp"

namespace mdd {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPartials = kMaxKePartials;  // blocks of a thermostatted post kernel

struct IntArgs {
    int64_t n, npad;
    double dt;
    double4* pos;
    double4* vel;
    double4* omega;
    double4* quat;
    const double* mass;
    const double4* build_pos;
    const double* f64;
    const float* f32;
    const double* t64;
    const float* t32;
    int rot;         // apply rotation half-steps
    int vv;          // apply Verlet translational half-steps
    int check;       // evaluate the rebuild decision
    double thr2;     // ((skin_fraction - 1) * min_skin)^2
    BoxP box;
    int box_exit;
    Status* st;
    int step;
    double* partial;
    int nparts;      // KE partials left by the last post kernel
    int q_in, q_out; // thermostat slots read / written
    double ndof, temp, q_mass;
    const SysP* sysp;       // batched ensemble: per-system box and threshold, or null
    const uint32_t* psys;
    const WallsP* walls;    // walls (device copy) and their state, or null: scheduled
    WallsDyn* wdyn;         // steps move them (PreInitialIntegration) and clear the sums
    int4* xq;               // fixed-point copy of the positions to keep current, or null
    double qs[3];           // its scale 2^32 / L per axis
    const int4* bxq;        // with xq on a periodic box: fixed-point build positions, or null
    double qu[3];           // L / 2^32 per axis
    double thr_band;        // |d2 - thr2| up to this: the double4 build position decides
};

__device__ inline int4 quantize_pos(const IntArgs& a, const double4& p) {
    return make_int4(quantize_axis(p.x, a.qs[0]), quantize_axis(p.y, a.qs[1]),
                     quantize_axis(p.z, a.qs[2]), 0);
}

__device__ inline void store_xq(const IntArgs& a, int64_t i, const double4& p) {
    a.xq[i] = quantize_pos(a, p);
}

template <typename F>
__device__ inline void load_f(const IntArgs& a, int64_t i, double& fx, double& fy, double& fz) {
    if constexpr (sizeof(F) == 8) {
        fx = a.f64[i];
        fy = a.f64[a.npad + i];
        fz = a.f64[2 * a.npad + i];
    } else {
        fx = double(a.f32[i]);
        fy = double(a.f32[a.npad + i]);
        fz = double(a.f32[2 * a.npad + i]);
    }
}

template <typename F>
__device__ inline void load_t(const IntArgs& a, int64_t i, double& tx, double& ty, double& tz) {
    if constexpr (sizeof(F) == 8) {
        tx = a.t64[i];
        ty = a.t64[a.npad + i];
        tz = a.t64[2 * a.npad + i];
    } else {
        tx = double(a.t32[i]);
        ty = double(a.t32[a.npad + i]);
        tz = double(a.t32[2 * a.npad + i]);
    }
}

// Sum over the 256 threads of a block in a fixed order; every thread gets
// the same bits (xor butterfly inside each wavefront: both partners add the
// same two operands; then the four wave sums in fixed order).
__device__ inline double block_sum(double v, double* lds4) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    return (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
}

// the KE partials (sum m v^2) thread t adds: t, t + 256, ... (at most
// kMaxPartials in all), loaded together so that one memory latency covers them
constexpr int kPartsPerThread = kMaxPartials / kThreads;
static_assert(kPartsPerThread * kThreads == kMaxPartials, "partials per thread");

__device__ inline void load_partials(const IntArgs& a, double (&v)[kPartsPerThread]) {
#pragma unroll
    for (int k = 0; k < kPartsPerThread; k++) {
        const int b = threadIdx.x + k * kThreads;
        v[k] = b < a.nparts ? a.partial[b] : 0.0;
    }
}

// total of the loaded partials, identical in every block (each thread adds
// its own in index order, then block_sum)
__device__ inline double sum_loaded(const IntArgs& a, const double (&v)[kPartsPerThread],
                                    double* lds4) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kPartsPerThread; k++)
        if (threadIdx.x + k * kThreads < a.nparts) s += v[k];
    return block_sum(s, lds4);
}

__device__ inline double sum_partials(const IntArgs& a, double* lds4) {
    double v[kPartsPerThread];
    load_partials(a, v);
    return sum_loaded(a, v, lds4);
}

// p_xi half update (md_thermostat/src/lib.rs:266 and 352)
__device__ inline double nh_term(const IntArgs& a, double ke) {
    return (a.dt / 2.0) * (2.0 * ke - a.ndof * a.temp);
}

// displacement_exceeded (neighbor.rs:564-601) + periodic box exit (651-666)
// for position p and the position b at the last build
__device__ inline void rebuild_check(const IntArgs& a, int64_t i, const double4& p,
                                     const double4& b) {
    const BoxP* box = &a.box;
    double thr2 = a.thr2;
    if (a.sysp) {
        const SysP& S = a.sysp[a.psys[i]];
        box = &S.box;
        thr2 = S.thr2;
    }
    double d[3] = {p.x - b.x, p.y - b.y, p.z - b.z};
    const double x[3] = {p.x, p.y, p.z};
    bool out = false;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (box->periodic[k]) {
            const double L = box->L[k], h = L * 0.5;
            if (d[k] > h)
                d[k] -= L;
            else if (d[k] < -h)
                d[k] += L;
            if (a.box_exit && (x[k] < box->lo[k] || x[k] >= box->hi[k])) out = true;
        }
    }
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > thr2 || out)
        atomicMin(&a.st->abort_step, a.step);
}

// rebuild_check from the fixed-point position and build position (every
// axis periodic). The int32 differences are the minimum image, and each
// coordinate is within one quantum of its double value, so d2 is within
// thr_band of rebuild_check's d2 near the threshold; outside the band both
// decide alike, inside it rebuild_check decides from the double4.
__device__ inline void rebuild_check_fixed(const IntArgs& a, int64_t i, const double4& p,
                                           const int4& bq) {
    const int4 q = quantize_pos(a, p);
    const double dx = double(int(uint32_t(q.x) - uint32_t(bq.x))) * a.qu[0];
    const double dy = double(int(uint32_t(q.y) - uint32_t(bq.y))) * a.qu[1];
    const double dz = double(int(uint32_t(q.z) - uint32_t(bq.z))) * a.qu[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (fabs(d2 - a.thr2) <= a.thr_band) {
        rebuild_check(a, i, p, a.build_pos[i]);
        return;
    }
    bool out = false;
    if (a.box_exit) {
        const double x[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (x[k] < a.box.lo[k] || x[k] >= a.box.hi[k]) out = true;
    }
    if (d2 > a.thr2 || out) atomicMin(&a.st->abort_step, a.step);
}

// orientation update of initial_rotation with the kicked omega w
// (rotational.rs:70-83)
__device__ inline void rotate_quat(const IntArgs& a, int64_t i, const double4& w) {
    const double dt = a.dt;
    const double omega_mag = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    const double angle = omega_mag * dt;
    if (angle > 1e-14) {
        const double inv = 1.0 / omega_mag;
        const double ax = w.x * inv, ay = w.y * inv, az = w.z * inv;
        const double half = angle * 0.5;
        const double s = sin(half);
        const double a0 = cos(half), a1 = ax * s, a2 = ay * s, a3 = az * s;
        const double4 b = a.quat[i];
        double4 q;
        q.x = a0 * b.x - a1 * b.y - a2 * b.z - a3 * b.w;
        q.y = a0 * b.y + a1 * b.x + a2 * b.w - a3 * b.z;
        q.z = a0 * b.z - a1 * b.w + a2 * b.x + a3 * b.y;
        q.w = a0 * b.w + a1 * b.z - a2 * b.y + a3 * b.x;
        a.quat[i] = q;
    }
}

// initial_rotation body (rotational.rs:59-84)
__device__ inline void rotate_initial(const IntArgs& a, int64_t i, double tx, double ty,
                                      double tz) {
    double4 w = a.omega[i];
    const double ii = w.w;
    if (ii == 0.0) return;
    const double dt = a.dt;
    w.x += 0.5 * dt * tx * ii;
    w.y += 0.5 * dt * ty * ii;
    w.z += 0.5 * dt * tz * ii;
    // a zero torque leaves the spin as it is: no store
    if (tx != 0.0 || ty != 0.0 || tz != 0.0) a.omega[i] = w;
    rotate_quat(a, i, w);
}

// final_rotation body (rotational.rs:87-100)
__device__ inline void rotate_final(const IntArgs& a, int64_t i, double tx, double ty,
                                    double tz) {
    double4 w = a.omega[i];
    const double ii = w.w;
    if (ii == 0.0) return;
    const double dt = a.dt;
    w.x += 0.5 * dt * tx * ii;
    w.y += 0.5 * dt * ty * ii;
    w.z += 0.5 * dt * tz * ii;
    if (tx != 0.0 || ty != 0.0 || tz != 0.0) a.omega[i] = w;
}

__device__ inline void kick_drift(const IntArgs& a, double4& v, double4& p, double fx, double fy,
                                  double fz) {
    const double dt = a.dt;
    const double h = 0.5 * dt * v.w;
    v.x += h * fx;
    v.y += h * fy;
    v.z += h * fz;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
}

// PreInitialIntegration + InitialIntegration of one step: thermostat
// (nh_pre_initial: p_xi update, rescale, kick, drift), Verlet kick + drift,
// rotation, then the rebuild decision.
template <typename F, bool NH>
__global__ void __launch_bounds__(kThreads) k_pre(IntArgs a) {
    __shared__ double lds4[4];
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    // the particle's loads and the KE partials are issued before the first
    // store and the reduction's barriers, so their latencies overlap
    const bool live = i < a.n;
    double4 v = make_double4(0.0, 0.0, 0.0, 0.0), p = v, b = v;
    int4 bq = make_int4(0, 0, 0, 0);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (live) {
        p = a.pos[i];
        if (NH || a.vv) {
            v = a.vel[i];
            load_f<F>(a, i, fx, fy, fz);
        }
        if (a.check) {
            if (a.bxq)
                bq = a.bxq[i];
            else
                b = a.build_pos[i];
        }
    }
    double part[kPartsPerThread];
    if (NH) load_partials(a, part);
    // uniform over the launch: only earlier kernels lower abort_step below step
    if (a.step >= 0 && a.st->abort_step < a.step) return;
    if (i == 0) {
        a.st->overlap_count = 0;
        if (a.wdyn && a.step >= 0) walls_advance(*a.walls, *a.wdyn, a.dt);
    }
    double scale = 1.0;
    if (NH) {
        const double ke = 0.5 * sum_loaded(a, part, lds4);
        const double term = nh_term(a, ke);
        double pxi = a.st->nh_p[a.q_in];
        if (a.st->nh_kind[a.q_in] == 0) pxi += term;  // post half of the previous step
        pxi += term;                                    // this step's pre half
        scale = exp(-a.dt / 2.0 * pxi / a.q_mass);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.st->nh_p[a.q_out] = pxi;
            a.st->nh_scale[a.q_out] = scale;
            a.st->nh_kind[a.q_out] = 0;
        }
    }
    if (!live) return;
    if (NH || a.vv) {
        if (NH) {
            v.x *= scale;
            v.y *= scale;
            v.z *= scale;
            kick_drift(a, v, p, fx, fy, fz);
        }
        if (a.vv) kick_drift(a, v, p, fx, fy, fz);
        a.vel[i] = v;
        a.pos[i] = p;
        if (a.xq) store_xq(a, i, p);
    }
    if (a.rot) {
        double tx, ty, tz;
        load_t<F>(a, i, tx, ty, tz);
        rotate_initial(a, i, tx, ty, tz);
    }
    if (a.check) {
        if (a.bxq)
            rebuild_check_fixed(a, i, p, bq);
        else
            rebuild_check(a, i, p, b);
    }
}

// FinalIntegration + PostFinalIntegration of one step: Verlet kick,
// rotation, thermostat kick + rescale with this step's factor and the
// per-block kinetic-energy partials (grid-stride over a fixed block count).
template <typename F, bool NH>
__global__ void __launch_bounds__(kThreads) k_post(IntArgs a) {
    __shared__ double lds4[4];
    if (a.step >= 0 && a.st->abort_step <= a.step) return;
    const double scale = NH ? a.st->nh_scale[a.q_in] : 1.0;
    double ke = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < a.n; i += stride) {
        double4 v = a.vel[i];
        if (NH || a.vv) {
            double fx, fy, fz;
            load_f<F>(a, i, fx, fy, fz);
            const double h = 0.5 * a.dt * v.w;
            if (a.vv) {
                v.x += h * fx;
                v.y += h * fy;
                v.z += h * fz;
            }
            if (NH) {
                v.x += h * fx;
                v.y += h * fy;
                v.z += h * fz;
                v.x *= scale;
                v.y *= scale;
                v.z *= scale;
            }
            a.vel[i] = v;
        }
        if (a.rot) {
            double tx, ty, tz;
            load_t<F>(a, i, tx, ty, tz);
            rotate_final(a, i, tx, ty, tz);
        }
        if (NH) ke += a.mass[i] * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    if (!NH) return;
    const double bs = block_sum(ke, lds4);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = bs;
}

// The final half step of step s - 1 and the first half step of step s in
// one pass, for runs without a thermostat (nothing global happens between
// them): Verlet kick (final_integration), kick + drift (initial_integration),
// final_rotation then initial_rotation on the same omega, then the rebuild
// decision. Each particle's arithmetic is the same as the two kernels'.
template <typename F>
__global__ void __launch_bounds__(kThreads) k_post_pre(IntArgs a) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (a.step >= 0 && a.st->abort_step < a.step) return;
    if (i == 0) {
        a.st->overlap_count = 0;
        if (a.wdyn && a.step >= 0) walls_advance(*a.walls, *a.wdyn, a.dt);
    }
    if (i >= a.n) return;
    // every load before the first store (the pointers may alias as far as
    // the compiler knows, so it would not move later loads above a store)
    double4 p = a.pos[i];
    double4 v = make_double4(0.0, 0.0, 0.0, 0.0), w = v, b = v;
    double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
    if (a.vv) {
        v = a.vel[i];
        load_f<F>(a, i, fx, fy, fz);
    }
    if (a.rot) {
        load_t<F>(a, i, tx, ty, tz);
        w = a.omega[i];
    }
    if (a.check) b = a.build_pos[i];
    if (a.vv) {
        const double h = 0.5 * a.dt * v.w;
        v.x += h * fx;
        v.y += h * fy;
        v.z += h * fz;
        kick_drift(a, v, p, fx, fy, fz);
        a.vel[i] = v;
        a.pos[i] = p;
        if (a.xq) store_xq(a, i, p);
    }
    if (a.rot) {
        const double ii = w.w;
        if (ii != 0.0) {
            const double dt = a.dt;
            w.x += 0.5 * dt * tx * ii;  // final_rotation of step s - 1
            w.y += 0.5 * dt * ty * ii;
            w.z += 0.5 * dt * tz * ii;
            w.x += 0.5 * dt * tx * ii;  // initial_rotation of step s
            w.y += 0.5 * dt * ty * ii;
            w.z += 0.5 * dt * tz * ii;
            // a zero torque leaves the spin as it is: no store
            if (tx != 0.0 || ty != 0.0 || tz != 0.0) a.omega[i] = w;
            rotate_quat(a, i, w);
        }
    }
    if (a.check) rebuild_check(a, i, p, b);
}

// KE partials of the current velocities, laid out like the post kernel's
__global__ void __launch_bounds__(kThreads) k_ke_partials(IntArgs a) {
    __shared__ double lds4[4];
    double ke = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < a.n; i += stride) {
        const double4 v = a.vel[i];
        ke += a.mass[i] * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    const double bs = block_sum(ke, lds4);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = bs;
}

// slot q <- Status::p_xi with the matching velocity scale. kind 1: a
// post-half value (the reference's p_xi between steps; a following pre
// kernel adds one KE term); kind 0: a pre-half value (a following finish
// adds the post-half term)
__global__ void k_nh_seed(IntArgs a, int q, int kind) {
    const double p = a.st->p_xi;
    a.st->nh_p[q] = p;
    a.st->nh_kind[q] = kind;
    a.st->nh_scale[q] = exp(-a.dt / 2.0 * p / a.q_mass);
}

// Status::p_xi <- slot q (after a standalone pre half-step)
__global__ void k_nh_pxi_from_slot(Status* st, int q) { st->p_xi = st->nh_p[q]; }

// finish the post half-update of slot q from the KE partials: Status::ke and
// Status::p_xi hold the reference's values between steps afterwards, and the
// slot becomes a post-half value (a following pre kernel adds one term)
__global__ void __launch_bounds__(kThreads) k_nh_finish(IntArgs a, int q) {
    __shared__ double lds4[4];
    const double ke = 0.5 * sum_partials(a, lds4);
    if (threadIdx.x != 0) return;
    double p = a.st->nh_p[q];
    if (a.st->nh_kind[q] == 0) p += nh_term(a, ke);
    a.st->nh_p[q] = p;
    a.st->nh_kind[q] = 1;
    a.st->ke = ke;
    a.st->p_xi = p;
}

// Thermo reduction, all on the device: per-block partials of
//   0 sum m v^2, 1-3 sum m v, 4 sum m, 5 sum m |v - vbar|^2,
//   6 half-list pairs (sum of the row counts nhalf), 7 contact bits (cmask)
// (kThermoCols per block), then one block sums each column in a fixed order
// (thread t takes blocks t, t + 256, ... in turn, then a fixed tree), so the
// totals do not depend on timing. Pass 1 leaves column 5 at 0 and the mean
// velocity vbar = sum m v / sum m in the scratch; pass 2 fills column 5 from
// it. Scratch layout in doubles: totals [0, 8), vbar [8, 11), partials from
// kThermoPartOff, kThermoCols per block.
constexpr int kThermoCols = 8;
constexpr int kThermoPartOff = 16;

__global__ void __launch_bounds__(kThreads) k_thermo(IntArgs a, const uint32_t* nhalf,
                                                     const uint64_t* cmask, double* scratch,
                                                     int pass) {
    __shared__ double lds4[4];
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    double* out = scratch + kThermoPartOff + size_t(blockIdx.x) * kThermoCols;
    if (pass == 1) {
        double q[kThermoCols] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (i < a.n) {
            const double4 v = a.vel[i];
            const double m = a.mass[i];
            q[0] = m * (v.x * v.x + v.y * v.y + v.z * v.z);
            q[1] = m * v.x;
            q[2] = m * v.y;
            q[3] = m * v.z;
            q[4] = m;
            if (nhalf) q[6] = double(nhalf[i]);
            if (cmask) q[7] = double(__popcll(cmask[i]));
        }
        for (int k = 0; k < kThermoCols; k++) {
            const double s = block_sum(q[k], lds4);
            if (threadIdx.x == 0) out[k] = s;
        }
    } else {
        double q = 0.0;
        if (i < a.n) {
            const double4 v = a.vel[i];
            const double m = a.mass[i];
            const double ux = v.x - scratch[8], uy = v.y - scratch[9], uz = v.z - scratch[10];
            q = m * (ux * ux + uy * uy + uz * uz);
        }
        const double s = block_sum(q, lds4);
        if (threadIdx.x == 0) out[5] = s;
    }
}

// one block: column totals over nb block partials (pass 1: every column and
// vbar; pass 2: column 5)
__global__ void __launch_bounds__(kThreads) k_thermo_total(double* scratch, int nb, int pass) {
    __shared__ double lds4[4];
    const double* parts = scratch + kThermoPartOff;
    for (int k = 0; k < kThermoCols; k++) {
        if (pass == 2 && k != 5) continue;
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += kThreads) s += parts[size_t(b) * kThermoCols + k];
        s = block_sum(s, lds4);
        if (threadIdx.x == 0) scratch[k] = s;
    }
    if (pass == 1 && threadIdx.x == 0) {
        const double m = scratch[4];
        scratch[8] = m > 0.0 ? scratch[1] / m : 0.0;
        scratch[9] = m > 0.0 ? scratch[2] / m : 0.0;
        scratch[10] = m > 0.0 ? scratch[3] / m : 0.0;
    }
}

// The static scratch of mdd_compute_thermo_noalloc: totals, vbar and the
// partials of at most kThermoStaticBlocks blocks (2 MiB), one copy per GPU
// for the life of the process. Kernels receive its address as looked up on
// each device (thermo_static_scratch); the symbol itself is never passed.
constexpr int64_t kThermoStaticBlocks = 32768;
static __device__ __align__(256) double scratch[kThermoPartOff +
                                                kThermoCols * kThermoStaticBlocks];

int post_blocks(const Ctx& c, bool nh) {
    const int nb = blocks_for(c.n, kThreads);
    return nh ? std::min(nb, kMaxPartials) : nb;
}

IntArgs make_args(Ctx& c, int step, bool check) {
    IntArgs a{};
    if (c.walls.count > 0 && c.wdyn.p) {
        a.walls = c.walls_dev.p;
        a.wdyn = c.wdyn.p;
    }
    a.n = c.n;
    a.npad = c.npad;
    a.dt = c.dt;
    a.pos = c.pos.p;
    a.vel = c.vel.p;
    a.omega = c.omega.p;
    a.quat = c.quat.p;
    a.mass = c.mass.p;
    a.build_pos = c.build_pos.p;
    a.f64 = c.f64.p;
    a.f32 = c.f32.p;
    a.t64 = c.t64.p;
    a.t32 = c.t32.p;
    a.rot = 0;
    a.vv = 0;
    a.check = check && c.list_valid && !c.list_injected ? 1 : 0;
    const double skin = (c.skin_fraction - 1.0) * c.min_skin;
    // every > 0 without check: the step count alone decides (the periodic
    // box exit still does, decide_rebuild, neighbor.rs:651-666)
    a.thr2 = (c.every > 0 && !c.check) ? INFINITY : skin * skin;
    a.box = c.box;
    a.box_exit = c.box_exit_rebuild;
    a.st = c.status.p;
    a.step = step;
    c.partial.resize(kMaxPartials);
    a.partial = c.partial.p;
    a.nparts = post_blocks(c, true);
    if (c.list_tile) {
        // tile lists: the fused force pass (or the post kernel) leaves
        // partials, summed into ke_total by launch_ke_reduce
        c.ke_total.resize(1);
        a.partial = c.ke_total.p;
        a.nparts = 1;
    }
    a.q_in = step >= 0 ? ((step - 1) & 1) : 0;
    a.q_out = step >= 0 ? (step & 1) : 1;
    a.ndof = c.nh_ndof;
    a.temp = c.nh_T;
    a.q_mass = c.nh_Q;
    a.sysp = c.sysp_or_null();
    a.psys = c.psys.p;
    if (c.xq_usable() && c.xq_valid) {
        a.xq = c.xq.p;
        for (int d = 0; d < 3; d++) a.qs[d] = 4294967296.0 / c.box.L[d];
        const bool periodic = c.box.periodic[0] && c.box.periodic[1] && c.box.periodic[2];
        if (c.bxq_valid && periodic && std::isfinite(a.thr2)) {
            // each coordinate differs from its double value by at most one
            // quantum u, so |d2 - d2_double| <= 2 sqrt(3) u |d| + 3 u^2;
            // the band holds that at |d| = thr with a margin of more than 2
            double u = 0.0;
            for (int d = 0; d < 3; d++) {
                a.qu[d] = c.box.L[d] / 4294967296.0;
                u = std::max(u, a.qu[d]);
            }
            a.bxq = c.bxq.p;
            a.thr_band = 8.0 * u * (std::sqrt(a.thr2) + u) + 1e-12 * a.thr2;
        }
    } else {
        c.xq_valid = false;  // these launches may move particles without it
    }
    return a;
}

template <bool NH>
void pre_impl(Ctx& c, IntArgs& a) {
    const int nb = blocks_for(c.n, kThreads);
    if (c.precision == 1)
        k_pre<float, NH><<<nb, kThreads, 0, c.stream>>>(a);
    else
        k_pre<double, NH><<<nb, kThreads, 0, c.stream>>>(a);
    MDD_HIP(hipGetLastError());
}

template <bool NH>
void post_impl(Ctx& c, IntArgs& a) {
    const int nb = post_blocks(c, NH);
    if (c.precision == 1)
        k_post<float, NH><<<nb, kThreads, 0, c.stream>>>(a);
    else
        k_post<double, NH><<<nb, kThreads, 0, c.stream>>>(a);
    MDD_HIP(hipGetLastError());
}

}  // namespace

// scheduled pre kernel of a step (all enabled integrators)
void launch_pre(Ctx& c, int step, bool check_rebuild) {
    IntArgs a = make_args(c, step, check_rebuild);
    a.rot = c.rot_on && c.model == 1;
    a.vv = c.vv_on;
    if (c.nh_on)
        pre_impl<true>(c, a);
    else
        pre_impl<false>(c, a);
}

// the deferred final half step of step - 1 with the first half of `step`
// (runs without a thermostat)
void launch_post_pre(Ctx& c, int step, bool check_rebuild) {
    IntArgs a = make_args(c, step, check_rebuild);
    a.rot = c.rot_on && c.model == 1;
    a.vv = c.vv_on;
    const int nb = blocks_for(c.n, kThreads);
    if (c.precision == 1)
        k_post_pre<float><<<nb, kThreads, 0, c.stream>>>(a);
    else
        k_post_pre<double><<<nb, kThreads, 0, c.stream>>>(a);
    MDD_HIP(hipGetLastError());
}

// scheduled post kernel of a step; its thermostat half uses slot step & 1
void launch_post(Ctx& c, int step) {
    if (c.last_force_fused) {
        // the force pass integrated already; cell tiles leave per-cell
        // partials, summed here; fused rows leave the post kernel's own
        if (c.nh_on && !c.fused_rows) launch_ke_reduce(c, step, c.fused_parts);
        return;
    }
    IntArgs a = make_args(c, step, false);
    a.partial = c.partial.p;
    a.rot = c.rot_on && c.model == 1;
    a.vv = c.vv_on;
    a.q_in = step & 1;
    if (c.nh_on)
        post_impl<true>(c, a);
    else
        post_impl<false>(c, a);
    if (c.nh_on && c.list_tile) launch_ke_reduce(c, step, post_blocks(c, true));
}

void launch_initial_integration(Ctx& c, int step, bool check_rebuild) {
    IntArgs a = make_args(c, step, check_rebuild);
    a.vv = 1;
    pre_impl<false>(c, a);
}

void launch_final_integration(Ctx& c) {
    IntArgs a = make_args(c, -1, false);
    a.vv = 1;
    post_impl<false>(c, a);
}

void launch_initial_rotation(Ctx& c) {
    IntArgs a = make_args(c, -1, false);
    a.rot = 1;
    pre_impl<false>(c, a);
}

void launch_final_rotation(Ctx& c) {
    IntArgs a = make_args(c, -1, false);
    a.rot = 1;
    post_impl<false>(c, a);
}

// KE partials of the current velocities (thermostat start)
void launch_ke_partials(Ctx& c) {
    IntArgs a = make_args(c, -1, false);
    a.partial = c.partial.p;
    k_ke_partials<<<post_blocks(c, true), kThreads, 0, c.stream>>>(a);
    MDD_HIP(hipGetLastError());
    if (c.list_tile) launch_ke_reduce(c, -1, post_blocks(c, true));
}

void launch_nh_seed(Ctx& c, int parity) {
    IntArgs a = make_args(c, -1, false);
    k_nh_seed<<<1, 1, 0, c.stream>>>(a, parity & 1, 1);
    MDD_HIP(hipGetLastError());
}

void launch_nh_finish(Ctx& c, int parity) {
    IntArgs a = make_args(c, -1, false);
    k_nh_finish<<<1, kThreads, 0, c.stream>>>(a, parity & 1);
    MDD_HIP(hipGetLastError());
}

// standalone nh_pre_initial: p_xi from Status, KE of the current velocities
void launch_nh_pre_standalone(Ctx& c) {
    launch_ke_partials(c);
    launch_nh_seed(c, 0);
    IntArgs a = make_args(c, -1, false);
    a.q_in = 0;
    a.q_out = 1;
    pre_impl<true>(c, a);
    k_nh_pxi_from_slot<<<1, 1, 0, c.stream>>>(c.status.p, 1);
    MDD_HIP(hipGetLastError());
}

// standalone nh_post_final: scale from Status::p_xi, then the KE update
void launch_nh_post_standalone(Ctx& c) {
    IntArgs a = make_args(c, -1, false);
    k_nh_seed<<<1, 1, 0, c.stream>>>(a, 0, 0);
    MDD_HIP(hipGetLastError());
    a.q_in = 0;
    a.partial = c.partial.p;
    post_impl<true>(c, a);
    if (c.list_tile) launch_ke_reduce(c, -1, post_blocks(c, true));
    launch_nh_finish(c, 0);
}

// out: [sum m v^2, sum m vx, sum m vy, sum m vz, sum m, sum m |v - vbar|^2]
namespace {

// both passes into `buf` (kThermoPartOff + kThermoCols * blocks doubles), the
// 8 totals copied to `out` (host); nothing is allocated
void thermo_into(Ctx& c, double* buf, double out[8]) {
    // only the particle count, velocities and masses (make_args would size
    // the step buffers)
    IntArgs a{};
    a.n = c.n;
    a.npad = c.npad;
    a.vel = c.vel.p;
    a.mass = c.mass.p;
    const int nb = blocks_for(c.n, kThreads);
    const uint32_t* nhalf = c.list_valid && c.nhalf.p ? c.nhalf.p : nullptr;
    const uint64_t* cmask = c.list_valid && c.model == 1 && c.cmask.p ? c.cmask.p : nullptr;
    for (int pass = 1; pass <= 2; pass++) {
        k_thermo<<<nb, kThreads, 0, c.stream>>>(a, nhalf, cmask, buf, pass);
        k_thermo_total<<<1, kThreads, 0, c.stream>>>(buf, nb, pass);
    }
    MDD_HIP(hipGetLastError());
    MDD_HIP(hipMemcpyAsync(out, buf, sizeof(double) * 8, hipMemcpyDeviceToHost, c.stream));
    MDD_HIP(hipStreamSynchronize(c.stream));
}

// address of `scratch` on each device, looked up once per device
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

void thermo_reduce(Ctx& c, double out[8]) {
    const int nb = blocks_for(c.n, kThreads);
    c.partial2.resize(size_t(kThermoPartOff) + size_t(nb) * kThermoCols);
    thermo_into(c, c.partial2.p, out);
}

// Look up the static scratch's device address on `device` (made current by
// the caller), once per device. Called when a context is created.
void init_thermo_static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw ApiError(-1, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    MDD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(p);
}

int64_t thermo_static_max_particles() { return kThermoStaticBlocks * kThreads; }

// thermo_reduce through the static scratch of the context's device. Calls
// are serialised within the process (the lock is held until the totals
// are on the host), so concurrent calls on one GPU do not overlap.
void thermo_reduce_static(Ctx& c, double out[8]) {
    if (c.n > thermo_static_max_particles())
        throw ApiError(-1, "mdd_compute_thermo_noalloc: " + std::to_string(c.n) +
                               " particles exceed the static scratch limit of " +
                               std::to_string(thermo_static_max_particles()));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    double* p = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch_ptr[c.device] : nullptr;
    if (!p) throw ApiError(-1, "static scratch not initialised on this device");
    thermo_into(c, p, out);
}

}  // namespace mdd
