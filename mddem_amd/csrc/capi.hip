// C ABI of the gfx950 particle path (include/mddem_hip.h) and the host step
// scheduler that runs the reference phase order (ParticleSimScheduleSet,
// crates/mddem_core/src/schedule.rs:22-52; schedule.dot) on one GPU.
//
// Step scheduling: a chunk of K steps is queued as [pre, force, post] x K
// with no host synchronisation. The pre kernel of each step takes the
// rebuild decision on the device; a step that needs a rebuild records itself
// in Status::abort_step and every later kernel of the chunk returns
// immediately. The host then reads the status once per chunk, rebuilds the
// neighbor list and resumes at the force pass of the aborted step, which is
// exactly where the reference rebuilds (decide_rebuild runs after the
// initial integration, neighbor.rs:639-674). Periodic spatial sorts
// (sort_atoms_by_bin every sort_every steps, neighbor.rs:697-701) are
// scheduled by the host.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

#include "ctx.hpp"
#include "mddem_hip.h"

using mdd::ApiError;
using mdd::Ctx;

struct mdd_ctx {
    Ctx c;
};

namespace mdd {
void neighbor_setup(Ctx& c);
}

namespace {

__global__ void k_status_init(mdd::Status* st) {
    st->abort_step = INT_MAX;
    st->error = 0;
    st->overflow = 0u;
    st->max_row = 0;
    st->max_half = 0;
    st->ghost_total = 0u;
    st->overlap_count = 0;
    st->nh_kind[0] = st->nh_kind[1] = 1;
    st->hist_records = 0u;
    st->listed = 0u;
    st->ke = 0.0;
    st->p_xi = 0.0;
    st->nh_p[0] = st->nh_p[1] = 0.0;
    st->nh_scale[0] = st->nh_scale[1] = 1.0;
}

__global__ void k_set_pxi(mdd::Status* st, double v) { st->p_xi = v; }

constexpr int kEventPool = 2048;
// a chunk after a rebuild runs to this multiple of the smoothed rebuild
// interval, plus one step
#ifndef MDD_CHUNK_OVERSHOOT
#define MDD_CHUNK_OVERSHOOT 0.85
#endif
constexpr double kChunkOvershoot = MDD_CHUNK_OVERSHOOT;

int fail(mdd_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->c.err = msg;
    return code;
}

template <class F>
int guarded(mdd_ctx* ctx, F&& body) {
    if (!ctx) return MDD_ERR_INVALID;
    try {
        MDD_HIP(hipSetDevice(ctx->c.device));
        body(ctx->c);
        ctx->c.err.clear();
        return MDD_OK;
    } catch (const ApiError& e) {
        return fail(ctx, e.code, e.what());
    } catch (const mdd::HipError& e) {
        return fail(ctx, MDD_ERR_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, MDD_ERR_CAPACITY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, MDD_ERR_INVALID, e.what());
    }
}

void commit_systems(Ctx& c);

void need_state(Ctx& c) {
    if (c.systems_dirty) commit_systems(c);
    if (c.n <= 0) throw ApiError(MDD_ERR_INVALID, "no particles uploaded");
}

void need_list(Ctx& c) {
    need_state(c);
    if (!c.list_valid) throw ApiError(MDD_ERR_INVALID, "neighbor list not built");
}

void sync_status(Ctx& c) {
    MDD_HIP(hipMemcpyAsync(c.status_host, c.status.p, sizeof(mdd::Status), hipMemcpyDeviceToHost,
                           c.stream));
    MDD_HIP(hipStreamSynchronize(c.stream));
}

// Called right after a status read. Timed force launches of steps at or
// past the chunk's abort step returned at once (the step is redone after
// the rebuild) and are not counted.
void harvest_profile(Ctx& c) {
    const int ab = c.status_host->abort_step;
    for (int64_t k = 0; k < c.prof_pending; k++) {
        if (c.prof_step[size_t(k)] >= ab) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.ev_pool[2 * k], c.ev_pool[2 * k + 1]) == hipSuccess) {
            c.prof_ms_sum += ms;
            c.prof_count++;
        }
    }
    c.prof_pending = 0;
    for (int64_t k = 0; k < c.rb_pending; k++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.rb_ev[3 * k + 1], c.rb_ev[3 * k + 2]) == hipSuccess) {
            c.rb_ms_sum += ms;
            c.rb_count++;
        }
        if (c.rb_has_gap[size_t(k)] &&
            hipEventElapsedTime(&ms, c.rb_ev[3 * k], c.rb_ev[3 * k + 1]) == hipSuccess) {
            c.rb_gap_ms_sum += ms;
            c.rb_gap_count++;
        }
    }
    c.rb_pending = 0;
}

void check_device_error(Ctx& c) {
    if (c.status_host->error) {
        const int e = c.status_host->error;
        if (e == MDD_ERR_OVERLAP)
            throw ApiError(MDD_ERR_OVERLAP,
                           "Over 500 excessive overlaps this step -- aborting. Check timestep or "
                           "initial configuration.");
        throw ApiError(e, "device reported an error");
    }
}

void setup_nose_hoover(Ctx& c) {
    // setup_nose_hoover (md_thermostat/src/lib.rs:195-243): ndof = 3N - 3,
    // Q = ndof T tau^2, p_xi = 0
    const double n = double(c.n);
    c.nh_ndof = 3.0 * n - 3.0;
    c.nh_Q = c.nh_ndof * c.nh_T * c.nh_tau * c.nh_tau;
    k_set_pxi<<<1, 1, 0, c.stream>>>(c.status.p, 0.0);
    MDD_HIP(hipGetLastError());
    c.nh_ready = true;
}

void run_steps(Ctx& c, int64_t nsteps) {
    need_state(c);
    if (c.model == MDD_MODEL_DEM && c.mat.n == 0 && c.nsys == 0)
        throw ApiError(MDD_ERR_INVALID, "materials not set");
    if (c.nsys > 0 && c.walls.count > 0)
        throw ApiError(MDD_ERR_UNSUPPORTED, "walls in a batched ensemble");
    if (c.nsys > 0 && c.ext_on)
        throw ApiError(MDD_ERR_UNSUPPORTED, "optional contact models in a batched ensemble");
    if (c.step + nsteps > INT_MAX - 2) throw ApiError(MDD_ERR_INVALID, "step count too large");
    if (!c.setup_done) mdd::neighbor_setup(c);
    if (c.nh_on && !c.nh_ready) setup_nose_hoover(c);
    if (nsteps == 0) return;
    if (c.nh_on) {
        // the first pre kernel reads the KE of the current velocities and
        // the current p_xi from the slot of the previous step's parity
        mdd::launch_ke_partials(c);
        mdd::launch_nh_seed(c, int(c.step & 1));
    }
    const int64_t end = c.step + nsteps;
    // Without a thermostat nothing global happens between a step's final
    // half step and the next step's first: the final half is deferred and
    // issued together with the next first half (one pass over the state).
    const bool defer = !c.nh_on;
    if (c.post_pending) {
        mdd::launch_post(c, int(c.step));
        c.post_pending = false;
    }
    auto first_half = [&](int s, bool check) {
        if (c.post_pending) {
            mdd::launch_post_pre(c, s, check);
            c.post_pending = false;
        } else {
            mdd::launch_pre(c, s, check);
        }
    };
    auto second_half = [&](int s) {
        if (defer && !c.last_force_fused)
            c.post_pending = true;
        else
            mdd::launch_post(c, s);
    };
    // resume at the force pass of step ab after the device stopped there:
    // a rebuild request from the drift, or a rebuild that did not fit
    auto resume = [&](int ab) {
        const bool retry = (c.status_host->overflow & 31u) != 0u;
        // the rebuild's first kernel clears the abort
        if (retry) {
            if (c.status_host->overflow & 4u) mdd::remove_outside(c);
            mdd::rebuild_grow(c, *c.status_host);
            if (++c.retries > 6) throw ApiError(MDD_ERR_CAPACITY, "neighbor rebuild did not fit");
        } else {
            c.retries = 0;
        }
        // a step that asks for a rebuild again right after its own rebuild
        // (without a capacity retry) would never advance: stop with an error
        if (!retry && ab == c.last_resume && ++c.repeats > 3)
            throw ApiError(MDD_ERR_STALLED, "rebuild requested again at the same step");
        if (ab != c.last_resume) c.repeats = 0;
        c.last_resume = ab;
        mdd::rebuild_async(c, ab, retry);
        mdd::launch_pair_forces(c, ab, false);
        second_half(ab);
        if (!retry) {
            // the next chunk ends near the expected next rebuild (0.85 of
            // the interval, plus one step): launches after a rebuild request
            // return at once but still cost about 2 us each, and an early
            // end costs one more host round trip
            const int64_t ran = int64_t(ab) - c.last_rebuild;
            if (ran > 0) {
                c.rb_interval = c.rb_interval > 0.0 ? 0.5 * c.rb_interval + 0.5 * double(ran)
                                                    : double(ran);
                c.chunk = int(std::min(1024.0, std::ceil(kChunkOvershoot * c.rb_interval) + 1.0));
            }
            c.last_rebuild = ab;
        }
        c.step = ab;
    };
    auto check_status = [&]() {
        sync_status(c);
        harvest_profile(c);
        check_device_error(c);
        return c.status_host->abort_step;
    };
    // needs_rebuild with every > 0 (neighbor.rs:605-627): steps_since_build
    // is 0 after the build of step b and grows by one at each later step's
    // check that keeps the list, so the count forces a rebuild at b + every + 1
    auto count_due = [&](int64_t s) {
        return c.every > 0 && s >= c.last_rebuild + int64_t(c.every) + 1;
    };
    for (;;) {
        while (c.step < end) {
            const int64_t s = c.step + 1;
            const bool forced = !c.list_valid || c.list_injected || c.list_lg != c.lg ||
                                c.list_tile != c.want_tile() || count_due(s);
            if (forced) {
                // FullRebuild step: integrate, rebuild, force, integrate
                first_half(int(s), false);
                mdd::rebuild_async(c, int(s), false);
                mdd::launch_pair_forces(c, int(s), false);
                second_half(int(s));
                c.step = s;
                c.last_rebuild = s;
                continue;
            }
            int64_t K = std::min<int64_t>(c.chunk, end - c.step);
            if (c.every > 0) K = std::min<int64_t>(K, c.last_rebuild + c.every - c.step);
            if (K < 1) K = 1;
            for (int64_t k = 1; k <= K; k++) {
                const int st = int(c.step + k);
                first_half(st, true);
                // sort_atoms_by_bin's periodic sort: a rebuild unless still sorted
                if (c.sort_every > 0 && st % c.sort_every == 0) mdd::launch_sort_check(c, st);
                mdd::launch_pair_forces(c, st, false);
                second_half(st);
            }
            if (c.prof_every > 0 && c.rb_mark) {
                MDD_HIP(hipEventRecord(c.rb_mark, c.stream));
                c.rb_chunk_mark = true;
            }
            const int ab = check_status();
            if (ab != INT_MAX) {
                resume(ab);
            } else {
                c.step += K;
                c.retries = 0;
                c.nghost = c.status_host->ghost_total;
                // no rebuild yet: the rest of the expected interval; past it,
                // short chunks that grow with the steps run since the rebuild;
                // with no interval known yet, doubling chunks
                const double since = double(c.step - c.last_rebuild);
                const double rest = kChunkOvershoot * c.rb_interval - since;
                if (c.rb_interval <= 0.0)
                    c.chunk = std::min(c.chunk * 2, 1024);
                else if (rest >= 1.0)
                    c.chunk = int(std::ceil(rest)) + 1;
                else
                    c.chunk = int(std::min(1024.0, std::max(2.0, std::ceil(0.25 * since))));
            }
        }
        const int ab = check_status();
        if (ab == INT_MAX) break;
        resume(ab);
    }
    c.nghost = c.status_host->ghost_total;
    if (c.post_pending) {
        mdd::launch_post(c, int(c.step));
        c.post_pending = false;
    }
    // post half of the last step's thermostat update -> Status::p_xi, ke
    if (c.nh_on) mdd::launch_nh_finish(c, int(c.step & 1));
    sync_status(c);
    harvest_profile(c);
    check_device_error(c);
}

void pack_upload(Ctx& c, int64_t n, const uint32_t* tag, const uint32_t* type, const double* pos,
                 const double* vel, const double* omega, const double* radius,
                 const double* cutr, const double* mass) {
    if (n <= 0 || n >= mdd::kMaxParticles)
        throw ApiError(MDD_ERR_INVALID, "particle count out of range");
    if (!tag || !pos || !vel || !cutr || !mass)
        throw ApiError(MDD_ERR_INVALID, "tag, pos, vel, cutoff_radius and mass are required");
    const bool dem = c.model == MDD_MODEL_DEM;
    if (dem && !radius) throw ApiError(MDD_ERR_INVALID, "DEM particles need a radius");
    int64_t cap = std::max<int64_t>(n, c.cap);
    c.cap = cap;
    c.n = n;
    c.xq_valid = false;
    c.npad = (cap + 63) / 64 * 64;
    const int64_t np = c.npad;
    c.pos.resize(np);
    c.vel.resize(np);
    c.tag.resize(np);
    c.type.resize(np);
    c.image.resize(np);
    c.mass.resize(np);
    c.cutr.resize(np);
    c.build_pos.resize(np);
    c.pos2.resize(np);
    c.vel2.resize(np);
    c.tag2.resize(np);
    c.type2.resize(np);
    c.image2.resize(np);
    c.mass2.resize(np);
    c.cutr2.resize(np);
    const bool mixed = c.precision == MDD_MIXED;
    if (mixed) {
        c.f32.resize(3 * np);
        c.f32b.resize(3 * np);
    } else {
        c.f64.resize(3 * np);
        c.f64b.resize(3 * np);
    }
    if (dem) {
        c.omega.resize(np);
        c.quat.resize(np);
        c.omega2.resize(np);
        c.quat2.resize(np);
        if (mixed) {
            c.t32.resize(3 * np);
            c.t32b.resize(3 * np);
        } else {
            c.t64.resize(3 * np);
            c.t64b.resize(3 * np);
        }
    }
    std::vector<double4> hp(n), hv(n), hw, hq;
    std::vector<uint32_t> ht(tag, tag + n), hty(n, 0u);
    if (type) std::copy(type, type + n, hty.begin());
    if (dem) {
        hw.resize(n);
        hq.resize(n);
    }
    for (int64_t i = 0; i < n; i++) {
        if (!(mass[i] > 0.0)) throw ApiError(MDD_ERR_INVALID, "masses must be positive");
        if (dem && hty[i] >= uint32_t(c.mat.n > 0 ? c.mat.n : mdd::kMaxMat))
            throw ApiError(MDD_ERR_INVALID, "particle material out of range");
        hp[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], dem ? radius[i] : 0.0);
        hv[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 1.0 / mass[i]);
        if (dem) {
            const double r = radius[i];
            // insert_single_particle (dem_atom/src/insert.rs:340-350)
            const double inv_i = 1.0 / (0.4 * mass[i] * r * r);
            hw[i] = make_double4(omega ? omega[3 * i] : 0.0, omega ? omega[3 * i + 1] : 0.0,
                                 omega ? omega[3 * i + 2] : 0.0, inv_i);
            hq[i] = make_double4(1.0, 0.0, 0.0, 0.0);
        }
    }
    h2d(c, c.pos.p, hp.data(), sizeof(double4) * n);
    h2d(c, c.vel.p, hv.data(), sizeof(double4) * n);
    h2d(c, c.tag.p, ht.data(), sizeof(uint32_t) * n);
    h2d(c, c.type.p, hty.data(), sizeof(uint32_t) * n);
    h2d(c, c.mass.p, mass, sizeof(double) * n);
    h2d(c, c.cutr.p, cutr, sizeof(double) * n);
    dzero(c, c.image.p, sizeof(int4) * np);
    h2d(c, c.build_pos.p, hp.data(), sizeof(double4) * n);
    c.bxq_valid = false;
    if (mixed)
        dzero(c, c.f32.p, sizeof(float) * 3 * np);
    else
        dzero(c, c.f64.p, sizeof(double) * 3 * np);
    if (dem) {
        h2d(c, c.omega.p, hw.data(), sizeof(double4) * n);
        h2d(c, c.quat.p, hq.data(), sizeof(double4) * n);
        if (mixed)
            dzero(c, c.t32.p, sizeof(float) * 3 * np);
        else
            dzero(c, c.t64.p, sizeof(double) * 3 * np);
    }
    c.list_valid = false;
    c.list_injected = false;
    c.setup_done = false;
    c.nh_ready = false;
    c.vir_valid = false;
    c.step = 0;
    c.rebuilds = 0;
    c.nghost = 0;
    c.chunk = 8;
    c.last_rebuild = 0;
    c.rb_interval = 0.0;
    k_status_init<<<1, 1, 0, c.stream>>>(c.status.p);
    MDD_HIP(hipStreamSynchronize(c.stream));
}

// the staged systems of a batched ensemble as one particle set
void commit_systems(Ctx& c) {
    c.systems_dirty = false;
    const int ns = int(c.staged.size());
    int64_t n = 0;
    for (const auto& S : c.staged) n += int64_t(S.tag.size());
    std::vector<uint32_t> tag;
    std::vector<uint32_t> type(static_cast<size_t>(n), 0u);
    std::vector<uint32_t> psys(static_cast<size_t>(n), 0u);
    std::vector<double> pos, vel, omega, radius, cutr, mass;
    tag.reserve(n);
    c.sys_first.assign(size_t(ns) + 1, 0);
    c.sys_h.assign(size_t(ns), mdd::SysP{});
    c.sys_rmin.assign(size_t(ns), 0.0);
    c.sys_rmax.assign(size_t(ns), 0.0);
    for (int s = 0; s < ns; s++) {
        const auto& S = c.staged[s];
        const int64_t k = int64_t(S.tag.size());
        std::fill(psys.begin() + c.sys_first[s], psys.begin() + c.sys_first[s] + k, uint32_t(s));
        c.sys_first[s + 1] = c.sys_first[s] + k;
        tag.insert(tag.end(), S.tag.begin(), S.tag.end());
        pos.insert(pos.end(), S.pos.begin(), S.pos.end());
        vel.insert(vel.end(), S.vel.begin(), S.vel.end());
        omega.insert(omega.end(), S.omega.begin(), S.omega.end());
        radius.insert(radius.end(), S.radius.begin(), S.radius.end());
        cutr.insert(cutr.end(), S.cutr.begin(), S.cutr.end());
        mass.insert(mass.end(), S.mass.begin(), S.mass.end());
        mdd::SysP& P = c.sys_h[s];
        for (int d = 0; d < 3; d++) {
            P.box.lo[d] = S.lo[d];
            P.box.hi[d] = S.hi[d];
            P.box.L[d] = S.hi[d] - S.lo[d];
            P.box.periodic[d] = 1;
        }
        double k4[4];
        mdd::pair_constants(S.youngs, S.poisson, S.restitution, S.friction, S.youngs, S.poisson,
                            S.restitution, S.friction, k4);
        P.beta = k4[0];
        P.mu = k4[1];
        P.eeff = k4[2];
        P.geff = k4[3];
        c.sys_rmin[s] = *std::min_element(S.cutr.begin(), S.cutr.end());
        c.sys_rmax[s] = *std::max_element(S.cutr.begin(), S.cutr.end());
    }
    pack_upload(c, n, tag.data(), type.data(), pos.data(), vel.data(), omega.data(),
                radius.data(), cutr.data(), mass.data());
    c.psys.resize(size_t(c.npad));
    h2d(c, c.psys.p, psys.data(), sizeof(uint32_t) * size_t(n));
    c.nsys = ns;
    c.setup_done = false;
}

template <class T>
std::vector<T> dl(Ctx& c, const mdd::DBuf<T>& b, size_t n) {
    std::vector<T> h(n);
    if (n) d2h(c, h.data(), b.p, sizeof(T) * n);
    return h;
}

}  // namespace

extern "C" {

int mdd_version(void) { return 1; }

int mdd_create(mdd_ctx** out, const mdd_params* p) {
    if (!out || !p) return MDD_ERR_INVALID;
    *out = nullptr;
    if (p->model != MDD_MODEL_LJ && p->model != MDD_MODEL_DEM) return MDD_ERR_INVALID;
    if (p->precision != MDD_FP64 && p->precision != MDD_MIXED) return MDD_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MDD_ERR_NO_DEVICE;
    if (p->device < 0 || p->device >= ndev) return MDD_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) return MDD_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MDD_ERR_NO_DEVICE;
    mdd_ctx* ctx = nullptr;
    try {
        ctx = new mdd_ctx();
    } catch (...) {
        return MDD_ERR_CAPACITY;
    }
    Ctx& c = ctx->c;
    c.device = p->device;
    c.model = p->model;
    c.precision = p->precision;
    c.cap = p->capacity > 0 ? p->capacity : 0;
    c.vv_on = p->model == MDD_MODEL_DEM;
    c.rot_on = p->model == MDD_MODEL_DEM;
    for (int d = 0; d < 3; d++) {
        c.box.lo[d] = 0.0;
        c.box.hi[d] = 1.0;
        c.box.L[d] = 1.0;
        c.box.periodic[d] = 1;
    }
    // LJForcePlugin default table (md_lj/src/lib.rs:313-315)
    c.lj1 = 48.0;
    c.lj2 = 24.0;
    c.lj_cut2 = 6.25;
    c.lj_e12 = 4.0;
    c.lj_e6 = 4.0;
    int rc = guarded(ctx, [&](Ctx& cc) {
        MDD_HIP(hipStreamCreateWithFlags(&cc.stream, hipStreamNonBlocking));
        cc.status.resize(1);
        MDD_HIP(hipHostMalloc(reinterpret_cast<void**>(&cc.status_host), sizeof(mdd::Status),
                              hipHostMallocDefault));
        std::memset(cc.status_host, 0, sizeof(mdd::Status));
        k_status_init<<<1, 1, 0, cc.stream>>>(cc.status.p);
        MDD_HIP(hipStreamSynchronize(cc.stream));
        mdd::init_thermo_static_scratch(cc.device);
    });
    if (rc != MDD_OK) {
        mdd_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return MDD_OK;
}

int mdd_destroy(mdd_ctx* ctx) {
    if (!ctx) return MDD_OK;
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (hipEvent_t e : c.ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c.rb_ev) (void)hipEventDestroy(e);
    if (c.rb_mark) (void)hipEventDestroy(c.rb_mark);
    mdd::release_graphs(c);
    if (c.status_host) (void)hipHostFree(c.status_host);
    c.status_host = nullptr;
    hipStream_t s = c.stream;
    delete ctx;  // device buffers free themselves
    if (s) (void)hipStreamDestroy(s);
    return MDD_OK;
}

const char* mdd_last_error(const mdd_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int mdd_set_box(mdd_ctx* ctx, const double lo[3], const double hi[3], const int periodic[3]) {
    return guarded(ctx, [&](Ctx& c) {
        if (!lo || !hi || !periodic) throw ApiError(MDD_ERR_INVALID, "null box argument");
        for (int d = 0; d < 3; d++) {
            if (!(hi[d] > lo[d])) throw ApiError(MDD_ERR_INVALID, "box hi must exceed lo");
            c.box.lo[d] = lo[d];
            c.box.hi[d] = hi[d];
            c.box.L[d] = hi[d] - lo[d];
            c.box.periodic[d] = periodic[d] ? 1 : 0;
        }
        c.setup_done = false;
        c.list_valid = false;
        c.xq_valid = false;
    });
}

int mdd_set_neighbor(mdd_ctx* ctx, double skin_fraction, double bin_size, int every, int check,
                     int sort_every, int newton) {
    return guarded(ctx, [&](Ctx& c) {
        if (!(skin_fraction >= 1.0)) throw ApiError(MDD_ERR_INVALID, "skin_fraction must be >= 1");
        if (every < 0) throw ApiError(MDD_ERR_INVALID, "every must be >= 0");
        c.skin_fraction = skin_fraction;
        c.bin_min_size = bin_size;
        c.every = every;
        c.check = check;
        c.sort_every = sort_every < 0 ? 0 : sort_every;
        c.newton = newton;
        c.setup_done = false;
        c.list_valid = false;
    });
}

int mdd_set_rebuild_on_box_exit(mdd_ctx* ctx, int enabled) {
    return guarded(ctx, [&](Ctx& c) { c.box_exit_rebuild = enabled ? 1 : 0; });
}

int mdd_set_virial(mdd_ctx* ctx, int every) {
    return guarded(ctx, [&](Ctx& c) {
        if (every < 0) throw ApiError(MDD_ERR_INVALID, "virial interval must be >= 0");
        c.virial_every = every;
    });
}

int mdd_set_dem_split(mdd_ctx* ctx, int mode) {
    return guarded(ctx, [&](Ctx& c) {
        if (mode < -1 || mode > 1) throw ApiError(MDD_ERR_INVALID, "split mode is -1, 0 or 1");
        c.dem_split = mode;
    });
}

int mdd_set_lj_positions(mdd_ctx* ctx, int mode) {
    return guarded(ctx, [&](Ctx& c) {
        if (mode != 0 && mode != 1) throw ApiError(MDD_ERR_INVALID, "position mode is 0 or 1");
        c.lj_fixed = mode == 1;
        c.xq_valid = false;
    });
}

int mdd_set_dem_tiles(mdd_ctx* ctx, int on) {
    return guarded(ctx, [&](Ctx& c) {
        c.tile_order = on != 0;
        if (c.tile_order && !c.tile_ok) c.list_valid = false;  // the next pass rebuilds
    });
}

int mdd_set_lj_post_fused(mdd_ctx* ctx, int on) {
    return guarded(ctx, [&](Ctx& c) { c.lj_post_fused = on != 0; });
}

int mdd_set_dem_wave(mdd_ctx* ctx, int on) {
    return guarded(ctx, [&](Ctx& c) { c.dem_wave = on != 0; });
}

int mdd_set_history_remap(mdd_ctx* ctx, int method) {
    return guarded(ctx, [&](Ctx& c) {
        if (method != 0 && method != 1) throw ApiError(MDD_ERR_INVALID, "remap method is 0 or 1");
        c.remap_sorted = method == 0;
    });
}

int mdd_set_history_mode(mdd_ctx* ctx, int mode) {
    return guarded(ctx, [&](Ctx& c) {
        if (mode != 0 && mode != 1) throw ApiError(MDD_ERR_INVALID, "history mode is 0 or 1");
        c.hist_by_owner = mode == 0 ? 1 : 0;
    });
}

int mdd_set_dt(mdd_ctx* ctx, double dt) {
    return guarded(ctx, [&](Ctx& c) {
        if (!(dt > 0.0)) throw ApiError(MDD_ERR_INVALID, "dt must be positive");
        c.dt = dt;
        if (c.nh_on) c.nh_ready = false;
    });
}

int mdd_set_materials(mdd_ctx* ctx, int n, const double* E, const double* nu, const double* e,
                      const double* mu) {
    return guarded(ctx, [&](Ctx& c) {
        if (n <= 0 || n > mdd::kMaxMat || !E || !nu || !e || !mu)
            throw ApiError(MDD_ERR_INVALID, "material table out of range");
        c.mat.n = n;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double k[4];
                mdd::pair_constants(E[i], nu[i], e[i], mu[i], E[j], nu[j], e[j], mu[j], k);
                c.mat.beta[i * n + j] = k[0];
                c.mat.mu[i * n + j] = k[1];
                c.mat.eeff[i * n + j] = k[2];
                c.mat.geff[i * n + j] = k[3];
            }
        mdd::upload_materials(c);
        // optional properties belong to the previous table: cleared
        std::memset(c.mext_host, 0, sizeof(c.mext_host));
        if (c.mext_dev.p) {
            c.mext_dev.resize(size_t(mdd::kNumExtras) * n * n);
            h2d(c, c.mext_dev.p, c.mext_host, sizeof(double) * mdd::kNumExtras * n * n);
        }
        if (c.ext_on) mdd::set_contact_path(c, c.cm_hooke != 0);
    });
}

int mdd_set_material_extras(mdd_ctx* ctx, int n, const double* values) {
    return guarded(ctx, [&](Ctx& c) {
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        if (c.mat.n == 0 || n != c.mat.n || !values)
            throw ApiError(MDD_ERR_INVALID, "set the materials first, one value per material");
        if (c.nsys > 0)
            throw ApiError(MDD_ERR_INVALID, "optional contact models need a single system");
        const int nn = n * n;
        auto v = [&](int x, int i) { return values[x * n + i]; };
        auto gmean = [](double a, double b) { return std::sqrt(a * b); };
        auto gmean0 = [](double a, double b) {
            return std::sqrt((a > 0.0 ? a : 0.0) * (b > 0.0 ? b : 0.0));
        };
        auto hmean = [](double a, double b, bool either) {
            if (a > 0.0 && b > 0.0) return 2.0 * a * b / (a + b);
            if (!either) return 0.0;
            return a > 0.0 ? a : b;
        };
        // MaterialTable::build_pair_tables, optional fields (dem_atom/src/lib.rs:425-510)
        bool any = false;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double* t = c.mext_host;
                const int k = i * n + j;
                t[0 * nn + k] = gmean(v(0, i), v(0, j));    // rolling friction
                t[1 * nn + k] = gmean0(v(1, i), v(1, j));   // twisting friction
                t[2 * nn + k] = gmean(v(2, i), v(2, j));    // cohesion energy
                t[3 * nn + k] = gmean(v(3, i), v(3, j));    // surface energy
                t[4 * nn + k] = hmean(v(4, i), v(4, j), false);  // Hooke k_n
                t[5 * nn + k] = hmean(v(5, i), v(5, j), false);  // Hooke k_t
                t[6 * nn + k] = hmean(v(6, i), v(6, j), true);   // SDS rolling stiffness
                t[7 * nn + k] = gmean0(v(7, i), v(7, j));   // SDS rolling damping
                t[8 * nn + k] = hmean(v(8, i), v(8, j), true);   // SDS twisting stiffness
                t[9 * nn + k] = gmean0(v(9, i), v(9, j));   // SDS twisting damping
            }
        for (int x = 0; x < mdd::kNumExtras * nn; x++) any = any || c.mext_host[x] != 0.0;
        c.mext_dev.resize(size_t(mdd::kNumExtras) * nn);
        h2d(c, c.mext_dev.p, c.mext_host, sizeof(double) * mdd::kNumExtras * nn);
        mdd::set_contact_path(c, any || c.cm_hooke);
    });
}

int mdd_set_contact_models(mdd_ctx* ctx, int hooke, int dmt, int rolling_sds, int twisting_sds) {
    return guarded(ctx, [&](Ctx& c) {
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        if (c.nsys > 0)
            throw ApiError(MDD_ERR_INVALID, "optional contact models need a single system");
        c.cm_hooke = hooke ? 1 : 0;
        c.cm_dmt = dmt ? 1 : 0;
        c.cm_roll_sds = rolling_sds ? 1 : 0;
        c.cm_twist_sds = twisting_sds ? 1 : 0;
        if (c.mext_dev.p == nullptr) {
            // no optional properties given: every table is zero
            const int nn = c.mat.n * c.mat.n;
            if (nn == 0) throw ApiError(MDD_ERR_INVALID, "set the materials first");
            std::memset(c.mext_host, 0, sizeof(c.mext_host));
            c.mext_dev.resize(size_t(mdd::kNumExtras) * nn);
            h2d(c, c.mext_dev.p, c.mext_host, sizeof(double) * mdd::kNumExtras * nn);
        }
        bool any = c.cm_hooke != 0;
        for (int x = 0; x < mdd::kNumExtras * c.mat.n * c.mat.n; x++)
            any = any || c.mext_host[x] != 0.0;
        mdd::set_contact_path(c, any);
    });
}

int mdd_set_lj(mdd_ctx* ctx, double epsilon, double sigma, double cutoff) {
    return guarded(ctx, [&](Ctx& c) {
        // LJPairCoeffs::from_params (md_lj/src/lib.rs:230-239)
        const double s2 = sigma * sigma;
        const double sigma6 = (s2 * s2) * s2;
        c.lj1 = 48.0 * epsilon * sigma6 * sigma6;
        c.lj2 = 24.0 * epsilon * sigma6;
        const double cs = cutoff * sigma;
        c.lj_cut2 = cs * cs;
        c.lj_e12 = 4.0 * epsilon * sigma6 * sigma6;
        c.lj_e6 = 4.0 * epsilon * sigma6;
        c.lj_eps = epsilon;
        c.lj_sigma = sigma;
    });
}

int mdd_set_nose_hoover(mdd_ctx* ctx, int enabled, double temperature, double coupling) {
    return guarded(ctx, [&](Ctx& c) {
        c.nh_on = enabled ? 1 : 0;
        c.nh_T = temperature;
        c.nh_tau = coupling;
        c.nh_ready = false;
        if (c.nh_on) c.vv_on = 0;  // NH fuses the Verlet half steps
    });
}

int mdd_set_integrators(mdd_ctx* ctx, int velocity_verlet, int rotation) {
    return guarded(ctx, [&](Ctx& c) {
        c.vv_on = velocity_verlet ? 1 : 0;
        c.rot_on = rotation ? 1 : 0;
    });
}

int mdd_set_gravity(mdd_ctx* ctx, int enabled, const double g[3]) {
    return guarded(ctx, [&](Ctx& c) {
        c.gravity_on = enabled ? 1 : 0;
        if (g)
            for (int d = 0; d < 3; d++) c.g[d] = g[d];
    });
}

}  // extern "C"

namespace {

// The wall table changes: the device copy is replaced and the current wall
// state (points, velocities, time, force sums) is kept for walls that stay;
// `fresh` walls start at their configured point and velocity.
void sync_walls(Ctx& c, int fresh_from) {
    mdd::WallsDyn d{};
    if (c.wdyn.p) {
        MDD_HIP(hipStreamSynchronize(c.stream));
        d2h(c, &d, c.wdyn.p, sizeof(d));
        MDD_HIP(hipStreamSynchronize(c.stream));
    }
    for (int k = fresh_from; k < c.walls.count; k++) {
        const mdd::Wall& w = c.walls.w[k];
        for (int q = 0; q < 3; q++) {
            d.w[k].p[q] = w.p[q];
            d.w[k].v[q] = w.motion == 1 ? w.v[q] : 0.0;
        }
        d.w[k].acc = 0ull;
    }
    c.walls_moving = false;
    for (int k = 0; k < c.walls.count; k++)
        c.walls_moving = c.walls_moving || (c.walls.w[k].kind == 0 && c.walls.w[k].motion != 0);
    c.walls_dev.resize(1);
    c.wdyn.resize(1);
    h2d(c, c.walls_dev.p, &c.walls, sizeof(c.walls));
    h2d(c, c.wdyn.p, &d, sizeof(d));
}

mdd::Wall& new_wall(Ctx& c, int material) {
    if (c.walls.count >= mdd::kMaxWalls) throw ApiError(MDD_ERR_INVALID, "too many walls");
    if (material < 0 || material >= mdd::kMaxMat)
        throw ApiError(MDD_ERR_INVALID, "wall material out of range");
    mdd::Wall& w = c.walls.w[c.walls.count];
    std::memset(&w, 0, sizeof(w));
    w.mat = material;
    w.active = 1;
    for (int d = 0; d < 3; d++) {
        w.blo[d] = -1.0e300;
        w.bhi[d] = 1.0e300;
    }
    return w;
}

mdd::Wall& wall_at(Ctx& c, int wall) {
    if (wall < 0 || wall >= c.walls.count) throw ApiError(MDD_ERR_INVALID, "no such wall");
    return c.walls.w[wall];
}

}  // namespace

extern "C" {

int mdd_add_wall(mdd_ctx* ctx, const double point[3], const double normal[3], int material,
                 const double blo[3], const double bhi[3]) {
    return guarded(ctx, [&](Ctx& c) {
        if (!point || !normal) throw ApiError(MDD_ERR_INVALID, "null wall argument");
        // WallPlugin::build normalises the plane normal (dem_wall/src/lib.rs:688-699)
        const double mag =
            std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
        if (!(mag > 1e-15)) throw ApiError(MDD_ERR_INVALID, "wall normal vector must be non-zero");
        mdd::Wall& w = new_wall(c, material);
        for (int d = 0; d < 3; d++) {
            w.p[d] = point[d];
            w.n[d] = normal[d] / mag;
            if (blo) w.blo[d] = blo[d];
            if (bhi) w.bhi[d] = bhi[d];
        }
        c.walls.count++;
        sync_walls(c, c.walls.count - 1);
    });
}

int mdd_add_cylinder_wall(mdd_ctx* ctx, int axis, double c0, double c1, double radius,
                          double lo, double hi, int inside, int material) {
    return guarded(ctx, [&](Ctx& c) {
        if (axis < 0 || axis > 2) throw ApiError(MDD_ERR_INVALID, "cylinder axis must be 0, 1 or 2");
        mdd::Wall& w = new_wall(c, material);
        w.kind = 1;
        w.axis = axis;
        w.c[0] = c0;
        w.c[1] = c1;
        w.radius = radius;
        w.lo = lo;
        w.hi = hi;
        w.inside = inside ? 1 : 0;
        c.walls.count++;
        sync_walls(c, c.walls.count - 1);
    });
}

int mdd_add_sphere_wall(mdd_ctx* ctx, const double center[3], double radius, int inside,
                        int material) {
    return guarded(ctx, [&](Ctx& c) {
        if (!center) throw ApiError(MDD_ERR_INVALID, "null wall argument");
        mdd::Wall& w = new_wall(c, material);
        w.kind = 2;
        for (int d = 0; d < 3; d++) w.c[d] = center[d];
        w.radius = radius;
        w.inside = inside ? 1 : 0;
        c.walls.count++;
        sync_walls(c, c.walls.count - 1);
    });
}

int mdd_add_region_wall(mdd_ctx* ctx, int nprim, const int* types, const double* params,
                        int inside, int material) {
    return guarded(ctx, [&](Ctx& c) {
        if (nprim < 1 || nprim > mdd::kMaxRegionPrims || !types || !params)
            throw ApiError(MDD_ERR_INVALID, "a region wall takes 1..8 primitives");
        for (int k = 0; k < nprim; k++) {
            if (types[k] < 0 || types[k] > 4)
                throw ApiError(MDD_ERR_INVALID, "region primitive type must be 0..4");
            const int axis = types[k] == 2 ? int(params[12 * k + 3])
                                           : (types[k] == 3 ? int(params[12 * k + 2]) : 0);
            if (axis < 0 || axis > 2) throw ApiError(MDD_ERR_INVALID, "region axis must be 0..2");
        }
        mdd::Wall& w = new_wall(c, material);
        w.kind = 3;
        w.nprim = nprim;
        for (int k = 0; k < nprim; k++) {
            w.ptype[k] = types[k];
            for (int q = 0; q < 12; q++) w.prm[k][q] = params[12 * k + q];
        }
        w.inside = inside ? 1 : 0;
        c.walls.count++;
        sync_walls(c, c.walls.count - 1);
    });
}

int mdd_set_wall_motion(mdd_ctx* ctx, int wall, int motion, const double* params) {
    return guarded(ctx, [&](Ctx& c) {
        mdd::Wall& w = wall_at(c, wall);
        if (w.kind != 0) throw ApiError(MDD_ERR_INVALID, "only plane walls move");
        if (motion < 0 || motion > 3) throw ApiError(MDD_ERR_INVALID, "wall motion must be 0..3");
        if (motion != 0 && !params) throw ApiError(MDD_ERR_INVALID, "null motion parameters");
        w.motion = motion;
        w.v[0] = w.v[1] = w.v[2] = 0.0;
        w.a0 = w.a1 = w.a2 = 0.0;
        if (motion == 1)
            for (int d = 0; d < 3; d++) w.v[d] = params[d];
        if (motion >= 2) {
            w.a0 = params[0];
            w.a1 = params[1];
            if (motion == 3) w.a2 = params[2];
        }
        // the wall restarts from its configured point
        sync_walls(c, wall);
    });
}

int mdd_set_wall_active(mdd_ctx* ctx, int wall, int active) {
    return guarded(ctx, [&](Ctx& c) {
        wall_at(c, wall).active = active ? 1 : 0;
        sync_walls(c, c.walls.count);
    });
}

int mdd_get_walls(mdd_ctx* ctx, double* force, double* point, double* velocity, int* count) {
    return guarded(ctx, [&](Ctx& c) {
        mdd::WallsDyn d{};
        if (c.wdyn.p) {
            MDD_HIP(hipStreamSynchronize(c.stream));
            d2h(c, &d, c.wdyn.p, sizeof(d));
            MDD_HIP(hipStreamSynchronize(c.stream));
        }
        for (int k = 0; k < c.walls.count; k++) {
            if (force)
                force[k] = double(static_cast<long long>(d.w[k].acc)) / mdd::kWallAccScale;
            for (int q = 0; q < 3; q++) {
                if (point) point[3 * k + q] = d.w[k].p[q];
                if (velocity) velocity[3 * k + q] = d.w[k].v[q];
            }
        }
        if (count) *count = c.walls.count;
    });
}

int mdd_clear_walls(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        c.walls.count = 0;
        c.walls_moving = false;
        c.wdyn.release();
        c.walls_dev.release();
    });
}

int mdd_upload_state(mdd_ctx* ctx, int64_t n, const uint32_t* tag, const uint32_t* type,
                     const double* pos, const double* vel, const double* omega,
                     const double* radius, const double* cutoff_radius, const double* mass) {
    return guarded(ctx, [&](Ctx& c) {
        c.staged.clear();
        c.systems_dirty = false;
        c.nsys = 0;
        pack_upload(c, n, tag, type, pos, vel, omega, radius, cutoff_radius, mass);
    });
}

int mdd_add_system(mdd_ctx* ctx, const double lo[3], const double hi[3], double youngs,
                   double poisson, double restitution, double friction, int64_t n,
                   const uint32_t* tag, const double* pos, const double* vel,
                   const double* omega, const double* radius, const double* cutoff_radius,
                   const double* mass) {
    return guarded(ctx, [&](Ctx& c) {
        if (c.model != MDD_MODEL_DEM)
            throw ApiError(MDD_ERR_UNSUPPORTED, "ensembles hold DEM systems");
        if (n <= 0 || !lo || !hi || !tag || !pos || !vel || !radius || !cutoff_radius || !mass)
            throw ApiError(MDD_ERR_INVALID, "bad system arguments");
        if (c.nsys == 0 && c.n > 0 && c.staged.empty())
            throw ApiError(MDD_ERR_INVALID, "context already holds a single system");
        Ctx::Staged S;
        for (int d = 0; d < 3; d++) {
            if (!(hi[d] > lo[d])) throw ApiError(MDD_ERR_INVALID, "box hi must exceed lo");
            S.lo[d] = lo[d];
            S.hi[d] = hi[d];
        }
        S.youngs = youngs;
        S.poisson = poisson;
        S.restitution = restitution;
        S.friction = friction;
        S.tag.assign(tag, tag + n);
        S.pos.assign(pos, pos + 3 * n);
        S.vel.assign(vel, vel + 3 * n);
        S.omega = omega ? std::vector<double>(omega, omega + 3 * n) : std::vector<double>(3 * n, 0.0);
        S.radius.assign(radius, radius + n);
        S.cutr.assign(cutoff_radius, cutoff_radius + n);
        S.mass.assign(mass, mass + n);
        c.staged.push_back(std::move(S));
        c.systems_dirty = true;
        c.list_valid = false;
    });
}

int mdd_nsystems(mdd_ctx* ctx) {
    if (!ctx) return MDD_ERR_INVALID;
    int ns = 0;
    const int rc = guarded(ctx, [&](Ctx& c) {
        need_state(c);
        ns = c.nsys;
    });
    return rc == MDD_OK ? ns : rc;
}

int mdd_system_range(mdd_ctx* ctx, int sys, int64_t* first, int64_t* count) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (sys < 0 || sys >= c.nsys) throw ApiError(MDD_ERR_INVALID, "no such system");
        if (first) *first = c.sys_first[sys];
        if (count) *count = c.sys_first[sys + 1] - c.sys_first[sys];
    });
}

int64_t mdd_nlocal(const mdd_ctx* ctx) { return ctx ? ctx->c.n : -1; }

int mdd_download_properties(mdd_ctx* ctx, uint32_t* type, double* mass, double* radius,
                            double* cutoff_radius) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n;
        if (type) {
            auto h = dl(c, c.type, n);
            std::copy(h.begin(), h.end(), type);
        }
        if (mass) {
            auto h = dl(c, c.mass, n);
            std::copy(h.begin(), h.end(), mass);
        }
        if (cutoff_radius) {
            auto h = dl(c, c.cutr, n);
            std::copy(h.begin(), h.end(), cutoff_radius);
        }
        if (radius) {
            auto h = dl(c, c.pos, n);
            for (int64_t i = 0; i < n; i++) radius[i] = h[i].w;
        }
    });
}

int mdd_download_state(mdd_ctx* ctx, uint32_t* tag, double* pos, double* vel, double* omega,
                       double* quat, int32_t* image) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n;
        if (tag) {
            auto t = dl(c, c.tag, n);
            std::copy(t.begin(), t.end(), tag);
        }
        if (pos) {
            auto h = dl(c, c.pos, n);
            for (int64_t i = 0; i < n; i++) {
                pos[3 * i] = h[i].x;
                pos[3 * i + 1] = h[i].y;
                pos[3 * i + 2] = h[i].z;
            }
        }
        if (vel) {
            auto h = dl(c, c.vel, n);
            for (int64_t i = 0; i < n; i++) {
                vel[3 * i] = h[i].x;
                vel[3 * i + 1] = h[i].y;
                vel[3 * i + 2] = h[i].z;
            }
        }
        if (omega) {
            if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "no angular state");
            auto h = dl(c, c.omega, n);
            for (int64_t i = 0; i < n; i++) {
                omega[3 * i] = h[i].x;
                omega[3 * i + 1] = h[i].y;
                omega[3 * i + 2] = h[i].z;
            }
        }
        if (quat) {
            if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "no orientation state");
            auto h = dl(c, c.quat, n);
            for (int64_t i = 0; i < n; i++) {
                quat[4 * i] = h[i].x;
                quat[4 * i + 1] = h[i].y;
                quat[4 * i + 2] = h[i].z;
                quat[4 * i + 3] = h[i].w;
            }
        }
        if (image) {
            auto h = dl(c, c.image, n);
            for (int64_t i = 0; i < n; i++) {
                image[3 * i] = h[i].x;
                image[3 * i + 1] = h[i].y;
                image[3 * i + 2] = h[i].z;
            }
        }
    });
}

int mdd_download_forces(mdd_ctx* ctx, double* force, double* torque) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n, np = c.npad;
        const bool mixed = c.precision == MDD_MIXED;
        auto take = [&](double* out, bool tq) {
            if (!out) return;
            if (tq && c.model != MDD_MODEL_DEM)
                throw ApiError(MDD_ERR_INVALID, "no torque state");
            if (mixed) {
                auto h = dl(c, tq ? c.t32 : c.f32, size_t(3 * np));
                for (int64_t i = 0; i < n; i++)
                    for (int d = 0; d < 3; d++) out[3 * i + d] = double(h[d * np + i]);
            } else {
                auto h = dl(c, tq ? c.t64 : c.f64, size_t(3 * np));
                for (int64_t i = 0; i < n; i++)
                    for (int d = 0; d < 3; d++) out[3 * i + d] = h[d * np + i];
            }
        };
        take(force, false);
        take(torque, true);
    });
}

int mdd_upload_forces(mdd_ctx* ctx, const double* force, const double* torque) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n, np = c.npad;
        const bool mixed = c.precision == MDD_MIXED;
        auto put = [&](const double* in, bool tq) {
            if (!in) return;
            if (tq && c.model != MDD_MODEL_DEM)
                throw ApiError(MDD_ERR_INVALID, "no torque state");
            if (mixed) {
                std::vector<float> h(size_t(3 * np), 0.f);
                for (int64_t i = 0; i < n; i++)
                    for (int d = 0; d < 3; d++) h[d * np + i] = float(in[3 * i + d]);
                h2d(c, (tq ? c.t32 : c.f32).p, h.data(), sizeof(float) * h.size());
            } else {
                std::vector<double> h(size_t(3 * np), 0.0);
                for (int64_t i = 0; i < n; i++)
                    for (int d = 0; d < 3; d++) h[d * np + i] = in[3 * i + d];
                h2d(c, (tq ? c.t64 : c.f64).p, h.data(), sizeof(double) * h.size());
            }
        };
        put(force, false);
        put(torque, true);
    });
}

int mdd_upload_quaternions(mdd_ctx* ctx, const double* quat) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (!quat) throw ApiError(MDD_ERR_INVALID, "null quaternions");
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "no orientation state");
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n;
        std::vector<double4> h(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; i++)
            h[i] = make_double4(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]);
        h2d(c, c.quat.p, h.data(), sizeof(double4) * size_t(n));
    });
}

int mdd_neighbor_setup(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        mdd::neighbor_setup(c);
    });
}

int mdd_bin_neighbor_list(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        mdd::rebuild(c);
        c.last_rebuild = c.step;
    });
}

int mdd_set_half_list(mdd_ctx* ctx, int64_t nlocal, const uint32_t* offsets,
                      const uint32_t* indices) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (nlocal != c.n || !offsets) throw ApiError(MDD_ERR_INVALID, "row count differs");
        if (offsets[c.n] > 0 && !indices) throw ApiError(MDD_ERR_INVALID, "null index array");
        std::vector<std::vector<uint32_t>> own(c.n), rev(c.n);
        for (int64_t i = 0; i < c.n; i++) {
            if (offsets[i + 1] < offsets[i]) throw ApiError(MDD_ERR_INVALID, "offsets decrease");
            for (uint32_t k = offsets[i]; k < offsets[i + 1]; k++) {
                const uint32_t j = indices[k];
                if (int64_t(j) >= c.n || int64_t(j) == i)
                    throw ApiError(MDD_ERR_INVALID, "half list index out of range");
                own[i].push_back(mdd::enc_entry(j, mdd::kCodeNone, false));
                rev[j].push_back(mdd::enc_entry(uint32_t(i), mdd::kCodeNone, true));
            }
        }
        int mx = 1;
        for (int64_t i = 0; i < c.n; i++) mx = std::max<int>(mx, int(own[i].size() + rev[i].size()));
        if (c.model == MDD_MODEL_DEM && mx > 64)
            throw ApiError(MDD_ERR_UNSUPPORTED, "DEM rows above 64 entries");
        const int G = 1 << c.lg;
        mx = (mx + G - 1) / G * G;
        c.maxnb = mx;
        const int64_t np = c.npad;
        std::vector<uint32_t> hn(size_t(mx) * np, 0u), cnt(np, 0u), half(np, 0u);
        for (int64_t i = 0; i < c.n; i++) {
            int k = 0;
            for (uint32_t e : own[i]) hn[mdd::row_addr(i, k++, np, c.lg)] = e;
            for (uint32_t e : rev[i]) hn[mdd::row_addr(i, k++, np, c.lg)] = e;
            cnt[i] = uint32_t(k);
            half[i] = uint32_t(own[i].size());
        }
        c.nbr.resize(hn.size());
        c.nnb.resize(np);
        c.nhalf.resize(np);
        h2d(c, c.nbr.p, hn.data(), sizeof(uint32_t) * hn.size());
        h2d(c, c.nnb.p, cnt.data(), sizeof(uint32_t) * np);
        h2d(c, c.nhalf.p, half.data(), sizeof(uint32_t) * np);
        if (c.model == MDD_MODEL_DEM) {
            c.hist.resize(size_t(c.hist_slots) * mx * np);
            c.cmask.resize(np);
            dzero(c, c.cmask.p, sizeof(uint64_t) * np);
        }
        c.list_valid = true;
        c.list_injected = true;
        c.listed_ok = false;
        c.list_tile = false;
        c.list_lg = c.lg;
        c.npairs = offsets[c.n];
    });
}

int64_t mdd_npairs(mdd_ctx* ctx) {
    if (!ctx || !ctx->c.list_valid) return 0;
    int64_t total = 0;
    int rc = guarded(ctx, [&](Ctx& c) {
        MDD_HIP(hipStreamSynchronize(c.stream));
        auto h = dl(c, c.nhalf, size_t(c.n));
        for (uint32_t v : h) total += v;
    });
    return rc == MDD_OK ? total : rc;
}

int mdd_download_pairs(mdd_ctx* ctx, int64_t cap, uint32_t* owner_tag, uint32_t* partner_tag,
                       uint8_t* code, int64_t* n_out) {
    return guarded(ctx, [&](Ctx& c) {
        need_list(c);
        if (c.list_tile) mdd::expand_tile_rows(c);
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n, np = c.npad;
        auto nn = dl(c, c.nnb, size_t(n));
        auto nb = dl(c, c.nbr, size_t(c.maxnb) * np);
        auto tg = dl(c, c.tag, size_t(n));
        int64_t k = 0;
        for (int64_t i = 0; i < n; i++) {
            const int cnt = std::min<int>(int(nn[i]), c.maxnb);
            for (int m = 0; m < cnt; m++) {
                const uint32_t e = nb[mdd::row_addr(i, m, np, c.list_lg)];
                if (mdd::entry_rev(e)) continue;
                if (k < cap) {
                    if (owner_tag) owner_tag[k] = tg[i];
                    if (partner_tag) partner_tag[k] = tg[mdd::entry_idx(e)];
                    if (code) code[k] = uint8_t(mdd::entry_code(e));
                }
                k++;
            }
        }
        if (n_out) *n_out = k;
    });
}

int mdd_lj_force(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_list(c);
        if (c.model != MDD_MODEL_LJ) throw ApiError(MDD_ERR_INVALID, "not an LJ context");
        mdd::launch_pair_forces(c, -1, false);
    });
}

int mdd_hertz_mindlin_contact_force(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_list(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        if (c.mat.n == 0 && c.nsys == 0) throw ApiError(MDD_ERR_INVALID, "materials not set");
        mdd::launch_pair_forces(c, -1, false);
        sync_status(c);
        check_device_error(c);
    });
}

int mdd_wall_contact_force(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        mdd::launch_walls_gravity(c, true, false);
    });
}

int mdd_wall_move(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) { mdd::launch_wall_systems(c, true, false); });
}

int mdd_wall_zero_force_accumulators(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) { mdd::launch_wall_systems(c, false, true); });
}

int mdd_apply_gravity(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        mdd::launch_walls_gravity(c, false, true);
    });
}

int mdd_initial_integration(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        mdd::launch_initial_integration(c, -1, false);
    });
}

int mdd_final_integration(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        mdd::launch_final_integration(c);
    });
}

int mdd_initial_rotation(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        mdd::launch_initial_rotation(c);
    });
}

int mdd_final_rotation(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        mdd::launch_final_rotation(c);
    });
}

int mdd_setup_nose_hoover(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        setup_nose_hoover(c);
    });
}

int mdd_nh_pre_initial(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (!c.nh_ready) setup_nose_hoover(c);
        mdd::launch_nh_pre_standalone(c);
    });
}

int mdd_nh_post_final(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (!c.nh_ready) throw ApiError(MDD_ERR_INVALID, "thermostat not set up");
        mdd::launch_nh_post_standalone(c);
    });
}

int mdd_run(mdd_ctx* ctx, int64_t nsteps) {
    return guarded(ctx, [&](Ctx& c) {
        if (nsteps < 0) throw ApiError(MDD_ERR_INVALID, "negative step count");
        run_steps(c, nsteps);
    });
}

int mdd_download_history(mdd_ctx* ctx, int64_t cap, uint32_t* tag_lo, uint32_t* tag_hi,
                         double* spring, int64_t* n_out) {
    return guarded(ctx, [&](Ctx& c) {
        need_list(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n, np = c.npad;
        const int64_t cs = int64_t(c.maxnb) * np;
        auto nn = dl(c, c.nnb, size_t(n));
        auto nb = dl(c, c.nbr, size_t(cs));
        auto tg = dl(c, c.tag, size_t(n));
        auto mk = dl(c, c.cmask, size_t(n));
        auto hs = dl(c, c.hist, size_t(3 * cs));
        int64_t k = 0;
        for (int64_t i = 0; i < n; i++) {
            const int cnt = std::min<int>(int(nn[i]), c.maxnb);
            for (int m = 0; m < cnt; m++) {
                if (!((mk[i] >> m) & 1ull)) continue;
                const int64_t slot = mdd::row_addr(i, m, np, c.list_lg);
                const uint32_t e = nb[slot];
                if (mdd::entry_rev(e)) continue;
                if (k < cap) {
                    const uint32_t a = tg[i], b = tg[mdd::entry_idx(e)];
                    if (tag_lo) tag_lo[k] = std::min(a, b);
                    if (tag_hi) tag_hi[k] = std::max(a, b);
                    if (spring)
                        for (int d = 0; d < 3; d++) spring[3 * k + d] = hs[d * cs + slot];
                }
                k++;
            }
        }
        if (n_out) *n_out = k;
    });
}

int mdd_download_history7(mdd_ctx* ctx, int64_t cap, uint32_t* tag_lo, uint32_t* tag_hi,
                          double* hist7, int64_t* n_out) {
    return guarded(ctx, [&](Ctx& c) {
        need_list(c);
        if (c.model != MDD_MODEL_DEM) throw ApiError(MDD_ERR_INVALID, "not a DEM context");
        MDD_HIP(hipStreamSynchronize(c.stream));
        const int64_t n = c.n, np = c.npad;
        const int64_t cs = int64_t(c.maxnb) * np;
        const int ns = c.hist_slots;
        auto nn = dl(c, c.nnb, size_t(n));
        auto nb = dl(c, c.nbr, size_t(cs));
        auto tg = dl(c, c.tag, size_t(n));
        auto mk = dl(c, c.cmask, size_t(n));
        auto hs = dl(c, c.hist, size_t(ns * cs));
        int64_t k = 0;
        for (int64_t i = 0; i < n; i++) {
            const int cnt = std::min<int>(int(nn[i]), c.maxnb);
            for (int m = 0; m < cnt; m++) {
                if (!((mk[i] >> m) & 1ull)) continue;
                const int64_t slot = mdd::row_addr(i, m, np, c.list_lg);
                const uint32_t e = nb[slot];
                if (mdd::entry_rev(e)) continue;
                if (k < cap) {
                    const uint32_t a = tg[i], b = tg[mdd::entry_idx(e)];
                    if (tag_lo) tag_lo[k] = std::min(a, b);
                    if (tag_hi) tag_hi[k] = std::max(a, b);
                    if (hist7)
                        for (int d = 0; d < 7; d++)
                            hist7[7 * k + d] = d < ns ? hs[d * cs + slot] : 0.0;
                }
                k++;
            }
        }
        if (n_out) *n_out = k;
    });
}

// the step observables from the device totals of thermo_reduce (sum m v^2,
// sum m v, sum m, sum m |v - vbar|^2, half-list pairs, contact bits)
static void thermo_fill(Ctx& c, const double s[8], mdd_thermo* out) {
    sync_status(c);
    out->step = c.step;
    out->rebuilds = c.rebuilds;
    out->nghost = c.nghost;
    out->ke = 0.5 * s[0];
    const double ndof = 3.0 * double(c.n) - 3.0;
    out->temperature = ndof > 0 ? 2.0 * out->ke / ndof : 0.0;
    out->momentum[0] = s[1];
    out->momentum[1] = s[2];
    out->momentum[2] = s[3];
    out->granular_temperature = s[4] > 0 ? s[5] / (3.0 * s[4]) : 0.0;
    out->p_xi = c.status_host->p_xi;
    if (c.list_valid) {
        out->npairs = int64_t(s[6]);
        if (c.model == MDD_MODEL_DEM) out->ncontacts = int64_t(s[7]) / 2;
    }
}

int mdd_compute_thermo_noalloc(mdd_ctx* ctx, mdd_thermo* out) {
    return guarded(ctx, [&](Ctx& c) {
        if (!out) throw ApiError(MDD_ERR_INVALID, "null output");
        if (c.systems_dirty)
            throw ApiError(MDD_ERR_INVALID, "systems added since the last step: their first "
                                            "set-up allocates (call mdd_run or mdd_compute_thermo)");
        if (c.n <= 0) throw ApiError(MDD_ERR_INVALID, "no particles uploaded");
        std::memset(out, 0, sizeof(*out));
        double s[8];
        mdd::thermo_reduce_static(c, s);
        thermo_fill(c, s, out);
    });
}

int mdd_compute_thermo(mdd_ctx* ctx, int with_energy, mdd_thermo* out) {
    return guarded(ctx, [&](Ctx& c) {
        need_state(c);
        if (!out) throw ApiError(MDD_ERR_INVALID, "null output");
        std::memset(out, 0, sizeof(*out));
        double s[8];
        mdd::thermo_reduce(c, s);
        thermo_fill(c, s, out);
        if (c.model == MDD_MODEL_DEM && c.vir_valid && c.vir.p) {
            // the contact virial of the last pass that computed it
            MDD_HIP(hipStreamSynchronize(c.stream));
            auto w = dl(c, c.vir, size_t(6) * size_t(c.npad));
            for (int k = 0; k < 6; k++) {
                double t = 0.0;
                for (int64_t i = 0; i < c.n; i++) t += w[size_t(k) * size_t(c.npad) + size_t(i)];
                out->virial[k] = t;
            }
        }
        if (with_energy && c.model == MDD_MODEL_LJ && c.list_valid) {
            mdd::launch_pair_forces(c, -1, true);
            MDD_HIP(hipStreamSynchronize(c.stream));
            auto h = dl(c, c.pe, size_t(c.n));
            double pe = 0.0;
            for (double v : h) pe += v;
            out->pe = pe;
            auto w = dl(c, c.vir, size_t(6) * size_t(c.npad));
            for (int k = 0; k < 6; k++) {
                double t = 0.0;
                for (int64_t i = 0; i < c.n; i++) t += w[size_t(k) * size_t(c.npad) + size_t(i)];
                out->virial[k] = t;
            }
            // compute_pressure (md_measure/src/lib.rs:605-651) with the tail
            // of setup_lj_tails (md_lj/src/lib.rs:397-443, one type)
            const double n = double(c.n);
            const double vol = c.box.L[0] * c.box.L[1] * c.box.L[2];
            const double rho = n / vol;
            const double rc = std::sqrt(c.lj_cut2);
            const double sigma3 = (c.lj_sigma * c.lj_sigma) * c.lj_sigma;
            const double rc3 = (rc * rc) * rc;
            const double rc9 = (rc3 * rc3) * rc3;
            const double sigma9 = (sigma3 * sigma3) * sigma3;
            const double pi = 3.141592653589793;
            const double p_tail = 0.0 + 1.0 * (16.0 / 3.0) * pi * rho * rho * c.lj_eps * sigma3 *
                                            (2.0 * sigma9 / (3.0 * rc9) - sigma3 / rc3);
            const double trace = out->virial[0] + out->virial[1] + out->virial[2];
            out->pressure = rho * out->temperature - trace / (3.0 * vol) + p_tail;
        }
    });
}

int mdd_synchronize(mdd_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) { MDD_HIP(hipStreamSynchronize(c.stream)); });
}

int mdd_set_profiling(mdd_ctx* ctx, int every) {
    return guarded(ctx, [&](Ctx& c) {
        c.prof_every = every > 0 ? every : 0;
        if (c.prof_every && c.ev_pool.empty()) {
            c.ev_pool.resize(2 * kEventPool);
            c.prof_step.assign(kEventPool, 0);
            for (auto& e : c.ev_pool) MDD_HIP(hipEventCreate(&e));
            c.rb_ev.resize(3 * 64);
            c.rb_has_gap.assign(64, 0);
            for (auto& e : c.rb_ev) MDD_HIP(hipEventCreate(&e));
            MDD_HIP(hipEventCreate(&c.rb_mark));
        }
        c.prof_ms_sum = 0;
        c.prof_count = 0;
        c.prof_pending = 0;
        c.force_launches = 0;
        c.prof_ctr = 0;
        c.rb_ms_sum = 0;
        c.rb_gap_ms_sum = 0;
        c.rb_count = 0;
        c.rb_gap_count = 0;
        c.rb_pending = 0;
        c.rb_chunk_mark = false;
    });
}

int mdd_rebuild_stats(mdd_ctx* ctx, double* avg_ms, double* avg_gap_ms, int64_t* count) {
    return guarded(ctx, [&](Ctx& c) {
        if (avg_ms) *avg_ms = c.rb_count ? c.rb_ms_sum / double(c.rb_count) : 0.0;
        if (avg_gap_ms)
            *avg_gap_ms = c.rb_gap_count ? c.rb_gap_ms_sum / double(c.rb_gap_count) : 0.0;
        if (count) *count = c.rb_count;
    });
}

int mdd_force_kernel_stats(mdd_ctx* ctx, double* avg_ms, int64_t* launches) {
    return guarded(ctx, [&](Ctx& c) {
        if (avg_ms) *avg_ms = c.prof_count ? c.prof_ms_sum / double(c.prof_count) : 0.0;
        if (launches) *launches = c.force_launches;
    });
}

int mdd_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return MDD_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return MDD_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return MDD_ERR_HIP;
    if (free_bytes) *free_bytes = int64_t(f);
    if (total_bytes) *total_bytes = int64_t(t);
    return MDD_OK;
}

int mdd_set_build_kernel(mdd_ctx* ctx, int mode) {
    return guarded(ctx, [&](Ctx& c) {
        if (mode < 0 || mode > 2) throw ApiError(MDD_ERR_INVALID, "build kernel mode must be 0..2");
        c.build_mode = mode;
    });
}

int mdd_set_launch_graphs(mdd_ctx* ctx, int enabled) {
    return guarded(ctx, [&](Ctx& c) {
        c.graphs_on = enabled != 0;
        if (!c.graphs_on) {
            MDD_HIP(hipStreamSynchronize(c.stream));
            mdd::release_graphs(c);
        }
    });
}

int mdd_launch_graph_stats(mdd_ctx* ctx, int64_t* replays, int64_t* captures) {
    return guarded(ctx, [&](Ctx& c) {
        if (replays) *replays = c.graph_hits;
        if (captures) *captures = c.graph_captures;
    });
}

int mdd_set_force_lanes(mdd_ctx* ctx, int log2_lanes) {
    return guarded(ctx, [&](Ctx& c) {
        if (log2_lanes < -1 || log2_lanes > 3)
            throw ApiError(MDD_ERR_INVALID, "log2 lanes per particle must be -1..3");
        // the next step rebuilds the list in the new layout
        c.tile_enabled = log2_lanes < 0;
        if (log2_lanes >= 0) c.lg = log2_lanes;
    });
}

}  // extern "C"
