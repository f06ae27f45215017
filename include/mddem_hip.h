/*
 * mddem_hip.h -- C ABI of the MI355X (gfx950) particle hot path.
 *
 * One context = one particle system on one GPU, driven by one host thread
 * (the reference's single-threaded scheduler; distinct contexts may live on
 * distinct threads). All host buffers are caller-owned; device memory is
 * owned by the context. Every function returns an int status: 0 = ok,
 * negative = error (MDD_ERR_*); mdd_last_error() gives the message. Nothing
 * unwinds or aborts across the ABI.
 *
 * Each entry point names the reference system / type it stands in for
 * (SueHeir/MDDEM v0.1.3, path:line). Array arguments are row-major: "n*3"
 * means x,y,z per particle.
 */
#ifndef MDDEM_HIP_H
#define MDDEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDD_OK 0
#define MDD_ERR_INVALID (-1)     /* bad argument or call order */
#define MDD_ERR_HIP (-2)         /* HIP runtime error */
#define MDD_ERR_OVERLAP (-3)     /* > 500 excessive overlaps in one step
                                    (crates/dem_granular/src/contact.rs:197-211) */
#define MDD_ERR_CAPACITY (-4)    /* allocation failed */
#define MDD_ERR_UNSUPPORTED (-5) /* configuration outside this path */
#define MDD_ERR_NO_DEVICE (-6)   /* no usable gfx950 device */
#define MDD_ERR_STALLED (-7)     /* a step asked for a rebuild again right after
                                    its own rebuild (internal inconsistency) */

enum { MDD_MODEL_LJ = 0, MDD_MODEL_DEM = 1 };
enum { MDD_FP64 = 0, MDD_MIXED = 1 };

typedef struct mdd_ctx mdd_ctx;

typedef struct {
    int device;        /* HIP device ordinal */
    int model;         /* MDD_MODEL_LJ (LJDefaultPlugins) or MDD_MODEL_DEM
                          (GranularDefaultPlugins) */
    int precision;     /* MDD_FP64: all FP64. MDD_MIXED: FP32 pair forces and
                          accumulation, FP64 positions, velocities, thermostat
                          and shear displacement */
    int64_t capacity;  /* max particles; 0 = size on first upload */
} mdd_params;

typedef struct {
    int64_t step;          /* completed steps */
    int64_t rebuilds;      /* neighbor rebuilds so far */
    int64_t npairs;        /* half-list pairs (reference pair count) */
    int64_t nghost;        /* periodic images at the last rebuild */
    int64_t ncontacts;     /* DEM pairs in contact at the last force pass */
    double ke;             /* 0.5 sum m v^2 (crates/mddem_core/src/atom.rs:489-502) */
    double temperature;    /* 2 KE / (3N - 3) */
    double pe;             /* LJ 12-6 energy, unshifted (computed on request) */
    double p_xi;           /* Nose-Hoover thermostat momentum */
    double momentum[3];    /* sum m v */
    double granular_temperature; /* sum m |v - vbar|^2 / (3 M)
                                    (crates/dem_granular/src/granular_temp.rs:44-134) */
    double virial[6];      /* LJ pair virial sum d (x) F: xx yy zz xy xz yz
                              (md_lj/src/lib.rs:510-585; computed on request) */
    double pressure;       /* rho T - tr(W) / (3 V) + P_tail (md_measure/src/lib.rs:
                              605-651, tails md_lj/src/lib.rs:397-443; on request) */
} mdd_thermo;

/* ---- lifetime (App / resources) ---------------------------------------- */
int mdd_create(mdd_ctx** ctx, const mdd_params* params);
int mdd_destroy(mdd_ctx* ctx);
const char* mdd_last_error(const mdd_ctx* ctx);
int mdd_version(void);

/* ---- configuration --------------------------------------------------------- */
/* Domain box and periodicity: [domain] section, Domain::update_derived
 * (crates/mddem_core/src/domain.rs:105-160) */
int mdd_set_box(mdd_ctx* ctx, const double lo[3], const double hi[3], const int periodic[3]);
/* [neighbor] section, NeighborConfig (crates/mddem_core/src/neighbor.rs:45-82)
 * with needs_rebuild's policies (neighbor.rs:605-627): every = 0 rebuilds on
 * the displacement check; every > 0 also every `every` steps (check = 1) or
 * only then (check = 0); sort_every forces a sorted rebuild every that many
 * steps (sort_atoms_by_bin, neighbor.rs:686-794); newton selects nothing here
 * (the pair set is the newton half list; forces are gathered) */
int mdd_set_neighbor(mdd_ctx* ctx, double skin_fraction, double bin_size, int every,
                     int check, int sort_every, int newton);
/* 1 (default): rebuild when any particle leaves the periodic box, as
 * decide_rebuild does (neighbor.rs:651-666). 0: wrap lazily at displacement
 * rebuilds only (pair images stay valid between rebuilds). */
int mdd_set_rebuild_on_box_exit(mdd_ctx* ctx, int enabled);
/* How a contact's shear history crosses a neighbor rebuild. 0 (default): as
 * ContactHistoryStore does (crates/dem_granular/src/tangential.rs:54-133,
 * contact.rs:358-364, 575-589): the spring lives in the list of the particle
 * that owns the pair in the half list and is found by partner tag, so it
 * survives a rebuild (sort_atoms_by_bin, neighbor.rs:686-794) only when the
 * same particle still owns the pair; when the pair moves to the partner's
 * row the spring restarts from zero. 1: the spring is keyed by the pair and
 * survives whichever side owns it. */
int mdd_set_history_mode(mdd_ctx* ctx, int mode);
/* How the history is carried across a rebuild: 0 (default) the old list's
 * contacts become records sorted by their 64-bit pair key (lower, higher
 * particle index of the new order) and each new row slot finds its pair by a
 * binary search; 1 each particle searches its own old row by partner tag.
 * The results are identical. No reference counterpart: a method choice. */
int mdd_set_history_remap(mdd_ctx* ctx, int method);
/* Atom::dt; [run] dt or calculate_delta_time (crates/dem_atom/src/insert.rs:1392-1431) */
int mdd_set_dt(mdd_ctx* ctx, double dt);
/* [[dem.materials]] + MaterialTable::build_pair_tables (crates/dem_atom/src/lib.rs:389-515) */
int mdd_set_materials(mdd_ctx* ctx, int n, const double* youngs, const double* poisson,
                      const double* restitution, const double* friction);
/* Optional MaterialConfig fields (crates/dem_atom/src/lib.rs:89-118) and their
 * pair tables (build_pair_tables, lib.rs:425-510): values holds 10 arrays of
 * n, in this order: rolling_friction, twisting_friction, cohesion_energy,
 * surface_energy, kn, kt, rolling_stiffness, rolling_damping,
 * twisting_stiffness, twisting_damping. Call after mdd_set_materials (which
 * clears them). Any nonzero value selects the contact kernel of the whole
 * model family (contact.rs:103-615, 622-990) with 7 history values per
 * contact; switching the history width drops the stored history. Single
 * systems only (not with mdd_add_system). */
int mdd_set_material_extras(mdd_ctx* ctx, int n, const double* values);
/* [materials] contact_model "hooke" (1) or "hertz" (0), adhesion_model "dmt"
 * (1) or "jkr" (0), rolling_model and twisting_model "sds" (1) or "constant"
 * (0) (crates/dem_atom/src/lib.rs:142-157; contact.rs:65-90 picks the kernel) */
int mdd_set_contact_models(mdd_ctx* ctx, int hooke, int dmt, int rolling_sds, int twisting_sds);
/* [lj] section, LJPairCoeffs::from_params (crates/md_lj/src/lib.rs:230-239) */
int mdd_set_lj(mdd_ctx* ctx, double epsilon, double sigma, double cutoff);
/* [thermostat] section, NoseHooverPlugin (crates/md_thermostat/src/lib.rs:166-243) */
int mdd_set_nose_hoover(mdd_ctx* ctx, int enabled, double temperature, double coupling);
/* VelocityVerletPlugin (crates/mddem_verlet/src/lib.rs:46-85) and
 * RotationalDynamicsPlugin (crates/dem_granular/src/rotational.rs:49-56) */
int mdd_set_integrators(mdd_ctx* ctx, int velocity_verlet, int rotation);
/* [gravity] section, GravityPlugin (crates/mddem_fixes/src/lib.rs:735-761) */
int mdd_set_gravity(mdd_ctx* ctx, int enabled, const double g[3]);
/* [[wall]] entries (WallPlugin::build, crates/dem_wall/src/lib.rs:546-757).
 * Walls share one index space in the order they are added (at most 12).
 * Plane: WallPlane (lib.rs:302-357), the normal is normalised (a zero normal
 * is an error), bounds may be NULL (unbounded). */
int mdd_add_wall(mdd_ctx* ctx, const double point[3], const double normal[3], int material,
                 const double bound_lo[3], const double bound_hi[3]);
/* type = "cylinder", WallCylinder (lib.rs:372-405): axis 0/1/2 = x/y/z,
 * center (c0, c1) in the plane normal to the axis, axial range [lo, hi] */
int mdd_add_cylinder_wall(mdd_ctx* ctx, int axis, double c0, double c1, double radius,
                          double lo, double hi, int inside, int material);
/* type = "sphere", WallSphere (lib.rs:411-428) */
int mdd_add_sphere_wall(mdd_ctx* ctx, const double center[3], double radius, int inside,
                        int material);
/* type = "region", WallRegion (lib.rs:435-447, forces 1170-1231): the
 * region tree's leaf primitives in depth-first order (Union and Intersect
 * both take the child surface nearest the particle, mddem_core/src/
 * region.rs:282-309, so any tree flattens to its leaves), at most 8, each
 * with 12 parameters (unused trailing ones zero):
 *   0 block min[3] max[3]; 1 sphere center[3] radius;
 *   2 cylinder center[2] radius axis lo hi; 3 cone center[2] axis rad_lo
 *   rad_hi lo hi; 4 plane point[3] normal[3] */
int mdd_add_region_wall(mdd_ctx* ctx, int nprim, const int* types, const double* params,
                        int inside, int material);
/* plane wall motion (WallMotion, lib.rs:262-293): 0 static, 1 constant
 * velocity (params: v[3]), 2 oscillate (amplitude, frequency), 3 servo
 * (target_force, max_velocity, gain). The wall restarts from its configured
 * point. Scheduled steps move the walls (wall_move, lib.rs:766-815). */
int mdd_set_wall_motion(mdd_ctx* ctx, int wall, int motion, const double* params);
/* Walls::deactivate_by_name (lib.rs:476-512), by index */
int mdd_set_wall_active(mdd_ctx* ctx, int wall, int active);
/* per wall: force_accumulator of the last force pass (sum of the scalar
 * contact forces, kept in fixed point at 2^-32 N), current point[3] and
 * velocity[3]; any output may be NULL; *count = number of walls */
int mdd_get_walls(mdd_ctx* ctx, double* force, double* point, double* velocity, int* count);
/* the wall systems one at a time (wall_move, lib.rs:766-815;
 * wall_zero_force_accumulators, lib.rs:821-834) */
int mdd_wall_move(mdd_ctx* ctx);
int mdd_wall_zero_force_accumulators(mdd_ctx* ctx);
int mdd_clear_walls(mdd_ctx* ctx);

/* ---- state ------------------------------------------------------------------ */
/* Replaces particle insertion into Atom + DemAtom (crates/dem_atom/src/insert.rs:323-355,
 * crates/md_lattice/src/lib.rs:336-360): per-particle tag, type (material or LJ
 * type), pos/vel/omega (n*3; omega may be NULL), radius (DEM; NULL for LJ),
 * cutoff_radius (neighbor skin radius) and mass. Forces start at zero and the
 * first step integrates with them, as the reference does. */
int mdd_upload_state(mdd_ctx* ctx, int64_t n, const uint32_t* tag, const uint32_t* type,
                     const double* pos, const double* vel, const double* omega,
                     const double* radius, const double* cutoff_radius, const double* mass);
int64_t mdd_nlocal(const mdd_ctx* ctx);

/* ---- batched ensemble ------------------------------------------------------ */
/* Append one independent, fully periodic DEM system (its own box [lo, hi),
 * single material, particles) to the context. All systems then step together:
 * one launch per phase covers every system, particles and bins are
 * offset-indexed by system, and a neighbor rebuild of every system happens
 * when any system's skin is spent. Each system follows the same per-system
 * physics as a context holding it alone (setup: neighbor_setup per system,
 * neighbor.rs:480-541; materials: MaterialTable::build_pair_tables,
 * crates/dem_atom/src/lib.rs:389-515). Tags only need to be unique within a
 * system. Replaces running many reference App instances, one per system. */
int mdd_add_system(mdd_ctx* ctx, const double lo[3], const double hi[3], double youngs,
                   double poisson, double restitution, double friction, int64_t n,
                   const uint32_t* tag, const double* pos, const double* vel,
                   const double* omega, const double* radius, const double* cutoff_radius,
                   const double* mass);
/* number of systems (0: a single system uploaded with mdd_upload_state) */
int mdd_nsystems(mdd_ctx* ctx);
/* system sys occupies particles [first, first + count) of every download */
int mdd_system_range(mdd_ctx* ctx, int sys, int64_t* first, int64_t* count);
/* current particle order (spatially sorted); any pointer may be NULL;
 * quat is n*4 (w,x,y,z) */
int mdd_download_state(mdd_ctx* ctx, uint32_t* tag, double* pos, double* vel, double* omega,
                       double* quat, int32_t* image);
int mdd_download_forces(mdd_ctx* ctx, double* force, double* torque);
/* per-particle constants in the current particle order (Atom::atom_type,
 * Atom::mass, DemAtom::radius, Atom::cutoff_radius; atom.rs:289-305,
 * dem_atom/src/lib.rs:531-557), for restart snapshots; any pointer may be
 * NULL (radius is 0 for LJ) */
int mdd_download_properties(mdd_ctx* ctx, uint32_t* type, double* mass, double* radius,
                            double* cutoff_radius);
int mdd_upload_forces(mdd_ctx* ctx, const double* force, const double* torque);
/* orientation quaternions n*4 (w,x,y,z) in the current particle order (the
 * upload order until the first step), as read_restart restores
 * DemAtom::quaternion (mddem_print/src/lib.rs:1238-1241); uploads start at
 * [1,0,0,0] as insert_single_particle does (dem_atom/src/insert.rs:350) */
int mdd_upload_quaternions(mdd_ctx* ctx, const double* quat);

/* ---- systems (one call = one reference system) ------------------------------ */
/* neighbor_setup + compute_bin_grid (crates/mddem_core/src/neighbor.rs:480-541, 290-371) */
int mdd_neighbor_setup(mdd_ctx* ctx);
/* One full rebuild now: pbc (domain.rs:413-459), sort_atoms_by_bin
 * (neighbor.rs:686-794), borders (comm.rs:449-570) and bin_neighbor_list
 * (neighbor.rs:810-1080). */
int mdd_bin_neighbor_list(mdd_ctx* ctx);
/* Inject a half list (CSR over particle indices, no images), the way the
 * reference unit tests build Neighbor by hand (e.g. contact.rs:1032-1034). */
int mdd_set_half_list(mdd_ctx* ctx, int64_t nlocal, const uint32_t* offsets,
                      const uint32_t* indices);
/* number of half-list pairs (Neighbor::neighbor_indices.len()) */
int64_t mdd_npairs(mdd_ctx* ctx);
/* Half-list pairs as (owner tag, partner tag, image code), code =
 * (sx+1)*9 + (sy+1)*3 + (sz+1) for the partner's periodic image shift;
 * Neighbor::pairs (neighbor.rs:172-227). Writes at most cap, returns the
 * total in *n. */
int mdd_download_pairs(mdd_ctx* ctx, int64_t cap, uint32_t* owner_tag, uint32_t* partner_tag,
                       uint8_t* code, int64_t* n);
/* lj_force (crates/md_lj/src/lib.rs:488-633); overwrites forces */
int mdd_lj_force(mdd_ctx* ctx);
/* hertz_mindlin_contact_force (crates/dem_granular/src/contact.rs:103-615) with the
 * shear history store (tangential.rs:54-133) keyed by canonical pair; overwrites
 * forces and torques */
int mdd_hertz_mindlin_contact_force(mdd_ctx* ctx);
/* wall_contact_force, plane walls (crates/dem_wall/src/lib.rs:848-1002); adds */
int mdd_wall_contact_force(mdd_ctx* ctx);
/* apply_gravity (crates/mddem_fixes/src/lib.rs:756-761); adds */
int mdd_apply_gravity(mdd_ctx* ctx);
/* initial_integration / final_integration (crates/mddem_verlet/src/lib.rs:97-154) */
int mdd_initial_integration(mdd_ctx* ctx);
int mdd_final_integration(mdd_ctx* ctx);
/* initial_rotation / final_rotation (crates/dem_granular/src/rotational.rs:59-100) */
int mdd_initial_rotation(mdd_ctx* ctx);
int mdd_final_rotation(mdd_ctx* ctx);
/* setup_nose_hoover, nh_pre_initial, nh_post_final
 * (crates/md_thermostat/src/lib.rs:195-353) */
int mdd_setup_nose_hoover(mdd_ctx* ctx);
int mdd_nh_pre_initial(mdd_ctx* ctx);
int mdd_nh_post_final(mdd_ctx* ctx);
/* nsteps full steps in the ParticleSimScheduleSet phase order
 * (crates/mddem_core/src/schedule.rs:22-52, schedule.dot); the App run loop */
int mdd_run(mdd_ctx* ctx, int64_t nsteps);

/* ---- shear history (ContactHistoryStore, tangential.rs:54-133) --------------- */
/* (min tag, max tag, spring[3]) in the canonical sign of the lower tag;
 * at most cap entries, total in *n */
int mdd_download_history(mdd_ctx* ctx, int64_t cap, uint32_t* tag_lo, uint32_t* tag_hi,
                         double* spring, int64_t* n);
/* the same with all 7 history values per contact (the [f64; 7] of
 * ContactHistoryStore, contact.rs:343-350): spring (3), rolling displacement
 * (3), twist angle (1), each stored as the reference stores it: sign * value,
 * sign = +1 when the owning particle has the lower tag (contact.rs:337-339);
 * zeros past the spring when no optional model is active */
int mdd_download_history7(mdd_ctx* ctx, int64_t cap, uint32_t* tag_lo, uint32_t* tag_hi,
                          double* hist7, int64_t* n);

/* ---- observables (print_thermo, mddem_print/src/lib.rs:595-700) -------------- */
/* VirialStress of the DEM contact force (mddem_core/src/virial.rs:19-120,
 * registered by HertzMindlinContactPlugin, crates/dem_granular/src/contact.rs:
 * 50-91, accumulated at contact.rs:338-345, 559-569, 962-970): the contact
 * force pass of every step whose number is a multiple of `every` (and every
 * standalone mdd_hertz_mindlin_contact_force call) also sums d (x) F over the
 * pairs, d = x_j - x_i, F the force on i from j; mdd_compute_thermo reports
 * the last such sum in mdd_thermo::virial. 0 (default): not computed.
 * Single systems only. */
int mdd_set_virial(mdd_ctx* ctx, int every);
int mdd_compute_thermo(mdd_ctx* ctx, int with_energy, mdd_thermo* out);
/* mdd_compute_thermo(ctx, 0, out) without any allocation. No counterpart in
 * the reference: for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops). Same kernels and the
 * same summation order, so the same values bit for bit; pe, virial and
 * pressure are left 0 (they need the energy pass). The per-block partial
 * sums go to a static 2 MiB device array (one per GPU, held until the
 * process ends), so at most 8,388,608 particles (MDD_ERR_INVALID beyond).
 * Calls on one GPU share that array: this library serialises them within a
 * process (each call returns only once its totals are on the host). */
int mdd_compute_thermo_noalloc(mdd_ctx* ctx, mdd_thermo* out);
/* block until queued work is done */
int mdd_synchronize(mdd_ctx* ctx);
/* Time every `every`-th scheduled pair-force launch with HIP events on the
 * context's stream (0 = off); resets the statistics. No reference
 * counterpart: measurement hook for the roofline report. */
int mdd_set_profiling(mdd_ctx* ctx, int every);
/* Average device time (ms) of the timed pair-force launches since the last
 * mdd_set_profiling, and the number of pair-force launches. */
int mdd_force_kernel_stats(mdd_ctx* ctx, double* avg_ms, int64_t* launches);
/* Average device time (ms) of the neighbor rebuilds timed since the last
 * mdd_set_profiling (first rebuild launch to the end of its last kernel, on
 * the context's stream), the average idle gap before a rebuild (end of the
 * step chunk that asked for it to its first launch: the host round trip),
 * and how many rebuilds were timed. */
int mdd_rebuild_stats(mdd_ctx* ctx, double* avg_ms, double* avg_gap_ms, int64_t* count);
/* Free and total device memory of a device (hipMemGetInfo), for admission by
 * footprint. No reference counterpart: serving-loop support. */
int mdd_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);
/* DEM force pass at one lane per particle (log2 lanes 0): -1 (default)
 * split by row length (particles with neighbors in the pair kernel, the
 * rest in a walls-and-gravity kernel) when fewer than half the particles
 * had neighbors at the last build the host has seen, else one kernel; 0
 * always one kernel; 1 always split. Identical forces. No reference
 * counterpart: launch-shape tuning. */
int mdd_set_dem_split(mdd_ctx* ctx, int mode);
/* DEM pair pass (single systems, rows of any lane count) in tile order:
 * on (default) the particles are visited band of bin columns by band (16
 * columns along y at each x, x after x), so that the partners gathered on
 * one XCD stay in its L2; off: index order. Identical forces. Turning it
 * on invalidates the list until the next rebuild builds the order. No
 * reference counterpart: launch-order tuning. */
int mdd_set_dem_tiles(mdd_ctx* ctx, int on);
/* DEM pair pass at one lane per particle (plain Hertz-Mindlin, no virial
 * on that step): on (default) each wavefront tests its rows' geometry
 * first and queues the pairs in contact, then runs the contact arithmetic
 * on 64 queued pairs at a time, one per lane; off: each lane runs its own
 * row's contacts. Identical forces, torques and springs. No reference
 * counterpart: launch-shape tuning. */
int mdd_set_dem_wave(mdd_ctx* ctx, int on);
/* Mixed LJ with the Nose-Hoover thermostat (single systems of at most
 * 262,144 particles, rows of 1, 2 or 4 lanes): on (default) each scheduled
 * step's final half (Verlet kick, thermostat rescale, kinetic-energy
 * partials) runs inside the row pass, right after each particle's force;
 * off: in its own kernel. Identical results. No reference counterpart:
 * launch-shape tuning. */
int mdd_set_lj_post_fused(mdd_ctx* ctx, int on);
/* LJ pair pass of MDD_MIXED contexts (rows, single system, every box
 * length at least twice the list cutoff): 1 (default) gathers fixed-point
 * positions (x * 2^32 / L per axis, one 16-byte load per neighbor; 32-bit
 * differences wrap to the minimum image) and computes the pair in FP32,
 * deciding pairs within 1e-6 of the cutoff in FP64 from the same fixed-
 * point differences; 0 gathers the FP64 positions and computes the
 * geometry in FP64. No reference counterpart: precision/speed choice of
 * the mixed path. */
int mdd_set_lj_positions(mdd_ctx* ctx, int mode);
/* Pair-force layout: log2 lanes per particle of the row kernels (0..3;
 * default 2), or -1 for the LJ cell-tile kernel (16-bit rows into each bin
 * cell's candidate list staged in LDS; LJ contexts only, others keep rows).
 * Changes the neighbor-row layout, so the next step rebuilds the list. No
 * reference counterpart: launch-shape tuning. */
int mdd_set_force_lanes(mdd_ctx* ctx, int log2_lanes);
/* Neighbor-row build kernel: 0 (default) picks by grid occupancy, 1 one
 * wavefront per bin cell, 2 one lane per particle or image (for sparse grids
 * with fewer particles than cells; not used for batched ensembles or LJ cell
 * tiles). Both produce identical rows. No reference counterpart: tuning. */
int mdd_set_build_kernel(mdd_ctx* ctx, int mode);
/* Issue each neighbor rebuild (wrap, spatial sort, images, bins, rows) as one
 * HIP graph launch (default on). The rebuild's launches are recorded with
 * their arguments; a graph captured from a launch sequence with identical
 * kernels, grids and argument bytes is replayed, otherwise the sequence is
 * captured anew. Results are identical either way. No reference counterpart:
 * launch-overhead tuning. */
int mdd_set_launch_graphs(mdd_ctx* ctx, int enabled);
/* Rebuilds issued by replaying a captured graph, and graphs captured. */
int mdd_launch_graph_stats(mdd_ctx* ctx, int64_t* replays, int64_t* captures);

#ifdef __cplusplus
}
#endif
#endif
