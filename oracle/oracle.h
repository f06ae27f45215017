/*
 * oracle.h -- CPU restatement of MDDEM's single-process per-step particle path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * path in mddem_amd/: only tests/, the package smoke check and the CPU
 * baseline leg of bench.py load it. The product never links it and never
 * falls back to it.
 *
 * Each function restates one reference system (SueHeir/MDDEM v0.1.3) and
 * cites the file:line it follows. The restatement keeps the reference's
 * operation order, its explicit fused multiply-adds (`mul_add` -> fma) and no
 * other contraction (built with -ffp-contract=off), so that on identical
 * inputs it reproduces the reference arithmetic.
 *
 * Documented departures (each is pinned or marked in the tests):
 *  - sort_atoms_by_bin uses a STABLE sort by bin key; the reference uses
 *    Rust's sort_unstable_by_key (crates/mddem_core/src/neighbor.rs:740),
 *    whose order among equal keys is not specified (parity unpinned there).
 *  - history_mode 1 keys shear history by the canonical pair (min tag,
 *    max tag) instead of per owner by partner tag
 *    (crates/dem_granular/src/contact.rs:358-364); mode 0 is the reference.
 *  - the contact force restates both reference kernels: Hertz-Mindlin
 *    (contact.rs:103-615) and Hooke (contact.rs:622-990), chosen by the
 *    global contact model as the plugin does (contact.rs:65-90).
 *  - the rebuild policies of needs_rebuild (every, check) and the periodic
 *    sort of sort_atoms_by_bin follow neighbor.rs:605-627, 686-794.
 */
#ifndef MDDEM_ORACLE_H
#define MDDEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

enum { ORC_MODEL_LJ = 0, ORC_MODEL_DEM = 1 };
enum { ORC_HIST_REFERENCE = 0, ORC_HIST_CANONICAL = 1 };
enum { ORC_STATE_FULL_REBUILD = 0, ORC_STATE_COMMUNICATE_ONLY = 1 };

orc_sys* orc_create(int model);
void orc_destroy(orc_sys* s);
/* last error message ("" when none); errors are the reference's panics /
 * process::exit paths turned into a code */
const char* orc_error(const orc_sys* s);

/* ---- configuration ---------------------------------------------------- */
void orc_set_box(orc_sys* s, const double lo[3], const double hi[3], const int periodic[3]);
void orc_set_neighbor(orc_sys* s, double skin_fraction, double bin_size, int every,
                      int check, int sort_every, int newton);
void orc_set_dt(orc_sys* s, double dt);
/* materials: arrays of n; builds the mixing tables (MaterialTable::build_pair_tables) */
int orc_set_materials(orc_sys* s, int n, const double* youngs, const double* poisson,
                      const double* restitution, const double* friction);
int orc_get_material_tables(const orc_sys* s, double* beta, double* friction,
                            double* e_eff, double* g_eff);
/* optional per-material properties (MaterialConfig, dem_atom/src/lib.rs:76-118),
 * one array of n per ORC_X_* in this order; zero disables each. Builds their
 * pair tables with the reference's mixing rules (lib.rs:417-511). */
enum {
    ORC_X_ROLLING_FRICTION = 0,
    ORC_X_TWISTING_FRICTION,
    ORC_X_COHESION_ENERGY,
    ORC_X_SURFACE_ENERGY,
    ORC_X_KN,
    ORC_X_KT,
    ORC_X_ROLLING_STIFFNESS,
    ORC_X_ROLLING_DAMPING,
    ORC_X_TWISTING_STIFFNESS,
    ORC_X_TWISTING_DAMPING,
    ORC_NX
};
int orc_set_material_extras(orc_sys* s, int n, const double* values /* ORC_NX * n */);
/* pair table of one ORC_X_* property (nmat * nmat), returns nmat */
int orc_get_extra_table(const orc_sys* s, int which, double* out);
/* contact_model "hooke" (else "hertz"), adhesion_model "dmt" (else "jkr"),
 * rolling_model / twisting_model "sds" (else "constant") */
void orc_set_contact_models(orc_sys* s, int hooke, int dmt, int rolling_sds, int twisting_sds);
void orc_set_lj(orc_sys* s, double epsilon, double sigma, double cutoff);
void orc_get_lj(const orc_sys* s, double out[3]); /* lj1, lj2, cutoff2 */
void orc_set_nose_hoover(orc_sys* s, int enabled, double temperature, double coupling);
void orc_set_integrators(orc_sys* s, int velocity_verlet, int rotation);
void orc_set_gravity(orc_sys* s, int enabled, const double g[3]);
/* walls (dem_wall/src/lib.rs:302-453); each returns the wall index or -1.
 * Plane: normal normalised, bounds may be NULL (unbounded). */
int orc_add_wall(orc_sys* s, const double point[3], const double normal[3], int material,
                 const double bound_lo[3], const double bound_hi[3]);
int orc_add_cylinder_wall(orc_sys* s, int axis, double c0, double c1, double radius, double lo,
                          double hi, int inside, int material);
int orc_add_sphere_wall(orc_sys* s, const double center[3], double radius, int inside,
                        int material);
/* type = "region" (WallRegion, dem_wall/src/lib.rs:435-447): the leaf
 * primitives of the region tree in depth-first order (Union and Intersect
 * both take the child surface nearest the point, region.rs:282-309, so any
 * tree flattens to its leaves), at most 8; 12 parameters each:
 *   0 block:    min[3], max[3]
 *   1 sphere:   center[3], radius
 *   2 cylinder: center[2], radius, axis (0/1/2), lo, hi
 *   3 cone:     center[2], axis, rad_lo, rad_hi, lo, hi
 *   4 plane:    point[3], normal[3] */
int orc_add_region_wall(orc_sys* s, int nprim, const int* types, const double* params,
                        int inside, int material);
/* Region::closest_point_on_surface of a region wall (region.rs:260-772):
 * signed distance (positive outside), point and outward normal */
double orc_region_surface(const orc_sys* s, int wall, const double pos[3], double point[3],
                          double normal[3]);
/* plane motion: 0 static, 1 constant velocity (params v[3]), 2 oscillate
 * (amplitude, frequency), 3 servo (target_force, max_velocity, gain) */
int orc_set_wall_motion(orc_sys* s, int wall, int motion, const double* params);
int orc_set_wall_active(orc_sys* s, int wall, int active);
/* per wall: force accumulator, current point and velocity (3 each); any may be NULL */
int orc_get_walls(const orc_sys* s, double* force_acc, double* point, double* vel);
void orc_wall_move(orc_sys* s);                        /* dem_wall lib.rs:766-815 */
void orc_wall_zero_force_accumulators(orc_sys* s);     /* dem_wall lib.rs:821-834 */
void orc_set_history_mode(orc_sys* s, int mode);
/* 0 (default): sort_atoms_by_bin as the reference runs it. 1: its
 * non-periodic permutation only on a rebuild step. With every > 0 the
 * reference re-evaluates needs_rebuild after decide_rebuild has counted the
 * step, so at the step before a count rebuild it permutes the atoms under
 * the old list (neighbor.rs:639-674, 715-717); 1 leaves that step alone. */
void orc_set_sort_only_on_rebuild(orc_sys* s, int on);

/* ---- state ---------------------------------------------------------------
 * pos/vel/omega: n*3 row-major; omega may be NULL. radius may be NULL for LJ.
 * cutoff_radius is the per-atom neighbor skin radius (DEM: radius; LJ:
 * the lattice `skin`). mass is per atom. Forces start at zero
 * (crates/dem_atom/src/insert.rs:342, crates/md_lattice/src/lib.rs:347). */
int orc_set_atoms(orc_sys* s, int n, const uint32_t* tag, const uint32_t* type,
                  const double* pos, const double* vel, const double* omega,
                  const double* radius, const double* cutoff_radius, const double* mass);
int orc_nlocal(const orc_sys* s);
int orc_nghost(const orc_sys* s);
/* copies locals (in current local order) into caller buffers; any may be NULL */
void orc_get_atoms(const orc_sys* s, uint32_t* tag, double* pos, double* vel, double* force,
                   double* omega, double* torque, double* quat, int32_t* image);
/* overwrite force / torque of locals (n*3); for single-system KATs */
void orc_set_forces(orc_sys* s, const double* force, const double* torque);

/* ---- systems (one call = one reference system) ---------------------------- */
void orc_neighbor_setup(orc_sys* s);                  /* neighbor.rs:480-541 */
int orc_bin_grid(const orc_sys* s, double* out);      /* diagnostics, see .c */
void orc_decide_rebuild(orc_sys* s);                  /* neighbor.rs:639-674 */
void orc_zero_all_forces(orc_sys* s);                 /* atom.rs:529-533 */
void orc_remove_ghost_atoms(orc_sys* s);              /* atom.rs:523-527 */
void orc_pbc(orc_sys* s);                             /* domain.rs:413-459 */
void orc_forward_comm(orc_sys* s);                    /* comm.rs:367-424 */
void orc_sort_atoms_by_bin(orc_sys* s);               /* neighbor.rs:686-794 */
void orc_borders(orc_sys* s);                         /* comm.rs:449-570 */
void orc_bin_neighbor_list(orc_sys* s);               /* neighbor.rs:810-1080 */
int orc_hertz_mindlin_contact_force(orc_sys* s);      /* contact.rs:103-615, 622-990 */
void orc_wall_contact_force(orc_sys* s);              /* dem_wall lib.rs:848-1241 */
void orc_apply_gravity(orc_sys* s);                   /* mddem_fixes lib.rs:756-761 */
void orc_lj_force(orc_sys* s);                        /* md_lj lib.rs:488-633 */
void orc_lj_virial(const orc_sys* s, double w[6]);   /* md_lj lib.rs:510-585 */
/* VirialStress of the contact force (dem_granular/src/contact.rs:338-345,
 * 559-569, 962-970; mddem_core/src/virial.rs:19-120): xx yy zz xy xz yz of
 * the last contact pass, accumulated while on */
void orc_set_virial(orc_sys* s, int on);
void orc_contact_virial(const orc_sys* s, double w[6]);
void orc_lj_tails(const orc_sys* s, double* e_tail, double* p_tail); /* md_lj 397-443 */
double orc_pressure(const orc_sys* s, const double w[6], double p_tail); /* md_measure 605-651 */
void orc_reverse_send_force(orc_sys* s);              /* comm.rs:579-656 */
void orc_initial_integration(orc_sys* s);             /* mddem_verlet lib.rs:97-123 */
void orc_final_integration(orc_sys* s);               /* mddem_verlet lib.rs:136-154 */
void orc_initial_rotation(orc_sys* s);                /* rotational.rs:59-84 */
void orc_final_rotation(orc_sys* s);                  /* rotational.rs:87-100 */
void orc_setup_nose_hoover(orc_sys* s);               /* md_thermostat lib.rs:195-243 */
void orc_nh_pre_initial(orc_sys* s);                  /* md_thermostat lib.rs:254-301 */
void orc_nh_post_final(orc_sys* s);                   /* md_thermostat lib.rs:313-353 */
double orc_compute_ke(const orc_sys* s);              /* atom.rs:489-502 */
/* print_granular_temperature (dem_granular/src/granular_temp.rs:60-99):
 * T_g = sum m |v - vbar|^2 / (3 M), KE, |p|, p_x, p_y, p_z */
void orc_granular_temperature(const orc_sys* s, double out[6]);
/* one full scheduled step in the reference phase order (schedule.dot); 0 ok */
int orc_step(orc_sys* s);
int orc_run(orc_sys* s, int64_t nsteps);

/* ---- inspection ------------------------------------------------------- */
int orc_state(const orc_sys* s);
double orc_p_xi(const orc_sys* s);
int64_t orc_rebuild_count(const orc_sys* s);
int64_t orc_overlap_warnings(const orc_sys* s);
/* inject a half list in CSR over local indices (reference unit-test style,
 * e.g. contact.rs:1032-1034); nlocal+1 offsets */
int orc_set_half_list(orc_sys* s, int nlocal, const uint32_t* offsets, const uint32_t* indices);
int64_t orc_npairs(const orc_sys* s);
/* half-list pairs in CSR order as (owner tag, partner tag, image code)
 * with image code = (sx+1)*9 + (sy+1)*3 + (sz+1), s = partner image shift */
int64_t orc_get_pairs(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                      uint8_t* code);
/* shear history entries: (owner tag, partner tag, s[3]) as stored (canonical
 * sign); returns count. In canonical mode owner tag = min tag. */
int64_t orc_history_count(const orc_sys* s);
int64_t orc_get_history(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                        double* spring);
int orc_set_history(orc_sys* s, int64_t n, const uint32_t* owner_tag,
                    const uint32_t* partner_tag, const double* spring);
/* the same with all 7 history values per entry: tangential spring (3),
 * rolling displacement (3), twist angle (1) */
int64_t orc_get_history7(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                         double* hist7);

/* reference setup helpers (host-side precompute rows of the path) */
double orc_dem_mass(double density, double radius);           /* insert.rs:340 */
double orc_dem_inv_inertia(double mass, double radius);        /* insert.rs:349 */
double orc_rayleigh_dt(double youngs, double poisson, double density, double radius);
                                                               /* insert.rs:1411-1427 */

#ifdef __cplusplus
}
#endif
#endif
