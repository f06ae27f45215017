/*
 * oracle.c -- CPU restatement of MDDEM's single-process per-step particle path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Built with -ffp-contract=off so
 * that every a*b+c is two roundings unless the reference writes mul_add,
 * which maps to fma() here. File:line citations are into SueHeir/MDDEM
 * v0.1.3.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_MAT 64
#define ORC_MAX_WALL 16
#define ORC_MAX_PRIM 8
/* dem_atom/src/lib.rs:63 (overridable only so that the closed-form tests can
 * show they reject a checker built with a wrong damping constant) */
#ifndef SQRT_5_3
#define SQRT_5_3 0.9128709291752768
#endif
/* dem_granular/src/lib.rs:93-108 */
#define TANGENTIAL_EPSILON 1e-30
#define LARGE_OVERLAP_WARN_THRESHOLD 0.90
#define MAX_OVERLAP_WARNINGS 500

/* per-contact history: tangential spring (3), rolling displacement (3) and
 * twist angle (1), stored in the lower tag's frame (tangential.rs:45-53) */
#define HSLOTS 7

typedef struct {
    uint32_t tag;
    double s[HSLOTS];
    int active;
} hist_entry;

typedef struct {
    hist_entry* e;
    int n, cap;
} hist_list;

typedef struct {
    uint64_t key;
    double s[HSLOTS];
    int used, active;
} cslot;

typedef struct {
    cslot* slots;
    uint64_t cap, count;
} cmap;

typedef struct {
    int* send;
    int nsend, capsend;
    int recv_start, recv_count;
    int to_self;     /* SingleProcessComm periodic self-send */
    double pbc_shift[3];
    int dim, sign;
} swap_t;

/* one wall (dem_wall/src/lib.rs:302-453): kind 0 plane, 1 cylinder, 2 sphere */
typedef struct {
    int kind, mat, active;
    double p[3], n[3], blo[3], bhi[3], v[3];
    /* plane motion (WallMotion, lib.rs:262-293): 0 static, 1 constant
     * velocity, 2 oscillate (a0 amplitude, a1 frequency), 3 servo (a0 target
     * force, a1 max velocity, a2 gain) */
    int motion;
    double a0, a1, a2, origin[3];
    /* cylinder: axis, center (c[0], c[1]), radius, axial lo / hi; sphere:
     * center c[0..2], radius */
    int axis, inside;
    double c[3], radius, lo, hi;
    double acc; /* force_accumulator */
    /* kind 3, region surface: leaf primitives of the region tree in
     * depth-first order (unions and intersections both take the child
     * surface nearest the point, region.rs:282-309, so any tree flattens to
     * its leaves); 12 parameters each, see orc_add_region_wall */
    int nprim;
    int ptype[ORC_MAX_PRIM];
    double prm[ORC_MAX_PRIM][12];
} wall_t;

struct orc_sys {
    int model;
    double lo[3], hi[3], size[3];
    int periodic[3];
    double dt;
    /* [neighbor] config (neighbor.rs:45-82) */
    double skin_fraction, bin_min_size;
    int every, check, sort_every, newton;
    /* neighbor state (neighbor.rs:87-159) */
    double max_cutoff, ghost_cutoff, bin_size[3], bin_origin[3];
    int bin_count[3], total_cells;
    int* stencil_fwd;
    int n_fwd, has_self;
    int have_uniform;
    double uniform_cutoff_sq, cached_min_skin;
    double (*last_build_pos)[3];
    int n_last_build; /* -1: cleared (Vec::clear) */
    int steps_since_build, sort_counter, state;
    int setup_done;
    uint32_t* offsets;
    uint32_t* indices;
    int64_t nidx, capidx;
    int cap_offsets;
    uint32_t *atom_cell, *bin_cnt, *bin_start, *sorted_atoms, *atom_sorted_idx;
    double (*sorted_pos)[3];
    int cap_bins, cap_cells;
    /* atoms: locals then ghosts (atom.rs:310-335 + DemAtom lib.rs:530-557) */
    int nlocal, nghost, cap;
    uint32_t *tag, *type;
    int32_t* origin;
    int8_t (*gshift)[3];
    double (*pos)[3], (*vel)[3], (*force)[3];
    double *cutoff_radius, *mass, *inv_mass;
    int32_t (*image)[3];
    double *radius, *inv_inertia;
    double (*quat)[4], (*omega)[3], (*torque)[3];
    hist_list* hist;
    int history_mode;
    /* 1: sort_atoms_by_bin permutes only on a rebuild step or a periodic
     * sort (see orc_set_sort_only_on_rebuild) */
    int sort_only_on_rebuild;
    /* VirialStress (mddem_core/src/virial.rs:19-68) of the last contact
     * force pass, accumulated when active */
    int virial_on;
    double virial[6];
    cmap cm;
    swap_t swaps[6];
    int nswaps;
    /* materials (dem_atom lib.rs:389-515) */
    int nmat;
    double youngs[ORC_MAX_MAT], poisson[ORC_MAX_MAT], restitution[ORC_MAX_MAT],
        friction[ORC_MAX_MAT];
    double* beta_ij;
    double* friction_ij;
    double* e_eff_ij;
    double* g_eff_ij;
    /* optional per-material properties and their pair tables, in the order
     * of ORC_X_* (zero unless set) */
    double extra[ORC_NX][ORC_MAX_MAT];
    double* extra_ij[ORC_NX];
    /* contact_model "hooke", adhesion_model "dmt", rolling_model "sds",
     * twisting_model "sds" (dem_atom/src/lib.rs:142-157) */
    int hooke, dmt, roll_sds, twist_sds;
    int nwalls;
    wall_t walls[ORC_MAX_WALL];
    double wall_time; /* Walls::time */
    int gravity_on;
    double gravity[3];
    /* LJ (md_lj lib.rs:230-239) */
    double lj1, lj2, lj_cut2, lj_eps, lj_sigma;
    /* Nose-Hoover (md_thermostat lib.rs:128-139) */
    int nh_on;
    double nh_T, nh_tau, nh_ndof, nh_Q, p_xi;
    int vv_on, rot_on;
    int64_t overlap_total, rebuilds, step;
    char err[256];
};

/* ---------------------------------------------------------------------- */

static void* xrealloc(void* p, size_t n) {
    void* q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return q;
}

static void set_err(orc_sys* s, const char* msg) {
    snprintf(s->err, sizeof(s->err), "%s", msg);
}

static void ensure_cap(orc_sys* s, int n) {
    if (n <= s->cap) return;
    int c = s->cap ? s->cap : 64;
    while (c < n) c *= 2;
    s->tag = xrealloc(s->tag, sizeof(uint32_t) * c);
    s->type = xrealloc(s->type, sizeof(uint32_t) * c);
    s->origin = xrealloc(s->origin, sizeof(int32_t) * c);
    s->gshift = xrealloc(s->gshift, sizeof(*s->gshift) * c);
    s->pos = xrealloc(s->pos, sizeof(*s->pos) * c);
    s->vel = xrealloc(s->vel, sizeof(*s->vel) * c);
    s->force = xrealloc(s->force, sizeof(*s->force) * c);
    s->cutoff_radius = xrealloc(s->cutoff_radius, sizeof(double) * c);
    s->mass = xrealloc(s->mass, sizeof(double) * c);
    s->inv_mass = xrealloc(s->inv_mass, sizeof(double) * c);
    s->image = xrealloc(s->image, sizeof(*s->image) * c);
    s->radius = xrealloc(s->radius, sizeof(double) * c);
    s->inv_inertia = xrealloc(s->inv_inertia, sizeof(double) * c);
    s->quat = xrealloc(s->quat, sizeof(*s->quat) * c);
    s->omega = xrealloc(s->omega, sizeof(*s->omega) * c);
    s->torque = xrealloc(s->torque, sizeof(*s->torque) * c);
    s->hist = xrealloc(s->hist, sizeof(hist_list) * c);
    for (int i = s->cap; i < c; i++) {
        s->hist[i].e = NULL;
        s->hist[i].n = s->hist[i].cap = 0;
    }
    s->cap = c;
}

static void hist_copy(hist_list* dst, const hist_list* src) {
    if (dst->cap < src->n) {
        dst->e = xrealloc(dst->e, sizeof(hist_entry) * src->n);
        dst->cap = src->n;
    }
    if (src->n) memcpy(dst->e, src->e, sizeof(hist_entry) * src->n);
    dst->n = src->n;
}

/* s holds ns values; the remaining slots start at zero */
static void hist_push(hist_list* h, uint32_t tag, const double* s, int ns, int active) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 4;
        h->e = xrealloc(h->e, sizeof(hist_entry) * h->cap);
    }
    h->e[h->n].tag = tag;
    for (int k = 0; k < HSLOTS; k++) h->e[h->n].s[k] = k < ns ? s[k] : 0.0;
    h->e[h->n].active = active;
    h->n++;
}

/* copy every per-atom field of atom src into slot dst (registry pack/unpack
 * semantics of comm.rs:256-299 without the f64 round trip of integers) */
static void copy_atom(orc_sys* s, int dst, int src) {
    s->tag[dst] = s->tag[src];
    s->type[dst] = s->type[src];
    s->origin[dst] = s->origin[src];
    memcpy(s->gshift[dst], s->gshift[src], sizeof(s->gshift[0]));
    memcpy(s->pos[dst], s->pos[src], sizeof(s->pos[0]));
    memcpy(s->vel[dst], s->vel[src], sizeof(s->vel[0]));
    memcpy(s->force[dst], s->force[src], sizeof(s->force[0]));
    s->cutoff_radius[dst] = s->cutoff_radius[src];
    s->mass[dst] = s->mass[src];
    s->inv_mass[dst] = s->inv_mass[src];
    memcpy(s->image[dst], s->image[src], sizeof(s->image[0]));
    s->radius[dst] = s->radius[src];
    s->inv_inertia[dst] = s->inv_inertia[src];
    memcpy(s->quat[dst], s->quat[src], sizeof(s->quat[0]));
    memcpy(s->omega[dst], s->omega[src], sizeof(s->omega[0]));
    memcpy(s->torque[dst], s->torque[src], sizeof(s->torque[0]));
    hist_copy(&s->hist[dst], &s->hist[src]);
}

/* ---- canonical pair-key history map ----------------------------------- */

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static void cmap_init(cmap* m, uint64_t cap) {
    uint64_t c = 16;
    while (c < cap) c <<= 1;
    m->slots = calloc(c, sizeof(cslot));
    if (!m->slots) abort();
    m->cap = c;
    m->count = 0;
}

static cslot* cmap_find(cmap* m, uint64_t key, int insert) {
    if (insert && 2 * (m->count + 1) > m->cap) {
        cmap old = *m;
        cmap_init(m, old.cap * 2);
        for (uint64_t i = 0; i < old.cap; i++)
            if (old.slots[i].used) {
                cslot* d = cmap_find(m, old.slots[i].key, 1);
                *d = old.slots[i];
            }
        free(old.slots);
    }
    uint64_t h = mix64(key) & (m->cap - 1);
    for (;;) {
        cslot* sl = &m->slots[h];
        if (!sl->used) {
            if (!insert) return NULL;
            sl->used = 1;
            sl->key = key;
            for (int k = 0; k < HSLOTS; k++) sl->s[k] = 0.0;
            sl->active = 0;
            m->count++;
            return sl;
        }
        if (sl->key == key) return sl;
        h = (h + 1) & (m->cap - 1);
    }
}

static void cmap_prune(cmap* m) {
    cmap old = *m;
    cmap_init(m, old.cap);
    for (uint64_t i = 0; i < old.cap; i++)
        if (old.slots[i].used && old.slots[i].active) {
            cslot* d = cmap_find(m, old.slots[i].key, 1);
            *d = old.slots[i];
        }
    free(old.slots);
}

/* ---------------------------------------------------------------------- */

orc_sys* orc_create(int model) {
    orc_sys* s = calloc(1, sizeof(orc_sys));
    if (!s) return NULL;
    s->model = model;
    for (int d = 0; d < 3; d++) {
        s->lo[d] = 0.0;
        s->hi[d] = 1.0;
        s->size[d] = 1.0;
        s->periodic[d] = 1;
    }
    s->dt = 1.0;
    s->skin_fraction = 1.0;
    s->bin_min_size = 1.0;
    s->every = 0;
    s->check = 1;
    s->sort_every = 1000;
    s->newton = 1;
    s->state = ORC_STATE_FULL_REBUILD;
    s->n_last_build = -1;
    s->cached_min_skin = 1.7976931348623157e308;
    s->vv_on = (model == ORC_MODEL_DEM);
    s->rot_on = (model == ORC_MODEL_DEM);
    s->history_mode = ORC_HIST_REFERENCE;
    cmap_init(&s->cm, 1024);
    /* LJForcePlugin default table (md_lj lib.rs:313-315) */
    orc_set_lj(s, 1.0, 1.0, 2.5);
    return s;
}

void orc_destroy(orc_sys* s) {
    if (!s) return;
    for (int i = 0; i < s->cap; i++) free(s->hist[i].e);
    free(s->hist);
    free(s->tag); free(s->type); free(s->origin); free(s->gshift);
    free(s->pos); free(s->vel); free(s->force); free(s->cutoff_radius);
    free(s->mass); free(s->inv_mass); free(s->image); free(s->radius);
    free(s->inv_inertia); free(s->quat); free(s->omega); free(s->torque);
    free(s->stencil_fwd); free(s->last_build_pos); free(s->offsets);
    free(s->indices); free(s->atom_cell); free(s->bin_cnt); free(s->bin_start);
    free(s->sorted_atoms); free(s->atom_sorted_idx); free(s->sorted_pos);
    for (int k = 0; k < 6; k++) free(s->swaps[k].send);
    free(s->beta_ij); free(s->friction_ij); free(s->e_eff_ij); free(s->g_eff_ij);
    for (int x = 0; x < ORC_NX; x++) free(s->extra_ij[x]);
    free(s->cm.slots);
    free(s);
}

const char* orc_error(const orc_sys* s) { return s->err; }

void orc_set_box(orc_sys* s, const double lo[3], const double hi[3], const int periodic[3]) {
    for (int d = 0; d < 3; d++) {
        s->lo[d] = lo[d];
        s->hi[d] = hi[d];
        /* Domain::update_derived (domain.rs) */
        s->size[d] = hi[d] - lo[d];
        s->periodic[d] = periodic[d] != 0;
    }
}

void orc_set_neighbor(orc_sys* s, double skin_fraction, double bin_size, int every, int check,
                      int sort_every, int newton) {
    /* neighbor_read_input (neighbor.rs:449-458) */
    s->skin_fraction = skin_fraction;
    s->bin_min_size = bin_size;
    s->every = every;
    s->check = check;
    s->sort_every = sort_every;
    s->newton = newton;
}

void orc_set_dt(orc_sys* s, double dt) { s->dt = dt; }

static double gmean(double a, double b) { return sqrt(a * b); }
static double gmean0(double a, double b) { return sqrt((a > 0.0 ? a : 0.0) * (b > 0.0 ? b : 0.0)); }
/* harmonic mean when both are set; `either` keeps the one that is set */
static double hmean(double a, double b, int either) {
    if (a > 0.0 && b > 0.0) return 2.0 * a * b / (a + b);
    if (!either) return 0.0;
    return a > 0.0 ? a : b;
}

/* MaterialTable::build_pair_tables, optional fields (dem_atom/src/lib.rs:425-510) */
static void build_extra_tables(orc_sys* s) {
    int n = s->nmat;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const int k = i * n + j;
#define X(w) s->extra[w][i], s->extra[w][j]
            s->extra_ij[ORC_X_ROLLING_FRICTION][k] = gmean(X(ORC_X_ROLLING_FRICTION));
            s->extra_ij[ORC_X_COHESION_ENERGY][k] = gmean(X(ORC_X_COHESION_ENERGY));
            s->extra_ij[ORC_X_SURFACE_ENERGY][k] = gmean(X(ORC_X_SURFACE_ENERGY));
            s->extra_ij[ORC_X_TWISTING_FRICTION][k] = gmean0(X(ORC_X_TWISTING_FRICTION));
            s->extra_ij[ORC_X_KN][k] = hmean(X(ORC_X_KN), 0);
            s->extra_ij[ORC_X_KT][k] = hmean(X(ORC_X_KT), 0);
            s->extra_ij[ORC_X_ROLLING_STIFFNESS][k] = hmean(X(ORC_X_ROLLING_STIFFNESS), 1);
            s->extra_ij[ORC_X_ROLLING_DAMPING][k] = gmean0(X(ORC_X_ROLLING_DAMPING));
            s->extra_ij[ORC_X_TWISTING_STIFFNESS][k] = hmean(X(ORC_X_TWISTING_STIFFNESS), 1);
            s->extra_ij[ORC_X_TWISTING_DAMPING][k] = gmean0(X(ORC_X_TWISTING_DAMPING));
#undef X
        }
}

int orc_set_materials(orc_sys* s, int n, const double* E, const double* nu, const double* e,
                      const double* mu) {
    if (n <= 0 || n > ORC_MAX_MAT) {
        set_err(s, "material count out of range");
        return -1;
    }
    s->nmat = n;
    for (int i = 0; i < n; i++) {
        s->youngs[i] = E[i];
        s->poisson[i] = nu[i];
        s->restitution[i] = e[i];
        s->friction[i] = mu[i];
    }
    free(s->beta_ij); free(s->friction_ij); free(s->e_eff_ij); free(s->g_eff_ij);
    for (int x = 0; x < ORC_NX; x++) {
        free(s->extra_ij[x]);
        s->extra_ij[x] = xrealloc(NULL, sizeof(double) * n * n);
    }
    s->beta_ij = xrealloc(NULL, sizeof(double) * n * n);
    s->friction_ij = xrealloc(NULL, sizeof(double) * n * n);
    s->e_eff_ij = xrealloc(NULL, sizeof(double) * n * n);
    s->g_eff_ij = xrealloc(NULL, sizeof(double) * n * n);
    /* MaterialTable::build_pair_tables (dem_atom/src/lib.rs:389-515) */
    const double PI = 3.141592653589793;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double e_ij = sqrt(e[i] * e[j]);
            double log_e = log(e_ij);
            s->beta_ij[i * n + j] = -log_e / sqrt(PI * PI + log_e * log_e);
            s->friction_ij[i * n + j] = sqrt(mu[i] * mu[j]);
            double nu_i = nu[i], nu_j = nu[j];
            s->e_eff_ij[i * n + j] =
                1.0 / ((1.0 - nu_i * nu_i) / E[i] + (1.0 - nu_j * nu_j) / E[j]);
            s->g_eff_ij[i * n + j] = 1.0 / (2.0 * (2.0 - nu_i) * (1.0 + nu_i) / E[i] +
                                            2.0 * (2.0 - nu_j) * (1.0 + nu_j) / E[j]);
        }
    build_extra_tables(s);
    return 0;
}

int orc_set_material_extras(orc_sys* s, int n, const double* values) {
    if (n != s->nmat) {
        set_err(s, "material extras need one value per material");
        return -1;
    }
    for (int x = 0; x < ORC_NX; x++)
        for (int i = 0; i < n; i++) s->extra[x][i] = values[x * n + i];
    build_extra_tables(s);
    return 0;
}

int orc_get_extra_table(const orc_sys* s, int which, double* out) {
    if (which < 0 || which >= ORC_NX || !s->extra_ij[which]) return -1;
    memcpy(out, s->extra_ij[which], sizeof(double) * s->nmat * s->nmat);
    return s->nmat;
}

void orc_set_contact_models(orc_sys* s, int hooke, int dmt, int rolling_sds, int twisting_sds) {
    s->hooke = hooke != 0;
    s->dmt = dmt != 0;
    s->roll_sds = rolling_sds != 0;
    s->twist_sds = twisting_sds != 0;
}

int orc_get_material_tables(const orc_sys* s, double* beta, double* friction, double* e_eff,
                            double* g_eff) {
    int nn = s->nmat * s->nmat;
    if (beta) memcpy(beta, s->beta_ij, sizeof(double) * nn);
    if (friction) memcpy(friction, s->friction_ij, sizeof(double) * nn);
    if (e_eff) memcpy(e_eff, s->e_eff_ij, sizeof(double) * nn);
    if (g_eff) memcpy(g_eff, s->g_eff_ij, sizeof(double) * nn);
    return s->nmat;
}

void orc_set_lj(orc_sys* s, double epsilon, double sigma, double cutoff) {
    /* LJPairCoeffs::from_params (md_lj/src/lib.rs:230-239); powi(3) is
     * (x*x)*x and powi(2) is x*x */
    double s2 = sigma * sigma;
    double sigma6 = (s2 * s2) * s2;
    s->lj1 = 48.0 * epsilon * sigma6 * sigma6;
    s->lj2 = 24.0 * epsilon * sigma6;
    double c = cutoff * sigma;
    s->lj_cut2 = c * c;
    s->lj_eps = epsilon;
    s->lj_sigma = sigma;
}

void orc_get_lj(const orc_sys* s, double out[3]) {
    out[0] = s->lj1;
    out[1] = s->lj2;
    out[2] = s->lj_cut2;
}

void orc_set_nose_hoover(orc_sys* s, int enabled, double temperature, double coupling) {
    s->nh_on = enabled;
    s->nh_T = temperature;
    s->nh_tau = coupling;
}

void orc_set_integrators(orc_sys* s, int velocity_verlet, int rotation) {
    s->vv_on = velocity_verlet;
    s->rot_on = rotation;
}

void orc_set_gravity(orc_sys* s, int enabled, const double g[3]) {
    s->gravity_on = enabled;
    for (int d = 0; d < 3; d++) s->gravity[d] = g[d];
}

int orc_add_wall(orc_sys* s, const double point[3], const double normal[3], int material,
                 const double blo[3], const double bhi[3]) {
    /* WallPlugin::build, plane walls (dem_wall/src/lib.rs:688-733): the
     * normal is normalised, a zero normal is an error */
    if (s->nwalls >= ORC_MAX_WALL) return -1;
    double mag = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    if (mag <= 1e-15) {
        set_err(s, "wall normal vector must be non-zero");
        return -1;
    }
    wall_t* w = &s->walls[s->nwalls++];
    memset(w, 0, sizeof(*w));
    for (int d = 0; d < 3; d++) {
        w->p[d] = point[d];
        w->origin[d] = point[d];
        w->n[d] = normal[d] / mag;
        w->blo[d] = blo ? blo[d] : -INFINITY;
        w->bhi[d] = bhi ? bhi[d] : INFINITY;
    }
    w->kind = 0;
    w->mat = material;
    w->active = 1;
    return s->nwalls - 1;
}

int orc_add_cylinder_wall(orc_sys* s, int axis, double c0, double c1, double radius, double lo,
                          double hi, int inside, int material) {
    if (s->nwalls >= ORC_MAX_WALL || axis < 0 || axis > 2) return -1;
    wall_t* w = &s->walls[s->nwalls++];
    memset(w, 0, sizeof(*w));
    w->kind = 1;
    w->axis = axis;
    w->c[0] = c0;
    w->c[1] = c1;
    w->radius = radius;
    w->lo = lo;
    w->hi = hi;
    w->inside = inside != 0;
    w->mat = material;
    w->active = 1;
    return s->nwalls - 1;
}

int orc_add_sphere_wall(orc_sys* s, const double center[3], double radius, int inside,
                        int material) {
    if (s->nwalls >= ORC_MAX_WALL) return -1;
    wall_t* w = &s->walls[s->nwalls++];
    memset(w, 0, sizeof(*w));
    w->kind = 2;
    for (int d = 0; d < 3; d++) w->c[d] = center[d];
    w->radius = radius;
    w->inside = inside != 0;
    w->mat = material;
    w->active = 1;
    return s->nwalls - 1;
}

int orc_add_region_wall(orc_sys* s, int nprim, const int* types, const double* params,
                        int inside, int material) {
    if (s->nwalls >= ORC_MAX_WALL || nprim < 1 || nprim > ORC_MAX_PRIM) return -1;
    wall_t* w = &s->walls[s->nwalls++];
    memset(w, 0, sizeof(*w));
    w->kind = 3;
    w->nprim = nprim;
    for (int k = 0; k < nprim; k++) {
        if (types[k] < 0 || types[k] > 4) {
            s->nwalls--;
            return -1;
        }
        w->ptype[k] = types[k];
        memcpy(w->prm[k], params + 12 * k, sizeof(double) * 12);
    }
    w->inside = inside != 0;
    w->mat = material;
    w->active = 1;
    return s->nwalls - 1;
}

/* ---- Region::closest_point_on_surface (mddem_core/src/region.rs:260-772) */
typedef struct {
    double point[3], normal[3], distance;
} surf_t;

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

static void decompose(const double* pos, const double* c, int axis, double* ax, double* d0,
                      double* d1) {
    if (axis == 0) {
        *ax = pos[0]; *d0 = pos[1] - c[0]; *d1 = pos[2] - c[1];
    } else if (axis == 1) {
        *ax = pos[1]; *d0 = pos[0] - c[0]; *d1 = pos[2] - c[1];
    } else {
        *ax = pos[2]; *d0 = pos[0] - c[0]; *d1 = pos[1] - c[1];
    }
}

static void compose(double ax, double r0, double r1, const double* c, int axis, double* out) {
    if (axis == 0) {
        out[0] = ax; out[1] = c[0] + r0; out[2] = c[1] + r1;
    } else if (axis == 1) {
        out[0] = c[0] + r0; out[1] = ax; out[2] = c[1] + r1;
    } else {
        out[0] = c[0] + r0; out[1] = c[1] + r1; out[2] = ax;
    }
}

static void compose_n(double ax, double n0, double n1, int axis, double* out) {
    if (axis == 0) {
        out[0] = ax; out[1] = n0; out[2] = n1;
    } else if (axis == 1) {
        out[0] = n0; out[1] = ax; out[2] = n1;
    } else {
        out[0] = n0; out[1] = n1; out[2] = ax;
    }
}

static double dist3(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* outside result: the candidate point, normal along pos - point */
static surf_t outside_result(const double* pos, const double* point, const double* normal) {
    surf_t r;
    memcpy(r.point, point, sizeof(r.point));
    double dx = pos[0] - point[0], dy = pos[1] - point[1], dz = pos[2] - point[2];
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < 1e-30) {
        memcpy(r.normal, normal, sizeof(r.normal));
        r.distance = 0.0;
    } else {
        r.normal[0] = dx / dist;
        r.normal[1] = dy / dist;
        r.normal[2] = dz / dist;
        r.distance = dist;
    }
    return r;
}

static surf_t closest_block(const double* pos, const double* mn, const double* mx) {
    /* region.rs:314-391 */
    surf_t r;
    double cl[3] = {clampd(pos[0], mn[0], mx[0]), clampd(pos[1], mn[1], mx[1]),
                    clampd(pos[2], mn[2], mx[2])};
    int inside = pos[0] >= mn[0] && pos[0] <= mx[0] && pos[1] >= mn[1] && pos[1] <= mx[1] &&
                 pos[2] >= mn[2] && pos[2] <= mx[2];
    if (!inside) {
        double dx = pos[0] - cl[0], dy = pos[1] - cl[1], dz = pos[2] - cl[2];
        double dist = sqrt(dx * dx + dy * dy + dz * dz);
        memcpy(r.point, cl, sizeof(cl));
        if (dist < 1e-30) {
            r.normal[0] = 0.0; r.normal[1] = 0.0; r.normal[2] = 1.0;
            r.distance = 0.0;
        } else {
            r.normal[0] = dx / dist; r.normal[1] = dy / dist; r.normal[2] = dz / dist;
            r.distance = dist;
        }
        return r;
    }
    double d[6] = {pos[0] - mn[0], mx[0] - pos[0], pos[1] - mn[1],
                   mx[1] - pos[1], pos[2] - mn[2], mx[2] - pos[2]};
    int best = 0;
    for (int k = 1; k < 6; k++)
        if (d[k] < d[best]) best = k;
    memcpy(r.point, pos, sizeof(r.point));
    r.normal[0] = r.normal[1] = r.normal[2] = 0.0;
    int axis = best / 2;
    r.normal[axis] = (best & 1) ? 1.0 : -1.0;
    r.point[axis] = (best & 1) ? mx[axis] : mn[axis];
    r.distance = -d[best];
    return r;
}

static surf_t closest_sphere(const double* pos, const double* c, double radius) {
    /* region.rs:393-415 */
    surf_t r;
    double dx = pos[0] - c[0], dy = pos[1] - c[1], dz = pos[2] - c[2];
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < 1e-30) {
        r.point[0] = c[0] + radius; r.point[1] = c[1]; r.point[2] = c[2];
        r.normal[0] = 1.0; r.normal[1] = 0.0; r.normal[2] = 0.0;
        r.distance = -radius;
        return r;
    }
    double inv = 1.0 / dist;
    double nx = dx * inv, ny = dy * inv, nz = dz * inv;
    r.point[0] = c[0] + nx * radius; r.point[1] = c[1] + ny * radius; r.point[2] = c[2] + nz * radius;
    r.normal[0] = nx; r.normal[1] = ny; r.normal[2] = nz;
    r.distance = dist - radius;
    return r;
}

static surf_t closest_cylinder(const double* pos, const double* c, double radius, int axis,
                               double lo, double hi) {
    /* region.rs:452-588 */
    double ax, d0, d1;
    decompose(pos, c, axis, &ax, &d0, &d1);
    double rd = sqrt(d0 * d0 + d1 * d1);
    double axc = clampd(ax, lo, hi);
    int ax_in = ax >= lo && ax <= hi;
    double cpt[3], cn[3], cdist;
    if (rd < 1e-30) {
        compose(axc, radius, 0.0, c, axis, cpt);
        compose_n(0.0, 1.0, 0.0, axis, cn);
        if (ax_in) {
            cdist = -radius;
        } else {
            double da = ax < lo ? lo - ax : ax - hi;
            cdist = sqrt(radius * radius + da * da);
        }
    } else {
        double inv = 1.0 / rd, n0 = d0 * inv, n1 = d1 * inv;
        compose(axc, n0 * radius, n1 * radius, c, axis, cpt);
        compose_n(0.0, n0, n1, axis, cn);
        if (ax_in) {
            cdist = rd - radius;
        } else {
            double da = ax < lo ? lo - ax : ax - hi;
            double dr = rd - radius;
            cdist = sqrt(dr * dr + da * da) * ((dr < 0.0 && da == 0.0) ? -1.0 : 1.0);
        }
    }
    double cap_r = rd <= radius ? rd : radius;
    double c0 = 0.0, c1 = 0.0;
    if (!(rd < 1e-30)) {
        double inv = 1.0 / rd;
        c0 = d0 * inv * cap_r;
        c1 = d1 * inv * cap_r;
    }
    double lpt[3], ln[3], hpt[3], hn[3];
    compose(lo, c0, c1, c, axis, lpt);
    compose_n(-1.0, 0.0, 0.0, axis, ln);
    compose(hi, c0, c1, c, axis, hpt);
    compose_n(1.0, 0.0, 0.0, axis, hn);
    double dl = dist3(pos, lpt), dh = dist3(pos, hpt);
    double lo_dist = (ax < lo || (ax_in && rd <= radius)) ? (ax <= lo ? lo - ax : dl) : dl;
    double hi_dist = (ax > hi || (ax_in && rd <= radius)) ? (ax >= hi ? ax - hi : dh) : dh;
    surf_t r;
    if (ax_in && rd <= radius) {
        double dc = fabs(radius - rd), dlo = ax - lo, dhi = hi - ax;
        if (dc <= dlo && dc <= dhi) {
            memcpy(r.point, cpt, sizeof(cpt)); memcpy(r.normal, cn, sizeof(cn));
            r.distance = rd - radius;
        } else if (dlo <= dhi) {
            memcpy(r.point, lpt, sizeof(lpt)); memcpy(r.normal, ln, sizeof(ln));
            r.distance = -(ax - lo);
        } else {
            memcpy(r.point, hpt, sizeof(hpt)); memcpy(r.normal, hn, sizeof(hn));
            r.distance = -(hi - ax);
        }
        return r;
    }
    double key[3] = {fabs(cdist), lo_dist, hi_dist};
    int best = 0;
    for (int k = 1; k < 3; k++)
        if (key[k] < key[best]) best = k;
    if (best == 0) return outside_result(pos, cpt, cn);
    if (best == 1) return outside_result(pos, lpt, ln);
    return outside_result(pos, hpt, hn);
}

static surf_t closest_cone(const double* pos, const double* c, int axis, double rlo, double rhi,
                           double lo, double hi) {
    /* region.rs:590-747 */
    double ax, d0, d1;
    decompose(pos, c, axis, &ax, &d0, &d1);
    double rd = sqrt(d0 * d0 + d1 * d1);
    double height = hi - lo;
    double n0 = 1.0, n1 = 0.0;
    if (!(rd < 1e-30)) {
        n0 = d0 / rd;
        n1 = d1 / rd;
    }
    int ax_in = ax >= lo && ax <= hi;
    double r_at;
    if (fabs(height) < 1e-30) {
        r_at = (rlo + rhi) * 0.5;
    } else {
        double t = clampd((ax - lo) / height, 0.0, 1.0);
        r_at = rlo + t * (rhi - rlo);
    }
    int inside = ax_in && rd <= r_at;
    /* lateral surface: nearest point of the segment (lo, rlo)-(hi, rhi) in
     * (axial, radial) coordinates */
    double sda = hi - lo, sdr = rhi - rlo;
    double sl2 = sda * sda + sdr * sdr;
    double t = sl2 < 1e-30 ? 0.5 : clampd(((ax - lo) * sda + (rd - rlo) * sdr) / sl2, 0.0, 1.0);
    double pa = lo + t * sda, pr = rlo + t * sdr;
    double da = ax - pa, dr = rd - pr;
    double d2 = sqrt(da * da + dr * dr);
    double cpt[3], cn[3];
    compose(pa, n0 * pr, n1 * pr, c, axis, cpt);
    double na = -sdr, nr = sda;
    double nl = sqrt(na * na + nr * nr);
    if (nl > 1e-30) {
        na /= nl;
        nr /= nl;
    } else {
        na = 0.0;
        nr = 1.0;
    }
    compose_n(na, nr * n0, nr * n1, axis, cn);
    double csigned = inside ? -d2 : d2;
    double rr = rd > 0.0 ? rd : 0.0;
    double lcr = rlo < rr ? rlo : rr;
    if (rlo < lcr) lcr = rlo;
    double hcr = rhi < rr ? rhi : rr;
    if (rhi < hcr) hcr = rhi;
    double lpt[3], ln[3], hpt[3], hn[3];
    compose(lo, n0 * lcr, n1 * lcr, c, axis, lpt);
    compose_n(-1.0, 0.0, 0.0, axis, ln);
    compose(hi, n0 * hcr, n1 * hcr, c, axis, hpt);
    compose_n(1.0, 0.0, 0.0, axis, hn);
    double dl = dist3(pos, lpt), dh = dist3(pos, hpt);
    surf_t r;
    if (inside) {
        double dlo = ax - lo, dhi = hi - ax;
        if (d2 <= dlo && d2 <= dhi) {
            memcpy(r.point, cpt, sizeof(cpt)); memcpy(r.normal, cn, sizeof(cn));
            r.distance = csigned;
        } else if (dlo <= dhi) {
            memcpy(r.point, lpt, sizeof(lpt)); memcpy(r.normal, ln, sizeof(ln));
            r.distance = -dlo;
        } else {
            memcpy(r.point, hpt, sizeof(hpt)); memcpy(r.normal, hn, sizeof(hn));
            r.distance = -dhi;
        }
        return r;
    }
    double key[3] = {d2, dl, dh};
    int best = 0;
    for (int k = 1; k < 3; k++)
        if (key[k] < key[best]) best = k;
    if (best == 0) return outside_result(pos, cpt, cn);
    if (best == 1) return outside_result(pos, lpt, ln);
    return outside_result(pos, hpt, hn);
}

static surf_t closest_plane(const double* pos, const double* pt, const double* nrm) {
    /* region.rs:749-772 */
    surf_t r;
    double mag = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double nx = 0.0, ny = 0.0, nz = 1.0;
    if (mag > 1e-30) {
        nx = nrm[0] / mag;
        ny = nrm[1] / mag;
        nz = nrm[2] / mag;
    }
    double dist = (pos[0] - pt[0]) * nx + (pos[1] - pt[1]) * ny + (pos[2] - pt[2]) * nz;
    r.point[0] = pos[0] - dist * nx;
    r.point[1] = pos[1] - dist * ny;
    r.point[2] = pos[2] - dist * nz;
    r.normal[0] = nx; r.normal[1] = ny; r.normal[2] = nz;
    r.distance = dist;
    return r;
}

static surf_t closest_prim(int type, const double* q, const double* pos) {
    switch (type) {
        case 0: return closest_block(pos, q, q + 3);
        case 1: return closest_sphere(pos, q, q[3]);
        case 2: return closest_cylinder(pos, q, q[2], (int)q[3], q[4], q[5]);
        case 3: return closest_cone(pos, q, (int)q[2], q[3], q[4], q[5], q[6]);
        default: return closest_plane(pos, q, q + 3);
    }
}

/* a region wall's surface: the leaf with the smallest |distance| (the
 * first of equals), as Union / Intersect choose (region.rs:282-309) */
static surf_t closest_region(const wall_t* w, const double* pos) {
    surf_t best = closest_prim(w->ptype[0], w->prm[0], pos);
    for (int k = 1; k < w->nprim; k++) {
        surf_t r = closest_prim(w->ptype[k], w->prm[k], pos);
        if (fabs(r.distance) < fabs(best.distance)) best = r;
    }
    return best;
}

double orc_region_surface(const orc_sys* s, int wall, const double pos[3], double point[3],
                          double normal[3]) {
    const wall_t* w = &s->walls[wall];
    surf_t r = closest_region(w, pos);
    if (point) memcpy(point, r.point, sizeof(r.point));
    if (normal) memcpy(normal, r.normal, sizeof(r.normal));
    return r.distance;
}

int orc_set_wall_motion(orc_sys* s, int wall, int motion, const double* params) {
    if (wall < 0 || wall >= s->nwalls || s->walls[wall].kind != 0 || motion < 0 || motion > 3)
        return -1;
    wall_t* w = &s->walls[wall];
    w->motion = motion;
    w->v[0] = w->v[1] = w->v[2] = 0.0;
    if (motion == 1)
        for (int d = 0; d < 3; d++) w->v[d] = params[d];
    if (motion >= 2) {
        w->a0 = params[0];
        w->a1 = params[1];
        w->a2 = motion == 3 ? params[2] : 0.0;
    }
    return 0;
}

int orc_set_wall_active(orc_sys* s, int wall, int active) {
    if (wall < 0 || wall >= s->nwalls) return -1;
    s->walls[wall].active = active != 0;
    return 0;
}

int orc_get_walls(const orc_sys* s, double* force_acc, double* point, double* vel) {
    for (int w = 0; w < s->nwalls; w++) {
        if (force_acc) force_acc[w] = s->walls[w].acc;
        for (int d = 0; d < 3; d++) {
            if (point) point[3 * w + d] = s->walls[w].p[d];
            if (vel) vel[3 * w + d] = s->walls[w].v[d];
        }
    }
    return s->nwalls;
}

void orc_wall_move(orc_sys* s) {
    /* wall_move (dem_wall/src/lib.rs:766-815), PreInitialIntegration */
    const double PI = 3.141592653589793;
    double dt = s->dt, time = s->wall_time;
    for (int k = 0; k < s->nwalls; k++) {
        wall_t* w = &s->walls[k];
        if (w->kind != 0 || !w->active) continue;
        if (w->motion == 1) {
            for (int d = 0; d < 3; d++) w->p[d] += w->v[d] * dt;
        } else if (w->motion == 2) {
            double phase = 2.0 * PI * w->a1 * (time + dt);
            double disp = w->a0 * sin(phase);
            for (int d = 0; d < 3; d++) w->p[d] = w->origin[d] + disp * w->n[d];
            double vel_mag = w->a0 * 2.0 * PI * w->a1 * cos(phase);
            for (int d = 0; d < 3; d++) w->v[d] = vel_mag * w->n[d];
        } else if (w->motion == 3) {
            double error = w->a0 - w->acc;
            double vel_mag = w->a2 * error;
            if (vel_mag < -w->a1) vel_mag = -w->a1;
            if (vel_mag > w->a1) vel_mag = w->a1;
            for (int d = 0; d < 3; d++) w->v[d] = vel_mag * w->n[d];
            for (int d = 0; d < 3; d++) w->p[d] += w->v[d] * dt;
        }
    }
    s->wall_time += dt;
}

void orc_wall_zero_force_accumulators(orc_sys* s) {
    /* wall_zero_force_accumulators (lib.rs:821-834), PreForce */
    for (int k = 0; k < s->nwalls; k++) s->walls[k].acc = 0.0;
}

void orc_set_history_mode(orc_sys* s, int mode) { s->history_mode = mode; }
void orc_set_sort_only_on_rebuild(orc_sys* s, int on) { s->sort_only_on_rebuild = on ? 1 : 0; }

double orc_dem_mass(double density, double radius) {
    /* insert_single_particle (dem_atom/src/insert.rs:340) */
    const double PI = 3.141592653589793;
    return density * 4.0 / 3.0 * PI * ((radius * radius) * radius);
}

double orc_dem_inv_inertia(double mass, double radius) {
    /* insert.rs:349 */
    return 1.0 / (0.4 * mass * radius * radius);
}

double orc_rayleigh_dt(double youngs, double poisson, double density, double radius) {
    /* calculate_delta_time (insert.rs:1411-1427), per particle, before x0.15 */
    const double PI = 3.141592653589793;
    double g = youngs / (2.0 * (1.0 + poisson));
    double alpha = 0.1631 * poisson + 0.876605;
    return PI * radius / alpha * sqrt(density / g);
}

int orc_set_atoms(orc_sys* s, int n, const uint32_t* tag, const uint32_t* type,
                  const double* pos, const double* vel, const double* omega,
                  const double* radius, const double* cutoff_radius, const double* mass) {
    ensure_cap(s, n);
    for (int i = 0; i < s->nlocal + s->nghost; i++) s->hist[i].n = 0;
    s->nlocal = n;
    s->nghost = 0;
    for (int i = 0; i < n; i++) {
        s->tag[i] = tag[i];
        s->type[i] = type ? type[i] : 0;
        s->origin[i] = 0;
        s->gshift[i][0] = s->gshift[i][1] = s->gshift[i][2] = 0;
        for (int d = 0; d < 3; d++) {
            s->pos[i][d] = pos[3 * i + d];
            s->vel[i][d] = vel ? vel[3 * i + d] : 0.0;
            s->force[i][d] = 0.0;
            s->image[i][d] = 0;
            s->omega[i][d] = omega ? omega[3 * i + d] : 0.0;
            s->torque[i][d] = 0.0;
        }
        s->cutoff_radius[i] = cutoff_radius[i];
        s->mass[i] = mass[i];
        s->inv_mass[i] = 1.0 / mass[i];
        s->radius[i] = radius ? radius[i] : cutoff_radius[i];
        s->inv_inertia[i] = radius ? orc_dem_inv_inertia(mass[i], radius[i]) : 0.0;
        s->quat[i][0] = 1.0;
        s->quat[i][1] = s->quat[i][2] = s->quat[i][3] = 0.0;
        s->hist[i].n = 0;
    }
    s->state = ORC_STATE_FULL_REBUILD;
    s->n_last_build = -1;
    s->steps_since_build = 0;
    s->sort_counter = 0;
    s->setup_done = 0;
    s->nidx = 0;
    s->step = 0;
    free(s->cm.slots);
    cmap_init(&s->cm, 1024);
    return 0;
}

int orc_nlocal(const orc_sys* s) { return s->nlocal; }
int orc_nghost(const orc_sys* s) { return s->nghost; }

void orc_get_atoms(const orc_sys* s, uint32_t* tag, double* pos, double* vel, double* force,
                   double* omega, double* torque, double* quat, int32_t* image) {
    for (int i = 0; i < s->nlocal; i++) {
        if (tag) tag[i] = s->tag[i];
        for (int d = 0; d < 3; d++) {
            if (pos) pos[3 * i + d] = s->pos[i][d];
            if (vel) vel[3 * i + d] = s->vel[i][d];
            if (force) force[3 * i + d] = s->force[i][d];
            if (omega) omega[3 * i + d] = s->omega[i][d];
            if (torque) torque[3 * i + d] = s->torque[i][d];
            if (image) image[3 * i + d] = s->image[i][d];
        }
        if (quat)
            for (int d = 0; d < 4; d++) quat[4 * i + d] = s->quat[i][d];
    }
}

void orc_set_forces(orc_sys* s, const double* force, const double* torque) {
    for (int i = 0; i < s->nlocal; i++)
        for (int d = 0; d < 3; d++) {
            if (force) s->force[i][d] = force[3 * i + d];
            if (torque) s->torque[i][d] = torque[3 * i + d];
        }
}

/* ---- neighbor setup: compute_bin_grid + neighbor_setup ------------------ */

static void compute_bin_grid(orc_sys* s) {
    /* compute_bin_grid (neighbor.rs:290-371), comm_size = 1 */
    double max_cutoff = s->max_cutoff;
    double required_bin = max_cutoff;
    double min_bin = s->bin_min_size > required_bin ? s->bin_min_size : required_bin;
    int n_int[3];
    for (int d = 0; d < 3; d++) {
        double q = floor(s->size[d] / min_bin);
        if (q < 1.0) q = 1.0;
        n_int[d] = (int)q;
        s->bin_size[d] = s->size[d] / (double)n_int[d];
    }
    double ghost_cut = s->ghost_cutoff > max_cutoff ? s->ghost_cutoff : max_cutoff;
    int sl[3];
    for (int d = 0; d < 3; d++) {
        int range = (int)ceil(max_cutoff / s->bin_size[d]);
        sl[d] = (int)ceil(ghost_cut / s->bin_size[d]) + range;
        s->bin_count[d] = n_int[d] + 2 * sl[d];
        s->bin_origin[d] = s->lo[d] - s->bin_size[d] * (double)sl[d];
    }
    int nx = s->bin_count[0], ny = s->bin_count[1], nz = s->bin_count[2];
    s->total_cells = nx * ny * nz;
    double cutoff_sq = max_cutoff * max_cutoff;
    free(s->stencil_fwd);
    s->stencil_fwd = xrealloc(NULL, sizeof(int) * (2 * sl[0] + 1) * (2 * sl[1] + 1) *
                                        (2 * sl[2] + 1));
    s->n_fwd = 0;
    s->has_self = 0;
    for (int dx = -sl[0]; dx <= sl[0]; dx++)
        for (int dy = -sl[1]; dy <= sl[1]; dy++)
            for (int dz = -sl[2]; dz <= sl[2]; dz++) {
                int ax = abs(dx) - 1, ay = abs(dy) - 1, az = abs(dz) - 1;
                double mdx = (double)(ax > 0 ? ax : 0) * s->bin_size[0];
                double mdy = (double)(ay > 0 ? ay : 0) * s->bin_size[1];
                double mdz = (double)(az > 0 ? az : 0) * s->bin_size[2];
                if (mdx * mdx + mdy * mdy + mdz * mdz < cutoff_sq) {
                    int off = dx * ny * nz + dy * nz + dz;
                    if (off > 0)
                        s->stencil_fwd[s->n_fwd++] = off;
                    else if (off == 0)
                        s->has_self = 1;
                }
            }
}

void orc_neighbor_setup(orc_sys* s) {
    /* neighbor_setup (neighbor.rs:480-541) */
    double max_skin = 0.0, min_skin = 1.7976931348623157e308;
    for (int i = 0; i < s->nlocal; i++) {
        if (s->cutoff_radius[i] > max_skin) max_skin = s->cutoff_radius[i];
        if (s->cutoff_radius[i] < min_skin) min_skin = s->cutoff_radius[i];
    }
    double max_cutoff = max_skin > 0.0 ? 2.0 * max_skin * s->skin_fraction : s->bin_min_size;
    double buffer = min_skin < 1.7976931348623157e308 * 0.5
                        ? (s->skin_fraction - 1.0) * min_skin
                        : 0.0;
    double ghost_cut = max_cutoff + 2.0 * buffer;
    if (ghost_cut < 0.0) ghost_cut = 0.0; /* .max(domain.ghost_cutoff = 0) */
    s->ghost_cutoff = ghost_cut;
    s->max_cutoff = max_cutoff;
    if (s->bin_min_size < max_cutoff) s->bin_min_size = max_cutoff;
    compute_bin_grid(s);
    s->setup_done = 1;
}

int orc_bin_grid(const orc_sys* s, double* out) {
    /* out: max_cutoff, ghost_cutoff, bin_size[3], bin_origin[3], bin_count[3],
     * n_forward, has_self */
    out[0] = s->max_cutoff;
    out[1] = s->ghost_cutoff;
    for (int d = 0; d < 3; d++) {
        out[2 + d] = s->bin_size[d];
        out[5 + d] = s->bin_origin[d];
        out[8 + d] = (double)s->bin_count[d];
    }
    out[11] = (double)s->n_fwd;
    out[12] = (double)s->has_self;
    return s->n_fwd;
}

/* ---- rebuild decision -------------------------------------------------- */

static void save_build_positions(orc_sys* s) {
    /* save_build_positions (neighbor.rs:545-561) */
    int n = s->nlocal;
    s->last_build_pos = xrealloc(s->last_build_pos, sizeof(*s->last_build_pos) * (n ? n : 1));
    memcpy(s->last_build_pos, s->pos, sizeof(*s->pos) * n);
    s->n_last_build = n;
    s->steps_since_build = 0;
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (int i = 0; i < n; i++) {
        double c = s->cutoff_radius[i];
        if (c < mn) mn = c;
        if (c > mx) mx = c;
    }
    s->cached_min_skin = mn;
    if (fabs(mx - mn) < 1e-15) {
        double cutoff = 2.0 * mn * s->skin_fraction;
        s->have_uniform = 1;
        s->uniform_cutoff_sq = cutoff * cutoff;
    } else {
        s->have_uniform = 0;
    }
}

static int displacement_exceeded(const orc_sys* s) {
    /* displacement_exceeded (neighbor.rs:564-601) */
    double per_atom_skin = (s->skin_fraction - 1.0) * s->cached_min_skin;
    double threshold_sq = per_atom_skin * per_atom_skin;
    double bx = s->size[0], by = s->size[1], bz = s->size[2];
    double hbx = bx * 0.5, hby = by * 0.5, hbz = bz * 0.5;
    for (int i = 0; i < s->n_last_build; i++) {
        double dx = s->pos[i][0] - s->last_build_pos[i][0];
        double dy = s->pos[i][1] - s->last_build_pos[i][1];
        double dz = s->pos[i][2] - s->last_build_pos[i][2];
        if (s->periodic[0]) {
            if (dx > hbx) dx -= bx; else if (dx < -hbx) dx += bx;
        }
        if (s->periodic[1]) {
            if (dy > hby) dy -= by; else if (dy < -hby) dy += by;
        }
        if (s->periodic[2]) {
            if (dz > hbz) dz -= bz; else if (dz < -hbz) dz += bz;
        }
        if (dx * dx + dy * dy + dz * dz > threshold_sq) return 1;
    }
    return 0;
}

static int needs_rebuild(const orc_sys* s) {
    /* needs_rebuild (neighbor.rs:605-627) */
    if (s->n_last_build != s->nlocal || s->nlocal == 0) return 1;
    if (s->state == ORC_STATE_FULL_REBUILD) return 1;
    if (s->every == 0) return displacement_exceeded(s);
    if (s->check) return s->steps_since_build >= s->every || displacement_exceeded(s);
    return s->steps_since_build >= s->every;
}

void orc_decide_rebuild(orc_sys* s) {
    /* decide_rebuild (neighbor.rs:639-674) */
    int need = needs_rebuild(s);
    if (!need) {
        for (int i = 0; i < s->nlocal && !need; i++)
            for (int d = 0; d < 3; d++)
                if (s->periodic[d] && (s->pos[i][d] < s->lo[d] || s->pos[i][d] >= s->hi[d])) {
                    need = 1;
                    break;
                }
    }
    if (need)
        s->state = ORC_STATE_FULL_REBUILD;
    else
        s->steps_since_build += 1;
}

/* ---- atoms / domain ---------------------------------------------------- */

void orc_zero_all_forces(orc_sys* s) {
    /* zero_all_forces (atom.rs:529-533) incl. DemAtom #[zero] torque */
    int n = s->nlocal + s->nghost;
    memset(s->force, 0, sizeof(*s->force) * n);
    memset(s->torque, 0, sizeof(*s->torque) * n);
}

void orc_remove_ghost_atoms(orc_sys* s) {
    /* remove_ghost_atoms (atom.rs:523-527) */
    for (int i = s->nlocal; i < s->nlocal + s->nghost; i++) s->hist[i].n = 0;
    s->nghost = 0;
}

static void swap_remove(orc_sys* s, int i) {
    int last = s->nlocal - 1;
    if (i != last) copy_atom(s, i, last);
    s->hist[last].n = 0;
    s->nlocal -= 1;
}

void orc_pbc(orc_sys* s) {
    /* pbc (domain.rs:413-459) with wrap_periodic (396-407) */
    int all = s->periodic[0] && s->periodic[1] && s->periodic[2];
    if (all) {
        for (int i = 0; i < s->nlocal; i++)
            for (int d = 0; d < 3; d++) {
                double high = s->lo[d] + s->size[d];
                if (s->pos[i][d] < s->lo[d]) {
                    s->pos[i][d] += s->size[d];
                    s->image[i][d] += -1;
                } else if (s->pos[i][d] >= high) {
                    s->pos[i][d] -= s->size[d];
                    s->image[i][d] += 1;
                }
            }
        return;
    }
    int removed = 0;
    for (int i = s->nlocal - 1; i >= 0; i--) {
        for (int d = 0; d < 3; d++) {
            if (s->periodic[d]) {
                double high = s->lo[d] + s->size[d];
                if (s->pos[i][d] < s->lo[d]) {
                    s->pos[i][d] += s->size[d];
                    s->image[i][d] += -1;
                } else if (s->pos[i][d] >= high) {
                    s->pos[i][d] -= s->size[d];
                    s->image[i][d] += 1;
                }
            } else if (s->pos[i][d] < s->lo[d] || s->pos[i][d] >= s->hi[d]) {
                swap_remove(s, i);
                removed++;
                break;
            }
        }
    }
    if (removed > 0) s->state = ORC_STATE_FULL_REBUILD;
}

/* ---- ghosts (SingleProcessComm self-send) -------------------------------- */

void orc_borders(orc_sys* s) {
    /* borders (comm.rs:449-570): maxneed = 1 per dim for one process; swap 0
     * sends atoms near the low face with +L, swap 1 atoms near the high face
     * with -L (periodic_swap from comm_setup, comm.rs:221-251). */
    if (s->nghost > 0) orc_remove_ghost_atoms(s);
    s->nswaps = 0;
    int scan_end = s->nlocal;
    double gc = s->ghost_cutoff;
    for (int dim = 0; dim < 3; dim++) {
        int dim_scan_end = scan_end;
        for (int sw = 0; sw < 2; sw++) {
            swap_t* w = &s->swaps[s->nswaps++];
            w->nsend = 0;
            w->dim = dim;
            w->sign = sw == 0 ? 1 : -1;
            w->to_self = s->periodic[dim];
            double offset = s->periodic[dim] ? (sw == 0 ? 1.0 : -1.0) : 0.0;
            w->pbc_shift[0] = w->pbc_shift[1] = w->pbc_shift[2] = 0.0;
            w->pbc_shift[dim] = offset * s->size[dim];
            w->recv_start = s->nlocal + s->nghost;
            w->recv_count = 0;
            for (int i = 0; i < dim_scan_end; i++) {
                double p = s->pos[i][dim];
                double cut = gc > 0.0 ? gc : s->cutoff_radius[i] * 4.0;
                int in_skin = sw == 0 ? (p < s->lo[dim] + cut) : (p >= s->hi[dim] - cut);
                if (!in_skin) continue;
                if (w->nsend == w->capsend) {
                    w->capsend = w->capsend ? 2 * w->capsend : 256;
                    w->send = xrealloc(w->send, sizeof(int) * w->capsend);
                }
                w->send[w->nsend++] = i;
                if (!w->to_self) continue; /* non-periodic dim: nobody receives */
                int g = s->nlocal + s->nghost;
                ensure_cap(s, g + 1);
                copy_atom(s, g, i);
                s->origin[g] = i;
                /* pack_atom_inner: pos + change_pos on every component */
                for (int d = 0; d < 3; d++) s->pos[g][d] = s->pos[i][d] + w->pbc_shift[d];
                s->gshift[g][dim] = (int8_t)(s->gshift[i][dim] + (sw == 0 ? 1 : -1));
                s->inv_mass[g] = 1.0 / s->mass[g];
                s->nghost++;
                w->recv_count++;
            }
        }
        scan_end = s->nlocal + s->nghost;
    }
}

void orc_forward_comm(orc_sys* s) {
    /* forward_comm (comm.rs:367-424) with compute_per_atom_offset (318-353) */
    for (int k = 0; k < s->nswaps; k++) {
        swap_t* w = &s->swaps[k];
        if (!w->to_self) continue;
        for (int m = 0; m < w->nsend; m++) {
            int idx = w->send[m];
            double offset[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < 3; d++) {
                if (w->pbc_shift[d] != 0.0) {
                    double sign = w->pbc_shift[d] > 0.0 ? 1.0 : -1.0;
                    double ghost_pos = s->pos[idx][d] + sign * s->size[d];
                    if (ghost_pos > s->lo[d] && ghost_pos < s->hi[d])
                        offset[d] = -sign * s->size[d];
                    else
                        offset[d] = sign * s->size[d];
                }
            }
            int g = w->recv_start + m;
            for (int d = 0; d < 3; d++) {
                s->pos[g][d] = s->pos[idx][d] + offset[d];
                s->vel[g][d] = s->vel[idx][d];
                s->omega[g][d] = s->omega[idx][d]; /* DemAtom #[forward] omega */
            }
        }
    }
}

void orc_reverse_send_force(orc_sys* s) {
    /* reverse_send_force (comm.rs:579-656), self-send branch */
    for (int k = s->nswaps - 1; k >= 0; k--) {
        swap_t* w = &s->swaps[k];
        if (!w->to_self) continue;
        for (int m = 0; m < w->recv_count; m++) {
            int g = w->recv_start + m;
            int o = s->origin[g];
            for (int d = 0; d < 3; d++) {
                s->force[o][d] += s->force[g][d];
                s->torque[o][d] += s->torque[g][d]; /* DemAtom #[reverse] torque */
            }
        }
    }
}

/* ---- spatial sort -------------------------------------------------------- */

static uint32_t cell_of(const orc_sys* s, const double p[3], double ibx, double iby,
                        double ibz) {
    int nx = s->bin_count[0], ny = s->bin_count[1], nz = s->bin_count[2];
    int cx = (int)floor((p[0] - s->bin_origin[0]) * ibx);
    int cy = (int)floor((p[1] - s->bin_origin[1]) * iby);
    int cz = (int)floor((p[2] - s->bin_origin[2]) * ibz);
    cx = cx < 0 ? 0 : (cx > nx - 1 ? nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy > ny - 1 ? ny - 1 : cy);
    cz = cz < 0 ? 0 : (cz > nz - 1 ? nz - 1 : cz);
    return (uint32_t)(cx * ny * nz + cy * nz + cz);
}

typedef struct {
    uint32_t key;
    int idx;
} keyidx;

static int cmp_keyidx(const void* a, const void* b) {
    const keyidx* x = a;
    const keyidx* y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

void orc_sort_atoms_by_bin(orc_sys* s) {
    /* sort_atoms_by_bin (neighbor.rs:686-794); stable order among equal keys */
    s->sort_counter += 1;
    int periodic_sort = s->sort_every > 0 && s->sort_counter >= s->sort_every;
    if (periodic_sort) s->sort_counter = 0;
    int n = s->nlocal;
    if (n == 0 || s->total_cells == 0) {
        if (periodic_sort) { /* all_reduce(0) = 0: nothing */ }
        return;
    }
    int rebuild_needed = needs_rebuild(s);
    /* With every > 0, decide_rebuild has just counted this step, so
     * needs_rebuild can turn true here while the step keeps its list: the
     * reference then permutes the atoms under a list that still holds the
     * old indices (neighbor.rs:715-717). sort_only_on_rebuild skips that
     * permutation. */
    if (s->sort_only_on_rebuild && s->state != ORC_STATE_FULL_REBUILD) rebuild_needed = 0;
    if (!periodic_sort && !rebuild_needed) return;
    double ibx = 1.0 / s->bin_size[0], iby = 1.0 / s->bin_size[1], ibz = 1.0 / s->bin_size[2];
    keyidx* ki = xrealloc(NULL, sizeof(keyidx) * n);
    for (int i = 0; i < n; i++) {
        ki[i].key = cell_of(s, s->pos[i], ibx, iby, ibz);
        ki[i].idx = i;
    }
    qsort(ki, n, sizeof(keyidx), cmp_keyidx);
    int already = 1;
    for (int i = 0; i < n; i++)
        if (ki[i].idx != i) {
            already = 0;
            break;
        }
    if (!already) {
        /* Atom::apply_permutation + registry.apply_permutation_all (atom.rs:383-394):
         * gather through scratch copies placed after the ghosts */
        int base = s->nlocal + s->nghost;
        ensure_cap(s, base + n);
        for (int i = 0; i < n; i++) copy_atom(s, base + i, ki[i].idx);
        for (int i = 0; i < n; i++) copy_atom(s, i, base + i);
        for (int i = 0; i < n; i++) s->hist[base + i].n = 0;
        if (s->n_last_build >= n) {
            double (*old)[3] = xrealloc(NULL, sizeof(*old) * n);
            memcpy(old, s->last_build_pos, sizeof(*old) * n);
            for (int i = 0; i < n; i++) memcpy(s->last_build_pos[i], old[ki[i].idx], sizeof(*old));
            free(old);
        }
        if (s->nghost > 0) {
            int* inv = xrealloc(NULL, sizeof(int) * n);
            for (int i = 0; i < n; i++) inv[ki[i].idx] = i;
            for (int g = n; g < n + s->nghost; g++)
                if (s->origin[g] < n) s->origin[g] = inv[s->origin[g]];
            /* sendlists hold pre-sort indices; the reference forces a full
             * borders rebuild on a periodic sort (neighbor.rs:781-793) */
            free(inv);
        }
    }
    free(ki);
    if (periodic_sort && !already) {
        s->n_last_build = -1;
        s->state = ORC_STATE_FULL_REBUILD;
    }
}

/* ---- half neighbor list ---------------------------------------------------- */

#define PUSH_INDEX(val)                                                          \
    do {                                                                         \
        if (nidx >= s->capidx) {                                                 \
            s->capidx = s->capidx ? s->capidx * 2 : 1024;                        \
            s->indices = xrealloc(s->indices, sizeof(uint32_t) * s->capidx);     \
        }                                                                        \
        s->indices[nidx++] = (val);                                              \
    } while (0)

void orc_bin_neighbor_list(orc_sys* s) {
    /* bin_neighbor_list (neighbor.rs:810-1080), newton = true */
    int nlocal = s->nlocal;
    int total = s->nlocal + s->nghost;
    s->state = ORC_STATE_COMMUNICATE_ONLY;
    save_build_positions(s);
    s->rebuilds++;
    int nx = s->bin_count[0], ny = s->bin_count[1], nz = s->bin_count[2];
    int cells = s->total_cells;
    double ibx = 1.0 / s->bin_size[0], iby = 1.0 / s->bin_size[1], ibz = 1.0 / s->bin_size[2];
    (void)nx; (void)ny; (void)nz;
    if (s->cap_bins < total) {
        s->cap_bins = total;
        s->atom_cell = xrealloc(s->atom_cell, sizeof(uint32_t) * total);
        s->sorted_atoms = xrealloc(s->sorted_atoms, sizeof(uint32_t) * total);
        s->atom_sorted_idx = xrealloc(s->atom_sorted_idx, sizeof(uint32_t) * total);
        s->sorted_pos = xrealloc(s->sorted_pos, sizeof(*s->sorted_pos) * total);
    }
    if (s->cap_cells < cells + 1) {
        s->cap_cells = cells + 1;
        s->bin_cnt = xrealloc(s->bin_cnt, sizeof(uint32_t) * (cells + 1));
        s->bin_start = xrealloc(s->bin_start, sizeof(uint32_t) * (cells + 1));
    }
    memset(s->bin_cnt, 0, sizeof(uint32_t) * cells);
    for (int i = 0; i < total; i++) {
        uint32_t c = cell_of(s, s->pos[i], ibx, iby, ibz);
        s->atom_cell[i] = c;
        s->bin_cnt[c] += 1;
    }
    s->bin_start[0] = 0;
    for (int c = 0; c < cells; c++) s->bin_start[c + 1] = s->bin_start[c] + s->bin_cnt[c];
    memcpy(s->bin_cnt, s->bin_start, sizeof(uint32_t) * cells);
    for (int i = 0; i < total; i++) {
        uint32_t c = s->atom_cell[i];
        uint32_t p = s->bin_cnt[c];
        s->sorted_atoms[p] = (uint32_t)i;
        s->atom_sorted_idx[i] = p;
        s->bin_cnt[c] = p + 1;
    }
    for (int m = 0; m < total; m++) memcpy(s->sorted_pos[m], s->pos[s->sorted_atoms[m]], sizeof(double) * 3);

    if (s->cap_offsets < nlocal + 1) {
        s->cap_offsets = nlocal + 1;
        s->offsets = xrealloc(s->offsets, sizeof(uint32_t) * (nlocal + 1));
    }
    int64_t nidx = 0;
    if (s->have_uniform) {
        double cutoff_sq = s->uniform_cutoff_sq;
        for (int i = 0; i < nlocal; i++) {
            s->offsets[i] = (uint32_t)nidx;
            uint32_t my_cell = s->atom_cell[i];
            const double* pi = s->pos[i];
            if (s->has_self) {
                uint32_t end = s->bin_start[my_cell + 1];
                for (uint32_t m = s->atom_sorted_idx[i] + 1; m < end; m++) {
                    const double* pj = s->sorted_pos[m];
                    double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
                    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    if (r2 < cutoff_sq) PUSH_INDEX(s->sorted_atoms[m]);
                }
            }
            for (int k = 0; k < s->n_fwd; k++) {
                int c = (int)my_cell + s->stencil_fwd[k];
                for (uint32_t m = s->bin_start[c]; m < s->bin_start[c + 1]; m++) {
                    const double* pj = s->sorted_pos[m];
                    double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
                    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    if (r2 < cutoff_sq) PUSH_INDEX(s->sorted_atoms[m]);
                }
            }
        }
    } else {
        double sf2 = s->skin_fraction * s->skin_fraction;
        for (int i = 0; i < nlocal; i++) {
            s->offsets[i] = (uint32_t)nidx;
            uint32_t my_cell = s->atom_cell[i];
            const double* pi = s->pos[i];
            double si = s->cutoff_radius[i];
            if (s->has_self) {
                uint32_t end = s->bin_start[my_cell + 1];
                for (uint32_t m = s->atom_sorted_idx[i] + 1; m < end; m++) {
                    uint32_t j = s->sorted_atoms[m];
                    const double* pj = s->sorted_pos[m];
                    double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
                    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    double sum = si + s->cutoff_radius[j];
                    if (r2 < sum * sum * sf2) PUSH_INDEX(j);
                }
            }
            for (int k = 0; k < s->n_fwd; k++) {
                int c = (int)my_cell + s->stencil_fwd[k];
                for (uint32_t m = s->bin_start[c]; m < s->bin_start[c + 1]; m++) {
                    uint32_t j = s->sorted_atoms[m];
                    const double* pj = s->sorted_pos[m];
                    double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
                    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    double sum = si + s->cutoff_radius[j];
                    if (r2 < sum * sum * sf2) PUSH_INDEX(j);
                }
            }
        }
    }
    s->offsets[nlocal] = (uint32_t)nidx;
    s->nidx = nidx;
}

int orc_set_half_list(orc_sys* s, int nlocal, const uint32_t* offsets, const uint32_t* indices) {
    if (nlocal != s->nlocal) {
        set_err(s, "half list row count differs from nlocal");
        return -1;
    }
    if (s->cap_offsets < nlocal + 1) {
        s->cap_offsets = nlocal + 1;
        s->offsets = xrealloc(s->offsets, sizeof(uint32_t) * (nlocal + 1));
    }
    memcpy(s->offsets, offsets, sizeof(uint32_t) * (nlocal + 1));
    int64_t n = offsets[nlocal];
    if (s->capidx < n) {
        s->capidx = n;
        s->indices = xrealloc(s->indices, sizeof(uint32_t) * (n ? n : 1));
    }
    if (n) memcpy(s->indices, indices, sizeof(uint32_t) * n);
    s->nidx = n;
    return 0;
}

int64_t orc_npairs(const orc_sys* s) { return s->nidx; }

int64_t orc_get_pairs(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                      uint8_t* code) {
    int64_t k = 0;
    if (!s->offsets) return 0;  /* no list built yet */
    for (int i = 0; i < s->nlocal; i++)
        for (uint32_t m = s->offsets[i]; m < s->offsets[i + 1]; m++, k++) {
            uint32_t j = s->indices[m];
            if (owner_tag) owner_tag[k] = s->tag[i];
            if (partner_tag) partner_tag[k] = s->tag[j];
            if (code)
                code[k] = (uint8_t)((s->gshift[j][0] + 1) * 9 + (s->gshift[j][1] + 1) * 3 +
                                    (s->gshift[j][2] + 1));
        }
    return k;
}

/* ---- forces ---------------------------------------------------------------- */

static double fmax0(double a) { return a > 0.0 ? a : 0.0; }

/* rolling resistance and twisting friction of one contact (contact.rs:427-557,
 * the same in the Hooke kernel 839-950); torques in the pair's frame are
 * added to i and subtracted from j; rd / tw are the rolling displacement and
 * twist angle in the local (i, j) frame, updated in place */
static void roll_twist(orc_sys* s, int i, int j, int k, double nx, double ny, double nz,
                       double f_n_mag, double r_eff, double* rd, double* tw) {
    const double dt = s->dt;
    double oix = s->omega[i][0], oiy = s->omega[i][1], oiz = s->omega[i][2];
    double ojx = s->omega[j][0], ojy = s->omega[j][1], ojz = s->omega[j][2];
    double orx = oix - ojx, ory = oiy - ojy, orz = oiz - ojz;
    double or_dot_n = orx * nx + ory * ny + orz * nz;
    double rollx = orx - or_dot_n * nx, rolly = ory - or_dot_n * ny, rollz = orz - or_dot_n * nz;
    double mu_r = s->extra_ij[ORC_X_ROLLING_FRICTION][k];
    double mu_tw = s->extra_ij[ORC_X_TWISTING_FRICTION][k];
    if (mu_r > 0.0) {
        double roll_mag = sqrt(rollx * rollx + rolly * rolly + rollz * rollz);
        double trx = 0.0, try_ = 0.0, trz = 0.0;
        int apply = 0;
        if (s->roll_sds) {
            double k_roll = s->extra_ij[ORC_X_ROLLING_STIFFNESS][k];
            double g_roll = s->extra_ij[ORC_X_ROLLING_DAMPING][k];
            double rd_dot_n = rd[0] * nx + rd[1] * ny + rd[2] * nz;
            rd[0] -= rd_dot_n * nx;
            rd[1] -= rd_dot_n * ny;
            rd[2] -= rd_dot_n * nz;
            rd[0] += rollx * dt;
            rd[1] += rolly * dt;
            rd[2] += rollz * dt;
            trx = -k_roll * rd[0] - g_roll * rollx;
            try_ = -k_roll * rd[1] - g_roll * rolly;
            trz = -k_roll * rd[2] - g_roll * rollz;
            double tr_mag = sqrt(trx * trx + try_ * try_ + trz * trz);
            double tau_max = mu_r * fabs(f_n_mag) * r_eff;
            if (tr_mag > tau_max && tr_mag > TANGENTIAL_EPSILON) {
                double sc = tau_max / tr_mag;
                trx *= sc;
                try_ *= sc;
                trz *= sc;
                if (k_roll > TANGENTIAL_EPSILON) {
                    rd[0] = (trx + g_roll * rollx) / (-k_roll);
                    rd[1] = (try_ + g_roll * rolly) / (-k_roll);
                    rd[2] = (trz + g_roll * rollz) / (-k_roll);
                }
            }
            apply = 1;
        } else if (roll_mag > 1e-30) {
            double tau_mag = mu_r * fabs(f_n_mag) * r_eff;
            double inv_roll = tau_mag / roll_mag;
            trx = -inv_roll * rollx;
            try_ = -inv_roll * rolly;
            trz = -inv_roll * rollz;
            apply = 1;
        }
        if (apply) {
            s->torque[i][0] += trx;
            s->torque[i][1] += try_;
            s->torque[i][2] += trz;
            s->torque[j][0] -= trx;
            s->torque[j][1] -= try_;
            s->torque[j][2] -= trz;
        }
    }
    if (mu_tw > 0.0) {
        double twist_vel = or_dot_n;
        double ttx = 0.0, tty = 0.0, ttz = 0.0;
        int apply = 0;
        if (s->twist_sds) {
            double k_tw = s->extra_ij[ORC_X_TWISTING_STIFFNESS][k];
            double g_tw = s->extra_ij[ORC_X_TWISTING_DAMPING][k];
            *tw += twist_vel * dt;
            double tau_twist = -k_tw * *tw - g_tw * twist_vel;
            double tau_max = mu_tw * fabs(f_n_mag) * r_eff;
            if (fabs(tau_twist) > tau_max) {
                /* f64::signum: +1 for +0.0, -1 for -0.0 */
                tau_twist = (signbit(tau_twist) ? -1.0 : 1.0) * tau_max;
                if (k_tw > TANGENTIAL_EPSILON) *tw = (tau_twist + g_tw * twist_vel) / (-k_tw);
            }
            ttx = tau_twist * nx;
            tty = tau_twist * ny;
            ttz = tau_twist * nz;
            apply = 1;
        } else if (fabs(twist_vel) > 1e-30) {
            double tau = mu_tw * fabs(f_n_mag) * r_eff;
            double sign_tw = twist_vel > 0.0 ? -1.0 : 1.0;
            ttx = sign_tw * tau * nx;
            tty = sign_tw * tau * ny;
            ttz = sign_tw * tau * nz;
            apply = 1;
        }
        if (apply) {
            s->torque[i][0] += ttx;
            s->torque[i][1] += tty;
            s->torque[i][2] += ttz;
            s->torque[j][0] -= ttx;
            s->torque[j][1] -= tty;
            s->torque[j][2] -= ttz;
        }
    }
}

/* VirialStress::add_pair (mddem_core/src/virial.rs:59-67): d = pos[j] -
 * pos[i], f = force on i from j */
static void virial_add_pair(double w[6], double dx, double dy, double dz, double fx, double fy,
                            double fz) {
    w[0] += dx * fx;
    w[1] += dy * fy;
    w[2] += dz * fz;
    w[3] += dx * fy;
    w[4] += dx * fz;
    w[5] += dy * fz;
}

void orc_set_virial(orc_sys* s, int on) { s->virial_on = on ? 1 : 0; }
void orc_contact_virial(const orc_sys* s, double w[6]) { memcpy(w, s->virial, sizeof(s->virial)); }

int orc_hertz_mindlin_contact_force(orc_sys* s) {
    /* hertz_mindlin_contact_force (dem_granular/src/contact.rs:103-615) or,
     * with contact_model "hooke", hooke_contact_force (contact.rs:622-990);
     * newton = true, no bonds or rigid bodies */
    const double PI = 3.141592653589793;
    int nlocal = s->nlocal;
    int nm = s->nmat;
    double dt = s->dt;
    int warnings = 0;
    const int hooke = s->hooke;
    if (s->history_mode == ORC_HIST_REFERENCE) {
        for (int i = 0; i < nlocal; i++)
            for (int k = 0; k < s->hist[i].n; k++) s->hist[i].e[k].active = 0;
    } else {
        for (uint64_t k = 0; k < s->cm.cap; k++) s->cm.slots[k].active = 0;
    }
    /* zero_virial_stress (virial.rs:112-120) when the virial is active */
    if (s->virial_on) memset(s->virial, 0, sizeof(s->virial));
    for (int i = 0; i < nlocal; i++) {
        for (uint32_t m = s->offsets[i]; m < s->offsets[i + 1]; m++) {
            int j = (int)s->indices[m];
            double r1 = s->radius[i], r2 = s->radius[j];
            double dx = s->pos[j][0] - s->pos[i][0];
            double dy = s->pos[j][1] - s->pos[i][1];
            double dz = s->pos[j][2] - s->pos[i][2];
            double dist_sq = dx * dx + dy * dy + dz * dz;
            double sum_r = r1 + r2;
            int mi = (int)s->type[i], mj = (int)s->type[j];
            const int kij = mi * nm + mj;
            double r_eff = (r1 * r2) / sum_r;
            double e_eff = s->e_eff_ij[kij];
            double surface_energy = hooke ? 0.0 : s->extra_ij[ORC_X_SURFACE_ENERGY][kij];
            int use_dmt = s->dmt;
            /* JKR pull-off distance (contact.rs:163-171) */
            double delta_pulloff = 0.0;
            if (surface_energy > 0.0 && !use_dmt)
                delta_pulloff = cbrt(PI * PI * surface_energy * surface_energy * r_eff /
                                     (4.0 * e_eff * e_eff));
            double interaction_r = sum_r + delta_pulloff;
            if (dist_sq >= interaction_r * interaction_r) continue;
            double distance = sqrt(dist_sq);
            if (distance == 0.0) continue;
            double r_min = r1 < r2 ? r1 : r2;
            double delta = sum_r - distance;
            if (0.5 * r_min < delta) delta = 0.5 * r_min;
            if (hooke && delta <= 0.0) continue;
            if (delta > 0.0 && distance / sum_r < LARGE_OVERLAP_WARN_THRESHOLD) {
                warnings++;
                s->overlap_total++;
                if (warnings > MAX_OVERLAP_WARNINGS) {
                    set_err(s, "Over 500 excessive overlaps this step -- aborting");
                    return -2;
                }
            }
            if (delta <= 0.0 && surface_energy <= 0.0) continue;
            double inv_dist = 1.0 / distance;
            double nx = dx * inv_dist, ny = dy * inv_dist, nz = dz * inv_dist;
            double g_eff = s->g_eff_ij[kij];
            double inv_m_i = s->inv_mass[i] > 0.0 ? s->inv_mass[i] : 1.0 / s->mass[i];
            double inv_m_j = s->inv_mass[j] > 0.0 ? s->inv_mass[j] : 1.0 / s->mass[j];
            double m_r = 1.0 / (inv_m_i + inv_m_j);
            double beta = s->beta_ij[kij];
            double mu = s->friction_ij[kij];
            double cohesion = s->extra_ij[ORC_X_COHESION_ENERGY][kij];
            int jkr_adhesion_only = surface_energy > 0.0 && !use_dmt && delta <= 0.0;
            double s_n = 0.0, k_n = 0.0, k_t = 0.0, gamma_n = 0.0;
            if (hooke) {
                k_n = s->extra_ij[ORC_X_KN][kij];
                k_t = s->extra_ij[ORC_X_KT][kij];
                gamma_n = 2.0 * beta * sqrt(k_n * m_r);
            } else if (delta > 0.0) {
                double sdr = sqrt(delta * r_eff);
                s_n = 2.0 * e_eff * sdr;
                k_n = 4.0 / 3.0 * e_eff * sdr;
                k_t = 8.0 * g_eff * sdr;
            }
            double oix = s->omega[i][0], oiy = s->omega[i][1], oiz = s->omega[i][2];
            double ojx = s->omega[j][0], ojy = s->omega[j][1], ojz = s->omega[j][2];
            double r1nx = r1 * nx, r1ny = r1 * ny, r1nz = r1 * nz;
            double vcix = s->vel[i][0] + (oiy * r1nz - oiz * r1ny);
            double vciy = s->vel[i][1] + (oiz * r1nx - oix * r1nz);
            double vciz = s->vel[i][2] + (oix * r1ny - oiy * r1nx);
            double r2nx = r2 * nx, r2ny = r2 * ny, r2nz = r2 * nz;
            double vcjx = s->vel[j][0] + (-ojy * r2nz + ojz * r2ny);
            double vcjy = s->vel[j][1] + (-ojz * r2nx + ojx * r2nz);
            double vcjz = s->vel[j][2] + (-ojx * r2ny + ojy * r2nx);
            double vrx = vcjx - vcix, vry = vcjy - vciy, vrz = vcjz - vciz;
            double v_n = vrx * nx + vry * ny + vrz * nz;
            double f_n_mag;
            if (hooke) {
                /* contact.rs:742-748 */
                if (cohesion > 0.0) {
                    double f_coh = cohesion * PI * delta * r_eff;
                    f_n_mag = k_n * delta - gamma_n * v_n - f_coh;
                } else {
                    f_n_mag = fmax0(k_n * delta - gamma_n * v_n);
                }
            } else if (surface_energy > 0.0 && use_dmt) {
                /* DMT (contact.rs:281-285) */
                double f_dmt = 2.0 * PI * surface_energy * r_eff;
                double f_diss_n = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                f_n_mag = k_n * delta - f_diss_n - f_dmt;
            } else if (surface_energy > 0.0) {
                /* JKR (contact.rs:286-296) */
                double f_adh = 1.5 * PI * surface_energy * r_eff;
                if (jkr_adhesion_only) {
                    f_n_mag = -f_adh;
                } else {
                    double f_diss_n = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                    f_n_mag = k_n * delta - f_diss_n - f_adh;
                }
            } else if (cohesion > 0.0) {
                /* SJKR (contact.rs:297-302) */
                double f_diss_n = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                double f_coh = cohesion * PI * delta * r_eff;
                f_n_mag = k_n * delta - f_diss_n - f_coh;
            } else {
                double f_diss_n = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                f_n_mag = fmax0(k_n * delta - f_diss_n);
            }
            double fnx = f_n_mag * nx, fny = f_n_mag * ny, fnz = f_n_mag * nz;
            s->force[i][0] -= fnx;
            s->force[i][1] -= fny;
            s->force[i][2] -= fnz;
            s->force[j][0] += fnx;
            s->force[j][1] += fny;
            s->force[j][2] += fnz;
            /* JKR adhesion-only: normal force only, no history touched
             * (contact.rs:320-331: the entry is not marked active); the
             * virial takes the normal force alone (contact.rs:338-345) */
            if (jkr_adhesion_only) {
                if (s->virial_on) virial_add_pair(s->virial, dx, dy, dz, -fnx, -fny, -fnz);
                continue;
            }
            double vtx = vrx - v_n * nx, vty = vry - v_n * ny, vtz = vrz - v_n * nz;
            uint32_t tag_i = s->tag[i], tag_j = s->tag[j];
            double sign = tag_i < tag_j ? 1.0 : -1.0;
            double stored[HSLOTS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            hist_entry* he = NULL;
            cslot* cs = NULL;
            if (s->history_mode == ORC_HIST_REFERENCE) {
                hist_list* hl = &s->hist[i];
                for (int k = 0; k < hl->n; k++)
                    if (hl->e[k].tag == tag_j) {
                        he = &hl->e[k];
                        break;
                    }
                if (he) memcpy(stored, he->s, sizeof(stored));
            } else {
                uint32_t a = tag_i < tag_j ? tag_i : tag_j, b = tag_i < tag_j ? tag_j : tag_i;
                cs = cmap_find(&s->cm, ((uint64_t)a << 32) | b, 1);
                memcpy(stored, cs->s, sizeof(stored));
            }
            double sx = sign * stored[0], sy = sign * stored[1], sz = sign * stored[2];
            double s_dot_n = sx * nx + sy * ny + sz * nz;
            sx -= s_dot_n * nx;
            sy -= s_dot_n * ny;
            sz -= s_dot_n * nz;
            sx += vtx * dt;
            sy += vty * dt;
            sz += vtz * dt;
            double s_mag = sqrt(sx * sx + sy * sy + sz * sz);
            double f_t_spring_mag = k_t * s_mag;
            double f_t_max = mu * fabs(f_n_mag);
            if (f_t_spring_mag > f_t_max && f_t_spring_mag > TANGENTIAL_EPSILON) {
                double sc = f_t_max / f_t_spring_mag;
                sx *= sc;
                sy *= sc;
                sz *= sc;
            }
            double gamma_t = 2.0 * SQRT_5_3 * beta * sqrt(k_t * m_r);
            double ftx = k_t * sx - gamma_t * vtx;
            double fty = k_t * sy - gamma_t * vty;
            double ftz = k_t * sz - gamma_t * vtz;
            double f_t_mag = sqrt(ftx * ftx + fty * fty + ftz * ftz);
            if (f_t_mag > f_t_max && f_t_mag > TANGENTIAL_EPSILON) {
                double sc = f_t_max / f_t_mag;
                ftx *= sc;
                fty *= sc;
                ftz *= sc;
            }
            double tix = r1ny * ftz - r1nz * fty;
            double tiy = r1nz * ftx - r1nx * ftz;
            double tiz = r1nx * fty - r1ny * ftx;
            double tjx = r2ny * ftz - r2nz * fty;
            double tjy = r2nz * ftx - r2nx * ftz;
            double tjz = r2nx * fty - r2ny * ftx;
            s->force[i][0] += ftx;
            s->force[i][1] += fty;
            s->force[i][2] += ftz;
            s->force[j][0] -= ftx;
            s->force[j][1] -= fty;
            s->force[j][2] -= ftz;
            s->torque[i][0] += tix;
            s->torque[i][1] += tiy;
            s->torque[i][2] += tiz;
            s->torque[j][0] += tjx;
            s->torque[j][1] += tjy;
            s->torque[j][2] += tjz;
            double rd[3] = {sign * stored[3], sign * stored[4], sign * stored[5]};
            double tw = sign * stored[6];
            roll_twist(s, i, j, kij, nx, ny, nz, f_n_mag, r_eff, rd, &tw);
            /* virial: force on i from j = -F_n + F_t, newton (contact.rs:
             * 559-569; Hooke 962-970) */
            if (s->virial_on)
                virial_add_pair(s->virial, dx, dy, dz, -fnx + ftx, -fny + fty, -fnz + ftz);
            double ns[HSLOTS] = {sign * sx,    sign * sy,    sign * sz, sign * rd[0],
                                 sign * rd[1], sign * rd[2], sign * tw};
            if (s->history_mode == ORC_HIST_REFERENCE) {
                if (he) {
                    memcpy(he->s, ns, sizeof(ns));
                    he->active = 1;
                } else {
                    hist_push(&s->hist[i], tag_j, ns, HSLOTS, 1);
                }
            } else {
                memcpy(cs->s, ns, sizeof(ns));
                cs->active = 1;
            }
        }
    }
    if (s->history_mode == ORC_HIST_REFERENCE) {
        for (int i = 0; i < nlocal; i++) {
            hist_list* hl = &s->hist[i];
            int w = 0;
            for (int k = 0; k < hl->n; k++)
                if (hl->e[k].active) hl->e[w++] = hl->e[k];
            hl->n = w;
        }
    } else {
        cmap_prune(&s->cm);
    }
    return 0;
}

/* Hertz normal force of a curved wall with optional SJKR cohesion
 * (cylinder and sphere walls, dem_wall/src/lib.rs:1080-1095, 1140-1155) */
static double curved_wall_force(const orc_sys* s, int i, int mat, double delta, double nx,
                                double ny, double nz) {
    const double PI = 3.141592653589793;
    int nm = s->nmat, k = (int)s->type[i] * nm + mat;
    double radius = s->radius[i];
    double r_eff = radius;
    double e_eff = s->e_eff_ij[k];
    double sdr = sqrt(delta * r_eff);
    double k_n = 4.0 / 3.0 * e_eff * sdr;
    double s_n = 2.0 * e_eff * sdr;
    double m_r = s->mass[i];
    double v_n = s->vel[i][0] * nx + s->vel[i][1] * ny + s->vel[i][2] * nz;
    double beta = s->beta_ij[k];
    double cohesion = s->extra_ij[ORC_X_COHESION_ENERGY][k];
    double f_diss = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
    if (cohesion > 0.0) return k_n * delta - f_diss - cohesion * PI * delta * r_eff;
    return fmax0(k_n * delta - f_diss);
}

void orc_wall_contact_force(orc_sys* s) {
    /* wall_contact_force (dem_wall/src/lib.rs:848-1241): plane walls with
     * JKR / DMT / SJKR and twisting friction, then cylinder and sphere walls */
    const double PI = 3.141592653589793;
    int nm = s->nmat;
    for (int w = 0; w < s->nwalls; w++) {
        wall_t* W = &s->walls[w];
        if (W->kind != 0 || !W->active) continue;
        double acc = 0.0;
        for (int i = 0; i < s->nlocal; i++) {
            double px = s->pos[i][0], py = s->pos[i][1], pz = s->pos[i][2];
            if (!(px >= W->blo[0] && px <= W->bhi[0] && py >= W->blo[1] && py <= W->bhi[1] &&
                  pz >= W->blo[2] && pz <= W->bhi[2]))
                continue;
            double dx = px - W->p[0], dy = py - W->p[1], dz = pz - W->p[2];
            double distance = dx * W->n[0] + dy * W->n[1] + dz * W->n[2];
            if (distance <= 0.0) continue;
            double radius = s->radius[i];
            int k = (int)s->type[i] * nm + W->mat;
            double r_eff = radius;
            double e_eff = s->e_eff_ij[k];
            double surface = s->extra_ij[ORC_X_SURFACE_ENERGY][k];
            int use_dmt = s->dmt;
            double d_po = 0.0;
            if (surface > 0.0 && !use_dmt)
                d_po = cbrt(PI * PI * surface * surface * r_eff / (4.0 * e_eff * e_eff));
            double delta = radius - distance;
            if (0.5 * radius < delta) delta = 0.5 * radius;
            if (delta <= 0.0 && surface <= 0.0) continue;
            if (delta < -d_po) continue;
            int adhesion_only = surface > 0.0 && !use_dmt && delta <= 0.0;
            double s_n = 0.0, k_n = 0.0;
            if (delta > 0.0) {
                double sdr = sqrt(delta * r_eff);
                s_n = 2.0 * e_eff * sdr;
                k_n = 4.0 / 3.0 * e_eff * sdr;
            }
            double m_r = s->mass[i];
            double vrx = s->vel[i][0] - W->v[0], vry = s->vel[i][1] - W->v[1],
                   vrz = s->vel[i][2] - W->v[2];
            double v_n = vrx * W->n[0] + vry * W->n[1] + vrz * W->n[2];
            double beta = s->beta_ij[k];
            double cohesion = s->extra_ij[ORC_X_COHESION_ENERGY][k];
            double f_net;
            if (surface > 0.0 && use_dmt) {
                double f_diss = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                f_net = k_n * delta - f_diss - 2.0 * PI * surface * r_eff;
            } else if (surface > 0.0) {
                double f_adh = 1.5 * PI * surface * r_eff;
                if (adhesion_only) {
                    f_net = -f_adh;
                } else {
                    double f_diss = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                    f_net = k_n * delta - f_diss - f_adh;
                }
            } else if (cohesion > 0.0) {
                double f_diss = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                f_net = k_n * delta - f_diss - cohesion * PI * delta * r_eff;
            } else {
                double f_diss = 2.0 * beta * SQRT_5_3 * sqrt(s_n * m_r) * v_n;
                f_net = fmax0(k_n * delta - f_diss);
            }
            s->force[i][0] += f_net * W->n[0];
            s->force[i][1] += f_net * W->n[1];
            s->force[i][2] += f_net * W->n[2];
            if (delta > 0.0) {
                double mu_tw = s->extra_ij[ORC_X_TWISTING_FRICTION][k];
                if (mu_tw > 0.0) {
                    double twist = s->omega[i][0] * W->n[0] + s->omega[i][1] * W->n[1] +
                                   s->omega[i][2] * W->n[2];
                    if (fabs(twist) > 1e-30) {
                        double tau = mu_tw * fabs(f_net) * r_eff;
                        double sign_tw = twist > 0.0 ? -1.0 : 1.0;
                        s->torque[i][0] += sign_tw * tau * W->n[0];
                        s->torque[i][1] += sign_tw * tau * W->n[1];
                        s->torque[i][2] += sign_tw * tau * W->n[2];
                    }
                }
            }
            acc += f_net;
        }
        W->acc += acc;
    }
    for (int w = 0; w < s->nwalls; w++) {
        wall_t* W = &s->walls[w];
        if (W->kind != 1 || !W->active) continue;
        double acc = 0.0;
        for (int i = 0; i < s->nlocal; i++) {
            const double* pos = s->pos[i];
            double radius = s->radius[i];
            double axial, d0, d1;
            if (W->axis == 0) {
                axial = pos[0];
                d0 = pos[1] - W->c[0];
                d1 = pos[2] - W->c[1];
            } else if (W->axis == 1) {
                axial = pos[1];
                d0 = pos[0] - W->c[0];
                d1 = pos[2] - W->c[1];
            } else {
                axial = pos[2];
                d0 = pos[0] - W->c[0];
                d1 = pos[1] - W->c[1];
            }
            if (axial < W->lo || axial > W->hi) continue;
            double radial = sqrt(d0 * d0 + d1 * d1);
            if (radial < 1e-30) continue;
            double inv_r = 1.0 / radial;
            double delta, n0, n1;
            if (W->inside) {
                delta = radius - (W->radius - radial);
                n0 = -d0 * inv_r;
                n1 = -d1 * inv_r;
            } else {
                delta = radius - (radial - W->radius);
                n0 = d0 * inv_r;
                n1 = d1 * inv_r;
            }
            double nx, ny, nz;
            if (W->axis == 0) {
                nx = 0.0; ny = n0; nz = n1;
            } else if (W->axis == 1) {
                nx = n0; ny = 0.0; nz = n1;
            } else {
                nx = n0; ny = n1; nz = 0.0;
            }
            if (delta <= 0.0) continue;
            if (0.5 * radius < delta) delta = 0.5 * radius;
            double f_net = curved_wall_force(s, i, W->mat, delta, nx, ny, nz);
            s->force[i][0] += f_net * nx;
            s->force[i][1] += f_net * ny;
            s->force[i][2] += f_net * nz;
            acc += f_net;
        }
        W->acc += acc;
    }
    for (int w = 0; w < s->nwalls; w++) {
        wall_t* W = &s->walls[w];
        if (W->kind != 2 || !W->active) continue;
        double acc = 0.0;
        for (int i = 0; i < s->nlocal; i++) {
            double dx = s->pos[i][0] - W->c[0], dy = s->pos[i][1] - W->c[1],
                   dz = s->pos[i][2] - W->c[2];
            double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < 1e-30) continue;
            double radius = s->radius[i];
            double inv = 1.0 / dist;
            double nx, ny, nz, delta;
            if (W->inside) {
                delta = radius - (W->radius - dist);
                nx = -dx * inv; ny = -dy * inv; nz = -dz * inv;
            } else {
                delta = radius - (dist - W->radius);
                nx = dx * inv; ny = dy * inv; nz = dz * inv;
            }
            if (delta <= 0.0) continue;
            if (0.5 * radius < delta) delta = 0.5 * radius;
            double f_net = curved_wall_force(s, i, W->mat, delta, nx, ny, nz);
            s->force[i][0] += f_net * nx;
            s->force[i][1] += f_net * ny;
            s->force[i][2] += f_net * nz;
            acc += f_net;
        }
        W->acc += acc;
    }
    /* region-surface walls (dem_wall/src/lib.rs:1170-1231) */
    for (int w = 0; w < s->nwalls; w++) {
        wall_t* W = &s->walls[w];
        if (W->kind != 3 || !W->active) continue;
        double acc = 0.0;
        for (int i = 0; i < s->nlocal; i++) {
            double radius = s->radius[i];
            surf_t sr = closest_region(W, s->pos[i]);
            double gap = W->inside ? -sr.distance : sr.distance;
            double delta = radius - gap;
            if (delta <= 0.0) continue;
            if (0.5 * radius < delta) delta = 0.5 * radius;
            double sg = W->inside ? -1.0 : 1.0;
            double nx = sg * sr.normal[0], ny = sg * sr.normal[1], nz = sg * sr.normal[2];
            double f_net = curved_wall_force(s, i, W->mat, delta, nx, ny, nz);
            s->force[i][0] += f_net * nx;
            s->force[i][1] += f_net * ny;
            s->force[i][2] += f_net * nz;
            acc += f_net;
        }
        W->acc += acc;
    }
}

void orc_apply_gravity(orc_sys* s) {
    /* apply_gravity (mddem_fixes/src/lib.rs:756-761) */
    if (!s->gravity_on) return;
    for (int i = 0; i < s->nlocal; i++) {
        s->force[i][0] += s->mass[i] * s->gravity[0];
        s->force[i][1] += s->mass[i] * s->gravity[1];
        s->force[i][2] += s->mass[i] * s->gravity[2];
    }
}

void orc_lj_force(orc_sys* s) {
    /* lj_force (md_lj/src/lib.rs:488-633), single type, virial inactive */
    double lj1 = s->lj1, lj2 = s->lj2, cut2 = s->lj_cut2;
    for (int i = 0; i < s->nlocal; i++) {
        double pix = s->pos[i][0], piy = s->pos[i][1], piz = s->pos[i][2];
        double fi0 = s->force[i][0], fi1 = s->force[i][1], fi2 = s->force[i][2];
        for (uint32_t m = s->offsets[i]; m < s->offsets[i + 1]; m++) {
            int j = (int)s->indices[m];
            double dx = s->pos[j][0] - pix, dy = s->pos[j][1] - piy, dz = s->pos[j][2] - piz;
            double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 >= cut2) continue;
            double r2inv = 1.0 / r2;
            double r6inv = r2inv * r2inv * r2inv;
            double fpair = r2inv * r6inv * fma(lj1, r6inv, -lj2);
            fi0 = fma(-fpair, dx, fi0);
            fi1 = fma(-fpair, dy, fi1);
            fi2 = fma(-fpair, dz, fi2);
            s->force[j][0] = fma(fpair, dx, s->force[j][0]);
            s->force[j][1] = fma(fpair, dy, s->force[j][1]);
            s->force[j][2] = fma(fpair, dz, s->force[j][2]);
        }
        s->force[i][0] = fi0;
        s->force[i][1] = fi1;
        s->force[i][2] = fi2;
    }
}

void orc_lj_virial(const orc_sys* s, double w[6]) {
    /* virial accumulation of lj_force (md_lj/src/lib.rs:510-585) over the
     * newton half list: W_ab = sum d_a F_b, d = x_j - x_i, F = -fpair d;
     * order xx, yy, zz, xy, xz, yz */
    double lj1 = s->lj1, lj2 = s->lj2, cut2 = s->lj_cut2;
    for (int k = 0; k < 6; k++) w[k] = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        double pix = s->pos[i][0], piy = s->pos[i][1], piz = s->pos[i][2];
        for (uint32_t m = s->offsets[i]; m < s->offsets[i + 1]; m++) {
            int j = (int)s->indices[m];
            double dx = s->pos[j][0] - pix, dy = s->pos[j][1] - piy, dz = s->pos[j][2] - piz;
            double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 >= cut2) continue;
            double r2inv = 1.0 / r2;
            double r6inv = r2inv * r2inv * r2inv;
            double fpair = r2inv * r6inv * fma(lj1, r6inv, -lj2);
            double fx = -fpair * dx, fy = -fpair * dy, fz = -fpair * dz;
            w[0] += dx * fx;
            w[1] += dy * fy;
            w[2] += dz * fz;
            w[3] += dx * fy;
            w[4] += dx * fz;
            w[5] += dy * fz;
        }
    }
}

void orc_lj_tails(const orc_sys* s, double* e_tail, double* p_tail) {
    /* setup_lj_tails (md_lj/src/lib.rs:397-443), one type (weight 1) */
    const double pi = 3.141592653589793;
    double n = (double)s->nlocal;
    double v = s->size[0] * s->size[1] * s->size[2]; /* Domain::volume (domain.rs:153) */
    double rho = n / v;
    double rc = sqrt(s->lj_cut2);
    double sigma3 = (s->lj_sigma * s->lj_sigma) * s->lj_sigma;
    double rc3 = (rc * rc) * rc;
    double rc9 = (rc3 * rc3) * rc3;
    double sigma6 = sigma3 * sigma3;
    double sigma9 = sigma6 * sigma3;
    double w = 1.0 * 1.0;
    *e_tail = 0.0 + w * (8.0 / 3.0) * pi * n * rho * s->lj_eps * sigma3 *
                        (sigma9 / (3.0 * rc9) - sigma3 / rc3);
    *p_tail = 0.0 + w * (16.0 / 3.0) * pi * rho * rho * s->lj_eps * sigma3 *
                        (2.0 * sigma9 / (3.0 * rc9) - sigma3 / rc3);
}

double orc_pressure(const orc_sys* s, const double w[6], double p_tail) {
    /* compute_pressure (md_measure/src/lib.rs:605-651):
     * P = rho T - tr(W) / (3 V) + P_tail, T = 2 KE / (3 N - 3) */
    double n = (double)s->nlocal;
    double v = s->size[0] * s->size[1] * s->size[2];
    double rho = n / v;
    double ke = 0.0;
    for (int i = 0; i < s->nlocal; i++)
        ke += s->mass[i] * (s->vel[i][0] * s->vel[i][0] + s->vel[i][1] * s->vel[i][1] +
                            s->vel[i][2] * s->vel[i][2]);
    ke *= 0.5;
    double ndof = 3.0 * n - 3.0;
    double temp = ndof > 0.0 ? 2.0 * ke / ndof : 0.0;
    double trace = w[0] + w[1] + w[2];
    return rho * temp - trace / (3.0 * v) + p_tail;
}

/* ---- integrators ------------------------------------------------------------ */

void orc_initial_integration(orc_sys* s) {
    /* initial_integration (mddem_verlet/src/lib.rs:97-123) */
    double dt = s->dt;
    for (int i = 0; i < s->nlocal; i++) {
        double h = 0.5 * dt * s->inv_mass[i];
        for (int d = 0; d < 3; d++) s->vel[i][d] += h * s->force[i][d];
        for (int d = 0; d < 3; d++) s->pos[i][d] += s->vel[i][d] * dt;
    }
}

void orc_final_integration(orc_sys* s) {
    /* final_integration (mddem_verlet/src/lib.rs:136-154) */
    double dt = s->dt;
    for (int i = 0; i < s->nlocal; i++) {
        double h = 0.5 * dt * s->inv_mass[i];
        for (int d = 0; d < 3; d++) s->vel[i][d] += h * s->force[i][d];
    }
}

void orc_initial_rotation(orc_sys* s) {
    /* initial_rotation (dem_granular/src/rotational.rs:59-84, helpers 26-41) */
    double dt = s->dt;
    for (int i = 0; i < s->nlocal; i++) {
        double ii = s->inv_inertia[i];
        if (ii == 0.0) continue;
        s->omega[i][0] += 0.5 * dt * s->torque[i][0] * ii;
        s->omega[i][1] += 0.5 * dt * s->torque[i][1] * ii;
        s->omega[i][2] += 0.5 * dt * s->torque[i][2] * ii;
        double ox = s->omega[i][0], oy = s->omega[i][1], oz = s->omega[i][2];
        double omega_mag = sqrt(ox * ox + oy * oy + oz * oz);
        double angle = omega_mag * dt;
        if (angle > 1e-14) {
            double inv = 1.0 / omega_mag;
            double ax = ox * inv, ay = oy * inv, az = oz * inv;
            double half = angle * 0.5;
            double sn = sin(half);
            double a[4] = {cos(half), ax * sn, ay * sn, az * sn};
            double* b = s->quat[i];
            double q0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
            double q1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
            double q2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
            double q3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
            b[0] = q0;
            b[1] = q1;
            b[2] = q2;
            b[3] = q3;
        }
    }
}

void orc_final_rotation(orc_sys* s) {
    /* final_rotation (rotational.rs:87-100) */
    double dt = s->dt;
    for (int i = 0; i < s->nlocal; i++) {
        double ii = s->inv_inertia[i];
        if (ii == 0.0) continue;
        s->omega[i][0] += 0.5 * dt * s->torque[i][0] * ii;
        s->omega[i][1] += 0.5 * dt * s->torque[i][1] * ii;
        s->omega[i][2] += 0.5 * dt * s->torque[i][2] * ii;
    }
}

double orc_compute_ke(const orc_sys* s) {
    /* compute_ke (mddem_core/src/atom.rs:489-502), no group mask */
    double ke = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        const double* v = s->vel[i];
        ke += s->mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    return 0.5 * ke;
}

void orc_granular_temperature(const orc_sys* s, double out[6]) {
    /* print_granular_temperature (dem_granular/src/granular_temp.rs:60-99):
     * out = T_g, KE, |p|, p_x, p_y, p_z */
    double mvx = 0.0, mvy = 0.0, mvz = 0.0, msum = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        mvx += s->mass[i] * s->vel[i][0];
        mvy += s->mass[i] * s->vel[i][1];
        mvz += s->mass[i] * s->vel[i][2];
        msum += s->mass[i];
    }
    const double ax = mvx / msum, ay = mvy / msum, az = mvz / msum;
    double vd = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        const double dx = s->vel[i][0] - ax, dy = s->vel[i][1] - ay, dz = s->vel[i][2] - az;
        /* powi(2) is one multiplication */
        vd += s->mass[i] * (dx * dx) + s->mass[i] * (dy * dy) + s->mass[i] * (dz * dz);
    }
    double ke = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        const double* v = s->vel[i];
        ke += 0.5 * s->mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    out[0] = vd / (3.0 * msum);
    out[1] = ke;
    out[2] = sqrt(mvx * mvx + mvy * mvy + mvz * mvz);
    out[3] = mvx;
    out[4] = mvy;
    out[5] = mvz;
}

void orc_setup_nose_hoover(orc_sys* s) {
    /* setup_nose_hoover (md_thermostat/src/lib.rs:195-243), no group */
    double n = (double)s->nlocal;
    double ndof = 3.0 * n - 3.0;
    s->nh_ndof = ndof;
    s->nh_Q = ndof * s->nh_T * s->nh_tau * s->nh_tau;
    s->p_xi = 0.0;
}

void orc_nh_pre_initial(orc_sys* s) {
    /* nh_pre_initial (md_thermostat/src/lib.rs:254-301) */
    double dt = s->dt;
    double ke = orc_compute_ke(s);
    s->p_xi += (dt / 2.0) * (2.0 * ke - s->nh_ndof * s->nh_T);
    double scale = exp(-dt / 2.0 * s->p_xi / s->nh_Q);
    for (int i = 0; i < s->nlocal; i++) {
        double* v = s->vel[i];
        v[0] *= scale;
        v[1] *= scale;
        v[2] *= scale;
        double h = 0.5 * dt * s->inv_mass[i];
        v[0] += h * s->force[i][0];
        v[1] += h * s->force[i][1];
        v[2] += h * s->force[i][2];
        s->pos[i][0] += v[0] * dt;
        s->pos[i][1] += v[1] * dt;
        s->pos[i][2] += v[2] * dt;
    }
}

void orc_nh_post_final(orc_sys* s) {
    /* nh_post_final (md_thermostat/src/lib.rs:313-353) */
    double dt = s->dt;
    double scale = exp(-dt / 2.0 * s->p_xi / s->nh_Q);
    for (int i = 0; i < s->nlocal; i++) {
        double* v = s->vel[i];
        double h = 0.5 * dt * s->inv_mass[i];
        v[0] += h * s->force[i][0];
        v[1] += h * s->force[i][1];
        v[2] += h * s->force[i][2];
        v[0] *= scale;
        v[1] *= scale;
        v[2] *= scale;
    }
    double ke = orc_compute_ke(s);
    s->p_xi += (dt / 2.0) * (2.0 * ke - s->nh_ndof * s->nh_T);
}

/* ---- scheduler: one step in the phase order of schedule.dot ------------------ */

int orc_step(orc_sys* s) {
    if (!s->setup_done) {
        orc_neighbor_setup(s);
        if (s->nh_on) orc_setup_nose_hoover(s);
    }
    /* PreInitialIntegration */
    if (s->nh_on) orc_nh_pre_initial(s);
    if (s->model == ORC_MODEL_DEM) orc_wall_move(s);
    /* InitialIntegration */
    if (s->vv_on) orc_initial_integration(s);
    if (s->rot_on) orc_initial_rotation(s);
    /* PostInitialIntegration */
    if (s->state == ORC_STATE_COMMUNICATE_ONLY) orc_decide_rebuild(s);
    orc_zero_all_forces(s);
    if (s->state == ORC_STATE_FULL_REBUILD) orc_remove_ghost_atoms(s);
    /* PreExchange */
    if (s->state == ORC_STATE_FULL_REBUILD) orc_pbc(s);
    /* PreNeighbor */
    if (s->state == ORC_STATE_COMMUNICATE_ONLY) orc_forward_comm(s);
    orc_sort_atoms_by_bin(s);
    if (s->state == ORC_STATE_FULL_REBUILD) orc_borders(s);
    /* Neighbor */
    if (s->state == ORC_STATE_FULL_REBUILD) orc_bin_neighbor_list(s);
    /* PreForce */
    if (s->model == ORC_MODEL_DEM) orc_wall_zero_force_accumulators(s);
    /* Force */
    if (s->model == ORC_MODEL_DEM) {
        if (orc_hertz_mindlin_contact_force(s) != 0) return -2;
        orc_wall_contact_force(s);
    } else {
        orc_lj_force(s);
    }
    orc_apply_gravity(s);
    /* PostForce */
    if (s->newton) orc_reverse_send_force(s);
    /* FinalIntegration */
    if (s->vv_on) orc_final_integration(s);
    if (s->rot_on) orc_final_rotation(s);
    /* PostFinalIntegration */
    if (s->nh_on) orc_nh_post_final(s);
    s->step++;
    return 0;
}

int orc_run(orc_sys* s, int64_t nsteps) {
    for (int64_t k = 0; k < nsteps; k++) {
        int rc = orc_step(s);
        if (rc) return rc;
    }
    return 0;
}

int orc_state(const orc_sys* s) { return s->state; }
double orc_p_xi(const orc_sys* s) { return s->p_xi; }
int64_t orc_rebuild_count(const orc_sys* s) { return s->rebuilds; }
int64_t orc_overlap_warnings(const orc_sys* s) { return s->overlap_total; }

int64_t orc_history_count(const orc_sys* s) {
    if (s->history_mode == ORC_HIST_CANONICAL) return (int64_t)s->cm.count;
    int64_t n = 0;
    for (int i = 0; i < s->nlocal; i++) n += s->hist[i].n;
    return n;
}

int64_t orc_get_history(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                        double* spring) {
    int64_t k = 0;
    if (s->history_mode == ORC_HIST_CANONICAL) {
        for (uint64_t m = 0; m < s->cm.cap; m++)
            if (s->cm.slots[m].used) {
                owner_tag[k] = (uint32_t)(s->cm.slots[m].key >> 32);
                partner_tag[k] = (uint32_t)(s->cm.slots[m].key & 0xffffffffu);
                memcpy(&spring[3 * k], s->cm.slots[m].s, sizeof(double) * 3);
                k++;
            }
        return k;
    }
    for (int i = 0; i < s->nlocal; i++)
        for (int m = 0; m < s->hist[i].n; m++, k++) {
            owner_tag[k] = s->tag[i];
            partner_tag[k] = s->hist[i].e[m].tag;
            memcpy(&spring[3 * k], s->hist[i].e[m].s, sizeof(double) * 3);
        }
    return k;
}

int64_t orc_get_history7(const orc_sys* s, uint32_t* owner_tag, uint32_t* partner_tag,
                         double* hist7) {
    int64_t k = 0;
    if (s->history_mode == ORC_HIST_CANONICAL) {
        for (uint64_t m = 0; m < s->cm.cap; m++)
            if (s->cm.slots[m].used) {
                owner_tag[k] = (uint32_t)(s->cm.slots[m].key >> 32);
                partner_tag[k] = (uint32_t)(s->cm.slots[m].key & 0xffffffffu);
                memcpy(&hist7[HSLOTS * k], s->cm.slots[m].s, sizeof(double) * HSLOTS);
                k++;
            }
        return k;
    }
    for (int i = 0; i < s->nlocal; i++)
        for (int m = 0; m < s->hist[i].n; m++, k++) {
            owner_tag[k] = s->tag[i];
            partner_tag[k] = s->hist[i].e[m].tag;
            memcpy(&hist7[HSLOTS * k], s->hist[i].e[m].s, sizeof(double) * HSLOTS);
        }
    return k;
}

int orc_set_history(orc_sys* s, int64_t n, const uint32_t* owner_tag, const uint32_t* partner_tag,
                    const double* spring) {
    if (s->history_mode == ORC_HIST_CANONICAL) {
        for (int64_t k = 0; k < n; k++) {
            uint32_t a = owner_tag[k] < partner_tag[k] ? owner_tag[k] : partner_tag[k];
            uint32_t b = owner_tag[k] < partner_tag[k] ? partner_tag[k] : owner_tag[k];
            cslot* c = cmap_find(&s->cm, ((uint64_t)a << 32) | b, 1);
            memcpy(c->s, &spring[3 * k], sizeof(double) * 3);
        }
        return 0;
    }
    for (int64_t k = 0; k < n; k++) {
        int found = -1;
        for (int i = 0; i < s->nlocal; i++)
            if (s->tag[i] == owner_tag[k]) {
                found = i;
                break;
            }
        if (found < 0) {
            set_err(s, "history owner tag not found");
            return -1;
        }
        hist_push(&s->hist[found], partner_tag[k], &spring[3 * k], 3, 0);
    }
    return 0;
}
