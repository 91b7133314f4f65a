/*
 * potg.h -- C ABI of the MI355X (gfx950) hot path of the pot treecode.
 *
 * One context holds one or more independent grain simulations
 * ("instances") resident in HBM.  A step advances every instance by one
 * time step: octree build (Morton keys with the reference's float64
 * midpoint recurrence, radix sort, bottom-up octree, monopole moments),
 * Barnes-Hut field walk, Boris push, grain crossing with exact integer
 * tallies, and SCEPTIC-like reinjection from a counter-based stream.
 *
 * Plain C types only; every call returns POTG_OK (0) or a negative
 * POTG_ERR_* code and leaves a message for potg_last_error().  A context
 * is used from one host thread at a time.
 *
 * Reference interfaces replaced (drewthomas/pot; SPH = pot-spheroid.c):
 *   build_tree(Proc*, Run)                         pot.c:1457, SPH:1478
 *   field_particle_feels_at(...)                   pot.c:1696, SPH:1794
 *   phi_at(...), phi_particle_feels(...)           pot.c:1818, 1934
 *   move_particles(Proc*, Run, Tree*, ..., t)      pot.c:2330, SPH:2541
 *   assess_field_accuracy(...) (direct sum)        pot.c:2198
 */
#ifndef POTG_H
#define POTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POTG_OK 0
#define POTG_ERR_ARG (-1)        /* bad argument or configuration */
#define POTG_ERR_HIP (-2)        /* HIP runtime failure */
#define POTG_ERR_OUT_OF_BOX (-3) /* particle outside [0,2R]^3 or NaN (pot.c:1503-1514) */
#define POTG_ERR_COINCIDENT (-4) /* two particles at one position (pot.c:1578-1586, 1799-1802) */
#define POTG_ERR_DEPTH (-5)      /* tree deeper than the walk stack (pot.c:1647-1652) */
#define POTG_ERR_NOMEM (-6)      /* device allocation failed */
#define POTG_ERR_REINJECT (-7)   /* SCEPTIC reinjection found no admissible entry in
                                    1000 tries: chi_b bars the species (the reference
                                    retries without end, pot.c:583-598) */

/* species codes, pot.c:286-290 */
#define POTG_DUST 0
#define POTG_ION 1
#define POTG_ELECTRON 2

/* grain geometry: sphere (pot.c) or spheroid with axis along x (SPH:111-121) */
#define POTG_SPHERE 0
#define POTG_PROLATE 1
#define POTG_OBLATE 2

/* particle integrator: Boris (BORIS, pot.c:130-132), velocity Verlet (the
   reference without BORIS) or Euler-Richardson (EULER_RICHARDSON,
   pot.c:122-128); the last two take a second field per step */
#define POTG_BORIS 0
#define POTG_VERLET 1
#define POTG_EULER_RICHARDSON 2

/* arithmetic of the field walk */
#define POTG_FP64 0  /* float64 everywhere */
#define POTG_MIXED 1 /* float32 walk with compensated sums; float64 push */
/* As POTG_MIXED, but the 64 targets of a wavefront share one interaction
   list built with the opening criterion applied to their bounding box:
   every cell the reference accepts for a target is taken whole or in
   parts, so forces are at least as accurate; no counterpart in the
   reference. */
#define POTG_MIXED_GROUP 2

/* The reference's 56-byte AoS particle record (pot.c:295-303); also the
   record format of state files (pot.c:3888). */
typedef struct potg_particle {
    double x, y, z, vx, vy, vz;
    int32_t species;
    int32_t pad;
} potg_particle;

/* Run parameters of one instance (pot.c:335-358, SPH:376-378) together
   with the compile-time switches of the reference as run-time values. */
typedef struct potg_config {
    uint64_t N;            /* -N: particles including the grain at index 0 */
    uint32_t r;            /* -r: domain radius in pixels */
    uint32_t geometry;     /* POTG_SPHERE / POTG_PROLATE / POTG_OBLATE */
    double dt;             /* -d: time step (s) */
    double mpp;            /* -m: metres per pixel */
    double soft;           /* -s: softening length (m) */
    double x_drift;        /* -x: ion drift velocity along x (m/s) */
    double a;              /* -a: grain equal-volume radius (m) */
    double settle;         /* -e: settling time before absorption (s) */
    double Te, Ti;         /* -E, -I: temperatures (K) */
    double A;              /* -A: spheroid aspect ratio */
    double theta;          /* OAP, the opening angle (pot.c:97) */
    double B[3];           /* EXTERNAL_B_{X,Y,Z} (T) (pot.c:55-57) */
    double max_lin_dev;    /* integrate_for_alpha tolerance (pot.c:436; SPH:457) */
    double mime;           /* MIME, ion/electron mass ratio (pot.c:60) */
    int32_t grain_enabled; /* absorbing grain (!DISABLE_GRAIN, pot.c:30) */
    int32_t sceptic;       /* SCEPTIC_REINJECTION (pot.c:180-182); must be 1 */
    uint64_t seed;         /* key of the counter-based reinjection stream */
    int32_t dipoles;       /* DIPOLES (pot.c:99-105): dipole moments and terms */
    int32_t integrator;    /* POTG_BORIS, POTG_VERLET or POTG_EULER_RICHARDSON
                              (pot.c:122-132); one value for every instance
                              of a context */
    int32_t out_bounce;    /* OUT_BOUNCE (pot.c:113-120): particles leaving the
                              domain bounce off its wall instead of being
                              reinjected */
    int32_t pad;
    double bounce_coe;     /* OUT_BOUNCE's value, the coefficient of restitution */
} potg_config;

/* Scalar state of an instance besides its particles. */
typedef struct potg_grain {
    int32_t q;           /* grain charge in e (pot.c:371) */
    uint32_t tot_absor_e; /* electrons absorbed so far (pot.c:395) */
    double m;            /* grain mass (kg) (pot.c:372) */
    double v[3];         /* grain velocity (parts[0].v) */
    double chi_b;        /* normalised edge potential (pot.c:397) */
    double epe[3];       /* electrostatic PE per species (pot.c:369) */
} potg_grain;

/* Per-instance result of one step (the reductions of pot.c:2813-2824). */
typedef struct potg_step_out {
    int32_t abs_q;        /* net charge absorbed this step (exact) */
    uint32_t abs_e;       /* electrons absorbed this step (exact) */
    uint32_t abs_i;       /* ions absorbed this step (exact) */
    uint32_t n_pushed;    /* particles stepped */
    uint32_t n_reinj_abs; /* reinjections after absorption */
    uint32_t n_reinj_out; /* reinjections after leaving the domain */
    uint32_t n_wall_inward; /* OUT_BOUNCE: outside the wall but moving inwards,
                               left alone (pot.c:2724-2728 warns of each) */
    uint32_t pad;
} potg_step_out;

/* Work counters accumulated since the last potg_stats reset. */
typedef struct potg_stats_t {
    uint64_t steps;
    uint64_t n_targets;  /* walk targets */
    uint64_t n_cell;     /* accepted cells with q != 0 */
    uint64_t n_leaf;     /* leaf-pair interactions */
    uint64_t n_open;     /* internal nodes opened (MAC failed) */
    uint64_t n_nodes;    /* tree nodes built (internal + leaves) */
    uint64_t n_particles;
    uint64_t max_depth;
    uint64_t n_range_loads; /* child ranges fetched by walk wavefronts */
    uint64_t n_child_visits; /* children visited by float32-walk wavefronts (64
                                targets each): (n_cell + n_leaf) / (64 n_child_visits)
                                is the fraction of lanes doing an interaction */
} potg_stats_t;

/* Launch parameters of the field walk (potg_tune picks them by timing).
   Until it or potg_set_tuning runs they follow the context's size, in
   64-target tiles per CU (half the particles are targets): the float32
   walk sweeps whole ranges with 64 targets per wavefront on 8 blocks per CU
   from 16 tiles per CU (about 0.5M particles), on 2 blocks per CU from 1
   tile per CU, and below that takes 8 targets per wavefront on 2 blocks
   per CU; the float64 walk takes 64 targets on 8 blocks per CU from 16
   tiles per CU, 4 below. */
typedef struct potg_tuning {
    uint32_t walk_blocks; /* blocks (4 wavefronts each) of the persistent walk */
    uint32_t group;       /* float32 walk: targets per wavefront, 8, 16 or 64;
                             with 8 or 16 the wavefront evaluates 8 or 4
                             children per target at once */
    uint32_t variant;     /* float32 walk with 64 targets: 0 visits a range's
                             cells until some lanes open one, 1 sweeps every
                             range whole before descending */
    uint32_t sort_tile;   /* keys per tile of the radix sort: 2048 or 4096 */
    double walk_ms;       /* potg_tune: the chosen walk's time (ms) */
} potg_tuning;

typedef struct potg_ctx potg_ctx;

/* Create a context for n_inst instances (configs[i]); precision is
   POTG_FP64 or POTG_MIXED.  Allocates every device buffer up front. */
int potg_create(const potg_config* configs, uint32_t n_inst, int32_t precision,
                potg_ctx** out);
int potg_destroy(potg_ctx* ctx);
const char* potg_last_error(void);

/* Check a particle array for instance configuration cfg as the reference's
   build_tree does before inserting (pot.c:1503-1514, with the NaN test of
   is_particle_position_nan, pot.c:1434-1446): particle 0 is the grain, the
   others are ions or electrons (POTG_ERR_ARG otherwise), and every position
   lies in the root cell [0, 2 r mpp]^3 with no NaN coordinate
   (POTG_ERR_OUT_OF_BOX otherwise, the reference's message in
   potg_last_error).  *bad_index (may be NULL) receives the first offending
   index.  Host only; potg_upload calls it. */
int potg_check_particles(const potg_config* cfg, const potg_particle* parts,
                         uint64_t* bad_index);

/* Copy one instance's particles (reference index order, grain at index 0)
   and grain state to the device, and set the step counter t (the number of
   steps already completed, pot.c:4261). */
int potg_upload(potg_ctx* ctx, uint32_t inst, const potg_particle* parts,
                const potg_grain* grain, uint64_t t_completed);
/* Copy one instance back, in reference index order. */
int potg_download(potg_ctx* ctx, uint32_t inst, potg_particle* parts, potg_grain* grain);

/* Advance every instance by one time step (run(), pot.c:4261-4388, with
   move_particles, pot.c:2330-2934).  When compute_epe is nonzero the
   potential-energy walks of pot.c:2771-2866 run too.  With out non-NULL
   the call waits for the step, reports kernel errors and fills n_inst
   records; with out NULL it only enqueues the step (errors surface at
   potg_sync). */
int potg_step(potg_ctx* ctx, int32_t compute_epe, potg_step_out* out);
/* Wait for all enqueued work and report any error a kernel raised. */
int potg_sync(potg_ctx* ctx);
/* Run `steps` steps as replays of one captured HIP graph (no host
   synchronisation between steps; no per-step outputs). */
int potg_run(potg_ctx* ctx, uint64_t steps);

/* Build the octree of the current state without stepping (build_tree). */
int potg_build_tree(potg_ctx* ctx);
/* Canonical listing of instance inst's tree: DFS pre-order over the
   non-empty nodes with children in the reference order (pot.c:324-331).
   Arrays of capacity cap; returns the node count in *n_out. */
int potg_tree_canon(potg_ctx* ctx, uint32_t inst, int32_t* depth, int32_t* is_leaf,
                    uint32_t* part, int32_t* q, uint32_t* aq, double* c, double* x1x2,
                    int32_t* nchild, int64_t cap, int64_t* n_out);

/* Dipole moments (x, y, z per node, about the centre of |q|) of instance
   inst's tree in the order of potg_tree_canon (DIPOLES, pot.c:1290-1307);
   zero for leaves and for instances without dipoles. */
int potg_tree_canon_dipoles(potg_ctx* ctx, uint32_t inst, double* p, int64_t cap,
                            int64_t* n_out);

/* Field at n points of instance inst using the current tree (built by
   potg_build_tree or the last step), as field_particle_feels_at does:
   exclude[k] is the reference index whose own leaf is skipped (UINT32_MAX
   for none).  counts, if non-NULL, receives 3 per point: accepted cells,
   leaf pairs, opened nodes.  Uses the context's precision. */
int potg_field_at(potg_ctx* ctx, uint32_t inst, const double* xyz, const uint32_t* exclude,
                  size_t n, double* E, uint32_t* counts);
/* Potential at n points (phi_at, pot.c:1818-1901). */
int potg_phi_at(potg_ctx* ctx, uint32_t inst, const double* xyz, size_t n, double* phi);
/* potg_phi_at for callers that cannot allocate: code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc.  No
   counterpart in the reference.  d_xyz (3 n doubles) and d_phi (n doubles)
   are device memory; the call allocates nothing (its scratch is a fixed
   array of about 0.5 MB per device, held until the process ends), does not
   synchronise and enqueues on `stream` (a hipStream_t; NULL for the
   context's stream).  n may be at most POTG_NOALLOC_MAX_POINTS, else
   POTG_ERR_ARG.  Calls that overlap on one GPU share the scratch and race.
   Results equal potg_phi_at's; errors a kernel raises surface at potg_sync. */
#define POTG_NOALLOC_MAX_POINTS 65536u
int potg_phi_at_noalloc(potg_ctx* ctx, uint32_t inst, const double* d_xyz, size_t n,
                        double* d_phi, void* stream);
/* Softened float64 direct sum at the positions of particles idx[0..n) of
   instance inst over all its other particles (E_ref of the force checks). */
int potg_direct_sum(potg_ctx* ctx, uint32_t inst, const uint32_t* idx, size_t n,
                    double* E);

/* Time the field walk of the coming step under candidate launch parameters
   (targets per wavefront, grid size), keep the fastest and report it in
   *out (out may be NULL).  The step itself is not taken: the state, step
   counter and work counters are left as they were.  No counterpart in the
   reference. */
int potg_tune(potg_ctx* ctx, potg_tuning* out);
/* Set or read the walk's launch parameters (set NULL to only read). */
int potg_set_tuning(potg_ctx* ctx, const potg_tuning* set, potg_tuning* get);

int potg_stats(potg_ctx* ctx, potg_stats_t* out);
int potg_stats_reset(potg_ctx* ctx);
/* What a SCEPTIC reinjection that finds no admissible entry does (the
   reference recurses without end there, pot.c:583-598).  Off (0, the
   default): the step returns POTG_ERR_REINJECT.  On: the particle's
   instance retires -- the particle is parked at rest inside the grain and,
   from the next step on, none of the instance's particles is stepped and
   its chi_b is no longer measured -- and the other instances go on.  For
   parameter sweeps, where one instance's failure should not end the batch.
   No counterpart in the reference. */
int potg_set_retire(potg_ctx* ctx, int32_t on);
/* Whether instance inst has retired (*retired = 1) and the step in which
   its reinjection failed (*step, may be NULL; 0 when not retired).
   potg_upload of the instance clears it. */
int potg_instance_retired(potg_ctx* ctx, uint32_t inst, int32_t* retired, uint64_t* step);

/* Interaction counters (n_cell, n_leaf, n_open, n_range_loads) of the step
   walks on (on != 0, the default) or off.  They are instrumentation: off,
   the float32 walk runs about 8-10% faster; point queries count either way.
   No counterpart in the reference. */
int potg_set_counting(potg_ctx* ctx, int32_t on);

/* Per-stage wall time (ms) averaged over the potg_step calls since timing
   was last enabled, from HIP events on the context's stream: [0] keys +
   sort, [1] tree build, [2] walk, [3] push + reinjection, [4] whole step.
   enable turns recording on or off for the steps that follow. */
int potg_stage_times(potg_ctx* ctx, int32_t enable, double* ms5);

#ifdef __cplusplus
}
#endif
#endif
