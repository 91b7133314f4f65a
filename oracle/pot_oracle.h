/*
 * pot_oracle.h -- CPU oracle for the pot per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C99 restatement of the
 * reference treecode's hot path (pot.c, pot-spheroid.c, prng.c, coord.c of
 * drewthomas/pot).  It exists to check the HIP implementation; nothing in the
 * product links it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic
 * types (double vs long double) and operation order follow the reference so
 * that, in "reference" mode (WELL512 stream, in-place half drift), a run is
 * bit-identical to the reference binary built from its own sources.
 */
#ifndef POT_ORACLE_H
#define POT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pot.c:286-290 */
enum { OR_DUST = 0, OR_ION = 1, OR_ELECTRON = 2 };
/* geometry of the grain: sphere (pot.c), prolate/oblate (pot-spheroid.c:111-121) */
enum { OR_SPHERE = 0, OR_PROLATE = 1, OR_OBLATE = 2 };
/* reinjection random stream: the reference WELL512 (prng.c:11-28) or the
   counter-based Philox4x32-10 stream the GPU uses */
enum { OR_RNG_WELL512 = 0, OR_RNG_PHILOX = 1 };
/* particle integrator: BORIS (pot.c:130-132), velocity Verlet (the
   reference's default without BORIS, pot.c:122-128) or EULER_RICHARDSON */
enum { OR_BORIS = 0, OR_VERLET = 1, OR_EULER_RICHARDSON = 2 };

/* pot.c:295-303: the reference's 56-byte AoS particle record */
typedef struct or_particle {
    double x, y, z, vx, vy, vz;
    int32_t species;
    int32_t pad;
} or_particle;

/* prng.h:17-24: WELL512 state, 84 bytes, byte-compatible with state files */
typedef struct or_prng {
    uint32_t a, b, c, d, index;
    uint32_t state[16];
} or_prng;

/* Run parameters (pot.c:335-358, pot-spheroid.c:376-378) plus the
   compile-time switches of the reference made runtime. */
typedef struct or_config {
    uint64_t N;          /* particles including the grain at index 0 */
    uint32_t r;          /* domain radius in pixels */
    uint32_t geometry;   /* OR_SPHERE / OR_PROLATE / OR_OBLATE */
    long double dt;      /* time step (s), parsed with strtold as -d */
    double mpp;          /* metres per pixel */
    double soft;         /* softening length (m) */
    double x_drift;      /* ion drift velocity along x (m/s) */
    double a;            /* grain (equal-volume) radius (m) */
    double settle;       /* settling time before absorption (s) */
    double Te, Ti;       /* temperatures (K) */
    double A;            /* spheroid aspect ratio */
    double theta;        /* opening angle OAP */
    double B[3];         /* EXTERNAL_B_{X,Y,Z} (T) */
    double max_lin_dev;  /* integrate_for_alpha tolerance: 0.01 pot.c:436, 0.1 SPH:457 */
    double mime;         /* MIME: 1836.15 pot.c:60, 1836.153 SPH:61 */
    int32_t grain_enabled;   /* !DISABLE_GRAIN */
    int32_t sceptic;         /* SCEPTIC_REINJECTION */
    int32_t inplace;         /* 1: reference in-place half drift feeds leaf sources */
    int32_t rng_kind;        /* OR_RNG_* */
    uint64_t philox_seed;    /* key of the counter-based stream */
    int32_t record_macro;    /* RECORD_MACRO_DATA switch (drives compute_epe cadence) */
    int32_t dipoles;         /* DIPOLES switch (pot.c:99-105): cell dipole moments */
    int32_t integrator;      /* OR_BORIS / OR_VERLET / OR_EULER_RICHARDSON */
    int32_t out_bounce;      /* OUT_BOUNCE switch (pot.c:113-120): bounce off the wall */
    double bounce_coe;       /* OUT_BOUNCE's value: coefficient of restitution */
} or_config;

/* Per-walk interaction counters. */
typedef struct or_walk_stats {
    uint64_t n_cell;   /* accepted cells with q != 0 (monopole evaluated) */
    uint64_t n_leaf;   /* leaf pair interactions evaluated */
    uint64_t n_open;   /* internal nodes whose MAC failed (opened) */
    uint64_t n_mac;    /* MAC tests on non-empty internal nodes */
} or_walk_stats;

/* Per-step tallies (pot.c:2345-2359). */
typedef struct or_step_out {
    int32_t abs_q;
    uint32_t abs_e;
    uint32_t abs_i;
    uint32_t n_pushed;
    uint32_t n_reinj_abs;    /* reinjections after grain absorption */
    uint32_t n_reinj_out;    /* reinjections after leaving the domain */
    uint32_t start;          /* first stepped index */
    uint32_t n_wall_inward;  /* OUT_BOUNCE: outside the wall but moving inwards
                                (check_particle_in_sim_area returns -1) */
} or_step_out;

typedef struct or_sim or_sim;

/* parse -d as the reference does (strtold), and the strtold-then-double
   options -E, -I, -e */
void or_config_set_dt(or_config* cfg, const char* text);
double or_strtold_to_double(const char* text);

or_sim* or_create(const or_config* cfg);
void or_destroy(or_sim* s);
const char* or_last_error(void);

/* particle state access (reference index order) */
void or_get_particles(const or_sim* s, or_particle* out);
void or_set_particles(or_sim* s, const or_particle* in);
or_particle* or_particles_ptr(or_sim* s);

/* scalar state: grain q, grain m, elapsed steps, chi_b, absorbed electrons */
void or_get_state(const or_sim* s, int32_t* q, long double* m, uint64_t* t,
                  double* chi_b, uint32_t* tot_e, long double epe[3]);
void or_set_state(or_sim* s, int32_t q, long double m, uint64_t t,
                  double chi_b, uint32_t tot_e);
void or_get_prng(const or_sim* s, or_prng* out);
void or_set_prng(or_sim* s, const or_prng* in);
void or_get_tss(const or_sim* s, uint32_t tss[3]);

/* initial state: seed WELL512 and inject all particles (pot.c:887-996) */
void or_init_prng(or_prng* st, const uint32_t seed[16]);
void or_init_particles(or_sim* s, const uint32_t seed[16]);

/* tree (pot.c:1457-1643) */
int or_build_tree(or_sim* s);
/* canonical DFS listing of the non-empty nodes, children in reference order */
int64_t or_tree_canon(const or_sim* s, int32_t* depth, int32_t* is_leaf,
                      uint32_t* part, int32_t* q, uint32_t* aq, double* c,
                      double* x1x2, int32_t* nchild, int64_t cap);
/* dipole moments of the canonical listing (zeros without DIPOLES) */
int64_t or_tree_canon_dipoles(const or_sim* s, double* p, int64_t cap);
uint32_t or_tree_used(const or_sim* s);

/* walks (pot.c:1696-1808, 1818-1901, 1934-2018; SPH:1689-1780, 1922-2084) */
void or_field_at(or_sim* s, uint32_t part, double x, double y, double z,
                 int32_t accum_double, double E[3], or_walk_stats* st);
double or_phi_at(or_sim* s, double x, double y, double z);
double or_phi_particle_feels(or_sim* s, uint32_t part);
void or_field_from_grain(const or_sim* s, uint32_t part, double x, double y,
                         double z, double E[3]);
double or_phi_from_grain(const or_sim* s, uint32_t part);
/* softened direct sum over every other particle (reference oracle of
   pot.c:2198-2237, with the grain's actual charge) */
void or_direct_sum(const or_sim* s, uint32_t part, double x, double y,
                   double z, double E[3]);
double or_measure_chi_b(or_sim* s);

/* one full time step of run() (pot.c:4259-4388) for one process */
int or_run_step(or_sim* s, or_step_out* out);
/* only the move (pot.c:2330-2934) with a caller-chosen compute_epe */
int or_move_particles(or_sim* s, int compute_epe, or_step_out* out);
int or_record_macro_this_step(const or_sim* s, uint64_t t);

/* state files, format C (pot.c:3816-4019) */
int or_write_state(const or_sim* s, const char* path);
int or_read_state(or_sim* s, const char* path);

/* samplers and helpers (prng.c, coord.c, pot.c:433-652) */
long double or_rnd(or_prng* st, long double maximum);
void or_nrv_pair(or_prng* st, double sd, double* n1, double* n2);
double or_u_no_flow_icdf(double chi, double q);
double or_u_icdf(double U, double chi, double q);
double or_integrate_for_alpha(double s, double chi_b, double u2, double Rol,
                              double max_lin_dev);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* reinjection driven by a caller-supplied list of uniforms in [0,1]
   (rnd(max) = max * u); returns the number of uniforms consumed, or -1 */
int or_sceptic_from_uniforms(const or_sim* s, or_particle* p, double chi_b,
                             const double* u, int n_u);
/* one SCEPTIC reinjection from the sim's own stream */
void or_reinject(or_sim* s, or_particle* p, uint32_t ref_idx);
/* grain crossing test (pot.c:2287-2325; SPH:2448-2540) */
int or_crosses_grain(const or_sim* s, const double p_new[3], const double d_pos[3],
                     const double grain[3]);

#ifdef __cplusplus
}
#endif
#endif
