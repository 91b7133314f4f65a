/*
 * pot_host.h -- host side of the drop-in pot programs: the WELL512 stream,
 * the initial particle load, the reference's run loop around the C ABI of
 * potg.h, and the output files.
 */
#ifndef POT_HOST_H
#define POT_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "potg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* WELL512 state, byte-compatible with the reference's PRNG record
   (prng.h:17-24) as written into state files. */
typedef struct pot_prng {
    uint32_t a, b, c, d, index;
    uint32_t state[16];
} pot_prng;

void pot_prng_seed(pot_prng* st, const uint32_t seed[16]);
/* uniform in [0, maximum] (prng.c:11-28) */
long double pot_rnd(pot_prng* st, long double maximum);
/* two normal variates of standard deviation sd (prng.c:32-47) */
void pot_nrv_pair(pot_prng* st, double sd, double* n1, double* n2);

/* Initial load of one instance (init_particles, pot.c:887-940, for one
   process): grain at the domain centre, ions at 1..(N-1)/2, electrons
   after them, each injected uniformly in the shell (a, R] with Maxwellian
   velocities (inject, pot.c:658-765).  The stream is seeded from seed and
   left in *st. */
int pot_init_particles(const potg_config* cfg, const uint32_t seed[16], potg_particle* parts,
                       pot_prng* st);

/* Settings of a run: the reference's Run record (pot.c:334-358) and, as
   run-time values, the compile-time switches it is built with. */
typedef struct pot_run {
    unsigned long N;
    unsigned int r;
    long double dt;
    double mpp;
    unsigned long run_for;  /* 0: no limit */
    double soft, x_drift, a, settle, tim_lim, Te, Ti;
    double A;               /* spheroid aspect ratio */
    char* sw_path;          /* -f */
    const char* sr_path;    /* -g / -R */
    int read_out;           /* -R */
    /* switches (pot.c:25-200, SPH:25-230) */
    int geometry;           /* POTG_SPHERE / POTG_PROLATE / POTG_OBLATE */
    double theta;           /* OAP */
    double B[3];            /* EXTERNAL_B_{X,Y,Z} */
    double mime;            /* MIME */
    int grain;              /* !DISABLE_GRAIN */
    int dipoles;            /* DIPOLES */
    int integrator;         /* POTG_BORIS / POTG_VERLET / POTG_EULER_RICHARDSON */
    int out_bounce;         /* OUT_BOUNCE */
    double bounce_coe;      /* its value, the coefficient of restitution */
    int tune;               /* time the walk's launch shapes (potg_tune) */
    int precision;          /* POTG_FP64 / POTG_MIXED */
    const char* seed_path;  /* RANDOMNESS_SOURCE */
    const char* out_dir;    /* OUTPUT_DIR */
} pot_run;

/* pot-run-init-info.txt line (pot.c:3512-3553) */
void pot_write_run_info(const char* path, const pot_run* ru);
/* pot-run-descrip.txt (pot.c:3583-3813); def holds the program defaults */
void pot_write_run_description(const char* path, const pot_run* ru, const pot_run* def);

/* Scalars of a state file besides the particles (pot.c:3870-3873). */
typedef struct pot_sim_state {
    int32_t grain_q;
    uint32_t tot_absor_e;
    long double grain_m;
    uint64_t t_completed;  /* steps completed */
    long double epe[3];
    double chi_b;
    pot_prng prng;         /* the WELL512 record carried by the file */
} pot_sim_state;

/* Format C state file of one worker (write_sim_state_to_file,
   pot.c:3816-3901); sceptic 0 stores 9999.99 for chi_b.  0 on success. */
int pot_write_state(const char* path, uint64_t N, const potg_particle* parts,
                    const pot_sim_state* s, int sceptic);
/* Read a format A, B or C state file of one worker
   (read_sim_state_from_file, pot.c:3904-4019).  Fields a format lacks keep
   the values *s held.  0 on success; messages go to stderr. */
int pot_read_state(const char* path, uint64_t N, potg_particle* parts, pot_sim_state* s);

/* What one macroscopic record reads (write_macroscopic_data_to_file,
   pot.c:3261-3506): particles in reference order after the step, the grain
   scalars and the potential energies of the last compute_epe step. */
typedef struct pot_macro_in {
    const potg_particle* parts;
    uint64_t N;
    uint32_t r;            /* domain radius in pixels */
    double mpp;
    long double dt;
    double mime;
    long double grain_m;
    int32_t grain_q;
    uint32_t tot_absor_e;
    long double epe[3];    /* indexed by species */
    uint64_t time_step;
} pot_macro_in;

/* column header written when the file is new (pot.c:5229-5234) */
const char* pot_macro_header(void);
/* Format one record (with its newline) into buf; returns its length, 0
   when N < 6 (no record), -1 when cap is too small. */
int pot_macro_format(const pot_macro_in* in, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
