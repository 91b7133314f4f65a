/*
 * pot_host.h -- host side of the drop-in pot programs: the WELL512 stream,
 * the initial particle load, the reference's run loop around the C ABI of
 * potg.h, and the output files.
 */
#ifndef POT_HOST_H
#define POT_HOST_H

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

#ifdef __cplusplus
}
#endif
#endif
