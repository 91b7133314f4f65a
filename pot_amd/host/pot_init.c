/*
 * pot_init.c -- WELL512 stream and the initial particle load on the host.
 *
 * The initial state is generated on the CPU exactly as the reference does
 * it, so a run started from the same 64-byte seed begins from the same
 * particles: WELL512 (prng.c:4-47) and init_particles/inject
 * (pot.c:658-765, 887-940).  Arithmetic types follow the reference (the
 * stream returns long double) because that is what fixes the bits.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "pot_host.h"

#define KB (1.38065e-23)
#define ME (9.109383e-31)
#define PI 3.14159265358979

void pot_prng_seed(pot_prng* st, const uint32_t seed[16])
{
    memcpy(st->state, seed, sizeof(st->state));
    st->index = 0;
}

/* WELL512 (Panneton, L'Ecuyer, Matsumoto 2006) as prng.c:11-28 runs it */
long double pot_rnd(pot_prng* st, long double maximum)
{
    const uint32_t i = st->index;
    uint32_t a = st->state[i];
    uint32_t c = st->state[(i + 13) & 15];
    const uint32_t b = a ^ c ^ (a << 16) ^ (c << 15);
    c = st->state[(i + 9) & 15];
    c ^= (c >> 11);
    st->state[i] = b ^ c;
    const uint32_t d = a ^ ((a << 5) & 0xda442d24UL);
    const uint32_t j = (i + 15) & 15;
    a = st->state[j];
    st->state[j] = a ^ b ^ d ^ (a << 2) ^ (b << 18) ^ (c << 28);
    st->index = j;
    st->a = a;
    st->b = b;
    st->c = c;
    st->d = d;
    return maximum * (st->state[j] / (long double)UINT32_MAX);
}

/* polar Box-Muller with a zero retry (prng.c:32-47) */
void pot_nrv_pair(pot_prng* st, double sd, double* n1, double* n2)
{
    const long double u2 = pot_rnd(st, 1.0L);
    long double u1;
    do {
        u1 = pot_rnd(st, 1.0L);
    } while (u1 == 0.0);
    const double r = sqrt(-2.0 * (double)logl(u1));
    const double th = 2.0 * PI * u2;
    *n1 = sd * r * cos(th);
    *n2 = sd * r * sin(th);
}

/* first injection of one particle in a spherical domain (pot.c:678-740) */
static void inject_initial(const potg_config* k, const long double mass[3], potg_particle* p,
                           pot_prng* st)
{
    double sigma = (p->species == POTG_ION) ? sqrt(KB * k->Ti / mass[POTG_ION])
                                            : sqrt(KB * k->Te / mass[POTG_ELECTRON]);
    if (!k->sceptic) sigma *= 0.85; /* pot.c:683-691 */
    const unsigned int r3 = k->r * k->r * k->r; /* unsigned arithmetic, pot.c:703 */
    double r = k->mpp * cbrt(pot_rnd(st, r3));
    if (k->grain_enabled || k->geometry != POTG_SPHERE)
        while (r <= k->a) r = k->mpp * cbrt(pot_rnd(st, r3));
    const double phi = pot_rnd(st, 2.0 * PI);
    const double theta = acos(pot_rnd(st, 2.0) - 1.0);
    p->x = (k->mpp * k->r) + (r * sin(theta) * cos(phi));
    p->y = (k->mpp * k->r) + (r * sin(theta) * sin(phi));
    p->z = (k->mpp * k->r) + (r * cos(theta));
    pot_nrv_pair(st, sigma, &p->vx, &p->vy);
    pot_nrv_pair(st, sigma, &p->vz, &p->vz);
    p->vx += k->x_drift;
}

int pot_init_particles(const potg_config* k, const uint32_t seed[16], potg_particle* parts,
                       pot_prng* st)
{
    const uint64_t N = k->N;
    const long double mass[3] = {1.67e-12, k->mime * ME, ME};
    uint64_t i;
    if (N < 1) return POTG_ERR_ARG;
    pot_prng_seed(st, seed);
    memset(parts, 0, sizeof(potg_particle) * N);
    parts[0].x = k->r * k->mpp;
    parts[0].y = parts[0].x;
    parts[0].z = parts[0].x;
    parts[0].species = POTG_DUST;
    const uint64_t n_i = (N - 1) / 2;
    for (i = 1; i < N; i++) parts[i].species = (i <= n_i) ? POTG_ION : POTG_ELECTRON;
    for (i = 1; i < N; i++) inject_initial(k, mass, &parts[i], st);
    return POTG_OK;
}
