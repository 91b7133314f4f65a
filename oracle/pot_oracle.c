/*
 * pot_oracle.c -- CPU oracle for the pot per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pot_oracle.h).  A C99 restatement of the
 * reference's hot path for one MPI process (P = 1):
 *   build_tree / compute_node_values      pot.c:1225-1402, 1457-1643
 *   field_particle_feels_at               pot.c:1696-1808, SPH:1794-1921
 *   phi_at / phi_particle_feels           pot.c:1818-1901, 1934-2018
 *   spheroid grain field / potential      SPH:1689-1780, 1922-1953, 2040-2084
 *   Boris push, crossing, tally, boundary pot.c:2330-2934, SPH:2448-2540
 *   inject / sceptically_reinject         pot.c:433-846
 *   WELL512 rnd, nrv_pair, inverse CDFs   prng.c:4-214
 *   rotations                             coord.c:6-59
 * The arithmetic keeps the reference's types (double vs long double) and
 * evaluation order, because "reference" mode must reproduce a run of the
 * reference binary bit for bit.  The restated expressions are necessarily
 * alike the reference's; the control structure (iterative retry instead of
 * recursion, explicit stream objects, run-time switches) is this file's own.
 *
 * Two switches select the semantics the GPU path uses instead:
 *   inplace = 0   leaf sources are the positions the tree was built from
 *                 (the reference drifts lower-indexed particles in place
 *                 before later walks, pot.c:2478-2485);
 *   rng_kind = 1  reinjection draws from a counter-based Philox4x32-10
 *                 stream keyed by (seed, particle, step) instead of WELL512.
 */
#define _GNU_SOURCE
#include "pot_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* physical constants, pot.c:278-282; PI from prng.h:9-14 (included first) */
#define KB (1.38065e-23)
#define EPSILON0 (8.85419e-12)
#define QE (1.6021766e-19)
#define ME (9.109383e-31)
#define PI 3.14159265358979
#define SQRT_PI 1.77245385090552
#define MAX_PATH_LEN 400 /* pot.c:196 */

static const int CHARGE[3] = {0, 1, -1}; /* pot.c:291 */

#define pythag(x, y, z) sqrt((x) * (x) + (y) * (y) + (z) * (z)) /* pot.c:270 */

/* reference Tree node (pot.c:306-332); children kept as an array in the
   reference order gxlylz, gxlygz, lxlylz, lxlygz, lxgylz, lxgygz, gxgylz,
   gxgygz */
typedef struct or_node {
    double x1, y1, z1, x2, y2, z2;
    unsigned int part;
    double cx, cy, cz;
    int q;
    unsigned int aq;
    double px, py, pz; /* dipole about (cx, cy, cz), DIPOLES only */
    unsigned int ch[8];
} or_node;

enum { GXLYLZ, GXLYGZ, LXLYLZ, LXLYGZ, LXGYLZ, LXGYGZ, GXGYLZ, GXGYGZ };

/* Philox stream for the counter-based reinjection draws */
typedef struct or_pstream {
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t buf[4];
    int pos;
} or_pstream;

typedef struct or_rng {
    int kind; /* 0 WELL512, 1 Philox, 2 caller list */
    or_prng* well;
    or_pstream ps;
    const double* u;
    int n_u;
    int used;
    int exhausted;
} or_rng;

struct or_sim {
    or_config ru;
    long double mass[3];
    unsigned int tss[3];
    or_particle* parts;
    or_particle* new_parts;
    or_particle* tree_parts; /* snapshot the tree was built from */
    or_prng ps;
    int dust_grain_q;
    long double dust_grain_m;
    long double epe[3];
    unsigned long t_steps;
    unsigned int tot_absor_e;
    double chi_b;
    or_node* root;
    unsigned int tr_alloc;
    unsigned int tr_used;
    long double* a_x1;
    long double* a_y1;
    long double* a_z1;
    double* prev_x;
    double* prev_y;
    double* prev_z;
    /* per-particle stream for the Philox mode (valid within one step) */
    or_pstream* cur_stream;
};

static __thread char g_err[256];
const char* or_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof(g_err), "%s", m); }

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy
   as 1, 2, 3"); round multipliers and Weyl key increments as published. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    int r;
    for (r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void pstream_init(or_pstream* s, uint64_t seed, uint32_t ref_idx, uint64_t step)
{
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->ctr[0] = ref_idx;
    s->ctr[1] = (uint32_t)step;
    s->ctr[2] = (uint32_t)(step >> 32);
    s->ctr[3] = 0;
    s->pos = 4;
}

static uint32_t pstream_next(or_pstream* s)
{
    if (s->pos == 4) {
        or_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[3]++;
        s->pos = 0;
    }
    return s->buf[s->pos++];
}

/* ------------------------------------------------------------------ */
/* WELL512 (prng.c:4-28) */
void or_init_prng(or_prng* state, const uint32_t seed[16])
{
    memcpy(state->state, seed, 16 * sizeof(uint32_t));
    state->index = 0;
}

long double or_rnd(or_prng* st, long double maximum)
{
    uint32_t* s = st->state;
    st->a = s[st->index];
    st->c = s[(st->index + 13) & 15];
    st->b = st->a ^ st->c ^ (st->a << 16) ^ (st->c << 15);
    st->c = s[(st->index + 9) & 15];
    st->c ^= (st->c >> 11);
    st->a = s[st->index];
    s[st->index] = st->b ^ st->c;
    st->d = st->a ^ ((st->a << 5) & 0xda442d24UL);
    st->index = (st->index + 15) & 15;
    st->a = s[st->index];
    s[st->index] = st->a ^ st->b ^ st->d ^ (st->a << 2) ^ (st->b << 18) ^ (st->c << 28);
    return maximum * (s[st->index] / (long double)UINT32_MAX);
}

/* uniform draw in [0, maximum] from whichever stream is active */
static long double draw(or_rng* g, long double maximum)
{
    if (g->kind == 0)
        return or_rnd(g->well, maximum);
    if (g->kind == 1) {
        uint32_t u = pstream_next(&g->ps);
        return (long double)((double)maximum * ((double)u / 4294967295.0));
    }
    if (g->used >= g->n_u) {
        g->exhausted = 1;
        return 0.5L * maximum;
    }
    return (long double)((double)maximum * g->u[g->used++]);
}

/* nrv_pair, prng.c:32-47 (polar Box-Muller with a zero-retry) */
static void nrv_pair_g(or_rng* g, double sd, double* nrv1, double* nrv2)
{
    double r;
    double theta;
    long double urv1;
    long double urv2 = draw(g, 1.0L);
    do {
        urv1 = draw(g, 1.0L);
    } while (urv1 == 0.0 && !g->exhausted);
    r = sqrt(-2.0 * (double)logl(urv1));
    theta = 2.0 * PI * urv2;
    *nrv1 = sd * r * cos(theta);
    *nrv2 = sd * r * sin(theta);
}

void or_nrv_pair(or_prng* st, double sd, double* n1, double* n2)
{
    or_rng g;
    memset(&g, 0, sizeof(g));
    g.kind = 0;
    g.well = st;
    nrv_pair_g(&g, sd, n1, n2);
}

/* prng.c:138-154 */
double or_u_no_flow_icdf(double chi, double q)
{
    double omc = 1.0 - chi;
    double p;
    double u = 1.29951;
    double usq;
    int it = 0;
    do {
        usq = u * u;
        p = 1.0 - ((usq + omc) * exp(-usq) / omc) - q;
        u -= p / (2.0 * u * exp(-usq) * (chi - usq) / -omc);
    } while ((p * p) > 1e-16 && ++it < 100000);
    return u;
}

/* prng.c:167-214 */
double or_u_icdf(double U, double chi, double q)
{
    const double Usq = U * U;
    const double psi = SQRT_PI * (1.0 + (2.0 * (Usq - chi))) / 2.0;
    const double p_den = 2.0 * ((U * exp(-Usq)) + (psi * erf(U)));
    double dp_den, dp_num, p, p_num, temp, u, usq;
    int it = 0;

    if (U > 2.1)
        u = U + (1.0 / U);
    else
        u = 1.2586147 + (0.221 * U) + (0.1903312 * U * U);
    dp_den = 2.0 * SQRT_PI * (U + (psi * exp(Usq) * erf(U)));
    do {
        usq = u * u;
        temp = -(U + u) * (U + u);
        p_num = (U - u) * exp(temp);
        p_num += (U + u) * exp((4.0 * U * u) + temp);
        p_num += psi * (erf(U - u) + erf(U + u));
        p = 1.0 - (p_num / p_den) - q;
        dp_num = (SQRT_PI * (1.0 + (2.0 * (Usq - usq)))) - (2.0 * psi);
        dp_num *= exp(-u * ((2.0 * U) + u)) - exp((2.0 * U * u) - usq);
        u -= p / (dp_num / dp_den);
    } while ((p * p) > 1e-16 && ++it < 100000);
    return u;
}

/* pot.c:433-492: crudely adaptive Simpson for Hutchinson's alpha */
double or_integrate_for_alpha(double s, double chi_b, double u2, double Rol, double max_lin_dev)
{
    const double chu2 = chi_b / u2;
    const double s2 = s * s;
    const double y_one = s / sqrt(1.0 - chu2 - s2);
    double alpha = 0.0;
    double xi[3] = {0.0, 0.5, 1.0};
    double y[3] = {s, -9.9, y_one};
    long guard = 0;

    if (s == 0.0)
        return 0.0;
    do {
        xi[2] = 1.0;
        y[2] = y_one;
        xi[1] = (xi[0] + xi[2]) / 2.0;
        y[1] = 1.0 - (chu2 * xi[1] * exp(Rol * (1.0 - (1.0 / xi[1]))));
        y[1] = 1.0 / sqrt((y[1] / s2) - (xi[1] * xi[1]));
        while (fabs(y[2] - y[1] - (y[1] - y[0])) > max_lin_dev) {
            xi[2] = xi[1];
            y[2] = y[1];
            xi[1] = (xi[0] + xi[2]) / 2.0;
            y[1] = 1.0 - (chu2 * xi[1] * exp(Rol * (1.0 - (1.0 / xi[1]))));
            y[1] = 1.0 / sqrt((y[1] / s2) - (xi[1] * xi[1]));
            if (++guard > 100000000L)
                return NAN;
        }
        alpha += (xi[2] - xi[0]) * (y[0] + (4.0 * y[1]) + y[2]) / 6.0;
        xi[0] = xi[2];
        y[0] = y[2];
    } while (xi[2] < 1.0);
    return alpha;
}

/* coord.c:6-14 */
static void velocity_in_cartesian_coords(const double* v, double phi, double theta,
                                         double* vx, double* vy, double* vz)
{
    *vx = (v[0] * cos(phi) * sin(theta)) - (v[1] * sin(phi)) + (v[2] * cos(phi) * cos(theta));
    *vy = (v[0] * sin(phi) * sin(theta)) + (v[1] * cos(phi)) + (v[2] * sin(phi) * cos(theta));
    *vz = (v[0] * cos(theta)) - (v[2] * sin(theta));
}

/* coord.c:21-35 */
static void rotate_vec_about_xy_dir(const double* d, double s, double c, double* x, double* y,
                                    double* z, double offset)
{
    const double omc = 1.0 - c;
    double t0, t1, t2;
    t0 = d[0] * ((d[0] * *x) + (d[1] * *y)) * omc;
    t0 += (*x * c) + (d[1] * *z * s);
    t1 = d[1] * ((d[0] * *x) + (d[1] * *y)) * omc;
    t1 += (*y * c) - (d[0] * *z * s);
    t2 = *z * c;
    t2 += ((d[0] * *y) - (d[1] * *x)) * s;
    *x = offset + t0;
    *y = offset + t1;
    *z = offset + t2;
}

/* coord.c:45-59 */
static void rotate_vec_about_dir(const double* d, double s, double c, double* x, double* y,
                                 double* z, double offset)
{
    const double omc = 1.0 - c;
    double t0, t1, t2;
    t0 = d[0] * ((d[0] * *x) + (d[1] * *y) + (d[2] * *z)) * omc;
    t0 += (*x * c) + (((d[1] * *z) - (d[2] * *y)) * s);
    t1 = d[1] * ((d[0] * *x) + (d[1] * *y) + (d[2] * *z)) * omc;
    t1 += (*y * c) + (((d[2] * *x) - (d[0] * *z)) * s);
    t2 = d[2] * ((d[0] * *x) + (d[1] * *y) + (d[2] * *z)) * omc;
    t2 += (*z * c) + (((d[0] * *y) - (d[1] * *x)) * s);
    *x = offset + t0;
    *y = offset + t1;
    *z = offset + t2;
}

/* ------------------------------------------------------------------ */
/* SCEPTIC-like reinjection, pot.c:496-652.  The reference retries by
   recursion when alpha is NaN, first undoing the electron renormalisation
   of chi_b (pot.c:591-598); the loop below keeps that exact sequence. */
static void sceptically_reinject_g(const or_sim* s, or_particle* part, or_rng* g, double chi_b_in)
{
    const or_config* ru = &s->ru;
    double a[2], alpha, b, c, co_z, phi, psi, q, Rol, U, u_mag, u2;
    double u[3], v[3], v_n;
    double chi_b = chi_b_in;
    int tries = 0;

    for (;;) {
        v_n = 2.0 * KB / s->mass[part->species];
        if (part->species == OR_ELECTRON) {
            v_n *= ru->Te;
            chi_b *= -ru->Ti / ru->Te;
        } else {
            v_n *= ru->Ti;
        }
        v_n = sqrt(v_n);

        if (ru->x_drift) {
            U = ru->x_drift / v_n;
            do {
                u_mag = or_u_icdf(U, chi_b, draw(g, 1.0));
                u2 = u_mag * u_mag;
            } while (u2 < chi_b && !g->exhausted);
            q = draw(g, 1.0);
            c = log(q + (1.0 - q) * exp(-4.0 * u_mag * U)) / (2.0 * u_mag * U);
            c += 1.0;
        } else {
            do {
                u_mag = or_u_no_flow_icdf(chi_b, draw(g, 1.0));
                u2 = u_mag * u_mag;
            } while (u2 < chi_b && !g->exhausted);
            c = draw(g, 2.0) - 1.0;
        }

        u[0] = 0.0;
        u[1] = 0.0;
        u[2] = u_mag;
        a[0] = 0.0;
        a[1] = 1.0;
        rotate_vec_about_xy_dir(a, c, sqrt(1.0 - (c * c)), &u[0], &u[1], &u[2], 0.0);
        a[0] = 1.0;
        a[1] = 0.0;
        psi = draw(g, 2.0 * PI);
        rotate_vec_about_xy_dir(a, sin(psi), cos(psi), &u[0], &u[1], &u[2], 0.0);

        b = sqrt(draw(g, 1.0 - (chi_b / u2)));

        Rol = QE * QE * (ru->N - 1) / 2.0;
        Rol /= (4.0 * PI / 3.0) * ru->mpp * ru->r * EPSILON0 * KB * ru->Te;
        Rol = sqrt(Rol);
        alpha = or_integrate_for_alpha(b, chi_b, u2, Rol, ru->max_lin_dev);
        if (isnan(alpha)) {
            if (part->species == OR_ELECTRON)
                chi_b /= -ru->Ti / ru->Te;
            if (++tries > 1000000 || g->exhausted)
                return;
            continue;
        }
        break;
    }

    phi = draw(g, 2.0 * PI);
    part->x = ru->mpp * ru->r * sin(alpha) * cos(phi);
    part->y = ru->mpp * ru->r * sin(alpha) * sin(phi);
    part->z = ru->mpp * ru->r * cos(alpha);

    a[0] = u[1];
    a[1] = -u[0];
    a[0] /= sqrt((u[0] * u[0]) + (u[1] * u[1]));
    a[1] /= sqrt((u[0] * u[0]) + (u[1] * u[1]));

    co_z = u[2] / u_mag;
    rotate_vec_about_xy_dir(a, sin(acos(co_z)), co_z, &part->x, &part->y, &part->z,
                            ru->mpp * ru->r);

    v[1] = 0.0;
    v[2] = u_mag * b;
    v[0] = -sqrt(u2 - chi_b - (v[2] * v[2]));
    velocity_in_cartesian_coords(v, phi, alpha, &part->vx, &part->vy, &part->vz);
    rotate_vec_about_xy_dir(a, sin(acos(co_z)), co_z, &part->vx, &part->vy, &part->vz, 0.0);

    part->vx *= v_n;
    part->vy *= v_n;
    part->vz *= v_n;
}

/* inject, pot.c:658-846, spherical branch (the configurations of this
   build are all spherical domains) */
static void inject_g(const or_sim* s, or_particle* part, int reinjection, or_rng* g)
{
    const or_config* ru = &s->ru;
    double sigma, phi, r, theta;
    double v[3];

    if (part->species == OR_ION)
        sigma = sqrt(KB * ru->Ti / s->mass[OR_ION]);
    else
        sigma = sqrt(KB * ru->Te / s->mass[OR_ELECTRON]);
    if (!ru->sceptic && !reinjection)
        sigma *= 0.85;

    if (!reinjection) {
        /* `ru.r * ru.r * ru.r` is unsigned int arithmetic (pot.c:703) */
        unsigned int r3 = ru->r * ru->r * ru->r;
        r = ru->mpp * cbrt(draw(g, r3));
        if (ru->grain_enabled || ru->geometry != OR_SPHERE) {
            while (r <= ru->a && !g->exhausted)
                r = ru->mpp * cbrt(draw(g, r3));
        }
    } else {
        r = ru->mpp * ru->r;
    }

    if (ru->sceptic && reinjection) {
        sceptically_reinject_g(s, part, g, s->chi_b);
        return;
    }
    phi = draw(g, 2.0 * PI);
    theta = acos(draw(g, 2.0) - 1.0);
    part->x = (ru->mpp * ru->r) + (r * sin(theta) * cos(phi));
    part->y = (ru->mpp * ru->r) + (r * sin(theta) * sin(phi));
    part->z = (ru->mpp * ru->r) + (r * cos(theta));

    if (!reinjection) {
        nrv_pair_g(g, sigma, &part->vx, &part->vy);
        nrv_pair_g(g, sigma, &part->vz, &part->vz);
    } else {
        nrv_pair_g(g, sigma, &v[0], &v[1]);
        nrv_pair_g(g, sigma, &v[2], &v[2]);
        if (v[0] > 0.0)
            v[0] = -v[0];
        velocity_in_cartesian_coords(v, phi, theta, &part->vx, &part->vy, &part->vz);
    }
    part->vx += ru->x_drift;
}

int or_sceptic_from_uniforms(const or_sim* s, or_particle* p, double chi_b, const double* u,
                             int n_u)
{
    or_rng g;
    memset(&g, 0, sizeof(g));
    g.kind = 2;
    g.u = u;
    g.n_u = n_u;
    sceptically_reinject_g(s, p, &g, chi_b);
    return g.exhausted ? -1 : g.used;
}

/* ------------------------------------------------------------------ */
/* command-line parsing helpers: -d is read with strtold (pot.c:4611);
   -E, -I and -e are read with strtold into double fields (pot.c:4625,
   4639, 4679) */
void or_config_set_dt(or_config* cfg, const char* text) { cfg->dt = strtold(text, NULL); }
double or_strtold_to_double(const char* text) { return (double)strtold(text, NULL); }

or_sim* or_create(const or_config* cfg)
{
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    size_t N = (size_t)cfg->N;
    if (!s) {
        set_err("out of memory");
        return NULL;
    }
    s->ru = *cfg;
    /* pot.c:292: MASS[] = { 1.67e-12, MIME * ME, ME } */
    s->mass[OR_DUST] = 1.67e-12;
    s->mass[OR_ION] = cfg->mime * ME;
    s->mass[OR_ELECTRON] = ME;
    /* pot.c:4157-4159 */
    s->tss[OR_ELECTRON] = 1;
    s->tss[OR_ION] = (int)sqrt(cfg->Te * s->mass[OR_ION] / (cfg->Ti * s->mass[OR_ELECTRON]));
    s->tss[OR_DUST] = s->tss[OR_ION] * (int)sqrt(0.1 * s->mass[OR_DUST] / s->mass[OR_ION]);
    s->parts = (or_particle*)calloc(N, sizeof(or_particle));
    s->new_parts = (or_particle*)calloc(N, sizeof(or_particle));
    s->tree_parts = (or_particle*)calloc(N, sizeof(or_particle));
    s->tr_alloc = (unsigned int)(9 * N + 9);
    s->root = (or_node*)malloc(sizeof(or_node) * s->tr_alloc);
    s->a_x1 = (long double*)malloc(3 * sizeof(long double) * N);
    s->prev_x = (double*)malloc(3 * sizeof(double) * N);
    if (!s->parts || !s->new_parts || !s->tree_parts || !s->root || !s->a_x1 || !s->prev_x) {
        or_destroy(s);
        set_err("out of memory");
        return NULL;
    }
    s->a_y1 = s->a_x1 + N;
    s->a_z1 = s->a_x1 + 2 * N;
    s->prev_y = s->prev_x + N;
    s->prev_z = s->prev_x + 2 * N;
    s->dust_grain_q = CHARGE[OR_DUST];
    s->dust_grain_m = s->mass[OR_DUST];
    s->chi_b = 0.0;
    return s;
}

void or_destroy(or_sim* s)
{
    if (!s)
        return;
    free(s->parts);
    free(s->new_parts);
    free(s->tree_parts);
    free(s->root);
    free(s->a_x1);
    free(s->prev_x);
    free(s);
}

void or_get_particles(const or_sim* s, or_particle* out)
{
    memcpy(out, s->parts, sizeof(or_particle) * s->ru.N);
}
void or_set_particles(or_sim* s, const or_particle* in)
{
    memcpy(s->parts, in, sizeof(or_particle) * s->ru.N);
}
or_particle* or_particles_ptr(or_sim* s) { return s->parts; }

void or_get_state(const or_sim* s, int32_t* q, long double* m, uint64_t* t, double* chi_b,
                  uint32_t* tot_e, long double epe[3])
{
    if (q) *q = s->dust_grain_q;
    if (m) *m = s->dust_grain_m;
    if (t) *t = s->t_steps;
    if (chi_b) *chi_b = s->chi_b;
    if (tot_e) *tot_e = s->tot_absor_e;
    if (epe) {
        epe[0] = s->epe[0];
        epe[1] = s->epe[1];
        epe[2] = s->epe[2];
    }
}

void or_set_state(or_sim* s, int32_t q, long double m, uint64_t t, double chi_b, uint32_t tot_e)
{
    s->dust_grain_q = q;
    s->dust_grain_m = m;
    s->t_steps = (unsigned long)t;
    s->chi_b = chi_b;
    s->tot_absor_e = tot_e;
}

void or_get_prng(const or_sim* s, or_prng* out) { *out = s->ps; }
void or_set_prng(or_sim* s, const or_prng* in) { s->ps = *in; }
void or_get_tss(const or_sim* s, uint32_t tss[3])
{
    tss[0] = s->tss[0];
    tss[1] = s->tss[1];
    tss[2] = s->tss[2];
}

/* init_particles, pot.c:887-940, for one worker process */
void or_init_particles(or_sim* s, const uint32_t seed[16])
{
    const or_config* ru = &s->ru;
    or_particle* parts = s->parts;
    unsigned long p_idx;
    unsigned int N_i;
    or_rng g;

    or_init_prng(&s->ps, seed);
    memset(&g, 0, sizeof(g));
    g.kind = 0;
    g.well = &s->ps;

    parts[0].x = ru->r * ru->mpp;
    parts[0].y = parts[0].x;
    parts[0].z = parts[0].x;
    parts[0].vx = 0.0;
    parts[0].vy = 0.0;
    parts[0].vz = 0.0;
    parts[0].species = OR_DUST;
    for (p_idx = 1; p_idx < ru->N; p_idx++)
        parts[p_idx].species = OR_ELECTRON;
    /* with a single worker the ions occupy indices 1..(N-1)/2 */
    N_i = (unsigned int)((ru->N - 1) / 2);
    for (p_idx = 1; p_idx <= N_i; p_idx++)
        parts[p_idx].species = OR_ION;
    for (p_idx = 1; p_idx < ru->N; p_idx++)
        inject_g(s, &parts[p_idx], 0, &g);
}

/* ------------------------------------------------------------------ */
/* init_node_children, pot.c:1225-1288 */
static void init_node_children(or_node* list, unsigned int node_no)
{
    or_node* tr = &list[node_no];
    or_node* ch[8];
    double mid_x, mid_y, mid_z;
    int i, k;
    for (i = 0; i < 8; i++)
        ch[i] = &list[tr->ch[i]];
    for (i = 0; i < 8; i++) {
        ch[i]->x1 = tr->x1;
        ch[i]->x2 = tr->x2;
        ch[i]->y1 = tr->y1;
        ch[i]->y2 = tr->y2;
        ch[i]->z1 = tr->z1;
        ch[i]->z2 = tr->z2;
        ch[i]->part = UINT_MAX;
        for (k = 0; k < 8; k++)
            ch[i]->ch[k] = 0;
    }
    mid_x = (tr->x1 + tr->x2) / 2;
    mid_y = (tr->y1 + tr->y2) / 2;
    mid_z = (tr->z1 + tr->z2) / 2;
    /* g = greater half (lower face moves to mid), l = lesser half */
    ch[GXLYLZ]->x1 = mid_x; ch[GXLYLZ]->y2 = mid_y; ch[GXLYLZ]->z2 = mid_z;
    ch[GXLYGZ]->x1 = mid_x; ch[GXLYGZ]->y2 = mid_y; ch[GXLYGZ]->z1 = mid_z;
    ch[LXLYLZ]->x2 = mid_x; ch[LXLYLZ]->y2 = mid_y; ch[LXLYLZ]->z2 = mid_z;
    ch[LXLYGZ]->x2 = mid_x; ch[LXLYGZ]->y2 = mid_y; ch[LXLYGZ]->z1 = mid_z;
    ch[LXGYLZ]->x2 = mid_x; ch[LXGYLZ]->y1 = mid_y; ch[LXGYLZ]->z2 = mid_z;
    ch[LXGYGZ]->x2 = mid_x; ch[LXGYGZ]->y1 = mid_y; ch[LXGYGZ]->z1 = mid_z;
    ch[GXGYLZ]->x1 = mid_x; ch[GXGYLZ]->y1 = mid_y; ch[GXGYLZ]->z2 = mid_z;
    ch[GXGYGZ]->x1 = mid_x; ch[GXGYGZ]->y1 = mid_y; ch[GXGYGZ]->z1 = mid_z;
}

/* child slot for a position against a node's midpoints (pot.c:1525-1553) */
static int octant(const or_particle* p, double mid_x, double mid_y, double mid_z)
{
    if (p->x < mid_x) {
        if (p->y < mid_y)
            return (p->z < mid_z) ? LXLYLZ : LXLYGZ;
        return (p->z < mid_z) ? LXGYLZ : LXGYGZ;
    }
    if (p->y < mid_y)
        return (p->z < mid_z) ? GXLYLZ : GXLYGZ;
    return (p->z < mid_z) ? GXGYLZ : GXGYGZ;
}

/* update_node_moments, pot.c:1290-1310 (the dipole lines are DIPOLES) */
static void update_node_moments(or_node* node, const or_node* chi, int dipoles)
{
    const double old_cx = node->cx, old_cy = node->cy, old_cz = node->cz;
    node->cx = (node->aq * node->cx + chi->aq * chi->cx) / (node->aq + chi->aq);
    node->cy = (node->aq * node->cy + chi->aq * chi->cy) / (node->aq + chi->aq);
    node->cz = (node->aq * node->cz + chi->aq * chi->cz) / (node->aq + chi->aq);
    if (dipoles) {
        node->px -= (node->cx - old_cx) * node->q;
        node->py -= (node->cy - old_cy) * node->q;
        node->pz -= (node->cz - old_cz) * node->q;
        node->px += chi->px - (node->cx - chi->cx) * chi->q;
        node->py += chi->py - (node->cy - chi->cy) * chi->q;
        node->pz += chi->pz - (node->cz - chi->cz) * chi->q;
    }
    node->q += chi->q;
    node->aq += chi->aq;
}

/* compute_node_values, pot.c:1312-1402 */
static void compute_node_values(or_sim* s, unsigned int node_no)
{
    or_node* node = &s->root[node_no];
    int k;
    if (node->part != UINT_MAX) {
        const or_particle* p = &s->tree_parts[node->part];
        node->q = (p->species == OR_DUST) ? s->dust_grain_q : CHARGE[p->species];
        node->aq = abs(node->q);
        node->cx = p->x;
        node->cy = p->y;
        node->cz = p->z;
        node->px = 0.0;
        node->py = 0.0;
        node->pz = 0.0;
    } else {
        node->q = 0;
        node->aq = 0;
        node->cx = 0.0;
        node->cy = 0.0;
        node->cz = 0.0;
        node->px = 0.0;
        node->py = 0.0;
        node->pz = 0.0;
        for (k = 0; k < 8; k++) {
            unsigned int c = s->root[node_no].ch[k];
            if (c) {
                compute_node_values(s, c);
                node = &s->root[node_no];
                if (s->root[c].aq)
                    update_node_moments(node, &s->root[c], s->ru.dipoles);
            }
        }
    }
}

/* build_tree, pot.c:1457-1643 (SPH:1516-1521 leaves the grain out) */
int or_build_tree(or_sim* s)
{
    const or_config* ru = &s->ru;
    or_node* root = s->root;
    or_node* tr;
    unsigned int cur, next, used = 0, part_no, k;
    double mid_x, mid_y, mid_z;
    const or_particle* part;
    long guard;

    memcpy(s->tree_parts, s->parts, sizeof(or_particle) * ru->N);

    root->x1 = 0.0;
    root->y1 = 0.0;
    root->z1 = 0.0;
    root->x2 = 2.0 * ru->r * ru->mpp;
    root->y2 = root->x2;
    root->z2 = root->y2;
    root->part = UINT_MAX;
    root->cx = 0.0;
    root->cy = 0.0;
    root->cz = 0.0;
    for (k = 0; k < 8; k++)
        root->ch[k] = 0;

    part_no = (ru->geometry != OR_SPHERE) ? 1 : 0;
    guard = 0;
    while (part_no < ru->N) {
        part = &s->tree_parts[part_no];
        if ((part->x < 0.0) || (part->x > root->x2) || (part->y < 0.0) ||
            (part->y > root->y2) || (part->z < 0.0) || (part->z > root->z2) || isnan(part->x) ||
            isnan(part->y) || isnan(part->z)) {
            snprintf(g_err, sizeof(g_err), "particle %u at (%g, %g, %g) not in simulation area",
                     part_no, part->x, part->y, part->z);
            return -1;
        }
        next = 0;
        do {
            cur = next;
            tr = &s->root[cur];
            mid_x = (tr->x1 + tr->x2) / 2.0;
            mid_y = (tr->y1 + tr->y2) / 2.0;
            mid_z = (tr->z1 + tr->z2) / 2.0;
            next = tr->ch[octant(part, mid_x, mid_y, mid_z)];
        } while (next);

        if (tr->part == UINT_MAX) {
            tr->part = part_no;
            part_no++;
            guard = 0;
        } else {
            const or_particle* old = &s->tree_parts[tr->part];
            if (old->x == part->x && old->y == part->y && old->z == part->z) {
                snprintf(g_err, sizeof(g_err), "particles %u & %u both have the position (%g, %g, %g)",
                         part_no, tr->part, part->x, part->y, part->z);
                return -2;
            }
            if (s->tr_alloc < (used + 9)) {
                unsigned int na = (unsigned int)(s->tr_alloc * 1.1) + 9;
                or_node* nr = (or_node*)realloc(s->root, sizeof(or_node) * na);
                if (!nr) {
                    set_err("cannot grow tree");
                    return -3;
                }
                s->root = nr;
                s->tr_alloc = na;
                root = s->root;
                tr = &s->root[cur];
            }
            for (k = 0; k < 8; k++)
                tr->ch[k] = ++used;
            init_node_children(s->root, cur);
            tr = &s->root[cur];
            s->root[tr->ch[octant(&s->tree_parts[tr->part], mid_x, mid_y, mid_z)]].part = tr->part;
            tr->part = UINT_MAX;
            if (++guard > 4000) {
                set_err("tree depth exceeded");
                return -2;
            }
        }
    }
    s->tr_used = used;
    compute_node_values(s, 0);
    return 0;
}

uint32_t or_tree_used(const or_sim* s) { return s->tr_used; }

int64_t or_tree_canon_dipoles(const or_sim* s, double* p, int64_t cap)
{
    unsigned int stack[4096];
    int sp = 0;
    int64_t n = 0;
    if (s->root[0].part == UINT_MAX && s->root[0].ch[0] == 0)
        return 0;
    stack[sp++] = 0;
    while (sp) {
        const or_node* nd = &s->root[stack[--sp]];
        int k;
        if (n >= cap)
            return -1;
        p[3 * n] = nd->px;
        p[3 * n + 1] = nd->py;
        p[3 * n + 2] = nd->pz;
        n++;
        if (nd->part == UINT_MAX && nd->ch[0])
            for (k = 7; k >= 0; k--) {
                const or_node* c = &s->root[nd->ch[k]];
                if (!(c->part == UINT_MAX && c->ch[0] == 0)) {
                    if (sp >= 4096)
                        return -2;
                    stack[sp++] = nd->ch[k];
                }
            }
    }
    return n;
}

static int node_empty(const or_node* n) { return n->part == UINT_MAX && n->ch[0] == 0; }

/* canonical listing: DFS pre-order over the non-empty nodes, children in
   reference order; leaves report their particle index */
int64_t or_tree_canon(const or_sim* s, int32_t* depth, int32_t* is_leaf, uint32_t* part,
                      int32_t* q, uint32_t* aq, double* c, double* x1x2, int32_t* nchild,
                      int64_t cap)
{
    unsigned int stack[4096];
    int32_t dstack[4096];
    int sp = 0;
    int64_t n = 0;
    stack[sp] = 0;
    dstack[sp++] = 0;
    if (node_empty(&s->root[0]))
        return 0;
    while (sp) {
        unsigned int id = stack[--sp];
        int32_t d = dstack[sp];
        const or_node* nd = &s->root[id];
        int k, cnt = 0;
        if (n >= cap)
            return -1;
        depth[n] = d;
        is_leaf[n] = nd->part != UINT_MAX;
        part[n] = nd->part;
        q[n] = nd->q;
        aq[n] = nd->aq;
        c[3 * n] = nd->cx;
        c[3 * n + 1] = nd->cy;
        c[3 * n + 2] = nd->cz;
        x1x2[2 * n] = nd->x1;
        x1x2[2 * n + 1] = nd->x2;
        if (nd->part == UINT_MAX && nd->ch[0]) {
            for (k = 7; k >= 0; k--) {
                if (!node_empty(&s->root[nd->ch[k]])) {
                    if (sp >= 4096)
                        return -2;
                    stack[sp] = nd->ch[k];
                    dstack[sp++] = d + 1;
                    cnt++;
                }
            }
        }
        nchild[n] = cnt;
        n++;
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* field_particle_feels_from_grain, SPH:1689-1780 (spheroid, axis along x) */
static void grain_field_ld(const or_sim* s, unsigned int part, double x, double y, double z,
                           long double* E)
{
    const or_config* ru = &s->ru;
    const or_particle* p = &s->parts[0];
    double rho2, rt, E_n, beta, theta, dx, dy, dz, c, phi_const;
    if (ru->geometry == OR_PROLATE) {
        c = (ru->a * sqrt(ru->A * ru->A - 1.0)) * pow(ru->A, -1.0 / 3.0);
        phi_const = (QE * s->dust_grain_q * acosh(ru->A)) /
                    (4.0 * PI * EPSILON0 * c * log(ru->A - sqrt(ru->A * ru->A - 1.0)));
    } else {
        c = (ru->a * sqrt(1.0 - ru->A * ru->A)) * pow(ru->A, -1.0 / 3.0);
        phi_const = (QE * s->dust_grain_q * acos(ru->A)) /
                    (8.0 * PI * EPSILON0 * c *
                     (atan((1.0 - sqrt(1.0 - ru->A * ru->A)) / ru->A) - 0.25 * PI));
    }
    if (!part) {
        E[0] = 0.0;
        E[1] = 0.0;
        E[2] = 0.0;
        return;
    }
    dx = x - p->x;
    dy = y - p->y;
    dz = z - p->z;
    rho2 = (dy * dy) + (dz * dz);
    if (ru->geometry == OR_PROLATE) {
        double sinh2;
        rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx +
                  2.0 * rho2 * c * c - 2.0 * dx * dx * c * c);
        sinh2 = (rho2 + dx * dx - c * c + rt) / (2.0 * c * c);
        E_n = -phi_const / sqrt(sinh2 * rt);
        beta = atan((-sinh2 * dx) / ((sinh2 + 1.0) * sqrt(rho2)));
    } else {
        double cosh2;
        rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx -
                  2.0 * rho2 * c * c + 2.0 * dx * dx * c * c);
        cosh2 = (rho2 + dx * dx + c * c + rt) / (2.0 * c * c);
        E_n = -phi_const / sqrt(cosh2 * rt);
        beta = atan((-cosh2 * dx) / ((cosh2 - 1.0) * sqrt(rho2)));
    }
    theta = atan(fabs(dz) / fabs(dy));
    E[0] = -E_n * sin(beta);
    E[1] = E_n * cos(beta) * cos(theta) * fabs(dy) / dy;
    E[2] = E_n * cos(beta) * sin(theta) * fabs(dz) / dz;
}

void or_field_from_grain(const or_sim* s, uint32_t part, double x, double y, double z,
                         double E[3])
{
    long double e[3] = {0, 0, 0};
    if (s->ru.geometry != OR_SPHERE)
        grain_field_ld(s, part, x, y, z, e);
    E[0] = (double)e[0];
    E[1] = (double)e[1];
    E[2] = (double)e[2];
}

/* grain potential at absolute coordinates (SPH phi_at, SPH:1924-1953) */
static double grain_phi_abs(const or_sim* s, double x, double y, double z)
{
    const or_config* ru = &s->ru;
    double c, phi_const, rho2, rt;
    if (ru->geometry == OR_PROLATE) {
        c = (ru->a * sqrt(ru->A * ru->A - 1.0)) / pow(ru->A, 1.0 / 3.0);
        phi_const = (QE * s->dust_grain_q * acosh(ru->A)) /
                    (4.0 * PI * EPSILON0 * c * log(ru->A - sqrt(ru->A * ru->A - 1.0)));
        rho2 = (y * y) + (z * z);
        rt = sqrt(rho2 * rho2 + x * x * x * x + c * c * c * c + 2.0 * rho2 * x * x +
                  2.0 * rho2 * c * c - 2.0 * x * x * c * c);
        return phi_const *
               log((sqrt(rho2 + x * x + c * c + rt) - c * sqrt(2.0)) / sqrt(rho2 + x * x - c * c + rt));
    }
    c = (ru->a * sqrt(1.0 - ru->A * ru->A)) * pow(ru->A, -1.0 / 3.0);
    phi_const = (QE * s->dust_grain_q * acos(ru->A)) /
                (4.0 * PI * EPSILON0 * c *
                 (atan((1.0 - sqrt(1.0 - ru->A * ru->A)) / ru->A) - 0.25 * PI));
    rho2 = (y * y) + (z * z);
    rt = sqrt(rho2 * rho2 + x * x * x * x + c * c * c * c + 2.0 * rho2 * x * x -
              2.0 * rho2 * c * c + 2.0 * x * x * c * c);
    return phi_const * (atan((sqrt(rho2 + x * x + c * c + rt) - c * sqrt(2.0)) /
                             sqrt(rho2 + x * x - c * c + rt)) -
                        0.25 * PI);
}

/* phi_particle_feels_from_grain, SPH:2040-2084 (grain-relative) */
double or_phi_from_grain(const or_sim* s, uint32_t part)
{
    const or_config* ru = &s->ru;
    const or_particle* p = &s->parts[0];
    double dx, dy, dz;
    if (!part || ru->geometry == OR_SPHERE)
        return 0.0;
    dx = s->parts[part].x - p->x;
    dy = s->parts[part].y - p->y;
    dz = s->parts[part].z - p->z;
    if (ru->geometry == OR_PROLATE) {
        double c = ru->a * sqrt(ru->A * ru->A - 1.0) / pow(ru->A, 1.0 / 3.0);
        double phi_const = (QE * s->dust_grain_q * acosh(ru->A)) /
                           (4.0 * PI * EPSILON0 * c * log(ru->A - sqrt(ru->A * ru->A - 1.0)));
        double rho2 = (dy * dy) + (dz * dz);
        double rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx +
                         2.0 * rho2 * c * c - 2.0 * dx * dx * c * c);
        return phi_const * log((sqrt(rho2 + dx * dx + c * c + rt) - c * sqrt(2.0)) /
                               sqrt(rho2 + dx * dx - c * c + rt));
    } else {
        double c = (ru->a * sqrt(1.0 - ru->A * ru->A)) * pow(ru->A, -1.0 / 3.0);
        double phi_const = (QE * s->dust_grain_q * acos(ru->A)) /
                           (4.0 * PI * EPSILON0 * c *
                            (atan((1.0 - sqrt(1.0 - ru->A * ru->A)) / ru->A) - 0.25 * PI));
        double rho2 = (dy * dy) + (dz * dz);
        double rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx -
                         2.0 * rho2 * c * c + 2.0 * dx * dx * c * c);
        return phi_const * (atan((sqrt(rho2 + dx * dx + c * c + rt) - c * sqrt(2.0)) /
                                 sqrt(rho2 + dx * dx - c * c + rt)) -
                            0.25 * PI);
    }
}

/* ------------------------------------------------------------------ */
/* field_particle_feels_at, pot.c:1696-1808 (SPH:1794-1921).  `src` holds
   the positions leaves read: the live array in the reference's in-place
   semantics, the build snapshot otherwise. */
#define FIELD_WALK_BODY(ACC)                                                              \
    do {                                                                                  \
        unsigned int path[MAX_PATH_LEN];                                                  \
        unsigned int path_len = 1;                                                        \
        path[0] = 0;                                                                      \
        do {                                                                              \
            const or_node* node = &root[path[path_len - 1]];                              \
            path_len--;                                                                   \
            if (node->part == UINT_MAX) {                                                 \
                dx = x - node->cx;                                                        \
                dy = y - node->cy;                                                        \
                dz = z - node->cz;                                                        \
                dist2 = (dx * dx) + (dy * dy) + (dz * dz);                                \
                if (node->ch[0]) st_mac++;                                    \
                if (((node->x2 - node->x1) * (node->x2 - node->x1)) < (oap2 * dist2)) {   \
                    if (ru->dipoles ? (node->aq != 0) : (node->q != 0)) {                 \
                        dist = sqrt(dist2);                                               \
                        E_scale = QE / (4 * PI * EPSILON0 * dist2 * dist);                \
                        if (ru->dipoles) {                                                \
                            double rpd2 = (dx * node->px) + (dy * node->py) + (dz * node->pz); \
                            rpd2 /= dist2;                                                \
                            ACC[0] += E_scale * ((node->q * dx) - node->px + (3 * dx * rpd2)); \
                            ACC[1] += E_scale * ((node->q * dy) - node->py + (3 * dy * rpd2)); \
                            ACC[2] += E_scale * ((node->q * dz) - node->pz + (3 * dz * rpd2)); \
                        } else {                                                          \
                            E_scale *= node->q;                                           \
                            ACC[0] += E_scale * dx;                                       \
                            ACC[1] += E_scale * dy;                                       \
                            ACC[2] += E_scale * dz;                                       \
                        }                                                                 \
                        st_cell++;                                                        \
                    }                                                                     \
                } else {                                                                  \
                    if ((path_len + 8) > MAX_PATH_LEN) {                                  \
                        set_err("particle tree path length exceeded");                    \
                        abort();                                                          \
                    }                                                                     \
                    if (node->ch[0]) {                                                    \
                        int k_;                                                           \
                        st_open++;                                                        \
                        for (k_ = 0; k_ < 8; k_++) path[path_len++] = node->ch[k_];       \
                    }                                                                     \
                }                                                                         \
            } else if (node->part != part) {                                              \
                const or_particle* p = &src[node->part];                                  \
                dx = x - p->x;                                                            \
                dy = y - p->y;                                                            \
                dz = z - p->z;                                                            \
                dist2 = (dx * dx) + (dy * dy) + (dz * dz);                                \
                if (dist2 != 0.0) {                                                       \
                    dist2 += ru->soft * ru->soft;                                         \
                    dist = sqrt(dist2);                                                   \
                    if (p->species == OR_DUST)                                            \
                        E_scale = QE * s->dust_grain_q / ((4 * PI * EPSILON0) * dist2 * dist); \
                    else                                                                  \
                        E_scale = QE * CHARGE[p->species] / ((4 * PI * EPSILON0) * dist2 * dist); \
                    ACC[0] += E_scale * dx;                                               \
                    ACC[1] += E_scale * dy;                                               \
                    ACC[2] += E_scale * dz;                                               \
                    st_leaf++;                                                            \
                } else {                                                                  \
                    set_err("multiple particles at one location");                        \
                    abort();                                                              \
                }                                                                         \
            }                                                                             \
        } while (path_len);                                                               \
    } while (0)

static void field_walk(or_sim* s, unsigned int part, double x, double y, double z,
                       const or_particle* src, long double* E, or_walk_stats* st)
{
    const or_config* ru = &s->ru;
    const or_node* root = s->root;
    const double oap2 = ru->theta * ru->theta;
    double dist, dist2, dx, dy, dz;
    long double E_scale;
    uint64_t st_cell = 0, st_leaf = 0, st_open = 0, st_mac = 0;

    E[0] = 0.0;
    E[1] = 0.0;
    E[2] = 0.0;
    if (ru->geometry != OR_SPHERE)
        grain_field_ld(s, part, x, y, z, E);
    if (!node_empty(&root[0]))
        FIELD_WALK_BODY(E);
    if (st) {
        st->n_cell += st_cell;
        st->n_leaf += st_leaf;
        st->n_open += st_open;
        st->n_mac += st_mac;
    }
}

/* the same walk with double accumulation (the arithmetic a GPU FP64 walk
   performs); the interaction set and order are identical */
static void field_walk_d(or_sim* s, unsigned int part, double x, double y, double z,
                         const or_particle* src, double* E, or_walk_stats* st)
{
    const or_config* ru = &s->ru;
    const or_node* root = s->root;
    const double oap2 = ru->theta * ru->theta;
    double dist, dist2, dx, dy, dz;
    double E_scale;
    uint64_t st_cell = 0, st_leaf = 0, st_open = 0, st_mac = 0;

    E[0] = 0.0;
    E[1] = 0.0;
    E[2] = 0.0;
    if (ru->geometry != OR_SPHERE) {
        long double e[3];
        grain_field_ld(s, part, x, y, z, e);
        E[0] = (double)e[0];
        E[1] = (double)e[1];
        E[2] = (double)e[2];
    }
    if (!node_empty(&root[0]))
        FIELD_WALK_BODY(E);
    if (st) {
        st->n_cell += st_cell;
        st->n_leaf += st_leaf;
        st->n_open += st_open;
        st->n_mac += st_mac;
    }
}

void or_field_at(or_sim* s, uint32_t part, double x, double y, double z, int32_t accum_double,
                 double E[3], or_walk_stats* st)
{
    const or_particle* src = s->ru.inplace ? s->parts : s->tree_parts;
    if (accum_double) {
        field_walk_d(s, part, x, y, z, src, E, st);
    } else {
        long double e[3];
        field_walk(s, part, x, y, z, src, e, st);
        E[0] = (double)e[0];
        E[1] = (double)e[1];
        E[2] = (double)e[2];
    }
}

/* phi_at, pot.c:1818-1901 (SPH:1922-2035 seeds the grain potential at
   absolute coordinates) */
double or_phi_at(or_sim* s, double x, double y, double z)
{
    const or_config* ru = &s->ru;
    const or_node* root = s->root;
    const or_particle* src = ru->inplace ? s->parts : s->tree_parts;
    const double oap2 = ru->theta * ru->theta;
    double dist, dist2, dx, dy, dz;
    double phi = (ru->geometry != OR_SPHERE) ? grain_phi_abs(s, x, y, z) : 0.0;
    unsigned int path[MAX_PATH_LEN];
    unsigned int path_len = 1;

    if (node_empty(&root[0]))
        return phi;
    path[0] = 0;
    do {
        const or_node* node = &root[path[path_len - 1]];
        if (node->part != UINT_MAX) {
            const or_particle* p = &src[node->part];
            dist2 = ((x - p->x) * (x - p->x)) + ((y - p->y) * (y - p->y)) + ((z - p->z) * (z - p->z));
            dist = sqrt(dist2 + (ru->soft * ru->soft));
            if (p->species == OR_DUST)
                phi += QE * s->dust_grain_q / (4 * PI * EPSILON0 * dist);
            else
                phi += QE * CHARGE[p->species] / (4 * PI * EPSILON0 * dist);
            path_len--;
        } else {
            dx = x - node->cx;
            dy = y - node->cy;
            dz = z - node->cz;
            dist2 = (dx * dx) + (dy * dy) + (dz * dz);
            if (((node->x2 - node->x1) * (node->x2 - node->x1)) < (oap2 * dist2)) {
                if (ru->dipoles && node->aq) {
                    double r_dot_p = (dx * node->px) + (dy * node->py) + (dz * node->pz);
                    dist = sqrt(dist2);
                    phi += (QE / (4 * PI * EPSILON0)) * ((node->q / dist) + (r_dot_p / (dist * dist2)));
                } else if (!ru->dipoles && node->q) {
                    phi += QE * node->q / (4 * PI * EPSILON0 * sqrt(dist2));
                }
                path_len--;
            } else {
                int k;
                path_len--;
                if ((path_len + 8) > MAX_PATH_LEN)
                    abort();
                if (node->ch[0])
                    for (k = 0; k < 8; k++)
                        path[path_len++] = node->ch[k];
            }
        }
    } while (path_len);
    return phi;
}

/* phi_particle_feels, pot.c:1934-2018 (SPH:2092-2180) */
static long double phi_particle_feels_ld(or_sim* s, uint32_t part)
{
    const or_config* ru = &s->ru;
    const or_node* root = s->root;
    const or_particle* src = ru->inplace ? s->parts : s->tree_parts;
    const double oap2 = ru->theta * ru->theta;
    double dist, dist2, dx, dy, dz;
    long double phi = 0.0;
    unsigned int path[MAX_PATH_LEN];
    unsigned int path_len = 1;

    if (ru->geometry != OR_SPHERE)
        phi = or_phi_from_grain(s, part);
    if (node_empty(&root[0]))
        return phi;
    path[0] = 0;
    do {
        const or_node* node = &root[path[path_len - 1]];
        path_len--;
        if (node->part == UINT_MAX) {
            dx = s->parts[part].x - node->cx;
            dy = s->parts[part].y - node->cy;
            dz = s->parts[part].z - node->cz;
            dist2 = (dx * dx) + (dy * dy) + (dz * dz);
            if (((node->x2 - node->x1) * (node->x2 - node->x1)) < (oap2 * dist2)) {
                if (ru->dipoles && node->aq) {
                    double r_dot_p = (dx * node->px) + (dy * node->py) + (dz * node->pz);
                    dist = sqrt(dist2);
                    phi += QE * ((node->q / dist) + (r_dot_p / (dist * dist2))) / (4 * PI * EPSILON0);
                } else if (!ru->dipoles && node->q) {
                    phi += QE * node->q / (4 * PI * EPSILON0 * sqrt(dist2));
                }
            } else {
                int k;
                if ((path_len + 8) > MAX_PATH_LEN)
                    abort();
                if (node->ch[0])
                    for (k = 0; k < 8; k++)
                        path[path_len++] = node->ch[k];
            }
        } else if (node->part != part) {
            const or_particle* p = &src[node->part];
            dx = s->parts[part].x - p->x;
            dy = s->parts[part].y - p->y;
            dz = s->parts[part].z - p->z;
            dist2 = (dx * dx) + (dy * dy) + (dz * dz);
            dist = sqrt(dist2 + (ru->soft * ru->soft));
            if (p->species == OR_DUST)
                phi += QE * s->dust_grain_q / (4 * PI * EPSILON0 * dist);
            else
                phi += QE * CHARGE[p->species] / (4 * PI * EPSILON0 * dist);
        }
    } while (path_len);
    return phi;
}

double or_phi_particle_feels(or_sim* s, uint32_t part)
{
    return (double)phi_particle_feels_ld(s, part);
}

/* softened direct sum in long double (the reference's own accuracy oracle,
   pot.c:2198-2237), with the grain's actual charge; the spheroid grain is
   analytic and not a point source */
void or_direct_sum(const or_sim* s, uint32_t part, double x, double y, double z, double E[3])
{
    const or_config* ru = &s->ru;
    long double acc[3] = {0, 0, 0};
    long double dist, dist2, dx, dy, dz, E_scale;
    uint64_t j;
    const or_particle* src = s->tree_parts;
    if (ru->geometry != OR_SPHERE)
        grain_field_ld(s, part, x, y, z, acc);
    for (j = (ru->geometry != OR_SPHERE) ? 1 : 0; j < ru->N; j++) {
        int qj;
        if (j == part)
            continue;
        qj = (src[j].species == OR_DUST) ? s->dust_grain_q : CHARGE[src[j].species];
        if (!qj)
            continue;
        dx = x - src[j].x;
        dy = y - src[j].y;
        dz = z - src[j].z;
        dist2 = (dx * dx) + (dy * dy) + (dz * dz) + (ru->soft * ru->soft);
        dist = sqrtl(dist2);
        E_scale = QE * qj / (4 * PI * EPSILON0 * dist2 * dist);
        acc[0] += E_scale * dx;
        acc[1] += E_scale * dy;
        acc[2] += E_scale * dz;
    }
    E[0] = (double)acc[0];
    E[1] = (double)acc[1];
    E[2] = (double)acc[2];
}

/* measure_chi_b, pot.c:2251-2279, for rank 0 of 1 worker */
double or_measure_chi_b(or_sim* s)
{
    const or_config* ru = &s->ru;
    const double x_grain = s->parts[0].x;
    const double y_grain = s->parts[0].y;
    const double z_grain = s->parts[0].z;
    const double zen = PI / 2.0;
    const double r = ru->r * ru->mpp;
    double azi = 2.0 * PI * 0 / 1;
    double phi = or_phi_at(s, x_grain + (r * cos(azi) * sin(zen)),
                           y_grain + (r * sin(azi) * sin(zen)), z_grain + (r * cos(zen)));
    return phi / (KB * ru->Ti / (CHARGE[OR_ION] * QE));
}

/* does_particle_cross_grain, pot.c:2287-2325; SPH:2448-2540 for spheroids */
static int crosses_grain(const or_config* ru, const or_particle* p, const double* d_pos,
                         const or_particle* gr)
{
    double s[3];
    if (ru->geometry != OR_SPHERE) {
        const double A = ru->A, a = ru->a;
        double x_min, x_prev, y_min, z_min;
        s[0] = gr->x - p->x;
        s[1] = gr->y - p->y;
        s[2] = gr->z - p->z;
        if ((s[0] * s[0] + (s[1] * s[1] + s[2] * s[2]) * A * A) <= (a * a * pow(A, 4.0 / 3.0)))
            return 1;
        x_min = (d_pos[1] * (d_pos[1] * s[0] - d_pos[0] * s[1]) +
                 d_pos[2] * (d_pos[2] * s[0] - d_pos[0] * s[2])) /
                ((d_pos[0] * d_pos[0] / (A * A)) + d_pos[1] * d_pos[1] + d_pos[2] * d_pos[2]);
        x_prev = s[0] - d_pos[0];
        if (((x_min > s[0]) && (x_min > x_prev)) || ((x_min < s[0]) && (x_min < x_prev)))
            return 0;
        y_min = d_pos[1] * (x_min - s[0]) / d_pos[0] + s[1];
        z_min = d_pos[2] * (x_min - s[0]) / d_pos[0] + s[2];
        return ((x_min * x_min) + (y_min * y_min + z_min * z_min) * A * A) <=
               (a * a * pow(A, 4.0 / 3.0));
    } else {
        const double a = ru->a;
        double alpha, d_dot_s, d_pos2, dist2, s2;
        s[0] = gr->x - p->x;
        s[1] = gr->y - p->y;
        s[2] = gr->z - p->z;
        s2 = (s[0] * s[0]) + (s[1] * s[1]) + (s[2] * s[2]);
        if (s2 <= (a * a))
            return 1;
        d_pos2 = (d_pos[0] * d_pos[0]) + (d_pos[1] * d_pos[1]) + (d_pos[2] * d_pos[2]);
        d_dot_s = (d_pos[0] * s[0]) + (d_pos[1] * s[1]) + (d_pos[2] * s[2]);
        alpha = d_dot_s / d_pos2;
        if ((alpha < -1.0) || (alpha > 0.0))
            return 0;
        dist2 = s2 - (alpha * d_dot_s);
        return dist2 <= (a * a);
    }
}

int or_crosses_grain(const or_sim* s, const double p_new[3], const double d_pos[3],
                     const double grain[3])
{
    or_particle p, g;
    memset(&p, 0, sizeof(p));
    memset(&g, 0, sizeof(g));
    p.x = p_new[0];
    p.y = p_new[1];
    p.z = p_new[2];
    g.x = grain[0];
    g.y = grain[1];
    g.z = grain[2];
    return crosses_grain(&s->ru, &p, d_pos, &g);
}

void or_reinject(or_sim* s, or_particle* p, uint32_t ref_idx)
{
    or_rng g;
    memset(&g, 0, sizeof(g));
    if (s->ru.rng_kind == OR_RNG_PHILOX) {
        g.kind = 1;
        pstream_init(&g.ps, s->ru.philox_seed, ref_idx, s->t_steps);
    } else {
        g.kind = 0;
        g.well = &s->ps;
    }
    inject_g(s, p, 1, &g);
}

/* ------------------------------------------------------------------ */
/* move_particles, pot.c:2330-2934, for a single worker with the BORIS
   integrator (the configurations of this build) */
/* OUT_BOUNCE containment of a spherical domain, check_particle_in_sim_area
   (pot.c:2036-2069): outside the wall and moving outwards, the particle's
   velocity is reflected in spherical components (radial by -coe, the others
   by coe) and it goes back to where it started the step; outside but moving
   inwards it is left alone and -1 reported (pot.c:2724-2728).  The angles
   are computed in long double and kept as double, as the reference's r[]. */
static int bounce_off_wall(const or_config* ru, or_particle* part, double prev_x, double prev_y,
                           double prev_z)
{
    const double rad = ru->r * ru->mpp;
    double r[3], v[3];
    r[0] = pythag(part->x - rad, part->y - rad, part->z - rad);
    if (r[0] > rad) {
        r[1] = atan2l(part->y - rad, part->x - rad);
        r[2] = acosl((part->z - rad) / r[0]);
        v[0] = part->vx * cos(r[1]) * sin(r[2]) + part->vy * sin(r[1]) * sin(r[2]) +
               part->vz * cos(r[2]);
        v[1] = -part->vx * sin(r[1]) + part->vy * cos(r[1]);
        v[2] = part->vx * cos(r[1]) * cos(r[2]) + part->vy * sin(r[1]) * cos(r[2]) -
               part->vz * sin(r[2]);
        if (v[0] > 0.0) {
            v[0] *= -ru->bounce_coe;
            v[1] *= ru->bounce_coe;
            v[2] *= ru->bounce_coe;
            velocity_in_cartesian_coords(v, r[1], r[2], &part->vx, &part->vy, &part->vz);
        } else {
            return -1;
        }
        part->x = prev_x;
        part->y = prev_y;
        part->z = prev_z;
    }
    return 0;
}

/* move_particles, pot.c:2330-2934 (P = 1: the gathers, the intermediate
   Bcast and the reductions are identities).  Boris (pot.c:2472-2531) takes
   one field per particle at its half-drifted position; velocity Verlet and
   Euler-Richardson (pot.c:2532-2554, 2598-2639) take one at the start
   position, write the intermediate state to the live array, and take a
   second there, with the tree built at the start of the step (its cells'
   moments unchanged, its leaves read live). */
int or_move_particles(or_sim* s, int compute_epe, or_step_out* out)
{
    const or_config* ru = &s->ru;
    const unsigned long time_steps = s->t_steps;
    const unsigned int last = (unsigned int)(ru->N - 1);
    const double B[3] = {ru->B[0], ru->B[1], ru->B[2]};
    const double B_mag = pythag(B[0], B[1], B[2]);
    const double b[3] = {B[0] / B_mag, B[1] / B_mag, B[2] / B_mag};
    const int boris = ru->integrator == OR_BORIS;
    double dephi;
    long double dt, E[3], qom, a_x2, a_y2, a_z2;
    unsigned int i, ims, start;
    or_particle *p, *p_n;
    unsigned int absorbed_e = 0, absorbed_i = 0, n_wall_inward = 0;
    long double absorbed_m = 0.0;
    int absorbed_q = 0;
    double delta_pos[3];
    long double dpx = 0.0, dpy = 0.0, dpz = 0.0;
    or_rng g;
    unsigned int n_reinj_abs = 0, n_reinj_out = 0;
    const or_particle* src;

    memcpy(s->new_parts, s->parts, sizeof(or_particle) * ru->N);
    src = ru->inplace ? s->parts : s->tree_parts;

    if (ru->sceptic && !(s->t_steps % 20)) {
        s->chi_b = or_measure_chi_b(s);
        s->chi_b /= 1;
    }

    for (start = 0; start <= last; start++)
        if (!(time_steps % s->tss[s->parts[start].species]))
            break;

    for (i = start; i <= last; i++) {
        p = &s->parts[i];
        s->prev_x[i - start] = p->x;
        s->prev_y[i - start] = p->y;
        s->prev_z[i - start] = p->z;
    }

    for (i = start; i <= last; i++) {
        p = &s->parts[i];
        dt = ru->dt * s->tss[p->species];
        if (boris) {
            /* half drift in place; with inplace = 0 the leaves keep
               reading the build snapshot, so only the target moves */
            p->x += 0.5 * dt * p->vx;
            p->y += 0.5 * dt * p->vy;
            p->z += 0.5 * dt * p->vz;
        }
        field_walk(s, i, p->x, p->y, p->z, src, E, NULL);
        ims = i - start;
        if (p->species == OR_DUST)
            qom = QE * s->dust_grain_q;
        else
            qom = QE * CHARGE[p->species];
        qom /= s->mass[p->species];
        if (boris) {
            s->a_x1[ims] = 0.5 * qom * E[0];
            s->a_y1[ims] = 0.5 * qom * E[1];
            s->a_z1[ims] = 0.5 * qom * E[2];
            s->new_parts[i].vx += dt * s->a_x1[ims];
            s->new_parts[i].vy += dt * s->a_y1[ims];
            s->new_parts[i].vz += dt * s->a_z1[ims];
            if (B_mag) {
                dephi = 2.0 * atanl(0.5 * dt * qom * B_mag);
                rotate_vec_about_dir(b, sin(dephi), cos(dephi), &s->new_parts[i].vx,
                                     &s->new_parts[i].vy, &s->new_parts[i].vz, 0.0);
            }
            s->new_parts[i].vx += dt * s->a_x1[ims];
            s->new_parts[i].vy += dt * s->a_y1[ims];
            s->new_parts[i].vz += dt * s->a_z1[ims];
            s->new_parts[i].x = p->x + (0.5 * dt * s->new_parts[i].vx);
            s->new_parts[i].y = p->y + (0.5 * dt * s->new_parts[i].vy);
            s->new_parts[i].z = p->z + (0.5 * dt * s->new_parts[i].vz);
        } else {
            s->a_x1[ims] = qom * (E[0] + p->vy * B[2] - p->vz * B[1]);
            s->a_y1[ims] = qom * (E[1] + p->vz * B[0] - p->vx * B[2]);
            s->a_z1[ims] = qom * (E[2] + p->vx * B[1] - p->vy * B[0]);
            s->new_parts[i].vx += 0.5 * dt * s->a_x1[ims];
            s->new_parts[i].vy += 0.5 * dt * s->a_y1[ims];
            s->new_parts[i].vz += 0.5 * dt * s->a_z1[ims];
            if (ru->integrator == OR_EULER_RICHARDSON) {
                s->new_parts[i].x += 0.5 * dt * p->vx;
                s->new_parts[i].y += 0.5 * dt * p->vy;
                s->new_parts[i].z += 0.5 * dt * p->vz;
            } else {
                s->new_parts[i].x += dt * (p->vx + (0.5 * dt * s->a_x1[ims]));
                s->new_parts[i].y += dt * (p->vy + (0.5 * dt * s->a_y1[ims]));
                s->new_parts[i].z += dt * (p->vz + (0.5 * dt * s->a_z1[ims]));
            }
        }
    }

    if (!boris) {
        /* the intermediate state becomes the live array (pot.c:2557-2591);
           the leaves the second walk reads follow it in both semantics */
        memcpy(s->parts, s->new_parts, sizeof(or_particle) * ru->N);
        if (!ru->inplace)
            memcpy(s->tree_parts, s->parts, sizeof(or_particle) * ru->N);
    }

    memset(&g, 0, sizeof(g));
    for (i = start; i <= last; i++) {
        p = &s->parts[i];
        p_n = &s->new_parts[i];
        ims = i - start;
        if (!boris) {
            field_walk(s, i, p->x, p->y, p->z, src, E, NULL);
            if (p->species == OR_DUST)
                qom = QE * s->dust_grain_q;
            else
                qom = QE * CHARGE[p->species];
            qom /= s->mass[p->species];
            a_x2 = qom * (E[0] + p->vy * B[2] - p->vz * B[1]);
            a_y2 = qom * (E[1] + p->vz * B[0] - p->vx * B[2]);
            a_z2 = qom * (E[2] + p->vx * B[1] - p->vy * B[0]);
            dt = ru->dt * s->tss[p->species];
            if (ru->integrator == OR_EULER_RICHARDSON) {
                p_n->x = (p->x - 0.5 * dt * (p->vx - 0.5 * dt * s->a_x1[ims])) + dt * p->vx;
                p_n->y = (p->y - 0.5 * dt * (p->vy - 0.5 * dt * s->a_y1[ims])) + dt * p->vy;
                p_n->z = (p->z - 0.5 * dt * (p->vz - 0.5 * dt * s->a_z1[ims])) + dt * p->vz;
                p_n->vx = (p->vx - 0.5 * dt * s->a_x1[ims]) + dt * a_x2;
                p_n->vy = (p->vy - 0.5 * dt * s->a_y1[ims]) + dt * a_y2;
                p_n->vz = (p->vz - 0.5 * dt * s->a_z1[ims]) + dt * a_z2;
            } else {
                p_n->vx += 0.5 * dt * a_x2;
                p_n->vy += 0.5 * dt * a_y2;
                p_n->vz += 0.5 * dt * a_z2;
            }
        }
        if (ru->rng_kind == OR_RNG_PHILOX) {
            g.kind = 1;
            pstream_init(&g.ps, ru->philox_seed, i, time_steps);
        } else {
            g.kind = 0;
            g.well = &s->ps;
        }
        if (ru->grain_enabled && (i > 0) && ((time_steps * ru->dt) > ru->settle)) {
            delta_pos[0] = p_n->x - s->prev_x[ims];
            delta_pos[1] = p_n->y - s->prev_y[ims];
            delta_pos[2] = p_n->z - s->prev_z[ims];
            if (crosses_grain(ru, p_n, delta_pos, &s->new_parts[0])) {
                if (p_n->species == OR_ELECTRON)
                    absorbed_e++;
                else
                    absorbed_i++;
                absorbed_m += s->mass[p_n->species];
                absorbed_q += CHARGE[p_n->species];
                dpx += s->mass[p_n->species] * p_n->vx;
                dpy += s->mass[p_n->species] * p_n->vy;
                dpz += s->mass[p_n->species] * p_n->vz;
                inject_g(s, p_n, 1, &g);
                n_reinj_abs++;
            }
        }
        if (ru->out_bounce) {
            if (bounce_off_wall(ru, p_n, s->prev_x[ims], s->prev_y[ims], s->prev_z[ims]) < 0)
                n_wall_inward++;
        } else {
            double rad = ru->r * ru->mpp;
            if (pythag(p_n->x - rad, p_n->y - rad, p_n->z - rad) > rad) {
                inject_g(s, p_n, 1, &g);
                n_reinj_out++;
            }
        }
    }

    if (compute_epe) {
        s->epe[OR_ELECTRON] = 0.0;
        s->epe[OR_ION] = 0.0;
        for (i = 1; i <= last; i++) {
            p_n = &s->new_parts[i];
            s->epe[p_n->species] += QE * CHARGE[p_n->species] * phi_particle_feels_ld(s, i);
        }
    }

    if (ru->grain_enabled) {
        s->dust_grain_m += absorbed_m;
        s->new_parts[0].vx += dpx / s->dust_grain_m;
        s->new_parts[0].vy += dpy / s->dust_grain_m;
        s->new_parts[0].vz += dpz / s->dust_grain_m;
        s->dust_grain_q += absorbed_q;
        s->tot_absor_e += absorbed_e;
        if (compute_epe) {
            if (!s->dust_grain_q)
                s->epe[OR_DUST] = 0.0;
            else
                s->epe[OR_DUST] = s->dust_grain_q * QE * phi_particle_feels_ld(s, 0);
        }
    }

    memcpy(s->parts, s->new_parts, sizeof(or_particle) * ru->N);

    if (out) {
        out->abs_q = absorbed_q;
        out->abs_e = absorbed_e;
        out->abs_i = absorbed_i;
        out->n_pushed = (start <= last) ? last - start + 1 : 0;
        out->n_reinj_abs = n_reinj_abs;
        out->n_reinj_out = n_reinj_out;
        out->start = start;
        out->n_wall_inward = n_wall_inward;
    }
    return 0;
}

/* record_macro_this_step, pot.c:4364-4369 */
int or_record_macro_this_step(const or_sim* s, uint64_t t)
{
    int rec;
    if (!s->ru.record_macro)
        return 0;
    rec = (t <= 200);
    if ((t * s->ru.dt) > (2.0 * s->ru.settle))
        rec |= !(t % 2000);
    else
        rec |= !(t % 200);
    return rec;
}

/* one iteration of run()'s loop, pot.c:4259-4388, without output */
int or_run_step(or_sim* s, or_step_out* out)
{
    int rc;
    s->t_steps++;
    rc = or_build_tree(s);
    if (rc)
        return rc;
    return or_move_particles(s, or_record_macro_this_step(s, s->t_steps), out);
}

/* ------------------------------------------------------------------ */
/* write_sim_state_to_file, pot.c:3816-3901, one worker */
int or_write_state(const or_sim* s, const char* path)
{
    FILE* fp = fopen(path, "wb");
    double chi_b = s->ru.sceptic ? s->chi_b : 9999.99;
    if (!fp) {
        set_err("cannot open state file for writing");
        return -1;
    }
    /* the reference prints t_steps - 1 from a counter already advanced past
       the last completed step (pot.c:3870-3873, 4261); this counter holds
       the completed steps themselves */
    fprintf(fp, "C %u %i %Lg %lu %Lg %Lg %Lg %u %g\n", 1u, s->dust_grain_q, s->dust_grain_m,
            s->t_steps, s->epe[0], s->epe[1], s->epe[2], s->tot_absor_e, chi_b);
    fwrite(&s->ps, sizeof(or_prng), 1, fp);
    fwrite(s->parts, sizeof(or_particle), s->ru.N, fp);
    fclose(fp);
    return 0;
}

/* read_sim_state_from_file, pot.c:3904-4019, format C, one worker */
int or_read_state(or_sim* s, const char* path)
{
    FILE* fp = fopen(path, "rb");
    unsigned int workers;
    double chi_b;
    unsigned long t;
    if (!fp) {
        set_err("cannot open state file");
        return -1;
    }
    if (getc(fp) != 'C' ||
        fscanf(fp, " %u %i %Lg %lu %Lg %Lg %Lg %u %lg\n", &workers, &s->dust_grain_q,
               &s->dust_grain_m, &t, &s->epe[0], &s->epe[1], &s->epe[2], &s->tot_absor_e,
               &chi_b) != 9) {
        fclose(fp);
        set_err("not a format C state file");
        return -2;
    }
    s->t_steps = t;
    s->chi_b = chi_b;
    if (workers != 1 || fread(&s->ps, sizeof(or_prng), 1, fp) != 1 ||
        fread(s->parts, sizeof(or_particle), s->ru.N, fp) != s->ru.N) {
        fclose(fp);
        set_err("bad state file body");
        return -3;
    }
    if (s->ps.index >= 16)
        s->ps.index %= 16;
    fclose(fp);
    return 0;
}
