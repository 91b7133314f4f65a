/*
 * pot_macro.c -- one record of the macroscopic summary file
 * (pot-macroscopic.dat, write_macroscopic_data_to_file, pot.c:3261-3506;
 * the same function in pot-spheroid.c:3483).
 *
 * The 35 columns are printed with the reference's conversions from
 * long double accumulators, so the arithmetic below keeps the reference's
 * types and the order in which it adds particles (index order, grain
 * excluded) -- a different order or a double accumulator changes the last
 * printed digit.  Two passes: means first, then central moments about
 * those means.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "pot_host.h"

#define KB_ (1.38065e-23)
#define EPSILON0_ (8.85419e-12)
#define QE_ (1.6021766e-19)
#define ME_ (9.109383e-31)
#define PI_ 3.14159265358979

/* electron and ion accumulators of the first pass */
typedef struct species_sums {
    unsigned int n;
    long double speed;     /* sum of |v| */
    long double ke;        /* sum of kinetic energy (electrons only) */
    long double v[3];      /* sum of velocity components */
    long double ang[3];    /* sum of (r - centre) x v */
} species_sums;

/* second pass: central moments, the ion kinetic energy and the
   electron position moments */
typedef struct species_moments {
    long double speed[4];  /* mean, then sums of powers 2..4 of (|v| - mean) */
    long double ke[4];
    double temp;           /* sum of |v - <v>|^2, a double as in the reference */
} species_moments;

const char* pot_macro_header(void)
{
    return "TIME KE PE KI PI KG PG XE XESD DGQ DL DN "
           "SE1 SE2 SE3 SE4 SI1 SI2 SI3 SI4 TE TI LMEX LMEY LMEZ "
           "LMIX LMIY LMIZ AMEX AMEY AMEZ AMIX AMIY AMIZ TOTE\n";
}

/* variance, skewness and excess kurtosis from sums of powers of the
   deviation (divide_and_adjust_central_moments, pot.c:3250-3259) */
static void finish_moments(long double m[4], unsigned int n)
{
    m[1] /= n;
    m[2] /= n * powl(m[1], 1.5);
    m[3] /= n * m[1] * m[1];
    m[3] -= 3.0;
}

static int put(char* buf, size_t cap, size_t* at, const char* fmt, ...)
    __attribute__((format(printf, 4, 5)));

#include <stdarg.h>
static int put(char* buf, size_t cap, size_t* at, const char* fmt, ...)
{
    va_list ap;
    int k;
    if (*at >= cap) return -1;
    va_start(ap, fmt);
    k = vsnprintf(buf + *at, cap - *at, fmt, ap);
    va_end(ap);
    if (k < 0 || (size_t)k >= cap - *at) return -1;
    *at += (size_t)k;
    return 0;
}

/* an energy column, or NA when it is neither normal nor zero (pot.c:276) */
static int put_energy(char* buf, size_t cap, size_t* at, long double e, long double per)
{
    if (isnormal(e) || fpclassify(e) == FP_ZERO) return put(buf, cap, at, "%Lg ", e / per);
    return put(buf, cap, at, "NA ");
}

int pot_macro_format(const pot_macro_in* in, char* buf, size_t cap)
{
    const long double mass[3] = {1.67e-12, in->mime * ME_, ME_};
    const double cent = in->r * in->mpp;
    const long double volume = (4.0 / 3.0) * PI_ * powl(in->r * in->mpp, 3.0);
    species_sums se, si;
    species_moments me, mi;
    long double ke_ions = 0.0, x_sum = 0.0, x_sq = 0.0, debye;
    long double lm_e[3], lm_i[3];
    size_t at = 0;
    uint64_t k;
    int c;

    if (in->N < 6) return 0; /* pot.c:3311-3318: no record for tiny runs */
    memset(&se, 0, sizeof(se));
    memset(&si, 0, sizeof(si));
    memset(&me, 0, sizeof(me));
    memset(&mi, 0, sizeof(mi));

    for (k = 1; k < in->N; k++) {
        const potg_particle* p = &in->parts[k];
        const long double v = sqrt(p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
        species_sums* s = (p->species == POTG_ELECTRON) ? &se : &si;
        if (p->species == POTG_ELECTRON) s->ke += 0.5 * mass[POTG_ELECTRON] * v * v;
        s->speed += v;
        s->n++;
        s->v[0] += p->vx;
        s->v[1] += p->vy;
        s->v[2] += p->vz;
        s->ang[0] += ((p->y - cent) * p->vz) - ((p->z - cent) * p->vy);
        s->ang[1] -= ((p->x - cent) * p->vz) - ((p->z - cent) * p->vx);
        s->ang[2] += ((p->x - cent) * p->vy) - ((p->y - cent) * p->vx);
    }
    me.ke[0] = se.ke / se.n;
    me.speed[0] = se.speed / se.n;
    mi.speed[0] = si.speed / si.n;
    for (c = 0; c < 3; c++) {
        se.v[c] /= se.n;
        si.v[c] /= si.n;
    }
    for (c = 0; c < 3; c++) {
        se.ang[c] *= mass[POTG_ELECTRON];
        si.ang[c] *= mass[POTG_ION];
        lm_e[c] = mass[POTG_ELECTRON] * se.v[c];
        lm_i[c] = mass[POTG_ION] * si.v[c];
    }

    for (k = 1; k < in->N; k++) {
        const potg_particle* p = &in->parts[k];
        const long double v = sqrt(p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
        const long double K = 0.5 * mass[p->species] * v * v;
        if (p->species == POTG_ELECTRON) {
            const long double dk = K - me.ke[0];
            const long double ds = v - me.speed[0];
            x_sum += p->x;
            x_sq += p->x * p->x;
            me.ke[1] += dk * dk;
            me.ke[2] += dk * dk * dk;
            me.ke[3] += dk * dk * dk * dk;
            me.speed[1] += ds * ds;
            me.speed[2] += ds * ds * ds;
            me.speed[3] += ds * ds * ds * ds;
            for (c = 0; c < 3; c++) {
                const double pv = c == 0 ? p->vx : (c == 1 ? p->vy : p->vz);
                me.temp += (pv - se.v[c]) * (pv - se.v[c]);
            }
        } else {
            const long double ds = v - mi.speed[0];
            ke_ions += K;
            mi.speed[1] += ds * ds;
            mi.speed[2] += ds * ds * ds;
            mi.speed[3] += ds * ds * ds * ds;
            for (c = 0; c < 3; c++) {
                const double pv = c == 0 ? p->vx : (c == 1 ? p->vy : p->vz);
                mi.temp += (pv - si.v[c]) * (pv - si.v[c]);
            }
        }
    }
    finish_moments(me.ke, se.n);
    finish_moments(me.speed, se.n);
    finish_moments(mi.speed, si.n);
    me.temp *= (mass[POTG_ELECTRON] / (3.0 * KB_)) / se.n;
    mi.temp *= (mass[POTG_ION] / (3.0 * KB_)) / si.n;
    debye = sqrtl(EPSILON0_ * KB_ * me.temp / ((se.n / volume) * QE_ * QE_));

    if (put(buf, cap, &at, "%.7Lg %Lg ", in->time_step * in->dt, me.ke[0])) return -1;
    if (put_energy(buf, cap, &at, in->epe[POTG_ELECTRON], se.n)) return -1;
    if (put(buf, cap, &at, "%Lg ", ke_ions / si.n)) return -1;
    if (put_energy(buf, cap, &at, in->epe[POTG_ION], si.n)) return -1;
    if (put(buf, cap, &at, "%Lg ",
            0.5 * in->grain_m *
                hypot(in->parts[0].vx, hypot(in->parts[0].vy, in->parts[0].vz))))
        return -1;
    if (put_energy(buf, cap, &at, in->epe[POTG_DUST], 1)) return -1;
    if (put(buf, cap, &at,
            "%.7Lg %.7Lg %i %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %g %g "
            "%Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %Lg %u\n",
            x_sum / se.n, sqrtl(x_sq / se.n - (x_sum / se.n) * x_sum / se.n), in->grain_q, debye,
            (4.0 * PI_ / 3.0) * debye * debye * debye * se.n / volume, me.speed[0], me.speed[1],
            me.speed[2], me.speed[3], mi.speed[0], mi.speed[1], mi.speed[2], mi.speed[3],
            me.temp, mi.temp, lm_e[0], lm_e[1], lm_e[2], lm_i[0], lm_i[1], lm_i[2], se.ang[0],
            se.ang[1], se.ang[2], si.ang[0], si.ang[1], si.ang[2], in->tot_absor_e))
        return -1;
    return (int)at;
}
