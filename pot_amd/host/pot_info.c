/*
 * pot_info.c -- the run's two descriptive files, written once at start-up
 * (run(), pot.c:4109-4112):
 *   pot-run-init-info.txt  one tab-separated line per run
 *                          (write_run_info_to_file, pot.c:3512-3553);
 *   pot-run-descrip.txt    the settings, with '*' on non-defaults, plasma
 *                          parameter estimates, with '!' on doubtful ones,
 *                          and the build's switches
 *                          (write_run_description_to_file, pot.c:3583-3813).
 * The estimates use the reference's formulas and types so the files read
 * the same for the same settings.
 */
#include <math.h>
#include <stdio.h>
#include <time.h>

#include "pot_host.h"

#define KB_ (1.38065e-23)
#define EPSILON0_ (8.85419e-12)
#define QE_ (1.6021766e-19)
#define ME_ (9.109383e-31)
#define PI_ 3.14159265358979

static long double ion_mass(const pot_run* ru) { return ru->mime * ME_; }

void pot_write_run_info(const char* path, const pot_run* ru)
{
    char when[50];
    time_t now = time(NULL);
    struct tm* tm_now;
    FILE* fp = fopen(path, "ab");
    if (!fp) {
        fprintf(stderr, "Can't open %s to write information about this run.\n", path);
        return;
    }
    tm_now = localtime(&now);
    if (!tm_now || !strftime(when, sizeof(when), "%Y-%m-%d %H:%M:%S", tm_now)) {
        fputs("Can't represent date & time as string of text.\n", stderr);
        fclose(fp);
        return;
    }
    fprintf(fp, "%s\t%lu\t%Lg\t%g\t%g\t%g\t%g\t%g\t%s\t%s\t%u\t%g\t%u\t%g\n", when, ru->N, ru->dt,
            ru->soft, ru->mpp, ru->Ti, ru->Te, ru->x_drift, __DATE__, __TIME__, ru->r, ru->a, 1u,
            ru->theta);
    fclose(fp);
}

/* charging time of a grain by OML currents at small t
   (early_oml_charge_time, pot.c:3555-3570) */
static double oml_charge_time(const pot_run* ru, double n_total)
{
    const long double mi = ion_mass(ru), me = ME_;
    const double nu = sqrt(mi * ru->Te / (me * ru->Ti));
    double tim = EPSILON0_ * sqrt(2.0 * PI_ * mi * KB_ * ru->Te * ru->Te / ru->Ti);
    tim /= (QE_ * QE_ * ru->a * n_total / 2.0);
    tim *= log((1.0 - nu) / (pow(nu, 1.0 - exp(-1.0)) - nu));
    return tim;
}

static char mark(int differs) { return differs ? '*' : ' '; }
static char warn(int doubtful) { return doubtful ? '!' : ' '; }

void pot_write_run_description(const char* path, const pot_run* ru, const pot_run* def)
{
    const int sph = ru->geometry != POTG_SPHERE;
    const long double mi = ion_mass(ru), me = ME_;
    FILE* fp = fopen(path, "ab");
    long double n, size, debye_len, wsr, e_colll, edct, eop, htcd;
    double debye_number, ccp, coulomb_log, charge_osc_rat;
    if (!fp) {
        fprintf(stderr, "Can't open %s to write this run's initial description.\n", path);
        return;
    }
    fprintf(fp, "%c total particle count:          %lu\n", mark(ru->N != def->N), ru->N);
    fprintf(fp, "%c time step (s):                 %Lg\n", mark(ru->dt != def->dt), ru->dt);
    fprintf(fp, "%c nominal elec. temperature (K): %g\n", mark(ru->Te != def->Te), ru->Te);
    fprintf(fp, "%c nominal ion temperature (K):   %g\n", mark(ru->Ti != def->Ti), ru->Ti);
    fprintf(fp, "%c softening parameter (m):       %g\n", mark(ru->soft != def->soft), ru->soft);
    fprintf(fp, "%c drift velocity (m*s^-1):       %g\n", mark(ru->x_drift != def->x_drift),
            ru->x_drift);
    fprintf(fp, "%c pixel size (m):                %g\n", mark(ru->mpp != def->mpp), ru->mpp);
    fprintf(fp, "%c simulation radius (px.):       %u\n", mark(ru->r != def->r), ru->r);
    if (sph) fprintf(fp, "%c aspect ratio:                  %g\n", mark(ru->A != def->A), ru->A);
    fprintf(fp, "%c tracked particles:             %hu\n", ' ', (unsigned short)0);
    fprintf(fp, "%c number of time steps:          %lu\n", mark(ru->run_for != def->run_for),
            ru->run_for);
    fprintf(fp, "%c grain radius (m):              %g\n", mark(ru->a != def->a), ru->a);
    fprintf(fp, "%c settling time (s):             %g\n\n", mark(ru->settle != def->settle),
            ru->settle);

    /* plasma parameter estimates (pot.c:3650-3676) */
    n = (ru->N - 1) / ((4.0 / 3.0) * PI_ * powl(ru->r * ru->mpp, 3.0));
    size = 2.0 * ru->r * ru->mpp;
    debye_len = sqrtl(EPSILON0_ * KB_ * ru->Te / ((n / 2.0) * QE_ * QE_));
    debye_number = (4.0 / 3.0) * PI_ * powl(debye_len, 3.0) * n / 2.0;
    wsr = powl(4.0 * PI_ * n / 3.0, -1.0 / 3.0);
    ccp = QE_ * 1 * QE_ * 1; /* |CHARGE[ELECTRON]| == |CHARGE[ION]| == 1 */
    ccp /= (4.0 * PI_ * EPSILON0_ * wsr * KB_ * ru->Te);
    coulomb_log = log(9.0 * debye_number);
    e_colll = 16.0 * PI_ * (n / 2.0) * powl(debye_len, 4.0) / coulomb_log;
    edct = ru->a / sqrtl(3.0 * KB_ * ru->Te / me);
    eop = 2.0 * PI_ * sqrtl(ME_ * EPSILON0_ / (n * QE_ * QE_ / 2.0));
    charge_osc_rat = oml_charge_time(ru, n) / eop;
    htcd = QE_ * QE_ / (12 * PI_ * EPSILON0_ * KB_);
    htcd /= (ru->Te > ru->Ti) ? ru->Te : ru->Ti;
    fprintf(fp,
            "  particle density (m^-3):    %Lg\n"
            "  Debye length (m):           %Lg\n"
            "  grain radius/DL:            %Lg\n"
            "  elec. oscil. period (s):    %Lg\n"
            "  early OML chg. time (s):    %g\n"
            "  Wigner-Seitz radius (m):    %Lg\n"
            "  normalized drift speed:     %Lg\n"
            "  ion traversal time (s):     %Lg\n"
            "  Coulomb logarithm:          %g\n"
            "  electron coll. len. (m):    %Lg\n",
            n, debye_len, ru->a / debye_len, eop, oml_charge_time(ru, n), wsr,
            ru->x_drift / sqrtl(KB_ * ru->Ti / mi), size / sqrtl(3.0 * KB_ * ru->Ti / mi),
            coulomb_log, e_colll);
    fprintf(fp, "%c electron Knudsen number:    %Lg\n", warn((e_colll / size) > 0.5), e_colll / size);
    fprintf(fp, "%c Debye number:               %g\n", warn(debye_number < 2.0), debye_number);
    fprintf(fp, "%c Coulomb coupling param.:    %g\n", warn(ccp > 1.0), ccp);
    fprintf(fp, "%c softening parameter/WSR:    %Lg\n", warn((ru->soft / wsr) > 1.0), ru->soft / wsr);
    fprintf(fp, "%c t. step/elec. cross time:   %Lg\n", warn((ru->dt / edct) > 0.5), ru->dt / edct);
    fprintf(fp, "%c DL/characteristic size:     %Lg\n", warn((debye_len / size) > 0.1),
            debye_len / size);
    fprintf(fp, "%c OML time/elec. osc. period: %g\n", warn(charge_osc_rat < 100.0), charge_osc_rat);
    fprintf(fp, "%c soft. par./head coll. dist: %Lg\n", warn((ru->soft / htcd) > 0.5),
            ru->soft / htcd);
    fprintf(fp, "%c soft. par./grain radius:    %g\n\n", warn((ru->soft / ru->a) > 0.1),
            ru->soft / ru->a);

    /* the reference's compile-time parameters, run-time values here */
    fprintf(fp,
            "opening angle parameter: %g\n"
            "ion/electron mass ratio: %Lg\n"
            "external B field (T):    (%g, %g, %g)\n"
            "spatial bin counts:      (%i, %i, %i)\n"
            "max. tree walk length:   %i\n"
            "grain mass (kg):         %Lg\n\n",
            ru->theta, mi / me, ru->B[0], ru->B[1], ru->B[2], 50, 4, 10, 400, (long double)1.67e-12);
    fputs(ru->grain ? "grain:                enabled\n" : "grain:                disabled\n", fp);
    fputs("OpenGL interface:     no\n", fp);
    fputs("trajectory length:    N/A\n", fp);
    if (ru->out_bounce)
        fprintf(fp, "containment:          bouncing (c.o.e. %g)\n", ru->bounce_coe);
    else
        fputs("containment:          reinjection (SCEPTIC-style)\n", fp);
    fputs("FP exceptions:        no\n", fp);
    fputs(ru->dipoles ? "multipole expansion:  dipole\n" : "multipole expansion:  monopole\n", fp);
    fputs("simulation shape:     spherical\n", fp);
    fputs(ru->integrator == POTG_EULER_RICHARDSON ? "numerical integrator: Euler-Richardson\n"
          : ru->integrator == POTG_VERLET ? "numerical integrator: velocity Verlet\n"
                                          : "numerical integrator: Boris\n",
          fp);
    fputs("pos. Pois. collec'n:  no\n", fp);
    fputs("Langmuir test wave:   no\n", fp);
    if (sph) {
        fputs("P-P interactions:     yes\n", fp);
        fputs(ru->geometry == POTG_PROLATE ? "Grain shape:          prolate spheroid\n\n"
                                           : "Grain shape:          oblate spheroid\n\n",
              fp);
    } else {
        fputs("P-P interactions:     yes\n\n", fp);
    }
    fprintf(fp, "output file dir.:     %s\ncompiled on:          %s %s\n", ru->out_dir, __DATE__,
            __TIME__);
    fclose(fp);
}
