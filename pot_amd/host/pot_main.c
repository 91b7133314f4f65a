/*
 * pot_main.c -- the drop-in programs `pot` (spherical grain, pot.c) and
 * `pot-spheroid` (spheroidal grain, pot-spheroid.c), one process on one
 * GPU.
 *
 * Same command line as the reference (set_run_params_from_args,
 * pot.c:4555-4850; -A from SPH:4800-4822), same defaults per program, same
 * files under the output directory: pot-run-init-info.txt,
 * pot-run-descrip.txt, pot-macroscopic.dat, pot-phi.dat, and format C state
 * files with -f / -g / -R.  The run loop is the reference's run()
 * (pot.c:4259-4494) with the step handed to the C ABI of potg.h: the
 * particles stay on the GPU and come back only for macroscopic records
 * and state files.
 *
 * The reference fixes some settings at compile time; here they are
 * long options:
 *   --theta X        OAP, the opening angle (pot.c:95-97)
 *   --bx/--by/--bz X EXTERNAL_B_{X,Y,Z} in tesla (pot.c:55-57)
 *   --mime X         MIME, ion/electron mass ratio (pot.c:60)
 *   --no-grain       DISABLE_GRAIN (pot.c:30)
 *   --dipoles        DIPOLES: dipole moments in the tree (pot.c:99-105)
 *   --verlet         velocity Verlet, the reference without BORIS
 *                    (pot.c:122-132)
 *   --euler-richardson  EULER_RICHARDSON (pot.c:122-128)
 *   --bounce COE     OUT_BOUNCE with coefficient of restitution COE:
 *                    particles bounce off the wall (pot.c:113-120)
 *   --tune           time the walk's launch shapes after the first steps
 *                    and keep the fastest (potg_tune); without it the
 *                    shape follows the particle count, so runs repeat
 *                    exactly
 *   --oblate         OBLATE instead of PROLATE (pot-spheroid only)
 *   --fp64           float64 walk (default: float32 walk with compensated
 *                    sums, float64 push)
 *   --seed-file P    RANDOMNESS_SOURCE: 64 bytes of PRNG seed (default
 *                    /dev/urandom)
 *   --output-dir D   OUTPUT_DIR (default ./output/)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pot_host.h"

#if defined(POT_GEOMETRY_SPHEROID)
#define SPHEROID_PROGRAM 1
#else
#define SPHEROID_PROGRAM 0
#endif

#define CELLS_R 50
#define CELLS_AZI 4
#define CELLS_ZEN 10
#define PI_ 3.14159265358979
#define DUST_MASS 1.67e-12

/* defaults of the two programs (pot.c:198-216, SPH:203-227) */
static void set_defaults(pot_run* ru)
{
    memset(ru, 0, sizeof(*ru));
    ru->N = 150001;
    ru->dt = 1e-12L;
    ru->mpp = 4e-7;
    ru->run_for = 0;
    ru->x_drift = 0.0;
    ru->a = 2.5e-6;
    ru->tim_lim = 0.0;
    ru->grain = 1;
    ru->dipoles = 0;
    ru->integrator = POTG_BORIS;
    ru->out_bounce = 0;
    ru->bounce_coe = 0.0;
    ru->tune = 0;
    ru->precision = POTG_MIXED;
    ru->seed_path = "/dev/urandom";
    ru->out_dir = "./output/";
    if (SPHEROID_PROGRAM) {
        ru->r = 654;
        ru->soft = 4e-10;
        ru->Te = ru->Ti = 11600;
        ru->settle = 0.0;
        ru->geometry = POTG_PROLATE;
        ru->A = 5.0;
        ru->theta = 1.0;
        ru->mime = 1836.153;
    } else {
        ru->r = 1000;
        ru->soft = 2e-8;
        ru->Te = ru->Ti = 220;
        ru->settle = 5e-7;
        ru->geometry = POTG_SPHERE;
        ru->A = 1.0;
        ru->theta = 0.8;
        ru->mime = 1836.15;
    }
}

static void usage(const char* prog)
{
    fprintf(stderr, "Usage: %s [OPTIONS]\n", prog);
    fputs("  -a [SIZE]     grain radius in metres\n"
          "  -d [NUMBER]   time step in seconds\n"
          "  -E [NUMBER]   initial electron temperature in Kelvin\n"
          "  -e [NUMBER]   initial settling/equilibration time\n"
          "  -f [PATH]     write final simulation state to [PATH]\n"
          "  -g [PATH]     read initial simulation state from [PATH]\n"
          "  -I [NUMBER]   initial ion temperature in Kelvin\n"
          "  -i [INTEGER]  number of time steps to run for\n"
          "  -l            show runtime settings as command-line arguments\n"
          "  -m [SIZE]     pixel length in metres\n"
          "  -N [INTEGER]  number of particles to simulate\n"
          "  -R [PATH]     read out the simulation state at [PATH]\n"
          "  -r [INTEGER]  simulation radius in pixels\n"
          "  -s [NUMBER]   softening parameter in metres\n"
          "  -w [NUMBER]   limit on wall-time spent running in minutes\n"
          "  -x [NUMBER]   drift velocity in x-direction\n",
          stderr);
    if (SPHEROID_PROGRAM) fputs("  -A [NUMBER]   aspect ratio of the spheroid\n", stderr);
    fputs("  --theta X  --bx X --by X --bz X  --mime X  --no-grain  --dipoles  --fp64\n"
          "  --verlet  --euler-richardson  --bounce COE  --tune\n"
          "  --seed-file PATH  --output-dir DIR", stderr);
    fputs(SPHEROID_PROGRAM ? "  --oblate\n" : "\n", stderr);
}

static int bad(const char* msg)
{
    fprintf(stderr, "%s\n", msg);
    return 2;
}

/* 0: run; 1: show usage; 2: error already shown; 3: list settings */
static int parse_args(int argc, char** argv, pot_run* ru)
{
    int listing = 0, A_given = 0;
    int n;
    for (n = 1; n < argc; n++) {
        const char* o = argv[n];
        if (o[0] != '-') return 1;
        if (o[1] == '-') {
            const char* v = NULL;
            if (!strcmp(o, "--no-grain")) {
                ru->grain = 0;
                continue;
            }
            if (!strcmp(o, "--dipoles")) {
                ru->dipoles = 1;
                continue;
            }
            if (!strcmp(o, "--fp64")) {
                ru->precision = POTG_FP64;
                continue;
            }
            if (!strcmp(o, "--tune")) {
                ru->tune = 1;
                continue;
            }
            if (!strcmp(o, "--verlet")) {
                ru->integrator = POTG_VERLET;
                continue;
            }
            if (!strcmp(o, "--euler-richardson")) {
                ru->integrator = POTG_EULER_RICHARDSON;
                continue;
            }
            if (!strcmp(o, "--oblate") && SPHEROID_PROGRAM) {
                ru->geometry = POTG_OBLATE;
                continue;
            }
            if (++n >= argc) return 1;
            v = argv[n];
            if (!strcmp(o, "--theta")) {
                ru->theta = strtod(v, NULL);
                if (!(ru->theta >= 0.0)) return bad("Opening angle must be non-negative.");
            } else if (!strcmp(o, "--bx")) {
                ru->B[0] = strtod(v, NULL);
            } else if (!strcmp(o, "--by")) {
                ru->B[1] = strtod(v, NULL);
            } else if (!strcmp(o, "--bz")) {
                ru->B[2] = strtod(v, NULL);
            } else if (!strcmp(o, "--mime")) {
                ru->mime = strtod(v, NULL);
                if (!(ru->mime > 0.0)) return bad("Mass ratio must be positive.");
            } else if (!strcmp(o, "--bounce")) {
                ru->out_bounce = 1;
                ru->bounce_coe = strtod(v, NULL);
                if (!(ru->bounce_coe >= 0.0))
                    return bad("Coefficient of restitution must be non-negative.");
            } else if (!strcmp(o, "--seed-file")) {
                ru->seed_path = v;
            } else if (!strcmp(o, "--output-dir")) {
                ru->out_dir = v;
            } else {
                return 1;
            }
            continue;
        }
        switch (o[1]) {
        case 'a':
            if (++n >= argc) return 1;
            ru->a = strtod(argv[n], NULL);
            if (ru->a <= 0.0) return bad("Grain radius must be positive.");
            if (ru->a == HUGE_VAL) return bad("Given grain radius is too huge!");
            break;
        case 'A':
            if (!SPHEROID_PROGRAM) return 1;
            if (++n >= argc) return 1;
            ru->A = strtod(argv[n], NULL);
            A_given = 1;
            break;
        case 'd':
            if (++n >= argc) return 1;
            ru->dt = strtold(argv[n], NULL);
            if (ru->dt <= 0.0) return bad("Time step must be positive.");
            if (ru->dt == HUGE_VALL) return bad("Given time step is too huge!");
            break;
        case 'E':
            if (++n >= argc) return 1;
            ru->Te = strtold(argv[n], NULL);
            if (ru->Te < 0.0) return bad("Electron temperature must be non-negative.");
            if (ru->Te == HUGE_VALL) return bad("Given electron temperature is too huge!");
            break;
        case 'e':
            if (++n >= argc) return 1;
            ru->settle = strtold(argv[n], NULL);
            if (ru->settle < 0.0) return bad("Settling time must be non-negative.");
            if (ru->settle == HUGE_VALL) return bad("Given settling time is too huge!");
            break;
        case 'f':
            if (++n >= argc) return 1;
            ru->sw_path = argv[n];
            break;
        case 'g':
            if (++n >= argc) return 1;
            ru->sr_path = argv[n];
            break;
        case 'I':
            if (++n >= argc) return 1;
            ru->Ti = strtold(argv[n], NULL);
            if (ru->Ti < 0.0) return bad("Ion temperature must be non-negative.");
            if (ru->Ti == HUGE_VALL) return bad("Given ion temperature is too huge!");
            break;
        case 'i':
            if (++n >= argc) return 1;
            if ((ru->run_for = strtoul(argv[n], NULL, 10)) == 0)
                return bad("Number of time steps to run for must be a valid nonzero integer.");
            break;
        case 'l':
            listing = 1;
            break;
        case 'm':
            if (++n >= argc) return 1;
            ru->mpp = strtod(argv[n], NULL);
            if (ru->mpp <= 0.0) return bad("Pixel size must be positive.");
            if (ru->mpp == HUGE_VAL) return bad("Given pixel size is too huge!");
            break;
        case 'N':
            if (++n >= argc) return 1;
            if ((ru->N = strtoul(argv[n], NULL, 10)) == 0)
                return bad("Number of particles must be a valid nonzero integer.");
            break;
        case 'R':
            if (++n >= argc) return 1;
            ru->read_out = 1;
            ru->sr_path = argv[n];
            break;
        case 'r':
            if (++n >= argc) return 1;
            if ((ru->r = (unsigned int)strtoul(argv[n], NULL, 10)) == 0)
                return bad("Simulation radius in pixels must be a valid nonzero integer.");
            break;
        case 's':
            if (++n >= argc) return 1;
            ru->soft = strtod(argv[n], NULL);
            if (ru->soft < 0.0) return bad("Softening parameter must be non-negative.");
            if (ru->soft == HUGE_VAL) return bad("Given softening parameter is too huge!");
            break;
        case 'w':
            if (++n >= argc) return 1;
            ru->tim_lim = strtod(argv[n], NULL);
            if (ru->tim_lim < 0.0) return bad("Wall-time limit must be non-negative.");
            if (ru->tim_lim == HUGE_VAL) return bad("Wall-time limit is too huge!");
            break;
        case 'x':
            if (++n >= argc) return 1;
            ru->x_drift = strtod(argv[n], NULL);
            if (ru->x_drift == HUGE_VAL) return bad("Given drift velocity is too huge!");
            break;
        default:
            return 1;
        }
    }
    /* the aspect ratio is checked against the shape chosen (SPH:4800-4822) */
    if (ru->geometry == POTG_OBLATE && !A_given) ru->A = 0.5; /* SPH:227 */
    if (ru->geometry == POTG_PROLATE && ru->A <= 1.0)
        return bad("Prolate aspect ratio must be greater than 1.");
    if (ru->geometry == POTG_OBLATE && ru->A >= 1.0)
        return bad("Oblate aspect ratio must be less than 1.");
    return listing ? 3 : 0;
}

/* output_args_for_run, pot.c:4876-4915 (SPH adds -A) */
static void output_args(const char* prog, const pot_run* ru)
{
    printf("Settings: %s -N %lu -r %u", prog, ru->N, ru->r);
    if (SPHEROID_PROGRAM) printf(" -A %g", ru->A);
    printf(" -d %Lg -m %g -s %g -E %g -I %g -x %g -a %g -e %g", ru->dt, ru->mpp, ru->soft, ru->Te,
           ru->Ti, ru->x_drift, ru->a, ru->settle);
    if (ru->run_for) printf(" -i %lu", ru->run_for);
    if (ru->sw_path) printf(" -f %s", ru->sw_path);
    if (ru->sr_path) printf(ru->read_out ? " -R %s" : " -g %s", ru->sr_path);
    if (ru->tim_lim) printf(" -w %g", ru->tim_lim);
    putchar('\n');
}

/* read_out_sim_state, pot.c:4948-4961 */
static void read_out(uint64_t N, const potg_particle* parts)
{
    const char letters[] = "die";
    uint64_t k;
    puts("NO S X Y Z VX VY VZ");
    for (k = 0; k < N; k++) {
        const potg_particle* p = &parts[k];
        printf("%u \"%c\" %g %g %g %g %g %g\n", (unsigned)k, letters[p->species], p->x, p->y, p->z,
               p->vx, p->vy, p->vz);
    }
}

/* record_macro_this_step, pot.c:4364-4369 */
static int record_macro(const pot_run* ru, unsigned long t)
{
    int rec = t <= 200;
    if ((t * ru->dt) > (2.0 * ru->settle))
        rec |= !(t % 2000);
    else
        rec |= !(t % 200);
    return rec;
}

/* the run's wall-time limit (-w, minutes) has passed */
static int over_time(const pot_run* ru, time_t start)
{
    if (!(ru->tim_lim > 0.0)) return 0;
    const time_t now = time(NULL);
    if (now < start) {
        fputs("0: system time seems to've gone backwards, wall-time limit unenforceable!\n",
              stderr);
        return 0;
    }
    return (now - start) > (60 * ru->tim_lim);
}

/* Steps t+1, t+2, ... (at most PLAIN_MAX) before the next one that writes a
   macro record (record_macro), a state file (every 50000 steps, before the
   step) or the phi file (every 20000, after it), is tuned (--tune, step 3)
   or may warn (OUT_BOUNCE), or that passes -i. */
#define PLAIN_MAX 200ul
static unsigned long plain_steps(const pot_run* ru, unsigned long t)
{
    if (ru->out_bounce) return 0;
    unsigned long k = 0;
    while (k < PLAIN_MAX) {
        const unsigned long s = t + 1 + k;
        if ((ru->run_for && s > ru->run_for) || record_macro(ru, s) ||
            (ru->sw_path && !(s % 50000)) || !(s % 20000) || (ru->tune && s <= 3))
            break;
        k++;
    }
    return k;
}

/* key of the counter-based reinjection stream: a mix of the WELL512 record,
   so a run resumed from its state file draws the same numbers */
static uint64_t stream_key(const pot_prng* ps)
{
    const uint32_t* w = (const uint32_t*)ps;
    uint64_t h = 0x9e3779b97f4a7c15ull;
    size_t k;
    for (k = 0; k < sizeof(pot_prng) / 4; k++) {
        uint64_t z = h ^ w[k];
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        h = z ^ (z >> 31);
    }
    return h;
}

static void out_path(char* buf, size_t cap, const pot_run* ru, const char* name)
{
    snprintf(buf, cap, "%s%s", ru->out_dir, name);
}

/* write_phi_data_to_file, pot.c:3001-3043: the potential on a spherical
   grid about the grain, from the tree of the step just taken */
static void write_phi(const pot_run* ru, potg_ctx* ctx, const potg_particle* grain,
                      unsigned long t)
{
    enum { NPT = CELLS_R * CELLS_AZI * (CELLS_ZEN + 1) };
    static double xyz[3 * NPT], phi[NPT], rr[NPT], az[NPT], ze[NPT];
    const double r_step = ru->r * ru->mpp / CELLS_R;
    char path[4096];
    FILE* fp;
    int ri, ai, zi, k = 0;
    out_path(path, sizeof(path), ru, "pot-phi.dat");
    fp = fopen(path, "ab");
    if (!fp) {
        fprintf(stderr, "Can't open %s to write electric potential data.\n", path);
        return;
    }
    for (ri = 1; ri <= CELLS_R; ri++)
        for (ai = 0; ai < CELLS_AZI; ai++)
            for (zi = 0; zi <= CELLS_ZEN; zi++, k++) {
                rr[k] = r_step * ri;
                az[k] = 2.0 * PI_ * ai / CELLS_AZI;
                ze[k] = PI_ * zi / CELLS_ZEN;
                xyz[3 * k] = grain->x + (rr[k] * cos(az[k]) * sin(ze[k]));
                xyz[3 * k + 1] = grain->y + (rr[k] * sin(az[k]) * sin(ze[k]));
                xyz[3 * k + 2] = grain->z + (rr[k] * cos(ze[k]));
            }
    if (potg_phi_at(ctx, 0, xyz, NPT, phi) != POTG_OK) {
        fprintf(stderr, "0:%lu: potential evaluation failed: %s\n", t, potg_last_error());
        fclose(fp);
        return;
    }
    for (k = 0; k < NPT; k++)
        fprintf(fp, "%Lg %g %g %g %g\n", t * ru->dt, rr[k], az[k], ze[k], phi[k]);
    fclose(fp);
}

static potg_config device_config(const pot_run* ru, uint64_t key)
{
    potg_config c;
    memset(&c, 0, sizeof(c));
    c.N = ru->N;
    c.r = ru->r;
    c.geometry = (uint32_t)ru->geometry;
    c.dt = (double)ru->dt;
    c.mpp = ru->mpp;
    c.soft = ru->soft;
    c.x_drift = ru->x_drift;
    c.a = ru->a;
    c.settle = ru->settle;
    c.Te = ru->Te;
    c.Ti = ru->Ti;
    c.A = ru->A;
    c.theta = ru->theta;
    c.B[0] = ru->B[0];
    c.B[1] = ru->B[1];
    c.B[2] = ru->B[2];
    c.max_lin_dev = ru->geometry != POTG_SPHERE ? 0.1 : 0.01; /* pot.c:436, SPH:457 */
    c.mime = ru->mime;
    c.grain_enabled = ru->grain;
    c.sceptic = 1;
    c.seed = key;
    c.dipoles = ru->dipoles;
    c.integrator = ru->integrator;
    c.out_bounce = ru->out_bounce;
    c.bounce_coe = ru->bounce_coe;
    return c;
}

/* everything the host keeps between steps */
typedef struct host_run {
    pot_run ru;
    potg_ctx* ctx;
    potg_particle* parts;
    pot_sim_state st;
    FILE* fp_macro;
} host_run;

/* copy the device state back into h->parts and h->st */
static int fetch(host_run* h)
{
    potg_grain g;
    int k;
    if (potg_download(h->ctx, 0, h->parts, &g) != POTG_OK) {
        fprintf(stderr, "0: can't copy the particles back: %s\n", potg_last_error());
        return 1;
    }
    h->st.grain_q = g.q;
    h->st.grain_m = g.m;
    h->st.tot_absor_e = g.tot_absor_e;
    h->st.chi_b = g.chi_b;
    for (k = 0; k < 3; k++) h->st.epe[k] = g.epe[k];
    return 0;
}

static void macro_record(host_run* h, unsigned long t)
{
    static char line[4096];
    pot_macro_in in;
    int n;
    if (!h->fp_macro) {
        fprintf(stderr, "Macroscopic summary data file not open for writing (time step %lu).\n", t);
        return;
    }
    memset(&in, 0, sizeof(in));
    in.parts = h->parts;
    in.N = h->ru.N;
    in.r = h->ru.r;
    in.mpp = h->ru.mpp;
    in.dt = h->ru.dt;
    in.mime = h->ru.mime;
    in.grain_m = h->st.grain_m;
    in.grain_q = h->st.grain_q;
    in.tot_absor_e = h->st.tot_absor_e;
    memcpy(in.epe, h->st.epe, sizeof(in.epe));
    in.time_step = t;
    n = pot_macro_format(&in, line, sizeof(line));
    if (n > 0) {
        fputs(line, h->fp_macro);
        if (fflush(h->fp_macro)) fputs("Can't force write of macroscopic summary data to disk.\n",
                                       stderr);
    }
}

/* run(), pot.c:4259-4494; returns a process exit status */
static int run(host_run* h, uint64_t t_completed)
{
    const pot_run* ru = &h->ru;
    unsigned long t = (unsigned long)t_completed;
    const time_t start = time(NULL);
    int rec = 0, failed = 0;
    potg_step_out so;
    char path[4096];

    out_path(path, sizeof(path), ru, "pot-run-init-info.txt");
    pot_write_run_info(path, ru);
    {
        pot_run def;
        set_defaults(&def);
        def.geometry = ru->geometry;
        if (ru->geometry == POTG_OBLATE) def.A = 0.5;
        out_path(path, sizeof(path), ru, "pot-run-descrip.txt");
        pot_write_run_description(path, ru, &def);
    }
    while (1) {
        /* the coming steps after which nothing is written, tuned or warned
           of run as replays of a captured step, with no host round trip
           between them; they end where such a step comes */
        const unsigned long k = plain_steps(ru, t);
        if (k >= 2) {
            if (over_time(ru, start)) {
                t++;  /* as the one-step path leaves it: the step not taken */
                break;
            }
            if (potg_run(h->ctx, k) != POTG_OK) {
                fprintf(stderr, "0:%lu-%lu: %s\n", t + 1, t + k, potg_last_error());
                failed = 1;
                break;
            }
            t += k;
            rec = 0;
            continue;
        }
        t++;
        if (ru->run_for && t > ru->run_for) break;
        if (over_time(ru, start)) break;
        if (ru->sw_path && !(t % 50000)) {
            if (fetch(h)) {
                failed = 1;
                break;
            }
            h->st.t_completed = t - 1;
            pot_write_state(ru->sw_path, ru->N, h->parts, &h->st, 1);
        }
        rec = record_macro(ru, t);
        if (potg_step(h->ctx, rec, ru->out_bounce ? &so : NULL) != POTG_OK ||
            potg_sync(h->ctx) != POTG_OK) {
            fprintf(stderr, "0:%lu: %s\n", t, potg_last_error());
            failed = 1;
            break;
        }
        /* once the first steps have warmed the caches: the walk's launch
           shape that is fastest for this run */
        if (ru->tune && t == 3) {
            potg_tuning tt;
            if (potg_tune(h->ctx, &tt) != POTG_OK) {
                fprintf(stderr, "0:%lu: %s\n", t, potg_last_error());
                failed = 1;
                break;
            }
        }
        /* the reference warns of each such particle (pot.c:2724-2728) */
        if (ru->out_bounce && so.n_wall_inward)
            fprintf(stderr, "0: warning: %u particle(s) in outer wall travelling inwards.\n",
                    so.n_wall_inward);
        if (!(t % 20000)) write_phi(ru, h->ctx, &h->parts[0], t);
        if (rec) {
            if (fetch(h)) {
                failed = 1;
                break;
            }
            macro_record(h, t);
        }
    }
    if (!failed && ru->sw_path) {
        if (fetch(h)) return EXIT_FAILURE;
        h->st.t_completed = t - 1;
        pot_write_state(ru->sw_path, ru->N, h->parts, &h->st, 1);
    }
    if (!failed && !rec) {
        /* final record of a run that did not end on a record step */
        if (fetch(h)) return EXIT_FAILURE;
        macro_record(h, t - 1);
    }
    return failed ? EXIT_FAILURE : EXIT_SUCCESS;
}

int main(int argc, char** argv)
{
    host_run h;
    uint32_t seed[16];
    int rc;
    potg_config cfg;
    potg_grain g;
    char path[4096];

    memset(&h, 0, sizeof(h));
    set_defaults(&h.ru);
    if (argc < 2) fputs("Run without arguments; attempting to use default settings.\n", stderr);
    switch (parse_args(argc, argv, &h.ru)) {
    case 1:
        usage(argv[0]);
        return EXIT_FAILURE;
    case 2:
        return EXIT_FAILURE;
    case 3:
        output_args(argv[0], &h.ru);
        return EXIT_SUCCESS;
    default:
        break;
    }
    if (h.ru.a >= (h.ru.r * h.ru.mpp)) {
        fprintf(stderr, "Grain radius (%g m) must be less than simulation region radius (%g m).\n",
                h.ru.a, h.ru.r * h.ru.mpp);
        return EXIT_FAILURE;
    }
    h.parts = malloc(sizeof(potg_particle) * h.ru.N);
    if (!h.parts) {
        fprintf(stderr, "0: can't allocate %.1f MiB for %lu particles' data.\n",
                sizeof(potg_particle) * h.ru.N / 1048576.0, h.ru.N);
        return EXIT_FAILURE;
    }
    h.st.grain_q = 0;
    h.st.grain_m = DUST_MASS;
    if (h.ru.sr_path) {
        if (pot_read_state(h.ru.sr_path, h.ru.N, h.parts, &h.st)) {
            free(h.parts);
            return EXIT_FAILURE;
        }
        if (h.ru.read_out) {
            read_out(h.ru.N, h.parts);
            free(h.parts);
            return EXIT_SUCCESS;
        }
        if (h.ru.run_for && h.st.t_completed >= h.ru.run_for) {
            printf("0: simulation's already run for %lu steps, exiting.\n", h.ru.run_for);
            free(h.parts);
            return EXIT_SUCCESS;
        }
    } else {
        FILE* fp = fopen(h.ru.seed_path, "rb");
        memset(seed, 0, sizeof(seed));
        if (!fp || fread(seed, sizeof(uint32_t), 16, fp) != 16) {
            fprintf(stderr, "0: can't read PRNG seed from %s, PRNG may behave inadequately.\n",
                    h.ru.seed_path);
            memset(seed, 0, sizeof(seed));
        }
        if (fp) fclose(fp);
        h.st.t_completed = 0;
        h.st.tot_absor_e = 0;
        h.st.chi_b = 0.0;
        memset(h.st.epe, 0, sizeof(h.st.epe));
    }

    cfg = device_config(&h.ru, 0);
    if (!h.ru.sr_path) {
        if (pot_init_particles(&cfg, seed, h.parts, &h.st.prng)) {
            fputs("0: initial particle load failed.\n", stderr);
            free(h.parts);
            return EXIT_FAILURE;
        }
    }
    cfg.seed = stream_key(&h.st.prng);
    if (potg_create(&cfg, 1, h.ru.precision, &h.ctx) != POTG_OK) {
        fprintf(stderr, "0: can't set up the GPU: %s\n", potg_last_error());
        free(h.parts);
        return EXIT_FAILURE;
    }
    potg_set_counting(h.ctx, 0);  /* the program reports no interaction counts */
    memset(&g, 0, sizeof(g));
    g.q = h.st.grain_q;
    g.m = (double)h.st.grain_m;
    g.tot_absor_e = h.st.tot_absor_e;
    g.chi_b = h.st.chi_b;
    g.v[0] = h.parts[0].vx;
    g.v[1] = h.parts[0].vy;
    g.v[2] = h.parts[0].vz;
    for (rc = 0; rc < 3; rc++) g.epe[rc] = (double)h.st.epe[rc];
    if (potg_upload(h.ctx, 0, h.parts, &g, h.st.t_completed) != POTG_OK) {
        fprintf(stderr, "0: can't load the particles: %s\n", potg_last_error());
        potg_destroy(h.ctx);
        free(h.parts);
        return EXIT_FAILURE;
    }
    if (h.ru.N >= 6) {
        out_path(path, sizeof(path), &h.ru, "pot-macroscopic.dat");
        h.fp_macro = fopen(path, "ab");
        if (!h.fp_macro)
            fprintf(stderr, "Can't open %s to write macroscopic data.\n", path);
        else if (!ftell(h.fp_macro))
            fputs(pot_macro_header(), h.fp_macro);
    }
    rc = run(&h, h.st.t_completed);
    if (h.fp_macro) fclose(h.fp_macro);
    potg_destroy(h.ctx);
    free(h.parts);
    return rc;
}
