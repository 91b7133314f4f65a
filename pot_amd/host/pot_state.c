/*
 * pot_state.c -- simulation state files of the drop-in programs
 * (write_sim_state_to_file / read_sim_state_from_file, pot.c:3816-4019).
 *
 * Format C: one text line
 *     C <workers> <grain q> <grain m> <completed steps> <epe0> <epe1> <epe2>
 *       <electrons absorbed> <chi_b>
 * then one 84-byte WELL512 record per worker and the N 56-byte particle
 * records in index order.  The programs here are one worker, so they write
 * one record and accept files written by one worker (pot.c:3968-3973).
 * Formats A and B, written by older versions, are read too.
 */
#include <stdio.h>
#include <string.h>

#include "pot_host.h"

int pot_write_state(const char* path, uint64_t N, const potg_particle* parts,
                    const pot_sim_state* s, int sceptic)
{
    FILE* fp = fopen(path, "wb");
    int bad = 0;
    if (!fp) {
        fprintf(stderr, "Failed to open simulation state file %s for writing.\n", path);
        return 1;
    }
    /* without SCEPTIC reinjection the reference stores 9999.99 (pot.c:3825-3829) */
    fprintf(fp, "C %u %i %Lg %lu %Lg %Lg %Lg %u %g\n", 1u, s->grain_q, s->grain_m,
            (unsigned long)s->t_completed, s->epe[0], s->epe[1], s->epe[2], s->tot_absor_e,
            sceptic ? s->chi_b : 9999.99);
    if (fwrite(&s->prng, sizeof(pot_prng), 1, fp) != 1) {
        fprintf(stderr, "Failed to write proc. 0's PRNG state to %s.\n", path);
        bad = 1;
    }
    if (fwrite(parts, sizeof(potg_particle), N, fp) != N) {
        fprintf(stderr, "Failed to write particle states to %s.\n", path);
        bad = 1;
    }
    if (fclose(fp)) {
        fprintf(stderr, "Failed to close simulation state output file %s.\n", path);
        bad = 1;
    }
    return bad;
}

int pot_read_state(const char* path, uint64_t N, potg_particle* parts, pot_sim_state* s)
{
    FILE* fp = fopen(path, "rb");
    unsigned int workers = 0;
    unsigned long t = 0;
    int first, ok;
    if (!fp) {
        fprintf(stderr, "0: failed to open simulation state file %s to read.\n", path);
        return 1;
    }
    first = getc(fp);
    if (first == 'C') {
        ok = fscanf(fp, " %u %i %Lg %lu %Lg %Lg %Lg %u %lg\n", &workers, &s->grain_q, &s->grain_m,
                    &t, &s->epe[0], &s->epe[1], &s->epe[2], &s->tot_absor_e, &s->chi_b) == 9;
        if (!ok) fprintf(stderr, "0: failed to read 9 assorted data values from %s.\n", path);
    } else if (first == 'B') {
        ok = fscanf(fp, " %u %i %lu %Lg %Lg %Lg %u\n", &workers, &s->grain_q, &t, &s->epe[0],
                    &s->epe[1], &s->epe[2], &s->tot_absor_e) == 7;
        if (!ok) fprintf(stderr, "0: failed to read 7 assorted data values from %s.\n", path);
    } else if (first >= '0' && first <= '9') {
        rewind(fp);
        ok = fscanf(fp, "%u %i %lu %Lg %Lg %Lg\n", &workers, &s->grain_q, &t, &s->epe[0],
                    &s->epe[1], &s->epe[2]) == 6;
        if (!ok) fprintf(stderr, "0: failed to read 6 assorted data values from %s.\n", path);
    } else {
        fprintf(stderr, "0: alleged state file %s isn't.\n", path);
        ok = 0;
    }
    if (ok && workers != 1) {
        fprintf(stderr, "0: conflicting worker process counts (%u vs. %u).\n", workers, 1u);
        ok = 0;
    }
    if (ok && fread(&s->prng, sizeof(pot_prng), 1, fp) != 1) {
        fprintf(stderr, "0: failed to read proc. 0's PRNG state from %s.\n", path);
        ok = 0;
    }
    if (ok && s->prng.index >= 16) {
        /* sanitize_prng, prng.c:117-125 */
        s->prng.index %= 16;
        fprintf(stderr, "0: read invalid PRNG state index, attempting to sanitize it.\n");
    }
    if (ok && fread(parts, sizeof(potg_particle), N, fp) != N) {
        fprintf(stderr, "0: failed to read particle states from %s.\n", path);
        ok = 0;
    }
    if (fclose(fp) && ok) {
        fprintf(stderr, "0: failed to close simulation state file %s.\n", path);
        ok = 0;
    }
    s->t_completed = t;
    return ok ? 0 : 1;
}
