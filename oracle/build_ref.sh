#!/bin/bash
# Build switch variants of the reference treecode from its own sources, as
# they lie under /root/reference, into oracle/_ref/ (binaries only).
#
# The reference selects behaviour with #define switches written into its
# source (pot.c:25-217, pot-spheroid.c:25-230).  Each variant deletes the
# shipped #define lines that must change and supplies the wanted values
# with -D; the edited text is piped straight into the compiler, so no copy
# of a reference source is written anywhere.  The build line is the
# reference Makefile's (Makefile:9): mpicc -Os -std=c99 ... -lm, with gcc
# as MPICH's back end.  The PRNG seed is read from ./seed.bin in the
# directory the binary runs in (RANDOMNESS_SOURCE, pot.c:192).
#
# Usage: oracle/build_ref.sh [REFERENCE_DIR]
set -euo pipefail

REF=${1:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
MPICC=${MPICC:-/opt/conda/bin/mpicc}

if [ ! -f "$REF/pot.c" ]; then
    echo "build_ref: no reference sources at $REF; skipping" >&2
    exit 0
fi
if [ ! -x "$MPICC" ]; then
    echo "build_ref: $MPICC not found; the reference needs MPI; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"

# lines of the shipped pot.c that the variants replace
POT_DEL=(-e '/^#define DISABLE_GRAIN$/d' -e '/^#define DEBUG_LANGMUIR 0.1$/d'
         -e '/^#define OAP 0.8$/d' -e '/^#define EXTERNAL_B_[XYZ] 0.0$/d'
         -e '/^#define RANDOMNESS_SOURCE/d')
# lines of the shipped pot-spheroid.c that the variants replace (GL viewer
# and its GLFW include removed, as the image has no GLFW)
SPH_DEL=(-e '/^#define GLFW_INCLUDE_GLU/d' -e '/^#include <GLFW\/glfw.h>/d'
         -e '/^#include <GL\/glu.h>/d' -e '/^#define GL$/d' -e '/^#define OAP 1.0$/d'
         -e '/^#define EXTERNAL_B_[XYZ] 0.0$/d' -e '/^#define RANDOMNESS_SOURCE/d')

build_pot() {  # name theta bx by bz [extra -D...]; NO_BORIS=1: velocity Verlet
    local name=$1 theta=$2 bx=$3 by=$4 bz=$5
    shift 5
    local del=("${POT_DEL[@]}")
    if [ -n "${NO_BORIS:-}" ]; then del+=(-e '/^#define BORIS$/d'); fi
    sed "${del[@]}" "$REF/pot.c" |
        MPICH_CC=gcc "$MPICC" -x c - "$REF/prng.c" -I"$REF" -std=c99 -Os \
            -DSPHERICAL -DSCEPTIC_REINJECTION -DOAP="$theta" \
            -DEXTERNAL_B_X="$bx" -DEXTERNAL_B_Y="$by" -DEXTERNAL_B_Z="$bz" \
            -DRANDOMNESS_SOURCE="${SEED_SOURCE:-\"seed.bin\"}" "$@" -o "$OUT/$name" -lm
}

build_sph() {  # name shape theta bx by bz [extra -D...]; NO_BORIS=1: velocity Verlet;
              # NO_MACRO=1: no pot-macroscopic.dat (RECORD_MACRO_DATA, SPH:68)
    local name=$1 shape=$2 theta=$3 bx=$4 by=$5 bz=$6
    shift 6
    # the shipped file defines PROLATE (SPH:113); OBLATE replaces it
    local extra=() shape_def=()
    if [ "$shape" = OBLATE ]; then extra=(-e '/^#define PROLATE$/d'); shape_def=(-DOBLATE); fi
    if [ -n "${NO_BORIS:-}" ]; then extra+=(-e '/^#define BORIS$/d'); fi
    if [ -n "${NO_MACRO:-}" ]; then extra+=(-e '/^#define RECORD_MACRO_DATA /d'); fi
    tr -d '\r' < "$REF/pot-spheroid.c" | sed "${SPH_DEL[@]}" "${extra[@]}" |
        MPICH_CC=gcc "$MPICC" -x c - "$REF/prng.c" -I"$REF" -std=c99 -Os \
            "${shape_def[@]}" -DOAP="$theta" \
            -DEXTERNAL_B_X="$bx" -DEXTERNAL_B_Y="$by" -DEXTERNAL_B_Z="$bz" \
            -DRANDOMNESS_SOURCE="${SEED_SOURCE:-\"seed.bin\"}" "$@" -o "$OUT/$name" -lm
}

# spherical grain, SCEPTIC reinjection, Boris
build_pot pot_sphere 0.5 0.0 0.0 0.0
# the same seeding from /dev/urandom, as shipped, for runs on several ranks (one
# seed file would give every rank the same stream, so coincident particles)
SEED_SOURCE='"/dev/urandom"' build_pot pot_sphere_mpi 0.5 0.0 0.0 0.0
build_pot pot_sphere_bz 0.5 0.0 0.0 50.0
build_pot pot_sphere_bx 0.5 30.0 0.0 0.0
build_pot pot_sphere_t08 0.8 0.0 0.0 0.0
# dipole moments and the dipole walk (DIPOLES, pot.c:99-105)
build_pot pot_sphere_dip 0.5 0.0 0.0 20.0 -DDIPOLES
# the other integrators and containment (pot.c:113-132): velocity Verlet
# (BORIS undefined), Euler-Richardson, bouncing off the wall (OUT_BOUNCE)
NO_BORIS=1 build_pot pot_sphere_verlet_bz 0.5 0.0 0.0 20.0
NO_BORIS=1 build_pot pot_sphere_er_bounce_bx 0.5 30.0 0.0 0.0 -DEULER_RICHARDSON -DOUT_BOUNCE=0.95
build_pot pot_sphere_bounce 0.5 0.0 0.0 0.0 -DOUT_BOUNCE=0.95
# benchmark workloads (pot_amd/workloads.py): sphere_2m_bz (B along z,
# r_ge = lambda_D / 2) and prolate_4m_tilted_b (B 30 degrees off the axis)
build_pot pot_sphere_2m_bz 0.5 0.0 0.0 3.72363
# the same for timing under mpiexec: every rank seeds from /dev/urandom, as
# the shipped pot.c does.  (With one seed file every rank draws the same
# reinjection stream, so ranks reinject particles onto the same points and
# the next tree build never separates them.)
SEED_SOURCE='"/dev/urandom"' build_pot pot_sphere_2m_bz_mpi 0.5 0.0 0.0 3.72363
# spheroidal grain (pot-spheroid.c; prolate_4m_tilted_b has B 30 degrees off the x axis)
build_sph pot_prolate PROLATE 0.5 43.30127018922193 0.0 25.0
build_sph pot_oblate OBLATE 0.5 0.0 0.0 0.0
build_sph pot_prolate_4m PROLATE 0.5 1.61238 0.0 0.93091
NO_BORIS=1 build_sph pot_prolate_verlet PROLATE 0.5 0.0 0.0 0.0
# prolate_4m_tilted_b for timing under mpiexec (bench.py's CPU leg): ranks
# seed from /dev/urandom, and the macroscopic record -- which on every one of
# the first 200 steps adds a potential walk over all particles (pot.c:2771-2795,
# 4364) that the GPU's timed steps do not take -- is compiled out.  One variant
# runs the reference's own algorithm (monopoles, OAP 0.5), the other the
# setting of the headline (DIPOLES, OAP 0.14).
SEED_SOURCE='"/dev/urandom"' NO_MACRO=1 build_sph pot_prolate_4m_mpi PROLATE 0.5 1.61238 0.0 0.93091
SEED_SOURCE='"/dev/urandom"' NO_MACRO=1 build_sph pot_prolate_4m_dip_mpi PROLATE 0.14 1.61238 0.0 0.93091 -DDIPOLES

echo "build_ref: built $(ls "$OUT" | wc -l) reference variants in $OUT"
