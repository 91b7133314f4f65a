// potg_dev.h -- device-side data layout and helpers shared by the gfx950
// kernels of the pot hot path.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace potg {

// Physical constants of the reference (pot.c:278-282; PI from prng.h:9-14,
// which pot.c includes before its own definition).
constexpr double KB = 1.38065e-23;
constexpr double EPSILON0 = 8.85419e-12;
constexpr double QE = 1.6021766e-19;
constexpr double ME = 9.109383e-31;
constexpr double PI = 3.14159265358979;
constexpr double SQRT_PI = 1.77245385090552;
// (4 * PI * EPSILON0) as the reference's expressions fold it (pot.c:1738, 1792)
constexpr double K4PE0 = (4.0 * PI) * EPSILON0;

constexpr int MAXD = 60;          // deepest leaf supported (the reference's 400-entry
                                  // walk stack allows about 57 levels, pot.c:194-196)
constexpr uint32_t LEAF = 0xFFFFFFFFu;   // NodeRec.cnt marker of a leaf
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t LEAF_BIT = 0x80000000u; // child-slot code of a leaf
constexpr int WAVE = 64;
constexpr int EPE_BLOCKS = 64;    // partial sums per instance of the PE reduction
// tree moments: blocks of the depth counting sort, and the band of depths
// given a launch per level (the others are done by one block)
constexpr uint32_t LVL_BLOCKS = 256;
constexpr int LVL_WIDE_LO = 12;   // depths >= this: one block
constexpr int LVL_WIDE_HI = 3;    // depths <= this: one block

// error codes raised by kernels (mirrors POTG_ERR_*)
constexpr int KERR_OUT_OF_BOX = 1;
constexpr int KERR_COINCIDENT = 2;
constexpr int KERR_DEPTH = 3;
constexpr int KERR_CAPACITY = 4;
constexpr int KERR_REINJECT = 5;

// Constant parameters of one instance (one grain simulation).
struct InstParams {
    double R;          // r * mpp: domain radius and centre coordinate
    double root_x2;    // 2.0 * r * mpp: upper face of the root box (pot.c:1476)
    double invR;       // 1 / R, for the float32 walk's normalised coordinates
    double dt;
    double soft2;      // soft * soft
    double theta2;     // OAP * OAP
    double open32;     // 1 / (theta2 R^2): float32 records store size2 * open32
    double a, A;
    double settle;
    double x_drift, Te, Ti;
    double B[3];
    double Bmag;
    double b[3];
    double max_lin_dev;
    double mass[3];    // dust, ion, electron
    double Rol;        // R / shielding length of integrate_for_alpha (pot.c:578-580)
    double grain[3];   // grain position (parts[0]), fixed
    double chi_pt[3];  // boundary point where chi_b is sampled (pot.c:2271-2275)
    double kTi_over_e; // KB * Ti / (CHARGE[ION] * QE) (pot.c:2278)
    double Nm1;        // N - 1 as a double
    uint32_t tss[3];
    uint32_t geometry;
    int32_t grain_enabled;
    int32_t sceptic;
    uint64_t seed;
    uint32_t offset;     // first global slot of this instance
    uint32_t N;          // particles including the grain
    uint32_t tree_begin; // first sorted slot of the instance's tree particles
    uint32_t tree_n;     // tree particles (N, or N - 1 when the grain is analytic)
    uint32_t first_idx[3]; // lowest reference index of each species
    int32_t dipoles;       // DIPOLES: dipole moments in the tree and the walks
    int32_t integrator;    // POTG_BORIS / POTG_VERLET / POTG_EULER_RICHARDSON
    int32_t out_bounce;    // OUT_BOUNCE: bounce off the wall
    double bounce_coe;     // its coefficient of restitution
};

// Mutable per-instance state on the device.
struct InstState {
    int32_t grain_q;
    uint32_t tot_e;
    double grain_m;
    double grain_v[3];
    double chi_b;
    double epe[3];
    uint32_t start;      // first stepped reference index of the current step
    uint32_t root;       // walk id of the tree root
    uint32_t grain_slot; // sorted slot of the grain (index 0) after the step's sort
    // per-step accumulators (cleared by the step's first kernel)
    int32_t acc_q;
    uint32_t acc_e, acc_i, n_pushed, n_reinj_abs, n_reinj_out, n_wall_inward;
    // absorbed momentum per axis, an exact sum in 128-bit two's-complement
    // fixed point (units of 2^-140 kg m/s; see add_fixed128)
    unsigned long long acc_p_lo[3];
    unsigned long long acc_p_hi[3];
    // outputs of the last step
    int32_t out_q;
    uint32_t out_e, out_i, out_pushed, out_reinj_abs, out_reinj_out, out_wall_inward;
    // retired by a failed reinjection (potg_set_retire): no longer stepped;
    // retired_t is the step in which it failed
    uint32_t retired;
    uint32_t retired_pad;
    unsigned long long retired_t;
};

// Context-wide device counters.
struct Globals {
    uint64_t t;            // steps completed
    uint32_t n_targets;    // target slots of the current step: each instance's
                           // targets start at a multiple of 64, the gaps hold NONE
    uint32_t n_internal;   // internal tree nodes
    uint32_t n_children;   // sum of child counts
    uint32_t tile_ctr;     // persistent-walk work counter
    uint32_t tile_ctr2;
    uint32_t n_reinj;      // reinjection queue length of the current step
    uint32_t n_targets_real; // particles stepped this step (targets without the gaps)
    int32_t err;
    uint32_t err_idx;
    uint32_t max_depth;
    unsigned long long n_cell, n_leaf, n_open, n_tgt_total, n_steps;
    unsigned long long n_pops;   // child ranges fetched by the walk (per wavefront)
    unsigned long long n_visits; // children visited by float32-walk wavefronts
    uint32_t retire;             // a failed reinjection retires its instance (1) or
                                 // fails the step (0, the default)
};

// Tree node records, children contiguous: an internal node's children are
// the records [first, first + cnt), ordered as the reference orders them.
struct __align__(16) NodeRec64 {
    double cx, cy, cz, size2;
    int32_t q;
    uint32_t first;
    uint32_t cnt;   // LEAF for a leaf
    uint32_t part;  // leaf: sorted slot of its particle
};
struct __align__(16) NodeRec32 {
    float cx, cy, cz;          // relative to the domain centre, in units of R
    float size2;               // box size squared over theta^2, in units of R^2
    float q;                   // net charge (exact: |q| < 2^24)
    uint32_t first;
    uint32_t cnt;
    uint32_t part;
};

// Child slot for a position against a node's midpoints, numbered in the
// reference's child order gxlylz, gxlygz, lxlylz, lxlygz, lxgylz, lxgygz,
// gxgylz, gxgygz (pot.c:324-331, 1525-1553).
__host__ __device__ inline int ref_digit(bool gx, bool gy, bool gz)
{
    const int g = gx ? (gy ? 3 : 0) : (gy ? 2 : 1);
    return 2 * g + (gz ? 1 : 0);
}

// One level of the reference's float64 midpoint recurrence (pot.c:1522-1524,
// 1261-1287): returns the child digit and narrows the box to that child.
__host__ __device__ inline int descend1(double px, double py, double pz, double& x1, double& x2,
                                        double& y1, double& y2, double& z1, double& z2)
{
    const double mx = (x1 + x2) / 2.0;
    const double my = (y1 + y2) / 2.0;
    const double mz = (z1 + z2) / 2.0;
    const bool gx = !(px < mx), gy = !(py < my), gz = !(pz < mz);
    if (gx) x1 = mx; else x2 = mx;
    if (gy) y1 = my; else y2 = my;
    if (gz) z1 = mz; else z2 = mz;
    return ref_digit(gx, gy, gz);
}

__device__ inline uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t rfl64(uint64_t v)
{
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Momentum tally of absorbed particles.  Float64 atomics would add in the
// order threads arrive, so the grain's velocity would differ between runs
// (and between a captured step and an eager one).  Each contribution p is
// instead scaled by 2^140 (exact), truncated to an integer and added to a
// 128-bit two's-complement accumulator with two 64-bit integer atomics:
// integer addition is associative, so the sum is exact and independent of
// the order.  |p| < 2^-13 kg m/s keeps the total below 2^127; the unit,
// 2^-140 kg m/s (about 7e-43), is far below an electron's momentum at any
// plasma speed (about 1e-25 at 1e5 m/s).
__device__ inline void add_fixed128(unsigned long long* lo, unsigned long long* hi, double p)
{
    const double x = fabs(ldexp(p, 140));
    const double xh = floor(x * 0x1p-64);
    // exact: x - xh 2^64 is a multiple of ulp(x) below 2^64
    const double xl = x - xh * 0x1p64;
    unsigned long long h = (unsigned long long)xh, l = (unsigned long long)xl;
    if (p < 0.0) {
        l = ~l + 1ull;
        h = ~h + (l == 0ull ? 1ull : 0ull);
    }
    const unsigned long long old = atomicAdd(lo, l);
    atomicAdd(hi, h + (old + l < old ? 1ull : 0ull));
}

// the accumulated 128-bit sum as a double (kg m/s)
__device__ inline double fixed128_value(unsigned long long lo, unsigned long long hi)
{
    const bool neg = (long long)hi < 0;
    if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0ull ? 1ull : 0ull);
    }
    const double mag = ((double)hi * 0x1p64 + (double)lo) * 0x1p-140;
    return neg ? -mag : mag;
}

// a kernel earlier in the step raised an error: the walk, push and
// reinjection kernels after it skip their work rather than compute with
// the state that raised it (NaN or out-of-box positions, coincident
// particles)
__device__ inline bool step_failed(const Globals* g)
{
    return *(volatile const int32_t*)&g->err != 0;
}

__device__ inline void raise_err(Globals* g, int code, uint32_t idx)
{
    if (atomicCAS(&g->err, 0, code) == 0) g->err_idx = idx;
}

}  // namespace potg
