// potg_push.hip -- Boris push, grain collection and reinjection on gfx950.
//
// One streaming pass over the particles due this step: Boris
// drift-kick-rotate-kick-drift in float64 (move_particles, pot.c:2472-2532),
// the grain-crossing test (pot.c:2641-2685 with does_particle_cross_grain,
// pot.c:2287-2325 / SPH:2448-2540), exact integer tallies of the absorbed
// charge and counts and the boundary test (pot.c:2071-2078); then, a
// wavefront per particle, SCEPTIC-like reinjection (pot.c:496-652) of the
// few that were absorbed or left.
// Reinjection draws come from a counter-based Philox4x32-10 stream keyed by
// (seed, particle, step), so the result does not depend on thread order.
#include "potg_grain.h"
#include "potg_internal.h"
#include "potg_prim.h"

namespace potg {

namespace {

constexpr int BT = 256;
inline uint32_t grid(uint32_t n) { return n ? (n + BT - 1) / BT : 1; }

// ---- counter-based stream -------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): published round
// multipliers and Weyl key increments.
__device__ inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// The stream of one particle within one step: counter (index, step lo,
// step hi, block), key = seed.  rnd(max) = max * u / (2^32 - 1) as prng.c:27.
struct Stream {
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t buf[4];
    int pos;
    __device__ Stream(uint64_t seed, uint32_t idx, uint64_t step)
    {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        ctr[0] = idx;
        ctr[1] = (uint32_t)step;
        ctr[2] = (uint32_t)(step >> 32);
        ctr[3] = 0;
        pos = 4;
    }
    __device__ double draw(double maximum)
    {
        if (pos == 4) {
            philox4x32_10(ctr, key, buf);
            ctr[3]++;
            pos = 0;
        }
        return maximum * ((double)buf[pos++] / 4294967295.0);
    }
};

// ---- SCEPTIC reinjection samplers (prng.c:138-214, pot.c:433-492) ----
__device__ double u_no_flow_icdf(double chi, double q)
{
    const double omc = 1.0 - chi;
    double p, u = 1.29951, usq;
    int it = 0;
    do {
        usq = u * u;
        p = 1.0 - ((usq + omc) * exp(-usq) / omc) - q;
        u -= p / (2.0 * u * exp(-usq) * (chi - usq) / -omc);
    } while ((p * p) > 1e-16 && ++it < 200);
    return u;
}

__device__ double u_icdf(double U, double chi, double q)
{
    const double Usq = U * U;
    const double psi = SQRT_PI * (1.0 + (2.0 * (Usq - chi))) / 2.0;
    const double p_den = 2.0 * ((U * exp(-Usq)) + (psi * erf(U)));
    double dp_den, dp_num, p, p_num, temp, u, usq;
    int it = 0;
    if (U > 2.1) u = U + (1.0 / U);
    else u = 1.2586147 + (0.221 * U) + (0.1903312 * U * U);
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
    } while ((p * p) > 1e-16 && ++it < 200);
    return u;
}

// lane j's double, in every lane
__device__ inline double readlane_d(double v, int j)
{
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// integrate_for_alpha (pot.c:433-492) by a whole wavefront, every lane
// with the same arguments and the same result.  The serial loop bisects a
// panel [xi0, xi2] until the integrand is linear within max_lin_dev at its
// midpoint: level l's midpoint is the previous level's halved towards xi0,
// and its test reads the integrand there and at the previous level.  Here
// lanes 0..IFA_LV-1 evaluate IFA_LV consecutive levels at once and the
// first level that passes is the panel; the midpoints, the integrand, the
// tests and the sum are the serial loop's float64 operations in its order,
// so the result is the serial one bit for bit.  The serial loop's guard of
// 4000000 bisections in all is kept.
constexpr int IFA_LV = 8;  // panels deeper than 7 levels are rare

__device__ double integrate_for_alpha(double s, double chi_b, double u2, double Rol,
                                      double max_lin_dev)
{
    const int lane = (int)__lane_id();
    const int kl = lane < IFA_LV ? lane : IFA_LV - 1;
    const double chu2 = chi_b / u2;
    const double s2 = s * s;
    const double y_one = s / sqrt(1.0 - chu2 - s2);
    double alpha = 0.0;
    double xi0 = 0.0, xi2 = 1.0;
    double y0 = s, y1, y2;
    long guard = 0;
    if (s == 0.0) return 0.0;
    do {
        // levels base .. base + IFA_LV - 1 below the panel top (top, ytop)
        double top = 1.0, ytop = y_one;
        for (long base = 0;; base += IFA_LV) {
            double xi1 = top;
            for (int j = 0; j <= kl; j++) xi1 = (xi0 + xi1) / 2.0;
            double y = 1.0 - (chu2 * xi1 * exp(Rol * (1.0 - (1.0 / xi1))));
            y = 1.0 / sqrt((y / s2) - (xi1 * xi1));
            const double xu = __shfl(xi1, lane > 0 ? lane - 1 : 0, WAVE);
            const double yu = __shfl(y, lane > 0 ? lane - 1 : 0, WAVE);
            const double xa = lane == 0 ? top : xu;   // the level above
            const double ya = lane == 0 ? ytop : yu;
            const uint64_t pass =
                ballot(lane < IFA_LV && !(fabs(ya - y - (y - y0)) > max_lin_dev));
            if (pass) {
                const int k = __ffsll((unsigned long long)pass) - 1;
                // level base + k took base + k bisections
                guard += base + k;
                if (guard > 4000000) return __builtin_nan("");
                xi2 = readlane_d(xa, k);
                y2 = readlane_d(ya, k);
                y1 = readlane_d(y, k);
                break;
            }
            if (guard + base + IFA_LV - 1 > 4000000) return __builtin_nan("");
            top = readlane_d(xi1, IFA_LV - 1);
            ytop = readlane_d(y, IFA_LV - 1);
        }
        alpha += (xi2 - xi0) * (y0 + (4.0 * y1) + y2) / 6.0;
        xi0 = xi2;
        y0 = y2;
    } while (xi2 < 1.0);
    return alpha;
}

// coord.c:21-35
__device__ inline void rot_xy(const double* d, double s, double c, double& x, double& y, double& z,
                              double offset)
{
    const double omc = 1.0 - c;
    double t0 = d[0] * ((d[0] * x) + (d[1] * y)) * omc;
    t0 += (x * c) + (d[1] * z * s);
    double t1 = d[1] * ((d[0] * x) + (d[1] * y)) * omc;
    t1 += (y * c) - (d[0] * z * s);
    double t2 = z * c;
    t2 += ((d[0] * y) - (d[1] * x)) * s;
    x = offset + t0;
    y = offset + t1;
    z = offset + t2;
}

// coord.c:45-59 (rotation by +dephi about b, the reference's sense)
__device__ inline void rot_dir(const double* d, double s, double c, double& x, double& y, double& z)
{
    const double omc = 1.0 - c;
    const double dot = (d[0] * x) + (d[1] * y) + (d[2] * z);
    double t0 = d[0] * dot * omc;
    t0 += (x * c) + (((d[1] * z) - (d[2] * y)) * s);
    double t1 = d[1] * dot * omc;
    t1 += (y * c) + (((d[2] * x) - (d[0] * z)) * s);
    double t2 = d[2] * dot * omc;
    t2 += (z * c) + (((d[0] * y) - (d[1] * x)) * s);
    x = t0;
    y = t1;
    z = t2;
}

// coord.c:6-14
__device__ inline void vel_cart(const double* v, double phi, double theta, double& vx, double& vy,
                                double& vz)
{
    vx = (v[0] * cos(phi) * sin(theta)) - (v[1] * sin(phi)) + (v[2] * cos(phi) * cos(theta));
    vy = (v[0] * sin(phi) * sin(theta)) + (v[1] * cos(phi)) + (v[2] * sin(phi) * cos(theta));
    vz = (v[0] * cos(theta)) - (v[2] * sin(theta));
}

// sceptically_reinject, pot.c:496-652.  The reference retries by recursion
// while alpha comes out NaN (pot.c:583-598), restoring chi_b as it does; here
// the retry is bounded.  When every try fails -- a chi_b that bars the
// species from entering, e.g. after a particle came within the softening
// length of chi_b's sample point, where the reference recurses without end --
// the particle is left as it was and false returned.
__device__ bool sceptic_reinject(const InstParams& p, int species, double chi_b_in, Stream& rs,
                                 double& x, double& y, double& z, double& vx, double& vy, double& vz)
{
    double a[2], alpha = 0.0, b = 0.0, c, co_z, phi, psi, q, U, u_mag = 1.0, u2 = 1.0, v_n = 0.0;
    double u[3] = {0.0, 0.0, 1.0}, v[3];
    double chi_b = chi_b_in;
    for (int tries = 0; tries < 1000; tries++) {
        v_n = 2.0 * KB / p.mass[species];
        if (species == 2) {
            v_n *= p.Te;
            chi_b *= -p.Ti / p.Te;
        } else {
            v_n *= p.Ti;
        }
        v_n = sqrt(v_n);
        if (p.x_drift) {
            U = p.x_drift / v_n;
            int it = 0;
            do {
                u_mag = u_icdf(U, chi_b, rs.draw(1.0));
                u2 = u_mag * u_mag;
            } while (u2 < chi_b && ++it < 100000);
            q = rs.draw(1.0);
            c = log(q + (1.0 - q) * exp(-4.0 * u_mag * U)) / (2.0 * u_mag * U);
            c += 1.0;
        } else {
            int it = 0;
            do {
                u_mag = u_no_flow_icdf(chi_b, rs.draw(1.0));
                u2 = u_mag * u_mag;
            } while (u2 < chi_b && ++it < 100000);
            c = rs.draw(2.0) - 1.0;
        }
        u[0] = 0.0;
        u[1] = 0.0;
        u[2] = u_mag;
        a[0] = 0.0;
        a[1] = 1.0;
        rot_xy(a, c, sqrt(1.0 - (c * c)), u[0], u[1], u[2], 0.0);
        a[0] = 1.0;
        a[1] = 0.0;
        psi = rs.draw(2.0 * PI);
        rot_xy(a, sin(psi), cos(psi), u[0], u[1], u[2], 0.0);
        b = sqrt(rs.draw(1.0 - (chi_b / u2)));
        alpha = integrate_for_alpha(b, chi_b, u2, p.Rol, p.max_lin_dev);
        if (!isnan(alpha)) break;
        if (species == 2) chi_b /= -p.Ti / p.Te;
    }
    if (isnan(alpha)) return false;
    phi = rs.draw(2.0 * PI);
    x = p.R * sin(alpha) * cos(phi);
    y = p.R * sin(alpha) * sin(phi);
    z = p.R * cos(alpha);
    a[0] = u[1];
    a[1] = -u[0];
    a[0] /= sqrt((u[0] * u[0]) + (u[1] * u[1]));
    a[1] /= sqrt((u[0] * u[0]) + (u[1] * u[1]));
    co_z = u[2] / u_mag;
    rot_xy(a, sin(acos(co_z)), co_z, x, y, z, p.R);
    v[1] = 0.0;
    v[2] = u_mag * b;
    v[0] = -sqrt(u2 - chi_b - (v[2] * v[2]));
    vel_cart(v, phi, alpha, vx, vy, vz);
    rot_xy(a, sin(acos(co_z)), co_z, vx, vy, vz, 0.0);
    vx *= v_n;
    vy *= v_n;
    vz *= v_n;
    return true;
}

// plain spherical reinjection at the boundary (inject, pot.c:710-765)
__device__ void plain_reinject(const InstParams& p, int species, Stream& rs, double& x, double& y,
                               double& z, double& vx, double& vy, double& vz)
{
    const double sigma = species == 1 ? sqrt(KB * p.Ti / p.mass[1]) : sqrt(KB * p.Te / p.mass[2]);
    const double r = p.R;
    const double phi = rs.draw(2.0 * PI);
    const double theta = acos(rs.draw(2.0) - 1.0);
    x = p.R + (r * sin(theta) * cos(phi));
    y = p.R + (r * sin(theta) * sin(phi));
    z = p.R + (r * cos(theta));
    double v[3], w;
    for (int k = 0; k < 2; k++) {
        const double urv2 = rs.draw(1.0);
        double urv1;
        int it = 0;
        do { urv1 = rs.draw(1.0); } while (urv1 == 0.0 && ++it < 64);
        const double rr = sqrt(-2.0 * log(urv1));
        const double th = 2.0 * PI * urv2;
        if (k == 0) {
            v[0] = sigma * rr * cos(th);
            v[1] = sigma * rr * sin(th);
        } else {
            w = sigma * rr * cos(th);
            v[2] = sigma * rr * sin(th);
            (void)w;
        }
    }
    if (v[0] > 0.0) v[0] = -v[0];
    vel_cart(v, phi, theta, vx, vy, vz);
    vx += p.x_drift;
}

// does_particle_cross_grain, pot.c:2287-2325 / SPH:2448-2540
__device__ bool crosses(const InstParams& p, double px, double py, double pz, const double* d)
{
    double s[3] = {p.grain[0] - px, p.grain[1] - py, p.grain[2] - pz};
    if (p.geometry != 0) {
        const double A = p.A, a = p.a;
        const double lim = a * a * pow(A, 4.0 / 3.0);
        if ((s[0] * s[0] + (s[1] * s[1] + s[2] * s[2]) * A * A) <= lim) return true;
        const double x_min = (d[1] * (d[1] * s[0] - d[0] * s[1]) + d[2] * (d[2] * s[0] - d[0] * s[2])) /
                             ((d[0] * d[0] / (A * A)) + d[1] * d[1] + d[2] * d[2]);
        const double x_prev = s[0] - d[0];
        if (((x_min > s[0]) && (x_min > x_prev)) || ((x_min < s[0]) && (x_min < x_prev))) return false;
        const double y_min = d[1] * (x_min - s[0]) / d[0] + s[1];
        const double z_min = d[2] * (x_min - s[0]) / d[0] + s[2];
        return ((x_min * x_min) + (y_min * y_min + z_min * z_min) * A * A) <= lim;
    }
    const double a = p.a;
    const double s2 = (s[0] * s[0]) + (s[1] * s[1]) + (s[2] * s[2]);
    if (s2 <= (a * a)) return true;
    const double d_pos2 = (d[0] * d[0]) + (d[1] * d[1]) + (d[2] * d[2]);
    const double d_dot_s = (d[0] * s[0]) + (d[1] * s[1]) + (d[2] * s[2]);
    const double alpha = d_dot_s / d_pos2;
    if ((alpha < -1.0) || (alpha > 0.0)) return false;
    return (s2 - (alpha * d_dot_s)) <= (a * a);
}

// OUT_BOUNCE containment of the spherical domain, check_particle_in_sim_area
// (pot.c:2036-2069): outside the wall and moving outwards, the velocity is
// reflected in spherical components (radial by -coe, the others by coe)
// and the particle goes back to where it started the step; outside and
// moving inwards it is left alone, and -1 reported (pot.c:2724-2728).
__device__ int bounce_off_wall(const InstParams& p, double& x, double& y, double& z, double& vx,
                               double& vy, double& vz, double px, double py, double pz)
{
    const double rad = p.R;
    const double dx = x - rad, dy = y - rad, dz = z - rad;
    const double r0 = sqrt(dx * dx + dy * dy + dz * dz);
    if (r0 > rad) {
        const double ph = atan2(dy, dx);
        const double th = acos(dz / r0);
        const double cp = cos(ph), sp = sin(ph), ct = cos(th), st = sin(th);
        double v[3];
        v[0] = vx * cp * st + vy * sp * st + vz * ct;
        v[1] = -vx * sp + vy * cp;
        v[2] = vx * cp * ct + vy * sp * ct - vz * st;
        if (!(v[0] > 0.0)) return -1;
        v[0] *= -p.bounce_coe;
        v[1] *= p.bounce_coe;
        v[2] *= p.bounce_coe;
        vel_cart(v, ph, th, vx, vy, vz);
        x = px;
        y = py;
        z = pz;
    }
    return 0;
}

// Velocity Verlet and Euler-Richardson, first half (pot.c:2532-2554): the
// acceleration at the start of the step (E and v x B), the velocity half
// step and the integrator's position update.  The intermediate state
// replaces the particle's state, as the reference's gather and broadcast
// replace its live array (pot.c:2557-2591); the acceleration and the start
// position are kept for the second half.
__global__ __launch_bounds__(BT) void k_half(const InstParams* __restrict__ P,
                                              const Globals* __restrict__ g, PState st,
                                              const uint32_t* __restrict__ targets,
                                              const double* __restrict__ E, double* __restrict__ a1,
                                              double* __restrict__ x0)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g->n_targets || step_failed(g)) return;
    const uint32_t i = targets[k];
    if (i == NONE) return;  // the gap before the next instance's targets
    const InstParams& p = P[st.inst[i]];
    const int s = st.spec[i];
    const double dts = p.dt * (double)p.tss[s];
    const double hdt = 0.5 * dts;
    const double qom = (QE * (s == 1 ? 1 : -1)) / p.mass[s];
    const double x = st.x[i], y = st.y[i], z = st.z[i];
    const double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
    const double ax = qom * (E[3 * k] + vy * p.B[2] - vz * p.B[1]);
    const double ay = qom * (E[3 * k + 1] + vz * p.B[0] - vx * p.B[2]);
    const double az = qom * (E[3 * k + 2] + vx * p.B[1] - vy * p.B[0]);
    x0[3 * k] = x;
    x0[3 * k + 1] = y;
    x0[3 * k + 2] = z;
    a1[3 * k] = ax;
    a1[3 * k + 1] = ay;
    a1[3 * k + 2] = az;
    st.vx[i] = vx + hdt * ax;
    st.vy[i] = vy + hdt * ay;
    st.vz[i] = vz + hdt * az;
    if (p.integrator == 2) {
        st.x[i] = x + hdt * vx;
        st.y[i] = y + hdt * vy;
        st.z[i] = z + hdt * vz;
    } else {
        st.x[i] = x + dts * (vx + hdt * ax);
        st.y[i] = y + dts * (vy + hdt * ay);
        st.z[i] = z + dts * (vz + hdt * az);
    }
}

__global__ __launch_bounds__(BT) void k_push(const InstParams* __restrict__ P, InstState* S,
                                              Globals* g, PState st,
                                              const uint32_t* __restrict__ targets,
                                              const double* __restrict__ E,
                                              const double* __restrict__ a1,
                                              const double* __restrict__ xs,
                                              uint32_t* __restrict__ queue,
                                              double* __restrict__ qprev, uint32_t qcap)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ntg = g->n_targets;
    const bool valid = k < ntg && !step_failed(g) && targets[k] != NONE;
    uint32_t inst = 0;
    if (valid) {
        const uint32_t i = targets[k];
        inst = st.inst[i];
        const InstParams& p = P[inst];
        const int s = st.spec[i];
        const uint64_t t = g->t;
        const double dts = p.dt * (double)p.tss[s];
        const double hdt = 0.5 * dts;
        const double qom = (QE * (s == 1 ? 1 : -1)) / p.mass[s];
        double x0, y0, z0, x, y, z, vx, vy, vz;
        if (p.integrator == 0) {
            // Boris (pot.c:2472-2531)
            x0 = st.x[i];
            y0 = st.y[i];
            z0 = st.z[i];
            vx = st.vx[i];
            vy = st.vy[i];
            vz = st.vz[i];
            const double xh = x0 + hdt * vx, yh = y0 + hdt * vy, zh = z0 + hdt * vz;
            const double ax = 0.5 * qom * E[3 * k], ay = 0.5 * qom * E[3 * k + 1],
                         az = 0.5 * qom * E[3 * k + 2];
            vx += dts * ax;
            vy += dts * ay;
            vz += dts * az;
            if (p.Bmag != 0.0) {
                const double dephi = 2.0 * atan(hdt * qom * p.Bmag);
                rot_dir(p.b, sin(dephi), cos(dephi), vx, vy, vz);
            }
            vx += dts * ax;
            vy += dts * ay;
            vz += dts * az;
            x = xh + hdt * vx;
            y = yh + hdt * vy;
            z = zh + hdt * vz;
        } else {
            // second half of velocity Verlet / Euler-Richardson
            // (pot.c:2598-2639): the field at the intermediate state, which
            // the state holds, and the first half's acceleration
            x0 = xs[3 * k];
            y0 = xs[3 * k + 1];
            z0 = xs[3 * k + 2];
            const double xm = st.x[i], ym = st.y[i], zm = st.z[i];
            const double vhx = st.vx[i], vhy = st.vy[i], vhz = st.vz[i];
            const double bx = qom * (E[3 * k] + vhy * p.B[2] - vhz * p.B[1]);
            const double by = qom * (E[3 * k + 1] + vhz * p.B[0] - vhx * p.B[2]);
            const double bz = qom * (E[3 * k + 2] + vhx * p.B[1] - vhy * p.B[0]);
            if (p.integrator == 2) {
                const double ax = a1[3 * k], ay = a1[3 * k + 1], az = a1[3 * k + 2];
                x = (xm - hdt * (vhx - hdt * ax)) + dts * vhx;
                y = (ym - hdt * (vhy - hdt * ay)) + dts * vhy;
                z = (zm - hdt * (vhz - hdt * az)) + dts * vhz;
                vx = (vhx - hdt * ax) + dts * bx;
                vy = (vhy - hdt * ay) + dts * by;
                vz = (vhz - hdt * az) + dts * bz;
            } else {
                x = xm;
                y = ym;
                z = zm;
                vx = vhx + hdt * bx;
                vy = vhy + hdt * by;
                vz = vhz + hdt * bz;
            }
        }

        // grain collection (pot.c:2641-2685) and the boundary test
        // (pot.c:2071-2078) or the bounce off the wall (pot.c:2036-2069);
        // the rare reinjections go to a queue so that the long SCEPTIC
        // sampling runs in wavefronts of its own (k_reinject)
        int fate = 0;  // 1: absorbed, 2: left the domain
        if (p.grain_enabled && ((double)t * p.dt) > p.settle) {
            const double d[3] = {x - x0, y - y0, z - z0};
            if (crosses(p, x, y, z, d)) {
                atomicAdd(&S[inst].acc_q, s == 1 ? 1 : -1);
                if (s == 2) atomicAdd(&S[inst].acc_e, 1u);
                else atomicAdd(&S[inst].acc_i, 1u);
                add_fixed128(&S[inst].acc_p_lo[0], &S[inst].acc_p_hi[0], p.mass[s] * vx);
                add_fixed128(&S[inst].acc_p_lo[1], &S[inst].acc_p_hi[1], p.mass[s] * vy);
                add_fixed128(&S[inst].acc_p_lo[2], &S[inst].acc_p_hi[2], p.mass[s] * vz);
                atomicAdd(&S[inst].n_reinj_abs, 1u);
                fate = 1;
            }
        }
        if (!fate) {
            if (p.out_bounce) {
                if (bounce_off_wall(p, x, y, z, vx, vy, vz, x0, y0, z0) < 0)
                    atomicAdd(&S[inst].n_wall_inward, 1u);
            } else {
                const double rad = p.R;
                const double dx = x - rad, dy = y - rad, dz = z - rad;
                if (sqrt(dx * dx + dy * dy + dz * dz) > rad) {
                    atomicAdd(&S[inst].n_reinj_out, 1u);
                    fate = 2;
                }
            }
        }
        if (fate) {
            const uint32_t q = atomicAdd(&g->n_reinj, 1u);
            if (q < qcap) {
                queue[q] = (i << 1) | (fate == 1 ? 1u : 0u);
                if (qprev) {
                    qprev[3 * (size_t)q] = x0;
                    qprev[3 * (size_t)q + 1] = y0;
                    qprev[3 * (size_t)q + 2] = z0;
                }
            }
        }
        st.x[i] = x;
        st.y[i] = y;
        st.z[i] = z;
        st.vx[i] = vx;
        st.vy[i] = vy;
        st.vz[i] = vz;
    }
}

// Reinjections queued by k_push: (slot << 1) | absorbed.  An absorbed
// particle is reinjected and then meets the boundary test like any other
// (pot.c:2676, 2693); one that left the domain is reinjected once.  The
// draws come from the particle's (seed, index, step) stream, so the order
// of the queue does not matter.  One wavefront per queued particle: the
// lanes repeat the sampling in step (same values in every lane) and share
// the integration; lane 0 stores.
__global__ __launch_bounds__(BT) void k_reinject(const InstParams* __restrict__ P, InstState* S,
                                                  Globals* g, PState st,
                                                  const uint32_t* __restrict__ queue,
                                                  const double* __restrict__ qprev, uint32_t qcap)
{
    const uint32_t n = step_failed(g) ? 0u : (g->n_reinj < qcap ? g->n_reinj : qcap);
    const uint32_t lane = __lane_id();
    const uint32_t nw = gridDim.x * (BT / WAVE);
    for (uint32_t e = blockIdx.x * (BT / WAVE) + threadIdx.x / WAVE; e < n; e += nw) {
        const uint32_t code = queue[e];
        const uint32_t i = code >> 1;
        const uint32_t inst = st.inst[i];
        const InstParams& p = P[inst];
        const int s = st.spec[i];
        Stream rs(p.seed, st.ref[i] - p.offset, g->t);
        const double chi_b = S[inst].chi_b;
        double x = st.x[i], y = st.y[i], z = st.z[i];
        double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
        bool ok = true;
        if (p.sceptic) ok = sceptic_reinject(p, s, chi_b, rs, x, y, z, vx, vy, vz);
        else plain_reinject(p, s, rs, x, y, z, vx, vy, vz);
        if ((code & 1u) && p.out_bounce) {
            // the absorbed particle's start position is where a bounce
            // returns it (pot.c:2686-2690)
            if (bounce_off_wall(p, x, y, z, vx, vy, vz, qprev[3 * (size_t)e],
                                qprev[3 * (size_t)e + 1], qprev[3 * (size_t)e + 2]) < 0 &&
                lane == 0)
                atomicAdd(&S[inst].n_wall_inward, 1u);
        } else if (code & 1u) {
            const double rad = p.R;
            const double dx = x - rad, dy = y - rad, dz = z - rad;
            if (ok && sqrt(dx * dx + dy * dy + dz * dz) > rad) {
                if (lane == 0) atomicAdd(&S[inst].n_reinj_out, 1u);
                if (p.sceptic) ok = sceptic_reinject(p, s, chi_b, rs, x, y, z, vx, vy, vz);
                else plain_reinject(p, s, rs, x, y, z, vx, vy, vz);
            }
        }
        if (!ok && !g->retire) {
            if (lane == 0) raise_err(g, KERR_REINJECT, st.ref[i]);
            continue;  // the particle keeps its state; the step reports the error
        }
        if (!ok) {
            // the instance retires: from the next step on none of its
            // particles is a target, so nothing of it moves again.  The
            // particle is parked at rest inside the grain, on the x axis at
            // a distance that its index makes unique (half the grain's
            // radius at most: inside the sphere and both spheroids), so the
            // instance's tree still builds
            if (lane == 0) {
                S[inst].retired = 1u;
                S[inst].retired_t = g->t;
                const double local = (double)(st.ref[i] - p.offset);
                x = p.R + 0.5 * p.a * (local + 1.0) / ((double)p.N + 1.0);
                y = p.R;
                z = p.R;
                vx = vy = vz = 0.0;
            }
        }
        if (lane == 0) {
            st.x[i] = x;
            st.y[i] = y;
            st.z[i] = z;
            st.vx[i] = vx;
            st.vy[i] = vy;
            st.vz[i] = vz;
        }
    }
}

__global__ void k_target_flags(const InstParams* __restrict__ P, const InstState* __restrict__ S,
                               PState st, uint32_t n_tree, uint32_t* __restrict__ flags)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tree) return;
    const uint32_t inst = st.inst[i];
    const uint32_t local = st.ref[i] - P[inst].offset;
    flags[i] = (st.spec[i] != 0 && local >= S[inst].start && !S[inst].retired) ? 1u : 0u;
}

// The step's targets grouped by instance, each instance's run starting at a
// multiple of 64: a wavefront's 64 targets then never span two instances,
// and an instance's tiles are the same whether it runs alone or in a batch,
// so its float32 sums -- whose grouping follows the tile's traversal -- are
// too.  One block: the padded run sizes are scanned over the instances,
// and each instance's gap is filled with NONE.
__global__ __launch_bounds__(1024) void k_target_bases(const InstParams* __restrict__ P,
                                                       uint32_t n_inst, uint32_t n_tree,
                                                       const uint32_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ tpos,
                                                       uint32_t* __restrict__ first,
                                                       uint32_t* __restrict__ base,
                                                       uint32_t* __restrict__ targets, Globals* g)
{
    __shared__ uint32_t sh[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t total = n_tree ? tpos[n_tree - 1] + flags[n_tree - 1] : 0u;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n_inst; c0 += 1024) {
        const uint32_t j = c0 + t;
        uint32_t cnt = 0, f = 0;
        if (j < n_inst) {
            // the instance's tree slots [tree_begin, tree_begin + tree_n)
            const uint32_t b0 = P[j].tree_begin, e0 = b0 + P[j].tree_n;
            f = b0 < n_tree ? tpos[b0] : total;
            cnt = (e0 < n_tree ? tpos[e0] : total) - f;
        }
        const uint32_t sz = (cnt + WAVE - 1) / WAVE * WAVE;
        sh[t] = sz;
        __syncthreads();
        for (uint32_t d = 1; d < 1024; d <<= 1) {
            const uint32_t v = t >= d ? sh[t - d] : 0u;
            __syncthreads();
            sh[t] += v;
            __syncthreads();
        }
        const uint32_t start = carry + sh[t] - sz;
        if (j < n_inst) {
            first[j] = f;
            base[j] = start;
            for (uint32_t q = start + cnt; q < start + sz; q++) targets[q] = NONE;
        }
        carry += sh[1023];
        __syncthreads();
    }
    if (t == 0) g->n_targets = carry;
}

__global__ void k_scatter_targets(PState st, uint32_t n_tree, const uint32_t* __restrict__ flags,
                                  const uint32_t* __restrict__ tpos,
                                  const uint32_t* __restrict__ first,
                                  const uint32_t* __restrict__ base, uint32_t* __restrict__ targets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tree || !flags[i]) return;
    const uint32_t inst = st.inst[i];
    targets[base[inst] + (tpos[i] - first[inst])] = i;
}

__global__ void k_finish(const InstParams* __restrict__ P, InstState* S, Globals* g, uint32_t n_inst,
                         int compute_epe, const double* __restrict__ epe_part,
                         const double* __restrict__ phi)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_inst) {
        InstState& s = S[k];
        const InstParams& p = P[k];
        if (compute_epe) {
            // electron and ion totals (pot.c:2775-2792), partials in a fixed order
            double se = 0.0, si = 0.0;
            for (int b = 0; b < EPE_BLOCKS; b++) {
                se += epe_part[2 * (k * EPE_BLOCKS + b)];
                si += epe_part[2 * (k * EPE_BLOCKS + b) + 1];
            }
            s.epe[2] = se;
            s.epe[1] = si;
        }
        if (p.grain_enabled) {
            // pot.c:2831-2836
            s.grain_m += (double)s.acc_i * p.mass[1] + (double)s.acc_e * p.mass[2];
            for (int c = 0; c < 3; c++)
                s.grain_v[c] += fixed128_value(s.acc_p_lo[c], s.acc_p_hi[c]) / s.grain_m;
            s.grain_q += s.acc_q;
            s.tot_e += s.acc_e;
            // the grain's own potential energy with its updated charge
            // (pot.c:2852-2863)
            if (compute_epe)
                s.epe[0] = s.grain_q ? ((double)s.grain_q * QE) * phi[s.grain_slot] : 0.0;
        }
        s.out_q = s.acc_q;
        s.out_e = s.acc_e;
        s.out_i = s.acc_i;
        // every index from the first stepped one to the last was pushed
        // (pot.c:2436-2440)
        // (none once the instance retired in an earlier step)
        const bool frozen = s.retired && s.retired_t < g->t;
        s.out_pushed = (s.start == NONE || frozen) ? 0u : p.N - s.start;
        s.out_reinj_abs = s.n_reinj_abs;
        s.out_reinj_out = s.n_reinj_out;
        s.out_wall_inward = s.n_wall_inward;
    }
    if (k == 0) {
        g->n_tgt_total += g->n_targets_real;
        g->n_steps += 1;
    }
}

// ---- upload / download --------------------------------------------------
struct AosParticle {
    double x, y, z, vx, vy, vz;
    int32_t species, pad;
};

__global__ void k_upload(PState st, const AosParticle* __restrict__ in, uint32_t n, uint32_t offset,
                         uint16_t inst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = offset + i;
    st.x[d] = in[i].x;
    st.y[d] = in[i].y;
    st.z[d] = in[i].z;
    st.vx[d] = in[i].vx;
    st.vy[d] = in[i].vy;
    st.vz[d] = in[i].vz;
    st.spec[d] = (uint8_t)in[i].species;
    st.inst[d] = inst;
    st.ref[d] = d;
}

__global__ void k_download(PState st, const uint32_t* __restrict__ inv, AosParticle* __restrict__ out,
                           uint32_t n, uint32_t offset)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = inv[offset + i];
    AosParticle r;
    r.x = st.x[j];
    r.y = st.y[j];
    r.z = st.z[j];
    r.vx = st.vx[j];
    r.vy = st.vy[j];
    r.vz = st.vz[j];
    r.species = st.spec[j];
    r.pad = 0;
    out[i] = r;
}

__global__ void k_inverse(PState st, uint32_t n, uint32_t* __restrict__ inv)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = st.ref[i];
    if (r < n) inv[r] = i;
}

// every slot starts as itself: reference layout, owned by its instance
__global__ void k_init_slots(PState st, const InstParams* __restrict__ P, uint32_t n_inst,
                             uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = n_inst - 1;
    while (lo < hi) {  // last instance whose offset <= i
        const uint32_t mid = (lo + hi + 1) / 2;
        if (P[mid].offset <= i) lo = mid; else hi = mid - 1;
    }
    st.x[i] = st.y[i] = st.z[i] = 0.0;
    st.vx[i] = st.vy[i] = st.vz[i] = 0.0;
    st.ref[i] = i;
    st.inst[i] = (uint16_t)lo;
    st.spec[i] = (i == P[lo].offset) ? 0 : 2;
}

// back to reference layout: slot r holds the particle whose reference slot is r
__global__ void k_unsort(PState src, PState dst, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = src.ref[i];
    if (r >= n) return;
    dst.x[r] = src.x[i];
    dst.y[r] = src.y[i];
    dst.z[r] = src.z[i];
    dst.vx[r] = src.vx[i];
    dst.vy[r] = src.vy[i];
    dst.vz[r] = src.vz[i];
    dst.ref[r] = r;
    dst.inst[r] = src.inst[i];
    dst.spec[r] = src.spec[i];
}

// ---- float64 direct sum (the force oracle on the GPU) --------------------
// Softened sum over every other particle of the instance, whatever order
// the particles are stored in (sources are gathered through inv).
__global__ __launch_bounds__(BT) void k_direct(const InstParams* __restrict__ P,
                                                const InstState* __restrict__ S, PState st,
                                                const uint32_t* __restrict__ inv, uint32_t inst,
                                                const uint32_t* __restrict__ tslots, uint32_t n,
                                                double* __restrict__ out)
{
    __shared__ double sx[BT], sy[BT], sz[BT];
    __shared__ int sq[BT];
    __shared__ uint32_t sslot[BT];
    const InstParams& p = P[inst];
    const int32_t gq = S[inst].grain_q;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n;
    const uint32_t self = valid ? tslots[k] : NONE;
    const double tx = valid ? st.x[self] : 0.0, ty = valid ? st.y[self] : 0.0,
                 tz = valid ? st.z[self] : 0.0;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (valid && p.geometry != 0) {
        double e3[3];
        grain_field(p, gq, tx, ty, tz, e3);
        ex = e3[0];
        ey = e3[1];
        ez = e3[2];
    }
    // the spheroid grain is analytic, not a point source (SPH:1516-1521)
    const uint32_t j_begin = p.geometry != 0 ? 1u : 0u;
    for (uint32_t j0 = j_begin; j0 < p.N; j0 += BT) {
        const uint32_t j = j0 + threadIdx.x;
        __syncthreads();
        if (j < p.N) {
            const uint32_t slot = inv[p.offset + j];
            sslot[threadIdx.x] = slot;
            sx[threadIdx.x] = st.x[slot];
            sy[threadIdx.x] = st.y[slot];
            sz[threadIdx.x] = st.z[slot];
            sq[threadIdx.x] = st.spec[slot] == 0 ? gq : (st.spec[slot] == 1 ? 1 : -1);
        } else {
            sslot[threadIdx.x] = NONE;
            sq[threadIdx.x] = 0;
        }
        __syncthreads();
        const uint32_t m = (p.N - j0) < BT ? (p.N - j0) : BT;
        for (uint32_t c = 0; c < m; c++) {
            const int q = sq[c];
            if (!q || sslot[c] == self) continue;
            const double dx = tx - sx[c], dy = ty - sy[c], dz = tz - sz[c];
            const double d2 = (dx * dx) + (dy * dy) + (dz * dz) + p.soft2;
            const double es = (QE * q) / ((K4PE0 * d2) * sqrt(d2));
            ex += es * dx;
            ey += es * dy;
            ez += es * dz;
        }
    }
    if (valid) {
        out[3 * k] = ex;
        out[3 * k + 1] = ey;
        out[3 * k + 2] = ez;
    }
}

}  // namespace

void launch_targets(Bufs& b, hipStream_t s)
{
    k_target_flags<<<grid(b.n_tree), BT, 0, s>>>(b.P, b.S, b.st, b.n_tree, b.flags);
    exclusive_scan(b.flags, b.tpos, b.n_tree, nullptr, &b.g->n_targets_real, nullptr, b.scan_tmp,
                   0, s);
    k_target_bases<<<1, 1024, 0, s>>>(b.P, b.n_inst, b.n_tree, b.flags, b.tpos, b.t_first,
                                      b.t_base, b.targets, b.g);
    k_scatter_targets<<<grid(b.n_tree), BT, 0, s>>>(b.st, b.n_tree, b.flags, b.tpos, b.t_first,
                                                    b.t_base, b.targets);
}

void launch_push(Bufs& b, hipStream_t s)
{
    k_push<<<grid(b.t_cap), BT, 0, s>>>(b.P, b.S, b.g, b.st, b.targets, b.E, b.a1, b.x0, b.queue,
                                        b.qprev, b.n_tree);
    k_reinject<<<256, BT, 0, s>>>(b.P, b.S, b.g, b.st, b.queue, b.qprev, b.n_tree);
}

void launch_half(Bufs& b, hipStream_t s)
{
    k_half<<<grid(b.t_cap), BT, 0, s>>>(b.P, b.g, b.st, b.targets, b.E, b.a1, b.x0);
}

void launch_finish_step(Bufs& b, int compute_epe, hipStream_t s)
{
    k_finish<<<grid(b.n_inst), BT, 0, s>>>(b.P, b.S, b.g, b.n_inst, compute_epe, b.epe_part, b.phi);
}

void launch_scatter_upload(Bufs& b, uint32_t inst, const double* aos, uint32_t n, hipStream_t s)
{
    uint32_t offset = 0;
    (void)hipMemcpyAsync(&offset, &b.P[inst].offset, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    k_upload<<<grid(n), BT, 0, s>>>(b.st, (const AosParticle*)aos, n, offset, (uint16_t)inst);
}

void launch_gather_download(Bufs& b, uint32_t inst, double* aos, uint32_t n, hipStream_t s)
{
    launch_build_inverse(b, s);
    uint32_t offset = 0;
    (void)hipMemcpyAsync(&offset, &b.P[inst].offset, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    k_download<<<grid(n), BT, 0, s>>>(b.st, b.inv, (AosParticle*)aos, n, offset);
}

void launch_build_inverse(Bufs& b, hipStream_t s)
{
    k_inverse<<<grid(b.n_total), BT, 0, s>>>(b.st, b.n_total, b.inv);
}

void launch_init_slots(Bufs& b, hipStream_t s)
{
    k_init_slots<<<grid(b.n_total), BT, 0, s>>>(b.st, b.P, b.n_inst, b.n_total);
    k_inverse<<<grid(b.n_total), BT, 0, s>>>(b.st, b.n_total, b.inv);
}

void launch_unsort(Bufs& b, hipStream_t s)
{
    k_unsort<<<grid(b.n_total), BT, 0, s>>>(b.st, b.alt, b.n_total);
    PState t = b.st;
    b.st = b.alt;
    b.alt = t;
}

void launch_direct_sum(Bufs& b, uint32_t inst, uint32_t n, hipStream_t s)
{
    k_direct<<<grid(n), BT, 0, s>>>(b.P, b.S, b.st, b.inv, inst, b.q_self, n, b.q_out);
}

}  // namespace potg
