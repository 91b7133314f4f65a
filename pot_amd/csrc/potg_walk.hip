// potg_walk.hip -- Barnes-Hut field and potential walks on gfx950.
//
// One wavefront walks the tree for 64 targets that are neighbours in the
// sort order.  The walk keeps the reference's per-particle opening test
// (field_particle_feels_at, pot.c:1718-1806): the wave shares one
// depth-first traversal whose stack entries carry a 64-bit lane mask; at a
// cell every lane in the mask applies its own MAC, lanes that accept it add
// the monopole, and lanes that reject it pass the children on with their
// mask.  Each lane therefore meets exactly the nodes, in exactly the order,
// that the reference visits for that particle.
//
// Memory: the children of a node are contiguous records, so a stack entry
// is a child range.  A range is fetched with one vector load (lane j holds
// child j) and its children are then visited from registers or LDS -- one
// memory round trip per opened node instead of one per child.  Ranges
// interrupted by a descent are pushed back (remaining siblings, same mask)
// on an LDS stack.
//
// FP64: every operation repeats the reference's float64 expressions, with
// float64 instead of x87 accumulation.  MIXED: coordinates relative to the
// domain centre in units of R, float32 arithmetic with a hardware reciprocal
// square root, Kahan-compensated per-particle sums, result scaled back in
// float64.
#include <atomic>

#include "potg_grain.h"
#include "potg_internal.h"

namespace potg {

namespace {

constexpr int BT = 256;
constexpr int NW = BT / WAVE;
constexpr int STK = MAXD + 4;

// KIND_EPE: the potential a particle feels, its own leaf skipped
// (phi_particle_feels, pot.c:1934-2018); KIND_PHI: the potential at a
// point, every leaf counted (phi_at, pot.c:1818-1901)
enum { KIND_FIELD = 0, KIND_PHI = 1, KIND_EPE = 2 };
// MODE_EPE: every particle slot, at the position the potential-energy sum
// reads (half-drifted when stepped this step, pot.c:2478-2480, 2784-2792)
enum { MODE_STEP = 0, MODE_POINTS = 1, MODE_EPE = 2 };

struct WalkIO {
    const uint32_t* targets;
    const uint32_t* n_targets_dev;
    const double* pts;
    const uint32_t* pt_inst;
    const uint32_t* pt_self;
    uint32_t n_pts;
    double* out;       // out: result; in, when acc is set: the grain term to add to
    int32_t acc;       // 0: the walk starts from zero and never reads out
    uint32_t* counts;
    uint32_t* tile_ctr;
    uint32_t n_rec;    // record capacity: child ranges must stay inside it
    const InstState* S;
    const double* dip64;   // record dipoles (DIPOLES instances), else null
    const float4* dip32;
};

__device__ inline void kahan_add(float& s, float& c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ inline float bcast_f(float v, uint32_t j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline uint32_t bcast_u(uint32_t v, uint32_t j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}
__device__ inline double bcast_d(double v, uint32_t j)
{
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// target of lane k: the half-drifted particle (Boris, pot.c:2478-2480), the
// particle where the state has it (velocity Verlet / Euler-Richardson, whose
// fields are taken at the start and the intermediate state, pot.c:2483-2485,
// 2603-2605), or a point
template <int MODE>
__device__ inline bool load_target(const InstParams* __restrict__ P, PState st, const WalkIO& io,
                                   uint32_t k, uint32_t ntg, uint32_t& inst, uint32_t& self,
                                   double& tx, double& ty, double& tz)
{
    inst = 0;
    self = NONE;
    tx = ty = tz = 0.0;
    if (k >= ntg) return false;
    if (MODE == MODE_STEP) {
        const uint32_t i = io.targets[k];
        if (i == NONE) return false;  // the gap before the next instance's targets
        self = i;
        inst = st.inst[i];
        if (P[inst].integrator == 0) {
            const double dts = P[inst].dt * (double)P[inst].tss[st.spec[i]];
            tx = st.x[i] + (0.5 * dts) * st.vx[i];
            ty = st.y[i] + (0.5 * dts) * st.vy[i];
            tz = st.z[i] + (0.5 * dts) * st.vz[i];
        } else {
            tx = st.x[i];
            ty = st.y[i];
            tz = st.z[i];
        }
    } else if (MODE == MODE_EPE) {
        self = k;
        inst = st.inst[k];
        const int sp = st.spec[k];
        // Boris half-drifted the stepped particles in place; the other
        // integrators' state already holds the intermediate state
        const bool stepped = sp != 0 && P[inst].integrator == 0 &&
                             st.ref[k] - P[inst].offset >= io.S[inst].start;
        const double hdt = stepped ? 0.5 * (P[inst].dt * (double)P[inst].tss[sp]) : 0.0;
        tx = stepped ? st.x[k] + hdt * st.vx[k] : st.x[k];
        ty = stepped ? st.y[k] + hdt * st.vy[k] : st.y[k];
        tz = stepped ? st.z[k] + hdt * st.vz[k] : st.z[k];
    } else {
        inst = io.pt_inst[k];
        self = io.pt_self[k];
        tx = io.pts[3 * k];
        ty = io.pts[3 * k + 1];
        tz = io.pts[3 * k + 2];
    }
    return true;
}

// ---- float64 walk (field or potential) --------------------------------
template <int MODE, int KIND, bool DIP, bool COUNT>
__global__ __launch_bounds__(BT) void k_walk64(const InstParams* __restrict__ P, Globals* g,
                                                PState st, const NodeRec64* __restrict__ rec,
                                                WalkIO io)
{
    __shared__ uint32_t s_first[NW][STK];
    __shared__ uint32_t s_cnt[NW][STK];
    __shared__ uint64_t s_mask[NW][STK];
    const int wv = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0, tot_pops = 0;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * WAVE >= ntg || rfl(step_failed(g) ? 1u : 0u)) break;
        const uint32_t k = tile * WAVE + lane;
        uint32_t inst, self;
        double tx, ty, tz;
        const bool valid = load_target<MODE>(P, st, io, k, ntg, inst, self, tx, ty, tz);
        const double theta2 = P[inst].theta2, soft2 = P[inst].soft2;
        double ex = 0.0, ey = 0.0, ez = 0.0, ph = 0.0;
        if (valid) {
            if (io.acc && KIND == KIND_FIELD) {
                ex = io.out[3 * k];
                ey = io.out[3 * k + 1];
                ez = io.out[3 * k + 2];
            } else if (io.acc) {
                ph = io.out[k];
            }
        }
        uint32_t ncell = 0, nleaf = 0, nopen = 0;

        uint64_t pending = ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint64_t grp = rfl64(ballot(valid && inst == gi));
            pending &= ~grp;
            const bool dip = DIP && P[gi].dipoles && io.dip64;
            uint32_t tf = gi, tc = 1;
            uint64_t tm = grp;
            int sp = 0;
            uint32_t loads = 0;
            for (;;) {
                if (tc == 0) {
                    if (sp == 0) break;
                    --sp;
                    tf = rfl(s_first[wv][sp]);
                    tc = rfl(s_cnt[wv][sp]);
                    tm = rfl64(s_mask[wv][sp]);
                }
                if (++loads > io.n_rec || tc > 8) {
                    raise_err(g, tc > 8 ? KERR_CAPACITY : KERR_DEPTH, tf);
                    break;
                }
                tot_pops++;
                // one load for the whole range: lane j holds child j
                NodeRec64 mine;
                double mpx = 0.0, mpy = 0.0, mpz = 0.0;
                if (lane < tc) mine = rec[tf + lane];
                else mine = NodeRec64{0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0};
                if (dip && lane < tc) {
                    mpx = io.dip64[3 * (size_t)(tf + lane)];
                    mpy = io.dip64[3 * (size_t)(tf + lane) + 1];
                    mpz = io.dip64[3 * (size_t)(tf + lane) + 2];
                }
                const bool act = (tm >> lane) & 1ull;
                uint32_t next_first = 0, next_cnt = 0;
                uint64_t next_mask = 0;
                int j = (int)tc - 1;
                for (; j >= 0; j--) {
                    const double cx = bcast_d(mine.cx, j), cy = bcast_d(mine.cy, j),
                                 cz = bcast_d(mine.cz, j);
                    const uint32_t cnt = bcast_u(mine.cnt, j);
                    const int q = (int)bcast_u((uint32_t)mine.q, j);
                    bool open = false;
                    if (act) {
                        const double dx = tx - cx, dy = ty - cy, dz = tz - cz;
                        const double d2 = (dx * dx) + (dy * dy) + (dz * dz);
                        if (cnt == LEAF) {
                            if (KIND == KIND_FIELD) {
                                if (bcast_u(mine.part, j) != self) {
                                    if (d2 != 0.0) {
                                        const double dd2 = d2 + soft2;
                                        const double dist = sqrt(dd2);
                                        const double es = (QE * q) / ((K4PE0 * dd2) * dist);
                                        ex += es * dx;
                                        ey += es * dy;
                                        ez += es * dz;
                                        nleaf++;
                                    } else {
                                        raise_err(g, KERR_COINCIDENT, self);
                                    }
                                }
                            } else if (KIND == KIND_PHI || bcast_u(mine.part, j) != self) {
                                ph += (QE * q) / (K4PE0 * sqrt(d2 + soft2));
                                nleaf++;
                            }
                        } else if (bcast_d(mine.size2, j) < theta2 * d2) {
                            // DIPOLES accepts every cell holding charge, net
                            // or not (pot.c:1732-1736); internal records
                            // carry |q| in `part`
                            if (dip ? bcast_u(mine.part, j) != 0 : q != 0) {
                                if (dip) {
                                    const double px = bcast_d(mpx, j), py = bcast_d(mpy, j),
                                                 pz = bcast_d(mpz, j);
                                    const double dist = sqrt(d2);
                                    if (KIND == KIND_FIELD) {
                                        // pot.c:1738-1745
                                        const double es = QE / ((K4PE0 * d2) * dist);
                                        const double rpd2 = ((dx * px) + (dy * py) + (dz * pz)) / d2;
                                        ex += es * ((q * dx) - px + ((3 * dx) * rpd2));
                                        ey += es * ((q * dy) - py + ((3 * dy) * rpd2));
                                        ez += es * ((q * dz) - pz + ((3 * dz) * rpd2));
                                    } else {
                                        const double rdp = (dx * px) + (dy * py) + (dz * pz);
                                        const double t = (q / dist) + (rdp / (dist * d2));
                                        // phi_at (pot.c:1862) and phi_particle_feels
                                        // (pot.c:1969) group the constants differently
                                        if (KIND == KIND_PHI) ph += (QE / K4PE0) * t;
                                        else ph += (QE * t) / K4PE0;
                                    }
                                } else if (KIND == KIND_FIELD) {
                                    const double dist = sqrt(d2);
                                    double es = QE / ((K4PE0 * d2) * dist);
                                    es *= q;
                                    ex += es * dx;
                                    ey += es * dy;
                                    ez += es * dz;
                                } else {
                                    ph += (QE * q) / (K4PE0 * sqrt(d2));
                                }
                                ncell++;
                            }
                        } else {
                            open = cnt != 0;
                        }
                    }
                    const uint64_t om = rfl64(ballot(open));
                    if (om) {
                        if (open) nopen++;
                        next_first = bcast_u(mine.first, j);
                        next_cnt = cnt;
                        next_mask = om;
                        break;
                    }
                }
                if (j < 0) {
                    tc = 0;
                    continue;
                }
                // descend into child j; its elder siblings wait on the stack
                if (next_first <= tf + (uint32_t)j || next_first + next_cnt > io.n_rec) {
                    raise_err(g, KERR_CAPACITY, tf + j);
                    break;
                }
                if (j > 0) {
                    if (sp >= STK) {
                        raise_err(g, KERR_DEPTH, tf);
                        break;
                    }
                    s_first[wv][sp] = tf;
                    s_cnt[wv][sp] = (uint32_t)j;
                    s_mask[wv][sp] = tm;
                    sp++;
                }
                tf = next_first;
                tc = next_cnt;
                tm = next_mask;
            }
        }
        if (valid) {
            if (KIND == KIND_FIELD) {
                io.out[3 * k] = ex;
                io.out[3 * k + 1] = ey;
                io.out[3 * k + 2] = ez;
            } else {
                io.out[k] = ph;
            }
            if (COUNT && io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        tot_cell += ncell;
        tot_leaf += nleaf;
        tot_open += nopen;
    }
    if (COUNT && KIND == KIND_FIELD) {
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
            tot_open += __shfl_xor(tot_open, d, WAVE);
        }
        if (lane == 0) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
            atomicAdd(&g->n_pops, tot_pops);
        }
    }
}

// ---- float32 compensated walk (field) ----------------------------------
//
// Same traversal and per-particle MAC as k_walk64, in float32 with
// coordinates relative to the domain centre in units of R.  A fetched child
// range (<= 8 children) is staged in LDS and visited in two passes: its
// leaves first, which every target in the mask takes (but for its own
// leaf) and which are never opened, then its cells, youngest first, each
// lane applying the MAC for its own target, until some lanes open a cell:
// the wavefront descends into it with those lanes, and the cells not yet
// visited wait on the stack (the leaves are done).  Each target gets
// exactly the reference's interaction set; only the summation order
// differs from the depth-first order.  The interactions of one range visit
// are summed in plain float32 and that sum enters the per-particle Kahan
// sum.
//
// Lane sets are 64-bit scalar masks and every per-child decision that is
// the same for all lanes (leaf or cell, which child next) is scalar, so a
// child step costs few scalar instructions besides its vector arithmetic;
// lanes not taking a child get zero coefficients and every lane adds
// a * d + b * p to its sums in place.  Interaction counts, when kept, are
// per-lane counters summed once per kernel: a scalar popcount per child
// cost the walk twice as much (the scalar unit is shared by the CU).
// one (target, node) interaction; returns true when the target opens the node.
// With dip (DIPOLES), cells holding any charge are accepted and pd is the
// cell's dipole in units of R (pot.c:1732-1745).
__device__ inline bool interact32(float tx, float ty, float tz, uint32_t self,
                                  float soft2, const NodeRec32& r, bool dip, float4 pd, float& ax,
                                  float& ay, float& az, uint32_t& ncell, uint32_t& nleaf)
{
    const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    if (r.cnt == LEAF) {
        if (r.part != self && d2 != 0.f) {
            const float inv = __builtin_amdgcn_rsqf(d2 + soft2);
            const float sc = r.q * (inv * inv * inv);
            ax = __builtin_fmaf(sc, dx, ax);
            ay = __builtin_fmaf(sc, dy, ay);
            az = __builtin_fmaf(sc, dz, az);
            nleaf++;
        }
        return false;
    }
    if (r.size2 < d2) {
        if (dip ? r.part != 0u : r.q != 0.f) {
            const float inv = __builtin_amdgcn_rsqf(d2);
            const float inv2 = inv * inv;
            const float inv3 = inv2 * inv;
            if (dip) {
                const float rpd2 =
                    __builtin_fmaf(dx, pd.x, __builtin_fmaf(dy, pd.y, dz * pd.z)) * inv2;
                const float t3 = 3.f * rpd2;
                ax = __builtin_fmaf(inv3, __builtin_fmaf(r.q + t3, dx, -pd.x), ax);
                ay = __builtin_fmaf(inv3, __builtin_fmaf(r.q + t3, dy, -pd.y), ay);
                az = __builtin_fmaf(inv3, __builtin_fmaf(r.q + t3, dz, -pd.z), az);
            } else {
                const float sc = r.q * inv3;
                ax = __builtin_fmaf(sc, dx, ax);
                ay = __builtin_fmaf(sc, dy, ay);
                az = __builtin_fmaf(sc, dz, az);
            }
            ncell++;
        }
        return false;
    }
    return r.cnt != 0;
}

// a scalar float's bits other than the sign, tested on the scalar unit (the
// barrier keeps the compiler from turning it into a float compare, which
// only the vector unit has)
__device__ inline bool nonzero_bits(uint32_t b)
{
    asm volatile("" : "+s"(b));
    return (b << 1) != 0u;
}

// one visit of a staged range: its leaves in `rem & lm`, then its cells in
// `rem & ~lm`, youngest first, until some lanes open a cell.  Returns that
// cell's index (with its range and the opening lanes in next_*) and leaves
// the cells still to visit in rem; -1 when the range is done.  hqm marks
// the children holding charge (the reference skips the others, pot.c:1732).
// a staged child: its record, then (DIPOLES kernels) its dipole, so that one
// LDS address serves both
template <bool DIPK>
struct Staged {
    NodeRec32 r;
    float4 d[DIPK ? 1 : 0];
};

template <bool D, bool CNT, bool DIPK>
__device__ inline int visit_range(const Staged<DIPK>* sr, uint32_t& rem, uint32_t lm,
                                  uint32_t hqm, uint64_t tm, float tx, float ty, float tz,
                                  uint32_t self, float soft2, float& px, float& py,
                                  float& pz, uint32_t& ncell, uint32_t& nleaf, uint32_t& nopen,
                                  uint32_t& next_first,
                                  uint32_t& next_cnt, uint64_t& next_mask)
{
    const uint32_t lane = __lane_id();
    const bool act = (tm >> lane) & 1ull;
    for (uint32_t m = rem & lm; m; m &= m - 1u) {
        const uint32_t j = (uint32_t)__builtin_ctz(m);
        const NodeRec32 r = sr[j].r;
        const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        // a target skips its own leaf (pot.c:1787); a leaf at float32
        // distance zero adds nothing
        const bool pn = r.part != self, dn = d2 != 0.f;
        const float inv = __builtin_amdgcn_rsqf(d2 + soft2);
        const bool use = act && pn && dn;
        const float ca = use ? r.q * (inv * inv * inv) : 0.f;
        px = __builtin_fmaf(ca, dx, px);
        py = __builtin_fmaf(ca, dy, py);
        pz = __builtin_fmaf(ca, dz, pz);
        if (CNT) nleaf += use ? 1u : 0u;
    }
    for (uint32_t c = rem & ~lm; c;) {
        const uint32_t j = 31u - (uint32_t)__builtin_clz(c);
        c &= ~(1u << j);
        const NodeRec32 r = sr[j].r;
        const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const bool acc = r.size2 < d2;
        const uint64_t am = ballot(acc);
        const bool use = act && acc;
        // lanes not taking the cell get zero coefficients (and nothing
        // infinite, d2 may be zero there)
        const float inv = use ? __builtin_amdgcn_rsqf(d2) : 0.f;
        const float inv2 = inv * inv;
        const float inv3 = inv2 * inv;
        if (D) {
            // DIPOLES (pot.c:1738-1745): q d / r^3 - p / r^3 + 3 (d.p) d / r^5
            const float4 pd = sr[j].d[0];
            const float rpd2 = __builtin_fmaf(dx, pd.x, __builtin_fmaf(dy, pd.y, dz * pd.z)) * inv2;
            const float ca = inv3 * __builtin_fmaf(3.f, rpd2, r.q);
            px = __builtin_fmaf(-inv3, pd.x, __builtin_fmaf(ca, dx, px));
            py = __builtin_fmaf(-inv3, pd.y, __builtin_fmaf(ca, dy, py));
            pz = __builtin_fmaf(-inv3, pd.z, __builtin_fmaf(ca, dz, pz));
        } else {
            // a cell without net charge adds zero
            const float ca = r.q * inv3;
            px = __builtin_fmaf(ca, dx, px);
            py = __builtin_fmaf(ca, dy, py);
            pz = __builtin_fmaf(ca, dz, pz);
        }
        if (CNT && ((hqm >> j) & 1u)) ncell += use ? 1u : 0u;
        const uint64_t om = tm & ~am;
        if (om) {
            if (CNT && ((om >> lane) & 1ull)) nopen++;
            next_first = rfl(r.first);
            next_cnt = rfl(r.cnt);
            next_mask = om;
            rem = c;
            return (int)j;
        }
    }
    rem = 0;
    return -1;
}

template <int MODE, bool DIP, bool COUNT>
__global__ __launch_bounds__(BT) void k_walk32(const InstParams* __restrict__ P, Globals* g,
                                                PState st, const NodeRec32* __restrict__ rec,
                                                WalkIO io)
{
    __shared__ uint32_t s_first[NW][STK];
    __shared__ uint32_t s_cnt[NW][STK];  // cnt | cells still to visit << 8
    __shared__ uint64_t s_mask[NW][STK];
    __shared__ Staged<DIP> s_rec[NW][8];
    const int wv = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    // per-lane interaction counts: for the point queries' own counts, and
    // for the context's totals when counting
    constexpr bool CNT = MODE == MODE_POINTS || COUNT;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0, tot_pops = 0;
    unsigned long long tot_vis = 0;  // children visited by the wavefront (counting only)

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * WAVE >= ntg || rfl(step_failed(g) ? 1u : 0u)) break;
        const uint32_t k = tile * WAVE + lane;
        uint32_t inst, self;
        double dtx, dty, dtz;
        const bool valid = load_target<MODE>(P, st, io, k, ntg, inst, self, dtx, dty, dtz);
        const InstParams& p = P[inst];
        const float tx = (float)((dtx - p.R) * p.invR);
        const float ty = (float)((dty - p.R) * p.invR);
        const float tz = (float)((dtz - p.R) * p.invR);
        const float soft2 = (float)(p.soft2 * p.invR * p.invR);
        float fx = 0.f, fy = 0.f, fz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        uint32_t ncell = 0, nleaf = 0, nopen = 0;
        uint32_t wopen = 0;  // lanes descending, as a scalar total

        uint64_t pending = ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint64_t grp = rfl64(ballot(valid && inst == gi));
            pending &= ~grp;
            const bool dip = DIP && P[gi].dipoles && io.dip32;
            // an empty instance's root has no children and no charge
            if (rfl(rec[gi].cnt) == 0u) continue;
            uint32_t tf = gi, tc = 1, rem = 1;
            uint64_t tm = grp;
            int sp = 0;
            uint32_t loads = 0;
            for (;;) {
                if (rem == 0) {
                    if (sp == 0) break;
                    --sp;
                    tf = rfl(s_first[wv][sp]);
                    const uint32_t cr = rfl(s_cnt[wv][sp]);
                    tc = cr & 0xffu;
                    rem = cr >> 8;
                    tm = rfl64(s_mask[wv][sp]);
                }
                if (++loads > io.n_rec || tc > 8) {
                    raise_err(g, tc > 8 ? KERR_CAPACITY : KERR_DEPTH, tf);
                    break;
                }
                tot_pops++;
                uint32_t lm = 0, hqm = 0;
                {
                    // lane j holds child j: one load for the whole range
                    NodeRec32 mr;
                    if (lane < tc) mr = rec[tf + lane];
                    else mr = NodeRec32{0.f, 0.f, 0.f, 0.f, 0.f, 0u, 0u, 0u};
                    if (lane < tc) {
                        s_rec[wv][lane].r = mr;
                        if constexpr (DIP) {
                            if (dip) s_rec[wv][lane].d[0] = io.dip32[tf + lane];
                        }
                    }
                    lm = (uint32_t)ballot(lane < tc && mr.cnt == LEAF);
                    // DIPOLES takes every cell holding charge, net or not
                    // (pot.c:1732-1736); internal records carry |q| in `part`
                    hqm = (uint32_t)ballot(lane < tc && ((DIP && dip) ? mr.part != 0u : mr.q != 0.f));
                }
                __builtin_amdgcn_wave_barrier();
                float px = 0.f, py = 0.f, pz = 0.f;  // this range visit's float32 partial
                uint32_t next_first = 0, next_cnt = 0;
                uint64_t next_mask = 0;
                const uint32_t rem_in = rem;
                const int j =
                    (DIP && dip)
                        ? visit_range<true, CNT, DIP>(s_rec[wv], rem, lm, hqm, tm, tx, ty,
                                                 tz, self, soft2, px, py, pz, ncell, nleaf,
                                                 nopen, next_first, next_cnt, next_mask)
                        : visit_range<false, CNT, DIP>(s_rec[wv], rem, lm, hqm, tm, tx, ty,
                                                  tz, self, soft2, px, py, pz, ncell,
                                                  nleaf, nopen, next_first, next_cnt, next_mask);
                kahan_add(fx, kx, px);
                kahan_add(fy, ky, py);
                kahan_add(fz, kz, pz);
                if (COUNT) tot_vis += (uint32_t)(__builtin_popcount(rem_in) - __builtin_popcount(rem));
                if (j < 0) continue;
                wopen += (uint32_t)__popcll(next_mask);
                if (next_first <= tf + (uint32_t)j || next_cnt - 1u > 7u ||
                    next_first + next_cnt > io.n_rec) {
                    raise_err(g, KERR_CAPACITY, tf + j);
                    break;
                }
                if (rem) {
                    if (sp >= STK) {
                        raise_err(g, KERR_DEPTH, tf);
                        break;
                    }
                    s_first[wv][sp] = tf;
                    s_cnt[wv][sp] = tc | (rem << 8);
                    s_mask[wv][sp] = tm;
                    sp++;
                }
                tf = next_first;
                tc = next_cnt;
                rem = (1u << tc) - 1u;
                tm = next_mask;
            }
        }
        if (valid) {
            const double K = QE / (K4PE0 * p.R * p.R);
            io.out[3 * k] = (io.acc ? io.out[3 * k] : 0.0) + K * ((double)fx - (double)kx);
            io.out[3 * k + 1] = (io.acc ? io.out[3 * k + 1] : 0.0) + K * ((double)fy - (double)ky);
            io.out[3 * k + 2] = (io.acc ? io.out[3 * k + 2] : 0.0) + K * ((double)fz - (double)kz);
            if (COUNT && io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        tot_cell += ncell;
        tot_leaf += nleaf;
        if (lane == 0) tot_open += wopen;
    }
    if (COUNT) {
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
        }
        if (lane == 0) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
            atomicAdd(&g->n_pops, tot_pops);
            atomicAdd(&g->n_visits, tot_vis);
        }
    }
}

// ---- float32 walk, whole-range sweeps -------------------------------------
//
// The traversal of k_walk32 with another visit order.  A freshly fetched
// range is swept whole: its leaves, then every one of its cells, youngest
// first, each lane applying its own MAC and the accepting lanes adding the
// cell; the cells that some lanes open are only noted.  Then the wavefront
// descends into the youngest opened cell with the lanes that opened it; the
// others wait on the stack as (range, opened cells, lanes).  When such an
// entry is popped, the next opened cell's record is read alone (a uniform
// load) and its opening lanes found again by the same MAC -- nothing is
// added, the range's interactions were all taken by the sweep.  The sweep
// has no exit in the middle, so each child costs its arithmetic and a few
// scalar mask operations; every target still gets exactly the reference's
// interaction set (the order of the sums differs from k_walk32's).
template <bool D, bool CNT, bool DIPK>
__device__ inline uint32_t sweep_range(const Staged<DIPK>* sr, uint32_t tc, uint32_t lm,
                                       uint32_t hqm, uint64_t tm, float tx, float ty, float tz,
                                       uint32_t self, float soft2, float& px, float& py, float& pz,
                                       uint32_t& ncell, uint32_t& nleaf, uint32_t& nopen,
                                       uint32_t& wopen, uint64_t& first_om)
{
    const uint32_t lane = __lane_id();
    // the lane's bit of the mask, used as a lane mask directly (no vector
    // shift and compare)
    const bool act = __builtin_amdgcn_inverse_ballot_w64(tm);
    const uint32_t all = (1u << tc) - 1u;
    const uint32_t lv = all & lm, cl = all & ~lm;
    // unrolled over the eight child slots (a scalar test skips a slot), so
    // that each slot's LDS reads take a constant offset from one base
#pragma unroll
    for (uint32_t j = 0; j < 8u; j++) {
        if (!((lv >> j) & 1u)) continue;
        const NodeRec32 r = sr[j].r;
        const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const bool pn = r.part != self, dn = d2 != 0.f;
        const float inv = __builtin_amdgcn_rsqf(d2 + soft2);
        const bool use = act && pn && dn;
        const float ca = use ? r.q * (inv * inv * inv) : 0.f;
        px = __builtin_fmaf(ca, dx, px);
        py = __builtin_fmaf(ca, dy, py);
        pz = __builtin_fmaf(ca, dz, pz);
        if (CNT) nleaf += use ? 1u : 0u;
    }
    uint32_t opened = 0;
    first_om = 0;
#pragma unroll
    for (int jj = 7; jj >= 0; jj--) {
        const uint32_t j = (uint32_t)jj;
        if (!((cl >> j) & 1u)) continue;
        const NodeRec32 r = sr[j].r;
        const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const bool acc = r.size2 < d2;
        const uint64_t am = ballot(acc);
        const bool use = act && acc;
        const float inv = use ? __builtin_amdgcn_rsqf(d2) : 0.f;
        const float inv2 = inv * inv;
        const float inv3 = inv2 * inv;
        if (D) {
            const float4 pd = sr[j].d[0];
            const float rpd2 = __builtin_fmaf(dx, pd.x, __builtin_fmaf(dy, pd.y, dz * pd.z)) * inv2;
            const float ca = __builtin_fmaf(3.f, rpd2, r.q);
            px = __builtin_fmaf(inv3, __builtin_fmaf(ca, dx, -pd.x), px);
            py = __builtin_fmaf(inv3, __builtin_fmaf(ca, dy, -pd.y), py);
            pz = __builtin_fmaf(inv3, __builtin_fmaf(ca, dz, -pd.z), pz);
        } else {
            const float ca = r.q * inv3;
            px = __builtin_fmaf(ca, dx, px);
            py = __builtin_fmaf(ca, dy, py);
            pz = __builtin_fmaf(ca, dz, pz);
        }
        if (CNT && ((hqm >> j) & 1u)) ncell += use ? 1u : 0u;
        const uint64_t om = tm & ~am;
        if (om) {
            if (CNT && ((om >> lane) & 1ull)) nopen++;
            wopen += (uint32_t)__popcll(om);
            if (!opened) first_om = om;  // the youngest opened cell, descended first
            opened |= 1u << j;
        }
    }
    return opened;
}

template <int MODE, bool DIP, bool COUNT>
__global__ __launch_bounds__(BT) void k_walk32s(const InstParams* __restrict__ P, Globals* g,
                                                 PState st, const NodeRec32* __restrict__ rec,
                                                 WalkIO io)
{
    __shared__ uint32_t s_first[NW][STK];
    __shared__ uint32_t s_cnt[NW][STK];  // cnt | opened cells still to descend << 8
    __shared__ uint64_t s_mask[NW][STK];
    __shared__ Staged<DIP> s_rec[NW][8];
    const int wv = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    constexpr bool CNT = MODE == MODE_POINTS || COUNT;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0, tot_pops = 0;
    unsigned long long tot_vis = 0;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * WAVE >= ntg || rfl(step_failed(g) ? 1u : 0u)) break;
        const uint32_t k = tile * WAVE + lane;
        uint32_t inst, self;
        double dtx, dty, dtz;
        const bool valid = load_target<MODE>(P, st, io, k, ntg, inst, self, dtx, dty, dtz);
        const InstParams& p = P[inst];
        const float tx = (float)((dtx - p.R) * p.invR);
        const float ty = (float)((dty - p.R) * p.invR);
        const float tz = (float)((dtz - p.R) * p.invR);
        const float soft2 = (float)(p.soft2 * p.invR * p.invR);
        float fx = 0.f, fy = 0.f, fz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        uint32_t ncell = 0, nleaf = 0, nopen = 0;
        uint32_t wopen = 0;

        uint64_t pending = ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint64_t grp = rfl64(ballot(valid && inst == gi));
            pending &= ~grp;
            const bool dip = DIP && P[gi].dipoles && io.dip32;
            if (rfl(rec[gi].cnt) == 0u) continue;
            // the range to sweep next: the root's
            uint32_t tf = gi, tc = 1;
            uint64_t tm = grp;
            int sp = 0;
            uint32_t loads = 0;
            for (;;) {
                if (++loads > io.n_rec || tc - 1u > 7u) {
                    raise_err(g, tc > 8 ? KERR_CAPACITY : KERR_DEPTH, tf);
                    break;
                }
                tot_pops++;
                uint32_t lm = 0, hqm = 0;
                {
                    // lane j stages child j; the lanes past the range read
                    // child 0 again (a clamped index: the load and the ballots
                    // need no branch) and their bits are masked off
                    const bool in = lane < tc;
                    const NodeRec32 mr = rec[tf + (in ? lane : 0u)];
                    if (in) {
                        s_rec[wv][lane].r = mr;
                        if constexpr (DIP) {
                            if (dip) s_rec[wv][lane].d[0] = io.dip32[tf + lane];
                        }
                    }
                    const uint32_t inm = (1u << tc) - 1u;
                    lm = (uint32_t)ballot(mr.cnt == LEAF) & inm;
                    hqm = (uint32_t)ballot((DIP && dip) ? mr.part != 0u : mr.q != 0.f) & inm;
                }
                __builtin_amdgcn_wave_barrier();
                float px = 0.f, py = 0.f, pz = 0.f;
                uint64_t om = 0;
                uint32_t opened =
                    (DIP && dip)
                        ? sweep_range<true, CNT, DIP>(s_rec[wv], tc, lm, hqm, tm, tx, ty, tz, self,
                                                      soft2, px, py, pz, ncell, nleaf, nopen,
                                                      wopen, om)
                        : sweep_range<false, CNT, DIP>(s_rec[wv], tc, lm, hqm, tm, tx, ty, tz,
                                                       self, soft2, px, py, pz, ncell, nleaf,
                                                       nopen, wopen, om);
                kahan_add(fx, kx, px);
                kahan_add(fy, ky, py);
                kahan_add(fz, kz, pz);
                if (COUNT) tot_vis += tc;
                uint32_t j = 0, next_first = 0, next_cnt = 0;
                if (opened) {
                    j = 31u - (uint32_t)__builtin_clz(opened);
                    opened &= ~(1u << j);
                    next_first = rfl(s_rec[wv][j].r.first);
                    next_cnt = rfl(s_rec[wv][j].r.cnt);
                } else {
                    // the next opened cell of a range waiting on the stack:
                    // its record alone, and the lanes that open it, found
                    // again by the sweep's own arithmetic
                    while (sp > 0 && om == 0) {
                        const uint32_t cr = rfl(s_cnt[wv][sp - 1]);
                        uint32_t rest = cr >> 8;
                        tf = rfl(s_first[wv][sp - 1]);
                        tm = rfl64(s_mask[wv][sp - 1]);
                        j = 31u - (uint32_t)__builtin_clz(rest);
                        rest &= ~(1u << j);
                        if (rest) s_cnt[wv][sp - 1] = (cr & 0xffu) | (rest << 8);
                        else --sp;
                        const NodeRec32 r = rec[tf + j];
                        const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
                        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        om = tm & ~ballot(r.size2 < d2);
                        next_first = rfl(r.first);
                        next_cnt = rfl(r.cnt);
                    }
                    if (om == 0) break;  // stack empty: the instance group is done
                }
                if (next_first <= tf + j || next_cnt - 1u > 7u || next_first + next_cnt > io.n_rec) {
                    raise_err(g, KERR_CAPACITY, tf + j);
                    break;
                }
                if (opened) {
                    if (sp >= STK) {
                        raise_err(g, KERR_DEPTH, tf);
                        break;
                    }
                    s_first[wv][sp] = tf;
                    s_cnt[wv][sp] = tc | (opened << 8);
                    s_mask[wv][sp] = tm;
                    sp++;
                }
                tf = next_first;
                tc = next_cnt;
                tm = om;
            }
        }
        if (valid) {
            const double K = QE / (K4PE0 * p.R * p.R);
            io.out[3 * k] = (io.acc ? io.out[3 * k] : 0.0) + K * ((double)fx - (double)kx);
            io.out[3 * k + 1] = (io.acc ? io.out[3 * k + 1] : 0.0) + K * ((double)fy - (double)ky);
            io.out[3 * k + 2] = (io.acc ? io.out[3 * k + 2] : 0.0) + K * ((double)fz - (double)kz);
            if (COUNT && io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        tot_cell += ncell;
        tot_leaf += nleaf;
        if (lane == 0) tot_open += wopen;
    }
    if (COUNT) {
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
        }
        if (lane == 0) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
            atomicAdd(&g->n_pops, tot_pops);
            atomicAdd(&g->n_visits, tot_vis);
        }
    }
}

// ---- grouped float32 walk ----------------------------------------------
//
// A wavefront serves a group of K targets (consecutive in key order, so
// close in space) and evaluates S = 64 / K children of a range at once:
// lane = slot * K + t pairs target t with child slot `slot`.  A range of
// up to 8 children takes ceil(cnt / S) passes; each lane fetches its own
// child record (S distinct records per pass, each shared by K lanes).
// The per-target MAC is the reference's, so every target gets exactly its
// own interaction set; a child opened by some of the group's targets is
// pushed once with the mask of exactly those targets.  Fewer targets per
// wavefront means the group's interaction lists overlap more, so fewer
// lanes idle than with 64 targets stepping through one child at a time;
// the group width K is chosen by potg_tune.
//
// Each lane sums its own (target, slot) contributions in float32, flushed
// every 4 ranges into a compensated sum; at the end the slots of a target
// are added in float64 across lanes.
template <int K, int MODE, bool DIP, bool COUNT>
__global__ __launch_bounds__(BT) void k_walkg(const InstParams* __restrict__ P, Globals* g,
                                               PState st, const NodeRec32* __restrict__ rec,
                                               WalkIO io)
{
    static_assert(K == 8 || K == 16, "group width");
    constexpr uint32_t S = WAVE / K;
    constexpr uint32_t KMASK = (1u << K) - 1u;
    // a visit pops one range and pushes at most 8, so the stack holds at
    // most 7 per level of the path below the root plus one
    constexpr int STKG = 7 * MAXD + 8;
    __shared__ uint32_t s_first[NW][STKG];
    __shared__ uint32_t s_cm[NW][STKG];  // cnt << 16 | target mask
    const int wv = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t t = lane % K, slot = lane / K;
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0, tot_pops = 0;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * K >= ntg || rfl(step_failed(g) ? 1u : 0u)) break;
        const uint32_t k = tile * K + t;
        uint32_t inst, self;
        double dtx, dty, dtz;
        const bool valid = load_target<MODE>(P, st, io, k, ntg, inst, self, dtx, dty, dtz);
        const InstParams& p = P[inst];
        const float tx = (float)((dtx - p.R) * p.invR);
        const float ty = (float)((dty - p.R) * p.invR);
        const float tz = (float)((dtz - p.R) * p.invR);
        const float soft2 = (float)(p.soft2 * p.invR * p.invR);
        float fx = 0.f, fy = 0.f, fz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        float px = 0.f, py = 0.f, pz = 0.f;
        uint32_t ncell = 0, nleaf = 0, nopen = 0;

        // the group's targets by instance (slot-0 lanes stand for them)
        uint32_t pending = (uint32_t)ballot(valid && slot == 0) & KMASK;
        while (pending) {
            const uint32_t lead = __ffs(pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint32_t grp = (uint32_t)ballot(valid && slot == 0 && inst == gi) & KMASK;
            pending &= ~grp;
            const bool dip = DIP && P[gi].dipoles && io.dip32;
            uint32_t tf = gi, tc = 1, tm = grp;
            int sp = 0;
            uint32_t loads = 0, visits = 0;
            bool fail = false;
            for (;;) {
                if (tc == 0) {
                    if (sp == 0) break;
                    --sp;
                    tf = rfl(s_first[wv][sp]);
                    const uint32_t cm = rfl(s_cm[wv][sp]);
                    tc = cm >> 16;
                    tm = cm & KMASK;
                }
                if (++loads > io.n_rec || tc > 8) {
                    raise_err(g, tc > 8 ? KERR_CAPACITY : KERR_DEPTH, tf);
                    break;
                }
                tot_pops++;
                const bool mine = (tm >> t) & 1u;
                for (uint32_t c0 = 0; c0 < tc; c0 += S) {
                    const uint32_t c = c0 + slot;
                    bool op = false;
                    uint32_t cf = 0, cn = 0;
                    if (mine && c < tc) {
                        const NodeRec32 r = rec[tf + c];
                        float4 pd = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (DIP && dip) pd = io.dip32[tf + c];
                        op = interact32(tx, ty, tz, self, soft2, r, DIP && dip, pd, px, py,
                                        pz, ncell, nleaf);
                        cf = r.first;
                        cn = r.cnt;
                    }
                    if (op) nopen++;
                    const uint64_t b = ballot(op);
                    // push this pass's opened children, ascending: the
                    // youngest ends on top and is visited first
                    for (uint32_t s2 = 0; s2 < S && c0 + s2 < tc; s2++) {
                        const uint32_t om = (uint32_t)(b >> (s2 * K)) & KMASK;
                        if (!om) continue;
                        const uint32_t src = s2 * K + (__ffs(om) - 1);
                        const uint32_t ccf = (uint32_t)__builtin_amdgcn_readlane((int)cf, src);
                        const uint32_t ccn = (uint32_t)__builtin_amdgcn_readlane((int)cn, src);
                        if (ccf <= tf + c0 + s2 || ccn > 8 || ccf + ccn > io.n_rec || sp >= STKG) {
                            raise_err(g, sp >= STKG ? KERR_DEPTH : KERR_CAPACITY, tf + c0 + s2);
                            fail = true;
                            break;
                        }
                        if (lane == 0) {
                            s_first[wv][sp] = ccf;
                            s_cm[wv][sp] = (ccn << 16) | om;
                        }
                        sp++;
                    }
                    if (fail) break;
                }
                if (fail) break;
                tc = 0;
                if ((++visits & 3u) == 0) {
                    kahan_add(fx, kx, px);
                    kahan_add(fy, ky, py);
                    kahan_add(fz, kz, pz);
                    px = py = pz = 0.f;
                }
            }
        }
        kahan_add(fx, kx, px);
        kahan_add(fy, ky, py);
        kahan_add(fz, kz, pz);
        // add the slots of each target (lanes t, t + K, ...) in float64
        double ex = (double)fx - (double)kx, ey = (double)fy - (double)ky,
               ez = (double)fz - (double)kz;
        for (uint32_t d = K; d < WAVE; d <<= 1) {
            ex += __shfl_xor(ex, (int)d, WAVE);
            ey += __shfl_xor(ey, (int)d, WAVE);
            ez += __shfl_xor(ez, (int)d, WAVE);
            ncell += __shfl_xor(ncell, (int)d, WAVE);
            nleaf += __shfl_xor(nleaf, (int)d, WAVE);
            nopen += __shfl_xor(nopen, (int)d, WAVE);
        }
        if (valid && slot == 0) {
            const double Kf = QE / (K4PE0 * p.R * p.R);
            io.out[3 * k] = (io.acc ? io.out[3 * k] : 0.0) + Kf * ex;
            io.out[3 * k + 1] = (io.acc ? io.out[3 * k + 1] : 0.0) + Kf * ey;
            io.out[3 * k + 2] = (io.acc ? io.out[3 * k + 2] : 0.0) + Kf * ez;
            if (COUNT && io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        if (slot == 0) {
            tot_cell += ncell;
            tot_leaf += nleaf;
            tot_open += nopen;
        }
    }
    if (COUNT) {
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
            tot_open += __shfl_xor(tot_open, d, WAVE);
        }
        if (lane == 0) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
            atomicAdd(&g->n_pops, tot_pops);
        }
    }
}

// ---- float32 walk with the group opening criterion --------------------
//
// (POTG_MIXED_GROUP)  The 64 targets of a wavefront share one interaction
// list.  A cell is accepted for all of them when size^2 < theta^2 dmin^2,
// dmin being the distance from the cell's centre of |q| to the bounding box
// of the targets; since dmin is at most each target's own distance, every
// cell accepted here the reference would accept for every target too, so
// each target's interactions refine the reference's (cells it would have
// taken whole are taken in parts): the force is at least as accurate, and
// the lanes are never idle.
//
// A wavefront's 64 targets are consecutive in key order, which can cross
// from one large cell into a distant one; so the targets are split into
// groups by the octree cell holding them at depth D, chosen so that such a
// cell holds a few dozen targets, and each group gets its own list.
//
// Phase A builds the list lane-parallel: up to 8 child ranges are taken
// off a stack at once, one child per lane, each lane applying the group
// criterion to its child; accepted cells and leaves are appended to an LDS
// list, opened cells' child ranges pushed.  Phase B, whenever the list
// fills or the walk ends, runs the list: every lane adds the interaction of
// each entry with its own target, the entry's record coming in through
// scalar loads.  The target's own leaf is skipped (pot.c:1787).
template <int MODE, bool DIP, bool COUNT>
__global__ __launch_bounds__(BT) void k_walk32grp(const InstParams* __restrict__ P, Globals* g,
                                                   PState st, const NodeRec32* __restrict__ rec,
                                                   WalkIO io)
{
    constexpr int NS = 1024;  // child ranges waiting: first | (cnt - 1) << 29
    constexpr int NL = 512;   // interaction list
    // while the stack is below this, up to 8 ranges are expanded at a time;
    // above it one (pure depth first), which adds at most 7 per level
    constexpr int SP_WIDE = NS - 7 * MAXD - 72;
    __shared__ uint32_t s_stk[NW][NS];
    __shared__ uint32_t s_list[NW][NL];
    const int wv = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0, tot_pops = 0;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * WAVE >= ntg || rfl(step_failed(g) ? 1u : 0u)) break;
        const uint32_t k = tile * WAVE + lane;
        uint32_t inst, self;
        double dtx, dty, dtz;
        const bool valid = load_target<MODE>(P, st, io, k, ntg, inst, self, dtx, dty, dtz);
        const InstParams& p = P[inst];
        const float tx = (float)((dtx - p.R) * p.invR);
        const float ty = (float)((dty - p.R) * p.invR);
        const float tz = (float)((dtz - p.R) * p.invR);
        const float soft2 = (float)(p.soft2 * p.invR * p.invR);
        float fx = 0.f, fy = 0.f, fz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        uint32_t ncell = 0, nleaf = 0, nopen = 0;
        // the target's octree cell at depth D: D = log8(tree particles / 32)
        uint32_t cell = 0;
        {
            const uint32_t nt = p.tree_n > 64u ? p.tree_n : 64u;
            int D = (31 - __clz(nt / 32u)) / 3;
            D = D < 1 ? 1 : (D > 9 ? 9 : D);
            const float sc = (float)(1u << D) * 0.5f;
            const uint32_t mx = (1u << D) - 1u;
            const uint32_t cx = (uint32_t)fminf(fmaxf((tx + 1.f) * sc, 0.f), (float)mx);
            const uint32_t cy = (uint32_t)fminf(fmaxf((ty + 1.f) * sc, 0.f), (float)mx);
            const uint32_t cz = (uint32_t)fminf(fmaxf((tz + 1.f) * sc, 0.f), (float)mx);
            cell = (cx << 20) | (cy << 10) | cz;
        }

        uint64_t pending = ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint32_t gc = rfl(__shfl(cell, lead, WAVE));
            const bool mine = valid && inst == gi && cell == gc;
            pending &= ~rfl64(ballot(mine));
            const bool dip = DIP && P[gi].dipoles && io.dip32;
            // bounding box of this instance's targets in the wavefront
            float lx = mine ? tx : 3.f, ly = mine ? ty : 3.f, lz = mine ? tz : 3.f;
            float hx = mine ? tx : -3.f, hy = mine ? ty : -3.f, hz = mine ? tz : -3.f;
            for (int d = 1; d < WAVE; d <<= 1) {
                lx = fminf(lx, __shfl_xor(lx, d, WAVE));
                ly = fminf(ly, __shfl_xor(ly, d, WAVE));
                lz = fminf(lz, __shfl_xor(lz, d, WAVE));
                hx = fmaxf(hx, __shfl_xor(hx, d, WAVE));
                hy = fmaxf(hy, __shfl_xor(hy, d, WAVE));
                hz = fmaxf(hz, __shfl_xor(hz, d, WAVE));
            }
            // the root's children: the root record is a leaf only for a
            // one-particle tree; an empty instance has no records to walk
            int sp = 0, nl = 0;
            {
                const NodeRec32 r0 = rec[gi];
                if (r0.cnt == LEAF) {
                    if (lane == 0) s_list[wv][0] = gi;
                    nl = 1;
                } else if (r0.cnt != 0) {
                    if (r0.cnt - 1u > 7u || r0.first + r0.cnt > io.n_rec) {
                        raise_err(g, KERR_CAPACITY, gi);
                        continue;
                    }
                    if (lane == 0) s_stk[wv][0] = r0.first | ((r0.cnt - 1u) << 29);
                    sp = 1;
                    nopen++;
                }
            }
            float px = 0.f, py = 0.f, pz = 0.f;
            uint32_t loads = 0;
            bool fail = false;
            while (sp > 0 || nl > 0) {
                // ---- phase A: expand ranges until the list is nearly full
                while (sp > 0 && nl <= NL - WAVE) {
                    if (++loads > io.n_rec) {
                        raise_err(g, KERR_DEPTH, gi);
                        fail = true;
                        break;
                    }
                    const int take = sp > SP_WIDE ? 1 : (sp < 8 ? sp : 8);
                    tot_pops += (unsigned long long)take;
                    const uint32_t ri = lane >> 3, ci = lane & 7u;  // range, child
                    uint32_t w = NONE;
                    if ((int)ri < take) {
                        const uint32_t fc = s_stk[wv][sp - 1 - (int)ri];
                        if (ci <= (fc >> 29)) w = (fc & 0x1fffffffu) + ci;
                    }
                    sp -= take;
                    bool to_list = false, to_stack = false;
                    uint32_t cf = 0, cn = 0;
                    if (w != NONE) {
                        const NodeRec32 r = rec[w];
                        if (r.cnt == LEAF) {
                            to_list = true;
                        } else {
                            const float ex = fmaxf(0.f, fmaxf(lx - r.cx, r.cx - hx));
                            const float ey = fmaxf(0.f, fmaxf(ly - r.cy, r.cy - hy));
                            const float ez = fmaxf(0.f, fmaxf(lz - r.cz, r.cz - hz));
                            const float dmin2 = __builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez));
                            if (r.size2 < dmin2)
                                to_list = dip ? r.part != 0u : r.q != 0.f;
                            else if (r.cnt != 0) {
                                to_stack = true;
                                cf = r.first;
                                cn = r.cnt;
                            }
                        }
                    }
                    if (to_stack && (cn - 1u > 7u || cf <= w || cf + cn > io.n_rec)) {
                        raise_err(g, KERR_CAPACITY, w);
                        to_stack = false;
                    }
                    const uint64_t bl = ballot(to_list), bs = ballot(to_stack);
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(bl >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bl, 0u));
                    if (to_list) s_list[wv][nl + (int)below] = w;
                    nl += __popcll(bl);
                    const int ns = __popcll(bs);
                    if (sp + ns > NS) {
                        raise_err(g, KERR_DEPTH, gi);
                        fail = true;
                        break;
                    }
                    const uint32_t sbelow = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(bs >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bs, 0u));
                    if (to_stack) s_stk[wv][sp + (int)sbelow] = cf | ((cn - 1u) << 29);
                    sp += ns;
                    nopen += (uint32_t)ns;
                }
                if (fail) break;
                __builtin_amdgcn_wave_barrier();
                // ---- phase B: every target against every listed entry,
                // the next entry's record requested before this one is used
                uint32_t wn = nl > 0 ? rfl(s_list[wv][0]) : 0u;
                NodeRec32 rn = rec[wn];  // uniform: scalar loads
                for (int e = 0; e < nl; e++) {
                    const uint32_t w = wn;
                    const NodeRec32 r = rn;
                    if (e + 1 < nl) {
                        wn = rfl(s_list[wv][e + 1]);
                        rn = rec[wn];
                    }
                    const float dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
                    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                    if (r.cnt == LEAF) {
                        const bool use = mine && r.part != self && d2 != 0.f;
                        const float inv = __builtin_amdgcn_rsqf(d2 + soft2);
                        const float sc = use ? r.q * (inv * inv * inv) : 0.f;
                        px = __builtin_fmaf(sc, dx, px);
                        py = __builtin_fmaf(sc, dy, py);
                        pz = __builtin_fmaf(sc, dz, pz);
                        nleaf += use ? 1u : 0u;
                    } else {
                        const float inv = __builtin_amdgcn_rsqf(d2);
                        const float inv2 = inv * inv;
                        const float inv3 = mine ? inv2 * inv : 0.f;
                        if (DIP && dip) {
                            const float4 pd = io.dip32[w];
                            const float rpd2 =
                                __builtin_fmaf(dx, pd.x, __builtin_fmaf(dy, pd.y, dz * pd.z)) * inv2;
                            const float qq = r.q + 3.f * rpd2;
                            px = __builtin_fmaf(inv3, __builtin_fmaf(qq, dx, -pd.x), px);
                            py = __builtin_fmaf(inv3, __builtin_fmaf(qq, dy, -pd.y), py);
                            pz = __builtin_fmaf(inv3, __builtin_fmaf(qq, dz, -pd.z), pz);
                        } else {
                            const float sc = r.q * inv3;
                            px = __builtin_fmaf(sc, dx, px);
                            py = __builtin_fmaf(sc, dy, py);
                            pz = __builtin_fmaf(sc, dz, pz);
                        }
                        ncell += mine ? 1u : 0u;
                    }
                    if ((e & 15) == 15) {
                        kahan_add(fx, kx, px);
                        kahan_add(fy, ky, py);
                        kahan_add(fz, kz, pz);
                        px = py = pz = 0.f;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                nl = 0;
            }
            kahan_add(fx, kx, px);
            kahan_add(fy, ky, py);
            kahan_add(fz, kz, pz);
        }
        if (valid) {
            const double Kf = QE / (K4PE0 * p.R * p.R);
            io.out[3 * k] = (io.acc ? io.out[3 * k] : 0.0) + Kf * ((double)fx - (double)kx);
            io.out[3 * k + 1] = (io.acc ? io.out[3 * k + 1] : 0.0) + Kf * ((double)fy - (double)ky);
            io.out[3 * k + 2] = (io.acc ? io.out[3 * k + 2] : 0.0) + Kf * ((double)fz - (double)kz);
            if (COUNT && io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        tot_cell += valid ? ncell : 0u;
        tot_leaf += valid ? nleaf : 0u;
        tot_open += lane == 0 ? nopen : 0u;
    }
    if (COUNT) {
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
            tot_open += __shfl_xor(tot_open, d, WAVE);
        }
        if (lane == 0) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
            atomicAdd(&g->n_pops, tot_pops);
        }
    }
}

// initial accumulator of every target: the spheroid grain's analytic field
// (SPH:1815-1818 seeds E with it); for phi_at, the grain potential at the
// absolute coordinates as phi_at uses them (SPH:1924-1953); for the
// potential a particle feels, the grain potential at its grain-relative
// position (SPH:2040-2084, 2100-2103).  Zero for the spherical grain, which
// is a leaf of the tree.
template <int MODE, int KIND>
__global__ void k_grain_init(const InstParams* __restrict__ P, const InstState* __restrict__ S,
                             PState st, WalkIO io)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    uint32_t inst, self;
    double tx, ty, tz;
    if (!load_target<MODE>(P, st, io, k, ntg, inst, self, tx, ty, tz)) return;
    const InstParams& p = P[inst];
    if (KIND == KIND_FIELD) {
        double e[3] = {0.0, 0.0, 0.0};
        if (p.geometry != 0) grain_field(p, S[inst].grain_q, tx, ty, tz, e);
        io.out[3 * k] = e[0];
        io.out[3 * k + 1] = e[1];
        io.out[3 * k + 2] = e[2];
    } else if (KIND == KIND_PHI) {
        io.out[k] = p.geometry != 0 ? grain_phi_abs(p, S[inst].grain_q, tx, ty, tz) : 0.0;
    } else {
        // grain-relative, and nothing for the grain itself (SPH:2040-2050)
        const bool grain = st.ref[k] == p.offset;
        io.out[k] = (p.geometry != 0 && !grain)
                        ? grain_phi_abs(p, S[inst].grain_q, tx - p.grain[0], ty - p.grain[1],
                                        tz - p.grain[2])
                        : 0.0;
    }
}

__global__ void k_reset_tiles(uint32_t* ctr) { *ctr = 0; }

}  // namespace

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t pr;
            if (hipGetDeviceProperties(&pr, dev) == hipSuccess) cus = pr.multiProcessorCount;
        }
    }
    return cus;
}

void launch_walk_reset(Bufs& b, hipStream_t s) { k_reset_tiles<<<1, 1, 0, s>>>(&b.g->tile_ctr); }

static WalkIO step_io(Bufs& b)
{
    WalkIO io{};
    io.targets = b.targets;
    io.n_targets_dev = &b.g->n_targets;
    io.out = b.E;
    io.acc = b.any_spheroid;
    io.tile_ctr = &b.g->tile_ctr;
    io.n_rec = b.n_rec;
    io.dip64 = b.dip64;
    io.dip32 = b.dip32;
    return io;
}

void launch_walk_init(Bufs& b, hipStream_t s)
{
    // only the spheroids' analytic grain has a term to start from
    if (!b.any_spheroid) return;
    WalkIO io = step_io(b);
    const uint32_t nt = b.t_cap ? b.t_cap : 1;
    k_grain_init<MODE_STEP, KIND_FIELD><<<(nt + BT - 1) / BT, BT, 0, s>>>(b.P, b.S, b.st, io);
}

// float32 field walk of the chosen shape: the group criterion
// (k_walk32grp), groups of 8 or 16 targets (k_walkg), or 64 targets stepping
// through the children (k_walk32); instantiated with dipole support only
// for contexts holding DIPOLES instances, and with interaction counting
// (COUNT) for point queries and for step walks while the context counts
template <int MODE, bool COUNT>
static void launch_walk32c(Bufs& b, int blocks, const WalkIO& io, hipStream_t s)
{
    const bool dip = b.dip32 != nullptr;
#define POTG_W32(KERNEL) KERNEL<<<blocks, BT, 0, s>>>(b.P, b.g, b.st, b.rec32, io)
    if (b.precision == POTG_MIXED_GROUP) {
        if (dip) POTG_W32((k_walk32grp<MODE, true, COUNT>));
        else POTG_W32((k_walk32grp<MODE, false, COUNT>));
    } else if (b.group == 64 && b.variant == 1) {
        if (dip) POTG_W32((k_walk32s<MODE, true, COUNT>));
        else POTG_W32((k_walk32s<MODE, false, COUNT>));
    } else if (b.group == 8) {
        if (dip) POTG_W32((k_walkg<8, MODE, true, COUNT>));
        else POTG_W32((k_walkg<8, MODE, false, COUNT>));
    } else if (b.group == 16) {
        if (dip) POTG_W32((k_walkg<16, MODE, true, COUNT>));
        else POTG_W32((k_walkg<16, MODE, false, COUNT>));
    } else {
        if (dip) POTG_W32((k_walk32<MODE, true, COUNT>));
        else POTG_W32((k_walk32<MODE, false, COUNT>));
    }
#undef POTG_W32
}

template <int MODE>
static void launch_walk32(Bufs& b, int blocks, const WalkIO& io, hipStream_t s)
{
    if (MODE == MODE_POINTS || b.count) launch_walk32c<MODE, true>(b, blocks, io, s);
    else launch_walk32c<MODE, false>(b, blocks, io, s);
}

template <int MODE, int KIND, bool COUNT>
static void launch_walk64c(Bufs& b, int blocks, const WalkIO& io, hipStream_t s)
{
    if (b.dip64)
        k_walk64<MODE, KIND, true, COUNT><<<blocks, BT, 0, s>>>(b.P, b.g, b.st, b.rec64, io);
    else
        k_walk64<MODE, KIND, false, COUNT><<<blocks, BT, 0, s>>>(b.P, b.g, b.st, b.rec64, io);
}

template <int MODE, int KIND>
static void launch_walk64(Bufs& b, int blocks, const WalkIO& io, hipStream_t s)
{
    if (MODE == MODE_POINTS || b.count) launch_walk64c<MODE, KIND, true>(b, blocks, io, s);
    else launch_walk64c<MODE, KIND, false>(b, blocks, io, s);
}

void launch_walk(Bufs& b, hipStream_t s)
{
    WalkIO io = step_io(b);
    if (b.precision)
        launch_walk32<MODE_STEP>(b, (int)b.walk_blocks, io, s);
    else
        launch_walk64<MODE_STEP, KIND_FIELD>(b, (int)b.walk_blocks, io, s);
}

void launch_points(Bufs& b, uint32_t n, int kind, hipStream_t s)
{
    WalkIO io{};
    io.pts = b.q_pts;
    io.pt_inst = b.q_inst;
    io.pt_self = b.q_self;
    io.n_pts = n;
    io.out = b.q_out;
    io.acc = 1;
    io.counts = kind == 0 ? b.q_counts : nullptr;
    io.tile_ctr = &b.g->tile_ctr2;
    io.n_rec = b.n_rec;
    io.dip64 = b.dip64;
    io.dip32 = b.dip32;
    k_reset_tiles<<<1, 1, 0, s>>>(&b.g->tile_ctr2);
    const uint32_t nb = (n + BT - 1) / BT ? (n + BT - 1) / BT : 1;
    const int blocks = (int)nb < (int)b.walk_blocks ? (int)nb : (int)b.walk_blocks;
    if (kind == 1) {
        k_grain_init<MODE_POINTS, KIND_PHI><<<nb, BT, 0, s>>>(b.P, b.S, b.st, io);
        launch_walk64<MODE_POINTS, KIND_PHI>(b, blocks, io, s);
    } else {
        k_grain_init<MODE_POINTS, KIND_FIELD><<<nb, BT, 0, s>>>(b.P, b.S, b.st, io);
        if (b.precision)
            launch_walk32<MODE_POINTS>(b, blocks, io, s);
        else
            launch_walk64<MODE_POINTS, KIND_FIELD>(b, blocks, io, s);
    }
}

// Potential energies of a compute_epe step (pot.c:2771-2792): the potential
// each particle feels, from the step's tree, at its half-drifted position,
// before the push moves it; then deterministic per-instance partial sums of
// QE * CHARGE * phi over electrons and ions (EPE_BLOCKS per instance, added
// in a fixed order by the step's last kernel).
__global__ __launch_bounds__(BT) void k_epe_partial(const InstParams* __restrict__ P, PState st,
                                                    const double* __restrict__ phi,
                                                    double* __restrict__ part)
{
    const uint32_t inst = blockIdx.y;
    const InstParams& p = P[inst];
    double se = 0.0, si = 0.0;
    for (uint32_t j = blockIdx.x * BT + threadIdx.x; j < p.tree_n; j += EPE_BLOCKS * BT) {
        const uint32_t k = p.tree_begin + j;
        const int sp = st.spec[k];
        if (sp == 2) se += -QE * phi[k];
        else if (sp == 1) si += QE * phi[k];
    }
    __shared__ double red[2][BT];
    red[0][threadIdx.x] = se;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int w = BT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (inst * EPE_BLOCKS + blockIdx.x)] = red[0][0];
        part[2 * (inst * EPE_BLOCKS + blockIdx.x) + 1] = red[1][0];
    }
}

void launch_epe(Bufs& b, hipStream_t s)
{
    WalkIO io{};
    io.n_pts = b.n_total;
    io.out = b.phi;
    io.acc = 1;
    io.tile_ctr = &b.g->tile_ctr2;
    io.n_rec = b.n_rec;
    io.S = b.S;
    io.dip64 = b.dip64;
    k_reset_tiles<<<1, 1, 0, s>>>(&b.g->tile_ctr2);
    const uint32_t nb = (b.n_total + BT - 1) / BT;
    const int blocks = (int)nb < (int)b.walk_blocks ? (int)nb : (int)b.walk_blocks;
    k_grain_init<MODE_EPE, KIND_EPE><<<nb, BT, 0, s>>>(b.P, b.S, b.st, io);
    launch_walk64<MODE_EPE, KIND_EPE>(b, blocks, io, s);
    k_epe_partial<<<dim3(EPE_BLOCKS, b.n_inst), BT, 0, s>>>(b.P, b.st, b.phi, b.epe_part);
}

// chi_b every 20 steps (pot.c:2409-2418): the potential at one boundary
// point per instance, before any particle moves; one lane walks (a handful
// of nodes at the domain edge)
__global__ __launch_bounds__(64) void k_chi_b(const InstParams* __restrict__ P, InstState* S,
                                              Globals* g, const NodeRec64* __restrict__ r64,
                                              const double* __restrict__ dip64, uint32_t n_inst,
                                              uint32_t n_rec)
{
    if ((g->t % 20) != 0 || step_failed(g)) return;
    const uint32_t inst = blockIdx.x;
    if (inst >= n_inst || threadIdx.x != 0 || S[inst].retired) return;
    const InstParams& p = P[inst];
    const double x = p.chi_pt[0], y = p.chi_pt[1], z = p.chi_pt[2];
    const double theta2 = p.theta2, soft2 = p.soft2;
    double ph = (p.geometry != 0) ? grain_phi_abs(p, S[inst].grain_q, x, y, z) : 0.0;
    uint32_t stk[8 * STK];
    int sp = 0;
    stk[sp++] = inst;
    uint32_t pops = 0;
    while (sp) {
        if (++pops > n_rec) {
            raise_err(g, KERR_DEPTH, inst);
            return;
        }
        const uint32_t w = stk[--sp];
        const NodeRec64 r = r64[w];
        const double dx = x - r.cx, dy = y - r.cy, dz = z - r.cz;
        const double d2 = (dx * dx) + (dy * dy) + (dz * dz);
        if (r.cnt == LEAF) {
            ph += (QE * r.q) / (K4PE0 * sqrt(d2 + soft2));
        } else if (r.size2 < theta2 * d2) {
            if (p.dipoles && dip64) {
                if (r.part) {  // |q| of the cell (pot.c:1856-1862)
                    const double* d = dip64 + 3 * (size_t)w;
                    const double dist = sqrt(d2);
                    const double rdp = (dx * d[0]) + (dy * d[1]) + (dz * d[2]);
                    ph += (QE / K4PE0) * ((r.q / dist) + (rdp / (dist * d2)));
                }
            } else if (r.q) {
                ph += (QE * r.q) / (K4PE0 * sqrt(d2));
            }
        } else {
            if (sp + (int)r.cnt > 8 * STK || (r.cnt && r.first <= w) || r.first + r.cnt > n_rec) {
                raise_err(g, KERR_DEPTH, w);
                return;
            }
            for (uint32_t c = 0; c < r.cnt; c++) stk[sp++] = r.first + c;
        }
    }
    S[inst].chi_b = ph / p.kTi_over_e;
}

void launch_chi_b(Bufs& b, hipStream_t s)
{
    k_chi_b<<<b.n_inst, 64, 0, s>>>(b.P, b.S, b.g, b.rec64, b.dip64, b.n_inst, b.n_rec);
}

// ---- point potentials with no allocation (potg_phi_at_noalloc) -----------
//
// The walk of a point query reads, besides the caller's points and result
// array, a tile counter and an instance and excluded-slot entry per point.
// For callers that cannot allocate, that scratch is this array: word 0 is
// the tile counter, then NOALLOC_MAX points' instances, then their excluded
// slots (all NONE: a potential counts every leaf).  Each device holds one
// copy until the process ends; the host reaches it only through the address
// hipGetSymbolAddress gives on that device.  Calls that overlap on one device
// share it and race.
constexpr uint32_t NOALLOC_HEAD = 64;  // words before the per-point arrays
constexpr uint32_t NOALLOC_WORDS = NOALLOC_HEAD + 2 * POTG_NOALLOC_MAX_POINTS;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {

__global__ void k_noalloc_prep(uint32_t* __restrict__ scr, uint32_t inst, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) scr[0] = 0u;  // the walk's tile counter starts from zero
    if (k < n) {
        scr[NOALLOC_HEAD + k] = inst;
        scr[NOALLOC_HEAD + POTG_NOALLOC_MAX_POINTS + k] = NONE;
    }
}

}  // namespace

int noalloc_scratch(uint32_t** out)
{
    // the address of this device's copy, looked up once per device
    constexpr int MAXDEV = 64;
    static std::atomic<uint32_t*> addr[MAXDEV];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) return -1;
    uint32_t* p = addr[dev].load();
    if (!p) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return -1;
        p = (uint32_t*)q;
        addr[dev].store(p);
    }
    *out = p;
    return 0;
}

void launch_phi_noalloc(Bufs& b, uint32_t* scr, uint32_t inst, const double* pts, uint32_t n,
                        double* out, hipStream_t s)
{
    WalkIO io{};
    io.pts = pts;
    io.pt_inst = scr + NOALLOC_HEAD;
    io.pt_self = scr + NOALLOC_HEAD + POTG_NOALLOC_MAX_POINTS;
    io.n_pts = n;
    io.out = out;
    io.acc = 1;
    io.tile_ctr = scr;
    io.n_rec = b.n_rec;
    io.dip64 = b.dip64;
    const uint32_t nb = (n + BT - 1) / BT ? (n + BT - 1) / BT : 1;
    const int blocks = (int)nb < (int)b.walk_blocks ? (int)nb : (int)b.walk_blocks;
    k_noalloc_prep<<<nb, BT, 0, s>>>(scr, inst, n);
    k_grain_init<MODE_POINTS, KIND_PHI><<<nb, BT, 0, s>>>(b.P, b.S, b.st, io);
    launch_walk64<MODE_POINTS, KIND_PHI>(b, blocks, io, s);
}

}  // namespace potg
