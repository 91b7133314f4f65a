// potg_walk.hip -- Barnes-Hut field and potential walks on gfx950.
//
// One wavefront walks the tree for 64 targets that are neighbours in the
// sort order.  The walk keeps the reference's per-particle opening test
// (field_particle_feels_at, pot.c:1718-1806): the wave shares one
// depth-first traversal whose stack entries carry a 64-bit lane mask; at a
// popped cell every lane in the mask applies its own MAC, lanes that accept
// it add the monopole and lanes that reject it pass the children on with
// their mask.  Each lane therefore visits exactly the nodes, in exactly the
// order, that the reference visits for that particle.  Stack entries are
// child ranges (children of a node are contiguous records), so the stack
// depth is the tree depth; it lives in LDS, the popped node index is
// wave-uniform and node records come through the scalar cache.
//
// FP64: every operation matches the reference's float64 expressions, with
// float64 instead of x87 accumulation.  MIXED: coordinates relative to the
// domain centre in units of R, float32 arithmetic, Kahan-compensated
// per-particle sums, and the result scaled back in float64.
#include "potg_grain.h"
#include "potg_internal.h"

namespace potg {

namespace {

constexpr int BT = 256;
constexpr int NW = BT / WAVE;
constexpr int STK = MAXD + 4;

enum { KIND_FIELD = 0, KIND_PHI = 1 };
enum { MODE_STEP = 0, MODE_POINTS = 1 };

struct WalkIO {
    const uint32_t* targets;
    const uint32_t* n_targets_dev;
    const double* pts;
    const uint32_t* pt_inst;
    const uint32_t* pt_self;
    uint32_t n_pts;
    double* out;
    uint32_t* counts;
    uint32_t* tile_ctr;
    uint32_t n_rec;   // record capacity: child ranges must stay inside it
};

__device__ inline int charge_of(uint8_t s, int32_t grain_q)
{
    return s == 0 ? grain_q : (s == 1 ? 1 : -1);
}

__device__ inline void kahan_add(float& s, float& c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int PREC, int MODE, int KIND>
__global__ __launch_bounds__(BT) void k_walk(const InstParams* __restrict__ P,
                                              const InstState* __restrict__ S, Globals* g,
                                              PState st, const NodeRec64* __restrict__ r64,
                                              const NodeRec32* __restrict__ r32, WalkIO io)
{
    __shared__ uint32_t s_first[NW][STK];
    __shared__ uint32_t s_cnt[NW][STK];
    __shared__ uint64_t s_mask[NW][STK];
    const int wv = threadIdx.x / WAVE;
    const int lane = __lane_id();
    const uint32_t ntg = (MODE == MODE_STEP) ? *io.n_targets_dev : io.n_pts;
    unsigned long long tot_cell = 0, tot_leaf = 0, tot_open = 0;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(io.tile_ctr, 1u);
        tile = rfl(__shfl(tile, 0, WAVE));
        if ((uint64_t)tile * WAVE >= ntg) break;
        const uint32_t k = tile * WAVE + lane;
        const bool valid = k < ntg;

        uint32_t inst = 0, self = NONE;
        double tx = 0.0, ty = 0.0, tz = 0.0;
        if (valid) {
            if (MODE == MODE_STEP) {
                const uint32_t i = io.targets[k];
                self = i;
                inst = st.inst[i];
                const double dts = P[inst].dt * (double)P[inst].tss[st.spec[i]];
                // half drift (pot.c:2478-2480)
                tx = st.x[i] + (0.5 * dts) * st.vx[i];
                ty = st.y[i] + (0.5 * dts) * st.vy[i];
                tz = st.z[i] + (0.5 * dts) * st.vz[i];
            } else {
                inst = io.pt_inst[k];
                self = io.pt_self[k];
                tx = io.pts[3 * k];
                ty = io.pts[3 * k + 1];
                tz = io.pts[3 * k + 2];
            }
        }
        const InstParams& p = P[inst];
        const double theta2 = p.theta2, soft2 = p.soft2;
        const int32_t gq = S[inst].grain_q;

        // accumulators
        double ex = 0.0, ey = 0.0, ez = 0.0, ph = 0.0;
        float fx = 0.f, fy = 0.f, fz = 0.f, cx_ = 0.f, cy_ = 0.f, cz_ = 0.f;
        float ftx = 0.f, fty = 0.f, ftz = 0.f, ftheta2 = 0.f, fsoft2 = 0.f;
        uint32_t ncell = 0, nleaf = 0, nopen = 0;
        if (valid && p.geometry != 0) {
            if (KIND == KIND_FIELD) {
                double e3[3];
                grain_field(p, gq, tx, ty, tz, e3);
                ex = e3[0];
                ey = e3[1];
                ez = e3[2];
            } else {
                ph = grain_phi_abs(p, gq, tx, ty, tz);
            }
        }
        if (PREC == 32) {
            ftx = (float)((tx - p.R) * p.invR);
            fty = (float)((ty - p.R) * p.invR);
            ftz = (float)((tz - p.R) * p.invR);
            ftheta2 = (float)theta2;
            fsoft2 = (float)(soft2 * p.invR * p.invR);
        }

        uint64_t pending = ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const uint32_t gi = rfl(__shfl(inst, lead, WAVE));
            const uint64_t grp = rfl64(ballot(valid && inst == gi));
            pending &= ~grp;

            uint32_t tf = gi, tc = 1;
            uint64_t tm = grp;
            int sp = 0;
            uint32_t pops = 0;
            for (;;) {
                if (++pops > io.n_rec) {
                    raise_err(g, KERR_DEPTH, tf);
                    break;
                }
                if (tc == 0) {
                    if (sp == 0) break;
                    --sp;
                    tf = rfl(s_first[wv][sp]);
                    tc = rfl(s_cnt[wv][sp]);
                    tm = rfl64(s_mask[wv][sp]);
                }
                const uint32_t w = tf + tc - 1;
                tc--;
                const bool act = (tm >> lane) & 1ull;
                bool open = false;
                uint32_t rfirst, rcnt;
                if (PREC == 64 || KIND == KIND_PHI) {
                    const NodeRec64 r = r64[w];
                    rfirst = r.first;
                    rcnt = r.cnt;
                    if (act) {
                        const double dx = tx - r.cx, dy = ty - r.cy, dz = tz - r.cz;
                        const double d2 = (dx * dx) + (dy * dy) + (dz * dz);
                        if (r.cnt == LEAF) {
                            if (KIND == KIND_FIELD) {
                                if (r.part != self) {
                                    if (d2 != 0.0) {
                                        const double dd2 = d2 + soft2;
                                        const double dist = sqrt(dd2);
                                        const double es = (QE * r.q) / ((K4PE0 * dd2) * dist);
                                        ex += es * dx;
                                        ey += es * dy;
                                        ez += es * dz;
                                        nleaf++;
                                    } else {
                                        raise_err(g, KERR_COINCIDENT, r.part);
                                    }
                                }
                            } else {
                                ph += (QE * r.q) / (K4PE0 * sqrt(d2 + soft2));
                                nleaf++;
                            }
                        } else if (r.size2 < theta2 * d2) {
                            if (r.q) {
                                if (KIND == KIND_FIELD) {
                                    const double dist = sqrt(d2);
                                    double es = QE / ((K4PE0 * d2) * dist);
                                    es *= r.q;
                                    ex += es * dx;
                                    ey += es * dy;
                                    ez += es * dz;
                                } else {
                                    ph += (QE * r.q) / (K4PE0 * sqrt(d2));
                                }
                                ncell++;
                            }
                        } else {
                            open = true;
                        }
                    }
                } else {
                    const NodeRec32 r = r32[w];
                    rfirst = r.first;
                    rcnt = r.cnt;
                    if (act) {
                        const float dx = ftx - r.cx, dy = fty - r.cy, dz = ftz - r.cz;
                        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        if (r.cnt == LEAF) {
                            if (r.part != self && d2 != 0.f) {
                                const float inv = __builtin_amdgcn_rsqf(d2 + fsoft2);
                                const float s = (float)r.q * (inv * inv * inv);
                                kahan_add(fx, cx_, s * dx);
                                kahan_add(fy, cy_, s * dy);
                                kahan_add(fz, cz_, s * dz);
                                nleaf++;
                            }
                        } else if (r.size2 < ftheta2 * d2) {
                            if (r.q) {
                                const float inv = __builtin_amdgcn_rsqf(d2);
                                const float s = (float)r.q * (inv * inv * inv);
                                kahan_add(fx, cx_, s * dx);
                                kahan_add(fy, cy_, s * dy);
                                kahan_add(fz, cz_, s * dz);
                                ncell++;
                            }
                        } else {
                            open = true;
                        }
                    }
                }
                const uint64_t om = rfl64(ballot(open));
                if (om && rcnt != LEAF && rcnt) {
                    // children lie after their parent and inside the records
                    if (rfirst <= w || rfirst + rcnt > io.n_rec) {
                        raise_err(g, KERR_CAPACITY, w);
                        break;
                    }
                    if (open) nopen++;
                    if (tc) {
                        if (sp >= STK) {
                            raise_err(g, KERR_DEPTH, w);
                            break;
                        }
                        s_first[wv][sp] = tf;
                        s_cnt[wv][sp] = tc;
                        s_mask[wv][sp] = tm;
                        sp++;
                    }
                    tf = rfirst;
                    tc = rcnt;
                    tm = om;
                }
            }
        }

        if (valid) {
            if (KIND == KIND_FIELD) {
                if (PREC == 32) {
                    const double K = QE / (K4PE0 * p.R * p.R);
                    ex += K * (double)fx;
                    ey += K * (double)fy;
                    ez += K * (double)fz;
                }
                io.out[3 * k] = ex;
                io.out[3 * k + 1] = ey;
                io.out[3 * k + 2] = ez;
            } else {
                io.out[k] = ph;
            }
            if (io.counts) {
                io.counts[3 * k] = ncell;
                io.counts[3 * k + 1] = nleaf;
                io.counts[3 * k + 2] = nopen;
            }
        }
        tot_cell += ncell;
        tot_leaf += nleaf;
        tot_open += nopen;
    }
    if (KIND == KIND_FIELD) {
        // one atomic per wave
        for (int d = 1; d < WAVE; d <<= 1) {
            tot_cell += __shfl_xor(tot_cell, d, WAVE);
            tot_leaf += __shfl_xor(tot_leaf, d, WAVE);
            tot_open += __shfl_xor(tot_open, d, WAVE);
        }
        if (lane == 0 && (tot_cell | tot_leaf | tot_open)) {
            atomicAdd(&g->n_cell, tot_cell);
            atomicAdd(&g->n_leaf, tot_leaf);
            atomicAdd(&g->n_open, tot_open);
        }
    }
}

__global__ void k_reset_tiles(uint32_t* ctr) { *ctr = 0; }

int walk_grid()
{
    static int grid = 0;
    if (!grid) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t pr;
            if (hipGetDeviceProperties(&pr, dev) == hipSuccess) cus = pr.multiProcessorCount;
        }
        grid = cus * 8;
    }
    return grid;
}

}  // namespace

void launch_walk(Bufs& b, hipStream_t s)
{
    WalkIO io{};
    io.targets = b.targets;
    io.n_targets_dev = &b.g->n_targets;
    io.out = b.E;
    io.tile_ctr = &b.g->tile_ctr;
    io.n_rec = b.n_rec;
    if (b.precision)
        k_walk<32, MODE_STEP, KIND_FIELD><<<walk_grid(), BT, 0, s>>>(b.P, b.S, b.g, b.st, b.rec64,
                                                                     b.rec32, io);
    else
        k_walk<64, MODE_STEP, KIND_FIELD><<<walk_grid(), BT, 0, s>>>(b.P, b.S, b.g, b.st, b.rec64,
                                                                     b.rec32, io);
}

void launch_points(Bufs& b, uint32_t n, int kind, hipStream_t s)
{
    WalkIO io{};
    io.pts = b.q_pts;
    io.pt_inst = b.q_inst;
    io.pt_self = b.q_self;
    io.n_pts = n;
    io.out = b.q_out;
    io.counts = kind == 0 ? b.q_counts : nullptr;
    io.tile_ctr = &b.g->tile_ctr2;
    io.n_rec = b.n_rec;
    k_reset_tiles<<<1, 1, 0, s>>>(&b.g->tile_ctr2);
    const int blocks = (int)((n + BT - 1) / BT) < walk_grid() ? (int)((n + BT - 1) / BT) : walk_grid();
    if (kind == 1)
        k_walk<64, MODE_POINTS, KIND_PHI><<<blocks ? blocks : 1, BT, 0, s>>>(b.P, b.S, b.g, b.st,
                                                                          b.rec64, b.rec32, io);
    else if (b.precision)
        k_walk<32, MODE_POINTS, KIND_FIELD><<<blocks ? blocks : 1, BT, 0, s>>>(b.P, b.S, b.g, b.st,
                                                                            b.rec64, b.rec32, io);
    else
        k_walk<64, MODE_POINTS, KIND_FIELD><<<blocks ? blocks : 1, BT, 0, s>>>(b.P, b.S, b.g, b.st,
                                                                            b.rec64, b.rec32, io);
}

// chi_b every 20 steps (pot.c:2409-2418): the potential at one boundary
// point per instance, before any particle moves
__global__ __launch_bounds__(BT) void k_chi_b(const InstParams* __restrict__ P, InstState* S,
                                              Globals* g, PState st,
                                              const NodeRec64* __restrict__ r64, uint32_t n_inst,
                                              uint32_t n_rec)
{
    if ((g->t % 20) != 0) return;
    const uint32_t inst = blockIdx.x;
    if (inst >= n_inst) return;
    // one lane walks; the potential walk is a handful of nodes
    if (threadIdx.x != 0) return;
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
            if (r.q) ph += (QE * r.q) / (K4PE0 * sqrt(d2));
        } else {
            if (sp + (int)r.cnt > 8 * STK || r.first <= w || r.first + r.cnt > n_rec) {
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
    k_chi_b<<<b.n_inst, 64, 0, s>>>(b.P, b.S, b.g, b.st, b.rec64, b.n_inst, b.n_rec);
}

}  // namespace potg
