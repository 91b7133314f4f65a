// potg_tree.hip -- octree build on gfx950.
//
// The reference inserts particles one at a time into an octree of leaf
// capacity 1 whose boxes come from the float64 midpoint recurrence
// mid = (x1 + x2) / 2 (build_tree, pot.c:1457-1643).  The node set of such a
// tree depends only on the positions, so it is rebuilt here in parallel:
//
//   1. keys: every particle descends key_levels levels of the same
//      recurrence; the digits (numbered in the reference's child order)
//      form a sort key behind the instance number;
//   2. an LSD radix sort orders the particles; runs of equal keys are
//      ordered by descending further (rare);
//   3. lcp[i] = number of leading digits particles i and i+1 share;
//      internal node (a, l) exists for every level l with
//      lcp[a-1] < l <= lcp[a] -- numbering the nodes by (a, l) gives a
//      pre-order; parents come from one binary search over the keys;
//   4. child slots, child counts and a scan lay the children of every node
//      out contiguously for the walk;
//   5. monopole moments are folded bottom-up, children in the reference
//      order (compute_node_values / update_node_moments, pot.c:1290-1402),
//      with the same float64 operations, so every node matches the
//      reference bit for bit.
#include "potg_internal.h"
#include "potg_prim.h"

namespace potg {

namespace {

constexpr int BT = 256;
inline uint32_t grid(uint32_t n) { return n ? (n + BT - 1) / BT : 1; }

__device__ inline int charge_of(uint8_t s, int32_t grain_q)
{
    return s == 0 ? grain_q : (s == 1 ? 1 : -1);
}

// digit of a position at depth `level` (1-based) and the box at that depth
__device__ inline int descend_to(double px, double py, double pz, double root_x2, int level,
                                 double& x1, double& x2)
{
    double y1 = 0.0, y2 = root_x2, z1 = 0.0, z2 = root_x2;
    int d = 0;
    x1 = 0.0;
    x2 = root_x2;
    for (int l = 0; l < level; l++) d = descend1(px, py, pz, x1, x2, y1, y2, z1, z2);
    return d;
}

// first depth at which two positions fall into different children; MAXD
// when they never separate (coincident)
__device__ inline int split_depth(double ax, double ay, double az, double bx, double by, double bz,
                                  double root_x2)
{
    double x1 = 0.0, x2 = root_x2, y1 = 0.0, y2 = root_x2, z1 = 0.0, z2 = root_x2;
    for (int l = 0; l < MAXD; l++) {
        const double mx = (x1 + x2) / 2.0, my = (y1 + y2) / 2.0, mz = (z1 + z2) / 2.0;
        const bool gax = !(ax < mx), gay = !(ay < my), gaz = !(az < mz);
        const bool gbx = !(bx < mx), gby = !(by < my), gbz = !(bz < mz);
        if (gax != gbx || gay != gby || gaz != gbz) return l;
        if (gax) x1 = mx; else x2 = mx;
        if (gay) y1 = my; else y2 = my;
        if (gaz) z1 = mz; else z2 = mz;
    }
    return MAXD;
}

// ---------------------------------------------------------------------
// One block: every thread reads the step counter before thread 0 advances
// it (the barrier), however many instances there are.
__global__ void k_begin_step(InstParams* __restrict__ P, InstState* __restrict__ S, Globals* g,
                             uint32_t n_inst)
{
    const uint64_t t = g->t + 1;  // the step being taken (pot.c:4261)
    for (uint32_t k = threadIdx.x; k < n_inst; k += blockDim.x) {
        // first stepped index: the lowest index whose species is due
        // (pot.c:2430-2434); the grain itself is never stepped here
        uint32_t start = NONE;
        for (int s = 1; s < 3; s++)
            if (P[k].first_idx[s] != NONE && (t % P[k].tss[s]) == 0 && P[k].first_idx[s] < start)
                start = P[k].first_idx[s];
        S[k].start = start;
        S[k].acc_q = 0;
        S[k].acc_e = 0;
        S[k].acc_i = 0;
        S[k].n_pushed = 0;
        S[k].n_reinj_abs = 0;
        S[k].n_reinj_out = 0;
        S[k].n_wall_inward = 0;
        for (int c = 0; c < 3; c++) {
            S[k].acc_p_lo[c] = 0ull;
            S[k].acc_p_hi[c] = 0ull;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        g->t = t;
        g->tile_ctr = 0;
        g->tile_ctr2 = 0;
        g->n_reinj = 0;
    }
}

__global__ void k_keys(const InstParams* __restrict__ P, PState st, uint32_t n, int levels,
                       uint64_t* __restrict__ key, uint32_t* __restrict__ perm, Globals* g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[i] = i;
    const uint32_t inst = st.inst[i];
    const InstParams& p = P[inst];
    if (p.geometry != 0 && st.spec[i] == 0) {
        key[i] = ~0ull;  // the analytic spheroid grain stays out of the tree (SPH:1516-1521)
        return;
    }
    const double px = st.x[i], py = st.y[i], pz = st.z[i];
    const double r2 = p.root_x2;
    if (!(px >= 0.0) || px > r2 || !(py >= 0.0) || py > r2 || !(pz >= 0.0) || pz > r2) {
        raise_err(g, KERR_OUT_OF_BOX, st.ref[i]);
        key[i] = (uint64_t)inst << (3 * levels);
        return;
    }
    double x1 = 0.0, x2 = r2, y1 = 0.0, y2 = r2, z1 = 0.0, z2 = r2;
    uint64_t k = inst;
    for (int l = 0; l < levels; l++) k = (k << 3) | (uint64_t)descend1(px, py, pz, x1, x2, y1, y2, z1, z2);
    key[i] = k;
}

// order runs of equal keys by their deeper digits (insertion sort per run)
__global__ void k_tiefix(const uint64_t* __restrict__ key, uint32_t* __restrict__ perm,
                         const InstParams* __restrict__ P, PState st, uint32_t n_tree, Globals* g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= n_tree) return;
    if (key[i] != key[i + 1] || (i > 0 && key[i - 1] == key[i])) return;
    uint32_t e = i + 1;
    while (e + 1 < n_tree && key[e + 1] == key[i]) e++;
    const double r2 = P[st.inst[perm[i]]].root_x2;
    for (uint32_t a = i + 1; a <= e; a++) {
        const uint32_t pa = perm[a];
        uint32_t b = a;
        while (b > i) {
            const uint32_t pb = perm[b - 1];
            // does pa sort before pb?  compare digits at the first split
            const int d = split_depth(st.x[pa], st.y[pa], st.z[pa], st.x[pb], st.y[pb], st.z[pb], r2);
            if (d >= MAXD) {
                raise_err(g, KERR_COINCIDENT, st.ref[pa]);
                return;
            }
            double x1, x2;
            const int da = descend_to(st.x[pa], st.y[pa], st.z[pa], r2, d + 1, x1, x2);
            const int db = descend_to(st.x[pb], st.y[pb], st.z[pb], r2, d + 1, x1, x2);
            if (da >= db) break;
            perm[b] = pb;
            b--;
        }
        perm[b] = pa;
    }
}

// the particles into sorted order; the grain of each instance notes its
// slot (the potential-energy pass reads the grain's potential there)
__global__ void k_permute(const uint32_t* __restrict__ perm, PState src, PState dst, uint32_t n,
                          const InstParams* __restrict__ P, InstState* __restrict__ S)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    dst.x[i] = src.x[j];
    dst.y[i] = src.y[j];
    dst.z[i] = src.z[j];
    dst.vx[i] = src.vx[j];
    dst.vy[i] = src.vy[j];
    dst.vz[i] = src.vz[j];
    const uint32_t r = src.ref[j];
    const uint16_t inst = src.inst[j];
    const uint8_t spec = src.spec[j];
    dst.ref[i] = r;
    dst.inst[i] = inst;
    dst.spec[i] = spec;
    if (spec == 0 && r == P[inst].offset) S[inst].grain_slot = i;
}

__global__ void k_lcp(const uint64_t* __restrict__ key, PState st, const InstParams* __restrict__ P,
                      uint32_t n_tree, int levels, int32_t* __restrict__ lcp,
                      uint32_t* __restrict__ cnt, Globals* g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = -1;
    if (i + 1 < n_tree && st.inst[i] == st.inst[i + 1]) {
        const uint64_t dmask = (levels >= 21) ? ~0ull >> 1 : ((1ull << (3 * levels)) - 1);
        const uint64_t diff = (key[i] ^ key[i + 1]) & dmask;
        if (diff) {
            const int p = 63 - __clzll(diff);
            v = levels - 1 - p / 3;
        } else {
            v = split_depth(st.x[i], st.y[i], st.z[i], st.x[i + 1], st.y[i + 1], st.z[i + 1],
                            P[st.inst[i]].root_x2);
            if (v >= MAXD - 1) raise_err(g, KERR_COINCIDENT, st.ref[i]);
        }
    }
    if (i < n_tree) lcp[i] = v;
    // deepest internal level + 1: reduced over the block, and the atomic
    // only where it raises the value (it rarely does after the first step)
    __shared__ uint32_t wmax[BT / WAVE];
    uint32_t m = (uint32_t)(v + 1);
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t o = __shfl_xor(m, d, WAVE);
        m = o > m ? o : m;
    }
    if (__lane_id() == 0) wmax[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BT / WAVE; w++) m = wmax[w] > m ? wmax[w] : m;
        if (m > *(volatile const uint32_t*)&g->max_depth) atomicMax(&g->max_depth, m);
    }
}

__global__ void k_counts(const int32_t* __restrict__ lcp, uint32_t n_tree, uint32_t* __restrict__ cnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tree) return;
    const int lo = (i == 0) ? -1 : lcp[i - 1];
    const int hi = lcp[i];
    cnt[i] = hi > lo ? (uint32_t)(hi - lo) : 0u;
}

__global__ void k_owners(const int32_t* __restrict__ lcp, const uint32_t* __restrict__ base,
                         uint32_t n_tree, uint32_t cap, uint32_t* __restrict__ owner,
                         int32_t* __restrict__ level, uint32_t* __restrict__ slots,
                         Globals* g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tree) return;
    const int lo = (i == 0) ? -1 : lcp[i - 1];
    const int hi = lcp[i];
    uint32_t k = base[i];
    for (int l = lo + 1; l <= hi; l++, k++) {
        if (k >= cap) {
            raise_err(g, KERR_CAPACITY, k);
            return;
        }
        owner[k] = i;
        level[k] = l;
#pragma unroll
        for (int s = 0; s < 8; s++) slots[8 * k + s] = NONE;
    }
}

__device__ inline uint32_t n_int_of(const Globals* g, uint32_t cap)
{
    const uint32_t n = g->n_internal;
    return n < cap ? n : cap;
}

// start of the level-L node that contains sorted slot p: the first slot in
// [tb, p] whose key is not below p's key prefix.  A backward galloping
// search brackets it, then bisection: the cells of the deep levels, where
// most nodes are, hold few slots.
__device__ inline uint32_t range_start(const uint64_t* key, const int32_t* lcp, uint32_t tb,
                                       uint32_t p, int L, int levels)
{
    if (L <= levels) {
        const int sh = 3 * (levels - L);
        const uint64_t low = (sh >= 64) ? ~0ull : ((1ull << sh) - 1ull);
        const uint64_t pk = key[p] & ~low;
        uint32_t lo = tb, hi = p;
        for (uint64_t step = 1; step <= (uint64_t)(p - tb); step <<= 1) {
            const uint32_t j = p - (uint32_t)step;
            if (key[j] < pk) {
                lo = j + 1;
                break;
            }
            hi = j;
        }
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (key[mid] < pk) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    uint32_t j = p;
    while (j > tb && lcp[j - 1] >= L) j--;
    return j;
}

__device__ inline uint32_t node_id(const int32_t* lcp, const uint32_t* base, uint32_t a, int l)
{
    const int lo = (a == 0) ? -1 : lcp[a - 1];
    return base[a] + (uint32_t)(l - lo - 1);
}

__device__ inline int digit_of(const uint64_t* key, PState st, uint32_t i, int l, int levels,
                               double root_x2)
{
    if (l <= levels) return (int)((key[i] >> (3 * (levels - l))) & 7u);
    double x1, x2;
    return descend_to(st.x[i], st.y[i], st.z[i], root_x2, l, x1, x2);
}

// every internal node (k < n_internal) and every leaf (tree slot) enters
// its parent's child slot at its digit; internal nodes also get their box
// size, roots their walk id
__global__ void k_link(const uint64_t* __restrict__ key, const int32_t* __restrict__ lcp,
                       const uint32_t* __restrict__ base, const uint32_t* __restrict__ owner,
                       const int32_t* __restrict__ level, PState st,
                       const InstParams* __restrict__ P, uint32_t n_tree, int levels,
                       uint32_t cap, Globals* g, uint32_t* __restrict__ slots,
                       uint32_t* __restrict__ wid_int, uint32_t* __restrict__ wid_leaf,
                       double* __restrict__ size2)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nint = n_int_of(g, cap);
    if (k < nint) {
        const uint32_t a = owner[k];
        const int l = level[k];
        const uint32_t inst = st.inst[a];
        const InstParams& p = P[inst];
        double x1, x2;
        const int d = descend_to(st.x[a], st.y[a], st.z[a], p.root_x2, l, x1, x2);
        size2[k] = (x2 - x1) * (x2 - x1);
        if (l == 0) {
            wid_int[k] = inst;
        } else {
            const int lo = (a == 0) ? -1 : lcp[a - 1];
            uint32_t par;
            if (l > lo + 1) {
                par = k - 1;
            } else {
                const uint32_t a2 = range_start(key, lcp, p.tree_begin, a, l - 1, levels);
                par = node_id(lcp, base, a2, l - 1);
            }
            if (par >= nint) raise_err(g, KERR_CAPACITY, k);
            else slots[8 * par + d] = k;
        }
    }
    if (k < n_tree) {
        const uint32_t i = k;
        const uint32_t inst = st.inst[i];
        const InstParams& p = P[inst];
        const int lo = (i == p.tree_begin) ? -1 : lcp[i - 1];
        const int hi = lcp[i];
        const int L = hi > lo ? hi : lo;
        if (L < 0) {
            wid_leaf[i] = inst;  // a one-particle tree: the leaf is the root
        } else {
            uint32_t par;
            if (hi > lo) {
                par = node_id(lcp, base, i, hi);
            } else {
                const uint32_t a2 = range_start(key, lcp, p.tree_begin, i - 1, L, levels);
                par = node_id(lcp, base, a2, L);
            }
            if (par >= nint) {
                raise_err(g, KERR_CAPACITY, i);
                wid_leaf[i] = inst;
            } else {
                slots[8 * par + digit_of(key, st, i, L + 1, levels, p.root_x2)] = LEAF_BIT | i;
            }
        }
    }
}

__global__ void k_ccount(const uint32_t* __restrict__ slots, const Globals* __restrict__ g,
                         uint32_t cap, uint32_t* __restrict__ ccount)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_int_of(g, cap)) return;
    uint32_t c = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) c += slots[8 * k + s] != NONE;
    ccount[k] = c;
}

__global__ void k_wids(const uint32_t* __restrict__ slots, const uint32_t* __restrict__ cfirst,
                       const Globals* __restrict__ g, uint32_t cap, uint32_t* __restrict__ wid_int,
                       uint32_t* __restrict__ wid_leaf)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_int_of(g, cap)) return;
    uint32_t w = cfirst[k];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const uint32_t c = slots[8 * k + s];
        if (c == NONE) continue;
        if (c & LEAF_BIT) wid_leaf[c & ~LEAF_BIT] = w;
        else wid_int[c] = w;
        w++;
    }
}

// roots of instances without tree particles: an empty cell
__global__ void k_empty_roots(const InstParams* __restrict__ P, uint32_t n_inst,
                              NodeRec64* __restrict__ r64, NodeRec32* __restrict__ r32)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_inst || P[k].tree_n != 0) return;
    NodeRec64 e;
    e.cx = e.cy = e.cz = 0.0;
    e.size2 = 0.0;
    e.q = 0;
    e.first = 0;
    e.cnt = 0;
    e.part = 0;
    r64[k] = e;
    if (r32) {
        NodeRec32 f;
        f.cx = f.cy = f.cz = 0.f;
        f.size2 = 0.f;
        f.q = 0.f;
        f.first = 0;
        f.cnt = 0;
        f.part = 0;
        r32[k] = f;
    }
}

__device__ inline void put32(NodeRec32* r32, uint32_t w, const InstParams& p, double cx, double cy,
                             double cz, double s2, int q, uint32_t first, uint32_t cnt, uint32_t part)
{
    NodeRec32 f;
    f.cx = (float)((cx - p.R) * p.invR);
    f.cy = (float)((cy - p.R) * p.invR);
    f.cz = (float)((cz - p.R) * p.invR);
    // the opening test of the float32 walks is size2 < d2: the box size
    // squared over theta^2, in units of R^2 (with theta = 0 every cell opens)
    f.size2 = s2 > 0.0 ? (float)(s2 * p.open32) : 0.f;
    f.q = (float)q;
    f.first = first;
    f.cnt = cnt;
    f.part = part;
    r32[w] = f;
}

// leaf records: every tree particle is a leaf of charge CHARGE[species]
// (the grain: its current charge, pot.c:1318-1328) and no dipole moment
__global__ void k_leaf_records(PState st, const InstParams* __restrict__ P,
                               const InstState* __restrict__ S, uint32_t n_tree, uint32_t n_rec,
                               const uint32_t* __restrict__ wid_leaf, NodeRec64* __restrict__ r64,
                               NodeRec32* __restrict__ r32, double* __restrict__ dip64,
                               float4* __restrict__ dip32)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tree) return;
    const uint32_t inst = st.inst[i];
    NodeRec64 e;
    e.cx = st.x[i];
    e.cy = st.y[i];
    e.cz = st.z[i];
    e.size2 = 0.0;
    e.q = charge_of(st.spec[i], S[inst].grain_q);
    e.first = 0;
    e.cnt = LEAF;
    e.part = i;
    const uint32_t w = wid_leaf[i];
    if (w >= n_rec) return;
    r64[w] = e;
    if (r32) put32(r32, w, P[inst], e.cx, e.cy, e.cz, 0.0, e.q, 0, LEAF, i);
    if (dip64) {
        dip64[3 * (size_t)w] = 0.0;
        dip64[3 * (size_t)w + 1] = 0.0;
        dip64[3 * (size_t)w + 2] = 0.0;
    }
    if (dip32) dip32[w] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Internal nodes bucketed by depth, so that each level's moment pass visits
// only that level's nodes: a counting sort without atomics.  Block b owns a
// contiguous chunk of the node ids; per-block level histograms are scanned
// level-major, then each block writes its nodes at its offsets.  Order
// within a level is free (the nodes of one level are independent).
__device__ inline void level_chunk(const Globals* g, uint32_t cap, uint32_t& lo, uint32_t& hi)
{
    const uint32_t nint = n_int_of(g, cap);
    const uint32_t chunk = (nint + LVL_BLOCKS - 1) / LVL_BLOCKS;
    lo = blockIdx.x * chunk;
    hi = lo + chunk < nint ? lo + chunk : nint;
    if (lo > hi) lo = hi;
}

__global__ __launch_bounds__(BT) void k_level_count(const int32_t* __restrict__ level,
                                                    const Globals* __restrict__ g, uint32_t cap,
                                                    uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[MAXD];
    if (threadIdx.x < MAXD) h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t lo, hi;
    level_chunk(g, cap, lo, hi);
    for (uint32_t k = lo + threadIdx.x; k < hi; k += BT) {
        const int32_t L = level[k];
        if (L >= 0 && L < MAXD) atomicAdd(&h[L], 1u);
    }
    __syncthreads();
    if (threadIdx.x < MAXD) hist[threadIdx.x * LVL_BLOCKS + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of the MAXD x LVL_BLOCKS histogram in level-major order;
// lstart[L] = first slot of level L, lstart[MAXD] = all internal nodes
__global__ __launch_bounds__(1024) void k_level_offsets(const uint32_t* __restrict__ hist,
                                                        uint32_t* __restrict__ off,
                                                        uint32_t* __restrict__ lstart)
{
    constexpr uint32_t N = MAXD * LVL_BLOCKS;
    constexpr uint32_t PER = (N + 1023) / 1024;
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    uint32_t sum = 0;
    for (uint32_t j = 0; j < PER; j++) {
        const uint32_t i = t * PER + j;
        if (i < N) sum += hist[i];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t v = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive prefix of this thread's run
    for (uint32_t j = 0; j < PER; j++) {
        const uint32_t i = t * PER + j;
        if (i < N) {
            off[i] = run;
            if (i % LVL_BLOCKS == 0) lstart[i / LVL_BLOCKS] = run;
            run += hist[i];
        }
    }
    if (t == 1023) lstart[MAXD] = part[1023];
}

__global__ __launch_bounds__(BT) void k_level_scatter(const int32_t* __restrict__ level,
                                                      const Globals* __restrict__ g, uint32_t cap,
                                                      const uint32_t* __restrict__ off,
                                                      uint32_t* __restrict__ by_level)
{
    __shared__ uint32_t h[MAXD], base[MAXD];
    uint32_t lo, hi;
    level_chunk(g, cap, lo, hi);
    if (threadIdx.x < MAXD) base[threadIdx.x] = off[threadIdx.x * LVL_BLOCKS + blockIdx.x];
    for (uint32_t k0 = lo; k0 < hi; k0 += BT) {
        if (threadIdx.x < MAXD) h[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t k = k0 + threadIdx.x;
        int32_t L = -1;
        uint32_t rank = 0;
        if (k < hi) {
            L = level[k];
            if (L >= 0 && L < MAXD) rank = atomicAdd(&h[L], 1u);
            else L = -1;
        }
        __syncthreads();
        if (L >= 0) {
            const uint32_t pos = base[L] + rank;
            if (pos < cap) by_level[pos] = k;
        }
        __syncthreads();
        if (threadIdx.x < MAXD) base[threadIdx.x] += h[threadIdx.x];
    }
}

// the moments of internal node k, its children folded in the reference
// order (compute_node_values / update_node_moments, pot.c:1290-1402):
// charge, |charge|, centre of |charge| and, for DIPOLES instances, the
// dipole moment about that centre.  The children are read from their
// records, contiguous in that order and finished before this level: the
// leaves by k_leaf_records, internal children by the deeper levels.
__device__ void level_node_moments(uint32_t k, PState st, const InstParams* __restrict__ P,
                                   uint32_t cap, const uint32_t* __restrict__ owner,
                                   const uint32_t* __restrict__ ccount,
                                   const uint32_t* __restrict__ cfirst,
                                   const uint32_t* __restrict__ wid_int,
                                   const double* __restrict__ size2,
                                   NodeRec64* __restrict__ r64, NodeRec32* __restrict__ r32,
                                   double* __restrict__ dip64, float4* __restrict__ dip32,
                                   uint32_t n_rec)
{
    if (k >= cap) return;
    const uint32_t w = wid_int[k];
    if (w >= n_rec) return;
    const uint32_t first = cfirst[k];
    const uint32_t end = first + ccount[k] < n_rec ? first + ccount[k] : n_rec;
    const uint32_t inst = st.inst[owner[k]];
    const bool dip = dip64 && P[inst].dipoles;
    double cx = 0.0, cy = 0.0, cz = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    int q = 0;
    uint32_t aq = 0;
    // every child's record in flight at once, then the serial fold
    NodeRec64 ch[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++)
        if (first + j < end) ch[j] = r64[first + j];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; jj++) {
        if (first + jj >= end) break;
        const uint32_t j = first + jj;
        const NodeRec64& c = ch[jj];
        const int qc = c.q;
        const uint32_t aqc = c.cnt == LEAF ? (uint32_t)abs(qc) : c.part;
        if (aqc) {
            // update_node_moments (pot.c:1290-1310), same float64 operations
            const double ox = cx, oy = cy, oz = cz;
            const double den = (double)(aq + aqc);
            cx = ((double)aq * cx + (double)aqc * c.cx) / den;
            cy = ((double)aq * cy + (double)aqc * c.cy) / den;
            cz = ((double)aq * cz + (double)aqc * c.cz) / den;
            if (dip) {
                // shift the dipole to the new centre, then add the child's
                // (zero for a leaf)
                px -= (cx - ox) * (double)q;
                py -= (cy - oy) * (double)q;
                pz -= (cz - oz) * (double)q;
                px += dip64[3 * (size_t)j] - (cx - c.cx) * (double)qc;
                py += dip64[3 * (size_t)j + 1] - (cy - c.cy) * (double)qc;
                pz += dip64[3 * (size_t)j + 2] - (cz - c.cz) * (double)qc;
            }
            q += qc;
            aq += aqc;
        }
    }
    NodeRec64 e;
    e.cx = cx;
    e.cy = cy;
    e.cz = cz;
    e.size2 = size2[k];
    e.q = q;
    e.first = first;
    e.cnt = ccount[k];
    e.part = aq;
    r64[w] = e;
    if (r32) put32(r32, w, P[inst], cx, cy, cz, e.size2, q, e.first, e.cnt, aq);
    if (dip64) {
        dip64[3 * (size_t)w] = px;
        dip64[3 * (size_t)w + 1] = py;
        dip64[3 * (size_t)w + 2] = pz;
    }
    // the float32 walk works in units of R: p / R
    if (dip32)
        dip32[w] = make_float4((float)(px * P[inst].invR), (float)(py * P[inst].invR),
                               (float)(pz * P[inst].invR), 0.f);
}

// the moments of the internal nodes at depth L; the deeper levels were
// finished by earlier launches
__global__ void k_level_moments(int L, PState st, const InstParams* __restrict__ P,
                                const Globals* __restrict__ g, uint32_t cap,
                                const uint32_t* __restrict__ owner,
                                const uint32_t* __restrict__ ccount,
                                const uint32_t* __restrict__ cfirst,
                                const uint32_t* __restrict__ wid_int,
                                const double* __restrict__ size2, NodeRec64* __restrict__ r64,
                                NodeRec32* __restrict__ r32, double* __restrict__ dip64,
                                float4* __restrict__ dip32, uint32_t n_rec,
                                const uint32_t* __restrict__ lstart,
                                const uint32_t* __restrict__ by_level)
{
    if ((uint32_t)L + 1 > g->max_depth) return;  // no internal node this deep
    const uint32_t lo = lstart[L], hi = lstart[L + 1];
    for (uint32_t idx = lo + blockIdx.x * blockDim.x + threadIdx.x; idx < hi;
         idx += gridDim.x * blockDim.x)
        level_node_moments(by_level[idx], st, P, cap, owner, ccount, cfirst, wid_int, size2, r64,
                           r32, dip64, dip32, n_rec);
}

// levels Lhi down to Llo by one block, a barrier between levels: for the
// thinly populated deep and shallow levels, where a launch per level would
// cost more than the work
__global__ __launch_bounds__(1024) void k_level_moments_block(
    int Lhi, int Llo, PState st, const InstParams* __restrict__ P, const Globals* __restrict__ g,
    uint32_t cap, const uint32_t* __restrict__ owner, const uint32_t* __restrict__ ccount,
    const uint32_t* __restrict__ cfirst, const uint32_t* __restrict__ wid_int,
    const double* __restrict__ size2, NodeRec64* __restrict__ r64, NodeRec32* __restrict__ r32,
    double* __restrict__ dip64, float4* __restrict__ dip32, uint32_t n_rec,
    const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ by_level)
{
    const uint32_t depth = g->max_depth;
    for (int L = Lhi; L >= Llo; L--) {
        if ((uint32_t)L + 1 > depth) continue;  // uniform over the block
        const uint32_t lo = lstart[L], hi = lstart[L + 1];
        for (uint32_t idx = lo + threadIdx.x; idx < hi; idx += blockDim.x)
            level_node_moments(by_level[idx], st, P, cap, owner, ccount, cfirst, wid_int, size2,
                               r64, r32, dip64, dip32, n_rec);
        // the next level up reads what this one wrote
        __syncthreads();
    }
}

}  // namespace

void launch_begin_step(Bufs& b, hipStream_t s)
{
    k_begin_step<<<1, BT, 0, s>>>(b.P, b.S, b.g, b.n_inst);
}

// the leaf records again, from the current state: between the two fields
// of velocity Verlet / Euler-Richardson the reference's walk reads its leaves
// from the live array, which then holds the intermediate state, while the
// cells keep the moments of the step's build (pot.c:1785-1804, 2557-2613)
void launch_leaf_refresh(Bufs& b, hipStream_t s)
{
    const uint32_t n = b.n_tree;
    if (!n) return;
    k_leaf_records<<<grid(n), BT, 0, s>>>(b.st, b.P, b.S, n, b.n_rec, b.wid_leaf, b.rec64,
                                          b.precision ? b.rec32 : nullptr, b.dip64,
                                          b.precision ? b.dip32 : nullptr);
}

void launch_keys_sort_permute(Bufs& b, hipStream_t s)
{
    k_keys<<<grid(b.n_total), BT, 0, s>>>(b.P, b.st, b.n_total, b.key_levels, b.key, b.perm, b.g);
    radix_sort_pairs(b.key, b.key_alt, b.perm, b.perm_alt, b.n_total, b.key_bits, b.sort_tmp,
                     (int)b.sort_ipt, s);
    k_tiefix<<<grid(b.n_tree), BT, 0, s>>>(b.key, b.perm, b.P, b.st, b.n_tree, b.g);
    k_permute<<<grid(b.n_total), BT, 0, s>>>(b.perm, b.st, b.alt, b.n_total, b.P, b.S);
    PState t = b.st;
    b.st = b.alt;
    b.alt = t;
}

void launch_tree(Bufs& b, hipStream_t s)
{
    const uint32_t n = b.n_tree;
    if (n) {
        k_lcp<<<grid(n), BT, 0, s>>>(b.key, b.st, b.P, n, b.key_levels, b.lcp, b.cnt, b.g);
        k_counts<<<grid(n), BT, 0, s>>>(b.lcp, n, b.cnt);
        exclusive_scan(b.cnt, b.base, n, nullptr, &b.g->n_internal, nullptr, b.scan_tmp, 0, s);
        const uint32_t cap = b.cap_int;
        k_owners<<<grid(n), BT, 0, s>>>(b.lcp, b.base, n, cap, b.owner, b.level, b.slots, b.g);
        k_link<<<grid(cap > n ? cap : n), BT, 0, s>>>(b.key, b.lcp, b.base, b.owner, b.level, b.st,
                                                       b.P, n, b.key_levels, cap, b.g, b.slots,
                                                       b.wid_int, b.wid_leaf, b.size2);
        k_ccount<<<grid(cap), BT, 0, s>>>(b.slots, b.g, cap, b.ccount);
        exclusive_scan(b.ccount, b.cfirst, cap, &b.g->n_internal, &b.g->n_children, nullptr,
                       b.scan_tmp, b.n_inst, s);
        k_wids<<<grid(cap), BT, 0, s>>>(b.slots, b.cfirst, b.g, cap, b.wid_int, b.wid_leaf);
    }
    k_empty_roots<<<grid(b.n_inst), BT, 0, s>>>(b.P, b.n_inst, b.rec64, b.precision ? b.rec32 : nullptr);
    if (n) {
        NodeRec32* r32 = b.precision ? b.rec32 : nullptr;
        float4* d32 = b.precision ? b.dip32 : nullptr;
        k_leaf_records<<<grid(n), BT, 0, s>>>(b.st, b.P, b.S, n, b.n_rec, b.wid_leaf, b.rec64,
                                              r32, b.dip64, d32);
        const uint32_t cap = b.cap_int;
        const uint32_t gb = grid(cap) < 2048u ? grid(cap) : 2048u;
        k_level_count<<<LVL_BLOCKS, BT, 0, s>>>(b.level, b.g, cap, b.lvl_hist);
        k_level_offsets<<<1, 1024, 0, s>>>(b.lvl_hist, b.lvl_off, b.lvl_start);
        k_level_scatter<<<LVL_BLOCKS, BT, 0, s>>>(b.level, b.g, cap, b.lvl_off, b.by_level);
        // deepest level first.  Levels below LVL_WIDE_LO hold few nodes in
        // any plasma (their cells are a few mean spacings wide or less), as
        // do those above LVL_WIDE_HI: one block each; the levels between
        // get a launch of their own
        k_level_moments_block<<<1, 1024, 0, s>>>(MAXD - 1, LVL_WIDE_LO, b.st, b.P, b.g, cap,
                                                 b.owner, b.ccount, b.cfirst, b.wid_int, b.size2,
                                                 b.rec64, r32, b.dip64, d32, b.n_rec, b.lvl_start,
                                                 b.by_level);
        for (int L = LVL_WIDE_LO - 1; L > LVL_WIDE_HI; L--)
            k_level_moments<<<gb, BT, 0, s>>>(L, b.st, b.P, b.g, cap, b.owner, b.ccount, b.cfirst,
                                              b.wid_int, b.size2, b.rec64, r32, b.dip64, d32,
                                              b.n_rec, b.lvl_start, b.by_level);
        k_level_moments_block<<<1, 1024, 0, s>>>(LVL_WIDE_HI, 0, b.st, b.P, b.g, cap, b.owner,
                                                 b.ccount, b.cfirst, b.wid_int, b.size2, b.rec64,
                                                 r32, b.dip64, d32, b.n_rec, b.lvl_start,
                                                 b.by_level);
    }
}

}  // namespace potg
