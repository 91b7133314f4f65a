// potg_api.hip -- the C ABI of include/potg.h: context, device memory,
// upload/download and the step pipeline on one HIP stream.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/potg.h"
#include "potg_internal.h"
#include "potg_prim.h"

using namespace potg;

namespace {

thread_local char g_err[512] = "";

void set_err(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return POTG_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

}  // namespace

struct potg_ctx {
    int device;
    hipStream_t stream;
    Bufs b;
    std::vector<potg_config> cfg;
    std::vector<InstParams> hp;   // host copy of the parameters
    std::vector<std::pair<void*, size_t>> allocs;
    hipEvent_t ev[6];
    int timing;
    double stage_ms[5];
    std::vector<hipEvent_t> tev;  // per-step stage events, a ring of TRING steps
    uint32_t t_next, t_count;
    hipGraphExec_t graph[2];
    int graph_ready;
    uint32_t parity;  // which of the two captured steps comes next
    void* h_aos;      // pinned staging
    void* d_aos;
    size_t aos_cap;
    std::vector<char> uploaded;
    double tuned_ms = 0.0;
};

namespace {

int dev_alloc(potg_ctx* c, void** p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_err("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        return POTG_ERR_NOMEM;
    }
    c->allocs.push_back({*p, bytes});
    return POTG_OK;
}

template <class T>
int alloc_n(potg_ctx* c, T** p, size_t n)
{
    return dev_alloc(c, (void**)p, n * sizeof(T));
}

// "instance i, particle k" for a global particle slot (instances are
// concatenated, each from its offset; k is the reference's index)
std::string slot_name(const potg_ctx* c, uint32_t idx)
{
    for (size_t i = 0; i < c->hp.size(); i++)
        if (idx >= c->hp[i].offset && idx - c->hp[i].offset < c->hp[i].N)
            return "instance " + std::to_string(i) + ", particle " +
                   std::to_string(idx - c->hp[i].offset);
    return "global slot " + std::to_string(idx);
}

int kernel_error(potg_ctx* c)
{
    int32_t err = 0;
    uint32_t idx = 0;
    HIPCHK(hipMemcpyAsync(&err, &c->b.g->err, sizeof(err), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&idx, &c->b.g->err_idx, sizeof(idx), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!err) return POTG_OK;
    int32_t zero = 0;
    HIPCHK(hipMemcpy(&c->b.g->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
    switch (err) {
    case KERR_OUT_OF_BOX:
        set_err("particle (%s) not in simulation area", slot_name(c, idx).c_str());
        return POTG_ERR_OUT_OF_BOX;
    case KERR_COINCIDENT:
        set_err("multiple particles at one location (%s)", slot_name(c, idx).c_str());
        return POTG_ERR_COINCIDENT;
    case KERR_CAPACITY:
        set_err("particle tree larger than its preallocated capacity (%u)", idx);
        return POTG_ERR_NOMEM;
    case KERR_REINJECT:
        set_err("particle (%s) found no admissible SCEPTIC reinjection in 1000 "
                "tries: chi_b bars its species from entering", slot_name(c, idx).c_str());
        return POTG_ERR_REINJECT;
    default:
        set_err("particle tree path length exceeded (node %u)", idx);
        return POTG_ERR_DEPTH;
    }
}

int validate(const potg_config& k, uint32_t i)
{
    if (k.N < 1 || k.N > 0x7fffffffULL) return set_err("instance %u: bad N", i), POTG_ERR_ARG;
    if (k.r == 0 || !(k.mpp > 0) || !(k.dt > 0) || !(k.a > 0) || !(k.soft >= 0) ||
        !(k.theta >= 0) || k.geometry > 2 || !(k.mime > 0))
        return set_err("instance %u: bad parameters", i), POTG_ERR_ARG;
    // the ion sub-cycling factor sqrt(Te mI / (Ti mE)) (pot.c:4157-4159),
    // the shielding length of the reinjection (pot.c:578-580) and chi_b's
    // kT_i/e (pot.c:2278) all divide by a temperature
    if (!(k.Te > 0) || !(k.Ti > 0) || !isfinite(k.Te) || !isfinite(k.Ti))
        return set_err("instance %u: temperatures must be positive and finite", i), POTG_ERR_ARG;
    if (k.geometry == POTG_PROLATE && !(k.A > 1.0))
        return set_err("instance %u: prolate aspect ratio must be greater than 1", i), POTG_ERR_ARG;
    if (k.geometry == POTG_OBLATE && !(k.A < 1.0 && k.A > 0.0))
        return set_err("instance %u: oblate aspect ratio must be less than 1", i), POTG_ERR_ARG;
    if (k.integrator < POTG_BORIS || k.integrator > POTG_EULER_RICHARDSON ||
        (k.out_bounce && !(k.bounce_coe >= 0.0)))
        return set_err("instance %u: bad integrator or bounce coefficient", i), POTG_ERR_ARG;
    return POTG_OK;
}

void fill_params(InstParams& p, const potg_config& k)
{
    memset(&p, 0, sizeof(p));
    p.R = k.r * k.mpp;
    p.root_x2 = 2.0 * k.r * k.mpp;
    p.invR = 1.0 / p.R;
    p.dt = k.dt;
    p.soft2 = k.soft * k.soft;
    p.theta2 = k.theta * k.theta;
    p.open32 = 1.0 / (p.theta2 * p.R * p.R);  // infinite for theta = 0: every cell opens
    p.a = k.a;
    p.A = k.A;
    p.settle = k.settle;
    p.x_drift = k.x_drift;
    p.Te = k.Te;
    p.Ti = k.Ti;
    for (int c = 0; c < 3; c++) p.B[c] = k.B[c];
    p.Bmag = sqrt(k.B[0] * k.B[0] + k.B[1] * k.B[1] + k.B[2] * k.B[2]);
    for (int c = 0; c < 3; c++) p.b[c] = p.Bmag != 0.0 ? k.B[c] / p.Bmag : 0.0;
    p.max_lin_dev = k.max_lin_dev;
    // MASS[] (pot.c:292) and tss[] (pot.c:4157-4159)
    const long double mD = 1.67e-12, mI = k.mime * ME, mE = ME;
    p.mass[0] = (double)mD;
    p.mass[1] = (double)mI;
    p.mass[2] = (double)mE;
    p.tss[2] = 1;
    p.tss[1] = (uint32_t)(int)sqrt((double)(k.Te * mI / (k.Ti * mE)));
    if (p.tss[1] == 0) p.tss[1] = 1;
    p.tss[0] = p.tss[1] * (uint32_t)(int)sqrt((double)(0.1 * mD / mI));
    // Rol of sceptically_reinject (pot.c:578-580)
    double Rol = QE * QE * (double)(k.N - 1) / 2.0;
    Rol /= (4.0 * PI / 3.0) * k.mpp * k.r * EPSILON0 * KB * k.Te;
    p.Rol = sqrt(Rol);
    p.kTi_over_e = KB * k.Ti / (1 * QE);
    p.Nm1 = (double)(k.N - 1);
    p.geometry = k.geometry;
    p.grain_enabled = k.grain_enabled;
    p.sceptic = k.sceptic;
    p.seed = k.seed;
    p.dipoles = k.dipoles != 0;
    p.integrator = k.integrator;
    p.out_bounce = k.out_bounce != 0;
    p.bounce_coe = k.bounce_coe;
    p.N = (uint32_t)k.N;
    for (int s = 0; s < 3; s++) p.first_idx[s] = NONE;
}

// chi_b sampling point of the one worker (measure_chi_b, pot.c:2257-2275)
void chi_point(InstParams& p)
{
    const double azi = 2.0 * PI * 0 / 1;
    const double zen = PI / 2.0;
    const double r = p.R;
    p.chi_pt[0] = p.grain[0] + (r * cos(azi) * sin(zen));
    p.chi_pt[1] = p.grain[1] + (r * sin(azi) * sin(zen));
    p.chi_pt[2] = p.grain[2] + (r * cos(zen));
}

int ensure_staging(potg_ctx* c, size_t bytes)
{
    if (bytes <= c->aos_cap) return POTG_OK;
    if (c->h_aos) (void)hipHostFree(c->h_aos);
    if (c->d_aos) (void)hipFree(c->d_aos);
    c->h_aos = nullptr;
    c->d_aos = nullptr;
    HIPCHK(hipHostMalloc(&c->h_aos, bytes, hipHostMallocDefault));
    HIPCHK(hipMalloc(&c->d_aos, bytes));
    c->aos_cap = bytes;
    return POTG_OK;
}

constexpr uint32_t TRING = 512;

void record(potg_ctx* c, int k)
{
    if (c->timing) (void)hipEventRecord(c->tev[(c->t_next % TRING) * 5 + k], c->stream);
}

void enqueue_step(potg_ctx* c, int compute_epe)
{
    Bufs& b = c->b;
    hipStream_t s = c->stream;
    record(c, 0);
    launch_begin_step(b, s);
    launch_keys_sort_permute(b, s);
    record(c, 1);
    launch_tree(b, s);
    launch_targets(b, s);
    launch_chi_b(b, s);
    launch_walk_init(b, s);
    record(c, 2);
    launch_walk(b, s);
    if (b.integrator != POTG_BORIS) {
        // velocity Verlet / Euler-Richardson: the second field, at the
        // intermediate state, with the leaves moved there
        launch_half(b, s);
        launch_leaf_refresh(b, s);
        launch_walk_init(b, s);
        launch_walk_reset(b, s);
        launch_walk(b, s);
    }
    if (compute_epe) launch_epe(b, s);
    record(c, 3);
    launch_push(b, s);
    launch_finish_step(b, compute_epe, s);
    record(c, 4);
    if (c->timing) {
        c->t_next++;
        if (c->t_count < TRING) c->t_count++;
    }
}

// average stage times over the steps recorded since timing was enabled
int read_stage_times(potg_ctx* c)
{
    if (!c->t_count) return POTG_OK;
    const uint32_t last = (c->t_next + TRING - 1) % TRING;
    HIPCHK(hipEventSynchronize(c->tev[last * 5 + 4]));
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t n = 0; n < c->t_count; n++) {
        const uint32_t r = (c->t_next + TRING - 1 - n) % TRING;
        float ms;
        for (int k = 0; k < 4; k++) {
            HIPCHK(hipEventElapsedTime(&ms, c->tev[r * 5 + k], c->tev[r * 5 + k + 1]));
            acc[k] += ms;
        }
        HIPCHK(hipEventElapsedTime(&ms, c->tev[r * 5], c->tev[r * 5 + 4]));
        acc[4] += ms;
    }
    for (int k = 0; k < 5; k++) c->stage_ms[k] = acc[k] / c->t_count;
    c->t_count = 0;
    return POTG_OK;
}

int all_uploaded(potg_ctx* c)
{
    for (size_t i = 0; i < c->uploaded.size(); i++)
        if (!c->uploaded[i]) {
            set_err("instance %zu has no particles uploaded", i);
            return POTG_ERR_ARG;
        }
    return POTG_OK;
}

void destroy_graphs(potg_ctx* c)
{
    if (c->graph_ready) {
        (void)hipGraphExecDestroy(c->graph[0]);
        (void)hipGraphExecDestroy(c->graph[1]);
        c->graph_ready = 0;
    }
}

}  // namespace

extern "C" {

const char* potg_last_error(void) { return g_err; }

int potg_create(const potg_config* configs, uint32_t n_inst, int32_t precision, potg_ctx** out)
{
    if (!configs || !out || n_inst == 0 || n_inst > 65535 ||
        (precision != POTG_FP64 && precision != POTG_MIXED && precision != POTG_MIXED_GROUP))
        return set_err("potg_create: bad arguments"), POTG_ERR_ARG;
    *out = nullptr;
    uint64_t n_total = 0, n_tree = 0;
    for (uint32_t i = 0; i < n_inst; i++) {
        int rc = validate(configs[i], i);
        if (rc) return rc;
        // the step's launch sequence depends on the integrator
        if (configs[i].integrator != configs[0].integrator)
            return set_err("potg_create: instances of one context need one integrator"),
                   POTG_ERR_ARG;
        n_total += configs[i].N;
        n_tree += configs[i].N - (configs[i].geometry != POTG_SPHERE ? 1 : 0);
    }
    if (n_total > 0x7fffffffULL) return set_err("too many particles"), POTG_ERR_ARG;

    potg_ctx* c = new potg_ctx();
    memset(c->ev, 0, sizeof(c->ev));
    c->timing = 0;
    c->graph_ready = 0;
    c->parity = 0;
    c->h_aos = c->d_aos = nullptr;
    c->aos_cap = 0;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        set_err("no HIP device available");
        delete c;
        return POTG_ERR_HIP;
    }
    for (int k = 0; k < 6; k++) (void)hipEventCreate(&c->ev[k]);
    c->tev.resize(TRING * 5);
    for (auto& e : c->tev) (void)hipEventCreate(&e);
    c->t_next = c->t_count = 0;
    c->cfg.assign(configs, configs + n_inst);
    Bufs& b = c->b;
    memset(&b, 0, sizeof(b));
    b.n_total = (uint32_t)n_total;
    b.n_tree = (uint32_t)n_tree;
    b.n_inst = n_inst;
    b.precision = precision;
    b.sort_ipt = 16;
    // this device's scratch of potg_phi_at_noalloc (its address, once per
    // device; the array itself is static)
    if (noalloc_scratch(&b.noalloc)) b.noalloc = nullptr;
    for (uint32_t i = 0; i < n_inst; i++) b.any_spheroid |= configs[i].geometry != POTG_SPHERE;
    b.count = 1;
    int ib = 0;
    while ((1u << ib) < n_inst) ib++;
    b.key_levels = (63 - ib) / 3 < 21 ? (63 - ib) / 3 : 21;
    b.key_bits = 64;
    // defaults until potg_tune or potg_set_tuning, by the number of 64-target
    // tiles per CU (half the particles are stepped in a step):
    // - 16 or more (about 0.5M particles and up): 64 targets per wavefront
    //   sweeping whole ranges, 8 blocks of 4 waves per CU;
    // - 1 to 16: the same walk on 2 blocks per CU, fewer wavefronts each
    //   with more tiles to take;
    // - below 1 (under about 32k particles) the tiles would leave most SIMDs
    //   idle: 8 targets per wavefront with 8 children evaluated at once, 2
    //   blocks per CU.
    // The float64 walk has one shape: 64 targets, 8 blocks per CU from 16
    // tiles per CU, 4 below.
    {
        const uint64_t cus = (uint64_t)device_cus();
        const uint64_t tiles = n_total / 128u;
        const bool large = tiles >= 16u * cus, tiny = tiles < cus;
        if (precision == POTG_MIXED) {
            b.walk_blocks = (large ? 8u : 2u) * (uint32_t)cus;
            b.group = tiny ? 8u : 64u;
            b.variant = tiny ? 0u : 1u;
        } else {
            b.walk_blocks = (large ? 8u : 4u) * (uint32_t)cus;
            b.group = 64u;
            b.variant = 0u;
        }
    }

    c->hp.resize(n_inst);
    uint32_t off = 0, tb = 0;
    for (uint32_t i = 0; i < n_inst; i++) {
        fill_params(c->hp[i], configs[i]);
        c->hp[i].offset = off;
        c->hp[i].tree_begin = tb;
        c->hp[i].tree_n = (uint32_t)configs[i].N - (configs[i].geometry != POTG_SPHERE ? 1 : 0);
        off += (uint32_t)configs[i].N;
        tb += c->hp[i].tree_n;
    }

    const size_t n = b.n_total, nt = b.n_tree ? b.n_tree : 1;
    int rc = 0;
    PState* S2[2] = {&b.st, &b.alt};
    for (int k = 0; k < 2 && !rc; k++) {
        PState& p = *S2[k];
        rc |= alloc_n(c, &p.x, n) | alloc_n(c, &p.y, n) | alloc_n(c, &p.z, n);
        rc |= alloc_n(c, &p.vx, n) | alloc_n(c, &p.vy, n) | alloc_n(c, &p.vz, n);
        rc |= alloc_n(c, &p.ref, n) | alloc_n(c, &p.inst, n) | alloc_n(c, &p.spec, n);
    }
    rc |= alloc_n(c, &b.P, n_inst) | alloc_n(c, &b.S, n_inst) | alloc_n(c, &b.g, 1);
    rc |= alloc_n(c, &b.key, n) | alloc_n(c, &b.key_alt, n);
    rc |= alloc_n(c, &b.perm, n) | alloc_n(c, &b.perm_alt, n);
    rc |= alloc_n(c, &b.sort_tmp, radix_temp_words((uint32_t)n));
    rc |= alloc_n(c, &b.scan_tmp, scan_temp_words((uint32_t)(8 * n + 64)));
    // internal nodes: about N/2 for uniform particles (single-child chains
    // between close pairs add a few percent); 2N + 64 leaves ample room and
    // every kernel checks the bound
    const size_t ci = 2 * nt + 64;
    b.cap_int = (uint32_t)ci;
    rc |= alloc_n(c, &b.lcp, nt) | alloc_n(c, &b.cnt, nt) | alloc_n(c, &b.base, nt);
    rc |= alloc_n(c, &b.owner, ci) | alloc_n(c, &b.level, ci);
    rc |= alloc_n(c, &b.slots, 8 * ci) | alloc_n(c, &b.ccount, ci) | alloc_n(c, &b.cfirst, ci);
    rc |= alloc_n(c, &b.wid_int, ci) | alloc_n(c, &b.wid_leaf, nt);
    rc |= alloc_n(c, &b.size2, ci);
    rc |= alloc_n(c, &b.lvl_hist, (size_t)MAXD * LVL_BLOCKS);
    rc |= alloc_n(c, &b.lvl_off, (size_t)MAXD * LVL_BLOCKS);
    rc |= alloc_n(c, &b.lvl_start, MAXD + 1) | alloc_n(c, &b.by_level, ci);
    b.n_rec = (uint32_t)(ci + nt + n_inst + 8);
    // the group walk packs a child range as first | (cnt - 1) << 29
    if (!rc && precision == POTG_MIXED_GROUP && b.n_rec >= (1u << 29)) {
        potg_destroy(c);
        return set_err("potg_create: POTG_MIXED_GROUP holds at most %u tree records "
                       "(about 178M particles)", (1u << 29) - 1),
               POTG_ERR_ARG;
    }

    rc |= alloc_n(c, &b.rec64, b.n_rec);
    if (precision != POTG_FP64) rc |= alloc_n(c, &b.rec32, b.n_rec);
    bool any_dip = false;
    for (uint32_t i = 0; i < n_inst; i++) any_dip |= configs[i].dipoles != 0;
    if (any_dip) {
        rc |= alloc_n(c, &b.dip64, 3 * (size_t)b.n_rec);
        if (precision != POTG_FP64) rc |= alloc_n(c, &b.dip32, b.n_rec);
    }
    const size_t tc = nt + (size_t)WAVE * n_inst;
    b.t_cap = (uint32_t)tc;
    rc |= alloc_n(c, &b.flags, nt) | alloc_n(c, &b.tpos, nt) | alloc_n(c, &b.targets, tc);
    rc |= alloc_n(c, &b.t_first, n_inst) | alloc_n(c, &b.t_base, n_inst);
    rc |= alloc_n(c, &b.queue, nt);
    rc |= alloc_n(c, &b.E, 3 * tc) | alloc_n(c, &b.inv, n);
    b.integrator = configs[0].integrator;
    if (b.integrator != POTG_BORIS) rc |= alloc_n(c, &b.a1, 3 * tc) | alloc_n(c, &b.x0, 3 * tc);
    bool any_bounce = false;
    for (uint32_t i = 0; i < n_inst; i++) any_bounce |= configs[i].out_bounce != 0;
    if (any_bounce) rc |= alloc_n(c, &b.qprev, 3 * nt);
    rc |= alloc_n(c, &b.phi, n) | alloc_n(c, &b.epe_part, 2 * (size_t)EPE_BLOCKS * n_inst);
    b.q_cap = (uint32_t)(n < 65536 ? 65536 : n);
    rc |= alloc_n(c, &b.q_pts, 3 * (size_t)b.q_cap) | alloc_n(c, &b.q_inst, b.q_cap);
    rc |= alloc_n(c, &b.q_self, b.q_cap) | alloc_n(c, &b.q_out, 3 * (size_t)b.q_cap);
    rc |= alloc_n(c, &b.q_counts, 3 * (size_t)b.q_cap);
    if (rc) {
        potg_destroy(c);
        return POTG_ERR_NOMEM;
    }
    // zeroed tree arrays: an entry the build has not written yet is an empty
    // cell or index 0, never an address outside the buffers
    bool clear_ok = true;
    for (size_t k = 0; k < c->allocs.size(); k++)
        if (hipMemset(c->allocs[k].first, 0, c->allocs[k].second) != hipSuccess) clear_ok = false;
    if (!clear_ok) {
        set_err("initial clear failed");
        potg_destroy(c);
        return POTG_ERR_HIP;
    }
    c->uploaded.assign(n_inst, 0);
    Globals g0;
    memset(&g0, 0, sizeof(g0));
    std::vector<InstState> s0(n_inst);
    memset(s0.data(), 0, sizeof(InstState) * n_inst);
    if (hipMemcpy(b.g, &g0, sizeof(g0), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.S, s0.data(), sizeof(InstState) * n_inst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.P, c->hp.data(), sizeof(InstParams) * n_inst, hipMemcpyHostToDevice) !=
            hipSuccess) {
        set_err("initial copies failed");
        potg_destroy(c);
        return POTG_ERR_HIP;
    }
    launch_init_slots(b, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_err("slot initialisation failed");
        potg_destroy(c);
        return POTG_ERR_HIP;
    }
    *out = c;
    return POTG_OK;
}

int potg_destroy(potg_ctx* c)
{
    if (!c) return POTG_OK;
    (void)hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    if (c->h_aos) (void)hipHostFree(c->h_aos);
    if (c->d_aos) (void)hipFree(c->d_aos);
    for (int k = 0; k < 6; k++)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    for (auto& e : c->tev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return POTG_OK;
}

int potg_check_particles(const potg_config* cfg, const potg_particle* parts, uint64_t* bad_index)
{
    if (!cfg || !parts) return set_err("potg_check_particles: bad arguments"), POTG_ERR_ARG;
    if (bad_index) *bad_index = 0;
    const uint64_t N = cfg->N;
    if (N < 1) return set_err("potg_check_particles: bad N"), POTG_ERR_ARG;
    if (parts[0].species != POTG_DUST) return set_err("particle 0 must be the grain"), POTG_ERR_ARG;
    const double x2 = 2.0 * cfg->r * cfg->mpp;  // the root cell's upper faces (pot.c:1476-1478)
    for (uint64_t i = 0; i < N; i++) {
        const potg_particle& q = parts[i];
        if (i && q.species != POTG_ION && q.species != POTG_ELECTRON) {
            if (bad_index) *bad_index = i;
            return set_err("particle %llu: species must be ion or electron",
                           (unsigned long long)i),
                   POTG_ERR_ARG;
        }
        // build_tree's test (pot.c:1503-1514), written so that a NaN
        // coordinate fails it too (is_particle_position_nan, pot.c:1434-1446)
        if (!(q.x >= 0.0 && q.x <= x2) || !(q.y >= 0.0 && q.y <= x2) ||
            !(q.z >= 0.0 && q.z <= x2)) {
            if (bad_index) *bad_index = i;
            return set_err("particle %llu at (%g, %g, %g) with velocity (%g, %g, %g) not in "
                           "simulation area.",
                           (unsigned long long)i, q.x, q.y, q.z, q.vx, q.vy, q.vz),
                   POTG_ERR_OUT_OF_BOX;
        }
    }
    return POTG_OK;
}

int potg_upload(potg_ctx* c, uint32_t inst, const potg_particle* parts, const potg_grain* grain,
                uint64_t t_completed)
{
    if (!c || !parts || !grain || inst >= c->b.n_inst)
        return set_err("potg_upload: bad arguments"), POTG_ERR_ARG;
    const potg_config& k = c->cfg[inst];
    const size_t N = k.N;
    {
        const int rc = potg_check_particles(&k, parts, nullptr);
        if (rc) return rc;
    }
    InstParams& p = c->hp[inst];
    for (int s = 0; s < 3; s++) p.first_idx[s] = NONE;
    for (size_t i = 1; i < N; i++) {
        const int s = parts[i].species;
        if (p.first_idx[s] == NONE) p.first_idx[s] = (uint32_t)i;
    }
    p.grain[0] = parts[0].x;
    p.grain[1] = parts[0].y;
    p.grain[2] = parts[0].z;
    chi_point(p);
    int rc = ensure_staging(c, N * sizeof(potg_particle));
    if (rc) return rc;
    destroy_graphs(c);
    memcpy(c->h_aos, parts, N * sizeof(potg_particle));
    HIPCHK(hipMemcpyAsync(c->d_aos, c->h_aos, N * sizeof(potg_particle), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(&c->b.P[inst], &p, sizeof(InstParams), hipMemcpyHostToDevice, c->stream));
    // reference layout first, so the instance's slots are [offset, offset + N)
    launch_unsort(c->b, c->stream);
    launch_scatter_upload(c->b, inst, (const double*)c->d_aos, (uint32_t)N, c->stream);
    launch_build_inverse(c->b, c->stream);
    InstState s;
    memset(&s, 0, sizeof(s));
    s.grain_q = grain->q;
    s.tot_e = grain->tot_absor_e;
    s.grain_m = grain->m;
    s.grain_v[0] = parts[0].vx;
    s.grain_v[1] = parts[0].vy;
    s.grain_v[2] = parts[0].vz;
    s.chi_b = grain->chi_b;
    for (int q = 0; q < 3; q++) s.epe[q] = grain->epe[q];
    HIPCHK(hipMemcpyAsync(&c->b.S[inst], &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(&c->b.g->t, &t_completed, sizeof(uint64_t), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->uploaded[inst] = 1;
    return POTG_OK;
}

int potg_download(potg_ctx* c, uint32_t inst, potg_particle* parts, potg_grain* grain)
{
    if (!c || inst >= c->b.n_inst) return set_err("potg_download: bad arguments"), POTG_ERR_ARG;
    const size_t N = c->cfg[inst].N;
    InstState s;
    HIPCHK(hipMemcpyAsync(&s, &c->b.S[inst], sizeof(s), hipMemcpyDeviceToHost, c->stream));
    if (parts) {
        int rc = ensure_staging(c, N * sizeof(potg_particle));
        if (rc) return rc;
        launch_gather_download(c->b, inst, (double*)c->d_aos, (uint32_t)N, c->stream);
        HIPCHK(hipMemcpyAsync(c->h_aos, c->d_aos, N * sizeof(potg_particle), hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (parts) {
        memcpy(parts, c->h_aos, N * sizeof(potg_particle));
        // the grain: fixed position, velocity from the absorbed momentum
        parts[0].vx = s.grain_v[0];
        parts[0].vy = s.grain_v[1];
        parts[0].vz = s.grain_v[2];
        parts[0].species = POTG_DUST;
    }
    if (grain) {
        grain->q = s.grain_q;
        grain->tot_absor_e = s.tot_e;
        grain->m = s.grain_m;
        for (int q = 0; q < 3; q++) grain->v[q] = s.grain_v[q];
        grain->chi_b = s.chi_b;
        for (int q = 0; q < 3; q++) grain->epe[q] = s.epe[q];
    }
    return POTG_OK;
}

int potg_step(potg_ctx* c, int32_t compute_epe, potg_step_out* out)
{
    if (!c) return set_err("potg_step: null context"), POTG_ERR_ARG;
    if (all_uploaded(c)) return POTG_ERR_ARG;
    enqueue_step(c, compute_epe != 0);
    c->parity ^= 1u;
    HIPCHK(hipGetLastError());
    if (!out) return POTG_OK;  // asynchronous: errors surface at potg_sync
    int rc = kernel_error(c);
    if (rc) return rc;
    std::vector<InstState> s(c->b.n_inst);
    HIPCHK(hipMemcpy(s.data(), c->b.S, sizeof(InstState) * c->b.n_inst, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < c->b.n_inst; i++) {
        memset(&out[i], 0, sizeof(potg_step_out));
        out[i].abs_q = s[i].out_q;
        out[i].abs_e = s[i].out_e;
        out[i].abs_i = s[i].out_i;
        out[i].n_pushed = s[i].out_pushed;
        out[i].n_reinj_abs = s[i].out_reinj_abs;
        out[i].n_reinj_out = s[i].out_reinj_out;
        out[i].n_wall_inward = s[i].out_wall_inward;
    }
    return POTG_OK;
}

int potg_sync(potg_ctx* c)
{
    if (!c) return set_err("potg_sync: null context"), POTG_ERR_ARG;
    HIPCHK(hipGetLastError());
    return kernel_error(c);
}

int potg_run(potg_ctx* c, uint64_t steps)
{
    if (!c) return set_err("potg_run: null context"), POTG_ERR_ARG;
    if (all_uploaded(c)) return POTG_ERR_ARG;
    if (!c->graph_ready) {
        // two graphs: the particle buffers swap roles every step
        const int timing = c->timing;
        c->timing = 0;
        for (int k = 0; k < 2; k++) {
            hipGraph_t gr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(c, 0);
            HIPCHK(hipStreamEndCapture(c->stream, &gr));
            HIPCHK(hipGraphInstantiate(&c->graph[k], gr, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(gr));
        }
        c->timing = timing;
        c->graph_ready = 1;
        c->parity = 0;
    }
    for (uint64_t s = 0; s < steps; s++) {
        HIPCHK(hipGraphLaunch(c->graph[c->parity], c->stream));
        c->parity ^= 1u;
        // host copy of the buffer roles follows the device
        PState t = c->b.st;
        c->b.st = c->b.alt;
        c->b.alt = t;
    }
    // the capture itself swapped the roles twice, so `st` is consistent
    return kernel_error(c);
}

int potg_build_tree(potg_ctx* c)
{
    if (!c) return set_err("potg_build_tree: null context"), POTG_ERR_ARG;
    if (all_uploaded(c)) return POTG_ERR_ARG;
    destroy_graphs(c);
    launch_keys_sort_permute(c->b, c->stream);
    launch_tree(c->b, c->stream);
    HIPCHK(hipGetLastError());
    return kernel_error(c);
}

}  // extern "C"

namespace {
// Visit instance inst's tree in canonical order (depth-first pre-order,
// children in the reference order): visit(n, walk id, record, depth).
template <class F>
int canon_walk(potg_ctx* c, uint32_t inst, int64_t cap, int64_t* n_out, const char* who, F visit)
{
    if (!c || inst >= c->b.n_inst || !n_out)
        return set_err("%s: bad arguments", who), POTG_ERR_ARG;
    Globals g;
    HIPCHK(hipMemcpy(&g, c->b.g, sizeof(g), hipMemcpyDeviceToHost));
    const size_t nrec = c->b.n_inst + g.n_children;
    if (nrec > c->b.n_rec) return set_err("%s: no tree", who), POTG_ERR_ARG;
    std::vector<NodeRec64> rec(nrec);
    HIPCHK(hipMemcpy(rec.data(), c->b.rec64, nrec * sizeof(NodeRec64), hipMemcpyDeviceToHost));
    std::vector<std::pair<uint32_t, int>> stk;
    int64_t n = 0;
    *n_out = 0;
    if (c->hp[inst].tree_n == 0) return POTG_OK;
    stk.push_back({inst, 0});
    while (!stk.empty()) {
        const uint32_t w = stk.back().first;
        const int d = stk.back().second;
        stk.pop_back();
        if (n >= cap) return set_err("%s: capacity", who), POTG_ERR_ARG;
        if (w >= nrec) return set_err("%s: bad record", who), POTG_ERR_ARG;
        const NodeRec64& r = rec[w];
        visit(n, w, r, d);
        n++;
        if (r.cnt != LEAF)
            for (int k = (int)r.cnt - 1; k >= 0; k--) stk.push_back({r.first + (uint32_t)k, d + 1});
    }
    *n_out = n;
    return POTG_OK;
}
}  // namespace

extern "C" {

int potg_tree_canon(potg_ctx* c, uint32_t inst, int32_t* depth, int32_t* is_leaf, uint32_t* part,
                    int32_t* q, uint32_t* aq, double* cc, double* size2, int32_t* nchild,
                    int64_t cap, int64_t* n_out)
{
    if (!c) return set_err("potg_tree_canon: null context"), POTG_ERR_ARG;
    std::vector<uint32_t> ref(c->b.n_total);
    HIPCHK(hipMemcpy(ref.data(), c->b.st.ref, ref.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t off = inst < c->b.n_inst ? c->hp[inst].offset : 0;
    return canon_walk(c, inst, cap, n_out, "potg_tree_canon",
                      [&](int64_t n, uint32_t, const NodeRec64& r, int d) {
                          const bool leaf = r.cnt == LEAF;
                          depth[n] = d;
                          is_leaf[n] = leaf;
                          part[n] = leaf ? ref[r.part] - off : 0xFFFFFFFFu;
                          q[n] = r.q;
                          aq[n] = leaf ? (uint32_t)abs(r.q) : r.part;
                          cc[3 * n] = r.cx;
                          cc[3 * n + 1] = r.cy;
                          cc[3 * n + 2] = r.cz;
                          size2[n] = leaf ? 0.0 : r.size2;
                          nchild[n] = leaf ? 0 : (int32_t)r.cnt;
                      });
}

int potg_tree_canon_dipoles(potg_ctx* c, uint32_t inst, double* p, int64_t cap, int64_t* n_out)
{
    if (!c || !p) return set_err("potg_tree_canon_dipoles: bad arguments"), POTG_ERR_ARG;
    std::vector<double> dip;
    if (c->b.dip64) {
        dip.resize(3 * (size_t)c->b.n_rec);
        HIPCHK(hipMemcpy(dip.data(), c->b.dip64, dip.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
    }
    return canon_walk(c, inst, cap, n_out, "potg_tree_canon_dipoles",
                      [&](int64_t n, uint32_t w, const NodeRec64&, int) {
                          for (int k = 0; k < 3; k++)
                              p[3 * n + k] = dip.empty() ? 0.0 : dip[3 * (size_t)w + k];
                      });
}

static int run_points(potg_ctx* c, uint32_t inst, const double* xyz, const uint32_t* exclude,
                      size_t n, int kind, double* out, uint32_t* counts)
{
    Bufs& b = c->b;
    std::vector<uint32_t> vi, vs;
    std::vector<uint32_t> inv(b.n_total);
    if (kind == 0 && exclude) {
        launch_build_inverse(b, c->stream);
        HIPCHK(hipMemcpyAsync(inv.data(), b.inv, inv.size() * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    const uint32_t off = c->hp[inst].offset;
    for (size_t s0 = 0; s0 < n; s0 += b.q_cap) {
        const uint32_t m = (uint32_t)((n - s0) < b.q_cap ? (n - s0) : b.q_cap);
        vi.assign(m, inst);
        vs.assign(m, NONE);
        if (kind == 0 && exclude)
            for (uint32_t k = 0; k < m; k++) {
                const uint32_t e = exclude[s0 + k];
                vs[k] = (e == 0xFFFFFFFFu || e >= c->cfg[inst].N) ? NONE : inv[off + e];
            }
        HIPCHK(hipMemcpyAsync(b.q_pts, xyz + 3 * s0, 3 * sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b.q_inst, vi.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b.q_self, vs.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, c->stream));
        launch_points(b, m, kind, c->stream);
        const size_t per = kind == 0 ? 3 : 1;
        HIPCHK(hipMemcpyAsync(out + per * s0, b.q_out, per * sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        if (counts && kind == 0)
            HIPCHK(hipMemcpyAsync(counts + 3 * s0, b.q_counts, 3 * sizeof(uint32_t) * m,
                                  hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return kernel_error(c);
}

static uint64_t spread3(uint64_t v)  // 21 bits to every third of 63
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

int potg_field_at(potg_ctx* c, uint32_t inst, const double* xyz, const uint32_t* exclude, size_t n,
                  double* E, uint32_t* counts)
{
    if (!c || inst >= c->b.n_inst || (n && (!xyz || !E)))
        return set_err("potg_field_at: bad arguments"), POTG_ERR_ARG;
    if (c->b.precision != POTG_MIXED_GROUP || n < 2)
        return run_points(c, inst, xyz, exclude, n, 0, E, counts);
    // the group criterion wants the 64 points of a wavefront close together:
    // walk the points in Morton order of their positions, then put the
    // results back in the caller's order
    const double scale = (double)(1u << 21) / c->hp[inst].root_x2;
    std::vector<std::pair<uint64_t, size_t>> ord(n);
    for (size_t k = 0; k < n; k++) {
        uint64_t u[3];
        for (int d = 0; d < 3; d++) {
            double f = xyz[3 * k + d] * scale;
            f = f < 0.0 ? 0.0 : (f > 2097151.0 ? 2097151.0 : f);
            u[d] = (uint64_t)f;
        }
        ord[k] = {spread3(u[0]) << 2 | spread3(u[1]) << 1 | spread3(u[2]), k};
    }
    std::sort(ord.begin(), ord.end());
    std::vector<double> pxyz(3 * n), pE(3 * n);
    std::vector<uint32_t> pex(exclude ? n : 0), pcnt(counts ? 3 * n : 0);
    for (size_t k = 0; k < n; k++) {
        for (int d = 0; d < 3; d++) pxyz[3 * k + d] = xyz[3 * ord[k].second + d];
        if (exclude) pex[k] = exclude[ord[k].second];
    }
    const int rc = run_points(c, inst, pxyz.data(), exclude ? pex.data() : nullptr, n, 0, pE.data(),
                              counts ? pcnt.data() : nullptr);
    if (rc) return rc;
    for (size_t k = 0; k < n; k++)
        for (int d = 0; d < 3; d++) {
            E[3 * ord[k].second + d] = pE[3 * k + d];
            if (counts) counts[3 * ord[k].second + d] = pcnt[3 * k + d];
        }
    return POTG_OK;
}

int potg_phi_at(potg_ctx* c, uint32_t inst, const double* xyz, size_t n, double* phi)
{
    if (!c || inst >= c->b.n_inst || (n && (!xyz || !phi)))
        return set_err("potg_phi_at: bad arguments"), POTG_ERR_ARG;
    return run_points(c, inst, xyz, nullptr, n, 1, phi, nullptr);
}

int potg_phi_at_noalloc(potg_ctx* c, uint32_t inst, const double* d_xyz, size_t n, double* d_phi,
                        void* stream)
{
    if (!c || inst >= c->b.n_inst || (n && (!d_xyz || !d_phi)))
        return set_err("potg_phi_at_noalloc: bad arguments"), POTG_ERR_ARG;
    if (n > POTG_NOALLOC_MAX_POINTS)
        return set_err("potg_phi_at_noalloc: %zu points, more than the %u its scratch holds", n,
                       POTG_NOALLOC_MAX_POINTS),
               POTG_ERR_ARG;
    if (!c->b.noalloc) return set_err("potg_phi_at_noalloc: no scratch on this device"), POTG_ERR_HIP;
    if (n == 0) return POTG_OK;
    launch_phi_noalloc(c->b, c->b.noalloc, inst, d_xyz, (uint32_t)n, d_phi,
                       stream ? (hipStream_t)stream : c->stream);
    HIPCHK(hipGetLastError());
    return POTG_OK;
}

int potg_direct_sum(potg_ctx* c, uint32_t inst, const uint32_t* idx, size_t n, double* E)
{
    if (!c || inst >= c->b.n_inst || (n && (!idx || !E)))
        return set_err("potg_direct_sum: bad arguments"), POTG_ERR_ARG;
    Bufs& b = c->b;
    std::vector<uint32_t> inv(b.n_total), vs;
    launch_build_inverse(b, c->stream);
    HIPCHK(hipMemcpyAsync(inv.data(), b.inv, inv.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t off = c->hp[inst].offset;
    for (size_t s0 = 0; s0 < n; s0 += b.q_cap) {
        const uint32_t m = (uint32_t)((n - s0) < b.q_cap ? (n - s0) : b.q_cap);
        vs.resize(m);
        for (uint32_t k = 0; k < m; k++) {
            if (idx[s0 + k] >= c->cfg[inst].N) return set_err("potg_direct_sum: index out of range"), POTG_ERR_ARG;
            vs[k] = inv[off + idx[s0 + k]];
        }
        HIPCHK(hipMemcpyAsync(b.q_self, vs.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, c->stream));
        launch_direct_sum(b, inst, m, c->stream);
        HIPCHK(hipMemcpyAsync(E + 3 * s0, b.q_out, 3 * sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return POTG_OK;
}

int potg_stats(potg_ctx* c, potg_stats_t* out)
{
    if (!c || !out) return set_err("potg_stats: bad arguments"), POTG_ERR_ARG;
    Globals g;
    HIPCHK(hipMemcpy(&g, c->b.g, sizeof(g), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->steps = g.n_steps;
    out->n_targets = g.n_tgt_total;
    out->n_cell = g.n_cell;
    out->n_leaf = g.n_leaf;
    out->n_open = g.n_open;
    out->n_nodes = (uint64_t)g.n_internal + c->b.n_tree;
    out->n_particles = c->b.n_total;
    out->max_depth = g.max_depth;
    out->n_range_loads = g.n_pops;
    out->n_child_visits = g.n_visits;
    return POTG_OK;
}

int potg_set_counting(potg_ctx* c, int32_t on)
{
    if (!c) return set_err("potg_set_counting: null context"), POTG_ERR_ARG;
    if ((on != 0) != (c->b.count != 0)) destroy_graphs(c);  // captured with the old kernels
    c->b.count = on != 0;
    return POTG_OK;
}

int potg_set_retire(potg_ctx* c, int32_t on)
{
    if (!c) return set_err("potg_set_retire: null context"), POTG_ERR_ARG;
    const uint32_t v = on != 0;
    // a device flag: captured steps read it, so they stay valid
    HIPCHK(hipMemcpyAsync(&c->b.g->retire, &v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return POTG_OK;
}

int potg_instance_retired(potg_ctx* c, uint32_t inst, int32_t* retired, uint64_t* step)
{
    if (!c || inst >= c->b.n_inst || !retired)
        return set_err("potg_instance_retired: bad arguments"), POTG_ERR_ARG;
    InstState s;
    HIPCHK(hipMemcpyAsync(&s, &c->b.S[inst], sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *retired = s.retired ? 1 : 0;
    if (step) *step = s.retired ? (uint64_t)s.retired_t : 0;
    return POTG_OK;
}

int potg_stats_reset(potg_ctx* c)
{
    if (!c) return set_err("potg_stats_reset: null context"), POTG_ERR_ARG;
    Globals g;
    HIPCHK(hipMemcpy(&g, c->b.g, sizeof(g), hipMemcpyDeviceToHost));
    g.n_cell = g.n_leaf = g.n_open = g.n_tgt_total = g.n_steps = g.n_pops = g.n_visits = 0;
    g.max_depth = 0;
    HIPCHK(hipMemcpy(c->b.g, &g, sizeof(g), hipMemcpyHostToDevice));
    return POTG_OK;
}

int potg_set_tuning(potg_ctx* c, const potg_tuning* set, potg_tuning* get)
{
    if (!c) return set_err("potg_set_tuning: null context"), POTG_ERR_ARG;
    if (set) {
        if (set->walk_blocks == 0 || set->walk_blocks > 65536 ||
            (set->group != 8 && set->group != 16 && set->group != 64) || set->variant > 1 ||
            (set->sort_tile != 0 && set->sort_tile != 2048 && set->sort_tile != 4096))
            return set_err("potg_set_tuning: walk_blocks must be 1..65536, group 8, 16 or 64, "
                           "variant 0 or 1, sort_tile 2048 or 4096 (0: unchanged)"),
                   POTG_ERR_ARG;
        c->b.walk_blocks = set->walk_blocks;
        c->b.group = set->group;
        c->b.variant = set->variant;
        if (set->sort_tile) c->b.sort_ipt = set->sort_tile / 256;
        destroy_graphs(c);  // captured steps hold the old launch parameters
    }
    if (get) {
        get->walk_blocks = c->b.walk_blocks;
        get->group = c->b.group;
        get->variant = c->b.variant;
        get->sort_tile = 256 * c->b.sort_ipt;
        get->walk_ms = c->tuned_ms;
    }
    return POTG_OK;
}

constexpr double TUNE_MARGIN = 1.01;

int potg_tune(potg_ctx* c, potg_tuning* out)
{
    if (!c) return set_err("potg_tune: null context"), POTG_ERR_ARG;
    if (all_uploaded(c)) return POTG_ERR_ARG;
    Bufs& b = c->b;
    hipStream_t s = c->stream;
    int rc = kernel_error(c);
    if (rc) return rc;
    Globals g0;
    HIPCHK(hipMemcpy(&g0, b.g, sizeof(g0), hipMemcpyDeviceToHost));
    // everything the coming step does before its walk
    launch_begin_step(b, s);
    launch_keys_sort_permute(b, s);
    launch_tree(b, s);
    launch_targets(b, s);
    HIPCHK(hipStreamSynchronize(s));
    rc = kernel_error(c);
    if (rc) return rc;
    const uint32_t cus = (uint32_t)device_cus();
    // candidates: targets per wavefront and visit order (float32 walk only:
    // 8, 16, or 64 in either order) x blocks per CU
    const uint32_t shapes[][2] = {{8, 0}, {16, 0}, {64, 0}, {64, 1}};
    const uint32_t bpc[] = {2, 4, 8, 16};
    const uint32_t keep_wb = b.walk_blocks, keep_gr = b.group, keep_var = b.variant;
    double best = 1e30, keep_ms = -1.0;
    uint32_t best_wb = keep_wb, best_gr = keep_gr, best_var = keep_var;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    for (uint32_t wb : bpc) {
        for (const auto& sh : shapes) {
            const uint32_t gr = sh[0];
            // the float64 and group-criterion walks have one shape: only the
            // grid varies
            if (b.precision != POTG_MIXED && (gr != 64 || sh[1] != 0)) continue;
            b.group = gr;
            b.variant = sh[1];
            b.walk_blocks = wb * cus;
            float ms_best = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                launch_walk_init(b, s);
                launch_walk_reset(b, s);
                HIPCHK(hipEventRecord(e0, s));
                launch_walk(b, s);
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < ms_best) ms_best = ms;
            }
            if (b.walk_blocks == keep_wb && gr == keep_gr && sh[1] == keep_var) keep_ms = ms_best;
            if (ms_best < best) {
                best = ms_best;
                best_wb = b.walk_blocks;
                best_gr = b.group;
                best_var = b.variant;
            }
        }
    }
    // the shape in use stays unless another is faster by more than the
    // timing noise (1%): equivalent shapes do not flap from run to run
    if (keep_ms >= 0.0 && keep_ms <= TUNE_MARGIN * best) {
        best = keep_ms;
        best_wb = keep_wb;
        best_gr = keep_gr;
        best_var = keep_var;
    }
    // the radix sort's tile: the keys, sort and permute of the coming step
    // again under each tile size (the state is sorted already, so the order
    // stays as it is)
    const uint32_t keep_ipt = b.sort_ipt;
    uint32_t best_ipt = keep_ipt;
    if (!rc) {
        float sbest = 1e30f, skeep = -1.f;
        for (uint32_t ipt : {8u, 16u}) {
            b.sort_ipt = ipt;
            float ms_best = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                HIPCHK(hipEventRecord(e0, s));
                launch_keys_sort_permute(b, s);
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < ms_best) ms_best = ms;
            }
            if (ipt == keep_ipt) skeep = ms_best;
            if (ms_best < sbest) {
                sbest = ms_best;
                best_ipt = ipt;
            }
        }
        if (skeep >= 0.f && skeep <= TUNE_MARGIN * sbest) best_ipt = keep_ipt;
        rc = kernel_error(c);
    }
    b.sort_ipt = rc ? keep_ipt : best_ipt;
    rc = kernel_error(c);
    // the step was not taken: step counter and work counters as before
    Globals g1;
    HIPCHK(hipMemcpy(&g1, b.g, sizeof(g1), hipMemcpyDeviceToHost));
    g1.t = g0.t;
    g1.n_cell = g0.n_cell;
    g1.n_leaf = g0.n_leaf;
    g1.n_open = g0.n_open;
    g1.n_pops = g0.n_pops;
    g1.n_visits = g0.n_visits;
    g1.n_tgt_total = g0.n_tgt_total;
    g1.n_steps = g0.n_steps;
    HIPCHK(hipMemcpy(b.g, &g1, sizeof(g1), hipMemcpyHostToDevice));
    b.walk_blocks = rc ? keep_wb : best_wb;
    b.group = rc ? keep_gr : best_gr;
    b.variant = rc ? keep_var : best_var;
    c->tuned_ms = rc ? 0.0 : best;
    destroy_graphs(c);
    if (rc) return rc;
    return potg_set_tuning(c, nullptr, out);
}

int potg_stage_times(potg_ctx* c, int32_t enable, double* ms5)
{
    if (!c) return set_err("potg_stage_times: null context"), POTG_ERR_ARG;
    int rc = read_stage_times(c);
    if (rc) return rc;
    c->timing = enable ? 1 : 0;
    if (ms5)
        for (int k = 0; k < 5; k++) ms5[k] = c->stage_ms[k];
    return POTG_OK;
}

}  // extern "C"
