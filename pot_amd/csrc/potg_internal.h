// potg_internal.h -- the context's device buffers and the kernel launchers
// of the step pipeline.  Not part of the C ABI.
#pragma once

#include "../../include/potg.h"
#include "potg_dev.h"

namespace potg {

// Particle state in sorted (octree) order, structure of arrays.
struct PState {
    double *x, *y, *z, *vx, *vy, *vz;
    uint32_t* ref;     // global slot in reference order (instance offset + index)
    uint16_t* inst;
    uint8_t* spec;
};

struct Bufs {
    uint32_t n_total;  // particle slots over all instances
    uint32_t n_tree;   // leading sorted slots that belong to the trees
    uint32_t n_inst;
    int precision;     // POTG_FP64 / POTG_MIXED / POTG_MIXED_GROUP
    int any_spheroid;  // some instance has the analytic spheroid grain
    int count;         // step walks keep the interaction counters (potg_set_counting)
    int key_levels;    // digits carried by a sort key
    int key_bits;      // bits of the key that are sorted
    uint32_t walk_blocks; // blocks of the persistent walk grid
    uint32_t group;    // float32 walk: targets per wavefront (8, 16 or 64)
    uint32_t variant;  // with 64: 0 visits cells until one opens, 1 sweeps whole ranges
    uint32_t sort_ipt; // radix-sort elements per thread: tiles of 256 x (8 or 16)
    InstParams* P;
    InstState* S;
    Globals* g;
    PState st, alt;
    uint64_t *key, *key_alt;
    uint32_t *perm, *perm_alt;
    uint32_t* sort_tmp;
    uint32_t* scan_tmp;
    // tree
    int32_t* lcp;
    uint32_t* cnt;
    uint32_t* base;
    uint32_t* owner;
    int32_t* level;
    uint32_t* slots;   // 8 per internal node; NONE when empty
    uint32_t* ccount;
    uint32_t* cfirst;
    uint32_t* wid_int;
    uint32_t* wid_leaf;
    uint32_t *lvl_hist, *lvl_off;  // per-block depth histograms and their scan
    uint32_t* lvl_start;           // first slot of each depth in by_level
    uint32_t* by_level;  // internal node ids grouped by depth
    double* size2;
    NodeRec64* rec64;
    NodeRec32* rec32;
    double* dip64;     // dipole of each record, 3 per record (DIPOLES only)
    float4* dip32;     // the same in units of R for the float32 walk
    uint32_t n_rec;    // record capacity
    uint32_t cap_int;  // internal-node capacity
    // targets of the walk and push: per instance a run starting at a
    // multiple of 64 (a wavefront's tile never mixes instances), so t_cap =
    // n_tree + 64 n_inst slots; t_first / t_base: each instance's first
    // target in the unpadded count and its run's start
    uint32_t* flags;
    uint32_t* tpos;
    uint32_t* targets;
    uint32_t t_cap;
    uint32_t *t_first, *t_base;
    uint32_t* queue;   // reinjection queue
    double* E;
    // velocity Verlet / Euler-Richardson: the first half's acceleration and
    // each target's start position; OUT_BOUNCE: the start position of each
    // queued reinjection
    int integrator;    // POTG_BORIS / POTG_VERLET / POTG_EULER_RICHARDSON
    double* a1;
    double* x0;
    double* qprev;
    uint32_t* inv;     // global reference slot -> sorted slot
    // potential energies (compute_epe steps)
    double* phi;       // potential felt by each sorted slot
    double* epe_part;  // EPE_BLOCKS x {electron, ion} partial sums per instance
    // point queries
    double* q_pts;
    uint32_t* q_inst;
    uint32_t* q_self;
    double* q_out;
    uint32_t* q_counts;
    uint32_t q_cap;
    uint32_t* noalloc;  // this device's potg_phi_at_noalloc scratch
};

// step stages (kernel launches only; no synchronisation)
void launch_begin_step(Bufs& b, hipStream_t s);
void launch_keys_sort_permute(Bufs& b, hipStream_t s);
void launch_tree(Bufs& b, hipStream_t s);
void launch_targets(Bufs& b, hipStream_t s);
void launch_chi_b(Bufs& b, hipStream_t s);
void launch_walk_init(Bufs& b, hipStream_t s);  // grain term of each target
void launch_walk(Bufs& b, hipStream_t s);
void launch_walk_reset(Bufs& b, hipStream_t s);  // rewind the walk's tile counter
int device_cus();
void launch_push(Bufs& b, hipStream_t s);
// velocity Verlet / Euler-Richardson: first half of the push (state becomes
// the intermediate state), then the leaf records follow it
void launch_half(Bufs& b, hipStream_t s);
void launch_leaf_refresh(Bufs& b, hipStream_t s);
// compute_epe nonzero: fold the potential-energy sums of launch_epe
void launch_finish_step(Bufs& b, int compute_epe, hipStream_t s);
void launch_epe(Bufs& b, hipStream_t s);
// upload / download helpers
void launch_scatter_upload(Bufs& b, uint32_t inst, const double* aos, uint32_t n, hipStream_t s);
void launch_gather_download(Bufs& b, uint32_t inst, double* aos, uint32_t n, hipStream_t s);
void launch_build_inverse(Bufs& b, hipStream_t s);
// set every slot to its reference identity (at creation)
void launch_init_slots(Bufs& b, hipStream_t s);
// permute the state back to reference layout (before an upload)
void launch_unsort(Bufs& b, hipStream_t s);
// point queries: n points of q_pts with q_inst/q_self; kind 0 field, 1 potential
void launch_points(Bufs& b, uint32_t n, int kind, hipStream_t s);
void launch_direct_sum(Bufs& b, uint32_t inst, uint32_t n, hipStream_t s);
// potg_phi_at_noalloc: this device's scratch array (0 on success), and the
// point-potential walk over the caller's device arrays using it
int noalloc_scratch(uint32_t** out);
void launch_phi_noalloc(Bufs& b, uint32_t* scr, uint32_t inst, const double* pts, uint32_t n,
                        double* out, hipStream_t s);

}  // namespace potg
