// potg_prim.hip -- scan, compaction and LSD radix sort for gfx950.
//
// Written for wave64: ranks inside a wavefront come from ballots, block
// partials through LDS.  Every routine takes an optional device-resident
// element count so that a captured step never needs a host round trip.
#include "potg_prim.h"

namespace potg {

namespace {

constexpr int BT = 256;          // threads per block
constexpr int IPT = 8;           // items per thread
constexpr int TILE = BT * IPT;   // elements per block

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// inclusive scan across a wavefront
__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t o = __shfl_up(v, d, WAVE);
        if ((int)lane_id() >= d) v += o;
    }
    return v;
}

// exclusive scan of one value per thread over a 256-thread block
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* sh_w, uint32_t* total)
{
    const int w = threadIdx.x / WAVE;
    uint32_t inc = wave_incl_scan(v);
    if (lane_id() == WAVE - 1) sh_w[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < BT / WAVE; k++) {
        if (k < w) off += sh_w[k];
        tot += sh_w[k];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

__device__ inline uint32_t eff_n(uint32_t n, const uint32_t* n_dev)
{
    if (n_dev) {
        uint32_t d = *n_dev;
        return d < n ? d : n;
    }
    return n;
}

__global__ __launch_bounds__(BT) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                     const uint32_t* n_dev,
                                                     uint32_t* __restrict__ part)
{
    __shared__ uint32_t sh_w[BT / WAVE];
    const uint32_t ne = eff_n(n, n_dev);
    const uint32_t base = blockIdx.x * TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        uint32_t i = base + k * BT + threadIdx.x;
        if (i < ne) s += in[i];
    }
    uint32_t tot;
    block_excl_scan(s, sh_w, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// single block: exclusive scan of the block partials; writes the total
__global__ __launch_bounds__(1024) void k_scan_parts(uint32_t* part, uint32_t nb, uint32_t* total,
                                                      uint32_t* total2)
{
    __shared__ uint32_t sh_w[1024 / WAVE];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        uint32_t i = b0 + threadIdx.x;
        uint32_t v = i < nb ? part[i] : 0;
        const int w = threadIdx.x / WAVE;
        uint32_t inc = wave_incl_scan(v);
        if (lane_id() == WAVE - 1) sh_w[w] = inc;
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (int k = 0; k < 1024 / WAVE; k++) {
            if (k < w) off += sh_w[k];
            tot += sh_w[k];
        }
        uint32_t c = carry;
        if (i < nb) part[i] = c + off + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (total) *total = carry;
        if (total2) *total2 = carry;
    }
}

__global__ __launch_bounds__(BT) void k_scan_down(const uint32_t* __restrict__ in, uint32_t n,
                                                   const uint32_t* n_dev,
                                                   const uint32_t* __restrict__ part,
                                                   uint32_t* __restrict__ out, uint32_t add)
{
    __shared__ uint32_t sh[TILE];
    __shared__ uint32_t sh_w[BT / WAVE];
    const uint32_t ne = eff_n(n, n_dev);
    const uint32_t base = blockIdx.x * TILE;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        uint32_t i = base + k * BT + threadIdx.x;
        sh[k * BT + threadIdx.x] = (i < ne) ? in[i] : 0;
    }
    __syncthreads();
    uint32_t loc[IPT];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        loc[k] = s;
        s += sh[threadIdx.x * IPT + k];
    }
    uint32_t off = block_excl_scan(s, sh_w, nullptr) + part[blockIdx.x] + add;
#pragma unroll
    for (int k = 0; k < IPT; k++) sh[threadIdx.x * IPT + k] = off + loc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        uint32_t i = base + k * BT + threadIdx.x;
        if (i < ne) out[i] = sh[k * BT + threadIdx.x];
    }
}

// ---- radix sort -------------------------------------------------------

__global__ __launch_bounds__(BT) void k_radix_hist(const uint64_t* __restrict__ keys, uint32_t n,
                                                    int shift, uint32_t nb,
                                                    uint32_t* __restrict__ ghist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        uint32_t i = base + k * BT + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    ghist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(BT) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint64_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, uint32_t n,
                                                       int shift, uint32_t nb,
                                                       const uint32_t* __restrict__ goff)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[BT / WAVE][256];
    const int w = threadIdx.x / WAVE;
    run[threadIdx.x] = goff[threadIdx.x * nb + blockIdx.x];
    const uint32_t base = blockIdx.x * TILE;
    for (int r = 0; r < IPT; r++) {
        for (int k = 0; k < BT / WAVE; k++) wcnt[k][threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = base + r * BT + threadIdx.x;
        const bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0;
        uint32_t val = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = __popcll(m & lanemask_lt());
        if (valid && rank == 0) wcnt[w][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pre = run[d];
            for (int k = 0; k < w; k++) pre += wcnt[k][d];
            kout[pre + rank] = key;
            vout[pre + rank] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < BT / WAVE; k++) add += wcnt[k][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

__global__ void k_iota(uint32_t* v, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void k_compact(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                          uint32_t n, uint32_t* __restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = i;
}

inline uint32_t nblocks(uint32_t n) { return (n + TILE - 1) / TILE; }

}  // namespace

size_t scan_temp_words(uint32_t n) { return nblocks(n) + 16; }

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* n_dev,
                    uint32_t* total, uint32_t* total2, uint32_t* temp, uint32_t add,
                    hipStream_t s)
{
    uint32_t nb = nblocks(n);
    if (nb == 0) nb = 1;
    k_scan_reduce<<<nb, BT, 0, s>>>(in, n, n_dev, temp);
    k_scan_parts<<<1, 1024, 0, s>>>(temp, nb, total, total2);
    k_scan_down<<<nb, BT, 0, s>>>(in, n, n_dev, temp, out, add);
}

size_t radix_temp_words(uint32_t n) { return 256ull * (nblocks(n) + 1) * 2 + 64; }

void radix_sort_pairs(uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt,
                      uint32_t n, int end_bit, uint32_t* temp, hipStream_t s)
{
    const uint32_t nb = nblocks(n) ? nblocks(n) : 1;
    uint32_t* hist = temp;
    uint32_t* scan_tmp = temp + 256ull * nb;
    for (int shift = 0; shift < end_bit; shift += 8) {
        k_radix_hist<<<nb, BT, 0, s>>>(keys, n, shift, nb, hist);
        exclusive_scan(hist, hist, 256u * nb, nullptr, nullptr, nullptr, scan_tmp, 0, s);
        k_radix_scatter<<<nb, BT, 0, s>>>(keys, vals, keys_alt, vals_alt, n, shift, nb, hist);
        uint64_t* tk = keys; keys = keys_alt; keys_alt = tk;
        uint32_t* tv = vals; vals = vals_alt; vals_alt = tv;
    }
}

void iota(uint32_t* v, uint32_t n, hipStream_t s)
{
    if (n) k_iota<<<(n + 255) / 256, 256, 0, s>>>(v, n);
}

void compact_flags(const uint32_t* flags, uint32_t* pos, uint32_t n, uint32_t* out,
                   uint32_t* count, uint32_t* temp, hipStream_t s)
{
    exclusive_scan(flags, pos, n, nullptr, count, nullptr, temp, 0, s);
    if (n) k_compact<<<(n + 255) / 256, 256, 0, s>>>(flags, pos, n, out);
}

}  // namespace potg
