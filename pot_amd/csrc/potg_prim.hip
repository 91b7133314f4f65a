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
//
// 8-bit digits, three launches per pass: the digit counts of every tile of
// 256 x RS_IPT elements (k_rs_count), a scan of each digit's counts over the
// tiles (k_rs_scan, one block per digit), and k_rs_scatter, which ranks its
// tile stably by digit with wave ballots, stages it in LDS in digit order
// and then stores every digit's run to its global place, so that
// neighbouring lanes store to neighbouring addresses.  The tile is 2048 or
// 4096 elements (RS_IPT 8 or 16, picked by potg_tune): a smaller tile gives
// more blocks, and so fills the device, when the keys are few.

constexpr int RS_WAVES = BT / WAVE;

inline uint32_t rs_blocks(uint32_t n, int ipt) { return (n + BT * ipt - 1) / (BT * ipt); }

template <int RS_IPT>
__global__ __launch_bounds__(BT) void k_rs_count(const uint64_t* __restrict__ keys, uint32_t n,
                                                  int shift, uint32_t nb,
                                                  uint32_t* __restrict__ cnt)
{
    constexpr int RS_TILE = BT * RS_IPT;
    __shared__ uint32_t h[RS_WAVES][256];
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) h[k][threadIdx.x] = 0;
    __syncthreads();
    const int w = threadIdx.x / WAVE;
    const uint32_t base = blockIdx.x * RS_TILE;
    uint32_t d[RS_IPT];
#pragma unroll
    for (int k = 0; k < RS_IPT; k++) {
        const uint32_t i = base + k * BT + threadIdx.x;
        d[k] = i < n ? (uint32_t)(keys[i] >> shift) & 255u : 256u;
    }
#pragma unroll
    for (int k = 0; k < RS_IPT; k++)
        if (d[k] < 256u) atomicAdd(&h[w][d[k]], 1u);
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) c += h[k][threadIdx.x];
    cnt[threadIdx.x * nb + blockIdx.x] = c;
}

// in place: cnt[d][b] becomes the count of digit d in the tiles before b;
// tot[d] = the count of digit d over all tiles
__global__ __launch_bounds__(BT) void k_rs_scan(uint32_t* __restrict__ cnt, uint32_t nb,
                                                 uint32_t* __restrict__ tot)
{
    __shared__ uint32_t sh_w[BT / WAVE];
    uint32_t* c = cnt + (size_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += BT * 4) {
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = b0 + threadIdx.x * 4 + k;
            v[k] = i < nb ? c[i] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t off = carry + block_excl_scan(sum, sh_w, &total);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = b0 + threadIdx.x * 4 + k;
            if (i < nb) c[i] = off;
            off += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <int RS_IPT>
__global__ __launch_bounds__(BT) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                    const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout, uint32_t n,
                                                    int shift, uint32_t nb,
                                                    const uint32_t* __restrict__ cnt,
                                                    const uint32_t* __restrict__ tot)
{
    constexpr int RS_TILE = BT * RS_IPT;
    constexpr int RS_WTILE = RS_TILE / RS_WAVES;  // contiguous elements per wavefront
    __shared__ uint64_t skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    __shared__ uint32_t wc[RS_WAVES][256];  // running digit counts per wavefront
    __shared__ uint32_t gdelta[256];        // global index - LDS index, per digit
    __shared__ uint32_t sh_w[BT / WAVE];

    const int w = threadIdx.x / WAVE;
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) wc[k][threadIdx.x] = 0;
    // where this tile's run of digit threadIdx.x starts in the output (the
    // scan's barriers also publish the zeroed counts)
    const uint32_t gstart =
        block_excl_scan(tot[threadIdx.x], sh_w, nullptr) + cnt[threadIdx.x * nb + blockIdx.x];

    // wavefront w ranks the elements [base, base + RS_WTILE) in rounds of 64
    const uint32_t base = blockIdx.x * RS_TILE + w * RS_WTILE;
    uint64_t key[RS_IPT];
    uint32_t val[RS_IPT], rnk[RS_IPT];
#pragma unroll
    for (int r = 0; r < RS_IPT; r++) {
        const uint32_t i = base + r * WAVE + lane;
        key[r] = i < n ? kin[i] : 0ull;
        val[r] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_IPT; r++) {
        const bool valid = base + r * WAVE + lane < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t below = __popcll(m & lanemask_lt());
        // every lane reads the count before the lowest lane of its digit
        // adds the round's run to it (one wavefront: LDS accesses in order)
        const uint32_t run = wc[w][d];
        rnk[r] = run + below;
        if (valid && below == 0) wc[w][d] = run + __popcll(m);
    }
    __syncthreads();
    {
        // per digit: start of each wavefront's run in the tile, and of the
        // tile's run in LDS (digit order)
        const uint32_t d = threadIdx.x;
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < RS_WAVES; k++) {
            const uint32_t c = wc[k][d];
            wc[k][d] = sum;
            sum += c;
        }
        const uint32_t lstart = block_excl_scan(sum, sh_w, nullptr);
#pragma unroll
        for (int k = 0; k < RS_WAVES; k++) wc[k][d] += lstart;
        gdelta[d] = gstart - lstart;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_IPT; r++) {
        if (base + r * WAVE + lane < n) {
            const uint32_t p = wc[w][(uint32_t)(key[r] >> shift) & 255u] + rnk[r];
            skey[p] = key[r];
            sval[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t t0 = blockIdx.x * RS_TILE;
    const uint32_t tile_n = n - t0 < (uint32_t)RS_TILE ? n - t0 : (uint32_t)RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_IPT; r++) {
        const uint32_t p = r * BT + threadIdx.x;
        if (p < tile_n) {
            const uint64_t k = skey[p];
            const uint32_t o = p + gdelta[(uint32_t)(k >> shift) & 255u];
            kout[o] = k;
            vout[o] = sval[p];
        }
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

// sized for the smaller tile, which needs more per-tile counts
size_t radix_temp_words(uint32_t n) { return 256ull * (rs_blocks(n, 8) + 1) + 64; }

void radix_sort_pairs(uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt,
                      uint32_t n, int end_bit, uint32_t* temp, int ipt, hipStream_t s)
{
    if (n == 0) return;
    ipt = ipt == 8 ? 8 : 16;
    const uint32_t nb = rs_blocks(n, ipt);
    uint32_t* cnt = temp;
    uint32_t* tot = temp + 256ull * nb;
    for (int shift = 0; shift < end_bit; shift += 8) {
        if (ipt == 8) k_rs_count<8><<<nb, BT, 0, s>>>(keys, n, shift, nb, cnt);
        else k_rs_count<16><<<nb, BT, 0, s>>>(keys, n, shift, nb, cnt);
        k_rs_scan<<<256, BT, 0, s>>>(cnt, nb, tot);
        if (ipt == 8)
            k_rs_scatter<8><<<nb, BT, 0, s>>>(keys, vals, keys_alt, vals_alt, n, shift, nb, cnt, tot);
        else
            k_rs_scatter<16><<<nb, BT, 0, s>>>(keys, vals, keys_alt, vals_alt, n, shift, nb, cnt,
                                                tot);
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
