// Stable LSD radix sort of (key, value) pairs, 8 bits per pass, for the
// batched drop-cutter.  It records no memset and no host synchronisation, so
// a hipGraph can capture it: every tile writes its whole histogram column,
// and the scatter derives all positions from the scanned histograms.
//
// A pass over n pairs:
//   k_rs_hist    one block per tile of kRsTile pairs: the tile's digit counts
//                -> hist[digit * ntiles + tile]
//   k_rs_scan_rows  one block per digit: exclusive scan of the digit's
//                tile counts in place (tiles in input order) and its total;
//   k_rs_scan_tops  exclusive scan of the 256 digit totals
//   k_rs_scatter one block per tile, four rounds of 256 pairs in input order:
//                rank among equal digits inside a wavefront from eight
//                ballots, then across the block's wavefronts and rounds, so
//                equal keys keep their input order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ocl_amd {

constexpr uint32_t kRsBlock = 256;
constexpr uint32_t kRsRounds = 4;
constexpr uint32_t kRsTile = kRsBlock * kRsRounds;

__global__ void __launch_bounds__(256) k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                 uint32_t* __restrict__ hist, uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRsTile;
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        const uint32_t i = base + r * kRsBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of each digit's row hist[d * ntiles .. (d + 1) * ntiles)
// in place, one block of 1024 threads per digit; the row totals go to
// rowsum[d].
__global__ void __launch_bounds__(1024) k_rs_scan_rows(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                       uint32_t* __restrict__ rowsum) {
    __shared__ uint32_t part[1024];
    uint32_t* v = hist + (size_t)blockIdx.x * ntiles;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (ntiles + 1023) / 1024;
    const uint32_t b = min(t * per, ntiles), e = min(b + per, ntiles);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += v[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t a = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += a;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023) rowsum[blockIdx.x] = part[1023];
}

// exclusive scan of the 256 row totals -> rowoff (one block of 256 threads)
__global__ void __launch_bounds__(256) k_rs_scan_tops(const uint32_t* __restrict__ rowsum,
                                                      uint32_t* __restrict__ rowoff) {
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x;
    const uint32_t mine = rowsum[t];
    part[t] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t a = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += a;
        __syncthreads();
    }
    rowoff[t] = part[t] - mine;
}

__global__ void __launch_bounds__(256) k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    uint32_t n, uint32_t shift, const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ rowoff, uint32_t ntiles,
                                                    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ uint32_t wcnt[4][256];  // per-wavefront digit counts of the round
    __shared__ uint32_t run[256];      // pairs of each digit placed by earlier rounds
    const uint32_t t = threadIdx.x, wv = t >> 6, ln = t & 63u;
    run[t] = rowoff[t] + off[t * ntiles + blockIdx.x];
    const uint32_t base = blockIdx.x * kRsTile;
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        for (uint32_t w = 0; w < 4; ++w) wcnt[w][t] = 0;
        __syncthreads();
        const uint32_t i = base + r * kRsBlock + t;
        const bool act = i < n;
        const uint32_t k = act ? kin[i] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        // lanes of this wavefront with the same digit
        unsigned long long peers = __ballot(act);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long b1 = __ballot(act && ((d >> bit) & 1u));
            peers &= ((d >> bit) & 1u) ? b1 : ~b1;
        }
        const uint32_t below = __popcll(peers & ((1ull << ln) - 1ull));
        const uint32_t leader = peers ? (uint32_t)(__ffsll((long long)peers) - 1) : 64u;
        if (act && ln == leader) wcnt[wv][d] = __popcll(peers);
        __syncthreads();
        if (act) {
            uint32_t pos = run[d] + below;
            for (uint32_t w = 0; w < wv; ++w) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = vin[i];
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}

// Scratch words needed by rs_sort for n pairs: hist (256 x ntiles), two
// ping-pong buffers of n keys and n values, row totals and row offsets.
constexpr size_t rs_scratch_words(uint32_t n) {
    const uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    return (size_t)256 * ntiles + 2ull * n + 512 + 64;
}

// Sorts (keys, vals) by the low `bits` of the key; the result ends up in
// keys_out / vals_out.  keys and vals are clobbered.  Launches only.
inline void rs_sort(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits, uint32_t* scratch,
                    uint32_t** keys_out, uint32_t** vals_out, hipStream_t st) {
    const uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    uint32_t* hist = scratch;
    uint32_t* k2 = scratch + (size_t)256 * ntiles;
    uint32_t* v2 = k2 + n;
    uint32_t* rowsum = v2 + n;
    uint32_t* rowoff = rowsum + 256;
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        k_rs_hist<<<ntiles, 256, 0, st>>>(ka, n, shift, hist, ntiles);
        k_rs_scan_rows<<<256, 1024, 0, st>>>(hist, ntiles, rowsum);
        k_rs_scan_tops<<<1, 256, 0, st>>>(rowsum, rowoff);
        k_rs_scatter<<<ntiles, 256, 0, st>>>(ka, va, n, shift, hist, rowoff, ntiles, kb, vb);
        uint32_t* tk = ka;
        ka = kb;
        kb = tk;
        uint32_t* tv = va;
        va = vb;
        vb = tv;
    }
    *keys_out = ka;
    *vals_out = va;
}

}  // namespace ocl_amd
