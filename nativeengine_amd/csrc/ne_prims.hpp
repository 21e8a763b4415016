// ne_prims.hpp -- device-wide primitives for the gfx950 step: exclusive scan
// and a stable LSD radix sort, both for u32 data whose element count may live
// in device memory (so a step never waits on the host to learn a count).
//
// wave64: intra-wave scans use DPP-backed shuffles (__shfl_up), ranking in
// the radix scatter uses 64-bit ballots.  Integer only: results are exact and
// independent of scheduling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nea {

constexpr int PRIM_THREADS = 256;          // 4 waves
constexpr int PRIM_WAVES = PRIM_THREADS / 64;

// Exclusive scan of one value per thread across a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* lds /*>=PRIM_WAVES+1*/) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < PRIM_WAVES; ++w) { uint32_t t = lds[w]; lds[w] = run; run += t; }
        lds[PRIM_WAVES] = run;
    }
    __syncthreads();
    uint32_t res = lds[wid] + x - v;
    *total = lds[PRIM_WAVES];
    __syncthreads();
    return res;
}

// n is min(*n_dev, n_max) when n_dev != nullptr, else n_max.
__device__ __forceinline__ uint32_t resolve_n(const uint32_t* n_dev, uint32_t n_max) {
    if (!n_dev) return n_max;
    uint32_t n = *n_dev;
    return n < n_max ? n : n_max;
}

// Phase 1: per-block sums over contiguous chunks.
__global__ void __launch_bounds__(PRIM_THREADS) k_scan_partial(const uint32_t* __restrict__ in, const uint32_t* n_dev,
                                                              uint32_t n_max, uint32_t chunk, uint32_t* __restrict__ partial) {
    __shared__ uint32_t lds[PRIM_WAVES + 1];
    const uint32_t n = resolve_n(n_dev, n_max);
    const uint32_t beg = blockIdx.x * chunk;
    const uint32_t end = min(beg + chunk, n);
    uint32_t s = 0;
    for (uint32_t t = beg + threadIdx.x; t < end; t += PRIM_THREADS) s += in[t];
    uint32_t tot;
    block_exclusive_scan(s, &tot, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Optional bookkeeping done by the scan's single-block phase on the grand
// total: counters[raw] = total, counters[eff] = min(total, cap), fault bit
// when it overflowed, and a list of counters to zero for the next stage.
struct ScanTotal {
    uint32_t* counters;
    int raw_slot, eff_slot, fault_slot;
    uint32_t cap, fault_bit;
    int zero_slot0, zero_slot1;
};

// Phase 2: one block scans the partials; writes the grand total to out[n].
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* __restrict__ partial, uint32_t nblocks,
                                                  const uint32_t* n_dev, uint32_t n_max, uint32_t* __restrict__ out,
                                                  ScanTotal tot) {
    __shared__ uint32_t buf[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        uint32_t t = base + threadIdx.x;
        uint32_t v = t < nblocks ? partial[t] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            uint32_t y = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += y;
            __syncthreads();
        }
        uint32_t incl = buf[threadIdx.x];
        if (t < nblocks) partial[t] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[resolve_n(n_dev, n_max)] = carry;
        if (tot.counters) {
            uint32_t r = carry;
            tot.counters[tot.raw_slot] = r;
            if (r > tot.cap) { atomicOr(&tot.counters[tot.fault_slot], tot.fault_bit); r = tot.cap; }
            tot.counters[tot.eff_slot] = r;
            if (tot.zero_slot0 >= 0) tot.counters[tot.zero_slot0] = 0u;
            if (tot.zero_slot1 >= 0) tot.counters[tot.zero_slot1] = 0u;
        }
    }
}

// Phase 3: rescan each chunk from its block offset.
__global__ void __launch_bounds__(PRIM_THREADS) k_scan_final(uint32_t* in, const uint32_t* n_dev,
                                                            uint32_t n_max, uint32_t chunk,
                                                            const uint32_t* __restrict__ partial, uint32_t* out,
                                                            int clear_in) {
    __shared__ uint32_t lds[PRIM_WAVES + 1];
    const uint32_t n = resolve_n(n_dev, n_max);
    const uint32_t beg = blockIdx.x * chunk;
    const uint32_t end = min(beg + chunk, n);
    uint32_t run = partial[blockIdx.x];
    for (uint32_t base = beg; base < end; base += PRIM_THREADS) {
        uint32_t t = base + threadIdx.x;
        uint32_t v = t < end ? in[t] : 0;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, &tot, lds);
        if (t < end) {
            if (clear_in) in[t] = 0u;   // leaves a zeroed histogram for the next use
            out[t] = run + ex;
        }
        run += tot;
    }
}

// Radix sort, 8-bit digits.
constexpr int RS_BINS = 256;

__global__ void __launch_bounds__(PRIM_THREADS) k_rs_hist(const uint32_t* __restrict__ keys, const uint32_t* n_dev,
                                                         uint32_t n_max, uint32_t chunk, int shift, uint32_t nblocks,
                                                         uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[RS_BINS];
    for (int d = threadIdx.x; d < RS_BINS; d += PRIM_THREADS) h[d] = 0;
    __syncthreads();
    const uint32_t n = resolve_n(n_dev, n_max);
    const uint32_t beg = blockIdx.x * chunk;
    const uint32_t end = min(beg + chunk, n);
    for (uint32_t t = beg + threadIdx.x; t < end; t += PRIM_THREADS) atomicAdd(&h[(keys[t] >> shift) & 255u], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < RS_BINS; d += PRIM_THREADS) hist[(uint32_t)d * nblocks + blockIdx.x] = h[d];
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Stable scatter: elements keep their input order within a digit.
__global__ void __launch_bounds__(PRIM_THREADS) k_rs_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                            const uint32_t* n_dev, uint32_t n_max, uint32_t chunk, int shift,
                                                            uint32_t nblocks, const uint32_t* __restrict__ hist_scanned,
                                                            uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t base[RS_BINS];
    __shared__ uint32_t wcnt[PRIM_WAVES][RS_BINS];
    for (int d = threadIdx.x; d < RS_BINS; d += PRIM_THREADS) base[d] = hist_scanned[(uint32_t)d * nblocks + blockIdx.x];
    const uint32_t n = resolve_n(n_dev, n_max);
    const uint32_t beg = blockIdx.x * chunk;
    const uint32_t end = min(beg + chunk, n);
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    for (uint32_t tb = beg; tb < end; tb += PRIM_THREADS) {
        for (int d = threadIdx.x; d < RS_BINS * PRIM_WAVES; d += PRIM_THREADS) (&wcnt[0][0])[d] = 0;
        __syncthreads();
        const uint32_t t = tb + threadIdx.x;
        const bool valid = t < end;
        uint32_t key = valid ? keys[t] : 0u;
        uint32_t val = valid ? vals[t] : 0u;
        uint32_t dig = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot(valid && ((dig >> b) & 1u));
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && rank == 0) wcnt[wid][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[dig];
            for (int w = 0; w < wid; ++w) off += wcnt[w][dig];
            keys_out[off + rank] = key;
            vals_out[off + rank] = val;
        }
        __syncthreads();
        for (int d = threadIdx.x; d < RS_BINS; d += PRIM_THREADS) {
            uint32_t s = 0;
            for (int w = 0; w < PRIM_WAVES; ++w) s += wcnt[w][d];
            base[d] += s;
        }
        __syncthreads();
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, const uint32_t* n_dev, uint32_t n_max) {
    const uint32_t n = resolve_n(n_dev, n_max);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, const uint32_t* n_dev, uint32_t n_max) {
    const uint32_t n = resolve_n(n_dev, n_max);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = t;
}

// Host-side drivers ---------------------------------------------------------

struct ScanPlan {
    uint32_t nblocks, chunk;
};

inline ScanPlan scan_plan(uint32_t n_max) {
    ScanPlan p;
    uint32_t nb = (n_max + 4095u) / 4096u;
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    p.nblocks = nb;
    p.chunk = (n_max + nb - 1) / nb;
    if (p.chunk == 0) p.chunk = 1;
    return p;
}

// out has n+1 entries: out[t] = sum(in[0..t)), out[n] = total.
// clear_in: zero the input as it is consumed (it must not alias out).
inline void scan_u32(uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_max, uint32_t* partial,
                     hipStream_t s, bool clear_in = false, ScanTotal tot = ScanTotal{nullptr, 0, 0, 0, 0, 0, -1, -1}) {
    ScanPlan p = scan_plan(n_max);
    k_scan_partial<<<p.nblocks, PRIM_THREADS, 0, s>>>(in, n_dev, n_max, p.chunk, partial);
    k_scan_top<<<1, 1024, 0, s>>>(partial, p.nblocks, n_dev, n_max, out, tot);
    k_scan_final<<<p.nblocks, PRIM_THREADS, 0, s>>>(in, n_dev, n_max, p.chunk, partial, out, clear_in ? 1 : 0);
}

// Sorts (keys, vals) by the low `bits` bits of keys, stably.  Ping-pongs
// between the two buffers; returns true if the result sits in (k1, v1).
inline bool radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* n_dev,
                           uint32_t n_max, int bits, uint32_t* hist /*256*nb+1*/, uint32_t* partial, hipStream_t s) {
    ScanPlan p = scan_plan(n_max);
    bool flip = false;
    for (int shift = 0; shift < bits; shift += 8) {
        uint32_t* ki = flip ? k1 : k0;
        uint32_t* vi = flip ? v1 : v0;
        uint32_t* ko = flip ? k0 : k1;
        uint32_t* vo = flip ? v0 : v1;
        k_rs_hist<<<p.nblocks, PRIM_THREADS, 0, s>>>(ki, n_dev, n_max, p.chunk, shift, p.nblocks, hist);
        scan_u32(hist, hist, nullptr, RS_BINS * p.nblocks, partial, s);
        k_rs_scatter<<<p.nblocks, PRIM_THREADS, 0, s>>>(ki, vi, n_dev, n_max, p.chunk, shift, p.nblocks, hist, ko, vo);
        flip = !flip;
    }
    return flip;
}

}  // namespace nea
