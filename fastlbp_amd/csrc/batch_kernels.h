/*
 * batch_kernels.h -- one PARALL sweep over many graphs at once.
 *
 * Graphs are concatenated: graph g's factors start at fac_base[g], its
 * variables at var_base[g] and its edge slots at edge_base[g] (64-bit
 * bases), and all indices stored inside a graph's arrays are graph-local
 * 32-bit values.  Work is cut into tiles of one graph each; a persistent
 * grid walks the tile list with a stride, so a launch uses exactly the
 * blocks that fit on the chip whatever the batch holds.
 *
 * Each graph keeps its own stop test: sweep k of a chunk runs on graph g
 * only if g was active when the chunk started, k is below its remaining
 * iteration budget, and its max belief change of sweep k-1 was > tol.  A
 * graph therefore stops after exactly the sweeps a plain solve of it would
 * run, and its messages are bit-identical to that solve.
 */
#ifndef FLBP_BATCH_KERNELS_H
#define FLBP_BATCH_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_kernels.h"

namespace {

using namespace flbp::dev;

struct BatchGraph {
    uint64_t fac_base, var_base, edge_base;   /* element offsets of graph g */
    uint64_t fac_off_base, var_off_base;      /* offsets into fac_off / var_off (F_g+1, V_g+1 entries) */
    uint32_t n_fac, n_var;
};

struct Tile {
    uint32_t graph;
    uint32_t first;    /* first local factor (factor tiles) or variable (var tiles) */
    uint32_t count;
};

/* per graph and sweep, kBatchSlots partial maxima one 64-byte line apart
 * (tile i of a graph folds into slot i mod kBatchSlots) */
constexpr int kBatchSlots = 8;

struct BatchCtl {
    unsigned long long *diff;       /* [chunk][G][kBatchSlots][kCtlPad] partial max |db| bits */
    const uint8_t *start_active;    /* [G] */
    const uint32_t *budget;         /* [G] sweeps left before maxiter */
    uint32_t G;
    double tol;
};

__device__ __forceinline__ unsigned long long *batch_slot(const BatchCtl &c, int k, uint32_t g, uint32_t j) {
    return c.diff + (((uint64_t)k * c.G + g) * kBatchSlots + j) * kCtlPad;
}

/* max of graph g's slots for sweep k; all lanes of the wave get it */
__device__ __forceinline__ unsigned long long batch_read(const BatchCtl &c, int k, uint32_t g) {
    const int lane = threadIdx.x & 63;
    const unsigned long long v =
        lane < kBatchSlots ? __atomic_load_n(batch_slot(c, k, g, lane), __ATOMIC_RELAXED) : 0ull;
    return wave_max_u64(v);
}

__device__ __forceinline__ bool graph_runs(const BatchCtl &c, uint32_t g, int k) {
    if (!c.start_active[g] || (uint32_t)k >= c.budget[g]) return false;
    if (k == 0) return true;
    return __longlong_as_double((long long)batch_read(c, k - 1, g)) > c.tol;
}

template <bool kDamp, int N>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_factor_pass(
    const BatchGraph *__restrict__ graphs, const Tile *__restrict__ tiles, uint32_t n_tiles,
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q,
    const double2 *__restrict__ fac_mask, const uint32_t *__restrict__ fac_vpos,
    const double2 *__restrict__ m_vf, double2 *__restrict__ m_fv, double damp, double keep,
    BatchCtl ctl, int k) {
    for (uint32_t ti = blockIdx.x; ti < n_tiles; ti += gridDim.x) {
        const Tile tile = tiles[ti];
        if (!graph_runs(ctl, tile.graph, k)) continue;
        const BatchGraph gr = graphs[tile.graph];
        const uint32_t *off = fac_off + gr.fac_off_base;
        const double2 *vf = m_vf + gr.edge_base;
        const uint32_t *vpos = fac_vpos + gr.edge_base;
        double2 *fv = m_fv + gr.edge_base;
        for (uint32_t lf = tile.first + threadIdx.x; lf < tile.first + tile.count; lf += kBlock) {
            const uint64_t f = gr.fac_base + lf;
            const uint32_t s = off[lf];
            const uint32_t meta = fac_meta[f];
            const double p = fac_p[f];
            const double q = (meta & 3u) ? fac_q[f] : 0.0;
            double mask0 = 1.0, mask1 = 1.0;
            if (meta & 4u) {
                const double2 mk = fac_mask[f];
                mask0 = mk.x;
                mask1 = mk.y;
            }
            factor_update<N, kDamp>(vf, vpos + s, fv + s, fv + s, (int)(off[lf + 1] - s), meta, p, q, mask0,
                                    mask1, damp, keep);
        }
    }
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_var_pass(
    const BatchGraph *__restrict__ graphs, const Tile *__restrict__ tiles, uint32_t n_tiles,
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf, double *__restrict__ belief,
    BatchCtl ctl, int k) {
    __shared__ unsigned long long red[kBlock / 64];
    for (uint32_t ti = blockIdx.x; ti < n_tiles; ti += gridDim.x) {
        const Tile tile = tiles[ti];
        if (!graph_runs(ctl, tile.graph, k)) continue;   /* uniform per block */
        const BatchGraph gr = graphs[tile.graph];
        const uint32_t *off = var_off + gr.var_off_base;
        const uint32_t *ve = var_edge + gr.edge_base;
        const double2 *fv = m_fv + gr.edge_base;
        double2 *vf = m_vf + gr.edge_base;
        double *bel = belief + gr.var_base;
        unsigned long long diff_bits = 0;
        for (uint32_t lv = tile.first + threadIdx.x; lv < tile.first + tile.count; lv += kBlock) {
            const uint32_t s = off[lv];
            const unsigned long long bits = var_update(ve + s, fv, vf + s, bel + lv, (int)(off[lv + 1] - s));
            diff_bits = bits > diff_bits ? bits : diff_bits;
        }
        diff_bits = wave_max_u64(diff_bits);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) red[wave] = diff_bits;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = red[0];
#pragma unroll
            for (int w = 1; w < kBlock / 64; w++) m = red[w] > m ? red[w] : m;
            if (m) {
                unsigned long long *slot = batch_slot(ctl, k, tile.graph, ti % kBatchSlots);
                if (m > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, m);
            }
        }
        __syncthreads();   /* red[] is reused by the next tile */
    }
}

/* compact the chunk's partial maxima into out[k][g] for the host */
__global__ __attribute__((unused)) void batch_compact(BatchCtl c, unsigned long long *out, int n) {
    const uint64_t waves = (uint64_t)n * c.G;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; w < waves;
         w += ((uint64_t)gridDim.x * blockDim.x) / 64) {
        const int k = (int)(w / c.G);
        const uint32_t g = (uint32_t)(w % c.G);
        const unsigned long long m = batch_read(c, k, g);
        if ((threadIdx.x & 63) == 0) out[w] = m;
    }
}

__global__ __attribute__((unused)) void batch_clear(unsigned long long *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace

#endif
