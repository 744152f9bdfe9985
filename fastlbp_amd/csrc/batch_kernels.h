/*
 * batch_kernels.h -- one PARALL sweep over many graphs at once.
 *
 * Graphs are concatenated: graph g's factors start at fac_base[g], its
 * variables at var_base[g] and its edge slots at edge_base[g] (64-bit
 * bases), and all indices stored inside a graph's arrays are graph-local
 * 32-bit values.  Work is cut into tiles of one graph each -- the staged
 * blocks of the plain solve (at most kBlock items and kStage edge slots,
 * with their work orders), in the pull or the scatter message layout of
 * staged_kernels.h -- and a persistent grid walks a tile list with a
 * stride, so a launch uses exactly the blocks that fit on the chip
 * whatever the batch holds.  The sweep kernels walk the list of the graphs
 * still running (its length read from device memory, so a captured chunk
 * stays valid when the host shortens the list between replays).
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

#include "staged_kernels.h"
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

template <bool kDamp, int N, int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_factor_pass(
    const BatchGraph *__restrict__ graphs, const Tile *__restrict__ tiles, const uint32_t *__restrict__ n_tiles_p,
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q,
    const double2 *__restrict__ fac_mask, const uint32_t *__restrict__ fac_vpos,
    const uint32_t *__restrict__ fac_work, const double2 *__restrict__ m_vf, double2 *m_fv,
    double2 *__restrict__ m_fv_vm, double damp, double keep, BatchCtl ctl, int k) {
    __shared__ FactorSmem sm;
    const uint32_t n_tiles = *n_tiles_p;
    for (uint32_t ti = blockIdx.x; ti < n_tiles; ti += gridDim.x) {
        const Tile tile = tiles[ti];
        if (!graph_runs(ctl, tile.graph, k)) continue;   /* uniform per block */
        const BatchGraph gr = graphs[tile.graph];
        const uint64_t eb = gr.edge_base, fb = gr.fac_base;
        factor_tile<kDamp, N, kLayout>(fac_off + gr.fac_off_base, fac_meta + fb, fac_p + fb, fac_q + fb,
                                              fac_mask + fb, fac_vpos + eb, fac_work + eb, m_vf + eb, m_fv + eb,
                                              m_fv + eb, m_fv_vm + eb, damp, keep, tile.first,
                                              tile.first + tile.count, sm);
        __syncthreads();   /* LDS is reused by the next tile */
    }
}

template <int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_var_pass(
    const BatchGraph *__restrict__ graphs, const Tile *__restrict__ tiles, const uint32_t *__restrict__ n_tiles_p,
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const uint16_t *__restrict__ var_work, const double2 *__restrict__ m_fv_in, double2 *__restrict__ m_vf,
    double *__restrict__ belief, BatchCtl ctl, int k) {
    __shared__ VarSmem sm;
    const uint32_t n_tiles = *n_tiles_p;
    /* each round of the loop covers tiles [base, base + span); in the pull
     * layout the round's tiles go to the XCDs in runs of kPullVarRun, as in
     * the plain var pass (465.7 vs 470.9 ms for the 768-graph batch) */
    for (uint32_t base = 0; base + blockIdx.x < n_tiles; base += gridDim.x) {
        const uint32_t left = n_tiles - base, span = left < gridDim.x ? left : gridDim.x;
        const uint32_t ti = base + (kLayout == kLayoutPull ? xcd_runs<kPullVarRun>(blockIdx.x, span) : blockIdx.x);
        const Tile tile = tiles[ti];
        if (!graph_runs(ctl, tile.graph, k)) continue;   /* uniform per block */
        const BatchGraph gr = graphs[tile.graph];
        const uint64_t eb = gr.edge_base;
        unsigned long long diff_bits =
            var_tile<kLayout>(var_off + gr.var_off_base, var_edge + eb, var_work + gr.var_base, m_fv_in + eb,
                              m_vf + eb, belief + gr.var_base, tile.first, tile.first + tile.count, sm);
        diff_bits = wave_max_u64(diff_bits);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) sm.red[wave] = diff_bits;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = sm.red[0];
#pragma unroll
            for (int w = 1; w < kBlock / 64; w++) m = sm.red[w] > m ? sm.red[w] : m;
            if (m) {
                unsigned long long *slot = batch_slot(ctl, k, tile.graph, ti % kBatchSlots);
                if (m > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, m);
            }
        }
        __syncthreads();   /* LDS is reused by the next tile */
    }
}

/* message layout moves over every factor tile (graph-local positions):
 * to_facmajor: dst[s] = src[vpos[s]]; to_varmajor: dst[vpos[s]] = src[s] */
template <bool kToFacMajor>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_permute(
    const BatchGraph *__restrict__ graphs, const Tile *__restrict__ tiles, uint32_t n_tiles,
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_vpos,
    const double2 *__restrict__ src, double2 *__restrict__ dst) {
    for (uint32_t ti = blockIdx.x; ti < n_tiles; ti += gridDim.x) {
        const Tile tile = tiles[ti];
        const BatchGraph gr = graphs[tile.graph];
        const uint32_t *off = fac_off + gr.fac_off_base;
        const uint32_t base = off[tile.first], n = off[tile.first + tile.count] - base;
        const uint64_t eb = gr.edge_base;
        for (uint32_t e = threadIdx.x; e < n; e += kBlock) {
            const uint64_t s = eb + base + e, v = eb + fac_vpos[s];
            if (kToFacMajor) dst[s] = src[v];
            else dst[v] = src[s];
        }
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
