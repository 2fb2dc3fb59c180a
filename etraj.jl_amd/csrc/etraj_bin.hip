// etraj_bin.hip -- PMD binning through LDS-privatised tiles.
//
// The integrator (with OutP::bin_cell set) writes one record per electron
// instead of adding to the PMD: the global cell (job * npix + cell, or
// kNoCell) and the contribution (CTMC: (W, 0); coherent: sqrt(W) e^{i phase}).
// Four kernels then add the records of a slab into the PMD:
//   k_tile_count    per workgroup, an LDS histogram of the records' tiles (a
//                   tile = 2^shift consecutive cells), added to tile_count;
//   k_tile_scan     one workgroup: tile_off = exclusive sums of the counts,
//                   chunk_off = exclusive sums of ceil(count / chunk), the
//                   scatter cursors, and tile_count zeroed for the next slab;
//   k_tile_scatter  per workgroup, the records sorted by tile (a range of
//                   each tile reserved with one global atomic per tile);
//   k_tile_accum    one workgroup per chunk of a tile's records: the tile is
//                   accumulated in LDS (FP64 LDS atomics, or exact 128-bit
//                   integer sums for ETRAJ_PMD_FIXED), then added to the PMD in
//                   HBM once per non-zero cell.
// The PMD sees one global atomic per (chunk, non-zero cell) instead of one per
// electron. Float sums are rounded in arrival order, as with direct atomics;
// fixed-point sums are exact, so the PMD does not depend on the path.
#include "etraj_engine.hpp"
#include "etraj_device.hpp"

#include <hip/hip_runtime.h>

namespace etraj {

constexpr int kBinThreads = 256;
constexpr int kScatterPerThread = 16;   // records per thread of k_tile_count / k_tile_scatter
constexpr int kScatterPerBlock = kBinThreads * kScatterPerThread;

__global__ __launch_bounds__(256) void k_bin_reset(uint32_t* __restrict__ cell, int64_t n,
                                                   uint32_t* __restrict__ tile_count, int ntile) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cell[i] = kNoCell;
    if (i < ntile) tile_count[i] = 0;
}

__global__ __launch_bounds__(256) void k_tile_count(const uint32_t* __restrict__ cell, int64_t n, int shift, int ntile,
                                                    uint32_t* __restrict__ tile_count) {
    extern __shared__ uint32_t hist[];
    for (int t = threadIdx.x; t < ntile; t += kBinThreads) hist[t] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kScatterPerBlock;
#pragma unroll 4
    for (int k = 0; k < kScatterPerThread; ++k) {
        const int64_t i = b0 + k * kBinThreads + threadIdx.x;
        if (i < n) {
            const uint32_t c = cell[i];
            if (c != kNoCell) atomicAdd(&hist[c >> shift], 1u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntile; t += kBinThreads)
        if (hist[t]) atomicAdd(&tile_count[t], hist[t]);
}

// One workgroup of 1024 threads; ntile <= kBinMaxTiles (8 tiles per thread).
__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t* __restrict__ tile_count, int ntile, int chunk,
                                                    uint32_t* __restrict__ tile_off, uint32_t* __restrict__ chunk_off,
                                                    uint32_t* __restrict__ cursor) {
    constexpr int kPer = kBinMaxTiles / 1024;
    __shared__ uint32_t ps[1024], pc[1024];
    const int t0 = threadIdx.x * kPer;
    uint32_t c[kPer], s = 0, q = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        c[k] = (t0 + k < ntile) ? tile_count[t0 + k] : 0u;
        s += c[k];
        q += (c[k] + chunk - 1) / chunk;
    }
    ps[threadIdx.x] = s;
    pc[threadIdx.x] = q;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {   // Hillis-Steele inclusive scans
        const uint32_t a = threadIdx.x >= (unsigned)off ? ps[threadIdx.x - off] : 0u;
        const uint32_t b = threadIdx.x >= (unsigned)off ? pc[threadIdx.x - off] : 0u;
        __syncthreads();
        ps[threadIdx.x] += a;
        pc[threadIdx.x] += b;
        __syncthreads();
    }
    uint32_t es = ps[threadIdx.x] - s, eq = pc[threadIdx.x] - q;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int t = t0 + k;
        if (t < ntile) {
            tile_off[t] = es;
            cursor[t] = es;
            chunk_off[t] = eq;
            tile_count[t] = 0;
        }
        es += c[k];
        eq += (c[k] + chunk - 1) / chunk;
    }
    if (threadIdx.x == 1023) {
        tile_off[ntile] = ps[1023];
        chunk_off[ntile] = pc[1023];
    }
}

__global__ __launch_bounds__(256) void k_tile_scatter(const uint32_t* __restrict__ cell, const double2* __restrict__ amp,
                                                      int64_t n, int shift, int ntile, uint32_t* __restrict__ cursor,
                                                      uint32_t* __restrict__ scell, double2* __restrict__ samp) {
    extern __shared__ uint32_t sh[];
    uint32_t* cnt = sh;            // records of this block per tile
    uint32_t* base = sh + ntile;   // their first slot in the sorted arrays
    for (int t = threadIdx.x; t < ntile; t += kBinThreads) cnt[t] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kScatterPerBlock;
    uint32_t c[kScatterPerThread], rank[kScatterPerThread];
#pragma unroll
    for (int k = 0; k < kScatterPerThread; ++k) {
        const int64_t i = b0 + k * kBinThreads + threadIdx.x;
        c[k] = i < n ? cell[i] : kNoCell;
        rank[k] = c[k] != kNoCell ? atomicAdd(&cnt[c[k] >> shift], 1u) : 0u;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntile; t += kBinThreads)
        if (cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScatterPerThread; ++k) {
        if (c[k] != kNoCell) {
            const int64_t i = b0 + k * kBinThreads + threadIdx.x;
            const uint32_t dst = base[c[k] >> shift] + rank[k];
            scell[dst] = c[k];
            samp[dst] = amp[i];
        }
    }
}

// W: words per cell value (1 CTMC, 2 coherent); FIXED: 128-bit integer sums.
// LDS layout: the tile's values (W doubles, or 2 W uint64 words, per cell),
// then one 32-bit hit counter per cell when B.hits is set.
template <int W, bool FIXED>
__global__ __launch_bounds__(256) void k_tile_accum(const uint32_t* __restrict__ scell, const double2* __restrict__ samp,
                                                    const uint32_t* __restrict__ tile_off,
                                                    const uint32_t* __restrict__ chunk_off, int ntile, int chunk,
                                                    int shift, int64_t total_cells, BinOut B) {
    extern __shared__ uint64_t lds[];
    const uint32_t b = blockIdx.x;
    if (b >= chunk_off[ntile]) return;
    int lo = 0, hi = ntile;   // the tile: the last t with chunk_off[t] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_off[mid] <= b) lo = mid; else hi = mid;
    }
    const int t = lo;
    const uint32_t r0 = tile_off[t] + (b - chunk_off[t]) * (uint32_t)chunk;
    const uint32_t r1 = min(r0 + (uint32_t)chunk, tile_off[t + 1]);
    const int T = 1 << shift;
    constexpr int kWords = FIXED ? 2 * W : W;   // 64-bit words per cell
    uint32_t* hits = reinterpret_cast<uint32_t*>(lds + (size_t)kWords * T);
    for (int i = threadIdx.x; i < kWords * T; i += kBinThreads) lds[i] = 0;
    if (B.hits)
        for (int i = threadIdx.x; i < T; i += kBinThreads) hits[i] = 0;
    __syncthreads();
    const uint32_t first = (uint32_t)t << shift;
    int nr = 0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBinThreads) {
        const uint32_t local = scell[i] - first;
        const double2 v = samp[i];
        if (FIXED) {
            bool ok1 = true, ok2 = true;
            fix_atomic_add(lds + 2 * W * local, fix_from_double(v.x, ok1));
            if (W == 2) fix_atomic_add(lds + 2 * W * local + 2, fix_from_double(v.y, ok2));
            nr += (ok1 && ok2) ? 0 : 1;
        } else {
            double* d = reinterpret_cast<double*>(lds) + W * local;
            unsafeAtomicAdd(d, v.x);
            if (W == 2) unsafeAtomicAdd(d + 1, v.y);
        }
        if (B.hits) atomicAdd(&hits[local], 1u);
    }
    if (FIXED && nr)
        atomicAdd((unsigned long long*)(B.cnt + ((int64_t)(first / B.npix) * kAccSlots + acc_slot()) * CNT_N +
                                        CNT_FIXRANGE),
                  (unsigned long long)nr);
    __syncthreads();
    for (int i = threadIdx.x; i < T; i += kBinThreads) {
        const int64_t g = (int64_t)first + i;
        if (g >= total_cells) break;
        if (FIXED) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const Fix128 q{lds[2 * (W * i + w)], lds[2 * (W * i + w) + 1]};
                fix_atomic_add(B.pmdq + 2 * (W * g + w), q);
            }
        } else {
            const double* d = reinterpret_cast<const double*>(lds) + W * i;
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (d[w] != 0.0) unsafeAtomicAdd(B.pmd + W * g + w, d[w]);
        }
        if (B.hits && hits[i]) atomicAdd((unsigned long long*)(B.hits + g), (unsigned long long)hits[i]);
    }
}

int bin_tile_shift(int words, bool fixed, bool hits, int max_shift) {
    const int per_cell = words * (fixed ? 16 : 8) + (hits ? 4 : 0);
    int shift = 0;
    while ((2 << shift) * per_cell <= kBinTileBytes && shift + 1 <= max_shift) ++shift;
    return shift;
}

hipError_t launch_bin_reset(hipStream_t st, uint32_t* cell, int64_t n, uint32_t* tile_count, int ntile) {
    const int64_t m = n > ntile ? n : ntile;
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_reset, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, cell, n, tile_count, ntile);
    return hipGetLastError();
}

hipError_t launch_bin_tiles(hipStream_t st, const BinCfg& cfg, const BinBufs& Bf, int64_t n, int64_t total_cells,
                            int words, bool fixed, const BinOut& out) {
    if (n <= 0) return hipSuccess;
    const int ntile = (int)((total_cells + (1ll << cfg.shift) - 1) >> cfg.shift);
    if (ntile < 1 || ntile > kBinMaxTiles || cfg.chunk < 1) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((n + kScatterPerBlock - 1) / kScatterPerBlock);
    hipLaunchKernelGGL(k_tile_count, dim3(blocks), dim3(kBinThreads), sizeof(uint32_t) * ntile, st, Bf.cell, n,
                       cfg.shift, ntile, Bf.tile_count);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, Bf.tile_count, ntile, cfg.chunk, Bf.tile_off,
                       Bf.chunk_off, Bf.cursor);
    hipLaunchKernelGGL(k_tile_scatter, dim3(blocks), dim3(kBinThreads), 2 * sizeof(uint32_t) * ntile, st, Bf.cell,
                       Bf.amp, n, cfg.shift, ntile, Bf.cursor, Bf.scell, Bf.samp);
    // chunks <= ceil(n / chunk) + ntile; the blocks past the last chunk leave at once
    const unsigned chunks = (unsigned)((n + cfg.chunk - 1) / cfg.chunk + ntile);
    const int T = 1 << cfg.shift;
    const size_t lds = (size_t)T * words * (fixed ? 16 : 8) + (out.hits ? (size_t)4 * T : 0);
    if (lds > (size_t)kBinTileBytes) return hipErrorInvalidValue;
#define ETRAJ_ACCUM(W_, F_)                                                                                        \
    hipLaunchKernelGGL((k_tile_accum<W_, F_>), dim3(chunks), dim3(kBinThreads), lds, st, Bf.scell, Bf.samp,           \
                       Bf.tile_off, Bf.chunk_off, ntile, cfg.chunk, cfg.shift, total_cells, out)
    if (words == 1) {
        if (fixed) ETRAJ_ACCUM(1, true); else ETRAJ_ACCUM(1, false);
    } else {
        if (fixed) ETRAJ_ACCUM(2, true); else ETRAJ_ACCUM(2, false);
    }
#undef ETRAJ_ACCUM
    return hipGetLastError();
}

}  // namespace etraj
