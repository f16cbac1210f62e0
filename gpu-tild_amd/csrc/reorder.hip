// reorder.hip -- periodic spatial reordering of a replica's particle arrays.
//
// The reference keeps particles in file order for the whole run (config_utils.cu:
// 5-364); its gather and spread then touch the grid in chain order.  Here, every K
// steps, the slots of a replica are permuted so that beads whose stencils start in
// the same grid row are neighbours in memory, rows ordered tile by tile (the tiles of
// the sorted-bin deposition, z-major).  A wave of the particle kernel then reads a
// few adjacent grid rows instead of 64 scattered ones, and a contiguous run of blocks
// covers one slab of the grid.
//
// Replicas on the tiled spread reuse the bucket order of the step just taken: each
// tile's beads sorted by site id (one LDS bitonic sort per tile).  Otherwise, and at
// creation, the permutation is a stable radix sort of the row keys (rocPRIM).  Both
// depend only on the positions and the previous order: runs stay bit-reproducible.  Every
// per-step result is independent of slot order (integer deposition, per-bead gather
// and bonds, noise keyed by the original site id `pid`), so trajectories are
// bit-identical with and without reordering.  Bond partners are renumbered through
// the inverse permutation.
#include <rocprim/device/device_radix_sort.hpp>

#include "device.hpp"
#include "launch.hpp"

namespace tild {

// Row key of each bead: tile-major, then the row inside the tile (a bijection of the
// Ny*Nz grid rows; with no tiles the row index itself).
__global__ void __launch_bounds__(256) k_sort_keys(const RepDev* __restrict__ reps, int rep,
                                                   uint32_t* __restrict__ keys, int* __restrict__ vals) {
    const RepDev& r = reps[rep];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= r.ns) return;
    const int D = r.dim;
    const float* x = r.X[r.step & 1];
    const int sy = stencil_node(x[i * D + 1], r.dx[1], r.Ny, r.P);
    const int sz = D == 3 ? stencil_node(x[i * D + 2], r.dx[2], r.Nz, r.P) : 0;
    uint32_t key;
    if (r.spread_mode == 2) {
        const int ty = sy / r.By, tz = sz / r.Bz;
        key = (uint32_t)(((tz * r.nty + ty) * r.Bz + sz % r.Bz) * r.By + sy % r.By);
    } else {
        key = (uint32_t)(sz * r.Ny + sy);
    }
    keys[i] = key;
    vals[i] = i;
}

// inv[old slot] = new slot
__global__ void __launch_bounds__(256) k_reorder_inv(int ns, const int* __restrict__ order, int* __restrict__ inv) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < ns) inv[order[p]] = p;
}

// Gather every per-bead array into the spare buffers in the new order.  Positions
// go to the free half of the position double buffer.
__device__ __forceinline__ void reorder_apply(const RepDev& r, int p, const int* __restrict__ order,
                                              const int* __restrict__ inv) {
    const int D = r.dim, i = order[p];
    const float* __restrict__ x = r.X[r.step & 1];
    float* __restrict__ xn = r.X[(r.step + 1) & 1];
    for (int d = 0; d < D; ++d) {
        xn[p * D + d] = x[i * D + d];
        r.XO_alt[p * D + d] = r.XO[i * D + d];
    }
    r.tp_alt[p] = r.tp[i];
    r.ikind_alt[p] = r.ikind[i];
    r.pid_alt[p] = r.pid[i];
    if (r.q) r.q_alt[p] = r.q[i];
    const int s = r.bond_shift, tmask = (1 << s) - 1;
    for (int b = 0; b < 3; ++b) {
        const int v = r.bond[i * 3 + b];
        r.bond_alt[p * 3 + b] = v < 0 ? -1 : (inv[v >> s] << s) | (v & tmask);
    }
}

__global__ void __launch_bounds__(256) k_reorder_apply(const RepDev* __restrict__ reps, int rep,
                                                       const int* __restrict__ order,
                                                       const int* __restrict__ inv) {
    const RepDev& r = reps[rep];
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < r.ns) reorder_apply(r, p, order, inv);
}

template <class T, class U>
__device__ inline void swap_ptr(T*& a, U*& b) {
    T* t = a;
    a = (T*)b;
    b = (U*)t;
}

// Make the reordered arrays live (one thread; stream-ordered after the gather).
__device__ __forceinline__ void reorder_swap(RepDev& r) {
    swap_ptr(r.X[0], r.X[1]);
    swap_ptr(r.XO, r.XO_alt);
    swap_ptr(r.tp, r.tp_alt);
    swap_ptr(r.ikind, r.ikind_alt);
    swap_ptr(r.pid, r.pid_alt);
    if (r.q) swap_ptr(r.q, r.q_alt);
    swap_ptr(r.bond, r.bond_alt);
}

__global__ void k_reorder_swap(RepDev* reps, int rep) { reorder_swap(reps[rep]); }

// ---- tiled replicas: reuse the last step's bucket order -----------------------------
// The step's binning left every bead's bucket (tile, halo class) and slot in tkey and
// the bucket offsets in toff, so the beads of tile t occupy [toff[4t], toff[4t+4]) of
// the bucket order.  Inside a tile the slot order is arbitrary (atomics); sorting each
// tile's beads by original site id makes the new order deterministic.

// src[bucket position] = slot
__device__ __forceinline__ void reorder_src(const RepDev& r, int i, int* __restrict__ src) {
    const long long kk = G((const long long*)r.tkey)[i];
    const int b = (int)(unsigned)(kk & 0xffffffffll), s = (int)(kk >> 32);
    G(src)[G((const int*)r.toff)[b] + s] = i;
}

__global__ void __launch_bounds__(256) k_reorder_src(const RepDev* __restrict__ reps, int rep, int* __restrict__ src) {
    const RepDev& r = reps[rep];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < r.ns) reorder_src(r, i, src);
}

constexpr int kTileSortMax = 2048;   // beads a tile sorts in LDS; larger tiles rank in place

// One block per tile: order[tile range] = the tile's slots sorted by site id.
__device__ __forceinline__ void reorder_tile_sort(const RepDev& r, int t, const int* __restrict__ src,
                                                  int* __restrict__ order) {
    __shared__ int key[kTileSortMax];
    __shared__ int val[kTileSortMax];
    const int a = G((const int*)r.toff)[4 * t], L = G((const int*)r.toff)[4 * t + 4] - a;
    gmem<const int> pid = G((const int*)r.pid);
    if (L > kTileSortMax) {
        // rare dense tile: the rank of each site id inside the tile places it
        for (int j = threadIdx.x; j < L; j += 1024) {
            const int sj = G(src)[a + j], kj = pid[sj];
            int rank = 0;
            for (int k = 0; k < L; ++k) rank += pid[G(src)[a + k]] < kj;
            G(order)[a + rank] = sj;
        }
        return;
    }
    int n2 = 64;
    while (n2 < L) n2 <<= 1;
    for (int j = threadIdx.x; j < n2; j += 1024) {
        if (j < L) {
            const int sj = G(src)[a + j];
            key[j] = pid[sj];
            val[j] = sj;
        } else {
            key[j] = 0x7fffffff;
            val[j] = -1;
        }
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1)          // bitonic sort, ascending site id
        for (int h = k >> 1; h > 0; h >>= 1) {
            for (int i = threadIdx.x; i < n2; i += 1024) {
                const int ixj = i ^ h;
                if (ixj > i) {
                    const bool up = (i & k) == 0;
                    const int ki = key[i], kx = key[ixj];
                    if ((ki > kx) == up) {
                        key[i] = kx; key[ixj] = ki;
                        const int v = val[i]; val[i] = val[ixj]; val[ixj] = v;
                    }
                }
            }
            __syncthreads();
        }
    for (int j = threadIdx.x; j < L; j += 1024) G(order)[a + j] = val[j];
}

__global__ void __launch_bounds__(1024) k_reorder_tile_sort(const RepDev* __restrict__ reps, int rep,
                                                           const int* __restrict__ src, int* __restrict__ order) {
    reorder_tile_sort(reps[rep], blockIdx.x, src, order);
}

// ---- sets of tiled replicas: the same kernels over work maps, scratch in RepDev
__global__ void __launch_bounds__(256) k_reorder_src_set(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int i = wi.blk * 256 + threadIdx.x;
    if (i < r.ns) reorder_src(r, i, r.rsrc);
}
__global__ void __launch_bounds__(1024) k_reorder_tile_sort_set(const RepDev* __restrict__ reps,
                                                               const WorkItem* __restrict__ wmap) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    reorder_tile_sort(r, wi.blk, r.rsrc, r.rorder);
}
__global__ void __launch_bounds__(256) k_reorder_inv_set(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int p = wi.blk * 256 + threadIdx.x;
    if (p < r.ns) G(r.rinv)[G((const int*)r.rorder)[p]] = p;
}
__global__ void __launch_bounds__(256) k_reorder_apply_set(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int p = wi.blk * 256 + threadIdx.x;
    if (p < r.ns) reorder_apply(r, p, r.rorder, r.rinv);
}
__global__ void __launch_bounds__(64) k_reorder_swap_set(RepDev* reps, const WorkItem* __restrict__ wmap) {
    if (threadIdx.x == 0) reorder_swap(reps[wmap[blockIdx.x].rep]);
}

#define CHECK_LAUNCH() do { hipError_t e__ = hipGetLastError(); if (e__ != hipSuccess) return e__; } while (0)

hipError_t launch_reorder_tiles(hipStream_t s, RepDev* reps, int rep, int ns, int ntiles, ReorderBufs& b) {
    const int nb = (ns + 255) / 256;
    int* src = b.vals[0];
    int* order = b.vals[1];
    hipLaunchKernelGGL(k_reorder_src, dim3(nb), dim3(256), 0, s, reps, rep, src);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_tile_sort, dim3(ntiles), dim3(1024), 0, s, reps, rep, src, order);
    CHECK_LAUNCH();
    int* inv = (int*)b.keys[0];
    hipLaunchKernelGGL(k_reorder_inv, dim3(nb), dim3(256), 0, s, ns, order, inv);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_apply, dim3(nb), dim3(256), 0, s, reps, rep, order, inv);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_swap, dim3(1), dim3(1), 0, s, reps, rep);
    CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_reorder_tiles_set(const LaunchCtx& beads, const LaunchCtx& tiles, const LaunchCtx& reps1) {
    if (!beads.n || !tiles.n || !reps1.n) return hipSuccess;
    hipLaunchKernelGGL(k_reorder_src_set, dim3(beads.n), dim3(256), 0, beads.stream, beads.reps, beads.wmap);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_tile_sort_set, dim3(tiles.n), dim3(1024), 0, tiles.stream, tiles.reps, tiles.wmap);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_inv_set, dim3(beads.n), dim3(256), 0, beads.stream, beads.reps, beads.wmap);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_apply_set, dim3(beads.n), dim3(256), 0, beads.stream, beads.reps, beads.wmap);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_swap_set, dim3(reps1.n), dim3(64), 0, reps1.stream, const_cast<RepDev*>(reps1.reps), reps1.wmap);
    CHECK_LAUNCH();
    return hipSuccess;
}

size_t reorder_tmp_bytes(int ns, int bits) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (const int*)nullptr, (int*)nullptr, (unsigned)ns, 0u, (unsigned)bits) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_reorder(hipStream_t s, RepDev* reps, int rep, int ns, int bits, ReorderBufs& b) {
    const int nb = (ns + 255) / 256;
    hipLaunchKernelGGL(k_sort_keys, dim3(nb), dim3(256), 0, s, reps, rep, b.keys[0], b.vals[0]);
    CHECK_LAUNCH();
    size_t tb = b.tmp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(b.tmp, tb, b.keys[0], b.keys[1], b.vals[0], b.vals[1], (unsigned)ns,
                                             0u, (unsigned)bits, s);
    if (e != hipSuccess) return e;
    int* inv = (int*)b.keys[0];   // the unsorted keys are no longer needed
    hipLaunchKernelGGL(k_reorder_inv, dim3(nb), dim3(256), 0, s, ns, b.vals[1], inv);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_apply, dim3(nb), dim3(256), 0, s, reps, rep, b.vals[1], inv);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reorder_swap, dim3(1), dim3(1), 0, s, reps, rep);
    CHECK_LAUNCH();
    return hipSuccess;
}

}  // namespace tild
