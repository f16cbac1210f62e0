// tile_dev.hpp -- bucket of a bead and slot claims in the bucket regions, shared by the
// bucket-claim kernel and the particle kernel (which claims the next step's slots from
// the positions it has just written).
//
// The grid is cut into full-x slabs of By x Bz rows.  A bead's bucket is (tile of its
// first stencil row, halo class).  Halo classes: a particle whose stencil rows also
// reach the next tile in y (yb) and/or z (zb) is tagged 0 none, 1 y only, 2 y and z,
// 3 z only.  This order makes every neighbour's needed subset contiguous:
// y-neighbour classes [1,3), z-neighbour [2,4), yz-neighbour [2,3).
#pragma once
#include "device.hpp"

namespace tild {

// Bucket of a bead from its y and z coordinates (z ignored in 2D).
__device__ __forceinline__ int tile_bucket_yz(const RepDev& r, float y, float z) {
    const int D = r.dim, P = r.P;
    const int sy = stencil_node(y, r.dx[1], r.Ny, P);
    const int sz = D == 3 ? stencil_node(z, r.dx[2], r.Nz, P) : 0;
    int my, mz;
    const int ty = div_small(sy, r.By, r.iBy, my), tz = div_small(sz, r.Bz, r.iBz, mz);
    const bool yb = my + P >= r.By;
    const bool zb = D == 3 && mz + P >= r.Bz;
    const int c = (yb && zb) ? 2 : (yb ? 1 : (zb ? 3 : 0));
    return (tz * r.nty + ty) * 4 + c;
}

// Order-1 stencil record of a bead (RepDev::srec): what the deposit needs, computed once
// by the bead's claim rather than by every tile that deposits it.
//   x  first stencil node along x | (my, mz) << 16, << 24: the first stencil row's
//      offsets inside the bead's tile
//   y  lower x weight (fixed_lower) | type << 23
//   z  lower y weight;  w  lower z weight (0 in 2D)
// The upper weights are the complements kFixedOne - lower.  Returns the bead's bucket
// (the same as tile_bucket_yz).  Requires Nx <= 65536, By, Bz <= 256 and fewer than 512
// types (the engine checks).
__device__ __forceinline__ int stencil_record(const RepDev& r, float4 rec, uint4& out) {
    const int D = r.dim;
    float W[3][2];
    const int nx0 = stencil_1d(rec.x, r.dx[0], r.idx[0], r.Nx, 1, W[0]);
    const int sy = stencil_1d(rec.y, r.dx[1], r.idx[1], r.Ny, 1, W[1]);
    const int sz = D == 3 ? stencil_1d(rec.z, r.dx[2], r.idx[2], r.Nz, 1, W[2]) : 0;
    int my, mz;
    const int ty = div_small(sy, r.By, r.iBy, my), tz = div_small(sz, r.Bz, r.iBz, mz);
    const bool yb = my + 1 >= r.By;
    const bool zb = D == 3 && mz + 1 >= r.Bz;
    const int c = (yb && zb) ? 2 : (yb ? 1 : (zb ? 3 : 0));
    out = make_uint4((uint32_t)nx0 | (uint32_t)my << 16 | (uint32_t)mz << 24,
                     fixed_lower(W[0][0]) | (uint32_t)__float_as_int(rec.w) << 23,
                     fixed_lower(W[1][0]), D == 3 ? fixed_lower(W[2][0]) : 0u);
    return (tz * r.nty + ty) * 4 + c;
}

__device__ __forceinline__ int tile_bucket(const RepDev& r, const float* xp, int i) {
    gmem<const float> x = G(xp);
    return tile_bucket_yz(r, x[i * r.dim + 1], r.dim == 3 ? x[i * r.dim + 2] : 0.f);
}

// Slot claims.  A block counts its beads in LDS over a window of kCountWin buckets
// that starts at the block's smallest bucket -- in a spatially ordered replica that
// is every bead of the block -- and claims slots from the counters `cnt` (bucket b at
// b << tshift) with one atomic per distinct bucket.  Beads beyond the window claim per
// wave: lanes that share a bucket use one atomic, issued by the lowest of them.  Slot
// order inside a bucket is arbitrary; the deposit does not depend on it.  One bead per
// thread of a 256-thread block (bucket b, 0x7fffffff for a thread without a bead);
// every thread of the block must call it.
constexpr int kCountWin = 256;

__device__ __forceinline__ void claim_slot(const RepDev& r, bool valid, int b, int* cnt, int& slot) {
    __shared__ int hist[kCountWin];
    __shared__ int wmin[4];
    const int lane = (int)__lane_id();
    int m = b;
    for (int k = threadIdx.x; k < kCountWin; k += 256) hist[k] = 0;
    for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o));
    if (lane == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    const int bmin = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
    const bool inwin = valid && b - bmin < kCountWin;
    const int local = inwin ? atomicAdd(&hist[b - bmin], 1) : 0;
    __syncthreads();
    for (int k = threadIdx.x; k < kCountWin; k += 256) {
        const int c = hist[k];
        if (c) hist[k] = gatomic_add(cnt + ((bmin + k) << r.tshift), c);   // first slot of the bucket
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    slot = inwin ? hist[b - bmin] + local : 0;
    unsigned long long todo = __ballot(valid && !inwin);
    while (todo) {   // wave-uniform: one round per distinct bucket beyond the window
        const int leader = __ffsll((long long)todo) - 1;
        const int bl = __shfl(b, leader);
        const unsigned long long same = __ballot(b == bl) & todo;
        int base = 0;
        if (lane == leader) base = gatomic_add(cnt + (bl << r.tshift), __popcll(same));
        base = __shfl(base, leader);
        if (b == bl && !inwin) slot = base + __popcll(same & below);
        todo &= ~same;
    }
}

// Slot claims of one wave, with no barrier: lanes that share a bucket take consecutive
// slots from one atomic, issued by the lowest of them, and all of the wave's atomics
// go out as one instruction.  Returns the lane's slot (0 for a lane without a bead).
__device__ __forceinline__ int claim_slot_wave(const RepDev& r, bool valid, int b, int* cnt) {
    const int lane = (int)__lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(valid);
    int rank = 0, n = 0, lead = lane;
    while (todo) {   // wave-uniform: one round per distinct bucket of the wave
        const int leader = __ffsll((long long)todo) - 1;
        const int bl = __builtin_amdgcn_readlane(b, leader);
        const bool mine = valid && b == bl;
        const unsigned long long same = __ballot(mine);
        if (mine) { rank = __popcll(same & below); n = __popcll(same); lead = leader; }
        todo &= ~same;
    }
    int base = 0;
    if (valid && rank == 0) base = gatomic_add(cnt + (b << r.tshift), n);
    return __shfl(base, lead) + rank;
}

// Slot counters of one step parity.
__device__ __forceinline__ int* claim_counters(const RepDev& r, int parity) {
    return r.tclaim + (((size_t)parity * r.nbuckets) << r.tshift);
}

// A bead's record and charge into its bucket's region for the step of the given parity:
// its stencil record (stencil_record) where the replica deposits those, else rec =
// (x, y, z, type).  A bead whose region is full goes to the overflow list, which the
// deposit reads too, so the densities stay exact whatever the occupancy.  The region
// bounds are requested before the slot claim, so that their latency hides behind the
// atomic.  Every lane of the wave calls it (valid false for a lane without a bead).
__device__ __forceinline__ void claim_record(const RepDev& r, bool valid, float4 rec, float qi, int parity) {
    int b = 0x7fffffff, cap = 0, base = 0;
    if (valid) {
        if (r.srec) {
            uint4 w;
            b = stencil_record(r, rec, w);
            rec = make_float4(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z), __uint_as_float(w.w));
        } else {
            b = tile_bucket_yz(r, rec.y, rec.z);
        }
        cap = G((const int*)r.tcap)[b];
        base = G((const int*)r.tbase)[b];
    }
    const int slot = claim_slot_wave(r, valid, b, claim_counters(r, parity));
    if (!valid) return;
    if (slot < cap) {
        const int pos = base + slot;
        gst(r.trec + pos, rec);
        if (r.trecq) G(r.trecq)[pos] = qi;
    } else {
        const int o = gatomic_add(r.tovf + parity * 32, 1);
        gst(r.tovf_rec + o, rec);
        G(r.tovf_b)[o] = b;
        if (r.trecq) G(r.tovf_q)[o] = qi;
    }
}

}  // namespace tild
