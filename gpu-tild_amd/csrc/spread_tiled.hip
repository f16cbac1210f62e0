// spread_tiled.hip -- sorted-bin, LDS-tiled particle-to-grid deposition fused with
// the x-axis R2C transform (replaces the global atomics of
// device_grid_utils.cu:232-343 with per-tile LDS accumulation).
//
// The grid is cut into full-x slabs of By x Bz rows.  A bead's bucket is (tile of its
// first stencil row, halo class).  Every step:
//   k_tile_claim    each bead claims a slot in its bucket and writes its (x, y, z,
//                   type) record into the bucket's region (an overflow list takes
//                   the beads of full regions); fused into the particle kernel of
//                   the previous step
//   k_tile_deposit  one block per tile: deposits its own particles and the halo
//                   particles of its three lower neighbours into 64-bit
//                   fixed-point LDS accumulators, converts its rows to densities
//                   and transforms them along x (R2C) straight from LDS.
// The regions are sized from an exact binning (k_tile_count, k_tile_scan,
// k_tile_regions) at creation and at every spatial reorder.
// Integer accumulation makes the densities independent of bucket order, tile
// shape and thread scheduling: they are bit-identical to the global-atomic path.
//
// After the first step of a run the particle kernel makes the claims of the next step
// (from the positions it has just written), so k_tile_claim runs only where no claims
// are in the regions: at the start of a run after a spatial reorder, new positions or
// a new replica.  Bucket and halo classes are defined in tile_dev.hpp.
#include "device.hpp"
#include "fft_dev.hpp"
#include "launch.hpp"
#include "tile_dev.hpp"

namespace tild {

// Exact binning (spatial reorder and region sizing): (bucket, slot) of every bead.
__global__ void __launch_bounds__(256) k_tile_count(const RepDev* __restrict__ reps,
                                                    const WorkItem* __restrict__ wmap, int koff) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const float* x = r.X[(r.step + koff) & 1];
    const int i = wi.blk * 256 + threadIdx.x;
    const int b = i < r.ns ? tile_bucket(r, x, i) : 0x7fffffff;
    int slot;
    claim_slot(r, i < r.ns, b, r.tcnt, slot);
    if (i < r.ns) G((long long*)r.tkey)[i] = (long long)(((unsigned long long)(unsigned)slot << 32) | (unsigned)b);
}

// Per-step binning at the start of a run: each bead claims a slot in its bucket and
// writes its record (x, y, z, type) straight into the bucket's region -- no scan, no
// second pass.
__global__ void __launch_bounds__(256) k_tile_claim(const RepDev* __restrict__ reps,
                                                    const WorkItem* __restrict__ wmap, int koff) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int step = r.step + koff;
    const int i = wi.blk * 256 + threadIdx.x;
    const bool valid = i < r.ns;
    const int D = r.dim;
    gmem<const float> x = G((const float*)r.X[step & 1]);
    float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
    float qi = 0.f;
    if (valid) {
        rec = make_float4(x[i * D], x[i * D + 1], D == 3 ? x[i * D + 2] : 0.f, __int_as_float(G(r.tp)[i]));
        if (r.trecq) qi = G(r.q)[i];
    }
    claim_record(r, valid, rec, qi, step & 1);
}

// Clears the slot counters and overflow count of one step parity (after a force
// evaluation that does not advance the step: the next step claims from them).
__global__ void k_tile_reset(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap, int koff) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int p = (r.step + koff) & 1;
    int* cnt = claim_counters(r, p);
    for (int b = threadIdx.x; b < r.nbuckets; b += blockDim.x) G(cnt)[b << r.tshift] = 0;
    if (threadIdx.x == 0) G(r.tovf)[p * 32] = 0;
}

// Exclusive scan of the bucket counts; clears the counts for the next step.  Each
// thread holds up to kScanPer consecutive counts in registers (their loads go out
// together); the block scan is a wave-level shuffle scan plus a scan of the 16 wave
// totals.  Larger bucket counts loop over the same in chunks.
constexpr int kScanPer = 8;

__device__ __forceinline__ int block_exclusive_scan_1024(int v, int* wsum, int& total) {
    const int lane = (int)__lane_id(), w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o);
            if (lane >= o) t += y;
        }
        if (lane < 16) wsum[16 + lane] = t;   // inclusive wave prefix
    }
    __syncthreads();
    total = wsum[31];
    const int before = w > 0 ? wsum[16 + w - 1] : 0;
    __syncthreads();   // wsum is reused by the next call
    return before + x - v;
}

__global__ void __launch_bounds__(1024) k_tile_scan(const RepDev* __restrict__ reps,
                                                    const WorkItem* __restrict__ wmap) {
    __shared__ int wsum[32];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int nb = r.nbuckets, sh = r.tshift;
    gmem<int> cnt = G(r.tcnt);
    gmem<int> off = G(r.toff);
    int run = 0;   // buckets before this chunk
    for (int c0 = 0; c0 < nb; c0 += 1024 * kScanPer) {
        const int b0 = c0 + threadIdx.x * kScanPer;
        int c[kScanPer];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            c[k] = b0 + k < nb ? cnt[(b0 + k) << sh] : 0;
            s += c[k];
        }
        int total;
        int pre = run + block_exclusive_scan_1024(s, wsum, total);
#pragma unroll
        for (int k = 0; k < kScanPer; ++k)
            if (b0 + k < nb) {
                off[b0 + k] = pre;
                pre += c[k];
                cnt[(b0 + k) << sh] = 0;
            }
        run += total;
    }
    if (threadIdx.x == 0) off[nb] = run;
}

// Bucket regions from an exact binning (toff): each bucket gets its count plus spare
// slots (by default a quarter plus 16), so that beads can drift between buckets for
// the steps until the next sizing without overflowing; tbase is the exclusive scan of
// the sizes.
__device__ __forceinline__ void tile_regions_body(const RepDev& r) {
    __shared__ int wsum[32];
    const int nb = r.nbuckets;
    gmem<const int> off = G((const int*)r.toff);
    gmem<int> cap = G(r.tcap);
    gmem<int> base = G(r.tbase);
    int run = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024 * kScanPer) {
        const int b0 = c0 + threadIdx.x * kScanPer;
        int c[kScanPer];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const int n = b0 + k < nb ? off[b0 + k + 1] - off[b0 + k] : 0;
            const int spare = r.tslack == 0 ? n / 4 + 16 : (r.tslack > 0 ? r.tslack : 0);
            c[k] = b0 + k < nb ? n + spare : 0;
            s += c[k];
        }
        int total;
        int pre = run + block_exclusive_scan_1024(s, wsum, total);
#pragma unroll
        for (int k = 0; k < kScanPer; ++k)
            if (b0 + k < nb) {
                cap[b0 + k] = c[k];
                base[b0 + k] = pre;
                pre += c[k];
            }
        run += total;
    }
    if (threadIdx.x == 0) base[nb] = run;
}

__global__ void __launch_bounds__(1024) k_tile_regions(const RepDev* __restrict__ reps, int rep) {
    tile_regions_body(reps[rep]);
}
__global__ void __launch_bounds__(1024) k_tile_regions_set(const RepDev* __restrict__ reps,
                                                          const WorkItem* __restrict__ wmap) {
    tile_regions_body(reps[wmap[blockIdx.x].rep]);
}

// Deposit one record into the tile's fixed-point LDS rows [f][lz][ly][x]; stencil
// rows outside the tile belong to a neighbouring tile and are skipped.  NP = P + 1
// points per axis, unrolled for the default order.
template <int NP>
__device__ __forceinline__ void deposit_record(const RepDev& r, const float4 rec, float qi,
                                               unsigned long long* patch, int y0, int z0, int By, int Bz) {
    const int D = r.dim, Nx = r.Nx, T = r.ntypes;
    const int tp = __float_as_int(rec.w);
    float W[3][5];
    int n0[3];
    n0[0] = stencil_1d(rec.x, r.dx[0], r.idx[0], Nx, NP - 1, W[0]);
    n0[1] = stencil_1d(rec.y, r.dx[1], r.idx[1], r.Ny, NP - 1, W[1]);
    if (D == 3) n0[2] = stencil_1d(rec.z, r.dx[2], r.idx[2], r.Nz, NP - 1, W[2]);
    else { n0[2] = 0; W[2][0] = 1.f; }
    int xx[NP];
    xx[0] = n0[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) xx[k] = wrap_next(xx[k - 1], Nx);
    const int nz = D == 3 ? NP : 1;
    const bool charged = r.trecq != nullptr;
    uint32_t qw[3][NP];   // order 1: axis weights in fixed point (fixed_weight2/3, fixed_charge)
    float fq[3][NP];
    const float qs = qi * r.fixq_s1;
    if constexpr (NP == 2) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            qw[d][0] = d < D ? fixed_lower(W[d][0]) : kFixedOne;
            qw[d][1] = kFixedOne - qw[d][0];
            fq[d][0] = d < D ? (float)qw[d][0] : 1.f;
            fq[d][1] = (float)qw[d][1];
        }
    }
    int zz = n0[2];
#pragma unroll
    for (int iz = 0; iz < NP; ++iz, zz = wrap_next(zz, r.Nz)) {
        if (iz >= nz) break;
        const int lz = zz - z0;
        if (lz < 0 || lz >= Bz) continue;
        int yy = n0[1];
#pragma unroll
        for (int iy = 0; iy < NP; ++iy, yy = wrap_next(yy, r.Ny)) {
            const int ly = yy - y0;
            if (ly < 0 || ly >= By) continue;
            unsigned long long* rowp = patch + ((size_t)(tp * Bz + lz) * By + ly) * Nx;
            unsigned long long* rowq = patch + ((size_t)(T * Bz + lz) * By + ly) * Nx;
#pragma unroll
            for (int ix = 0; ix < NP; ++ix) {
                if constexpr (NP == 2) {
                    atomicAdd(rowp + xx[ix], D == 3 ? fixed_weight3(fixed_pair(qw[0][ix], qw[1][iy]), qw[2][iz], r.fix_s3)
                                                    : fixed_weight2(qw[0][ix], qw[1][iy], r.fix_s2));
                    if (charged) atomicAdd(rowq + xx[ix], fixed_charge(fq[0][ix], fq[1][iy], fq[2][iz], qs));
                } else {
                    const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                    atomicAdd(rowp + xx[ix], (unsigned long long)__float2ll_rn(w * r.fix_scale));
                    if (charged) atomicAdd(rowq + xx[ix], (unsigned long long)__float2ll_rn((w * qi) * r.fixq_scale));
                }
            }
        }
    }
}

// Block-uniform values of the stencil-record deposit, read once before the record loop
// (the compiler reloads fields of the replica after every LDS atomic otherwise).
struct SrecTile {
    int Nx, By, Bz, ntypes, sh;
    unsigned long long half;   // node weight (a b + half) >> sh: fixed_weight3 (3D), fixed_weight2 (2D)
    float qsc;                 // fixq_s1
    bool charged;
    __device__ __forceinline__ SrecTile(const RepDev& r, int D)
        : Nx(r.Nx), By(r.By), Bz(r.Bz), ntypes(r.ntypes), sh(D == 3 ? r.fix_s3 : r.fix_s2),
          half(sh > 0 ? 1ull << (sh - 1) : 0ull), qsc(r.fixq_s1), charged(r.trecq != nullptr) {}
};

// Deposit one order-1 stencil record (stencil_record).  The record's first stencil row
// sits at (my + dy, mz + dz) in this tile: dy, dz are 0 for the tile's own buckets and
// -By, -Bz for the halo buckets of its lower neighbours.  The same integer weights and
// charges as the position-record path and the global-atomic spread.
template <int D>
__device__ __forceinline__ void deposit_srec(const SrecTile& c, const float4 rec, float qi, unsigned long long* patch,
                                             int dy, int dz) {
    const uint32_t w0 = __float_as_uint(rec.x), w1 = __float_as_uint(rec.y);
    int xx[2];
    xx[0] = (int)(w0 & 0xffffu);
    xx[1] = wrap_next(xx[0], c.Nx);
    const int ly0 = (int)((w0 >> 16) & 0xffu) + dy, lz0 = (int)(w0 >> 24) + dz;
    const int tp = (int)(w1 >> 23);
    uint32_t qx[2], qy[2], qz[2];
    qx[0] = w1 & 0x7fffffu; qx[1] = kFixedOne - qx[0];
    qy[0] = __float_as_uint(rec.z); qy[1] = kFixedOne - qy[0];
    qz[0] = D == 3 ? __float_as_uint(rec.w) : kFixedOne; qz[1] = kFixedOne - qz[0];
    const float qs = qi * c.qsc;
    uint32_t qxy[2][2];   // [iy][ix]
    if constexpr (D == 3) {
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) qxy[iy][ix] = fixed_pair(qx[ix], qy[iy]);
    }
#pragma unroll
    for (int iz = 0; iz < (D == 3 ? 2 : 1); ++iz) {
        const int lz = lz0 + iz;
        if ((unsigned)lz >= (unsigned)c.Bz) continue;
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
            const int ly = ly0 + iy;
            if ((unsigned)ly >= (unsigned)c.By) continue;
            const uint32_t row = ((uint32_t)(tp * c.Bz + lz) * (uint32_t)c.By + (uint32_t)ly) * (uint32_t)c.Nx;
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) {
                const unsigned long long v = D == 3 ? fixed_node(qxy[iy][ix], qz[iz], c.half, c.sh)
                                                    : fixed_node(qx[ix], qy[iy], c.half, c.sh);
                atomicAdd(patch + row + xx[ix], v);
                if (c.charged) {
                    const uint32_t rowq = ((uint32_t)(c.ntypes * c.Bz + lz) * (uint32_t)c.By + (uint32_t)ly) * (uint32_t)c.Nx;
                    atomicAdd(patch + rowq + xx[ix],
                              fixed_charge((float)qx[ix], (float)qy[iy], D == 3 ? (float)qz[iz] : 1.f, qs));
                }
            }
        }
    }
}

__device__ __forceinline__ void deposit_record_any(const RepDev& r, const float4 rec, float qi,
                                                   unsigned long long* patch, int y0, int z0, int By, int Bz) {
    const int D = r.dim, P = r.P, Nx = r.Nx, T = r.ntypes;
    const int tp = __float_as_int(rec.w);
    float W[3][5];
    int n0[3];
    n0[0] = stencil_1d(rec.x, r.dx[0], r.idx[0], Nx, P, W[0]);
    n0[1] = stencil_1d(rec.y, r.dx[1], r.idx[1], r.Ny, P, W[1]);
    if (D == 3) n0[2] = stencil_1d(rec.z, r.dx[2], r.idx[2], r.Nz, P, W[2]);
    else { n0[2] = 0; W[2][0] = 1.f; }
    const int nz = D == 3 ? P + 1 : 1;
    for (int iz = 0, zz = n0[2]; iz < nz; ++iz, zz = wrap_next(zz, r.Nz)) {
        const int lz = zz - z0;
        if (lz < 0 || lz >= Bz) continue;
        for (int iy = 0, yy = n0[1]; iy <= P; ++iy, yy = wrap_next(yy, r.Ny)) {
            const int ly = yy - y0;
            if (ly < 0 || ly >= By) continue;
            unsigned long long* rowp = patch + ((size_t)(tp * Bz + lz) * By + ly) * Nx;
            unsigned long long* rowq = patch + ((size_t)(T * Bz + lz) * By + ly) * Nx;
            for (int ix = 0, xx = n0[0]; ix <= P; ++ix, xx = wrap_next(xx, Nx)) {
                const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                atomicAdd(rowp + xx, (unsigned long long)__float2ll_rn(w * r.fix_scale));
                if (r.trecq)
                    atomicAdd(rowq + xx, (unsigned long long)__float2ll_rn((w * qi) * r.fixq_scale));
            }
        }
    }
}

// Record j of a block's source regions taken as one index range: j + shift[k] for the
// last region k with start[k] <= j (start ascending, both block-uniform), summed from
// the steps dsh[s] = shift[s] - shift[s - 1] (a chain of selects between the shifts
// themselves is compiled into an indexed load from scratch memory).
__device__ __forceinline__ int region_pos(int j, const int (&start)[9], const int (&dsh)[9]) {
    int sh = dsh[0];
#pragma unroll
    for (int s = 1; s < 9; ++s) sh += j >= start[s] ? dsh[s] : 0;
    return j + sh;
}

// Loads records j0, j0 + NT, j0 + 2 NT, j0 + 3 NT of the block's index range (zeros past
// its end) and their charges.
template <int NT>
__device__ __forceinline__ void load_srec4(const RepDev& r, int j0, int total, const int (&start)[9],
                                           const int (&dsh)[9], bool charged, float4 (&rec)[4], float (&qi)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = j0 + k * NT;
        rec[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        qi[k] = 0.f;
        if (j < total) {
            const int p = region_pos(j, start, dsh);
            rec[k] = gld(r.trec + p);
            if (charged) qi[k] = G((const float*)r.trecq)[p];
        }
    }
}

// Deposit one tile's rows and transform them along x.  NT = 512 for tiles of about
// 1,000 beads plus their halo share (three blocks of 512 fit a CU; on the 640,000-bead
// melt 17.5 us, against 19.6 with 256 threads and 22.8 with 1,024, where one block
// fills a CU); NT = 256 for small tiles, where a larger block mostly idles.
template <int kDepositThreads>
__global__ void __launch_bounds__(kDepositThreads) k_tile_deposit(const RepDev* __restrict__ reps,
                                                              const WorkItem* __restrict__ wmap, int koff) {
    extern __shared__ unsigned long long tile_lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int D = r.dim, P = r.P, Nx = r.Nx, Nxh = r.Nxh;
    const int By = r.By, Bz = r.Bz, nf = r.nf;
    const int t = wi.blk, ty = t % r.nty, tz = t / r.nty;
    const int y0 = ty * By, z0 = tz * Bz;
    const int rows = nf * By * Bz;
    const int par = (r.step + koff) & 1;
    unsigned long long* patch = tile_lds;                         // [f][lz][ly][x]
    float2* fbuf = (float2*)(tile_lds + (size_t)rows * Nx);
    // source buckets: own tile (all classes) and the lower neighbours' halo classes;
    // each bucket's records are the filled part of its region.  Their counts and
    // bounds are requested first: the LDS rows are cleared while they are in flight.
    const int tym = (ty + r.nty - 1) % r.nty;
    const int tzm = (tz + r.ntz - 1) % r.ntz;
    const int own = tz * r.nty + ty, ny_ = tz * r.nty + tym, nz_ = tzm * r.nty + ty, nyz = tzm * r.nty + tym;
    int bk[9], nbk = 0;   // 4 own + 2 (y) + 2 (z) + 1 (yz)
    for (int c = 0; c < 4; ++c) bk[nbk++] = own * 4 + c;
    bk[nbk++] = ny_ * 4 + 1; bk[nbk++] = ny_ * 4 + 2;
    if (D == 3) { bk[nbk++] = nz_ * 4 + 2; bk[nbk++] = nz_ * 4 + 3; bk[nbk++] = nyz * 4 + 2; }
    gmem<const int> cnt = G((const int*)claim_counters(r, par));
    gmem<const int> base = G((const int*)r.tbase);
    gmem<const int> cap = G((const int*)r.tcap);
    int vcnt[9], vcap[9], vbase[9];
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        const int bb = s < nbk ? bk[s] : 0;
        vcnt[s] = cnt[bb << r.tshift];
        vcap[s] = cap[bb];
        vbase[s] = base[bb];
    }
    const int novf = G((const int*)r.tovf)[par * 32];
    // the other parity's counters were last used by the previous step: clear this
    // tile's for the next one
    if (threadIdx.x < 4) G(claim_counters(r, par ^ 1))[(t * 4 + threadIdx.x) << r.tshift] = 0;
    if (t == 0 && threadIdx.x == 0) G(r.tovf)[(par ^ 1) * 32] = 0;
    for (int k = threadIdx.x; k < rows * Nx; k += kDepositThreads) patch[k] = 0ull;
    const int T = r.ntypes;
    const int Ny = r.Ny, M = r.M, plane_fwd = r.plane_fwd;
    // where the plane pass transforms the densities the FFT scratch is free: it holds
    // the first density element of each tile row (f, lz, ly)
    int* const rowoff = (int*)fbuf;
    if (plane_fwd)
        for (int rr = threadIdx.x; rr < rows; rr += kDepositThreads) {
            const int ly = rr % By, lz = (rr / By) % Bz, f = rr / (By * Bz);
            rowoff[rr] = f * M + ((z0 + lz) * Ny + (y0 + ly)) * Nx;
        }
    // the filled parts of the source regions as one index range: record j of the block
    // is j + shift[k] for the last region k with start[k] <= j (all threads busy, no
    // partial sweep per region)
    int start[9], dsh[9], total = 0, prev = 0;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        const int n = s < nbk ? min(vcnt[s], vcap[s]) : 0;
        const int shift = vbase[s] - total;
        // block-uniform: kept in scalar registers
        start[s] = __builtin_amdgcn_readfirstlane(total);
        dsh[s] = __builtin_amdgcn_readfirstlane(shift - prev);
        prev = shift;
        total += n;
    }
    if (r.srec) {
        // stencil records: the region decides the row offsets (own 0-3, y neighbour
        // 4-5, z neighbour 6-7, yz neighbour 8); four records per thread are in flight
        // at a time, the first four across the barrier that ends the clearing
        const SrecTile c(r, D);
        float4 rec[4];
        float qi[4];
        load_srec4<kDepositThreads>(r, threadIdx.x, total, start, dsh, c.charged, rec, qi);
        __syncthreads();
        for (int j0 = threadIdx.x; j0 < total; j0 += 4 * kDepositThreads) {
            if (j0 > (int)threadIdx.x) load_srec4<kDepositThreads>(r, j0, total, start, dsh, c.charged, rec, qi);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int j = j0 + k * kDepositThreads;
                if (j >= total) break;
                const bool hy = (j >= start[4] && j < start[6]) || j >= start[8];
                const int dy = hy ? -By : 0, dz = j >= start[6] ? -Bz : 0;
                if (D == 3) deposit_srec<3>(c, rec[k], qi[k], patch, dy, dz);
                else deposit_srec<2>(c, rec[k], qi[k], patch, dy, 0);
            }
        }
    } else {
        __syncthreads();
        for (int j = threadIdx.x; j < total; j += kDepositThreads) {
            const int p = region_pos(j, start, dsh);
            const float4 rec = gld(r.trec + p);
            const float qi = r.trecq ? G((const float*)r.trecq)[p] : 0.f;
            if (P == 1) deposit_record<2>(r, rec, qi, patch, y0, z0, By, Bz);
            else deposit_record_any(r, rec, qi, patch, y0, z0, By, Bz);
        }
    }
    // beads whose region was full: the same bucket test over the overflow list
    for (int j = threadIdx.x; j < novf; j += kDepositThreads) {
        const int bb = G((const int*)r.tovf_b)[j];
        int src = -1;
#pragma unroll
        for (int s = 0; s < 9; ++s) src = s < nbk && bb == bk[s] ? s : src;
        if (src < 0) continue;
        const float4 rec = gld(r.tovf_rec + j);
        const float qi = r.trecq ? G((const float*)r.tovf_q)[j] : 0.f;
        if (r.srec) {
            const int dy = (src >= 4 && src < 6) || src == 8 ? -By : 0, dz = src >= 6 ? -Bz : 0;
            if (D == 3) deposit_srec<3>(SrecTile(r, 3), rec, qi, patch, dy, dz);
            else deposit_srec<2>(SrecTile(r, 2), rec, qi, patch, dy, 0);
        } else if (P == 1) {
            deposit_record<2>(r, rec, qi, patch, y0, z0, By, Bz);
        } else {
            deposit_record_any(r, rec, qi, patch, y0, z0, By, Bz);
        }
    }
    __syncthreads();
    // densities of the tile rows -> rho and the x-axis R2C (pairs of rows)
    const int npairs = (rows + 1) / 2;
    const int ldb = npairs | 1;
    const FastDiv dNx(Nx), dNxh(Nxh), dBy(By), dBz(Bz);
    // block-uniform values read once: a lane-dependent choice between two fields of the
    // replica is compiled into a vector load of the replica record, once per row
    const double inv_t = r.fix_inv, inv_q = r.fixq_inv;
    const int off32 = r.grid_off32;
    float* const rho = r.rho;
    const int trows = T * By * Bz;   // rows below are particle densities, the rest charge
    if (plane_fwd && off32 && (Nx & 1) == 0) {
        // densities only, two x nodes per task: (row, node pair) digits stepped, the
        // row's first element from the table
        const int np = Nx >> 1;
        const FastDiv dNp(np);
        Digits2 gq(threadIdx.x, kDepositThreads, dNp);
        for (int task = threadIdx.x; task < rows * np; task += kDepositThreads, gq.step()) {
            const int q = gq.lo, rr = gq.hi;
            const double inv = rr < trows ? inv_t : inv_q;
            const ulonglong2 a = *(const ulonglong2*)(patch + (size_t)rr * Nx + 2 * q);
            const float2 v = make_float2((float)((double)(long long)a.x * inv), (float)((double)(long long)a.y * inv));
            gst(at32((float2*)rho, (uint32_t)(rowoff[rr] >> 1) + (uint32_t)q), v);
        }
        return;
    }
    // task = j * Nx + n; rows 2j and 2j + 1 are (f, lz, ly) with ly fastest
    Digits2 gt(threadIdx.x, kDepositThreads, dNx);
    for (int task = threadIdx.x; task < npairs * Nx; task += kDepositThreads, gt.step()) {
        const int n = gt.lo, j = gt.hi;
        float v[2] = {0.f, 0.f};
        int ly, lz;
        int f = dBz.div(dBy.div(2 * j, ly), lz);
        for (int h = 0; h < 2; ++h) {
            const int rr = 2 * j + h;
            if (rr >= rows) break;
            if (h) {   // the next row: ly + 1 with carries
                ++ly;
                if (ly == By) { ly = 0; if (++lz == Bz) { lz = 0; ++f; } }
            }
            const double inv = f < T ? inv_t : inv_q;
            v[h] = (float)((double)(long long)patch[(size_t)rr * Nx + n] * inv);
            const int node = ((z0 + lz) * Ny + (y0 + ly)) * Nx + n;
            if (off32) G(at32(rho, (uint32_t)f * (uint32_t)M + (uint32_t)node))[0] = v[h];
            else G(rho)[(size_t)f * M + node] = v[h];
        }
        if (!plane_fwd) fbuf[n * ldb + j] = make_float2(v[0], v[1]);
    }
    if (plane_fwd) return;   // the plane pass transforms the densities
    __syncthreads();
    const float2* res = lds_fft<kDepositThreads>(fbuf, fbuf + (size_t)Nx * ldb, ldb, npairs, r.ax[0]);
    for (int task = threadIdx.x; task < npairs * Nxh; task += kDepositThreads) {
        int k;
        const int j = dNxh.div(task, k);
        const float2 Zk = res[k * ldb + j];
        const float2 Zm = res[(k == 0 ? 0 : Nx - k) * ldb + j];
        for (int h = 0; h < 2; ++h) {
            const int rr = 2 * j + h;
            if (rr >= rows) break;
            int ly, lz;
            const int f = dBz.div(dBy.div(rr, ly), lz);
            const size_t row = ((size_t)f * r.Nz + (z0 + lz)) * r.Ny + (y0 + ly);
            gst(r.H + row * Nxh + k, h == 0 ? make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y))
                                            : make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x)));
        }
    }
}

#define CHECK_LAUNCH() do { hipError_t e__ = hipGetLastError(); if (e__ != hipSuccess) return e__; } while (0)

hipError_t launch_tile_count(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    hipLaunchKernelGGL(k_tile_count, dim3(c.n), dim3(256), 0, c.stream, c.reps, c.wmap, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_scan(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    hipLaunchKernelGGL(k_tile_scan, dim3(c.n), dim3(1024), 0, c.stream, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_claim(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_tile_claim, dim3(c.n), dim3(256), 0, c.reps, c.wmap, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_reset(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    hipLaunchKernelGGL(k_tile_reset, dim3(c.n), dim3(256), 0, c.stream, c.reps, c.wmap, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_regions(hipStream_t s, const RepDev* reps, int rep) {
    hipLaunchKernelGGL(k_tile_regions, dim3(1), dim3(1024), 0, s, reps, rep);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_regions_set(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    hipLaunchKernelGGL(k_tile_regions_set, dim3(c.n), dim3(1024), 0, c.stream, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_tile_deposit(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    if (c.threads >= 512) TILD_LAUNCH(c, k_tile_deposit<512>, dim3(c.n), dim3(512), c.lds, c.reps, c.wmap, koff);
    else TILD_LAUNCH(c, k_tile_deposit<256>, dim3(c.n), dim3(256), c.lds, c.reps, c.wmap, koff);
    CHECK_LAUNCH(); return hipSuccess;
}

}  // namespace tild
