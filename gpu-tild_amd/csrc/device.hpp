// device.hpp -- device-side descriptors and helpers shared by the TILD kernels.
//
// Every kernel is launched over the work of one or more replicas: block b of a
// launch looks up (replica, local block) in a per-phase map, and reads that
// replica's geometry, plans and buffers from a RepDev descriptor in device memory.
// A single-replica step is the same launches with a one-entry map.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tild {

constexpr int kMaxTypes = 16;      // particle types per replica
constexpr int kMaxStages = 12;     // FFT stages per axis
constexpr int kFftThreads = 256;   // threads of every FFT block
constexpr int kFftElems = 8;       // complex values per thread per FFT call (LDS budget)
constexpr int kMaxRamp = 8;        // ramped pair styles per replica
constexpr int kMaxQuant = 128;     // property quantities per replica

// Mixed-radix Stockham plan of one axis.  Stage s has radix rad[s] and
// p_s = prod(rad[0..s-1]); its twiddles w_{p R}^{r k} (forward sign, k < p, 1 <= r < R)
// start at tw + tw_off[s], index k*(R-1) + (r-1).
struct AxisPlan {
    int N;
    int nst;
    int rad[kMaxStages];
    int tw_off[kMaxStages];
    int ntw;            // twiddles in the table (at most N)
    const float2* tw;
};
// room for the twiddles of an N-point plan (ntw <= tw_elems(N))
__host__ __device__ inline int tw_elems(int N) { return N > 1 ? N : 1; }

// Per-replica descriptor (device memory).
struct RepDev {
    // geometry (initialize.cu:32-48)
    int dim, Nx, Ny, Nz, Nxh, M, Mh, ns, ntypes, nf, P;
    float L[3], Lh[3], dx[3], gvol, kfac[3];   // kfac = 2 pi / L
    float idx[3];                              // 1 / dx (FP32), for the spline weights
    // fixed-point densities: acc = round(w * 2^F), rho = acc * inv
    float fix_scale, fixq_scale;
    double fix_inv, fixq_inv;
    int fix_s2, fix_s3;        // order-1 integer weights: 44 - F and 52 - F (fixed_weight2/3)
    float fixq_s1;             // order-1 charge weights: fixq_scale 2^-22 per axis (fixed_charge)
    // FFT plans (x, y, z) and chunk sizes of each pass
    AxisPlan ax[3];
    AxisPlan axr;  // inverse plan of the last axis: a two-stage plan reversed, else ax[dim - 1]
    int nl_x;      // line pairs per block, forward x pass
    int nl_mid;    // batch columns per block, middle (y) pass
    int nl_last;   // batch columns per block, fused last-axis pass
    int nr_c2r;    // rows per block, inverse x pass
    int mid_elems; // complex values per block of the middle-axis passes (0: kFftBudget)
    int plane_fwd; // 1: the tiled deposit writes rho only; x R2C + y FFT run per plane
    int last_dirs; // directions per inverse batch of the fused last-axis pass (1..dim)
    int last_reg;  // 1: the fused last-axis pass runs register-resident (k_last_reg)
    int grad_reg;  // 1: the y inverse of the gradient spectra runs register-resident (k_grad_reg)
    int plane_reg; // 1: the forward plane pass runs register-resident (k_plane_reg)
    int plane_inv; // 1: y inverse and x C2R run as one register-resident plane pass (k_plane_inv)
    // particle state
    float* X[2];            // ns*dim, double-buffered by step parity
    float* XO;              // ns*dim, GJF previous position
    const int* tp;          // ns
    const float* q;         // ns or null
    const int* bond;        // ns*3 slots: (partner << bond_shift) | bond type; -1 = empty
    int bond_shift;         // bits of the bond type (0 for one type)
    const float* bond_k;    // per bond type
    const float* bond_r0;
    const signed char* ikind;  // ns: 0 none, 1 EM, 2 GJF
    float delt, noise_mag;
    int any_gjf;            // 1: some site integrates with GJF
    float gjf_b[kMaxTypes], gjf_a[kMaxTypes], gjf_c1[kMaxTypes], gjf_c2[kMaxTypes],
          gjf_noise[kMaxTypes];
    float* fout;            // ns*dim (optional)
    int* pid;               // ns: original site id of each slot (noise key, host order)
    int *rsrc, *rorder, *rinv;   // ns each: scratch of the tiled spatial reorder
    // spare buffers of the periodic spatial reorder, swapped with the live arrays
    float* XO_alt;
    int* tp_alt;
    float* q_alt;
    int* bond_alt;
    signed char* ikind_alt;
    int* pid_alt;
    // grids
    unsigned long long* acc;   // nf*M fixed-point accumulators
    float* rho;                // nf*M densities (x fastest)
    float2* H;                 // nf*Mh half spectra
    float2* Gk;                // nf*dim*Mh half spectra of the gathered fields
    float* G3;                 // nf*M*3 gather records (g_x, g_y, g_z), zero where rho <= 0
    int g3_off32;              // 1: nf*M*12 bytes of records fit 32-bit byte offsets
    int grid_off32;            // 1: every grid buffer (rho, H, Gk, G3) fits 32-bit byte offsets
    const float* Ctab;         // coupling tables C_IJ/M (I <= J) + charge table, blocked as
                               // [column block][last-axis index][row][nl_last columns]
    int ctab_rows;             // rows of Ctab: T(T+1)/2 (+1 with charges)
    const float* FPtab;        // T*dim*M net static field-phase forces per type (or null)
    int ctab_charge;           // row of the charge kernel in Ctab (or -1)
    // ramped pair styles (pair_style.cu:308-328): Ctab row ramp_row0 + j holds ramped
    // style j's table per unit prefactor, added to pair row ramp_pair[j] with weight
    // A_j(k) - ramp_A0[j] at 0-based step k, where A_j(0) = ramp_A0[j] and
    // A_j(k) = ramp_c0[j] + ramp_c1[j] * (k + 1) after the first step
    int nramp, ramp_row0;
    int ramp_pair[kMaxRamp];
    double ramp_A0[kMaxRamp], ramp_c0[kMaxRamp], ramp_c1[kMaxRamp];
    // property tables
    int nstyles, npairs, nq;
    int st_t1[64], st_t2[64];
    int pr_t1[32], pr_t2[32];
    const float* Etab;         // nstyles*Mh : u_s(k) * gvol / M
    const float2* Vtab;        // npairs*nP*Mh : vir_k * gvol / (V D M)
    const float* Eltab;        // Mh : charge potential / M
    double* partials;          // nq * nchunk partial sums
    int nchunk;
    double* bpart;             // 7 * ceil(ns/256) bond-property partials
    // sorted-bin deposition: full-x slab tiles of By x Bz rows (By | Ny, Bz | Nz)
    int spread_mode;           // 1 global atomics, 2 LDS-tiled sorted bins
    int By, Bz, nty, ntz, ntiles, nbuckets;
    float iBy, iBz;            // 1 / By, 1 / Bz (FP32) for the bucket divisions
    int tshift;                // log2 of the counter stride: one 128-byte line per bucket
    // exact binning (spatial reorder, capacity sizing): counts, offsets, (bucket, slot)
    int* tcnt;                 // nbuckets counters (zero between uses), bucket b at b << tshift
    int* toff;                 // nbuckets + 1 exclusive offsets
    int2* tkey;                // ns (bucket, slot)
    // per-step binning into fixed bucket regions (sized at creation and each reorder)
    int* tclaim;               // 2 x (nbuckets << tshift) slot counters, by step parity
    int* tbase;                // nbuckets + 1: first record of each bucket's region
    int* tcap;                 // nbuckets: region sizes
    int tslack;                // spare slots per region (0: count/4 + 16, < 0: none)
    int srec;                  // 1: order-1 stencil records (stencil_record), 0: (x, y, z, type)
    float4* trec;              // records in bucket regions (16 bytes, either kind)
    float* trecq;              // their charges (charged decks)
    int* tovf;                 // overflow counts by step parity, at [0] and [32]
    float4* tovf_rec;          // ns records that found their region full
    int* tovf_b;               // their buckets
    float* tovf_q;             // their charges
    // counters
    int step;                  // completed steps; X[step & 1] is current
    uint32_t seed;
    int nP;
};

struct WorkItem { int rep, blk; };

// Global-memory view of a pointer read from a descriptor.  Such pointers are generic
// to the compiler, which then emits flat_* accesses that also wait on LDS traffic;
// with the address space spelled out it emits global_* instructions.
template <class T> using gmem = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ gmem<T> G(T* p) { return (gmem<T>)p; }
// Device-scope atomic add on global memory (a global_atomic instruction; through a
// generic pointer it is a flat atomic, which also waits on LDS traffic).
__device__ __forceinline__ int gatomic_add(int* p, int v) {
    return __hip_atomic_fetch_add(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// float2 / float4 are classes and cannot be copied out of an address-space-qualified
// lvalue; these go through the compiler's native vector types instead.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 gld(const float2* p) {
    const f2v v = *(gmem<const f2v>)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 gld(const float4* p) {
    const f4v v = *(gmem<const f4v>)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(float2* p, float2 a) { *(gmem<f2v>)p = f2v{a.x, a.y}; }
// three consecutive floats (one gather record); merged into one 12-byte access
__device__ __forceinline__ float3 gld3(const float* p) {
    gmem<const float> q = G(p);
    return make_float3(q[0], q[1], q[2]);
}
__device__ __forceinline__ int3 gld3(const int* p) {
    gmem<const int> q = G(p);
    return make_int3(q[0], q[1], q[2]);
}
__device__ __forceinline__ void gst3(float* p, float a, float b, float c) {
    gmem<float> q = G(p);
    q[0] = a; q[1] = b; q[2] = c;
}
__device__ __forceinline__ void gst(float4* p, float4 a) { *(gmem<f4v>)p = f4v{a.x, a.y, a.z, a.w}; }

// Weight of ramped style j's coupling row at 0-based step k (pair_style.cu:308-328, see
// RepDev::ramp_*): A_j(k) - A_j(0).
__device__ __forceinline__ float ramp_weight(const RepDev& r, int j, int k) {
    const double A = k == 0 ? r.ramp_A0[j] : r.ramp_c0[j] + r.ramp_c1[j] * (double)(k + 1);
    return (float)(A - r.ramp_A0[j]);
}

// ---------------------------------------------------------------- Philox / normals
// Philox4x32-10 (Salmon et al., SC'11).
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32 x 32 -> 64-bit multiply (v_mad_u64_u32) gives both halves of each product
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

// Four standard normals for (seed, site, step): Box-Muller on (k + 1/2) / 2^23
// uniforms, which are exact in FP32.  The transcendentals are the hardware ones:
// v_log_f32 (log2), v_sqrt_f32 and v_sin/v_cos_f32, whose argument is in revolutions,
// i.e. exactly u2 for the angle 2 pi u2 -- a few instructions instead of the library
// routines, with about 1e-6 absolute error on the normals.
__device__ inline void normals4(uint32_t seed, uint32_t site, uint32_t step, float z[4]) {
    uint32_t c[4] = {site, step, 0u, 0u};
    philox4x32_10(c, seed, 0u);
    const float ln2 = 0.69314718055994530942f;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const float u1 = ((float)(c[2 * p] >> 9) + 0.5f) * (1.0f / 8388608.0f);
        const float u2 = ((float)(c[2 * p + 1] >> 9) + 0.5f) * (1.0f / 8388608.0f);
        const float r = __builtin_amdgcn_sqrtf(-2.0f * ln2 * __builtin_amdgcn_logf(u1));
        z[2 * p] = r * __builtin_amdgcn_cosf(u2);
        z[2 * p + 1] = r * __builtin_amdgcn_sinf(u2);
    }
}

// ---------------------------------------------------------------- stencil
// PME B-spline weights of orders 0-4 (device_grid_utils.cu:3-49), including the
// order-4 term sx3 = sx2*sx2 of the reference.
__device__ __forceinline__ void spline_weights(float gdx, float iH, float* W, int P) {
#pragma clang fp contract(off)
    float sx = gdx * iH, sx2, sx3, sx4;
    if (P == 0) { W[0] = 1.f; }
    else if (P == 1) { W[0] = 0.5f - sx; W[1] = 0.5f + sx; }
    else if (P == 2) {
        sx2 = sx * sx;
        W[0] = 0.125f - 0.5f * sx + 0.5f * sx2; W[1] = 0.75f - sx2; W[2] = 0.125f + 0.5f * sx + 0.5f * sx2;
    } else if (P == 3) {
        sx2 = sx * sx; sx3 = sx2 * sx;
        W[0] = (1.f - 6.f * sx + 12.f * sx2 - 8.f * sx3) / 48.f;
        W[1] = (23.f - 30.f * sx - 12.f * sx2 + 24.f * sx3) / 48.f;
        W[2] = (23.f + 30.f * sx - 12.f * sx2 - 24.f * sx3) / 48.f;
        W[3] = (1.f + 6.f * sx + 12.f * sx2 + 8.f * sx3) / 48.f;
    } else {
        sx2 = sx * sx; sx3 = sx2 * sx2; sx4 = sx2 * sx2;
        W[0] = (1.f - 8.f * sx + 24.f * sx2 - 32.f * sx3 + 16.f * sx4) / 384.f;
        W[1] = (19.f - 44.f * sx + 24.f * sx2 + 16.f * sx3 - 16.f * sx4) / 96.f;
        W[2] = (115.f - 120.f * sx2 + 48.f * sx4) / 192.f;
        W[3] = (19.f + 44.f * sx + 24.f * sx2 - 16.f * sx3 - 16.f * sx4) / 96.f;
        W[4] = (1.f + 8.f * sx + 24.f * sx2 + 32.f * sx3 + 16.f * sx4) / 384.f;
    }
}

// First stencil node n0 in [0, N) of a bin g.  Positions stay within [0, L], so
// g - P/2 lies in [-N, 2N) and one add or subtract wraps it; the integer remainder
// (a long division sequence on the GPU) is only the fallback for anything else.
__device__ __forceinline__ int wrap_node(int n, int N) {
    if ((unsigned)n < (unsigned)N) return n;
    if (n < 0 && n >= -N) return n + N;
    if (n >= N && n < 2 * N) return n - N;
    n %= N;
    return n < 0 ? n + N : n;
}

// n / d and n % d for 0 <= n < 2^22, d >= 1, from inv = 1/d in FP32: the float
// quotient is within one of the true quotient, and one correction step fixes it.  The
// product q*d stays below 2^24, so the 24-bit multiply (full rate, where the 32-bit one
// is quarter rate) is exact; the correction is selects, not branches.
__device__ __forceinline__ int div_small(int n, int d, float inv, int& rem) {
    int q = (int)((float)n * inv);
    const int m = n - (int)__umul24((unsigned)q, (unsigned)d);
    const bool lo = m < 0, hi = m >= d;
    q += (int)hi - (int)lo;
    rem = lo ? m + d : (hi ? m - d : m);
    return q;
}

// A divisor of the in-block index arithmetic: n / d and n % d for 0 <= n < 2^22 from
// the hardware reciprocal (1 ulp, so the float quotient is still within one of the
// true one), in a few instructions instead of the integer division sequence.
struct FastDiv {
    int d;
    float inv;
    __device__ __forceinline__ explicit FastDiv(int dd) : d(dd), inv(__builtin_amdgcn_rcpf((float)dd)) {}
    __device__ __forceinline__ int div(int n, int& rem) const { return div_small(n, d, inv, rem); }
};

// Element i of a buffer addressed as its base plus a 32-bit byte offset (buffers below
// 4 GiB): the uniform base stays in scalar registers and each lane carries one 32-bit
// offset, where a 64-bit index costs 64-bit multiply-adds per element.
template <class T>
__device__ __forceinline__ T* at32(T* base, uint32_t i) {
    return (T*)((char*)base + i * (uint32_t)sizeof(T));
}

// Mixed-radix digits of a strided loop index: i = (hi * A + mid) * B + lo for
// i = t, t + S, t + 2S, ...  One division pair at set-up; each step adds the digits of
// S with carries (adds and selects), where dividing i anew costs two FastDivs.
// Requires A, B >= 1; hi may grow past any bound (the loop tests it).
struct Digits3 {
    int lo, mid, hi;
    int sl, sm, sh, A, B;
    __device__ __forceinline__ Digits3(int t, int S, const FastDiv& dB, const FastDiv& dA) : A(dA.d), B(dB.d) {
        hi = dA.div(dB.div(t, lo), mid);
        sh = dA.div(dB.div(S, sl), sm);
    }
    __device__ __forceinline__ void step() {
        lo += sl;
        const bool c0 = lo >= B;
        lo = c0 ? lo - B : lo;
        mid += sm + (int)c0;
        const bool c1 = mid >= A;
        mid = c1 ? mid - A : mid;
        hi += sh + (int)c1;
    }
};

// Two-digit form of the same: i = hi * B + lo.
struct Digits2 {
    int lo, hi;
    int sl, sh, B;
    __device__ __forceinline__ Digits2(int t, int S, const FastDiv& dB) : B(dB.d) {
        hi = dB.div(t, lo);
        sh = dB.div(S, sl);
    }
    __device__ __forceinline__ void step() {
        lo += sl;
        const bool c = lo >= B;
        lo = c ? lo - B : lo;
        hi += sh + (int)c;
    }
};

// Bin of one coordinate (device_grid_utils.cu:136-146).  The bin uses IEEE FP32
// division exactly as the reference, so bins match bit for bit.
__device__ __forceinline__ int stencil_bin(float x, float dx, int P) {
    return (P & 1) == 0 ? (int)((x + 0.5f * dx) / dx) : (int)(x / dx);
}

// First stencil node (wrapped into [0, N)) and weights of one coordinate.  The
// weights take the offset times 1/dx: at most an ulp from the reference's quotient,
// and a multiply instead of a division sequence.  Contraction into FMAs is off here
// and in the weights, so that every kernel that inlines them (the deposits of both
// spread paths, the gather) rounds the weights the same way whatever surrounds them.
__device__ __forceinline__ int stencil_1d(float x, float dx, float idx, int N, int P, float* W) {
#pragma clang fp contract(off)
    const int g = stencil_bin(x, dx, P);
    const float gdx = (P & 1) == 0 ? x - (float)g * dx : x - ((float)g + 0.5f) * dx;
    spline_weights(gdx, idx, W, P);
    return wrap_node(g - P / 2, N);
}

// First stencil node only (bucket keys, bins): no weights.
__device__ __forceinline__ int stencil_node(float x, float dx, int N, int P) {
    return wrap_node(stencil_bin(x, dx, P) - P / 2, N);
}

__device__ inline int wrap_next(int n, int N) { return (n + 1 == N) ? 0 : n + 1; }

// ---------------------------------------------------------------- fixed-point weights
// Order-1 deposition in integers: the lower weight W0 = 0.5 - sx of each axis is rounded
// to q0 = W0 2^22 (fixed_lower) and the upper weight is its complement 2^22 - q0, so the
// pair sums to one exactly.  A node's weight Wx Wy Wz 2^F is formed from 32 x 32 -> 64-bit
// products: (qx qy + 2^13) >> 14 (at most 2^30 + 2^9, scale 2^30), times qz plus the
// rounding half, shifted right by 52 - F.  Two 64-bit multiply-adds and two shifts per
// node, where rounding the FP32 product Wx Wy Wz 2^F to a 64-bit integer takes a dozen
// instructions.  The integer weight is within about 2^-22 of the exact product (FP32
// rounds it within 2^-23 relative).  Both spread paths call these, so their densities
// stay bit-identical.
constexpr uint32_t kFixedOne = 1u << 22;
__device__ __forceinline__ uint32_t fixed_lower(float w0) {
    // round to nearest; a negative roundoff of 0.5 - sx saturates to 0, one above 1
    // is clamped to one
    return min(__float2uint_rn(w0 * 4194304.0f), kFixedOne);
}
__device__ __forceinline__ uint32_t fixed_pair(uint32_t qx, uint32_t qy) {
    return (uint32_t)(((uint64_t)qx * qy + (1ull << 13)) >> 14);
}
__device__ __forceinline__ unsigned long long fixed_node(uint32_t a, uint32_t b, unsigned long long half, int sh) {
    return ((uint64_t)a * b + half) >> sh;
}
__device__ __forceinline__ unsigned long long fixed_weight3(uint32_t qxy, uint32_t qz, int s3) {
    return fixed_node(qxy, qz, 1ull << (s3 - 1), s3);
}
__device__ __forceinline__ unsigned long long fixed_weight2(uint32_t qx, uint32_t qy, int s2) {
    return fixed_node(qx, qy, s2 > 0 ? 1ull << (s2 - 1) : 0ull, s2);
}
// Charge of an order-1 node: the FP32 product of the quantised axis weights (qz = 1 in
// 2D) times qs = charge * fixq_s1, rounded to a 64-bit integer.
__device__ __forceinline__ unsigned long long fixed_charge(float fqx, float fqy, float fqz, float qs) {
#pragma clang fp contract(off)
    return (unsigned long long)__float2ll_rn(((fqx * fqy) * fqz) * qs);
}

}  // namespace tild
