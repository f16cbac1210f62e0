// last_reg.hip -- fused last-axis pass with register-resident columns (CDNA4 / gfx950).
//
// The same operation as k_last_fused (kernels.hip) -- last-axis forward FFT of every
// field, the k-space coupling w_I = sum_J C_IJ rho_J (+ the charge row), and the
// inverse transforms of w and -i k_la w -- for last axes with a two-stage plan
// N = R1 R2 (R1 >= R2: 160 = 16 x 10, 80 = 10 x 8, 64 = 8 x 8, 40 = 8 x 5).  Each
// thread keeps its Stockham columns in registers between the stages, so a block makes
// two LDS exchanges and two barriers, where the LDS-staged form stages every value
// through LDS and ends each of its nine phases with a barrier:
//
//   A  thread (b, i), i < R2: rows i + R2 r (r < R1) of column b of every field, read
//      from HBM straight into registers; DFT_R1; exchange.
//   B  thread (b, i), i < R1: twiddles w_N^{i r}, DFT_R2 -> rho_f(k) at k = i + R1 q;
//      the coupling in registers; the inverse lines conj(w) and conj(-i k w) are
//      transformed by the Stockham plan R2 x R1, whose first stage (radix R2, no
//      twiddles) reads exactly the values k = i + R1 q this thread holds; exchange.
//   C  thread (b, j), j < R2: twiddles w_N^{j r}, DFT_R1 -> rows j + R2 q; store.
//
// The transforms and sums are those of the LDS-staged pass (whose inverse also runs the
// reversed plan RepDev::axr); the two agree to rounding, since the compiler contracts
// multiplies into the following additions differently where one keeps in registers
// what the other stages through LDS.  The result does not depend on the number of
// columns per block.
#include "device.hpp"
#include "fft_dev.hpp"
#include "launch.hpp"

namespace tild {

namespace {

template <class T>
__device__ __forceinline__ T* grid_at(T* base, bool o32, size_t i) {
    return o32 ? at32(base, (uint32_t)i) : base + i;
}

// Products that the LDS-staged pass writes to LDS before its next stage adds them:
// kept out of FMA contraction with those additions, as they are there.
__device__ __forceinline__ float2 scaled(float s, float2 a) {
#pragma clang fp contract(off)
    return make_float2(s * a.x, s * a.y);
}
__device__ __forceinline__ float2 grad_line(float k, float2 a) {
#pragma clang fp contract(off)
    return make_float2(k * a.y, k * a.x);
}

// One block: nb (<= NT / R1) adjacent columns q0 .. q0 + nb - 1 of the last-axis lines,
// TT types and CH charge fields (NF = TT + CH fields).
template <int R1, int R2, int NT, int TT, int CH>
__device__ __forceinline__ void last_reg_block(const RepDev& r, int blk, float2* lds, const float* rw) {
    constexpr int N = R1 * R2, NF = TT + CH, NR = TT * (TT + 1) / 2;
    const int la = r.dim - 1, D = r.dim;
    const int Q = la == 2 ? r.Ny * r.Nxh : r.Nxh;   // plane size = line stride
    const int nl = r.nl_last;                       // columns per block (<= NT / R1)
    const int q0 = blk * nl;
    const int nb = min(nl, Q - q0);
    const int ldx = NF * nl + 1, ldy = 2 * NF * nl + 1;
    float2* X1 = lds;                               // N x ldx: forward exchange
    float2* Y1 = lds + (size_t)N * ldx;             // N x ldy: inverse exchange
    const float2* tw1 = r.ax[la].tw + r.ax[la].tw_off[1];   // w_N^{i r}, i < R1, 1 <= r < R2
    const float2* tw2 = r.axr.tw + r.axr.tw_off[1];         // w_N^{j r}, j < R2, 1 <= r < R1
    const bool o32 = r.grid_off32;
    const FastDiv dnl(nl);
    int b, i;
    i = dnl.div((int)threadIdx.x, b);
    const bool liveB = i < R1 && b < nb;

    // coupling rows of this thread's phase-B nodes, requested before anything else
    const int nrows = r.ctab_rows, nlc = nl;
    gmem<const float> cb = G(r.Ctab) + (size_t)blk * N * nrows * nlc + b;
    float c[R2][NR], cq[R2];
    if (liveB) {
#pragma unroll
        for (int q = 0; q < R2; ++q) {
            gmem<const float> cn = cb + (size_t)(i + R1 * q) * nrows * nlc;
#pragma unroll
            for (int k = 0; k < NR; ++k) c[q][k] = cn[k * nlc];
            cq[q] = CH ? cn[r.ctab_charge * nlc] : 0.f;
        }
    }

    // ---- A: rows i + R2 r of every field, DFT_R1 (Stockham stage 1: out[i R1 + q])
    if (i < R2 && b < nb) {
        float2 v[NF][R1];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int rr = 0; rr < R1; ++rr)
                v[f][rr] = gld(grid_at((const float2*)r.H, o32, (size_t)f * r.Mh + (size_t)(i + R2 * rr) * Q + q0 + b));
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            Dft<R1>::run(v[f]);
#pragma unroll
            for (int q = 0; q < R1; ++q) X1[(i * R1 + q) * ldx + f * nl + b] = v[f][q];
        }
    }
    __syncthreads();

    // ---- B: stage 2 (radix R2, twiddles w_N^{i r}) -> rho_f(k), k = i + R1 q
    if (liveB) {
        float2 w[NF][R2];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int rr = 0; rr < R2; ++rr) w[f][rr] = X1[(i + R1 * rr) * ldx + f * nl + b];
        float2 t[R2];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) t[rr] = gld(tw1 + i * (R2 - 1) + rr - 1);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int rr = 1; rr < R2; ++rr) w[f][rr] = cmul(w[f][rr], t[rr]);
            Dft<R2>::run(w[f]);
        }
        // coupling: w_I = sum_J C_IJ rho_J over the symmetric rows pidx(lo, hi) =
        // lo TT - lo (lo - 1) / 2 + (hi - lo); ramped styles add their weighted rows;
        // charges w_c = G_c rho_c
#pragma unroll
        for (int q = 0; q < R2; ++q) {
            const int n = i + R1 * q;
            for (int j = 0; j < r.nramp; ++j) {   // uniform trip count
                const float v = rw[j] * cb[((size_t)n * nrows + r.ramp_row0 + j) * nlc];
#pragma unroll
                for (int k = 0; k < NR; ++k) c[q][k] += r.ramp_pair[j] == k ? v : 0.f;
            }
            float2 rh[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) rh[f] = w[f][q];
#pragma unroll
            for (int I = 0; I < TT; ++I) {
                float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                for (int J = 0; J < TT; ++J) {
                    const int lo = I < J ? I : J, hi = I < J ? J : I;
                    const float cc = c[q][lo * TT - lo * (lo - 1) / 2 + (hi - lo)];
                    acc.x += cc * rh[J].x; acc.y += cc * rh[J].y;
                }
                w[I][q] = acc;
            }
            if (CH) w[TT][q] = scaled(cq[q], rh[TT]);
        }
        // inverse lines (the inverse is conj(FFT(conj))): o = 0 conj(w) = (w.x, -w.y),
        // o = 1 conj(-i k_la w) = (k w.y, k w.x), k_la = 0 on the Nyquist plane;
        // first stage of the R2 x R1 plan: radix R2, out[i R2 + q]
#pragma unroll
        for (int o = 0; o < 2; ++o) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float2 z[R2];
#pragma unroll
                for (int q = 0; q < R2; ++q) {
                    const int n = i + R1 * q;
                    const float2 a = w[f][q];
                    if (o == 0) z[q] = make_float2(a.x, -a.y);
                    else {
                        const float kl = 2 * n == N ? 0.f : r.kfac[la] * (float)(2 * n < N ? n : n - N);
                        z[q] = grad_line(kl, a);
                    }
                }
                Dft<R2>::run(z);
#pragma unroll
                for (int q = 0; q < R2; ++q) Y1[(i * R2 + q) * ldy + (o * NF + f) * nl + b] = z[q];
            }
        }
    }
    __syncthreads();

    // ---- C: stage 2 of the R2 x R1 plan (radix R1, twiddles w_N^{j r}, including the
    // unit ones of j = 0 as the staged pass multiplies them) -> rows j + R2 q
    if (i < R2 && b < nb) {
        const int j = i;
        float2 t[R1];
#pragma unroll
        for (int rr = 1; rr < R1; ++rr) t[rr] = gld(tw2 + j * (R1 - 1) + rr - 1);
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const int slot = o == 0 ? 0 : la;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float2 z[R1];
#pragma unroll
                for (int rr = 0; rr < R1; ++rr) z[rr] = Y1[(j + R2 * rr) * ldy + (o * NF + f) * nl + b];
#pragma unroll
                for (int rr = 1; rr < R1; ++rr) z[rr] = cmul(z[rr], t[rr]);
                Dft<R1>::run(z);
                const size_t base = (size_t)(f * D + slot) * r.Mh + q0 + b;
#pragma unroll
                for (int q = 0; q < R1; ++q)
                    gst(grid_at(r.Gk, o32, base + (size_t)(j + R2 * q) * Q), conjf2(z[q]));
            }
        }
    }
}

template <int R1, int R2, int NT>
__device__ __forceinline__ void last_reg_types(const RepDev& r, int blk, float2* lds, const float* rw) {
    const bool ch = r.nf > r.ntypes;
    switch (r.ntypes * 2 + (ch ? 1 : 0)) {
        case 2: last_reg_block<R1, R2, NT, 1, 0>(r, blk, lds, rw); break;
        case 3: last_reg_block<R1, R2, NT, 1, 1>(r, blk, lds, rw); break;
        case 4: last_reg_block<R1, R2, NT, 2, 0>(r, blk, lds, rw); break;
        case 5: last_reg_block<R1, R2, NT, 2, 1>(r, blk, lds, rw); break;
        case 6: last_reg_block<R1, R2, NT, 3, 0>(r, blk, lds, rw); break;
        case 7: last_reg_block<R1, R2, NT, 3, 1>(r, blk, lds, rw); break;
        default: break;   // excluded by last_reg_plan on the host
    }
}

}  // namespace

template <int NT>
__global__ void __launch_bounds__(NT) k_last_reg(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap,
                                                 int koff) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const AxisPlan& pl = r.ax[r.dim - 1];
    const int la = r.dim - 1;
    const int N = la == 2 ? r.Nz : r.Ny;
    // ramp weights after both exchange buffers
    const int ldx = r.nf * r.nl_last + 1, ldy = 2 * r.nf * r.nl_last + 1;
    float* rw = (float*)(lds + (size_t)N * (ldx + ldy));
    if ((int)threadIdx.x < r.nramp) rw[threadIdx.x] = ramp_weight(r, threadIdx.x, r.step + koff);
    __syncthreads();
    switch (pl.rad[0] * 32 + pl.rad[1]) {   // uniform per block
        case 16 * 32 + 10: last_reg_types<16, 10, NT>(r, wi.blk, lds, rw); break;
        case 10 * 32 + 8: last_reg_types<10, 8, NT>(r, wi.blk, lds, rw); break;
        case 8 * 32 + 8: last_reg_types<8, 8, NT>(r, wi.blk, lds, rw); break;
        case 8 * 32 + 5: last_reg_types<8, 5, NT>(r, wi.blk, lds, rw); break;
        default: break;   // excluded by last_reg_plan on the host
    }
}

bool last_reg_plan(const AxisPlan& pl, int ntypes, int nf) {
    if (pl.nst != 2 || ntypes < 1 || ntypes > 3 || nf > ntypes + 1) return false;
    const int k = pl.rad[0] * 32 + pl.rad[1];
    return k == 16 * 32 + 10 || k == 10 * 32 + 8 || k == 8 * 32 + 8 || k == 8 * 32 + 5;
}

hipError_t launch_last_reg(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_last_reg<kLastRegThreads>, dim3(c.n), dim3(kLastRegThreads), c.lds, c.reps, c.wmap, koff);
    return hipGetLastError();
}

hipError_t configure_last_reg() {
    return hipFuncSetAttribute((const void*)k_last_reg<kLastRegThreads>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPlaneLds);
}

}  // namespace tild
