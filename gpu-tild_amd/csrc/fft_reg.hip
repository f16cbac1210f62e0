// fft_reg.hip -- FFT passes of the step with register-resident columns (CDNA4 / gfx950):
// the forward plane pass (k_plane_reg), the fused last-axis pass (k_last_reg) and the y
// inverse of the gradient spectra (k_grad_reg).  Each takes the axes whose
// Stockham plan has two stages R1 x R2 (R1 >= R2) among 16 x 10, 10 x 8, 8 x 8 and 8 x 5;
// a thread keeps its columns in registers between the stages, so a pass makes one to
// three LDS exchanges where the LDS-staged forms (kernels.hip) stage every value through
// LDS with a barrier after each phase.  (The inverse x transform into the gather
// records stays LDS-staged: its register forms measured slower, see the design notes.)
//
// k_last_reg: the same operation as k_last_fused (kernels.hip) -- last-axis forward FFT of every
// field, the k-space coupling w_I = sum_J C_IJ rho_J (+ the charge row), and the
// inverse transforms of w and -i k_la w -- for last axes with a two-stage plan
// N = R1 R2 (R1 >= R2: 160 = 16 x 10, 80 = 10 x 8, 64 = 8 x 8, 40 = 8 x 5).  Each
// thread keeps its Stockham columns in registers between the stages, so a block makes
// two LDS exchanges and two barriers, where the LDS-staged form stages every value
// through LDS and ends each of its nine phases with a barrier:
//
//   A  task (b, i < R2, field): rows i + R2 r (r < R1) of column b, read from HBM
//      straight into registers; DFT_R1; exchange.
//   B  thread (b, i < R1, h < 2): twiddles w_N^{i r}, DFT_R2 -> rho_f(k) at
//      k = i + R1 q; the coupling in registers; inverse direction h (conj(w) or
//      conj(-i k w)) is transformed by the Stockham plan R2 x R1, whose first stage
//      (radix R2, no twiddles) reads exactly the values k = i + R1 q this thread holds;
//      exchange.
//   C  task (b, j < R2, line): twiddles w_N^{j r}, DFT_R1 -> rows j + R2 q; store.
//
// The transforms and sums are those of the LDS-staged pass (whose inverse also runs the
// reversed plan RepDev::axr); the two agree to rounding, since the compiler contracts
// multiplies into the following additions differently where one keeps in registers
// what the other stages through LDS.
#include "device.hpp"
#include "fft_dev.hpp"
#include "launch.hpp"

namespace tild {

namespace {

bool reg_plan_pair(int k) {   // k = R1 * 32 + R2
    switch (k) {
#define TILD_CASE(A, B) case A * 32 + B: return true;
        TILD_REG_PLANS(TILD_CASE)
#undef TILD_CASE
        default: return false;
    }
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

// One block: NB = NT / (2 R1) adjacent columns q0 .. q0 + NB - 1 of the last-axis lines
// (fewer in the last block), TT types and CH charge fields (NF = TT + CH fields).  Two
// threads share each (column, i) of the middle stage: both run the forward stage and
// the coupling, and each transforms one of the two inverse directions.
// LDS: the forward exchange (N x ldx), the inverse exchange (N x ldy), both stages'
// twiddles, then the ramp weights (last_reg_lds_bytes).
template <int R1, int R2, int NT, int TT, int CH>
__device__ __forceinline__ void last_reg_block(const RepDev& r, int blk, float2* lds, int koff) {
    constexpr int N = R1 * R2, NF = TT + CH, NR = TT * (TT + 1) / 2, NB = NT / (2 * R1);
    constexpr int ldx = NF * NB + 1, ldy = 2 * NF * NB + 1;
    constexpr int NT1 = R1 * (R2 - 1), NT2 = R2 * (R1 - 1);
    const int la = r.dim - 1, D = r.dim;
    const int Q = la == 2 ? r.Ny * r.Nxh : r.Nxh;   // plane size = line stride
    const int q0 = blk * NB;
    const int nb = min(NB, Q - q0);
    float2* X1 = lds;                               // N x ldx: forward exchange
    float2* Y1 = lds + N * ldx;                     // N x ldy: inverse exchange
    float2* TW1 = Y1 + N * ldy;                     // w_N^{i r}, i < R1, 1 <= r < R2 (forward stage 2)
    float2* TW2 = TW1 + NT1;                        // w_N^{j r}, j < R2, 1 <= r < R1 (inverse stage 2)
    float* rw = (float*)(TW2 + NT2);                // ramp weights of this step
    // phase B: thread (b, i, h), h = the inverse direction it transforms
    const int b = (int)threadIdx.x % NB, i = ((int)threadIdx.x / NB) % R1, h = (int)threadIdx.x / (NB * R1);
    // grid buffers below 4 GiB (required on the host): 32-bit element offsets
    const uint32_t Qu = (uint32_t)Q, Mhu = (uint32_t)r.Mh;
    const bool liveB = h < 2 && b < nb;

    // set-up, visible after the first barrier: twiddles and ramp weights into LDS
    {
        const float2* tw1 = r.ax[la].tw + r.ax[la].tw_off[1];
        const float2* tw2 = r.axr.tw + r.axr.tw_off[1];
        for (int t = threadIdx.x; t < NT1; t += NT) TW1[t] = gld(tw1 + t);
        for (int t = threadIdx.x; t < NT2; t += NT) TW2[t] = gld(tw2 + t);
        if ((int)threadIdx.x < r.nramp) rw[threadIdx.x] = ramp_weight(r, threadIdx.x, r.step + koff);
    }
    // coupling rows of this thread's phase-B nodes, requested before anything else
    const int nrows = r.ctab_rows;
    // element (n, row) of this thread's column: cb0 + (n nrows + row) NB, from the
    // uniform table base (every offset 32-bit)
    const uint32_t cb0 = (uint32_t)blk * (uint32_t)(N * nrows * NB) + (uint32_t)b;
    auto ctab = [&](int n, int row) { return G(at32(r.Ctab, cb0 + (uint32_t)((n * nrows + row) * NB)))[0]; };
    // up to three types every row of the thread's nodes is requested at block start
    // and held through the forward stages; with four (ten rows per node) each node's
    // rows are requested where the coupling uses them, which keeps the pass within 256
    // registers
    constexpr bool kEarly = NR <= 6;
    constexpr int QE = kEarly ? R2 : 1;
    float c[QE][NR], cq[QE];
    auto load_rows = [&](int q, int slot) {
#pragma unroll
        for (int k = 0; k < NR; ++k) c[slot][k] = ctab(i + R1 * q, k);
        cq[slot] = CH ? ctab(i + R1 * q, r.ctab_charge) : 0.f;
    };
    if (kEarly && liveB) {
#pragma unroll
        for (int q = 0; q < R2; ++q) load_rows(q, q < QE ? q : 0);
    }

    // ---- A: task (b, i < R2, f): rows i + R2 r of field f, DFT_R1 (Stockham stage 1:
    // out[i R1 + q])
    // lanes: column, then field, then i (the fields' LDS rows are NB values apart,
    // which spreads a wave's writes over the banks)
    for (int task = threadIdx.x; task < NB * R2 * NF; task += NT) {
        const int ba = task % NB, f = (task / NB) % NF, ia = task / (NB * NF);
        if (ba >= nb) continue;
        float2 v[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr)
            v[rr] = gld(at32((const float2*)r.H, (uint32_t)f * Mhu + (uint32_t)(ia + R2 * rr) * Qu + (uint32_t)(q0 + ba)));
        Dft<R1>::run(v);
#pragma unroll
        for (int q = 0; q < R1; ++q) X1[(ia * R1 + q) * ldx + f * NB + ba] = v[q];
    }
    __syncthreads();

    // ---- B: stage 2 (radix R2, twiddles w_N^{i r}) -> rho_f(k), k = i + R1 q
    if (liveB) {
        float2 w[NF][R2];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int rr = 0; rr < R2; ++rr) w[f][rr] = X1[(i + R1 * rr) * ldx + f * NB + b];
        float2 t[R2];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) t[rr] = TW1[i * (R2 - 1) + rr - 1];
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
            const int cs = kEarly ? q : 0;     // slot of this node's rows
            if (!kEarly) load_rows(q, 0);
            for (int j = 0; j < r.nramp; ++j) {   // uniform trip count
                const float v = rw[j] * ctab(n, r.ramp_row0 + j);
#pragma unroll
                for (int k = 0; k < NR; ++k) c[cs][k] += r.ramp_pair[j] == k ? v : 0.f;
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
                    const float cc = c[cs][lo * TT - lo * (lo - 1) / 2 + (hi - lo)];
                    acc.x += cc * rh[J].x; acc.y += cc * rh[J].y;
                }
                w[I][q] = acc;
            }
            if (CH) w[TT][q] = scaled(cq[cs], rh[TT]);
        }
        // this thread's inverse lines (the inverse is conj(FFT(conj))): h = 0
        // conj(w) = (w.x, -w.y), h = 1 conj(-i k_la w) = (k w.y, k w.x), k_la = 0 on the
        // Nyquist plane; first stage of the R2 x R1 plan: radix R2, out[i R2 + q]
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            float2 z[R2];
#pragma unroll
            for (int q = 0; q < R2; ++q) {
                const int n = i + R1 * q;
                const float2 a = w[f][q];
                const float kl = 2 * n == N ? 0.f : r.kfac[la] * (float)(2 * n < N ? n : n - N);
                z[q] = h == 0 ? make_float2(a.x, -a.y) : grad_line(kl, a);
            }
            Dft<R2>::run(z);
#pragma unroll
            for (int q = 0; q < R2; ++q) Y1[(i * R2 + q) * ldy + (h * NF + f) * NB + b] = z[q];
        }
    }
    __syncthreads();

    // ---- C: stage 2 of the R2 x R1 plan (radix R1, twiddles w_N^{j r}, including the
    // unit ones of j = 0 as the staged pass multiplies them) -> rows j + R2 q.  One task
    // per (column, j, line), spread over all threads.
    // lanes: column, then line, then j (lines NB values apart in LDS: a wave's reads of
    // one row fall in distinct banks)
    for (int task = threadIdx.x; task < NB * R2 * 2 * NF; task += NT) {
        const int bc = task % NB, rest = task / NB, line = rest % (2 * NF), j = rest / (2 * NF);   // line = o NF + f
        if (bc >= nb) continue;
        const int o = line >= NF ? 1 : 0, f = line - o * NF;
        float2 z[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) z[rr] = Y1[(j + R2 * rr) * ldy + line * NB + bc];
#pragma unroll
        for (int rr = 1; rr < R1; ++rr) z[rr] = cmul(z[rr], TW2[j * (R1 - 1) + rr - 1]);
        Dft<R1>::run(z);
        const uint32_t o0 = (uint32_t)(f * D + (o == 0 ? 0 : la)) * Mhu + (uint32_t)j * Qu + (uint32_t)(q0 + bc);
        const uint32_t step = (uint32_t)R2 * Qu;
#pragma unroll
        for (int q = 0; q < R1; ++q) gst(at32(r.Gk, o0 + (uint32_t)q * step), conjf2(z[q]));
    }
}

template <int R1, int R2, int NT, int TT, int CH, int CLS>
__device__ __forceinline__ void last_reg_variant_block(const RepDev& r, int blk, float2* lds, int koff) {
    if constexpr (last_reg_variant(R1, R2, TT, CH) && last_reg_class(TT) == CLS)
        last_reg_block<R1, R2, NT, TT, CH>(r, blk, lds, koff);
}

template <int R1, int R2, int NT, int CLS>
__device__ __forceinline__ void last_reg_types(const RepDev& r, int blk, float2* lds, int koff) {
    const bool ch = r.nf > r.ntypes;
    switch (r.ntypes * 2 + (ch ? 1 : 0)) {
        case 2: last_reg_variant_block<R1, R2, NT, 1, 0, CLS>(r, blk, lds, koff); break;
        case 3: last_reg_variant_block<R1, R2, NT, 1, 1, CLS>(r, blk, lds, koff); break;
        case 4: last_reg_variant_block<R1, R2, NT, 2, 0, CLS>(r, blk, lds, koff); break;
        case 5: last_reg_variant_block<R1, R2, NT, 2, 1, CLS>(r, blk, lds, koff); break;
        case 6: last_reg_variant_block<R1, R2, NT, 3, 0, CLS>(r, blk, lds, koff); break;
        case 7: last_reg_variant_block<R1, R2, NT, 3, 1, CLS>(r, blk, lds, koff); break;
        case 8: last_reg_variant_block<R1, R2, NT, 4, 0, CLS>(r, blk, lds, koff); break;
        case 9: last_reg_variant_block<R1, R2, NT, 4, 1, CLS>(r, blk, lds, koff); break;
        default: break;   // excluded by last_reg_plan on the host
    }
}

}  // namespace

// CLS: the register class of the type counts the instantiation holds (last_reg_class),
// so that the four-type variants' registers do not lower the others' occupancy
template <int NT, int CLS>
__global__ void __launch_bounds__(NT) k_last_reg(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap,
                                                 int koff) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const AxisPlan& pl = r.ax[r.dim - 1];
    switch (pl.rad[0] * 32 + pl.rad[1]) {   // uniform per block
#define TILD_CASE(A, B) case A * 32 + B: last_reg_types<A, B, NT, CLS>(r, wi.blk, lds, koff); break;
        TILD_REG_PLANS(TILD_CASE)
#undef TILD_CASE
        default: break;   // excluded by last_reg_plan on the host
    }
}

// ------------------------------------------------------------------ y inverse (3D)
// The y-axis inverse of the gradient spectra (the operation of k_mid_grad, kernels.hip)
// for two-stage y plans R1 x R2, columns in registers between the stages: per (field,
// z) plane and NB = NT / R1 adjacent x columns,
//   A  thread (b, i), i < R2: rows i + R2 r of U (slot 0) and V (slot 2), the three
//      lines conj(U), conj(-i k_y U), conj(V), DFT_R1; exchange;
//   B  thread (b, i), i < R1: twiddles w_Ny^{i r}, DFT_R2 -> rows i + R1 q of the three
//      outputs (slots 0, 1, 2).
// One barrier.  A block reads its columns of slots 0 and 2 before it writes any slot,
// and no other block touches them.
namespace {
template <int R1, int R2, int NT>
__device__ __forceinline__ void grad_reg_block(const RepDev& r, int blk, float2* lds) {
    constexpr int N = R1 * R2, NB = NT / R1, ld = 3 * NB + 1, NT1 = R1 * (R2 - 1);
    const int Nxh = r.Nxh;
    const int nchunk = (Nxh + NB - 1) / NB;
    const int o = blk / nchunk, c = blk - o * nchunk;   // o = f Nz + z
    const int f = o / r.Nz, z = o - f * r.Nz;
    const int b0 = c * NB, nb = min(NB, Nxh - b0);
    float2* X1 = lds;                 // N x ld
    float2* TW1 = X1 + N * ld;        // w_Ny^{i r}, i < R1, 1 <= r < R2
    const int b = (int)threadIdx.x % NB, i = (int)threadIdx.x / NB;
    {
        const float2* tw1 = r.ax[1].tw + r.ax[1].tw_off[1];
        for (int t = threadIdx.x; t < NT1; t += NT) TW1[t] = gld(tw1 + t);
    }
    // element (n, column b) of slot s: 32-bit offsets from the uniform base
    const uint32_t Mhu = (uint32_t)r.Mh, Nxu = (uint32_t)Nxh;
    const uint32_t plane = (uint32_t)z * (uint32_t)N * Nxu + (uint32_t)(b0 + b);
    const uint32_t sU = (uint32_t)(f * 3 + 0) * Mhu + plane, sY = (uint32_t)(f * 3 + 1) * Mhu + plane,
                   sV = (uint32_t)(f * 3 + 2) * Mhu + plane;
    if (i < R2 && b < nb) {
        float2 u[R1], v[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const uint32_t n = (uint32_t)(i + R2 * rr) * Nxu;
            u[rr] = gld(at32((const float2*)r.Gk, sU + n));
            v[rr] = gld(at32((const float2*)r.Gk, sV + n));
        }
        float2 L[3][R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const int n = i + R2 * rr;
            const float ky = 2 * n == N ? 0.f : r.kfac[1] * (float)(2 * n < N ? n : n - N);
            // lines of the conj(FFT(conj)) inverse: conj(u), conj(-i ky u), conj(v)
            L[0][rr] = make_float2(u[rr].x, -u[rr].y);
            L[1][rr] = grad_line(ky, u[rr]);
            L[2][rr] = make_float2(v[rr].x, -v[rr].y);
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            Dft<R1>::run(L[l]);
#pragma unroll
            for (int q = 0; q < R1; ++q) X1[(i * R1 + q) * ld + l * NB + b] = L[l][q];
        }
    }
    __syncthreads();
    if (i < R1 && b < nb) {
        float2 t[R2];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) t[rr] = TW1[i * (R2 - 1) + rr - 1];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            float2 w[R2];
#pragma unroll
            for (int rr = 0; rr < R2; ++rr) w[rr] = X1[(i + R1 * rr) * ld + l * NB + b];
#pragma unroll
            for (int rr = 1; rr < R2; ++rr) w[rr] = cmul(w[rr], t[rr]);
            Dft<R2>::run(w);
            const uint32_t s0 = l == 0 ? sU : (l == 1 ? sY : sV);
#pragma unroll
            for (int q = 0; q < R2; ++q)
                gst(at32(r.Gk, s0 + (uint32_t)(i + R1 * q) * Nxu), conjf2(w[q]));
        }
    }
}
}  // namespace

template <int NT>
__global__ void __launch_bounds__(NT) k_grad_reg(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const AxisPlan& pl = r.ax[1];
    switch (pl.rad[0] * 32 + pl.rad[1]) {   // uniform per block
#define TILD_CASE(A, B) case A * 32 + B: grad_reg_block<A, B, NT>(r, wi.blk, lds); break;
        TILD_REG_PLANS(TILD_CASE)
#undef TILD_CASE
        default: break;   // excluded by grad_reg_plan on the host
    }
}

// ------------------------------------------------------------------ plane forward (3D)
// x R2C and y FFT of one (field, z) density plane (the operation of k_plane_r2c,
// kernels.hip) for square planes Nx = Ny with a two-stage plan R1 x R2:
//   A  task (pair j, i < R2): rows 2j and 2j+1 at x = i + R2 r as one complex line, DFT_R1;
//   B  task (j, i < R1): twiddles, DFT_R2 -> Z_j(k), k = i + R1 q; exchange;
//   C  task (column k < Nxh, i < R2): the half spectra of rows y = i + R2 r unpacked on
//      the fly from Z_{y/2}(k) and Z_{y/2}(N - k), DFT_R1;
//   D  task (k, i < R1): twiddles, DFT_R2 -> H(y = i + R1 q, k), stored.
// Three barriers, where the staged pass has seven; loads and stores go straight to and
// from registers.
namespace {
__device__ __forceinline__ float half_of(float a) {
#pragma clang fp contract(off)
    return 0.5f * a;
}

template <int R1, int R2, int NT>
__device__ __forceinline__ void plane_reg_block(const RepDev& r, int blk, float2* lds) {
    constexpr int N = R1 * R2, NP = N / 2, NXH = N / 2 + 1, ldp = NP | 1, ldc = NXH | 1;
    constexpr int NT1 = R1 * (R2 - 1), LDA = (N * ldp > N * ldc) ? N * ldp : N * ldc;
    const int z = blk % r.Nz, f = blk / r.Nz;
    float2* Xa = lds;                  // x stage-1 output, later the y stage-1 output (N x ldc)
    float2* Xb = Xa + LDA;             // Z: x transforms of the row pairs (N x ldp)
    float2* TWx = Xb + N * ldp;        // w_N^{i r} of the x plan and of the y plan
    float2* TWy = TWx + NT1;
    {
        const float2* tx = r.ax[0].tw + r.ax[0].tw_off[1];
        const float2* ty = r.ax[1].tw + r.ax[1].tw_off[1];
        for (int t = threadIdx.x; t < NT1; t += NT) { TWx[t] = gld(tx + t); TWy[t] = gld(ty + t); }
    }
    const uint32_t pbase = (uint32_t)f * (uint32_t)r.M + (uint32_t)z * (uint32_t)(N * N);
    // A: pairs of rows as complex lines, stage 1 of the x plan
    for (int task = threadIdx.x; task < NP * R2; task += NT) {
        const int i = task % R2, j = task / R2;
        float2 v[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const uint32_t o = pbase + (uint32_t)(2 * j * N + i + R2 * rr);
            v[rr] = make_float2(G(at32((const float*)r.rho, o))[0], G(at32((const float*)r.rho, o + N))[0]);
        }
        Dft<R1>::run(v);
#pragma unroll
        for (int q = 0; q < R1; ++q) Xa[(i * R1 + q) * ldp + j] = v[q];
    }
    __syncthreads();
    // B: stage 2 of the x plan -> Z_j(k)
    for (int task = threadIdx.x; task < NP * R1; task += NT) {
        const int i = task % R1, j = task / R1;
        float2 v[R2];
#pragma unroll
        for (int rr = 0; rr < R2; ++rr) v[rr] = Xa[(i + R1 * rr) * ldp + j];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) v[rr] = cmul(v[rr], TWx[i * (R2 - 1) + rr - 1]);
        Dft<R2>::run(v);
#pragma unroll
        for (int q = 0; q < R2; ++q) Xb[(i + R1 * q) * ldp + j] = v[q];
    }
    __syncthreads();
    // C: unpack rows y = i + R2 r of column k (even y: the real row's spectrum
    // (Z_k + conj Z_{N-k}) / 2, odd y: the imaginary row's (Z_k - conj Z_{N-k}) / 2i),
    // stage 1 of the y plan
    for (int task = threadIdx.x; task < NXH * R2; task += NT) {
        const int k = task % NXH, i = task / NXH;
        const int km = k == 0 ? 0 : N - k;
        float2 v[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const int y = i + R2 * rr, j = y >> 1;
            const float2 Zk = Xb[k * ldp + j], Zm = Xb[km * ldp + j];
            v[rr] = (y & 1) == 0 ? make_float2(half_of(Zk.x + Zm.x), half_of(Zk.y - Zm.y))
                                 : make_float2(half_of(Zk.y + Zm.y), half_of(Zm.x - Zk.x));
        }
        Dft<R1>::run(v);
#pragma unroll
        for (int q = 0; q < R1; ++q) Xa[(i * R1 + q) * ldc + k] = v[q];
    }
    __syncthreads();
    // D: stage 2 of the y plan -> H(y, k)
    const uint32_t hbase = ((uint32_t)f * (uint32_t)r.Nz + (uint32_t)z) * (uint32_t)(N * NXH);
    for (int task = threadIdx.x; task < NXH * R1; task += NT) {
        const int k = task % NXH, i = task / NXH;
        float2 v[R2];
#pragma unroll
        for (int rr = 0; rr < R2; ++rr) v[rr] = Xa[(i + R1 * rr) * ldc + k];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) v[rr] = cmul(v[rr], TWy[i * (R2 - 1) + rr - 1]);
        Dft<R2>::run(v);
#pragma unroll
        for (int q = 0; q < R2; ++q) gst(at32(r.H, hbase + (uint32_t)((i + R1 * q) * NXH + k)), v[q]);
    }
}
}  // namespace

template <int NT>
__global__ void __launch_bounds__(NT) k_plane_reg(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const AxisPlan& pl = r.ax[0];
    switch (pl.rad[0] * 32 + pl.rad[1]) {   // uniform per block
#define TILD_CASE(A, B) case A * 32 + B: plane_reg_block<A, B, NT>(r, wi.blk, lds); break;
        TILD_REG_PLANS(TILD_CASE)
#undef TILD_CASE
        default: break;   // excluded by plane_reg_plan on the host
    }
}

// ------------------------------------------------------------------ plane inverse (3D)
// The y inverse of the gradient spectra and the x C2R into gather records of one
// (field, z) plane in one block (the operations of k_grad_reg and k_x_c2r, kernels.hip,
// without the half spectra of the y inverses going through HBM), for square planes
// Nx = Ny = N with a two-stage plan R1 x R2 (an odd N pairs its last row's z derivative
// with an all-zero line):
//   A  task (column k < N/2 + 1, i < R2): rows i + R2 r of U and V (slots 0 and 2),
//      the lines conj(U), conj(-i k_y U), conj(V), DFT_R1;
//   B  task (k, i < R1): twiddles, DFT_R2 -> the y inverses Y_l(y, k), y = i + R1 q,
//      written back over the rows the task read;
//   C  task (x line c, i < R2): the complex lines of the x inverse -- per row y the
//      x derivative (-i k_x applied) plus i times the y derivative, per row pair the
//      z derivatives of rows 2p and 2p + 1 -- extended to the full spectrum by
//      Hermitian symmetry at k = i + R2 r, DFT_R1; held across a barrier, then written
//      over the y inverses;
//   D  task (c, i < R1): twiddles, DFT_R2 -> the real lines at x = i + R1 q, held
//      across a barrier, then laid out as the plane's records in LDS and copied out.
// Six barriers; the y and x transforms and the conjugation and Hermitian rules are
// those of the two staged passes, which the records match to rounding (the x lines
// pair components differently).  Records carry no density mask (see k_x_c2r).
namespace {
template <int R1, int R2, int NT>
__device__ __forceinline__ void plane_inv_block(const RepDev& r, int blk, float2* lds) {
    constexpr int N = R1 * R2, NXH = N / 2 + 1, NT1 = R1 * (R2 - 1);
    constexpr int LA = (3 * NXH) | 1;          // row stride of the y inverses: [y][line][k]
    constexpr int NCL = N + (N + 1) / 2;       // x lines: N rows + the row pairs
    constexpr int LC = NCL | 1;                // row stride of the x exchange
    constexpr int CT = (NCL * R2 + NT - 1) / NT;   // stage C tasks per thread
    constexpr int DT = (NCL * R1 + NT - 1) / NT;   // stage D tasks per thread
    static_assert(LC <= LA && 3 * N * N <= 2 * N * LA, "plane_inv_block: exchange and records within the y buffer");
    const int z = blk % r.Nz, f = blk / r.Nz;
    float2* X = lds;                   // N x LA: y stages, then the x exchange (N x LC)
    float2* TW = X + N * LA;           // w_N^{i r}, i < R1, 1 <= r < R2 (same plan on x and y)
    {
        const float2* tw = r.ax[1].tw + r.ax[1].tw_off[1];
        for (int t = threadIdx.x; t < NT1; t += NT) TW[t] = gld(tw + t);
    }
    const uint32_t Mhu = (uint32_t)r.Mh;
    const uint32_t plane = (uint32_t)z * (uint32_t)(N * NXH);
    const uint32_t sU = (uint32_t)(f * 3 + 0) * Mhu + plane, sV = (uint32_t)(f * 3 + 2) * Mhu + plane;
    const float kfy = r.kfac[1], kfx = r.kfac[0];
    // A: y stage 1 of the three lines of column k
    for (int task = threadIdx.x; task < NXH * R2; task += NT) {
        const int k = task % NXH, i = task / NXH;
        float2 u[R1], v[R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const uint32_t o = (uint32_t)((i + R2 * rr) * NXH + k);
            u[rr] = gld(at32((const float2*)r.Gk, sU + o));
            v[rr] = gld(at32((const float2*)r.Gk, sV + o));
        }
        float2 L[3][R1];
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const int n = i + R2 * rr;
            const float ky = 2 * n == N ? 0.f : kfy * (float)(2 * n < N ? n : n - N);
            // lines of the conj(FFT(conj)) inverse: conj(u), conj(-i ky u), conj(v)
            L[0][rr] = make_float2(u[rr].x, -u[rr].y);
            L[1][rr] = grad_line(ky, u[rr]);
            L[2][rr] = make_float2(v[rr].x, -v[rr].y);
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            Dft<R1>::run(L[l]);
#pragma unroll
            for (int q = 0; q < R1; ++q) X[(i * R1 + q) * LA + l * NXH + k] = L[l][q];
        }
    }
    __syncthreads();
    // B: y stage 2; Y_l(y, k) = conj of the result, over the rows the task read
    for (int task = threadIdx.x; task < NXH * R1; task += NT) {
        const int k = task % NXH, i = task / NXH;
        float2 t[R2];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) t[rr] = TW[i * (R2 - 1) + rr - 1];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            float2 w[R2];
#pragma unroll
            for (int rr = 0; rr < R2; ++rr) w[rr] = X[(i + R1 * rr) * LA + l * NXH + k];
#pragma unroll
            for (int rr = 1; rr < R2; ++rr) w[rr] = cmul(w[rr], t[rr]);
            Dft<R2>::run(w);
#pragma unroll
            for (int q = 0; q < R2; ++q) X[(i + R1 * q) * LA + l * NXH + k] = conjf2(w[q]);
        }
    }
    __syncthreads();
    // C: x stage 1.  Line c < N is row y = c: A = -i k_x Y_0, B = Y_1; line N + p pairs
    // A = Y_2(2p), B = Y_2(2p + 1).  Z = A + i B with Z_{N-k} = conj A_k + i conj B_k;
    // the self-conjugate columns (0, N/2) drop their imaginary parts; conj(Z) is
    // transformed (the inverse as conj(FFT(conj))).
    float2 hold[CT][R1];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int task = (int)threadIdx.x + ct * NT;
        if (task >= NCL * R2) break;
        const int c = task % NCL, i = task / NCL;
        const bool rowline = c < N;
        const int ya = rowline ? c : 2 * (c - N), yb = rowline ? c : ya + 1;
        const int la = rowline ? 0 : 2, lb = rowline ? 1 : 2;
#pragma unroll
        for (int rr = 0; rr < R1; ++rr) {
            const int kn = i + R2 * rr;                 // full-spectrum index
            const bool upper = kn >= NXH;
            const int kk = upper ? N - kn : kn;         // its half-spectrum column
            float2 A = X[ya * LA + la * NXH + kk];
            float2 B = yb < N ? X[yb * LA + lb * NXH + kk] : make_float2(0.f, 0.f);
            const bool self = kk == 0 || 2 * kk == N;
            if (rowline) {   // the x derivative: -i k_x (0 on the Nyquist column)
                const float kx = self && kk != 0 ? 0.f : kfx * (float)kk;
                A = make_float2(kx * A.y, -kx * A.x);
            }
            if (self) { A.y = 0.f; B.y = 0.f; }
            hold[ct][rr] = upper ? make_float2(A.x + B.y, A.y - B.x) : make_float2(A.x - B.y, -(A.y + B.x));
        }
        Dft<R1>::run(hold[ct]);
    }
    __syncthreads();
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int task = (int)threadIdx.x + ct * NT;
        if (task >= NCL * R2) break;
        const int c = task % NCL, i = task / NCL;
#pragma unroll
        for (int q = 0; q < R1; ++q) X[(i * R1 + q) * LC + c] = hold[ct][q];
    }
    __syncthreads();
    // D: x stage 2 -> the real lines A = res.x, B = -res.y at x = i + R1 q, held across
    // a barrier and laid out as the plane's records [y][x](g_x, g_y, g_z) over the
    // exchange; the plane's records are one contiguous run of HBM, copied out in 16-byte
    // pieces
    float2 out[DT][R2];
#pragma unroll
    for (int dt = 0; dt < DT; ++dt) {
        const int task = (int)threadIdx.x + dt * NT;
        if (task >= NCL * R1) break;
        const int i = task % R1, c = task / R1;   // adjacent lanes: adjacent x
#pragma unroll
        for (int rr = 0; rr < R2; ++rr) out[dt][rr] = X[(i + R1 * rr) * LC + c];
#pragma unroll
        for (int rr = 1; rr < R2; ++rr) out[dt][rr] = cmul(out[dt][rr], TW[i * (R2 - 1) + rr - 1]);
        Dft<R2>::run(out[dt]);
    }
    __syncthreads();
    float* rec = (float*)X;
#pragma unroll
    for (int dt = 0; dt < DT; ++dt) {
        const int task = (int)threadIdx.x + dt * NT;
        if (task >= NCL * R1) break;
        const int i = task % R1, c = task / R1;   // adjacent lanes: adjacent x
#pragma unroll
        for (int q = 0; q < R2; ++q) {
            const int x = i + R1 * q;
            if (c < N) {   // row y = c: (g_x, g_y)
                rec[(c * N + x) * 3 + 0] = out[dt][q].x;
                rec[(c * N + x) * 3 + 1] = -out[dt][q].y;
            } else {       // rows 2p and 2p + 1: g_z
                const int y0 = 2 * (c - N);
                rec[(y0 * N + x) * 3 + 2] = out[dt][q].x;
                if (y0 + 1 < N) rec[((y0 + 1) * N + x) * 3 + 2] = -out[dt][q].y;
            }
        }
    }
    __syncthreads();
    const uint32_t rec0 = (uint32_t)f * (uint32_t)r.M + (uint32_t)z * (uint32_t)(N * N);
    if constexpr (N % 2 == 0) {   // 12 rec0 bytes, 16-byte aligned
        float4* dst = (float4*)((char*)r.G3 + ((rec0 << 3) + (rec0 << 2)));
        const float4* src = (const float4*)rec;
        for (int t = threadIdx.x; t < 3 * N * N / 4; t += NT) gst(dst + t, src[t]);
    } else {
        float* dst = (float*)((char*)r.G3 + ((rec0 << 3) + (rec0 << 2)));
        for (int t = threadIdx.x; t < 3 * N * N; t += NT) G(dst)[t] = rec[t];
    }
}
}  // namespace

template <int NT>
__global__ void __launch_bounds__(NT) k_plane_inv(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const AxisPlan& pl = r.ax[1];
    switch (pl.rad[0] * 32 + pl.rad[1]) {   // uniform per block
#define TILD_CASE(A, B) \
    case A * 32 + B: if constexpr ((A * B / 2 + 1) * (A > B ? A : B) <= NT) plane_inv_block<A, B, NT>(r, wi.blk, lds); break;
        TILD_REG_PLANS(TILD_CASE)
#undef TILD_CASE
        default: break;   // excluded by plane_inv_plan on the host
    }
}

bool plane_inv_plan(const AxisPlan& px, const AxisPlan& py) {
    if (!(px.N == py.N && grad_reg_plan(py))) return false;
    const int rmax = py.rad[0] > py.rad[1] ? py.rad[0] : py.rad[1];
    return (py.N / 2 + 1) * rmax <= kPlaneInvThreads && plane_inv_lds_bytes(py) <= kPlaneLds / 2;
}

hipError_t launch_plane_inv(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_plane_inv<kPlaneInvThreads>, dim3(c.n), dim3(kPlaneInvThreads), c.lds, c.reps, c.wmap);
    return hipGetLastError();
}

hipError_t configure_plane_inv() {
    return hipFuncSetAttribute((const void*)k_plane_inv<kPlaneInvThreads>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(kPlaneLds / 2));
}

bool plane_reg_plan(const AxisPlan& px, const AxisPlan& py) {
    return px.N == py.N && grad_reg_plan(px);
}

hipError_t launch_plane_reg(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_plane_reg<kPlaneRegThreads>, dim3(c.n), dim3(kPlaneRegThreads), c.lds, c.reps, c.wmap);
    return hipGetLastError();
}

hipError_t configure_plane_reg() {
    return hipFuncSetAttribute((const void*)k_plane_reg<kPlaneRegThreads>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPlaneLds);
}

bool grad_reg_plan(const AxisPlan& pl) {
    if (pl.nst != 2) return false;
    const int k = pl.rad[0] * 32 + pl.rad[1];
    return reg_plan_pair(k);
}

hipError_t launch_grad_reg(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_grad_reg<kGradRegThreads>, dim3(c.n), dim3(kGradRegThreads), c.lds, c.reps, c.wmap);
    return hipGetLastError();
}

bool last_reg_plan(const AxisPlan& pl, int ntypes, int nf) {
    if (pl.nst != 2 || ntypes < 1 || ntypes > 4 || nf > ntypes + 1) return false;
    if (!last_reg_variant(pl.rad[0], pl.rad[1], ntypes, nf - ntypes)) return false;
    const int k = pl.rad[0] * 32 + pl.rad[1];
    return reg_plan_pair(k);
}

hipError_t launch_last_reg(const LaunchCtx& c, int koff, int cls) {
    if (!c.n) return hipSuccess;
    if (cls == 1) TILD_LAUNCH(c, (k_last_reg<kLastRegThreads, 1>), dim3(c.n), dim3(kLastRegThreads), c.lds, c.reps, c.wmap, koff);
    else TILD_LAUNCH(c, (k_last_reg<kLastRegThreads, 0>), dim3(c.n), dim3(kLastRegThreads), c.lds, c.reps, c.wmap, koff);
    return hipGetLastError();
}

hipError_t configure_last_reg() {
    hipError_t e = hipFuncSetAttribute((const void*)k_last_reg<kLastRegThreads, 0>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPlaneLds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_last_reg<kLastRegThreads, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPlaneLds);
}

}  // namespace tild
