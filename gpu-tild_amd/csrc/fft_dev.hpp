// fft_dev.hpp -- in-LDS mixed-radix Stockham FFT executed by one thread block.
//
// Data: nb independent lines of length N, element n of line b at buf[n*ldb + b].
// Only the forward transform (e^{-2 pi i nk/N}) is implemented here; callers form
// the inverse as conj(FFT(conj(x))) by conjugating on load and store.
//
// Each stage (radix R, p = product of earlier radices) reads one LDS buffer and
// writes the other (Stockham autosort, no bit reversal); a thread keeps only one
// butterfly in registers at a time:
//     u_r = x[i + r N/R] w_{pR}^{r k},  k = i mod p,   y[(i - k) R + k + q p] = DFT_R(u)_q.
#pragma once
#include "device.hpp"

namespace tild {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }   // a * (-i)
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Forward DFT of small sizes (2, 3, 4, 5, 7, 8, 10, 16), in place on v[0..R-1].
template <int R> struct Dft;

template <> struct Dft<2> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    }
};
template <> struct Dft<3> {
    __device__ __forceinline__ static void run(float2* v) {
        const float s3 = 0.86602540378443864676f;   // sin(2 pi / 3)
        float2 t = cadd(v[1], v[2]);
        float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        float2 d = csub(v[1], v[2]);
        float2 e = make_float2(s3 * d.y, -s3 * d.x);  // -i sin(2pi/3) d
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, e);
        v[2] = csub(m, e);
    }
};
template <> struct Dft<4> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
    }
};
template <> struct Dft<8> {
    __device__ __forceinline__ static void run(float2* v) {
        const float h = 0.70710678118654752440f;
        float2 a[4] = {v[0], v[2], v[4], v[6]}, b[4] = {v[1], v[3], v[5], v[7]};
        Dft<4>::run(a); Dft<4>::run(b);
        // twiddles w8^k, k = 0..3: 1, (h,-h), -i, (-h,-h)
        float2 b1 = make_float2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
        float2 b2 = mul_mi(b[2]);
        float2 b3 = make_float2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
        v[0] = cadd(a[0], b[0]); v[4] = csub(a[0], b[0]);
        v[1] = cadd(a[1], b1);   v[5] = csub(a[1], b1);
        v[2] = cadd(a[2], b2);   v[6] = csub(a[2], b2);
        v[3] = cadd(a[3], b3);   v[7] = csub(a[3], b3);
    }
};
// 16 = 4 x 4: four DFT-4 over the stride-4 subsequences, twiddles w16^(n2 k1), four
// DFT-4 across them; output k1 + 4 k2.
template <> struct Dft<16> {
    __device__ __forceinline__ static void run(float2* v) {
        const float c1 = 0.92387953251128675613f;   // cos(pi/8)
        const float s1 = 0.38268343236508977173f;   // sin(pi/8)
        const float h = 0.70710678118654752440f;
        float2 y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            float2 a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
            Dft<4>::run(a);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
        }
        // w16^1 = (c1,-s1), w16^2 = (h,-h), w16^3 = (s1,-c1), w16^4 = -i, w16^6 = (-h,-h), w16^9 = (-c1,s1)
        y[1][1] = cmul(y[1][1], make_float2(c1, -s1));
        y[1][2] = make_float2(h * (y[1][2].x + y[1][2].y), h * (y[1][2].y - y[1][2].x));
        y[1][3] = cmul(y[1][3], make_float2(s1, -c1));
        y[2][1] = make_float2(h * (y[2][1].x + y[2][1].y), h * (y[2][1].y - y[2][1].x));
        y[2][2] = mul_mi(y[2][2]);
        y[2][3] = make_float2(h * (y[2][3].y - y[2][3].x), -h * (y[2][3].x + y[2][3].y));
        y[3][1] = cmul(y[3][1], make_float2(s1, -c1));
        y[3][2] = make_float2(h * (y[3][2].y - y[3][2].x), -h * (y[3][2].x + y[3][2].y));
        y[3][3] = cmul(y[3][3], make_float2(-c1, s1));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            Dft<4>::run(a);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k2];
        }
    }
};

// Odd prime radices: real/imaginary symmetric pairs (Rader-free direct form).
template <int R> struct DftOdd {
    __device__ __forceinline__ static void run(float2* v, const float* c, const float* s) {
        // c[j] = cos(2 pi j / R), s[j] = sin(2 pi j / R), j = 1..(R-1)/2
        float2 out[R];
        float2 sum = v[0];
#pragma unroll
        for (int j = 1; j < R; ++j) sum = cadd(sum, v[j]);
        out[0] = sum;
#pragma unroll
        for (int q = 1; q <= (R - 1) / 2; ++q) {
            float2 re = v[0], im = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 1; j <= (R - 1) / 2; ++j) {
                int e = (j * q) % R;
                int ee = e <= (R - 1) / 2 ? e : R - e;
                float cc = c[ee];
                float ss = e <= (R - 1) / 2 ? s[ee] : -s[ee];
                float2 ap = cadd(v[j], v[R - j]), am = csub(v[j], v[R - j]);
                re.x += cc * ap.x; re.y += cc * ap.y;
                im.x += ss * am.x; im.y += ss * am.y;
            }
            // X_q = re - i*im ; X_{R-q} = re + i*im
            out[q] = make_float2(re.x + im.y, re.y - im.x);
            out[R - q] = make_float2(re.x - im.y, re.y + im.x);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = out[j];
    }
};
template <> struct Dft<5> {
    __device__ __forceinline__ static void run(float2* v) {
        const float c[3] = {1.f, 0.30901699437494742410f, -0.80901699437494742410f};
        const float s[3] = {0.f, 0.95105651629515357212f, 0.58778525229247312917f};
        DftOdd<5>::run(v, c, s);
    }
};
// 10 = 2 x 5 by the prime-factor mapping (2 and 5 are coprime, so no twiddles
// between the two): inputs n = (5 n1 + 2 n2) mod 10 form five DFT-2 pairs, then two
// DFT-5 over n2 give outputs k = (5 k1 + 6 k2) mod 10.
template <> struct Dft<10> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a[5], b[5];
#pragma unroll
        for (int n2 = 0; n2 < 5; ++n2) {
            const float2 x0 = v[(2 * n2) % 10], x1 = v[(5 + 2 * n2) % 10];
            a[n2] = cadd(x0, x1);   // k1 = 0
            b[n2] = csub(x0, x1);   // k1 = 1
        }
        Dft<5>::run(a);
        Dft<5>::run(b);
#pragma unroll
        for (int k2 = 0; k2 < 5; ++k2) {
            v[(6 * k2) % 10] = a[k2];
            v[(5 + 6 * k2) % 10] = b[k2];
        }
    }
};
template <> struct Dft<7> {
    __device__ __forceinline__ static void run(float2* v) {
        const float c[4] = {1.f, 0.62348980185873353053f, -0.22252093395631440429f, -0.90096886790241912624f};
        const float s[4] = {0.f, 0.78183148246802980871f, 0.97492791218182360702f, 0.43388373911755812048f};
        DftOdd<7>::run(v, c, s);
    }
};

template <int R, int NT>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ src, float2* __restrict__ dst, int ldb, int nb,
                                 int N, int p, const float2* __restrict__ tw) {
    const int m = N / R, nbf = m * nb;
    const FastDiv dnb(nb), dvp(p);
    const int sst = m * ldb, dst_st = p * ldb;   // uniform strides: one add per LDS address
    // element offsets stay below 2^24 (LDS), so the 24-bit multiply is exact
    for (int task = threadIdx.x; task < nbf; task += NT) {
        int b, k;
        const int i = dnb.div(task, b);
        dvp.div(i, k);
        const float2* sp = src + ((int)__umul24((unsigned)i, (unsigned)ldb) + b);
        float2* op = dst + ((int)__umul24((unsigned)((i - k) * R + k), (unsigned)ldb) + b);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sp[r * sst];
        if (p > 1) {
            const float2* w = tw + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], gld(w + r - 1));
        }
        Dft<R>::run(v);
#pragma unroll
        for (int q = 0; q < R; ++q) op[q * dst_st] = v[q];
    }
}

// The same stage with the line length N and the span P of the earlier radices known
// at compile time: k = i mod P divides by a constant, and m = N / R is a constant.
// Same arithmetic, in the same order, as fft_stage.
template <int R, int NT, int N, int P>
__device__ __forceinline__ void fft_stage_c(const float2* __restrict__ src, float2* __restrict__ dst, int ldb, int nb,
                                            const float2* __restrict__ tw) {
    constexpr int m = N / R;
    const int nbf = m * nb;
    const FastDiv dnb(nb);
    const int sst = m * ldb, dst_st = P * ldb;
    for (int task = threadIdx.x; task < nbf; task += NT) {
        int b;
        const int i = dnb.div(task, b);
        const int k = i % P;
        const float2* sp = src + ((int)__umul24((unsigned)i, (unsigned)ldb) + b);
        float2* op = dst + ((int)__umul24((unsigned)((i - k) * R + k), (unsigned)ldb) + b);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sp[r * sst];
        if constexpr (P > 1) {
            const float2* w = tw + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], gld(w + r - 1));
        }
        Dft<R>::run(v);
#pragma unroll
        for (int q = 0; q < R; ++q) op[q * dst_st] = v[q];
    }
}

// Two-stage transform with both radices known at compile time (N = R1 R2, the plans
// of the 80- and 160-point axes): no plan is read and no radix switched on per stage.
// Returns A, as lds_fft does for two stages.
template <int R1, int R2, int NT>
__device__ __forceinline__ float2* lds_fft2(float2* A, float2* B, int ldb, int nb, const AxisPlan& pl) {
    fft_stage_c<R1, NT, R1 * R2, 1>(A, B, ldb, nb, pl.tw + pl.tw_off[0]);
    __syncthreads();
    fft_stage_c<R2, NT, R1 * R2, R1>(B, A, ldb, nb, pl.tw + pl.tw_off[1]);
    __syncthreads();
    return A;
}

// Forward FFT of nb lines held in A (element n of line b at A[n*ldb + b]), using B
// as the ping-pong partner.  Requires N*nb <= kFftBudget and a synchronised block
// (A fully written).  Returns the buffer holding the result (A or B); every stage
// ends with a barrier, so the result is visible to the whole block on return.
// Always inlined: the buffers are then known to be LDS (ds_* instructions, not flat).
// NT = threads of the block.
template <int NT = kFftThreads>
__device__ __forceinline__ float2* lds_fft(float2* A, float2* B, int ldb, int nb, const AxisPlan& pl) {
    if (pl.nst == 2) {   // the plans of 80 and 160 points and their reverses (uniform branch)
        if (pl.rad[0] == 10 && pl.rad[1] == 8) return lds_fft2<10, 8, NT>(A, B, ldb, nb, pl);
        if (pl.rad[0] == 16 && pl.rad[1] == 10) return lds_fft2<16, 10, NT>(A, B, ldb, nb, pl);
        if (pl.rad[0] == 8 && pl.rad[1] == 10) return lds_fft2<8, 10, NT>(A, B, ldb, nb, pl);
        if (pl.rad[0] == 10 && pl.rad[1] == 16) return lds_fft2<10, 16, NT>(A, B, ldb, nb, pl);
    }
    float2* src = A;
    float2* dst = B;
    int p = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        const float2* tw = pl.tw + pl.tw_off[s];
        switch (R) {
            case 2: fft_stage<2, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 3: fft_stage<3, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 4: fft_stage<4, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 5: fft_stage<5, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 7: fft_stage<7, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 10: fft_stage<10, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            case 16: fft_stage<16, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
            default: fft_stage<8, NT>(src, dst, ldb, nb, pl.N, p, tw); break;
        }
        __syncthreads();
        float2* t = src; src = dst; dst = t;
        p *= R;
    }
    return src;
}

}  // namespace tild
