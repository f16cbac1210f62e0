/*
 * ora_util.c -- oracle primitives: Philox4x32-10, Box-Muller normals, FP64 FFT.
 * TEST INFRASTRUCTURE ONLY (see tild_oracle.h).
 *
 * Parity status:
 *   ora_philox4x32_10  pinned: Random123 known-answer vectors (tests/golden/philox_kat.json).
 *   ora_normals        pinned: moments/KS statistics; it replaces cuRAND's XORWOW
 *                      (cuda_random_posits.cu:22-30, device_EM_integrator.cu:27), whose
 *                      values have no golden vectors -- reference noise values are
 *                      "parity unpinned" by design.
 *   ora_fft            pinned: naive DFT on small grids, numpy pocketfft FP64 golden
 *                      vectors, Parseval.  Convention = cuFFT C2C as used at
 *                      pair_style.cu:97/118 (forward e^{-ikr}, unnormalised inverse).
 */
#include "tild_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.14159265358979323846

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
   1, 2, 3", SC'11.  Multipliers and Weyl key increments from that paper. */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform in (0,1) from the top 23 bits, (k + 1/2) / 2^23: exactly representable in
   FP32 and FP64, so the product (FP32) and the oracle (FP64) start Box-Muller from the
   same value. */
static double u23(uint32_t v) { return ((double)(v >> 9) + 0.5) * (1.0 / 8388608.0); }

void ora_normals(uint32_t seed, uint32_t particle, uint32_t step, uint32_t stream, double z[4]) {
    uint32_t ctr[4] = {particle, step, stream, 0u};
    uint32_t key[2] = {seed, 0u};
    uint32_t o[4];
    ora_philox4x32_10(ctr, key, o);
    for (int p = 0; p < 2; ++p) {
        double u1 = u23(o[2 * p]), u2 = u23(o[2 * p + 1]);
        double r = sqrt(-2.0 * log(u1));
        /* angle 2*pi*u2 (Box & Muller 1958) */
        z[2 * p]     = r * cos(2.0 * ORA_PI * u2);
        z[2 * p + 1] = r * sin(2.0 * ORA_PI * u2);
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 mixed-radix Stockham FFT (any N; radices 4,2,3,5,7 then generic primes).
   Stage s with radix R and p = product of earlier radices:
     for i in [0, N/R): k = i mod p; u_r = x[i + r N/R] * w_{pR}^{r k};
     U = DFT_R(u); y[(i-k)R + k + q p] = U_q.                                  */
typedef struct {
    int n, nst, rad[64];
    double* tw[64];   /* per stage: p*R complex twiddles w^{r k}, index k*R + r */
    double* dft[64];  /* per stage: R*R complex DFT matrix, index q*R + r */
    double *buf0, *buf1;
} fft_plan;

static int factorize(int n, int* rad) {
    int k = 0;
    while (n % 4 == 0) { rad[k++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[k++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2) {
        if ((long)p * p > n) { rad[k++] = n; break; }
        while (n % p == 0) { rad[k++] = p; n /= p; }
    }
    return k;
}

static void plan_free(fft_plan* P) {
    for (int s = 0; s < P->nst; ++s) { free(P->tw[s]); free(P->dft[s]); }
    free(P->buf0); free(P->buf1);
}

static int plan_make(fft_plan* P, int n, int sign) {
    memset(P, 0, sizeof(*P));
    P->n = n;
    P->nst = n > 1 ? factorize(n, P->rad) : 0;
    P->buf0 = (double*)malloc(sizeof(double) * 2 * n);
    P->buf1 = (double*)malloc(sizeof(double) * 2 * n);
    if (!P->buf0 || !P->buf1) return -2;
    int p = 1;
    for (int s = 0; s < P->nst; ++s) {
        int R = P->rad[s];
        P->tw[s] = (double*)malloc(sizeof(double) * 2 * (size_t)p * R);
        P->dft[s] = (double*)malloc(sizeof(double) * 2 * (size_t)R * R);
        if (!P->tw[s] || !P->dft[s]) return -2;
        for (int k = 0; k < p; ++k)
            for (int r = 0; r < R; ++r) {
                double ang = sign * 2.0 * ORA_PI * (double)((long)r * k) / (double)((long)p * R);
                P->tw[s][2 * (k * R + r)] = cos(ang);
                P->tw[s][2 * (k * R + r) + 1] = sin(ang);
            }
        for (int q = 0; q < R; ++q)
            for (int r = 0; r < R; ++r) {
                double ang = sign * 2.0 * ORA_PI * (double)((r * q) % R) / (double)R;
                P->dft[s][2 * (q * R + r)] = cos(ang);
                P->dft[s][2 * (q * R + r) + 1] = sin(ang);
            }
        p *= R;
    }
    return 0;
}

/* buf0, buf1: 2n doubles of scratch each (per thread when lines run in parallel) */
static void plan_exec(const fft_plan* P, double* data, double* buf0, double* buf1) {
    int n = P->n;
    if (n <= 1) return;
    double *src = buf0, *dst = buf1;
    memcpy(src, data, sizeof(double) * 2 * n);
    double ur[2 * 64], uo[2 * 64];
    int p = 1;
    for (int s = 0; s < P->nst; ++s) {
        int R = P->rad[s], m = n / R;
        const double* tw = P->tw[s];
        const double* D = P->dft[s];
        for (int i = 0; i < m; ++i) {
            int k = i % p;
            for (int r = 0; r < R; ++r) {
                double a = src[2 * (i + r * m)], b = src[2 * (i + r * m) + 1];
                double c = tw[2 * (k * R + r)], sn = tw[2 * (k * R + r) + 1];
                ur[2 * r] = a * c - b * sn;
                ur[2 * r + 1] = a * sn + b * c;
            }
            for (int q = 0; q < R; ++q) {
                double sr = 0, si = 0;
                for (int r = 0; r < R; ++r) {
                    double c = D[2 * (q * R + r)], sn = D[2 * (q * R + r) + 1];
                    sr += ur[2 * r] * c - ur[2 * r + 1] * sn;
                    si += ur[2 * r] * sn + ur[2 * r + 1] * c;
                }
                uo[2 * q] = sr; uo[2 * q + 1] = si;
            }
            int j = (i - k) * R + k;
            for (int q = 0; q < R; ++q) {
                dst[2 * (j + q * p)] = uo[2 * q];
                dst[2 * (j + q * p) + 1] = uo[2 * q + 1];
            }
        }
        double* t = src; src = dst; dst = t;
        p *= R;
    }
    memcpy(data, src, sizeof(double) * 2 * n);
}

int ora_fft1d(double* data, int n, int sign) {
    if (n <= 0) return -1;
    fft_plan P;
    int rc = plan_make(&P, n, sign);
    for (int s = 0; s < P.nst; ++s) if (P.rad[s] > 64) rc = -3;
    if (rc == 0) plan_exec(&P, data, P.buf0, P.buf1);
    plan_free(&P);
    return rc;
}

int ora_fft(double* data, int dim, const int* Nx, int sign) {
    int N[3] = {Nx[0], dim > 1 ? Nx[1] : 1, dim > 2 ? Nx[2] : 1};
    int rc = 0;
    for (int ax = 0; ax < dim && rc == 0; ++ax) {
        fft_plan P;
        rc = plan_make(&P, N[ax], sign);
        for (int s = 0; s < P.nst; ++s) if (P.rad[s] > 64) rc = -3;
        if (rc == 0) {
            size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)N[0] : (size_t)N[0] * N[1]);
            size_t total = (size_t)N[0] * N[1] * N[2];
            long nlines = (long)(total / N[ax]);
            int n = N[ax];
            /* lines are independent: each thread transforms whole lines with its own
               buffers, so the result does not depend on the thread count */
            #pragma omp parallel
            {
                double* tl = (double*)malloc(sizeof(double) * 6 * (size_t)n);
                #pragma omp for schedule(static)
                for (long l = 0; l < nlines; ++l) {
                    /* line start: enumerate all indices with coordinate ax == 0 */
                    size_t base;
                    if (ax == 0) base = (size_t)l * N[0];
                    else if (ax == 1) base = ((size_t)l % N[0]) + ((size_t)l / N[0]) * (size_t)N[0] * N[1];
                    else base = (size_t)l;
                    double* ln = tl;
                    for (int q = 0; q < n; ++q) {
                        ln[2 * q] = data[2 * (base + q * stride)];
                        ln[2 * q + 1] = data[2 * (base + q * stride) + 1];
                    }
                    plan_exec(&P, ln, tl + 2 * (size_t)n, tl + 4 * (size_t)n);
                    for (int q = 0; q < n; ++q) {
                        data[2 * (base + q * stride)] = ln[2 * q];
                        data[2 * (base + q * stride) + 1] = ln[2 * q + 1];
                    }
                }
                free(tl);
            }
        }
        plan_free(&P);
    }
    return rc;
}

/* Worker threads of the oracle's parallel loops (FFT lines, per-node and per-site
   loops).  Every parallel loop writes disjoint outputs and every reduction stays
   serial, so results are bit-identical for any thread count. */
void ora_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int ora_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
