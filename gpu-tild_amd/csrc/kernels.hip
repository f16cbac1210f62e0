// kernels.hip -- per-step kernels of the TILD engine (CDNA4 / gfx950).
//
// Step pipeline for each replica (3D; 2D drops the middle passes):
//   k_spread        particles -> fixed-point density accumulators (deterministic)
//   k_x_r2c         accumulators -> rho (real) -> x-axis R2C half spectra H
//   k_mid (fwd)     y-axis FFT of H
//   k_last_fused    z-axis FFT, k-space coupling sum_J C_IJ(k) rho_J(k), -i k_d,
//                   inverse z-axis FFT -> Gk
//   k_mid (inv)     inverse y-axis FFT of Gk
//   k_x_c2r         inverse x-axis C2R -> float4 (g_x, g_y, g_z, rho) records
//   k_particles     gather + bonds + EM/GJF update (double-buffered positions); on
//                   LDS-tiled replicas also the next step's bucket claims
// Log steps add k_kprops (energies/virial from the spectra), k_bond_props and
// k_reduce (fixed-order FP64 sums).
#include "device.hpp"
#include "fft_dev.hpp"
#include "launch.hpp"
#include "tile_dev.hpp"

namespace tild {

// ------------------------------------------------------------------ spread
// One thread per particle deposits its (P+1)^D stencil into 64-bit fixed-point
// accumulators with integer atomics: integer addition is associative, so the
// result is independent of thread order (unlike float atomicAdd,
// device_grid_utils.cu:173-331).
__global__ void __launch_bounds__(256) k_spread(const RepDev* __restrict__ reps,
                                                const WorkItem* __restrict__ wmap, int koff) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int i = wi.blk * 256 + threadIdx.x;
    if (i >= r.ns) return;
    const int D = r.dim, P = r.P;
    const float* x = r.X[(r.step + koff) & 1];
    float W[3][5];
    int n0[3];
    n0[2] = 0; W[2][0] = 1.f;
    for (int d = 0; d < D; ++d) {
        int N = d == 0 ? r.Nx : (d == 1 ? r.Ny : r.Nz);
        n0[d] = stencil_1d(x[i * D + d], r.dx[d], r.idx[d], N, P, W[d]);
    }
    const int t = r.tp[i];
    unsigned long long* acc = r.acc + (size_t)t * r.M;
    const bool charged = r.q != nullptr;
    const float qi = charged ? r.q[i] : 0.f;
    unsigned long long* accq = r.acc + (size_t)r.ntypes * r.M;
    const int nz = D == 3 ? P + 1 : 1;
    // order 1 in the integers of the tiled deposit: lower axis weights rounded to 2^22,
    // upper ones their complements (fixed_lower, fixed_weight2/3, fixed_charge)
    uint32_t q[3][2];
    float fq[3][2];
    if (P == 1) {
        for (int d = 0; d < 3; ++d) {
            q[d][0] = d < D ? fixed_lower(W[d][0]) : kFixedOne;
            q[d][1] = kFixedOne - q[d][0];
            fq[d][0] = d < D ? (float)q[d][0] : 1.f;
            fq[d][1] = (float)q[d][1];
        }
    }
    const float qs = qi * r.fixq_s1;
    int iz0 = n0[2];
    for (int iz = 0, nzz = iz0; iz < nz; ++iz, nzz = wrap_next(nzz, r.Nz)) {
        for (int iy = 0, nyy = n0[1]; iy <= P; ++iy, nyy = wrap_next(nyy, r.Ny)) {
            const int rowb = (nzz * r.Ny + nyy) * r.Nx;
            for (int ix = 0, nxx = n0[0]; ix <= P; ++ix, nxx = wrap_next(nxx, r.Nx)) {
                if (P == 1) {
                    const unsigned long long v = D == 3 ? fixed_weight3(fixed_pair(q[0][ix], q[1][iy]), q[2][iz], r.fix_s3)
                                                        : fixed_weight2(q[0][ix], q[1][iy], r.fix_s2);
                    atomicAdd(acc + rowb + nxx, v);
                    if (charged) atomicAdd(accq + rowb + nxx, fixed_charge(fq[0][ix], fq[1][iy], fq[2][iz], qs));
                    continue;
                }
                // Wx*Wy*Wz in the reference's product order (device_grid_utils.cu:329)
                const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                atomicAdd(acc + rowb + nxx, (unsigned long long)__float2ll_rn(w * r.fix_scale));
                if (charged) {
                    long long vq = __float2ll_rn((w * qi) * r.fixq_scale);
                    atomicAdd(accq + rowb + nxx, (unsigned long long)vq);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ x R2C
// Two real rows a, b are transformed as one complex row z = a + ib; then
// A_k = (Z_k + conj Z_{-k})/2, B_k = (Z_k - conj Z_{-k})/(2i), k in [0, N/2].
// FROM_ACC: rows come from the fixed-point accumulators (converted, written to rho
// and cleared); otherwise from rho.
template <bool FROM_ACC>
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(8))) k_x_r2c(const RepDev* __restrict__ reps,
                                                       const WorkItem* __restrict__ wmap,
                                                       int nrows_override, const float* __restrict__ in_override,
                                                       float2* __restrict__ out_override) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int N = r.Nx, Nxh = r.Nxh;
    const int rows_per_field = r.M / r.Nx;
    const int nrows = nrows_override > 0 ? nrows_override : r.nf * rows_per_field;
    const int npairs = (nrows + 1) / 2;
    const int nl = r.nl_x;
    const int p0 = wi.blk * nl;
    const int nb = min(nl, npairs - p0);
    const int ldb = nb | 1;
    gmem<const float> rin = G(in_override ? in_override : (const float*)r.rho);
    float2* Hout = out_override ? out_override : r.H;
    gmem<unsigned long long> acc = G(r.acc);
    gmem<float> rho_out = G(r.rho);
    const FastDiv dN(N);
    const int type_rows = r.ntypes * rows_per_field;   // rows below are particle densities
    // read once: a lane-dependent choice between the two fields is compiled into a
    // vector load of the replica record
    const double inv_t = r.fix_inv, inv_q = r.fixq_inv;
    for (int task = threadIdx.x; task < nb * N; task += kFftThreads) {
        int n;
        const int b = dN.div(task, n);
        const int ra = 2 * (p0 + b), rb = ra + 1;
        float a, bb = 0.f;
        if (FROM_ACC) {
            size_t ia = (size_t)ra * N + n;
            a = (float)((double)(long long)acc[ia] * (ra < type_rows ? inv_t : inv_q));
            acc[ia] = 0ull;
            rho_out[ia] = a;
            if (rb < nrows) {
                size_t ib = (size_t)rb * N + n;
                bb = (float)((double)(long long)acc[ib] * (rb < type_rows ? inv_t : inv_q));
                acc[ib] = 0ull;
                rho_out[ib] = bb;
            }
        } else {
            a = rin[(size_t)ra * N + n];
            if (rb < nrows) bb = rin[(size_t)rb * N + n];
        }
        lds[n * ldb + b] = make_float2(a, bb);
    }
    __syncthreads();
    const float2* res = lds_fft(lds, lds + (size_t)N * ldb, ldb, nb, r.ax[0]);
    const FastDiv dNxh(Nxh);
    for (int task = threadIdx.x; task < nb * Nxh; task += kFftThreads) {
        int k;
        const int b = dNxh.div(task, k);
        const float2 Zk = res[k * ldb + b];
        const float2 Zm = res[(k == 0 ? 0 : N - k) * ldb + b];
        const int ra = 2 * (p0 + b), rb = ra + 1;
        gst(Hout + (size_t)ra * Nxh + k, make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y)));
        if (rb < nrows) gst(Hout + (size_t)rb * Nxh + k, make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x)));
    }
}

// ------------------------------------------------------------------ strided pass
// FFT along axis `axis` (1 = y, 2 = z) of nfb half-spectrum fields stored in buf
// (x fastest, Nxh columns).  Lines are batched over the contiguous columns so that
// every global access is a run of consecutive complex values.
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(8))) k_mid(const RepDev* __restrict__ reps,
                                                     const WorkItem* __restrict__ wmap,
                                                     int axis, int use_gk, int nfb_override,
                                                     int inverse, float2* __restrict__ buf_override) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    MidGeom g = mid_geom(r, axis, nfb_override > 0 ? nfb_override : (use_gk ? r.nf * r.dim : r.nf));
    float2* buf = buf_override ? buf_override : (use_gk ? r.Gk : r.H);
    const int o = wi.blk / g.nchunk, c = wi.blk % g.nchunk;
    const int b0 = c * g.nl;
    const int nb = min(g.nl, g.batch - b0);
    const int ldb = nb | 1;
    float2* base = buf + (size_t)o * g.ostride + b0;
    const FastDiv dnb(nb);
    for (int task = threadIdx.x; task < nb * g.N; task += kFftThreads) {
        int b;
        const int n = dnb.div(task, b);
        float2 v = gld(base + (size_t)n * g.lstride + b);
        lds[n * ldb + b] = inverse ? conjf2(v) : v;
    }
    __syncthreads();
    const float2* res = lds_fft(lds, lds + (size_t)g.N * ldb, ldb, nb, r.ax[axis]);
    for (int task = threadIdx.x; task < nb * g.N; task += kFftThreads) {
        int b;
        const int n = dnb.div(task, b);
        float2 v = res[n * ldb + b];
        gst(base + (size_t)n * g.lstride + b, inverse ? conjf2(v) : v);
    }
}

// ------------------------------------------------------------------ fused last axis

// k-space coupling of one block's nodes for TT types (the same sums, in the same order,
// as the general loop of k_last_fused): w_I = sum_J C_IJ rho_J with the symmetric table
// rows pidx(lo, hi) = lo*TT - lo*(lo-1)/2 + (hi - lo), and the charge row.  Ramped
// styles add their rows, weighted, to their pairs' coefficients.
template <int TT, int NT>
__device__ __forceinline__ void couple_nodes(gmem<const float> cb, int nrows, int nlc, int nb, int N, int ldb, int ldw,
                                             int nf, int ctab_charge, const FastDiv& dnb, const float2* A, float2* Wb,
                                             const RepDev& r, const float* rw) {
    constexpr int NR = TT * (TT + 1) / 2;
    const bool charged = nf > TT;
    Digits2 gc(threadIdx.x, NT, dnb);
    for (int task = threadIdx.x; task < nb * N; task += NT, gc.step()) {
        const int b = gc.lo, n = gc.hi;
        const int row = n * ldb + b, wrow = n * ldw + b;
        gmem<const float> cn = cb + (n * nrows * nlc + b);   // offsets inside the block's run
        float c[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) c[k] = cn[k * nlc];
        const float cq = charged ? cn[ctab_charge * nlc] : 0.f;
        for (int j = 0; j < r.nramp; ++j) {   // uniform trip count
            const float v = rw[j] * cn[(r.ramp_row0 + j) * nlc];
#pragma unroll
            for (int k = 0; k < NR; ++k) c[k] += r.ramp_pair[j] == k ? v : 0.f;
        }
        float2 rh[TT];
#pragma unroll
        for (int J = 0; J < TT; ++J) rh[J] = A[row + J * nb];
#pragma unroll
        for (int I = 0; I < TT; ++I) {
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int J = 0; J < TT; ++J) {
                const int lo = I < J ? I : J, hi = I < J ? J : I;
                const float cc = c[lo * TT - lo * (lo - 1) / 2 + (hi - lo)];
                acc.x += cc * rh[J].x; acc.y += cc * rh[J].y;
            }
            Wb[wrow + I * nb] = acc;
        }
        if (charged) {
            const float2 rq = A[row + TT * nb];
            Wb[wrow + TT * nb] = make_float2(cq * rq.x, cq * rq.y);
        }
    }
}

// Last-axis forward FFT of every field, the k-space coupling
//   w_I(k) = sum_J C_IJ(k) rho_J(k)       (C includes 1/M and counts each style in
//                                          both orientations, pair_style.cu:81-196)
//   w_c(k) = G_c(k) rho_c(k)              (smeared charges, double smearing,
//                                          device_array_utils.cu:199-241)
// then the gradient spectra -i k_d w(k) (k_d = 0 on the d-Nyquist plane, where the
// reference's term is anti-Hermitian and drops out of Re IFFT) start their inverse
// transform: along the last axis la, w and -i k_la w go to Gk[f*dim] and
// Gk[f*dim + la]; the factors of the other directions are constant along those lines
// and are applied by the passes of their own axes.
template <int NT>
__global__ void __launch_bounds__(NT) k_last_fused(const RepDev* __restrict__ reps,
                                                            const WorkItem* __restrict__ wmap,
                                                            int store_spectrum, int koff) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int la = r.dim - 1;                       // last axis
    const int N = la == 2 ? r.Nz : r.Ny;
    const int Q = la == 2 ? r.Ny * r.Nxh : r.Nxh;   // plane size = line stride
    const int nf = r.nf, T = r.ntypes;
    const int q0 = wi.blk * r.nl_last;
    const int nb = min(r.nl_last, Q - q0);
    const int nlines = nf * nb;
    const int ldb = nlines | 1;                          // forward lines
    const int ldw = nlines;                              // coupled lines (unpadded)
    const int D = r.dim, DB = r.last_dirs;               // inverse outputs per batch (1 or 2)
    const int ldg = (nlines * DB) | 1;                   // inverse lines of one batch
    // LDS: Wb [N x ldw] coupled fields, then two [N x ldg] buffers that first hold the
    // forward transform's ping-pong pair and then the batched inverse transform's
    float2* Wb = lds;
    float2* E0 = lds + (size_t)N * ldw;
    float2* E1 = E0 + (size_t)N * ldg;
    float* rw = (float*)(E1 + (size_t)N * ldg);          // ramp weights of this step
    if ((int)threadIdx.x < r.nramp) rw[threadIdx.x] = ramp_weight(r, threadIdx.x, r.step + koff);
    const FastDiv dnb(nb), dnf(nf);
    const bool o32 = r.grid_off32;   // field offsets fit 32-bit byte offsets
    Digits3 gl(threadIdx.x, NT, dnb, dnf);
    for (int task = threadIdx.x; task < nlines * N; task += NT, gl.step()) {
        const int b = gl.lo, f = gl.mid, n = gl.hi;
        E0[n * ldb + f * nb + b] =
            o32 ? gld(at32((const float2*)r.H, (uint32_t)f * (uint32_t)r.Mh + (uint32_t)(n * Q + q0 + b)))
                : gld(r.H + (size_t)f * r.Mh + (size_t)n * Q + q0 + b);
    }
    __syncthreads();
    const float2* A = lds_fft<NT>(E0, E1, ldb, nlines, r.ax[la]);   // spectra rho_f(k)
    // coupling w_I = sum_J C_IJ rho_J (C symmetric, rows I <= J), charge w_c = G_c rho_c;
    // this block's table rows are one contiguous run
    const int nrows = r.ctab_rows, nlc = r.nl_last;

    gmem<const float> cb = G(r.Ctab) + (size_t)wi.blk * N * nrows * nlc;
    switch (T) {   // up to four types: unrolled, every table row of a node requested at once
        case 1: couple_nodes<1, NT>(cb, nrows, nlc, nb, N, ldb, ldw, nf, r.ctab_charge, dnb, A, Wb, r, rw); break;
        case 2: couple_nodes<2, NT>(cb, nrows, nlc, nb, N, ldb, ldw, nf, r.ctab_charge, dnb, A, Wb, r, rw); break;
        case 3: couple_nodes<3, NT>(cb, nrows, nlc, nb, N, ldb, ldw, nf, r.ctab_charge, dnb, A, Wb, r, rw); break;
        case 4: couple_nodes<4, NT>(cb, nrows, nlc, nb, N, ldb, ldw, nf, r.ctab_charge, dnb, A, Wb, r, rw); break;
        default: break;
    }
    for (int task = threadIdx.x; T > 4 && task < nb * N; task += NT) {
        int b;
        const int n = dnb.div(task, b);
        const int row = n * ldb + b, wrow = n * ldw + b;
        for (int I = 0; I < T; ++I) {
            float2 acc = make_float2(0.f, 0.f);
            for (int J = 0; J < T; ++J) {
                const int lo = I < J ? I : J, hi = I < J ? J : I;
                const int pidx = lo * T - lo * (lo - 1) / 2 + (hi - lo);
                float c = cb[(n * nrows + pidx) * nlc + b];
                for (int j = 0; j < r.nramp; ++j)
                    if (r.ramp_pair[j] == pidx) c += rw[j] * cb[(n * nrows + r.ramp_row0 + j) * nlc + b];
                const float2 rh = A[row + J * nb];
                acc.x += c * rh.x; acc.y += c * rh.y;
            }
            Wb[wrow + I * nb] = acc;
        }
        if (nf > T) {
            const float c = cb[(n * nrows + r.ctab_charge) * nlc + b];
            const float2 rh = A[row + T * nb];
            Wb[wrow + T * nb] = make_float2(c * rh.x, c * rh.y);
        }
    }
    __syncthreads();
    // Two inverse lines per field: w itself, and -i k_la w (k = 0 on the Nyquist plane).
    // The factors -i k_d of the other directions are constant along these lines, so
    // the inverse transform of w serves all of them; each is applied in the pass of its
    // own axis (k_mid_grad for y, k_x_c2r for x).  DB outputs per batch.
    const int slot1 = la;                                // Gk slot of the -i k_la w line
    for (int o0 = 0; o0 < 2; o0 += DB) {
        const int no = min(DB, 2 - o0), nlg = nlines * no;
        Digits2 gp(threadIdx.x, NT, dnb);
        for (int task = threadIdx.x; task < nb * N; task += NT, gp.step()) {
            const int b = gp.lo, n = gp.hi;
            const float kl = 2 * n == N ? 0.f : r.kfac[la] * (float)(2 * n < N ? n : n - N);
            for (int oo = 0; oo < no; ++oo) {
                const float kd = o0 + oo == 0 ? 1.f : kl;
                for (int f = 0; f < nf; ++f) {
                    const float2 w = Wb[n * ldw + f * nb + b];
                    // the inverse is taken as conj(FFT(conj)): conj(w) = (w.x, -w.y) and
                    // conj(-i k w) = (k w.y, k w.x)
                    E0[n * ldg + (f * no + oo) * nb + b] =
                        o0 + oo == 0 ? make_float2(w.x, -w.y) : make_float2(kd * w.y, kd * w.x);
                }
            }
        }
        __syncthreads();
        const float2* res = lds_fft<NT>(E0, E1, ldg, nlg, r.axr);   // two-stage plans reversed
        Digits3 gs(threadIdx.x, NT, dnb, FastDiv(nf * no));
        for (int task = threadIdx.x; task < nlg * N; task += NT, gs.step()) {
            const int b = gs.lo, fo = gs.mid, n = gs.hi;
            const int oo = no == 2 ? (fo & 1) : 0, f = no == 2 ? (fo >> 1) : fo;   // no is 1 or 2
            const int slot = o0 + oo == 0 ? 0 : slot1;
            const float2 v = conjf2(res[n * ldg + fo * nb + b]);
            if (o32) gst(at32(r.Gk, (uint32_t)(f * D + slot) * (uint32_t)r.Mh + (uint32_t)(n * Q + q0 + b)), v);
            else gst(r.Gk + (size_t)(f * D + slot) * r.Mh + (size_t)n * Q + q0 + b, v);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ y inverse (3D)
// y-axis inverse of the gradient spectra.  Per field, slot 0 holds U (the last-axis
// inverse of w) and slot 2 holds V (that of -i k_z w); the pass writes the y-inverses
// of U (the x derivative: -i k_x follows in the x pass), of -i k_y U and of V to slots
// 0, 1 and 2.  One block: nb columns of one (field, z) plane, all three line sets in
// one batch; it reads its columns of slots 0 and 2 before it writes any slot.
template <int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8)))
k_mid_grad(const RepDev* __restrict__ reps, const WorkItem* __restrict__ wmap) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const MidGeom g = mid_grad_geom(r);
    const int o = wi.blk / g.nchunk, c = wi.blk % g.nchunk;   // o = f * Nz + z
    const int f = o / r.Nz, z = o - f * r.Nz;
    const int N = r.Ny, Nxh = r.Nxh;
    const int b0 = c * g.nl;
    const int nb = min(g.nl, g.batch - b0);
    const int ldb = (3 * nb) | 1;
    const size_t plane = ((size_t)z * N) * Nxh + b0;
    float2* U = r.Gk + (size_t)(f * 3 + 0) * r.Mh + plane;
    float2* Y = r.Gk + (size_t)(f * 3 + 1) * r.Mh + plane;
    float2* V = r.Gk + (size_t)(f * 3 + 2) * r.Mh + plane;
    const FastDiv dnb(nb);
    Digits2 gl(threadIdx.x, NT, dnb);
    for (int task = threadIdx.x; task < nb * N; task += NT, gl.step()) {
        const int b = gl.lo, n = gl.hi;
        const float ky = 2 * n == N ? 0.f : r.kfac[1] * (float)(2 * n < N ? n : n - N);
        const float2 u = gld(U + (n * Nxh + b)), v = gld(V + (n * Nxh + b));
        // lines of the conj(FFT(conj)) inverse: conj(u), conj(-i ky u), conj(v)
        lds[n * ldb + b] = make_float2(u.x, -u.y);
        lds[n * ldb + nb + b] = make_float2(ky * u.y, ky * u.x);
        lds[n * ldb + 2 * nb + b] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    const float2* res = lds_fft<NT>(lds, lds + (size_t)N * ldb, ldb, 3 * nb, r.ax[1]);
    Digits3 gs(threadIdx.x, NT, dnb, FastDiv(3));
    for (int task = threadIdx.x; task < 3 * nb * N; task += NT, gs.step()) {
        const int b = gs.lo, sl = gs.mid, n = gs.hi;
        float2* out = sl == 0 ? U : (sl == 1 ? Y : V);
        gst(out + (n * Nxh + b), conjf2(res[n * ldb + sl * nb + b]));
    }
}

// ------------------------------------------------------------------ x C2R
// Inverse x-axis transform of half-spectrum rows.  Two real output lines share one
// complex FFT: Z_k = A_k + i B_k with the Hermitian extension Z_{N-k} = conj A_k +
// i conj B_k; the imaginary parts of the self-conjugate columns (k = 0, N/2) are
// dropped, which is the Hermitian projection the reference's Re(C2C inverse) takes.
// RECORDS: output float4 (g_0, g_1, g_2, rho_f) per field and node (fields are lines
// f*dim + d of Gk); otherwise plain real rows to out_override.
// The gather skips nodes whose density of the bead's type is not positive
// (device_grid_utils.cu:424-452).  Without field phases that test holds by construction
// at every node whose weight the deposit resolves: the bead's own deposit makes that
// node's density of its type positive, the weights being never negative.  A node it
// does not resolve (order 1: an axis weight below 2^-23; any order: a node weight below
// 2^-F, F the fractional bits of the fixed-point densities, at least 23 below 10^11
// beads) adds under 2^-23 of its record to the force.  So the records of these passes
// carry no mask and the pass reads no densities (8 MB a step on the 640,000-bead melt);
// the field-phase path, which needs rho itself, keeps the test.
// The gather records of one block of the inverse x pass (see k_x_c2r): D lines per
// field unrolled, buffers addressed by 32-bit offsets; no field phases.
template <int D, int NT>
__device__ __forceinline__ void c2r_records(const RepDev& r, const float2* res, int ldb, int nc, int rows, int row0,
                                            int N, const FastDiv& dN, const FastDiv& dnf) {
    for (int task = threadIdx.x; task < rows * r.nf * N; task += NT) {
        int x, f;
        const int rr = dnf.div(dN.div(task, x), f);
        float g[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int line = f * D + d;
            const float2 z = res[x * ldb + rr * nc + line / 2];
            g[d] = (line & 1) ? -z.y : z.x;      // conj of the forward result
        }
        const uint32_t fn = (uint32_t)f * (uint32_t)r.M + (uint32_t)(row0 + rr) * (uint32_t)N + (uint32_t)x;
        // no density mask (see above)
        gst3((float*)((char*)r.G3 + ((fn << 3) + (fn << 2))), g[0], g[1], g[2]);   // 12 fn bytes
    }
}

// The gather-record form of the inverse x pass for NF fields in D dimensions (buffers
// below 4 GiB, no field phases), with the line bookkeeping known at compile time: the
// same loads, arithmetic and stores as the general body of k_x_c2r (bit-identical
// records), with one run-time division per task instead of two and no run-time line
// masks.
template <int NF, int D, int NT>
__device__ __forceinline__ void x_c2r_records(const RepDev& r, const WorkItem wi, float2* lds) {
    constexpr int NL = NF * D, NC = (NL + 1) / 2;
    const int N = r.Nx, Nxh = r.Nxh;
    const int rows_per_field = r.M / r.Nx;
    const int row0 = wi.blk * r.nr_c2r;
    const int rows = min(r.nr_c2r, rows_per_field - row0);
    const int nb = rows * NC;
    const int ldb = nb | 1;
    const float2* Gin = (const float2*)r.Gk;
    const bool even = (N & 1) == 0;
    const uint32_t plane = (uint32_t)rows_per_field * (uint32_t)Nxh;   // one line's half spectra
    const float kf = r.kfac[0];
    const FastDiv dNxh(Nxh);
    for (int task = threadIdx.x; task < nb * Nxh; task += NT) {
        int kk;
        const int b = dNxh.div(task, kk);
        const int rr = b / NC, j = b - rr * NC;
        const int la = 2 * j, lb = la + 1;
        const uint32_t ia = (uint32_t)la * plane + (uint32_t)(row0 + rr) * (uint32_t)Nxh + (uint32_t)kk;
        float2 A = gld(at32(Gin, ia)), Bv = make_float2(0.f, 0.f);
        if (lb < NL) Bv = gld(at32(Gin, ia + plane));
        const bool self = kk == 0 || (even && 2 * kk == N);   // self-conjugate column
        // lines f*D + 0 hold the x derivative without its factor -i k_x
        const float kx = self && kk != 0 ? 0.f : kf * (float)kk;
        if (la % D == 0) A = make_float2(kx * A.y, -kx * A.x);
        if (lb < NL && lb % D == 0) Bv = make_float2(kx * Bv.y, -kx * Bv.x);
        if (self) { A.y = 0.f; Bv.y = 0.f; }
        lds[kk * ldb + b] = make_float2(A.x - Bv.y, -(A.y + Bv.x));
        if (!self && N - kk >= Nxh) lds[(N - kk) * ldb + b] = make_float2(A.x + Bv.y, A.y - Bv.x);
    }
    __syncthreads();
    const float2* res = lds_fft<NT>(lds, lds + (size_t)N * ldb, ldb, nb, r.ax[0]);
    const FastDiv dN(N);
    for (int task = threadIdx.x; task < rows * NF * N; task += NT) {
        int x;
        const int rest = dN.div(task, x);
        const int rr = rest / NF, f = rest - rr * NF;
        float g[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int line = f * D + d;
            const float2 z = res[x * ldb + rr * NC + line / 2];
            g[d] = (line & 1) ? -z.y : z.x;      // conj of the forward result
        }
        const uint32_t fn = (uint32_t)f * (uint32_t)r.M + (uint32_t)(row0 + rr) * (uint32_t)N + (uint32_t)x;
        // no density mask (see above)
        gst3((float*)((char*)r.G3 + ((fn << 3) + (fn << 2))), g[0], g[1], g[2]);   // 12 fn bytes
    }
}

template <bool RECORDS, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) k_x_c2r(const RepDev* __restrict__ reps,
                                                       const WorkItem* __restrict__ wmap,
                                                       int nlines_override, const float2* __restrict__ in_override,
                                                       float* __restrict__ out_override) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    if (RECORDS && r.grid_off32 && !r.FPtab) {   // the common layouts, per block (uniform)
        if (r.dim == 3 && r.nf == 2) { x_c2r_records<2, 3, NT>(r, wi, lds); return; }
        if (r.dim == 3 && r.nf == 5) { x_c2r_records<5, 3, NT>(r, wi, lds); return; }
        if (r.dim == 2 && r.nf == 2) { x_c2r_records<2, 2, NT>(r, wi, lds); return; }
    }
    const int N = r.Nx, Nxh = r.Nxh;
    const int rows_per_field = r.M / r.Nx;
    const int nlines = RECORDS ? r.nf * r.dim : nlines_override;     // real lines per row
    const int nc = (nlines + 1) / 2;                                  // complex lines per row
    const int nr = r.nr_c2r;
    const int row0 = wi.blk * nr;
    const int rows = min(nr, rows_per_field - row0);
    const int nb = rows * nc;
    const int ldb = nb | 1;
    const float2* Gin = RECORDS ? (const float2*)r.Gk : in_override;
    const bool even = (N & 1) == 0;
    const FastDiv dN(N), dNxh(Nxh), dnc(nc);
    // lines f*dim + 0 hold the x derivative without its factor -i k_x (constant along
    // the y and z lines it came through): it is applied here, to the half spectrum
    unsigned long long xlines = 0ull;
    if (RECORDS)
        for (int L = 0; L < nlines; L += r.dim) xlines |= 1ull << L;
    // each half-spectrum value is read once and fills column kk and its Hermitian
    // partner N - kk
    // (about two tasks per thread: a division pair each is cheaper than the set-up of
    // digit counters here)
    const bool o32 = RECORDS && r.grid_off32;
    for (int task = threadIdx.x; task < nb * Nxh; task += NT) {
        int kk, j;
        const int b = dNxh.div(task, kk);
        const int row = row0 + dnc.div(b, j);
        const int la = 2 * j, lb = la + 1;
        float2 A, Bv = make_float2(0.f, 0.f);
        if (o32) {
            const uint32_t ia = ((uint32_t)la * (uint32_t)rows_per_field + (uint32_t)row) * (uint32_t)Nxh + (uint32_t)kk;
            A = gld(at32(Gin, ia));
            if (lb < nlines) Bv = gld(at32(Gin, ia + (uint32_t)rows_per_field * (uint32_t)Nxh));
        } else {
            A = gld(Gin + ((size_t)la * rows_per_field + row) * Nxh + kk);
            if (lb < nlines) Bv = gld(Gin + ((size_t)lb * rows_per_field + row) * Nxh + kk);
        }
        const bool self = kk == 0 || (even && 2 * kk == N);   // self-conjugate column
        if (RECORDS) {
            const float kx = self && kk != 0 ? 0.f : r.kfac[0] * (float)kk;
            if ((xlines >> la) & 1ull) A = make_float2(kx * A.y, -kx * A.x);
            if ((xlines >> lb) & 1ull) Bv = make_float2(kx * Bv.y, -kx * Bv.x);
        }
        if (self) { A.y = 0.f; Bv.y = 0.f; }
        // Z = A + i B ; store conj(Z) for the conj(FFT(conj)) inverse.  Column N - kk
        // holds the Hermitian extension conj(A) + i conj(B).
        lds[kk * ldb + b] = make_float2(A.x - Bv.y, -(A.y + Bv.x));
        if (!self && N - kk >= Nxh) lds[(N - kk) * ldb + b] = make_float2(A.x + Bv.y, A.y - Bv.x);
    }
    __syncthreads();
    const float2* res = lds_fft<NT>(lds, lds + (size_t)N * ldb, ldb, nb, r.ax[0]);
    if (RECORDS && o32 && !r.FPtab) {
        // records of buffers below 4 GiB: dimension unrolled, 32-bit offsets
        const FastDiv dnf(r.nf);
        if (r.dim == 3) c2r_records<3, NT>(r, res, ldb, nc, rows, row0, N, dN, dnf);
        else c2r_records<2, NT>(r, res, ldb, nc, rows, row0, N, dN, dnf);
    } else if (RECORDS) {
        const int D = r.dim;
        const FastDiv dnf(r.nf);
        for (int task = threadIdx.x; task < rows * r.nf * N; task += NT) {
            int x, f;
            const int rest = dN.div(task, x), rr = dnf.div(rest, f);
            const int row = row0 + rr;
            float g[3] = {0.f, 0.f, 0.f};
            for (int d = 0; d < D; ++d) {
                const int line = f * D + d;
                const float2 z = res[x * ldb + rr * nc + line / 2];
                g[d] = (line & 1) ? -z.y : z.x;      // conj of the forward result
            }
            const size_t node = (size_t)row * N + x;
            if (f < r.ntypes) {
                const float rho = G(r.rho)[(size_t)f * r.M + node];
                if (r.FPtab)
                    // field phases (pair_style_fieldphases.cu:54-86): the grid force gains
                    // +-rho^2 f_d (d_prepareFieldForce then d_accumulateGridForce multiply
                    // by rho twice), i.e. rho f_d per unit density
                    for (int d = 0; d < D; ++d) g[d] += rho * G(r.FPtab)[((size_t)f * D + d) * r.M + node];
                // the gather skips nodes with rho <= 0 (device_grid_utils.cu:424-452):
                // a zero record adds nothing
                if (!(rho > 0.f)) g[0] = g[1] = g[2] = 0.f;
            }
            gst3(r.G3 + ((size_t)f * r.M + node) * 3, g[0], g[1], g[2]);
        }
    } else {
        const FastDiv dnl(nlines);
        for (int task = threadIdx.x; task < rows * nlines * N; task += NT) {
            int x, line;
            const int rest = dN.div(task, x), rr = dnl.div(rest, line);
            const float2 z = res[x * ldb + rr * nc + line / 2];
            G(out_override)[((size_t)line * rows_per_field + row0 + rr) * N + x] = (line & 1) ? -z.y : z.x;
        }
    }
}

// ------------------------------------------------------------------ plane forward
// x-axis R2C and y-axis FFT of one (field, z) density plane in a single block (3D,
// even Ny): rows (y, y+1) share one complex x transform -- the pairing of the tiled
// deposit with even tile heights, so H is bit-identical to deposit + separate y pass
// -- and the unpacked half spectra are transformed along y in LDS before one write.
template <int NT>
__global__ void __launch_bounds__(NT) k_plane_r2c(const RepDev* __restrict__ reps,
                                                  const WorkItem* __restrict__ wmap) {
    extern __shared__ float2 lds[];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int Nx = r.Nx, Ny = r.Ny, Nxh = r.Nxh;
    const int z = wi.blk % r.Nz, f = wi.blk / r.Nz;
    const int np = Ny / 2, ldp = np | 1, ldc = Nxh | 1;
    const size_t bsz = max((size_t)Nx * ldp, (size_t)Ny * ldc);
    float2* A = lds;
    float2* B = lds + bsz;
    gmem<const float> rho = G((const float*)r.rho) + (size_t)f * r.M + (size_t)z * Ny * Nx;
    const FastDiv dNx(Nx), dNxh(Nxh);
    Digits2 gl(threadIdx.x, NT, dNx);
    for (int task = threadIdx.x; task < np * Nx; task += NT, gl.step()) {
        const int n = gl.lo, j = gl.hi;
        A[n * ldp + j] = make_float2(rho[(size_t)(2 * j) * Nx + n], rho[(size_t)(2 * j + 1) * Nx + n]);
    }
    __syncthreads();
    float2* Z = lds_fft<NT>(A, B, ldp, np, r.ax[0]);
    float2* C = (Z == A) ? B : A;
    Digits2 gm(threadIdx.x, NT, dNxh);
    for (int task = threadIdx.x; task < np * Nxh; task += NT, gm.step()) {
        const int k = gm.lo, j = gm.hi;
        const float2 Zk = Z[k * ldp + j];
        const float2 Zm = Z[(k == 0 ? 0 : Nx - k) * ldp + j];
        C[(2 * j) * ldc + k] = make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y));
        C[(2 * j + 1) * ldc + k] = make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x));
    }
    __syncthreads();
    const float2* Y = lds_fft<NT>(C, Z, ldc, Nxh, r.ax[1]);
    float2* H = r.H + ((size_t)f * r.Nz + z) * Ny * Nxh;
    Digits2 gs(threadIdx.x, NT, dNxh);
    for (int task = threadIdx.x; task < Ny * Nxh; task += NT, gs.step()) {
        const int k = gs.lo, y = gs.hi;
        gst(H + (size_t)y * Nxh + k, Y[y * ldc + k]);
    }
}

// ------------------------------------------------------------------ particles
// Gather (device_grid_utils.cu:346-452) + harmonic bonds (device_bonds.cu:4-60) +
// EM (device_EM_integrator.cu:5-36) / GJF (device_GJF_integrator.cu:4-62).
// The gather reads g = all_f/rho directly; nodes with rho <= 0 contribute nothing
// (device_grid_utils.cu:640).  Positions are read from X[step&1] and written to
// X[(step+1)&1], so bonds always see x_n.
//
// The bead body is specialised on the dimension D and on the stencil width NP = P + 1
// of the orders P = 0 and 1, so every per-axis array stays in registers and every
// stencil loop is unrolled; NP = 0 is the any-order body (P from the descriptor).
// The per-bead loads are issued before the first of them is used, the three bond
// partners together after the gather (before it they would hold registers through
// the eight gather records and cost a wave per SIMD).

// Stencil sum acc += sum_nodes w * record over the records of one field, which start
// at element fb of G3.  The rho > 0 mask is already in the records (zero where
// rho <= 0): adding w*0 leaves every sum unchanged, and the loads stay unconditional
// (a branch per node would serialise them).  NP > 0: unrolled, every record requested
// before the first is used.  OFF32: every record lies within 4 GiB of G3, so a record
// is addressed by the uniform base plus a 32-bit byte offset (one register per load
// in flight instead of two).
template <int NP, bool OFF32>
__device__ __forceinline__ float3 gather_rec(const float* G3, size_t fb, int node) {
    if constexpr (OFF32)
        return gld3((const float*)((const char*)G3 + ((uint32_t)fb + (uint32_t)node) * 12u));
    else
        return gld3(G3 + (fb + (size_t)node) * 3);
}

template <int D, int NP, bool OFF32>
__device__ __forceinline__ void gather_sum(const RepDev& r, int P, const float (&W)[3][5], const int (&n0)[3],
                                           size_t fb, float (&acc)[3]) {
    const float* G3 = r.G3;
    if constexpr (NP > 0) {
        int zz[NP], yy[NP], xx[NP];
        zz[0] = n0[2]; yy[0] = n0[1]; xx[0] = n0[0];
#pragma unroll
        for (int k = 1; k < NP; ++k) {
            zz[k] = wrap_next(zz[k - 1], r.Nz);
            yy[k] = wrap_next(yy[k - 1], r.Ny);
            xx[k] = wrap_next(xx[k - 1], r.Nx);
        }
        constexpr int nzl = D == 3 ? NP : 1;
#pragma unroll
        for (int iz = 0; iz < nzl; ++iz) {
#pragma unroll
            for (int iy = 0; iy < NP; ++iy) {
                const int rowb = (zz[iz] * r.Ny + yy[iy]) * r.Nx;
#pragma unroll
                for (int ix = 0; ix < NP; ++ix) {
                    const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                    const float3 g = gather_rec<NP, OFF32>(G3, fb, rowb + xx[ix]);
                    acc[0] += w * g.x; acc[1] += w * g.y; acc[2] += w * g.z;
                }
            }
        }
    } else {
        const int nz = D == 3 ? P + 1 : 1;
        for (int iz = 0, nzz = n0[2]; iz < nz; ++iz, nzz = wrap_next(nzz, r.Nz)) {
            for (int iy = 0, nyy = n0[1]; iy <= P; ++iy, nyy = wrap_next(nyy, r.Ny)) {
                const int rowb = (nzz * r.Ny + nyy) * r.Nx;
                for (int ix = 0, nxx = n0[0]; ix <= P; ++ix, nxx = wrap_next(nxx, r.Nx)) {
                    const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                    const float3 g = gather_rec<NP, OFF32>(G3, fb, rowb + nxx);
                    acc[0] += w * g.x; acc[1] += w * g.y; acc[2] += w * g.z;
                }
            }
        }
    }
}

// The unrolled gather in two halves, so that other work can run while the records are
// in flight: the loads (records in the order gather_sum adds them) ...
template <int D, int NP, bool OFF32>
__device__ __forceinline__ void gather_load(const RepDev& r, const int (&n0)[3], size_t fb,
                                            float3 (&g)[(D == 3 ? NP : 1) * NP * NP]) {
    int zz[NP], yy[NP], xx[NP];
    zz[0] = n0[2]; yy[0] = n0[1]; xx[0] = n0[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) {
        zz[k] = wrap_next(zz[k - 1], r.Nz);
        yy[k] = wrap_next(yy[k - 1], r.Ny);
        xx[k] = wrap_next(xx[k - 1], r.Nx);
    }
    constexpr int nzl = D == 3 ? NP : 1;
#pragma unroll
    for (int iz = 0; iz < nzl; ++iz)
#pragma unroll
        for (int iy = 0; iy < NP; ++iy) {
            const int rowb = (zz[iz] * r.Ny + yy[iy]) * r.Nx;
#pragma unroll
            for (int ix = 0; ix < NP; ++ix) g[(iz * NP + iy) * NP + ix] = gather_rec<NP, OFF32>(r.G3, fb, rowb + xx[ix]);
        }
}

// ... and the weighted sum, in gather_sum's order (bit-identical to it)
template <int D, int NP>
__device__ __forceinline__ void gather_acc(const float (&W)[3][5], const float3 (&g)[(D == 3 ? NP : 1) * NP * NP],
                                           float (&acc)[3]) {
    constexpr int nzl = D == 3 ? NP : 1;
#pragma unroll
    for (int iz = 0; iz < nzl; ++iz)
#pragma unroll
        for (int iy = 0; iy < NP; ++iy)
#pragma unroll
            for (int ix = 0; ix < NP; ++ix) {
                const float w = D == 3 ? W[0][ix] * W[1][iy] * W[2][iz] : W[0][ix] * W[1][iy];
                const float3 v = g[(iz * NP + iy) * NP + ix];
                acc[0] += w * v.x; acc[1] += w * v.y; acc[2] += w * v.z;
            }
}

// Gathered force of one bead: its type's records, plus q times the charge records
// (a second sweep, so that no more than one stencil of records is in flight).
template <int D, int NP, bool OFF32>
__device__ __forceinline__ void gather_nodes(const RepDev& r, int P, const float (&W)[3][5], const int (&n0)[3],
                                             int t, float qi, bool charged, float (&f)[3]) {
    gather_sum<D, NP, OFF32>(r, P, W, n0, (size_t)t * r.M, f);
    if (charged) {
        __builtin_amdgcn_sched_barrier(0);
        float fc[3] = {0.f, 0.f, 0.f};
        gather_sum<D, NP, OFF32>(r, P, W, n0, (size_t)r.ntypes * r.M, fc);
#pragma unroll
        for (int d = 0; d < D; ++d) f[d] += qi * fc[d];
    }
}

template <int D>
__device__ __forceinline__ void load_pos(const float* p, float (&v)[3]) {
    if constexpr (D == 3) {
        const float3 a = gld3(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z;
    } else {
        gmem<const float> q = G(p);
        v[0] = q[0]; v[1] = q[1]; v[2] = 0.f;
    }
}

template <int D>
__device__ __forceinline__ void store_pos(float* p, const float (&v)[3]) {
    if constexpr (D == 3) {
        gst3(p, v[0], v[1], v[2]);
    } else {
        gmem<float> q = G(p);
        q[0] = v[0]; q[1] = v[1];
    }
}

// Minimum image with one wrap, as device_bonds.cu: selects, not branches.
__device__ __forceinline__ float min_image(float dr, float L, float Lh) {
    const float lo = dr + L, hi = dr - L;
    return dr > Lh ? hi : (dr < -Lh ? lo : dr);
}

// A bead's state for the next step's bucket claim: new position, type and charge.
struct BeadNext {
    float4 rec;   // (x, y, z, type bits)
    float q;
};

template <bool INTEGRATE, int D, int NP, bool OFF32>
__device__ __forceinline__ void particle_bead(const RepDev& r, int i, int write_forces, int step, BeadNext& nx) {
    const int P = NP > 0 ? NP - 1 : r.P;
    const float* xp = r.X[step & 1];
    // per-bead loads
    float xi[3];
    load_pos<D>(xp + (size_t)i * D, xi);
    const int t = G(r.tp)[i];
    const bool charged = r.q != nullptr;
    const float qi = charged ? G(r.q)[i] : 0.f;
    const int3 bv = gld3(r.bond + (size_t)i * 3);
    // bonds, every bond evaluated from both ends (no atomics).  Slots fill from the
    // front and the first empty one ends the list; an empty slot loads the bead itself
    // and bond type 0 (both exist) and contributes nothing, so the loads of all three
    // partners go out together with no branch.
    // the site id goes out first: the noise needs only it, and can then be computed
    // while the partners are in flight
    uint32_t site = 0;
    if constexpr (INTEGRATE) site = (uint32_t)G(r.pid)[i];   // original site id: slot order does not matter
    const int bs = r.bond_shift, tmask = (1 << bs) - 1;
    const int js[3] = {bv.x >> bs, bv.y >> bs, bv.z >> bs}, ts[3] = {bv.x & tmask, bv.y & tmask, bv.z & tmask};
    bool live[3];
    live[0] = bv.x >= 0;
    live[1] = live[0] && bv.y >= 0;
    live[2] = live[1] && bv.z >= 0;
    float xj[3][3], kb[3], r0[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        load_pos<D>(xp + (size_t)(live[b] ? js[b] : i) * D, xj[b]);
        const int tb = live[b] ? ts[b] : 0;
        kb[b] = G(r.bond_k)[tb];
        r0[b] = G(r.bond_r0)[tb];
    }
    // integrator inputs go out with the partners
    int kind = 0;
    float xo[3] = {0.f, 0.f, 0.f}, ga = 0.f, gc1 = 0.f, gc2 = 0.f, gnm = 0.f;
    if constexpr (INTEGRATE) {
        kind = G(r.ikind)[i];
        if (r.any_gjf) {
            load_pos<D>(r.XO + (size_t)i * D, xo);
            ga = r.gjf_a[t];
            gc1 = r.gjf_c1[t];
            gc2 = r.gjf_c2[t];
            gnm = r.gjf_noise[t];
        }
    }
    // stencil
    float W[3][5];
    int n0[3] = {0, 0, 0};
    W[2][0] = 1.f;
    n0[0] = stencil_1d(xi[0], r.dx[0], r.idx[0], r.Nx, P, W[0]);
    n0[1] = stencil_1d(xi[1], r.dx[1], r.idx[1], r.Ny, P, W[1]);
    if constexpr (D == 3) n0[2] = stencil_1d(xi[2], r.dx[2], r.idx[2], r.Nz, P, W[2]);
    // gather; with the unrolled stencil the noise is computed while the type's records
    // are in flight
    float f[3] = {0.f, 0.f, 0.f};
    float z[4] = {0.f, 0.f, 0.f, 0.f}, zo[4] = {0.f, 0.f, 0.f, 0.f};
    const uint32_t stepc = (uint32_t)(step + 1);
    if constexpr (NP > 0) {
        float3 g[(D == 3 ? NP : 1) * NP * NP];
        gather_load<D, NP, OFF32>(r, n0, (size_t)t * r.M, g);
        if constexpr (INTEGRATE) {
            normals4(r.seed, site, stepc, z);
            if (r.any_gjf && stepc > 1u) normals4(r.seed, site, stepc - 1u, zo);
        }
        gather_acc<D, NP>(W, g, f);
        if (charged) {
            __builtin_amdgcn_sched_barrier(0);
            float fc[3] = {0.f, 0.f, 0.f};
            gather_sum<D, NP, OFF32>(r, P, W, n0, (size_t)r.ntypes * r.M, fc);
#pragma unroll
            for (int d = 0; d < D; ++d) f[d] += qi * fc[d];
        }
    } else {
        gather_nodes<D, NP, OFF32>(r, P, W, n0, t, qi, charged, f);
        if constexpr (INTEGRATE) {
            normals4(r.seed, site, stepc, z);
            if (r.any_gjf && stepc > 1u) normals4(r.seed, site, stepc - 1u, zo);
        }
    }
    float lf[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        float dr[3] = {0.f, 0.f, 0.f}, r2 = 0.f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            dr[d] = min_image(xi[d] - xj[b][d], r.L[d], r.Lh[d]);
            r2 += dr[d] * dr[d];
        }
        // 2 k (|dr| - r0) / |dr| from one reciprocal square root
        const float ir = __builtin_amdgcn_rsqf(fmaxf(r2, 1.0e-4f));
        const float mf = (live[b] && r2 > 1.0e-4f) ? 2.0f * kb[b] * (r2 * ir - r0[b]) * ir : 0.f;
#pragma unroll
        for (int d = 0; d < D; ++d) lf[d] -= mf * dr[d];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) f[d] += lf[d];
    if (write_forces && r.fout) store_pos<D>(r.fout + (size_t)i * D, f);
    if constexpr (INTEGRATE) {
        // kind 0 keeps x, 1 is EM (device_EM_integrator.cu:5-36), 2 is GJF
        // (device_GJF_integrator.cu:25-58) in displacement form: with
        // b = m/(m + dt/2D) and a = (1 - dt/2Dm)/(1 + dt/2Dm), 2b - a = 1 exactly, so
        //   x' = 2b x - a xo + c1 f + c2 (beta + beta_old)
        //      = x + [a (x - xo) + c1 f + c2 (beta + beta_old)].
        // XO holds the unwrapped displacement x - xo, which removes the reference's
        // minimum-image unwrap of xo and keeps FP32 rounding off the large terms.
        float v[3] = {0.f, 0.f, 0.f}, dn[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float nn = gnm * z[d], no = gnm * zo[d];
            dn[d] = ga * xo[d] + gc1 * f[d] + gc2 * (nn + no);
            const float em = xi[d] + r.delt * f[d] + r.noise_mag * z[d];
            float u = kind == 2 ? xi[d] + dn[d] : em;
            u = u > r.L[d] ? u - r.L[d] : (u < 0.0f ? u + r.L[d] : u);
            v[d] = kind == 0 ? xi[d] : u;
        }
        store_pos<D>(r.X[(step + 1) & 1] + (size_t)i * D, v);
        if (kind == 2) store_pos<D>(r.XO + (size_t)i * D, dn);
        nx.rec = make_float4(v[0], v[1], v[2], __int_as_float(t));
        nx.q = qi;
    }
}

// CLAIM (integrating steps of LDS-tiled replicas): the block then claims the next
// step's bucket slots from the positions it has just written, which takes the bucket
// claim pass out of every step but the first of a run.
template <bool INTEGRATE, bool CLAIM>
__global__ void __launch_bounds__(256) k_particles(const RepDev* __restrict__ reps,
                                                   const WorkItem* __restrict__ wmap,
                                                   int write_forces, int koff) {
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int i = wi.blk * 256 + threadIdx.x;
    const bool valid = i < r.ns;
    if (!CLAIM && !valid) return;
    const int step = r.step + koff;
    BeadNext nx{make_float4(0.f, 0.f, 0.f, 0.f), 0.f};
    // unrolled bodies: the default order P = 1 on grids whose records fit 32-bit
    // offsets, and P = 0; any other case takes the looped body
    if (valid) {
        if (r.P == 1 && r.g3_off32) {
            if (r.dim == 3) particle_bead<INTEGRATE, 3, 2, true>(r, i, write_forces, step, nx);
            else particle_bead<INTEGRATE, 2, 2, true>(r, i, write_forces, step, nx);
        } else if (r.P == 0) {
            if (r.dim == 3) particle_bead<INTEGRATE, 3, 1, false>(r, i, write_forces, step, nx);
            else particle_bead<INTEGRATE, 2, 1, false>(r, i, write_forces, step, nx);
        } else {
            if (r.dim == 3) particle_bead<INTEGRATE, 3, 0, false>(r, i, write_forces, step, nx);
            else particle_bead<INTEGRATE, 2, 0, false>(r, i, write_forces, step, nx);
        }
    }
    if constexpr (CLAIM) {
        if (r.spread_mode == 2) claim_record(r, valid, nx.rec, nx.q, (step + 1) & 1);
    }
}

// First stencil node of every coordinate (the bins of device_grid_utils.cu:136-146).
__global__ void __launch_bounds__(256) k_bins(const RepDev* __restrict__ reps, int rep, int* __restrict__ out) {
    const RepDev& r = reps[rep];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= r.ns) return;
    const float* x = r.X[r.step & 1];
    for (int d = 0; d < r.dim; ++d) {
        int N = d == 0 ? r.Nx : (d == 1 ? r.Ny : r.Nz);
        out[i * r.dim + d] = stencil_node(x[i * r.dim + d], r.dx[d], N, r.P);
    }
}

// ------------------------------------------------------------------ properties
// Non-bonded energies and virial by Parseval over the stored half spectra H:
//   U_s  = gvol/M sum_k w_k u_s(k) Re(conj rho_t1 rho_t2)     (pair_style.cu:204-233)
//   P_pc = gvol/(V D M) sum_k w_k Re(conj rho_t1 V_pc(k) rho_t2)   (pair_style.cu:235-260)
//   Uel  = 1/M sum_k w_k phi(k) |rho_c|^2                     (calc_properties.cu:106-116)
// with w_k = 1 on the self-conjugate x columns and 2 elsewhere.  grid.y = quantity.
__global__ void __launch_bounds__(256) k_kprops(const RepDev* __restrict__ reps,
                                                const WorkItem* __restrict__ wmap) {
    __shared__ double red[256];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int qn = blockIdx.y;
    if (qn >= r.nq) return;
    const int nP = r.nP;
    double s = 0.0;
    const int per = (r.Mh + r.nchunk - 1) / r.nchunk;
    const int k0 = wi.blk * per, k1 = min(r.Mh, k0 + per);
    const bool even = (r.Nx & 1) == 0;
    for (int idx = k0 + threadIdx.x; idx < k1; idx += 256) {
        const int kx = idx % r.Nxh;
        const double w = (kx == 0 || (even && 2 * kx == r.Nx)) ? 1.0 : 2.0;
        if (qn < r.nstyles) {
            const float2 a = r.H[(size_t)r.st_t1[qn] * r.Mh + idx], b = r.H[(size_t)r.st_t2[qn] * r.Mh + idx];
            const double re = (double)a.x * b.x + (double)a.y * b.y;
            s += w * (double)r.Etab[(size_t)qn * r.Mh + idx] * re;
        } else if (qn < r.nstyles + r.npairs * nP) {
            const int pq = qn - r.nstyles, p = pq / nP, c = pq % nP;
            const float2 a = r.H[(size_t)r.pr_t1[p] * r.Mh + idx], b = r.H[(size_t)r.pr_t2[p] * r.Mh + idx];
            const float2 v = r.Vtab[((size_t)p * nP + c) * r.Mh + idx];
            // Re(conj(a) v b)
            const double ab_re = (double)a.x * b.x + (double)a.y * b.y;
            const double ab_im = (double)a.x * b.y - (double)a.y * b.x;
            s += w * ((double)v.x * ab_re - (double)v.y * ab_im);
        } else {
            const float2 c = r.H[(size_t)r.ntypes * r.Mh + idx];
            s += w * (double)r.Eltab[idx] * ((double)c.x * c.x + (double)c.y * c.y);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) r.partials[(size_t)qn * r.nchunk + wi.blk] = red[0];
}

// Bond energy and virial at the current positions (bonds.cu:3-52): each bond once
// from its lower-index end, no r^2 guard.  Quantities: Ubond, vir[0..nP-1].
__global__ void __launch_bounds__(256) k_bond_props(const RepDev* __restrict__ reps,
                                                    const WorkItem* __restrict__ wmap) {
    __shared__ double red[7][256];
    const WorkItem wi = wmap[blockIdx.x];
    const RepDev& r = reps[wi.rep];
    const int i = wi.blk * 256 + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < r.ns) {
        const int D = r.dim;
        const float* x = r.X[r.step & 1];
        for (int b = 0; b < 3; ++b) {
            const int v = r.bond[i * 3 + b];
            if (v < 0) break;
            const int j = v >> r.bond_shift;
            if (j < i) continue;
            const int bt = v & ((1 << r.bond_shift) - 1);
            double dr[3] = {0, 0, 0}, r2 = 0.0;
            for (int d = 0; d < D; ++d) {
                float df = x[i * D + d] - x[j * D + d];
                if (df > r.Lh[d]) df -= r.L[d];
                else if (df < -r.Lh[d]) df += r.L[d];
                dr[d] = df;
                r2 += dr[d] * dr[d];
            }
            const double rr = sqrt(r2), del = rr - (double)r.bond_r0[bt];
            acc[0] += del * del * (double)r.bond_k[bt];
            const double mf = 2.0 * (double)r.bond_k[bt] * del / rr;
            acc[1] += -mf * dr[0] * dr[0];
            acc[2] += -mf * dr[1] * dr[1];
            if (D == 2) acc[3] += -mf * dr[0] * dr[1];
            else {
                acc[3] += -mf * dr[2] * dr[2];
                acc[4] += -mf * dr[0] * dr[1];
                acc[5] += -mf * dr[0] * dr[2];
                acc[6] += -mf * dr[1] * dr[2];
            }
        }
    }
    for (int q = 0; q < 7; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 7; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    const int nblk = (r.ns + 255) / 256;
    if (threadIdx.x < 7) r.bpart[(size_t)threadIdx.x * nblk + wi.blk] = red[threadIdx.x][0];
}

// Fixed-order sums of per-block partials: out[q] = sum_b in[q*n + b].
__global__ void __launch_bounds__(256) k_reduce_rows(const double* __restrict__ in, int n,
                                                     double* __restrict__ out) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += 256) s += in[(size_t)q * n + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// Step counters of the replicas that just integrated nsteps steps.  The kernels of a
// run of steps read the counter plus their step's offset in the run, so one launch
// per run advances it.  (Counting the particle blocks done inside k_particles instead
// costs each block a device-scope atomic round trip, which measured slower.)
__global__ void k_advance(RepDev* reps, const int* __restrict__ which, int n, int nsteps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) reps[which[k]].step += nsteps;
}

// Virial source fields of one pair style (d_initVirial, device_array_utils.cu:456-494):
// vir_c(r) = dr_a(r) f_b(r) with dr the minimum image of the node position
// (wrap test in FP32 as the reference), components xx yy [zz] xy [xz yz].
__global__ void k_virial_source(const RepDev* __restrict__ reps, int rep,
                                const float* __restrict__ fr, float* __restrict__ vr) {
    const RepDev& r = reps[rep];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.M) return;
    const int D = r.dim;
    int nn[3];
    nn[0] = i % r.Nx;
    nn[1] = (i / r.Nx) % r.Ny;
    nn[2] = D == 3 ? i / (r.Nx * r.Ny) : 0;
    float dr[3];
    for (int j = 0; j < D; ++j) {
        float v = (float)nn[j] * r.dx[j];
        if (v > r.Lh[j]) v -= r.L[j];
        else if (v < -r.Lh[j]) v += r.L[j];
        dr[j] = v;
    }
    float f[3];
    for (int d = 0; d < D; ++d) f[d] = fr[(size_t)d * r.M + i];
    const size_t M = r.M;
    vr[0 * M + i] = dr[0] * f[0];
    vr[1 * M + i] = dr[1] * f[1];
    if (D == 2) vr[2 * M + i] = dr[1] * f[0];
    else {
        vr[2 * M + i] = dr[2] * f[2];
        vr[3 * M + i] = dr[1] * f[0];
        vr[4 * M + i] = dr[2] * f[0];
        vr[5 * M + i] = dr[2] * f[1];
    }
}

// ------------------------------------------------------------------ host launchers
#define CHECK_LAUNCH() do { hipError_t e__ = hipGetLastError(); if (e__ != hipSuccess) return e__; } while (0)

hipError_t launch_spread(const LaunchCtx& c, int koff) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_spread, dim3(c.n), dim3(256), 0, c.reps, c.wmap, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_x_r2c(const LaunchCtx& c, bool from_acc, int nrows_override, const float* in, float2* out) {
    if (!c.n) return hipSuccess;
    if (from_acc) TILD_LAUNCH(c, k_x_r2c<true>, dim3(c.n), dim3(kFftThreads), c.lds, c.reps, c.wmap, nrows_override, in, out);
    else TILD_LAUNCH(c, k_x_r2c<false>, dim3(c.n), dim3(kFftThreads), c.lds, c.reps, c.wmap, nrows_override, in, out);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_mid(const LaunchCtx& c, int axis, int use_gk, int nfb, int inverse, float2* buf) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_mid, dim3(c.n), dim3(kFftThreads), c.lds, c.reps, c.wmap, axis, use_gk, nfb, inverse, buf);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_last_fused(const LaunchCtx& c, int store_spectrum, int koff) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_last_fused<kLastThreads>, dim3(c.n), dim3(kLastThreads), c.lds, c.reps, c.wmap, store_spectrum, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_x_c2r(const LaunchCtx& c, bool records, int nlines, const float2* in, float* out) {
    if (!c.n) return hipSuccess;
    if (records) TILD_LAUNCH(c, (k_x_c2r<true, kC2rThreads>), dim3(c.n), dim3(kC2rThreads), c.lds, c.reps, c.wmap, nlines, in, out);
    else TILD_LAUNCH(c, (k_x_c2r<false, kFftThreads>), dim3(c.n), dim3(kFftThreads), c.lds, c.reps, c.wmap, nlines, in, out);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t configure_kernels() {
    hipError_t e = hipFuncSetAttribute((const void*)k_last_fused<kLastThreads>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPlaneLds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_plane_r2c<kPlaneThreads>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPlaneLds);
}
hipError_t launch_plane_r2c(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_plane_r2c<kPlaneThreads>, dim3(c.n), dim3(kPlaneThreads), c.lds, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_mid_grad(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    TILD_LAUNCH(c, k_mid_grad<kGradThreads>, dim3(c.n), dim3(kGradThreads), c.lds, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_particles(const LaunchCtx& c, bool integrate, int write_forces, int koff, bool claim) {
    if (!c.n) return hipSuccess;
    if (integrate && claim)
        TILD_LAUNCH(c, (k_particles<true, true>), dim3(c.n), dim3(256), 0, c.reps, c.wmap, write_forces, koff);
    else if (integrate)
        TILD_LAUNCH(c, (k_particles<true, false>), dim3(c.n), dim3(256), 0, c.reps, c.wmap, write_forces, koff);
    else
        TILD_LAUNCH(c, (k_particles<false, false>), dim3(c.n), dim3(256), 0, c.reps, c.wmap, write_forces, koff);
    CHECK_LAUNCH(); return hipSuccess;
}
// avg_sk (Compute.cu:83-117): acc(k) += rho_t(k) rho_t(k) / M over the half spectrum,
// the complex square of the unnormalised transform (d_multiplyComplex with both
// operands rho(k), device_array_utils.cu:314-326 -- not |rho(k)|^2), summed in FP64.
__global__ void __launch_bounds__(256) k_sk_accumulate(const RepDev* __restrict__ reps, int rep, int type,
                                                       double* __restrict__ acc) {
    const RepDev& r = reps[rep];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= r.Mh) return;
    const float2 h = gld(r.H + (size_t)type * r.Mh + k);
    const double x = h.x, y = h.y, inv = 1.0 / (double)r.M;
    gmem<double> a = G(acc) + 2 * (size_t)k;
    a[0] += (x * x - y * y) * inv;
    a[1] += 2.0 * x * y * inv;
}

hipError_t launch_sk_accumulate(hipStream_t s, const RepDev* reps, int rep, int type, int Mh, double* acc) {
    hipLaunchKernelGGL(k_sk_accumulate, dim3((Mh + 255) / 256), dim3(256), 0, s, reps, rep, type, acc);
    return hipGetLastError();
}

hipError_t launch_kprops(const LaunchCtx& c, int nq_max) {
    if (!c.n || nq_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kprops, dim3(c.n, nq_max), dim3(256), 0, c.stream, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_bond_props(const LaunchCtx& c) {
    if (!c.n) return hipSuccess;
    hipLaunchKernelGGL(k_bond_props, dim3(c.n), dim3(256), 0, c.stream, c.reps, c.wmap);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_reduce_rows(hipStream_t s, const double* in, int nrows, int n, double* out) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_rows, dim3(nrows), dim3(256), 0, s, in, n, out);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_advance(hipStream_t s, RepDev* reps, const int* which, int n, int nsteps, hipEvent_t t0, hipEvent_t t1) {
    if (n <= 0) return hipSuccess;
    const LaunchCtx c{nullptr, nullptr, 0, 0, s, t0, t1};
    TILD_LAUNCH(c, k_advance, dim3((n + 255) / 256), dim3(256), 0, reps, which, n, nsteps);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_bins(hipStream_t s, const RepDev* reps, int rep, int ns, int* out) {
    hipLaunchKernelGGL(k_bins, dim3((ns + 255) / 256), dim3(256), 0, s, reps, rep, out);
    CHECK_LAUNCH(); return hipSuccess;
}
hipError_t launch_virial_source(hipStream_t s, const RepDev* reps, int rep, int M, const float* fr, float* vr) {
    hipLaunchKernelGGL(k_virial_source, dim3((M + 255) / 256), dim3(256), 0, s, reps, rep, fr, vr);
    CHECK_LAUNCH(); return hipSuccess;
}

}  // namespace tild
