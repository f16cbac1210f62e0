/*
 * ora_step.c -- FP64 oracle of the reference per-step routines.
 * TEST INFRASTRUCTURE ONLY (see tild_oracle.h).
 *
 * Each function names the reference routine it restates.  Parity status (the tests
 * that pin it: tests/test_oracle.py, tests/test_oracle_pins.py):
 *   derived constants      initialize.cu:32-48             pinned (Coacervate data.dat)
 *   get_k / unstack        device_array_utils.cu:9-82      pinned (data.dat; closed forms)
 *   spline weights         device_grid_utils.cu:3-49       pinned (cardinal B-splines by
 *                                                          Cox-de Boor in 3D and 2D; CIC
 *                                                          exactness; order-4 bug kept)
 *   spread                 device_grid_utils.cu:232-343    pinned (data.dat; sum rule 3D/2D)
 *   pair-style tables      pair_style_gaussian.cu:72-90, pair_style_erf.cu:122-157,
 *                          pair_style_gaussian_erf.cu:82-115  pinned (real-space energies
 *                                                          of the smoothed sphere / overlap)
 *   CalcForces             pair_style.cu:81-196            pinned (momentum conservation,
 *                                                          two-particle -grad u in 3D and
 *                                                          2D, exact for beads on nodes;
 *                                                          -dU/dx convergence)
 *   CalcCharges            pair_style_charges.cu:21-111    pinned (two-charge smeared
 *                                                          Coulomb force; data.dat Uelec)
 *   gather                 device_grid_utils.cu:346-452    pinned (the force pins above)
 *   d_bonds                device_bonds.cu:4-60            pinned (finite difference of
 *                                                          host_bonds energy; Newton III)
 *   EM                     device_EM_integrator.cu:5-36    pinned: noise off, one step is
 *                                                          explicit Euler x + delt f and a
 *                                                          harmonic dimer decays as
 *                                                          (1 - 4 k delt)^n; noise on, the
 *                                                          dimer's bond variance is the
 *                                                          discrete Ornstein-Uhlenbeck value
 *                                                          and free diffusion gives 2 delt
 *   GJF                    device_GJF_integrator.cu:4-62   pinned: noise off, a harmonic
 *                                                          dimer follows the closed-form
 *                                                          solution of the GJF recurrence
 *                                                          (three mass / diffusivity pairs);
 *                                                          noise on, <r^2> = 3/(2k) (GJF's
 *                                                          exact harmonic sampling, D = 1)
 *                                                          and the underdamped free MSD.
 *                                                          Noise values replace cuRAND (see
 *                                                          ora_util.c)
 *   CalcEnergy/CalcVirial  pair_style.cu:204-276           pinned (data.dat; continuum pair
 *                                                          energy and virial r_a F_b / (V D)
 *                                                          for beads on nodes, 3D and 2D)
 *   host_bonds             bonds.cu:3-52                   pinned (data.dat Ubond, P)
 *   electrostatic energy   calc_properties.cu:106-116      pinned (data.dat Upe)
 *   ramp Update            pair_style.cu:308-328           pinned (closed-form amplitude)
 *   NList                  nlist.cu:13-228                 pinned (brute-force O(n^2) sets)
 */
#include "ora_internal.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ERR(...) do { snprintf(err, errlen, __VA_ARGS__); return -1; } while (0)

/* array_utils.cu:43-64 / device_array_utils.cu:9-32 */
static void unstack(const ora_sys* s, int id, int* nn) {
    if (s->dim == 2) { nn[1] = id / s->Nx[0]; nn[0] = id % s->Nx[0]; nn[2] = 0; }
    else {
        nn[2] = id / (s->Nx[1] * s->Nx[0]);
        int r = id - nn[2] * s->Nx[0] * s->Nx[1];
        nn[1] = r / s->Nx[0]; nn[0] = r % s->Nx[0];
    }
}

/* device_array_utils.cu:60-82 (FP64, exact 2*pi) */
static double get_k(const ora_sys* s, int id, double* k) {
    int nn[3];
    unstack(s, id, nn);
    double k2 = 0.0;
    for (int i = 0; i < s->dim; ++i) {
        if ((double)nn[i] < (double)s->Nx[i] / 2.0) k[i] = 2.0 * ORA_PI * nn[i] / s->L[i];
        else k[i] = 2.0 * ORA_PI * (double)(nn[i] - s->Nx[i]) / s->L[i];
        k2 += k[i] * k[i];
    }
    return k2;
}

/* device_grid_utils.cu:3-49, including the order-4 sx3 = sx2*sx2 defect (:39) */
static void spline_weights(double gdx, double H, double* W, int P) {
    double sx = gdx / H, sx2, sx3, sx4;
    if (P == 0) W[0] = 1.0;
    else if (P == 1) { W[0] = 0.5 - sx; W[1] = 0.5 + sx; }
    else if (P == 2) {
        sx2 = sx * sx;
        W[0] = 0.125 - 0.5 * sx + 0.5 * sx2; W[1] = 0.75 - sx2; W[2] = 0.125 + 0.5 * sx + 0.5 * sx2;
    } else if (P == 3) {
        sx2 = sx * sx; sx3 = sx2 * sx;
        W[0] = (1. - 6. * sx + 12. * sx2 - 8. * sx3) / 48.;
        W[1] = (23. - 30. * sx - 12. * sx2 + 24. * sx3) / 48.;
        W[2] = (23. + 30. * sx - 12. * sx2 - 24. * sx3) / 48.;
        W[3] = (1. + 6. * sx + 12. * sx2 + 8. * sx3) / 48.;
    } else {
        sx2 = sx * sx; sx3 = sx2 * sx2; sx4 = sx2 * sx2;   /* reference defect kept */
        W[0] = (1. - 8. * sx + 24. * sx2 - 32. * sx3 + 16. * sx4) / 384.;
        W[1] = (19. - 44. * sx + 24. * sx2 + 16. * sx3 - 16. * sx4) / 96.;
        W[2] = (115. - 120. * sx2 + 48. * sx4) / 192.;
        W[3] = (19. + 44. * sx + 24. * sx2 - 16. * sx3 - 16. * sx4) / 96.;
        W[4] = (1. + 8. * sx + 24. * sx2 + 32. * sx3 + 16. * sx4) / 384.;
    }
}

static double* cz(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* pair-style k-space tables and virial setup */
static void style_tables(ora_sys* s, ora_style* st) {
    int M = s->M, D = s->dim, nP = s->nP;
    st->u_k = cz(2 * (size_t)M);
    st->f_k = cz(2 * (size_t)M * D);
    st->f_r = cz((size_t)M * D);
    st->vir_k = cz(2 * (size_t)M * nP);
    double A = st->A, s2 = st->sigma2, Rp = st->Rp;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        double kv[3];
        double k2 = get_k(s, i, kv), u;
        if (st->kind == ORA_GAUSS) {                   /* pair_style_gaussian.cu:82-83 */
            u = A * exp(-k2 * s2 / 2.0);
        } else {
            double k = sqrt(k2), step;
            if (k2 == 0) step = 4.0 * ORA_PI * Rp * Rp * Rp / 3.0;
            else step = 4.0 * ORA_PI * (sin(Rp * k) - Rp * k * cos(Rp * k)) / (k2 * k);
            if (st->kind == ORA_ERF)                   /* pair_style_erf.cu:135-150 */
                u = (k2 == 0) ? A * step * step : A * exp(-k2 * s2 * 0.5) * step * step;
            else                                       /* pair_style_gaussian_erf.cu:96-107 */
                u = (k2 == 0) ? A * step : A * exp(-k2 * s2 * 0.5) * step;
        }
        st->u_k[2 * i] = u; st->u_k[2 * i + 1] = 0.0;
        for (int d = 0; d < D; ++d) {
            st->f_k[2 * (i * D + d)] = 0.0;
            st->f_k[2 * (i * D + d) + 1] = -kv[d] * u;
        }
    }
    /* real-space force: Re IFFT(f_k_d), pair_style_gaussian.cu:34-38 */
    double* c = s->cpx1;
    for (int d = 0; d < D; ++d) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) { c[2 * i] = st->f_k[2 * (i * D + d)]; c[2 * i + 1] = st->f_k[2 * (i * D + d) + 1]; }
        ora_fft(c, D, s->Nx, +1);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) st->f_r[i * D + d] = c[2 * i];
    }
    /* InitializeVirial, pair_style.cu:262-276 + d_initVirial device_array_utils.cu:456-494 */
    double* vr = cz((size_t)M * nP);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        int nn[3]; unstack(s, i, nn);
        double dr[3];
        for (int j = 0; j < D; ++j) {
            /* node position and its minimum image against the origin, wrap test in FP32 */
            float rf = (float)nn[j] * s->dxf[j];
            float df = rf - 0.0f;
            double dv = (double)nn[j] * s->dx[j];
            if (df > s->Lhf[j]) dv -= s->L[j];
            else if (df < -s->Lhf[j]) dv += s->L[j];
            dr[j] = dv;
        }
        const double* fr = st->f_r + (size_t)i * D;
        for (int j = 0; j < D; ++j) {
            vr[(size_t)i * nP + j] = dr[j] * fr[j];
            if (D == 2 && j == 0) vr[(size_t)i * nP + 2] = dr[1] * fr[0];
            if (D == 3 && j == 0) { vr[(size_t)i * nP + 3] = dr[1] * fr[0]; vr[(size_t)i * nP + 4] = dr[2] * fr[0]; }
            if (D == 3 && j == 1) vr[(size_t)i * nP + 5] = dr[2] * fr[1];
        }
    }
    for (int cc = 0; cc < nP; ++cc) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) { c[2 * i] = vr[(size_t)i * nP + cc]; c[2 * i + 1] = 0.0; }
        ora_fft(c, D, s->Nx, -1);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {
            st->vir_k[2 * ((size_t)i * nP + cc)] = c[2 * i] / M;
            st->vir_k[2 * ((size_t)i * nP + cc) + 1] = c[2 * i + 1] / M;
        }
    }
    free(vr);
}


/* init_device_fieldphase, pair_style_fieldphases.cu:89-185: the force table f_r(node)
   of a static field w(r) at node positions r = float(n) dx (FP32, d_get_r,
   device_array_utils.cu:34-41), arguments 2 PI n_periods / L with the reference's FP32
   PI; trigonometry in FP64.  Phases: 0 lamellar along dir, 2 cylinders along dir, 3
   gyroid (3D).  Phase 1 (BCC) has no branch in the reference (its table stays
   uninitialised) and is rejected, as are the combinations that index past Dim. */
static int field_tables(ora_sys* s, ora_style* st, char* err, int errlen) {
    int D = s->dim, M = s->M;
    const double PIf = (double)3.141592654f;
    if (st->phase == 1) { snprintf(err, errlen, "fieldphase phase 1 (BCC) is not defined by the reference"); return -1; }
    if (st->phase != 0 && st->phase != 2 && st->phase != 3) { snprintf(err, errlen, "unknown fieldphase phase %d", st->phase); return -1; }
    if (st->phase == 0 && (st->dir < 0 || st->dir >= D)) { snprintf(err, errlen, "fieldphase direction out of range"); return -1; }
    if (st->phase == 2 && (st->dir < 0 || st->dir > 2 || (D == 2 && st->dir != 2))) { snprintf(err, errlen, "fieldphase cylinder direction out of range"); return -1; }
    if (st->phase == 3 && D != 3) { snprintf(err, errlen, "fieldphase gyroid needs Dim 3"); return -1; }
    st->f_r = (double*)calloc((size_t)M * D, sizeof(double));
    double A = st->A;
    for (int i = 0; i < M; ++i) {
        int nn[3];
        unstack(s, i, nn);
        double r[3] = {0, 0, 0};
        for (int j = 0; j < D; ++j) r[j] = (double)((float)nn[j] * s->dxf[j]);
        double* f = st->f_r + (size_t)i * D;
        if (st->phase == 0) {
            double a = 2.0 * PIf * (double)st->n_periods / (double)s->Lf[st->dir];
            f[st->dir] = A * a * cos(a * r[st->dir]);
        } else if (st->phase == 2) {
            int d1 = 0, d2 = 1;
            if (st->dir == d1) d1 = 2;
            else if (st->dir == d2) d2 = 2;
            double a1 = 2.0 * PIf * (double)st->n_periods / (double)s->Lf[d1];
            double a2 = 2.0 * PIf * (double)st->n_periods / (double)s->Lf[d2];
            if (D == 3) f[st->dir] = 0.0;
            f[d1] = A * a1 * sin(a1 * r[d1]) * cos(a2 * r[d2]);
            f[d2] = A * a2 * sin(a2 * r[d2]) * cos(a1 * r[d1]);
        } else {
            double a[3], c[3], sn[3];
            for (int j = 0; j < 3; ++j) {
                a[j] = 2.0 * PIf * (double)st->n_periods / (double)s->Lf[j];
                c[j] = cos(a[j] * r[j]);
                sn[j] = sin(a[j] * r[j]);
            }
            double e = sn[0] * c[1] + sn[1] * c[2] + sn[2] * c[0];
            f[0] = -A * e * a[0] * (c[0] * c[1] - sn[2] * sn[0]);
            f[1] = -A * e * a[1] * (c[1] * c[2] - sn[0] * sn[1]);
            f[2] = -A * e * a[2] * (c[2] * c[0] - sn[1] * sn[2]);
        }
    }
    return 0;
}

/* FieldPhase::CalcForces, pair_style_fieldphases.cu:54-86 with d_prepareFieldForce
   (device_array_utils.cu:410-423) and d_accumulateGridForce (:362-372).  QUIRK: the
   prepared product rho_t f_d is multiplied by rho_t once more on accumulation, so the
   grid force is +rho_1^2 f_d on type1 and -rho_2^2 f_d on type2, and after the gather's
   division by rho a bead feels rho(node) f_d(node). */
static void calc_forces_field(ora_sys* s, const ora_style* st) {
    int M = s->M, D = s->dim;
    double* Fd[3] = {s->fx, s->fy, s->fz};
    for (int d = 0; d < D; ++d)
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {
            double fr = st->f_r[(size_t)i * D + d];
            double ra = s->rho[(size_t)st->t1 * M + i], rb = s->rho[(size_t)st->t2 * M + i];
            Fd[d][(size_t)st->t1 * M + i] += ra * (ra * fr);
            Fd[d][(size_t)st->t2 * M + i] += rb * (-rb * fr);
        }
}

int ora_setup(ora_sys* s, char* err, int errlen) {
    int D = s->dim;
    s->M = 1; s->S = 1; s->gvolf = 1.0f;
    for (int j = 0; j < D; ++j) {
        if (s->Nx[j] <= 0) ERR("grid size Nx[%d] = %d is invalid", j, s->Nx[j]);
        s->M *= s->Nx[j];
        s->dxf[j] = s->Lf[j] / (float)s->Nx[j];
        s->gvolf *= s->dxf[j];
        s->S *= (s->pmeorder + 1);
    }
    s->nP = (D == 2) ? 3 : 6;
    s->V = 1.0; s->gvol = 1.0;
    for (int j = 0; j < 3; ++j) { s->L[j] = s->Lf[j]; s->Lh[j] = s->Lhf[j]; s->dx[j] = s->dxf[j]; }
    for (int j = 0; j < D; ++j) { s->V *= s->L[j]; s->gvol *= s->dx[j]; }
    s->delt = s->delt_f;
    s->noise_mag = sqrt(2.0 * s->delt);
    size_t M = s->M, T = s->ntypes;
    s->rho = cz(T * M); s->fx = cz(T * M); s->fy = cz(T * M); s->fz = cz(T * M);
    s->rho_c = cz(M); s->phi = cz(M); s->efield = cz(D * M);
    s->fxc = cz(M); s->fyc = cz(M); s->fzc = cz(M);
    s->cpx1 = cz(2 * M); s->cpx2 = cz(2 * M);
    s->gridW = cz((size_t)s->ns * s->S);
    s->gridI = (int*)calloc((size_t)s->ns * s->S, sizeof(int));
    s->f = cz((size_t)s->ns * D);
    s->prev_noise = cz((size_t)s->ns * D);
    for (int k = 0; k < s->nst; ++k) {
        if (s->st[k].kind == ORA_FIELDPHASE) {
            if (field_tables(s, &s->st[k], err, errlen)) return -1;
        } else {
            style_tables(s, &s->st[k]);
        }
    }
    memcpy(s->xo, s->x, sizeof(double) * s->ns * D);   /* device_comm_utils.cu:15-16 */
    return 0;
}

/* d_charge_grid / d_charge_grid_charges, device_grid_utils.cu:111-343 (FP32 bins,
   FP64 weights).  W3 = Wx*Wy*Wz/gvol stored with the node for the gather. */
static void spread(ora_sys* s) {
    int D = s->dim, P = s->pmeorder, M = s->M, S = s->S, T = s->ntypes;
    memset(s->rho, 0, sizeof(double) * (size_t)T * M);
    memset(s->rho_c, 0, sizeof(double) * (size_t)M);
    int shift = P / 2;
    for (int i = 0; i < s->ns; ++i) {
        double W[3][6];
        int g[3];
        for (int j = 0; j < D; ++j) {
            float xf = (float)s->x[i * D + j], dxf = s->dxf[j];
            double gdx;
            if (P % 2 == 0) {
                float t = xf + 0.5f * dxf;
                g[j] = (int)(t / dxf);
                gdx = (double)xf - (double)g[j] * s->dx[j];
            } else {
                g[j] = (int)(xf / dxf);
                gdx = (double)xf - ((double)g[j] + 0.5) * s->dx[j];
            }
            spline_weights(gdx, s->dx[j], W[j], P);
        }
        int ct = 0, t = s->tp[i];
        int nz_n = (D == 3) ? P + 1 : 1;
        for (int ix = 0; ix <= P; ++ix) {
            int n0 = g[0] + ix - shift;
            if (n0 < 0) n0 += s->Nx[0]; else if (n0 >= s->Nx[0]) n0 -= s->Nx[0];
            for (int iy = 0; iy <= P; ++iy) {
                int n1 = g[1] + iy - shift;
                if (n1 < 0) n1 += s->Nx[1]; else if (n1 >= s->Nx[1]) n1 -= s->Nx[1];
                for (int iz = 0; iz < nz_n; ++iz) {
                    int mi; double w3;
                    if (D == 3) {
                        int n2 = g[2] + iz - shift;
                        if (n2 < 0) n2 += s->Nx[2]; else if (n2 >= s->Nx[2]) n2 -= s->Nx[2];
                        mi = n0 + (n1 + n2 * s->Nx[1]) * s->Nx[0];
                        w3 = W[0][ix] * W[1][iy] * W[2][iz] / s->gvol;
                    } else {
                        mi = n1 * s->Nx[0] + n0;
                        w3 = W[0][ix] * W[1][iy] / s->gvol;
                    }
                    s->rho[(size_t)t * M + mi] += w3;
                    if (s->do_charges) s->rho_c[mi] += w3 * s->q[i];
                    s->gridI[(size_t)i * S + ct] = mi;
                    s->gridW[(size_t)i * S + ct] = w3;
                    ct++;
                }
            }
        }
    }
}

/* PairStyle::CalcForces, pair_style.cu:81-196 */
static void calc_forces_style(ora_sys* s, ora_style* st) {
    int M = s->M, D = s->dim;
    double *c1 = s->cpx1, *c2 = s->cpx2;
    double* Fd[3] = {s->fx, s->fy, s->fz};
    for (int pass = 0; pass < 2; ++pass) {
        int src = pass == 0 ? st->t2 : st->t1, dst = pass == 0 ? st->t1 : st->t2;
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) { c2[2 * i] = s->rho[(size_t)src * M + i]; c2[2 * i + 1] = 0.0; }
        ora_fft(c2, D, s->Nx, -1);
        for (int d = 0; d < D; ++d) {
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < M; ++i) {        /* d_prepareForceKSpace :298-312 */
                double fr = st->f_k[2 * (i * D + d)], fi = st->f_k[2 * (i * D + d) + 1];
                c1[2 * i] = (fr * c2[2 * i] - fi * c2[2 * i + 1]) / M;
                c1[2 * i + 1] = (fr * c2[2 * i + 1] + fi * c2[2 * i]) / M;
            }
            ora_fft(c1, D, s->Nx, +1);
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < M; ++i)          /* d_accumulateGridForce :362-372 */
                Fd[d][(size_t)dst * M + i] += s->rho[(size_t)dst * M + i] * c1[2 * i];
        }
    }
}

/* Charges::CalcCharges, pair_style_charges.cu:21-111; kernels device_array_utils.cu:186-279 */
static void calc_charges(ora_sys* s) {
    int M = s->M, D = s->dim;
    double *c1 = s->cpx1, *c2 = s->cpx2;
    double lb = s->bjerrum_f, a = s->smear_f;
    for (int i = 0; i < M; ++i) { c1[2 * i] = s->rho_c[i]; c1[2 * i + 1] = 0.0; }
    memcpy(c2, c1, sizeof(double) * 2 * M);
    ora_fft(c2, D, s->Nx, -1);
    for (int i = 0; i < 2 * M; ++i) c2[i] /= M;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {                /* d_prepareElectrostaticPotential */
        double kv[3];
        double k2 = get_k(s, i, kv);
        if (k2 != 0) {
            double g = 4.0 * ORA_PI * lb / k2 * exp(-k2 / (2.0 * a * a));
            c1[2 * i] = c2[2 * i] * g; c1[2 * i + 1] = c2[2 * i + 1] * g;
        } else { c1[2 * i] = 0.0; c1[2 * i + 1] = 0.0; }
    }
    double* Fc[3] = {s->fxc, s->fyc, s->fzc};
    for (int d = 0; d < D; ++d) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {            /* d_prepareElectricField :225-241 */
            double kv[3];
            double k2 = get_k(s, i, kv);
            double e = exp(-k2 / (2.0 * a * a));
            c2[2 * i] = kv[d] * c1[2 * i + 1] * e;
            c2[2 * i + 1] = -kv[d] * c1[2 * i] * e;
        }
        ora_fft(c2, D, s->Nx, +1);
        for (int i = 0; i < M; ++i) { Fc[d][i] = c2[2 * i]; s->efield[(size_t)d * M + i] = c2[2 * i]; }
    }
    ora_fft(c1, D, s->Nx, +1);
    for (int i = 0; i < M; ++i) s->phi[i] = c1[2 * i];
}

/* gathers, device_grid_utils.cu:346-452; charge gather first (forces.cu:144-169) */
static void gather(ora_sys* s) {
    int D = s->dim, S = s->S, M = s->M;
    memset(s->f, 0, sizeof(double) * s->ns * D);
    double* Fd[3] = {s->fx, s->fy, s->fz};
    double* Fc[3] = {s->fxc, s->fyc, s->fzc};
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ns; ++i) {
        if (s->do_charges)
            for (int m = 0; m < S; ++m) {
                int gi = s->gridI[(size_t)i * S + m];
                double w = s->gridW[(size_t)i * S + m];
                for (int d = 0; d < D; ++d) s->f[i * D + d] += Fc[d][gi] * w * s->gvol * s->q[i];
            }
        int t = s->tp[i];
        for (int m = 0; m < S; ++m) {
            int gi = s->gridI[(size_t)i * S + m];
            double w = s->gridW[(size_t)i * S + m];
            size_t ti = (size_t)t * M + gi;
            if (s->rho[ti] > 0.0)
                for (int d = 0; d < D; ++d) s->f[i * D + d] += Fd[d][ti] * w * s->gvol / s->rho[ti];
        }
    }
}

/* d_bonds, device_bonds.cu:4-60 */
static void bonds(ora_sys* s) {
    int D = s->dim;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ns; ++i) {
        double lf[3] = {0, 0, 0};
        for (int b = 0; b < s->n_bonds[i]; ++b) {
            int j = s->bonded_to[i * ORA_MAX_BONDS + b], bt = s->bond_type[i * ORA_MAX_BONDS + b];
            double dr[3], r2 = 0.0;
            for (int d = 0; d < D; ++d) {
                dr[d] = s->x[i * D + d] - s->x[j * D + d];
                if (dr[d] > s->Lh[d]) dr[d] -= s->L[d];
                else if (dr[d] < -s->Lh[d]) dr[d] += s->L[d];
                r2 += dr[d] * dr[d];
            }
            double mf = 0.0;
            if (r2 > 1.0e-4) { double r = sqrt(r2); mf = 2.0 * s->bond_kf[bt] * (r - s->bond_reqf[bt]) / r; }
            for (int d = 0; d < D; ++d) lf[d] -= mf * dr[d];
        }
        for (int d = 0; d < D; ++d) s->f[i * D + d] += lf[d];
    }
}

/* forces(), forces.cu:42-196 */
void ora_forces(ora_sys* s) {
    size_t TM = (size_t)s->ntypes * s->M;
    memset(s->fx, 0, sizeof(double) * TM);
    memset(s->fy, 0, sizeof(double) * TM);
    memset(s->fz, 0, sizeof(double) * TM);
    spread(s);
    static const int order[3] = {ORA_GAUSS, ORA_ERF, ORA_GAUSS_ERF};
    for (int o = 0; o < 3; ++o) {
        for (int k = 0; k < s->nst; ++k)
            if (s->st[k].kind == order[o]) calc_forces_style(s, &s->st[k]);
        if (o == 1)   /* forces.cu:108-126: Gaussians, Erfs, FieldPhases, GaussianErfs */
            for (int k = 0; k < s->nst; ++k)
                if (s->st[k].kind == ORA_FIELDPHASE) calc_forces_field(s, &s->st[k]);
    }
    if (s->do_charges) calc_charges(s);
    gather(s);
    if (s->n_total_bonds > 0) bonds(s);
    s->forces_done = 1;
}

/* Integrator::Integrate_2 (integrator.cu:24-39) with d_EM_integrator
   (device_EM_integrator.cu:5-36) and d_GJF_integrator (device_GJF_integrator.cu:4-62).
   Noise: standard normals from Philox at counter (site, step, 0, 0). */
void ora_integrate(ora_sys* s) {
    int D = s->dim;
    double* old = s->prev_noise;                 /* zero at setup, initialize.cu:120 */
    for (int k = 0; k < s->ninteg; ++k) {
        const ora_group* g = &s->grp[s->integ[k].group];
        #pragma omp parallel for schedule(static)
        for (int l = 0; l < g->nsites; ++l) {
            int i = g->index[l];
            double z[4];
            ora_normals(s->seed, (uint32_t)i, (uint32_t)s->step, 0u, z);
            if (s->integ[k].kind == 0) {
                for (int j = 0; j < D; ++j) {
                    double* xp = &s->x[i * D + j];
                    *xp = *xp + s->delt * s->f[i * D + j] + s->noise_mag * z[j];
                    if (*xp > s->L[j]) *xp -= s->L[j];
                    else if (*xp < 0.0) *xp += s->L[j];
                }
            } else {
                int t = s->tp[i];
                double m = s->massf[t], Df = s->difff[t], dt = s->delt;
                double b = m / (m + dt / (2.0 * Df));
                double a = (1.0 - dt / (2.0 * Df * m)) / (1.0 + dt / (2.0 * Df * m));
                for (int j = 0; j < D; ++j) {
                    int ai = i * D + j;
                    double nn = s->noise_mag / Df * z[j];
                    double xt = s->x[ai];
                    if (s->xo[ai] - s->x[ai] > s->L[j] / 2.0) s->xo[ai] -= s->L[j];
                    else if (s->xo[ai] - s->x[ai] < -s->L[j] / 2.0) s->xo[ai] += s->L[j];
                    s->x[ai] = 2.0 * b * s->x[ai] - a * s->xo[ai] + b * dt * dt / m * s->f[ai]
                               + b * dt / (2.0 * m) * (nn + old[ai]);
                    s->xo[ai] = xt;
                    old[ai] = nn;
                    if (s->x[ai] > s->L[j]) s->x[ai] -= s->L[j];
                    else if (s->x[ai] < 0.0) s->x[ai] += s->L[j];
                }
            }
        }
    }
}

/* PairStyle::Update, pair_style.cu:308-328, after step `step` (1-based): the table u_k
 * and the force tables f_k are multiplied by new/cur, with
 *   delta = 1000 (A0 - A_final) / max_steps,  new = 1000 A0 + delta step + delta,
 *   cur = 1000 Re u_k[0]   (the current k = 0 value, not the current prefactor);
 * vir_k is left as it is.  The reference does this in FP32; here in FP64. */
static void update_style(ora_sys* s, ora_style* st, int step) {
    if (!st->ramp || st->kind == ORA_FIELDPHASE) return;
    int M = s->M, D = s->dim;
    double delta_A = (1000.0 * (st->A - st->A_final)) / (double)s->max_steps;
    double cur_A = 1000.0 * st->A + delta_A * (double)step;
    double new_A = cur_A + delta_A;
    cur_A = st->u_k[0] * 1000.0;
    double r = new_A / cur_A;
    for (int i = 0; i < 2 * M; ++i) st->u_k[i] *= r;
    for (int i = 0; i < 2 * M * D; ++i) st->f_k[i] *= r;
}

void ora_run(ora_sys* s, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {   /* main.cu:106-117 */
        /* update_pairstyles() of the previous step (main.cu:199), applied when the next
         * step starts, so that properties evaluated between steps see the table of the
         * step just taken, as the reference's log does (calc_properties precedes
         * update_pairstyles in the step) */
        if (s->step > 0)
            for (int j = 0; j < s->nst; ++j) update_style(s, &s->st[j], s->step);
        s->step += 1;
        ora_forces(s);
        ora_integrate(s);
    }
}

/* Compute::doCompute for avg_sk, Compute.cu:83-117: d_prepareDensity (the type's
 * density as a complex grid), cufftExecC2C forward, d_multiplyComplex(rho_k, rho_k)
 * -- (a.x b.x - a.y b.y, a.x b.y + a.y b.x) / M with a = b, the quirk of squaring
 * rho(k) rather than taking |rho(k)|^2 -- summed over the calls. */
void ora_sk_accumulate(ora_sys* s, int t) {
    const int M = s->M;
    if (!s->forces_done) ora_forces(s);
    if (!s->sk) {
        s->sk = calloc((size_t)s->ntypes * 2 * M, sizeof(double));
        s->sk_count = calloc((size_t)s->ntypes, sizeof(int));
    }
    double* c = s->cpx1;
    for (int i = 0; i < M; ++i) { c[2 * i] = s->rho[(size_t)t * M + i]; c[2 * i + 1] = 0.0; }
    ora_fft(c, s->dim, s->Nx, -1);
    double* acc = s->sk + (size_t)t * 2 * M;
    for (int i = 0; i < M; ++i) {
        const double x = c[2 * i], y = c[2 * i + 1];
        acc[2 * i] += (x * x - y * y) / M;
        acc[2 * i + 1] += (x * y + y * x) / M;
    }
    s->sk_count[t] += 1;
}

/* Compute::writeResults, Compute.cu:121-137: the sum over the number of calls */
int ora_sk_average(const ora_sys* s, int t, double* out) {
    const int M = s->M;
    const int n = s->sk ? s->sk_count[t] : 0;
    for (int i = 0; i < 2 * M; ++i) out[i] = n > 0 ? s->sk[(size_t)t * 2 * M + i] / n : 0.0;
    return n;
}

/* PairStyle::CalcEnergy / CalcVirial, pair_style.cu:204-260; integ_trapPBC integ_utils.cu:3-13 */
static void style_energy_virial(ora_sys* s, ora_style* st) {
    int M = s->M, D = s->dim, nP = s->nP;
    double *c1 = s->cpx1, *c2 = s->cpx2;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) { c2[2 * i] = s->rho[(size_t)st->t2 * M + i]; c2[2 * i + 1] = 0.0; }
    ora_fft(c2, D, s->Nx, -1);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        double ur = st->u_k[2 * i], ui = st->u_k[2 * i + 1];
        c1[2 * i] = (ur * c2[2 * i] - ui * c2[2 * i + 1]) / M;
        c1[2 * i + 1] = (ur * c2[2 * i + 1] + ui * c2[2 * i]) / M;
    }
    ora_fft(c1, D, s->Nx, +1);
    double sum = 0.0;
    for (int i = 0; i < M; ++i) sum += c1[2 * i] * s->rho[(size_t)st->t1 * M + i];
    st->energy = sum * s->gvol;
    for (int i = 0; i < 2 * M; ++i) c2[i] /= M;           /* d_divideByDimension */
    for (int c = 0; c < nP; ++c) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {                       /* d_prepareVirial :496-507 */
            double vr = st->vir_k[2 * ((size_t)i * nP + c)], vi = st->vir_k[2 * ((size_t)i * nP + c) + 1];
            c1[2 * i] = vr * c2[2 * i] - vi * c2[2 * i + 1];
            c1[2 * i + 1] = vi * c2[2 * i] + vr * c2[2 * i + 1];
        }
        ora_fft(c1, D, s->Nx, +1);
        double sv = 0.0;
        for (int i = 0; i < M; ++i) sv += c1[2 * i] * s->rho[(size_t)st->t1 * M + i];
        st->total_vir[c] = sv * s->gvol / s->V / (double)D;
    }
}

/* host_bonds, bonds.cu:3-52 (no r^2 guard; each bond once with id2 >= id1) */
static void host_bonds(ora_sys* s) {
    int D = s->dim;
    s->Ubond = 0.0;
    for (int c = 0; c < 6; ++c) s->bondVir[c] = 0.0;
    for (int i = 0; i < s->ns; ++i)
        for (int b = 0; b < s->n_bonds[i]; ++b) {
            int j = s->bonded_to[i * ORA_MAX_BONDS + b];
            if (j < i) continue;
            int bt = s->bond_type[i * ORA_MAX_BONDS + b];
            double dr[3] = {0, 0, 0}, r2 = 0.0;
            for (int d = 0; d < D; ++d) {
                dr[d] = s->x[i * D + d] - s->x[j * D + d];
                if (dr[d] > s->Lh[d]) dr[d] -= s->L[d];
                else if (dr[d] < -s->Lh[d]) dr[d] += s->L[d];
                r2 += dr[d] * dr[d];
            }
            double r = sqrt(r2), del = r - s->bond_reqf[bt];
            s->Ubond += del * del * s->bond_kf[bt];
            double mf = 2.0 * s->bond_kf[bt] * del / r;
            s->bondVir[0] += -mf * dr[0] * dr[0];
            s->bondVir[1] += -mf * dr[1] * dr[1];
            if (D == 2) s->bondVir[2] += -mf * dr[0] * dr[1];
            else {
                s->bondVir[2] += -mf * dr[2] * dr[2];
                s->bondVir[3] += -mf * dr[0] * dr[1];
                s->bondVir[4] += -mf * dr[0] * dr[2];
                s->bondVir[5] += -mf * dr[1] * dr[2];
            }
        }
    for (int c = 0; c < s->nP; ++c) s->bondVir[c] *= 1.0 / s->V / (double)D;
}

/* calc_properties, calc_properties.cu:20-116.  Non-bonded terms use the densities of
   the last ora_forces() call; bonds use the current positions (calc_bondProps :49-59). */
int ora_properties(ora_sys* s, double* out, int nout) {
    if (!s->forces_done) ora_forces(s);
    static const int order[3] = {ORA_GAUSS, ORA_ERF, ORA_GAUSS_ERF};
    s->Upe = 0.0;
    for (int c = 0; c < 6; ++c) s->Ptens[c] = 0.0;
    int n = 9;
    double es[256];
    int nes = 0;
    for (int o = 0; o < 3; ++o)
        for (int k = 0; k < s->nst; ++k)
            if (s->st[k].kind == order[o]) {
                style_energy_virial(s, &s->st[k]);
                s->Upe += s->st[k].energy;
                for (int c = 0; c < s->nP; ++c) s->Ptens[c] += s->st[k].total_vir[c];
                if (nes < 256) es[nes++] = s->st[k].energy;
            }
    s->Uelec = 0.0;
    if (s->do_charges) {
        for (int i = 0; i < s->M; ++i) s->Uelec += s->phi[i] * s->rho_c[i];
        s->Upe += s->Uelec;
    }
    host_bonds(s);
    s->Upe += s->Ubond;
    for (int c = 0; c < s->nP; ++c) s->Ptens[c] += s->bondVir[c];
    double vals[9 + 256];
    vals[0] = s->Upe; vals[1] = s->Ubond; vals[2] = s->Uelec;
    for (int c = 0; c < 6; ++c) vals[3 + c] = s->Ptens[c];
    for (int k = 0; k < nes; ++k) vals[n + k] = es[k];
    int tot = n + nes;
    for (int k = 0; k < nout && k < tot; ++k) out[k] = vals[k];
    return tot;
}

/* NList::Initialize / AssignCells / BuildNList, nlist.cu:13-228.  The reference's slot
   order is nondeterministic (atomics); here each list is sorted ascending.  A site at
   exactly x == L maps to cell 0 (the reference indexes one past the end, nlist.cu:217). */
int ora_nlist(ora_sys* s, double rc, int group, int* cell_of, int* counts, int* neigh,
              int neigh_cap, int* cell_max_out, int* ncells_out) {
    int D = s->dim;
    if (group < 0 || group >= s->ngrp) return -1;
    const ora_group* g = &s->grp[group];
    float rcf = (float)rc;
    int nc[3] = {1, 1, 1}, total = 1;
    float Lc[3] = {1, 1, 1}, cell_vol = 1.0f;
    for (int j = 0; j < D; ++j) {
        nc[j] = (int)(s->Lf[j] / rcf);
        if (nc[j] < 1) return -2;
        Lc[j] = s->Lf[j] / (float)nc[j];
        total *= nc[j];
        cell_vol *= Lc[j];
    }
    int cmax = (int)((float)(s->ns * cell_vol / s->Vf) * 7.5);
    if (cell_max_out) *cell_max_out = cmax;
    if (ncells_out) *ncells_out = total;
    int* ccount = (int*)calloc(total, sizeof(int));
    int* cinds = (int*)calloc((size_t)total * (cmax > 0 ? cmax : 1), sizeof(int));
    for (int i = 0; i < s->ns; ++i) { counts[i] = 0; cell_of[i] = -1; }
    int overflow = 0;
    for (int l = 0; l < g->nsites; ++l) {                      /* d_assign_cells */
        int id = g->index[l], cid[3] = {0, 0, 0};
        for (int j = 0; j < D; ++j) {
            cid[j] = (int)((float)s->x[id * D + j] / Lc[j]);
            if (cid[j] >= nc[j]) cid[j] -= nc[j];
            if (cid[j] < 0) cid[j] += nc[j];
        }
        int cn = (D == 2) ? cid[0] + nc[0] * cid[1] : cid[0] + (cid[1] + cid[2] * nc[1]) * nc[0];
        if (ccount[cn] < cmax) cinds[(size_t)cn * cmax + ccount[cn]] = id; else overflow = 1;
        ccount[cn]++;
        cell_of[id] = cn;
    }
    if (overflow) { free(ccount); free(cinds); return -3; }
    float rc2 = rcf * rcf;
    for (int l = 0; l < g->nsites; ++l) {                      /* d_make_nlist */
        int id = g->index[l], cn = cell_of[id], nn[3];
        if (D == 2) { nn[1] = cn / nc[0]; nn[0] = cn % nc[0]; nn[2] = 1; }
        else { nn[2] = cn / (nc[1] * nc[0]); int r = cn - nn[2] * nc[0] * nc[1]; nn[1] = r / nc[0]; nn[0] = r % nc[0]; }
        int cnt = 0;
        for (int ix = nn[0] - 1; ix <= nn[0] + 1; ++ix) {
            int nix = ix < 0 ? nc[0] - 1 : (ix == nc[0] ? 0 : ix);
            for (int iy = nn[1] - 1; iy <= nn[1] + 1; ++iy) {
                int niy = iy < 0 ? nc[1] - 1 : (iy == nc[1] ? 0 : iy);
                for (int iz = nn[2] - 1; iz <= nn[2] + 1; ++iz) {
                    int niz = iz;
                    if (D == 2) niz = 0;
                    else niz = iz < 0 ? nc[2] - 1 : (iz == nc[2] ? 0 : iz);
                    int ncn = (D == 2) ? nix + nc[0] * niy : nix + (niy + niz * nc[1]) * nc[0];
                    for (int q = 0; q < ccount[ncn]; ++q) {
                        int nid = cinds[(size_t)ncn * cmax + q];
                        float r2 = 0.0f;
                        for (int j = 0; j < D; ++j) {
                            float dr = (float)s->x[id * D + j] - (float)s->x[nid * D + j];
                            if (dr < -s->Lhf[j]) dr += s->Lf[j];
                            else if (dr > s->Lhf[j]) dr -= s->Lf[j];
                            r2 += dr * dr;
                        }
                        if (r2 < rc2) {
                            if (cnt < neigh_cap) neigh[(size_t)id * neigh_cap + cnt] = nid;
                            cnt++;
                        }
                    }
                    if (D == 2) break;
                }
            }
        }
        counts[id] = cnt;
        int m = cnt < neigh_cap ? cnt : neigh_cap;
        int* a = neigh + (size_t)id * neigh_cap;
        for (int p = 1; p < m; ++p) { int v = a[p], q = p - 1; while (q >= 0 && a[q] > v) { a[q + 1] = a[q]; --q; } a[q + 1] = v; }
    }
    free(ccount); free(cinds);
    return 0;
}

/* ------------------------------------------------------------------------ */
ora_sys* ora_create(const char* text, const char* base_dir, uint32_t seed_override,
                    int use_seed_override, char* err, int errlen) {
    ora_sys* s = (ora_sys*)calloc(1, sizeof(ora_sys));
    if (!s) return NULL;
    if (err && errlen > 0) err[0] = 0;
    char dummy[8];
    if (!err) { err = dummy; errlen = sizeof dummy; }
    if (ora_parse(s, text, base_dir, err, errlen) || ora_setup(s, err, errlen)) { ora_destroy(s); return NULL; }
    if (use_seed_override) s->seed = seed_override;
    return s;
}

void ora_destroy(ora_sys* s) {
    if (!s) return;
    for (int k = 0; k < s->nst; ++k) { free(s->st[k].u_k); free(s->st[k].f_k); free(s->st[k].f_r); free(s->st[k].vir_k); free(s->st[k].u_r); }
    free(s->st);
    for (int g = 0; g < s->ngrp; ++g) free(s->grp[g].index);
    free(s->grp); free(s->integ);
    free(s->x); free(s->xo); free(s->tp); free(s->molec); free(s->q); free(s->massf); free(s->difff);
    free(s->n_bonds); free(s->bonded_to); free(s->bond_type); free(s->bond_kf); free(s->bond_reqf);
    free(s->rho); free(s->fx); free(s->fy); free(s->fz); free(s->rho_c); free(s->phi); free(s->efield);
    free(s->fxc); free(s->fyc); free(s->fzc); free(s->cpx1); free(s->cpx2); free(s->gridW); free(s->gridI); free(s->f); free(s->prev_noise);
    free(s->sk); free(s->sk_count);
    free(s);
}

int ora_get_int(const ora_sys* s, const char* k) {
    if (!strcmp(k, "ns")) return s->ns;
    if (!strcmp(k, "dim")) return s->dim;
    if (!strcmp(k, "ntypes")) return s->ntypes;
    if (!strcmp(k, "M")) return s->M;
    if (!strcmp(k, "Nx0")) return s->Nx[0];
    if (!strcmp(k, "Nx1")) return s->Nx[1];
    if (!strcmp(k, "Nx2")) return s->Nx[2];
    if (!strcmp(k, "pmeorder")) return s->pmeorder;
    if (!strcmp(k, "S")) return s->S;
    if (!strcmp(k, "nP")) return s->nP;
    if (!strcmp(k, "step")) return s->step;
    if (!strcmp(k, "do_charges")) return s->do_charges;
    if (!strcmp(k, "nstyles")) return s->nst;
    if (!strcmp(k, "n_total_bonds")) return s->n_total_bonds;
    if (!strcmp(k, "nbond_types")) return s->nbond_types;
    if (!strcmp(k, "max_steps")) return s->max_steps;
    if (!strcmp(k, "log_freq")) return s->log_freq;
    if (!strcmp(k, "ngroups")) return s->ngrp;
    if (!strcmp(k, "nintegrators")) return s->ninteg;
    if (!strcmp(k, "seed")) return (int)s->seed;
    return -1;
}

double ora_get_double(const ora_sys* s, const char* k) {
    if (!strcmp(k, "L0")) return s->L[0];
    if (!strcmp(k, "L1")) return s->L[1];
    if (!strcmp(k, "L2")) return s->L[2];
    if (!strcmp(k, "dx0")) return s->dx[0];
    if (!strcmp(k, "dx1")) return s->dx[1];
    if (!strcmp(k, "dx2")) return s->dx[2];
    if (!strcmp(k, "gvol")) return s->gvol;
    if (!strcmp(k, "V")) return s->V;
    if (!strcmp(k, "delt")) return s->delt;
    if (!strcmp(k, "noise_mag")) return s->noise_mag;
    if (!strcmp(k, "bjerrum")) return s->bjerrum_f;
    if (!strcmp(k, "smear")) return s->smear_f;
    return NAN;
}

/* Test-only settings.  "noise_mag": the thermal-noise amplitude sqrt(2 delt) of the
   EM and GJF updates (0 switches the noise off, for closed-form checks of the drift
   terms).  Returns 0, or -1 for an unknown key. */
int ora_set_double(ora_sys* s, const char* k, double v) {
    if (!strcmp(k, "noise_mag")) { s->noise_mag = v; return 0; }
    return -1;
}

void ora_get_positions(const ora_sys* s, double* out) { memcpy(out, s->x, sizeof(double) * s->ns * s->dim); }
void ora_set_positions(ora_sys* s, const double* in) {
    memcpy(s->x, in, sizeof(double) * s->ns * s->dim);
    memcpy(s->xo, in, sizeof(double) * s->ns * s->dim);
    s->forces_done = 0;
}
void ora_get_types(const ora_sys* s, int* out) { memcpy(out, s->tp, sizeof(int) * s->ns); }
void ora_get_molecules(const ora_sys* s, int* out) { memcpy(out, s->molec, sizeof(int) * s->ns); }
void ora_get_charges(const ora_sys* s, double* out) { memcpy(out, s->q, sizeof(double) * s->ns); }
void ora_get_bonds(const ora_sys* s, int* nb, int* bt, int* ty) {
    memcpy(nb, s->n_bonds, sizeof(int) * s->ns);
    memcpy(bt, s->bonded_to, sizeof(int) * s->ns * ORA_MAX_BONDS);
    memcpy(ty, s->bond_type, sizeof(int) * s->ns * ORA_MAX_BONDS);
}
int ora_group_sites(const ora_sys* s, int g, int* out) {
    if (g < 0 || g >= s->ngrp) return -1;
    if (out) memcpy(out, s->grp[g].index, sizeof(int) * s->grp[g].nsites);
    return s->grp[g].nsites;
}
void ora_get_forces(const ora_sys* s, double* out) { memcpy(out, s->f, sizeof(double) * s->ns * s->dim); }
void ora_get_density(const ora_sys* s, int t, double* out) { memcpy(out, s->rho + (size_t)t * s->M, sizeof(double) * s->M); }
void ora_get_charge_density(const ora_sys* s, double* out) { memcpy(out, s->rho_c, sizeof(double) * s->M); }
void ora_get_grid_force(const ora_sys* s, int t, int d, double* out) {
    const double* F = d == 0 ? s->fx : (d == 1 ? s->fy : s->fz);
    memcpy(out, F + (size_t)t * s->M, sizeof(double) * s->M);
}
void ora_get_electric_field(const ora_sys* s, int d, double* out) { memcpy(out, s->efield + (size_t)d * s->M, sizeof(double) * s->M); }
void ora_get_grid_stencil(const ora_sys* s, int* inds, double* W) {
    memcpy(inds, s->gridI, sizeof(int) * (size_t)s->ns * s->S);
    memcpy(W, s->gridW, sizeof(double) * (size_t)s->ns * s->S);
}
