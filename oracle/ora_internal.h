/* ora_internal.h -- private state of the FP64 oracle.  TEST INFRASTRUCTURE ONLY. */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H
#include "tild_oracle.h"
#include <stddef.h>

#define ORA_MAX_BONDS 3      /* globals.h:26 */
#define ORA_PI 3.14159265358979323846

enum { ORA_GAUSS = 0, ORA_ERF = 1, ORA_FIELDPHASE = 2, ORA_GAUSS_ERF = 3 };

typedef struct {
    int kind, t1, t2;
    double A, A_final, sigma2, Rp;   /* A, sigma^2 hold the values the parser stores */
    int ramp, phase, dir, n_periods;
    double* u_k;    /* M complex, interleaved re,im */
    double* f_k;    /* M*D complex, [idx*D+d] */
    double* f_r;    /* M*D real: Re IFFT(f_k) (pair_style_gaussian.cu:34-38) */
    double* vir_k;  /* M*nP complex, [idx*nP+c] */
    double* u_r;    /* fieldphase: real-space field */
    double energy, total_vir[6];
} ora_style;

typedef struct { char name[64]; int nsites; int* index; } ora_group;
typedef struct { int kind; /* 0 EM, 1 GJF */ int group; } ora_integ;

struct ora_sys {
    /* input-deck state (read_input.cu:536-562 defaults) */
    int dim, Nx[3], pmeorder, max_steps, log_freq, grid_freq, traj_freq, bin_freq,
        struc_freq, skip_steps, threads;
    uint32_t seed;
    float delt_f, noise_mag_f;
    int do_charges;
    float bjerrum_f, smear_f;
    /* data file */
    int ns, n_total_bonds, n_total_angles, ntypes, nbond_types, nangle_types;
    float Lf[3], Lhf[3], Vf;
    double* x;       /* ns*D positions (FP64 state; file values are FP32) */
    double* xo;      /* GJF previous positions */
    int *tp, *molec;
    double* q;
    float* massf;    /* per type */
    float* difff;
    int *n_bonds, *bonded_to, *bond_type;
    float *bond_kf, *bond_reqf;
    /* styles, groups, integrators */
    ora_style* st; int nst;
    ora_group* grp; int ngrp;
    ora_integ* integ; int ninteg;
    /* derived (initialize.cu:32-48) */
    int M, S, nP;
    float dxf[3], gvolf;
    double L[3], Lh[3], V, dx[3], gvol, delt, noise_mag;
    /* grids & scratch */
    double *rho, *fx, *fy, *fz;          /* T*M each */
    double *rho_c, *phi, *efield;        /* M, M, D*M */
    double *fxc, *fyc, *fzc;             /* M each */
    double *cpx1, *cpx2;                 /* M complex */
    double* gridW; int* gridI;           /* ns*S */
    double* f;                           /* ns*D */
    double* prev_noise;                  /* ns*D GJF previous noise */
    int step;
    int forces_done;
    /* properties */
    double Upe, Ubond, Uelec, Ptens[6], bondVir[6];
    /* avg_sk computes: per type, sum of rho(k) rho(k) / M (M complex) and counts */
    double* sk;
    int* sk_count;
};

/* parser (ora_input.c) */
int ora_parse(ora_sys* s, const char* text, const char* base_dir, char* err, int errlen);
/* setup after parsing (ora_step.c) */
int ora_setup(ora_sys* s, char* err, int errlen);
void ora_free_all(ora_sys* s);

#endif
