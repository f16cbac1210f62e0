/*
 * tild_oracle.h -- FP64 CPU restatement of the reference TILD per-step routines.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it.  The product library never
 * links, loads or calls it, and nothing here is ever the measured path.
 *
 * Conventions (all follow the reference, cited per function in the .c files):
 *   - integer indexing (grid bins, cell ids, bond tables) is computed with FP32
 *     semantics from FP32 positions and FP32 grid spacings, exactly as the
 *     reference kernels do, so that it can be compared bit-exactly;
 *   - all floating arithmetic after indexing is FP64;
 *   - grids are x-fastest: idx = nx + (ny + nz*Ny)*Nx; particles are AoS x[i*D+d];
 *   - FFTs are unnormalised C2C, forward e^{-ikr}, inverse e^{+ikr} (cuFFT convention);
 *   - thermal noise is Philox4x32-10 (Salmon et al., SC'11) + Box-Muller, keyed by
 *     (seed, 0) with counter (particle, step, integrator, 0): the same stream the
 *     product draws, replacing the reference's cuRAND XORWOW (whose values are not
 *     reproduced: that part of the reference has no golden vectors).
 *
 * Parity status per function is listed in the header comment of each .c file.
 */
#ifndef TILD_ORACLE_H
#define TILD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_sys ora_sys;

/* --- primitives ------------------------------------------------------- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 4 standard normals for (seed, particle, step, stream) */
void ora_normals(uint32_t seed, uint32_t particle, uint32_t step, uint32_t stream, double z[4]);
/* in-place unnormalised C2C FFT of an interleaved (re,im) FP64 grid, x-fastest */
int ora_fft(double* data, int dim, const int* Nx, int sign);
/* 1-D FFT helper for tests */
int ora_fft1d(double* data, int n, int sign);
/* worker threads of the parallel loops; results are bit-identical for any count */
void ora_set_threads(int n);
int  ora_get_threads(void);

/* --- system ----------------------------------------------------------- */
/* parse an input deck (read_input.cu grammar); data files resolve against base_dir */
ora_sys* ora_create(const char* input_text, const char* base_dir, uint32_t seed_override,
                    int use_seed_override, char* err, int errlen);
void ora_destroy(ora_sys* s);

int    ora_get_int(const ora_sys* s, const char* key);
double ora_get_double(const ora_sys* s, const char* key);
/* test-only settings: "noise_mag" (0 = noise off); 0 on success, -1 unknown key */
int    ora_set_double(ora_sys* s, const char* key, double v);

void ora_get_positions(const ora_sys* s, double* out);       /* ns*D */
void ora_set_positions(ora_sys* s, const double* in);        /* also resets GJF xo = x */
void ora_get_types(const ora_sys* s, int* out);              /* ns */
void ora_get_molecules(const ora_sys* s, int* out);          /* ns */
void ora_get_charges(const ora_sys* s, double* out);         /* ns (0 when uncharged) */
void ora_get_bonds(const ora_sys* s, int* n_bonds, int* bonded_to, int* bond_type); /* ns, ns*3, ns*3 */
int  ora_group_sites(const ora_sys* s, int group, int* out); /* returns nsites */

/* forces.cu:42 -- spread, k-space pair forces, charges, gather, bonds */
void ora_forces(ora_sys* s);
void ora_get_forces(const ora_sys* s, double* out);          /* ns*D */
void ora_get_density(const ora_sys* s, int type, double* out);       /* M */
void ora_get_charge_density(const ora_sys* s, double* out);          /* M */
void ora_get_grid_force(const ora_sys* s, int type, int dir, double* out); /* M, rho*conv */
void ora_get_electric_field(const ora_sys* s, int dir, double* out); /* M */
void ora_get_grid_stencil(const ora_sys* s, int* inds, double* W);   /* ns*S each */

/* calc_properties.cu:20 -- out: [Upe, Ubond, Uelec, Ptens[6], n_style energies...] */
int  ora_properties(ora_sys* s, double* out, int nout);

/* Compute::doCompute, avg_sk (Compute.cu:83-117): adds rho_t(k) rho_t(k) / M of the
   densities of the last force evaluation (the complex square of the unnormalised
   forward transform, not |rho(k)|^2) to type t's sum.  ora_sk_average writes the
   average over all M wavevectors (interleaved re, im, x fastest; zeros before the
   first evaluation, Compute.cu:121-137) and returns the evaluation count. */
void ora_sk_accumulate(ora_sys* s, int type);
int  ora_sk_average(const ora_sys* s, int type, double* out);

/* integrator.cu:24 Integrate_2 for every integrator at the current step */
void ora_integrate(ora_sys* s);
/* main.cu:106-201 loop body (forces, integrate), nsteps times; step counter advances */
void ora_run(ora_sys* s, int nsteps);

/* nlist.cu:13-228 cell list + neighbour list for a group; returns 0 on success.
   counts[ns] (0 for sites outside the group) and neighbours sorted ascending in
   neigh[ns*cell_max]; cell_of[ns] = cell index. */
int  ora_nlist(ora_sys* s, double rc, int group, int* cell_of, int* counts, int* neigh,
               int neigh_cap, int* cell_max_out, int* ncells_out);

#ifdef __cplusplus
}
#endif
#endif
