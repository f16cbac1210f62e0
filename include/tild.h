/*
 * tild.h -- C ABI of the MI355X TILD step engine (libtild.so).
 *
 * The reference (rar-ensemble/gpu-tild) is a monolithic CUDA executable whose
 * per-step path is driven through process globals and the PairStyle / Group /
 * Integrator / NList classes.  This header replaces that path with an explicit,
 * re-entrant boundary: plain pointers and sizes, status codes instead of exit(),
 * no torch or HIP types in any signature.  Each entry point names the reference
 * interface it stands in for.
 *
 * Threading: one host thread per engine; all device work of an engine runs on its
 * own HIP stream.  Engines are independent (no process globals).
 */
#ifndef TILD_H
#define TILD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference calls die() -> exit(1) on every error
   (die.cu:3-15; e.g. read_input.cu:517-519, forces.cu:61-66). */
typedef enum {
    TILD_OK = 0,
    TILD_ERR_PARSE = 1,       /* input deck or data file malformed */
    TILD_ERR_CONFIG = 2,      /* deck is well formed but not runnable (missing grid, bad group...) */
    TILD_ERR_OOM = 3,         /* device or host allocation failed */
    TILD_ERR_DEVICE = 4,      /* HIP runtime error */
    TILD_ERR_NAN = 5,         /* NaN energy at a property evaluation (main.cu:250-291) */
    TILD_ERR_ARG = 6,         /* bad argument (size mismatch, null pointer, index range) */
    TILD_ERR_BUDGET = 7,      /* pool admission: job can never fit the HBM budget */
    TILD_ERR_UNSUPPORTED = 8  /* feature of the reference outside this engine's scope */
} tild_status;

#define TILD_MAX_STYLES 64

typedef struct tild_system tild_system;   /* parsed deck, host only */
typedef struct tild_engine tild_engine;   /* one device-resident replica */

/* ---------------------------------------------------------------- host only --- */

/* Parse an input deck with the reference grammar and read its data file.
   Replaces read_input() (read_input.cu:12-533) with read_config / read_charge_config
   (config_utils.cu:5-364); relative file names resolve against base_dir.
   No GPU is touched.  On failure *out is NULL and tild_last_error() says why. */
tild_status tild_system_load(const char* input_text, const char* base_dir, tild_system** out);
void        tild_system_free(tild_system* sys);

/* Message of the last failed call on this thread (never NULL). */
const char* tild_last_error(void);

/* Scalar facts about a parsed deck.  Keys: "ns", "dim", "ntypes", "M", "Nx0", "Nx1",
   "Nx2", "pmeorder", "n_styles", "n_gaussians", "n_erfs", "n_gaussian_erfs",
   "n_fieldphases", "do_charges", "n_bonds_total", "nbond_types", "max_steps", "log_freq", "traj_freq",
   "grid_freq", "binary_freq", "n_groups", "n_integrators", "seed", "L0", "L1", "L2",
   "dx0", "dx1", "dx2", "gvol", "V", "delt", "noise_mag", "bjerrum", "smear".  Mirrors
   the globals of globals.h:47-80 after initialize() (initialize.cu:32-48). */
tild_status tild_system_info(const tild_system* sys, const char* key, double* value);
/* Text-valued deck settings.  Keys: "dump_name" (traj_name, read_input.cu:504-512;
   default traj.lammpstrj, :561).  Copies a NUL-terminated string of at most cap bytes. */
tild_status tild_system_info_text(const tild_system* sys, const char* key, char* buf, size_t cap);

/* Per-site arrays of a parsed deck; n is the element count of the buffer(s). */
tild_status tild_system_get_positions(const tild_system* sys, float* x, size_t n);       /* ns*Dim */
tild_status tild_system_get_types(const tild_system* sys, int* tp, size_t n);           /* ns, 0-based */
tild_status tild_system_get_molecules(const tild_system* sys, int* mol, size_t n);      /* ns, 0-based */
tild_status tild_system_get_charges(const tild_system* sys, float* q, size_t n);        /* ns */
/* Bond tables exactly as config_utils.cu:96-125 builds them: n_bonds[ns],
   bonded_to[ns*3], bond_type[ns*3] (MAX_BONDS = 3, globals.h:26). */
tild_status tild_system_get_bonds(const tild_system* sys, int* n_bonds, int* bonded_to,
                                  int* bond_type, size_t ns);
/* Site list of group g (make_group_all / make_group_type, group.cu:30-60). */
/* Compute `i` of the deck (`compute avg_sk T [freq F] [wait W]`, Compute.cu:20-75):
   0-based type, frequency (default 100) and wait (default 0) in steps. */
tild_status tild_system_compute(const tild_system* sys, int i, int* type, int* freq, int* wait);
tild_status tild_system_group_sites(const tild_system* sys, int group, int* sites,
                                    size_t cap, int* nsites);

/* ---------------------------------------------------------------- engine ------ */

typedef struct {
    int device;          /* HIP device ordinal */
    int use_seed;        /* nonzero: override the deck's rand_seed (read_input.cu:450-454) */
    uint32_t seed;
    int spread_mode;     /* 0 auto, 1 global fixed-point atomics, 2 LDS-tiled sorted bins */
    int tile_target;     /* tiled spread: aim for about this many tiles (0 = 512) */
    int tile_by, tile_bz;/* tiled spread: force a By x Bz tile shape (0 = free) */
    int reorder_every;   /* spatial reorder of the particle arrays: 0 automatic (every 60
                            steps once the gather records outgrow L2, or for tiled
                            replicas of 4,096 beads or more), n > 0 every n steps,
                            < 0 never.  Results do not depend on it. */
    int fft_mid_elems;   /* FFT block geometry, 0 = default: complex values per block of */
    int fft_last_lines;  /* the y pass, columns per block of the fused z pass, rows per  */
    int fft_c2r_rows;    /* block of the inverse x pass.  Results do not depend on them. */
    int fft_plane;       /* 3D forward plane pass (x R2C + y FFT per plane, for tiled
                            replicas with even Ny and even tile height): >= 0 used where
                            it applies (default), < 0 separate passes.  Results do not
                            depend on it. */
    int fft_last_dirs;   /* inverse outputs per batch of the fused z pass, 1 or 2 (0 = 1) */
    int tile_slack;      /* tiled spread: spare slots per bucket region beyond the count at
                            the last sizing (0 = count/4 + 16; < 0 none, so that drifting
                            beads take the overflow path).  Results do not depend on it. */
    int autotune;        /* nonzero: each replica's block geometry (tile count, deposit
                            block size, plane pass, FFT line batching) comes from the
                            tuning cache for its class, or from a measured sweep on a
                            scratch copy of the replica that is then cached (file
                            $TILD_TUNE_CACHE, else ~/.cache/tild/tune.tsv); the explicit
                            geometry settings above are ignored.  Results do not depend
                            on it. */
    int deposit_threads; /* tiled spread: deposit block size, 256 or 512 (0 = by the
                            beads per tile).  Results do not depend on it. */
    int fft_staged;      /* nonzero: the forward plane pass, the fused z pass and the y
                            inverse stage every value through LDS even where their
                            register-resident forms apply (two-stage plans 16x10, 10x8,
                            8x8, 8x5 of the axis; the plane pass square planes, the z
                            pass at most four types, two on a 160-point axis).  The forms agree to rounding (about 1e-7
                            relative), not bit for bit: the compiler fuses multiplies into
                            additions where one keeps in registers what the other stages
                            through LDS. */
} tild_opts;

/* Properties of calc_properties() (calc_properties.cu:20-116): FP64 reductions. */
typedef struct {
    double Upe;                  /* sum of style energies + Uelec + Ubond */
    double Ubond;                /* host_bonds (bonds.cu:3-52) at current positions */
    double Uelec;                /* calc_electrostatic_energy (calc_properties.cu:106-116) */
    double Ptens[6];             /* xx yy zz xy xz yz (3D) | xx yy xy (2D) */
    int n_styles;                /* energies below: gaussians, erfs, gaussian_erfs in deck order */
    double U_style[TILD_MAX_STYLES];
    int step;
    int nan_flag;                /* a style energy is NaN (main.cu:250-288) */
} tild_props;

/* Allocate device state for one replica, build k-space tables and upload the
   particles.  Replaces initialize() (initialize.cu:20-85) + allocate_device_memory
   (:93-222) + allocate_device_potentials (:224-244) + send_box_params_to_device
   (device_comm_utils.cu:90-140). */
tild_status tild_engine_create(const tild_system* sys, const tild_opts* opts, tild_engine** out);
void        tild_engine_destroy(tild_engine* e);
const char* tild_engine_last_error(const tild_engine* e);

/* forces() at the current positions (forces.cu:42-196): spread, k-space pair
   forces and charges, gather, bonds.  Leaves the particle forces readable. */
tild_status tild_engine_forces(tild_engine* e);

/* n iterations of the main-loop body (main.cu:106-117): step += 1, forces(x_n),
   Integrate_2 for every integrator (integrator.cu:24-39) -> x_{n+1}.  Small replicas
   (beads + grid nodes up to 32,768) replay runs of 60 steps from a captured hipGraph;
   larger ones, whose kernels outlast their launches, launch eagerly. */
tild_status tild_engine_step(tild_engine* e, int n);

/* calc_properties() after the last forces()/step: non-bonded energies and virial
   from the densities of that force evaluation, bonds at the current positions
   (the reference's mixed-time convention, calc_properties.cu:49-59). */
tild_status tild_engine_properties(tild_engine* e, tild_props* out);

/* Host <-> device copies (cuda_collect_x / send_3n_to_device,
   device_comm_utils.cu:5-31; cuda_collect_f :82-87; cuda_collect_rho :33-44).
   set_positions also resets the GJF previous position to x (device_comm_utils.cu:15-16). */
tild_status tild_engine_get_positions(tild_engine* e, float* x, size_t n);       /* ns*Dim */
tild_status tild_engine_set_positions(tild_engine* e, const float* x, size_t n); /* ns*Dim */
tild_status tild_engine_get_forces(tild_engine* e, float* f, size_t n);          /* ns*Dim */
/* Density of type t on the grid (x fastest); t == ntypes returns the charge density. */
tild_status tild_engine_get_density(tild_engine* e, int type, float* rho, size_t n);
/* Gathered field of type t (t == ntypes: electric field), direction d: the grid
   quantity the gather interpolates, i.e. all_f{x,y,z}[t]/rho[t] of
   device_grid_utils.cu:424-452, or E_d of pair_style_charges.cu:71-79. */
tild_status tild_engine_get_grid_field(tild_engine* e, int type, int dir, float* g, size_t n);
/* First stencil node of every coordinate, ns*Dim ints: the grid bin the reference
   computes from FP32 x / dx (device_grid_utils.cu:136-146, :153-226), wrapped. */
tild_status tild_engine_get_bins(tild_engine* e, int* base, size_t n);
/* nsteps of tild_engine_step, each kernel launched with a HIP event pair bound to its
   own start and stop (hipExtLaunchKernel) on the engine's stream; phase_ms[k] receives
   the mean per-step duration of phase k (nphase >= 9), one kernel each: 0 bucket-claim
   pass (only where a run starts without claims; global-atomic spread on that path),
   1 tile deposit (accumulator conversion + x R2C on the atomic path), 2 forward plane
   pass (y FFT without it), 3 fused last-axis k-space pass, 4 inverse y-axis pass,
   5 x-axis C2R, 6 gather+bonds+integrate+next-step claims, 7 step counter; 8 the
   spatial reorders (an event pair around them), amortised over the steps. */
tild_status tild_engine_profile(tild_engine* e, int nsteps, double* phase_ms, int nphase);
tild_status tild_engine_get_step(tild_engine* e, int* step);
/* Exact restart.  The reference's resume frame holds positions only
   (config_utils.cu:405-454, written at main.cu:205-207), so a GJF run does not resume
   exactly there.  Here the thermal noise is keyed by (seed, site, step), so the step
   counter and the GJF displacement x - xo of every site (ns*Dim floats, site order;
   unused for EM sites) restore a run bit for bit together with the positions
   (tild_engine_get/set_positions) and the same seed.  set_state after set_positions. */
tild_status tild_engine_get_state(tild_engine* e, int* step, float* xo_disp, size_t n);
tild_status tild_engine_set_state(tild_engine* e, int step, const float* xo_disp, size_t n);

/* Cell-list neighbour lists of a site group at the current positions
   (NList::Initialize + AssignCells + BuildNList, nlist.cu:13-228; the reference builds
   it on the device for its group name).  sites/nsites: the group's site ids (NULL = all
   sites; tild_system_group_sites lists a deck group).  Outputs, by site id: cell_of[ns]
   (x-fastest cell index, -1 outside the group), counts[ns] and neigh[ns*cap], each list
   sorted ascending and including the site itself.  cap >= *cell_max (the reference's
   CELL_MAX, nlist.cu:49-51); neigh == NULL only returns *cell_max and *ncells.  Cells hold
   CELL_MAX sites and lists up to cap: more sites in a cell than CELL_MAX, or a list longer
   than cap, is TILD_ERR_CONFIG.  The reference's lists hold CELL_MAX and it writes past
   them, so a cap above CELL_MAX returns correct lists where the reference overruns. */
tild_status tild_engine_nlist(tild_engine* e, float rc, const int* sites, int nsites,
                              int* cell_of, int* counts, int* neigh, size_t cap,
                              int* cell_max, int* ncells);
tild_status tild_engine_synchronize(tild_engine* e);
/* Block geometry in use (no counterpart in the reference: its launch shapes are fixed,
   M_Grid/M_Block and ns_Grid/ns_Block, initialize.cu:62-69).  tuned = 1 when the
   autotuner chose it (from_cache = 1 when read from the tuning cache rather than
   measured), with the tile target it picked and its measured kernel time per step. */
typedef struct {
    int tile_target;     /* tuner's tile target (0 when not tuned) */
    int tiles;           /* tiles of the tiled spread (0: global-atomic spread) */
    int deposit_threads; /* requested deposit block size (0: by beads per tile) */
    int plane;           /* 1: forward plane pass */
    int last_lines, last_dirs, mid_elems, c2r_rows;   /* FFT line batching */
    int last_reg;        /* 1: register-resident fused z pass */
    int grad_reg;        /* 1: register-resident y inverse */
    int plane_reg;       /* 1: register-resident forward plane pass */
    int plane_inv;       /* 1: y inverse and x C2R as one register-resident plane pass */
    int tuned, from_cache;
    double step_ms;      /* kernel time per step measured by the tuner (0 if not tuned) */
} tild_geometry;
tild_status tild_engine_geometry(tild_engine* e, tild_geometry* out);
/* Compute::doCompute for avg_sk (Compute.cu:83-117): adds rho_t(k) rho_t(k) / M -- the
   complex square of the unnormalised transform, as d_multiplyComplex forms it, not
   |rho_t(k)|^2 -- of the densities of the last force evaluation (after
   tild_engine_step, those of the positions the last step started from, which is what
   the reference's computes see) to the running sum of compute slot `slot` (0..63, one
   per Compute object; a slot keeps the type of its first call), in FP64 on the device. */
tild_status tild_engine_sk_accumulate(tild_engine* e, int slot, int type);
/* Compute::writeResults (Compute.cu:121-137): slot's running average for all M
   wavevectors, x fastest, interleaved (re, im) in sk[2*M]; zeros before the first
   evaluation.  *count (may be NULL) receives the number of evaluations. */
tild_status tild_engine_sk_average(tild_engine* e, int slot, float* sk, size_t n, int* count);

/* ---------------------------------------------------------------- replica batches */

/* Many independent replicas (uneven boxes, grids, chain counts, pair styles) resident
   on one device.  tild_batch_step advances a set of them together: every per-step
   kernel is launched once for the whole set (blocks map to (replica, tile) work
   items), and with use_graph sets of up to 32,768 beads + grid nodes replay runs of
   60 steps captured once per set in a hipGraph (larger sets launch eagerly).  The reference has no counterpart (one system per process,
   main.cu:36-227); each replica follows the same per-step semantics as tild_engine. */
typedef struct tild_batch tild_batch;
tild_status tild_batch_create(const tild_opts* opts, tild_batch** out);
void        tild_batch_destroy(tild_batch* b);
const char* tild_batch_last_error(const tild_batch* b);
tild_status tild_batch_add(tild_batch* b, const tild_system* sys, uint32_t seed, int* replica);
tild_status tild_batch_remove(tild_batch* b, int replica);
/* replicas == NULL/nrep == 0: every replica of the batch */
tild_status tild_batch_step(tild_batch* b, const int* replicas, int nrep, int nsteps, int use_graph);
tild_status tild_batch_properties(tild_batch* b, int replica, tild_props* out);
tild_status tild_batch_get_positions(tild_batch* b, int replica, float* x, size_t n);
tild_status tild_batch_set_positions(tild_batch* b, int replica, const float* x, size_t n);
tild_status tild_batch_get_density(tild_batch* b, int replica, int type, float* rho, size_t n);
tild_status tild_batch_replica_step(tild_batch* b, int replica, int* step);
tild_status tild_batch_geometry(tild_batch* b, int replica, tild_geometry* out);
/* as tild_engine_profile, over a replica set */
tild_status tild_batch_profile(tild_batch* b, const int* replicas, int nrep, int nsteps,
                               double* phase_ms, int nphase);

/* ---------------------------------------------------------------- job server */

/* Continuous batching of independent simulation jobs on one device.  A job is a deck
   (+ base directory for its data file), a Philox seed and a step count; the server
   parses it on the submitting thread, admits queued jobs in arrival order while their
   device memory fits the HBM budget, advances all running jobs together in chunks of
   chunk_steps (each chunk one tild_batch_step of the running set, cut short at the
   nearest job end), and retires finished jobs with their final properties and
   positions, freeing their memory for the queue.  A job's trajectory does not depend
   on what runs beside it.  The reference has no counterpart (one system per process,
   main.cu:36-227). */
typedef struct tild_server tild_server;
typedef struct {
    int device;
    int chunk_steps;      /* steps between admission points; 0 = 100 */
    int max_replicas;     /* running-job cap; 0 = budget only */
    int use_graph;        /* replay captured runs of steps for small running sets */
    size_t hbm_budget;    /* bytes of device memory for jobs; 0 = 90% of free memory */
    int step_work;        /* 0: every running job advances in every chunk.  > 0: each chunk
                             advances the running jobs with the least remaining work
                             (steps left x (beads + grid nodes)), taken in that order until
                             their work per step reaches step_work thousand beads + nodes;
                             the rest stay resident and wait.  Short jobs then finish
                             sooner while the device stays as busy. */
    int reserved[7];
} tild_server_opts;

enum { TILD_JOB_QUEUED = 0, TILD_JOB_RUNNING = 1, TILD_JOB_DONE = 2, TILD_JOB_FAILED = 3 };

typedef struct {
    int submitted, queued, running, done, failed;
    size_t bytes_in_use, peak_bytes, budget;
    double busy_s;        /* wall time spent stepping */
    double elapsed_s;     /* since tild_server_create */
    double bead_steps;    /* sum over chunks of (beads running) x (steps) */
    double p50_s, p99_s;  /* time from submission to result of finished jobs */
    double admit_s;       /* worker time setting up admitted jobs (tables, uploads) */
    double retire_s;      /* worker time copying out and releasing finished jobs */
    double wait50_s, wait99_s;  /* finished jobs: submission to admission (queueing) */
    double svc50_s, svc99_s;    /* finished jobs: admission to result (service) */
} tild_server_stats;

tild_status tild_server_create(const tild_server_opts* opts, tild_server** out);
/* Stops the worker after its current chunk; unfinished jobs are dropped. */
void        tild_server_destroy(tild_server* s);
/* TILD_ERR_BUDGET: the job alone exceeds the budget.  Parse errors as tild_system_load. */
tild_status tild_server_submit(tild_server* s, const char* input_text, const char* base_dir,
                               uint32_t seed, int nsteps, int* job);
tild_status tild_server_poll(tild_server* s, int job, int* state);
/* timeout_s < 0: no limit.  A failed job returns its error. */
tild_status tild_server_wait(tild_server* s, int job, double timeout_s);
tild_status tild_server_drain(tild_server* s, double timeout_s);
/* Final properties, positions (ns*Dim, may be NULL) and time-to-result of a done job. */
tild_status tild_server_result(tild_server* s, int job, tild_props* props, float* x, size_t n,
                               double* latency_s);
tild_status tild_server_stats_get(tild_server* s, tild_server_stats* out);

/* ---------------------------------------------------------------- building blocks */

/* Batched single-precision real<->complex 2D/3D transforms on device memory with the
   engine's Stockham kernels (the reference uses cuFFT C2C plans, initialize.cu:217-220).
   Layout: x fastest; complex output has Nx[0]/2+1 columns.  Forward is e^{-ikr},
   inverse unnormalised.  `stream` is a hipStream_t (NULL = default stream). */
tild_status tild_fft_r2c(int dim, const int* Nx, int batch, const float* d_in,
                         float* d_out, void* stream);
tild_status tild_fft_c2r(int dim, const int* Nx, int batch, const float* d_in,
                         float* d_out, void* stream);
/* tild_fft_c2r without any device allocation, for callers that cannot allocate (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
   counterpart in the reference.  Its scratch (descriptor, twiddles, work map and the
   work copy of the batch*(Nx[0]/2+1)*Nx[1]*Nx[2] input spectra) is a fixed 32 MiB
   device array per GPU, held until the process ends; transforms that need more return
   TILD_ERR_ARG.  Calls that overlap on one GPU share that array and race.  Same
   results as tild_fft_c2r, bit for bit. */
tild_status tild_fft_c2r_noalloc(int dim, const int* Nx, int batch, const float* d_in,
                                 float* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
