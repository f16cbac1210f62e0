// launch.hpp -- block geometry of every pass (shared by kernels and host launchers)
// and the host-side launcher declarations.
#pragma once
#include "device.hpp"

#include <hip/hip_ext.h>

namespace tild {

// Largest batch of complex values one FFT block transforms at once (two LDS
// ping-pong buffers of this size).
constexpr int kFftBudget = kFftElems * kFftThreads;        // 2048
constexpr size_t kMaxLds = 65536;

// Strided pass along y (axis 1) or z (axis 2) of nfb half-spectrum fields.
struct MidGeom {
    int N, lstride, batch, nouter, ostride, nl, nchunk;
};
__host__ __device__ inline MidGeom mid_geom(const RepDev& r, int axis, int nfb) {
    MidGeom g;
    const int Nz = r.dim == 3 ? r.Nz : 1;
    if (axis == 1) {
        g.N = r.Ny; g.lstride = r.Nxh; g.batch = r.Nxh;
        g.nouter = nfb * Nz; g.ostride = r.Ny * r.Nxh;
    } else {
        g.N = r.Nz; g.lstride = r.Ny * r.Nxh; g.batch = r.Ny * r.Nxh;
        g.nouter = nfb; g.ostride = r.Nz * r.Ny * r.Nxh;
    }
    int nl = (r.mid_elems > 0 ? r.mid_elems : kFftBudget) / g.N;
    if (nl < 1) nl = 1;
    if (nl > g.batch) nl = g.batch;
    g.nchunk = (g.batch + nl - 1) / nl;
    g.nl = (g.batch + g.nchunk - 1) / g.nchunk;     // balanced chunks
    return g;
}
// 3D y-inverse of the gradient spectra: y lines of (field, z) planes, three line sets
// per column in one batch
__host__ __device__ inline MidGeom mid_grad_geom(const RepDev& r) {
    MidGeom g;
    g.N = r.Ny; g.lstride = r.Nxh; g.batch = r.Nxh;
    g.nouter = r.nf * r.Nz; g.ostride = r.Ny * r.Nxh;
    int nl = (r.mid_elems > 0 ? r.mid_elems : kFftBudget) / (3 * g.N);
    if (nl < 1) nl = 1;
    if (nl > g.batch) nl = g.batch;
    g.nchunk = (g.batch + nl - 1) / nl;
    g.nl = (g.batch + g.nchunk - 1) / g.nchunk;     // balanced chunks
    return g;
}
// one FFT buffer of N x nlines complex values, and the ping-pong pair a pass needs
inline size_t fft_buf_bytes(int N, int nlines) { return (size_t)N * (size_t)(nlines | 1) * sizeof(float2); }
inline size_t lds_bytes(int N, int nlines) { return 2 * fft_buf_bytes(N, nlines); }
// fused last-axis pass: coupled fields (nf*nl lines, unpadded: only read and written
// element by element) + two buffers of nf*DB*nl lines (DB inverse outputs per batch)
// + the ramp weights of the step
inline size_t last_lds_bytes(int N, int nf, int DB, int nl) {
    return (size_t)N * nf * nl * sizeof(float2) + 2 * fft_buf_bytes(N, nf * DB * nl) + kMaxRamp * sizeof(float);
}

struct LaunchCtx {
    const RepDev* reps;
    const WorkItem* wmap;
    int n;              // blocks
    size_t lds;         // dynamic LDS bytes
    hipStream_t stream;
    // step profile: events that take the kernel's own start and stop times
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int threads = 0;    // block size where a pass has more than one (0: its default)
};

// Launch on c.stream; with c.t0 set, through hipExtLaunchKernel, whose events take
// the start and stop times of the kernel itself.
#define TILD_LAUNCH(c, kernel, grid, block, lds, ...)                                                   \
    do {                                                                                               \
        if ((c).t0) hipExtLaunchKernelGGL(kernel, grid, block, lds, (c).stream, (c).t0, (c).t1, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(kernel, grid, block, lds, (c).stream, __VA_ARGS__);                   \
    } while (0)

hipError_t launch_spread(const LaunchCtx& c, int koff);
hipError_t launch_x_r2c(const LaunchCtx& c, bool from_acc, int nrows_override, const float* in, float2* out);
hipError_t launch_mid(const LaunchCtx& c, int axis, int use_gk, int nfb, int inverse, float2* buf);
hipError_t launch_last_fused(const LaunchCtx& c, int store_spectrum, int koff);
hipError_t launch_x_c2r(const LaunchCtx& c, bool records, int nlines, const float2* in, float* out);
// claim: integrating steps of LDS-tiled replicas also claim the next step's records
hipError_t launch_particles(const LaunchCtx& c, bool integrate, int write_forces, int koff, bool claim);
hipError_t launch_advance(hipStream_t s, RepDev* reps, const int* which, int n, int nsteps,
                          hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
constexpr int kPlaneThreads = 512;                      // block of the plane forward pass
constexpr int kLastThreads = 256;                       // block of the fused last-axis pass
constexpr int kGradThreads = 256;                       // block of the y inverse pass
constexpr int kC2rThreads = 256;                        // block of the inverse x pass (records)
constexpr size_t kPlaneLds = 163840;                    // LDS one block may use (MI355X)
hipError_t launch_mid_grad(const LaunchCtx& c);
// Two-stage Stockham plans R1 x R2 the register-resident passes are compiled for (the
// benchmark workloads' axes: 160, 80, 64, 40, and the sweep's and server's 128, 50, 48,
// 32, 30, 25 and 24 points).
#define TILD_REG_PLANS(X) X(16, 10) X(10, 8) X(8, 8) X(8, 5) X(16, 8) X(10, 5) X(16, 3) X(8, 4) X(10, 3) X(5, 5) X(8, 3)
// Fused last-axis pass with register-resident columns (fft_reg.hip): last axes with one
// of those plans, at most four types (two on axes whose first radix is 16) plus charges.  Blocks of kLastRegThreads threads hold nl_last = kLastRegThreads / (2 R1)
// columns; results agree with k_last_fused to rounding.
constexpr int kLastRegThreads = 256;
// the variants compiled in (plan R1 x R2, types, charge field): those that fit 256
// registers per thread, so that every variant runs two waves per SIMD (160-point axes
// up to two types)
__host__ __device__ constexpr bool last_reg_variant(int R1, int R2, int TT, int CH) {
    return R1 == 16 ? TT <= 2 : TT <= 4;
}
// the kernel instantiation of a type count: up to three types (class 0), four (class 1,
// up to 256 registers, in a kernel of its own so that the others keep theirs)
__host__ __device__ constexpr int last_reg_class(int TT) { return TT >= 4 ? 1 : 0; }
bool last_reg_plan(const AxisPlan& pl, int ntypes, int nf);
inline int last_reg_lines(const AxisPlan& pl) { return kLastRegThreads / (2 * pl.rad[0]); }
// the two exchange buffers, both stages' twiddles (2N - R1 - R2) and the ramp weights
inline size_t last_reg_lds_bytes(const AxisPlan& pl, int nf) {
    const size_t N = (size_t)pl.N, nl = (size_t)last_reg_lines(pl);
    return (N * (3 * (size_t)nf * nl + 2) + 2 * N) * sizeof(float2) + kMaxRamp * sizeof(float);
}
hipError_t launch_last_reg(const LaunchCtx& c, int koff, int cls);
// The y inverse of the gradient spectra with register-resident columns (3D, two-stage
// y plans as above): blocks of kGradRegThreads threads, grad_reg_lines(plan) columns
// each, grad_reg_chunks per (field, z) plane.
constexpr int kGradRegThreads = 128;
bool grad_reg_plan(const AxisPlan& pl);
inline int grad_reg_lines(const AxisPlan& pl) { return kGradRegThreads / pl.rad[0]; }
inline int grad_reg_chunks(const AxisPlan& pl, int Nxh) { return (Nxh + grad_reg_lines(pl) - 1) / grad_reg_lines(pl); }
inline size_t grad_reg_lds_bytes(const AxisPlan& pl) {
    return ((size_t)pl.N * (3 * (size_t)grad_reg_lines(pl) + 1) + (size_t)pl.N) * sizeof(float2);
}
hipError_t launch_grad_reg(const LaunchCtx& c);
// The forward plane pass with register-resident columns (square planes, two-stage
// plans as above): one block of kPlaneRegThreads threads per (field, z) plane.
constexpr int kPlaneRegThreads = 512;
bool plane_reg_plan(const AxisPlan& px, const AxisPlan& py);
inline size_t plane_reg_lds_bytes(const AxisPlan& pl) {
    const size_t N = (size_t)pl.N, ldp = (size_t)((pl.N / 2) | 1), ldc = (size_t)((pl.N / 2 + 1) | 1);
    return (N * (ldp > ldc ? ldp : ldc) + N * ldp + 2 * N) * sizeof(float2);
}
hipError_t launch_plane_reg(const LaunchCtx& c);
hipError_t configure_plane_reg();
// Register-resident inverse plane pass (y inverse + x C2R into the gather records) of
// square planes with a two-stage plan: one block of kPlaneInvThreads threads per (field,
// z) plane, two blocks to a CU.
constexpr int kPlaneInvThreads = 512;
inline size_t plane_inv_lds_bytes(const AxisPlan& pl) {
    const size_t N = (size_t)pl.N, la = (size_t)((3 * (pl.N / 2 + 1)) | 1);
    return (N * la + (size_t)pl.rad[0] * (pl.rad[1] - 1)) * sizeof(float2);
}
bool plane_inv_plan(const AxisPlan& px, const AxisPlan& py);
hipError_t launch_plane_inv(const LaunchCtx& c);
hipError_t configure_plane_inv();
hipError_t configure_last_reg();
hipError_t launch_plane_r2c(const LaunchCtx& c);
// LDS of the forward plane pass: two buffers holding either the Ny/2 row pairs of
// length Nx or the Ny x Nxh half spectra
inline size_t plane_r2c_lds(const RepDev& h) {
    const size_t a = (size_t)h.Nx * (size_t)((h.Ny / 2) | 1), b = (size_t)h.Ny * (size_t)(h.Nxh | 1);
    return 2 * (a > b ? a : b) * sizeof(float2);
}
// One-time kernel attributes (dynamic LDS above 64 KiB for the plane pass).
hipError_t configure_kernels();
hipError_t launch_kprops(const LaunchCtx& c, int nq_max);
hipError_t launch_sk_accumulate(hipStream_t s, const RepDev* reps, int rep, int type, int Mh, double* acc);
hipError_t launch_bond_props(const LaunchCtx& c);
hipError_t launch_reduce_rows(hipStream_t s, const double* in, int nrows, int n, double* out);
constexpr int kCountBlock = 256;                        // beads per block of the bucket claims
hipError_t launch_tile_count(const LaunchCtx& c, int koff);
hipError_t launch_tile_scan(const LaunchCtx& c);
hipError_t launch_tile_claim(const LaunchCtx& c, int koff);
hipError_t launch_tile_reset(const LaunchCtx& c, int koff);
hipError_t launch_tile_regions(hipStream_t s, const RepDev* reps, int rep);
// c.threads: 512 for sets whose tiles hold many beads, else 256
hipError_t launch_tile_deposit(const LaunchCtx& c, int koff);
hipError_t launch_bins(hipStream_t s, const RepDev* reps, int rep, int ns, int* out);
hipError_t launch_virial_source(hipStream_t s, const RepDev* reps, int rep, int M, const float* fr, float* vr);

// Buffers of the periodic spatial reorder (one set per replica that reorders).
struct ReorderBufs {
    uint32_t* keys[2] = {nullptr, nullptr};   // ns row keys (unsorted, sorted)
    int* vals[2] = {nullptr, nullptr};        // ns slot ids (identity, new order)
    void* tmp = nullptr;                      // radix-sort scratch
    size_t tmp_bytes = 0;
};
size_t reorder_tmp_bytes(int ns, int bits);

// Neighbour-list build (device buffers, all indexed by site id except cell_slots).
struct NlistArgs {
    const unsigned char* member;   // ns group-membership flags
    int* cell_of;                  // ns cell of each member (-1 outside the group)
    int* counts;                   // ns list lengths
    int* neigh;                    // ns * cap sorted neighbour site ids
    int* cell_count;               // ncells
    int* cell_slots;               // ncells * cmax slots
    int* overflow;                 // 1: a cell exceeded cmax, 2: a list exceeded cap
    float Lc[3];
    int nc[3];
    float rc;
    int cmax, cap;
};
hipError_t launch_nlist(hipStream_t s, const RepDev* reps, int rep, int ns, const NlistArgs& a);
// Sort the replica's beads by grid row (stable), permute every per-bead array and
// swap the live and spare buffers in the device descriptor.
hipError_t launch_reorder(hipStream_t s, RepDev* reps, int rep, int ns, int bits, ReorderBufs& b);
// Tiled replicas: each tile's beads (from an exact binning, tkey and toff) sorted by
// site id, then the same permutation.
hipError_t launch_reorder_tiles(hipStream_t s, RepDev* reps, int rep, int ns, int ntiles, ReorderBufs& b);
// The same for a set of tiled replicas at once: `beads` maps 256-bead blocks, `tiles`
// one block per tile and `reps1` one block per replica (scratch in RepDev::rsrc etc.).
hipError_t launch_reorder_tiles_set(const LaunchCtx& beads, const LaunchCtx& tiles, const LaunchCtx& reps1);
// Bucket regions (k_tile_regions) of every replica of a set, one block each.
hipError_t launch_tile_regions_set(const LaunchCtx& reps1);

}  // namespace tild
