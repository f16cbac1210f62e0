// engine.hpp -- device-resident TILD replicas and their step pipeline.
//
// A Batch owns a HIP stream, an array of replica descriptors in device memory and
// the per-phase work maps.  A single engine (C ABI tild_engine) is a Batch with one
// replica; batched sweeps advance every active replica with the same ~8 launches
// per step.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tild.h"
#include "device.hpp"
#include "launch.hpp"
#include "system.hpp"

namespace tild {

constexpr int kXcds = 8;         // accelerator complex dies of an MI355X
constexpr int kGraphSteps = 60;  // integrating steps per graph launch: one spatial-reorder interval
constexpr long long kGraphMaxWork = 32768;   // beads + grid nodes of a set above which steps launch eagerly
constexpr int kProfPhases = 9;   // claim, deposit, forward plane, last_fused, y_inv, x_c2r, particles, advance, reorder

// Work-map phases.  *A: replicas on the global-atomic spread; T*: replicas on the
// LDS-tiled spread; XP: x-axis R2C of stored densities (properties).
// MIDY_H/MIDY_G: y-axis passes of the step (3D replicas); PY: forward y-axis pass
// of stored spectra for properties (every replica: it is the last axis in 2D).
// PLANE: y-inverse + x C2R of 3D replicas whose planes fit one block's LDS (they
// skip MIDY_G and C2R); PLANEF: x R2C + y FFT of tiled 3D replicas with even Ny
// (their deposit writes densities only and they skip MIDY_H).
enum Phase { PH_PART = 0, PH_SPREADA, PH_XA, PH_XP, PH_SCAN, PH_TILE, PH_MIDY_H, PH_MIDY_G,
             PH_MIDZ_H, PH_PY, PH_LAST, PH_C2R, PH_PLANEF, PH_TCOUNT, PH_KPROPS, PH_RTILE, PH_LASTR, PH_MIDY_GR, PH_PLANER, PH_LASTR4, PH_PLANEI, PH_COUNT };

struct PhaseMap {
    std::vector<WorkItem> host;
    WorkItem* dev = nullptr;
    size_t lds = 0;
    int n = 0;
    int threads = 0;   // block size where a pass has more than one (0: its default)
    int cap = 0;       // work items the device buffer holds
};

// A ramped pair style: its index and ramp schedule (RepDev::ramp_*).
struct RampInfo {
    int style;
    double A0, c0, c1;
};

constexpr int kMaxSkSlots = 64;   // avg_sk computes per replica

// Block geometry of a replica: settings that change only how the work of a step is cut
// into blocks, never a result -- the tile count of the tiled spread (tile shape), the
// deposit block size, the forward plane pass, and the line batching of the FFT passes
// (0 = the built-in default for each).
struct Geometry {
    int tile_target = 512, dep_threads = 0, plane = 0, last_lines = 0, last_dirs = 0, mid_elems = 0, c2r_rows = 0;
};
// What the autotuner picked for a replica: the geometry, its measured kernel time per
// step, and whether it came from the tuning cache.
struct TuneResult {
    Geometry g;
    double step_ms = 0.0;
    bool from_cache = false, tuned = false;
};

struct Replica {
    RepDev h{};                       // host copy of the device descriptor
    std::vector<void*> allocs;        // owned device allocations
    int ns = 0, dim = 3, ntypes = 0, nf = 0, M = 0, Mh = 0, nP = 6;
    int nstyles = 0, npairs = 0, nq = 0;
    float L[3] = {1, 1, 1};
    double V = 1.0;
    int nblk_part = 0;
    size_t bytes = 0;                 // device memory held
    bool active = true;
    int step = 0;                     // host mirror of h.step
    PhaseMap maps[PH_COUNT];          // single-replica work maps
    double* d_results = nullptr;      // nq + 7 results of the last property evaluation
    double* d_bpart = nullptr;        // 7 * nblk_part bond partials
    float* d_fout = nullptr;
    // periodic spatial reorder of the particle arrays (0 = never)
    int reorder_every = 0, since_sort = 0, sort_bits = 0;
    bool permuted = false;            // slots no longer in site order (getters un-permute)
    // LDS-tiled replicas: the current step's records are in the bucket regions (made
    // by the last integrating step's particle kernel, or by a bucket-claim pass)
    bool claimed = false;
    std::vector<RampInfo> ramps;
    ReorderBufs rb;
    // avg_sk accumulators, one per compute slot: sum of rho(k) rho(k) / M over the half
    // spectrum (FP64 re, im), the type it follows and the number of evaluations summed
    std::vector<double*> sk_buf;
    std::vector<int> sk_type, sk_count;
    int dep_threads = 0;              // deposit block size asked for (0: by bead density)
    TuneResult tune;                  // the autotuner's pick (tune.tuned false: not tuned)
};

class Batch {
public:
    explicit Batch(int device) : device_(device) {}
    // spread mode for replicas added from now on: 0 auto, 1 global atomics, 2 tiled;
    // target block count and forced tile shape (0 = free) of the tiled mode.
    void set_spread(int mode, int target_tiles = 512, int by = 0, int bz = 0, int slack = 0) {
        spread_mode_ = mode; tile_target_ = target_tiles; tile_by_ = by; tile_bz_ = bz; tile_slack_ = slack;
    }
    // spatial reorder interval of replicas added from now on: > 0 every n steps,
    // 0 never, < 0 automatic (every 20 steps when the gather records outgrow L2)
    void set_reorder(int every) { reorder_every_ = every; }
    // FFT block geometry of replicas added from now on (0 = default): complex values
    // per middle-axis block, lines per fused last-axis block, rows per inverse-x block
    struct FftTuning { int mid_elems = 0, last_lines = 0, c2r_rows = 0, plane = 0, last_dirs = 0, staged = 0; };
    void set_fft_tuning(const FftTuning& t) { tune_ = t; }
    // Deposit block size of replicas added from now on (0: by bead density, else 256 or 512).
    void set_deposit_threads(int n) { dep_threads_ = n; }
    // Autotuning of replicas added from now on: each replica's block geometry is taken
    // from the tuning cache for its class (device, grid, field count, order, bead-count
    // octave), or measured by a sweep over the candidates and stored there.  The
    // explicit tile and FFT settings above are then ignored.
    void set_autotune(bool on) { autotune_ = on; }
    ~Batch();
    int init(std::string& err);

    // Build a replica from a parsed deck; returns its id (>= 0) or a negative tild_status.
    int add(const System& S, uint32_t seed, std::string& err);
    void remove(int id);
    int size() const { return (int)reps_.size(); }
    Replica* rep(int id) { return (id >= 0 && id < (int)reps_.size() && reps_[id]) ? reps_[id].get() : nullptr; }
    static size_t estimate_bytes(const System& S);

    // forces() of one replica at its current positions (no integration).
    int forces(int id, std::string& err);
    // n steps of every replica in `ids` (empty = all active), one launch per phase;
    // with use_graph the step is captured once per replica set and replayed.
    int step(const std::vector<int>& ids, int n, std::string& err, bool use_graph = false);
    // n steps with an event pair around every launch; ms[k] = mean ms of phase k.
    int profile(const std::vector<int>& ids, int n, double* ms, int nph, std::string& err);
    // calc_properties() of one replica.
    int properties(int id, tild_props* out, std::string& err);
    // Compute::doCompute for avg_sk (Compute.cu:83-117): add rho_t(k) rho_t(k) / M of
    // the densities of the last force evaluation to compute slot `slot`'s running sum
    // (each compute keeps its own sum, as each Compute object does)
    int sk_accumulate(int id, int slot, int type, std::string& err);
    // the running average over all M wavevectors (x fastest, interleaved re, im; zero
    // before the first evaluation, as Compute::writeResults) and the evaluation count
    int sk_average(int id, int slot, float* out, size_t n, int* count, std::string& err);

    int get_positions(int id, float* x, size_t n, std::string& err);
    int set_positions(int id, const float* x, size_t n, std::string& err);
    // exact restart state: step counter and GJF displacement x - xo in site order
    int get_state(int id, int* step, float* xo, size_t n, std::string& err);
    int set_state(int id, int step, const float* xo, size_t n, std::string& err);
    int get_forces(int id, float* f, size_t n, std::string& err);
    int get_density(int id, int type, float* rho, size_t n, std::string& err);
    int get_grid_field(int id, int type, int dir, float* g, size_t n, std::string& err);
    int get_bins(int id, int* base, size_t n, std::string& err);
    // Neighbour lists of a site group at the current positions (sites == null: all).
    int nlist(int id, float rc, const int* sites, int nsites, int* cell_of, int* counts, int* neigh,
              size_t cap, int* cell_max, int* ncells, std::string& err);
    int sync(std::string& err);
    hipStream_t stream() const { return stream_; }

    // Work maps of an arbitrary replica set (cached by set).
    struct SetMaps {
        PhaseMap maps[PH_COUNT];
        int* d_ids = nullptr;
        int nids = 0;
        hipGraphExec_t graph_run = nullptr;   // kGraphSteps captured integrating steps
    };
    SetMaps* maps_for(const std::vector<int>& ids, std::string& err);

    // Step pipeline of a replica set with prepared maps (graph-capturable).  The
    // records of the first step must be claimed (ensure_claimed); integrating steps
    // leave the next step's records claimed.
    int enqueue_steps(SetMaps& m, int nsteps, bool integrate, bool write_forces, std::string& err);
    // Claims the current step's records of every replica of the set that has none
    // (a replica whose claims are stale is cleared first).
    // With t0/t1 the claim launch binds them to its start and stop; *launched says
    // whether it ran.
    int ensure_claimed(SetMaps& m, const std::vector<int>& ids, std::string& err,
                       hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr, bool* launched = nullptr);

private:
    // the sweep behind set_autotune (autotune.cpp)
    int tune_geometry(const System& S, uint32_t seed, TuneResult& out, std::string& err);
    // kernel ms per replica of a step of g; *passes: the FFT-pass forms g runs, set by the
    // first call (< 0 on entry), and any later g running other forms measures +inf
    double measure_geometry(const System& S, uint32_t seed, const Geometry& g, std::string& err, int* passes);
    int add_with(const System& S, uint32_t seed, const Geometry& g, std::string& err);
    int add_plain(const System& S, uint32_t seed, std::string& err);
    bool autotune_ = false, tuning_ = false;
    int dep_threads_ = 0;
    int device_;
    int spread_mode_ = 0, tile_target_ = 512, tile_by_ = 0, tile_bz_ = 0, tile_slack_ = 0, reorder_every_ = -1;
    FftTuning tune_;
    hipStream_t stream_ = nullptr;
    // add() builds a replica on its own stream, so that a caller (the job server) can
    // build new replicas while steps of the others run on stream_
    hipStream_t setup_stream_ = nullptr;
    std::vector<std::unique_ptr<Replica>> reps_;
    RepDev* d_reps_ = nullptr;
    int d_cap_ = 0;
    std::map<std::vector<int>, std::unique_ptr<SetMaps>> set_cache_;

    int upload_desc(int id, std::string& err);
    // reorder_only: only the phases of a set-wide spatial reorder, into device buffers
    // that are reused while large enough (the set of replicas due changes from one
    // reorder to the next, so these maps are not cached per set)
    int build_maps(const std::vector<int>& ids, PhaseMap* maps, std::string& err, bool reorder_only = false);
    PhaseMap reorder_maps_[PH_COUNT];
    int build_tables(Replica& R, const System& S, std::string& err);
    int forward_spectra(int id, std::string& err);   // rho -> H (properties)
    // spatial reorder of one replica; after_step: the last step's tile binning matches
    // the current slots and is reused
    int reorder(int id, std::string& err, bool after_step = false);
    // the after-step reorder of several tiled replicas at once (one launch per kernel)
    int reorder_set(const std::vector<int>& ids, std::string& err);
    void mirror_reorder(Replica& R);   // the host copy of the descriptor follows the swaps
    int run_steps(SetMaps& m, const std::vector<int>& ids, int n, bool use_graph, std::string& err);
    int capture_steps(SetMaps& m, int nsteps, hipGraphExec_t* out, std::string& err);
    int tile_regions(int id, bool bin_only, std::string& err);
    // drops the claimed records of one replica (positions or regions change): clears
    // the slot counters of its current step parity
    int unclaim(int id, std::string& err);
    static void drop_graphs(SetMaps& m);
    int fetch_pid(const Replica& R, std::vector<int>& pid, std::string& err);
    void free_maps(PhaseMap* maps);
};

// Plan of one FFT axis: radices and twiddle table (forward sign).
int make_axis_plan(int N, AxisPlan& pl, std::vector<float2>& tw, std::string& err);
// The same without allocating: tw has room for max(N, 1) values, *ntw receives the
// count (N - 1, or 1 for N = 1).  Returns nullptr, or why N has no plan.
const char* plan_axis(int N, AxisPlan& pl, float2* tw, int* ntw);
// plan of the given radices (product N) and its twiddles (at most N entries)
void fill_plan(int N, const int* rad, int nst, AxisPlan& pl, float2* tw, int* ntw);
// Message of the last failed C ABI call on this thread (tild_last_error).
void set_last_error(const std::string& msg);

// Tile shape of the sorted-bin deposition (By | Ny, Bz | Nz, halo below one tile,
// LDS and FFT-budget limits).  Returns false when no shape fits.  `target_tiles`
// steers the choice; want_by/want_bz > 0 force a shape (autotuner sweeps).
bool choose_tiles(int dim, const int* Nx, int P, int nf, int target_tiles, int want_by, int want_bz,
                  int& By, int& Bz);
size_t tile_lds_bytes(int Nx, int nf, int By, int Bz);

}  // namespace tild
