// autotune.cpp -- block-geometry autotuner of the batch (Batch::set_autotune).
//
// A replica's block geometry (tile count of the tiled spread, deposit block size, the
// forward plane pass, the line batching of the FFT passes) never changes a result: the
// densities are integer sums, and every FFT line is transformed by the same arithmetic
// whatever block holds it.  So the fastest geometry can be picked by measurement.  For
// each replica class -- device, grid, field count, stencil order, bead-count half-octave
// and spread mode -- the tuner measures the kernel time of a step of scratch copies of
// the deck (HIP events bound to every launch; small replicas as a set of copies, the
// way a sweep runs them) under the built-in default, then sweeps one setting at a time
// over its candidates, keeping a change only when it is at least 2 % faster.  Results are kept for the process and appended to a tab-separated cache
// file ($TILD_TUNE_CACHE, else $XDG_CACHE_HOME/tild/tune.tsv or ~/.cache/tild/tune.tsv),
// which later processes read instead of sweeping again.
#include <sys/stat.h>

#include <cerrno>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace tild {

namespace {

std::mutex g_mu;
std::map<std::string, TuneResult> g_cache;
bool g_loaded = false;

std::string cache_path() {
    if (const char* p = std::getenv("TILD_TUNE_CACHE")) return p;
    if (const char* x = std::getenv("XDG_CACHE_HOME")) if (*x) return std::string(x) + "/tild/tune.tsv";
    if (const char* h = std::getenv("HOME")) if (*h) return std::string(h) + "/.cache/tild/tune.tsv";
    return "";
}

// mkdir -p of the directory part of `path` (failures surface when the file is opened)
void make_parent_dirs(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
}

// one line: key \t tile_target dep_threads plane last_lines last_dirs mid_elems c2r_rows \t ms
void load_cache_locked() {
    if (g_loaded) return;
    g_loaded = true;
    const std::string path = cache_path();
    if (path.empty()) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t a = line.find('\t'), b = a == std::string::npos ? a : line.find('\t', a + 1);
        if (b == std::string::npos) continue;
        TuneResult r;
        std::istringstream is(line.substr(a + 1, b - a - 1));
        Geometry& g = r.g;
        if (!(is >> g.tile_target >> g.dep_threads >> g.plane >> g.last_lines >> g.last_dirs >> g.mid_elems >> g.c2r_rows))
            continue;
        r.step_ms = std::atof(line.c_str() + b + 1);
        r.tuned = true;
        g_cache[line.substr(0, a)] = r;   // later lines win
    }
}

void store_locked(const std::string& key, const TuneResult& r) {
    g_cache[key] = r;
    const std::string path = cache_path();
    if (path.empty()) return;
    make_parent_dirs(path);
    if (FILE* f = std::fopen(path.c_str(), "a")) {
        const Geometry& g = r.g;
        std::fprintf(f, "%s\t%d %d %d %d %d %d %d\t%.6f\n", key.c_str(), g.tile_target, g.dep_threads, g.plane,
                     g.last_lines, g.last_dirs, g.mid_elems, g.c2r_rows, r.step_ms);
        std::fclose(f);
    }
}

std::string class_key(int device, const System& S, int spread_mode) {
    hipDeviceProp_t p{};
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) { arch = p.gcnArchName; cus = p.multiProcessorCount; }
    const int oct2 = (int)std::floor(2.0 * std::log2((double)std::max(S.ns, 1)));
    std::ostringstream k;
    k << arch << "|cu" << cus << "|d" << S.Dim << "|" << S.Nx[0] << "x" << S.Nx[1] << "x" << (S.Dim == 3 ? S.Nx[2] : 1)
      << "|nf" << (S.ntypes + (S.charges.enabled ? 1 : 0)) << "|P" << S.pmeorder << "|ns2log" << oct2
      << "|spread" << spread_mode;
    return k.str();
}

}  // namespace

int Batch::add_with(const System& S, uint32_t seed, const Geometry& g, std::string& err) {
    const int tt = tile_target_, dt = dep_threads_;
    const FftTuning ft = tune_;
    tile_target_ = g.tile_target;
    dep_threads_ = g.dep_threads;
    tune_ = FftTuning{g.mid_elems, g.last_lines, g.c2r_rows, g.plane, g.last_dirs, 0};
    const int id = add_plain(S, seed, err);
    tile_target_ = tt;
    dep_threads_ = dt;
    tune_ = ft;
    return id;
}

// Kernel time of one step (ms per replica) of scratch copies of S with geometry g, or
// +inf when the geometry does not fit the deck.  Small replicas are measured as a set
// of copies (up to kTuneCopies, about kTuneBeads beads in all), as a sweep runs them:
// alone, a small grid leaves most of the device idle and favours the geometry with the
// most blocks, which is not the fastest once many replicas share each launch.
constexpr int kTuneCopies = 32;
constexpr double kTuneBeads = 4.0e5;

double Batch::measure_geometry(const System& S, uint32_t seed, const Geometry& g, std::string& err, int* passes) {
    constexpr double kInf = std::numeric_limits<double>::infinity();
    const int copies = std::max(1, std::min(kTuneCopies, (int)std::lround(kTuneBeads / std::max(S.ns, 1))));
    std::vector<int> ids;
    std::string e;
    tuning_ = true;
    for (int k = 0; k < copies; ++k) {
        const int id = add_with(S, seed + (uint32_t)k, g, e);
        if (id < 0) {
            tuning_ = false;
            for (int r : ids) remove(r);
            if (id == -TILD_ERR_DEVICE || id == -TILD_ERR_OOM) err = e;
            return kInf;
        }
        ids.push_back(id);
    }
    tuning_ = false;
    // which forms of the FFT passes the geometry runs: a geometry that switches between
    // forms that agree only to rounding is not a candidate (tuning changes no result)
    {
        const RepDev& h = reps_[ids[0]]->h;
        const int sig = h.plane_reg | (h.last_reg << 1) | (h.grad_reg << 2) | (h.plane_inv << 3);
        if (*passes < 0) *passes = sig;
        else if (sig != *passes) {
            for (int r : ids) remove(r);
            return kInf;
        }
    }
    double ms[kProfPhases] = {};
    double t = kInf;
    if (step(ids, 3, e, false) == 0 && profile(ids, 6, ms, kProfPhases, e) == 0) {
        t = 0.0;
        for (int k = 0; k < kProfPhases - 1; ++k) t += ms[k];   // kernels; the last phase is the reorders
        t /= copies;
    } else {
        err = e;
    }
    for (int r : ids) remove(r);
    return t;
}

int Batch::tune_geometry(const System& S, uint32_t seed, TuneResult& out, std::string& err) {
    const std::string key = class_key(device_, S, spread_mode_);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_cache_locked();
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            out = it->second;
            out.from_cache = true;
            return 0;
        }
    }
    Geometry best;
    int passes = -1;
    double best_t = measure_geometry(S, seed, best, err, &passes);
    if (!std::isfinite(best_t)) {
        if (err.empty()) err = "autotune: the default block geometry does not run this deck";
        return -TILD_ERR_DEVICE;
    }
    const bool tiled_ok = spread_mode_ != 1;
    // The register-resident passes are taken wherever the deck allows them: the fused
    // z pass has no batching of inverse directions to tune, and the forward plane pass
    // stays on where its register form applies (the separate passes agree with it to
    // rounding only, and tuning changes no result).
    bool reg_ok = false, plane_reg = false, plane_inv = false;
    {
        const int la = S.Dim - 1, nf = S.ntypes + (S.charges.enabled ? 1 : 0);
        AxisPlan lp, px, py;
        std::vector<float2> tw((size_t)std::max({S.Nx[0], S.Nx[1], S.Nx[la], 1}));
        int ntw = 0;
        reg_ok = plan_axis(S.Nx[la], lp, tw.data(), &ntw) == nullptr && last_reg_plan(lp, S.ntypes, nf);
        plane_reg = S.Dim == 3 && plan_axis(S.Nx[0], px, tw.data(), &ntw) == nullptr &&
                    plan_axis(S.Nx[1], py, tw.data(), &ntw) == nullptr && plane_reg_plan(px, py);
        // the inverse plane pass leaves no x C2R rows to batch
        plane_inv = plane_reg && S.Fields.empty() && plane_inv_plan(px, py) && px.rad[0] == py.rad[0] &&
                    px.rad[1] == py.rad[1];
    }
    struct Axis { int Geometry::*field; std::vector<int> cand; bool on; };
    const std::vector<Axis> axes = {
        {&Geometry::plane, {0, -1}, S.Dim == 3 && tiled_ok && !plane_reg},
        {&Geometry::tile_target, {128, 256, 512, 1024}, tiled_ok},
        {&Geometry::dep_threads, {256, 512}, tiled_ok},
        {&Geometry::last_lines, {0, 2, 4, 8, 16}, true},
        {&Geometry::last_dirs, {1, 2}, true},
        {&Geometry::mid_elems, {0, 1024}, true},
        {&Geometry::c2r_rows, {0, 1, 2, 4, 8, 16}, !plane_inv},
    };
    for (const Axis& a : axes) {
        if (!a.on) continue;
        if (a.field == &Geometry::last_dirs && reg_ok) continue;
        for (int c : a.cand) {
            if (best.*a.field == c) continue;
            Geometry g = best;
            g.*a.field = c;
            std::string e;
            const double t = measure_geometry(S, seed, g, e, &passes);
            if (!e.empty()) { err = e; return -TILD_ERR_DEVICE; }
            if (t < 0.98 * best_t) { best = g; best_t = t; }
        }
    }
    out = TuneResult{best, best_t, false, true};
    std::lock_guard<std::mutex> lk(g_mu);
    store_locked(key, out);
    return 0;
}

}  // namespace tild
