// engine.cpp -- host side of the TILD step engine: replica setup, k-space tables,
// work maps and the per-step launch sequence.
#include "engine.hpp"

#include <functional>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

namespace tild {

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            err = std::string(#call) + " failed: " + hipGetErrorString(e_);           \
            return -TILD_ERR_DEVICE;                                                  \
        }                                                                             \
    } while (0)

// ---------------------------------------------------------------- FFT plans
namespace {
// Depth-first search over non-increasing radix sequences (largest radices tried
// first): as few stages as possible (every stage is an LDS round trip ending with a
// barrier, and the passes are latency-bound), then the smallest largest radix
// (radices spread evenly).  Fixed arrays: N <= 2048 has at most 11 prime factors.
struct RadixSearch {
    static constexpr int kDepth = 16;
    int best[kDepth], nbest = -1, cur[kDepth], ncur = 0;
    void run(int n) {
        if (n == 1) {
            if (nbest < 0 || ncur < nbest || (ncur == nbest && cur[0] < best[0])) {
                for (int i = 0; i < ncur; ++i) best[i] = cur[i];
                nbest = ncur;
            }
            return;
        }
        if ((nbest >= 0 && ncur + 1 > nbest) || ncur == kDepth) return;
        static const int kRadices[] = {16, 10, 8, 7, 5, 4, 3, 2};
        for (int R : kRadices)
            if (n % R == 0 && (ncur == 0 || R <= cur[ncur - 1])) {
                cur[ncur++] = R;
                run(n / R);
                --ncur;
            }
    }
};
}  // namespace

const char* plan_axis(int N, AxisPlan& pl, float2* tw, int* ntw) {
    std::memset(&pl, 0, sizeof(pl));
    pl.N = N;
    if (N < 1 || N > kFftBudget) return "FFT length outside [1, 2048]";
    RadixSearch rs;
    if (N > 1) rs.run(N);
    else rs.nbest = 0;
    if (rs.nbest < 0) return "FFT length has a prime factor above 7";
    if (rs.nbest > kMaxStages) return "too many FFT stages";
    fill_plan(N, rs.best, rs.nbest, pl, tw, ntw);
    return nullptr;
}

void fill_plan(int N, const int* rad, int nst, AxisPlan& pl, float2* tw, int* ntw) {
    std::memset(&pl, 0, sizeof(pl));
    pl.N = N;
    pl.nst = nst;
    int p = 1, n = 0;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = rad[s];
        pl.rad[s] = R;
        pl.tw_off[s] = n;
        for (int k = 0; k < p; ++k)
            for (int r = 1; r < R; ++r) {
                const double ang = -2.0 * kPi * (double)r * (double)k / (double)(p * R);
                tw[n++] = make_float2((float)std::cos(ang), (float)std::sin(ang));
            }
        p *= R;
    }
    if (n == 0) tw[n++] = make_float2(1.f, 0.f);
    pl.ntw = n;
    *ntw = n;
}

int make_axis_plan(int N, AxisPlan& pl, std::vector<float2>& tw, std::string& err) {
    tw.assign((size_t)std::max(N, 1), make_float2(0.f, 0.f));
    int ntw = 0;
    if (const char* why = plan_axis(N, pl, tw.data(), &ntw)) {
        err = std::string(why) + " (N = " + std::to_string(N) + ")";
        tw.clear();
        return -TILD_ERR_UNSUPPORTED;
    }
    tw.resize(ntw);
    return 0;
}

// ---------------------------------------------------------------- tiles
size_t tile_lds_bytes(int Nx, int nf, int By, int Bz) {
    size_t rows = (size_t)nf * By * Bz;
    return rows * Nx * sizeof(unsigned long long) + lds_bytes(Nx, (int)((rows + 1) / 2));
}

bool choose_tiles(int dim, const int* Nx, int P, int nf, int target_tiles, int want_by, int want_bz,
                  int& By, int& Bz) {
    const int Ny = Nx[1], Nz = dim == 3 ? Nx[2] : 1;
    double best = 1e300;
    bool found = false;
    for (int by = P + 1; by <= Ny / 2; ++by) {
        if (Ny % by) continue;
        if (want_by > 0 && by != want_by) continue;
        // even By keeps the x-transform row pairs (y, y+1) of the global-atomic path, so
        // the tile shape does not touch a single bit of the result
        if (want_by <= 0 && Ny % 2 == 0 && by % 2) continue;
        for (int bz = (dim == 3 ? P + 1 : 1); bz <= (dim == 3 ? Nz / 2 : 1); ++bz) {
            if (Nz % bz) continue;
            if (want_bz > 0 && dim == 3 && bz != want_bz) continue;
            const int rows = nf * by * bz;
            if ((long)Nx[0] * ((rows + 1) / 2) > kFftBudget) continue;
            if (tile_lds_bytes(Nx[0], nf, by, bz) > kMaxLds) continue;
            const int ntiles = (Ny / by) * (Nz / bz);
            // prefer about target_tiles blocks, then square-ish tiles
            double score = std::fabs(std::log((double)ntiles / target_tiles)) + 0.05 * std::fabs(std::log((double)by / bz));
            if (dim == 2) score = std::fabs(std::log((double)ntiles / target_tiles));
            if (score < best) { best = score; By = by; Bz = bz; found = true; }
        }
    }
    return found;
}

// ---------------------------------------------------------------- helpers
static int fixed_bits(double bound) {
    // integer bits needed for sums up to `bound` (with sign) -> fractional bits
    int ib = (int)std::ceil(std::log2(bound + 1.0)) + 1;
    int F = 62 - ib;
    return std::max(16, std::min(44, F));
}

template <class T>
static int dalloc(Replica& R, T** p, size_t count, std::string& err, bool zero = false) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIPCHK(hipMalloc((void**)p, bytes));
    R.allocs.push_back((void*)*p);
    R.bytes += bytes;
    if (zero) HIPCHK(hipMemset(*p, 0, bytes));
    return 0;
}

Batch::~Batch() {
    if (device_ >= 0) (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto& kv : set_cache_) {
        free_maps(kv.second->maps);
        if (kv.second->d_ids) (void)hipFree(kv.second->d_ids);
        drop_graphs(*kv.second);
    }
    free_maps(reorder_maps_);
    for (auto& r : reps_)
        if (r) {
            free_maps(r->maps);
            for (void* p : r->allocs) (void)hipFree(p);
        }
    if (d_reps_) (void)hipFree(d_reps_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (setup_stream_) (void)hipStreamDestroy(setup_stream_);
}

void Batch::free_maps(PhaseMap* maps) {
    for (int p = 0; p < PH_COUNT; ++p)
        if (maps[p].dev) { (void)hipFree(maps[p].dev); maps[p].dev = nullptr; maps[p].cap = 0; }
}

int Batch::init(std::string& err) {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device_ < 0 || device_ >= ndev) { err = "no HIP device " + std::to_string(device_); return -TILD_ERR_DEVICE; }
    HIPCHK(hipSetDevice(device_));
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&setup_stream_, hipStreamNonBlocking));
    HIPCHK(configure_kernels());
    HIPCHK(configure_last_reg());
    HIPCHK(configure_plane_reg());
    HIPCHK(configure_plane_inv());
    return 0;
}

size_t Batch::estimate_bytes(const System& S) {
    size_t D = S.Dim, ns = S.ns, M = S.M;
    size_t Nxh = S.Nx[0] / 2 + 1, Mh = M / S.Nx[0] * Nxh;
    size_t nf = S.ntypes + (S.charges.enabled ? 1 : 0);
    size_t T = S.ntypes, nst = S.styles().size();
    // X[2], XO, fout; type, charge, bond tables, integrator kind; sorted-bin keys and records
    // site ids; spare arrays, keys and sort scratch of the spatial reorder
    size_t part = ns * (4 * D * 4 + 4 + 4 + 24 + 1) + ns * (8 + 20 + 28) + ns * 4 + ns * (4 * D + 100);
    size_t grids = nf * M * (8 + 4 + 12) + nf * Mh * 8 * (1 + D);
    size_t tabs = Mh * 4 * (T * (T + 1) / 2 + 1 + nst + 1) + Mh * 8 * (D == 3 ? 6 : 3) * nst;
    if (!S.Fields.empty()) tabs += T * D * M * 4;
    return part + grids + tabs + (1 << 20);
}

// ---------------------------------------------------------------- replica setup
int Batch::add(const System& S, uint32_t seed, std::string& err) {
    if (!autotune_ || tuning_) return add_plain(S, seed, err);
    HIPCHK(hipSetDevice(device_));
    TuneResult t;
    if (int rc = tune_geometry(S, seed, t, err)) return rc;
    const int id = add_with(S, seed, t.g, err);
    if (id >= 0) reps_[id]->tune = t;
    return id;
}

int Batch::add_plain(const System& S, uint32_t seed, std::string& err) {
    HIPCHK(hipSetDevice(device_));
    // every device operation of the build goes to the set-up stream (and the build ends
    // by synchronising it), so it overlaps whatever runs on the step stream
    struct StreamSwap {
        hipStream_t &a, &b;
        StreamSwap(hipStream_t& x, hipStream_t& y) : a(x), b(y) { std::swap(a, b); }
        ~StreamSwap() { std::swap(a, b); }
    } swap_streams(stream_, setup_stream_);
    auto Rp = std::make_unique<Replica>();
    Replica& R = *Rp;
    RepDev& h = R.h;
    std::memset(&h, 0, sizeof(h));
    const int D = S.Dim;
    if (S.ntypes > kMaxTypes) { err = "more than 16 particle types"; return -TILD_ERR_UNSUPPORTED; }
    if (S.ns <= 0) { err = "no particles"; return -TILD_ERR_CONFIG; }
    h.dim = D; h.Nx = S.Nx[0]; h.Ny = S.Nx[1]; h.Nz = D == 3 ? S.Nx[2] : 1;
    h.Nxh = h.Nx / 2 + 1;
    h.M = S.M; h.Mh = h.Nxh * h.Ny * h.Nz;
    h.ns = S.ns; h.ntypes = S.ntypes; h.nf = S.ntypes + (S.charges.enabled ? 1 : 0); h.P = S.pmeorder;
    h.nP = S.n_P_comps;
    for (int d = 0; d < 3; ++d) { h.L[d] = S.L[d]; h.Lh[d] = S.Lh[d]; h.dx[d] = S.dx[d]; h.idx[d] = 1.0f / S.dx[d]; }
    h.gvol = S.gvol;
    for (int d = 0; d < D; ++d) h.kfac[d] = (float)(2.0 * kPi / (double)S.L[d]);
    R.ns = S.ns; R.dim = D; R.ntypes = S.ntypes; R.nf = h.nf; R.M = h.M; R.Mh = h.Mh; R.nP = h.nP;
    for (int d = 0; d < 3; ++d) R.L[d] = S.L[d];
    R.V = 1.0;
    for (int d = 0; d < D; ++d) R.V *= (double)S.L[d];

    // fixed point: |w| <= 2 for every supported order; sums over all sites
    double qmax = 0.0;
    for (float qq : S.q) qmax = std::max(qmax, (double)std::fabs(qq));
    int F = fixed_bits(2.0 * S.ns), Fq = fixed_bits(2.0 * S.ns * std::max(qmax, 1e-30));
    h.fix_scale = std::ldexp(1.0f, F);
    h.fix_s2 = 44 - F;   // F <= 44
    h.fix_s3 = 52 - F;
    h.fixq_scale = std::ldexp(1.0f, Fq);
    h.fixq_s1 = std::ldexp(1.0f, Fq - 22 * D);   // order-1 charges: 2^-22 per quantised axis weight
    h.fix_inv = std::ldexp(1.0, -F) / (double)S.gvol;
    h.fixq_inv = std::ldexp(1.0, -Fq) / (double)S.gvol;

    // FFT plans
    std::vector<float2> twall;
    int toff[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) {
        std::vector<float2> tw;
        int rc = make_axis_plan(S.Nx[a], h.ax[a], tw, err);
        if (rc) return rc;
        toff[a] = (int)twall.size();
        twall.insert(twall.end(), tw.begin(), tw.end());
    }
    // inverse plan of the last axis: a two-stage plan R1 x R2 runs reversed (R2 x R1),
    // whose first stage reads the spectrum values k = i + R1 q that one column of the
    // forward transform ends with (the register-resident fused pass keeps them)
    const int lax = D - 1;
    int toff_r = toff[lax];
    h.axr = h.ax[lax];
    if (h.ax[lax].nst == 2) {
        const int rr[2] = {h.ax[lax].rad[1], h.ax[lax].rad[0]};
        std::vector<float2> tw((size_t)h.ax[lax].N);
        int ntw = 0;
        fill_plan(h.ax[lax].N, rr, 2, h.axr, tw.data(), &ntw);
        toff_r = (int)twall.size();
        twall.insert(twall.end(), tw.begin(), tw.begin() + ntw);
    }
    float2* d_tw;
    if (int rc = dalloc(R, &d_tw, twall.size(), err)) return rc;
    HIPCHK(hipMemcpy(d_tw, twall.data(), twall.size() * sizeof(float2), hipMemcpyHostToDevice));
    for (int a = 0; a < D; ++a) h.ax[a].tw = d_tw + toff[a];
    h.axr.tw = d_tw + toff_r;
    if (D == 2) h.ax[2] = h.ax[1];

    // chunking of each pass
    const int rows = h.M / h.Nx;
    const int npairs_x = (h.nf * rows + 1) / 2;
    h.nl_x = std::max(1, std::min(kFftBudget / h.Nx, npairs_x));
    const int la = D - 1, Nla = la == 2 ? h.Nz : h.Ny, Q = la == 2 ? h.Ny * h.Nxh : h.Nxh;
    h.last_dirs = tune_.last_dirs > 0 ? std::min(tune_.last_dirs, 2) : 1;
    if (Nla * h.nf > kFftBudget || last_lds_bytes(Nla, h.nf, h.last_dirs, 1) > kPlaneLds) { err = "grid too large for the fused k-space pass"; return -TILD_ERR_UNSUPPORTED; }
    // about 1,280 complex values per field set per block measured best on the
    // 80x80x160 melt (fewer, fuller blocks lose to the per-block latency chain)
    int nl = std::max(1, std::min(kFftBudget, 1280) / (Nla * h.nf));
    while (nl > 1 && last_lds_bytes(Nla, h.nf, h.last_dirs, nl) > kPlaneLds) --nl;
    h.nl_last = std::min(nl, Q);
    const int nc = (h.nf * D + 1) / 2;
    if (nc * h.Nx > kFftBudget) { err = "grid too large for the inverse x pass"; return -TILD_ERR_UNSUPPORTED; }
    // about 1,000 complex values per block measured best for the inverse x pass
    h.nr_c2r = std::max(1, std::min(1024 / (nc * h.Nx), rows));
    // optional block geometry (batching of independent lines: results do not change)
    if (tune_.last_lines > 0) {
        h.nl_last = std::min(tune_.last_lines, Q);
        while (h.nl_last > 1 && last_lds_bytes(Nla, h.nf, h.last_dirs, h.nl_last) > kPlaneLds) --h.nl_last;
    }
    // the register-resident form of the fused pass where it applies: as many columns
    // per block as its threads hold (the coupling table is blocked to match)
    // per (column, index) the pass addresses spectra and table by 32-bit offsets
    const bool off32 = (size_t)h.nf * D * h.Mh * 8 <= (size_t)0xFFFFFFFFu &&
                       (size_t)h.Mh * 80 <= (size_t)0xFFFFFFFFu;   // table: at most 16 rows, padded
    h.last_reg = !tune_.staged && off32 && last_reg_plan(h.ax[la], h.ntypes, h.nf) &&
                 last_reg_lds_bytes(h.ax[la], h.nf) <= kPlaneLds ? 1 : 0;
    if (h.last_reg) h.nl_last = last_reg_lines(h.ax[la]);
    h.grad_reg = !tune_.staged && D == 3 && off32 && grad_reg_plan(h.ax[1]) ? 1 : 0;
    if (tune_.c2r_rows > 0) {
        h.nr_c2r = std::min(tune_.c2r_rows, rows);
        while (h.nr_c2r > 1 && (size_t)h.nr_c2r * nc * h.Nx > (size_t)kFftBudget) --h.nr_c2r;
    }
    if (tune_.mid_elems > 0) h.mid_elems = std::min(kFftBudget, tune_.mid_elems);
    h.nl_mid = mid_geom(h, 1, h.nf).nl;

    // particle state
    const size_t nsD = (size_t)S.ns * D;
    if (int rc = dalloc(R, &h.X[0], nsD, err)) return rc;
    if (int rc = dalloc(R, &h.X[1], nsD, err)) return rc;
    if (int rc = dalloc(R, &h.XO, nsD, err)) return rc;
    HIPCHK(hipMemcpy(h.X[0], S.x.data(), nsD * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h.X[1], S.x.data(), nsD * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(h.XO, 0, nsD * 4));   // xo = x (device_comm_utils.cu:15-16): zero displacement
    int* d_tp;
    if (int rc = dalloc(R, &d_tp, S.ns, err)) return rc;
    HIPCHK(hipMemcpy(d_tp, S.tp.data(), S.ns * 4, hipMemcpyHostToDevice));
    h.tp = d_tp;
    if (S.charges.enabled) {
        float* d_q;
        if (int rc = dalloc(R, &d_q, S.ns, err)) return rc;
        HIPCHK(hipMemcpy(d_q, S.q.data(), S.ns * 4, hipMemcpyHostToDevice));
        h.q = d_q;
    }
    {
        // one int per bond slot: partner << s | type, with s the bits of the largest
        // bond type (0 for one type), -1 for an empty slot; partners need 31 - s bits
        int s = 0;
        while ((1 << s) < std::max(S.nbond_types, 1)) ++s;
        if (S.ns - 1 > (int)(0x7fffffffu >> s)) { err = "too many sites for the bond-type count"; return -TILD_ERR_UNSUPPORTED; }
        h.bond_shift = s;
        std::vector<int> bv((size_t)S.ns * 3, -1);
        for (int i = 0; i < S.ns; ++i)
            for (int b = 0; b < S.n_bonds[i]; ++b)
                bv[(size_t)i * 3 + b] = (S.bonded_to[(size_t)i * kMaxBonds + b] << s) | S.bond_type[(size_t)i * kMaxBonds + b];
        int* d_bv;
        float *d_bk, *d_br;
        if (int rc = dalloc(R, &d_bv, bv.size(), err)) return rc;
        if (int rc = dalloc(R, &d_bk, S.bond_k.size(), err)) return rc;
        if (int rc = dalloc(R, &d_br, S.bond_req.size(), err)) return rc;
        HIPCHK(hipMemcpy(d_bv, bv.data(), bv.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_bk, S.bond_k.data(), S.bond_k.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_br, S.bond_req.data(), S.bond_req.size() * 4, hipMemcpyHostToDevice));
        h.bond = d_bv; h.bond_k = d_bk; h.bond_r0 = d_br;
    }
    // integrators (integrator.cu:24-39): per-site kind; a site may belong to one integrator
    {
        std::vector<signed char> kind(S.ns, 0);
        for (const auto& I : S.Integrators) {
            const Group& g = S.Groups[I.group_index];
            for (int i : g.index) {
                if (kind[i]) { err = "site " + std::to_string(i + 1) + " is in more than one integrator group"; return -TILD_ERR_UNSUPPORTED; }
                kind[i] = I.name == "EM" ? 1 : 2;
            }
        }
        signed char* d_k;
        if (int rc = dalloc(R, &d_k, S.ns, err)) return rc;
        HIPCHK(hipMemcpy(d_k, kind.data(), S.ns, hipMemcpyHostToDevice));
        h.ikind = d_k;
        h.any_gjf = std::find(kind.begin(), kind.end(), (signed char)2) != kind.end() ? 1 : 0;
    }
    h.delt = S.delt;
    h.noise_mag = S.noise_mag;
    for (int t = 0; t < S.ntypes; ++t) {   // device_GJF_integrator.cu:25-28, 36, 46-48 in FP32
        float m = S.mass[t], Df = S.Diff[t], dt = S.delt;
        h.gjf_b[t] = m / (m + dt / (2.0f * Df));
        h.gjf_a[t] = (1.0f - dt / (2.0f * Df * m)) / (1.0f + dt / (2.0f * Df * m));
        h.gjf_c1[t] = h.gjf_b[t] * (dt * dt) / m;
        h.gjf_c2[t] = h.gjf_b[t] * dt / (2.0f * m);
        h.gjf_noise[t] = S.noise_mag / Df;
    }
    if (int rc = dalloc(R, &R.d_fout, nsD, err, true)) return rc;
    h.fout = R.d_fout;
    {
        std::vector<int> ident(S.ns);
        for (int i = 0; i < S.ns; ++i) ident[i] = i;
        if (int rc = dalloc(R, &h.pid, S.ns, err)) return rc;
        HIPCHK(hipMemcpy(h.pid, ident.data(), S.ns * 4, hipMemcpyHostToDevice));
    }
    h.seed = seed;
    h.step = 0;

    // grids
    const size_t M = h.M, Mh = h.Mh, nf = h.nf;
    if (int rc = dalloc(R, &h.acc, nf * M, err, true)) return rc;
    if (int rc = dalloc(R, &h.rho, nf * M, err, true)) return rc;
    if (int rc = dalloc(R, &h.H, nf * Mh, err, true)) return rc;
    if (int rc = dalloc(R, &h.Gk, nf * D * Mh, err, true)) return rc;
    if (int rc = dalloc(R, &h.G3, nf * M * 3, err, true)) return rc;
    h.g3_off32 = nf * M * 12 <= (size_t)0xFFFFFFFFu ? 1 : 0;
    h.grid_off32 = h.g3_off32 && (size_t)nf * D * Mh * 8 <= (size_t)0xFFFFFFFFu ? 1 : 0;
    R.nblk_part = (S.ns + 255) / 256;
    if (int rc = dalloc(R, &R.d_bpart, 7 * (size_t)R.nblk_part, err, true)) return rc;
    h.bpart = R.d_bpart;

    // spread mode: LDS-tiled sorted bins when a tile shape fits, else global atomics
    h.spread_mode = 1;
    if (spread_mode_ != 1) {
        int By = 0, Bz = 0;
        if (choose_tiles(D, S.Nx, S.pmeorder, h.nf, tile_target_, tile_by_, tile_bz_, By, Bz)) {
            h.spread_mode = 2;
            h.By = By; h.Bz = Bz;
            h.iBy = 1.0f / (float)By; h.iBz = 1.0f / (float)Bz;
            h.tslack = tile_slack_;
            h.nty = h.Ny / By; h.ntz = h.Nz / Bz;
            h.ntiles = h.nty * h.ntz;
            // order 1: the claims write stencil records where their fields fit
            h.srec = S.pmeorder == 1 && h.Nx <= 65536 && By <= 256 && Bz <= 256 && S.ntypes < 512 ? 1 : 0;
            h.nbuckets = h.ntiles * 4;
            // Every bucket counter on a line of its own: device-scope atomics on one line
            // serialise, and a bucket line is claimed by every block of its tiles.
            h.tshift = h.nbuckets <= (1 << 16) ? 5 : 0;
            if (int rc = dalloc(R, &h.tcnt, (size_t)h.nbuckets << h.tshift, err, true)) return rc;
            if (int rc = dalloc(R, &h.toff, (size_t)h.nbuckets + 1, err, true)) return rc;
            if (int rc = dalloc(R, &h.tkey, S.ns, err)) return rc;
            if (int rc = dalloc(R, &h.tclaim, (size_t)2 * h.nbuckets << h.tshift, err, true)) return rc;
            if (int rc = dalloc(R, &h.tbase, (size_t)h.nbuckets + 1, err, true)) return rc;
            if (int rc = dalloc(R, &h.tcap, (size_t)h.nbuckets, err, true)) return rc;
            // regions hold a bucket's count plus the slack (k_tile_regions), at most this many
            const size_t nrec = (size_t)S.ns + (tile_slack_ > 0 ? (size_t)tile_slack_ * h.nbuckets
                                                                : (tile_slack_ == 0 ? S.ns / 4 + 16 * (size_t)h.nbuckets : 0));
            if (int rc = dalloc(R, &h.trec, nrec, err)) return rc;
            if (S.charges.enabled) { if (int rc = dalloc(R, &h.trecq, nrec, err)) return rc; }
            if (int rc = dalloc(R, &h.tovf, 64, err, true)) return rc;
            if (int rc = dalloc(R, &h.tovf_rec, S.ns, err)) return rc;
            if (int rc = dalloc(R, &h.tovf_b, S.ns, err)) return rc;
            if (S.charges.enabled) { if (int rc = dalloc(R, &h.tovf_q, S.ns, err)) return rc; }
        } else if (spread_mode_ == 2) {
            err = "no tile shape fits this grid for the tiled spread";
            return -TILD_ERR_UNSUPPORTED;
        }
    }

    // static field-phase forces, summed per type: +f on type1, -f on type2
    if (!S.Fields.empty()) {
        std::vector<float> fp((size_t)S.ntypes * D * M, 0.0f), fr;
        for (const FieldPhase& F : S.Fields) {
            fieldphase_force_table(S, F, fr);
            for (int d = 0; d < D; ++d)
                for (size_t i = 0; i < M; ++i) {
                    fp[((size_t)F.type1 * D + d) * M + i] += fr[i * D + d];
                    fp[((size_t)F.type2 * D + d) * M + i] -= fr[i * D + d];
                }
        }
        float* d_fp;
        if (int rc = dalloc(R, &d_fp, fp.size(), err)) return rc;
        HIPCHK(hipMemcpy(d_fp, fp.data(), fp.size() * 4, hipMemcpyHostToDevice));
        h.FPtab = d_fp;
    }

    // spatial reorder: worthwhile once the gather records no longer fit one XCD's L2;
    // every 60 steps measured best on the 640,000-bead melt (+0.45 % over 40; 20 and
    // 80 lose, 80 because the gather locality decays).  Tiled replicas of 4,096 beads
    // or more reorder too: replicas run together share the L2s, and sets due at the
    // same step reorder in one launch per kernel (a 256-replica sweep of 17,000 beads
    // each: 1.35 -> 1.30 ms per step)
    const bool big = (size_t)h.nf * M * 3 * sizeof(float) > ((size_t)8 << 20);
    R.reorder_every = reorder_every_ >= 0 ? reorder_every_ : ((big || (h.spread_mode == 2 && S.ns >= 4096)) ? 60 : 0);
    if (R.reorder_every > 0) {
        const uint32_t maxkey = (uint32_t)(h.Ny * h.Nz - 1);
        R.sort_bits = 1;
        while (R.sort_bits < 32 && (maxkey >> R.sort_bits)) ++R.sort_bits;
        if (int rc = dalloc(R, &h.XO_alt, nsD, err)) return rc;
        if (int rc = dalloc(R, &h.tp_alt, S.ns, err)) return rc;
        if (S.charges.enabled) { if (int rc = dalloc(R, &h.q_alt, S.ns, err)) return rc; }
        if (int rc = dalloc(R, &h.bond_alt, (size_t)S.ns * 3, err)) return rc;
        if (int rc = dalloc(R, &h.ikind_alt, S.ns, err)) return rc;
        if (int rc = dalloc(R, &h.pid_alt, S.ns, err)) return rc;
        for (int k = 0; k < 2; ++k) {
            if (int rc = dalloc(R, &R.rb.keys[k], S.ns, err)) return rc;
            if (int rc = dalloc(R, &R.rb.vals[k], S.ns, err)) return rc;
        }
        h.rsrc = R.rb.vals[0]; h.rorder = R.rb.vals[1]; h.rinv = (int*)R.rb.keys[0];
        R.rb.tmp_bytes = reorder_tmp_bytes(S.ns, R.sort_bits);
        if (!R.rb.tmp_bytes) { err = "radix sort sizing failed"; return -TILD_ERR_DEVICE; }
        unsigned char* tmp;
        if (int rc = dalloc(R, &tmp, R.rb.tmp_bytes, err)) return rc;
        R.rb.tmp = tmp;
    }

    // forward plane pass: tiled 3D replicas with even Ny and even tile height (the
    // row pairing (y, y+1) of the tiled x transform), planes within one block's LDS
    h.plane_fwd = 0;
    if (D == 3 && h.spread_mode == 2 && tune_.plane >= 0 && h.Ny % 2 == 0 && h.By % 2 == 0 &&
        plane_r2c_lds(h) <= kPlaneLds)
        h.plane_fwd = 1;
    // its register-resident form: square planes with a two-stage plan
    h.plane_reg = h.plane_fwd && !tune_.staged && h.grid_off32 && plane_reg_plan(h.ax[0], h.ax[1]) &&
                  plane_reg_lds_bytes(h.ax[0]) <= kPlaneLds ? 1 : 0;
    // the y inverse and the x C2R as one pass per (field, z) plane: square planes with an
    // even two-stage plan, no field phases (whose records need the densities)
    h.plane_inv = D == 3 && !tune_.staged && h.grid_off32 && !h.FPtab && plane_inv_plan(h.ax[0], h.ax[1]) &&
                  h.ax[0].rad[0] == h.ax[1].rad[0] && h.ax[0].rad[1] == h.ax[1].rad[1] ? 1 : 0;

    int id = -1;
    for (int k = 0; k < (int)reps_.size(); ++k) if (!reps_[k]) { id = k; break; }
    if (id < 0) { id = (int)reps_.size(); reps_.emplace_back(); }
    reps_[id] = std::move(Rp);
    Replica& RR = *reps_[id];
    RR.dep_threads = dep_threads_;
    if (int rc = upload_desc(id, err)) { remove(id); return rc; }
    if (int rc = build_maps({id}, RR.maps, err)) { remove(id); return rc; }
    if (int rc = build_tables(RR, S, err)) { remove(id); return rc; }
    if (int rc = upload_desc(id, err)) { remove(id); return rc; }
    if (RR.reorder_every > 0) {
        if (int rc = reorder(id, err)) { remove(id); return rc; }
    }
    if (int rc = tile_regions(id, false, err)) { remove(id); return rc; }
    HIPCHK(hipStreamSynchronize(stream_));
    return id;
}

// Tiled replicas: exact binning of the current positions (tkey, toff), then, with
// `regions`, each bucket's region in the per-step record array sized from it.
int Batch::tile_regions(int id, bool bin_only, std::string& err) {
    Replica& R = *reps_[id];
    if (R.h.spread_mode != 2) return 0;
    if (int rc = unclaim(id, err)) return rc;
    PhaseMap* mp = R.maps;
    auto ctx = [&](Phase p) { return LaunchCtx{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_, nullptr, nullptr, mp[p].threads}; };
    HIPCHK(launch_tile_count(ctx(PH_TCOUNT), 0));
    HIPCHK(launch_tile_scan(ctx(PH_SCAN)));
    if (!bin_only) HIPCHK(launch_tile_regions(stream_, d_reps_, id));
    return 0;
}

int Batch::unclaim(int id, std::string& err) {
    Replica& R = *reps_[id];
    if (!R.claimed) return 0;
    PhaseMap* mp = R.maps;
    HIPCHK(launch_tile_reset(LaunchCtx{d_reps_, mp[PH_SCAN].dev, mp[PH_SCAN].n, 0, stream_}, 0));
    R.claimed = false;
    return 0;
}

int Batch::ensure_claimed(SetMaps& m, const std::vector<int>& ids, std::string& err,
                          hipEvent_t t0, hipEvent_t t1, bool* launched) {
    bool all = true;
    for (int id : ids) all = all && reps_[id]->claimed;
    if (launched) *launched = !all && m.maps[PH_TCOUNT].n > 0;
    if (all) return 0;
    // one claim pass over the set: replicas that hold claims drop them first
    for (int id : ids)
        if (int rc = unclaim(id, err)) return rc;
    HIPCHK(launch_tile_claim(LaunchCtx{d_reps_, m.maps[PH_TCOUNT].dev, m.maps[PH_TCOUNT].n, 0, stream_, t0, t1}, 0));
    for (int id : ids) reps_[id]->claimed = true;
    return 0;
}

int Batch::reorder(int id, std::string& err, bool after_step) {
    Replica& R = *reps_[id];
    if (int rc = unclaim(id, err)) return rc;   // the slots move
    if (after_step && R.h.spread_mode == 2) {   // bin, sort each tile by site id, permute
        if (int rc = tile_regions(id, true, err)) return rc;
        HIPCHK(launch_reorder_tiles(stream_, d_reps_, id, R.ns, R.h.ntiles, R.rb));
        HIPCHK(launch_tile_regions(stream_, d_reps_, id));   // bins do not move with the slots
    } else {
        HIPCHK(launch_reorder(stream_, d_reps_, id, R.ns, R.sort_bits, R.rb));
    }
    mirror_reorder(R);
    return 0;
}

void Batch::mirror_reorder(Replica& R) {
    // mirror the device-side swaps in the host copy of the descriptor
    RepDev& h = R.h;
    std::swap(h.X[0], h.X[1]);
    std::swap(h.XO, h.XO_alt);
    { const int* t = h.tp; h.tp = h.tp_alt; h.tp_alt = const_cast<int*>(t); }
    { const signed char* t = h.ikind; h.ikind = h.ikind_alt; h.ikind_alt = const_cast<signed char*>(t); }
    std::swap(h.pid, h.pid_alt);
    if (h.q) { const float* t = h.q; h.q = h.q_alt; h.q_alt = const_cast<float*>(t); }
    { const int* t = h.bond; h.bond = h.bond_alt; h.bond_alt = const_cast<int*>(t); }
    R.permuted = true;
    R.since_sort = 0;
}

int Batch::reorder_set(const std::vector<int>& ids, std::string& err) {
    if (int rc = build_maps(ids, reorder_maps_, err, true)) return rc;
    PhaseMap* mp = reorder_maps_;
    auto ctx = [&](Phase p) { return LaunchCtx{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_}; };
    // the claims in the regions move with the slots: clear the current parity's
    // counters of every replica (zero already where there were none)
    HIPCHK(launch_tile_reset(ctx(PH_SCAN), 0));
    for (int id : ids) reps_[id]->claimed = false;
    HIPCHK(launch_tile_count(ctx(PH_TCOUNT), 0));
    HIPCHK(launch_tile_scan(ctx(PH_SCAN)));
    HIPCHK(launch_reorder_tiles_set(ctx(PH_PART), ctx(PH_RTILE), ctx(PH_SCAN)));
    HIPCHK(launch_tile_regions_set(ctx(PH_SCAN)));
    for (int id : ids) mirror_reorder(*reps_[id]);
    return 0;
}

int Batch::fetch_pid(const Replica& R, std::vector<int>& pid, std::string& err) {
    pid.resize(R.ns);
    HIPCHK(hipMemcpyAsync(pid.data(), R.h.pid, (size_t)R.ns * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

void Batch::remove(int id) {
    if (id < 0 || id >= (int)reps_.size() || !reps_[id]) return;
    (void)hipStreamSynchronize(stream_);
    free_maps(reps_[id]->maps);
    for (void* p : reps_[id]->allocs) (void)hipFree(p);
    reps_[id].reset();
    for (auto it = set_cache_.begin(); it != set_cache_.end();) {
        if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) {
            free_maps(it->second->maps);
            if (it->second->d_ids) (void)hipFree(it->second->d_ids);
            drop_graphs(*it->second);
            it = set_cache_.erase(it);
        } else ++it;
    }
}

int Batch::upload_desc(int id, std::string& err) {
    if (id >= d_cap_) {
        // captured graphs hold the descriptor array's address: drop them
        for (auto& kv : set_cache_) drop_graphs(*kv.second);
        int cap = std::max(256, d_cap_);
        while (cap <= id) cap *= 2;
        RepDev* nd;
        HIPCHK(hipMalloc((void**)&nd, sizeof(RepDev) * cap));
        if (d_reps_) {
            HIPCHK(hipMemcpyAsync(nd, d_reps_, sizeof(RepDev) * d_cap_, hipMemcpyDeviceToDevice, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            HIPCHK(hipFree(d_reps_));
        }
        d_reps_ = nd;
        d_cap_ = cap;
    }
    Replica& R = *reps_[id];
    R.h.step = R.step;
    HIPCHK(hipMemcpyAsync(d_reps_ + id, &R.h, sizeof(RepDev), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

// ---------------------------------------------------------------- work maps
int Batch::build_maps(const std::vector<int>& ids, PhaseMap* maps, std::string& err, bool reorder_only) {
    for (int p = 0; p < PH_COUNT; ++p) { maps[p].host.clear(); maps[p].lds = 0; maps[p].n = 0; }
    for (int id : ids) {
        const Replica& R = *reps_[id];
        const RepDev& h = R.h;
        const int D = h.dim, rows = h.M / h.Nx;
        auto add = [&](Phase ph, int nblk, size_t lds) {
            for (int b = 0; b < nblk; ++b) maps[ph].host.push_back(WorkItem{id, b});
            maps[ph].lds = std::max(maps[ph].lds, lds);
        };
        add(PH_PART, R.nblk_part, 0);
        if (reorder_only) {   // the passes of reorder_set (tiled replicas)
            add(PH_TCOUNT, (h.ns + kCountBlock - 1) / kCountBlock, 0);
            add(PH_SCAN, 1, 0);
            add(PH_RTILE, h.ntiles, 0);
            continue;
        }
        const int npairs = (h.nf * rows + 1) / 2;
        const int nbx = (npairs + h.nl_x - 1) / h.nl_x;
        add(PH_XP, nbx, lds_bytes(h.Nx, h.nl_x));
        if (h.spread_mode == 2) {
            add(PH_TCOUNT, (h.ns + kCountBlock - 1) / kCountBlock, 0);
            add(PH_SCAN, 1, 0);
            add(PH_TILE, h.ntiles, tile_lds_bytes(h.Nx, h.nf, h.By, h.Bz));
            if (R.reorder_every > 0) add(PH_RTILE, h.ntiles, 0);
        } else {
            add(PH_SPREADA, R.nblk_part, 0);
            add(PH_XA, nbx, lds_bytes(h.Nx, h.nl_x));
        }
        {
            MidGeom g = mid_geom(h, 1, h.nf);
            add(PH_PY, g.nouter * g.nchunk, lds_bytes(g.N, g.nl));
            if (D == 3 && h.plane_reg) add(PH_PLANER, h.nf * h.Nz, plane_reg_lds_bytes(h.ax[0]));
            else if (D == 3 && h.plane_fwd) add(PH_PLANEF, h.nf * h.Nz, plane_r2c_lds(h));
            else if (D == 3) add(PH_MIDY_H, g.nouter * g.nchunk, lds_bytes(g.N, g.nl));
        }
        if (D == 3) {
            MidGeom g = mid_grad_geom(h);
            if (h.plane_inv) add(PH_PLANEI, h.nf * h.Nz, plane_inv_lds_bytes(h.ax[1]));
            else if (h.grad_reg) add(PH_MIDY_GR, h.nf * h.Nz * grad_reg_chunks(h.ax[1], h.Nxh), grad_reg_lds_bytes(h.ax[1]));
            else add(PH_MIDY_G, g.nouter * g.nchunk, lds_bytes(g.N, 3 * g.nl));
            MidGeom gz = mid_geom(h, 2, h.nf);
            add(PH_MIDZ_H, gz.nouter * gz.nchunk, lds_bytes(gz.N, gz.nl));
        }
        const int la = D - 1, Nla = la == 2 ? h.Nz : h.Ny, Q = la == 2 ? h.Ny * h.Nxh : h.Nxh;
        if (h.last_reg) add(last_reg_class(h.ntypes) ? PH_LASTR4 : PH_LASTR, (Q + h.nl_last - 1) / h.nl_last,
                            last_reg_lds_bytes(h.ax[la], h.nf));
        else add(PH_LAST, (Q + h.nl_last - 1) / h.nl_last, last_lds_bytes(Nla, h.nf, h.last_dirs, h.nl_last));
        const int nc = (h.nf * D + 1) / 2;
        if (!h.plane_inv) add(PH_C2R, (rows + h.nr_c2r - 1) / h.nr_c2r, lds_bytes(h.Nx, h.nr_c2r * nc));
        add(PH_KPROPS, h.nchunk > 0 ? h.nchunk : 1, 0);
    }
    // Deposit block size: 512 threads when the set's tiles hold ~600 beads or more on
    // average (one or two records per thread), else 256 (small tiles leave most of a
    // larger block idle, and the deposit of a sweep of small replicas took twice as long)
    {
        long long beads = 0, tiles = 0;
        int want = -1;   // the block size the set's tiled replicas ask for, if they agree
        for (int id : ids)
            if (reps_[id]->h.spread_mode == 2) {
                beads += reps_[id]->ns; tiles += reps_[id]->h.ntiles;
                const int w = reps_[id]->dep_threads;
                want = want == -1 ? w : (want == w ? want : 0);
            }
        maps[PH_TILE].threads = want == 256 || want == 512 ? want : (tiles > 0 && beads >= 600 * tiles ? 512 : 256);
    }
    // Workgroups go to the 8 XCDs round-robin (launch block w runs on XCD w % 8).  Every
    // pass hands each XCD a contiguous run of work items: for the particle-order passes
    // one slab of a spatially ordered replica, so that its L2 holds that slab's grid
    // records; for the FFT passes adjacent column batches, whose rows share cache lines
    // that a single L2 then fetches once.
    for (Phase ph : {PH_PART, PH_TCOUNT, PH_SPREADA, PH_TILE, PH_XA, PH_XP, PH_MIDY_H, PH_MIDY_G, PH_MIDY_GR,
                     PH_MIDZ_H, PH_PY, PH_LAST, PH_LASTR, PH_LASTR4, PH_C2R, PH_PLANEF, PH_PLANER, PH_PLANEI}) {
        std::vector<WorkItem>& v = maps[ph].host;
        const size_t n = v.size(), q = n / kXcds, rm = n % kXcds;
        if (n < 2 * (size_t)kXcds) continue;
        std::vector<WorkItem> o(n);
        for (size_t w = 0; w < n; ++w) {
            const size_t x = w % kXcds, k = w / kXcds;
            o[w] = v[x * q + std::min(x, rm) + k];
        }
        v.swap(o);
    }
    for (int p = 0; p < PH_COUNT; ++p) {
        maps[p].n = (int)maps[p].host.size();
        if (maps[p].lds > ((p == PH_PLANEF || p == PH_PLANER || p == PH_PLANEI || p == PH_LAST || p == PH_LASTR || p == PH_LASTR4) ? kPlaneLds : (size_t)kMaxLds)) { err = "pass needs more LDS than a block may use"; return -TILD_ERR_UNSUPPORTED; }
        if (maps[p].dev && (!reorder_only || maps[p].cap < maps[p].n)) {
            // reused buffers are last read by the previous reorder's launches
            HIPCHK(hipStreamSynchronize(stream_));
            (void)hipFree(maps[p].dev);
            maps[p].dev = nullptr;
            maps[p].cap = 0;
        }
        if (maps[p].n) {
            if (!maps[p].dev) {
                const int cap = reorder_only ? std::max(maps[p].n, 2 * maps[p].cap) : maps[p].n;
                HIPCHK(hipMalloc((void**)&maps[p].dev, (size_t)cap * sizeof(WorkItem)));
                maps[p].cap = cap;
            }
            // from pageable host memory: staged before the call returns, so the host
            // vector may change afterwards; ordered on the stream after the previous
            // reorder's launches
            HIPCHK(hipMemcpyAsync(maps[p].dev, maps[p].host.data(), maps[p].n * sizeof(WorkItem),
                                  hipMemcpyHostToDevice, stream_));
        }
    }
    if (!reorder_only) HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

Batch::SetMaps* Batch::maps_for(const std::vector<int>& ids, std::string& err) {
    auto it = set_cache_.find(ids);
    if (it != set_cache_.end()) return it->second.get();
    auto m = std::make_unique<SetMaps>();
    if (build_maps(ids, m->maps, err)) return nullptr;
    m->nids = (int)ids.size();
    if (hipMalloc((void**)&m->d_ids, std::max<size_t>(1, ids.size()) * sizeof(int)) != hipSuccess ||
        hipMemcpy(m->d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        err = "work map allocation failed";
        return nullptr;
    }
    SetMaps* raw = m.get();
    set_cache_[ids] = std::move(m);
    return raw;
}

// ---------------------------------------------------------------- tables
// k^2 of half-spectrum index idx = (z*Ny + y)*Nxh + kx with the reference's
// wavenumber convention (device_array_utils.cu:60-82), in FP64.
static inline double k2_of(const RepDev& h, size_t idx, double* kv) {
    int kx = (int)(idx % h.Nxh);
    size_t rest = idx / h.Nxh;
    int ky = (int)(rest % h.Ny);
    int kz = (int)(rest / h.Ny);
    int n[3] = {kx, ky, kz}, N[3] = {h.Nx, h.Ny, h.Nz};
    double k2 = 0.0;
    for (int d = 0; d < h.dim; ++d) {
        int s = (2 * n[d] < N[d]) ? n[d] : n[d] - N[d];
        kv[d] = 2.0 * kPi * (double)s / (double)h.L[d];
        k2 += kv[d] * kv[d];
    }
    return k2;
}

// f(begin, end) over [0, n) in contiguous ranges on up to 16 host threads (the k-space
// tables of a large grid take most of a replica's set-up time on one thread).
template <class F>
static void parallel_ranges(size_t n, F f) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({hw, 16, n / 8192 + 1});
    if (nt <= 1) { f((size_t)0, n); return; }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 0; t < nt; ++t) {
        const size_t b = t * per, e = std::min(n, b + per);
        if (b < e) th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

int Batch::build_tables(Replica& R, const System& S, std::string& err) {
    RepDev& h = R.h;
    const int T = S.ntypes, D = h.dim, nP = h.nP;
    const size_t Mh = h.Mh;
    const double M = (double)h.M, gvol = (double)S.gvol;
    auto styles = S.styles();
    R.nstyles = (int)styles.size();
    if (R.nstyles > 64) { err = "more than 64 pair styles"; return -TILD_ERR_UNSUPPORTED; }
    // ordered type pairs of the styles (virial tables are summed per pair)
    std::vector<std::pair<int, int>> pairs;
    std::vector<int> pair_of(styles.size());
    for (size_t s = 0; s < styles.size(); ++s) {
        std::pair<int, int> pq(styles[s]->type1, styles[s]->type2);
        auto it = std::find(pairs.begin(), pairs.end(), pq);
        if (it == pairs.end()) { pairs.push_back(pq); pair_of[s] = (int)pairs.size() - 1; }
        else pair_of[s] = (int)(it - pairs.begin());
    }
    if (pairs.size() > 32) { err = "more than 32 type pairs"; return -TILD_ERR_UNSUPPORTED; }
    R.npairs = (int)pairs.size();
    h.nstyles = R.nstyles; h.npairs = R.npairs;
    for (size_t s = 0; s < styles.size(); ++s) { h.st_t1[s] = styles[s]->type1; h.st_t2[s] = styles[s]->type2; }
    for (size_t p = 0; p < pairs.size(); ++p) { h.pr_t1[p] = pairs[p].first; h.pr_t2[p] = pairs[p].second; }
    R.nq = R.nstyles + R.npairs * nP + (S.charges.enabled ? 1 : 0);
    h.nq = R.nq;
    h.nchunk = (int)std::max<size_t>(1, std::min<size_t>(512, (Mh + 4095) / 4096));

    // coupling tables C_IJ (I <= J) / M, then the charge kernel, then one row per ramped
    // style: its table per unit prefactor, weighted per step by the kernel
    const int ncp = T * (T + 1) / 2;
    std::vector<int> ramped;
    for (size_t s = 0; s < styles.size(); ++s) if (styles[s]->ramp) ramped.push_back((int)s);
    if ((int)ramped.size() > kMaxRamp) { err = "more than 8 ramped pair styles"; return -TILD_ERR_UNSUPPORTED; }
    if (!ramped.empty() && S.max_steps <= 0) { err = "a ramp needs max_steps > 0"; return -TILD_ERR_UNSUPPORTED; }
    const int nrows = ncp + (S.charges.enabled ? 1 : 0) + (int)ramped.size();
    h.ctab_charge = S.charges.enabled ? ncp : -1;
    std::vector<float> ctab((size_t)nrows * Mh), etab((size_t)std::max(1, R.nstyles) * Mh), eltab;
    if (S.charges.enabled) eltab.resize(Mh);
    auto pidx = [&](int I, int J) { if (I > J) std::swap(I, J); return I * T - I * (I - 1) / 2 + (J - I); };
    // Ramp schedule of pair_style.cu:308-328.  After step s (1-based) the reference sets
    // the table to new/cur times itself, with new = 1000 A0 + delta (s + 1), delta =
    // 1000 (A0 - A_final) / max_steps, and cur = 1000 u[k=0]: the amplitude becomes
    // (A0 + (A0 - A_final)(s + 1) / max_steps) / u1(0), with u1 the table per unit
    // prefactor (u1(0) = 1 for a Gaussian, (4 pi Rp^3 / 3)^2 for erf, 4 pi Rp^3 / 3 for
    // gaussian_erf).  The ramp therefore moves away from A_final, and the virial
    // tables keep A0, as in the reference.  Step k (0-based) uses the table of the
    // update after step k, none at k = 0.
    h.nramp = (int)ramped.size();
    h.ramp_row0 = ncp + (S.charges.enabled ? 1 : 0);
    R.ramps.clear();
    for (int j = 0; j < h.nramp; ++j) {
        const PairStyle* ps = styles[ramped[j]];
        const double A0 = ps->initial_prefactor, u10 = ps->kspace(0.0, 1.0);
        h.ramp_pair[j] = pidx(ps->type1, ps->type2);
        h.ramp_A0[j] = A0;
        h.ramp_c0[j] = A0 / u10;
        h.ramp_c1[j] = (A0 - ps->final_prefactor) / ((double)S.max_steps * u10);
        R.ramps.push_back({ramped[j], A0, h.ramp_c0[j], h.ramp_c1[j]});
    }
    parallel_ranges(Mh, [&](size_t i0, size_t i1) {
    std::vector<double> c(ncp);
    for (size_t idx = i0; idx < i1; ++idx) {
        double kv[3];
        double k2 = k2_of(h, idx, kv);
        std::fill(c.begin(), c.end(), 0.0);
        for (size_t s = 0; s < styles.size(); ++s) {
            const PairStyle* ps = styles[s];
            double u = ps->kspace(k2, ps->initial_prefactor);
            // each style acts on type1 from type2 and on type2 from type1 (pair_style.cu:81-196)
            c[pidx(ps->type1, ps->type2)] += (ps->type1 == ps->type2) ? 2.0 * u : u;
            etab[s * Mh + idx] = (float)(u * gvol / M);
        }
        for (int p = 0; p < ncp; ++p) ctab[(size_t)p * Mh + idx] = (float)(c[p] / M);
        for (int j = 0; j < h.nramp; ++j) {
            const PairStyle* ps = styles[ramped[j]];
            const double u1 = ps->kspace(k2, 1.0) * (ps->type1 == ps->type2 ? 2.0 : 1.0);
            ctab[(size_t)(h.ramp_row0 + j) * Mh + idx] = (float)(u1 / M);
        }
        if (S.charges.enabled) {
            double g1 = S.charges.potential(k2);
            double a = S.charges.smearing_length;
            ctab[(size_t)ncp * Mh + idx] = (float)(g1 * std::exp(-k2 / (2.0 * a * a)) / M);
            eltab[idx] = (float)(g1 / M);
        }
    }
    });
    // The fused last-axis pass reads the coupling rows of nl_last adjacent columns for
    // every last-axis index; store them in exactly that order, block by block
    // ([column block][last-axis index][row][column]), so each block streams one
    // contiguous run instead of nl_last-wide pieces of every row.
    const int la = h.dim - 1, Nla = la == 2 ? h.Nz : h.Ny;
    const size_t Q = la == 2 ? (size_t)h.Ny * h.Nxh : (size_t)h.Nxh;
    const size_t nl = (size_t)h.nl_last, nqb = (Q + nl - 1) / nl;
    std::vector<float> cblk(nqb * Nla * nrows * nl, 0.0f);
    for (size_t qb = 0; qb < nqb; ++qb)
        for (int n = 0; n < Nla; ++n)
            for (int p = 0; p < nrows; ++p)
                for (size_t b = 0; b < nl; ++b) {
                    const size_t q = qb * nl + b;
                    if (q < Q) cblk[((qb * Nla + n) * nrows + p) * nl + b] = ctab[(size_t)p * Mh + (size_t)n * Q + q];
                }
    h.ctab_rows = nrows;
    float *d_ctab, *d_etab, *d_eltab = nullptr;
    if (int rc = dalloc(R, &d_ctab, cblk.size(), err)) return rc;
    if (int rc = dalloc(R, &d_etab, etab.size(), err)) return rc;
    HIPCHK(hipMemcpy(d_ctab, cblk.data(), cblk.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_etab, etab.data(), etab.size() * 4, hipMemcpyHostToDevice));
    if (S.charges.enabled) {
        if (int rc = dalloc(R, &d_eltab, Mh, err)) return rc;
        HIPCHK(hipMemcpy(d_eltab, eltab.data(), Mh * 4, hipMemcpyHostToDevice));
    }
    h.Ctab = d_ctab; h.Etab = d_etab; h.Eltab = d_eltab;
    double* d_part;
    if (int rc = dalloc(R, &d_part, (size_t)std::max(1, R.nq) * h.nchunk, err, true)) return rc;
    h.partials = d_part;
    if (int rc = dalloc(R, &R.d_results, (size_t)R.nq + 8, err, true)) return rc;

    // virial tables (pair_style.cu:262-276), on the device:
    //   f_d(r) = IFFT(-i k_d u(k)), vir_c(r) = dr_a f_b, V_c = FFT(vir_c) gvol/(V D M^2)
    float2* d_vtab;
    if (int rc = dalloc(R, &d_vtab, std::max<size_t>(1, (size_t)R.npairs * nP * Mh), err, true)) return rc;
    h.Vtab = d_vtab;
    if (R.nstyles > 0) {
        int id = -1;
        for (int k = 0; k < (int)reps_.size(); ++k) if (reps_[k].get() == &R) id = k;
        if (int rc = upload_desc(id, err)) return rc;
        float2 *tmpK, *tmpV;
        float *fr, *vr;
        // stream-ordered temporaries: freeing them does not wait for other streams
        HIPCHK(hipMallocAsync((void**)&tmpK, (size_t)D * Mh * sizeof(float2), stream_));
        HIPCHK(hipMallocAsync((void**)&tmpV, (size_t)nP * Mh * sizeof(float2), stream_));
        HIPCHK(hipMallocAsync((void**)&fr, (size_t)D * h.M * sizeof(float), stream_));
        HIPCHK(hipMallocAsync((void**)&vr, (size_t)nP * h.M * sizeof(float), stream_));
        std::vector<float2> hk((size_t)D * Mh), hv((size_t)nP * Mh);
        std::vector<double> vacc((size_t)R.npairs * nP * Mh * 2, 0.0);
        const double scale = gvol / (R.V * (double)D * M * M);
        const int rows = h.M / h.Nx;
        // generic pass maps for this replica
        auto iota = [&](int nblk, WorkItem** dm) -> int {
            std::vector<WorkItem> v(nblk);
            for (int b = 0; b < nblk; ++b) v[b] = WorkItem{id, b};
            HIPCHK(hipMallocAsync((void**)dm, nblk * sizeof(WorkItem), stream_));
            HIPCHK(hipMemcpyAsync(*dm, v.data(), nblk * sizeof(WorkItem), hipMemcpyHostToDevice, stream_));
            HIPCHK(hipStreamSynchronize(stream_));   // v is released on return
            return 0;
        };
        int rc = 0;
        std::vector<WorkItem*> tmp_maps;
        auto run_mid = [&](int axis, int nfb, int inverse, float2* buf) -> int {
            MidGeom g = mid_geom(h, axis, nfb);
            WorkItem* dm;
            if (int e = iota(g.nouter * g.nchunk, &dm)) return e;
            tmp_maps.push_back(dm);
            LaunchCtx c{d_reps_, dm, g.nouter * g.nchunk, lds_bytes(g.N, g.nl), stream_};
            HIPCHK(launch_mid(c, axis, 0, nfb, inverse, buf));
            return 0;
        };
        for (size_t s = 0; s < styles.size() && !rc; ++s) {
            const PairStyle* ps = styles[s];
            parallel_ranges(Mh, [&](size_t i0, size_t i1) {
                for (size_t idx = i0; idx < i1; ++idx) {
                    double kv[3];
                    double k2 = k2_of(h, idx, kv);
                    double u = ps->kspace(k2, ps->initial_prefactor);
                    int n[3] = {(int)(idx % h.Nxh), (int)((idx / h.Nxh) % h.Ny), (int)(idx / h.Nxh / h.Ny)};
                    int N[3] = {h.Nx, h.Ny, h.Nz};
                    for (int d = 0; d < D; ++d) {
                        double kd = (2 * n[d] == N[d]) ? 0.0 : kv[d];
                        // -i k_d u (k_d = 0 on the d-Nyquist plane)
                        hk[(size_t)d * Mh + idx] = make_float2(0.0f, (float)(-kd * u));
                    }
                }
            });
            HIPCHK(hipMemcpyAsync(tmpK, hk.data(), hk.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
            if (D == 3) { if ((rc = run_mid(2, D, 1, tmpK))) break; }
            if ((rc = run_mid(1, D, 1, tmpK))) break;
            {
                const int nc = (D + 1) / 2;
                int nr = std::max(1, std::min(kFftBudget / (nc * h.Nx), rows));
                // x_c2r uses h.nr_c2r; the generic call needs its own row chunk
                RepDev hh = h;
                hh.nr_c2r = nr;
                HIPCHK(hipMemcpyAsync(d_reps_ + id, &hh, sizeof(RepDev), hipMemcpyHostToDevice, stream_));
                WorkItem* dm;
                int nblk = (rows + nr - 1) / nr;
                if ((rc = iota(nblk, &dm))) break;
                tmp_maps.push_back(dm);
                LaunchCtx c{d_reps_, dm, nblk, lds_bytes(h.Nx, nr * nc), stream_};
                HIPCHK(launch_x_c2r(c, false, D, tmpK, fr));
                HIPCHK(launch_virial_source(stream_, d_reps_, id, h.M, fr, vr));
                // forward R2C of the nP virial fields
                const int npairs = (nP * rows + 1) / 2;
                int nlx = std::max(1, std::min(kFftBudget / h.Nx, npairs));
                hh.nl_x = nlx;
                HIPCHK(hipMemcpyAsync(d_reps_ + id, &hh, sizeof(RepDev), hipMemcpyHostToDevice, stream_));
                int nbx = (npairs + nlx - 1) / nlx;
                if ((rc = iota(nbx, &dm))) break;
                tmp_maps.push_back(dm);
                LaunchCtx cx{d_reps_, dm, nbx, lds_bytes(h.Nx, nlx), stream_};
                HIPCHK(launch_x_r2c(cx, false, nP * rows, vr, tmpV));
                HIPCHK(hipMemcpyAsync(d_reps_ + id, &h, sizeof(RepDev), hipMemcpyHostToDevice, stream_));
            }
            if ((rc = run_mid(1, nP, 0, tmpV))) break;
            if (D == 3) { if ((rc = run_mid(2, nP, 0, tmpV))) break; }
            HIPCHK(hipMemcpyAsync(hv.data(), tmpV, hv.size() * sizeof(float2), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            const int p = pair_of[s];
            for (int cc = 0; cc < nP; ++cc)
                for (size_t idx = 0; idx < Mh; ++idx) {
                    size_t o = (((size_t)p * nP + cc) * Mh + idx) * 2;
                    vacc[o] += hv[(size_t)cc * Mh + idx].x * scale;
                    vacc[o + 1] += hv[(size_t)cc * Mh + idx].y * scale;
                }
        }
        for (WorkItem* m : tmp_maps) (void)hipFreeAsync(m, stream_);
        (void)hipFreeAsync(tmpK, stream_); (void)hipFreeAsync(tmpV, stream_);
        (void)hipFreeAsync(fr, stream_); (void)hipFreeAsync(vr, stream_);
        (void)hipStreamSynchronize(stream_);
        if (rc) return rc;
        std::vector<float2> vt((size_t)R.npairs * nP * Mh);
        for (size_t k = 0; k < vt.size(); ++k) vt[k] = make_float2((float)vacc[2 * k], (float)vacc[2 * k + 1]);
        HIPCHK(hipMemcpy(d_vtab, vt.data(), vt.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    // kprops map for this replica (nchunk known now)
    return build_maps({(int)(std::find_if(reps_.begin(), reps_.end(), [&](const std::unique_ptr<Replica>& p) { return p.get() == &R; }) - reps_.begin())}, R.maps, err);
}

// ---------------------------------------------------------------- step pipeline
// nsteps consecutive steps (1 when not integrating): the kernels of step k read the
// device step counter plus k, and one launch at the end advances the counter.
int Batch::enqueue_steps(SetMaps& m, int nsteps, bool integrate, bool write_forces, std::string& err) {
    PhaseMap* mp = m.maps;
    auto ctx = [&](Phase p) { return LaunchCtx{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_, nullptr, nullptr, mp[p].threads}; };
    for (int k = 0; k < nsteps; ++k) {
        HIPCHK(launch_spread(ctx(PH_SPREADA), k));
        HIPCHK(launch_x_r2c(ctx(PH_XA), true, 0, nullptr, nullptr));
        HIPCHK(launch_tile_deposit(ctx(PH_TILE), k));
        HIPCHK(launch_plane_r2c(ctx(PH_PLANEF)));
        HIPCHK(launch_plane_reg(ctx(PH_PLANER)));
        HIPCHK(launch_mid(ctx(PH_MIDY_H), 1, 0, 0, 0, nullptr));
        HIPCHK(launch_last_fused(ctx(PH_LAST), 0, k));
        HIPCHK(launch_last_reg(ctx(PH_LASTR), k, 0));
        HIPCHK(launch_last_reg(ctx(PH_LASTR4), k, 1));
        HIPCHK(launch_mid_grad(ctx(PH_MIDY_G)));
        HIPCHK(launch_grad_reg(ctx(PH_MIDY_GR)));
        HIPCHK(launch_x_c2r(ctx(PH_C2R), true, 0, nullptr, nullptr));
        HIPCHK(launch_plane_inv(ctx(PH_PLANEI)));
        // integrating steps claim the next step's records; a force evaluation leaves
        // positions and records as they are
        HIPCHK(launch_particles(ctx(PH_PART), integrate, write_forces ? 1 : 0, k, integrate));
    }
    if (integrate) HIPCHK(launch_advance(stream_, d_reps_, m.d_ids, m.nids, nsteps));
    return 0;
}

int Batch::forces(int id, std::string& err) {
    Replica* R = rep(id);
    if (!R) { err = "no such replica"; return -TILD_ERR_ARG; }
    HIPCHK(hipSetDevice(device_));
    SetMaps* m = maps_for({id}, err);
    if (!m) return -TILD_ERR_DEVICE;
    if (int rc = ensure_claimed(*m, {id}, err)) return rc;
    return enqueue_steps(*m, 1, false, true, err);
}

int Batch::step(const std::vector<int>& ids_in, int n, std::string& err, bool use_graph) {
    std::vector<int> ids = ids_in;
    if (ids.empty())
        for (int k = 0; k < (int)reps_.size(); ++k) if (reps_[k] && reps_[k]->active) ids.push_back(k);
    if (ids.empty() || n <= 0) return 0;
    HIPCHK(hipSetDevice(device_));
    SetMaps* m = maps_for(ids, err);
    if (!m) return -TILD_ERR_DEVICE;
    // Graphs pay where launches would outlast the kernels: sets of small replicas.  A
    // set with more work per step keeps the device busy with eager launches, and a graph
    // launch leaves a gap of its own (below): on the 24,000-bead, 640,000-bead and
    // 81,000-bead decks, eager runs were 1-2 % faster than 60-step graphs.
    long long work = 0;
    for (int id : ids) work += (long long)reps_[id]->ns + reps_[id]->h.M;
    use_graph = use_graph && work <= kGraphMaxWork;
    if (use_graph && n >= kGraphSteps && !m->graph_run) {
        // Every kernel argument is constant across steps: step counters and position
        // parity live in the device descriptors and are advanced by the run's last
        // node, so kGraphSteps captured steps serve every run of that many.  Each graph
        // launch leaves a gap of about 40 us on the device (32 us before the next run
        // and 9 us before its counter update, on the 640,000-bead melt); eagerly
        // launched steps leave none once the kernels outlast the launches, so runs
        // shorter than kGraphSteps are launched eagerly (the melt: 10-step graphs took
        // 103.2 us per step, 60-step graphs 102.0, single-step graphs 108.1).
        if (int rc = capture_steps(*m, kGraphSteps, &m->graph_run, err)) return rc;
    }
    // chunks end where a replica of the set is due for its spatial reorder; each
    // replica reorders at multiples of its own interval, whatever set it runs in
    int done = 0;
    while (done < n) {
        int chunk = n - done;
        for (int id : ids) {
            const Replica& R = *reps_[id];
            if (R.reorder_every > 0) chunk = std::min(chunk, std::max(1, R.reorder_every - R.since_sort));
        }
        if (int rc = ensure_claimed(*m, ids, err)) return rc;
        if (int rc = run_steps(*m, ids, chunk, use_graph, err)) return rc;
        done += chunk;
        std::vector<int> due;
        bool all_tiled = true;
        for (int id : ids) {
            Replica& R = *reps_[id];
            if (R.reorder_every > 0 && (R.since_sort += chunk) >= R.reorder_every) {
                due.push_back(id);
                all_tiled = all_tiled && R.h.spread_mode == 2;
            }
        }
        if (due.size() > 1 && all_tiled) {   // one launch per kernel for the whole set
            if (int rc = reorder_set(due, err)) return rc;
        } else {
            for (int id : due)
                if (int rc = reorder(id, err, true)) return rc;
        }
    }
    for (int id : ids) reps_[id]->h.step = reps_[id]->step;
    return 0;
}

int Batch::capture_steps(SetMaps& m, int nsteps, hipGraphExec_t* out, std::string& err) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    rc = enqueue_steps(m, nsteps, true, false, err);
    hipError_t ec = hipStreamEndCapture(stream_, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    HIPCHK(ec);
    hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ei);
    return 0;
}

void Batch::drop_graphs(SetMaps& m) {
    if (m.graph_run) { (void)hipGraphExecDestroy(m.graph_run); m.graph_run = nullptr; }
}

int Batch::run_steps(SetMaps& m, const std::vector<int>& ids, int n, bool use_graph, std::string& err) {
    int s = 0;
    if (use_graph && m.graph_run)
        for (; s + kGraphSteps <= n; s += kGraphSteps) HIPCHK(hipGraphLaunch(m.graph_run, stream_));
    if (n > s)
        if (int rc = enqueue_steps(m, n - s, true, false, err)) return rc;
    for (int id : ids) reps_[id]->step += n;
    return 0;
}

int Batch::profile(const std::vector<int>& ids, int n, double* ms, int nph, std::string& err) {
    if (n <= 0 || !ms || nph < kProfPhases) { err = "profile: bad arguments"; return -TILD_ERR_ARG; }
    HIPCHK(hipSetDevice(device_));
    SetMaps* m = maps_for(ids, err);
    if (!m) return -TILD_ERR_DEVICE;
    // phases 0-7 are kernels, timed by events bound to each launch's own start and stop
    // (at most two kernels per phase: the two spread paths of a mixed set); phase 8,
    // the spatial reorders, by an event pair around them
    constexpr int kSlots = 3, kKernelPhases = kProfPhases - 1;
    hipEvent_t ev[kKernelPhases][kSlots][2], rv[2];
    for (auto& a : ev) for (auto& b : a) for (auto& e : b) HIPCHK(hipEventCreate(&e));
    for (auto& e : rv) HIPCHK(hipEventCreate(&e));
    std::vector<double> acc(kProfPhases, 0.0);
    PhaseMap* mp = m->maps;
    int rc = 0;
    for (int s = 0; s < n && !rc; ++s) {
        bool used[kKernelPhases][kSlots] = {};
        auto ctx = [&](Phase p, int ph, int slot) {
            LaunchCtx c{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_, nullptr, nullptr, mp[p].threads};
            if (c.n > 0) { c.t0 = ev[ph][slot][0]; c.t1 = ev[ph][slot][1]; used[ph][slot] = true; }
            return c;
        };
        hipError_t e0 = hipSuccess;
        auto ok = [&](hipError_t e) { if (e0 == hipSuccess) e0 = e; };
        ok(launch_spread(ctx(PH_SPREADA, 0, 0), 0));
        if (e0 == hipSuccess && ensure_claimed(*m, ids, err, ev[0][1][0], ev[0][1][1], &used[0][1]))   // after a reorder
            e0 = hipErrorUnknown;
        ok(launch_x_r2c(ctx(PH_XA, 1, 0), true, 0, nullptr, nullptr));
        ok(launch_tile_deposit(ctx(PH_TILE, 1, 1), 0));
        ok(launch_plane_r2c(ctx(PH_PLANEF, 2, 0)));
        ok(launch_plane_reg(ctx(PH_PLANER, 2, 2)));
        ok(launch_mid(ctx(PH_MIDY_H, 2, 1), 1, 0, 0, 0, nullptr));
        ok(launch_last_fused(ctx(PH_LAST, 3, 0), 0, 0));
        ok(launch_last_reg(ctx(PH_LASTR, 3, 1), 0, 0));
        ok(launch_last_reg(ctx(PH_LASTR4, 3, 2), 0, 1));
        ok(launch_mid_grad(ctx(PH_MIDY_G, 4, 0)));
        ok(launch_grad_reg(ctx(PH_MIDY_GR, 4, 1)));
        ok(launch_x_c2r(ctx(PH_C2R, 5, 0), true, 0, nullptr, nullptr));
        ok(launch_plane_inv(ctx(PH_PLANEI, 5, 1)));
        ok(launch_particles(ctx(PH_PART, 6, 0), true, 0, 0, true));
        ok(launch_advance(stream_, d_reps_, m->d_ids, m->nids, 1, ev[7][0][0], ev[7][0][1]));
        used[7][0] = m->nids > 0;
        // amortised spatial reorders, as step() runs them: one launch per kernel for a
        // set of tiled replicas due together
        std::vector<int> due;
        bool all_tiled = true;
        for (int id : ids) {
            Replica& R = *reps_[id];
            if (R.reorder_every > 0 && ++R.since_sort >= R.reorder_every) {
                due.push_back(id);
                all_tiled = all_tiled && R.h.spread_mode == 2;
            }
        }
        const bool reordered = e0 == hipSuccess && !due.empty();
        if (reordered) {
            ok(hipEventRecord(rv[0], stream_));
            if (due.size() > 1 && all_tiled) {
                if (reorder_set(due, err)) e0 = hipErrorUnknown;
            } else {
                for (int id : due)
                    if (e0 == hipSuccess && reorder(id, err, true)) e0 = hipErrorUnknown;
            }
            ok(hipEventRecord(rv[1], stream_));
        }
        ok(hipStreamSynchronize(stream_));
        if (e0 != hipSuccess) { err = std::string("profile: ") + hipGetErrorString(e0); rc = -TILD_ERR_DEVICE; break; }
        for (int k = 0; k < kKernelPhases; ++k)
            for (int j = 0; j < kSlots; ++j) {
                if (!used[k][j]) continue;
                float t = 0.f;
                if (hipEventElapsedTime(&t, ev[k][j][0], ev[k][j][1]) != hipSuccess) {
                    err = "profile: kernel timing unavailable"; rc = -TILD_ERR_DEVICE; break;
                }
                acc[k] += t;
            }
        if (reordered) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, rv[0], rv[1]);
            acc[kKernelPhases] += t;
        }
        for (int id : ids) reps_[id]->step += 1;
    }
    for (int id : ids) reps_[id]->h.step = reps_[id]->step;
    for (auto& a : ev) for (auto& b : a) for (auto& e : b) (void)hipEventDestroy(e);
    for (auto& e : rv) (void)hipEventDestroy(e);
    for (int k = 0; k < kProfPhases; ++k) ms[k] = acc[k] / n;
    return rc;
}

int Batch::forward_spectra(int id, std::string& err) {
    Replica& R = *reps_[id];
    PhaseMap* mp = R.maps;
    auto ctx = [&](Phase p) { return LaunchCtx{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_, nullptr, nullptr, mp[p].threads}; };
    HIPCHK(launch_x_r2c(ctx(PH_XP), false, 0, nullptr, nullptr));
    HIPCHK(launch_mid(ctx(PH_PY), 1, 0, 0, 0, nullptr));
    if (R.dim == 3) HIPCHK(launch_mid(ctx(PH_MIDZ_H), 2, 0, 0, 0, nullptr));
    return 0;
}

int Batch::properties(int id, tild_props* out, std::string& err) {
    Replica* Rp = rep(id);
    if (!Rp || !out) { err = "bad replica or output"; return -TILD_ERR_ARG; }
    Replica& R = *Rp;
    HIPCHK(hipSetDevice(device_));
    if (int rc = forward_spectra(id, err)) return rc;
    PhaseMap* mp = R.maps;
    auto ctx = [&](Phase p) { return LaunchCtx{d_reps_, mp[p].dev, mp[p].n, mp[p].lds, stream_, nullptr, nullptr, mp[p].threads}; };
    HIPCHK(launch_kprops(ctx(PH_KPROPS), R.nq));
    HIPCHK(launch_reduce_rows(stream_, R.h.partials, R.nq, R.h.nchunk, R.d_results));
    HIPCHK(launch_bond_props(ctx(PH_PART)));
    HIPCHK(launch_reduce_rows(stream_, R.d_bpart, 7, R.nblk_part, R.d_results + R.nq));
    std::vector<double> res((size_t)R.nq + 7);
    HIPCHK(hipMemcpyAsync(res.data(), R.d_results, res.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    std::memset(out, 0, sizeof(*out));
    const int nP = R.nP;
    out->n_styles = R.nstyles;
    double Ust = 0.0;
    out->nan_flag = 0;
    // ramped styles: the energy of the table in force at the last step taken
    // (0-based k = step - 1; the initial table before any step)
    for (const RampInfo& ri : R.ramps) {
        const int k = std::max(R.step - 1, 0);
        const double A = k == 0 ? ri.A0 : ri.c0 + ri.c1 * (double)(k + 1);
        res[ri.style] *= A / ri.A0;
    }
    for (int s = 0; s < R.nstyles && s < TILD_MAX_STYLES; ++s) {
        out->U_style[s] = res[s];
        Ust += res[s];
        if (std::isnan(res[s])) out->nan_flag = 1;
    }
    double P[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 0; p < R.npairs; ++p)
        for (int c = 0; c < nP; ++c) P[c] += res[R.nstyles + p * nP + c];
    out->Uelec = (R.nf > R.ntypes) ? res[R.nstyles + R.npairs * nP] : 0.0;
    out->Ubond = res[R.nq];
    const double inv = 1.0 / R.V / (double)R.dim;
    for (int c = 0; c < nP; ++c) out->Ptens[c] = P[c] + res[R.nq + 1 + c] * inv;
    out->Upe = Ust + out->Uelec + out->Ubond;
    out->step = R.step;
    return 0;
}

// ---------------------------------------------------------------- avg_sk compute
int Batch::sk_accumulate(int id, int slot, int type, std::string& err) {
    Replica* Rp = rep(id);
    if (!Rp || type < 0 || type >= Rp->ntypes || slot < 0 || slot >= kMaxSkSlots) {
        err = "sk_accumulate: bad replica, slot or type";
        return -TILD_ERR_ARG;
    }
    Replica& R = *Rp;
    HIPCHK(hipSetDevice(device_));
    if ((int)R.sk_buf.size() <= slot) {
        R.sk_buf.resize(slot + 1, nullptr);
        R.sk_type.resize(slot + 1, -1);
        R.sk_count.resize(slot + 1, 0);
    }
    if (!R.sk_buf[slot]) {
        if (int rc = dalloc(R, &R.sk_buf[slot], (size_t)R.Mh * 2, err, true)) return rc;
        R.sk_type[slot] = type;
    }
    if (R.sk_type[slot] != type) { err = "sk_accumulate: slot holds another type's sum"; return -TILD_ERR_ARG; }
    if (int rc = forward_spectra(id, err)) return rc;
    HIPCHK(launch_sk_accumulate(stream_, d_reps_, id, type, R.Mh, R.sk_buf[slot]));
    R.sk_count[slot] += 1;
    return 0;
}

int Batch::sk_average(int id, int slot, float* out, size_t n, int* count, std::string& err) {
    Replica* Rp = rep(id);
    if (!Rp || slot < 0 || slot >= kMaxSkSlots || !out || n != 2 * (size_t)Rp->M) {
        err = "sk_average: bad replica, slot or size";
        return -TILD_ERR_ARG;
    }
    Replica& R = *Rp;
    const int cnt = slot < (int)R.sk_count.size() ? R.sk_count[slot] : 0;
    if (count) *count = cnt;
    if (cnt == 0) { std::fill(out, out + n, 0.f); return 0; }
    HIPCHK(hipSetDevice(device_));
    std::vector<double> h((size_t)R.Mh * 2);
    HIPCHK(hipMemcpyAsync(h.data(), R.sk_buf[slot], h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    // full grid from the half spectrum: rho(-k) = conj rho(k) for real rho, so the
    // square at -k is the conjugate of the square at k
    const int Nx = R.h.Nx, Ny = R.h.Ny, Nz = R.dim == 3 ? R.h.Nz : 1, Nxh = R.h.Nxh;
    for (int z = 0; z < Nz; ++z)
        for (int y = 0; y < Ny; ++y)
            for (int x = 0; x < Nx; ++x) {
                const bool lower = x < Nxh;
                const int hx = lower ? x : Nx - x, hy = lower ? y : (Ny - y) % Ny, hz = lower ? z : (Nz - z) % Nz;
                const size_t hi = ((size_t)hz * Ny + hy) * Nxh + hx;
                const size_t fi = ((size_t)z * Ny + y) * Nx + x;
                out[2 * fi] = (float)(h[2 * hi] / cnt);
                out[2 * fi + 1] = (float)((lower ? h[2 * hi + 1] : -h[2 * hi + 1]) / cnt);
            }
    return 0;
}

// ---------------------------------------------------------------- host copies
// out[site] = in[slot] for slot p holding site pid[p] (identity when pid is empty).
template <class T>
static void unpermute(const T* in, T* out, size_t n, const std::vector<int>& pid, int D) {
    if (pid.empty()) { std::copy(in, in + n, out); return; }
    for (size_t p = 0; p < pid.size(); ++p)
        for (int d = 0; d < D; ++d) out[(size_t)pid[p] * D + d] = in[p * D + d];
}
int Batch::get_positions(int id, float* x, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !x || n != (size_t)R->ns * R->dim) { err = "get_positions: size mismatch"; return -TILD_ERR_ARG; }
    if (!R->permuted) {
        HIPCHK(hipMemcpyAsync(x, R->h.X[R->step & 1], n * 4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        return 0;
    }
    std::vector<float> tmp(n);
    std::vector<int> pid;
    HIPCHK(hipMemcpyAsync(tmp.data(), R->h.X[R->step & 1], n * 4, hipMemcpyDeviceToHost, stream_));
    if (int rc = fetch_pid(*R, pid, err)) return rc;
    unpermute(tmp.data(), x, n, pid, R->dim);
    return 0;
}

int Batch::set_positions(int id, const float* x, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !x || n != (size_t)R->ns * R->dim) { err = "set_positions: size mismatch"; return -TILD_ERR_ARG; }
    std::vector<float> tmp;
    const float* src = x;
    if (R->permuted) {   // slot p holds site pid[p]
        std::vector<int> pid;
        if (int rc = fetch_pid(*R, pid, err)) return rc;
        tmp.resize(n);
        const int D = R->dim;
        for (int p = 0; p < R->ns; ++p)
            for (int d = 0; d < D; ++d) tmp[(size_t)p * D + d] = x[(size_t)pid[p] * D + d];
        src = tmp.data();
    }
    HIPCHK(hipMemcpyAsync(R->h.X[R->step & 1], src, n * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemsetAsync(R->h.XO, 0, n * 4, stream_));   // xo = x
    if (int rc = tile_regions(id, false, err)) return rc;   // new positions, new bucket sizes
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Batch::get_state(int id, int* step, float* xo, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !step || !xo || n != (size_t)R->ns * R->dim) { err = "get_state: size mismatch"; return -TILD_ERR_ARG; }
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), R->h.XO, n * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    std::vector<int> pid;
    if (R->permuted) { if (int rc = fetch_pid(*R, pid, err)) return rc; }
    unpermute(tmp.data(), xo, n, pid, R->dim);
    *step = R->step;
    return 0;
}

int Batch::set_state(int id, int step, const float* xo, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !xo || step < 0 || n != (size_t)R->ns * R->dim) { err = "set_state: bad arguments"; return -TILD_ERR_ARG; }
    // the claims in the regions belong to the current step's parity, and the other
    // parity's counters are only cleared by the next step's deposit: clear both, here
    // for the current parity and after the step changes for the new one
    if (int rc = unclaim(id, err)) return rc;
    const bool tiled = R->h.spread_mode == 2;
    PhaseMap* mp = R->maps;
    const LaunchCtx rctx{d_reps_, mp[PH_SCAN].dev, mp[PH_SCAN].n, 0, stream_};
    if (tiled) HIPCHK(launch_tile_reset(rctx, 0));
    std::vector<float> tmp;
    const float* src = xo;
    if (R->permuted) {   // slot p holds site pid[p]
        std::vector<int> pid;
        if (int rc = fetch_pid(*R, pid, err)) return rc;
        tmp.resize(n);
        const int D = R->dim;
        for (int p = 0; p < R->ns; ++p)
            for (int d = 0; d < D; ++d) tmp[(size_t)p * D + d] = xo[(size_t)pid[p] * D + d];
        src = tmp.data();
    }
    HIPCHK(hipMemcpyAsync(R->h.XO, src, n * 4, hipMemcpyHostToDevice, stream_));
    // the positions live in the buffer of the step's parity
    if ((step & 1) != (R->step & 1))
        HIPCHK(hipMemcpyAsync(R->h.X[step & 1], R->h.X[R->step & 1], n * 4, hipMemcpyDeviceToDevice, stream_));
    R->step = step;
    if (int rc = upload_desc(id, err)) return rc;
    if (tiled) HIPCHK(launch_tile_reset(rctx, 0));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Batch::get_forces(int id, float* f, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !f || n != (size_t)R->ns * R->dim) { err = "get_forces: size mismatch"; return -TILD_ERR_ARG; }
    if (!R->permuted) {
        HIPCHK(hipMemcpyAsync(f, R->d_fout, n * 4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        return 0;
    }
    std::vector<float> tmp(n);
    std::vector<int> pid;
    HIPCHK(hipMemcpyAsync(tmp.data(), R->d_fout, n * 4, hipMemcpyDeviceToHost, stream_));
    if (int rc = fetch_pid(*R, pid, err)) return rc;
    unpermute(tmp.data(), f, n, pid, R->dim);
    return 0;
}

int Batch::get_density(int id, int type, float* rho, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !rho || n != (size_t)R->M || type < 0 || type >= R->nf) { err = "get_density: bad type or size"; return -TILD_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(rho, R->h.rho + (size_t)type * R->M, n * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Batch::get_grid_field(int id, int type, int dir, float* g, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !g || n != (size_t)R->M || type < 0 || type >= R->nf || dir < 0 || dir >= R->dim) {
        err = "get_grid_field: bad type, direction or size";
        return -TILD_ERR_ARG;
    }
    std::vector<float> tmp((size_t)R->M * 3);
    HIPCHK(hipMemcpyAsync(tmp.data(), R->h.G3 + (size_t)type * R->M * 3, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int i = 0; i < R->M; ++i) g[i] = tmp[(size_t)i * 3 + dir];
    return 0;
}

int Batch::get_bins(int id, int* out, size_t n, std::string& err) {
    Replica* R = rep(id);
    if (!R || !out || n != (size_t)R->ns * R->dim) { err = "get_bins: size mismatch"; return -TILD_ERR_ARG; }
    int* d;
    HIPCHK(hipMalloc((void**)&d, n * sizeof(int)));
    std::vector<int> tmp(n);
    hipError_t e1 = launch_bins(stream_, d_reps_, id, R->ns, d);
    hipError_t e2 = hipMemcpyAsync(tmp.data(), d, n * sizeof(int), hipMemcpyDeviceToHost, stream_);
    hipError_t e3 = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    HIPCHK(e1); HIPCHK(e2); HIPCHK(e3);
    std::vector<int> pid;
    if (R->permuted) { if (int rc = fetch_pid(*R, pid, err)) return rc; }
    unpermute(tmp.data(), out, n, pid, R->dim);
    return 0;
}

int Batch::nlist(int id, float rc, const int* sites, int nsites, int* cell_of, int* counts, int* neigh,
                 size_t cap, int* cell_max, int* ncells, std::string& err) {
    Replica* R = rep(id);
    if (!R || !(rc > 0.0f)) { err = "nlist: bad replica or cut-off"; return -TILD_ERR_ARG; }
    const int D = R->dim, ns = R->ns;
    // cell geometry in FP32 exactly as NList::Initialize (nlist.cu:36-51)
    NlistArgs a{};
    int total = 1;
    float cell_vol = 1.0f, V = 1.0f;
    for (int j = 0; j < D; ++j) V *= R->L[j];
    for (int j = 0; j < 3; ++j) { a.nc[j] = 1; a.Lc[j] = 1.0f; }
    for (int j = 0; j < D; ++j) {
        a.nc[j] = (int)(R->L[j] / rc);
        if (a.nc[j] < 1) { err = "nlist: cut-off longer than the box"; return -TILD_ERR_ARG; }
        a.Lc[j] = R->L[j] / (float)a.nc[j];
        total *= a.nc[j];
        cell_vol *= a.Lc[j];
    }
    a.cmax = (int)((float)((float)ns * cell_vol / V) * 7.5);
    a.rc = rc;
    if (cell_max) *cell_max = a.cmax;
    if (ncells) *ncells = total;
    if (!neigh) return 0;   // size query
    if (!cell_of || !counts || cap < (size_t)std::max(a.cmax, 1)) { err = "nlist: output buffers too small"; return -TILD_ERR_ARG; }
    if (a.cmax < 1) { err = "nlist: cell capacity below one site"; return -TILD_ERR_CONFIG; }
    // cells hold the reference's CELL_MAX sites; lists hold what the caller has room
    // for (the reference's hold CELL_MAX and are overrun beyond it)
    a.cap = (int)std::min(cap, (size_t)1 << 20);
    std::vector<unsigned char> mem(ns, sites ? 0 : 1);
    if (sites)
        for (int k = 0; k < nsites; ++k) {
            if (sites[k] < 0 || sites[k] >= ns) { err = "nlist: site id out of range"; return -TILD_ERR_ARG; }
            mem[sites[k]] = 1;
        }
    HIPCHK(hipSetDevice(device_));
    unsigned char* d_mem = nullptr;
    int *d_cell_of = nullptr, *d_counts = nullptr, *d_neigh = nullptr, *d_cc = nullptr, *d_cs = nullptr, *d_ovf = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)d_mem, (void*)d_cell_of, (void*)d_counts, (void*)d_neigh, (void*)d_cc, (void*)d_cs, (void*)d_ovf})
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess && x != hipSuccess) e = x; };
    chk(hipMalloc((void**)&d_mem, ns));
    chk(hipMalloc((void**)&d_cell_of, (size_t)ns * 4));
    chk(hipMalloc((void**)&d_counts, (size_t)ns * 4));
    chk(hipMalloc((void**)&d_neigh, (size_t)ns * a.cap * 4));
    chk(hipMalloc((void**)&d_cc, (size_t)total * 4));
    chk(hipMalloc((void**)&d_cs, (size_t)total * a.cmax * 4));
    chk(hipMalloc((void**)&d_ovf, 4));
    if (e == hipSuccess) {
        chk(hipMemcpyAsync(d_mem, mem.data(), ns, hipMemcpyHostToDevice, stream_));
        chk(hipMemsetAsync(d_cc, 0, (size_t)total * 4, stream_));
        chk(hipMemsetAsync(d_ovf, 0, 4, stream_));
        a.member = d_mem; a.cell_of = d_cell_of; a.counts = d_counts; a.neigh = d_neigh;
        a.cell_count = d_cc; a.cell_slots = d_cs; a.overflow = d_ovf;
        chk(launch_nlist(stream_, d_reps_, id, ns, a));
    }
    int ovf = 0;
    std::vector<int> nb;
    if (e == hipSuccess) {
        nb.resize((size_t)ns * a.cap);
        chk(hipMemcpyAsync(&ovf, d_ovf, 4, hipMemcpyDeviceToHost, stream_));
        chk(hipMemcpyAsync(cell_of, d_cell_of, (size_t)ns * 4, hipMemcpyDeviceToHost, stream_));
        chk(hipMemcpyAsync(counts, d_counts, (size_t)ns * 4, hipMemcpyDeviceToHost, stream_));
        chk(hipMemcpyAsync(nb.data(), d_neigh, nb.size() * 4, hipMemcpyDeviceToHost, stream_));
        chk(hipStreamSynchronize(stream_));
    }
    cleanup();
    if (e != hipSuccess) { err = std::string("nlist: ") + hipGetErrorString(e); return -TILD_ERR_DEVICE; }
    if (ovf & 1) { err = "nlist: more sites in a cell than CELL_MAX (nlist.cu:49-51)"; return -TILD_ERR_CONFIG; }
    if (ovf & 2) { err = "nlist: a neighbour list is longer than cap"; return -TILD_ERR_CONFIG; }
    for (int i = 0; i < ns; ++i) {
        const int c = std::min(counts[i], a.cap);
        std::copy(nb.begin() + (size_t)i * a.cap, nb.begin() + (size_t)i * a.cap + c, neigh + (size_t)i * cap);
    }
    return 0;
}

int Batch::sync(std::string& err) {
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

}  // namespace tild
