// abi.cpp -- extern "C" boundary of libtild.so (declared in include/tild.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/tild.h"
#include "engine.hpp"
#include "system.hpp"

struct tild_system { tild::System sys; };
struct tild_engine {
    std::unique_ptr<tild::Batch> batch;
    int id = -1;
    std::string err;
};
struct tild_batch {
    std::unique_ptr<tild::Batch> batch;
    std::string err;
};

namespace {
thread_local std::string g_last_error;
tild_status fail(tild_status s, const std::string& m) { g_last_error = m; return s; }
tild_status from_rc(int rc, std::string& err, tild_engine* e) {
    if (rc >= 0) return TILD_OK;
    if (e) e->err = err;
    g_last_error = err;
    return (tild_status)(-rc);
}
}  // namespace

namespace tild {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace tild

extern "C" {

const char* tild_last_error(void) { return g_last_error.c_str(); }

tild_status tild_system_load(const char* input_text, const char* base_dir, tild_system** out) {
    if (!out || !input_text) return fail(TILD_ERR_ARG, "null argument");
    *out = nullptr;
    auto s = std::make_unique<tild_system>();
    std::string err;
    int rc = tild::load_system(input_text, base_dir ? base_dir : "", s->sys, err);
    if (rc) return fail((tild_status)rc, err);
    *out = s.release();
    return TILD_OK;
}

void tild_system_free(tild_system* sys) { delete sys; }

tild_status tild_system_info(const tild_system* s, const char* key, double* v) {
    if (!s || !key || !v) return fail(TILD_ERR_ARG, "null argument");
    const tild::System& S = s->sys;
    std::string k(key);
    if (k == "ns") *v = S.ns;
    else if (k == "dim") *v = S.Dim;
    else if (k == "ntypes") *v = S.ntypes;
    else if (k == "M") *v = S.M;
    else if (k == "Nx0") *v = S.Nx[0];
    else if (k == "Nx1") *v = S.Nx[1];
    else if (k == "Nx2") *v = S.Dim > 2 ? S.Nx[2] : 1;
    else if (k == "pmeorder") *v = S.pmeorder;
    else if (k == "n_styles") *v = (double)S.styles().size();
    else if (k == "n_gaussians") *v = (double)S.Gausses.size();
    else if (k == "n_erfs") *v = (double)S.Erfs.size();
    else if (k == "n_gaussian_erfs") *v = (double)S.GaussErfs.size();
    else if (k == "n_fieldphases") *v = (double)S.Fields.size();
    else if (k == "do_charges") *v = S.charges.enabled ? 1 : 0;
    else if (k == "n_bonds_total") *v = S.n_total_bonds;
    else if (k == "nbond_types") *v = S.nbond_types;
    else if (k == "max_steps") *v = S.max_steps;
    else if (k == "log_freq") *v = S.log_freq;
    else if (k == "traj_freq") *v = S.traj_freq;
    else if (k == "grid_freq") *v = S.grid_freq;
    else if (k == "binary_freq") *v = S.bin_freq;
    else if (k == "n_groups") *v = (double)S.Groups.size();
    else if (k == "n_integrators") *v = (double)S.Integrators.size();
    else if (k == "n_computes") *v = (double)S.computes.size();
    else if (k == "seed") *v = S.rand_seed;
    else if (k == "L0") *v = S.L[0];
    else if (k == "L1") *v = S.L[1];
    else if (k == "L2") *v = S.L[2];
    else if (k == "dx0") *v = S.dx[0];
    else if (k == "dx1") *v = S.dx[1];
    else if (k == "dx2") *v = S.dx[2];
    else if (k == "gvol") *v = S.gvol;
    else if (k == "V") *v = S.V;
    else if (k == "delt") *v = S.delt;
    else if (k == "noise_mag") *v = S.noise_mag;
    else if (k == "bjerrum") *v = S.charges.bjerrum_length;
    else if (k == "smear") *v = S.charges.smearing_length;
    else return fail(TILD_ERR_ARG, "unknown key " + k);
    return TILD_OK;
}

tild_status tild_system_info_text(const tild_system* s, const char* key, char* buf, size_t cap) {
    if (!s || !key || !buf || cap == 0) return fail(TILD_ERR_ARG, "null argument");
    std::string k(key), v;
    if (k == "dump_name") v = s->sys.dump_name;
    else return fail(TILD_ERR_ARG, "unknown key " + k);
    if (v.size() + 1 > cap) return fail(TILD_ERR_ARG, "buffer too small");
    std::memcpy(buf, v.c_str(), v.size() + 1);
    return TILD_OK;
}

tild_status tild_system_get_positions(const tild_system* s, float* x, size_t n) {
    if (!s || !x || n != s->sys.x.size()) return fail(TILD_ERR_ARG, "size mismatch");
    std::memcpy(x, s->sys.x.data(), n * sizeof(float));
    return TILD_OK;
}
tild_status tild_system_get_types(const tild_system* s, int* tp, size_t n) {
    if (!s || !tp || n != (size_t)s->sys.ns) return fail(TILD_ERR_ARG, "size mismatch");
    std::memcpy(tp, s->sys.tp.data(), n * sizeof(int));
    return TILD_OK;
}
tild_status tild_system_get_molecules(const tild_system* s, int* mol, size_t n) {
    if (!s || !mol || n != (size_t)s->sys.ns) return fail(TILD_ERR_ARG, "size mismatch");
    std::memcpy(mol, s->sys.molecID.data(), n * sizeof(int));
    return TILD_OK;
}
tild_status tild_system_get_charges(const tild_system* s, float* q, size_t n) {
    if (!s || !q || n != (size_t)s->sys.ns) return fail(TILD_ERR_ARG, "size mismatch");
    std::memcpy(q, s->sys.q.data(), n * sizeof(float));
    return TILD_OK;
}
tild_status tild_system_get_bonds(const tild_system* s, int* nb, int* bt, int* ty, size_t ns) {
    if (!s || !nb || !bt || !ty || ns != (size_t)s->sys.ns) return fail(TILD_ERR_ARG, "size mismatch");
    std::memcpy(nb, s->sys.n_bonds.data(), ns * sizeof(int));
    std::memcpy(bt, s->sys.bonded_to.data(), ns * 3 * sizeof(int));
    std::memcpy(ty, s->sys.bond_type.data(), ns * 3 * sizeof(int));
    return TILD_OK;
}
tild_status tild_system_compute(const tild_system* s, int i, int* type, int* freq, int* wait) {
    if (!s || i < 0 || i >= (int)s->sys.computes.size()) return fail(TILD_ERR_ARG, "no such compute");
    const tild::SkCompute& c = s->sys.computes[i];
    if (type) *type = c.type;
    if (freq) *freq = c.freq;
    if (wait) *wait = c.wait;
    return TILD_OK;
}

tild_status tild_system_group_sites(const tild_system* s, int g, int* sites, size_t cap, int* nsites) {
    if (!s || g < 0 || g >= (int)s->sys.Groups.size()) return fail(TILD_ERR_ARG, "no such group");
    const auto& idx = s->sys.Groups[g].index;
    if (nsites) *nsites = (int)idx.size();
    if (sites) {
        if (cap < idx.size()) return fail(TILD_ERR_ARG, "buffer too small");
        std::memcpy(sites, idx.data(), idx.size() * sizeof(int));
    }
    return TILD_OK;
}

tild_status tild_engine_create(const tild_system* s, const tild_opts* opts, tild_engine** out) {
    if (!s || !out) return fail(TILD_ERR_ARG, "null argument");
    *out = nullptr;
    tild_opts o{};
    if (opts) o = *opts;
    auto e = std::make_unique<tild_engine>();
    e->batch = std::make_unique<tild::Batch>(o.device);
    e->batch->set_spread(o.spread_mode, o.tile_target > 0 ? o.tile_target : 512, o.tile_by, o.tile_bz, o.tile_slack);
    e->batch->set_reorder(o.reorder_every == 0 ? -1 : (o.reorder_every < 0 ? 0 : o.reorder_every));
    e->batch->set_fft_tuning({o.fft_mid_elems, o.fft_last_lines, o.fft_c2r_rows, o.fft_plane, o.fft_last_dirs,
                                   o.fft_staged ? 1 : 0});
    e->batch->set_deposit_threads(o.deposit_threads);
    e->batch->set_autotune(o.autotune != 0);
    std::string err;
    int rc = e->batch->init(err);
    if (rc < 0) return from_rc(rc, err, nullptr);
    uint32_t seed = o.use_seed ? o.seed : s->sys.rand_seed;
    int id = e->batch->add(s->sys, seed, err);
    if (id < 0) return from_rc(id, err, nullptr);
    e->id = id;
    *out = e.release();
    return TILD_OK;
}

void tild_engine_destroy(tild_engine* e) { delete e; }

const char* tild_engine_last_error(const tild_engine* e) { return e ? e->err.c_str() : g_last_error.c_str(); }

tild_status tild_engine_forces(tild_engine* e) {
    if (!e) return fail(TILD_ERR_ARG, "null engine");
    std::string err;
    int rc = e->batch->forces(e->id, err);
    if (rc >= 0) rc = e->batch->sync(err);
    return from_rc(rc, err, e);
}

tild_status tild_engine_step(tild_engine* e, int n) {
    if (!e || n < 0) return fail(TILD_ERR_ARG, "bad argument");
    std::string err;
    // several steps replay one captured step graph (same kernels, same results)
    int rc = e->batch->step({e->id}, n, err, n >= 2);
    if (rc >= 0) rc = e->batch->sync(err);
    return from_rc(rc, err, e);
}

tild_status tild_engine_properties(tild_engine* e, tild_props* out) {
    if (!e || !out) return fail(TILD_ERR_ARG, "null argument");
    std::string err;
    int rc = e->batch->properties(e->id, out, err);
    tild_status st = from_rc(rc, err, e);
    if (st == TILD_OK && out->nan_flag) { e->err = "NaN energy"; return TILD_ERR_NAN; }
    return st;
}

#define ENGINE_COPY(call)                                   \
    if (!e) return fail(TILD_ERR_ARG, "null engine");       \
    std::string err;                                        \
    int rc = e->batch->call;                                \
    return from_rc(rc, err, e);

tild_status tild_engine_get_positions(tild_engine* e, float* x, size_t n) { ENGINE_COPY(get_positions(e->id, x, n, err)) }
tild_status tild_engine_set_positions(tild_engine* e, const float* x, size_t n) { ENGINE_COPY(set_positions(e->id, x, n, err)) }
tild_status tild_engine_get_forces(tild_engine* e, float* f, size_t n) { ENGINE_COPY(get_forces(e->id, f, n, err)) }
tild_status tild_engine_get_density(tild_engine* e, int t, float* rho, size_t n) { ENGINE_COPY(get_density(e->id, t, rho, n, err)) }
tild_status tild_engine_get_grid_field(tild_engine* e, int t, int d, float* g, size_t n) { ENGINE_COPY(get_grid_field(e->id, t, d, g, n, err)) }
tild_status tild_engine_synchronize(tild_engine* e) { ENGINE_COPY(sync(err)) }
tild_status tild_engine_sk_accumulate(tild_engine* e, int slot, int t) { ENGINE_COPY(sk_accumulate(e->id, slot, t, err)) }
tild_status tild_engine_sk_average(tild_engine* e, int slot, float* sk, size_t n, int* count) {
    ENGINE_COPY(sk_average(e->id, slot, sk, n, count, err))
}
tild_status tild_engine_get_bins(tild_engine* e, int* base, size_t n) { ENGINE_COPY(get_bins(e->id, base, n, err)) }

tild_status tild_engine_profile(tild_engine* e, int nsteps, double* phase_ms, int nphase) {
    if (!e) return fail(TILD_ERR_ARG, "null engine");
    std::string err;
    int rc = e->batch->profile({e->id}, nsteps, phase_ms, nphase, err);
    return from_rc(rc, err, e);
}

// ------------------------------------------------------------------ batch
tild_status tild_batch_create(const tild_opts* opts, tild_batch** out) {
    if (!out) return fail(TILD_ERR_ARG, "null argument");
    *out = nullptr;
    tild_opts o{};
    if (opts) o = *opts;
    auto b = std::make_unique<tild_batch>();
    b->batch = std::make_unique<tild::Batch>(o.device);
    b->batch->set_spread(o.spread_mode, o.tile_target > 0 ? o.tile_target : 512, o.tile_by, o.tile_bz, o.tile_slack);
    b->batch->set_reorder(o.reorder_every == 0 ? -1 : (o.reorder_every < 0 ? 0 : o.reorder_every));
    b->batch->set_fft_tuning({o.fft_mid_elems, o.fft_last_lines, o.fft_c2r_rows, o.fft_plane, o.fft_last_dirs,
                                   o.fft_staged ? 1 : 0});
    b->batch->set_deposit_threads(o.deposit_threads);
    b->batch->set_autotune(o.autotune != 0);
    std::string err;
    int rc = b->batch->init(err);
    if (rc < 0) return from_rc(rc, err, nullptr);
    *out = b.release();
    return TILD_OK;
}

void tild_batch_destroy(tild_batch* b) { delete b; }

const char* tild_batch_last_error(const tild_batch* b) { return b ? b->err.c_str() : g_last_error.c_str(); }

#define BATCH_CALL(call)                                    \
    if (!b) return fail(TILD_ERR_ARG, "null batch");        \
    std::string err;                                        \
    int rc = b->batch->call;                                \
    if (rc < 0) { b->err = err; g_last_error = err; return (tild_status)(-rc); } \
    return TILD_OK;

tild_status tild_batch_add(tild_batch* b, const tild_system* s, uint32_t seed, int* replica) {
    if (!b || !s || !replica) return fail(TILD_ERR_ARG, "null argument");
    std::string err;
    int id = b->batch->add(s->sys, seed, err);
    if (id < 0) { b->err = err; g_last_error = err; return (tild_status)(-id); }
    *replica = id;
    return TILD_OK;
}

tild_status tild_batch_remove(tild_batch* b, int replica) {
    if (!b || !b->batch->rep(replica)) return fail(TILD_ERR_ARG, "no such replica");
    b->batch->remove(replica);
    return TILD_OK;
}

tild_status tild_batch_step(tild_batch* b, const int* replicas, int nrep, int nsteps, int use_graph) {
    if (!b || nsteps < 0 || nrep < 0 || (nrep > 0 && !replicas)) return fail(TILD_ERR_ARG, "bad argument");
    std::vector<int> ids(replicas, replicas + nrep);
    for (int id : ids) if (!b->batch->rep(id)) return fail(TILD_ERR_ARG, "no such replica");
    std::string err;
    int rc = b->batch->step(ids, nsteps, err, use_graph != 0);
    if (rc >= 0) rc = b->batch->sync(err);
    if (rc < 0) { b->err = err; g_last_error = err; return (tild_status)(-rc); }
    return TILD_OK;
}

tild_status tild_batch_properties(tild_batch* b, int replica, tild_props* out) {
    if (!out) return fail(TILD_ERR_ARG, "null argument");
    BATCH_CALL(properties(replica, out, err))
}
tild_status tild_batch_get_positions(tild_batch* b, int r, float* x, size_t n) { BATCH_CALL(get_positions(r, x, n, err)) }
tild_status tild_batch_set_positions(tild_batch* b, int r, const float* x, size_t n) { BATCH_CALL(set_positions(r, x, n, err)) }
tild_status tild_batch_get_density(tild_batch* b, int r, int t, float* rho, size_t n) { BATCH_CALL(get_density(r, t, rho, n, err)) }
tild_status tild_batch_replica_step(tild_batch* b, int r, int* step) {
    if (!b || !step || !b->batch->rep(r)) return fail(TILD_ERR_ARG, "no such replica");
    *step = b->batch->rep(r)->step;
    return TILD_OK;
}
// Block geometry of a replica (the tuner's pick, or the settings it was built with).
static tild_status geometry_of(tild::Replica* R, tild_geometry* out) {
    if (!R || !out) return fail(TILD_ERR_ARG, "no such replica");
    std::memset(out, 0, sizeof(*out));
    const tild::RepDev& h = R->h;
    out->tile_target = R->tune.tuned ? R->tune.g.tile_target : 0;
    out->tiles = h.spread_mode == 2 ? h.ntiles : 0;
    out->deposit_threads = R->dep_threads;
    out->plane = h.plane_fwd;
    out->last_lines = h.nl_last; out->last_dirs = h.last_dirs; out->mid_elems = h.mid_elems;
    out->c2r_rows = h.nr_c2r;
    out->last_reg = h.last_reg;
    out->grad_reg = h.grad_reg;
    out->plane_reg = h.plane_reg;
    out->plane_inv = h.plane_inv;
    out->tuned = R->tune.tuned ? 1 : 0;
    out->from_cache = R->tune.from_cache ? 1 : 0;
    out->step_ms = R->tune.step_ms;
    return TILD_OK;
}
tild_status tild_batch_geometry(tild_batch* b, int r, tild_geometry* out) {
    if (!b) return fail(TILD_ERR_ARG, "null batch");
    return geometry_of(b->batch->rep(r), out);
}
tild_status tild_engine_geometry(tild_engine* e, tild_geometry* out) {
    if (!e) return fail(TILD_ERR_ARG, "null engine");
    return geometry_of(e->batch->rep(e->id), out);
}

tild_status tild_batch_profile(tild_batch* b, const int* replicas, int nrep, int nsteps, double* ms, int nphase) {
    if (!b || nrep < 0 || (nrep > 0 && !replicas)) return fail(TILD_ERR_ARG, "bad argument");
    std::vector<int> ids(replicas, replicas + nrep);
    BATCH_CALL(profile(ids, nsteps, ms, nphase, err))
}

tild_status tild_engine_nlist(tild_engine* e, float rc, const int* sites, int nsites, int* cell_of, int* counts,
                              int* neigh, size_t cap, int* cell_max, int* ncells) {
    if (!e || nsites < 0 || (nsites > 0 && !sites)) return fail(TILD_ERR_ARG, "bad argument");
    std::string err;
    int rc_ = e->batch->nlist(e->id, rc, sites, nsites, cell_of, counts, neigh, cap, cell_max, ncells, err);
    return from_rc(rc_, err, e);
}

tild_status tild_engine_get_state(tild_engine* e, int* step, float* xo_disp, size_t n) {
    ENGINE_COPY(get_state(e->id, step, xo_disp, n, err))
}
tild_status tild_engine_set_state(tild_engine* e, int step, const float* xo_disp, size_t n) {
    ENGINE_COPY(set_state(e->id, step, xo_disp, n, err))
}

tild_status tild_engine_get_step(tild_engine* e, int* step) {
    if (!e || !step) return fail(TILD_ERR_ARG, "null argument");
    *step = e->batch->rep(e->id)->step;
    return TILD_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ job server
#include "server.hpp"

struct tild_server {
    std::unique_ptr<tild::Server> srv;
};

extern "C" {

tild_status tild_server_create(const tild_server_opts* opts, tild_server** out) {
    if (!out) return fail(TILD_ERR_ARG, "null argument");
    *out = nullptr;
    tild_server_opts o{};
    if (opts) o = *opts;
    if (o.chunk_steps < 0 || o.max_replicas < 0) return fail(TILD_ERR_ARG, "bad server options");
    auto s = std::make_unique<tild_server>();
    s->srv = std::make_unique<tild::Server>(o);
    std::string err;
    int rc = s->srv->start(err);
    if (rc < 0) return from_rc(rc, err, nullptr);
    *out = s.release();
    return TILD_OK;
}

void tild_server_destroy(tild_server* s) { delete s; }

#define SERVER_CALL(call)                                   \
    if (!s) return fail(TILD_ERR_ARG, "null server");       \
    std::string err;                                        \
    int rc = s->srv->call;                                  \
    if (rc < 0) return from_rc(rc, err, nullptr);           \
    return TILD_OK;

tild_status tild_server_submit(tild_server* s, const char* text, const char* base, uint32_t seed, int nsteps,
                               int* job) {
    if (!text || !job) return fail(TILD_ERR_ARG, "null argument");
    SERVER_CALL(submit(text, base ? base : "", seed, nsteps, job, err))
}
tild_status tild_server_poll(tild_server* s, int job, int* state) {
    if (!state) return fail(TILD_ERR_ARG, "null argument");
    if (!s) return fail(TILD_ERR_ARG, "null server");
    if (s->srv->poll(job, state) < 0) return fail(TILD_ERR_ARG, "no such job");
    return TILD_OK;
}
tild_status tild_server_wait(tild_server* s, int job, double timeout_s) { SERVER_CALL(wait(job, timeout_s, err)) }
tild_status tild_server_drain(tild_server* s, double timeout_s) { SERVER_CALL(drain(timeout_s, err)) }
tild_status tild_server_result(tild_server* s, int job, tild_props* props, float* x, size_t n, double* lat) {
    SERVER_CALL(result(job, props, x, n, lat, err))
}
tild_status tild_server_stats_get(tild_server* s, tild_server_stats* out) {
    if (!s || !out) return fail(TILD_ERR_ARG, "null argument");
    s->srv->stats(out);
    return TILD_OK;
}

}  // extern "C"
