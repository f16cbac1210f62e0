// fft_api.cpp -- tild_fft_r2c / tild_fft_c2r: the engine's Stockham passes on caller
// device buffers (used to cross-check the transforms against rocFFT).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tild.h"
#include "engine.hpp"

namespace {

struct TmpDev {
    std::vector<void*> p;
    ~TmpDev() { for (void* q : p) (void)hipFree(q); }
    template <class T> hipError_t alloc(T** ptr, size_t n) {
        hipError_t e = hipMalloc((void**)ptr, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) p.push_back((void*)*ptr);
        return e;
    }
};

#define TRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return TILD_ERR_DEVICE; } while (0)

// Descriptor with only geometry and plans filled in.
tild_status make_geom(int dim, const int* Nx, tild::RepDev& h, TmpDev& tmp, hipStream_t s) {
    using namespace tild;
    if ((dim != 2 && dim != 3) || !Nx) return TILD_ERR_ARG;
    std::memset(&h, 0, sizeof(h));
    h.dim = dim; h.Nx = Nx[0]; h.Ny = Nx[1]; h.Nz = dim == 3 ? Nx[2] : 1;
    if (h.Nx < 1 || h.Ny < 1 || h.Nz < 1) return TILD_ERR_ARG;
    h.Nxh = h.Nx / 2 + 1;
    h.M = h.Nx * h.Ny * h.Nz; h.Mh = h.Nxh * h.Ny * h.Nz;
    std::vector<float2> all;
    int off[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        std::vector<float2> tw;
        std::string err;
        if (make_axis_plan(Nx[a], h.ax[a], tw, err)) return TILD_ERR_UNSUPPORTED;
        off[a] = (int)all.size();
        all.insert(all.end(), tw.begin(), tw.end());
    }
    float2* d_tw;
    TRY(tmp.alloc(&d_tw, all.size()));
    TRY(hipMemcpyAsync(d_tw, all.data(), all.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    for (int a = 0; a < dim; ++a) h.ax[a].tw = d_tw + off[a];
    return TILD_OK;
}

tild_status run_map(int nblk, tild::WorkItem** dm, TmpDev& tmp, hipStream_t s) {
    std::vector<tild::WorkItem> v(nblk);
    for (int b = 0; b < nblk; ++b) v[b] = tild::WorkItem{0, b};
    TRY(tmp.alloc(dm, nblk));
    TRY(hipMemcpyAsync(*dm, v.data(), nblk * sizeof(tild::WorkItem), hipMemcpyHostToDevice, s));
    return TILD_OK;
}

tild_status mid_pass(tild::RepDev* d_h, const tild::RepDev& h, int axis, int nfb, int inverse, float2* buf,
                     TmpDev& tmp, hipStream_t s) {
    using namespace tild;
    MidGeom g = mid_geom(h, axis, nfb);
    WorkItem* dm;
    if (tild_status st = run_map(g.nouter * g.nchunk, &dm, tmp, s)) return st;
    LaunchCtx c{d_h, dm, g.nouter * g.nchunk, lds_bytes(g.N, g.nl), s};
    TRY(launch_mid(c, axis, 0, nfb, inverse, buf));
    return TILD_OK;
}

}  // namespace

extern "C" tild_status tild_fft_r2c(int dim, const int* Nx, int batch, const float* d_in, float* d_out, void* stream) {
    using namespace tild;
    if (batch < 1 || !d_in || !d_out) return TILD_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    TmpDev tmp;
    RepDev h;
    if (tild_status st = make_geom(dim, Nx, h, tmp, s)) return st;
    const int rows = h.M / h.Nx, npairs = (batch * rows + 1) / 2;
    h.nl_x = std::max(1, std::min(kFftBudget / h.Nx, npairs));
    RepDev* d_h;
    TRY(tmp.alloc(&d_h, 1));
    TRY(hipMemcpyAsync(d_h, &h, sizeof(h), hipMemcpyHostToDevice, s));
    WorkItem* dm;
    int nblk = (npairs + h.nl_x - 1) / h.nl_x;
    if (tild_status st = run_map(nblk, &dm, tmp, s)) return st;
    float2* out = (float2*)d_out;
    LaunchCtx c{d_h, dm, nblk, lds_bytes(h.Nx, h.nl_x), s};
    TRY(launch_x_r2c(c, false, batch * rows, d_in, out));
    if (tild_status st = mid_pass(d_h, h, 1, batch, 0, out, tmp, s)) return st;
    if (dim == 3) if (tild_status st = mid_pass(d_h, h, 2, batch, 0, out, tmp, s)) return st;
    TRY(hipStreamSynchronize(s));
    return TILD_OK;
}

extern "C" tild_status tild_fft_c2r(int dim, const int* Nx, int batch, const float* d_in, float* d_out, void* stream) {
    using namespace tild;
    if (batch < 1 || !d_in || !d_out) return TILD_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    TmpDev tmp;
    RepDev h;
    if (tild_status st = make_geom(dim, Nx, h, tmp, s)) return st;
    const int rows = h.M / h.Nx, nc = (batch + 1) / 2;
    if (nc * h.Nx > kFftBudget) {
        // transform in groups of lines that fit one block
        return TILD_ERR_UNSUPPORTED;
    }
    h.nr_c2r = std::max(1, std::min(kFftBudget / (nc * h.Nx), rows));
    RepDev* d_h;
    TRY(tmp.alloc(&d_h, 1));
    TRY(hipMemcpyAsync(d_h, &h, sizeof(h), hipMemcpyHostToDevice, s));
    float2* work;
    TRY(tmp.alloc(&work, (size_t)batch * h.Mh));
    TRY(hipMemcpyAsync(work, d_in, (size_t)batch * h.Mh * sizeof(float2), hipMemcpyDeviceToDevice, s));
    if (dim == 3) if (tild_status st = mid_pass(d_h, h, 2, batch, 1, work, tmp, s)) return st;
    if (tild_status st = mid_pass(d_h, h, 1, batch, 1, work, tmp, s)) return st;
    WorkItem* dm;
    int nblk = (rows + h.nr_c2r - 1) / h.nr_c2r;
    if (tild_status st = run_map(nblk, &dm, tmp, s)) return st;
    LaunchCtx c{d_h, dm, nblk, lds_bytes(h.Nx, h.nr_c2r * nc), s};
    TRY(launch_x_c2r(c, false, batch, work, d_out));
    TRY(hipStreamSynchronize(s));
    return TILD_OK;
}
