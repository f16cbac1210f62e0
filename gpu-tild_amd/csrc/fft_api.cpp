// fft_api.cpp -- tild_fft_r2c / tild_fft_c2r: the engine's Stockham passes on caller
// device buffers (used to cross-check the transforms against rocFFT).
#include <cstring>
#include <mutex>
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

tild_status arg_error(tild_status s, const std::string& why) {
    tild::set_last_error(why);
    return s;
}

// Geometry and twiddles of a standalone transform: axis a's twiddles at tw + off[a]
// (tw has room for Nx[0] + Nx[1] + Nx[2] values); h.ax[a].tw is left for the caller.
tild_status plan_geom(int dim, const int* Nx, tild::RepDev& h, float2* tw, int off[3]) {
    using namespace tild;
    if ((dim != 2 && dim != 3) || !Nx) return arg_error(TILD_ERR_ARG, "fft: dim must be 2 or 3");
    std::memset(&h, 0, sizeof(h));
    h.dim = dim; h.Nx = Nx[0]; h.Ny = Nx[1]; h.Nz = dim == 3 ? Nx[2] : 1;
    if (h.Nx < 1 || h.Ny < 1 || h.Nz < 1) return arg_error(TILD_ERR_ARG, "fft: grid sizes must be positive");
    h.Nxh = h.Nx / 2 + 1;
    h.M = h.Nx * h.Ny * h.Nz; h.Mh = h.Nxh * h.Ny * h.Nz;
    int n = 0;
    for (int a = 0; a < dim; ++a) {
        off[a] = n;
        int ntw = 0;
        if (const char* why = plan_axis(Nx[a], h.ax[a], tw + n, &ntw)) return arg_error(TILD_ERR_UNSUPPORTED, why);
        n += ntw;
    }
    return TILD_OK;
}

tild_status upload_geom(tild::RepDev& h, const float2* tw, const int off[3], int ntw, tild::RepDev* d_h,
                        float2* d_tw, hipStream_t s) {
    for (int a = 0; a < h.dim; ++a) h.ax[a].tw = d_tw + off[a];
    TRY(hipMemcpyAsync(d_tw, tw, (size_t)ntw * sizeof(float2), hipMemcpyHostToDevice, s));
    TRY(hipMemcpyAsync(d_h, &h, sizeof(h), hipMemcpyHostToDevice, s));
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

// ---------------------------------------------------------------- inverse transform
// Block counts of the three passes of an inverse transform (z pass in 3D, y pass, x
// C2R) and the complex values of its work copy; h.nr_c2r is set.
struct C2rShape {
    int nblk[3];
    size_t work;
};

tild_status c2r_shape(tild::RepDev& h, int batch, C2rShape& sh) {
    using namespace tild;
    const int rows = h.M / h.Nx, nc = (batch + 1) / 2;
    // one block of the x pass holds nc complex lines of every row it transforms
    if ((long long)nc * h.Nx > kFftBudget) return arg_error(TILD_ERR_UNSUPPORTED, "fft_c2r: batch x Nx too large for one block");
    h.nr_c2r = std::max(1, std::min(kFftBudget / (nc * h.Nx), rows));
    const MidGeom gz = mid_geom(h, 2, batch), gy = mid_geom(h, 1, batch);
    sh.nblk[0] = h.dim == 3 ? gz.nouter * gz.nchunk : 0;
    sh.nblk[1] = gy.nouter * gy.nchunk;
    sh.nblk[2] = (rows + h.nr_c2r - 1) / h.nr_c2r;
    sh.work = (size_t)batch * h.Mh;
    return TILD_OK;
}

// Enqueue z (3D) and y inverse passes in place on `work` (already holding the input
// spectra) and the x C2R into d_out; maps[k] are identity work maps of at least
// sh.nblk[k] blocks.
tild_status c2r_enqueue(const tild::RepDev& h, const C2rShape& sh, tild::RepDev* d_h, tild::WorkItem* const maps[3],
                        int batch, float2* work, float* d_out, hipStream_t s) {
    using namespace tild;
    if (h.dim == 3) {
        const MidGeom g = mid_geom(h, 2, batch);
        LaunchCtx c{d_h, maps[0], sh.nblk[0], lds_bytes(g.N, g.nl), s};
        TRY(launch_mid(c, 2, 0, batch, 1, work));
    }
    {
        const MidGeom g = mid_geom(h, 1, batch);
        LaunchCtx c{d_h, maps[1], sh.nblk[1], lds_bytes(g.N, g.nl), s};
        TRY(launch_mid(c, 1, 0, batch, 1, work));
    }
    const int nc = (batch + 1) / 2;
    LaunchCtx c{d_h, maps[2], sh.nblk[2], lds_bytes(h.Nx, h.nr_c2r * nc), s};
    TRY(launch_x_c2r(c, false, batch, work, d_out));
    return TILD_OK;
}

}  // namespace

// ---------------------------------------------------------------- no-allocation inverse
// Device scratch of tild_fft_c2r_noalloc: descriptor, twiddles, one identity work map
// and the work copy of the spectra, at 256-byte aligned offsets.  Each device holds its
// own copy for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the symbol's name is a host-side
// stand-in, not a device address).  Calls that overlap on one device share it and race.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

__global__ void k_identity_map(tild::WorkItem* m, int n) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) m[b] = tild::WorkItem{0, b};
}

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch`, looked up once per device.
tild_status scratch_address(unsigned char** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
        return arg_error(TILD_ERR_DEVICE, "fft_c2r_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return arg_error(TILD_ERR_DEVICE, "fft_c2r_noalloc: hipGetSymbolAddress(scratch) failed");
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return TILD_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" tild_status tild_fft_c2r_noalloc(int dim, const int* Nx, int batch, const float* d_in, float* d_out,
                                            void* stream) {
    using namespace tild;
    if (batch < 1 || !d_in || !d_out) return arg_error(TILD_ERR_ARG, "fft_c2r_noalloc: bad argument");
    hipStream_t s = (hipStream_t)stream;
    RepDev h;
    float2 tw[3 * 2048];
    int off[3] = {0, 0, 0};
    if (tild_status st = plan_geom(dim, Nx, h, tw, off)) return st;
    C2rShape sh;
    if (tild_status st = c2r_shape(h, batch, sh)) return st;
    const int ntw = off[dim - 1] + std::max(Nx[dim - 1] - 1, 1);
    const int nmap = std::max(sh.nblk[0], std::max(sh.nblk[1], sh.nblk[2]));
    // layout: descriptor | twiddles | identity map | work copy
    const size_t o_tw = align256(sizeof(RepDev));
    const size_t o_map = o_tw + align256((size_t)ntw * sizeof(float2));
    const size_t o_work = o_map + align256((size_t)nmap * sizeof(WorkItem));
    const size_t need = o_work + sh.work * sizeof(float2);
    if (need > kNoAllocScratchBytes)
        return arg_error(TILD_ERR_ARG, "fft_c2r_noalloc: transform needs " + std::to_string(need) +
                                           " bytes of scratch, more than the " +
                                           std::to_string(kNoAllocScratchBytes) + " available; use tild_fft_c2r");
    unsigned char* base;
    if (tild_status st = scratch_address(&base)) return st;
    RepDev* d_h = (RepDev*)base;
    float2* d_tw = (float2*)(base + o_tw);
    WorkItem* map = (WorkItem*)(base + o_map);
    float2* work = (float2*)(base + o_work);
    if (tild_status st = upload_geom(h, tw, off, ntw, d_h, d_tw, s)) return st;
    hipLaunchKernelGGL(k_identity_map, dim3((nmap + 255) / 256), dim3(256), 0, s, map, nmap);
    TRY(hipGetLastError());
    TRY(hipMemcpyAsync(work, d_in, sh.work * sizeof(float2), hipMemcpyDeviceToDevice, s));
    WorkItem* maps[3] = {map, map, map};
    if (tild_status st = c2r_enqueue(h, sh, d_h, maps, batch, work, d_out, s)) return st;
    // the host arrays staged above live on this frame: wait before returning
    TRY(hipStreamSynchronize(s));
    return TILD_OK;
}

extern "C" tild_status tild_fft_r2c(int dim, const int* Nx, int batch, const float* d_in, float* d_out, void* stream) {
    using namespace tild;
    if (batch < 1 || !d_in || !d_out) return arg_error(TILD_ERR_ARG, "fft_r2c: bad argument");
    hipStream_t s = (hipStream_t)stream;
    TmpDev tmp;
    RepDev h;
    float2 tw[3 * 2048];
    int off[3] = {0, 0, 0};
    if (tild_status st = plan_geom(dim, Nx, h, tw, off)) return st;
    float2* d_tw;
    const int ntw = off[dim - 1] + std::max(Nx[dim - 1] - 1, 1);
    TRY(tmp.alloc(&d_tw, ntw));
    const int rows = h.M / h.Nx, npairs = (batch * rows + 1) / 2;
    h.nl_x = std::max(1, std::min(kFftBudget / h.Nx, npairs));
    RepDev* d_h;
    TRY(tmp.alloc(&d_h, 1));
    if (tild_status st = upload_geom(h, tw, off, ntw, d_h, d_tw, s)) return st;
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
    if (batch < 1 || !d_in || !d_out) return arg_error(TILD_ERR_ARG, "fft_c2r: bad argument");
    hipStream_t s = (hipStream_t)stream;
    TmpDev tmp;
    RepDev h;
    float2 tw[3 * 2048];
    int off[3] = {0, 0, 0};
    if (tild_status st = plan_geom(dim, Nx, h, tw, off)) return st;
    C2rShape sh;
    if (tild_status st = c2r_shape(h, batch, sh)) return st;
    const int ntw = off[dim - 1] + std::max(Nx[dim - 1] - 1, 1);
    float2* d_tw;
    RepDev* d_h;
    TRY(tmp.alloc(&d_tw, ntw));
    TRY(tmp.alloc(&d_h, 1));
    if (tild_status st = upload_geom(h, tw, off, ntw, d_h, d_tw, s)) return st;
    WorkItem* maps[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < 3; ++k)
        if (sh.nblk[k] > 0) if (tild_status st = run_map(sh.nblk[k], &maps[k], tmp, s)) return st;
    float2* work;
    TRY(tmp.alloc(&work, sh.work));
    TRY(hipMemcpyAsync(work, d_in, sh.work * sizeof(float2), hipMemcpyDeviceToDevice, s));
    if (tild_status st = c2r_enqueue(h, sh, d_h, maps, batch, work, d_out, s)) return st;
    TRY(hipStreamSynchronize(s));
    return TILD_OK;
}
