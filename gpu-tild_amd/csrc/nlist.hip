// nlist.hip -- cell-list neighbour lists of a site group (NList, nlist.cu:13-228).
//
// Cells: ncells_d = int(L_d / rc), Lcell_d = L_d / ncells_d, cell = int(x_d / Lcell_d)
// in FP32 (nlist.cu:36-40, 217), x fastest; a site exactly at x == L goes to cell 0
// (the reference indexes one cell past the end there).  Each member scans the 27
// (2D: 9) surrounding cells in the reference's order and keeps every member j with
// FP32 min-image r^2 < rc^2, itself included (nlist.cu:118-203); with fewer than 3
// cells along an axis a cell is visited more than once and its sites repeat, as in
// the reference.  The reference fills lists in atomic order; here every list is
// sorted ascending by site id, so the result is deterministic.  Lists are indexed by
// site id whatever the slot order of the replica.
#include "device.hpp"
#include "launch.hpp"

namespace tild {

__device__ inline int nl_axis_cell(float x, float Lc, int nc) {
    int c = (int)(x / Lc);
    if (c >= nc) c -= nc;
    if (c < 0) c += nc;
    return c;
}

__global__ void __launch_bounds__(256) k_nl_cells(const RepDev* __restrict__ reps, int rep, NlistArgs a) {
    const RepDev& r = reps[rep];
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= r.ns) return;
    const int site = r.pid[p];
    if (!a.member[site]) { a.cell_of[site] = -1; a.counts[site] = 0; return; }
    const int D = r.dim;
    const float* x = r.X[r.step & 1];
    int cid[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) cid[d] = nl_axis_cell(x[p * D + d], a.Lc[d], a.nc[d]);
    const int cn = D == 2 ? cid[0] + a.nc[0] * cid[1] : cid[0] + (cid[1] + cid[2] * a.nc[1]) * a.nc[0];
    a.cell_of[site] = cn;
    const int k = atomicAdd(a.cell_count + cn, 1);
    if (k < a.cmax) a.cell_slots[(size_t)cn * a.cmax + k] = p;
    else atomicOr(a.overflow, 1);
}

__global__ void __launch_bounds__(256) k_nl_build(const RepDev* __restrict__ reps, int rep, NlistArgs a) {
    const RepDev& r = reps[rep];
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= r.ns) return;
    const int site = r.pid[p];
    if (!a.member[site]) return;
    const int D = r.dim;
    const float* x = r.X[r.step & 1];
    const int cn = a.cell_of[site];
    int nn[3];
    if (D == 2) { nn[1] = cn / a.nc[0]; nn[0] = cn % a.nc[0]; nn[2] = 1; }
    else { nn[2] = cn / (a.nc[1] * a.nc[0]); const int rr = cn - nn[2] * a.nc[0] * a.nc[1]; nn[1] = rr / a.nc[0]; nn[0] = rr % a.nc[0]; }
    const float rc2 = a.rc * a.rc;
    float xi[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < D; ++d) xi[d] = x[p * D + d];
    int* out = a.neigh + (size_t)site * a.cap;
    int cnt = 0;
    for (int ix = nn[0] - 1; ix <= nn[0] + 1; ++ix) {
        const int nix = ix < 0 ? a.nc[0] - 1 : (ix == a.nc[0] ? 0 : ix);
        for (int iy = nn[1] - 1; iy <= nn[1] + 1; ++iy) {
            const int niy = iy < 0 ? a.nc[1] - 1 : (iy == a.nc[1] ? 0 : iy);
            for (int iz = nn[2] - 1; iz <= nn[2] + 1; ++iz) {
                const int niz = D == 2 ? 0 : (iz < 0 ? a.nc[2] - 1 : (iz == a.nc[2] ? 0 : iz));
                const int ncn = D == 2 ? nix + a.nc[0] * niy : nix + (niy + niz * a.nc[1]) * a.nc[0];
                const int m = min(a.cell_count[ncn], a.cmax);
                for (int k = 0; k < m; ++k) {
                    const int q = a.cell_slots[(size_t)ncn * a.cmax + k];
                    float r2 = 0.f;
                    for (int d = 0; d < D; ++d) {
                        float dr = xi[d] - x[q * D + d];
                        if (dr < -r.Lh[d]) dr += r.L[d];
                        else if (dr > r.Lh[d]) dr -= r.L[d];
                        r2 += dr * dr;
                    }
                    if (r2 < rc2) {
                        if (cnt < a.cap) out[cnt] = r.pid[q];
                        ++cnt;
                    }
                }
                if (D == 2) break;
            }
        }
    }
    a.counts[site] = cnt;
    if (cnt > a.cap) { atomicOr(a.overflow, 2); return; }
    // insertion sort of the list (short: at most cap entries)
    for (int i = 1; i < cnt; ++i) {
        const int v = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; --j; }
        out[j + 1] = v;
    }
}

#define CHECK_LAUNCH() do { hipError_t e__ = hipGetLastError(); if (e__ != hipSuccess) return e__; } while (0)

hipError_t launch_nlist(hipStream_t s, const RepDev* reps, int rep, int ns, const NlistArgs& a) {
    const int nb = (ns + 255) / 256;
    hipLaunchKernelGGL(k_nl_cells, dim3(nb), dim3(256), 0, s, reps, rep, a);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_nl_build, dim3(nb), dim3(256), 0, s, reps, rep, a);
    CHECK_LAUNCH();
    return hipSuccess;
}

}  // namespace tild
