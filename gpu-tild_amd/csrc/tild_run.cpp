// tild_run.cpp -- command-line driver over the C ABI: `tild-run -in input`.
//
// Runs a deck the way the reference executable does (main.cu:36-227): forces and
// properties at step 0, then max_steps steps with the trajectory, grid, binary and
// log outputs at their frequencies, in the reference's file formats:
//   traj.lammpstrj (or traj_name)   write_lammps_traj, config_utils.cu:366-402
//   data.dat + stdout log lines     main.cu:60-73, 230-293
//   rho<t>.dat                      write_grid_data, io_utils.cu:7-31
//   grid_densities.bin, positions.bin  init_binary_output / write_binary, io_utils.cu:85-134
//   avg_sk_<id>.dat                 Compute::writeResults / write_kspace_data,
//                                   Compute.cu:121-137, io_utils.cu:60-83
// Between output steps the engine advances in one call, so no host round trip
// happens on steps that write nothing.  Extra options: -device N, -seed S (overrides
// rand_seed).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "tild.h"

namespace {

struct Run {
    tild_system* sys = nullptr;
    tild_engine* eng = nullptr;
    int ns = 0, dim = 3, ntypes = 0, M = 0, nP = 6, do_charges = 0;
    int Nx[3] = {1, 1, 1};
    float L[3] = {1, 1, 1}, dx[3] = {1, 1, 1};
    double V = 1;
    int max_steps = 0, log_freq = 0, traj_freq = 0, grid_freq = 0, bin_freq = 0;
    int ng = 0, ne = 0, nge = 0, nfp = 0;
    std::vector<int> tp, mol;
    std::vector<float> q, x, rho;
    std::string dump_name;
    std::ofstream dout;
    struct Comp { int type, freq, wait; };
    std::vector<Comp> computes;
    std::vector<float> sk;   // 2*M
};

[[noreturn]] void die(const std::string& what) {
    std::cerr << "tild-run: " << what << ": " << tild_last_error() << std::endl;
    std::exit(1);
}

double info(const Run& R, const char* key) {
    double v = 0;
    if (tild_system_info(R.sys, key, &v) != TILD_OK) die(std::string("system info ") + key);
    return v;
}

void collect_x(Run& R) {
    if (tild_engine_get_positions(R.eng, R.x.data(), R.x.size()) != TILD_OK) die("get_positions");
}

void collect_rho(Run& R) {
    for (int t = 0; t < R.ntypes; ++t)
        if (tild_engine_get_density(R.eng, t, R.rho.data() + (size_t)t * R.M, R.M) != TILD_OK) die("get_density");
}

void write_traj(Run& R, int step) {
    FILE* f = std::fopen(R.dump_name.c_str(), step == 0 ? "w" : "a");
    if (!f) die("cannot open " + R.dump_name);
    std::fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n", step, R.ns);
    std::fprintf(f, "ITEM: BOX BOUNDS pp pp pp\n");
    std::fprintf(f, "%f %f\n%f %f\n%f %f\n", 0.0, (double)R.L[0], 0.0, (double)R.L[1],
                 R.dim == 3 ? 0.0 : 1.0, R.dim == 3 ? (double)R.L[2] : 1.0);
    std::fprintf(f, R.do_charges ? "ITEM: ATOMS id type mol x y z q\n" : "ITEM: ATOMS id type mol x y z\n");
    for (int i = 0; i < R.ns; ++i) {
        std::fprintf(f, "%d %d %d  ", i + 1, R.tp[i] + 1, R.mol[i] + 1);
        for (int d = 0; d < R.dim; ++d) std::fprintf(f, "%f ", (double)R.x[(size_t)i * R.dim + d]);
        for (int d = R.dim; d < 3; ++d) std::fprintf(f, "%f", 0.0);
        if (R.do_charges) std::fprintf(f, " %f", (double)R.q[i]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

// write_kspace_data (io_utils.cu:60-83): every wavevector but k = 0, its components
// (get_k, array_utils.cu:17-41: FP32, 2 pi n / L folded at N / 2), then |S|, |k|,
// Re S and Im S; 2D grids get a blank line after each x row.
void write_kspace(const Run& R, const char* name, const float* sk) {
    FILE* f = std::fopen(name, "w");
    if (!f) die(std::string("cannot open ") + name);
    const float pi2 = 6.2831853071f;
    for (int i = 1; i < R.M; ++i) {
        int n[3] = {i % R.Nx[0], (i / R.Nx[0]) % R.Nx[1], i / (R.Nx[0] * R.Nx[1])};
        float kv[3], k2 = 0.f;
        for (int d = 0; d < R.dim; ++d) {
            kv[d] = (double)n[d] < (double)R.Nx[d] / 2.0 ? pi2 * (float)n[d] / R.L[d]
                                                          : pi2 * (float)(n[d] - R.Nx[d]) / R.L[d];
            k2 += kv[d] * kv[d];
        }
        for (int d = 0; d < R.dim; ++d) std::fprintf(f, "%f ", (double)kv[d]);
        const float re = sk[2 * i], im = sk[2 * i + 1];
        std::fprintf(f, "%1.5e %1.5e %1.5e %1.5e\n", (double)std::hypot(re, im), (double)std::sqrt(k2), (double)re,
                     (double)im);
        if (R.dim == 2 && n[0] == R.Nx[0] - 1) std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void write_grid(const Run& R, const char* name, const float* dat) {
    FILE* f = std::fopen(name, "w");
    if (!f) die(std::string("cannot open ") + name);
    for (int i = 0; i < R.M; ++i) {
        int n[3] = {i % R.Nx[0], (i / R.Nx[0]) % R.Nx[1], i / (R.Nx[0] * R.Nx[1])};
        for (int d = 0; d < R.dim; ++d) std::fprintf(f, "%f ", (double)(R.dx[d] * (float)n[d]));
        std::fprintf(f, "%1.8e \n", (double)dat[i]);
        if (R.dim == 2 && n[0] == R.Nx[0] - 1) std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void init_binary(const Run& R) {
    FILE* f = std::fopen("grid_densities.bin", "wb");
    if (!f) die("cannot open grid_densities.bin");
    std::fwrite(&R.dim, sizeof(int), 1, f);
    std::fwrite(R.Nx, sizeof(int), R.dim, f);
    std::fwrite(R.L, sizeof(float), R.dim, f);
    std::fwrite(&R.ntypes, sizeof(int), 1, f);
    std::fclose(f);
    f = std::fopen("positions.bin", "wb");
    if (!f) die("cannot open positions.bin");
    std::fwrite(&R.ns, sizeof(int), 1, f);
    std::fwrite(&R.dim, sizeof(int), 1, f);
    std::fwrite(R.L, sizeof(float), 3, f);
    std::fwrite(R.tp.data(), sizeof(int), R.ns, f);
    std::fwrite(R.mol.data(), sizeof(int), R.ns, f);
    std::fwrite(&R.do_charges, sizeof(int), 1, f);
    if (R.do_charges) std::fwrite(R.q.data(), sizeof(float), R.ns, f);
    std::fclose(f);
}

void write_binary(const Run& R) {
    FILE* f = std::fopen("grid_densities.bin", "ab");
    if (!f) die("cannot append to grid_densities.bin");
    std::fwrite(R.rho.data(), sizeof(float), (size_t)R.M * R.ntypes, f);
    std::fclose(f);
    f = std::fopen("positions.bin", "ab");
    if (!f) die("cannot append to positions.bin");
    std::fwrite(R.x.data(), sizeof(float), R.x.size(), f);
    std::fclose(f);
}

// Log line on stdout and data.dat; returns 1 when a style energy is NaN.
int print_timestep(Run& R, int step, const tild_props& p) {
    int die_flag = 0;
    std::cout << "Step " << step << " of " << R.max_steps << " ";
    if (R.do_charges) std::cout << " Electrostatic Energy: " << (float)p.Uelec;
    std::cout << " U/V: " << (float)(p.Upe / R.V) << " Ubond: " << (float)p.Ubond << " Pdiags: " << (float)p.Ptens[0]
              << " " << (float)p.Ptens[1] << " ";
    if (R.dim == 3) std::cout << (float)p.Ptens[2] << " ";
    R.dout << step << " " << (float)p.Upe << " " << (float)p.Ubond << " ";
    for (int c = 0; c < R.nP; ++c) R.dout << (float)p.Ptens[c] << " ";
    int s = 0;
    const char* label[3] = {"Ugauss: ", "Uerf: ", "UGaussErf: "};
    const int count[3] = {R.ng, R.ne, R.nge};
    for (int k = 0; k < 3; ++k) {
        if (!count[k]) continue;
        std::cout << label[k];
        for (int j = 0; j < count[k]; ++j, ++s) {
            std::cout << (float)p.U_style[s] << " ";
            R.dout << (float)p.U_style[s] << " ";
            if (std::isnan(p.U_style[s])) die_flag = 1;
        }
    }
    // field phases: the reference prints their energy members, which nothing computes
    // (calc_properties.cu:62-77 sums only the k-space styles); they are reported as 0
    if (R.nfp) {
        std::cout << "UFieldPhase: ";
        for (int j = 0; j < R.nfp; ++j) { std::cout << 0 << " "; R.dout << 0 << " "; }
    }
    R.dout << std::endl;
    std::cout << std::endl;
    return die_flag;
}

int next_multiple(int step, int freq) { return freq > 0 ? (step / freq + 1) * freq : 1 << 30; }

}  // namespace

int main(int argc, char** argv) {
    const time_t t_in = std::time(nullptr);
    std::string input;
    int device = 0;
    long long seed = -1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-in" && i + 1 < argc) input = argv[++i];
        else if (a == "-device" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (a == "-seed" && i + 1 < argc) seed = std::atoll(argv[++i]);
        else { std::cerr << "usage: tild-run -in input [-device N] [-seed S]" << std::endl; return 2; }
    }
    if (input.empty()) { std::cerr << "usage: tild-run -in input [-device N] [-seed S]" << std::endl; return 2; }
    std::ifstream in(input);
    if (!in) { std::cerr << "tild-run: cannot open " << input << std::endl; return 1; }
    std::stringstream ss;
    ss << in.rdbuf();

    Run R;
    // data files resolve against the working directory, as in the reference
    if (tild_system_load(ss.str().c_str(), ".", &R.sys) != TILD_OK) die("input");
    R.ns = (int)info(R, "ns"); R.dim = (int)info(R, "dim"); R.ntypes = (int)info(R, "ntypes");
    R.M = (int)info(R, "M"); R.do_charges = (int)info(R, "do_charges");
    R.Nx[0] = (int)info(R, "Nx0"); R.Nx[1] = (int)info(R, "Nx1"); R.Nx[2] = (int)info(R, "Nx2");
    R.L[0] = (float)info(R, "L0"); R.L[1] = (float)info(R, "L1"); R.L[2] = (float)info(R, "L2");
    R.dx[0] = (float)info(R, "dx0"); R.dx[1] = (float)info(R, "dx1"); R.dx[2] = (float)info(R, "dx2");
    R.V = info(R, "V");
    R.nP = R.dim == 3 ? 6 : 3;
    R.max_steps = (int)info(R, "max_steps"); R.log_freq = (int)info(R, "log_freq");
    R.traj_freq = (int)info(R, "traj_freq"); R.grid_freq = (int)info(R, "grid_freq");
    R.bin_freq = (int)info(R, "binary_freq");
    R.ng = (int)info(R, "n_gaussians"); R.ne = (int)info(R, "n_erfs"); R.nge = (int)info(R, "n_gaussian_erfs");
    R.nfp = (int)info(R, "n_fieldphases");
    char name[4096];
    if (tild_system_info_text(R.sys, "dump_name", name, sizeof(name)) != TILD_OK) die("dump_name");
    R.dump_name = name;
    R.tp.resize(R.ns); R.mol.resize(R.ns); R.q.assign(R.ns, 0.f);
    R.x.resize((size_t)R.ns * R.dim);
    R.rho.resize((size_t)R.M * R.ntypes);
    if (tild_system_get_types(R.sys, R.tp.data(), R.ns) != TILD_OK) die("types");
    if (tild_system_get_molecules(R.sys, R.mol.data(), R.ns) != TILD_OK) die("molecules");
    if (R.do_charges && tild_system_get_charges(R.sys, R.q.data(), R.ns) != TILD_OK) die("charges");
    for (int i = 0, nc = (int)info(R, "n_computes"); i < nc; ++i) {
        Run::Comp c{};
        if (tild_system_compute(R.sys, i, &c.type, &c.freq, &c.wait) != TILD_OK) die("compute");
        R.computes.push_back(c);
    }
    R.sk.resize(2 * (size_t)R.M);

    tild_opts o{};
    o.device = device;
    if (seed >= 0) { o.use_seed = 1; o.seed = (uint32_t)seed; }
    if (tild_engine_create(R.sys, &o, &R.eng) != TILD_OK) die("engine");
    if (R.bin_freq != 0) init_binary(R);

    collect_x(R);
    write_traj(R, 0);
    R.dout.open("data.dat");
    R.dout << "# step Upe Ubond ";
    if (R.dim == 2) R.dout << "Pxx Pyy Pxy ";
    else R.dout << " Pxx Pyy Pzz Pxy Pxz Pyz ";
    for (int i = 0; i < R.ng; ++i) R.dout << "Ugauss[" << i << "] ";
    for (int i = 0; i < R.ne; ++i) R.dout << "Uerf[" << i << "] ";
    for (int i = 0; i < R.nge; ++i) R.dout << "UGuassErf[" << i << "] ";
    R.dout << std::endl;

    tild_props p{};
    if (tild_engine_forces(R.eng) != TILD_OK) die("forces");
    if (tild_engine_properties(R.eng, &p) != TILD_OK) die("properties");
    if (R.grid_freq > 0) {
        collect_rho(R);
        for (int t = 0; t < R.ntypes; ++t) {
            char nm[32];
            std::snprintf(nm, sizeof(nm), "rho%d.dat", t);
            write_grid(R, nm, R.rho.data() + (size_t)t * R.M);
        }
    }
    int die_flag = print_timestep(R, 0, p);
    const time_t t_init = std::time(nullptr);
    double t_io = 0;

    int step = 0;
    while (step < R.max_steps && !die_flag) {
        int next = R.max_steps;
        for (int f : {R.log_freq, R.traj_freq, R.grid_freq, R.bin_freq}) next = std::min(next, next_multiple(step, f));
        for (const Run::Comp& c : R.computes) next = std::min(next, next_multiple(step, c.freq));
        if (tild_engine_step(R.eng, next - step) != TILD_OK) die("step");
        step = next;
        // computes follow the integration (main.cu:123-131), on the densities the step
        // started from
        for (size_t i = 0; i < R.computes.size(); ++i) {
            const Run::Comp& c = R.computes[i];
            if (step > c.wait && step % c.freq == 0 && tild_engine_sk_accumulate(R.eng, (int)i, c.type) != TILD_OK)
                die("compute");
        }
        const time_t io0 = std::time(nullptr);
        if (R.traj_freq > 0 && step % R.traj_freq == 0) { collect_x(R); write_traj(R, step); }
        if (R.grid_freq > 0 && step % R.grid_freq == 0) {
            collect_rho(R);
            for (int t = 0; t < R.ntypes; ++t) {
                char nm[32];
                std::snprintf(nm, sizeof(nm), "rho%d.dat", t);
                write_grid(R, nm, R.rho.data() + (size_t)t * R.M);
            }
        }
        if (R.bin_freq > 0 && step % R.bin_freq == 0) { collect_rho(R); collect_x(R); write_binary(R); }
        if (R.log_freq > 0 && step % R.log_freq == 0) {
            if (tild_engine_properties(R.eng, &p) != TILD_OK) die("properties");
            die_flag = print_timestep(R, step, p);
            for (size_t i = 0; i < R.computes.size(); ++i) {   // main.cu:182-184
                if (step <= R.computes[i].wait) continue;
                if (tild_engine_sk_average(R.eng, (int)i, R.sk.data(), R.sk.size(), nullptr) != TILD_OK) die("compute");
                char nm[32];
                std::snprintf(nm, sizeof(nm), "avg_sk_%d.dat", (int)i);
                write_kspace(R, nm, R.sk.data());
            }
        }
        t_io += (double)(std::time(nullptr) - io0);
    }
    collect_x(R);
    write_traj(R, step);
    tild_engine_destroy(R.eng);
    tild_system_free(R.sys);

    const int dt = (int)(std::time(nullptr) - t_in), di = (int)(t_init - t_in), dio = (int)t_io;
    std::cout << "Total run time: " << dt / 60 << "m" << dt % 60 << "sec" << std::endl;
    std::cout << "Total init time: " << di / 60 << "m" << di % 60 << "sec" << std::endl;
    std::cout << "I/O + Comm time: " << dio / 60 << "m" << dio % 60 << "sec" << std::endl;
    return 0;
}
