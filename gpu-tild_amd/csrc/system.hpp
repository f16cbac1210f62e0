// system.hpp -- host-side model of a TILD deck: the reference's PairStyle / Group /
// Integrator vocabulary (pair_style.h:9-37, group.h:11-27, integrator.h:13-24) held
// by value in one re-entrant object instead of process globals (globals.h:47-184).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace tild {

constexpr int kMaxBonds = 3;          // MAX_BONDS, globals.h:26
constexpr double kPi = 3.14159265358979323846;

enum class StyleKind { Gaussian = 0, Erf = 1, GaussianErf = 2 };

// Base of the k-space pair styles (pair_style.h:9-37).  A style couples the
// densities of type1 and type2 through a potential known in Fourier space.
class PairStyle {
public:
    int type1 = 0, type2 = 0;
    double sigma_squared = 0.0;      // squared at parse time, read_input.cu:329-331
    double Rp = 0.0;
    double initial_prefactor = 0.0, final_prefactor = 0.0;
    bool ramp = false;
    double energy = 0.0;
    double total_vir[6] = {0, 0, 0, 0, 0, 0};

    virtual ~PairStyle() = default;
    virtual StyleKind kind() const = 0;
    virtual const char* keyword() const = 0;
    // u(k) as a function of k^2, evaluated in FP64 with amplitude `A`.
    virtual double kspace(double k2, double A) const = 0;
};

// u(k) = A exp(-k^2 sigma^2 / 2)   (pair_style_gaussian.cu:72-90)
class Gaussian : public PairStyle {
public:
    void Initialize_Gaussian(float Ao, float sigma_squared_in, int typ_A, int typ_B);
    StyleKind kind() const override { return StyleKind::Gaussian; }
    const char* keyword() const override { return "gaussian"; }
    double kspace(double k2, double A) const override;
};

// u(k) = A exp(-k^2 sigma^2/2) s(k)^2, s(k) = 4 pi (sin kRp - kRp cos kRp)/k^3
// (pair_style_erf.cu:122-157)
class Erf : public PairStyle {
public:
    void Initialize_Erf(float Ao, float sigma2, float Rp_in, int typ_A, int typ_B);
    StyleKind kind() const override { return StyleKind::Erf; }
    const char* keyword() const override { return "erf"; }
    double kspace(double k2, double A) const override;
};

// u(k) = A exp(-k^2 sigma^2/2) s(k)   (pair_style_gaussian_erf.cu:82-115)
class GaussianErf : public PairStyle {
public:
    void Initialize_Gaussian_Erf(float Ao, float sigma2, float Rp_in, int typ_A, int typ_B);
    StyleKind kind() const override { return StyleKind::GaussianErf; }
    const char* keyword() const override { return "gaussian_erf"; }
    double kspace(double k2, double A) const override;
};

// Smeared charges (pair_style_charges.h, device_array_utils.cu:199-241).
struct Charges {
    bool enabled = false;
    float bjerrum_length = 0.0f;
    float smearing_length = 0.0f;
    // potential kernel 4 pi lB / k^2 exp(-k^2/(2 a^2)), zero at k = 0
    double potential(double k2) const;
};

// Static field coupling (pair_style_fieldphases.h): type1 feels +w(r), type2 -w(r).
// phase 0 lamellar along dir, 2 cylinders along dir, 3 gyroid (3D).
struct FieldPhase {
    int type1 = 0, type2 = 0;
    float initial_prefactor = 0.0f, final_prefactor = 0.0f;
    int phase = 0, dir = 0, n_periods = 1;
    bool ramp = false;
};

struct Group {                        // group.h:11-27
    std::string name, command_line;
    std::vector<int> index;
    int nsites() const { return (int)index.size(); }
};

struct Integrator {                   // integrator.h:13-24
    std::string name;                 // "EM" or "GJF"
    std::string group_name;
    int group_index = -1;
};

// s(k) of a sphere of radius Rp, with the k -> 0 limit 4 pi Rp^3 / 3.
double sphere_kspace(double k2, double Rp);

// `compute avg_sk T [freq F] [wait W]` (Compute.cu:20-75): the running average of
// rho_T(k) rho_T(k) / M, evaluated every `freq` steps once step > wait.
struct SkCompute {
    int type = 0;        // 0-based
    int freq = 100, wait = 0;
};

struct System {
    // deck (read_input.cu:536-562 defaults)
    int Dim = 2;
    int Nx[3] = {35, 35, 0};
    int pmeorder = 1;
    float delt = 0.005f;
    float noise_mag = 0.0f;           // sqrtf(2 delt), read_input.cu:93
    int max_steps = 100000, log_freq = 2500, grid_freq = 0, traj_freq = 0, bin_freq = 5000,
        struc_freq = 0, skip_steps = 20000, threads = 512;
    uint32_t rand_seed = 0;
    bool seed_given = false;
    std::string dump_name = "traj.lammpstrj";
    Charges charges;

    // data file
    int ns = 0, ntypes = 0, nbond_types = 0, n_total_bonds = 0;
    float L[3] = {1, 1, 1}, Lh[3] = {0.5f, 0.5f, 0.5f}, V = 1.0f;
    std::vector<float> x;             // ns*Dim
    std::vector<int> tp, molecID;     // 0-based
    std::vector<float> q;             // ns (zero when uncharged)
    std::vector<float> mass, Diff;    // per type
    std::vector<int> n_bonds, bonded_to, bond_type;   // ns, ns*3, ns*3
    std::vector<float> bond_k, bond_req;

    std::vector<std::unique_ptr<Gaussian>> Gausses;
    std::vector<std::unique_ptr<Erf>> Erfs;
    std::vector<std::unique_ptr<GaussianErf>> GaussErfs;
    std::vector<FieldPhase> Fields;
    std::vector<Group> Groups;
    std::vector<Integrator> Integrators;
    std::vector<SkCompute> computes;
    std::vector<std::string> notes;   // accepted-but-ignored keywords

    // derived, initialize.cu:32-48
    int M = 0, grid_per_partic = 1, n_P_comps = 3;
    float dx[3] = {1, 1, 1}, gvol = 1.0f;

    // All k-space styles in force/energy order: Gaussians, Erfs, GaussianErfs
    // (forces.cu:107-126, calc_properties.cu:62-77).
    std::vector<const PairStyle*> styles() const;
    int derive(std::string& err);
};

// Force table fr[M*Dim] of one field phase at the grid nodes, in FP32 as
// init_device_fieldphase computes it (pair_style_fieldphases.cu:89-185).
void fieldphase_force_table(const System& S, const FieldPhase& F, std::vector<float>& fr);

// Parse `text` (reference grammar) and its data file(s).  0 on success.
// Returns 1 for malformed input (TILD_ERR_PARSE), 2 for non-runnable decks
// (TILD_ERR_CONFIG), 8 for features outside the engine (TILD_ERR_UNSUPPORTED).
int load_system(const std::string& text, const std::string& base_dir, System& sys,
                std::string& err);

}  // namespace tild
