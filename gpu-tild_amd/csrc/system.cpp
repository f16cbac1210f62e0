// system.cpp -- input-deck grammar and data-file reader of the TILD engine.
//
// The grammar is the reference's (read_input.cu:23-523): keywords are read word by
// word from each non-blank line whose first character is not '#'; style blocks
// ("n_gaussians K" etc.) consume the K following lines; order rules for read_data,
// charges, bond, diffusivity, group and integrator are enforced as in the reference.
// The data file is the reference's LAMMPS-like format (config_utils.cu:5-364), read
// here by section headers rather than fixed line offsets; numeric conversion follows
// the reference: plain decks parse coordinates in double and round to float
// (config_utils.cu:34-84), charged decks parse floats directly (config_utils.cu:214-267).
#include "system.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tild {

static constexpr int kParse = 1, kConfig = 2, kUnsupported = 8;

// ---------------------------------------------------------------- pair styles
void Gaussian::Initialize_Gaussian(float Ao, float s2, int tA, int tB) {
    initial_prefactor = final_prefactor = Ao;
    sigma_squared = s2;
    type1 = tA; type2 = tB;
}
double Gaussian::kspace(double k2, double A) const { return A * std::exp(-k2 * sigma_squared / 2.0); }

double sphere_kspace(double k2, double Rp) {
    if (k2 == 0.0) return 4.0 * kPi * Rp * Rp * Rp / 3.0;
    double k = std::sqrt(k2), x = Rp * k;
    if (x < 1e-2) {
        // series of 4 pi (sin x - x cos x)/k^3 = (4 pi Rp^3/3)(1 - x^2/10 + x^4/280 - x^6/15120)
        double x2 = x * x;
        return 4.0 * kPi * Rp * Rp * Rp / 3.0 * (1.0 - x2 / 10.0 + x2 * x2 / 280.0 - x2 * x2 * x2 / 15120.0);
    }
    return 4.0 * kPi * (std::sin(x) - x * std::cos(x)) / (k2 * k);
}

void Erf::Initialize_Erf(float Ao, float s2, float Rp_in, int tA, int tB) {
    initial_prefactor = final_prefactor = Ao;
    sigma_squared = s2; Rp = Rp_in;
    type1 = tA; type2 = tB;
}
double Erf::kspace(double k2, double A) const {
    double s = sphere_kspace(k2, Rp);
    return A * std::exp(-k2 * sigma_squared * 0.5) * s * s;
}

void GaussianErf::Initialize_Gaussian_Erf(float Ao, float s2, float Rp_in, int tA, int tB) {
    initial_prefactor = final_prefactor = Ao;
    sigma_squared = s2; Rp = Rp_in;
    type1 = tA; type2 = tB;
}
double GaussianErf::kspace(double k2, double A) const {
    return A * std::exp(-k2 * sigma_squared * 0.5) * sphere_kspace(k2, Rp);
}

double Charges::potential(double k2) const {
    if (k2 == 0.0) return 0.0;
    double a = smearing_length;
    return 4.0 * kPi * (double)bjerrum_length / k2 * std::exp(-k2 / (2.0 * a * a));
}

std::vector<const PairStyle*> System::styles() const {
    std::vector<const PairStyle*> v;
    for (auto& p : Gausses) v.push_back(p.get());
    for (auto& p : Erfs) v.push_back(p.get());
    for (auto& p : GaussErfs) v.push_back(p.get());
    return v;
}

int System::derive(std::string& err) {
    M = 1; grid_per_partic = 1; gvol = 1.0f;
    for (int j = 0; j < Dim; ++j) {
        if (Nx[j] <= 0) { err = "grid size N" + std::string(1, "xyz"[j]) + " is not set"; return kConfig; }
        M *= Nx[j];
        dx[j] = L[j] / float(Nx[j]);
        gvol *= dx[j];
        grid_per_partic *= (pmeorder + 1);
    }
    n_P_comps = (Dim == 2) ? 3 : 6;
    for (const FieldPhase& F : Fields) {
        // pair_style_fieldphases.cu:89-185 defines phases 0, 2 and 3; phase 1 (BCC) has
        // no branch there and leaves the field uninitialised
        if (F.phase == 1) { err = "fieldphase phase 1 (BCC) is not defined by the reference"; return kUnsupported; }
        if (F.phase != 0 && F.phase != 2 && F.phase != 3) { err = "unknown fieldphase phase " + std::to_string(F.phase); return kConfig; }
        if (F.phase == 0 && (F.dir < 0 || F.dir >= Dim)) { err = "fieldphase direction out of range"; return kConfig; }
        if (F.phase == 2 && (F.dir < 0 || F.dir > 2 || (Dim == 2 && F.dir != 2))) { err = "fieldphase cylinder direction out of range (2D needs 2)"; return kConfig; }
        if (F.phase == 3 && Dim != 3) { err = "fieldphase gyroid needs Dim 3"; return kUnsupported; }
    }
    return 0;
}

void fieldphase_force_table(const System& S, const FieldPhase& F, std::vector<float>& fr) {
    const int D = S.Dim, M = S.M;
    const float PI = 3.141592654f;     // globals.h:29
    const float Ao = F.initial_prefactor;
    fr.assign((size_t)M * D, 0.0f);
    for (int ind = 0; ind < M; ++ind) {
        const int n[3] = {ind % S.Nx[0], (ind / S.Nx[0]) % S.Nx[1], D == 3 ? ind / (S.Nx[0] * S.Nx[1]) : 0};
        float r[3] = {0.f, 0.f, 0.f};
        for (int j = 0; j < D; ++j) r[j] = float(n[j]) * S.dx[j];   // d_get_r, device_array_utils.cu:34-41
        float* f = fr.data() + (size_t)ind * D;
        if (F.phase == 0) {
            const float a = 2.0f * PI * float(F.n_periods) / S.L[F.dir];
            f[F.dir] = Ao * a * std::cos(a * r[F.dir]);
        } else if (F.phase == 2) {
            int d1 = 0, d2 = 1;
            if (F.dir == d1) d1 = 2;
            else if (F.dir == d2) d2 = 2;
            const float a1 = 2.0f * PI * float(F.n_periods) / S.L[d1];
            const float a2 = 2.0f * PI * float(F.n_periods) / S.L[d2];
            if (D == 3) f[F.dir] = 0.0f;
            f[d1] = Ao * a1 * std::sin(a1 * r[d1]) * std::cos(a2 * r[d2]);
            f[d2] = Ao * a2 * std::sin(a2 * r[d2]) * std::cos(a1 * r[d1]);
        } else {
            float a[3], c[3], sn[3];
            for (int j = 0; j < 3; ++j) {
                a[j] = 2.0f * PI * float(F.n_periods) / S.L[j];
                c[j] = std::cos(a[j] * r[j]);
                sn[j] = std::sin(a[j] * r[j]);
            }
            const float e = sn[0] * c[1] + sn[1] * c[2] + sn[2] * c[0];
            f[0] = -Ao * e * a[0] * (c[0] * c[1] - sn[2] * sn[0]);
            f[1] = -Ao * e * a[1] * (c[1] * c[2] - sn[0] * sn[1]);
            f[2] = -Ao * e * a[2] * (c[2] * c[0] - sn[1] * sn[2]);
        }
    }
}

// ---------------------------------------------------------------- helpers
namespace {

bool to_int(const std::string& s, int& v) {           // std::stoi semantics
    const char* b = s.c_str(); char* e = nullptr;
    errno = 0;
    long r = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE) return false;
    v = (int)r; return true;
}
bool to_float(const std::string& s, float& v) {       // std::stof semantics
    const char* b = s.c_str(); char* e = nullptr;
    float r = std::strtof(b, &e);
    if (e == b) return false;
    v = r; return true;
}
std::string join(const std::string& base, const std::string& name) {
    if (name.empty() || name[0] == '/' || base.empty()) return name;
    return base + "/" + name;
}
bool read_file(const std::string& path, std::string& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::ostringstream ss; ss << in.rdbuf();
    out = ss.str();
    return true;
}
std::vector<std::string> split_lines(const std::string& t) {
    std::vector<std::string> v;
    size_t p = 0;
    while (p <= t.size()) {
        size_t n = t.find('\n', p);
        if (n == std::string::npos) { if (p < t.size()) v.push_back(t.substr(p)); break; }
        v.push_back(t.substr(p, n - p));
        p = n + 1;
    }
    return v;
}
std::vector<std::string> words(const std::string& line) {
    std::vector<std::string> w;
    std::istringstream is(line);
    std::string s;
    while (is >> s) w.push_back(s);
    return w;
}
bool blank(const std::string& l) {
    for (char c : l) if (!std::isspace((unsigned char)c)) return false;
    return true;
}

// Data file reader (config_utils.cu:5-364 format).
int read_data(System& S, const std::string& path, bool charged, std::string& err) {
    std::string text;
    if (!read_file(path, text)) { err = "Failed to open " + path + "!"; return kParse; }
    auto L = split_lines(text);
    int n_angles = 0;
    bool have[3] = {false, false, false};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    float lof[3] = {0, 0, 0}, hif[3] = {1, 1, 1};
    size_t i = 1;  // line 0 is a title
    S.ns = -1; S.ntypes = -1; S.n_total_bonds = 0; S.nbond_types = 0;
    // header: until the first section keyword
    for (; i < L.size(); ++i) {
        auto w = words(L[i]);
        if (w.empty()) continue;
        if (w[0] == "Masses" || w[0] == "Atoms" || w[0] == "Bonds") break;
        int v;
        if (w.size() >= 3 && w[1] == "atom" && w[2] == "types") { if (!to_int(w[0], v)) return err = "bad atom types line", kParse; S.ntypes = v; }
        else if (w.size() >= 3 && w[1] == "bond" && w[2] == "types") { if (!to_int(w[0], v)) return err = "bad bond types line", kParse; S.nbond_types = v; }
        else if (w.size() >= 3 && w[1] == "angle" && w[2] == "types") { /* not used */ }
        else if (w.size() >= 2 && w[1] == "atoms") { if (!to_int(w[0], v)) return err = "bad atoms line", kParse; S.ns = v; }
        else if (w.size() >= 2 && w[1] == "bonds") { if (!to_int(w[0], v)) return err = "bad bonds line", kParse; S.n_total_bonds = v; }
        else if (w.size() >= 2 && w[1] == "angles") { if (!to_int(w[0], v)) return err = "bad angles line", kParse; n_angles = v; }
        else if (w.size() >= 4 && w[2].size() == 3 && w[3].size() == 3 && w[2].compare(1, 2, "lo") == 0 &&
                 w[3].compare(1, 2, "hi") == 0 && w[2][0] >= 'x' && w[2][0] <= 'z' && w[3][0] == w[2][0]) {
            int d = w[2][0] - 'x';
            if (charged) {
                if (!to_float(w[0], lof[d]) || !to_float(w[1], hif[d])) return err = "bad box line", kParse;
            } else {
                lo[d] = std::strtod(w[0].c_str(), nullptr); hi[d] = std::strtod(w[1].c_str(), nullptr);
            }
            have[d] = true;
        }
    }
    if (S.ns <= 0 || S.ntypes <= 0) { err = "malformed data header in " + path; return kParse; }
    if (!have[0] || !have[1] || (S.Dim > 2 && !have[2])) { err = "box bounds missing in " + path; return kParse; }
    for (int d = 0; d < 3; ++d) {
        if (d < 2 || S.Dim > 2) S.L[d] = charged ? (float)(hif[d] - lof[d]) : (float)(hi[d] - lo[d]);
        else S.L[d] = 1.0f;
    }
    S.V = 1.0f;
    for (int d = 0; d < S.Dim; ++d) { S.Lh[d] = 0.5f * S.L[d]; S.V *= S.L[d]; }

    const int ns = S.ns, D = S.Dim;
    S.x.assign((size_t)ns * D, 0.f);
    S.tp.assign(ns, 0); S.molecID.assign(ns, 0); S.q.assign(ns, 0.f);
    S.mass.assign(S.ntypes, 0.f); S.Diff.assign(S.ntypes, 1.0f);   // initialize.cu:312-313
    S.n_bonds.assign(ns, 0);
    S.bonded_to.assign((size_t)ns * kMaxBonds, 0);
    S.bond_type.assign((size_t)ns * kMaxBonds, 0);
    int nbt = S.nbond_types > 0 ? S.nbond_types : 1;
    S.bond_k.assign(nbt, 0.f); S.bond_req.assign(nbt, 0.f);

    auto next_data = [&](size_t& j) -> bool {
        while (j < L.size() && blank(L[j])) ++j;
        return j < L.size();
    };
    std::vector<char> seen(ns, 0);
    for (; i < L.size(); ++i) {
        auto w = words(L[i]);
        if (w.empty()) continue;
        const std::string sec = w[0];
        if (sec == "Masses") {
            ++i;
            for (int k = 0; k < S.ntypes; ++k, ++i) {
                if (!next_data(i)) return err = "Premature end of Masses", kParse;
                auto m = words(L[i]);
                int t; float mf;
                if (m.size() < 2 || !to_int(m[0], t) || t < 1 || t > S.ntypes) return err = "bad Masses line", kParse;
                if (charged) { if (!to_float(m[1], mf)) return err = "bad mass", kParse; }
                else mf = (float)std::strtod(m[1].c_str(), nullptr);
                S.mass[t - 1] = mf;
            }
            --i;
        } else if (sec == "Atoms") {
            ++i;
            size_t need = 3 + (charged ? 1 : 0) + D;
            for (int k = 0; k < ns; ++k, ++i) {
                if (!next_data(i)) return err = "Premature end of input.conf!", kParse;
                auto a = words(L[i]);
                int id, mol, ty;
                if (a.size() < need || !to_int(a[0], id) || !to_int(a[1], mol) || !to_int(a[2], ty))
                    return err = "malformed atom line " + std::to_string(k + 1), kParse;
                if (id < 1 || id > ns || ty < 1 || ty > S.ntypes)
                    return err = "atom id/type out of range at atom line " + std::to_string(k + 1), kParse;
                int idx = id - 1;
                seen[idx] = 1;
                S.molecID[idx] = mol - 1;
                S.tp[idx] = ty - 1;
                size_t c = 3;
                if (charged) { float qv; if (!to_float(a[c++], qv)) return err = "bad charge", kParse; S.q[idx] = qv; }
                for (int d = 0; d < D; ++d) {
                    if (charged) { float v; if (!to_float(a[c + d], v)) return err = "bad coordinate", kParse; S.x[(size_t)idx * D + d] = v; }
                    else S.x[(size_t)idx * D + d] = (float)std::strtod(a[c + d].c_str(), nullptr);
                }
            }
            --i;
        } else if (sec == "Bonds") {
            ++i;
            for (int k = 0; k < S.n_total_bonds; ++k, ++i) {
                if (!next_data(i)) return err = "Premature end of input.conf!", kParse;
                auto b = words(L[i]);
                int bid, bt, i1, i2;
                if (b.size() < 4 || !to_int(b[0], bid) || !to_int(b[1], bt) || !to_int(b[2], i1) || !to_int(b[3], i2))
                    return err = "malformed bond line " + std::to_string(k + 1), kParse;
                --bt; --i1; --i2;
                if (i2 < i1) std::swap(i1, i2);   // config_utils.cu:111-115
                if (i1 < 0 || i2 >= ns || bt < 0 || bt >= nbt) return err = "bond index out of range", kParse;
                if (S.n_bonds[i1] >= kMaxBonds || S.n_bonds[i2] >= kMaxBonds)
                    return err = "more than 3 bonds on one site", kParse;
                S.bonded_to[(size_t)i1 * kMaxBonds + S.n_bonds[i1]] = i2;
                S.bond_type[(size_t)i1 * kMaxBonds + S.n_bonds[i1]] = bt;
                S.n_bonds[i1]++;
                S.bonded_to[(size_t)i2 * kMaxBonds + S.n_bonds[i2]] = i1;
                S.bond_type[(size_t)i2 * kMaxBonds + S.n_bonds[i2]] = bt;
                S.n_bonds[i2]++;
            }
            --i;
        } else if (sec == "Angles") {
            break;
        }
    }
    for (int k = 0; k < ns; ++k)
        if (!seen[k]) { err = "atom " + std::to_string(k + 1) + " missing from Atoms section"; return kParse; }
    if (n_angles > 0) { err = "Angles not yet set up on device!"; return kUnsupported; }  // initialize.cu:166-168
    return 0;
}

// lammpstrj resume frame (config_utils.cu:405-454)
int read_resume(System& S, const std::string& path, std::string& err) {
    std::string text;
    if (!read_file(path, text)) { err = "Failed to open " + path + "!"; return kParse; }
    auto L = split_lines(text);
    if (L.size() < 9) return err = "resume file too short", kParse;
    int n;
    if (!to_int(words(L[3]).empty() ? std::string() : words(L[3])[0], n) || n != S.ns)
        return err = "number of sites in resume file doesn't match input!", kParse;
    for (int i = 0; i < S.ns; ++i) {
        if (9 + (size_t)i >= L.size()) return err = "resume file truncated", kParse;
        auto w = words(L[9 + i]);
        if (w.size() < 3 + (size_t)S.Dim) return err = "malformed resume line", kParse;
        for (int d = 0; d < S.Dim; ++d) {
            float v;
            if (!to_float(w[3 + d], v)) return err = "malformed resume coordinate", kParse;
            S.x[(size_t)i * S.Dim + d] = v;
        }
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------- deck parser
int load_system(const std::string& text, const std::string& base_dir, System& S, std::string& err) {
    S.noise_mag = std::sqrt(2.0f * S.delt);
    bool config_read = false, charge_flag = false;
    auto lines = split_lines(text);
    size_t li = 0;
    auto bad = [&](const std::string& m) { err = m; return kParse; };
    while (li < lines.size()) {
        std::string line = lines[li++];
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string w;
        while (iss >> w) {
            std::string a, b2;
            int iv; float fv;
            auto geti = [&](int& v) { return (bool)(iss >> a) && to_int(a, v); };
            auto getf = [&](float& v) { return (bool)(iss >> a) && to_float(a, v); };
            if (w == "binary_freq") { if (!geti(iv)) return bad("binary_freq needs an integer"); S.bin_freq = iv; }
            else if (w == "bond") {
                if (!config_read) return bad("Error in input file, bond keyword before config read");
                int bt; float k, r0;
                if (!geti(bt)) return bad("bond needs a type");
                if (!(iss >> b2) || b2 != "harmonic") return bad("You fool! That bond type is not supported!");
                if (!getf(k) || !getf(r0)) return bad("bond needs k and r0");
                if (bt < 1 || bt > (int)S.bond_k.size()) return bad("bond type out of range");
                S.bond_k[bt - 1] = k; S.bond_req[bt - 1] = r0;
            }
            else if (w == "charges") {
                if (config_read) return bad("charges keyword must come before read_data command!");
                charge_flag = true;
                if (!getf(fv)) return bad("charges needs a Bjerrum length");
                S.charges.bjerrum_length = fv;
                if (!getf(fv)) return bad("charges needs a smearing length");
                S.charges.smearing_length = fv;
                S.charges.enabled = true;
            }
            else if (w == "compute") {
                // Compute::Initialize (Compute.cu:20-75): avg_sk is the only style; the
                // type is checked against the types read so far (none before read_data)
                std::string style;
                iss >> style;
                if (style != "avg_sk") return bad("Undefined compute style!");
                SkCompute c;
                if (!geti(iv)) return bad("Invalid type to calculate avg_sk");
                c.type = iv - 1;
                if (c.type < 0 || !config_read || c.type >= S.ntypes) return bad("Invalid type to calculate avg_sk");
                while (iss >> a) {
                    if (a == "freq") { if (!geti(iv) || iv < 1) return bad("compute avg_sk: bad freq"); c.freq = iv; }
                    else if (a == "wait") { if (!geti(iv)) return bad("compute avg_sk: bad wait"); c.wait = iv; }
                }
                S.computes.push_back(c);
            }
            else if (w == "delt") { if (!getf(fv)) return bad("delt needs a number"); S.delt = fv; S.noise_mag = std::sqrt(2.f * S.delt); }
            else if (w == "diffusivity") {
                if (!config_read) return bad("Error in input file, diffusivities defined before config read");
                if (!geti(iv)) return bad("diffusivity needs a type");
                if (iv > S.ntypes || iv < 1) return bad("Invalid type for diffusivity!");
                if (!getf(fv)) return bad("diffusivity needs a value");
                S.Diff[iv - 1] = fv;
            }
            else if (w == "Dim") { if (!geti(iv)) return bad("Dim needs an integer"); if (iv != 2 && iv != 3) return bad("Dim must be 2 or 3"); S.Dim = iv; }
            else if (w == "grid_freq") { if (!geti(iv)) return bad("grid_freq needs an integer"); S.grid_freq = iv; }
            else if (w == "group") {
                if (!config_read) return bad("Must read the configuration file before defining a group!");
                Group g;
                if (!(iss >> g.name)) return bad("group needs a name");
                if (!(iss >> b2)) return bad("group needs a style");
                g.command_line = line;
                if (b2 != "type") { err = "ERROR: Group type " + b2 + " not supported!"; return kParse; }
                if (!geti(iv)) return bad("group type needs an integer");
                for (int i = 0; i < S.ns; ++i) if (S.tp[i] == iv - 1) g.index.push_back(i);
                S.Groups.push_back(std::move(g));
            }
            else if (w == "integrator") {
                if (!config_read) return bad("Integrator must be defined after read_data command!");
                Integrator I;
                if (!(iss >> I.group_name) || !(iss >> I.name)) return bad("integrator needs a group and a name");
                if (I.name != "EM" && I.name != "GJF") { err = I.name + " is not a valid integrator!"; return kParse; }
                for (size_t g = 0; g < S.Groups.size(); ++g) if (S.Groups[g].name == I.group_name) I.group_index = (int)g;
                if (I.group_index < 0) { err = "Group " + I.group_name + " not found to integrate!"; return kConfig; }
                S.Integrators.push_back(I);
            }
            else if (w == "log_freq") { if (!geti(iv)) return bad("log_freq needs an integer"); S.log_freq = iv; }
            else if (w == "max_steps") { if (!geti(iv)) return bad("max_steps needs an integer"); S.max_steps = iv; }
            else if (w == "n_gaussians" || w == "n_erfs" || w == "n_gaussian_erfs" || w == "n_fieldphases") {
                if (!geti(iv) || iv < 0) return bad(w + " needs a count");
                std::string want = w.substr(2, w.size() - 3);   // gaussian / erf / gaussian_erf / fieldphase
                if (w == "n_gaussians") { S.Gausses.clear(); }
                else if (w == "n_erfs") { S.Erfs.clear(); }
                else if (w == "n_gaussian_erfs") { S.GaussErfs.clear(); }
                else { S.Fields.clear(); }
                for (int k = 0; k < iv; ++k) {
                    std::string sl = li < lines.size() ? lines[li] : std::string();
                    ++li;
                    auto t = words(sl);
                    if (t.empty() || t[0] != want) { err = "Read: " + (t.empty() ? std::string() : t[0]) + " but expected '" + want + "'"; return kParse; }
                    if (want == "fieldphase") {   // read_input.cu:237-297
                        FieldPhase F;
                        int t1, t2;
                        if (t.size() < 7 || !to_int(t[1], t1) || !to_int(t[2], t2) || !to_float(t[3], F.initial_prefactor) ||
                            !to_int(t[4], F.phase) || !to_int(t[5], F.dir) || !to_int(t[6], F.n_periods))
                            { err = "malformed fieldphase line: " + sl; return kParse; }
                        if (t1 < 1 || t2 < 1 || t1 > S.ntypes || t2 > S.ntypes) { err = "Invalid particle type in FieldPhase potential input!"; return kParse; }
                        F.type1 = t1 - 1; F.type2 = t2 - 1;
                        F.final_prefactor = F.initial_prefactor;
                        if (t.size() > 7 && t[7] == "ramp") {
                            if (t.size() <= 8 || !to_float(t[8], F.final_prefactor)) { err = "ramp needs a final prefactor"; return kParse; }
                            F.ramp = true;
                            err = "ramped pair-style prefactors are not supported by this engine";
                            return kUnsupported;
                        }
                        S.Fields.push_back(F);
                        continue;
                    }
                    size_t need = (want == "gaussian") ? 5 : 6;
                    int t1, t2; float A, p1, p2 = 0;
                    if (t.size() < need || !to_int(t[1], t1) || !to_int(t[2], t2) || !to_float(t[3], A) || !to_float(t[4], p1) ||
                        (need == 6 && !to_float(t[5], p2)))
                        { err = "malformed " + want + " line: " + sl; return kParse; }
                    if (t1 < 1 || t2 < 1 || t1 > S.ntypes || t2 > S.ntypes) { err = "Invalid particle type in " + want + " input!"; return kParse; }
                    bool ramp = false; float Afin = A;
                    if (t.size() > need && t[need] == "ramp") {
                        if (t.size() <= need + 1 || !to_float(t[need + 1], Afin)) { err = "ramp needs a final prefactor"; return kParse; }
                        ramp = true;
                    }
                    PairStyle* ps;
                    if (want == "gaussian") {
                        auto g = std::make_unique<Gaussian>();
                        double s2 = (double)p1; s2 *= s2;       // read_input.cu:329-331
                        g->Initialize_Gaussian(A, 0.f, t1 - 1, t2 - 1); g->sigma_squared = s2;
                        ps = g.get(); S.Gausses.push_back(std::move(g));
                    } else if (want == "erf") {
                        auto g = std::make_unique<Erf>();
                        double s2 = (double)p2; s2 *= s2;       // read_input.cu:205-210
                        g->Initialize_Erf(A, 0.f, p1, t1 - 1, t2 - 1); g->sigma_squared = s2;
                        ps = g.get(); S.Erfs.push_back(std::move(g));
                    } else {
                        auto g = std::make_unique<GaussianErf>();
                        double s2 = (double)p2; s2 *= s2;       // read_input.cu:386-391
                        g->Initialize_Gaussian_Erf(A, 0.f, p1, t1 - 1, t2 - 1); g->sigma_squared = s2;
                        ps = g.get(); S.GaussErfs.push_back(std::move(g));
                    }
                    ps->ramp = ramp; ps->final_prefactor = Afin;
                    // the reference's update divides by the current k = 0 value
                    // (pair_style.cu:318-324): a zero prefactor has no ramp
                    if (ramp && A == 0.0f) { err = "ramp from a zero prefactor"; return kUnsupported; }
                }
            }
            else if (w == "Nx") { if (!geti(iv)) return bad("Nx needs an integer"); S.Nx[0] = iv; }
            else if (w == "Ny") { if (!geti(iv)) return bad("Ny needs an integer"); S.Nx[1] = iv; }
            else if (w == "Nz") { if (!geti(iv)) return bad("Nz needs an integer"); if (S.Dim > 2) S.Nx[2] = iv; }
            else if (w == "pmeorder") {
                if (!geti(iv)) return bad("pmeorder needs an integer");
                if (iv > 4 || iv < 0) return bad("PMEORDER greater than 4 not supported!");
                S.pmeorder = iv;
            }
            else if (w == "rand_seed") { if (!geti(iv)) return bad("rand_seed needs an integer"); S.rand_seed = (uint32_t)iv; S.seed_given = true; }
            else if (w == "read_data") {
                if (!(iss >> a)) return bad("read_data needs a file name");
                int rc = read_data(S, join(base_dir, a), charge_flag, err);
                if (rc) return rc;
                config_read = true;
                S.Groups.clear();
                Group all; all.name = "all"; all.command_line = "NONE";   // group.cu:30-42
                all.index.resize(S.ns);
                for (int i = 0; i < S.ns; ++i) all.index[i] = i;
                S.Groups.push_back(std::move(all));
            }
            else if (w == "read_resume") {
                if (!(iss >> a)) return bad("read_resume needs a file name");
                if (!config_read) return bad("read_resume before read_data");
                int rc = read_resume(S, join(base_dir, a), err);
                if (rc) return rc;
            }
            else if (w == "skip_steps") { if (!geti(iv)) return bad("skip_steps needs an integer"); S.skip_steps = iv; }
            else if (w == "struc_freq") { if (!geti(iv)) return bad("struc_freq needs an integer"); S.struc_freq = iv; }
            else if (w == "threads") { if (!geti(iv)) return bad("threads needs an integer"); S.threads = iv; }
            else if (w == "traj_freq") { if (!geti(iv)) return bad("traj_freq needs an integer"); S.traj_freq = iv; }
            else if (w == "traj_name") { S.dump_name.clear(); iss >> S.dump_name; }
            else { err = "Invalid keyword: " + w; return kParse; }
        }
    }
    if (!config_read) { err = "no read_data command in input"; return kConfig; }
    if (!S.seed_given) S.rand_seed = (uint32_t)time(nullptr);   // read_input.cu:554
    return S.derive(err);
}

}  // namespace tild
