// Host side of the C ABI that needs no GPU: initial collapse states, the
// 172-byte packed final-state file (write with rotation, read with exact size
// check) and the final-state comparator with its failures CSV.  Formats and
// tolerances are the reference's (reproducer.cpp:7406-7431, 7488-7505,
// 8036-8142, 8289-8578).
#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "../../include/warp_reactor.h"
#include "wr_eos.hpp"

static_assert(sizeof(wr_burn_state) == 136, "burn_t layout");
static_assert(sizeof(wr_integrator_stats) == 56, "IntegratorStats layout");
static_assert(sizeof(wr_collapse_state) == 216, "CollapseState layout");
static_assert(sizeof(wr_packed_final_state) == 172, "PackedFinalState layout");

namespace {

const char* const kSpeciesNames[WR_NUM_SPEC] = {"E",  "Hp", "H",  "Hm",  "Dp",  "D",   "H2p",
                                                "Dm", "H2", "HDp", "HD", "HEpp", "HEp", "HE"};
// Comparator tolerances (reproducer.cpp:7409-7413, 7429-7431).
constexpr double kSpeciesRtol[WR_NUM_SPEC] = {1.0e-3, 1.0e-3, 1.1e-3, 1.0e-3, 1.0e-4, 10.0, 1.0e-3,
                                              1.0e-4, 3.0e-4, 1.0e-4, 10.0, 1.0e-4, 1.0e-4, 1.0e-4};
constexpr double kRtolSpec = 1.0e-4, kAtolSpec = 1.0e-4, kAtolEnergy = 1.0e-6;
constexpr double kThermoRtol = 2.0e-4;

bool deuterium_bearing(int n) { return n == 4 || n == 5 || n == 7 || n == 9 || n == 10; }

bool nearly_equal(double v, double r, double rtol, double atol) {
    return std::abs(v - r) <= atol + rtol * std::abs(r);
}

double relative_error(double v, double r, double atol) {
    const double denom = std::max(std::abs(r), atol);
    return std::abs(v - r) / denom;
}

// std::setprecision(17) << std::scientific << (long double)value
std::string sci(double v) {
    char buf[64];
    snprintf(buf, sizeof(buf), "%.17Le", static_cast<long double>(v));
    return buf;
}

std::string csv_escape(const std::string& value) {
    if (value.find_first_of(",\"\n\r") == std::string::npos) return value;
    std::string out = "\"";
    for (char ch : value) {
        if (ch == '"') out.push_back('"');
        out.push_back(ch);
    }
    out.push_back('"');
    return out;
}

struct Failure {
    int32_t cell, i, j, k;
    std::string field, actual, reference, rel, rtol, atol;
};

bool file_exists(const std::string& path, bool* error) {
    struct stat st;
    if (stat(path.c_str(), &st) == 0) return true;
    *error = (errno != ENOENT);
    return false;
}

}  // namespace

extern "C" {

// make_initial_state / make_collapse_state (reproducer.cpp:7424-7426, 7533-7546)
int wr_make_collapse_states(wr_collapse_state* cells, int num_cells) {
    if (!cells || num_cells < 0) return WR_ERR_ARGUMENT;
    static const double n0[WR_NUM_SPEC] = {1.0e-4, 1.0e-4, 1.0e0, 1.0e-40, 1.0e-40, 1.0e-40, 1.0e-40,
                                          1.0e-40, 1.0e-6, 1.0e-40, 1.0e-40, 1.0e-40, 1.0e-40, 0.0775};
    wr_collapse_state s;
    memset(&s, 0, sizeof(s));
    double xn[WR_NUM_SPEC];
    memcpy(xn, n0, sizeof(xn));
    const double T = 1.0e2;
    const double rho = wr::density_of(xn);
    wr::normalize(xn, rho);
    s.current.rho = rho;
    s.current.T = T;
    s.current.e = wr::eos_energy(xn, T);
    memcpy(s.current.xn, xn, sizeof(xn));
    s.time = 0.0;
    s.density_driver = rho;
    for (int c = 0; c < num_cells; ++c) cells[c] = s;
    return WR_SUCCESS;
}

// Synthetic initial states of the benchmark recipes (no counterpart in the
// reference, whose only initial state is make_collapse_state).  Cell c draws
// three SplitMix64 outputs h1 = splitmix64(seed ^ c), h2 = splitmix64(h1),
// h3 = splitmix64(h2) and uniforms u = (h >> 11) 2^-53 (u1 shifted into
// (0, 1]).  Box-Muller: z = sqrt(-2 ln u1) cos(2 pi u2); hydrogen density
// n_H = exp(sigma z) cm^-3; T = t_min (t_max^... log-uniform when t_max >
// t_min: T = t_min exp(u3 ln(t_max / t_min))).  Composition per n_H: H+ 1e-4,
// H2 1e-6, H the rest of the hydrogen, He by mass fraction Y = 0.24
// (n_He / n_H = Y / (4 (1 - Y))), D/H 3e-5, all other species 1e-40;
// electrons by charge balance.  Then as make_initial_state
// (reproducer.cpp:7533-7541): rho from the number densities, normalise,
// e from eos_rt; the density driver starts at rho.
int wr_make_recipe_states(wr_collapse_state* cells, int num_cells, uint64_t seed,
                          double sigma_ln_nh, double t_min, double t_max) {
    if (!cells || num_cells < 0 || !(sigma_ln_nh >= 0.0) || !std::isfinite(sigma_ln_nh) ||
        !(t_min > 0.0) || !(t_max >= t_min) || !std::isfinite(t_max))
        return WR_ERR_ARGUMENT;
    const double two_pi = 2.0 * wr::kPi;
    const double he_per_h = 0.24 / (4.0 * (1.0 - 0.24));
    for (int c = 0; c < num_cells; ++c) {
        const uint64_t h1 = wr::splitmix64(seed ^ static_cast<uint64_t>(c));
        const uint64_t h2 = wr::splitmix64(h1);
        const uint64_t h3 = wr::splitmix64(h2);
        const double u1 = static_cast<double>((h1 >> 11U) + 1U) * 0x1p-53;
        const double u2 = static_cast<double>(h2 >> 11U) * 0x1p-53;
        const double u3 = static_cast<double>(h3 >> 11U) * 0x1p-53;
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
        const double nh = std::exp(sigma_ln_nh * z);
        const double T = (t_max > t_min) ? t_min * std::exp(u3 * std::log(t_max / t_min)) : t_min;
        double xn[WR_NUM_SPEC];
        for (int n = 0; n < WR_NUM_SPEC; ++n) xn[n] = 1.0e-40 * nh;
        xn[1] = 1.0e-4 * nh;
        xn[8] = 1.0e-6 * nh;
        xn[2] = nh - xn[1] - 2.0 * xn[8];
        xn[5] = 3.0e-5 * nh;
        xn[13] = he_per_h * nh;
        wr::balance_charge(xn);
        wr_collapse_state s;
        memset(&s, 0, sizeof(s));
        const double rho = wr::density_of(xn);
        wr::normalize(xn, rho);
        s.current.rho = rho;
        s.current.T = T;
        s.current.e = wr::eos_energy(xn, T);
        memcpy(s.current.xn, xn, sizeof(xn));
        s.density_driver = rho;
        cells[c] = s;
    }
    return WR_SUCCESS;
}

// make_packed_final_state (reproducer.cpp:8289-8305) for the cell at `index`.
static wr_packed_final_state pack_one(const wr_collapse_state& s, int index, int grid_dim) {
    wr_packed_final_state p;
    memset(&p, 0, sizeof(p));
    p.cell = index;
    p.i = index % grid_dim;
    p.j = (index / grid_dim) % grid_dim;
    p.k = index / (grid_dim * grid_dim);
    p.completed_steps = s.completed_steps;
    p.time = s.time;
    p.density_driver = s.density_driver;
    p.rho = s.current.rho;
    p.T = s.current.T;
    p.e = s.current.e;
    memcpy(p.xn, s.current.xn, sizeof(p.xn));
    return p;
}

int wr_pack_final_states(const wr_collapse_state* cells, int num_cells, int grid_dim,
                         wr_packed_final_state* out) {
    if ((!cells || !out) && num_cells > 0) return WR_ERR_ARGUMENT;
    if (num_cells < 0 || grid_dim <= 0) return WR_ERR_ARGUMENT;
    for (int c = 0; c < num_cells; ++c) out[c] = pack_one(cells[c], c, grid_dim);
    return WR_SUCCESS;
}

// rotate_existing_output + write_final_states (reproducer.cpp:8307-8378)
int wr_write_final_states(const wr_collapse_state* cells, int num_cells, int grid_dim,
                          const char* path) {
    if (!path || (!cells && num_cells > 0) || num_cells < 0 || grid_dim <= 0) return WR_ERR_ARGUMENT;
    const std::string p(path);
    bool err = false;
    if (file_exists(p, &err)) {
        std::random_device seed;
        std::mt19937 gen(seed());
        std::uniform_int_distribution<int> dist(0, 999999);
        bool moved = false;
        for (int attempt = 0; attempt < 100 && !moved; ++attempt) {
            char suffix[16];
            snprintf(suffix, sizeof(suffix), "%06d", dist(gen));
            const std::string backup = p + ".old." + suffix;
            bool e2 = false;
            if (file_exists(backup, &e2)) continue;
            if (e2) {
                fprintf(stderr, "failed to check backup final-state file: %s\n", backup.c_str());
                return WR_ERR_IO;
            }
            if (rename(p.c_str(), backup.c_str()) != 0) {
                fprintf(stderr, "failed to move existing final-state output file from %s to %s (%s)\n",
                        p.c_str(), backup.c_str(), strerror(errno));
                return WR_ERR_IO;
            }
            printf("existing final states moved to: %s\n", backup.c_str());
            moved = true;
        }
        if (!moved) {
            fprintf(stderr, "failed to choose a unique backup filename for: %s\n", p.c_str());
            return WR_ERR_IO;
        }
    } else if (err) {
        fprintf(stderr, "failed to check final-state output file: %s\n", p.c_str());
        return WR_ERR_IO;
    }
    FILE* f = fopen(p.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "failed to open final-state output file: %s\n", p.c_str());
        return WR_ERR_IO;
    }
    std::vector<wr_packed_final_state> buf(4096);
    for (int c0 = 0; c0 < num_cells; c0 += 4096) {
        const int m = std::min(4096, num_cells - c0);
        for (int k = 0; k < m; ++k) buf[k] = pack_one(cells[c0 + k], c0 + k, grid_dim);
        if (fwrite(buf.data(), sizeof(wr_packed_final_state), m, f) != static_cast<size_t>(m)) {
            fprintf(stderr, "failed to write final-state output file: %s\n", p.c_str());
            fclose(f);
            return WR_ERR_IO;
        }
    }
    if (fclose(f) != 0) {
        fprintf(stderr, "failed to write final-state output file: %s\n", p.c_str());
        return WR_ERR_IO;
    }
    return WR_SUCCESS;
}

// read_final_states (reproducer.cpp:8380-8414)
int wr_read_final_states(const char* path, int expected_records, wr_packed_final_state* out) {
    if (!path || expected_records < 0 || (!out && expected_records > 0)) return WR_ERR_ARGUMENT;
    FILE* f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "failed to open final-state comparison file: %s\n", path);
        return WR_ERR_IO;
    }
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        fprintf(stderr, "failed to determine size of final-state comparison file: %s\n", path);
        return WR_ERR_IO;
    }
    const long size = ftell(f);
    const unsigned long long expected =
        static_cast<unsigned long long>(expected_records) * sizeof(wr_packed_final_state);
    if (size < 0 || static_cast<unsigned long long>(size) != expected) {
        fprintf(stderr,
                "final-state comparison file has %ld bytes, expected %llu bytes (%d packed records "
                "of %zu bytes)\n",
                size, expected, expected_records, sizeof(wr_packed_final_state));
        fclose(f);
        return WR_ERR_IO;
    }
    fseek(f, 0, SEEK_SET);
    const size_t got = expected_records ? fread(out, sizeof(wr_packed_final_state), expected_records, f) : 0;
    fclose(f);
    if (got != static_cast<size_t>(expected_records)) {
        fprintf(stderr, "failed to read final-state comparison file: %s\n", path);
        return WR_ERR_IO;
    }
    return WR_SUCCESS;
}

// compare_final_states (reproducer.cpp:8416-8549) + failures CSV (8100-8142)
int wr_compare_final_states(const wr_collapse_state* cells, int num_cells, int grid_dim,
                            const char* path, const char* csv_path, wr_comparison_summary* out) {
    if (!cells || !path || !out || num_cells <= 0 || grid_dim <= 0) return WR_ERR_ARGUMENT;
    std::vector<wr_packed_final_state> ref(static_cast<size_t>(num_cells));
    int st = wr_read_final_states(path, num_cells, ref.data());
    if (st != WR_SUCCESS) return st;
    memset(out, 0, sizeof(*out));
    out->pass = 1;
    std::string first;
    std::vector<Failure> failures;
    std::vector<wr_packed_final_state> act(static_cast<size_t>(num_cells));
    wr_pack_final_states(cells, num_cells, grid_dim, act.data());
    auto add = [&](const wr_packed_final_state& a, const std::string& field, const std::string& av,
                   const std::string& rv, const std::string& rel, const std::string& rt,
                   const std::string& at) {
        failures.push_back({a.cell, a.i, a.j, a.k, field, av, rv, rel, rt, at});
    };
    for (int c = 0; c < num_cells; ++c) {
        const auto& a = act[c];
        const auto& e = ref[c];
        const bool meta = a.cell == e.cell && a.i == e.i && a.j == e.j && a.k == e.k &&
                          a.completed_steps == e.completed_steps;
        if (!meta && first.empty()) {
            char buf[512];
            snprintf(buf, sizeof(buf),
                     "metadata mismatch at cell %d (actual cell/i/j/k/steps %d/%d/%d/%d/%d, "
                     "reference %d/%d/%d/%d/%d)",
                     c, a.cell, a.i, a.j, a.k, a.completed_steps, e.cell, e.i, e.j, e.k,
                     e.completed_steps);
            first = buf;
        }
        out->pass = out->pass && meta;
        const int32_t av[5] = {a.cell, a.i, a.j, a.k, a.completed_steps};
        const int32_t ev[5] = {e.cell, e.i, e.j, e.k, e.completed_steps};
        const char* names[5] = {"cell", "i", "j", "k", "completed_steps"};
        for (int q = 0; q < 5; ++q) {
            if (av[q] != ev[q]) add(a, names[q], std::to_string(av[q]), std::to_string(ev[q]), "", "", "");
        }
        for (int n = 0; n < WR_NUM_SPEC; ++n) {
            const double v = a.xn[n], r = e.xn[n];
            const double rel = relative_error(v, r, kAtolSpec);
            double& mx = deuterium_bearing(n) ? out->max_deuterium_species_rel_error
                                              : out->max_non_deuterium_species_rel_error;
            mx = std::max(mx, rel);
            if (deuterium_bearing(n)) continue;
            const bool ok = nearly_equal(v, r, kSpeciesRtol[n], kAtolSpec);
            if (!ok && first.empty()) {
                char buf[512];
                snprintf(buf, sizeof(buf),
                         "species %s mismatch at cell %d (actual %s, reference %s, relative error "
                         "%s, rtol %g, atol %g)",
                         kSpeciesNames[n], c, sci(v).c_str(), sci(r).c_str(), sci(rel).c_str(),
                         kSpeciesRtol[n], kAtolSpec);
                first = buf;
            }
            out->pass = out->pass && ok;
            if (!ok) add(a, kSpeciesNames[n], sci(v), sci(r), sci(rel), sci(kSpeciesRtol[n]), sci(kAtolSpec));
        }
        const double trel = relative_error(a.T, e.T, kAtolSpec);
        const double erel = relative_error(a.e, e.e, kAtolEnergy);
        const double rrel = relative_error(a.rho, e.rho, kAtolSpec);
        out->max_thermodynamic_rel_error =
            std::max(out->max_thermodynamic_rel_error, std::max(trel, erel));
        out->max_rho_rel_error = std::max(out->max_rho_rel_error, rrel);
        const bool tok = nearly_equal(a.T, e.T, kThermoRtol, kAtolSpec);
        const bool eok = nearly_equal(a.e, e.e, kThermoRtol, kAtolEnergy);
        const bool rok = nearly_equal(a.rho, e.rho, kRtolSpec, kAtolSpec);
        if ((!tok || !eok || !rok) && first.empty()) {
            char buf[512];
            snprintf(buf, sizeof(buf), "thermodynamic mismatch at cell %d (T rel %s, e rel %s, rho rel %s)",
                     c, sci(trel).c_str(), sci(erel).c_str(), sci(rrel).c_str());
            first = buf;
        }
        out->pass = out->pass && tok && eok && rok;
        if (!tok) add(a, "T", sci(a.T), sci(e.T), sci(trel), sci(kThermoRtol), sci(kAtolSpec));
        if (!eok) add(a, "e", sci(a.e), sci(e.e), sci(erel), sci(kThermoRtol), sci(kAtolEnergy));
        if (!rok) add(a, "rho", sci(a.rho), sci(e.rho), sci(rrel), sci(kRtolSpec), sci(kAtolSpec));
    }
    out->failures = static_cast<int32_t>(failures.size());
    snprintf(out->first_failure, sizeof(out->first_failure), "%s", first.c_str());
    if (!out->pass && csv_path && !failures.empty()) {
        std::stable_sort(failures.begin(), failures.end(),
                         [](const Failure& l, const Failure& r) { return l.cell < r.cell; });
        FILE* f = fopen(csv_path, "w");
        if (!f) {
            fprintf(stderr, "failed to open comparison-failures CSV file: %s\n", csv_path);
            return WR_ERR_IO;
        }
        fprintf(f, "cell,i,j,k,field,actual,reference,relative_error,rtol,atol\n");
        for (const auto& x : failures) {
            fprintf(f, "%d,%d,%d,%d,%s,%s,%s,%s,%s,%s\n", x.cell, x.i, x.j, x.k,
                    csv_escape(x.field).c_str(), csv_escape(x.actual).c_str(),
                    csv_escape(x.reference).c_str(), csv_escape(x.rel).c_str(),
                    csv_escape(x.rtol).c_str(), csv_escape(x.atol).c_str());
        }
        if (fclose(f) != 0) {
            fprintf(stderr, "failed to write comparison-failures CSV file: %s\n", csv_path);
            return WR_ERR_IO;
        }
    }
    return WR_SUCCESS;
}

}  // extern "C"
