// wr_reproducer: command-line front end with the reference reproducer's
// interface and output (reproducer.cpp:8201-8277, 8591-8703) on the gfx950
// library.  The driver policy is the reference's HIP build's
// (gridwide-independent) unless --driver-policy asks for its CPU build's
// (legacy-local: each cell steps with its own free-fall dt, 8628-8648) or for
// the Mojo front end's gridwide-cta32 (one step controller per 32-cell tile,
// reproducer.mojo:2829-3093; prints the limiting cell as that front end does,
// 3663-3664).
//
//   wr_reproducer [--grid N] [--perturb|--no-perturb]
//                 [--compare-final-state FILE|--no-compare-final-state]
//                 [--driver-policy gridwide-independent|legacy-local|gridwide-cta32]
//                 [--write-final-state FILE]
//
// Writes final_states_grid<N>_hip.bin (_hip_legacy_local.bin, _hip_cta32.bin
// for the other policies; 172-byte records) for N > 1, prints the final state
// for N = 1, exit status 0 on success and 1 on failure.  --write-final-state
// FILE (the option of the reference's Mojo front end, reproducer.mojo:
// 3283-3340, 3529-3535) also writes the packed records to FILE and a sidecar
// FILE.manifest naming the record format, cell count, driver policy and
// network source, and refuses to overwrite either file.
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "warp_reactor.h"

namespace {

const char* const kSpeciesNames[WR_NUM_SPEC] = {"E",  "Hp", "H",  "Hm",  "Dp",  "D",   "H2p",
                                                "Dm", "H2", "HDp", "HD", "HEpp", "HEp", "HE"};
const char* const kBackend = "hip";

struct Options {
    int grid_dim = 64;
    bool perturb = true;
    bool show_help = false;
    int policy = WR_POLICY_GRIDWIDE;
    std::string compare_path;
    std::string write_path;
};

#ifndef WR_NETWORK_SHA256
#define WR_NETWORK_SHA256 "unknown"
#endif

void print_usage(const char* program) {
    fprintf(stderr,
            "usage: %s [--grid N] [--perturb|--no-perturb]"
            " [--compare-final-state FILE|--no-compare-final-state]"
            " [--driver-policy gridwide-independent|legacy-local|gridwide-cta32]"
            " [--write-final-state FILE]\n",
            program);
}

// Policy names: as --driver-policy spells them, or as the manifest tags them
// (the Mojo front end tags its own policy gridwide-cta32-v1, reproducer.mojo:3299).
const char* policy_name(int policy, bool manifest) {
    if (policy == WR_POLICY_LEGACY_LOCAL) return "legacy-local";
    if (policy == WR_POLICY_GRIDWIDE_CTA32) return manifest ? "gridwide-cta32-v1" : "gridwide-cta32";
    return "gridwide-independent";
}

bool file_exists(const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fclose(f);
    return true;
}

// The Mojo front end's final-state output (reproducer.mojo:3318-3340): the
// packed records in cell order, never over an existing file, and a sidecar
// manifest of key=value lines.
bool write_final_state_file(const std::vector<wr_collapse_state>& cells, int grid_dim,
                            const std::string& path, int policy) {
    const std::string manifest = path + ".manifest";
    if (file_exists(path) || file_exists(manifest)) {
        fprintf(stderr, "refusing to overwrite final-state output: %s\n", path.c_str());
        return false;
    }
    const int n = static_cast<int>(cells.size());
    std::vector<wr_packed_final_state> packed(cells.size());
    if (wr_pack_final_states(cells.data(), n, grid_dim, packed.data()) != WR_SUCCESS) {
        fprintf(stderr, "%s\n", wr_last_error());
        return false;
    }
    FILE* f = fopen(path.c_str(), "wb");
    const size_t bytes = packed.size() * sizeof(wr_packed_final_state);
    if (!f || fwrite(packed.data(), 1, bytes, f) != bytes || fclose(f) != 0) {
        fprintf(stderr, "failed to write final-state output: %s\n", path.c_str());
        return false;
    }
    FILE* m = fopen(manifest.c_str(), "w");
    if (!m) {
        fprintf(stderr, "failed to write final-state manifest: %s\n", manifest.c_str());
        return false;
    }
    fprintf(m,
            "format=packed-final-state-v1\ncell_count=%d\ngrid_dim=%d\n", n, grid_dim);
    // the Mojo manifest's tile size and policy tag for its own policy (3293-3307)
    if (policy == WR_POLICY_GRIDWIDE_CTA32) fprintf(m, "tile_cells=32\n");
    fprintf(m, "policy=%s\nnetwork_source_sha256=%s\nbackend=%s-gfx950\nabi=%d\n",
            policy_name(policy, true), WR_NETWORK_SHA256, kBackend, wr_abi_version());
    if (fclose(m) != 0) return false;
    printf("final-state output: %s (%d packed records, manifest %s)\n", path.c_str(), n,
           manifest.c_str());
    return true;
}

bool parse_positive_int(const char* text, int& value) {
    errno = 0;
    char* end = nullptr;
    const long parsed = strtol(text, &end, 10);
    if (errno != 0 || end == text || *end != '\0' || parsed <= 0 || parsed > INT_MAX) return false;
    value = static_cast<int>(parsed);
    return true;
}

bool parse_args(int argc, char** argv, Options& o) {
    const std::string prefix = "--compare-final-state=";
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--help" || arg == "-h") {
            o.show_help = true;
            return true;
        }
        if (arg == "--grid") {
            if (i + 1 >= argc || !parse_positive_int(argv[++i], o.grid_dim)) {
                print_usage(argv[0]);
                return false;
            }
            continue;
        }
        if (arg == "--perturb") {
            o.perturb = true;
            continue;
        }
        if (arg == "--no-perturb") {
            o.perturb = false;
            continue;
        }
        if (arg == "--compare-final-state") {
            if (i + 1 >= argc) {
                print_usage(argv[0]);
                return false;
            }
            o.compare_path = argv[++i];
            continue;
        }
        if (arg == "--driver-policy") {
            const std::string v = i + 1 < argc ? argv[++i] : "";
            if (v == "gridwide-independent") {
                o.policy = WR_POLICY_GRIDWIDE;
            } else if (v == "legacy-local") {
                o.policy = WR_POLICY_LEGACY_LOCAL;
            } else if (v == "gridwide-cta32") {
                o.policy = WR_POLICY_GRIDWIDE_CTA32;
            } else {
                print_usage(argv[0]);
                return false;
            }
            continue;
        }
        if (arg == "--write-final-state") {
            if (i + 1 >= argc || argv[i + 1][0] == '\0') {
                print_usage(argv[0]);
                return false;
            }
            o.write_path = argv[++i];
            continue;
        }
        if (arg == "--no-compare-final-state") {
            o.compare_path.clear();
            continue;
        }
        if (arg.rfind(prefix, 0) == 0) {
            o.compare_path = arg.substr(prefix.size());
            if (o.compare_path.empty()) {
                print_usage(argv[0]);
                return false;
            }
            continue;
        }
        print_usage(argv[0]);
        return false;
    }
    return true;
}

template <typename T>
long double median_of(std::vector<T> v, T& mn, T& mx) {
    std::sort(v.begin(), v.end());
    mn = v.front();
    mx = v.back();
    const size_t n = v.size();
    return (n % 2 == 1) ? static_cast<long double>(v[n / 2])
                        : (static_cast<long double>(v[n / 2 - 1]) + static_cast<long double>(v[n / 2])) / 2.0L;
}

void print_value_summary(const char* label, std::vector<long double> v) {
    long double mn, mx;
    const long double med = median_of(std::move(v), mn, mx);
    printf("%s: [%.17Lg, %.17Lg, %.17Lg]\n", label, mn, med, mx);
}

void print_scientific_summary(const char* label, std::vector<long double> v) {
    long double mn, mx;
    const long double med = median_of(std::move(v), mn, mx);
    printf("%s: [%.17Le, %.17Le, %.17Le]\n", label, mn, med, mx);
}

struct CounterSummary {
    uint64_t min, max;
    long double median;
};

CounterSummary summarize_counter(const std::vector<wr_collapse_state>& cells, size_t which) {
    std::vector<uint64_t> v;
    v.reserve(cells.size());
    for (const auto& c : cells) v.push_back(reinterpret_cast<const uint64_t*>(&c.stats)[which]);
    CounterSummary s;
    s.median = median_of(std::move(v), s.min, s.max);
    return s;
}

// The reference's CPU driver loop (reproducer.cpp:8628-8648) or the Mojo
// front end's gridwide-cta32 loop (reproducer.mojo:3604-3621) on the device:
// context API with that policy; HIP and API errors are reported on stderr and
// fail the run.  *limiting: the (dt, lowest cell) holder of the last prepare.
int run_policy(std::vector<wr_collapse_state>& cells, bool perturb, int policy, int& completed,
               int& limiting) {
    const int n = static_cast<int>(cells.size());
    wr_context* ctx = nullptr;
    int r = wr_context_create(0, n, &ctx);
    if (r == WR_SUCCESS) r = wr_set_policy(ctx, policy);
    if (r == WR_SUCCESS) r = wr_upload_cells(ctx, cells.data(), n);
    if (r == WR_SUCCESS) r = wr_run(ctx, perturb ? 1 : 0, 1000, 0, 0.0);
    wr_run_info info{};
    if (r > WR_ERR_HIP && ctx) {
        int d = wr_download_cells(ctx, cells.data(), n);
        if (d == WR_SUCCESS) d = wr_run_info_get(ctx, &info);
        if (d != WR_SUCCESS) r = d;
        completed = info.completed_global_steps;
        limiting = info.limiting_cell;
    }
    if (r <= WR_ERR_HIP) fprintf(stderr, "%s\n", wr_last_error());
    wr_context_destroy(ctx);
    return r;
}

int finish(bool pass, bool emit_status) {
    fflush(stderr);
    if (emit_status) printf("\n%s\n", pass ? "PASS" : "FAIL");
    fflush(stdout);
    return pass ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse_args(argc, argv, o)) return finish(false, false);
    if (o.show_help) {
        print_usage(argv[0]);
        return finish(true, false);
    }
    const long long n3 = static_cast<long long>(o.grid_dim) * o.grid_dim * o.grid_dim;
    if (n3 <= 0 || n3 > INT_MAX) {
        fprintf(stderr, "grid dimension is too large: %d\n", o.grid_dim);
        return finish(false, false);
    }
    const int num_cells = static_cast<int>(n3);
    std::vector<wr_collapse_state> cells(static_cast<size_t>(num_cells));
    wr_make_collapse_states(cells.data(), num_cells);

    int completed = 0, limiting = -1;
    const auto t0 = std::chrono::steady_clock::now();
    const int result = o.policy != WR_POLICY_GRIDWIDE
                           ? run_policy(cells, o.perturb, o.policy, completed, limiting)
                           : wr_run_cells(cells.data(), num_cells, o.perturb ? 1 : 0, &completed);
    const auto t1 = std::chrono::steady_clock::now();
    const double elapsed = std::chrono::duration<double>(t1 - t0).count();
    if (result != WR_SUCCESS) return finish(false, false);

    wr_integrator_stats total{};
    for (const auto& c : cells) {
        total.internal_steps += c.stats.internal_steps;
        total.rhs_calls += c.stats.rhs_calls;
        total.jacobian_calls += c.stats.jacobian_calls;
        total.decompositions += c.stats.decompositions;
        total.linear_solves += c.stats.linear_solves;
        total.accepted_steps += c.stats.accepted_steps;
        total.rejected_steps += c.stats.rejected_steps;
    }
    const wr_collapse_state& rep = cells.front();

    printf("Primordial chemistry collapse grid with ROS2S\n");
    printf("grid: %d^3 (%d cells)\n", o.grid_dim, num_cells);
    printf("perturbations: %s\n", o.perturb ? "enabled" : "disabled");
    printf("driver policy: %s\n", policy_name(o.policy, false));
    printf("completed global collapse steps: %d\n", completed);
    // no limiting cell: the Mojo front end prints its Int.MAX sentinel
    if (o.policy == WR_POLICY_GRIDWIDE_CTA32)
        printf("limiting cell: %lld\n", limiting >= 0 ? (long long)limiting : LLONG_MAX);
    if (num_cells == 1) {
        printf("representative cell completed steps: %d\n", rep.completed_steps);
        printf("representative cell physical time: %.17Le\n", static_cast<long double>(rep.time));
        printf("representative cell density driver: %.17g\n", rep.density_driver);
    } else {
        std::vector<long double> steps, times, drivers;
        for (const auto& c : cells) {
            steps.push_back(c.completed_steps);
            times.push_back(c.time);
            drivers.push_back(c.density_driver);
        }
        print_value_summary("cell completed steps", std::move(steps));
        print_scientific_summary("cell physical time", std::move(times));
        print_value_summary("cell density driver", std::move(drivers));
    }
    printf("wall time: %.17g s\n", elapsed);
    if (num_cells == 1) {
        printf("ROS2S internal steps: %llu\n", (unsigned long long)total.internal_steps);
        printf("ROS2S rhs calls: %llu\n", (unsigned long long)total.rhs_calls);
        printf("ROS2S jacobian calls: %llu\n", (unsigned long long)total.jacobian_calls);
        printf("ROS2S decompositions: %llu\n", (unsigned long long)total.decompositions);
        printf("ROS2S linear solves: %llu\n", (unsigned long long)total.linear_solves);
        printf("ROS2S accepted/rejected: %llu/%llu\n\n", (unsigned long long)total.accepted_steps,
               (unsigned long long)total.rejected_steps);
    } else {
        const char* labels[5] = {"ROS2S internal steps", "ROS2S rhs calls", "ROS2S jacobian calls",
                                 "ROS2S decompositions", "ROS2S linear solves"};
        for (size_t k = 0; k < 5; ++k) {
            const CounterSummary s = summarize_counter(cells, k);
            printf("%s: [%llu, %.17Lg, %llu]\n", labels[k], (unsigned long long)s.min, s.median,
                   (unsigned long long)s.max);
        }
        const CounterSummary a = summarize_counter(cells, 5), r = summarize_counter(cells, 6);
        printf("ROS2S accepted/rejected: [%llu/%llu, %.17Lg/%.17Lg, %llu/%llu]\n\n",
               (unsigned long long)a.min, (unsigned long long)r.min, a.median, r.median,
               (unsigned long long)a.max, (unsigned long long)r.max);
    }

    bool comparison_pass = true, comparison_done = false;
    if (!o.compare_path.empty()) {
        comparison_done = true;
        wr_comparison_summary sum;
        const std::string csv = std::string("final_state_comparison_failures_") + kBackend + ".csv";
        const int st = wr_compare_final_states(cells.data(), num_cells, o.grid_dim,
                                               o.compare_path.c_str(), csv.c_str(), &sum);
        if (st != WR_SUCCESS) {
            comparison_pass = false;
        } else {
            printf("final-state comparison file: %s\n", o.compare_path.c_str());
            printf("final-state comparison: %s (max deuterium-bearing species rel error %.17Le, max "
                   "non-deuterium-bearing species rel error %.17Le, max thermodynamic rel error "
                   "%.17Le, max rho rel error %.17Le)\n",
                   sum.pass ? "PASS" : "FAIL", (long double)sum.max_deuterium_species_rel_error,
                   (long double)sum.max_non_deuterium_species_rel_error,
                   (long double)sum.max_thermodynamic_rel_error, (long double)sum.max_rho_rel_error);
            if (!sum.pass && sum.first_failure[0]) printf("first comparison failure: %s\n", sum.first_failure);
            comparison_pass = sum.pass != 0;
        }
    }

    if (!o.write_path.empty() &&
        !write_final_state_file(cells, o.grid_dim, o.write_path, o.policy))
        return finish(false, comparison_done);
    if (o.grid_dim > 1) {
        const std::string file = "final_states_grid" + std::to_string(o.grid_dim) + "_" + kBackend +
                                 (o.policy == WR_POLICY_LEGACY_LOCAL    ? "_legacy_local"
                                  : o.policy == WR_POLICY_GRIDWIDE_CTA32 ? "_cta32"
                                                                         : "") +
                                 ".bin";
        if (wr_write_final_states(cells.data(), num_cells, o.grid_dim, file.c_str()) != WR_SUCCESS)
            return finish(false, comparison_done);
        printf("final states: %s (%d packed records, %zu bytes each)\n", file.c_str(), num_cells,
               sizeof(wr_packed_final_state));
    } else {
        printf("rho: %.17g\n", rep.current.rho);
        printf("T: %.17g\n", rep.current.T);
        printf("Eint: %.17g\n", rep.current.e);
        for (int n = 0; n < WR_NUM_SPEC; ++n) printf("%s: %.17g\n", kSpeciesNames[n], rep.current.xn[n]);
    }
    return finish(comparison_pass, comparison_done);
}
