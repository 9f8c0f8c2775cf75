// Host/device building blocks shared by the gfx950 kernels and the host
// library: network constants, EOS, species helpers, the perturbation hash and
// the free-fall timestep (reproducer.cpp:199-339, 7548-7644).  FP64, operand
// order as in the reference; compile with -ffp-contract=off.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WR_HD __host__ __device__ __forceinline__
#else
#define WR_HD inline
#endif

namespace wr {


constexpr int kNumSpec = 14;
constexpr int kNeqs = 15;
constexpr int kSuccess = 1, kBadInputs = -1, kDtUnderflow = -2, kTooManySteps = -4,
              kLuError = -7;

// ---- network constants (reproducer.cpp:199-253) -----------------------------
WR_HD constexpr double species_mass(int n) {
    return n == 0 ? 9.10938188e-28 : n == 1 ? 1.67262158e-24 : n == 2 ? 1.67353251819e-24
         : n == 3 ? 1.67444345638e-24 : n == 4 ? 3.34512158e-24 : n == 5 ? 3.34603251819e-24
         : n == 6 ? 3.34615409819e-24 : n == 7 ? 3.34694345638e-24 : n == 8 ? 3.34706503638e-24
         : n == 9 ? 5.01865409819e-24 : n == 10 ? 5.01956503638e-24 : n == 11 ? 6.69024316e-24
         : n == 12 ? 6.69115409819e-24 : 6.69206503638e-24;
}
WR_HD constexpr double species_gamma(int n) {
    return (n == 6 || n == 8 || n == 9 || n == 10) ? 1.4 : 5.0 / 3.0;
}
constexpr double kAvogadro = 6.02214076e+23;
constexpr double kBoltzmann = 1.3806490000000002e-16;
constexpr double kProtonMass = 1.67262192595e-24;
constexpr double kGasConstant = kAvogadro * kBoltzmann;
constexpr double kFloor = 1.0e-100;
constexpr double kGrav = 6.674e-8;
constexpr double kPi = 3.141592653589793238462643383279502884;

// std::max / std::min exactly as the reference evaluates them.
WR_HD double rmax(double a, double b) { return (a < b) ? b : a; }
WR_HD double rmin(double a, double b) { return (b < a) ? b : a; }

// ---- EOS (reproducer.cpp:270-339) -------------------------------------------
WR_HD double density_of(const double (&xn)[kNumSpec]) {
    double rho = 0.0;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) rho += xn[n] * species_mass(n);
    return rho;
}

// eos_sums_from_number_densities: returns sum_gammasinv * gasconstant * sum_abarinv
// pieces as the reference combines them.
WR_HD void eos_sums(const double (&xn)[kNumSpec], double& sa, double& sg) {
    double rhotot = 0.0;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) rhotot += xn[n] * species_mass(n);
    sa = 0.0;
    sg = 0.0;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) {
        sa += xn[n];
        sg += (xn[n] * kProtonMass / rhotot) * (1.0 / (species_gamma(n) - 1.0));
    }
    sa *= kProtonMass / rhotot;
    sg /= sa;
}

// eos_re: T from e (reproducer.cpp:311-314)
WR_HD double eos_temperature(const double (&xn)[kNumSpec], double e) {
    double sa, sg;
    eos_sums(xn, sa, sg);
    return e / (sg * kGasConstant * sa);
}

// eos_rt: e from T (reproducer.cpp:306-309)
WR_HD double eos_energy(const double (&xn)[kNumSpec], double T) {
    double sa, sg;
    eos_sums(xn, sa, sg);
    return sg * sa * kGasConstant * T;
}

// balance_charge (reproducer.cpp:316-319)
WR_HD void balance_charge(double (&x)[kNumSpec]) {
    x[0] = -x[3] - x[7] + x[1] + x[12] + x[6] + x[4] + x[9] + 2.0 * x[11];
}

// normalize_number_densities_to_density (reproducer.cpp:321-332)
WR_HD void normalize(double (&xn)[kNumSpec], double rho) {
    double mf[kNumSpec];
    double sum = 0.0;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) {
        mf[n] = species_mass(n) * xn[n] / rho;
        sum += mf[n];
    }
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) {
        mf[n] /= sum;
        xn[n] = mf[n] * rho / species_mass(n);
    }
}

// floor_and_normalize_number_densities (reproducer.cpp:334-339)
WR_HD void floor_normalize(double (&xn)[kNumSpec], double rho) {
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] = rmax(xn[n], kFloor);
    normalize(xn, rho);
}

// burn_state_from_y (reproducer.cpp:7029-7038): floored X and T(e, X).
WR_HD double state_from_y(const double (&y)[kNeqs], double (&X)[kNumSpec]) {
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) X[n] = rmax(y[n], kFloor);
    return eos_temperature(X, y[kNumSpec]);
}

// ---- per-cell driver pieces (reproducer.cpp:7548-7644) ---------------------------
WR_HD uint64_t splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31U);
}

WR_HD double perturbation_factor(int cell, int step) {
    const uint64_t seed = (static_cast<uint64_t>(static_cast<uint32_t>(cell)) << 32U) ^
                          (static_cast<uint64_t>(static_cast<uint32_t>(step)) << 16U);
    const uint64_t bits = splitmix64(seed) >> 11U;
    const double unit = static_cast<double>(bits) * (1.0 / static_cast<double>(1ULL << 53U));
    return 1.0 + 0.2 * (2.0 * unit - 1.0);
}

WR_HD bool valid_positive(double v) { return v > 0.0 && isfinite(v); }

// collapse_timestep (reproducer.cpp:7633-7644) given the number densities.
WR_HD double collapse_timestep(const double (&xn)[kNumSpec]) {
    const double rho = density_of(xn);
    if (!valid_positive(rho)) return -1.0;
    const double tff = sqrt(kPi * 3.0 / (32.0 * rho * kGrav));
    if (!valid_positive(tff)) return -1.0;
    const double dt = 1.0e-1 * tff;
    return valid_positive(dt) ? dt : -1.0;
}

// Post-step sequence (reproducer.cpp:7773-7776, also 7576-7579):
// floor/normalize, charge balance, floor/normalize, T from e.
WR_HD double post_step(double (&xn)[kNumSpec], double rho, double e) {
    floor_normalize(xn, rho);
    balance_charge(xn);
    floor_normalize(xn, rho);
    return eos_temperature(xn, e);
}

}  // namespace wr
