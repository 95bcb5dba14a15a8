// C++ host layer over the C ABI: the run result, its batch statistics and its JSON form.
//
// RunResult keeps the reference's fields and meanings (include/tissuemc/result.hpp:15-59) so
// the JSON a run writes is a superset of the reference's schema (src/io/json_out.cpp:164-306).
// Additions: T_r, A(r,z), the exit-angle and time-of-flight profiles, per-bin standard errors of
// the profiles, boundary-step counts and the launch shape. The GPU's integer tallies (units of 2^-32) are
// summed as integers, then scaled.
#pragma once

#include <cstdint>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "json_doc.h"
#include "tissuemc_hip.h"

namespace tmc {

// Same roles as the reference's ConfigError (bad input or options, CLI exit 1) and
// runtime_error (device or API failure).
class ConfigError : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

struct BatchResult {
    uint64_t first_photon = 0;
    uint64_t photons = 0;
    double rd = 0, tt = 0;
    std::vector<double> absorbed;
    uint64_t events = 0, anomalies = 0, boundary_steps = 0;
};

struct Estimate {
    double mean = 0;
    double se = 0;  // NaN with fewer than two batches
};

struct RunResult {
    tmc_spec spec{};
    std::string engine = "hip";
    std::string precision = "float32";
    uint64_t seed = 0;
    int threads = 1;
    double specular = 0;
    std::vector<BatchResult> batches;
    std::vector<double> a_z, rd_r, t_r;           // summed over batches
    std::vector<double> a_z_se, rd_r_se, t_r_se;  // per-bin standard error of the per-photon mean
    std::vector<double> a_rz;                     // [nz*nr] or empty
    std::vector<double> rd_a, t_a;                // [na] exit-angle profiles, or empty
    std::vector<double> rd_ra, t_ra;              // [nr*na], row-major in radius, or empty
    std::vector<double> rd_t, t_t;                // [nt] time-of-flight profiles, or empty
    double dt_ns = 0;                             // their bin width
    double wall_seconds = 0, kernel_seconds = 0;
    std::string device;
    std::string tally_precision = "int64 fixed point (2^-32)";
    int rung = 0;
    bool profiles = true;
    int groups = 0, waves_per_group = 0, lds_bytes = 0, chunk = 0, arz_tile_z = 0, arz_tile_r = 0;

    uint64_t photons() const;
    uint64_t events() const;
    uint64_t anomalies() const;
    uint64_t boundary_steps() const;
    Estimate rd() const;
    Estimate tt() const;
    Estimate absorbed() const;
    Estimate conservation_residual() const;
};

// Batch-means estimate of E[value per photon] from per-batch sums (the estimator of the
// reference's json_out.cpp:37-64): mean = sum of the sums / sum of the photons; standard error =
// sample standard deviation of the per-batch means / sqrt(batches), NaN with fewer than two
// batches. The spread is accumulated in one pass with Welford's update.
class BatchMeans {
  public:
    void add(double sum, uint64_t photons);
    Estimate estimate() const;

  private:
    double total_ = 0.0, center_ = 0.0, m2_ = 0.0;
    uint64_t photons_ = 0, n_ = 0;
};

Estimate batch_estimate(const std::vector<double> &sums, const std::vector<uint64_t> &photons);

// Run one spec on the GPU through the C ABI; throws ConfigError / std::runtime_error.
RunResult run_hip(const tmc_spec &spec, const tmc_opts &opts);

// Parse an .mci file through the C ABI; throws ConfigError naming the file and line.
std::vector<tmc_spec> parse_mci_file(const std::string &path);

// The JSON document of one run (the reference's schema plus this engine's additions) and its
// text.
json::Value run_doc(const RunResult &run);
void write_json(const RunResult &run, std::ostream &out);

// MCML's ASCII output (.mco, "A1" format): the echoed input, then RAT, A_l, A_z, Rd_r, Rd_a,
// Tt_r, Tt_a, A_rz, Rd_ra, Tt_ra with MCML's normalizations (per photon; per cm of depth, cm^2 of
// ring area, sr of solid angle). Needs the A(r,z) and angle profiles; throws ConfigError without.
void write_mco(const RunResult &run, std::ostream &out);

}  // namespace tmc
