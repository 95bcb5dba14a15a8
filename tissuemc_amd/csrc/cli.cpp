// `tissuemc` command-line host: the reference's `run` surface (src/cli/main.cpp:18-204) on the
// MI355X engine. Same input (.mci), same summary lines, same JSON (a superset), same exit
// codes: 0 ok, 1 error or anomalies, 2 usage.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <string>

#include "run_result.h"

namespace {

void usage(std::FILE *to) {
    std::fprintf(to,
                 "usage:\n"
                 "  tissuemc --version\n"
                 "  tissuemc run <input.mci> [options]\n"
                 "\n"
                 "options for 'run':\n"
                 "  --photons N          override the photon count in the input file\n"
                 "  --batches B          number of batches for the standard error (default 10)\n"
                 "  --seed S             64-bit random seed (default 1)\n"
                 "  --engine E           hip (default): the MI355X engine\n"
                 "  --precision P        float32 (the GPU transport precision)\n"
                 "  --device D           HIP device ordinal (default 0)\n"
                 "  --arz                also tally A(r,z)\n"
                 "  --angles             also tally the exit-angle profiles (na bins of the .mci)\n"
                 "  --aggregate          combine a wave's equal-bin profile adds before the LDS atomics\n"
                 "  --time NT DT         also tally R_d(t), T_t(t): NT bins of DT ns of flight time\n"
                 "  --mco                write MCML's .mco file for each run, named as in the .mci\n"
                 "                       (implies --arz --angles)\n"
                 "  --no-profiles        diagnostic: skip the depth and radius profiles\n"
                 "  --waves-per-group W  launch shape (default automatic)\n"
                 "  --groups G           persistent workgroups (default one per compute unit)\n"
                 "  --chunk C            photon ids claimed per wave at a time\n"
                 "  --arz-tile Z R       A(r,z) bins kept in LDS (default automatic)\n"
                 "  --stream W           Philox counter word 3 (default 0, the reference streams)\n"
                 "  --autotune F         pick the launch shape from pilot runs at fraction F of the\n"
                 "                       photons (0 < F <= 1); results do not depend on the shape\n"
                 "  --output FILE        write full results as JSON\n");
}

unsigned long long parse_count(const char *name, const char *text) {
    char *end = nullptr;
    const unsigned long long v = std::strtoull(text, &end, 10);
    if (end == text || *end != '\0' || text[0] == '-') {
        std::fprintf(stderr, "tissuemc: %s expects a non-negative integer, got '%s'\n", name, text);
        std::exit(2);
    }
    return v;
}

double parse_fraction(const char *name, const char *text) {
    char *end = nullptr;
    const double v = std::strtod(text, &end);
    if (end == text || *end != '\0' || !(v > 0.0) || v > 1.0) {
        std::fprintf(stderr, "tissuemc: %s expects a fraction in (0, 1], got '%s'\n", name, text);
        std::exit(2);
    }
    return v;
}

int run_command(int argc, char **argv) {
    if (argc < 3) {
        usage(stderr);
        return 2;
    }
    const std::string input = argv[2];
    tmc_opts opt;
    tmc_default_opts(&opt);
    std::string output;
    double tune_fraction = 0.0;
    bool write_mco_files = false;
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "tissuemc: %s needs a value\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--photons") {
            opt.photons = parse_count("--photons", value("--photons"));
        } else if (a == "--batches") {
            opt.batches = static_cast<int32_t>(parse_count("--batches", value("--batches")));
        } else if (a == "--seed") {
            opt.seed = parse_count("--seed", value("--seed"));
        } else if (a == "--engine") {
            const std::string e = value("--engine");
            if (e != "hip") {
                std::fprintf(stderr,
                             "tissuemc: --engine must be hip, got '%s' (this build is the MI355X "
                             "engine; the float64 CPU engine is the test oracle)\n",
                             e.c_str());
                return 2;
            }
        } else if (a == "--precision") {
            const std::string p = value("--precision");
            if (p != "float32") {
                std::fprintf(stderr, "tissuemc: the GPU transports in float32, got '%s'\n", p.c_str());
                return 2;
            }
        } else if (a == "--device") {
            opt.device = static_cast<int32_t>(parse_count("--device", value("--device")));
        } else if (a == "--arz") {
            opt.flags |= TMC_FLAG_ARZ;
        } else if (a == "--angles") {
            opt.flags |= TMC_FLAG_ANGLES;
        } else if (a == "--aggregate") {
            opt.flags |= TMC_FLAG_AGGREGATE;
        } else if (a == "--time") {
            opt.flags |= TMC_FLAG_TIME;
            opt.nt = static_cast<int32_t>(parse_count("--time", value("--time")));
            const std::string dt = value("--time");
            char *end = nullptr;
            opt.dt = std::strtod(dt.c_str(), &end);
            if (end == dt.c_str() || *end != '\0' || !(opt.dt > 0.0)) {
                std::fprintf(stderr, "tissuemc: --time needs a bin count and a bin width > 0 (ns)\n");
                return 2;
            }
        } else if (a == "--mco") {
            write_mco_files = true;
            opt.flags |= TMC_FLAG_ARZ | TMC_FLAG_ANGLES;
        } else if (a == "--no-profiles") {
            opt.flags |= TMC_FLAG_NO_PROFILES;
        } else if (a == "--waves-per-group") {
            opt.waves_per_group = static_cast<int32_t>(parse_count(a.c_str(), value(a.c_str())));
        } else if (a == "--groups") {
            opt.groups = static_cast<int32_t>(parse_count(a.c_str(), value(a.c_str())));
        } else if (a == "--chunk") {
            opt.chunk = static_cast<int32_t>(parse_count(a.c_str(), value(a.c_str())));
        } else if (a == "--arz-tile") {
            opt.arz_tile_z = static_cast<int32_t>(parse_count(a.c_str(), value(a.c_str())));
            opt.arz_tile_r = static_cast<int32_t>(parse_count(a.c_str(), value(a.c_str())));
        } else if (a == "--stream") {
            opt.stream = static_cast<uint32_t>(parse_count(a.c_str(), value(a.c_str())));
        } else if (a == "--autotune") {
            tune_fraction = parse_fraction(a.c_str(), value(a.c_str()));
        } else if (a == "--output") {
            output = value("--output");
        } else {
            std::fprintf(stderr, "tissuemc: unknown option '%s'\n", a.c_str());
            usage(stderr);
            return 2;
        }
    }
    if ((opt.flags & (TMC_FLAG_ARZ | TMC_FLAG_ANGLES | TMC_FLAG_AGGREGATE | TMC_FLAG_TIME)) &&
        (opt.flags & TMC_FLAG_NO_PROFILES)) {
        std::fprintf(stderr, "tissuemc: --arz, --angles, --mco, --aggregate and --time cannot be "
                             "combined with --no-profiles\n");
        return 2;
    }

    const auto runs = tmc::parse_mci_file(input);
    std::ofstream json;
    if (!output.empty()) {
        json.open(output);
        if (!json) {
            std::fprintf(stderr, "tissuemc: cannot write '%s'\n", output.c_str());
            return 1;
        }
        json << "{\"runs\":[";
    }
    int status = 0;
    for (size_t r = 0; r < runs.size(); ++r) {
        tmc_opts run_opt = opt;
        if (tune_fraction > 0.0) {
            tmc_tune_report rep;
            const int code = tmc_autotune(&runs[r], &opt, tune_fraction, &run_opt, &rep);
            if (code == TMC_ERR_CONFIG) throw tmc::ConfigError(tmc_last_error());
            if (code != TMC_OK) throw std::runtime_error(tmc_last_error());
            const tmc_tune_entry &b = rep.tried[rep.best];
            std::printf("run %zu: autotune chose %d waves x %d workgroups (%d shapes, pilot %.3g, "
                        "%.2f ms",
                        r + 1, b.waves_per_group, b.groups, rep.n_tried, tune_fraction,
                        b.kernel_seconds * 1e3);
            if (rep.confirm_fraction > 0)
                std::printf("; confirmed at %.3g, %.2f ms", rep.confirm_fraction,
                            b.confirm_seconds * 1e3);
            std::printf(")\n");
        }
        const tmc::RunResult res = tmc::run_hip(runs[r], run_opt);
        const auto rd = res.rd(), tt = res.tt(), ab = res.absorbed();
        const auto cons = res.conservation_residual();
        std::printf("run %zu: %llu photons, %zu batches, %.2f s (%s, %s)\n", r + 1,
                    static_cast<unsigned long long>(res.photons()), res.batches.size(),
                    res.wall_seconds, res.engine.c_str(), res.precision.c_str());
        std::printf("  device %s, kernel %.3f s, %s totals\n", res.device.c_str(), res.kernel_seconds,
                    res.tally_precision.c_str());
        std::printf("  specular R_sp = %.6f\n", res.specular);
        std::printf("  diffuse  R_d  = %.6f +/- %.6f\n", rd.mean, rd.se);
        std::printf("  transmit T_t  = %.6f +/- %.6f\n", tt.mean, tt.se);
        std::printf("  absorbed  A   = %.6f +/- %.6f\n", ab.mean, ab.se);
        std::printf("  conservation residual = %+.2e +/- %.2e\n", cons.mean, cons.se);
        if (res.anomalies() != 0) {
            std::fprintf(stderr,
                         "tissuemc: %llu anomalies (non-finite weight, non-unit direction or step "
                         "limit); results are not trustworthy\n",
                         static_cast<unsigned long long>(res.anomalies()));
            status = 1;
        }
        if (json.is_open()) {
            if (r) json << ',';
            tmc::write_json(res, json);
        }
        if (write_mco_files) {
            const std::string name = runs[r].output_name[0] ? runs[r].output_name : "tissuemc.mco";
            std::ofstream mco(name);
            if (!mco) {
                std::fprintf(stderr, "tissuemc: cannot write '%s'\n", name.c_str());
                return 1;
            }
            tmc::write_mco(res, mco);
            std::printf("  wrote %s\n", name.c_str());
        }
    }
    if (json.is_open()) {
        json << "]}\n";
        json.flush();
        if (!json) {
            std::fprintf(stderr, "tissuemc: error writing '%s'\n", output.c_str());
            return 1;
        }
    }
    return status;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        if (argc >= 2 && std::strcmp(argv[1], "--version") == 0) {
            std::printf("tissuemc 0.1.0 (MI355X engine)\n");
            return 0;
        }
        if (argc >= 2 && std::strcmp(argv[1], "run") == 0) return run_command(argc, argv);
        usage(stderr);
        return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tissuemc: error: %s\n", e.what());
        return 1;
    }
}
