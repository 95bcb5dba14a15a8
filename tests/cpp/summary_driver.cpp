// Test driver: builds a RunResult from per-batch values read on stdin and writes its JSON
// document to stdout through the host layer's own writer (no device work).
//
// stdin: "specular batches layers", then one line per batch:
//        "first_photon photons rd tt absorbed[0..layers-1] events anomalies"
#include <cstdio>
#include <iostream>

#include "run_result.h"

int main() {
    tmc::RunResult r;
    size_t batches = 0, layers = 0;
    if (!(std::cin >> r.specular >> batches >> layers)) return 2;
    r.spec.num_layers = static_cast<int32_t>(layers);
    for (size_t i = 0; i < layers; ++i) r.spec.layers[i] = tmc_layer{1.4, 1.0, 100.0, 0.9, 0.1};
    r.spec.n_above = r.spec.n_below = 1.0;
    r.spec.dz = r.spec.dr = 0.01;
    r.spec.nz = r.spec.nr = 1;
    r.precision = "float64";
    r.a_z.assign(1, 0.0);
    r.rd_r.assign(1, 0.0);
    r.t_r.assign(1, 0.0);
    r.a_z_se = r.rd_r_se = r.t_r_se = r.a_z;
    for (size_t b = 0; b < batches; ++b) {
        tmc::BatchResult br;
        std::cin >> br.first_photon >> br.photons >> br.rd >> br.tt;
        br.absorbed.resize(layers);
        for (double &a : br.absorbed) std::cin >> a;
        std::cin >> br.events >> br.anomalies;
        r.batches.push_back(br);
    }
    if (!std::cin) return 2;
    tmc::write_json(r, std::cout);
    std::cout << "\n";
    return 0;
}
