// Random layer stacks for the batched workload and the surrogate's data (tmc_generate_stacks).
// Stack k is a pure function of (seed, k), so any range of the sequence can be generated on its
// own and two ranges never overlap.
#include <cmath>
#include <cstring>

#include "device_rng.h"
#include "engine.h"

namespace {

constexpr uint32_t kStackTag = 0x53544B31u;  // counter word 3 of the generator's blocks

double unit(uint32_t x) { return (static_cast<double>(x) + 0.5) * 0x1p-32; }
double log_uniform(double lo, double hi, uint32_t x) {
    return std::exp(std::log(lo) + unit(x) * (std::log(hi) - std::log(lo)));
}
double uniform(double lo, double hi, uint32_t x) { return lo + unit(x) * (hi - lo); }

}  // namespace

extern "C" {

void tmc_default_stack_dist(tmc_stack_dist *d) {
    if (!d) return;
    d->min_layers = 1;
    d->max_layers = 5;
    d->mua_lo = 0.01;
    d->mua_hi = 10.0;
    d->mus_lo = 10.0;
    d->mus_hi = 500.0;
    d->g_lo = 0.6;
    d->g_hi = 0.95;
    d->n_lo = 1.33;
    d->n_hi = 1.5;
    d->d_lo = 0.005;
    d->d_hi = 0.5;
    d->photons_lo = 1e4;
    d->photons_hi = 1e6;
    d->n_ambient = 1.0;
}

int tmc_generate_stacks(uint64_t seed, uint64_t first_id, int32_t count, const tmc_stack_dist *dist,
                        tmc_stack *out) {
    if (!dist || (!out && count > 0) || count < 0) return TMC_ERR_CONFIG;
    const tmc_stack_dist &d = *dist;
    if (d.min_layers < 1 || d.max_layers > TMC_MAX_LAYERS || d.min_layers > d.max_layers ||
        !(d.photons_lo >= 1) || !(d.photons_hi >= d.photons_lo) || !(d.mua_lo > 0) ||
        !(d.mus_lo > 0) || !(d.d_lo > 0))
        return TMC_ERR_CONFIG;
    const uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
    for (int32_t i = 0; i < count; ++i) {
        const uint64_t id = first_id + static_cast<uint64_t>(i);
        const uint32_t c0 = static_cast<uint32_t>(id), c1 = static_cast<uint32_t>(id >> 32);
        tmc_stack &s = out[i];
        std::memset(&s, 0, sizeof s);
        const tmc::Block4 h = tmc::philox4x32_10(c0, c1, 0u, kStackTag, k0, k1);
        const int span = d.max_layers - d.min_layers + 1;
        int nl = d.min_layers + static_cast<int>(unit(h.v0) * span);
        if (nl > d.max_layers) nl = d.max_layers;
        s.num_layers = nl;
        s.photons = static_cast<uint64_t>(std::llround(log_uniform(d.photons_lo, d.photons_hi, h.v1)));
        s.n_above = s.n_below = d.n_ambient;
        for (int l = 0; l < nl; ++l) {
            const tmc::Block4 p = tmc::philox4x32_10(c0, c1, 1u + l, kStackTag, k0, k1);
            const tmc::Block4 q = tmc::philox4x32_10(c0, c1, 17u + l, kStackTag, k0, k1);
            tmc_layer &L = s.layers[l];
            L.mua = log_uniform(d.mua_lo, d.mua_hi, p.v0);
            L.mus = log_uniform(d.mus_lo, d.mus_hi, p.v1);
            L.g = uniform(d.g_lo, d.g_hi, p.v2);
            L.n = uniform(d.n_lo, d.n_hi, p.v3);
            L.d = log_uniform(d.d_lo, d.d_hi, q.v0);
        }
    }
    return TMC_OK;
}

}  // extern "C"
