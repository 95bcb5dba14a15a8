// Internal C++ interface between the C ABI (capi.cpp) and the HIP engine (transport.hip).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include "tissuemc_hip.h"

namespace tmc {

constexpr int kMaxLayers = TMC_MAX_LAYERS;

// An error with the C-ABI status it maps to (TMC_ERR_CONFIG / _DEVICE / _RUNTIME).
class EngineError : public std::runtime_error {
  public:
    EngineError(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

// Throws EngineError(TMC_ERR_CONFIG) naming the offending field, as the reference's
// validate() does (include/tissuemc/config.hpp:38-66).
void validate_spec(const tmc_spec &spec);

// Runs the transport kernel for one run and fills `out` (capi.cpp checks pointers first).
void run_transport(const tmc_spec &spec, const tmc_opts &opts, tmc_out *out);
// The same with every device buffer taken from a static per-device scratch array of
// static_scratch_bytes(): no device allocation at all.
void run_transport_static(const tmc_spec &spec, const tmc_opts &opts, tmc_out *out);
size_t static_scratch_bytes();

// Launch-shape choices shared by single and batched runs (0 / -1 = automatic).
struct tmc_tuning {
    int waves_per_group = 0, groups = 0, chunk = 0, arz_tile_z = -1, arz_tile_r = -1;
    uint64_t step_limit = 0;
    size_t lds_budget = 0;  // LDS one workgroup may use; 0 = the device's per-workgroup limit
};

// Batched runs (opaque to callers of the C ABI).
struct BatchedPlan;
BatchedPlan *batched_create(int n_stacks, const tmc_batched_opts &opts);
void batched_execute(BatchedPlan &plan, const tmc_stack *stacks, int n_stacks, uint64_t seed,
                     tmc_batched_out *out);
void batched_destroy(BatchedPlan *plan);

// Launch-shape autotuning by pilot runs (tmc_autotune, tmc_batched_autotune).
void autotune_single(const tmc_spec &spec, const tmc_opts &opts, double pilot_fraction,
                     tmc_opts *tuned, tmc_tune_report *report);
void batched_autotune(BatchedPlan &plan, const tmc_stack *stacks, int n_stacks, uint64_t seed,
                      double pilot_fraction, tmc_tune_report *report);

// The surrogate model (mlp.hip).
class MlpModel;
MlpModel *mlp_create(const tmc_mlp_config &c);
void mlp_destroy(MlpModel *m);
void mlp_set_data(MlpModel &m, const float *x, const float *y, int64_t n, const float *w);
double mlp_train(MlpModel &m, int64_t steps, float *losses);
void mlp_gradients(MlpModel &m, const int64_t *rows, float *grads, float *loss);
double mlp_predict(MlpModel &m, const float *x, int64_t n, float *mean, float *spread, float *per_head);
void mlp_answer(MlpModel &m, const tmc_surrogate_gate &gate, const tmc_stack *stacks, int64_t n,
                tmc_surrogate_answers *out);
void mlp_get_params(MlpModel &m, float *out);
void mlp_set_params(MlpModel &m, const float *in);
void mlp_set_lr(MlpModel &m, float lr);
long long mlp_param_count(const MlpModel &m);
uint64_t mlp_flops_per_step(const MlpModel &m);
uint64_t mlp_flops_per_query(const MlpModel &m);

// Device-side Philox stream, for the parity hook.
void philox_stream(uint64_t seed, uint64_t id, uint32_t word3, int64_t n, uint32_t *out, int device);

bool device_available();

}  // namespace tmc
