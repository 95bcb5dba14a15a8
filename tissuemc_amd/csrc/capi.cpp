// The C ABI (include/tissuemc_hip.h): error mapping, run validation and the .mci reader.
//
// The .mci reader accepts the MCML input format the reference reads (src/io/mci.cpp:97-171),
// with the same messages; run validation rules and messages follow
// include/tissuemc/config.hpp:38-66.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "engine.h"

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

template <class F> int guarded(F &&f) {
    try {
        g_last_error.clear();
        f();
        return TMC_OK;
    } catch (const tmc::EngineError &e) {
        return set_error(e.code(), e.what());
    } catch (const std::bad_alloc &) {
        return set_error(TMC_ERR_RUNTIME, "out of host memory");
    } catch (const std::exception &e) {
        return set_error(TMC_ERR_RUNTIME, e.what());
    }
}

[[noreturn]] void config_fail(const std::string &msg) { throw tmc::EngineError(TMC_ERR_CONFIG, msg); }

// ------------------------------------------------------------------ .mci reader
//
// An .mci file is a fixed sequence of records, one per content line ('#' starts a comment;
// lines with nothing else are skipped):
//   version | run count | per run: output name + format, photons, dz dr, nz nr na, layer count,
//   n above, one line per layer (n mua mus g d), n below.
// The reader walks the text once with a cursor that yields content lines as views into the text,
// and a record declares how many fields it has and what it is called in messages. Field values
// are converted by kind: reals must be finite, counts must be base-10 integers (MCML's manual
// asks for "5", not "5.0"). Every message names the line it is about; the wording is the
// reference parser's (src/io/mci.cpp), which users' scripts may match.

[[noreturn]] void fail_at(int line, const std::string &msg) {
    config_fail("line " + std::to_string(line) + ": " + msg);
}

struct Record {
    int line = 0;
    std::vector<std::string_view> fields;

    std::string text(size_t i) const { return std::string(fields[i]); }

    double real(size_t i, const std::string &name) const {
        const std::string t = text(i);
        char *end = nullptr;
        errno = 0;
        const double v = std::strtod(t.c_str(), &end);
        const bool whole = end != t.c_str() && *end == '\0';
        if (!whole || errno == ERANGE || !std::isfinite(v))
            fail_at(line, name + ": '" + t + "' is not a finite number");
        return v;
    }

    long long count(size_t i, const std::string &name) const {
        const std::string t = text(i);
        char *end = nullptr;
        errno = 0;
        const long long v = std::strtoll(t.c_str(), &end, 10);
        if (end == t.c_str() || *end != '\0' || errno == ERANGE)
            fail_at(line, name + ": '" + t + "' is not an integer");
        return v;
    }
};

class RecordCursor {
  public:
    explicit RecordCursor(std::string_view text) : rest_(text) { advance(); }

    // The next record, which must hold exactly `fields` values.
    Record expect(size_t fields, const char *what) {
        if (!has_next_)
            fail_at(last_line_, std::string("unexpected end of file, expected ") + what);
        Record r = std::move(next_);
        advance();
        if (r.fields.size() != fields)
            fail_at(r.line, std::string(what) + ": expected " + std::to_string(fields) +
                                " value(s), found " + std::to_string(r.fields.size()));
        return r;
    }

    bool at_end() const { return !has_next_; }
    int next_line() const { return next_.line; }

  private:
    // Moves to the next line with content, splitting it into whitespace-separated fields.
    void advance() {
        has_next_ = false;
        while (!rest_.empty() && !has_next_) {
            const size_t nl = rest_.find('\n');
            std::string_view line = rest_.substr(0, nl);
            rest_ = nl == std::string_view::npos ? std::string_view() : rest_.substr(nl + 1);
            ++line_no_;
            line = line.substr(0, line.find('#'));
            Record r;
            r.line = line_no_;
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
                size_t j = i;
                while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
                if (j > i) r.fields.push_back(line.substr(i, j - i));
                i = j;
            }
            if (!r.fields.empty()) {
                next_ = std::move(r);
                has_next_ = true;
                last_line_ = line_no_;
            }
        }
    }

    std::string_view rest_;
    int line_no_ = 0, last_line_ = 0;
    Record next_;
    bool has_next_ = false;
};

constexpr long long kMaxBins = 1000000;

tmc_spec read_run(RecordCursor &in) {
    tmc_spec s;
    std::memset(&s, 0, sizeof s);
    const Record out = in.expect(2, "output filename and format");
    std::snprintf(s.output_name, sizeof s.output_name, "%s", out.text(0).c_str());
    if (out.fields[1] != "A")
        fail_at(out.line, "output format must be 'A' (ASCII), found '" + out.text(1) + "'");

    const Record np = in.expect(1, "number of photons");
    const long long photons = np.count(0, "number of photons");
    if (photons < 1) fail_at(np.line, "number of photons must be >= 1");
    s.photons = static_cast<uint64_t>(photons);

    const Record widths = in.expect(2, "dz and dr");
    s.dz = widths.real(0, "dz");
    s.dr = widths.real(1, "dr");

    const Record grid = in.expect(3, "grid counts (nz nr na)");
    long long counts[3];
    const char *names[3] = {"nz", "nr", "na"};
    for (size_t k = 0; k < 3; ++k) counts[k] = grid.count(k, names[k]);
    for (long long c : counts)
        if (c < 1 || c > kMaxBins) fail_at(grid.line, "grid counts must be integers in [1, 1000000]");
    s.nz = static_cast<int32_t>(counts[0]);
    s.nr = static_cast<int32_t>(counts[1]);
    s.na = static_cast<int32_t>(counts[2]);

    const Record nl = in.expect(1, "number of layers");
    const long long layers = nl.count(0, "number of layers");
    if (layers < 1 || layers > tmc::kMaxLayers)
        fail_at(nl.line, "number of layers must be in [1, " + std::to_string(tmc::kMaxLayers) + "]");
    s.num_layers = static_cast<int32_t>(layers);

    s.n_above = in.expect(1, "refractive index above").real(0, "n above");
    static const char *const kLayerFields[5] = {" n", " mua", " mus", " g", " thickness"};
    for (int i = 0; i < s.num_layers; ++i) {
        const Record L = in.expect(5, "layer (n mua mus g d)");
        const std::string tag = "layer " + std::to_string(i + 1);
        double v[5];
        for (size_t k = 0; k < 5; ++k) v[k] = L.real(k, tag + kLayerFields[k]);
        s.layers[i] = tmc_layer{v[0], v[1], v[2], v[3], v[4]};
    }
    const Record below = in.expect(1, "refractive index below");
    s.n_below = below.real(0, "n below");
    try {
        tmc::validate_spec(s);
    } catch (const tmc::EngineError &e) {
        fail_at(below.line, std::string("invalid run: ") + e.what());
    }
    return s;
}

std::vector<tmc_spec> parse_text(std::string_view text) {
    RecordCursor in(text);
    const Record version = in.expect(1, "file version");
    if (version.real(0, "file version") != 1.0)
        fail_at(version.line, "unsupported file version '" + version.text(0) + "' (expected 1.0)");
    const Record nr = in.expect(1, "number of runs");
    const long long n = nr.count(0, "number of runs");
    if (n < 1) fail_at(nr.line, "number of runs must be >= 1");
    std::vector<tmc_spec> runs;
    for (long long r = 0; r < n; ++r) runs.push_back(read_run(in));
    if (!in.at_end()) fail_at(in.next_line(), "unexpected extra content after the last run");
    return runs;
}

void copy_runs(const std::vector<tmc_spec> &runs, tmc_spec *out, int32_t max_runs, int32_t *n_runs) {
    if (n_runs) *n_runs = static_cast<int32_t>(runs.size());
    if (static_cast<int64_t>(runs.size()) > max_runs)
        config_fail("the file holds " + std::to_string(runs.size()) + " runs; room for " +
                    std::to_string(max_runs));
    for (size_t i = 0; i < runs.size(); ++i) out[i] = runs[i];
}

}  // namespace

namespace tmc {

void validate_spec(const tmc_spec &run) {
    if (run.photons == 0) config_fail("number of photons must be > 0");
    if (!(run.dz > 0) || !(run.dr > 0)) config_fail("dz and dr must be > 0");
    if (run.nz < 1 || run.nr < 1 || run.na < 1) config_fail("grid bin counts must be >= 1");
    if (!(run.n_above >= 1.0) || !(run.n_below >= 1.0))
        config_fail("ambient refractive indices must be >= 1");
    if (run.num_layers < 1) config_fail("at least one layer is required");
    if (run.num_layers > kMaxLayers)
        config_fail("at most " + std::to_string(kMaxLayers) + " layers are supported");
    for (int i = 0; i < run.num_layers; ++i) {
        const tmc_layer &L = run.layers[i];
        const std::string where = "layer " + std::to_string(i + 1) + ": ";
        if (!(L.n >= 1.0)) config_fail(where + "n must be >= 1");
        if (!(L.mua >= 0) || !(L.mus >= 0)) config_fail(where + "mua and mus must be >= 0");
        if (!(L.mua + L.mus > 0)) config_fail(where + "clear layers (mua = mus = 0) are not supported");
        if (!(L.g > -1.0 && L.g < 1.0)) config_fail(where + "g must be in (-1, 1)");
        if (!(L.d > 0) || !std::isfinite(L.d)) config_fail(where + "thickness must be finite and > 0");
    }
}

}  // namespace tmc

extern "C" {

int tmc_hip_available(void) { return tmc::device_available() ? 1 : 0; }

const char *tmc_last_error(void) { return g_last_error.c_str(); }

double tmc_fixed_point_unit(void) { return 0x1p-32; }

void tmc_default_opts(tmc_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->seed = 1;
    o->batches = 10;
    o->arz_tile_z = -1;
    o->arz_tile_r = -1;
}

int tmc_validate_spec(const tmc_spec *spec) {
    return guarded([&] {
        if (!spec) config_fail("spec is NULL");
        tmc::validate_spec(*spec);
    });
}

int tmc_hip_run(const tmc_spec *spec, const tmc_opts *opts, tmc_out *out) {
    return guarded([&] {
        if (!spec || !opts || !out) config_fail("spec, opts and out must not be NULL");
        if (!out->rd || !out->tt || !out->absorbed || !out->events || !out->anomalies)
            config_fail("tmc_out: rd, tt, absorbed, events and anomalies are required");
        tmc::run_transport(*spec, *opts, out);
    });
}

int tmc_hip_run_static(const tmc_spec *spec, const tmc_opts *opts, tmc_out *out) {
    return guarded([&] {
        if (!spec || !opts || !out) config_fail("spec, opts and out must not be NULL");
        if (!out->rd || !out->tt || !out->absorbed || !out->events || !out->anomalies)
            config_fail("tmc_out: rd, tt, absorbed, events and anomalies are required");
        tmc::run_transport_static(*spec, *opts, out);
    });
}

uint64_t tmc_static_scratch_bytes(void) { return tmc::static_scratch_bytes(); }

int tmc_parse_mci_text(const char *text, tmc_spec *runs, int32_t max_runs, int32_t *n_runs) {
    return guarded([&] {
        if (!text || !runs) config_fail("text and runs must not be NULL");
        copy_runs(parse_text(text), runs, max_runs, n_runs);
    });
}

int tmc_parse_mci_file(const char *path, tmc_spec *runs, int32_t max_runs, int32_t *n_runs) {
    return guarded([&] {
        if (!path || !runs) config_fail("path and runs must not be NULL");
        std::ifstream f(path);
        if (!f) config_fail(std::string("cannot open '") + path + "'");
        std::stringstream ss;
        ss << f.rdbuf();
        try {
            copy_runs(parse_text(ss.str()), runs, max_runs, n_runs);
        } catch (const tmc::EngineError &e) {
            throw tmc::EngineError(e.code(), std::string(path) + ": " + e.what());
        }
    });
}

struct tmc_batched_plan {
    tmc::BatchedPlan *impl;
};

int tmc_batched_create(int32_t n_stacks, const tmc_batched_opts *opts, tmc_batched_plan **plan) {
    return guarded([&] {
        if (!opts || !plan) config_fail("opts and plan must not be NULL");
        *plan = nullptr;
        auto *h = new tmc_batched_plan{nullptr};
        try {
            h->impl = tmc::batched_create(n_stacks, *opts);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

int tmc_autotune(const tmc_spec *spec, const tmc_opts *opts, double pilot_fraction,
                 tmc_opts *tuned, tmc_tune_report *report) {
    return guarded([&] {
        if (!spec || !opts || !tuned) config_fail("spec, opts and tuned must not be NULL");
        tmc::autotune_single(*spec, *opts, pilot_fraction, tuned, report);
    });
}

int tmc_batched_autotune(tmc_batched_plan *plan, const tmc_stack *stacks, int32_t n_stacks,
                         uint64_t seed, double pilot_fraction, tmc_tune_report *report) {
    return guarded([&] {
        if (!plan || !plan->impl || !stacks) config_fail("plan and stacks must not be NULL");
        tmc::batched_autotune(*plan->impl, stacks, n_stacks, seed, pilot_fraction, report);
    });
}

int tmc_batched_execute(tmc_batched_plan *plan, const tmc_stack *stacks, int32_t n_stacks,
                        uint64_t seed, tmc_batched_out *out) {
    return guarded([&] {
        if (!plan || !plan->impl || !stacks || !out) config_fail("plan, stacks and out must not be NULL");
        if (!out->rd || !out->tt || !out->absorbed || !out->events || !out->anomalies)
            config_fail("tmc_batched_out: rd, tt, absorbed, events and anomalies are required");
        tmc::batched_execute(*plan->impl, stacks, n_stacks, seed, out);
    });
}

void tmc_batched_destroy(tmc_batched_plan *plan) {
    if (!plan) return;
    tmc::batched_destroy(plan->impl);
    delete plan;
}

int tmc_hip_run_batched(const tmc_stack *stacks, int32_t n_stacks, const tmc_batched_opts *opts,
                        tmc_batched_out *out) {
    tmc_batched_plan *plan = nullptr;
    int rc = tmc_batched_create(n_stacks, opts, &plan);
    if (rc != TMC_OK) return rc;
    rc = tmc_batched_execute(plan, stacks, n_stacks, opts->seed, out);
    const std::string err = g_last_error;
    tmc_batched_destroy(plan);
    g_last_error = err;
    return rc;
}

struct tmc_mlp {
    tmc::MlpModel *impl;
};

void tmc_mlp_default_config(tmc_mlp_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->heads = 4;
    c->inputs = 30;
    c->hidden = 512;
    c->depth = 4;
    c->outputs = 202;
    c->batch = 4096;
    c->lr = 1e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
    c->weight_decay = 0.0f;
    c->seed = 1;
    c->device = 0;
}

int tmc_mlp_create(const tmc_mlp_config *config, tmc_mlp **model) {
    return guarded([&] {
        if (!config || !model) config_fail("config and model must not be NULL");
        *model = nullptr;
        auto *h = new tmc_mlp{nullptr};
        try {
            h->impl = tmc::mlp_create(*config);
        } catch (...) {
            delete h;
            throw;
        }
        *model = h;
    });
}

void tmc_mlp_destroy(tmc_mlp *model) {
    if (!model) return;
    tmc::mlp_destroy(model->impl);
    delete model;
}

int64_t tmc_mlp_param_count(const tmc_mlp *model) {
    return model && model->impl ? tmc::mlp_param_count(*model->impl) : -1;
}

#define TMC_NEED_MODEL() \
    if (!model || !model->impl) config_fail("model must not be NULL")

int tmc_mlp_set_data(tmc_mlp *model, const float *x, const float *y, int64_t n, const float *w) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!x || !y) config_fail("x and y must not be NULL");
        tmc::mlp_set_data(*model->impl, x, y, n, w);
    });
}

int tmc_mlp_train(tmc_mlp *model, int64_t steps, float *losses, double *seconds) {
    return guarded([&] {
        TMC_NEED_MODEL();
        const double s = tmc::mlp_train(*model->impl, steps, losses);
        if (seconds) *seconds = s;
    });
}

int tmc_mlp_gradients(tmc_mlp *model, const int64_t *rows, float *grads, float *loss) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!rows || !grads) config_fail("rows and grads must not be NULL");
        tmc::mlp_gradients(*model->impl, rows, grads, loss);
    });
}

int tmc_mlp_predict(tmc_mlp *model, const float *x, int64_t n, float *mean, float *spread,
                    float *per_head, double *seconds) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!x && n > 0) config_fail("x must not be NULL");
        if (n < 0) config_fail("n must be >= 0");
        const double s = tmc::mlp_predict(*model->impl, x, n, mean, spread, per_head);
        if (seconds) *seconds = s;
    });
}

int tmc_mlp_answer(tmc_mlp *model, const tmc_surrogate_gate *gate, const tmc_stack *stacks,
                   int64_t n, tmc_surrogate_answers *out) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!gate || !out) config_fail("gate and out must not be NULL");
        if (n < 0) config_fail("n must be >= 0");
        if (n > 0 && (!stacks || !out->rd || !out->tt || !out->route))
            config_fail("stacks and out->rd, tt, route must not be NULL");
        if (!gate->y_mean || !gate->y_std) config_fail("gate->y_mean and y_std must not be NULL");
        tmc::mlp_answer(*model->impl, *gate, stacks, n, out);
    });
}

int tmc_mlp_get_params(tmc_mlp *model, float *params) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!params) config_fail("params must not be NULL");
        tmc::mlp_get_params(*model->impl, params);
    });
}

int tmc_mlp_set_params(tmc_mlp *model, const float *params) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!params) config_fail("params must not be NULL");
        tmc::mlp_set_params(*model->impl, params);
    });
}

int tmc_mlp_set_lr(tmc_mlp *model, float lr) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (!(lr >= 0.0f)) config_fail("the learning rate must be >= 0");
        tmc::mlp_set_lr(*model->impl, lr);
    });
}

int tmc_mlp_flops(const tmc_mlp *model, uint64_t *per_step, uint64_t *per_query) {
    return guarded([&] {
        TMC_NEED_MODEL();
        if (per_step) *per_step = tmc::mlp_flops_per_step(*model->impl);
        if (per_query) *per_query = tmc::mlp_flops_per_query(*model->impl);
    });
}

int tmc_hip_philox_stream(uint64_t seed, uint64_t photon_id, uint32_t word3, int64_t n,
                          uint32_t *out, int32_t device) {
    return guarded([&] {
        if (!out && n > 0) config_fail("out must not be NULL");
        tmc::philox_stream(seed, photon_id, word3, n, out, device);
    });
}

}  // extern "C"
