#include "run_result.h"

#include <cmath>
#include <cstdio>
#include <numeric>

namespace tmc {

namespace {

constexpr double kUnit = 0x1p-32;  // weight of one tally unit

[[noreturn]] void rethrow(int code) {
    const std::string msg = tmc_last_error();
    if (code == TMC_ERR_CONFIG) throw ConfigError(msg);
    throw std::runtime_error(msg);
}

double absorbed_sum(const BatchResult &b) {
    return std::accumulate(b.absorbed.begin(), b.absorbed.end(), 0.0);
}

// Batch-means estimate of one per-batch quantity over a run's batches.
template <class F> Estimate estimate_over(const std::vector<BatchResult> &bs, F value) {
    BatchMeans m;
    for (const BatchResult &b : bs) m.add(value(b), b.photons);
    return m.estimate();
}

template <class F> uint64_t count_over(const std::vector<BatchResult> &bs, F field) {
    return std::accumulate(bs.begin(), bs.end(), uint64_t{0},
                           [&](uint64_t n, const BatchResult &b) { return n + field(b); });
}

}  // namespace

void BatchMeans::add(double sum, uint64_t photons) {
    total_ += sum;
    photons_ += photons;
    // Welford's update of the mean and the sum of squared deviations of the per-batch means
    const double x = sum / static_cast<double>(photons);
    ++n_;
    const double d = x - center_;
    center_ += d / static_cast<double>(n_);
    m2_ += d * (x - center_);
}

Estimate BatchMeans::estimate() const {
    Estimate e;
    e.mean = total_ / static_cast<double>(photons_);
    e.se = n_ < 2 ? std::nan("")
                  : std::sqrt(m2_ / static_cast<double>(n_ - 1) / static_cast<double>(n_));
    return e;
}

Estimate batch_estimate(const std::vector<double> &sums, const std::vector<uint64_t> &photons) {
    BatchMeans m;
    for (size_t i = 0; i < sums.size(); ++i) m.add(sums[i], photons[i]);
    return m.estimate();
}

uint64_t RunResult::photons() const {
    return count_over(batches, [](const BatchResult &b) { return b.photons; });
}
uint64_t RunResult::events() const {
    return count_over(batches, [](const BatchResult &b) { return b.events; });
}
uint64_t RunResult::anomalies() const {
    return count_over(batches, [](const BatchResult &b) { return b.anomalies; });
}
uint64_t RunResult::boundary_steps() const {
    return count_over(batches, [](const BatchResult &b) { return b.boundary_steps; });
}
Estimate RunResult::rd() const {
    return estimate_over(batches, [](const BatchResult &b) { return b.rd; });
}
Estimate RunResult::tt() const {
    return estimate_over(batches, [](const BatchResult &b) { return b.tt; });
}
Estimate RunResult::absorbed() const { return estimate_over(batches, absorbed_sum); }
Estimate RunResult::conservation_residual() const {
    // every launched photon's weight ends as specular, diffuse reflectance, transmittance or
    // absorption: the residual estimates (sum of the four) - 1
    Estimate e = estimate_over(batches, [](const BatchResult &b) { return b.rd + b.tt + absorbed_sum(b); });
    e.mean += specular - 1.0;
    return e;
}

std::vector<tmc_spec> parse_mci_file(const std::string &path) {
    std::vector<tmc_spec> runs(64);
    int32_t n = 0;
    int code = tmc_parse_mci_file(path.c_str(), runs.data(), static_cast<int32_t>(runs.size()), &n);
    if (code != TMC_OK && n > static_cast<int32_t>(runs.size())) {
        runs.resize(static_cast<size_t>(n));
        code = tmc_parse_mci_file(path.c_str(), runs.data(), n, &n);
    }
    if (code != TMC_OK) rethrow(code);
    runs.resize(static_cast<size_t>(n));
    return runs;
}

RunResult run_hip(const tmc_spec &spec_in, const tmc_opts &opts) {
    tmc_spec spec = spec_in;
    if (opts.photons != 0) spec.photons = opts.photons;
    if (opts.batches < 1) throw ConfigError("batches must be >= 1");
    const size_t B = static_cast<size_t>(opts.batches);
    const size_t L = static_cast<size_t>(spec.num_layers > 0 ? spec.num_layers : 0);
    const size_t nz = static_cast<size_t>(spec.nz > 0 ? spec.nz : 0);
    const size_t nr = static_cast<size_t>(spec.nr > 0 ? spec.nr : 0);
    const bool arz = (opts.flags & TMC_FLAG_ARZ) != 0;
    const bool angles = (opts.flags & TMC_FLAG_ANGLES) != 0;
    const size_t na = static_cast<size_t>(spec.na > 0 ? spec.na : 0);
    const bool time = (opts.flags & TMC_FLAG_TIME) != 0;
    const size_t nt = static_cast<size_t>(opts.nt > 0 ? opts.nt : 0);
    std::vector<uint64_t> rd(B), tt(B), ab(B * L), ev(B), an(B), bs(B), az(B * nz), rr(B * nr),
        tr(B * nr), a_rz(arz ? nz * nr : 0), ra(angles ? na : 0), ta(angles ? na : 0),
        rra(angles ? nr * na : 0), tra(angles ? nr * na : 0), rt(time ? nt : 0),
        ttt(time ? nt : 0);
    tmc_out out{};
    out.rd = rd.data();
    out.tt = tt.data();
    out.absorbed = ab.data();
    out.events = ev.data();
    out.anomalies = an.data();
    out.boundary_steps = bs.data();
    out.az = az.data();
    out.rr = rr.data();
    out.tr = tr.data();
    out.arz = arz ? a_rz.data() : nullptr;
    if (angles) {
        out.ra = ra.data();
        out.ta = ta.data();
        out.rra = rra.data();
        out.tra = tra.data();
    }
    if (time) {
        out.rt = rt.data();
        out.tt_t = ttt.data();
    }
    const int code = tmc_hip_run(&spec, &opts, &out);
    if (code != TMC_OK) rethrow(code);

    RunResult r;
    r.spec = spec;
    r.seed = opts.seed;
    r.specular = out.specular;
    r.kernel_seconds = out.kernel_seconds;
    r.wall_seconds = out.wall_seconds;
    r.device = out.device;
    r.profiles = (opts.flags & TMC_FLAG_NO_PROFILES) == 0;
    r.groups = out.groups;
    r.waves_per_group = out.waves_per_group;
    r.lds_bytes = out.lds_bytes;
    r.chunk = out.chunk;
    r.arz_tile_z = out.arz_tile_z;
    r.arz_tile_r = out.arz_tile_r;
    const uint64_t base = spec.photons / B, extra = spec.photons % B;
    for (size_t b = 0; b < B; ++b) {
        BatchResult br;
        br.first_photon = b * base + (b < extra ? b : extra);
        br.photons = base + (b < extra ? 1 : 0);
        br.rd = static_cast<double>(rd[b]) * kUnit;
        br.tt = static_cast<double>(tt[b]) * kUnit;
        for (size_t l = 0; l < L; ++l) br.absorbed.push_back(static_cast<double>(ab[b * L + l]) * kUnit);
        br.events = ev[b];
        br.anomalies = an[b];
        br.boundary_steps = bs[b];
        r.batches.push_back(std::move(br));
    }
    // Profiles: exact integer sums over batches, then per-bin batch-means standard errors.
    auto profile = [&](const std::vector<uint64_t> &pb, size_t n, std::vector<double> &tot,
                       std::vector<double> &se) {
        tot.assign(n, 0.0);
        se.assign(n, 0.0);
        std::vector<double> sums(B);
        std::vector<uint64_t> ph(B);
        for (size_t b = 0; b < B; ++b) ph[b] = r.batches[b].photons;
        for (size_t i = 0; i < n; ++i) {
            uint64_t s = 0;
            for (size_t b = 0; b < B; ++b) {
                s += pb[b * n + i];
                sums[b] = static_cast<double>(pb[b * n + i]) * kUnit;
            }
            tot[i] = static_cast<double>(s) * kUnit;
            se[i] = batch_estimate(sums, ph).se;
        }
    };
    profile(az, nz, r.a_z, r.a_z_se);
    profile(rr, nr, r.rd_r, r.rd_r_se);
    profile(tr, nr, r.t_r, r.t_r_se);
    if (arz) {
        r.a_rz.resize(nz * nr);
        for (size_t i = 0; i < nz * nr; ++i) r.a_rz[i] = static_cast<double>(a_rz[i]) * kUnit;
    }
    auto scaled = [](const std::vector<uint64_t> &v) {
        std::vector<double> d(v.size());
        for (size_t i = 0; i < v.size(); ++i) d[i] = static_cast<double>(v[i]) * kUnit;
        return d;
    };
    if (angles) {
        r.rd_a = scaled(ra);
        r.t_a = scaled(ta);
        r.rd_ra = scaled(rra);
        r.t_ra = scaled(tra);
    }
    if (time) {
        r.rd_t = scaled(rt);
        r.t_t = scaled(ttt);
        r.dt_ns = opts.dt;
    }
    return r;
}

// ------------------------------------------------------------------ JSON
//
// The document of one run: the reference's keys in its order (src/io/json_out.cpp:164-306, the
// schema users read), then this engine's additions.

namespace {

json::Value estimate_doc(const Estimate &e) {
    return json::Value::record().set("mean", e.mean).set("se", e.se);
}

json::Value layer_doc(const tmc_layer &L) {
    return json::Value::record()
        .set("n", L.n).set("mua", L.mua).set("mus", L.mus).set("g", L.g).set("d", L.d);
}

json::Value batch_doc(const BatchResult &b) {
    return json::Value::record()
        .set("first_photon", b.first_photon)
        .set("photons", b.photons)
        .set("rd", b.rd)
        .set("tt", b.tt)
        .set("absorbed", json::Value::list_of(b.absorbed))
        .set("events", b.events)
        .set("anomalies", b.anomalies)
        .set("boundary_steps", b.boundary_steps);
}

}  // namespace

json::Value run_doc(const RunResult &run) {
    using json::Value;
    const tmc_spec &s = run.spec;
    Value layers = Value::list();
    for (int i = 0; i < s.num_layers; ++i) layers.push(layer_doc(s.layers[i]));
    Value batches = Value::list();
    for (const BatchResult &b : run.batches) batches.push(batch_doc(b));
    Value doc = Value::record();
    doc.set("tissuemc_version", "0.1.0")
        .set("engine", run.engine)
        .set("precision", run.precision)
        .set("seed", run.seed)
        .set("threads", run.threads)
        .set("photons", run.photons())
        .set("output_name", std::string(s.output_name))
        .set("n_above", s.n_above)
        .set("n_below", s.n_below)
        .set("layers", std::move(layers))
        .set("grid", Value::record().set("dz", s.dz).set("dr", s.dr).set("nz", s.nz).set("nr", s.nr))
        .set("specular", run.specular)
        .set("wall_seconds", run.wall_seconds)
        .set("kernel_seconds", run.kernel_seconds)
        .set("device", run.device)
        .set("tally_precision", run.tally_precision)
        .set("rung", run.rung)
        .set("profiles", run.profiles)
        .set("events", run.events())
        .set("anomalies", run.anomalies())
        .set("boundary_steps", run.boundary_steps())
        .set("launch", Value::record()
                           .set("groups", run.groups)
                           .set("waves_per_group", run.waves_per_group)
                           .set("lds_bytes", run.lds_bytes)
                           .set("chunk", run.chunk)
                           .set("arz_tile", Value::list().push(run.arz_tile_z).push(run.arz_tile_r)))
        .set("summary", Value::record()
                            .set("rd", estimate_doc(run.rd()))
                            .set("tt", estimate_doc(run.tt()))
                            .set("absorbed", estimate_doc(run.absorbed()))
                            .set("conservation_residual", estimate_doc(run.conservation_residual())))
        .set("batches", std::move(batches))
        .set("a_z", Value::list_of(run.a_z))
        .set("rd_r", Value::list_of(run.rd_r))
        .set("t_r", Value::list_of(run.t_r))
        .set("profile_se", Value::record()
                               .set("a_z", Value::list_of(run.a_z_se))
                               .set("rd_r", Value::list_of(run.rd_r_se))
                               .set("t_r", Value::list_of(run.t_r_se)));
    if (!run.a_rz.empty()) doc.set("a_rz", Value::list_of(run.a_rz));
    if (!run.rd_t.empty())
        doc.set("time_dt_ns", run.dt_ns).set("rd_t", Value::list_of(run.rd_t)).set("t_t", Value::list_of(run.t_t));
    if (!run.rd_a.empty())
        doc.set("rd_a", Value::list_of(run.rd_a))
            .set("t_a", Value::list_of(run.t_a))
            .set("rd_ra", Value::list_of(run.rd_ra))
            .set("t_ra", Value::list_of(run.t_ra));
    return doc;
}

void write_json(const RunResult &run, std::ostream &out) { run_doc(run).write(out); }

// ------------------------------------------------------------------ MCML .mco

namespace {

void put(std::ostream &o, const char *fmt, double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, fmt, v);
    o << buf;
}

// One value per line, as MCML writes its 1-D arrays.
void column(std::ostream &o, const char *header, const std::vector<double> &v) {
    o << header << "\n";
    for (double x : v) {
        put(o, "%12.4E", x);
        o << "\n";
    }
    o << "\n";
}

// Five values per line, as MCML writes its 2-D arrays.
void rows_of_five(std::ostream &o, const std::vector<double> &v) {
    for (size_t i = 0; i < v.size(); ++i) {
        put(o, "%12.4E ", v[i]);
        if (i % 5 == 4) o << "\n";
    }
    o << "\n\n";
}

}  // namespace

void write_mco(const RunResult &run, std::ostream &o) {
    const tmc_spec &s = run.spec;
    const size_t nz = static_cast<size_t>(s.nz), nr = static_cast<size_t>(s.nr),
                 na = static_cast<size_t>(s.na), L = static_cast<size_t>(s.num_layers);
    if (run.a_rz.size() != nz * nr || run.rd_a.size() != na || run.rd_ra.size() != nr * na)
        throw ConfigError("an .mco file needs the A(r,z) and exit-angle profiles (--arz --angles)");
    const double pi = 3.14159265358979323846;
    const double N = static_cast<double>(run.photons());
    const double dz = s.dz, dr = s.dr, da = pi / (2.0 * static_cast<double>(na));
    // ring area 2 pi (ir + 1/2) dr^2 (exact for an annulus) and solid angle of a cone shell
    // 2 pi (cos a - cos(a + da)) = 4 pi sin((ia + 1/2) da) sin(da / 2)
    auto area = [&](size_t ir) { return 2.0 * pi * (static_cast<double>(ir) + 0.5) * dr * dr; };
    auto omega = [&](size_t ia) {
        return 4.0 * pi * std::sin((static_cast<double>(ia) + 0.5) * da) * std::sin(0.5 * da);
    };

    char buf[160];
    o << "A1 \t# Version number of the file format.\n\n"
         "####\n# Data categories include: \n# InParm, RAT, \n"
         "# A_l, A_z, Rd_r, Rd_a, Tt_r, Tt_a, \n# A_rz, Rd_ra, Tt_ra \n####\n\n";
    std::snprintf(buf, sizeof buf, "# User time: %8.2G sec = %8.2G hr.  Simulation time of this run.\n\n",
                  run.wall_seconds, run.wall_seconds / 3600.0);
    o << buf;
    o << "InParm \t\t\t# Input parameters. cm is used.\n";
    o << (s.output_name[0] ? s.output_name : "tissuemc.mco") << " \tA\t\t# output file name, ASCII.\n";
    o << static_cast<unsigned long long>(s.photons) << " \t\t\t# No. of photons\n";
    std::snprintf(buf, sizeof buf, "%G\t%G\t\t# dz, dr [cm]\n", s.dz, s.dr);
    o << buf;
    std::snprintf(buf, sizeof buf, "%d\t%d\t%d\t# No. of dz, dr, da.\n\n", s.nz, s.nr, s.na);
    o << buf;
    std::snprintf(buf, sizeof buf, "%d\t\t\t\t\t# Number of layers\n", s.num_layers);
    o << buf << "#n\tmua\tmus\tg\td\t# One line for each layer\n";
    std::snprintf(buf, sizeof buf, "%G\t\t\t\t\t# n for medium above\n", s.n_above);
    o << buf;
    for (size_t l = 0; l < L; ++l) {
        const tmc_layer &y = s.layers[l];
        std::snprintf(buf, sizeof buf, "%G\t%G\t%G\t%G\t%G\t# layer %zu\n", y.n, y.mua, y.mus, y.g,
                      y.d, l + 1);
        o << buf;
    }
    std::snprintf(buf, sizeof buf, "%G\t\t\t\t\t# n for medium below\n\n", s.n_below);
    o << buf;

    double rd = 0, tt = 0;
    std::vector<double> a_l(L, 0.0);
    for (const BatchResult &b : run.batches) {
        rd += b.rd;
        tt += b.tt;
        for (size_t l = 0; l < L && l < b.absorbed.size(); ++l) a_l[l] += b.absorbed[l];
    }
    double a = 0;
    for (double &v : a_l) {
        a += v;
        v /= N;
    }
    o << "RAT #Reflectance, absorption, transmission. \n";
    put(o, "%-14.6G", run.specular);
    o << " \t#Specular reflectance [-]\n";
    put(o, "%-14.6G", rd / N);
    o << " \t#Diffuse reflectance [-]\n";
    put(o, "%-14.6G", a / N);
    o << " \t#Absorbed fraction [-]\n";
    put(o, "%-14.6G", tt / N);
    o << " \t#Transmittance [-]\n\n";

    o << "A_l #Absorption as a function of layer. [-]\n";
    for (double v : a_l) {
        put(o, "%12.4G", v);
        o << "\n";
    }
    o << "\n";

    std::vector<double> v;
    v.assign(nz, 0.0);
    for (size_t i = 0; i < nz; ++i) v[i] = run.a_z[i] / (N * dz);
    column(o, "A_z #A[0], [1],..A[nz-1]. [1/cm]", v);
    v.assign(nr, 0.0);
    for (size_t i = 0; i < nr; ++i) v[i] = run.rd_r[i] / (N * area(i));
    column(o, "Rd_r #Rd[0], [1],..Rd[nr-1]. [1/cm2]", v);
    v.assign(na, 0.0);
    for (size_t i = 0; i < na; ++i) v[i] = run.rd_a[i] / (N * omega(i));
    column(o, "Rd_a #Rd[0], [1],..Rd[na-1]. [sr-1]", v);
    v.assign(nr, 0.0);
    for (size_t i = 0; i < nr; ++i) v[i] = run.t_r[i] / (N * area(i));
    column(o, "Tt_r #Tt[0], [1],..Tt[nr-1]. [1/cm2]", v);
    v.assign(na, 0.0);
    for (size_t i = 0; i < na; ++i) v[i] = run.t_a[i] / (N * omega(i));
    column(o, "Tt_a #Tt[0], [1],..Tt[na-1]. [sr-1]", v);

    o << "# A[r][z]. [1/cm3]\n# A[0][0], [0][1],..[0][nz-1]\n# A[1][0], [1][1],..[1][nz-1]\n"
         "# ...\n# A[nr-1][0], [nr-1][1],..[nr-1][nz-1]\nA_rz\n";
    v.assign(nr * nz, 0.0);
    for (size_t ir = 0; ir < nr; ++ir)
        for (size_t iz = 0; iz < nz; ++iz)
            v[ir * nz + iz] = run.a_rz[iz * nr + ir] / (N * area(ir) * dz);
    rows_of_five(o, v);
    for (int t = 0; t < 2; ++t) {
        const std::vector<double> &src = t == 0 ? run.rd_ra : run.t_ra;
        const char *q = t == 0 ? "Rd" : "Tt";
        std::snprintf(buf, sizeof buf,
                      "# %s[r][angle]. [1/(cm2sr)].\n# %s[0][0], [0][1],..[0][na-1]\n"
                      "# %s[1][0], [1][1],..[1][na-1]\n# ...\n"
                      "# %s[nr-1][0], [nr-1][1],..[nr-1][na-1]\n%s_ra\n",
                      q, q, q, q, q);
        o << buf;
        v.assign(nr * na, 0.0);
        for (size_t ir = 0; ir < nr; ++ir)
            for (size_t ia = 0; ia < na; ++ia)
                v[ir * na + ia] = src[ir * na + ia] / (N * area(ir) * omega(ia));
        rows_of_five(o, v);
    }
}

}  // namespace tmc
