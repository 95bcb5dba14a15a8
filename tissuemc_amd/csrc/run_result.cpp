#include "run_result.h"

#include <cmath>
#include <cstdio>

namespace tmc {

namespace {

constexpr double kUnit = 0x1p-32;  // weight of one tally unit

[[noreturn]] void rethrow(int code) {
    const std::string msg = tmc_last_error();
    if (code == TMC_ERR_CONFIG) throw ConfigError(msg);
    throw std::runtime_error(msg);
}

double absorbed_sum(const BatchResult &b) {
    double s = 0;
    for (double a : b.absorbed) s += a;
    return s;
}

template <class F> Estimate estimate_of(const std::vector<BatchResult> &bs, F f) {
    std::vector<double> sums;
    std::vector<uint64_t> ph;
    for (const auto &b : bs) {
        sums.push_back(f(b));
        ph.push_back(b.photons);
    }
    return batch_estimate(sums, ph);
}

}  // namespace

Estimate batch_estimate(const std::vector<double> &sums, const std::vector<uint64_t> &photons) {
    double total = 0;
    uint64_t n_ph = 0;
    for (size_t i = 0; i < sums.size(); ++i) {
        total += sums[i];
        n_ph += photons[i];
    }
    Estimate e;
    e.mean = total / static_cast<double>(n_ph);
    const size_t n = sums.size();
    if (n < 2) {
        e.se = std::nan("");
        return e;
    }
    double center = 0;
    for (size_t i = 0; i < n; ++i) center += sums[i] / static_cast<double>(photons[i]);
    center /= static_cast<double>(n);
    double ss = 0;
    for (size_t i = 0; i < n; ++i) {
        const double d = sums[i] / static_cast<double>(photons[i]) - center;
        ss += d * d;
    }
    e.se = std::sqrt(ss / static_cast<double>(n - 1) / static_cast<double>(n));
    return e;
}

uint64_t RunResult::photons() const {
    uint64_t n = 0;
    for (const auto &b : batches) n += b.photons;
    return n;
}
uint64_t RunResult::events() const {
    uint64_t n = 0;
    for (const auto &b : batches) n += b.events;
    return n;
}
uint64_t RunResult::anomalies() const {
    uint64_t n = 0;
    for (const auto &b : batches) n += b.anomalies;
    return n;
}
uint64_t RunResult::boundary_steps() const {
    uint64_t n = 0;
    for (const auto &b : batches) n += b.boundary_steps;
    return n;
}
Estimate RunResult::rd() const { return estimate_of(batches, [](const BatchResult &b) { return b.rd; }); }
Estimate RunResult::tt() const { return estimate_of(batches, [](const BatchResult &b) { return b.tt; }); }
Estimate RunResult::absorbed() const { return estimate_of(batches, absorbed_sum); }
Estimate RunResult::conservation_residual() const {
    Estimate e = estimate_of(batches, [](const BatchResult &b) { return b.rd + b.tt + absorbed_sum(b); });
    e.mean = e.mean + specular - 1.0;
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

namespace {

class Json {
  public:
    explicit Json(std::ostream &o) : o_(o) {}
    void raw(const char *s) { o_ << s; }
    void str(const std::string &s) {
        o_ << '"';
        for (char c : s) {
            if (c == '"' || c == '\\') {
                o_ << '\\' << c;
            } else if (static_cast<unsigned char>(c) < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", static_cast<unsigned>(c));
                o_ << buf;
            } else {
                o_ << c;
            }
        }
        o_ << '"';
    }
    // 17 significant digits round-trip a double; JSON has no NaN, so it is written as null.
    void num(double v) {
        if (!std::isfinite(v)) {
            o_ << "null";
            return;
        }
        char buf[40];
        std::snprintf(buf, sizeof buf, "%.17g", v);
        o_ << buf;
    }
    void integer(uint64_t v) { o_ << v; }
    void key(const char *k) {
        str(k);
        o_ << ':';
    }
    void array(const std::vector<double> &v) {
        o_ << '[';
        for (size_t i = 0; i < v.size(); ++i) {
            if (i) o_ << ',';
            num(v[i]);
        }
        o_ << ']';
    }
    void estimate(const Estimate &e) {
        o_ << '{';
        key("mean");
        num(e.mean);
        o_ << ',';
        key("se");
        num(e.se);
        o_ << '}';
    }

  private:
    std::ostream &o_;
};

}  // namespace

void write_json(const RunResult &run, std::ostream &out) {
    Json j(out);
    const tmc_spec &s = run.spec;
    j.raw("{");
    j.key("tissuemc_version");
    j.str("0.1.0");
    j.raw(",");
    j.key("engine");
    j.str(run.engine);
    j.raw(",");
    j.key("precision");
    j.str(run.precision);
    j.raw(",");
    j.key("seed");
    j.integer(run.seed);
    j.raw(",");
    j.key("threads");
    j.integer(static_cast<uint64_t>(run.threads));
    j.raw(",");
    j.key("photons");
    j.integer(run.photons());
    j.raw(",");
    j.key("output_name");
    j.str(s.output_name);
    j.raw(",");
    j.key("n_above");
    j.num(s.n_above);
    j.raw(",");
    j.key("n_below");
    j.num(s.n_below);
    j.raw(",");
    j.key("layers");
    j.raw("[");
    for (int i = 0; i < s.num_layers; ++i) {
        const tmc_layer &L = s.layers[i];
        j.raw(i ? ",{" : "{");
        j.key("n");
        j.num(L.n);
        j.raw(",");
        j.key("mua");
        j.num(L.mua);
        j.raw(",");
        j.key("mus");
        j.num(L.mus);
        j.raw(",");
        j.key("g");
        j.num(L.g);
        j.raw(",");
        j.key("d");
        j.num(L.d);
        j.raw("}");
    }
    j.raw("],");
    j.key("grid");
    j.raw("{");
    j.key("dz");
    j.num(s.dz);
    j.raw(",");
    j.key("dr");
    j.num(s.dr);
    j.raw(",");
    j.key("nz");
    j.integer(static_cast<uint64_t>(s.nz));
    j.raw(",");
    j.key("nr");
    j.integer(static_cast<uint64_t>(s.nr));
    j.raw("},");
    j.key("specular");
    j.num(run.specular);
    j.raw(",");
    j.key("wall_seconds");
    j.num(run.wall_seconds);
    j.raw(",");
    j.key("kernel_seconds");
    j.num(run.kernel_seconds);
    j.raw(",");
    j.key("device");
    j.str(run.device);
    j.raw(",");
    j.key("tally_precision");
    j.str(run.tally_precision);
    j.raw(",");
    j.key("rung");
    j.integer(static_cast<uint64_t>(run.rung));
    j.raw(",");
    j.key("profiles");
    j.raw(run.profiles ? "true" : "false");
    j.raw(",");
    j.key("events");
    j.integer(run.events());
    j.raw(",");
    j.key("anomalies");
    j.integer(run.anomalies());
    j.raw(",");
    j.key("boundary_steps");
    j.integer(run.boundary_steps());
    j.raw(",");
    j.key("launch");
    j.raw("{");
    j.key("groups");
    j.integer(static_cast<uint64_t>(run.groups));
    j.raw(",");
    j.key("waves_per_group");
    j.integer(static_cast<uint64_t>(run.waves_per_group));
    j.raw(",");
    j.key("lds_bytes");
    j.integer(static_cast<uint64_t>(run.lds_bytes));
    j.raw(",");
    j.key("chunk");
    j.integer(static_cast<uint64_t>(run.chunk));
    j.raw(",");
    j.key("arz_tile");
    j.raw("[");
    j.integer(static_cast<uint64_t>(run.arz_tile_z));
    j.raw(",");
    j.integer(static_cast<uint64_t>(run.arz_tile_r));
    j.raw("]}");
    j.raw(",");
    j.key("summary");
    j.raw("{");
    j.key("rd");
    j.estimate(run.rd());
    j.raw(",");
    j.key("tt");
    j.estimate(run.tt());
    j.raw(",");
    j.key("absorbed");
    j.estimate(run.absorbed());
    j.raw(",");
    j.key("conservation_residual");
    j.estimate(run.conservation_residual());
    j.raw("},");
    j.key("batches");
    j.raw("[");
    for (size_t i = 0; i < run.batches.size(); ++i) {
        const BatchResult &b = run.batches[i];
        j.raw(i ? ",{" : "{");
        j.key("first_photon");
        j.integer(b.first_photon);
        j.raw(",");
        j.key("photons");
        j.integer(b.photons);
        j.raw(",");
        j.key("rd");
        j.num(b.rd);
        j.raw(",");
        j.key("tt");
        j.num(b.tt);
        j.raw(",");
        j.key("absorbed");
        j.array(b.absorbed);
        j.raw(",");
        j.key("events");
        j.integer(b.events);
        j.raw(",");
        j.key("anomalies");
        j.integer(b.anomalies);
        j.raw(",");
        j.key("boundary_steps");
        j.integer(b.boundary_steps);
        j.raw("}");
    }
    j.raw("],");
    j.key("a_z");
    j.array(run.a_z);
    j.raw(",");
    j.key("rd_r");
    j.array(run.rd_r);
    j.raw(",");
    j.key("t_r");
    j.array(run.t_r);
    j.raw(",");
    j.key("profile_se");
    j.raw("{");
    j.key("a_z");
    j.array(run.a_z_se);
    j.raw(",");
    j.key("rd_r");
    j.array(run.rd_r_se);
    j.raw(",");
    j.key("t_r");
    j.array(run.t_r_se);
    j.raw("}");
    if (!run.a_rz.empty()) {
        j.raw(",");
        j.key("a_rz");
        j.array(run.a_rz);
    }
    if (!run.rd_t.empty()) {
        j.raw(",");
        j.key("time_dt_ns");
        j.num(run.dt_ns);
        j.raw(",");
        j.key("rd_t");
        j.array(run.rd_t);
        j.raw(",");
        j.key("t_t");
        j.array(run.t_t);
    }
    if (!run.rd_a.empty()) {
        const std::pair<const char *, const std::vector<double> *> angle[] = {
            {"rd_a", &run.rd_a}, {"t_a", &run.t_a}, {"rd_ra", &run.rd_ra}, {"t_ra", &run.t_ra}};
        for (const auto &kv : angle) {
            j.raw(",");
            j.key(kv.first);
            j.array(*kv.second);
        }
    }
    j.raw("}");
}

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
