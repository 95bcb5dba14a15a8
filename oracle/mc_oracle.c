/* ORACLE — test infrastructure, not product code.
 *
 * A plain-C restatement of the reference's float64/float32 CPU engine (Raghava-Dwivedi/tissuemc),
 * used only by tests/, by the smoke check and by bench.py's CPU-baseline leg, as the checker.
 * Nothing in the product links or calls it.
 *
 * What each part follows (reference file:line):
 *   Philox4x32-10 block          include/tissuemc/philox.hpp:26-57
 *   per-photon stream            include/tissuemc/philox.hpp:62-83 (ctr = {id_lo, id_hi, block, w3})
 *   open-interval uniforms       include/tissuemc/philox.hpp:127-140
 *   batch plan                   include/tissuemc/batching.hpp:11-29
 *   run validation, critical cos include/tissuemc/config.hpp:38-66, 71-73
 *   Neumaier compensated sum     include/tissuemc/accumulate.hpp:15-30
 *   tally binning and checks     src/cpu/engine.cpp:58-78
 *   batch-parallel engine        src/cpu/engine.cpp:119-214 (workers claim batches; results are
 *                                per batch, so they do not depend on the thread count)
 *   physics                      mc_physics.inc (include/tissuemc/physics.hpp)
 *
 * How it is pinned (tests/test_oracle_*.py):
 *   - Philox against the three Random123 known-answer vectors (tests/cpp/test_philox.cpp:24-30);
 *   - whole-engine float64 and float32 results against golden JSON written by the reference's own
 *     CPU engine, built from its sources by oracle/Makefile (tests/golden/); the float64 path must
 *     agree bit for bit on this toolchain;
 *   - the recorded validation outputs (results/validation-cpu.json) and event count of
 *     results/cpu-scaling.json;
 *   - closed forms: Fresnel against the textbook s/p formulas, HG moments, the purely absorbing
 *     slab's multiple-reflection series, exponential step lengths.
 * T_r, A_rz and per-batch profiles have no reference counterpart: they are pinned by exact
 * identities (sum over r of T_r equals T_t, of A_rz equals A_z, per batch) and by the radial
 * symmetry shared with rd_r, which the golden data do pin.
 *
 * Extension beyond the reference: counter word 3 (reserved 0 in the reference) can carry a
 * stream word for the batched multi-stack engine; 0 reproduces the reference stream exactly.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_MAX_LAYERS 16

typedef struct {
    double n, mua, mus, g, d;
} orc_layer;

typedef struct {
    uint64_t photons;
    double dz, dr;
    int32_t nz, nr, na;
    double n_above, n_below;
    int32_t num_layers;
    orc_layer layers[ORC_MAX_LAYERS];
} orc_spec;

typedef struct {
    uint64_t seed;
    int32_t batches;
    int32_t threads;   /* 0 = one per hardware thread */
    int32_t precision; /* 64 or 32 */
    uint32_t stream;   /* counter word 3; 0 = the reference stream */
    uint64_t photons;  /* nonzero overrides spec->photons */
    int32_t nt;        /* time-of-flight bins (with out->rt); no reference counterpart */
    double dt;         /* their width [ns] */
} orc_opts;

/* Caller-owned output arrays; B = batches, L = layers. Any profile pointer may be NULL. */
typedef struct {
    uint64_t *first, *count;           /* [B] */
    double *rd, *tt;                   /* [B] */
    double *absorbed;                  /* [B*L] */
    uint64_t *events, *anomalies;      /* [B] */
    uint64_t *boundary_steps;          /* [B] steps that ended on a boundary */
    double *az;                        /* [B*nz] */
    double *rr, *tr;                   /* [B*nr] */
    double *arz;                       /* [B*nz*nr] */
    double *ra, *ta;                   /* [B*na] exit-angle profiles, or NULL */
    double *rra, *tra;                 /* [B*nr*na] (radius, angle), or NULL with ra */
    double *rt, *tt_t;                 /* [B*nt] time-of-flight profiles, or NULL */
    double specular;
    int32_t threads_used;
    double wall_seconds;
} orc_out;

/* ---------------------------------------------------------------- Philox4x32-10 */

static const uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
static const uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int i = 0; i < 10; ++i) {
        if (i > 0) {
            k0 += kW0;
            k1 += kW1;
        }
        const uint64_t p0 = (uint64_t)kM0 * c0;
        const uint64_t p1 = (uint64_t)kM1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

typedef struct {
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t buf[4];
    int idx;
} orc_rng;

static void rng_init(orc_rng *r, uint64_t seed, uint64_t id, uint32_t word3) {
    r->ctr[0] = (uint32_t)id;
    r->ctr[1] = (uint32_t)(id >> 32);
    r->ctr[2] = 0u;
    r->ctr[3] = word3;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->idx = 4;
}

static inline uint32_t rng_next(orc_rng *r) {
    if (r->idx == 4) {
        orc_philox4x32_10(r->ctr, r->key, r->buf);
        ++r->ctr[2];
        r->idx = 0;
    }
    return r->buf[r->idx++];
}

double orc_u32_to_double(uint32_t x) { return ((double)x + 0.5) * 0x1p-32; }
float orc_u32_to_float(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 0x1p-23f; }

static inline double uniform_d(orc_rng *r) { return orc_u32_to_double(rng_next(r)); }
static inline float uniform_f(orc_rng *r) { return orc_u32_to_float(rng_next(r)); }

/* The first n words of the stream (seed, id, word3). */
void orc_stream_u32(uint64_t seed, uint64_t id, uint32_t word3, int64_t n, uint32_t *out) {
    orc_rng r;
    rng_init(&r, seed, id, word3);
    for (int64_t i = 0; i < n; ++i) out[i] = rng_next(&r);
}

/* ---------------------------------------------------------------- batch plan */

typedef struct {
    uint64_t base, extra, batches;
} orc_plan;

static orc_plan plan_make(uint64_t photons, uint64_t batches) {
    orc_plan p = {photons / batches, photons % batches, batches};
    return p;
}
static uint64_t plan_count(const orc_plan *p, uint64_t b) { return p->base + (b < p->extra ? 1 : 0); }
static uint64_t plan_first(const orc_plan *p, uint64_t b) {
    return b * p->base + (b < p->extra ? b : p->extra);
}
uint64_t orc_batch_of(uint64_t photons, uint64_t batches, uint64_t id) {
    const orc_plan p = plan_make(photons, batches);
    const uint64_t boundary = p.extra * (p.base + 1);
    return id < boundary ? id / (p.base + 1) : p.extra + (id - boundary) / p.base;
}

/* ---------------------------------------------------------------- tallies */

typedef struct {
    double sum, comp;
} csum;

static inline void csum_add(csum *s, double x) {
    const double t = s->sum + x;
    if (fabs(s->sum) >= fabs(x)) {
        s->comp += (s->sum - t) + x;
    } else {
        s->comp += (x - t) + s->sum;
    }
    s->sum = t;
}
static inline double csum_value(const csum *s) { return s->sum + s->comp; }

typedef struct {
    csum rd, tt;
    csum absorbed[ORC_MAX_LAYERS];
    csum *az, *rr, *tr, *arz;
    csum *ra, *ta, *rra, *tra;
    csum *rt, *tt_t;
    uint64_t events, anomalies, boundary_steps;
} orc_tally;

static inline size_t orc_bin(double x, double width, int count) {
    if (!(x > 0)) return 0;
    const double i = x / width;
    return i >= (double)count ? (size_t)(count - 1) : (size_t)i;
}

double orc_critical_cos(double ni, double nt) { return ni > nt ? sqrt(1.0 - (nt * nt) / (ni * ni)) : 0.0; }

/* ---------------------------------------------------------------- physics, both precisions */

#define REAL double
#define SFX d
#define SQRT sqrt
#define LOG log
#define COS cos
#define SIN sin
#define FABS fabs
#define PI 3.14159265358979323846
#define COSZERO (1.0 - 1e-12)
#define COS90D 1e-6
#define WEIGHT_THRESHOLD 1e-4
#define CHANCE 0.1
#include "mc_physics.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef LOG
#undef COS
#undef SIN
#undef FABS
#undef PI
#undef COSZERO
#undef COS90D
#undef WEIGHT_THRESHOLD
#undef CHANCE

#define REAL float
#define SFX f
#define SQRT sqrtf
#define LOG logf
#define COS cosf
#define SIN sinf
#define FABS fabsf
#define PI 3.14159265358979323846f
#define COSZERO (1.0f - 1e-6f)
#define COS90D 1e-6f
#define WEIGHT_THRESHOLD 1e-4f
#define CHANCE 0.1f
#include "mc_physics.inc"

/* ---------------------------------------------------------------- validation (config.hpp:38-66) */

static int fail_msg(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", msg);
    return 1;
}

int orc_validate(const orc_spec *run, char *err, int errlen) {
    char buf[160];
    if (run->photons == 0) return fail_msg(err, errlen, "number of photons must be > 0");
    if (!(run->dz > 0) || !(run->dr > 0)) return fail_msg(err, errlen, "dz and dr must be > 0");
    if (run->nz < 1 || run->nr < 1 || run->na < 1)
        return fail_msg(err, errlen, "grid bin counts must be >= 1");
    if (!(run->n_above >= 1.0) || !(run->n_below >= 1.0))
        return fail_msg(err, errlen, "ambient refractive indices must be >= 1");
    if (run->num_layers < 1) return fail_msg(err, errlen, "at least one layer is required");
    if (run->num_layers > ORC_MAX_LAYERS) {
        snprintf(buf, sizeof buf, "at most %d layers are supported", ORC_MAX_LAYERS);
        return fail_msg(err, errlen, buf);
    }
    for (int i = 0; i < run->num_layers; ++i) {
        const orc_layer *L = &run->layers[i];
        const char *why = NULL;
        if (!(L->n >= 1.0))
            why = "n must be >= 1";
        else if (!(L->mua >= 0) || !(L->mus >= 0))
            why = "mua and mus must be >= 0";
        else if (!(L->mua + L->mus > 0))
            why = "clear layers (mua = mus = 0) are not supported";
        else if (!(L->g > -1.0 && L->g < 1.0))
            why = "g must be in (-1, 1)";
        else if (!(L->d > 0) || !isfinite(L->d))
            why = "thickness must be finite and > 0";
        if (why) {
            snprintf(buf, sizeof buf, "layer %d: %s", i + 1, why);
            return fail_msg(err, errlen, buf);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- engine (engine.cpp:119-214) */

typedef struct {
    const orc_spec *run;
    const orc_opts *opt;
    orc_out *out;
    orc_plan plan;
    geo_t_d geo_d;
    geo_t_f geo_f;
    atomic_ullong next_batch;
} orc_job;

static void run_batch(orc_job *job, uint64_t b) {
    const orc_spec *run = job->run;
    orc_out *out = job->out;
    const int L = run->num_layers, nz = run->nz, nr = run->nr;
    orc_tally t;
    memset(&t, 0, sizeof t);
    t.az = calloc((size_t)nz, sizeof(csum));
    t.rr = calloc((size_t)nr, sizeof(csum));
    t.tr = calloc((size_t)nr, sizeof(csum));
    if (out->arz) t.arz = calloc((size_t)nz * (size_t)nr, sizeof(csum));
    const int na = run->na;
    if (out->ra) {
        t.ra = calloc((size_t)na, sizeof(csum));
        t.ta = calloc((size_t)na, sizeof(csum));
        t.rra = calloc((size_t)nr * (size_t)na, sizeof(csum));
        t.tra = calloc((size_t)nr * (size_t)na, sizeof(csum));
    }
    const int nt = job->opt->nt;
    if (out->rt) {
        t.rt = calloc((size_t)nt, sizeof(csum));
        t.tt_t = calloc((size_t)nt, sizeof(csum));
    }
    const uint64_t first = plan_first(&job->plan, b), count = plan_count(&job->plan, b);
    for (uint64_t id = first; id < first + count; ++id) {
        orc_rng rng;
        rng_init(&rng, job->opt->seed, id, job->opt->stream);
        if (job->opt->precision == 32)
            simulate_photon_f(&job->geo_f, &rng, &t);
        else
            simulate_photon_d(&job->geo_d, &rng, &t);
    }
    out->first[b] = first;
    out->count[b] = count;
    out->rd[b] = csum_value(&t.rd);
    out->tt[b] = csum_value(&t.tt);
    for (int l = 0; l < L; ++l) out->absorbed[b * (uint64_t)L + (uint64_t)l] = csum_value(&t.absorbed[l]);
    out->events[b] = t.events;
    out->anomalies[b] = t.anomalies;
    if (out->boundary_steps) out->boundary_steps[b] = t.boundary_steps;
    if (out->az)
        for (int i = 0; i < nz; ++i) out->az[b * (uint64_t)nz + (uint64_t)i] = csum_value(&t.az[i]);
    if (out->rr)
        for (int i = 0; i < nr; ++i) out->rr[b * (uint64_t)nr + (uint64_t)i] = csum_value(&t.rr[i]);
    if (out->tr)
        for (int i = 0; i < nr; ++i) out->tr[b * (uint64_t)nr + (uint64_t)i] = csum_value(&t.tr[i]);
    if (out->arz) {
        const size_t n = (size_t)nz * (size_t)nr;
        for (size_t i = 0; i < n; ++i) out->arz[b * n + i] = csum_value(&t.arz[i]);
    }
    if (out->ra) {
        for (int i = 0; i < na; ++i) {
            out->ra[b * (uint64_t)na + (uint64_t)i] = csum_value(&t.ra[i]);
            out->ta[b * (uint64_t)na + (uint64_t)i] = csum_value(&t.ta[i]);
        }
        const size_t n = (size_t)nr * (size_t)na;
        for (size_t i = 0; i < n; ++i) {
            out->rra[b * n + i] = csum_value(&t.rra[i]);
            out->tra[b * n + i] = csum_value(&t.tra[i]);
        }
    }
    free(t.az);
    free(t.rr);
    free(t.tr);
    free(t.arz);
    free(t.ra);
    free(t.ta);
    if (out->rt) {
        for (int i = 0; i < nt; ++i) {
            out->rt[b * (uint64_t)nt + (uint64_t)i] = csum_value(&t.rt[i]);
            out->tt_t[b * (uint64_t)nt + (uint64_t)i] = csum_value(&t.tt_t[i]);
        }
    }
    free(t.rra);
    free(t.tra);
    free(t.rt);
    free(t.tt_t);
}

static void *worker(void *arg) {
    orc_job *job = arg;
    const uint64_t nb = (uint64_t)job->opt->batches;
    for (;;) {
        const uint64_t b = atomic_fetch_add_explicit(&job->next_batch, 1, memory_order_relaxed);
        if (b >= nb) return NULL;
        run_batch(job, b);
    }
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_run(const orc_spec *spec_in, const orc_opts *opt, orc_out *out, char *err, int errlen) {
    orc_spec run = *spec_in;
    if (opt->photons != 0) run.photons = opt->photons;
    if (orc_validate(&run, err, errlen)) return 1;
    if (opt->batches < 1) return fail_msg(err, errlen, "batches must be >= 1");
    if ((uint64_t)opt->batches > run.photons)
        return fail_msg(err, errlen, "batches cannot exceed the number of photons");
    if (opt->threads < 0 || opt->threads > 1024)
        return fail_msg(err, errlen, "threads must be in [0, 1024] (0 = one per hardware thread)");
    if (opt->precision != 64 && opt->precision != 32)
        return fail_msg(err, errlen, "precision must be 64 or 32");
    if (out->ra && (!out->ta || !out->rra || !out->tra))
        return fail_msg(err, errlen, "angle profiles need ra, ta, rra and tra");
    if (out->ra && run.na < 1) return fail_msg(err, errlen, "angle profiles need na >= 1");
    if (out->rt && !out->tt_t) return fail_msg(err, errlen, "time profiles need rt and tt_t");
    if (out->rt && (opt->nt < 1 || !(opt->dt > 0.0)))
        return fail_msg(err, errlen, "time profiles need nt >= 1 and dt > 0");
    int threads = opt->threads;
    if (threads == 0) {
        const long hw = sysconf(_SC_NPROCESSORS_ONLN);
        threads = hw > 0 ? (int)hw : 1;
    }
    if (threads > opt->batches) threads = opt->batches;

    const double t0 = now_seconds();
    orc_job job;
    job.run = &run;
    job.opt = opt;
    job.out = out;
    job.plan = plan_make(run.photons, (uint64_t)opt->batches);
    make_geometry_d(&run, &job.geo_d);
    make_geometry_f(&run, &job.geo_f);
    /* a time bin of dt ns is c dt cm of optical path (c = 29.9792458 cm/ns) */
    job.geo_d.nt = job.geo_f.nt = out->rt ? opt->nt : 0;
    job.geo_d.dt_opl = job.geo_f.dt_opl = out->rt ? 29.9792458 * opt->dt : 1.0;
    atomic_init(&job.next_batch, 0);
    out->specular = specular_d(&job.geo_d);

    pthread_t *pool = calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 1; t < threads; ++t) pthread_create(&pool[t], NULL, worker, &job);
    worker(&job);
    for (int t = 1; t < threads; ++t) pthread_join(pool[t], NULL);
    free(pool);
    out->threads_used = threads;
    out->wall_seconds = now_seconds() - t0;
    return 0;
}

/* ---------------------------------------------------------------- single-function hooks for tests */

double orc_fresnel(double n1, double n2, double ca1, double *ca2, int precision) {
    if (precision == 32) {
        float c2 = 0.0f;
        const float r = fresnel_f((float)n1, (float)n2, (float)ca1, &c2);
        *ca2 = c2;
        return r;
    }
    return fresnel_d(n1, n2, ca1, ca2);
}

/* n draws of cos(theta) from the HG sampler on the stream (seed, id). */
void orc_hg_samples(double g, uint64_t seed, uint64_t id, int64_t n, double *out, int precision) {
    orc_rng rng;
    rng_init(&rng, seed, id, 0);
    for (int64_t i = 0; i < n; ++i)
        out[i] = precision == 32 ? (double)sample_hg_cos_f((float)g, &rng) : sample_hg_cos_d(g, &rng);
}

/* Repeated spins from an initial direction; reports the worst |norm - 1| and the mean cosine
 * between successive directions. restart_every > 0 resets the direction to (0, 0, -1). */
void orc_spin_walk(double g, uint64_t seed, uint64_t id, int64_t n, const double u0[3],
                   int64_t restart_every, int precision, double *worst_norm_err, double *mean_cos) {
    orc_rng rng;
    rng_init(&rng, seed, id, 0);
    double worst = 0, sum_cos = 0;
    photon_t_d pd;
    photon_t_f pf;
    memset(&pd, 0, sizeof pd);
    memset(&pf, 0, sizeof pf);
    pd.ux = u0[0], pd.uy = u0[1], pd.uz = u0[2];
    pf.ux = (float)u0[0], pf.uy = (float)u0[1], pf.uz = (float)u0[2];
    for (int64_t i = 0; i < n; ++i) {
        double ux, uy, uz, vx, vy, vz;
        if (precision == 32) {
            ux = pf.ux, uy = pf.uy, uz = pf.uz;
            spin_f(&pf, (float)g, &rng);
            vx = pf.ux, vy = pf.uy, vz = pf.uz;
        } else {
            ux = pd.ux, uy = pd.uy, uz = pd.uz;
            spin_d(&pd, g, &rng);
            vx = pd.ux, vy = pd.uy, vz = pd.uz;
        }
        sum_cos += ux * vx + uy * vy + uz * vz;
        const double e = fabs(sqrt(vx * vx + vy * vy + vz * vz) - 1.0);
        if (e > worst) worst = e;
        if (restart_every > 0 && i % restart_every == 0) {
            pd.ux = 0, pd.uy = 0, pd.uz = -1;
            pf.ux = 0, pf.uy = 0, pf.uz = -1;
        }
    }
    *worst_norm_err = worst;
    *mean_cos = sum_cos / (double)n;
}

/* n fresh step lengths in layer 0 of the run. */
void orc_step_lengths(const orc_spec *run, uint64_t seed, uint64_t id, int64_t n, double *out) {
    geo_t_d geo;
    make_geometry_d(run, &geo);
    orc_rng rng;
    rng_init(&rng, seed, id, 0);
    photon_t_d p;
    launch_d(&p, &geo);
    for (int64_t i = 0; i < n; ++i) {
        p.sleft = 0;
        step_size_d(&p, &geo, &rng);
        out[i] = p.s;
    }
}
