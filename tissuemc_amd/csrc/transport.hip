// MCML photon transport for MI355X (gfx950): persistent wave64 lanes, Philox streams in
// registers, FP32 hop/drop/spin, integer fixed-point tallies in LDS.
//
// What the kernel computes is the reference's photon history (include/tissuemc/physics.hpp:
// launch :58, step_size :149, hit_boundary :162, hop :180, drop :188, spin :116, cross_or_not
// :199, roulette :250, step_once :264) and its scoring (include/tissuemc/gpu_tally.hpp:51-94:
// bins clamp at both ends, a finite weight and a unit direction are checked at every score).
// How it computes it is this file's own design:
//
// * One workgroup per compute unit, W waves each. A wave is the unit of scheduling: it owns
//   one batch at a time, claims photon ids of that batch in chunks with one global atomic, and
//   hands ids to its dead lanes (ballot + mbcnt), so lanes relaunch as soon as their photon ends.
//   When its batch has no ids left, the wave drains its live photons, adds its LDS tally slot into
//   the batch's HBM arrays and moves to the next batch that still has ids.
// * Tallies are integers in units of 2^-32 (a float32 weight >= 2^-8 converts exactly). Each
//   wave has a private LDS slot for its current batch (scalars and A_z, scored at every event).
//   R_r and T_r are scored once per photon and go straight to the batch's HBM arrays. The
//   workgroup shares an LDS tile of A(r,z) as 32-bit counters: an add that wraps a counter
//   carries 2^32 into the HBM bin (the LDS atomic returns the old value, so the carry is exact),
//   which halves the tile's footprint so that it covers most of the grid; bins outside the tile
//   go to HBM atomics. Integer sums do not depend on order, so a seed gives bit-identical
//   tallies on every run.
// * Random words: each lane buffers two Philox blocks (8 words). At the top of every step it
//   refills if fewer than 5 words are left, so the at-most-4 draws of a step never refill in the
//   middle of divergent code; the words of the step are rotated into r0..r3 once.
// * Math uses the gfx950 transcendental units directly: v_log_f32 (log2), v_sin_f32/v_cos_f32
//   (argument in revolutions, so the azimuth 2*pi*xi needs no multiply), v_sqrt, v_rsq, v_rcp.
//   The file is compiled without automatic FMA contraction and every fused multiply-add is
//   written out, so that all instantiations of the kernel follow bit-identical photon paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "device_rng.h"
#include "engine.h"

namespace tmc {

// ------------------------------------------------------------------ per-layer record

enum LayerField {
    kMut = 0,    // mua + mus
    kInvMut,     // 1 / mut
    kNl2InvMut,  // -ln(2) / mut: step = log2(xi) * this
    kAlbedoA,    // mua / mut: fraction of the weight absorbed per event
    kZ0,
    kZ1,
    kHgA,  // 1 - g^2
    kHgB,  // 1 - g
    kHgC,  // 2 g
    kHgD,  // 1 + g^2
    kHgE,  // 1 / (2 g), 0 when g == 0
    kN,
    kNUp,    // index above the layer
    kNDown,  // index below the layer
    kCcTop,  // cosine of the critical angle going up (0 if none)
    kCcBot,  // going down
    kLayerFields
};
static_assert(kLayerFields == 16, "layer record is 16 floats");

constexpr int kCachedFields = 11;  // kMut..kHgE live in registers between layer changes

// Slot layout (u64 words) of one wave's tally for its current batch.
enum SlotScalar { kSlotRd = 0, kSlotTt, kSlotEvents, kSlotAnomalies, kSlotBoundary, kSlotAbs };

struct KernelArgs {
    float layers[kMaxLayers * kLayerFields];
    int32_t num_layers, nz, nr, batches;
    float inv_dz, inv_dr, w0;
    uint32_t key0, key1, word3;
    uint64_t base, extra;  // batch plan: batch b holds base (+1 if b < extra) photons
    uint64_t *next;        // [B] ids claimed so far in each batch
    uint64_t *rd, *tt, *absorbed, *events, *anomalies, *bsteps, *az, *rr, *tr, *arz;
    int32_t slot_words;  // u64 words per wave slot
    int32_t off_az;
    int32_t tile_z, tile_r, tile_off;  // A(r,z) LDS tile: [tile_z x tile_r] u32 at word tile_off
    int32_t layer_off;                 // word offset of the layer table (floats) in LDS
    uint32_t chunk;
    uint32_t step_limit;
};

__device__ __forceinline__ uint64_t plan_first(const KernelArgs &a, uint64_t b) {
    return b * a.base + (b < a.extra ? b : a.extra);
}
__device__ __forceinline__ uint64_t plan_count(const KernelArgs &a, uint64_t b) {
    return a.base + (b < a.extra ? 1 : 0);
}

__device__ __forceinline__ uint64_t readfirstlane_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Weight -> fixed point (units of 2^-32), rounding to nearest. Weights never exceed 1; a weight
// of exactly 1 (a ballistic photon under matched boundaries) is 2^32.
__device__ __forceinline__ uint64_t to_fixed(float w) {
    const float y = w * 4294967296.0f;
    const uint32_t lo = static_cast<uint32_t>(__builtin_rintf(y));
    return y >= 4294967296.0f ? (1ull << 32) : static_cast<uint64_t>(lo);
}

__device__ __forceinline__ int bin_of(float x, float inv_width, int count) {
    if (!(x > 0.0f)) return 0;
    const float i = x * inv_width;
    return i >= static_cast<float>(count) ? count - 1 : static_cast<int>(i);
}

__device__ __forceinline__ void lds_add(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void hbm_add(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The score checks of the reference (gpu_tally.hpp:87-94): a NaN weight or a direction whose
// squared norm is off by more than 1e-5 counts as an anomaly.
__device__ __forceinline__ uint32_t bad_score(float ux, float uy, float uz, float w) {
    const float n2 = __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, uz * uz));
    const float dev = __builtin_fabsf(n2 - 1.0f);
    return (!(w == w) || !(dev <= 1e-5f)) ? 1u : 0u;
}

// Add a wave's slot into the HBM arrays of batch b and clear it. All 64 lanes take part.
__device__ void flush_slot(const KernelArgs &a, uint64_t *slot, uint64_t b, int lane) {
    const int L = a.num_layers;
    for (int j = lane; j < a.slot_words; j += 64) {
        const uint64_t v = slot[j];
        if (v == 0) continue;
        slot[j] = 0;
        uint64_t *dst;
        if (j == kSlotRd) dst = a.rd + b;
        else if (j == kSlotTt) dst = a.tt + b;
        else if (j == kSlotEvents) dst = a.events + b;
        else if (j == kSlotAnomalies) dst = a.anomalies + b;
        else if (j == kSlotBoundary) dst = a.bsteps + b;
        else if (j < kSlotAbs + L) dst = a.absorbed + b * L + (j - kSlotAbs);
        else dst = a.az + b * a.nz + (j - a.off_az);
        hbm_add(dst, v);
    }
}

template <bool kProfiles, bool kArz>
__global__ __launch_bounds__(1024) void transport_kernel(const KernelArgs a) {
    extern __shared__ uint64_t lds[];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wid = static_cast<int>(threadIdx.x >> 6);
    const int waves = static_cast<int>(blockDim.x >> 6);

    // Clear the wave slots and the A(r,z) tile; copy the layer table into LDS.
    const int words = waves * a.slot_words + (kArz ? (a.tile_z * a.tile_r + 1) / 2 : 0);
    for (int i = static_cast<int>(threadIdx.x); i < words; i += static_cast<int>(blockDim.x))
        lds[i] = 0;
    float *ltab = reinterpret_cast<float *>(lds + a.layer_off);
    for (int i = static_cast<int>(threadIdx.x); i < a.num_layers * kLayerFields;
         i += static_cast<int>(blockDim.x))
        ltab[i] = a.layers[i];
    __syncthreads();

    uint64_t *slot = lds + wid * a.slot_words;
    uint32_t *tile = reinterpret_cast<uint32_t *>(lds + a.tile_off);
    const int last_layer = a.num_layers - 1;

    // ---- wave-uniform scheduling state
    const uint64_t total_waves = static_cast<uint64_t>(gridDim.x) * waves;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * waves + wid;
    uint64_t b = (gw * static_cast<uint64_t>(a.batches)) / total_waves;
    uint64_t first_b = plan_first(a, b), count_b = plan_count(a, b);
    uint64_t cur = 0, end = 0;
    bool draining = false;

    // ---- lane state
    bool alive = false;
    float x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 1, w = 0, sleft = 0;
    int layer = 0;
    float P[kCachedFields];
#pragma unroll
    for (int k = 0; k < kCachedFields; ++k) P[k] = 0.0f;
    uint32_t id_lo = 0, id_hi = 0, blk = 0, idx = 8;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, q6 = 0, q7 = 0;
    uint32_t ev = 0, an = 0, bs = 0, nstep = 0;
    uint64_t acc = 0;

    for (;;) {
        // ---- relaunch: hand the next ids of the wave's batch to lanes without a photon
        uint64_t need = __ballot(!alive);
        while (need != 0 && !draining) {
            if (cur >= end) {
                uint64_t off = 0;
                if (lane == 0)
                    off = __hip_atomic_fetch_add(a.next + b, static_cast<uint64_t>(a.chunk),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                off = readfirstlane_u64(off);
                if (off >= count_b) {
                    draining = true;
                    break;
                }
                cur = first_b + off;
                end = first_b + std::min<uint64_t>(off + a.chunk, count_b);
            }
            const uint32_t n = static_cast<uint32_t>(__popcll(need));
            const uint64_t avail = end - cur;
            const uint32_t m = avail < n ? static_cast<uint32_t>(avail) : n;
            const uint32_t rank = lane_rank(need);
            if (!alive && rank < m) {
                const uint64_t id = cur + rank;
                id_lo = static_cast<uint32_t>(id);
                id_hi = static_cast<uint32_t>(id >> 32);
                const Block4 blk0 = philox4x32_10(id_lo, id_hi, 0u, a.word3, a.key0, a.key1);
                q4 = blk0.v0, q5 = blk0.v1, q6 = blk0.v2, q7 = blk0.v3;
                blk = 1;
                idx = 4;
                x = y = z = 0.0f;
                ux = uy = 0.0f;
                uz = 1.0f;
                w = a.w0;
                sleft = 0.0f;
                layer = 0;
#pragma unroll
                for (int k = 0; k < kCachedFields; ++k) P[k] = ltab[k];
                nstep = 0;
                alive = true;
            }
            cur += m;
            need = __ballot(!alive);
        }

        if (draining && __ballot(alive) == 0) {
            // The batch is done for this wave: publish its slot, then find a batch with ids left.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            flush_slot(a, slot, b, lane);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            bool found = false;
            for (int k = 1; k <= a.batches && !found; ++k) {
                const uint64_t bb = (b + static_cast<uint64_t>(k)) % static_cast<uint64_t>(a.batches);
                const uint64_t cnt = plan_count(a, bb);
                uint64_t off = 0;
                if (lane == 0)
                    off = __hip_atomic_fetch_add(a.next + bb, static_cast<uint64_t>(a.chunk),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                off = readfirstlane_u64(off);
                if (off < cnt) {
                    b = bb;
                    first_b = plan_first(a, bb);
                    count_b = cnt;
                    cur = first_b + off;
                    end = first_b + std::min<uint64_t>(off + a.chunk, cnt);
                    found = true;
                }
            }
            if (!found) break;
            draining = false;
            continue;
        }

        if (!alive) continue;

        // ---- one hop-drop-spin step (physics.hpp:264-277)
        if (idx >= 4) {
            q0 = q4, q1 = q5, q2 = q6, q3 = q7;
            const Block4 nb = philox4x32_10(id_lo, id_hi, blk, a.word3, a.key0, a.key1);
            q4 = nb.v0, q5 = nb.v1, q6 = nb.v2, q7 = nb.v3;
            ++blk;
            idx -= 4;
        }
        // r_j = word idx + j of the 8 buffered words (idx in 0..3)
        const bool i1 = (idx & 1u) != 0, i2 = (idx & 2u) != 0;
        const uint32_t r0 = i2 ? (i1 ? q3 : q2) : (i1 ? q1 : q0);
        const uint32_t r1 = i2 ? (i1 ? q4 : q3) : (i1 ? q2 : q1);
        const uint32_t r2 = i2 ? (i1 ? q5 : q4) : (i1 ? q3 : q2);
        const uint32_t r3 = i2 ? (i1 ? q6 : q5) : (i1 ? q4 : q3);

        // step_size (physics.hpp:149-158)
        const bool fresh = sleft == 0.0f;
        float s = fresh ? __builtin_amdgcn_logf(unit_open(r0)) * P[kNl2InvMut] : sleft * P[kInvMut];
        uint32_t c = fresh ? 1u : 0u;
        sleft = 0.0f;

        // hit_boundary (physics.hpp:162-178)
        const float zb = uz > 0.0f ? P[kZ1] : P[kZ0];
        const float dl = (zb - z) * __builtin_amdgcn_rcpf(uz);
        const bool hit = (uz != 0.0f) && (s > dl);
        if (hit) {
            sleft = (s - dl) * P[kMut];
            s = dl;
        }
        // hop (physics.hpp:180-184)
        x = __builtin_fmaf(s, ux, x);
        y = __builtin_fmaf(s, uy, y);
        z = __builtin_fmaf(s, uz, z);

        bool died = false;
        if (hit) {
            // cross_or_not (physics.hpp:199-246) with Fresnel (physics.hpp:71-98)
            ++bs;
            const float *lt = ltab + layer * kLayerFields;
            const bool up = uz < 0.0f;
            const float ni = lt[kN];
            const float nt = up ? lt[kNUp] : lt[kNDown];
            const float cc = up ? lt[kCcTop] : lt[kCcBot];
            const float ca1 = up ? -uz : uz;
            float r = 1.0f, ca2 = 0.0f;
            if (ca1 > cc) {
                if (ni == nt) {
                    r = 0.0f;
                    ca2 = ca1;
                } else if (ca1 > 1.0f - 1e-6f) {
                    const float t = (nt - ni) * __builtin_amdgcn_rcpf(nt + ni);
                    r = t * t;
                    ca2 = ca1;
                } else if (ca1 < 1e-6f) {
                    r = 1.0f;
                } else {
                    const float sa1 = __builtin_amdgcn_sqrtf(__builtin_fmaf(-ca1, ca1, 1.0f));
                    const float sa2 = ni * sa1 * __builtin_amdgcn_rcpf(nt);
                    if (sa2 >= 1.0f) {
                        r = 1.0f;
                    } else {
                        ca2 = __builtin_amdgcn_sqrtf(__builtin_fmaf(-sa2, sa2, 1.0f));
                        const float cap = __builtin_fmaf(ca1, ca2, -(sa1 * sa2));
                        const float cam = __builtin_fmaf(ca1, ca2, sa1 * sa2);
                        const float sap = __builtin_fmaf(sa1, ca2, ca1 * sa2);
                        const float sam = __builtin_fmaf(sa1, ca2, -(ca1 * sa2));
                        r = 0.5f * sam * sam * __builtin_fmaf(cam, cam, cap * cap) *
                            __builtin_amdgcn_rcpf(sap * sap * cam * cam);
                    }
                }
            }
            const float xi = unit_open(c ? r1 : r0);
            ++c;
            if (xi > r) {
                const float ratio = ni * __builtin_amdgcn_rcpf(nt);
                ux *= ratio;
                uy *= ratio;
                uz = up ? -ca2 : ca2;
                const bool exit_top = up && layer == 0;
                const bool exit_bot = !up && layer == last_layer;
                if (exit_top || exit_bot) {
                    an += bad_score(ux, uy, uz, w);
                    const uint64_t qw = to_fixed(w);
                    lds_add(slot + (exit_top ? kSlotRd : kSlotTt), qw);
                    if (kProfiles) {
                        const float rr = __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, y * y));
                        const int ir = bin_of(rr, a.inv_dr, a.nr);
                        hbm_add((exit_top ? a.rr : a.tr) + b * a.nr + ir, qw);
                    }
                    died = true;
                } else {
                    if (acc != 0) {
                        lds_add(slot + kSlotAbs + layer, acc);
                        acc = 0;
                    }
                    layer += up ? -1 : 1;
                    const float *nl = ltab + layer * kLayerFields;
#pragma unroll
                    for (int k = 0; k < kCachedFields; ++k) P[k] = nl[k];
                }
            } else {
                uz = -uz;
            }
        } else {
            // drop (physics.hpp:188-193) and its score (gpu_tally.hpp:51-56)
            const float dwa = w * P[kAlbedoA];
            w -= dwa;
            ++ev;
            an += bad_score(ux, uy, uz, dwa);
            const uint64_t q = to_fixed(dwa);
            acc += q;
            if (kProfiles) {
                const int iz = bin_of(z, a.inv_dz, a.nz);
                lds_add(slot + a.off_az + iz, q);
                if (kArz) {
                    const float rr = __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, y * y));
                    const int ir = bin_of(rr, a.inv_dr, a.nr);
                    uint64_t *hbm_bin = a.arz + static_cast<uint64_t>(iz) * a.nr + ir;
                    if (iz < a.tile_z && ir < a.tile_r) {
                        const uint32_t lo = static_cast<uint32_t>(q);
                        const uint32_t old = __hip_atomic_fetch_add(
                            tile + iz * a.tile_r + ir, lo, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WORKGROUP);
                        const uint64_t carry = (q >> 32) + (old + lo < old ? 1u : 0u);
                        if (carry != 0) hbm_add(hbm_bin, carry << 32);
                    } else {
                        hbm_add(hbm_bin, q);
                    }
                }
            }
            // spin (physics.hpp:116-144) with the HG sampler (physics.hpp:101-113)
            const float xi_hg = unit_open(c ? r1 : r0);
            const float xi_az = unit_open(c ? r2 : r1);
            c += 2;
            float cost;
            if (P[kHgC] == 0.0f) {
                cost = __builtin_fmaf(2.0f, xi_hg, -1.0f);
            } else {
                const float t = P[kHgA] * __builtin_amdgcn_rcpf(__builtin_fmaf(P[kHgC], xi_hg, P[kHgB]));
                cost = __builtin_fmaf(-t, t, P[kHgD]) * P[kHgE];
                cost = __builtin_fminf(__builtin_fmaxf(cost, -1.0f), 1.0f);
            }
            const float s2 = __builtin_fmaf(-cost, cost, 1.0f);
            const float sint = s2 > 0.0f ? __builtin_amdgcn_sqrtf(s2) : 0.0f;
            const float cosp = __builtin_amdgcn_cosf(xi_az);  // cos(2 pi xi)
            const float sinp = __builtin_amdgcn_sinf(xi_az);  // sin(2 pi xi)
            float nx, ny, nz_;
            if (__builtin_fabsf(uz) > 1.0f - 1e-6f) {
                nx = sint * cosp;
                ny = sint * sinp;
                nz_ = uz >= 0.0f ? cost : -cost;
            } else {
                const float t2 = __builtin_fmaf(-uz, uz, 1.0f);
                const float rt = __builtin_amdgcn_rsqf(t2);
                const float temp = t2 * rt;
                const float sr = sint * rt;
                nx = __builtin_fmaf(sr, __builtin_fmaf(ux * uz, cosp, -(uy * sinp)), ux * cost);
                ny = __builtin_fmaf(sr, __builtin_fmaf(uy * uz, cosp, ux * sinp), uy * cost);
                nz_ = __builtin_fmaf(-(sint * cosp), temp, uz * cost);
            }
            const float inv = __builtin_amdgcn_rsqf(__builtin_fmaf(nx, nx, __builtin_fmaf(ny, ny, nz_ * nz_)));
            ux = nx * inv;
            uy = ny * inv;
            uz = nz_ * inv;
        }
        // roulette (physics.hpp:250-259, 274-276)
        if (!died && w < 1e-4f) {
            if (w == 0.0f) {
                died = true;
            } else {
                const uint32_t rw = c == 1 ? r1 : (c == 2 ? r2 : r3);
                ++c;
                if (unit_open(rw) < 0.1f)
                    w = w / 0.1f;
                else
                    died = true;
            }
        }
        idx += c;
        if (++nstep >= a.step_limit && !died) {
            ++an;  // a photon this long is reported, not silently dropped
            died = true;
        }
        if (died) {
            if (acc != 0) lds_add(slot + kSlotAbs + layer, acc);
            if (ev != 0) lds_add(slot + kSlotEvents, ev);
            if (bs != 0) lds_add(slot + kSlotBoundary, bs);
            if (an != 0) lds_add(slot + kSlotAnomalies, an);
            acc = 0;
            ev = bs = an = 0;
            alive = false;
        }
    }

    if (kArz) {
        __syncthreads();
        const int tn = a.tile_z * a.tile_r;
        for (int i = static_cast<int>(threadIdx.x); i < tn; i += static_cast<int>(blockDim.x)) {
            const uint64_t v = tile[i];
            if (v != 0)
                hbm_add(a.arz + static_cast<uint64_t>(i / a.tile_r) * a.nr + (i % a.tile_r), v);
        }
    }
}

__global__ void philox_stream_kernel(uint32_t key0, uint32_t key1, uint32_t id_lo, uint32_t id_hi,
                                     uint32_t word3, int64_t n, uint32_t *out) {
    const int64_t blocks = (n + 3) / 4;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < blocks;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const Block4 v = philox4x32_10(id_lo, id_hi, static_cast<uint32_t>(k), word3, key0, key1);
        const uint32_t words[4] = {v.v0, v.v1, v.v2, v.v3};
        for (int j = 0; j < 4; ++j)
            if (4 * k + j < n) out[4 * k + j] = words[j];
    }
}

// ------------------------------------------------------------------ host side

#define HIP_TRY(call)                                                                          \
    do {                                                                                       \
        const hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                                  \
            throw EngineError(TMC_ERR_RUNTIME, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                   " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

namespace {

struct DeviceBuffer {
    void *p = nullptr;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) {
        if (bytes) HIP_TRY(hipMalloc(&p, bytes));
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

double critical_cos(double ni, double nt) { return ni > nt ? std::sqrt(1.0 - (nt * nt) / (ni * ni)) : 0.0; }

// The reference's make_geometry (config.hpp:363-390) and the per-layer constants the kernel keeps.
void fill_layers(const tmc_spec &s, float *rec) {
    double zacc = 0.0;
    for (int i = 0; i < s.num_layers; ++i) {
        const tmc_layer &L = s.layers[i];
        const double n_top = i == 0 ? s.n_above : s.layers[i - 1].n;
        const double n_bot = i + 1 == s.num_layers ? s.n_below : s.layers[i + 1].n;
        float *r = rec + i * kLayerFields;
        const float mua = static_cast<float>(L.mua), mus = static_cast<float>(L.mus);
        const float g = static_cast<float>(L.g);
        const float mut = mua + mus;
        r[kMut] = mut;
        r[kInvMut] = 1.0f / mut;
        r[kNl2InvMut] = static_cast<float>(-std::log(2.0) / static_cast<double>(mut));
        r[kAlbedoA] = mua / mut;
        r[kZ0] = static_cast<float>(zacc);
        zacc += L.d;
        r[kZ1] = static_cast<float>(zacc);
        r[kHgA] = 1.0f - g * g;
        r[kHgB] = 1.0f - g;
        r[kHgC] = 2.0f * g;
        r[kHgD] = 1.0f + g * g;
        r[kHgE] = g != 0.0f ? 1.0f / (2.0f * g) : 0.0f;
        r[kN] = static_cast<float>(L.n);
        r[kNUp] = static_cast<float>(n_top);
        r[kNDown] = static_cast<float>(n_bot);
        r[kCcTop] = static_cast<float>(critical_cos(L.n, n_top));
        r[kCcBot] = static_cast<float>(critical_cos(L.n, n_bot));
    }
}

struct DeviceInfo {
    int id = 0;
    int cus = 0;
    size_t lds_per_block = 0;
    std::string name;
};

DeviceInfo device_info(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw EngineError(TMC_ERR_DEVICE, "no usable HIP device found");
    if (device < 0 || device >= count)
        throw EngineError(TMC_ERR_CONFIG, "device " + std::to_string(device) + " does not exist (" +
                                              std::to_string(count) + " present)");
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    DeviceInfo d;
    d.id = device;
    d.cus = prop.multiProcessorCount;
    d.lds_per_block = prop.sharedMemPerBlock;
    d.name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ", " +
             std::to_string(prop.multiProcessorCount) + " CUs)";
    return d;
}

template <bool P, bool R> void set_lds_limit(size_t bytes) {
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&transport_kernel<P, R>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

}  // namespace

void run_transport(const tmc_spec &spec_in, const tmc_opts &opt, tmc_out *out) {
    const auto t0 = std::chrono::steady_clock::now();
    tmc_spec spec = spec_in;
    if (opt.photons != 0) spec.photons = opt.photons;
    validate_spec(spec);
    if (opt.batches < 1) throw EngineError(TMC_ERR_CONFIG, "batches must be >= 1");
    if (static_cast<uint64_t>(opt.batches) > spec.photons)
        throw EngineError(TMC_ERR_CONFIG, "batches cannot exceed the number of photons");
    const bool profiles = (opt.flags & TMC_FLAG_NO_PROFILES) == 0;
    const bool want_arz = (opt.flags & TMC_FLAG_ARZ) != 0;
    if (want_arz && !profiles)
        throw EngineError(TMC_ERR_CONFIG, "A(r,z) is a profile: it cannot be combined with no-profiles");
    if (want_arz && out->arz == nullptr)
        throw EngineError(TMC_ERR_CONFIG, "A(r,z) requested but tmc_out.arz is NULL");

    const DeviceInfo dev = device_info(opt.device);
    const int L = spec.num_layers, nz = spec.nz, nr = spec.nr;
    const uint64_t B = static_cast<uint64_t>(opt.batches);

    // ---- launch shape and LDS layout
    KernelArgs a;
    std::memset(&a, 0, sizeof a);
    fill_layers(spec, a.layers);
    a.num_layers = L;
    a.nz = nz;
    a.nr = nr;
    a.batches = opt.batches;
    a.inv_dz = 1.0f / static_cast<float>(spec.dz);
    a.inv_dr = 1.0f / static_cast<float>(spec.dr);
    {
        const float n0 = static_cast<float>(spec.n_above), n1 = static_cast<float>(spec.layers[0].n);
        const float r = (n0 - n1) / (n0 + n1);
        a.w0 = 1.0f - r * r;  // launch weight in float, as physics.hpp:62 computes it
    }
    a.key0 = static_cast<uint32_t>(opt.seed);
    a.key1 = static_cast<uint32_t>(opt.seed >> 32);
    a.word3 = opt.stream;
    a.base = spec.photons / B;
    a.extra = spec.photons % B;
    a.step_limit = opt.step_limit ? static_cast<uint32_t>(std::min<uint64_t>(opt.step_limit, 0xffffffffu))
                                  : 1000000000u;

    a.off_az = kSlotAbs + L;
    a.slot_words = a.off_az + (profiles ? nz : 0);

    const size_t lds_cap = dev.lds_per_block;
    const size_t layer_bytes = static_cast<size_t>(L) * kLayerFields * sizeof(float);
    int wpg = opt.waves_per_group > 0 ? opt.waves_per_group : 16;
    if (wpg > 16) throw EngineError(TMC_ERR_CONFIG, "waves_per_group must be <= 16");
    auto slots_bytes = [&](int w) { return static_cast<size_t>(w) * a.slot_words * 8; };
    while (wpg > 1 && slots_bytes(wpg) + layer_bytes > lds_cap) --wpg;
    if (slots_bytes(wpg) + layer_bytes > lds_cap)
        throw EngineError(TMC_ERR_CONFIG,
                          "the profile grid is too large for one wave's LDS tally (" +
                              std::to_string(slots_bytes(1)) + " bytes; the device has " +
                              std::to_string(lds_cap) + " per workgroup)");
    int tile_z = 0, tile_r = 0;
    if (want_arz) {
        const size_t free_bins = (lds_cap - slots_bytes(wpg) - layer_bytes) / 4;  // u32 counters
        if (opt.arz_tile_z >= 0 && opt.arz_tile_r >= 0) {
            tile_z = std::min(opt.arz_tile_z, nz);
            tile_r = std::min(opt.arz_tile_r, nr);
            if (static_cast<size_t>(tile_z) * tile_r > free_bins)
                throw EngineError(TMC_ERR_CONFIG, "A(r,z) tile does not fit in LDS");
        } else {
            // Depth bins below the stack never receive weight; cover every reachable depth and
            // as many radii from the axis outward as fit (deposits thin out with r).
            double depth = 0.0;
            for (int i = 0; i < L; ++i) depth += spec.layers[i].d;
            const double reach = std::floor(depth / spec.dz) + 2.0;
            const int nz_eff = reach < static_cast<double>(nz) ? static_cast<int>(reach) : nz;
            tile_z = nz_eff;
            tile_r = static_cast<int>(std::min<size_t>(nr, free_bins / std::max(tile_z, 1)));
            if (tile_r < std::min(nr, 32)) {
                tile_r = static_cast<int>(std::min<size_t>(nr, std::max<size_t>(1, static_cast<size_t>(std::sqrt(static_cast<double>(free_bins))))));
                tile_z = static_cast<int>(std::min<size_t>(nz_eff, free_bins / std::max(tile_r, 1)));
            }
        }
    }
    a.tile_z = tile_z;
    a.tile_r = tile_r;
    a.tile_off = wpg * a.slot_words;
    a.layer_off = a.tile_off + (tile_z * tile_r + 1) / 2;
    const size_t lds_bytes = static_cast<size_t>(a.layer_off) * 8 + layer_bytes;

    const int groups = opt.groups > 0 ? opt.groups : dev.cus;
    const uint64_t lanes = static_cast<uint64_t>(groups) * wpg * 64;
    uint32_t chunk = opt.chunk > 0 ? static_cast<uint32_t>(opt.chunk) : 0;
    if (chunk == 0) {
        // Enough claims per wave to balance the tail, few enough to keep the atomic rare.
        const uint64_t per_wave = spec.photons / std::max<uint64_t>(lanes / 64, 1);
        chunk = static_cast<uint32_t>(std::clamp<uint64_t>(per_wave / 8, 64, 4096));
    }
    a.chunk = chunk;

    // ---- device buffers
    HIP_TRY(hipSetDevice(dev.id));
    hipStream_t stream;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{stream};
    const size_t nB = B, nBL = B * L, nBz = B * nz, nBr = B * nr;
    const size_t n_arz = want_arz ? static_cast<size_t>(nz) * nr : 0;
    const size_t total_words = B /*next*/ + 6 * nB + nBL + nBz + 2 * nBr + n_arz;
    DeviceBuffer buf(total_words * 8);
    uint64_t *p = buf.as<uint64_t>();
    a.next = p; p += nB;
    a.rd = p; p += nB;
    a.tt = p; p += nB;
    a.events = p; p += nB;
    a.anomalies = p; p += nB;
    a.bsteps = p; p += nB;
    p += nB;  // spare, keeps the layout 16-byte aligned for every array
    a.absorbed = p; p += nBL;
    a.az = p; p += nBz;
    a.rr = p; p += nBr;
    a.tr = p; p += nBr;
    a.arz = want_arz ? p : nullptr;
    HIP_TRY(hipMemsetAsync(buf.p, 0, total_words * 8, stream));

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};

    const dim3 grid(static_cast<unsigned>(groups)), block(static_cast<unsigned>(wpg * 64));
    HIP_TRY(hipEventRecord(e0, stream));
    if (!profiles) {
        set_lds_limit<false, false>(lds_bytes);
        hipLaunchKernelGGL((transport_kernel<false, false>), grid, block, lds_bytes, stream, a);
    } else if (want_arz) {
        set_lds_limit<true, true>(lds_bytes);
        hipLaunchKernelGGL((transport_kernel<true, true>), grid, block, lds_bytes, stream, a);
    } else {
        set_lds_limit<true, false>(lds_bytes);
        hipLaunchKernelGGL((transport_kernel<true, false>), grid, block, lds_bytes, stream, a);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));

    auto d2h = [&](uint64_t *dst, const uint64_t *src, size_t n) {
        if (dst && n) HIP_TRY(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, stream));
    };
    d2h(out->rd, a.rd, nB);
    d2h(out->tt, a.tt, nB);
    d2h(out->events, a.events, nB);
    d2h(out->anomalies, a.anomalies, nB);
    d2h(out->boundary_steps, a.bsteps, nB);
    d2h(out->absorbed, a.absorbed, nBL);
    d2h(out->az, a.az, nBz);
    d2h(out->rr, a.rr, nBr);
    d2h(out->tr, a.tr, nBr);
    if (want_arz) d2h(out->arz, a.arz, n_arz);
    HIP_TRY(hipStreamSynchronize(stream));

    {
        const double n0 = spec.n_above, n1 = spec.layers[0].n;
        const double r = (n0 - n1) / (n0 + n1);
        out->specular = r * r;  // host double, as the reference reports it
    }
    out->kernel_seconds = static_cast<double>(ms) / 1000.0;
    std::snprintf(out->device, sizeof out->device, "%s", dev.name.c_str());
    out->groups = groups;
    out->waves_per_group = wpg;
    out->lds_bytes = static_cast<int32_t>(lds_bytes);
    out->chunk = static_cast<int32_t>(chunk);
    out->arz_tile_z = tile_z;
    out->arz_tile_r = tile_r;
    out->wall_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void philox_stream(uint64_t seed, uint64_t id, uint32_t word3, int64_t n, uint32_t *out,
                   int device) {
    if (n < 0) throw EngineError(TMC_ERR_CONFIG, "n must be >= 0");
    if (n == 0) return;
    const DeviceInfo dev = device_info(device);
    (void)dev;
    DeviceBuffer d(static_cast<size_t>(n) * 4);
    const int64_t blocks = (n + 3) / 4;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((blocks + 255) / 256, 1024));
    hipLaunchKernelGGL(philox_stream_kernel, dim3(grid), dim3(256), 0, 0,
                       static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32),
                       static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), word3, n,
                       d.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, d.p, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost));
}

bool device_available() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

}  // namespace tmc
