// MCML photon transport for MI355X (gfx950): persistent wave64 lanes, Philox streams in
// registers, FP32 hop/drop/spin, integer fixed-point tallies in LDS.
//
// What the kernel computes is the reference's photon history (include/tissuemc/physics.hpp:
// launch :58, step_size :149, hit_boundary :162, hop :180, drop :188, spin :116, cross_or_not
// :199, roulette :250, step_once :264) and its scoring (include/tissuemc/gpu_tally.hpp:51-94:
// bins clamp at both ends, a finite weight and a unit direction are checked at every score).
// How it computes it is this file's own design:
//
// * Work comes in units: a range of photon ids of one layer stack, with the rows its scalars
//   and profiles add into. A single run is one stack cut into its batches; a batched run is many
//   stacks, each cut into batches. Ids are claimed within a unit in chunks, one global atomic
//   per chunk.
// * One workgroup per compute unit, W waves each. A wave owns one unit at a time and hands ids
//   of the unit to its dead lanes (ballot + mbcnt), so lanes relaunch as soon as a photon ends.
//   Waves take units from a global cursor first; once the cursor has passed the last unit, a wave
//   looks for units that still have ids with a 64-lane parallel scan of the claim counters, so
//   large units end up shared by many waves. When its unit has no ids left, the wave drains its
//   live photons and adds its LDS slot into the unit's HBM rows.
// * Tallies are integers in units of 2^-32 (a float32 weight >= 2^-8 converts exactly). Each
//   wave has a private LDS slot (scalars and A_z, scored at every event) and a private copy of
//   its stack's layer table. R_r and T_r are scored once per photon, straight into HBM. For a
//   single stack the workgroup also shares an LDS tile of A(r,z) as 32-bit counters: an add that
//   wraps a counter carries 2^32 into the HBM bin (the LDS atomic returns the old value, so the
//   carry is exact), which halves the tile's footprint so that it covers most of the grid; bins
//   outside the tile go to HBM atomics. Integer sums do not depend on order, so a seed gives
//   bit-identical tallies on every run and for every launch shape.
// * Random words: each lane buffers two Philox blocks (8 words). At the top of every step it
//   refills if fewer than 5 words are left, so the at-most-4 draws of a step never refill in the
//   middle of divergent code; the words of the step are rotated into r0..r3 once.
// * Loop-carried lane state is only ever changed under wave-uniform control flow or by select:
//   state written inside a divergent branch makes the compiler copy it on every iteration.
// * Math uses the gfx950 transcendental units directly: v_log_f32 (log2), v_sin_f32/v_cos_f32
//   (argument in revolutions, so the azimuth 2*pi*xi needs no multiply), v_sqrt, v_rsq, v_rcp.
//   The file is compiled without automatic FMA contraction and every fused multiply-add is
//   written out, so that all instantiations of the kernel follow bit-identical photon paths.
// * Variants are template parameters, so an option that is off costs the step loop nothing:
//   kProfiles (A_z, R_r, T_r; off for the reference's no-profiles diagnostic), kArz (the A(r,z)
//   tile), kAgg (lane aggregation of the A_z and A(r,z) adds, TMC_FLAG_AGGREGATE) and kExit
//   (the exit-angle and time-of-flight profiles, TMC_FLAG_ANGLES and TMC_FLAG_TIME, which are
//   then each a run-time switch on the escape path). Even a branch that is never taken costs the
//   step loop registers: its kernel arguments stay live across the loop, and past the scalar
//   register file they are spilled to and reloaded from vector lanes on the hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "device_rng.h"
#include "engine.h"

namespace tmc {

// ------------------------------------------------------------------ records

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
    kCcTop,    // cosine of the critical angle going up (0 if none)
    kCcBot,    // going down
    kEtaUp,    // n / n_above: Snell's ratio sin(t) / sin(i) going up
    kEtaDown,  // n / n_below
    kR0Up,     // Fresnel reflectance at normal incidence going up, ((n - n_above) / (n + n_above))^2
    kR0Down,   // going down
    kLayerFields
};
static_assert(kLayerFields == 20, "layer record is 20 floats");

// The step reads a layer's record as 16-byte LDS loads; the table starts 16-byte aligned.
// Pointers into LDS are kept in address space 3, so that their arithmetic stays 32-bit.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const f32x4 lds_f32x4;
typedef __attribute__((address_space(3))) const float lds_float;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
struct LayerRec {
    float f[kLayerFields];
};
// Reads 16-byte groups [kFirst, kFirst + kCount) of a layer's record into r. The step reads the
// first three (stepping, absorption, scattering, n) and a boundary crossing the last two, so the
// crossing's eight values are not held in registers through every step.
template <int kFirst, int kCount>
__device__ __forceinline__ void load_layer(LayerRec &r, uint32_t ltab, int layer) {
    const uint32_t at = ltab + __umul24(static_cast<uint32_t>(layer), kLayerFields * sizeof(float));
    const lds_f32x4 *p = (const lds_f32x4 *)(uintptr_t)at;
#pragma unroll
    for (int k = kFirst; k < kFirst + kCount; ++k) {
        const f32x4 v = p[k];
        r.f[4 * k] = v.x;
        r.f[4 * k + 1] = v.y;
        r.f[4 * k + 2] = v.z;
        r.f[4 * k + 3] = v.w;
    }
}
static_assert(kN < 12 && kNUp >= 12, "the step's groups end with n, the crossing's begin with n_above");

// Stack record (floats): this header, then kLayerFields floats per layer.
enum StackHeader { kHdrW0 = 0, kHdrInvDz, kHdrInvDr, kHdrLayers, kStackHeader };

// Slot layout (u64 words) of one wave's tally for its current unit: these scalars, then
// absorbed[lmax], then A_z[nz].
enum SlotScalar { kSlotRd = 0, kSlotTt, kSlotEvents, kSlotAnomalies, kSlotBoundary, kSlotAbs };

constexpr double kLightCmPerNs = 29.9792458;  // speed of light in vacuum

// A unit of work: a range of photon ids of one stack and the rows its results add into.
struct UnitRec {
    uint64_t first;  // first photon id (ids index the stack's stream)
    uint64_t count;
    uint32_t stack;  // stack record
    uint32_t prow;   // profile row (A_z, R_r, T_r)
    uint32_t srow;   // scalar row (rd, tt, absorbed, counters)
    uint32_t word3;  // Philox counter word 3 of the stack's streams
};
static_assert(sizeof(UnitRec) == 32, "UnitRec is 32 bytes");

// The rows a wave's slot is flushed into (ColdRows order): read from device memory when a unit
// ends, so that the step loop does not keep these pointers in scalar registers.
enum ColdRow { kColdRd = 0, kColdTt, kColdEvents, kColdAnomalies, kColdBsteps, kColdAbsorbed, kColdAz, kColdRows };
static_assert(int(kColdRd) == int(kSlotRd) && int(kColdTt) == int(kSlotTt) &&
                  int(kColdEvents) == int(kSlotEvents) && int(kColdAnomalies) == int(kSlotAnomalies) &&
                  int(kColdBsteps) == int(kSlotBoundary) && int(kColdAbsorbed) == int(kSlotAbs),
              "the slot's scalars and the flush rows share their order");

struct KernelArgs {
    const UnitRec *units;
    const float *stacks;
    const uint64_t *key;  // [1] seed, in device memory so that a captured graph can be replayed
    uint64_t *const *cold;  // [kColdRows] the per-batch rows below, for flush_slot
    uint64_t *next;       // [U] ids claimed so far in each unit
    uint64_t *cursor;     // [1] units handed out so far
    uint64_t *rd, *tt, *absorbed, *events, *anomalies, *bsteps, *az, *rr, *tr, *arz;
    uint64_t *ra, *ta, *rra, *tra;  // exit-angle profiles [na], [nr x na]; na == 0: none
    uint64_t *rt, *ttt;             // time-resolved R_d(t), T_t(t) [nt]; nt == 0: none
    int32_t nt;
    float inv_dopl;                 // 1 / (c dt): optical path (cm) -> time bin
    uint32_t n_units, stack_stride;
    int32_t lmax, abs_stride, nz, nr, na;
    float inv_da;  // 2 na / pi
    int32_t slot_words, wave_words;    // u64 words: the slot, and slot + layer table, per wave
                                       // (the table starts at the even word after the slot)
    int32_t tile_z, tile_r, tile_off;  // A(r,z) LDS tile: [tile_z x tile_r] u32 at word tile_off
    // Exit-angle and time-of-flight bins of the workgroup, [ra na][ta na][rt nt][ttt nt] u64 at
    // word 0 (totals over all batches, flushed once at the end); the wave regions follow them
    int32_t ext_words;
    int32_t lds_words;  // u64 words of LDS the workgroup uses (cleared at start)
    uint32_t chunk;
    uint32_t step_limit;
};

__device__ __forceinline__ uint64_t readfirstlane_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// A wave-uniform 64-bit mask as this lane's predicate: the mask is the condition register
// itself, no vector instruction.
__device__ __forceinline__ bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// The active lanes where p holds. Free when p is one comparison (its result is the mask); any
// other boolean is first written to a vector register and compared again (two instructions).
__device__ __forceinline__ uint64_t wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Weight -> fixed point (units of 2^-32), rounding to nearest. Weights never exceed 1; a weight
// of exactly 1 (a ballistic photon under matched boundaries) is 2^32.
__device__ __forceinline__ uint64_t to_fixed(float w) {
    const float y = w * 4294967296.0f;
    const uint32_t lo = static_cast<uint32_t>(__builtin_rintf(y));
    return y >= 4294967296.0f ? (1ull << 32) : static_cast<uint64_t>(lo);
}

// Bin of x >= 0 with both ends clamped (gpu_tally.hpp:51-56): x <= 0 and NaN give bin 0, x past
// the grid the last bin. The clamp is one v_med3_f32 in float (last = count - 1, exact below
// 2^24; with a NaN operand med3 returns the minimum of the others, 0), so the conversion only
// ever sees values in [0, count - 1].
__device__ __forceinline__ int bin_of(float x, float inv_width, float last) {
    const float i = __builtin_amdgcn_fmed3f(x * inv_width, 0.0f, last);
    return static_cast<int>(i);
}

__device__ __forceinline__ void lds_add(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void hbm_add(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same through a pointer marked global: an address formed in vector registers otherwise
// makes a flat atomic, which also counts against the LDS wait counter, so the next LDS wait
// would wait for the HBM atomic too.
typedef __attribute__((address_space(1))) uint64_t global_u64;
__device__ __forceinline__ void hbm_add(global_u64 *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ lane aggregation
//
// TMC_FLAG_AGGREGATE: the wave-level counterpart of the reference's aggregated_add
// (gpu_tally.hpp:200-234, match_any + shuffles over a 32-lane warp). Wave64 has no match
// instruction, so groups are peeled with ballots: the lowest lane still to score names its bin,
// a ballot finds the lanes with that bin, a DPP reduction sums their weights, and that one lane
// makes a single atomic add for the group. kAggPeels groups are peeled per score; the lanes left
// after that add one by one. Integer sums do not depend on how they are grouped, so the tallies
// are bit-identical to the plain path's. Every lane of the wave must be active.
constexpr int kAggPeels = 1;

// Lanes of a wave that wait at a layer boundary before the wave makes their crossings.
constexpr int kCrossBatch = 6;

// Sum of v over the wave's 64 lanes (every lane active), returned wave-uniform. DPP steps:
// row_shr 1, 2, 3 of v (zero past the start of a 16-lane row), then row_shr 4 into banks 1-3 and
// row_shr 8 into banks 2-3 of the running sum (lane 15 of each row now holds the row's sum),
// row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3: lane 63 holds the total.
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    const int x = static_cast<int>(v);
    int t = x;
    t += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
    t += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
    t += __builtin_amdgcn_update_dpp(0, x, 0x113, 0xf, 0xf, true);
    t += __builtin_amdgcn_update_dpp(0, t, 0x114, 0xf, 0xe, false);
    t += __builtin_amdgcn_update_dpp(0, t, 0x118, 0xf, 0xc, false);
    t += __builtin_amdgcn_update_dpp(0, t, 0x142, 0xa, 0xf, false);
    t += __builtin_amdgcn_update_dpp(0, t, 0x143, 0xc, 0xf, false);
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(t, 63));
}

// Sum of fixed-point weights q <= 2^32 over the wave, exact: summed as 16-bit low parts and
// high parts (each sum < 2^23), then recombined.
__device__ __forceinline__ uint64_t wave_sum_fixed(uint64_t q) {
    const uint32_t lo = wave_sum_u32(static_cast<uint32_t>(q) & 0xffffu);
    const uint32_t hi = wave_sum_u32(static_cast<uint32_t>(q >> 16));
    return (static_cast<uint64_t>(hi) << 16) + lo;
}

// Adds q into row[key] for the lanes with `score`, peeling kAggPeels groups first.
__device__ __forceinline__ void aggregate_lds_add(uint64_t *row, bool score, int key, uint64_t q,
                                                  int lane) {
    bool pending = score;
    uint64_t rest = __ballot(pending);
    for (int p = 0; p < kAggPeels && rest != 0; ++p) {
        const int leader = static_cast<int>(__builtin_ctzll(rest));
        const int kb = __builtin_amdgcn_readlane(key, leader);
        const bool in = pending && key == kb;
        const uint64_t sum = wave_sum_fixed(in ? q : 0ull);
        if (lane == leader) lds_add(row + kb, sum);
        pending = pending && !in;
        rest &= ~__ballot(in);
    }
    if (pending) lds_add(row + key, q);
}

// A(r,z) tile add of v < 2^38 into the 32-bit counter t, whose HBM bin is `bin`: the low word
// goes to the counter, the high word and the counter's wrap (the atomic returns the value it
// added to) to the HBM bin.
__device__ __forceinline__ void tile_add(uint32_t *tile, uint64_t *arz, uint32_t t, uint32_t bin,
                                         uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(v);
    const uint32_t old =
        __hip_atomic_fetch_add(tile + t, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint64_t carry = (v >> 32) + (old + lo < old ? 1u : 0u);
    if (carry != 0) hbm_add(arz + bin, carry << 32);
}

__device__ __forceinline__ void aggregate_tile_add(uint32_t *tile, uint64_t *arz, bool score,
                                                   uint32_t t, uint32_t bin, uint64_t q, int lane) {
    bool pending = score;
    uint64_t rest = __ballot(pending);
    for (int p = 0; p < kAggPeels && rest != 0; ++p) {
        const int leader = static_cast<int>(__builtin_ctzll(rest));
        const uint32_t tb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t), leader));
        const uint32_t gb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bin), leader));
        const bool in = pending && t == tb;
        const uint64_t sum = wave_sum_fixed(in ? q : 0ull);
        if (lane == leader) tile_add(tile, arz, tb, gb, sum);
        pending = pending && !in;
        rest &= ~__ballot(in);
    }
    if (pending) tile_add(tile, arz, t, bin, q);
}

// The score checks of the reference (gpu_tally.hpp:87-94): a NaN weight or a direction whose
// squared norm is off by more than 1e-5 counts as an anomaly.
__device__ __forceinline__ uint32_t bad_score(float ux, float uy, float uz, float w) {
    const float n2 = __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, uz * uz));
    const float dev = __builtin_fabsf(n2 - 1.0f);
    return (!(w == w) || !(dev <= 1e-5f)) ? 1u : 0u;
}

// The wave's current unit (wave-uniform values).
struct WaveUnit {
    uint64_t cur, end;  // ids of the current chunk not handed out yet
    uint32_t u, stack, prow, srow, word3;
    float w0, inv_dz, inv_dr;
    int nl;
};

// Claim the next chunk of unit u for this wave; false if u has no ids left.
__device__ __forceinline__ bool claim_chunk(const KernelArgs &a, uint32_t u, int lane, WaveUnit &wu) {
    uint64_t off = 0;
    if (lane == 0)
        off = __hip_atomic_fetch_add(a.next + u, static_cast<uint64_t>(a.chunk), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    off = readfirstlane_u64(off);
    UnitRec rec = a.units[u];
    // The two words the step loop uses become scalars here, so the wait for their load is made
    // here, once per chunk. A loaded value first used inside the loop (word3 in every Philox
    // block) makes the compiler wait at that use on every step, with vmcnt(0), which also waits
    // for every HBM atomic the wave still has in flight. (Measured: no change in time on the
    // three-layer, tissue and batched workloads, where other waves cover those waits; two fewer
    // VGPRs.)
    rec.word3 = __builtin_amdgcn_readfirstlane(rec.word3);
    rec.prow = __builtin_amdgcn_readfirstlane(rec.prow);
    if (off >= rec.count) return false;
    wu.u = u;
    wu.cur = rec.first + off;
    wu.end = rec.first + std::min<uint64_t>(off + a.chunk, rec.count);
    wu.stack = rec.stack;
    wu.prow = rec.prow;
    wu.srow = rec.srow;
    wu.word3 = rec.word3;
    return true;
}

// Find a unit with ids left: the cursor hands each unit out once; after that, a parallel scan of
// the claim counters, 64 units per round, from a wave-specific offset.
__device__ bool acquire_unit(const KernelArgs &a, int lane, uint32_t scan_start, bool &cursor_done,
                             WaveUnit &wu) {
    while (!cursor_done) {
        uint64_t c = 0;
        if (lane == 0)
            c = __hip_atomic_fetch_add(a.cursor, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        c = readfirstlane_u64(c);
        if (c >= a.n_units) {
            cursor_done = true;
            break;
        }
        if (claim_chunk(a, static_cast<uint32_t>(c), lane, wu)) return true;
    }
    for (uint32_t k = 0; k < a.n_units; k += 64) {
        const uint32_t i = k + static_cast<uint32_t>(lane);
        uint32_t u = 0;
        bool has = false;
        if (i < a.n_units) {
            u = (scan_start + i) % a.n_units;
            const uint64_t claimed =
                __hip_atomic_load(a.next + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            has = claimed < a.units[u].count;
        }
        uint64_t m = __ballot(has);
        while (m != 0) {
            const int l = __builtin_ctzll(m);
            const uint32_t uc =
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u), l));
            if (claim_chunk(a, uc, lane, wu)) return true;
            m &= m - 1;
        }
    }
    return false;
}

// Copy the unit's stack into the wave's LDS layer table, and its header into registers.
__device__ __forceinline__ void load_stack(const KernelArgs &a, WaveUnit &wu, float *ltab, int lane) {
    const float *rec = a.stacks + static_cast<uint64_t>(wu.stack) * a.stack_stride;
    wu.w0 = rec[kHdrW0];
    wu.inv_dz = rec[kHdrInvDz];
    wu.inv_dr = rec[kHdrInvDr];
    wu.nl = static_cast<int>(rec[kHdrLayers]);
    // The empty asm uses these four here, so the wait for their loads is made here and not in
    // the step loop (see claim_chunk).
    asm volatile("" : "+v"(wu.w0), "+v"(wu.inv_dz), "+v"(wu.inv_dr), "+v"(wu.nl));
    for (int j = lane; j < wu.nl * kLayerFields; j += 64) ltab[j] = rec[kStackHeader + j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

// Add a wave's slot into the HBM rows of its unit and clear it. All 64 lanes take part. The row
// pointers are loaded here, through a pointer the compiler cannot see through, so that they are
// not hoisted into scalar registers that the step loop would have to keep.
__device__ __forceinline__ void flush_slot(const KernelArgs &a, uint64_t *slot, const WaveUnit &wu,
                                           int lane) {
    uint64_t *const *cold = a.cold;
    asm volatile("" : "+s"(cold));
    typedef const __attribute__((address_space(1))) uintptr_t global_addr;
    global_addr *table = (global_addr *)(uintptr_t)cold;
    global_u64 *rows[kColdRows];
#pragma unroll
    for (int k = 0; k < kColdRows; ++k) rows[k] = (global_u64 *)table[k];
    const int L = a.lmax;
    for (int j = lane; j < a.slot_words; j += 64) {
        const uint64_t v = slot[j];
        if (v == 0) continue;
        slot[j] = 0;
        global_u64 *dst;
        if (j < kSlotAbs) {
            dst = rows[j] + wu.srow;  // kSlotRd .. kSlotBoundary are kColdRd .. kColdBsteps
        } else if (j < kSlotAbs + L) {
            dst = rows[kColdAbsorbed] + static_cast<uint64_t>(wu.srow) * a.abs_stride + (j - kSlotAbs);
        } else {
            dst = rows[kColdAz] + static_cast<uint64_t>(wu.prow) * a.nz + (j - (kSlotAbs + L));
        }
        hbm_add(dst, v);
    }
}

template <bool kProfiles, bool kArz, bool kAgg, bool kExit>
__global__ __launch_bounds__(1024) void transport_kernel(const KernelArgs a) {
    extern __shared__ uint64_t lds[];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wid = static_cast<int>(threadIdx.x >> 6);
    const int waves = static_cast<int>(blockDim.x >> 6);

    // Clear the wave regions, the A(r,z) tile and the exit-angle / time bins.
    for (int i = static_cast<int>(threadIdx.x); i < a.lds_words; i += static_cast<int>(blockDim.x))
        lds[i] = 0;
    __syncthreads();

    uint64_t *slot = lds + a.ext_words + wid * a.wave_words;
    float *ltab = reinterpret_cast<float *>(slot + ((a.slot_words + 1) & ~1));
    const uint32_t ltab_lds = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_float *)ltab));
    uint32_t *tile = reinterpret_cast<uint32_t *>(lds + a.tile_off);
    const int off_az = kSlotAbs + a.lmax;
    const float last_z = static_cast<float>(a.nz - 1), last_r = static_cast<float>(a.nr - 1);
    // The A(r,z) row lengths (grid and LDS tile) and the tile's LDS address, held in vector
    // registers: the step loop's scalar registers are full, and from a spill slot each would be
    // a v_readlane per event.
    uint32_t nr_v = static_cast<uint32_t>(a.nr), tile_r_v = static_cast<uint32_t>(a.tile_r);
    uint32_t tile_v = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_u32 *)tile));
    uint32_t expo_v = kUnitExponent;  // unit_open's exponent bits
    asm volatile("" : "+v"(nr_v), "+v"(tile_r_v), "+v"(tile_v), "+v"(expo_v));
    // Likewise the A(r,z), R_r and T_r base addresses, used in branches of the step.
    uint64_t arz_v = reinterpret_cast<uintptr_t>(a.arz), rr_v = reinterpret_cast<uintptr_t>(a.rr),
             tr_v = reinterpret_cast<uintptr_t>(a.tr);
    asm volatile("" : "+v"(arz_v), "+v"(rr_v), "+v"(tr_v));
    global_u64 *const arz = (global_u64 *)(uintptr_t)arz_v;
    const float last_a = static_cast<float>(a.na - 1);
    const float last_t = static_cast<float>(a.nt - 1);
    const uint64_t seed = a.key[0];
    const uint32_t key0 = static_cast<uint32_t>(seed), key1 = static_cast<uint32_t>(seed >> 32);

    // ---- wave-uniform scheduling state
    const uint32_t total_waves = gridDim.x * static_cast<uint32_t>(waves);
    const uint32_t gw = blockIdx.x * static_cast<uint32_t>(waves) + static_cast<uint32_t>(wid);
    const uint32_t scan_start =
        static_cast<uint32_t>((static_cast<uint64_t>(gw) * a.n_units) / total_waves);
    bool cursor_done = false;
    WaveUnit wu;
    bool have = acquire_unit(a, lane, scan_start, cursor_done, wu);
    if (have) load_stack(a, wu, ltab, lane);
    bool draining = false;

    // ---- lane state. Which lanes hold a photon (alive_m), and which of those stopped on a layer
    // boundary whose crossing is not made yet (pend_m), are wave-uniform masks in scalar
    // registers rather than per-lane booleans: the wave's votes on them are then free.
    uint64_t alive_m = 0, pend_m = 0;
    float x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 1, w = 0, sleft = 0;
    int layer = 0;
    uint32_t id_lo = 0, id_hi = 0, blk = 0, idx = 8;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, q6 = 0, q7 = 0;
    uint32_t ev = 0, an = 0, bs = 0;
    uint64_t acc = 0;
    float opl = 0;  // kExit: optical path length so far (cm): sum of n s over the hops

    while (have) {
        for (;;) {
            // ---- relaunch: hand the next ids of the wave's unit to lanes without a photon
            uint64_t need = ~alive_m;  // workgroups are whole waves: every lane is active here
            while (need != 0 && !draining) {
                if (wu.cur >= wu.end && !claim_chunk(a, wu.u, lane, wu)) {
                    draining = true;
                    break;
                }
                const uint32_t n = static_cast<uint32_t>(__popcll(need));
                const uint64_t avail = wu.end - wu.cur;
                const uint32_t m = avail < n ? static_cast<uint32_t>(avail) : n;
                const uint32_t rank = lane_rank(need);
                // launch (physics.hpp:58-66), by select on every lane: see the note on loop state.
                const uint64_t start_m = need & wave_mask(rank < m);
                const bool start = lane_in(start_m);
                const uint64_t id = wu.cur + rank;
                const uint32_t lo = static_cast<uint32_t>(id), hi = static_cast<uint32_t>(id >> 32);
                const Block4 blk0 = philox4x32_10(lo, hi, 0u, wu.word3, key0, key1);
                id_lo = start ? lo : id_lo;
                id_hi = start ? hi : id_hi;
                q4 = start ? blk0.v0 : q4;
                q5 = start ? blk0.v1 : q5;
                q6 = start ? blk0.v2 : q6;
                q7 = start ? blk0.v3 : q7;
                blk = start ? 1u : blk;
                idx = start ? 4u : idx;
                x = start ? 0.0f : x;
                y = start ? 0.0f : y;
                z = start ? 0.0f : z;
                ux = start ? 0.0f : ux;
                uy = start ? 0.0f : uy;
                uz = start ? 1.0f : uz;
                w = start ? wu.w0 : w;
                sleft = start ? 0.0f : sleft;
                if (kExit) opl = start ? 0.0f : opl;
                layer = start ? 0 : layer;
                acc = start ? 0ull : acc;
                ev = start ? 0u : ev;
                bs = start ? 0u : bs;
                an = start ? 0u : an;
                pend_m &= ~start_m;
                alive_m |= start_m;
                wu.cur += m;
                need = ~alive_m;
            }

            if (draining && alive_m == 0) break;
            const bool alive = lane_in(alive_m);

            // ---- one hop-drop-spin step (physics.hpp:264-277). Every lane runs it, so that no
            // divergent branch encloses the loop-carried state; a lane without a photon (only while
            // its wave drains a unit) steps stale values and scores nothing. The layer's constants
            // are read from the wave's LDS table at every step rather than carried in registers.
            LayerRec P;
            load_layer<0, 3>(P, ltab_lds, layer);
            {
                // Refill when fewer than 4 buffered words are left. In a wave some lane needs
                // a block at nearly every step, so every lane computes its next block and keeps
                // it by select: a branch would cost the same block plus copies of the buffer.
                const Block4 nb = philox4x32_10(id_lo, id_hi, blk, wu.word3, key0, key1);
                const bool take = idx >= 4;
                q0 = take ? q4 : q0;
                q1 = take ? q5 : q1;
                q2 = take ? q6 : q2;
                q3 = take ? q7 : q3;
                q4 = take ? nb.v0 : q4;
                q5 = take ? nb.v1 : q5;
                q6 = take ? nb.v2 : q6;
                q7 = take ? nb.v3 : q7;
                blk += take ? 1u : 0u;
                idx -= take ? 4u : 0u;
            }
            // r_j = word idx + j of the 8 buffered words (idx in 0..3)
            const bool i1 = (idx & 1u) != 0, i2 = (idx & 2u) != 0;
            const uint32_t r0 = i2 ? (i1 ? q3 : q2) : (i1 ? q1 : q0);
            const uint32_t r1 = i2 ? (i1 ? q4 : q3) : (i1 ? q2 : q1);
            const uint32_t r2 = i2 ? (i1 ? q5 : q4) : (i1 ? q3 : q2);
            const uint32_t r3 = i2 ? (i1 ? q6 : q5) : (i1 ? q4 : q3);

            // step_size (physics.hpp:149-158). A lane waiting at a boundary does not step: its
            // step length is 0, it draws no words and keeps its remaining step, and it takes no
            // part in the drop below.
            const bool go = !lane_in(pend_m);
            const bool fresh = sleft == 0.0f;
            float s = fresh ? __builtin_amdgcn_logf(unit_open(r0, expo_v)) * P.f[kNl2InvMut]
                            : sleft * P.f[kInvMut];
            uint32_t c = go && fresh ? 1u : 0u;
            s = go ? s : 0.0f;
            sleft = go ? 0.0f : sleft;

            // hit_boundary (physics.hpp:162-178)
            const float zb = uz > 0.0f ? P.f[kZ1] : P.f[kZ0];
            const float dl = (zb - z) * __builtin_amdgcn_rcpf(uz);
            const uint64_t hit_m = ~pend_m & wave_mask(uz != 0.0f) & wave_mask(s > dl);
            const bool hit = lane_in(hit_m);  // go && uz != 0 && s > dl
            if (hit) {
                sleft = (s - dl) * P.f[kMut];
                s = dl;
            }
            // hop (physics.hpp:180-184)
            x = __builtin_fmaf(s, ux, x);
            y = __builtin_fmaf(s, uy, y);
            z = __builtin_fmaf(s, uz, z);
            // time of flight: the hop's optical path, n s, added as a product then a sum (the
            // oracle's float32 arithmetic, so both bin the same photons alike)
            if (kExit) opl = opl + s * P.f[kN];

            bool died = false;
            // With kAgg the event's A_z and A(r,z)-tile scores are made after the branch, where
            // every lane of the wave is active.
            bool agg_z = false, agg_t = false;
            int agg_iz = 0;
            uint32_t agg_ti = 0, agg_bin = 0;
            uint64_t agg_q = 0;
            if (!hit && go) {
                // drop (physics.hpp:188-193) and its score (gpu_tally.hpp:51-56)
                const float dwa = w * P.f[kAlbedoA];
                w -= dwa;
                ++ev;
                an += bad_score(ux, uy, uz, dwa);
                const uint64_t q = to_fixed(dwa);
                acc += q;
                // A(r,z) tile add: its carry is resolved after the spin, so that the LDS atomic's
                // returned value is not waited for right away.
                bool tile_pending = false;
                uint32_t tile_old = 0, tile_bin = 0;
                if (kProfiles && alive) {
                    const int iz = bin_of(z, wu.inv_dz, last_z);
                    if (kAgg) {
                        agg_z = true;
                        agg_iz = iz;
                        agg_q = q;
                    } else {
                        lds_add(slot + off_az + iz, q);
                    }
                    if (kArz) {
                        const float rr = __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, y * y));
                        const int ir = bin_of(rr, wu.inv_dr, last_r);
                        tile_bin = __umul24(static_cast<uint32_t>(iz), nr_v) +
                                   static_cast<uint32_t>(ir);
                        if (iz < a.tile_z && ir < a.tile_r) {
                            const uint32_t t = __umul24(static_cast<uint32_t>(iz), tile_r_v) +
                                               static_cast<uint32_t>(ir);
                            if (kAgg) {
                                agg_t = true;
                                agg_ti = t;
                                agg_bin = tile_bin;
                            } else {
                                tile_old = __hip_atomic_fetch_add((lds_u32 *)(uintptr_t)(tile_v + 4u * t),
                                                                  static_cast<uint32_t>(q), __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
                                tile_pending = true;
                            }
                        } else {
                            hbm_add(arz + tile_bin, q);
                        }
                    }
                }
                // spin (physics.hpp:116-144) with the HG sampler (physics.hpp:101-113)
                const float xi_hg = unit_open(c ? r1 : r0, expo_v);
                const float xi_az = unit_open(c ? r2 : r1, expo_v);
                c += 2;
                float cost;
                if (P.f[kHgC] == 0.0f) {
                    cost = __builtin_fmaf(2.0f, xi_hg, -1.0f);
                } else {
                    const float den = __builtin_fmaf(P.f[kHgC], xi_hg, P.f[kHgB]);
                    const float t = P.f[kHgA] * __builtin_amdgcn_rcpf(den);
                    cost = __builtin_fmaf(-t, t, P.f[kHgD]) * P.f[kHgE];
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
                const float n2 = __builtin_fmaf(nx, nx, __builtin_fmaf(ny, ny, nz_ * nz_));
                const float inv = __builtin_amdgcn_rsqf(n2);
                ux = nx * inv;
                uy = ny * inv;
                uz = nz_ * inv;
                if (kArz && tile_pending) {
                    // A counter that wrapped carries 2^32 into the HBM bin (exact: the atomic
                    // returned the value it added to). The test is the add's carry-out and the
                    // q >= 2^32 comparison to_fixed made, both masks: one vector instruction.
                    uint32_t sum;
                    const bool wrapped = __builtin_add_overflow(tile_old, static_cast<uint32_t>(q), &sum);
                    if (wrapped || q >= (1ull << 32))
                        hbm_add(arz + tile_bin, ((q >> 32) + (wrapped ? 1u : 0u)) << 32);
                }
            }
            if (kAgg) {
                aggregate_lds_add(slot + off_az, agg_z, agg_iz, agg_q, lane);
                if (kArz) aggregate_tile_add(tile, a.arz, agg_t, agg_ti, agg_bin, agg_q, lane);
            }
            // cross_or_not, deferred: a lane that reached a boundary waits there until
            // kCrossBatch lanes of the wave wait (or every live lane waits, or the wave drains),
            // and then they cross together. The crossing branch is divergent code that the whole
            // wave issues; batching crossings issues it a few times less often. Each photon still
            // makes the same crossing with the same random word, only later.
            pend_m |= hit_m & alive_m;
            if (pend_m != 0 &&
                (__popcll(pend_m) >= kCrossBatch || pend_m == alive_m || draining)) {
                if (lane_in(pend_m)) {
                    // cross_or_not (physics.hpp:199-246) with Fresnel (physics.hpp:71-98)
                    ++bs;
                    load_layer<3, 2>(P, ltab_lds, layer);
                    const bool up = uz < 0.0f;
                    const float ni = P.f[kN];
                    const float nt = up ? P.f[kNUp] : P.f[kNDown];
                    const float cc = up ? P.f[kCcTop] : P.f[kCcBot];
                    const float eta = up ? P.f[kEtaUp] : P.f[kEtaDown];
                    const float ca1 = up ? -uz : uz;
                    float r = 1.0f, ca2 = 0.0f;
                    if (ca1 > cc) {
                        if (ni == nt) {
                            r = 0.0f;
                            ca2 = ca1;
                        } else if (ca1 > 1.0f - 1e-6f) {
                            r = up ? P.f[kR0Up] : P.f[kR0Down];
                            ca2 = ca1;
                        } else if (ca1 < 1e-6f) {
                            r = 1.0f;
                        } else {
                            const float sa1 = __builtin_amdgcn_sqrtf(__builtin_fmaf(-ca1, ca1, 1.0f));
                            const float sa2 = sa1 * eta;
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
                    const float xi = unit_open(c ? r1 : r0, expo_v);
                    ++c;
                    if (xi > r) {
                        ux *= eta;
                        uy *= eta;
                        uz = up ? -ca2 : ca2;
                        const bool exit_top = up && layer == 0;
                        const bool exit_bot = !up && layer == wu.nl - 1;
                        if (exit_top || exit_bot) {
                            an += bad_score(ux, uy, uz, w);
                            if (alive) {
                                const uint64_t qw = to_fixed(w);
                                lds_add(slot + (exit_top ? kSlotRd : kSlotTt), qw);
                                if (kProfiles) {
                                    const float rr =
                                        __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, y * y));
                                    const int ir = bin_of(rr, wu.inv_dr, last_r);
                                    global_u64 *row =
                                        (global_u64 *)(uintptr_t)(exit_top ? rr_v : tr_v) +
                                        static_cast<uint64_t>(wu.prow) * a.nr;
                                    hbm_add(row + ir, qw);
                                    if (kExit && a.na > 0) {
                                        // MCML's exit angle: from the normal, acos of the
                                        // escaping |uz| (= ca2), in na bins of pi / (2 na)
                                        const int ia = bin_of(acosf(ca2), a.inv_da, last_a);
                                        lds_add(lds + (exit_top ? 0 : a.na) + ia, qw);
                                        hbm_add((exit_top ? a.rra : a.tra) +
                                                    __umul24(static_cast<uint32_t>(ir),
                                                             static_cast<uint32_t>(a.na)) +
                                                    ia,
                                                qw);
                                    }
                                    if (kExit && a.nt > 0) {
                                        // time of flight opl / c, in nt bins of dt, clamped
                                        const int it = bin_of(opl, a.inv_dopl, last_t);
                                        lds_add(lds + 2 * a.na + (exit_top ? 0 : a.nt) + it, qw);
                                    }
                                }
                            }
                            died = true;
                        } else {
                            if (alive && acc != 0) lds_add(slot + kSlotAbs + layer, acc);
                            acc = 0;
                            layer += up ? -1 : 1;
                        }
                    } else {
                        uz = -uz;
                    }
                }
                pend_m = 0;
            }
            const bool waits = lane_in(pend_m);  // still at a boundary, its crossing not made yet
            // roulette (physics.hpp:250-259, 274-276), after the step's crossing if it had one
            if (!died && !waits && w < 1e-4f) {
                if (w == 0.0f) {
                    died = true;
                } else {
                    const uint32_t rw = c == 1 ? r1 : (c == 2 ? r2 : r3);
                    ++c;
                    if (rw < kUnitBelowTenth) {  // unit_open(rw) < CHANCE = 0.1f
                        // w / 0.1f, correctly rounded: the product by 10 and one correction
                        // by its exact residual (4 instructions against 11 for the division)
                        // equal the IEEE quotient for every float w in [2^-64, 1e-4), checked
                        // exhaustively; below 2^-125 they would not (the residual underflows),
                        // so weights under 2^-64 (albedos below ~1e-15) take the division.
                        if (w >= 0x1p-64f) {
                            const float q10 = w * 10.0f;
                            w = __builtin_fmaf(__builtin_fmaf(-q10, 0.1f, w), 10.0f, q10);
                        } else {
                            w = w / 0.1f;
                        }
                    } else {
                        died = true;
                    }
                }
            }
            idx += c;
            // Steps made so far: with no crossing pending, every step was a drop (ev) or
            // stopped on a boundary it has since crossed (bs).
            if (ev + bs >= a.step_limit && !died && !waits) {
                ++an;  // a photon this long is reported, not silently dropped
                died = true;
            }
            if (died && alive) {
                if (acc != 0) lds_add(slot + kSlotAbs + layer, acc);
                if (ev != 0) lds_add(slot + kSlotEvents, ev);
                if (bs != 0) lds_add(slot + kSlotBoundary, bs);
                if (an != 0) lds_add(slot + kSlotAnomalies, an);
            }
            alive_m &= ~wave_mask(died);
        }
        // The unit is done for this wave: publish its slot, then find another unit.
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        flush_slot(a, slot, wu, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        have = acquire_unit(a, lane, scan_start, cursor_done, wu);
        if (have) load_stack(a, wu, ltab, lane);
        layer = 0;  // keeps the idle lanes' table index inside the new stack
        draining = false;
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
    if (kProfiles && a.ext_words > 0) {
        __syncthreads();
        for (int i = static_cast<int>(threadIdx.x); i < a.ext_words; i += static_cast<int>(blockDim.x)) {
            const uint64_t v = lds[i];
            if (v == 0) continue;
            const int na2 = 2 * a.na;
            uint64_t *dst = i < a.na ? a.ra + i
                          : i < na2 ? a.ta + (i - a.na)
                          : i < na2 + a.nt ? a.rt + (i - na2)
                                           : a.ttt + (i - na2 - a.nt);
            hbm_add(dst, v);
        }
    }
}

// Clears n words. Used instead of a memset so that the clear can be captured in a graph.
__global__ void clear_words(uint64_t *p, uint64_t n) {
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = 0;
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

#define HIP_TRY(call)                                                                              \
    do {                                                                                           \
        const hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                      \
            throw EngineError(TMC_ERR_RUNTIME, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                   " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

namespace {

double critical_cos(double ni, double nt) {
    return ni > nt ? std::sqrt(1.0 - (nt * nt) / (ni * ni)) : 0.0;
}

// One stack record: header and per-layer constants, derived like the reference's make_geometry
// (config.hpp:363-390: depths accumulated in double, then cast).
void fill_stack(int nl, double n_above, double n_below, const tmc_layer *layers, double dz,
                double dr, float *rec) {
    {
        const float n0 = static_cast<float>(n_above), n1 = static_cast<float>(layers[0].n);
        const float r = (n0 - n1) / (n0 + n1);
        rec[kHdrW0] = 1.0f - r * r;  // launch weight in float, as physics.hpp:62 computes it
    }
    rec[kHdrInvDz] = 1.0f / static_cast<float>(dz);
    rec[kHdrInvDr] = 1.0f / static_cast<float>(dr);
    rec[kHdrLayers] = static_cast<float>(nl);
    double zacc = 0.0;
    for (int i = 0; i < nl; ++i) {
        const tmc_layer &L = layers[i];
        const double n_top = i == 0 ? n_above : layers[i - 1].n;
        const double n_bot = i + 1 == nl ? n_below : layers[i + 1].n;
        float *r = rec + kStackHeader + i * kLayerFields;
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
        r[kEtaUp] = static_cast<float>(L.n / n_top);
        r[kEtaDown] = static_cast<float>(L.n / n_bot);
        const double t_up = (n_top - L.n) / (n_top + L.n), t_down = (n_bot - L.n) / (n_bot + L.n);
        r[kR0Up] = static_cast<float>(t_up * t_up);
        r[kR0Down] = static_cast<float>(t_down * t_down);
    }
}

struct DeviceInfo {
    int id = 0;
    int cus = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
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
    d.lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                         : prop.sharedMemPerBlock;
    d.name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ", " +
             std::to_string(prop.multiProcessorCount) + " CUs)";
    return d;
}

template <bool P, bool R, bool A, bool T> const void *kernel_ptr() {
    return reinterpret_cast<const void *>(&transport_kernel<P, R, A, T>);
}

// A stream and its two timing events, reused across runs: creating a stream costs about 0.1 ms,
// as much as a small run's copies. A Resident takes a set from its device's pool and returns it
// when it is destroyed; the sets live until the process ends.
struct StreamSet {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
std::mutex g_stream_mutex;
std::vector<std::vector<StreamSet>> g_stream_pool;  // per device ordinal

// A set for `device`, which must be the current device.
StreamSet take_streams(int device) {
    {
        std::lock_guard<std::mutex> lock(g_stream_mutex);
        if (g_stream_pool.size() > static_cast<size_t>(device) && !g_stream_pool[device].empty()) {
            const StreamSet s = g_stream_pool[device].back();
            g_stream_pool[device].pop_back();
            return s;
        }
    }
    StreamSet s;
    HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    if (hipEventCreate(&s.e0) != hipSuccess || hipEventCreate(&s.e1) != hipSuccess) {
        if (s.e0) (void)hipEventDestroy(s.e0);
        (void)hipStreamDestroy(s.stream);
        throw EngineError(TMC_ERR_RUNTIME, "hipEventCreate failed");
    }
    return s;
}

// Back to the pool, unless the stream was left capturing a graph (an error during capture):
// such a set is destroyed.
void give_streams(int device, const StreamSet &s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s.stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) {
        (void)hipEventDestroy(s.e0);
        (void)hipEventDestroy(s.e1);
        (void)hipStreamDestroy(s.stream);
        return;
    }
    std::lock_guard<std::mutex> lock(g_stream_mutex);
    if (g_stream_pool.size() <= static_cast<size_t>(device)) g_stream_pool.resize(static_cast<size_t>(device) + 1);
    g_stream_pool[device].push_back(s);
}

}  // namespace

// ------------------------------------------------------------------ resident engine state

// Everything the kernel needs for one problem shape, on the host.
struct Problem {
    std::vector<UnitRec> units;
    std::vector<float> stacks;  // n_stacks * stack_stride floats
    uint32_t stack_stride = 0;
    int lmax = 1, abs_stride = 1, nz = 1, nr = 1;
    size_t n_srows = 0, n_prows = 0;
    bool profiles = true, arz = false;
    bool aggregate = false;  // lane aggregation of the A_z / A(r,z) adds
    int na = 0;  // exit-angle bins tallied (0: none)
    int nt = 0;  // time bins tallied (0: none)
    double dt = 0.0;  // their width (ns)
    uint64_t photons = 0;
    int reach_z = 1;  // depth bins that can receive weight (for the A(r,z) tile)
};

// A fixed region of device memory that a Resident carves its buffers from instead of allocating
// them (the no-allocation entry point): 256-byte aligned pieces, in order, from `base`.
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

// Device bytes a Resident for problem `p` takes from an Arena: the unit table, the stack table,
// the key and the cleared output range, each rounded up to 256 bytes.
size_t arena_bytes(const Problem &p) {
    auto up = [](size_t b) { return (std::max<size_t>(b, 8) + 255) / 256 * 256; };
    const size_t U = p.units.size(), S = p.n_srows, PR = p.n_prows;
    const size_t out_words = 1 + U + 5 * S + S * p.abs_stride +
                             (p.profiles ? PR * (p.nz + 2 * p.nr) : 0) +
                             (p.arz ? static_cast<size_t>(p.nz) * p.nr : 0) +
                             (p.na > 0 ? 2 * static_cast<size_t>(p.na) * (1 + p.nr) : 0) +
                             2 * static_cast<size_t>(p.nt);
    return up(U * sizeof(UnitRec)) + up(p.stacks.size() * sizeof(float)) + up(8 * (1 + kColdRows)) +
           up(out_words * 8);
}

// Device memory, kernel arguments and launch shape for one problem shape. A batched plan keeps
// one of these and replays a captured graph (clear + transport) on it. With an Arena, buffers
// come from the arena instead of hipMalloc, work goes to the null stream and no events are
// created (the kernel is timed on the host around its completion).
class Resident {
  public:
    Resident(const DeviceInfo &dev, const Problem &p, const tmc_tuning &t, Arena *arena = nullptr)
        : dev_(dev), arena_(arena) {
        std::memset(&a_, 0, sizeof a_);
        a_.n_units = static_cast<uint32_t>(p.units.size());
        a_.stack_stride = p.stack_stride;
        a_.lmax = p.lmax;
        a_.abs_stride = p.abs_stride;
        a_.nz = p.nz;
        a_.nr = p.nr;
        a_.step_limit = t.step_limit ? static_cast<uint32_t>(std::min<uint64_t>(t.step_limit, 0xffffffffu))
                                     : 1000000000u;
        profiles_ = p.profiles;
        arz_ = p.arz;
        agg_ = p.aggregate;
        exit_ = p.na > 0 || p.nt > 0;
        a_.slot_words = kSlotAbs + p.lmax + (p.profiles ? p.nz : 0);
        a_.wave_words = ((a_.slot_words + 1) & ~1) + p.lmax * kLayerFields / 2;

        const size_t cap = t.lds_budget > 0 ? std::min<size_t>(t.lds_budget, dev.lds_per_block)
                                            : dev.lds_per_block;
        int wpg = t.waves_per_group > 0 ? t.waves_per_group : 16;
        if (wpg > 16) throw EngineError(TMC_ERR_CONFIG, "waves_per_group must be <= 16");
        auto wave_bytes = [&](int w) { return static_cast<size_t>(w) * a_.wave_words * 8; };
        while (wpg > 1 && wave_bytes(wpg) > cap) --wpg;
        if (wave_bytes(wpg) > cap)
            throw EngineError(TMC_ERR_CONFIG,
                              "the profile grid is too large for one wave's LDS tally (" +
                                  std::to_string(wave_bytes(1)) + " bytes; the device has " +
                                  std::to_string(cap) + " per workgroup)");
        // The exit-angle and time bins of the workgroup (a few KB) sit in LDS ahead of the wave
        // slots: every escape adds to one of a few hot bins, which HBM atomics serialise.
        const size_t ext_words = 2 * (static_cast<size_t>(p.na) + static_cast<size_t>(p.nt));
        const size_t ext_bytes = ext_words * 8;
        while (wpg > 1 && wave_bytes(wpg) + ext_bytes > cap) --wpg;
        if (wave_bytes(wpg) + ext_bytes > cap)
            throw EngineError(TMC_ERR_CONFIG,
                              "the exit-angle and time bins (2 (na + nt) = " + std::to_string(ext_words) +
                                  " counters) do not fit in LDS beside one wave's tally");
        int tile_z = 0, tile_r = 0;
        if (p.arz) {
            const size_t free_bins = (cap - wave_bytes(wpg) - ext_bytes) / 4;  // u32 counters
            if (t.arz_tile_z >= 0 && t.arz_tile_r >= 0) {
                tile_z = std::min(t.arz_tile_z, p.nz);
                tile_r = std::min(t.arz_tile_r, p.nr);
                if (static_cast<size_t>(tile_z) * tile_r > free_bins)
                    throw EngineError(TMC_ERR_CONFIG, "A(r,z) tile does not fit in LDS");
            } else {
                // Cover every reachable depth and as many radii from the axis as fit (deposits
                // thin out with r); fall back to a square-ish tile when that is too narrow.
                tile_z = p.reach_z;
                tile_r = static_cast<int>(std::min<size_t>(p.nr, free_bins / std::max(tile_z, 1)));
                if (tile_r < std::min(p.nr, 32)) {
                    tile_r = static_cast<int>(std::min<size_t>(
                        p.nr, std::max<size_t>(1, static_cast<size_t>(std::sqrt(static_cast<double>(free_bins))))));
                    tile_z = static_cast<int>(std::min<size_t>(p.reach_z, free_bins / std::max(tile_r, 1)));
                }
            }
        }
        a_.tile_z = tile_z;
        a_.tile_r = tile_r;
        a_.ext_words = static_cast<int32_t>(ext_words);
        a_.tile_off = a_.ext_words + wpg * a_.wave_words;
        a_.lds_words = a_.tile_off + static_cast<int32_t>((static_cast<size_t>(tile_z) * tile_r + 1) / 2);
        lds_bytes_ = static_cast<size_t>(a_.lds_words) * 8;
        wpg_ = wpg;
        groups_ = t.groups > 0 ? t.groups : dev.cus;
        const uint64_t waves_total = static_cast<uint64_t>(groups_) * wpg;
        // Ids claimed per wave at a time: about 64 chunks per wave, at most 1024 ids. The last
        // chunks set how unevenly the waves finish (three-layer stack, 1e8 photons: 230 ms with
        // chunks of 3051, 219 with 381-763); below a few hundred ids the claims themselves start
        // to cost (batched stacks: 703 ms at 1024, 738 at 256).
        a_.chunk = t.chunk > 0 ? static_cast<uint32_t>(t.chunk)
                               : static_cast<uint32_t>(std::clamp<uint64_t>(p.photons / (waves_total * 64), 64, 1024));

        // Device memory: inputs, then everything the clear kernel zeroes in one range.
        const size_t U = p.units.size(), S = p.n_srows, PR = p.n_prows;
        n_units_ = U;
        n_stacks_floats_ = p.stacks.size();
        // the inputs in one region (one copy per upload): the unit table, the stack table, then
        // the key buffer (the seed, then the flush rows' pointers), each at a 256-byte offset
        auto up256 = [](size_t b) { return (std::max<size_t>(b, 8) + 255) / 256 * 256; };
        in_stacks_off_ = up256(U * sizeof(UnitRec));
        in_key_off_ = in_stacks_off_ + up256(p.stacks.size() * sizeof(float));
        in_bytes_ = in_key_off_ + 8 * (1 + kColdRows);
        char *in = static_cast<char *>(alloc(in_bytes_));
        units_ = in;
        stacks_ = in + in_stacks_off_;
        key_ = in + in_key_off_;
        out_words_ = 1 /*cursor*/ + U /*next*/ + 5 * S + S * p.abs_stride +
                     (p.profiles ? PR * (p.nz + 2 * p.nr) : 0) +
                     (p.arz ? static_cast<size_t>(p.nz) * p.nr : 0) +
                     (p.na > 0 ? 2 * static_cast<size_t>(p.na) * (1 + p.nr) : 0) +
                     2 * static_cast<size_t>(p.nt);
        out_ = alloc(out_words_ * 8);
        uint64_t *q = static_cast<uint64_t *>(out_);
        a_.cursor = q; q += 1;
        a_.next = q; q += U;
        a_.rd = q; q += S;
        a_.tt = q; q += S;
        a_.events = q; q += S;
        a_.anomalies = q; q += S;
        a_.bsteps = q; q += S;
        a_.absorbed = q; q += S * p.abs_stride;
        if (p.profiles) {
            a_.az = q; q += PR * p.nz;
            a_.rr = q; q += PR * p.nr;
            a_.tr = q; q += PR * p.nr;
        }
        if (p.arz) {
            a_.arz = q;
            q += static_cast<size_t>(p.nz) * p.nr;
        }
        a_.na = p.na;
        a_.inv_da = p.na > 0 ? static_cast<float>(2.0 * p.na / M_PI) : 0.0f;
        if (p.na > 0) {
            a_.ra = q; q += p.na;
            a_.ta = q; q += p.na;
            a_.rra = q; q += static_cast<size_t>(p.nr) * p.na;
            a_.tra = q; q += static_cast<size_t>(p.nr) * p.na;
        }
        a_.nt = p.nt;
        // a time bin of dt ns is c dt cm of optical path (c = 29.9792458 cm/ns)
        a_.inv_dopl = p.nt > 0 ? static_cast<float>(1.0 / (kLightCmPerNs * p.dt)) : 0.0f;
        if (p.nt > 0) {
            a_.rt = q; q += p.nt;
            a_.ttt = q; q += p.nt;
        }
        a_.units = static_cast<const UnitRec *>(units_);
        a_.stacks = static_cast<const float *>(stacks_);
        a_.key = static_cast<const uint64_t *>(key_);
        a_.cold = reinterpret_cast<uint64_t *const *>(static_cast<uint64_t *>(key_) + 1);
        key_words_[0] = 0;
        key_words_[1 + kColdRd] = reinterpret_cast<uintptr_t>(a_.rd);
        key_words_[1 + kColdTt] = reinterpret_cast<uintptr_t>(a_.tt);
        key_words_[1 + kColdEvents] = reinterpret_cast<uintptr_t>(a_.events);
        key_words_[1 + kColdAnomalies] = reinterpret_cast<uintptr_t>(a_.anomalies);
        key_words_[1 + kColdBsteps] = reinterpret_cast<uintptr_t>(a_.bsteps);
        key_words_[1 + kColdAbsorbed] = reinterpret_cast<uintptr_t>(a_.absorbed);
        key_words_[1 + kColdAz] = reinterpret_cast<uintptr_t>(a_.az);
        n_srows_ = S;
        n_prows_ = PR;

        if (!arena_) {
            const StreamSet ss = take_streams(dev_.id);
            stream_ = ss.stream;
            e0_ = ss.e0;
            e1_ = ss.e1;
        }
        const void *k = kernel();
        HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds_bytes_)));
    }

    ~Resident() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (stream_) {
            (void)hipStreamSynchronize(stream_);
            give_streams(dev_.id, StreamSet{stream_, e0_, e1_});
        }
        for (void *p : allocs_) (void)hipFree(p);
    }
    Resident(const Resident &) = delete;
    Resident &operator=(const Resident &) = delete;

    void upload(const Problem &p, uint64_t seed) {
        if (p.units.size() != n_units_ || p.stacks.size() != n_stacks_floats_)
            throw EngineError(TMC_ERR_CONFIG, "the plan was made for another number of stacks");
        seed_ = seed;
        key_words_[0] = seed;
        // packed as the device region is laid out, then one copy (each copy from pageable host
        // memory costs 10-30 us whatever its size)
        host_in_.assign(in_bytes_, 0);
        std::memcpy(host_in_.data(), p.units.data(), p.units.size() * sizeof(UnitRec));
        std::memcpy(host_in_.data() + in_stacks_off_, p.stacks.data(), p.stacks.size() * sizeof(float));
        std::memcpy(host_in_.data() + in_key_off_, key_words_, sizeof key_words_);
        HIP_TRY(hipMemcpyAsync(units_, host_in_.data(), in_bytes_, hipMemcpyHostToDevice, stream_));
    }

    // Clear the outputs and run the transport kernel; returns the kernel time in seconds (HIP
    // events on this stream). With use_graph the two launches are captured once and replayed.
    double run(bool use_graph) {
        if (use_graph) {
            if (!graph_exec_) {
                HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                enqueue_clear();
                enqueue_transport();
                HIP_TRY(hipStreamEndCapture(stream_, &graph_));
                HIP_TRY(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
            }
            HIP_TRY(hipEventRecord(e0_, stream_));
            HIP_TRY(hipGraphLaunch(graph_exec_, stream_));
            HIP_TRY(hipEventRecord(e1_, stream_));
        } else if (!e0_) {
            // no events (arena mode): host clock from the launch to its completion
            enqueue_clear();
            HIP_TRY(hipStreamSynchronize(stream_));
            const auto t0 = std::chrono::steady_clock::now();
            enqueue_transport();
            HIP_TRY(hipStreamSynchronize(stream_));
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        } else {
            enqueue_clear();
            HIP_TRY(hipEventRecord(e0_, stream_));
            enqueue_transport();
            HIP_TRY(hipEventRecord(e1_, stream_));
        }
        HIP_TRY(hipEventSynchronize(e1_));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0_, e1_));
        return static_cast<double>(ms) / 1000.0;
    }

    // Copy results to caller buffers (any may be null).
    void download(uint64_t *rd, uint64_t *tt, uint64_t *absorbed, uint64_t *events,
                  uint64_t *anomalies, uint64_t *bsteps, uint64_t *az, uint64_t *rr, uint64_t *tr,
                  uint64_t *arz, uint64_t *ra = nullptr, uint64_t *ta = nullptr,
                  uint64_t *rra = nullptr, uint64_t *tra = nullptr, uint64_t *rt = nullptr,
                  uint64_t *ttt = nullptr) {
        // The outputs are one range of device memory from rd on. Up to kOneCopyBytes it comes
        // back in one copy and is spread over the caller's buffers on the host (each copy from
        // device to pageable memory costs 10-30 us whatever its size); larger outputs go
        // straight to the caller's buffers, array by array.
        const uint64_t *base = a_.rd;
        const size_t span = out_words_ - static_cast<size_t>(base - static_cast<const uint64_t *>(out_));
        const bool one_copy = span * 8 <= kOneCopyBytes;
        if (one_copy) {
            host_out_.resize(span);
            HIP_TRY(hipMemcpyAsync(host_out_.data(), base, span * 8, hipMemcpyDeviceToHost, stream_));
            HIP_TRY(hipStreamSynchronize(stream_));
        }
        auto d2h = [&](uint64_t *dst, const uint64_t *src, size_t n) {
            if (!dst || !src || !n) return;
            if (one_copy)
                std::memcpy(dst, host_out_.data() + (src - base), n * 8);
            else
                HIP_TRY(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, stream_));
        };
        const size_t S = n_srows_, PR = n_prows_;
        d2h(rd, a_.rd, S);
        d2h(tt, a_.tt, S);
        d2h(absorbed, a_.absorbed, S * a_.abs_stride);
        d2h(events, a_.events, S);
        d2h(anomalies, a_.anomalies, S);
        d2h(bsteps, a_.bsteps, S);
        if (profiles_) {
            d2h(az, a_.az, PR * a_.nz);
            d2h(rr, a_.rr, PR * a_.nr);
            d2h(tr, a_.tr, PR * a_.nr);
        } else {
            if (az) std::memset(az, 0, PR * a_.nz * 8);
            if (rr) std::memset(rr, 0, PR * a_.nr * 8);
            if (tr) std::memset(tr, 0, PR * a_.nr * 8);
        }
        if (arz_) d2h(arz, a_.arz, static_cast<size_t>(a_.nz) * a_.nr);
        if (a_.na > 0) {
            d2h(ra, a_.ra, a_.na);
            d2h(ta, a_.ta, a_.na);
            d2h(rra, a_.rra, static_cast<size_t>(a_.nr) * a_.na);
            d2h(tra, a_.tra, static_cast<size_t>(a_.nr) * a_.na);
        }
        if (a_.nt > 0) {
            d2h(rt, a_.rt, a_.nt);
            d2h(ttt, a_.ttt, a_.nt);
        }
        HIP_TRY(hipStreamSynchronize(stream_));
    }

    int groups() const { return groups_; }
    int waves_per_group() const { return wpg_; }
    size_t lds_bytes() const { return lds_bytes_; }
    uint32_t chunk() const { return a_.chunk; }
    int tile_z() const { return a_.tile_z; }
    int tile_r() const { return a_.tile_r; }
    const DeviceInfo &device() const { return dev_; }

  private:
    void *alloc(size_t bytes) {
        if (arena_) {
            const size_t need = (std::max<size_t>(bytes, 8) + 255) / 256 * 256;
            if (arena_->used + need > arena_->size)
                throw EngineError(TMC_ERR_CONFIG, "the static device scratch is too small for this run");
            void *p = arena_->base + arena_->used;
            arena_->used += need;
            return p;
        }
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, std::max<size_t>(bytes, 8)));
        allocs_.push_back(p);
        return p;
    }
    const void *kernel() const {
        if (!profiles_) return kernel_ptr<false, false, false, false>();
        static const void *const with_profiles[8] = {
            kernel_ptr<true, false, false, false>(), kernel_ptr<true, false, false, true>(),
            kernel_ptr<true, false, true, false>(),  kernel_ptr<true, false, true, true>(),
            kernel_ptr<true, true, false, false>(),  kernel_ptr<true, true, false, true>(),
            kernel_ptr<true, true, true, false>(),   kernel_ptr<true, true, true, true>()};
        return with_profiles[(arz_ ? 4 : 0) + (agg_ ? 2 : 0) + (exit_ ? 1 : 0)];
    }
    void enqueue_clear() {
        const unsigned grid = static_cast<unsigned>(std::min<size_t>((out_words_ + 255) / 256, 4096));
        hipLaunchKernelGGL(clear_words, dim3(grid), dim3(256), 0, stream_,
                           static_cast<uint64_t *>(out_), static_cast<uint64_t>(out_words_));
        HIP_TRY(hipGetLastError());
    }
    void enqueue_transport() {
        const dim3 grid(static_cast<unsigned>(groups_)), block(static_cast<unsigned>(wpg_ * 64));
        void *args[] = {&a_};
        HIP_TRY(hipLaunchKernel(kernel(), grid, block, args, lds_bytes_, stream_));
    }

    DeviceInfo dev_;
    Arena *arena_ = nullptr;
    KernelArgs a_;
    bool profiles_ = true, arz_ = false, agg_ = false, exit_ = false;
    int wpg_ = 16, groups_ = 0;
    size_t lds_bytes_ = 0, out_words_ = 0, n_units_ = 0, n_stacks_floats_ = 0, n_srows_ = 0, n_prows_ = 0;
    void *units_ = nullptr, *stacks_ = nullptr, *key_ = nullptr, *out_ = nullptr;
    size_t in_bytes_ = 0, in_stacks_off_ = 0, in_key_off_ = 0;
    std::vector<char> host_in_;       // the packed inputs of the last upload
    std::vector<uint64_t> host_out_;  // the outputs of the last download (one-copy case)
    static constexpr size_t kOneCopyBytes = size_t{8} << 20;
    uint64_t seed_ = 0;
    uint64_t key_words_[1 + kColdRows] = {};  // host copy of the key buffer (must outlive the copy)
    std::vector<void *> allocs_;
    hipStream_t stream_ = nullptr;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
};

// ------------------------------------------------------------------ single run

namespace {

// Checks a single run's spec and options; returns the spec with the photon override applied.
tmc_spec checked_single(const tmc_spec &spec_in, const tmc_opts &opt) {
    tmc_spec spec = spec_in;
    if (opt.photons != 0) spec.photons = opt.photons;
    validate_spec(spec);
    if (opt.batches < 1) throw EngineError(TMC_ERR_CONFIG, "batches must be >= 1");
    if (static_cast<uint64_t>(opt.batches) > spec.photons)
        throw EngineError(TMC_ERR_CONFIG, "batches cannot exceed the number of photons");
    const bool profiles = (opt.flags & TMC_FLAG_NO_PROFILES) == 0;
    if ((opt.flags & TMC_FLAG_ARZ) != 0 && !profiles)
        throw EngineError(TMC_ERR_CONFIG, "A(r,z) is a profile: it cannot be combined with no-profiles");
    if ((opt.flags & TMC_FLAG_AGGREGATE) != 0 && !profiles)
        throw EngineError(TMC_ERR_CONFIG, "lane aggregation applies to the profiles: it cannot be "
                                          "combined with no-profiles");
    if ((opt.flags & TMC_FLAG_TIME) != 0) {
        if (!profiles)
            throw EngineError(TMC_ERR_CONFIG,
                              "the time-resolved profiles cannot be combined with no-profiles");
        if (opt.nt < 1) throw EngineError(TMC_ERR_CONFIG, "the time-resolved profiles need nt >= 1");
        if (!(opt.dt > 0.0) || !std::isfinite(opt.dt))
            throw EngineError(TMC_ERR_CONFIG, "the time bin width dt must be > 0 (ns)");
    }
    if ((opt.flags & TMC_FLAG_ANGLES) != 0) {
        if (!profiles)
            throw EngineError(TMC_ERR_CONFIG,
                              "the angle profiles cannot be combined with no-profiles");
        if (spec.na < 1) throw EngineError(TMC_ERR_CONFIG, "the angle profiles need na >= 1");
    }
    return spec;
}

// One stack cut into the reference's batches (BatchPlan::first_of), one unit per batch.
Problem single_problem(const tmc_spec &spec, const tmc_opts &opt) {
    Problem p;
    const int L = spec.num_layers;
    const uint64_t B = static_cast<uint64_t>(opt.batches);
    p.lmax = L;
    p.abs_stride = L;
    p.nz = spec.nz;
    p.nr = spec.nr;
    p.profiles = (opt.flags & TMC_FLAG_NO_PROFILES) == 0;
    p.arz = (opt.flags & TMC_FLAG_ARZ) != 0;
    p.aggregate = (opt.flags & TMC_FLAG_AGGREGATE) != 0;
    p.na = (opt.flags & TMC_FLAG_ANGLES) != 0 ? spec.na : 0;
    p.nt = (opt.flags & TMC_FLAG_TIME) != 0 ? opt.nt : 0;
    p.dt = opt.dt;
    p.photons = spec.photons;
    p.n_srows = B;
    p.n_prows = B;
    p.stack_stride = kStackHeader + L * kLayerFields;
    p.stacks.assign(p.stack_stride, 0.0f);
    fill_stack(L, spec.n_above, spec.n_below, spec.layers, spec.dz, spec.dr, p.stacks.data());
    double depth = 0.0;
    for (int i = 0; i < L; ++i) depth += spec.layers[i].d;
    const double reach = std::floor(depth / spec.dz) + 2.0;
    p.reach_z = reach < static_cast<double>(spec.nz) ? static_cast<int>(reach) : spec.nz;
    const uint64_t base = spec.photons / B, extra = spec.photons % B;
    for (uint64_t b = 0; b < B; ++b) {
        UnitRec u;
        u.first = b * base + (b < extra ? b : extra);
        u.count = base + (b < extra ? 1 : 0);
        u.stack = 0;
        u.prow = static_cast<uint32_t>(b);
        u.srow = static_cast<uint32_t>(b);
        u.word3 = opt.stream;
        p.units.push_back(u);
    }
    return p;
}

tmc_tuning tuning_of(const tmc_opts &opt) {
    tmc_tuning t;
    t.waves_per_group = opt.waves_per_group;
    t.groups = opt.groups;
    t.chunk = opt.chunk;
    t.arz_tile_z = opt.arz_tile_z;
    t.arz_tile_r = opt.arz_tile_r;
    t.step_limit = opt.step_limit;
    return t;
}

}  // namespace

namespace {

void check_out(const tmc_opts &opt, const tmc_out *out) {
    if ((opt.flags & TMC_FLAG_ARZ) != 0 && out->arz == nullptr)
        throw EngineError(TMC_ERR_CONFIG, "A(r,z) requested but tmc_out.arz is NULL");
    if ((opt.flags & TMC_FLAG_ANGLES) != 0 && (!out->ra || !out->ta || !out->rra || !out->tra))
        throw EngineError(TMC_ERR_CONFIG, "angle profiles requested but tmc_out.ra/ta/rra/tra is NULL");
    if ((opt.flags & TMC_FLAG_TIME) != 0 && (!out->rt || !out->tt_t))
        throw EngineError(TMC_ERR_CONFIG, "time-resolved profiles requested but tmc_out.rt/tt_t is NULL");
}

// Upload, run and download one single run on `r`; fills every field of `out`.
void run_single_on(Resident &r, const DeviceInfo &dev, const Problem &p, const tmc_spec &spec,
                   const tmc_opts &opt, tmc_out *out) {
    const bool want_arz = (opt.flags & TMC_FLAG_ARZ) != 0;
    const bool want_angles = (opt.flags & TMC_FLAG_ANGLES) != 0;
    const bool want_time = (opt.flags & TMC_FLAG_TIME) != 0;
    r.upload(p, opt.seed);
    out->kernel_seconds = r.run(false);
    r.download(out->rd, out->tt, out->absorbed, out->events, out->anomalies, out->boundary_steps,
               out->az, out->rr, out->tr, want_arz ? out->arz : nullptr,
               want_angles ? out->ra : nullptr, want_angles ? out->ta : nullptr,
               want_angles ? out->rra : nullptr, want_angles ? out->tra : nullptr,
               want_time ? out->rt : nullptr, want_time ? out->tt_t : nullptr);
    {
        const double n0 = spec.n_above, n1 = spec.layers[0].n;
        const double rr = (n0 - n1) / (n0 + n1);
        out->specular = rr * rr;  // host double, as the reference reports it
    }
    std::snprintf(out->device, sizeof out->device, "%s", dev.name.c_str());
    out->groups = r.groups();
    out->waves_per_group = r.waves_per_group();
    out->lds_bytes = static_cast<int32_t>(r.lds_bytes());
    out->chunk = static_cast<int32_t>(r.chunk());
    out->arz_tile_z = r.tile_z();
    out->arz_tile_r = r.tile_r();
}

// ---- the no-allocation entry point's device scratch
//
// kScratchBytes of device memory per GPU, part of this library's code object: every GPU the
// library runs on holds it from the library's first use there until the process ends. In host code `scratch` names a host-side stand-in, not device memory:
// its device address is taken once per device with hipGetSymbolAddress, with that device
// current, and only that address is used.
constexpr size_t kScratchWords = size_t{2} << 20;
constexpr size_t kScratchBytes = kScratchWords * 8;  // 16 MiB
static __device__ __align__(256) uint64_t scratch[kScratchWords];

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr;  // per device ordinal; null until set up

// The device address of `scratch` on the current device `device`.
void *static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr.size() <= static_cast<size_t>(device)) g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
    void *&ptr = g_scratch_ptr[static_cast<size_t>(device)];
    if (!ptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw EngineError(TMC_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr = p;
    }
    return ptr;
}

}  // namespace

void run_transport(const tmc_spec &spec_in, const tmc_opts &opt, tmc_out *out) {
    const auto t0 = std::chrono::steady_clock::now();
    const tmc_spec spec = checked_single(spec_in, opt);
    check_out(opt, out);
    const DeviceInfo dev = device_info(opt.device);
    const Problem p = single_problem(spec, opt);
    Resident r(dev, p, tuning_of(opt));
    run_single_on(r, dev, p, spec, opt, out);
    out->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t static_scratch_bytes() { return kScratchBytes; }

void run_transport_static(const tmc_spec &spec_in, const tmc_opts &opt, tmc_out *out) {
    const auto t0 = std::chrono::steady_clock::now();
    const tmc_spec spec = checked_single(spec_in, opt);
    check_out(opt, out);
    const Problem p = single_problem(spec, opt);
    const size_t need = arena_bytes(p);
    if (need > kScratchBytes)
        throw EngineError(TMC_ERR_CONFIG, "this run needs " + std::to_string(need) +
                                              " bytes of device scratch; the static buffer holds " +
                                              std::to_string(kScratchBytes) + " (use tmc_hip_run)");
    const DeviceInfo dev = device_info(opt.device);
    Arena arena;
    arena.base = static_cast<char *>(static_scratch(dev.id));
    arena.size = kScratchBytes;
    Resident r(dev, p, tuning_of(opt), &arena);
    run_single_on(r, dev, p, spec, opt, out);
    out->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------ batched runs

struct BatchedPlan {
    DeviceInfo dev;
    Problem shape;
    tmc_batched_opts opt;
    std::unique_ptr<Resident> res;
};

namespace {

void check_stack(const tmc_stack &s, int i) {
    tmc_spec spec;
    std::memset(&spec, 0, sizeof spec);
    spec.photons = s.photons;
    spec.dz = spec.dr = 1.0;
    spec.nz = spec.nr = spec.na = 1;
    spec.n_above = s.n_above;
    spec.n_below = s.n_below;
    spec.num_layers = s.num_layers;
    std::memcpy(spec.layers, s.layers, sizeof spec.layers);
    try {
        validate_spec(spec);
    } catch (const EngineError &e) {
        throw EngineError(TMC_ERR_CONFIG, "stack " + std::to_string(i) + ": " + e.what());
    }
}

// Units of a batched run: stack s, batch b covers the ids of the reference's BatchPlan over the
// stack's photon budget; the stream of stack s uses counter word 3 = s. Units are listed from
// the largest budget down, so the longest work starts first.
void batched_units(const tmc_stack *stacks, int n, int batches, Problem &p) {
    p.units.clear();
    for (int s = 0; s < n; ++s) {
        const uint64_t N = stacks[s].photons, B = static_cast<uint64_t>(batches);
        const uint64_t base = N / B, extra = N % B;
        for (uint64_t b = 0; b < B; ++b) {
            UnitRec u;
            u.first = b * base + (b < extra ? b : extra);
            u.count = base + (b < extra ? 1 : 0);
            u.stack = static_cast<uint32_t>(s);
            u.prow = static_cast<uint32_t>(s);
            u.srow = static_cast<uint32_t>(s * B + b);
            u.word3 = static_cast<uint32_t>(s);
            p.units.push_back(u);
        }
    }
    std::stable_sort(p.units.begin(), p.units.end(),
                     [](const UnitRec &x, const UnitRec &y) { return x.count > y.count; });
    p.photons = 0;
    for (int s = 0; s < n; ++s) p.photons += stacks[s].photons;
}

void batched_stacks(const tmc_stack *stacks, int n, const tmc_batched_opts &o, Problem &p) {
    p.stacks.assign(static_cast<size_t>(n) * p.stack_stride, 0.0f);
    for (int s = 0; s < n; ++s)
        fill_stack(stacks[s].num_layers, stacks[s].n_above, stacks[s].n_below, stacks[s].layers,
                   o.dz, o.dr, p.stacks.data() + static_cast<size_t>(s) * p.stack_stride);
}

}  // namespace

BatchedPlan *batched_create(int n_stacks, const tmc_batched_opts &o) {
    if (n_stacks < 1) throw EngineError(TMC_ERR_CONFIG, "n_stacks must be >= 1");
    if (o.batches_per_stack < 1) throw EngineError(TMC_ERR_CONFIG, "batches_per_stack must be >= 1");
    if (!(o.dz > 0) || !(o.dr > 0) || o.nz < 1 || o.nr < 1)
        throw EngineError(TMC_ERR_CONFIG, "dz, dr must be > 0 and nz, nr >= 1");
    if (static_cast<uint64_t>(n_stacks) * o.batches_per_stack > 0xffffffffull)
        throw EngineError(TMC_ERR_CONFIG, "too many stacks x batches");
    if ((o.flags & (TMC_FLAG_ARZ | TMC_FLAG_ANGLES | TMC_FLAG_TIME)) != 0)
        throw EngineError(TMC_ERR_CONFIG, "batched runs tally A_z, R_r and T_r per stack; A(r,z), "
                                          "the angle and the time-resolved profiles are single-run "
                                          "options");
    auto plan = std::make_unique<BatchedPlan>();
    plan->dev = device_info(o.device);
    plan->opt = o;
    Problem &p = plan->shape;
    p.lmax = kMaxLayers;
    p.abs_stride = kMaxLayers;
    p.nz = o.nz;
    p.nr = o.nr;
    p.profiles = (o.flags & TMC_FLAG_NO_PROFILES) == 0;
    p.arz = false;
    if ((o.flags & TMC_FLAG_AGGREGATE) != 0 && !p.profiles)
        throw EngineError(TMC_ERR_CONFIG, "lane aggregation applies to the profiles: it cannot be "
                                          "combined with no-profiles");
    p.aggregate = (o.flags & TMC_FLAG_AGGREGATE) != 0;
    p.n_srows = static_cast<size_t>(n_stacks) * o.batches_per_stack;
    p.n_prows = static_cast<size_t>(n_stacks);
    p.stack_stride = kStackHeader + kMaxLayers * kLayerFields;
    p.units.resize(static_cast<size_t>(n_stacks) * o.batches_per_stack);
    p.stacks.resize(static_cast<size_t>(n_stacks) * p.stack_stride);
    p.photons = o.photons_hint ? o.photons_hint : static_cast<uint64_t>(n_stacks) * 100000ull;
    tmc_tuning t;
    t.waves_per_group = o.waves_per_group;
    t.groups = o.groups;
    t.chunk = o.chunk;
    t.arz_tile_z = t.arz_tile_r = 0;
    t.step_limit = o.step_limit;
    plan->res = std::make_unique<Resident>(plan->dev, p, t);
    return plan.release();
}

void batched_execute(BatchedPlan &plan, const tmc_stack *stacks, int n_stacks, uint64_t seed,
                     tmc_batched_out *out) {
    const auto t0 = std::chrono::steady_clock::now();
    Problem &p = plan.shape;
    if (static_cast<size_t>(n_stacks) != p.n_prows)
        throw EngineError(TMC_ERR_CONFIG, "the plan was made for " + std::to_string(p.n_prows) +
                                              " stacks, got " + std::to_string(n_stacks));
    for (int s = 0; s < n_stacks; ++s) {
        check_stack(stacks[s], s);
        if (stacks[s].photons < static_cast<uint64_t>(plan.opt.batches_per_stack))
            throw EngineError(TMC_ERR_CONFIG, "stack " + std::to_string(s) +
                                                  ": batches cannot exceed the number of photons");
    }
    batched_units(stacks, n_stacks, plan.opt.batches_per_stack, p);
    batched_stacks(stacks, n_stacks, plan.opt, p);
    plan.res->upload(p, seed);
    out->kernel_seconds = plan.res->run((plan.opt.flags & TMC_FLAG_NO_GRAPH) == 0);
    plan.res->download(out->rd, out->tt, out->absorbed, out->events, out->anomalies,
                       out->boundary_steps, out->az, out->rr, out->tr, nullptr);
    std::snprintf(out->device, sizeof out->device, "%s", plan.dev.name.c_str());
    out->groups = plan.res->groups();
    out->waves_per_group = plan.res->waves_per_group();
    out->lds_bytes = static_cast<int32_t>(plan.res->lds_bytes());
    out->chunk = static_cast<int32_t>(plan.res->chunk());
    out->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void batched_destroy(BatchedPlan *plan) { delete plan; }

// ------------------------------------------------------------------ launch-shape autotuning

namespace {

// Candidate shapes: waves per workgroup x workgroups per compute unit. One 16-wave workgroup
// gives one workgroup-wide A(r,z) tile the most LDS; more, smaller workgroups split the LDS and
// spread the claim traffic. 4 waves per SIMD is what the transport kernel's registers allow with
// one workgroup per CU; 20 waves per CU (5 per SIMD) is the register limit.
struct Shape {
    int wpg, per_cu;
};
constexpr Shape kShapes[] = {{16, 1}, {12, 1}, {8, 1}, {8, 2}, {10, 2}, {6, 2},
                             {4, 4},  {5, 4},  {4, 2}, {4, 5}};

double check_fraction(double f) {
    if (!(f > 0.0) || f > 1.0) throw EngineError(TMC_ERR_CONFIG, "pilot_fraction must be in (0, 1]");
    return f;
}

uint64_t scaled_budget(uint64_t photons, double f, uint64_t at_least) {
    const double v = std::ceil(static_cast<double>(photons) * f);
    return std::max<uint64_t>(at_least, static_cast<uint64_t>(v));
}

// Times a pilot at every candidate shape, then confirms the choice at a larger scale; returns
// the fastest tuning (base plus shape) and fills the report. The tile of a single run follows
// from the LDS budget of the shape.
//
// Pilot launches are short (about a dozen photons per lane at 1/32 of the three-layer run), so
// their time is dominated by the last, longest photon walks and their ranking need not hold at
// the full budget: on the three-layer run 12x1 beat 16x1 by 4% in the pilot and then ran 14%
// slower at 1e8 photons. So every shape whose pilot lies within kContendFrac of the fastest pilot
// is re-timed, with the default shape (listed first), at `confirm` (8 times the pilot fraction,
// at most the full budget), and a shape replaces the default only if it is more than kTuneMargin
// faster there.
constexpr double kContendFrac = 0.10;
constexpr double kTuneMargin = 0.03;
constexpr int kMaxContenders = 3;

struct Candidate {
    tmc_tuning tuning;
    int entry;  // index in the report
};

double time_shape(const DeviceInfo &dev, const Problem &p, const tmc_tuning &t, uint64_t seed,
                  int launches, std::unique_ptr<Resident> *keep) {
    auto res = std::make_unique<Resident>(dev, p, t);
    res->upload(p, seed);
    double tmin = 0.0;
    for (int k = 0; k < launches; ++k) {
        const double dt = res->run(false);
        tmin = k == 0 ? dt : std::min(tmin, dt);
    }
    if (keep) *keep = std::move(res);
    return tmin;
}

template <class MakeProblem>
tmc_tuning tune_shapes(const DeviceInfo &dev, MakeProblem make_problem, const tmc_tuning &base,
                       uint64_t seed, double fraction, tmc_tune_report *rep) {
    tmc_tune_report local;
    tmc_tune_report &r = rep ? *rep : local;
    std::memset(&r, 0, sizeof r);
    r.pilot_fraction = fraction;
    r.best = -1;
    const Problem pilot = make_problem(fraction);
    std::vector<Candidate> fits;
    for (const Shape &sh : kShapes) {
        if (r.n_tried >= TMC_TUNE_MAX) break;
        tmc_tuning t = base;
        t.waves_per_group = sh.wpg;
        t.groups = dev.cus * sh.per_cu;
        t.lds_budget = dev.lds_per_cu / static_cast<size_t>(sh.per_cu);
        t.arz_tile_z = t.arz_tile_r = -1;
        std::unique_ptr<Resident> res;
        double tmin;
        try {
            tmin = time_shape(dev, pilot, t, seed, 3, &res);
        } catch (const EngineError &e) {
            if (e.code() == TMC_ERR_CONFIG) continue;  // does not fit this shape's LDS budget
            throw;
        }
        if (res->waves_per_group() != sh.wpg) continue;  // the slots did not fit: shrunk
        if (pilot.arz) {
            t.arz_tile_z = res->tile_z();
            t.arz_tile_r = res->tile_r();
        }
        tmc_tune_entry &e = r.tried[r.n_tried];
        e.waves_per_group = sh.wpg;
        e.groups = t.groups;
        e.lds_bytes = static_cast<int32_t>(res->lds_bytes());
        e.arz_tile_z = res->tile_z();
        e.arz_tile_r = res->tile_r();
        e.kernel_seconds = tmin;
        fits.push_back({t, r.n_tried});
        ++r.n_tried;
    }
    if (fits.empty()) throw EngineError(TMC_ERR_CONFIG, "no launch shape fits this problem's LDS tallies");

    // Contenders: the default (first fitting shape) and the shapes whose pilot is close to the
    // fastest pilot, fastest first.
    double fastest = r.tried[fits[0].entry].kernel_seconds;
    for (const Candidate &c : fits) fastest = std::min(fastest, r.tried[c.entry].kernel_seconds);
    std::vector<Candidate> others;
    for (size_t i = 1; i < fits.size(); ++i)
        if (r.tried[fits[i].entry].kernel_seconds <= fastest * (1.0 + kContendFrac)) others.push_back(fits[i]);
    std::sort(others.begin(), others.end(), [&](const Candidate &a, const Candidate &b) {
        return r.tried[a.entry].kernel_seconds < r.tried[b.entry].kernel_seconds;
    });
    if (others.size() > kMaxContenders - 1) others.resize(kMaxContenders - 1);
    r.best = fits[0].entry;
    tmc_tuning best = fits[0].tuning;
    if (others.empty()) return best;

    r.confirm_fraction = std::min(1.0, 8.0 * fraction);
    const Problem big = make_problem(r.confirm_fraction);
    const double t_default = time_shape(dev, big, fits[0].tuning, seed, 2, nullptr);
    r.tried[fits[0].entry].confirm_seconds = t_default;
    double best_t = t_default;
    for (const Candidate &c : others) {
        const double t = time_shape(dev, big, c.tuning, seed, 2, nullptr);
        r.tried[c.entry].confirm_seconds = t;
        if (t < t_default * (1.0 - kTuneMargin) && t < best_t) {
            best_t = t;
            best = c.tuning;
            r.best = c.entry;
        }
    }
    return best;
}

}  // namespace

void autotune_single(const tmc_spec &spec_in, const tmc_opts &opt, double pilot_fraction,
                     tmc_opts *tuned, tmc_tune_report *report) {
    const double f = check_fraction(pilot_fraction);
    const tmc_spec spec = checked_single(spec_in, opt);
    const DeviceInfo dev = device_info(opt.device);
    tmc_opts po = opt;
    po.photons = 0;
    auto make_problem = [&](double fraction) {
        tmc_spec ps = spec;
        ps.photons = scaled_budget(spec.photons, fraction, static_cast<uint64_t>(opt.batches));
        return single_problem(ps, po);
    };
    const tmc_tuning best = tune_shapes(dev, make_problem, tuning_of(opt), opt.seed, f, report);
    *tuned = opt;
    tuned->waves_per_group = best.waves_per_group;
    tuned->groups = best.groups;
    tuned->arz_tile_z = best.arz_tile_z;
    tuned->arz_tile_r = best.arz_tile_r;
}

void batched_autotune(BatchedPlan &plan, const tmc_stack *stacks, int n_stacks, uint64_t seed,
                      double pilot_fraction, tmc_tune_report *report) {
    const double f = check_fraction(pilot_fraction);
    if (static_cast<size_t>(n_stacks) != plan.shape.n_prows)
        throw EngineError(TMC_ERR_CONFIG, "the plan was made for " + std::to_string(plan.shape.n_prows) +
                                              " stacks, got " + std::to_string(n_stacks));
    const uint64_t Bs = static_cast<uint64_t>(plan.opt.batches_per_stack);
    for (int s = 0; s < n_stacks; ++s) {
        check_stack(stacks[s], s);
        if (stacks[s].photons < Bs)
            throw EngineError(TMC_ERR_CONFIG, "stack " + std::to_string(s) +
                                                  ": batches cannot exceed the number of photons");
    }
    auto make_problem = [&](double fraction) {
        std::vector<tmc_stack> ps(stacks, stacks + n_stacks);
        for (int s = 0; s < n_stacks; ++s)
            ps[static_cast<size_t>(s)].photons = scaled_budget(stacks[s].photons, fraction, Bs);
        Problem pilot = plan.shape;
        batched_units(ps.data(), n_stacks, plan.opt.batches_per_stack, pilot);
        batched_stacks(ps.data(), n_stacks, plan.opt, pilot);
        return pilot;
    };
    tmc_tuning base;
    base.chunk = plan.opt.chunk;
    base.step_limit = plan.opt.step_limit;
    base.arz_tile_z = base.arz_tile_r = 0;
    const tmc_tuning best = tune_shapes(plan.dev, make_problem, base, seed, f, report);
    plan.res.reset();
    plan.res = std::make_unique<Resident>(plan.dev, plan.shape, best);
    plan.opt.waves_per_group = best.waves_per_group;
    plan.opt.groups = best.groups;
}

// ------------------------------------------------------------------ parity hook

void philox_stream(uint64_t seed, uint64_t id, uint32_t word3, int64_t n, uint32_t *out,
                   int device) {
    if (n < 0) throw EngineError(TMC_ERR_CONFIG, "n must be >= 0");
    if (n == 0) return;
    device_info(device);
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, static_cast<size_t>(n) * 4));
    const int64_t blocks = (n + 3) / 4;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((blocks + 255) / 256, 1024));
    hipLaunchKernelGGL(philox_stream_kernel, dim3(grid), dim3(256), 0, 0,
                       static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32),
                       static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), word3, n,
                       static_cast<uint32_t *>(d));
    const hipError_t e = hipGetLastError();
    const hipError_t c = hipMemcpy(out, d, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIP_TRY(e);
    HIP_TRY(c);
}

bool device_available() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

}  // namespace tmc
