// The learned surrogate: an ensemble of MLPs trained and evaluated with hand-written BF16 MFMA
// kernels (FP32 accumulation) on gfx950. No reference counterpart.
//
// Network (per ensemble head): inputs -> depth x [hidden, GELU] -> outputs, GELU in its tanh
// form. The forward and input-gradient products are `gemm_nt`: C[m, n] = sum_k A[m, k] * B[n, k]
// with both operands K-contiguous in memory (the layout the MFMA wants per lane); the weight
// gradient, whose K is the batch, is `gemm_tn_group`: C[m, n] = sum_k A[k, m] * B[k, n], reading the
// activations and gradients as the other two wrote them and transposing in LDS:
//   forward          Z = H_prev . W^T     gemm_nt  A = H_prev [batch x in],  B = Wt [out x in]
//   input gradient   dH = dZ . W          gemm_nt  A = dZ [batch x out],     B = W  [in x out]
//   weight gradient  dW = dZ^T . H_prev   gemm_tn_group  A = dZ [batch x out], B = H_prev [batch x in]
// so Adam keeps bf16 weights in both orientations (Wt and W) and nothing else is transposed.
// Heads are independent networks batched along gridDim.z. Matrices are padded to multiples of
// 128 with zeros that stay zero (zero inputs, zero loss weights). Weight gradients reduce over the
// batch (K = batch), so they are split along K across workgroups; bias gradients are column sums
// of dZ, made 64 rows at a time in the epilogue that writes dZ. Adam sums both kinds of partials
// in a fixed order: training is bit-reproducible.
//
// MFMA: v_mfma_f32_16x16x32_bf16. Lane l holds A row l%16 and B column l%16, k = 8*(l/16)..+7;
// the 16x16 result has column l%16, rows 4*(l/16)..+3 (checked on the device). A workgroup of
// four waves computes a 128x128 tile, each wave 64x64 (4x4 MFMA tiles), K in steps of 32, operands
// staged through a ring of LDS stages (see gemm_nt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_rng.h"
#include "engine.h"

namespace tmc {
namespace mlp {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef uint16_t bf16;  // storage type

__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(bf16, static_cast<__bf16>(f));  // v_cvt_pk_bf16_f32, nearest even
#else
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<bf16>(u >> 16);
#endif
}
// Two floats -> two bf16 in one word (lo in the low half), nearest even: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float float2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2_t{lo, hi}, bf16x2_t));
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;
constexpr float kLog2E = 1.4426950408889634f;

// GELU in its tanh form, 0.5 z (1 + tanh u) with u = sqrt(2/pi) (z + 0.044715 z^3), evaluated
// as z s with s = 0.5 (1 + tanh u) = 1 / (1 + e^(-2u)) and e^(-2u) = 2^(z (e1 + e3 z^2)): one
// v_exp_f32 and one v_rcp_f32.
constexpr float kGeluE1 = -2.0f * kGeluC * kLog2E;
constexpr float kGeluE3 = -2.0f * kGeluC * kGeluA * kLog2E;
__device__ __forceinline__ float gelu_s(float z) {
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(kGeluE3, z * z, kGeluE1));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
// d/dz: s + z (1 - tanh^2 u) u' / 2 = s + 2 z s (1 - s) sqrt(2/pi) (1 + 3 a z^2)
__device__ __forceinline__ float gelu_grad_s(float z, float s) {
    const float du = __builtin_fmaf(3.0f * kGeluA * 2.0f * kGeluC, z * z, 2.0f * kGeluC);
    return __builtin_fmaf(z * s * (1.0f - s), du, s);
}

constexpr int kMaxMats = 8;  // weight matrices of a network (layers)

// kInferHidden is kFwdHidden for inference: no Z and no transposed copy, so its results need no
// LDS staging (below).
enum Epilogue { kFwdHidden = 0, kFwdOut = 1, kBwdData = 2, kInferHidden = 4 };

struct GemmArgs {
    const bf16 *A;
    const bf16 *B;
    int M, N, K, lda, ldb;
    long long sA, sB;  // head strides (elements)
    int splits;        // split-K factor (weight gradients): tile z = head * splits + split
    // epilogue operands (head strides in elements)
    const float *bias;
    long long sBias;
    float *Cf;  // fp32 output [M x N] (Y for kFwdOut, dW partials for the weight gradients)
    long long sCf;
    long long sSplit;  // weight gradients: stride between split partials
    bf16 *Cb;   // bf16 output [M x N] (H or dZ)
    long long sCb;  // head stride of Cb (M * N)
    bf16 *Gd;         // kFwdHidden (training): GELU'(Z) as bf16, M x N in fragment order
    const bf16 *Gin;  // kBwdData: the GELU'(Z) of this layer, as kFwdHidden wrote it
    float *Bp;        // kBwdData: the bias gradient's partials, the column sums of dZ over each
    long long sBp;    // 64 rows: [M / 64 x ldBp] per head (head stride sBp), columns from 0
    int ldBp;
};

// Fragment order: the layout in which the training forward hands GELU'(Z) to the backward data
// product of the same layer. Both compute the same M x N output on the same 128 x 128 tiles and
// wave and fragment split, so a lane's four values (rows mb .. mb + 3 of column n) are stored
// where the other kernel's lane will read them: 8 contiguous bytes per lane, 512 per wave and
// fragment, instead of 4-byte values strided by rows. Offset in elements: tile (row-major over
// the N / 128 column tiles), then wave, fragment (i, j), lane.
__device__ __forceinline__ long long frag_off(int mblk, int nblk, int N, int wave, int i, int j,
                                              int lane) {
    const long long t = static_cast<long long>(mblk >> 7) * (N >> 7) + (nblk >> 7);
    return (((t * 4 + wave) * 16 + i * 4 + j) * 64 + lane) * 4;
}

// C[m, n] = sum_k A[m, k] B[n, k] on a 128 x 128 tile per workgroup: four waves in a 2 x 2
// arrangement, each accumulating 64 x 64 as 4 x 4 MFMA tiles of 16 x 16.
//
// K advances in stages of 32. A stage's A and B tiles (128 x 32 bf16 each) go from global memory
// straight into LDS (global_load_lds_dwordx4: each lane's 16 bytes land at the wave's LDS base +
// 16 x lane, no registers involved), into a ring of kStages stages, so the loads run
// kStages - 1 stages ahead of the multiplications. One barrier per stage, after the wave has
// waited for its own loads of that stage (vmcnt); it also tells every wave that the ring slot the
// next loads overwrite has been read. Each wave then reads its MFMA fragments from LDS (16 bytes
// per lane). Rows are 64 bytes (four 16-byte chunks), and chunk c of row r is stored at chunk
// c ^ s(r), s(r) = ((r >> 1) ^ (r >> 2)) & 3: the 16 rows of a fragment land in distinct banks
// whether the LDS serves 8 or 16 lanes of a 16-byte read per cycle. The loads pick their global
// chunk to match, since a lane's destination in LDS is fixed by its lane number.
//
// The training epilogues, which write bf16 H / dZ, stage the tile through LDS (34 KB, inside
// the 64 KB of their four-stage ring: two workgroups per CU). The others store straight from
// registers and use only the ring; for those the MFMA operands are swapped (the tile is computed
// as B A^T), so that a lane holds four consecutive columns of one row: 16 bytes of fp32, or 8
// bytes of bf16, per store. Their ring has three stages (48 KB): three workgroups per CU, which
// the registers also allow.
constexpr int kBK = 32;                          // K per stage
constexpr int kStageElems = 2 * 128 * kBK;       // bf16 elements of one stage (A then B)
constexpr int kPitch = 136;                      // bf16 row pitch of the epilogue's output tiles
constexpr int kGemmLds = 128 * kPitch * 2;       // bytes: the epilogue's output tile
constexpr bool lds_epilogue(int epi) { return epi == kFwdHidden || epi == kBwdData; }
constexpr bool swapped(int epi) { return epi == kInferHidden || epi == kFwdOut; }
constexpr int stages_of(int epi) { return lds_epilogue(epi) ? 4 : 3; }
constexpr int lds_bytes_of(int epi) {
    return lds_epilogue(epi) && kGemmLds > stages_of(epi) * kStageElems * 2
               ? kGemmLds
               : stages_of(epi) * kStageElems * 2;
}

__device__ __forceinline__ int swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

typedef __attribute__((address_space(1))) const void global_cvoid;
typedef __attribute__((address_space(3))) void lds_void;

// Wait for this wave's global loads until at most N are in flight, and for its LDS accesses, then
// meet the other waves. Written as asm so that the compiler neither waits for every load (a
// fence would) nor moves LDS accesses across it.
template <int N>
__device__ __forceinline__ void stage_sync() {
    static_assert(N >= 0 && N <= 63, "vmcnt is a 6-bit count");
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}
// stage_sync for `ahead` stages (four loads each) still allowed in flight, 0 <= ahead <= kMax
// (wave-uniform)
template <int kMax>
__device__ __forceinline__ void stage_sync_ahead(int ahead) {
    if constexpr (kMax == 0) {
        stage_sync<0>();
    } else {
        if (ahead >= kMax) stage_sync<4 * kMax>();
        else stage_sync_ahead<kMax - 1>(ahead);
    }
}

template <int kEpi>
__global__ __launch_bounds__(256) void gemm_nt(const GemmArgs g) {
    static_assert(swapped(kEpi) || lds_epilogue(kEpi), "an epilogue of each kind");
    constexpr int kStages = stages_of(kEpi);
    extern __shared__ __attribute__((aligned(16))) bf16 smem[];  // staging ring, then T
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int z = static_cast<int>(blockIdx.z);
    const int h = z / g.splits, split = z % g.splits;
    // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin in launch order, so
    // hardware tile L runs on XCD L % 8. Give XCD k the contiguous logical tiles
    // [k T / 8, (k + 1) T / 8): the column tiles of one row block then run on one XCD, which
    // reads that block of A from HBM once into its own L2 instead of once per column tile.
    const int NX = static_cast<int>(gridDim.x), T = NX * static_cast<int>(gridDim.y);
    int tile = static_cast<int>(blockIdx.x) + NX * static_cast<int>(blockIdx.y);
    if (T % 8 == 0) tile = (tile % 8) * (T / 8) + tile / 8;
    const int mblk = (tile / NX) * 128, nblk = (tile % NX) * 128;
    const int m0 = mblk + (wave >> 1) * 64;
    const int n0 = nblk + (wave & 1) * 64;
    const int kspan = g.K / g.splits;
    const int kbeg = split * kspan;
    const int nk = kspan / kBK;
    const bf16 *A = g.A + h * g.sA;
    const bf16 *B = g.B + h * g.sB;

    // loads: wave w moves rows 32 w + 16 i + lane / 4 (i = 0, 1) of A and of B; lane's LDS slot
    // is chunk lane % 4 of its row, so it fetches the chunk stored there
    const int lrow0 = 32 * wave + (lane >> 2), lrow1 = lrow0 + 16;
    const int lc0 = ((lane & 3) ^ swz(lrow0)) * 8, lc1 = ((lane & 3) ^ swz(lrow1)) * 8;
    const bf16 *a0 = A + static_cast<long long>(mblk + lrow0) * g.lda + kbeg + lc0;
    const bf16 *a1 = A + static_cast<long long>(mblk + lrow1) * g.lda + kbeg + lc1;
    const bf16 *b0 = B + static_cast<long long>(nblk + lrow0) * g.ldb + kbeg + lc0;
    const bf16 *b1 = B + static_cast<long long>(nblk + lrow1) * g.ldb + kbeg + lc1;
    const int wdst = 32 * wave * kBK;  // the wave's first row in a tile (elements)
    // fragments: lane holds row (lane % 16), k chunk (lane / 16)
    const int r = lane & 15, kq = lane >> 4;
    const int ch = (kq ^ swz(r)) * 8;
    const int arow = ((wave >> 1) * 64 + r) * kBK + ch, brow = ((wave & 1) * 64 + r) * kBK + ch;

    // Swapped tiles start from the bias (a lane's four columns are consecutive), which takes the
    // bias add out of the epilogue.
    floatx4 acc[4][4];
    const int g4 = (lane >> 4) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        floatx4 init = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (swapped(kEpi))
            init = *reinterpret_cast<const floatx4 *>(g.bias + h * g.sBias + n0 + g4 + 16 * j);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][j] = init;
    }

#define TMC_GEMM_ISSUE(st)                                                                        \
    do {                                                                                          \
        bf16 *As_ = smem + ((st) % kStages) * kStageElems + wdst, *Bs_ = As_ + 128 * kBK;         \
        const int ko = (st) * kBK;                                                                \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(a0 + ko), (lds_void *)As_, 16, 0, 0);   \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(a1 + ko), (lds_void *)(As_ + 16 * kBK), \
                                         16, 0, 0);                                               \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b0 + ko), (lds_void *)Bs_, 16, 0, 0);   \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b1 + ko), (lds_void *)(Bs_ + 16 * kBK), \
                                         16, 0, 0);                                               \
    } while (0)

    for (int st = 0; st < kStages - 1 && st < nk; ++st) TMC_GEMM_ISSUE(st);
    for (int kt = 0; kt < nk; ++kt) {
        // stages issued after kt so far: min(kStages - 2, nk - 1 - kt), four loads each
        const int ahead = nk - 1 - kt < kStages - 2 ? nk - 1 - kt : kStages - 2;
        stage_sync_ahead<kStages - 2>(ahead);
        if (kt + kStages - 1 < nk) TMC_GEMM_ISSUE(kt + kStages - 1);
        const bf16 *As = smem + (kt % kStages) * kStageElems, *Bs = As + 128 * kBK;
        bf16x8 fa[4], fb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            fa[i] = *reinterpret_cast<const bf16x8 *>(As + arow + i * 16 * kBK);
            fb[i] = *reinterpret_cast<const bf16x8 *>(Bs + brow + i * 16 * kBK);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = swapped(kEpi)
                                ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0)
                                : __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
#undef TMC_GEMM_ISSUE
    if constexpr (swapped(kEpi)) {
        // In each 16 x 16 tile the lane holds row m = m0 + 16 i + (lane % 16) and the four
        // columns nb .. nb + 3, nb = n0 + 16 j + 4 (lane / 16), bias included. Every value the
        // ring held has been read, and nothing here touches LDS.
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long row = static_cast<long long>(m0 + 16 * i + r) * g.N + n0 + g4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if constexpr (kEpi == kInferHidden) {
                    const floatx4 &z = acc[i][j];
                    *reinterpret_cast<uint2 *>(g.Cb + h * g.sCb + row + 16 * j) =
                        uint2{pack_bf16x2(gelu(z[0]), gelu(z[1])), pack_bf16x2(gelu(z[2]), gelu(z[3]))};
                } else {  // kFwdOut
                    *reinterpret_cast<floatx4 *>(g.Cf + h * g.sCf + row + 16 * j) = acc[i][j];
                }
            }
        }
        return;
    }
    stage_sync<0>();  // every wave is done with the ring before the epilogue reuses the LDS

    // Epilogue: in each 16 x 16 tile the lane holds column n, rows mb .. mb + 3.
    if constexpr (kEpi == kFwdHidden || kEpi == kBwdData) {
        // bf16 results go through LDS (free after the last stage's barrier): written row-major
        // into T, then read back in 16-byte row pieces, so that every global store is coalesced.
        // The pitch of 136 puts the four row groups of a fragment 16 banks apart.
        bf16 *T = smem;
        bf16 *Cb = g.Cb + h * g.sCb;
        const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = wn + 16 * j + r, rb = wm + 16 * i + (lane >> 4) * 4;
                const int n = nblk + col;
                const long long fo = h * g.sCb + frag_off(mblk, nblk, g.N, wave, i, j, lane);
                float v[4];
                if (kEpi == kFwdHidden) {
                    const float b = g.bias[h * g.sBias + n];
                    float gd[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float zv = acc[i][j][q] + b;
                        const float s = gelu_s(zv);
                        v[q] = zv * s;
                        gd[q] = gelu_grad_s(zv, s);
                    }
                    if (g.Gd)
                        *reinterpret_cast<uint2 *>(g.Gd + fo) =
                            uint2{pack_bf16x2(gd[0], gd[1]), pack_bf16x2(gd[2], gd[3])};
                } else {  // kBwdData
                    const uint2 gg = *reinterpret_cast<const uint2 *>(g.Gin + fo);
                    v[0] = acc[i][j][0] * from_bf16(static_cast<bf16>(gg.x));
                    v[1] = acc[i][j][1] * from_bf16(static_cast<bf16>(gg.x >> 16));
                    v[2] = acc[i][j][2] * from_bf16(static_cast<bf16>(gg.y));
                    v[3] = acc[i][j][3] * from_bf16(static_cast<bf16>(gg.y >> 16));
                }
                const uint32_t p01 = pack_bf16x2(v[0], v[1]), p23 = pack_bf16x2(v[2], v[3]);
                T[rb * kPitch + col] = static_cast<bf16>(p01);
                T[(rb + 1) * kPitch + col] = static_cast<bf16>(p01 >> 16);
                T[(rb + 2) * kPitch + col] = static_cast<bf16>(p23);
                T[(rb + 3) * kPitch + col] = static_cast<bf16>(p23 >> 16);
            }
        }
        __syncthreads();
        if (kEpi == kBwdData && g.Bp) {
            // the bias gradient's partials: column sums of the bf16 dZ over each 64 rows of the
            // tile, in row order (thread -> column tid % 128, half tid / 128)
            const int col = tid & 127, half = tid >> 7;
            float cs = 0.0f;
            for (int rr = 0; rr < 64; ++rr) cs += from_bf16(T[(half * 64 + rr) * kPitch + col]);
            g.Bp[h * g.sBp + static_cast<long long>((mblk >> 6) + half) * g.ldBp + nblk + col] = cs;
        }
        // 128 rows x 16 pieces of 8 bf16, eight pieces per thread
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = tid + 256 * k, row = c >> 4, piece = (c & 15) * 8;
            *reinterpret_cast<uint4 *>(Cb + static_cast<long long>(mblk + row) * g.N + nblk + piece) =
                *reinterpret_cast<const uint4 *>(T + row * kPitch + piece);
        }
    }
}

// stage_sync for `ahead` stages of kPer loads each still allowed in flight, 0 <= ahead <= kMax
template <int kMax, int kPer>
__device__ __forceinline__ void stage_sync_ahead_n(int ahead) {
    if constexpr (kMax == 0) {
        stage_sync<0>();
    } else {
        if (ahead >= kMax) stage_sync<kPer * kMax>();
        else stage_sync_ahead_n<kMax - 1, kPer>(ahead);
    }
}

// The inference products on 128 x 256 tiles (N a multiple of 256): the loop and epilogues of
// gemm_nt<kInferHidden> (bias in the accumulators, swapped operands, GELU, 8-byte bf16 stores)
// and gemm_nt<kFwdOut> (16-byte fp32 stores), with each wave computing 64 x 128 (4 x 8 MFMA
// tiles) instead of 64 x 64. A stage's MFMA work
// per wave doubles against the same barrier and half again as many fragment reads (12 for 32
// MFMAs, against 8 for 16). A stage holds 128 rows of A then 256 of B (24 KB), three stages:
// two workgroups per CU.
constexpr int kWideStages = 3;
constexpr int kWideStageElems = (128 + 256) * kBK;
constexpr int kWideLds = kWideStages * kWideStageElems * 2;  // bytes

template <int kEpi>
__global__ __launch_bounds__(256, 2) void gemm_infer_wide(const GemmArgs g) {
    static_assert(kEpi == kInferHidden || kEpi == kFwdOut, "the swapped inference epilogues");
    extern __shared__ __attribute__((aligned(16))) bf16 smem[];
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int h = static_cast<int>(blockIdx.z);
    const int NX = static_cast<int>(gridDim.x), T = NX * static_cast<int>(gridDim.y);
    int tile = static_cast<int>(blockIdx.x) + NX * static_cast<int>(blockIdx.y);
    if (T % 8 == 0) tile = (tile % 8) * (T / 8) + tile / 8;  // XCD-aware, as in gemm_nt
    const int mblk = (tile / NX) * 128, nblk = (tile % NX) * 256;
    const int m0 = mblk + (wave >> 1) * 64;
    const int n0 = nblk + (wave & 1) * 128;
    const int nk = g.K / kBK;
    const bf16 *A = g.A + h * g.sA;
    const bf16 *B = g.B + h * g.sB;

    // loads: wave w moves rows 32 w + 16 i + lane / 4 (i = 0, 1) of A and 64 w + 16 i + lane / 4
    // (i = 0..3) of B; the lane's LDS slot is chunk lane % 4 of its row (rows of 64 bytes, the
    // chunk swizzle of gemm_nt), so it fetches the chunk stored there
    const int q4 = lane >> 2;
    const int lc = ((lane & 3) ^ swz(q4)) * 8;  // swz depends on the row mod 16, = q4 here
    const bf16 *a0 = A + static_cast<long long>(mblk + 32 * wave + q4) * g.lda + lc;
    const bf16 *b0 = B + static_cast<long long>(nblk + 64 * wave + q4) * g.ldb + lc;
    const long long a16 = 16LL * g.lda, b16 = 16LL * g.ldb;
    // fragments: lane holds row (lane % 16), k chunk (lane / 16)
    const int r = lane & 15, kq = lane >> 4;
    const int ch = (kq ^ swz(r)) * 8;
    const int arow = ((wave >> 1) * 64 + r) * kBK + ch, brow = ((wave & 1) * 128 + r) * kBK + ch;

    floatx4 acc[4][8];
    const int g4 = (lane >> 4) * 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const floatx4 init = *reinterpret_cast<const floatx4 *>(g.bias + h * g.sBias + n0 + g4 + 16 * j);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][j] = init;
    }

#define TMC_WIDE_ISSUE(st)                                                                        \
    do {                                                                                          \
        bf16 *S_ = smem + ((st) % kWideStages) * kWideStageElems;                                 \
        bf16 *As_ = S_ + 32 * wave * kBK, *Bs_ = S_ + 128 * kBK + 64 * wave * kBK;                \
        const int ko = (st) * kBK;                                                                \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(a0 + ko), (lds_void *)As_, 16, 0, 0);   \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(a0 + a16 + ko), (lds_void *)(As_ + 16 * kBK), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b0 + ko), (lds_void *)Bs_, 16, 0, 0);   \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b0 + b16 + ko), (lds_void *)(Bs_ + 16 * kBK), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b0 + 2 * b16 + ko), (lds_void *)(Bs_ + 32 * kBK), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((global_cvoid *)(b0 + 3 * b16 + ko), (lds_void *)(Bs_ + 48 * kBK), 16, 0, 0); \
    } while (0)

    for (int st = 0; st < kWideStages - 1 && st < nk; ++st) TMC_WIDE_ISSUE(st);
    for (int kt = 0; kt < nk; ++kt) {
        const int ahead = nk - 1 - kt < kWideStages - 2 ? nk - 1 - kt : kWideStages - 2;
        stage_sync_ahead_n<kWideStages - 2, 6>(ahead);
        if (kt + kWideStages - 1 < nk) TMC_WIDE_ISSUE(kt + kWideStages - 1);
        const bf16 *As = smem + (kt % kWideStages) * kWideStageElems, *Bs = As + 128 * kBK;
        bf16x8 fa[4], fb[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[i] = *reinterpret_cast<const bf16x8 *>(As + arow + i * 16 * kBK);
#pragma unroll
        for (int j = 0; j < 8; ++j) fb[j] = *reinterpret_cast<const bf16x8 *>(Bs + brow + j * 16 * kBK);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
#undef TMC_WIDE_ISSUE
    // the lane holds row m0 + 16 i + (lane % 16) and the four columns n0 + 16 j + 4 (lane / 16) ..
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long row = static_cast<long long>(m0 + 16 * i + r) * g.N + n0 + g4;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const floatx4 &zz = acc[i][j];
            if constexpr (kEpi == kInferHidden)
                *reinterpret_cast<uint2 *>(g.Cb + h * g.sCb + row + 16 * j) =
                    uint2{pack_bf16x2(gelu(zz[0]), gelu(zz[1])), pack_bf16x2(gelu(zz[2]), gelu(zz[3]))};
            else
                *reinterpret_cast<floatx4 *>(g.Cf + h * g.sCf + row + 16 * j) = zz;
        }
    }
}

// The weight gradient without transposed copies: C[m, n] = sum_k A[k, m] B[k, n] with A = dZ
// [batch x out] and B = H [batch x in] as the forward and backward products wrote them (K = the
// batch, the row index of both), split along K across workgroups, fp32 partials.
// Tiles and waves as in gemm_nt: 128 x 128 per workgroup, 64 x 64 per wave, K in stages of 32.
//
// A stage holds 32 k-rows x 128 columns of each operand, loaded row by row from global memory
// (16-byte loads, 16 lanes a row of 256 bytes, coalesced) and written to LDS; the 16-byte
// chunk c of row r is stored at chunk c ^ s(r), s(r) = 2 (r & 3) + 8 ((r >> 3) & 1). The MFMA wants 8 consecutive k of one column
// per lane, which ds_read_b64_tr_b16 delivers from this layout: within each 16 lanes, lane i
// receives element i % 4 of the 8 bytes that lane 4 j + i / 4 addresses, as its element j
// (j = 0..3). So lane L of group g addresses row 8 g + L / 4 (then 8 g + 4 + L / 4), columns
// c0 + 4 (L % 4) .. + 3 of the fragment's 16 columns c0 .. c0 + 15, and receives column c0 + L at
// k = 8 g .. 8 g + 3 (then 8 g + 4 .. 8 g + 7): the fragment the MFMA reads. Rows are 256 bytes,
// one pass over the 64 banks, so without the swizzle every row would use the same banks. With
// it, the four rows a group reads in one instruction take four different pairs of chunks of
// one 128-byte half of the bank space, and groups 2 h and 2 h + 1 (rows 8 apart), which the
// LDS serves in the same cycle, take opposite halves: no conflicts (counters: half the LDS
// cycles went to bank conflicts with the swizzle 2 (r & 7), which gave both groups the same
// banks).
typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) shortx4 lds_shortx4;
typedef __attribute__((address_space(3))) char lds_char;
constexpr int kTnLds = 2 * kStageElems * 2;  // bytes: two stages
__device__ __forceinline__ int tn_swz(int r) { return 2 * (r & 3) + 8 * ((r >> 3) & 1); }

// Wait for this wave's LDS accesses, then meet the other waves (no wait for global loads, which
// stay in flight across it).
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One 128 x 128 tile of one (head, split) of the product g: logical workgroup `lid` of the
// product's N / 128 x M / 128 x (heads x splits) tiles, column tiles fastest.
//
// The K loop: a stage's operands come from global memory into registers two stages ahead of their
// use (16 bytes x 4 per lane), are written to one of two LDS slots one stage ahead, and the
// fragments of stage kt + 1 are read from LDS (ds_read_b64_tr_b16, the builtin) in between stage
// kt's MFMAs, one read after each MFMA, into a second set of registers: the LDS latency is
// hidden behind the MFMAs, and the barrier each stage (after the stage's writes) also tells every
// wave that the slot the next writes overwrite has been read. The reads are the builtin, which
// the compiler waits for exactly; the staging is through registers rather than global-to-LDS
// copies because with those in flight the compiler waits for every copy before each LDS read.
// The MFMA sequence (operands, K order) is that of a plain loop over the stages.
__device__ __forceinline__ void gemm_tn_tile(const GemmArgs &g, int lid) {
    extern __shared__ __attribute__((aligned(16))) bf16 smem[];
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int NX = g.N / 128, T = NX * (g.M / 128);
    const int z = lid / T, tile = lid % T;
    const int h = z / g.splits, split = z % g.splits;
    const int mblk = (tile / NX) * 128, nblk = (tile % NX) * 128;
    const int kspan = g.K / g.splits;
    const int kbeg = split * kspan;
    const int nk = kspan / kBK;  // >= 1
    const bf16 *A = g.A + h * g.sA;
    const bf16 *B = g.B + h * g.sB;

    // loads: wave w moves rows 8 w + 4 q + lane / 16 (q = 0, 1) of A and of B; the lane's LDS
    // slot is chunk lane % 16 of its row, so it fetches the chunk stored there
    const int lr0 = 8 * wave + (lane >> 4), lr1 = lr0 + 4;
    const int lc0 = ((lane & 15) ^ tn_swz(lr0)) * 8, lc1 = ((lane & 15) ^ tn_swz(lr1)) * 8;
    const bf16 *a0 = A + static_cast<long long>(kbeg + lr0) * g.lda + mblk + lc0;
    const bf16 *a1 = A + static_cast<long long>(kbeg + lr1) * g.lda + mblk + lc1;
    const bf16 *b0 = B + static_cast<long long>(kbeg + lr0) * g.ldb + nblk + lc0;
    const bf16 *b1 = B + static_cast<long long>(kbeg + lr1) * g.ldb + nblk + lc1;
    const int wdst = 8 * wave * 128 + 8 * lane;  // the lane's 16 bytes in a stage's A tile (elements)

    // transposed fragment reads: lane L of group gq reads rows 8 gq + L / 4 (+ 4), columns
    // c0 + 4 (L % 4), c0 = the fragment's first column (a multiple of 16)
    const int gq = lane >> 4, L = lane & 15;
    const int fr0 = 8 * gq + (L >> 2), fr1 = fr0 + 4;
    const int fcol = 4 * (L & 3);  // column offset inside the fragment
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    auto tr_off = [&](int row, int c0) {  // byte offset in a tile of (row, c0 + fcol)
        const int col = c0 + fcol;
        return static_cast<uint32_t>(2 * (row * 128 + (((col >> 3) ^ tn_swz(row)) << 3) + (col & 7)));
    };
    // reads 0..3: k = 8 gq .. + 3 of A's fragments 0..3; 4..7: k = 8 gq + 4 .. + 7; 8..15: B's
    uint32_t off[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        off[i] = tr_off(fr0, wm + 16 * i);
        off[4 + i] = tr_off(fr1, wm + 16 * i);
        off[8 + i] = tr_off(fr0, wn + 16 * i) + 32 * 128 * 2;
        off[12 + i] = tr_off(fr1, wn + 16 * i) + 32 * 128 * 2;
    }
    lds_char *const lds = reinterpret_cast<lds_char *>((lds_void *)smem);

    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    struct Regs {
        uint4 a0, a1, b0, b1;
    };
    auto gload = [&](int st) {
        const long long ko = static_cast<long long>(st) * kBK;
        Regs r;
        r.a0 = *reinterpret_cast<const uint4 *>(a0 + ko * g.lda);
        r.a1 = *reinterpret_cast<const uint4 *>(a1 + ko * g.lda);
        r.b0 = *reinterpret_cast<const uint4 *>(b0 + ko * g.ldb);
        r.b1 = *reinterpret_cast<const uint4 *>(b1 + ko * g.ldb);
        return r;
    };
    auto swrite = [&](int slot, const Regs &r) {
        bf16 *As_ = smem + slot * kStageElems + wdst, *Bs_ = As_ + 32 * 128;
        *reinterpret_cast<uint4 *>(As_) = r.a0;
        *reinterpret_cast<uint4 *>(As_ + 4 * 128) = r.a1;
        *reinterpret_cast<uint4 *>(Bs_) = r.b0;
        *reinterpret_cast<uint4 *>(Bs_ + 4 * 128) = r.b1;
    };
    auto rd = [&](int slot, int u) {
        return __builtin_amdgcn_ds_read_tr16_b64_v4i16(
            reinterpret_cast<lds_shortx4 *>(lds + slot * kStageElems * 2 + off[u]));
    };
    // The operands are swapped (the tile is computed as B^T A, its transpose), so that in each
    // 16 x 16 tile the lane holds row m = lane % 16 and four consecutive columns.
    auto mfma = [&](const shortx4 (&f)[16], int i, int j) {
        const bf16x8 fa = __builtin_bit_cast(bf16x8, __builtin_shufflevector(f[i], f[4 + i], 0, 1, 2, 3, 4, 5, 6, 7));
        const bf16x8 fb = __builtin_bit_cast(bf16x8, __builtin_shufflevector(f[8 + j], f[12 + j], 0, 1, 2, 3, 4, 5, 6, 7));
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb, fa, acc[i][j], 0, 0, 0);
    };
    auto mfma_all = [&](const shortx4 (&f)[16]) {
#pragma unroll
        for (int t = 0; t < 16; ++t) mfma(f, t >> 2, t & 3);
    };
    // this stage's MFMAs (fragments f) with the next stage's reads (from `slot`, into nf)
    auto mfma_read = [&](const shortx4 (&f)[16], shortx4 (&nf)[16], int slot) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            mfma(f, t >> 2, t & 3);
            nf[t] = rd(slot, t);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
        }
    };
    // Stage s goes to LDS slot s % 2. At the top of each loop iteration (kt even): F0 holds
    // stage kt's fragments, Y stage kt + 1's operands, X stage kt + 2's (loads in flight). Loads
    // past the last stage re-read the last one (in bounds, unused).
    const int last = nk - 1;
    auto clampk = [&](int st) { return st < last ? st : last; };
    Regs X = gload(0), Y = gload(clampk(1));
    swrite(0, X);
    X = gload(clampk(2));
    lds_barrier();
    shortx4 F0[16], F1[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) F0[u] = rd(0, u);
    int kt = 0;
    for (; kt + 2 < last; kt += 2) {
        swrite(1, Y);
        Y = gload(kt + 3);
        lds_barrier();
        mfma_read(F0, F1, 1);
        swrite(0, X);
        X = gload(clampk(kt + 4));
        lds_barrier();
        mfma_read(F1, F0, 0);
    }
    // stages kt .. last remain (last - kt is 0, 1 or 2): Y holds stage kt + 1, X stage kt + 2
    if (kt == last) {
        mfma_all(F0);
    } else {
        swrite(1, Y);
        lds_barrier();
        mfma_read(F0, F1, 1);
        if (kt + 1 == last) {
            mfma_all(F1);
        } else {
            swrite(0, X);
            lds_barrier();
            mfma_read(F1, F0, 0);
            mfma_all(F0);
        }
    }
    float *C = g.Cf + h * g.sCf + split * g.sSplit;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long row = static_cast<long long>(mblk + wm + 16 * i + (lane & 15)) * g.N;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            *reinterpret_cast<floatx4 *>(C + row + nblk + wn + 16 * j + 4 * (lane >> 4)) = acc[i][j];
    }
}

// Every weight gradient of a training step in one launch: product m takes the logical
// workgroups [first[m], first[m + 1]). XCD-aware over the whole grid: workgroup L (launch order)
// runs on XCD L % 8, and XCD k takes the contiguous logical workgroups [k W / 8, (k + 1) W / 8),
// i.e. whole (head, split) slices, whose tiles then share that slice's columns of dZ and H in
// one L2. One launch instead of one per layer lets the small products (the first layer's 64
// tiles) run beside the others instead of alone.
struct TnGroup {
    GemmArgs g[kMaxMats];
    int first[kMaxMats + 1];
    int n;
};
__global__ __launch_bounds__(256, 2) void gemm_tn_group(const TnGroup grp) {
    const int W = static_cast<int>(gridDim.x);
    int lid = static_cast<int>(blockIdx.x);
    if (W % 8 == 0) lid = (lid % 8) * (W / 8) + lid / 8;
    int m = 0;
    while (m + 1 < grp.n && lid >= grp.first[m + 1]) ++m;
    gemm_tn_tile(grp.g[m], lid - grp.first[m]);
}


// Batch assembly for a block of 16 rows of head h: row m takes dataset sample idx (Philox of
// (step, m, h) mod n); writes the standardized features as bf16 X [B x in_pad] and the targets
// padded to out_pad as fp32 Y [B x out_pad]. Consecutive threads take consecutive columns, so
// every access is coalesced. Only the first in_used columns of X are written, zero past `in`:
// the rest stay as allocated (zero). The rows are random, so each read is a separate trip to memory: a thread issues all of
// its reads before its first store, and waits for them once rather than once per row. in_used is
// a multiple of 32 and at most in_pad.
constexpr int kGatherRows = 16;
__global__ __launch_bounds__(256) void gather_batch(const float *__restrict__ xd,
                                                    const float *__restrict__ yd, int64_t n,
                                                    int in, int out, int in_used, int in_pad,
                                                    int out_pad, int B,
                                                    const uint64_t *step, uint32_t key0,
                                                    uint32_t key1, bf16 *__restrict__ X,
                                                    float *__restrict__ Y,
                                                    const int64_t *fixed_rows) {
    __shared__ int64_t rows[kGatherRows];
    const int tid = static_cast<int>(threadIdx.x);
    const int m0 = static_cast<int>(blockIdx.x) * kGatherRows;
    const int h = static_cast<int>(blockIdx.y);
    if (tid < kGatherRows) {
        const int m = m0 + tid;
        int64_t idx;
        if (fixed_rows) {
            idx = fixed_rows[m];
        } else {
            const Block4 v = philox4x32_10(static_cast<uint32_t>(m), static_cast<uint32_t>(h),
                                           static_cast<uint32_t>(step[0]), 0x4D4C5031u, key0, key1);
            idx = static_cast<int64_t>((static_cast<uint64_t>(v.v0) * static_cast<uint64_t>(n)) >> 32);
        }
        rows[tid] = idx;
    }
    __syncthreads();
    bf16 *Xh = X + static_cast<long long>(h) * B * in_pad;
    float *Yh = Y + static_cast<long long>(h) * B * out_pad;
    for (int j0 = 0; j0 < in_used; j0 += 32) {
        // 16 rows x 32 columns: thread -> column tid % 32, rows tid / 32 and tid / 32 + 8
        const int j = j0 + (tid & 31), i0 = tid >> 5;
        const float v0 = j < in ? xd[rows[i0] * in + j] : 0.0f;
        const float v1 = j < in ? xd[rows[i0 + 8] * in + j] : 0.0f;
        const bf16 b0 = to_bf16(v0), b1 = to_bf16(v1);
        Xh[static_cast<long long>(m0 + i0) * in_pad + j] = b0;
        Xh[static_cast<long long>(m0 + i0 + 8) * in_pad + j] = b1;
    }
    for (int j = tid; j < out_pad; j += 256) {
        float v[kGatherRows];
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) v[i] = j < out ? yd[rows[i] * out + j] : 0.0f;
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) Yh[static_cast<long long>(m0 + i) * out_pad + j] = v[i];
    }
}

// Weighted squared error over the real outputs on a 64-row x 64-column tile of head h: writes dY
// (bf16, coalesced), the tile's share of the loss into part[h][tile] (summed in a fixed order by
// loss_reduce), and the output bias gradient's partials, the column sums of the bf16 dY over the
// tile's rows in row order (through an LDS tile), into bp[h][row block][column] (row stride
// ld_bp, head stride s_bp).
__global__ __launch_bounds__(256) void loss_grad(const float *__restrict__ Yp,
                                                 const float *__restrict__ Yt,
                                                 const float *__restrict__ w, int B, int out_pad,
                                                 float inv_count, bf16 *__restrict__ dY,
                                                 float *__restrict__ part,
                                                 float *__restrict__ bp, long long s_bp, int ld_bp) {
    __shared__ bf16 tile[64][66];
    __shared__ float red[256];
    const int tx = static_cast<int>(threadIdx.x & 63u), ty = static_cast<int>(threadIdx.x >> 6);
    const int m0 = static_cast<int>(blockIdx.x) * 64, j0 = static_cast<int>(blockIdx.y) * 64;
    const int h = static_cast<int>(blockIdx.z);
    const long long hb = static_cast<long long>(h) * B;
    const float wj = w[j0 + tx];
    // a thread's 16 rows: every read issued before the first store
    float d[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const long long e = (hb + m0 + ty + 4 * k) * out_pad + j0 + tx;
        d[k] = Yp[e] - Yt[e];
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = ty + 4 * k;
        const long long e = (hb + m0 + i) * out_pad + j0 + tx;
        acc += wj * d[k] * d[k];
        const bf16 g = to_bf16(2.0f * wj * d[k] * inv_count);
        dY[e] = g;
        tile[tx][i] = g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (ty == 0) {
        float cs = 0.0f;
        for (int i = 0; i < 64; ++i) cs += from_bf16(tile[tx][i]);
        bp[h * s_bp + static_cast<long long>(blockIdx.x) * ld_bp + j0 + tx] = cs;
    }
    for (int off = 128; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[(static_cast<long long>(h) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// The batch loss of each head: its n tile partials summed in a fixed order, written to
// loss[(step % ring) * heads + h].
__global__ void loss_reduce(const float *part, int n, float inv_count, float *loss,
                            const uint64_t *step, int heads, uint64_t ring) {
    __shared__ float red[256];
    const int h = static_cast<int>(blockIdx.x);
    float s = 0.0f;
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x))
        s += part[static_cast<long long>(h) * n + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = static_cast<int>(blockDim.x) / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[static_cast<long long>(step[0] % ring) * heads + h] = red[0] * inv_count;
}

struct MatInfo {
    long long off;  // offset of the matrix in a head's parameter block (floats)
    int out, in;    // padded shape [out x in]
    long long boff;  // offset of the bias
    long long wt_off, w_off;  // offsets of the bf16 copies in a head's bf16 block
    int bb;                   // first column of the bias in the bias-gradient partials
};
constexpr int kMaxSplits = 8;  // split-K pieces of a weight gradient

struct AdamArgs {
    float *p, *g, *m, *v;  // [heads x P]
    const float *part;     // [heads x splits x P]: split-K partials of the matrix gradients
    int splits;
    const float *bpart;    // [heads x bp_rows x nb]: the bias gradients' partials (64 rows each)
    int bp_rows, nb;
    int grad_only;         // write the summed gradients to g instead of updating
    bf16 *wb;              // [heads x Pb] bf16 copies (Wt then W per matrix)
    long long P, Pb;
    int n_mats;
    MatInfo mats[kMaxMats];
    float beta1, beta2, eps, wd;
    const float *lr;  // device memory, so a schedule can change it between graph replays
    const uint64_t *step;
};

// Adam (Kingma & Ba 2015) with bias correction for one element: parameter p, gradient g,
// moments m and v in, their new values out.
__device__ __forceinline__ void adam_update(const AdamArgs &a, float &p, float g, float &m,
                                            float &v, float c1, float c2) {
    const float gr = g + a.wd * p;
    m = a.beta1 * m + (1.0f - a.beta1) * gr;
    v = a.beta2 * v + (1.0f - a.beta2) * gr * gr;
    p = p - a.lr[0] * (m * c1) / (__builtin_sqrtf(v * c2) + a.eps);
}

// Adam on every parameter of one head (blockIdx.y). Block k < n_mats updates the bias of matrix
// k, its gradient the sum of its partials in row-block order (these blocks go first: each sums
// dozens of partials per element). Blocks n_mats + t take tile t of the weight matrices (64 x 64,
// the tiles of all matrices in order): the gradient is the sum of the matrix's split-K partials
// in split order, and the block writes the bf16 copies the products read: W [out x in] row by
// row like the fp32 parameters, and its transpose [in x out] through an LDS tile, so that every
// global access is coalesced. With grad_only the blocks write
// the summed gradients to g and update nothing: the same sums, for a caller that asks for the
// gradient.
constexpr int kAdamTile = 64;
__global__ __launch_bounds__(256) void adam_params(const AdamArgs a) {
    __shared__ bf16 tile[kAdamTile][kAdamTile + 2];  // [i][o]
    const int h = static_cast<int>(blockIdx.y);
    const float tt = static_cast<float>(a.step[0] + 1);
    const float c1 = 1.0f / (1.0f - __builtin_powf(a.beta1, tt));
    const float c2 = 1.0f / (1.0f - __builtin_powf(a.beta2, tt));
    float *p = a.p + h * a.P, *m = a.m + h * a.P, *v = a.v + h * a.P;
    const int tid = static_cast<int>(threadIdx.x);
    int t = static_cast<int>(blockIdx.x);
    if (t < a.n_mats) {
        const MatInfo mi = a.mats[t];
        const float *bp = a.bpart + static_cast<long long>(h) * a.bp_rows * a.nb + mi.bb;
        for (int o = tid; o < mi.out; o += static_cast<int>(blockDim.x)) {
            const long long e = mi.boff + o;
            // sixteen partial loads in flight at a time, added in row-block order
            float ge = 0.0f;
            int rb = 0;
            for (; rb + 16 <= a.bp_rows; rb += 16) {
                float pv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) pv[u] = bp[static_cast<long long>(rb + u) * a.nb + o];
#pragma unroll
                for (int u = 0; u < 16; ++u) ge += pv[u];
            }
            for (; rb < a.bp_rows; ++rb) ge += bp[static_cast<long long>(rb) * a.nb + o];
            if (a.grad_only) {
                a.g[h * a.P + e] = ge;
                continue;
            }
            float pe = p[e], me = m[e], ve = v[e];
            adam_update(a, pe, ge, me, ve, c1, c2);
            p[e] = pe;
            m[e] = me;
            v[e] = ve;
        }
        return;
    }
    t -= a.n_mats;
    int k = 0;
    for (; k < a.n_mats; ++k) {
        const int n = (a.mats[k].out / kAdamTile) * (a.mats[k].in / kAdamTile);
        if (t < n) break;
        t -= n;
    }
    const MatInfo mi = a.mats[k];
    const int tiles_i = mi.in / kAdamTile;
    const int o0 = (t / tiles_i) * kAdamTile, i0 = (t % tiles_i) * kAdamTile;
    const float *part = a.part + static_cast<long long>(h) * a.splits * a.P;
    bf16 *wb = a.wb + h * a.Pb;
    // rows o0 + r, columns i0 + cg .. cg + 3 as 16-byte accesses (every offset is a multiple of 4
    // floats): thread -> column group tid % 16, rows tid / 16 + 16 q; the partial loads of its
    // four rows (at most kMaxSplits each) are in flight together
    const int cg = (tid & 15) * 4, rq = tid >> 4;
    floatx4 g[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) g[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int sp = 0; sp < kMaxSplits; ++sp) {
        if (sp < a.splits) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                g[q] += *reinterpret_cast<const floatx4 *>(
                    part + sp * a.P + mi.off + static_cast<long long>(o0 + rq + 16 * q) * mi.in + i0 + cg);
        }
    }
    if (a.grad_only) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<floatx4 *>(a.g + h * a.P + mi.off +
                                         static_cast<long long>(o0 + rq + 16 * q) * mi.in + i0 + cg) = g[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = rq + 16 * q;
        const long long rel = static_cast<long long>(o0 + r) * mi.in + i0 + cg;
        floatx4 pe = *reinterpret_cast<const floatx4 *>(p + mi.off + rel);
        floatx4 me = *reinterpret_cast<const floatx4 *>(m + mi.off + rel);
        floatx4 ve = *reinterpret_cast<const floatx4 *>(v + mi.off + rel);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float pj = pe[j], mj = me[j], vj = ve[j];
            adam_update(a, pj, g[q][j], mj, vj, c1, c2);
            pe[j] = pj;
            me[j] = mj;
            ve[j] = vj;
        }
        *reinterpret_cast<floatx4 *>(p + mi.off + rel) = pe;
        *reinterpret_cast<floatx4 *>(m + mi.off + rel) = me;
        *reinterpret_cast<floatx4 *>(v + mi.off + rel) = ve;
        const uint32_t b01 = pack_bf16x2(pe[0], pe[1]), b23 = pack_bf16x2(pe[2], pe[3]);
        *reinterpret_cast<uint2 *>(wb + mi.wt_off + rel) = uint2{b01, b23};
        tile[cg][r] = static_cast<bf16>(b01);
        tile[cg + 1][r] = static_cast<bf16>(b01 >> 16);
        tile[cg + 2][r] = static_cast<bf16>(b23);
        tile[cg + 3][r] = static_cast<bf16>(b23 >> 16);
    }
    __syncthreads();
    const int c = tid & 63;
    for (int r = tid >> 6; r < kAdamTile; r += 4)  // transposed: row i0 + r, column o0 + c
        wb[mi.w_off + static_cast<long long>(i0 + r) * mi.out + o0 + c] = tile[r][c];
}

__global__ void bump_step(uint64_t *step) { step[0] += 1; }

// Copies the bf16 weights from fp32 masters (after initialisation or set_params). Uses the
// parameter and matrix fields of AdamArgs.
__global__ void refresh_bf16(const AdamArgs a) {
    const int h = static_cast<int>(blockIdx.y);
    for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < a.P;
         e += static_cast<long long>(gridDim.x) * blockDim.x) {
        for (int k = 0; k < a.n_mats; ++k) {
            const MatInfo &mi = a.mats[k];
            const long long rel = e - mi.off;
            if (rel >= 0 && rel < static_cast<long long>(mi.out) * mi.in) {
                const int o = static_cast<int>(rel / mi.in), i = static_cast<int>(rel % mi.in);
                const bf16 b = to_bf16(a.p[h * a.P + e]);
                a.wb[h * a.Pb + mi.wt_off + rel] = b;
                a.wb[h * a.Pb + mi.w_off + static_cast<long long>(i) * mi.out + o] = b;
            }
        }
    }
}

}  // namespace mlp
}  // namespace tmc

namespace tmc {
namespace mlp {

// Inputs for prediction: the same rows for every head, bf16, zero-padded.
// Queries -> bf16 network inputs X [B x in_pad] (rows past `rows` and columns past `in` are
// zero), shared by every head: one thread per 8 columns, one 16-byte store. Only the first
// in_used columns (a multiple of 8: the columns the first layer multiplies) are written; the
// others stay as allocated, zero.
__global__ void load_inputs(const float *xf, int rows, int in, int in_used, int in_pad, int B,
                            bf16 *X) {
    const int per_row = in_used / 8;
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(B) * per_row) return;
    const int m = static_cast<int>(t / per_row), j0 = static_cast<int>(t % per_row) * 8;
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = j0 + 2 * k;
        const bool ok = m < rows;
        const uint32_t lo = ok && j < in ? to_bf16(xf[static_cast<long long>(m) * in + j]) : 0u;
        const uint32_t hi = ok && j + 1 < in ? to_bf16(xf[static_cast<long long>(m) * in + j + 1]) : 0u;
        w[k] = lo | (hi << 16);
    }
    *reinterpret_cast<uint4 *>(X + static_cast<long long>(m) * in_pad + j0) = uint4{w[0], w[1], w[2], w[3]};
}

// Ensemble mean of every output and spread (standard deviation over heads) of the first
// n_spread (<= 4), per query row m < rows:
// four rows per block, 64 threads per row, four consecutive outputs per thread (16-byte loads;
// out_pad is a multiple of 4 at least out rounded up to 4).
__global__ __launch_bounds__(256) void ensemble_reduce(const float *Yp, int heads, int B, int out_pad,
                                                       int out, int n_spread, int rows, float *mean,
                                                       float *spread) {
    const int m = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (m >= rows) return;
    for (int c = static_cast<int>(threadIdx.x & 63) * 4; c < out; c += 256) {
        floatx4 s = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int h = 0; h < heads; ++h)
            s += *reinterpret_cast<const floatx4 *>(Yp + (static_cast<long long>(h) * B + m) * out_pad + c);
        const floatx4 mu = s / static_cast<float>(heads);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (c + q < out) mean[static_cast<long long>(m) * out + c + q] = mu[q];
        if (c < n_spread) {  // two passes: no cancellation when the heads nearly agree
            floatx4 ss = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int h = 0; h < heads; ++h) {
                const floatx4 d =
                    *reinterpret_cast<const floatx4 *>(Yp + (static_cast<long long>(h) * B + m) * out_pad + c) - mu;
                ss += d * d;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c + q < n_spread)
                    spread[static_cast<long long>(m) * n_spread + c + q] =
                        heads > 1 ? __builtin_sqrtf(ss[q] / (heads - 1)) : 0.0f;
        }
    }
}

// ---- fused ensemble inference
//
// The whole network of every head for 128 query rows in one workgroup of eight waves (one
// workgroup per CU): the activations stay in LDS from layer to layer, a 128 x 512 bf16 tile, and
// only the inputs come in and each head's output layer goes out. Each wave computes 1/8 of a
// layer's columns for all 128 rows, one half of the rows at a time, with the same MFMA as the
// layer-by-layer kernels (swapped operands, bias-seeded accumulators, K in steps of 32 in
// order), so its results are the same bit for bit. Its weight fragments come straight from
// global memory (L2) into registers, one K step ahead, from a copy of the weights in fragment
// order (pack_fragments): a wave's fragment of 16 output rows x 32 K is 1 KB of consecutive
// bytes, whole cache lines, where the [out x in] layout would give it sixteen half lines. A
// lane's 16 bytes are the MFMA fragment itself, so the K loop has no LDS staging and no barrier.
// The activation fragments are read from LDS (rows padded by 16 bytes: the 16 rows of a fragment
// fall in distinct bank groups).
constexpr int kFusedRows = 128;
constexpr int kFusedHid = 512;               // hidden width the kernel is built for
constexpr int kFusedOut = 256;               // padded output width
constexpr int kFusedK0 = 32;                 // K of the first layer (the inputs, padded)
constexpr int kFusedPitch = kFusedHid + 8;   // bf16 pitch of the activation tile
constexpr int kFusedInPitch = kFusedK0 + 8;  // bf16 pitch of the input tile
constexpr int kFusedLds = kFusedRows * (kFusedPitch + kFusedInPitch) * 2;  // bytes

// Weights in fragment order, per matrix [out x in] with the first `kuse` columns multiplied:
// element (n, k) at ((n / 16) (kuse / 32) + k / 32) 512 + lane 8 + k % 8, lane = 16 (k / 8 % 4)
// + n % 16. The fragment a lane loads for a 16-row block and a K step is its 8 elements there,
// and the wave's 64 lanes read 1 KB of consecutive bytes.
struct FragArgs {
    const bf16 *wb;  // [out x in] copies (head stride sWb), matrix m at wt[m]
    long long sWb;
    bf16 *wf;  // fragment-order copies (head stride sWf), matrix m at fo[m]
    long long sWf;
    long long wt[kMaxMats], fo[kMaxMats];
    int out[kMaxMats], in[kMaxMats], kuse[kMaxMats];
};

// One thread per 16-byte fragment piece; blockIdx.y = matrix, blockIdx.z = head.
__global__ __launch_bounds__(256) void pack_fragments(const FragArgs f) {
    const int m = static_cast<int>(blockIdx.y), h = static_cast<int>(blockIdx.z);
    const long long pieces = static_cast<long long>(f.out[m]) * (f.kuse[m] / 8);
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= pieces) return;
    const int lane = static_cast<int>(c & 63);
    const long long blk = c >> 6;  // (n / 16) (kuse / 32) + k / 32
    const int kts = f.kuse[m] / kBK;
    const int n = static_cast<int>(blk / kts) * 16 + (lane & 15);
    const int k = static_cast<int>(blk % kts) * kBK + (lane >> 4) * 8;
    *reinterpret_cast<uint4 *>(f.wf + h * f.sWf + f.fo[m] + c * 8) =
        *reinterpret_cast<const uint4 *>(f.wb + h * f.sWb + f.wt[m] + static_cast<long long>(n) * f.in[m] + k);
}

struct FusedArgs {
    const bf16 *X;  // network inputs [rows x in_pad] (the first kFusedK0 columns are used)
    int in_pad;
    const bf16 *W;  // fragment-order weights of head 0 (pack_fragments); head h at W + h sW,
    long long sW;   // layer l at wt[l]
    const float *bias;  // biases of head 0 (head stride sBias), layer l's at bo[l]
    long long sBias;
    long long wt[kMaxMats], bo[kMaxMats];
    int depth, heads;
    float *Yp;  // [heads x rows x kFusedOut] fp32 outputs, head stride sYp
    long long sYp;
};

// One layer's products for the wave's columns n0 .. n0 + 16 JF - 1 and one half of the rows
// (64): acc = bias + A . W^T with A [64 x 32 NK] in LDS (row pitch PITCH) and W [N x 32 NK] in
// global memory in fragment order. The K loop is unrolled into groups of one activation fragment
// (16 rows x 32 K) and its JF MFMAs, issued in this order: the fragment for three groups ahead
// is read from LDS and, at the first group of a K step, the weights of the next step are loaded,
// then the group's MFMAs, then a share of `pend`'s UNITS units of other work (the other half's
// epilogue), spread evenly over the groups so that its VALU work issues while the MFMAs run. The
// schedule barrier after each group keeps the compiler from sinking the loads next to their use,
// where their latency would stall the wave, or from bunching the epilogue work.
template <int JF, int NK, int PITCH, int UNITS, class Pend>
__device__ __forceinline__ void fused_half(const bf16 *As, const bf16 *W, int n0, const float *bias,
                                           floatx4 (&acc)[4][JF], int lane, const Pend &pend) {
    constexpr int kAhead = 3, kRing = 4, kGroups = 4 * NK;
    const int r = lane & 15, kq = lane >> 4, g4 = kq * 4;
#pragma unroll
    for (int j = 0; j < JF; ++j) {
        const floatx4 init = *reinterpret_cast<const floatx4 *>(bias + n0 + 16 * j + g4);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][j] = init;
    }
    const bf16 *wl[JF];
#pragma unroll
    for (int j = 0; j < JF; ++j) wl[j] = W + (static_cast<long long>((n0 / 16 + j) * NK) * 64 + lane) * 8;
    const bf16 *al = As + r * PITCH + kq * 8;
    auto a_frag = [&](int g) {
        return *reinterpret_cast<const bf16x8 *>(al + 16 * (g & 3) * PITCH + (g >> 2) * kBK);
    };
    bf16x8 b[2][JF], fa[kRing];
#pragma unroll
    for (int j = 0; j < JF; ++j) b[0][j] = *reinterpret_cast<const bf16x8 *>(wl[j]);
#pragma unroll
    for (int g = 0; g < kAhead && g < kGroups; ++g) fa[g % kRing] = a_frag(g);
    static_assert(UNITS == 0 || UNITS % kGroups == 0 || kGroups % UNITS == 0, "units spread evenly");
#pragma clang loop unroll(full)
    for (int g = 0; g < kGroups; ++g) {
        const int kt = g >> 2, i = g & 3;
        if (i == 0 && kt + 1 < NK) {
#pragma unroll
            for (int j = 0; j < JF; ++j)
                b[(kt + 1) & 1][j] = *reinterpret_cast<const bf16x8 *>(wl[j] + (kt + 1) * 512);
        }
        if (g + kAhead < kGroups) fa[(g + kAhead) % kRing] = a_frag(g + kAhead);
#pragma unroll
        for (int j = 0; j < JF; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[kt & 1][j], fa[g % kRing], acc[i][j], 0, 0, 0);
        if constexpr (UNITS >= kGroups) {
#pragma unroll
            for (int u = 0; u < UNITS / kGroups; ++u) pend(g * (UNITS / kGroups) + u);
        } else if constexpr (UNITS > 0) {
            if (g % (kGroups / UNITS) == kGroups / UNITS - 1) pend(g / (kGroups / UNITS));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Epilogue units for fused_half's `pend`. GeluPending: unit v of 64 is the GELU of value v % 4 of
// the lane's four in 16 x 16 tile t = v / 4 of a hidden layer's half (i = t / 4, j = t % 4); at
// the fourth the tile's bf16 values are stored into the activation tile at `dst` (the lane's row
// and first column of that half), so that a group's MFMAs share the issue slots with one GELU.
// OutputPending: unit u of 8 is the fp32 store of one output tile.
struct NoPending {
    __device__ void operator()(int) const {}
};
struct GeluPending {
    const floatx4 (&acc)[4][4];
    bf16 *dst;
    mutable float done[4];
    __device__ __forceinline__ void operator()(int v) const {
        const int t = v >> 2, q = v & 3;
        done[q] = gelu(acc[t >> 2][t & 3][q]);
        if (q == 3)
            *reinterpret_cast<uint2 *>(dst + 16 * (t >> 2) * kFusedPitch + 16 * (t & 3)) =
                uint2{pack_bf16x2(done[0], done[1]), pack_bf16x2(done[2], done[3])};
    }
};
struct OutputPending {
    const floatx4 (&acc)[4][2];
    float *dst;
    __device__ __forceinline__ void operator()(int u) const {
        *reinterpret_cast<floatx4 *>(dst + 16 * (u >> 1) * kFusedOut + 16 * (u & 1)) = acc[u >> 1][u & 1];
    }
};

// Each layer runs in two phases, one per half of the rows, and in each phase the wave's MFMAs for
// one half interleave with the VALU epilogue (GELU and bf16 stores) of the other half's previous
// phase. Phase A of layer l multiplies rows 0-63 of layer l - 1 and writes GELU(rows 64-127 of
// layer l - 1); phase B multiplies rows 64-127 and writes GELU(rows 0-63 of layer l). A barrier
// follows every phase: the rows a phase writes are the ones the phase before read, and the rows
// a phase reads are the ones the phase before wrote. Per output value the products, their K
// order, the bias seeding and the GELU are those of the layer-by-layer kernels.
__global__ __launch_bounds__(512, 1) void infer_fused(const FusedArgs a) {
    extern __shared__ __attribute__((aligned(16))) bf16 smem[];
    bf16 *Xs = smem;                                 // activations [128 x kFusedPitch]
    bf16 *Ins = smem + kFusedRows * kFusedPitch;     // inputs [128 x kFusedInPitch]
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g4 = (lane >> 4) * 4;
    const long long row0 = static_cast<long long>(blockIdx.x) * kFusedRows;
    {
        const int rr = tid >> 2, c = (tid & 3) * 8;  // 512 threads: 128 rows x 4 pieces of 16 bytes
        *reinterpret_cast<uint4 *>(Ins + rr * kFusedInPitch + c) =
            *reinterpret_cast<const uint4 *>(a.X + (row0 + rr) * a.in_pad + c);
    }
    __syncthreads();
    const int nh = wave * (kFusedHid / 8), no = wave * (kFusedOut / 8);
    // the lane's row and first column in each half of the activation tile
    bf16 *const x_lo = Xs + r * kFusedPitch + nh + g4, *const x_hi = x_lo + 64 * kFusedPitch;
    constexpr int kHalf = 64 * kFusedPitch, kInHalf = 64 * kFusedInPitch;
    floatx4 lo[4][4], hi[4][4];  // a hidden layer's rows 0-63 and 64-127
    for (int h = 0; h < a.heads; ++h) {
        const bf16 *W = a.W + h * a.sW;
        const float *B = a.bias + h * a.sBias;
        fused_half<4, kFusedK0 / kBK, kFusedInPitch, 0>(Ins, W + a.wt[0], nh, B + a.bo[0], lo, lane, NoPending{});
        __syncthreads();
        fused_half<4, kFusedK0 / kBK, kFusedInPitch, 64>(Ins + kInHalf, W + a.wt[0], nh, B + a.bo[0], hi, lane,
                                                          GeluPending{lo, x_lo, {}});
        __syncthreads();
        for (int l = 1; l < a.depth; ++l) {
            fused_half<4, kFusedHid / kBK, kFusedPitch, 64>(Xs, W + a.wt[l], nh, B + a.bo[l], lo, lane,
                                                             GeluPending{hi, x_hi, {}});
            __syncthreads();
            fused_half<4, kFusedHid / kBK, kFusedPitch, 64>(Xs + kHalf, W + a.wt[l], nh, B + a.bo[l], hi, lane,
                                                             GeluPending{lo, x_lo, {}});
            __syncthreads();
        }
        floatx4 olo[4][2], ohi[4][2];
        float *Y = a.Yp + h * a.sYp + (row0 + r) * kFusedOut + no + g4;  // the lane's row and first column
        fused_half<2, kFusedHid / kBK, kFusedPitch, 64>(Xs, W + a.wt[a.depth], no, B + a.bo[a.depth], olo, lane,
                                                         GeluPending{hi, x_hi, {}});
        __syncthreads();
        fused_half<2, kFusedHid / kBK, kFusedPitch, 8>(Xs + kHalf, W + a.wt[a.depth], no, B + a.bo[a.depth], ohi,
                                                        lane, OutputPending{olo, Y});
#pragma unroll
        for (int u = 0; u < 8; ++u) OutputPending{ohi, Y + 64 * kFusedOut}(u);
        __syncthreads();
    }
}

// ---- surrogate answers: stacks in, gated R_d / T_t out

constexpr int kSlots = 5;                         // layer slots of the encoding
constexpr int kRawPerStack = 1 + kSlots * 5;      // layer count, then (mua, mus, g, n, d) per slot
constexpr int kAnswerInputs = kSlots * 5 + kSlots;  // 25 encoded parameters + 5 slot masks

struct EncodeArgs {
    const double *raw;  // [Bq x kRawPerStack]
    int rows, Bq, in_pad;
    int min_layers, max_layers;
    double center[5], half[5];  // encoding: (v - center) / half
    double lo[5], hi[5];        // the domain box (log10 for mua, mus, d)
    bf16 *X;                    // [Bq x in_pad]: the first 32 columns are written
    uint8_t *dom;               // [Bq]
};

// One thread per query row: the 30 network inputs (in double, then float, then bf16, as the
// host encoder and load_inputs do) and the domain flag. Rows past `rows` get zero inputs.
__global__ __launch_bounds__(256) void encode_stacks(const EncodeArgs a) {
    const int m = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m >= a.Bq) return;
    float f[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) f[j] = 0.0f;
    bool ok = false;
    if (m < a.rows) {
        const double *r = a.raw + static_cast<long long>(m) * kRawPerStack;
        const int nl = static_cast<int>(r[0]);
        ok = nl >= a.min_layers && nl <= a.max_layers;
        for (int i = 0; i < kSlots; ++i) {
            if (i >= nl) continue;
            const double *q = r + 1 + 5 * i;
            const double v[5] = {q[0] > 0.0 ? log10(q[0]) : -99.0, q[1] > 0.0 ? log10(q[1]) : -99.0,
                                 q[2], q[3], log10(q[4])};
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                f[5 * i + j] = static_cast<float>((v[j] - a.center[j]) / a.half[j]);
                ok = ok && v[j] >= a.lo[j] - 1e-9 && v[j] <= a.hi[j] + 1e-9;
            }
            f[25 + i] = 1.0f;
        }
    }
    uint4 *dst = reinterpret_cast<uint4 *>(a.X + static_cast<long long>(m) * a.in_pad);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = 8 * c + 2 * k;
            w[k] = static_cast<uint32_t>(to_bf16(f[j])) | (static_cast<uint32_t>(to_bf16(f[j + 1])) << 16);
        }
        dst[c] = uint4{w[0], w[1], w[2], w[3]};
    }
    a.dom[m] = ok ? 1 : 0;
}

struct FinishArgs {
    const float *mean;    // [rows x outputs] ensemble means (standardized)
    const float *spread;  // [rows x 2] ensemble standard deviations of outputs 0, 1
    const uint8_t *dom;
    const float *y_mean, *y_std;  // [outputs]
    int rows, outputs;
    double floor_rd, floor_tt, tau_rd, tau_tt;
    double *res;     // [rows x 4]: rd, tt, rd_spread, tt_spread
    uint8_t *route;  // [rows]
};

// One thread per row: de-standardize R_d and T_t (float multiply, then add: no contraction to an
// FMA, whose single rounding would differ from the host's), undo the log10 (the value is
// max(10^y - floor, 0); its spread is the ensemble's spread of y times d value / dy =
// ln 10 (value + floor)), and route the row to Monte Carlo if it is out of domain or a relative
// spread exceeds its threshold (R_d's relative to max(R_d, 1e-3) tau_rd, T_t's relative to
// max(T_t, 1e-2) tau_tt).
__global__ __launch_bounds__(256) void finish_answers(const FinishArgs a) {
#pragma clang fp contract(off)
    const int m = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m >= a.rows) return;
    double v[2], sp[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const float y = a.mean[static_cast<long long>(m) * a.outputs + q] * a.y_std[q] + a.y_mean[q];
        const double fl = q == 0 ? a.floor_rd : a.floor_tt;
        v[q] = fmax(exp10(static_cast<double>(y)) - fl, 0.0);
        sp[q] = (v[q] + fl) * 2.302585092994046 * static_cast<double>(a.spread[2 * m + q]) *
                static_cast<double>(a.y_std[q]);
    }
    const double rel_rd = sp[0] / fmax(fabs(v[0]), 1e-3);
    const double rel_tt = sp[1] / fmax(fabs(v[1]), 1e-2);
    double *o = a.res + 4ll * m;
    o[0] = v[0];
    o[1] = v[1];
    o[2] = sp[0];
    o[3] = sp[1];
    a.route[m] = (!a.dom[m] || rel_rd > a.tau_rd || rel_tt > a.tau_tt) ? 1 : 0;
}

// The other outputs' values, max(10^y - floor, 0), as float: one thread per (row, output).
__global__ __launch_bounds__(256) void finish_profiles(const float *mean, const float *y_mean, const float *y_std,
                                                       int rows, int outputs, double floor, float *prof) {
#pragma clang fp contract(off)
    const int width = outputs - 2;
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(rows) * width) return;
    const long long m = t / width;
    const int c = static_cast<int>(t % width) + 2;
    const float y = mean[m * outputs + c] * y_std[c] + y_mean[c];
    prof[t] = static_cast<float>(fmax(exp10(static_cast<double>(y)) - floor, 0.0));
}

}  // namespace mlp

#define MLP_TRY(call)                                                                          \
    do {                                                                                       \
        const hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                                  \
            throw EngineError(TMC_ERR_RUNTIME, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

using mlp::bf16;

// A few host threads that split a loop between them, the calling thread taking a share: the
// packing of a chunk of stacks, which reads the caller's stack array (memory-bound, a few hundred
// bytes a stack) faster from several cores than from one.
class HostTeam {
  public:
    explicit HostTeam(int n) : size_(std::max(n, 1)) {
        for (int t = 1; t < size_; ++t) workers_.emplace_back([this, t] { serve(t); });
    }
    ~HostTeam() {
        {
            std::lock_guard<std::mutex> g(m_);
            quit_ = true;
            ++gen_;
        }
        go_.notify_all();
        for (auto &w : workers_) w.join();
    }
    // f(begin, end) over [0, rows) in size_ contiguous pieces; returns when all are done
    void run(int64_t rows, const std::function<void(int64_t, int64_t)> &f) {
        job_ = [&](int t) { f(rows * t / size_, rows * (t + 1) / size_); };
        {
            std::lock_guard<std::mutex> g(m_);
            pending_ = size_ - 1;
            ++gen_;
        }
        go_.notify_all();
        job_(0);
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return pending_ == 0; });
    }

  private:
    void serve(int t) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                go_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (quit_) return;
            }
            job_(t);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    int size_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable go_, done_;
    std::function<void(int)> job_;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool quit_ = false;
};

class MlpModel {
  public:
    static int pad128(int x) { return (x + 127) / 128 * 128; }

    explicit MlpModel(const tmc_mlp_config &c) : cfg_(c) {
        if (c.heads < 1 || c.heads > 64) throw EngineError(TMC_ERR_CONFIG, "heads must be in [1, 64]");
        if (c.depth < 1 || c.depth > mlp::kMaxMats - 1)
            throw EngineError(TMC_ERR_CONFIG, "depth must be in [1, 7]");
        if (c.hidden < 128 || c.hidden % 128) throw EngineError(TMC_ERR_CONFIG, "hidden must be a multiple of 128");
        if (c.batch < 128 || c.batch % 128) throw EngineError(TMC_ERR_CONFIG, "batch must be a multiple of 128");
        if (c.inputs < 1 || c.outputs < 1) throw EngineError(TMC_ERR_CONFIG, "inputs and outputs must be >= 1");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            throw EngineError(TMC_ERR_DEVICE, "no usable HIP device found");
        if (c.device < 0 || c.device >= count) throw EngineError(TMC_ERR_CONFIG, "no such device");
        MLP_TRY(hipSetDevice(c.device));
        // the GEMM's LDS (staging tiles, then the epilogue's output tiles) exceeds the 64 KB that a
        // launch gets without asking
        set_lds<mlp::kFwdHidden>();
        set_lds<mlp::kFwdOut>();
        set_lds<mlp::kBwdData>();
        set_lds<mlp::kInferHidden>();
        MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp::gemm_tn_group),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, mlp::kTnLds));
        MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp::gemm_infer_wide<mlp::kInferHidden>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, mlp::kWideLds));
        MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp::gemm_infer_wide<mlp::kFwdOut>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, mlp::kWideLds));
        H_ = c.heads;
        in_pad_ = pad128(c.inputs);
        k0_ = (c.inputs + mlp::kBK - 1) / mlp::kBK * mlp::kBK;
        out_pad_ = pad128(c.outputs);
        hid_ = c.hidden;
        B_ = c.batch;
        // The fused inference kernel (one workgroup of 128 rows per CU, activations in LDS) serves
        // networks of hidden width 512 with at most 32 inputs and 129..256 outputs; a chunk of
        // queries is one workgroup per CU. TMC_MLP_INFER=layered picks the layer-by-layer kernels.
        {
            const char *pick = std::getenv("TMC_MLP_INFER");
            const bool layered = pick && std::strcmp(pick, "layered") == 0;
            fused_ = !layered && hid_ == mlp::kFusedHid && k0_ == mlp::kFusedK0 && out_pad_ == mlp::kFusedOut;
        }
        int cus = 0;
        MLP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        Bq_ = std::max(B_, fused_ ? std::max(cus, 1) * mlp::kFusedRows : 16384);
        if (fused_)
            MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp::infer_fused),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, mlp::kFusedLds));
        // parameter layout per head: W [out x in] then b [out], per matrix
        long long off = 0, boff = 0;
        nb_ = 0;
        for (int k = 0; k <= c.depth; ++k) {
            mlp::MatInfo mi;
            mi.bb = nb_;
            mi.in = k == 0 ? in_pad_ : hid_;
            mi.out = k == c.depth ? out_pad_ : hid_;
            mi.off = off;
            off += static_cast<long long>(mi.out) * mi.in;
            mi.boff = off;
            off += mi.out;
            mi.wt_off = boff;
            boff += static_cast<long long>(mi.out) * mi.in;
            mi.w_off = boff;
            boff += static_cast<long long>(mi.out) * mi.in;
            nb_ += mi.out;
            mats_.push_back(mi);
            real_in_.push_back(k == 0 ? c.inputs : hid_);
            real_out_.push_back(k == c.depth ? c.outputs : hid_);
        }
        P_ = off;
        Pb_ = boff;
        MLP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        MLP_TRY(hipEventCreate(&e0_));
        MLP_TRY(hipEventCreate(&e1_));
        p_ = alloc<float>(H_ * P_);
        g_ = alloc<float>(H_ * P_);
        // weight gradients split the batch (their K) into pieces of 1024: measured at batch 4096,
        // 1024 beats 256, 512, 2048 and no split (0.447 ms a step against 0.498, 0.458, 0.492,
        // 0.601): fewer partials to reduce, and still 256 workgroups for a hidden layer's product.
        // Each split is a whole number of GEMM stages; Adam sums at most kMaxSplits partials.
        splits_ = std::clamp(B_ / 1024, 1, mlp::kMaxSplits);
        while (B_ % (mlp::kBK * splits_) != 0) --splits_;
        part_ = alloc<float>(static_cast<long long>(H_) * splits_ * P_);
        bpart_ = alloc<float>(static_cast<long long>(H_) * (B_ / 64) * nb_);
        m_ = alloc<float>(H_ * P_);
        v_ = alloc<float>(H_ * P_);
        wb_ = alloc<bf16>(H_ * Pb_);
        step_ = alloc<uint64_t>(1);
        lr_ = alloc<float>(1);
        lossw_ = alloc<float>(out_pad_);
        loss_ = alloc<float>(kLossRing * H_);
        lpart_ = alloc<float>(static_cast<long long>(H_) * (B_ / 64) * (out_pad_ / 64));
        const long long BH = static_cast<long long>(B_) * hid_;
        X_ = alloc<bf16>(H_ * B_ * in_pad_);
        Yt_ = alloc<float>(H_ * B_ * out_pad_);
        Yp_ = alloc<float>(H_ * static_cast<long long>(Bq_) * out_pad_);
        dY_ = alloc<bf16>(H_ * B_ * out_pad_);
        for (int l = 0; l < c.depth; ++l) {
            G_.push_back(alloc<bf16>(H_ * BH));
            Hs_.push_back(alloc<bf16>(H_ * BH));
        }
        for (int l = 0; l < c.depth; ++l) dZ_.push_back(alloc<bf16>(H_ * BH));
        const long long QH = static_cast<long long>(Bq_) * hid_;
        Xq_ = alloc<bf16>(H_ * static_cast<long long>(Bq_) * in_pad_);
        Hq_[0] = alloc<bf16>(H_ * QH);
        Hq_[1] = alloc<bf16>(H_ * QH);
        xq_ = alloc<float>(static_cast<long long>(Bq_) * c.inputs);
        if (fused_) {
            for (size_t m = 0; m < mats_.size(); ++m) {
                frag_off_.push_back(Pf_);
                Pf_ += static_cast<long long>(mats_[m].out) * (m == 0 ? k0_ : mats_[m].in);
            }
            wf_ = alloc<bf16>(H_ * Pf_);
        }
        mean_ = alloc<float>(static_cast<long long>(Bq_) * c.outputs);
        spread_ = alloc<float>(static_cast<long long>(Bq_) * 2);
        init_params();
        std::vector<float> w(out_pad_, 0.0f);
        for (int j = 0; j < c.outputs; ++j) w[j] = 1.0f;
        MLP_TRY(hipMemcpy(lossw_, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        MLP_TRY(hipMemset(step_, 0, 8));
        set_lr(cfg_.lr);
        MLP_TRY(hipMemset(loss_, 0, static_cast<size_t>(kLossRing) * H_ * 4));
    }

    ~MlpModel() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (e0_) (void)hipEventDestroy(e0_);
        if (e1_) (void)hipEventDestroy(e1_);
        for (int i = 0; i < 2; ++i) {
            if (ans_k0_[i]) (void)hipEventDestroy(ans_k0_[i]);
            if (ans_k1_[i]) (void)hipEventDestroy(ans_k1_[i]);
            if (ans_done_[i]) (void)hipEventDestroy(ans_done_[i]);
            if (ans_h2d_[i]) (void)hipEventDestroy(ans_h2d_[i]);
            if (ans_used_[i]) (void)hipEventDestroy(ans_used_[i]);
        }
        if (ans_copy_) (void)hipStreamDestroy(ans_copy_);
        if (stream_) (void)hipStreamDestroy(stream_);
        for (void *q : allocs_) (void)hipFree(q);
        for (void *q : pinned_) (void)hipHostFree(q);
    }

    long long logical_params() const {
        long long n = 0;
        for (size_t k = 0; k < mats_.size(); ++k)
            n += static_cast<long long>(real_out_[k]) * real_in_[k] + real_out_[k];
        return n;
    }

    void set_data(const float *x, const float *y, int64_t n, const float *loss_w) {
        if (n < 1) throw EngineError(TMC_ERR_CONFIG, "the training set is empty");
        if (xd_) (void)hipFree(xd_);
        if (yd_) (void)hipFree(yd_);
        xd_ = yd_ = nullptr;
        MLP_TRY(hipMalloc(&xd_, static_cast<size_t>(n) * cfg_.inputs * 4));
        MLP_TRY(hipMalloc(&yd_, static_cast<size_t>(n) * cfg_.outputs * 4));
        MLP_TRY(hipMemcpy(xd_, x, static_cast<size_t>(n) * cfg_.inputs * 4, hipMemcpyHostToDevice));
        MLP_TRY(hipMemcpy(yd_, y, static_cast<size_t>(n) * cfg_.outputs * 4, hipMemcpyHostToDevice));
        n_data_ = n;
        if (loss_w) {
            std::vector<float> w(out_pad_, 0.0f);
            for (int j = 0; j < cfg_.outputs; ++j) w[j] = loss_w[j];
            MLP_TRY(hipMemcpy(lossw_, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        }
        if (graph_exec_) {  // the data pointers are baked into the captured step
            (void)hipGraphExecDestroy(graph_exec_);
            (void)hipGraphDestroy(graph_);
            graph_exec_ = nullptr;
            graph_ = nullptr;
        }
    }

    // Run `steps` optimisation steps (one captured graph per step, replayed). Returns seconds.
    double train(int64_t steps, float *losses) {
        if (!xd_) throw EngineError(TMC_ERR_CONFIG, "no training data (tmc_mlp_set_data)");
        if (steps < 0 || steps > kLossRing) throw EngineError(TMC_ERR_CONFIG, "steps must be in [0, 65536]");
        uint64_t step0 = 0;
        MLP_TRY(hipMemcpy(&step0, step_, 8, hipMemcpyDeviceToHost));
        if (!graph_exec_) {
            MLP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            enqueue_step(nullptr, true);
            MLP_TRY(hipStreamEndCapture(stream_, &graph_));
            MLP_TRY(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        }
        MLP_TRY(hipEventRecord(e0_, stream_));
        if (steps > 0) frag_stale_ = true;
        for (int64_t s = 0; s < steps; ++s) MLP_TRY(hipGraphLaunch(graph_exec_, stream_));
        MLP_TRY(hipEventRecord(e1_, stream_));
        MLP_TRY(hipEventSynchronize(e1_));
        float ms = 0.0f;
        MLP_TRY(hipEventElapsedTime(&ms, e0_, e1_));
        if (losses && steps > 0) {
            std::vector<float> ring(static_cast<size_t>(kLossRing) * H_);
            MLP_TRY(hipMemcpy(ring.data(), loss_, ring.size() * 4, hipMemcpyDeviceToHost));
            for (int64_t s = 0; s < steps; ++s) {
                const int64_t slot = static_cast<int64_t>((step0 + s) % kLossRing);
                for (int h = 0; h < H_; ++h) losses[s * H_ + h] = ring[slot * H_ + h];
            }
        }
        return ms / 1000.0;
    }

    // One forward/backward on the given dataset rows (no update): gradients [heads x logical].
    void gradients(const int64_t *rows, float *grads, float *loss) {
        if (!xd_) throw EngineError(TMC_ERR_CONFIG, "no training data (tmc_mlp_set_data)");
        for (int i = 0; i < B_; ++i)
            if (rows[i] < 0 || rows[i] >= n_data_)
                throw EngineError(TMC_ERR_CONFIG, "row " + std::to_string(rows[i]) + " is outside the training set");
        int64_t *drows = nullptr;
        MLP_TRY(hipMalloc(&drows, static_cast<size_t>(B_) * 8));
        MLP_TRY(hipMemcpy(drows, rows, static_cast<size_t>(B_) * 8, hipMemcpyHostToDevice));
        uint64_t step0 = 0;
        MLP_TRY(hipMemcpy(&step0, step_, 8, hipMemcpyDeviceToHost));
        const int64_t slot = static_cast<int64_t>(step0 % kLossRing);
        enqueue_step(drows, false);
        MLP_TRY(hipStreamSynchronize(stream_));
        (void)hipFree(drows);
        std::vector<float> g(static_cast<size_t>(H_) * P_);
        MLP_TRY(hipMemcpy(g.data(), g_, g.size() * 4, hipMemcpyDeviceToHost));
        unpad(g.data(), grads);
        if (loss) {
            std::vector<float> l(H_);
            MLP_TRY(hipMemcpy(l.data(), loss_ + slot * H_, l.size() * 4, hipMemcpyDeviceToHost));
            for (int h = 0; h < H_; ++h) loss[h] = l[h];
        }
    }

    // Ensemble prediction: mean of every output and the spread of the first two. Returns the
    // device time (upload excluded, reductions included) in seconds.
    double predict(const float *x, int64_t n, float *mean, float *spread, float *per_head) {
        double total = 0.0;
        for (int64_t r0 = 0; r0 < n; r0 += Bq_) {
            const int rows = static_cast<int>(std::min<int64_t>(Bq_, n - r0));
            MLP_TRY(hipMemcpyAsync(xq_, x + r0 * cfg_.inputs, static_cast<size_t>(rows) * cfg_.inputs * 4,
                                   hipMemcpyHostToDevice, stream_));
            MLP_TRY(hipEventRecord(e0_, stream_));
            hipLaunchKernelGGL(mlp::load_inputs, dim3((Bq_ * (k0_ / 8) + 255) / 256), dim3(256), 0,
                               stream_, xq_, rows, cfg_.inputs, k0_, in_pad_, Bq_, Xq_);
            infer(rows);
            hipLaunchKernelGGL(mlp::ensemble_reduce, dim3((rows + 3) / 4), dim3(256), 0, stream_, Yp_, H_, Bq_,
                               out_pad_, cfg_.outputs, 2, rows, mean_, spread_);
            MLP_TRY(hipGetLastError());
            MLP_TRY(hipEventRecord(e1_, stream_));
            if (mean)
                MLP_TRY(hipMemcpyAsync(mean + r0 * cfg_.outputs, mean_, static_cast<size_t>(rows) * cfg_.outputs * 4,
                                       hipMemcpyDeviceToHost, stream_));
            if (spread)
                MLP_TRY(hipMemcpyAsync(spread + r0 * 2, spread_, static_cast<size_t>(rows) * 2 * 4,
                                       hipMemcpyDeviceToHost, stream_));
            if (per_head)
                for (int h = 0; h < H_; ++h)
                    MLP_TRY(hipMemcpy2DAsync(per_head + (static_cast<int64_t>(h) * n + r0) * cfg_.outputs,
                                             cfg_.outputs * 4, Yp_ + static_cast<long long>(h) * Bq_ * out_pad_,
                                             out_pad_ * 4, cfg_.outputs * 4, rows, hipMemcpyDeviceToHost, stream_));
            MLP_TRY(hipStreamSynchronize(stream_));
            float ms = 0.0f;
            MLP_TRY(hipEventElapsedTime(&ms, e0_, e1_));
            total += ms / 1000.0;
        }
        return total;
    }

    // Gated surrogate answers for stacks, end to end (tmc_mlp_answer): chunks of Bq_ stacks,
    // the host packing chunk k + 1 into pinned memory (a few threads) while its copy to the
    // device runs on a stream of its own and the device encodes, runs and gates chunk k; two
    // pinned slots and two device input buffers alternate.
    void answer(const tmc_surrogate_gate &gate, const tmc_stack *stacks, int64_t n, tmc_surrogate_answers *out) {
        const auto t_start = std::chrono::steady_clock::now();
        if (cfg_.inputs != mlp::kAnswerInputs)
            throw EngineError(TMC_ERR_CONFIG, "surrogate answers need a network with 30 inputs");
        if (cfg_.outputs < 2) throw EngineError(TMC_ERR_CONFIG, "surrogate answers need >= 2 outputs");
        const tmc_stack_dist &d = gate.box;
        const double lo_lin[5] = {d.mua_lo, d.mus_lo, d.g_lo, d.n_lo, d.d_lo};
        const double hi_lin[5] = {d.mua_hi, d.mus_hi, d.g_hi, d.n_hi, d.d_hi};
        mlp::EncodeArgs ea{};
        for (int j = 0; j < 5; ++j) {
            const bool logp = j == 0 || j == 1 || j == 4;
            if (!(hi_lin[j] > lo_lin[j]) || (logp && !(lo_lin[j] > 0.0)))
                throw EngineError(TMC_ERR_CONFIG, "gate box: every range needs lo < hi (and lo > 0 for mua, mus, d)");
            ea.lo[j] = logp ? std::log10(lo_lin[j]) : lo_lin[j];
            ea.hi[j] = logp ? std::log10(hi_lin[j]) : hi_lin[j];
            ea.center[j] = (ea.lo[j] + ea.hi[j]) / 2.0;
            ea.half[j] = (ea.hi[j] - ea.lo[j]) / 2.0;
        }
        ea.min_layers = d.min_layers;
        ea.max_layers = d.max_layers;
        const bool want_prof = out->profiles != nullptr && cfg_.outputs > 2;
        answer_buffers(want_prof);
        MLP_TRY(hipMemcpyAsync(ans_ymean_, gate.y_mean, static_cast<size_t>(cfg_.outputs) * 4, hipMemcpyHostToDevice, stream_));
        MLP_TRY(hipMemcpyAsync(ans_ystd_, gate.y_std, static_cast<size_t>(cfg_.outputs) * 4, hipMemcpyHostToDevice, stream_));
        MLP_TRY(hipStreamSynchronize(stream_));

        const int width = cfg_.outputs - 2;
        const int64_t chunks = (n + Bq_ - 1) / Bq_;
        double device = 0.0;
        auto rows_of = [&](int64_t k) { return static_cast<int>(std::min<int64_t>(Bq_, n - k * Bq_)); };
        // results of chunk k (its slot's events have completed) to the caller's arrays
        auto drain = [&](int64_t k) {
            const int slot = static_cast<int>(k & 1), rows = rows_of(k);
            MLP_TRY(hipEventSynchronize(ans_done_[slot]));
            float ms = 0.0f;
            MLP_TRY(hipEventElapsedTime(&ms, ans_k0_[slot], ans_k1_[slot]));
            device += ms / 1000.0;
            const double *res = ans_pin_res_[slot];
            const uint8_t *flags = reinterpret_cast<const uint8_t *>(res + 4ll * Bq_);
            const int64_t r0 = k * Bq_;
            for (int m = 0; m < rows; ++m) {
                out->rd[r0 + m] = res[4 * m];
                out->tt[r0 + m] = res[4 * m + 1];
                if (out->rd_spread) out->rd_spread[r0 + m] = res[4 * m + 2];
                if (out->tt_spread) out->tt_spread[r0 + m] = res[4 * m + 3];
                out->route[r0 + m] = flags[m];
                if (out->in_domain) out->in_domain[r0 + m] = flags[Bq_ + m];
            }
            if (want_prof)
                std::memcpy(out->profiles + r0 * width, ans_pin_prof_[slot], static_cast<size_t>(rows) * width * 4);
        };
        // packing threads: only for more than one chunk's worth, at most 8 (the copy is memory-bound)
        const unsigned hw = std::thread::hardware_concurrency();
        std::unique_ptr<HostTeam> team;
        if (n > Bq_ / 2) team = std::make_unique<HostTeam>(static_cast<int>(std::clamp(hw, 1u, 8u)));
        for (int64_t k = 0; k < chunks; ++k) {
            const int slot = static_cast<int>(k & 1), rows = rows_of(k);
            if (k >= 2) drain(k - 2);
            // pack: layer count, then the five parameters of the first kSlots layers (0 if absent)
            double *raw = ans_pin_raw_[slot];
            const tmc_stack *first = stacks + k * Bq_;
            auto pack = [&](int64_t m0, int64_t m1) {
                for (int64_t m = m0; m < m1; ++m) {
                    const tmc_stack &st = first[m];
                    double *r = raw + m * mlp::kRawPerStack;
                    r[0] = static_cast<double>(st.num_layers);
                    for (int i = 0; i < mlp::kSlots; ++i) {
                        double *q = r + 1 + 5 * i;
                        if (i < st.num_layers && i < TMC_MAX_LAYERS) {
                            const tmc_layer &L = st.layers[i];
                            q[0] = L.mua; q[1] = L.mus; q[2] = L.g; q[3] = L.n; q[4] = L.d;
                        } else {
                            q[0] = q[1] = q[2] = q[3] = q[4] = 0.0;
                        }
                    }
                }
            };
            if (team) team->run(rows, pack);
            else pack(0, rows);
            // the copy waits until the encoding of chunk k - 2 has read this device buffer
            MLP_TRY(hipStreamWaitEvent(ans_copy_, ans_used_[slot], 0));
            MLP_TRY(hipMemcpyAsync(ans_raw_[slot], raw, static_cast<size_t>(rows) * mlp::kRawPerStack * 8,
                                   hipMemcpyHostToDevice, ans_copy_));
            MLP_TRY(hipEventRecord(ans_h2d_[slot], ans_copy_));
            MLP_TRY(hipStreamWaitEvent(stream_, ans_h2d_[slot], 0));
            MLP_TRY(hipEventRecord(ans_k0_[slot], stream_));
            ea.raw = ans_raw_[slot];
            ea.rows = rows;
            ea.Bq = Bq_;
            ea.in_pad = in_pad_;
            ea.X = Xq_;
            ea.dom = ans_dom_;
            hipLaunchKernelGGL(mlp::encode_stacks, dim3((Bq_ + 255) / 256), dim3(256), 0, stream_, ea);
            MLP_TRY(hipEventRecord(ans_used_[slot], stream_));
            infer(rows);
            hipLaunchKernelGGL(mlp::ensemble_reduce, dim3((rows + 3) / 4), dim3(256), 0, stream_, Yp_, H_, Bq_,
                               out_pad_, cfg_.outputs, 2, rows, mean_, spread_);
            mlp::FinishArgs fa{mean_, spread_, ans_dom_, ans_ymean_, ans_ystd_, rows, cfg_.outputs,
                               gate.floor_rd, gate.floor_tt, gate.tau_rd, gate.tau_tt, ans_res_,
                               ans_route_};
            hipLaunchKernelGGL(mlp::finish_answers, dim3((rows + 255) / 256), dim3(256), 0, stream_, fa);
            if (want_prof) {
                const long long cells = static_cast<long long>(rows) * width;
                hipLaunchKernelGGL(mlp::finish_profiles, dim3(static_cast<unsigned>((cells + 255) / 256)), dim3(256), 0,
                                   stream_, mean_, ans_ymean_, ans_ystd_, rows, cfg_.outputs, gate.floor, ans_prof_);
            }
            MLP_TRY(hipGetLastError());
            MLP_TRY(hipEventRecord(ans_k1_[slot], stream_));
            double *res = ans_pin_res_[slot];
            uint8_t *flags = reinterpret_cast<uint8_t *>(res + 4ll * Bq_);
            MLP_TRY(hipMemcpyAsync(res, ans_res_, static_cast<size_t>(rows) * 32, hipMemcpyDeviceToHost, stream_));
            MLP_TRY(hipMemcpyAsync(flags, ans_route_, static_cast<size_t>(rows), hipMemcpyDeviceToHost, stream_));
            MLP_TRY(hipMemcpyAsync(flags + Bq_, ans_dom_, static_cast<size_t>(rows), hipMemcpyDeviceToHost, stream_));
            if (want_prof)
                MLP_TRY(hipMemcpyAsync(ans_pin_prof_[slot], ans_prof_, static_cast<size_t>(rows) * width * 4,
                                       hipMemcpyDeviceToHost, stream_));
            MLP_TRY(hipEventRecord(ans_done_[slot], stream_));
        }
        for (int64_t k = std::max<int64_t>(0, chunks - 2); k < chunks; ++k) drain(k);
        out->device_seconds = device;
        out->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    }

    void get_params(float *out) {
        std::vector<float> p(static_cast<size_t>(H_) * P_);
        MLP_TRY(hipMemcpy(p.data(), p_, p.size() * 4, hipMemcpyDeviceToHost));
        unpad(p.data(), out);
    }

    void set_params(const float *in) {
        std::vector<float> p(static_cast<size_t>(H_) * P_, 0.0f);
        pad(in, p.data());
        MLP_TRY(hipMemcpy(p_, p.data(), p.size() * 4, hipMemcpyHostToDevice));
        refresh_bf16();
    }

    void set_lr(float lr) { MLP_TRY(hipMemcpy(lr_, &lr, 4, hipMemcpyHostToDevice)); }
    int heads() const { return H_; }
    // Algorithmic FLOPs, on the network's real widths (the kernels also multiply padding).
    uint64_t flops_per_step() const {
        uint64_t f = 0;
        for (size_t k = 0; k < mats_.size(); ++k)
            f += 2ull * B_ * static_cast<uint64_t>(real_out_[k]) * real_in_[k] * (k == 0 ? 2 : 3);  // fwd + dW (+ dX)
        return f * H_;
    }
    uint64_t flops_per_query() const {
        uint64_t f = 0;
        for (size_t k = 0; k < mats_.size(); ++k) f += 2ull * real_out_[k] * real_in_[k];
        return f * H_;
    }

  private:
    static constexpr int kLossRing = 65536;

    template <class T> T *alloc(long long n) {
        void *q = nullptr;
        MLP_TRY(hipMalloc(&q, static_cast<size_t>(std::max<long long>(n, 1)) * sizeof(T)));
        MLP_TRY(hipMemset(q, 0, static_cast<size_t>(std::max<long long>(n, 1)) * sizeof(T)));
        allocs_.push_back(q);
        return static_cast<T *>(q);
    }

    // Logical (unpadded) <-> padded parameter layouts; per head: W [out x in] then b [out].
    void unpad(const float *padded, float *logical) const {
        long long o = 0;
        for (int h = 0; h < H_; ++h)
            for (size_t k = 0; k < mats_.size(); ++k) {
                const auto &mi = mats_[k];
                const float *src = padded + h * P_;
                for (int r = 0; r < real_out_[k]; ++r)
                    for (int c = 0; c < real_in_[k]; ++c) logical[o++] = src[mi.off + static_cast<long long>(r) * mi.in + c];
                for (int r = 0; r < real_out_[k]; ++r) logical[o++] = src[mi.boff + r];
            }
    }
    void pad(const float *logical, float *padded) const {
        long long o = 0;
        for (int h = 0; h < H_; ++h)
            for (size_t k = 0; k < mats_.size(); ++k) {
                const auto &mi = mats_[k];
                float *dst = padded + h * P_;
                for (int r = 0; r < real_out_[k]; ++r)
                    for (int c = 0; c < real_in_[k]; ++c) dst[mi.off + static_cast<long long>(r) * mi.in + c] = logical[o++];
                for (int r = 0; r < real_out_[k]; ++r) dst[mi.boff + r] = logical[o++];
            }
    }

    // He-style initialisation from Philox (seed, head, element): uniform in +-sqrt(6 / fan_in).
    void init_params() {
        std::vector<float> p(static_cast<size_t>(H_) * P_, 0.0f);
        const uint32_t k0 = static_cast<uint32_t>(cfg_.seed), k1 = static_cast<uint32_t>(cfg_.seed >> 32);
        for (int h = 0; h < H_; ++h)
            for (size_t k = 0; k < mats_.size(); ++k) {
                const auto &mi = mats_[k];
                const float a = std::sqrt(6.0f / static_cast<float>(real_in_[k]));
                for (int r = 0; r < real_out_[k]; ++r)
                    for (int c = 0; c < real_in_[k]; ++c) {
                        const uint64_t e = static_cast<uint64_t>(r) * mi.in + c;
                        const Block4 v = philox4x32_10(static_cast<uint32_t>(e), static_cast<uint32_t>(k),
                                                       static_cast<uint32_t>(h), 0x494E4954u, k0, k1);
                        const float u = (static_cast<float>(v.v0 >> 8) + 0.5f) * 0x1p-24f;
                        p[static_cast<size_t>(h) * P_ + mi.off + e] = a * (2.0f * u - 1.0f);
                    }
            }
        MLP_TRY(hipMemcpy(p_, p.data(), p.size() * 4, hipMemcpyHostToDevice));
        refresh_bf16();
    }

    void refresh_bf16() {
        frag_stale_ = true;
        mlp::AdamArgs a;
        std::memset(&a, 0, sizeof a);
        a.p = p_;
        a.wb = wb_;
        a.P = P_;
        a.Pb = Pb_;
        a.n_mats = static_cast<int>(mats_.size());
        for (size_t k = 0; k < mats_.size(); ++k) a.mats[k] = mats_[k];
        hipLaunchKernelGGL(mlp::refresh_bf16, dim3(256, H_), dim3(256), 0, stream_, a);
        MLP_TRY(hipGetLastError());
        MLP_TRY(hipStreamSynchronize(stream_));
    }

    template <int kEpi> static void set_lds() {
        MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp::gemm_nt<kEpi>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    mlp::lds_bytes_of(kEpi)));
    }
    // The 128 x 256 inference kernel takes N a multiple of 256 (the others use gemm_nt).
    static bool wide_ok(const mlp::GemmArgs &g) {
        return g.N % 256 == 0 && g.M % 128 == 0 && g.K % mlp::kBK == 0;
    }
    template <int kEpi> void gemm_wide(const mlp::GemmArgs &g) {
        hipLaunchKernelGGL((mlp::gemm_infer_wide<kEpi>), dim3(g.N / 256, g.M / 128, H_), dim3(256),
                           mlp::kWideLds, stream_, g);
        MLP_TRY(hipGetLastError());
    }
    template <int kEpi> void gemm(mlp::GemmArgs g) {
        if (g.splits < 1) g.splits = 1;
        if (g.M % 128 || g.N % 128 || g.K % (mlp::kBK * g.splits))
            throw EngineError(TMC_ERR_RUNTIME, "gemm_nt shape is not a multiple of the tile");
        hipLaunchKernelGGL((mlp::gemm_nt<kEpi>), dim3(g.N / 128, g.M / 128, H_ * g.splits), dim3(256),
                           mlp::lds_bytes_of(kEpi), stream_, g);
        MLP_TRY(hipGetLastError());
    }

    // dW [out x in] of matrix mi = dZ^T . H with dZ [B x out] (head stride s_dz) and H, the
    // layer's input, [B x in] (head stride s_h), split along the batch into partials that
    // adam_params sums.
    // The weight gradient of matrix mi, dW = dZ^T H (K = the batch, split in splits_ pieces),
    // as one product of a gemm_tn_group launch.
    mlp::GemmArgs weight_grad_args(const mlp::MatInfo &mi, const bf16 *dz, long long s_dz,
                                   const bf16 *h, long long s_h) const {
        mlp::GemmArgs g = base_args();
        g.A = dz;
        g.lda = mi.out;
        g.sA = s_dz;
        g.B = h;
        g.ldb = mi.in;
        g.sB = s_h;
        g.M = mi.out;
        g.N = mi.in;
        g.K = B_;
        g.splits = splits_;
        g.Cf = part_ + mi.off;
        g.sCf = static_cast<long long>(splits_) * P_;
        g.sSplit = P_;
        if (g.M % 128 || g.N % 128 || g.K % (mlp::kBK * g.splits))
            throw EngineError(TMC_ERR_RUNTIME, "weight gradient shape is not a multiple of the tile");
        return g;
    }

    // Every matrix's weight gradient in one launch (products in the order given).
    void weight_grads(const std::vector<mlp::GemmArgs> &gs) {
        mlp::TnGroup grp;
        std::memset(&grp, 0, sizeof grp);
        if (gs.empty() || gs.size() > static_cast<size_t>(mlp::kMaxMats))
            throw EngineError(TMC_ERR_RUNTIME, "weight gradient group of an unsupported size");
        int total = 0;
        for (size_t m = 0; m < gs.size(); ++m) {
            grp.g[m] = gs[m];
            grp.first[m] = total;
            total += (gs[m].N / 128) * (gs[m].M / 128) * H_ * gs[m].splits;
        }
        grp.first[gs.size()] = total;
        grp.n = static_cast<int>(gs.size());
        hipLaunchKernelGGL(mlp::gemm_tn_group, dim3(static_cast<unsigned>(total)), dim3(256), mlp::kTnLds,
                           stream_, grp);
        MLP_TRY(hipGetLastError());
    }

    mlp::GemmArgs base_args() const {
        mlp::GemmArgs g;
        std::memset(&g, 0, sizeof g);
        return g;
    }

    // Every head's outputs for the first `rows` queries of Xq_ into Yp_ [head x Bq_ x out_pad_]:
    // the fused kernel where the network's shape allows it, else the layer-by-layer products
    // over the whole chunk.
    void infer(int rows) {
        if (!fused_) {
            forward(Xq_, Bq_, Hq_, nullptr, true);
            return;
        }
        if (frag_stale_) {
            mlp::FragArgs f;
            std::memset(&f, 0, sizeof f);
            f.wb = wb_;
            f.sWb = Pb_;
            f.wf = wf_;
            f.sWf = Pf_;
            long long most = 0;
            for (size_t m = 0; m < mats_.size(); ++m) {
                f.wt[m] = mats_[m].wt_off;
                f.fo[m] = frag_off_[m];
                f.out[m] = mats_[m].out;
                f.in[m] = mats_[m].in;
                f.kuse[m] = m == 0 ? k0_ : mats_[m].in;
                most = std::max(most, static_cast<long long>(f.out[m]) * (f.kuse[m] / 8));
            }
            hipLaunchKernelGGL(mlp::pack_fragments, dim3(static_cast<unsigned>((most + 255) / 256),
                               static_cast<unsigned>(mats_.size()), H_), dim3(256), 0, stream_, f);
            MLP_TRY(hipGetLastError());
            frag_stale_ = false;
        }
        mlp::FusedArgs a;
        std::memset(&a, 0, sizeof a);
        a.X = Xq_;
        a.in_pad = in_pad_;
        a.W = wf_;
        a.sW = Pf_;
        a.bias = p_;
        a.sBias = P_;
        for (int l = 0; l <= cfg_.depth; ++l) {
            a.wt[l] = frag_off_[l];
            a.bo[l] = mats_[l].boff;
        }
        a.depth = cfg_.depth;
        a.heads = H_;
        a.Yp = Yp_;
        a.sYp = static_cast<long long>(Bq_) * out_pad_;
        const int tiles = (rows + mlp::kFusedRows - 1) / mlp::kFusedRows;
        if (tiles * mlp::kFusedRows > Bq_) throw EngineError(TMC_ERR_RUNTIME, "fused inference: rows exceed the chunk");
        hipLaunchKernelGGL(mlp::infer_fused, dim3(tiles), dim3(512), mlp::kFusedLds, stream_, a);
        MLP_TRY(hipGetLastError());
    }

    // Forward through every layer: X [Bn x in_pad] per head -> Yp_. With `infer`, activations
    // alternate between two buffers and neither Z nor Ht is kept.
    void forward(bf16 *X, int Bn, bf16 *const *Hbuf, bf16 *const *Gbuf, bool infer) {
        const long long BH = static_cast<long long>(Bn) * hid_;
        const bf16 *in = X;
        int in_dim = in_pad_;
        for (int l = 0; l < cfg_.depth; ++l) {
            const auto &mi = mats_[l];
            mlp::GemmArgs g = base_args();
            g.A = in;
            g.lda = in_dim;
            // inference feeds every head the same queries
            g.sA = infer && l == 0 ? 0 : static_cast<long long>(Bn) * in_dim;
            g.B = wb_ + mi.wt_off;
            g.ldb = mi.in;
            g.sB = Pb_;
            g.M = Bn;
            g.N = hid_;
            // the first layer's inputs are zero past k0_ (padding): its products stop there
            g.K = l == 0 ? k0_ : in_dim;
            g.bias = p_ + mi.boff;
            g.sBias = P_;
            g.Gd = infer ? nullptr : Gbuf[l];
            g.Cb = infer ? Hbuf[l % 2] : Hbuf[l];
            g.sCb = BH;
            if (infer && wide_ok(g)) {
                gemm_wide<mlp::kInferHidden>(g);
            } else if (infer) {
                gemm<mlp::kInferHidden>(g);
            } else {
                gemm<mlp::kFwdHidden>(g);
            }
            in = g.Cb;
            in_dim = hid_;
        }
        const auto &mo = mats_[cfg_.depth];
        mlp::GemmArgs g = base_args();
        g.A = in;
        g.lda = hid_;
        g.sA = BH;
        g.B = wb_ + mo.wt_off;
        g.ldb = hid_;
        g.sB = Pb_;
        g.M = Bn;
        g.N = out_pad_;
        g.K = hid_;
        g.bias = p_ + mo.boff;
        g.sBias = P_;
        g.Cf = Yp_;
        g.sCf = static_cast<long long>(Bn) * out_pad_;
        if (infer && wide_ok(g))
            gemm_wide<mlp::kFwdOut>(g);
        else
            gemm<mlp::kFwdOut>(g);
    }

    // One training step (gather, forward, loss, backward, and Adam if `update`).
    void enqueue_step(const int64_t *fixed_rows, bool update) {
        const uint32_t k0 = static_cast<uint32_t>(cfg_.seed ^ 0x5DEECE66DULL);
        const uint32_t k1 = static_cast<uint32_t>((cfg_.seed >> 32) ^ 0xB5297A4DULL);
        hipLaunchKernelGGL(mlp::gather_batch, dim3(B_ / mlp::kGatherRows, H_), dim3(256), 0, stream_, xd_,
                           yd_, n_data_,
                           cfg_.inputs, cfg_.outputs, k0_, in_pad_, out_pad_, B_, step_, k0, k1, X_,
                           Yt_, fixed_rows);
        MLP_TRY(hipGetLastError());
        forward(X_, B_, Hs_.data(), G_.data(), false);
        const float inv_count = 1.0f / (static_cast<float>(B_) * cfg_.outputs);
        hipLaunchKernelGGL(mlp::loss_grad, dim3(B_ / 64, out_pad_ / 64, H_), dim3(256), 0, stream_, Yp_,
                           Yt_, lossw_, B_, out_pad_, inv_count, dY_, lpart_,
                           bpart_ + mats_[cfg_.depth].bb, static_cast<long long>(B_ / 64) * nb_, nb_);
        MLP_TRY(hipGetLastError());
        hipLaunchKernelGGL(mlp::loss_reduce, dim3(H_), dim3(256), 0, stream_, lpart_,
                           (B_ / 64) * (out_pad_ / 64), inv_count, loss_, step_, H_,
                           static_cast<uint64_t>(kLossRing));
        MLP_TRY(hipGetLastError());
        const long long BH = static_cast<long long>(B_) * hid_;
        const int D = cfg_.depth;
        // the backward data chain (dZ of every hidden layer kept for the weight gradients), then
        // every weight gradient in one launch
        std::vector<mlp::GemmArgs> wg;
        {
            const auto &mo = mats_[D];
            wg.push_back(weight_grad_args(mo, dY_, static_cast<long long>(B_) * out_pad_, Hs_[D - 1], BH));
            mlp::GemmArgs d = base_args();
            d.A = dY_;
            d.lda = out_pad_;
            d.sA = static_cast<long long>(B_) * out_pad_;
            d.B = wb_ + mo.w_off;
            d.ldb = out_pad_;
            d.sB = Pb_;
            d.M = B_;
            d.N = hid_;
            d.K = out_pad_;
            d.Gin = G_[D - 1];
            set_bias_partials(d, mats_[D - 1]);
            d.Cb = dZ_[D - 1];
            d.sCb = BH;
            gemm<mlp::kBwdData>(d);
        }
        for (int l = D - 1; l >= 0; --l) {
            const auto &mi = mats_[l];
            wg.push_back(weight_grad_args(mi, dZ_[l], BH, l == 0 ? X_ : Hs_[l - 1],
                                          l == 0 ? static_cast<long long>(B_) * in_pad_ : BH));
            if (l > 0) {
                mlp::GemmArgs d = base_args();
                d.A = dZ_[l];
                d.lda = hid_;
                d.sA = BH;
                d.B = wb_ + mi.w_off;
                d.ldb = hid_;
                d.sB = Pb_;
                d.M = B_;
                d.N = hid_;
                d.K = hid_;
                d.Gin = G_[l - 1];
                set_bias_partials(d, mats_[l - 1]);
                d.Cb = dZ_[l - 1];
                d.sCb = BH;
                gemm<mlp::kBwdData>(d);
            }
        }
        weight_grads(wg);
        // Adam, or (without `update`) the summed gradients into g_
        {
            mlp::AdamArgs a;
            std::memset(&a, 0, sizeof a);
            a.p = p_;
            a.g = g_;
            a.part = part_;
            a.splits = splits_;
            a.bpart = bpart_;
            a.bp_rows = B_ / 64;
            a.nb = nb_;
            a.grad_only = update ? 0 : 1;
            a.m = m_;
            a.v = v_;
            a.wb = wb_;
            a.P = P_;
            a.Pb = Pb_;
            a.n_mats = static_cast<int>(mats_.size());
            for (size_t k = 0; k < mats_.size(); ++k) a.mats[k] = mats_[k];
            a.lr = lr_;
            a.beta1 = cfg_.beta1;
            a.beta2 = cfg_.beta2;
            a.eps = cfg_.eps;
            a.wd = cfg_.weight_decay;
            a.step = step_;
            int tiles = 0;
            for (const auto &mi : mats_) tiles += (mi.out / mlp::kAdamTile) * (mi.in / mlp::kAdamTile);
            hipLaunchKernelGGL(mlp::adam_params, dim3(tiles + static_cast<int>(mats_.size()), H_),
                               dim3(256), 0, stream_, a);
            MLP_TRY(hipGetLastError());
        }
        if (update) {
            hipLaunchKernelGGL(mlp::bump_step, dim3(1), dim3(1), 0, stream_, step_);
            MLP_TRY(hipGetLastError());
        }
    }

    // The bias-gradient partials of layer mi, written by the backward data product that makes
    // its dZ.
    void set_bias_partials(mlp::GemmArgs &d, const mlp::MatInfo &mi) const {
        d.Bp = bpart_ + mi.bb;
        d.sBp = static_cast<long long>(B_ / 64) * nb_;
        d.ldBp = nb_;
    }

    tmc_mlp_config cfg_;
    int H_ = 1, in_pad_ = 128, out_pad_ = 128, hid_ = 128, B_ = 128, Bq_ = 128;
    bool fused_ = false;  // inference through infer_fused
    // its weights in fragment order (pack_fragments), per head Pf_ elements, matrix m at
    // frag_off_[m]; stale after every parameter change until the next inference repacks them
    bf16 *wf_ = nullptr;
    long long Pf_ = 0;
    std::vector<long long> frag_off_;
    bool frag_stale_ = true;
    int k0_ = 128;  // K of the first layer: the inputs rounded up to a stage (32)
    long long P_ = 0, Pb_ = 0;
    std::vector<mlp::MatInfo> mats_;
    std::vector<int> real_in_, real_out_;
    std::vector<void *> allocs_;
    // surrogate answers (allocated on first use): device buffers, pinned slots, events
    double *ans_raw_[2] = {nullptr, nullptr}, *ans_res_ = nullptr;
    hipStream_t ans_copy_ = nullptr;  // the answers' input copies
    hipEvent_t ans_h2d_[2] = {nullptr, nullptr}, ans_used_[2] = {nullptr, nullptr};
    uint8_t *ans_dom_ = nullptr, *ans_route_ = nullptr;
    float *ans_prof_ = nullptr, *ans_ymean_ = nullptr, *ans_ystd_ = nullptr;
    double *ans_pin_raw_[2] = {nullptr, nullptr}, *ans_pin_res_[2] = {nullptr, nullptr};
    float *ans_pin_prof_[2] = {nullptr, nullptr};
    hipEvent_t ans_k0_[2] = {nullptr, nullptr}, ans_k1_[2] = {nullptr, nullptr}, ans_done_[2] = {nullptr, nullptr};
    std::vector<void *> pinned_;

    void answer_buffers(bool profiles) {
        if (!ans_res_) {
            MLP_TRY(hipStreamCreateWithFlags(&ans_copy_, hipStreamNonBlocking));
            ans_res_ = alloc<double>(4ll * Bq_);
            ans_dom_ = alloc<uint8_t>(Bq_);
            ans_route_ = alloc<uint8_t>(Bq_);
            ans_ymean_ = alloc<float>(cfg_.outputs);
            ans_ystd_ = alloc<float>(cfg_.outputs);
            for (int i = 0; i < 2; ++i) {
                ans_raw_[i] = alloc<double>(static_cast<long long>(Bq_) * mlp::kRawPerStack);
                MLP_TRY(hipEventCreateWithFlags(&ans_h2d_[i], hipEventDisableTiming));
                MLP_TRY(hipEventCreateWithFlags(&ans_used_[i], hipEventDisableTiming));
                ans_pin_raw_[i] = pinned<double>(static_cast<size_t>(Bq_) * mlp::kRawPerStack);
                ans_pin_res_[i] = pinned<double>(4 * static_cast<size_t>(Bq_) + (2 * static_cast<size_t>(Bq_) + 7) / 8);
                MLP_TRY(hipEventCreate(&ans_k0_[i]));
                MLP_TRY(hipEventCreate(&ans_k1_[i]));
                MLP_TRY(hipEventCreateWithFlags(&ans_done_[i], hipEventDisableTiming));
            }
        }
        if (profiles && !ans_prof_) {
            const size_t cells = static_cast<size_t>(Bq_) * (cfg_.outputs - 2);
            ans_prof_ = alloc<float>(static_cast<long long>(cells));
            for (int i = 0; i < 2; ++i) ans_pin_prof_[i] = pinned<float>(cells);
        }
    }
    template <class T> T *pinned(size_t n) {
        void *q = nullptr;
        MLP_TRY(hipHostMalloc(&q, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
        pinned_.push_back(q);
        return static_cast<T *>(q);
    }
    int splits_ = 1;
    float *bpart_ = nullptr;  // bias-gradient partials [H x B / 64 x nb_]
    int nb_ = 0;              // biases per head
    float *part_ = nullptr;
    float *p_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr, *lossw_ = nullptr, *loss_ = nullptr;
    float *lpart_ = nullptr;  // per-tile loss partials [heads x tiles]
    bf16 *wb_ = nullptr;
    uint64_t *step_ = nullptr;
    float *lr_ = nullptr;
    bf16 *X_ = nullptr, *dY_ = nullptr;
    float *Yt_ = nullptr, *Yp_ = nullptr;
    std::vector<bf16 *> G_;  // GELU'(Z) of each hidden layer (bf16, fragment order)
    std::vector<bf16 *> Hs_;
    std::vector<bf16 *> dZ_;  // dZ of each hidden layer [heads x B x hidden]
    bf16 *Xq_ = nullptr, *Hq_[2] = {nullptr, nullptr};
    float *xq_ = nullptr, *mean_ = nullptr, *spread_ = nullptr;
    float *xd_ = nullptr, *yd_ = nullptr;
    int64_t n_data_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;

  public:
    void free_data() {
        if (xd_) (void)hipFree(xd_);
        if (yd_) (void)hipFree(yd_);
        xd_ = yd_ = nullptr;
    }
};

MlpModel *mlp_create(const tmc_mlp_config &c) { return new MlpModel(c); }
void mlp_destroy(MlpModel *m) {
    if (m) m->free_data();
    delete m;
}
void mlp_set_data(MlpModel &m, const float *x, const float *y, int64_t n, const float *w) { m.set_data(x, y, n, w); }
double mlp_train(MlpModel &m, int64_t steps, float *losses) { return m.train(steps, losses); }
void mlp_gradients(MlpModel &m, const int64_t *rows, float *grads, float *loss) { m.gradients(rows, grads, loss); }
double mlp_predict(MlpModel &m, const float *x, int64_t n, float *mean, float *spread, float *per_head) {
    return m.predict(x, n, mean, spread, per_head);
}
void mlp_answer(MlpModel &m, const tmc_surrogate_gate &gate, const tmc_stack *stacks, int64_t n,
                tmc_surrogate_answers *out) {
    m.answer(gate, stacks, n, out);
}
void mlp_get_params(MlpModel &m, float *out) { m.get_params(out); }
void mlp_set_params(MlpModel &m, const float *in) { m.set_params(in); }
void mlp_set_lr(MlpModel &m, float lr) { m.set_lr(lr); }
long long mlp_param_count(const MlpModel &m) { return m.logical_params(); }
uint64_t mlp_flops_per_step(const MlpModel &m) { return m.flops_per_step(); }
uint64_t mlp_flops_per_query(const MlpModel &m) { return m.flops_per_query(); }

}  // namespace tmc
