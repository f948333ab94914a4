// afd_mlp_device.hpp -- wave-level BF16 MFMA multilayer perceptron (5 -> 64 -> 64 -> 2, SiLU).
//
// One wavefront evaluates 64 samples as two tiles of 32 with v_mfma_f32_32x32x16_bf16.
// Activations are kept transposed (features x samples) so that samples run along the
// lanes: a layer is  Out^T (rows x 32 samples) = W (rows x K) * In^T (K x 32 samples)
// with the weights as the A operand (pre-packed in lane order, BF16) and the activations
// as the B operand.  MFMA operand layouts (32x32x16, 64 lanes, l = lane):
//   A (32 x 16):  lane l holds A[l % 32][8 (l / 32) + e],  e = 0..7
//   B (16 x 32):  lane l holds B[8 (l / 32) + e][l % 32]
//   D (32 x 32):  lane l holds D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32],  i = 0..15
// A D-layout result becomes the next layer's B operand after one exchange of four values
// between lanes l and l ^ 32 per 16-row K block.  Accumulation is FP32.
// The biases b1 and b2 ride in the MFMAs as a split pair b = hi + lo of BF16 values (hi the
// nearest BF16 of b, lo that of the remainder: 16 significant bits) against constant-one input
// rows: b1 in columns 5, 6 of W1 (inputs 5, 6 are one), b2 in an extra K block whose rows 0, 1
// are one.  No bias lives in registers across the layers.
#pragma once
#include <hip/hip_runtime.h>

#include "afd_device.hpp"

namespace afd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Packed device weights: NBLK blocks of 64 lanes x 8 BF16 (A operands), then FP32 biases.
enum MlpBlock {
    BLK_W1 = 0,    // [W1 b1_hi b1_lo 0 ...] [64 x 16]: m-tile 0..1, k-block 0 (inputs 5, 6 are one)
    BLK_W2 = 2,    // W2 [64 x 64]: m-tile 0..1 x k-block 0..3  (index 2 + 4 mt + kb)
    BLK_W3 = 10,   // W3 [32 x 64]: k-block 0..3 (rows 2..31 zero)
    BLK_W3T = 14,  // W3^T [64 x 16]: m-tile 0..1 (columns 2..15 zero)
    BLK_W2T = 16,  // W2^T [64 x 64]: m-tile 0..1 x k-block 0..3
    BLK_B2 = 24,   // [b2_hi b2_lo 0 ...] [64 x 16]: m-tile 0..1, against the ones block
    MLP_NBLK = 26
};
constexpr int MLP_BIAS_FLOATS = 64 + 64 + 2;

struct MlpDevW {
    const bf16x8 *blk;  // MLP_NBLK * 64 vectors
    const float *bias;  // b1[64], b2[64], b3[2]
};

AFD_D MlpDevW mlp_view(const void *w) {
    // an opaque copy of the pointer: a kernel that evaluates the network inside a loop would
    // otherwise hoist the 26 loop-invariant weight operands (104 VGPRs) out of it and spill
    asm volatile("" : "+s"(w));
    MlpDevW v;
    v.blk = (const bf16x8 *)w;
    v.bias = (const float *)((const char *)w + (size_t)MLP_NBLK * 64 * 16);
    return v;
}

// Sources of the forward pass's weight operands: device memory (MlpGlobal) or a workgroup's LDS
// copy (MlpShared).  blk(b, lane) is the A operand of block b (MlpBlock numbering) at a lane.
struct MlpGlobal {
    MlpDevW w;
    AFD_D bf16x8 blk(int b, int lane) const { return w.blk[b * 64 + lane]; }
    AFD_D float b3(int i) const { return w.bias[128 + i]; }
};

// The blocks the forward pass reads (W1, W2, W3, B2: 16 KB) and b3, staged in LDS by the
// workgroups of the learned stage update.  Read from device memory, every wave's inference
// fetched them again (32 KB a wave), as much traffic through the caches as the update's state.
constexpr int MLP_FWD_BLK = 16;
struct MlpLds {
    bf16x8 blk[MLP_FWD_BLK][64];
    float b3[2];
};
// LDS slot of a forward block: W1, W2, W3 (blocks 0..13) in place, B2 (24, 25) at 14, 15
AFD_D int mlp_lds_slot(int b) { return b >= BLK_B2 ? b - BLK_B2 + 14 : b; }

struct MlpShared {
    const MlpLds *s;
    AFD_D bf16x8 blk(int b, int lane) const { return s->blk[mlp_lds_slot(b)][lane]; }
    AFD_D float b3(int i) const { return s->b3[i]; }
};

// Copies the forward blocks and b3 of the packed weights w to s, in two halves: mlp_stage_load
// issues a thread's loads into registers, mlp_stage_store writes them to LDS and ends with a
// barrier.  Called by every thread of a workgroup of NT threads (tid in [0, NT)).  The loads are
// all issued before the first store, so that a workgroup waits for one memory round trip, not
// one per copy; between the halves the workgroup can issue its own loads, whose round trip then
// overlaps the weights'.
template <int NT> struct MlpStaged {
    static_assert((MLP_FWD_BLK * 64) % NT == 0, "whole copies per thread");
    static constexpr int PER = MLP_FWD_BLK * 64 / NT;
    bf16x8 v[PER];
    float b3;
};

template <int NT> AFD_D void mlp_stage_load(const void *w, MlpStaged<NT> &st, int tid) {
    const bf16x8 *g = (const bf16x8 *)w;
    const float *bias = (const float *)((const char *)w + (size_t)MLP_NBLK * 64 * 16);
#pragma unroll
    for (int r = 0; r < MlpStaged<NT>::PER; ++r) {
        const int i = tid + r * NT, slot = i >> 6, l = i & 63;
        const int b = slot < 14 ? slot : BLK_B2 + (slot - 14);
        st.v[r] = g[b * 64 + l];
    }
    st.b3 = tid < 2 ? bias[128 + tid] : 0.0f;
}

template <int NT> AFD_D void mlp_stage_store(const MlpStaged<NT> &st, MlpLds &s, int tid) {
#pragma unroll
    for (int r = 0; r < MlpStaged<NT>::PER; ++r) {
        const int i = tid + r * NT;
        s.blk[i >> 6][i & 63] = st.v[r];
    }
    if (tid < 2) s.b3[tid] = st.b3;
    __syncthreads();
}

template <int NT> AFD_D void mlp_stage_lds(const void *w, MlpLds &s, int tid) {
    MlpStaged<NT> st;
    mlp_stage_load<NT>(w, st, tid);
    mlp_stage_store<NT>(st, s, tid);
}

AFD_D floatx16 mfma_bf16(bf16x8 a, bf16x8 b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

AFD_D floatx16 zero16() {
    floatx16 z;
#pragma unroll
    for (int i = 0; i < 16; ++i) z[i] = 0.0f;
    return z;
}

AFD_D int drow(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

AFD_D float silu(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
// d/dx silu(x) = s (1 + x (1 - s)), s = sigmoid(x)
AFD_D float dsilu(float x) {
    float s = __builtin_amdgcn_rcpf(1.0f + __expf(-x));
    return s * (1.0f + x * (1.0f - s));
}

AFD_D bf16x8 pack8(float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7) {
    bf16x8 r;
    r[0] = (__bf16)a0; r[1] = (__bf16)a1; r[2] = (__bf16)a2; r[3] = (__bf16)a3;
    r[4] = (__bf16)a4; r[5] = (__bf16)a5; r[6] = (__bf16)a6; r[7] = (__bf16)a7;
    return r;
}

// D layout of a 64-row activation (two 32-row m-tiles) -> B operands of its four 16-row K blocks.
// Lanes < 32 keep their `lo` rows and take the upper lanes' `lo`; lanes >= 32 take the lower
// lanes' `hi` and keep their `hi`: one v_permlane32_swap per value pair (it exchanges the upper
// half of its first operand with the lower half of its second), after which every lane packs
// (first, second) alike.
AFD_D void d_to_b(const floatx16 (&d)[2], int lane, bf16x8 (&bop)[4]) {
    (void)lane;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // lo: rows 16kk + 4 (l/32) + r; hi: rows 16kk + 8 + 4 (l/32) + r
                const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(d[mt][8 * kk + r]),
                                                                  __float_as_uint(d[mt][8 * kk + 4 + r]), false, false);
                a[r] = __uint_as_float(sw[0]);
                b[r] = __uint_as_float(sw[1]);
            }
            bop[2 * mt + kk] = pack8(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]);
        }
    }
}

// Forward pass of one 32-sample tile.  xb: B operand of the inputs (lanes < 32 hold the five
// features of sample l, lanes >= 32 zeros).  Returns pre-activations z1, z2 (bias included),
// the BF16 B operands of h1, h2 and the output tile o (rows 0, 1 = outputs at lanes < 32).
struct MlpTile {
    floatx16 z1[2], z2[2], o;
    bf16x8 h1b[4], h2b[4];
};

// B operand of the ones block: K rows 0 and 1 are one for every sample column.
AFD_D bf16x8 ones_operand(bool upper) {
    return upper ? pack8(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f) : pack8(1.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
}

template <class WS> AFD_D void mlp_tile_forward(const WS &w, int lane, bf16x8 xb, MlpTile &t) {
    floatx16 h[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        t.z1[mt] = mfma_bf16(w.blk(BLK_W1 + mt, lane), xb, zero16());
#pragma unroll
        for (int i = 0; i < 16; ++i) h[mt][i] = silu(t.z1[mt][i]);
    }
    d_to_b(h, lane, t.h1b);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        floatx16 acc = mfma_bf16(w.blk(BLK_B2 + mt, lane), ones_operand(lane >= 32), zero16());
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma_bf16(w.blk(BLK_W2 + 4 * mt + kb, lane), t.h1b[kb], acc);
#pragma unroll
        for (int i = 0; i < 16; ++i) h[mt][i] = silu(acc[i]);
        t.z2[mt] = acc;
    }
    d_to_b(h, lane, t.h2b);
    floatx16 o = zero16();
    o[0] = w.b3(0);
    o[1] = w.b3(1);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) o = mfma_bf16(w.blk(BLK_W3 + kb, lane), t.h2b[kb], o);
    t.o = o;
}

// B operand of the inputs: the five features and the two ones of the b1 columns.
AFD_D bf16x8 input_operand(const float x[5], bool upper) {
    return upper ? pack8(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f) : pack8(x[0], x[1], x[2], x[3], x[4], 1.f, 1.f, 0.f);
}

// Whole-wave inference: lane l supplies the features of its sample, gets its two outputs.
// Must be reached by all 64 lanes of the wave.
template <class WS> AFD_D void mlp_infer_wave(const WS &w, const float x[5], float y[2]) {
    const int lane = threadIdx.x & 63;
    const bool upper = lane >= 32;
    float xs[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) xs[f] = __shfl_xor(x[f], 32);  // tile 1: samples 32..63 to lanes 0..31
    MlpTile t0, t1;
    mlp_tile_forward(w, lane, input_operand(x, upper), t0);
    // one tile after the other: interleaved by the scheduler, the two tiles' activations took the
    // stage update to 163 registers (3 waves per SIMD); one at a time it needs 121 (4 waves)
    __builtin_amdgcn_sched_barrier(0);
    // the second tile loads the weights again rather than keep the first tile's live: an opaque
    // copy of the lane index keeps the compiler from reusing the first tile's loads
    int lane1 = lane;
    asm volatile("" : "+v"(lane1));
    mlp_tile_forward(w, lane1, input_operand(xs, upper), t1);
    float a0 = __shfl_xor(t1.o[0], 32), a1 = __shfl_xor(t1.o[1], 32);
    y[0] = upper ? a0 : t0.o[0];
    y[1] = upper ? a1 : t0.o[1];
}

// ------------------------------------------------------------------ features and targets
// Scaled inputs from the conserved state (D, S, tau, B): five O(1) numbers.
//   x0 = log10 D / 4, x1 = log10(tau / D) / 4, x2 = log10(|S| / D + 1e-10) / 4,
//   x3 = log10(B^2 / D + 1e-10) / 4, x4 = S.B / (|S| |B|)
// Outputs: y0 = log10(p / D) / 4, y1 = log10 W.
AFD_D bool mlp_features(const double *U, float x[5]) {
    double D = U[0], tau = U[4];
    double S2 = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
    double B2 = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
    double SB = U[1] * U[5] + U[2] * U[6] + U[3] * U[7];
    bool ok = D > 0.0 && tau > 0.0 && isfinite(D) && isfinite(tau) && isfinite(S2) && isfinite(B2);
    if (!ok) {
        for (int f = 0; f < 5; ++f) x[f] = 0.0f;
        return false;
    }
    // FP32 from here: the features only need FP32 resolution (log2 / rcp / sqrt instructions)
    const float c = 0.30102999566f * 0.25f;  // log10(2) / 4
    const float Df = (float)D, S2f = (float)S2, B2f = (float)B2;
    const float rD = __builtin_amdgcn_rcpf(Df);
    const float lD = __log2f(Df);
    const float Sm = __builtin_amdgcn_sqrtf(S2f), Bm = __builtin_amdgcn_sqrtf(B2f);
    x[0] = lD * c;
    x[1] = (__log2f((float)tau) - lD) * c;
    x[2] = __log2f(Sm * rD + 1e-10f) * c;
    x[3] = __log2f(B2f * rD + 1e-10f) * c;
    const float den = Sm * Bm;
    x[4] = den > 0.0f ? (float)SB * __builtin_amdgcn_rcpf(den) : 0.0f;
    for (int f = 0; f < 5; ++f)
        if (!isfinite(x[f])) ok = false;
    return ok;
}

// Z0 = rho h W^2 = D W + Gamma/(Gamma-1) p W^2 from the network outputs
AFD_D double mlp_guess_Z(double D, double gfac, const float y[2]) {
    const float l2_10 = 3.32192809489f;  // log2(10)
    double p = D * (double)__builtin_amdgcn_exp2f(4.0f * y[0] * l2_10);
    double W = (double)__builtin_amdgcn_exp2f(y[1] * l2_10);
    if (!(W >= 1.0)) W = 1.0;
    return D * W + gfac * p * W * W;
}

}  // namespace afd
