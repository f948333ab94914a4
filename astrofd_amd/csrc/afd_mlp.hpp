// afd_mlp.hpp -- storage of the learned con2prim guess (MLP 5 -> 64 -> 64 -> 2, SiLU).
//
// Packed FP32 master layout (4674 values):
//   W1[64][5], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
// The device keeps the FP32 master and a BF16 copy of the weight matrices laid out as
// MFMA A-operands (the biases stay FP32).
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <functional>

#include "afd_kernels.hpp"

namespace afd {

constexpr int MLP_IN = 5, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;  // 4674
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN, MLP_OFF_W2 = MLP_OFF_B1 + MLP_H,
              MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H, MLP_OFF_W3 = MLP_OFF_B2 + MLP_H,
              MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_H;

struct MlpDevice {
    float *master = nullptr;  // MLP_NPARAM FP32
    void *w = nullptr;        // packed inference weights (BF16 matrices + FP32 biases)
    bool loaded = false;
};

// BF16 inference block: W1 padded to [64][16], W2 [64][64], W3 padded to [32][64] (rows 2..31 zero),
// then FP32 biases b1[64], b2[64], b3[2].
constexpr int MLP_W1P = 64 * 16, MLP_W2P = 64 * 64, MLP_W3P = 32 * 64;
constexpr size_t MLP_BF16_BYTES = (size_t)(MLP_W1P + MLP_W2P + MLP_W3P) * 2;
constexpr size_t MLP_DEV_BYTES = MLP_BF16_BYTES + (size_t)(64 + 64 + 2) * 4;

int mlp_alloc(MlpDevice &m, const std::function<int(void **, size_t)> &alloc);
hipError_t mlp_upload(MlpDevice &m, const float *host_master, hipStream_t s);

struct C2PBatchExtra {
    const void *mlp;  // packed inference weights, or null for the plain solve
};
void launch_c2p_batch(const BatchArgs &a, const C2PBatchExtra &x, hipStream_t s);

}  // namespace afd
