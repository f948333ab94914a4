// afd_mlp.hpp -- the learned con2prim guess: storage, training and inference launchers.
//
// MLP 5 -> 64 -> 64 -> 2 with SiLU.  Packed FP32 master layout (4674 values):
//   W1[64][5], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
// The device keeps the FP32 master, Adam moments, a gradient buffer and the BF16 copy of
// the weight matrices laid out as MFMA A operands (afd_mlp_device.hpp) followed by the
// FP32 biases.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "afd_kernels.hpp"
#include "afd_mlp_device.hpp"

namespace afd {

constexpr int MLP_IN = 5, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;  // 4674
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN, MLP_OFF_W2 = MLP_OFF_B1 + MLP_H,
              MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H, MLP_OFF_W3 = MLP_OFF_B2 + MLP_H,
              MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_H;
constexpr size_t MLP_DEV_BYTES = (size_t)MLP_NBLK * 64 * 16 + (size_t)MLP_BIAS_FLOATS * 4;

struct MlpDevice {
    float *master = nullptr;   // MLP_NPARAM FP32
    void *w = nullptr;         // packed inference weights
    float *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    double *loss = nullptr;    // per training step
    float *X = nullptr, *Y = nullptr;  // dataset (n x 5 features, n x 2 targets)
    long n_data = 0, cap_data = 0;
    long adam_t = 0;
    bool loaded = false;
};

int mlp_alloc(MlpDevice &m, const std::function<int(void **, size_t)> &alloc);
hipError_t mlp_upload(MlpDevice &m, const float *host_master, hipStream_t s);

struct C2PBatchExtra {
    const void *mlp;  // packed inference weights, or null for the plain solve
};
void launch_c2p_batch(const BatchArgs &a, const C2PBatchExtra &x, hipStream_t s);

// dataset from primitive states: p2c, the solver's own con2prim, features and targets
struct MlpDataArgs {
    const double *prim;  // n x 9
    float *X, *Y;
    int *valid;
    long n;
    double gamma, gfac, p_max, v_max;
};
hipError_t mlp_prepare_dataset(const MlpDataArgs &a, hipStream_t s);
hipError_t mlp_forward(const void *w, const float *X, long n, float *Y, hipStream_t s);
// gradient of loss = mean over samples of sum over outputs (y - t)^2 accumulated into grad (zeroed here)
hipError_t mlp_loss_grad(const void *w, const float *X, const float *Y, long n, float *grad, double *loss,
                         hipStream_t s);
hipError_t mlp_adam(float *master, float *m, float *v, const float *grad, long t, float lr, hipStream_t s);
hipError_t mlp_pack(const float *master, void *w, hipStream_t s);

}  // namespace afd
