// afd_mlp.hip -- weight upload of the learned guess and the c2p batch launcher.
#include <cstring>
#include <vector>

#include "afd_mlp.hpp"

namespace afd {

int mlp_alloc(MlpDevice &m, const std::function<int(void **, size_t)> &alloc) {
    int rc = alloc((void **)&m.master, MLP_NPARAM * 4);
    if (rc) return rc;
    return alloc(&m.w, MLP_DEV_BYTES);
}

static uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

hipError_t mlp_upload(MlpDevice &m, const float *w, hipStream_t s) {
    std::vector<unsigned char> blob(MLP_DEV_BYTES, 0);
    uint16_t *W1 = (uint16_t *)blob.data();
    uint16_t *W2 = W1 + MLP_W1P;
    uint16_t *W3 = W2 + MLP_W2P;
    float *b = (float *)(blob.data() + MLP_BF16_BYTES);
    for (int h = 0; h < 64; ++h)
        for (int k = 0; k < 16; ++k) W1[h * 16 + k] = k < MLP_IN ? to_bf16(w[MLP_OFF_W1 + h * MLP_IN + k]) : 0;
    for (int h = 0; h < 64; ++h)
        for (int k = 0; k < 64; ++k) W2[h * 64 + k] = to_bf16(w[MLP_OFF_W2 + h * 64 + k]);
    for (int o = 0; o < 32; ++o)
        for (int k = 0; k < 64; ++k) W3[o * 64 + k] = o < MLP_OUT ? to_bf16(w[MLP_OFF_W3 + o * 64 + k]) : 0;
    for (int h = 0; h < 64; ++h) b[h] = w[MLP_OFF_B1 + h];
    for (int h = 0; h < 64; ++h) b[64 + h] = w[MLP_OFF_B2 + h];
    for (int o = 0; o < 2; ++o) b[128 + o] = w[MLP_OFF_B3 + o];
    hipError_t e = hipMemcpyAsync(m.master, w, MLP_NPARAM * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(m.w, blob.data(), blob.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);  // blob is a host temporary
}

void launch_c2p_batch(const BatchArgs &a, const C2PBatchExtra &x, hipStream_t s) {
    (void)x;
    hipLaunchKernelGGL(k_c2p_batch, grid1d(a.n), dim3(BLOCK), 0, s, a);
}

}  // namespace afd
