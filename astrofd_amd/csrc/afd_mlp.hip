// afd_mlp.hip -- kernels of the learned con2prim guess: weight packing, dataset preparation,
// inference, the fused forward/backward gradient kernel (BF16 MFMA, FP32 accumulation),
// Adam, and the con2prim batch with the learned initial guess.
#include <cstring>
#include <vector>

#include "afd_mlp.hpp"

namespace afd {

int mlp_alloc(MlpDevice &m, const std::function<int(void **, size_t)> &alloc) {
    int rc;
    if ((rc = alloc((void **)&m.master, MLP_NPARAM * 4))) return rc;
    if ((rc = alloc(&m.w, MLP_DEV_BYTES))) return rc;
    if ((rc = alloc((void **)&m.adam_m, MLP_NPARAM * 4))) return rc;
    if ((rc = alloc((void **)&m.adam_v, MLP_NPARAM * 4))) return rc;
    if ((rc = alloc((void **)&m.grad, MLP_NPARAM * 4))) return rc;
    return alloc((void **)&m.loss, 8 * 65536);
}

// Master FP32 weights -> BF16 A-operand blocks (lane order, afd_mlp_device.hpp) + FP32 biases.
// A bias b as hi + lo: hi = the nearest BF16 of b, lo = the nearest BF16 of b - hi (exact in FP32).
AFD_D float bias_part(float b, int which) {
    const float hi = (float)(__bf16)b;
    return which == 0 ? hi : b - hi;
}

__global__ __launch_bounds__(BLOCK) void k_mlp_pack(const float *master, void *wout) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    __bf16 *blk = (__bf16 *)wout;
    float *bias = (float *)((char *)wout + (size_t)MLP_NBLK * 64 * 16);
    if (t < MLP_NBLK * 64 * 8) {
        const int e = t % 8, lane = (t / 8) % 64, b = t / 512;
        const int m = lane % 32, k = 8 * (lane / 32) + e;
        float v = 0.0f;
        if (b >= BLK_B2) {  // b2 as the split pair hi + lo in K columns 0, 1
            const float bb = master[MLP_OFF_B2 + 32 * (b - BLK_B2) + m];
            if (k < 2) v = bias_part(bb, k);
        } else if (b < BLK_W2) {
            int mt = b - BLK_W1;
            if (k < MLP_IN) v = master[MLP_OFF_W1 + (32 * mt + m) * MLP_IN + k];
            else if (k < MLP_IN + 2) v = bias_part(master[MLP_OFF_B1 + 32 * mt + m], k - MLP_IN);
        } else if (b < BLK_W3) {
            int idx = b - BLK_W2, mt = idx / 4, kb = idx % 4;
            v = master[MLP_OFF_W2 + (32 * mt + m) * 64 + 16 * kb + k];
        } else if (b < BLK_W3T) {
            int kb = b - BLK_W3;
            if (m < MLP_OUT) v = master[MLP_OFF_W3 + m * 64 + 16 * kb + k];
        } else if (b < BLK_W2T) {
            int mt = b - BLK_W3T;
            if (k < MLP_OUT) v = master[MLP_OFF_W3 + k * 64 + 32 * mt + m];
        } else {
            int idx = b - BLK_W2T, mt = idx / 4, kb = idx % 4;
            v = master[MLP_OFF_W2 + (16 * kb + k) * 64 + 32 * mt + m];
        }
        blk[t] = (__bf16)v;
    }
    if (t < MLP_BIAS_FLOATS) {
        if (t < 64) bias[t] = master[MLP_OFF_B1 + t];
        else if (t < 128) bias[t] = master[MLP_OFF_B2 + t - 64];
        else bias[t] = master[MLP_OFF_B3 + t - 128];
    }
}

hipError_t mlp_pack(const float *master, void *w, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NBLK * 64 * 8 + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, master, w);
    return hipGetLastError();
}

hipError_t mlp_upload(MlpDevice &m, const float *w, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(m.master, w, MLP_NPARAM * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(m.adam_m, 0, MLP_NPARAM * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(m.adam_v, 0, MLP_NPARAM * 4, s);
    if (e == hipSuccess) e = mlp_pack(m.master, m.w, s);
    m.adam_t = 0;
    return e;
}

// ------------------------------------------------------------------ dataset
// (conserved -> primitive) pairs of the solver's own con2prim: p2c of the sampled primitives,
// the FP64 RMHD con2prim, features of the conserved state and scaled targets (log p/D, log W).
__global__ __launch_bounds__(BLOCK) void k_mlp_dataset(MlpDataArgs a) {
    long m = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (m >= a.n) return;
    double q[9], U[9], out[9];
    for (int v = 0; v < 9; ++v) q[v] = a.prim[9 * m + v];
    rmhd_p2c_cell(a.gfac, a.p_max, a.v_max, q, U);
    RmhdC2PParams cp = {a.gamma, a.gfac, a.p_max, a.v_max};
    C2PStats st = {0, 0, 0};
    rmhd_c2p(cp, U, out, st);
    float x[5];
    bool ok = mlp_features(U, x);
    double v2 = out[1] * out[1] + out[2] * out[2] + out[3] * out[3];
    double W = 1.0 / sqrt(1.0 - (v2 < 1.0 ? v2 : 1.0 - 1e-14));
    float y0 = (float)(log10(out[4] / U[0]) * 0.25), y1 = (float)log10(W);
    // only pairs where the con2prim converged by Newton to a valid state: elsewhere it returns the
    // bisection result or the hydro guess, which no initial guess changes
    ok = ok && isfinite(y0) && isfinite(y1) && (st.flags & 6) == 0;
    for (int f = 0; f < 5; ++f) a.X[5 * m + f] = ok ? x[f] : 0.0f;
    a.Y[2 * m] = ok ? y0 : 0.0f;
    a.Y[2 * m + 1] = ok ? y1 : 0.0f;
    a.valid[m] = ok ? 1 : 0;
}

hipError_t mlp_prepare_dataset(const MlpDataArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_dataset, grid1d(a.n), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ inference (test hook)
__global__ __launch_bounds__(BLOCK) void k_mlp_forward(const void *w, const float *X, long n, float *Y) {
    long sidx = (long)blockIdx.x * BLOCK + threadIdx.x;
    float x[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, y[2];
    if (sidx < n)
        for (int f = 0; f < 5; ++f) x[f] = X[5 * sidx + f];
    mlp_infer_wave(MlpGlobal{mlp_view(w)}, x, y);  // every lane of the wave participates
    if (sidx < n) {
        Y[2 * sidx] = y[0];
        Y[2 * sidx + 1] = y[1];
    }
}

hipError_t mlp_forward(const void *w, const float *X, long n, float *Y, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_forward, grid1d(n), dim3(BLOCK), 0, s, w, X, n, Y);
    return hipGetLastError();
}

// ------------------------------------------------------------------ forward + backward
// One wavefront per block.  For each 32-sample tile: forward (MFMA), output error, backprop of
// the activations through W3^T and W2^T (MFMA), then the weight gradients as GEMMs over the
// sample dimension (MFMA) from BF16 tiles staged in LDS as [feature][sample].
constexpr int LDS_S = 40;  // row stride in BF16 elements (32 samples + pad, 80 bytes)

__global__ __launch_bounds__(64) void k_mlp_train_grad(const void *wp, const float *X, const float *Yt, long n,
                                                       float inv_n, float *grad, double *loss_out) {
    __shared__ __attribute__((aligned(16))) __bf16 sH1[64 * LDS_S], sH2[64 * LDS_S], sDZ1[64 * LDS_S],
        sDZ2[64 * LDS_S], sX[32 * LDS_S], sDO[32 * LDS_S];
    const MlpDevW w = mlp_view(wp);
    const int lane = threadIdx.x;
    const bool upper = lane >= 32;
    const int col = lane & 31;
    for (int r = 0; r < 32; ++r) {  // rows of X beyond the 5 features and of dO beyond the 2 outputs stay 0
        sX[r * LDS_S + col] = (__bf16)0.0f;
        sDO[r * LDS_S + col] = (__bf16)0.0f;
    }
    floatx16 gW2[2][2], gW1[2], gW3[2];
    for (int a = 0; a < 2; ++a) {
        gW1[a] = zero16();
        gW3[a] = zero16();
        for (int b = 0; b < 2; ++b) gW2[a][b] = zero16();
    }
    float gb1[2][16], gb2[2][16], gb3[2] = {0.f, 0.f};
    for (int a = 0; a < 2; ++a)
        for (int i = 0; i < 16; ++i) gb1[a][i] = gb2[a][i] = 0.f;
    double lsum = 0.0;
    __syncthreads();
    for (long base = (long)blockIdx.x * 64; base < n; base += (long)gridDim.x * 64) {
        const long sidx = base + lane;
        float x[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, y[2] = {0.f, 0.f};
        if (sidx < n) {
            for (int f = 0; f < 5; ++f) x[f] = X[5 * sidx + f];
            y[0] = Yt[2 * sidx];
            y[1] = Yt[2 * sidx + 1];
        }
        float xs[5], ys[2];
        for (int f = 0; f < 5; ++f) xs[f] = __shfl_xor(x[f], 32);
        ys[0] = __shfl_xor(y[0], 32);
        ys[1] = __shfl_xor(y[1], 32);
        for (int t = 0; t < 2; ++t) {
            const float *xt = t ? xs : x;
            const float *yt = t ? ys : y;
            const bool tlive = base + 32 * t + col < n;
            MlpTile T;
            mlp_tile_forward(MlpGlobal{w}, lane, input_operand(xt, upper), T);
            float d0 = 0.f, d1 = 0.f;
            if (!upper && tlive) {
                float e0 = T.o[0] - yt[0], e1 = T.o[1] - yt[1];
                d0 = 2.0f * inv_n * e0;
                d1 = 2.0f * inv_n * e1;
                lsum += (double)(e0 * e0 + e1 * e1) * inv_n;
                gb3[0] += d0;
                gb3[1] += d1;
            }
            const bf16x8 dob = upper ? pack8(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f) : pack8(d0, d1, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
            floatx16 dz2[2], dz1[2];
            for (int mt = 0; mt < 2; ++mt) {
                dz2[mt] = mfma_bf16(w.blk[(BLK_W3T + mt) * 64 + lane], dob, zero16());
                for (int i = 0; i < 16; ++i) dz2[mt][i] *= dsilu(T.z2[mt][i]);
            }
            bf16x8 dz2b[4];
            d_to_b(dz2, lane, dz2b);
            for (int mt = 0; mt < 2; ++mt) {
                floatx16 acc = zero16();
                for (int kb = 0; kb < 4; ++kb) acc = mfma_bf16(w.blk[(BLK_W2T + 4 * mt + kb) * 64 + lane], dz2b[kb], acc);
                for (int i = 0; i < 16; ++i) acc[i] *= dsilu(T.z1[mt][i]);
                dz1[mt] = acc;
            }
            for (int mt = 0; mt < 2; ++mt)
                for (int i = 0; i < 16; ++i) {
                    int row = 32 * mt + drow(i, lane);
                    gb2[mt][i] += dz2[mt][i];
                    gb1[mt][i] += dz1[mt][i];
                    sH1[row * LDS_S + col] = (__bf16)silu(T.z1[mt][i]);
                    sH2[row * LDS_S + col] = (__bf16)silu(T.z2[mt][i]);
                    sDZ2[row * LDS_S + col] = (__bf16)dz2[mt][i];
                    sDZ1[row * LDS_S + col] = (__bf16)dz1[mt][i];
                }
            if (!upper) {
                for (int f = 0; f < 5; ++f) sX[f * LDS_S + col] = (__bf16)xt[f];
                sDO[0 * LDS_S + col] = (__bf16)d0;
                sDO[1 * LDS_S + col] = (__bf16)d1;
            }
            __syncthreads();
            for (int kb = 0; kb < 2; ++kb) {
                const int cb = 16 * kb + 8 * (lane >> 5);
                bf16x8 aDZ2[2], aDZ1[2], bH1[2], bH2[2];
                for (int mt = 0; mt < 2; ++mt) {
                    aDZ2[mt] = *(const bf16x8 *)&sDZ2[(32 * mt + col) * LDS_S + cb];
                    aDZ1[mt] = *(const bf16x8 *)&sDZ1[(32 * mt + col) * LDS_S + cb];
                    bH1[mt] = *(const bf16x8 *)&sH1[(32 * mt + col) * LDS_S + cb];
                    bH2[mt] = *(const bf16x8 *)&sH2[(32 * mt + col) * LDS_S + cb];
                }
                const bf16x8 aDO = *(const bf16x8 *)&sDO[col * LDS_S + cb];
                const bf16x8 bX = *(const bf16x8 *)&sX[col * LDS_S + cb];
                for (int mt = 0; mt < 2; ++mt) {
                    for (int nt = 0; nt < 2; ++nt) gW2[mt][nt] = mfma_bf16(aDZ2[mt], bH1[nt], gW2[mt][nt]);
                    gW1[mt] = mfma_bf16(aDZ1[mt], bX, gW1[mt]);
                }
                for (int nt = 0; nt < 2; ++nt) gW3[nt] = mfma_bf16(aDO, bH2[nt], gW3[nt]);
            }
            __syncthreads();
        }
    }
    // flush the wave's partial gradients
    for (int mt = 0; mt < 2; ++mt)
        for (int i = 0; i < 16; ++i) {
            int row = 32 * mt + drow(i, lane);
            for (int nt = 0; nt < 2; ++nt) atomicAdd(&grad[MLP_OFF_W2 + row * 64 + 32 * nt + col], gW2[mt][nt][i]);
            if (col < MLP_IN) atomicAdd(&grad[MLP_OFF_W1 + row * MLP_IN + col], gW1[mt][i]);
        }
    for (int nt = 0; nt < 2; ++nt)
        for (int i = 0; i < 16; ++i) {
            int row = drow(i, lane);
            if (row < MLP_OUT) atomicAdd(&grad[MLP_OFF_W3 + row * 64 + 32 * nt + col], gW3[nt][i]);
        }
    for (int mt = 0; mt < 2; ++mt)
        for (int i = 0; i < 16; ++i) {
            float v1 = gb1[mt][i], v2 = gb2[mt][i];
            for (int off = 16; off > 0; off >>= 1) {
                v1 += __shfl_xor(v1, off);
                v2 += __shfl_xor(v2, off);
            }
            if (col == 0) {
                int row = 32 * mt + drow(i, lane);
                atomicAdd(&grad[MLP_OFF_B1 + row], v1);
                atomicAdd(&grad[MLP_OFF_B2 + row], v2);
            }
        }
    for (int j = 0; j < 2; ++j) {
        float v = gb3[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) atomicAdd(&grad[MLP_OFF_B3 + j], v);
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) atomicAdd(loss_out, lsum);
}

hipError_t mlp_loss_grad(const void *w, const float *X, const float *Y, long n, float *grad, double *loss,
                         hipStream_t s) {
    hipError_t e = hipMemsetAsync(grad, 0, MLP_NPARAM * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(loss, 0, 8, s);
    if (e != hipSuccess) return e;
    long chunks = (n + 63) / 64;
    int blocks = (int)(chunks < 2048 ? chunks : 2048);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_train_grad, dim3(blocks), dim3(64), 0, s, w, X, Y, n, 1.0f / (float)n, grad, loss);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Adam (Kingma & Ba 2015)
__global__ __launch_bounds__(BLOCK) void k_mlp_adam(float *p, float *m, float *v, const float *g, int n, float lr,
                                                     float bc1, float bc2) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = 0.9f * m[i] + 0.1f * gi;
    float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

hipError_t mlp_adam(float *master, float *m, float *v, const float *grad, long t, float lr, hipStream_t s) {
    float bc1 = 1.0f - powf(0.9f, (float)t), bc2 = 1.0f - powf(0.999f, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, master, m, v, grad,
                       MLP_NPARAM, lr, bc1, bc2);
    return hipGetLastError();
}

// ------------------------------------------------------------------ con2prim batch
__global__ __launch_bounds__(BLOCK) void k_c2p_batch_learned(BatchArgs a, const void *mlp) {
    long m = (long)blockIdx.x * BLOCK + threadIdx.x;
    const bool live = m < a.n;
    double U[9], q[9];
    for (int v = 0; v < 9; ++v) U[v] = live ? a.in[9 * m + v] : 0.0;
    C2PStats st = {0, 0, 0};
    RmhdC2PParams cp = {a.gamma, a.gfac, a.p_max, a.v_max};
    float x[5], y[2];
    bool fok = mlp_features(U, x) && live;
    mlp_infer_wave(MlpGlobal{mlp_view(mlp)}, x, y);
    bool done = fok && rmhd_c2p_from_guess(cp, U, mlp_guess_Z(U[0], a.gfac, y), q, st);
    if (live && !done) rmhd_c2p(cp, U, q, st);
    if (live) {
        for (int v = 0; v < 9; ++v) a.out[9 * m + v] = q[v];
        if (a.iters) a.iters[m] = st.hydro_iters + st.newton_iters;
    }
    if (a.stats) wave_add_stats(a.stats, st);
}

void launch_c2p_batch(const BatchArgs &a, const C2PBatchExtra &x, hipStream_t s) {
    if (x.mlp && a.physics == 1)
        hipLaunchKernelGGL(k_c2p_batch_learned, grid1d(a.n), dim3(BLOCK), 0, s, a, x.mlp);
    else
        hipLaunchKernelGGL(k_c2p_batch, grid1d(a.n), dim3(BLOCK), 0, s, a);
}

}  // namespace afd
