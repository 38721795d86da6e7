// mlp.hpp -- launch interface of the surrogate MLP kernels (mlp.hip).
// bfloat16 values are stored as their 16-bit patterns (uint16_t).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tg {

constexpr int MLP_IN = 8;     // input features (6 used, 2 zero)
constexpr int MLP_H = 128;    // hidden width
constexpr int MLP_NOUT = 3;   // d log10 T, d HII fraction, log10 estimated relative error
// bf16 weight blob: W1[128][8] | W2 | W3 | W4 [128][128] | W5[NOUT][128]
constexpr long long MLP_NW = (long long)MLP_H * MLP_IN + 3LL * MLP_H * MLP_H + (long long)MLP_NOUT * MLP_H;
// fp32 bias blob: b1 | b2 | b3 | b4 [128] | b5[NOUT]
constexpr long long MLP_NB = 4LL * MLP_H + MLP_NOUT;

enum GemmMode { GEMM_FWD_HIDDEN = 0, GEMM_FWD_OUT = 1, GEMM_BWD_DATA = 2, GEMM_SPLIT_K = 3 };

// C[M][N] = A[M][K] . B[K][N]; element (m,k) of A at A[m*sam + k*sak],
// element (k,n) of B at B[k*sbk + n*sbn].
struct GemmArgs {
    const uint16_t* A;
    long long sam, sak;
    const uint16_t* B;
    long long sbk, sbn;
    int M, N, K;
    int mode;
    int kchunk;        // SPLIT_K: K range per grid z index (multiple of 32)
    int pad;
    const float* bias; // FWD_*: [N]
    float* Z;          // FWD_*: fp32 out [M][ldz]
    long long ldz;
    uint16_t* H;       // FWD_HIDDEN: bf16 SiLU(Z); BWD_DATA: bf16 dZ_prev
    long long ldh;
    const float* Zprev;  // BWD_DATA: pre-activation of the previous layer
    long long ldzp;
    float* partial;    // SPLIT_K: [ksplit][M][N]; BWD_DATA: bias partials [row tiles][N]
};

hipError_t launch_gemm(const GemmArgs& a, int ksplit, hipStream_t stream);
hipError_t launch_reduce_partials(const float* partial, int nz, long long n, float* out, int accumulate,
                                  hipStream_t stream);
hipError_t launch_loss_grad(const float* y, const float* t, int B, int nout, float m0, float m1, float m2,
                            int err_head, float* dy32, uint16_t* dy16, float* loss_partial, hipStream_t stream);
hipError_t launch_colsum(const float* x, int B, int n, float* partial, hipStream_t stream);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint16_t* p16, long long n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float wd, hipStream_t stream);
hipError_t launch_to_bf16(const float* x, uint16_t* y, long long n, hipStream_t stream);
hipError_t launch_mlp_infer(const uint16_t* wts, const float* bias, const float* feat, long long n, float* y,
                            int nblocks, hipStream_t stream);
// as launch_mlp_infer with the cell count read from device memory (<= n_max)
hipError_t launch_mlp_infer_dev(const uint16_t* wts, const float* bias, const float* feat,
                                const unsigned long long* n_dev, long long n_max, float* y, int nblocks,
                                hipStream_t stream);

}  // namespace tg
