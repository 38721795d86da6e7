// mlp_api.cpp -- C ABI of the surrogate MLP: parameters, fused inference,
// and training (forward, loss, backward and Adam) sequenced over the GEMM
// kernels of mlp.hip.  Weight gradients are split-K partial products reduced
// in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/torchgpu.h"
#include "mlp.hpp"

using tg::GemmArgs;
using tg::MLP_H;
using tg::MLP_IN;
using tg::MLP_NB;
using tg::MLP_NOUT;
using tg::MLP_NW;

struct torchgpu_mlp {
    hipStream_t stream = nullptr;
    int max_batch = 0;
    int nsm = 256;
    long long step = 0;
    float* W = nullptr;   // fp32 master weights [MLP_NW]
    float* b = nullptr;   // fp32 biases [MLP_NB]
    uint16_t* W16 = nullptr;
    float *gW = nullptr, *gb = nullptr;
    float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
    // batch workspace
    uint16_t* H[5] = {};  // H0 [B][8], H1..H4 [B][128] (bf16)
    float* Z[6] = {};     // Z1..Z4 [B][128], Z5 = y [B][3] (fp32)
    uint16_t* dZ[5] = {}; // dZ1..dZ4 [B][128] (bf16)
    float* dy32 = nullptr;
    uint16_t* dy16 = nullptr;
    float* feat32 = nullptr;
    float* partial = nullptr;   // split-K partials
    float* bpartial = nullptr;  // bias partials
    float* loss_partial = nullptr;
    int ksplit = 32;
};

namespace {

constexpr long long OFF_W[6] = {0, 0, MLP_H * MLP_IN, MLP_H * MLP_IN + MLP_H * MLP_H,
                                MLP_H * MLP_IN + 2LL * MLP_H * MLP_H, MLP_H * MLP_IN + 3LL * MLP_H * MLP_H};
constexpr long long OFF_B[6] = {0, 0, MLP_H, 2 * MLP_H, 3 * MLP_H, 4 * MLP_H};
constexpr int DIN[6] = {0, MLP_IN, MLP_H, MLP_H, MLP_H, MLP_H};
constexpr int DOUT[6] = {0, MLP_H, MLP_H, MLP_H, MLP_H, MLP_NOUT};

#define HIP_CHECK(call)                                      \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) {                              \
            std::fprintf(stderr, "torchgpu mlp: %s: %s\n", #call, hipGetErrorString(e_)); \
            return TORCHGPU_ERR_HIP;                         \
        }                                                    \
    } while (0)

GemmArgs gemm(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk, long long sbn, int M,
              int N, int K, int mode) {
    GemmArgs a;
    std::memset(&a, 0, sizeof a);
    a.A = A; a.sam = sam; a.sak = sak;
    a.B = B; a.sbk = sbk; a.sbn = sbn;
    a.M = M; a.N = N; a.K = K;
    a.mode = mode;
    return a;
}

// forward of the training path: H0 = bf16(feat), Z/H per layer, y = Z5
int forward_train(torchgpu_mlp* m, const float* feat, int B) {
    HIP_CHECK(tg::launch_to_bf16(feat, m->H[0], (long long)B * MLP_IN, m->stream));
    for (int L = 1; L <= 5; ++L) {
        const int din = DIN[L], dout = DOUT[L];
        GemmArgs a = gemm(m->H[L - 1], din, 1, m->W16 + OFF_W[L], 1, din, B, dout, din,
                          L < 5 ? tg::GEMM_FWD_HIDDEN : tg::GEMM_FWD_OUT);
        a.bias = m->b + OFF_B[L];
        a.Z = m->Z[L];
        a.ldz = dout;
        if (L < 5) {
            a.H = m->H[L];
            a.ldh = dout;
        }
        HIP_CHECK(tg::launch_gemm(a, 1, m->stream));
    }
    return TORCHGPU_OK;
}

int backward(torchgpu_mlp* m, int B) {
    const int kchunk = ((B + m->ksplit - 1) / m->ksplit + 31) / 32 * 32;
    const int nz = (B + kchunk - 1) / kchunk;
    const int rowtiles = (B + 63) / 64;
    // output layer: gW5 = dy^T H4, gb5 = column sums of dy
    {
        GemmArgs a = gemm(m->dy16, 1, MLP_NOUT, m->H[4], MLP_H, 1, MLP_NOUT, MLP_H, B, tg::GEMM_SPLIT_K);
        a.kchunk = kchunk;
        a.partial = m->partial;
        HIP_CHECK(tg::launch_gemm(a, nz, m->stream));
        HIP_CHECK(tg::launch_reduce_partials(m->partial, nz, (long long)MLP_NOUT * MLP_H, m->gW + OFF_W[5], 0, m->stream));
        HIP_CHECK(tg::launch_colsum(m->dy32, B, MLP_NOUT, m->bpartial, m->stream));
        HIP_CHECK(tg::launch_reduce_partials(m->bpartial, (B + 255) / 256, MLP_NOUT, m->gb + OFF_B[5], 0, m->stream));
    }
    // dZ4 = bf16((dy W5) * SiLU'(Z4)); then down the hidden layers
    const uint16_t* up = m->dy16;
    int dup = MLP_NOUT;
    for (int L = 5; L >= 2; --L) {
        // G_{L-1} = dZ_L . W_L  (A = dZ_L [B][dout], B(k=o, n=i) = W_L[o][i])
        GemmArgs a = gemm(up, dup, 1, m->W16 + OFF_W[L], DIN[L], 1, B, DIN[L], dup, tg::GEMM_BWD_DATA);
        a.H = m->dZ[L - 1];
        a.ldh = DIN[L];
        a.Zprev = m->Z[L - 1];
        a.ldzp = DIN[L];
        a.partial = m->bpartial;
        HIP_CHECK(tg::launch_gemm(a, 1, m->stream));
        HIP_CHECK(tg::launch_reduce_partials(m->bpartial, rowtiles, DIN[L], m->gb + OFF_B[L - 1], 0, m->stream));
        // gW_{L-1} = dZ_{L-1}^T H_{L-2}
        const int L1 = L - 1;
        GemmArgs w = gemm(m->dZ[L1], 1, DOUT[L1], m->H[L1 - 1], DIN[L1], 1, DOUT[L1], DIN[L1], B, tg::GEMM_SPLIT_K);
        w.kchunk = kchunk;
        w.partial = m->partial;
        HIP_CHECK(tg::launch_gemm(w, nz, m->stream));
        HIP_CHECK(tg::launch_reduce_partials(m->partial, nz, (long long)DOUT[L1] * DIN[L1], m->gW + OFF_W[L1], 0,
                                             m->stream));
        up = m->dZ[L1];
        dup = DOUT[L1];
    }
    return TORCHGPU_OK;
}

int loss_grad(torchgpu_mlp* m, const float* feat, const float* target, int B, const float* w3, int err_head,
              float* loss) {
    if (B <= 0 || B > m->max_batch) return TORCHGPU_ERR_INVALID_ARG;
    int rc = forward_train(m, feat, B);
    if (rc) return rc;
    const float w0 = w3 ? w3[0] : 1.f, w1 = w3 ? w3[1] : 1.f, w2 = w3 ? w3[2] : 1.f;
    HIP_CHECK(tg::launch_loss_grad(m->Z[5], target, B, MLP_NOUT, w0, w1, w2, err_head ? 2 : -1, m->dy32, m->dy16,
                                   m->loss_partial, m->stream));
    rc = backward(m, B);
    if (rc) return rc;
    if (loss) {
        const int nb = (B + 255) / 256;
        std::vector<float> lp(nb);
        HIP_CHECK(hipMemcpyAsync(lp.data(), m->loss_partial, nb * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIP_CHECK(hipStreamSynchronize(m->stream));
        double s = 0;
        for (float v : lp) s += v;
        *loss = (float)(s / B);
    }
    return TORCHGPU_OK;
}

}  // namespace

extern "C" {

int torchgpu_mlp_create(int max_batch, void* stream, torchgpu_mlp** out) {
    if (!out || max_batch <= 0) return TORCHGPU_ERR_INVALID_ARG;
    *out = nullptr;
    torchgpu_mlp* m = new torchgpu_mlp();
    m->max_batch = max_batch;
    m->stream = (hipStream_t)stream;  // NULL: the null (default) stream
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        m->nsm = prop.multiProcessorCount;
    const long long Bm = max_batch;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, bytes) == hipSuccess; };
    alloc((void**)&m->W, MLP_NW * 4);
    alloc((void**)&m->b, MLP_NB * 4);
    alloc((void**)&m->W16, MLP_NW * 2);
    alloc((void**)&m->gW, MLP_NW * 4);
    alloc((void**)&m->gb, MLP_NB * 4);
    alloc((void**)&m->mW, MLP_NW * 4);
    alloc((void**)&m->vW, MLP_NW * 4);
    alloc((void**)&m->mb, MLP_NB * 4);
    alloc((void**)&m->vb, MLP_NB * 4);
    alloc((void**)&m->H[0], Bm * MLP_IN * 2);
    for (int L = 1; L <= 4; ++L) {
        alloc((void**)&m->H[L], Bm * MLP_H * 2);
        alloc((void**)&m->Z[L], Bm * MLP_H * 4);
        alloc((void**)&m->dZ[L], Bm * MLP_H * 2);
    }
    alloc((void**)&m->Z[5], Bm * MLP_NOUT * 4);
    alloc((void**)&m->dy32, Bm * MLP_NOUT * 4);
    alloc((void**)&m->dy16, Bm * MLP_NOUT * 2);
    alloc((void**)&m->partial, (size_t)m->ksplit * MLP_H * MLP_H * 4);
    alloc((void**)&m->bpartial, ((Bm + 63) / 64) * MLP_H * 4);
    alloc((void**)&m->loss_partial, ((Bm + 255) / 256) * 4);
    if (!ok) {
        torchgpu_mlp_destroy(m);
        return TORCHGPU_ERR_HIP;
    }
    std::vector<float> zw(MLP_NW, 0.f), zb(MLP_NB, 0.f);
    for (float* p : {m->mW, m->vW}) (void)hipMemcpy(p, zw.data(), MLP_NW * 4, hipMemcpyHostToDevice);
    for (float* p : {m->mb, m->vb}) (void)hipMemcpy(p, zb.data(), MLP_NB * 4, hipMemcpyHostToDevice);
    *out = m;
    return TORCHGPU_OK;
}

void torchgpu_mlp_destroy(torchgpu_mlp* m) {
    if (!m) return;
    (void)hipStreamSynchronize(m->stream);
    std::vector<void*> ptrs = {m->W, m->b, m->W16, m->gW, m->gb, m->mW, m->vW, m->mb, m->vb, m->Z[5], m->dy32,
                               m->dy16, m->feat32, m->partial, m->bpartial, m->loss_partial, m->H[0]};
    for (int L = 1; L <= 4; ++L) {
        ptrs.push_back(m->H[L]);
        ptrs.push_back(m->Z[L]);
        ptrs.push_back(m->dZ[L]);
    }
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
}

int torchgpu_mlp_set_params(torchgpu_mlp* m, const float* W, const float* b) {
    if (!m || !W || !b) return TORCHGPU_ERR_INVALID_ARG;
    HIP_CHECK(hipMemcpyAsync(m->W, W, MLP_NW * 4, hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipMemcpyAsync(m->b, b, MLP_NB * 4, hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(tg::launch_to_bf16(m->W, m->W16, MLP_NW, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    m->step = 0;
    return TORCHGPU_OK;
}

int torchgpu_mlp_get_params(torchgpu_mlp* m, float* W, float* b) {
    if (!m) return TORCHGPU_ERR_INVALID_ARG;
    if (W) HIP_CHECK(hipMemcpyAsync(W, m->W, MLP_NW * 4, hipMemcpyDeviceToHost, m->stream));
    if (b) HIP_CHECK(hipMemcpyAsync(b, m->b, MLP_NB * 4, hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    return TORCHGPU_OK;
}

int torchgpu_mlp_device_weights(torchgpu_mlp* m, const uint16_t** w16, const float** b) {
    if (!m || !w16 || !b) return TORCHGPU_ERR_INVALID_ARG;
    *w16 = m->W16;
    *b = m->b;
    return TORCHGPU_OK;
}

int torchgpu_mlp_forward(torchgpu_mlp* m, const float* feat, int64_t B, float* y) {
    if (!m || B < 0 || (B > 0 && (!feat || !y))) return TORCHGPU_ERR_INVALID_ARG;
    HIP_CHECK(tg::launch_mlp_infer(m->W16, m->b, feat, B, y, m->nsm, m->stream));
    return TORCHGPU_OK;
}

int torchgpu_mlp_loss_grad(torchgpu_mlp* m, const float* feat, const float* target, int B, const float* w3,
                           int err_head, float* loss, float* gW, float* gb) {
    if (!m || !feat || !target) return TORCHGPU_ERR_INVALID_ARG;
    int rc = loss_grad(m, feat, target, B, w3, err_head, loss);
    if (rc) return rc;
    if (gW) HIP_CHECK(hipMemcpyAsync(gW, m->gW, MLP_NW * 4, hipMemcpyDeviceToHost, m->stream));
    if (gb) HIP_CHECK(hipMemcpyAsync(gb, m->gb, MLP_NB * 4, hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    return TORCHGPU_OK;
}

int torchgpu_mlp_train_step(torchgpu_mlp* m, const float* feat, const float* target, int B, const float* w3,
                            int err_head, float lr, float weight_decay, float* loss) {
    if (!m || !feat || !target) return TORCHGPU_ERR_INVALID_ARG;
    int rc = loss_grad(m, feat, target, B, w3, err_head, loss);
    if (rc) return rc;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - (float)std::pow((double)b1, (double)m->step);
    const float bc2 = 1.f - (float)std::pow((double)b2, (double)m->step);
    HIP_CHECK(tg::launch_adam(m->W, m->gW, m->mW, m->vW, m->W16, MLP_NW, lr, b1, b2, eps, bc1, bc2, weight_decay,
                              m->stream));
    HIP_CHECK(tg::launch_adam(m->b, m->gb, m->mb, m->vb, nullptr, MLP_NB, lr, b1, b2, eps, bc1, bc2, 0.f, m->stream));
    return TORCHGPU_OK;
}

}  // extern "C"
