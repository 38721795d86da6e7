// mlp.hip -- BF16 MFMA kernels of the cooling-surrogate MLP (gfx950).
//
// Network: 8 input features (6 used) -> 4 hidden layers of width 128 (SiLU)
// -> NOUT = 3 outputs (d log10 T, d HII fraction, log10 of the estimated
// relative T error).  Every product takes bfloat16 operands and accumulates
// in FP32 with v_mfma_f32_16x16x32_bf16.
//
// Training uses one tiled GEMM kernel (64x64x32 tiles, 4 waves, each wave a
// 2x2 block of 16x16 MFMA tiles) with arbitrary operand strides staged
// through LDS, and fused epilogues:
//   FWD_HIDDEN  Z = acc + b (FP32, kept for the backward pass), H = bf16(SiLU(Z))
//   FWD_OUT     y = acc + b
//   BWD_DATA    dZ_prev = bf16(acc * SiLU'(Z_prev)), bias-gradient partials
//   SPLIT_K     per-K-chunk partial products (weight gradients), reduced in
//               a fixed order by reduce_partials -> deterministic
// plus the loss/gradient kernel and Adam.
//
// Inference (mlp_infer_kernel) is one fused kernel: all weights in LDS, each
// wave carries 32 cells through the five layers with the activations kept in
// registers.  Layers are computed transposed (D = W . H^T, output features
// along MFMA rows, cells along columns) so that the FP32 accumulator of one
// layer, after SiLU and rounding, is directly the B operand of the next: the
// k-order it implies inside each 32-wide K step is matched by storing every
// weight matrix in LDS with its input columns permuted the same way.
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace tg {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef uint16_t bf16_t;

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu_f(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
// float -> bfloat16 bits, round to nearest even (NaN -> quiet NaN)
__device__ __forceinline__ bf16_t to_bf16(float x) {
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t x) { return __uint_as_float((unsigned int)x << 16); }
// Inference activation: SiLU z / (1 + e^-z) from the hardware base-2
// exponential and reciprocal (about 1 ulp each, where the training kernels
// use the correctly rounded expf and division), rounded to bfloat16 by the
// hardware conversion (round to nearest even, as to_bf16).  The result
// keeps 8 significant bits, so the two forms differ at most in its last bit.
__device__ __forceinline__ bf16_t silu_bf16_fast(float z) {
    const float e = __builtin_amdgcn_exp2f(-1.4426950408889634f * z);
    const float a = z * __builtin_amdgcn_rcpf(1.0f + e);
    return __builtin_bit_cast(bf16_t, (__bf16)a);
}
__device__ __forceinline__ bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

// ---------------------------------------------------------------------------
// Generic GEMM: C[M][N] = A[M][K] . B[K][N]
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs a) {
    __shared__ __attribute__((aligned(16))) bf16_t As[BM][LDK];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[BN][LDK];
    __shared__ float red[4][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    int kbeg = 0, kend = a.K;
    if (a.mode == GEMM_SPLIT_K) {
        kbeg = blockIdx.z * a.kchunk;
        kend = min(a.K, kbeg + a.kchunk);
    }
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // global -> LDS mapping: 8 consecutive k per thread
    const bool a_kfast = (a.sak == 1);
    const int am = a_kfast ? (tid >> 2) : (tid & 63);
    const int akc = a_kfast ? (tid & 3) : (tid >> 6);
    const bool b_kfast = (a.sbk == 1);
    const int bn = b_kfast ? (tid >> 2) : (tid & 63);
    const int bkc = b_kfast ? (tid & 3) : (tid >> 6);

    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        {
            const int m = m0 + am;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + akc * 8 + j;
                As[am][akc * 8 + j] = (m < a.M && k < kend) ? a.A[(long long)m * a.sam + (long long)k * a.sak] : (bf16_t)0;
            }
            const int n = n0 + bn;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + bkc * 8 + j;
                Bs[bn][bkc * 8 + j] = (n < a.N && k < kend) ? a.B[(long long)k * a.sbk + (long long)n * a.sbn] : (bf16_t)0;
            }
        }
        __syncthreads();
        const int q = lane >> 4, r = lane & 15;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 av = as_bf16x8(*reinterpret_cast<const u16x8*>(&As[(wave >> 1) * 32 + i * 16 + r][q * 8]));
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8 bv = as_bf16x8(*reinterpret_cast<const u16x8*>(&Bs[(wave & 1) * 32 + j * 16 + r][q * 8]));
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // epilogue: lane holds C[row 4q+e][col r] of each 16x16 tile
    const int q = lane >> 4, r = lane & 15;
    float colsum[2] = {0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + (wave >> 1) * 32 + i * 16 + q * 4 + e;
                const int n = n0 + (wave & 1) * 32 + j * 16 + r;
                if (m >= a.M || n >= a.N) continue;
                const float v = acc[i][j][e];
                switch (a.mode) {
                case GEMM_FWD_HIDDEN: {
                    const float z = v + a.bias[n];
                    a.Z[(long long)m * a.ldz + n] = z;
                    a.H[(long long)m * a.ldh + n] = to_bf16(silu_f(z));
                    break;
                }
                case GEMM_FWD_OUT:
                    a.Z[(long long)m * a.ldz + n] = v + a.bias[n];
                    break;
                case GEMM_BWD_DATA: {
                    const float dz = v * dsilu_f(a.Zprev[(long long)m * a.ldzp + n]);
                    a.H[(long long)m * a.ldh + n] = to_bf16(dz);
                    colsum[j] += dz;
                    break;
                }
                default:  // GEMM_SPLIT_K
                    a.partial[((long long)blockIdx.z * a.M + m) * a.N + n] = v;
                }
            }
    if (a.mode == GEMM_BWD_DATA) {
        // bias-gradient partial of this row tile: fixed-order sum over rows
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            float s = colsum[j];
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (q == 0) red[wave][(wave & 1) * 32 + j * 16 + r] = s;
        }
        __syncthreads();
        if (tid < BN && n0 + tid < a.N) {
            const int half = tid >> 5;  // waves {half, half + 2} own columns [32 half, 32 half + 32)
            a.partial[(long long)blockIdx.x * a.N + n0 + tid] = red[half][tid] + red[half + 2][tid];
        }
    }
}

// out[i] = sum_z partial[z][i] (z ascending), i < n
__global__ void reduce_partials_kernel(const float* __restrict__ partial, int nz, long long n, float* __restrict__ out,
                                       int accumulate) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < nz; ++z) s += partial[(long long)z * n + i];
        out[i] = accumulate ? out[i] + s : s;
    }
}

// Loss 0.5/B sum m_o (y - t)^2; dy = m (y - t) / B.  When err_head >= 0 the
// target of output err_head is log10(|y0 - t0| * ln10 + 1e-5) of the current
// prediction (no gradient through it).
__global__ void loss_grad_kernel(const float* __restrict__ y, const float* __restrict__ t, int B, int nout,
                                 float m0, float m1, float m2, int err_head, float* __restrict__ dy32,
                                 bf16_t* __restrict__ dy16, float* __restrict__ loss_partial) {
    __shared__ float red[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (b < B) {
        const float ms[3] = {m0, m1, m2};
        for (int o = 0; o < nout; ++o) {
            float target = t[(long long)b * nout + o];
            if (o == err_head) target = log10f(fabsf(y[(long long)b * nout] - t[(long long)b * nout]) * 2.302585093f + 1e-5f);
            const float d = y[(long long)b * nout + o] - target;
            const float mm = o < 3 ? ms[o] : 1.f;
            l += 0.5f * mm * d * d;
            const float g = mm * d / (float)B;
            dy32[(long long)b * nout + o] = g;
            dy16[(long long)b * nout + o] = to_bf16(g);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// column sums of a [B][n] fp32 matrix into partial[rowtile][n] (rowtile = 256 rows)
__global__ void colsum_kernel(const float* __restrict__ x, int B, int n, float* __restrict__ partial) {
    const int c = threadIdx.x;
    if (c >= n) return;
    float s = 0.f;
    const int b0 = blockIdx.x * 256, b1 = min(B, b0 + 256);
    for (int b = b0; b < b1; ++b) s += x[(long long)b * n + c];
    partial[(long long)blockIdx.x * n + c] = s;
}

// Adam on fp32 master parameters; writes the bf16 copy used by the GEMMs.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, bf16_t* __restrict__ p16, long long n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        p[i] -= upd;
        if (p16) p16[i] = to_bf16(p[i]);
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, bf16_t* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}

// ---------------------------------------------------------------------------
// Fused inference
// ---------------------------------------------------------------------------
// k-permutation of one 32-wide K step: lane group q holds, after a layer's
// 16x16 D tiles 2s and 2s+1, the features 32s + 4q + e (e < 4) and
// 32s + 16 + 4q + (e - 4) (e >= 4) as its 8 B-operand entries.
__device__ __forceinline__ int kperm(int q, int e) { return e < 4 ? 4 * q + e : 16 + 4 * q + (e - 4); }

struct InferLDS {
    bf16_t w1[MLP_H][8];              // layer 1, plain columns (features 0..7)
    bf16_t wh[3][MLP_H][MLP_H];       // layers 2-4, columns permuted per K step
    bf16_t wo[16][MLP_H];             // output layer (rows >= NOUT zero), permuted
    float b[4][MLP_H];
    float bo[16];
};

// Stores W[out][in] (fp32 master, row-major) into LDS with columns permuted:
// dst[o][32 s + 8 q + e] = W[o][32 s + kperm(q, e)].
__device__ __forceinline__ void load_perm(bf16_t (*dst)[MLP_H], const bf16_t* __restrict__ src, int rows, int rows_pad,
                                          int tid, int nthreads) {
    for (int idx = tid; idx < rows_pad * MLP_H; idx += nthreads) {
        const int o = idx / MLP_H, c = idx % MLP_H;
        const int s = c >> 5, q = (c >> 3) & 3, e = c & 7;
        const int srcc = 32 * s + kperm(q, e);
        dst[o][c] = o < rows ? src[o * MLP_H + srcc] : (bf16_t)0;
    }
}

// One hidden layer for NT cell tiles: acc[mt][nt] = W . Hin (+ bias), then
// SiLU -> bf16 into hout (the next B operands).
template <int NT>
__device__ __forceinline__ void hidden_layer(const bf16_t (*w)[MLP_H], const float* bias, const bf16x8 (&hin)[4][NT],
                                             bf16x8 (&hout)[4][NT], int lane) {
    const int q = lane >> 4, r = lane & 15;
#pragma unroll
    for (int s = 0; s < 4; ++s) {      // output K steps of the NEXT layer = pairs of M tiles here
        f32x4 acc[2][NT];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[h][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int mt = 2 * s + h;
                const bf16x8 av = as_bf16x8(*reinterpret_cast<const u16x8*>(&w[mt * 16 + r][ks * 32 + q * 8]));
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[h][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, hin[ks][nt], acc[h][nt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            u16x8 o;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                o[e] = silu_bf16_fast(acc[0][nt][e] + bias[(2 * s) * 16 + 4 * q + e]);
                o[4 + e] = silu_bf16_fast(acc[1][nt][e] + bias[(2 * s + 1) * 16 + 4 * q + e]);
            }
            hout[s][nt] = as_bf16x8(o);
        }
    }
}

// features: [ncells][8] fp32 (already normalised); y: [ncells][NOUT] fp32.
// cells: optional index list (features/outputs addressed through it).
constexpr int INFER_THREADS = 512;

__global__ __launch_bounds__(INFER_THREADS) void mlp_infer_kernel(const bf16_t* __restrict__ wts, const float* __restrict__ bias,
                                                       const float* __restrict__ feat, long long n,
                                                       const unsigned long long* __restrict__ n_dev,
                                                       float* __restrict__ y) {
    __shared__ InferLDS s;
    if (n_dev) n = (long long)*n_dev;
    // a block with no cells (the grid is sized for the largest queue) leaves
    // before staging the weights
    if ((long long)blockIdx.x * (INFER_THREADS / 64) * 32 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // weight blob layout (bf16): W1[128][8] | W2 | W3 | W4 [128][128] | W5[NOUT][128]
    for (int i = tid; i < MLP_H * 8; i += INFER_THREADS) s.w1[i / 8][i % 8] = wts[i];
    for (int l = 0; l < 3; ++l)
        load_perm(s.wh[l], wts + MLP_H * 8 + l * MLP_H * MLP_H, MLP_H, MLP_H, tid, INFER_THREADS);
    load_perm(s.wo, wts + MLP_H * 8 + 3 * MLP_H * MLP_H, MLP_NOUT, 16, tid, INFER_THREADS);
    for (int i = tid; i < 4 * MLP_H; i += INFER_THREADS) s.b[i / MLP_H][i % MLP_H] = bias[i];
    if (tid < 16) s.bo[tid] = tid < MLP_NOUT ? bias[4 * MLP_H + tid] : 0.f;
    __syncthreads();

    constexpr int NT = 2;  // 32 cells per wave per iteration
    const int q = lane >> 4, r = lane & 15;
    const long long ntiles = (n + 16 * NT - 1) / (16 * NT);
    constexpr int NW = INFER_THREADS / 64;
    for (long long tile = (long long)blockIdx.x * NW + wave; tile < ntiles; tile += (long long)gridDim.x * NW) {
        // keeps the (loop-invariant) LDS weight fragments from being hoisted
        // out of the tile loop into registers
        asm volatile("" ::: "memory");
        const long long c0 = tile * 16 * NT;
        // layer 1 B operand: lane group 0 carries the 8 features of cell r
        bf16x8 x1[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const long long c = c0 + nt * 16 + r;
            u16x8 xv;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float v = 0.f;
                if (q == 0 && c < n) v = feat[c * 8 + e];
                xv[e] = to_bf16(v);
            }
            x1[nt] = as_bf16x8(xv);
        }
        bf16x8 ha[4][NT], hb[4][NT];
        // layer 1: 8 M tiles, one K step (only k < 8 non-zero)
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
            f32x4 acc[2][NT];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int mt = 2 * s2 + h;
                u16x8 aw;
#pragma unroll
                for (int e = 0; e < 8; ++e) aw[e] = q == 0 ? s.w1[mt * 16 + r][e] : (bf16_t)0;
                const bf16x8 av = as_bf16x8(aw);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[h][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, x1[nt], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                u16x8 o;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    o[e] = silu_bf16_fast(acc[0][nt][e] + s.b[0][(2 * s2) * 16 + 4 * q + e]);
                    o[4 + e] = silu_bf16_fast(acc[1][nt][e] + s.b[0][(2 * s2 + 1) * 16 + 4 * q + e]);
                }
                ha[s2][nt] = as_bf16x8(o);
            }
        }
        hidden_layer<NT>(s.wh[0], s.b[1], ha, hb, lane);
        hidden_layer<NT>(s.wh[1], s.b[2], hb, ha, lane);
        hidden_layer<NT>(s.wh[2], s.b[3], ha, hb, lane);
        // output layer: one M tile (rows 0..15, NOUT used)
        f32x4 acc[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8 av = as_bf16x8(*reinterpret_cast<const u16x8*>(&s.wo[r][ks * 32 + q * 8]));
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, hb[ks][nt], acc[nt], 0, 0, 0);
        }
        if (q == 0) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const long long c = c0 + nt * 16 + r;
                if (c < n) {
#pragma unroll
                    for (int e = 0; e < MLP_NOUT; ++e) y[c * MLP_NOUT + e] = acc[nt][e] + s.bo[e];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_gemm(const GemmArgs& a, int ksplit, hipStream_t stream) {
    dim3 grid((a.M + BM - 1) / BM, (a.N + BN - 1) / BN, a.mode == GEMM_SPLIT_K ? ksplit : 1);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* partial, int nz, long long n, float* out, int accumulate,
                                  hipStream_t stream) {
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)nb), dim3(256), 0, stream, partial, nz, n, out, accumulate);
    return hipGetLastError();
}

hipError_t launch_loss_grad(const float* y, const float* t, int B, int nout, float m0, float m1, float m2, int err_head,
                            float* dy32, uint16_t* dy16, float* loss_partial, hipStream_t stream) {
    hipLaunchKernelGGL(loss_grad_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, y, t, B, nout, m0, m1, m2,
                       err_head, dy32, dy16, loss_partial);
    return hipGetLastError();
}

hipError_t launch_colsum(const float* x, int B, int n, float* partial, hipStream_t stream) {
    hipLaunchKernelGGL(colsum_kernel, dim3((B + 255) / 256), dim3(64), 0, stream, x, B, n, partial);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint16_t* p16, long long n, float lr,
                       float b1, float b2, float eps, float bc1, float bc2, float wd, hipStream_t stream) {
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)nb), dim3(256), 0, stream, p, g, m, v, p16, n, lr, b1, b2, eps,
                       bc1, bc2, wd);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, uint16_t* y, long long n, hipStream_t stream) {
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)nb), dim3(256), 0, stream, x, y, n);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const uint16_t* wts, const float* bias, const float* feat, long long n, float* y,
                            int nblocks, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    long long ntiles = (n + 31) / 32;
    long long nb = (ntiles + INFER_THREADS / 64 - 1) / (INFER_THREADS / 64);
    if (nb > nblocks) nb = nblocks;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)nb), dim3(INFER_THREADS), 0, stream, wts, bias, feat, n,
                       (const unsigned long long*)nullptr, y);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_dev(const uint16_t* wts, const float* bias, const float* feat,
                                const unsigned long long* n_dev, long long n_max, float* y, int nblocks,
                                hipStream_t stream) {
    if (n_max <= 0) return hipSuccess;
    long long ntiles = (n_max + 31) / 32;
    long long nb = (ntiles + INFER_THREADS / 64 - 1) / (INFER_THREADS / 64);
    if (nb > nblocks) nb = nblocks;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)nb), dim3(INFER_THREADS), 0, stream, wts, bias, feat, n_max,
                       n_dev, y);
    return hipGetLastError();
}

}  // namespace tg
