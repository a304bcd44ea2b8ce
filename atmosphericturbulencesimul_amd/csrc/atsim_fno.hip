// atsim_fno.hip -- kernels of the Fourier-neural-operator surrogate of the
// Fresnel layer loop (no counterpart in the reference).
//
// Every linear map of the network -- the truncated DFTs along x and y, the
// per-mode complex channel mixing, the 1x1 channel maps -- and every product
// of its backward pass is one strided, batched GEMM
//     C[b] = alpha * A[b] (M x K) * B[b] (K x N) + beta * C[b]
// with arbitrary element strides, so transposes, real/imaginary views and
// channel-major / mode-major layouts cost nothing.  Operands are FP32 in
// memory and rounded to BF16 (round to nearest even) when staged into LDS;
// the products run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Tile: 64 x 64 outputs per 256-thread workgroup (4 waves in 2 x 2, each a
// 32 x 32 block of 2 x 2 MFMA 16x16 tiles), K step 32.  MFMA 16x16x32 bf16
// operand layout (one wave): lane l holds A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16]; the accumulator holds D[4 (l / 16) .. +3][l % 16].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atsim_kernels.hpp"
#include "atsim_launch.hpp"

namespace atsim {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float x) {
    // round to nearest even on the bit pattern (finite inputs)
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const uint16_t h = (uint16_t)(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}

constexpr int GBM = 64, GBN = 64, GBK = 32;

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    // LDS tiles, K-contiguous for both operands: As[m][k], Bs[n][k] (+8 pad)
    __shared__ __bf16 As[GBM][GBK + 8];
    __shared__ __bf16 Bs[GBN][GBK + 8];
    const long b = blockIdx.z;
    const long m0 = (long)blockIdx.y * GBM, n0 = (long)blockIdx.x * GBN;
    const float* A = g.A + b * g.sab;
    const float* B = g.B + b * g.sbb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const bool a_kfast = g.sak == 1, b_kfast = g.sbk == 1;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long k0 = 0; k0 < g.K; k0 += GBK) {
        // stage 64 x 32 elements of each operand, 8 per thread; consecutive
        // threads walk the operand's unit-stride dimension when it has one
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int idx = tid + e * 256;
            int r = idx >> 5, kk = idx & 31;
            if (!a_kfast) { r = idx & 63; kk = idx >> 6; }
            long gm = m0 + r, gk = k0 + kk;
            const float va = (gm < g.M && gk < g.K) ? A[gm * g.sam + gk * g.sak] : 0.f;
            As[r][kk] = to_bf16(va);
            r = idx >> 5; kk = idx & 31;
            if (!b_kfast) { r = idx & 63; kk = idx >> 6; }
            const long gn = n0 + r;
            gk = k0 + kk;
            const float vb = (gn < g.N && gk < g.K) ? B[gk * g.sbk + gn * g.sbn] : 0.f;
            Bs[r][kk] = to_bf16(vb);
        }
        __syncthreads();
        const int row = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            bf16x8 a;
#pragma unroll
            for (int q = 0; q < 8; q++) a[q] = As[wm + i * 16 + row][kq + q];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                bf16x8 bv;
#pragma unroll
                for (int q = 0; q < 8; q++) bv[q] = Bs[wn + j * 16 + row][kq + q];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* C = g.C + b * g.scb;
    const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long gm = m0 + wm + i * 16 + rq + q, gn = n0 + wn + j * 16 + col;
                if (gm < g.M && gn < g.N) {
                    float* p = C + gm * g.scm + gn * g.scn;
                    const float v = g.alpha * acc[i][j][q];
                    *p = g.beta == 0.f ? v : fmaf(g.beta, *p, v);
                }
            }
}

hipError_t launch_gemm_bf16(const GemmArgs& g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return hipSuccess;
    const long gy = (g.M + GBM - 1) / GBM, gx = (g.N + GBN - 1) / GBN;
    if (gy > 65535 || g.batch > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)g.batch);
    k_gemm_bf16<<<grid, 256, 0, st>>>(g);
    return hipGetLastError();
}

// ------------------------------------------------------------------ elementwise
// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
__device__ __forceinline__ float gelu_t(float x, float* dgelu) {
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float u = k0 * fmaf(k1 * x * x, x, x);
    const float t = tanhf(u);
    if (dgelu) *dgelu = 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * k0 * fmaf(3.f * k1, x * x, 1.f);
    return 0.5f * x * (1.f + t);
}

// y = act(a + b + bias[channel]) over [S][C][P]; pre (optional) keeps a + b + bias
__global__ void k_add_bias_act(const float* __restrict__ a, const float* __restrict__ b,
                               const float* __restrict__ bias, float* __restrict__ pre, float* __restrict__ y,
                               long C, long P, long n, int gelu) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long c = (e / P) % C;
        float v = a[e] + (b ? b[e] : 0.f) + (bias ? bias[c] : 0.f);
        if (pre) pre[e] = v;
        y[e] = gelu ? gelu_t(v, nullptr) : v;
    }
}

// dx = dy * gelu'(pre)
__global__ void k_gelu_bwd(const float* __restrict__ pre, const float* __restrict__ dy, float* __restrict__ dx,
                           long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        float d;
        gelu_t(pre[e], &d);
        dx[e] = dy[e] * d;
    }
}

// db[c] (+)= sum over s, p of dy[s][c][p]; one workgroup per channel
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ dy, float* __restrict__ db, long S,
                                                   long C, long P, int accumulate) {
    const long c = blockIdx.x;
    double acc = 0.0;
    for (long s = 0; s < S; s++)
        for (long p = threadIdx.x; p < P; p += 256) acc += dy[(s * C + c) * P + p];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = (float)(accumulate ? db[c] + red[0] : red[0]);
}

// mean squared error over n elements: partial sums of (pred - target)^2 and
// grad = 2 (pred - target) / n
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                                             float* __restrict__ grad, long n, double* __restrict__ partial) {
    double acc = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256L) {
        const float d = pred[e] - target[e];
        acc += (double)d * d;
        if (grad) grad[e] = 2.f * d / (float)n;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction.  Parameters are kept in FP32;
// the GEMMs round them to BF16 when they stage them
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float ge = g[e];
        const float me = fmaf(b1, m[e] - ge, ge);
        const float ve = fmaf(b2, v[e] - ge * ge, ge * ge);
        m[e] = me;
        v[e] = ve;
        w[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
    }
}

// out[i] (+)= sum_b part[b n + i]: the split-K partial products of the
// weight gradients, summed in a fixed order
__global__ void k_reduce_batches(const float* __restrict__ part, float* __restrict__ out, long nb, long n,
                                 int accumulate) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float acc = accumulate ? out[i] : 0.f;
        for (long b = 0; b < nb; b++) acc += part[b * n + i];
        out[i] = acc;
    }
}

// Training / serving pairs of the surrogate from one step of the simulator.
// screens [D][L][P] (sampled layer screens), s[L] phase multipliers of the
// wavelength; for sample s0 + d:
//   x[l][s][p] = s_l screen_l / xs                    (network input)
//   t[0][s][p] = log(amp) / cs, t[1][s][p] = (pha - sum_l s_l screen_l) / ps
// (targets only when pha and amp are given), channel-major over Stot samples.
__global__ void k_nn_pairs(const float* __restrict__ screens, const double* __restrict__ sc,
                           const float* __restrict__ pha, const float* __restrict__ amp, int L, long P, long Stot,
                           long s0, float* __restrict__ x, float* __restrict__ t, double xs, double cs, double ps) {
    const int d = blockIdx.y;
    const long s = s0 + d;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        double phi = 0.0;
        for (int l = 0; l < L; l++) {
            const double v = sc[l] * (double)screens[((long)d * L + l) * P + p];
            phi += v;
            x[((long)l * Stot + s) * P + p] = (float)(v / xs);
        }
        if (t && pha && amp) {
            const float a = fmaxf(amp[(long)d * P + p], 1e-30f);
            t[s * P + p] = (float)(log((double)a) / cs);
            t[(Stot + s) * P + p] = (float)(((double)pha[(long)d * P + p] - phi) / ps);
        }
    }
}

// network output y [2][S][P] -> amp = exp(cs y0), pha = xs sum_l x_l + ps y1
__global__ void k_nn_unpack(const float* __restrict__ y, const float* __restrict__ x, int L, long S, long P,
                            double xs, double cs, double ps, float* __restrict__ amp, float* __restrict__ pha) {
    const long n = S * P;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        double phi = 0.0;
        for (int l = 0; l < L; l++) phi += (double)x[(long)l * n + e];
        amp[e] = (float)exp(cs * (double)y[e]);
        pha[e] = (float)(xs * phi + ps * (double)y[n + e]);
    }
}

static unsigned grid_for(long n) {
    long g = (n + 255) / 256;
    return (unsigned)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

hipError_t launch_add_bias_act(const float* a, const float* b, const float* bias, float* pre, float* y, long C,
                               long P, long n, int gelu, hipStream_t st) {
    k_add_bias_act<<<grid_for(n), 256, 0, st>>>(a, b, bias, pre, y, C, P, n, gelu);
    return hipGetLastError();
}
hipError_t launch_gelu_bwd(const float* pre, const float* dy, float* dx, long n, hipStream_t st) {
    k_gelu_bwd<<<grid_for(n), 256, 0, st>>>(pre, dy, dx, n);
    return hipGetLastError();
}
hipError_t launch_bias_grad(const float* dy, float* db, long S, long C, long P, int accumulate, hipStream_t st) {
    k_bias_grad<<<(unsigned)C, 256, 0, st>>>(dy, db, S, C, P, accumulate);
    return hipGetLastError();
}
hipError_t launch_mse(const float* pred, const float* target, float* grad, long n, double* partial, int nblk,
                      hipStream_t st) {
    k_mse<<<nblk, 256, 0, st>>>(pred, target, grad, n, partial);
    return hipGetLastError();
}
hipError_t launch_reduce_batches(const float* part, float* out, long nb, long n, int accumulate, hipStream_t st) {
    k_reduce_batches<<<grid_for(n), 256, 0, st>>>(part, out, nb, n, accumulate);
    return hipGetLastError();
}
hipError_t launch_nn_pairs(const float* screens, const double* sc, const float* pha, const float* amp, long D, long L,
                           long P, long Stot, long s0, float* x, float* t, double xs, double cs, double ps,
                           hipStream_t st) {
    long g = (P + 255) / 256;
    dim3 grid((unsigned)(g < 2048 ? g : 2048), (unsigned)D);
    k_nn_pairs<<<grid, 256, 0, st>>>(screens, sc, pha, amp, (int)L, P, Stot, s0, x, t, xs, cs, ps);
    return hipGetLastError();
}
hipError_t launch_nn_unpack(const float* y, const float* x, long L, long S, long P, double xs, double cs, double ps,
                            float* amp, float* pha, hipStream_t st) {
    k_nn_unpack<<<grid_for(S * P), 256, 0, st>>>(y, x, (int)L, S, P, xs, cs, ps, amp, pha);
    return hipGetLastError();
}
hipError_t launch_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st) {
    k_adam<<<grid_for(n), 256, 0, st>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace atsim
