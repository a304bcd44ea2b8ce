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
// 32 x 32 block of 2 x 2 MFMA 16x16 tiles), K step 64 (two MFMA k-steps per
// staging).  The epilogue can add a per-row bias, keep the pre-activation and
// apply GELU, so a 1x1 channel map, its bias and its activation are one pass.  MFMA 16x16x32 bf16
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

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
__device__ __forceinline__ float gelu_t(float x, float* dgelu) {
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float u = k0 * fmaf(k1 * x * x, x, x);
    const float t = tanhf(u);
    if (dgelu) *dgelu = 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * k0 * fmaf(3.f * k1, x * x, 1.f);
    return 0.5f * x * (1.f + t);
}

// Stage an RT (rows) x BK (k) tile of an operand into T[row][k] as BF16.
// mode 1: unit stride along k (float4 loads along k); mode 2: unit stride
// along the rows (float4 loads along rows); 0: scalar loads, consecutive
// threads along k.  Vector modes are chosen on the host only when the
// extents and strides are multiples of 4 and the base is 16-byte aligned.
template <int RT, int BK>
__device__ __forceinline__ void stage_tile(__bf16 (*T)[BK + 8], const float* __restrict__ X, long r0, long R,
                                           long srow, long k0, long K, long sk, int mode, int tid) {
    static_assert((RT * BK) % 1024 == 0, "tile must split into float4 per thread");
    if (mode == 1) {
#pragma unroll
        for (int e = 0; e < RT * BK / 4 / 256; e++) {
            const int idx = tid + e * 256;
            const int r = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
            const long gr = r0 + r, gk = k0 + kq;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gr < R && gk < K) v = *reinterpret_cast<const float4*>(X + gr * srow + gk);
            bf16x4 h = {to_bf16(v.x), to_bf16(v.y), to_bf16(v.z), to_bf16(v.w)};
            *reinterpret_cast<bf16x4*>(&T[r][kq]) = h;
        }
    } else if (mode == 2) {
#pragma unroll
        for (int e = 0; e < RT * BK / 4 / 256; e++) {
            const int idx = tid + e * 256;
            const int kk = idx / (RT / 4), rq = (idx % (RT / 4)) * 4;
            const long gr = r0 + rq, gk = k0 + kk;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gr < R && gk < K) v = *reinterpret_cast<const float4*>(X + gr + gk * sk);
            T[rq + 0][kk] = to_bf16(v.x);
            T[rq + 1][kk] = to_bf16(v.y);
            T[rq + 2][kk] = to_bf16(v.z);
            T[rq + 3][kk] = to_bf16(v.w);
        }
    } else {
#pragma unroll
        for (int e = 0; e < RT * BK / 256; e++) {
            const int idx = tid + e * 256;
            const int r = idx / BK, kk = idx % BK;
            const long gr = r0 + r, gk = k0 + kk;
            T[r][kk] = to_bf16((gr < R && gk < K) ? X[gr * srow + gk * sk] : 0.f);
        }
    }
}

// BM x BN outputs per 256-thread workgroup; the 4 waves form WM x (4 / WM)
// and each holds (BM / WM) x (BN / (4 / WM)) outputs as TI x TJ MFMA 16x16
// tiles.  Shapes: 64 x 64 (general), 32 x 256 (few rows: 1x1 channel maps),
// 256 x 32 (few columns: DFTs onto the retained modes).
template <int BM, int BN, int BK, int WM>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    constexpr int WN = 4 / WM, WTM = BM / WM, WTN = BN / WN, TI = WTM / 16, TJ = WTN / 16;
    __shared__ __bf16 As[BM][BK + 8];
    __shared__ __bf16 Bs[BN][BK + 8];
    const long b = blockIdx.z;
    const long m0 = (long)blockIdx.y * BM, n0 = (long)blockIdx.x * BN;
    const float* A = g.A + b * g.sab;
    const float* B = g.B + b * g.sbb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WN) * WTM, wn = (wave % WN) * WTN;
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long k0 = 0; k0 < g.K; k0 += BK) {
        stage_tile<BM, BK>(As, A, m0, g.M, g.sam, k0, g.K, g.sak, g.avec, tid);
        stage_tile<BN, BK>(Bs, B, n0, g.N, g.sbn, k0, g.K, g.sbk, g.bvec, tid);
        __syncthreads();
        const int row = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            bf16x8 bv[TJ];
#pragma unroll
            for (int j = 0; j < TJ; j++) bv[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + row][ks + kq]);
#pragma unroll
            for (int i = 0; i < TI; i++) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + row][ks + kq]);
#pragma unroll
                for (int j = 0; j < TJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // epilogue: v = alpha AB + beta C + bias[m]; pre = v; C = act(v)
    float* C = g.C + b * g.scb;
    float* Pre = g.pre ? g.pre + b * g.scb : nullptr;
    const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long gm = m0 + wm + i * 16 + rq + q, gn = n0 + wn + j * 16 + col;
                if (gm < g.M && gn < g.N) {
                    const long o = gm * g.scm + gn * g.scn;
                    float v = g.alpha * acc[i][j][q];
                    if (g.beta != 0.f) v = fmaf(g.beta, C[o], v);
                    if (g.bias) v += g.bias[gm];
                    if (Pre) Pre[o] = v;
                    C[o] = g.act ? gelu_t(v, nullptr) : v;
                }
            }
}

// vector staging mode of an operand (see stage_tile)
static int vec_mode(const float* X, long rows, long srow, long K, long sk, long sbatch) {
    const bool aligned = ((uintptr_t)X & 15) == 0 && (sbatch & 3) == 0;
    if (aligned && sk == 1 && (K & 3) == 0 && (srow & 3) == 0) return 1;
    if (aligned && srow == 1 && (rows & 3) == 0 && (sk & 3) == 0) return 2;
    return 0;
}

template <int BM, int BN, int BK, int WM>
static hipError_t gemm_launch(const GemmArgs& g, hipStream_t st) {
    const long gy = (g.M + BM - 1) / BM, gx = (g.N + BN - 1) / BN;
    if (gy > 65535 || g.batch > 65535 || gx > 0x7FFFFFFFL) return hipErrorInvalidValue;
    dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)g.batch);
    k_gemm_bf16<BM, BN, BK, WM><<<grid, 256, 0, st>>>(g);
    return hipGetLastError();
}

hipError_t launch_gemm_bf16(const GemmArgs& g0, hipStream_t st) {
    if (g0.M <= 0 || g0.N <= 0 || g0.batch <= 0) return hipSuccess;
    GemmArgs g = g0;
    g.avec = vec_mode(g.A, g.M, g.sam, g.K, g.sak, g.sab);
    g.bvec = vec_mode(g.B, g.N, g.sbn, g.K, g.sbk, g.sbb);
    const bool k32 = g.K <= 32;
    if (g.M <= 32 && g.N > 64)
        return k32 ? gemm_launch<32, 256, 32, 1>(g, st) : gemm_launch<32, 256, 64, 1>(g, st);
    if (g.N <= 32 && g.M > 64)
        return k32 ? gemm_launch<256, 32, 32, 4>(g, st) : gemm_launch<256, 32, 64, 4>(g, st);
    return k32 ? gemm_launch<64, 64, 32, 2>(g, st) : gemm_launch<64, 64, 64, 2>(g, st);
}

// ------------------------------------------------------------------ elementwise

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

// partial[chunk][c] = sum of dy[s][c][p] over this workgroup's chunk of
// the S * P values of channel c (FP64); the chunks are summed by
// k_reduce_batches in a fixed order
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ dy, float* __restrict__ partial, long S,
                                                   long C, long P) {
    const long c = blockIdx.x, nchunk = gridDim.y, chunk = blockIdx.y;
    const long n = S * P, per = (n + nchunk - 1) / nchunk;
    const long e0 = chunk * per, e1 = min(n, e0 + per);
    double acc = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += 256) {
        const long s = e / P, p = e - s * P;
        acc += dy[(s * C + c) * P + p];
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[chunk * C + c] = (float)red[0];
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
// screens [D][L][N*N] (sampled layer screens), s[L] phase multipliers of the
// wavelength; for sample s0 + d, channel-major over Stot samples:
//   x[l][s][p]     = v_l / xs,  v_l = s_l screen_l            (phase, l < L)
//   x[L + l][s][p] = lap(v_l) / xl                            (5-point Laplacian,
//                                                              edge pixels clamped)
// and, when pha and amp are given, the smooth complex field perturbation
// u = amp e^{i (pha - Phi)}, Phi = sum_l v_l:
//   t[0][s][p] = (Re u - 1) / cs,  t[1][s][p] = Im u / ps.
__global__ void k_nn_pairs(const float* __restrict__ screens, const double* __restrict__ sc,
                           const float* __restrict__ pha, const float* __restrict__ amp, int L, int N, long Stot,
                           long s0, float* __restrict__ x, float* __restrict__ t, double xs, double xl, double cs,
                           double ps) {
    const int d = blockIdx.y;
    const long s = s0 + d, P = (long)N * N;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        const int j = (int)(p / N), i = (int)(p - (long)j * N);
        const long pl = j * (long)N + max(i - 1, 0), pr = j * (long)N + min(i + 1, N - 1);
        const long pu = max(j - 1, 0) * (long)N + i, pd = min(j + 1, N - 1) * (long)N + i;
        double phi = 0.0;
        for (int l = 0; l < L; l++) {
            const float* sl = screens + ((long)d * L + l) * P;
            const double v = sc[l] * (double)sl[p];
            const double lap = sc[l] * ((double)sl[pl] + (double)sl[pr] + (double)sl[pu] + (double)sl[pd]) - 4.0 * v;
            phi += v;
            x[((long)l * Stot + s) * P + p] = (float)(v / xs);
            x[((long)(L + l) * Stot + s) * P + p] = (float)(lap / xl);
        }
        if (t && pha && amp) {
            const double a = (double)amp[(long)d * P + p];
            const double dp = (double)pha[(long)d * P + p] - phi;
            t[s * P + p] = (float)((a * cos(dp) - 1.0) / cs);
            t[(Stot + s) * P + p] = (float)(a * sin(dp) / ps);
        }
    }
}

// network output y [2][S][P] -> u = (1 + cs y0) + i ps y1; amp = |u|,
// pha = xs sum_{l < L} x_l + arg u
__global__ void k_nn_unpack(const float* __restrict__ y, const float* __restrict__ x, int L, long S, long P,
                            double xs, double cs, double ps, float* __restrict__ amp, float* __restrict__ pha) {
    const long n = S * P;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        double phi = 0.0;
        for (int l = 0; l < L; l++) phi += (double)x[(long)l * n + e];
        const double ur = 1.0 + cs * (double)y[e], ui = ps * (double)y[n + e];
        amp[e] = (float)sqrt(ur * ur + ui * ui);
        pha[e] = (float)(xs * phi + atan2(ui, ur));
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
hipError_t launch_bias_grad(const float* dy, float* partial, long S, long C, long P, int nchunk, hipStream_t st) {
    k_bias_grad<<<dim3((unsigned)C, (unsigned)nchunk), 256, 0, st>>>(dy, partial, S, C, P);
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
                           long N, long Stot, long s0, float* x, float* t, double xs, double xl, double cs, double ps,
                           hipStream_t st) {
    long g = (N * N + 255) / 256;
    dim3 grid((unsigned)(g < 2048 ? g : 2048), (unsigned)D);
    k_nn_pairs<<<grid, 256, 0, st>>>(screens, sc, pha, amp, (int)L, (int)N, Stot, s0, x, t, xs, xl, cs, ps);
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
