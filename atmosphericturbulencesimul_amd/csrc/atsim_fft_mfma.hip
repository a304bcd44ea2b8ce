// atsim_fft_mfma.hip -- the mixed-precision variant of the 2048-point row
// transform: radix-16 Stockham stages as DFT-16 matrix products on the matrix
// cores (v_mfma_f32_16x16x32_f16, FP32 accumulation) with a hi + lo split of
// both operands, the radix-8 stage and all twiddles in FP32 on the vector
// ALUs.  No counterpart in the reference; it serves the precision / speed
// comparison of the two arithmetic paths (atsim_fft2d_variant).
//
// A radix-16 stage over the 128 vectors of a row: each thread twiddles its
// vector (16 complex points, j = t) and writes it to LDS as 32 reals [re 0..15,
// im 0..15]; each wave then takes groups of 16 vectors as the B operand
// (K = 32 reals x N = 16 vectors) of
//     [Yr; Yi] = [[Fr, -Fi], [Fi, Fr]] [Xr; Xi],   F[m][k] = e^{DIR 2 pi i m k / 16},
// as two 16-row blocks.  Each operand is split as x = hi + lo with hi = f16(x)
// and lo = f16(x - hi) after scaling every vector by a power of two that brings
// its largest |value| into [1, 2) (so lo stays a normal FP16 number), and a
// block is hi*hi + hi*lo + lo*hi: three MFMAs, the lo*lo term (2^-22 relative)
// dropped.  The products land in the Stockham output positions of the stage.
//
// MFMA 16x16x32 operand layout (one wave): lane l holds A[l % 16][8 (l / 16)
// .. +7] and B[8 (l / 16) .. +7][l % 16]; the accumulator holds
// D[4 (l / 16) .. +3][l % 16].
#include <hip/hip_runtime.h>
#include <math.h>

#include "atsim_fft.hpp"
#include "atsim_launch.hpp"

namespace atsim {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int MN = 2048;          // row length of the variant
constexpr int MT = MN / 16;       // threads (= DFT-16 vectors of a radix-16 stage) per row
constexpr int MROWS = 2;          // rows per workgroup
constexpr int VSTR = 36;          // floats per vector slot in LDS (16-byte aligned rows)

// the 16 complex values of vector j = t at a stage of radix 16 and span NS,
// times the stage's twiddles w^(r k), k = j mod NS (the same products as
// stage_compute, without its butterfly)
template <int NS, int DIR>
__device__ __forceinline__ void twiddle16(float2* v, int t, const float2* __restrict__ tw) {
    if constexpr (NS > 1) {
        const int k = t % NS;
        constexpr int STEP = MN / (NS * 16);
        float2 p[4];
#pragma unroll
        for (int e = 0; e < 4; e++) p[e] = twiddle<MN, DIR>(tw, (1 << e) * k * STEP);
#pragma unroll
        for (int r = 1; r < 16; r++) {
            float2 w = make_float2(1.f, 0.f);
            bool first = true;
#pragma unroll
            for (int e = 3; e >= 0; e--)
                if (r & (1 << e)) {
                    w = first ? p[e] : cmul(w, p[e]);
                    first = false;
                }
            v[r] = cmul(v[r], w);
        }
    }
}

// hi / lo halves of 8 scaled values
__device__ __forceinline__ void split8(const float* x, float s, f16x8& hi, f16x8& lo) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float y = x[i] * s;
        const _Float16 h = (_Float16)y;
        hi[i] = h;
        lo[i] = (_Float16)(y - (float)h);
    }
}

// One radix-16 stage of span NS on the matrix cores: vectors come from vec
// (row region, MT x VSTR floats), outputs go to the Stockham positions of xch
// (row region, padded).  Called by all threads of the workgroup between
// barriers; wave wr (0, 1) of the row takes groups 4 wr .. 4 wr + 3.
template <int NS>
__device__ __forceinline__ void mfma_stage(const float* vec, float2* xch, int wr, int lane, const f16x8& at_h,
                                           const f16x8& at_l, const f16x8& ab_h, const f16x8& ab_l) {
    const int n = lane & 15, k0 = (lane >> 4) * 8, m0 = (lane >> 4) * 4;
#pragma unroll
    for (int gi = 0; gi < 4; gi++) {
        const int j = (wr * 4 + gi) * 16 + n;
        const float4* src = reinterpret_cast<const float4*>(vec + j * VSTR + k0);
        const float4 a = src[0], b = src[1];
        const float x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        float mx = 0.f;
#pragma unroll
        for (int i = 0; i < 8; i++) mx = fmaxf(mx, fabsf(x[i]));
        mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        // power-of-two scale taking the vector's largest |value| into [1, 2)
        const float s = mx > 0.f ? ldexpf(1.f, -ilogbf(mx)) : 1.f;
        f16x8 bh, bl;
        split8(x, s, bh, bl);
        f32x4 top = {0.f, 0.f, 0.f, 0.f}, bot = {0.f, 0.f, 0.f, 0.f};
        top = __builtin_amdgcn_mfma_f32_16x16x32_f16(at_h, bh, top, 0, 0, 0);
        top = __builtin_amdgcn_mfma_f32_16x16x32_f16(at_h, bl, top, 0, 0, 0);
        top = __builtin_amdgcn_mfma_f32_16x16x32_f16(at_l, bh, top, 0, 0, 0);
        bot = __builtin_amdgcn_mfma_f32_16x16x32_f16(ab_h, bh, bot, 0, 0, 0);
        bot = __builtin_amdgcn_mfma_f32_16x16x32_f16(ab_h, bl, bot, 0, 0, 0);
        bot = __builtin_amdgcn_mfma_f32_16x16x32_f16(ab_l, bh, bot, 0, 0, 0);
        const float inv = 1.f / s;
        const int base = (j / NS) * NS * 16 + j % NS;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int pos = base + (m0 + q) * NS;
            xch[pos + (pos >> 4)] = make_float2(top[q] * inv, bot[q] * inv);
        }
    }
}

// write the twiddled vector of thread t to its LDS slot
__device__ __forceinline__ void put_vec(float* vec, int t, const float2* v) {
    float4* d = reinterpret_cast<float4*>(vec + t * VSTR);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        d[i] = make_float4(v[4 * i].x, v[4 * i + 1].x, v[4 * i + 2].x, v[4 * i + 3].x);
        d[4 + i] = make_float4(v[4 * i].y, v[4 * i + 1].y, v[4 * i + 2].y, v[4 * i + 3].y);
    }
}

}  // namespace

// Row transforms of length 2048 (sign DIR) with the radix-16 stages on the
// matrix cores; rows MROWS * blockIdx.x .. of data (nrows a multiple of MROWS).
template <int DIR>
__global__ void __launch_bounds__(MROWS * MT) k_row_fft_mfma(float2* __restrict__ data, const float2* __restrict__ tw) {
    using P = Plan<MN>;
    static_assert(P::S == 3 && P::R0 == 8, "plan [8, 16, 16] / [16, 16, 8]");
    __shared__ float2 xch_all[MROWS * P::LDS];
    __shared__ __attribute__((aligned(16))) float vec_all[MROWS * MT * VSTR];
    const int lr = threadIdx.x / MT, t = threadIdx.x % MT;
    const int lane = threadIdx.x & 63, wr = (threadIdx.x >> 6) & 1;
    float2* xch = xch_all + lr * P::LDS;
    float* vec = vec_all + lr * MT * VSTR;
    float2* rowp = data + ((size_t)blockIdx.x * MROWS + lr) * MN;
    // DFT-16 operands of this lane: A_top = [Fr | -Fi], A_bot = [Fi | Fr],
    // row m = lane % 16, columns k = 8 (lane / 16) .. +7 (k < 16: Xr, else Xi)
    f16x8 at_h, at_l, ab_h, ab_l;
    {
        const int m = lane & 15, k0 = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = k0 + i, kk = k & 15;
            const double a = 2.0 * M_PI * (double)((m * kk) & 15) / 16.0;
            const double fr = cos(a), fi = DIR * sin(a);
            const double top = k < 16 ? fr : -fi, bot = k < 16 ? fi : fr;
            const _Float16 th = (_Float16)top, bh = (_Float16)bot;
            at_h[i] = th;
            at_l[i] = (_Float16)(top - (double)th);
            ab_h[i] = bh;
            ab_l[i] = (_Float16)(bot - (double)bh);
        }
    }
    float2 v[16];
    if constexpr (DIR < 0) {
        // forward [8, 16, 16]: radix 8 on the vector ALUs
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = rowp[reg_index<MN, 8>(t, q)];
        stage_compute<MN, 8, 1, DIR>(v, t, tw);
        stage_store<MN, 8, 1, 1>(v, xch, t, 0);
        __syncthreads();
        stage_load<MN, 16, 1>(v, xch, t, 0);
        twiddle16<8, DIR>(v, t, tw);
        put_vec(vec, t, v);
        __syncthreads();
        mfma_stage<8>(vec, xch, wr, lane, at_h, at_l, ab_h, ab_l);
        __syncthreads();
        stage_load<MN, 16, 1>(v, xch, t, 0);
        twiddle16<128, DIR>(v, t, tw);
        put_vec(vec, t, v);
        __syncthreads();
        mfma_stage<128>(vec, xch, wr, lane, at_h, at_l, ab_h, ab_l);
        __syncthreads();
    } else {
        // inverse [16, 16, 8]
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = rowp[t + r * MT];
        put_vec(vec, t, v);
        __syncthreads();
        mfma_stage<1>(vec, xch, wr, lane, at_h, at_l, ab_h, ab_l);
        __syncthreads();
        stage_load<MN, 16, 1>(v, xch, t, 0);
        twiddle16<16, DIR>(v, t, tw);
        put_vec(vec, t, v);
        __syncthreads();
        mfma_stage<16>(vec, xch, wr, lane, at_h, at_l, ab_h, ab_l);
        __syncthreads();
        stage_load<MN, 8, 1>(v, xch, t, 0);
        stage_compute<MN, 8, 256, DIR>(v, t, tw);
        __syncthreads();
        stage_store<MN, 8, 256, 1>(v, xch, t, 0);
        __syncthreads();
    }
    // natural order in xch: coalesced copy out
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int p = t + i * MT;
        rowp[p] = xch[p + (p >> 4)];
    }
}

// out[b][x][y] = in[b][y][x] for n x n images, 32 x 32 tiles through LDS
__global__ void __launch_bounds__(256) k_transpose(const float2* __restrict__ in, float2* __restrict__ out, int n) {
    __shared__ float2 tile[32][33];
    const size_t img = (size_t)blockIdx.z * n * n;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) tile[k][tx] = in[img + (size_t)(y0 + k) * n + x0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8) out[img + (size_t)(x0 + k) * n + y0 + tx] = tile[tx][k];
}

bool fft_mfma_supported(long n) { return n == MN; }

hipError_t launch_row_fft_mfma(float2* data, long nrows, int dir, const float2* tw, hipStream_t st) {
    if (nrows % MROWS || nrows / MROWS > (1L << 31) - 1) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(nrows / MROWS);
    if (dir < 0) k_row_fft_mfma<-1><<<grid, MROWS * MT, 0, st>>>(data, tw);
    else k_row_fft_mfma<+1><<<grid, MROWS * MT, 0, st>>>(data, tw);
    return hipGetLastError();
}

hipError_t launch_transpose(const float2* in, float2* out, long n, long nbatch, hipStream_t st) {
    if (n % 32 || nbatch > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)(n / 32), (unsigned)(n / 32), (unsigned)nbatch);
    k_transpose<<<grid, 256, 0, st>>>(in, out, (int)n);
    return hipGetLastError();
}

}  // namespace atsim
