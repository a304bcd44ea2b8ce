// atsim_kernels.hip -- CDNA4 (gfx950) kernels of the wavefront path.
//
//  k_row / k_col   LDS Stockham FFT passes along x / y.  OP_FRESNEL fuses
//                  FFT -> transfer-function multiply -> IFFT in one pass
//                  (the separable form of Fresnel_propagate_wavefront,
//                  WFpropagate.c:68-152); k_row can also multiply the field by
//                  the screen phase of a group of layers first, sampling the
//                  master screens itself (AtmosphericTurbulence.c:2759-3412).
//  k_row_fresnel   the frame's x pass over several wavelengths of one row:
//                  optionally the x factor of the gap before a layer group,
//                  the group's screen, the x factor of the gap after it.
//  k_col_screen(_split)  the frame's y pass: y factor, screen (from sampled
//                  phase planes), y factor.  With the x and y factors of a
//                  gap commuting, gaps alternate x-first / y-first and a
//                  frame makes one pass over the fields per gap.
//  k_layer_phase   FP64 phase planes of the layers column passes apply.
//  k_permut        quadrant swap (fft.c:570-792).
//  k_noise         Philox4x32-10 + Box-Muller complex Gaussian noise.
//  k_synth_filter  von Karman spectral filter on the noise, written in the
//                  quadrant-swapped layout (AtmosphericTurbulence.c:643-784).
//  k_lag_sums, k_strfn  FP64 structure-function reductions
//                  (AT.c:790-847, 2246-2270).
//  k_extract       crop, amplitude, phase, geometric-phase unwrap
//                  (AT.c:3418-3476), block sums of the science-lambda rule
//                  (AT.c:3523-3625).
//  k_block_fix / k_block_apply  the +-2 pi block consistency heuristic and
//                  the final unwrap (AT.c:3627-4262).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "atsim_fft.hpp"
#include "atsim_kernels.hpp"
#include "atsim_launch.hpp"

namespace atsim {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int wrap_m(int x, int M) {
    x += (x < 0) ? M : 0;
    x -= (x >= M) ? M : 0;
    x -= (x >= M) ? M : 0;
    return x;
}

// x mod M for any int x (staging loops)
__device__ __forceinline__ int mod_m(int x, int M) {
    int r = x % M;
    return r < 0 ? r + M : r;
}

// separable 4 x 4 interpolation of layer l's master around (x0 + i, y0 + j), FP64
__device__ __forceinline__ double sample_layer(const float* __restrict__ m, int M, const LayerSample& p,
                                               int i, int j) {
    const int xb = p.x0 + i - 1, yb = p.y0 + j - 1;
    const int x0 = wrap_m(xb, M), x1 = wrap_m(xb + 1, M), x2 = wrap_m(xb + 2, M), x3 = wrap_m(xb + 3, M);
    double v = 0.0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const float* r = m + (size_t)wrap_m(yb + b, M) * M;
        double h = p.wx[0] * (double)r[x0];
        h = fma(p.wx[1], (double)r[x1], h);
        h = fma(p.wx[2], (double)r[x2], h);
        h = fma(p.wx[3], (double)r[x3], h);
        v = fma(p.wy[b], h, v);
    }
    return v;
}

// e^{i ph} for an FP64 phase.  The phase is reduced in FP64 to a quarter
// turn q and a remainder t = ph/2pi - q/4 in [-1/8, 1/8] revolutions (so the
// float remainder carries ~2e-8 rad of rounding, not the ~2e-7 of an angle
// near pi); sin(2 pi t) and cos(2 pi t) are their Taylor series through
// t^9 / t^10 (truncation below 3e-8 on |2 pi t| <= pi/4), evaluated in
// FP32 Horner form, then rotated by q quarter turns.
__device__ __forceinline__ float2 expi_d(double ph) {
    const double rr = ph * 0.15915494309189533577;  // revolutions
    const double q4 = rint(rr * 4.0);
    const float t = (float)fma(q4, -0.25, rr);
    const int qi = (int)q4 & 3;  // quarter turns mod 4 (two's complement)
    const float t2 = t * t;
    float s = fmaf(t2, 42.058693944897634f, -76.70585975306136f);
    s = fmaf(t2, s, 81.60524927607504f);
    s = fmaf(t2, s, -41.341702240399755f);
    s = fmaf(t2, s, 6.283185307179586f);
    s *= t;
    float c = fmaf(t2, -26.426256783374388f, 60.24464137187664f);
    c = fmaf(t2, c, -85.45681720669371f);
    c = fmaf(t2, c, 64.93939402266828f);
    c = fmaf(t2, c, -19.739208802178716f);
    c = fmaf(t2, c, 1.0f);
    // (cos, sin)(a + qi pi/2)
    float sr = (qi & 1) ? c : s;
    float cr = (qi & 1) ? s : c;
    if (qi & 2) sr = -sr;
    if ((qi + 1) & 2) cr = -cr;
    return make_float2(cr, sr);
}

// e^{i (pi/2) qt} for an FP64 phase in quarter turns (the row pass keeps its
// sampled screen pre-scaled by 2/pi, so a wavelength's phase costs one
// product): qt = n + t with n = rint(qt), |t| <= 1/2; sin and cos of
// (pi/2) t are degree-7 / degree-8 polynomials in t (least-squares fits of
// the absolute error on |t| <= 1/2, written here; 1.1e-7 and 8.6e-8 in FP32
// Horner form), then rotated by n quarter turns.
__device__ __forceinline__ float2 expi_q(double qt) {
    const double q4 = rint(qt);
    const float t = (float)(qt - q4);
    const int qi = (int)q4 & 3;  // quarter turns mod 4 (two's complement)
    const float t2 = t * t;
    float s = fmaf(t2, -0.004592288751155138f, 0.07967590540647507f);
    s = fmaf(t2, s, -0.6459629535675049f);
    s = fmaf(t2, s, 1.570796251296997f);
    s *= t;
    float c = fmaf(t2, 0.0009036321425810456f, -0.02086007222533226f);
    c = fmaf(t2, c, 0.25366920232772827f);
    c = fmaf(t2, c, -1.2337005138397217f);
    c = fmaf(t2, c, 1.0f);
    float sr = (qi & 1) ? c : s;
    float cr = (qi & 1) ? s : c;
    if (qi & 2) sr = -sr;
    if ((qi + 1) & 2) cr = -cr;
    return make_float2(cr, sr);
}

__device__ __forceinline__ double screen_phase(const ScreenArgs& sa, int d, int w, int i, int j) {
    double ph = 0.0;
    for (int q = 0; q < sa.lcount; q++) {
        const int l = sa.lfirst + q;
        const LayerSample p = sa.ls[d * sa.L + l];
        const float* m = sa.masters + (size_t)l * sa.M * sa.M;
        ph = fma(sa.s[w * sa.L + l], sample_layer(m, (int)sa.M, p, i, j), ph);
    }
    return ph;
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx b runs on
// XCD b % 8).  xcd_swizzle maps b to a tile index so that 4 consecutive
// tiles run on the same XCD at about the same time and share its L2 (the
// 4 column tiles of one 128-byte line, or 4 adjacent rows that read the same
// master rows).  Identity unless total is a multiple of 32.
__device__ __forceinline__ int xcd_swizzle(int b, int total) {
    if (total & 31) return b;
    const int xcd = b & 7, r = b >> 3;
    return ((((r >> 2) << 3) | xcd) << 2) | (r & 3);
}
// the same with chunks of K consecutive indices per XCD (K a power of two)
template <int K>
__device__ __forceinline__ int xcd_chunk(int b, int total) {
    if (total % (8 * K)) return b;
    const int xcd = b & 7, r = b >> 3;
    return ((r / K) * 8 + xcd) * K + r % K;
}
// each XCD takes one contiguous eighth of the launch (identity unless total
// is a multiple of 8)
__device__ __forceinline__ int xcd_slab(int b, int total) {
    if (total & 7) return b;
    return (b & 7) * (total >> 3) + (b >> 3);
}

// Raw buffer access (vector memory): a resource over [p, p + bytes), a
// per-lane byte offset and a scalar byte offset.  Accesses past `bytes`
// read zero and drop the store.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    const u32x2 x = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return make_float2(__uint_as_float(x.x), __uint_as_float(x.y));
}
__device__ __forceinline__ double buf_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    const u32x2 x = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return __hiloint2double((int)x.y, (int)x.x);
}
__device__ __forceinline__ void buf_store2(float2 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    u32x2 x;
    x.x = __float_as_uint(v.x);
    x.y = __float_as_uint(v.y);
    __builtin_amdgcn_raw_buffer_store_b64(x, r, voff, soff, 0);
}

// the value of x, opaque to the optimiser: index arithmetic derived from it
// is recomputed instead of kept live from an earlier use of x (an empty asm
// statement with no instructions)
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// most layers a frame can hold (per-frame layer parameters are staged in LDS)
constexpr int MAX_LAYERS = 64;

// (long)(x + 1000.5) - 1000, the reference's rounding of unwrap counts: the
// C conversion truncates toward zero, which trunc() does exactly in FP64
// (no 64-bit integer conversion)
__device__ __forceinline__ double unwrap_count(double x) { return trunc(x + 1000.5) - 1000.0; }

// Raise a kernel's dynamic LDS limit to what the 160 KB of a gfx950 CU leaves
// beside its static LDS.  Errors are consumed here so that they do not
// surface from the next hipGetLastError.
static void allow_dynamic_lds(const void* kernel) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, kernel) == hipSuccess) {
        const int room = 160 * 1024 - (int)fa.sharedSizeBytes;
        if (room > 0) (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, room);
    }
    (void)hipGetLastError();
}


// Master-row staging.  A workgroup samples a layer over one output row
// segment: it needs the 4 master rows ys..ys+3 and the columns xs-1 .. of its
// pixels, all taken modulo M.  The row bases are uniform (scalar registers);
// each lane loads whole columns, so a staged element costs one load, one LDS
// write and no per-element index arithmetic.

// x mod M for 0 <= x < 2M (the second test only fires for masters narrower
// than a staged segment)
__device__ __forceinline__ int wrap_col(int x, int M) {
    x = x < M ? x : x - M;
    return x < M ? x : x % M;
}

// the 4 rows of column x (already wrapped) of layer m, rows ys.. (uniform)
__device__ __forceinline__ void load_col4(float (&t)[4], const float* __restrict__ m, int M, int ys, int x) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
        int y = ys + b;
        y = y < M ? y : y - M;
        t[b] = m[(size_t)y * M + x];
    }
}

// columns [0, NC) of the 4 rows of one layer into S[b * stride + c];
// UNR column slots per lane are loaded before any is written
template <int UNR>
__device__ __forceinline__ void stage_rows4(float* S, int stride, const float* __restrict__ m, int M, int ys,
                                            int xs, int NC, int tid, int nthr) {
    ys = __builtin_amdgcn_readfirstlane(ys);
    xs = __builtin_amdgcn_readfirstlane(xs);
    for (int c0 = tid; c0 < NC; c0 += UNR * nthr) {
        float t[UNR][4];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int c = c0 + u * nthr;
            if (c < NC) load_col4(t[u], m, M, ys, wrap_col(xs + c, M));
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int c = c0 + u * nthr;
            if (c < NC) {
#pragma unroll
                for (int b = 0; b < 4; b++) S[b * stride + c] = t[u][b];
            }
        }
    }
}

// tail columns [NC0, NC0 + nt) of the 4 rows of nl layers (row bases ysq[q],
// column bases xsq[q]), one element per lane: S[(q * 4 + b) * stride + c]
__device__ __forceinline__ void stage_tail(float* S, int stride, const float* __restrict__ masters,
                                           size_t layer_stride, int M, const int* ysq, const int* xsq, int nl,
                                           int NC0, int nt, int tid, int nthr) {
    for (int e = tid; e < nl * 4 * nt; e += nthr) {
        const int r = e / nt, c = NC0 + (e - r * nt);
        const int q = r >> 2, b = r & 3;
        int y = ysq[q] + b;
        y = y < M ? y : y - M;
        S[r * stride + c] = masters[(size_t)q * layer_stride + (size_t)y * M + wrap_col(xsq[q] + c, M)];
    }
}

// ------------------------------------------------------------------ configs
template <int N>
struct RowCfg {
    static constexpr int T = N / 16;
    static constexpr int ROWS0 = 256 / T > 0 ? 256 / T : 1;
    static constexpr int ROWS = ROWS0 < N ? ROWS0 : N;
    static constexpr int THREADS = ROWS * T;
};
template <int N>
struct ColCfg {
    static constexpr int C0 = N <= 1024 ? 16 : (N == 2048 ? 8 : (N == 4096 ? 4 : 2));
    static constexpr int C = C0 < N ? C0 : N;
    static constexpr int THREADS = C * (N / 16);
};

// ------------------------------------------------------------------ passes
template <int N, int OP, int DIR, bool GEN>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
k_row(float2* __restrict__ data, const float2* __restrict__ tf, int W, const float2* __restrict__ tw,
      ScreenArgs sa) {
    using P = Plan<N>;
    constexpr int T = P::T, ROWS = RowCfg<N>::ROWS;
    constexpr int RI = (OP == OP_FRESNEL) ? P::first(-1) : P::first(DIR);
    constexpr int RO = (OP == OP_FRESNEL) ? P::last(+1) : P::last(DIR);
    __shared__ float2 lds[ROWS * P::LDS];
    const int lr = threadIdx.x / T, t = threadIdx.x - lr * T;
    const int row = blockIdx.x * ROWS + lr;
    const int f = blockIdx.y;
    float2* rowp = data + ((size_t)f * N + row) * N;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = GEN ? make_float2(1.f, 0.f) : rowp[reg_index<N, RI>(t, q)];
    if (sa.lcount > 0) {
        const int w = f % sa.W, d = f / sa.W;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const double ph = screen_phase(sa, d, w, reg_index<N, RI>(t, q), row);
            v[q] = cmul(v[q], expi_d(ph));
        }
    }
    float2* L = lds + lr * P::LDS;
    if constexpr (OP == OP_FRESNEL) {
        fft_regs<N, -1, 1>(v, L, t, 0, tw);
        const float2* h = tf + (size_t)(f % W) * N;
        constexpr int RL = P::last(-1);
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], h[reg_index<N, RL>(t, q)]);
        fft_regs<N, +1, 1>(v, L, t, 0, tw);
    } else {
        fft_regs<N, DIR, 1>(v, L, t, 0, tw);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) rowp[reg_index<N, RO>(t, q)] = v[q];
}

template <int N, int OP, int DIR, int C>
__global__ void __launch_bounds__(C * (N / 16))
k_col(float2* __restrict__ data, const float2* __restrict__ tf, int W, const float2* __restrict__ tw, int wlo,
      int wcnt) {
    using P = Plan<N>;
    constexpr int RI = (OP == OP_FRESNEL) ? P::first(-1) : P::first(DIR);
    constexpr int RO = (OP == OP_FRESNEL) ? P::last(+1) : P::last(DIR);
    __shared__ float2 lds[C * P::LDSC];
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int ntile = N / C;
    const int tile = xcd_swizzle(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int fi = tile / ntile;
    // field of launch index fi: all fields, or wavelengths wlo.. of every direction
    const int f = wcnt > 0 ? (fi / wcnt) * W + wlo + fi % wcnt : fi;
    const int col = (tile - fi * ntile) * C + c;
    float2* base = data + (size_t)f * N * N + col;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = base[(size_t)reg_index<N, RI>(t, q) * N];
    if constexpr (OP == OP_FRESNEL) {
        fft_regs<N, -1, C>(v, lds, t, c, tw);
        const float2* h = tf + (size_t)(f % W) * N;
        constexpr int RL = P::last(-1);
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], h[reg_index<N, RL>(t, q)]);
        fft_regs<N, +1, C>(v, lds, t, c, tw);
    } else {
        fft_regs<N, DIR, C>(v, lds, t, c, tw);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) base[(size_t)reg_index<N, RO>(t, q) * N] = v[q];
}

// Column Fresnel pass with E = 32 elements per thread (N / 32 threads per
// column, C * N / 32 per workgroup) and float exchanges: C columns take
// C * Plan::LDSC * 4 B of LDS.  C = 8 (72 KB at N = 2048) lets two
// workgroups share a CU; C = 16 (144 KB) reads 128-byte row segments.  At
// most 128 VGPRs (4 waves per SIMD).
template <int N, int C, int E>
__global__ void __launch_bounds__(C * (N / E)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_split(float2* __restrict__ data, const float2* __restrict__ tf, int W, const float2* __restrict__ tw, int wlo,
            int wcnt) {
    using P = Plan<N>;
    constexpr int RI = P::first(-1), RO = P::last(+1), RL = P::last(-1);
    __shared__ float lds[C * P::LDSC];
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int ntile = N / C;
    const int tile = xcd_swizzle(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int fi = tile / ntile;
    const int f = wcnt > 0 ? (fi / wcnt) * W + wlo + fi % wcnt : fi;
    const int col = (tile - fi * ntile) * C + c;
    // buffer addressing: one lane offset (row t, column col) and the
    // register slot's row as a constant scalar offset
    const auto fr = buf_rsrc(data + (size_t)f * N * N, (uint32_t)N * N * 8);
    const uint32_t lo = ((uint32_t)t * N + col) * 8;
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = buf_load2(fr, lo, (reg_index<N, RI, E>(0, q) * N) * 8);
    fft_regs_split<N, -1, C, E>(v, lds, t, c, tw);
    const auto hr = buf_rsrc(tf + (size_t)(f % W) * N, (uint32_t)N * 8);
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = cmul(v[q], buf_load2(hr, (uint32_t)t * 8, reg_index<N, RL, E>(0, q) * 8));
    fft_regs_split<N, +1, C, E>(v, lds, t, c, tw);
#pragma unroll
    for (int q = 0; q < E; q++) buf_store2(v[q], fr, lo, (reg_index<N, RO, E>(0, q) * N) * 8);
}

// Column pass around a layer group (the x and y factors of a gap commute, so
// the gaps alternate between x-first and y-first and one pass over the field
// applies the y factor of the gap before the group, the group's screen and
// the y factor of the gap after it).  Workgroup: C columns of one field.  The
// group's layer phases come sampled (FP64 quarter turns, one plane per layer
// and direction, k_layer_phase): the phase of a pixel is sum_l s_wl V_l,
// applied as one e^{i phase}.  The fields of one column tile are adjacent
// launch indices on one XCD, so they read the phase planes through its L2.
template <int N, int C>
__global__ void __launch_bounds__(C * (N / 16))
k_col_screen(float2* __restrict__ data, const float2* __restrict__ tfp, const float2* __restrict__ tfc,
             const float2* __restrict__ tw, ScreenArgs sa, const double* __restrict__ vq, int slot0) {
    using P = Plan<N>;
    constexpr int RI = P::first(-1), RL = P::last(-1);
    __shared__ float2 lds[C * P::LDSC];
    static_assert(P::first(-1) == P::last(+1), "the inverse plan must end in the forward plan's layout");
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int nf = gridDim.y;
    // 16 consecutive launch indices per XCD: with 8 fields, both 64-byte
    // halves of a 128-byte line of every field and of the phase planes are
    // read on one XCD
    const int lin = xcd_chunk<16>(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int ct = lin / nf, fi = lin - ct * nf;
    const int nw = sa.wcnt > 0 ? sa.wcnt : sa.W;
    const int d = fi / nw, w = (sa.wcnt > 0 ? sa.wlo : 0) + (fi - d * nw);
    const int f = d * sa.W + w;
    const int D = nf / nw;
    const int col0 = ct * C;
    // buffer addressing: one lane offset (row t, column col0 + c), each
    // register slot's row as a constant scalar offset (no per-slot address
    // registers live across the kernel)
    const auto fr = buf_rsrc(data + (size_t)f * N * N, (uint32_t)N * N * 8);
    const uint32_t lo = ((uint32_t)t * N + col0 + c) * 8;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = buf_load2(fr, lo, (reg_index<N, RI>(0, q) * N) * 8);
    fft_regs<N, -1, C>(v, lds, t, c, tw);
    {
        const auto hr = buf_rsrc(tfp + (size_t)w * N, (uint32_t)N * 8);
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], buf_load2(hr, (uint32_t)t * 8, reg_index<N, RL>(0, q) * 8));
    }
    fft_regs<N, +1, C>(v, lds, t, c, tw);
    if (sa.lcount > 0) {
        double ph[16];
#pragma unroll
        for (int q = 0; q < 16; q++) ph[q] = 0.0;
        for (int k = 0; k < sa.lcount; k++) {
            const double sw = sa.s[(size_t)w * sa.L + sa.lfirst + k];
            const auto vr = buf_rsrc(vq + ((size_t)(slot0 + k) * D + d) * N * N, (uint32_t)N * N * 8);
#pragma unroll
            for (int q = 0; q < 16; q++) ph[q] = fma(sw, buf_load_f64(vr, lo, (reg_index<N, RI>(0, q) * N) * 8), ph[q]);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], expi_q(ph[q]));
    }
    // the second factor's twiddle and LDS indices are recomputed, not kept
    // live through the screen from the first factor
    const int t2 = opaque(t), c2 = opaque(c);
    fft_regs<N, -1, C>(v, lds, t2, c2, tw);
    {
        const auto hr = buf_rsrc(tfc + (size_t)w * N, (uint32_t)N * 8);
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], buf_load2(hr, (uint32_t)t2 * 8, reg_index<N, RL>(0, q) * 8));
    }
    fft_regs<N, +1, C>(v, lds, t2, c2, tw);
#pragma unroll
    for (int q = 0; q < 16; q++) buf_store2(v[q], fr, lo, (reg_index<N, RI>(0, q) * N) * 8);
}

// The same column pass with E = 32 elements per thread (N / 32 threads per
// column) exchanging through float LDS: C = 8 columns take 8 Plan::LDSC 4 B
// (72 KB at N = 2048), so two workgroups share a CU and one's barriers and
// loads overlap the other's arithmetic.  At most 128 VGPRs (4 waves per
// SIMD).
template <int N, int C, int E>
__global__ void __launch_bounds__(C * (N / E)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_screen_split(float2* __restrict__ data, const float2* __restrict__ tfp, const float2* __restrict__ tfc,
                   const float2* __restrict__ tw, ScreenArgs sa, const double* __restrict__ vq, int slot0) {
    using P = Plan<N>;
    constexpr int RI = P::first(-1), RL = P::last(-1);
    __shared__ float lds[C * P::LDSC];
    static_assert(P::first(-1) == P::last(+1), "the inverse plan must end in the forward plan's layout");
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int nf = gridDim.y;
    // launch indices run in groups of 4 adjacent column tiles (256 bytes of
    // every row) of one field, fields next, and each XCD takes a contiguous
    // eighth of them: the tiles that share 128-byte lines, and every field's
    // phase-plane tiles, are read through one XCD's L2 at about the same
    // time (2048^2 x 8 fields: 262-265 us and 1.20-1.22 x the algorithmic
    // bytes, against 266 us and 1.27 x with single tiles and 285 us with
    // 16-index chunks)
    constexpr int TG = 4;
    static_assert((N / C) % TG == 0, "column tiles come in groups of 4");
    const int lin = xcd_slab(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int h = lin % TG, rest = lin / TG;
    const int fi = rest % nf, ct = (rest / nf) * TG + h;
    const int nw = sa.wcnt > 0 ? sa.wcnt : sa.W;
    const int d = fi / nw, w = (sa.wcnt > 0 ? sa.wlo : 0) + (fi - d * nw);
    const int f = d * sa.W + w;
    const int D = nf / nw;
    const int col0 = ct * C;
    const auto fr = buf_rsrc(data + (size_t)f * N * N, (uint32_t)N * N * 8);
    const uint32_t lo = ((uint32_t)t * N + col0 + c) * 8;
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = buf_load2(fr, lo, (reg_index<N, RI, E>(0, q) * N) * 8);
    fft_regs_split<N, -1, C, E, true>(v, lds, t, c, tw);
    {
        const auto hr = buf_rsrc(tfp + (size_t)w * N, (uint32_t)N * 8);
#pragma unroll
        for (int q = 0; q < E; q++) v[q] = cmul(v[q], buf_load2(hr, (uint32_t)t * 8, reg_index<N, RL, E>(0, q) * 8));
    }
    fft_regs_split<N, +1, C, E, true>(v, lds, t, c, tw);
    // the group's phase sum_l s_wl V_l, 8 register slots at a time
#pragma unroll
    for (int q0 = 0; q0 < E; q0 += 8) {
        double ph[8];
#pragma unroll
        for (int u = 0; u < 8; u++) ph[u] = 0.0;
        for (int k = 0; k < sa.lcount; k++) {
            const double sw = sa.s[(size_t)w * sa.L + sa.lfirst + k];
            const auto vr = buf_rsrc(vq + ((size_t)(slot0 + k) * D + d) * N * N, (uint32_t)N * N * 8);
#pragma unroll
            for (int u = 0; u < 8; u++)
                ph[u] = fma(sw, buf_load_f64(vr, lo, (reg_index<N, RI, E>(0, q0 + u) * N) * 8), ph[u]);
        }
        if (sa.lcount > 0) {
#pragma unroll
            for (int u = 0; u < 8; u++) v[q0 + u] = cmul(v[q0 + u], expi_q(ph[u]));
        }
    }
    const int t2 = opaque(t), c2 = opaque(c);
    fft_regs_split<N, -1, C, E, true>(v, lds, t2, c2, tw);
    {
        const auto hr = buf_rsrc(tfc + (size_t)w * N, (uint32_t)N * 8);
#pragma unroll
        for (int q = 0; q < E; q++) v[q] = cmul(v[q], buf_load2(hr, (uint32_t)t2 * 8, reg_index<N, RL, E>(0, q) * 8));
    }
    fft_regs_split<N, +1, C, E, true>(v, lds, t2, c2, tw);
#pragma unroll
    for (int q = 0; q < E; q++) buf_store2(v[q], fr, lo, (reg_index<N, RI, E>(0, q) * N) * 8);
}

// Sampled phase of layer slot_layer[slot] for direction d, in FP64 quarter
// turns (x 2 / pi): vq[((slot D + d) N + j) N + i].  A workgroup makes PR
// consecutive rows j0 .. j0 + PR - 1 from master rows y0 + j0 - 1 ..
// y0 + j0 + PR + 2 (PR + 3 <= 8 staged rows); every pixel sums its 4 x 4
// taps in the order of the row pass (x taps per master row, then the rows).
constexpr int PHASE_ROWS = 4;
__global__ void __launch_bounds__(512) k_layer_phase(double* __restrict__ vq, const float* __restrict__ masters,
                                                     const LayerSample* __restrict__ ls,
                                                     const int* __restrict__ slot_layer, int N, int M, int L) {
    constexpr int PR = PHASE_ROWS;
    static_assert(PR + 3 <= 8, "two 4-row stagings");
    extern __shared__ float2 lds[];
    float* S = reinterpret_cast<float*>(lds);  // [8][N + 4]
    // each XCD makes a contiguous eighth of the row blocks (adjacent blocks
    // stage overlapping master rows)
    const int nb = N / PR;
    const int j0 = xcd_slab(blockIdx.x, nb) * PR, d = blockIdx.y, slot = blockIdx.z, D = gridDim.y;
    const int SW = N + 4;
    const int l = slot_layer[slot];
    const LayerSample* p = ls + (size_t)d * L + l;
    const float* m = masters + (size_t)l * M * M;
    int ys[2];
    ys[0] = mod_m(p->y0 + j0 - 1, M);
    ys[1] = mod_m(p->y0 + j0 + 3, M);
    int xs[2];
    xs[0] = xs[1] = mod_m(p->x0 - 1, M);
    stage_rows4<4>(S, SW, m, M, ys[0], xs[0], N, threadIdx.x, blockDim.x);
    stage_rows4<4>(S + 4 * SW, SW, m, M, ys[1], xs[1], N, threadIdx.x, blockDim.x);
    stage_tail(S, SW, m, 0, M, ys, xs, 2, N, 3, threadIdx.x, blockDim.x);
    __syncthreads();
    // per column: the x-filtered value of each staged row once, then the PR
    // outputs from 4 of them each (the same FP64 operations in the same order
    // as one pixel at a time)
    double* out = vq + (((size_t)slot * D + d) * N + j0) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double h[PR + 3];
#pragma unroll
        for (int r = 0; r < PR + 3; r++) {
            const float* q = S + r * SW + i;
            double x = p->wx[0] * (double)q[0];
            x = fma(p->wx[1], (double)q[1], x);
            x = fma(p->wx[2], (double)q[2], x);
            h[r] = fma(p->wx[3], (double)q[3], x);
        }
#pragma unroll
        for (int rr = 0; rr < PR; rr++) {
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < 4; b++) v = fma(p->wy[b], h[rr + b], v);
            out[(size_t)rr * N + i] = v * 0.63661977236758134308;
        }
    }
}

// Fresnel x pass for up to MW wavelength fields of the same direction and
// spatial row j: the screen phase of the layer group is sampled once per
// pixel (FP64, master rows staged in LDS) and shared by the wavelengths,
// then each field is multiplied by e^{i s_w v}, transformed along x,
// multiplied by its transfer factor and transformed back.
template <int N>
struct LamCfg {
    static constexpr int T = N / 16;
    static constexpr int MW0 = 16384 / N;
    static constexpr int MW = MW0 < 8 ? (MW0 < 1 ? 1 : MW0) : 8;  // most wavelengths per workgroup
    // default: at most 512 threads and N (N + N/16) 8 / 2 bytes of LDS per
    // workgroup, so that two workgroups share a CU and one's loads and
    // stores overlap the other's arithmetic (2048: 172 vs 202 us for 8)
    static constexpr int MWD0 = 8192 / N;
    static constexpr int MWD = MWD0 < MW ? (MWD0 < 1 ? 1 : MWD0) : MW;
    static constexpr int SEG = N < 2048 ? N : 2048;
};

// PRE: the row first gets the x factor tfp of the gap before the group
// (forward transform, multiply, inverse transform; the inverse leaves the
// registers in the layout the forward transform reads), then the screen and
// the x factor tf of the gap after it.
template <int N, bool GEN, bool PRE>
__global__ void __launch_bounds__(LamCfg<N>::MW * (N / 16))
k_row_fresnel(float2* __restrict__ data, const float2* __restrict__ tf, const float2* __restrict__ tw,
              ScreenArgs sa, const float2* __restrict__ tfp) {
    using P = Plan<N>;
    constexpr int T = P::T, MW = LamCfg<N>::MW, SEG = LamCfg<N>::SEG, SW = SEG + 4;
    constexpr int RI = P::first(-1), RO = P::last(+1), RL = P::last(-1);
    extern __shared__ float2 lds[];  // max(mwc * P::LDS float2, screen staging)
    float* S = reinterpret_cast<float*>(lds);            // [4][SW] staged master rows
    double* V = reinterpret_cast<double*>(S + 4 * SW);  // [SEG] sampled screen
    const int mwc = blockDim.x / T;                      // wavelengths per workgroup (<= MW)
    const int wend = sa.wcnt > 0 ? sa.wlo + sa.wcnt : sa.W;  // end of this launch's wavelength range
    // 16 adjacent rows (which stage overlapping master rows) run on one XCD
    const int j = xcd_chunk<16>(blockIdx.x, N), d = blockIdx.y, w0 = (sa.wcnt > 0 ? sa.wlo : 0) + blockIdx.z * mwc;
    const int nw = min(mwc, wend - w0);
    const int lw = threadIdx.x / T, t = threadIdx.x - lw * T;
    const bool active = lw < nw;
    const int w = min(w0 + lw, wend - 1);
    const int f = d * sa.W + w;
    const int nthr = blockDim.x;
    const int M = (int)sa.M;
    float2* rowp = data + ((size_t)f * N + j) * N;
    float2 v[16];
    // the field is needed only after the screen: issue its loads first
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = (GEN || !active) ? make_float2(1.f, 0.f) : rowp[reg_index<N, RI>(t, q)];
    static_assert(RI == RO, "the inverse plan must end in the layout the forward plan starts from");
    if constexpr (PRE) {
        float2* Lp = lds + lw * P::LDS;
        fft_regs<N, -1, 1, true>(v, Lp, t, 0, tw);
        const float2* hp = tfp + (size_t)w * N;
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = cmul(v[q], hp[reg_index<N, RL>(t, q)]);
        fft_regs<N, +1, 1, true>(v, Lp, t, 0, tw);
    }
    __shared__ LayerSample sls[MAX_LAYERS];
    __shared__ double ss[8][MAX_LAYERS];
    for (int k = threadIdx.x; k < sa.lcount; k += nthr) sls[k] = sa.ls[d * sa.L + sa.lfirst + k];
    for (int k = threadIdx.x; k < mwc * sa.lcount; k += nthr) {
        const int lw2 = k / sa.lcount, l = k - lw2 * sa.lcount;
        ss[lw2][l] = sa.s[min(w0 + lw2, wend - 1) * sa.L + sa.lfirst + l];
    }
    // each layer of the group multiplies the field by its own e^{i s v}: no
    // per-element phase accumulators stay live across the staging
    for (int k = 0; k < sa.lcount; k++) {
        const int l = sa.lfirst + k;
        __syncthreads();
        const LayerSample& p = sls[k];
        const float* m = sa.masters + (size_t)l * M * M;
        const double sw = ss[lw][k];
        for (int seg0 = 0; seg0 < N; seg0 += SEG) {
            __syncthreads();
            const int ys = mod_m(p.y0 + j - 1, M), xst = mod_m(p.x0 + seg0 - 1, M);  // uniform
            stage_rows4<2>(S, SW, m, M, ys, xst, SEG, threadIdx.x, nthr);
            stage_tail(S, SW, m, 0, M, &ys, &xst, 1, SEG, 3, threadIdx.x, nthr);
            __syncthreads();
            for (int i = threadIdx.x; i < SEG; i += nthr) {
                double v = 0.0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const float* r = S + b * SW + i;
                    double h = p.wx[0] * (double)r[0];
                    h = fma(p.wx[1], (double)r[1], h);
                    h = fma(p.wx[2], (double)r[2], h);
                    h = fma(p.wx[3], (double)r[3], h);
                    v = fma(p.wy[b], h, v);
                }
                V[i] = v * 0.63661977236758134308;  // in quarter turns (2 / pi)
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int idx = reg_index<N, RI>(t, q) - seg0;
                if (SEG == N || (idx >= 0 && idx < SEG)) v[q] = cmul(v[q], expi_q(sw * V[idx]));
            }
        }
    }
    float2* Lr = lds + lw * P::LDS;
    const int t2 = PRE ? opaque(t) : t;  // with PRE: indices recomputed, not kept live through the screen
    fft_regs<N, -1, 1, true>(v, Lr, t2, 0, tw);
    const float2* h = tf + (size_t)w * N;
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = cmul(v[q], h[reg_index<N, RL>(t2, q)]);
    fft_regs<N, +1, 1, true>(v, Lr, t2, 0, tw);
    if (active) {
#pragma unroll
        for (int q = 0; q < 16; q++) rowp[reg_index<N, RO>(t, q)] = v[q];
    }
}

template <int N>
static hipError_t row_fresnel_dispatch(bool gen, float2* data, int D, const float2* tf, const float2* tw,
                                       const ScreenArgs& sa, int mw, hipStream_t st, const float2* tfp) {
    constexpr int MW = LamCfg<N>::MW, SEG = LamCfg<N>::SEG;
    int mwc = (mw > 0 && mw < MW) ? mw : (mw >= MW ? MW : LamCfg<N>::MWD);
    const int nwl = sa.wcnt > 0 ? sa.wcnt : sa.W;  // wavelengths of this launch
    if (nwl < mwc) mwc = nwl;
    size_t lds = (size_t)mwc * Plan<N>::LDS * sizeof(float2);
    const size_t staging = (size_t)16 * (SEG + 4) + (size_t)8 * SEG;
    if (lds < staging) lds = staging;
    dim3 grid(N, D, (nwl + mwc - 1) / mwc), block(mwc * (N / 16));
    static bool attr = false;
    if (!attr) {
        allow_dynamic_lds((const void*)k_row_fresnel<N, true, false>);
        allow_dynamic_lds((const void*)k_row_fresnel<N, false, false>);
        allow_dynamic_lds((const void*)k_row_fresnel<N, false, true>);
        attr = true;
    }
    if (gen && tfp) return hipErrorInvalidValue;
    if (gen) k_row_fresnel<N, true, false><<<grid, block, lds, st>>>(data, tf, tw, sa, nullptr);
    else if (tfp) k_row_fresnel<N, false, true><<<grid, block, lds, st>>>(data, tf, tw, sa, tfp);
    else k_row_fresnel<N, false, false><<<grid, block, lds, st>>>(data, tf, tw, sa, nullptr);
    return hipGetLastError();
}

template <int N>
static hipError_t row_dispatch(int op, int dir, bool gen, float2* data, long F, const float2* tf, int W,
                               const float2* tw, const ScreenArgs& sa, hipStream_t st) {
    dim3 grid(N / RowCfg<N>::ROWS, (unsigned)F), block(RowCfg<N>::THREADS);
    if (op == OP_FRESNEL) {
        if (gen) k_row<N, OP_FRESNEL, 0, true><<<grid, block, 0, st>>>(data, tf, W, tw, sa);
        else k_row<N, OP_FRESNEL, 0, false><<<grid, block, 0, st>>>(data, tf, W, tw, sa);
    } else if (dir < 0) {
        k_row<N, OP_FFT, -1, false><<<grid, block, 0, st>>>(data, tf, W, tw, sa);
    } else {
        k_row<N, OP_FFT, +1, false><<<grid, block, 0, st>>>(data, tf, W, tw, sa);
    }
    return hipGetLastError();
}

// column tile widths available for n (C * n/16 <= 1024 threads, C * LDSC <= 144 KB)
template <int N, int C>
constexpr bool col_ok() {
    return C <= N && C * (N / 16) <= 1024 && (long)C * Plan<N>::LDSC * 8 <= 147456;
}

template <int N, int C>
static hipError_t col_launch(int op, int dir, float2* data, long F, const float2* tf, int W, const float2* tw,
                             int wlo, int wcnt, hipStream_t st) {
    if constexpr (!col_ok<N, C>()) {
        return hipErrorInvalidValue;
    } else {
        dim3 grid(N / C, (unsigned)F), block(C * (N / 16));
        if (op == OP_FRESNEL) k_col<N, OP_FRESNEL, 0, C><<<grid, block, 0, st>>>(data, tf, W, tw, wlo, wcnt);
        else if (dir < 0) k_col<N, OP_FFT, -1, C><<<grid, block, 0, st>>>(data, tf, W, tw, 0, 0);
        else k_col<N, OP_FFT, +1, C><<<grid, block, 0, st>>>(data, tf, W, tw, 0, 0);
        return hipGetLastError();
    }
}

template <int N>
static int col_default_c() { return ColCfg<N>::C; }

template <int N>
static hipError_t col_dispatch(int op, int dir, float2* data, long F, const float2* tf, int W, const float2* tw,
                               int c, int e, int wlo, int wcnt, hipStream_t st) {
    if constexpr (N >= 256 && N <= 2048) {
        if (e == 32 && op == OP_FRESNEL && (c <= 0 || c == 8)) {
            dim3 grid(N / 8, (unsigned)F), block(8 * (N / 32));
            k_col_split<N, 8, 32><<<grid, block, 0, st>>>(data, tf, W, tw, wlo, wcnt);
            return hipGetLastError();
        }
        if (e == 32 && op == OP_FRESNEL && c == 16) {
            dim3 grid(N / 16, (unsigned)F), block(16 * (N / 32));
            k_col_split<N, 16, 32><<<grid, block, 0, st>>>(data, tf, W, tw, wlo, wcnt);
            return hipGetLastError();
        }
    }
    if (c <= 0) c = col_default_c<N>();
    switch (c) {
        case 2: return col_launch<N, 2>(op, dir, data, F, tf, W, tw, wlo, wcnt, st);
        case 4: return col_launch<N, 4>(op, dir, data, F, tf, W, tw, wlo, wcnt, st);
        case 8: return col_launch<N, 8>(op, dir, data, F, tf, W, tw, wlo, wcnt, st);
        case 16: return col_launch<N, 16>(op, dir, data, F, tf, W, tw, wlo, wcnt, st);
        default: return hipErrorInvalidValue;
    }
}

bool supported_n(long n) { return n >= 16 && n <= 8192 && (n & (n - 1)) == 0; }

#define ATSIM_FOR_EACH_N(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

hipError_t launch_row(long n, int op, int dir, bool gen, float2* data, long F, const float2* tf,
                      const float2* tw, const ScreenArgs& sa, hipStream_t st) {
    const int W = sa.W > 0 ? sa.W : 1;
    switch (n) {
#define X(NN) \
    case NN: return row_dispatch<NN>(op, dir, gen, data, F, tf, W, tw, sa, st);
        ATSIM_FOR_EACH_N(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_row_fresnel(long n, bool gen, float2* data, int D, const float2* tf, const float2* tw,
                              const ScreenArgs& sa, hipStream_t st, int mw, const float2* tfp) {
    switch (n) {
#define X(NN) \
    case NN: return row_fresnel_dispatch<NN>(gen, data, D, tf, tw, sa, mw, st, tfp);
        ATSIM_FOR_EACH_N(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_col(long n, int op, int dir, float2* data, long F, const float2* tf, int W,
                      const float2* tw, hipStream_t st, int c, int e, int wlo, int wcnt) {
    switch (n) {
#define X(NN) \
    case NN: return col_dispatch<NN>(op, dir, data, F, tf, W > 0 ? W : 1, tw, c, e, wlo, wcnt, st);
        ATSIM_FOR_EACH_N(X)
#undef X
        default: return hipErrorInvalidValue;
    }
}

bool col_screen_supported(long n) { return supported_n(n) && n <= 4096; }

hipError_t launch_layer_phase(double* vq, const float* masters, const LayerSample* ls, const int* slot_layer,
                              int nslots, long N, long M, long L, long D, hipStream_t st) {
    if (nslots <= 0) return hipSuccess;
    static bool attr = false;
    if (!attr) {
        allow_dynamic_lds((const void*)k_layer_phase);
        attr = true;
    }
    if (N % PHASE_ROWS) return hipErrorInvalidValue;
    dim3 grid((unsigned)(N / PHASE_ROWS), (unsigned)D, (unsigned)nslots);
    k_layer_phase<<<grid, 512, (size_t)8 * (N + 4) * sizeof(float), st>>>(vq, masters, ls, slot_layer, (int)N,
                                                                           (int)M, (int)L);
    return hipGetLastError();
}

template <int N>
static hipError_t col_screen_split(float2* data, int D, const float2* tfp, const float2* tf, const float2* tw,
                                   const ScreenArgs& sa, const double* vq, int slot0, hipStream_t st) {
    const int nw = sa.wcnt > 0 ? sa.wcnt : sa.W;
    dim3 grid(N / 8, (unsigned)(D * nw)), block(8 * (N / 32));
    k_col_screen_split<N, 8, 32><<<grid, block, 0, st>>>(data, tfp, tf, tw, sa, vq, slot0);
    return hipGetLastError();
}

hipError_t launch_col_screen(long n, float2* data, int D, const float2* tfp, const float2* tf, const float2* tw,
                             const ScreenArgs& sa, const double* vq, int slot0, hipStream_t st, int e) {
    if (!tfp || !tf || D <= 0 || (sa.lcount > 0 && !vq)) return hipErrorInvalidValue;
    const int nw = sa.wcnt > 0 ? sa.wcnt : sa.W;
    if (e != 16 && n >= 256 && n <= 2048) {
        switch (n) {
            case 256: return col_screen_split<256>(data, D, tfp, tf, tw, sa, vq, slot0, st);
            case 512: return col_screen_split<512>(data, D, tfp, tf, tw, sa, vq, slot0, st);
            case 1024: return col_screen_split<1024>(data, D, tfp, tf, tw, sa, vq, slot0, st);
            default: return col_screen_split<2048>(data, D, tfp, tf, tw, sa, vq, slot0, st);
        }
    }
    switch (n) {
#define X(NN)                                                                                         \
    case NN: {                                                                                        \
        constexpr int C = ColCfg<NN>::C;                                                              \
        dim3 grid(NN / C, (unsigned)(D * nw)), block(C * (NN / 16));                                  \
        k_col_screen<NN, C><<<grid, block, 0, st>>>(data, tfp, tf, tw, sa, vq, slot0);                \
        return hipGetLastError();                                                                     \
    }
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------ permut
__global__ void k_permut(float2* __restrict__ d, long n0, long n1, long nbatch) {
    const long half = n0 * (n1 / 2);
    const long total = half * nbatch;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long b = e / half, r = e - b * half;
        const long jj = r / n0, ii = r - jj * n0;
        float2* img = d + b * n0 * n1;
        const long a = jj * n0 + ii;
        const long p = (jj + n1 / 2) * n0 + (ii + n0 / 2) % n0;
        float2 t = img[a];
        img[a] = img[p];
        img[p] = t;
    }
}

hipError_t launch_permut(float2* data, long n0, long n1, long nbatch, hipStream_t st) {
    long total = n0 * (n1 / 2) * nbatch;
    int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
    if (blocks < 1) blocks = 1;
    k_permut<<<blocks, 256, 0, st>>>(data, n0, n1, nbatch);
    return hipGetLastError();
}

// ------------------------------------------------------------------ transforms of any size
// A 1D transform of length L along one axis of a batch of images is run on
// contiguous rows of a padded scratch [rows][M] (M a power of two): k_line_gather
// copies line l, element n (n < L) from src + base(l) + n estride, times
// chirp[n] when given, and zeroes n in [L, M); k_line_scatter writes element
// k < L back, times chirp[k] when given.  base(l) = (l / per_img) img_size +
// (l % per_img) lstep.  Lines are the fast index of the launch so that lines
// adjacent in memory (the columns of an image) load together.
__global__ void k_line_gather(const float2* __restrict__ src, float2* __restrict__ dst, long nlines, long L, long M,
                              long per_img, long img_size, long lstep, long estride,
                              const float2* __restrict__ chirp) {
    const long total = nlines * M;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long n = e / nlines, l = e - n * nlines;
        float2 v = make_float2(0.f, 0.f);
        if (n < L) {
            const long base = (l / per_img) * img_size + (l % per_img) * lstep;
            v = src[base + n * estride];
            if (chirp) v = cmul(v, chirp[n]);
        }
        dst[l * M + n] = v;
    }
}

__global__ void k_line_scatter(const float2* __restrict__ src, float2* __restrict__ dst, long nlines, long L, long M,
                               long per_img, long img_size, long lstep, long estride,
                               const float2* __restrict__ chirp) {
    const long total = nlines * L;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long k = e / nlines, l = e - k * nlines;
        float2 v = src[l * M + k];
        if (chirp) v = cmul(v, chirp[k]);
        dst[(l / per_img) * img_size + (l % per_img) * lstep + k * estride] = v;
    }
}

// data[r][k] *= h[k] for rows r < nrows of length M
__global__ void k_row_mul(float2* __restrict__ data, const float2* __restrict__ h, long nrows, long M) {
    const long total = nrows * M;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
        data[e] = cmul(data[e], h[e % M]);
}

// out[b][p] = in[b][sigma[p]] over nbatch images of n elements
__global__ void k_index_gather(const float2* __restrict__ in, float2* __restrict__ out, const int* __restrict__ sigma,
                               long n, long nbatch) {
    const long total = n * nbatch;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long b = e / n, p = e - b * n;
        out[e] = in[b * n + sigma[p]];
    }
}

// The Fresnel factor of Fresnel_propagate_wavefront (WFpropagate.c:68-152) on
// the spectrum of any n0 x n1 image: permut, times e^{-i coeff ((ii - n0/2)^2 +
// (jj - n1/2)^2)} / (n0 n1) at the permuted position (ii, jj), permut.  With
// sigma the gather map of one permut (out[p] = in[sigma(p)]): out[p] =
// in[sigma(sigma(p))] h(sigma(p)).  sigma null: even sizes, where permut is
// the half-size shift and sigma(sigma(p)) = p.  Angle and factor in FP64.
__global__ void k_fresnel_any(const float2* __restrict__ in, float2* __restrict__ out, long n0, long n1, long nbatch,
                              double coeff, const int* __restrict__ sigma) {
    const long n = n0 * n1, total = n * nbatch;
    const double inv = 1.0 / ((double)n0 * (double)n1);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long b = e / n, p = e - b * n;
        long q, src;
        if (sigma) {
            q = sigma[p];
            src = sigma[q];
        } else {
            const long jj = p / n0, ii = p - jj * n0;
            q = ((jj + n1 / 2) % n1) * n0 + (ii + n0 / 2) % n0;
            src = p;
        }
        const long jq = q / n0, iq = q - jq * n0;
        const long di = iq - n0 / 2, dj = jq - n1 / 2;
        const double ang = -coeff * (double)(di * di + dj * dj);
        double sn, cs;
        sincos(ang, &sn, &cs);
        const float2 v = in[b * n + src];
        const double re = (double)v.x * inv, im = (double)v.y * inv;
        out[e] = make_float2((float)(re * cs - im * sn), (float)(re * sn + im * cs));
    }
}

static unsigned grid_for(long total) {
    long g = (total + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

hipError_t launch_line_gather(const float2* src, float2* dst, long nlines, long L, long M, long per_img, long img_size,
                              long lstep, long estride, const float2* chirp, hipStream_t st) {
    k_line_gather<<<grid_for(nlines * M), 256, 0, st>>>(src, dst, nlines, L, M, per_img, img_size, lstep, estride,
                                                         chirp);
    return hipGetLastError();
}
hipError_t launch_line_scatter(const float2* src, float2* dst, long nlines, long L, long M, long per_img,
                               long img_size, long lstep, long estride, const float2* chirp, hipStream_t st) {
    k_line_scatter<<<grid_for(nlines * L), 256, 0, st>>>(src, dst, nlines, L, M, per_img, img_size, lstep, estride,
                                                          chirp);
    return hipGetLastError();
}
hipError_t launch_row_mul(float2* data, const float2* h, long nrows, long M, hipStream_t st) {
    k_row_mul<<<grid_for(nrows * M), 256, 0, st>>>(data, h, nrows, M);
    return hipGetLastError();
}
hipError_t launch_index_gather(const float2* in, float2* out, const int* sigma, long n, long nbatch, hipStream_t st) {
    k_index_gather<<<grid_for(n * nbatch), 256, 0, st>>>(in, out, sigma, n, nbatch);
    return hipGetLastError();
}
hipError_t launch_fresnel_any(const float2* in, float2* out, long n0, long n1, long nbatch, double coeff,
                              const int* sigma, hipStream_t st) {
    k_fresnel_any<<<grid_for(n0 * n1 * nbatch), 256, 0, st>>>(in, out, n0, n1, nbatch, coeff, sigma);
    return hipGetLastError();
}

// ------------------------------------------------------------------ noise
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__global__ void k_noise(float2* __restrict__ out, long n, uint32_t k0, uint32_t k1) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        uint32_t c[4] = {(uint32_t)e, (uint32_t)((uint64_t)e >> 32), 0x0A75u, 0u};
        philox10(c, k0, k1);
        const uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
        const uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
        const double u1 = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (double)b * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        double s, co;
        sincos(2.0 * M_PI * u2, &s, &co);
        out[e] = make_float2((float)(r * co), (float)(r * s));
    }
}

hipError_t launch_noise(float2* out, long n, uint64_t seed, uint32_t stream, hipStream_t st) {
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    if (blocks < 1) blocks = 1;
    k_noise<<<blocks, 256, 0, st>>>(out, n, (uint32_t)seed, (uint32_t)(seed >> 32) ^ stream);
    return hipGetLastError();
}

// ------------------------------------------------------------------ synthesis
// storage index p holds centred index (p + s/2) mod s after the quadrant swap
__global__ void k_synth_filter(const float2* __restrict__ noise, float2* __restrict__ out, long size,
                               double f0, double fi, double rlim) {
    const long n = size * size, h = size / 2;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long py = e / size, px = e - py * size;
        const long cx = (px + h) % size, cy = (py + h) % size;
        const double dx = (double)(cx - h), dy = (double)(cy - h);
        const double r2 = dx * dx + dy * dy;
        double A = 0.0;
        const double tmpd = sqrt(r2 + f0 * f0);
        const bool inside = rlim > 0.0 && sqrt(r2) < rlim;
        if (!inside && tmpd > 0.0) A = sqrt(exp(-r2 / fi / fi)) / pow(tmpd, 11.0 / 6.0);
        if (cx == h && cy == h) A = 0.0;
        const float2 z = noise[cy * size + cx];
        out[e] = make_float2((float)(A * (double)z.x), (float)(A * (double)z.y));
    }
}

hipError_t launch_synth_filter(const float2* noise, float2* out, long size, double f0, double fi, double rlim,
                               hipStream_t st) {
    long n = size * size;
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    k_synth_filter<<<blocks, 256, 0, st>>>(noise, out, size, f0, fi, rlim);
    return hipGetLastError();
}

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NV], double* __restrict__ out) {
    __shared__ double red[NV][256 / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) red[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[threadIdx.x][q];
        out[threadIdx.x] = s;
    }
}

__global__ void __launch_bounds__(256) k_lag_sums(const float2* __restrict__ c, long size, double* partial) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long n = size * size;
    const int la[4] = {1, 1, 2, 2}, lb[4] = {1, 2, 1, 2};
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long y = e / size, x = e - y * size;
        const float2 p = c[e];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long x2 = (x + la[q]) % size, y2 = (y + lb[q]) % size;
            const float2 o = c[y2 * size + x2];
            const double dr = (double)p.x - (double)o.x, di = (double)p.y - (double)o.y;
            acc[q] = fma(dr, dr, acc[q]);
            acc[4 + q] = fma(di, di, acc[4 + q]);
        }
    }
    block_reduce_store<8>(acc, partial + (size_t)blockIdx.x * 8);
}

hipError_t launch_lag_sums(const float2* c, long size, double* partial, int nblocks, hipStream_t st) {
    k_lag_sums<<<nblocks, 256, 0, st>>>(c, size, partial);
    return hipGetLastError();
}

__global__ void k_split_scale(const float2* __restrict__ c, float* __restrict__ o1, float* __restrict__ o2,
                              long n, double s1, double s2) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float2 p = c[e];
        o1[e] = (float)((double)p.x * s1);
        if (o2) o2[e] = (float)((double)p.y * s2);
    }
}

hipError_t launch_split_scale(const float2* c, float* o1, float* o2, long n, double s1, double s2,
                              hipStream_t st) {
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    k_split_scale<<<blocks, 256, 0, st>>>(c, o1, o2, n, s1, s2);
    return hipGetLastError();
}

// sum of (s(x, y) - s(x + a, y + b))^2; circular, or over x < size - a
// (non circular in x, as the r0 measure) when circular == 0 (b must be 0)
__global__ void __launch_bounds__(256) k_strfn(const float* __restrict__ s, long size, long a, long b,
                                              int circular, double* partial) {
    double acc[1] = {0.0};
    const long nx = circular ? size : size - a;
    const long n = nx * size;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long y = e / nx, x = e - y * nx;
        const long x2 = (x + a) % size, y2 = ((y + b) % size + size) % size;
        const double d = (double)s[y * size + x] - (double)s[y2 * size + x2];
        acc[0] = fma(d, d, acc[0]);
    }
    block_reduce_store<1>(acc, partial + blockIdx.x);
}

hipError_t launch_strfn(const float* s, long size, long a, long b, int circular, double* partial, int nblocks,
                        hipStream_t st) {
    k_strfn<<<nblocks, 256, 0, st>>>(s, size, a, b, circular, partial);
    return hipGetLastError();
}

__global__ void k_scale(float* __restrict__ s, long n, double scale) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        s[e] = (float)((double)s[e] * scale);
}

hipError_t launch_scale(float* s, long n, double scale, hipStream_t st) {
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    if (blocks < 1) blocks = 1;
    k_scale<<<blocks, 256, 0, st>>>(s, n, scale);
    return hipGetLastError();
}

__global__ void k_cmul(float2* __restrict__ a, const float2* __restrict__ b, long n, long nbatch) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n * nbatch; e += (long)gridDim.x * blockDim.x) {
        const float2 x = a[e], y = b[e % n];
        // products in FP64 as Fresnel_propagate_wavefront1 (WFpropagate.c:302-310)
        a[e] = make_float2((float)((double)x.x * y.x - (double)x.y * y.y),
                           (float)((double)x.x * y.y + (double)x.y * y.x));
    }
}

hipError_t launch_cmul(float2* a, const float2* b, long n, long nbatch, hipStream_t st) {
    long tot = n * nbatch;
    int blocks = (int)((tot + 255) / 256 < 16384 ? (tot + 255) / 256 : 16384);
    if (blocks < 1) blocks = 1;
    k_cmul<<<blocks, 256, 0, st>>>(a, b, n, nbatch);
    return hipGetLastError();
}

__global__ void k_amp_pha(const float2* __restrict__ e, float* __restrict__ amp, float* __restrict__ pha, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double re = e[i].x, im = e[i].y;
        amp[i] = (float)sqrt(re * re + im * im);
        pha[i] = (float)atan2(im, re);
    }
}

hipError_t launch_amp_pha(const float2* e, float* amp, float* pha, long n, hipStream_t st) {
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    if (blocks < 1) blocks = 1;
    k_amp_pha<<<blocks, 256, 0, st>>>(e, amp, pha, n);
    return hipGetLastError();
}

// Init_Fresnel_propagate_wavefront (WFpropagate.c:166-258), written directly
// at the quadrant-swapped position (WFpropagate.c:258).
__global__ void k_tf_init(float2* __restrict__ C, long size, double coeff, double fpmaskrad) {
    const long n = size * size;
    const double co1 = (double)size * (double)size;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long jj = e / size, ii = e - jj * size;
        double Re = 0.0, Im = 0.0;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const double x = 1.0 * ii + 1.0 * (0.5 + i - 2) / 4 - (double)(size / 2);
                const double y = 1.0 * jj + 1.0 * (0.5 + j - 2) / 4 - (double)(size / 2);
                const double sq = x * x + y * y;
                if (sqrt(sq) < fpmaskrad) {
                    double s, c;
                    sincos(-coeff * sq, &s, &c);
                    Re += c;
                    Im += s;
                }
            }
        double Amp = sqrt(Re * Re + Im * Im) / 16.0;
        const double Pha = atan2(Im, Re);
        Amp = (Amp > 0.3) ? 1.0 : 0.0;
        const long dx = (ii + size / 2) % size, dy = (jj + size / 2) % size;
        C[dy * size + dx] = make_float2((float)(Amp * cos(Pha) / co1), (float)(Amp * sin(Pha) / co1));
    }
}

hipError_t launch_tf_init(float2* C, long size, double coeff, double fpmaskrad, hipStream_t st) {
    long n = size * size;
    int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    k_tf_init<<<blocks, 256, 0, st>>>(C, size, coeff, fpmaskrad);
    return hipGetLastError();
}

// atan2(y, x) in FP32: t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] with the
// hardware reciprocal, atan t = t P(t^2) with a degree-7 polynomial in t^2
// (least-squares fit of the absolute error on [0, 1], written here; 1.3e-7
// rad in FP32 Horner form), then the octant: pi/2 - a when |y| > |x|, pi - a
// when x < 0, and the sign of y.  Total error below 4e-7 rad (the float
// spacing near pi is 2.4e-7; float atan2 itself errs by up to 3.4e-7 on
// random Gaussian arguments); (0, 0) -> 0.
__device__ __forceinline__ float atan2_poly(float y, float x) {
    // no contracted products: every kernel that inlines it gets the same bits
#pragma clang fp contract(off)
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
    const float s = t * t;
    float p = -0.004054518882185221f;
    p = fmaf(p, s, 0.021862762048840523f);
    p = fmaf(p, s, -0.05591200292110443f);
    p = fmaf(p, s, 0.09642169624567032f);
    p = fmaf(p, s, -0.13908617198467255f);
    p = fmaf(p, s, 0.19946561753749847f);
    p = fmaf(p, s, -0.33329859375953674f);
    p = fmaf(p, s, 0.9999993443489075f);
    float r = p * t;
    if (ay > ax) r = 1.5707963267948966f - r;
    if (x < 0.f) r = 3.141592653589793f - r;
    return copysignf(r, y);
}

// The steps of the extraction that the block rule's final update repeats
// for the blocks it visits when the remainders are not stored; written
// without contracted products so that both kernels get the same bits.
// Factored geometric phase of one wavelength from the planes: the last
// group's part PL = a vl and the total PHI = a vr + PL.
__device__ __forceinline__ void planes_phase(double a, double vr, double vl, double& phi, double& pl) {
#pragma clang fp contract(off)
    pl = a * vl;
    phi = fma(a, vr, pl);
}
// arg of the field rotated by the last group's screen, not yet in (-pi, pi]
__device__ __forceinline__ double rotated_arg(float2 E, double rot) {
#pragma clang fp contract(off)
    return (double)atan2_poly(E.y, E.x) + rot;
}
// the block rule's whole turns n = rint((PHI - ph) / 2 pi) and FP64
// remainder r = PHI - ph - 2 pi n
__device__ __forceinline__ double block_turns(double phi, double ph, double& r) {
#pragma clang fp contract(off)
    const double dd = phi - ph;
    const double n = rint(dd * 0.15915494309189533577);
    r = fma(-6.283185307179586477, n, dd);
    return n;
}

// Inclusive prefix sum over the 64 lanes of a wave: shifts by 1, 2, 4, 8
// lanes within each 16-lane row, then row 0's and row 2's last lane added to
// rows 1 and 3 (row_bcast:15) and lane 31 added to rows 2 and 3
// (row_bcast:31).  Lanes a shift leaves without a source add zero.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_shift_add(float x) {
    const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false);
    return x + __int_as_float(y);
}
__device__ __forceinline__ float wave_inclusive_sum(float x) {
    x = dpp_shift_add<0x111, 0xf>(x);  // row_shr:1
    x = dpp_shift_add<0x112, 0xf>(x);  // row_shr:2
    x = dpp_shift_add<0x114, 0xf>(x);  // row_shr:4
    x = dpp_shift_add<0x118, 0xf>(x);  // row_shr:8
    x = dpp_shift_add<0x142, 0xa>(x);  // row_bcast:15
    x = dpp_shift_add<0x143, 0xc>(x);  // row_bcast:31
    return x;
}

// ------------------------------------------------------------------ extraction
// One workgroup per tile of 4 rows x 64 pixels (one wave per row, one pixel
// per lane) and direction x wavelength chunk of <= 8.  For every layer the 7
// master rows under the tile are staged in LDS, each x-filtered value (FP64)
// is made once and shared by the 4 rows that read it, and the pixel's sample
// is shared by the chunk's wavelengths: geometric phase PHI_w = sum_l s_wl v_l
// and the screen of the last layer group, applied here after the last
// Fresnel gap.
//
// MODE 0 does all of that.  When the phase multipliers factor as
// s_wl = a_w b_l, the sums need the layers only through two planes per
// direction, Vr = sum b_l v_l over the layers before the last group and Vl
// over the last group: k_geo_planes writes them, and MODE 2 reads them
// instead of sampling (PHI_w = a_w Vr + a_w Vl).
constexpr int EXT_W = 64;              // pixels per tile row (one wave)
constexpr int EXT_R = 4;               // rows per tile
constexpr int EXT_T = EXT_W * EXT_R;   // threads
constexpr int EXT_SR = EXT_R + 3;      // staged master rows
constexpr int EXT_SW = EXT_W + 4;      // staged row stride (floats)
constexpr int EXT_LC = 5;              // layers staged per LDS fill
constexpr int EXT_BC = 8;              // block columns a tile row can touch (blocks are >= 20 pixels wide)
template <bool FULL, int MODE>
__global__ void __launch_bounds__(EXT_T) k_extract(ExtractArgs a) {
    static_assert(MODE == 0 || MODE == 2, "sampling or planes");
    constexpr bool SAMPLE = MODE == 0;
    __shared__ float S[EXT_LC][EXT_SR][EXT_SW];   // staged master rows
    __shared__ double H[EXT_LC][EXT_SR][EXT_W];   // their x-filtered values
    __shared__ double bs[2][8][2][EXT_BC];        // [block row - jb0][w][cos, sin][block column - b0]
    __shared__ unsigned bm[2][8][EXT_BC];         // bits of the block's largest |r|, same indices
    __shared__ LayerSample sls[MAX_LAYERS];       // this direction's layer samples
    __shared__ double ss[MAX_LAYERS][8];          // phase scales of this chunk's wavelengths (0 past Wc)
    __shared__ int ysq[MAX_LAYERS], xsq[MAX_LAYERS];  // wrapped first staged row / column per layer
    __shared__ int um[8];                         // unwrap rule of each wavelength of the chunk
    const int nchunks = (a.W + 7) / 8;
    const int d = blockIdx.y / nchunks, chunk = blockIdx.y - d * nchunks;
    const int w_begin = chunk * 8;
    const int Wc = min(8, a.W - w_begin);
    const int nrows = FULL ? a.N : a.Nout, ncols = nrows;
    // tiles: with a multiple of 8 tile columns each XCD takes a stripe of
    // adjacent tile columns and walks it down the rows, so the master lines
    // that neighbouring tiles share (3 rows below, 3 columns right) are read
    // through one L2 close together in time
    const int ntx = (ncols + EXT_W - 1) / EXT_W;
    int tx, ty;
    if ((ntx & 7) == 0) {
        const int sw = ntx >> 3, idx = blockIdx.x >> 3;
        ty = idx / sw;
        tx = (blockIdx.x & 7) * sw + (idx - ty * sw);
    } else {
        ty = blockIdx.x / ntx;
        tx = blockIdx.x - ty * ntx;
    }
    // wave index (uniform: the tile row and the staging loops stay scalar)
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int o = (a.N - a.Nout) / 2;
    const int r0 = ty * EXT_R, row = r0 + wv;
    const int c0 = tx * EXT_W;
    const bool valid = row < nrows && c0 + lane < ncols;
    const int jf0 = FULL ? r0 : r0 + o;  // full-grid row of the tile's first row
    const int jf = jf0 + wv;
    const int jj = jf - o;
    const bool rowin = row < nrows && jj >= 0 && jj < a.Nout;
    const int xfirst = FULL ? c0 : c0 + o;  // full-grid x of the tile's first pixel
    const int ifl = xfirst + lane;
    const int ii = ifl - o;
    const bool colin = valid && ii >= 0 && ii < a.Nout;
    const int xs = a.xs;
    const int jj2 = rowin ? (int)((double)jj / a.Nout * xs) : -1;
    // block row of the tile's first row in the output (the 4 rows span at
    // most 2 block rows) and block column of its first pixel in the output
    const int jjs = max(jf0 - o, 0), iis = max(xfirst - o, 0);
    const int jb0 = (xs > 0 && jjs < a.Nout) ? (int)((double)jjs / a.Nout * xs) : 0;
    const int b0 = (xs > 0 && iis < a.Nout) ? (int)((double)iis / a.Nout * xs) : 0;
    const int jb = jj2 - jb0;
    const bool blockrow = rowin && jj2 >= 0 && jj2 < xs && jb >= 0 && jb < 2 && a.amplitude && a.bsum != nullptr;
    // block column of this pixel (AT.c:3611: (long)(1.0*ii/Nout*xs)); -1 if it adds to no block
    const int ii2 = (blockrow && colin) ? a.blk[ii] : -1;
    const int L = a.all.L;
    const int M = (int)a.all.M;
    // per-frame layer parameters, read once into LDS (global loads of uniform
    // values would otherwise be issued per lane inside the layer loop)
    if (SAMPLE) {
        for (int k = threadIdx.x; k < L; k += EXT_T) {
            const LayerSample p = a.all.ls[d * L + k];
            sls[k] = p;
            ysq[k] = mod_m(p.y0 + jf0 - 1, M);
            xsq[k] = mod_m(p.x0 + xfirst - 1, M);
        }
        for (int k = threadIdx.x; k < 8 * L; k += EXT_T) {
            const int w = k / L, l = k - w * L;
            ss[l][w] = w < Wc ? a.all.s[(w_begin + w) * L + l] : 0.0;
        }
    }
    {
        if (threadIdx.x < 8) um[threadIdx.x] = threadIdx.x < Wc ? a.unwrap_mode[w_begin + threadIdx.x] : 0;
        for (int k = threadIdx.x; k < 2 * 8 * 2 * EXT_BC; k += EXT_T) (&bs[0][0][0][0])[k] = 0.0;
        for (int k = threadIdx.x; k < 2 * 8 * EXT_BC; k += EXT_T) (&bm[0][0][0])[k] = 0u;
    }
    __syncthreads();
    // the fields are needed only after the layers: issue their loads first
    float2 Ein[8];
#pragma unroll
    for (int w = 0; w < 8; w++)
        Ein[w] = (valid && w < Wc && a.fields)
                     ? a.fields[((size_t)(d * a.W + w_begin + w) * a.N + jf) * a.N + ifl]
                     : make_float2(1.f, 0.f);
    double PHI[8], PL[8];
#pragma unroll
    for (int w = 0; w < 8; w++) { PHI[w] = 0.0; PL[w] = 0.0; }
    // planes of the factored sums: [D][N][N] at full-grid positions
    const size_t gi = ((size_t)d * a.N + jf) * a.N + ifl;
    if (MODE == 2 && valid) {
        const double vr = a.vr[gi], vl = a.vl[gi];
#pragma unroll
        for (int w = 0; w < 8; w++) planes_phase(w < Wc ? a.geo_a[w_begin + w] : 0.0, vr, vl, PHI[w], PL[w]);
    }
    for (int l0 = 0; SAMPLE && l0 < L; l0 += EXT_LC) {
        const int nl = min(EXT_LC, L - l0);
        __syncthreads();  // the previous fill's values are read
        // lane c of wave wv stages column c of the 7 rows of layers wv, wv + 4
        for (int q = wv; q < nl; q += EXT_R) {
            const int ys = __builtin_amdgcn_readfirstlane(ysq[l0 + q]);
            const int x = wrap_col(__builtin_amdgcn_readfirstlane(xsq[l0 + q]) + lane, M);
            const float* m = a.all.masters + (size_t)(l0 + q) * M * M;
            float t[EXT_SR];
#pragma unroll
            for (int r = 0; r < EXT_SR; r++) t[r] = m[(size_t)wrap_col(ys + r, M) * M + x];
#pragma unroll
            for (int r = 0; r < EXT_SR; r++) S[q][r][lane] = t[r];
        }
        // the 3 columns right of the tile, one element per thread
        for (int e = threadIdx.x; e < nl * EXT_SR * 3; e += EXT_T) {
            const int qr = e / 3, c = EXT_W + (e - qr * 3);
            const int q = qr / EXT_SR, r = qr - q * EXT_SR;
            S[q][r][c] = a.all.masters[(size_t)(l0 + q) * M * M + (size_t)wrap_col(ysq[l0 + q] + r, M) * M +
                                       wrap_col(xsq[l0 + q] + c, M)];
        }
        __syncthreads();
        // x-filtered values: wave wv makes staged (layer, row) pairs wv, wv + 4, ...
        for (int k = wv; k < nl * EXT_SR; k += EXT_R) {
            const int q = k / EXT_SR, r = k - q * EXT_SR;  // uniform
            const LayerSample& p = sls[l0 + q];
            const float* s = &S[q][r][lane];
            double h = p.wx[0] * (double)s[0];
            h = fma(p.wx[1], (double)s[1], h);
            h = fma(p.wx[2], (double)s[2], h);
            H[q][r][lane] = fma(p.wx[3], (double)s[3], h);
        }
        __syncthreads();
        if (valid) {
            // the pixel's FP64 sample of staged layer q: its 4 x-filtered rows
            auto sample = [&](int q) {
                const LayerSample& p = sls[l0 + q];
                double v = 0.0;
#pragma unroll
                for (int b = 0; b < 4; b++) v = fma(p.wy[b], H[q][wv + b][lane], v);
                return v;
            };
            // staged layers lo .. hi - 1 belong to the last group (uniform
            // bounds).  Two loops with straight bodies: one branch per layer
            // would make the compiler copy the accumulators it does not
            // update.  Wavelengths past Wc have zero scales: no per-wavelength
            // tests.
            const int lo = min(max(a.last_first - l0, 0), nl);
            const int hi = min(max(a.last_first + a.last_count - l0, lo), nl);
            for (int k = 0; k < nl - (hi - lo); k++) {
                const int q = k < lo ? k : k + (hi - lo);
                const double v = sample(q);
                const double* sl = ss[l0 + q];
#pragma unroll
                for (int w = 0; w < 8; w++) PHI[w] = fma(sl[w], v, PHI[w]);
            }
            for (int q = lo; q < hi; q++) {
                const double v = sample(q);
                const double* sl = ss[l0 + q];
#pragma unroll
                for (int w = 0; w < 8; w++) PL[w] = fma(sl[w], v, PL[w]);
            }
        }
    }
    // the total geometric phase: the other layers' sum plus the last group's
    // (from the planes it is the total already)
    if (MODE == 0) {
#pragma unroll
        for (int w = 0; w < 8; w++) PHI[w] += PL[w];
    }
    const double inv2pi = 0.15915494309189533577;
    // block column of lane + 1 (-1: none, or the last lane) and whether this
    // lane ends a run of one block column
    int next_key = __shfl_down(ii2, 1, 64);
    const bool run_end = ii2 >= 0 && (lane == 63 || next_key != ii2);
    if (lane == 63 || next_key == ii2) next_key = -1;
    // unwrap rules of the chunk as a uniform bit mask (bit w: block rule)
    int umask = 0;
#pragma unroll
    for (int w = 0; w < 8; w++) umask |= (um[w] != 0) << w;
    umask = __builtin_amdgcn_readfirstlane(umask);
    // output element of the chunk's first field; the next field is a plane further
    const size_t plane = (size_t)a.Nout * a.Nout;
    const size_t o0 = ((size_t)(d * a.W + w_begin) * a.Nout + (rowin ? jj : 0)) * a.Nout + (colin ? ii : 0);
#pragma unroll
    for (int w = 0; w < 8; w++) {
        if (w >= Wc) continue;
        const int f = d * a.W + w_begin + w;
        const bool block_rule = (umask >> w) & 1;
        float cx = 0.f, cy = 0.f;  // this pixel's contribution to its block sums
        if (valid) {
            // The last group's screen only rotates the field: arg(E e^{i PL})
            // is arg E + PL modulo 2 pi and |E| does not change, so the
            // rotation is applied to the FP64 phase (one add) unless the
            // rotated field itself is an output.
            float2 E = Ein[w];
            double rot = 0.0;
            if (a.amplitude && a.last_count > 0) {
                if (FULL && a.field_out) E = cmul(E, expi_d(PL[w]));
                else rot = PL[w];
            }
            if (FULL && a.field_out) a.field_out[((size_t)f * a.N + jf) * a.N + ifl] = E;
            if (rowin && colin) {
                const size_t oi = o0 + (size_t)w * plane;
                if (!a.amplitude) {
                    a.pha[oi] = (float)PHI[w];
                    a.amp[oi] = 1.f;
                } else {
                    // |E| with the hardware square root (1 ulp); arg E with
                    // the polynomial arctangent (< 4e-7 rad)
                    a.amp[oi] = __builtin_amdgcn_sqrtf(E.x * E.x + E.y * E.y);
                    // arg of the rotated field, not yet brought into (-pi, pi]
                    const double ph = rotated_arg(E, rot);
                    if (!block_rule) {
                        // this rule adds whole turns of PHI alone to the phase
                        // in (-pi, pi] (AT.c:3456), so the rotated phase is
                        // wrapped first
                        const double phw = fma(-6.283185307179586477, rint(ph * inv2pi), ph);
                        a.pha[oi] = (float)(phw + 2.0 * M_PI * unwrap_count(PHI[w] * inv2pi));
                    } else {
                        // whole turns of PHI - atan2 go into the phase now, the
                        // FP64 remainder (rounded to float) waits for the
                        // block's peak; the final count of the block rule is
                        // n + count((r - peak) / 2 pi)
                        double r;
                        const double n = block_turns(PHI[w], ph, r);
                        a.pha[oi] = (float)fma(6.283185307179586477, n, ph);
                        const float rf = (float)r;
                        if (a.dres) a.dres[oi] = rf;
                        if (ii2 >= 0) {
                            // contribution to the block's mean direction,
                            // hardware sin/cos (the block sums only feed the
                            // atan2 of a 400-pixel sum)
                            cx = __cosf(rf);
                            cy = __sinf(rf);
                            // the block's largest |r| (the bits of a
                            // non-negative float order like the float)
                            if (a.rmax && ii2 - b0 < EXT_BC)
                                atomicMax(&bm[jb][w][ii2 - b0], __float_as_uint(fabsf(rf)));
                        }
                    }
                }
            }
        }
        // per-block sums over the wave (one tile row): with P the wave's
        // inclusive prefix sum, a run of lanes s..e of one block sums to
        // P[e] - P[s-1].  Block columns are non-decreasing along the valid
        // lanes and lanes outside every block add zero, so the last lane e of
        // each run adds P[e] to its block and subtracts it from the block of
        // lane e + 1.
        if (blockrow && block_rule) {
            const double px = (double)wave_inclusive_sum(cx), py = (double)wave_inclusive_sum(cy);
            if (run_end) {
                if (ii2 - b0 < EXT_BC) {
                    atomicAdd(&bs[jb][w][0][ii2 - b0], px);
                    atomicAdd(&bs[jb][w][1][ii2 - b0], py);
                }
                if (next_key >= 0 && next_key - b0 < EXT_BC) {
                    atomicAdd(&bs[jb][w][0][next_key - b0], -px);
                    atomicAdd(&bs[jb][w][1][next_key - b0], -py);
                }
            }
        }
    }
    __syncthreads();
    if (xs <= 0 || !a.amplitude || a.bsum == nullptr) return;
    const int nb = xs * xs;
    for (int k = threadIdx.x; k < 2 * 8 * EXT_BC; k += EXT_T) {
        const int j = k / (8 * EXT_BC), w = (k / EXT_BC) & 7, lb = k % EXT_BC;
        if (w >= Wc || um[w] == 0 || jb0 + j >= xs || b0 + lb >= xs) continue;
        if (bs[j][w][0][lb] == 0.0 && bs[j][w][1][lb] == 0.0) continue;
        const int f = d * a.W + w_begin + w;
        double* dst = a.bsum + ((size_t)f * nb + (size_t)(jb0 + j) * xs + b0 + lb) * 2;
        atomicAdd(dst, bs[j][w][0][lb]);
        atomicAdd(dst + 1, bs[j][w][1][lb]);
    }
    if (a.rmax == nullptr) return;
    for (int k = threadIdx.x; k < 2 * 8 * EXT_BC; k += EXT_T) {
        const int j = k / (8 * EXT_BC), w = (k / EXT_BC) & 7, lb = k % EXT_BC;
        if (w >= Wc || um[w] == 0 || jb0 + j >= xs || b0 + lb >= xs || bm[j][w][lb] == 0u) continue;
        const int f = d * a.W + w_begin + w;
        atomicMax(a.rmax + (size_t)f * nb + (size_t)(jb0 + j) * xs + b0 + lb, bm[j][w][lb]);
    }
}

// The factored geometric phase's planes for the extraction's MODE 2, and
// optionally (vq) the column passes' phase planes of the layers slot_layer
// in the same sweep (k_layer_phase's values: the same samples in FP64
// quarter turns), over the whole grid:
// vr / vl [D][N][N] (full-grid positions) = sum_l b_l v_l over the layers
// outside the last group [lf, lf + lc) and over that group, for full-grid
// rows r_first .. r_first + nrows - 1 and columns c_first .. + ncols - 1.
// A workgroup makes 8 rows x 256 columns: per layer it stages the 11 master
// rows under them (LDS, 11 KB; the next layer's loads issued before this
// layer's arithmetic) and each thread makes its column's 11 x-filtered
// values once and its 8 samples (the taps and order of the extraction's
// sampling).  With a multiple of 8 tile columns each XCD walks its stripe of
// tile columns down the rows (tiles share 3 master rows).
constexpr int GEO_W = 256, GEO_R = 8, GEO_SR = GEO_R + 3;
__global__ void __launch_bounds__(GEO_W)
k_geo_planes(double* __restrict__ vr, double* __restrict__ vl, const float* __restrict__ masters,
             const LayerSample* __restrict__ ls, const double* __restrict__ geo_b, int N, int M, int L, int lf, int lc,
             int r_first, int nrows, int c_first, int ncols, double* __restrict__ vq,
             const int* __restrict__ slot_layer, int nslots, double* __restrict__ zero_p, long zero_n) {
    __shared__ float S[GEO_SR][GEO_W + 4];
    // the extraction's block sums of this frame start from zero (cleared here
    // rather than by a launch of their own)
    for (long e = ((long)blockIdx.y * gridDim.x + blockIdx.x) * GEO_W + threadIdx.x; e < zero_n;
         e += (long)gridDim.x * gridDim.y * GEO_W)
        zero_p[e] = 0.0;
    const int ntx = (ncols + GEO_W - 1) / GEO_W;
    int tx, ty;
    if ((ntx & 7) == 0) {
        const int sw = ntx >> 3, idx = blockIdx.x >> 3;
        ty = idx / sw;
        tx = (blockIdx.x & 7) * sw + (idx - ty * sw);
    } else {
        ty = blockIdx.x / ntx;
        tx = blockIdx.x - ty * ntx;
    }
    const int d = blockIdx.y, t = threadIdx.x;
    const int jf0 = r_first + ty * GEO_R, cf0 = c_first + tx * GEO_W;
    const bool colin = tx * GEO_W + t < ncols;
    double acc[GEO_R];
#pragma unroll
    for (int q = 0; q < GEO_R; q++) acc[q] = 0.0;
    // layer order: the layers outside the last group, then the group
    const int nrest = L - lc;
    auto layer_at = [&](int k) { return k < nrest ? (k < lf ? k : k + lc) : lf + (k - nrest); };
    // this thread's staged values of one layer (its column of the rows, and
    // one of the 3 columns right of the tile for t < 3 GEO_SR), loaded one
    // layer ahead of their use
    float v[GEO_SR], vt = 0.f;
    auto fetch = [&](int l) {
        const LayerSample& p = ls[d * L + l];
        const int ys = mod_m(p.y0 + jf0 - 1, M), xs0 = mod_m(p.x0 + cf0 - 1, M);  // uniform
        const float* m = masters + (size_t)l * M * M;
        const int x = wrap_col(xs0 + t, M);
#pragma unroll
        for (int r = 0; r < GEO_SR; r++) v[r] = m[(size_t)wrap_col(ys + r, M) * M + x];
        if (t < GEO_SR * 3) {
            const int r = t / 3, c = GEO_W + (t - r * 3);
            vt = m[(size_t)wrap_col(ys + r, M) * M + wrap_col(xs0 + c, M)];
        }
    };
    // stage the fetched layer k, fetch layer k + 1, then this column's
    // GEO_R samples of layer k into acc
    auto layer = [&](int k, double(&acc)[GEO_R]) {
        const int l = layer_at(k);
        __syncthreads();  // the previous layer's staged rows are read
#pragma unroll
        for (int r = 0; r < GEO_SR; r++) S[r][t] = v[r];
        if (t < GEO_SR * 3) S[t / 3][GEO_W + t % 3] = vt;
        __syncthreads();
        if (k + 1 < L) fetch(layer_at(k + 1));
        const LayerSample& p = ls[d * L + l];
        double h[GEO_SR];
#pragma unroll
        for (int r = 0; r < GEO_SR; r++) {
            const float* q = &S[r][t];
            double x = p.wx[0] * (double)q[0];
            x = fma(p.wx[1], (double)q[1], x);
            x = fma(p.wx[2], (double)q[2], x);
            h[r] = fma(p.wx[3], (double)q[3], x);
        }
        const double b = geo_b[l];
        // the phase plane of a layer that a column pass applies, if this
        // launch makes them too (uniform)
        int slot = -1;
        for (int q = 0; vq != nullptr && q < nslots; q++)
            if (slot_layer[q] == l) slot = q;
        double* pq = slot >= 0 ? vq + (((size_t)slot * gridDim.y + d) * N + jf0) * N + cf0 + t : nullptr;
#pragma unroll
        for (int rr = 0; rr < GEO_R; rr++) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) s = fma(p.wy[q], h[rr + q], s);
            acc[rr] = fma(b, s, acc[rr]);
            if (pq != nullptr && colin && ty * GEO_R + rr < nrows) pq[(size_t)rr * N] = s * 0.63661977236758134308;
        }
    };
    // the sums of this column's rows into plane pl, then restarted
    auto put = [&](double* pl) {
#pragma unroll
        for (int rr = 0; rr < GEO_R; rr++) {
            if (colin && ty * GEO_R + rr < nrows) pl[((size_t)d * N + jf0 + rr) * N + cf0 + t] = acc[rr];
            acc[rr] = 0.0;
        }
    };
    if (L > 0) fetch(layer_at(0));
    // the layers before the last group, then the group, one accumulator set
    for (int k = 0; k < nrest; k++) layer(k, acc);
    put(vr);
    for (int k = nrest; k < L; k++) layer(k, acc);
    put(vl);
}

hipError_t launch_extract_geometry(const ExtractArgs& a, bool full, hipStream_t st, double* vq,
                                   const int* slot_layer, int nslots, double* zero_p, long zero_n) {
    if (a.vr == nullptr || a.vl == nullptr || a.geo_b == nullptr) return hipErrorInvalidValue;
    if (vq != nullptr && (slot_layer == nullptr || nslots <= 0)) return hipErrorInvalidValue;
    // the column passes' phase planes cover the whole grid
    const bool whole = full || vq != nullptr;
    const int n = whole ? a.N : a.Nout, first = whole ? 0 : (a.N - a.Nout) / 2;
    const long tiles = (long)((n + GEO_W - 1) / GEO_W) * ((n + GEO_R - 1) / GEO_R);
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    dim3 grid((unsigned)tiles, a.D), block(GEO_W);
    k_geo_planes<<<grid, block, 0, st>>>(a.vr, a.vl, a.all.masters, a.all.ls, a.geo_b, a.N, (int)a.all.M, a.all.L,
                                         a.last_first, a.last_count, first, n, first, n, vq, slot_layer, nslots,
                                         zero_p, zero_p ? zero_n : 0);
    return hipGetLastError();
}

hipError_t launch_extract(const ExtractArgs& a, bool full, hipStream_t st) {
    const int n = full ? a.N : a.Nout;
    const long tiles = (long)((n + EXT_W - 1) / EXT_W) * ((n + EXT_R - 1) / EXT_R);
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    dim3 grid((unsigned)tiles, a.D * ((a.W + 7) / 8)), block(EXT_T);
    if (a.vr == nullptr) {
        if (full) k_extract<true, 0><<<grid, block, 0, st>>>(a);
        else k_extract<false, 0><<<grid, block, 0, st>>>(a);
    } else {
        if (full) k_extract<true, 2><<<grid, block, 0, st>>>(a);
        else k_extract<false, 2><<<grid, block, 0, st>>>(a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ block rule
// The twelve neighbour relations of the block fix-up (AT.c:3667-3903): block
// (ii2+dx1, jj2+dy1) vs (ii2+dx2, jj2+dy2) for ii2 < xs-xm, jj2 < ys-ym.
__constant__ int c_pairs[12][6] = {
    {0, 0, 1, 0, 1, 0}, {0, 0, 0, 1, 0, 1}, {0, 0, 1, 1, 1, 1}, {0, 1, 1, 0, 1, 1},
    {0, 0, 2, 0, 2, 0}, {0, 2, 0, 0, 0, 2}, {1, 2, 0, 0, 1, 2}, {0, 2, 1, 0, 1, 2},
    {2, 2, 0, 0, 2, 2}, {0, 2, 2, 0, 2, 2}, {0, 1, 2, 0, 2, 1}, {2, 1, 0, 0, 2, 1}};

// the fix-up count of block b as the reference's float image holds it
// (WFprecision = 0, AT.c:3667-3903): +-0.2 / +-0.15 for every neighbour
// relation whose peaks differ by more than pi.  The steps reach b relation by
// relation and, within a relation, in the loop order (ii2 outer, jj2 inner)
// of the pair's base block; each step is a double add rounded to float, so
// the order shows in the last bits (a count that is 0.5 in exact arithmetic
// can come out at 0.50000006 and pass plim = 0.5).
template <class Peak>
__device__ __forceinline__ float block_count_at(Peak peak, int i2, int j2, int xs) {
    const double pb = peak(i2, j2);
    float c = 0.f;
    for (int q = 0; q < 12; q++) {
        const int dx1 = c_pairs[q][0], dy1 = c_pairs[q][1], dx2 = c_pairs[q][2], dy2 = c_pairs[q][3];
        const int xm = c_pairs[q][4], ym = c_pairs[q][5];
        const double cq = q < 4 ? 0.2 : 0.15;
        // b as the second block has the smaller base when the pair's offset
        // (d2 - d1) is positive in x, or zero in x and positive in y
        const int ox = dx2 - dx1, oy = dy2 - dy1;
        const bool second_first = ox > 0 || (ox == 0 && oy > 0);
        for (int t = 0; t < 2; t++) {
            const bool as_first = (t == 0) ? !second_first : second_first;
            const int x = i2 - (as_first ? dx1 : dx2), y = j2 - (as_first ? dy1 : dy2);
            if (x < 0 || y < 0 || x >= xs - xm || y >= xs - ym) continue;
            const double po = as_first ? peak(x + dx2, y + dy2) : peak(x + dx1, y + dy1);
            const double pv1 = as_first ? pb : po, pv2 = as_first ? po : pb;
            if (pv2 > pv1 + M_PI) c = (float)((double)c + (as_first ? cq : -cq));
            if (pv1 > pv2 + M_PI) c = (float)((double)c + (as_first ? -cq : cq));
        }
    }
    return c;
}
__device__ __forceinline__ float block_count(const double* B, int b, int xs) {
    return block_count_at([&](int x, int y) { return B[y * xs + x]; }, b % xs, b / xs, xs);
}

// The first round of the fix-up, spread over the whole GPU (one round is all
// that realistic fields take): the peaks atan2 of the block sums, each
// block's count and, per field, the largest |count| as the bits of a
// non-negative float (ordered like unsigned integers; cmax is zero before).
// A workgroup takes a 16 x 16 tile of blocks: it makes the peaks of the tile
// and of the 2 blocks around it (every relation reaches 2 blocks away) into
// LDS, writes its own, and counts from LDS.
constexpr int PC_T = 16, PC_H = PC_T + 4;
__global__ void __launch_bounds__(PC_T * PC_T)
k_block_peaks_counts(const double* __restrict__ bsum, double* __restrict__ bin, double* __restrict__ ch,
                     const int* __restrict__ flist, int xs, unsigned* __restrict__ cmax) {
    __shared__ double T[PC_H][PC_H];
    __shared__ float wmax[PC_T * PC_T / 64];
    const int f = flist[blockIdx.y];
    const size_t nb = (size_t)xs * xs;
    const int ntx = (xs + PC_T - 1) / PC_T;
    const int tx0 = (blockIdx.x % ntx) * PC_T, ty0 = (blockIdx.x / ntx) * PC_T;
    const double* sf = bsum + (size_t)f * nb * 2;
    for (int e = threadIdx.x; e < PC_H * PC_H; e += PC_T * PC_T) {
        const int hy = e / PC_H, hx = e - hy * PC_H;
        const int x = tx0 - 2 + hx, y = ty0 - 2 + hy;
        T[hy][hx] = (x >= 0 && y >= 0 && x < xs && y < xs)
                        ? atan2(sf[((size_t)y * xs + x) * 2 + 1], sf[((size_t)y * xs + x) * 2]) : 0.0;
    }
    __syncthreads();
    const int i = threadIdx.x % PC_T, j = threadIdx.x / PC_T;
    const int i2 = tx0 + i, j2 = ty0 + j;
    float a = 0.f;
    if (i2 < xs && j2 < xs) {
        const size_t b = (size_t)j2 * xs + i2;
        bin[(size_t)f * nb + b] = T[j + 2][i + 2];
        const float c = block_count_at([&](int x, int y) { return T[y - ty0 + 2][x - tx0 + 2]; }, i2, j2, xs);
        reinterpret_cast<float*>(ch)[(size_t)f * nb + b] = c;
        a = fabsf(c);
    }
    // the workgroup's largest |count|, then one atomic per workgroup (atomics
    // on one field's word serialise in L2)
    for (int o = 32; o > 0; o >>= 1) a = fmaxf(a, __shfl_down(a, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(&cmax[blockIdx.y], __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

__global__ void __launch_bounds__(256) k_block_counts(const double* __restrict__ bin, double* __restrict__ ch,
                                                      const int* __restrict__ flist, int xs,
                                                      unsigned* __restrict__ cmax) {
    const int f = flist[blockIdx.y];
    const int nb = xs * xs;
    const int b = blockIdx.x * 256 + threadIdx.x;
    float a = 0.f;
    if (b < nb) {
        const float c = block_count(bin + (size_t)f * nb, b, xs);
        reinterpret_cast<float*>(ch)[(size_t)f * nb + b] = c;
        a = fabsf(c);
    }
    // the workgroup's largest |count|, then one atomic per workgroup (atomics
    // on one field's word serialise in L2)
    for (int o = 32; o > 0; o >>= 1) a = fmaxf(a, __shfl_down(a, o, 64));
    __shared__ float wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(&cmax[blockIdx.y], __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

// The iterative fix-up from the first round's peaks, counts and largest
// count (k_block_peaks_counts); further rounds on one workgroup per field,
// the block grid in LDS.
__global__ void __launch_bounds__(1024) k_block_fix(double* __restrict__ bin, double* __restrict__ ch,
                                                    const int* __restrict__ flist, int xs, int use_lds,
                                                    const unsigned* __restrict__ cmax) {
    __shared__ double redd[16];
    __shared__ int redi[16];
    extern __shared__ double dyn[];  // [nb] B (double), [nb] ch (float) when use_lds
    const int f = flist[blockIdx.x];
    const int nb = xs * xs;
    double* Bg = bin + (size_t)f * nb;
    double* B = use_lds ? dyn : Bg;
    float* Cg = reinterpret_cast<float*>(ch) + (size_t)f * nb;
    float* Cc = use_lds ? reinterpret_cast<float*>(dyn + nb) : Cg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // no first-round count above 0.5: the first round changes no peak
    // (plim = max(m - 0.001, 0.5)) and ends the loop, so the peaks in bin
    // are final (uniform per workgroup)
    if (__uint_as_float(cmax[blockIdx.x]) <= 0.5f) return;
    if (use_lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            B[b] = Bg[b];
            Cc[b] = Cg[b];
        }
    __syncthreads();
    int chiter = 0, chcnt0cnt = 0;
    double plim = 1.0;
    while (plim > 0.51 && chiter < 1000 && chcnt0cnt < 5) {
        chiter++;
        double m;
        if (chiter == 1) {
            m = (double)__uint_as_float(cmax[blockIdx.x]);  // the first round's counts are in Cc
        } else {
            double mx = 0.0;
            for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                const float c = block_count(B, b, xs);
                Cc[b] = c;
                mx = fmax(mx, (double)fabsf(c));
            }
            for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
            if (lane == 0) redd[wv] = mx;
            __syncthreads();
            m = 0.0;
            for (int q = 0; q < nw; q++) m = fmax(m, redd[q]);
        }
        plim = m - 0.001;
        if (plim < 0.5) plim = 0.5;
        int cnt = 0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int i2 = b % xs, j2 = b / xs;
            const bool interior = i2 > 1 && j2 > 1 && i2 < xs - 2 && j2 < xs - 2;
            const double c = (double)Cc[b];
            if (c > plim) { if (interior) cnt++; B[b] += 2.0 * M_PI; }
            if (c < -plim) { if (interior) cnt++; B[b] -= 2.0 * M_PI; }
        }
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
        __syncthreads();
        if (lane == 0) redi[wv] = cnt;
        __syncthreads();
        int chcnt = 0;
        for (int q = 0; q < nw; q++) chcnt += redi[q];
        if (chcnt == 0) chcnt0cnt++;
        else chcnt0cnt = 0;
        __syncthreads();
    }
    if (use_lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) Bg[b] = B[b];
}

hipError_t launch_block_fix(double* bsum, double* bin, double* ch, unsigned* cmax, const int* field_list,
                            int nfields, int xs, hipStream_t st) {
    if (nfields <= 0 || xs <= 0) return hipSuccess;
    const int nb = xs * xs;
    const int nt = (xs + PC_T - 1) / PC_T;
    k_block_peaks_counts<<<dim3(nt * nt, nfields), PC_T * PC_T, 0, st>>>(bsum, bin, ch, field_list, xs, cmax);
    const size_t need = (size_t)nb * (sizeof(double) + sizeof(float));
    const int use_lds = need <= 150 * 1024;
    static bool attr = false;
    if (!attr) {
        allow_dynamic_lds((const void*)k_block_fix);
        attr = true;
    }
    k_block_fix<<<nfields, 1024, use_lds ? need : 0, st>>>(bin, ch, field_list, xs, use_lds, cmax);
    return hipGetLastError();
}

hipError_t launch_block_counts(const double* bin, float* ch, unsigned* cmax, const int* field_list, int nfields,
                               int xs, hipStream_t st) {
    if (nfields <= 0 || xs <= 0) return hipSuccess;
    dim3 grid((xs * xs + 255) / 256, nfields);
    k_block_counts<<<grid, 256, 0, st>>>(bin, reinterpret_cast<double*>(ch), field_list, xs, cmax);
    return hipGetLastError();
}

// pha = atan2 + 2 pi (n + count((r - peak) / 2 pi)) with the block's peak
// (AT.c:4240-4262: count((PHI - atan2 - peak) / 2 pi), the whole turns n of
// PHI - atan2 already in pha).  APPLY_ROWS output rows of one field per
// workgroup, their remainder loads issued together.  Only the remainder map
// is read in full: pha is read and written only where a count is not zero
// (a zero count leaves pha's bits unchanged: fma(2 pi, +0, p) = p), which
// for a block's pixels around its peak is rare.
constexpr int APPLY_ROWS = 4;
__device__ __forceinline__ double apply_count(float r, double peak) {
    // r is a float-rounded remainder, so (r - peak) / (2 pi) as a product with
    // 1 / (2 pi) changes no decision that the rounding of r leaves alone
    return unwrap_count(((double)r - peak) * 0.15915494309189533577);
}
__global__ void __launch_bounds__(256) k_block_apply(float* __restrict__ pha, const float* __restrict__ dres,
                                                     const double* __restrict__ bin, const int* __restrict__ flist,
                                                     const int* __restrict__ blk, int Nout, int xs) {
    const int f = flist[blockIdx.y];
    const int j0 = blockIdx.x * APPLY_ROWS;
    const int nr = min(APPLY_ROWS, Nout - j0);
    // every pixel lies in a block ((long)(1.0 i / Nout xs) < xs); no blocks: peak 0
    auto brow = [&](int jj) { return bin + ((size_t)f * xs + (xs > 0 ? blk[jj] : 0)) * xs; };
    if ((Nout & 3) == 0 && (reinterpret_cast<uintptr_t>(pha) & 15) == 0) {  // 16-byte rows
        for (int i4 = threadIdx.x; i4 < Nout / 4; i4 += 256) {
            const int4 b = reinterpret_cast<const int4*>(blk)[i4];
            float4 r[APPLY_ROWS];
#pragma unroll
            for (int q = 0; q < APPLY_ROWS; q++)
                if (q < nr) r[q] = reinterpret_cast<const float4*>(dres + ((size_t)f * Nout + j0 + q) * Nout)[i4];
#pragma unroll
            for (int q = 0; q < APPLY_ROWS; q++) {
                if (q >= nr) continue;
                const double* br = brow(j0 + q);
                const double n0 = apply_count(r[q].x, xs > 0 ? br[b.x] : 0.0);
                const double n1 = apply_count(r[q].y, xs > 0 ? br[b.y] : 0.0);
                const double n2 = apply_count(r[q].z, xs > 0 ? br[b.z] : 0.0);
                const double n3 = apply_count(r[q].w, xs > 0 ? br[b.w] : 0.0);
                if (n0 != 0.0 || n1 != 0.0 || n2 != 0.0 || n3 != 0.0) {
                    float4* pp = reinterpret_cast<float4*>(pha + ((size_t)f * Nout + j0 + q) * Nout) + i4;
                    float4 p = *pp;
                    p.x = (float)fma(2.0 * M_PI, n0, (double)p.x);
                    p.y = (float)fma(2.0 * M_PI, n1, (double)p.y);
                    p.z = (float)fma(2.0 * M_PI, n2, (double)p.z);
                    p.w = (float)fma(2.0 * M_PI, n3, (double)p.w);
                    *pp = p;
                }
            }
        }
        return;
    }
    for (int q = 0; q < nr; q++) {
        const double* br = brow(j0 + q);
        float* prow = pha + ((size_t)f * Nout + j0 + q) * Nout;
        const float* rrow = dres + ((size_t)f * Nout + j0 + q) * Nout;
        for (int ii = threadIdx.x; ii < Nout; ii += 256) {
            const double n = apply_count(rrow[ii], xs > 0 ? br[blk[ii]] : 0.0);
            if (n != 0.0) prow[ii] = (float)fma(2.0 * M_PI, n, (double)prow[ii]);
        }
    }
}

// The same update where it can change a phase: a block's pixels all keep
// their phase when the counts of -R and R are zero, R the largest |r| of the
// block's stored remainders (the count is non-decreasing in r), so one lane
// per block tests that from the extraction's R, and the wave then visits
// only the blocks that fail it (pixels as apply_count, in any order: each is
// independent).  bstart[b] is the first output row / column of block b,
// bstart[xs] = Nout.
__global__ void __launch_bounds__(64) k_block_apply_sparse(float* __restrict__ pha, const float* __restrict__ dres,
                                                           const double* __restrict__ bin,
                                                           const unsigned* __restrict__ rmax,
                                                           const int* __restrict__ flist,
                                                           const int* __restrict__ bstart, int Nout, int xs,
                                                           RemainderSource rs) {
    const int f = flist[blockIdx.y];
    const int nb = xs * xs;
    const int lane = threadIdx.x;
    const int b = blockIdx.x * 64 + lane;
    bool need = false;
    if (b < nb) {
        const float R = __uint_as_float(rmax[(size_t)f * nb + b]);
        const double peak = bin[(size_t)f * nb + b];
        need = apply_count(-R, peak) != 0.0 || apply_count(R, peak) != 0.0;
    }
    unsigned long long m = __ballot(need);
    while (m) {
        const int bb = blockIdx.x * 64 + (__ffsll((long long)m) - 1);
        m &= m - 1;
        const int by = bb / xs, bx = bb - by * xs;
        const int r0 = bstart[by], c0 = bstart[bx];
        const int w = bstart[bx + 1] - c0, n = (bstart[by + 1] - r0) * w;
        const double peak = bin[(size_t)f * nb + bb];
        for (int e = lane; e < n; e += 64) {
            const int q = e / w, jj = r0 + q, ii = c0 + (e - q * w);
            const size_t o = ((size_t)f * Nout + jj) * Nout + ii;
            float rf;
            if (rs.fields) {
                // the extraction's remainder again, from the final field and
                // the geometric phase's planes at the full-grid pixel
                const int dd = f / rs.W, jf = jj + rs.o, ifl = ii + rs.o;
                const size_t gi = ((size_t)dd * rs.N + jf) * rs.N + ifl;
                double phi, pl, r;
                planes_phase(rs.geo_a[f - dd * rs.W], rs.vr[gi], rs.vl[gi], phi, pl);
                const float2 E = rs.fields[((size_t)f * rs.N + jf) * rs.N + ifl];
                block_turns(phi, rotated_arg(E, rs.rot ? pl : 0.0), r);
                rf = (float)r;
            } else {
                rf = dres[o];
            }
            const double cnt = apply_count(rf, peak);
            if (cnt != 0.0) pha[o] = (float)fma(2.0 * M_PI, cnt, (double)pha[o]);
        }
    }
}

hipError_t launch_block_apply_sparse(float* pha, const float* dres, const double* bin, const unsigned* rmax,
                                     const int* field_list, const int* bstart, int nfields, int Nout, int xs,
                                     const RemainderSource& rs, hipStream_t st) {
    if (nfields <= 0 || xs <= 0) return hipSuccess;
    if (rs.fields == nullptr && dres == nullptr) return hipErrorInvalidValue;
    dim3 grid((xs * xs + 63) / 64, nfields);
    k_block_apply_sparse<<<grid, 64, 0, st>>>(pha, dres, bin, rmax, field_list, bstart, Nout, xs, rs);
    return hipGetLastError();
}

hipError_t launch_block_apply(float* pha, const float* dres, const double* bin, const int* field_list,
                              const int* blk, int nfields, int Nout, int xs, hipStream_t st) {
    if (nfields <= 0 || Nout <= 0) return hipSuccess;
    dim3 grid((Nout + APPLY_ROWS - 1) / APPLY_ROWS, nfields);
    k_block_apply<<<grid, 256, 0, st>>>(pha, dres, bin, field_list, blk, Nout, xs);
    return hipGetLastError();
}

// ------------------------------------------------------------------ layer screens
// out[(d L + l) N^2 + j N + i] = sampled screen of layer l (FP64 sampler,
// rounded to float) at the positions of the last frame: the per-layer phase
// stack the surrogate network reads.
__global__ void k_layer_screens(const float* __restrict__ masters, const LayerSample* __restrict__ ls,
                                float* __restrict__ out, int N, int M, int L) {
    const int l = blockIdx.y, d = blockIdx.z;
    const long n = (long)N * N;
    const LayerSample p = ls[d * L + l];
    const float* m = masters + (size_t)l * M * M;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int j = (int)(e / N), i = (int)(e - (long)j * N);
        out[((long)d * L + l) * n + e] = (float)sample_layer(m, M, p, i, j);
    }
}

hipError_t launch_layer_screens(const float* masters, const LayerSample* ls, float* out, long N, long M, long L,
                                long D, hipStream_t st) {
    long g = (N * N + 255) / 256;
    dim3 grid((unsigned)(g < 4096 ? g : 4096), (unsigned)L, (unsigned)D);
    k_layer_screens<<<grid, 256, 0, st>>>(masters, ls, out, (int)N, (int)M, (int)L);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fill
// zero n doubles (a kernel, not a memset, so that it can sit in a captured
// frame graph)
// ------------------------------------------------------------------ frame statistics
// Per field of one frame: sums of the output phase, its square, the intensity
// I = amp^2 and I^2 over the Nout^2 output pixels, accumulated in FP64.  The
// first kernel reduces a contiguous slice of one field per workgroup (wave
// reduction by DPP-free xor shuffles, then across the 4 waves in LDS) to
// part[f][g][4]; the second adds a field's STAT_G slices in a fixed order, so
// the sums do not depend on scheduling.
constexpr int STAT_G = 128;  // slices per field
constexpr int STAT_T = 256;

__device__ __forceinline__ double wave_sum_f64(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ void __launch_bounds__(STAT_T) k_stats_partial(const float* __restrict__ pha, const float* __restrict__ amp,
                                                          long n, double* __restrict__ part) {
    const int f = blockIdx.y, g = blockIdx.x;
    // slices of whole float4 groups, so every slice starts 16-byte aligned
    const long n4 = n >> 2;
    const long per = (n4 + STAT_G - 1) / STAT_G;
    const long q0 = min(n4, g * per), q1 = min(n4, q0 + per);
    const float* P = pha + (size_t)f * n;
    const float* A = amp + (size_t)f * n;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    const bool vec = ((n & 3) == 0) && ((reinterpret_cast<uintptr_t>(P) | reinterpret_cast<uintptr_t>(A)) & 15) == 0;
    if (vec) {
        for (long q = q0 + threadIdx.x; q < q1; q += STAT_T) {
            const float4 p = reinterpret_cast<const float4*>(P)[q];
            const float4 a = reinterpret_cast<const float4*>(A)[q];
            const double pv[4] = {p.x, p.y, p.z, p.w};
            const double av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double I = av[k] * av[k];
                s1 += pv[k];
                s2 = fma(pv[k], pv[k], s2);
                s3 += I;
                s4 = fma(I, I, s4);
            }
        }
    } else {
        // scalar pixels of the same slices; the last slice also takes the n % 4 tail
        const long e0 = q0 * 4, e1 = (g == STAT_G - 1) ? n : q1 * 4;
        for (long e = e0 + threadIdx.x; e < e1; e += STAT_T) {
            const double pv = P[e], av = A[e];
            const double I = av * av;
            s1 += pv;
            s2 = fma(pv, pv, s2);
            s3 += I;
            s4 = fma(I, I, s4);
        }
    }
    __shared__ double red[STAT_T / 64][4];
    s1 = wave_sum_f64(s1);
    s2 = wave_sum_f64(s2);
    s3 = wave_sum_f64(s3);
    s4 = wave_sum_f64(s4);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = s1;
        red[wv][1] = s2;
        red[wv][2] = s3;
        red[wv][3] = s4;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int k = 0; k < STAT_T / 64; k++) t += red[k][threadIdx.x];
        part[((size_t)f * STAT_G + g) * 4 + threadIdx.x] = t;
    }
}

__global__ void __launch_bounds__(64) k_stats_final(const double* __restrict__ part, double* __restrict__ out) {
    const int f = blockIdx.x;
    // lane = 4 g' + k: lane sums value k of slices g', g' + 16, ...
    const int k = threadIdx.x & 3, g0 = threadIdx.x >> 2;
    double t = 0.0;
    for (int g = g0; g < STAT_G; g += 16) t += part[((size_t)f * STAT_G + g) * 4 + k];
    for (int o = 32; o >= 4; o >>= 1) t += __shfl_xor(t, o, 64);
    if (threadIdx.x < 4) out[(size_t)f * 4 + k] = t;
}

hipError_t launch_stats(const float* pha, const float* amp, long F, long n, double* part, double* out,
                        hipStream_t st) {
    if (F <= 0 || n <= 0) return hipSuccess;
    k_stats_partial<<<dim3(STAT_G, (unsigned)F), STAT_T, 0, st>>>(pha, amp, n, part);
    k_stats_final<<<(unsigned)F, 64, 0, st>>>(part, out);
    return hipGetLastError();
}

__global__ void k_zero_f64(double* __restrict__ p, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0.0;
}

hipError_t launch_zero_f64(double* p, long n, hipStream_t st) {
    long g = (n + 255) / 256;
    k_zero_f64<<<(unsigned)(g < 4096 ? (g > 0 ? g : 1) : 4096), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace atsim
