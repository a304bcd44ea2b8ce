// atsim_fft.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A length-N transform (N = 2^n, 16 <= N <= 8192) is split into S stages of
// radix <= 16: the forward plan is [R0, 16, ..., 16] and the inverse plan is
// the reverse [16, ..., 16, R0] with R0 = N / 16^(S-1).  Each of the T = N/16
// threads of one transform holds 16 complex values in registers; at a stage of
// radix R it performs 16/R butterflies j = t + b*T.  Stage s (Ns = product of
// the radices before it) reads x[j + r*N/R], applies twiddles w^(r*k)
// (k = j mod Ns, w = e^{dir 2 pi i/(Ns R)}), does a radix-R DFT and writes
// y[(j/Ns)*Ns*R + k + r*Ns] (Stockham autosort, natural-order output).
//
// Register layout contract: before the first stage and after the last stage,
// register v[b*R + r] of thread t holds element t + b*T + r*N/R where R is the
// radix of that stage.  Because the inverse plan starts with the radix the
// forward plan ends with, a forward transform, an element-wise multiply and an
// inverse transform chain in registers without an LDS exchange in between.
//
// Twiddles come from a table tw[m] = e^{-2 pi i m/N} (m < N) built in FP64 on
// the host and rounded to float; the inverse uses the conjugate.
#pragma once
#include <hip/hip_runtime.h>

namespace atsim {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by e^{DIR i pi/2}: DIR=-1 -> -i, DIR=+1 -> +i
template <int DIR>
__device__ __forceinline__ float2 mul_qi(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos / sin of pi*m/8, m = 0..9
__device__ __forceinline__ float2 w16(int m, int dir) {
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f, H = 0.70710678118654752f;
    float c, s;
    switch (m) {
        case 0: c = 1.f; s = 0.f; break;
        case 1: c = C1; s = S1; break;
        case 2: c = H; s = H; break;
        case 3: c = S1; s = C1; break;
        case 4: c = 0.f; s = 1.f; break;
        case 6: c = -H; s = H; break;
        case 9: c = -C1; s = -S1; break;
        default: c = 1.f; s = 0.f; break;
    }
    return make_float2(c, dir < 0 ? -s : s);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = mul_qi<DIR>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// 8 = 2 x 4: n = n1 + 2 n2, k = 4 k1 + k2
template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    constexpr float H = 0.70710678118654752f;
    // o_k *= e^{DIR 2 pi i k / 8}
    o1 = cmul(o1, make_float2(H, DIR < 0 ? -H : H));
    o2 = mul_qi<DIR>(o2);
    o3 = cmul(o3, make_float2(-H, DIR < 0 ? -H : H));
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// 16 = 4 x 4: n = n1 + 4 n2, k = 4 k1 + k2
template <int DIR>
__device__ __forceinline__ void dft16(float2* v) {
    float2 a[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) {
        a[n1][0] = v[n1]; a[n1][1] = v[n1 + 4]; a[n1][2] = v[n1 + 8]; a[n1][3] = v[n1 + 12];
        dft4<DIR>(a[n1][0], a[n1][1], a[n1][2], a[n1][3]);
    }
    // internal twiddles w16^(n1 k2): m = 4 is -+i, m = 2 and 6 are (+-1 -+ i)/sqrt 2
    constexpr float H = 0.70710678118654752f;
#pragma unroll
    for (int n1 = 1; n1 < 4; n1++)
#pragma unroll
        for (int k2 = 1; k2 < 4; k2++) {
            const int m = n1 * k2;
            float2& x = a[n1][k2];
            if (m == 4) {
                x = mul_qi<DIR>(x);
            } else if (m == 2) {  // (1 + DIR i) H
                x = DIR < 0 ? make_float2((x.x + x.y) * H, (x.y - x.x) * H)
                            : make_float2((x.x - x.y) * H, (x.y + x.x) * H);
            } else if (m == 6) {  // (-1 + DIR i) H
                x = DIR < 0 ? make_float2((x.y - x.x) * H, -(x.x + x.y) * H)
                            : make_float2(-(x.x + x.y) * H, (x.x - x.y) * H);
            } else {
                x = cmul(x, w16(m, DIR));
            }
        }
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
        float2 b0 = a[0][k2], b1 = a[1][k2], b2 = a[2][k2], b3 = a[3][k2];
        dft4<DIR>(b0, b1, b2, b3);
        v[k2] = b0; v[k2 + 4] = b1; v[k2 + 8] = b2; v[k2 + 12] = b3;
    }
}

template <int R, int DIR>
__device__ __forceinline__ void dftR(float2* v) {
    if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<DIR>(v);
    else dft16<DIR>(v);
}

// ---------------------------------------------------------------- plans
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
template <int N>
struct Plan {
    static constexpr int LOG = ilog2c(N);
    static constexpr int S = (LOG + 3) / 4;                    // stages
    static constexpr int R0 = N >> (4 * (S - 1));              // first forward radix
    static constexpr int T = N / 16;                           // threads per transform
    static_assert((1 << LOG) == N && N >= 16 && N <= 8192, "N must be a power of two in [16, 8192]");
    // radix of stage s in direction DIR (forward [R0,16..], inverse [..16,R0])
    static constexpr int radix(int s, int dir) { return dir < 0 ? (s == 0 ? R0 : 16) : (s == S - 1 ? R0 : 16); }
    static constexpr int ns(int s, int dir) {  // product of radices of earlier stages
        int p = 1;
        for (int q = 0; q < s; q++) p *= radix(q, dir);
        return p;
    }
    static constexpr int first(int dir) { return radix(0, dir); }
    static constexpr int last(int dir) { return radix(S - 1, dir); }
    static constexpr int LDS = N + N / 16;                     // padded LDS elements per transform
    // the same for C > 1 interleaved transforms (xpad: N / 8 pad slots when
    // the forward plan starts with radix 8)
    static constexpr int LDSC = R0 == 8 ? N + N / 8 : N + N / 16;
};

// Pad shift of exchange S (after stage S) of a transform in direction DIR
// whose LDS region interleaves C transforms: one pad slot every 2^shift
// elements.  The gfx950 LDS has 32 four-byte banks and serves 32 lanes (b32)
// or 16 lanes (b64) per cycle; with C = 8 those lanes are 8 transforms x 4
// (or 2) consecutive threads, so their positions must differ mod 4 (mod 2).
// The stores of a forward plan's first radix-8 stage are 8 elements apart:
// a pad every 8 elements separates them (every 16 put two threads on one
// bank); every other exchange is conflict-free with a pad every 16.
template <int N, int DIR, int C, int S>
__device__ __forceinline__ constexpr int xpad() {
    return (C > 1 && DIR < 0 && S == 0 && Plan<N>::R0 == 8) ? 3 : 4;
}

template <int N, int DIR>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int m) {
    float2 w = tw[m];
    return DIR < 0 ? w : make_float2(w.x, -w.y);
}

// one stage's twiddle + butterfly on registers.  The R-1 twiddles w^r of a
// butterfly are composed from the table values w^1, w^2, w^4, w^8 (at most
// three complex products each) instead of R-1 table loads.  E is the number
// of elements a thread holds (16 or 32): T = N / E threads per transform,
// E / R butterflies per thread.
template <int N, int R, int NS, int DIR, int E = 16>
__device__ __forceinline__ void stage_compute(float2* v, int t, const float2* __restrict__ tw) {
    constexpr int T = N / E;
#pragma unroll
    for (int b = 0; b < E / R; b++) {
        if constexpr (NS > 1) {
            const int j = t + b * T;
            const int k = j % NS;
            constexpr int STEP = N / (NS * R);
            float2 p[4];  // w^1, w^2, w^4, w^8
#pragma unroll
            for (int e = 0; e < 4; e++)
                if ((1 << e) < R) p[e] = twiddle<N, DIR>(tw, (1 << e) * k * STEP);
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = make_float2(1.f, 0.f);
                bool first = true;
#pragma unroll
                for (int e = 3; e >= 0; e--)
                    if (r & (1 << e)) {
                        w = first ? p[e] : cmul(w, p[e]);
                        first = false;
                    }
                v[b * R + r] = cmul(v[b * R + r], w);
            }
        }
        dftR<R, DIR>(v + b * R);
        // with 32 elements per thread, one butterfly at a time keeps the
        // temporaries of a single butterfly live
        if constexpr (E > 16) __builtin_amdgcn_sched_barrier(0);
    }
}

// Per-stage twiddle tables.  Stage s >= 1 of a plan (radix R, NS = product of
// the earlier radices, STEP = N / (NS R)) reads w^(2^e k STEP) for e with
// 2^e < R and k = j mod NS.  Gathered into [e][k] order they are contiguous in
// k, so that the lanes of a wave, whose k are consecutive, read whole cache
// lines (the N-entry table read at stride 2^e STEP touches up to 2^e times
// as many).  The tables follow the N entries of tw[]: forward stages first,
// then inverse, each stage's [e][k] block in stage order.
__host__ __device__ constexpr int tw_radix(int n, int s, int dir) {
    const int S = (ilog2c(n) + 3) / 4, R0 = n >> (4 * (S - 1));
    return dir < 0 ? (s == 0 ? R0 : 16) : (s == S - 1 ? R0 : 16);
}
__host__ __device__ constexpr int tw_ns(int n, int s, int dir) {
    int p = 1;
    for (int q = 0; q < s; q++) p *= tw_radix(n, q, dir);
    return p;
}
__host__ __device__ constexpr int tw_terms(int r) { return r >= 16 ? 4 : r >= 8 ? 3 : r >= 4 ? 2 : 1; }
__host__ __device__ constexpr int tw_stage_size(int n, int s, int dir) {
    return s == 0 ? 0 : tw_terms(tw_radix(n, s, dir)) * tw_ns(n, s, dir);
}
// offset of stage s of direction dir in the tables (from tw + n)
__host__ __device__ constexpr int tw_stage_off(int n, int dir, int s) {
    const int S = (ilog2c(n) + 3) / 4;
    int o = 0;
    if (dir > 0)
        for (int q = 0; q < S; q++) o += tw_stage_size(n, q, -1);
    for (int q = 0; q < s; q++) o += tw_stage_size(n, q, dir);
    return o;
}
__host__ __device__ constexpr int tw_tables_size(int n) { return tw_stage_off(n, +1, (ilog2c(n) + 3) / 4); }

// stage_compute with the twiddles from the stage's table (tw: the N-entry
// table followed by the per-stage tables): w^1, w^2, w^4, w^8 loaded, the
// others composed as stage_compute does -- the same values, so the same
// results.  (Loading all R - 1 twiddles of a butterfly from [r][k] tables
// instead saves the 11 products but costs more than it saves: 2048^2 fused
// row pass 272 us against 242 us.)
template <int N, int R, int NS, int DIR, int E, int S>
__device__ __forceinline__ void stage_compute_tab(float2* v, int t, const float2* __restrict__ tw) {
    constexpr int T = N / E;
    constexpr int OFF = N + tw_stage_off(N, DIR, S);
#pragma unroll
    for (int b = 0; b < E / R; b++) {
        if constexpr (NS > 1) {
            const int k = (t + b * T) % NS;
            float2 p[4];  // w^1, w^2, w^4, w^8
#pragma unroll
            for (int e = 0; e < 4; e++)
                if ((1 << e) < R) p[e] = twiddle<N, DIR>(tw, OFF + e * NS + k);
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = make_float2(1.f, 0.f);
                bool first = true;
#pragma unroll
                for (int e = 3; e >= 0; e--)
                    if (r & (1 << e)) {
                        w = first ? p[e] : cmul(w, p[e]);
                        first = false;
                    }
                v[b * R + r] = cmul(v[b * R + r], w);
            }
        }
        dftR<R, DIR>(v + b * R);
        if constexpr (E > 16) __builtin_amdgcn_sched_barrier(0);
    }
}

// Padded LDS positions (one pad slot every 2^PS elements) as one per-thread
// base plus compile-time offsets (immediate offsets of the LDS instructions,
// no per-element address registers).  For every plan here (all N, E = 16 or
// 32, every stage, PS = 4, and PS = 3 after a first radix-8 stage) the pad
// of a store position P0 + r NS is pad(P0) + pad(r NS), and the pad of a
// load position t + m (m = b T + r N / R) is pad(t) + pad(m): the carries
// never cross a 2^PS-element boundary.
template <int N, int R, int NS, int E, int PS = 4>
__device__ __forceinline__ int store_base(int t, int b) {
    constexpr int T = N / E;
    const int j = t + b * T;
    const int p0 = (j / NS) * NS * R + j % NS;
    return p0 + (p0 >> PS);
}
template <int R, int NS, int PS = 4>
__device__ __forceinline__ constexpr int store_off(int r) { return r * NS + ((r * NS) >> PS); }
template <int N, int R, int E, int PS = 4>
__device__ __forceinline__ constexpr int load_off(int b, int r) {
    return (b * (N / E) + r * (N / R)) + ((b * (N / E) + r * (N / R)) >> PS);
}

template <int N, int R, int NS, int C, int E = 16, int PS = 4>
__device__ __forceinline__ void stage_store(const float2* v, float2* lds, int t, int c) {
#pragma unroll
    for (int b = 0; b < E / R; b++) {
        float2* p = lds + store_base<N, R, NS, E, PS>(t, b) * C + c;
#pragma unroll
        for (int r = 0; r < R; r++) p[store_off<R, NS, PS>(r) * C] = v[b * R + r];
    }
}

template <int N, int R, int C, int E = 16, int PS = 4>
__device__ __forceinline__ void stage_load(float2* v, const float2* lds, int t, int c) {
    const float2* p = lds + (t + (t >> PS)) * C + c;
#pragma unroll
    for (int b = 0; b < E / R; b++)
#pragma unroll
        for (int r = 0; r < R; r++) v[b * R + r] = p[load_off<N, R, E, PS>(b, r) * C];
}

// stage s >= 1 of a transform: twiddles from the N-entry table, or (TAB) from
// the stage's own table that follows it
template <int N, int R, int NS, int DIR, int E, int S, bool TAB>
__device__ __forceinline__ void stage_compute_sel(float2* v, int t, const float2* __restrict__ tw) {
    if constexpr (TAB) stage_compute_tab<N, R, NS, DIR, E, S>(v, t, tw);
    else stage_compute<N, R, NS, DIR, E>(v, t, tw);
}

// Full transform on registers.  All threads of the workgroup must call it
// (it contains __syncthreads); `lds` is this transform group's region.  TAB:
// tw carries the per-stage tables after its N entries.
template <int N, int DIR, int C, bool TAB = false>
__device__ __forceinline__ void fft_regs(float2* v, float2* lds, int t, int c, const float2* __restrict__ tw) {
    using P = Plan<N>;
    stage_compute<N, P::radix(0, DIR), P::ns(0, DIR), DIR>(v, t, tw);
    if constexpr (P::S > 1) {
        __syncthreads();
        stage_store<N, P::radix(0, DIR), P::ns(0, DIR), C, 16, xpad<N, DIR, C, 0>()>(v, lds, t, c);
        __syncthreads();
        stage_load<N, P::radix(1, DIR), C, 16, xpad<N, DIR, C, 0>()>(v, lds, t, c);
        stage_compute_sel<N, P::radix(1, DIR), P::ns(1, DIR), DIR, 16, 1, TAB>(v, t, tw);
    }
    if constexpr (P::S > 2) {
        __syncthreads();
        stage_store<N, P::radix(1, DIR), P::ns(1, DIR), C, 16, xpad<N, DIR, C, 1>()>(v, lds, t, c);
        __syncthreads();
        stage_load<N, P::radix(2, DIR), C, 16, xpad<N, DIR, C, 1>()>(v, lds, t, c);
        stage_compute_sel<N, P::radix(2, DIR), P::ns(2, DIR), DIR, 16, 2, TAB>(v, t, tw);
    }
    if constexpr (P::S > 3) {
        __syncthreads();
        stage_store<N, P::radix(2, DIR), P::ns(2, DIR), C, 16, xpad<N, DIR, C, 2>()>(v, lds, t, c);
        __syncthreads();
        stage_load<N, P::radix(3, DIR), C, 16, xpad<N, DIR, C, 2>()>(v, lds, t, c);
        stage_compute_sel<N, P::radix(3, DIR), P::ns(3, DIR), DIR, 16, 3, TAB>(v, t, tw);
    }
}

// index held by register slot q = b*R + r in the layout of a stage of radix R
template <int N, int R, int E = 16>
__device__ __forceinline__ int reg_index(int t, int q) {
    return t + (q / R) * (N / E) + (q % R) * (N / R);
}

// The same exchange through an LDS array of floats, one component at a time
// (real parts, then imaginary parts): half the LDS of a float2 exchange for
// two more barriers per stage.
template <int N, int R, int NS, int C, int E, int PART, int PS = 4>
__device__ __forceinline__ void stage_store_part(const float2* v, float* lds, int t, int c) {
#pragma unroll
    for (int b = 0; b < E / R; b++) {
        float* p = lds + store_base<N, R, NS, E, PS>(t, b) * C + c;
#pragma unroll
        for (int r = 0; r < R; r++) p[store_off<R, NS, PS>(r) * C] = PART ? v[b * R + r].y : v[b * R + r].x;
    }
}

template <int N, int R, int C, int E, int PART, int PS = 4>
__device__ __forceinline__ void stage_load_part(float2* v, const float* lds, int t, int c) {
    const float* p = lds + (t + (t >> PS)) * C + c;
#pragma unroll
    for (int b = 0; b < E / R; b++)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float x = p[load_off<N, R, E, PS>(b, r) * C];
            if (PART) v[b * R + r].y = x;
            else v[b * R + r].x = x;
        }
}

template <int N, int DIR, int C, int E, int S>
__device__ __forceinline__ void exchange_split(float2* v, float* lds, int t, int c) {
    using P = Plan<N>;
    constexpr int PS = xpad<N, DIR, C, S>();
    __syncthreads();
    stage_store_part<N, P::radix(S, DIR), P::ns(S, DIR), C, E, 0, PS>(v, lds, t, c);
    __syncthreads();
    stage_load_part<N, P::radix(S + 1, DIR), C, E, 0, PS>(v, lds, t, c);
    __syncthreads();
    stage_store_part<N, P::radix(S, DIR), P::ns(S, DIR), C, E, 1, PS>(v, lds, t, c);
    __syncthreads();
    stage_load_part<N, P::radix(S + 1, DIR), C, E, 1, PS>(v, lds, t, c);
}

// Full transform on registers, E elements per thread, exchanges through a
// float LDS region of C interleaved transforms (C * Plan<N>::LDS floats).
template <int N, int DIR, int C, int E, bool TAB = false>
__device__ __forceinline__ void fft_regs_split(float2* v, float* lds, int t, int c, const float2* __restrict__ tw) {
    using P = Plan<N>;
    stage_compute<N, P::radix(0, DIR), P::ns(0, DIR), DIR, E>(v, t, tw);
    if constexpr (P::S > 1) {
        exchange_split<N, DIR, C, E, 0>(v, lds, t, c);
        stage_compute_sel<N, P::radix(1, DIR), P::ns(1, DIR), DIR, E, 1, TAB>(v, t, tw);
    }
    if constexpr (P::S > 2) {
        exchange_split<N, DIR, C, E, 1>(v, lds, t, c);
        stage_compute_sel<N, P::radix(2, DIR), P::ns(2, DIR), DIR, E, 2, TAB>(v, t, tw);
    }
    if constexpr (P::S > 3) {
        exchange_split<N, DIR, C, E, 2>(v, lds, t, c);
        stage_compute_sel<N, P::radix(3, DIR), P::ns(3, DIR), DIR, E, 3, TAB>(v, t, tw);
    }
}

}  // namespace atsim
