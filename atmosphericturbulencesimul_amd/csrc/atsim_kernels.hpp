// atsim_kernels.hpp -- kernel argument structures shared by the kernels and
// the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atsim {

// Per (direction, layer) sampling parameters of one frame: integer base
// position (x0, y0) in [0, M) and separable interpolation weights for taps at
// offsets -1, 0, +1, +2.  Built on the host in FP64 from the cumulative layer
// offsets (AtmosphericTurbulence.c:2700-2755, samplers 2761-2922).
struct LayerSample {
    int x0, y0;
    int pad0, pad1;
    double wx[4];
    double wy[4];
};

// Screen-phase source: phase(i, j) of field f = sum over layers l in
// [lfirst, lfirst + lcount) of s[w][l] * interp_l(i, j), w = f % W,
// direction = f / W.
struct ScreenArgs {
    const float* masters;    // L x M x M
    const LayerSample* ls;   // D x L
    const double* s;         // W x L
    long M;
    int L, W, lfirst, lcount;
    // wavelengths wlo .. wlo + wcnt - 1 of every direction (wcnt 0: all W)
    int wlo, wcnt;
};

struct ExtractArgs {
    const float2* fields;    // [F][N][N] final-but-one field, or null (field == 1)
    float2* field_out;       // [F][N][N] or null
    float* pha;              // [F][Nout][Nout]
    float* amp;              // [F][Nout][Nout]
    // block-rule fields: pha holds atan2 + 2 pi n and dres the remainder
    // r = PHI - atan2 - 2 pi n in [-pi, pi], n = rint((PHI - atan2) / 2 pi)
    float* dres;             // [F][Nout][Nout] (or null)
    const int* blk;          // [Nout] block index (long)(1.0 * i / Nout * xs) of column / row i
    double* bsum;            // [F][nb][2] block sums of cos/sin(PHI - atan2)
    unsigned* rmax;          // [F][nb] bits of each block's largest |r| (float), or null
    const int* unwrap_mode;  // [W]
    ScreenArgs all;          // every layer: the geometric phase PHI
    int last_first, last_count;  // super-layer group whose screen is applied here
    int N, Nout, W, D;
    int amplitude;           // 0: phase only (pha = PHI, amp = 1)
    int xs;                  // Nout / 20 blocks per side
    // factored phase multipliers s_wl = geo_a[w] geo_b[l]: the planes vr, vl
    // ([D][N][N], FP64) of sum_l geo_b[l] v_l over the layers before the last
    // group and over the last group, at full-grid positions; vr null: the
    // extraction samples the layers itself
    const double* geo_a;     // [W]
    const double* geo_b;     // [L]
    double* vr;
    double* vl;
    int w_begin;             // first field of this chunk of <= 8 wavelengths
};

// Where the block rule's final update takes the remainders r of the blocks
// it visits: fields null, the extraction's FP32 map; otherwise they are made
// again from the final fields [F][N][N] and the factored geometric phase's
// planes vr, vl [D][N][N] with multipliers geo_a [W] (the output crop starts
// at full-grid row / column o), as the extraction made them; rot: the
// extraction rotates by the last group's screen (it has layers).
struct RemainderSource {
    const float2* fields;
    const double* vr;
    const double* vl;
    const double* geo_a;
    int N, W, o, rot;
};

}  // namespace atsim
