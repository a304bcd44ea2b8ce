// atsim_launch.hpp -- host-side launchers of the HIP kernels (atsim_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atsim_kernels.hpp"

namespace atsim {

enum { OP_FFT = 0, OP_FRESNEL = 1 };

bool supported_n(long n);

// Row pass over F fields of n x n: optional screen multiply (sa.lcount > 0)
// before the transform; op OP_FRESNEL = FFT_x, x h(kx) (table tf + (f % W) n),
// IFFT_x; OP_FFT = one transform of sign dir.  gen: the input field is 1
// (nothing is read).
hipError_t launch_row(long n, int op, int dir, bool gen, float2* data, long F, const float2* tf,
                      const float2* tw, const ScreenArgs& sa, hipStream_t st);
// Fresnel x pass of every (direction, wavelength) field with the screen
// phase of the layer group sa.lfirst .. + sa.lcount (sampled once per pixel
// and direction); gen: the incoming field is 1.  Tables tf + w n.
// mw: wavelengths per workgroup (0 = largest that fits).  tfp (not with
// gen): the x factor of the gap before the group is applied first (tables
// tfp + w n).  tw: the n twiddles followed by the per-stage tables
// (tw_tables_size(n) entries, tw_stage_off).
hipError_t launch_row_fresnel(long n, bool gen, float2* data, int D, const float2* tf, const float2* tw,
                              const ScreenArgs& sa, hipStream_t st, int mw = 0, const float2* tfp = nullptr);
// Column pass around a layer group: y factor tfp of the gap before the
// group, the group's screen phase (layers sa.lfirst .. + sa.lcount, their
// sampled phases in planes slot0 .. of vq), y factor tf of the gap after it;
// every field of the wavelength range of sa (tables + w n).  n <= 4096.
// e: elements per thread (0 = default: 32 with float exchanges for
// 256 <= n <= 2048, else 16; 16 forces 16).  tw as for launch_row_fresnel
// (the 32-element passes read the per-stage tables).
hipError_t launch_col_screen(long n, float2* data, int D, const float2* tfp, const float2* tf, const float2* tw,
                             const ScreenArgs& sa, const double* vq, int slot0, hipStream_t st, int e = 0);
// Sampled phase planes (FP64 quarter turns) of layers slot_layer[0 ..
// nslots) for every direction: vq[((slot D + d) n + j) n + i].
hipError_t launch_layer_phase(double* vq, const float* masters, const LayerSample* ls, const int* slot_layer,
                              int nslots, long n, long M, long L, long D, hipStream_t st);
bool col_screen_supported(long n);
// Column pass: same transforms along y; no screen.
// F fields, or (wcnt > 0) F = D * wcnt fields: wavelengths wlo .. wlo + wcnt - 1
// of each direction.  c: columns per workgroup (0 = default for n); e = 32: the Fresnel pass
// holds 32 elements per thread and exchanges through float LDS (256 <= n <= 2048)
hipError_t launch_col(long n, int op, int dir, float2* data, long F, const float2* tf, int W,
                      const float2* tw, hipStream_t st, int c = 0, int e = 0, int wlo = 0, int wcnt = 0);

hipError_t launch_permut(float2* data, long n0, long n1, long nbatch, hipStream_t st);
// mixed-precision row transforms (atsim_fft_mfma.hip): length 2048, radix-16
// stages on the matrix cores (FP16 hi + lo operands, FP32 accumulation);
// nrows a multiple of 2
bool fft_mfma_supported(long n);
hipError_t launch_row_fft_mfma(float2* data, long nrows, int dir, const float2* tw, hipStream_t st);
// out[b][x][y] = in[b][y][x], n a multiple of 32
hipError_t launch_transpose(const float2* in, float2* out, long n, long nbatch, hipStream_t st);
// transforms of any size: lines of length L (< = M) between strided images and
// a padded scratch [nlines][M]; base of line l = (l / per_img) img_size +
// (l % per_img) lstep, elements estride apart; chirp (nullable) multiplies
hipError_t launch_line_gather(const float2* src, float2* dst, long nlines, long L, long M, long per_img, long img_size,
                              long lstep, long estride, const float2* chirp, hipStream_t st);
hipError_t launch_line_scatter(const float2* src, float2* dst, long nlines, long L, long M, long per_img,
                               long img_size, long lstep, long estride, const float2* chirp, hipStream_t st);
hipError_t launch_row_mul(float2* data, const float2* h, long nrows, long M, hipStream_t st);
// out[b][p] = in[b][sigma[p]] (images of n elements)
hipError_t launch_index_gather(const float2* in, float2* out, const int* sigma, long n, long nbatch, hipStream_t st);
// Fresnel factor of any n0 x n1 spectrum with the reference's permut pair
// (sigma: gather map of one permut, null for even sizes); out != in
hipError_t launch_fresnel_any(const float2* in, float2* out, long n0, long n1, long nbatch, double coeff,
                              const int* sigma, hipStream_t st);
hipError_t launch_noise(float2* out, long n, uint64_t seed, uint32_t stream, hipStream_t st);
hipError_t launch_synth_filter(const float2* noise, float2* out, long size, double f0, double fi,
                               double rlim, hipStream_t st);
// partial sums of squared differences at lags (1,1),(1,2),(2,1),(2,2) of Re
// and Im: partial[block][8]
hipError_t launch_lag_sums(const float2* c, long size, double* partial, int nblocks, hipStream_t st);
hipError_t launch_split_scale(const float2* c, float* o1, float* o2, long n, double s1, double s2,
                              hipStream_t st);
hipError_t launch_strfn(const float* s, long size, long a, long b, int circular, double* partial,
                        int nblocks, hipStream_t st);
hipError_t launch_scale(float* s, long n, double scale, hipStream_t st);
hipError_t launch_cmul(float2* a, const float2* b, long n, long nbatch, hipStream_t st);
hipError_t launch_amp_pha(const float2* e, float* amp, float* pha, long n, hipStream_t st);
hipError_t launch_tf_init(float2* C, long size, double coeff, double fpmaskrad, hipStream_t st);

// the planes a.vr, a.vl of the factored geometric phase, and with vq the
// column passes' phase planes of the nslots layers slot_layer as
// launch_layer_phase makes them; zero_n doubles at zero_p are cleared
hipError_t launch_extract_geometry(const ExtractArgs& a, bool full, hipStream_t st, double* vq = nullptr,
                                   const int* slot_layer = nullptr, int nslots = 0, double* zero_p = nullptr,
                                   long zero_n = 0);
hipError_t launch_extract(const ExtractArgs& a, bool full, hipStream_t st);
// cmax: nfields unsigned, zero on entry (the first round's largest |count|
// per field)
hipError_t launch_block_fix(double* bsum, double* bin, double* ch, unsigned* cmax, const int* field_list,
                            int nfields, int xs, hipStream_t st);
// first-round counts alone (float, per field nb = xs^2), with cmax as above
hipError_t launch_block_counts(const double* bin, float* ch, unsigned* cmax, const int* field_list, int nfields,
                               int xs, hipStream_t st);
hipError_t launch_block_apply(float* pha, const float* dres, const double* bin, const int* field_list,
                              const int* blk, int nfields, int Nout, int xs, hipStream_t st);
// the same update, visiting only the blocks whose largest |r| (rmax, [F][nb]
// float bits from the extraction) lets a count be non-zero, with r from dres
// or made again as rs says; bstart[b]: first output row / column of block b
// (xs + 1 entries)
hipError_t launch_block_apply_sparse(float* pha, const float* dres, const double* bin, const unsigned* rmax,
                                     const int* field_list, const int* bstart, int nfields, int Nout, int xs,
                                     const RemainderSource& rs, hipStream_t st);

hipError_t launch_zero_f64(double* p, long n, hipStream_t st);
// per-field sums of pha, pha^2, amp^2, amp^4 over n pixels of F fields:
// out[F][4]; part holds F * 128 * 4 doubles of partial sums
constexpr long STATS_PART_PER_FIELD = 128 * 4;
hipError_t launch_stats(const float* pha, const float* amp, long F, long n, double* part, double* out,
                        hipStream_t st);
hipError_t launch_layer_screens(const float* masters, const LayerSample* ls, float* out, long N, long M, long L,
                                long D, hipStream_t st);

// surrogate network (atsim_fno.hip)
struct GemmArgs {
    long M, N, K, batch;
    float alpha, beta;
    const float* A;
    long sab, sam, sak;
    const float* B;
    long sbb, sbk, sbn;
    float* C;
    long scb, scm, scn;
    const float* bias;  // per-row bias (or null)
    float* pre;         // pre-activation output with C's strides (or null)
    int act;            // 1: GELU
    int avec, bvec;     // staging modes, set by launch_gemm_bf16
};
hipError_t launch_gemm_bf16(const GemmArgs& g, hipStream_t st);
hipError_t launch_add_bias_act(const float* a, const float* b, const float* bias, float* pre, float* y, long C,
                               long P, long n, int gelu, hipStream_t st);
hipError_t launch_gelu_bwd(const float* pre, const float* dy, float* dx, long n, hipStream_t st);
hipError_t launch_bias_grad(const float* dy, float* partial, long S, long C, long P, int nchunk, hipStream_t st);
hipError_t launch_mse(const float* pred, const float* target, float* grad, long n, double* partial, int nblk,
                      hipStream_t st);
hipError_t launch_reduce_batches(const float* part, float* out, long nb, long n, int accumulate, hipStream_t st);
hipError_t launch_nn_pairs(const float* screens, const double* sc, const float* pha, const float* amp, long D, long L,
                           long N, long Stot, long s0, float* x, float* t, double xs, double xl, double cs, double ps,
                           hipStream_t st);
hipError_t launch_nn_unpack(const float* y, const float* x, long L, long S, long P, double xs, double cs, double ps,
                            float* amp, float* pha, hipStream_t st);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st);

}  // namespace atsim
