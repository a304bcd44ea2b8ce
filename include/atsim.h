/*
 * atsim.h -- C ABI of the MI355X wavefront library (libatsim.so).
 *
 * Plain C types only: device buffers are `void *` device pointers (HIP
 * allocations on the context's device), sizes are `long`, status is `int`
 * (0 = success, < 0 = error; the message is returned by atsim_error()).
 * Nothing here exits the process (the reference's C functions print and
 * exit on fatal errors; the registry-level compat layer in atsim_compat.h
 * keeps that behaviour on top of these calls).
 *
 * Complex data are interleaved (re, im) float32 pairs, row-major with x
 * fastest, exactly the reference's complex_float layout (CLIcore.h:130-140)
 * so host images can be copied to the device byte for byte.
 *
 * Sizes: the HIP path handles square power-of-two grids 16 <= N <= 8192.
 */
#ifndef ATSIM_H
#define ATSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atsim_ctx atsim_ctx;
typedef struct atsim_sim atsim_sim;

#define ATSIM_OK 0
#define ATSIM_EINVAL (-1)
#define ATSIM_EHIP (-2)
#define ATSIM_ENOMEM (-3)
#define ATSIM_EUNSUPPORTED (-4)

/* ---------------------------------------------------------------- context */
/* One context = one device + one HIP stream.  Not thread safe; use one
 * context per host thread. */
int atsim_ctx_create(int device, atsim_ctx **out);
void atsim_ctx_destroy(atsim_ctx *ctx);
const char *atsim_error(const atsim_ctx *ctx);
int atsim_sync(atsim_ctx *ctx);
/* HIP stream (hipStream_t) all work of this context is queued on */
void *atsim_stream(atsim_ctx *ctx);
/* device memory of the context's device */
int atsim_malloc(atsim_ctx *ctx, void **ptr, size_t nbytes);
int atsim_free(atsim_ctx *ctx, void *ptr);
/* synchronous copy between host and/or device buffers (unified addressing) */
int atsim_copy(atsim_ctx *ctx, void *dst, const void *src, size_t nbytes);
/* Pinned host memory (hipHostMalloc) for asynchronous copies */
int atsim_host_alloc(atsim_ctx *ctx, void **ptr, size_t nbytes);
int atsim_host_free(atsim_ctx *ctx, void *ptr);
/* Asynchronous copy on the context's copy stream, ordered after all work
 * queued so far on its main stream, so that frame k's outputs travel while
 * frame k+1 computes.  slot (0..3) names the copy: atsim_copy_wait(slot)
 * blocks the host until it has landed; atsim_stream_wait_copy(slot) makes
 * later main-stream work (a frame that overwrites the source) wait for it.
 * No counterpart in the reference (its frames are copied synchronously). */
int atsim_copy_async(atsim_ctx *ctx, int slot, void *dst, const void *src, size_t nbytes);
int atsim_copy_wait(atsim_ctx *ctx, int slot);
int atsim_stream_wait_copy(atsim_ctx *ctx, int slot);
/* Launch-shape tuning of this context (0 restores the built-in default):
 *   "row_mw"  wavelength fields per row-pass workgroup (1..8; default
 *             8192 / n, at most 8: two workgroups per CU)
 *   "col_c"   columns per column-pass workgroup (2, 4, 8, 16)
 *   "col_e"   elements per thread of the column Fresnel pass (16, or 32:
 *             8 columns per workgroup exchanging through half the LDS).
 *             32 applies only to Fresnel column passes with 256 <= n <= 2048
 *             and col_c 0, 8 or 16; plain FFT passes, other sizes and
 *             col_c 2 or 4 run the 16-element kernel (the value is kept and
 *             atsim_get_tuning returns it)
 *   "streams" S = 2..4 (default 2): the Fresnel passes of S contiguous sets
 *             of the wavelengths run as S chains on S streams (set k started
 *             after set k - 1's first row pass) and join before the
 *             extraction; 1: one stream.  Frames under atsim_sim_profile use
 *             one stream.
 *   "gap_passes" 1 (default): one pass over the fields per Fresnel gap --
 *             the x and y factors of a gap commute, so gaps alternate
 *             between x-first and y-first and a row (column) pass applies
 *             the x (y) factor of the gap before a layer group, the group's
 *             screen and the x (y) factor of the gap after it (n <= 4096,
 *             at least 2 gaps; the screens of the column passes' layers are
 *             sampled once per frame and direction into FP64 planes of
 *             n^2 x 8 bytes each); 2: a row pass (screen, x factor) and a
 *             column pass (y factor) per gap
 *   "col2_e"  elements per thread of the column passes that apply a screen
 *             (16, or 32: 8 columns per workgroup exchanging through half
 *             the LDS, the default for 256 <= n <= 2048)
 *   "graph"   1: atsim_sim_frame replays a HIP graph of the frame's kernels,
 *             captured at the second frame (and again when the output
 *             pointers change, or at the second frame after a tuning
 *             change); the position upload stays outside the graph
 *   "extract_planes" 1 (default): when the phase multipliers factor
 *             exactly as s_wl = a_w b_l (a_w = s_w0, b_l = s_0l / s_00, as
 *             the lambda_ref / lambda rows do), the layers' screens are
 *             summed into two FP64 planes per direction (sum b_l v_l over the
 *             layers before the last group and over the last group, 16 n^2
 *             bytes) by a kernel that runs beside the Fresnel passes, and
 *             the extraction reads them (PHI_w = a_w Vr + a_w Vl) instead of
 *             sampling every layer; 2: the extraction samples the layers
 *   "block_apply" 1 (default): the final update of the block unwrap rule
 *             visits only the blocks where a count can be non-zero (from
 *             each block's largest |r|, kept by the extraction); 2: every
 *             pixel.  Both give the same bits.
 *   "pipeline" 1: frames of a simulation overlap.  Frame k + 1's Fresnel
 *             passes run on a stream of the simulation's own beside frame
 *             k's extraction and block fix-up, which stay on the context
 *             stream after the passes of their frame; two sets of fields,
 *             planes and block sums alternate (allocated at the first such
 *             frame).  A frame's outputs are complete when the context
 *             stream reaches them, and are the same bits as with 0 (the
 *             default: every frame on the context stream).  The passes wait
 *             for the context stream when an atsim call that writes device
 *             memory came since the simulation's previous frame; device
 *             writes of the caller's own to the masters (through
 *             atsim_sim_master's pointer) must be complete before the next
 *             atsim_sim_frame call.  Profiled frames run without it.  Leave
 *             it at 0 in a loop that hands every frame to the host through
 *             atsim_copy_async: the copy stream can share a hardware queue
 *             with the passes' stream, whose next frame then waits for the
 *             copy.
 * Returns 0 / < 0; atsim_get_tuning returns the value or < 0. */
int atsim_set_tuning(atsim_ctx *ctx, const char *key, int value);
int atsim_get_tuning(atsim_ctx *ctx, const char *key);
/* 1 if the library was built with code for the current device */
int atsim_device_ok(atsim_ctx *ctx);

/* ---------------------------------------------------------------- FFT */
/* In-place unnormalised 2D c2c over nbatch contiguous n x n images.
 * sign = -1 forward, +1 backward: the reference's do2dfft / do2dffti
 * (src/fft/fft.c:1298-1315, FFTW plan over the swapped axes fft.c:1178-1190). */
int atsim_fft2d(atsim_ctx *ctx, void *d_data, long n, long nbatch, int sign);
/* The 2D transform of atsim_fft2d in one of two arithmetic variants:
 * 0 = the FP32 Stockham transform (atsim_fft2d); 1 = radix-16 stages as
 * DFT-16 matrix products on the matrix cores (v_mfma_f32_16x16x32_f16) with
 * FP16 hi + lo operands (three products, FP32 accumulation, every vector
 * scaled by a power of two first) and the radix-8 stage and twiddles in
 * FP32, rows then columns through two transposes; variant 1 is built for
 * n = 2048.  atsim_row_fft_variant transforms nrows contiguous rows of length
 * n (nrows a multiple of n) in either variant.  No counterpart in the
 * reference: the precision / speed comparison of the two paths. */
int atsim_fft2d_variant(atsim_ctx *ctx, void *d_data, long n, long nbatch, int sign, int variant);
int atsim_row_fft_variant(atsim_ctx *ctx, void *d_data, long n, long nrows, int sign, int variant);
/* 2D c2c of any nbatch contiguous n0 x n1 images (n0 = x size, fastest):
 * the reference's FFT_do2dfft on any shape (FFTW plan over the swapped axes,
 * fft.c:1145-1209), unnormalised, sign as atsim_fft2d.  Square power-of-two
 * images take atsim_fft2d; otherwise each axis is transformed on its own,
 * power-of-two lengths directly and other lengths by Bluestein's chirp
 * convolution through power-of-two transforms (lengths up to 4096, or powers
 * of two up to 8192).  Allocates temporary device memory. */
int atsim_fft2d_nm(atsim_ctx *ctx, void *d_data, long n0, long n1, long nbatch, int sign);
/* Quadrant swap of nbatch complex n0 x n1 images, permut (fft.c:570-792),
 * any size: for odd sizes the reference's swap loops are applied as their
 * gather map. */
int atsim_permut(atsim_ctx *ctx, void *d_data, long n0, long n1, long nbatch);

/* ---------------------------------------------------------------- Fresnel */
/* Angular-spectrum Fresnel step of Fresnel_propagate_wavefront
 * (src/WFpropagate/WFpropagate.c:68-152): out = IFFT[FFT[in] exp(-i pi z
 * lambda (kx^2+ky^2)/(dx n)^2)/n^2], signed frequencies, on nbatch n x n
 * fields.  d_in may equal d_out. */
int atsim_fresnel(atsim_ctx *ctx, const void *d_in, void *d_out, long n, long nbatch,
                  double pupil_scale, double z, double lambda);
/* atsim_fresnel without allocating: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The transfer factor
 * and twiddle tables go to a fixed 128 KB device array of the library
 * (n <= 8192; larger sizes return ATSIM_EUNSUPPORTED) that every call on a
 * device shares: calls that overlap on one device race.  Asynchronous on the
 * context stream (atsim_sync waits); same results as atsim_fresnel. */
int atsim_fresnel_noalloc(atsim_ctx *ctx, const void *d_in, void *d_out, long n, long nbatch,
                          double pupil_scale, double z, double lambda);
/* Fresnel_propagate_wavefront on any n0 x n1 shape (WFpropagate.c:68-152):
 * FFT, permut, e^{-i coeff ((ii - n0/2)^2 + (jj - n1/2)^2)} / (n0 n1) with
 * coeff = pi z lambda / (dx n0)^2 (n0 for both axes, as the reference),
 * permut, inverse FFT.  Square power-of-two images take atsim_fresnel. */
int atsim_fresnel_nm(atsim_ctx *ctx, const void *d_in, void *d_out, long n0, long n1, long nbatch,
                     double pupil_scale, double z, double lambda);
/* Init_Fresnel_propagate_wavefront (WFpropagate.c:166-261): fills the
 * anti-aliased, permuted transfer function C (n x n complex, device). */
int atsim_fresnel_tf_init(atsim_ctx *ctx, void *d_C, long n, double pupil_scale, double z,
                          double lambda, double fpmaskrad);
/* Fresnel_propagate_wavefront1 (WFpropagate.c:271-330): out = IFFT[FFT[in] C]. */
int atsim_fresnel_tf_apply(atsim_ctx *ctx, const void *d_in, void *d_out, const void *d_C,
                           long n, long nbatch);
/* Fresnel_propagate_cube (WFpropagate.c:337-401): amplitude and phase cubes
 * (n x n x nz float, device) for z_k = zstart + (zend - zstart) k / nz. */
int atsim_fresnel_cube(atsim_ctx *ctx, const void *d_in, float *d_amp, float *d_pha, long n,
                       double pupil_scale, double zstart, double zend, long nz, double lambda);

/* ---------------------------------------------------------------- noise */
/* Philox4x32-10 Gaussian complex noise (re, im ~ N(0,1) independent), element
 * e of stream s uses counter (e_lo, e_hi, 0x0A75, 0) and key (seed_lo,
 * seed_hi ^ s); Box-Muller in FP64.  Replaces the reference's clock-seeded
 * make_rnd (src/image_gen/image_gen.c:1634-1740). */
int atsim_noise_c64(atsim_ctx *ctx, void *d_out, long n, uint64_t seed, uint32_t stream);

/* ---------------------------------------------------------------- screens */
/* make_master_turbulence_screen (AtmosphericTurbulence.c:643-855) with the
 * noise injected: C = A(r) z with the von Karman amplitude A, quadrant swap,
 * forward FFT, out1 = Re/sqrt(C1), out2 = Im/sqrt(C2), where C is the 4-lag
 * structure-function fit (AT.c:790-847).  d_noise: size^2 complex in the
 * reference's centred layout.  d_out1/d_out2: size^2 float (d_out2 may be
 * NULL).  C1/C2 (host) receive the fit coefficients.  rlim > 0 zeroes the
 * spectrum inside rlim (the RLIM variable, AT.c:671-736). */
int atsim_make_master(atsim_ctx *ctx, const void *d_noise, long size, double outerscale_px,
                      double innerscale_px, double rlim, float *d_out1, float *d_out2,
                      double *C1, double *C2);
/* Structure function at lag (a, b), circular mean of squared differences,
 * accumulated in FP64 (the quantity of fft_structure_function, info.c:1670). */
int atsim_strfn(atsim_ctx *ctx, const float *d_screen, long size, long a, long b, double *out);
/* r0 [pixels] from the x-lag dpix structure function, non circular
 * (AT.c:2246-2270). */
int atsim_r0_measure(atsim_ctx *ctx, const float *d_screen, long size, long dpix, double *r0);
/* d_screen *= scale (float storage, AT.c:2274-2285 and 2339-2351). */
int atsim_scale(atsim_ctx *ctx, float *d_screen, long n, double scale);

/* ---------------------------------------------------------------- layers */
/* Super-layer binning (AT.c:2067-2150).  alt must be ascending.  Writes
 * slayer_alt[<=L] and super_layer_index[L]; returns the number of
 * super-layers (>= 1) or < 0 on error. */
long atsim_superlayers(const double *alt, const double *cn2, long L, double bin,
                       double *slayer_alt, long *super_layer_index);

/* ---------------------------------------------------------------- simulation */
typedef struct {
    long N;        /* WF_RAW_SIZE: propagation grid (power of 2) */
    long Nout;     /* WFsize: centred output crop, Nout <= N */
    long M;        /* MASTER_SIZE: master screen size, N <= M */
    long L;        /* layers (ascending altitude), at most 64 */
    long W;        /* wavelengths (fields per direction) */
    long D;        /* directions */
    double dx;     /* PUPIL_SCALE [m/pixel] */
    double dt;     /* WFTIME_STEP [s] */
    double zangle; /* ZENITH_ANGLE [rad] */
    int bicubic;   /* 1: Catmull-Rom (reference default, AT.c:1605); 0: bilinear */
    int fresnel;   /* FRESNEL_PROPAGATION */
    int amplitude; /* WAVEFRONT_AMPLITUDE */
    const double *slayer_alt;       /* super-layer altitudes [m] */
    const long *super_layer_index;  /* L */
    const double *lambda;           /* W wavelengths [m] */
    const double *scoeff;           /* W x L phase multipliers (1.0 for the reference lambda) */
    const int *unwrap_mode;         /* W: 0 reference-lambda rule (AT.c:3456), 1 block rule (AT.c:3549-4262) */
    const double *xpos0;            /* D x L initial x positions [pixels] */
    const double *ypos0;            /* D x L initial y positions [pixels] */
    const double *vx;               /* L [pixels/s] */
    const double *vy;               /* L [pixels/s] */
} atsim_sim_desc;

int atsim_sim_create(atsim_ctx *ctx, const atsim_sim_desc *desc, atsim_sim **out);
void atsim_sim_destroy(atsim_sim *sim);
/* device pointer of layer l's M x M float master (write it directly, or use
 * atsim_sim_set_master) */
float *atsim_sim_master(atsim_sim *sim, long layer);
/* copy an M x M float master (host or device pointer) into layer `layer` */
int atsim_sim_set_master(atsim_sim *sim, long layer, const float *src);
/* Fill every layer's master with make_master_turbulence_screen from the
 * Philox noise stream `layer` of `seed` (outer/inner scales in pixels per
 * layer, AT.c:2204-2226). */
int atsim_sim_synth_masters(atsim_sim *sim, uint64_t seed, const double *outerscale_px,
                            const double *innerscale_px);
/* Per-run normalisation (AT.c:2246-2363): r0 of every master from its
 * 50-pixel x structure function, all masters scaled by (mean r0 /
 * r0_goal)^(5/6), then layer l by sqrt(cn2[l] / cos zangle).  r0_measured
 * (nullable, L) receives the r0 values before scaling [pixels]. */
int atsim_sim_normalize_masters(atsim_sim *sim, double r0_goal_px, const double *cn2, double zangle,
                                double *r0_measured);
/* Advance every layer position by one time step (AT.c:2700-2755) and compute
 * one frame for all D x W fields: d_pha / d_amp are device float arrays
 * [D][W][Nout][Nout] (phase in rad, unwrapped; amplitude); d_field (nullable)
 * receives the final complex fields [D][W][N][N]. */
int atsim_sim_frame(atsim_sim *sim, float *d_pha, float *d_amp, void *d_field);
/* One frame of n simulations of one context (any grid sizes): every
 * simulation advances its layers and computes its frame as atsim_sim_frame
 * does, into d_pha[k] / d_amp[k], each simulation on its own stream forked
 * from and joined to the context stream, so that grids of different sizes
 * run side by side (launched eagerly; no captured graph).  No counterpart in
 * the reference (one process per grid). */
int atsim_sims_frame(atsim_sim **sims, long n, float **d_pha, float **d_amp);
/* Use the masters of `owner` (same context, M and L) instead of this
 * simulation's own, which are freed: simulations of several grid sizes over
 * one set of masters.  The owner must outlive this simulation. */
int atsim_sim_use_masters(atsim_sim *sim, const atsim_sim *owner);
/* Replace the current FP64 positions [D][L] [pixels]; the next frame
 * advances from them by v dt (the reference's REALTIME mode sets positions
 * from the clock, AT.c:2700-2704). */
int atsim_sim_set_positions(atsim_sim *sim, const double *xpos, const double *ypos);
/* current FP64 positions [D][L] (host arrays, nullable) */
int atsim_sim_positions(atsim_sim *sim, double *xpos, double *ypos);
/* number of Fresnel gaps per frame */
long atsim_sim_gaps(atsim_sim *sim);
/* Per-launch HIP-event timing of the frame kernels, for benchmarking.
 * enable = 1 starts (and clears) the record, 0 stops it.  kind: 0 row pass
 * with one factor (screen + x factor, or x factor alone), 1 column pass with
 * one factor, 2 extraction, 3 block unwrap, 4 row pass with two x factors
 * and a screen, 5 column pass with two y factors and a screen, 6 sampling of
 * the phase planes that column passes apply, 7 output statistics, 8 the
 * geometric-phase planes of the extraction ("extract_planes"; with one pass
 * per gap the same sweep makes the phase planes, and kind 6 records none);
 * returns
 * the summed duration [ms] and the number of launches recorded since the
 * last enable. */
int atsim_sim_profile(atsim_sim *sim, int enable);
int atsim_sim_profile_read(atsim_sim *sim, int kind, double *ms, long *count);

/* Block sums of the science-lambda unwrap rule of the last frame: out [D][W]
 * [xs][xs][2] (host, n = D W xs xs 2 doubles, xs = Nout / 20) holds, per
 * block (jj2, ii2), sum cos(PHI - pha) and sum sin(PHI - pha) over its pixels
 * -- the bre / bim sums of AtmosphericTurbulence.c:3523-3625, before the
 * block fix-up; zero for fields of the reference-lambda rule.  Synchronises
 * the context stream.  A diagnostic for tests, with no entry point in the
 * reference. */
int atsim_sim_block_sums(atsim_sim *sim, double *out, long n);

/* Autotuner of the launch shapes for this simulation's sizes: times `frames`
 * frames (after two untimed ones) under each candidate of a coordinate
 * descent over the tuning keys gap_passes, streams, row_mw, col2_e, col_c
 * and col_e (two sweeps, values the kernels support at this grid size; a
 * candidate replaces the current set only when it is more than 1 % faster)
 * and leaves the fastest set on the context.  Frames are written to d_pha /
 * d_amp (as atsim_sim_frame); the layer positions are restored afterwards,
 * so the series continues as if the search had not run.  report (nullable)
 * receives one line per candidate, "key=value ... ms=<per frame>", and a
 * final "best ..." line.  No counterpart in the reference. */
int atsim_sim_autotune(atsim_sim *sim, float *d_pha, float *d_amp, int frames, char *report, long report_len);

/* Output statistics of every frame (the OPD / scintillation statistics
 * stream; no counterpart in the reference, which leaves them to
 * post-processing).  With enable = 1 every following frame also reduces its
 * outputs per field, in FP64 over the Nout^2 pixels; 0 stops it (the default).
 * atsim_sim_stats writes, for the last frame, out [D][W][4] (host, n >= 4 D W):
 * OPD piston <pha> lambda / 2 pi [m], OPD rms [m], mean intensity <amp^2> and
 * scintillation index <I^2> / <I>^2 - 1.  Synchronises the context stream. */
int atsim_sim_stats_enable(atsim_sim *sim, int enable);
int atsim_sim_stats(atsim_sim *sim, double *out, long n);

/* The block fix-up of the science-lambda rule on given block sums: bsum
 * [xs][xs][2] (host; sum cos, sum sin per block), peaks [xs][xs] (host) the
 * block peaks after the reference's iterative +-2 pi consistency heuristic
 * (AtmosphericTurbulence.c:3627-4262), as atsim_sim_frame computes them.
 * xs <= 4096.  A test entry point with no counterpart in the reference. */
int atsim_block_fixup(atsim_ctx *ctx, const double *bsum, long xs, double *peaks);
/* First-round counts of the block fix-up for given block peaks [xs][xs]
 * (host): counts [xs][xs] (host) as the reference's float image holds them
 * with WFprecision = 0 (AtmosphericTurbulence.c:3667-3903: each +-0.2 /
 * +-0.15 rounded to float in the reference's loop order).  xs <= 4096.  A
 * test entry point with no counterpart in the reference. */
int atsim_block_counts(atsim_ctx *ctx, const double *peaks, long xs, float *counts);

/* Advance every layer by one time step without computing a frame (the
 * positions a following atsim_sim_layer_screens samples).  No counterpart in
 * the reference: the surrogate's frames skip the layer loop. */
int atsim_sim_advance(atsim_sim *sim);
/* Per-layer sampled screens at the positions of the last frame (the FP64
 * sampler of the frame loop, rounded to float): d_out [D][L][N][N], radians
 * at the reference wavelength before Scoeff.  No counterpart in the
 * reference: the input of the surrogate network. */
int atsim_sim_layer_screens(atsim_sim *sim, float *d_out);

/* ---------------------------------------------------------------- surrogate network
 * Building blocks of the Fourier-neural-operator surrogate of the layer loop
 * (no counterpart in the reference).  All buffers are device FP32.
 *
 * atsim_gemm_bf16: for b < batch, C_b = alpha A_b B_b + beta C_b with
 * A_b (M x K) at A + b sab + m sam + k sak, B_b (K x N) at B + b sbb + k sbk
 * + n sbn, C_b (M x N) at C + b scb + m scm + n scn (element strides).  The
 * operands are rounded to BF16 and multiplied on the matrix cores with FP32
 * accumulation.  batch and ceil(M / 64) at most 65535.  Strides are
 * non-negative (ATSIM_EINVAL otherwise) and otherwise trusted: the library
 * does not know the allocations, so the caller guarantees that every
 * addressed element of A, B, C (and bias, pre) lies inside its buffer (the
 * Python front end, fno.py, checks its operands against their tensors). */
int atsim_gemm_bf16(atsim_ctx *ctx, long M, long N, long K, long batch, float alpha,
                    const float *A, long sab, long sam, long sak,
                    const float *B, long sbb, long sbk, long sbn, float beta,
                    float *C, long scb, long scm, long scn);
/* atsim_gemm_bf16 with an epilogue: v = alpha A B + beta C + bias[m] (bias
 * may be NULL); pre (may be NULL, C's layout) receives v; C receives GELU(v)
 * if gelu != 0, else v. */
int atsim_gemm_bf16_ex(atsim_ctx *ctx, long M, long N, long K, long batch, float alpha,
                       const float *A, long sab, long sam, long sak,
                       const float *B, long sbb, long sbk, long sbn, float beta,
                       float *C, long scb, long scm, long scn, const float *bias, float *pre, int gelu);
/* y = act(a + b + bias[c]) over n = S*C*P elements laid out [S][C][P]; b, bias
 * and pre may be NULL; pre keeps the activation's input; act = GELU (tanh
 * form) if gelu != 0, else identity. */
int atsim_nn_add_bias_act(atsim_ctx *ctx, const float *a, const float *b, const float *bias, float *pre,
                          float *y, long C, long P, long n, int gelu);
/* dx = dy * GELU'(pre) */
int atsim_nn_gelu_bwd(atsim_ctx *ctx, const float *pre, const float *dy, float *dx, long n);
/* db[c] = sum_{s,p} dy[s][c][p] (added to db if accumulate) */
int atsim_nn_bias_grad(atsim_ctx *ctx, const float *dy, float *db, long S, long C, long P, int accumulate);
/* mean squared error; grad (may be NULL) = 2 (pred - target) / n; loss (may
 * be NULL; synchronises the stream) = mean of (pred - target)^2 */
int atsim_nn_mse(atsim_ctx *ctx, const float *pred, const float *target, float *grad, long n, double *loss);
/* Pairs of the surrogate from one simulator step: screens [D][L][N*N] (from
 * atsim_sim_layer_screens), d_scoeff [L] (device, FP64) the wavelength's
 * phase multipliers; writes samples s0 .. s0+D-1 of Stot, channel-major
 * [2L][Stot][N*N]: x[l] = v_l / xs with v_l = scoeff_l screen_l, x[L+l] =
 * 5-point Laplacian of v_l / xl (edge pixels clamped); when pha and amp (each
 * [D][N*N]) are given, t [2][Stot][N*N] = ((Re u - 1) / cs, Im u / ps) with
 * u = amp e^{i (pha - sum_l v_l)}. */
int atsim_nn_make_pairs(atsim_ctx *ctx, const float *screens, const double *d_scoeff, const float *pha,
                        const float *amp, long D, long L, long N, float *x, float *t, long Stot, long s0,
                        double xs, double xl, double cs, double ps);
/* network output y [2][S][P] -> u = (1 + cs y0) + i ps y1, amp = |u|,
 * pha = xs sum_{l < L} x_l + arg u (x: the network input, phase channels first) */
int atsim_nn_unpack(atsim_ctx *ctx, const float *y, const float *x, long L, long S, long P, double xs, double cs,
                    double ps, float *amp, float *pha);
/* out[i] = sum_{b < nb} part[b n + i] (added to out if accumulate) */
int atsim_nn_reduce_batches(atsim_ctx *ctx, const float *part, float *out, long nb, long n, int accumulate);
/* one Adam step (bias-corrected, step >= 1) on n parameters */
int atsim_nn_adam(atsim_ctx *ctx, float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                  float b2, float eps, long step);

#ifdef __cplusplus
}
#endif
#endif
