/*
 * atsim_oracle.h -- CPU restatement of the reference wavefront path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker (or the timed CPU baseline).
 *
 * Everything is FP64 (the reference's `mkwfs <lambda> 1` semantics).  An
 * optional FP32-storage mode (`prec_f32` arguments) rounds every stored
 * array element to float the way the reference's WFprecision=0 path stores
 * into float images while doing the per-pixel arithmetic in double
 * (AtmosphericTurbulence.c:1472-1615; COREMOD_arith.c:2746-2790).
 *
 * Complex arrays are interleaved (re, im) doubles, row-major with x (size[0])
 * fastest: element (ii, jj) of an n0 x n1 image is at index jj*n0 + ii, as in
 * the reference image registry (CLIcore.h:130-140, 210-243).
 *
 * Pinning status of each function is recorded in the header comment of
 * atsim_oracle.c.
 */
#ifndef ATSIM_ORACLE_H
#define ATSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- FFT (own implementation; FFTW sign and scaling) -------- */
/* Unnormalised 1D/2D c2c.  sign = -1 forward, +1 backward (fft.c:1298-1315). */
int  orc_fft1d(double *data, long n, int sign);
int  orc_fft2d(double *data, long n0, long n1, int sign);
/* brute-force O(n^2) DFT used to pin orc_fft1d */
void orc_dft1d_naive(const double *in, double *out, long n, int sign);

/* quadrant swap (fft.c:570-792), 2D, in place */
void orc_permut_c(double *data, long n0, long n1);
void orc_permut_r(double *data, long n0, long n1);

/* ---------------- Fresnel propagation (WFpropagate.c) -------------------- */
/* Fresnel_propagate_wavefront (WFpropagate.c:68-152): in -> out (may alias) */
int  orc_fresnel(const double *in, double *out, long n0, long n1,
                 double pupil_scale, double z, double lambda);
/* Init_Fresnel_propagate_wavefront (WFpropagate.c:166-261): size^2 TF, permuted */
void orc_fresnel_init_tf(double *C, long size, double pupil_scale, double z,
                         double lambda, double fpmaskrad);
/* Fresnel_propagate_wavefront1 (WFpropagate.c:271-330) */
int  orc_fresnel1(const double *in, double *out, const double *C, long n0, long n1);
/* Fresnel_propagate_cube (WFpropagate.c:337-401): amp/pha cubes n0*n1*nz */
int  orc_fresnel_cube(const double *in, double *amp, double *pha, long n0, long n1,
                      double pupil_scale, double zstart, double zend, long nz,
                      double lambda);

/* ---------------- noise (injected at the boundary) ----------------------- */
/* Philox4x32-10 keyed (seed, stream); counter = element index.  Writes
 * n complex N(0,1) values (re, im independent, Box-Muller in FP64), rounded
 * to float when out_f32 != NULL (interleaved re, im). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_noise_c64(uint64_t seed, uint32_t stream, long n, double *out_f64, float *out_f32);

/* ---------------- master screen synthesis (AtmosphericTurbulence.c:643-855) */
/* Spectral amplitude filter A(ii,jj) in the reference's centred layout. */
void orc_vk_filter(double *A, long size, double outerscale_px, double innerscale_px,
                   double rlim);
/* Structure function D at lag (a, b), circular, direct definition (info.c:1670-1693). */
double orc_strfn_direct(const double *s, long size, long a, long b);
/* Same quantity through the reference's FFT route (fft.c:1793-1848 + info.c:1670-1693),
 * full size x size map. */
void orc_strfn_fft(const double *s, long size, double *D);
/* C coefficient from the 4-lag log fit (AtmosphericTurbulence.c:794-847). */
double orc_strfn_coeff(const double *s, long size);
/* make_master_turbulence_screen with injected noise z (interleaved complex,
 * centred layout, size^2).  out1 = Re/sqrt(C1), out2 = Im/sqrt(C2). */
int orc_make_master(const double *noise, long size, double outerscale_px,
                    double innerscale_px, double rlim, int prec_f32,
                    double *out1, double *out2, double *C1, double *C2);

/* ---------------- per-run normalisation (AtmosphericTurbulence.c:2246-2363) */
/* r0 in pixels measured from the x lag dpix (non circular) */
double orc_r0_measure(const double *m, long size, long dpix);

/* ---------------- layer set-up (AtmosphericTurbulence.c:1947-2197) ------- */
/* Merge layers closer than bin into super-layers.  alt sorted as given.
 * Writes slayer_alt[<=L], super_layer_index[L]; returns number of super-layers. */
long orc_superlayers(const double *alt, const double *cn2, long L, double bin,
                     double *slayer_alt, long *super_layer_index);

/* ---------------- frame simulation --------------------------------------- */
typedef struct {
    long N;          /* WF_RAW_SIZE */
    long Nout;       /* WFsize */
    long M;          /* MASTER_SIZE */
    long L;          /* number of layers */
    long W;          /* number of fields (wavelengths) */
    double dx;       /* PUPIL_SCALE [m] */
    double dt;       /* WFTIME_STEP [s] */
    double zangle;   /* ZENITH_ANGLE [rad] */
    int bicubic;     /* 1 Catmull-Rom (reference default), 0 bilinear */
    int fresnel;     /* FRESNEL_PROPAGATION */
    int amplitude;   /* WAVEFRONT_AMPLITUDE */
    int prec_f32;    /* 1: float storage as WFprecision=0 */
    const double *masters;  /* L x M x M, already normalised and weighted */
    const double *slayer_alt;      /* super-layer altitudes */
    const long *super_layer_index; /* L */
    const double *lambda;          /* W wavelengths [m] */
    const double *scoeff;          /* W x L phase multipliers (1 for the reference lambda) */
    const int *unwrap_mode;        /* W: 0 = reference-lambda rule, 1 = block rule */
    double *xpos, *ypos;           /* L, updated in place (FP64 cumulative) */
    const double *vx, *vy;         /* L, pixels per second */
} orc_sim;

/* Advance one frame (positions first, as AtmosphericTurbulence.c:2700-2755)
 * and write per field: the complex field E (N x N, may be NULL), the
 * geometric phase PHI (N x N, may be NULL), output phase pha and amplitude
 * amp (Nout x Nout each, may be NULL).  Returns 0 on success. */
int orc_sim_frame(orc_sim *s, double *E, double *PHI, double *pha, double *amp);

/* Position update + wrap for one layer (AtmosphericTurbulence.c:2706-2755) */
void orc_advance_position(double *xpos, double *ypos, double vx, double vy,
                          double dt, long M);
/* Sample one layer at pixel (ii, jj) (AtmosphericTurbulence.c:2761-2922) */
double orc_sample(const double *master, long M, double xpos, double ypos,
                  long ii, long jj, int bicubic);

/* reference-lambda extraction + unwrap (AtmosphericTurbulence.c:3418-3476) */
void orc_extract_simple(const double *E, const double *PHI, long N, long Nout,
                        double *pha, double *amp);
/* science-lambda extraction + block unwrap (AtmosphericTurbulence.c:3523-4262).
 * Returns the iteration count of the fix-up loop. */
long orc_extract_block(const double *E, const double *PHI, long N, long Nout,
                       int prec_f32, double *pha, double *amp, double *peak_out);
/* The block fix-up loop alone (AtmosphericTurbulence.c:3627-4216) on an
 * xs x ys grid of block peaks, in place; ch_first (nullable) receives the
 * first round's counts; max_rounds <= 1000 caps the loop (1000 in the
 * reference).  Returns the number of rounds. */
long orc_block_fix(double *bin, long xs, long ys, int prec_f32, double *ch_first, long max_rounds);

#ifdef __cplusplus
}
#endif
#endif
