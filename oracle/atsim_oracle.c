/*
 * atsim_oracle.c -- CPU restatement of the reference wavefront path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  The product never links or calls it.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root; AT.c = src/AtmosphericTurbulence/AtmosphericTurbulence.c,
 * WFp.c = src/WFpropagate/WFpropagate.c, fft.c = src/fft/fft.c).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_fft1d / orc_fft2d   brute-force DFT (n <= 64), numpy.fft (pocketfft),
 *                           Parseval, impulse <-> constant, single-bin plane wave.
 *   orc_permut_*            fftshift for even sizes; swap-loop semantics on odd
 *                           sizes checked against a literal loop transcription.
 *   orc_fresnel             unitarity, z = 0 identity, additivity in z, uniform
 *                           field invariance, closed-form phase of an integer-
 *                           frequency plane wave, and an independent numpy
 *                           angular-spectrum implementation.
 *   orc_fresnel_init_tf     the separable closed form of the 4x4 sub-sample
 *                           average (pairs of offsets +-0.375, +-0.125 give
 *                           cos factors) with the 0.3 cut-off, the small-
 *                           coeff phase -coeff (r^2 + 0.15625), and at z = 0
 *                           the mask rule "at least 5 of 16 sub-samples
 *                           strictly inside FPMASKRAD" over many rim radii.
 *   orc_fresnel1 / cube     the exact table reproduces orc_fresnel; cube
 *                           slices equal single propagations.
 *   orc_philox4x32_10       Random123 known-answer vectors.
 *   orc_noise_c64           moments (mean 0, var 1 per component), independence.
 *   orc_vk_filter           closed form at chosen pixels, DC = 0.
 *   orc_strfn_*             direct definition == FFT route; closed form on a
 *                           sinusoid.
 *   orc_make_master         linearity in the noise, spectrum of the output
 *                           equals the filter (single noise bin), D(r) power law.
 *   orc_r0_measure          closed form on a linear ramp.
 *   orc_superlayers         hand-worked profiles.
 *   orc_sample              exact at integer shifts, bilinear exact on affine
 *                           fields, Catmull-Rom exact on quadratic fields,
 *                           periodic wrap.
 *   orc_sim_frame           single layer without propagation == e^{i v} of the
 *                           sampled screen; unitarity of the field energy
 *                           through gaps; equality with a numpy restatement of
 *                           the layer loop on small cases.
 *   orc_extract_simple      identity when |PHI - pha| < pi; 2 pi multiples.
 *   orc_extract_block       identity on smooth fields; hand-built block
 *                           inconsistency fixed by the +-2 pi heuristic.
 *   orc_block_fix           first-round counts equal a brute-force enumeration
 *                           of all block pairs at Chebyshev distance 1 (0.2)
 *                           and 2 (0.15) on random grids; a single planted
 *                           pair at every offset and grid position; the plim
 *                           rule (max - 0.001, floor 0.5) on hand-built count
 *                           maps; the interior rule and the five quiet rounds
 *                           from oscillating blocks inside / outside the
 *                           interior (1000 / 5 rounds); float counts in the
 *                           reference's per-block step order
 *                           (tests/test_oracle_block.py).
 */
#include "atsim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double rf(double x) { return (double)(float)x; }

/* ========================================================================
 * FFT -- own implementation.  FFTW conventions: unnormalised, sign -1 =
 * forward, +1 = backward; the reference calls fftw_plan_dft_2d with the
 * first two axes swapped so the row-major 2D DFT is over (size[1], size[0])
 * (fft.c:1178-1183, 1190).
 * ======================================================================== */

static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

typedef struct {
    long n;
    int sign;
    double *tw;  /* n/2 complex twiddles e^{sign 2 pi i k / n} */
    long *rev;   /* bit reversal (pow2) */
    double *tab; /* n complex e^{sign 2 pi i k / n} (non pow2 naive DFT) */
} fft_plan;

static int plan_make(fft_plan *p, long n, int sign)
{
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->sign = sign;
    if (is_pow2(n)) {
        long h = n / 2 > 0 ? n / 2 : 1;
        p->tw = (double *)malloc(sizeof(double) * 2 * h);
        p->rev = (long *)malloc(sizeof(long) * n);
        if (!p->tw || !p->rev) return -1;
        for (long k = 0; k < n / 2; k++) {
            double a = 2.0 * M_PI * (double)k / (double)n;
            p->tw[2 * k] = cos(a);
            p->tw[2 * k + 1] = sign * sin(a);
        }
        int lg = 0;
        while ((1L << lg) < n) lg++;
        for (long i = 0; i < n; i++) {
            long r = 0;
            for (int b = 0; b < lg; b++)
                if (i & (1L << b)) r |= 1L << (lg - 1 - b);
            p->rev[i] = r;
        }
    } else {
        p->tab = (double *)malloc(sizeof(double) * 2 * n);
        if (!p->tab) return -1;
        for (long k = 0; k < n; k++) {
            double a = 2.0 * M_PI * (double)k / (double)n;
            p->tab[2 * k] = cos(a);
            p->tab[2 * k + 1] = sign * sin(a);
        }
    }
    return 0;
}

static void plan_free(fft_plan *p)
{
    free(p->tw);
    free(p->rev);
    free(p->tab);
    memset(p, 0, sizeof(*p));
}

/* contiguous interleaved complex vector x[n], scratch of n complex */
static void plan_exec(const fft_plan *p, double *x, double *scratch)
{
    long n = p->n;
    if (n <= 1) return;
    if (p->rev) {
        for (long i = 0; i < n; i++) {
            long r = p->rev[i];
            if (r > i) {
                double t0 = x[2 * i], t1 = x[2 * i + 1];
                x[2 * i] = x[2 * r];
                x[2 * i + 1] = x[2 * r + 1];
                x[2 * r] = t0;
                x[2 * r + 1] = t1;
            }
        }
        for (long len = 2; len <= n; len <<= 1) {
            long half = len >> 1, step = n / len;
            for (long i = 0; i < n; i += len) {
                for (long k = 0; k < half; k++) {
                    double wr = p->tw[2 * k * step], wi = p->tw[2 * k * step + 1];
                    double *a = x + 2 * (i + k), *b = x + 2 * (i + k + half);
                    double vr = b[0] * wr - b[1] * wi;
                    double vi = b[0] * wi + b[1] * wr;
                    b[0] = a[0] - vr;
                    b[1] = a[1] - vi;
                    a[0] += vr;
                    a[1] += vi;
                }
            }
        }
    } else {
        for (long k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (long j = 0; j < n; j++) {
                long m = (long)(((unsigned long long)j * (unsigned long long)k) % (unsigned long long)n);
                double wr = p->tab[2 * m], wi = p->tab[2 * m + 1];
                sr += x[2 * j] * wr - x[2 * j + 1] * wi;
                si += x[2 * j] * wi + x[2 * j + 1] * wr;
            }
            scratch[2 * k] = sr;
            scratch[2 * k + 1] = si;
        }
        memcpy(x, scratch, sizeof(double) * 2 * n);
    }
}

void orc_dft1d_naive(const double *in, double *out, long n, int sign)
{
    for (long k = 0; k < n; k++) {
        double sr = 0.0, si = 0.0;
        for (long j = 0; j < n; j++) {
            long m = (long)(((unsigned long long)j * (unsigned long long)k) % (unsigned long long)n);
            double a = 2.0 * M_PI * (double)m / (double)n;
            double wr = cos(a), wi = sign * sin(a);
            sr += in[2 * j] * wr - in[2 * j + 1] * wi;
            si += in[2 * j] * wi + in[2 * j + 1] * wr;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
}

int orc_fft1d(double *data, long n, int sign)
{
    fft_plan p;
    if (plan_make(&p, n, sign)) return -1;
    double *scr = (double *)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
    if (!scr) { plan_free(&p); return -1; }
    plan_exec(&p, data, scr);
    free(scr);
    plan_free(&p);
    return 0;
}

int orc_fft2d(double *data, long n0, long n1, int sign)
{
    fft_plan p0, p1;
    if (plan_make(&p0, n0, sign)) return -1;
    if (plan_make(&p1, n1, sign)) { plan_free(&p0); return -1; }
    long nmax = n0 > n1 ? n0 : n1;
    const long B = 16; /* columns per block */
    double *scr = (double *)malloc(sizeof(double) * 2 * nmax);
    double *blk = (double *)malloc(sizeof(double) * 2 * n1 * B);
    if (!scr || !blk) { free(scr); free(blk); plan_free(&p0); plan_free(&p1); return -1; }
    for (long j = 0; j < n1; j++) plan_exec(&p0, data + 2 * j * n0, scr);
    for (long i0 = 0; i0 < n0; i0 += B) {
        long nb = (n0 - i0) < B ? (n0 - i0) : B;
        for (long j = 0; j < n1; j++)
            for (long b = 0; b < nb; b++) {
                blk[2 * (b * n1 + j)] = data[2 * (j * n0 + i0 + b)];
                blk[2 * (b * n1 + j) + 1] = data[2 * (j * n0 + i0 + b) + 1];
            }
        for (long b = 0; b < nb; b++) plan_exec(&p1, blk + 2 * b * n1, scr);
        for (long j = 0; j < n1; j++)
            for (long b = 0; b < nb; b++) {
                data[2 * (j * n0 + i0 + b)] = blk[2 * (b * n1 + j)];
                data[2 * (j * n0 + i0 + b) + 1] = blk[2 * (b * n1 + j) + 1];
            }
    }
    free(scr);
    free(blk);
    plan_free(&p0);
    plan_free(&p1);
    return 0;
}

/* ========================================================================
 * permut -- quadrant swap, fft.c:570-792 (2D branch, e.g. fft.c:698-711).
 * First loop swaps (ii, jj) with (ii+xh, jj+yh) for jj < yh; second swaps
 * (ii, jj) with (ii+xh, jj-yh) for jj >= yh; ii < xh in both.  For even
 * sizes this is fftshift (an involution).
 * ======================================================================== */
void orc_permut_c(double *d, long n0, long n1)
{
    long xh = n0 / 2, yh = n1 / 2;
    for (long jj = 0; jj < yh; jj++)
        for (long ii = 0; ii < xh; ii++) {
            long a = jj * n0 + ii, b = (jj + yh) * n0 + (ii + xh);
            double t0 = d[2 * a], t1 = d[2 * a + 1];
            d[2 * a] = d[2 * b]; d[2 * a + 1] = d[2 * b + 1];
            d[2 * b] = t0; d[2 * b + 1] = t1;
        }
    for (long jj = yh; jj < n1; jj++)
        for (long ii = 0; ii < xh; ii++) {
            long a = jj * n0 + ii, b = (jj - yh) * n0 + (ii + xh);
            double t0 = d[2 * a], t1 = d[2 * a + 1];
            d[2 * a] = d[2 * b]; d[2 * a + 1] = d[2 * b + 1];
            d[2 * b] = t0; d[2 * b + 1] = t1;
        }
}

void orc_permut_r(double *d, long n0, long n1)
{
    long xh = n0 / 2, yh = n1 / 2;
    for (long jj = 0; jj < yh; jj++)
        for (long ii = 0; ii < xh; ii++) {
            long a = jj * n0 + ii, b = (jj + yh) * n0 + (ii + xh);
            double t = d[a]; d[a] = d[b]; d[b] = t;
        }
    for (long jj = yh; jj < n1; jj++)
        for (long ii = 0; ii < xh; ii++) {
            long a = jj * n0 + ii, b = (jj - yh) * n0 + (ii + xh);
            double t = d[a]; d[a] = d[b]; d[b] = t;
        }
}

/* ========================================================================
 * Fresnel propagation
 * ======================================================================== */

/* Fresnel_propagate_wavefront, WFp.c:68-152: FFT(-1) -> permut -> multiply by
 * exp(-i coeff ((ii-n0/2)^2 + (jj-n1/2)^2)) / (n0 n1) with
 * coeff = pi z lambda / (dx n0)^2 (n0 used for both axes, WFp.c:97) ->
 * permut -> FFT(+1). */
int orc_fresnel(const double *in, double *out, long n0, long n1,
                double pupil_scale, double z, double lambda)
{
    long n = n0 * n1;
    double *t = (double *)malloc(sizeof(double) * 2 * n);
    if (!t) return -1;
    memcpy(t, in, sizeof(double) * 2 * n);
    if (orc_fft2d(t, n0, n1, -1)) { free(t); return -1; }
    orc_permut_c(t, n0, n1);
    double coeff = M_PI * z * lambda / (pupil_scale * n0) / (pupil_scale * n0);
    double co1 = 1.0 * n0 * n1;
    long n0h = n0 / 2;
    for (long jj = 0; jj < n1; jj++) {
        long jj2 = (jj - n1 / 2) * (jj - n1 / 2);
        for (long ii = 0; ii < n0; ii++) {
            long k = jj * n0 + ii;
            long ii2 = ii - n0h;
            double sqdist = (double)(ii2 * ii2 + jj2);
            double angle = -coeff * sqdist;
            double re = t[2 * k] / co1, im = t[2 * k + 1] / co1;
            double c = cos(angle), s = sin(angle);
            t[2 * k] = re * c - im * s;
            t[2 * k + 1] = re * s + im * c;
        }
    }
    orc_permut_c(t, n0, n1);
    if (orc_fft2d(t, n0, n1, 1)) { free(t); return -1; }
    memcpy(out, t, sizeof(double) * 2 * n);
    free(t);
    return 0;
}

/* Init_Fresnel_propagate_wavefront, WFp.c:166-261 (double branch): 4 x 4
 * sub-sampled TF average, amplitude thresholded at 0.3 to 0/1, phase kept,
 * divided by size^2, then permut. */
void orc_fresnel_init_tf(double *C, long size, double pupil_scale, double z,
                         double lambda, double fpmaskrad)
{
    const long BIN = 4;
    double coeff = M_PI * z * lambda / (pupil_scale * size) / (pupil_scale * size);
    double co1 = 1.0 * size * size;
    for (long ii = 0; ii < size; ii++)
        for (long jj = 0; jj < size; jj++) {
            double Re = 0.0, Im = 0.0;
            for (long i = 0; i < BIN; i++)
                for (long j = 0; j < BIN; j++) {
                    double x = 1.0 * ii + 1.0 * (0.5 + i - BIN / 2) / BIN - size / 2;
                    double y = 1.0 * jj + 1.0 * (0.5 + j - BIN / 2) / BIN - size / 2;
                    double sqdist = x * x + y * y;
                    double angle = -coeff * sqdist;
                    if (sqrt(sqdist) < fpmaskrad) {
                        Re += cos(angle);
                        Im += sin(angle);
                    }
                }
            double Amp = sqrt(Re * Re + Im * Im) / BIN / BIN;
            double Pha = atan2(Im, Re);
            Amp = (Amp > 0.3) ? 1.0 : 0.0;
            C[2 * (jj * size + ii)] = Amp * cos(Pha) / co1;
            C[2 * (jj * size + ii) + 1] = Amp * sin(Pha) / co1;
        }
    orc_permut_c(C, size, size);
}

/* Fresnel_propagate_wavefront1, WFp.c:271-330: FFT(-1), multiply by C, FFT(+1). */
int orc_fresnel1(const double *in, double *out, const double *C, long n0, long n1)
{
    long n = n0 * n1;
    double *t = (double *)malloc(sizeof(double) * 2 * n);
    if (!t) return -1;
    memcpy(t, in, sizeof(double) * 2 * n);
    if (orc_fft2d(t, n0, n1, -1)) { free(t); return -1; }
    for (long k = 0; k < n; k++) {
        double re = t[2 * k], im = t[2 * k + 1];
        double rr = C[2 * k], ri = C[2 * k + 1];
        t[2 * k] = re * rr - im * ri;
        t[2 * k + 1] = re * ri + im * rr;
    }
    if (orc_fft2d(t, n0, n1, 1)) { free(t); return -1; }
    memcpy(out, t, sizeof(double) * 2 * n);
    free(t);
    return 0;
}

/* Fresnel_propagate_cube, WFp.c:337-401: z_k = zstart + (zend - zstart) k / nz. */
int orc_fresnel_cube(const double *in, double *amp, double *pha, long n0, long n1,
                     double pupil_scale, double zstart, double zend, long nz,
                     double lambda)
{
    long n = n0 * n1;
    double *t = (double *)malloc(sizeof(double) * 2 * n);
    if (!t) return -1;
    for (long kk = 0; kk < nz; kk++) {
        double zprop = zstart + (zend - zstart) * kk / nz;
        if (orc_fresnel(in, t, n0, n1, pupil_scale, zprop, lambda)) { free(t); return -1; }
        for (long k = 0; k < n; k++) {
            double re = t[2 * k], im = t[2 * k + 1];
            amp[kk * n + k] = sqrt(re * re + im * im);
            pha[kk * n + k] = atan2(im, re);
        }
    }
    free(t);
    return 0;
}

/* ========================================================================
 * Noise: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), used in place of
 * the reference's clock-seeded rand()/gsl generators (statistic.c:91-111,
 * CLIcore.c:696-698) so that the GPU and the CPU see the same input.
 * ======================================================================== */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* element e of stream `stream`: counter (e_lo, e_hi, 0x0A75, 0), key
 * (seed_lo, seed_hi ^ stream); two 53-bit uniforms; Box-Muller in FP64. */
void orc_noise_c64(uint64_t seed, uint32_t stream, long n, double *out_f64, float *out_f32)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ stream};
    for (long e = 0; e < n; e++) {
        uint32_t ctr[4] = {(uint32_t)e, (uint32_t)((uint64_t)e >> 32), 0x0A75u, 0u};
        uint32_t x[4];
        orc_philox4x32_10(ctr, key, x);
        uint64_t a = (((uint64_t)x[0] << 32) | x[1]) >> 11;
        uint64_t b = (((uint64_t)x[2] << 32) | x[3]) >> 11;
        double u1 = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
        double u2 = (double)b * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        double t = 2.0 * M_PI * u2;
        double re = r * cos(t), im = r * sin(t);
        if (out_f64) { out_f64[2 * e] = re; out_f64[2 * e + 1] = im; }
        if (out_f32) { out_f32[2 * e] = (float)re; out_f32[2 * e + 1] = (float)im; }
    }
}

/* ========================================================================
 * Master screen synthesis, AT.c:643-855
 * ======================================================================== */

/* Spectral amplitude: sqrt(exp(-r^2/fi^2)) / (r^2 + f0^2)^(11/12), with
 * f0 = size / outerscale (AT.c:679), fi = (5.92/(2 pi)) size / innerscale
 * (AT.c:680); r measured from (size/2, size/2) (AT.c:701-702, 752-753); the
 * RLIM variable zeroes the denominator inside rlim (AT.c:704-709, here the
 * amplitude is set to 0 there); DC pixel = 0 (AT.c:778). */
void orc_vk_filter(double *A, long size, double outerscale_px, double innerscale_px,
                   double rlim)
{
    double f0 = 1.0 * size / outerscale_px;
    double fi = (5.92 / (2.0 * M_PI)) * size / innerscale_px;
    for (long jj = 0; jj < size; jj++)
        for (long ii = 0; ii < size; ii++) {
            double dx = 1.0 * ii - size / 2;
            double dy = 1.0 * jj - size / 2;
            double tmpd = sqrt(dx * dx + dy * dy + f0 * f0);
            if (rlim > 0.0 && sqrt(dx * dx + dy * dy) < rlim) tmpd = 0.0;
            double isc = exp(-(dx * dx + dy * dy) / fi / fi);
            double a = 0.0;
            if (tmpd > 0.0) a = sqrt(isc) / pow(tmpd, 11.0 / 6.0);
            A[jj * size + ii] = a;
        }
    A[(size / 2) * size + size / 2] = 0.0;
}

/* D(a, b) = mean over pixels of (s(x) - s(x + (a, b)))^2, circular; this is
 * what the FFT route of fft_structure_function computes (info.c:1670-1693,
 * fft.c:1793-1848); the map element [b*size + a] holds lag (a, b). */
double orc_strfn_direct(const double *s, long size, long a, long b)
{
    double tot = 0.0;
    for (long jj = 0; jj < size; jj++) {
        long j2 = ((jj + b) % size + size) % size;
        for (long ii = 0; ii < size; ii++) {
            long i2 = ((ii + a) % size + size) % size;
            double d = s[jj * size + ii] - s[j2 * size + i2];
            tot += d * d;
        }
    }
    return tot / ((double)size * (double)size);
}

/* FFT route: AC = Re FFT(|FFT(s)|^2 / (sqrt(n) n)); D = (AC - AC[0]) (-2/sqrt(n)). */
void orc_strfn_fft(const double *s, long size, double *D)
{
    long n = size * size;
    double *t = (double *)malloc(sizeof(double) * 2 * n);
    for (long k = 0; k < n; k++) { t[2 * k] = s[k]; t[2 * k + 1] = 0.0; }
    orc_fft2d(t, size, size, -1);
    double sc = 1.0 / sqrt((double)n) / (1.0 * n);
    for (long k = 0; k < n; k++) {
        double a2 = t[2 * k] * t[2 * k] + t[2 * k + 1] * t[2 * k + 1];
        t[2 * k] = a2 * sc;
        t[2 * k + 1] = 0.0;
    }
    orc_fft2d(t, size, size, -1);
    double v0 = -t[0];
    for (long k = 0; k < n; k++) D[k] = (t[2 * k] + v0) * (-2.0 / sqrt((double)n));
    free(t);
}

/* AT.c:794-816: C = 10^(mean over (ii,jj) in {1,2}^2 of
 * log10 D(ii,jj) - (5/3) log10 sqrt(ii^2+jj^2)). */
double orc_strfn_coeff(const double *s, long size)
{
    const long Dlim = 3;
    double value = 0.0;
    long cnt = 0;
    for (long ii = 1; ii < Dlim; ii++)
        for (long jj = 1; jj < Dlim; jj++) {
            value += log10(orc_strfn_direct(s, size, ii, jj)) -
                     5.0 / 3.0 * log10(sqrt((double)(ii * ii + jj * jj)));
            cnt++;
        }
    return pow(10.0, value / cnt);
}

/* AT.c:643-855 with the noise injected: C = A z, permut, forward FFT,
 * out1 = Re * (1/sqrt(C1)), out2 = Im * (1/sqrt(C2)) (AT.c:849-850). */
int orc_make_master(const double *noise, long size, double outerscale_px,
                    double innerscale_px, double rlim, int prec_f32,
                    double *out1, double *out2, double *C1, double *C2)
{
    long n = size * size;
    double *A = (double *)malloc(sizeof(double) * n);
    double *c = (double *)malloc(sizeof(double) * 2 * n);
    if (!A || !c) { free(A); free(c); return -1; }
    orc_vk_filter(A, size, outerscale_px, innerscale_px, rlim);
    for (long k = 0; k < n; k++) {
        double a = prec_f32 ? rf(A[k]) : A[k];
        c[2 * k] = a * noise[2 * k];
        c[2 * k + 1] = a * noise[2 * k + 1];
        if (prec_f32) { c[2 * k] = rf(c[2 * k]); c[2 * k + 1] = rf(c[2 * k + 1]); }
    }
    free(A);
    orc_permut_c(c, size, size);
    if (orc_fft2d(c, size, size, -1)) { free(c); return -1; }
    for (long k = 0; k < n; k++) {
        out1[k] = prec_f32 ? rf(c[2 * k]) : c[2 * k];
        out2[k] = prec_f32 ? rf(c[2 * k + 1]) : c[2 * k + 1];
    }
    free(c);
    double c1 = orc_strfn_coeff(out1, size);
    double c2 = orc_strfn_coeff(out2, size);
    double s1 = 1.0 / sqrt(c1), s2 = 1.0 / sqrt(c2);
    for (long k = 0; k < n; k++) {
        out1[k] *= s1;
        out2[k] *= s2;
        if (prec_f32) { out1[k] = rf(out1[k]); out2[k] = rf(out2[k]); }
    }
    if (C1) *C1 = c1;
    if (C2) *C2 = c2;
    return 0;
}

/* AT.c:2246-2270: D50 over ii < size - dpix (x lag only, non circular),
 * r0 = dpix ((tot/cnt)/6.88)^(-3/5). */
double orc_r0_measure(const double *m, long size, long dpix)
{
    double tot = 0.0;
    long cnt = 0;
    for (long ii = 0; ii < size - dpix; ii++)
        for (long jj = 0; jj < size; jj++) {
            double p1 = m[jj * size + ii];
            double p2 = m[jj * size + ii + dpix];
            tot += (p1 - p2) * (p1 - p2);
            cnt++;
        }
    return 1.0 * dpix * pow((tot / cnt) / 6.88, -3.0 / 5.0);
}

/* ========================================================================
 * Super-layers, AT.c:2067-2150
 * ======================================================================== */
long orc_superlayers(const double *alt, const double *cn2, long L, double bin,
                     double *sa, long *sli)
{
    double *sc = (double *)malloc(sizeof(double) * (L > 0 ? L : 1));
    long S = L;
    for (long i = 0; i < L; i++) { sa[i] = alt[i]; sc[i] = cn2[i]; }
    for (;;) {
        double minaltd = alt[L - 1];
        long index = 0;
        for (long i = 0; i < S - 1; i++) {
            double v = sa[i + 1] - sa[i];
            if (v < minaltd) { minaltd = v; index = i; }
        }
        if (minaltd > bin || S == 1) break;
        sa[index] = (sc[index] * sa[index] + sc[index + 1] * sa[index + 1]) /
                    (sc[index] + sc[index + 1]);
        sc[index] = sc[index] + sc[index + 1];
        for (long i = index + 1; i < S - 1; i++) { sa[i] = sa[i + 1]; sc[i] = sc[i + 1]; }
        S -= 1;
    }
    free(sc);
    long nsep = S - 1;
    double *sep = (double *)malloc(sizeof(double) * (nsep > 0 ? nsep : 1));
    for (long i = 0; i < nsep; i++) sep[i] = 0.5 * (sa[i] + sa[i + 1]);
    for (long l = 0; l < L; l++) {
        long index = 0;
        /* AT.c:2141-2146; the reference reads sep[nsep] on its last
         * iteration, which the following test overrides, so it is skipped. */
        for (long i = 0; i + 1 < nsep; i++)
            if (sep[i] < alt[l] && sep[i + 1] > alt[l]) index = i + 1;
        if (nsep > 0 && alt[l] > sep[nsep - 1]) index = nsep;
        sli[l] = index;
    }
    free(sep);
    return S;
}

/* ========================================================================
 * Per-frame loop, AT.c:2561-3413
 * ======================================================================== */
void orc_advance_position(double *xpos, double *ypos, double vx, double vy,
                          double dt, long M)
{
    *xpos += vx * dt;
    *ypos += vy * dt;
    while (*xpos < 0) *xpos += 1.0 * M;
    while (*xpos > 1.0 * M) *xpos -= 1.0 * M;
    while (*ypos < 0) *ypos += 1.0 * M;
    while (*ypos > 1.0 * M) *ypos -= 1.0 * M;
}

/* Catmull-Rom (Keys, a = -1/2) cubic weights for taps at offsets -1, 0, 1, 2.
 * The reference spells the same interpolant out as 16 polynomial
 * coefficients a_ij (AT.c:2884-2910); the tensor product of these weights is
 * that polynomial. */
static void cr_weights(double t, double w[4])
{
    double t2 = t * t, t3 = t2 * t;
    w[0] = 0.5 * (-t3 + 2.0 * t2 - t);
    w[1] = 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0);
    w[2] = 0.5 * (-3.0 * t3 + 4.0 * t2 + t);
    w[3] = 0.5 * (t3 - t2);
}

double orc_sample(const double *F, long M, double xpos, double ypos,
                  long ii, long jj, int bicubic)
{
    double iimf = fmod(xpos + ii, 1.0 * M);
    double jjmf = fmod(ypos + jj, 1.0 * M);
    long iim = (long)iimf, jjm = (long)jjmf;
    if (!bicubic) {
        /* AT.c:2761-2790 */
        double fx = iimf - iim, fy = jjmf - jjm;
        long iim1 = iim + 1, jjm1 = jjm + 1;
        if (iim == M) iim = 0;
        if (jjm == M) jjm = 0;
        if (iim1 > M - 1) iim1 -= M;
        if (jjm1 > M - 1) jjm1 -= M;
        double v = (1.0 - fx) * (1.0 - fy) * F[jjm * M + iim];
        v += (1.0 - fx) * fy * F[jjm1 * M + iim];
        v += fx * fy * F[jjm1 * M + iim1];
        v += fx * (1.0 - fy) * F[jjm * M + iim1];
        return v;
    }
    /* AT.c:2803-2912 */
    double x = iimf - iim, y = jjmf - jjm;
    long ix[4], jy[4];
    for (int k = 0; k < 4; k++) {
        long a = iim - 1 + k, b = jjm - 1 + k;
        if (a > M - 1) a -= M;
        if (a < 0) a += M;
        if (b > M - 1) b -= M;
        if (b < 0) b += M;
        ix[k] = a;
        jy[k] = b;
    }
    double wx[4], wy[4];
    cr_weights(x, wx);
    cr_weights(y, wy);
    double v = 0.0;
    for (int b = 0; b < 4; b++) {
        double row = 0.0;
        for (int a = 0; a < 4; a++) row += wx[a] * F[jy[b] * M + ix[a]];
        v += wy[b] * row;
    }
    return v;
}

static double unwrap_count(double x)
{
    /* (long)(x + 1000.5) - 1000, AT.c:3456 / 4241 */
    return (double)((long)(x + 1000.5) - 1000);
}

int orc_sim_frame(orc_sim *s, double *E_out, double *PHI_out, double *pha, double *amp)
{
    long N = s->N, Nout = s->Nout, M = s->M, L = s->L, W = s->W;
    long n = N * N;
    double *E = E_out ? E_out : (double *)malloc(sizeof(double) * 2 * n * W);
    double *PHI = PHI_out ? PHI_out : (double *)malloc(sizeof(double) * n * W);
    if (!E || !PHI) return -1;
    /* AT.c:2566-2633 */
    for (long k = 0; k < n * W; k++) { E[2 * k] = 1.0; E[2 * k + 1] = 0.0; PHI[k] = 0.0; }
    for (long layer = L - 1; layer != -1; layer--) {
        /* AT.c:2666-2692 */
        if (layer != L - 1 && s->super_layer_index[layer + 1] != s->super_layer_index[layer] &&
            s->fresnel && s->amplitude) {
            double dz = (s->slayer_alt[s->super_layer_index[layer + 1]] -
                         s->slayer_alt[s->super_layer_index[layer]]) / cos(s->zangle);
            for (long w = 0; w < W; w++) {
                double *Ew = E + 2 * n * w;
                if (orc_fresnel(Ew, Ew, N, N, s->dx, dz, s->lambda[w])) return -1;
                if (s->prec_f32)
                    for (long k = 0; k < 2 * n; k++) Ew[k] = rf(Ew[k]);
            }
        }
        orc_advance_position(&s->xpos[layer], &s->ypos[layer], s->vx[layer], s->vy[layer],
                             s->dt, M);
        const double *F = s->masters + (size_t)layer * M * M;
        for (long ii = 0; ii < N; ii++)
            for (long jj = 0; jj < N; jj++) {
                double v = orc_sample(F, M, s->xpos[layer], s->ypos[layer], ii, jj, s->bicubic);
                long k = jj * N + ii;
                for (long w = 0; w < W; w++) {
                    double val = v * s->scoeff[w * L + layer];
                    double *P = PHI + n * w;
                    P[k] += val;
                    if (s->prec_f32) P[k] = rf(P[k]);
                    if (s->amplitude) {
                        double *Ew = E + 2 * n * w;
                        double re = Ew[2 * k], im = Ew[2 * k + 1];
                        double c = cos(val), sn = sin(val);
                        Ew[2 * k] = re * c - im * sn;
                        Ew[2 * k + 1] = re * sn + im * c;
                        if (s->prec_f32) { Ew[2 * k] = rf(Ew[2 * k]); Ew[2 * k + 1] = rf(Ew[2 * k + 1]); }
                    }
                }
            }
    }
    if (pha || amp) {
        long no = Nout * Nout;
        for (long w = 0; w < W; w++) {
            double *pw = pha ? pha + no * w : NULL;
            double *aw = amp ? amp + no * w : NULL;
            double *tp = pw ? NULL : (double *)malloc(sizeof(double) * no);
            double *ta = aw ? NULL : (double *)malloc(sizeof(double) * no);
            if (!s->amplitude) {
                long o = (N - Nout) / 2;
                for (long jj = 0; jj < Nout; jj++)
                    for (long ii = 0; ii < Nout; ii++) {
                        (pw ? pw : tp)[jj * Nout + ii] = PHI[n * w + (jj + o) * N + ii + o];
                        (aw ? aw : ta)[jj * Nout + ii] = 1.0;
                    }
            } else if (s->unwrap_mode[w] == 0) {
                orc_extract_simple(E + 2 * n * w, PHI + n * w, N, Nout, pw ? pw : tp, aw ? aw : ta);
            } else {
                orc_extract_block(E + 2 * n * w, PHI + n * w, N, Nout, s->prec_f32,
                                  pw ? pw : tp, aw ? aw : ta, NULL);
            }
            free(tp);
            free(ta);
        }
    }
    if (!E_out) free(E);
    if (!PHI_out) free(PHI);
    return 0;
}

/* AT.c:3418-3476: centred crop, amp = |E|, pha = atan2 + 2 pi ((long)(PHI/2pi + 1000.5) - 1000). */
void orc_extract_simple(const double *E, const double *PHI, long N, long Nout,
                        double *pha, double *amp)
{
    long o = (N - Nout) / 2;
    for (long ii = 0; ii < Nout; ii++)
        for (long jj = 0; jj < Nout; jj++) {
            long k = (jj + o) * N + ii + o;
            double re = E[2 * k], im = E[2 * k + 1];
            double phi = PHI[k];
            amp[jj * Nout + ii] = sqrt(re * re + im * im);
            double p = atan2(im, re);
            pha[jj * Nout + ii] = p + 2.0 * M_PI * unwrap_count(phi / 2.0 / M_PI);
        }
}

/* The twelve neighbour relations of the block fix-up (AT.c:3667-3903):
 * block (ii2+dx1, jj2+dy1) vs (ii2+dx2, jj2+dy2) for ii2 < xs - xm,
 * jj2 < ys - ym, weight c. */
static const struct { int dx1, dy1, dx2, dy2, xm, ym; double c; } BLK_PAIRS[12] = {
    {0, 0, 1, 0, 1, 0, 0.2},  {0, 0, 0, 1, 0, 1, 0.2},  {0, 0, 1, 1, 1, 1, 0.2},
    {0, 1, 1, 0, 1, 1, 0.2},  {0, 0, 2, 0, 2, 0, 0.15}, {0, 2, 0, 0, 0, 2, 0.15},
    {1, 2, 0, 0, 1, 2, 0.15}, {0, 2, 1, 0, 1, 2, 0.15}, {2, 2, 0, 0, 2, 2, 0.15},
    {0, 2, 2, 0, 2, 2, 0.15}, {0, 1, 2, 0, 2, 1, 0.15}, {2, 1, 0, 0, 2, 1, 0.15},
};

/* AT.c:3627-4216: the iterative +-2 pi consistency fix-up on an xs x ys grid
 * of block peaks bin[jj2 * xs + ii2], in place.  Each round clears the
 * counts, adds +-c for every relation of BLK_PAIRS whose peaks differ by more
 * than pi (the block with the larger peak gets -c, the other +c), in the
 * reference's loop order (relation, then ii2, then jj2), sets
 * plim = max(max|count| - 0.001, 0.5) and moves every block whose |count|
 * exceeds plim by -+2 pi; a round that moves no interior block (ii2, jj2 in
 * [2, size - 3]) counts towards the five that end the loop.  With prec_f32
 * the counts and peaks are floats as with WFprecision = 0 (every += rounds).
 * ch_first (nullable, xs * ys) receives the first round's counts.
 * max_rounds caps the loop (the reference's cap is 1000; a smaller one stops
 * the loop early for tests).  Returns the number of rounds. */
long orc_block_fix(double *bin, long xs, long ys, int prec_f32, double *ch_first, long max_rounds)
{
    long nb = xs * ys;
    double *ch = (double *)calloc(nb > 0 ? nb : 1, sizeof(double));
    if (prec_f32)
        for (long b = 0; b < nb; b++) bin[b] = rf(bin[b]);
    long chcnt = 1, chcnt0cnt = 0, chiter = 0;
    double plim = 1.0;
    if (max_rounds <= 0 || max_rounds > 1000) max_rounds = 1000;
    while (plim > 0.51 && chiter < max_rounds && chcnt0cnt < 5) {
        chiter++;
        chcnt = 0;
        for (long b = 0; b < nb; b++) ch[b] = 0.0;
        for (int q = 0; q < 12; q++) {
            double c = BLK_PAIRS[q].c;
            for (long ii2 = 0; ii2 < xs - BLK_PAIRS[q].xm; ii2++)
                for (long jj2 = 0; jj2 < ys - BLK_PAIRS[q].ym; jj2++) {
                    long i1 = (jj2 + BLK_PAIRS[q].dy1) * xs + ii2 + BLK_PAIRS[q].dx1;
                    long i2 = (jj2 + BLK_PAIRS[q].dy2) * xs + ii2 + BLK_PAIRS[q].dx2;
                    double pv1 = bin[i1], pv2 = bin[i2];
                    if (pv2 > pv1 + M_PI) {
                        ch[i1] += c;
                        if (prec_f32) ch[i1] = rf(ch[i1]);
                        ch[i2] -= c;
                        if (prec_f32) ch[i2] = rf(ch[i2]);
                    }
                    if (pv1 > pv2 + M_PI) {
                        ch[i1] -= c;
                        if (prec_f32) ch[i1] = rf(ch[i1]);
                        ch[i2] += c;
                        if (prec_f32) ch[i2] = rf(ch[i2]);
                    }
                }
        }
        if (chiter == 1 && ch_first)
            for (long b = 0; b < nb; b++) ch_first[b] = ch[b];
        plim = 0.0;
        for (long b = 0; b < nb; b++)
            if (fabs(ch[b]) > plim) plim = fabs(ch[b]);
        plim -= 0.001;
        if (plim < 0.5) plim = 0.5;
        for (long ii2 = 0; ii2 < xs; ii2++)
            for (long jj2 = 0; jj2 < ys; jj2++) {
                long b = jj2 * xs + ii2;
                int interior = (ii2 > 1) && (jj2 > 1) && (ii2 < xs - 2) && (jj2 < ys - 2);
                if (ch[b] > plim) {
                    if (interior) chcnt++;
                    bin[b] += 2.0 * M_PI;
                    if (prec_f32) bin[b] = rf(bin[b]);
                }
                if (ch[b] < -plim) {
                    if (interior) chcnt++;
                    bin[b] -= 2.0 * M_PI;
                    if (prec_f32) bin[b] = rf(bin[b]);
                }
            }
        if (chcnt == 0) chcnt0cnt++;
        else chcnt0cnt = 0;
    }
    free(ch);
    return chiter;
}

/* AT.c:3523-4262. */
long orc_extract_block(const double *E, const double *PHI, long N, long Nout,
                       int prec_f32, double *pha, double *amp, double *peak_out)
{
    long o = (N - Nout) / 2;
    long xs = Nout / 20, ys = Nout / 20;
    long nb = xs * ys;
    double *bre = (double *)calloc(nb > 0 ? nb : 1, sizeof(double));
    double *bim = (double *)calloc(nb > 0 ? nb : 1, sizeof(double));
    double *bin = (double *)calloc(nb > 0 ? nb : 1, sizeof(double));
    for (long ii = 0; ii < Nout; ii++)
        for (long jj = 0; jj < Nout; jj++) {
            long k = (jj + o) * N + ii + o;
            double re = E[2 * k], im = E[2 * k + 1];
            amp[jj * Nout + ii] = sqrt(re * re + im * im);
            double p = atan2(im, re);
            pha[jj * Nout + ii] = p;
            long ii2 = (long)(1.0 * ii / Nout * xs);
            long jj2 = (long)(1.0 * jj / Nout * ys);
            if (ii2 < xs && jj2 < ys) {
                long b = jj2 * xs + ii2;
                bre[b] += cos(PHI[k] - p);
                bim[b] += sin(PHI[k] - p);
                if (prec_f32) { bre[b] = rf(bre[b]); bim[b] = rf(bim[b]); }
            }
        }
    for (long b = 0; b < nb; b++) {
        bin[b] = atan2(bim[b], bre[b]);
        if (prec_f32) bin[b] = rf(bin[b]);
    }
    long chiter = orc_block_fix(bin, xs, ys, prec_f32, NULL, 1000);
    double peakpha = 0.0;
    for (long ii = 0; ii < Nout; ii++)
        for (long jj = 0; jj < Nout; jj++) {
            long ii2 = (long)(1.0 * ii / Nout * xs);
            long jj2 = (long)(1.0 * jj / Nout * ys);
            if (ii2 < xs && jj2 < ys) peakpha = bin[jj2 * xs + ii2];
            long k = (jj + o) * N + ii + o;
            double p = pha[jj * Nout + ii];
            pha[jj * Nout + ii] = p + 2.0 * M_PI * unwrap_count((PHI[k] - p - peakpha) / 2.0 / M_PI);
        }
    if (peak_out)
        for (long b = 0; b < nb; b++) peak_out[b] = bin[b];
    free(bre); free(bim); free(bin);
    return chiter;
}
