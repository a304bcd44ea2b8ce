/* compat_ops.c -- the reference's fft / WFpropagate / screen entry points on
 * registry images, computed by the HIP library.
 *
 *   do2dfft, do2dffti, permut            src/fft/fft.c:570-792, 1145-1315
 *   Fresnel_propagate_wavefront          src/WFpropagate/WFpropagate.c:68-152
 *   Init_Fresnel_propagate_wavefront     WFpropagate.c:166-261
 *   Fresnel_propagate_wavefront1         WFpropagate.c:271-330
 *   Fresnel_propagate_cube               WFpropagate.c:337-401
 *   make_master_turbulence_screen        AtmosphericTurbulence.c:643-855
 *
 * Images are staged to the device as complex float (COMPLEX_DOUBLE images are
 * converted on the way in and out: the device path computes in FP32).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "atsim_host.h"

static unsigned long long g_seed = 0xA75ULL;
static unsigned int g_stream = 0;
static double g_rlim = 0.0;

void atsim_compat_set_seed(unsigned long long seed) {
    g_seed = seed;
    g_stream = 0;
}
void atsim_compat_set_rlim(double rlim_px) { g_rlim = rlim_px > 0 ? rlim_px : 0.0; }

static int is_complex(int t) { return t == ATSIM_COMPLEX_FLOAT || t == ATSIM_COMPLEX_DOUBLE; }

/* device copy (complex float) of a complex image; *n0, *n1, *nb its shape */
static void *to_device_c64(atsim_image *im, long *n0, long *n1, long *nb) {
    atsim_ctx *c = atsim_compat_ctx();
    if (!c) return NULL;
    atsim_image_metadata *md = im->md;
    *n0 = md->size[0];
    *n1 = md->naxis > 1 ? md->size[1] : 1;
    *nb = md->naxis > 2 ? md->size[2] : 1;
    const long n = md->nelement;
    void *d = NULL;
    if (atsim_malloc(c, &d, (size_t)n * 8)) return NULL;
    if (md->atype == ATSIM_COMPLEX_FLOAT) {
        atsim_copy(c, d, im->array, (size_t)n * 8);
    } else {
        float *h = (float *)malloc((size_t)n * 8);
        const double *src = (const double *)im->array;
        for (long k = 0; k < 2 * n; k++) h[k] = (float)src[k];
        atsim_copy(c, d, h, (size_t)n * 8);
        free(h);
    }
    return d;
}

static void from_device_c64(atsim_image *im, const void *d) {
    atsim_ctx *c = atsim_compat_ctx();
    const long n = im->md->nelement;
    if (im->md->atype == ATSIM_COMPLEX_FLOAT) {
        atsim_copy(c, im->array, d, (size_t)n * 8);
    } else {
        float *h = (float *)malloc((size_t)n * 8);
        atsim_copy(c, h, d, (size_t)n * 8);
        double *dst = (double *)im->array;
        for (long k = 0; k < 2 * n; k++) dst[k] = h[k];
        free(h);
    }
}

static long new_like(const char *name, atsim_image *src, int atype) {
    long ID = image_ID(name);
    if (ID != -1) delete_image_ID(name);
    long sz[3] = {src->md->size[0], src->md->size[1], src->md->size[2]};
    return create_image_ID(name, src->md->naxis, sz, atype, 0, 0);
}

static int check_square_pow2(const char *who, long n0, long n1) {
    if (n0 != n1 || n0 < 16 || n0 > 8192 || (n0 & (n0 - 1)))
        return atsim_compat_fail(0, "%s: the device path needs a square power-of-two image in [16, 8192] (got %ld x %ld)",
                                 who, n0, n1);
    return 0;
}

static long fft_common(const char *in_name, const char *out_name, int sign) {
    atsim_image *in = atsim_reg_get(image_ID(in_name));
    if (!in) return atsim_compat_fail(0, "do2dfft: image \"%s\" does not exist", in_name);
    if (!is_complex(in->md->atype)) return atsim_compat_fail(0, "do2dfft: image \"%s\" is not complex", in_name);
    long n0, n1, nb;
    void *d = to_device_c64(in, &n0, &n1, &nb);
    if (!d) return -1;
    atsim_ctx *c = atsim_compat_ctx();
    /* any shape, as FFTW plans it (fft.c:1190) */
    int rc = atsim_fft2d_nm(c, d, n0, n1, nb, sign);
    long ID = -1;
    if (rc == 0) {
        ID = new_like(out_name, in, in->md->atype);
        atsim_image *out = atsim_reg_get(ID);
        if (out) from_device_c64(out, d);
    }
    atsim_free(c, d);
    if (rc) return atsim_compat_fail(0, "do2dfft: %s", atsim_error(c));
    return ID;
}

long do2dfft(const char *in_name, const char *out_name) { return fft_common(in_name, out_name, -1); }
long do2dffti(const char *in_name, const char *out_name) { return fft_common(in_name, out_name, +1); }

int permut(const char *ID_name) {
    atsim_image *im = atsim_reg_get(image_ID(ID_name));
    if (!im) return atsim_compat_fail(0, "permut: image \"%s\" does not exist", ID_name);
    atsim_ctx *c = atsim_compat_ctx();
    if (!c) return -1;
    atsim_image_metadata *md = im->md;
    const long n0 = md->size[0], n1 = md->naxis > 1 ? md->size[1] : 1, nb = md->naxis > 2 ? md->size[2] : 1;
    const long n = md->nelement;
    /* real images travel as (value, 0) pairs */
    float *h = (float *)malloc((size_t)n * 8);
    for (long k = 0; k < n; k++) {
        switch (md->atype) {
            case ATSIM_COMPLEX_FLOAT: h[2 * k] = ((float *)im->array)[2 * k]; h[2 * k + 1] = ((float *)im->array)[2 * k + 1]; break;
            case ATSIM_COMPLEX_DOUBLE: h[2 * k] = (float)((double *)im->array)[2 * k]; h[2 * k + 1] = (float)((double *)im->array)[2 * k + 1]; break;
            case ATSIM_FLOAT: h[2 * k] = ((float *)im->array)[k]; h[2 * k + 1] = 0.f; break;
            case ATSIM_DOUBLE: h[2 * k] = (float)((double *)im->array)[k]; h[2 * k + 1] = 0.f; break;
            default: free(h); return atsim_compat_fail(0, "permut: unsupported type %d", md->atype);
        }
    }
    void *d = NULL;
    int rc = atsim_malloc(c, &d, (size_t)n * 8);
    if (!rc) rc = atsim_copy(c, d, h, (size_t)n * 8);
    if (!rc) rc = atsim_permut(c, d, n0, n1, nb);
    if (!rc) rc = atsim_copy(c, h, d, (size_t)n * 8);
    atsim_free(c, d);
    if (rc) {
        free(h);
        return atsim_compat_fail(0, "permut: %s", atsim_error(c));
    }
    /* a device permut moves whole elements: values come back bit for bit
     * (double images were already float-representable only if they came from
     * float data; keep their full values by permuting a copy on the index) */
    if (md->atype == ATSIM_DOUBLE || md->atype == ATSIM_COMPLEX_DOUBLE) {
        /* the device moved the float copies; apply the same move to the
         * double values through the index travelling in the imaginary slot */
        free(h);
        float *idx = (float *)malloc((size_t)n * 8);
        for (long k = 0; k < n; k++) {
            idx[2 * k] = (float)(k & 0xFFFFFF);
            idx[2 * k + 1] = (float)(k >> 24);
        }
        rc = atsim_malloc(c, &d, (size_t)n * 8);
        if (!rc) rc = atsim_copy(c, d, idx, (size_t)n * 8);
        if (!rc) rc = atsim_permut(c, d, n0, n1, nb);
        if (!rc) rc = atsim_copy(c, idx, d, (size_t)n * 8);
        atsim_free(c, d);
        if (rc) {
            free(idx);
            return atsim_compat_fail(0, "permut: %s", atsim_error(c));
        }
        const size_t es = atsim_type_size(md->atype);
        char *copy = (char *)malloc((size_t)n * es);
        memcpy(copy, im->array, (size_t)n * es);
        for (long k = 0; k < n; k++) {
            long src = (long)idx[2 * k] + ((long)idx[2 * k + 1] << 24);
            memcpy((char *)im->array + k * es, copy + src * es, es);
        }
        free(copy);
        free(idx);
        return 0;
    }
    for (long k = 0; k < n; k++) {
        if (md->atype == ATSIM_COMPLEX_FLOAT) {
            ((float *)im->array)[2 * k] = h[2 * k];
            ((float *)im->array)[2 * k + 1] = h[2 * k + 1];
        } else {
            ((float *)im->array)[k] = h[2 * k];
        }
    }
    free(h);
    return 0;
}

int init_WFpropagate(void) { return 0; }

int Fresnel_propagate_wavefront(const char *in, const char *out, double PUPIL_SCALE, double z, double lambda) {
    atsim_image *im = atsim_reg_get(image_ID(in));
    if (!im) return atsim_compat_fail(0, "Fresnel_propagate_wavefront: image \"%s\" does not exist", in);
    if (!is_complex(im->md->atype)) return atsim_compat_fail(0, "Fresnel_propagate_wavefront: \"%s\" is not complex", in);
    long n0, n1, nb;
    void *d = to_device_c64(im, &n0, &n1, &nb);
    if (!d) return -1;
    atsim_ctx *c = atsim_compat_ctx();
    int rc = atsim_fresnel_nm(c, d, d, n0, n1, nb, PUPIL_SCALE, z, lambda);
    if (rc == 0) {
        long ID = new_like(out, im, im->md->atype);
        atsim_image *o = atsim_reg_get(ID);
        if (o) from_device_c64(o, d);
    }
    atsim_free(c, d);
    if (rc) return atsim_compat_fail(0, "Fresnel_propagate_wavefront: %s", atsim_error(c));
    return 0;
}

int Init_Fresnel_propagate_wavefront(const char *Cim, long size, double PUPIL_SCALE, double z, double lambda,
                                     double FPMASKRAD, int Precision) {
    atsim_ctx *c = atsim_compat_ctx();
    if (!c) return -1;
    if (size <= 0 || (size & 1)) return atsim_compat_fail(0, "Init_Fresnel_propagate_wavefront: even size required");
    void *d = NULL;
    int rc = atsim_malloc(c, &d, (size_t)size * size * 8);
    if (!rc) rc = atsim_fresnel_tf_init(c, d, size, PUPIL_SCALE, z, lambda, FPMASKRAD);
    if (rc) {
        atsim_free(c, d);
        return atsim_compat_fail(0, "Init_Fresnel_propagate_wavefront: %s", atsim_error(c));
    }
    if (image_ID(Cim) != -1) delete_image_ID(Cim);
    long ID = Precision == 0 ? create_2DCimage_ID(Cim, size, size) : create_2DCimage_ID_double(Cim, size, size);
    atsim_image *o = atsim_reg_get(ID);
    if (o) from_device_c64(o, d);
    atsim_free(c, d);
    return 0;
}

int Fresnel_propagate_wavefront1(const char *in, const char *out, const char *Cin) {
    atsim_image *im = atsim_reg_get(image_ID(in));
    atsim_image *C = atsim_reg_get(image_ID(Cin));
    if (!im || !C) return atsim_compat_fail(0, "Fresnel_propagate_wavefront1: missing image");
    if (!is_complex(im->md->atype) || !is_complex(C->md->atype))
        return atsim_compat_fail(0, "Fresnel_propagate_wavefront1: complex images required");
    if (check_square_pow2("Fresnel_propagate_wavefront1", im->md->size[0], im->md->size[1])) return -1;
    long n0, n1, nb, c0, c1, cb;
    void *d = to_device_c64(im, &n0, &n1, &nb);
    void *dc = to_device_c64(C, &c0, &c1, &cb);
    atsim_ctx *c = atsim_compat_ctx();
    int rc = (!d || !dc) ? -1 : (c0 != n0 || c1 != n1) ? -2 : atsim_fresnel_tf_apply(c, d, d, dc, n0, nb);
    if (rc == 0) {
        long ID = new_like(out, im, im->md->atype);
        atsim_image *o = atsim_reg_get(ID);
        if (o) from_device_c64(o, d);
    }
    atsim_free(c, d);
    atsim_free(c, dc);
    if (rc) return atsim_compat_fail(0, "Fresnel_propagate_wavefront1: %s", rc == -2 ? "size mismatch" : atsim_error(c));
    return 0;
}

long Fresnel_propagate_cube(const char *IDcin_name, const char *IDout_name_amp, const char *IDout_name_pha,
                            double PUPIL_SCALE, double zstart, double zend, long NBzpts, double lambda) {
    atsim_image *im = atsim_reg_get(image_ID(IDcin_name));
    if (!im || !is_complex(im->md->atype)) return atsim_compat_fail(0, "Fresnel_propagate_cube: complex input required");
    if (check_square_pow2("Fresnel_propagate_cube", im->md->size[0], im->md->size[1])) return -1;
    long n0, n1, nb;
    void *d = to_device_c64(im, &n0, &n1, &nb);
    atsim_ctx *c = atsim_compat_ctx();
    void *da = NULL, *dp = NULL;
    const size_t cube = (size_t)n0 * n1 * NBzpts;
    int rc = !d ? -1 : atsim_malloc(c, &da, cube * 4);
    if (!rc) rc = atsim_malloc(c, &dp, cube * 4);
    if (!rc) rc = atsim_fresnel_cube(c, d, (float *)da, (float *)dp, n0, PUPIL_SCALE, zstart, zend, NBzpts, lambda);
    if (!rc) {
        const int dbl = im->md->atype == ATSIM_COMPLEX_DOUBLE;
        long IDa = dbl ? create_3Dimage_ID_double(IDout_name_amp, n0, n1, NBzpts) : create_3Dimage_ID(IDout_name_amp, n0, n1, NBzpts);
        long IDp = dbl ? create_3Dimage_ID_double(IDout_name_pha, n0, n1, NBzpts) : create_3Dimage_ID(IDout_name_pha, n0, n1, NBzpts);
        float *h = (float *)malloc(cube * 4);
        for (int k = 0; k < 2; k++) {
            long ID = k ? IDp : IDa;
            atsim_copy(c, h, k ? dp : da, cube * 4);
            void *arr = atsim_image_data(ID);
            for (size_t i = 0; i < cube; i++) {
                if (dbl) ((double *)arr)[i] = h[i];
                else ((float *)arr)[i] = h[i];
            }
        }
        free(h);
    }
    atsim_free(c, d);
    atsim_free(c, da);
    atsim_free(c, dp);
    if (rc) return atsim_compat_fail(0, "Fresnel_propagate_cube: %s", atsim_error(c));
    return image_ID(IDout_name_amp);
}

/* noise stream g_stream of g_seed on the device, filter, FFT, normalise */
int make_master_turbulence_screen(const char *ID_name1, const char *ID_name2, long size, float outerscale,
                                  float innerscale, long WFprecision) {
    atsim_ctx *c = atsim_compat_ctx();
    if (!c) return -1;
    if (size < 16 || size > 8192 || (size & (size - 1)))
        return atsim_compat_fail(0, "make_master_turbulence_screen: size %ld must be a power of two in [16, 8192]", size);
    const size_t n = (size_t)size * size;
    void *dz = NULL, *d1 = NULL, *d2 = NULL;
    double C1 = 0, C2 = 0;
    int rc = atsim_malloc(c, &dz, n * 8);
    if (!rc) rc = atsim_malloc(c, &d1, n * 4);
    if (!rc) rc = atsim_malloc(c, &d2, n * 4);
    if (!rc) rc = atsim_noise_c64(c, dz, (long)n, g_seed, g_stream++);
    if (!rc) rc = atsim_make_master(c, dz, size, (double)outerscale, (double)innerscale, g_rlim, (float *)d1, (float *)d2, &C1, &C2);
    if (!rc) {
        printf("C1, C2 =   %f %f\n", C1, C2);
        float *h = (float *)malloc(n * 4);
        for (int k = 0; k < 2; k++) {
            const char *name = k ? ID_name2 : ID_name1;
            if (image_ID(name) != -1) delete_image_ID(name);
            long ID = WFprecision == 0 ? create_2Dimage_ID(name, size, size) : create_2Dimage_ID_double(name, size, size);
            atsim_copy(c, h, k ? d2 : d1, n * 4);
            void *arr = atsim_image_data(ID);
            for (size_t i = 0; i < n; i++) {
                if (WFprecision == 0) ((float *)arr)[i] = h[i];
                else ((double *)arr)[i] = h[i];
            }
        }
        free(h);
    }
    atsim_free(c, dz);
    atsim_free(c, d1);
    atsim_free(c, d2);
    if (rc) return atsim_compat_fail(0, "make_master_turbulence_screen: %s", atsim_error(c));
    return 0;
}
