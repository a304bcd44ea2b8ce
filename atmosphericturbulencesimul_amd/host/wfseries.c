/* wfseries.c -- make_AtmosphericTurbulence_wavefront_series (the `mkwfs`
 * command) on the MI355X library.
 *
 * Steps, in the reference's order (AtmosphericTurbulence.c:1433-4455):
 *   configuration (1654) -> output cubes (1889-1918) -> master files
 *   t%03ld_%ld_{f,d}.fits counted / loaded / synthesised (1925-1942,
 *   2204-2226) -> profile, Cn2 normalisation (1947-2022) -> super-layers
 *   (2067-2150) -> offsets and velocities (2152-2197) -> r0 renormalisation
 *   (2246-2363) -> shared-memory images (2443-2483) -> SKIP_EXISTING
 *   (2489-2512) -> frame loop (2561-4345) -> FITS cubes per time span
 *   (4349-4404).
 * Every frame's layer loop, Fresnel gaps, extraction and unwrapping run on
 * the GPU (atsim_sim_frame) for the reference wavelength and, when
 * MAKE_SWAVEFRONT = 1, the science wavelength.
 *
 * The atmosphere model (AtmosphereModel.c, NRLMSISE-00 and refractive-index
 * tables) is outside this library: refractivity is Edlen's standard-air
 * dispersion scaled by an exponential density profile (scale height 8 km),
 * which gives the science-wavelength phase factor Scoeff = (lref / ls)
 * N(ls) / N(lref) (AT.c:2661-2663) and the refraction offsets of AT.c:2172-2190.
 */
#define _GNU_SOURCE
#include <fcntl.h>
#include <math.h>
#include <semaphore.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "atsim_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static char g_conffile[200] = "WFsim.conf"; /* AtmosphericTurbulence.c:65 */

int init_AtmosphericTurbulence(void) { return 0; }

int AtmosphericTurbulence_change_configuration_file(const char *fname) {
    snprintf(g_conffile, sizeof g_conffile, "%s", fname);
    return 0;
}

/* Edlen standard-air refractivity (n - 1) at sea level for lambda [m], scaled
 * by exp(-h / 8000 m) */
double atsim_air_refractivity(double h, double lambda) {
    const double s2 = 1.0 / (lambda * 1e6 * lambda * 1e6);
    const double n0 = 0.0000834213 + 0.0240603 / (130.0 - s2) + 0.00015997 / (38.9 - s2);
    return n0 * exp(-h / 8000.0);
}

/* sum over 1 m steps from the site to the layer of the refraction slope
 * difference (AtmosphericTurbulence.c:2172-2181) */
static double refraction_offset(double site, double alt, double zangle, double lambda) {
    const double sz = sin(zangle);
    double r = 0.0;
    for (double h = site; h < alt; h += 1.0) {
        const double n = 1.0 + atsim_air_refractivity(h, lambda);
        r += sz / sqrt(n * n - sz * sz) - sz / sqrt(1.0 - sz * sz);
    }
    return r;
}

static int copy_cube_slice(long ID, long frame, const float *src, long no) {
    void *arr = atsim_image_data(ID);
    if (!arr) return -1;
    if (atsim_image_type(ID) == ATSIM_FLOAT) memcpy((float *)arr + frame * no, src, sizeof(float) * no);
    else
        for (long i = 0; i < no; i++) ((double *)arr)[frame * no + i] = src[i];
    return 0;
}

static void shm_write(long ID, const float *src, long no, double scale, int tkw, double t) {
    atsim_image *im = atsim_reg_get(ID);
    if (!im) return;
    im->md->write = 1;
    if (tkw && im->kw) {
        im->kw[0].value.numf = t;
    }
    if (im->md->atype == ATSIM_FLOAT)
        for (long i = 0; i < no; i++) ((float *)im->array)[i] = (float)(src[i] * scale);
    else
        for (long i = 0; i < no; i++) ((double *)im->array)[i] = src[i] * scale;
    __atomic_thread_fence(__ATOMIC_RELEASE);
    im->md->cnt0++;
    im->md->write = 0;
    atsim_image_post(ID);
}

static long shm_image(const char *name, long n, int dbl, int timekw) {
    long sz[2] = {n, n};
    if (image_ID(name) != -1) delete_image_ID(name);
    long ID = create_image_ID(name, 2, sz, dbl ? ATSIM_DOUBLE : ATSIM_FLOAT, 1, timekw ? 1 : 0);
    atsim_image *im = atsim_reg_get(ID);
    if (im && timekw && im->kw) {
        snprintf(im->kw[0].name, sizeof im->kw[0].name, "TIME");
        im->kw[0].type = 'D';
        im->kw[0].value.numf = 0.0;
        snprintf(im->kw[0].comment, sizeof im->kw[0].comment, "Physical time [sec]");
    }
    return ID;
}

int make_AtmosphericTurbulence_wavefront_series(float slambdaum, long WFprecision) {
    atsim_conf cf;
    atsim_conf_defaults(&cf);
    if (atsim_conf_read(g_conffile, &cf)) return -1;
    const double SLAMBDA = 1.0e-6 * slambdaum;
    const double dx = cf.pupil_scale, dt = cf.time_step, lref = cf.lambda, zang = cf.zangle;
    const long N = cf.raw_size, Nout = cf.wfsize, M = cf.master_size;
    /* AT.c:1853-1870: WF_RAW_SIZE / WFsize must be 1, 2, 4 or 8 */
    if (Nout <= 0 || N % Nout || (N / Nout != 1 && N / Nout != 2 && N / Nout != 4 && N / Nout != 8))
        return atsim_compat_fail(0, "ERROR: unknown contraction factor");
    const long NBFRAMES = (long)(1.0 * cf.time_span / cf.time_step + 0.5);
    if (NBFRAMES <= 0) return atsim_compat_fail(0, "no frames (TIME_SPAN / WFTIME_STEP)");
    const long no = Nout * Nout;
    const int dbl = WFprecision != 0;
    printf("Making the wavefront series ... N = %ld  Nout = %ld  M = %ld  frames = %ld\n", N, Nout, M, NBFRAMES);

    /* output cubes (AT.c:1889-1918), amp 1 and phase 0 */
    const char *cube_names[4] = {"outarraypha", "outarrayamp", "outsarraypha", "outsarrayamp"};
    long cube[4];
    for (int k = 0; k < 4; k++) {
        if (image_ID(cube_names[k]) != -1) delete_image_ID(cube_names[k]);
        cube[k] = dbl ? create_3Dimage_ID_double(cube_names[k], Nout, Nout, NBFRAMES)
                      : create_3Dimage_ID(cube_names[k], Nout, Nout, NBFRAMES);
        if (cube[k] < 0) return -1;
        void *arr = atsim_image_data(cube[k]);
        for (long i = 0; i < no * NBFRAMES; i++) {
            const double v = (k & 1) ? 1.0 : 0.0;
            if (dbl) ((double *)arr)[i] = v;
            else ((float *)arr)[i] = (float)v;
        }
    }

    /* profile and Cn2 normalisation (AT.c:1947-2022) */
    atsim_profile pr;
    if (atsim_profile_read(cf.prof_file, cf.site_alt, &pr)) return -1;
    const long L = pr.n;
    if (L <= 0) {
        atsim_profile_free(&pr);
        return atsim_compat_fail(1, "no turbulence layer above the site in \"%s\"", cf.prof_file);
    }
    double tot = 0.0;
    for (long l = 0; l < L; l++) tot += pr.cn2[l];
    for (long l = 0; l < L; l++) pr.cn2[l] *= 1.0 / tot;

    /* master screens (AT.c:1925-1942, 2200-2240) */
    long nbm = 0;
    char fname[512], name[64];
    for (;;) {
        snprintf(fname, sizeof fname, dbl ? "t%03ld_%ld_d.fits" : "t%03ld_%ld_f.fits", nbm, M);
        if (!file_exists(fname)) break;
        nbm++;
    }
    printf("%ld turbulence master files\n", nbm);
    if (nbm < L) nbm = L;
    atsim_compat_set_seed(cf.seed);
    for (long i = 0; i < nbm; i++) {
        snprintf(fname, sizeof fname, dbl ? "t%03ld_%ld_d.fits" : "t%03ld_%ld_f.fits", i, M);
        snprintf(name, sizeof name, "TM%ld", i);
        if (load_fits(fname, name, 0) == -1) {
            /* a 4-column profile has no scales (the reference then reads
             * uninitialised values): no outer scale is an infinite one, no
             * inner scale no inner cut-off */
            const double L0px = pr.outer[i] > 0 ? pr.outer[i] / dx : 1e30;
            const double l0px = pr.inner[i] > 0 ? pr.inner[i] / dx : 1e-30;
            printf("CREATING %s   (%f - %f)\n", fname, L0px, l0px);
            if (make_master_turbulence_screen(name, "tursctmp", M, (float)L0px, (float)l0px, WFprecision))
                return -1;
            char out[520];
            snprintf(out, sizeof out, "!%s", fname);
            if (dbl) save_db_fits(name, out);
            else save_fl_fits(name, out);
            delete_image_ID("tursctmp");
        }
    }

    /* super-layers (AT.c:2067-2150) */
    double *sa = (double *)malloc(sizeof(double) * L);
    long *sli = (long *)malloc(sizeof(long) * L);
    long S = atsim_superlayers(pr.alt, pr.cn2, L, cf.fresnel_bin, sa, sli);
    if (S < 1) {
        free(sa); free(sli); atsim_profile_free(&pr);
        return atsim_compat_fail(0, "profile altitudes must be ascending");
    }
    printf("%ld super layers\n", S);

    /* wavelengths: reference lambda (Scoeff 1, AT.c:3456 unwrap), science lambda */
    const long W = cf.make_swavefront == 1 ? 2 : 1;
    double lam[2] = {lref, SLAMBDA};
    double *scoeff = (double *)malloc(sizeof(double) * W * L);
    int unwrap[2] = {0, 1};
    for (long l = 0; l < L; l++) {
        scoeff[l] = 1.0;
        if (W == 2)
            scoeff[L + l] = lref / SLAMBDA * atsim_air_refractivity(pr.alt[l], SLAMBDA) / atsim_air_refractivity(pr.alt[l], lref);
    }
    /* offsets and velocities (AT.c:2152-2197) */
    double *x0 = (double *)malloc(sizeof(double) * L), *y0 = (double *)malloc(sizeof(double) * L);
    double *xr = (double *)malloc(sizeof(double) * L), *yr = (double *)malloc(sizeof(double) * L);
    double *vx = (double *)malloc(sizeof(double) * L), *vy = (double *)malloc(sizeof(double) * L);
    for (long l = 0; l < L; l++) {
        const double roff = zang != 0.0 ? refraction_offset(cf.site_alt, pr.alt[l], zang, lref) : 0.0;
        const double roffs = zang != 0.0 ? refraction_offset(cf.site_alt, pr.alt[l], zang, SLAMBDA) : 0.0;
        x0[l] = 0.05 * M + cf.source_x * pr.alt[l] / dx;
        y0[l] = 0.05 * M + (roffs - roff) / dx + cf.source_y * pr.alt[l] / dx;
        xr[l] = cf.source_x * pr.alt[l] / dx; /* REALTIME origin (AT.c:2193-2194) */
        yr[l] = cf.source_y * pr.alt[l] / dx;
        vx[l] = pr.speed[l] * cos(pr.dir[l]) / dx;
        vy[l] = pr.speed[l] * sin(pr.dir[l]) / dx;
    }

    atsim_ctx *c = atsim_compat_ctx();
    atsim_sim *sim = NULL;
    atsim_sim_desc d;
    memset(&d, 0, sizeof d);
    d.N = N; d.Nout = Nout; d.M = M; d.L = L; d.W = W; d.D = 1;
    d.dx = dx; d.dt = dt; d.zangle = zang;
    d.bicubic = cf.bicubic; d.fresnel = cf.fresnel; d.amplitude = cf.amplitude == 1;
    d.slayer_alt = sa; d.super_layer_index = sli; d.lambda = lam; d.scoeff = scoeff; d.unwrap_mode = unwrap;
    d.xpos0 = x0; d.ypos0 = y0; d.vx = vx; d.vy = vy;
    if (c) atsim_set_tuning(c, "graph", 1); /* same output buffers every frame: replay the frame's graph */
    int rc = c ? atsim_sim_create(c, &d, &sim) : -1;
    if (rc) {
        atsim_compat_fail(0, "simulation set-up failed: %s", c ? atsim_error(c) : "no GPU");
        goto done;
    }

    /* masters to the device, r0 of every master, renormalisation (AT.c:2246-2363) */
    {
        double r0tot = 0.0;
        float *h = (float *)malloc(sizeof(float) * M * M);
        void *dtmp = NULL;
        rc = atsim_malloc(c, &dtmp, sizeof(float) * M * M);
        for (long i = 0; i < nbm && !rc; i++) {
            snprintf(name, sizeof name, "TM%ld", i);
            long ID = image_ID(name);
            if (ID < 0 || atsim_image_size(ID, 0) != M || atsim_image_size(ID, 1) != M) {
                rc = atsim_compat_fail(0, "master %s missing or not %ld x %ld", name, M, M);
                break;
            }
            void *arr = atsim_image_data(ID);
            for (long k = 0; k < M * M; k++)
                h[k] = atsim_image_type(ID) == ATSIM_FLOAT ? ((float *)arr)[k] : (float)((double *)arr)[k];
            float *dst = i < L ? atsim_sim_master(sim, i) : (float *)dtmp;
            rc = atsim_copy(c, dst, h, sizeof(float) * M * M);
            double r0 = 0.0;
            if (!rc) rc = atsim_r0_measure(c, dst, M, 50, &r0);
            printf("TURBULENCE MASTER %ld    r0 = %g pix\n", i, r0);
            r0tot += r0;
        }
        atsim_free(c, dtmp);
        free(h);
        if (!rc) {
            const double r0 = r0tot / nbm;
            const double r0goal = lref / (cf.seeing / 3600.0 / 180.0 * M_PI) / dx;
            const double tmp1 = pow(r0 / r0goal, 5.0 / 6.0);
            printf("r0 = %g pix -> %g pix\n", r0, r0goal);
            for (long l = 0; l < L && !rc; l++) {
                rc = atsim_scale(c, atsim_sim_master(sim, l), M * M, tmp1);
                if (!rc) rc = atsim_scale(c, atsim_sim_master(sim, l), M * M, sqrt(pr.cn2[l] / cos(zang)));
            }
        }
        if (rc) {
            atsim_compat_fail(0, "master normalisation failed: %s", atsim_error(c));
            goto done;
        }
    }

    /* shared-memory images (AT.c:2443-2483) */
    long shm_pha = -1, shm_amp = -1, shm_spha = -1, shm_samp = -1;
    if (cf.shm_output == 1) {
        shm_pha = shm_image("atmwfpha", Nout, dbl, 0);
        shm_amp = shm_image("atmwfamp", Nout, dbl, 0);
    }
    if (cf.shm_soutput == 1) {
        char nm[160];
        snprintf(nm, sizeof nm, "%spha", cf.shm_sprefix);
        shm_spha = shm_image(nm, Nout, dbl, 1);
        snprintf(nm, sizeof nm, "%samp", cf.shm_sprefix);
        shm_samp = shm_image(nm, Nout, dbl, 1);
    }

    /* SKIP_EXISTING (AT.c:2489-2512) */
    long start_tspan = 0;
    if (cf.skip_existing) {
        for (;;) {
            snprintf(fname, sizeof fname, "%s%08ld.%09ld.pha.fits", cf.swf_prefix, start_tspan, (long)(SLAMBDA * 1e12 + 0.5));
            if (!file_exists(fname)) break;
            start_tspan++;
        }
    }
    printf("Start TSPAN = %ld\n", start_tspan);

    {
        /* two output sets: frame k computes into set k % 2 while the copy of
         * frame k-1 travels to pinned host memory on the copy stream; the
         * host writes cubes and shared memory for frame k-1 meanwhile */
        const size_t ob = sizeof(float) * W * no;
        void *dpha[2] = {NULL, NULL}, *damp[2] = {NULL, NULL};
        float *hp[2] = {NULL, NULL}, *ha[2] = {NULL, NULL};
        for (int b = 0; b < 2 && !rc; b++) {
            rc = atsim_malloc(c, &dpha[b], ob);
            if (!rc) rc = atsim_malloc(c, &damp[b], ob);
            if (!rc) rc = atsim_host_alloc(c, (void **)&hp[b], ob);
            if (!rc) rc = atsim_host_alloc(c, (void **)&ha[b], ob);
        }
        sem_t *waitsem = NULL;
        if (cf.waitforsem == 1) {
            char sn[256];
            snprintf(sn, sizeof sn, "/%s_sem00", cf.waitsemimname);
            waitsem = sem_open(sn, 0);
            if (waitsem == SEM_FAILED) {
                waitsem = NULL;
                printf("WARNING: semaphore %s not found, not waiting\n", sn);
            }
        }
        double scoeff_shm = 1.0; /* AT.c:4286-4296 */
        if (cf.shm_soutputm == 1) scoeff_shm = SLAMBDA / 2.0 / M_PI;
        else if (cf.shm_soutputm == 2) scoeff_shm = SLAMBDA / 2.0 / M_PI * 1e6;
        long k = 0;                    /* frames launched */
        long pend_frame = -1, pend_tspan = 0;
        double pend_t = 0.0;
        /* host side of the frame in set b: cube slices and shared memory */
        #define FINISH_PENDING()                                                                          \
            do {                                                                                          \
                if (pend_frame >= 0 && !rc) {                                                             \
                    const int b = (int)((k - 1) & 1);                                                     \
                    rc = atsim_copy_wait(c, 2 * b);                                                       \
                    if (!rc) rc = atsim_copy_wait(c, 2 * b + 1);                                          \
                    if (!rc) {                                                                            \
                        const float *p_ = hp[b], *a_ = ha[b];                                             \
                        copy_cube_slice(cube[0], pend_frame, p_, no);                                     \
                        copy_cube_slice(cube[1], pend_frame, a_, no);                                     \
                        if (W == 2) {                                                                     \
                            copy_cube_slice(cube[2], pend_frame, p_ + no, no);                            \
                            copy_cube_slice(cube[3], pend_frame, a_ + no, no);                            \
                        }                                                                                 \
                        if (shm_pha >= 0) {                                                               \
                            shm_write(shm_pha, p_, no, 1.0, 0, pend_t);                                   \
                            if (cf.amplitude == 1) shm_write(shm_amp, a_, no, 1.0, 0, pend_t);            \
                        }                                                                                 \
                        if (shm_spha >= 0 && W == 2) {                                                    \
                            shm_write(shm_spha, p_ + no, no, scoeff_shm, 1, pend_t);                      \
                            if (cf.amplitude == 1) shm_write(shm_samp, a_ + no, no, 1.0, 1, pend_t);      \
                        }                                                                                 \
                        printf("\rFrame %4ld/%4ld, File %6ld/%6ld  [TIME = %10.4f s]  ", pend_frame, NBFRAMES, \
                               pend_tspan, cf.nb_tspan, pend_t);                                          \
                        fflush(stdout);                                                                   \
                    }                                                                                     \
                    pend_frame = -1;                                                                      \
                }                                                                                         \
            } while (0)
        for (long tspan = start_tspan; tspan < cf.nb_tspan && !rc; tspan++) {
            for (long frame = 0; frame < NBFRAMES && !rc; frame++) {
                if (cf.simtdelay > 0) usleep((useconds_t)cf.simtdelay);
                if (waitsem) sem_wait(waitsem);
                double t = (tspan * NBFRAMES + frame) * dt;
                if (cf.realtime == 1) {
                    struct timespec ts;
                    clock_gettime(CLOCK_REALTIME, &ts);
                    t = (1.0 * ts.tv_sec + 1.0e-9 * ts.tv_nsec) * cf.realtimefactor;
                    /* positions at time t (AT.c:2700-2704); the frame advances them by v dt */
                    double *px = (double *)malloc(sizeof(double) * L), *py = (double *)malloc(sizeof(double) * L);
                    for (long l = 0; l < L; l++) {
                        px[l] = fmod(xr[l] + vx[l] * (t - dt), 1.0 * M);
                        py[l] = fmod(yr[l] + vy[l] * (t - dt), 1.0 * M);
                        if (px[l] < 0) px[l] += M;
                        if (py[l] < 0) py[l] += M;
                    }
                    atsim_sim_set_positions(sim, px, py);
                    free(px);
                    free(py);
                }
                const int b = (int)(k & 1);
                /* set b was last copied out by frame k - 2 */
                if (k >= 2) rc = atsim_stream_wait_copy(c, 2 * b);
                if (!rc && k >= 2) rc = atsim_stream_wait_copy(c, 2 * b + 1);
                if (!rc) rc = atsim_sim_frame(sim, (float *)dpha[b], (float *)damp[b], NULL);
                if (!rc) rc = atsim_copy_async(c, 2 * b, hp[b], dpha[b], ob);
                if (!rc) rc = atsim_copy_async(c, 2 * b + 1, ha[b], damp[b], ob);
                if (rc) break;
                FINISH_PENDING();
                k++;
                pend_frame = frame;
                pend_tspan = tspan;
                pend_t = t;
            }
            FINISH_PENDING();
            if (rc) break;
            const long lpm = (long)(1.0e12 * SLAMBDA + 0.5);
            char f1[600], f2[600];
            if (cf.wfoutput == 1) { /* AT.c:4349-4366 */
                snprintf(f1, sizeof f1, "!%s%08ld.%09ld.pha.fits", cf.wf_prefix, tspan, lpm);
                snprintf(f2, sizeof f2, "!%s%08ld.%09ld.amp.fits", cf.wf_prefix, tspan, lpm);
                if (dbl) save_db_fits("outarraypha", f1);
                else save_fl_fits("outarraypha", f1);
                if (cf.amplitude == 1) {
                    if (dbl) save_db_fits("outarrayamp", f2);
                    else save_fl_fits("outarrayamp", f2);
                }
            } else if (cf.wfoutput == 0) { /* empty files, AT.c:4367-4380 */
                snprintf(f1, sizeof f1, "%s%08ld.%09ld.pha.fits", cf.wf_prefix, tspan, lpm);
                FILE *fp = fopen(f1, "a");
                if (fp) fclose(fp);
                if (cf.amplitude == 1) {
                    snprintf(f2, sizeof f2, "%s%08ld.%09ld.amp.fits", cf.wf_prefix, tspan, lpm);
                    fp = fopen(f2, "a");
                    if (fp) fclose(fp);
                }
            }
            if (W == 2 && cf.swf_write2disk == 1) { /* AT.c:4383-4404 */
                snprintf(f1, sizeof f1, "!%s%08ld.%09ld.pha.fits", cf.swf_prefix, tspan, lpm);
                snprintf(f2, sizeof f2, "!%s%08ld.%09ld.amp.fits", cf.swf_prefix, tspan, lpm);
                if (dbl) save_db_fits("outsarraypha", f1);
                else save_fl_fits("outsarraypha", f1);
                if (cf.amplitude == 1) {
                    if (dbl) save_db_fits("outsarrayamp", f2);
                    else save_fl_fits("outsarrayamp", f2);
                }
            }
        }
        #undef FINISH_PENDING
        printf("\n");
        if (waitsem) sem_close(waitsem);
        atsim_sync(c);
        for (int b = 0; b < 2; b++) {
            atsim_copy_wait(c, 2 * b);
            atsim_copy_wait(c, 2 * b + 1);
            atsim_host_free(c, hp[b]);
            atsim_host_free(c, ha[b]);
            atsim_free(c, dpha[b]);
            atsim_free(c, damp[b]);
        }
        if (rc) atsim_compat_fail(0, "frame loop failed: %s", atsim_error(c));
    }

done:
    if (sim) atsim_sim_destroy(sim);
    free(sa); free(sli); free(scoeff); free(x0); free(y0); free(xr); free(yr); free(vx); free(vy);
    atsim_profile_free(&pr);
    return rc ? -1 : 0;
}
