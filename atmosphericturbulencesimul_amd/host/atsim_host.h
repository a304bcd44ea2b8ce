/* atsim_host.h -- internal declarations of the C host layer (compat API,
 * registry, FITS, shared memory, configuration, CLI). */
#ifndef ATSIM_HOST_H
#define ATSIM_HOST_H

#include <semaphore.h>
#include <stdint.h>
#include <time.h>

#include "atsim.h"
#include "atsim_compat.h"

/* Shared-memory image header, byte-compatible with the reference's
 * IMAGE_METADATA (CLIcore.h:179-201, 200 bytes on x86-64). */
typedef struct {
    char name[80];
    long naxis;
    long size[3];
    long nelement;
    int atype;
    double creation_time;
    double last_access;
    struct timespec wtime;
    int shared;
    int write;
    int status;
    long cnt0;
    long cnt1;
    long NBkw;
} atsim_image_metadata;

/* IMAGE_KEYWORD (CLIcore.h:165-175, 120 bytes) */
typedef struct {
    char name[16];
    char type; /* N unused, L long, D double, S string */
    union {
        long numl;
        double numf;
        char valstr[16];
    } value;
    char comment[80];
} atsim_keyword_t;

#define ATSIM_NSEM 10

typedef struct {
    int used;
    char name[80];
    atsim_image_metadata *md; /* points into the shm map or to md_local */
    atsim_image_metadata md_local;
    void *array;
    atsim_keyword_t *kw;
    /* shared memory */
    int shared;
    int shmfd;
    size_t memsize;
    void *map;
    sem_t *sem[ATSIM_NSEM];
    int nsem;
} atsim_image;

atsim_image *atsim_reg_get(long ID);
size_t atsim_type_size(int atype);
/* post every semaphore of a shared image (value capped) */
void atsim_image_post(long ID);

/* errors: print, then exit (reference behaviour) or return -1 */
int atsim_compat_fail(int exit_code, const char *fmt, ...);
/* lazily created GPU context of the compat layer */
atsim_ctx *atsim_compat_ctx(void);
const char *atsim_compat_shm_dir(void);

/* configuration (AtmosphericTurbulence.c:1233-1427) */
typedef struct {
    float lambda;   /* TURBULENCE_REF_WAVEL [m] */
    float seeing;   /* TURBULENCE_SEEING [arcsec] */
    char prof_file[200];
    float zangle;
    float source_x, source_y;
    int wfoutput;
    char wf_prefix[200];
    int shm_output;
    int make_swavefront;
    int swf_write2disk;
    char swf_prefix[200];
    int shm_soutput;
    char shm_sprefix[100];
    int shm_soutputm;
    float pupil_scale;
    long wfsize;
    int realtime;
    double realtimefactor;
    float time_step;
    float time_span;
    long nb_tspan;
    long simtdelay;
    int waitforsem;
    char waitsemimname[100];
    int skip_existing;
    long raw_size;
    long master_size;
    int amplitude;
    int fresnel;
    float fresnel_bin;
    double site_alt;     /* SITE_ALT (AtmosphereModel.c:3907) */
    int bicubic;         /* ATSIM_SAMPLER: 1 bicubic (default), 0 bilinear */
    unsigned long long seed; /* ATSIM_NOISE_SEED */
} atsim_conf;

void atsim_conf_defaults(atsim_conf *c);
int atsim_conf_read(const char *fname, atsim_conf *c);

/* turbulence profile: 8 columns alt cn2 speed dir L0 l0 sigmaW Lwind
 * (AtmosphericTurbulence.c:1951-2004); layers below site_alt - 0.1 dropped */
typedef struct {
    long n;
    double *alt, *cn2, *speed, *dir, *outer, *inner, *sigmaw, *lwind;
} atsim_profile;
int atsim_profile_read(const char *fname, double site_alt, atsim_profile *p);
void atsim_profile_free(atsim_profile *p);

/* standard-air refractivity N(h, lambda) (see wfseries.c) */
double atsim_air_refractivity(double h, double lambda);

#endif
