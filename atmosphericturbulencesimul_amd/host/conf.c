/* conf.c -- WFsim.conf / turbul.prof parsing.
 *
 * read_config_parameter follows COREMOD_tools.c:629-690: every line is read
 * as "%s %s"; the last line whose first token equals the key wins; a missing
 * key yields the content "-" (so atof gives 0).  atsim_conf_read reads the
 * keys of AtmosphericTurbulence_ReadConf (AtmosphericTurbulence.c:1233-1427)
 * into the same types (float for the physical scalars) plus SITE_ALT
 * (AtmosphereModel.c:3907) and two keys of this library that the reference
 * ignores: ATSIM_SAMPLER (bicubic / bilinear) and ATSIM_NOISE_SEED.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "atsim_host.h"

int read_config_parameter_exists(const char *config_file, const char *keyword) {
    FILE *fp = fopen(config_file, "r");
    if (!fp) {
        atsim_compat_fail(0, "cannot open file \"%s\"", config_file);
        return 0;
    }
    char line[1000], keyw[200];
    int read = 0;
    while (!read && fgets(line, sizeof line, fp)) {
        keyw[0] = 0;
        if (sscanf(line, "%199s", keyw) == 1 && strcmp(keyw, keyword) == 0) read = 1;
    }
    fclose(fp);
    return read;
}

int read_config_parameter(const char *config_file, const char *keyword, char *content) {
    FILE *fp = fopen(config_file, "r");
    if (!fp) {
        atsim_compat_fail(0, "cannot open file \"%s\"", config_file);
        strcpy(content, "-");
        return 0;
    }
    char line[1000], keyw[200], cont[200];
    int read = 0;
    strcpy(content, "---");
    while (fgets(line, sizeof line, fp)) {
        keyw[0] = cont[0] = 0;
        int k = sscanf(line, "%199s %199s", keyw, cont);
        if (k >= 1 && strcmp(keyw, keyword) == 0) {
            strcpy(content, k == 2 ? cont : "");
            read = 1;
        }
    }
    fclose(fp);
    if (!read) strcpy(content, "-");
    return read;
}

void atsim_conf_defaults(atsim_conf *c) {
    memset(c, 0, sizeof *c);
    c->wfoutput = 1;          /* AtmosphericTurbulence.c:113 */
    c->realtimefactor = 1.0;  /* AtmosphericTurbulence.c:136 */
    c->bicubic = 1;           /* BICUBIC = 1, AtmosphericTurbulence.c:1605 */
    c->seed = 0xA75ULL;
}

static double kd(const char *f, const char *k) {
    char s[256];
    read_config_parameter(f, k, s);
    return atof(s);
}
static float kf(const char *f, const char *k) {
    char s[256];
    read_config_parameter(f, k, s);
    return (float)atof(s);
}
static long kl(const char *f, const char *k) {
    char s[256];
    read_config_parameter(f, k, s);
    return atol(s);
}
static int ki(const char *f, const char *k) {
    char s[256];
    read_config_parameter(f, k, s);
    return atoi(s);
}
static void ks(const char *f, const char *k, char *dst, size_t n) {
    char s[256];
    read_config_parameter(f, k, s);
    snprintf(dst, n, "%s", s);
}

int atsim_conf_read(const char *f, atsim_conf *c) {
    FILE *fp = fopen(f, "r");
    if (!fp) return atsim_compat_fail(0, "cannot open configuration file \"%s\"", f);
    fclose(fp);
    c->lambda = (float)(kd(f, "TURBULENCE_REF_WAVEL") * 0.000001);
    c->seeing = kf(f, "TURBULENCE_SEEING");
    ks(f, "TURBULENCE_PROF_FILE", c->prof_file, sizeof c->prof_file);
    c->zangle = kf(f, "ZENITH_ANGLE");
    c->source_x = kf(f, "SOURCE_XPOS");
    c->source_y = kf(f, "SOURCE_YPOS");
    if (read_config_parameter_exists(f, "WFOUTPUT")) c->wfoutput = ki(f, "WFOUTPUT");
    ks(f, "WF_FILE_PREFIX", c->wf_prefix, sizeof c->wf_prefix);
    if (read_config_parameter_exists(f, "SHM_OUTPUT")) c->shm_output = ki(f, "SHM_OUTPUT");
    c->make_swavefront = ki(f, "MAKE_SWAVEFRONT");
    c->swf_write2disk = ki(f, "SWF_WRITE2DISK");
    ks(f, "SWF_FILE_PREFIX", c->swf_prefix, sizeof c->swf_prefix);
    if (read_config_parameter_exists(f, "SHM_SOUTPUT")) c->shm_soutput = ki(f, "SHM_SOUTPUT");
    ks(f, "SHM_SPREFIX", c->shm_sprefix, sizeof c->shm_sprefix);
    c->shm_soutputm = ki(f, "SHM_SOUTPUTM");
    c->pupil_scale = kf(f, "PUPIL_SCALE");
    c->wfsize = kl(f, "WFsize");
    c->realtime = ki(f, "REALTIME");
    c->realtimefactor = kf(f, "REALTIMEFACTOR");
    c->time_step = kf(f, "WFTIME_STEP");
    c->time_span = kf(f, "TIME_SPAN");
    c->nb_tspan = kl(f, "NB_TSPAN");
    c->simtdelay = kl(f, "SIMTDELAY");
    c->waitforsem = ki(f, "WAITFORSEM");
    ks(f, "WAITSEMIMNAME", c->waitsemimname, sizeof c->waitsemimname);
    c->skip_existing = read_config_parameter_exists(f, "SKIP_EXISTING") ? 1 : 0;
    c->raw_size = kl(f, "WF_RAW_SIZE");
    c->master_size = kl(f, "MASTER_SIZE");
    c->amplitude = ki(f, "WAVEFRONT_AMPLITUDE");
    c->fresnel = c->amplitude == 1 ? ki(f, "FRESNEL_PROPAGATION") : 0;
    c->fresnel_bin = kf(f, "FRESNEL_PROPAGATION_BIN");
    c->site_alt = read_config_parameter_exists(f, "SITE_ALT") ? kd(f, "SITE_ALT") : 0.0;
    if (read_config_parameter_exists(f, "ATSIM_SAMPLER")) {
        char s[256];
        read_config_parameter(f, "ATSIM_SAMPLER", s);
        c->bicubic = strcmp(s, "bilinear") == 0 || strcmp(s, "0") == 0 ? 0 : 1;
    }
    if (read_config_parameter_exists(f, "ATSIM_NOISE_SEED")) {
        char s[256];
        read_config_parameter(f, "ATSIM_NOISE_SEED", s);
        c->seed = strtoull(s, NULL, 0);
    }
    return 0;
}

void atsim_profile_free(atsim_profile *p) {
    free(p->alt); free(p->cn2); free(p->speed); free(p->dir);
    free(p->outer); free(p->inner); free(p->sigmaw); free(p->lwind);
    memset(p, 0, sizeof *p);
}

int atsim_profile_read(const char *fname, double site_alt, atsim_profile *p) {
    memset(p, 0, sizeof *p);
    FILE *fp = fopen(fname, "r");
    if (!fp) return atsim_compat_fail(1, "Cannot open turbulence profile file \"%s\"", fname);
    char line[2000], word[200];
    long n = 0;
    while (fgets(line, sizeof line, fp))
        if (sscanf(line, "%199s", word) == 1 && isdigit((unsigned char)word[0])) n++;
    rewind(fp);
    double **cols[8] = {&p->alt, &p->cn2, &p->speed, &p->dir, &p->outer, &p->inner, &p->sigmaw, &p->lwind};
    for (int k = 0; k < 8; k++) *cols[k] = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    long l = 0;
    while (fgets(line, sizeof line, fp)) {
        if (sscanf(line, "%199s", word) != 1 || !isdigit((unsigned char)word[0])) continue;
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        sscanf(line, "%lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
        if (v[0] > site_alt - 0.1) {  /* AtmosphericTurbulence.c:1988 */
            for (int k = 0; k < 8; k++) (*cols[k])[l] = v[k];
            l++;
        }
    }
    fclose(fp);
    p->n = l;
    return 0;
}
