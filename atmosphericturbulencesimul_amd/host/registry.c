/* registry.c -- named image registry with optional shared-memory backing.
 *
 * Follows the reference's contract (COREMOD_memory.c:1172-1907): images are
 * looked up by name; shared images live in <shmdir>/<name>.im.shm laid out as
 * [IMAGE_METADATA | data | NBkw IMAGE_KEYWORD] (COREMOD_memory.c:1551-1624,
 * 1716-1744) with 10 named semaphores <name>_sem00..09
 * (COREMOD_memory.c:4071-4126).  The reference places DOUBLE shared data at
 * map + sizeof(IMAGE) and never sets its keyword pointer; here every type
 * uses the FLOAT layout (data right after the header, keywords after data).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "atsim_host.h"

#define ATSIM_MAX_IMAGES 5000 /* NB_MAX_IMAGE, CLIcore.c:1003 */

static atsim_image g_img[ATSIM_MAX_IMAGES];
static int g_exit_on_error = 1;
static char g_shm_dir[512] = ""; /* empty: $ATSIM_SHM_DIR, else /tmp */

void atsim_compat_exit_on_error(int enable) { g_exit_on_error = enable; }
void atsim_compat_set_shm_dir(const char *dir) { snprintf(g_shm_dir, sizeof g_shm_dir, "%s", dir ? dir : ""); }
const char *atsim_compat_shm_dir(void) {
    if (g_shm_dir[0]) return g_shm_dir;
    const char *e = getenv("ATSIM_SHM_DIR");
    return e && *e ? e : "/tmp";
}

int atsim_compat_fail(int exit_code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "ERROR: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    fflush(stderr);
    if (g_exit_on_error) exit(exit_code);
    return -1;
}

static atsim_ctx *g_ctx = NULL;
atsim_ctx *atsim_compat_ctx(void) {
    if (!g_ctx) {
        const char *e = getenv("ATSIM_DEVICE");
        int dev = e ? atoi(e) : 0;
        int rc = atsim_ctx_create(dev, &g_ctx);
        if (rc) {
            g_ctx = NULL;
            atsim_compat_fail(1, "cannot open GPU %d for the wavefront library (status %d)", dev, rc);
            return NULL;
        }
    }
    return g_ctx;
}

size_t atsim_type_size(int atype) {
    switch (atype) {
        case ATSIM_CHAR: return 1;
        case ATSIM_INT: return sizeof(int);
        case ATSIM_FLOAT: return sizeof(float);
        case ATSIM_DOUBLE: return sizeof(double);
        case ATSIM_COMPLEX_FLOAT: return 2 * sizeof(float);
        case ATSIM_COMPLEX_DOUBLE: return 2 * sizeof(double);
        case ATSIM_USHORT: return sizeof(unsigned short);
        case ATSIM_LONG: return sizeof(long);
        default: return 0;
    }
}

atsim_image *atsim_reg_get(long ID) {
    if (ID < 0 || ID >= ATSIM_MAX_IMAGES || !g_img[ID].used) return NULL;
    return &g_img[ID];
}

long image_ID(const char *name) {
    if (!name) return -1;
    for (long i = 0; i < ATSIM_MAX_IMAGES; i++)
        if (g_img[i].used && strcmp(g_img[i].name, name) == 0) return i;
    return -1;
}

static double now_s(struct timespec *ts) {
    clock_gettime(CLOCK_REALTIME, ts);
    return 1.0 * ts->tv_sec + 1e-9 * ts->tv_nsec;
}

static int open_sems(atsim_image *im) {
    char sname[256];
    for (int s = 0; s < ATSIM_NSEM; s++) {
        snprintf(sname, sizeof sname, "/%s_sem%02d", im->name, s);
        sem_t *p = sem_open(sname, O_CREAT, 0644, 0);
        if (p == SEM_FAILED) {
            im->sem[s] = NULL;
            continue;
        }
        im->sem[s] = p;
    }
    im->nsem = ATSIM_NSEM;
    return 0;
}

long create_image_ID(const char *name, long naxis, long *size, int atype, int shared, int NBkw) {
    if (!name || !size || naxis < 1 || naxis > 3) {
        atsim_compat_fail(0, "create_image_ID: bad arguments for \"%s\"", name ? name : "(null)");
        return -1;
    }
    size_t tsz = atsim_type_size(atype);
    if (!tsz) {
        atsim_compat_fail(0, "create_image_ID: unknown type %d", atype);
        return -1;
    }
    long ID = image_ID(name);
    if (ID != -1) {
        atsim_image *im = &g_img[ID];
        if (im->md->atype != atype || im->md->naxis != naxis) {
            atsim_compat_fail(0, "Pre-existing image \"%s\" has wrong type or axes", name);
            return -1;
        }
        return ID;
    }
    for (ID = 0; ID < ATSIM_MAX_IMAGES && g_img[ID].used; ID++) {
    }
    if (ID == ATSIM_MAX_IMAGES) {
        atsim_compat_fail(0, "image registry full");
        return -1;
    }
    atsim_image *im = &g_img[ID];
    memset(im, 0, sizeof *im);
    snprintf(im->name, sizeof im->name, "%s", name);
    long nelement = 1;
    for (long i = 0; i < naxis; i++) nelement *= size[i];
    if (shared) {
        size_t memsize = sizeof(atsim_image_metadata) + (size_t)nelement * tsz + (size_t)NBkw * sizeof(atsim_keyword_t);
        char fname[1024];
        snprintf(fname, sizeof fname, "%s/%s.im.shm", atsim_compat_shm_dir(), name);
        int fd = open(fname, O_RDWR | O_CREAT | O_TRUNC, (mode_t)0600);
        if (fd == -1) {
            atsim_compat_fail(0, "cannot open %s: %s", fname, strerror(errno));
            return -1;
        }
        if (ftruncate(fd, (off_t)memsize) != 0) {
            close(fd);
            atsim_compat_fail(0, "cannot size %s", fname);
            return -1;
        }
        void *map = mmap(0, memsize, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
        if (map == MAP_FAILED) {
            close(fd);
            atsim_compat_fail(0, "cannot map %s", fname);
            return -1;
        }
        memset(map, 0, memsize);
        im->map = map;
        im->memsize = memsize;
        im->shmfd = fd;
        im->shared = 1;
        im->md = (atsim_image_metadata *)map;
        im->array = (char *)map + sizeof(atsim_image_metadata);
        im->kw = NBkw > 0 ? (atsim_keyword_t *)((char *)im->array + (size_t)nelement * tsz) : NULL;
        open_sems(im);
    } else {
        im->md = &im->md_local;
        im->array = calloc((size_t)nelement, tsz);
        if (!im->array) {
            atsim_compat_fail(0, "memory allocation failed for \"%s\" (%ld elements)", name, nelement);
            return -1;
        }
        im->kw = NBkw > 0 ? (atsim_keyword_t *)calloc((size_t)NBkw, sizeof(atsim_keyword_t)) : NULL;
    }
    atsim_image_metadata *md = im->md;
    snprintf(md->name, sizeof md->name, "%s", name);
    md->naxis = naxis;
    for (long i = 0; i < 3; i++) md->size[i] = i < naxis ? size[i] : 0;
    md->nelement = nelement;
    md->atype = atype;
    md->last_access = now_s(&md->wtime);
    md->creation_time = md->last_access;
    md->shared = shared ? 1 : 0;
    md->write = 0;
    md->status = 0;
    md->cnt0 = 0;
    md->cnt1 = 0;
    md->NBkw = NBkw;
    for (long k = 0; k < NBkw && im->kw; k++) im->kw[k].type = 'N';
    im->used = 1;
    return ID;
}

long create_2Dimage_ID(const char *name, long xsize, long ysize) {
    long s[2] = {xsize, ysize};
    return create_image_ID(name, 2, s, ATSIM_FLOAT, 0, 0);
}
long create_2Dimage_ID_double(const char *name, long xsize, long ysize) {
    long s[2] = {xsize, ysize};
    return create_image_ID(name, 2, s, ATSIM_DOUBLE, 0, 0);
}
long create_2DCimage_ID(const char *name, long xsize, long ysize) {
    long s[2] = {xsize, ysize};
    return create_image_ID(name, 2, s, ATSIM_COMPLEX_FLOAT, 0, 0);
}
long create_2DCimage_ID_double(const char *name, long xsize, long ysize) {
    long s[2] = {xsize, ysize};
    return create_image_ID(name, 2, s, ATSIM_COMPLEX_DOUBLE, 0, 0);
}
long create_3Dimage_ID(const char *name, long xsize, long ysize, long zsize) {
    long s[3] = {xsize, ysize, zsize};
    return create_image_ID(name, 3, s, ATSIM_FLOAT, 0, 0);
}
long create_3Dimage_ID_double(const char *name, long xsize, long ysize, long zsize) {
    long s[3] = {xsize, ysize, zsize};
    return create_image_ID(name, 3, s, ATSIM_DOUBLE, 0, 0);
}

long delete_image_ID(const char *name) {
    long ID = image_ID(name);
    if (ID == -1) return -1;
    atsim_image *im = &g_img[ID];
    if (im->shared) {
        for (int s = 0; s < im->nsem; s++)
            if (im->sem[s]) sem_close(im->sem[s]);
        munmap(im->map, im->memsize);
        close(im->shmfd);
    } else {
        free(im->array);
        free(im->kw);
    }
    memset(im, 0, sizeof *im);
    return ID;
}

long chname_image_ID(const char *old_name, const char *new_name) {
    long ID = image_ID(old_name);
    if (ID == -1 || image_ID(new_name) != -1) return -1;
    snprintf(g_img[ID].name, sizeof g_img[ID].name, "%s", new_name);
    if (!g_img[ID].shared) snprintf(g_img[ID].md->name, sizeof g_img[ID].md->name, "%s", new_name);
    return ID;
}

void *atsim_image_data(long ID) {
    atsim_image *im = atsim_reg_get(ID);
    return im ? im->array : NULL;
}
int atsim_image_type(long ID) {
    atsim_image *im = atsim_reg_get(ID);
    return im ? im->md->atype : -1;
}
long atsim_image_naxis(long ID) {
    atsim_image *im = atsim_reg_get(ID);
    return im ? im->md->naxis : -1;
}
long atsim_image_size(long ID, int axis) {
    atsim_image *im = atsim_reg_get(ID);
    if (!im || axis < 0 || axis > 2) return -1;
    return im->md->size[axis];
}
void *atsim_image_keyword(long ID, long k) {
    atsim_image *im = atsim_reg_get(ID);
    if (!im || !im->kw || k < 0 || k >= im->md->NBkw) return NULL;
    return &im->kw[k];
}
long atsim_image_cnt0(long ID) {
    atsim_image *im = atsim_reg_get(ID);
    return im ? im->md->cnt0 : -1;
}

void atsim_image_post(long ID) {
    atsim_image *im = atsim_reg_get(ID);
    if (!im || !im->shared) return;
    for (int s = 0; s < im->nsem; s++) {
        if (!im->sem[s]) continue;
        int v = 0;
        sem_getvalue(im->sem[s], &v);
        if (v < 10) sem_post(im->sem[s]);
    }
}
