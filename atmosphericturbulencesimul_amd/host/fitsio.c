/* fitsio.c -- FITS primary-image writer/reader without cfitsio.
 *
 * Writes what save_fl_fits / save_db_fits produce for the wavefront outputs
 * (COREMOD_iofits.c:781-1070): a primary HDU with BITPIX -32 or -64, NAXIS
 * 2 or 3, NAXISn = image sizes (x first), big-endian data, header and data
 * padded to 2880-byte records.  A leading '!' in the file name means
 * overwrite, as with cfitsio.  The reader accepts BITPIX 8/16/32/-32/-64
 * and applies BSCALE / BZERO as cfitsio does.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "atsim_host.h"

int file_exists(const char *file_name) {
    struct stat st;
    return (file_name && stat(file_name, &st) == 0) ? 1 : 0;
}

static void card(char *hdr, int *n, const char *fmt, ...) {
    char buf[128];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    char *c = hdr + 80 * (*n);
    memset(c, ' ', 80);
    memcpy(c, buf, strlen(buf) < 80 ? strlen(buf) : 80);
    (*n)++;
}

static void swap_bytes(unsigned char *p, size_t n, size_t w) {
    for (size_t i = 0; i < n; i++) {
        unsigned char *q = p + i * w;
        for (size_t a = 0, b = w - 1; a < b; a++, b--) {
            unsigned char t = q[a];
            q[a] = q[b];
            q[b] = t;
        }
    }
}

static int save_fits(const char *ID_name, const char *file_name, int bitpix) {
    long ID = image_ID(ID_name);
    atsim_image *im = atsim_reg_get(ID);
    if (!im) return atsim_compat_fail(0, "save_fits: image \"%s\" does not exist", ID_name);
    const char *fname = file_name[0] == '!' ? file_name + 1 : file_name;
    if (file_name[0] != '!' && file_exists(fname))
        return atsim_compat_fail(0, "save_fits: file \"%s\" exists (prefix '!' to overwrite)", fname);
    atsim_image_metadata *md = im->md;
    const long n = md->nelement;
    const size_t w = bitpix == -32 ? 4 : 8;
    unsigned char *buf = (unsigned char *)malloc((size_t)n * w + 1);
    if (!buf) return atsim_compat_fail(0, "save_fits: out of memory");
    for (long i = 0; i < n; i++) {
        double v;
        switch (md->atype) {
            case ATSIM_FLOAT: v = ((float *)im->array)[i]; break;
            case ATSIM_DOUBLE: v = ((double *)im->array)[i]; break;
            case ATSIM_INT: v = ((int *)im->array)[i]; break;
            case ATSIM_LONG: v = (double)((long *)im->array)[i]; break;
            case ATSIM_USHORT: v = ((unsigned short *)im->array)[i]; break;
            case ATSIM_CHAR: v = ((char *)im->array)[i]; break;
            default: free(buf); return atsim_compat_fail(0, "save_fits: complex image \"%s\"", ID_name);
        }
        if (w == 4) {
            float f = (float)v;
            memcpy(buf + i * 4, &f, 4);
        } else {
            memcpy(buf + i * 8, &v, 8);
        }
    }
    swap_bytes(buf, (size_t)n, w);
    char hdr[2880 * 2];
    memset(hdr, ' ', sizeof hdr);
    int nc = 0;
    card(hdr, &nc, "SIMPLE  =                    T / file does conform to FITS standard");
    card(hdr, &nc, "BITPIX  = %20d / number of bits per data pixel", bitpix);
    card(hdr, &nc, "NAXIS   = %20ld / number of data axes", md->naxis);
    for (long a = 0; a < md->naxis; a++) card(hdr, &nc, "NAXIS%ld  = %20ld / length of data axis %ld", a + 1, md->size[a], a + 1);
    card(hdr, &nc, "EXTEND  =                    T / FITS dataset may contain extensions");
    card(hdr, &nc, "END");
    const size_t hlen = ((size_t)nc * 80 + 2879) / 2880 * 2880;
    FILE *fp = fopen(fname, "wb");
    if (!fp) {
        free(buf);
        return atsim_compat_fail(0, "save_fits: cannot write \"%s\": %s", fname, strerror(errno));
    }
    size_t dlen = (size_t)n * w, pad = (2880 - dlen % 2880) % 2880;
    int ok = fwrite(hdr, 1, hlen, fp) == hlen && fwrite(buf, 1, dlen, fp) == dlen;
    static const char zeros[2880];
    if (ok && pad) ok = fwrite(zeros, 1, pad, fp) == pad;
    fclose(fp);
    free(buf);
    return ok ? 0 : atsim_compat_fail(0, "save_fits: short write to \"%s\"", fname);
}

int save_fl_fits(const char *ID_name, const char *file_name) { return save_fits(ID_name, file_name, -32); }
int save_db_fits(const char *ID_name, const char *file_name) { return save_fits(ID_name, file_name, -64); }

static long card_int(const char *c, const char *key, int *found) {
    if (strncmp(c, key, strlen(key)) != 0 || c[8] != '=') return 0;
    *found = 1;
    return strtol(c + 10, NULL, 10);
}

long load_fits(const char *file_name, const char *ID_name, int errcode) {
    FILE *fp = fopen(file_name, "rb");
    if (!fp) {
        if (errcode) fprintf(stderr, "load_fits: cannot open \"%s\"\n", file_name);
        return -1;
    }
    long bitpix = 0, naxis = 0, size[3] = {1, 1, 1};
    double bscale = 1.0, bzero = 0.0;
    int end = 0;
    char rec[2880];
    while (!end) {
        if (fread(rec, 1, 2880, fp) != 2880) {
            fclose(fp);
            return -1;
        }
        for (int i = 0; i < 36 && !end; i++) {
            const char *c = rec + 80 * i;
            int f = 0;
            if (strncmp(c, "END", 3) == 0 && (c[3] == ' ')) end = 1;
            long v = card_int(c, "BITPIX  ", &f);
            if (f) bitpix = v;
            f = 0;
            v = card_int(c, "NAXIS   ", &f);
            if (f) naxis = v;
            if (strncmp(c, "BSCALE  = ", 10) == 0) bscale = strtod(c + 10, NULL);
            if (strncmp(c, "BZERO   = ", 10) == 0) bzero = strtod(c + 10, NULL);
            for (int a = 0; a < 3; a++) {
                char key[9];
                snprintf(key, sizeof key, "NAXIS%d  ", a + 1);
                f = 0;
                v = card_int(c, key, &f);
                if (f) size[a] = v;
            }
        }
    }
    if (naxis < 1 || naxis > 3 || (bitpix != 8 && bitpix != 16 && bitpix != 32 && bitpix != -32 && bitpix != -64)) {
        fclose(fp);
        return -1;
    }
    long n = 1;
    for (long a = 0; a < naxis; a++) n *= size[a];
    const size_t w = (size_t)(bitpix < 0 ? -bitpix : bitpix) / 8;
    unsigned char *buf = (unsigned char *)malloc((size_t)n * w + 1);
    if (!buf || fread(buf, 1, (size_t)n * w, fp) != (size_t)n * w) {
        free(buf);
        fclose(fp);
        return -1;
    }
    fclose(fp);
    swap_bytes(buf, (size_t)n, w);
    const int atype = bitpix == -64 ? ATSIM_DOUBLE : ATSIM_FLOAT;
    long ID = image_ID(ID_name);
    if (ID != -1) delete_image_ID(ID_name);
    ID = create_image_ID(ID_name, naxis, size, atype, 0, 0);
    if (ID < 0) {
        free(buf);
        return -1;
    }
    void *arr = atsim_image_data(ID);
    for (long i = 0; i < n; i++) {
        double v;
        switch (bitpix) {
            case -64: memcpy(&v, buf + i * 8, 8); break;
            case -32: { float f; memcpy(&f, buf + i * 4, 4); v = f; } break;
            case 32: { int32_t q; memcpy(&q, buf + i * 4, 4); v = q; } break;
            case 16: { int16_t q; memcpy(&q, buf + i * 2, 2); v = q; } break;
            default: v = buf[i]; break;
        }
        v = v * bscale + bzero;
        if (atype == ATSIM_DOUBLE) ((double *)arr)[i] = v;
        else ((float *)arr)[i] = (float)v;
    }
    free(buf);
    return ID;
}
