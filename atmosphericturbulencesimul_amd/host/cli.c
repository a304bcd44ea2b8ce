/* cli.c -- the commands of the reference's command line for this path.
 *
 *   mkwfs      <lambda [um]> <precision 0|1>        AtmosphericTurbulence.c:183-189, 337-344
 *   mkmastert  <out1> <out2> <size> <outer [px]> <inner [px]>
 *                                                    AtmosphericTurbulence.c:201-210, 355-362
 *   fresnelpw  <in> <out> <pupil scale> <z> <lambda> WFpropagate.c:26-36, 47-53
 *   dofft      <in> <out>                            fft.c:122-131, 205-212
 *   permut     <image>                               fft.c:88-94, 235-242
 *   loadfits   <file> <name>                         COREMOD_iofits.c:155
 *   saveflfits / savedbfits / savefits <name> <file> COREMOD_iofits.c:164-191
 *   rm <name>, listim, readconf <file>, exit / quit
 *
 * Argument checks follow CLI_checkarg (CLIcore.c:1630-1853): a float argument
 * accepts integers; an integer argument rounds a float (long)(x + 0.5); a
 * string argument refuses numbers and existing image names; an image
 * argument refuses numbers (a name that is no image passes, as the
 * reference's case 4 falls through into case 5) and the command then reports
 * the missing image.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "atsim_host.h"

#define MAXTOK 16

typedef struct {
    int type; /* 1 float, 2 integer, 3 string, 4 image */
    double numf;
    long numl;
    char str[256];
} token;

static void classify(const char *s, token *t) {
    snprintf(t->str, sizeof t->str, "%s", s);
    char *end = NULL;
    long l = strtol(s, &end, 10);
    if (*s && end && *end == 0) {
        t->type = 2;
        t->numl = l;
        t->numf = (double)l;
        return;
    }
    double f = strtod(s, &end);
    if (*s && end && *end == 0) {
        t->type = 1;
        t->numf = f;
        return;
    }
    t->type = image_ID(s) != -1 ? 4 : 3;
}

/* 0 if token k has the wanted type (converting numbers as the reference does) */
static int check(token *tk, int k, int ntok, int want) {
    if (k >= ntok) {
        printf("missing argument %d\n", k);
        return 1;
    }
    token *t = &tk[k];
    switch (want) {
        case 1:
            if (t->type == 1 || t->type == 2) {
                t->type = 1;
                return 0;
            }
            printf("arg %d (\"%s\") should be a floating point number\n", k, t->str);
            return 1;
        case 2:
            if (t->type == 1) {
                printf("converting floating point arg %d to integer\n", k);
                t->numl = (long)(t->numf + 0.5);
                t->type = 2;
            }
            if (t->type == 2) return 0;
            printf("arg %d (\"%s\") should be integer\n", k, t->str);
            return 1;
        case 3:
            if (t->type == 3) return 0;
            printf("arg %d (\"%s\") should be a string that is not an image\n", k, t->str);
            return 1;
        case 4:
            if (t->type == 3 || t->type == 4) return 0;
            printf("arg %d (\"%s\") should be an image\n", k, t->str);
            return 1;
    }
    return 1;
}

static void listim(void) {
    printf("ID    NAME                 TYPE  SIZE\n");
    for (long ID = 0; ID < 5000; ID++) {
        atsim_image *im = atsim_reg_get(ID);
        if (!im) continue;
        printf("%4ld  %-20s %4d  %ld x %ld x %ld%s\n", ID, im->name, im->md->atype, im->md->size[0],
               im->md->naxis > 1 ? im->md->size[1] : 1, im->md->naxis > 2 ? im->md->size[2] : 1,
               im->shared ? "  (shm)" : "");
    }
}

int atsim_cli_execute(const char *line) {
    char buf[2048];
    snprintf(buf, sizeof buf, "%s", line ? line : "");
    char *hash = strchr(buf, '#');
    if (hash) *hash = 0;
    token tk[MAXTOK];
    int n = 0;
    for (char *save = NULL, *p = strtok_r(buf, " \t\r\n", &save); p && n < MAXTOK; p = strtok_r(NULL, " \t\r\n", &save))
        classify(p, &tk[n++]);
    if (n == 0) return 0;
    const char *cmd = tk[0].str;
    if (!strcmp(cmd, "exit") || !strcmp(cmd, "quit")) return 1;
    if (!strcmp(cmd, "mkwfs")) {
        if (check(tk, 1, n, 1) + check(tk, 2, n, 2)) return -1;
        return make_AtmosphericTurbulence_wavefront_series((float)tk[1].numf, tk[2].numl) ? -1 : 0;
    }
    if (!strcmp(cmd, "mkmastert")) {
        if (check(tk, 1, n, 3) + check(tk, 2, n, 3) + check(tk, 3, n, 2) + check(tk, 4, n, 1) + check(tk, 5, n, 1))
            return -1;
        return make_master_turbulence_screen(tk[1].str, tk[2].str, tk[3].numl, (float)tk[4].numf, (float)tk[5].numf, 0)
                   ? -1 : 0;
    }
    if (!strcmp(cmd, "fresnelpw")) {
        if (check(tk, 1, n, 4) + check(tk, 2, n, 3) + check(tk, 3, n, 1) + check(tk, 4, n, 1) + check(tk, 5, n, 1))
            return -1;
        return Fresnel_propagate_wavefront(tk[1].str, tk[2].str, tk[3].numf, tk[4].numf, tk[5].numf) ? -1 : 0;
    }
    if (!strcmp(cmd, "dofft")) {
        if (check(tk, 1, n, 4) + check(tk, 2, n, 3)) return -1;
        return do2dfft(tk[1].str, tk[2].str) < 0 ? -1 : 0;
    }
    if (!strcmp(cmd, "permut")) {
        if (check(tk, 1, n, 4)) return -1;
        return permut(tk[1].str) ? -1 : 0;
    }
    if (!strcmp(cmd, "loadfits")) {
        if (n < 3) {
            printf("loadfits <file> <name>\n");
            return -1;
        }
        return load_fits(tk[1].str, tk[2].str, 1) == -1 ? -1 : 0;
    }
    if (!strcmp(cmd, "saveflfits") || !strcmp(cmd, "savefits") || !strcmp(cmd, "savedbfits")) {
        if (check(tk, 1, n, 4) || n < 3) return -1;
        return (!strcmp(cmd, "savedbfits") ? save_db_fits(tk[1].str, tk[2].str) : save_fl_fits(tk[1].str, tk[2].str)) ? -1 : 0;
    }
    if (!strcmp(cmd, "rm")) {
        for (int k = 1; k < n; k++) delete_image_ID(tk[k].str);
        return 0;
    }
    if (!strcmp(cmd, "listim")) {
        listim();
        return 0;
    }
    if (!strcmp(cmd, "readconf")) {
        if (n < 2) return -1;
        return AtmosphericTurbulence_change_configuration_file(tk[1].str);
    }
    printf("unknown command \"%s\"\n", cmd);
    return -1;
}
