/*
 * atsim_compat.h -- the reference's C entry points of the wavefront path,
 * same names, argument meaning and return types, implemented on the MI355X
 * library (libatsim.so).  Operands are images named in this library's
 * registry (the reference's data.image[] contract, CLIcore.h:210-243); the
 * registry functions below keep the reference's names.
 *
 * Errors: the reference prints a message and calls exit(0) / exit(1).  These
 * functions do the same by default; atsim_compat_exit_on_error(0) makes them
 * print and return -1 instead (used by the tests).
 */
#ifndef ATSIM_COMPAT_H
#define ATSIM_COMPAT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* image type codes (CLIcore.h:143-150) */
#define ATSIM_CHAR 1
#define ATSIM_INT 2
#define ATSIM_FLOAT 3
#define ATSIM_DOUBLE 4
#define ATSIM_COMPLEX_FLOAT 5
#define ATSIM_COMPLEX_DOUBLE 6
#define ATSIM_USHORT 7
#define ATSIM_LONG 8

/* ------------------------------------------------------------ registry */
/* image_ID (COREMOD_memory.c:1172-1203): index of a named image or -1 */
long image_ID(const char *name);
/* create_image_ID (COREMOD_memory.c:1519-1907): shared = 1 backs the image
 * by /tmp/<name>.im.shm in the reference's IMAGE_METADATA layout with NBkw
 * keywords and 10 POSIX semaphores <name>_sem00..09 */
long create_image_ID(const char *name, long naxis, long *size, int atype, int shared, int NBkw);
long create_2Dimage_ID(const char *name, long xsize, long ysize);
long create_2Dimage_ID_double(const char *name, long xsize, long ysize);
long create_2DCimage_ID(const char *name, long xsize, long ysize);
long create_2DCimage_ID_double(const char *name, long xsize, long ysize);
long create_3Dimage_ID(const char *name, long xsize, long ysize, long zsize);
long create_3Dimage_ID_double(const char *name, long xsize, long ysize, long zsize);
long delete_image_ID(const char *name);
/* chname_image_ID (COREMOD_memory.c:3188-3207) */
long chname_image_ID(const char *old_name, const char *new_name);
/* accessors of this registry */
void *atsim_image_data(long ID);
int atsim_image_type(long ID);
long atsim_image_naxis(long ID);
long atsim_image_size(long ID, int axis);
/* keyword slot k of a shared image (NULL if none): name[16], type, value, comment */
void *atsim_image_keyword(long ID, long k);
long atsim_image_cnt0(long ID);

/* ------------------------------------------------------------ FITS */
/* save_fl_fits / save_db_fits (COREMOD_iofits.c:929-1070, 781-910): a
 * leading '!' overwrites; BITPIX -32 / -64 */
int save_fl_fits(const char *ID_name, const char *file_name);
int save_db_fits(const char *ID_name, const char *file_name);
/* load_fits (COREMOD_iofits.c:403-780): BITPIX -32 -> FLOAT, -64 -> DOUBLE,
 * 8/16/32 -> FLOAT; returns the ID or -1 */
long load_fits(const char *file_name, const char *ID_name, int errcode);
int file_exists(const char *file_name);

/* ------------------------------------------------------------ config */
/* read_config_parameter (COREMOD_tools.c:629-690): last occurrence wins,
 * content "-" and return 0 when the key is absent */
int read_config_parameter(const char *config_file, const char *keyword, char *content);
int read_config_parameter_exists(const char *config_file, const char *keyword);

/* ------------------------------------------------------------ fft (fft.h:19-31) */
long do2dfft(const char *in_name, const char *out_name);
long do2dffti(const char *in_name, const char *out_name);
int permut(const char *ID_name);

/* ------------------------------------------------------------ WFpropagate (WFpropagate.h:5-16) */
int init_WFpropagate(void);
int Fresnel_propagate_wavefront(const char *in, const char *out, double PUPIL_SCALE, double z, double lambda);
int Init_Fresnel_propagate_wavefront(const char *Cim, long size, double PUPIL_SCALE, double z, double lambda,
                                     double FPMASKRAD, int Precision);
int Fresnel_propagate_wavefront1(const char *in, const char *out, const char *Cin);
long Fresnel_propagate_cube(const char *IDcin_name, const char *IDout_name_amp, const char *IDout_name_pha,
                            double PUPIL_SCALE, double zstart, double zend, long NBzpts, double lambda);

/* ------------------------------------------------------------ AtmosphericTurbulence (AtmosphericTurbulence.h:5-44) */
int init_AtmosphericTurbulence(void);
int AtmosphericTurbulence_change_configuration_file(const char *fname);
int make_master_turbulence_screen(const char *ID_name1, const char *ID_name2, long size, float outerscale,
                                  float innerscale, long WFprecision);
int make_AtmosphericTurbulence_wavefront_series(float slambdaum, long WFprecision);

/* ------------------------------------------------------------ library controls (no reference counterpart) */
/* 1 (default): fatal errors print and exit like the reference; 0: print and return -1 */
void atsim_compat_exit_on_error(int enable);
/* Philox seed of the noise behind make_master_turbulence_screen (the reference
 * seeds from the clock, CLIcore.c:696-698); each call uses the next stream */
void atsim_compat_set_seed(unsigned long long seed);
/* the RLIM CLI variable of make_master_turbulence_screen (AtmosphericTurbulence.c:671-677); <= 0 clears */
void atsim_compat_set_rlim(double rlim_px);
/* directory of the shared-memory image files (default /tmp, CLIcore.h:45) */
void atsim_compat_set_shm_dir(const char *dir);
/* run one CLI command line (mkwfs, mkmastert, fresnelpw, ...); returns 0 on success */
int atsim_cli_execute(const char *line);

#ifdef __cplusplus
}
#endif
#endif
