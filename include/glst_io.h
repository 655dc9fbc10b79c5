/*
 * glst_io.h — CHARMM PSF/COR readers and the coordinate recentring of the
 * reference driver (utils/io.cpp:16-117, utils/utils.cpp:35-58), host only.
 */
#ifndef GLST_IO_H
#define GLST_IO_H

#include <stddef.h>
#include <stdint.h>

#include "glst_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* get_natom_psf (utils/io.cpp:35-52): the count in columns 0-10 of the first
 * line whose columns 11-16 read "!NATOM"; 0 when there is none. */
glst_status_t glst_psf_natom(const char *path, uint64_t *natom);

/* read_charmm_psf (utils/io.cpp:87-117): charge of each atom, columns
 * 50-63 of the natom lines after the !NATOM line.  Fails (GLST_ERR_IO) when
 * the file's count differs from natom. */
glst_status_t glst_read_psf_charges(const char *path, uint64_t natom,
                                    double *qc);

/* read_charmm_cor (utils/io.cpp:54-85): skips '*' title lines, checks the
 * count in columns 0-9, then x/y/z from columns 40-59, 60-79, 80-99. */
glst_status_t glst_read_cor(const char *path, uint64_t natom, double *rx,
                            double *ry, double *rz);

/* recenter (utils/utils.cpp:35-58): shift so each coordinate's minimum is 0. */
glst_status_t glst_recenter(uint64_t natom, double *rx, double *ry,
                            double *rz);

#ifdef __cplusplus
}
#endif

#endif /* GLST_IO_H */
