/*
 * attract_io.h — ATTRACT grid files (the `.grid` binaries of ATTRACT's
 * make-grid tool) as the scoring path's descriptors.
 *
 * att_gridfile_read replaces the reference's readGridFromGridFile
 * (fileIO/readFile.cpp:260-422 over attract::Grid::read_std,
 * fileIO/grid_orig.cpp:85-301): same format (64-bit, standard, not torque),
 * same checks, same conversion — inner grid at the file's spacing and origin,
 * outer grid at twice the spacing and origin - gridextension * spacing, one
 * grid per alphabet type in ascending order with the electrostatic grid moved
 * to the front, empty voxels as zeros, neighbour lists with 0-based starts.
 * att_gridfile_write is the inverse (grid_orig.cpp:303+ write_std): every
 * voxel carries potentials, energy/gradient records in voxel order.
 * Shared-memory grid files (shm key != -1) are refused.
 *
 * Library: libattract_service.so (host code only).
 */
#ifndef ATTRACT_IO_H_
#define ATTRACT_IO_H_

#include <stddef.h>
#include <stdint.h>

#include "attract_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct att_gridfile att_gridfile;

/* The header of an ATTRACT grid file (attract::Grid, grid_orig.h:70-100). */
typedef struct {
    double gridspacing;          /* inner spacing; the outer grid has twice it */
    int gridextension;           /* the outer grid starts gridextension * gridspacing below the inner */
    double plateaudis, neighbourdis;
    int natoms;
    double pivot[3];
    unsigned char alphabet[99];  /* alphabet[t-1] != 0: type t has a grid */
    unsigned alphabetsize;       /* filled by read; computed by write */
} att_gridfile_info;

att_status att_gridfile_read(const char* path, att_gridfile** out);
void att_gridfile_free(att_gridfile* g);
/* Message of the calling thread's last failed att_gridfile_* call. */
const char* att_gridfile_last_error(void);

/* Header and descriptors; the arrays stay owned by g (valid until
 * att_gridfile_free).  Any output may be NULL. */
att_status att_gridfile_get(const att_gridfile* g, att_gridfile_info* info, att_grid_desc* inner,
                            att_grid_desc* outer, att_nlgrid_desc* nl);

/* Writes descriptors in ATTRACT's format.  The outer grid must have spacing
 * (float)(2 * gridspacing) and origin (float)(inner origin - gridextension *
 * gridspacing), the inner spacing (float)gridspacing, numGrids = the number of
 * alphabet types + 1, the neighbour grid the inner geometry.  rec_xyz (SoA,
 * natoms beads; may be NULL) sets each neighbour's ATTRACT type (1 within
 * plateaudis of the voxel point, else 2); without it every neighbour is
 * written as type 2. */
att_status att_gridfile_write(const char* path, const att_gridfile_info* info,
                              const att_grid_desc* inner, const att_grid_desc* outer,
                              const att_nlgrid_desc* nl, const float* rec_xyz);

#ifdef __cplusplus
}
#endif

#endif
