/*
 * attract_formats.h — the reference's input and output files on the host:
 * reduced PDB bodies, normal-mode files, attract.par force fields, DOF files
 * (header with pivots and centering, one block of poses per "#" line) with
 * their conversion to receptor-frame ligand DOFs, and the result text the
 * reference's scoring and minimization apps print.
 *
 * Each reader follows the reference's reader line by line in what it accepts
 * and how it converts text to numbers (same stream extraction, same float
 * arithmetic), so the arrays equal the reference's bit for bit; where the
 * reference reads uninitialized memory or exits on malformed input, these
 * return an error instead.  Library: libattract_service.so (host code only).
 */
#ifndef ATTRACT_FORMATS_H_
#define ATTRACT_FORMATS_H_

#include <stddef.h>
#include <stdint.h>

#include "attract_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the calling thread's last failed call of this header. */
const char* att_formats_last_error(void);

/* --- bodies ----------------------------------------------------------------
 * att_body_read_pdb: readProteinFromPDB (fileIO/readFile.cpp:60-76,108-170):
 * columns 30-38, 38-46, 46-54 (x, y, z), 54-59 (type), 59-67 (charge) of every
 * non-empty line; a last line without a newline is not read (the reference's
 * loop stops at end of file before using it).
 * att_body_read_modes: readHMMode (readFile.cpp:187-258): per mode a line
 * "index eigenvalue" then one "x y z" line per bead; each mode is divided by
 * its norm; the force constant is eigenvalue^2.
 * att_body_pivotize: Protein::pivotize / auto_pivotize (model/Protein.tpp:
 * 130-167); pivot NULL = the centroid.  The pivot used is returned.
 * att_body_map_types: the mapped (grid) types of a ligand from a receptor
 * grid alphabet (model/TypeMap.h:52-60): alphabet[i] -> i+1, 99 -> 0, 0 -> 0;
 * n_alphabet = 0: identity (the default map).  Unmapped types are an error. */
typedef struct att_body att_body;
att_status att_body_read_pdb(const char* path, att_body** out);
att_status att_body_read_modes(att_body* b, const char* path, unsigned numModes);
att_status att_body_pivotize(att_body* b, const float* pivot, float* pivot_out);
att_status att_body_map_types(att_body* b, const unsigned* alphabet, unsigned n_alphabet);
/* Descriptor over the body's arrays (valid until the body changes or is freed). */
att_status att_body_desc(const att_body* b, att_protein_desc* out);
void att_body_free(att_body* b);

/* --- force field: readParamTableFromFile (readFile.cpp:439-527) -----------
 * "shape numTypes swiOn swiOff", then rbc, abc (numTypes^2 each) and ipon,
 * derived to {rc, ac, ipon, rmin2, emin}.  capacity = entries of `table`. */
att_status att_read_params(const char* path, att_pair_params* table, unsigned capacity,
                           unsigned* numTypes, int* potShape, float* swiOn, float* swiOff);

/* --- grid alphabet files: readGridAlphabetFromFile (readFile.cpp:901-926) ----
 * One non-negative integer (a bead type) per line, in the order that
 * att_body_map_types takes.  With out NULL only *n is set (size query);
 * capacity = entries of `out`. */
att_status att_read_alphabet(const char* path, unsigned* out, unsigned capacity, unsigned* n);

/* --- DOF files: readDOFHeader (readFile.cpp:766-871) + readDOF (:579-651) ---
 * then DOFConverter[_Modes] (model/DOFConverter.h) and transformDOF_glob2rec
 * (model/DOFTransform.cpp:36-84) for a receptor-ligand pair. */
typedef struct att_dof_file att_dof_file;
typedef struct {
    unsigned numMolecules;      /* lines per pose block */
    size_t numPoses;
    int autoPivot;              /* "#pivot auto" */
    unsigned numPivots;         /* explicit pivots (at most 2 are kept) */
    float pivots[2][3];
    int centeredReceptor, centeredLigands;
} att_dof_info;
att_status att_dof_read(const char* path, att_dof_file** out);
att_status att_dof_get_info(const att_dof_file* f, att_dof_info* info);
/* One line of the file as readDOF reads it (readFile.cpp:579-651): the
 * values of molecule `molecule` in pose block `pose` into values[0..15]
 * (phi, ssi, rot, x, y, z, then the mode amplitudes), their number (6 plus
 * the amplitudes) into *count. */
att_status att_dof_values(const att_dof_file* f, unsigned molecule, size_t pose, double* values,
                          unsigned* count);
/* Ligand DOFs in the receptor frame (receptor pose folded into the ligand's,
 * pivots and centering applied) in float, as att_dof6 (modes = 0) or
 * att_dof6m records (modes = 1: receptor mode amplitudes from the receptor's
 * line, ligand amplitudes from the ligand's; numModesRec/Lig of them, the rest
 * 0).  Pivots: the file's, or for "#pivot auto" the caller's (the bodies'
 * centroids, att_body_pivotize with NULL). */
att_status att_dof_ligand(const att_dof_file* f, const float* pivotRec, const float* pivotLig,
                          int modes, unsigned numModesRec, unsigned numModesLig, void* out);
void att_dof_free(att_dof_file* f);

/* --- results: operator<< of Result_6D / Result_6D_Modes (model/Types_6D.tpp:
 * 39-67, Types_6D_Modes.tpp:49-88), one record per pose followed by a newline,
 * as the apps print them (apps/sc/scATTRACT.tpp:45-47).  Writes at most cap
 * bytes (no terminator) and the full length to *len. */
att_status att_format_results(const void* results, size_t n, int modes, unsigned numModesRec,
                              unsigned numModesLig, char* buf, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif

#endif
