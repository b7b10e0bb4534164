/* Whole-case runs over the reference's files, on the GPU.
 *
 * A case directory holds Nemoh.cal and the body mesh files, as for the
 * reference.  nemoh_case_run replaces the reference's preProcessor and
 * Solver executables (preProcessor/Main.f90:26-337, Solver/NEMOH.f90:26-175):
 * it reads and writes the same files, record for record in the same text
 * formats, so either stage can be mixed with the reference's own (e.g. the
 * reference preProcessor's files solved here, or these files post-processed
 * by the reference postProcessor).
 *
 *   stages & NEMOH_STAGE_PRE:   mesh/L10.dat, L12.dat, Mesh.tec,
 *     Integration.dat, Kochin.dat, Freesurface.dat, Normalvelocities.dat,
 *     results/FKForce.dat, FKForce.tec, index.dat.  Body conditions, force
 *     weights and Froude-Krylov pressures come from the device
 *     (nemoh_core_set_dofs / nemoh_core_body_conditions) on the mesh files'
 *     own coordinates, as the preProcessor computes them.
 *   stages & NEMOH_STAGE_SOLVE: reads the files above as the Solver does
 *     (the mesh from L12.dat, i.e. coordinates rounded to 7 digits),
 *     solves every problem on the device (all problems of one frequency per
 *     call) and writes results/Forces.dat and, where the problems ask for
 *     them, Kochin.<n>.dat, freesurface.<n>.dat and pressure.<n>.dat.
 *   stages & NEMOH_STAGE_POST:  the postProcessor (postProcessor/Main.f90:
 *     21-101, Common/Results.f90:132-251, postProcessor/IRF.f90:66-151) on
 *     the Solver's files: results/RadiationCoefficients.tec,
 *     DiffractionForce.tec, ExcitationForce.tec, CA.dat, CM.dat, Fe.dat and,
 *     when Nemoh.cal's IRF switch is 1, IRF.tec.  Host work only.
 *
 * precision: NEMOH_FP64 or NEMOH_MIXED (include/nemoh_core.h).  Returns
 * NEMOH_OK or an error code, with a message in msg (msglen bytes). */
#ifndef NEMOH_CASE_H
#define NEMOH_CASE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { NEMOH_STAGE_PRE = 1, NEMOH_STAGE_SOLVE = 2, NEMOH_STAGE_POST = 4 };

int nemoh_case_run(const char *case_dir, int stages, int precision, char *msg, int msglen);

/* The host-only part of the preProcessor stage (no device needed): mesh/
 * L10.dat, L12.dat, Mesh.tec, Kochin.dat, Freesurface.dat and
 * results/index.dat. */
int nemoh_case_write_mesh_files(const char *case_dir, char *msg, int msglen);

/* Formats n reals as one list-directed WRITE of the reference's Fortran
 * runtime (records of at most 79 characters, shortest round-trip digits)
 * into out (cap bytes, NUL-terminated); returns the length or -1 if it does
 * not fit.  For checking the writers. */
int nemoh_format_list(int n, const double *v, char *out, size_t cap);
/* The edit descriptor Ew.d of the same runtime (0.ddd form, "E+dd" or
 * "+ddd" exponent) into out; returns the length or -1. */
int nemoh_format_e(double v, int w, int d, char *out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
