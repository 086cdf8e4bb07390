/*
 * lattice.h -- the openQCD-1.4 pieces the measurement path calls, supplied by
 * libwloop_amd.so for the standalone build (openQCD is not part of it).
 * Names and meaning follow openQCD: geometry() builds ipt/iup/idn, udfld()
 * returns the double-precision gauge field (4*VOLUME su3_dble; link U(x,mu)
 * for odd x at [8*(ix-VOLUME/2)+2*mu], for even x at the odd neighbour
 * x+mu, slot +2*mu+1 -- modules/lnk/lnkcom.c:840-859), set_bc() zeroes the
 * time links at x0 = N0-1 for open boundary conditions.
 *
 * The gauge field also lives on the GPU; the library keeps the two copies
 * coherent: udfld() hands out the host copy (downloading it first if the
 * device holds newer links) and the next GPU operation uploads it again.
 */
#ifndef LATTICE_H
#define LATTICE_H

#include <stdio.h>
#include "su3.h"
#include "global.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces openQCD's compile-time L0..L3 (global.h); extents must be even */
void set_lattice_sizes(int n0, int n1, int n2, int n3);
void geometry(void);

/* boundary conditions: 0 open, 3 periodic (the two the reference supports,
 * README:22); the other openQCD parameters are not used by the path */
void set_bc_type(int type);
int bc_type(void);
void set_bc(void);

su3_dble *udfld(void);
void copy_bnd_ud(void);

#define UPDATED_UD 1
void set_flags(int event);

/* openQCD-style error reporting: print to stdout and abort */
void error(int test, int no, const char *name, const char *format, ...);
void error_root(int test, int no, const char *name, const char *format, ...);
void error_chk(void);
void message(const char *format, ...);

/* GPU device used by the module entry points (default 0) */
void set_wloop_device(int device);

/* configurations (openQCD archive.h names).  Exported format: 4 int32 lattice
 * sizes, the average plaquette (double), then for every odd point x in
 * lexicographic order (x3 fastest) the 8 links U(x,0),U(x-0,0),...,U(x,3),
 * U(x-3,3), little endian.  import_cnfg checks sizes and plaquette. */
void import_cnfg(const char *in);
void export_cnfg(const char *out);
/* lattice sizes from the header of an exported configuration */
void cnfg_sizes(const char *in, int n[4]);
/* Imported (per-process) format, as ms10 -noexp reads it (main/ms10.c:706-713):
 * 16 native ints (process grid, local sizes, process block sizes, process
 * coordinates), the local plaquette sum (double), then udfld() as it lies in
 * memory.  One process owns the lattice here, so the grid is 1x1x1x1.
 * read_cnfg checks the header and the plaquette sum. */
void read_cnfg(const char *in);
void write_cnfg(const char *out);
/* local lattice sizes from the header of an imported configuration */
void cnfg_sizes_imported(const char *in, int n[4]);
/* No reference counterpart: a seeded synthetic configuration in udfld(),
 * generated on the GPU (U = exp(i eps H), Haar gauge transform, set_bc) */
void random_cnfg(unsigned long long seed, double eps);
/* average plaquette Re tr / 3 over all plaquettes (open bc: time plaquettes
 * at x0 = N0-1 left out) */
double avg_plaq(void);

#ifdef __cplusplus
}
#endif
#endif
