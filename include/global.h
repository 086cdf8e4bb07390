/*
 * global.h -- runtime lattice geometry for the standalone build.
 *
 * openQCD-1.4 fixes the lattice and process grid at compile time in its
 * global.h (main/README.ms10:36-37).  This library measures lattices of any
 * size on one GPU, so the same names resolve to runtime values set by
 * set_lattice_sizes() (lattice.h).  One process owns the whole lattice:
 * NPROC = 1, no boundary buffers (BNDRY = 0), every index wraps periodically.
 */
#ifndef GLOBAL_H
#define GLOBAL_H

extern int wl_dims[4];
extern int wl_volume;

#define NPROC0 1
#define NPROC1 1
#define NPROC2 1
#define NPROC3 1
#define NPROC 1
#define NPROC0_BLK 1
#define NPROC1_BLK 1
#define NPROC2_BLK 1
#define NPROC3_BLK 1

#define L0 (wl_dims[0])
#define L1 (wl_dims[1])
#define L2 (wl_dims[2])
#define L3 (wl_dims[3])
#define VOLUME (wl_volume)

#define FACE0 0
#define FACE1 0
#define FACE2 0
#define FACE3 0
#define BNDRY 0

/* the openQCD interface the library supplies (the pieces the path calls) */
#define openQCD_RELEASE "openQCD-1.4"

#define NAME_SIZE 128
#define ALIGN 6

/* even points first: ipt[x3+L3*x2+L2*L3*x1+L1*L2*L3*x0] in [0,VOLUME/2) for
 * even x0+x1+x2+x3, in [VOLUME/2,VOLUME) for odd; iup/idn: neighbours */
extern int *ipt;
extern int (*iup)[4];
extern int (*idn)[4];
extern int *map;
extern int cpr[4];

#endif
