/*
 * records.h -- the binary files of an ms10 run and the bookkeeping of which
 * configurations a run has processed.
 *
 * Formats (main/README.ms10:112-159, main/ms10.c:86-156): every number is
 * stored little endian, whatever the host.
 *   .dat  header of 5 int32 (wls, msl, mwlt, mwlr, N0), then per
 *         configuration one int32 (its number) and wls doubles (W in wl_idx
 *         order).
 *   .par  the boundary-condition record (int32 type, 10 doubles: cG, cG',
 *         cF, cF', phi[2][3]) followed by the [Wloop] record of
 *         write_wloop_parms (wloop_parms.c:343-382).
 */
#ifndef MS10_RECORDS_H
#define MS10_RECORDS_H

#include <stdio.h>

/* n int32 / doubles to or from f; the return value is the count moved */
int le_write_ints(FILE *f, const int *v, int n);
int le_read_ints(FILE *f, int *v, int n);
int le_write_doubles(FILE *f, const double *v, int n);
int le_read_doubles(FILE *f, double *v, int n);

typedef struct {
   int wls, msl, mwlt, mwlr, n0;
} dat_header;

/* 1 on success */
int dat_header_put(FILE *f, const dat_header *h);
int dat_header_get(FILE *f, dat_header *h);
/* one record; dat_record_get returns 1 (read), 0 (end of file) or -1 (a
 * truncated record) */
int dat_record_put(FILE *f, int nc, const double *w, int wls);
int dat_record_get(FILE *f, int *nc, double *w, int wls);

/* the boundary-condition record of openQCD's write_bc_parms for the types
 * supported here (0 open, 3 periodic: cG = cG' = cF = cF' = 1, phi = 0) */
int par_bc_put(FILE *f, int type);
/* 1 if the record in f is the one par_bc_put writes for `type`, 0 if it
 * differs, -1 if it is incomplete */
int par_bc_matches(FILE *f, int type);

/* Configuration numbers in the order a run processed them.  A run steps
 * through first, first+step, ...: `uneven` is set once a number breaks the
 * spacing the first two established. */
typedef struct {
   int count, first, last, spacing, uneven;
} cnfg_series;

void series_reset(cnfg_series *s);
void series_add(cnfg_series *s, int nc);

#endif
