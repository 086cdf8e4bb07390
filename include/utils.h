/*
 * utils.h -- openQCD-1.4 utility routines used by the wloop parameter code
 * and the ms10 driver (reference wloop_parms.c:218-269, ms10.c:194-526):
 * input-file sections and tagged lines, command-line options, file helpers,
 * little-endian binary records.  Same names and meaning as openQCD's
 * utils/mutils.c and utils/endian.c.
 */
#ifndef UTILS_H
#define UTILS_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t stdint_t;

#define UNKNOWN_ENDIAN 0
#define LITTLE_ENDIAN_ORDER 1
#define BIG_ENDIAN_ORDER 2
#ifndef LITTLE_ENDIAN
#define LITTLE_ENDIAN LITTLE_ENDIAN_ORDER
#endif
#ifndef BIG_ENDIAN
#define BIG_ENDIAN BIG_ENDIAN_ORDER
#endif

#define IMAX(n, m) (((n) > (m)) ? (n) : (m))

/* sections "[title]" and "tag value..." lines of the input read from stdin;
 * text after '#' is a comment; a tag must appear exactly once per section */
long find_section(const char *title);
/* like find_section, but returns -1 if the section is absent */
long find_optional_section(const char *title);
int read_line(const char *tag, const char *format, ...);
void read_iprms(const char *tag, int n, int *iprms);
void read_dprms(const char *tag, int n, double *dprms);
/* re-read stdin on the next find_section (after freopen of stdin) */
void reset_input(void);

int find_opt(int argc, char *argv[], const char *opt);
int fdigits(double x);
int name_size(const char *format, ...);
void check_dir_root(const char *dir);
int copy_file(const char *in, const char *out);
int endianness(void);
void bswap_int(int n, void *a);
void bswap_double(int n, void *a);
int safe_mod(int x, int y);

#ifdef __cplusplus
}
#endif
#endif
