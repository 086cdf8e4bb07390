/*
 * utils.c -- input-file parsing and file utilities (utils.h).
 */
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "lattice.h"
#include "utils.h"

static char **lines = NULL;
static int nlines = 0;
static int sec_first = -1, sec_last = -1;

void reset_input(void)
{
   for (int i = 0; i < nlines; i++) free(lines[i]);
   free(lines);
   lines = NULL;
   nlines = 0;
   sec_first = sec_last = -1;
}

static void load_input(void)
{
   if (lines != NULL) return;
   int cap = 64;
   char buf[4096];
   lines = malloc(sizeof(char *) * cap);
   error(lines == NULL, 1, "find_section [utils.c]", "Out of memory");
   while (fgets(buf, sizeof(buf), stdin) != NULL) {
      char *h = strchr(buf, '#');
      if (h) *h = '\0';
      if (nlines == cap) {
         cap *= 2;
         lines = realloc(lines, sizeof(char *) * cap);
         error(lines == NULL, 1, "find_section [utils.c]", "Out of memory");
      }
      lines[nlines++] = strdup(buf);
   }
}

static int is_section(const char *l, char *title, size_t n)
{
   while (*l == ' ' || *l == '\t') l++;
   if (*l != '[') return 0;
   const char *e = strchr(l, ']');
   if (!e) return 0;
   size_t len = (size_t)(e - l - 1);
   if (len >= n) len = n - 1;
   memcpy(title, l + 1, len);
   title[len] = '\0';
   return 1;
}

long find_optional_section(const char *title)
{
   load_input();
   char t[256];
   for (int i = 0; i < nlines; i++)
      if (is_section(lines[i], t, sizeof(t)) && strcmp(t, title) == 0) return find_section(title);
   return -1;
}

long find_section(const char *title)
{
   load_input();
   int found = -1, count = 0;
   char t[256];
   for (int i = 0; i < nlines; i++)
      if (is_section(lines[i], t, sizeof(t)) && strcmp(t, title) == 0) {
         if (found < 0) found = i;
         count++;
      }
   error(count == 0, 1, "find_section [utils.c]", "Section [%s] not found", title);
   error(count > 1, 1, "find_section [utils.c]", "Section [%s] occurs more than once", title);
   sec_first = found + 1;
   sec_last = nlines;
   for (int i = sec_first; i < nlines; i++)
      if (is_section(lines[i], t, sizeof(t))) {
         sec_last = i;
         break;
      }
   return (long)found;
}

/* the text after the tag in the current section */
static const char *tag_value(const char *tag)
{
   error(sec_first < 0, 1, "read_line [utils.c]", "No section selected before reading %s", tag);
   const char *val = NULL;
   int count = 0;
   size_t n = strlen(tag);
   for (int i = sec_first; i < sec_last; i++) {
      const char *l = lines[i];
      while (*l == ' ' || *l == '\t') l++;
      if (strncmp(l, tag, n) == 0 && (l[n] == ' ' || l[n] == '\t' || l[n] == '\n' || l[n] == '\0')) {
         val = l + n;
         count++;
      }
   }
   error(count == 0, 1, "read_line [utils.c]", "Line with tag %s not found", tag);
   error(count > 1, 1, "read_line [utils.c]", "Tag %s occurs more than once", tag);
   return val;
}

int read_line(const char *tag, const char *format, ...)
{
   const char *v = tag_value(tag);
   va_list ap;
   va_start(ap, format);
   int nconv = 0;
   const char *f = format;
   while (*f) if (*f++ == '%') nconv++;
   int ir = vsscanf(v, format, ap);
   va_end(ap);
   error(ir != nconv, 1, "read_line [utils.c]", "Read error in line with tag %s", tag);
   return ir;
}

void read_iprms(const char *tag, int n, int *iprms)
{
   const char *v = tag_value(tag);
   for (int i = 0; i < n; i++) {
      char *end;
      long x = strtol(v, &end, 10);
      error(end == v, 1, "read_iprms [utils.c]", "Read error in line with tag %s", tag);
      iprms[i] = (int)x;
      v = end;
   }
}

void read_dprms(const char *tag, int n, double *dprms)
{
   const char *v = tag_value(tag);
   for (int i = 0; i < n; i++) {
      char *end;
      double x = strtod(v, &end);
      error(end == v, 1, "read_dprms [utils.c]", "Read error in line with tag %s", tag);
      dprms[i] = x;
      v = end;
   }
}

int find_opt(int argc, char *argv[], const char *opt)
{
   for (int k = 1; k < argc; k++)
      if (strcmp(argv[k], opt) == 0) return k;
   return 0;
}

/* smallest n such that %.nf reproduces x to a few units of DBL_EPSILON */
int fdigits(double x)
{
   char buf[64];
   for (int n = 0; n < 17; n++) {
      snprintf(buf, sizeof(buf), "%.*f", n, x);
      double y = atof(buf);
      if (fabs(x - y) <= 4.0 * DBL_EPSILON * fabs(x)) return n;
   }
   return 16;
}

int name_size(const char *format, ...)
{
   va_list ap;
   va_start(ap, format);
   int n = vsnprintf(NULL, 0, format, ap);
   va_end(ap);
   return n;
}

void check_dir_root(const char *dir)
{
   struct stat st;
   error((stat(dir, &st) != 0) || !S_ISDIR(st.st_mode), 1, "check_dir_root [utils.c]",
         "Unable to access directory %s", dir);
}

int copy_file(const char *in, const char *out)
{
   FILE *fi = fopen(in, "rb");
   error(fi == NULL, 1, "copy_file [utils.c]", "Unable to open %s", in);
   FILE *fo = fopen(out, "wb");
   error(fo == NULL, 1, "copy_file [utils.c]", "Unable to open %s", out);
   char buf[65536];
   size_t n;
   while ((n = fread(buf, 1, sizeof(buf), fi)) > 0)
      error(fwrite(buf, 1, n, fo) != n, 1, "copy_file [utils.c]", "Write error on %s", out);
   fclose(fi);
   fclose(fo);
   return 0;
}

int endianness(void)
{
   uint32_t x = 0x01020304u;
   unsigned char *c = (unsigned char *)&x;
   if (c[0] == 4) return LITTLE_ENDIAN;
   if (c[0] == 1) return BIG_ENDIAN;
   return UNKNOWN_ENDIAN;
}

static void swap_bytes(int n, int size, void *a)
{
   unsigned char *p = a;
   for (int i = 0; i < n; i++, p += size)
      for (int j = 0; j < size / 2; j++) {
         unsigned char t = p[j];
         p[j] = p[size - 1 - j];
         p[size - 1 - j] = t;
      }
}

void bswap_int(int n, void *a) { swap_bytes(n, 4, a); }
void bswap_double(int n, void *a) { swap_bytes(n, 8, a); }

int safe_mod(int x, int y)
{
   int r = x % y;
   return (r < 0) ? r + y : r;
}
