/*
 * records.c -- little-endian records of the ms10 files and the series of
 * processed configuration numbers (see records.h).
 *
 * Numbers are encoded byte by byte from their integer images, so the files
 * are the same on little- and big-endian hosts without a byte swap pass.
 */
#include "records.h"

#include <stdint.h>
#include <string.h>

static void put_u32(unsigned char *b, uint32_t v)
{
   for (int i = 0; i < 4; i++) b[i] = (unsigned char)(v >> (8 * i));
}

static uint32_t get_u32(const unsigned char *b)
{
   uint32_t v = 0;
   for (int i = 0; i < 4; i++) v |= (uint32_t)b[i] << (8 * i);
   return v;
}

static void put_u64(unsigned char *b, uint64_t v)
{
   for (int i = 0; i < 8; i++) b[i] = (unsigned char)(v >> (8 * i));
}

static uint64_t get_u64(const unsigned char *b)
{
   uint64_t v = 0;
   for (int i = 0; i < 8; i++) v |= (uint64_t)b[i] << (8 * i);
   return v;
}

int le_write_ints(FILE *f, const int *v, int n)
{
   int done = 0;
   for (; done < n; done++) {
      unsigned char b[4];
      put_u32(b, (uint32_t)(int32_t)v[done]);
      if (fwrite(b, 1, 4, f) != 4) break;
   }
   return done;
}

int le_read_ints(FILE *f, int *v, int n)
{
   int done = 0;
   for (; done < n; done++) {
      unsigned char b[4];
      if (fread(b, 1, 4, f) != 4) break;
      v[done] = (int)(int32_t)get_u32(b);
   }
   return done;
}

int le_write_doubles(FILE *f, const double *v, int n)
{
   int done = 0;
   for (; done < n; done++) {
      uint64_t u;
      unsigned char b[8];
      memcpy(&u, v + done, 8);
      put_u64(b, u);
      if (fwrite(b, 1, 8, f) != 8) break;
   }
   return done;
}

int le_read_doubles(FILE *f, double *v, int n)
{
   int done = 0;
   for (; done < n; done++) {
      unsigned char b[8];
      if (fread(b, 1, 8, f) != 8) break;
      uint64_t u = get_u64(b);
      memcpy(v + done, &u, 8);
   }
   return done;
}

int dat_header_put(FILE *f, const dat_header *h)
{
   int v[5] = {h->wls, h->msl, h->mwlt, h->mwlr, h->n0};
   return le_write_ints(f, v, 5) == 5;
}

int dat_header_get(FILE *f, dat_header *h)
{
   int v[5];
   if (le_read_ints(f, v, 5) != 5) return 0;
   h->wls = v[0];
   h->msl = v[1];
   h->mwlt = v[2];
   h->mwlr = v[3];
   h->n0 = v[4];
   return 1;
}

int dat_record_put(FILE *f, int nc, const double *w, int wls)
{
   return le_write_ints(f, &nc, 1) == 1 && le_write_doubles(f, w, wls) == wls;
}

int dat_record_get(FILE *f, int *nc, double *w, int wls)
{
   if (le_read_ints(f, nc, 1) != 1) return 0;
   return le_read_doubles(f, w, wls) == wls ? 1 : -1;
}

static void bc_values(double d[10])
{
   for (int i = 0; i < 10; i++) d[i] = (i < 4) ? 1.0 : 0.0;
}

int par_bc_put(FILE *f, int type)
{
   double d[10];
   bc_values(d);
   return le_write_ints(f, &type, 1) == 1 && le_write_doubles(f, d, 10) == 10;
}

int par_bc_matches(FILE *f, int type)
{
   int t;
   double d[10], want[10];
   if (le_read_ints(f, &t, 1) != 1 || le_read_doubles(f, d, 10) != 10) return -1;
   bc_values(want);
   int same = (t == type);
   for (int i = 0; i < 10; i++) same = same && (d[i] == want[i]);
   return same;
}

void series_reset(cnfg_series *s) { memset(s, 0, sizeof(*s)); }

void series_add(cnfg_series *s, int nc)
{
   if (s->count == 0)
      s->first = nc;
   else if (s->count == 1)
      s->spacing = nc - s->first;
   else if (nc - s->last != s->spacing)
      s->uneven = 1;
   s->last = nc;
   s->count++;
}
