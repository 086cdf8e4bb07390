/*
 * ms10.c -- Wilson-loop measurement program over libwloop_amd.so, with the
 * command line, input file and output files of wloop's main/ms10.c.
 *
 *   ms10 -i <input file> [-noexp] [-a]
 *
 * Input sections and tags as main/README.ms10:41-73 ([Run name], [Directories],
 * [Configurations], [Boundary conditions], [Wloop]); -noexp reads the imported
 * per-process files loc_dir/<name>n<no>_0 instead of cnfg_dir/<name>n<no>.
 * Optional sections the reference does not read:
 *   [Lattice]    size N0 N1 N2 N3    lattice extents (else from the first
 *                                    configuration file)
 *   [Synthetic]  seed <int>          configurations generated on the GPU
 *                eps  <double>       (seed + configuration number) instead
 *                                    of read from files
 *   [Device]     precision 0|1       0: FP64 (default), 1: mixed
 *                device <int>        GPU index
 * Output files (records.h): <name>.ms10.log and .ms10.dat with their ~
 * backups, <name>.ms10.par; -a continues a run whose .log and .dat agree;
 * <name>.ms10.end in log_dir stops the run after the current configuration.
 *
 * The program is a sequence of phases over one run_t; each phase returns the
 * next.  Until the log is opened, everything printed goes to STARTUP_ERROR
 * (removed if empty), where error_root leaves its message before exiting.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "smear.h"
#include "utils.h"
#include "wloop.h"
#include "records.h"

#define PATH_SIZE (2 * NAME_SIZE + 32)

typedef enum {
   PH_COMMAND_LINE,
   PH_INPUT_FILE,
   PH_PARAMETER_FILE,
   PH_DATA_FILE,
   PH_LOG_HEADER,
   PH_CONFIGURATION,
   PH_CLOSE,
   PH_END
} phase_t;

typedef enum { SRC_EXPORTED, SRC_IMPORTED, SRC_SYNTHETIC } source_t;

typedef struct {
   /* command line */
   const char *input;
   int noexp, append, endian;
   /* input file */
   char name[NAME_SIZE], log_dir[NAME_SIZE], dat_dir[NAME_SIZE], cnfg_dir[NAME_SIZE],
        loc_dir[NAME_SIZE];
   int first, last, step, bc_type;
   source_t source;
   long long seed;
   double eps;
   int precision, device, size[4];
   wloop_parms_t wl;
   dat_header head;
   /* files */
   char log_file[PATH_SIZE], par_file[PATH_SIZE], dat_file[PATH_SIZE], end_file[PATH_SIZE];
   FILE *out;        /* stdout: STARTUP_ERROR, then the log file */
   /* measurement loop */
   int nc, measured;
   double seconds;
   double *w;
} run_t;

static double seconds_now(void)
{
   struct timespec t;
   clock_gettime(CLOCK_MONOTONIC, &t);
   return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* "<dir>/<name><suffix>", checked against the buffer */
static void file_name(char *dst, const char *dir, const char *name, const char *suffix)
{
   int n = snprintf(dst, PATH_SIZE, "%s/%s%s", dir, name, suffix);
   error_root((n < 0) || (n >= PATH_SIZE), 1, "file_name [ms10.c]", "File name %s/%s%s is too long",
              dir, name, suffix);
}

static void backup(const char *file)
{
   char save[PATH_SIZE + 2];
   snprintf(save, sizeof(save), "%s~", file);
   copy_file(file, save);
}

/* the file configuration nc is read from */
static void cnfg_name(const run_t *r, int nc, char *dst)
{
   char suffix[64];
   if (r->source == SRC_IMPORTED) {
      snprintf(suffix, sizeof(suffix), "n%d_0", nc);
      file_name(dst, r->loc_dir, r->name, suffix);
   } else {
      snprintf(suffix, sizeof(suffix), "n%d", nc);
      file_name(dst, r->cnfg_dir, r->name, suffix);
   }
}

/* ---- phases ---------------------------------------------------------- */

static phase_t on_command_line(run_t *r, int argc, char *argv[])
{
   r->out = freopen("STARTUP_ERROR", "w", stdout);
   int at = find_opt(argc, argv, "-i");
   r->endian = endianness();
   error_root((at == 0) || (at == argc - 1), 1, "read_infile [ms10.c]",
              "Syntax: ms10 -i <input file> [-noexp] [-a]");
   error_root(r->endian == UNKNOWN_ENDIAN, 1, "read_infile [ms10.c]", "Machine has unknown endianness");
   r->input = argv[at + 1];
   r->noexp = find_opt(argc, argv, "-noexp") != 0;
   r->append = find_opt(argc, argv, "-a") != 0;
   return PH_INPUT_FILE;
}

/* lattice extents: [Lattice] or the header of the first configuration */
static void lattice_extents(run_t *r)
{
   if (r->size[0] > 0) {
      set_lattice_sizes(r->size[0], r->size[1], r->size[2], r->size[3]);
      return;
   }
   error_root(r->source == SRC_SYNTHETIC, 1, "set_sizes [ms10.c]", "[Synthetic] needs a [Lattice] section");
   char f[PATH_SIZE];
   int n[4];
   cnfg_name(r, r->first, f);
   if (r->source == SRC_IMPORTED)
      cnfg_sizes_imported(f, n);
   else
      cnfg_sizes(f, n);
   set_lattice_sizes(n[0], n[1], n[2], n[3]);
}

static phase_t on_input_file(run_t *r)
{
   error_root(freopen(r->input, "r", stdin) == NULL, 1, "read_infile [ms10.c]",
              "Unable to open input file");
   reset_input();

   find_section("Run name");
   read_line("name", "%127s", r->name);
   find_section("Configurations");
   read_line("first", "%d", &r->first);
   read_line("last", "%d", &r->last);
   read_line("step", "%d", &r->step);
   error_root((r->last < r->first) || (r->step < 1) || ((r->last - r->first) % r->step != 0), 1,
              "read_dirs [ms10.c]", "Improper configuration range");

   r->source = r->noexp ? SRC_IMPORTED : SRC_EXPORTED;
   if (find_optional_section("Synthetic") >= 0) {
      read_line("seed", "%lld", &r->seed);
      read_line("eps", "%lf", &r->eps);
      if (!r->noexp) r->source = SRC_SYNTHETIC;
   }
   find_section("Directories");
   read_line("log_dir", "%127s", r->log_dir);
   read_line("dat_dir", "%127s", r->dat_dir);
   if (r->source == SRC_IMPORTED)
      read_line("loc_dir", "%127s", r->loc_dir);
   else if (r->source == SRC_EXPORTED)
      read_line("cnfg_dir", "%127s", r->cnfg_dir);

   if (find_optional_section("Device") >= 0) {
      read_line("precision", "%d", &r->precision);
      read_line("device", "%d", &r->device);
      error_root((r->precision != 0) && (r->precision != 1), 1, "read_extra [ms10.c]",
                 "precision must be 0 (FP64) or 1 (mixed)");
   }
   if (find_optional_section("Lattice") >= 0) read_iprms("size", 4, r->size);

   find_section("Boundary conditions");
   read_line("type", "%d", &r->bc_type);
   error_root((r->bc_type != 0) && (r->bc_type != 3), 1, "read_bc_parms [ms10.c]",
              "Only open (0) and periodic (3) boundary conditions are supported");

   check_dir_root(r->log_dir);
   check_dir_root(r->dat_dir);
   file_name(r->log_file, r->log_dir, r->name, ".ms10.log");
   file_name(r->end_file, r->log_dir, r->name, ".ms10.end");
   file_name(r->par_file, r->dat_dir, r->name, ".ms10.par");
   file_name(r->dat_file, r->dat_dir, r->name, ".ms10.dat");
   if (r->source != SRC_SYNTHETIC) {
      char f[PATH_SIZE];
      cnfg_name(r, r->last, f);   /* fails now if the longest file name does not fit */
   }

   set_bc_type(r->bc_type);
   lattice_extents(r);
   set_wloop_device(r->device);
   set_wloop_precision(r->precision);
   read_wloop_parms();
   r->wl = wloop_parms();
   r->head = (dat_header){r->wl.wls, r->wl.msl, r->wl.mwlt, r->wl.mwlr, NPROC0 * L0};
   fclose(stdin);
   return PH_PARAMETER_FILE;
}

/* .par: written for a new run, compared for a continuation */
static phase_t on_parameter_file(run_t *r)
{
   FILE *f = fopen(r->par_file, r->append ? "rb" : "wb");
   error_root(f == NULL, 1, "read_infile [ms10.c]", "Unable to open parameter file");
   if (r->append) {
      int m = par_bc_matches(f, r->bc_type);
      error_root(m < 0, 1, "read_bc_parms [ms10.c]", "Incorrect read count");
      error_root(m == 0, 1, "read_bc_parms [ms10.c]", "Parameters do not match");
      check_wloop_parms(f);
   } else {
      error_root(!par_bc_put(f, r->bc_type), 1, "read_bc_parms [ms10.c]", "Incorrect write count");
      write_wloop_parms(f);
   }
   fclose(f);
   if (!r->append) backup(r->par_file);
   return PH_DATA_FILE;
}

/* the configurations a log file reports as processed; an unfinished
 * configuration at its end is an error */
static void log_series(const run_t *r, cnfg_series *s)
{
   FILE *f = fopen(r->log_file, "r");
   error_root(f == NULL, 1, "check_old_log [ms10.c]", "Unable to open log file");
   char text[NAME_SIZE];
   int bad_line = 0, open_cnfg = 0;
   series_reset(s);
   while (fgets(text, sizeof(text), f) != NULL) {
      int n[8];
      if (strstr(text, "process grid") != NULL) {
         bad_line |= sscanf(text, "%dx%dx%dx%d process grid, %dx%dx%dx%d", n, n + 1, n + 2, n + 3,
                            n + 4, n + 5, n + 6, n + 7) != 8;
      } else if (strstr(text, "fully processed") != NULL) {
         if (sscanf(text, "Configuration no %d", n) == 1)
            series_add(s, n[0]);
         else
            bad_line = 1;
         open_cnfg = 0;
      } else if (strstr(text, "Configuration no") != NULL) {
         open_cnfg = 1;
      }
   }
   fclose(f);
   error_root(bad_line, 1, "check_old_log [ms10.c]", "Incorrect read count");
   error_root(s->uneven, 1, "check_old_log [ms10.c]", "Configuration numbers are not equally spaced");
   error_root(open_cnfg || s->count == 0, 1, "check_old_log [ms10.c]",
              "Log file extends beyond the last configuration save");
}

/* the configurations a data file holds records of */
static void dat_series(const run_t *r, cnfg_series *s)
{
   FILE *f = fopen(r->dat_file, "rb");
   error_root(f == NULL, 1, "check_old_dat [ms10.c]", "Unable to open data file");
   dat_header h;
   error_root(!dat_header_get(f, &h), 1, "check_file_head [ms10.c]", "Incorrect read count");
   error_root(memcmp(&h, &r->head, sizeof(h)) != 0, 1, "check_file_head [ms10.c]",
              "Unexpected value of dn,nn,tmax or eps");
   series_reset(s);
   int nc, got;
   while ((got = dat_record_get(f, &nc, r->w, r->head.wls)) == 1) series_add(s, nc);
   fclose(f);
   error_root(got < 0, 1, "read_data [ms10.c]", "Read error or incomplete data record");
   error_root(s->count == 0, 1, "check_old_dat [ms10.c]", "No data records found");
   error_root(s->uneven, 1, "check_old_dat [ms10.c]", "Configuration numbers are not equally spaced");
}

static phase_t on_data_file(run_t *r)
{
   if (r->append) {
      cnfg_series lg, dt;
      log_series(r, &lg);
      dat_series(r, &dt);
      error_root((lg.first != dt.first) || (lg.last != dt.last) || (lg.spacing != dt.spacing), 1,
                 "check_old_dat [ms10.c]", "Configuration range is not as reported in the log file");
      error_root((lg.count > 1) && (lg.spacing != r->step), 1, "check_files [ms10.c]",
                 "Continuation run:\nPrevious run had a different configuration separation");
      error_root(r->first != lg.last + r->step, 1, "check_files [ms10.c]",
                 "Continuation run:\nConfiguration range does not continue the previous one");
      return PH_LOG_HEADER;
   }
   FILE *lf = fopen(r->log_file, "r"), *df = fopen(r->dat_file, "rb");
   if (lf) fclose(lf);
   if (df) fclose(df);
   error_root((lf != NULL) || (df != NULL), 1, "check_files [ms10.c]",
              "Attempt to overwrite old *.log or *.dat file");
   df = fopen(r->dat_file, "wb");
   error_root(df == NULL, 1, "check_files [ms10.c]", "Unable to open data file");
   error_root(!dat_header_put(df, &r->head), 1, "write_file_head [ms10.c]", "Incorrect write count");
   fclose(df);
   return PH_LOG_HEADER;
}

static phase_t on_log_header(run_t *r)
{
   long written = ftell(r->out);
   fclose(r->out);
   if (written == 0L) remove("STARTUP_ERROR");
   r->out = freopen(r->log_file, r->append ? "a" : "w", stdout);
   error_root(r->out == NULL, 1, "print_info [ms10.c]", "Unable to open log file");

   printf("\n%s", r->append ? "Continuation run\n\n"
                            : "Computation of Wilson loops\n--------------------------------------\n\n");
   printf("Wloop version %s\n", wloop_RELEASE);
   printf("openQCD version %s\n", openQCD_RELEASE);
   printf("MI355X build: libwloop_amd (HIP/gfx950), %s arithmetic\n",
          r->precision ? "mixed FP32/FP64" : "FP64");
   printf("The machine is %s endian\n", r->endian == LITTLE_ENDIAN ? "little" : "big");
   if (r->source == SRC_SYNTHETIC)
      printf("Configurations are generated on the GPU (seed %lld + no, eps %.6g)\n\n", r->seed, r->eps);
   else
      printf("Configurations are read in %s file format\n\n",
             r->source == SRC_IMPORTED ? "imported" : "exported");
   if (!r->append) {
      printf("%dx%dx%dx%d lattice, %dx%dx%dx%d local lattice\n", NPROC0 * L0, NPROC1 * L1,
             NPROC2 * L2, NPROC3 * L3, L0, L1, L2, L3);
      printf("%dx%dx%dx%d process grid, %dx%dx%dx%d process block size\n\n", NPROC0, NPROC1, NPROC2,
             NPROC3, NPROC0_BLK, NPROC1_BLK, NPROC2_BLK, NPROC3_BLK);
      printf("%s boundary conditions\n\n", r->bc_type == 0 ? "Open" : "Periodic");
      print_wloop_parms();
   }
   printf("Configurations no %d -> %d in steps of %d\n\n", r->first, r->last, r->step);
   fflush(r->out);
   geometry();
   r->nc = r->first;
   return PH_CONFIGURATION;
}

/* measure configuration r->nc (main/ms10.c:633-646 for what a measurement is) */
static phase_t on_configuration(run_t *r)
{
   double t0 = seconds_now();
   printf("Configuration no %d\n", r->nc);
   if (r->source == SRC_SYNTHETIC) {
      random_cnfg((unsigned long long)(r->seed + r->nc), r->eps);
   } else {
      char f[PATH_SIZE];
      cnfg_name(r, r->nc, f);
      if (r->source == SRC_IMPORTED)
         read_cnfg(f);
      else
         import_cnfg(f);
   }
   hyp_links(r->wl.alpha_action[0], r->wl.alpha_action[1], r->wl.alpha_action[2], r->wl.proj,
             r->wl.proj_iter, 0);
   wloop_sum(1);

   FILE *df = fopen(r->dat_file, "ab");
   error_root(df == NULL, 1, "save_data [ms10.c]", "Unable to open data file");
   wloop_parms_t now = wloop_parms();
   error_root(!dat_record_put(df, r->nc, now.wl_mat, r->head.wls), 1, "write_data [ms10.c]",
              "Incorrect write count");
   fclose(df);

   double dt = seconds_now() - t0;
   r->seconds += dt;
   r->measured++;
   error_chk();
   printf("Configuration no %d fully processed in %.2e sec ", r->nc, dt);
   printf("(average = %.2e sec)\n\n", r->seconds / (double)r->measured);
   fflush(r->out);
   backup(r->log_file);
   backup(r->dat_file);

   FILE *ef = fopen(r->end_file, "r");
   if (ef != NULL) {
      fclose(ef);
      remove(r->end_file);
      printf("End flag set, run stopped\n\n");
      return PH_CLOSE;
   }
   r->nc += r->step;
   return (r->nc <= r->last) ? PH_CONFIGURATION : PH_CLOSE;
}

static phase_t on_close(run_t *r)
{
   error_chk();
   fflush(r->out);
   backup(r->log_file);
   fclose(r->out);
   return PH_END;
}

int main(int argc, char *argv[])
{
   static run_t run;
   phase_t ph = PH_COMMAND_LINE;
   while (ph != PH_END) {
      switch (ph) {
         case PH_COMMAND_LINE: ph = on_command_line(&run, argc, argv); break;
         case PH_INPUT_FILE:
            ph = on_input_file(&run);
            run.w = malloc(sizeof(double) * (size_t)run.head.wls);
            error(run.w == NULL, 1, "main [ms10.c]", "Unable to allocate data arrays");
            break;
         case PH_PARAMETER_FILE: ph = on_parameter_file(&run); break;
         case PH_DATA_FILE: ph = on_data_file(&run); break;
         case PH_LOG_HEADER: ph = on_log_header(&run); break;
         case PH_CONFIGURATION: ph = on_configuration(&run); break;
         case PH_CLOSE: ph = on_close(&run); break;
         case PH_END: break;
      }
   }
   free(run.w);
   return 0;
}
