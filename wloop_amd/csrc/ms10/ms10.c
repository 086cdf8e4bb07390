/*
 * ms10.c -- Wilson-loop measurement driver, a drop-in for wloop's main/ms10.c
 * with the measurement on the GPU (libwloop_amd.so).
 *
 *   ms10 -i <input file> [-noexp] [-a]
 *
 * Input file: the sections and tags of main/README.ms10:41-73
 * ([Run name] [Directories] [Configurations] [Boundary conditions] [Wloop]),
 * plus optional sections the reference does not read:
 *   [Lattice]    size N0 N1 N2 N3    lattice extents (else taken from the
 *                                    header of the first configuration)
 *   [Synthetic]  seed <int>          configurations are generated on the GPU
 *                eps  <double>       (seed + configuration number) instead of
 *                                    read from cnfg_dir
 *   [Device]     precision 0|1       0: FP64 (default), 1: mixed
 *                device <int>        GPU index
 * Output: <name>.ms10.log, .dat, .par and their ~ backups, records as
 * ms10.c:86-156 (header 5 int32; per configuration int32 nc and wls doubles,
 * little endian); -a continues a run after the checks of ms10.c:374-526;
 * <name>.ms10.end in log_dir stops the run after the current configuration.
 * -noexp (imported per-process files) is not supported: one process owns the
 * whole lattice here.
 */
#include <stdint.h>
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

#define N0 (NPROC0 * L0)
#define N1 (NPROC1 * L1)
#define N2 (NPROC2 * L2)
#define N3 (NPROC3 * L3)

static struct
{
   int wls, msl, wlmt, wlmr, tmax;
} file_head;

static struct
{
   int nc;
   double *wl_mat;
} data;

static int noexp, append, endian, first, last, step, bc;
static int synthetic, precision, device, lattice[4];
static long long synth_seed;
static double synth_eps;
static char line[NAME_SIZE];
static char log_dir[NAME_SIZE], dat_dir[NAME_SIZE], loc_dir[NAME_SIZE], cnfg_dir[NAME_SIZE];
/* file names: lengths are checked against NAME_SIZE before they are built */
#define PATH_SIZE (2 * NAME_SIZE + 32)
static char log_file[PATH_SIZE], log_save[PATH_SIZE + 2], end_file[PATH_SIZE];
static char par_file[PATH_SIZE], par_save[PATH_SIZE + 2], dat_file[PATH_SIZE], dat_save[PATH_SIZE + 2];
static char cnfg_file[PATH_SIZE], nbase[NAME_SIZE];
static FILE *fin = NULL, *flog = NULL, *fdat = NULL, *fend = NULL;
static wloop_parms_t wl;

static double wtime(void)
{
   struct timespec t;
   clock_gettime(CLOCK_MONOTONIC, &t);
   return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void write_file_head(void)
{
   stdint_t istd[5] = {file_head.wls, file_head.msl, file_head.wlmt, file_head.wlmr, file_head.tmax};
   if (endian == BIG_ENDIAN) bswap_int(5, istd);
   error_root(fwrite(istd, sizeof(stdint_t), 5, fdat) != 5, 1, "write_file_head [ms10.c]",
              "Incorrect write count");
}

static void check_file_head(void)
{
   stdint_t istd[5];
   error_root(fread(istd, sizeof(stdint_t), 5, fdat) != 5, 1, "check_file_head [ms10.c]",
              "Incorrect read count");
   if (endian == BIG_ENDIAN) bswap_int(5, istd);
   error_root((istd[0] != file_head.wls) || (istd[1] != file_head.msl) || (istd[2] != file_head.wlmt) ||
              (istd[3] != file_head.wlmr) || (istd[4] != file_head.tmax), 1,
              "check_file_head [ms10.c]", "Unexpected value of dn,nn,tmax or eps");
}

static void write_data(void)
{
   stdint_t istd = data.nc;
   if (endian == BIG_ENDIAN) bswap_int(1, &istd);
   size_t iw = fwrite(&istd, sizeof(stdint_t), 1, fdat);
   for (int i = 0; i < file_head.wls; i++) {
      double d = data.wl_mat[i];
      if (endian == BIG_ENDIAN) bswap_double(1, &d);
      iw += fwrite(&d, sizeof(double), 1, fdat);
   }
   error_root(iw != (size_t)(1 + file_head.wls), 1, "write_data [ms10.c]", "Incorrect write count");
}

static int read_data(void)
{
   stdint_t istd;
   if (fread(&istd, sizeof(stdint_t), 1, fdat) != 1) return 0;
   if (endian == BIG_ENDIAN) bswap_int(1, &istd);
   data.nc = istd;
   size_t ir = 1;
   for (int i = 0; i < file_head.wls; i++) {
      double d;
      ir += fread(&d, sizeof(double), 1, fdat);
      if (endian == BIG_ENDIAN) bswap_double(1, &d);
      data.wl_mat[i] = d;
   }
   error_root(ir != (size_t)(1 + file_head.wls), 1, "read_data [ms10.c]",
              "Read error or incomplete data record");
   return 1;
}

static void read_dirs(void)
{
   find_section("Run name");
   read_line("name", "%127s", nbase);
   find_section("Directories");
   read_line("log_dir", "%127s", log_dir);
   read_line("dat_dir", "%127s", dat_dir);
   if (noexp) {
      read_line("loc_dir", "%127s", loc_dir);
      cnfg_dir[0] = '\0';
   } else {
      if (find_optional_section("Synthetic") < 0) {
         find_section("Directories");
         read_line("cnfg_dir", "%127s", cnfg_dir);
      }
      loc_dir[0] = '\0';
   }
   find_section("Configurations");
   read_line("first", "%d", &first);
   read_line("last", "%d", &last);
   read_line("step", "%d", &step);
   error_root((last < first) || (step < 1) || (((last - first) % step) != 0), 1, "read_dirs [ms10.c]",
              "Improper configuration range");
}

static void read_extra(void)
{
   synthetic = 0;
   precision = 0;
   device = 0;
   lattice[0] = 0;
   if (find_optional_section("Synthetic") >= 0) {
      synthetic = 1;
      read_line("seed", "%lld", &synth_seed);
      read_line("eps", "%lf", &synth_eps);
   }
   if (find_optional_section("Device") >= 0) {
      read_line("precision", "%d", &precision);
      read_line("device", "%d", &device);
      error_root((precision != 0) && (precision != 1), 1, "read_extra [ms10.c]",
                 "precision must be 0 (FP64) or 1 (mixed)");
   }
   if (find_optional_section("Lattice") >= 0) {
      int n[4];
      read_iprms("size", 4, n);
      for (int m = 0; m < 4; m++) lattice[m] = n[m];
   }
}

static void setup_files(void)
{
   error_root(noexp, 1, "setup_files [ms10.c]",
              "-noexp is not supported: one process owns the whole lattice");
   if (!synthetic)
      error_root(name_size("%s/%sn%d", cnfg_dir, nbase, last) >= NAME_SIZE, 1, "setup_files [ms10.c]",
                 "cnfg_dir name is too long");
   check_dir_root(log_dir);
   check_dir_root(dat_dir);
   error_root(name_size("%s/%s.ms10.log~", log_dir, nbase) >= NAME_SIZE, 1, "setup_files [ms10.c]",
              "log_dir name is too long");
   error_root(name_size("%s/%s.ms10.dat~", dat_dir, nbase) >= NAME_SIZE, 1, "setup_files [ms10.c]",
              "dat_dir name is too long");
   snprintf(log_file, sizeof(log_file), "%s/%s.ms10.log", log_dir, nbase);
   snprintf(par_file, sizeof(par_file), "%s/%s.ms10.par", dat_dir, nbase);
   snprintf(dat_file, sizeof(dat_file), "%s/%s.ms10.dat", dat_dir, nbase);
   snprintf(end_file, sizeof(end_file), "%s/%s.ms10.end", log_dir, nbase);
   snprintf(log_save, sizeof(log_save), "%s~", log_file);
   snprintf(par_save, sizeof(par_save), "%s~", par_file);
   snprintf(dat_save, sizeof(dat_save), "%s~", dat_file);
}

/* boundary-condition record of the .par file: type (int32), then
 * cG, cG', cF, cF' and phi[2][3] (10 doubles), as openQCD's write_bc_parms */
static void bc_record(stdint_t *istd, double *dstd)
{
   istd[0] = bc;
   for (int i = 0; i < 4; i++) dstd[i] = 1.0;
   for (int i = 4; i < 10; i++) dstd[i] = 0.0;
}

static void read_bc_parms(void)
{
   stdint_t istd[1], ref_i[1];
   double dstd[10], ref_d[10];
   find_section("Boundary conditions");
   read_line("type", "%d", &bc);
   error_root((bc != 0) && (bc != 3), 1, "read_bc_parms [ms10.c]",
              "Only open (0) and periodic (3) boundary conditions are supported");
   set_bc_type(bc);
   bc_record(ref_i, ref_d);
   if (append) {
      size_t ir = fread(istd, sizeof(stdint_t), 1, fdat);
      ir += fread(dstd, sizeof(double), 10, fdat);
      error_root(ir != 11, 1, "read_bc_parms [ms10.c]", "Incorrect read count");
      if (endian == BIG_ENDIAN) {
         bswap_int(1, istd);
         bswap_double(10, dstd);
      }
      int ie = (istd[0] != ref_i[0]);
      for (int i = 0; i < 10; i++) ie |= (dstd[i] != ref_d[i]);
      error_root(ie, 1, "read_bc_parms [ms10.c]", "Parameters do not match");
   } else {
      if (endian == BIG_ENDIAN) {
         bswap_int(1, ref_i);
         bswap_double(10, ref_d);
      }
      size_t iw = fwrite(ref_i, sizeof(stdint_t), 1, fdat);
      iw += fwrite(ref_d, sizeof(double), 10, fdat);
      error_root(iw != 11, 1, "read_bc_parms [ms10.c]", "Incorrect write count");
   }
}

static void set_sizes(void)
{
   int n[4];
   if (lattice[0] > 0) {
      for (int m = 0; m < 4; m++) n[m] = lattice[m];
   } else {
      error_root(synthetic, 1, "set_sizes [ms10.c]", "[Synthetic] needs a [Lattice] section");
      snprintf(cnfg_file, sizeof(cnfg_file), "%s/%sn%d", cnfg_dir, nbase, first);
      cnfg_sizes(cnfg_file, n);
   }
   set_lattice_sizes(n[0], n[1], n[2], n[3]);
}

static void read_wl_parms(void)
{
   read_wloop_parms();
   wl = wloop_parms();
   file_head.msl = wl.msl;
   file_head.wls = wl.wls;
   file_head.tmax = N0;
   file_head.wlmr = wl.mwlr;
   file_head.wlmt = wl.mwlt;
   if (append)
      check_wloop_parms(fdat);
   else
      write_wloop_parms(fdat);
}

static void read_infile(int argc, char *argv[])
{
   flog = freopen("STARTUP_ERROR", "w", stdout);
   int ifile = find_opt(argc, argv, "-i");
   endian = endianness();
   error_root((ifile == 0) || (ifile == (argc - 1)), 1, "read_infile [ms10.c]",
              "Syntax: ms10 -i <input file> [-noexp] [-a]");
   error_root(endian == UNKNOWN_ENDIAN, 1, "read_infile [ms10.c]", "Machine has unknown endianness");
   noexp = find_opt(argc, argv, "-noexp");
   append = find_opt(argc, argv, "-a");
   fin = freopen(argv[ifile + 1], "r", stdin);
   error_root(fin == NULL, 1, "read_infile [ms10.c]", "Unable to open input file");
   reset_input();

   read_extra();
   read_dirs();
   setup_files();
   fdat = fopen(par_file, append ? "rb" : "wb");
   error_root(fdat == NULL, 1, "read_infile [ms10.c]", "Unable to open parameter file");
   read_bc_parms();
   set_sizes();
   set_wloop_device(device);
   set_wloop_precision(precision);
   read_wl_parms();
   fclose(fin);
   fclose(fdat);
   if (append == 0) copy_file(par_file, par_save);
}

static void check_old_log(int *fst, int *lst, int *stp)
{
   int ie = 0, ic = 0, isv = 0, fc = 0, lc = 0, dc = 0, pc = 0;
   fend = fopen(log_file, "r");
   error_root(fend == NULL, 1, "check_old_log [ms10.c]", "Unable to open log file");
   while (fgets(line, NAME_SIZE, fend) != NULL) {
      if (strstr(line, "process grid") != NULL) {
         int np[4], bp[4];
         if (sscanf(line, "%dx%dx%dx%d process grid, %dx%dx%dx%d", np, np + 1, np + 2, np + 3, bp,
                    bp + 1, bp + 2, bp + 3) != 8)
            ie |= 0x1;
      } else if (strstr(line, "fully processed") != NULL) {
         pc = lc;
         if (sscanf(line, "Configuration no %d", &lc) == 1) {
            ic += 1;
            isv = 1;
         } else
            ie |= 0x1;
         if (ic == 1)
            fc = lc;
         else if (ic == 2)
            dc = lc - fc;
         else if ((ic > 2) && (lc != (pc + dc)))
            ie |= 0x2;
      } else if (strstr(line, "Configuration no") != NULL)
         isv = 0;
   }
   fclose(fend);
   error_root((ie & 0x1) != 0x0, 1, "check_old_log [ms10.c]", "Incorrect read count");
   error_root((ie & 0x2) != 0x0, 1, "check_old_log [ms10.c]", "Configuration numbers are not equally spaced");
   error_root(isv == 0, 1, "check_old_log [ms10.c]", "Log file extends beyond the last configuration save");
   *fst = fc;
   *lst = lc;
   *stp = dc;
}

static void check_old_dat(int fst, int lst, int stp)
{
   int ie = 0, ic = 0, fc = 0, lc = 0, dc = 0, pc = 0;
   fdat = fopen(dat_file, "rb");
   error_root(fdat == NULL, 1, "check_old_dat [ms10.c]", "Unable to open data file");
   check_file_head();
   while (read_data() == 1) {
      pc = lc;
      lc = data.nc;
      ic += 1;
      if (ic == 1)
         fc = lc;
      else if (ic == 2)
         dc = lc - fc;
      else if ((ic > 2) && (lc != (pc + dc)))
         ie |= 0x1;
   }
   fclose(fdat);
   error_root(ic == 0, 1, "check_old_dat [ms10.c]", "No data records found");
   error_root((ie & 0x1) != 0x0, 1, "check_old_dat [ms10.c]", "Configuration numbers are not equally spaced");
   error_root((fst != fc) || (lst != lc) || (stp != dc), 1, "check_old_dat [ms10.c]",
              "Configuration range is not as reported in the log file");
}

static void check_files(void)
{
   if (append) {
      int fst, lst, stp;
      check_old_log(&fst, &lst, &stp);
      check_old_dat(fst, lst, stp);
      error_root((fst != lst) && (stp != step), 1, "check_files [ms10.c]",
                 "Continuation run:\nPrevious run had a different configuration separation");
      error_root(first != lst + step, 1, "check_files [ms10.c]",
                 "Continuation run:\nConfiguration range does not continue the previous one");
   } else {
      fin = fopen(log_file, "r");
      fdat = fopen(dat_file, "rb");
      error_root((fin != NULL) || (fdat != NULL), 1, "check_files [ms10.c]",
                 "Attempt to overwrite old *.log or *.dat file");
      fdat = fopen(dat_file, "wb");
      error_root(fdat == NULL, 1, "check_files [ms10.c]", "Unable to open data file");
      write_file_head();
      fclose(fdat);
   }
}

static void print_info(void)
{
   long ip = ftell(flog);
   fclose(flog);
   if (ip == 0L) remove("STARTUP_ERROR");
   flog = freopen(log_file, append ? "a" : "w", stdout);
   error_root(flog == NULL, 1, "print_info [ms10.c]", "Unable to open log file");
   printf("\n");
   if (append)
      printf("Continuation run\n\n");
   else {
      printf("Computation of Wilson loops\n");
      printf("--------------------------------------\n\n");
   }
   printf("Wloop version %s\n", wloop_RELEASE);
   printf("MI355X build: libwloop_amd (HIP/gfx950), %s arithmetic\n",
          precision ? "mixed FP32/FP64" : "FP64");
   printf("The machine is %s endian\n", endian == LITTLE_ENDIAN ? "little" : "big");
   if (synthetic)
      printf("Configurations are generated on the GPU (seed %lld + no, eps %.6g)\n\n", synth_seed, synth_eps);
   else
      printf("Configurations are read in exported file format\n\n");
   if (append == 0) {
      printf("%dx%dx%dx%d lattice, ", N0, N1, N2, N3);
      printf("%dx%dx%dx%d local lattice\n", L0, L1, L2, L3);
      printf("%dx%dx%dx%d process grid, ", NPROC0, NPROC1, NPROC2, NPROC3);
      printf("%dx%dx%dx%d process block size\n\n", NPROC0_BLK, NPROC1_BLK, NPROC2_BLK, NPROC3_BLK);
      if (bc == 0)
         printf("Open boundary conditions\n\n");
      else
         printf("Periodic boundary conditions\n\n");
      print_wloop_parms();
   }
   printf("Configurations no %d -> %d in steps of %d\n\n", first, last, step);
   fflush(flog);
}

static void set_data(int nc)
{
   data.nc = nc;
   hyp_links(wl.alpha_action[0], wl.alpha_action[1], wl.alpha_action[2], wl.proj, wl.proj_iter, 0);
   wloop_sum(1);
   wl = wloop_parms();
   memcpy(data.wl_mat, wl.wl_mat, sizeof(double) * (size_t)file_head.wls);
}

static void save_data(void)
{
   fdat = fopen(dat_file, "ab");
   error_root(fdat == NULL, 1, "save_data [ms10.c]", "Unable to open data file");
   write_data();
   fclose(fdat);
}

static void check_endflag(int *iend)
{
   fend = fopen(end_file, "r");
   if (fend != NULL) {
      fclose(fend);
      remove(end_file);
      *iend = 1;
      printf("End flag set, run stopped\n\n");
   } else
      *iend = 0;
}

int main(int argc, char *argv[])
{
   read_infile(argc, argv);
   data.wl_mat = malloc(sizeof(double) * (size_t)file_head.wls);
   error(data.wl_mat == NULL, 1, "main [ms10.c]", "Unable to allocate data arrays");
   check_files();
   print_info();
   geometry();

   int iend = 0;
   double wtavg = 0.0;
   for (int nc = first; (iend == 0) && (nc <= last); nc += step) {
      double wt1 = wtime();
      printf("Configuration no %d\n", nc);
      if (synthetic) {
         random_cnfg((unsigned long long)(synth_seed + nc), synth_eps);
      } else {
         snprintf(cnfg_file, sizeof(cnfg_file), "%s/%sn%d", cnfg_dir, nbase, nc);
         import_cnfg(cnfg_file);
      }
      set_data(nc);
      save_data();
      double wt2 = wtime();
      wtavg += (wt2 - wt1);
      error_chk();
      printf("Configuration no %d fully processed in %.2e sec ", nc, wt2 - wt1);
      printf("(average = %.2e sec)\n\n", wtavg / (double)((nc - first) / step + 1));
      fflush(flog);
      copy_file(log_file, log_save);
      copy_file(dat_file, dat_save);
      check_endflag(&iend);
   }
   error_chk();
   fflush(flog);
   copy_file(log_file, log_save);
   fclose(flog);
   free(data.wl_mat);
   return 0;
}
