/*
 * lib_io.c -- data_in / data_out of the 2D program.
 *
 *   example_io, flu_var_count, flu_var_count_line, flu_var_read, time_plot_read
 *                            <- src/file_io/io_control.c:40-335
 *   initialize_2D            <- src/file_io/file_2D_in.c:28-179
 *   file_2D_write, file_2D_write_POINT_TEC
 *                            <- src/file_io/file_2D_out.c:19-151
 *   DispPro, CreateDir       <- src/tools/sys_pro.c:37-100
 *
 * Formats: inputs are whitespace-separated numbers made only of [0-9.eE-]
 * (a '+' is illegal), one grid line (y index) per text line; outputs are
 * "%.10g\t" tables, frames separated by two blank lines, and Tecplot POINT
 * zones "FLU_VAR_<t+eps %.8g>.tec".
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "hydrocode_dropin.h"

#define N_MAX_2D 5        /* frames kept in memory (file_2D_in.c:17) */

static int number_char(int ch) { return ch == '-' || ch == '.' || ch == 'e' || ch == 'E' || isdigit(ch); }

void DispPro(const double pro, const int step)
{
    static double pro_print = 0.0;
    if (pro >= pro_print) {
        char bar[51];
        const long done = lround(pro / 2);
        for (int j = 0; j < 50; ++j)
            bar[j] = j < done ? '+' : '-';
        bar[50] = '\0';
        printf("\r%s  %6.2f%%   STEP=%-8d", bar, pro, step);
        fflush(stdout);
        pro_print += 0.1;
    }
}

int CreateDir(const char *path)
{
    if (access(path, W_OK) == 0)
        return -1;
    char tmp[FILENAME_MAX + 40];
    const size_t len = strlen(path);
    if (len >= sizeof tmp)
        return 1;
    for (size_t k = 1; k <= len; ++k) {
        if (path[k] == '/' || path[k] == '\0') {
            memcpy(tmp, path, k);
            tmp[k] = '\0';
            if (access(tmp, F_OK) != 0)
                mkdir(tmp, S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
        }
    }
    return access(path, W_OK) == 0 ? 0 : 1;
}

void example_io(const char *example, char *add_mkdir, const int i_or_o)
{
    static int announced = 0;
    const int dim = (int)config[0], el = (int)config[8], order = (int)config[9];
    const char *dname = dim == 1 ? "one-dim/" : dim == 2 ? "two-dim/" : dim == 3 ? "three-dim/" : NULL;
    if (!dname) {
        fprintf(stderr, "Strange computational dimension!\n");
        exit(2);
    }
    if (i_or_o == 0) {
        const char *frame = el == 0 ? "EUL_" : el == 1 ? "LAG_" : el == 2 ? "ALE_" : NULL;
        if (!frame) {
            fprintf(stderr, "Strange description method of fluid motion!\n");
            exit(2);
        }
        sprintf(add_mkdir, "../../data_out/%s%s%d_order/%s", dname, frame, order, example);
        if (CreateDir(add_mkdir) == 1) {
            fprintf(stderr, "Output directory '%s' construction failed!\n", add_mkdir);
            exit(1);
        } else if (!announced) {
            printf("Output directory '%s' has been constructed.\n", add_mkdir);
            announced = 1;
        }
    } else {
        sprintf(add_mkdir, "../../data_in/%s%s", dname, example);
        if (access(add_mkdir, R_OK) == -1) {
            fprintf(stderr, "Input directory '%s' is nonexistent or unreadable!\n", add_mkdir);
            exit(1);
        }
    }
    strcat(add_mkdir, "/");
}

int flu_var_count(FILE *fp, const char *add)
{
    int num = 0, in_number = 0, ch;
    while ((ch = getc(fp)) != EOF) {
        if (number_char(ch))
            in_number = 1;
        else if (!isspace(ch)) {
            fprintf(stderr, "Input contains illegal character(ASCII=%d, flag=%d) in the file '%s'!\n", ch, in_number, add);
            return 0;
        } else if (in_number) {
            num++;
            in_number = 0;
        }
    }
    rewind(fp);
    return num;
}

int flu_var_count_line(FILE *fp, const char *add, int *n_x)
{
    int line = 0, column = 0, in_number = 0, ch;
    do {
        ch = getc(fp);
        if (ch == '\n' || ch == EOF) {
            if (in_number)
                ++column;
            in_number = 0;
            if (column) {
                if (!line)
                    *n_x = column;
                else if (column != *n_x) {
                    printf("Error in input data file '%s', line=%d, column=%d, n_x=%d\n", add, line, column, *n_x);
                    return 0;
                }
                ++line;
                column = 0;
            }
        } else if (number_char(ch))
            in_number = 1;
        else if (!isspace(ch)) {
            printf("Input contains illigal character(ASCII=%d, flag=%d) in the file '%s', line=%d!\n", ch, in_number, add, line);
            return 0;
        } else if (in_number) {
            ++column;
            in_number = 0;
        }
    } while (ch != EOF);
    rewind(fp);
    return line;
}

/* Reads `num` numbers; a number is stored when the whitespace after it is
 * read (a final number with no trailing whitespace is not stored, as in the
 * reference).  Returns 0, or the 1-based index of the bad entry. */
int flu_var_read(FILE *fp, double *U, const int num)
{
    char number[100];
    int idx = 0, j = 0, ch;
    while ((ch = getc(fp)) != EOF) {
        if (isspace(ch) && idx) {
            number[idx] = '\0';
            idx = 0;
            char *end;
            errno = 0;
            const double val = strtod(number, &end);
            if (errno == ERANGE || *end != '\0') {
                printf("The %dth entry in the initial data file is not a double-precision floats.\n", j + 1);
                return j + 1;
            }
            if (j == num) {
                printf("Error on the initial data file reading!\n");
                return j;
            }
            U[j++] = val;
        } else if (number_char(ch) && idx < (int)sizeof number - 1) {
            number[idx++] = (char)ch;
        }
    }
    return 0;
}

static int compare_double(const void *a, const void *b)
{
    const double r = *(const double *)a - *(const double *)b;
    return r > 0.0 ? 1 : (r < 0.0 ? -1 : 0);
}

int time_plot_read(const char *add_in, const int N_max, int *N_plot, double **time_plot)
{
    char add[FILENAME_MAX + 40];
    FILE *fp;
    snprintf(add, sizeof add, "%stime_plot.txt", add_in);
    if ((fp = fopen(add, "r")) == NULL) {
        snprintf(add, sizeof add, "%stime_plot.dat", add_in);
        fp = fopen(add, "r");
    }
    if (fp == NULL) {
        printf("No time data file for plotting! Only the initial data and final result will be plotted.\n");
        *N_plot = 2;
    } else {
        *N_plot = flu_var_count(fp, add) + 2;
    }
    if (*N_plot < 2) {
        printf("Error in counting time data file for plotting!\n");
        if (fp) fclose(fp);
        exit(2);
    }
    *time_plot = (double *)malloc((size_t)(*N_plot) * sizeof(double));
    if (*time_plot == NULL) {
        printf("NOT enough memory! time_plot[]\n");
        exit(5);
    }
    (*time_plot)[0] = 0.0;
    (*time_plot)[*N_plot - 1] = config[1];
    if (fp) {
        if (flu_var_read(fp, *time_plot + 1, *N_plot - 2)) {
            fclose(fp);
            exit(2);
        }
        printf("Load time data file for plotting! Plot time step is %d.\n", *N_plot - 2);
        qsort(*time_plot, (size_t)(*N_plot - 1), sizeof(double), compare_double);
        fclose(fp);
    }
    return N_max < *N_plot ? N_max : *N_plot;
}

/* one initial-data file: count its lines/columns, check them against config[3], [13], [14], read it */
static double *read_field(const char *add, FILE *fp)
{
    int column = 0;
    const int line = flu_var_count_line(fp, add, &column);
    const int num = line * column;
    if (num < 1) {
        printf("Error in counting fluid variables in initial data file! %s\n", add);
        fclose(fp);
        exit(2);
    }
    if (isinf(config[3]))
        config[3] = (double)num;
    if (isinf(config[13]))
        config[13] = (double)column;
    if (isinf(config[14]))
        config[14] = (double)line;
    else if (num != (int)config[3] || column != (int)config[13] || line != (int)config[14]) {
        printf("Input unequal! %s\n", add);
        printf(" num=%d, num_cell=%d; column=%d, n_x=%d; line=%d, n_y=%d.\n", num, (int)config[3], column,
               (int)config[13], line, (int)config[14]);
        exit(2);
    }
    double *f = (double *)malloc((size_t)num * sizeof(double));
    if (f == NULL) {
        printf("NOT enough memory! %s\n", add);
        exit(5);
    }
    if (flu_var_read(fp, f, num)) {
        fclose(fp);
        exit(2);
    }
    return f;
}

static double *open_field(const char *add_in, const char *name)
{
    char add[FILENAME_MAX + 40];
    FILE *fp;
    snprintf(add, sizeof add, "%s%s.txt", add_in, name);
    if ((fp = fopen(add, "r")) == NULL) {
        snprintf(add, sizeof add, "%s%s.dat", add_in, name);
        fp = fopen(add, "r");
    }
    if (fp == NULL) {
        printf("Cannot open initial data file: %s!\n", name);
        exit(1);
    }
    double *f = read_field(add, fp);
    fclose(fp);
    return f;
}

struct flu_var initialize_2D(const char *name, int *N, int *N_plot, double **time_plot)
{
    struct flu_var FV0 = {NULL, NULL, NULL, NULL};
    char add_in[FILENAME_MAX + 40];
    example_io(name, add_in, 1);
    configurate(add_in);
    printf("  delta_x\t= %g\n", config[10]);
    printf("  delta_y\t= %g\n", config[11]);
    printf("  bondary_x\t= %d\n", (int)config[17]);
    printf("  bondary_y\t= %d\n", (int)config[18]);
    *N = time_plot_read(add_in, N_MAX_2D, N_plot, time_plot);
    FV0.RHO = open_field(add_in, "RHO");
    FV0.U = open_field(add_in, "U");
    FV0.V = open_field(add_in, "V");
    FV0.P = open_field(add_in, "P");
    printf("'%s' data initialized, line = %d, column = %d.\n", add_in, (int)config[14], (int)config[13]);
    return FV0;
}

static double cell_centre(double **A, int j, int i)
{
    return 0.25 * (A[j][i] + A[j][i + 1] + A[j + 1][i] + A[j + 1][i + 1]);
}

static void write_table(const char *add_out, const char *name, int n_x, int n_y, int N,
                        const struct cell_var_stru CV[], double **A, int which)
{
    char file[FILENAME_MAX + 40];
    snprintf(file, sizeof file, "%s%s.dat", add_out, name);
    FILE *fp = fopen(file, "w");
    if (fp == NULL) {
        printf("Cannot open solution output file: %s!\n", name);
        exit(1);
    }
    for (int k = 0; k < N; ++k) {
        for (int i = 0; i < n_y; ++i) {
            for (int j = 0; j < n_x; ++j) {
                double val;
                switch (which) {
                case 0: val = CV[k].RHO[j][i]; break;
                case 1: val = CV[k].U[j][i]; break;
                case 2: val = CV[k].V[j][i]; break;
                case 3: val = CV[k].P[j][i]; break;
                case 4: val = CV[k].E[j][i]; break;
                default: val = cell_centre(A, j, i); break;
                }
                fprintf(fp, "%.10g\t", val);
            }
            fprintf(fp, "\n");
        }
        fprintf(fp, "\n\n");
    }
    fclose(fp);
}

void file_2D_write(const int n_x, const int n_y, const int N, const struct cell_var_stru CV[],
                   double **X, double **Y, const double *cpu_time, const char *problem, const double time_plot[])
{
    char add_out[FILENAME_MAX + 40];
    example_io(problem, add_out, 0);
    static const char *names[5] = {"RHO", "U", "V", "P", "E"};
    for (int f = 0; f < 5; ++f)
        write_table(add_out, names[f], n_x, n_y, N, CV, NULL, f);
    write_table(add_out, "X", n_x, n_y, N, CV, X, 5);
    write_table(add_out, "Y", n_x, n_y, N, CV, Y, 5);

    char file[FILENAME_MAX + 40];
    snprintf(file, sizeof file, "%stime_plot.dat", add_out);
    FILE *fp = fopen(file, "w");
    if (fp == NULL) {
        printf("Cannot open solution output file: time_plot!\n");
        exit(1);
    }
    for (int k = 0; k < N; ++k)
        fprintf(fp, "%.10g\n", time_plot[k]);
    fclose(fp);
    config_write(add_out, cpu_time, problem);
}

void file_2D_write_POINT_TEC(const int n_x, const int n_y, const int N, const struct cell_var_stru CV[],
                             double **X, double **Y, const double *cpu_time, const char *problem,
                             const double time_plot[])
{
    (void)cpu_time;
    char add_out[FILENAME_MAX + 40], file[FILENAME_MAX + 80];
    example_io(problem, add_out, 0);
    snprintf(file, sizeof file, "%sFLU_VAR_%.8g.tec", add_out, time_plot[N - 1] + config[4]);
    FILE *fp = fopen(file, "w");
    if (fp == NULL) {
        fprintf(stderr, "Cannot open solution output TECPLOT file of '%s'!\n", problem);
        exit(1);
    }
    fprintf(fp, "TITLE = \"FE-Volume Point Data\"\n");
    fprintf(fp, "VARIABLES = \"X\", \"Y\", \"P\", \"RHO\", \"U\", \"V\", \"E\"\n");
    for (int k = 0; k < N; ++k) {
        fprintf(fp, "ZONE I=%d, J=%d, SOLUTIONTIME=%.10g, DATAPACKING=POINT\n", n_x, n_y, time_plot[k]);
        for (int i = 0; i < n_y; ++i)
            for (int j = 0; j < n_x; ++j)
                fprintf(fp, "%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t\n", cell_centre(X, j, i),
                        cell_centre(Y, j, i), CV[k].P[j][i], CV[k].RHO[j][i], CV[k].U[j][i], CV[k].V[j][i],
                        CV[k].E[j][i]);
        fprintf(fp, "\n");
    }
    fclose(fp);
}

/* ---- full-precision frames (no counterpart in the reference) ----------------
 * FLU_VAR.f64, little-endian IEEE doubles: "HGRPF64\0", int32 version (1),
 * n_x, n_y, N; time_plot[N]; then for each frame k the fields RHO U V P E,
 * each n_x*n_y doubles in CV order (j = x index outer, i = y index inner). */
static const char F64_MAGIC[8] = {'H', 'G', 'R', 'P', 'F', '6', '4', '\0'};

int file_2D_write_f64(const char *file, const int n_x, const int n_y, const int N, const struct cell_var_stru CV[],
                      const double time_plot[])
{
    FILE *fp = fopen(file, "wb");
    if (fp == NULL) {
        printf("Cannot open full-precision output file: %s\n", file);
        return 1;
    }
    const int head[4] = {1, n_x, n_y, N};
    int ok = fwrite(F64_MAGIC, 1, 8, fp) == 8 && fwrite(head, sizeof(int), 4, fp) == 4 &&
             fwrite(time_plot, sizeof(double), (size_t)N, fp) == (size_t)N;
    for (int k = 0; ok && k < N; ++k) {
        double **f[5] = {CV[k].RHO, CV[k].U, CV[k].V, CV[k].P, CV[k].E};
        for (int q = 0; ok && q < 5; ++q)
            for (int j = 0; ok && j < n_x; ++j)
                ok = fwrite(f[q][j], sizeof(double), (size_t)n_y, fp) == (size_t)n_y;
    }
    if (fclose(fp) != 0) ok = 0;
    if (!ok) printf("Write error in full-precision output file: %s\n", file);
    return ok ? 0 : 1;
}

int file_2D_read_f64(const char *file, int *n_x, int *n_y, int *N, double **time_plot, double **frames)
{
    FILE *fp = fopen(file, "rb");
    if (fp == NULL) return 1;
    char magic[8];
    int head[4];
    int ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, F64_MAGIC, 8) == 0 && fread(head, sizeof(int), 4, fp) == 4 &&
             head[0] == 1 && head[1] > 0 && head[2] > 0 && head[3] > 0;
    double *tp = NULL, *fr = NULL;
    if (ok) {
        const size_t cells = (size_t)head[1] * (size_t)head[2];
        tp = (double *)malloc((size_t)head[3] * sizeof(double));
        fr = (double *)malloc((size_t)head[3] * 5 * cells * sizeof(double));
        ok = tp && fr && fread(tp, sizeof(double), (size_t)head[3], fp) == (size_t)head[3] &&
             fread(fr, sizeof(double), (size_t)head[3] * 5 * cells, fp) == (size_t)head[3] * 5 * cells;
    }
    fclose(fp);
    if (!ok) {
        free(tp);
        free(fr);
        printf("Not a readable full-precision frame file: %s\n", file);
        return 2;
    }
    *n_x = head[1]; *n_y = head[2]; *N = head[3];
    *time_plot = tp;
    *frames = fr;
    return 0;
}
