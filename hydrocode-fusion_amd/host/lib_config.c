/*
 * lib_config.c -- config[] and its sources: command line, config.txt, defaults.
 *
 *   config[]        <- double config[N_CONF]  src/hydrocode_2D/hydrocode.c:123
 *   arg_preprocess  <- src/file_io/terminal_io.c:26-98
 *   configurate     <- src/file_io/config_handle.c:162-243 (config_read, config_check)
 *   config_write    <- src/file_io/config_handle.c:253-300  (log.dat)
 *
 * Precedence is the reference's: CLI "n=C" > config.txt > defaults; config.txt
 * only fills entries still INFINITY.
 */
#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrocode_dropin.h"

double config[N_CONF];

void arg_preprocess(const int argc_least, const int argc, char *argv[], char *scheme)
{
    printf("\n");
    for (int k = 0; k < argc; k++)
        printf("%s ", argv[k]);
    printf("\n\nTEST:\n %s\n", argc > 1 ? argv[1] : "");
    if (argc < argc_least) {
        printf("Test Beginning: ARGuments Counter %d is less than %d\n", argc, argc_least);
        exit(4);
    }
    printf("Test Beginning: ARGuments Counter = %d\n", argc);
    printf("Order[_Scheme]: %s\n", argv[3]);
    errno = 0;
    const unsigned long order = strtoul(argv[3], &scheme, 10);
    if (*scheme == '_')
        scheme++;
    else if (*scheme != '\0' || errno == ERANGE) {
        printf("No order or Wrog scheme!\n");
        exit(4);
    }
    config[9] = (double)order;

    printf("Configurating:\n");
    for (int k = argc_least + 1; k < argc; k++) {
        char *end;
        errno = 0;
        const unsigned long j = strtoul(argv[k], &end, 10);
        if (errno == ERANGE || *end != '=' || j >= N_CONF) {
            printf("Configuration error in ARGument variable %d! ERROR before '='!\n", k);
            exit(4);
        }
        errno = 0;
        const double val = strtod(end + 1, &end);
        if (errno == ERANGE || *end != '\0') {
            printf("Configuration error in ARGument variable %d! ERROR after '='!\n", k);
            exit(4);
        }
        config[j] = val;
        printf("%3lu-th configuration: %g (ARGument)\n", j, val);
    }
}

/* defaults and range checks (config_handle.c:39-151) */
static void config_check(void)
{
    const int dim = (int)config[0];
    printf("  dimension\t= %d\n", dim);
    if (isfinite(config[1]) && config[1] >= 0.0) {
        config[5] = isfinite(config[5]) ? config[5] : (double)INT_MAX;
        printf("  total time\t= %g\n", config[1]);
    } else if (!isfinite(config[5])) {
        fprintf(stderr, "The total time or the maximum number of time steps must be setted properly!\n");
        exit(2);
    } else {
        config[1] = INFINITY;
        if (isfinite(config[16])) {
            printf("  total time\t= %g * %d = %g\n", config[16], (int)config[5], config[16] * (int)config[5]);
            printf("  delta_t\t= %g\n", config[16]);
        }
    }
    printf("  time step\t= %d\n", (int)config[5]);

    if (isinf(config[4]))
        config[4] = EPS;
    const double eps = config[4];
    if (eps < 0.0 || eps > 0.01) {
        fprintf(stderr, "eps(%f) should in (0, 0.01)!\n", eps);
        exit(2);
    }
    printf("  eps\t\t= %g\n", eps);

    if (isinf(config[6]))
        config[6] = 1.4;
    else if (config[6] < 1.0 + eps) {
        fprintf(stderr, "The constant of the perfect gas(%f) should be larger than 1.0!\n", config[6]);
        exit(2);
    }
    printf("  gamma\t\t= %g\n", config[6]);

    if (isinf(config[7]))
        config[7] = dim == 1 ? 0.9 : 0.45;
    else if (config[7] > 1.0 - eps) {
        fprintf(stderr, "The CFL number(%f) should be smaller than 1.0.\n", config[7]);
        exit(2);
    }
    printf("  CFL number\t= %g\n", config[7]);

    if (isinf(config[41]))
        config[41] = 1.9;
    else if (config[41] < -eps || config[41] > 2.0) {
        fprintf(stderr, "The parameter in minmod limiter(%f) should in [0, 2)!\n", config[41]);
        exit(2);
    }
    if (isinf(config[110]))
        config[110] = 0.72;
    else if (config[110] < eps) {
        fprintf(stderr, "The specific heat at constant volume(%f) should be larger than 0.0!\n", config[110]);
        exit(2);
    }
    static const struct { int idx; double val; } defaults[] = {
        {2, 1}, {8, 0}, {30, 0}, {31, 0}, {32, 1}, {33, 0}, {40, 0}, {42, 1},
        {53, 0}, {60, 0}, {61, 0}, {70, 0}, {210, 0}, {211, 0}, {212, 0}};
    for (size_t k = 0; k < sizeof defaults / sizeof defaults[0]; ++k)
        if (!isfinite(config[defaults[k].idx]))
            config[defaults[k].idx] = defaults[k].val;
    if (!isfinite(config[20]))
        config[20] = config[10];
}

/* "index value" lines; anything else is a comment (config_handle.c:162-200) */
static int config_read(FILE *fp)
{
    char line[200];
    int line_num = 1;
    while (fgets(line, sizeof line, fp) != NULL) {
        char *end;
        const int i = (int)strtoul(line, &end, 10);
        while (isspace((unsigned char)*end))
            end++;
        if (0 < i && i < N_CONF) {
            errno = 0;
            const double tmp = strtod(end, NULL);
            if (errno == ERANGE) {
                fprintf(stderr, "Value range error of %d-th configuration in line %d of configuration file!\n", i, line_num);
                return 1;
            } else if (isinf(config[i])) {
                config[i] = tmp;
                printf("%3d-th configuration: %g\n", i, tmp);
            } else if (fabs(config[i] - tmp) > EPS) {
                printf("%3d-th configuration is repeatedly assigned with %g and %g(abandon)!\n", i, config[i], tmp);
            }
        } else if (i != 0 || (*end != '#' && *end != '\0')) {
            fprintf(stderr, "Warning: unknown row occurrs in line %d of configuration file!\n", line_num);
        }
        line_num++;
    }
    if (ferror(fp)) {
        fprintf(stderr, "Read error occurrs in configuration file!\n");
        return 0;
    }
    return 1;
}

void configurate(const char *add_in)
{
    char add[FILENAME_MAX + 40];
    FILE *fp;
    snprintf(add, sizeof add, "%sconfig.txt", add_in);
    if ((fp = fopen(add, "r")) == NULL) {
        snprintf(add, sizeof add, "%sconfig.dat", add_in);
        fp = fopen(add, "r");
    }
    if (fp == NULL) {
        printf("Cannot open configuration data file!\n");
        perror(add_in);
        exit(1);
    }
    if (config_read(fp) == 0) {
        fclose(fp);
        exit(2);
    }
    fclose(fp);
    printf("Configurated:\n");
    config_check();
}

void config_write(const char *add_out, const double *cpu_time, const char *name)
{
    (void)cpu_time;
    char file[FILENAME_MAX + 40];
    const int dim = (int)config[0];
    snprintf(file, sizeof file, "%slog.dat", add_out);
    FILE *fp = fopen(file, "w");
    if (fp == NULL) {
        printf("Cannot open log output file!\n");
        exit(1);
    }
    fprintf(fp, "%s is initialized with %d grids.\n\n", name, (int)config[3]);
    fprintf(fp, "Configurated:\n");
    fprintf(fp, "dim\t\t= %d\n", dim);
    if (isfinite(config[1]))
        fprintf(fp, "t_all\t= %d\n", (int)config[1]);
    else if (isfinite(config[16]))
        fprintf(fp, "tau\t\t= %g\n", config[16]);
    fprintf(fp, "eps\t\t= %g\n", config[4]);
    fprintf(fp, "gamma\t= %g\n", config[6]);
    fprintf(fp, "CFL\t\t= %g\n", config[7]);
    fprintf(fp, "h\t\t= %g\n", config[10]);
    fprintf(fp, "bond\t= %d\n", (int)config[17]);
    if (dim == 2) {
        fprintf(fp, "h_y\t\t= %g\n", config[11]);
        fprintf(fp, "bond_y\t= %d\n", (int)config[18]);
    }
    fprintf(fp, "\nA total of %d time steps are computed.\n", (int)config[5]);
    fclose(fp);
}
