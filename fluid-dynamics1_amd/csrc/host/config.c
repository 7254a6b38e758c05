/*
 * config.c -- the reference's key = value configuration format (reference src/config.c:47-361).
 *
 * File format, as the reference reads it: lines starting with '#', ';' or a newline are skipped;
 * other lines are parsed as "<key>=<value>" with the value a single whitespace-free token (text
 * after it, such as a trailing comment, is ignored); keys and values are trimmed; numbers go
 * through atof/atoi (a malformed value becomes 0); an unknown key prints a warning and is
 * skipped.  Defaults are the reference's (src/config.c:47-87).  load_config_ext() reads the same
 * format and also accepts the keys this build adds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "config.h"

Config load_default_config(void)
{
    Config c;
    memset(&c, 0, sizeof(c));
    c.Re = 1000.0;
    c.Lx = 1;
    c.Ly = 1;
    c.nx = 64;
    c.ny = 64;
    c.dt = 0.005;
    c.tf = 20.0;
    c.max_co = 1.0;
    c.order = 6;
    c.poisson_max_it = 10000;
    c.poisson_tol = 1E-3;
    c.output_interval = 10;
    c.poisson_type = 2;
    c.openmp_enabled = 0;
    c.u4 = 1.0;
    return c;
}

static void trim(char *s)
{
    char *b = s;
    while (*b == ' ' || *b == '\t') b++;
    if (b != s) memmove(s, b, strlen(b) + 1);
    size_t n = strlen(s);
    while (n > 1 && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\n' || s[n - 1] == '\r')) s[--n] = '\0';
}

/* returns 1 if the key is one of the reference's keys (and stores it) */
static int base_key(Config *c, const char *k, const char *v)
{
    struct { const char *name; double *d; int *i; } keys[] = {
        {"Re", &c->Re, NULL}, {"Lx", NULL, &c->Lx}, {"Ly", NULL, &c->Ly}, {"nx", NULL, &c->nx},
        {"ny", NULL, &c->ny}, {"dt", &c->dt, NULL}, {"tf", &c->tf, NULL}, {"max_co", &c->max_co, NULL},
        {"order", NULL, &c->order}, {"poisson_max_it", NULL, &c->poisson_max_it},
        {"poisson_tol", &c->poisson_tol, NULL}, {"output_interval", NULL, &c->output_interval},
        {"poisson_type", NULL, &c->poisson_type}, {"openmp_enabled", NULL, &c->openmp_enabled},
        {"ui", &c->ui, NULL}, {"vi", &c->vi, NULL}, {"u1", &c->u1, NULL}, {"u2", &c->u2, NULL},
        {"u3", &c->u3, NULL}, {"u4", &c->u4, NULL}, {"v1", &c->v1, NULL}, {"v2", &c->v2, NULL},
        {"v3", &c->v3, NULL}, {"v4", &c->v4, NULL},
    };
    for (size_t q = 0; q < sizeof(keys) / sizeof(keys[0]); q++) {
        if (strcmp(k, keys[q].name) == 0) {
            if (keys[q].d) *keys[q].d = atof(v);
            else *keys[q].i = atoi(v);
            return 1;
        }
    }
    return 0;
}

static int ext_key(ConfigExt *x, const char *k, const char *v)
{
    if (strcmp(k, "sor_omega") == 0) { x->sor_omega = atof(v); x->has_sor_omega = 1; return 1; }
    if (strcmp(k, "nsteps") == 0) { x->nsteps = atoi(v); return 1; }
    if (strcmp(k, "precision") == 0) {
        x->precision = (strcmp(v, "fp32dc") == 0 || strcmp(v, "1") == 0) ? 1 : 0;
        return 1;
    }
    if (strcmp(k, "inner_rtol") == 0) { x->inner_rtol = atof(v); return 1; }
    if (strcmp(k, "sweeps_per_launch") == 0) { x->sweeps_per_launch = atoi(v); return 1; }
    if (strcmp(k, "diagnostics") == 0) { x->diagnostics = atoi(v); return 1; }
    if (strcmp(k, "device") == 0) { x->device = atoi(v); return 1; }
    if (strcmp(k, "use_graph") == 0) { x->use_graph = atoi(v); return 1; }
    if (strcmp(k, "warm_start") == 0) { x->warm_start = atoi(v); return 1; }
    if (strcmp(k, "output_async") == 0) { x->output_async = atoi(v); return 1; }
    if (strcmp(k, "transport_fp32") == 0) { x->transport_fp32 = atoi(v); return 1; }
    if (strcmp(k, "centerlines") == 0) { x->centerlines = atoi(v); return 1; }
    if (strcmp(k, "inner_beta") == 0) {
        /* -1 marks "reference": resolved with the grid size once the whole file is read */
        x->inner_beta = strcmp(v, "reference") == 0 ? -1.0 : atof(v);
        return 1;
    }
    return 0;
}

static int parse_file(const char *filename, Config *c, ConfigExt *x)
{
    FILE *f = fopen(filename, "r");
    if (!f) {
        printf("Error: Could not open configuration file '%s'\n", filename);
        printf("Using default configuration.\n");
        return 0;
    }
    printf("Loading configuration from: %s\n", filename);
    char line[256], key[64], value[64];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '\n' || line[0] == '#' || line[0] == ';') continue;
        if (sscanf(line, "%63[^=]=%63s", key, value) != 2) continue;
        trim(key);
        trim(value);
        if (base_key(c, key, value)) continue;
        if (x && ext_key(x, key, value)) continue;
        printf("Warning: Unknown configuration parameter '%s'\n", key);
    }
    fclose(f);
    printf("Configuration loaded successfully.\n");
    return 1;
}

Config load_config_from_file(const char *filename)
{
    Config c = load_default_config();
    parse_file(filename, &c, NULL);
    return c;
}

ConfigExt load_config_ext(const char *filename)
{
    ConfigExt x;
    memset(&x, 0, sizeof(x));
    x.base = load_default_config();
    x.inner_rtol = 1e-5;
    x.diagnostics = 1;
    x.output_async = 1;
    x.transport_fp32 = 1;
    if (filename) parse_file(filename, &x.base, &x);
    return x;
}

void print_config(const Config *c)
{
    printf("\n=== Simulation Configuration ===\n");
    printf("Physical Parameters:\n");
    printf("  Reynolds number (Re): %.2f\n", c->Re);
    printf("  Domain length (Lx): %d\n", c->Lx);
    printf("  Domain width (Ly): %d\n", c->Ly);
    printf("\nNumerical Parameters:\n");
    printf("  Grid points x (nx): %d\n", c->nx);
    printf("  Grid points y (ny): %d\n", c->ny);
    printf("  Time step (dt): %.6f\n", c->dt);
    printf("  Final time (tf): %.2f\n", c->tf);
    printf("  Max Courant number: %.2f\n", c->max_co);
    printf("  Finite difference order: %d\n", c->order);
    printf("  Poisson max iterations: %d\n", c->poisson_max_it);
    printf("  Poisson tolerance: %.2E\n", c->poisson_tol);
    printf("  Output interval: %d\n", c->output_interval);
    printf("  Poisson solver type: %d\n", c->poisson_type);
    printf("\nBoundary Conditions:\n");
    printf("  Internal u field (ui): %.2f\n", c->ui);
    printf("  Internal v field (vi): %.2f\n", c->vi);
    printf("  u boundaries (u1,u2,u3,u4): %.2f, %.2f, %.2f, %.2f\n", c->u1, c->u2, c->u3, c->u4);
    printf("  v boundaries (v1,v2,v3,v4): %.2f, %.2f, %.2f, %.2f\n", c->v1, c->v2, c->v3, c->v4);
    const double pi = 3.14159265359;
    const double beta = 0.5 * (2 / (1 + sin(pi / (c->nx + 1))) + 2 / (1 + sin(pi / (c->ny + 1))));
    printf("\nDerived Parameters:\n");
    printf("  Grid spacing dx: %.6f\n", (double)c->Lx / c->nx);
    printf("  Grid spacing dy: %.6f\n", (double)c->Ly / c->ny);
    printf("  SOR parameter (beta): %.6f\n", beta);
    printf("  Maximum iterations: %d\n", (int)((c->tf / c->dt) - 1));
    printf("================================\n\n");
}

void print_usage(const char *program_name)
{
    printf("Usage: %s [config_file] [output_folder]\n", program_name);
    printf("\n  config_file    key = value file (defaults are used without one)\n");
    printf("  output_folder  results go to ./output/<output_folder>/ (default ./output/)\n");
    printf("\nEvery step runs on the GPU.  Keys added to the reference's set:\n");
    printf("  sor_omega, nsteps, precision (fp64 | fp32dc), inner_rtol, inner_beta,\n");
    printf("  sweeps_per_launch, diagnostics, device, use_graph, warm_start, output_async,\n");
    printf("  transport_fp32, centerlines\n");
    printf("\nFor the list of parameters, run with --help-config\n");
}

void print_openmp_status(const Config *c)
{
    printf("\n=== Execution ===\n");
    printf("Configuration openmp_enabled = %d is accepted and ignored: the step runs on the GPU.\n", c->openmp_enabled);
    printf("=================\n\n");
}
