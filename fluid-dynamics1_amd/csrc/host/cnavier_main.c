/*
 * cnavier_main.c -- the cnavier driver on one MI355X.
 *
 * Same command line, configuration file, log file and VTK output as the reference program
 * (src/main.c:27-481), with the time loop (src/main.c:276-451) running device-resident through
 * the cnv C ABI: the state never leaves the GPU except for VTK output every output_interval
 * steps.  Log lines keep the reference's format so that logs can be diffed; the per-step time is
 * wall-clock time (the reference logs clock(), i.e. process CPU time).
 *
 *   cnavier [config_file] [output_folder]      cnavier --help | --help-config
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "cnv.h"
#include "config.h"
#include "linearalg.h"
#include "poisson.h"
#include "utils.h"

static void logf_(FILE *f, const char *fmt, ...)
{
    if (!f) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
    fflush(f);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void help_config(void)
{
    printf("Complete list of configuration parameters:\n\n");
    printf("Physical:   Re (1000.0), Lx (1), Ly (1)\n");
    printf("Numerical:  nx (64), ny (64), dt (0.005), tf (20.0), max_co (1.0), order (6: 2, 4 or 6),\n");
    printf("            poisson_max_it (10000), poisson_tol (1E-3), output_interval (10),\n");
    printf("            poisson_type (2: 1 = Gauss-Seidel, 2 = SOR), openmp_enabled (ignored)\n");
    printf("Boundary:   ui, vi (interior start, 0.0); u1/v1 column 0, u2/v2 last column,\n");
    printf("            u3/v3 row 0, u4/v4 last row (u4 = 1.0 is the moving lid)\n");
    printf("Added keys: sor_omega (SOR relaxation; default 0.5*(2/(1+sin(PI/(nx+1)))+2/(1+sin(PI/(ny+1))))),\n");
    printf("            nsteps (default (int)(tf/dt-1)+1), precision (fp64 | fp32dc), inner_rtol (1e-5),\n");
    printf("            inner_beta (fp32dc inner relaxation: 0 = sor_omega, or \"reference\"),\n");
    printf("            sweeps_per_launch (0 = tuned default), diagnostics (1), device (0), use_graph (0),\n");
    printf("            warm_start (0; 1 = each Poisson solve starts from the previous psi, not 0),\n");
    printf("            output_async (1 = VTK files written by a second thread while the next steps run),\n");
    printf("            transport_fp32 (1 = fp32dc runs the transport and velocity stencils in FP32;\n");
    printf("            0 = in FP64)\n");
}

static mtrx to_mtrx(const double *a, int n)
{
    mtrx A = initm(n, n);
    for (int i = 0; i < n; i++) memcpy(A.M[i], a + (size_t)i * n, sizeof(double) * n);
    return A;
}

/* The four VTK files of one output step, in the reference's order (src/main.c:429-436). */
/* The centreline profiles of step `step`: u(x = 0.5, y) is column c = n/2 of u and v(x, y = 0.5)
 * row c of v (node j sits at x = j*dx, src/main.c:138-139, so x = 0.5 is node n/2 for even n;
 * the nearest node below it for odd n), one line per node k with its coordinate k*h, full
 * precision. */
static void write_centerlines(const double *u, const double *v, int n, double h, int step, const char *dir)
{
    char name[512];
    snprintf(name, sizeof(name), "%s/centerlines-%d.csv", dir, step);
    FILE *f = fopen(name, "w");
    if (!f) {
        printf("\nError while opening file: %s\n", name);
        return;
    }
    const int c = n / 2;
    fprintf(f, "# step %d, n %d: u(x=%.17g, y) at column %d, v(x, y=%.17g) at row %d\n", step, n, c * h, c,
            c * h, c);
    fprintf(f, "k,coord,u_center,v_center\n");
    for (int k = 0; k < n; k++)
        fprintf(f, "%d,%.17g,%.17g,%.17g\n", k, k * h, u[(size_t)k * n + c], v[(size_t)c * n + k]);
    fclose(f);
}

static void write_vtk_set(const double *w, const double *psi, const double *u, const double *v, int n,
                          const char *dir)
{
    const double *order[4] = {psi, w, u, v};
    char *titles[4] = {"stream-function", "vorticity", "x-velocity", "y-velocity"};
    for (int q = 0; q < 4; q++) {
        mtrx A = to_mtrx(order[q], n);
        printvtk(A, titles[q], dir);
        A.M = freem(A);
    }
}

/* output_async: one snapshot's files, written on a thread of its own while the next steps run
 * (the snapshot was queued by cnv_snapshot_begin; one writer at a time keeps the files' order) */
typedef struct {
    cnv_ctx *ctx;
    int n;
    const char *dir;
    int rc;
    pthread_t thread;
    int running;
    int centerlines;  /* also write the centreline profiles, of step `step`, spacing h */
    int step;
    double h;
} vtk_writer;

static void *vtk_writer_main(void *arg)
{
    vtk_writer *wr = (vtk_writer *)arg;
    const double *w, *psi, *u, *v;
    wr->rc = cnv_snapshot_wait(wr->ctx, &w, &psi, &u, &v);
    if (wr->rc == CNV_OK) {
        write_vtk_set(w, psi, u, v, wr->n, wr->dir);
        if (wr->centerlines) write_centerlines(u, v, wr->n, wr->h, wr->step, wr->dir);
    }
    return NULL;
}

/* waits for the writer of the previous snapshot, if any; returns its status */
static int vtk_writer_join(vtk_writer *wr)
{
    if (!wr->running) return CNV_OK;
    pthread_join(wr->thread, NULL);
    wr->running = 0;
    if (wr->rc != CNV_OK) printf("Error: %s\n", cnv_last_error(NULL));
    return wr->rc;
}

static int vtk_writer_start(vtk_writer *wr, int step)
{
    int rc = vtk_writer_join(wr);
    if (rc != CNV_OK) return rc;
    wr->step = step;  /* the previous writer, which read it, has been joined */
    rc = cnv_snapshot_begin(wr->ctx);
    if (rc != CNV_OK) {
        printf("Error: %s\n", cnv_last_error(wr->ctx));
        return rc;
    }
    if (pthread_create(&wr->thread, NULL, vtk_writer_main, wr) == 0) {
        wr->running = 1;
        return CNV_OK;
    }
    vtk_writer_main(wr);  /* no thread available: write the files here */
    if (wr->rc != CNV_OK) printf("Error: %s\n", cnv_last_error(NULL));
    return wr->rc;
}

int main(int argc, char *argv[])
{
    const double t_start = now_s();
    char output_dir[256] = "./output";
    const char *cfg_file = NULL;
    if (argc == 1) {
        printf("No configuration file specified. Using default values.\n");
    } else if (argc == 2) {
        if (strcmp(argv[1], "--help") == 0 || strcmp(argv[1], "-h") == 0) {
            print_usage(argv[0]);
            return 0;
        }
        if (strcmp(argv[1], "--help-config") == 0) {
            help_config();
            return 0;
        }
        cfg_file = argv[1];
    } else if (argc == 3) {
        cfg_file = argv[1];
        snprintf(output_dir, sizeof(output_dir), "./output/%s", argv[2]);
        printf("Using output directory: %s\n", output_dir);
    } else {
        printf("Error: Too many arguments.\n");
        print_usage(argv[0]);
        return 1;
    }
    ConfigExt cx = load_config_ext(cfg_file);
    Config cfg = cx.base;
    print_config(&cfg);
    print_openmp_status(&cfg);

    const int n = cfg.nx;
    if (cfg.nx != cfg.ny || n < 8) {
        printf("Error: the GPU step needs nx == ny >= 8 (got %d x %d)\n", cfg.nx, cfg.ny);
        return 1;
    }
    const double beta = cx.has_sor_omega ? cx.sor_omega
                                         : 0.5 * (2 / (1 + sin(PI / (cfg.nx + 1))) + 2 / (1 + sin(PI / (cfg.ny + 1))));
    printf("Poisson SOR parameter: %lf\n", beta);
    const double dx = (double)cfg.Lx / cfg.nx, dy = (double)cfg.Ly / cfg.ny;
    const int it_max = (int)((cfg.tf / cfg.dt) - 1);
    const int nsteps = cx.nsteps > 0 ? cx.nsteps : it_max + 1;
    /* Courant check of src/main.c:165-174 (it tests u1, as the reference does) */
    const double r1 = cfg.u1 * cfg.dt / dx, r2 = cfg.u1 * cfg.dt / dy;
    if (r1 > cfg.max_co || r2 > cfg.max_co) {
        printf("Unstable Solution!\n");
        printf("r1: %lf\n", r1);
        printf("r2: %lf\n", r2);
        return 1;
    }

    cnv_params p;
    cnv_default_params(&p, n);
    p.order = cfg.order;
    p.poisson_type = cfg.poisson_type;
    p.precision = cx.precision ? CNV_PREC_FP32_DC : CNV_PREC_FP64;
    p.itmax = cfg.poisson_max_it;
    p.device = cx.device;
    p.dx = dx;
    p.dy = dy;
    p.Re = cfg.Re;
    p.dt = cfg.dt;
    p.beta = beta;
    p.tol = cfg.poisson_tol;
    p.ui = cfg.ui;
    p.vi = cfg.vi;
    p.u1 = cfg.u1; p.u2 = cfg.u2; p.u3 = cfg.u3; p.u4 = cfg.u4;
    p.v1 = cfg.v1; p.v2 = cfg.v2; p.v3 = cfg.v3; p.v4 = cfg.v4;
    p.inner_rtol = cx.inner_rtol;
    p.inner_beta = cx.inner_beta < 0 ? cnv_reference_beta(cfg.nx, cfg.ny) : cx.inner_beta;
    p.sweeps_per_launch = cx.sweeps_per_launch;
    p.diagnostics = cx.diagnostics;
    p.use_graph = cx.use_graph;
    p.warm_start = cx.warm_start;
    p.transport_fp32 = cx.transport_fp32;
    cnv_ctx *ctx = NULL;
    int rc = cnv_create(&p, &ctx);
    if (rc != CNV_OK) {
        printf("Error: %s\n", cnv_last_error(NULL));
        return 1;
    }
    char dev[128] = "";
    cnv_device_name(p.device, dev, sizeof(dev));
    printf("Device: %s, %s\n", dev, p.precision ? "FP32 sweeps in FP64 defect correction" : "FP64");

    /* log file ./output/logs/<run>.txt (src/main.c:228-273) */
    char run_name[256] = "default", log_filename[1024];
    const char *prefix = "./output/";
    if (strstr(output_dir, prefix) && strlen(output_dir) > strlen(prefix))
        snprintf(run_name, sizeof(run_name), "%s", output_dir + strlen(prefix));
    snprintf(log_filename, sizeof(log_filename), "./output/logs/%s.txt", run_name);
    create_output_directory("./output/logs");
    FILE *log_file = fopen(log_filename, "w");
    if (!log_file) {
        printf("Warning: Could not create log file %s. Logging to console instead.\n", log_filename);
    } else {
        printf("Logging simulation progress to: %s\n", log_filename);
        fprintf(log_file, "=== Fluid Dynamics Simulation Log ===\n");
        fprintf(log_file, "Run name: %s\n", run_name);
        fprintf(log_file, "Output directory: %s\n", output_dir);
        fprintf(log_file, "Reynolds number: %.2f\n", cfg.Re);
        fprintf(log_file, "Grid size: %dx%d\n", cfg.nx, cfg.ny);
        fprintf(log_file, "Time step: %.6f\n", cfg.dt);
        fprintf(log_file, "Final time: %.2f\n", cfg.tf);
        fprintf(log_file, "Max iterations: %d\n", nsteps);
        fprintf(log_file, "OpenMP enabled: %s\n", cfg.openmp_enabled ? "Yes" : "No");
        fprintf(log_file, "=====================================\n\n");
        fflush(log_file);
    }

    const size_t cells = (size_t)n * n;
    double *host = NULL;  /* synchronous output only */
    if (!cx.output_async) {
        host = (double *)malloc(sizeof(double) * cells * 4);
        if (!host) {
            printf("** Error: insufficient memory **");
            return 1;
        }
    }
    vtk_writer writer = {ctx, n, output_dir, CNV_OK, 0, 0, cx.centerlines, 0, dx};
    const double t_sim = now_s();
    for (int t = 0; t < nsteps; t++) {
        const double t0 = now_s();
        cnv_step_stats st;
        rc = cnv_step(ctx, 1, &st);
        /* a non-finite Poisson update (CNV_DIVERGED) fails the way the reference's solve fails at
           itmax (src/poisson.c:280-284), with the same message */
        if (rc == CNV_NOT_CONVERGED || rc == CNV_DIVERGED) {
            logf_(log_file, "Error: maximum number of iterations achieved for Poisson equation.\n");
            printf("Error: maximum number of iterations achieved for Poisson equation.\n");
            vtk_writer_join(&writer);
            cnv_destroy(ctx);
            return 1;
        }
        if (rc != CNV_OK) {
            printf("Error: %s\n", cnv_last_error(ctx));
            vtk_writer_join(&writer);
            cnv_destroy(ctx);
            return 1;
        }
        const double t1 = now_s();
        const double iter_time = t1 - t0, elapsed = t1 - t_sim;
        const double remaining = elapsed / (t + 1) * (it_max - t);
        logf_(log_file, "Poisson equation solved with %d iterations - root-sum-of-squares error: %E\n", st.k, st.e);
        logf_(log_file, "Iteration: %d | ", t);
        logf_(log_file, "Time: %lf | ", (double)t * cfg.dt);
        logf_(log_file, "Progress: %.2lf%% | ", (double)100 * t / it_max);
        logf_(log_file, "Iter time: %.3f s\n", iter_time);
        logf_(log_file, "Continuity max: %E | ", st.cont_max);
        logf_(log_file, "Continuity min: %E | ", st.cont_min);
        logf_(log_file, "Elapsed: %.1f s | ", elapsed);
        if (t > 0)
            logf_(log_file, "Est. remaining: %.1f s\n", remaining);
        else
            logf_(log_file, "Est. remaining: -- s\n");
        if (cfg.output_interval > 0 && t % cfg.output_interval == 0) {
            if (cx.output_async) {
                if (vtk_writer_start(&writer, t) != CNV_OK) {
                    vtk_writer_join(&writer);
                    cnv_destroy(ctx);
                    return 1;
                }
            } else {
                double *w = host, *psi = host + cells, *u = host + 2 * cells, *v = host + 3 * cells;
                if (cnv_get_state(ctx, w, psi, u, v) != CNV_OK) {
                    printf("Error: %s\n", cnv_last_error(ctx));
                    cnv_destroy(ctx);
                    return 1;
                }
                write_vtk_set(w, psi, u, v, n, output_dir);
                if (cx.centerlines) write_centerlines(u, v, n, dx, t, output_dir);
            }
        }
    }
    const int wrc = vtk_writer_join(&writer);
    cnv_destroy(ctx);
    free(host);
    if (wrc != CNV_OK) return 1;
    printf("Simulation complete!\n");
    const double t_end = now_s();
    const double total_sim = t_end - t_sim;
    logf_(log_file, "\n=== Timing Summary ===\n");
    logf_(log_file, "Setup time: %.4f seconds\n", t_sim - t_start);
    logf_(log_file, "Simulation time: %.2f seconds\n", total_sim);
    logf_(log_file, "Total program time: %.2f seconds\n", t_end - t_start);
    logf_(log_file, "Average iteration time: %.4f seconds\n", total_sim / nsteps);
    logf_(log_file, "Iterations completed: %d\n", nsteps);
    logf_(log_file, "======================\n");
    if (log_file) {
        fclose(log_file);
        printf("Log saved to: %s\n", log_filename);
    }
    return 0;
}
