/* config.h -- simulation configuration, same struct layout and functions as the reference
 * (reference include/config.h:4-43), plus an extended loader for the keys this build adds. */
#ifndef CONFIG_H
#define CONFIG_H

typedef struct {
    double Re;
    int Lx;
    int Ly;
    int nx;
    int ny;
    double dt;
    double tf;
    double max_co;
    int order;
    int poisson_max_it;
    double poisson_tol;
    int output_interval;
    int poisson_type;     /* 1 = Gauss-Seidel, 2 = SOR */
    int openmp_enabled;
    double ui, vi;
    double u1, u2, u3, u4;
    double v1, v2, v3, v4;
} Config;

Config load_default_config(void);
Config load_config_from_file(const char *filename);  /* unknown keys warn and are ignored */
void print_config(const Config *config);
void print_usage(const char *program_name);
void print_openmp_status(const Config *config);

/* Keys this build adds (each optional).  The reference's loader warns about them and ignores
 * them, so one file drives both programs. */
typedef struct {
    Config base;
    double sor_omega;       /* sor_omega: SOR relaxation (default: the reference formula)      */
    int nsteps;             /* nsteps: number of time steps (default: (int)(tf/dt - 1) + 1)   */
    int precision;          /* precision = fp64 | fp32dc                                        */
    double inner_rtol;      /* inner_rtol: FP32 fold ratio of the defect-correction form       */
    int sweeps_per_launch;  /* sweeps_per_launch: SOR sweeps per kernel launch (0 = default)   */
    int diagnostics;        /* diagnostics: 1 = continuity max/min every step                  */
    int device;             /* device: HIP device ordinal                                        */
    int use_graph;          /* use_graph: 1 = replay each FP64 step from captured HIP graphs    */
    double inner_beta;      /* inner_beta: relaxation of the inner FP32 solves of fp32dc (0 =
                               sor_omega; "reference" = the reference's formula for nx, ny)     */
    int warm_start;         /* warm_start: 1 = each Poisson solve starts from the previous psi  */
    int output_async;       /* output_async: 1 = VTK files written by a second thread while the
                               next steps run (default); 0 = between steps, as the reference     */
    int transport_fp32;     /* transport_fp32: 1 (default) = fp32dc runs the transport and
                               velocity stencils in FP32; 0 = in FP64                            */
    int centerlines;        /* centerlines: 1 = at each output step also write the centreline
                               profiles u(x = 0.5, y) and v(x, y = 0.5) to
                               <output>/centerlines-<step>.csv (default 0: the reference's files
                               only)                                                             */
    int has_sor_omega;
} ConfigExt;

ConfigExt load_config_ext(const char *filename);     /* filename may be NULL: defaults */

#endif
