/*
 * utils.c -- VTK output and helpers with the reference's C API (reference include/utils.h:8-10).
 * printvtk writes the reference's format (src/utils.c:38-100): ASCII STRUCTURED_POINTS, values
 * printed with %.6lf, one matrix row per line, appended to <dir>/<title>-1-<count>.vtk where
 * count is a single process-wide counter advanced by every call.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "utils.h"

double randdouble(double min, double max)
{
    return min + (double)rand() / ((double)RAND_MAX / (max - min));
}

void create_output_directory(const char *output_dir)
{
    struct stat st;
    if (stat(output_dir, &st) == 0) return;
    char path[512];
    snprintf(path, sizeof(path), "%s", output_dir);
    for (char *p = path + 1; *p; p++) {  /* mkdir -p */
        if (*p != '/') continue;
        *p = '\0';
        mkdir(path, 0755);
        *p = '/';
    }
    mkdir(path, 0755);
    printf("Created output directory: %s\n", output_dir);
}

void printvtk(mtrx A, char *title, const char *output_dir)
{
    static int count = 0;
    if (A.M == NULL) {
        printf("\n** Error: Aborting program **\n");
        exit(1);
    }
    if (A.m < 1 || A.n < 1) {
        printf("\n** Error: Invalid parameter **\n");
        exit(1);
    }
    create_output_directory(output_dir);
    char name[512];
    snprintf(name, sizeof(name), "%s/%s-1-%d.vtk", output_dir, title, count);
    FILE *pf = fopen(name, "a");
    if (!pf) {
        printf("\nError while opening file: %s\n", name);
        exit(1);
    }
    printf("%s\n", name);
    fprintf(pf, "# vtk DataFile Version 2.0\ntest\nASCII\nDATASET STRUCTURED_POINTS\n");
    fprintf(pf, "DIMENSIONS %d %d 1\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA %d\n", A.m, A.n, A.m * A.n);
    fprintf(pf, "SCALARS values float\nLOOKUP_TABLE default");
    for (int i = 0; i < A.m; i++) {
        fputc('\n', pf);
        for (int j = 0; j < A.n; j++) fprintf(pf, j ? " %.6lf" : "%.6lf", A.M[i][j]);
    }
    fclose(pf);
    count++;
}
