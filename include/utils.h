/* utils.h -- output helpers, same C API as the reference (reference include/utils.h:8-10). */
#ifndef UTILS_H_INCLUDED
#define UTILS_H_INCLUDED

#include "linearalg.h"

double randdouble(double min, double max);
/* Appends matrix A to <output_dir>/<title>-1-<count>.vtk as ASCII STRUCTURED_POINTS with %.6lf
 * values; count is one process-wide counter shared by all titles (src/utils.c:38-100). */
void printvtk(mtrx A, char *title, const char *output_dir);
void create_output_directory(const char *output_dir);

#endif
