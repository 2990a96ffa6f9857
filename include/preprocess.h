// C++ host mirror of the reference's preprocess entry points
// (src/preprocess/preprocess.h:21-37).  The NICSLU context of the reference
// (SNicsLU*) is replaced by a glu_analysis* the callee creates; free it with
// glu_analysis_destroy().  ax/ai/ap receive malloc'd copies of the
// permuted + scaled matrix (my_DumpA's output); the caller frees them.
// Return 0 on success, -1 on failure (the reference's convention).
#ifndef GLU_MIRROR_PREPROCESS_H
#define GLU_MIRROR_PREPROCESS_H

#include "glu_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

int preprocess(const char *matrixName, glu_analysis **an, double **ax, unsigned int **ai,
               unsigned int **ap);

int preprocess_from_arrays(unsigned int n, unsigned int nnz, const double *ax_in,
                           const unsigned int *ai_in, const unsigned int *ap_in, glu_analysis **an,
                           double **ax_out, unsigned int **ai_out, unsigned int **ap_out);

/* Column-sorted 1-based triplet file (optional '%' comment lines, then
 * "rows cols nnz") to 0-based CSC, the reader of
 * src/nicslu/util/nicslu_readfile.c:98-228.  Returns 0, -4 (cannot open) or
 * -7 (invalid). */
int glu_read_triplet_column(const char *file, unsigned int *n, unsigned int *nnz, double **ax,
                            unsigned int **ai, unsigned int **ap);

#ifdef __cplusplus
}
#endif

#endif
