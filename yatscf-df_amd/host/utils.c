#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "utils.h"
#include "yatscf_gpu.h"

void *ys_aligned_malloc(size_t bytes)
{
    void *p = NULL;
    if (bytes == 0) bytes = 64;
    if (posix_memalign(&p, 64, bytes) != 0) return NULL;
    return p;
}

void ys_aligned_free(void *p) { free(p); }

void copy_matrix_block(double *dst, const int ldd, double *src, const int lds, const int nrows, const int ncols)
{
    for (int r = 0; r < nrows; r++) memcpy(dst + (size_t) r * ldd, src + (size_t) r * lds, sizeof(double) * ncols);
}

double get_wtime_sec(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double) tv.tv_sec + (double) tv.tv_usec * 1e-6;
}

void print_mat(double *mat, const int ldm, const int nrows, const int ncols, const char *mat_name)
{
    printf("%s:\n", mat_name);
    for (int i = 0; i < nrows; i++) {
        for (int j = 0; j < ncols; j++) printf("% 10.6lf ", mat[(size_t) i * ldm + j]);
        printf("\n");
    }
}

void ys_check_gpu(int status, const char *what)
{
    if (status == YGPU_OK) return;
    fprintf(stderr, "TinySCF: %s failed (status %d): %s\n", what, status, ygpu_last_error());
    fflush(stderr);
    abort();
}
