#include <setjmp.h>
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

static jmp_buf *ys_jmp = NULL;
static char ys_msg[1536];

const char *ys_error_message(void) { return ys_msg; }

/* failure handler of the device library's ygpu_fail (a failed call inside a drop-in TinySCF
 * function): back to the installed jump target with the status */
static void ys_failure_jump(int status, const char *where, const char *message)
{
    snprintf(ys_msg, sizeof(ys_msg), "%s failed (status %d): %s", where, status, message);
    if (ys_jmp) longjmp(*ys_jmp, status);
}

void ys_set_error_jump(void *jb)
{
    ys_jmp = (jmp_buf *) jb;
    (void) ygpu_set_failure_handler(jb ? ys_failure_jump : NULL);
}

/* A failed device call ends the current TinySCF function: with a jump target
 * installed (callers that handle errors, such as the Python bindings) control
 * returns there with the status; otherwise the process aborts with the
 * message, keeping the reference's void-function contract. */
void ys_check_gpu(int status, const char *what)
{
    if (status == YGPU_OK) return;
    snprintf(ys_msg, sizeof(ys_msg), "%s failed (status %d): %s", what, status, ygpu_last_error());
    if (ys_jmp) longjmp(*ys_jmp, status);
    fprintf(stderr, "TinySCF: %s\n", ys_msg);
    fflush(stderr);
    abort();
}
