#ifndef _YATSCF_UTILS_H_
#define _YATSCF_UTILS_H_

#include <stddef.h>

/* 64-byte aligned allocation, as the reference (utils.h:6-7) */
void *ys_aligned_malloc(size_t bytes);
void  ys_aligned_free(void *p);
#define ALIGN64B_MALLOC(x) ys_aligned_malloc(x)
#define ALIGN64B_FREE(x)   ys_aligned_free(x)
#define DBL_SIZE           sizeof(double)
#define INT_SIZE           sizeof(int)

void copy_matrix_block(double *dst, const int ldd, double *src, const int lds, const int nrows, const int ncols);
double get_wtime_sec(void);
void print_mat(double *mat, const int ldm, const int nrows, const int ncols, const char *mat_name);

/* abort with the device library's message when a C-ABI call fails */
void ys_check_gpu(int status, const char *what);
void ys_set_error_jump(void *jmp_buf_ptr);
const char *ys_error_message(void);
#define YS_GPU(call) ys_check_gpu((call), #call)

#endif
