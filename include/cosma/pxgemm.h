// ScaLAPACK p?gemm entry points served by cosma_amd (library
// libcosma_amd_pxgemm.so, which interposes on ScaLAPACK's symbols).
//
// Same symbols and argument lists as src/cosma/pxgemm.h:6-107: four spellings
// per type (plain, trailing underscore, upper case, upper case with
// underscore), 19 pointer arguments, descriptors of 9 ints
// (desc[1] context, [2..3] global size, [4..5] block size, [6..7] source
// process, [8] local leading dimension).  Complex types are passed as
// interleaved (re, im) pairs of float / double.
//
// p[d]gemm runs on the MI355X when the BLACS grid of the descriptors is 1 x 1
// (the only grid a single rank has).  Problems with min(m, n, k) below
// COSMA_DIM_THRESHOLD, and the s/c/z types, are handed to the next
// definition of the symbol in the process (ScaLAPACK's), as the reference
// does for small problems (src/cosma/pxgemm.cpp:8-17); without one they are an
// error.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
void psgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void pdgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void pcgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void pzgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void PSGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void PDGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void PCGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void PZGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void psgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void pdgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void pcgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void pzgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void PSGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void PDGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void PCGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void PZGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

#ifdef __cplusplus
}
#endif
