// ScaLAPACK-style GEMM under COSMA's own prefix, in libcosma_amd.so.
//
// Same symbols as src/cosma/prefixed_pxgemm.h:11-91 (cosma_p?gemm and the
// trailing-underscore / upper-case spellings).  cosma_pdgemm runs on the
// MI355X (1 x 1 BLACS grid); problems below COSMA_DIM_THRESHOLD and the s/c/z
// types are passed to ScaLAPACK's pdgemm_ / psgemm_ / pcgemm_ / pzgemm_ when
// the process has it (src/cosma/prefixed_pxgemm.cpp:47-50), else they are an
// error.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
void cosma_psgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void cosma_pdgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void cosma_pcgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void cosma_pzgemm(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void cosma_psgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void cosma_pdgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void cosma_pcgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void cosma_pzgemm_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void COSMA_PSGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void COSMA_PDGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void COSMA_PCGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void COSMA_PZGEMM(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

void COSMA_PSGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void COSMA_PDGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);
void COSMA_PCGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const float *alpha, const float *a, const int *ia, const int *ja, const int *desca,
        const float *b, const int *ib, const int *jb, const int *descb, const float *beta,
        float *c, const int *ic, const int *jc, const int *descc);
void COSMA_PZGEMM_(const char *trans_a, const char *trans_b, const int *m, const int *n, const int *k,
        const double *alpha, const double *a, const int *ia, const int *ja, const int *desca,
        const double *b, const int *ib, const int *jb, const int *descb, const double *beta,
        double *c, const int *ic, const int *jc, const int *descc);

#ifdef __cplusplus
}
#endif
