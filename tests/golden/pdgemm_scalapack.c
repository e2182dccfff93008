/*
 * Runs ScaLAPACK's pdgemm on a 1 x 1 BLACS grid for one case and writes the
 * whole local C array (lldc x cols doubles, column-major) to a file.
 * Fixture generator for tests/golden/pdgemm_*.npz (run by
 * make_pdgemm_golden.py in the build container, against the MKL ScaLAPACK
 * and MPICH found there); not part of the library or of any GPU run.
 *
 * Arguments: ta tb m n k alpha beta ia ja ib jb ic jc mb nb pad out
 *   Global arrays: A is (op=='N' ? m x k : k x m) plus ia-1+2 extra rows and
 *   ja-1+2 extra columns around the submatrix (same for B, C); lld = rows + pad.
 *   Array X (seed s) element i = u(splitmix64(s * 0x9E3779B97F4A7C15 + i)),
 *   u(x) = ((x >> 11) * 2^-53) * 2 - 1, over the whole lld x cols array;
 *   seeds 1, 2, 3 for A, B, C.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void Cblacs_get(int, int, int *);
void Cblacs_gridinit(int *, const char *, int, int);
void Cblacs_gridexit(int);
void pdgemm_(const char *, const char *, const int *, const int *, const int *, const double *, const double *,
             const int *, const int *, const int *, const double *, const int *, const int *, const int *,
             const double *, double *, const int *, const int *, const int *);

static uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double *make(int64_t n, uint64_t seed) {
    double *x = malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        x[i] = (double)(splitmix64(seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    return x;
}

int main(int argc, char **argv) {
    if (argc != 18) {
        fprintf(stderr, "usage: ta tb m n k alpha beta ia ja ib jb ic jc mb nb pad out\n");
        return 2;
    }
    MPI_Init(&argc, &argv);
    const char ta = argv[1][0], tb = argv[2][0];
    const int m = atoi(argv[3]), n = atoi(argv[4]), k = atoi(argv[5]);
    const double alpha = atof(argv[6]), beta = atof(argv[7]);
    const int ia = atoi(argv[8]), ja = atoi(argv[9]), ib = atoi(argv[10]), jb = atoi(argv[11]);
    const int ic = atoi(argv[12]), jc = atoi(argv[13]), mb = atoi(argv[14]), nb = atoi(argv[15]);
    const int pad = atoi(argv[16]);
    const int na = (ta == 'N' || ta == 'n'), nbt = (tb == 'N' || tb == 'n');
    const int ar = (na ? m : k) + ia + 1, ac = (na ? k : m) + ja + 1;
    const int br = (nbt ? k : n) + ib + 1, bc = (nbt ? n : k) + jb + 1;
    const int cr = m + ic + 1, cc = n + jc + 1;
    const int lda = ar + pad, ldb = br + pad, ldc = cr + pad;
    double *A = make((int64_t)lda * ac, 1), *B = make((int64_t)ldb * bc, 2), *C = make((int64_t)ldc * cc, 3);
    int ctxt;
    Cblacs_get(-1, 0, &ctxt);
    Cblacs_gridinit(&ctxt, "C", 1, 1);
    const int desca[9] = {1, ctxt, ar, ac, mb, nb, 0, 0, lda};
    const int descb[9] = {1, ctxt, br, bc, mb, nb, 0, 0, ldb};
    const int descc[9] = {1, ctxt, cr, cc, mb, nb, 0, 0, ldc};
    pdgemm_(&ta, &tb, &m, &n, &k, &alpha, A, &ia, &ja, desca, B, &ib, &jb, descb, &beta, C, &ic, &jc, descc);
    FILE *f = fopen(argv[17], "wb");
    if (!f || fwrite(C, sizeof(double), (size_t)ldc * cc, f) != (size_t)ldc * cc) return 1;
    fclose(f);
    Cblacs_gridexit(ctxt);
    MPI_Finalize();
    return 0;
}
