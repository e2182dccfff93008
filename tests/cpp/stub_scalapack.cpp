// Stand-in for the next p?gemm definition in the process (ScaLAPACK's), used
// by the interposition tests: every p{s,d,c,z}gemm spelling counts its calls
// in stub_scalapack_calls and writes -7777 into the first element of C.
extern "C" {

int stub_scalapack_calls = 0;

#define STUB(NAME, T)                                                                                         \
    void NAME(const char *, const char *, const int *, const int *, const int *, const T *, const T *,        \
              const int *, const int *, const int *, const T *, const int *, const int *, const int *,        \
              const T *, T *c, const int *, const int *, const int *) {                                       \
        ++stub_scalapack_calls;                                                                               \
        c[0] = (T)-7777;                                                                                      \
    }
#define STUB4(lo, LO, T) STUB(lo, T) STUB(lo##_, T) STUB(LO, T) STUB(LO##_, T)

STUB4(pdgemm, PDGEMM, double)
STUB4(psgemm, PSGEMM, float)
STUB4(pcgemm, PCGEMM, float)
STUB4(pzgemm, PZGEMM, double)

}  // extern "C"
