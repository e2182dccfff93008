// libcosma_amd_pxgemm.so: ScaLAPACK's p?gemm symbols served by cosma_amd.
//
// Loaded before ScaLAPACK (linked first or LD_PRELOADed), these definitions
// take pdgemm calls; the next definition in the process (ScaLAPACK's, found
// with dlsym(RTLD_NEXT)) takes what cosma_amd does not run: problems with
// min(m, n, k) below COSMA_DIM_THRESHOLD and the s/c/z types
// (src/cosma/pxgemm.cpp:8-17 and its INTERPOSE_C_VOID definitions).  The
// symbols are strong here; the reference makes them weak aliases so that a
// static ScaLAPACK can win at link time instead.
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <dlfcn.h>

#include <string>

#include <cosma/cosma_pxgemm.hpp>
#include <cosma/pxgemm.h>

#include "front.hpp"

namespace {

#define D_ARGS                                                                                                \
    const char *ta, const char *tb, const int *m, const int *n, const int *k, const double *alpha,            \
        const double *a, const int *ia, const int *ja, const int *desca, const double *b, const int *ib,      \
        const int *jb, const int *descb, const double *beta, double *c, const int *ic, const int *jc,         \
        const int *descc
#define S_ARGS                                                                                                \
    const char *ta, const char *tb, const int *m, const int *n, const int *k, const float *alpha,             \
        const float *a, const int *ia, const int *ja, const int *desca, const float *b, const int *ib,        \
        const int *jb, const int *descb, const float *beta, float *c, const int *ic, const int *jc,           \
        const int *descc
#define PASS ta, tb, m, n, k, alpha, a, ia, ja, desca, b, ib, jb, descb, beta, c, ic, jc, descc

using d_fn = void (*)(D_ARGS);
using s_fn = void (*)(S_ARGS);

// Load address of this library, to recognise a lookup that lands in it.
const void *self_base() {
    static const void *base = [] {
        Dl_info info{};
        return dladdr(reinterpret_cast<const void *>(&self_base), &info) ? info.dli_fbase : nullptr;
    }();
    return base;
}

// The definition of `name` after this library in the lookup order, or
// nullptr.  dlsym(RTLD_NEXT) must run here: it searches the objects after
// the one that calls it, so called from libcosma_amd.so it would start the
// search in the wrong place (and could return these very functions).
template <typename F>
F next_symbol(const char *name) {
    void *p = dlsym(RTLD_NEXT, name);
    if (p) {
        Dl_info info{};
        if (dladdr(p, &info) && info.dli_fbase == self_base()) p = nullptr;
    }
    return reinterpret_cast<F>(p);
}

void do_pdgemm(const char *name, D_ARGS) {
    if (cosma::is_problem_too_small(*m, *n, *k)) {
        if (auto f = next_symbol<d_fn>(name)) {
            f(PASS);
            return;
        }
    }
    cosma::pxgemm<double>(*ta, *tb, *m, *n, *k, *alpha, a, *ia, *ja, desca, b, *ib, *jb, descb, *beta, c, *ic, *jc,
                          descc);
}

template <typename F, typename... Args>
void pass_on(const char *name, Args... args) {
    if (auto f = next_symbol<F>(name)) {
        f(args...);
        return;
    }
    cosma::detail::fail(std::string(name) + ": cosma_amd implements double precision only and no ScaLAPACK " +
                        name + " follows it in the process");
}

}  // namespace

extern "C" {

void pdgemm(D_ARGS) { do_pdgemm("pdgemm", PASS); }
void pdgemm_(D_ARGS) { do_pdgemm("pdgemm_", PASS); }
void PDGEMM(D_ARGS) { do_pdgemm("PDGEMM", PASS); }
void PDGEMM_(D_ARGS) { do_pdgemm("PDGEMM_", PASS); }

void psgemm(S_ARGS) { pass_on<s_fn>("psgemm", PASS); }
void psgemm_(S_ARGS) { pass_on<s_fn>("psgemm_", PASS); }
void PSGEMM(S_ARGS) { pass_on<s_fn>("PSGEMM", PASS); }
void PSGEMM_(S_ARGS) { pass_on<s_fn>("PSGEMM_", PASS); }

void pcgemm(S_ARGS) { pass_on<s_fn>("pcgemm", PASS); }
void pcgemm_(S_ARGS) { pass_on<s_fn>("pcgemm_", PASS); }
void PCGEMM(S_ARGS) { pass_on<s_fn>("PCGEMM", PASS); }
void PCGEMM_(S_ARGS) { pass_on<s_fn>("PCGEMM_", PASS); }

void pzgemm(D_ARGS) { pass_on<d_fn>("pzgemm", PASS); }
void pzgemm_(D_ARGS) { pass_on<d_fn>("pzgemm_", PASS); }
void PZGEMM(D_ARGS) { pass_on<d_fn>("PZGEMM", PASS); }
void PZGEMM_(D_ARGS) { pass_on<d_fn>("PZGEMM_", PASS); }

}  // extern "C"
