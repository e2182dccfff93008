// Driver for the ScaLAPACK interposition tests (tests/test_interpose.py).
//
//   interpose_test <call>     call = pdgemm_ | pdgemm | PDGEMM_ | psgemm_ | pzgemm_
//
// Calls the given symbol once on a 2 x 2 x 2 problem (1 x 1 grid
// descriptors) and prints one line:
//   "stub <calls> <c0>"   the next definition (tests/cpp/stub_scalapack.cpp)
//                         took the call;
//   "throws <text>"       the call raised std::runtime_error;
//   "returned <c0>"       the call returned without reaching the stub.
// Which library serves the symbol is decided by how the executable was
// linked (or what was preloaded); the tests build several link orders.
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include <cosma/pxgemm.h>

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: interpose_test <call>\n");
        return 2;
    }
    const std::string call = argv[1];
    const int two = 2, one = 1, zero = 0;
    const int desc[9] = {1, zero, two, two, two, two, zero, zero, two};
    double a[4] = {1, 2, 3, 4}, b[4] = {1, 0, 0, 1}, c[4] = {0, 0, 0, 0};
    float af[8] = {1, 2, 3, 4, 5, 6, 7, 8}, bf[8] = {1, 0, 0, 1, 0, 0, 0, 0}, cf[8] = {0};
    const double alpha = 1.0, beta = 0.0;
    const float alphaf = 1.0f, betaf = 0.0f;
    double c0 = 0.0;
    try {
        if (call == "pdgemm_") {
            pdgemm_("N", "N", &two, &two, &two, &alpha, a, &one, &one, desc, b, &one, &one, desc, &beta, c, &one,
                    &one, desc);
            c0 = c[0];
        } else if (call == "pdgemm") {
            pdgemm("N", "N", &two, &two, &two, &alpha, a, &one, &one, desc, b, &one, &one, desc, &beta, c, &one,
                   &one, desc);
            c0 = c[0];
        } else if (call == "PDGEMM_") {
            PDGEMM_("N", "N", &two, &two, &two, &alpha, a, &one, &one, desc, b, &one, &one, desc, &beta, c, &one,
                    &one, desc);
            c0 = c[0];
        } else if (call == "psgemm_") {
            psgemm_("N", "N", &two, &two, &two, &alphaf, af, &one, &one, desc, bf, &one, &one, desc, &betaf, cf,
                    &one, &one, desc);
            c0 = cf[0];
        } else if (call == "pzgemm_") {
            pzgemm_("N", "N", &two, &two, &two, &alpha, a, &one, &one, desc, b, &one, &one, desc, &beta, c, &one,
                    &one, desc);
            c0 = c[0];
        } else {
            std::fprintf(stderr, "unknown call %s\n", call.c_str());
            return 2;
        }
    } catch (const std::runtime_error &e) {
        std::printf("throws %s\n", e.what());
        return 0;
    }
    const int *calls = static_cast<const int *>(dlsym(RTLD_DEFAULT, "stub_scalapack_calls"));
    if (calls && *calls > 0) std::printf("stub %d %g\n", *calls, c0);
    else std::printf("returned %g\n", c0);
    return 0;
}
