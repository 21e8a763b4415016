// atan2 rounded to nearest as the hinge limit takes it (nativeengine_amd/
// csrc/ne_atan2.hpp, host build) beside the host C library's atan2:
//   as a program, reads "y x" hex-double pairs from stdin and prints both
//   results per line in hex;
//   as a shared library (-shared -DNE_ATAN2_LIB), atan2rn_batch fills both
//   for arrays.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>

#include "../../nativeengine_amd/csrc/ne_atan2.hpp"

extern "C" void atan2rn_batch(const double* y, const double* x, double* rn, double* lib, size_t n) {
    for (size_t k = 0; k < n; ++k) {
        rn[k] = nea::atan2_rn(y[k], x[k]);
        lib[k] = std::atan2(y[k], x[k]);
    }
}

#ifndef NE_ATAN2_LIB
int main() {
    char a[64], b[64];
    while (std::scanf("%63s %63s", a, b) == 2) {
        const double y = std::strtod(a, nullptr), x = std::strtod(b, nullptr);
        std::printf("%a %a\n", nea::atan2_rn(y, x), std::atan2(y, x));
    }
    return 0;
}
#endif
