/*
 * oracle_math.cpp -- TEST INFRASTRUCTURE ONLY: host evaluation of the shared
 * math library (pct_math.h) with the oracle's compiler flags, so tests can
 * compare it with numpy/glibc (accuracy) and with the device (bit equality).
 */
#include <stdint.h>
#include "../pct_amd/csrc/pct_math.h"

extern "C" void oracle_math(int func, const double* in, const double* in2, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        double a = in[i], b = in2 ? in2[i] : 0.0, r = 0.0;
        switch (func) {
            case 0: r = pct_log(a); break;
            case 1: r = pct_sin(a); break;
            case 2: r = pct_cos(a); break;
            case 3: r = pct_tan(a); break;
            case 4: r = pct_atan(a); break;
            case 5: r = pct_atan2(a, b); break;
            case 6: r = (double)pct_sinf((float)a); break;
            case 7: r = (double)pct_cosf((float)a); break;
            case 8: r = (double)pct_atan2f((float)a, (float)b); break;
            case 9: r = pct_sqrt(a); break;
            case 10: r = a / b; break;
            case 11: r = a / b; break;      /* the device's shared-reciprocal quotient */
            case 12: r = pct_log(a); break; /* the device's log with that division */
            default: r = 0.0; break;
        }
        out[i] = r;
    }
}
