// Host build of the product's FP64 per-pair algebra (reml_math.h) for CPU tests.
#define FFS_HD
#include "../../fastfuncstuff_amd/csrc/reml_math.h"

extern "C" {
double rm_pair_coefficients(double a, double b, int T, const int32_t* gap, int white, double* s2, double* c,
                            double* alpha) {
    return remlk::pair_coefficients(a, b, T, gap, white, s2, c, alpha);
}
void rm_whiten_forward(int T, const double* s2, const double* c, const double* alpha, const int32_t* gap,
                       const double* y, double* w) {
    remlk::whiten_forward(T, s2, c, alpha, gap, y, 1, w, 1);
}
void rm_whiten_adjoint(int T, const double* s2, const double* c, const double* alpha, const int32_t* gap,
                       const double* u, double* g) {
    remlk::whiten_adjoint(T, s2, c, alpha, gap, u, 1, g, 1);
}
int rm_cholesky(int p, double* A) { return remlk::cholesky_inplace(p, A); }
void rm_cholesky_inverse(int p, const double* L, double* Ainv, double* tmp) {
    remlk::cholesky_inverse(p, L, Ainv, tmp);
}
int rm_gauss_jordan(int p, const double* A, double* inv, double* work) {
    return remlk::gauss_jordan_inverse(p, A, inv, work);
}
}
