// ref_shim.cc -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Exposes the reference's own host-compilable helpers through a C ABI so that
// tests can compare the oracle's restatement with them bit for bit.  Nothing
// from the reference is copied: the recipe in oracle/Makefile compiles this file
// with -I pointing at gpucsl/pc/cuda/helpers of the reference checkout, under
// the reference's own -DGTEST host mode (tests/cuda/helper_test.cc:1 includes
// math_helpers.cu the same way).  Output goes to oracle/_ref/ only.
#include "math_helpers.cu"
#include "set_helpers.cu"
#include "gaussian_helpers.cu"

// gaussian_ci_test<level, n> (gaussian_helpers.cu:240-262) on a packed
// (level+2) x (level+2) correlation matrix, v_i = 0, v_j = 1, S = {2..level+1}.
template <int L> static double z_packed(const double *C) {
    double M[(L + 2) * (L + 2)];
    for (int t = 0; t < (L + 2) * (L + 2); ++t) M[t] = C[t];
    int S[L];
    for (int t = 0; t < L; ++t) S[t] = 2 + t;
    return gaussian_ci_test<L, L + 2>(0u, 1u, M, S);
}

extern "C" {

int ref_binomial(int n, int k) { return binomial_coefficient(n, k); }
double ref_poz(double z) { return poz(z); }
double ref_pchisq(double x, int df) { return pchisq(x, df); }

// calculate_sepset_indices<level> (set_helpers.cu:25-49) for level 1..5.
int ref_sepset_indices(int level, int neighbours_count, int index, int *out) {
    switch (level) {
    case 1: calculate_sepset_indices<1>(neighbours_count, index, out); return 0;
    case 2: calculate_sepset_indices<2>(neighbours_count, index, out); return 0;
    case 3: calculate_sepset_indices<3>(neighbours_count, index, out); return 0;
    case 4: calculate_sepset_indices<4>(neighbours_count, index, out); return 0;
    case 5: calculate_sepset_indices<5>(neighbours_count, index, out); return 0;
    default: return -1;
    }
}

double ref_gauss_z_packed(int level, const double *C) {
    switch (level) {
    case 1: return z_packed<1>(C);
    case 2: return z_packed<2>(C);
    case 3: return z_packed<3>(C);
    case 4: return z_packed<4>(C);
    case 5: return z_packed<5>(C);
    default: return -1.0;
    }
}

// compute_inverse_matrix<3>, cholesky_decomposition<3>, compute_m_2_inv<3>
// (gaussian_helpers.cu:123-218); the inputs are copied because the first
// overwrites its argument.
void ref_inverse3(const double *A, double *out) {
    double M[9];
    for (int t = 0; t < 9; ++t) M[t] = A[t];
    compute_inverse_matrix<3>(M, out);
}
void ref_cholesky3(const double *A, double *out) {
    double M[9];
    for (int t = 0; t < 9; ++t) M[t] = A[t];
    cholesky_decomposition<3>(M, out);
}
void ref_m2_inv3(const double *A, double *out) {
    double M[9];
    for (int t = 0; t < 9; ++t) M[t] = A[t];
    compute_m_2_inv<3>(M, out);
}
}
