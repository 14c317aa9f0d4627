// Harness that compiles the reference's host math header
// (VCL_Volume_Scanner/STAC/Util/_yooji_math_inlines.h, included from where it
// lies; the path is passed in REF_MATH_HEADER by the Makefile) into
// oracle/_ref/libref_math.so, so tests can compare the oracle's pose algebra
// with the reference's own FP32 code: d_cme_Compute_Matrix_Exponential
// (:115-194), d_im_Inverse_Matrix_4x4 (:404-465), d_mmm_Multiply_Matrix_Matrix
// (:229-246), d_niv_Norm_Infinity_Vector (:57-62).
// TEST INFRASTRUCTURE ONLY; the header needs -fpermissive (it calls helpers
// before declaring them, which MSVC 2013 accepted).
#include <cmath>
#include <cstdlib>
using std::abs;
using std::cos;
using std::sin;
using std::sqrt;
#include REF_MATH_HEADER

extern "C" {
void ref_expmap(const float x[6], float T[16]) {
    float in[6];
    for (int i = 0; i < 6; ++i) in[i] = x[i];
    d_cme_Compute_Matrix_Exponential<float>(in, T);
}
int ref_inverse4x4(const float A[16], float B[16]) {
    float in[16];
    for (int i = 0; i < 16; ++i) in[i] = A[i];
    return d_im_Inverse_Matrix_4x4(in, B) ? 1 : 0;
}
void ref_matmul4(const float A[16], const float B[16], float C[16]) {
    float a[16], b[16];
    for (int i = 0; i < 16; ++i) { a[i] = A[i]; b[i] = B[i]; }
    d_mmm_Multiply_Matrix_Matrix<float>(a, b, 4, 4, 4, C);
}
float ref_norm_inf(const float x[6]) {
    float in[6], n;
    for (int i = 0; i < 6; ++i) in[i] = x[i];
    d_niv_Norm_Infinity_Vector<float>(in, 6, n);
    return n;
}
}
