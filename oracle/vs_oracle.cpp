// vs_oracle.cpp — CPU restatement of the reference's fusion stages.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP pipeline (tests/, the
// smoke check, the CPU-baseline leg of the benchmark). Never linked into the
// product.
//
// Build with -ffp-contract=off: every FP32 expression below is evaluated in
// the reference's operation order, one IEEE rounding per operation. Values
// are carried in `R` (double for the oracle, float for the CPU baseline);
// index/mask-deciding expressions always use float.
//
// Semantics decisions shared with the HIP kernels (each applied identically):
//  Q0  index/mask maths in FP32, fixed op order, no contraction;
//      normalised(v) = v * (1.0f / sqrtf(dot(v,v))), dot = (x*x + y*y) + z*z.
//  Q1  bilateral taps are point samples with clamp-to-edge addressing.
//  Q2  Vector3f::toInt(residual) = truncation toward zero, residual = v - (float)(int)v.
//  Q3  the brick mask is cleared every frame; samples n = -4..1 with break.
//  Q4  integration follows g_ctv (projective z difference, ROUNDF, capped weight).
//  Q5  the raycast's "state == SEARCH_COARSE" no-op is reproduced (fine never
//      reverts to coarse); coarse step = 0.9*mu/voxel.
//  Q6  ROUNDF(v)/8 precedence quirk reproduced (negative branch not divided).
//  Q7  optional minimum-correspondence gate, default 0 = none (GPU reference).
//  Q8  exp(0) handled by the series branch.
//  Q10 weight cap is a parameter.
//  Q11 ROUNDF of a non-finite or |v| >= 2^31 value is "out of range".
//  Q12 the raycast cone test compares cosines: the ray is culled when
//      dot(rd, n_cc) < sqrt(1 - (r/d)^2) (no culling when r >= d), where the
//      reference compares acosf(dot) > asin(r/d).
//  Q13 the march is capped at total_step_max / min(step_scale, 1) + 64
//      iterations: the most steps a ray can take while every step moves it
//      forward by at least min(step_scale, 1) voxels, plus 64. Only a
//      fine-search ray that keeps stepping back (a -1.0 voxel behind the
//      surface) reaches it; the reference loop never ends for such a ray
//      (total_step sums signed step lengths, tsdf_renderer.cu:544-560).
//  Q14 bilinear taps that the reference skips (zero residual) contribute 0.
//  Q15 a Cholesky pivot that is not >= 1e-8 (including NaN) fails the solve.
//  Q16 Vector3f::toUChar() (ITMMath, not vendored) = ROUND half away from zero,
//      then clamp to [0, 255] (the ORUtils definition).
//  Q17 g_ctvc ignores the return of d_gii_Get_Interpolated_Intensity, so a
//      channel whose bilinear tap fails (outside the image or a participating
//      intensity <= 0) averages in a stale or uninitialised value; here that
//      channel keeps its stored value.
//  Q18 __float2int_rn of a non-finite or |v| >= 2^31 value is "out of range".
#include "vs_oracle.h"

// OR_MUTATION (0 = none) builds a deliberately wrong oracle for the pin
// tests' own check: each pin must fail on the mutant that breaks the function
// it pins (oracle/Makefile target `mutants`, tests/test_oracle_mutants.py).
//   1  A13: samples n = -1..1 (the CPU twin's range) instead of -4..1
//   2  A13: ROUNDF(v)/8 with the macro parenthesised (no precedence quirk)
//   3  A17: one of the 32 gradient taps read one voxel off along x
//   4  A18: shading factor 0.75 instead of 0.8
#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif
// OR_MARCH_CAP_MULT (1 = decision Q13) scales the march's iteration cap, to
// check on the test scenes that the cap decides no ray (a build with a cap
// 16x longer gives the same maps).
#ifndef OR_MARCH_CAP_MULT
#define OR_MARCH_CAP_MULT 1
#endif

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// FP32 helpers with the reference's operation order
// (STAC/Util/_yooji_cuda_math.h:205-302, _yooji_cuda_defines.h:53)
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
inline f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 scale(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(f3 a) { return sqrtf(dot(a, a)); }
inline f3 normalised(f3 v) {
    float n = 1.0f / sqrtf(dot(v, v));
    return {v.x * n, v.y * n, v.z * n};
}
// d_t_Transform (_yooji_cuda_math.h:229-235)
inline f3 transform(const float *T, f3 p) {
    return {T[0] * p.x + T[1] * p.y + T[2] * p.z + T[3], T[4] * p.x + T[5] * p.y + T[6] * p.z + T[7],
            T[8] * p.x + T[9] * p.y + T[10] * p.z + T[11]};
}
// d_r_Rotate (_yooji_cuda_math.h:238-244)
inline f3 rotate(const float *T, f3 p) {
    return {T[0] * p.x + T[1] * p.y + T[2] * p.z, T[4] * p.x + T[5] * p.y + T[6] * p.z,
            T[8] * p.x + T[9] * p.y + T[10] * p.z};
}
// d_bp_Back_Project (_yooji_cuda_math.h:212-218)
inline f3 back_project(float u, float v, const float *K, float d) {
    return {d * (u - K[2]) / K[0], d * (v - K[3]) / K[1], d};
}
// d_p_Project (_yooji_cuda_math.h:205-210)
inline void project(f3 p, const float *K, float &u, float &v) {
    u = K[0] * p.x / p.z + K[2];
    v = K[1] * p.y / p.z + K[3];
}
// ROUNDF (_yooji_cuda_defines.h:53): round half away from zero via truncation
// of v -+ 0.5f. Q11: non-finite or out-of-int-range -> false.
inline bool roundf_ref(float v, int &out) {
    if (!(v > -2147483520.0f && v < 2147483520.0f)) return false;
    out = (v < 0.0f) ? (int)(v - 0.5f) : (int)(v + 0.5f);
    return true;
}
// d_gcc_Get_Camera_Center (_yooji_cuda_math.h:247-262)
inline f3 camera_centre(const float *T) {
    float tx = T[3], ty = T[7], tz = T[11];
    return {-T[0] * tx - T[4] * ty - T[8] * tz, -T[1] * tx - T[5] * ty - T[9] * tz,
            -T[2] * tx - T[6] * ty - T[10] * tz};
}

struct Vol {
    int W, H, D;
    f3 origin;
    float vox_inv;
    float voxel;
    float mu, max_w;
    int BW, BH, BD;
};
Vol make_vol(const or_volume *v) {
    Vol o;
    o.W = v->dim[0];
    o.H = v->dim[1];
    o.D = v->dim[2];
    o.origin = {v->origin[0], v->origin[1], v->origin[2]};
    o.vox_inv = 1.0f / v->voxel;
    o.voxel = v->voxel;
    o.mu = v->mu;
    o.max_w = v->max_w;
    o.BW = (o.W + 7) / 8;
    o.BH = (o.H + 7) / 8;
    o.BD = (o.D + 7) / 8;
    return o;
}
// d_gpv_Get_Position_in_Voxel (_yooji_cuda_math.h:265-283)
inline bool world_to_voxel(const Vol &V, f3 p, f3 &vox) {
    vox.x = (p.x - V.origin.x) * V.vox_inv - 0.5f;
    vox.y = (p.y - V.origin.y) * V.vox_inv - 0.5f;
    vox.z = (p.z - V.origin.z) * V.vox_inv - 0.5f;
    return !(vox.x < -0.5f || vox.x >= (float)V.W - 0.5f || vox.y < -0.5f || vox.y >= (float)V.H - 0.5f ||
             vox.z < -0.5f || vox.z >= (float)V.D - 0.5f);
}
// d_gpw_Get_Position_in_World (_yooji_cuda_math.h:285-302); validity unused on the path
inline f3 voxel_to_world(const Vol &V, f3 vox) {
    return {(vox.x + 0.5f) / V.vox_inv + V.origin.x, (vox.y + 0.5f) / V.vox_inv + V.origin.y,
            (vox.z + 0.5f) / V.vox_inv + V.origin.z};
}

inline int nthreads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// A2-A3 bilateral (z_yooji_cuda_img_proc.cu:86-141, 275-286)
// ---------------------------------------------------------------------------
void gaussian_weights(float delta, int radius, float *g) {
    // updateGaussian: fGaussian[i] = expf(-(x*x) / (2*delta*delta)), x = i - radius
    for (int i = 0; i < 2 * radius + 1; ++i) {
        float x = (float)(i - radius);
        g[i] = expf(-(x * x) / (2 * delta * delta));
    }
}

template <class R>
void bilateral(const uint16_t *mm, int W, int H, float sigma_s, float sigma_r, int radius, R *out) {
    float g[16];
    gaussian_weights(sigma_s, radius, g);
    const R inv2s2 = R(1) / (R(2) * R(sigma_r) * R(sigma_r));
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const float c = (float)mm[y * W + x] * 0.001f;
            R sum = 0, t = 0;
            bool zero = false;
            for (int i = -radius; i <= radius && !zero; ++i) {
                int yy = std::min(std::max(y + i, 0), H - 1);
                for (int j = -radius; j <= radius; ++j) {
                    int xx = std::min(std::max(x + j, 0), W - 1);
                    float p = (float)mm[yy * W + xx] * 0.001f;
                    if (p <= 0.0f) { zero = true; break; }
                    R diff = R(p) - R(c);
                    R f = R(g[i + radius]) * R(g[j + radius]) * std::exp(-diff * diff * inv2s2);
                    t += f * diff;  // centred accumulation: sum w (d - c)
                    sum += f;
                }
            }
            out[y * W + x] = zero ? R(0) : R(c) + t / sum;
        }
    }
}

// ---------------------------------------------------------------------------
// A4 pyramid (z_yooji_image_processor.cpp:368-407, 567-596, 857-940)
// ---------------------------------------------------------------------------
template <class R>
void pyr_down(const R *in, int W, int H, R *out) {
    const int W2 = W / 2, H2 = H / 2;
#pragma omp parallel for schedule(static)
    for (int j2 = 0; j2 < H2; ++j2) {
        const int j = 2 * j2 + 1;
        for (int i2 = 0; i2 < W2; ++i2) {
            const int i = 2 * i2 + 1;
            R v = 0;
            if (i >= 2 && i <= W - 3 && j >= 2 && j <= H - 3) {
                R h[5];
                for (int r = -2; r <= 2; ++r) {
                    const R *row = in + (size_t)(j + r) * W;
                    h[r + 2] = R(1) * row[i - 2] + R(1) * row[i + 2] + R(4) * row[i - 1] + R(4) * row[i + 1] +
                               R(6) * row[i];
                }
                v = (R(1) * h[0] + R(1) * h[4] + R(4) * h[1] + R(4) * h[3] + R(6) * h[2]) / R(256);
            }
            out[(size_t)j2 * W2 + i2] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// A6-A7 vertex / normal maps (z_yooji_image_processor.cpp:119-227)
// ---------------------------------------------------------------------------
template <class R>
void vnmap(const float *depth, int W, int H, const float *K, int inplace, R *vtx, R *nrm) {
    // Vertices and every decision (validity, |dx| and |dy| thresholds, the
    // grazing test) in FP32; the normal's value in R.
    const size_t n = (size_t)W * H;
    std::vector<float> V(3 * n, -100.0f);
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t t = (size_t)i * W + j;
            float d = depth[t];
            if (d <= 0.0f) continue;
            V[3 * t + 0] = d * ((float)j - K[2]) / K[0];
            V[3 * t + 1] = d * ((float)i - K[3]) / K[1];
            V[3 * t + 2] = d;
        }
    std::vector<float> out(V);  // vertex output (in-place invalidation lands here)
    const float *P = inplace ? out.data() : V.data();
    std::vector<R> N(3 * n, R(-100));
    for (int i = 1; i < H - 1; ++i)
        for (int j = 1; j < W - 1; ++j) {
            size_t t = (size_t)i * W + j;
            const float *c = P + 3 * t, *x1 = P + 3 * (t - 1), *x2 = P + 3 * (t + 1);
            const float *y1 = P + 3 * (t - W), *y2 = P + 3 * (t + W);
            if (c[0] == -100.0f || x1[0] == -100.0f || x2[0] == -100.0f || y1[0] == -100.0f || y2[0] == -100.0f)
                continue;
            float dx = x2[0] - x1[0];
            if (std::fabs(dx) < 1.0e-7f) continue;
            float dy = y2[1] - y1[1];
            if (std::fabs(dy) < 1.0e-7f) continue;
            float gx = (x2[2] - x1[2]) / dx, gy = (y2[2] - y1[2]) / dy;
            float nn = sqrtf(gx * gx + gy * gy + 1.0f);
            float nz = -1.0f / nn;
            if (-nz <= 0.5f) {  // -n_z < cos(60 deg), compared in double by the reference
                for (int k = 0; k < 3; ++k) out[3 * t + k] = -100.0f;
            } else {
                R gxr = R(x2[2] - x1[2]) / R(dx), gyr = R(y2[2] - y1[2]) / R(dy);
                R nr = std::sqrt(gxr * gxr + gyr * gyr + R(1));
                N[3 * t + 0] = gxr / nr;
                N[3 * t + 1] = gyr / nr;
                N[3 * t + 2] = R(-1) / nr;
            }
        }
    for (size_t k = 0; k < 3 * n; ++k) {
        vtx[k] = out[k];
        nrm[k] = N[k];
    }
}

// ---------------------------------------------------------------------------
// A8-A9 ICP summand + sums (z_yooji_cuda_pose_tracker.cu:279-395, 548-697,
// bilinear d_giv/d_gin _yooji_cuda_math.h:365-481)
// ---------------------------------------------------------------------------
// Bilinear vertex taps; FP32 because the result gates the correspondence.
inline bool interp_vertex(float xf, float yf, int W, int H, const float *m, f3 &out) {
    // reference bounds test; written so that a NaN coordinate is rejected (Q11)
    if (!(xf >= 0.0f && xf <= float(W - 1) && yf >= 0.0f && yf <= float(H - 1))) return false;
    int x = (int)xf, y = (int)yf;
    float rx = xf - (float)x, ry = yf - (float)y;
    f3 v1{0, 0, 0}, v2{0, 0, 0}, v3{0, 0, 0}, v4{0, 0, 0};
    const float *p = m + 3 * ((size_t)y * W + x);
    if (p[0] == -100.0f) return false;
    v1 = {p[0], p[1], p[2]};
    if (rx > 0.0f) {
        p = m + 3 * ((size_t)y * W + x + 1);
        if (p[0] == -100.0f) return false;
        v2 = {p[0], p[1], p[2]};
    }
    if (ry > 0.0f) {
        p = m + 3 * ((size_t)(y + 1) * W + x);
        if (p[0] == -100.0f) return false;
        v3 = {p[0], p[1], p[2]};
    }
    if (rx > 0.0f && ry > 0.0f) {
        p = m + 3 * ((size_t)(y + 1) * W + x + 1);
        if (p[0] == -100.0f) return false;
        v4 = {p[0], p[1], p[2]};
    }
    float ix = 1.0f - rx, iy = 1.0f - ry;
    out = add(scale(iy, add(scale(ix, v1), scale(rx, v2))), scale(ry, add(scale(ix, v3), scale(rx, v4))));
    return true;
}
// Bilinear normal taps; validity in FP32, the interpolated value in R.
template <class R>
inline bool interp_normal(float xf, float yf, int W, int H, const float *m, R n[3]) {
    // reference bounds test; written so that a NaN coordinate is rejected (Q11)
    if (!(xf >= 0.0f && xf <= float(W - 1) && yf >= 0.0f && yf <= float(H - 1))) return false;
    int x = (int)xf, y = (int)yf;
    float rx = xf - (float)x, ry = yf - (float)y;
    R t[4][3] = {};
    const float *p = m + 3 * ((size_t)y * W + x);
    if (p[0] == -100.0f) return false;
    for (int k = 0; k < 3; ++k) t[0][k] = p[k];
    if (rx > 0.0f) {
        p = m + 3 * ((size_t)y * W + x + 1);
        if (p[0] == -100.0f) return false;
        for (int k = 0; k < 3; ++k) t[1][k] = p[k];
    }
    if (ry > 0.0f) {
        p = m + 3 * ((size_t)(y + 1) * W + x);
        if (p[0] == -100.0f) return false;
        for (int k = 0; k < 3; ++k) t[2][k] = p[k];
    }
    if (rx > 0.0f && ry > 0.0f) {
        p = m + 3 * ((size_t)(y + 1) * W + x + 1);
        if (p[0] == -100.0f) return false;
        for (int k = 0; k < 3; ++k) t[3][k] = p[k];
    }
    R ix = R(1) - R(rx), iy = R(1) - R(ry);
    for (int k = 0; k < 3; ++k)
        n[k] = iy * (ix * t[0][k] + R(rx) * t[1][k]) + R(ry) * (ix * t[2][k] + R(rx) * t[3][k]);
    R len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] /= len;
    return true;
}

// One pixel's summand; returns false when the pixel has no correspondence.
template <class R>
inline bool icp_summand(int x, int y, const float *d1, int W, int H, const float *K, const float *mv,
                        const float *mn, const float *T21, float dist2, R A[6], R &b) {
    float td1 = d1[(size_t)y * W + x];
    if (td1 < 1.0e-6f) return false;
    f3 tp = back_project((float)x, (float)y, K, td1);
    f3 p = transform(T21, tp);
    float u, v;
    project(p, K, u, v);
    f3 v0;
    if (!interp_vertex(u, v, W, H, mv, v0)) return false;
    float dx = v0.x - p.x, dy = v0.y - p.y, dz = v0.z - p.z;
    float td0 = dx * dx + dy * dy + dz * dz;
    if (td0 > dist2) return false;
    R n[3];
    if (!interp_normal<R>(u, v, W, H, mn, n)) return false;
    const R px = p.x, py = p.y, pz = p.z;
    A[0] = pz * n[1] + (-py) * n[2];
    A[1] = (-pz) * n[0] + px * n[2];
    A[2] = py * n[0] + (-px) * n[1];
    A[3] = -n[0];
    A[4] = -n[1];
    A[5] = -n[2];
    b = n[0] * (R(v0.x) - px) + n[1] * (R(v0.y) - py) + n[2] * (R(v0.z) - pz);
    return true;
}

// Rows are summed in parallel, row partials then added in row order, so the
// result does not depend on the thread count.
template <class R>
long long icp_system(const float *d1, int W, int H, const float *K, const float *mv, const float *mn,
                     const float *T21, float dist2, R ATA[36], R ATb[6], uint8_t *mask) {
    std::vector<R> rows((size_t)H * 42, R(0));
    std::vector<long long> cnt(H, 0);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        R *acc = &rows[(size_t)y * 42];
        for (int x = 0; x < W; ++x) {
            R A[6], b;
            bool ok = icp_summand<R>(x, y, d1, W, H, K, mv, mn, T21, dist2, A, b);
            if (mask) mask[(size_t)y * W + x] = ok ? 1 : 0;
            if (!ok) continue;
            cnt[y]++;
            for (int i = 0; i < 6; ++i) {
                for (int j = 0; j < 6; ++j) acc[i * 6 + j] += A[i] * A[j];
                acc[36 + i] += A[i] * b;
            }
        }
    }
    for (int k = 0; k < 36; ++k) ATA[k] = 0;
    for (int k = 0; k < 6; ++k) ATb[k] = 0;
    long long n = 0;
    for (int y = 0; y < H; ++y) {
        for (int k = 0; k < 36; ++k) ATA[k] += rows[(size_t)y * 42 + k];
        for (int k = 0; k < 6; ++k) ATb[k] += rows[(size_t)y * 42 + 36 + k];
        n += cnt[y];
    }
    return n;
}

// ---------------------------------------------------------------------------
// A10 Cholesky + solve (_yooji_cuda_math.h:489-592), A11 pose algebra
// ---------------------------------------------------------------------------
// rel > 0: also fail (status 8, the build's relative pivot guard) when a
// squared pivot L_jj^2 < rel * A_jj. Returns 0, 4 (absolute pivot) or 8.
int cholesky_status(const double *A, const double *b, double rel, double *x) {
    double L[36] = {0};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            if (i == j) {
                double s = 0;
                for (int k = 0; k < j; ++k) s += L[j * 6 + k] * L[j * 6 + k];
                L[j * 6 + j] = std::sqrt(A[j * 6 + j] - s);
            } else if (std::fabs(L[j * 6 + j]) >= 1e-8) {
                double s = 0;
                for (int k = 0; k < j; ++k) s += L[i * 6 + k] * L[j * 6 + k];
                L[i * 6 + j] = (A[i * 6 + j] - s) / L[j * 6 + j];
            } else {
                L[i * 6 + j] = 0.0;
            }
        }
    for (int i = 0; i < 6; ++i)
        if (!(std::fabs(L[i * 6 + i]) >= 1e-8)) return 4;
    if (rel > 0.0)
        for (int i = 0; i < 6; ++i)
            if (L[i * 6 + i] * L[i * 6 + i] < rel * A[i * 6 + i]) return 8;
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0;
        for (int k = 0; k < i; ++k) s += L[i * 6 + k] * y[k];
        y[i] = (b[i] - s) / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = 0;
        for (int k = i + 1; k < 6; ++k) s += L[k * 6 + i] * x[k];
        x[i] = (y[i] - s) / L[i * 6 + i];
    }
    return 0;
}
bool cholesky_solve(const double *A, const double *b, double *x) { return cholesky_status(A, b, 0.0, x) == 0; }

// SE(3) exponential of the twist x = [w; v] (Ma et al., "An Invitation to 3D
// Vision", eq. 2.30, as used by z_uitme_..._Matrix_Exponential,
// pose_tracker.cu:50-149): R = I + A[w]x + B[w]x^2, t = (I + B[w]x + C[w]x^2) v
// with A = sin(t)/t, B = (1-cos t)/t^2, C = (t - sin t)/t^3, t = |w|.
void expmap(const double *x, double *T) {
    const double w0 = x[0], w1 = x[1], w2 = x[2];
    const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = std::sqrt(th2);
    double A, B, C;
    if (th < 1e-4) {
        A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        C = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        A = std::sin(th) / th;
        B = (1.0 - std::cos(th)) / th2;
        C = (th - std::sin(th)) / (th2 * th);
    }
    const double W[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
    double W2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += W[r * 3 + k] * W[k * 3 + c];
            W2[r * 3 + c] = s;
        }
    for (int r = 0; r < 3; ++r) {
        double t = 0;
        for (int c = 0; c < 3; ++c) {
            double I = (r == c) ? 1.0 : 0.0;
            T[r * 4 + c] = I + A * W[r * 3 + c] + B * W2[r * 3 + c];
            t += (I + B * W[r * 3 + c] + C * W2[r * 3 + c]) * x[3 + c];
        }
        T[r * 4 + 3] = t;
    }
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

void rigid_inverse(const double *T, double *Ti) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ti[r * 4 + c] = T[c * 4 + r];
        Ti[r * 4 + 3] = -(T[0 * 4 + r] * T[3] + T[1 * 4 + r] * T[7] + T[2 * 4 + r] * T[11]);
    }
    Ti[12] = Ti[13] = Ti[14] = 0.0;
    Ti[15] = 1.0;
}

void matmul4(const double *A, const double *B, double *C) {
    double t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += A[r * 4 + k] * B[k * 4 + c];
            t[r * 4 + c] = s;
        }
    std::memcpy(C, t, sizeof(t));
}

void intrinsics_pyramid(const float *K0, int levels, float *K) {
    float fx = K0[0], fy = K0[1], px = K0[2], py = K0[3];
    for (int l = 0; l < levels; ++l) {
        K[4 * l + 0] = fx;
        K[4 * l + 1] = fy;
        K[4 * l + 2] = px;
        K[4 * l + 3] = py;
        fx = 0.5f * fx;
        fy = 0.5f * fy;
        px = 0.5f * (px - 0.5f);
        py = 0.5f * (py - 0.5f);
    }
}

// rotation angle of a rotation matrix (degrees)
double rotation_deg(const double *T) {
    double c = (T[0] + T[5] + T[10] - 1.0) * 0.5;
    c = std::max(-1.0, std::min(1.0, c));
    return std::acos(c) * 180.0 / 3.14159265358979323846;
}

template <class R>
int icp_track(const float *const *d1, int W0, int H0, const float *K, int levels, const int *iters,
              const float *const *mv, const float *const *mn, float dist2, int min_valid, float max_m,
              float max_deg, float rel_pivot, const double *T_init, double *T_gc, int *iters_done) {
    double T12[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, T21[16];
    std::memcpy(T21, T12, sizeof(T12));
    int nlev = levels;
    if (T_init) {
        std::memcpy(T12, T_init, sizeof(T12));
        rigid_inverse(T12, T21);
        nlev = 1;
    }
    int solves = 0;
    for (int n = nlev - 1; n >= 0; --n) {
        const int W = W0 >> n, H = H0 >> n;
        for (int it = 0; it < iters[n]; ++it) {
            float T21f[16];
            for (int k = 0; k < 16; ++k) T21f[k] = (float)T21[k];
            R ATAr[36], ATbr[6];
            long long cnt = icp_system<R>(d1[n], W, H, K + 4 * n, mv[n], mn[n], T21f, dist2, ATAr, ATbr, nullptr);
            ++solves;
            if (min_valid > 0 && cnt < min_valid) {
                if (iters_done) *iters_done = solves;
                return 7;
            }
            double ATA[36], ATb[6], x[6];
            for (int k = 0; k < 36; ++k) ATA[k] = ATAr[k];
            for (int k = 0; k < 6; ++k) ATb[k] = ATbr[k];
            const int cs = cholesky_status(ATA, ATb, (double)rel_pivot, x);
            if (cs != 0) {
                if (iters_done) *iters_done = solves;
                return cs;
            }
            double ninf = 0;
            for (int k = 0; k < 6; ++k) ninf = std::max(ninf, std::fabs(x[k]));
            if (ninf < 5.0e-7) break;
            double E[16];
            expmap(x, E);
            matmul4(E, T12, T12);
            rigid_inverse(T12, T21);
        }
    }
    if (iters_done) *iters_done = solves;
    if (max_m > 0 || max_deg > 0) {
        double tn = std::sqrt(T12[3] * T12[3] + T12[7] * T12[7] + T12[11] * T12[11]);
        if ((max_m > 0 && tn > max_m) || (max_deg > 0 && rotation_deg(T12) > max_deg)) return 5;
    }
    matmul4(T12, T_gc, T_gc);
    return 0;
}

// ---------------------------------------------------------------------------
// A13-A14 integrate (z_yooji_cuda_volume_integrator.cu:84-262, 1006-1010)
// ---------------------------------------------------------------------------
// g_fvsc_Find_Visible_Sub_Cubes (:84-152) after the per-frame mask clear
// (:1006-1010): per valid pixel, 6 samples along the ray (n = -4..1)
void mark_bricks(const uint16_t *mm, int W, int H, const float *K, const float *Tcg, const Vol &V,
                 uint8_t *mask) {
    std::memset(mask, 0, (size_t)V.BW * V.BH * V.BD);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float td = (float)mm[(size_t)y * W + x] * 0.001f;
            if (!(td > 0.0f)) continue;
            f3 p = transform(Tcg, back_project((float)x, (float)y, K, td));
            f3 pe = transform(Tcg, back_project((float)x, (float)y, K, 100.0f));
            f3 dir = normalised(sub(pe, p));
            for (int n = (OR_MUTATION == 1 ? -1 : -4); n <= 1; ++n) {
                f3 q = add(p, scale((float)n * V.mu, dir));
                f3 tv;
                if (!world_to_voxel(V, q, tv)) continue;
                int bxyz[3];
                const float c[3] = {tv.x, tv.y, tv.z};
                for (int k = 0; k < 3; ++k) {
                    // ROUNDF(v)/8 with the macro's precedence (Q6)
#if OR_MUTATION == 2
                    bxyz[k] = ((c[k] < 0.0f) ? (int)(c[k] - 0.5f) : (int)(c[k] + 0.5f)) / 8;
#else
                    bxyz[k] = (c[k] < 0.0f) ? (int)(c[k] - 0.5f) : (int)(c[k] + 0.5f) / 8;
#endif
                }
                if (bxyz[0] < 0 || bxyz[0] >= V.BW || bxyz[1] < 0 || bxyz[1] >= V.BH || bxyz[2] < 0 ||
                    bxyz[2] >= V.BD)
                    break;
                mask[((size_t)bxyz[2] * V.BH + bxyz[1]) * V.BW + bxyz[0]] = 1;  // benign race: same value
            }
        }
}

template <class R>
void integrate(const uint16_t *mm, int W, int H, const float *K, const float *Tgc, const float *Tcg,
               const or_volume *vol, R *F, R *Wt, uint8_t *bmask, uint8_t *upd, long long *counts) {
    const Vol V = make_vol(vol);
    const size_t nb = (size_t)V.BW * V.BH * V.BD;
    std::vector<uint8_t> local(nb, 0);
    uint8_t *mask = bmask ? bmask : local.data();
    mark_bricks(mm, W, H, K, Tcg, V, mask);
    long long vis = 0, nupd = 0;
    for (size_t b = 0; b < nb; ++b) vis += mask[b];
    const size_t N = (size_t)V.W * V.H * V.D;
    if (upd) std::memset(upd, 0, N);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nupd)
    for (long long b = 0; b < (long long)nb; ++b) {
        if (!mask[b]) continue;
        const int bx = (int)(b % V.BW), by = (int)((b / V.BW) % V.BH), bz = (int)(b / ((long long)V.BW * V.BH));
        for (int zs = 0; zs < 8; ++zs)
            for (int ys = 0; ys < 8; ++ys)
                for (int xs = 0; xs < 8; ++xs) {
                    const int xc = bx * 8 + xs, yc = by * 8 + ys, zc = bz * 8 + zs;
                    f3 g = voxel_to_world(V, {(float)xc, (float)yc, (float)zc});
                    f3 p = transform(Tgc, g);
                    float u, v;
                    project(p, K, u, v);
                    int iu, iv;
                    if (!roundf_ref(u, iu) || !roundf_ref(v, iv)) continue;
                    if (iu < 0 || iu > W - 1 || iv < 0 || iv > H - 1) continue;
                    float dm = (float)mm[(size_t)iv * W + iu] * 0.001f;
                    if (!(dm > 0.0f)) continue;
                    float td = dm - p.z;
                    if (td < -V.mu) continue;
                    const size_t idx = ((size_t)zc * V.H + yc) * V.W + xc;
                    R tsd = std::min(R(1), R(td) / R(V.mu));
                    R oldW = Wt[idx], newW = oldW + R(1);
                    F[idx] = (oldW * F[idx] + tsd) / newW;
                    Wt[idx] = std::min(newW, R(V.max_w));
                    if (upd) upd[idx] = 1;
                    ++nupd;
                }
    }
    if (counts) {
        counts[0] = vis;
        counts[1] = nupd;
    }
}

// ---------------------------------------------------------------------------
// Production integrate with colour: g_ctvc_Compute_TSDF_of_Voxels_with_Color
// (z_yooji_cuda_volume_integrator.cu:265-440) as driven by
// cdtocc_Convert_Depth_to_TSDF_on_Cube_with_Color (:976-1031) with
// in_flag_on_rgb = true (z_yooji_object_scanner.cpp:1274-1281): the colour
// frame is registered to the depth frame. Differences from g_ctv: pixel by
// __float2int_rn, SDF along the ray (depth - |g - C| / sqrt(xn^2 + yn^2 + 1)),
// voxels at the weight cap are frozen, the average multiplies by 1/newW, and
// each colour channel is a running average of its bilinear sample
// (d_gii_Get_Interpolated_Intensity, _yooji_cuda_math.h:305-332), stored as
// uchar. Colour arithmetic is FP32 in the reference's order (Q0), so colour
// volumes compare bit for bit.
// ---------------------------------------------------------------------------
// __float2int_rn (round half to even), Q18
inline bool rint_ref(float v, int &out) {
    if (!(v > -2147483520.0f && v < 2147483520.0f)) return false;
    out = (int)std::nearbyintf(v);
    return true;
}
// d_gii_Get_Interpolated_Intensity: bilinear over the 2x2 neighbourhood of an
// 8-bit plane; a participating tap <= 0 or a position outside fails
inline bool gii(const uint8_t *plane, int W, int H, float xf, float yf, float &out) {
    if (xf < 0.0f || xf > float(W - 1) || yf < 0.0f || yf > float(H - 1)) return false;
    float d2 = 0.0f, d3 = 0.0f, d4 = 0.0f;
    const int x = (int)xf, y = (int)yf;
    const float rx = xf - (float)x, ry = yf - (float)y;
    const float d1 = (float)plane[y * W + x];
    if (d1 <= 0.0f) return false;
    if (rx > 0.0f) {
        d2 = (float)plane[y * W + x + 1];
        if (d2 <= 0.0f) return false;
    }
    if (ry > 0.0f) {
        d3 = (float)plane[(y + 1) * W + x];
        if (d3 <= 0.0f) return false;
    }
    if (rx > 0.0f && ry > 0.0f) {
        d4 = (float)plane[(y + 1) * W + x + 1];
        if (d4 <= 0.0f) return false;
    }
    out = (1.0f - ry) * ((1.0f - rx) * d1 + rx * d2) + ry * ((1.0f - rx) * d3 + rx * d4);
    return true;
}

template <class R>
void integrate_color(const uint16_t *mm, const uint8_t *rgb, int W, int H, const float *K, const float *Tgc,
                     const float *Tcg, const or_volume *vol, R *F, R *Wt, uint8_t *C, uint8_t *bmask,
                     uint8_t *upd, long long *counts) {
    const Vol V = make_vol(vol);
    const size_t nb = (size_t)V.BW * V.BH * V.BD;
    std::vector<uint8_t> local(nb, 0);
    uint8_t *mask = bmask ? bmask : local.data();
    mark_bricks(mm, W, H, K, Tcg, V, mask);  // the brick pass is shared with g_ctv
    long long vis = 0, nupd = 0;
    for (size_t b = 0; b < nb; ++b) vis += mask[b];
    const size_t N = (size_t)V.W * V.H * V.D;
    if (upd) std::memset(upd, 0, N);
    const f3 cam = {Tcg[3], Tcg[7], Tcg[11]};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nupd)
    for (long long b = 0; b < (long long)nb; ++b) {
        if (!mask[b]) continue;
        const int bx = (int)(b % V.BW), by = (int)((b / V.BW) % V.BH), bz = (int)(b / ((long long)V.BW * V.BH));
        for (int zs = 0; zs < 8; ++zs)
            for (int ys = 0; ys < 8; ++ys)
                for (int xs = 0; xs < 8; ++xs) {
                    const int xc = bx * 8 + xs, yc = by * 8 + ys, zc = bz * 8 + zs;
                    const f3 g = voxel_to_world(V, {(float)xc, (float)yc, (float)zc});
                    const f3 p = transform(Tgc, g);
                    float u, v;
                    project(p, K, u, v);
                    int x, y;
                    if (!rint_ref(u, x) || !rint_ref(v, y)) continue;
                    if (x < 0 || x > W - 1 || y < 0 || y > H - 1) continue;
                    const float dm = (float)mm[(size_t)y * W + x] * 0.001f;
                    if (!(dm > 0.0f)) continue;
                    // lambda: the pixel's ray at unit depth (d_bp_Back_Project of the int pixel)
                    const f3 r1 = {1.0f * ((float)x - K[2]) / K[0], 1.0f * ((float)y - K[3]) / K[1], 1.0f};
                    const float lambda_inv = 1.0f / sqrtf(r1.x * r1.x + r1.y * r1.y + 1.0f);
                    const f3 dg = sub(g, cam);
                    const float norm = sqrtf(dg.x * dg.x + dg.y * dg.y + dg.z * dg.z);
                    const float sdf = dm - lambda_inv * norm;
                    if (sdf < -V.mu) continue;
                    const size_t idx = ((size_t)zc * V.H + yc) * V.W + xc;
                    const float oldWf = (float)Wt[idx];
                    if (oldWf >= V.max_w) continue;  // frozen voxel
                    const float newWf = oldWf + 1.0f, inewWf = 1.0f / newWf;
                    const R tsd = std::min(R(1), R(sdf) / R(V.mu));
                    const R oldW = Wt[idx];
                    F[idx] = (oldW * F[idx] + tsd) * (R(1) / (oldW + R(1)));
                    Wt[idx] = oldW + R(1);
                    for (int k = 0; k < 3; ++k) {
                        float sample;
                        if (!gii(rgb + (size_t)k * W * H, W, H, u, v, sample)) continue;  // Q17
                        const size_t ci = (size_t)k * N + idx;
                        const float val = (oldWf * (float)C[ci] + sample) * inewWf;
                        C[ci] = (uint8_t)std::fmax(0.0f, std::fmin(255.0f, val));
                    }
                    if (upd) upd[idx] = 1;
                    ++nupd;
                }
    }
    if (counts) {
        counts[0] = vis;
        counts[1] = nupd;
    }
}

// ---------------------------------------------------------------------------
// A16-A18 raycast (z_yooji_cuda_tsdf_renderer.cu:68-153, 288-635, 1069-1209,
// 1642-1700)
// ---------------------------------------------------------------------------
struct Cast {
    const float *F, *Wt;
    Vol V;
};
// d_gtvu: nearest tap; invalid (outside or weight <= 0) reads as 1.0
inline bool tap_nearest(const Cast &c, f3 vox, float &out) {
    out = 1.0f;
    int x, y, z;
    if (!roundf_ref(vox.x, x) || !roundf_ref(vox.y, y) || !roundf_ref(vox.z, z)) return false;
    if (x < 0 || x >= c.V.W || y < 0 || y >= c.V.H || z < 0 || z >= c.V.D) return false;
    size_t idx = ((size_t)z * c.V.H + y) * c.V.W + x;
    if (!(c.Wt[idx] > 0.0f)) return false;
    out = c.F[idx];
    return true;
}
inline bool tap_int(const Cast &c, int x, int y, int z, float &out) {
    return tap_nearest(c, {(float)x, (float)y, (float)z}, out);
}
// Q2: truncation toInt
inline void to_int(f3 v, int o[3], f3 &res) {
    o[0] = (int)v.x;
    o[1] = (int)v.y;
    o[2] = (int)v.z;
    res = {v.x - (float)o[0], v.y - (float)o[1], v.z - (float)o[2]};
}
inline bool toint_ok(f3 v) {
    return std::fabs(v.x) < 2147483520.0f && std::fabs(v.y) < 2147483520.0f && std::fabs(v.z) < 2147483520.0f;
}
// d_gtvi: trilinear over 8 taps; any invalid tap -> false with out = 1.0
inline bool tap_trilinear(const Cast &c, f3 vox, float &out) {
    out = 1.0f;
    if (!toint_ok(vox)) return false;
    int o[3];
    f3 r;
    to_int(vox, o, r);
    float f[4], b[4];
    static const int dx[4] = {0, 1, 0, 1}, dy[4] = {0, 0, 1, 1};
    for (int i = 0; i < 4; ++i)
        if (!tap_int(c, o[0] + dx[i], o[1] + dy[i], o[2], f[i])) return false;
    for (int i = 0; i < 4; ++i)
        if (!tap_int(c, o[0] + dx[i], o[1] + dy[i], o[2] + 1, b[i])) return false;
    float fi = (1.0f - r.y) * ((1.0f - r.x) * f[0] + r.x * f[1]) + r.y * ((1.0f - r.x) * f[2] + r.x * f[3]);
    float bi = (1.0f - r.y) * ((1.0f - r.x) * b[0] + r.x * b[1]) + r.y * ((1.0f - r.x) * b[2] + r.x * b[3]);
    out = (1.0f - r.z) * fi + r.z * bi;
    return true;
}

struct RayParams {
    float K[4];
    float Tgc[16], Tcg[16];
    f3 cam, ncc;
    float dist_cc, r_cube, cos_max, step_scale, total_max;
    int max_iter;
};
RayParams ray_params(const Vol &V, const float *K, const float *Tgc, const float *Tcg) {
    RayParams P;
    std::memcpy(P.K, K, sizeof(P.K));
    std::memcpy(P.Tgc, Tgc, sizeof(P.Tgc));
    std::memcpy(P.Tcg, Tcg, sizeof(P.Tcg));
    P.cam = camera_centre(Tgc);
    // cube centre = origin + dim*voxel/2 (z_yooji_object_scanner.cpp:306)
    const float vox = V.voxel;
    f3 cen = {V.origin.x + (float)V.W * vox / 2.0f, V.origin.y + (float)V.H * vox / 2.0f,
              V.origin.z + (float)V.D * vox / 2.0f};
    f3 vcc = sub(cen, P.cam);
    P.ncc = normalised(vcc);
    P.dist_cc = length(vcc);
    P.r_cube = length(sub(V.origin, cen));
    float s = P.r_cube / P.dist_cc;
    P.cos_max = (s < 1.0f) ? sqrtf(1.0f - s * s) : -2.0f;
    P.step_scale = 0.9f * V.mu * V.vox_inv;
    P.total_max = 2.0f * P.r_cube * V.vox_inv;
    P.max_iter = (int)(P.total_max / std::min(P.step_scale, 1.0f)) * OR_MARCH_CAP_MULT + 64;
    return P;
}

// d_crtc_Cast_Ray_on_TSDF_Cube_NEW (tsdf_renderer.cu:466-635). FP32 march.
inline bool cast_ray(const Cast &c, const RayParams &P, float px, float py, f3 &hit, long long *st) {
    enum { COARSE, FINE, BEHIND };
    const float dist_min = P.dist_cc - P.r_cube, dist_max = P.dist_cc + P.r_cube;
    f3 pe = transform(P.Tcg, back_project(px, py, P.K, dist_max));
    f3 rd = normalised(sub(pe, P.cam));
    if (dot(rd, P.ncc) < P.cos_max) return false;
    f3 ps = add(P.cam, scale(dist_min, rd));
    f3 vox;
    world_to_voxel(c.V, ps, vox);
    float tsdf;
    int state;
    if (st) st[0]++;
    bool valid = tap_nearest(c, vox, tsdf);
    if (!valid) state = COARSE;
    else if (tsdf >= 0.0f) state = FINE;
    else return false;
    float total = 0.0f;
    int iter = 0;
    while (state != BEHIND) {
        if (++iter > P.max_iter) return false;
        float step = (state == COARSE) ? P.step_scale
                                       : ((tsdf > 0.0f) ? std::max(P.step_scale * tsdf, 1.0f)
                                                        : std::min(P.step_scale * tsdf, -1.0f));
        vox = add(vox, scale(step, rd));
        total += step;
        if (total > P.total_max) break;
        if (st) st[0]++;
        tap_nearest(c, vox, tsdf);
        if (tsdf > -1.0f && tsdf < 1.0f) {
            if (st) st[1]++;
            tap_trilinear(c, vox, tsdf);
            if (tsdf < 0.0f) {
                if (state == COARSE) break;
                state = BEHIND;
            } else if (tsdf == 0.0f) {
                hit = voxel_to_world(c.V, vox);
                return true;
            } else if (state == COARSE) {
                state = FINE;
            }
        }
    }
    if (state != BEHIND) return false;
    float step = P.step_scale * tsdf;
    vox = add(vox, scale(step, rd));
    if (st) st[1]++;
    tap_trilinear(c, vox, tsdf);
    step = P.step_scale * tsdf;
    vox = add(vox, scale(step, rd));
    hit = voxel_to_world(c.V, vox);
    return true;
}

// d_csnt_Compute_Surface_Normal_from_TSDF (tsdf_renderer.cu:288-464). Taps and
// the zero-gradient test in FP32; the normal's value in R.
template <class R>
inline bool surface_normal(const Cast &c, f3 p, R n[3]) {
    f3 vox;
    if (!world_to_voxel(c.V, p, vox)) return false;
    int o[3];
    f3 r;
    to_int(vox, o, r);
    static const int dx1[4] = {0, 1, 0, 1}, dy1[4] = {0, 0, 1, 1};
#if OR_MUTATION == 3
    static const int dx2[12] = {0, 1, -1, 0, 1, 1, -1, 0, 1, 2, 0, 1};
#else
    static const int dx2[12] = {0, 1, -1, 0, 1, 2, -1, 0, 1, 2, 0, 1};
#endif
    static const int dy2[12] = {-1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2};
    float fm[4], fr[12], bk[12], bm[4];
    for (int i = 0; i < 4; ++i)
        if (!tap_int(c, o[0] + dx1[i], o[1] + dy1[i], o[2] - 1, fm[i])) return false;
    for (int i = 0; i < 12; ++i)
        if (!tap_int(c, o[0] + dx2[i], o[1] + dy2[i], o[2], fr[i])) return false;
    for (int i = 0; i < 12; ++i)
        if (!tap_int(c, o[0] + dx2[i], o[1] + dy2[i], o[2] + 1, bk[i])) return false;
    for (int i = 0; i < 4; ++i)
        if (!tap_int(c, o[0] + dx1[i], o[1] + dy1[i], o[2] + 2, bm[i])) return false;
    // gradient in FP32 (decides the zero test) and in R (the value)
    auto grad = [&](auto rx, auto ry, auto rz, auto &gx, auto &gy, auto &gz) {
        using T = std::decay_t<decltype(rx)>;
        const T one = T(1);
        T a4[4], b4[4], a2[2], b2[2];
        auto L = [&](T w, float u, float v) { return (one - w) * T(u) + w * T(v); };
        a4[0] = L(rz, fr[2], bk[2]); a4[1] = L(rz, fr[3], bk[3]); a4[2] = L(rz, fr[4], bk[4]); a4[3] = L(rz, fr[5], bk[5]);
        b4[0] = L(rz, fr[6], bk[6]); b4[1] = L(rz, fr[7], bk[7]); b4[2] = L(rz, fr[8], bk[8]); b4[3] = L(rz, fr[9], bk[9]);
        a2[0] = (one - rx) * a4[0] + rx * a4[1];
        a2[1] = (one - rx) * b4[0] + rx * b4[1];
        b2[0] = (one - rx) * a4[2] + rx * a4[3];
        b2[1] = (one - rx) * b4[2] + rx * b4[3];
        gx = (one - ry) * b2[0] + ry * b2[1];
        gx -= (one - ry) * a2[0] + ry * a2[1];
        a4[0] = L(rz, fr[0], bk[0]); a4[1] = L(rz, fr[3], bk[3]); a4[2] = L(rz, fr[7], bk[7]); a4[3] = L(rz, fr[10], bk[10]);
        b4[0] = L(rz, fr[1], bk[1]); b4[1] = L(rz, fr[4], bk[4]); b4[2] = L(rz, fr[8], bk[8]); b4[3] = L(rz, fr[11], bk[11]);
        a2[0] = (one - ry) * a4[0] + ry * a4[1];
        a2[1] = (one - ry) * b4[0] + ry * b4[1];
        b2[0] = (one - ry) * a4[2] + ry * a4[3];
        b2[1] = (one - ry) * b4[2] + ry * b4[3];
        gy = (one - rx) * b2[0] + rx * b2[1];
        gy -= (one - rx) * a2[0] + rx * a2[1];
        a4[0] = L(ry, fm[0], fm[2]); a4[1] = L(ry, fr[3], fr[7]); a4[2] = L(ry, bk[3], bk[7]); a4[3] = L(ry, bm[0], bm[2]);
        b4[0] = L(ry, fm[1], fm[3]); b4[1] = L(ry, fr[4], fr[8]); b4[2] = L(ry, bk[4], bk[8]); b4[3] = L(ry, bm[1], bm[3]);
        a2[0] = (one - rz) * a4[0] + rz * a4[1];
        a2[1] = (one - rz) * b4[0] + rz * b4[1];
        b2[0] = (one - rz) * a4[2] + rz * a4[3];
        b2[1] = (one - rz) * b4[2] + rz * b4[3];
        gz = (one - rx) * b2[0] + rx * b2[1];
        gz -= (one - rx) * a2[0] + rx * a2[1];
    };
    float gxf, gyf, gzf;
    grad(r.x, r.y, r.z, gxf, gyf, gzf);
    if (gxf == 0.0f && gyf == 0.0f && gzf == 0.0f) return false;
    R gx, gy, gz;
    grad(R(r.x), R(r.y), R(r.z), gx, gy, gz);
    R len = std::sqrt(gx * gx + gy * gy + gz * gz);
    if (!(len > R(0))) return false;
    n[0] = gx / len;
    n[1] = gy / len;
    n[2] = gz / len;
    return true;
}

// d_grvi_Get_RGB_Value_Interpolated (z_yooji_cuda_tsdf_renderer.cu:232-284)
// with d_grvu (:192-229) taps: trilinear colour over the 8 corners of
// toInt(vox); a tap outside the grid or without weight fails the whole
// lookup, which then leaves (0, 0, 0). Result rounded and clamped (Q16).
inline void grvi(const Cast &c, const uint8_t *C, f3 vox, uint8_t out[3]) {
    out[0] = out[1] = out[2] = 0;
    if (!toint_ok(vox)) return;
    int o[3];
    f3 r;
    to_int(vox, o, r);
    static const int dx[4] = {0, 1, 0, 1}, dy[4] = {0, 0, 1, 1};
    const size_t N = (size_t)c.V.W * c.V.H * c.V.D;
    float fr[4][3], bk[4][3];
    for (int zz = 0; zz < 2; ++zz)
        for (int i = 0; i < 4; ++i) {
            const int x = o[0] + dx[i], y = o[1] + dy[i], z = o[2] + zz;
            if (x < 0 || x >= c.V.W || y < 0 || y >= c.V.H || z < 0 || z >= c.V.D) return;
            const size_t idx = ((size_t)z * c.V.H + y) * c.V.W + x;
            if (!(c.Wt[idx] > 0.0f)) return;
            for (int k = 0; k < 3; ++k) (zz ? bk : fr)[i][k] = (float)C[k * N + idx];
        }
    for (int k = 0; k < 3; ++k) {
        const float f = (1.0f - r.y) * ((1.0f - r.x) * fr[0][k] + r.x * fr[1][k]) +
                        r.y * ((1.0f - r.x) * fr[2][k] + r.x * fr[3][k]);
        const float b = (1.0f - r.y) * ((1.0f - r.x) * bk[0][k] + r.x * bk[1][k]) +
                        r.y * ((1.0f - r.x) * bk[2][k] + r.x * bk[3][k]);
        const float v = (1.0f - r.z) * f + r.z * b;
        const int q = (v < 0.0f) ? (int)(v - 0.5f) : (int)(v + 0.5f);
        out[k] = (uint8_t)std::min(255, std::max(0, q));
    }
}

// ---------------------------------------------------------------------------
// Input pre-steps. z_tdrcc_Transform_Depth_to_Rgb_Camera_Coordinates
// (z_yooji_object_scanner.cpp:2429-2476) with convert_pixel_coord
// (z_yooji_matrix_rgbd.cpp:101-118) and the CKvYooji_Intrinsics /
// CKvYooji_Extrinsics maths (z_yooji_camera_parameters.cpp:89-108, 678-694);
// god_Get_Object_Depth's cube test (z_yooji_object_extractor.cpp:357-410).
// Both move or zero depth values without changing them, so they run on the
// uint16 millimetres (depth = mm * 0.001f as A1).
// ---------------------------------------------------------------------------
void depth_reproject(const uint16_t *src, int sw, int sh, const float *Ks, const float *T, uint16_t *dst, int dw,
                     int dh, const float *Kd) {
    std::memset(dst, 0, sizeof(uint16_t) * (size_t)dw * dh);
    for (int j = 0; j < sh; ++j)
        for (int i = 0; i < sw; ++i) {
            const uint16_t raw = src[(size_t)j * sw + i];
            const float td = (float)raw * 0.001f;
            if (!(td > 0.0f)) continue;
            const f3 p = {td * ((float)i - Ks[2]) / Ks[0], td * ((float)j - Ks[3]) / Ks[1], td};
            const f3 q = transform(T, p);
            const float u = Kd[0] * q.x / q.z + Kd[2], v = Kd[1] * q.y / q.z + Kd[3];
            int tx, ty;
            if (!roundf_ref(u, tx) || !roundf_ref(v, ty)) continue;
            if (tx < 0 || tx > dw - 1 || ty < 0 || ty > dh - 1) continue;
            uint16_t &o = dst[(size_t)ty * dw + tx];
            // "if(out > td || out <= 0) out = td": the nearest depth wins
            if ((float)o * 0.001f > td || o == 0) o = raw;
        }
}

// g_uvv_Update_Valid_Volume (z_yooji_cuda_volume_integrator.cu:718-799): bricks
// of the mask with no valid voxel are reset (TSDF 1, weight 0, colour 0) and
// leave the mask. Returns the number reset.
long long volume_carve(double *F, double *Wt, uint8_t *C, uint8_t *bmask, const uint8_t *valid, const Vol &V) {
    const size_t N = (size_t)V.W * V.H * V.D;
    long long carved = 0;
    for (int bz = 0; bz < V.BD; ++bz)
        for (int by = 0; by < V.BH; ++by)
            for (int bx = 0; bx < V.BW; ++bx) {
                const size_t b = ((size_t)bz * V.BH + by) * V.BW + bx;
                if (!bmask[b]) continue;
                int cnt = 0;
                for (int z = bz * 8; z < bz * 8 + 8; ++z)
                    for (int y = by * 8; y < by * 8 + 8; ++y)
                        for (int x = bx * 8; x < bx * 8 + 8; ++x) cnt += valid[((size_t)z * V.H + y) * V.W + x] != 0;
                if (cnt > 0) continue;
                bmask[b] = 0;
                ++carved;
                for (int z = bz * 8; z < bz * 8 + 8; ++z)
                    for (int y = by * 8; y < by * 8 + 8; ++y)
                        for (int x = bx * 8; x < bx * 8 + 8; ++x) {
                            const size_t i = ((size_t)z * V.H + y) * V.W + x;
                            F[i] = 1.0;
                            Wt[i] = 0.0;
                            if (C)
                                for (int k = 0; k < 3; ++k) C[k * N + i] = 0;
                        }
            }
    return carved;
}

void depth_crop(uint16_t *mm, int W, int H, const float *K, const float *Tcg, const Vol &V) {
    const float sx = (float)V.W * V.voxel, sy = (float)V.H * V.voxel, sz = (float)V.D * V.voxel;
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const float td = (float)mm[(size_t)j * W + i] * 0.001f;
            const f3 p = {td * ((float)i - K[2]) / K[0], td * ((float)j - K[3]) / K[1], td};
            const f3 g = transform(Tcg, p);
            if (g.x < V.origin.x || g.x > V.origin.x + sx || g.y < V.origin.y || g.y > V.origin.y + sy ||
                g.z < V.origin.z || g.z > V.origin.z + sz)
                mm[(size_t)j * W + i] = 0;
        }
}

template <class R>
void raycast(const float *F, const float *Wt, const or_volume *vol, int W, int H, const float *K, const float *Tgc,
             const float *Tcg, R *depth, R *vtx, R *nrm, uint8_t *img, long long *stats, const uint8_t *C = nullptr,
             uint8_t *tex = nullptr) {
    Cast c{F, Wt, make_vol(vol)};
    const RayParams P = ray_params(c.V, K, Tgc, Tcg);
    long long s0 = 0, s1 = 0, s2 = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : s0, s1, s2)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const size_t t = (size_t)y * W + x;
            long long st[2] = {0, 0};
            f3 p;
            R n[3];
            bool ok = cast_ray(c, P, (float)x, (float)y, p, st) && surface_normal<R>(c, p, n);
            s0 += st[0];
            s1 += st[1];
            if (ok) {
                ++s2;
                const R px = p.x, py = p.y, pz = p.z;
                R pc[3], nc[3];
                for (int r = 0; r < 3; ++r) {
                    pc[r] = R(Tgc[4 * r]) * px + R(Tgc[4 * r + 1]) * py + R(Tgc[4 * r + 2]) * pz + R(Tgc[4 * r + 3]);
                    nc[r] = R(Tgc[4 * r]) * n[0] + R(Tgc[4 * r + 1]) * n[1] + R(Tgc[4 * r + 2]) * n[2];
                }
                if (depth) depth[t] = pc[2];
                for (int k = 0; k < 3; ++k) {
                    if (vtx) vtx[3 * t + k] = pc[k];
                    if (nrm) nrm[3 * t + k] = nc[k];
                }
                // light (0,0,-1) (z_yooji_object_scanner.cpp:1313)
                R sa = -nc[2];
                R v = (R(OR_MUTATION == 4 ? 0.75 : 0.8) * sa + R(0.2)) * R(255);
                v = std::max(R(0), std::min(R(255), v));
                if (img) img[t] = (uint8_t)v;
                if (tex) {
                    f3 pv;
                    world_to_voxel(c.V, p, pv);
                    uint8_t rgb[3];
                    grvi(c, C, pv, rgb);
                    for (int k = 0; k < 3; ++k) tex[(size_t)k * W * H + t] = rgb[k];
                }
            } else {
                if (depth) depth[t] = R(0);
                for (int k = 0; k < 3; ++k) {
                    if (vtx) vtx[3 * t + k] = R(-100);
                    if (nrm) nrm[3 * t + k] = R(-100);
                }
                if (img) img[t] = 0;
                if (tex)
                    for (int k = 0; k < 3; ++k) tex[(size_t)k * W * H + t] = 0;
            }
        }
    }
    if (stats) {
        stats[0] = s0;
        stats[1] = s1;
        stats[2] = s2;
        stats[3] = (long long)W * H;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// A19 per-frame driver (z_yooji_object_scanner.cpp:472-629, 1265-1356)
// ---------------------------------------------------------------------------
struct or_session {
    or_session_desc d;
    float K[12];
    double Tgc[16];
    int frame = 0;
    std::vector<double> Fd, Wd;
    std::vector<float> Ff, Wf;
    std::vector<float> filt[3];           // filtered pyramid (float view)
    std::vector<float> mdepth[3], mvtx[3], mnrm[3];
};

namespace {
template <class R>
int fuse_frame_impl(or_session *s, const uint16_t *mm, double *ms) {
    using clk = std::chrono::steady_clock;
    const or_session_desc &d = s->d;
    const int W = d.width, H = d.height;
    auto t0 = clk::now();
    // 1 filter
    std::vector<R> lev[3];
    lev[0].resize((size_t)W * H);
    bilateral<R>(mm, W, H, d.bil_sigma_s, d.bil_sigma_r, d.bil_radius, lev[0].data());
    auto t1 = clk::now();
    // 2 pyramid
    for (int l = 1; l < d.levels; ++l) {
        lev[l].resize((size_t)(W >> l) * (H >> l));
        pyr_down<R>(lev[l - 1].data(), W >> (l - 1), H >> (l - 1), lev[l].data());
    }
    for (int l = 0; l < d.levels; ++l) s->filt[l].assign(lev[l].begin(), lev[l].end());
    auto t2 = clk::now();
    // 3 track (frames > 0)
    int status = 0;
    if (s->frame > 0) {
        const float *dp[3], *mv[3], *mn[3];
        for (int l = 0; l < d.levels; ++l) {
            dp[l] = s->filt[l].data();
            mv[l] = s->mvtx[l].data();
            mn[l] = s->mnrm[l].data();
        }
        status = icp_track<R>(dp, W, H, s->K, d.levels, d.iters, mv, mn, d.icp_dist2, d.icp_min_valid,
                              d.max_motion_m, d.max_motion_deg, d.icp_rel_pivot, nullptr, s->Tgc, nullptr);
    }
    auto t3 = clk::now();
    if (status != 0) {
        if (ms) {
            ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
            ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
            ms[3] = ms[4] = 0;
        }
        s->frame++;
        return status;
    }
    // 4 integrate raw depth at the current pose
    float Tgcf[16], Tcgf[16];
    double Tcg[16];
    rigid_inverse(s->Tgc, Tcg);
    for (int k = 0; k < 16; ++k) {
        Tgcf[k] = (float)s->Tgc[k];
        Tcgf[k] = (float)Tcg[k];
    }
    R *F = d.fp32 ? (R *)s->Ff.data() : (R *)s->Fd.data();
    R *Wt = d.fp32 ? (R *)s->Wf.data() : (R *)s->Wd.data();
    integrate<R>(mm, W, H, s->K, Tgcf, Tcgf, &d.vol, F, Wt, nullptr, nullptr, nullptr);
    auto t4 = clk::now();
    // 5 raycast every level into the model maps
    const size_t N = (size_t)d.vol.dim[0] * d.vol.dim[1] * d.vol.dim[2];
    const float *Fv, *Wv;
    std::vector<float> Fc, Wc;
    if (d.fp32) {
        Fv = s->Ff.data();
        Wv = s->Wf.data();
    } else {
        Fc.resize(N);
        Wc.resize(N);
        for (size_t i = 0; i < N; ++i) {
            Fc[i] = (float)s->Fd[i];
            Wc[i] = (float)s->Wd[i];
        }
        Fv = Fc.data();
        Wv = Wc.data();
    }
    for (int l = 0; l < d.levels; ++l) {
        const int Wl = W >> l, Hl = H >> l;
        std::vector<R> dep((size_t)Wl * Hl), vt((size_t)Wl * Hl * 3), nr((size_t)Wl * Hl * 3);
        raycast<R>(Fv, Wv, &d.vol, Wl, Hl, s->K + 4 * l, Tgcf, Tcgf, dep.data(), vt.data(), nr.data(), nullptr,
                   nullptr);
        s->mdepth[l].assign(dep.begin(), dep.end());
        s->mvtx[l].assign(vt.begin(), vt.end());
        s->mnrm[l].assign(nr.begin(), nr.end());
    }
    auto t5 = clk::now();
    if (ms) {
        ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
        ms[3] = std::chrono::duration<double, std::milli>(t4 - t3).count();
        ms[4] = std::chrono::duration<double, std::milli>(t5 - t4).count();
    }
    s->frame++;
    return 0;
}
}  // namespace

extern "C" {

void or_depth_to_m(const uint16_t *mm, long long n, float *out) {
    for (long long i = 0; i < n; ++i) out[i] = (float)mm[i] * 0.001f;
}
void or_gaussian_weights(float sigma_s, int radius, float *g) { gaussian_weights(sigma_s, radius, g); }
void or_bilateral_d(const uint16_t *mm, int W, int H, float ss, float sr, int r, double *out) {
    bilateral<double>(mm, W, H, ss, sr, r, out);
}
void or_bilateral_f(const uint16_t *mm, int W, int H, float ss, float sr, int r, float *out) {
    bilateral<float>(mm, W, H, ss, sr, r, out);
}
void or_pyr_down_d(const double *in, int W, int H, double *out) { pyr_down<double>(in, W, H, out); }
void or_pyr_down_f(const float *in, int W, int H, float *out) { pyr_down<float>(in, W, H, out); }
void or_intrinsics_pyramid(const float K0[4], int levels, float *K) { intrinsics_pyramid(K0, levels, K); }
void or_vnmap_d(const float *depth, int W, int H, const float K[4], int inplace, double *vtx, double *nrm) {
    vnmap<double>(depth, W, H, K, inplace, vtx, nrm);
}
long long or_icp_system_d(const float *d1, int W, int H, const float K[4], const float *mv, const float *mn,
                          const float T21[16], float dist2, double ATA[36], double ATb[6], uint8_t *mask) {
    return icp_system<double>(d1, W, H, K, mv, mn, T21, dist2, ATA, ATb, mask);
}
int or_cholesky_solve(const double A[36], const double b[6], double x[6]) { return cholesky_solve(A, b, x) ? 1 : 0; }
void or_expmap(const double x[6], double T[16]) { expmap(x, T); }
void or_rigid_inverse(const double T[16], double Ti[16]) { rigid_inverse(T, Ti); }
void or_matmul4(const double A[16], const double B[16], double C[16]) { matmul4(A, B, C); }
int or_icp_track_d(const float *const *d1, int W0, int H0, const float *K, int levels, const int iters[3],
                   const float *const *mv, const float *const *mn, float dist2, int min_valid, float max_m,
                   float max_deg, float rel_pivot, const double *T_init, double T_gc[16], int *iters_done) {
    return icp_track<double>(d1, W0, H0, K, levels, iters, mv, mn, dist2, min_valid, max_m, max_deg, rel_pivot,
                             T_init, T_gc, iters_done);
}
int or_cholesky_status(const double A[36], const double b[6], double rel, double x[6]) {
    return cholesky_status(A, b, rel, x);
}
void or_integrate_d(const uint16_t *mm, int W, int H, const float K[4], const float Tgc[16], const float Tcg[16],
                    const or_volume *vol, double *F, double *Wt, uint8_t *bm, uint8_t *upd, long long counts[2]) {
    integrate<double>(mm, W, H, K, Tgc, Tcg, vol, F, Wt, bm, upd, counts);
}
void or_integrate_f(const uint16_t *mm, int W, int H, const float K[4], const float Tgc[16], const float Tcg[16],
                    const or_volume *vol, float *F, float *Wt, uint8_t *bm, uint8_t *upd, long long counts[2]) {
    integrate<float>(mm, W, H, K, Tgc, Tcg, vol, F, Wt, bm, upd, counts);
}
void or_raycast_d(const float *F, const float *Wt, const or_volume *vol, int W, int H, const float K[4],
                  const float Tgc[16], const float Tcg[16], double *depth, double *vtx, double *nrm, uint8_t *img,
                  long long stats[4]) {
    raycast<double>(F, Wt, vol, W, H, K, Tgc, Tcg, depth, vtx, nrm, img, stats);
}
void or_integrate_color_d(const uint16_t *mm, const uint8_t *rgb, int W, int H, const float K[4], const float Tgc[16],
                          const float Tcg[16], const or_volume *vol, double *F, double *Wt, uint8_t *C,
                          uint8_t *bm, uint8_t *upd, long long counts[2]) {
    integrate_color<double>(mm, rgb, W, H, K, Tgc, Tcg, vol, F, Wt, C, bm, upd, counts);
}
void or_raycast_color_d(const float *F, const float *Wt, const uint8_t *C, const or_volume *vol, int W, int H,
                        const float K[4], const float Tgc[16], const float Tcg[16], double *depth, double *vtx,
                        double *nrm, uint8_t *img, uint8_t *tex) {
    raycast<double>(F, Wt, vol, W, H, K, Tgc, Tcg, depth, vtx, nrm, img, nullptr, C, tex);
}
void or_depth_reproject(const uint16_t *src, int sw, int sh, const float K_src[4], const float T_dst_src[16],
                        uint16_t *dst, int dw, int dh, const float K_dst[4]) {
    depth_reproject(src, sw, sh, K_src, T_dst_src, dst, dw, dh, K_dst);
}
long long or_volume_carve(double *F, double *Wt, uint8_t *C, uint8_t *brick_mask, const uint8_t *valid,
                          const or_volume *vol) {
    return volume_carve(F, Wt, C, brick_mask, valid, make_vol(vol));
}
void or_depth_crop(uint16_t *mm, int W, int H, const float K[4], const float T_cg[16], const or_volume *vol) {
    depth_crop(mm, W, H, K, T_cg, make_vol(vol));
}
void or_raycast_f(const float *F, const float *Wt, const or_volume *vol, int W, int H, const float K[4],
                  const float Tgc[16], const float Tcg[16], float *depth, float *vtx, float *nrm, uint8_t *img,
                  long long stats[4]) {
    raycast<float>(F, Wt, vol, W, H, K, Tgc, Tcg, depth, vtx, nrm, img, stats);
}

or_session *or_session_create(const or_session_desc *d, const double T_gc0[16]) {
    or_session *s = new or_session();
    s->d = *d;
    const float K0[4] = {d->fx, d->fy, d->cx, d->cy};
    intrinsics_pyramid(K0, 3, s->K);
    if (T_gc0) std::memcpy(s->Tgc, T_gc0, sizeof(s->Tgc));
    else {
        std::memset(s->Tgc, 0, sizeof(s->Tgc));
        s->Tgc[0] = s->Tgc[5] = s->Tgc[10] = s->Tgc[15] = 1.0;
    }
    const size_t N = (size_t)d->vol.dim[0] * d->vol.dim[1] * d->vol.dim[2];
    if (d->fp32) {
        s->Ff.assign(N, 1.0f);
        s->Wf.assign(N, 0.0f);
    } else {
        s->Fd.assign(N, 1.0);
        s->Wd.assign(N, 0.0);
    }
    return s;
}
void or_session_destroy(or_session *s) { delete s; }
int or_fuse_frame(or_session *s, const uint16_t *mm, double stage_ms[5]) {
    return s->d.fp32 ? fuse_frame_impl<float>(s, mm, stage_ms) : fuse_frame_impl<double>(s, mm, stage_ms);
}
void or_session_pose(const or_session *s, double T[16]) { std::memcpy(T, s->Tgc, sizeof(s->Tgc)); }
void or_session_maps(const or_session *s, int l, float *depth, float *vtx, float *nrm) {
    if (depth) std::copy(s->mdepth[l].begin(), s->mdepth[l].end(), depth);
    if (vtx) std::copy(s->mvtx[l].begin(), s->mvtx[l].end(), vtx);
    if (nrm) std::copy(s->mnrm[l].begin(), s->mnrm[l].end(), nrm);
}
void or_session_volume(const or_session *s, float *F, float *Wt) {
    if (s->d.fp32) {
        std::copy(s->Ff.begin(), s->Ff.end(), F);
        std::copy(s->Wf.begin(), s->Wf.end(), Wt);
    } else {
        for (size_t i = 0; i < s->Fd.size(); ++i) {
            F[i] = (float)s->Fd[i];
            Wt[i] = (float)s->Wd[i];
        }
    }
}
void or_session_filtered(const or_session *s, int l, float *out) {
    std::copy(s->filt[l].begin(), s->filt[l].end(), out);
}
int or_num_threads(void) { return nthreads(); }

}  // extern "C"
