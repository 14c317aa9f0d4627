// vs_pre.hip — input pre-steps around the filter: reprojection of a depth
// frame into another camera (z_tdrcc) and the crop of depths outside the
// object volume (god). One thread per pixel; both are memory-bound streams.
#include "vs_device.hpp"
#include "vs_launch.hpp"

namespace vs {

// z_tdrcc_Transform_Depth_to_Rgb_Camera_Coordinates (z_yooji_object_scanner.cpp:2429-2476)
// with convert_pixel_coord (z_yooji_matrix_rgbd.cpp:101-118): back-project the
// source pixel (CKvYooji_Intrinsics::back_project, z_yooji_camera_parameters.cpp:104-108),
// transform into the destination camera (CKvYooji_Extrinsics::transform, :678-685),
// project (:89-96), ROUNDF, and keep the nearest source depth per destination
// pixel. The value written is the source depth itself, so the operator is
// exact on integer millimetres; the nearest one is an atomic minimum on its
// bit pattern (non-negative values order like their bits), which makes the
// result independent of the order of the writes.
template <class T>
__device__ __forceinline__ float depth_value(T v);
template <>
__device__ __forceinline__ float depth_value<uint16_t>(uint16_t v) { return depth_m(v); }
template <>
__device__ __forceinline__ float depth_value<float>(float v) { return v; }
__device__ __forceinline__ uint32_t depth_key(uint16_t v) { return v; }
__device__ __forceinline__ uint32_t depth_key(float v) { return __float_as_uint(v); }

template <class T>
__global__ void k_depth_reproject(const T *__restrict__ src, int sw, int sh, Mat16f Ks4, Mat16f Tds, uint32_t *key,
                                  int dw, int dh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sw * sh) return;
    const T raw = src[i];
    const float td = depth_value<T>(raw);
    if (!(td > 0.0f)) return;
    const float *Ks = Ks4.m, *Kd = Ks4.m + 4, *M = Tds.m;
    const float x = (float)(i % sw), y = (float)(i / sw);
    const f3 p = {td * (x - Ks[2]) / Ks[0], td * (y - Ks[3]) / Ks[1], td};
    const f3 q = transform(M, p);
    const float u = Kd[0] * q.x / q.z + Kd[2], v = Kd[1] * q.y / q.z + Kd[3];
    int tx, ty;
    if (!roundf_ref(u, tx) || !roundf_ref(v, ty)) return;
    if (tx < 0 || tx > dw - 1 || ty < 0 || ty > dh - 1) return;
    atomicMin(key + (size_t)ty * dw + tx, depth_key(raw));
}

__global__ void k_depth_key_fill(uint32_t *key, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        key[i] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void key_to_depth(uint32_t k, uint16_t &out) { out = k == 0xFFFFFFFFu ? 0 : (uint16_t)k; }
__device__ __forceinline__ void key_to_depth(uint32_t k, float &out) {
    out = k == 0xFFFFFFFFu ? 0.0f : __uint_as_float(k);
}
template <class T>
__global__ void k_depth_key_finish(const uint32_t *key, T *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        key_to_depth(key[i], dst[i]);
}

template <class T>
static void reproject_t(const void *src, int sw, int sh, const float Ks[4], const float Kd[4], const float M[16],
                        uint32_t *key, void *dst, int dw, int dh, hipStream_t s) {
    Mat16f k{}, m{};
    for (int i = 0; i < 4; ++i) {
        k.m[i] = Ks[i];
        k.m[4 + i] = Kd[i];
    }
    for (int i = 0; i < 16; ++i) m.m[i] = M[i];
    const size_t nd = (size_t)dw * dh, ns = (size_t)sw * sh;
    hipLaunchKernelGGL(k_depth_key_fill, dim3(1024), dim3(256), 0, s, key, nd);
    hipLaunchKernelGGL(k_depth_reproject<T>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, (const T *)src, sw,
                       sh, k, m, key, dw, dh);
    hipLaunchKernelGGL(k_depth_key_finish<T>, dim3(1024), dim3(256), 0, s, key, (T *)dst, nd);
}
void launch_depth_reproject(int is_float, const void *src, int sw, int sh, const float Ks[4], const float Kd[4],
                            const float T[16], uint32_t *key, void *dst, int dw, int dh, hipStream_t s) {
    if (is_float)
        reproject_t<float>(src, sw, sh, Ks, Kd, T, key, dst, dw, dh, s);
    else
        reproject_t<uint16_t>(src, sw, sh, Ks, Kd, T, key, dst, dw, dh, s);
}

// god_Get_Object_Depth (z_yooji_object_extractor.cpp:357-410), its cube test:
// back-project the integer pixel with its depth, map it to the world with
// the camera-to-world transform (CKvYooji_Extrinsics::transform_inv,
// z_yooji_camera_parameters.cpp:687-694) and zero the depth when the point
// leaves [origin, origin + size] on any axis. T_cg is the session's FP32
// pose inverse unless one is given.
template <class T>
__global__ void k_depth_crop(const SessParams *__restrict__ SP, T *depth, int w, int h, Mat16f K4, Mat16f Tcg,
                             int use_session_pose) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const SessParams &P = SP[0];
    const float td = depth_value<T>(depth[i]);
    const float *K = K4.m;
    const float *M = use_session_pose ? P.st->T_cg_f : Tcg.m;
    const f3 p = {td * ((float)(i % w) - K[2]) / K[0], td * ((float)(i / w) - K[3]) / K[1], td};
    const f3 g = transform(M, p);
    // size_cube = dim * voxel (z_yooji_object_extractor.cpp:388), per axis
    const float sx = (float)P.vdim[0] * P.voxel, sy = (float)P.vdim[1] * P.voxel, sz = (float)P.vdim[2] * P.voxel;
    if (g.x < P.origin[0] || g.x > P.origin[0] + sx || g.y < P.origin[1] || g.y > P.origin[1] + sy ||
        g.z < P.origin[2] || g.z > P.origin[2] + sz)
        depth[i] = T(0);
}
void launch_depth_crop(const SessParams *P, int is_float, void *depth, int w, int h, const float K[4],
                       const float *Tcg, hipStream_t s) {
    Mat16f k{}, m{};
    for (int i = 0; i < 4; ++i) k.m[i] = K[i];
    if (Tcg)
        for (int i = 0; i < 16; ++i) m.m[i] = Tcg[i];
    const unsigned blocks = (unsigned)(((size_t)w * h + 255) / 256);
    if (is_float)
        hipLaunchKernelGGL(k_depth_crop<float>, dim3(blocks), dim3(256), 0, s, P, (float *)depth, w, h, k, m,
                           Tcg ? 0 : 1);
    else
        hipLaunchKernelGGL(k_depth_crop<uint16_t>, dim3(blocks), dim3(256), 0, s, P, (uint16_t *)depth, w, h, k, m,
                           Tcg ? 0 : 1);
}

}  // namespace vs
