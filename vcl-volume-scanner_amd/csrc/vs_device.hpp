// vs_device.hpp — session layout in HBM and the FP32 device helpers shared by
// every kernel of the fusion pipeline.
//
// All HIP sources are compiled with -ffp-contract=off: the helpers below
// evaluate index- and mask-deciding expressions in FP32 in the reference's
// operation order (STAC/Util/_yooji_cuda_math.h:205-302), one IEEE rounding
// per operation, so indices and masks are bit-identical to any other
// implementation that does the same. Value arithmetic that decides nothing
// uses fmaf where it pays.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vs {

constexpr int kMaxLevels = 3;
constexpr int kBrick = 8;           // brick edge (voxels); a brick is 512 voxels, stored contiguously
constexpr int kBrickVox = 512;
constexpr int kTileW = 32;          // image tile of one 256-thread workgroup: 32 x 8 pixels
constexpr int kTileH = 8;
constexpr int kTileThreads = kTileW * kTileH;
constexpr int kIcpVals = 29;        // 21 JtJ (upper triangle) + 6 Jtr + count + squared residual
// ICP tiles summed by one workgroup, per pyramid level: a fixed grouping of
// each session's tiles, so the sums do not depend on the batch. (Level 0
// takes 8 under the ILP-maximising schedule the ICP is built with: ICP 1.89
// -> 1.84 ms per step; under the default schedule 4 was best.)
#ifndef VS_ICP_GROUP0
#define VS_ICP_GROUP0 8
#endif
#ifndef VS_ICP_GROUP1
#define VS_ICP_GROUP1 4
#endif
#ifndef VS_ICP_GROUP2
#define VS_ICP_GROUP2 4
#endif
__host__ __device__ constexpr int icp_group(int level) {
    return level == 0 ? VS_ICP_GROUP0 : (level == 1 ? VS_ICP_GROUP1 : VS_ICP_GROUP2);
}
__host__ __device__ constexpr int icp_groups(int tiles, int level) { return (tiles + icp_group(level) - 1) / icp_group(level); }
// Superbricks: cubes of 2^3 bricks (16^3 voxels). The raycast's distance
// field (below) is kept per superbrick and copied to each raycast
// workgroup's LDS, whose 16 KB hold the 32^3 superbricks of a 512^3 grid;
// smaller cells bound the band regions more tightly (a band superbrick of
// 32^3 voxels left the march ~6 brick-sized boxes per ray inside it).
constexpr int kSuperShift = 1;                   // log2(bricks per superbrick edge)
constexpr int kSuper = kBrick << kSuperShift;    // superbrick edge (voxels)
// Superbrick distance field: per superbrick the L-infinity distance (in
// superbricks, capped at kDistMax) to the nearest band superbrick, 4 bits
// each; sessions of up to kDistCells superbricks (512^3 voxels) have one.
constexpr int kDistMax = 15;
constexpr int kDistCells = 32768;

struct LevelDesc {
    int W, H;
    int tiles_x, tiles_y;
    float K[4];  // fx, fy, cx, cy
};

// Mutable per-session state; lives in HBM, touched only by kernels.
struct alignas(16) SessState {
    double T_gc[16];   // FP64 master pose (world -> camera)
    double T_cg[16];
    double T_12[16];   // tracker increment camera t0 -> t1
    double T_21[16];
    float T_gc_f[16];  // FP32 copies the kernels consume
    float T_cg_f[16];
    float T_21_f[16];
    float cam[3];      // camera centre (FP32, from T_gc_f)
    float ncc[3];      // unit vector camera -> cube centre
    float dist_cc;     // |cube centre - camera|
    float cos_max;     // raycast cone threshold
    int frames;        // frames fused so far
    int sdist_dirty;   // a superbrick band bit was set since the distance field was built
    int active;        // the session has a frame in the current step
    int do_track;      // this frame runs the tracker
    int icp_failed;    // 0 running, else a vs_status code
    int conv_level;    // level whose iterations converged (-1 none)
    int fuse_ok;       // this frame integrates and raycasts
    int last_status;   // status of the last frame
    int solves;        // Gauss-Newton solves of the last frame
    int force_single;  // T_init given: only level 0
    unsigned vis_count;         // visible bricks of the last integrate
    unsigned long long upd_count;  // voxels updated by the last integrate
};

// One element of a model vertex or normal map: x, y, z (12 bytes, AoS like
// the reference's float3 maps); -100 in x marks an invalid pixel.
struct MapXYZ {
    float x, y, z;
};
static_assert(sizeof(MapXYZ) == 12, "model map records are 3 packed floats");

// Immutable per-session parameters plus buffer pointers; one per session in
// a device array, so a launch can serve a batch of sessions.
struct alignas(16) SessParams {
    LevelDesc lev[kMaxLevels];
    int levels;
    int iters[kMaxLevels];
    float icp_dist2;
    int icp_min_valid;
    float max_motion_m, max_motion_deg;
    float icp_rel_pivot;  // relative Cholesky pivot guard (0 = off)
    int raw_f32;          // raw frame holds float metres (else uint16 mm)
    int vdim[3];        // voxels per axis
    int bdim[3];        // bricks per axis
    float origin[3];
    float vox_inv;      // 1.0f / voxel (FP32)
    float voxel;        // voxel edge (m)
    float mu, max_w;
    float cube_cen[3];
    float r_cube, step_scale, total_max;
    int max_iter;
    float bil_w[49];    // spatial weights g[i] * g[j] (the reference's FP32 product), (2r+1)^2 row-major
    float bil_nk;       // -log2(e) / (2 sigma_r^2): the range weight is 2^(bil_nk (d - c)^2)
    int bil_radius;
    // buffers (device)
    const void *raw;         // W*H depth, level-0 size: uint16 mm, or float m when raw_f32
    float *filt[kMaxLevels];
    // Model vertex and normal maps, interleaved: pixel i's vertex is record
    // 2i, its normal record 2i + 1 (-100 invalid), so the ICP's bilinear
    // taps of two neighbouring pixels are 48 contiguous bytes; one padding
    // pixel after the last keeps such a pair inside the buffer.
    MapXYZ *mvn[kMaxLevels];
    uint8_t *mimg[kMaxLevels];
    void *vox;               // brick-major voxel records (value and weight, VoxRec)
    uint32_t *bvis;          // visible-brick bits of the current frame
    uint8_t *vmask;          // visible-brick flags of the current frame, one byte per brick
                             // (zero-padded to a multiple of 128 bricks)
    uint32_t *bocc;          // bricks holding any weight > 0 (monotone)
    // Band bits (set-only, so a superset): a brick whose bit is clear holds
    // only voxels of weight 0 or value exactly 1.0, which the raycast's march
    // steps through like empty space; a superbrick (2^3 bricks = 16^3 voxels)
    // whose bit is clear holds no band brick.
    uint32_t *bband;
    uint32_t *sband;
    uint8_t *sdist;          // superbrick distance field, 2 superbricks per byte (low nibble first)
    uint8_t *sdtmp;          // its rebuild's scratch: one byte per superbrick
    int sdim[3];             // superbricks per axis
    int sband_words;         // 32-bit words of sband
    double *icp_part;        // ICP group partials (kIcpVals per group of level-0 tiles)
    // World coordinate of each voxel centre along each axis, the reference's
    // FP32 expression ((float)i + 0.5f) / vox_inv + origin[a] (d_gpw,
    // _yooji_cuda_math.h:285-302), evaluated once per session: wtab[a][i]
    const float *wtab[3];
    SessState *st;
    int tsdf_type, weight_type;
    int sid;                 // index in the context's table
    // VS_INTEGRATE_COLOR sessions (integ_mode 1); null otherwise
    int integ_mode;
    uint32_t *vcol;          // brick-major RGB (bits 0-7 R, 8-15 G, 16-23 B)
    const uint8_t *rgb;      // colour frame: R, G, B planes of the level-0 size
    uint8_t *mtex[kMaxLevels];  // texture maps: R, G, B planes per level
};

// a 4x4 matrix (or a few small vectors) passed to a kernel by value
struct Mat16f {
    float m[16];
};

struct WorkItem {
    int sess;  // index into the SessParams array of the launch
    int tile;  // tile (image stages) within the session's level
};

// A work item beyond its session's tile count (slot tables keep work lists
// sized for the largest frame so their launch grid never changes).
__device__ __forceinline__ bool tile_out(const LevelDesc &L, int tile) { return tile >= L.tiles_x * L.tiles_y; }

// ---------------------------------------------------------------------------
// FP32 helpers, reference operation order
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 scale(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 normalised(f3 v) {
    float n = 1.0f / sqrtf(dot(v, v));
    return {v.x * n, v.y * n, v.z * n};
}
__device__ __forceinline__ f3 transform(const float *T, f3 p) {
    return {T[0] * p.x + T[1] * p.y + T[2] * p.z + T[3], T[4] * p.x + T[5] * p.y + T[6] * p.z + T[7],
            T[8] * p.x + T[9] * p.y + T[10] * p.z + T[11]};
}
__device__ __forceinline__ f3 back_project(float u, float v, const float *K, float d) {
    return {d * (u - K[2]) / K[0], d * (v - K[3]) / K[1], d};
}
__device__ __forceinline__ void project(f3 p, const float *K, float &u, float &v) {
    u = K[0] * p.x / p.z + K[2];
    v = K[1] * p.y / p.z + K[3];
}
// ROUNDF (reference _yooji_cuda_defines.h:53) with out-of-range -> false.
__device__ __forceinline__ bool roundf_ref(float v, int &out) {
    if (!(v > -2147483520.0f && v < 2147483520.0f)) return false;
    out = (v < 0.0f) ? (int)(v - 0.5f) : (int)(v + 0.5f);
    return true;
}
// ---------------------------------------------------------------------------
// Address-space views. Buffer pointers held in SessParams (and any pointer
// loaded from memory) are generic, and accesses through generic pointers
// compile to FLAT instructions: 64-bit per-lane addresses, and each one
// counts against both the vector-memory and the LDS/scalar wait counters, so
// a wait for any scalar load also waits for every outstanding voxel tap.
// Every buffer the kernels touch lives in HBM, so they read and write it
// through these global-qualified views (global_load / global_store).
// ---------------------------------------------------------------------------
#define VS_GLOBAL __attribute__((address_space(1)))
#define VS_LDS __attribute__((address_space(3)))
typedef float vs_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t vs_u4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ T ldg(const T *p, size_t i) {
    return ((const VS_GLOBAL T *)p)[i];
}
template <class T>
__device__ __forceinline__ void stg(T *p, size_t i, T v) {
    ((VS_GLOBAL T *)p)[i] = v;
}
__device__ __forceinline__ float4 ldg4(const float4 *p, size_t i) {
    const vs_f4 v = ((const VS_GLOBAL vs_f4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg4(float4 *p, size_t i, float x, float y, float z, float w) {
    ((VS_GLOBAL vs_f4 *)p)[i] = vs_f4{x, y, z, w};
}
// (field by field: one 12-byte access after the compiler merges them)
__device__ __forceinline__ MapXYZ ldg3(const MapXYZ *p, size_t i) {
    const VS_GLOBAL MapXYZ *q = (const VS_GLOBAL MapXYZ *)p + i;
    return MapXYZ{q->x, q->y, q->z};
}
__device__ __forceinline__ void stg3(MapXYZ *p, size_t i, float x, float y, float z) {
    VS_GLOBAL MapXYZ *q = (VS_GLOBAL MapXYZ *)p + i;
    q->x = x;
    q->y = y;
    q->z = z;
}
// pixel i of a level's interleaved model maps
__device__ __forceinline__ MapXYZ ld_mvtx(const MapXYZ *mvn, size_t i) { return ldg3(mvn, 2 * i); }
__device__ __forceinline__ MapXYZ ld_mnrm(const MapXYZ *mvn, size_t i) { return ldg3(mvn, 2 * i + 1); }
__device__ __forceinline__ void st_mvtx(MapXYZ *mvn, size_t i, float x, float y, float z) { stg3(mvn, 2 * i, x, y, z); }
__device__ __forceinline__ void st_mnrm(MapXYZ *mvn, size_t i, float x, float y, float z) {
    stg3(mvn, 2 * i + 1, x, y, z);
}

// The top 3 rows of a row-major 4x4 FP32 matrix in HBM (16-byte aligned),
// loaded as three 16-byte vectors into registers.
struct M34 {
    float m[12];
};
__device__ __forceinline__ M34 ld_m34(const float *T) {
    M34 r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const vs_f4 v = ((const VS_GLOBAL vs_f4 *)T)[i];
        r.m[4 * i] = v.x;
        r.m[4 * i + 1] = v.y;
        r.m[4 * i + 2] = v.z;
        r.m[4 * i + 3] = v.w;
    }
    return r;
}

__device__ __forceinline__ float depth_m(uint16_t mm) { return (float)mm * 0.001f; }
// Depth in metres of raw-frame pixel i: uint16 millimetres * 0.001f (0 stays
// invalid), or a float-metre frame, where any value that is not a finite
// positive number (0, negative, NaN, infinite) reads as 0, i.e. invalid.
__device__ __forceinline__ float raw_depth(const SessParams &P, size_t i) {
    if (P.raw_f32) {
        const float v = ldg((const float *)P.raw, i);
        return (v > 0.0f && v <= 3.402823466e38f) ? v : 0.0f;
    }
    return depth_m(ldg((const uint16_t *)P.raw, i));
}
// __float2int_rn (round half to even); non-finite or |v| >= 2^31 is out of range
__device__ __forceinline__ bool rint_ref(float v, int &out) {
    if (!(v > -2147483520.0f && v < 2147483520.0f)) return false;
    out = (int)rintf(v);
    return true;
}

__device__ __forceinline__ bool world_to_voxel(const SessParams &P, f3 p, f3 &vox) {
    vox.x = (p.x - P.origin[0]) * P.vox_inv - 0.5f;
    vox.y = (p.y - P.origin[1]) * P.vox_inv - 0.5f;
    vox.z = (p.z - P.origin[2]) * P.vox_inv - 0.5f;
    return !(vox.x < -0.5f || vox.x >= (float)P.vdim[0] - 0.5f || vox.y < -0.5f || vox.y >= (float)P.vdim[1] - 0.5f ||
             vox.z < -0.5f || vox.z >= (float)P.vdim[2] - 0.5f);
}
__device__ __forceinline__ f3 voxel_to_world(const SessParams &P, f3 vox) {
    return {(vox.x + 0.5f) / P.vox_inv + P.origin[0], (vox.y + 0.5f) / P.vox_inv + P.origin[1],
            (vox.z + 0.5f) / P.vox_inv + P.origin[2]};
}

// Brick-major voxel address: brick (x>>3, y>>3, z>>3), then z,y,x inside.
__device__ __forceinline__ size_t vox_index(const SessParams &P, int x, int y, int z) {
    size_t b = ((size_t)(z >> 3) * P.bdim[1] + (y >> 3)) * P.bdim[0] + (x >> 3);
    return b * kBrickVox + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7));
}
// The same in 32 bits, for voxel-record arrays below 4 GB (the caller
// checks): the loads then take a 32-bit byte offset from the array base held
// in scalar registers instead of a 64-bit address per lane.
__device__ __forceinline__ uint32_t vox_index32(const SessParams &P, int x, int y, int z) {
    const uint32_t b = ((uint32_t)(z >> 3) * (uint32_t)P.bdim[1] + (uint32_t)(y >> 3)) * (uint32_t)P.bdim[0] +
                       (uint32_t)(x >> 3);
    return b * kBrickVox + (uint32_t)(((z & 7) << 6) | ((y & 7) << 3) | (x & 7));
}
__device__ __forceinline__ int brick_of(const SessParams &P, int x, int y, int z) {
    return ((z >> 3) * P.bdim[1] + (y >> 3)) * P.bdim[0] + (x >> 3);
}
// superbrick of brick (bx, by, bz)
__device__ __forceinline__ int super_of_brick(const SessParams &P, int bx, int by, int bz) {
    return ((bz >> kSuperShift) * P.sdim[1] + (by >> kSuperShift)) * P.sdim[0] + (bx >> kSuperShift);
}

// Voxel records. A voxel's value and weight share one record, so a lookup
// is one load and one address: 4 bytes for FP16 value + u8 weight (value in
// bits 0-15, weight in bits 16-23, bits 24-31 zero), 8 bytes for the other
// forms (FP32 value in word 0 or FP16 value in bits 0-15 of word 0; FP32
// weight in word 1 or u8 weight in bits 0-7 of word 1). The conversions are
// the ones separate planes used: FP16 values round to nearest, u8 weights
// are the truncated float.
typedef unsigned int vs_u2 __attribute__((ext_vector_type(2)));
template <class TV, class TW>
struct VoxRec {
    typedef vs_u2 T;  // 8-byte forms
    static __device__ __forceinline__ float val(T r) {
        if constexpr (sizeof(TV) == 4) return __uint_as_float(r.x);
        else return __half2float(__ushort_as_half((unsigned short)(r.x & 0xFFFFu)));
    }
    static __device__ __forceinline__ float w(T r) {
        if constexpr (sizeof(TW) == 4) return __uint_as_float(r.y);
        else return (float)(r.y & 0xFFu);
    }
    static __device__ __forceinline__ bool has_w(T r) {
        if constexpr (sizeof(TW) == 4) return __uint_as_float(r.y) > 0.0f;
        else return (r.y & 0xFFu) != 0u;
    }
    static __device__ __forceinline__ T make(float v, float wt) {
        T r;
        if constexpr (sizeof(TV) == 4) r.x = __float_as_uint(v);
        else r.x = (unsigned)__half_as_ushort(__float2half_rn(v));
        if constexpr (sizeof(TW) == 4) r.y = __float_as_uint(wt);
        else r.y = (unsigned)(uint8_t)wt;
        return r;
    }
};
template <>
struct VoxRec<__half, uint8_t> {
    typedef uint32_t T;
    static __device__ __forceinline__ float val(T r) { return __half2float(__ushort_as_half((unsigned short)(r & 0xFFFFu))); }
    static __device__ __forceinline__ float w(T r) { return (float)((r >> 16) & 0xFFu); }
    static __device__ __forceinline__ bool has_w(T r) { return (r & 0x00FF0000u) != 0u; }
    static __device__ __forceinline__ T make(float v, float wt) {
        return (uint32_t)__half_as_ushort(__float2half_rn(v)) | ((uint32_t)(uint8_t)wt << 16);
    }
};
__host__ __device__ constexpr int vox_rec_bytes(int tsdf_f16, int weight_u8) { return (tsdf_f16 && weight_u8) ? 4 : 8; }

template <class TV, class TW>
__device__ __forceinline__ typename VoxRec<TV, TW>::T ld_rec(const void *base, size_t i) {
    return ((const VS_GLOBAL typename VoxRec<TV, TW>::T *)base)[i];
}
// the same with a 32-bit byte offset from the base (records below 4 GB)
template <class TV, class TW>
__device__ __forceinline__ typename VoxRec<TV, TW>::T ld_rec32(const void *base, uint32_t i) {
    typedef typename VoxRec<TV, TW>::T T;
    return *(const VS_GLOBAL T *)((const VS_GLOBAL char *)base + i * (uint32_t)sizeof(T));
}
template <class TV, class TW>
__device__ __forceinline__ void st_rec(void *base, size_t i, typename VoxRec<TV, TW>::T r) {
    ((VS_GLOBAL typename VoxRec<TV, TW>::T *)base)[i] = r;
}
// value / weight of record i, and a record written from floats
template <class TV, class TW>
__device__ __forceinline__ float ld_val(const void *base, size_t i) {
    return VoxRec<TV, TW>::val(ld_rec<TV, TW>(base, i));
}
template <class TV, class TW>
__device__ __forceinline__ float ld_w(const void *base, size_t i) {
    return VoxRec<TV, TW>::w(ld_rec<TV, TW>(base, i));
}
template <class TV, class TW>
__device__ __forceinline__ void st_vw(void *base, size_t i, float v, float w) {
    st_rec<TV, TW>(base, i, VoxRec<TV, TW>::make(v, w));
}

}  // namespace vs
