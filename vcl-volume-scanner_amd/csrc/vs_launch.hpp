// vs_launch.hpp — host-side launchers of the fusion kernels. Every launcher
// takes the device array of session parameters and a work list; n is the
// number of work items (one workgroup each) unless stated otherwise.
#pragma once

#include "vs_device.hpp"

namespace vs {

// image stages (vs_img.hip)
constexpr int kDepthChunk = 64;
struct DepthChunk {  // passed by value: device frame pointers and the batch slots they go to
    const uint16_t *src[kDepthChunk];
    int slot[kDepthChunk];
    int n;
};
void launch_gather_depth(const SessParams *P, const DepthChunk &ch, hipStream_t s);
// n 32-bit words at p set to v
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
void launch_bilateral(const SessParams *P, const WorkItem *work, int n, hipStream_t s);
void launch_pyr_down(const SessParams *P, const WorkItem *work, int n, int level, hipStream_t s);
void launch_vnmap(const SessParams *P, const WorkItem *work, int n, int level, float *vtx, float *nrm, hipStream_t s);
// single session (P[0]): after launch_vnmap, the reference loop's in-place order
void launch_vnmap_inplace(const SessParams *P, int level, float *vtx, float *nrm, hipStream_t s);

// tracker (vs_icp.hip); nsess = sessions in the batch
void launch_track_setup(const SessParams *P, int nsess, const double *T_init, hipStream_t s);
// slot-table form: active[s] != 0 selects the slots that fuse this step
constexpr int kMaxSlots = 256;
void launch_track_setup_active(const SessParams *P, int nsess, const int *active, hipStream_t s);
void launch_set_active(int *active, int nslots, const uint32_t bits[kMaxSlots / 32], hipStream_t s);
void launch_icp_iter(const SessParams *P, const WorkItem *work, int n, int nsess, int level, int iter,
                     hipStream_t s);
void launch_track_finalize(const SessParams *P, int nsess, hipStream_t s);
// one system, no solve: writes 29 doubles (+ optional mask) for diagnostics
void launch_icp_system(const SessParams *P, const WorkItem *work, int n, int level, const float *T21,
                       double *out_part, uint8_t *mask, hipStream_t s);
// the 29 sums of nt tile partials, in tile order, into out (one workgroup)
void launch_icp_sum_tiles(const double *part, int nt, double *out, hipStream_t s);

// volume stages (vs_vol.hip)
struct BrickList {
    int2 *items;        // (session, brick)
    unsigned *count;    // device counter
    unsigned capacity;
};
void launch_vis_clear(const SessParams *P, int nsess, BrickList bl, hipStream_t s);
// marks the visible bricks' bits, then compacts every session's bits into
// the brick list (max_words: the longest bitmap of the nsess sessions)
void launch_vis_mark(const SessParams *P, const WorkItem *work, int n, int nsess, uint32_t max_words, BrickList bl,
                     hipStream_t s);
// sd_slices: the most superbricks along z of any session (distance field)
void launch_integrate(const SessParams *P, int nsess, BrickList bl, int tsdf_type, int weight_type, int grid,
                      int sd_slices, hipStream_t s);
// o32: every session's voxel records are below 4 GB (the march and the
// shading then address them with 32-bit offsets)
void launch_raycast(const SessParams *P, const WorkItem *work, int n, int level, int tsdf_type, int weight_type,
                    bool o32, hipStream_t s);
// a single session's model maps of `level` (npx pixels) into caller buffers
// (each nullable): depth, vertex and normal (3 floats per pixel), image
void launch_maps_out(const SessParams *P, int level, size_t npx, float *depth, float *vtx, float *nrm, uint8_t *img,
                     hipStream_t s);
// Several pyramid levels in one launch: segment k of the grid is level
// level[k]'s work list (passed by value).
struct LevelWork {
    const WorkItem *w[kMaxLevels];
    int n[kMaxLevels];
    int level[kMaxLevels];
    int nseg;
    int blocks() const {
        int t = 0;
        for (int k = 0; k < nseg; ++k) t += n[k];
        return t;
    }
};
LevelWork level_work(const WorkItem *const work[kMaxLevels], const int n[kMaxLevels], int levels);
// superbrick distance field of sessions P[0..nsess) (only those that gained
// a band superbrick unless force); sd_slices >= every session's sdim[2]
void launch_sdist(const SessParams *P, int nsess, int sd_slices, int force, hipStream_t s);
// Ray block of one wave in the raycast kernels (same results for each).
enum RayBlock { kRay8x8 = 0, kRay16x4 = 1, kRay32x2 = 2 };
// o32 as for launch_raycast
void launch_raycast_march(const SessParams *P, const LevelWork &lw, int tsdf_type, int weight_type,
                          unsigned long long *stats, int ray_block, bool o32, hipStream_t s);
// waves: the waves per SIMD its registers aim for (4 or 5); o32 as for launch_raycast
void launch_raycast_shade(const SessParams *P, const LevelWork &lw, int tsdf_type, int weight_type,
                          unsigned long long *stats, int ray_block, int waves, bool o32, hipStream_t s);
// volume maintenance (single session, P points at its parameters)
void launch_volume_init(const SessParams *P, size_t nvox, int tsdf_type, int weight_type, uint32_t nwords,
                        int sd_slices, hipStream_t s);
void launch_volume_to_linear(const SessParams *P, size_t nvox, int tsdf_type, int weight_type, float *F, float *Wt,
                             hipStream_t s);
void launch_volume_from_linear(const SessParams *P, size_t nvox, int tsdf_type, int weight_type, const float *F,
                               const float *Wt, hipStream_t s);
void launch_occupancy_rebuild(const SessParams *P, uint32_t nbricks, int tsdf_type, int weight_type, int sd_slices,
                              hipStream_t s);
void launch_color_to_linear(const SessParams *P, size_t nvox, uint8_t *rgb, hipStream_t s);
// input pre-steps (vs_pre.hip); is_float selects float metres, else uint16 mm
void launch_depth_reproject(int is_float, const void *src, int sw, int sh, const float Ks[4], const float Kd[4],
                            const float T[16], uint32_t *key, void *dst, int dw, int dh, hipStream_t s);
void launch_depth_crop(const SessParams *P, int is_float, void *depth, int w, int h, const float K[4],
                       const float *Tcg, hipStream_t s);
// g_uvv carving: reset visible bricks with no valid voxel; *carved += their count
void launch_carve(const SessParams *P, int tsdf_type, int weight_type, const uint8_t *valid, unsigned *carved,
                  hipStream_t s);
// snapshots: copy the listed bricks into (unpack = 0) or out of (1) a packed buffer
void launch_brick_pack(const SessParams *P, const uint32_t *idx, unsigned n, int value_bytes, int weight_bytes,
                       int color_bytes, uint8_t *buf, int unpack, hipStream_t s);

}  // namespace vs
