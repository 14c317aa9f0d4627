/*
 * vs_oracle.h — CPU restatement of the reference's per-frame fusion stages.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP pipeline is
 * compared against (tests/, the smoke check and the CPU-baseline leg of the
 * benchmark); nothing in the product links or calls it.
 *
 * Two instantiations of one templated code base:
 *   *_d : the oracle. Values (filtered depth, TSDF, ICP sums, solve, pose,
 *         maps) in FP64. Expressions that decide an index or a mask (pixel and
 *         voxel indices, validity tests, march decisions) are evaluated in FP32
 *         in the reference's operation order with no contraction, so that
 *         indices and masks can be compared bit for bit.
 *   *_f : the reference algorithm in FP32 with OpenMP: the CPU baseline.
 *
 * Layouts follow the reference: images row-major, vertex/normal maps AoS
 * float3, volumes linear x-fastest (k*W*H + j*W + i), poses row-major 4x4.
 */
#ifndef VS_ORACLE_H
#define VS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim[3];
    float origin[3];
    float voxel;      /* edge (m); vox_inv = 1.0f/voxel in FP32 */
    float mu;
    float max_w;
} or_volume;

/* A1: uint16 mm -> float metres (mm * 0.001f). */
void or_depth_to_m(const uint16_t *mm, long long n, float *out);

/* A2-A3 bilateral (point taps, clamp-to-edge, any tap <= 0 -> 0). */
void or_bilateral_d(const uint16_t *mm, int W, int H, float sigma_s, float sigma_r, int radius, double *out);
void or_bilateral_f(const uint16_t *mm, int W, int H, float sigma_s, float sigma_r, int radius, float *out);
/* the 1-D spatial weights exactly as the reference computes them (FP32) */
void or_gaussian_weights(float sigma_s, int radius, float *g);

/* A4: one pyramid step W x H -> (W/2) x (H/2). */
void or_pyr_down_d(const double *in, int W, int H, double *out);
void or_pyr_down_f(const float *in, int W, int H, float *out);
/* A5: per-level intrinsics {fx,fy,cx,cy} for `levels` levels (FP32). */
void or_intrinsics_pyramid(const float K0[4], int levels, float *K_out);

/* A6-A7: vertex and normal maps (AoS float3); `inplace` = 1 reproduces the
 * reference CPU loop's in-place vertex invalidation order. */
void or_vnmap_d(const float *depth, int W, int H, const float K[4], int inplace, double *vtx, double *nrm);

/* A8-A9: normal equations of one ICP iteration. mask (nullable) gets 1 per
 * valid correspondence. Returns the valid count. */
long long or_icp_system_d(const float *depth1, int W, int H, const float K[4], const float *mvtx, const float *mnrm,
                          const float T21[16], float dist2, double ATA[36], double ATb[6], uint8_t *mask);

/* A10: reference Cholesky + solve semantics in FP64; returns 1 on success. */
int or_cholesky_solve(const double ATA[36], const double ATb[6], double x[6]);
/* The same with the build's relative pivot guard (rel > 0): returns 0 (solved),
 * 4 (a pivot below 1e-8, the reference's failure) or 8 (L_jj^2 < rel * A_jj). */
int or_cholesky_status(const double ATA[36], const double ATb[6], double rel, double x[6]);
/* A11: SE(3) exponential of x = [w; v] (FP64), rigid inverse, 4x4 product. */
void or_expmap(const double x[6], double T[16]);
void or_rigid_inverse(const double T[16], double Ti[16]);
void or_matmul4(const double A[16], const double B[16], double C[16]);

/* A12: coarse-to-fine tracking. depth1[l], mvtx[l], mnrm[l] per level; K has
 * 4 floats per level. T_gc (in/out). Returns a vs_status code (0 ok, 4 lost,
 * 5 motion rejected, 7 too few points, 8 relative pivot guard). iters_done
 * (nullable) = total solves. rel_pivot as or_cholesky_status (0 = off). */
int or_icp_track_d(const float *const *depth1, int W0, int H0, const float *K, int levels, const int iters[3],
                   const float *const *mvtx, const float *const *mnrm, float dist2, int min_valid,
                   float max_motion_m, float max_motion_deg, float rel_pivot, const double *T_init_01,
                   double T_gc[16], int *iters_done);

/* A13-A14: brick visibility + TSDF update. brick_mask and upd_mask (nullable)
 * are outputs; counts[0] = visible bricks, counts[1] = updated voxels. */
void or_integrate_d(const uint16_t *mm, int W, int H, const float K[4], const float Tgc[16], const float Tcg[16],
                    const or_volume *vol, double *F, double *Wt, uint8_t *brick_mask, uint8_t *upd_mask,
                    long long counts[2]);
void or_integrate_f(const uint16_t *mm, int W, int H, const float K[4], const float Tgc[16], const float Tcg[16],
                    const or_volume *vol, float *F, float *Wt, uint8_t *brick_mask, uint8_t *upd_mask,
                    long long counts[2]);

/* A16-A18: raycast one level. F, Wt are float volumes. stats (nullable):
 * [0] nearest taps, [1] trilinear evaluations, [2] hits, [3] rays cast. */
void or_raycast_d(const float *F, const float *Wt, const or_volume *vol, int W, int H, const float K[4],
                  const float Tgc[16], const float Tcg[16], double *depth, double *vtx, double *nrm, uint8_t *img,
                  long long stats[4]);
void or_raycast_f(const float *F, const float *Wt, const or_volume *vol, int W, int H, const float K[4],
                  const float Tgc[16], const float Tcg[16], float *depth, float *vtx, float *nrm, uint8_t *img,
                  long long stats[4]);

/* Production integrate with colour (g_ctvc, z_yooji_cuda_volume_integrator.cu:265-440,
 * registered colour frame as cdtocc(..., flag_on_rgb = true)): rgb = 3 planes of
 * W*H uint8 (R, G, B) registered to the depth frame; C = 3 planes of the
 * colour volume (uint8, linear x-fastest). Counts as or_integrate_d. */
void or_integrate_color_d(const uint16_t *mm, const uint8_t *rgb, int W, int H, const float K[4], const float Tgc[16],
                          const float Tcg[16], const or_volume *vol, double *F, double *Wt, uint8_t *C,
                          uint8_t *brick_mask, uint8_t *upd_mask, long long counts[2]);
/* Raycast with the texture map: tex = 3 planes of W*H uint8 (d_grvi at each hit,
 * z_yooji_cuda_tsdf_renderer.cu:1166-1184), 0 for misses. */
void or_raycast_color_d(const float *F, const float *Wt, const uint8_t *C, const or_volume *vol, int W, int H,
                        const float K[4], const float Tgc[16], const float Tcg[16], double *depth, double *vtx,
                        double *nrm, uint8_t *img, uint8_t *tex);

/* Input pre-steps on uint16 mm frames: z_tdrcc reprojection into the
 * destination camera (nearest depth wins, 0 where none lands) and the
 * god_Get_Object_Depth volume crop (in place). */
void or_depth_reproject(const uint16_t *src, int sw, int sh, const float K_src[4], const float T_dst_src[16],
                        uint16_t *dst, int dw, int dh, const float K_dst[4]);
void or_depth_crop(uint16_t *mm, int W, int H, const float K[4], const float T_cg[16], const or_volume *vol);

/* g_uvv volume carving: bricks of brick_mask (in/out) with no valid voxel are
 * reset (F 1, W 0, colour 0 when C is given) and leave the mask. Returns the
 * number of bricks reset. valid: one byte per voxel, linear x-fastest. */
long long or_volume_carve(double *F, double *Wt, uint8_t *C, uint8_t *brick_mask, const uint8_t *valid,
                          const or_volume *vol);

/* A19: the whole per-frame pipeline (so_Scan_Object). */
typedef struct or_session or_session;
typedef struct {
    int width, height;
    float fx, fy, cx, cy;
    int levels;
    int iters[3];
    float icp_dist2;
    int icp_min_valid;
    float max_motion_m, max_motion_deg;
    or_volume vol;
    float bil_sigma_s, bil_sigma_r;
    int bil_radius;
    int fp32; /* 1 = FP32/OpenMP baseline instantiation, 0 = FP64 oracle */
    float icp_rel_pivot; /* relative pivot guard of the solve, 0 = off */
} or_session_desc;
or_session *or_session_create(const or_session_desc *d, const double T_gc0[16]);
void or_session_destroy(or_session *s);
/* Returns the frame's status (0 = fused). stage_ms (nullable) gets 5 timings. */
int or_fuse_frame(or_session *s, const uint16_t *mm, double stage_ms[5]);
void or_session_pose(const or_session *s, double T_gc[16]);
void or_session_maps(const or_session *s, int level, float *depth, float *vtx, float *nrm);
void or_session_volume(const or_session *s, float *F, float *Wt);
void or_session_filtered(const or_session *s, int level, float *out);
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
