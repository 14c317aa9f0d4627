/*
 * vscan.h — C ABI of the MI355X-native per-frame depth-fusion pipeline.
 *
 * This is the drop-in boundary for the scanner's RGB-D GPU engine: the four
 * operator classes and three containers that the reference drives once per
 * frame (reference: VCL_Volume_Scanner/STAC/Engine/GPU/_yooji_cuda_img_proc.h:2-29,
 * _yooji_cuda_pose_tracker.h:2-42, _yooji_cuda_volume_integrator.h:2-52,
 * _yooji_cuda_tsdf_renderer.h:2-96; containers STAC/Object/GPU/_yooji_cuda_object_cube.h:9-101,
 * _yooji_cuda_tracking_state.h:4-168, _yooji_cuda_rgbd_frame.h:4-91).
 *
 * Differences from the reference boundary, all deliberate:
 *   - opaque handles instead of C++ classes; no process-global state (the
 *     reference keeps camera/volume parameters in per-translation-unit
 *     __constant__ singletons, so it can run one session per process);
 *   - every call takes an explicit stream (a hipStream_t passed as void*;
 *     NULL selects the session's own stream) and is asynchronous unless it
 *     says otherwise;
 *   - status codes instead of bool;
 *   - the Gauss-Newton loop of the tracker runs on the device: no host round
 *     trip per iteration, the pose stays resident in HBM (FP64 master copy).
 *
 * Conventions (same as the reference):
 *   - depth input is uint16 millimetres, converted to metres as mm * 0.001f
 *     (STAC/Util/z_yooji_frame_capture.cpp:597,626), 0 = invalid, or float
 *     metres as the reference's operators take them (vs_depth_units);
 *   - T_gc maps world ("global") points into the camera frame, row-major 4x4;
 *     T_cg is its inverse;
 *   - voxel (i,j,k) has its centre at origin + (i+0.5)*voxel; the linear index
 *     used by vs_volume_download/upload is k*W*H + j*W + i (x fastest);
 *   - invalid vertex/normal = (-100,-100,-100), invalid depth = 0.
 */
#ifndef VSCAN_H
#define VSCAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vs_ctx vs_ctx;
typedef struct vs_session vs_session;
typedef void *vs_stream; /* hipStream_t */

typedef enum {
    VS_OK = 0,
    VS_EINVAL = 1,          /* bad argument or descriptor */
    VS_ENOMEM = 2,          /* device allocation failed */
    VS_EHIP = 3,            /* a HIP runtime call failed */
    VS_TRACK_LOST = 4,      /* Cholesky pivot below 1e-8 (or non-finite): pose kept, frame not fused */
    VS_MOTION_REJECTED = 5, /* optional motion gate fired: pose kept, frame not fused */
    VS_BUDGET = 6,          /* HBM budget exhausted (serving admission) */
    VS_TOO_FEW_POINTS = 7,  /* optional minimum-correspondence gate fired */
    VS_SINGULAR = 8         /* optional relative pivot guard fired (icp_rel_pivot): the normal equations
                               are nearly rank deficient (e.g. one plane in view); pose kept, frame not fused */
} vs_status;

typedef enum { VS_TSDF_F32 = 0, VS_TSDF_F16 = 1 } vs_tsdf_type;
typedef enum { VS_WEIGHT_F32 = 0, VS_WEIGHT_U8 = 1 } vs_weight_type;
/* Integration operator:
 *   VS_INTEGRATE_DEPTH: g_ctv_Compute_TSDF_of_Voxels (z_yooji_cuda_volume_integrator.cu:154-262),
 *     the depth operator of cdtoc (:936-972);
 *   VS_INTEGRATE_COLOR: the production g_ctvc_Compute_TSDF_of_Voxels_with_Color (:265-440) as
 *     cdtocc (:976-1031) drives it from so_Scan_Object (z_yooji_object_scanner.cpp:1274-1281)
 *     with a colour frame registered to the depth frame: SDF along the pixel ray,
 *     voxels at the weight cap frozen, an RGB running average per voxel, and a
 *     rendered texture map per pyramid level (tsdf_renderer.cu:1166-1184). */
typedef enum { VS_INTEGRATE_DEPTH = 0, VS_INTEGRATE_COLOR = 1 } vs_integrate_mode;
/* Units of the session's depth frames (every entry point that takes a depth
 * frame of the session's size reads it in these units):
 *   VS_DEPTH_MM_U16: uint16 millimetres, converted as mm * 0.001f (the capture
 *     path, z_yooji_frame_capture.cpp:597,626); 0 = invalid;
 *   VS_DEPTH_M_F32: float metres, as the reference's operators take them
 *     (bfd_Bilateral_Filter_Depth's host_in, cdtoc's depth_dev); any value that
 *     is not a finite positive number (0, negative, NaN, infinite) is invalid.
 * Frames passed through the uint16_t pointers of vs_set_depth, vs_fuse_frame,
 * vs_fuse_batch and vs_server_step are read in the session's units (cast a
 * float frame's pointer). */
typedef enum { VS_DEPTH_MM_U16 = 0, VS_DEPTH_M_F32 = 1 } vs_depth_units;

/* One scan session: camera, pyramid, tracker, volume and filter settings.
 * Replaces CKvYooji_Scanner_Parameter (STAC/Object/CPU/_yooji_scanner_parameter.h:4-71)
 * plus the per-operator ip_Initialize_Parameters calls
 * (z_yooji_cuda_pose_tracker.cu:735-785, z_yooji_cuda_volume_integrator.cu:822-869,
 *  z_yooji_cuda_tsdf_renderer.cu:1577-1639) and the bilateral globals
 * (z_yooji_cuda_img_proc.cu:42-45). */
typedef struct {
    int width, height;           /* level-0 depth frame size (pixels) */
    float fx, fy, cx, cy;        /* level-0 intrinsics; level l uses fx/2^l, c <- (c-0.5)/2 per level */
    int levels;                  /* pyramid levels, 1..3 */
    int iters[3];                /* Gauss-Newton iterations per level, index 0 = finest */
    float icp_dist2;             /* correspondence gate on squared point distance (m^2) */
    int icp_min_valid;           /* minimum valid correspondences per solve (0 = no gate, as the GPU reference) */
    float max_motion_m;          /* optional motion gate: translation limit (m), <= 0 disables */
    float max_motion_deg;        /* optional motion gate: rotation limit (deg), <= 0 disables */
    int vol_dim[3];              /* voxels per axis; each a multiple of 8 */
    float origin[3];             /* world position of the volume's min corner (m) */
    float voxel;                 /* voxel edge length (m) */
    float mu;                    /* truncation distance (m) */
    float max_weight;            /* running-average weight cap */
    int tsdf_type;               /* vs_tsdf_type */
    int weight_type;             /* vs_weight_type */
    float bil_sigma_s;           /* bilateral spatial sigma (pixels) */
    float bil_sigma_r;           /* bilateral range sigma (m) */
    int bil_radius;              /* bilateral half window (1..3) */
    int integrate_mode;          /* vs_integrate_mode */
    int depth_units;             /* vs_depth_units of the session's depth frames */
    float icp_rel_pivot;         /* relative pivot guard of the ICP solve: VS_SINGULAR when a squared
                                    Cholesky pivot L_jj^2 < icp_rel_pivot * A_jj; 0 = off (the reference
                                    tests only the absolute pivot, L_jj < 1e-8) */
} vs_session_desc;

/* Fill *d with the reference's defaults for a W x H camera and an N^3 cube of
 * edge `cube_m` metres: 3 levels, iterations {10,7,7}, mu = 5*voxel... Callers
 * normally overwrite the fields their workload names. */
void vs_session_desc_default(vs_session_desc *d, int width, int height, int n, float cube_m);

/* ---- context ---------------------------------------------------------- */
vs_status vs_ctx_create(int device, vs_ctx **out);
vs_status vs_ctx_destroy(vs_ctx *ctx);
const char *vs_status_string(int status);
const char *vs_last_error(void); /* thread-local text of the last failure */

/* ---- sessions (GKvObjectCube + GKvTrackingState + GKvRgbdFrame) -------- */
/* Device bytes one session holds (volume, masks, frame and model buffers). */
size_t vs_session_bytes(const vs_session_desc *desc);
/* Allocate a session with TSDF = 1, weight = 0 (z_yooji_object_scanner.cpp:343-347)
 * and pose T_gc0 (NULL = identity). */
vs_status vs_session_create(vs_ctx *ctx, const vs_session_desc *desc, const double T_gc0[16], vs_session **out);
vs_status vs_session_destroy(vs_session *s);
/* Clear the volume and frame counter and set the pose again (stream-ordered). */
vs_status vs_session_reset(vs_session *s, const double T_gc0[16], vs_stream stream);
vs_status vs_session_desc_get(const vs_session *s, vs_session_desc *out);
/* Block until all work queued on the session's own stream has finished. */
vs_status vs_session_sync(vs_session *s);

/* ---- per-stage operators ---------------------------------------------- */
/* Copy one uint16 mm frame into the session's raw-frame buffer. */
vs_status vs_set_depth(vs_session *s, const uint16_t *depth_mm, int is_host, vs_stream stream);
/* VS_INTEGRATE_COLOR sessions: copy the colour frame registered to the depth
 * frame, 3 planes (R, G, B) of width*height uint8 as the reference's
 * g_img_color_t1_, used by the next integration (stream-ordered: give the
 * stream of the next step, or synchronise). */
vs_status vs_set_color(vs_session *s, const uint8_t *rgb_planar, int is_host, vs_stream stream);
/* A2-A3: bilateral filter of the raw frame into pyramid level 0 (metres).
 * Replaces LGKvImageProcessor::bfd_Bilateral_Filter_Depth (z_yooji_cuda_img_proc.cu:394-417). */
vs_status vs_bilateral(vs_session *s, vs_stream stream);
/* A2-A3 with the reference operator's buffers
 * (LGKvImageProcessor::bfd_Bilateral_Filter_Depth(float *dev_out, const float *host_in, w, h),
 * z_yooji_cuda_img_proc.cu:394-417, and its _Host variant :420-442): depth_in is
 * a frame in the session's depth units, on the host when in_is_host, else on
 * the device; NULL filters the frame already in the session's raw buffer. The
 * result goes to pyramid level 0 and, when depth_out_dev is not NULL, also to
 * that device buffer of width*height floats (metres). */
vs_status vs_bilateral_filter_depth(vs_session *s, const void *depth_in, int in_is_host, float *depth_out_dev,
                                    vs_stream stream);
/* A4: levels 1..L-1 from level 0 (5x5 binomial blur + decimation), as
 * cpdd_Construct_Pyramids_Down_Depth (z_yooji_image_processor.cpp:368-407). */
vs_status vs_pyramid(vs_session *s, vs_stream stream);
/* A6-A7: vertex and normal maps of the frame at `level` into device buffers
 * of W_l*H_l*3 floats each (cpm_Compute_Point_Map / enm_Estimate_Normal_Map,
 * z_yooji_image_processor.cpp:119-227). Not used by vs_fuse_frame, which
 * back-projects inline like the reference GPU tracker. */
vs_status vs_vertex_normal_maps(vs_session *s, int level, float *vtx_dev, float *nrm_dev, vs_stream stream);
/* A6-A7 in the reference CPU loop's in-place order (enm_Estimate_Normal_Map,
 * z_yooji_image_processor.cpp:155-227): a grazing pixel's vertex becomes -100
 * before the pixels after it in raster order are processed, so a pixel whose
 * left or upper neighbour was invalidated gets no normal (and, if grazing
 * itself, keeps its vertex). vs_vertex_normal_maps evaluates every pixel on
 * the input map instead (order-independent); the two differ only next to
 * grazing pixels. */
vs_status vs_vertex_normal_maps_inplace(vs_session *s, int level, float *vtx_dev, float *nrm_dev,
                                        vs_stream stream);
/* A8-A12: projective point-to-plane ICP of the frame pyramid against the
 * model maps of the previous raycast; updates the resident pose.
 * Replaces LGKvPoseTracker::tp_Track_Pose (z_yooji_cuda_pose_tracker.cu:790-1042).
 * T_01_init (nullable, row-major, maps camera t0 into camera t1) forces a
 * single level-0 pass, as the reference's in_pose_init_t0_t1. The status of
 * the solve is reported by vs_get_frame_status after the stream completes. */
vs_status vs_track(vs_session *s, const double *T_01_init, vs_stream stream);
/* A13-A15: brick visibility + TSDF update with the raw frame at the current pose.
 * Replaces LGKvVolumeIntegrator::cdtoc_Convert_Depth_to_TSDF_on_Cube
 * (z_yooji_cuda_volume_integrator.cu:936-972) with the per-frame brick-mask
 * clear of cdtocc (:1006-1010). */
vs_status vs_integrate(vs_session *s, vs_stream stream);
/* A13-A15 on a caller's device frame, as cdtoc_Convert_Depth_to_TSDF_on_Cube(cube,
 * const float *depth_dev, w, h, T_gc_dev, T_cg_dev) (z_yooji_cuda_volume_integrator.cu:936-972):
 * depth_dev (session's size and depth units; NULL = the raw buffer) is copied
 * into the raw buffer, then integrated at the session's current pose. */
vs_status vs_integrate_depth(vs_session *s, const void *depth_dev, vs_stream stream);
/* A16-A18: raycast the volume at `level` into the session's model maps.
 * Replaces LGKvRendererTSDF::rmi_Render_Maps_for_ICP (z_yooji_cuda_tsdf_renderer.cu:1642-1700). */
vs_status vs_raycast(vs_session *s, int level, vs_stream stream);
/* A16-A18 into caller-owned device buffers, as rmi_Render_Maps_for_ICP(cube, w, h,
 * T_gc_dev, T_cg_dev, cam_cen, light, level, io_map_depth, io_map_vertex,
 * io_map_normals, io_img_normals) (z_yooji_cuda_tsdf_renderer.cu:1642-1700): raycasts
 * `level` at the session's current pose into its model maps (which the next
 * track reads) and writes them to the buffers given (each may be NULL):
 * depth W_l*H_l floats (camera z, 0 = miss), vertex and normal W_l*H_l*3
 * floats (camera frame, -100 = miss), shaded normal image W_l*H_l bytes.
 * Asynchronous on `stream`. */
vs_status vs_render_maps(vs_session *s, int level, float *depth_dev, float *vtx_dev, float *nrm_dev,
                         uint8_t *img_dev, vs_stream stream);
/* A19: one fused frame (filter -> pyramid -> track (frames > 0) -> integrate
 * raw depth -> raycast every level), as so_Scan_Object
 * (z_yooji_object_scanner.cpp:472-629). Asynchronous; see vs_get_frame_status.
 * depth_mm NULL fuses the frame already in the session's raw buffer
 * (vs_set_depth or a pre-step). */
vs_status vs_fuse_frame(vs_session *s, const uint16_t *depth_mm, int is_host, vs_stream stream);
/* Many sessions, one frame each, in one batched launch per stage. Sessions
 * must share tsdf/weight types. depths[i] are device pointers unless is_host;
 * a NULL depths[i] (or depths NULL) fuses the frame already in that session's
 * raw buffer. Host frames are copied before the call returns (the caller may
 * reuse its buffers at once) and uploaded by one transfer on a copy stream of
 * the library's own, which overlaps the previous step's GPU work. */
vs_status vs_fuse_batch(vs_ctx *ctx, vs_session *const *sessions, const uint16_t *const *depths,
                        int n, int is_host, vs_stream stream);

/* ---- input pre-steps ------------------------------------------------------
 * The CPU steps either side of the filter in the reference's input path, on
 * the device. Depth buffers are device pointers: uint16 millimetres, or float
 * metres when *_is_float. A NULL frame pointer selects the session's own
 * raw-frame buffer (the session's size, intrinsics and depth units, which
 * *_is_float must then match), which a following vs_fuse_frame /
 * vs_fuse_batch / vs_server_step with a NULL depth fuses as it is. */
/* z_tdrcc_Transform_Depth_to_Rgb_Camera_Coordinates (z_yooji_object_scanner.cpp:2429-2476, with
 * CKvYooji_MatrixRgbD::convert_pixel_coord, z_yooji_matrix_rgbd.cpp:101-118): forward-project the
 * source frame (src_w x src_h, intrinsics K_src) through T_dst_src (row-major, source camera to the
 * session's camera) into a frame of the session's size and intrinsics, keeping the nearest source
 * depth per pixel and 0 where none lands. dst_is_float must equal src_is_float (the depth value is
 * moved, not converted). */
vs_status vs_depth_reproject(vs_session *s, const void *src, int src_w, int src_h, const float K_src[4],
                             const float T_dst_src[16], int src_is_float, void *dst, vs_stream stream);
/* god_Get_Object_Depth (z_yooji_object_extractor.cpp:357-410), its volume test: zero, in place,
 * every depth whose back-projection (intrinsics K) lies outside the session's volume box
 * [origin, origin + dim*voxel] in the world, using T_cg (camera to world, row-major) or, when NULL,
 * the session's current pose as the device holds it (stream-ordered). The hand-colour
 * segmentation of that function is not part of this operator. */
vs_status vs_depth_crop_to_volume(vs_session *s, void *depth, int w, int h, const float K[4], const float *T_cg,
                                  int is_float, vs_stream stream);

/* ---- state access (synchronous unless noted) --------------------------- */
/* Status of the last tracked frame (vs_status value) and frames fused so far. */
vs_status vs_get_frame_status(vs_session *s, int *last_status, int *frames_fused, int *icp_iters);
vs_status vs_get_pose(vs_session *s, double T_gc[16]);
vs_status vs_set_pose(vs_session *s, const double T_gc[16], vs_stream stream);
/* Pyramid level `level` of the filtered frame (W_l*H_l floats, metres). */
vs_status vs_get_filtered(vs_session *s, int level, float *host);
vs_status vs_set_filtered(vs_session *s, int level, const float *host);
/* Model maps written by vs_raycast: depth W_l*H_l, vertex/normal W_l*H_l*3
 * (AoS, camera frame), shaded normal image W_l*H_l. Any pointer may be NULL. */
vs_status vs_get_maps(vs_session *s, int level, float *depth, float *vtx, float *nrm, uint8_t *img);
/* VS_INTEGRATE_COLOR sessions: the texture map of the last raycast at `level`,
 * 3 planes of W_l*H_l uint8 (d_grvi at each hit, 0 elsewhere). */
vs_status vs_get_texture(vs_session *s, int level, uint8_t *rgb_planar);
vs_status vs_set_maps(vs_session *s, int level, const float *vtx, const float *nrm);
/* Whole volume in the reference's linear layout; tsdf/weight converted to
 * float. brick_mask: (W/8)(H/8)(D/8) bytes of the last frame's visibility. */
vs_status vs_volume_download(vs_session *s, float *tsdf, float *weight, uint8_t *brick_mask);
vs_status vs_volume_upload(vs_session *s, const float *tsdf, const float *weight);
/* VS_INTEGRATE_COLOR sessions: the colour volume, 3 planes of W*H*D uint8 in
 * the linear layout (the reference's GKvObjectCube rgb planes). */
vs_status vs_volume_download_color(vs_session *s, uint8_t *rgb_planar);
/* Normal equations of one ICP iteration at `level` for the float pose
 * T_21 (camera t1 -> camera t0): 36 + 6 doubles and the valid count, plus an
 * optional W_l*H_l validity mask. Diagnostic and parity entry point. */
vs_status vs_icp_system(vs_session *s, int level, const float T_21[16], double ATA[36], double ATb[6],
                        long long *valid, uint8_t *mask_host);
/* vs_icp_system without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops): no counterpart
 * in the reference. The tile partials live in a fixed device array of the
 * library (8192 tiles of 32x8 pixels, 1.9 MB per GPU, held until the process
 * ends); a level with more tiles (over 2,097,152 pixels) is refused with
 * VS_EINVAL. Asynchronous on `stream` (NULL: the session's stream), ordered
 * after earlier work on that stream. sums_dev: 29 doubles on the device, JtJ's
 * upper triangle row by row (21), Jtr (6), the valid count and the sum of
 * squared residuals, bit-identical to vs_icp_system's; mask_dev (nullable):
 * W_l*H_l bytes on the device, 1 per valid correspondence. Every call on one
 * GPU shares the array, so calls that overlap on a GPU (other streams, other
 * threads, other contexts) race. */
vs_status vs_icp_system_noalloc(vs_session *s, int level, const float T_21[16], double *sums_dev,
                                uint8_t *mask_dev, vs_stream stream);
/* Per-stage kernel times of the last vs_fuse_frame on this session (ms):
 * [bilateral, pyramid, track, integrate, raycast]. Valid after sync when
 * timing was enabled with vs_session_set_timing. */
vs_status vs_session_set_timing(vs_session *s, int enable);
vs_status vs_get_stage_ms(vs_session *s, float ms[5]);
/* Same for batches of this context (vs_fuse_batch): when enabled, steps are
 * launched eagerly with events between kernels, and vs_ctx_stage_ms returns the
 * last step's [track reset + bilateral, pyramid, ICP, brick visibility,
 * integrate, raycast march (all levels, one launch), raycast shading (all
 * levels, one launch), 0] in ms (waits for the step). */
vs_status vs_ctx_set_timing(vs_ctx *ctx, int enable);
vs_status vs_ctx_stage_ms(vs_ctx *ctx, float ms[8]);

/* Launch-shape autotuning (no counterpart in the reference, whose launches
 * have fixed shapes). Three launches of a step have interchangeable shapes:
 * the ray block of a wave in the raycast march (8x8, 16x4 or 32x2 rays), the
 * ray block (8x8 or 32x2) and register budget (4 or 5 waves per SIMD) of
 * its shading, and the workgroup count of the integrate (1024 to 32768);
 * every shape gives bit-identical results. With
 * VS_TUNE_ONLINE, a batch (vs_fuse_batch, a server's slot table, a session's
 * own vs_fuse_frame) whose size class (largest frame, largest volume edge
 * and session count rounded up to powers of two, storage types) has no
 * choice yet runs its first 12 steps eagerly with the candidates in turn,
 * timed per stage (those steps ignore vs_ctx_set_timing), then keeps the
 * fastest shape per stage for that class (a ray block other than the fixed
 * one only when it is at least 3% faster: the tuning frames are the first
 * ones, with volumes still filling). VS_TUNE_CACHED applies the shapes
 * settled so far (in this context or read from the cache file) and tunes
 * nothing. VS_TUNE_OFF (the default) tunes nothing and applies nothing: batches
 * not settled earlier use fixed shapes (march 8x8, shade 8x8, integrate 8192
 * workgroups, 2048 for a session's own steps). cache_path (nullable) names a
 * text file the choices are read from now and written to after each tuning. */
typedef enum { VS_TUNE_OFF = 0, VS_TUNE_ONLINE = 1, VS_TUNE_CACHED = 2 } vs_tune_mode;
vs_status vs_ctx_set_tuning(vs_ctx *ctx, int mode, const char *cache_path);
/* The context's settled shapes, one size class per line:
 * "<class>|<march block>|<shade block>|<integrate workgroups>|<means>", where
 * <means> is "m8x8,m16x4,m32x2|s8x8,s32x2,s8x8w5,s32x2w5|i1024,i2048,i4096,i8192,i16384,i32768" (mean stage
 * ms per candidate, -1 for one not timed) or "cached" (from the file). Writes
 * at most len bytes including the terminating zero; *needed (nullable) = the
 * full length + 1. */
vs_status vs_ctx_tuning(vs_ctx *ctx, char *buf, size_t len, size_t *needed);
/* Counts from the last integrate: visible bricks and updated voxels. */
vs_status vs_get_integrate_counts(vs_session *s, long long *visible_bricks, long long *updated_voxels);
/* Tap statistics of one instrumented raycast at `level` with the current pose
 * and volume, marching every step as the reference does (synchronous):
 *   [0] nearest taps (d_gtvu), [1] trilinear evaluations (d_gtvi), [2] hits,
 *   [3] nearest taps outside the grid, [4] nearest taps of voxels without weight,
 *   [5] bytes the nearest taps request, [6] bytes the trilinear evaluations
 *   request, [7] bytes the hits' 32-tap normals (d_csnt) request.
 * Bytes follow the reference's lookups (z_yooji_cuda_tsdf_renderer.cu:68-153,288-464):
 * nothing outside the grid, the weight alone for a voxel without weight,
 * weight + value otherwise, and a stencil stops at its first failing tap.
 * They price the raycast's algorithmic bytes; no counterpart in the reference. */
vs_status vs_raycast_counts(vs_session *s, int level, unsigned long long counts[8]);

/* Volume carving: g_uvv_Update_Valid_Volume (z_yooji_cuda_volume_integrator.cu:718-799,
 * LGKvVolumeIntegrator::uvv_Update_Valid_Volume :1035-1060). valid = one byte per voxel in the
 * linear x-fastest layout (the reference's GKvVolumeBool, e.g. a visual hull), device pointer
 * unless is_host. Every brick of the last frame's visibility mask with no valid voxel is reset to
 * TSDF 1, weight 0, colour 0 and leaves the mask. *carved (nullable) = bricks reset
 * (synchronous). */
vs_status vs_volume_carve(vs_session *s, const uint8_t *valid, int is_host, unsigned *carved);

/* ---- snapshots: save / restore / evict ----------------------------------
 * Replaces itsdfcd_Import_TSDF_Cube_from_Device + stsdfc_Save_TSDF_Cube
 * (z_yooji_object_scanner.cpp:2687-2745), whose .vcl files use the closed
 * KAISION format, with an open sparse format. Layout (little endian):
 *   vs_snapshot_header
 *   uint32 brick index [bricks_stored]      (brick (bx,by,bz) = (bz*BH + by)*BW + bx)
 *   values  [bricks_stored][512]            (value_bytes each: FP32 or FP16)
 *   weights [bricks_stored][512]            (weight_bytes each: FP32 or uint8)
 *   colours [bricks_stored][512]            (color_bytes = 4: R, G, B, 0; absent when 0)
 * Inside a brick voxels are ordered z, y, x (x fastest). Bricks not stored
 * hold TSDF 1, weight 0, colour 0 (the initial state). A restored session
 * continues exactly as the saved one would have: its model maps are
 * re-rendered at the saved pose. */
typedef struct {
    char magic[8];           /* "VSCANSNP" */
    uint32_t version;        /* 2 */
    uint32_t header_bytes;   /* sizeof(vs_snapshot_header) */
    vs_session_desc desc;
    double T_gc[16];
    int32_t frames;          /* frames fused so far */
    uint32_t bricks_total;
    uint32_t bricks_stored;
    uint32_t value_bytes, weight_bytes, color_bytes;
} vs_snapshot_header;
/* Bytes of the session's snapshot now (synchronous). */
vs_status vs_session_snapshot_bytes(vs_session *s, size_t *bytes);
/* Write the snapshot into buf (capacity bytes); *bytes = bytes written, or the
 * size needed when buf is NULL or too small (VS_EINVAL then). Synchronous. */
vs_status vs_session_snapshot(vs_session *s, void *buf, size_t capacity, size_t *bytes);
/* A new session of ctx from a snapshot (synchronous). */
vs_status vs_session_restore(vs_ctx *ctx, const void *buf, size_t bytes, vs_session **out);

/* ---- serving: slot table with continuous batching ------------------------
 * No counterpart in the reference (one scan per process there). A server owns
 * up to max_slots sessions of one storage form. vs_server_open admits a
 * session when a slot is free and the bytes of all open sessions stay within
 * hbm_budget (0 = no limit); otherwise it returns VS_BUDGET and allocates
 * nothing. vs_server_step fuses one frame for each listed slot in one
 * graph-captured launch sequence; the other open slots keep their volume, pose
 * and maps untouched. Opening or closing a slot does not wait for the steps in
 * flight: its effects (slot parameters, work lists in the second of two
 * buffers, freeing a closed session's memory) are enqueued in stream order
 * after them, and the next step captures the new launch sequence once; steps
 * between changes replay it with a per-step active mask. Each session's
 * result is the same as
 * vs_fuse_frame on it alone. Sessions stay usable through vs_server_session
 * (pose, maps, volume download); closing destroys them. */
typedef struct vs_server vs_server;
typedef struct {
    int max_slots;          /* 1..256 */
    int tsdf_type;          /* vs_tsdf_type of every session */
    int weight_type;        /* vs_weight_type of every session */
    unsigned long long hbm_budget; /* bytes (vs_session_bytes summed), 0 = no limit */
} vs_server_desc;
vs_status vs_server_create(vs_ctx *ctx, const vs_server_desc *desc, vs_server **out);
vs_status vs_server_destroy(vs_server *srv);
vs_status vs_server_open(vs_server *srv, const vs_session_desc *desc, const double T_gc0[16], int *slot);
vs_status vs_server_close(vs_server *srv, int slot);
vs_session *vs_server_session(vs_server *srv, int slot);
/* depths[i] is slot slots[i]'s frame (device pointer unless is_host; host
 * frames as for vs_fuse_batch); each slot at most once per step. Asynchronous
 * on `stream` (NULL: the server's). */
vs_status vs_server_step(vs_server *srv, const int *slots, const uint16_t *const *depths, int n, int is_host,
                         vs_stream stream);
/* Open slots and the session bytes they hold. */
vs_status vs_server_usage(vs_server *srv, int *open_slots, unsigned long long *bytes);
/* Eviction and resume: vs_server_evict writes the slot's snapshot into buf
 * (as vs_session_snapshot) and closes the slot; vs_server_open_snapshot
 * admits a session from a snapshot under the same slot and budget rules as
 * vs_server_open. */
vs_status vs_server_evict(vs_server *srv, int slot, void *buf, size_t capacity, size_t *bytes);
vs_status vs_server_open_snapshot(vs_server *srv, const void *buf, size_t bytes, int *slot);

#ifdef __cplusplus
}
#endif
#endif /* VSCAN_H */
