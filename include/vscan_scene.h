/*
 * vscan_scene.h — seeded synthetic RGB-D depth streams.
 *
 * Replaces the reference's capture layer (CKvYooji_FrameCapture,
 * STAC/Util/z_yooji_frame_capture.cpp:581-633, and the OpenNI/Kinect grabber),
 * which has no input data in the repository: frames are rendered from an
 * analytic scene (ground plane, spheres, axis-aligned boxes) seen from a
 * camera on a circular or helical orbit, with the axial noise model
 * sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, millimetre quantisation and
 * random dropout. Everything is deterministic in (seed, frame, pixel).
 *
 * World frame: z up, ground plane z = 0. Camera frame: x right, y down,
 * z forward. Depth is the camera-frame z of the first hit, in millimetres.
 */
#ifndef VSCAN_SCENE_H
#define VSCAN_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VS_PRIM_PLANE = 0, VS_PRIM_SPHERE = 1, VS_PRIM_BOX = 2 };

typedef struct {
    int kind;     /* VS_PRIM_* */
    double p[6];  /* plane: half extent in x/y (p[0]); sphere: centre xyz, radius;
                     box: min xyz, max xyz */
} vs_prim;

#define VS_SCENE_MAX_PRIMS 64

typedef struct {
    int width, height;
    double fx, fy, cx, cy;
    int n_prims;
    vs_prim prims[VS_SCENE_MAX_PRIMS];
    /* orbit: horizontal radius goes linearly radius0 -> radius1 over n_frames,
       height = height0 + height_amp * sin(2*pi*frame/n_frames) */
    double radius0, radius1, height0, height_amp;
    double deg_per_frame, phase0_deg;
    double target[3];
    int n_frames;
    /* sensor */
    uint64_t seed;
    double dropout;   /* probability a valid pixel reads 0 */
    int noise;        /* 0 = exact depth (still quantised to mm) */
    double max_range; /* metres; farther hits read 0 */
} vs_scene;

/* Presets mirroring the workloads: config 0 = 320x240 sphere+box scene on a
 * 0.8 m circular orbit (1 deg/frame, 30 frames); 1 = same scene at 640x480 on
 * a helical orbit (radius 0.6 -> 1.0 m, +-10 cm, 300 frames); 2 = config 1
 * plus 20 random boxes/spheres of 2-15 cm; 3 = a random-primitive scene and
 * orbit drawn from `seed` (batched sessions), at width x height; 4 = the
 * camera and orbit of config 0 over the scene of config 2. */
void vs_scene_preset(int config, uint64_t seed, int width, int height, vs_scene *out);
/* Ground-truth world->camera pose of `frame`, row-major 4x4. */
void vs_scene_pose(const vs_scene *sc, int frame, double T_gc[16]);
/* Render one noisy uint16 mm frame (OpenMP over rows). */
void vs_scene_render(const vs_scene *sc, int frame, uint16_t *depth_mm);
/* Render with an explicit pose instead of the orbit (for warp tests). */
void vs_scene_render_pose(const vs_scene *sc, const double T_gc[16], int frame, uint16_t *depth_mm);
/* Exact (noise-free, unquantised) depth in metres, 0 where nothing is hit. */
void vs_scene_render_exact(const vs_scene *sc, const double T_gc[16], double *depth_m);
/* Signed distance from a world point to the scene surface (negative inside
 * an object or below the plane); used to build analytic TSDF volumes. */
double vs_scene_sdf(const vs_scene *sc, const double p[3]);

#ifdef __cplusplus
}
#endif
#endif /* VSCAN_SCENE_H */
