// Seeded analytic RGB-D scene renderer: the synthetic stand-in for the
// reference's frame capture (STAC/Util/z_yooji_frame_capture.cpp:581-633).
// Host-only C++; everything is a pure function of (scene, frame, pixel).
#include "vscan_scene.h"

#include <cmath>
#include <cstring>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in (0, 1]: 53 random bits.
inline double u01(uint64_t h) { return ((h >> 11) + 1) * (1.0 / 9007199254740992.0); }

struct Rng {
    uint64_t s;
    double next() { s = splitmix64(s); return u01(s); }
    double range(double a, double b) { return a + (b - a) * next(); }
};

void add_plane(vs_scene *sc, double half) {
    vs_prim &p = sc->prims[sc->n_prims++];
    p.kind = VS_PRIM_PLANE;
    std::memset(p.p, 0, sizeof(p.p));
    p.p[0] = half;
}
void add_sphere(vs_scene *sc, double x, double y, double z, double r) {
    vs_prim &p = sc->prims[sc->n_prims++];
    p.kind = VS_PRIM_SPHERE;
    p.p[0] = x; p.p[1] = y; p.p[2] = z; p.p[3] = r; p.p[4] = p.p[5] = 0;
}
void add_box(vs_scene *sc, double x0, double y0, double z0, double x1, double y1, double z1) {
    vs_prim &p = sc->prims[sc->n_prims++];
    p.kind = VS_PRIM_BOX;
    p.p[0] = x0; p.p[1] = y0; p.p[2] = z0; p.p[3] = x1; p.p[4] = y1; p.p[5] = z1;
}

// Base scene of the plain workloads: a 0.15 m sphere and a 0.2 x 0.1 x 0.1 m
// box resting on the ground plane.
void base_scene(vs_scene *sc) {
    add_plane(sc, 1.5);
    add_sphere(sc, 0.12, 0.06, 0.15, 0.15);
    add_box(sc, -0.30, -0.12, 0.0, -0.10, -0.02, 0.10);
}

void random_prims(vs_scene *sc, uint64_t seed, int n, double spread) {
    Rng rng{splitmix64(seed ^ 0x5CE11EULL)};
    for (int i = 0; i < n && sc->n_prims < VS_SCENE_MAX_PRIMS; ++i) {
        double x = rng.range(-spread, spread), y = rng.range(-spread, spread);
        if (rng.next() < 0.5) {
            double r = 0.5 * rng.range(0.02, 0.15);
            add_sphere(sc, x, y, r, r);
        } else {
            double dx = rng.range(0.02, 0.15), dy = rng.range(0.02, 0.15), dz = rng.range(0.02, 0.15);
            add_box(sc, x - dx / 2, y - dy / 2, 0.0, x + dx / 2, y + dy / 2, dz);
        }
    }
}

void set_camera(vs_scene *sc, int w, int h) {
    sc->width = w;
    sc->height = h;
    // fx = fy = 525 at 640x480, scaled with the width; principal point at the
    // image centre minus half a pixel (cx = W/2 - 0.5).
    sc->fx = sc->fy = 525.0 * w / 640.0;
    sc->cx = w * 0.5 - 0.5;
    sc->cy = h * 0.5 - 0.5;
}

// Ray/primitive intersection; returns the smallest t > tmin or -1.
double hit_prim(const vs_prim &p, const double o[3], const double d[3]) {
    const double tmin = 1e-6;
    if (p.kind == VS_PRIM_PLANE) {
        if (std::fabs(d[2]) < 1e-15) return -1;
        double t = -o[2] / d[2];
        if (t <= tmin) return -1;
        double x = o[0] + t * d[0], y = o[1] + t * d[1];
        if (std::fabs(x) > p.p[0] || std::fabs(y) > p.p[0]) return -1;
        return t;
    }
    if (p.kind == VS_PRIM_SPHERE) {
        double oc[3] = {o[0] - p.p[0], o[1] - p.p[1], o[2] - p.p[2]};
        double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - p.p[3] * p.p[3];
        double disc = b * b - a * c;
        if (disc < 0) return -1;
        double s = std::sqrt(disc);
        double t = (-b - s) / a;
        if (t > tmin) return t;
        t = (-b + s) / a;
        return t > tmin ? t : -1;
    }
    // box: slab method
    double t0 = -1e300, t1 = 1e300;
    for (int k = 0; k < 3; ++k) {
        double lo = p.p[k], hi = p.p[k + 3];
        if (std::fabs(d[k]) < 1e-15) {
            if (o[k] < lo || o[k] > hi) return -1;
            continue;
        }
        double a = (lo - o[k]) / d[k], b = (hi - o[k]) / d[k];
        if (a > b) { double tmp = a; a = b; b = tmp; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return -1;
    }
    if (t0 > tmin) return t0;
    return t1 > tmin ? t1 : -1;
}

// Exact camera-frame depth of pixel (u, v) for pose T_gc.
double pixel_depth(const vs_scene *sc, const double T[16], const double C[3], int u, int v) {
    double dc[3] = {(u - sc->cx) / sc->fx, (v - sc->cy) / sc->fy, 1.0};
    // world direction = R^T dc (R = upper-left 3x3 of T_gc)
    double dw[3];
    for (int k = 0; k < 3; ++k) dw[k] = T[0 * 4 + k] * dc[0] + T[1 * 4 + k] * dc[1] + T[2 * 4 + k] * dc[2];
    double best = -1;
    for (int i = 0; i < sc->n_prims; ++i) {
        double t = hit_prim(sc->prims[i], C, dw);
        if (t > 0 && (best < 0 || t < best)) best = t;
    }
    return best;  // dc.z == 1, so t is the camera-frame depth
}

void camera_centre(const double T[16], double C[3]) {
    for (int k = 0; k < 3; ++k) C[k] = -(T[0 * 4 + k] * T[3] + T[1 * 4 + k] * T[7] + T[2 * 4 + k] * T[11]);
}

void render(const vs_scene *sc, const double T[16], int frame, uint16_t *out) {
    double C[3];
    camera_centre(T, C);
    const int W = sc->width, H = sc->height;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < H; ++v) {
        for (int u = 0; u < W; ++u) {
            const int idx = v * W + u;
            double z = pixel_depth(sc, T, C, u, v);
            uint16_t mm = 0;
            if (z > 0 && z <= sc->max_range) {
                uint64_t h = splitmix64(sc->seed * 0x100000001B3ull ^ splitmix64((uint64_t)frame * 0x9E3779B1ull + (uint64_t)idx));
                double zn = z;
                if (sc->noise) {
                    double u1 = u01(h), u2 = u01(splitmix64(h ^ 0xA5A5A5A5ull));
                    double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
                    zn = z + sigma * g;
                }
                double drop = u01(splitmix64(h ^ 0x3C3C3C3C3C3Cull));
                long q = std::lround(zn * 1000.0);
                if (drop >= sc->dropout && q > 0 && q < 65536) mm = (uint16_t)q;
            }
            out[idx] = mm;
        }
    }
}

}  // namespace

extern "C" {

void vs_scene_preset(int config, uint64_t seed, int width, int height, vs_scene *sc) {
    std::memset(sc, 0, sizeof(*sc));
    sc->seed = seed;
    sc->dropout = 0.02;
    sc->noise = 1;
    sc->max_range = 4.0;
    sc->target[0] = sc->target[1] = sc->target[2] = 0.0;
    sc->phase0_deg = 0.0;
    switch (config) {
    case 0:  // plain, CPU-runnable
        set_camera(sc, width > 0 ? width : 320, height > 0 ? height : 240);
        base_scene(sc);
        sc->radius0 = sc->radius1 = 0.8;
        sc->height0 = 0.45;
        sc->height_amp = 0.0;
        sc->deg_per_frame = 1.0;
        sc->n_frames = 30;
        break;
    case 4:  // plain camera and orbit of config 0 over the 20-primitive scene
        set_camera(sc, width > 0 ? width : 320, height > 0 ? height : 240);
        base_scene(sc);
        random_prims(sc, seed, 20, 0.40);
        sc->radius0 = sc->radius1 = 0.8;
        sc->height0 = 0.45;
        sc->height_amp = 0.0;
        sc->deg_per_frame = 1.0;
        sc->n_frames = 30;
        break;
    case 1:  // plain at 640x480, helical orbit
    case 2:  // precision form: add 20 random primitives
        set_camera(sc, width > 0 ? width : 640, height > 0 ? height : 480);
        base_scene(sc);
        if (config == 2) random_prims(sc, seed, 20, 0.40);
        sc->radius0 = 0.6;
        sc->radius1 = 1.0;
        sc->height0 = 0.45;
        sc->height_amp = 0.10;
        sc->deg_per_frame = 1.0;
        sc->n_frames = 300;
        break;
    case 3:
    default: {  // batched sessions: random scene and orbit from the seed
        // the base objects plus 20 random ones keep every view of the orbit
        // constrained in all six degrees of freedom (a view of the plane and a
        // few small objects alone leaves the point-to-plane system nearly
        // rank 3, and the tracker drifts)
        set_camera(sc, width > 0 ? width : 640, height > 0 ? height : 480);
        base_scene(sc);
        random_prims(sc, seed, 20, 0.40);
        Rng rng{splitmix64(seed ^ 0x0B17ull)};
        sc->radius0 = rng.range(0.6, 0.9);
        sc->radius1 = sc->radius0 + rng.range(-0.1, 0.1);
        sc->height0 = rng.range(0.35, 0.6);
        sc->height_amp = rng.range(0.0, 0.1);
        sc->deg_per_frame = rng.range(0.6, 1.2);
        sc->phase0_deg = rng.range(0.0, 360.0);
        sc->n_frames = 100;
        break;
    }
    }
}

void vs_scene_pose(const vs_scene *sc, int frame, double T[16]) {
    const double pi = 3.14159265358979323846;
    double n = sc->n_frames > 1 ? (double)sc->n_frames : 1.0;
    double r = sc->radius0 + (sc->radius1 - sc->radius0) * frame / n;
    double h = sc->height0 + sc->height_amp * std::sin(2.0 * pi * frame / n);
    double phi = (sc->phase0_deg + sc->deg_per_frame * frame) * pi / 180.0;
    double C[3] = {r * std::cos(phi), r * std::sin(phi), h};
    double f[3] = {sc->target[0] - C[0], sc->target[1] - C[1], sc->target[2] - C[2]};
    double fn = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double &x : f) x /= fn;
    // right = forward x up(0,0,1); down = forward x right
    double rt[3] = {f[1], -f[0], 0.0};
    double rn = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1]);
    rt[0] /= rn; rt[1] /= rn;
    double dn[3] = {f[1] * rt[2] - f[2] * rt[1], f[2] * rt[0] - f[0] * rt[2], f[0] * rt[1] - f[1] * rt[0]};
    const double *rows[3] = {rt, dn, f};
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) T[i * 4 + k] = rows[i][k];
        T[i * 4 + 3] = -(rows[i][0] * C[0] + rows[i][1] * C[1] + rows[i][2] * C[2]);
    }
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

void vs_scene_render(const vs_scene *sc, int frame, uint16_t *depth_mm) {
    double T[16];
    vs_scene_pose(sc, frame, T);
    render(sc, T, frame, depth_mm);
}

void vs_scene_render_pose(const vs_scene *sc, const double T_gc[16], int frame, uint16_t *depth_mm) {
    render(sc, T_gc, frame, depth_mm);
}

void vs_scene_render_exact(const vs_scene *sc, const double T[16], double *depth_m) {
    double C[3];
    camera_centre(T, C);
    const int W = sc->width, H = sc->height;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u) {
            double z = pixel_depth(sc, T, C, u, v);
            depth_m[v * W + u] = z > 0 ? z : 0.0;
        }
}

double vs_scene_sdf(const vs_scene *sc, const double p[3]) {
    double best = 1e300;
    for (int i = 0; i < sc->n_prims; ++i) {
        const vs_prim &q = sc->prims[i];
        double d;
        if (q.kind == VS_PRIM_PLANE) {
            d = p[2];
        } else if (q.kind == VS_PRIM_SPHERE) {
            double dx = p[0] - q.p[0], dy = p[1] - q.p[1], dz = p[2] - q.p[2];
            d = std::sqrt(dx * dx + dy * dy + dz * dz) - q.p[3];
        } else {
            double e[3], out2 = 0.0, inside = -1e300;
            for (int k = 0; k < 3; ++k) {
                double c = 0.5 * (q.p[k] + q.p[k + 3]), hk = 0.5 * (q.p[k + 3] - q.p[k]);
                e[k] = std::fabs(p[k] - c) - hk;
                if (e[k] > 0) out2 += e[k] * e[k];
                if (e[k] > inside) inside = e[k];
            }
            d = out2 > 0 ? std::sqrt(out2) : inside;
        }
        if (d < best) best = d;
    }
    return best;
}

}  // extern "C"
