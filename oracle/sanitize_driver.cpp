// Runs every stage of the oracle on a seeded synthetic stream, for a build
// with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
// `sanitize`; tests/test_oracle_sanitize.py). Exits 0 when every stage ran;
// the sanitizers abort the process on the first finding.
// TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vs_oracle.h"
#include "vscan_scene.h"

int main() {
    // the plain workload's camera over the 20-primitive scene, at 160x120 and
    // a 48^3 volume (small: the sanitizers slow everything down)
    vs_scene sc;
    vs_scene_preset(4, 1, 160, 120, &sc);
    const int W = sc.width, H = sc.height;
    or_session_desc d;
    std::memset(&d, 0, sizeof(d));
    d.width = W;
    d.height = H;
    d.fx = (float)sc.fx;
    d.fy = (float)sc.fy;
    d.cx = (float)sc.cx;
    d.cy = (float)sc.cy;
    d.levels = 3;
    d.iters[0] = 10;
    d.iters[1] = 5;
    d.iters[2] = 4;
    d.icp_dist2 = 0.055f * 0.055f;
    d.vol.dim[0] = d.vol.dim[1] = d.vol.dim[2] = 48;
    d.vol.origin[0] = d.vol.origin[1] = -0.5f;
    d.vol.origin[2] = -0.3f;
    d.vol.voxel = 1.0f / 48;
    d.vol.mu = 3.0f / 48;
    d.vol.max_w = 100.0f;
    d.bil_sigma_s = 4.0f;
    d.bil_sigma_r = 0.03f;
    d.bil_radius = 2;
    d.icp_rel_pivot = 1e-4f;
    std::vector<uint16_t> mm((size_t)W * H);
    double T0[16];
    vs_scene_pose(&sc, 0, T0);
    for (int fp32 = 0; fp32 < 2; ++fp32) {
        d.fp32 = fp32;
        or_session *s = or_session_create(&d, T0);
        for (int k = 0; k < 6; ++k) {
            vs_scene_render(&sc, k, mm.data());
            double ms[5];
            const int st = or_fuse_frame(s, mm.data(), ms);
            std::printf("fp32=%d frame %d status %d\n", fp32, k, st);
        }
        std::vector<float> dm((size_t)W * H), v((size_t)W * H * 3), n((size_t)W * H * 3), f((size_t)W * H);
        for (int l = 0; l < 3; ++l) {
            or_session_maps(s, l, dm.data(), v.data(), n.data());
            or_session_filtered(s, l, f.data());
        }
        std::vector<float> F(48 * 48 * 48), Wt(48 * 48 * 48);
        or_session_volume(s, F.data(), Wt.data());
        or_session_destroy(s);
    }
    // stand-alone operators
    vs_scene_render(&sc, 0, mm.data());
    std::vector<double> fd((size_t)W * H), vtx((size_t)W * H * 3), nrm((size_t)W * H * 3);
    or_bilateral_d(mm.data(), W, H, 4.0f, 0.03f, 3, fd.data());
    std::vector<float> ff((size_t)W * H);
    for (size_t i = 0; i < ff.size(); ++i) ff[i] = (float)fd[i];
    const float K[4] = {d.fx, d.fy, d.cx, d.cy};
    or_vnmap_d(ff.data(), W, H, K, 0, vtx.data(), nrm.data());
    or_vnmap_d(ff.data(), W, H, K, 1, vtx.data(), nrm.data());
    std::vector<double> half((size_t)(W / 2) * (H / 2));
    or_pyr_down_d(fd.data(), W, H, half.data());
    const size_t N = 48 * 48 * 48;
    std::vector<double> F(N, 1.0), Wt(N, 0.0);
    std::vector<uint8_t> bm(6 * 6 * 6), upd(N), C(3 * N, 0), rgb((size_t)3 * W * H, 128), valid(N, 1);
    long long counts[2];
    float Tf[16], Tcf[16];
    double Tc[16];
    or_rigid_inverse(T0, Tc);
    for (int k = 0; k < 16; ++k) {
        Tf[k] = (float)T0[k];
        Tcf[k] = (float)Tc[k];
    }
    or_integrate_d(mm.data(), W, H, K, Tf, Tcf, &d.vol, F.data(), Wt.data(), bm.data(), upd.data(), counts);
    or_integrate_color_d(mm.data(), rgb.data(), W, H, K, Tf, Tcf, &d.vol, F.data(), Wt.data(), C.data(), bm.data(),
                         upd.data(), counts);
    std::vector<float> F32(N), W32(N);
    for (size_t i = 0; i < N; ++i) {
        F32[i] = (float)F[i];
        W32[i] = (float)Wt[i];
    }
    std::vector<uint8_t> img((size_t)W * H), tex((size_t)3 * W * H);
    std::vector<double> depth((size_t)W * H);
    long long stats[4];
    or_raycast_d(F32.data(), W32.data(), &d.vol, W, H, K, Tf, Tcf, depth.data(), vtx.data(), nrm.data(), img.data(),
                 stats);
    or_raycast_color_d(F32.data(), W32.data(), C.data(), &d.vol, W, H, K, Tf, Tcf, depth.data(), vtx.data(), nrm.data(),
                       img.data(), tex.data());
    std::vector<uint16_t> rep((size_t)W * H);
    const float I4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    or_depth_reproject(mm.data(), W, H, K, I4, rep.data(), W, H, K);
    or_depth_crop(rep.data(), W, H, K, Tcf, &d.vol);
    valid[0] = 0;
    or_volume_carve(F.data(), Wt.data(), C.data(), bm.data(), valid.data(), &d.vol);
    double A[36] = {0}, b[6] = {0}, x[6];
    for (int i = 0; i < 6; ++i) A[i * 6 + i] = 1.0;
    or_cholesky_status(A, b, 1e-4, x);
    std::printf("sanitized oracle run complete\n");
    return 0;
}
