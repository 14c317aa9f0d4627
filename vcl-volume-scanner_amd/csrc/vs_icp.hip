// vs_icp.hip — projective point-to-plane ICP with the Gauss-Newton loop on
// the device (A8-A12).
//
// Reference: LGKvPoseTracker::tp_Track_Pose (z_yooji_cuda_pose_tracker.cu:790-1042)
// launches one summand+reduction kernel per iteration, copies ~200 KB of
// block partials to the host, sums them in FP32, solves on the host and copies
// the pose back. Here each iteration is two launches: k_icp_iter writes one
// FP64 partial of the 29 sums per group of tiles, then k_icp_solve (one
// workgroup per session) sums them in FP64 in a fixed order, solves the 6x6
// system (Cholesky, FP64), applies the SE(3) update and publishes the FP32
// pose the next iteration reads. Convergence (|x|_inf < 5e-7) and failure are
// flags the following launches test, so the host enqueues a fixed launch
// sequence that can be captured in a graph.
#include "vs_device.hpp"
#include "vs_launch.hpp"

namespace vs {

struct Mat16d {
    double m[16];
};

// ---------------------------------------------------------------------------
// FP64 pose algebra (one thread)
// ---------------------------------------------------------------------------
__device__ void matmul4d(const double *A, const double *B, double *C) {
    double t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s = fma(A[r * 4 + k], B[k * 4 + c], s);
            t[r * 4 + c] = s;
        }
    for (int k = 0; k < 16; ++k) C[k] = t[k];
}
// Inverse of a rigid transform: [R^T, -R^T t]. Equals the reference's general
// 4x4 inverse (d_im_Inverse_Matrix_4x4, _yooji_math_inlines.h:404-465) on
// rigid inputs.
__device__ void rigid_inverse(const double *T, double *Ti) {
    double o[16];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) o[r * 4 + c] = T[c * 4 + r];
        o[r * 4 + 3] = -(T[0 * 4 + r] * T[3] + T[1 * 4 + r] * T[7] + T[2 * 4 + r] * T[11]);
    }
    o[12] = o[13] = o[14] = 0.0;
    o[15] = 1.0;
    for (int k = 0; k < 16; ++k) Ti[k] = o[k];
}
// SE(3) exponential of x = [w; v] (the twist update of
// z_uitme_Update_Incremental_Tracking_with_Matrix_Exponential, pose_tracker.cu:50-149),
// with a series branch for small |w| (the reference divides by |w|).
__device__ void expmap(const double *x, double *T) {
    const double w0 = x[0], w1 = x[1], w2 = x[2];
    const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
    double A, B, C;
    if (th < 1e-4) {
        A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        C = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        A = sin(th) / th;
        B = (1.0 - cos(th)) / th2;
        C = (th - sin(th)) / (th2 * th);
    }
    const double W[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
    double W2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) W2[r * 3 + c] = W[r * 3] * W[c] + W[r * 3 + 1] * W[3 + c] + W[r * 3 + 2] * W[6 + c];
    for (int r = 0; r < 3; ++r) {
        double t = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double I = (r == c) ? 1.0 : 0.0;
            T[r * 4 + c] = I + A * W[r * 3 + c] + B * W2[r * 3 + c];
            t += (I + B * W[r * 3 + c] + C * W2[r * 3 + c]) * x[3 + c];
        }
        T[r * 4 + 3] = t;
    }
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}
// Cholesky and triangular solves with the reference's pivot rule
// (d_lld_LL_Decomposition / d_sls_Solve_Linear_System_using_LLD,
// _yooji_cuda_math.h:489-592): entries below a pivot < 1e-8 are zeroed and the
// solve fails (status 4) if any pivot is below 1e-8 (or not a number). With
// rel > 0 it also fails (status 8) when a squared pivot L_jj^2 is below
// rel * A_jj: column j of J is then nearly a combination of the columns
// before it (a plane leaves 3 of the 6 motions unconstrained), a near-rank
// deficiency the absolute test cannot see at FP32 scales. The ratio is
// invariant to the units of the rotation and translation columns.
// Column by column, each below-diagonal entry and substitution step scaled by
// the column's reciprocal pivot (6 FP64 divisions instead of 33 on the
// solve's one-lane chain; the same sums in the same order).
// Returns 0 on success.
__device__ int cholesky_solve(const double *A, const double *b, double rel, double *x) {
    double L[36], inv[6];
    for (int k = 0; k < 36; ++k) L[k] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < j; ++k) s += L[j * 6 + k] * L[j * 6 + k];
        L[j * 6 + j] = sqrt(A[j * 6 + j] - s);
        inv[j] = 1.0 / L[j * 6 + j];
        const bool piv = fabs(L[j * 6 + j]) >= 1e-8;
        for (int i = j + 1; i < 6; ++i) {
            double t = 0.0;
            for (int k = 0; k < j; ++k) t += L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = piv ? (A[i * 6 + j] - t) * inv[j] : 0.0;
        }
    }
    for (int i = 0; i < 6; ++i)
        if (!(fabs(L[i * 6 + i]) >= 1e-8)) return 4;
    if (rel > 0.0)
        for (int i = 0; i < 6; ++i)
            if (L[i * 6 + i] * L[i * 6 + i] < rel * A[i * 6 + i]) return 8;
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int k = 0; k < i; ++k) s += L[i * 6 + k] * y[k];
        y[i] = (b[i] - s) * inv[i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = 0.0;
        for (int k = i + 1; k < 6; ++k) s += L[k * 6 + i] * x[k];
        x[i] = (y[i] - s) * inv[i];
    }
    return 0;
}

__device__ void publish_T21(SessState *st) {
    rigid_inverse(st->T_12, st->T_21);
    for (int k = 0; k < 16; ++k) st->T_21_f[k] = (float)st->T_21[k];
}

// ---------------------------------------------------------------------------
// per-pixel summand (d_cssfp_Compute_Single_Summand_with_Forward_Params,
// pose_tracker.cu:279-395; bilinear taps d_giv/d_gin, _yooji_cuda_math.h:365-481)
// ---------------------------------------------------------------------------
// The 2x2 neighbourhood of (u, v) in a model map. A tap takes part only when
// its weight is positive (rx > 0 for the right column, ry > 0 for the lower
// row); a participating tap at -100 rejects the pixel. The maps interleave
// vertex and normal records, so a row's two pixels x0, x0 + 1 are 48
// contiguous bytes (vertex, normal, vertex, normal), read as three 16-byte
// loads: an unused right column (rx == 0) is read and ignored (x0 + 1 <= W,
// and the padding pixel after the last covers the last row), an unused
// lower row re-reads the upper one. Every address is inside the buffer, so
// all taps are loaded together, before any test.
struct Quad {
    MapXYZ a, b, c, d;
};
typedef vs_f4 vs_f4_a4 __attribute__((aligned(4)));  // 24-byte records: 8-byte aligned
struct Row {
    vs_f4 q0, q1, q2;  // vertex x0 | normal x0 | vertex x0 + 1 | normal x0 + 1, 3 floats each
};
__device__ __forceinline__ Row load_row(const MapXYZ *mvn, size_t i) {
    const VS_GLOBAL vs_f4_a4 *p = (const VS_GLOBAL vs_f4_a4 *)(mvn + 2 * i);
    return {p[0], p[1], p[2]};
}
__device__ __forceinline__ Quad vertex_quad(const Row &r0, const Row &r1) {
    return {{r0.q0.x, r0.q0.y, r0.q0.z}, {r0.q1.z, r0.q1.w, r0.q2.x}, {r1.q0.x, r1.q0.y, r1.q0.z},
            {r1.q1.z, r1.q1.w, r1.q2.x}};
}
__device__ __forceinline__ Quad normal_quad(const Row &r0, const Row &r1) {
    return {{r0.q0.w, r0.q1.x, r0.q1.y}, {r0.q2.y, r0.q2.z, r0.q2.w}, {r1.q0.w, r1.q1.x, r1.q1.y},
            {r1.q2.y, r1.q2.z, r1.q2.w}};
}
// Taps are picked by selects, not branches, so that no tap's load is sunk
// into a branch of its own (and waited for alone).
__device__ __forceinline__ bool bilerp(const Quad &q, float rx, float ry, f3 &out) {
    const bool ux = rx > 0.0f, uy = ry > 0.0f, uxy = ux & uy;
    const bool ok = (q.a.x != -100.0f) & (!ux | (q.b.x != -100.0f)) & (!uy | (q.c.x != -100.0f)) &
                    (!uxy | (q.d.x != -100.0f));
    const f3 t1 = {q.a.x, q.a.y, q.a.z};
    const f3 t2 = {ux ? q.b.x : 0.0f, ux ? q.b.y : 0.0f, ux ? q.b.z : 0.0f};
    const f3 t3 = {uy ? q.c.x : 0.0f, uy ? q.c.y : 0.0f, uy ? q.c.z : 0.0f};
    const f3 t4 = {uxy ? q.d.x : 0.0f, uxy ? q.d.y : 0.0f, uxy ? q.d.z : 0.0f};
    const float ix = 1.0f - rx, iy = 1.0f - ry;
    out = add(scale(iy, add(scale(ix, t1), scale(rx, t2))), scale(ry, add(scale(ix, t3), scale(rx, t4))));
    return ok;
}

// The summand of pixel (x, y) as a branch-free sequence: every test of the
// reference is evaluated (same FP32 expressions, so the same decisions) and
// folded into `ok`; the model taps use coordinates clamped into the map, so
// every load is in range whatever the tests say, and a rejected pixel's
// Jacobian and residual are zero. In two halves, so that a caller can issue
// the next pixel's model taps (icp_pre) before it evaluates this pixel's
// (icp_post); the operations are the same either way.
struct IcpPre {
    f3 p;            // the pixel's vertex in the model's frame
    float rx, ry;    // bilinear weights of the taps
    bool ok;         // depth valid, projection inside the map
    Row r0, r1;      // model vertex and normal taps of the two rows (loads in flight)
};
// td1: the filtered depth at (x, y), loaded by the caller
__device__ __forceinline__ IcpPre icp_pre(const SessParams &P, int l, int x, int y, float td1, const float *T21) {
    const LevelDesc &L = P.lev[l];
    const int W = L.W, H = L.H;
    IcpPre r;
    r.p = transform(T21, back_project((float)x, (float)y, L.K, td1));
    float u, v;
    project(r.p, L.K, u, v);
    const bool inside = u >= 0.0f && u <= float(W - 1) && v >= 0.0f && v <= float(H - 1);
    r.ok = !(td1 < 1.0e-6f) && inside;
    const float uc = inside ? u : 0.0f, vc = inside ? v : 0.0f;
    const int x0 = (int)uc, y0 = (int)vc;
    r.rx = uc - (float)x0;
    r.ry = vc - (float)y0;
    const size_t i0 = (size_t)y0 * W + x0, dj = r.ry > 0.0f ? (size_t)W : 0;
    r.r0 = load_row(P.mvn[l], i0);
    r.r1 = load_row(P.mvn[l], i0 + dj);
    return r;
}
__device__ __forceinline__ bool icp_post(const SessParams &P, const IcpPre &r, float A[6], float &b) {
    const f3 p = r.p;
    bool ok = r.ok;
    f3 v0, n;
    ok &= bilerp(vertex_quad(r.r0, r.r1), r.rx, r.ry, v0);
    const float dx = v0.x - p.x, dy = v0.y - p.y, dz = v0.z - p.z;
    ok &= !(dx * dx + dy * dy + dz * dz > P.icp_dist2);
    ok &= bilerp(normal_quad(r.r0, r.r1), r.rx, r.ry, n);
    n = normalised(n);
    A[0] = ok ? p.z * n.y + (-p.y) * n.z : 0.0f;
    A[1] = ok ? (-p.z) * n.x + p.x * n.z : 0.0f;
    A[2] = ok ? p.y * n.x + (-p.x) * n.y : 0.0f;
    A[3] = ok ? -n.x : 0.0f;
    A[4] = ok ? -n.y : 0.0f;
    A[5] = ok ? -n.z : 0.0f;
    b = ok ? n.x * dx + n.y * dy + n.z * dz : 0.0f;
    return ok;
}
__device__ __forceinline__ bool icp_summand_d(const SessParams &P, int l, int x, int y, float td1, const float *T21,
                                              float A[6], float &b) {
    return icp_post(P, icp_pre(P, l, x, y, td1, T21), A, b);
}
__device__ __forceinline__ bool icp_summand(const SessParams &P, int l, int x, int y, const float *T21, float A[6],
                                            float &b) {
    return icp_summand_d(P, l, x, y, ldg(P.filt[l], (size_t)y * P.lev[l].W + x), T21, A, b);
}

// Adds one pixel's 29 terms (JtJ upper triangle, Jtr, count, r^2) to the
// thread's FP32 sums; a rejected pixel has A = 0, b = 0 and adds zeros.
__device__ __forceinline__ void add_terms(bool ok, const float A[6], float b, float acc[kIcpVals]) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j, ++k) acc[k] = fmaf(A[i], A[j], acc[k]);
#pragma unroll
    for (int i = 0; i < 6; ++i, ++k) acc[k] = fmaf(A[i], b, acc[k]);
    acc[k++] += ok ? 1.0f : 0.0f;
    acc[k] = fmaf(b, b, acc[k]);
}

// Block sum of the 29 terms through LDS: every thread stores its 29 FP32
// sums (term-major rows of 257 words, so that the stores of a wave and the
// reads below spread over the LDS banks), 8 threads per term each add a fixed
// chunk of 32 threads' values in FP32 (in a fixed order), and the 8 chunk sums
// of a term are added in FP64 in chunk order. Result in red64 (shared), valid
// after the call for all threads. (29 wave-wide DPP trees cost ~11 VALU
// instructions per term and wave; this costs ~4 per term and thread.)
constexpr int kTermRow = kTileThreads + 1;
constexpr int kChunks = 8, kChunk = kTileThreads / kChunks;
__device__ __forceinline__ void block_terms(float v[kIcpVals], double *red64) {
    __shared__ float rows[kIcpVals * kTermRow];
    __shared__ double part[kIcpVals][kChunks];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kIcpVals; ++k) rows[k * kTermRow + t] = v[k];
    __syncthreads();
    if (t < kIcpVals * kChunks) {
        const int k = t / kChunks, c = t % kChunks;
        const float *row = rows + k * kTermRow + c * kChunk;
        float s = 0.0f;
#pragma unroll 8
        for (int i = 0; i < kChunk; ++i) s += row[(i + 4 * c) & (kChunk - 1)];
        part[k][c] = (double)s;
    }
    __syncthreads();
    if (t < kIcpVals) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) s += part[t][c];
        red64[t] = s;
    }
    __syncthreads();
}

// VS_ICP_MFMA (1 in the default build; build.py --variant icp_lds builds the
// per-thread FMA sums and LDS chunk sums instead): the block's 29 sums from
// the matrix cores. For each tile of the group, each pixel's row
// x = [A (6), b, ok] goes to LDS (2 KB per wave) and the wave runs
// v_mfma_f32_16x16x4f32 over its 64 rows, 8 pixels per instruction:
// the 16x16 product packs two 8x8 Gram blocks (two sets of 4 pixels) on its
// diagonal, with x read as both operands (lane l supplies feature l % 8 of
// pixel 4 * ((l % 16) / 8) + l / 16 of the instruction's 8), so
// G = sum x x^T holds J^T J, J^T r, sum r^2 and the count. The two diagonal
// blocks are added across lanes (lane l and l + 40 hold matching entries),
// and the 4 waves' blocks are added in FP64 in wave order. Without the 29
// per-thread accumulators and the 30 KB of LDS rows the kernel holds 64
// registers and 9 KB of LDS, so 8 waves per SIMD are resident instead of 5:
// the iteration waits on its dependent gathers (depth, then the model taps),
// and the extra waves hide them (ICP stage 2.61 -> 2.22 ms per 64-session
// step, with the same VALU work and bytes fetched).
#ifndef VS_ICP_MFMA
#define VS_ICP_MFMA 1
#endif
#if VS_ICP_MFMA
typedef float vs_f4x __attribute__((ext_vector_type(4)));
// entry (i, j) of the 8x8 Gram block that holds term k of the 29
__device__ __forceinline__ void term_ij(int k, int &i, int &j) {
    if (k < 21) {
        int r = 0, n = 6;
        while (k >= n) {
            k -= n;
            ++r;
            --n;
        }
        i = r;
        j = r + k;
    } else if (k < 27) {
        i = k - 21;
        j = 6;
    } else if (k == 27) {
        i = j = 7;
    } else {
        i = j = 6;
    }
}
#endif

__device__ __forceinline__ bool icp_skip(const SessParams &P, const SessState *st, int level, int iter) {
    return !ldg(&st->do_track, 0) || ldg(&st->icp_failed, 0) || ldg(&st->conv_level, 0) == level ||
           iter >= P.iters[level] || (ldg(&st->force_single, 0) && level != 0);
}

// One Gauss-Newton iteration at `level`, first half: the 29 sums of every
// group of G = icp_group(level) consecutive 32x8 tiles of every session of the work
// list (items are (session, group)), one FP64 partial per group, summed on
// the matrix cores (above) or, with VS_ICP_MFMA=0, by per-thread FP32
// accumulation over the group's tiles and one block reduction (block_terms).
// No cross-workgroup synchronisation here: on a
// multi-XCD device an agent-scope release per workgroup would write back L2
// each time, so the reduction and the solve are the next launch
// (k_icp_solve).
template <int G>
__global__ __launch_bounds__(kTileThreads) void k_icp_iter(const SessParams *__restrict__ SP,
                                                           const WorkItem *__restrict__ work, int level, int iter) {
    __shared__ float T21s[16];
    __shared__ double red64[kIcpVals];
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    const SessState *st = P.st;
    if (level >= P.levels || icp_skip(P, st, level, iter)) return;
    const LevelDesc &L = P.lev[level];
    const int ntiles = L.tiles_x * L.tiles_y;
    const int t0 = wi.tile * G;
    if (t0 >= ntiles) return;
    if (threadIdx.x < 16) T21s[threadIdx.x] = ldg(st->T_21_f, threadIdx.x);
    __syncthreads();
#if !VS_ICP_MFMA
    float acc[kIcpVals];
#pragma unroll
    for (int k = 0; k < kIcpVals; ++k) acc[k] = 0.0f;
#endif
    // the group's depths first (independent loads; a pixel outside the
    // frame re-reads pixel 0 and is skipped), then the pixels in order
    int xs[G], ys[G];
    bool in[G];
    float td[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int tile = t0 + g;
        xs[g] = (tile % L.tiles_x) * kTileW + threadIdx.x % kTileW;
        ys[g] = (tile / L.tiles_x) * kTileH + threadIdx.x / kTileW;
        in[g] = tile < ntiles && xs[g] < L.W && ys[g] < L.H;
        td[g] = ldg(P.filt[level], in[g] ? (size_t)ys[g] * L.W + xs[g] : (size_t)0);
    }
#if VS_ICP_MFMA
    __shared__ float feat[kTileThreads / 64][64][8];
    __shared__ float gw[kTileThreads / 64][8][8];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = ((lane & 15) >> 3) * 4 + (lane >> 4), fi = lane & 7;
    vs_f4x g4 = {0.0f, 0.0f, 0.0f, 0.0f};
    // one pixel ahead: the next tile's model taps are in flight while this
    // tile's summand is evaluated (a pixel outside the frame has depth 0 and
    // adds zeros)
    IcpPre nxt = icp_pre(P, level, xs[0], ys[0], in[0] ? td[0] : 0.0f, T21s);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        float A[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, b = 0.0f;
        const IcpPre cur = nxt;
        if (g + 1 < G) nxt = icp_pre(P, level, xs[g + 1], ys[g + 1], in[g + 1] ? td[g + 1] : 0.0f, T21s);
        const bool ok = icp_post(P, cur, A, b);
        // the wave's 64 rows of this tile. A wave reads only its own rows, and
        // the LDS performs one wave's accesses in the order they are issued,
        // so no workgroup barrier is needed between the writes and the reads
        // (nor before the next tile's writes); the compiler keeps their order
        // (the rows may alias) and wave_barrier keeps it from moving them.
        __builtin_amdgcn_wave_barrier();
        float *f = feat[wv][lane];
#pragma unroll
        for (int k = 0; k < 6; ++k) f[k] = A[k];
        f[6] = b;
        f[7] = ok ? 1.0f : 0.0f;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const float a = feat[wv][8 * m + px][fi];
            g4 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, a, g4, 0, 0, 0);
        }
    }
    // lane l holds D[4 (l / 16) + r][l % 16]; the second block's entry
    // D[i + 8][j + 8] of the first block's D[i][j] sits in lane l + 40
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float other = __shfl_down(g4[r], 40, 64);
        if ((lane >> 4) < 2 && (lane & 15) < 8) gw[wv][4 * (lane >> 4) + r][lane & 15] = g4[r] + other;
    }
    __syncthreads();
    if (threadIdx.x < kIcpVals) {
        int i, j;
        term_ij(threadIdx.x, i, j);
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kTileThreads / 64; ++w) s += (double)gw[w][i][j];
        stg(P.icp_part, (size_t)wi.tile * kIcpVals + threadIdx.x, s);
    }
#else
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (in[g]) {
            float A[6], b;
            const bool ok = icp_summand_d(P, level, xs[g], ys[g], td[g], T21s, A, b);
            add_terms(ok, A, b, acc);
        }
    }
    block_terms(acc, red64);
    if (threadIdx.x < kIcpVals) stg(P.icp_part, (size_t)wi.tile * kIcpVals + threadIdx.x, red64[threadIdx.x]);
#endif
}

// Second half: one wave per session sums the group partials in FP64 in a
// fixed order, solves the 6x6 system and applies the SE(3) update. Lane t
// adds groups t, t + 64, ...; then the 29 sums (padded to 32) are reduced
// across the wave by recursive halving: at each of five exchanges a lane
// sends half of the sums it still holds to its partner (lane ^ 32, ^ 16, ...,
// ^ 2) and adds the partner's half of its own, so 16 + 8 + 4 + 2 + 1 values
// cross lanes instead of 29 x 6 for 29 separate butterflies; a last exchange
// with lane ^ 1 completes sum v = lane / 2, and lane 0 reads the 29 totals
// with readlane. The order is fixed, so the result does not depend on the
// batch or the schedule. (A 256-thread block with a butterfly per sum spent
// about twice as long in this part.)
constexpr int kSolveThreads = 64;
__global__ __launch_bounds__(kSolveThreads) void k_icp_solve(const SessParams *__restrict__ SP, int nsess, int level,
                                                             int iter) {
    const int s = blockIdx.x;
    if (s >= nsess) return;
    const SessParams &P = SP[s];
    SessState *st = P.st;
    if (level >= P.levels || icp_skip(P, st, level, iter)) return;
    const LevelDesc &L = P.lev[level];
    const int ngroups = icp_groups(L.tiles_x * L.tiles_y, level);
    const int lane = threadIdx.x;
    double acc[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) acc[k] = 0.0;
    for (int t = lane; t < ngroups; t += kSolveThreads) {
        const double *pp = P.icp_part + (size_t)t * kIcpVals;
#pragma unroll
        for (int k = 0; k < kIcpVals; ++k) acc[k] += ldg(pp, k);
    }
#pragma unroll
    for (int h = 16; h >= 1; h >>= 1) {  // partner lane ^ 2h; keep the upper h sums when that bit is set
        const bool up = (lane & (2 * h)) != 0;
#pragma unroll
        for (int i = 0; i < h; ++i) {
            const double send = up ? acc[i] : acc[h + i];
            const double keep = up ? acc[h + i] : acc[i];
            acc[i] = keep + __shfl_xor(send, 2 * h, 64);
        }
    }
    const double own = acc[0] + __shfl_xor(acc[0], 1, 64);  // sum (lane / 2) over all 64 lanes
    double tot[kIcpVals];
    {
        const int lo = __double2loint(own), hi = __double2hiint(own);
#pragma unroll
        for (int k = 0; k < kIcpVals; ++k)
            tot[k] = __hiloint2double(__builtin_amdgcn_readlane(hi, 2 * k), __builtin_amdgcn_readlane(lo, 2 * k));
    }
    if (lane != 0) return;
    double ATA[36], ATb[6], x6[6];
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            ATA[i * 6 + j] = ATA[j * 6 + i] = tot[k++];
        }
    for (int i = 0; i < 6; ++i) ATb[i] = tot[k++];
    const double cnt = tot[21 + 6];
    // state through global-qualified accesses (a chain of generic accesses
    // in one thread waits on both the vector-memory and the scalar counters)
    double T12[16];
    for (int i = 0; i < 16; ++i) T12[i] = ldg(st->T_12, i);
    stg(&st->solves, 0, ldg(&st->solves, 0) + 1);
    if (P.icp_min_valid > 0 && cnt < (double)P.icp_min_valid) {
        stg(&st->icp_failed, 0, 7);
        return;
    }
    const int cs = cholesky_solve(ATA, ATb, (double)P.icp_rel_pivot, x6);
    if (cs != 0) {
        stg(&st->icp_failed, 0, cs);
        return;
    }
    double ninf = 0.0;
    for (int i = 0; i < 6; ++i) ninf = fmax(ninf, fabs(x6[i]));
    if (ninf < 5.0e-7) {
        stg(&st->conv_level, 0, level);
        return;
    }
    double E[16], T21[16];
    expmap(x6, E);
    matmul4d(E, T12, T12);
    rigid_inverse(T12, T21);
    for (int i = 0; i < 16; ++i) {
        stg(st->T_12, i, T12[i]);
        stg(st->T_21, i, T21[i]);
        stg(st->T_21_f, i, (float)T21[i]);
    }
}

// Per-frame / per-track reset of the tracker state (one thread per session).
// `active` (nullable: all active) marks the sessions of a slot table that have
// a frame in this step; the others keep their state untouched by the step.
__global__ void k_track_setup(const SessParams *__restrict__ SP, int nsess, Mat16d Tinit, int has_init,
                              int force_track, const int *__restrict__ active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsess) return;
    SessState *st = SP[s].st;
    st->active = active ? active[s] : 1;
    if (!st->active) {
        st->do_track = 0;
        st->fuse_ok = 0;
        return;
    }
    st->do_track = force_track ? 1 : (st->frames > 0);
    st->icp_failed = 0;
    st->conv_level = -1;
    st->solves = 0;
    st->force_single = has_init;
    for (int k = 0; k < 16; ++k) st->T_12[k] = has_init ? Tinit.m[k] : ((k % 5) == 0 ? 1.0 : 0.0);
    publish_T21(st);
}

__device__ void set_pose_derived(const SessParams &P, SessState *st) {
    rigid_inverse(st->T_gc, st->T_cg);
    for (int k = 0; k < 16; ++k) {
        st->T_gc_f[k] = (float)st->T_gc[k];
        st->T_cg_f[k] = (float)st->T_cg[k];
    }
    // raycast constants from the FP32 pose (d_gcc_Get_Camera_Center and
    // rmi_Render_Maps_for_ICP, tsdf_renderer.cu:1667-1675)
    const float *T = st->T_gc_f;
    const float tx = T[3], ty = T[7], tz = T[11];
    const f3 cam = {-T[0] * tx - T[4] * ty - T[8] * tz, -T[1] * tx - T[5] * ty - T[9] * tz,
                    -T[2] * tx - T[6] * ty - T[10] * tz};
    const f3 vcc = sub({P.cube_cen[0], P.cube_cen[1], P.cube_cen[2]}, cam);
    const f3 ncc = normalised(vcc);
    const float d = sqrtf(dot(vcc, vcc));
    const float s = P.r_cube / d;
    st->cam[0] = cam.x; st->cam[1] = cam.y; st->cam[2] = cam.z;
    st->ncc[0] = ncc.x; st->ncc[1] = ncc.y; st->ncc[2] = ncc.z;
    st->dist_cc = d;
    st->cos_max = (s < 1.0f) ? sqrtf(1.0f - s * s) : -2.0f;
}

__device__ double rotation_deg(const double *T) {
    double c = (T[0] + T[5] + T[10] - 1.0) * 0.5;
    c = fmax(-1.0, fmin(1.0, c));
    return acos(c) * 57.295779513082320876798;
}

// Apply the tracked increment (T_gc <- T_12 T_gc, pose_tracker.cu:992-998) or
// record the failure; decides whether the frame is integrated.
__global__ void k_track_finalize(const SessParams *__restrict__ SP, int nsess) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsess) return;
    const SessParams &P = SP[s];
    SessState *st = P.st;
    if (!st->active) {
        st->fuse_ok = 0;
        return;
    }
    int status = 0;
    if (st->do_track) {
        if (st->icp_failed) {
            status = st->icp_failed;
        } else {
            const double *T = st->T_12;
            const double tn = sqrt(T[3] * T[3] + T[7] * T[7] + T[11] * T[11]);
            if ((P.max_motion_m > 0.0f && tn > P.max_motion_m) ||
                (P.max_motion_deg > 0.0f && rotation_deg(T) > P.max_motion_deg)) {
                status = 5;
            } else {
                matmul4d(st->T_12, st->T_gc, st->T_gc);
                set_pose_derived(P, st);
            }
        }
    }
    st->last_status = status;
    st->fuse_ok = (status == 0);
    if (status == 0) st->frames += 1;
}

// Pose set from the host: derived copies only.
__global__ void k_pose_derived(const SessParams *__restrict__ SP) {
    if (threadIdx.x == 0 && blockIdx.x == 0) set_pose_derived(SP[0], SP[0].st);
}

// Normal equations of one iteration without a solve (diagnostics/parity).
__global__ __launch_bounds__(kTileThreads) void k_icp_system(const SessParams *__restrict__ SP,
                                                             const WorkItem *__restrict__ work, int level, Mat16f T21,
                                                             double *out_part, uint8_t *mask) {
    __shared__ double red64[kIcpVals];
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    const LevelDesc &L = P.lev[level];
    const int x = (wi.tile % L.tiles_x) * kTileW + threadIdx.x % kTileW;
    const int y = (wi.tile / L.tiles_x) * kTileH + threadIdx.x / kTileW;
    float A[6] = {0, 0, 0, 0, 0, 0}, b = 0.0f;
    bool ok = false;
    if (x < L.W && y < L.H) {
        ok = icp_summand(P, level, x, y, T21.m, A, b);
        if (mask) mask[(size_t)y * L.W + x] = ok ? 1 : 0;
    }
    float v[kIcpVals];
#pragma unroll
    for (int k = 0; k < kIcpVals; ++k) v[k] = 0.0f;
    add_terms(ok, A, b, v);
    block_terms(v, red64);
    if (threadIdx.x < kIcpVals) out_part[(size_t)wi.tile * kIcpVals + threadIdx.x] = red64[threadIdx.x];
}

// Standalone integrate/raycast calls always run: mark the frame fusable.
__global__ void k_fuse_ok(const SessParams *__restrict__ SP, int nsess) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsess) {
        SP[s].st->fuse_ok = 1;
        SP[s].st->active = 1;
    }
}
void launch_fuse_ok(const SessParams *P, int nsess, hipStream_t s) {
    hipLaunchKernelGGL(k_fuse_ok, dim3((nsess + 63) / 64), dim3(64), 0, s, P, nsess);
}

void launch_track_setup(const SessParams *P, int nsess, const double *T_init, hipStream_t s) {
    Mat16d m{};
    int has = T_init != nullptr;
    if (has)
        for (int k = 0; k < 16; ++k) m.m[k] = T_init[k];
    // force_track: an explicit T_init call (vs_track) always tracks
    hipLaunchKernelGGL(k_track_setup, dim3((nsess + 63) / 64), dim3(64), 0, s, P, nsess, m, has, 0,
                       (const int *)nullptr);
}
void launch_track_setup_forced(const SessParams *P, int nsess, const double *T_init, hipStream_t s) {
    Mat16d m{};
    int has = T_init != nullptr;
    if (has)
        for (int k = 0; k < 16; ++k) m.m[k] = T_init[k];
    hipLaunchKernelGGL(k_track_setup, dim3((nsess + 63) / 64), dim3(64), 0, s, P, nsess, m, has, 1,
                       (const int *)nullptr);
}
void launch_track_setup_active(const SessParams *P, int nsess, const int *active, hipStream_t s) {
    Mat16d m{};
    hipLaunchKernelGGL(k_track_setup, dim3((nsess + 63) / 64), dim3(64), 0, s, P, nsess, m, 0, 0, active);
}

// Writes a slot table's active flags from a bit mask passed by value, so a
// captured step reads them from device memory.
struct ActiveMask {
    uint32_t bits[kMaxSlots / 32];
};
__global__ void k_set_active(int *active, int nslots, ActiveMask m) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) active[s] = (m.bits[s >> 5] >> (s & 31)) & 1u;
}
void launch_set_active(int *active, int nslots, const uint32_t bits[kMaxSlots / 32], hipStream_t s) {
    ActiveMask m;
    for (int k = 0; k < kMaxSlots / 32; ++k) m.bits[k] = bits[k];
    hipLaunchKernelGGL(k_set_active, dim3((nslots + 255) / 256), dim3(256), 0, s, active, nslots, m);
}
void launch_icp_iter(const SessParams *P, const WorkItem *work, int n, int nsess, int level, int iter,
                     hipStream_t s) {
    if (n <= 0) return;
    static_assert(kMaxLevels == 3, "one launch case per level");
    // one instantiation per distinct group size in use
    if (level == 0)
        hipLaunchKernelGGL(k_icp_iter<icp_group(0)>, dim3(n), dim3(kTileThreads), 0, s, P, work, level, iter);
    else if (level == 1)
        hipLaunchKernelGGL(k_icp_iter<icp_group(1)>, dim3(n), dim3(kTileThreads), 0, s, P, work, level, iter);
    else
        hipLaunchKernelGGL(k_icp_iter<icp_group(2)>, dim3(n), dim3(kTileThreads), 0, s, P, work, level, iter);
    hipLaunchKernelGGL(k_icp_solve, dim3(nsess), dim3(kSolveThreads), 0, s, P, nsess, level, iter);
}
void launch_track_finalize(const SessParams *P, int nsess, hipStream_t s) {
    hipLaunchKernelGGL(k_track_finalize, dim3((nsess + 63) / 64), dim3(64), 0, s, P, nsess);
}
void launch_pose_derived(const SessParams *P, hipStream_t s) {
    hipLaunchKernelGGL(k_pose_derived, dim3(1), dim3(64), 0, s, P);
}
// Sum of one ICP system's tile partials in tile order: thread k adds term k
// of tiles 0, 1, ..., nt-1 in FP64, the order of vs_icp_system's host loop,
// so both give the same bits.
__global__ void k_icp_sum_tiles(const double *__restrict__ part, int nt, double *__restrict__ out) {
    const int k = threadIdx.x;
    if (k >= kIcpVals) return;
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += ldg(part, (size_t)t * kIcpVals + k);
    stg(out, k, s);
}
void launch_icp_sum_tiles(const double *part, int nt, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_icp_sum_tiles, dim3(1), dim3(64), 0, s, part, nt, out);
}
void launch_icp_system(const SessParams *P, const WorkItem *work, int n, int level, const float *T21,
                       double *out_part, uint8_t *mask, hipStream_t s) {
    Mat16f m;
    for (int k = 0; k < 16; ++k) m.m[k] = T21[k];
    if (n > 0)
        hipLaunchKernelGGL(k_icp_system, dim3(n), dim3(kTileThreads), 0, s, P, work, level, m, out_part, mask);
}

}  // namespace vs
