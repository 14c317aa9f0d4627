// vs_vol.hip — volume stages: brick visibility (A13), TSDF integration (A14),
// raycast of the ICP model maps (A16-A18), and volume maintenance.
//
// Volume layout in HBM: brick-major. The cube is cut into 8^3 bricks; brick
// (bx,by,bz) has index (bz*BH + by)*BW + bx and its 512 voxels are stored
// contiguously, z-major then y then x inside the brick, each as one record of
// its value and weight (VoxRec: 4 bytes for FP16 + u8, 8 otherwise; a lookup
// is one load). A brick is 2 or 4 KB, so one workgroup streams a whole brick
// with fully coalesced accesses, and the 2x2x2 trilinear and 32-tap gradient
// stencils of the raycast mostly stay inside one or two bricks.
//
// Per-session bitmaps: `bvis` (bricks touched by the current frame, rebuilt
// each frame from the per-brick flags `vmask`), `bocc` (bricks holding any
// voxel of weight > 0; weights never
// decrease, so the bit is set once and kept) and the band bits `bband` /
// `sband` (bricks / superbricks that may hold a voxel of weight > 0 and a value
// other than 1.0; set when an update writes such a value, and kept).
#include "vs_device.hpp"
#include "vs_launch.hpp"

namespace vs {

// ---------------------------------------------------------------------------
// A13: visible bricks (g_fvsc_Find_Visible_Sub_Cubes, volume_integrator.cu:84-152)
// ---------------------------------------------------------------------------
// Clears the list counter and, for every session that fuses this frame, its
// visible-brick flags.
__global__ void k_vis_clear(const SessParams *__restrict__ SP, int nsess, BrickList bl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *bl.count = 0u;
    const int s = blockIdx.y;
    if (s >= nsess) return;
    const SessParams &P = SP[s];
    if (!P.st->active) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.st->vis_count = 0u;
        P.st->upd_count = 0ull;
    }
    if (!P.st->fuse_ok) return;
    const uint32_t nb = (uint32_t)P.bdim[0] * P.bdim[1] * P.bdim[2];
    const uint32_t n16 = (nb + 127) / 128 * 8;  // 16-byte words of the padded flags
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n16; w += gridDim.x * blockDim.x)
        stg((vs_u4 *)P.vmask, w, vs_u4{0u, 0u, 0u, 0u});
}

// Flags the bricks hit by the samples p + n*mu*dir, n = -4..1, of every valid
// pixel's ray: one byte per brick, set to 1 by plain stores (the reference's
// bool mask), so the bricks that many pixels hit need neither atomics nor a
// deduplication pass; a lane whose brick equals its left neighbour's for the
// same sample leaves the store to that lane (runs of ~10 neighbouring pixels
// share a brick). The brick list and the visible-brick bits are built from
// the flags by k_vis_list. (A device-scope atomic per new brick on a bitmap,
// and a returning atomic on one list counter per image tile, had cost this
// launch most of its time.)
__global__ __launch_bounds__(kTileThreads) void k_vis_mark(const SessParams *__restrict__ SP,
                                                           const WorkItem *__restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    SessState *st = P.st;
    const LevelDesc &L = P.lev[0];
    if (!st->fuse_ok || tile_out(L, wi.tile)) return;
    const int x = (wi.tile % L.tiles_x) * kTileW + threadIdx.x % kTileW;
    const int y = (wi.tile / L.tiles_x) * kTileH + threadIdx.x / kTileW;
    const float td = (x < L.W && y < L.H) ? raw_depth(P, (size_t)y * L.W + x) : 0.0f;
    uint32_t bb[6];
    for (int k = 0; k < 6; ++k) bb[k] = 0xFFFFFFFFu;
    if (td > 0.0f) {
        const M34 Tc = ld_m34(st->T_cg_f);
        const float *Tcg = Tc.m;
        const f3 p = transform(Tcg, back_project((float)x, (float)y, L.K, td));
        const f3 pe = transform(Tcg, back_project((float)x, (float)y, L.K, 100.0f));
        const f3 dir = normalised(sub(pe, p));
        for (int n = -4; n <= 1; ++n) {
            const f3 q = add(p, scale((float)n * P.mu, dir));
            f3 tv;
            if (!world_to_voxel(P, q, tv)) continue;
            // ROUNDF(v)/8 with the macro's precedence: the negative branch is not divided
            // (the positive branch's integer is >= 0, so / 8 is a shift)
            const int bx = (tv.x < 0.0f) ? (int)(tv.x - 0.5f) : (int)((unsigned)(int)(tv.x + 0.5f) >> 3);
            const int by = (tv.y < 0.0f) ? (int)(tv.y - 0.5f) : (int)((unsigned)(int)(tv.y + 0.5f) >> 3);
            const int bz = (tv.z < 0.0f) ? (int)(tv.z - 0.5f) : (int)((unsigned)(int)(tv.z + 0.5f) >> 3);
            if (bx < 0 || bx >= P.bdim[0] || by < 0 || by >= P.bdim[1] || bz < 0 || bz >= P.bdim[2]) break;
            bb[n + 4] = ((uint32_t)bz * P.bdim[1] + by) * P.bdim[0] + bx;  // 0xFFFFFFFF: no brick
        }
    }
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t b = bb[k];
        const uint32_t left = __shfl_up(b, 1, 64);
        if (b != 0xFFFFFFFFu && (lane == 0 || b != left)) stg(P.vmask, b, (uint8_t)1);
    }
}

// Turns the sessions' visible-brick flags into the visible-brick bits and
// the batch's brick list. Workgroup (c, s) takes bricks [c, c + 1) *
// kVisListBricks of session s: each thread packs the flags of 128
// consecutive bricks (eight 16-byte loads) into 4 bit words and stores them,
// then the workgroup ranks the set bits (popcounts and a prefix sum),
// reserves their range with one atomic on the list counter (8 workgroups for
// a 512^3 session) and writes (session, brick) in brick order, each thread
// its own bricks.
constexpr int kVisListWordsPerThread = 4;
constexpr uint32_t kVisListBricks = 32u * kVisListWordsPerThread * kTileThreads;

__global__ __launch_bounds__(kTileThreads) void k_vis_list(const SessParams *__restrict__ SP, BrickList bl) {
    __shared__ unsigned wsum[kTileThreads / 64];
    __shared__ unsigned base;
    const int sess = blockIdx.y;
    const SessParams &P = SP[sess];
    SessState *st = P.st;
    if (!ldg(&st->fuse_ok, 0)) return;
    const uint32_t nb = (uint32_t)P.bdim[0] * P.bdim[1] * P.bdim[2], nw = (nb + 31) / 32;
    const uint32_t b0 = blockIdx.x * kVisListBricks + threadIdx.x * (32u * kVisListWordsPerThread);
    if (blockIdx.x * kVisListBricks >= nb) return;
    uint32_t w[kVisListWordsPerThread] = {0u, 0u, 0u, 0u};
    if (b0 < nb) {  // flags padded to 128 bricks: the 128 read here are in the buffer
        const vs_u4 *f = (const vs_u4 *)(P.vmask + b0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const vs_u4 v = ldg(f, q);
            const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t x = e[r];  // flags of bricks b0 + 16q + 4r .. +3, one byte each (0 or 1)
                const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
                w[q >> 1] |= nib << (16 * (q & 1) + 4 * r);
            }
        }
    }
    const uint32_t w0 = b0 / 32u;
    unsigned c = 0;
#pragma unroll
    for (int k = 0; k < kVisListWordsPerThread; ++k) {
        if (w0 + k < nw) stg(P.bvis, w0 + k, w[k]);
        c += (unsigned)__popc(w[k]);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned v = __shfl_up(inc, off, 64);
        if (lane >= off) inc += v;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    unsigned before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kTileThreads / 64; ++k) {
        before += (k < wv) ? wsum[k] : 0u;
        total += wsum[k];
    }
    if (total == 0) return;
    if (threadIdx.x == 0) {
        base = atomicAdd(bl.count, total);
        atomicAdd(&st->vis_count, total);
    }
    __syncthreads();
    unsigned pos = base + before + inc - c;
#pragma unroll
    for (int k = 0; k < kVisListWordsPerThread; ++k) {
        uint32_t bits = w[k];
        while (bits) {
            const int j = __ffs(bits) - 1;
            bits &= bits - 1u;
            if (pos < bl.capacity) bl.items[pos] = make_int2(sess, (int)((w0 + k) * 32u + (uint32_t)j));
            ++pos;
        }
    }
}

// ---------------------------------------------------------------------------
// A14: TSDF update of the listed bricks (g_ctv_Compute_TSDF_of_Voxels,
// volume_integrator.cu:154-262). Persistent: each workgroup walks the brick
// list with a grid stride; 256 threads x 2 voxels = one brick per step.
// Value and weight are loaded only for voxels that pass the projective test.
// ---------------------------------------------------------------------------
// d_gii_Get_Interpolated_Intensity (_yooji_cuda_math.h:305-332): bilinear
// over the 2x2 neighbourhood of an 8-bit plane; a participating tap <= 0 or
// a position outside the plane fails.
__device__ __forceinline__ bool gii(const uint8_t *plane, int W, int H, float xf, float yf, float &out) {
    if (xf < 0.0f || xf > float(W - 1) || yf < 0.0f || yf > float(H - 1)) return false;
    float d2 = 0.0f, d3 = 0.0f, d4 = 0.0f;
    const int x = (int)xf, y = (int)yf;
    const float rx = xf - (float)x, ry = yf - (float)y;
    const float d1 = (float)ldg(plane, (size_t)y * W + x);
    if (d1 <= 0.0f) return false;
    if (rx > 0.0f) {
        d2 = (float)ldg(plane, (size_t)y * W + x + 1);
        if (d2 <= 0.0f) return false;
    }
    if (ry > 0.0f) {
        d3 = (float)ldg(plane, (size_t)(y + 1) * W + x);
        if (d3 <= 0.0f) return false;
    }
    if (rx > 0.0f && ry > 0.0f) {
        d4 = (float)ldg(plane, (size_t)(y + 1) * W + x + 1);
        if (d4 <= 0.0f) return false;
    }
    out = (1.0f - ry) * ((1.0f - rx) * d1 + rx * d2) + ry * ((1.0f - rx) * d3 + rx * d4);
    return true;
}

// One voxel of the production operator g_ctvc_Compute_TSDF_of_Voxels_with_Color
// (z_yooji_cuda_volume_integrator.cu:265-440, registered colour frame): pixel
// by __float2int_rn, SDF along the pixel's ray, voxels at the weight cap
// frozen, average by 1/newW, and each colour channel a running average of its
// bilinear sample (a channel whose tap fails keeps its value). Returns true
// when the voxel was updated.
template <class TV, class TW>
__device__ __forceinline__ bool integrate_voxel_color(const SessParams &P, const float *T, const float *Tcg, f3 g,
                                                      void *vb, uint32_t *col, int li, bool &band) {
    const LevelDesc &L = P.lev[0];
    const f3 p = transform(T, g);
    float u, v;
    project(p, L.K, u, v);
    int x, y;
    if (!rint_ref(u, x) || !rint_ref(v, y)) return false;
    if (x < 0 || x > L.W - 1 || y < 0 || y > L.H - 1) return false;
    const float dm = raw_depth(P, (size_t)y * L.W + x);
    if (!(dm > 0.0f)) return false;
    const float r1x = 1.0f * ((float)x - L.K[2]) / L.K[0], r1y = 1.0f * ((float)y - L.K[3]) / L.K[1];
    const float lambda_inv = 1.0f / sqrtf(r1x * r1x + r1y * r1y + 1.0f);
    const float dx = g.x - Tcg[3], dy = g.y - Tcg[7], dz = g.z - Tcg[11];
    const float norm = sqrtf(dx * dx + dy * dy + dz * dz);
    const float sdf = dm - lambda_inv * norm;
    if (sdf < -P.mu) return false;
    typedef VoxRec<TV, TW> R;
    const typename R::T rec = ld_rec<TV, TW>(vb, li);
    const float w = R::w(rec);
    if (w >= P.max_w) return false;  // frozen voxel
    const float nw = w + 1.0f, inw = 1.0f / nw;
    const float tsd = fminf(1.0f, sdf / P.mu);
    const float nv = (w * R::val(rec) + tsd) * inw;
    band |= nv != 1.0f;
    st_rec<TV, TW>(vb, li, R::make(nv, nw));
    uint32_t c = ldg(col, li);
    const size_t plane = (size_t)L.W * L.H;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float smp;
        if (!gii(P.rgb + k * plane, L.W, L.H, u, v, smp)) continue;
        const float val = (w * (float)((c >> (8 * k)) & 255u) + smp) * inw;
        const uint32_t q = (uint32_t)(uint8_t)fmaxf(0.0f, fminf(255.0f, val));
        c = (c & ~(255u << (8 * k))) | (q << (8 * k));
    }
    stg(col, li, c);
    return true;
}

// Four consecutive voxel records of a brick (x, x+1, x+2, x+3 of one row, x a
// multiple of 4): one 16-byte access (4-byte records) or two (8-byte).
template <class TV, class TW>
__device__ __forceinline__ void ld_rec4(const void *base, size_t i, typename VoxRec<TV, TW>::T r[4]) {
    typedef typename VoxRec<TV, TW>::T T;
    const VS_GLOBAL vs_u4 *p = (const VS_GLOBAL vs_u4 *)((const T *)base + i);
    if constexpr (sizeof(T) == 4) {
        const vs_u4 v = p[0];
        r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
    } else {
        const vs_u4 a = p[0], b = p[1];
        r[0] = T{a.x, a.y}; r[1] = T{a.z, a.w}; r[2] = T{b.x, b.y}; r[3] = T{b.z, b.w};
    }
}
template <class TV, class TW>
__device__ __forceinline__ void st_rec4(void *base, size_t i, const typename VoxRec<TV, TW>::T r[4]) {
    typedef typename VoxRec<TV, TW>::T T;
    VS_GLOBAL vs_u4 *p = (VS_GLOBAL vs_u4 *)((T *)base + i);
    if constexpr (sizeof(T) == 4) {
        p[0] = vs_u4{r[0], r[1], r[2], r[3]};
    } else {
        p[0] = vs_u4{r[0].x, r[0].y, r[1].x, r[1].y};
        p[1] = vs_u4{r[2].x, r[2].y, r[3].x, r[3].y};
    }
}

// The integrate walks the brick list in chunks of kIntegChunk consecutive
// entries. The list comes from k_vis_mark in (session, image tile) order, so
// a chunk's bricks mostly belong to one session and lie behind one small
// patch of its depth frame. Workgroup g runs on XCD g % 8 (workgroups are
// dealt to the eight XCDs in turn), and the workgroups of one XCD take the
// chunks of one contiguous eighth of the list: the depth rows that
// neighbouring bricks read are then fetched into that XCD's L2 once, not
// into eight of them.
constexpr int kIntegChunk = 8;
constexpr int kXcds = 8;

// One workgroup step takes two bricks, 128 threads each; a thread updates
// four voxels of one brick row (x0..x0+3, y, z), so a wave covers 256
// consecutive voxels (z-slices of 64). Voxel arithmetic per voxel as
// g_ctv_Compute_TSDF_of_Voxels (volume_integrator.cu:154-262). Voxel records
// are read and written in whole 128-voxel groups (32 threads, 4 records
// each): a group with any updated voxel loads all of its records and stores
// them all back, unchanged ones as they were, so its memory lines are written whole
// (a partly written line costs the memory a read-modify-write); a group
// without updates touches neither. Per-brick updated counts and band flags
// go to LDS; after each chunk one barrier, then thread k < 8 sets brick k's
// occupancy and band bits and thread 0 adds the chunk's counts.
template <class TV, class TW>
__global__ __launch_bounds__(kTileThreads) void k_integrate(const SessParams *__restrict__ SP, BrickList bl) {
    const unsigned n = min(*bl.count, bl.capacity);
    const unsigned nch = (n + kIntegChunk - 1) / kIntegChunk;
    const unsigned G = gridDim.x;
    const unsigned nx = (G % kXcds == 0) ? kXcds : 1u;
    const unsigned xcd = blockIdx.x % nx, R = G / nx, r = blockIdx.x / nx;
    const unsigned c_lo = (unsigned)((unsigned long long)nch * xcd / nx);
    const unsigned c_hi = (unsigned)((unsigned long long)nch * (xcd + 1) / nx);
    const int t = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(t >> 7);  // which brick of the pair (wave-uniform)
    const int lt = t & 127, lane = t & 63, wib = (t >> 6) & 1;
    const int xl = (lt & 1) * 4, yl = (lt >> 1) & 7, zl = lt >> 4;
    const int li = zl * 64 + yl * 8 + xl;
    // [chunk parity][brick of the chunk][wave of the brick]
    __shared__ unsigned s_upd[2][kIntegChunk][2], s_band[2][kIntegChunk][2];
    __shared__ int2 s_item[2][kIntegChunk];
    int parity = 0;
    int cur = -1;  // session whose pose this wave holds
    M34 Tg = {};
    // updated-voxel count of the current run of same-session bricks (thread 0)
    int run_sess = -1;
    unsigned long long run_upd = 0ull;
    for (unsigned j = c_lo + r; j < c_hi; j += R) {
        const unsigned b0 = j * kIntegChunk;
        const int nb = (int)min((unsigned)kIntegChunk, n - b0);
        for (int k = half; k < nb; k += 2) {
            const int2 it = bl.items[b0 + k];
            const SessParams &P = SP[it.x];
            if (it.x != cur) {
                cur = it.x;
                Tg = ld_m34(P.st->T_gc_f);
            }
            const float *T = Tg.m;
            const int b = it.y;
            const int bx = b % P.bdim[0], by = (b / P.bdim[0]) % P.bdim[1], bz = b / (P.bdim[0] * P.bdim[1]);
            void *vb = (typename VoxRec<TV, TW>::T *)P.vox + (size_t)b * kBrickVox;
            const LevelDesc &L = P.lev[0];
            // voxel-centre world coordinates from the session's tables
            const float gy = ldg(P.wtab[1], by * kBrick + yl), gz = ldg(P.wtab[2], bz * kBrick + zl);
            float gxs[4];
            {
                const vs_f4 g4 = *(const VS_GLOBAL vs_f4 *)(P.wtab[0] + bx * kBrick + xl);
                gxs[0] = g4.x;
                gxs[1] = g4.y;
                gxs[2] = g4.z;
                gxs[3] = g4.w;
            }
            int nupd = 0;
            bool band = false;  // an updated voxel's new value is not exactly 1.0
            if (P.integ_mode == 1) {
                const M34 Tc = ld_m34(P.st->T_cg_f);
                uint32_t *col = P.vcol + (size_t)b * kBrickVox;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    nupd += integrate_voxel_color<TV, TW>(P, T, Tc.m, {gxs[q], gy, gz}, vb, col, li + q, band);
            } else {
                float tsd[4];
                bool up[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const f3 p = transform(T, {gxs[q], gy, gz});
                    float u, v;
                    project(p, L.K, u, v);
                    // the pixel tests combined with & (no nest of branches); the
                    // depth is read at pixel 0 for a voxel outside the image
                    const bool fin = (int)(fabsf(u) < 2147483520.0f) & (int)(fabsf(v) < 2147483520.0f);
                    const float uc = fin ? u : 0.0f, vc = fin ? v : 0.0f;
                    const int iu = uc < 0.0f ? (int)(uc - 0.5f) : (int)(uc + 0.5f);  // ROUNDF
                    const int iv = vc < 0.0f ? (int)(vc - 0.5f) : (int)(vc + 0.5f);
                    const bool ok = (int)fin & (int)((unsigned)iu < (unsigned)L.W) & (int)((unsigned)iv < (unsigned)L.H);
                    const float d0 = raw_depth(P, ok ? (size_t)iv * L.W + iu : 0);
                    const float dm = ok ? d0 : 0.0f;
                    const float td = dm - p.z;
                    up[q] = (int)ok & (int)(dm > 0.0f) & (int)!(td < -P.mu);
                    tsd[q] = fminf(1.0f, __fdividef(td, P.mu));
                    nupd += up[q];
                }
                const unsigned long long any = __ballot(nupd > 0);
                if ((any >> (lane & 32)) & 0xFFFFFFFFull) {
                    typedef VoxRec<TV, TW> R;
                    typename R::T r[4];
                    ld_rec4<TV, TW>(vb, li, r);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (!up[q]) continue;
                        const float w = R::w(r[q]), nw = w + 1.0f;
                        const float nv = __fdividef(fmaf(w, R::val(r[q]), tsd[q]), nw);
                        band |= nv != 1.0f;
                        r[q] = R::make(nv, fminf(nw, P.max_w));
                    }
                    st_rec4<TV, TW>(vb, li, r);
                }
            }
            const unsigned c1 = (unsigned)__popcll(__ballot(nupd >= 1)), c2 = (unsigned)__popcll(__ballot(nupd >= 2));
            const unsigned c3 = (unsigned)__popcll(__ballot(nupd >= 3)), c4 = (unsigned)__popcll(__ballot(nupd >= 4));
            const bool wband = __ballot(band) != 0ull;
            if (lane == 0) {
                s_upd[parity][k][wib] = c1 + c2 + c3 + c4;
                s_band[parity][k][wib] = wband;
                if (wib == 0) s_item[parity][k] = it;
            }
        }
        __syncthreads();
        if (t < nb) {
            const unsigned upd = s_upd[parity][t][0] + s_upd[parity][t][1];
            if (upd > 0) {
                // occupancy and band bits are only ever set during this
                // launch, and earlier launches' writes (clears included) are
                // visible at its start: a plain (cached) read that shows a
                // bit proves it set, and a stale 0 costs one redundant atomic
                const int2 it = s_item[parity][t];
                const SessParams &P = SP[it.x];
                const int b = it.y;
                const bool cb = s_band[parity][t][0] | s_band[parity][t][1];
                const int bx = b % P.bdim[0], by = (b / P.bdim[0]) % P.bdim[1], bz = b / (P.bdim[0] * P.bdim[1]);
                const uint32_t bb = 1u << (b & 31);
                const int sb = super_of_brick(P, bx, by, bz);
                const uint32_t sbb = 1u << (sb & 31);
                const uint32_t wo = ldg(P.bocc, b >> 5), wbb = ldg(P.bband, b >> 5), ws = ldg(P.sband, sb >> 5);
                if (!(wo & bb)) atomicOr(P.bocc + (b >> 5), bb);
                if (cb && !(wbb & bb)) atomicOr(P.bband + (b >> 5), bb);
                if (cb && !(ws & sbb) && !(atomicOr(P.sband + (sb >> 5), sbb) & sbb))
                    stg(&P.st->sdist_dirty, 0, 1);  // a new band superbrick: rebuild the distance field
            }
        }
        if (t == 0) {
            for (int k = 0; k < nb; ++k) {
                const unsigned upd = s_upd[parity][k][0] + s_upd[parity][k][1];
                const int sess = s_item[parity][k].x;
                if (sess != run_sess) {
                    if (run_upd) atomicAdd(&SP[run_sess].st->upd_count, run_upd);
                    run_sess = sess;
                    run_upd = 0ull;
                }
                run_upd += (unsigned long long)upd;
            }
        }
        parity ^= 1;
    }
    if (t == 0 && run_upd) atomicAdd(&SP[run_sess].st->upd_count, run_upd);
}

// ---------------------------------------------------------------------------
// Superbrick distance field of the raycast's skip boxes: for every superbrick
// the L-infinity distance, in superbricks, to the nearest band superbrick,
// capped at kDistMax (0 = a band superbrick; cells outside the grid count as
// free), so the (2d-1)^3 superbricks centred on one at distance d hold no
// band brick. Separable: the distance along x within each row, then
// d'(p) = min over |t| < kDistMax of max(|t|, d(p + t*e)) along y and then z.
// Rebuilt only for sessions that gained a band superbrick (or when forced),
// over many workgroups: pass 1, one per (session, z-slice of superbricks),
// does x and y in LDS and leaves one byte per superbrick in the session's
// scratch; pass 2, one per (session, 2048 superbricks), does z from the
// scratch and packs two cells per byte of the field (byte-aligned chunks, so
// no two workgroups write one byte); pass 3 clears the rebuild flags, which
// every workgroup of passes 1 and 2 has read by then.
// ---------------------------------------------------------------------------
constexpr int kSdistThreads = 1024, kSdistChunk = 2 * kSdistThreads;

__device__ __forceinline__ bool sdist_todo(const SessParams &P, int force) {
    const int n = P.sdim[0] * P.sdim[1] * P.sdim[2];
    return n > 0 && n <= kDistCells && (force || ldg(&P.st->sdist_dirty, 0));
}

// min of rd(c) and, over 0 < |t| < kDistMax with pos + t in [0, len), of
// max(|t|, rd(c + t * step)); rd(i) reads cell i of the line's array
template <class R>
__device__ __forceinline__ int dist_line(R rd, int c, int pos, int len, int step) {
    int d = rd(c);
    for (int t = 1; t < kDistMax && t < d; ++t) {
        if (pos - t >= 0) d = min(d, max(t, rd(c - t * step)));
        if (pos + t < len) d = min(d, max(t, rd(c + t * step)));
    }
    return d;
}

__global__ __launch_bounds__(kSdistThreads) void k_sdist_xy(const SessParams *__restrict__ SP, int force) {
    __shared__ uint8_t A[kDistCells], B[kDistCells];
    const SessParams &P = SP[blockIdx.x];
    const int z = blockIdx.y;
    if (z >= P.sdim[2] || !sdist_todo(P, force)) return;
    const int nx = P.sdim[0], ny = P.sdim[1], m = nx * ny, base = z * m;  // m <= kDistCells
    for (int c = threadIdx.x; c < m; c += kSdistThreads) {
        const int g = base + c;
        A[c] = ((ldg(P.sband, g >> 5) >> (g & 31)) & 1u) ? 0 : kDistMax;
    }
    __syncthreads();
    const auto ra = [&](int i) { return (int)A[i]; };
    for (int c = threadIdx.x; c < m; c += kSdistThreads) B[c] = (uint8_t)dist_line(ra, c, c % nx, nx, 1);
    __syncthreads();
    const auto rb = [&](int i) { return (int)B[i]; };
    for (int c = threadIdx.x; c < m; c += kSdistThreads)
        stg(P.sdtmp, (size_t)(base + c), (uint8_t)dist_line(rb, c, c / nx, ny, nx));
}

__global__ __launch_bounds__(kSdistThreads) void k_sdist_z(const SessParams *__restrict__ SP, int force) {
    const SessParams &P = SP[blockIdx.x];
    if (!sdist_todo(P, force)) return;
    const int m = P.sdim[0] * P.sdim[1], nz = P.sdim[2], n = m * nz;
    const int c0 = blockIdx.y * kSdistChunk + 2 * threadIdx.x;
    if (c0 >= n) return;
    const uint8_t *tmp = P.sdtmp;
    const auto rt = [&](int i) { return (int)ldg(tmp, (size_t)i); };
    uint8_t out = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = c0 + h;
        if (c < n) out |= (uint8_t)(dist_line(rt, c, c / m, nz, m) << (4 * h));
    }
    stg(P.sdist, (size_t)(c0 >> 1), out);
}

__global__ void k_sdist_done(const SessParams *__restrict__ SP, int nsess) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsess) stg(&SP[s].st->sdist_dirty, 0, 0);
}

void launch_sdist(const SessParams *P, int nsess, int sd_slices, int force, hipStream_t s) {
    if (nsess <= 0 || sd_slices <= 0) return;
    hipLaunchKernelGGL(k_sdist_xy, dim3(nsess, sd_slices), dim3(kSdistThreads), 0, s, P, force);
    hipLaunchKernelGGL(k_sdist_z, dim3(nsess, kDistCells / kSdistChunk), dim3(kSdistThreads), 0, s, P, force);
    hipLaunchKernelGGL(k_sdist_done, dim3((nsess + 255) / 256), dim3(256), 0, s, P, nsess);
}

// ---------------------------------------------------------------------------
// A16-A18: raycast (d_crtc_Cast_Ray_on_TSDF_Cube_NEW + d_csnt + g_rmi,
// tsdf_renderer.cu:68-153, 288-635, 1069-1209). FP32 march with the
// reference's step rules. After its first tap a ray's march depends only on
// the values its taps read, not on their validity, and every voxel of a
// brick whose band bit is clear reads 1.0 (weight 0 reads as "no data" = 1.0,
// the others hold exactly 1.0): the march reads such bricks as 1.0 without
// touching the volume, and skips the steps that stay inside them.
// ---------------------------------------------------------------------------
// Range tests of the march's lookups, the three axes evaluated together
// (bitwise, no short-circuit). finite3: every coordinate inside the range
// where the reference's int conversions are defined (|v| < 2147483520);
// below3: 0 <= a < na on every axis, unsigned compares.
__device__ __forceinline__ bool finite3(f3 v) {
    constexpr float kLim = 2147483520.0f;
    return (int)(fabsf(v.x) < kLim) & (int)(fabsf(v.y) < kLim) & (int)(fabsf(v.z) < kLim);
}
__device__ __forceinline__ bool below3(int x, int y, int z, int nx, int ny, int nz) {
    return (int)((unsigned)x < (unsigned)nx) & (int)((unsigned)y < (unsigned)ny) & (int)((unsigned)z < (unsigned)nz);
}

template <class TV, class TW>
struct Vol {
    const SessParams &P;
    const void *vx;  // voxel records
    // superbrick distance field: the workgroup's LDS copy (4 bits per
    // superbrick). A session without one gets zeros (every superbrick counts
    // as a band one and only the brick bits decide); the index is masked to
    // the copy's size either way.
    const VS_LDS uint8_t *sdist;
    // of the last brick looked up (a ray's consecutive taps mostly stay in
    // one brick; the bitmaps are read-only during the launch), in one
    // register: brick index in bits 0-24 (< 2^24 bricks, so the initial
    // 2^25 - 1 matches none), its superbrick's distance in bits 25-28, bit 31
    // = the brick is a band one
    static constexpr unsigned kIdx = 0x01FFFFFFu, kBrickBand = 1u << 31;
    mutable unsigned cbf = kIdx;
    __device__ __forceinline__ int cb() const { return (int)(cbf & kIdx); }
    __device__ __forceinline__ int dist() const { return (int)((cbf >> 25) & 15u); }
    __device__ __forceinline__ bool band(int x, int y, int z) const {
        // brick_of and super_of_brick of an in-grid voxel with 24-bit
        // products (< 2^24 bricks)
        const unsigned b = __umul24(__umul24((unsigned)z >> 3, (unsigned)P.bdim[1]) + ((unsigned)y >> 3),
                                    (unsigned)P.bdim[0]) + ((unsigned)x >> 3);
        if (b != (cbf & kIdx)) {
            constexpr int ss = 3 + kSuperShift;
            const unsigned sb = __umul24(__umul24((unsigned)z >> ss, (unsigned)P.sdim[1]) + ((unsigned)y >> ss),
                                         (unsigned)P.sdim[0]) + ((unsigned)x >> ss);
            const unsigned d = (sdist[(sb >> 1) & (kDistCells / 2 - 1)] >> ((sb & 1u) * 4u)) & 15u;
            unsigned f = b | (d << 25);
            if (d == 0u && ((ldg(P.bband, b >> 5) >> (b & 31)) & 1u)) f |= kBrickBand;
            cbf = f;
        }
        return cbf & kBrickBand;
    }
    __device__ __forceinline__ bool in_grid(int x, int y, int z) const {
        return x >= 0 && x < P.vdim[0] && y >= 0 && y < P.vdim[1] && z >= 0 && z < P.vdim[2];
    }
    // Value of an in-grid voxel. Weight and value are loaded together (no
    // dependency between the two loads); no weight reads as "no data" (1.0).
    __device__ __forceinline__ bool load(size_t idx, float &out) const {
        typedef VoxRec<TV, TW> R;
        const typename R::T r = ld_rec<TV, TW>(vx, idx);
        const bool ok = R::has_w(r);
        out = ok ? R::val(r) : 1.0f;
        return ok;
    }
    // voxel (x, y, z) in range; O32: 32-bit offsets (records below 4 GB)
    template <bool O32>
    __device__ __forceinline__ bool load_at(int x, int y, int z, float &out) const {
        if (!O32) return load(vox_index(P, x, y, z), out);
        typedef VoxRec<TV, TW> R;
        const typename R::T r = ld_rec32<TV, TW>(vx, vox_index32(P, x, y, z));
        const bool ok = R::has_w(r);
        out = ok ? R::val(r) : 1.0f;
        return ok;
    }
    // voxel b * 512 + l (brick index b, in-brick offset l < 512); O32 as for
    // load_at
    template <bool O32>
    __device__ __forceinline__ bool load_bl(unsigned b, unsigned l, float &out) const {
        if (!O32) return load((size_t)b * kBrickVox + l, out);
        typedef VoxRec<TV, TW> R;
        const typename R::T r = ld_rec32<TV, TW>(vx, (b << 9) | l);
        const bool ok = R::has_w(r);
        out = ok ? R::val(r) : 1.0f;
        return ok;
    }
    // d_gtvu, exactly: nearest voxel by ROUNDF; true when it has weight.
    __device__ __forceinline__ bool nearest_valid(f3 v, float &out) const {
        int x, y, z;
        out = 1.0f;
        if (!roundf_ref(v.x, x) || !roundf_ref(v.y, y) || !roundf_ref(v.z, z)) return false;
        if (!in_grid(x, y, z)) return false;
        return load(vox_index(P, x, y, z), out);
    }
    // d_gtvu's value as the march uses it: the value of the nearest voxel by
    // ROUNDF, 1.0 in a brick without band bit. Returns 1 (a band brick's
    // value), 0 (1.0 read without a load) or -1 when the rounded position
    // lies outside the grid on an axis along which the ray moves away or not
    // at all: every later position of a ray that keeps stepping forward is
    // then outside as well. For 0 it also reports what the lookup learnt
    // about the neighbourhood: size > 0 when every voxel of the box
    // [lo, lo + size)^3 (a superbrick or brick without band bit) reads 1.0,
    // size = -1 when the position is outside the grid (moving towards it),
    // size = 0 otherwise.
    template <bool O32>
    __device__ __forceinline__ int nearest(f3 v, f3 rd, float &out, int3 &lo, int &size) const {
        out = 1.0f;
        size = 0;
        // the tests of the three axes are combined with & rather than &&:
        // one exec-mask branch instead of a nest of six, each with its own
        // mask bookkeeping (the march loop's instruction issue is its bound)
        if (!finite3(v)) return 0;
        const int x = v.x < 0.0f ? (int)(v.x - 0.5f) : (int)(v.x + 0.5f);  // ROUNDF
        const int y = v.y < 0.0f ? (int)(v.y - 0.5f) : (int)(v.y + 0.5f);
        const int z = v.z < 0.0f ? (int)(v.z - 0.5f) : (int)(v.z + 0.5f);
        if (!below3(x, y, z, P.vdim[0], P.vdim[1], P.vdim[2])) {
            if ((int)(x < 0 && !(rd.x > 0.0f)) | (int)(x >= P.vdim[0] && !(rd.x < 0.0f)) |
                (int)(y < 0 && !(rd.y > 0.0f)) | (int)(y >= P.vdim[1] && !(rd.y < 0.0f)) |
                (int)(z < 0 && !(rd.z > 0.0f)) | (int)(z >= P.vdim[2] && !(rd.z < 0.0f)))
                return -1;
            size = -1;
            return 0;
        }
        if (!band(x, y, z)) {
            // d = 0: the brick (a brick without band bit in a band
            // superbrick); d > 0: the (2d-1)^3 superbricks centred here,
            // which hold no band brick
            const int d = dist();
            const int sh = d == 0 ? 3 : 3 + kSuperShift, r = max(d - 1, 0), unit = 1 << sh;
            lo = make_int3(((x >> sh) - r) * unit, ((y >> sh) - r) * unit, ((z >> sh) - r) * unit);
            size = (2 * r + 1) * unit;
            return 0;
        }
        return load_bl<O32>((unsigned)cb(), (unsigned)(((z & 7) << 6) | ((y & 7) << 3) | (x & 7)), out) ? 1 : 0;
    }
    // ---- instrumented form only: the bytes the reference's own lookups
    // request, tap by tap, in its order and with its early exits.
    // d_gtvu (tsdf_renderer.cu:68-97) at integer voxel (x, y, z): nothing
    // outside the grid, the weight of a voxel without weight, weight and
    // value otherwise; true when the tap succeeds.
    __device__ __forceinline__ bool ref_tap(int x, int y, int z, unsigned long long &bytes) const {
        if (!in_grid(x, y, z)) return false;
        const bool w = VoxRec<TV, TW>::has_w(ld_rec<TV, TW>(vx, vox_index(P, x, y, z)));
        bytes += sizeof(TW);
        if (!w) return false;
        bytes += sizeof(TV);
        return true;
    }
    // a nearest tap (ROUNDF of the position); cls = 0 outside the grid, 1 no
    // weight, 2 valid
    __device__ __forceinline__ int ref_nearest(f3 v, unsigned long long &bytes) const {
        int x, y, z;
        if (!roundf_ref(v.x, x) || !roundf_ref(v.y, y) || !roundf_ref(v.z, z) || !in_grid(x, y, z)) return 0;
        return ref_tap(x, y, z, bytes) ? 2 : 1;
    }
    // d_gtvi (:100-153): 4 taps of the front plane, then 4 of the back plane,
    // up to the first failing one
    __device__ __forceinline__ void ref_trilinear(f3 v, unsigned long long &bytes) const {
        if (!(fabsf(v.x) < 2147483520.0f && fabsf(v.y) < 2147483520.0f && fabsf(v.z) < 2147483520.0f)) return;
        const int ox = (int)v.x, oy = (int)v.y, oz = (int)v.z;
        for (int k = 0; k < 8; ++k)
            if (!ref_tap(ox + (k & 1), oy + ((k >> 1) & 1), oz + (k >> 2), bytes)) return;
    }
    // d_csnt (:288-464): foremost 4, front 12, back 12, backmost 4 taps, up
    // to the first failing one
    __device__ __forceinline__ void ref_normal(f3 p, unsigned long long &bytes) const {
        f3 v;
        if (!world_to_voxel(P, p, v)) return;
        const int ox = (int)v.x, oy = (int)v.y, oz = (int)v.z;
        const int dx1[4] = {0, 1, 0, 1}, dy1[4] = {0, 0, 1, 1};
        const int dx2[12] = {0, 1, -1, 0, 1, 2, -1, 0, 1, 2, 0, 1};
        const int dy2[12] = {-1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2};
        for (int i = 0; i < 4; ++i)
            if (!ref_tap(ox + dx1[i], oy + dy1[i], oz - 1, bytes)) return;
        for (int i = 0; i < 12; ++i)
            if (!ref_tap(ox + dx2[i], oy + dy2[i], oz, bytes)) return;
        for (int i = 0; i < 12; ++i)
            if (!ref_tap(ox + dx2[i], oy + dy2[i], oz + 1, bytes)) return;
        for (int i = 0; i < 4; ++i)
            if (!ref_tap(ox + dx1[i], oy + dy1[i], oz + 2, bytes)) return;
    }
    // d_gtvi: trilinear with toInt = truncation; any invalid tap -> false,
    // 1.0. The 2x2x2 stencil is range-checked as a whole and its 8 values and
    // weights are loaded together; a voxel of an empty brick has weight 0, so
    // no occupancy lookup is needed here.
    template <bool O32>
    __device__ __forceinline__ bool trilinear(f3 v, float &out) const {
        out = 1.0f;
        if (!finite3(v)) return false;
        const int ox = (int)v.x, oy = (int)v.y, oz = (int)v.z;
        // 0 <= o and o + 1 < vdim on every axis
        if (!below3(ox, oy, oz, P.vdim[0] - 1, P.vdim[1] - 1, P.vdim[2] - 1)) return false;
        const float rx = v.x - (float)ox, ry = v.y - (float)oy, rz = v.z - (float)oz;
        // vox_index of the 8 corners from per-axis parts: brick index
        // bz*bdim1*bdim0 + by*bdim0 + bx (24-bit products: < 2^24 bricks)
        // and in-brick offset lz*64 + ly*8 + lx of the two coordinates on
        // each axis
        const unsigned bd0 = (unsigned)P.bdim[0], bd01 = (unsigned)(P.bdim[0] * P.bdim[1]);
        unsigned Bx[2], By[2], Bz[2], Lx[2], Ly[2], Lz[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            Bx[a] = (unsigned)(ox + a) >> 3;
            By[a] = __umul24((unsigned)(oy + a) >> 3, bd0);
            Bz[a] = __umul24((unsigned)(oz + a) >> 3, bd01);
            Lx[a] = (unsigned)(ox + a) & 7u;
            Ly[a] = ((unsigned)(oy + a) & 7u) << 3;
            Lz[a] = ((unsigned)(oz + a) & 7u) << 6;
        }
        float f[8];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const unsigned b = Bz[k >> 2] + By[(k >> 1) & 1] + Bx[k & 1];
            const unsigned l = Lz[k >> 2] | Ly[(k >> 1) & 1] | Lx[k & 1];
            ok &= load_bl<O32>(b, l, f[k]);
        }
        if (!ok) return false;
        const float fi = (1.0f - ry) * ((1.0f - rx) * f[0] + rx * f[1]) + ry * ((1.0f - rx) * f[2] + rx * f[3]);
        const float bi = (1.0f - ry) * ((1.0f - rx) * f[4] + rx * f[5]) + ry * ((1.0f - rx) * f[6] + rx * f[7]);
        out = (1.0f - rz) * fi + rz * bi;
        return true;
    }
};

struct RayConst {
    float K[4];
    f3 cam, ncc;
    float dist_cc, cos_max, r_cube, step_scale, total_max;
    int max_iter;
};

// Exact fast-forward of the march through a region whose voxels all read
// 1.0. While every tap reads 1.0 the march repeats the same step S:
// position += S*rd (the same FP32 product every time), total += S, iter += 1,
// and nothing else changes. These return how many further steps certainly
// keep the ROUNDF position inside the region: the bounds carry a 0.25-voxel
// margin, far above both the rounding drift of at most 64 FP32 additions of
// positions below 2^13 voxels (< 0.02 voxel) and the error of the reciprocal
// estimates (relative 2^-22 of at most 64 steps), so each skipped step's tap
// would have read 1.0 too. The skipped additions are performed one by
// one (a closed form per binade costs more VALU than the
// handful of additions a box skips on average), four per loop trip: the
// loop's own counter, compare and branch cost a wave as much as the adds.
constexpr int kSkipMax = 64;
constexpr float kSkipMargin = 0.25f;

// rounded position inside [lo, lo + size)^3 clipped to the grid:
// v in [lo - 0.5, hi + 0.5) on each axis
__device__ __forceinline__ int steps_in_box(f3 v, f3 d, int3 lo, int size, const int *vdim) {
    float k = (float)kSkipMax;
    const float vv[3] = {v.x, v.y, v.z}, dd[3] = {d.x, d.y, d.z};
    const int ll[3] = {lo.x, lo.y, lo.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float hi = (float)min(ll[a] + size - 1, vdim[a] - 1);
        const float bound = dd[a] > 0.0f ? hi + 0.5f - kSkipMargin : (float)ll[a] - 0.5f + kSkipMargin;
        const float t = (bound - vv[a]) * __builtin_amdgcn_rcpf(dd[a]);
        k = dd[a] != 0.0f ? fminf(k, t) : k;
    }
    return k >= 1.0f ? (int)k : 0;
}

// rounded position outside the grid: it stays outside while any axis does
// (below: v <= -0.5; above: v >= vdim - 0.5)
__device__ __forceinline__ int steps_outside(f3 v, f3 d, const int *vdim) {
    float k = 0.0f;
    const float vv[3] = {v.x, v.y, v.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool below = vv[a] <= -0.5f && dd[a] > 0.0f;
        const bool above = vv[a] >= (float)vdim[a] - 0.5f && dd[a] < 0.0f;
        const float bound = below ? -0.5f - kSkipMargin : (float)vdim[a] - 0.5f + kSkipMargin;
        const float t = (bound - vv[a]) * __builtin_amdgcn_rcpf(dd[a]);
        k = (below || above) ? fmaxf(k, t) : k;
    }
    k = fminf(k, (float)kSkipMax);
    return k >= 1.0f ? (int)k : 0;
}

// Tap statistics of the instrumented march (the reference's lookups: what
// it would request, not what this kernel loads).
struct RayStats {
    unsigned n_near = 0, n_tri = 0, n_out = 0, n_w0 = 0;
    unsigned long long b_near = 0, b_tri = 0;
    template <class TV, class TW>
    __device__ __forceinline__ void nearest(const Vol<TV, TW> &V, f3 v) {
        ++n_near;
        const int c = V.ref_nearest(v, b_near);
        n_out += c == 0;
        n_w0 += c == 1;
    }
    template <class TV, class TW>
    __device__ __forceinline__ void trilinear(const Vol<TV, TW> &V, f3 v) {
        ++n_tri;
        V.ref_trilinear(v, b_tri);
    }
};

template <class TV, class TW, bool STATS, bool O32>
__device__ __forceinline__ bool cast_ray(const Vol<TV, TW> &V, const RayConst &C, const float *Tcg, float px,
                                         float py, f3 &hit, RayStats &rs) {
    const float dist_min = C.dist_cc - C.r_cube, dist_max = C.dist_cc + C.r_cube;
    const f3 pe = transform(Tcg, back_project(px, py, C.K, dist_max));
    const f3 rd = normalised(sub(pe, C.cam));
    if (dot(rd, C.ncc) < C.cos_max) return false;
    const f3 ps = add(C.cam, scale(dist_min, rd));
    f3 vox;
    world_to_voxel(V.P, ps, vox);
    float tsdf;
    if (STATS) rs.nearest(V, vox);
    const bool valid = V.nearest_valid(vox, tsdf);
    bool coarse;
    if (!valid) coarse = true;
    else if (tsdf >= 0.0f) coarse = false;
    else return false;
    bool behind = false;
    float total = 0.0f;
    int iter = 0;
    // Brent cycle detection on the fine-search state (position, tsdf): the
    // next state is a function of it alone (the volume does not change), and
    // neither a hit nor the surface-behind exit lies on a cycle, so a state
    // seen before means the ray can only end as a miss at the iteration or
    // distance cap. Fine-mode rays that bounce across a voxel of F = -1.0
    // (common with FP16 values) do this; ending them at once saves up to
    // max_iter iterations each. Coarse steps only move forward, and the mode
    // never returns to coarse, so coarse states need no check.
    // The saved state is refreshed after 1, 2, 4, ... steps; the step count
    // since the last save and log2 of the interval share one register (a
    // match is a genuine repeat whatever the schedule).
    unsigned sv_x = 0xFFFFFFFFu, sv_y = 0xFFFFFFFFu, sv_z = 0xFFFFFFFFu, sv_t = 0xFFFFFFFFu, brent = 0;
    while (true) {
        if (!STATS && !coarse) {
            const unsigned bx = __float_as_uint(vox.x), by = __float_as_uint(vox.y), bz = __float_as_uint(vox.z),
                           bt = __float_as_uint(tsdf);
            if (((bx ^ sv_x) | (by ^ sv_y) | (bz ^ sv_z) | (bt ^ sv_t)) == 0u) return false;
            const bool save = (brent & 0xFFFFu) == (1u << (brent >> 16));
            sv_x = save ? bx : sv_x;
            sv_y = save ? by : sv_y;
            sv_z = save ? bz : sv_z;
            sv_t = save ? bt : sv_t;
            brent = (save ? ((brent >> 16) + 1) << 16 : brent) + 1u;
        }
        // fine step: max(s*tsdf, 1) for tsdf > 0, else min(s*tsdf, -1), as
        // +-max(|s*tsdf|, 1) (s > 0); both exits give a miss, so they share
        // one test
        ++iter;
        const float sm = fmaxf(fabsf(C.step_scale * tsdf), 1.0f);
        const float step = coarse ? C.step_scale : (tsdf > 0.0f ? sm : -sm);
        vox = add(vox, scale(step, rd));
        total += step;
        if ((int)(iter > C.max_iter) | (int)(total > C.total_max)) return false;
        if (STATS) rs.nearest(V, vox);
        // A ray that has left the grid for good can only end as a miss
        // (every later tap reads 1.0 and the step stays positive); the
        // instrumented form keeps marching to count the reference's taps.
        int3 lo;
        int size;
        if (V.template nearest<O32>(vox, rd, tsdf, lo, size) < 0 && !STATS) return false;
        if (!STATS && size != 0) {
            // 1.0 here and in the region around: skip the steps that stay
            // inside it (tsdf is 1.0, so the next step is the same S each time)
            const float S = coarse ? C.step_scale : fmaxf(C.step_scale * tsdf, 1.0f);
            const f3 d = scale(S, rd);
            const int k = size > 0 ? steps_in_box(vox, d, lo, size, V.P.vdim) : steps_outside(vox, d, V.P.vdim);
            {  // k = 0 changes nothing
                iter += k;
                int j = k;
                for (; j >= 4; j -= 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        vox = add(vox, d);
                        total += S;
                    }
                }
                if (j & 2) {
                    vox = add(vox, d);
                    total += S;
                    vox = add(vox, d);
                    total += S;
                }
                if (j & 1) {
                    vox = add(vox, d);
                    total += S;
                }
                if ((int)(iter > C.max_iter) | (int)(total > C.total_max)) return false;
            }
            continue;
        }
        if (fabsf(tsdf) < 1.0f) {
            if (STATS) rs.trilinear(V, vox);
            V.template trilinear<O32>(vox, tsdf);
            if (tsdf < 0.0f) {
                if (coarse) return false;
                behind = true;
                break;
            } else if (tsdf == 0.0f) {
                hit = voxel_to_world(V.P, vox);
                return true;
            } else if (coarse) {
                coarse = false;
            }
        }
    }
    (void)behind;
    float step = C.step_scale * tsdf;
    vox = add(vox, scale(step, rd));
    if (STATS) rs.trilinear(V, vox);
    V.template trilinear<O32>(vox, tsdf);
    step = C.step_scale * tsdf;
    vox = add(vox, scale(step, rd));
    hit = voxel_to_world(V.P, vox);
    return true;
}

// d_csnt_Compute_Surface_Normal_from_TSDF: central differences of trilinear
// values from 32 taps around the hit.
template <bool O32, class TV, class TW>
__device__ __forceinline__ bool surface_normal(const Vol<TV, TW> &V, f3 p, f3 &n) {
    f3 vox;
    if (!world_to_voxel(V.P, p, vox)) return false;
    const int ox = (int)vox.x, oy = (int)vox.y, oz = (int)vox.z;
    const float rx = vox.x - (float)ox, ry = vox.y - (float)oy, rz = vox.z - (float)oz;
    // front/back slices: 12 taps (dx2,dy2); foremost/backmost: 4 taps (dx1,dy1)
    const int dx1[4] = {0, 1, 0, 1}, dy1[4] = {0, 0, 1, 1};
    const int dx2[12] = {0, 1, -1, 0, 1, 2, -1, 0, 1, 2, 0, 1};
    const int dy2[12] = {-1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2};
    // every tap lies in [o-1, o+2]^3 and each of those extremes is used, so
    // the stencil is valid only if that box is inside the grid; the 32 taps
    // are then loaded without per-tap branches (all-or-nothing result)
    if (!V.in_grid(ox - 1, oy - 1, oz - 1) || !V.in_grid(ox + 2, oy + 2, oz + 2)) return false;
    float fm[4], fr[12], bk[12], bm[4];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) ok &= V.template load_at<O32>(ox + dx1[i], oy + dy1[i], oz - 1, fm[i]);
#pragma unroll
    for (int i = 0; i < 12; ++i) ok &= V.template load_at<O32>(ox + dx2[i], oy + dy2[i], oz, fr[i]);
#pragma unroll
    for (int i = 0; i < 12; ++i) ok &= V.template load_at<O32>(ox + dx2[i], oy + dy2[i], oz + 1, bk[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) ok &= V.template load_at<O32>(ox + dx1[i], oy + dy1[i], oz + 2, bm[i]);
    if (!ok) return false;
    auto Lz = [&](int i) { return (1.0f - rz) * fr[i] + rz * bk[i]; };
    float a4[4], b4[4], a2[2], b2[2];
    a4[0] = Lz(2); a4[1] = Lz(3); a4[2] = Lz(4); a4[3] = Lz(5);
    b4[0] = Lz(6); b4[1] = Lz(7); b4[2] = Lz(8); b4[3] = Lz(9);
    a2[0] = (1.0f - rx) * a4[0] + rx * a4[1];
    a2[1] = (1.0f - rx) * b4[0] + rx * b4[1];
    b2[0] = (1.0f - rx) * a4[2] + rx * a4[3];
    b2[1] = (1.0f - rx) * b4[2] + rx * b4[3];
    float gx = (1.0f - ry) * b2[0] + ry * b2[1];
    gx -= (1.0f - ry) * a2[0] + ry * a2[1];
    a4[0] = Lz(0); a4[1] = Lz(3); a4[2] = Lz(7); a4[3] = Lz(10);
    b4[0] = Lz(1); b4[1] = Lz(4); b4[2] = Lz(8); b4[3] = Lz(11);
    a2[0] = (1.0f - ry) * a4[0] + ry * a4[1];
    a2[1] = (1.0f - ry) * b4[0] + ry * b4[1];
    b2[0] = (1.0f - ry) * a4[2] + ry * a4[3];
    b2[1] = (1.0f - ry) * b4[2] + ry * b4[3];
    float gy = (1.0f - rx) * b2[0] + rx * b2[1];
    gy -= (1.0f - rx) * a2[0] + rx * a2[1];
    a4[0] = (1.0f - ry) * fm[0] + ry * fm[2];
    a4[1] = (1.0f - ry) * fr[3] + ry * fr[7];
    a4[2] = (1.0f - ry) * bk[3] + ry * bk[7];
    a4[3] = (1.0f - ry) * bm[0] + ry * bm[2];
    b4[0] = (1.0f - ry) * fm[1] + ry * fm[3];
    b4[1] = (1.0f - ry) * fr[4] + ry * fr[8];
    b4[2] = (1.0f - ry) * bk[4] + ry * bk[8];
    b4[3] = (1.0f - ry) * bm[1] + ry * bm[3];
    a2[0] = (1.0f - rz) * a4[0] + rz * a4[1];
    a2[1] = (1.0f - rz) * b4[0] + rz * b4[1];
    b2[0] = (1.0f - rz) * a4[2] + rz * a4[3];
    b2[1] = (1.0f - rz) * b4[2] + rz * b4[3];
    float gz = (1.0f - rx) * b2[0] + rx * b2[1];
    gz -= (1.0f - rx) * a2[0] + rx * a2[1];
    if (gx == 0.0f && gy == 0.0f && gz == 0.0f) return false;
    n = normalised({gx, gy, gz});
    return true;
}

// d_grvi_Get_RGB_Value_Interpolated (z_yooji_cuda_tsdf_renderer.cu:232-284)
// with its d_grvu taps (:192-229): trilinear colour over the 8 corners of
// toInt(vox), each channel rounded half away from zero and clamped
// (Vector3f::toUChar); a corner outside the grid or without weight makes the
// lookup fail, which leaves (0, 0, 0) (the caller ignores the return,
// :1171-1175). The stencil is range-checked as a whole and its 16 loads are
// issued together.
template <class TV, class TW>
__device__ __forceinline__ uint32_t rgb_interpolated(const Vol<TV, TW> &V, const uint32_t *col, f3 v) {
    if (!(fabsf(v.x) < 2147483520.0f && fabsf(v.y) < 2147483520.0f && fabsf(v.z) < 2147483520.0f)) return 0u;
    const int ox = (int)v.x, oy = (int)v.y, oz = (int)v.z;
    if (!V.in_grid(ox, oy, oz) || !V.in_grid(ox + 1, oy + 1, oz + 1)) return 0u;
    const float rx = v.x - (float)ox, ry = v.y - (float)oy, rz = v.z - (float)oz;
    uint32_t c[8];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const size_t idx = vox_index(V.P, ox + (k & 1), oy + ((k >> 1) & 1), oz + (k >> 2));
        ok &= VoxRec<TV, TW>::has_w(ld_rec<TV, TW>(V.vx, idx));
        c[k] = ldg(col, idx);
    }
    if (!ok) return 0u;
    uint32_t out = 0u;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float f[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = (float)((c[k] >> (8 * ch)) & 255u);
        const float fi = (1.0f - ry) * ((1.0f - rx) * f[0] + rx * f[1]) + ry * ((1.0f - rx) * f[2] + rx * f[3]);
        const float bi = (1.0f - ry) * ((1.0f - rx) * f[4] + rx * f[5]) + ry * ((1.0f - rx) * f[6] + rx * f[7]);
        const float r = (1.0f - rz) * fi + rz * bi;
        const int q = (r < 0.0f) ? (int)(r - 0.5f) : (int)(r + 0.5f);
        out |= (uint32_t)min(255, max(0, q)) << (8 * ch);
    }
    return out;
}

__device__ __forceinline__ void write_texture(const SessParams &P, int level, size_t t, uint32_t rgb) {
    const size_t plane = (size_t)P.lev[level].W * P.lev[level].H;
    stg(P.mtex[level], t, (uint8_t)(rgb & 255u));
    stg(P.mtex[level], plane + t, (uint8_t)((rgb >> 8) & 255u));
    stg(P.mtex[level], 2 * plane + t, (uint8_t)((rgb >> 16) & 255u));
}

// The raycast runs as two launches per level. k_raycast marches every ray
// (g_rmi's d_crtc part) with a small register footprint, so 8 waves per SIMD
// hide the march's dependent taps; a hit leaves its world point in the vertex
// map with w = 1, a miss writes its final sentinels. k_raycast_shade then
// evaluates the 32-tap gradient of every hit with all its loads in flight at
// once and writes the final vertex / normal / shade (camera frame).
// Work item and level of this workgroup in a multi-level launch.
__device__ __forceinline__ WorkItem level_item(const LevelWork &lw, int &level) {
    int b = blockIdx.x;
    int s = 0;
    while (s + 1 < lw.nseg && b >= lw.n[s]) {
        b -= lw.n[s];
        ++s;
    }
    level = lw.level[s];
    return lw.w[s][b];
}

// Pixel of thread t inside its 32x8 tile when each wave takes a block of
// BW x (64 / BW) rays (RayBlock: 8x8, 16x4 or 32x2). Neighbouring rays take
// similar paths, so squarer blocks diverge less; which block is fastest for a
// stage and workload is the autotuner's choice (the results are the same).
template <int RB>
__device__ __forceinline__ void ray_pixel(int t, int &px, int &py) {
    constexpr int BW = RB == kRay8x8 ? 8 : RB == kRay16x4 ? 16 : 32, BH = 64 / BW, PER_ROW = kTileW / BW;
    const int w = t >> 6, lane = t & 63;
    px = (w % PER_ROW) * BW + lane % BW;
    py = (w / PER_ROW) * BH + lane / BW;
}

// 8 waves per SIMD (at most 64 VGPRs): the march is VALU-bound with dependent
// taps, and the extra waves hide the taps' latency better than the few
// registers the compiler would add beyond 64.
// O32: the lookups take 32-bit offsets from the voxel-record base (every
// session of the launch has its records below 4 GB), as in k_raycast_shade.
template <class TV, class TW, bool STATS, int RB, bool O32>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_raycast(
    const SessParams *__restrict__ SP, LevelWork lw, unsigned long long *stats) {
    __shared__ uint32_t s_dist[kDistCells / 8];
    int level;
    const WorkItem wi = level_item(lw, level);
    const SessParams &P = SP[wi.sess];
    const SessState *st = P.st;
    const LevelDesc &L = P.lev[level];
    if (!st->fuse_ok || level >= P.levels || tile_out(L, wi.tile)) return;
    const int ncell = P.sdim[0] * P.sdim[1] * P.sdim[2];
    if (ncell <= kDistCells) {
        const uint32_t *src = (const uint32_t *)P.sdist;
        for (int i = threadIdx.x; i < (ncell + 7) / 8; i += kTileThreads) s_dist[i] = ldg(src, i);
    } else {
        for (int i = threadIdx.x; i < kDistCells / 8; i += kTileThreads) s_dist[i] = 0u;
    }
    __syncthreads();
    int px, py;
    ray_pixel<RB>(threadIdx.x, px, py);
    const int x = (wi.tile % L.tiles_x) * kTileW + px;
    const int y = (wi.tile / L.tiles_x) * kTileH + py;
    if (x >= L.W || y >= L.H) return;
    RayConst C;
#pragma unroll
    for (int k = 0; k < 4; ++k) C.K[k] = L.K[k];
    C.cam = {st->cam[0], st->cam[1], st->cam[2]};
    C.ncc = {st->ncc[0], st->ncc[1], st->ncc[2]};
    C.dist_cc = st->dist_cc;
    C.cos_max = st->cos_max;
    C.r_cube = P.r_cube;
    C.step_scale = P.step_scale;
    C.total_max = P.total_max;
    C.max_iter = P.max_iter;
    const Vol<TV, TW> V{P, P.vox, (const VS_LDS uint8_t *)s_dist};
    RayStats rs;
    f3 p;
    const M34 Tc = ld_m34(st->T_cg_f);
    const bool ok = cast_ray<TV, TW, STATS, O32>(V, C, Tc.m, (float)x, (float)y, p, rs);
    const size_t t = (size_t)y * L.W + x;
    if (ok) {
        st_mvtx(P.mvn[level], t, p.x, p.y, p.z);  // world hit; the shading makes it the camera-frame vertex
    } else {
        st_mvtx(P.mvn[level], t, -100.0f, -100.0f, -100.0f);
        st_mnrm(P.mvn[level], t, -100.0f, -100.0f, -100.0f);
        stg(P.mimg[level], t, (uint8_t)0);
        if (P.mtex[level]) write_texture(P, level, t, 0u);
    }
    if (STATS) {
        atomicAdd(stats + 0, (unsigned long long)rs.n_near);
        atomicAdd(stats + 1, (unsigned long long)rs.n_tri);
        atomicAdd(stats + 3, (unsigned long long)rs.n_out);
        atomicAdd(stats + 4, (unsigned long long)rs.n_w0);
        atomicAdd(stats + 5, rs.b_near);
        atomicAdd(stats + 6, rs.b_tri);
    }
}

// WV: the waves per SIMD the register allocation aims for (4: the compiler's
// own choice, ~126 VGPRs with all 64 stencil loads in flight; 5: 91, fewer
// loads in flight); the same results for each, so the autotuner picks. (6
// waves, 80 VGPRs, spills a few registers and was slower still.) O32: the
// stencil's loads take 32-bit offsets from the array bases (every session
// of the launch has its voxel records below 4 GB): 62 VGPRs instead of 101 at
// WV 4, the same results.
template <class TV, class TW, bool STATS, int RB, int WV, bool O32>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(WV == 4 ? 1 : WV, WV == 4 ? 8 : WV)))
void k_raycast_shade(
    const SessParams *__restrict__ SP, LevelWork lw, unsigned long long *stats) {
    int level;
    const WorkItem wi = level_item(lw, level);
    const SessParams &P = SP[wi.sess];
    const SessState *st = P.st;
    const LevelDesc &L = P.lev[level];
    if (!st->fuse_ok || level >= P.levels || tile_out(L, wi.tile)) return;
    int px, py;
    ray_pixel<RB>(threadIdx.x, px, py);
    const int x = (wi.tile % L.tiles_x) * kTileW + px;
    const int y = (wi.tile / L.tiles_x) * kTileH + py;
    if (x >= L.W || y >= L.H) return;
    const size_t t = (size_t)y * L.W + x;
    // the march's output (it always runs just before): a world hit point,
    // or -100 for a miss, which is final already
    const MapXYZ h = ld_mvtx(P.mvn[level], t);
    if (h.x == -100.0f) return;
    const Vol<TV, TW> V{P, P.vox, nullptr};
    const f3 p = {h.x, h.y, h.z};
    if (STATS) {
        unsigned long long b = 0;
        V.ref_normal(p, b);
        atomicAdd(stats + 7, b);
    }
    f3 n;
    if (surface_normal<O32>(V, p, n)) {
        const M34 Tg = ld_m34(st->T_gc_f);
        const float *T = Tg.m;
        const f3 pc = transform(T, p);
        const f3 nc = {T[0] * n.x + T[1] * n.y + T[2] * n.z, T[4] * n.x + T[5] * n.y + T[6] * n.z,
                       T[8] * n.x + T[9] * n.y + T[10] * n.z};
        st_mvtx(P.mvn[level], t, pc.x, pc.y, pc.z);
        st_mnrm(P.mvn[level], t, nc.x, nc.y, nc.z);
        // light (0,0,-1): n.l = 0*nx + 0*ny - nz (z_yooji_object_scanner.cpp:1313)
        const float sa = nc.x * 0.0f + nc.y * 0.0f + nc.z * -1.0f;
        stg(P.mimg[level], t, (uint8_t)(fmaxf(0.0f, fminf(255.0f, (0.8f * sa + 0.2f) * 255.0f))));
        if (P.mtex[level]) {
            f3 pv;
            world_to_voxel(P, p, pv);
            write_texture(P, level, t, rgb_interpolated(V, P.vcol, pv));
        }
        if (STATS) atomicAdd(stats + 2, 1ull);
    } else {
        st_mvtx(P.mvn[level], t, -100.0f, -100.0f, -100.0f);
        st_mnrm(P.mvn[level], t, -100.0f, -100.0f, -100.0f);
        stg(P.mimg[level], t, (uint8_t)0);
        if (P.mtex[level]) write_texture(P, level, t, 0u);
    }
}

// ---------------------------------------------------------------------------
// volume maintenance
// ---------------------------------------------------------------------------
template <class TV, class TW>
__global__ void k_volume_init(const SessParams *__restrict__ SP, size_t nvox, uint32_t nwords) {
    const SessParams &P = SP[0];
    const typename VoxRec<TV, TW>::T empty = VoxRec<TV, TW>::make(1.0f, 0.0f);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
        st_rec<TV, TW>(P.vox, i, empty);
        if (P.vcol) stg(P.vcol, i, 0u);
    }
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += (size_t)gridDim.x * blockDim.x) {
        stg(P.bvis, w, 0u);
        stg(P.bocc, w, 0u);
        stg(P.bband, w, 0u);
        if (w < (size_t)P.sband_words) stg(P.sband, w, 0u);
    }
}

__global__ void k_sband_clear(const SessParams *__restrict__ SP) {
    const SessParams &P = SP[0];
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < P.sband_words; w += gridDim.x * blockDim.x) stg(P.sband, (size_t)w, 0u);
}

// brick-major -> linear x-fastest, converted to float
template <class TV, class TW>
__global__ void k_volume_to_linear(const SessParams *__restrict__ SP, size_t nvox, float *F, float *Wt) {
    const SessParams &P = SP[0];
    const size_t WH = (size_t)P.vdim[0] * P.vdim[1];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i / WH), y = (int)((i / P.vdim[0]) % P.vdim[1]), x = (int)(i % P.vdim[0]);
        const size_t j = vox_index(P, x, y, z);
        const typename VoxRec<TV, TW>::T r = ld_rec<TV, TW>(P.vox, j);
        if (F) F[i] = VoxRec<TV, TW>::val(r);
        if (Wt) Wt[i] = VoxRec<TV, TW>::w(r);
    }
}
// Whole bricks between the volume planes and a packed array (snapshots):
// values, then weights, then colours of the listed bricks, one workgroup per
// brick, 16-byte vectors (every offset is a multiple of 512 bytes).
__device__ __forceinline__ void copy16(uint8_t *dst, const uint8_t *src, size_t bytes) {
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    for (size_t k = threadIdx.x; k < bytes / 16; k += blockDim.x) d[k] = s[k];
}
__global__ void k_brick_pack(const SessParams *__restrict__ SP, const uint32_t *idx, unsigned n, int sv, int sw,
                             int sc, uint8_t *out, int unpack) {
    const SessParams &P = SP[0];
    const size_t bv = (size_t)kBrickVox * sv, bw = (size_t)kBrickVox * sw, bc = (size_t)kBrickVox * sc;
    const bool r4 = vox_rec_bytes(sv == 2, sw == 1) == 4;
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
        const size_t b = idx[i];
        uint8_t *pv = out + i * bv, *pw = out + n * bv + i * bw, *pc = out + n * (bv + bw) + i * bc;
        uint8_t *vc = sc ? (uint8_t *)P.vcol + b * bc : nullptr;
        // values and weights as separate planes of raw storage bits: the
        // record's value bits (bits 0-15 or word 0) and weight bits (bits
        // 16-23, or word 1)
        for (int v = threadIdx.x; v < kBrickVox; v += blockDim.x) {
            const size_t j = b * kBrickVox + v;
            if (r4) {
                uint32_t *rec = (uint32_t *)P.vox + j;
                if (unpack)
                    stg(rec, 0, (uint32_t)ldg((const uint16_t *)pv, v) | ((uint32_t)ldg(pw, v) << 16));
                else {
                    const uint32_t r = ldg(rec, 0);
                    stg((uint16_t *)pv, v, (uint16_t)(r & 0xFFFFu));
                    stg(pw, v, (uint8_t)(r >> 16));
                }
            } else {
                uint32_t *rec = (uint32_t *)P.vox + 2 * j;
                if (unpack) {
                    stg(rec, 0, sv == 4 ? ldg((const uint32_t *)pv, v) : (uint32_t)ldg((const uint16_t *)pv, v));
                    stg(rec, 1, sw == 4 ? ldg((const uint32_t *)pw, v) : (uint32_t)ldg(pw, v));
                } else {
                    const uint32_t r0 = ldg(rec, 0), r1 = ldg(rec, 1);
                    if (sv == 4) stg((uint32_t *)pv, v, r0);
                    else stg((uint16_t *)pv, v, (uint16_t)r0);
                    if (sw == 4) stg((uint32_t *)pw, v, r1);
                    else stg(pw, v, (uint8_t)r1);
                }
            }
        }
        if (sc) {
            if (unpack) copy16(vc, pc, bc);
            else copy16(pc, vc, bc);
        }
    }
}
void launch_brick_pack(const SessParams *P, const uint32_t *idx, unsigned n, int sv, int sw, int sc, uint8_t *buf,
                       int unpack, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_brick_pack, dim3(n < 4096u ? n : 4096u), dim3(256), 0, s, P, idx, n, sv, sw, sc, buf,
                           unpack);
}

// g_uvv_Update_Valid_Volume (z_yooji_cuda_volume_integrator.cu:718-799): every
// brick of the last frame's visibility mask whose voxels are all invalid in
// `valid` (linear x-fastest, one byte per voxel) is reset to the initial
// state (TSDF 1, weight 0, colour 0) and leaves the mask; it also leaves the
// occupancy bits, since it holds no weight any more. One workgroup per brick,
// 2 voxels per thread.
template <class TV, class TW>
__global__ __launch_bounds__(kTileThreads) void k_carve(const SessParams *__restrict__ SP, const uint8_t *valid,
                                                        unsigned *carved) {
    const SessParams &P = SP[0];
    const uint32_t nb = (uint32_t)P.bdim[0] * P.bdim[1] * P.bdim[2];
    const int t = threadIdx.x;
    const int xl = t & 7, yl = (t >> 3) & 7, zl = t >> 6;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        if (!((ldg(P.bvis, b >> 5) >> (b & 31)) & 1u)) continue;
        const int bx = b % P.bdim[0], by = (b / P.bdim[0]) % P.bdim[1], bz = b / (P.bdim[0] * P.bdim[1]);
        const int x = bx * kBrick + xl, y = by * kBrick + yl;
        bool any = false;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int z = bz * kBrick + zl + 4 * h;
            any |= ldg(valid, ((size_t)z * P.vdim[1] + y) * P.vdim[0] + x) != 0;
        }
        if (__syncthreads_or(any)) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int li = t + kTileThreads * h;
            st_vw<TV, TW>(P.vox, (size_t)b * kBrickVox + li, 1.0f, 0.0f);
            if (P.vcol) stg(P.vcol, (size_t)b * kBrickVox + li, 0u);
        }
        if (t == 0) {
            atomicAnd(P.bvis + (b >> 5), ~(1u << (b & 31)));
            atomicAnd(P.bocc + (b >> 5), ~(1u << (b & 31)));
            atomicAnd(P.bband + (b >> 5), ~(1u << (b & 31)));
            atomicAdd(carved, 1u);
        }
    }
}

// brick-major RGB -> three linear x-fastest planes
__global__ void k_color_to_linear(const SessParams *__restrict__ SP, size_t nvox, uint8_t *rgb) {
    const SessParams &P = SP[0];
    const size_t WH = (size_t)P.vdim[0] * P.vdim[1];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i / WH), y = (int)((i / P.vdim[0]) % P.vdim[1]), x = (int)(i % P.vdim[0]);
        const uint32_t c = ldg(P.vcol, vox_index(P, x, y, z));
        rgb[i] = (uint8_t)(c & 255u);
        rgb[nvox + i] = (uint8_t)((c >> 8) & 255u);
        rgb[2 * nvox + i] = (uint8_t)((c >> 16) & 255u);
    }
}
void launch_color_to_linear(const SessParams *P, size_t nvox, uint8_t *rgb, hipStream_t s) {
    hipLaunchKernelGGL(k_color_to_linear, dim3(2048), dim3(256), 0, s, P, nvox, rgb);
}

template <class TV, class TW>
__global__ void k_volume_from_linear(const SessParams *__restrict__ SP, size_t nvox, const float *F,
                                     const float *Wt) {
    const SessParams &P = SP[0];
    const size_t WH = (size_t)P.vdim[0] * P.vdim[1];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i / WH), y = (int)((i / P.vdim[0]) % P.vdim[1]), x = (int)(i % P.vdim[0]);
        const size_t j = vox_index(P, x, y, z);
        st_vw<TV, TW>(P.vox, j, F[i], Wt[i]);
    }
}
// bocc and the band bits from the volume (after an upload or a restore): one
// workgroup per brick; sband is cleared first
template <class TV, class TW>
__global__ __launch_bounds__(kTileThreads) void k_occupancy(const SessParams *__restrict__ SP, uint32_t nbricks) {
    const SessParams &P = SP[0];
    for (uint32_t b = blockIdx.x; b < nbricks; b += gridDim.x) {
        bool any = false, band = false;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const size_t i = (size_t)b * kBrickVox + threadIdx.x + kTileThreads * h;
            const typename VoxRec<TV, TW>::T r = ld_rec<TV, TW>(P.vox, i);
            const bool w = VoxRec<TV, TW>::has_w(r);
            any |= w;
            band |= w && VoxRec<TV, TW>::val(r) != 1.0f;
        }
        const int cany = __syncthreads_or(any), cband = __syncthreads_or(band);
        if (threadIdx.x == 0) {
            const uint32_t bb = 1u << (b & 31);
            if (cany) atomicOr(P.bocc + (b >> 5), bb);
            else atomicAnd(P.bocc + (b >> 5), ~bb);
            if (cband) {
                atomicOr(P.bband + (b >> 5), bb);
                const int bx = b % P.bdim[0], by = (b / P.bdim[0]) % P.bdim[1], bz = b / (P.bdim[0] * P.bdim[1]);
                const int sb = super_of_brick(P, bx, by, bz);
                atomicOr(P.sband + (sb >> 5), 1u << (sb & 31));
            } else {
                atomicAnd(P.bband + (b >> 5), ~bb);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_vis_clear(const SessParams *P, int nsess, BrickList bl, hipStream_t s) {
    hipLaunchKernelGGL(k_vis_clear, dim3(8, nsess), dim3(256), 0, s, P, nsess, bl);
}
void launch_vis_mark(const SessParams *P, const WorkItem *work, int n, int nsess, uint32_t max_words, BrickList bl,
                     hipStream_t s) {
    if (n <= 0 || nsess <= 0 || max_words == 0) return;
    hipLaunchKernelGGL(k_vis_mark, dim3(n), dim3(kTileThreads), 0, s, P, work);
    const uint32_t per = kVisListBricks / 32u;  // bitmap words per k_vis_list workgroup
    hipLaunchKernelGGL(k_vis_list, dim3((max_words + per - 1) / per, nsess), dim3(kTileThreads), 0, s, P, bl);
}

#define VS_DISPATCH(tt, wt, FN, ...)                                                \
    do {                                                                            \
        if ((tt) == 0 && (wt) == 0) FN<float, float>(__VA_ARGS__);                  \
        else if ((tt) == 0 && (wt) == 1) FN<float, uint8_t>(__VA_ARGS__);           \
        else if ((tt) == 1 && (wt) == 0) FN<__half, float>(__VA_ARGS__);            \
        else FN<__half, uint8_t>(__VA_ARGS__);                                      \
    } while (0)

template <class TV, class TW>
static void integrate_t(const SessParams *P, BrickList bl, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_integrate<TV, TW>), dim3(grid), dim3(kTileThreads), 0, s, P, bl);
}
void launch_integrate(const SessParams *P, int nsess, BrickList bl, int tt, int wt, int grid, int sd_slices,
                      hipStream_t s) {
    VS_DISPATCH(tt, wt, integrate_t, P, bl, grid, s);
    launch_sdist(P, nsess, sd_slices, 0, s);
}

template <class TV, class TW>
static void march_t(const SessParams *P, const LevelWork &lw, unsigned long long *stats, int rb, bool o32,
                    hipStream_t s) {
    const int n = lw.blocks();
    if (n <= 0) return;
#define VS_MARCH(RB_)                                                                                          \
    do {                                                                                                       \
        if (o32)                                                                                               \
            hipLaunchKernelGGL((k_raycast<TV, TW, false, RB_, true>), dim3(n), dim3(kTileThreads), 0, s, P, lw, \
                               stats);                                                                         \
        else                                                                                                   \
            hipLaunchKernelGGL((k_raycast<TV, TW, false, RB_, false>), dim3(n), dim3(kTileThreads), 0, s, P, lw, \
                               stats);                                                                         \
    } while (0)
    if (stats)
        hipLaunchKernelGGL((k_raycast<TV, TW, true, kRay8x8, false>), dim3(n), dim3(kTileThreads), 0, s, P, lw, stats);
    else if (rb == kRay16x4)
        VS_MARCH(kRay16x4);
    else if (rb == kRay32x2)
        VS_MARCH(kRay32x2);
    else
        VS_MARCH(kRay8x8);
#undef VS_MARCH
}
template <class TV, class TW>
static void shade_t(const SessParams *P, const LevelWork &lw, unsigned long long *stats, int rb, int waves,
                    bool o32, hipStream_t s) {
    const int n = lw.blocks();
    if (n <= 0) return;
#define VS_SHADE(RB_, WV_)                                                                                      \
    do {                                                                                                        \
        if (o32)                                                                                                \
            hipLaunchKernelGGL((k_raycast_shade<TV, TW, false, RB_, WV_, true>), dim3(n), dim3(kTileThreads), 0, \
                               s, P, lw, stats);                                                                \
        else                                                                                                    \
            hipLaunchKernelGGL((k_raycast_shade<TV, TW, false, RB_, WV_, false>), dim3(n), dim3(kTileThreads),  \
                               0, s, P, lw, stats);                                                             \
    } while (0)
    if (stats)
        hipLaunchKernelGGL((k_raycast_shade<TV, TW, true, kRay32x2, 4, false>), dim3(n), dim3(kTileThreads), 0, s, P,
                           lw, stats);
    else if (rb == kRay8x8 && waves == 5)
        VS_SHADE(kRay8x8, 5);
    else if (rb == kRay8x8)
        VS_SHADE(kRay8x8, 4);
    else if (waves == 5)
        VS_SHADE(kRay32x2, 5);
    else
        VS_SHADE(kRay32x2, 4);
#undef VS_SHADE
}
LevelWork level_work(const WorkItem *const work[kMaxLevels], const int n[kMaxLevels], int levels) {
    LevelWork lw{};
    for (int l = levels - 1; l >= 0; --l)  // coarse levels first: their launches are the least parallel
        if (n[l] > 0) {
            lw.w[lw.nseg] = work[l];
            lw.n[lw.nseg] = n[l];
            lw.level[lw.nseg] = l;
            ++lw.nseg;
        }
    return lw;
}
void launch_raycast_march(const SessParams *P, const LevelWork &lw, int tt, int wt, unsigned long long *stats,
                          int rb, bool o32, hipStream_t s) {
    VS_DISPATCH(tt, wt, march_t, P, lw, stats, rb, o32, s);
}
void launch_raycast_shade(const SessParams *P, const LevelWork &lw, int tt, int wt, unsigned long long *stats,
                          int rb, int waves, bool o32, hipStream_t s) {
    VS_DISPATCH(tt, wt, shade_t, P, lw, stats, rb, waves, o32, s);
}
static LevelWork one_level(const WorkItem *work, int n, int level) {
    LevelWork lw{};
    if (n > 0) {
        lw.w[0] = work;
        lw.n[0] = n;
        lw.level[0] = level;
        lw.nseg = 1;
    }
    return lw;
}
void launch_raycast(const SessParams *P, const WorkItem *work, int n, int level, int tt, int wt, bool o32,
                    hipStream_t s) {
    const LevelWork lw = one_level(work, n, level);
    launch_raycast_march(P, lw, tt, wt, nullptr, kRay8x8, o32, s);
    launch_raycast_shade(P, lw, tt, wt, nullptr, kRay8x8, 4, o32, s);
}
void launch_raycast_stats(const SessParams *P, const WorkItem *work, int n, int level, int tt, int wt,
                          unsigned long long *stats, hipStream_t s) {
    const LevelWork lw = one_level(work, n, level);
    launch_raycast_march(P, lw, tt, wt, stats, kRay8x8, false, s);
    launch_raycast_shade(P, lw, tt, wt, stats, kRay32x2, 4, false, s);
}

static int grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    return (int)(g > 8192 ? 8192 : (g == 0 ? 1 : g));
}
template <class TV, class TW>
static void init_t(const SessParams *P, size_t nvox, uint32_t nwords, hipStream_t s) {
    hipLaunchKernelGGL((k_volume_init<TV, TW>), dim3(grid_for(nvox)), dim3(256), 0, s, P, nvox, nwords);
}
void launch_volume_init(const SessParams *P, size_t nvox, int tt, int wt, uint32_t nwords, int sd_slices,
                        hipStream_t s) {
    VS_DISPATCH(tt, wt, init_t, P, nvox, nwords, s);
    launch_sdist(P, 1, sd_slices, 1, s);
}
template <class TV, class TW>
static void to_lin_t(const SessParams *P, size_t nvox, float *F, float *Wt, hipStream_t s) {
    hipLaunchKernelGGL((k_volume_to_linear<TV, TW>), dim3(grid_for(nvox)), dim3(256), 0, s, P, nvox, F, Wt);
}
void launch_volume_to_linear(const SessParams *P, size_t nvox, int tt, int wt, float *F, float *Wt, hipStream_t s) {
    VS_DISPATCH(tt, wt, to_lin_t, P, nvox, F, Wt, s);
}
template <class TV, class TW>
static void from_lin_t(const SessParams *P, size_t nvox, const float *F, const float *Wt, hipStream_t s) {
    hipLaunchKernelGGL((k_volume_from_linear<TV, TW>), dim3(grid_for(nvox)), dim3(256), 0, s, P, nvox, F, Wt);
}
void launch_volume_from_linear(const SessParams *P, size_t nvox, int tt, int wt, const float *F, const float *Wt,
                               hipStream_t s) {
    VS_DISPATCH(tt, wt, from_lin_t, P, nvox, F, Wt, s);
}
template <class TV, class TW>
static void occupancy_t(const SessParams *P, uint32_t nbricks, int g, hipStream_t s) {
    hipLaunchKernelGGL((k_occupancy<TV, TW>), dim3(g), dim3(kTileThreads), 0, s, P, nbricks);
}
void launch_occupancy_rebuild(const SessParams *P, uint32_t nbricks, int tt, int wt, int sd_slices, hipStream_t s) {
    int g = (int)(nbricks > 16384 ? 16384 : nbricks);
    if (g <= 0) return;
    hipLaunchKernelGGL(k_sband_clear, dim3(16), dim3(256), 0, s, P);
    VS_DISPATCH(tt, wt, occupancy_t, P, nbricks, g, s);
    launch_sdist(P, 1, sd_slices, 1, s);
}

template <class TV, class TW>
static void carve_t(const SessParams *P, const uint8_t *valid, unsigned *carved, hipStream_t s) {
    hipLaunchKernelGGL((k_carve<TV, TW>), dim3(4096), dim3(kTileThreads), 0, s, P, valid, carved);
}
void launch_carve(const SessParams *P, int tt, int wt, const uint8_t *valid, unsigned *carved, hipStream_t s) {
    VS_DISPATCH(tt, wt, carve_t, P, valid, carved, s);
}

// The session's model maps of `level` written into caller buffers (any may be
// null), in the layout of rmi_Render_Maps_for_ICP's outputs
// (z_yooji_cuda_tsdf_renderer.cu:1642-1656, g_rmi :1069-1209): depth = camera
// z of the hit or 0, vertex and normal 3 floats per pixel (-100 for a miss),
// normal image one byte per pixel. Grid-stride over the pixels.
__global__ __launch_bounds__(256) void k_maps_out(const SessParams *__restrict__ SP, int level, float *depth,
                                                  float *vtx, float *nrm, uint8_t *img) {
    const SessParams &P = SP[0];
    const size_t n = (size_t)P.lev[level].W * P.lev[level].H;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const MapXYZ v = ld_mvtx(P.mvn[level], i);
        if (depth) stg(depth, i, v.x == -100.0f ? 0.0f : v.z);
        if (vtx) {
            stg(vtx, 3 * i, v.x);
            stg(vtx, 3 * i + 1, v.y);
            stg(vtx, 3 * i + 2, v.z);
        }
        if (nrm) {
            const MapXYZ q = ld_mnrm(P.mvn[level], i);
            stg(nrm, 3 * i, q.x);
            stg(nrm, 3 * i + 1, q.y);
            stg(nrm, 3 * i + 2, q.z);
        }
        if (img) stg(img, i, ldg(P.mimg[level], i));
    }
}
void launch_maps_out(const SessParams *P, int level, size_t npx, float *depth, float *vtx, float *nrm, uint8_t *img,
                     hipStream_t s) {
    const size_t nb = (npx + 255) / 256;
    const int g = (int)(nb < 1024 ? nb : 1024);
    if (g > 0) hipLaunchKernelGGL(k_maps_out, dim3(g), dim3(256), 0, s, P, level, depth, vtx, nrm, img);
}

}  // namespace vs
