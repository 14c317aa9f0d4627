// vs_img.hip — image stages of the fusion step: bilateral depth filter (A2-A3),
// depth pyramid (A4) and the vertex/normal-map operator (A6-A7).
//
// Launch shape: one 256-thread workgroup (4 wave64) per 32 x 8 output tile,
// tiles of every session of a batch concatenated in one work list, so a
// single launch serves the whole batch. Rows of 32 pixels keep global loads
// and stores coalesced (a wave covers two 128-byte rows of a float image).
#include "vs_device.hpp"
#include "vs_launch.hpp"

namespace vs {

constexpr int kHalo = 3;  // largest bilateral radius supported

// The 2R+1 x 2R+1 window of one pixel (tile coordinates of the centre): sum
// of w and of w (d - c), and whether a tap is <= 0. The radius is a
// compile-time constant (1, 2 or 3, the radii a session accepts), so the
// loops unroll. The spatial products g[i] g[j] come from the session
// parameters (uniform, in scalar registers) and the range weight is one
// hardware base-2 exponential of the pre-scaled squared difference.
template <int R, int TW>
__device__ __forceinline__ void bil_window(const float (*tile)[TW], int cy, int cx, const float *wij, float c,
                                           float nk, float &s, float &t, bool &zero) {
    constexpr int r = R, n = 2 * R + 1;
#pragma unroll
    for (int i = -r; i <= r; ++i) {
#pragma unroll
        for (int j = -r; j <= r; ++j) {
            const float p = tile[cy + i][cx + j];
            zero |= (p <= 0.0f);
            const float d = p - c;
            const float w = wij[(i + r) * n + (j + r)] * __builtin_amdgcn_exp2f((d * d) * nk);
            t = fmaf(w, d, t);
            s += w;
        }
    }
}

// Bilateral filter of the raw uint16 mm frame into pyramid level 0 (metres).
// Reference: d_bilateral_filter (z_yooji_cuda_img_proc.cu:103-141): weight
// g[i] g[j] exp(-(d - c)^2 / (2 sigma_r^2)), output 0 if any tap <= 0. Taps are
// point samples with clamp-to-edge addressing; the sum is accumulated about
// the centre value (c + sum w (d - c) / sum w), which is the same weighted mean
// with a smaller rounding error.
__global__ __launch_bounds__(kTileThreads) void k_bilateral(const SessParams *__restrict__ SP,
                                                            const WorkItem *__restrict__ work) {
    __shared__ float tile[kTileH + 2 * kHalo][kTileW + 2 * kHalo];
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    const LevelDesc &L = P.lev[0];
    if (tile_out(L, wi.tile) || !P.st->active) return;
    const int W = L.W, H = L.H;
    const int x0 = (wi.tile % L.tiles_x) * kTileW, y0 = (wi.tile / L.tiles_x) * kTileH;
    const int R = P.bil_radius;
    constexpr int TW = kTileW + 2 * kHalo, TH = kTileH + 2 * kHalo;
    for (int i = threadIdx.x; i < TW * TH; i += kTileThreads) {
        int ly = i / TW, lx = i - ly * TW;
        int gx = min(max(x0 + lx - kHalo, 0), W - 1);
        int gy = min(max(y0 + ly - kHalo, 0), H - 1);
        tile[ly][lx] = raw_depth(P, (size_t)gy * W + gx);
    }
    __syncthreads();
    const int lx = threadIdx.x % kTileW, ly = threadIdx.x / kTileW;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H) return;
    const float c = tile[ly + kHalo][lx + kHalo];
    float s = 0.0f, t = 0.0f;
    bool zero = false;
    const float nk = P.bil_nk;
    const int cy = ly + kHalo, cx = lx + kHalo;
    // session creation accepts radii 1..3 only; each radius has its own
    // complete path to the store
    switch (R) {
    case 1:
        bil_window<1, TW>(tile, cy, cx, P.bil_w, c, nk, s, t, zero);
        stg(P.filt[0], (size_t)y * W + x, zero ? 0.0f : c + t / s);
        break;
    case 2:
        bil_window<2, TW>(tile, cy, cx, P.bil_w, c, nk, s, t, zero);
        stg(P.filt[0], (size_t)y * W + x, zero ? 0.0f : c + t / s);
        break;
    case 3:
        bil_window<3, TW>(tile, cy, cx, P.bil_w, c, nk, s, t, zero);
        stg(P.filt[0], (size_t)y * W + x, zero ? 0.0f : c + t / s);
        break;
    }
}

// One pyramid step, level l-1 -> l: the separable [1 4 6 4 1]^2/256 blur of
// level l-1 sampled at (2i+1, 2j+1), 0 where the 5x5 window leaves the image.
// Reference: cpdd_Construct_Pyramids_Down_Depth + sdg5 + ddh
// (z_yooji_image_processor.cpp:368-407, 857-940, 567-596). Only the decimated
// pixels are computed.
__global__ __launch_bounds__(kTileThreads) void k_pyr_down(const SessParams *__restrict__ SP,
                                                           const WorkItem *__restrict__ work, int l) {
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    const LevelDesc &L = P.lev[l];
    if (l >= P.levels || tile_out(L, wi.tile) || !P.st->active) return;
    const int Wi = P.lev[l - 1].W, Hi = P.lev[l - 1].H;
    const int x = (wi.tile % L.tiles_x) * kTileW + threadIdx.x % kTileW;
    const int y = (wi.tile / L.tiles_x) * kTileH + threadIdx.x / kTileW;
    if (x >= L.W || y >= L.H) return;
    const int i = 2 * x + 1, j = 2 * y + 1;
    float v = 0.0f;
    if (i >= 2 && i <= Wi - 3 && j >= 2 && j <= Hi - 3) {
        const float *in = P.filt[l - 1];
        float h[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const size_t row = (size_t)(j + r - 2) * Wi;
            h[r] = 1.0f * ldg(in, row + i - 2) + 1.0f * ldg(in, row + i + 2) + 4.0f * ldg(in, row + i - 1) +
                   4.0f * ldg(in, row + i + 1) + 6.0f * ldg(in, row + i);
        }
        v = (1.0f * h[0] + 1.0f * h[4] + 4.0f * h[1] + 4.0f * h[3] + 6.0f * h[2]) / (16.0f * 16.0f);
    }
    stg(P.filt[l], (size_t)y * L.W + x, v);
}

// Vertex and normal maps of pyramid level l (AoS float3, -100 invalid).
// Reference: cpm_Compute_Point_Map / enm_Estimate_Normal_Map
// (z_yooji_image_processor.cpp:119-227): n = normalize(dz/dx, dz/dy, -1) from
// central differences, 1-pixel border skipped, |dx| or |dy| < 1e-7 rejected,
// and grazing normals (-n_z <= cos 60 deg) invalidate vertex and normal.
// Evaluated on the input vertex map (order-independent form).
__global__ __launch_bounds__(kTileThreads) void k_vnmap(const SessParams *__restrict__ SP,
                                                        const WorkItem *__restrict__ work, int l, float *vtx,
                                                        float *nrm) {
    const WorkItem wi = work[blockIdx.x];
    const SessParams &P = SP[wi.sess];
    const LevelDesc &L = P.lev[l];
    const int W = L.W, H = L.H;
    const int x = (wi.tile % L.tiles_x) * kTileW + threadIdx.x % kTileW;
    const int y = (wi.tile / L.tiles_x) * kTileH + threadIdx.x / kTileW;
    if (x >= W || y >= H) return;
    const float *D = P.filt[l];
    const float *K = L.K;
    auto vert = [&](int xx, int yy, f3 &v) -> bool {
        float d = ldg(D, (size_t)yy * W + xx);
        if (d <= 0.0f) return false;
        v = {d * ((float)xx - K[2]) / K[0], d * ((float)yy - K[3]) / K[1], d};
        return true;
    };
    const size_t t = (size_t)y * W + x;
    f3 c;
    bool vc = vert(x, y, c);
    float o[6] = {-100.f, -100.f, -100.f, -100.f, -100.f, -100.f};
    if (vc) {
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
    }
    if (vc && x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
        f3 x1, x2, y1, y2;
        if (vert(x - 1, y, x1) && vert(x + 1, y, x2) && vert(x, y - 1, y1) && vert(x, y + 1, y2)) {
            float dx = x2.x - x1.x, dy = y2.y - y1.y;
            if (fabsf(dx) >= 1.0e-7f && fabsf(dy) >= 1.0e-7f) {
                float gx = (x2.z - x1.z) / dx, gy = (y2.z - y1.z) / dy;
                float nn = sqrtf(gx * gx + gy * gy + 1.0f);
                float nx = gx / nn, ny = gy / nn, nz = -1.0f / nn;
                if (-nz <= 0.5f) {
                    o[0] = o[1] = o[2] = -100.f;
                } else {
                    o[3] = nx; o[4] = ny; o[5] = nz;
                }
            }
        }
    }
    vtx[3 * t + 0] = o[0]; vtx[3 * t + 1] = o[1]; vtx[3 * t + 2] = o[2];
    nrm[3 * t + 0] = o[3]; nrm[3 * t + 1] = o[4]; nrm[3 * t + 2] = o[5];
}

// The reference CPU loop's in-place order (enm_Estimate_Normal_Map,
// z_yooji_image_processor.cpp:155-227, which writes the grazing invalidation
// into the vertex map it is still reading): pixels after a grazing pixel in
// raster order see its vertex as -100. A grazing pixel is invalidated only
// when neither its left nor its upper neighbour was:
//     inv(i, j) = c(i, j) & !inv(i, j-1) & !inv(i-1, j)
// where c marks the grazing pixels of the order-independent pass (k_vnmap,
// which has written vtx/nrm already). A grazing pixel that is not
// invalidated keeps its vertex (its normal was never computed), and a pixel
// whose left or upper neighbour is invalidated loses its normal. One wave
// walks the rows: it reads a row's flags as 64-bit ballots, resolves the
// recurrence along the row word by word (runs of c alternate, starting
// invalidated unless the pixel before was), and applies the fix-ups.
__global__ __launch_bounds__(64) void k_vnmap_inplace(const SessParams *__restrict__ SP, int l, float *vtx,
                                                      float *nrm) {
    const SessParams &P = SP[0];
    const LevelDesc &L = P.lev[l];
    const int W = L.W, H = L.H, lane = threadIdx.x;
    const int nw = (W + 63) >> 6;  // <= 128 words (frames are at most 8192 wide)
    __shared__ unsigned long long s_up[128], s_c[128], s_n[128], s_inv[128];
    for (int k = lane; k < nw; k += 64) s_up[k] = 0ull;
    const float *D = P.filt[l];
    const float *K = L.K;
    __syncthreads();
    for (int i = 1; i < H - 1; ++i) {
        for (int k = 0; k < nw; ++k) {
            const int j = (k << 6) + lane;
            const bool in = j >= 1 && j < W - 1;
            const size_t t = (size_t)i * W + j;
            const float vx = in ? vtx[3 * t] : 0.0f;
            const float nx = in ? nrm[3 * t] : -100.0f;
            const float d = in ? ldg(D, t) : 0.0f;
            const unsigned long long cm = __ballot(in && vx == -100.0f && d > 0.0f);
            const unsigned long long nm = __ballot(in && nx != -100.0f);
            if (lane == 0) {
                s_c[k] = cm;
                s_n[k] = nm;
            }
        }
        __syncthreads();
        if (lane == 0) {
            bool carry = false;  // inv of the pixel before the word
            for (int k = 0; k < nw; ++k) {
                unsigned long long rest = s_c[k] & ~s_up[k], inv = 0ull;
                while (rest) {
                    const int s0 = __builtin_ctzll(rest);
                    const unsigned long long from = rest >> s0;
                    const int len = (~from == 0ull) ? 64 - s0 : __builtin_ctzll(~from);
                    const unsigned long long run = (len >= 64 ? ~0ull : ((1ull << len) - 1ull)) << s0;
                    const int first = s0 + ((s0 == 0 && carry) ? 1 : 0);
                    inv |= run & ((first & 1) ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull);
                    rest &= ~run;
                }
                s_inv[k] = inv;
                carry = (inv >> 63) & 1ull;
            }
        }
        __syncthreads();
        for (int k = 0; k < nw; ++k) {
            const int j = (k << 6) + lane;
            if (!(j >= 1 && j < W - 1)) continue;
            const unsigned long long inv = s_inv[k];
            const bool c = (s_c[k] >> lane) & 1ull, nv = (s_n[k] >> lane) & 1ull;
            const bool here = (inv >> lane) & 1ull;
            const bool left = lane ? ((inv >> (lane - 1)) & 1ull) : (k ? ((s_inv[k - 1] >> 63) & 1ull) : false);
            const bool up = (s_up[k] >> lane) & 1ull;
            const size_t t = (size_t)i * W + j;
            if (c && !here) {
                const float d = ldg(D, t);
                vtx[3 * t + 0] = d * ((float)j - K[2]) / K[0];
                vtx[3 * t + 1] = d * ((float)i - K[3]) / K[1];
                vtx[3 * t + 2] = d;
            } else if (nv && (left || up)) {
                nrm[3 * t + 0] = -100.0f;
                nrm[3 * t + 1] = -100.0f;
                nrm[3 * t + 2] = -100.0f;
            }
        }
        __syncthreads();
        for (int k = lane; k < nw; k += 64) s_up[k] = s_inv[k];
        __syncthreads();
    }
}

// Copies the caller's device-resident frames into the sessions' raw-frame
// buffers (in each session's depth units, copied as 16-bit halves): one
// launch for a whole batch (blockIdx.y = session of the chunk), 16-byte
// vector loads/stores where the frame size allows.
__global__ __launch_bounds__(256) void k_gather_depth(const SessParams *__restrict__ SP, DepthChunk ch) {
    const int s = blockIdx.y;
    if (s >= ch.n) return;
    const SessParams &P = SP[ch.slot[s]];
    const size_t npx = (size_t)P.lev[0].W * P.lev[0].H * (P.raw_f32 ? 2 : 1);
    const uint16_t *src = ch.src[s];
    uint16_t *dst = (uint16_t *)const_cast<void *>(P.raw);
    const bool vec = ((npx & 7) == 0) && ((((uintptr_t)src) & 15) == 0);
    if (vec) {
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        uint4 *d4 = reinterpret_cast<uint4 *>(dst);
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx / 8; i += (size_t)gridDim.x * blockDim.x)
            d4[i] = s4[i];
    } else {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += (size_t)gridDim.x * blockDim.x)
            dst[i] = src[i];
    }
}

// n 32-bit words at p set to v (session set-up: map sentinels).
__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        stg(p, i, v);
}
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_gather_depth(const SessParams *P, const DepthChunk &ch, hipStream_t s) {
    if (ch.n > 0) hipLaunchKernelGGL(k_gather_depth, dim3(64, ch.n), dim3(256), 0, s, P, ch);
}

void launch_bilateral(const SessParams *P, const WorkItem *work, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_bilateral, dim3(n), dim3(kTileThreads), 0, s, P, work);
}
void launch_pyr_down(const SessParams *P, const WorkItem *work, int n, int l, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_pyr_down, dim3(n), dim3(kTileThreads), 0, s, P, work, l);
}
void launch_vnmap(const SessParams *P, const WorkItem *work, int n, int l, float *vtx, float *nrm, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_vnmap, dim3(n), dim3(kTileThreads), 0, s, P, work, l, vtx, nrm);
}
void launch_vnmap_inplace(const SessParams *P, int l, float *vtx, float *nrm, hipStream_t s) {
    hipLaunchKernelGGL(k_vnmap_inplace, dim3(1), dim3(64), 0, s, P, l, vtx, nrm);
}

}  // namespace vs
