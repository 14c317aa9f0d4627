// vs_api.hip — host side of the C ABI declared in include/vscan.h: contexts,
// sessions (one HBM allocation each), batches and the per-frame launch
// sequence, optionally replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "vs_device.hpp"
#include "vs_launch.hpp"
#include "vscan.h"

namespace vs {
void launch_track_setup_forced(const SessParams *P, int nsess, const double *T_init, hipStream_t s);
void launch_pose_derived(const SessParams *P, hipStream_t s);
void launch_raycast_stats(const SessParams *P, const WorkItem *work, int n, int level, int tt, int wt,
                          unsigned long long *stats, hipStream_t s);
void launch_fuse_ok(const SessParams *P, int nsess, hipStream_t s);
}  // namespace vs

using namespace vs;

// Scratch of vs_icp_system_noalloc: the per-tile partial sums of one ICP
// system (kIcpVals doubles per 32x8 tile), for levels of up to
// kNoAllocMaxTiles tiles (2,097,152 pixels). A file-scope device array: each
// GPU holds its own copy until the process ends, and every call on that GPU
// shares it. Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (vs_ctx::scratch,
// taken in vs_ctx_create); the array's name is not a device address on the
// host.
constexpr int kNoAllocMaxTiles = 8192;
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxTiles * kIcpVals];

namespace {

thread_local std::string g_err;

vs_status fail(vs_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(VS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

struct Layout {
    size_t state, raw, filt[kMaxLevels], mvn[kMaxLevels], mimg[kMaxLevels];
    size_t vox, bvis, vmask, bocc, bband, sband, sdist, sdtmp, icp_part, diag_part, diag_mask, lists, params, stats, keys;
    size_t vcol, rgb, mtex[kMaxLevels];  // colour sessions only (0 otherwise)
    size_t wtab;                          // voxel-centre world coordinates, x then y then z
    size_t total;
    size_t nvox;
    uint32_t nbricks, nwords, nsuper, nswords;
    int sdim[3];
    int tiles[kMaxLevels];
};

int tiles_of(int W, int H) { return ((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH); }

// Bytes of a batch's list region: per level the (slot, tile) items of every
// tile and the (slot, group) items of every ICP tile group, then the brick
// list and its counter.
size_t lists_bytes(const size_t ntile[kMaxLevels], const size_t ngroup[kMaxLevels], size_t bricks) {
    size_t off = 0;
    for (int l = 0; l < kMaxLevels; ++l) {
        off = align_up(off + std::max<size_t>(ntile[l], 1) * sizeof(WorkItem));
        off = align_up(off + std::max<size_t>(ngroup[l], 1) * sizeof(WorkItem));
    }
    off = align_up(off + std::max<size_t>(bricks, 1) * sizeof(int2));
    return align_up(off + 16);
}

Layout make_layout(const vs_session_desc &d) {
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes);
        return o;
    };
    const int W = d.width, H = d.height;
    L.nvox = (size_t)d.vol_dim[0] * d.vol_dim[1] * d.vol_dim[2];
    L.nbricks = (uint32_t)((d.vol_dim[0] / 8) * (d.vol_dim[1] / 8) * (d.vol_dim[2] / 8));
    L.nwords = (L.nbricks + 31) / 32;
    for (int k = 0; k < 3; ++k) L.sdim[k] = (d.vol_dim[k] / kBrick + (1 << kSuperShift) - 1) >> kSuperShift;
    L.nsuper = (uint32_t)(L.sdim[0] * L.sdim[1] * L.sdim[2]);
    L.nswords = (L.nsuper + 31) / 32;
    const size_t sv = d.tsdf_type == VS_TSDF_F16 ? 2 : 4, sw = d.weight_type == VS_WEIGHT_U8 ? 1 : 4;
    L.state = take(sizeof(SessState));
    L.params = take(sizeof(SessParams));
    L.raw = take((size_t)W * H * (d.depth_units == VS_DEPTH_M_F32 ? 4 : 2));
    for (int l = 0; l < kMaxLevels; ++l) {
        const int Wl = W >> l, Hl = H >> l;
        const size_t n = (size_t)Wl * Hl;
        L.tiles[l] = l < d.levels ? tiles_of(Wl, Hl) : 0;
        L.filt[l] = take(n * 4);
        L.mvn[l] = take((n + 1) * 2 * sizeof(MapXYZ));  // vertex + normal per pixel, one padding pixel
        L.mimg[l] = take(n);
    }
    // value and weight of a voxel in one record (4 bytes for FP16 + u8, 8
    // otherwise)
    L.vox = take(L.nvox * vox_rec_bytes(sv == 2, sw == 1));
    L.bvis = take((size_t)L.nwords * 4);
    L.vmask = take(((size_t)L.nbricks + 127) / 128 * 128);
    L.bocc = take((size_t)L.nwords * 4);
    L.bband = take((size_t)L.nwords * 4);
    L.sband = take((size_t)L.nswords * 4);
    L.sdist = take(((size_t)L.nsuper + 7) / 8 * 4);  // 4 bits per superbrick, whole words
    L.sdtmp = take(L.nsuper);
    L.icp_part = take((size_t)L.tiles[0] * kIcpVals * 8);
    L.diag_part = take((size_t)L.tiles[0] * kIcpVals * 8);
    L.diag_mask = take((size_t)W * H);
    {
        size_t nt[kMaxLevels], ng[kMaxLevels];
        for (int l = 0; l < kMaxLevels; ++l) {
            nt[l] = (size_t)L.tiles[l];
            ng[l] = (size_t)icp_groups(L.tiles[l], l);
        }
        L.lists = take(lists_bytes(nt, ng, L.nbricks));
    }
    L.stats = take(64);
    L.keys = take((size_t)W * H * 4);  // depth reprojection z-buffer
    L.wtab = take((size_t)(d.vol_dim[0] + d.vol_dim[1] + d.vol_dim[2]) * 4);
    if (d.integrate_mode == VS_INTEGRATE_COLOR) {
        L.vcol = take(L.nvox * 4);
        L.rgb = take((size_t)W * H * 3);
        for (int l = 0; l < kMaxLevels; ++l) L.mtex[l] = take((size_t)(W >> l) * (H >> l) * 3);
    }
    L.total = off;
    return L;
}

vs_status check_desc(const vs_session_desc *d) {
    if (!d) return fail(VS_EINVAL, "null descriptor");
    if (d->width < 16 || d->height < 16 || d->width > 8192 || d->height > 8192)
        return fail(VS_EINVAL, "frame size %dx%d out of range", d->width, d->height);
    if (d->levels < 1 || d->levels > kMaxLevels) return fail(VS_EINVAL, "levels must be 1..3");
    for (int l = 0; l < d->levels; ++l)
        if (d->iters[l] < 0 || d->iters[l] > 100) return fail(VS_EINVAL, "iters out of range");
    for (int k = 0; k < 3; ++k)
        if (d->vol_dim[k] < 8 || d->vol_dim[k] % 8 || d->vol_dim[k] > 2048)
            return fail(VS_EINVAL, "vol_dim[%d]=%d must be a multiple of 8 in [8, 2048]", k, d->vol_dim[k]);
    if (!(d->voxel > 0) || !(d->mu > 0) || !(d->max_weight >= 1)) return fail(VS_EINVAL, "voxel/mu/max_weight");
    if (d->weight_type == VS_WEIGHT_U8 && d->max_weight > 255) return fail(VS_EINVAL, "u8 weight cap > 255");
    if (d->tsdf_type < 0 || d->tsdf_type > 1 || d->weight_type < 0 || d->weight_type > 1)
        return fail(VS_EINVAL, "storage type");
    if (d->integrate_mode != VS_INTEGRATE_DEPTH && d->integrate_mode != VS_INTEGRATE_COLOR)
        return fail(VS_EINVAL, "integrate_mode %d", d->integrate_mode);
    if (d->bil_radius < 1 || d->bil_radius > 3 || !(d->bil_sigma_s > 0) || !(d->bil_sigma_r > 0))
        return fail(VS_EINVAL, "bilateral parameters");
    if (!(d->fx > 0) || !(d->fy > 0)) return fail(VS_EINVAL, "focal length");
    if (d->depth_units != VS_DEPTH_MM_U16 && d->depth_units != VS_DEPTH_M_F32)
        return fail(VS_EINVAL, "depth_units %d", d->depth_units);
    if (!(d->icp_rel_pivot >= 0.0f && d->icp_rel_pivot < 1.0f))
        return fail(VS_EINVAL, "icp_rel_pivot must be in [0, 1)");
    return VS_OK;
}

// Bytes of one depth frame of the session's size in its units.
size_t frame_bytes(const vs_session_desc &d) {
    return (size_t)d.width * d.height * (d.depth_units == VS_DEPTH_M_F32 ? 4 : 2);
}

}  // namespace

// ---------------------------------------------------------------------------
// context, batches
// ---------------------------------------------------------------------------
// Events around the stages of one frame step:
//   [0,1) track setup + bilateral   [1,2) pyramid    [2,3) ICP + finalize
//   [3,4) brick clear + mark        [4,5) integrate  [5,6) raycast march, all levels
//   [6,7) raycast shading, all levels                [7,8) empty
constexpr int kMarks = 9;
struct Timer {
    hipEvent_t ev[kMarks] = {};
    bool used = false;
    hipError_t init() {
        for (auto &e : ev)
            if (!e) {
                hipError_t r = hipEventCreate(&e);
                if (r != hipSuccess) return r;
            }
        return hipSuccess;
    }
    void destroy() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------------------
// launch shapes and their autotuner
// ---------------------------------------------------------------------------
// The shape of three launches of the step: the ray block of a wave in the
// raycast march and in its shading (RayBlock), and the workgroup count of the
// persistent integrate. Every shape gives the same results; which is fastest
// depends on the workload, so a context can pick them per size class of
// batch (frame size, volume size, storage types, session count).
// The shading's candidates: ray block and the waves per SIMD its register
// allocation aims for (4: the compiler's own choice, 80 VGPRs with 32-bit
// stencil offsets or ~126 with 64-bit ones, every stencil load in flight; 5:
// at most 102).
struct ShadeOpt {
    int block, waves;
    const char *name;
};
constexpr ShadeOpt kShadeOpts[] = {
    {kRay8x8, 4, "8x8"}, {kRay32x2, 4, "32x2"}, {kRay8x8, 5, "8x8w5"}, {kRay32x2, 5, "32x2w5"}};
struct Shape {
    int march = kRay8x8, shade = 0;  // shade: index into kShadeOpts
    int integ = -1;  // index into kIntegGrids; -1: the batch's default grid
};
constexpr int kMarchOpts[] = {kRay8x8, kRay16x4, kRay32x2};
constexpr int kIntegGrids[] = {1024, 2048, 4096, 8192, 16384, 32768};
constexpr int kNumMarch = 3, kNumShade = 4, kNumInteg = 6;  // kTuneSteps is a multiple of each
const char *const kRayBlockName[] = {"8x8", "16x4", "32x2"};
// Online tuning: the first kTuneSteps steps of a batch whose size class has
// no choice yet run eagerly and timed, step i with march candidate i % 3,
// shade candidate i % 4 and integrate grid i % 6 (each candidate of each
// stage is timed 2 to 4 times, on consecutive frames); then each stage keeps
// the candidate with the lowest mean time (for the ray blocks: unless the
// first candidate is within kTuneMargin of it) and the step is captured with
// it.
constexpr int kTuneSteps = 12;
constexpr double kTuneMargin = 0.03;  // see tune_step
static_assert(kTuneSteps % kNumMarch == 0 && kTuneSteps % kNumShade == 0 && kTuneSteps % kNumInteg == 0,
              "every candidate gets the same number of tuning steps");
struct TuneState {
    int step = 0;           // tuning steps enqueued
    static constexpr int kCand = std::max(kNumMarch, std::max(kNumShade, kNumInteg));
    double sum[3][kCand] = {};  // [march, shade, integrate][candidate] ms
    int cnt[3][kCand] = {};
    int ci[3] = {0, 0, 0};  // candidates of the step in flight
    bool pending = false;   // that step's events are not read yet
    Timer timer;
    ~TuneState() { timer.destroy(); }
};

// Frames that a batched step receives in host memory are copied by the CPU
// into one page-locked area and uploaded by a single DMA transfer on a copy
// stream of the library's own, then gathered into the sessions' raw-frame
// buffers on the step's stream. Two areas (host and device) are used in
// turn: the upload that one call enqueues waits only for the gather that
// read the same device area two calls before, so it runs while the GPU
// still fuses the previous step. (One transfer per frame costs the step
// about 1 ms more at 64 frames of 614 KB, even from page-locked memory.)
struct HostStage {
    hipStream_t cs = nullptr;
    char *pinned[2] = {nullptr, nullptr};  // host areas
    char *area[2] = {nullptr, nullptr};    // device areas
    size_t bytes = 0;  // capacity of each area
    hipEvent_t copied[2] = {nullptr, nullptr}, read[2] = {nullptr, nullptr};
    bool used[2] = {false, false};  // copied[g] and read[g] recorded since the areas were (re)allocated
    int next = 0;
};

struct Batch {
    int n = 0, tt = 0, wt = 0, levels = 0;
    int max_iters[kMaxLevels] = {0, 0, 0};
    int sd_slices = 1;         // most superbricks along z of a session (distance field passes)
    SessParams *dP = nullptr;  // device array
    WorkItem *dWork[kMaxLevels] = {nullptr, nullptr, nullptr};   // (slot, tile)
    int nWork[kMaxLevels] = {0, 0, 0};
    WorkItem *dGroup[kMaxLevels] = {nullptr, nullptr, nullptr};  // (slot, ICP tile group)
    int nGroup[kMaxLevels] = {0, 0, 0};
    BrickList bl{};
    uint32_t bricks = 0;    // capacity of the brick list
    uint32_t max_words = 0; // longest visible-brick bitmap of a member (32-bit words)
    bool o32 = false;       // every member's voxel records are below 4 GB
    int integ_grid = 1;     // default integrate grid
    Shape shape;
    std::string size_class;
    bool tuned = false;     // shape settled (from the tuner or its cache)
    std::unique_ptr<TuneState> tune;
    void *owned = nullptr;  // allocation owned by a context batch (null for a session's own)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<vs_session *> members;
    const int *active = nullptr;  // slot table: device flags of the slots with a frame this step
    // Parts of a large batch, each with its own lists: a captured step runs
    // them as independent branches of the graph, so one part's launch tails
    // (the last workgroups of a launch, a long ray) overlap the other parts'
    // work. Per-session results do not depend on the grouping.
    std::vector<std::unique_ptr<Batch>> sub;
    HostStage hs;  // host-memory frames of this batch's steps
};

struct vs_ctx {
    int device = 0;
    double *scratch = nullptr;  // this device's copy of `scratch` (vs_icp_system_noalloc)
    std::map<std::vector<vs_session *>, std::unique_ptr<Batch>> batches;
    std::vector<vs_session *> sessions;
    std::vector<vs_server *> servers;
    bool use_graph = true;
    bool timing = false;
    Timer timer;
    int tune_mode = VS_TUNE_OFF;
    std::string tune_path;                 // cache file ("" = none)
    std::map<std::string, Shape> tuned;    // size class -> shape
    std::map<std::string, std::string> tune_ms;  // size class -> candidate means
};

struct vs_session {
    vs_ctx *ctx = nullptr;
    vs_session_desc d{};
    Layout L{};
    char *mem = nullptr;
    SessParams hp{};
    SessState *st = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t last = nullptr;
    Batch self;
    bool timing = false;
    Timer timer;
    vs_server *server = nullptr;  // slot table holding this session, if any
    int slot = -1;
    bool pooled = false;  // mem from the stream-ordered allocator (slot-table sessions)
};

// Slot table of the serving form. B spans all max_slots slots: B.dP[i] holds
// the parameters of the session in slot i, or for a free slot a zero
// parameter block (no levels, no tiles) whose state pointer is that slot's
// own placeholder state, so every kernel of a step sees an inactive session
// and returns before touching any buffer. Work lists and the brick list cover
// the open slots only and are rebuilt when a slot opens or closes.
struct vs_server {
    vs_ctx *ctx = nullptr;
    vs_server_desc d{};
    std::vector<vs_session *> slot;
    hipStream_t stream = nullptr;
    hipStream_t last = nullptr;
    hipEvent_t order = nullptr;
    Batch B;
    char *fixed = nullptr;   // params[max_slots], placeholder states, active flags
    SessState *ghost = nullptr;
    int *active = nullptr;
    // Work lists, brick list and brick counter of the current slot layout, in
    // one of two device buffers: a new layout is written into the buffer the
    // steps in flight do not read. Each has a pinned host copy the upload is
    // made from, and an event after that upload.
    char *lists[2] = {nullptr, nullptr};
    size_t lists_bytes[2] = {0, 0};
    char *lstage[2] = {nullptr, nullptr};
    size_t lstage_bytes[2] = {0, 0};
    hipEvent_t lstage_ev[2] = {nullptr, nullptr};
    int gen = 0;  // buffer of the current layout
    // Slot parameter uploads: a ring of pinned blocks, each with an event
    // after its upload.
    SessParams *pstage = nullptr;
    hipEvent_t *pstage_ev = nullptr;
    int pstage_n = 0, pstage_next = 0;
    // Step graphs of earlier layouts, destroyed once the steps that may still
    // run them have completed (the event after the last of them).
    struct Retired {
        hipGraph_t graph;
        hipGraphExec_t gexec;
        hipEvent_t done;
    };
    std::vector<Retired> retired;
    bool dirty = true;
    unsigned long long bytes = 0;
};

namespace {

// Host copy of a batch's work lists (see lists_bytes for the device layout).
struct ListSet {
    std::vector<WorkItem> tile[kMaxLevels], group[kMaxLevels];
    uint32_t bricks = 0, max_words = 0;
    size_t max_array = 0;  // largest voxel-record array of a member (bytes)
    int levels = 0, max_iters[kMaxLevels] = {0, 0, 0};
    int count = 0, max_w = 0, max_h = 0, max_dim = 0;
    void add(int slot, const vs_session *s) {
        ++count;
        max_words = std::max(max_words, s->L.nwords);
        max_array = std::max(max_array, s->L.nvox * vox_rec_bytes(s->d.tsdf_type == VS_TSDF_F16,
                                                                  s->d.weight_type == VS_WEIGHT_U8));
        max_w = std::max(max_w, s->d.width);
        max_h = std::max(max_h, s->d.height);
        for (int k = 0; k < 3; ++k) max_dim = std::max(max_dim, s->d.vol_dim[k]);
        levels = std::max(levels, s->d.levels);
        for (int l = 0; l < kMaxLevels; ++l) {
            max_iters[l] = std::max(max_iters[l], s->hp.iters[l]);
            for (int t = 0; t < s->L.tiles[l]; ++t) tile[l].push_back({slot, t});
            for (int g = 0; g < icp_groups(s->L.tiles[l], l); ++g) group[l].push_back({slot, g});
        }
        bricks += s->L.nbricks;
    }
    size_t bytes() const {
        size_t nt[kMaxLevels], ng[kMaxLevels];
        for (int l = 0; l < kMaxLevels; ++l) {
            nt[l] = tile[l].size();
            ng[l] = group[l].size();
        }
        return lists_bytes(nt, ng, bricks);
    }
    // Points the batch at `mem` (bytes() long): work lists, then the brick
    // list, then the brick counter; returns the offset of the brick list (the
    // work lists are the bytes before it).
    size_t point(Batch &B, char *mem) const {
        size_t off = 0;
        for (int l = 0; l < kMaxLevels; ++l) {
            B.dWork[l] = (WorkItem *)(mem + off);
            B.nWork[l] = (int)tile[l].size();
            off = align_up(off + std::max<size_t>(tile[l].size(), 1) * sizeof(WorkItem));
            B.dGroup[l] = (WorkItem *)(mem + off);
            B.nGroup[l] = (int)group[l].size();
            off = align_up(off + std::max<size_t>(group[l].size(), 1) * sizeof(WorkItem));
        }
        const size_t head = off;
        B.bl.items = (int2 *)(mem + off);
        off = align_up(off + std::max<size_t>(bricks, 1) * sizeof(int2));
        B.bl.count = (unsigned *)(mem + off);
        B.bl.capacity = bricks;
        return head;
    }
    // Bytes of the work lists (the part of the layout a host copy uploads)
    // plus 16 for the zeroed brick counter.
    size_t head_bytes() const {
        size_t off = 0;
        for (int l = 0; l < kMaxLevels; ++l) {
            off = align_up(off + std::max<size_t>(tile[l].size(), 1) * sizeof(WorkItem));
            off = align_up(off + std::max<size_t>(group[l].size(), 1) * sizeof(WorkItem));
        }
        return off + 16;
    }
    // Points the batch at `mem` (bytes() long), uploads the lists and clears
    // the brick counter; sets the iteration bounds and the integrate grid.
    vs_status place(Batch &B, char *mem, int integ_max) const {
        point(B, mem);
        for (int l = 0; l < kMaxLevels; ++l) {
            if (!tile[l].empty())
                HIP_TRY(hipMemcpy(B.dWork[l], tile[l].data(), tile[l].size() * sizeof(WorkItem), hipMemcpyHostToDevice));
            if (!group[l].empty())
                HIP_TRY(hipMemcpy(B.dGroup[l], group[l].data(), group[l].size() * sizeof(WorkItem),
                                  hipMemcpyHostToDevice));
        }
        HIP_TRY(hipMemset(B.bl.count, 0, 16));
        return settle(B, integ_max);
    }
    // The same in stream order on h, from `stage` (pinned, head_bytes() long,
    // not to be rewritten before the uploads have run).
    vs_status place_async(Batch &B, char *mem, char *stage, hipStream_t h, int integ_max) const {
        const size_t head = point(B, mem);
        for (int l = 0; l < kMaxLevels; ++l) {
            if (!tile[l].empty())
                std::memcpy(stage + ((char *)B.dWork[l] - mem), tile[l].data(), tile[l].size() * sizeof(WorkItem));
            if (!group[l].empty())
                std::memcpy(stage + ((char *)B.dGroup[l] - mem), group[l].data(), group[l].size() * sizeof(WorkItem));
        }
        std::memset(stage + head, 0, 16);
        HIP_TRY(hipMemcpyAsync(mem, stage, head, hipMemcpyHostToDevice, h));
        HIP_TRY(hipMemcpyAsync(B.bl.count, stage + head, 16, hipMemcpyHostToDevice, h));
        return settle(B, integ_max);
    }
    // Iteration bounds, integrate grid and size class of the batch.
    vs_status settle(Batch &B, int integ_max) const {
        B.levels = levels;
        B.sd_slices = (max_dim / kBrick + (1 << kSuperShift) - 1) >> kSuperShift;
        for (int l = 0; l < kMaxLevels; ++l) B.max_iters[l] = max_iters[l];
        B.bricks = bricks;
        B.max_words = max_words;
        B.o32 = max_array < (size_t(1) << 32);
        B.integ_grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(bricks, (uint32_t)integ_max));
        // size class: largest frame, largest volume edge and session count
        // rounded up to powers of two, storage types
        auto pow2 = [](int v) {
            int p = 1;
            while (p < v) p <<= 1;
            return p;
        };
        char buf[96];
        snprintf(buf, sizeof(buf), "%dx%d v%d %s+%s n%d", max_w, max_h, pow2(max_dim), B.tt ? "f16" : "f32",
                 B.wt ? "u8" : "f32", pow2(count));
        if (B.size_class != buf) {
            B.size_class = buf;
            B.tuned = false;
            B.tune.reset();
        }
        return VS_OK;
    }
};

// Workgroups of the integrate launch for the batch's shape.
int integ_grid(const Batch &B) {
    if (B.shape.integ < 0) return B.integ_grid;
    return (int)std::max<uint32_t>(1, std::min<uint32_t>(B.bricks, (uint32_t)kIntegGrids[B.shape.integ]));
}

void free_batch_graph(Batch &b) {
    if (b.gexec) hipGraphExecDestroy(b.gexec);
    if (b.graph) hipGraphDestroy(b.graph);
    b.gexec = nullptr;
    b.graph = nullptr;
}

void release_stage(HostStage &S) {
    if (S.cs) hipStreamSynchronize(S.cs);
    for (int g = 0; g < 2; ++g) {
        if (S.read[g]) hipEventSynchronize(S.read[g]);
        if (S.area[g]) hipFree(S.area[g]);
        if (S.pinned[g]) hipHostFree(S.pinned[g]);
        if (S.copied[g]) hipEventDestroy(S.copied[g]);
        if (S.read[g]) hipEventDestroy(S.read[g]);
    }
    if (S.cs) hipStreamDestroy(S.cs);
    S = HostStage();
}

// Graph, allocation and halves of a context batch.
void release_batch(Batch &b) {
    release_stage(b.hs);
    free_batch_graph(b);
    if (b.owned) hipFree(b.owned);
    b.owned = nullptr;
    for (auto &h : b.sub) release_batch(*h);
    b.sub.clear();
}

void drop_batches_with(vs_ctx *ctx, vs_session *s) {
    for (auto it = ctx->batches.begin(); it != ctx->batches.end();) {
        if (std::find(it->first.begin(), it->first.end(), s) != it->first.end()) {
            release_batch(*it->second);
            it = ctx->batches.erase(it);
        } else {
            ++it;
        }
    }
}

// Waits for the server's steps in flight (used only where the caller needs
// their results: eviction, destruction).
vs_status server_quiesce(vs_server *v) {
    if (v->last && v->last != v->stream) HIP_TRY(hipStreamSynchronize(v->last));
    HIP_TRY(hipStreamSynchronize(v->stream));
    return VS_OK;
}

// The stream whose end follows every step enqueued so far (steps on other
// streams are chained to their predecessor by an event). Changes to the slot
// table are enqueued there: they take effect after the steps in flight and
// before the next one, and the host does not wait.
hipStream_t server_order(const vs_server *v) { return v->last ? v->last : v->stream; }

// Destroys the step graphs whose last launch has completed (all of them with
// wait = true, after a quiesce).
void server_reap(vs_server *v, bool wait) {
    auto &r = v->retired;
    for (size_t k = 0; k < r.size();) {
        if (wait || hipEventQuery(r[k].done) == hipSuccess) {
            if (r[k].gexec) hipGraphExecDestroy(r[k].gexec);
            if (r[k].graph) hipGraphDestroy(r[k].graph);
            hipEventDestroy(r[k].done);
            r[k] = r.back();
            r.pop_back();
        } else {
            ++k;
        }
    }
}

// Takes the current step graph out of use; it is destroyed after the steps
// enqueued so far (which may launch it) have completed.
vs_status server_retire_graph(vs_server *v) {
    Batch &B = v->B;
    if (!B.gexec && !B.graph) return VS_OK;
    vs_server::Retired r{B.graph, B.gexec, nullptr};
    B.graph = nullptr;
    B.gexec = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    v->retired.push_back(r);
    HIP_TRY(hipEventRecord(r.done, server_order(v)));
    return VS_OK;
}

// Parameter block of slot i: the open session's, or the inactive placeholder.
// Uploaded in stream order after the steps in flight (which may still read
// the slot's previous block), from a pinned block of the ring.
vs_status server_write_slot(vs_server *v, int i) {
    SessParams p{};
    if (v->slot[i]) {
        p = v->slot[i]->hp;
    } else {
        p.st = v->ghost + i;
    }
    const int k = v->pstage_next;
    v->pstage_next = (k + 1) % v->pstage_n;
    HIP_TRY(hipEventSynchronize(v->pstage_ev[k]));  // that block's earlier upload has run
    v->pstage[k] = p;
    const hipStream_t h = server_order(v);
    HIP_TRY(hipMemcpyAsync(v->B.dP + i, v->pstage + k, sizeof(p), hipMemcpyHostToDevice, h));
    HIP_TRY(hipEventRecord(v->pstage_ev[k], h));
    return VS_OK;
}

// Removes a session from its slot (the session itself stays alive); the
// slot's placeholder block takes effect after the steps in flight.
vs_status server_forget(vs_session *s) {
    vs_server *v = s->server;
    if (!v) return VS_OK;
    const int i = s->slot;
    v->slot[i] = nullptr;
    v->bytes -= s->L.total;
    s->server = nullptr;
    s->slot = -1;
    v->dirty = true;
    return server_write_slot(v, i);
}

// Work lists (slot, tile) of the open slots, the brick list sized for their
// bricks, and the iteration bounds of the launch sequence, written into the
// list buffer the steps in flight do not read, in stream order after them;
// retires the captured graph so the next step captures the new sequence. The
// host waits for nothing but, in the rare case of two layout changes within
// the steps in flight, the upload from the same pinned copy two layouts ago.
vs_status server_rebuild(vs_server *v) {
    Batch &B = v->B;
    vs_status r = server_retire_graph(v);
    if (r != VS_OK) return r;
    ListSet ls;
    for (int i = 0; i < (int)v->slot.size(); ++i)
        if (v->slot[i]) ls.add(i, v->slot[i]);
    const size_t need = ls.bytes();
    const int g = v->gen ^ 1;
    const hipStream_t h = server_order(v);
    if (need > v->lists_bytes[g]) {
        if (v->lists[g]) HIP_TRY(hipFreeAsync(v->lists[g], h));
        v->lists[g] = nullptr;
        v->lists_bytes[g] = 0;
        const size_t want = need + need / 2;
        hipError_t e = hipMallocAsync((void **)&v->lists[g], want, h);
        if (e != hipSuccess) return fail(VS_ENOMEM, "server lists (%zu bytes): %s", want, hipGetErrorString(e));
        v->lists_bytes[g] = want;
    }
    HIP_TRY(hipEventSynchronize(v->lstage_ev[g]));
    const size_t head = ls.head_bytes();
    if (head > v->lstage_bytes[g]) {
        if (v->lstage[g]) HIP_TRY(hipHostFree(v->lstage[g]));
        v->lstage[g] = nullptr;
        v->lstage_bytes[g] = 0;
        const size_t want = head + head / 2;
        hipError_t e = hipHostMalloc((void **)&v->lstage[g], want, hipHostMallocDefault);
        if (e != hipSuccess) return fail(VS_ENOMEM, "server list staging (%zu bytes): %s", want, hipGetErrorString(e));
        v->lstage_bytes[g] = want;
    }
    r = ls.place_async(B, v->lists[g], v->lstage[g], h, 8192);
    if (r != VS_OK) return r;
    HIP_TRY(hipEventRecord(v->lstage_ev[g], h));
    v->gen = g;
    v->dirty = false;
    return VS_OK;
}

// Slot and budget check for a session of descriptor d; *slot = a free slot.
vs_status server_admit(vs_server *v, const vs_session_desc *d, int *slot) {
    vs_status cs = check_desc(d);
    if (cs != VS_OK) return cs;
    if (d->tsdf_type != v->d.tsdf_type || d->weight_type != v->d.weight_type)
        return fail(VS_EINVAL, "session storage types differ from the server's");
    const unsigned long long need = make_layout(*d).total;
    if (v->d.hbm_budget && v->bytes + need > v->d.hbm_budget)
        return fail(VS_BUDGET, "session needs %llu bytes; %llu of %llu in use", need, v->bytes, v->d.hbm_budget);
    int i = 0;
    while (i < (int)v->slot.size() && v->slot[i]) ++i;
    if (i == (int)v->slot.size()) return fail(VS_BUDGET, "all %d slots open", (int)v->slot.size());
    *slot = i;
    return VS_OK;
}

// Puts a new session (initialised: its own stream is idle) into free slot i
// (destroys it on failure). Its parameter block takes effect after the steps
// in flight.
vs_status server_attach(vs_server *v, vs_session *s, int i) {
    v->slot[i] = s;
    s->server = v;
    s->slot = i;
    v->bytes += s->L.total;
    v->dirty = true;
    vs_status r = server_write_slot(v, i);
    if (r != VS_OK) vs_session_destroy(s);
    return r;
}

hipStream_t pick(vs_session *s, vs_stream st) {
    hipStream_t h = st ? (hipStream_t)st : s->stream;
    s->last = h;
    return h;
}

// The per-frame launch sequence of a batch (so_Scan_Object order).
void enqueue_step(const Batch &B, hipStream_t s, Timer *timed) {
    auto mark = [&](int i) {
        if (timed) (void)hipEventRecord(timed->ev[i], s);
    };
    mark(0);
    if (B.active)
        launch_track_setup_active(B.dP, B.n, B.active, s);
    else
        launch_track_setup(B.dP, B.n, nullptr, s);
    launch_bilateral(B.dP, B.dWork[0], B.nWork[0], s);
    mark(1);
    for (int l = 1; l < B.levels; ++l) launch_pyr_down(B.dP, B.dWork[l], B.nWork[l], l, s);
    mark(2);
    for (int l = B.levels - 1; l >= 0; --l)
        for (int it = 0; it < B.max_iters[l]; ++it) launch_icp_iter(B.dP, B.dGroup[l], B.nGroup[l], B.n, l, it, s);
    launch_track_finalize(B.dP, B.n, s);
    mark(3);
    launch_vis_clear(B.dP, B.n, B.bl, s);
    launch_vis_mark(B.dP, B.dWork[0], B.nWork[0], B.n, B.max_words, B.bl, s);
    mark(4);
    launch_integrate(B.dP, B.n, B.bl, B.tt, B.wt, integ_grid(B), B.sd_slices, s);
    mark(5);
    const LevelWork lw = level_work(B.dWork, B.nWork, B.levels);
    launch_raycast_march(B.dP, lw, B.tt, B.wt, nullptr, B.shape.march, B.o32, s);
    mark(6);
    launch_raycast_shade(B.dP, lw, B.tt, B.wt, nullptr, kShadeOpts[B.shape.shade].block,
                         kShadeOpts[B.shape.shade].waves, B.o32, s);
    mark(7);
    mark(8);
    if (timed) timed->used = true;
}

std::string shape_text(const Shape &h, const Batch &B) {
    char buf[96];
    snprintf(buf, sizeof(buf), "march=%s shade=%s integrate=%d", kRayBlockName[h.march], kShadeOpts[h.shade].name,
             h.integ < 0 ? B.integ_grid : kIntegGrids[h.integ]);
    return buf;
}

// Writes the context's settled shapes to its cache file, one size class a
// line: "<class>|<march block>|<shade block>|<integrate grid>".
void save_tuning(const vs_ctx *ctx) {
    if (ctx->tune_path.empty()) return;
    FILE *f = fopen(ctx->tune_path.c_str(), "w");
    if (!f) return;
    fprintf(f, "# vscan launch shapes v2: size class|march ray block|shade ray block[w waves]|integrate workgroups\n");
    for (const auto &kv : ctx->tuned)
        fprintf(f, "%s|%s|%s|%d\n", kv.first.c_str(), kRayBlockName[kv.second.march], kShadeOpts[kv.second.shade].name,
                kIntegGrids[kv.second.integ]);
    fclose(f);
}

int block_index(const char *name) {
    for (int i = 0; i < 3; ++i)
        if (!strcmp(name, kRayBlockName[i])) return i;
    return -1;
}

void load_tuning(vs_ctx *ctx) {
    FILE *f = fopen(ctx->tune_path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        char *fld[4];
        int nf = 0;
        for (char *p = strtok(line, "|\r\n"); p && nf < 4; p = strtok(nullptr, "|\r\n")) fld[nf++] = p;
        if (nf != 4) continue;
        Shape h;
        h.march = block_index(fld[1]);
        h.shade = -1;
        for (int i = 0; i < kNumShade; ++i)
            if (!strcmp(fld[2], kShadeOpts[i].name)) h.shade = i;
        const int g = atoi(fld[3]);
        for (int i = 0; i < kNumInteg; ++i)
            if (kIntegGrids[i] == g) h.integ = i;
        if (h.march < 0 || h.shade < 0 || h.integ < 0) continue;
        ctx->tuned[fld[0]] = h;
    }
    fclose(f);
}

// One step of online tuning: reads the stage times of the previous tuning
// step, then enqueues the next candidate step (eager, timed), or after
// kTuneSteps settles the batch's shape and leaves the step to the caller.
vs_status tune_step(vs_ctx *ctx, Batch &B, hipStream_t s) {
    if (!B.tune) B.tune.reset(new TuneState());
    TuneState &T = *B.tune;
    if (T.pending) {
        HIP_TRY(hipEventSynchronize(T.timer.ev[kMarks - 1]));
        float ms[3];
        HIP_TRY(hipEventElapsedTime(&ms[0], T.timer.ev[5], T.timer.ev[6]));  // march
        HIP_TRY(hipEventElapsedTime(&ms[1], T.timer.ev[6], T.timer.ev[7]));  // shade
        HIP_TRY(hipEventElapsedTime(&ms[2], T.timer.ev[4], T.timer.ev[5]));  // integrate
        for (int k = 0; k < 3; ++k) {
            T.sum[k][T.ci[k]] += ms[k];
            T.cnt[k][T.ci[k]] += 1;
        }
        T.pending = false;
    }
    if (T.step < kTuneSteps) {
        T.ci[0] = T.step % kNumMarch;
        T.ci[1] = T.step % kNumShade;
        T.ci[2] = T.step % kNumInteg;
        B.shape.march = kMarchOpts[T.ci[0]];
        B.shape.shade = T.ci[1];
        B.shape.integ = T.ci[2];
        HIP_TRY(T.timer.init());
        enqueue_step(B, s, &T.timer);
        HIP_TRY(hipGetLastError());
        T.pending = true;
        ++T.step;
        return VS_OK;
    }
    const int nopt[3] = {kNumMarch, kNumShade, kNumInteg};
    int best[3] = {0, 0, 0};
    std::string ms;
    for (int k = 0; k < 3; ++k) {
        double bm = 1e30;
        for (int c = 0; c < nopt[k]; ++c) {
            const double m = T.cnt[k][c] ? T.sum[k][c] / T.cnt[k][c] : 1e30;
            if (m < bm) {
                bm = m;
                best[k] = c;
            }
            char b[32];
            snprintf(b, sizeof(b), "%s%.4f", c ? "," : (k ? "|" : ""), T.cnt[k][c] ? m : -1.0);
            ms += b;
        }
        // The tuning steps are the sessions' first frames, whose volumes are
        // still filling: ray blocks within a few percent of each other there
        // can rank the other way once the surface is built (measured on the
        // 64-session workload: march 16x4 0.5% ahead of 8x8 while tuning, 3%
        // behind it afterwards; shading 32x2 level with 8x8 while tuning, 9%
        // behind it afterwards). A ray block replaces the first candidate
        // (8x8 for both) only when it is clearly faster.
        if (k < 2 && T.cnt[k][0] && bm > T.sum[k][0] / T.cnt[k][0] * (1.0 - kTuneMargin)) best[k] = 0;
    }
    B.shape.march = kMarchOpts[best[0]];
    B.shape.shade = best[1];
    B.shape.integ = best[2];
    B.tuned = true;
    B.tune.reset();
    free_batch_graph(B);
    ctx->tuned[B.size_class] = B.shape;
    ctx->tune_ms[B.size_class] = ms;
    save_tuning(ctx);
    return VS_OK;
}

vs_status run_step(vs_ctx *ctx, Batch &B, hipStream_t s, Timer *timed) {
    if (ctx->tune_mode != VS_TUNE_OFF && !B.tuned) {
        auto it = ctx->tuned.find(B.size_class);
        if (it != ctx->tuned.end()) {
            B.shape = it->second;
            B.tuned = true;
            free_batch_graph(B);
        } else if (ctx->tune_mode == VS_TUNE_ONLINE) {
            vs_status r = tune_step(ctx, B, s);
            if (r != VS_OK || !B.tuned) return r;  // a tuning step was enqueued
        }
    }
    if (timed || !ctx->use_graph) {
        if (timed) HIP_TRY(timed->init());
        enqueue_step(B, s, timed);
        HIP_TRY(hipGetLastError());
        return VS_OK;
    }
    if (!B.gexec) {
        // one branch per half of a split batch (fork after an event on the
        // capture stream, join by events back into it), else one chain
        const size_t nb = B.sub.size();
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        std::vector<hipStream_t> side(nb > 1 ? nb - 1 : 0, nullptr);
        std::vector<hipEvent_t> ev(nb, nullptr);
        hipError_t e = hipSuccess;
        for (auto &h : side)
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&h, hipStreamNonBlocking);
        for (auto &x : ev)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        hipGraph_t g = nullptr;
        if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            if (nb == 0) {
                enqueue_step(B, cap, nullptr);
            } else {
                e = hipEventRecord(ev[0], cap);
                for (size_t k = 1; k < nb && e == hipSuccess; ++k) e = hipStreamWaitEvent(side[k - 1], ev[0], 0);
                for (size_t k = 0; k < nb && e == hipSuccess; ++k) {
                    B.sub[k]->shape = B.shape;
                    enqueue_step(*B.sub[k], k == 0 ? cap : side[k - 1], nullptr);
                }
                for (size_t k = 1; k < nb && e == hipSuccess; ++k) {
                    e = hipEventRecord(ev[k], side[k - 1]);
                    if (e == hipSuccess) e = hipStreamWaitEvent(cap, ev[k], 0);
                }
            }
            const hipError_t ee = hipStreamEndCapture(cap, &g);
            if (e == hipSuccess) e = ee;
        }
        for (auto h : side)
            if (h) hipStreamDestroy(h);
        for (auto x : ev)
            if (x) hipEventDestroy(x);
        hipStreamDestroy(cap);
        if (e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return fail(VS_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        }
        B.graph = g;
        HIP_TRY(hipGraphInstantiate(&B.gexec, g, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(B.gexec, s));
    return VS_OK;
}

vs_status sync_last(vs_session *s) {
    HIP_TRY(hipStreamSynchronize(s->last ? s->last : s->stream));
    if (s->last != s->stream) HIP_TRY(hipStreamSynchronize(s->stream));
    return VS_OK;
}

// A NULL frame keeps the raw buffer as it is (set by vs_set_depth or a pre-step).
vs_status copy_depth(vs_session *s, const void *depth, int is_host, hipStream_t st) {
    if (!depth) return VS_OK;
    const size_t bytes = frame_bytes(s->d);
    HIP_TRY(hipMemcpyAsync((void *)s->hp.raw, depth, bytes, is_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                           st));
    return VS_OK;
}

// Host-memory frames of a batched step, frame k for batch slot slot[k]
// (slot = k when null) of session sess[k]; a null frame keeps that session's
// raw frame as it is. The CPU copies the frames into the stage's next host
// area (several threads for a large batch), one transfer moves the area to
// the device on the copy stream, and one gather per 64 frames on stream h
// moves them into the raw-frame buffers once the transfer has run. The
// caller's buffers are free again when this returns.
vs_status stage_host_frames(HostStage &S, const SessParams *dP, vs_session *const *sess, const int *slot,
                            const uint16_t *const *depths, int n, hipStream_t h) {
    struct Job {
        const void *src;
        size_t off, bytes;
        int slot;
    };
    std::vector<Job> jobs;
    size_t need = 0;
    for (int k = 0; k < n; ++k) {
        if (!depths[k]) continue;
        const size_t bytes = frame_bytes(sess[k]->d);
        jobs.push_back({depths[k], need, bytes, slot ? slot[k] : k});
        need += (bytes + 255) & ~(size_t)255;
    }
    if (need == 0) return VS_OK;
    if (!S.cs) {
        HIP_TRY(hipStreamCreateWithFlags(&S.cs, hipStreamNonBlocking));
        for (int g = 0; g < 2; ++g) {
            HIP_TRY(hipEventCreateWithFlags(&S.copied[g], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&S.read[g], hipEventDisableTiming));
        }
    }
    if (need > S.bytes) {
        HIP_TRY(hipStreamSynchronize(S.cs));
        for (int g = 0; g < 2; ++g) {
            if (S.used[g]) HIP_TRY(hipEventSynchronize(S.read[g]));
            S.used[g] = false;
            if (S.area[g]) HIP_TRY(hipFree(S.area[g]));
            if (S.pinned[g]) HIP_TRY(hipHostFree(S.pinned[g]));
            S.area[g] = S.pinned[g] = nullptr;
        }
        S.bytes = 0;
        for (int g = 0; g < 2; ++g) {
            HIP_TRY(hipMalloc((void **)&S.area[g], need));
            HIP_TRY(hipHostMalloc((void **)&S.pinned[g], need, hipHostMallocDefault));
        }
        S.bytes = need;
    }
    const int g = S.next;
    S.next ^= 1;
    // the transfer that last read this host area has run
    if (S.used[g]) HIP_TRY(hipEventSynchronize(S.copied[g]));
    char *dst = S.pinned[g];
    const int nt = (int)std::min<size_t>({(size_t)8, std::max(1u, std::thread::hardware_concurrency()),
                                          std::max<size_t>(1, need >> 22)});  // a thread per 4 MB, at most 8
    auto copy_range = [&](int t) {
        for (size_t j = t; j < jobs.size(); j += nt) memcpy(dst + jobs[j].off, jobs[j].src, jobs[j].bytes);
    };
    if (nt > 1) {
        // threads that cannot be started leave their share to this one
        std::vector<std::thread> pool;
        std::vector<int> mine{0};
        for (int t = 1; t < nt; ++t) {
            try {
                pool.emplace_back(copy_range, t);
            } catch (const std::exception &) {
                mine.push_back(t);
            }
        }
        for (int t : mine) copy_range(t);
        for (auto &th : pool) th.join();
    } else {
        copy_range(0);
    }
    if (S.used[g]) HIP_TRY(hipStreamWaitEvent(S.cs, S.read[g], 0));  // the gather that read this device area
    HIP_TRY(hipMemcpyAsync(S.area[g], dst, need, hipMemcpyHostToDevice, S.cs));
    HIP_TRY(hipEventRecord(S.copied[g], S.cs));
    HIP_TRY(hipStreamWaitEvent(h, S.copied[g], 0));
    DepthChunk ch{};
    for (const Job &j : jobs) {
        ch.src[ch.n] = (const uint16_t *)(S.area[g] + j.off);
        ch.slot[ch.n] = j.slot;
        if (++ch.n == kDepthChunk) {
            launch_gather_depth(dP, ch, h);
            ch.n = 0;
        }
    }
    launch_gather_depth(dP, ch, h);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(S.read[g], h));
    S.used[g] = true;
    return VS_OK;
}

}  // namespace

extern "C" {

const char *vs_status_string(int st) {
    switch (st) {
    case VS_OK: return "ok";
    case VS_EINVAL: return "invalid argument";
    case VS_ENOMEM: return "out of device memory";
    case VS_EHIP: return "HIP runtime error";
    case VS_TRACK_LOST: return "tracking lost (singular system)";
    case VS_MOTION_REJECTED: return "motion rejected";
    case VS_BUDGET: return "HBM budget exhausted";
    case VS_TOO_FEW_POINTS: return "too few correspondences";
    case VS_SINGULAR: return "singular normal equations (relative pivot guard)";
    default: return "unknown status";
    }
}
const char *vs_last_error(void) { return g_err.c_str(); }

void vs_session_desc_default(vs_session_desc *d, int width, int height, int n, float cube_m) {
    std::memset(d, 0, sizeof(*d));
    d->width = width;
    d->height = height;
    d->fx = d->fy = 525.0f * width / 640.0f;
    d->cx = width * 0.5f - 0.5f;
    d->cy = height * 0.5f - 0.5f;
    d->levels = 3;
    d->iters[0] = 10;
    d->iters[1] = 7;
    d->iters[2] = 7;
    d->icp_dist2 = (0.055f * cube_m) * (0.055f * cube_m);
    d->vol_dim[0] = d->vol_dim[1] = d->vol_dim[2] = n;
    d->origin[0] = d->origin[1] = d->origin[2] = -0.5f * cube_m;
    d->voxel = cube_m / n;
    d->mu = 5.0f * cube_m / n;
    d->max_weight = 100.0f;
    d->tsdf_type = VS_TSDF_F32;
    d->weight_type = VS_WEIGHT_F32;
    d->bil_sigma_s = 4.0f;
    d->bil_sigma_r = 0.1f;
    d->bil_radius = 2;
    d->integrate_mode = VS_INTEGRATE_DEPTH;
    d->depth_units = VS_DEPTH_MM_U16;
    d->icp_rel_pivot = 0.0f;
}

vs_status vs_ctx_create(int device, vs_ctx **out) {
    if (!out) return fail(VS_EINVAL, "null out");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(VS_EINVAL, "device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    double *scr = nullptr;
    HIP_TRY(hipGetSymbolAddress((void **)&scr, HIP_SYMBOL(scratch)));
    if (!scr) return fail(VS_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
    vs_ctx *c = new vs_ctx();
    c->device = device;
    c->scratch = scr;
    const char *ng = std::getenv("VS_NO_GRAPH");
    c->use_graph = !(ng && ng[0] == '1');
    *out = c;
    return VS_OK;
}

vs_status vs_ctx_destroy(vs_ctx *ctx) {
    if (!ctx) return VS_OK;
    hipSetDevice(ctx->device);
    std::vector<vs_server *> vv = ctx->servers;
    for (vs_server *v : vv) vs_server_destroy(v);
    for (auto &kv : ctx->batches) release_batch(*kv.second);
    ctx->batches.clear();
    std::vector<vs_session *> ss = ctx->sessions;
    for (vs_session *s : ss) vs_session_destroy(s);
    ctx->timer.destroy();
    delete ctx;
    return VS_OK;
}

size_t vs_session_bytes(const vs_session_desc *d) {
    if (check_desc(d) != VS_OK) return 0;
    return make_layout(*d).total;
}

}  // extern "C"

namespace {
// A session whose memory comes from the stream-ordered allocator when
// `pooled` (slot-table sessions: closing one frees it in stream order after
// the steps in flight, where hipFree would wait for the whole device).
vs_status create_session(vs_ctx *ctx, const vs_session_desc *desc, const double T_gc0[16], bool pooled,
                         vs_session **out) {
    if (!ctx || !out) return fail(VS_EINVAL, "null argument");
    vs_status cs = check_desc(desc);
    if (cs != VS_OK) return cs;
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<vs_session> s(new vs_session());
    s->ctx = ctx;
    s->d = *desc;
    s->L = make_layout(*desc);
    const Layout &L = s->L;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->last = s->stream;
    hipError_t e = pooled ? hipMallocAsync((void **)&s->mem, L.total, s->stream) : hipMalloc((void **)&s->mem, L.total);
    if (e == hipSuccess && pooled) e = hipStreamSynchronize(s->stream);  // usable from any stream from here on
    if (e != hipSuccess) {
        hipStreamDestroy(s->stream);
        return fail(VS_ENOMEM, "device allocation of %zu bytes failed: %s", L.total, hipGetErrorString(e));
    }
    s->pooled = pooled;
    char *m = s->mem;
    SessParams &P = s->hp;
    std::memset(&P, 0, sizeof(P));
    const vs_session_desc &d = *desc;
    // per-level geometry; intrinsics halve per level with c <- (c - 0.5)/2
    // (pose_tracker.cu:764-776, tsdf_renderer.cu:1611-1623)
    float fx = d.fx, fy = d.fy, px = d.cx, py = d.cy;
    for (int l = 0; l < kMaxLevels; ++l) {
        LevelDesc &ld = P.lev[l];
        ld.W = d.width >> l;
        ld.H = d.height >> l;
        ld.tiles_x = (ld.W + kTileW - 1) / kTileW;
        ld.tiles_y = (ld.H + kTileH - 1) / kTileH;
        ld.K[0] = fx;
        ld.K[1] = fy;
        ld.K[2] = px;
        ld.K[3] = py;
        fx = 0.5f * fx;
        fy = 0.5f * fy;
        px = 0.5f * (px - 0.5f);
        py = 0.5f * (py - 0.5f);
        P.iters[l] = l < d.levels ? d.iters[l] : 0;
    }
    P.levels = d.levels;
    P.icp_dist2 = d.icp_dist2;
    P.icp_min_valid = d.icp_min_valid;
    P.max_motion_m = d.max_motion_m;
    P.max_motion_deg = d.max_motion_deg;
    P.icp_rel_pivot = d.icp_rel_pivot;
    P.raw_f32 = d.depth_units == VS_DEPTH_M_F32;
    for (int k = 0; k < 3; ++k) {
        P.vdim[k] = d.vol_dim[k];
        P.bdim[k] = d.vol_dim[k] / kBrick;
        P.origin[k] = d.origin[k];
    }
    P.vox_inv = 1.0f / d.voxel;
    P.voxel = d.voxel;
    P.mu = d.mu;
    P.max_w = d.max_weight;
    // cube centre = origin + dim*voxel/2 (z_yooji_object_scanner.cpp:306); radius = |origin - centre|
    for (int k = 0; k < 3; ++k) P.cube_cen[k] = d.origin[k] + (float)d.vol_dim[k] * d.voxel / 2.0f;
    {
        const float ex = d.origin[0] - P.cube_cen[0], ey = d.origin[1] - P.cube_cen[1], ez = d.origin[2] - P.cube_cen[2];
        P.r_cube = sqrtf(ex * ex + ey * ey + ez * ez);
    }
    P.step_scale = 0.9f * P.mu * P.vox_inv;
    P.total_max = 2.0f * P.r_cube * P.vox_inv;
    // the march's iteration cap: the most steps of a ray that only moves
    // forward (by at least min(step_scale, 1) voxels a step) before its
    // distance cap, plus 64; only a ray that keeps stepping back behind a
    // surface reaches it, and the reference loop would not end for such a
    // ray (tsdf_renderer.cu:544-560 sums signed step lengths)
    P.max_iter = (int)(P.total_max / std::min(P.step_scale, 1.0f)) + 64;
    // bilateral spatial weights exactly as updateGaussian (img_proc.cu:275-286),
    // and their products g[i] g[j] as d_bilateral_filter forms them (:126-134)
    {
        const int n = 2 * d.bil_radius + 1;
        float g[2 * 3 + 1];
        for (int i = 0; i < n; ++i) {
            float x = (float)(i - d.bil_radius);
            g[i] = expf(-(x * x) / (2 * d.bil_sigma_s * d.bil_sigma_s));
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) P.bil_w[i * n + j] = g[i] * g[j];
    }
    // exp(-x / (2 sigma_r^2)) = 2^(x * -log2(e) / (2 sigma_r^2)): one scaling
    // on the host in place of two multiplications per tap
    P.bil_nk = (float)(-1.4426950408889634 / (2.0 * (double)d.bil_sigma_r * (double)d.bil_sigma_r));
    P.bil_radius = d.bil_radius;
    P.raw = m + L.raw;
    for (int l = 0; l < kMaxLevels; ++l) {
        P.filt[l] = (float *)(m + L.filt[l]);
        P.mvn[l] = (MapXYZ *)(m + L.mvn[l]);
        P.mimg[l] = (uint8_t *)(m + L.mimg[l]);
    }
    P.vox = m + L.vox;
    P.bvis = (uint32_t *)(m + L.bvis);
    P.vmask = (uint8_t *)(m + L.vmask);
    P.bocc = (uint32_t *)(m + L.bocc);
    P.bband = (uint32_t *)(m + L.bband);
    P.sband = (uint32_t *)(m + L.sband);
    P.sdist = (uint8_t *)(m + L.sdist);
    P.sdtmp = (uint8_t *)(m + L.sdtmp);
    for (int k = 0; k < 3; ++k) P.sdim[k] = L.sdim[k];
    P.sband_words = (int)L.nswords;
    P.icp_part = (double *)(m + L.icp_part);
    {
        float *wt = (float *)(m + L.wtab);
        P.wtab[0] = wt;
        P.wtab[1] = wt + d.vol_dim[0];
        P.wtab[2] = wt + d.vol_dim[0] + d.vol_dim[1];
        // the same FP32 operations as the device would do (division, then
        // addition, each rounded once; this file is built without contraction)
        std::vector<float> h((size_t)d.vol_dim[0] + d.vol_dim[1] + d.vol_dim[2]);
        size_t o = 0;
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < d.vol_dim[a]; ++i) {
                volatile float q = ((float)i + 0.5f) / P.vox_inv;
                h[o++] = q + P.origin[a];
            }
        HIP_TRY(hipMemcpy(wt, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    P.st = (SessState *)(m + L.state);
    P.tsdf_type = d.tsdf_type;
    P.weight_type = d.weight_type;
    P.integ_mode = d.integrate_mode;
    if (d.integrate_mode == VS_INTEGRATE_COLOR) {
        P.vcol = (uint32_t *)(m + L.vcol);
        P.rgb = (const uint8_t *)(m + L.rgb);
        for (int l = 0; l < kMaxLevels; ++l) P.mtex[l] = l < d.levels ? (uint8_t *)(m + L.mtex[l]) : nullptr;
    }
    P.sid = (int)ctx->sessions.size();
    s->st = P.st;
    // self batch
    Batch &B = s->self;
    B.n = 1;
    B.tt = d.tsdf_type;
    B.wt = d.weight_type;
    B.dP = (SessParams *)(m + L.params);
    {
        ListSet ls;
        ls.add(0, s.get());
        vs_status r = ls.place(B, m + L.lists, 2048);
        if (r != VS_OK) return r;
    }
    B.members = {s.get()};
    HIP_TRY(hipMemcpy(B.dP, &P, sizeof(P), hipMemcpyHostToDevice));
    // state: pose, counters; volume F = 1, W = 0 (vs_session_reset); empty
    // pyramid and images, model maps at the -100 sentinel; all on the
    // session's stream (one wait at the end of set-up)
    HIP_TRY(hipMemsetAsync(m + L.state, 0, sizeof(SessState), s->stream));
    const uint32_t kMiss = 0xC2C80000u;  // -100.0f
    for (int l = 0; l < kMaxLevels; ++l) {
        const size_t n = (size_t)(d.width >> l) * (d.height >> l);
        HIP_TRY(hipMemsetAsync(m + L.filt[l], 0, n * 4, s->stream));
        HIP_TRY(hipMemsetAsync(m + L.mimg[l], 0, n, s->stream));
        if (P.mtex[l]) HIP_TRY(hipMemsetAsync(P.mtex[l], 0, n * 3, s->stream));
        launch_fill_u32((uint32_t *)(m + L.mvn[l]), (n + 1) * 6, kMiss, s->stream);
    }
    HIP_TRY(hipGetLastError());
    vs_session *raw = s.release();
    ctx->sessions.push_back(raw);
    vs_status r = vs_session_reset(raw, T_gc0, nullptr);
    if (r != VS_OK) {
        vs_session_destroy(raw);
        return r;
    }
    HIP_TRY(hipStreamSynchronize(raw->stream));
    *out = raw;
    return VS_OK;
}
}  // namespace

extern "C" {

vs_status vs_session_create(vs_ctx *ctx, const vs_session_desc *desc, const double T_gc0[16], vs_session **out) {
    return create_session(ctx, desc, T_gc0, false, out);
}

vs_status vs_session_destroy(vs_session *s) {
    if (!s) return VS_OK;
    hipSetDevice(s->ctx->device);
    if (s->server) {
        // A slot-table session: its slot takes the placeholder block after
        // the steps in flight, and its memory is freed in stream order after
        // them; the host waits for neither.
        vs_server *v = s->server;
        server_forget(s);
        drop_batches_with(s->ctx, s);
        free_batch_graph(s->self);
        s->timer.destroy();
        if (s->stream) hipStreamDestroy(s->stream);
        if (s->mem) {
            if (s->pooled) {
                hipFreeAsync(s->mem, server_order(v));
            } else {
                (void)server_quiesce(v);
                hipFree(s->mem);
            }
        }
        auto &vs = s->ctx->sessions;
        vs.erase(std::remove(vs.begin(), vs.end(), s), vs.end());
        delete s;
        return VS_OK;
    }
    (void)sync_last(s);
    drop_batches_with(s->ctx, s);
    free_batch_graph(s->self);
    s->timer.destroy();
    if (s->stream) hipStreamDestroy(s->stream);
    if (s->mem) {
        if (s->pooled) {
            hipFreeAsync(s->mem, nullptr);
            hipStreamSynchronize(nullptr);
        } else {
            hipFree(s->mem);
        }
    }
    auto &v = s->ctx->sessions;
    v.erase(std::remove(v.begin(), v.end(), s), v.end());
    delete s;
    return VS_OK;
}

vs_status vs_session_reset(vs_session *s, const double T_gc0[16], vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    launch_volume_init(s->self.dP, s->L.nvox, s->d.tsdf_type, s->d.weight_type, s->L.nwords, s->L.sdim[2], h);
    SessState init{};
    for (int k = 0; k < 16; ++k) init.T_gc[k] = T_gc0 ? T_gc0[k] : ((k % 5) == 0 ? 1.0 : 0.0);
    init.conv_level = -1;
    init.active = 1;
    HIP_TRY(hipMemcpyAsync(s->st, &init, sizeof(init), hipMemcpyHostToDevice, h));
    launch_pose_derived(s->self.dP, h);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h));
    return VS_OK;
}

vs_status vs_session_desc_get(const vs_session *s, vs_session_desc *out) {
    if (!s || !out) return fail(VS_EINVAL, "null argument");
    *out = s->d;
    return VS_OK;
}

vs_status vs_session_sync(vs_session *s) {
    if (!s) return fail(VS_EINVAL, "null session");
    return sync_last(s);
}

vs_status vs_set_color(vs_session *s, const uint8_t *rgb_planar, int is_host, vs_stream stream) {
    if (!s || !rgb_planar) return fail(VS_EINVAL, "null argument");
    if (!s->hp.rgb) return fail(VS_EINVAL, "session was not created with VS_INTEGRATE_COLOR");
    const size_t bytes = (size_t)s->d.width * s->d.height * 3;
    HIP_TRY(hipMemcpyAsync((void *)s->hp.rgb, rgb_planar, bytes,
                           is_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, pick(s, stream)));
    return VS_OK;
}

vs_status vs_depth_reproject(vs_session *s, const void *src, int src_w, int src_h, const float K_src[4],
                             const float T_dst_src[16], int src_is_float, void *dst, vs_stream stream) {
    if (!s || !src || !K_src || !T_dst_src) return fail(VS_EINVAL, "null argument");
    if (src_w < 1 || src_h < 1 || src_w > 8192 || src_h > 8192) return fail(VS_EINVAL, "source size %dx%d", src_w, src_h);
    if (!dst && (src_is_float != 0) != (s->d.depth_units == VS_DEPTH_M_F32))
        return fail(VS_EINVAL, "src_is_float does not match the depth units of the session's raw buffer");
    if (!(K_src[0] != 0.0f) || !(K_src[1] != 0.0f)) return fail(VS_EINVAL, "source focal length");
    const float *Kd = s->hp.lev[0].K;
    launch_depth_reproject(src_is_float, src, src_w, src_h, K_src, Kd, T_dst_src,
                           (uint32_t *)(s->mem + s->L.keys), dst ? dst : (void *)s->hp.raw, s->d.width, s->d.height,
                           pick(s, stream));
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_depth_crop_to_volume(vs_session *s, void *depth, int w, int h, const float K[4], const float *T_cg,
                                  int is_float, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    if (!depth) {
        if ((is_float != 0) != (s->d.depth_units == VS_DEPTH_M_F32))
            return fail(VS_EINVAL, "is_float does not match the depth units of the session's raw buffer");
        depth = (void *)s->hp.raw;
        w = s->d.width;
        h = s->d.height;
        K = s->hp.lev[0].K;
    }
    if (!K || w < 1 || h < 1 || w > 8192 || h > 8192) return fail(VS_EINVAL, "frame %dx%d / intrinsics", w, h);
    launch_depth_crop(s->self.dP, is_float, depth, w, h, K, T_cg, pick(s, stream));
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_set_depth(vs_session *s, const uint16_t *depth_mm, int is_host, vs_stream stream) {
    if (!s || !depth_mm) return fail(VS_EINVAL, "null argument");
    return copy_depth(s, depth_mm, is_host, pick(s, stream));
}

vs_status vs_bilateral(vs_session *s, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    launch_fuse_ok(s->self.dP, 1, h);
    launch_bilateral(s->self.dP, s->self.dWork[0], s->self.nWork[0], h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_bilateral_filter_depth(vs_session *s, const void *depth_in, int in_is_host, float *depth_out_dev,
                                    vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    vs_status r = copy_depth(s, depth_in, in_is_host, h);
    if (r != VS_OK) return r;
    launch_fuse_ok(s->self.dP, 1, h);
    launch_bilateral(s->self.dP, s->self.dWork[0], s->self.nWork[0], h);
    HIP_TRY(hipGetLastError());
    if (depth_out_dev)
        HIP_TRY(hipMemcpyAsync(depth_out_dev, s->hp.filt[0], (size_t)s->d.width * s->d.height * 4,
                               hipMemcpyDeviceToDevice, h));
    return VS_OK;
}

vs_status vs_pyramid(vs_session *s, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    launch_fuse_ok(s->self.dP, 1, h);
    for (int l = 1; l < s->d.levels; ++l) launch_pyr_down(s->self.dP, s->self.dWork[l], s->self.nWork[l], l, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_vertex_normal_maps(vs_session *s, int level, float *vtx_dev, float *nrm_dev, vs_stream stream) {
    if (!s || !vtx_dev || !nrm_dev) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    hipStream_t h = pick(s, stream);
    launch_vnmap(s->self.dP, s->self.dWork[level], s->self.nWork[level], level, vtx_dev, nrm_dev, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_vertex_normal_maps_inplace(vs_session *s, int level, float *vtx_dev, float *nrm_dev,
                                        vs_stream stream) {
    if (!s || !vtx_dev || !nrm_dev) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    hipStream_t h = pick(s, stream);
    launch_vnmap(s->self.dP, s->self.dWork[level], s->self.nWork[level], level, vtx_dev, nrm_dev, h);
    launch_vnmap_inplace(s->self.dP, level, vtx_dev, nrm_dev, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_track(vs_session *s, const double *T_01_init, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    const Batch &B = s->self;
    launch_track_setup_forced(B.dP, 1, T_01_init, h);
    for (int l = B.levels - 1; l >= 0; --l)
        for (int it = 0; it < B.max_iters[l]; ++it) launch_icp_iter(B.dP, B.dGroup[l], B.nGroup[l], B.n, l, it, h);
    launch_track_finalize(B.dP, 1, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_integrate(vs_session *s, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    const Batch &B = s->self;
    launch_fuse_ok(B.dP, 1, h);
    launch_vis_clear(B.dP, 1, B.bl, h);
    launch_vis_mark(B.dP, B.dWork[0], B.nWork[0], 1, B.max_words, B.bl, h);
    launch_integrate(B.dP, B.n, B.bl, B.tt, B.wt, integ_grid(B), B.sd_slices, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_integrate_depth(vs_session *s, const void *depth_dev, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    vs_status r = copy_depth(s, depth_dev, 0, pick(s, stream));
    if (r != VS_OK) return r;
    return vs_integrate(s, stream);
}

vs_status vs_render_maps(vs_session *s, int level, float *depth_dev, float *vtx_dev, float *nrm_dev,
                         uint8_t *img_dev, vs_stream stream) {
    vs_status r = vs_raycast(s, level, stream);
    if (r != VS_OK) return r;
    const size_t npx = (size_t)(s->d.width >> level) * (s->d.height >> level);
    if (depth_dev || vtx_dev || nrm_dev || img_dev)
        launch_maps_out(s->self.dP, level, npx, depth_dev, vtx_dev, nrm_dev, img_dev, s->last);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_raycast(vs_session *s, int level, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    hipStream_t h = pick(s, stream);
    const Batch &B = s->self;
    launch_fuse_ok(B.dP, 1, h);
    launch_raycast(B.dP, B.dWork[level], B.nWork[level], level, B.tt, B.wt, B.o32, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_fuse_frame(vs_session *s, const uint16_t *depth_mm, int is_host, vs_stream stream) {
    if (!s) return fail(VS_EINVAL, "null session");
    hipStream_t h = pick(s, stream);
    vs_status r = copy_depth(s, depth_mm, is_host, h);
    if (r != VS_OK) return r;
    return run_step(s->ctx, s->self, h, s->timing ? &s->timer : nullptr);
}

namespace {
// Batches of at least kSplitMin sessions run as kSplitParts parts in a
// captured step (64 sessions of 640x480 / 512^3: 10.70 ms per step as one
// chain, 10.37 as 2 parts, 10.15 as 3, 10.21 as 4; with faster kernels
// 8.30 as 2, 8.15-8.21 as 3, 8.16-8.17 as 4 and 8.74-8.79 as 5, more
// branches than the 4 hardware queues a process gets; later still 6.32 as
// 2, 6.23 as 3 and 6.27 as 4).
constexpr int kSplitMin = 16;
#ifndef VS_SPLIT_PARTS
#define VS_SPLIT_PARTS 3
#endif
constexpr int kSplitParts = VS_SPLIT_PARTS;

// A context batch over `members`: its own parameter array and lists.
vs_status make_ctx_batch(const std::vector<vs_session *> &members, Batch &B) {
    const int n = (int)members.size();
    B.n = n;
    B.tt = members[0]->d.tsdf_type;
    B.wt = members[0]->d.weight_type;
    B.members = members;
    std::vector<SessParams> hp(n);
    ListSet ls;
    for (int i = 0; i < n; ++i) {
        hp[i] = members[i]->hp;
        ls.add(i, members[i]);
    }
    const size_t offP = 0, offL = align_up(n * sizeof(SessParams)), total = offL + ls.bytes();
    char *mem = nullptr;
    hipError_t e = hipMalloc((void **)&mem, total);
    if (e != hipSuccess) return fail(VS_ENOMEM, "batch allocation (%zu bytes): %s", total, hipGetErrorString(e));
    B.owned = mem;
    B.dP = (SessParams *)(mem + offP);
    HIP_TRY(hipMemcpy(B.dP, hp.data(), n * sizeof(SessParams), hipMemcpyHostToDevice));
    return ls.place(B, mem + offL, 8192);
}
}  // namespace

vs_status vs_fuse_batch(vs_ctx *ctx, vs_session *const *sessions, const uint16_t *const *depths, int n, int is_host,
                        vs_stream stream) {
    if (!ctx || !sessions || n <= 0) return fail(VS_EINVAL, "bad batch arguments");
    std::vector<vs_session *> key(sessions, sessions + n);
    for (int i = 0; i < n; ++i) {
        if (!key[i] || key[i]->ctx != ctx) return fail(VS_EINVAL, "session %d not in context", i);
        if (key[i]->d.tsdf_type != key[0]->d.tsdf_type || key[i]->d.weight_type != key[0]->d.weight_type)
            return fail(VS_EINVAL, "sessions of one batch must share storage types");
        for (int j = 0; j < i; ++j)
            if (key[j] == key[i]) return fail(VS_EINVAL, "session repeated in batch");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t h = stream ? (hipStream_t)stream : key[0]->stream;
    auto it = ctx->batches.find(key);
    Batch *B;
    if (it == ctx->batches.end()) {
        std::unique_ptr<Batch> nb(new Batch());
        vs_status r = make_ctx_batch(key, *nb);
        if (r == VS_OK && n >= kSplitMin) {
            for (int g = 0; g < kSplitParts && r == VS_OK; ++g) {
                const std::vector<vs_session *> part(key.begin() + (size_t)n * g / kSplitParts,
                                                     key.begin() + (size_t)n * (g + 1) / kSplitParts);
                nb->sub.emplace_back(new Batch());
                r = make_ctx_batch(part, *nb->sub.back());
            }
        }
        if (r != VS_OK) {
            release_batch(*nb);
            return r;
        }
        B = nb.get();
        ctx->batches[key] = std::move(nb);
    } else {
        B = it->second.get();
    }
    if (depths && is_host) {
        vs_status r = stage_host_frames(B->hs, B->dP, key.data(), nullptr, depths, n, h);
        if (r != VS_OK) return r;
    } else if (depths) {
        // device-resident frames: one gather launch per 64 sessions that have one
        DepthChunk ch{};
        for (int i = 0; i < n; ++i) {
            if (!depths[i]) continue;
            ch.src[ch.n] = depths[i];
            ch.slot[ch.n] = i;
            if (++ch.n == kDepthChunk) {
                launch_gather_depth(B->dP, ch, h);
                ch.n = 0;
            }
        }
        launch_gather_depth(B->dP, ch, h);
        HIP_TRY(hipGetLastError());
    }
    for (int i = 0; i < n; ++i) key[i]->last = h;
    return run_step(ctx, *B, h, ctx->timing ? &ctx->timer : nullptr);
}

vs_status vs_get_frame_status(vs_session *s, int *last_status, int *frames_fused, int *icp_iters) {
    if (!s) return fail(VS_EINVAL, "null session");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    SessState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof(st), hipMemcpyDeviceToHost));
    if (last_status) *last_status = st.last_status;
    if (frames_fused) *frames_fused = st.frames;
    if (icp_iters) *icp_iters = st.solves;
    return VS_OK;
}

vs_status vs_get_pose(vs_session *s, double T_gc[16]) {
    if (!s || !T_gc) return fail(VS_EINVAL, "null argument");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    HIP_TRY(hipMemcpy(T_gc, s->st->T_gc, 16 * sizeof(double), hipMemcpyDeviceToHost));
    return VS_OK;
}

vs_status vs_set_pose(vs_session *s, const double T_gc[16], vs_stream stream) {
    if (!s || !T_gc) return fail(VS_EINVAL, "null argument");
    hipStream_t h = pick(s, stream);
    HIP_TRY(hipStreamSynchronize(h));
    HIP_TRY(hipMemcpy(s->st->T_gc, T_gc, 16 * sizeof(double), hipMemcpyHostToDevice));
    launch_pose_derived(s->self.dP, h);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h));
    return VS_OK;
}

vs_status vs_get_filtered(vs_session *s, int level, float *host) {
    if (!s || !host) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level);
    HIP_TRY(hipMemcpy(host, s->hp.filt[level], n * 4, hipMemcpyDeviceToHost));
    return VS_OK;
}

vs_status vs_set_filtered(vs_session *s, int level, const float *host) {
    if (!s || !host) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level);
    HIP_TRY(hipMemcpy(s->hp.filt[level], host, n * 4, hipMemcpyHostToDevice));
    return VS_OK;
}

vs_status vs_get_maps(vs_session *s, int level, float *depth, float *vtx, float *nrm, uint8_t *img) {
    if (!s) return fail(VS_EINVAL, "null session");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level);
    // the device maps interleave vertex and normal records; the ABI's maps
    // are x, y, z per pixel, one map each
    std::vector<MapXYZ> v(2 * n);
    HIP_TRY(hipMemcpy(v.data(), s->hp.mvn[level], 2 * n * sizeof(MapXYZ), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        const MapXYZ a = v[2 * i], b = v[2 * i + 1];
        if (vtx) std::memcpy(vtx + 3 * i, &a, sizeof(MapXYZ));
        if (nrm) std::memcpy(nrm + 3 * i, &b, sizeof(MapXYZ));
        if (depth) depth[i] = (a.x == -100.0f) ? 0.0f : a.z;
    }
    if (img) HIP_TRY(hipMemcpy(img, s->hp.mimg[level], n, hipMemcpyDeviceToHost));
    return VS_OK;
}

vs_status vs_set_maps(vs_session *s, int level, const float *vtx, const float *nrm) {
    if (!s || !vtx || !nrm) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level);
    std::vector<MapXYZ> v(2 * n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&v[2 * i], vtx + 3 * i, sizeof(MapXYZ));
        std::memcpy(&v[2 * i + 1], nrm + 3 * i, sizeof(MapXYZ));
    }
    HIP_TRY(hipMemcpy(s->hp.mvn[level], v.data(), 2 * n * sizeof(MapXYZ), hipMemcpyHostToDevice));
    return VS_OK;
}

vs_status vs_get_texture(vs_session *s, int level, uint8_t *rgb_planar) {
    if (!s || !rgb_planar) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    if (!s->hp.mtex[level]) return fail(VS_EINVAL, "session was not created with VS_INTEGRATE_COLOR");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level) * 3;
    HIP_TRY(hipMemcpy(rgb_planar, s->hp.mtex[level], n, hipMemcpyDeviceToHost));
    return VS_OK;
}

vs_status vs_volume_download_color(vs_session *s, uint8_t *rgb_planar) {
    if (!s || !rgb_planar) return fail(VS_EINVAL, "null argument");
    if (!s->hp.vcol) return fail(VS_EINVAL, "session was not created with VS_INTEGRATE_COLOR");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = s->L.nvox;
    uint8_t *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, n * 3);
    if (e != hipSuccess) return fail(VS_ENOMEM, "download buffer: %s", hipGetErrorString(e));
    launch_color_to_linear(s->self.dP, n, d, s->stream);
    e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(rgb_planar, d, n * 3, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(VS_EHIP, "colour download: %s", hipGetErrorString(e));
    return VS_OK;
}

vs_status vs_volume_carve(vs_session *s, const uint8_t *valid, int is_host, unsigned *carved) {
    if (!s || !valid) return fail(VS_EINVAL, "null argument");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t n = s->L.nvox;
    char *tmp = nullptr;
    hipError_t e = hipMalloc((void **)&tmp, 256 + (is_host ? n : 0));
    if (e != hipSuccess) return fail(VS_ENOMEM, "carve buffer: %s", hipGetErrorString(e));
    unsigned *cnt = (unsigned *)tmp;
    const uint8_t *dv = valid;
    e = hipMemset(cnt, 0, 4);
    if (e == hipSuccess && is_host) {
        e = hipMemcpy(tmp + 256, valid, n, hipMemcpyHostToDevice);
        dv = (const uint8_t *)(tmp + 256);
    }
    if (e == hipSuccess) {
        launch_carve(s->self.dP, s->d.tsdf_type, s->d.weight_type, dv, cnt, s->stream);
        e = hipStreamSynchronize(s->stream);
    }
    unsigned c = 0;
    if (e == hipSuccess) e = hipMemcpy(&c, cnt, 4, hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e != hipSuccess) return fail(VS_EHIP, "carve: %s", hipGetErrorString(e));
    if (carved) *carved = c;
    return VS_OK;
}

// ---------------------------------------------------------------------------
// snapshots
// ---------------------------------------------------------------------------
static void snapshot_sizes(const vs_session_desc &d, uint32_t &sv, uint32_t &sw, uint32_t &sc) {
    sv = d.tsdf_type == VS_TSDF_F16 ? 2 : 4;
    sw = d.weight_type == VS_WEIGHT_U8 ? 1 : 4;
    sc = d.integrate_mode == VS_INTEGRATE_COLOR ? 4 : 0;
}

// Occupied bricks (bocc bits) of a session, ascending.
static vs_status occupied_bricks(vs_session *s, std::vector<uint32_t> &out) {
    std::vector<uint32_t> bits(s->L.nwords);
    HIP_TRY(hipMemcpy(bits.data(), s->hp.bocc, bits.size() * 4, hipMemcpyDeviceToHost));
    out.clear();
    for (uint32_t b = 0; b < s->L.nbricks; ++b)
        if ((bits[b >> 5] >> (b & 31)) & 1u) out.push_back(b);
    return VS_OK;
}

vs_status vs_session_snapshot_bytes(vs_session *s, size_t *bytes) {
    if (!s || !bytes) return fail(VS_EINVAL, "null argument");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    std::vector<uint32_t> bl;
    r = occupied_bricks(s, bl);
    if (r != VS_OK) return r;
    uint32_t sv, sw, sc;
    snapshot_sizes(s->d, sv, sw, sc);
    *bytes = sizeof(vs_snapshot_header) + bl.size() * (4 + (size_t)kBrickVox * (sv + sw + sc));
    return VS_OK;
}

vs_status vs_session_snapshot(vs_session *s, void *buf, size_t capacity, size_t *bytes) {
    if (!s || !bytes) return fail(VS_EINVAL, "null argument");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    std::vector<uint32_t> bl;
    r = occupied_bricks(s, bl);
    if (r != VS_OK) return r;
    uint32_t sv, sw, sc;
    snapshot_sizes(s->d, sv, sw, sc);
    const size_t n = bl.size(), packed = n * (size_t)kBrickVox * (sv + sw + sc);
    const size_t need = sizeof(vs_snapshot_header) + n * 4 + packed;
    *bytes = need;
    if (!buf || capacity < need) return buf ? fail(VS_EINVAL, "snapshot needs %zu bytes, buffer has %zu", need, capacity)
                                            : VS_OK;
    SessState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof(st), hipMemcpyDeviceToHost));
    vs_snapshot_header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "VSCANSNP", 8);
    h.version = 2;
    h.header_bytes = sizeof(h);
    h.desc = s->d;
    std::memcpy(h.T_gc, st.T_gc, sizeof(h.T_gc));
    h.frames = st.frames;
    h.bricks_total = s->L.nbricks;
    h.bricks_stored = (uint32_t)n;
    h.value_bytes = sv;
    h.weight_bytes = sw;
    h.color_bytes = sc;
    char *out = (char *)buf;
    std::memcpy(out, &h, sizeof(h));
    std::memcpy(out + sizeof(h), bl.data(), n * 4);
    if (n) {
        char *dev = nullptr;
        hipError_t e = hipMalloc((void **)&dev, align_up(n * 4) + packed);
        if (e != hipSuccess) return fail(VS_ENOMEM, "snapshot buffer: %s", hipGetErrorString(e));
        e = hipMemcpy(dev, bl.data(), n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_brick_pack(s->self.dP, (const uint32_t *)dev, (unsigned)n, (int)sv, (int)sw, (int)sc,
                              (uint8_t *)dev + align_up(n * 4), 0, s->stream);
            e = hipStreamSynchronize(s->stream);
        }
        if (e == hipSuccess) e = hipMemcpy(out + sizeof(h) + n * 4, dev + align_up(n * 4), packed, hipMemcpyDeviceToHost);
        hipFree(dev);
        if (e != hipSuccess) return fail(VS_EHIP, "snapshot: %s", hipGetErrorString(e));
    }
    return VS_OK;
}

// Header checks of a snapshot of `bytes` bytes; returns its descriptor.
static vs_status snapshot_check(const void *buf, size_t bytes, vs_snapshot_header &h) {
    if (!buf || bytes < sizeof(vs_snapshot_header)) return fail(VS_EINVAL, "snapshot too short");
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, "VSCANSNP", 8) != 0 || h.version != 2 || h.header_bytes != sizeof(h))
        return fail(VS_EINVAL, "not a version-2 snapshot");
    vs_status cs = check_desc(&h.desc);
    if (cs != VS_OK) return cs;
    uint32_t sv, sw, sc;
    snapshot_sizes(h.desc, sv, sw, sc);
    const Layout L = make_layout(h.desc);
    if (h.value_bytes != sv || h.weight_bytes != sw || h.color_bytes != sc || h.bricks_total != L.nbricks ||
        h.bricks_stored > L.nbricks)
        return fail(VS_EINVAL, "snapshot sizes do not match its descriptor");
    const size_t need = sizeof(h) + (size_t)h.bricks_stored * (4 + (size_t)kBrickVox * (sv + sw + sc));
    if (bytes < need) return fail(VS_EINVAL, "snapshot truncated (%zu of %zu bytes)", bytes, need);
    const uint32_t *idx = (const uint32_t *)((const char *)buf + sizeof(h));
    for (uint32_t i = 0; i < h.bricks_stored; ++i)
        if (idx[i] >= L.nbricks) return fail(VS_EINVAL, "snapshot brick index %u out of range", idx[i]);
    return VS_OK;
}

}  // extern "C"

namespace {
// vs_session_restore; `pooled` as create_session.
vs_status restore_session(vs_ctx *ctx, const void *buf, size_t bytes, bool pooled, vs_session **out) {
    if (!ctx || !out) return fail(VS_EINVAL, "null argument");
    vs_snapshot_header h;
    vs_status r = snapshot_check(buf, bytes, h);
    if (r != VS_OK) return r;
    vs_session *s = nullptr;
    r = create_session(ctx, &h.desc, h.T_gc, pooled, &s);
    if (r != VS_OK) return r;
    const size_t n = h.bricks_stored;
    const size_t packed = n * (size_t)kBrickVox * (h.value_bytes + h.weight_bytes + h.color_bytes);
    hipError_t e = hipSuccess;
    if (n) {
        char *dev = nullptr;
        e = pooled ? hipMallocAsync((void **)&dev, align_up(n * 4) + packed, s->stream)
                   : hipMalloc((void **)&dev, align_up(n * 4) + packed);
        if (e == hipSuccess && pooled) e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess) {
            const char *in = (const char *)buf + sizeof(h);
            e = hipMemcpy(dev, in, n * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(dev + align_up(n * 4), in + n * 4, packed, hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                launch_brick_pack(s->self.dP, (const uint32_t *)dev, (unsigned)n, (int)h.value_bytes,
                                  (int)h.weight_bytes, (int)h.color_bytes, (uint8_t *)dev + align_up(n * 4), 1,
                                  s->stream);
                launch_occupancy_rebuild(s->self.dP, s->L.nbricks, s->d.tsdf_type, s->d.weight_type, s->L.sdim[2], s->stream);
                e = hipStreamSynchronize(s->stream);
            }
            if (pooled) {
                hipFreeAsync(dev, s->stream);
            } else {
                hipFree(dev);
            }
        }
    }
    if (e == hipSuccess) {
        const int32_t frames = h.frames;
        e = hipMemcpy((char *)s->st + offsetof(SessState, frames), &frames, sizeof(frames), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        vs_session_destroy(s);
        return fail(VS_EHIP, "restore: %s", hipGetErrorString(e));
    }
    // the model maps the next frame tracks against: render them at the saved pose
    for (int l = 0; l < s->d.levels && r == VS_OK; ++l) r = vs_raycast(s, l, nullptr);
    if (r == VS_OK) r = sync_last(s);
    if (r != VS_OK) {
        vs_session_destroy(s);
        return r;
    }
    *out = s;
    return VS_OK;
}
}  // namespace

extern "C" {

vs_status vs_session_restore(vs_ctx *ctx, const void *buf, size_t bytes, vs_session **out) {
    return restore_session(ctx, buf, bytes, false, out);
}

vs_status vs_server_evict(vs_server *v, int slot, void *buf, size_t capacity, size_t *bytes) {
    if (!v) return fail(VS_EINVAL, "null server");
    if (slot < 0 || slot >= (int)v->slot.size() || !v->slot[slot]) return fail(VS_EINVAL, "slot %d not open", slot);
    vs_status r = server_quiesce(v);
    if (r == VS_OK) r = vs_session_snapshot(v->slot[slot], buf, capacity, bytes);
    if (r != VS_OK || !buf) return r;
    return vs_server_close(v, slot);
}

vs_status vs_server_open_snapshot(vs_server *v, const void *buf, size_t bytes, int *slot) {
    if (!v || !slot) return fail(VS_EINVAL, "null argument");
    vs_snapshot_header h;
    vs_status r = snapshot_check(buf, bytes, h);
    if (r != VS_OK) return r;
    int i = -1;
    r = server_admit(v, &h.desc, &i);
    if (r != VS_OK) return r;
    vs_session *s = nullptr;
    r = restore_session(v->ctx, buf, bytes, true, &s);
    if (r != VS_OK) return r;
    r = server_attach(v, s, i);
    if (r != VS_OK) return r;
    *slot = i;
    return VS_OK;
}

vs_status vs_volume_download(vs_session *s, float *tsdf, float *weight, uint8_t *brick_mask) {
    if (!s) return fail(VS_EINVAL, "null session");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t N = s->L.nvox;
    if (tsdf || weight) {
        float *tmp = nullptr;
        HIP_TRY(hipMalloc((void **)&tmp, 2 * N * sizeof(float)));
        launch_volume_to_linear(s->self.dP, N, s->d.tsdf_type, s->d.weight_type, tmp, tmp + N, s->stream);
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && tsdf) e = hipMemcpy(tsdf, tmp, N * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && weight) e = hipMemcpy(weight, tmp + N, N * 4, hipMemcpyDeviceToHost);
        hipFree(tmp);
        if (e != hipSuccess) return fail(VS_EHIP, "volume download: %s", hipGetErrorString(e));
    }
    if (brick_mask) {
        std::vector<uint32_t> bits(s->L.nwords);
        HIP_TRY(hipMemcpy(bits.data(), s->hp.bvis, bits.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t b = 0; b < s->L.nbricks; ++b) brick_mask[b] = (bits[b >> 5] >> (b & 31)) & 1u;
    }
    return VS_OK;
}

vs_status vs_volume_upload(vs_session *s, const float *tsdf, const float *weight) {
    if (!s || !tsdf || !weight) return fail(VS_EINVAL, "null argument");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const size_t N = s->L.nvox;
    float *tmp = nullptr;
    HIP_TRY(hipMalloc((void **)&tmp, 2 * N * sizeof(float)));
    hipError_t e = hipMemcpy(tmp, tsdf, N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(tmp + N, weight, N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_volume_from_linear(s->self.dP, N, s->d.tsdf_type, s->d.weight_type, tmp, tmp + N, s->stream);
        launch_occupancy_rebuild(s->self.dP, s->L.nbricks, s->d.tsdf_type, s->d.weight_type, s->L.sdim[2], s->stream);
        e = hipStreamSynchronize(s->stream);
    }
    hipFree(tmp);
    if (e != hipSuccess) return fail(VS_EHIP, "volume upload: %s", hipGetErrorString(e));
    return VS_OK;
}

vs_status vs_icp_system(vs_session *s, int level, const float T_21[16], double ATA[36], double ATb[6],
                        long long *valid, uint8_t *mask_host) {
    if (!s || !T_21) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    const int nt = s->L.tiles[level];
    double *part = (double *)(s->mem + s->L.diag_part);
    uint8_t *mask = mask_host ? (uint8_t *)(s->mem + s->L.diag_mask) : nullptr;
    launch_icp_system(s->self.dP, s->self.dWork[level], nt, level, T_21, part, mask, s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<double> h((size_t)nt * kIcpVals);
    HIP_TRY(hipMemcpy(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost));
    double tot[kIcpVals] = {0};
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < kIcpVals; ++k) tot[k] += h[(size_t)t * kIcpVals + k];
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            if (ATA) ATA[i * 6 + j] = ATA[j * 6 + i] = tot[k];
            ++k;
        }
    for (int i = 0; i < 6; ++i)
        if (ATb) ATb[i] = tot[k + i];
    if (valid) *valid = (long long)llround(tot[27]);
    if (mask_host) {
        const size_t n = (size_t)(s->d.width >> level) * (s->d.height >> level);
        HIP_TRY(hipMemcpy(mask_host, mask, n, hipMemcpyDeviceToHost));
    }
    return VS_OK;
}

vs_status vs_icp_system_noalloc(vs_session *s, int level, const float T_21[16], double *sums_dev,
                                uint8_t *mask_dev, vs_stream stream) {
    if (!s || !T_21 || !sums_dev) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    const int nt = s->L.tiles[level];
    if (nt > kNoAllocMaxTiles)
        return fail(VS_EINVAL, "level %d has %d tiles of 32x8 pixels; vs_icp_system_noalloc holds at most %d", level,
                    nt, kNoAllocMaxTiles);
    hipStream_t h = pick(s, stream);
    launch_icp_system(s->self.dP, s->self.dWork[level], nt, level, T_21, s->ctx->scratch, mask_dev, h);
    launch_icp_sum_tiles(s->ctx->scratch, nt, sums_dev, h);
    HIP_TRY(hipGetLastError());
    return VS_OK;
}

vs_status vs_session_set_timing(vs_session *s, int enable) {
    if (!s) return fail(VS_EINVAL, "null session");
    s->timing = enable != 0;
    return VS_OK;
}

vs_status vs_get_stage_ms(vs_session *s, float ms[5]) {
    if (!s || !ms) return fail(VS_EINVAL, "null argument");
    if (!s->timing || !s->timer.used) return fail(VS_EINVAL, "timing not enabled");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    float k[kMarks - 1];
    for (int i = 0; i < kMarks - 1; ++i) HIP_TRY(hipEventElapsedTime(&k[i], s->timer.ev[i], s->timer.ev[i + 1]));
    ms[0] = k[0];                 // bilateral (with the tracker reset)
    ms[1] = k[1];                 // pyramid
    ms[2] = k[2];                 // track
    ms[3] = k[3] + k[4];          // integrate (visibility + update)
    ms[4] = k[5] + k[6] + k[7];   // raycast, all levels
    return VS_OK;
}

vs_status vs_ctx_set_tuning(vs_ctx *ctx, int mode, const char *cache_path) {
    if (!ctx) return fail(VS_EINVAL, "null context");
    if (mode != VS_TUNE_OFF && mode != VS_TUNE_ONLINE && mode != VS_TUNE_CACHED)
        return fail(VS_EINVAL, "tuning mode %d", mode);
    ctx->tune_mode = mode;
    ctx->tune_path = cache_path ? cache_path : "";
    if (!ctx->tune_path.empty()) load_tuning(ctx);
    return VS_OK;
}

vs_status vs_ctx_tuning(vs_ctx *ctx, char *buf, size_t len, size_t *needed) {
    if (!ctx) return fail(VS_EINVAL, "null context");
    std::string out;
    for (const auto &kv : ctx->tuned) {
        const Shape &h = kv.second;
        char line[160];
        snprintf(line, sizeof(line), "%s|%s|%s|%d|", kv.first.c_str(), kRayBlockName[h.march],
                 kShadeOpts[h.shade].name, kIntegGrids[h.integ]);
        out += line;
        auto m = ctx->tune_ms.find(kv.first);
        out += m != ctx->tune_ms.end() ? m->second : std::string("cached");
        out += "\n";
    }
    if (needed) *needed = out.size() + 1;
    if (buf && len) {
        const size_t n = std::min(len - 1, out.size());
        memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return VS_OK;
}

vs_status vs_ctx_set_timing(vs_ctx *ctx, int enable) {
    if (!ctx) return fail(VS_EINVAL, "null context");
    ctx->timing = enable != 0;
    return VS_OK;
}

vs_status vs_ctx_stage_ms(vs_ctx *ctx, float ms[8]) {
    if (!ctx || !ms) return fail(VS_EINVAL, "null argument");
    if (!ctx->timing || !ctx->timer.used) return fail(VS_EINVAL, "timing not enabled");
    HIP_TRY(hipEventSynchronize(ctx->timer.ev[kMarks - 1]));
    for (int i = 0; i < kMarks - 1; ++i) HIP_TRY(hipEventElapsedTime(&ms[i], ctx->timer.ev[i], ctx->timer.ev[i + 1]));
    return VS_OK;
}

vs_status vs_get_integrate_counts(vs_session *s, long long *visible_bricks, long long *updated_voxels) {
    if (!s) return fail(VS_EINVAL, "null session");
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    SessState st;
    HIP_TRY(hipMemcpy(&st, s->st, sizeof(st), hipMemcpyDeviceToHost));
    if (visible_bricks) *visible_bricks = st.vis_count;
    if (updated_voxels) *updated_voxels = (long long)st.upd_count;
    return VS_OK;
}

// Tap statistics of one instrumented raycast (layout in vscan.h).
vs_status vs_raycast_counts(vs_session *s, int level, unsigned long long counts[8]) {
    if (!s || !counts) return fail(VS_EINVAL, "null argument");
    if (level < 0 || level >= s->d.levels) return fail(VS_EINVAL, "level %d", level);
    vs_status r = sync_last(s);
    if (r != VS_OK) return r;
    unsigned long long *d = (unsigned long long *)(s->mem + s->L.stats);
    HIP_TRY(hipMemset(d, 0, 64));
    const Batch &B = s->self;
    launch_fuse_ok(B.dP, 1, s->stream);
    launch_raycast_stats(B.dP, B.dWork[level], B.nWork[level], level, B.tt, B.wt, d, s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(counts, d, 64, hipMemcpyDeviceToHost));
    return VS_OK;
}

// ---------------------------------------------------------------------------
// serving: slot table
// ---------------------------------------------------------------------------
vs_status vs_server_create(vs_ctx *ctx, const vs_server_desc *desc, vs_server **out) {
    if (!ctx || !desc || !out) return fail(VS_EINVAL, "null argument");
    if (desc->max_slots < 1 || desc->max_slots > kMaxSlots)
        return fail(VS_EINVAL, "max_slots %d outside 1..%d", desc->max_slots, kMaxSlots);
    if (desc->tsdf_type < 0 || desc->tsdf_type > 1 || desc->weight_type < 0 || desc->weight_type > 1)
        return fail(VS_EINVAL, "storage type");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<vs_server> v(new vs_server());
    v->ctx = ctx;
    v->d = *desc;
    const int n = desc->max_slots;
    v->slot.assign(n, nullptr);
    const size_t offP = 0, offG = align_up(n * sizeof(SessParams)), offA = offG + align_up(n * sizeof(SessState));
    const size_t total = offA + align_up(n * sizeof(int));
    hipError_t e = hipMalloc((void **)&v->fixed, total);
    if (e != hipSuccess) return fail(VS_ENOMEM, "server tables (%zu bytes): %s", total, hipGetErrorString(e));
    HIP_TRY(hipMemset(v->fixed, 0, total));
    v->B.n = n;
    v->B.tt = desc->tsdf_type;
    v->B.wt = desc->weight_type;
    v->B.dP = (SessParams *)(v->fixed + offP);
    v->ghost = (SessState *)(v->fixed + offG);
    v->active = (int *)(v->fixed + offA);
    v->B.active = v->active;
    HIP_TRY(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&v->order, hipEventDisableTiming));
    v->last = v->stream;
    for (int g = 0; g < 2; ++g) HIP_TRY(hipEventCreateWithFlags(&v->lstage_ev[g], hipEventDisableTiming));
    v->pstage_n = 2 * n + 8;
    e = hipHostMalloc((void **)&v->pstage, v->pstage_n * sizeof(SessParams), hipHostMallocDefault);
    if (e != hipSuccess) return fail(VS_ENOMEM, "server parameter staging: %s", hipGetErrorString(e));
    v->pstage_ev = new hipEvent_t[v->pstage_n];
    for (int k = 0; k < v->pstage_n; ++k) HIP_TRY(hipEventCreateWithFlags(&v->pstage_ev[k], hipEventDisableTiming));
    for (int i = 0; i < n; ++i) {
        vs_status r = server_write_slot(v.get(), i);
        if (r != VS_OK) {
            hipFree(v->fixed);
            return r;
        }
    }
    HIP_TRY(hipStreamSynchronize(v->stream));
    ctx->servers.push_back(v.get());
    *out = v.release();
    return VS_OK;
}

vs_status vs_server_destroy(vs_server *v) {
    if (!v) return VS_OK;
    hipSetDevice(v->ctx->device);
    (void)server_quiesce(v);
    for (int i = 0; i < (int)v->slot.size(); ++i)
        if (v->slot[i]) vs_session_destroy(v->slot[i]);
    for (int g = 0; g < 2; ++g)
        if (v->lists[g]) hipFreeAsync(v->lists[g], server_order(v));
    (void)server_quiesce(v);
    server_reap(v, true);
    release_stage(v->B.hs);
    free_batch_graph(v->B);
    for (int g = 0; g < 2; ++g) {
        if (v->lstage[g]) hipHostFree(v->lstage[g]);
        if (v->lstage_ev[g]) hipEventDestroy(v->lstage_ev[g]);
    }
    if (v->pstage) hipHostFree(v->pstage);
    if (v->pstage_ev) {
        for (int k = 0; k < v->pstage_n; ++k)
            if (v->pstage_ev[k]) hipEventDestroy(v->pstage_ev[k]);
        delete[] v->pstage_ev;
    }
    if (v->fixed) hipFree(v->fixed);
    if (v->order) hipEventDestroy(v->order);
    if (v->stream) hipStreamDestroy(v->stream);
    auto &vv = v->ctx->servers;
    vv.erase(std::remove(vv.begin(), vv.end(), v), vv.end());
    delete v;
    return VS_OK;
}

vs_status vs_server_open(vs_server *v, const vs_session_desc *desc, const double T_gc0[16], int *slot) {
    if (!v || !desc || !slot) return fail(VS_EINVAL, "null argument");
    int i = -1;
    vs_status r = server_admit(v, desc, &i);
    if (r != VS_OK) return r;
    vs_session *s = nullptr;
    r = create_session(v->ctx, desc, T_gc0, true, &s);
    if (r != VS_OK) return r;
    r = server_attach(v, s, i);
    if (r != VS_OK) return r;
    *slot = i;
    return VS_OK;
}

vs_status vs_server_close(vs_server *v, int slot) {
    if (!v) return fail(VS_EINVAL, "null server");
    if (slot < 0 || slot >= (int)v->slot.size() || !v->slot[slot]) return fail(VS_EINVAL, "slot %d not open", slot);
    return vs_session_destroy(v->slot[slot]);
}

vs_session *vs_server_session(vs_server *v, int slot) {
    if (!v || slot < 0 || slot >= (int)v->slot.size()) return nullptr;
    return v->slot[slot];
}

vs_status vs_server_step(vs_server *v, const int *slots, const uint16_t *const *depths, int n, int is_host,
                         vs_stream stream) {
    if (!v || !slots || n <= 0) return fail(VS_EINVAL, "bad step arguments");
    const int ns = (int)v->slot.size();
    if (n > ns) return fail(VS_EINVAL, "%d frames for %d slots", n, ns);
    uint32_t bits[kMaxSlots / 32] = {};
    for (int k = 0; k < n; ++k) {
        const int i = slots[k];
        if (i < 0 || i >= ns || !v->slot[i]) return fail(VS_EINVAL, "slot %d not open", i);
        if ((bits[i >> 5] >> (i & 31)) & 1u) return fail(VS_EINVAL, "slot %d repeated in step", i);
        bits[i >> 5] |= 1u << (i & 31);
    }
    HIP_TRY(hipSetDevice(v->ctx->device));
    server_reap(v, false);
    if (v->dirty) {
        vs_status r = server_rebuild(v);
        if (r != VS_OK) return r;
    }
    hipStream_t h = stream ? (hipStream_t)stream : v->stream;
    if (h != v->last) {
        // steps share the slot table's lists: order this one after the previous
        HIP_TRY(hipEventRecord(v->order, v->last));
        HIP_TRY(hipStreamWaitEvent(h, v->order, 0));
    }
    if (depths && is_host) {
        std::vector<vs_session *> ss(n);
        for (int k = 0; k < n; ++k) ss[k] = v->slot[slots[k]];
        vs_status r = stage_host_frames(v->B.hs, v->B.dP, ss.data(), slots, depths, n, h);
        if (r != VS_OK) return r;
    } else if (depths) {
        DepthChunk ch{};
        for (int k = 0; k < n; ++k) {
            if (!depths[k]) continue;
            ch.src[ch.n] = depths[k];
            ch.slot[ch.n] = slots[k];
            if (++ch.n == kDepthChunk) {
                launch_gather_depth(v->B.dP, ch, h);
                ch.n = 0;
            }
        }
        launch_gather_depth(v->B.dP, ch, h);
    }
    launch_set_active(v->active, ns, bits, h);
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < n; ++k) v->slot[slots[k]]->last = h;
    v->last = h;
    return run_step(v->ctx, v->B, h, v->ctx->timing ? &v->ctx->timer : nullptr);
}

vs_status vs_server_usage(vs_server *v, int *open_slots, unsigned long long *bytes) {
    if (!v) return fail(VS_EINVAL, "null server");
    if (open_slots) {
        int c = 0;
        for (vs_session *s : v->slot) c += s != nullptr;
        *open_slots = c;
    }
    if (bytes) *bytes = v->bytes;
    return VS_OK;
}

}  // extern "C"
