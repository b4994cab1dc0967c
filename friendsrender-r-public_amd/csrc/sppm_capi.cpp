// sppm_capi.cpp -- C ABI of the SPPM pass core (include/sppm_amd.h).
//
// Host side of the drop-in boundary.  Scene setup mirrors Scene::CreateScene
// (Scene.cpp:13-218) after the OBJ has been read: facet records, light
// registration in facet order (Scene.cpp:145-154), textured-light tables
// (LightSampler.cpp:30-66) and the light pick table (LightSampler.cpp:6-28)
// are built here; the closest-hit structure is built on the GPU.  All per-pass
// work runs in hand-written HIP kernels; this file only allocates, uploads and
// launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sppm_amd.h"
#include "sppm_guide.h"
#include "sppm_host_math.h"
#include "sppm_kernels.h"
#include "sppm_mtjump.h"
#include "sppm_pass.h"
#include "sppm_stripes.h"
#include "sppm_wave.h"

using namespace sppm;
namespace H = sppm_host;

static thread_local std::string g_err;
static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define HIPOK(x)                                                                                     \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(SPPM_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_));  \
    } while (0)

template <typename T>
static int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * n);
    if (e != hipSuccess) return fail(SPPM_E_NOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * n, hipGetErrorString(e));
    return 0;
}
template <typename T>
static void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

struct SceneBuffers {
    bool ready = false;
    BvhResult bvh;
    float* facet_nrm = nullptr;
    float2* facet_uv = nullptr;
    int* facet_attr = nullptr;
    DevAttr* attrs = nullptr;
    TexDesc* tex = nullptr;
    float4* texels = nullptr;
    float* lpick_prob = nullptr;
    int* lpick_left = nullptr;
    int* lpick_right = nullptr;
    DevLight* lights = nullptr;
    float4* ltab_uvw = nullptr;
    float4* ltab_L = nullptr;
    float* ltab_prob = nullptr;
    int* ltab_left = nullptr;
    int* ltab_right = nullptr;
    int* facet_light = nullptr;
    SceneDev dev{};
    float bmin[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
    int stack_cap = 0;
    uint32_t* spill = nullptr;  // traversal stack entries beyond TRACE_LDS
    uint32_t* spill_eye = nullptr;  // the same for the eye pass, which runs beside the photon pass
    int spill_blocks = 0;
    int* next_path = nullptr;   // striped path counters of the photon pass, then pixel counters of the eye pass
    std::vector<DevLight> hlights;  // host copy of the area lights
    std::vector<float4> h_tuvw, h_tL;  // host copy of the textured lights' 1000-point tables
    float h_lpick_avg = 1.0f;           // TableSampler average of the light pick table
    TraceSpill trace_spill() const { return TraceSpill{spill, spill_blocks, stack_cap}; }
    TraceSpill trace_spill_eye() const { return TraceSpill{spill_eye, spill_blocks, stack_cap}; }
    void release() {
        free_bvh(&bvh);
        dfree(spill); dfree(spill_eye); dfree(next_path);
        dfree(facet_nrm); dfree(facet_uv); dfree(facet_attr); dfree(attrs); dfree(tex); dfree(texels);
        dfree(lpick_prob); dfree(lpick_left); dfree(lpick_right); dfree(lights);
        dfree(ltab_uvw); dfree(ltab_L); dfree(ltab_prob); dfree(ltab_left); dfree(ltab_right); dfree(facet_light);
        ready = false;
    }
};

struct RenderBuffers {
    bool ready = false;
    int N = 0, W = 0, H = 0, np = 0;
    SlotsDev slots{};
    int* off = nullptr;
    int* nstored = nullptr;   // device scalar
    int* scan_tmp = nullptr;
    int cell_cap = 0;
    int4* cell_gaps = nullptr;  // long runs of empty cells of the grid build
    int* cell_ngap = nullptr;
    int* cell_start = nullptr;
    PhotonSortBuffers psort{};  // stable sort of the flattened photons by cell
    StagedPhotons staged{};     // the pass's flattened photons, out of the slots
    float4 *gpos = nullptr, *gflux = nullptr, *gdir = nullptr;
    int photon_cap = 0;
    GridParams* gp = nullptr;
    HitDev hits{};       // the hit points of the current pass: hitbuf[pass % 2]
    // Pass k's eye pass writes buffer k % 2, so in passes in a row it waits
    // only for the gather two passes back, not for the previous one.
    HitDev hitbuf[2]{};
    SharedDev shared{};
    int* counts = nullptr;
    int* list_off = nullptr;
    int* list_scan_tmp = nullptr;
    int* lists = nullptr;
    long long lists_cap = 0;
    uint32_t* words = nullptr;
    uint32_t* cam_states = nullptr;  // cam_K windows of 624 raw words (camera stream segments)
    uint32_t* cam_phi = nullptr;     // jump polynomial over the other cam_K - 1 segments
    int cam_K = 0;
    long long cam_L = 0;
    unsigned long long* rstat = nullptr;  // (radius sum in 2^-32 units, hit points) of the last gather
    TileRange* tiles = nullptr;
    int ntiles = 0;
    int* chunkp = nullptr;
    int* nch = nullptr;
    int* chunk_off = nullptr;
    int* tile_scan_tmp = nullptr;
    float4* partial = nullptr;
    GroupBuffers groups{};
    int* pass_dev = nullptr;   // passes done (device copy of the pass counter, read by graph replays);
                               // [1], [2]: the same counted on the photon and eye streams of a replay
    int* path_cell = nullptr;  // guided emission: emission cell per path, 2 x N (block = step parity)
    float* path_ratio = nullptr;  // guided emission: q / p_plain per path, like path_cell
    unsigned long long* pdep = nullptr;  // guided emission: deposited luminance per flattened photon (guide_fix)
    // guided emission: this pass's photon offsets and stored count (N + 1),
    // copied by the build for the gather's deposits and the training records,
    // which run while the next pass stages its photons into off / nstored
    int* doff = nullptr;
    uint32_t* flags = nullptr;
    Counters* ctr = nullptr;
    // closest-hit ray and tie counts, striped (STRIPES lines of RAY_STRIPE
    // words: photon rays, eye rays, photon ties, eye ties); readers add the
    // lines' sums to the Counters record
    unsigned long long* raystripes = nullptr;
    float4* image = nullptr;
    bool have_lists = false;
    // record_neighbours 2: (hit point, photon) pairs appended by the production gather
    int2* pairs = nullptr;
    unsigned long long* npairs = nullptr;
    long long pairs_cap = 0;
    // wavefront photon pass
    WaveBuffers wave{};
    int wave_bounces = 0;      // 0: the megakernel photon pass
    int wave_lanes = 0;        // lanes per ray of the trace kernel
    int wave_blocks = 0;       // trace kernel grid
    void release() {
        dfree(pairs); dfree(npairs);
        dfree(wave.path.rng); dfree(wave.path.org); dfree(wave.path.dir); dfree(wave.path.ndir);
        dfree(wave.path.rad); dfree(wave.path.thr); dfree(wave.rays[0]); dfree(wave.rays[1]); dfree(wave.hits);
        dfree(wave.qctl); dfree(wave.spill);
        wave_bounces = 0;
        pairs_cap = 0;
        dfree(slots.pos); dfree(slots.flux); dfree(slots.dir); dfree(slots.valid); dfree(slots.count);
        dfree(off); dfree(nstored); dfree(scan_tmp); dfree(cell_gaps); dfree(cell_ngap); dfree(cell_start);
        dfree(psort.keys); dfree(psort.vals); dfree(psort.tmp_k); dfree(psort.tmp_v); dfree(psort.hist);
        dfree(staged.rec);
        dfree(gpos); dfree(gflux); dfree(gdir); dfree(gp);
        for (HitDev& h : hitbuf) {
            dfree(h.pos); dfree(h.nrm); dfree(h.out); dfree(h.weight); dfree(h.mat); dfree(h.absorb); dfree(h.direct);
        }
        hits = HitDev{};
        dfree(shared.rad); dfree(shared.cnt); dfree(shared.direct); dfree(counts); dfree(list_off); dfree(list_scan_tmp); dfree(lists);
        dfree(words); dfree(cam_states); dfree(cam_phi); dfree(rstat); dfree(tiles); dfree(chunkp); dfree(nch); dfree(chunk_off); dfree(tile_scan_tmp); dfree(partial); dfree(groups.keys); dfree(groups.vals); dfree(groups.tmp_k); dfree(groups.tmp_v); dfree(groups.hist); dfree(groups.nvalid); dfree(groups.chunkq); dfree(groups.stat); dfree(pass_dev); dfree(path_cell); dfree(path_ratio); dfree(pdep); dfree(doff); dfree(flags); dfree(ctr); dfree(raystripes); dfree(image);
        ready = false;
    }
};

// Per-kernel timing: an event pair per launch from a pool; elapsed times are
// resolved lazily (no synchronisation inside a pass).
struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool, pending;
    std::pair<hipEvent_t, hipEvent_t> cur{nullptr, nullptr};
    double total_ms = 0.0;
    long long n = 0;
    void start(hipStream_t st) {
        if (pool.empty()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
            pool.push_back({a, b});
        }
        cur = pool.back();
        pool.pop_back();
        (void)hipEventRecord(cur.first, st);
    }
    void stop(hipStream_t st) {
        if (!cur.first) return;
        (void)hipEventRecord(cur.second, st);
        pending.push_back(cur);
        cur = {nullptr, nullptr};
    }
    void resolve() {
        for (auto& e : pending) {
            float ms = 0.0f;
            if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
                total_ms += ms;
                n += 1;
            }
            pool.push_back(e);
        }
        pending.clear();
    }
    void reset() {
        resolve();
        total_ms = 0.0;
        n = 0;
    }
    void destroy() {
        resolve();
        for (auto& e : pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        pool.clear();
    }
    double avg() const { return n ? total_ms / n : 0.0; }
};

// Several contexts' passes in one graph: one branch per context (two streams
// each), forked from and joined into the batch's origin stream.  `done` is
// recorded after every launch; member contexts wait for it before reading or
// freeing what the graph uses.
// One multi-scene launch: jobs [first, first + count) of its table, block
// offsets at bstart + boff, `grid` blocks; key = BVH width * 2 + flag.
struct BatchLaunch {
    int first, count, key, boff, grid;
};
struct sppm_batch {
    int device = 0;
    int mode = SPPM_BATCH_AUTO;
    std::vector<sppm_ctx*> ctx;
    std::vector<hipStream_t> s1, s2;
    std::vector<hipEvent_t> join;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool launched = false;
    bool broken = false;  // a member was destroyed: the graph is gone
    // multi-scene mode: the launches, their job tables and block offsets,
    // the camera segments' offsets and the path kernels' spill area
    std::vector<BatchLaunch> pl, el, gl;  // photon, eye and gather launches
    int cap = 0, nseg = 0;
    PhotonJob* pjobs = nullptr;
    EyeJob* ejobs = nullptr;
    GatherJob* gjobs = nullptr;
    CamJob* cjobs = nullptr;
    int* bstart = nullptr;
    int* kstart = nullptr;
    uint32_t* spill = nullptr;
};

struct sppm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;
    // Passes in a row run their photon passes on stream3: a plain pass's
    // photon pass follows the previous pass's staging (ev_staged: its photons
    // copied out of the slots), and then stages its own photons there once the
    // previous build has read the last staged ones (ev_built); a guided
    // pass's photon pass waits for the previous build (which publishes the
    // proposal it samples) and is staged by its own build.  Either way it runs
    // beside the previous pass's gather.  lead_ok says that nothing but
    // sppm_pass has queued work since.
    hipStream_t stream3 = nullptr;
    hipEvent_t ev_built = nullptr, ev_photons = nullptr, ev_staged = nullptr, ev_published = nullptr;
    hipEvent_t ev_gathered[2] = {nullptr, nullptr};  // the gather of the last pass of each parity (its hit points)
    bool lead_ok = false;
    bool own_stream = false;
    // Camera sample words (UniformSample stream) are generated on stream2 one
    // pass ahead into two buffers: block k (the words of the k-th eye pass)
    // lands in buffer k % 2 and is overwritten only after eye pass k has run.
    hipEvent_t ev_words[2] = {nullptr, nullptr}, ev_eye[2] = {nullptr, nullptr};
    hipEvent_t ev_gfork = nullptr, ev_gjoin = nullptr;  // camera-word branch inside a captured pass
    // Captured passes (sppm_graph_capture), replayed as the eager passes run:
    // the photon pass on stream3 beside the previous gather, the eye pass on
    // stream2 (word and hit-point buffer b = pass % 2), then there the next
    // pass's camera words (into buffer 1 - b, beside the gather), build and
    // gather on the main stream.  graph / graph_exec is the gather of even
    // passes (hit-point buffer 0), graph_odd / graph_exec_odd of odd ones;
    // graph_passes counts the replays.
    hipGraph_t graph = nullptr, graph_odd = nullptr;
    hipGraphExec_t graph_exec = nullptr, graph_exec_odd = nullptr;
    // The photons are staged in a graph of their own (g_stage / x_stage) on
    // stream3 after the photon pass; guided passes publish the proposal in
    // one (g_publish / x_publish) on the main stream before the build.
    hipGraph_t g_photon = nullptr, g_stage = nullptr, g_publish = nullptr, g_build = nullptr,
               g_eye[2] = {nullptr, nullptr}, g_words[2] = {nullptr, nullptr};
    hipGraphExec_t x_photon = nullptr, x_stage = nullptr, x_publish = nullptr, x_build = nullptr,
                   x_eye[2] = {nullptr, nullptr}, x_words[2] = {nullptr, nullptr};
    long long graph_passes = 0;
    // The batch whose graph holds this context's buffers (sppm_batch_add), and
    // whether it has launched since: readbacks and synchronisation wait for
    // the batch's last launch, setup changes are refused while it holds them.
    sppm_batch* batch = nullptr;
    // Hash of the scene, camera, parameters and emission setup the render
    // was begun with (written into checkpoints, checked when loading one).
    uint64_t setup_hash = 0, scene_hash = 0;
    int tuned_blocks = 0;                                // path-kernel grid chosen by the autotuner (0: default)
    bool eye_recorded[2] = {false, false};
    long long words_gen = 0, words_used = 0;
    SceneBuffers scene;
    RenderBuffers rb;
    sppm_camera_desc cam{};
    bool have_cam = false;
    CamDev camdev{};
    float* dia_prob = nullptr;  // thin-lens diaphragm table (CamDev::dia_*)
    float4* dia_uv = nullptr;
    sppm_params params{};
    bool have_params = false;
    int pass = 0;
    int split_step = 0;  // step of the last sppm_pass_photons (its gather's training records)
    Timer t_photons, t_build, t_eye, t_gather, t_guide;
    // learned emission proposal
    GuideBuffers guide;
    bool guided = false;
    sppm_guide_params gparams{};
    GuideEmit emit{};
};

// Waits for the last launch of the batch whose graph uses c's buffers.
static int batch_wait(sppm_ctx* c) {
    if (c->batch && c->batch->launched) HIPOK(hipEventSynchronize(c->batch->done));
    return SPPM_OK;
}
// Setup changes and the context's own passes are refused while a batch
// holds its buffers (the batch graph would replay them, or freed memory).
static int not_in_batch(sppm_ctx* c, const char* what) {
    if (c->batch) return fail(SPPM_E_STATE, "%s: the context belongs to a batch (destroy the batch first)", what);
    return SPPM_OK;
}
// A captured pass holds the scene, render and proposal buffers by value.
static void drop_graph(sppm_ctx* c) {
    for (hipGraphExec_t* x :
         {&c->graph_exec, &c->graph_exec_odd, &c->x_photon, &c->x_stage, &c->x_publish, &c->x_build, &c->x_eye[0],
          &c->x_eye[1], &c->x_words[0], &c->x_words[1]}) {
        if (*x) (void)hipGraphExecDestroy(*x);
        *x = nullptr;
    }
    for (hipGraph_t* g : {&c->graph, &c->graph_odd, &c->g_photon, &c->g_stage, &c->g_publish, &c->g_build,
                          &c->g_eye[0], &c->g_eye[1], &c->g_words[0], &c->g_words[1]}) {
        if (*g) (void)hipGraphDestroy(*g);
        *g = nullptr;
    }
    c->graph_passes = 0;
}
// FNV-1a over bytes (setup fingerprints of checkpoints).
static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
static const uint64_t FNV_BASIS = 0xcbf29ce484222325ull;

// ---- the fixed staging buffer of sppm_trace_closest_static ----------------
// One array per GPU, `scratch`, holding the rays and hits of one call:
// origins, directions and t/u/v (3 floats per ray each), facet ids, then the
// flag word.  In host code the name `scratch` only stands in for the device
// array; the device address of each GPU's copy comes from
// hipGetSymbolAddress, resolved once per device (sppm_create), and host code
// uses only that address.
constexpr size_t SCRATCH_RAYS = SPPM_TRACE_STATIC_MAX_RAYS;
constexpr size_t SCRATCH_FLAG_OFF = (SCRATCH_RAYS * 40 + 255) & ~(size_t)255;
constexpr size_t SCRATCH_BYTES = SCRATCH_FLAG_OFF + 256;
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_dev[SCRATCH_DEVICES] = {};

// Device address of `scratch` on `device`, which must be current.
static int scratch_resolve(int device, char** out) {
    if (device < 0 || device >= SCRATCH_DEVICES) return fail(SPPM_E_ARG, "device %d beyond %d", device, SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SPPM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_dev[device] = p;
    }
    if (out) *out = static_cast<char*>(g_scratch_dev[device]);
    return SPPM_OK;
}

// Closest hits of n host rays through device buffers the caller provides:
// org, dir 3n floats, fac n ints, tuv 3n floats (t, u, v blocks), fl 1 word.
// Results go straight into the caller's host arrays (t, u, v may be NULL).
static int trace_rays(sppm_ctx* c, int n, const float* org, const float* dir, float tmin, float tmax, int32_t* facet,
                      float* t, float* u, float* v, float* dorg, float* ddir, int* dfac, float* dtuv, uint32_t* dfl) {
    hipStream_t st = c->stream;
    HIPOK(hipMemcpyAsync(dorg, org, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(ddir, dir, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
    launch_set_u32(dfl, 0u, st);
    float *dt = dtuv, *du = dtuv + n, *dv = dtuv + 2 * (size_t)n;
    launch_trace(c->scene.dev, n, dorg, ddir, tmin, tmax, c->scene.trace_spill(), dfac, t ? dt : nullptr,
                 u ? du : nullptr, v ? dv : nullptr, dfl, st);
    HIPOK(hipGetLastError());
    uint32_t fl = 0;
    HIPOK(hipMemcpyAsync(facet, dfac, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    if (t) HIPOK(hipMemcpyAsync(t, dt, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (u) HIPOK(hipMemcpyAsync(u, du, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (v) HIPOK(hipMemcpyAsync(v, dv, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(&fl, dfl, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    if (fl & 4u) return fail(SPPM_E_OVERFLOW, "traversal stack overflow");
    return SPPM_OK;
}

extern "C" {

int sppm_abi_version(void) { return SPPM_ABI_VERSION; }
const char* sppm_last_error(void) { return g_err.c_str(); }

int sppm_create(sppm_ctx** out, int device, void* stream) {
    if (!out) return fail(SPPM_E_ARG, "out is NULL");
    int ndev = 0;
    HIPOK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SPPM_E_ARG, "device %d out of range (%d devices)", device, ndev);
    HIPOK(hipSetDevice(device));
    if (int rc0 = scratch_resolve(device, nullptr)) return rc0;  // the device's staging array, once per device
    sppm_ctx* c = new sppm_ctx();
    c->device = device;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return fail(SPPM_E_HIP, "hipStreamCreate failed");
        }
        c->own_stream = true;
    }
    HIPOK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPOK(hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking));
    HIPOK(hipEventCreateWithFlags(&c->ev_built, hipEventDisableTiming));
    HIPOK(hipEventCreateWithFlags(&c->ev_photons, hipEventDisableTiming));
    HIPOK(hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming));
    HIPOK(hipEventCreateWithFlags(&c->ev_published, hipEventDisableTiming));
    for (hipEvent_t& e : c->ev_gathered) {
        HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPOK(hipEventRecord(e, c->stream));
    }
    for (int b = 0; b < 2; ++b) {
        HIPOK(hipEventCreateWithFlags(&c->ev_words[b], hipEventDisableTiming));
        HIPOK(hipEventCreateWithFlags(&c->ev_eye[b], hipEventDisableTiming));
    }
    HIPOK(hipEventCreateWithFlags(&c->ev_gfork, hipEventDisableTiming));
    HIPOK(hipEventCreateWithFlags(&c->ev_gjoin, hipEventDisableTiming));
    c->params.alpha = 0.9f;
    c->params.first_radius = 0.015f;
    c->params.photons_per_pass = 0;
    c->params.bvh_width = 8;
    *out = c;
    return SPPM_OK;
}

int sppm_destroy(sppm_ctx* c) {
    if (!c) return SPPM_OK;
    (void)hipSetDevice(c->device);
    if (c->batch) {
        // the batch's graph uses this context's buffers: wait for it, then
        // leave the batch, which can no longer launch
        sppm_batch* b = c->batch;
        (void)batch_wait(c);
        for (size_t k = 0; k < b->ctx.size(); ++k)
            if (b->ctx[k] == c) b->ctx[k] = nullptr;
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->broken = true;
        c->batch = nullptr;
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream3);
    c->rb.release();
    c->scene.release();
    for (Timer* t : {&c->t_photons, &c->t_build, &c->t_eye, &c->t_gather, &c->t_guide}) t->destroy();
    guide_free(c->guide);
    dfree(c->dia_prob);
    dfree(c->dia_uv);
    for (int b = 0; b < 2; ++b) {
        (void)hipEventDestroy(c->ev_words[b]);
        (void)hipEventDestroy(c->ev_eye[b]);
    }
    (void)hipEventDestroy(c->ev_gfork);
    (void)hipEventDestroy(c->ev_gjoin);
    (void)hipEventDestroy(c->ev_built);
    (void)hipEventDestroy(c->ev_photons);
    (void)hipEventDestroy(c->ev_staged);
    (void)hipEventDestroy(c->ev_published);
    for (hipEvent_t e : c->ev_gathered) (void)hipEventDestroy(e);
    drop_graph(c);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream3);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPPM_OK;
}

// ---------------------------------------------------------------------------
// Scene upload
// ---------------------------------------------------------------------------
int sppm_scene_upload(sppm_ctx* c, const sppm_scene_desc* d, int bvh_width_flags) {
    if (!c || !d) return fail(SPPM_E_ARG, "NULL argument");
    const bool lbvh = (bvh_width_flags & SPPM_BVH_LBVH) != 0;
    const int bvh_width = bvh_width_flags & ~SPPM_BVH_LBVH;
    if (bvh_width != 2 && bvh_width != 4 && bvh_width != 8) return fail(SPPM_E_ARG, "bvh_width must be 2, 4 or 8");
    const int nf = d->nfacets;
    if (nf <= 0 || !d->pos || !d->nrm || !d->uv || !d->attr) return fail(SPPM_E_SCENE, "empty or incomplete facet arrays");
    if (nf >= (1 << 27)) return fail(SPPM_E_SCENE, "too many facets (%d)", nf);
    if (d->nattrs <= 0 || !d->attrs) return fail(SPPM_E_SCENE, "no attributes");
    for (int f = 0; f < nf; ++f)
        if (d->attr[f] < 0 || d->attr[f] >= d->nattrs)
            return fail(SPPM_E_SCENE, "facet %d has attribute %d outside [0,%d)", f, d->attr[f], d->nattrs);
    for (int a = 0; a < d->nattrs; ++a) {
        const sppm_attr& A = d->attrs[a];
        if (A.kind < SPPM_MAT_DIFFUSE || A.kind > SPPM_MAT_GLOSS_TRANSMIT) return fail(SPPM_E_SCENE, "attribute %d: unknown kind %d", a, A.kind);
        for (int t : {A.tex_albedo, A.tex_roughness, A.tex_bump})
            if (t < -1 || t >= d->ntextures) return fail(SPPM_E_SCENE, "attribute %d: texture id %d out of range", a, t);
    }
    if (d->has_ibl && (d->ibl_texture < 0 || d->ibl_texture >= d->ntextures))
        return fail(SPPM_E_SCENE, "environment map texture id %d out of range", d->ibl_texture);
    for (int t = 0; t < d->ntextures; ++t)
        if (d->textures[t].width <= 0 || d->textures[t].height <= 0 || !d->textures[t].rgba)
            return fail(SPPM_E_SCENE, "texture %d is empty", t);
    {
        int rc0 = not_in_batch(c, "sppm_scene_upload");
        if (rc0) return rc0;
    }
    HIPOK(hipSetDevice(c->device));
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipStreamSynchronize(c->stream2));
    HIPOK(hipStreamSynchronize(c->stream3));
    // the render buffers (photon slots, hit points) and any captured pass
    // belong to the old scene: the render must be begun again
    c->lead_ok = false;
    c->rb.ready = false;
    drop_graph(c);
    c->scene.release();
    SceneBuffers& S = c->scene;
    {
        uint64_t h = FNV_BASIS;
        h = fnv1a(h, &d->nfacets, sizeof d->nfacets);
        h = fnv1a(h, d->pos, sizeof(float) * 9 * (size_t)nf);
        h = fnv1a(h, d->nrm, sizeof(float) * 9 * (size_t)nf);
        h = fnv1a(h, d->uv, sizeof(float) * 9 * (size_t)nf);
        h = fnv1a(h, d->attr, sizeof(int32_t) * (size_t)nf);
        h = fnv1a(h, d->attrs, sizeof(sppm_attr) * (size_t)d->nattrs);
        for (int t = 0; t < d->ntextures; ++t) {
            h = fnv1a(h, &d->textures[t].width, 2 * sizeof(int32_t));
            h = fnv1a(h, d->textures[t].rgba, sizeof(float) * 4 * (size_t)d->textures[t].width * d->textures[t].height);
        }
        h = fnv1a(h, &d->has_ibl, sizeof d->has_ibl);
        h = fnv1a(h, &d->ibl_texture, sizeof d->ibl_texture);
        c->scene_hash = h;  // the BVH width is left out: closest hits do not depend on it
    }
    const hipStream_t st = c->stream;

    // ---- lights, in facet order (Scene.cpp:145-154) ----
    std::vector<DevLight> lights;
    std::vector<float4> tuvw, tL;
    std::vector<float> tprob;
    std::vector<int> tleft, tright;
    for (int f = 0; f < nf; ++f) {
        const sppm_attr& A = d->attrs[d->attr[f]];
        H::V3 rad = H::ld(A.radiance);
        if (!(rad.x + rad.y + rad.z > 0.00001f)) continue;
        const float* P = d->pos + 9 * (size_t)f;
        H::V3 p0 = H::ld(P), p1 = H::ld(P + 3), p2 = H::ld(P + 6);
        float area = 0.5f * H::norm(H::cross(p1 - p0, p2 - p0));
        const float* Nn = d->nrm + 9 * (size_t)f;
        H::V3 nrm = H::normalized((H::ld(Nn) + H::ld(Nn + 3)) + H::ld(Nn + 6));
        DevLight L{};
        L.facet = f;
        L.area = area;
        L.normal[0] = nrm.x; L.normal[1] = nrm.y; L.normal[2] = nrm.z;
        memcpy(L.p, P, sizeof(float) * 9);
        H::V3 flux;
        if (A.tex_albedo >= 0) {  // LightSampler::CreateAreaLight, textured branch
            const sppm_texture& T = d->textures[A.tex_albedo];
            const float* UV = d->uv + 9 * (size_t)f;
            const int NS = 1000;
            std::vector<H::V3> uvw(NS), Lv(NS);
            H::V3 total = H::v3(0, 0, 0);
            for (int i = 0; i < NS; ++i) {
                float w0 = (float)i / (float)NS, w1 = H::halton(2, i), w2 = H::halton(3, i);
                if (w1 > w0) { w0 *= 0.5f; w1 -= w0; }
                else { w1 *= 0.5f; w0 -= w1; }
                H::V3 tuv = (H::ld(UV) * w0 + H::ld(UV + 3) * w1) + H::ld(UV + 6) * (1.0f - w0 - w1);
                float col[4];
                H::bilinear(T.rgba, T.width, T.height, tuv.x, tuv.y, col);
                H::V3 l = H::v3(col[0], col[1], col[2]) * rad;
                if (H::norm2(l) < 0.01f) l = H::v3(0, 0, 0);
                uvw[i] = H::v3(w0, w1, w2);
                Lv[i] = l;
                total = total + l;
            }
            if ((double)H::norm2(total) < 1e-6) {
                Lv[0] = H::v3(1, 1, 1) * (float)1e-6;
                uvw[0] = H::v3(0, 0, 0);
            } else {
                total = total * (2.0f * H::PI * area / (float)NS);
            }
            flux = total;
            std::vector<float> vals(NS), prob;
            std::vector<int> left, right;
            for (int i = 0; i < NS; ++i) vals[i] = H::norm(Lv[i]);
            float avg = 0.0f;
            H::build_alias(vals, prob, left, right, avg);
            L.textured = 1;
            L.tab_off = (int)tprob.size();
            L.tab_avg = avg;
            for (int i = 0; i < NS; ++i) {
                tuvw.push_back(make_float4(uvw[i].x, uvw[i].y, uvw[i].z, vals[i]));
                tL.push_back(make_float4(Lv[i].x, Lv[i].y, Lv[i].z, 0.0f));
                tprob.push_back(prob[i]);
                tleft.push_back(left[i]);
                tright.push_back(right[i]);
            }
        } else {
            flux = rad * (2.0f * H::PI * area);
            L.textured = 0;
            L.tab_off = 0;
            L.tab_avg = 1.0f;
        }
        L.flux[0] = flux.x; L.flux[1] = flux.y; L.flux[2] = flux.z;
        L.flux_norm = H::norm(flux);
        lights.push_back(L);
    }
    if (lights.empty()) return fail(SPPM_E_SCENE, "scene has no emitting facet (Radiance sum > 1e-5)");
    // LightSampler::GetAreaLightFromFacetID (LightSampler.cpp:98-106) as a table
    std::vector<int> facet_light(nf, -1);
    for (size_t i = 0; i < lights.size(); ++i) facet_light[lights[i].facet] = (int)i;
    std::vector<float> lvals(lights.size()), lprob;
    std::vector<int> lleft, lright;
    for (size_t i = 0; i < lights.size(); ++i) lvals[i] = lights[i].flux_norm;
    float lavg = 0.0f;
    H::build_alias(lvals, lprob, lleft, lright, lavg);

    // ---- attributes ----
    std::vector<DevAttr> attrs(d->nattrs);
    for (int a = 0; a < d->nattrs; ++a) {
        const sppm_attr& A = d->attrs[a];
        DevAttr& D = attrs[a];
        D.kind = A.kind; D.index = A.index; D.opaque = A.if_opaque ? 1 : 0; D.is_sppm = A.is_sppm ? 1 : 0;
        for (int k = 0; k < 3; ++k) { D.albedo[k] = A.albedo[k]; D.radiance[k] = A.radiance[k]; }
        D.rough = A.roughness; D.refl = A.reflectance; D.absorb = A.extinction;
        D.tex_albedo = A.tex_albedo; D.tex_rough = A.tex_roughness; D.tex_bump = A.tex_bump;
        D.radiance_norm2 = H::norm2(H::ld(A.radiance));
        D.delta = A.if_delta ? 1 : 0;
    }
    // ---- textures (+ the environment map after IBL::Init's conversion
    // 0.4 * c^2, LightSampler.cpp:22-26, appended as one more texture) ----
    const int ntex_all = d->ntextures + (d->has_ibl ? 1 : 0);
    std::vector<TexDesc> texd(std::max(1, ntex_all));
    long long ntexel = 0;
    for (int t = 0; t < ntex_all; ++t) {
        const sppm_texture& T = d->textures[t < d->ntextures ? t : d->ibl_texture];
        texd[t].width = T.width;
        texd[t].height = T.height;
        texd[t].offset = ntexel;
        ntexel += (long long)T.width * T.height;
    }
    std::vector<float4> ibl_conv;
    if (d->has_ibl) {
        const sppm_texture& T = d->textures[d->ibl_texture];
        ibl_conv.resize((size_t)T.width * T.height);
        for (size_t k = 0; k < ibl_conv.size(); ++k) {
            const float* c = T.rgba + 4 * k;
            ibl_conv[k] = make_float4(0.4f * (c[0] * c[0]), 0.4f * (c[1] * c[1]), 0.4f * (c[2] * c[2]), 0.4f * 0.0f);
        }
    }
    // ---- facets ----
    std::vector<float2> uv2(3 * (size_t)nf);
    for (size_t f = 0; f < (size_t)nf; ++f)
        for (int k = 0; k < 3; ++k) uv2[3 * f + k] = make_float2(d->uv[9 * f + 3 * k], d->uv[9 * f + 3 * k + 1]);
    // scene box for the photon grid
    float bmin[3] = {3e38f, 3e38f, 3e38f}, bmax[3] = {-3e38f, -3e38f, -3e38f};
    for (size_t i = 0; i < (size_t)nf * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            float x = d->pos[3 * i + k];
            bmin[k] = std::min(bmin[k], x);
            bmax[k] = std::max(bmax[k], x);
        }
    for (int k = 0; k < 3; ++k) {
        S.bmin[k] = bmin[k] - 0.01f;
        S.ext[k] = (bmax[k] - bmin[k]) + 0.02f;
    }

    int rc = 0;
    float* dpos = nullptr;
    if ((rc = dalloc(&dpos, 9 * (size_t)nf))) return rc;
    if ((rc = dalloc(&S.facet_nrm, 9 * (size_t)nf)) || (rc = dalloc(&S.facet_uv, 3 * (size_t)nf)) ||
        (rc = dalloc(&S.facet_attr, nf)) || (rc = dalloc(&S.attrs, attrs.size())) || (rc = dalloc(&S.tex, texd.size())) ||
        (rc = dalloc(&S.texels, (size_t)std::max(1LL, ntexel))) || (rc = dalloc(&S.lpick_prob, lprob.size())) ||
        (rc = dalloc(&S.lpick_left, lleft.size())) || (rc = dalloc(&S.lpick_right, lright.size())) ||
        (rc = dalloc(&S.lights, lights.size())) || (rc = dalloc(&S.ltab_uvw, std::max<size_t>(1, tuvw.size()))) ||
        (rc = dalloc(&S.ltab_L, std::max<size_t>(1, tL.size()))) || (rc = dalloc(&S.ltab_prob, std::max<size_t>(1, tprob.size()))) ||
        (rc = dalloc(&S.ltab_left, std::max<size_t>(1, tleft.size()))) || (rc = dalloc(&S.ltab_right, std::max<size_t>(1, tright.size()))) ||
        (rc = dalloc(&S.facet_light, nf))) {
        dfree(dpos);
        S.release();
        return rc;
    }
    HIPOK(hipMemcpyAsync(dpos, d->pos, sizeof(float) * 9 * nf, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.facet_nrm, d->nrm, sizeof(float) * 9 * nf, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.facet_uv, uv2.data(), sizeof(float2) * 3 * nf, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.facet_attr, d->attr, sizeof(int) * nf, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.attrs, attrs.data(), sizeof(DevAttr) * attrs.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.tex, texd.data(), sizeof(TexDesc) * texd.size(), hipMemcpyHostToDevice, st));
    for (int t = 0; t < d->ntextures; ++t)
        HIPOK(hipMemcpyAsync(S.texels + texd[t].offset, d->textures[t].rgba,
                             sizeof(float4) * (size_t)texd[t].width * texd[t].height, hipMemcpyHostToDevice, st));
    if (d->has_ibl) {
        HIPOK(hipMemcpyAsync(S.texels + texd[d->ntextures].offset, ibl_conv.data(), sizeof(float4) * ibl_conv.size(),
                             hipMemcpyHostToDevice, st));
        HIPOK(hipStreamSynchronize(st));  // ibl_conv is a local
    }
    HIPOK(hipMemcpyAsync(S.lpick_prob, lprob.data(), sizeof(float) * lprob.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.lpick_left, lleft.data(), sizeof(int) * lleft.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.lpick_right, lright.data(), sizeof(int) * lright.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.lights, lights.data(), sizeof(DevLight) * lights.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(S.facet_light, facet_light.data(), sizeof(int) * nf, hipMemcpyHostToDevice, st));
    S.hlights = lights;
    S.h_tuvw = tuvw;
    S.h_tL = tL;
    S.h_lpick_avg = lavg;
    if (!tuvw.empty()) {
        HIPOK(hipMemcpyAsync(S.ltab_uvw, tuvw.data(), sizeof(float4) * tuvw.size(), hipMemcpyHostToDevice, st));
        HIPOK(hipMemcpyAsync(S.ltab_L, tL.data(), sizeof(float4) * tL.size(), hipMemcpyHostToDevice, st));
        HIPOK(hipMemcpyAsync(S.ltab_prob, tprob.data(), sizeof(float) * tprob.size(), hipMemcpyHostToDevice, st));
        HIPOK(hipMemcpyAsync(S.ltab_left, tleft.data(), sizeof(int) * tleft.size(), hipMemcpyHostToDevice, st));
        HIPOK(hipMemcpyAsync(S.ltab_right, tright.data(), sizeof(int) * tright.size(), hipMemcpyHostToDevice, st));
    }
    char err[256] = {0};
    if (build_bvh(dpos, nf, bvh_width, !lbvh, st, &S.bvh, err, sizeof err) != 0) {
        dfree(dpos);
        S.release();
        return fail(SPPM_E_HIP, "BVH build failed: %s", err);
    }
    HIPOK(hipStreamSynchronize(st));
    dfree(dpos);

    SceneDev& D = S.dev;
    D.width = bvh_width;
    D.nnodes = S.bvh.nnodes;
    D.nodes = S.bvh.nodes;
    D.ntri = nf;
    D.tri = S.bvh.tri;
    D.nfacets = nf;
    D.facet_nrm = S.facet_nrm;
    D.facet_uv = S.facet_uv;
    D.facet_attr = S.facet_attr;
    D.attrs = S.attrs;
    D.tex = S.tex;
    D.texels = S.texels;
    D.ibl_tex = d->has_ibl ? d->ntextures : -1;
    D.nlights = (int)lights.size();
    D.lpick_avg = lavg;
    D.lpick_prob = S.lpick_prob;
    D.lpick_left = S.lpick_left;
    D.lpick_right = S.lpick_right;
    D.lights = S.lights;
    D.ltab_uvw = S.ltab_uvw;
    D.ltab_L = S.ltab_L;
    D.ltab_prob = S.ltab_prob;
    D.ltab_left = S.ltab_left;
    D.ltab_right = S.ltab_right;
    D.facet_light = S.facet_light;
    // traversal stack bound: each wide level pushes at most W entries
    S.stack_cap = bvh_width * (S.bvh.depth + 1) + 4;
    S.spill_blocks = trace_grid_blocks(bvh_width);
    if (S.spill_blocks <= 0) {
        S.release();
        return fail(SPPM_E_HIP, "cannot query the device for the path kernel grid");
    }
    {
        const size_t per_lane = S.stack_cap > TRACE_LDS ? (size_t)(S.stack_cap - TRACE_LDS) : 1;
        int rc;
        if ((rc = dalloc(&S.spill, per_lane * (size_t)S.spill_blocks * 64)) ||
            (rc = dalloc(&S.spill_eye, per_lane * (size_t)S.spill_blocks * 64)) || (rc = dalloc(&S.next_path, 2 * (size_t)STRIPE_INTS))) {
            S.release();
            return rc;
        }
    }
    S.ready = true;
    return SPPM_OK;
}

// ---------------------------------------------------------------------------
// Camera (iCamera ctor, Camera.h:34-53)
// ---------------------------------------------------------------------------
int sppm_camera_set(sppm_ctx* c, const sppm_camera_desc* cd) {
    if (!c || !cd) return fail(SPPM_E_ARG, "NULL argument");
    if (int rc = not_in_batch(c, "sppm_camera_set")) return rc;
    if (cd->type != SPPM_CAM_PINHOLE && cd->type != SPPM_CAM_THINLENS) return fail(SPPM_E_ARG, "unknown camera type");
    if (cd->width <= 0 || cd->height <= 0) return fail(SPPM_E_ARG, "bad resolution %dx%d", cd->width, cd->height);
    if ((long long)cd->width * cd->height >= (1LL << 28)) return fail(SPPM_E_ARG, "resolution too large");
    if (cd->sampler < SPPM_SAMPLER_UNIFORM || cd->sampler > SPPM_SAMPLER_HAMMERSLEY) return fail(SPPM_E_ARG, "unknown sampler");
    if (cd->sampler == SPPM_SAMPLER_HAMMERSLEY && cd->sampler_nsample <= 0) return fail(SPPM_E_ARG, "Hammersley NumSample must be > 0");
    c->cam = *cd;
    CamDev& C = c->camdev;
    C.type = cd->type;
    C.W = cd->width;
    C.H = cd->height;
    H::V3 pos = H::ld(cd->pos);
    H::V3 dir = H::normalized(H::ld(cd->target) - pos);
    float fov = H::PI * cd->fov_deg / 360.0f;
    C.PW = (float)(2.0 * (double)cd->dist_to_lens * tan((double)fov) * (double)cd->aspect);
    C.PH = (float)(2.0 * (double)cd->dist_to_lens * tan((double)fov));
    C.LW = C.PW;
    C.LH = C.PH;
    H::V3 e1, e2;
    if ((double)dir.y > 0.99) {
        e1 = H::v3(1, 0, 0);
        e2 = H::v3(0, 0, 1);
    } else {
        e1 = H::normalized(H::cross(H::v3(0, 1, 0), dir));
        e2 = H::normalized(H::cross(dir, e1));
    }
    C.pos[0] = pos.x; C.pos[1] = pos.y; C.pos[2] = pos.z;
    C.dir[0] = dir.x; C.dir[1] = dir.y; C.dir[2] = dir.z;
    C.e1[0] = e1.x; C.e1[1] = e1.y; C.e1[2] = e1.z;
    C.e2[0] = e2.x; C.e2[1] = e2.y; C.e2[2] = e2.z;
    C.F = cd->F;
    C.D = cd->dist_to_lens;
    C.sampler = cd->sampler;
    C.dia_n = 0;
    dfree(c->dia_prob);
    dfree(c->dia_uv);
    C.dia_prob = nullptr;
    C.dia_uv = nullptr;
    c->have_cam = true;
    c->rb.ready = false;  // resolution may have changed
    return SPPM_OK;
}

// ThinLensCamera(..., DiaphragmImageFileName) (Camera.h:136-158): the alias
// table over the image's pixels, weight 1 - red, and the open-pixel share
// that selects plain lens sampling (Camera.h:106-107).
int sppm_camera_diaphragm(sppm_ctx* c, int32_t width, int32_t height, const float* rgba) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    if (int rc = not_in_batch(c, "sppm_camera_diaphragm")) return rc;
    if (!c->have_cam) return fail(SPPM_E_STATE, "no camera set (sppm_camera_set first)");
    HIPOK(hipSetDevice(c->device));
    CamDev& C = c->camdev;
    C.dia_n = 0;
    dfree(c->dia_prob);
    dfree(c->dia_uv);
    C.dia_prob = nullptr;
    C.dia_uv = nullptr;
    c->rb.ready = false;
    if (!rgba) return SPPM_OK;
    if (c->cam.type != SPPM_CAM_THINLENS) return fail(SPPM_E_ARG, "a diaphragm needs a thin-lens camera");
    if (width <= 0 || height <= 0 || (long long)width * height > (1LL << 26))
        return fail(SPPM_E_ARG, "diaphragm image %dx%d: sides > 0, at most 2^26 pixels", width, height);
    const int n = width * height;
    std::vector<float> L(n);
    int open = 0;
    for (int i = 0; i < n; ++i) {
        const float r = rgba[4 * (size_t)i];
        if ((double)r < 1e-6) ++open;
        L[i] = (float)(1.0 - (double)r);
    }
    const float hole = (float)open / (float)n;
    if ((double)hole > 0.999) return SPPM_OK;  // plain lens sampling
    std::vector<float> prob;
    std::vector<int> left, right;
    float avg = 0.0f;
    H::build_alias(L, prob, left, right, avg);
    if (!(avg > 0.0f)) return fail(SPPM_E_ARG, "the diaphragm image has no open pixel (all red >= 1)");
    auto lens_u = [&](int i) { return (float)(i % width) / (float)width; };
    auto lens_v = [&](int i) { return (float)(1.0 - (double)((float)(i / width) / (float)height)); };
    std::vector<float4> uv(n);
    for (int k = 0; k < n; ++k) {
        const int l = left[k], r = right[k] < 0 ? l : right[k];
        uv[k] = make_float4(lens_u(l), lens_v(l), lens_u(r), lens_v(r));
    }
    int rc;
    if ((rc = dalloc(&c->dia_prob, (size_t)n)) || (rc = dalloc(&c->dia_uv, (size_t)n))) return rc;
    HIPOK(hipMemcpy(c->dia_prob, prob.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(c->dia_uv, uv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    C.dia_n = n;
    C.dia_prob = c->dia_prob;
    C.dia_uv = c->dia_uv;
    return SPPM_OK;
}

int sppm_params_set(sppm_ctx* c, const sppm_params* p) {
    if (!c || !p) return fail(SPPM_E_ARG, "NULL argument");
    if (int rc = not_in_batch(c, "sppm_params_set")) return rc;
    if (p->photons_per_pass <= 0) return fail(SPPM_E_ARG, "photons_per_pass must be > 0");
    if ((long long)p->photons_per_pass * (NSLOT - 1) > RADIX_SORT_MAX_N)
        return fail(SPPM_E_ARG, "photons_per_pass %d: at most %d (stored photons are sorted with 30-bit counts)",
                    p->photons_per_pass, RADIX_SORT_MAX_N / (NSLOT - 1));
    if (!(p->alpha > 0.0f && p->alpha <= 1.0f)) return fail(SPPM_E_ARG, "alpha must be in (0, 1]");
    if (!(p->first_radius > 0.0f)) return fail(SPPM_E_ARG, "first_radius must be > 0");
    if (p->eye_nee < SPPM_NEE_OFF || p->eye_nee > SPPM_NEE_KEEP) return fail(SPPM_E_ARG, "eye_nee must be 0, 1 or 2");
    if (p->record_neighbours < 0 || p->record_neighbours > 2) return fail(SPPM_E_ARG, "record_neighbours must be 0, 1 or 2");
    if (p->step_offset < 0) return fail(SPPM_E_ARG, "step_offset must be >= 0");
    if (p->photon_kernel != SPPM_PHOTON_WAVEFRONT && p->photon_kernel != SPPM_PHOTON_MEGAKERNEL)
        return fail(SPPM_E_ARG, "photon_kernel must be SPPM_PHOTON_MEGAKERNEL (0) or SPPM_PHOTON_WAVEFRONT (1)");
    if (p->wave_bounces < 0 || p->wave_bounces > 32) return fail(SPPM_E_ARG, "wave_bounces must be in [0, 32]");
    if (p->wave_lanes != 0 && !wave_lanes_ok(p->bvh_width > 0 ? p->bvh_width : 8, p->wave_lanes))
        return fail(SPPM_E_ARG, "wave_lanes must be 0, 2, 4 or 8 and at most the BVH width");
    if (p->record_neighbours == 2 && p->max_neighbours <= 0)
        return fail(SPPM_E_ARG, "record_neighbours 2 needs max_neighbours > 0 (pairs recorded per pass)");
    c->params = *p;
    c->have_params = true;
    c->rb.ready = false;
    return SPPM_OK;
}

// initializeSharedPoints (SPPM.cpp:356-373) + buffers of the pass loop
// Camera stream segments: the block of a pass (6 words per pixel) is cut
// into K segments of L words, K the largest divisor of the block up to 16.
// Segment k starts k * L words into the stream; after a pass every segment
// jumps over the K - 1 others.  Every jump costs the same (~10^4 shifted
// windows), so fewer, longer segments cost less in all; one wave emits a
// segment, and the words are made a pass ahead of their use.
static int init_camera_stream(sppm_ctx* c) {
    RenderBuffers& R = c->rb;
    const long long block = 6LL * R.np;
    int K = 1;
    for (int k = 16; k >= 1; --k)
        if (block % k == 0) {
            K = k;
            break;
        }
    const long long L = block / K;
    int rc = dalloc(&R.cam_states, 2 * (size_t)K * MT_WORDS);  // the windows, then scratch for the kernels
    if (rc) return rc;
    std::vector<uint32_t> win((size_t)K * MT_WORDS);
    mt_window_seed((uint32_t)c->cam.sampler_id, win.data());
    for (int k = 1; k < K; ++k) {
        std::copy(win.begin() + (size_t)(k - 1) * MT_WORDS, win.begin() + (size_t)k * MT_WORDS,
                  win.begin() + (size_t)k * MT_WORDS);
        mt_window_advance(win.data() + (size_t)k * MT_WORDS, (uint64_t)L);
    }
    uint32_t phi[MT_WORDS];
    try {
        mt_jump_poly((uint64_t)(K - 1) * (uint64_t)L, phi);
    } catch (const std::exception& e) {
        return fail(SPPM_E_STATE, "camera stream jump polynomial: %s", e.what());
    }
    HIPOK(hipMemcpy(R.cam_states, win.data(), win.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(R.cam_phi, phi, sizeof phi, hipMemcpyHostToDevice));
    R.cam_K = K;
    R.cam_L = L;
    return SPPM_OK;
}

int sppm_begin(sppm_ctx* c) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    if (!c->have_cam) return fail(SPPM_E_STATE, "no camera set");
    if (!c->have_params) return fail(SPPM_E_STATE, "no params set");
    if (int rc = not_in_batch(c, "sppm_begin")) return rc;
    c->lead_ok = false;
    HIPOK(hipSetDevice(c->device));
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipStreamSynchronize(c->stream2));
    HIPOK(hipStreamSynchronize(c->stream3));
    c->words_gen = c->words_used = 0;
    c->eye_recorded[0] = c->eye_recorded[1] = false;
    drop_graph(c);
    RenderBuffers& R = c->rb;
    R.release();
    const int N = c->params.photons_per_pass;
    const int W = c->cam.width, Hh = c->cam.height, np = W * Hh;
    R.N = N; R.W = W; R.H = Hh; R.np = np;
    R.photon_cap = N * (NSLOT - 1);  // checked against RADIX_SORT_MAX_N by sppm_params_set
    R.cell_cap = 1 << 22;
    const size_t NS = (size_t)NSLOT * N;
    int rc = 0;
    // tiles of the recording gather (8x8 pixels) or hit-point groups of the production gather
    const int ntile_max = std::max(gather_tile_count(W, Hh), gather_group_count(np));
    R.slots.npaths = N;
    if ((rc = dalloc(&R.slots.pos, NS)) || (rc = dalloc(&R.slots.flux, NS)) || (rc = dalloc(&R.slots.dir, NS)) ||
        (rc = dalloc(&R.slots.valid, NS)) || (rc = dalloc(&R.slots.count, N)) || (rc = dalloc(&R.off, N)) ||
        (rc = dalloc(&R.nstored, 1)) || (rc = dalloc(&R.scan_tmp, scan_tmp_entries(N))) ||
        (rc = dalloc(&R.cell_gaps, cell_gap_entries(R.cell_cap))) || (rc = dalloc(&R.cell_ngap, 1)) ||
        (rc = dalloc(&R.cell_start, R.cell_cap + 1)) || (rc = dalloc(&R.psort.keys, R.photon_cap)) ||
        (rc = dalloc(&R.psort.vals, R.photon_cap)) || (rc = dalloc(&R.psort.tmp_k, R.photon_cap)) ||
        (rc = dalloc(&R.psort.tmp_v, R.photon_cap)) || (rc = dalloc(&R.staged.rec, 3 * (size_t)R.photon_cap)) ||
        (rc = dalloc(&R.psort.hist, radix_hist_entries(R.photon_cap))) ||
        (rc = dalloc(&R.gpos, R.photon_cap)) || (rc = dalloc(&R.gflux, R.photon_cap)) ||
        (rc = dalloc(&R.gdir, R.photon_cap)) || (rc = dalloc(&R.gp, 1)) || (rc = dalloc(&R.shared.rad, np)) ||
        (rc = dalloc(&R.shared.cnt, np)) || (rc = dalloc(&R.shared.direct, np)) ||
        (rc = dalloc(&R.counts, np + 1)) || (rc = dalloc(&R.list_off, np + 1)) ||
        (rc = dalloc(&R.list_scan_tmp, scan_tmp_entries(np + 1))) ||
        (rc = dalloc(&R.words, 2 * 6 * (size_t)np)) || (rc = dalloc(&R.cam_phi, MT_WORDS)) ||
        (rc = dalloc(&R.rstat, 2)) || (rc = dalloc(&R.flags, 1)) || (rc = dalloc(&R.ctr, 1)) ||
        (rc = dalloc(&R.raystripes, (size_t)STRIPES * RAY_STRIPE)) ||
        (rc = dalloc((char**)&R.tiles, gather_tile_bytes() * (size_t)ntile_max)) ||
        (rc = dalloc(&R.chunkp, 1)) || (rc = dalloc(&R.nch, ntile_max + 1)) ||
        (rc = dalloc(&R.chunk_off, ntile_max + 1)) ||
        (rc = dalloc(&R.tile_scan_tmp, scan_tmp_entries(ntile_max + 1))) ||
        (rc = dalloc(&R.partial, (size_t)64 * (GATHER_CHUNK_CAP + ntile_max))) ||
        (rc = dalloc(&R.groups.keys, np)) || (rc = dalloc(&R.groups.vals, np)) || (rc = dalloc(&R.groups.tmp_k, np)) ||
        (rc = dalloc(&R.groups.tmp_v, np)) || (rc = dalloc(&R.groups.hist, radix_hist_entries(np))) ||
        (rc = dalloc(&R.groups.nvalid, 1)) || (rc = dalloc(&R.groups.chunkq, STRIPE_INTS)) || (rc = dalloc(&R.groups.stat, 3 * (size_t)gather_group_count(np))) ||
        (rc = dalloc(&R.pass_dev, 3)) ||
        (rc = dalloc(&R.image, np))) {
        R.release();
        return rc;
    }
    for (HitDev& h : R.hitbuf) {
        if ((rc = dalloc(&h.pos, np)) || (rc = dalloc(&h.nrm, np)) || (rc = dalloc(&h.out, np)) ||
            (rc = dalloc(&h.weight, np)) || (rc = dalloc(&h.mat, np)) || (rc = dalloc(&h.absorb, np)) ||
            (rc = dalloc(&h.direct, np))) {
            R.release();
            return rc;
        }
        h.np = np;
    }
    R.hits = R.hitbuf[0];
    if (c->params.photon_kernel == SPPM_PHOTON_WAVEFRONT) {
        const int B = c->params.wave_bounces > 0 ? c->params.wave_bounces : WAVE_BOUNCES_DEFAULT;
        const int width = c->scene.dev.width;
        int lanes = c->params.wave_lanes > 0 ? c->params.wave_lanes : WAVE_LANES_DEFAULT;
        if (lanes > width) lanes = width;
        if (!wave_lanes_ok(width, lanes)) {
            R.release();
            return fail(SPPM_E_ARG, "wave_lanes %d with a BVH of width %d", lanes, width);
        }
        R.wave_lanes = lanes;
        R.wave_blocks = wave_trace_blocks(width, lanes);
        if (R.wave_blocks <= 0) {
            R.release();
            return fail(SPPM_E_HIP, "cannot query the device for the trace kernel grid");
        }
        R.wave.cap = c->scene.stack_cap;
        WaveBuffers& w = R.wave;
        const size_t n = (size_t)N;
        if ((rc = dalloc(&w.path.rng, n)) || (rc = dalloc(&w.path.org, n)) || (rc = dalloc(&w.path.dir, n)) ||
            (rc = dalloc(&w.path.ndir, n)) || (rc = dalloc(&w.path.rad, n)) || (rc = dalloc(&w.path.thr, n)) ||
            (rc = dalloc(&w.rays[0], 2 * wave_queue_entries(N))) ||
            (rc = dalloc(&w.rays[1], 2 * wave_queue_entries(N))) || (rc = dalloc(&w.hits, wave_queue_entries(N))) ||
            (rc = dalloc(&w.qctl, wave_ctl_ints(B))) ||
            (rc = dalloc(&w.spill, wave_spill_entries(lanes, R.wave_blocks, R.wave.cap)))) {
            R.release();
            return rc;
        }
        R.wave_bounces = B;
    }
    if (c->params.record_neighbours == 2) {
        R.pairs_cap = c->params.max_neighbours;
        if ((rc = dalloc(&R.pairs, (size_t)R.pairs_cap)) || (rc = dalloc(&R.npairs, 2))) {
            R.release();
            return rc;
        }
    }
    if (c->guided) {
        if ((rc = dalloc(&R.path_cell, 2 * (size_t)N)) || (rc = dalloc(&R.path_ratio, 2 * (size_t)N)) ||
            (rc = dalloc(&R.pdep, R.photon_cap)) || (rc = dalloc(&R.doff, (size_t)N + 1))) {
            R.release();
            return rc;
        }
        c->emit.path_cell = R.path_cell;
        c->emit.path_ratio = c->gparams.target == SPPM_GUIDE_TARGET_RMS ? R.path_ratio : nullptr;
    }
    HIPOK(hipMemsetAsync(R.counts, 0, sizeof(int) * (size_t)(np + 1), c->stream));
    launch_init_slots(R.slots, c->stream);
    // device pass counter: steps are pass + 1 + step_offset (graph replays read it)
    launch_set_u32(reinterpret_cast<uint32_t*>(R.pass_dev), (uint32_t)c->params.step_offset, c->stream);
    launch_init_shared(R.shared, np, c->params.first_radius, c->stream);
    // UniformSample(ID): the mt19937(ID) stream in K segments per pass
    if (c->cam.sampler == SPPM_SAMPLER_UNIFORM) {
        rc = init_camera_stream(c);
        if (rc) {
            R.release();
            return rc;
        }
    }
    R.ntiles = gather_tile_count(W, Hh);
    launch_set_u64x2(R.rstat, (unsigned long long)((double)c->params.first_radius * 4294967296.0), 1ull, c->stream);
    launch_set_u32(R.flags, 0u, c->stream);
    launch_set_u32(reinterpret_cast<uint32_t*>(R.nstored), 0u, c->stream);
    {
        Counters z{};
        HIPOK(hipMemcpyAsync(R.ctr, &z, sizeof z, hipMemcpyHostToDevice, c->stream));
        HIPOK(hipMemsetAsync(R.raystripes, 0, sizeof(unsigned long long) * STRIPES * RAY_STRIPE, c->stream));
    }
    HIPOK(hipGetLastError());
    HIPOK(hipStreamSynchronize(c->stream));
    c->pass = 0;
    for (Timer* t : {&c->t_photons, &c->t_build, &c->t_eye, &c->t_gather}) t->reset();
    R.ready = true;
    R.have_lists = false;
    return SPPM_OK;
}

static int check_ready(sppm_ctx* c) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    c->lead_ok = false;
    if (!c->rb.ready) return fail(SPPM_E_STATE, "sppm_begin has not been called since the last setup change");
    return batch_wait(c);
}
// check_ready for the calls that queue pass work on the context itself
static int check_own_pass(sppm_ctx* c, const char* what) {
    int rc = check_ready(c);
    if (rc) return rc;
    return not_in_batch(c, what);
}

static void fill_fixed_sample(sppm_ctx* c, int step) {
    CamDev& C = c->camdev;
    if (c->cam.sampler == SPPM_SAMPLER_HALTON) {
        C.fixed_smp[0] = H::halton(2, step);
        C.fixed_smp[1] = H::halton(3, step);
        C.fixed_smp[2] = H::halton(5, step);
    } else if (c->cam.sampler == SPPM_SAMPLER_HAMMERSLEY) {
        C.fixed_smp[0] = (float)step / (float)c->cam.sampler_nsample;
        C.fixed_smp[1] = H::halton(2, step);
        C.fixed_smp[2] = H::halton(3, step);
    }
}

static GridDev grid_dev(sppm_ctx* c) {
    GridDev G{};
    G.cell_start = c->rb.cell_start;
    G.pos = c->rb.gpos;
    G.flux = c->rb.gflux;
    G.dir = c->rb.gdir;
    return G;
}

// ---- split steps --------------------------------------------------------
// traversal spill area and the path-kernel grid (the autotuner may halve it)
static TraceSpill trace_spill(const sppm_ctx* c) {
    TraceSpill sp = c->scene.trace_spill();
    if (c->tuned_blocks > 0 && c->tuned_blocks < sp.blocks) sp.blocks = c->tuned_blocks;
    return sp;
}
// the eye pass's own spill area: in a full pass it runs beside the photon pass
static TraceSpill trace_spill_eye(const sppm_ctx* c) {
    TraceSpill sp = c->scene.trace_spill_eye();
    if (c->tuned_blocks > 0 && c->tuned_blocks < sp.blocks) sp.blocks = c->tuned_blocks;
    return sp;
}
// device addresses of the striped ray counters (ties two words further)
static unsigned long long* photon_rays(RenderBuffers& R) { return R.raystripes; }
static unsigned long long* eye_rays(RenderBuffers& R) { return R.raystripes + 1; }
// the Counters record with the striped ray and tie counts added
static int read_counters(sppm_ctx* c, Counters& h) {
    RenderBuffers& R = c->rb;
    std::vector<unsigned long long> rs((size_t)STRIPES * RAY_STRIPE);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(&h, R.ctr, sizeof h, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(rs.data(), R.raystripes, sizeof(unsigned long long) * rs.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < STRIPES; ++k) {
        const unsigned long long* l = rs.data() + (size_t)k * RAY_STRIPE;
        h.photon_rays += l[0];
        h.eye_rays += l[1];
        h.photon_ties += l[2];
        h.eye_ties += l[3];
    }
    return SPPM_OK;
}
// beside_gather: the photon pass runs beside the previous pass's gather and
// takes 10/16 of the path-kernel grid, so the gather's kernels find free wave
// slots instead of waiting for the persistent photon waves to drain.  On the
// full-material workload (1M photons, 1024^2) a pass takes 3.95 ms with
// 10/16, 4.03 with 9/16, 4.29 with 8/16, 3.94-4.47 with 11/16 and 4.04-4.58
// with 12/16 (two steady states each), and 4.52 with the full grid, whose
// photon waves starve the gather.
static void photon_pass(sppm_ctx* c, int step, const int* step_dev, hipStream_t st, bool beside_gather) {
    RenderBuffers& R = c->rb;
    TraceSpill sp = trace_spill(c);
    if (beside_gather) sp.blocks = sp.blocks * 10 / 16 > 0 ? sp.blocks * 10 / 16 : 1;
    if (R.wave_bounces > 0) {
        int tb = R.wave_blocks;
        if (beside_gather) tb = tb * 10 / 16 > 0 ? tb * 10 / 16 : 1;
        launch_photon_wave(c->scene.dev, R.slots, R.N, step, step_dev, R.wave, R.wave_bounces, R.wave_lanes, tb, sp,
                           c->guided ? &c->emit : nullptr, R.flags, photon_rays(R), st);
    } else {
        launch_photon(c->scene.dev, R.slots, R.N, step, step_dev, sp, c->scene.next_path,
                      c->guided ? &c->emit : nullptr, R.flags, photon_rays(R), st);
    }
}
static void do_photons(sppm_ctx* c, int step, hipStream_t st, bool beside_gather = false) {
    photon_pass(c, step, nullptr, st, beside_gather);
}
// Training records of the pass (deposit histogram), and on training passes
// (every train_interval-th step; every replay of a captured pass) Adam steps
// on the decayed histogram and the new proposal.  The emission of pass s
// samples the proposal published by the build of pass s - 1, i.e. the one
// trained after pass s - 2: a pass's photon pass can then run beside the
// previous pass's gather and training.  Each photon carries the PDF of the
// proposal that drew it, so the estimate stays unbiased.
static void do_guide(sppm_ctx* c, hipStream_t st, int step, const int* step_dev) {
    RenderBuffers& R = c->rb;
    const bool rms = c->gparams.target == SPPM_GUIDE_TARGET_RMS;
    launch_guide_deposit(R.doff, R.doff + R.N, R.pdep, R.path_cell, rms ? R.path_ratio : nullptr, R.N, step, step_dev,
                         c->guide.hpass, st);
    const int every = c->gparams.train_interval > 1 ? c->gparams.train_interval : 1;
    if (!step_dev && step % every != 0) return;  // the deposits gather in the pass histogram until then
    guide_fold_hist(c->guide, c->gparams.history_decay, st);
    const float* w = c->guide.hist;
    if (rms) {
        guide_rms_target(c->guide, st);
        w = c->guide.tgt;
    }
    guide_hist_sum(c->guide, w, st);
    for (int k = 0; k < c->gparams.train_steps; ++k) {
        guide_gradient(c->guide, w, st);
        guide_adam(c->guide, c->gparams.learning_rate, st);
    }
    guide_infer(c->guide, st);
}
// The build's first step: the pass's flattened photons out of the slots
// (after it the next photon pass may rewrite them).
static void do_stage(sppm_ctx* c, hipStream_t st) {
    RenderBuffers& R = c->rb;
    exclusive_scan(R.slots.count, R.off, R.N, R.nstored, R.scan_tmp, st);
    launch_photon_stage(R.slots, R.off, R.staged, st);
}
// Guided emission: once the photon pass that read the emission tables is
// done, and before the next one, the proposal trained last (after the
// previous pass's gather) becomes the next photon pass's.
static void do_publish(sppm_ctx* c, hipStream_t st) {
    if (c->guided) guide_publish(c->guide, st);
}
// The rest of the build: grid, keys, sort and the cell-sorted records.
static void do_build_grid(sppm_ctx* c, hipStream_t st) {
    RenderBuffers& R = c->rb;
    if (c->guided) launch_copy_offsets(R.off, R.nstored, R.N, R.doff, st);
    launch_build_grid(R.staged, R.nstored, R.ctr, R.gp, R.rstat, c->scene.bmin, c->scene.ext, R.cell_cap,
                      R.cell_start, R.cell_gaps, R.cell_ngap, R.psort, R.photon_cap, R.gpos, R.gflux, R.gdir, st);
}
static void do_build(sppm_ctx* c, hipStream_t st) {
    do_stage(c, st);
    do_publish(c, st);
    do_build_grid(c, st);
}
// Generates camera word blocks up to and including block k on stream2, in
// stream order (each block continues the shared mt19937 stream).
static int words_generate_upto(sppm_ctx* c, long long k) {
    if (c->cam.sampler != SPPM_SAMPLER_UNIFORM) return SPPM_OK;
    const size_t blk = 6 * (size_t)c->rb.np;
    while (c->words_gen <= k) {
        const int b = (int)(c->words_gen & 1);
        if (c->eye_recorded[b]) HIPOK(hipStreamWaitEvent(c->stream2, c->ev_eye[b], 0));
        launch_camera_words(c->rb.cam_states, c->rb.cam_phi, c->rb.cam_K, c->rb.cam_L, c->rb.words + b * blk, c->stream2);
        HIPOK(hipEventRecord(c->ev_words[b], c->stream2));
        c->words_gen += 1;
    }
    return SPPM_OK;
}
// Eye pass on stream st with the next camera word block.
static int do_eye(sppm_ctx* c, int step, hipStream_t st, Timer* t) {
    fill_fixed_sample(c, step);
    const bool uni = c->cam.sampler == SPPM_SAMPLER_UNIFORM;
    const long long u = c->words_used;
    const int b = (int)(u & 1);
    if (uni) {
        int rc = words_generate_upto(c, u);
        if (rc) return rc;
        HIPOK(hipStreamWaitEvent(st, c->ev_words[b], 0));
    }
    if (t) t->start(st);
    launch_eye(c->scene.dev, c->camdev, c->rb.hits, step, nullptr, trace_spill_eye(c),
               c->rb.words + (size_t)b * 6 * (size_t)c->rb.np, c->rb.flags, eye_rays(c->rb), c->scene.next_path + STRIPE_INTS,
               c->params.eye_nee, st);
    if (t) t->stop(st);
    if (uni) {
        HIPOK(hipEventRecord(c->ev_eye[b], st));
        c->eye_recorded[b] = true;
        c->words_used = u + 1;
    }
    return SPPM_OK;
}
static int do_gather(sppm_ctx* c, hipStream_t st) {
    RenderBuffers& R = c->rb;
    GridDev G = grid_dev(c);
    const float alpha = c->params.alpha;
    if (c->params.record_neighbours == 1) {
        launch_gather(R.hits, R.shared, G, R.gp, R.W, R.H, alpha, 1, R.counts, nullptr, nullptr, R.rstat, R.ctr, R.flags, st);
        int total = 0;
        // counts[np] is kept at zero, so list_off[np] is the total
        exclusive_scan(R.counts, R.list_off, R.np + 1, nullptr, R.list_scan_tmp, st);
        HIPOK(hipMemcpyAsync(&total, R.list_off + R.np, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        if ((long long)total > R.lists_cap) {
            dfree(R.lists);
            int rc = dalloc(&R.lists, (size_t)total + 1);
            if (rc) return rc;
            R.lists_cap = total;
        }
        launch_set_u64x2(R.rstat, 0ull, 0ull, st);
        launch_gather(R.hits, R.shared, G, R.gp, R.W, R.H, alpha, 2, R.counts, R.list_off, R.lists, R.rstat,
                      R.ctr, R.flags, st);
        R.have_lists = true;
    } else {
        launch_set_u64x2(R.rstat, 0ull, 0ull, st);
        if (c->guided) launch_deposit_clear(R.pdep, R.doff + R.N, st);
        const PairRecord rec{R.pairs, R.npairs, R.pairs_cap};
        launch_gather_balanced(R.hits, R.shared, G, R.gp, R.np, c->scene.bmin, c->scene.ext, R.groups, alpha, R.tiles,
                               GATHER_CHUNK_CAP, R.chunkp, R.nch, R.chunk_off, R.tile_scan_tmp, R.partial, R.counts,
                               R.rstat, R.ctr, R.flags, c->guided ? R.pdep : nullptr, rec, st);
        R.have_lists = R.pairs != nullptr;
    }
    if (c->scene.dev.ibl_tex >= 0 || c->params.eye_nee != SPPM_NEE_OFF) launch_direct_light(R.hits, R.shared, R.np, st);
    return SPPM_OK;
}

int sppm_pass_photons(sppm_ctx* c, int step, int32_t* emitted) {
    int rc = check_own_pass(c, "sppm_pass_photons");
    if (rc) return rc;
    if (c->graph_exec) return fail(SPPM_E_STATE, "split pass steps are not available after sppm_graph_capture");
    c->t_photons.start(c->stream);
    do_photons(c, step, c->stream);
    c->t_photons.stop(c->stream);
    c->split_step = step;
    HIPOK(hipGetLastError());
    if (emitted) *emitted = c->rb.N;
    return SPPM_OK;
}
int sppm_pass_build(sppm_ctx* c, int32_t* stored) {
    int rc = check_own_pass(c, "sppm_pass_build");
    if (rc) return rc;
    if (c->graph_exec) return fail(SPPM_E_STATE, "split pass steps are not available after sppm_graph_capture");
    c->t_build.start(c->stream);
    do_build(c, c->stream);
    c->t_build.stop(c->stream);
    HIPOK(hipGetLastError());
    if (stored) {
        HIPOK(hipMemcpyAsync(stored, c->rb.nstored, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPOK(hipStreamSynchronize(c->stream));
    }
    return SPPM_OK;
}
int sppm_pass_eye(sppm_ctx* c, int step) {
    int rc = check_own_pass(c, "sppm_pass_eye");
    if (rc) return rc;
    if (c->graph_exec) return fail(SPPM_E_STATE, "split pass steps are not available after sppm_graph_capture");
    rc = do_eye(c, step, c->stream, &c->t_eye);
    if (rc) return rc;
    HIPOK(hipGetLastError());
    return SPPM_OK;
}
int sppm_pass_gather(sppm_ctx* c) {
    int rc = check_own_pass(c, "sppm_pass_gather");
    if (rc) return rc;
    if (c->graph_exec) return fail(SPPM_E_STATE, "split pass steps are not available after sppm_graph_capture");
    c->t_gather.start(c->stream);
    rc = do_gather(c, c->stream);
    if (rc) return rc;
    c->t_gather.stop(c->stream);
    if (c->guided) do_guide(c, c->stream, c->split_step, nullptr);
    HIPOK(hipGetLastError());
    return SPPM_OK;
}


// Per-pass statistics (synchronises the stream).
static int read_pass_stats(sppm_ctx* c, sppm_pass_stats* stats) {
    hipStream_t st = c->stream;
    memset(stats, 0, sizeof *stats);
    int stored = 0;
    uint32_t flags = 0;
    unsigned long long rs[2] = {0, 0};
    HIPOK(hipMemcpyAsync(&stored, c->rb.nstored, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(&flags, c->rb.flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(rs, c->rb.rstat, sizeof(rs), hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    for (Timer* t : {&c->t_photons, &c->t_build, &c->t_eye, &c->t_gather}) t->resolve();
    float r = rs[1] ? (float)(((double)rs[0] / 4294967296.0) / (double)rs[1]) : 0.0f;
    stats->photons_emitted = c->rb.N;
    stats->photons_stored = stored;
    stats->mean_radius = r;
    stats->hit_points = (int32_t)rs[1];
    stats->flags = (int)flags;
    stats->ms_photons = (float)c->t_photons.avg();
    stats->ms_build = (float)c->t_build.avg();
    stats->ms_eye = (float)c->t_eye.avg();
    stats->ms_gather = (float)c->t_gather.avg();
    if (c->params.check_numeric && (flags & 1u)) return fail(SPPM_E_NUMERIC, "NaN/Inf radiance after pass %d", c->pass);
    return SPPM_OK;
}

int sppm_pass(sppm_ctx* c, sppm_pass_stats* stats) {
    // passes in a row: this photon pass need not wait for the previous
    // build or gather (guided emission samples the proposal published right
    // before the previous build, not the one the previous pass trains)
    const bool lead = c && c->lead_ok;
    int rc = check_own_pass(c, "sppm_pass");
    if (rc) return rc;
    const int step = c->pass + 1 + c->params.step_offset;
    hipStream_t st = c->stream;
    if (c->graph_exec) {
        // the eager pass's streams and waits, each of its steps a graph launch
        const int b = (int)(c->graph_passes & 1);
        c->rb.hits = c->rb.hitbuf[b];
        if (lead) {
            HIPOK(hipStreamWaitEvent(c->stream2, c->ev_gathered[b], 0));
        } else {
            HIPOK(hipEventRecord(c->ev_gfork, st));
            HIPOK(hipStreamWaitEvent(c->stream2, c->ev_gfork, 0));
        }
        if (c->graph_passes == 0) HIPOK(hipGraphLaunch(c->x_words[0], c->stream2));  // the first pass's words
        HIPOK(hipGraphLaunch(c->x_eye[b], c->stream2));
        HIPOK(hipEventRecord(c->ev_gjoin, c->stream2));
        HIPOK(hipGraphLaunch(c->x_words[1 - b], c->stream2));  // the next pass's, beside the gather
        HIPOK(hipStreamWaitEvent(c->stream3,
                                 !lead ? c->ev_gfork : c->guided ? c->ev_published : c->ev_staged, 0));
        HIPOK(hipGraphLaunch(c->x_photon, c->stream3));
        if (lead) HIPOK(hipStreamWaitEvent(c->stream3, c->ev_built, 0));
        HIPOK(hipGraphLaunch(c->x_stage, c->stream3));
        HIPOK(hipEventRecord(c->ev_staged, c->stream3));
        HIPOK(hipEventRecord(c->ev_photons, c->stream3));
        HIPOK(hipStreamWaitEvent(st, c->ev_photons, 0));
        if (c->x_publish) {
            HIPOK(hipGraphLaunch(c->x_publish, st));
            HIPOK(hipEventRecord(c->ev_published, st));
        }
        HIPOK(hipGraphLaunch(c->x_build, st));
        HIPOK(hipEventRecord(c->ev_built, st));
        HIPOK(hipStreamWaitEvent(st, c->ev_gjoin, 0));
        HIPOK(hipGraphLaunch(b ? c->graph_exec_odd : c->graph_exec, st));
        HIPOK(hipEventRecord(c->ev_gathered[b], st));
        c->graph_passes += 1;
        c->pass += 1;
        c->lead_ok = true;
        return stats ? read_pass_stats(c, stats) : SPPM_OK;
    }
    // The eye pass runs on the side stream beside the photon pass and the
    // grid build (createHitpoints runs beside the kd-tree build in the
    // reference too, SPPM.cpp:809-838) into the hit-point buffer of this
    // pass's parity: in passes in a row it waits for the gather that last
    // read that buffer (two passes back), otherwise for what the main stream
    // has queued so far; then for this pass's camera words.  The next pass's
    // words follow it there.
    const int hb = (int)(c->pass & 1);
    c->rb.hits = c->rb.hitbuf[hb];
    if (lead) {
        HIPOK(hipStreamWaitEvent(c->stream2, c->ev_gathered[hb], 0));
    } else {
        HIPOK(hipEventRecord(c->ev_gfork, st));
        HIPOK(hipStreamWaitEvent(c->stream2, c->ev_gfork, 0));
    }
    rc = do_eye(c, step, c->stream2, &c->t_eye);
    if (rc) return rc;
    HIPOK(hipEventRecord(c->ev_gjoin, c->stream2));
    rc = words_generate_upto(c, c->words_used);
    if (rc) return rc;
    // The photon pass rewrites the photon slots, which the previous pass
    // staged; the previous gather reads only the grid built from them.  A
    // guided photon pass waits for the previous pass's publish instead: its
    // emission tables, and the path-cell block it rewrites (read by the
    // training records of the pass before).
    hipStream_t sp = st;
    if (lead) {
        sp = c->stream3;
        HIPOK(hipStreamWaitEvent(sp, c->guided ? c->ev_published : c->ev_staged, 0));
    }
    c->t_photons.start(sp);
    do_photons(c, step, sp, lead);
    c->t_photons.stop(sp);
    if (lead) {
        // the previous build has read the staged photons and their offsets
        // (and copied them for the gather's deposits and training records)
        HIPOK(hipStreamWaitEvent(sp, c->ev_built, 0));
        do_stage(c, sp);
        HIPOK(hipEventRecord(c->ev_staged, sp));
        HIPOK(hipEventRecord(c->ev_photons, sp));
        HIPOK(hipStreamWaitEvent(st, c->ev_photons, 0));
    }
    c->t_build.start(st);
    if (!lead) {
        do_stage(c, st);
        HIPOK(hipEventRecord(c->ev_staged, st));
    }
    if (c->guided) {
        do_publish(c, st);
        HIPOK(hipEventRecord(c->ev_published, st));
    }
    do_build_grid(c, st);
    c->t_build.stop(st);
    HIPOK(hipEventRecord(c->ev_built, st));
    HIPOK(hipStreamWaitEvent(st, c->ev_gjoin, 0));
    c->t_gather.start(st);
    rc = do_gather(c, st);
    if (rc) return rc;
    c->t_gather.stop(st);
    if (c->guided) {
        c->t_guide.start(st);
        do_guide(c, st, step, nullptr);
        c->t_guide.stop(st);
    }
    launch_add_i32(c->rb.pass_dev, 1, st);
    HIPOK(hipEventRecord(c->ev_gathered[hb], st));
    HIPOK(hipGetLastError());
    c->pass += 1;
    c->lead_ok = true;
    return stats ? read_pass_stats(c, stats) : SPPM_OK;
}

// Captures what body() queues on st into *g and instantiates it as *x.
static int capture_on(hipStream_t st, hipGraph_t* g, hipGraphExec_t* x, const std::function<int()>& body) {
    HIPOK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = body();
    hipGraph_t gg = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &gg);
    if (rc) {
        if (gg) (void)hipGraphDestroy(gg);
        return rc;
    }
    if (e != hipSuccess) return fail(SPPM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (hipGraphInstantiate(x, gg, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(gg);
        return fail(SPPM_E_HIP, "hipGraphInstantiate failed");
    }
    *g = gg;
    return SPPM_OK;
}

// ---- graph-captured passes -------------------------------------------------
// One pass as a capturable sequence (a batch branch): the pass's camera
// words and the eye pass on st2 (forked from st) while the photon pass and
// the build run on st, then, joined, the gather, guided training and the
// device pass counter.  Every size and step is read on the device, so a
// replay continues where the previous one ended; memory is cleared by
// kernels, not by memset nodes.
static int enqueue_pass_graph(sppm_ctx* c, hipStream_t st, hipStream_t st2) {
    RenderBuffers& R = c->rb;
    HIPOK(hipEventRecord(c->ev_gfork, st));
    HIPOK(hipStreamWaitEvent(st2, c->ev_gfork, 0));
    launch_camera_words(R.cam_states, R.cam_phi, R.cam_K, R.cam_L, R.words, st2);
    launch_eye(c->scene.dev, c->camdev, R.hits, 0, R.pass_dev, trace_spill_eye(c), R.words, R.flags, eye_rays(R),
               c->scene.next_path + STRIPE_INTS, c->params.eye_nee, st2);
    HIPOK(hipEventRecord(c->ev_gjoin, st2));
    photon_pass(c, 0, R.pass_dev, st, false);
    do_build(c, st);
    HIPOK(hipStreamWaitEvent(st, c->ev_gjoin, 0));
    int rc = do_gather(c, st);
    if (rc) return rc;
    if (c->guided) do_guide(c, st, 0, R.pass_dev);
    launch_add_i32(R.pass_dev, 1, st);
    return SPPM_OK;
}

static int graph_ok(sppm_ctx* c, const sppm_batch* member_of = nullptr) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->batch && c->batch != member_of) return not_in_batch(c, "graph capture");
    if (c->pass != 0 || c->words_gen != 0)
        return fail(SPPM_E_STATE, "capture a pass right after sppm_begin (pass %d done)", c->pass);
    if (c->cam.sampler != SPPM_SAMPLER_UNIFORM)
        return fail(SPPM_E_STATE, "graph-captured passes need the UniformSample camera stream");
    if (c->params.record_neighbours) return fail(SPPM_E_STATE, "graph-captured passes need record_neighbours 0");
    if (c->guided && c->gparams.train_interval > 1)
        return fail(SPPM_E_STATE, "graph-captured passes train after every pass (guide train_interval 1)");
    return SPPM_OK;
}

int sppm_graph_capture(sppm_ctx* c) {
    int rc = graph_ok(c);
    if (rc) return rc;
    if ((rc = sppm_sync(c))) return rc;
    RenderBuffers& R = c->rb;
    // The photon and eye streams count their passes from where the main one
    // is, one behind: their graphs add one first, so the counter's kernel
    // runs before the pass kernel (after it, it would wait for free wave
    // slots behind the gather's waves).
    const int before = c->pass + c->params.step_offset - 1;  // the main counter reads pass + step_offset
    HIPOK(hipMemcpy(R.pass_dev + 1, &before, sizeof(int), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(R.pass_dev + 2, &before, sizeof(int), hipMemcpyHostToDevice));
    const size_t blk = 6 * (size_t)R.np;
    if ((rc = capture_on(c->stream3, &c->g_photon, &c->x_photon, [&]() {
             launch_add_i32(R.pass_dev + 1, 1, c->stream3);
             photon_pass(c, 0, R.pass_dev + 1, c->stream3, true);
             return SPPM_OK;
         })) ||
        (rc = capture_on(c->stream3, &c->g_stage, &c->x_stage, [&]() {
             do_stage(c, c->stream3);
             return SPPM_OK;
         })) ||
        (c->guided && (rc = capture_on(c->stream, &c->g_publish, &c->x_publish, [&]() {
             do_publish(c, c->stream);
             return SPPM_OK;
         }))) ||
        (rc = capture_on(c->stream, &c->g_build, &c->x_build, [&]() {
             do_build_grid(c, c->stream);
             return SPPM_OK;
         }))) {
        drop_graph(c);
        return rc;
    }
    // even and odd passes: the gather of hit-point buffer b
    for (int b = 0; b < 2 && !rc; ++b) {
        R.hits = R.hitbuf[b];
        rc = capture_on(c->stream, b ? &c->graph_odd : &c->graph, b ? &c->graph_exec_odd : &c->graph_exec, [&]() {
            int r = do_gather(c, c->stream);
            if (!r && c->guided) do_guide(c, c->stream, 0, R.pass_dev);
            launch_add_i32(R.pass_dev, 1, c->stream);
            return r;
        });
    }
    R.hits = R.hitbuf[0];
    if (rc) {
        drop_graph(c);
        return rc;
    }
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = capture_on(c->stream2, &c->g_eye[b], &c->x_eye[b], [&]() {
            launch_add_i32(R.pass_dev + 2, 1, c->stream2);
            launch_eye(c->scene.dev, c->camdev, R.hitbuf[b], 0, R.pass_dev + 2, trace_spill_eye(c), R.words + b * blk,
                       R.flags, eye_rays(R), c->scene.next_path + STRIPE_INTS, c->params.eye_nee, c->stream2);
            return SPPM_OK;
        });
        if (!rc)
            rc = capture_on(c->stream2, &c->g_words[b], &c->x_words[b], [&]() {
                launch_camera_words(R.cam_states, R.cam_phi, R.cam_K, R.cam_L, R.words + b * blk, c->stream2);
                return SPPM_OK;
            });
    }
    if (rc) {
        drop_graph(c);
        return rc;
    }
    c->graph_passes = 0;
    return SPPM_OK;
}

}  // extern "C"


extern "C" {

// ---- autotuner ------------------------------------------------------------------
// Times `passes` passes (after one warm-up pass) for each BVH width {2, 4, 8}
// x path-kernel grid {device-filling, half}, keeps the fastest, and leaves the
// context re-begun with it.  The closest hit is width-independent, so the
// choice changes time only.
// ms per pass of `passes` passes after a warm-up pass, everything begun anew
static int tune_time(sppm_ctx* c, int passes, hipEvent_t e0, hipEvent_t e1, float* ms) {
    int rc;
    // warm-up pass, finished on every stream before the timed passes start
    // (the next photon pass may otherwise begin before e0)
    if ((rc = sppm_begin(c)) || (rc = sppm_pass(c, nullptr)) || (rc = sppm_sync(c))) return rc;
    c->lead_ok = false;
    HIPOK(hipEventRecord(e0, c->stream));
    for (int p = 0; p < passes; ++p)
        if ((rc = sppm_pass(c, nullptr))) return rc;
    HIPOK(hipEventRecord(e1, c->stream));
    HIPOK(hipEventSynchronize(e1));
    HIPOK(hipEventElapsedTime(ms, e0, e1));
    *ms /= (float)passes;
    return SPPM_OK;
}
int sppm_autotune(sppm_ctx* c, const sppm_scene_desc* scene, int passes, sppm_tune_result* out) {
    if (!c || !scene || !out) return fail(SPPM_E_ARG, "NULL argument");
    if (passes < 1) return fail(SPPM_E_ARG, "passes < 1");
    if (!c->have_cam || !c->have_params) return fail(SPPM_E_STATE, "camera and params must be set before tuning");
    if (int rc0 = not_in_batch(c, "sppm_autotune")) return rc0;
    c->lead_ok = false;
    memset(out, 0, sizeof *out);
    const int widths[3] = {2, 4, 8}, builders[2] = {0, SPPM_BVH_LBVH};
    float best = 3.4e38f;
    int bw = 8, bf = 0, rc = SPPM_OK;
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    const int kernel0 = c->params.photon_kernel;
    c->params.photon_kernel = SPPM_PHOTON_MEGAKERNEL;
    // stage 1: builder x width, device-filling grid, megakernel
    c->tuned_blocks = 0;
    for (int b = 0; b < 2 && !rc; ++b)
        for (int i = 0; i < 3 && !rc; ++i) {
            float ms = 0.0f;
            if ((rc = sppm_scene_upload(c, scene, widths[i] | builders[b]))) break;
            if ((rc = tune_time(c, passes, e0, e1, &ms))) break;
            out->trial_ms[b * 3 + i] = ms;
            if (ms < best) {
                best = ms;
                bw = widths[i];
                bf = builders[b];
            }
        }
    // stage 2: grid; stage 3: photon kernel
    int bb = 0, bk = SPPM_PHOTON_MEGAKERNEL;
    if (!rc) rc = sppm_scene_upload(c, scene, bw | bf);
    const int full = c->scene.spill_blocks;
    float gbest = 3.4e38f;
    for (int g = 0; g < 3 && !rc; ++g) {
        float ms = 0.0f;
        c->tuned_blocks = g == 0 ? 0 : (g == 1 ? full * 3 / 4 : full / 2);
        if ((rc = tune_time(c, passes, e0, e1, &ms))) break;
        out->grid_ms[g] = ms;
        if (ms < gbest) {
            gbest = ms;
            bb = c->tuned_blocks;
        }
    }
    c->tuned_blocks = bb;
    for (int k = 0; k < 2 && !rc; ++k) {
        float ms = 0.0f;
        c->params.photon_kernel = k == 0 ? SPPM_PHOTON_MEGAKERNEL : SPPM_PHOTON_WAVEFRONT;
        if ((rc = tune_time(c, passes, e0, e1, &ms))) break;
        out->kernel_ms[k] = ms;
        if (k == 0 || ms < out->kernel_ms[0]) bk = c->params.photon_kernel;
    }
    c->params.photon_kernel = bk;
    best = out->kernel_ms[bk == SPPM_PHOTON_MEGAKERNEL ? 0 : 1];
    // photons per pass: measured, reported, not applied
    const int n0 = c->params.photons_per_pass;
    const int cand[3] = {std::max(1, n0 / 2), n0,
                         (int)std::min<long long>(2LL * n0, RADIX_SORT_MAX_N / (NSLOT - 1))};
    float bestps = 0.0f;
    for (int q = 0; q < 3 && !rc; ++q) {
        float ms = 0.0f;
        c->params.photons_per_pass = cand[q];
        if ((rc = tune_time(c, passes, e0, e1, &ms))) break;
        out->photons_per_s[q] = ms > 0.0f ? (float)cand[q] / (ms * 1e-3f) : 0.0f;
        if (out->photons_per_s[q] > bestps) {
            bestps = out->photons_per_s[q];
            out->photons_suggested = cand[q];
        }
    }
    c->params.photons_per_pass = n0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        c->params.photon_kernel = kernel0;
        return rc;
    }
    c->params.bvh_width = bw;
    if ((rc = sppm_begin(c))) return rc;
    out->bvh_width = bw;
    out->bvh_builder = bf;
    out->photon_kernel = bk;
    out->grid_blocks = bb ? bb : c->scene.spill_blocks;
    out->ms_per_pass = best;
    return SPPM_OK;
}

int sppm_batch_create(sppm_batch** out, int device) {
    if (!out) return fail(SPPM_E_ARG, "out is NULL");
    HIPOK(hipSetDevice(device));
    sppm_batch* b = new sppm_batch();
    b->device = device;
    if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(SPPM_E_HIP, "batch stream / event creation failed");
    }
    *out = b;
    return SPPM_OK;
}

int sppm_batch_add(sppm_batch* b, sppm_ctx* c) {
    if (!b || !c) return fail(SPPM_E_ARG, "NULL argument");
    if (b->exec) return fail(SPPM_E_STATE, "batch already captured");
    if (c->device != b->device) return fail(SPPM_E_ARG, "context is on device %d, batch on %d", c->device, b->device);
    if (b->broken) return fail(SPPM_E_STATE, "a member of this batch was destroyed");
    int rc = graph_ok(c);
    if (rc) return rc;
    if (c->graph_exec) return fail(SPPM_E_STATE, "the context has its own captured pass");
    hipStream_t x = nullptr, y = nullptr;
    hipEvent_t j = nullptr;
    HIPOK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    HIPOK(hipStreamCreateWithFlags(&y, hipStreamNonBlocking));
    HIPOK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
    b->ctx.push_back(c);
    c->batch = b;
    b->s1.push_back(x);
    b->s2.push_back(y);
    b->join.push_back(j);
    return SPPM_OK;
}

int sppm_batch_mode(sppm_batch* b, int mode) {
    if (!b) return fail(SPPM_E_ARG, "NULL batch");
    if (b->exec) return fail(SPPM_E_STATE, "batch already captured");
    if (mode != SPPM_BATCH_MULTI && mode != SPPM_BATCH_BRANCHES && mode != SPPM_BATCH_AUTO)
        return fail(SPPM_E_ARG, "mode must be SPPM_BATCH_MULTI (0), SPPM_BATCH_BRANCHES (1) or SPPM_BATCH_AUTO (2)");
    b->mode = mode;
    return SPPM_OK;
}

// Block offsets of one multi-scene launch: grid blocks dealt to the scenes in
// proportion to their work (at least the scenes' order is kept; a scene may
// get none and is then reached by waves that finished their own).
static void deal_blocks(const std::vector<double>& work, int grid, std::vector<int>& out) {
    double tot = 0.0;
    for (double w : work) tot += w;
    out.push_back(0);
    double acc = 0.0;
    for (size_t k = 0; k < work.size(); ++k) {
        acc += work[k];
        int v = tot > 0.0 ? (int)std::llround((double)grid * acc / tot) : (int)((k + 1) * grid / work.size());
        v = std::min(std::max(v, out.back()), grid);
        out.push_back(k + 1 == work.size() ? grid : v);
    }
}

// The batch pass as one launch per phase for all scenes, pipelined over two
// streams like the eager passes: stream X runs the photon passes (graph
// "photons": per-scene counter clears, then one photon launch per group of
// scenes sharing the kernel's template arguments -- BVH width, guided
// emission; scenes on the wavefront kernels trace in their own branch) and
// the grid builds (graph "build", per scene); stream Y runs the camera words
// and the eye passes (graph "eye": two camera launches and one eye launch
// per width / NEE group) and the gathers (graph "gather": per-scene
// planning, one chunk launch, per-scene sums, training, pass counter).  A
// pass's build waits for the previous pass's gather (it rewrites the grid
// that gather reads) and its gather for its build, so the photon pass of
// pass p + 1 runs beside the gather of pass p.  Stream X counts its passes
// in its own device counters (the steps of the photon launches).
static int hip_step(hipError_t e, const char* what) {
    return e == hipSuccess ? SPPM_OK : fail(SPPM_E_HIP, "batch: %s: %s", what, hipGetErrorString(e));
}
// Per-scene work as concurrent graph branches from stream st: body(k, branch
// stream).  The scenes' setup kernels are small and latency bound; run side
// by side they share the device.
static int batch_branches(sppm_batch* b, hipStream_t st, const std::function<void(int, hipStream_t)>& body) {
    const int n = (int)b->ctx.size();
    int r = hip_step(hipEventRecord(b->fork, st), "fork");
    for (int k = 0; k < n && !r; ++k) {
        if ((r = hip_step(hipStreamWaitEvent(b->s1[k], b->fork, 0), "fork wait"))) break;
        body(k, b->s1[k]);
        r = hip_step(hipEventRecord(b->join[k], b->s1[k]), "branch end");
    }
    for (int k = 0; k < n && !r; ++k) r = hip_step(hipStreamWaitEvent(st, b->join[k], 0), "join");
    return r;
}

// The batch pass as one launch per phase for all scenes, in one graph:
//   per scene: clear the work counters;
//   one photon launch per group of scenes sharing the kernel's template
//     arguments (BVH width, guided emission), the megakernel for every scene;
//   per scene: the grid build;
//   the camera words of all scenes (two launches), one eye launch per BVH
//     width / NEE group;
//   per scene: the gather's planning;
//   one gather chunk launch;
//   per scene: the gather's sums, direct light, training, pass counter.
// The phases run one after the other: each of the big launches fills the
// device by itself (run side by side, measured, they only slowed each other
// and starved the small setup kernels).
static int prepare_multi(sppm_batch* b) {
    const int n = (int)b->ctx.size();
    std::vector<PhotonJob> pj;
    std::vector<EyeJob> ej;
    std::vector<GatherJob> gj;
    std::vector<CamJob> cj;
    std::vector<int> pkey, ekey, kst{0};
    int cap = 1;
    int rc;
    for (int k = 0; k < n; ++k) {
        sppm_ctx* c = b->ctx[k];
        RenderBuffers& R = c->rb;
        cap = std::max(cap, c->scene.stack_cap);
        {
            // every scene's photon pass joins the shared megakernel launch,
            // whatever photon kernel the context chose for its own passes
            // (the result is the same)
            PhotonJob J{};
            J.S = c->scene.dev;
            J.slots = R.slots;
            J.N = R.N;
            J.step_dev = R.pass_dev;
            J.next_path = c->scene.next_path;
            if (c->guided) J.G = c->emit;
            J.flags = R.flags;
            J.rays = photon_rays(R);
            pj.push_back(J);
            pkey.push_back(c->scene.dev.width * 2 + (c->guided ? 1 : 0));
        }
        EyeJob E{};
        E.S = c->scene.dev;
        E.C = c->camdev;
        E.H = R.hits;
        E.keep = c->params.eye_nee == SPPM_NEE_KEEP ? 1 : 0;
        E.step_dev = R.pass_dev;
        E.words = R.words;
        E.queue = c->scene.next_path + STRIPE_INTS;
        E.flags = R.flags;
        E.rays = eye_rays(R);
        ej.push_back(E);
        ekey.push_back(c->scene.dev.width * 2 + (c->params.eye_nee != SPPM_NEE_OFF ? 1 : 0));
        GatherJob J{};
        J.H = R.hits;
        J.SH = R.shared;
        J.G = grid_dev(c);
        J.gp = R.gp;
        J.order = gather_order(R.groups, R.np);
        J.nvalid = R.groups.nvalid;
        J.tiles = R.tiles;
        J.ntiles = gather_group_count(R.np);
        J.chunk_off = R.chunk_off;
        J.chunkp = R.chunkp;
        J.queue = R.groups.chunkq;
        J.partial = R.partial;
        J.pdep = c->guided ? R.pdep : nullptr;
        J.rec = PairRecord{R.pairs, R.npairs, R.pairs_cap};
        gj.push_back(J);
        cj.push_back(CamJob{R.cam_states, R.cam_phi, R.cam_K, R.cam_L, R.words});
        kst.push_back(kst.back() + R.cam_K);
    }
    // jobs sorted by group (stable: scenes keep their order inside a group),
    // grid blocks dealt to a group's scenes in proportion to their work
    std::vector<int> bst;
    auto plan = [&](const std::vector<int>& key, const std::vector<double>& work, int fixed_grid, std::vector<int>& o,
                    std::vector<BatchLaunch>& L) {
        o.resize(key.size());
        L.clear();
        for (size_t q = 0; q < o.size(); ++q) o[q] = (int)q;
        std::stable_sort(o.begin(), o.end(), [&](int a, int c) { return key[a] < key[c]; });
        for (size_t i = 0; i < o.size();) {
            size_t j = i;
            while (j < o.size() && key[o[j]] == key[o[i]]) ++j;
            const int grid = fixed_grid > 0 ? fixed_grid : trace_grid_blocks(key[o[i]] / 2);
            std::vector<double> w;
            for (size_t q = i; q < j; ++q) w.push_back(work[o[q]]);
            L.push_back(BatchLaunch{(int)i, (int)(j - i), key[o[i]], (int)bst.size(), grid});
            deal_blocks(w, grid, bst);
            i = j;
        }
    };
    std::vector<double> pw, ew;
    for (const PhotonJob& J : pj) pw.push_back((double)J.N);
    for (int k = 0; k < n; ++k) ew.push_back((double)b->ctx[k]->rb.np);
    std::vector<int> po, eo, go;
    plan(pkey, pw, 0, po, b->pl);
    plan(ekey, ew, 0, eo, b->el);
    plan(std::vector<int>(n, 0), ew, gather_wave_count(), go, b->gl);
    int grid_max = 1;
    for (const BatchLaunch& l : b->pl) grid_max = std::max(grid_max, l.grid);
    for (const BatchLaunch& l : b->el) grid_max = std::max(grid_max, l.grid);
    b->cap = cap;
    b->nseg = kst.back();
    dfree(b->pjobs); dfree(b->ejobs); dfree(b->gjobs); dfree(b->cjobs); dfree(b->bstart); dfree(b->kstart);
    dfree(b->spill);
    // the photon and eye launches run one after the other: one spill area
    const size_t per_lane = cap > TRACE_LDS ? (size_t)(cap - TRACE_LDS) : 1;
    if ((rc = dalloc(&b->pjobs, std::max<size_t>(pj.size(), 1))) || (rc = dalloc(&b->ejobs, ej.size())) ||
        (rc = dalloc(&b->gjobs, gj.size())) || (rc = dalloc(&b->cjobs, cj.size())) ||
        (rc = dalloc(&b->bstart, bst.size())) || (rc = dalloc(&b->kstart, kst.size())) ||
        (rc = dalloc(&b->spill, per_lane * (size_t)grid_max * 64)))
        return rc;
    std::vector<PhotonJob> pjs;
    for (int q : po) pjs.push_back(pj[q]);
    std::vector<EyeJob> ejs;
    for (int q : eo) ejs.push_back(ej[q]);
    std::vector<GatherJob> gjs;
    for (int q : go) gjs.push_back(gj[q]);
    if ((!pjs.empty() && hipMemcpy(b->pjobs, pjs.data(), sizeof(PhotonJob) * pjs.size(), hipMemcpyHostToDevice)) ||
        hipMemcpy(b->ejobs, ejs.data(), sizeof(EyeJob) * ejs.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(b->gjobs, gjs.data(), sizeof(GatherJob) * gjs.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(b->cjobs, cj.data(), sizeof(CamJob) * cj.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(b->kstart, kst.data(), sizeof(int) * kst.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(b->bstart, bst.data(), sizeof(int) * bst.size(), hipMemcpyHostToDevice))
        return fail(SPPM_E_HIP, "batch job tables upload");
    return SPPM_OK;
}
static int enqueue_multi(sppm_batch* b, hipStream_t st) {
    int r = batch_branches(b, st, [&](int k, hipStream_t bs) {
        launch_zero_i32(b->ctx[k]->scene.next_path, 2 * STRIPE_INTS, bs);
    });
    if (r) return r;
    for (const BatchLaunch& l : b->pl) {
        TraceSpill sp{b->spill, l.grid, b->cap};
        launch_photon_multi(l.key / 2, (l.key & 1) != 0, b->pjobs + l.first, l.count, b->bstart + l.boff, sp, st);
    }
    if ((r = batch_branches(b, st, [&](int k, hipStream_t bs) { do_build(b->ctx[k], bs); }))) return r;
    launch_camera_words_multi(b->cjobs, b->kstart, (int)b->ctx.size(), b->nseg, st);
    for (const BatchLaunch& l : b->el) {
        TraceSpill sp{b->spill, l.grid, b->cap};
        launch_eye_multi(l.key / 2, (l.key & 1) != 0, b->ejobs + l.first, l.count, b->bstart + l.boff, sp, st);
    }
    r = batch_branches(b, st, [&](int k, hipStream_t bs) {
        sppm_ctx* c = b->ctx[k];
        RenderBuffers& R = c->rb;
        launch_set_u64x2(R.rstat, 0ull, 0ull, bs);
        if (c->guided) launch_deposit_clear(R.pdep, R.doff + R.N, bs);
        const PairRecord rec{R.pairs, R.npairs, R.pairs_cap};
        (void)launch_gather_plan(R.hits, R.shared, grid_dev(c), R.gp, R.np, c->scene.bmin, c->scene.ext, R.groups,
                                 R.tiles, GATHER_CHUNK_CAP, R.chunkp, R.nch, R.chunk_off, R.tile_scan_tmp, R.counts,
                                 R.ctr, rec, bs);
    });
    if (r) return r;
    for (const BatchLaunch& l : b->gl) launch_gather_chunks_multi(b->gjobs + l.first, l.count, b->bstart + l.boff, st);
    return batch_branches(b, st, [&](int k, hipStream_t bs) {
        sppm_ctx* c = b->ctx[k];
        RenderBuffers& R = c->rb;
        launch_gather_finish(R.hits, R.shared, R.np, gather_order(R.groups, R.np), R.groups, c->params.alpha,
                             R.chunk_off, R.partial, R.counts, R.rstat, R.ctr, R.flags, bs);
        if (c->scene.dev.ibl_tex >= 0 || c->params.eye_nee != SPPM_NEE_OFF)
            launch_direct_light(R.hits, R.shared, R.np, bs);
        if (c->guided) do_guide(c, bs, 0, R.pass_dev);
        launch_add_i32(R.pass_dev, 1, bs);
    });
}
static int capture_multi(sppm_batch* b) {
    int rc = prepare_multi(b);
    if (rc) return rc;
    return capture_on(b->origin, &b->graph, &b->exec, [&]() { return enqueue_multi(b, b->origin); });
}

int sppm_batch_capture(sppm_batch* b) {
    if (!b || b->ctx.empty()) return fail(SPPM_E_ARG, "empty batch");
    if (b->exec) return fail(SPPM_E_STATE, "batch already captured");
    HIPOK(hipSetDevice(b->device));
    if (b->broken) return fail(SPPM_E_STATE, "a member of this batch was destroyed");
    for (sppm_ctx* c : b->ctx) {
        int rc = graph_ok(c, b);
        if (rc) return rc;
        HIPOK(hipStreamSynchronize(c->stream));
    }
    if (b->mode == SPPM_BATCH_AUTO) {
        double photons = 0.0;
        for (sppm_ctx* c : b->ctx) photons += (double)c->rb.N;
        b->mode = photons < (double)SPPM_BATCH_AUTO_PHOTONS * (double)b->ctx.size() ? SPPM_BATCH_MULTI
                                                                                       : SPPM_BATCH_BRANCHES;
    }
    if (b->mode == SPPM_BATCH_MULTI) return capture_multi(b);
    HIPOK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    int rc = SPPM_OK;
    if (hipEventRecord(b->fork, b->origin) != hipSuccess) rc = fail(SPPM_E_HIP, "fork record");
    for (size_t k = 0; k < b->ctx.size() && !rc; ++k) {
        if (hipStreamWaitEvent(b->s1[k], b->fork, 0) != hipSuccess) { rc = fail(SPPM_E_HIP, "fork wait"); break; }
        rc = enqueue_pass_graph(b->ctx[k], b->s1[k], b->s2[k]);
        if (rc) break;
        if (hipEventRecord(b->join[k], b->s1[k]) != hipSuccess ||
            hipStreamWaitEvent(b->origin, b->join[k], 0) != hipSuccess)
            rc = fail(SPPM_E_HIP, "join");
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->origin, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(SPPM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(SPPM_E_HIP, "hipGraphInstantiate failed");
    }
    b->graph = g;
    return SPPM_OK;
}

int sppm_batch_get_mode(sppm_batch* b, int32_t* mode) {
    if (!b || !mode) return fail(SPPM_E_ARG, "NULL argument");
    *mode = b->mode;
    return SPPM_OK;
}

int sppm_batch_pass(sppm_batch* b) {
    if (!b) return fail(SPPM_E_ARG, "NULL batch");
    if (b->broken) return fail(SPPM_E_STATE, "a member of this batch was destroyed");
    if (!b->exec) return fail(SPPM_E_STATE, "batch not captured");
    HIPOK(hipGraphLaunch(b->exec, b->origin));
    HIPOK(hipEventRecord(b->done, b->origin));
    b->launched = true;
    for (sppm_ctx* c : b->ctx) c->pass += 1;
    return SPPM_OK;
}

int sppm_batch_sync(sppm_batch* b) {
    if (!b) return fail(SPPM_E_ARG, "NULL batch");
    HIPOK(hipStreamSynchronize(b->origin));
    return SPPM_OK;
}

int sppm_batch_destroy(sppm_batch* b) {
    if (!b) return SPPM_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->origin);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        if (b->ctx[k] && b->ctx[k]->batch == b) b->ctx[k]->batch = nullptr;
        (void)hipStreamDestroy(b->s1[k]);
        (void)hipStreamDestroy(b->s2[k]);
        (void)hipEventDestroy(b->join[k]);
    }
    dfree(b->pjobs); dfree(b->ejobs); dfree(b->gjobs); dfree(b->cjobs); dfree(b->bstart); dfree(b->kstart);
    dfree(b->spill);
    (void)hipEventDestroy(b->fork);
    (void)hipEventDestroy(b->done);
    (void)hipStreamDestroy(b->origin);
    delete b;
    return SPPM_OK;
}

int sppm_counters(sppm_ctx* c, int64_t* hit_points, int64_t* neighbours, int64_t* photons_stored,
                  int64_t* candidates, int64_t* passes) {
    int rc = check_ready(c);
    if (rc) return rc;
    Counters h{};
    if ((rc = read_counters(c, h))) return rc;
    if (hit_points) *hit_points = (int64_t)h.hit_points;
    if (neighbours) *neighbours = (int64_t)h.neighbours;
    if (photons_stored) *photons_stored = (int64_t)h.stored;
    if (candidates) *candidates = (int64_t)h.candidates;
    if (passes) *passes = c->pass;
    return SPPM_OK;
}

int sppm_ray_counts(sppm_ctx* c, int64_t* photon_rays, int64_t* eye_rays) {
    int rc = check_ready(c);
    if (rc) return rc;
    Counters h{};
    if ((rc = read_counters(c, h))) return rc;
    if (photon_rays) *photon_rays = (int64_t)h.photon_rays;
    if (eye_rays) *eye_rays = (int64_t)h.eye_rays;
    return SPPM_OK;
}

int sppm_tie_counts(sppm_ctx* c, int64_t* photon_ties, int64_t* eye_ties) {
    int rc = check_ready(c);
    if (rc) return rc;
    Counters h{};
    if ((rc = read_counters(c, h))) return rc;
    if (photon_ties) *photon_ties = (int64_t)h.photon_ties;
    if (eye_ties) *eye_ties = (int64_t)h.eye_ties;
    return SPPM_OK;
}

int sppm_sync(sppm_ctx* c) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    if (int rc = batch_wait(c)) return rc;
    HIPOK(hipStreamSynchronize(c->stream3));
    HIPOK(hipStreamSynchronize(c->stream2));
    HIPOK(hipStreamSynchronize(c->stream));
    return SPPM_OK;
}

// ---- checkpoint / resume --------------------------------------------------------
// Everything a pass carries to the next: the photon slots (a path's stale slot
// is flattened again, SPPM.cpp:812-822), the shared hit points, the grid's
// radius statistic, the counters and flags, the camera stream (segment
// windows, pre-generated word blocks and their counters), the pass counter,
// and the learned proposal's weights, Adam state and histogram.  Layout: a
// header, then the arrays in the order of ckpt_sections.
namespace {
struct CkptHeader {
    char magic[8];
    int32_t version;
    int32_t N, W, H, np, nslot, cam_K;
    int64_t cam_L;
    int32_t guided, nparams, pass;
    int64_t words_gen, words_used;
    uint64_t setup;  // fingerprint of scene, camera, parameters and emission setup
};
// What decides the render besides the saved state: the scene, the camera,
// alpha / r0 / photons / eye NEE, and the guided-emission settings (the
// weights themselves are saved).  BVH width and recording do not change it.
uint64_t setup_fingerprint(const sppm_ctx* c) {
    uint64_t h = c->scene_hash;
    h = fnv1a(h, &c->cam, sizeof c->cam);
    h = fnv1a(h, &c->params.alpha, sizeof(float));
    h = fnv1a(h, &c->params.first_radius, sizeof(float));
    h = fnv1a(h, &c->params.photons_per_pass, sizeof(int32_t));
    h = fnv1a(h, &c->params.eye_nee, sizeof(int32_t));
    h = fnv1a(h, &c->params.step_offset, sizeof(int32_t));
    const int32_t g = c->guided ? 1 : 0;
    h = fnv1a(h, &g, sizeof g);
    if (c->guided) {
        const sppm_guide_params& q = c->gparams;
        h = fnv1a(h, &q.mix_plain, sizeof(float));
        h = fnv1a(h, &q.learning_rate, sizeof(float));
        h = fnv1a(h, &q.train_steps, sizeof(int32_t));
        h = fnv1a(h, &q.history_decay, sizeof(float));
        h = fnv1a(h, &q.hidden, 3 * sizeof(int32_t));  // hidden, layers, fp16
        h = fnv1a(h, &q.train_interval, sizeof(int32_t));
    }
    return h;
}
struct CkptSection {
    void* dev;
    size_t bytes;
};
}  // namespace

static std::vector<CkptSection> ckpt_sections(sppm_ctx* c) {
    RenderBuffers& R = c->rb;
    const size_t NS = (size_t)NSLOT * R.N;
    std::vector<CkptSection> v = {
        {R.slots.pos, sizeof(float4) * NS}, {R.slots.flux, sizeof(float4) * NS}, {R.slots.dir, sizeof(float4) * NS},
        {R.slots.valid, NS}, {R.shared.rad, sizeof(float4) * R.np}, {R.shared.cnt, sizeof(int2) * R.np},
        {R.shared.direct, sizeof(float4) * R.np},
        {R.rstat, 2 * sizeof(unsigned long long)}, {R.ctr, sizeof(Counters)},
        {R.raystripes, sizeof(unsigned long long) * STRIPES * RAY_STRIPE}, {R.flags, sizeof(uint32_t)},
        {R.pass_dev, sizeof(int)}, {R.words, sizeof(uint32_t) * 2 * 6 * (size_t)R.np}};
    if (R.cam_states) v.push_back({R.cam_states, sizeof(uint32_t) * (size_t)R.cam_K * MT_WORDS});
    if (c->guided) {
        GuideBuffers& g = c->guide;
        v.push_back({g.theta, sizeof(float) * g.nparams});
        v.push_back({g.m, sizeof(float) * g.nparams});
        v.push_back({g.v, sizeof(float) * g.nparams});
        v.push_back({g.hist, sizeof(float) * GUIDE_CELLS});
        v.push_back({g.hpass, sizeof(unsigned long long) * GUIDE_CELLS});
        v.push_back({g.eprob, sizeof(float) * GUIDE_CELLS});
        v.push_back({g.ecdf, sizeof(float) * GUIDE_CELLS});
        v.push_back({g.scal, sizeof(float) * 8});
    }
    return v;
}

static CkptHeader ckpt_header(sppm_ctx* c) {
    CkptHeader h{};
    memcpy(h.magic, "SPPMCKPT", 8);
    h.version = 3;
    h.setup = setup_fingerprint(c);
    h.N = c->rb.N; h.W = c->rb.W; h.H = c->rb.H; h.np = c->rb.np; h.nslot = NSLOT;
    h.cam_K = c->rb.cam_K; h.cam_L = c->rb.cam_L;
    h.guided = c->guided ? 1 : 0;
    h.nparams = c->guided ? c->guide.nparams : 0;
    h.pass = c->pass;
    h.words_gen = c->words_gen; h.words_used = c->words_used;
    return h;
}

int sppm_checkpoint_size(sppm_ctx* c, int64_t* bytes) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!bytes) return fail(SPPM_E_ARG, "bytes is NULL");
    size_t n = sizeof(CkptHeader);
    for (const CkptSection& s : ckpt_sections(c)) n += s.bytes;
    *bytes = (int64_t)n;
    return SPPM_OK;
}

int sppm_checkpoint_save(sppm_ctx* c, void* buf, int64_t bytes) {
    int64_t need = 0;
    int rc = sppm_checkpoint_size(c, &need);
    if (rc) return rc;
    if (!buf || bytes < need) return fail(SPPM_E_ARG, "checkpoint buffer: %lld bytes needed", (long long)need);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipStreamSynchronize(c->stream2));
    const CkptHeader h = ckpt_header(c);
    char* p = static_cast<char*>(buf);
    memcpy(p, &h, sizeof h);
    p += sizeof h;
    for (const CkptSection& s : ckpt_sections(c)) {
        HIPOK(hipMemcpy(p, s.dev, s.bytes, hipMemcpyDeviceToHost));
        p += s.bytes;
    }
    return SPPM_OK;
}

int sppm_checkpoint_load(sppm_ctx* c, const void* buf, int64_t bytes) {
    int64_t need = 0;
    int rc = sppm_checkpoint_size(c, &need);
    if (rc) return rc;
    if (!buf || bytes < (int64_t)sizeof(CkptHeader)) return fail(SPPM_E_ARG, "checkpoint buffer too small");
    CkptHeader h;
    memcpy(&h, buf, sizeof h);
    const CkptHeader mine = ckpt_header(c);
    if (memcmp(h.magic, "SPPMCKPT", 8) || h.version != 3) return fail(SPPM_E_ARG, "not a checkpoint of this format");
    if (h.N != mine.N || h.W != mine.W || h.H != mine.H || h.np != mine.np || h.nslot != mine.nslot ||
        h.cam_K != mine.cam_K || h.cam_L != mine.cam_L || h.guided != mine.guided || h.nparams != mine.nparams)
        return fail(SPPM_E_STATE, "checkpoint was taken with another resolution, photon count, camera sampler or "
                    "emission setup");
    if (h.setup != mine.setup)
        return fail(SPPM_E_STATE, "checkpoint was taken with another scene, camera, alpha / radius / NEE setting or "
                    "guided-emission setting");
    if (bytes < need) return fail(SPPM_E_ARG, "checkpoint buffer: %lld bytes needed", (long long)need);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipStreamSynchronize(c->stream2));
    const char* p = static_cast<const char*>(buf) + sizeof h;
    for (const CkptSection& s : ckpt_sections(c)) {
        HIPOK(hipMemcpy(s.dev, p, s.bytes, hipMemcpyHostToDevice));
        p += s.bytes;
    }
    c->pass = h.pass;
    c->words_gen = h.words_gen;
    c->words_used = h.words_used;
    c->eye_recorded[0] = c->eye_recorded[1] = false;  // both streams are idle
    c->rb.have_lists = false;
    if (c->guided) {
        guide_restore(c->guide, c->stream);
        HIPOK(hipStreamSynchronize(c->stream));
    }
    return SPPM_OK;
}

int sppm_kernel_ms(sppm_ctx* c, const char* name, float* ms) {
    if (!c || !name || !ms) return fail(SPPM_E_ARG, "NULL argument");
    Timer* t = nullptr;
    if (!strcmp(name, "reset")) {
        for (Timer* x : {&c->t_photons, &c->t_build, &c->t_eye, &c->t_gather, &c->t_guide}) x->reset();
        *ms = 0.0f;
        return SPPM_OK;
    }
    if (!strcmp(name, "photons")) t = &c->t_photons;
    else if (!strcmp(name, "build")) t = &c->t_build;
    else if (!strcmp(name, "eye")) t = &c->t_eye;
    else if (!strcmp(name, "gather")) t = &c->t_gather;
    else if (!strcmp(name, "guide")) t = &c->t_guide;
    if (!t) return fail(SPPM_E_ARG, "unknown kernel name '%s'", name);
    t->resolve();
    *ms = (float)t->avg();
    return SPPM_OK;
}

int sppm_reconstruct(sppm_ctx* c, int64_t total_photons, int32_t passes, float* rgba) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!rgba) return fail(SPPM_E_ARG, "rgba is NULL");
    launch_reconstruct(c->rb.shared, c->rb.np, (long long)total_photons, passes, c->rb.image, c->stream);
    HIPOK(hipGetLastError());
    HIPOK(hipMemcpyAsync(rgba, c->rb.image, sizeof(float4) * c->rb.np, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return SPPM_OK;
}

// ---- debug readback ---------------------------------------------------------
int sppm_read_photons(sppm_ctx* c, int32_t cap, float* pos, float* flux, float* indir, float* index, float* pdf,
                      int32_t* count) {
    int rc = check_ready(c);
    if (rc) return rc;
    RenderBuffers& R = c->rb;
    int n = 0;
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(&n, R.nstored, sizeof(int), hipMemcpyDeviceToHost));
    if (count) *count = n;
    if (!pos) return SPPM_OK;
    if (n > cap) return fail(SPPM_E_ARG, "capacity %d < %d photons", cap, n);
    float4 *fp = nullptr, *ff = nullptr, *fd = nullptr;
    if ((rc = dalloc(&fp, n)) || (rc = dalloc(&ff, n)) || (rc = dalloc(&fd, n))) {
        dfree(fp); dfree(ff); dfree(fd);
        return rc;
    }
    launch_photon_flat(R.slots, R.off, fp, ff, fd, c->stream);
    std::vector<float4> hp(n), hf(n), hd(n);
    HIPOK(hipMemcpyAsync(hp.data(), fp, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipMemcpyAsync(hf.data(), ff, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipMemcpyAsync(hd.data(), fd, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    dfree(fp); dfree(ff); dfree(fd);
    for (int i = 0; i < n; ++i) {
        pos[3 * i] = hp[i].x; pos[3 * i + 1] = hp[i].y; pos[3 * i + 2] = hp[i].z;
        flux[3 * i] = hf[i].x; flux[3 * i + 1] = hf[i].y; flux[3 * i + 2] = hf[i].z;
        indir[3 * i] = hd[i].x; indir[3 * i + 1] = hd[i].y; indir[3 * i + 2] = hd[i].z;
        index[i] = hp[i].w;
        pdf[i] = hf[i].w;
    }
    return SPPM_OK;
}

int sppm_read_hitpoints(sppm_ctx* c, int32_t* type, int32_t* facet, float* pos, float* weight, float* outdir,
                        float* normal) {
    int rc = check_ready(c);
    if (rc) return rc;
    RenderBuffers& R = c->rb;
    const int np = R.np;
    std::vector<float4> p(np), w(np), o(np), n(np);
    std::vector<int2> m(np);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(p.data(), R.hits.pos, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(w.data(), R.hits.weight, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(o.data(), R.hits.out, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(n.data(), R.hits.nrm, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(m.data(), R.hits.mat, sizeof(int2) * np, hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) {
        if (type) type[i] = (int)p[i].w;
        if (facet) facet[i] = m[i].y;
        if (pos) { pos[3 * i] = p[i].x; pos[3 * i + 1] = p[i].y; pos[3 * i + 2] = p[i].z; }
        if (weight) { weight[3 * i] = w[i].x; weight[3 * i + 1] = w[i].y; weight[3 * i + 2] = w[i].z; }
        if (outdir) { outdir[3 * i] = o[i].x; outdir[3 * i + 1] = o[i].y; outdir[3 * i + 2] = o[i].z; }
        if (normal) { normal[3 * i] = n[i].x; normal[3 * i + 1] = n[i].y; normal[3 * i + 2] = n[i].z; }
    }
    return SPPM_OK;
}

int sppm_read_shared(sppm_ctx* c, float* radiance, float* radius, int32_t* num_photon, int32_t* num_gather) {
    int rc = check_ready(c);
    if (rc) return rc;
    RenderBuffers& R = c->rb;
    const int np = R.np;
    std::vector<float4> r(np);
    std::vector<int2> k(np);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(r.data(), R.shared.rad, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(k.data(), R.shared.cnt, sizeof(int2) * np, hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) {
        if (radiance) { radiance[3 * i] = r[i].x; radiance[3 * i + 1] = r[i].y; radiance[3 * i + 2] = r[i].z; }
        if (radius) radius[i] = r[i].w;
        if (num_photon) num_photon[i] = k[i].x;
        if (num_gather) num_gather[i] = k[i].y;
    }
    return SPPM_OK;
}

int sppm_read_direct(sppm_ctx* c, float* hit_direct, float* shared_direct, int32_t* num_direct) {
    int rc = check_ready(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->stream));
    const size_t np = (size_t)c->rb.np;
    std::vector<float4> h(np), sh(np);
    HIPOK(hipMemcpy(h.data(), c->rb.hits.direct, sizeof(float4) * np, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(sh.data(), c->rb.shared.direct, sizeof(float4) * np, hipMemcpyDeviceToHost));
    const bool ibl = c->scene.dev.ibl_tex >= 0 || c->params.eye_nee != SPPM_NEE_OFF;
    for (size_t i = 0; i < np; ++i) {
        if (hit_direct) {  // written by the eye pass only with an environment light
            hit_direct[3 * i] = ibl ? h[i].x : 0.0f;
            hit_direct[3 * i + 1] = ibl ? h[i].y : 0.0f;
            hit_direct[3 * i + 2] = ibl ? h[i].z : 0.0f;
        }
        if (shared_direct) {
            shared_direct[3 * i] = sh[i].x;
            shared_direct[3 * i + 1] = sh[i].y;
            shared_direct[3 * i + 2] = sh[i].z;
        }
        if (num_direct) memcpy(&num_direct[i], &sh[i].w, sizeof(int32_t));  // NumDirectLight, int bits
    }
    return SPPM_OK;
}

int sppm_read_path_counts(sppm_ctx* c, int32_t* counts) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!counts) return fail(SPPM_E_ARG, "NULL output");
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(counts, c->rb.slots.count, sizeof(int) * c->rb.N, hipMemcpyDeviceToHost));
    return SPPM_OK;
}

int sppm_read_grid(sppm_ctx* c, int32_t* dims, float* geom, int64_t cell_cap, int32_t* cell_start,
                   int64_t photon_cap, int32_t* order) {
    int rc = check_ready(c);
    if (rc) return rc;
    RenderBuffers& R = c->rb;
    GridParams g{};
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(&g, R.gp, sizeof g, hipMemcpyDeviceToHost));
    if (dims) {
        dims[0] = g.nx; dims[1] = g.ny; dims[2] = g.nz; dims[3] = g.ncells; dims[4] = g.nphotons;
    }
    if (geom) {
        geom[0] = g.bmin[0]; geom[1] = g.bmin[1]; geom[2] = g.bmin[2]; geom[3] = g.cell;
    }
    if (cell_start) {
        if (cell_cap < (int64_t)g.ncells + 1) return fail(SPPM_E_ARG, "capacity %lld < %d cell starts",
                                                          (long long)cell_cap, g.ncells + 1);
        HIPOK(hipMemcpy(cell_start, R.cell_start, sizeof(int) * ((size_t)g.ncells + 1), hipMemcpyDeviceToHost));
    }
    if (order) {
        if (photon_cap < (int64_t)g.nphotons) return fail(SPPM_E_ARG, "capacity %lld < %d photons",
                                                          (long long)photon_cap, g.nphotons);
        std::vector<float4> d((size_t)g.nphotons);
        if (g.nphotons) HIPOK(hipMemcpy(d.data(), R.gdir, sizeof(float4) * d.size(), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < d.size(); ++j) {
            int32_t v;
            memcpy(&v, &d[j].w, sizeof v);  // the record carries its flattened index in dir.w
            order[j] = v;
        }
    }
    return SPPM_OK;
}

int sppm_read_counts(sppm_ctx* c, int32_t* m) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!m) return fail(SPPM_E_ARG, "NULL output");
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(m, c->rb.counts, sizeof(int) * c->rb.np, hipMemcpyDeviceToHost));
    return SPPM_OK;
}

int sppm_read_neighbours(sppm_ctx* c, int64_t cap, int64_t* offsets, int32_t* idx, int64_t* total) {
    int rc = check_ready(c);
    if (rc) return rc;
    RenderBuffers& R = c->rb;
    if (!R.have_lists) return fail(SPPM_E_STATE, "no neighbour lists (set params.record_neighbours and run a gather)");
    const int np = R.np;
    if (R.pairs) {
        // pairs of the production gather: grouped by hit point, sorted
        unsigned long long n = 0;
        HIPOK(hipStreamSynchronize(c->stream));
        HIPOK(hipMemcpy(&n, R.npairs, sizeof n, hipMemcpyDeviceToHost));
        if ((long long)n > R.pairs_cap)
            return fail(SPPM_E_OVERFLOW, "%llu neighbour pairs, more than max_neighbours %lld", n, R.pairs_cap);
        if (total) *total = (int64_t)n;
        if (!offsets) return SPPM_OK;
        if ((long long)n > cap) return fail(SPPM_E_ARG, "capacity %lld < %llu neighbours", (long long)cap, n);
        std::vector<int2> pr(n);
        if (n) HIPOK(hipMemcpy(pr.data(), R.pairs, sizeof(int2) * n, hipMemcpyDeviceToHost));
        std::sort(pr.begin(), pr.end(), [](const int2& a, const int2& b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
        std::fill(offsets, offsets + np + 1, (int64_t)0);
        for (size_t k = 0; k < n; ++k) {
            if (pr[k].x < 0 || pr[k].x >= np) return fail(SPPM_E_STATE, "recorded pair with hit point %d", pr[k].x);
            offsets[pr[k].x + 1] += 1;
            idx[k] = pr[k].y;
        }
        for (int i = 0; i < np; ++i) offsets[i + 1] += offsets[i];
        return SPPM_OK;
    }
    std::vector<int> cnt(np);
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(cnt.data(), R.counts, sizeof(int) * np, hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (int i = 0; i < np; ++i) tot += cnt[i];
    if (total) *total = tot;
    if (!offsets) return SPPM_OK;
    if (tot > cap) return fail(SPPM_E_ARG, "capacity %lld < %lld neighbours", (long long)cap, (long long)tot);
    offsets[0] = 0;
    for (int i = 0; i < np; ++i) offsets[i + 1] = offsets[i] + cnt[i];
    if (tot > 0) HIPOK(hipMemcpy(idx, R.lists, sizeof(int) * tot, hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) std::sort(idx + offsets[i], idx + offsets[i + 1]);
    return SPPM_OK;
}

int sppm_scene_stats(sppm_ctx* c, sppm_scene_info* info) {
    if (!c || !info) return fail(SPPM_E_ARG, "NULL argument");
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    const SceneBuffers& S = c->scene;
    info->nfacets = S.dev.ntri;
    info->nlights = S.dev.nlights;
    info->bvh_width = S.bvh.width;
    info->bvh_nodes = S.bvh.nnodes;
    info->bvh_depth = S.bvh.depth;
    info->stack_lds = S.stack_cap < TRACE_LDS ? S.stack_cap : TRACE_LDS;
    info->stack_global = S.stack_cap > TRACE_LDS ? S.stack_cap - TRACE_LDS : 0;
    return SPPM_OK;
}

// ---- learned emission proposal ------------------------------------------------
// He-uniform hidden layers from a splitmix64 stream, zero output layer: the
// first proposal is uniform over the cells.
static void guide_init_params(const GuideNet& net, uint32_t seed, std::vector<float>& th) {
    th.assign(guide_nparams(net), 0.0f);
    uint64_t x = 0x9E3779B97F4A7C15ull * ((uint64_t)seed + 1u);
    auto next = [&]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 1; l <= net.layers; ++l) {
        const double lim = sqrt(6.0 / guide_in(net, l));  // fan-in of the layer
        for (int k = 0; k < guide_in(net, l) * net.hidden; ++k)
            th[guide_w_off(net, l) + k] = (float)((2.0 * next() - 1.0) * lim);
    }
}

// The proposal's light frame: exactly one rectangle (two facets).  Untextured:
// equal radiance, so the plain emission is uniform on it.  Textured: both
// facets emit from their 1000-point tables.
static int guide_light_frame(sppm_ctx* c, GuideEmit& E) {
    const auto& Ls = c->scene.hlights;
    if (Ls.size() != 2 || Ls[0].textured != Ls[1].textured)
        return fail(SPPM_E_SCENE, "guided emission needs exactly one rectangular light (2 facets, both textured or "
                    "both not); scene has %d emitting facets", (int)Ls.size());
    std::vector<H::V3> v;
    for (const DevLight& L : Ls)
        for (int k = 0; k < 3; ++k) {
            H::V3 p = H::ld(L.p + 3 * k);
            bool dup = false;
            for (const H::V3& q : v) dup = dup || H::norm2(p - q) < 1e-12f;
            if (!dup) v.push_back(p);
        }
    if (v.size() != 4) return fail(SPPM_E_SCENE, "guided emission: the light's two facets do not form a quad");
    int a = 0, b = 1;
    float best = -1.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j)
            if (H::norm2(v[i] - v[j]) > best) { best = H::norm2(v[i] - v[j]); a = i; b = j; }
    int o[2], k = 0;
    for (int i = 0; i < 4; ++i)
        if (i != a && i != b) o[k++] = i;
    const H::V3 P0 = v[a];
    H::V3 e1 = v[o[0]] - P0, e2 = v[o[1]] - P0;
    const H::V3 ln = H::ld(Ls[0].normal);
    if (H::dot(H::cross(e1, e2), ln) < 0.0f) std::swap(e1, e2);
    const float l1 = sqrtf(H::norm2(e1)), l2 = sqrtf(H::norm2(e2));
    if (fabsf(H::dot(e1, e2)) > 1e-4f * l1 * l2) return fail(SPPM_E_SCENE, "guided emission: the light is not a rectangle");
    const H::V3 n = H::normalized(H::cross(e1, e2));
    if (H::dot(n, ln) < 0.999f) return fail(SPPM_E_SCENE, "guided emission: light normal differs from the facet normals");
    const float r0 = Ls[0].flux[0] / Ls[0].area, r1 = Ls[1].flux[0] / Ls[1].area;
    if (!Ls[0].textured && fabsf(r0 - r1) > 1e-5f * fabsf(r0))
        return fail(SPPM_E_SCENE, "guided emission: the light facets differ in radiance");
    E.textured = Ls[0].textured;
    const H::V3 t1 = e1 / l1, t2 = H::cross(n, t1);
    for (int q = 0; q < 3; ++q) {
        E.P0[q] = (&P0.x)[q]; E.e1[q] = (&e1.x)[q]; E.e2[q] = (&e2.x)[q];
        E.n[q] = (&n.x)[q]; E.t1[q] = (&t1.x)[q]; E.t2[q] = (&t2.x)[q];
        E.rad[q] = Ls[0].flux[q] / (2.0f * H::PI * Ls[0].area);  // as the plain emission recovers it
    }
    E.inv_e1 = 1.0f / H::norm2(e1);
    E.inv_e2 = 1.0f / H::norm2(e2);
    E.inv_area = 1.0f / (l1 * l2);
    return SPPM_OK;
}

// The plain emission's probability of each cell: the position cell's share of
// the light (uniform area, or the table points' plain probabilities for a
// textured light) times the cosine lobe's mass in the direction cell,
// integral of cos / pi d(omega) = pz / (2 pi |P|^4) ds dt over the cell of the
// hemi-octahedral map (32 x 32 midpoint samples per cell; the lobe's total is
// normalised to 1).
static void guide_plain_cells(const std::vector<float>& pos_mass, std::vector<float>& pc) {
    std::vector<double> dir((size_t)GUIDE_GS * GUIDE_GT, 0.0);
    const int Q = 32;
    double tot = 0.0;
    for (int is = 0; is < GUIDE_GS; ++is)
        for (int it = 0; it < GUIDE_GT; ++it) {
            double acc = 0.0;
            for (int a = 0; a < Q; ++a)
                for (int b = 0; b < Q; ++b) {
                    const double s = -1.0 + (is + (a + 0.5) / Q) * (2.0 / GUIDE_GS);
                    const double t = -1.0 + (it + (b + 0.5) / Q) * (2.0 / GUIDE_GT);
                    const double px = 0.5 * (s + t), py = 0.5 * (s - t);
                    const double pz = std::max(0.0, 1.0 - std::fabs(px) - std::fabs(py));
                    const double P2 = px * px + py * py + pz * pz;
                    acc += pz / (2.0 * 3.14159265358979323846 * P2 * P2);
                }
            dir[(size_t)is * GUIDE_GT + it] = acc * (2.0 / GUIDE_GS) * (2.0 / GUIDE_GT) / (Q * Q);
            tot += dir[(size_t)is * GUIDE_GT + it];
        }
    double mtot = 0.0;
    for (float m : pos_mass) mtot += m;
    pc.assign(GUIDE_CELLS, 0.0f);
    for (int p = 0; p < GUIDE_GU * GUIDE_GV; ++p)
        for (int d = 0; d < GUIDE_GS * GUIDE_GT; ++d)
            pc[(size_t)p * GUIDE_GS * GUIDE_GT + d] = (float)(pos_mass[p] / mtot * dir[d] / tot);
}

// Textured light: the table points grouped by position cell.  Point k of
// facet f has plain probability pi = P(pick f) * value_k / (1000 * average)
// (the pick and the point tables, LightSampler.cpp:68-76); its position cell
// is fixed here once, so both mixture components use the same cell.
static int guide_light_points(sppm_ctx* c, const GuideEmit& E, std::vector<float>& mass_out) {
    const auto& Ls = c->scene.hlights;
    const auto& S = c->scene;
    const int ncell = GUIDE_GU * GUIDE_GV;
    std::vector<int> pt_cell(S.h_tuvw.size(), -1);
    std::vector<std::vector<std::pair<int, int>>> cells(ncell);
    std::vector<std::vector<double>> probs(ncell);
    const H::V3 P0 = H::v3(E.P0[0], E.P0[1], E.P0[2]), e1 = H::v3(E.e1[0], E.e1[1], E.e1[2]),
                e2 = H::v3(E.e2[0], E.e2[1], E.e2[2]);
    for (int li = 0; li < 2; ++li) {
        const DevLight& L = Ls[li];
        const double pick = (double)L.flux_norm / ((double)S.h_lpick_avg * (double)Ls.size());
        const H::V3 p0 = H::ld(L.p), p1 = H::ld(L.p + 3), p2 = H::ld(L.p + 6);
        for (int k = 0; k < 1000; ++k) {
            const int j = L.tab_off + k;
            const float4 w = S.h_tuvw[j];
            const double pi = pick * (double)w.w / ((double)L.tab_avg * 1000.0);
            const H::V3 pos = (p0 * w.x + p1 * w.y) + p2 * (1.0f - w.x - w.y);
            const H::V3 rel = pos - P0;
            const float u = std::min(std::max(H::dot(rel, e1) * E.inv_e1, 0.0f), 0.99999994f);
            const float v = std::min(std::max(H::dot(rel, e2) * E.inv_e2, 0.0f), 0.99999994f);
            const int iu = std::min((int)(u * (float)GUIDE_GU), GUIDE_GU - 1);
            const int iv = std::min((int)(v * (float)GUIDE_GV), GUIDE_GV - 1);
            const int cell = iu * GUIDE_GV + iv;
            pt_cell[j] = cell;
            if (pi > 0.0) {
                cells[cell].push_back({li, j});
                probs[cell].push_back(pi);
            }
        }
    }
    std::vector<int> start(ncell + 1, 0);
    std::vector<int2> pts;
    std::vector<float> cum, mass(ncell, 0.0f);
    for (int cc = 0; cc < ncell; ++cc) {
        start[cc] = (int)pts.size();
        double run = 0.0;
        for (size_t q = 0; q < cells[cc].size(); ++q) {
            run += probs[cc][q];
            pts.push_back(make_int2(cells[cc][q].first, cells[cc][q].second));
            cum.push_back((float)run);
        }
        mass[cc] = cum.empty() || cells[cc].empty() ? 0.0f : cum.back();
    }
    start[ncell] = (int)pts.size();
    mass_out = mass;
    GuideBuffers& g = c->guide;
    dfree(g.pt_cell); dfree(g.pcell_mass); dfree(g.pc_start); dfree(g.pc_pt); dfree(g.pc_cum);
    g.pt_cell = nullptr; g.pcell_mass = nullptr; g.pc_start = nullptr; g.pc_pt = nullptr; g.pc_cum = nullptr;
    int rc = 0;
    if ((rc = dalloc(&g.pt_cell, pt_cell.size())) || (rc = dalloc(&g.pcell_mass, ncell)) ||
        (rc = dalloc(&g.pc_start, ncell + 1)) || (rc = dalloc(&g.pc_pt, std::max<size_t>(1, pts.size()))) ||
        (rc = dalloc(&g.pc_cum, std::max<size_t>(1, cum.size()))))
        return rc;
    HIPOK(hipMemcpy(g.pt_cell, pt_cell.data(), sizeof(int) * pt_cell.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(g.pcell_mass, mass.data(), sizeof(float) * ncell, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(g.pc_start, start.data(), sizeof(int) * (ncell + 1), hipMemcpyHostToDevice));
    if (!pts.empty()) {
        HIPOK(hipMemcpy(g.pc_pt, pts.data(), sizeof(int2) * pts.size(), hipMemcpyHostToDevice));
        HIPOK(hipMemcpy(g.pc_cum, cum.data(), sizeof(float) * cum.size(), hipMemcpyHostToDevice));
    }
    return SPPM_OK;
}

int sppm_guide_set(sppm_ctx* c, const sppm_guide_params* p) {
    if (!c || !p) return fail(SPPM_E_ARG, "NULL argument");
    c->lead_ok = false;
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    if (int rc0 = not_in_batch(c, "sppm_guide_set")) return rc0;
    HIPOK(hipSetDevice(c->device));
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipStreamSynchronize(c->stream2));
    HIPOK(hipStreamSynchronize(c->stream3));
    // a captured pass holds the proposal's buffers and the emission mode
    drop_graph(c);
    if (!p->enable) {
        c->guided = false;
        return SPPM_OK;
    }
    if (!(p->mix_plain >= 0.25f && p->mix_plain <= 1.0f)) return fail(SPPM_E_ARG, "mix_plain must be in [0.25, 1]");
    if (!(p->learning_rate > 0.0f) || p->train_steps < 0) return fail(SPPM_E_ARG, "bad learning rate / train steps");
    if (!(p->history_decay >= 0.0f && p->history_decay < 1.0f)) return fail(SPPM_E_ARG, "history_decay must be in [0, 1)");
    if (p->train_interval < 0) return fail(SPPM_E_ARG, "train_interval must be >= 0");
    if (p->target != SPPM_GUIDE_TARGET_MEAN && p->target != SPPM_GUIDE_TARGET_RMS)
        return fail(SPPM_E_ARG, "target must be SPPM_GUIDE_TARGET_MEAN (0) or SPPM_GUIDE_TARGET_RMS (1)");
    if (c->params.record_neighbours == 1)
        return fail(SPPM_E_STATE, "guided emission needs the production gather (record_neighbours 0 or 2)");
    GuideNet net;
    net.hidden = p->hidden ? p->hidden : 64;
    net.layers = p->layers ? p->layers : 4;
    net.fp16 = p->fp16 ? 1 : 0;
    if (!guide_net_ok(net)) return fail(SPPM_E_ARG, "network: hidden must be 64 or 128, layers 4 or 6");
    GuideEmit E{};
    int rc = guide_light_frame(c, E);
    if (rc) return rc;
    if (c->guide.theta && (c->guide.net.hidden != net.hidden || c->guide.net.layers != net.layers ||
                           c->guide.net.fp16 != net.fp16))
        guide_free(c->guide);
    c->guide.net = net;
    if (!c->guide.theta && guide_alloc(c->guide)) return fail(SPPM_E_NOMEM, "guide buffers");
    std::vector<float> pos_mass((size_t)GUIDE_GU * GUIDE_GV, 1.0f);  // plain probability of each position cell
    if (E.textured) {
        if ((rc = guide_light_points(c, E, pos_mass))) return rc;
        E.pt_cell = c->guide.pt_cell;
        E.pcell_mass = c->guide.pcell_mass;
        E.pc_start = c->guide.pc_start;
        E.pc_pt = c->guide.pc_pt;
        E.pc_cum = c->guide.pc_cum;
    } else {
        GuideBuffers& g = c->guide;
        dfree(g.pt_cell); dfree(g.pcell_mass); dfree(g.pc_start); dfree(g.pc_pt); dfree(g.pc_cum);
        g.pt_cell = nullptr; g.pcell_mass = nullptr; g.pc_start = nullptr; g.pc_pt = nullptr; g.pc_cum = nullptr;
    }
    {
        // the plain emission's probability of every cell (RMS target)
        std::vector<float> pc;
        guide_plain_cells(pos_mass, pc);
        GuideBuffers& g = c->guide;
        if ((!g.pcell && (rc = dalloc(&g.pcell, GUIDE_CELLS))) || (!g.tgt && (rc = dalloc(&g.tgt, GUIDE_CELLS))))
            return rc;
        HIPOK(hipMemcpy(g.pcell, pc.data(), sizeof(float) * GUIDE_CELLS, hipMemcpyHostToDevice));
    }
    std::vector<float> th;
    guide_init_params(net, p->seed, th);
    if (guide_upload_params(c->guide, th.data(), c->stream)) return fail(SPPM_E_HIP, "guide upload");
    guide_clear_hist(c->guide, c->stream);
    guide_infer(c->guide, c->stream);
    guide_publish(c->guide, c->stream);
    HIPOK(hipGetLastError());
    E.w = p->mix_plain;
    E.prob = c->guide.eprob;
    E.cdf = c->guide.ecdf;
    E.path_cell = c->rb.path_cell;
    c->emit = E;
    c->gparams = *p;
    c->guided = true;
    if (c->rb.ready && !c->rb.path_cell) {
        RenderBuffers& R = c->rb;
        if ((rc = dalloc(&R.path_cell, 2 * (size_t)R.N)) || (rc = dalloc(&R.path_ratio, 2 * (size_t)R.N)) ||
            (rc = dalloc(&R.pdep, R.photon_cap)) || (rc = dalloc(&R.doff, (size_t)R.N + 1)))
            return rc;
        c->emit.path_cell = R.path_cell;
    }
    c->emit.path_ratio = p->target == SPPM_GUIDE_TARGET_RMS ? c->rb.path_ratio : nullptr;
    HIPOK(hipStreamSynchronize(c->stream));
    return SPPM_OK;
}

static int guide_ready(sppm_ctx* c) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    c->lead_ok = false;
    if (!c->guide.theta) return fail(SPPM_E_STATE, "guided emission not set up (sppm_guide_set)");
    HIPOK(hipSetDevice(c->device));
    return batch_wait(c);
}

int sppm_guide_weights(sppm_ctx* c, float* out, const float* in) {
    int rc = guide_ready(c);
    if (rc) return rc;
    if (in) {
        if (guide_upload_params(c->guide, in, c->stream)) return fail(SPPM_E_HIP, "guide upload");
        guide_infer(c->guide, c->stream);
        guide_publish(c->guide, c->stream);
    }
    if (out) HIPOK(hipMemcpyAsync(out, c->guide.theta, sizeof(float) * c->guide.nparams, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return SPPM_OK;
}

int sppm_guide_gradient(sppm_ctx* c, const float* hist, float* loss, float* grad) {
    int rc = guide_ready(c);
    if (rc) return rc;
    if (hist) HIPOK(hipMemcpyAsync(c->guide.hist, hist, sizeof(float) * GUIDE_CELLS, hipMemcpyHostToDevice, c->stream));
    guide_hist_sum(c->guide, c->guide.hist, c->stream);
    guide_gradient(c->guide, c->guide.hist, c->stream);
    float sc[3];
    HIPOK(hipMemcpyAsync(sc, c->guide.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPOK(hipMemcpyAsync(grad, c->guide.grad, sizeof(float) * c->guide.nparams, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    if (loss) *loss = sc[2];
    return SPPM_OK;
}

int sppm_guide_train(sppm_ctx* c, const float* hist, int steps) {
    int rc = guide_ready(c);
    if (rc) return rc;
    if (steps < 0) return fail(SPPM_E_ARG, "steps < 0");
    if (hist) HIPOK(hipMemcpyAsync(c->guide.hist, hist, sizeof(float) * GUIDE_CELLS, hipMemcpyHostToDevice, c->stream));
    const float lr = c->gparams.learning_rate > 0.0f ? c->gparams.learning_rate : 1e-3f;
    guide_hist_sum(c->guide, c->guide.hist, c->stream);
    for (int k = 0; k < steps; ++k) {
        guide_gradient(c->guide, c->guide.hist, c->stream);
        guide_adam(c->guide, lr, c->stream);
    }
    guide_infer(c->guide, c->stream);
    guide_publish(c->guide, c->stream);
    HIPOK(hipGetLastError());
    HIPOK(hipStreamSynchronize(c->stream));
    return SPPM_OK;
}

int sppm_guide_tables(sppm_ctx* c, float* prob, float* cdf) {
    int rc = guide_ready(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->stream));
    if (prob) HIPOK(hipMemcpy(prob, c->guide.eprob, sizeof(float) * GUIDE_CELLS, hipMemcpyDeviceToHost));
    if (cdf) HIPOK(hipMemcpy(cdf, c->guide.ecdf, sizeof(float) * GUIDE_CELLS, hipMemcpyDeviceToHost));
    return SPPM_OK;
}

int sppm_guide_path_cells(sppm_ctx* c, int32_t* cells) {
    int rc = guide_ready(c);
    if (rc) return rc;
    if (!cells) return fail(SPPM_E_ARG, "NULL output");
    if (c->pass < 1) return fail(SPPM_E_STATE, "no pass rendered since sppm_begin");
    // the emission of step s writes block s & 1; the last pass ran step pass + step_offset
    const long long step = (long long)c->pass + c->params.step_offset;
    HIPOK(hipStreamSynchronize(c->stream));
    HIPOK(hipMemcpy(cells, c->rb.path_cell + (size_t)(step & 1) * (size_t)c->rb.N, sizeof(int) * (size_t)c->rb.N,
                    hipMemcpyDeviceToHost));
    return SPPM_OK;
}

int sppm_guide_read(sppm_ctx* c, float* prob, float* hist) {
    int rc = guide_ready(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->stream));
    if (prob) HIPOK(hipMemcpy(prob, c->guide.prob, sizeof(float) * GUIDE_CELLS, hipMemcpyDeviceToHost));
    if (hist) HIPOK(hipMemcpy(hist, c->guide.hist, sizeof(float) * GUIDE_CELLS, hipMemcpyDeviceToHost));
    return SPPM_OK;
}

int sppm_trace_closest(sppm_ctx* c, int32_t n, const float* org, const float* dir, float tmin, float tmax,
                       int32_t* facet, float* t, float* u, float* v) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    c->lead_ok = false;
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    if (int rc0 = batch_wait(c)) return rc0;  // the batch's passes use the traversal spill area
    if (n <= 0) return SPPM_OK;
    if (!org || !dir || !facet) return fail(SPPM_E_ARG, "NULL argument");
    HIPOK(hipSetDevice(c->device));
    float *dorg = nullptr, *ddir = nullptr, *dtuv = nullptr;
    int* dfac = nullptr;
    uint32_t* dfl = nullptr;
    int rc = 0;
    if ((rc = dalloc(&dorg, 3 * (size_t)n)) || (rc = dalloc(&ddir, 3 * (size_t)n)) || (rc = dalloc(&dtuv, 3 * (size_t)n)) ||
        (rc = dalloc(&dfac, n)) || (rc = dalloc(&dfl, 1))) {
        dfree(dorg); dfree(ddir); dfree(dtuv); dfree(dfac); dfree(dfl);
        return rc;
    }
    rc = trace_rays(c, n, org, dir, tmin, tmax, facet, t, u, v, dorg, ddir, dfac, dtuv, dfl);
    dfree(dorg); dfree(ddir); dfree(dtuv); dfree(dfac); dfree(dfl);
    return rc;
}

int sppm_trace_closest_static(sppm_ctx* c, int32_t n, const float* org, const float* dir, float tmin, float tmax,
                              int32_t* facet, float* t, float* u, float* v) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    c->lead_ok = false;
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    if (int rc0 = batch_wait(c)) return rc0;
    if (n <= 0) return SPPM_OK;
    if ((size_t)n > SCRATCH_RAYS)
        return fail(SPPM_E_ARG, "sppm_trace_closest_static: %d rays, more than the %d its fixed buffer holds "
                    "(split the batch, or use sppm_trace_closest)", n, (int)SCRATCH_RAYS);
    if (!org || !dir || !facet) return fail(SPPM_E_ARG, "NULL argument");
    HIPOK(hipSetDevice(c->device));
    char* base = nullptr;
    if (int rc0 = scratch_resolve(c->device, &base)) return rc0;
    float* dorg = reinterpret_cast<float*>(base);
    float* ddir = dorg + 3 * SCRATCH_RAYS;
    float* dtuv = ddir + 3 * SCRATCH_RAYS;
    int* dfac = reinterpret_cast<int*>(dtuv + 3 * SCRATCH_RAYS);
    uint32_t* dfl = reinterpret_cast<uint32_t*>(base + SCRATCH_FLAG_OFF);
    return trace_rays(c, n, org, dir, tmin, tmax, facet, t, u, v, dorg, ddir, dfac, dtuv, dfl);
}

int sppm_trace_occluded(sppm_ctx* c, int32_t n, const float* org, const float* dir, float tmin, const float* tmax,
                        float translucency, int32_t* occluded) {
    if (!c) return fail(SPPM_E_ARG, "NULL ctx");
    c->lead_ok = false;
    if (!c->scene.ready) return fail(SPPM_E_STATE, "no scene uploaded");
    if (int rc0 = batch_wait(c)) return rc0;  // the batch's passes use the traversal spill area
    if (n <= 0) return SPPM_OK;
    if (!org || !dir || !tmax || !occluded) return fail(SPPM_E_ARG, "NULL argument");
    HIPOK(hipSetDevice(c->device));
    float *dorg = nullptr, *ddir = nullptr, *dmax = nullptr;
    int* docc = nullptr;
    uint32_t* dfl = nullptr;
    int rc = 0;
    if ((rc = dalloc(&dorg, 3 * (size_t)n)) || (rc = dalloc(&ddir, 3 * (size_t)n)) || (rc = dalloc(&dmax, n)) ||
        (rc = dalloc(&docc, n)) || (rc = dalloc(&dfl, 1))) {
        dfree(dorg); dfree(ddir); dfree(dmax); dfree(docc); dfree(dfl);
        return rc;
    }
    hipStream_t st = c->stream;
    HIPOK(hipMemcpyAsync(dorg, org, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(ddir, dir, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(dmax, tmax, sizeof(float) * n, hipMemcpyHostToDevice, st));
    launch_set_u32(dfl, 0u, st);
    launch_occluded(c->scene.dev, n, dorg, ddir, tmin, dmax, translucency, c->scene.trace_spill(), docc, dfl, st);
    HIPOK(hipGetLastError());
    uint32_t fl = 0;
    HIPOK(hipMemcpyAsync(occluded, docc, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(&fl, dfl, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    dfree(dorg); dfree(ddir); dfree(dmax); dfree(docc); dfree(dfl);
    if (fl & 4u) return fail(SPPM_E_OVERFLOW, "traversal stack overflow or more than 256 translucent layers");
    return SPPM_OK;
}

}  // extern "C"
