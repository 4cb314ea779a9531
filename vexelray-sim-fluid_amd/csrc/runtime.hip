// The C ABI of libvexelhip.so (include/vexel/vxr.h): simulations as sets of named, resident
// device buffers on one HIP stream, and the pass lists that run over them.
//
// It stands where the reference's SupirVast runtime stands under the gui runners:
// PassRunner.gpu/run/write/read/clear/finish over a Buffered step (core/particle/FlipStep.java,
// core/particle/Flip3Step.java) and KernelHandle.dispatch over the rotated shallow-water
// bindings (gui/PatchSimulation.java:197-213).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/vexel/vxr.h"
#include "common.h"

namespace {

thread_local std::string g_error = "no error";

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(VXR_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

const char* kFields3[17] = {"x", "y", "z", "u", "v", "w", "m", "j", "c00", "c01", "c02", "c10", "c11", "c12", "c20",
                            "c21", "c22"};
const char* kGrid3[7] = {"gm", "gmu", "gmv", "gmw", "gu", "gv", "gw"};
const char* kFields2[10] = {"x", "y", "u", "v", "m", "j", "c00", "c01", "c10", "c11"};
const char* kGrid2[5] = {"gm", "gmu", "gmv", "gu", "gv"};

constexpr int64_t kChunk = 2048;  // particle arrays are padded to whole chunks

int64_t round_up(int64_t n, int64_t m) { return (n + m - 1) / m * m; }

int bits_for(int64_t values) {  // bits to hold 0 .. values-1
    int b = 0;
    while ((1LL << b) < values) b++;
    return b;
}

}  // namespace

namespace vxr {

struct Timer {
    std::string only;  // times this kernel alone when set
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<double, int64_t>> totals;
    hipEvent_t open = nullptr;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    int drain() {
        for (auto& p : pending) {
            float ms = 0;
            if (hipEventSynchronize(p.second.second) != hipSuccess) return 1;
            hipEventElapsedTime(&ms, p.second.first, p.second.second);
            auto& t = totals[p.first];
            t.first += ms;
            t.second += 1;
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        pending.clear();
        return 0;
    }
    ~Timer() {
        drain();
        for (hipEvent_t e : pool) hipEventDestroy(e);
    }
};

void timed_begin(const Launch& l, const char* kernel) {
    if (!l.timer) return;
    if (!l.timer->only.empty() && l.timer->only != kernel) return;
    l.timer->open = l.timer->take();
    if (l.timer->open) hipEventRecord(l.timer->open, l.stream);
}

void timed_end(const Launch& l, const char* kernel) {
    if (!l.timer || !l.timer->open) return;
    hipEvent_t stop = l.timer->take();
    if (!stop) return;
    hipEventRecord(stop, l.stream);
    l.timer->pending.push_back({kernel, {l.timer->open, stop}});
    l.timer->open = nullptr;
}

int check_launch(const char* kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, kernel);
    return 0;
}

}  // namespace vxr

struct Buf {
    void* ptr = nullptr;
    int64_t len = 0;    // logical words
    int64_t alloc = 0;  // allocated words
};

struct vxr_sim {
    int kind = 0;
    int nx = 0, ny = 0, nz = 0;
    int64_t count = 0, capacity = 0;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;
    vxr::Timer* timer = nullptr;
    std::vector<void*> allocations;

    // MPM
    std::vector<float*> p, p_alt;  // particle fields, and the buffers a sort permutes into
    std::vector<float*> g;         // grid fields
    int32_t* params = nullptr;
    int64_t param_words = 0;
    int32_t* keys = nullptr;
    int32_t* index = nullptr;
    int32_t* scratch = nullptr;
    int32_t* starts = nullptr;  // 2D: cells + 1
    int64_t nodes = 0;
    float *level = nullptr, *target = nullptr, *stay = nullptr, *view = nullptr;  // 3D level of detail
    int32_t* home = nullptr;                    // 3D: the position of each slot (p[17] is the slot of each particle)
    int32_t* active = nullptr;                  // 3D: particles before the first inactive one
    bool tension = false;                       // 3D surface tension (VXR_FLAG_TENSION)
    float* colour[2] = {};                      // c0, c1: the blurred colour, in turn
    float* stress[6] = {};                      // txx tyy tzz txy txz tyz
    uint8_t* near = nullptr;                    // the blur/stress region flags, one byte a brick
    // How many nodes c0 and c1 may have spread past the touched or live bricks of the last region
    // computation (one a blur; -1 unknown), and whether a pass since then wrote past the region.
    int spread[2] = {-1, -1};
    bool stale = true;
    uint8_t* touched = nullptr;  // 3D: bricks P2G reached, and bricks holding velocities
    uint8_t* live = nullptr;
    int64_t bricks = 0;
    // 3D P2G by layers (option p2g = 3): the tables of the last sort, which laid the particles out
    // layer by layer, and whether they still cover every particle with mass (no host write of
    // positions, masses or slots, no lod pass, no clear since)
    vxr::Layers layers{};
    bool layers_valid = false;
    // 3D P2G in fixed point (option p2g = 2): the 64-bit sums of gm gmu gmv gmw, made when the
    // option is first set
    unsigned long long* q3[4] = {};

    // SWE
    float* state[2][3] = {};
    int32_t* speed[3] = {};
    int32_t* budget[2] = {};
    int32_t* clamped = nullptr;
    int64_t dispatched = 0;

    // 2D scenes as a batch (VXR_KIND_MPM2_BATCH, or one scene whose P2G is the fixed-point one)
    std::vector<vxr::Scene2> scenes;
    vxr::Scene2* d_scenes = nullptr;
    int32_t *pscene = nullptr, *gscene = nullptr, *sscene = nullptr;
    int64_t pchunks = 0, gchunks = 0, schunks = 0;
    uint8_t *btouched = nullptr, *blive = nullptr;  // 2D fixed-point path: a flag each 64 nodes
    int32_t* sinfo = nullptr;    // batch: per starts chunk, what the layout fixes (4 words)
    int32_t* sbounds = nullptr;  // batch: per starts chunk, its first particle (the sort's scratch)
    unsigned long long* q[3] = {};
    int cell_bits = 0, scene_bits = 0;
    int64_t starts_len = 0;

    // ParticleSimulation.advance's step counter, and the graphs it replays
    int64_t steps = 0;
    bool flipped = false;  // whether p holds the second set of particle buffers (each sort swaps)
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;  // (sort first, steps, flipped)

    int32_t* diag = nullptr;  // diagnostics scratch, made on first use

    std::map<std::string, int64_t> options;

    vxr::Launch launch() const { return vxr::Launch{stream, timer}; }
};

namespace {

int alloc(vxr_sim* s, int64_t words, void** out) {
    void* ptr = nullptr;
    const int64_t bytes = words * 4 > 0 ? round_up(words * 4, 256) : 256;
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        return fail(e == hipErrorOutOfMemory ? VXR_ENOMEM : VXR_EDEVICE, "hipMalloc of %lld bytes: %s",
                    (long long)bytes, hipGetErrorString(e));
    }
    s->allocations.push_back(ptr);
    *out = ptr;
    return 0;
}

// Resolves a buffer name to its current address and logical length.
bool lookup(vxr_sim* s, const std::string& name, Buf* b) {
    if (s->kind == VXR_KIND_MPM2_BATCH) {
        for (int f = 0; f < 10; ++f) {
            if (name == kFields2[f]) {
                *b = Buf{s->p[f], s->capacity, s->capacity};
                return true;
            }
        }
        for (int f = 0; f < 5; ++f) {
            if (name == kGrid2[f]) {
                *b = Buf{s->g[f], s->nodes, s->nodes};
                return true;
            }
        }
        if (name == "params") *b = Buf{s->params, s->param_words, s->param_words};
        else if (name == "keys") *b = Buf{s->keys, s->capacity, s->capacity};
        else if (name == "starts") *b = Buf{s->starts, s->starts_len, s->starts_len};
        else return false;
        return true;
    }
    if (s->kind == VXR_KIND_MPM3 || s->kind == VXR_KIND_MPM2) {
        const bool three = s->kind == VXR_KIND_MPM3;
        const char** fields = three ? kFields3 : kFields2;
        const char** grid = three ? kGrid3 : kGrid2;
        const int nf = three ? 17 : 10, ng = three ? 7 : 5;
        for (int f = 0; f < nf; ++f) {
            if (name == fields[f]) {
                *b = Buf{s->p[f], s->count, s->capacity};
                return true;
            }
        }
        for (int f = 0; f < ng; ++f) {
            if (name == grid[f]) {
                *b = Buf{s->g[f], s->nodes, s->nodes};
                return true;
            }
        }
        if (name == "params") {
            *b = Buf{s->params, s->param_words, s->param_words};
            return true;
        }
        if (name == "keys") {
            *b = Buf{s->keys, s->count, s->capacity};
            return true;
        }
        if (three && (name == "level" || name == "target")) {
            const int64_t groups = s->count / 8;
            *b = Buf{name == "level" ? s->level : s->target, groups, groups};
            return true;
        }
        if (three && name == "slot") {
            *b = Buf{s->p[17], s->count, s->capacity};
            return true;
        }
        if (three && name == "active") {
            *b = Buf{s->active, 1, 1};
            return true;
        }
        if (three && name == "stay") {
            *b = Buf{s->stay, s->count, s->capacity};
            return true;
        }
        if (!three && s->tension) {
            const char* tn[5] = {"c0", "c1", "txx", "tyy", "txy"};
            for (int k = 0; k < 5; ++k) {
                if (name == tn[k]) {
                    *b = Buf{k < 2 ? s->colour[k] : s->stress[k - 2], s->nodes, s->nodes};
                    return true;
                }
            }
        }
        if (three && s->tension) {
            const char* tn[8] = {"c0", "c1", "txx", "tyy", "tzz", "txy", "txz", "tyz"};
            for (int k = 0; k < 8; ++k) {
                if (name == tn[k]) {
                    *b = Buf{k < 2 ? s->colour[k] : s->stress[k - 2], s->nodes, s->nodes};
                    return true;
                }
            }
        }
        if (three && name == "view") {
            *b = Buf{s->view, 5, 8};
            return true;
        }
        if (!three && name == "starts") {
            const int64_t cells = (int64_t)(s->nx - 1) * (s->ny - 1);
            *b = Buf{s->starts, cells + 1, cells + 1};
            return true;
        }
        return false;
    }
    const int64_t cells = (int64_t)s->nx * s->ny;
    const int slot2 = (int)(s->dispatched % 2), slot3 = (int)(s->dispatched % 3);
    if (name == "h") *b = Buf{s->state[slot2][0], cells, cells};
    else if (name == "hu") *b = Buf{s->state[slot2][1], cells, cells};
    else if (name == "hv") *b = Buf{s->state[slot2][2], cells, cells};
    else if (name == "params") *b = Buf{s->params, 5, 5};
    else if (name == "speed") *b = Buf{s->speed[slot3], 1, 1};
    else if (name == "budget") *b = Buf{s->budget[slot2], 1, 1};
    else if (name == "clamped") *b = Buf{s->clamped, 1, 1};
    else return false;
    return true;
}

// A node outside the tension region is at least 9 nodes from every touched or live brick: the
// colour may spread 8 nodes past them before a pass writing it must run dense, 7 before the
// stress (which reads one node further) must.
constexpr int kColourReach = 8;

vxr::Mpm3Args mpm3_args(vxr_sim* s) {
    vxr::Mpm3Args a;
    a.nx = s->nx;
    a.ny = s->ny;
    a.nz = s->nz;
    a.count = s->count;
    a.capacity = s->capacity;
    for (int f = 0; f < 17; ++f) a.p[f] = s->p[f];
    for (int f = 0; f < 7; ++f) a.g[f] = s->g[f];
    a.params = s->params;
    a.touched = s->touched;
    a.live = s->live;
    a.near = s->near;
    a.slot = (int32_t*)s->p[17];
    a.home = s->home;
    a.active = s->active;
    return a;
}

// Every brick flagged: the next clear and grid passes cover the whole grid (after the host wrote
// grid fields, or cleared everything). Never inside a captured graph.
int mark_all_bricks(vxr_sim* s) {
    if (s->btouched) {
        HIP_TRY(hipMemsetAsync(s->btouched, 1, s->gchunks * 4, s->stream));  // four flags a chunk
        HIP_TRY(hipMemsetAsync(s->blive, 1, s->gchunks * 4, s->stream));
    }
    if (!s->touched) return 0;
    HIP_TRY(hipMemsetAsync(s->touched, 1, s->bricks, s->stream));
    HIP_TRY(hipMemsetAsync(s->live, 1, s->bricks, s->stream));
    if (s->near) HIP_TRY(hipMemsetAsync(s->near, 3, s->bricks, s->stream));
    return 0;
}

vxr::Mpm2Args mpm2_args(vxr_sim* s) {
    vxr::Mpm2Args a;
    a.nx = s->nx;
    a.ny = s->ny;
    a.count = s->count;
    a.capacity = s->capacity;
    for (int f = 0; f < 10; ++f) a.p[f] = s->p[f];
    for (int f = 0; f < 5; ++f) a.g[f] = s->g[f];
    a.params = s->params;
    return a;
}

int64_t opt(vxr_sim* s, const char* key) { return s->options[key]; }

vxr::Batch2Args batch2_args(vxr_sim* s) {
    vxr::Batch2Args a;
    a.scenes = s->d_scenes;
    a.pscene = s->pscene;
    a.gscene = s->gscene;
    a.sscene = s->sscene;
    a.pchunks = s->pchunks;
    a.gchunks = s->gchunks;
    a.schunks = s->schunks;
    for (int f = 0; f < 10; ++f) a.p[f] = s->p[f];
    for (int f = 0; f < 5; ++f) a.g[f] = s->g[f];
    for (int f = 0; f < 3; ++f) a.q[f] = s->q[f];
    a.params = s->params;
    a.cell_bits = s->cell_bits;
    a.touched = s->btouched;
    a.live = s->blive;
    a.sinfo = s->sinfo;
    a.bounds = s->sbounds;
    return a;
}

bool fixed_p2g(vxr_sim* s) {
    return s->kind == VXR_KIND_MPM2_BATCH || (s->kind == VXR_KIND_MPM2 && s->options["p2g"] == 2);
}

int batch_sort_pass(vxr_sim* s) {
    const vxr::Launch l = s->launch();
    const vxr::Batch2Args a = batch2_args(s);
    const int blocks = (int)opt(s, "blocks");
    int rc = vxr::batch2_keys(l, a, s->keys, s->index, blocks);
    if (!rc) rc = vxr::radix_sort(l, s->keys, s->index, s->capacity, s->cell_bits + s->scene_bits,
                                  (int)opt(s, "radix_bits"), s->scratch);
    if (!rc) rc = vxr::permute_fields(l, s->p.data(), s->p_alt.data(), 10, s->index, s->capacity);
    if (rc) return rc;
    std::swap(s->p, s->p_alt);
    s->flipped = !s->flipped;
    return vxr::batch2_starts(l, batch2_args(s), s->keys, s->starts, blocks);
}

int sort_pass(vxr_sim* s) {
    if (s->kind == VXR_KIND_MPM2_BATCH) return batch_sort_pass(s);
    const vxr::Launch l = s->launch();
    const bool three = s->kind == VXR_KIND_MPM3;
    int rc = three ? vxr::mpm3_keys(l, mpm3_args(s), s->keys, s->index) : vxr::mpm2_keys(l, mpm2_args(s), s->keys, s->index);
    if (rc) return rc;
    int bits = 0;
    if (three) {
        bits = vxr::mpm3_key_bits(s->nx, s->ny, s->nz);
    } else {
        const int64_t keys = (int64_t)(s->nx - 1) * (s->ny - 1) + 1;
        while ((1LL << bits) < keys) bits++;
    }
    if ((rc = vxr::radix_sort(l, s->keys, s->index, s->count, bits, (int)opt(s, "radix_bits"), s->scratch))) return rc;
    const int nf = three ? 18 : 10;  // 3D: the 17 fields and the slot
    const bool by_layers = three && opt(s, "p2g") == 3;
    const int32_t* order = s->index;
    if (by_layers) {
        // the radix sort is done with its scratch: the new order's index and keys go there
        if ((rc = vxr::mpm3_layers(l, mpm3_args(s), s->keys, s->index, s->scratch, s->layers))) return rc;
        order = s->scratch;
    }
    if ((rc = vxr::permute_fields(l, s->p.data(), s->p_alt.data(), nf, order, s->count))) return rc;
    std::swap(s->p, s->p_alt);
    s->flipped = !s->flipped;
    if (three) s->layers_valid = by_layers;
    if (three && (rc = vxr::mpm3_after_sort(l, mpm3_args(s), s->keys))) return rc;
    if (!three) {
        const int32_t cells = (s->nx - 1) * (s->ny - 1);
        if ((rc = vxr::mpm2_starts(l, s->keys, s->count, cells, s->starts))) return rc;
    }
    return 0;
}

int run_pass(vxr_sim* s, const std::string& pass) {
    const vxr::Launch l = s->launch();
    if (pass == "sort") return sort_pass(s);
    if (s->kind == VXR_KIND_MPM3) {
        const vxr::Mpm3Args a = mpm3_args(s);
        const vxr::Fixed3 fx{{s->q3[0], s->q3[1], s->q3[2], s->q3[3]}};
        if (pass == "clear") return opt(s, "p2g") == 2 ? vxr::mpm3_clear_fixed(l, a, fx) : vxr::mpm3_clear(l, a);
        if (pass == "scatter") {
            const int64_t p2g = opt(s, "p2g");
            if (p2g == 0) return vxr::mpm3_scatter_atomic(l, a);
            if (p2g == 2) return vxr::mpm3_scatter_fixed(l, a, fx);
            if (p2g == 3 && s->layers_valid) return vxr::mpm3_scatter_layers(l, a, s->layers);
            return vxr::mpm3_scatter_tiled(l, a, (int)opt(s, "run"), opt(s, "p2g_ints") != 0);
        }
        if (pass == "grid") {
            if (!s->tension) return vxr::mpm3_grid(l, a);
            // the colour the last blur wrote: blur k writes c[k % 2], and there are six
            if (int rc = vxr::mpm3_grid_tension(l, a, s->stress, s->colour[1])) return rc;
            return vxr::mpm3_flags(l, a);
        }
        if (s->tension && pass.size() == 5 && pass.compare(0, 4, "blur") == 0 && pass[4] >= '0' && pass[4] <= '5') {
            // blur k reads gm (k = 0) or the colour blur k - 1 wrote, and writes c[k % 2]
            const int k = pass[4] - '0';
            int from = 0;
            if (k == 0) {
                if (int rc = vxr::tension3_region(l, a, s->stale)) return rc;
                s->stale = false;
            } else {
                from = s->spread[(k - 1) % 2];
            }
            const bool dense = from < 0 || from + 1 > kColourReach;
            s->spread[k % 2] = dense ? -1 : from + 1;
            if (dense) s->stale = true;
            const float* src = k == 0 ? a.g[0] : s->colour[(k - 1) % 2];
            return vxr::tension3_blur(l, a, src, s->colour[k % 2], k == 0, dense);
        }
        if (s->tension && pass == "stress") {
            const bool dense = s->spread[1] < 0 || s->spread[1] + 1 > kColourReach;
            if (dense) s->stale = true;
            return vxr::tension3_stress(l, a, s->colour[1], s->stress, dense);
        }
        if (pass == "lod") {
            s->layers_valid = false;  // it moves mass between slots
            return vxr::mpm3_lod(l, a, s->level, s->target, s->stay);
        }
        if (pass == "lod_target") return vxr::mpm3_lod_target(l, a, s->level, s->target, s->view);
        if (pass == "advect") {
            const int64_t g2p = opt(s, "g2p");
            return g2p == 0 ? vxr::mpm3_advect(l, a) : vxr::mpm3_advect_tiled(l, a, (int)opt(s, "run"), g2p == 2);
        }
    }
    if (s->kind == VXR_KIND_MPM2 && s->tension) {
        // FlipStep(..., tension)'s passes (core/particle/FlipStep.java:125-139): blur k reads gm
        // (k = 0) or the colour blur k - 1 wrote and writes c[k % 2]; the stress reads the last
        // colour; the grid update adds the capillary force. With the fixed-point P2G the sums are
        // decoded into gm gmu gmv before the first blur.
        const vxr::Mpm2Args a = mpm2_args(s);
        if (pass.size() == 5 && pass.compare(0, 4, "blur") == 0 && pass[4] >= '0' && pass[4] <= '5') {
            const int k = pass[4] - '0';
            if (k == 0 && fixed_p2g(s))
                if (int rc = vxr::batch2_decode(l, batch2_args(s), (int)opt(s, "blocks"))) return rc;
            const float* src = k == 0 ? a.g[0] : s->colour[(k - 1) % 2];
            return vxr::tension2_blur(l, a, src, s->colour[k % 2], k == 0);
        }
        if (pass == "stress") return vxr::tension2_stress(l, a, s->colour[1], s->stress);
        if (pass == "grid") {
            if (int rc = vxr::mpm2_grid_tension(l, a, s->stress, s->colour[1])) return rc;
            return fixed_p2g(s) ? vxr::batch2_flags(l, batch2_args(s)) : 0;
        }
    }
    if (s->kind != VXR_KIND_MPM3 && fixed_p2g(s)) {
        const vxr::Batch2Args b = batch2_args(s);
        const int blocks = (int)opt(s, "blocks");
        if (pass == "clear") return vxr::batch2_clear(l, b, blocks);
        if (pass == "scatter") return vxr::batch2_scatter(l, b, (int)opt(s, "run"), blocks);
        if (pass == "grid") return vxr::batch2_grid(l, b, blocks);
        if (pass == "advect")
            return s->kind == VXR_KIND_MPM2 ? vxr::mpm2_advect(l, mpm2_args(s)) : vxr::batch2_advect(l, b, blocks);
    } else if (s->kind == VXR_KIND_MPM2) {
        const vxr::Mpm2Args a = mpm2_args(s);
        if (pass == "clear") return vxr::mpm2_clear(l, a);
        if (pass == "scatter")
            return opt(s, "p2g") == 0 ? vxr::mpm2_scatter_atomic(l, a) : vxr::mpm2_scatter_runs(l, a, (int)opt(s, "run"));
        if (pass == "grid") return vxr::mpm2_grid(l, a);
        if (pass == "advect") return vxr::mpm2_advect(l, a);
    }
    return fail(VXR_EINVAL, "no pass named '%s' for this simulation", pass.c_str());
}

// ------------------------------------------------------------------ advance, graphs

const char* const kStep[4] = {"clear", "scatter", "grid", "advect"};

// Flip3Step's step with tension (core/particle/Flip3Step.java:88-109)
const char* const kTensionStep[10] = {"clear", "scatter", "blur0", "blur1", "blur2", "blur3", "blur4", "blur5",
                                      "stress", "grid"};

int run_step(vxr_sim* s) {
    if (s->tension) {
        for (const char* p : kTensionStep)
            if (int rc = run_pass(s, p)) return rc;
        return run_pass(s, "advect");
    }
    for (const char* p : kStep)
        if (int rc = run_pass(s, p)) return rc;
    return 0;
}
constexpr int kMaxSegment = 64;  // steps one captured graph holds at most

// A segment: a sort when `sort`, then `n` steps.
int run_segment(vxr_sim* s, bool sort, int n) {
    if (sort)
        if (int rc = sort_pass(s)) return rc;
    for (int i = 0; i < n; ++i)
        if (int rc = run_step(s)) return rc;
    return 0;
}

// With the "graph" option (and timing off) a segment is captured once into a hipGraph per
// (sort, n, which set of particle buffers is current) and replayed from then on: the same
// kernels with the same arguments, one launch for the lot. Every pass clears with kernels, so
// no memset node is recorded.
int replay_segment(vxr_sim* s, bool sort, int n) {
    if (!opt(s, "graph") || s->timer) return run_segment(s, sort, n);
    // The tension passes' region bookkeeping is host state, read while a segment is recorded: a
    // segment recorded while it is stale would replay whole-grid passes for good. Run eagerly
    // until a step has rebuilt the region.
    if (s->tension && s->kind == VXR_KIND_MPM3 && s->stale) return run_segment(s, sort, n);
    const auto key = std::make_tuple((int)sort, n, (int)s->flipped | (s->layers_valid ? 2 : 0));
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        const std::vector<float*> p0 = s->p, alt0 = s->p_alt;
        const bool flipped0 = s->flipped, layers0 = s->layers_valid;
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int rc = run_segment(s, sort, n);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        // capture ran nothing: put back the buffer sets the sort swapped while recording
        s->p = p0;
        s->p_alt = alt0;
        s->flipped = flipped0;
        s->layers_valid = layers0;
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
        it = s->graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, s->stream));
    if (sort) {
        std::swap(s->p, s->p_alt);
        s->flipped = !s->flipped;
        if (s->kind == VXR_KIND_MPM3) s->layers_valid = opt(s, "p2g") == 3;
    }
    return 0;
}

void drop_graphs(vxr_sim* s) {
    if (s->stream) hipStreamSynchronize(s->stream);
    for (auto& g : s->graphs) hipGraphExecDestroy(g.second);
    s->graphs.clear();
}

// ------------------------------------------------------------------ autotuning

// Milliseconds of `reps` runs of `body`, best of two rounds, on the simulation's stream.
template <class Body>
int time_ms(vxr_sim* s, int reps, Body body, float* out) {
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    float best = 3.0e38f;
    int rc = 0;
    for (int round = 0; round < 2 && !rc; ++round) {
        if (hipEventRecord(a, s->stream) != hipSuccess) rc = fail(VXR_EDEVICE, "hipEventRecord");
        for (int r = 0; r < reps && !rc; ++r) rc = body();
        if (!rc && hipEventRecord(b, s->stream) != hipSuccess) rc = fail(VXR_EDEVICE, "hipEventRecord");
        if (!rc && hipEventSynchronize(b) != hipSuccess) rc = fail(VXR_EDEVICE, "hipEventSynchronize");
        float ms = 0;
        if (!rc) hipEventElapsedTime(&ms, a, b);
        if (!rc && ms < best) best = ms;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    *out = best;
    return rc;
}

// Picks the fastest value of `key` among `values` for `body`, leaving the option set to it.
template <class Body>
int tune_option(vxr_sim* s, const char* key, std::initializer_list<int64_t> values, int reps, Body body) {
    int64_t best_v = opt(s, key);
    float best_ms = 3.0e38f;
    for (int64_t v : values) {
        s->options[key] = v;
        float ms = 0;
        if (int rc = time_ms(s, reps, body, &ms)) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best_v = v;
        }
    }
    s->options[key] = best_v;
    return 0;
}

// Times passes that leave the simulation as it was: clear + scatter + grid (they rewrite the
// grid from the particles), the sort's keys and radix passes on the key scratch (the particles
// keep their order), and for a patch the step into the spare slot, its clamp count restored.
int autotune(vxr_sim* s, int reps) {
    const vxr::Launch l = s->launch();
    auto scatter = [&]() -> int {
        for (const char* p : {"clear", "scatter", "grid"})
            if (int rc = run_pass(s, p)) return rc;
        return 0;
    };
    auto keys_and_sort = [&]() -> int {
        int rc = 0;
        int bits = 0;
        int64_t n = s->count;
        if (s->kind == VXR_KIND_MPM2_BATCH) {
            rc = vxr::batch2_keys(l, batch2_args(s), s->keys, s->index, (int)opt(s, "blocks"));
            bits = s->cell_bits + s->scene_bits;
            n = s->capacity;
        } else if (s->kind == VXR_KIND_MPM3) {
            rc = vxr::mpm3_keys(l, mpm3_args(s), s->keys, s->index);
            bits = vxr::mpm3_key_bits(s->nx, s->ny, s->nz);
        } else {
            rc = vxr::mpm2_keys(l, mpm2_args(s), s->keys, s->index);
            bits = bits_for((int64_t)(s->nx - 1) * (s->ny - 1) + 1);
        }
        if (rc) return rc;
        return vxr::radix_sort(l, s->keys, s->index, n, bits, (int)opt(s, "radix_bits"), s->scratch);
    };
    if (s->kind == VXR_KIND_SWE) {
        int32_t clamped = 0;
        HIP_TRY(hipMemcpy(&clamped, s->clamped, 4, hipMemcpyDeviceToHost));
        vxr::SweArgs a{};
        a.nx = s->nx;
        a.ny = s->ny;
        a.edges[0] = (s->flags & VXR_FLAG_OPEN_WEST) ? 1 : 0;
        a.edges[1] = (s->flags & VXR_FLAG_OPEN_EAST) ? 1 : 0;
        a.edges[2] = (s->flags & VXR_FLAG_OPEN_SOUTH) ? 1 : 0;
        a.edges[3] = (s->flags & VXR_FLAG_OPEN_NORTH) ? 1 : 0;
        a.params = s->params;
        a.clamped = s->clamped;
        const int64_t k = s->dispatched;
        for (int f = 0; f < 3; ++f) {
            a.in[f] = s->state[k % 2][f];
            a.out[f] = s->state[(k + 1) % 2][f];
        }
        a.speed_in = s->speed[k % 3];
        a.speed_out = s->speed[(k + 1) % 3];
        a.speed_clear = s->speed[(k + 2) % 3];
        a.budget_in = s->budget[k % 2];
        a.budget_out = s->budget[(k + 1) % 2];
        int rc = tune_option(s, "rows", {2, 4, 8, 16, 32, 48, 64, 128}, reps, [&]() {
            return vxr::swe_step(l, a, (int)opt(s, "rows"), opt(s, "hll_fast") != 0);
        });
        HIP_TRY(hipMemcpy(s->clamped, &clamped, 4, hipMemcpyHostToDevice));
        return rc;
    }
    int rc = 0;
    if (opt(s, "p2g") >= 1) rc = tune_option(s, "run", {1, 2, 4}, reps, scatter);
    if (!rc && fixed_p2g(s))
        rc = tune_option(s, "blocks", {0, 1024, 2048, 4096, 8192, 16384}, reps, [&]() {
            for (const char* p : kStep)
                if (p != std::string("advect"))
                    if (int r = run_pass(s, p)) return r;
            return 0;
        });
    if (!rc) rc = tune_option(s, "radix_bits", {4, 5, 6, 7, 8}, reps, keys_and_sort);
    HIP_TRY(hipStreamSynchronize(s->stream));
    return rc;
}

int clear_all(vxr_sim* s) {
    // hipMemsetAsync is fine here: these calls are never part of a captured graph.
    if (s->kind == VXR_KIND_SWE) {
        const int64_t cells = (int64_t)s->nx * s->ny;
        for (int k = 0; k < 2; ++k)
            for (int f = 0; f < 3; ++f) HIP_TRY(hipMemsetAsync(s->state[k][f], 0, cells * 4, s->stream));
        for (int k = 0; k < 3; ++k) HIP_TRY(hipMemsetAsync(s->speed[k], 0, 4, s->stream));
        for (int k = 0; k < 2; ++k) HIP_TRY(hipMemsetAsync(s->budget[k], 0, 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->clamped, 0, 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->params, 0, 5 * 4, s->stream));
        s->dispatched = 0;
    } else {
        for (float* f : s->p) HIP_TRY(hipMemsetAsync(f, 0, s->capacity * 4, s->stream));
        for (float* f : s->p_alt) HIP_TRY(hipMemsetAsync(f, 0, s->capacity * 4, s->stream));
        for (float* f : s->g) HIP_TRY(hipMemsetAsync(f, 0, round_up(s->nodes, 4) * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->params, 0, s->param_words * 4, s->stream));
        if (int rc = mark_all_bricks(s)) return rc;
        if (s->kind == VXR_KIND_MPM3) {
            if (int rc = vxr::mpm3_reset_slots(s->launch(), mpm3_args(s))) return rc;
            HIP_TRY(hipMemsetAsync(s->layers.meta, 0, 8 * 4, s->stream));
            s->layers_valid = false;
        }
        if (s->tension && s->kind == VXR_KIND_MPM3) {
            // what the full passes give a grid without mass: +0 colour and diagonal stress, -0
            // off-diagonal stress
            for (float* f : s->colour) HIP_TRY(hipMemsetAsync(f, 0, round_up(s->nodes, 4) * 4, s->stream));
            for (int k = 0; k < 6; ++k)
                HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)s->stress[k], k < 3 ? 0u : 0x80000000u,
                                          round_up(s->nodes, 4), s->stream));
            s->spread[0] = s->spread[1] = 0;
            s->stale = true;
        } else if (s->tension) {
            // 2D: the passes are dense and rewrite every node; zero, as the reference's clear leaves them
            for (float* f : s->colour) HIP_TRY(hipMemsetAsync(f, 0, round_up(s->nodes, 4) * 4, s->stream));
            for (int k = 0; k < 3; ++k) HIP_TRY(hipMemsetAsync(s->stress[k], 0, round_up(s->nodes, 4) * 4, s->stream));
        }
        for (unsigned long long* q : s->q3)
            if (q) HIP_TRY(hipMemsetAsync(q, 0, round_up(s->nodes, 4) * 8, s->stream));
        if (s->stay) {
            const int64_t groups = std::max<int64_t>(s->count / 8, 1);
            HIP_TRY(hipMemsetAsync(s->level, 0, groups * 4, s->stream));
            HIP_TRY(hipMemsetAsync(s->target, 0, groups * 4, s->stream));
            HIP_TRY(hipMemsetAsync(s->stay, 0, s->capacity * 4, s->stream));
            HIP_TRY(hipMemsetAsync(s->view, 0, 8 * 4, s->stream));
        }
        HIP_TRY(hipMemsetAsync(s->keys, 0, s->capacity * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->index, 0, s->capacity * 4, s->stream));
        if (s->starts) HIP_TRY(hipMemsetAsync(s->starts, 0, s->starts_len * 4, s->stream));
        for (unsigned long long* qf : s->q)
            if (qf) HIP_TRY(hipMemsetAsync(qf, 0, s->gchunks * 256 * 8, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

void destroy(vxr_sim* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    for (auto& g : s->graphs) hipGraphExecDestroy(g.second);
    delete s->timer;
    for (void* ptr : s->allocations) hipFree(ptr);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

struct OptionSpec {
    const char* key;
    int64_t lo, hi, def;
};
const OptionSpec kOptions[] = {
    {"p2g", 0, 3, 1},   {"g2p", 0, 2, 2},        {"run", 1, 4, 4},   {"radix_bits", 4, 8, 8},
    {"rows", 1, 256, 32}, {"blocks", 0, 1 << 20, 0}, {"graph", 0, 1, 0}, {"hll_fast", 0, 1, 0},
    {"p2g_ints", 0, 1, 1},
};



// Lays scenes out (particles padded to kBatchAlign, nodes and starts to 256) and fills the
// chunk counts; the device tables come later, in alloc_batch.
void layout_batch(vxr_sim* s) {
    int64_t po = 0, go = 0, so = 0;
    int64_t max_cells = 0;
    for (vxr::Scene2& sc : s->scenes) {
        sc.cells = (sc.nx - 1) * (sc.ny - 1);
        sc.poff = po;
        sc.goff = go;
        sc.soff = so;
        po += round_up(sc.count, vxr::kBatchAlign);
        go += round_up((int64_t)sc.nx * sc.ny, 256);
        so += round_up((int64_t)sc.cells + 1, 256);
        max_cells = std::max<int64_t>(max_cells, sc.cells);
    }
    s->pchunks = po / vxr::kBatchAlign;
    s->gchunks = go / 256;
    s->schunks = so / 256;
    s->cell_bits = bits_for(max_cells + 1);
    s->scene_bits = bits_for((int64_t)s->scenes.size());
}

int alloc_batch(vxr_sim* s) {
    void* ptr = nullptr;
    const int64_t nscenes = (int64_t)s->scenes.size();
    int rc = alloc(s, nscenes * (int64_t)(sizeof(vxr::Scene2) / 4), &ptr);
    if (rc) return rc;
    s->d_scenes = (vxr::Scene2*)ptr;
    std::vector<int32_t> pt(s->pchunks), gt(s->gchunks), st(s->schunks);
    for (int64_t k = 0; k < nscenes; ++k) {
        const vxr::Scene2& sc = s->scenes[k];
        const int64_t p1 = k + 1 < nscenes ? s->scenes[k + 1].poff : s->pchunks * vxr::kBatchAlign;
        const int64_t g1 = k + 1 < nscenes ? s->scenes[k + 1].goff : s->gchunks * 256;
        const int64_t s1 = k + 1 < nscenes ? s->scenes[k + 1].soff : s->schunks * 256;
        for (int64_t c = sc.poff / vxr::kBatchAlign; c < p1 / vxr::kBatchAlign; ++c) pt[c] = (int32_t)k;
        for (int64_t c = sc.goff / 256; c < g1 / 256; ++c) gt[c] = (int32_t)k;
        for (int64_t c = sc.soff / 256; c < s1 / 256; ++c) st[c] = (int32_t)k;
    }
    if ((rc = alloc(s, std::max<int64_t>(s->pchunks, 1), &ptr))) return rc;
    s->pscene = (int32_t*)ptr;
    if ((rc = alloc(s, std::max<int64_t>(s->gchunks, 1), &ptr))) return rc;
    s->gscene = (int32_t*)ptr;
    if ((rc = alloc(s, std::max<int64_t>(s->schunks, 1), &ptr))) return rc;
    s->sscene = (int32_t*)ptr;
    for (int f = 0; f < 3; ++f) {
        if ((rc = alloc(s, s->gchunks * 256 * 2, &ptr))) return rc;
        s->q[f] = (unsigned long long*)ptr;
    }
    // a byte for each 64 nodes: a word a chunk
    if ((rc = alloc(s, std::max<int64_t>(s->gchunks, 1), &ptr))) return rc;
    s->btouched = (uint8_t*)ptr;
    if ((rc = alloc(s, std::max<int64_t>(s->gchunks, 1), &ptr))) return rc;
    s->blive = (uint8_t*)ptr;
    if ((rc = alloc(s, std::max<int64_t>(s->schunks, 1), &ptr))) return rc;
    s->sbounds = (int32_t*)ptr;
    if ((rc = alloc(s, std::max<int64_t>(s->schunks, 1) * 4, &ptr))) return rc;
    s->sinfo = (int32_t*)ptr;
    std::vector<int32_t> info(std::max<int64_t>(s->schunks, 1) * 4, 0);
    for (int64_t k = 0; k < nscenes; ++k) {
        const vxr::Scene2& sc = s->scenes[k];
        for (int64_t first = 0; first <= sc.cells; first += 256) {
            int32_t* w = &info[(sc.soff + first) / 256 * 4];
            const bool last = first + 256 > sc.cells;
            w[0] = (int32_t)(((uint32_t)k << s->cell_bits) | (uint32_t)first);
            w[1] = (int32_t)sc.poff;
            w[2] = (int32_t)(last ? sc.cells + 1 - first : 256) | (last ? 1 << 16 : 0);
            w[3] = (int32_t)(sc.poff + sc.count);
        }
    }
    HIP_TRY(hipMemcpy(s->sinfo, info.data(), info.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->d_scenes, s->scenes.data(), nscenes * sizeof(vxr::Scene2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->pscene, pt.data(), pt.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->gscene, gt.data(), gt.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->sscene, st.data(), st.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

extern "C" {

const char* vxr_last_error(void) { return g_error.c_str(); }

int vxr_device_count(int* n) {
    if (!n) return fail(VXR_EINVAL, "n is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = e == hipSuccess ? c : 0;
    return VXR_OK;
}

int vxr_create(int kind, const int32_t dims[3], int64_t particles, uint32_t flags, vxr_sim** out) {
    if (!out || !dims) return fail(VXR_EINVAL, "null argument");
    *out = nullptr;
    const uint32_t physics = VXR_FLAG_TENSION | VXR_FLAG_HEAT | VXR_FLAG_CONVECTION | VXR_FLAG_RELAX | VXR_FLAG_PUMP |
                             VXR_FLAG_PLASTIC;
    const uint32_t built = kind == VXR_KIND_MPM3 || kind == VXR_KIND_MPM2 ? VXR_FLAG_TENSION : 0u;
    if (flags & physics & ~built)
        return fail(VXR_EUNSUPPORTED, "of the optional physics only surface tension is built (not heat, convection, "
                                      "relax, pump or plastic)");
    const uint32_t edges = VXR_FLAG_OPEN_WEST | VXR_FLAG_OPEN_EAST | VXR_FLAG_OPEN_SOUTH | VXR_FLAG_OPEN_NORTH;
    if (flags & ~(physics | edges)) return fail(VXR_EINVAL, "unknown flags 0x%x", flags & ~(physics | edges));

    vxr_sim* s = new vxr_sim;
    s->kind = kind;
    s->flags = flags;
    for (const OptionSpec& o : kOptions) s->options[o.key] = o.def;
    int rc = 0;
    if (kind == VXR_KIND_MPM2_BATCH) {
        delete s;
        return fail(VXR_EINVAL, "a batch of scenes is made by vxr_create_batch");
    }
    if (kind == VXR_KIND_SWE) {
        if (dims[0] < 1 || dims[1] < 1) {
            delete s;
            return fail(VXR_EINVAL, "a patch needs at least one cell each way, got %d x %d", dims[0], dims[1]);
        }
        s->nx = dims[0];
        s->ny = dims[1];
    } else if (kind == VXR_KIND_MPM3 || kind == VXR_KIND_MPM2) {
        const bool three = kind == VXR_KIND_MPM3;
        if (dims[0] < 4 || dims[1] < 4 || (three && dims[2] < 4)) {
            delete s;
            return fail(VXR_EINVAL, "a grid needs at least four nodes each way");
        }
        if (particles < 1) {
            delete s;
            return fail(VXR_EINVAL, "a step needs particles, got %lld", (long long)particles);
        }
        if (particles > (int64_t)INT32_MAX - kChunk) {
            delete s;
            return fail(VXR_EINVAL, "at most 2^31 - 2049 particles, got %lld", (long long)particles);
        }
        s->nx = dims[0];
        s->ny = dims[1];
        s->nz = three ? dims[2] : 1;
        s->count = particles;
        s->capacity = round_up(particles, kChunk);
        s->nodes = (int64_t)s->nx * s->ny * s->nz;
        if ((double)s->nodes > 2.0e9) {
            delete s;
            return fail(VXR_EINVAL, "grid of %lld nodes is too large", (long long)s->nodes);
        }
    } else {
        delete s;
        return fail(VXR_EINVAL, "unknown kind %d", kind);
    }

    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipStreamCreate");
    }
    void* ptr = nullptr;
    if (kind == VXR_KIND_SWE) {
        const int64_t cells = (int64_t)s->nx * s->ny;
        for (int k = 0; k < 2 && !rc; ++k)
            for (int f = 0; f < 3 && !rc; ++f) {
                rc = alloc(s, cells, &ptr);
                s->state[k][f] = (float*)ptr;
            }
        for (int k = 0; k < 3 && !rc; ++k) {
            rc = alloc(s, 1, &ptr);
            s->speed[k] = (int32_t*)ptr;
        }
        for (int k = 0; k < 2 && !rc; ++k) {
            rc = alloc(s, 1, &ptr);
            s->budget[k] = (int32_t*)ptr;
        }
        if (!rc) rc = alloc(s, 1, &ptr), s->clamped = (int32_t*)ptr;
        if (!rc) rc = alloc(s, 5, &ptr), s->params = (int32_t*)ptr;
    } else {
        const bool three = kind == VXR_KIND_MPM3;
        const int nf = three ? 18 : 10, ng = three ? 7 : 5;  // 3D: the 17 fields and the slot
        for (int f = 0; f < nf && !rc; ++f) {
            rc = alloc(s, s->capacity, &ptr);
            s->p.push_back((float*)ptr);
            if (!rc) rc = alloc(s, s->capacity, &ptr), s->p_alt.push_back((float*)ptr);
        }
        for (int f = 0; f < ng && !rc; ++f) {
            rc = alloc(s, round_up(s->nodes, 4), &ptr);
            s->g.push_back((float*)ptr);
        }
        s->param_words = three ? 7 : 28;
        if (!rc) rc = alloc(s, s->param_words, &ptr), s->params = (int32_t*)ptr;
        if (!rc) rc = alloc(s, s->capacity, &ptr), s->keys = (int32_t*)ptr;
        if (!rc) rc = alloc(s, s->capacity, &ptr), s->index = (int32_t*)ptr;
        // 3D: the radix scratch also holds the layer passes' new index and block counts afterwards
        const int64_t sort_words = std::max<int64_t>(vxr::radix_scratch_words(s->capacity, 8),
                                                     three ? vxr::mpm3_layer_scratch_words(s->capacity) : 0);
        if (!rc) rc = alloc(s, sort_words, &ptr), s->scratch = (int32_t*)ptr;
        if (!rc && !three && (flags & VXR_FLAG_TENSION)) {
            // FlipStep(..., tension)'s c0 c1 txx tyy txy (core/particle/FlipStep.java:111-115)
            s->tension = true;
            s->stale = false;  // the 2D passes are dense: no region to keep
            for (int k = 0; k < 2 && !rc; ++k) rc = alloc(s, round_up(s->nodes, 4), &ptr), s->colour[k] = (float*)ptr;
            for (int k = 0; k < 3 && !rc; ++k) rc = alloc(s, round_up(s->nodes, 4), &ptr), s->stress[k] = (float*)ptr;
        }
        if (!rc && three && (flags & VXR_FLAG_TENSION)) {
            s->tension = true;
            for (int k = 0; k < 2 && !rc; ++k) rc = alloc(s, round_up(s->nodes, 4), &ptr), s->colour[k] = (float*)ptr;
            for (int k = 0; k < 6 && !rc; ++k) rc = alloc(s, round_up(s->nodes, 4), &ptr), s->stress[k] = (float*)ptr;
            if (!rc) rc = alloc(s, (vxr::brick_count3(s->nx, s->ny, s->nz) + 3) / 4, &ptr), s->near = (uint8_t*)ptr;
        }
        if (!rc && three) {
            const int64_t groups = std::max<int64_t>(s->count / 8, 1);
            rc = alloc(s, groups, &ptr), s->level = (float*)ptr;
            if (!rc) rc = alloc(s, groups, &ptr), s->target = (float*)ptr;
            if (!rc) rc = alloc(s, s->capacity, &ptr), s->stay = (float*)ptr;
            if (!rc) rc = alloc(s, 8, &ptr), s->view = (float*)ptr;
        }
        if (!rc && three) {
            rc = alloc(s, s->capacity, &ptr), s->home = (int32_t*)ptr;
            if (!rc) rc = alloc(s, 1, &ptr), s->active = (int32_t*)ptr;
        }
        if (!rc && three) {
            vxr::Layers& L = s->layers;
            L.cap = vxr::mpm3_layer_capacity(s->nx, s->ny, s->nz);
            const int64_t cells = L.cap * vxr::mpm3_layer_cells();
            rc = alloc(s, L.cap + 1, &ptr), L.layer = (int32_t*)ptr;
            if (!rc) rc = alloc(s, L.cap + 1, &ptr), L.start = (int32_t*)ptr;
            if (!rc) rc = alloc(s, cells, &ptr), L.count = (int32_t*)ptr;
            if (!rc) rc = alloc(s, (cells + 1) / 2, &ptr), L.cell = (uint16_t*)ptr;
            if (!rc) rc = alloc(s, 8, &ptr), L.meta = (int32_t*)ptr;
        }
        if (!rc && three) {
            s->bricks = vxr::brick_count3(s->nx, s->ny, s->nz);
            rc = alloc(s, (s->bricks + 3) / 4, &ptr), s->touched = (uint8_t*)ptr;
            if (!rc) rc = alloc(s, (s->bricks + 3) / 4, &ptr), s->live = (uint8_t*)ptr;
        }
        if (!rc && !three) {
            s->starts_len = (int64_t)(s->nx - 1) * (s->ny - 1) + 1;
            rc = alloc(s, s->starts_len, &ptr), s->starts = (int32_t*)ptr;
        }
        if (!rc && !three) {
            vxr::Scene2 one{};
            one.nx = s->nx;
            one.ny = s->ny;
            one.count = s->count;
            s->scenes.push_back(one);
            layout_batch(s);
            rc = alloc_batch(s);
        }
    }
    if (!rc) rc = clear_all(s);
    if (rc) {
        destroy(s);
        return rc;
    }
    *out = s;
    return VXR_OK;
}

int vxr_create_batch(int nscenes, const int32_t* nx, const int32_t* ny, const int64_t* particles, uint32_t flags,
                     vxr_sim** out) {
    if (!out || !nx || !ny || !particles) return fail(VXR_EINVAL, "null argument");
    *out = nullptr;
    if (nscenes < 1) return fail(VXR_EINVAL, "a batch needs at least one scene, got %d", nscenes);
    const uint32_t physics = VXR_FLAG_TENSION | VXR_FLAG_HEAT | VXR_FLAG_CONVECTION | VXR_FLAG_RELAX | VXR_FLAG_PUMP |
                             VXR_FLAG_PLASTIC;
    if (flags & physics)
        return fail(VXR_EUNSUPPORTED, "tension, heat, convection, relax, pump and plastic are not built");
    if (flags) return fail(VXR_EINVAL, "unknown flags 0x%x", flags);
    vxr_sim* s = new vxr_sim;
    s->kind = VXR_KIND_MPM2_BATCH;
    for (const OptionSpec& o : kOptions) s->options[o.key] = o.def;
    s->options["p2g"] = 2;
    for (int k = 0; k < nscenes; ++k) {
        if (nx[k] < 4 || ny[k] < 4 || particles[k] < 1) {
            delete s;
            return fail(VXR_EINVAL, "scene %d: a grid needs at least four nodes each way and a step particles, got "
                        "%d x %d and %lld", k, nx[k], ny[k], (long long)particles[k]);
        }
        if ((int64_t)nx[k] * ny[k] > (1LL << 28)) {
            delete s;
            return fail(VXR_EINVAL, "scene %d: grid of %d x %d is too large", k, nx[k], ny[k]);
        }
        vxr::Scene2 sc{};
        sc.nx = nx[k];
        sc.ny = ny[k];
        sc.count = particles[k];
        s->scenes.push_back(sc);
    }
    layout_batch(s);
    s->capacity = s->pchunks * vxr::kBatchAlign;
    for (const vxr::Scene2& sc : s->scenes) s->count += sc.count;
    s->nodes = s->gchunks * 256;
    s->starts_len = s->schunks * 256;
    s->param_words = 28LL * nscenes;
    if (s->capacity > (int64_t)INT32_MAX - vxr::kBatchAlign || s->cell_bits + s->scene_bits > 31) {
        const int64_t cap = s->capacity;
        const int bits = s->cell_bits + s->scene_bits;
        delete s;
        return fail(VXR_EINVAL, "batch too large: %lld particle slots, %d key bits", (long long)cap, bits);
    }
    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipStreamCreate");
    }
    int rc = 0;
    void* ptr = nullptr;
    for (int f = 0; f < 10 && !rc; ++f) {
        rc = alloc(s, s->capacity, &ptr);
        s->p.push_back((float*)ptr);
        if (!rc) rc = alloc(s, s->capacity, &ptr), s->p_alt.push_back((float*)ptr);
    }
    for (int f = 0; f < 5 && !rc; ++f) {
        rc = alloc(s, s->nodes, &ptr);
        s->g.push_back((float*)ptr);
    }
    if (!rc) rc = alloc(s, s->param_words, &ptr), s->params = (int32_t*)ptr;
    if (!rc) rc = alloc(s, s->capacity, &ptr), s->keys = (int32_t*)ptr;
    if (!rc) rc = alloc(s, s->capacity, &ptr), s->index = (int32_t*)ptr;
    if (!rc) rc = alloc(s, vxr::radix_scratch_words(s->capacity, 8), &ptr), s->scratch = (int32_t*)ptr;
    if (!rc) rc = alloc(s, s->starts_len, &ptr), s->starts = (int32_t*)ptr;
    if (!rc) rc = alloc_batch(s);
    if (!rc) rc = clear_all(s);
    if (rc) {
        destroy(s);
        return rc;
    }
    *out = s;
    return VXR_OK;
}

int vxr_batch_layout(vxr_sim* s, int64_t* particle_offsets, int64_t* node_offsets, int64_t* start_offsets) {
    if (!s || !particle_offsets || !node_offsets || !start_offsets) return fail(VXR_EINVAL, "null argument");
    if (s->kind != VXR_KIND_MPM2_BATCH) return fail(VXR_ESTATE, "vxr_batch_layout is for a batch of scenes");
    const size_t n = s->scenes.size();
    for (size_t k = 0; k < n; ++k) {
        particle_offsets[k] = s->scenes[k].poff;
        node_offsets[k] = s->scenes[k].goff;
        start_offsets[k] = s->scenes[k].soff;
    }
    particle_offsets[n] = s->pchunks * vxr::kBatchAlign;
    node_offsets[n] = s->gchunks * 256;
    start_offsets[n] = s->schunks * 256;
    return VXR_OK;
}

int vxr_scene_count(vxr_sim* s, int* n) {
    if (!s || !n) return fail(VXR_EINVAL, "null argument");
    *n = s->kind == VXR_KIND_MPM2_BATCH ? (int)s->scenes.size() : 1;
    return VXR_OK;
}

int vxr_destroy(vxr_sim* s) {
    destroy(s);
    return VXR_OK;
}

int vxr_write(vxr_sim* s, const char* name, const void* words, int64_t n) {
    if (!s || !name || (!words && n)) return fail(VXR_EINVAL, "null argument");
    Buf b;
    if (!lookup(s, name, &b)) return fail(VXR_EINVAL, "no buffer named '%s'", name);
    const bool prefix_ok = std::string(name) == "params";
    if (n != b.len && !(prefix_ok && n >= 0 && n <= b.len))
        return fail(VXR_EINVAL, "buffer '%s' holds %lld words, got %lld", name, (long long)b.len, (long long)n);
    const bool three = s->kind == VXR_KIND_MPM3;
    if (three && b.ptr == (void*)s->active) return fail(VXR_EINVAL, "buffer 'active' is read-only");
    if (three && b.ptr == (void*)s->p[17]) {
        // the slots must be a permutation of 0 .. count - 1: the level of detail indexes through them
        const int32_t* v = static_cast<const int32_t*>(words);
        std::vector<uint8_t> seen((size_t)n, 0);
        for (int64_t i = 0; i < n; ++i) {
            if (v[i] < 0 || v[i] >= n || seen[(size_t)v[i]])
                return fail(VXR_EINVAL, "'slot' must be a permutation of 0..%lld; word %lld is %d", (long long)n - 1,
                            (long long)i, v[i]);
            seen[(size_t)v[i]] = 1;
        }
    }
    HIP_TRY(hipMemcpyAsync(b.ptr, words, n * 4, hipMemcpyHostToDevice, s->stream));
    if ((three || s->btouched) && name[0] == 'g' && name[1] != 0)  // a grid field: sweep it all next time
        if (int rc = mark_all_bricks(s)) return rc;
    if (three && (b.ptr == (void*)s->p[17] || b.ptr == (void*)s->p[6]))
        // new slots, or new masses: home from the slots, and P2G and G2P visit every particle
        if (int rc = vxr::mpm3_rebuild_home(s->launch(), mpm3_args(s))) return rc;
    if (three && (b.ptr == (void*)s->p[0] || b.ptr == (void*)s->p[1] || b.ptr == (void*)s->p[2] ||
                  b.ptr == (void*)s->p[6] || b.ptr == (void*)s->p[17] || b.ptr == (void*)s->keys))
        s->layers_valid = false;  // P2G goes by the sorted-segmented schedule until the next sort
    bool tension_buffer = false;
    for (float* f : s->stress) tension_buffer = tension_buffer || (f && b.ptr == (void*)f);
    if (three && s->tension && (b.ptr == s->colour[0] || b.ptr == s->colour[1] || tension_buffer)) {
        // tension buffers: unknown content, so the next passes run dense
        if (b.ptr == s->colour[0]) s->spread[0] = -1;
        if (b.ptr == s->colour[1]) s->spread[1] = -1;
        s->stale = true;
        if (int rc = mark_all_bricks(s)) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return VXR_OK;
}

int vxr_read(vxr_sim* s, const char* name, void* words, int64_t n) {
    if (!s || !name || (!words && n)) return fail(VXR_EINVAL, "null argument");
    Buf b;
    if (!lookup(s, name, &b)) return fail(VXR_EINVAL, "no buffer named '%s'", name);
    if (n < 0 || n > b.len)
        return fail(VXR_EINVAL, "buffer '%s' holds %lld words, asked %lld", name, (long long)b.len, (long long)n);
    HIP_TRY(hipMemcpyAsync(words, b.ptr, n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return VXR_OK;
}

int vxr_length(vxr_sim* s, const char* name, int64_t* n) {
    if (!s || !name || !n) return fail(VXR_EINVAL, "null argument");
    Buf b;
    if (!lookup(s, name, &b)) return fail(VXR_EINVAL, "no buffer named '%s'", name);
    *n = b.len;
    return VXR_OK;
}

int vxr_run(vxr_sim* s, int group, int repeat) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    if (s->kind == VXR_KIND_SWE) return fail(VXR_ESTATE, "a shallow-water patch steps by vxr_swe_dispatch");
    if (repeat < 0) return fail(VXR_EINVAL, "negative repeat");
    for (int r = 0; r < repeat; ++r) {
        if (group == VXR_GROUP_STEP) {
            if (int rc = run_step(s)) return rc;
        } else if (group == VXR_GROUP_SORT) {
            if (int rc = sort_pass(s)) return rc;
        } else if (group == VXR_GROUP_LOD) {
            if (s->kind != VXR_KIND_MPM3) return fail(VXR_ESTATE, "the level of detail is the 3D step's");
            if (int rc = run_pass(s, "lod_target")) return rc;
            if (int rc = run_pass(s, "lod")) return rc;
        } else {
            return fail(VXR_EINVAL, "unknown group %d", group);
        }
    }
    return VXR_OK;
}

int vxr_run_pass(vxr_sim* s, const char* pass, int repeat) {
    if (!s || !pass) return fail(VXR_EINVAL, "null argument");
    if (s->kind == VXR_KIND_SWE) return fail(VXR_ESTATE, "a shallow-water patch steps by vxr_swe_dispatch");
    for (int r = 0; r < repeat; ++r)
        if (int rc = run_pass(s, pass)) return rc;
    return VXR_OK;
}

int vxr_advance(vxr_sim* s, int count, int sort_every, int64_t first_step) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    if (s->kind == VXR_KIND_SWE) return fail(VXR_ESTATE, "a shallow-water patch steps by vxr_swe_dispatch");
    if (count < 0 || sort_every < 0 || first_step < 0)
        return fail(VXR_EINVAL, "count, sort_every and first_step must not be negative");
    int64_t step = first_step;
    int left = count;
    while (left > 0) {
        const bool sort = sort_every > 0 && step % sort_every == 0;
        int64_t n = left;
        if (sort_every > 0) n = std::min<int64_t>(n, sort_every - step % sort_every);
        n = std::min<int64_t>(n, kMaxSegment);
        if (int rc = replay_segment(s, sort, (int)n)) return rc;
        step += n;
        left -= (int)n;
    }
    return VXR_OK;
}

int vxr_autotune(vxr_sim* s, int reps) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    if (reps < 1) return fail(VXR_EINVAL, "reps must be positive, got %d", reps);
    if (s->timer) return fail(VXR_ESTATE, "turn timing off before tuning");
    drop_graphs(s);
    return autotune(s, reps);
}

int vxr_fixed_scatter(void* stream, int n, int64_t count, const float* x, const float* y, const float* u,
                      const float* v, const float* m, int mass_bits, int velocity_bits, int schedule, int32_t* grid_m,
                      int32_t* grid_mu, int32_t* grid_mv) {
    if (n < 2) return fail(VXR_EINVAL, "a grid needs at least two nodes each way, got %d", n);
    if (count < 0) return fail(VXR_EINVAL, "negative particle count");
    if (count > 0 && (!x || !y || !u || !v || !m)) return fail(VXR_EINVAL, "null particle field");
    if (!grid_m || !grid_mu || !grid_mv) return fail(VXR_EINVAL, "null grid");
    if (mass_bits < 0 || velocity_bits < 0 || mass_bits > 30 || velocity_bits > 30)
        return fail(VXR_EINVAL, "bits out of range: %d, %d", mass_bits, velocity_bits);
    if (schedule != 0 && schedule != 1) return fail(VXR_EINVAL, "schedule is 0 (direct) or 1 (segmented)");
    const vxr::Launch l{(hipStream_t)stream, nullptr};
    return vxr::fixed_scatter(l, n, count, x, y, u, v, m, mass_bits, velocity_bits, schedule ? 4 : 0, grid_m, grid_mu,
                              grid_mv);
}

int vxr_swe_diagnostics(vxr_sim* s, double g, double dry, double dt, double dx, double dy, double out[9]) {
    if (!s || !out) return fail(VXR_EINVAL, "null argument");
    if (s->kind != VXR_KIND_SWE) return fail(VXR_ESTATE, "vxr_swe_diagnostics is for a shallow-water patch");
    void* ptr = nullptr;
    if (!s->diag) {
        if (int rc = alloc(s, vxr::diag_scratch_words(), &ptr)) return rc;
        s->diag = (int32_t*)ptr;
    }
    const int slot = (int)(s->dispatched % 2);
    return vxr::swe_diagnostics(s->launch(), s->state[slot][0], s->state[slot][1], s->state[slot][2],
                                (int64_t)s->nx * s->ny, g, dry, dt, dx, dy, s->diag, out);
}

int vxr_particle_diagnostics(vxr_sim* s, double speed, double sound, double dt, double out[7]) {
    if (!s || !out) return fail(VXR_EINVAL, "null argument");
    if (s->kind == VXR_KIND_SWE) return fail(VXR_ESTATE, "vxr_particle_diagnostics is for particles");
    void* ptr = nullptr;
    if (!s->diag) {
        if (int rc = alloc(s, vxr::diag_scratch_words(), &ptr)) return rc;
        s->diag = (int32_t*)ptr;
    }
    const float* j = s->p[s->kind == VXR_KIND_MPM3 ? 7 : 5];
    const int64_t count = s->kind == VXR_KIND_MPM2_BATCH ? s->capacity : s->count;
    // a batch's pad particles hold J = 0 and count as not valid; they are left out below
    int64_t pads = 0;
    if (s->kind == VXR_KIND_MPM2_BATCH) pads = s->capacity - s->count;
    int rc = vxr::particle_diagnostics(s->launch(), j, count, s->capacity, s->keys, s->index, s->scratch,
                                       (int)opt(s, "radix_bits"), s->diag, speed, sound, dt, out);
    if (!rc && pads) out[6] -= (double)pads;
    return rc;
}

int vxr_particle_diagnostics_noalloc(void* stream, const float* j, int64_t count, double speed, double sound,
                                     double dt, double out[7]) {
    if (!j || !out) return fail(VXR_EINVAL, "null argument");
    const int64_t limit = vxr::particle_diagnostics_static_limit();
    if (count < 1 || count > limit)
        return fail(VXR_EINVAL, "vxr_particle_diagnostics_noalloc takes 1 to %lld particles (its fixed scratch), got %lld",
                    (long long)limit, (long long)count);
    const char* why = "";
    const int rc = vxr::particle_diagnostics_static(vxr::Launch{(hipStream_t)stream, nullptr}, j, count, speed, sound,
                                                    dt, out, &why);
    if (rc == 3 && *why) return fail(VXR_EDEVICE, "%s", why);
    return rc;
}

int vxr_swe_dispatch(vxr_sim* s, int count) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    if (s->kind != VXR_KIND_SWE) return fail(VXR_ESTATE, "vxr_swe_dispatch is for a shallow-water patch");
    if (count < 0) return fail(VXR_EINVAL, "negative count");
    vxr::SweArgs a{};
    a.nx = s->nx;
    a.ny = s->ny;
    a.edges[0] = (s->flags & VXR_FLAG_OPEN_WEST) ? 1 : 0;
    a.edges[1] = (s->flags & VXR_FLAG_OPEN_EAST) ? 1 : 0;
    a.edges[2] = (s->flags & VXR_FLAG_OPEN_SOUTH) ? 1 : 0;
    a.edges[3] = (s->flags & VXR_FLAG_OPEN_NORTH) ? 1 : 0;
    a.params = s->params;
    a.clamped = s->clamped;
    const int rows = (int)opt(s, "rows");
    const bool fast = opt(s, "hll_fast") != 0;
    for (int i = 0; i < count; ++i) {
        const int64_t k = s->dispatched;
        for (int f = 0; f < 3; ++f) {
            a.in[f] = s->state[k % 2][f];
            a.out[f] = s->state[(k + 1) % 2][f];
        }
        a.speed_in = s->speed[k % 3];
        a.speed_out = s->speed[(k + 1) % 3];
        a.speed_clear = s->speed[(k + 2) % 3];
        a.budget_in = s->budget[k % 2];
        a.budget_out = s->budget[(k + 1) % 2];
        if (int rc = vxr::swe_step(s->launch(), a, rows, fast)) return rc;
        s->dispatched++;
    }
    return VXR_OK;
}

int vxr_clear(vxr_sim* s) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    return clear_all(s);
}

int vxr_finish(vxr_sim* s) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return VXR_OK;
}

int vxr_device_ptr(vxr_sim* s, const char* name, void** ptr) {
    if (!s || !name || !ptr) return fail(VXR_EINVAL, "null argument");
    Buf b;
    if (!lookup(s, name, &b)) return fail(VXR_EINVAL, "no buffer named '%s'", name);
    *ptr = b.ptr;
    return VXR_OK;
}

int vxr_stream(vxr_sim* s, void** stream) {
    if (!s || !stream) return fail(VXR_EINVAL, "null argument");
    *stream = (void*)s->stream;
    return VXR_OK;
}

int vxr_set_option(vxr_sim* s, const char* key, int64_t value) {
    if (!s || !key) return fail(VXR_EINVAL, "null argument");
    for (const OptionSpec& o : kOptions) {
        if (std::string(key) == o.key) {
            if (value < o.lo || value > o.hi)
                return fail(VXR_EINVAL, "option '%s' takes %lld..%lld, got %lld", key, (long long)o.lo, (long long)o.hi,
                            (long long)value);
            if (std::string(key) == "run" && value == 3)
                return fail(VXR_EINVAL, "option 'run' takes 1, 2 or 4, got 3");
            if (std::string(key) == "p2g" && value == 2 && s->kind == VXR_KIND_SWE)
                return fail(VXR_EINVAL, "p2g is for particle simulations");
            if (std::string(key) == "p2g" && value == 2 && s->kind == VXR_KIND_MPM3 && !s->q3[0]) {
                // the fixed-point sums, made once and zeroed (the option change below marks every
                // brick, so the first clear zeroes them again over the whole grid)
                const int64_t words = 2 * round_up(s->nodes, 4);
                for (int f = 0; f < 4; ++f) {
                    void* ptr = nullptr;
                    if (int rc = alloc(s, words, &ptr)) return rc;
                    s->q3[f] = (unsigned long long*)ptr;
                    HIP_TRY(hipMemsetAsync(ptr, 0, words * 4, s->stream));
                }
            }
            if (std::string(key) == "p2g" && value != 2 && s->kind == VXR_KIND_MPM2_BATCH)
                return fail(VXR_EINVAL, "a batch of scenes always scatters in fixed point (p2g 2)");
            drop_graphs(s);
            const bool p2g_changed = std::string(key) == "p2g" && s->options[key] != value;
            s->options[key] = value;
            // the float P2G paths clear and update the grid in full: the chunk flags start over
            if (p2g_changed)
                if (int rc = mark_all_bricks(s)) return rc;
            return VXR_OK;
        }
    }
    return fail(VXR_EINVAL, "no option named '%s'", key);
}

int vxr_get_option(vxr_sim* s, const char* key, int64_t* value) {
    if (!s || !key || !value) return fail(VXR_EINVAL, "null argument");
    auto it = s->options.find(key);
    if (it == s->options.end()) return fail(VXR_EINVAL, "no option named '%s'", key);
    *value = it->second;
    return VXR_OK;
}

int vxr_timing(vxr_sim* s, int enable) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    HIP_TRY(hipStreamSynchronize(s->stream));
    delete s->timer;
    s->timer = enable ? new vxr::Timer : nullptr;
    return VXR_OK;
}

int vxr_timing_only(vxr_sim* s, const char* kernel) {
    if (!s || !kernel || !kernel[0]) return fail(VXR_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    delete s->timer;
    s->timer = new vxr::Timer;
    s->timer->only = kernel;
    return VXR_OK;
}

int vxr_kernel_time(vxr_sim* s, const char* kernel, double* total_ms, int64_t* launches) {
    if (!s || !kernel || !total_ms || !launches) return fail(VXR_EINVAL, "null argument");
    if (!s->timer) return fail(VXR_ESTATE, "timing is off");
    if (s->timer->drain()) return fail(VXR_EDEVICE, "event synchronize failed");
    auto it = s->timer->totals.find(kernel);
    *total_ms = it == s->timer->totals.end() ? 0.0 : it->second.first;
    *launches = it == s->timer->totals.end() ? 0 : it->second.second;
    return VXR_OK;
}

}  // extern "C"
