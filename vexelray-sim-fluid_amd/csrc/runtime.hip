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
#include <map>
#include <string>
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

}  // namespace

namespace vxr {

struct Timer {
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

void timed_begin(const Launch& l, const char*) {
    if (!l.timer) return;
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

    // SWE
    float* state[2][3] = {};
    int32_t* speed[3] = {};
    int32_t* budget[2] = {};
    int32_t* clamped = nullptr;
    int64_t dispatched = 0;

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
    return a;
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

int sort_pass(vxr_sim* s) {
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
    const int nf = three ? 17 : 10;
    if ((rc = vxr::permute_fields(l, s->p.data(), s->p_alt.data(), nf, s->index, s->count))) return rc;
    std::swap(s->p, s->p_alt);
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
        if (pass == "clear") return vxr::mpm3_clear(l, a);
        if (pass == "scatter")
            return opt(s, "p2g") == 0 ? vxr::mpm3_scatter_atomic(l, a) : vxr::mpm3_scatter_tiled(l, a, (int)opt(s, "run"));
        if (pass == "grid") return vxr::mpm3_grid(l, a);
        if (pass == "advect") {
            const int64_t g2p = opt(s, "g2p");
            return g2p == 0 ? vxr::mpm3_advect(l, a) : vxr::mpm3_advect_tiled(l, a, (int)opt(s, "run"), g2p == 2);
        }
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
        HIP_TRY(hipMemsetAsync(s->keys, 0, s->capacity * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->index, 0, s->capacity * 4, s->stream));
        if (s->starts) {
            const int64_t cells = (int64_t)(s->nx - 1) * (s->ny - 1);
            HIP_TRY(hipMemsetAsync(s->starts, 0, (cells + 1) * 4, s->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

void destroy(vxr_sim* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
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
    {"p2g", 0, 1, 1}, {"g2p", 0, 2, 2}, {"run", 1, 4, 4}, {"radix_bits", 4, 8, 8}, {"rows", 4, 256, 32},
};

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
    if (flags & physics)
        return fail(VXR_EUNSUPPORTED, "tension, heat, convection, relax, pump and plastic are not built");
    const uint32_t edges = VXR_FLAG_OPEN_WEST | VXR_FLAG_OPEN_EAST | VXR_FLAG_OPEN_SOUTH | VXR_FLAG_OPEN_NORTH;
    if (flags & ~(physics | edges)) return fail(VXR_EINVAL, "unknown flags 0x%x", flags & ~(physics | edges));

    vxr_sim* s = new vxr_sim;
    s->kind = kind;
    s->flags = flags;
    for (const OptionSpec& o : kOptions) s->options[o.key] = o.def;
    int rc = 0;
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
        const int nf = three ? 17 : 10, ng = three ? 7 : 5;
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
        if (!rc) rc = alloc(s, vxr::radix_scratch_words(s->capacity, 8), &ptr), s->scratch = (int32_t*)ptr;
        if (!rc && !three) rc = alloc(s, (int64_t)(s->nx - 1) * (s->ny - 1) + 1, &ptr), s->starts = (int32_t*)ptr;
    }
    if (!rc) rc = clear_all(s);
    if (rc) {
        destroy(s);
        return rc;
    }
    *out = s;
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
    HIP_TRY(hipMemcpyAsync(b.ptr, words, n * 4, hipMemcpyHostToDevice, s->stream));
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
    static const char* step[4] = {"clear", "scatter", "grid", "advect"};
    for (int r = 0; r < repeat; ++r) {
        if (group == VXR_GROUP_STEP) {
            for (const char* p : step)
                if (int rc = run_pass(s, p)) return rc;
        } else if (group == VXR_GROUP_SORT) {
            if (int rc = sort_pass(s)) return rc;
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

int vxr_swe_dispatch(vxr_sim* s, int count) {
    if (!s) return fail(VXR_EINVAL, "null simulation");
    if (s->kind != VXR_KIND_SWE) return fail(VXR_ESTATE, "vxr_swe_dispatch is for a shallow-water patch");
    if (count < 0) return fail(VXR_EINVAL, "negative count");
    vxr::SweArgs a;
    a.nx = s->nx;
    a.ny = s->ny;
    a.edges[0] = (s->flags & VXR_FLAG_OPEN_WEST) ? 1 : 0;
    a.edges[1] = (s->flags & VXR_FLAG_OPEN_EAST) ? 1 : 0;
    a.edges[2] = (s->flags & VXR_FLAG_OPEN_SOUTH) ? 1 : 0;
    a.edges[3] = (s->flags & VXR_FLAG_OPEN_NORTH) ? 1 : 0;
    a.params = s->params;
    a.clamped = s->clamped;
    const int rows = (int)opt(s, "rows");
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
        if (int rc = vxr::swe_step(s->launch(), a, rows)) return rc;
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
            s->options[key] = value;
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
