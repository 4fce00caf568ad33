// vg_runtime.hip -- errors, device memory plumbing, kernel timing and the
// host-side grid geometry shared by every operator.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "vg_internal.h"

namespace vg {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

vg_status device_error(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
    return e == hipErrorOutOfMemory ? VG_ENOMEM : VG_EDEVICE;
}

// ------------------------------------------------------------------ profiling
namespace {
struct PendingLaunch {
    std::string name;
    hipEvent_t start, stop;
};
struct Tally {
    double ms = 0.0;
    uint64_t count = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<std::string> g_prof_only;  // name prefixes to time; empty: every kernel
std::vector<PendingLaunch *> g_pending;
std::vector<hipEvent_t> g_free_events;  // recorded and read events, reused
std::map<std::string, Tally> g_tally;

hipEvent_t take_event_locked() {
    if (!g_free_events.empty()) {
        hipEvent_t e = g_free_events.back();
        g_free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain_locked() {
    for (PendingLaunch *p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p->stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p->start, p->stop) == hipSuccess) {
            Tally &t = g_tally[p->name];
            t.ms += ms;
            t.count += 1;
        }
        g_free_events.push_back(p->start);
        g_free_events.push_back(p->stop);
        delete p;
    }
    g_pending.clear();
}
}  // namespace

bool profiling_enabled() { return g_prof_on; }

bool profile_wanted(const char *name) {
    if (!g_prof_on) return false;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (g_prof_only.empty()) return true;
    for (const std::string &p : g_prof_only)
        if (strncmp(name, p.c_str(), p.size()) == 0) return true;
    return false;
}

// A launch whose events cannot be created or recorded is left untimed (the
// profiler is a diagnostic: it never fails the launch it brackets).
void profile_begin(const char *name, hipStream_t s, void **token) {
    *token = nullptr;
    auto *p = new PendingLaunch{name, nullptr, nullptr};
    std::lock_guard<std::mutex> lk(g_prof_mu);
    p->start = take_event_locked();
    p->stop = take_event_locked();
    if (p->start && p->stop && hipEventRecord(p->start, s) == hipSuccess) {
        *token = p;
        return;
    }
    if (p->start) g_free_events.push_back(p->start);
    if (p->stop) g_free_events.push_back(p->stop);
    delete p;
}

void profile_end(void *token, hipStream_t s) {
    auto *p = static_cast<PendingLaunch *>(token);
    if (!p) return;
    const bool recorded = hipEventRecord(p->stop, s) == hipSuccess;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (recorded) {
        g_pending.push_back(p);
        return;
    }
    g_free_events.push_back(p->start);
    g_free_events.push_back(p->stop);
    delete p;
}

// ------------------------------------------------------------------ scratch
vg_status salloc(void **p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    if (bytes == 0) return VG_OK;
    VG_HIP(hipMallocAsync(p, bytes, s));
    return VG_OK;
}

vg_status sfree(void *p, hipStream_t s) {
    if (p) VG_HIP(hipFreeAsync(p, s));
    return VG_OK;
}

}  // namespace vg

using namespace vg;

// ------------------------------------------------------------------ C-ABI
VG_API const char *vg_last_error(void) { return g_last_error.c_str(); }

VG_API const char *vg_version(void) { return "vortegrid-amd 0.1 (gfx950)"; }

VG_API vg_status vg_device_count(int *count) {
    VG_REQUIRE(count, "count is null");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) { *count = 0; return device_error(e, "hipGetDeviceCount", __FILE__, __LINE__); }
    return VG_OK;
}

VG_API vg_status vg_malloc(void **ptr, size_t bytes) {
    VG_REQUIRE(ptr, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) return VG_OK;
    VG_HIP(hipMalloc(ptr, bytes));
    return VG_OK;
}

VG_API vg_status vg_free(void *ptr) {
    if (ptr) VG_HIP(hipFree(ptr));
    return VG_OK;
}

VG_API vg_status vg_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return VG_OK;
    VG_REQUIRE(dst && src, "null pointer");
    VG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return VG_OK;
}

VG_API vg_status vg_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return VG_OK;
    VG_REQUIRE(dst && src, "null pointer");
    VG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return VG_OK;
}

VG_API vg_status vg_memset(void *dst, int value, size_t bytes) {
    if (bytes == 0) return VG_OK;
    VG_REQUIRE(dst, "null pointer");
    VG_HIP(hipMemset(dst, value, bytes));
    return VG_OK;
}

VG_API vg_status vg_stream_create(vg_stream *s) {
    VG_REQUIRE(s, "stream out is null");
    hipStream_t h;
    VG_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = h;
    return VG_OK;
}

VG_API vg_status vg_stream_destroy(vg_stream s) {
    if (s) VG_HIP(hipStreamDestroy((hipStream_t)s));
    return VG_OK;
}

VG_API vg_status vg_stream_synchronize(vg_stream s) {
    VG_HIP(hipStreamSynchronize((hipStream_t)s));
    return VG_OK;
}

VG_API vg_status vg_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
    return VG_OK;
}

VG_API vg_status vg_profile_select(const char *prefixes) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_only.clear();
    if (!prefixes) return VG_OK;
    std::string cur;
    for (const char *c = prefixes;; ++c) {
        if (*c == ',' || *c == '\0') {
            if (!cur.empty()) g_prof_only.push_back(cur);
            cur.clear();
            if (*c == '\0') break;
        } else if (*c != ' ') {
            cur += *c;
        }
    }
    return VG_OK;
}

VG_API vg_status vg_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    drain_locked();
    g_tally.clear();
    return VG_OK;
}

VG_API vg_status vg_profile_get(const char *kernel, double *total_ms, uint64_t *count) {
    VG_REQUIRE(kernel && total_ms && count, "null argument");
    std::lock_guard<std::mutex> lk(g_prof_mu);
    drain_locked();
    auto it = g_tally.find(kernel);
    *total_ms = it == g_tally.end() ? 0.0 : it->second.ms;
    *count = it == g_tally.end() ? 0 : it->second.count;
    return VG_OK;
}

VG_API vg_status vg_profile_dump(char *buf, size_t cap) {
    VG_REQUIRE(buf && cap > 0, "null buffer");
    std::lock_guard<std::mutex> lk(g_prof_mu);
    drain_locked();
    std::string out;
    char line[256];
    for (auto &kv : g_tally) {
        snprintf(line, sizeof line, "%s %llu %.6f\n", kv.first.c_str(),
                 (unsigned long long)kv.second.count, kv.second.ms);
        out += line;
    }
    snprintf(buf, cap, "%s", out.c_str());
    return VG_OK;
}

VG_API vg_status vg_define_shape(vg_grid_desc *g, size_t n, const float vmin[3],
                                 const float vmax[3], int pow2) {
    VG_REQUIRE(g && vmin && vmax, "null argument");
    define_shape(g, n, vmin, vmax, pow2);
    return VG_OK;
}

VG_API vg_status vg_decimate(vg_grid_desc *dst, const vg_grid_desc *src, uint32_t dec) {
    VG_REQUIRE(dst && src && dec > 0, "bad argument");
    VG_REQUIRE(src->np[0] >= 2 && src->np[1] >= 2 && src->np[2] >= 2, "source grid is empty");
    decimate(dst, src, dec);
    return VG_OK;
}

VG_API vg_status vg_nested_layers(const vg_grid_desc *src, vg_grid_desc *layers, uint32_t *dec,
                                  uint32_t max_layers, uint32_t *depth) {
    VG_REQUIRE(src && depth, "null argument");
    VG_REQUIRE(src->np[0] >= 2 && src->np[1] >= 2 && src->np[2] >= 2, "source grid is empty");
    Nested ng;
    nested_init(&ng, *src);
    *depth = ng.depth;
    for (uint32_t l = 0; l < ng.depth && l < max_layers; ++l) {
        if (layers) layers[l] = ng.layer[l];
        if (dec)
            for (int a = 0; a < 3; ++a) dec[3 * l + a] = ng.dec[l][a];
    }
    return VG_OK;
}

VG_API vg_status vg_finalize_bounding_box(float mn[3], float mx[3], float vorton_radius,
                                          size_t num_vortons, vg_grid_desc *grid_template) {
#pragma clang fp contract(off)
    VG_REQUIRE(mn && mx && grid_template, "null argument");
    if (num_vortons == 0) {
        const float z[3] = {0.f, 0.f, 0.f};
        define_shape(grid_template, 0, z, z, 1);
        return VG_OK;
    }
    const float margin = (1.0f + 2.0f) * vorton_radius;
    const float nudge = margin * 1.0f;
    for (int a = 0; a < 3; ++a) {
        mn[a] -= nudge;
        mx[a] += nudge;
    }
    define_shape(grid_template, num_vortons, mn, mx, 1);
    return VG_OK;
}

VG_API vg_status vg_auto_mollify(const vg_grid_desc *g, float vorton_radius, float *spread,
                                 float *circ) {
#pragma clang fp contract(off)
    VG_REQUIRE(g && spread && circ && vorton_radius > 0.f, "bad argument");
    const float one_third = 1.0f / 3.0f;
    const float cell_volume = g->cell[0] * g->cell[1] * g->cell[2];
    const float cgl = 1.0f * pow_rounded(cell_volume, one_third);
    float s = cgl / vorton_radius;
    if (s < 1.0f) s = 1.0f;
    *spread = s;
    *circ = 1.0f / (s * s * s);
    return VG_OK;
}
