// capi.hip -- runtime plumbing of the C ABI: errors, device selection,
// memory, streams and events.  Kernels live in square.hip, fft.hip and
// sieve.hip.
#include "gf2t_internal.hpp"
#include "host_poly.hpp"
#include "../../include/gf2t.h"
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <mutex>
#include <string>

namespace gf2t {
static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error '%s' in %s (%s:%d)", hipGetErrorString(e), what, file, line);
    return e == hipErrorOutOfMemory ? GF2T_ENOMEM : GF2T_EHIP;
}
}  // namespace gf2t

extern "C" const char *gf2t_last_error(void) { return gf2t::g_err; }

extern "C" int gf2t_version(void) { return 1; }

extern "C" int gf2t_device_count(int *count) {
    if (!count) return GF2T_EINVAL;
    GF2T_HIP(hipGetDeviceCount(count));
    return GF2T_OK;
}

int gf2t_noalloc_setup();   // fft.hip: per-device scratch of gf2t_mulmod_noalloc

namespace {
// The autotuner's choices: "key value" lines (gf2t_tune keys, '#' starts a
// comment) from $GF2T_TUNE_FILE, or else gf2t_tune.txt beside this
// library; GF2T_TUNE_FILE set to the empty string skips them.  A knob
// whose environment variable is set keeps the environment's value.
// Applied once per process, by the first gf2t_init.
void load_tune_file() {
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path;
        if (const char *e = getenv("GF2T_TUNE_FILE")) {
            if (!*e) return;
            path = e;
        } else {
            Dl_info info;
            if (!dladdr((const void *)&gf2t_version, &info) || !info.dli_fname) return;
            std::string lib = info.dli_fname;
            size_t sl = lib.rfind('/');
            path = (sl == std::string::npos ? std::string(".") : lib.substr(0, sl)) +
                   "/gf2t_tune.txt";
        }
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return;
        static const char *env_of[][2] = {{"sq_fuse", "GF2T_SQ_FUSE"}, {"sq_ppt", "GF2T_SQ_PPT"},
                                          {"sq_order", "GF2T_SQ_ORDER"}, {"sq_nt", "GF2T_SQ_NT"},
                                          {"gcd_dev_bits", "GF2T_GCD_DEV_BITS"}};
        char line[256], key[96];
        long long v;
        while (fgets(line, sizeof line, f)) {
            if (line[0] == '#' || sscanf(line, "%95s %lld", key, &v) != 2) continue;
            bool env = false;
            for (auto &kv : env_of)
                if (!strcmp(key, kv[0]) && getenv(kv[1])) env = true;
            if (!env && gf2t_tune(key, (int64_t)v) != GF2T_OK)
                fprintf(stderr, "gf2t: %s: ignoring '%s %lld' (%s)\n", path.c_str(), key, v,
                        gf2t_last_error());
        }
        fclose(f);
    });
}
}  // namespace

extern "C" int gf2t_init(int device) {
    int n = 0;
    GF2T_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        gf2t::set_error("gf2t_init: device %d of %d", device, n);
        return GF2T_EINVAL;
    }
    GF2T_HIP(hipSetDevice(device));
    // GF2T_SYNC=block: host threads waiting for the device sleep instead of
    // spinning (for many gcd threads sharing the host's cores)
    if (const char *e = getenv("GF2T_SYNC"))
        if (!strcmp(e, "block")) (void)hipSetDeviceFlags(hipDeviceScheduleBlockingSync);
    GF2T_HIP(hipFree(nullptr));   // create the context now
    int rc = gf2t_noalloc_setup();
    if (rc) return rc;
    load_tune_file();
    // with a device present, large gcd products go to the device
    if (!gf2t_host::device_product_bits()) {
        const char *e = getenv("GF2T_GCD_DEV_BITS");
        gf2t_host::set_device_product_bits(e ? strtoull(e, nullptr, 10) : (1ULL << 16));
    }
    return GF2T_OK;
}

extern "C" int gf2t_device_name(char *buf, size_t cap) {
    if (!buf || !cap) return GF2T_EINVAL;
    int dev = 0;
    GF2T_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    GF2T_HIP(hipGetDeviceProperties(&p, dev));
    snprintf(buf, cap, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return GF2T_OK;
}

extern "C" int gf2t_malloc(void **dptr, size_t bytes) {
    if (!dptr) return GF2T_EINVAL;
    *dptr = nullptr;
    if (!bytes) bytes = 8;
    GF2T_HIP(hipMalloc(dptr, bytes));
    return GF2T_OK;
}

extern "C" int gf2t_free(void *dptr) {
    if (dptr) GF2T_HIP(hipFree(dptr));
    return GF2T_OK;
}

extern "C" int gf2t_mem_info(size_t *free_bytes, size_t *total_bytes) {
    size_t f = 0, t = 0;
    GF2T_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GF2T_OK;
}

extern "C" int gf2t_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return GF2T_OK;
    if (stream) {
        GF2T_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
        GF2T_HIP(hipStreamSynchronize((hipStream_t)stream));
    } else {
        GF2T_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }
    return GF2T_OK;
}

extern "C" int gf2t_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return GF2T_OK;
    if (stream) {
        GF2T_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
        GF2T_HIP(hipStreamSynchronize((hipStream_t)stream));
    } else {
        GF2T_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    return GF2T_OK;
}

extern "C" int gf2t_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return GF2T_OK;
    GF2T_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return GF2T_OK;
}

namespace {
__global__ void k_fill_bytes(unsigned char *p, size_t n, unsigned char v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u64(u64 *p, size_t n, u64 v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
}  // namespace

// Fill by kernel rather than hipMemsetAsync so the call is safe inside a
// captured HIP graph.
extern "C" int gf2t_memset(void *dptr, int value, size_t bytes, void *stream) {
    if (!bytes) return GF2T_OK;
    if (!dptr) return GF2T_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    unsigned char v = (unsigned char)value;
    if (((uintptr_t)dptr & 7) == 0 && (bytes & 7) == 0) {
        u64 w = 0x0101010101010101ULL * v;
        size_t n = bytes / 8;
        unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, st, (u64 *)dptr, n, w);
    } else {
        unsigned blocks = (unsigned)std::min<size_t>((bytes + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_bytes, dim3(blocks), dim3(256), 0, st, (unsigned char *)dptr,
                           bytes, v);
    }
    GF2T_CHECK_LAUNCH("gf2t_memset");
    return GF2T_OK;
}

extern "C" int gf2t_stream_create(void **stream) {
    if (!stream) return GF2T_EINVAL;
    hipStream_t s;
    GF2T_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return GF2T_OK;
}

extern "C" int gf2t_stream_destroy(void *stream) {
    if (stream) GF2T_HIP(hipStreamDestroy((hipStream_t)stream));
    return GF2T_OK;
}

// ---- HIP graphs of captured call sequences
struct gf2t_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int gf2t_graph_begin(void *stream) {
    if (!stream) {
        gf2t::set_error("gf2t_graph_begin: capture needs a stream from gf2t_stream_create");
        return GF2T_EINVAL;
    }
    GF2T_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return GF2T_OK;
}

extern "C" int gf2t_graph_end(void *stream, gf2t_graph **out) {
    if (!stream || !out) return GF2T_EINVAL;
    *out = nullptr;
    hipGraph_t g = nullptr;
    GF2T_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (rc != hipSuccess) {
        hipGraphDestroy(g);
        return gf2t::hip_fail(rc, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    *out = new gf2t_graph{g, e};
    return GF2T_OK;
}

extern "C" int gf2t_graph_launch(gf2t_graph *g, void *stream) {
    if (!g || !g->exec) return GF2T_EINVAL;
    GF2T_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return GF2T_OK;
}

extern "C" int gf2t_graph_nodes(const gf2t_graph *g, size_t *nodes) {
    if (!g || !nodes) return GF2T_EINVAL;
    *nodes = 0;
    GF2T_HIP(hipGraphGetNodes(g->graph, nullptr, nodes));
    return GF2T_OK;
}

extern "C" void gf2t_graph_destroy(gf2t_graph *g) {
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
}

extern "C" int gf2t_stream_sync(void *stream) {
    GF2T_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GF2T_OK;
}

extern "C" int gf2t_device_sync(void) {
    GF2T_HIP(hipDeviceSynchronize());
    return GF2T_OK;
}

extern "C" int gf2t_event_create(void **event) {
    if (!event) return GF2T_EINVAL;
    hipEvent_t e;
    GF2T_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return GF2T_OK;
}

extern "C" int gf2t_event_destroy(void *event) {
    if (event) GF2T_HIP(hipEventDestroy((hipEvent_t)event));
    return GF2T_OK;
}

extern "C" int gf2t_event_record(void *event, void *stream) {
    GF2T_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return GF2T_OK;
}

extern "C" int gf2t_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms) return GF2T_EINVAL;
    GF2T_HIP(hipEventSynchronize((hipEvent_t)stop));
    GF2T_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GF2T_OK;
}
