// libdqcrt.so: host runtime of the MI355X DogQC backend (C ABI in
// include/dogqc_amd.h).
//
// Replaces, from the reference, the host half of every generated program:
// mmdb column mapping (dogqc/include/mappedmalloc.h:69-82), device allocation
// and copies (dogqc/gpuio.py:5-58), running the compiled query
// (dogqc/codegen.py:209-218) and the exit()-on-error policy (util.h:2-5),
// which here becomes return codes plus an automatic resize-and-retry when a
// hash table, multimap or result buffer turns out too small.
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "dogqc_amd.h"
#include "dqc_runtime_internal.h"

static_assert(sizeof(dqc_plan_tuning) == 1640, "dqc_plan_tuning layout");
static_assert(sizeof(dqc_plan_status) == 856, "dqc_plan_status layout");
static_assert(sizeof(dqc_plan_layout) == 1904, "dqc_plan_layout layout");
static_assert(sizeof(dqc_plan_inputs) == 24, "dqc_plan_inputs layout");

using dqcrt::Module;
using dqcrt::fail;
using dqcrt::hipFail;

namespace {
thread_local std::string g_err;
}  // namespace

namespace dqcrt {

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hipFail(hipError_t e, const char* what) {
    return fail(DQC_EHIP, "%s: %s", what, hipGetErrorString(e));
}

bool adjustForRetry(const dqc_plan_layout& L, const dqc_plan_status& st, dqc_plan_tuning* tun) {
    bool retry = false;
    if (st.flags & DQC_ST_HT_OVERFLOW) {
        for (int i = 0; i < L.nhts; ++i)
            if (st.ht_overflow_mask & (1u << i)) tun->ht_capacity[i] = std::max<int64_t>(64, L.ht_capacity[i] * 4);
        retry = true;
    }
    if (st.flags & DQC_ST_MM_OVERFLOW) {
        for (int i = 0; i < L.nhts; ++i)
            if (st.mm_overflow_mask & (1u << i))
                tun->mm_entries[i] = std::max<int64_t>((int64_t)st.mm_total[i], L.mm_entries[i] + 1);
        retry = true;
    }
    if (st.result_count > (uint64_t)L.result_capacity && L.nresult > 0) {
        tun->result_capacity = (int64_t)st.result_count;
        retry = true;
    }
    for (int i = 0; i < L.ntemps && i < DQC_MAX_TEMPS; ++i) {
        if (st.temp_count[i] > (uint64_t)L.temp_capacity[i]) {
            tun->temp_capacity[i] = (int64_t)st.temp_count[i];
            retry = true;
        }
    }
    for (int i = 0; i < L.nhts; ++i) {
        if (L.split_capacity[i] > 0 && st.split_count[i] > (uint64_t)L.split_capacity[i]) {
            tun->split_capacity[i] = (int64_t)st.split_count[i];
            retry = true;
        }
    }
    return retry;
}

}  // namespace dqcrt

extern "C" {

const char* dqc_last_error(void) { return g_err.c_str(); }

int dqc_device_init(int device, int32_t* compute_units, uint64_t* hbm_bytes) {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(DQC_EINVAL, "device %d not present (%d devices)", device, n);
    HIP_CHECK(hipSetDevice(device), "hipSetDevice");
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    size_t freeB = 0, totalB = 0;
    HIP_CHECK(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    if (compute_units) *compute_units = cus;
    if (hbm_bytes) *hbm_bytes = totalB;
    return DQC_OK;
}

int dqc_malloc(void** ptr, uint64_t bytes) {
    if (!ptr) return fail(DQC_EINVAL, "dqc_malloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(DQC_ENOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    return DQC_OK;
}

int dqc_free(void* ptr) {
    if (ptr) HIP_CHECK(hipFree(ptr), "hipFree");
    return DQC_OK;
}

int dqc_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return DQC_OK;
}

int dqc_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return DQC_OK;
}

int dqc_memset(void* dst, int value, uint64_t bytes) {
    if (bytes) HIP_CHECK(hipMemset(dst, value, bytes), "hipMemset");
    return DQC_OK;
}

int dqc_stream_create(void** stream) {
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = (void*)s;
    return DQC_OK;
}

int dqc_stream_destroy(void* stream) {
    if (stream) HIP_CHECK(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
    return DQC_OK;
}

int dqc_stream_sync(void* stream) {
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return DQC_OK;
}

int dqc_event_create(void** event) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e), "hipEventCreate");
    *event = (void*)e;
    return DQC_OK;
}

int dqc_event_destroy(void* event) {
    if (event) HIP_CHECK(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
    return DQC_OK;
}

int dqc_event_record(void* event, void* stream) {
    HIP_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
    return DQC_OK;
}

int dqc_event_elapsed_ms(void* start, void* stop, float* ms) {
    HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
    return DQC_OK;
}

int dqc_mmdb_load(const char* path, void** dptr, uint64_t* bytes) {
    if (!path || !dptr) return fail(DQC_EINVAL, "dqc_mmdb_load: null argument");
    *dptr = nullptr;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(DQC_EIO, "cannot open %s", path);
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return fail(DQC_EIO, "cannot stat %s", path);
    }
    uint64_t filesize = 0;
    if (read(fd, &filesize, 8) != 8 || filesize < 9 || filesize > (uint64_t)sb.st_size) {
        close(fd);
        return fail(DQC_EIO, "bad mmdb header in %s", path);
    }
    uint64_t payload = filesize - 9;
    void* map = mmap(nullptr, filesize, PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return fail(DQC_EIO, "cannot mmap %s", path);
    int rc = dqc_malloc(dptr, payload);
    if (rc == DQC_OK) rc = dqc_memcpy_h2d(*dptr, (const char*)map + 8, payload);
    munmap(map, filesize);
    if (rc != DQC_OK) {
        if (*dptr) dqc_free(*dptr);
        *dptr = nullptr;
        return rc;
    }
    if (bytes) *bytes = payload;
    return DQC_OK;
}

int dqc_module_load(const char* so_path, void** module) {
    if (!so_path || !module) return fail(DQC_EINVAL, "dqc_module_load: null argument");
    void* h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return fail(DQC_ELOAD, "dlopen %s: %s", so_path, dlerror());
    Module* m = new Module();
    m->handle = h;
    m->describe = (const char* (*)(void))dlsym(h, "dqc_plan_describe");
    m->layout = (int (*)(const dqc_plan_inputs*, const dqc_plan_tuning*, dqc_plan_layout*))dlsym(h, "dqc_plan_layout_of");
    m->enqueue = (int (*)(const dqc_plan_inputs*, const dqc_plan_tuning*, void*, uint64_t, void*, void* const*))dlsym(
        h, "dqc_plan_enqueue");
    if (!m->describe || !m->layout || !m->enqueue) {
        dlclose(h);
        delete m;
        return fail(DQC_ELOAD, "%s does not export the dqc_plan ABI", so_path);
    }
    *module = m;
    return DQC_OK;
}

int dqc_module_unload(void* module) {
    Module* m = (Module*)module;
    if (!m) return DQC_OK;
    dlclose(m->handle);
    delete m;
    return DQC_OK;
}

const char* dqc_module_describe(void* module) { return module ? ((Module*)module)->describe() : ""; }

int dqc_query_layout(void* module, const dqc_plan_inputs* in, const dqc_plan_tuning* tun, dqc_plan_layout* out) {
    if (!module) return fail(DQC_EINVAL, "null module");
    int rc = ((Module*)module)->layout(in, tun, out);
    if (rc) return fail(rc, "plan layout failed (%d): input binding does not match the plan", rc);
    return DQC_OK;
}

int dqc_query_enqueue(void* module, const dqc_plan_inputs* in, const dqc_plan_tuning* tun, void* ws,
                      uint64_t ws_bytes, void* stream, void* const* step_events) {
    if (!module) return fail(DQC_EINVAL, "null module");
    int rc = ((Module*)module)->enqueue(in, tun, ws, ws_bytes, stream, step_events);
    if (rc) return fail(rc, "plan enqueue failed (%d)", rc);
    return DQC_OK;
}

int dqc_query_run(void* module, const dqc_plan_inputs* in, dqc_plan_tuning* tun, void* stream,
                  int32_t max_retries, dqc_query_result* res) {
    if (!module || !in || !tun || !res) return fail(DQC_EINVAL, "dqc_query_run: null argument");
    memset(res, 0, sizeof(*res));
    Module* m = (Module*)module;
    for (int attempt = 0;; ++attempt) {
        dqc_plan_layout L;
        int rc = m->layout(in, tun, &L);
        if (rc) return fail(rc, "plan layout failed (%d)", rc);
        if (res->ws == nullptr || res->ws_bytes < L.ws_bytes) {
            if (res->ws) dqc_free(res->ws);
            res->ws = nullptr;
            rc = dqc_malloc(&res->ws, L.ws_bytes);
            if (rc) return rc;
            res->ws_bytes = L.ws_bytes;
        }
        rc = m->enqueue(in, tun, res->ws, res->ws_bytes, stream, nullptr);
        if (rc) return fail(rc, "plan enqueue failed (%d)", rc);
        HIP_CHECK(hipMemcpyAsync(&res->status, res->ws, sizeof(dqc_plan_status), hipMemcpyDeviceToHost,
                                 (hipStream_t)stream), "status copy");
        HIP_CHECK(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        res->layout = L;
        res->retries = attempt;
        const dqc_plan_status& st = res->status;
        const bool retry = dqcrt::adjustForRetry(L, st, tun);
        if (!retry) return DQC_OK;
        if (attempt >= max_retries)
            return fail(DQC_EOVERFLOW, "tables still too small after %d retries (flags 0x%x)", attempt, st.flags);
    }
}

int dqc_query_release(dqc_query_result* res) {
    if (!res) return DQC_OK;
    int rc = dqc_free(res->ws);
    res->ws = nullptr;
    res->ws_bytes = 0;
    return rc;
}

int dqc_graph_create(void* module, const dqc_plan_inputs* in, const dqc_plan_tuning* tun, void* ws,
                     uint64_t ws_bytes, void* stream, void** graph_exec) {
    if (!module || !graph_exec) return fail(DQC_EINVAL, "dqc_graph_create: null argument");
    hipStream_t s = (hipStream_t)stream;
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    int rc = ((Module*)module)->enqueue(in, tun, ws, ws_bytes, stream, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return fail(rc, "plan enqueue failed during capture (%d)", rc);
    }
    if (e != hipSuccess) return hipFail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hipFail(e, "hipGraphInstantiate");
    *graph_exec = (void*)x;
    return DQC_OK;
}

int dqc_graph_launch(void* graph_exec, void* stream) {
    HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
    return DQC_OK;
}

int dqc_graph_destroy(void* graph_exec) {
    if (graph_exec) HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
    return DQC_OK;
}

}  // extern "C"
