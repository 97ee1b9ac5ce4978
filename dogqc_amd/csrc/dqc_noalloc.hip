// dqc_query_run_noalloc: a plan run that allocates no memory (part of
// libdqcrt.so, C ABI in include/dogqc_amd.h).
//
// dqc_query_run allocates the plan's workspace (status block, hash tables,
// multimaps, temp tables, long-list queues, partition segments, result
// columns) with hipMalloc.  Callers that cannot allocate (code under stream
// capture, tight serving loops, hosts that forbid it) use this entry point
// instead: the workspace is a fixed device array of DQC_NOALLOC_WS_BYTES
// compiled into the library, one copy per GPU, held until the process ends.
// The caller passes the input columns and one device buffer per result
// column; the result rows are copied there at the end of the run.
//
// The array is shared by every call on a GPU and keeps the last call's
// contents.  Every run starts with the plan's own init kernel, which clears
// the status block and fills every table region, so nothing of an earlier
// call is read.  Two calls that overlapped on one GPU would race on the
// array; a per-device lock held for the whole call (the call waits for its
// stream before returning) serialises calls from different host threads.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "dogqc_amd.h"
#include "dqc_runtime_internal.h"

using dqcrt::fail;
using dqcrt::hipFail;
using dqcrt::Module;

// In host code `scratch` names a host-side stand-in; the device address of
// this GPU's copy comes from hipGetSymbolAddress (scratchOf below).
static __device__ __align__(256) unsigned char scratch[DQC_NOALLOC_WS_BYTES];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_initLock;
void* g_scratch[kMaxDevices];       // device address of `scratch`, per device
std::mutex g_runLock[kMaxDevices];  // one run at a time per device

// Device address of the current device's copy of `scratch` (looked up once
// per device, with that device current).
int scratchOf(int dev, void** out) {
    std::lock_guard<std::mutex> g(g_initLock);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hipFail(e, "hipGetSymbolAddress(scratch)");
        if (p == nullptr) return fail(DQC_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return DQC_OK;
}

int ctypeBytes(const char* s, size_t n) {
    struct W { const char* name; int bytes; };
    static const W table[] = {{"int8_t", 1}, {"char", 1}, {"int32_t", 4}, {"uint32_t", 4}, {"float", 4},
                              {"int64_t", 8}, {"uint64_t", 8}, {"double", 8}, {"dqc::str_offs", 16},
                              {"dqc::str_t", 16}};
    for (const W& w : table)
        if (strlen(w.name) == n && strncmp(w.name, s, n) == 0) return w.bytes;
    return -1;
}

// Byte widths of the result columns, read from the module's describe() JSON
// (its "result" array of {"name", "type", "ctype", ...} objects, written by
// the translator).  Returns the number of columns, or -1 on a malformed
// description.
int resultWidths(const char* js, int* widths, int maxCols) {
    // the key "result" (a string "result" elsewhere is not followed by ':')
    const char* p = js;
    for (;;) {
        p = strstr(p, "\"result\"");
        if (!p) return -1;
        p += 8;
        while (*p == ' ') ++p;
        if (*p == ':') break;
    }
    ++p;
    while (*p == ' ') ++p;
    if (*p != '[') return -1;
    ++p;
    int n = 0, depth = 0;
    bool wantValue = false;
    while (*p) {
        const char c = *p;
        if (c == '"') {
            const char* s = ++p;
            while (*p && *p != '"') p += (*p == '\\' && p[1]) ? 2 : 1;
            if (!*p) return -1;
            const size_t len = (size_t)(p - s);
            ++p;
            if (wantValue) {
                if (n >= maxCols) return -1;
                widths[n] = ctypeBytes(s, len);
                if (widths[n] < 0) return -1;
                ++n;
                wantValue = false;
            } else if (depth == 1 && len == 5 && strncmp(s, "ctype", 5) == 0) {
                while (*p == ' ' || *p == ':') ++p;
                wantValue = *p == '"';
            }
            continue;
        }
        if (c == '{' || c == '[') ++depth;
        if (c == '}' || c == ']') {
            if (depth == 0) return n;  // end of the result array
            --depth;
        }
        ++p;
    }
    return -1;
}

}  // namespace

extern "C" int dqc_query_run_noalloc(void* module, const dqc_plan_inputs* in, dqc_plan_tuning* tun, void* stream,
                                     int32_t max_retries, void* const* out_cols, int64_t out_rows,
                                     dqc_query_result* res) {
    if (!module || !in || !tun || !res) return fail(DQC_EINVAL, "dqc_query_run_noalloc: null argument");
    if (out_rows < 0) return fail(DQC_EINVAL, "dqc_query_run_noalloc: negative output row count");
    memset(res, 0, sizeof(*res));
    Module* m = (Module*)module;
    int widths[DQC_MAX_RESULT_COLS];
    const int ncols = resultWidths(m->describe(), widths, DQC_MAX_RESULT_COLS);
    if (ncols < 0) return fail(DQC_EINVAL, "dqc_query_run_noalloc: unreadable result schema in the plan module");
    if (ncols > 0 && !out_cols) return fail(DQC_EINVAL, "dqc_query_run_noalloc: null output column array");
    for (int c = 0; c < ncols; ++c)
        if (!out_cols[c]) return fail(DQC_EINVAL, "dqc_query_run_noalloc: output column %d is null", c);

    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(DQC_EINVAL, "device %d beyond the supported %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> hold(g_runLock[dev]);
    void* ws = nullptr;
    int rc = scratchOf(dev, &ws);
    if (rc) return rc;

    hipStream_t s = (hipStream_t)stream;
    for (int attempt = 0;; ++attempt) {
        dqc_plan_layout L;
        rc = m->layout(in, tun, &L);
        if (rc) return fail(rc, "plan layout failed (%d)", rc);
        if (L.ws_bytes > (uint64_t)DQC_NOALLOC_WS_BYTES)
            return fail(DQC_ENOMEM,
                        "dqc_query_run_noalloc: the plan needs a %llu-byte workspace, beyond the %llu-byte "
                        "static scratch (use dqc_query_run)",
                        (unsigned long long)L.ws_bytes, (unsigned long long)DQC_NOALLOC_WS_BYTES);
        if (L.nresult != ncols) return fail(DQC_EINVAL, "result schema and layout disagree (%d vs %d columns)",
                                            ncols, L.nresult);
        rc = m->enqueue(in, tun, ws, (uint64_t)DQC_NOALLOC_WS_BYTES, stream, nullptr);
        if (rc) return fail(rc, "plan enqueue failed (%d)", rc);
        HIP_CHECK(hipMemcpyAsync(&res->status, ws, sizeof(dqc_plan_status), hipMemcpyDeviceToHost, s),
                  "status copy");
        HIP_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
        res->layout = L;
        res->retries = attempt;
        if (!dqcrt::adjustForRetry(L, res->status, tun)) break;
        if (attempt >= max_retries)
            return fail(DQC_EOVERFLOW, "tables still too small after %d retries (flags 0x%x)", attempt,
                        res->status.flags);
    }
    const uint64_t n = res->status.result_count;
    if (n > (uint64_t)out_rows)
        return fail(DQC_EOVERFLOW, "dqc_query_run_noalloc: the result has %llu rows, the output columns hold %lld",
                    (unsigned long long)n, (long long)out_rows);
    if (n > 0) {
        const unsigned char* W = (const unsigned char*)ws;
        for (int c = 0; c < ncols; ++c)
            HIP_CHECK(hipMemcpyAsync(out_cols[c], W + res->layout.result_offset[c], n * (uint64_t)widths[c],
                                     hipMemcpyDeviceToDevice, s),
                      "result column copy");
        HIP_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    return DQC_OK;
}
