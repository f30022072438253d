/*
 * scratch.cpp - device workspaces of the entry points whose reference
 * interface has no workspace argument.
 *
 * The reference-interface gradient, statistics and metric entry points take
 * only their float outputs (gradients.cu:486-530, utils.cu:258-345).  Here
 * they accumulate first into exact fixed-point lines (fix_split,
 * sad_device.h) and then write the floats, so their results do not depend
 * on the order of the atomics.  The lines live in a workspace owned by the
 * library, one per (device, stream): calls on one stream are ordered, calls
 * on different streams use different workspaces.  A workspace grows when a
 * call needs more (the stream is drained before the old one is freed) and is
 * kept for the life of the process; the table that holds them is never
 * destroyed, so no HIP call runs from a static destructor after the runtime
 * has shut down.
 */
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "sad_device.h"

namespace sadh {

namespace {

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct WorkspaceTable {
    std::mutex mu;
    std::map<std::pair<int, hipStream_t>, Workspace> ws;
};

WorkspaceTable& table()
{
    static WorkspaceTable* t = new WorkspaceTable();   /* intentionally never destroyed */
    return *t;
}

}  // namespace

int stream_workspace(hipStream_t stream, size_t bytes, void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, std::string("workspace: hipGetDevice: ") + hipGetErrorString(e));
    WorkspaceTable& T = table();
    std::lock_guard<std::mutex> lock(T.mu);
    Workspace& w = T.ws[std::make_pair(dev, stream)];
    if (w.bytes < bytes) {
        if (w.ptr) {
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return fail((int)e, std::string("workspace: drain: ") + hipGetErrorString(e));
            (void)hipFree(w.ptr);
            w.ptr = nullptr;
            w.bytes = 0;
        }
        const size_t want = bytes + bytes / 4 + 4096;   /* some headroom for the next, larger call */
        e = hipMalloc(&w.ptr, want);
        if (e != hipSuccess) {
            w.ptr = nullptr;
            return fail((int)e, "workspace: hipMalloc(" + std::to_string(want) + " B) failed");
        }
        w.bytes = want;
    }
    *out = w.ptr;
    return SADH_OK;
}

}  // namespace sadh
