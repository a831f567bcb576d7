// HIP events on a context's stream, for timing submitted work from the host without a framework:
// the elapsed time between two events recorded on the stream the kernels run on.
#include <hip/hip_runtime.h>

#include "../../include/fort_scope.h"

extern "C" {

int ftc_event_create(void **event)
{
    hipEvent_t e = nullptr;
    if (!event || hipEventCreate(&e) != hipSuccess) return FORT_SCOPE_RESOURCE;
    *event = e;
    return FORT_SCOPE_OK;
}

int ftc_event_record(void *event, void *stream)
{
    return hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)) == hipSuccess
               ? FORT_SCOPE_OK
               : FORT_SCOPE_EXECUTION;
}

int ftc_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!ms) return FORT_SCOPE_ARGUMENT;
    if (hipEventSynchronize(static_cast<hipEvent_t>(stop)) != hipSuccess) return FORT_SCOPE_EXECUTION;
    return hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)) == hipSuccess
               ? FORT_SCOPE_OK
               : FORT_SCOPE_EXECUTION;
}

int ftc_event_destroy(void *event)
{
    return hipEventDestroy(static_cast<hipEvent_t>(event)) == hipSuccess ? FORT_SCOPE_OK : FORT_SCOPE_ARGUMENT;
}

// The number of visible devices; 0 without a GPU.
int ftc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // extern "C"
