// streams.hip -- plain stream / event / graph helpers for the host layer
// (default-priority streams only).
#include <hip/hip_runtime.h>
#include "../shk.h"

#define CHK(call)                                                           \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      {                                                                     \
        shk_set_error ("%s: %s", #call, hipGetErrorString (e_));            \
        return SHK_EHIP;                                                    \
      }                                                                     \
  } while (0)

extern "C" int shk_stream_create (void **s)
{
  hipStream_t h = nullptr;
  CHK (hipStreamCreateWithFlags (&h, hipStreamNonBlocking));
  *s = (void *) h;
  return SHK_OK;
}

extern "C" void shk_stream_destroy (void *s)
{
  if (s) (void) hipStreamDestroy ((hipStream_t) s);
}

extern "C" int shk_stream_sync (void *s)
{
  CHK (hipStreamSynchronize ((hipStream_t) s));
  return SHK_OK;
}

extern "C" int shk_event_create (void **e)
{
  hipEvent_t h = nullptr;
  CHK (hipEventCreateWithFlags (&h, hipEventDisableTiming));
  *e = (void *) h;
  return SHK_OK;
}

extern "C" void shk_event_destroy (void *e)
{
  if (e) (void) hipEventDestroy ((hipEvent_t) e);
}

extern "C" int shk_event_record (void *e, void *s)
{
  CHK (hipEventRecord ((hipEvent_t) e, (hipStream_t) s));
  return SHK_OK;
}

extern "C" int shk_stream_wait (void *s, void *e)
{
  CHK (hipStreamWaitEvent ((hipStream_t) s, (hipEvent_t) e, 0));
  return SHK_OK;
}

extern "C" int shk_stream_capture_begin (void *s)
{
  CHK (hipStreamBeginCapture ((hipStream_t) s, hipStreamCaptureModeThreadLocal));
  return SHK_OK;
}

extern "C" int shk_stream_capture_end (void *s, void **exec)
{
  hipGraph_t g = nullptr;
  CHK (hipStreamEndCapture ((hipStream_t) s, &g));
  hipGraphExec_t x = nullptr;
  const hipError_t rc = hipGraphInstantiate (&x, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy (g);
  CHK (rc);
  *exec = (void *) x;
  return SHK_OK;
}

extern "C" void shk_stream_capture_abort (void *s)
{
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture ((hipStream_t) s, &g) == hipSuccess && g) (void) hipGraphDestroy (g);
  (void) hipGetLastError ();
}

extern "C" int shk_graph_launch_on (void *exec, void *s)
{
  CHK (hipGraphLaunch ((hipGraphExec_t) exec, (hipStream_t) s));
  return SHK_OK;
}
