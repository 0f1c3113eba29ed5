// Graph capture against the library's other HIP calls.
//
// A stream capture in progress can be invalidated by HIP calls that another
// host thread makes meanwhile (a service factorizes one handle while its
// solve thread runs another; seen even with thread-local capture mode on a
// private non-blocking stream).  So a capture runs while no other host
// thread is inside the library: every exported entry point that makes HIP
// calls holds the lock shared for its duration (ApiShared), and the capture
// holds it exclusively (CaptureExclusive; host-side enqueue only, a few
// milliseconds).  Nested entry points share the outer hold.
#pragma once

namespace ssids_hip {

struct ApiShared {
  ApiShared();
  ~ApiShared();
  ApiShared(const ApiShared &) = delete;
  ApiShared &operator=(const ApiShared &) = delete;
};

// Taken inside an entry point: gives up this thread's shared hold, waits for
// the other threads' calls to return, and takes the shared hold back after.
struct CaptureExclusive {
  CaptureExclusive();
  ~CaptureExclusive();
  CaptureExclusive(const CaptureExclusive &) = delete;
  CaptureExclusive &operator=(const CaptureExclusive &) = delete;

 private:
  bool nested_;
};

}  // namespace ssids_hip
