// Internal host-side helpers shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/fastpli_amd.h"

namespace fpl {

// Error carrier: every entry point fills `msg` and returns a status code.
struct Status {
   int code = FPL_OK;
   std::string msg;
   bool ok() const { return code == FPL_OK; }
};

inline Status make_status(int code, const char *fmt, ...) {
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   Status s;
   s.code = code;
   s.msg = buf;
   return s;
}

inline int finish(const Status &s, char *msg, size_t msg_len) {
   if (msg && msg_len > 0) {
      size_t n = s.msg.size() < msg_len - 1 ? s.msg.size() : msg_len - 1;
      memcpy(msg, s.msg.data(), n);
      msg[n] = '\0';
   }
   return s.code;
}

#define FPL_HIP(expr)                                                         \
   do {                                                                       \
      hipError_t e_ = (expr);                                                 \
      if (e_ != hipSuccess)                                                   \
         return ::fpl::make_status(FPL_E_DEVICE, "%s failed: %s", #expr,      \
                                   hipGetErrorString(e_));                    \
   } while (0)

#define FPL_TRY(expr)                                                         \
   do {                                                                       \
      ::fpl::Status s_ = (expr);                                              \
      if (!s_.ok()) return s_;                                                \
   } while (0)

// Grow-only device scratch buffers, one set per (device, stream).  Each
// call of an entry point owns the set of its stream for the call's duration;
// concurrent calls on different streams use different sets.
enum Slot : int {
   WS_BOXES = 0,   // generation: fiber + bundle AABBs (ordered u64)
   WS_FIBER_OK,    // generation: per fiber flags
   WS_SEGS,        // generation: segment records
   WS_SEG_COUNT,   // generation: per segment tile count
   WS_SEG_OFF,     // generation: exclusive scan of counts
   WS_KEYS_A,      // generation: (tile) keys
   WS_VALS_A,      // generation: (segment) values
   WS_KEYS_B,
   WS_VALS_B,
   WS_TILE_RANGE,  // generation: per tile [begin, end)
   WS_CUB_TEMP,    // hipcub temporary storage
   WS_SMALL,       // small per-call tables (bundles, layers, tilts, props)
   WS_FLAGS,       // warning / error words
   WS_PACKED,      // simulation: packed (label, axis) records in [x][z][y] order
   WS_PACKED_B,    // simulation: second array of the angle storage
   WS_ANALYSIS,    // optic: two image-sized temporaries
   WS_ADJOINT,     // adjoint march: checkpoints + the gradient table
   WS_NUM_SLOTS
};

Status ws_get(void *stream, int slot, size_t bytes, void **ptr);

// Device bytes the workspace of `stream` holds, and its release (the stream
// must be idle).
size_t ws_bytes(void *stream);
void ws_free(void *stream);

// The (tile, segment) pair count of the stream's last generation of a scene
// shape (points, fibers, nx, ny, nz; 0: none), which the next generation of
// that shape assumes instead of reading its own back mid-call (a serving slot
// alternates between scenes: one hint per stream would mostly be wrong).
int64_t pair_hint(void *stream, const std::array<int64_t, 5> &scene);
void set_pair_hint(void *stream, const std::array<int64_t, 5> &scene, int64_t n);

// Pinned host staging for the small device->host readbacks of a call.
Status host_stage(void *stream, size_t bytes, void **ptr);

// Optional kernel timing (fpl_profile_enable).  Usage around one launch:
//   KernelTimer kt("march", stream);  launch;  kt.stop();
// and once the launch is queued for good: kt.collect() hands the two events
// over; their elapsed time is read when fpl_profile_read asks for it.
struct KernelTimer {
   const char *name;
   hipStream_t stream;
   hipEvent_t a = nullptr, b = nullptr;
   bool on = false;
   KernelTimer(const char *n, hipStream_t s);
   void stop();
   void collect();
   ~KernelTimer();
};

} // namespace fpl
