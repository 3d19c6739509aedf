// Library-wide entry points and the per-stream device workspace.
#include <array>
#include <cmath>
#include <map>
#include <string>
#include <mutex>
#include <utility>
#include <vector>

#include "fpl_internal.h"
#include "fpl_records.h"

namespace fpl {
namespace {

struct Workspace {
   void *ptr[WS_NUM_SLOTS] = {};
   size_t cap[WS_NUM_SLOTS] = {};
   void *host = nullptr;
   size_t host_cap = 0;
   // generation: the (tile, segment) pair count of the last call per scene shape
   std::map<std::array<int64_t, 5>, int64_t> pair_hints;
};

std::mutex g_mutex;
std::map<std::pair<int, void *>, Workspace> g_ws;

Workspace &ws_for(void *stream) {
   int dev = 0;
   (void)hipGetDevice(&dev);
   std::lock_guard<std::mutex> lock(g_mutex);
   return g_ws[{dev, stream}];
}

} // namespace

int64_t pair_hint(void *stream, const std::array<int64_t, 5> &key) {
   Workspace &w = ws_for(stream);
   auto it = w.pair_hints.find(key);
   return it == w.pair_hints.end() ? 0 : it->second;
}

void set_pair_hint(void *stream, const std::array<int64_t, 5> &key, int64_t n) {
   Workspace &w = ws_for(stream);
   if (w.pair_hints.size() >= 256) w.pair_hints.clear();
   w.pair_hints[key] = n;
}

Status ws_get(void *stream, int slot, size_t bytes, void **ptr) {
   Workspace &w = ws_for(stream);
   if (bytes == 0) bytes = 1;
   if (w.cap[slot] < bytes) {
      if (w.ptr[slot]) {
         // the stream may still use the old buffer
         FPL_HIP(hipStreamSynchronize((hipStream_t)stream));
         FPL_HIP(hipFree(w.ptr[slot]));
         w.ptr[slot] = nullptr;
         w.cap[slot] = 0;
      }
      // grow by 1/8 to avoid re-allocating on every slightly larger call
      size_t want = bytes + bytes / 8;
      want = (want + 255) & ~size_t(255);
      hipError_t e = hipMalloc(&w.ptr[slot], want);
      if (e != hipSuccess) {
         (void)hipGetLastError();
         w.ptr[slot] = nullptr;
         return make_status(FPL_E_NOMEM, "device allocation of %zu bytes failed: %s", want,
                            hipGetErrorString(e));
      }
      w.cap[slot] = want;
   }
   *ptr = w.ptr[slot];
   return Status();
}

size_t ws_bytes(void *stream) {
   Workspace &w = ws_for(stream);
   size_t n = 0;
   for (int k = 0; k < WS_NUM_SLOTS; k++) n += w.cap[k];
   return n;
}

void ws_free(void *stream) {
   Workspace &w = ws_for(stream);
   (void)hipStreamSynchronize((hipStream_t)stream);
   for (int k = 0; k < WS_NUM_SLOTS; k++) {
      if (w.ptr[k]) (void)hipFree(w.ptr[k]);
      w.ptr[k] = nullptr;
      w.cap[k] = 0;
   }
}

Status host_stage(void *stream, size_t bytes, void **ptr) {
   Workspace &w = ws_for(stream);
   if (w.host_cap < bytes) {
      if (w.host) {
         FPL_HIP(hipStreamSynchronize((hipStream_t)stream));
         FPL_HIP(hipHostFree(w.host));
         w.host = nullptr;
         w.host_cap = 0;
      }
      size_t want = bytes < 4096 ? 4096 : bytes;
      FPL_HIP(hipHostMalloc(&w.host, want, hipHostMallocDefault));
      w.host_cap = want;
   }
   *ptr = w.host;
   return Status();
}

namespace {
std::mutex g_prof_mutex;
bool g_prof_on = false;
std::map<std::string, std::pair<double, int64_t>> g_prof;
// timing events are reused (creating and destroying two events per launch
// held up the launches that follow); a timer that is never collected (an
// error path) destroys its events
std::vector<hipEvent_t> g_event_pool;

bool take_event(hipEvent_t *e) {
   {
      std::lock_guard<std::mutex> lock(g_prof_mutex);
      if (!g_event_pool.empty()) {
         *e = g_event_pool.back();
         g_event_pool.pop_back();
         return true;
      }
   }
   return hipEventCreate(e) == hipSuccess;
}

// Finished timers wait here (their events recorded, not yet read): reading
// an elapsed time at every launch costs host time between the launches of a
// step, so the times are read when they are asked for (fpl_profile_read) or
// when many have gathered.
struct PendingTime {
   const char *name;
   hipEvent_t a, b;
};
std::vector<PendingTime> g_pending;

// caller holds g_prof_mutex
void drain_pending() {
   for (const PendingTime &p : g_pending) {
      float ms = 0.0f;
      if (hipEventSynchronize(p.b) == hipSuccess &&
          hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
         auto &e = g_prof[p.name];
         e.first += ms;
         e.second += 1;
         g_event_pool.push_back(p.a);
         g_event_pool.push_back(p.b);
      } else {
         (void)hipEventDestroy(p.a);
         (void)hipEventDestroy(p.b);
      }
   }
   g_pending.clear();
}
} // namespace

KernelTimer::KernelTimer(const char *n, hipStream_t s) : name(n), stream(s) {
   {
      std::lock_guard<std::mutex> lock(g_prof_mutex);
      on = g_prof_on;
   }
   if (!on) return;
   if (!take_event(&a) || !take_event(&b)) {
      on = false;
      return;
   }
   (void)hipEventRecord(a, stream);
}

void KernelTimer::stop() {
   if (on) (void)hipEventRecord(b, stream);
}

void KernelTimer::collect() {
   if (!on) return;
   std::lock_guard<std::mutex> lock(g_prof_mutex);
   g_pending.push_back({name, a, b});
   a = b = nullptr;
   if (g_pending.size() >= 1024) drain_pending();
   on = false;
}

KernelTimer::~KernelTimer() {
   if (a) (void)hipEventDestroy(a);
   if (b) (void)hipEventDestroy(b);
}

} // namespace fpl

extern "C" {

void fpl_profile_enable(int enable) {
   std::lock_guard<std::mutex> lock(fpl::g_prof_mutex);
   fpl::g_prof_on = enable != 0;
   if (enable) {
      fpl::drain_pending();
      fpl::g_prof.clear();
   }
}

int fpl_profile_read(const char *name, double *total_ms, int64_t *launches) {
   std::lock_guard<std::mutex> lock(fpl::g_prof_mutex);
   fpl::drain_pending();
   auto it = fpl::g_prof.find(name ? name : "");
   if (it == fpl::g_prof.end()) {
      if (total_ms) *total_ms = 0;
      if (launches) *launches = 0;
      return -1;
   }
   if (total_ms) *total_ms = it->second.first;
   if (launches) *launches = it->second.second;
   return 0;
}

const char *fpl_version(void) { return "fastpli_amd 0.1.0 (gfx950)"; }

void fpl_release_workspace(void) {
   int dev = 0;
   (void)hipGetDevice(&dev);
   std::lock_guard<std::mutex> lock(fpl::g_mutex);
   for (auto it = fpl::g_ws.begin(); it != fpl::g_ws.end();) {
      if (it->first.first != dev) {
         ++it;
         continue;
      }
      (void)hipStreamSynchronize((hipStream_t)it->first.second);
      for (int s = 0; s < fpl::WS_NUM_SLOTS; s++)
         if (it->second.ptr[s]) (void)hipFree(it->second.ptr[s]);
      if (it->second.host) (void)hipHostFree(it->second.host);
      it = fpl::g_ws.erase(it);
   }
}

size_t fpl_record_bytes(int32_t field_storage) {
   return field_storage < 0 || field_storage > 2 ? 0 : fpl::rec::record_bytes(field_storage);
}

size_t fpl_fields_codes_offset(const int64_t dim[3], int32_t field_storage) {
   if (!dim || dim[0] <= 0 || dim[1] <= 0 || dim[2] <= 0 || field_storage < 0 ||
       field_storage > 2)
      return 0;
   const size_t n = (size_t)dim[0] * (size_t)dim[1] * (size_t)dim[2];
   return (n * fpl::rec::record_bytes(field_storage) + 255) & ~(size_t)255;
}

size_t fpl_fields_bytes(const int64_t dim[3], int32_t field_storage) {
   const size_t off = fpl_fields_codes_offset(dim, field_storage);
   if (off == 0) return 0;
   const size_t n = (size_t)dim[0] * (size_t)dim[1] * (size_t)dim[2];
   const size_t sb = fpl::rec::second_bytes(field_storage);
   return sb ? off + n * sb : n * fpl::rec::record_bytes(field_storage);
}

int64_t fpl_num_properties(const fpl_fiber_bundles *fbs) {
   if (!fbs) return 1;
   int64_t max_layer = 0;
   for (int64_t b = 0; b < fbs->n_bundles; b++) {
      int64_t n = fbs->layer_ptr[b + 1] - fbs->layer_ptr[b];
      if (n > max_layer) max_layer = n;
   }
   return fbs->n_bundles * max_layer + 1;
}

// z-march iteration count of simulator.cpp:259-261: every ray starts at
// z = 0 + 0.5 * step_z and stops once floor(z) leaves [0, nz).
int64_t fpl_count_voxel_steps(const fpl_setup *setup, const int64_t dim[3], double theta,
                              double phi) {
   (void)phi;
   double c = std::cos(theta);
   double st;
   {
      // vm::rot_pi_cases::cos (vemath.hpp:510-520)
      double v = std::fmod(theta, 2 * M_PI);
      if (v == 0)
         st = 1;
      else if (v == M_PI_2 || v == M_PI + M_PI_2 || v == -M_PI_2 || v == -(M_PI + M_PI_2))
         st = 0;
      else if (v == M_PI || v == -M_PI)
         st = -1;
      else
         st = c;
   }
   const double step_z = st * setup->step_size;
   // the bound fpl_simulate puts on one ray's z-steps (-1: the tilt is rejected)
   if (!((double)dim[2] / std::fabs(step_z) <= (double)((int64_t)1 << 24))) return -1;
   double z = 0.0 + step_z * 0.5;
   int64_t n = 0;
   for (; std::floor(z) >= 0 && std::floor(z) < (double)dim[2]; z = z + step_z) {
      n++;
      if (n > (int64_t)1 << 40) break;
   }
   return n * dim[0] * dim[1] * (int64_t)setup->n_rho;
}

} // extern "C"
