// Serving loop: many run_simulation requests on one device at once.
//
// A request is the march of one registered tissue under its own setup, tilt
// list and property table.  The tissue's fields are packed once at
// registration (the same pack kernel as fpl_simulate) and stay resident.
//
// A worker thread owns the device side.  Pending requests are admitted in
// arrival order into free stream slots while the HBM budget allows (resident
// tissues + footprints of the requests in flight + the new one); each admitted
// request runs on its slot's stream as a replay of a captured hipGraph
//     H2D parameter block -> flag reset kernel -> march kernel -> D2H warnings
// whose buffers are the slot's (so one graph per slot and request shape is
// captured once and replayed), then a device-to-device copy of the slot's
// images into the request's output buffer and an event.  The worker polls the
// events of busy slots, retires finished requests and admits new ones as slots
// and memory free up, so requests that arrive while others run overlap with
// them on the device.
//
// The graphs hold no memset nodes: flags are cleared by a kernel.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

#include "fpl_internal.h"
#include "fpl_sim.h"

struct fpl_server; // public opaque handle (defined below)

namespace fpl {
namespace serve {

using Clock = std::chrono::steady_clock;

struct Tissue {
   int64_t dim[3] = {0, 0, 0};
   int storage = 0;
   void *packed = nullptr;
   uint32_t *packed_b = nullptr;
   int32_t lmin = 0, lmax = 0;
   int64_t bytes = 0;
   int64_t users = 0; // requests pending or in flight
};

struct Req {
   int64_t ticket = 0, tissue = 0;
   sim::Block block;
   double *out = nullptr;
   size_t out_bytes = 0;
   int64_t footprint = 0;
   Clock::time_point t0;
   bool done = false;
   Status status;
   uint32_t warn = 0;
   double latency = 0;
};

using GraphKey = std::tuple<int64_t, int64_t, int64_t, int>; // tissue, tilts, props, rho

struct Slot {
   hipStream_t stream = nullptr;
   hipEvent_t ev = nullptr;
   void *d_block = nullptr, *h_block = nullptr;
   size_t block_cap = 0;
   double *d_images = nullptr;
   size_t img_cap = 0;
   uint32_t *d_flags = nullptr, *h_flags = nullptr;
   std::map<GraphKey, hipGraphExec_t> graphs;
   std::shared_ptr<Req> req;
};

} // namespace serve
} // namespace fpl

struct fpl_server {
   int device = 0;
   int64_t budget = 0, resident = 0, inflight = 0;
   hipStream_t reg_stream = nullptr;
   void *reg_flags = nullptr, *reg_host = nullptr;
   std::vector<fpl::serve::Slot> slots;
   std::map<int64_t, fpl::serve::Tissue> tissues;
   std::deque<std::shared_ptr<fpl::serve::Req>> pending;
   std::map<int64_t, std::shared_ptr<fpl::serve::Req>> reqs;
   std::mutex mu;
   std::condition_variable cv_work, cv_done;
   std::thread worker;
   bool stop = false;
   int64_t next_ticket = 1, next_tissue = 1;
   int64_t n_submitted = 0, n_done = 0, n_graphs = 0, n_replays = 0;
};

namespace fpl {
namespace serve {

struct DeviceGuard { // makes `dev` current for a scope
   int prev = -1;
   explicit DeviceGuard(int dev) {
      if (hipGetDevice(&prev) != hipSuccess) prev = -1;
      (void)hipSetDevice(dev);
   }
   ~DeviceGuard() {
      if (prev >= 0) (void)hipSetDevice(prev);
   }
};

Status grow(Slot &s, size_t block_bytes, size_t img_bytes) {
   if (block_bytes > s.block_cap) {
      for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.second);
      s.graphs.clear();
      if (s.d_block) (void)hipFree(s.d_block);
      if (s.h_block) (void)hipHostFree(s.h_block);
      s.d_block = s.h_block = nullptr;
      s.block_cap = 0;
      const size_t cap = block_bytes * 2;
      if (hipMalloc(&s.d_block, cap) != hipSuccess)
         return make_status(FPL_E_NOMEM, "serving: slot parameter buffer");
      if (hipHostMalloc(&s.h_block, cap, hipHostMallocDefault) != hipSuccess)
         return make_status(FPL_E_NOMEM, "serving: pinned parameter buffer");
      s.block_cap = cap;
   }
   if (img_bytes > s.img_cap) {
      for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.second);
      s.graphs.clear();
      if (s.d_images) (void)hipFree(s.d_images);
      s.d_images = nullptr;
      s.img_cap = 0;
      if (hipMalloc(&s.d_images, img_bytes) != hipSuccess)
         return make_status(FPL_E_NOMEM, "serving: slot image buffer");
      s.img_cap = img_bytes;
   }
   return Status();
}

// Captures the request shape's graph on the slot's stream.
Status capture(fpl_server *S, Slot &s, const Tissue &T, const Req &r, hipGraphExec_t *out) {
   hipGraph_t g = nullptr;
   FPL_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed));
   hipMemcpyAsync(s.d_block, s.h_block, r.block.bytes.size(), hipMemcpyHostToDevice, s.stream);
   sim::clear_flags(s.d_flags, (int32_t *)(s.d_flags + 4), s.stream);
   sim::march(r.block, s.d_block, T.packed, T.packed_b, s.d_images, s.d_flags, s.stream);
   hipMemcpyAsync(s.h_flags, s.d_flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream);
   const hipError_t e = hipStreamEndCapture(s.stream, &g);
   if (e != hipSuccess || !g)
      return make_status(FPL_E_DEVICE, "serving: graph capture failed: %s", hipGetErrorString(e));
   const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
   (void)hipGraphDestroy(g);
   if (ei != hipSuccess)
      return make_status(FPL_E_DEVICE, "serving: graph instantiation failed: %s",
                         hipGetErrorString(ei));
   S->n_graphs++;
   return Status();
}

Status launch(fpl_server *S, Slot &s, Req &r) {
   const Tissue &T = S->tissues.at(r.tissue);
   FPL_TRY(grow(s, r.block.bytes.size(), r.out_bytes));
   memcpy(s.h_block, r.block.bytes.data(), r.block.bytes.size());
   const GraphKey key{r.tissue, r.block.n_tilts, r.block.n_props, r.block.n_rho};
   auto it = s.graphs.find(key);
   if (it == s.graphs.end()) {
      hipGraphExec_t ge = nullptr;
      FPL_TRY(capture(S, s, T, r, &ge));
      it = s.graphs.emplace(key, ge).first;
   }
   FPL_HIP(hipGraphLaunch(it->second, s.stream));
   S->n_replays++;
   FPL_HIP(hipMemcpyAsync(r.out, s.d_images, r.out_bytes, hipMemcpyDeviceToDevice, s.stream));
   FPL_HIP(hipEventRecord(s.ev, s.stream));
   return Status();
}

void retire(fpl_server *S, const std::shared_ptr<Req> &r) {
   r->latency = std::chrono::duration<double>(Clock::now() - r->t0).count();
   r->done = true;
   auto t = S->tissues.find(r->tissue);
   if (t != S->tissues.end()) t->second.users--;
   S->n_done++;
   S->cv_done.notify_all();
}

void loop(fpl_server *S) {
   (void)hipSetDevice(S->device);
   std::unique_lock<std::mutex> lk(S->mu);
   for (;;) {
      // admit in arrival order
      for (auto &s : S->slots) {
         if (s.req || S->pending.empty()) continue;
         std::shared_ptr<Req> r = S->pending.front();
         if (S->resident + S->inflight + r->footprint > S->budget) {
            bool busy = false;
            for (auto &o : S->slots) busy = busy || (bool)o.req;
            if (busy) break; // wait for memory to free up
            S->pending.pop_front();
            r->status = make_status(FPL_E_NOMEM, "serving: request needs %lld bytes, budget %lld "
                                    "with %lld resident", (long long)r->footprint,
                                    (long long)S->budget, (long long)S->resident);
            retire(S, r);
            continue;
         }
         S->pending.pop_front();
         Status st = launch(S, s, *r);
         if (!st.ok()) {
            r->status = st;
            retire(S, r);
            continue;
         }
         S->inflight += r->footprint;
         s.req = r;
      }
      // retire finished requests
      bool busy = false;
      for (auto &s : S->slots) {
         if (!s.req) continue;
         const hipError_t q = hipEventQuery(s.ev);
         if (q == hipErrorNotReady) {
            busy = true;
            continue;
         }
         std::shared_ptr<Req> r = s.req;
         s.req.reset();
         S->inflight -= r->footprint;
         if (q != hipSuccess)
            r->status = make_status(FPL_E_DEVICE, "serving: request failed: %s",
                                    hipGetErrorString(q));
         else
            r->warn = s.h_flags[0] & 0x7fffffffu;
         retire(S, r);
      }
      if (S->stop && !busy) break;
      if (busy) {
         lk.unlock();
         std::this_thread::sleep_for(std::chrono::microseconds(10));
         lk.lock();
      } else if (S->pending.empty() && !S->stop) {
         S->cv_work.wait(lk);
      }
   }
   // fail what is still queued
   while (!S->pending.empty()) {
      std::shared_ptr<Req> r = S->pending.front();
      S->pending.pop_front();
      r->status = make_status(FPL_E_INVALID, "serving: server destroyed before the request ran");
      retire(S, r);
   }
}

Status create(int device, int64_t budget, int32_t max_inflight, fpl_server **out) {
   if (!out) return make_status(FPL_E_INVALID, "null argument");
   *out = nullptr;
   if (max_inflight < 1 || max_inflight > 16)
      return make_status(FPL_E_INVALID, "max_inflight must be in [1, 16]");
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
      return make_status(FPL_E_DEVICE, "no HIP device %d", device);
   DeviceGuard dg(device);
   std::unique_ptr<fpl_server> S(new fpl_server());
   S->device = device;
   if (budget <= 0) {
      size_t fr = 0, tot = 0;
      FPL_HIP(hipMemGetInfo(&fr, &tot));
      budget = (int64_t)(0.9 * (double)fr);
   }
   S->budget = budget;
   FPL_HIP(hipStreamCreateWithFlags(&S->reg_stream, hipStreamNonBlocking));
   FPL_HIP(hipMalloc(&S->reg_flags, 64));
   FPL_HIP(hipHostMalloc(&S->reg_host, 64, hipHostMallocDefault));
   S->slots.resize(max_inflight);
   for (auto &s : S->slots) {
      FPL_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
      FPL_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
      FPL_HIP(hipMalloc(&s.d_flags, 64));
      FPL_HIP(hipHostMalloc(&s.h_flags, 64, hipHostMallocDefault));
   }
   fpl_server *raw = S.release();
   raw->worker = std::thread(loop, raw);
   *out = raw;
   return Status();
}

Status add_tissue(fpl_server *S, const int64_t dim[3], const int32_t *label, const float *vector,
                  int32_t storage, int64_t *id) {
   if (!S || !dim || !label || !vector || !id) return make_status(FPL_E_INVALID, "null argument");
   if (storage < 0 || storage > 2)
      return make_status(FPL_E_INVALID, "field_storage must be 0 (f32), 1 (f16) or 2 (angle16)");
   if (dim[0] <= 0 || dim[1] <= 0 || dim[2] <= 0 || dim[0] >= (1 << 30) || dim[1] >= (1 << 30) ||
       dim[2] >= (1 << 30))
      return make_status(FPL_E_INVALID, "bad tissue dimensions");
   DeviceGuard dg(S->device);
   Tissue T;
   for (int k = 0; k < 3; k++) T.dim[k] = dim[k];
   T.storage = storage;
   const size_t n_vox = (size_t)(dim[0] * dim[1] * dim[2]);
   T.bytes = (int64_t)(sim::record_bytes(storage) * n_vox + (storage == 2 ? 4 * n_vox : 0));
   {
      std::lock_guard<std::mutex> g(S->mu);
      if (S->resident + S->inflight + T.bytes > S->budget)
         return make_status(FPL_E_NOMEM, "serving: tissue of %lld bytes exceeds the budget",
                            (long long)T.bytes);
      S->resident += T.bytes; // reserved
   }
   Status st;
   do {
      if (hipMalloc(&T.packed, sim::record_bytes(storage) * n_vox) != hipSuccess) {
         st = make_status(FPL_E_NOMEM, "serving: tissue fields");
         break;
      }
      if (storage == 2 && hipMalloc((void **)&T.packed_b, 4 * n_vox) != hipSuccess) {
         st = make_status(FPL_E_NOMEM, "serving: tissue angle codes");
         break;
      }
      uint32_t *fl = (uint32_t *)S->reg_flags;
      int32_t *mm = (int32_t *)((char *)S->reg_flags + 16);
      sim::clear_flags(fl, mm, S->reg_stream);
      sim::pack(storage, dim, label, vector, T.packed, T.packed_b, mm, S->reg_stream);
      hipError_t e = hipMemcpyAsync(S->reg_host, mm, 8, hipMemcpyDeviceToHost, S->reg_stream);
      if (e == hipSuccess) e = hipStreamSynchronize(S->reg_stream);
      if (e != hipSuccess) {
         st = make_status(FPL_E_DEVICE, "serving: packing failed: %s", hipGetErrorString(e));
         break;
      }
      T.lmin = ((int32_t *)S->reg_host)[0];
      T.lmax = ((int32_t *)S->reg_host)[1];
      // label < 0 and the storage's label width are properties of the tissue;
      // the property-table bound is checked per request
      st = sim::check_labels(T.lmin, T.lmax, (int64_t)T.lmax + 1, storage);
   } while (false);
   std::lock_guard<std::mutex> g(S->mu);
   if (!st.ok()) {
      if (T.packed) (void)hipFree(T.packed);
      if (T.packed_b) (void)hipFree(T.packed_b);
      S->resident -= T.bytes;
      return st;
   }
   *id = S->next_tissue++;
   S->tissues[*id] = T;
   return Status();
}

Status remove_tissue(fpl_server *S, int64_t id) {
   if (!S) return make_status(FPL_E_INVALID, "null argument");
   DeviceGuard dg(S->device);
   std::lock_guard<std::mutex> g(S->mu);
   auto it = S->tissues.find(id);
   if (it == S->tissues.end()) return make_status(FPL_E_INVALID, "unknown tissue %lld", (long long)id);
   if (it->second.users > 0)
      return make_status(FPL_E_INVALID, "tissue %lld is used by queued requests", (long long)id);
   for (auto &s : S->slots)
      for (auto gi = s.graphs.begin(); gi != s.graphs.end();) {
         if (std::get<0>(gi->first) == id) {
            (void)hipGraphExecDestroy(gi->second);
            gi = s.graphs.erase(gi);
         } else {
            ++gi;
         }
      }
   (void)hipFree(it->second.packed);
   if (it->second.packed_b) (void)hipFree(it->second.packed_b);
   S->resident -= it->second.bytes;
   S->tissues.erase(it);
   return Status();
}

Status submit(fpl_server *S, const fpl_request *q, int64_t *ticket) {
   if (!S || !q || !ticket || !q->setup) return make_status(FPL_E_INVALID, "null argument");
   if (!q->images) return make_status(FPL_E_INVALID, "null array");
   auto r = std::make_shared<Req>();
   int64_t dim[3];
   int storage;
   int32_t lmin, lmax;
   {
      std::lock_guard<std::mutex> g(S->mu);
      auto it = S->tissues.find(q->tissue);
      if (it == S->tissues.end())
         return make_status(FPL_E_INVALID, "unknown tissue %lld", (long long)q->tissue);
      for (int k = 0; k < 3; k++) dim[k] = it->second.dim[k];
      storage = it->second.storage;
      lmin = it->second.lmin;
      lmax = it->second.lmax;
   }
   fpl_setup setup = *q->setup;
   setup.field_storage = storage;
   FPL_TRY(sim::prepare(&setup, dim, q->props, q->n_props, q->tilts, q->n_tilts, r->block));
   FPL_TRY(sim::check_labels(lmin, lmax, q->n_props, storage));
   r->tissue = q->tissue;
   r->out = q->images;
   r->out_bytes = sizeof(double) * (size_t)(q->n_tilts * dim[0] * dim[1] * setup.n_rho);
   r->footprint = (int64_t)(r->out_bytes + r->block.bytes.size());
   r->t0 = Clock::now();
   std::lock_guard<std::mutex> g(S->mu);
   if (S->stop) return make_status(FPL_E_INVALID, "serving: server is shutting down");
   auto it = S->tissues.find(q->tissue);
   if (it == S->tissues.end())
      return make_status(FPL_E_INVALID, "unknown tissue %lld", (long long)q->tissue);
   it->second.users++;
   r->ticket = S->next_ticket++;
   S->reqs[r->ticket] = r;
   S->pending.push_back(r);
   S->n_submitted++;
   *ticket = r->ticket;
   S->cv_work.notify_one();
   return Status();
}

Status wait(fpl_server *S, int64_t ticket, double timeout_s, uint32_t *warn, double *latency) {
   if (!S) return make_status(FPL_E_INVALID, "null argument");
   std::unique_lock<std::mutex> lk(S->mu);
   auto it = S->reqs.find(ticket);
   if (it == S->reqs.end())
      return make_status(FPL_E_INVALID, "unknown ticket %lld", (long long)ticket);
   std::shared_ptr<Req> r = it->second;
   if (timeout_s < 0) {
      S->cv_done.wait(lk, [&] { return r->done; });
   } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                   [&] { return r->done; })) {
      return make_status(FPL_E_PENDING, "request %lld is not finished", (long long)ticket);
   }
   S->reqs.erase(ticket);
   if (warn) *warn = r->warn;
   if (latency) *latency = r->latency;
   return r->status;
}

Status destroy(fpl_server *S) {
   if (!S) return Status();
   {
      std::lock_guard<std::mutex> g(S->mu);
      S->stop = true;
      S->cv_work.notify_all();
   }
   if (S->worker.joinable()) S->worker.join();
   DeviceGuard dg(S->device);
   for (auto &s : S->slots) {
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.second);
      if (s.d_block) (void)hipFree(s.d_block);
      if (s.h_block) (void)hipHostFree(s.h_block);
      if (s.d_images) (void)hipFree(s.d_images);
      if (s.d_flags) (void)hipFree(s.d_flags);
      if (s.h_flags) (void)hipHostFree(s.h_flags);
      if (s.ev) (void)hipEventDestroy(s.ev);
      if (s.stream) (void)hipStreamDestroy(s.stream);
   }
   for (auto &t : S->tissues) {
      (void)hipFree(t.second.packed);
      if (t.second.packed_b) (void)hipFree(t.second.packed_b);
   }
   if (S->reg_stream) (void)hipStreamDestroy(S->reg_stream);
   if (S->reg_flags) (void)hipFree(S->reg_flags);
   if (S->reg_host) (void)hipHostFree(S->reg_host);
   delete S;
   return Status();
}

} // namespace serve
} // namespace fpl

extern "C" int fpl_server_create(int device, int64_t hbm_budget, int32_t max_inflight,
                                 fpl_server **server, char *msg, size_t msg_len) {
   return fpl::finish(fpl::serve::create(device, hbm_budget, max_inflight, server), msg, msg_len);
}

extern "C" int fpl_server_add_tissue(fpl_server *server, const int64_t dim[3],
                                     const int32_t *label, const float *vector,
                                     int32_t field_storage, int64_t *tissue, char *msg,
                                     size_t msg_len) {
   return fpl::finish(fpl::serve::add_tissue(server, dim, label, vector, field_storage, tissue),
                      msg, msg_len);
}

extern "C" int fpl_server_remove_tissue(fpl_server *server, int64_t tissue, char *msg,
                                        size_t msg_len) {
   return fpl::finish(fpl::serve::remove_tissue(server, tissue), msg, msg_len);
}

extern "C" int fpl_server_submit(fpl_server *server, const fpl_request *request, int64_t *ticket,
                                 char *msg, size_t msg_len) {
   return fpl::finish(fpl::serve::submit(server, request, ticket), msg, msg_len);
}

extern "C" int fpl_server_wait(fpl_server *server, int64_t ticket, double timeout_s,
                               uint32_t *warn, double *latency_s, char *msg, size_t msg_len) {
   return fpl::finish(fpl::serve::wait(server, ticket, timeout_s, warn, latency_s), msg, msg_len);
}

extern "C" int fpl_server_stats(fpl_server *server, int64_t stats[6]) {
   if (!server || !stats) return FPL_E_INVALID;
   std::lock_guard<std::mutex> g(server->mu);
   int64_t busy = 0;
   for (auto &s : server->slots) busy += s.req ? 1 : 0;
   stats[0] = server->n_submitted;
   stats[1] = server->n_done;
   stats[2] = busy;
   stats[3] = (int64_t)server->pending.size();
   stats[4] = server->n_graphs;
   stats[5] = server->n_replays;
   return FPL_OK;
}

extern "C" int fpl_server_destroy(fpl_server *server) {
   return fpl::serve::destroy(server).code;
}
