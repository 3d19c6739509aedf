// Serving loop: many simulation requests on one device at once.
//
// Two kinds of request:
//   * march of a registered tissue (its fields packed once at registration and
//     resident) under the request's own setup, property table and tilt list;
//   * a fiber model: generation of the request's own tissue (straight into
//     march records in the slot's buffer, fpl_generate_fields) and then its
//     march, i.e. Simpli.generate_tissue + run_simulation of every tilt.
//
// The server owns `max_inflight` slots.  Each slot has a stream, device
// buffers that grow to the largest request it ran, and a host thread that
// runs one request at a time to completion: generation (which reads a few
// words back mid-call) and then the march as a replay of a hipGraph captured
// once per (slot, request shape),
//     H2D parameter block -> flag reset kernel -> march kernel -> D2H warnings,
// a device-to-device copy of the slot's images into the request's buffer and a
// stream synchronisation.  Requests of different slots overlap on the device.
//
// Admission is in arrival order: a free slot takes the oldest pending request
// when the tissues, the device memory all slots hold, and what this slot must
// grow by to run the request fit the HBM budget.  Otherwise it waits for a busy
// slot to finish (one that already holds enough memory may then take the
// request); when every slot is idle and the request still does not fit, idle
// slots release their buffers, and a request larger than the whole budget
// fails with FPL_E_NOMEM.
//
// A request starts after the work its caller had queued on the submitting
// stream (an event recorded at submit), so its output buffer may still be in
// use there.  The graphs hold no memset nodes: flags are cleared by a kernel.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "fpl_gen.h"
#include "fpl_internal.h"
#include "fpl_records.h"
#include "fpl_sim.h"

struct fpl_server; // public opaque handle (defined below)

namespace fpl {
namespace serve {

using Clock = std::chrono::steady_clock;

struct Tissue {
   fpl_fields F{};
   void *mem = nullptr;
   int64_t bytes = 0;
   int64_t users = 0; // requests pending or in flight
};

// device buffers a slot must hold to run a request
enum Buf { B_BLOCK = 0, B_IMAGES, B_POINTS, B_FPTR, B_FIELDS, B_NUM };

struct Req {
   int64_t ticket = 0;
   bool model = false;
   // tissue request
   int64_t tissue = 0;
   sim::Block block;
   fpl_fields tf{}; // the tissue's fields, read at admission
   // model request: the fiber model (points and fiber offsets borrowed from
   // the caller until the request is done, the tables copied) and the volume
   fpl_volume vol{};
   const double *points = nullptr;
   const int64_t *fiber_ptr = nullptr;
   int64_t n_points = 0, n_fibers = 0;
   std::vector<double> lscale, ldn, lmu;
   std::vector<int64_t> bundle_ptr, layer_ptr;
   std::vector<uint8_t> lmodel;
   int64_t n_props = 0;
   double bg_mu = 0.0;
   fpl_setup setup{};
   std::vector<double> rho, tilts;
   // both
   int64_t dim[3] = {0, 0, 0};
   int storage = 0;
   int64_t n_tilts = 0;
   double *out = nullptr;
   size_t need[B_NUM] = {};
   size_t scratch = 0; // estimated generation workspace
   hipEvent_t after = nullptr;
   Clock::time_point t0;
   bool done = false;
   Status status;
   uint32_t warn = 0;
   double latency = 0;
};

// (model?, tissue, nx, ny, nz, storage, tilts, property table in LDS?): what
// the captured launches depend on.  The parameter upload copies the slot's
// whole block capacity, so property-table size and rotation count (values in
// the block) need no graphs of their own.
using GraphKey = std::tuple<int, int64_t, int64_t, int64_t, int64_t, int, int64_t, bool>;

struct Slot {
   int index = 0;
   hipStream_t stream = nullptr;
   void *buf[B_NUM] = {};
   size_t cap[B_NUM] = {};
   void *h_block = nullptr;
   size_t h_block_cap = 0;
   uint32_t *d_flags = nullptr, *h_flags = nullptr;
   std::map<GraphKey, hipGraphExec_t> graphs;
   std::shared_ptr<Req> req;
   int64_t held = 0; // device bytes: buffers + the stream's generation workspace
   std::thread th;
};

} // namespace serve
} // namespace fpl

struct fpl_server {
   int device = 0;
   int64_t budget = 0, resident = 0, held = 0;
   hipStream_t reg_stream = nullptr;
   std::vector<std::unique_ptr<fpl::serve::Slot>> slots;
   std::map<int64_t, fpl::serve::Tissue> tissues;
   std::deque<std::shared_ptr<fpl::serve::Req>> pending;
   std::map<int64_t, std::shared_ptr<fpl::serve::Req>> reqs;
   std::mutex mu;
   std::condition_variable cv_work, cv_done;
   bool stop = false;
   int64_t next_ticket = 1, next_tissue = 1;
   int64_t n_submitted = 0, n_done = 0, n_graphs = 0, n_replays = 0, n_generated = 0;
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

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Bytes the slot grows by to run r (buffers below the request's needs, and
// the generation workspace below its estimate).
int64_t growth(const Slot &s, const Req &r) {
   int64_t g = 0;
   for (int k = 0; k < B_NUM; k++)
      if (r.need[k] > s.cap[k]) g += (int64_t)(r.need[k] - s.cap[k]);
   const size_t ws = ws_bytes(s.stream);
   if (r.scratch > ws) g += (int64_t)(r.scratch - ws);
   return g;
}

int64_t slot_bytes(const Slot &s) {
   int64_t n = (int64_t)ws_bytes(s.stream);
   for (int k = 0; k < B_NUM; k++) n += (int64_t)s.cap[k];
   return n;
}

void drop_graphs(fpl_server *S, Slot &s) {
   (void)S;
   for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.second);
   s.graphs.clear();
}

// Frees an idle slot's buffers and workspace (caller holds the lock).
void shrink(fpl_server *S, Slot &s) {
   (void)hipStreamSynchronize(s.stream);
   drop_graphs(S, s);
   for (int k = 0; k < B_NUM; k++) {
      if (s.buf[k]) (void)hipFree(s.buf[k]);
      s.buf[k] = nullptr;
      s.cap[k] = 0;
   }
   ws_free(s.stream);
   S->held -= s.held;
   s.held = 0;
}

// Grows the slot's buffers to the request's needs (the stream is idle).
Status grow(fpl_server *S, Slot &s, const Req &r) {
   for (int k = 0; k < B_NUM; k++) {
      if (r.need[k] <= s.cap[k]) continue;
      drop_graphs(S, s); // they hold the old pointers
      if (s.buf[k]) (void)hipFree(s.buf[k]);
      s.buf[k] = nullptr;
      s.cap[k] = 0;
      if (hipMalloc(&s.buf[k], r.need[k]) != hipSuccess) {
         (void)hipGetLastError();
         s.buf[k] = nullptr;
         return make_status(FPL_E_NOMEM, "serving: slot buffer of %zu bytes", r.need[k]);
      }
      s.cap[k] = r.need[k];
   }
   if (s.cap[B_BLOCK] > s.h_block_cap) { // the graphs upload the whole capacity
      if (s.h_block) (void)hipHostFree(s.h_block);
      s.h_block = nullptr;
      s.h_block_cap = 0;
      if (hipHostMalloc(&s.h_block, s.cap[B_BLOCK], hipHostMallocDefault) != hipSuccess)
         return make_status(FPL_E_NOMEM, "serving: pinned parameter buffer");
      s.h_block_cap = s.cap[B_BLOCK];
   }
   return Status();
}

// Captures the march of a request shape on the slot's stream.
Status capture(fpl_server *S, Slot &s, const sim::Block &B, const void *rec_a,
               const uint32_t *rec_b, hipGraphExec_t *out) {
   hipGraph_t g = nullptr;
   FPL_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed));
   hipMemcpyAsync(s.buf[B_BLOCK], s.h_block, s.cap[B_BLOCK], hipMemcpyHostToDevice, s.stream);
   sim::clear_flags(s.d_flags, (int32_t *)(s.d_flags + 4), s.stream);
   sim::march(B, s.buf[B_BLOCK], rec_a, rec_b, (double *)s.buf[B_IMAGES], s.d_flags, s.stream);
   hipMemcpyAsync(s.h_flags, s.d_flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream);
   const hipError_t e = hipStreamEndCapture(s.stream, &g);
   if (e != hipSuccess || !g)
      return make_status(FPL_E_DEVICE, "serving: graph capture failed: %s", hipGetErrorString(e));
   const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
   (void)hipGraphDestroy(g);
   if (ei != hipSuccess)
      return make_status(FPL_E_DEVICE, "serving: graph instantiation failed: %s",
                         hipGetErrorString(ei));
   std::lock_guard<std::mutex> lk(S->mu);
   S->n_graphs++;
   return Status();
}

// The march of B over the given fields as a graph replay, then the copy of
// the images into the request's buffer; returns the march's flag word.
Status march(fpl_server *S, Slot &s, Req &r, const sim::Block &B, const void *rec_a,
             const uint32_t *rec_b, uint32_t *flags) {
   memcpy(s.h_block, B.bytes.data(), B.bytes.size());
   const GraphKey key{r.model ? 1 : 0,   r.model ? 0 : r.tissue, r.dim[0],
                      r.dim[1],          r.dim[2],               r.storage,
                      B.n_tilts,         sim::props_in_lds(B)};
   auto it = s.graphs.find(key);
   if (it == s.graphs.end()) {
      hipGraphExec_t ge = nullptr;
      FPL_TRY(capture(S, s, B, rec_a, rec_b, &ge));
      it = s.graphs.emplace(key, ge).first;
   }
   FPL_HIP(hipGraphLaunch(it->second, s.stream));
   {
      std::lock_guard<std::mutex> lk(S->mu);
      S->n_replays++;
   }
   const size_t img = sizeof(double) * (size_t)(r.n_tilts * r.dim[0] * r.dim[1] * B.n_rho);
   FPL_HIP(hipMemcpyAsync(r.out, s.buf[B_IMAGES], img, hipMemcpyDeviceToDevice, s.stream));
   FPL_HIP(hipStreamSynchronize(s.stream));
   *flags = s.h_flags[0];
   return Status();
}

// One request on the slot, synchronously.
Status run_request(fpl_server *S, Slot &s, Req &r) {
   FPL_TRY(grow(S, s, r));
   if (r.after) FPL_HIP(hipStreamWaitEvent(s.stream, r.after, 0));
   uint32_t f = 0;
   if (!r.model) {
      FPL_TRY(march(S, s, r, r.block, r.tf.records, r.tf.codes, &f));
   } else {
      const int64_t np = r.n_points, nf = r.n_fibers;
      if (np > 0)
         FPL_HIP(hipMemcpyAsync(s.buf[B_POINTS], r.points, sizeof(double) * 4 * (size_t)np,
                                hipMemcpyHostToDevice, s.stream));
      FPL_HIP(hipMemcpyAsync(s.buf[B_FPTR], r.fiber_ptr, sizeof(int64_t) * (size_t)(nf + 1),
                             hipMemcpyHostToDevice, s.stream));
      fpl_fiber_bundles fb;
      fb.n_bundles = (int64_t)r.bundle_ptr.size() - 1;
      fb.n_fibers = nf;
      fb.n_points = np;
      fb.points = (const double *)s.buf[B_POINTS];
      fb.fiber_ptr = (const int64_t *)s.buf[B_FPTR];
      fb.bundle_ptr = r.bundle_ptr.data();
      fb.layer_ptr = r.layer_ptr.data();
      fb.layer_scale = r.lscale.data();
      fb.layer_dn = r.ldn.data();
      fb.layer_mu = r.lmu.data();
      fb.layer_model = r.lmodel.data();
      const size_t off = align256((size_t)(r.dim[0] * r.dim[1] * r.dim[2]) *
                                  rec::record_bytes(r.storage));
      void *rec_a = s.buf[B_FIELDS];
      uint32_t *rec_b =
          r.storage == rec::ST_ANGLE16 ? (uint32_t *)((char *)s.buf[B_FIELDS] + off) : nullptr;
      const gen::Outputs out{nullptr, nullptr, rec_a, rec_b, r.storage};
      std::vector<double> props(2 * (size_t)r.n_props, 0.0);
      uint32_t gw = 0;
      int32_t lmax = 0;
      FPL_TRY(gen::run(&r.vol, &fb, out, props.data(), &gw, &lmax, s.stream));
      props[1] = r.bg_mu; // Simpli.generate_tissue sets the background absorption
      {
         std::lock_guard<std::mutex> lk(S->mu);
         S->n_generated++;
      }
      sim::Block B;
      r.setup.filter_rotations = r.rho.data();
      r.setup.field_storage = r.storage;
      FPL_TRY(sim::prepare(&r.setup, r.dim, props.data(), r.n_props, r.tilts.data(), r.n_tilts,
                           B));
      FPL_TRY(sim::check_labels(0, lmax, r.n_props, r.storage));
      FPL_TRY(march(S, s, r, B, rec_a, rec_b, &f));
      f |= gw;
   }
   if (f & 0x80000000u) return make_status(FPL_E_LABEL_MAX, "label exceed properties.size()");
   r.warn = f & 0x7fffffffu;
   return Status();
}

void retire(fpl_server *S, const std::shared_ptr<Req> &r) { // caller holds the lock
   r->latency = std::chrono::duration<double>(Clock::now() - r->t0).count();
   r->done = true;
   if (r->after) {
      (void)hipEventDestroy(r->after);
      r->after = nullptr;
   }
   if (!r->model) {
      auto t = S->tissues.find(r->tissue);
      if (t != S->tissues.end()) t->second.users--;
   }
   S->n_done++;
   S->cv_done.notify_all();
}

// Admission decision of free slot `me` for the oldest pending request (caller
// holds the lock): 1 take it, 0 wait, -1 it can never run (rejected).
int admit(fpl_server *S, Slot &me, const Req &r) {
   const int64_t base = S->resident + S->held;
   if (base + growth(me, r) <= S->budget) return 1;
   bool busy = false;
   for (auto &o : S->slots) {
      if (o.get() == &me) continue;
      if (o->req) {
         busy = true;
      } else if (base + growth(*o, r) <= S->budget) {
         return 0; // another free slot can take it without exceeding the budget
      }
   }
   if (busy) return 0; // memory is held by running requests' slots
   // every slot idle and none fits: release the idle slots' memory
   for (auto &o : S->slots) shrink(S, *o);
   return S->resident + growth(me, r) <= S->budget ? 1 : -1;
}

void slot_loop(fpl_server *S, Slot *me) {
   (void)hipSetDevice(S->device);
   std::unique_lock<std::mutex> lk(S->mu);
   for (;;) {
      if (S->pending.empty()) {
         if (S->stop) break;
         S->cv_work.wait(lk);
         continue;
      }
      std::shared_ptr<Req> r = S->pending.front();
      const int a = admit(S, *me, *r);
      if (a == 0) {
         S->cv_work.wait(lk);
         continue;
      }
      S->pending.pop_front();
      if (a < 0) {
         int64_t need = (int64_t)r->scratch;
         for (int k = 0; k < B_NUM; k++) need += (int64_t)r->need[k];
         r->status = make_status(FPL_E_NOMEM, "serving: request needs %lld bytes, budget %lld "
                                 "with %lld resident", (long long)need, (long long)S->budget,
                                 (long long)S->resident);
         retire(S, r);
         S->cv_work.notify_all();
         continue;
      }
      const int64_t g = growth(*me, *r);
      S->held += g; // reserved until the slot's real size is known
      me->held += g;
      me->req = r;
      if (!r->model) r->tf = S->tissues.at(r->tissue).F; // in use: not removed meanwhile
      // graphs of removed tissues (only this thread touches its slot's graphs)
      for (auto gi = me->graphs.begin(); gi != me->graphs.end();) {
         if (std::get<0>(gi->first) == 0 && !S->tissues.count(std::get<1>(gi->first))) {
            (void)hipGraphExecDestroy(gi->second);
            gi = me->graphs.erase(gi);
         } else {
            ++gi;
         }
      }
      lk.unlock();
      Status st = run_request(S, *me, *r);
      if (!st.ok()) (void)hipStreamSynchronize(me->stream); // nothing of it may still run
      lk.lock();
      const int64_t real = slot_bytes(*me);
      S->held += real - me->held;
      me->held = real;
      me->req.reset();
      r->status = st;
      retire(S, r);
      S->cv_work.notify_all(); // a waiting slot may take the next request now
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
   for (int k = 0; k < max_inflight; k++) {
      std::unique_ptr<Slot> s(new Slot());
      s->index = k;
      FPL_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      FPL_HIP(hipMalloc(&s->d_flags, 64));
      FPL_HIP(hipHostMalloc(&s->h_flags, 64, hipHostMallocDefault));
      S->slots.push_back(std::move(s));
   }
   fpl_server *raw = S.release();
   for (auto &s : raw->slots) s->th = std::thread(slot_loop, raw, s.get());
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
   T.F.storage = storage;
   const size_t n_vox = (size_t)(dim[0] * dim[1] * dim[2]);
   const size_t off = align256(rec::record_bytes(storage) * n_vox);
   T.bytes = (int64_t)(off + rec::second_bytes(storage) * n_vox);
   {
      std::lock_guard<std::mutex> g(S->mu);
      if (S->resident + S->held + T.bytes > S->budget)
         return make_status(FPL_E_NOMEM, "serving: tissue of %lld bytes exceeds the budget",
                            (long long)T.bytes);
      S->resident += T.bytes; // reserved
   }
   Status st;
   do {
      if (hipMalloc(&T.mem, (size_t)T.bytes) != hipSuccess) {
         (void)hipGetLastError();
         T.mem = nullptr;
         st = make_status(FPL_E_NOMEM, "serving: tissue fields");
         break;
      }
      T.F.records = T.mem;
      T.F.codes = storage == rec::ST_ANGLE16 ? (uint32_t *)((char *)T.mem + off) : nullptr;
      st = sim::pack_fields(dim, label, vector, &T.F, S->reg_stream);
      if (!st.ok()) break;
      // label < 0 and the storage's label width are properties of the tissue;
      // the property-table bound is checked per request
      st = sim::check_labels(T.F.label_min, T.F.label_max, (int64_t)T.F.label_max + 1, storage);
   } while (false);
   std::lock_guard<std::mutex> g(S->mu);
   if (!st.ok()) {
      if (T.mem) (void)hipFree(T.mem);
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
   // the slots drop their graphs of it when they take their next request
   (void)hipFree(it->second.mem);
   S->resident -= it->second.bytes;
   S->tissues.erase(it);
   return Status();
}

// Records the event the request waits for on the caller's stream.
Status record_after(Req &r, void *stream) {
   FPL_HIP(hipEventCreateWithFlags(&r.after, hipEventDisableTiming));
   const hipError_t e = hipEventRecord(r.after, (hipStream_t)stream);
   if (e != hipSuccess) {
      (void)hipEventDestroy(r.after);
      r.after = nullptr;
      return make_status(FPL_E_DEVICE, "serving: event on the caller's stream: %s",
                         hipGetErrorString(e));
   }
   return Status();
}

Status enqueue(fpl_server *S, const std::shared_ptr<Req> &r, int64_t *ticket) {
   std::lock_guard<std::mutex> g(S->mu);
   if (S->stop) return make_status(FPL_E_INVALID, "serving: server is shutting down");
   if (!r->model) {
      auto it = S->tissues.find(r->tissue);
      if (it == S->tissues.end())
         return make_status(FPL_E_INVALID, "unknown tissue %lld", (long long)r->tissue);
      it->second.users++;
   }
   r->ticket = S->next_ticket++;
   S->reqs[r->ticket] = r;
   S->pending.push_back(r);
   S->n_submitted++;
   *ticket = r->ticket;
   S->cv_work.notify_all();
   return Status();
}

Status submit(fpl_server *S, const fpl_request *q, int64_t *ticket) {
   if (!S || !q || !ticket || !q->setup) return make_status(FPL_E_INVALID, "null argument");
   if (!q->images) return make_status(FPL_E_INVALID, "null array");
   auto r = std::make_shared<Req>();
   int32_t lmin, lmax;
   {
      std::lock_guard<std::mutex> g(S->mu);
      auto it = S->tissues.find(q->tissue);
      if (it == S->tissues.end())
         return make_status(FPL_E_INVALID, "unknown tissue %lld", (long long)q->tissue);
      for (int k = 0; k < 3; k++) r->dim[k] = it->second.F.dim[k];
      r->storage = it->second.F.storage;
      lmin = it->second.F.label_min;
      lmax = it->second.F.label_max;
   }
   fpl_setup setup = *q->setup;
   setup.field_storage = r->storage;
   FPL_TRY(sim::prepare(&setup, r->dim, q->props, q->n_props, q->tilts, q->n_tilts, r->block));
   FPL_TRY(sim::check_labels(lmin, lmax, q->n_props, r->storage));
   r->tissue = q->tissue;
   r->n_tilts = q->n_tilts;
   r->out = q->images;
   r->need[B_BLOCK] = align256(r->block.bytes.size());
   r->need[B_IMAGES] =
       align256(sizeof(double) * (size_t)(q->n_tilts * r->dim[0] * r->dim[1] * setup.n_rho));
   DeviceGuard dg(S->device);
   FPL_TRY(record_after(*r, q->stream));
   r->t0 = Clock::now();
   Status st = enqueue(S, r, ticket);
   if (!st.ok() && r->after) (void)hipEventDestroy(r->after);
   return st;
}

Status submit_model(fpl_server *S, const fpl_model_request *q, int64_t *ticket) {
   if (!S || !q || !ticket || !q->setup) return make_status(FPL_E_INVALID, "null argument");
   if (!q->images) return make_status(FPL_E_INVALID, "null array");
   const fpl_fiber_bundles &fb = q->bundles;
   if (fb.n_bundles < 0 || fb.n_fibers < 0 || fb.n_points < 0 || !fb.bundle_ptr || !fb.layer_ptr ||
       (fb.n_points > 0 && !fb.points) || !fb.fiber_ptr)
      return make_status(FPL_E_INVALID, "null argument");
   const fpl_volume &v = q->volume;
   for (int k = 0; k < 3; k++)
      if (v.dim[k] <= 0)
         return make_status(FPL_E_INVALID, "dim.global[any] <= 0: [%lld,%lld,%lld]",
                            (long long)v.dim[0], (long long)v.dim[1], (long long)v.dim[2]);
   const int32_t storage = q->setup->field_storage;
   if (storage < 0 || storage > 2)
      return make_status(FPL_E_INVALID, "field_storage must be 0 (f32), 1 (f16) or 2 (angle16)");
   auto r = std::make_shared<Req>();
   r->model = true;
   r->vol = v;
   for (int k = 0; k < 3; k++) r->dim[k] = v.dim[k];
   r->storage = storage;
   const int64_t nb = fb.n_bundles, nl = fb.layer_ptr[nb];
   r->points = fb.points;
   r->fiber_ptr = fb.fiber_ptr;
   r->n_points = fb.n_points;
   r->n_fibers = fb.n_fibers;
   r->bundle_ptr.assign(fb.bundle_ptr, fb.bundle_ptr + nb + 1);
   r->layer_ptr.assign(fb.layer_ptr, fb.layer_ptr + nb + 1);
   if (nl > 0) {
      if (!fb.layer_scale || !fb.layer_dn || !fb.layer_mu || !fb.layer_model)
         return make_status(FPL_E_INVALID, "null argument");
      r->lscale.assign(fb.layer_scale, fb.layer_scale + nl);
      r->ldn.assign(fb.layer_dn, fb.layer_dn + nl);
      r->lmu.assign(fb.layer_mu, fb.layer_mu + nl);
      r->lmodel.assign(fb.layer_model, fb.layer_model + nl);
   }
   r->n_props = fpl_num_properties(&fb);
   if (r->n_props > rec::label_limit(storage))
      return make_status(FPL_E_INVALID, "label %lld does not fit the %s field storage",
                         (long long)r->n_props, storage == rec::ST_F16 ? "16-bit" : "15-bit");
   r->bg_mu = q->background_absorption;
   r->setup = *q->setup;
   r->rho.assign(q->setup->filter_rotations,
                 q->setup->filter_rotations + (q->setup->n_rho > 0 ? q->setup->n_rho : 0));
   r->setup.filter_rotations = r->rho.data();
   if (!q->tilts && q->n_tilts > 0) return make_status(FPL_E_INVALID, "null argument");
   r->tilts.assign(q->tilts, q->tilts + 2 * (q->n_tilts > 0 ? q->n_tilts : 0));
   r->n_tilts = q->n_tilts;
   // the setup checks of run_simulation now (with a placeholder property
   // table: the generated one is checked again before the march)
   {
      std::vector<double> props(2 * (size_t)r->n_props, 0.0);
      sim::Block B;
      FPL_TRY(sim::prepare(&r->setup, r->dim, props.data(), r->n_props, r->tilts.data(),
                           r->n_tilts, B));
      r->need[B_BLOCK] = align256(B.bytes.size());
   }
   r->out = q->images;
   const size_t n_vox = (size_t)(v.dim[0] * v.dim[1] * v.dim[2]);
   r->need[B_IMAGES] =
       align256(sizeof(double) * (size_t)(r->n_tilts * v.dim[0] * v.dim[1] * r->setup.n_rho));
   r->need[B_POINTS] = align256(sizeof(double) * 4 * (size_t)fb.n_points + 1);
   r->need[B_FPTR] = align256(sizeof(int64_t) * (size_t)(fb.n_fibers + 1));
   r->need[B_FIELDS] =
       align256(align256(n_vox * rec::record_bytes(storage)) + n_vox * rec::second_bytes(storage));
   r->scratch = gen::workspace_bytes(fb.n_points, fb.n_fibers, nb, v.dim);
   DeviceGuard dg(S->device);
   FPL_TRY(record_after(*r, q->stream));
   r->t0 = Clock::now();
   Status st = enqueue(S, r, ticket);
   if (!st.ok() && r->after) (void)hipEventDestroy(r->after);
   return st;
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
      // fail what is still queued; the slots finish what they run
      while (!S->pending.empty()) {
         std::shared_ptr<Req> r = S->pending.front();
         S->pending.pop_front();
         r->status = make_status(FPL_E_INVALID, "serving: server destroyed before the request ran");
         retire(S, r);
      }
      S->cv_work.notify_all();
   }
   for (auto &s : S->slots)
      if (s->th.joinable()) s->th.join();
   DeviceGuard dg(S->device);
   for (auto &s : S->slots) {
      (void)hipStreamSynchronize(s->stream);
      drop_graphs(S, *s);
      for (int k = 0; k < B_NUM; k++)
         if (s->buf[k]) (void)hipFree(s->buf[k]);
      ws_free(s->stream);
      if (s->h_block) (void)hipHostFree(s->h_block);
      if (s->d_flags) (void)hipFree(s->d_flags);
      if (s->h_flags) (void)hipHostFree(s->h_flags);
      (void)hipStreamDestroy(s->stream);
   }
   for (auto &t : S->tissues) (void)hipFree(t.second.mem);
   if (S->reg_stream) {
      ws_free(S->reg_stream);
      (void)hipStreamDestroy(S->reg_stream);
   }
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

extern "C" int fpl_server_submit_model(fpl_server *server, const fpl_model_request *request,
                                       int64_t *ticket, char *msg, size_t msg_len) {
   return fpl::finish(fpl::serve::submit_model(server, request, ticket), msg, msg_len);
}

extern "C" int fpl_server_wait(fpl_server *server, int64_t ticket, double timeout_s,
                               uint32_t *warn, double *latency_s, char *msg, size_t msg_len) {
   return fpl::finish(fpl::serve::wait(server, ticket, timeout_s, warn, latency_s), msg, msg_len);
}

extern "C" int fpl_server_stats(fpl_server *server, int64_t stats[9]) {
   if (!server || !stats) return FPL_E_INVALID;
   std::lock_guard<std::mutex> g(server->mu);
   int64_t busy = 0;
   for (auto &s : server->slots) busy += s->req ? 1 : 0;
   stats[0] = server->n_submitted;
   stats[1] = server->n_done;
   stats[2] = busy;
   stats[3] = (int64_t)server->pending.size();
   stats[4] = server->n_graphs;
   stats[5] = server->n_replays;
   stats[6] = server->n_generated;
   stats[7] = server->held;
   stats[8] = server->resident;
   return FPL_OK;
}

extern "C" int fpl_server_destroy(fpl_server *server) {
   return fpl::serve::destroy(server).code;
}
