// Request server: continuous batching of asqtad EVEN/ODD solves over
// gauge configurations of several lattices that stay resident on the device
// (the QOPX_serve_* entry points).
//
// A host worker thread owns all device work.  Each (configuration, parity)
// that has requests gets an engine: a BatchSolver with `capacity` slots on a
// stream of its own, so engines of different lattices and configurations run
// concurrently on the GPU.  The worker loops over
//   collect  - engines whose iteration chunk has finished report solved
//              slots; each solution is written to the caller's buffer and
//              the slot is freed;
//   admit    - queued requests take free slots of idle engines (a new engine
//              is created when the HBM budget and the device's free memory
//              allow it, evicting idle engines of other configurations if
//              needed);
//   launch   - every idle engine with running slots starts its next chunk.
// A request therefore joins a batch that is already iterating, at the next
// chunk boundary, and leaves it as soon as it has converged.  Slot results
// do not depend on the other slots of the batch (fixed-order per-slot
// reductions), so a request gives the same answer alone and in the mix.
//
// With precision 64 the worker instead solves one request at a time with the
// FP64 reference algorithm (the one-at-a-time FP64 loop served solves are
// compared against).
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "batch.h"
#include "handles.h"

namespace qop {
namespace {

double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3 };

struct Request {
  int cfg = 0, eo = 0, max_iter = 0;
  double mass = 0, tol = 0;
  const double *src = nullptr;
  double *out = nullptr;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int state = RQ_QUEUED;
  int iter = 0, nupdate = 0, reason = 0;
  double true_rsq = 0;
};

struct Config {
  Lattice *lat = nullptr;
  Links links;
  size_t bytes = 0;
};

struct Engine {
  int cfg = 0, eo = 0;
  std::unique_ptr<BatchSolver> bs;  // null for the FP64 one-at-a-time engine
  hipStream_t st = nullptr;
  bool own_stream = false;
  DevVec stage;           // FP64 full vector: source in, solution out
  double *raw = nullptr;  // device raw vector [V][3][2]
  std::vector<long> slot_req;
  int nloaded = 0;
  size_t bytes = 0;
  double last_used = 0;
};

size_t link_bytes(const Lattice *lat, int prec) {
  const size_t n = (size_t)2 * 16 * 9 * lat->Vh;
  if (prec == P64) return n * 16;
  if (prec == P32) return n * 8;
  return n * 4 + (size_t)2 * 16 * lat->Vh * sizeof(float);
}

size_t engine_bytes(const Lattice *lat, int cap, int prec) {
  const size_t vh = lat->Vh;
  const size_t fp64 = 48 * vh;                                  // one parity, FP64
  const size_t sl = prec == P16 ? 12 * vh + 4 * vh : 24 * vh;   // one parity, sloppy
  const size_t nch = (vh + 63) / 64;
  size_t per_slot = 5 * fp64 + 3 * sl + 24 * vh + 3 * nch * sizeof(double) + 256;
  if (prec == P64) per_slot = 0;
  return (size_t)cap * per_slot + 2 * fp64 /* stage */ + 2 * fp64 /* raw */ + (1u << 20);
}

}  // namespace

class Server {
 public:
  Server(int capacity, int prec, size_t budget) : cap_(capacity), prec_(prec), budget_(budget) {
    size_t fr = 0, tot = 0;
    QOP_CHECK_HIP(hipMemGetInfo(&fr, &tot));
    reserve_ = std::min<size_t>((size_t)2 << 30, tot / 32);
    if (!budget_) budget_ = fr > reserve_ ? fr - reserve_ : 0;
    worker_ = std::thread([this] { run(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    worker_.join();
    for (auto &e : engines_) drop_engine(*e);
    for (auto &c : cfgs_) c->links.free_all();
  }

  int add_config(const int L[4], double *const U[4], double *fat[4], double *lng[4],
                 const QOP_asqtad_coeffs_t *c, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
    auto cf = std::make_unique<Config>();
    cf->lat = lattice_get(L);
    cf->bytes = link_bytes(cf->lat, P64) + (prec_ == P64 ? 0 : link_bytes(cf->lat, prec_));
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (used_ + cf->bytes > budget_) return -1;
      used_ += cf->bytes;
    }
    cf->links.lat = cf->lat;
    const bool fit = guard_oom("QOPX_serve_add_config", nullptr, [&] {
      if (U) links_from_gauge64(cf->links, U, c, r0, bc, sign);
      else links_from_raw64(cf->links, fat, lng);
      if (prec_ != P64) cf->links.get(prec_);
    });
    if (!fit) {
      sync();
      cf->links.free_all();
      std::lock_guard<std::mutex> lk(mu_);
      used_ -= cf->bytes;
      return -1;
    }
    std::lock_guard<std::mutex> lk(mu_);
    cfgs_.push_back(std::move(cf));
    return (int)cfgs_.size() - 1;
  }

  long submit(int cfg, int eo, double mass, double tol, int max_iter, const double *src,
              double *out) {
    std::lock_guard<std::mutex> lk(mu_);
    if (cfg < 0 || cfg >= (int)cfgs_.size() || (eo != 0 && eo != 1) || !src || !out ||
        !(tol > 0) || stop_)
      return -1;
    const long id = next_id_++;
    Request &r = reqs_[id];
    r.cfg = cfg;
    r.eo = eo;
    r.mass = mass;
    r.tol = tol;
    r.max_iter = max_iter > 0 ? max_iter : 20000;
    r.src = src;
    r.out = out;
    r.t_submit = now();
    queue_.push_back(id);
    arrivals_++;
    cv_work_.notify_all();
    return id;
  }

  // 0 converged, 1 stopped unconverged, -1 unknown or failed request
  int wait(long id, QOP_resid_arg_t *res, double *latency) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -1;
    cv_done_.wait(lk, [&] { return it->second.state >= RQ_DONE; });
    Request r = it->second;
    reqs_.erase(it);
    if (r.state != RQ_DONE) return -1;
    if (res) {
      res->final_rsq = r.true_rsq;
      res->final_rel = 0;
      res->final_iter = r.iter;
      res->final_restart = r.nupdate;
    }
    if (latency) *latency = r.t_done - r.t_submit;
    return r.reason == 1 ? 0 : 1;
  }

  int test(long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -1;
    return it->second.state >= RQ_DONE ? 1 : 0;
  }

  void stats(double *s) {
    std::lock_guard<std::mutex> lk(mu_);
    int running = 0;
    for (auto &kv : reqs_) running += kv.second.state == RQ_RUNNING;
    s[0] = (double)completed_;
    s[1] = (double)queue_.size();
    s[2] = running;
    s[3] = (double)nengines_;
    s[4] = (double)used_;
    s[5] = (double)budget_;
    s[6] = (double)chunks_;
    s[7] = (double)failed_;
  }

 private:
  void run() {
    for (;;) {
      bool progressed = false;
      for (auto &e : engines_) progressed |= collect(*e);
      progressed |= admit();
      bool inflight = false;
      for (auto &e : engines_) {
        if (!e->bs) continue;
        if (!e->bs->in_flight() && e->bs->nactive() > 0) {
          e->bs->launch_chunk();
          chunks_++;
          progressed = true;
        }
        inflight |= e->bs->in_flight();
      }
      bool loaded = false;
      for (auto &e : engines_) loaded |= e->nloaded > 0;
      std::unique_lock<std::mutex> lk(mu_);
      if (stop_ && !loaded && (queue_.empty() || !progressed)) {
        // shutting down: what is still queued could not be admitted
        for (long id : queue_) {
          reqs_[id].state = RQ_FAILED;
          failed_++;
        }
        queue_.clear();
        cv_done_.notify_all();
        return;
      }
      if (progressed) continue;
      if (!inflight && !loaded) {
        // nothing on the device: sleep until a request arrives
        const long seen = arrivals_;
        cv_work_.wait_for(lk, std::chrono::milliseconds(5),
                          [&] { return stop_ || arrivals_ != seen; });
      } else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }

  // Finished slots of an engine whose chunk is done: write results out.
  bool collect(Engine &e) {
    if (!e.bs || e.nloaded == 0 || !e.bs->chunk_done()) return false;
    std::vector<int> done;
    e.bs->collect(done);
    const Lattice *lat = e.stage.lat;
    const long Vh = lat->Vh, b = e.eo ? Vh : 0;
    for (int slot : done) {
      SlotResult r;
      e.bs->result(slot, e.stage, true, &r);
      launch_vec_to_raw(lat, e.stage, e.raw, e.eo, e.st);
      const long id = e.slot_req[slot];
      double *out;
      {
        std::lock_guard<std::mutex> lk(mu_);
        out = reqs_[id].out;
      }
      QOP_CHECK_HIP(hipMemcpyAsync(out + 6 * b, e.raw + 6 * b, (size_t)Vh * 6 * sizeof(double),
                                   hipMemcpyDeviceToHost, e.st));
      QOP_CHECK_HIP(hipStreamSynchronize(e.st));
      e.bs->release(slot);
      e.slot_req[slot] = 0;
      e.nloaded--;
      e.last_used = now();
      std::lock_guard<std::mutex> lk(mu_);
      Request &q = reqs_[id];
      q.iter = r.iter;
      q.nupdate = r.nupdate;
      q.reason = r.reason;
      q.true_rsq = r.true_rsq;
      q.t_done = now();
      q.state = RQ_DONE;
      completed_++;
    }
    if (!done.empty()) cv_done_.notify_all();
    return !done.empty();
  }

  Engine *find_engine(int cfg, int eo) {
    for (auto &e : engines_)
      if (e->cfg == cfg && (e->eo == eo || !e->bs)) return e.get();
    return nullptr;
  }

  void drop_engine(Engine &e) {
    e.bs.reset();
    devvec_free(e.stage);
    if (e.raw) QOP_CHECK_HIP(hipFree(e.raw));
    e.raw = nullptr;
    if (e.st && e.own_stream) QOP_CHECK_HIP(hipStreamDestroy(e.st));
    e.st = nullptr;
  }

  // Create an engine for (cfg, eo) if memory allows, evicting idle engines
  // (least recently used first).  Returns null if it cannot be made now.
  Engine *make_engine(int cfg, int eo, bool *never) {
    Config &c = *cfgs_[cfg];
    const size_t need = engine_bytes(c.lat, cap_, prec_);
    *never = false;
    for (;;) {
      size_t used;
      {
        std::lock_guard<std::mutex> lk(mu_);
        used = used_;
      }
      size_t fr = 0, tot = 0;
      QOP_CHECK_HIP(hipMemGetInfo(&fr, &tot));
      if (used + need <= budget_ && fr >= need + reserve_) break;
      // evict the least recently used idle engine
      int victim = -1;
      for (int i = 0; i < (int)engines_.size(); i++) {
        Engine &e = *engines_[i];
        if (e.nloaded || (e.bs && e.bs->in_flight())) continue;
        if (victim < 0 || e.last_used < engines_[victim]->last_used) victim = i;
      }
      if (victim < 0) {
        // only a request that cannot fit beside the resident links never runs
        size_t links = 0;
        for (auto &cf : cfgs_) links += cf->bytes;
        *never = links + need > budget_;
        return nullptr;
      }
      drop_engine(*engines_[victim]);
      {
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= engines_[victim]->bytes;
      }
      engines_.erase(engines_.begin() + victim);
      nengines_--;
    }
    auto e = std::make_unique<Engine>();
    e->cfg = cfg;
    e->eo = eo;
    e->bytes = need;
    {
      std::lock_guard<std::mutex> lk(mu_);
      used_ += need;
    }
    try {
      devvec_alloc(e->stage, c.lat, P64);
      dev_alloc(&e->raw, (size_t)c.lat->V * 6 * sizeof(double));
      if (prec_ != P64) {
        QOP_CHECK_HIP(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
        e->own_stream = true;
        e->bs = std::make_unique<BatchSolver>(c.links, eo, cap_, prec_, 0.1, e->st);
        e->slot_req.assign(cap_, 0);
      } else {
        e->st = stream();  // the reference solver runs on the library stream
        e->slot_req.assign(1, 0);
      }
    } catch (const DeviceOutOfMemory &) {
      // the device holds less than the budget promised (another process, or
      // fragmentation): give the memory back; the next admission pass evicts
      // idle engines first, and a request with no engine of ours to evict
      // fails
      drop_engine(*e);
      {
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= need;
      }
      *never = engines_.empty();
      return nullptr;
    }
    sync();
    e->last_used = now();
    engines_.push_back(std::move(e));
    nengines_++;
    return engines_.back().get();
  }

  // Load queued requests into free slots of idle engines (FIFO per engine).
  bool admit() {
    std::vector<long> ids;
    {
      std::lock_guard<std::mutex> lk(mu_);
      ids.assign(queue_.begin(), queue_.end());
    }
    bool any = false;
    for (long id : ids) {
      Request q;
      {
        std::lock_guard<std::mutex> lk(mu_);
        q = reqs_[id];
      }
      Engine *e = find_engine(q.cfg, q.eo);
      if (!e) {
        bool never = false;
        e = make_engine(q.cfg, q.eo, &never);
        if (!e) {
          if (never) fail(id);
          continue;
        }
      }
      if (!e->bs) {
        solve_fp64(*e, id, q);
        any = true;
        continue;
      }
      if (e->bs->in_flight()) continue;
      const int slot = e->bs->free_slot();
      if (slot < 0) continue;
      const Lattice *lat = e->stage.lat;
      const long Vh = lat->Vh, b = q.eo ? Vh : 0;
      QOP_CHECK_HIP(hipMemcpyAsync(e->raw + 6 * b, q.src + 6 * b, (size_t)Vh * 6 * sizeof(double),
                                   hipMemcpyHostToDevice, e->st));
      launch_raw_to_vec(lat, e->raw, e->stage, q.eo, e->st);
      e->bs->load(slot, e->stage, nullptr, q.mass, q.tol, q.max_iter);
      e->slot_req[slot] = id;
      e->nloaded++;
      e->last_used = now();
      std::lock_guard<std::mutex> lk(mu_);
      remove_queued(id);
      Request &r = reqs_[id];
      r.state = RQ_RUNNING;
      r.t_admit = now();
      any = true;
    }
    return any;
  }

  // The FP64 one-at-a-time loop: reference algorithm, library stream.
  void solve_fp64(Engine &e, long id, const Request &q) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      remove_queued(id);
      Request &r = reqs_[id];
      r.state = RQ_RUNNING;
      r.t_admit = now();
    }
    Config &c = *cfgs_[q.cfg];
    const Lattice *lat = c.lat;
    const long Vh = lat->Vh, b = q.eo ? Vh : 0;
    hipStream_t st = stream();
    QOP_CHECK_HIP(hipMemcpyAsync(e.raw + 6 * b, q.src + 6 * b, (size_t)Vh * 6 * sizeof(double),
                                 hipMemcpyHostToDevice, st));
    launch_raw_to_vec(lat, e.raw, e.stage, q.eo, st);
    InvArg ia{1e9, q.max_iter, q.max_iter, 5, q.eo};
    ResArg ra{q.tol * q.tol, 0, 0, 0, 0, 0};
    const bool fit = guard_oom("QOPX_serve (FP64 solve)", nullptr, [&] {
      TmpVec out(c.lat, P64);
      invert(c.links, &ia, &ra, q.mass, out.v, e.stage, nullptr, nullptr);
      launch_vec_to_raw(lat, out.v, e.raw, q.eo, st);
      QOP_CHECK_HIP(hipMemcpyAsync(q.out + 6 * b, e.raw + 6 * b, (size_t)Vh * 6 * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
      sync();
    });
    if (!fit) {
      fail(id);
      return;
    }
    std::lock_guard<std::mutex> lk(mu_);
    Request &r = reqs_[id];
    r.iter = ra.final_iter;
    r.nupdate = ra.final_restart;
    r.reason = ra.final_rsq <= ra.rsqmin ? 1 : 2;
    r.true_rsq = ra.final_rsq;
    r.t_done = now();
    r.state = RQ_DONE;
    completed_++;
    cv_done_.notify_all();
  }

  void fail(long id) {
    std::lock_guard<std::mutex> lk(mu_);
    remove_queued(id);
    Request &r = reqs_[id];
    r.state = RQ_FAILED;
    r.t_done = now();
    failed_++;
    cv_done_.notify_all();
  }

  void remove_queued(long id) {  // mu_ held
    for (auto it = queue_.begin(); it != queue_.end(); ++it)
      if (*it == id) {
        queue_.erase(it);
        return;
      }
  }

  const int cap_, prec_;
  size_t budget_, reserve_ = 0, used_ = 0;
  std::vector<std::unique_ptr<Config>> cfgs_;
  std::vector<std::unique_ptr<Engine>> engines_;  // worker thread only
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<long> queue_;
  std::map<long, Request> reqs_;
  long next_id_ = 1;
  long completed_ = 0, failed_ = 0, arrivals_ = 0;
  std::atomic<long> chunks_{0}, nengines_{0};  // written by the worker only
  bool stop_ = false;
  std::thread worker_;
};

}  // namespace qop

struct QOPX_Server_struct {
  qop::Server *s;
};

extern "C" {

QOPX_Server *QOPX_serve_create(int capacity, int prec, unsigned long long max_bytes) {
  if (capacity < 1 || capacity > 256) return nullptr;
  if (prec != 16 && prec != 32 && prec != 64) return nullptr;
  const int p = prec == 64 ? qop::P64 : prec == 32 ? qop::P32 : qop::P16;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return nullptr;
  return new QOPX_Server{new qop::Server(capacity, p, (size_t)max_bytes)};
}

int QOPX_serve_add_gauge(QOPX_Server *s, const int L[4], QOP_D_Real *U[4],
                         QOP_asqtad_coeffs_t *coeffs, int *r0, QOP_bc_t *bc,
                         QOP_staggered_sign_t *sign) {
  if (!s || !L || !U || !coeffs) return -1;
  return s->s->add_config(L, U, nullptr, nullptr, coeffs, r0, bc, sign);
}

int QOPX_serve_add_links(QOPX_Server *s, const int L[4], QOP_D_Real *fat[4], QOP_D_Real *lng[4]) {
  if (!s || !L || !fat) return -1;
  return s->s->add_config(L, nullptr, fat, lng, nullptr, nullptr, nullptr, nullptr);
}

long QOPX_serve_submit(QOPX_Server *s, int config, QOP_evenodd_t evenodd, double mass, double tol,
                       int max_iter, const QOP_D_Real *src, QOP_D_Real *out) {
  if (!s) return -1;
  return s->s->submit(config, (int)evenodd, mass, tol, max_iter, src, out);
}

int QOPX_serve_wait(QOPX_Server *s, long request, QOP_resid_arg_t *res, double *latency) {
  if (!s) return -1;
  return s->s->wait(request, res, latency);
}

int QOPX_serve_test(QOPX_Server *s, long request) {
  if (!s) return -1;
  return s->s->test(request);
}

void QOPX_serve_stats(QOPX_Server *s, double stats[8]) {
  if (s && stats) s->s->stats(stats);
}

void QOPX_serve_destroy(QOPX_Server *s) {
  if (!s) return;
  delete s->s;
  delete s;
}

}  // extern "C"
