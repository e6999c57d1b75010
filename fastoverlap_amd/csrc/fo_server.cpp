// Continuous-batching server for alignment requests (spherical and periodic).
//
// Requests are copied at submit and queued in arrival order.  Each of
// `streams` worker threads owns a HIP stream and a device arena; whenever a
// worker is free it forms the next batch from the class (kind + grid shape)
// of the oldest pending request: requests of that class are admitted in
// arrival order, up to max_batch pairs, while the device memory the workers'
// arenas hold stays within the byte budget.  When other batches are running
// and the batch is not full, the worker waits up to max_wait_ms for more
// arrivals of that class.  Batches on different workers overlap on the
// device.  Every per-pair computation is done by per-pair workgroups with
// fixed reduction orders, so a result does not depend on the batch it was
// served in.
//
// Admission by HBM footprint: each request carries `est`, an upper bound of
// what it adds to any batch of its class (the device bytes of a one-pair
// batch, with the per-pair selection scratch sized for the largest degree
// the class allows), so a batch's exact footprint never exceeds the sum of
// its requests' estimates.  `committed` counts the bytes the arenas hold; a
// worker admits a batch only if the arena it would grow to fits the budget
// beside the other arenas.  With a budget set, a worker with nothing to run
// releases its arena, so memory moves to where requests are waiting; a
// request that cannot fit even alone, with no other arena held, fails with
// FO_ERR_NOMEM while the server keeps serving.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fastoverlap_hip.h"
#include "fo_host.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  long long id = 0;
  fo_request r{};  // pointer fields re-targeted to the owned vectors
  std::vector<double> pos1, pos2;
  std::vector<int> gsizes, gatoms;
  int key_a = 0, key_b = 0;  // class: spherical (grid_n, invert), periodic (nk, F)
  size_t est = 0;            // upper bound of the bytes it adds to a batch of its class
  Clock::time_point t_submit, t_admit, t_done;
  bool done = false;
  fo_result res{};
  std::vector<double> X2;
  std::vector<int> perm;
};

bool same_class(const Req& a, const Req& b) {
  return a.r.kind == b.r.kind && a.key_a == b.key_a && a.key_b == b.key_b;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

int next_fast_len(int target) {
  if (target <= 6) return target;
  if (!(target & (target - 1))) return target;
  long long best = 1LL << 40;
  for (long long p5 = 1; p5 < 2LL * target; p5 *= 5)
    for (long long p35 = p5; p35 < 2LL * target; p35 *= 3) {
      long long N = p35;
      while (N < target) N *= 2;
      best = std::min(best, N);
    }
  return (int)best;
}

}  // namespace

struct fo_server {
  fo_context* ctx = nullptr;
  fo_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<long long, std::shared_ptr<Req>> reqs;
  std::deque<std::shared_ptr<Req>> pending;
  size_t committed = 0, peak = 0;  // bytes held by the worker arenas
  int running = 0;                 // batches admitted and not finished
  long long next_id = 1, completed = 0, batches = 0;
  bool stop = false;
  std::vector<std::thread> workers;
};

namespace {

// Build the batch descriptor arrays for a list of same-class requests.
struct SphDescBuf {
  std::vector<int> natoms, L, ngroups, gsizes, gatoms;
  std::vector<double> scale, pos1, pos2;
  fo_sph_batch_desc d{};
};
struct PerDescBuf {
  std::vector<int> natoms, ngroups, gsizes, gatoms;
  std::vector<double> box, scale, pos1, pos2;
  fo_per_batch_desc d{};
};

void fill_groups(const Req& q, std::vector<int>& ngroups, std::vector<int>& gsizes,
                 std::vector<int>& gatoms) {
  if (q.gsizes.empty()) {
    ngroups.push_back(1);
    gsizes.push_back(q.r.natoms);
    for (int a = 0; a < q.r.natoms; ++a) gatoms.push_back(a);
  } else {
    ngroups.push_back((int)q.gsizes.size());
    gsizes.insert(gsizes.end(), q.gsizes.begin(), q.gsizes.end());
    gatoms.insert(gatoms.end(), q.gatoms.begin(), q.gatoms.end());
  }
}

void make_sph_desc(const std::vector<std::shared_ptr<Req>>& v, SphDescBuf& b) {
  for (const auto& q : v) {
    b.natoms.push_back(q->r.natoms);
    b.L.push_back(q->r.L);
    b.scale.push_back(q->r.scale);
    fill_groups(*q, b.ngroups, b.gsizes, b.gatoms);
    b.pos1.insert(b.pos1.end(), q->pos1.begin(), q->pos1.end());
    b.pos2.insert(b.pos2.end(), q->pos2.begin(), q->pos2.end());
  }
  b.d.npairs = (int)v.size();
  b.d.grid_n = v[0]->key_a;
  b.d.natoms = b.natoms.data();
  b.d.L = b.L.data();
  b.d.scale = b.scale.data();
  b.d.ngroups = b.ngroups.data();
  b.d.group_sizes = b.gsizes.data();
  b.d.group_atoms = b.gatoms.data();
  b.d.flags = v[0]->key_b ? 0 : FO_SPH_NO_INVERT;
}

void make_per_desc(const std::vector<std::shared_ptr<Req>>& v, PerDescBuf& b) {
  for (const auto& q : v) {
    b.natoms.push_back(q->r.natoms);
    for (int k = 0; k < 3; ++k) b.box.push_back(q->r.box[k]);
    b.scale.push_back(q->r.scale);
    fill_groups(*q, b.ngroups, b.gsizes, b.gatoms);
    b.pos1.insert(b.pos1.end(), q->pos1.begin(), q->pos1.end());
    b.pos2.insert(b.pos2.end(), q->pos2.begin(), q->pos2.end());
  }
  b.d.npairs = (int)v.size();
  b.d.nk = v[0]->key_a;
  b.d.F = v[0]->key_b;
  b.d.natoms = b.natoms.data();
  b.d.box = b.box.data();
  b.d.scale = b.scale.data();
  b.d.ngroups = b.ngroups.data();
  b.d.group_sizes = b.gsizes.data();
  b.d.group_atoms = b.gatoms.data();
  b.d.niter = 10;
  b.d.flags = 0;
}

// A worker's device memory.  Its capacity is accounted in the server's
// `committed` bytes by the worker (under the server lock) before it grows.
struct Arena {
  fo_context* ctx = nullptr;
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return FO_OK;
    release();
    int rc = fo::ctx_alloc_bytes(ctx, bytes, &p);
    if (rc == FO_OK) cap = bytes;
    return rc;
  }
  void release() {
    if (p) fo::ctx_free_bytes(ctx, p, cap);
    p = nullptr;
    cap = 0;
  }
  ~Arena() { release(); }
};

// Run one batch; fills each request's result.  Returns a status code.
int run_batch(fo_server* s, std::vector<std::shared_ptr<Req>>& v, hipStream_t stream, Arena& ar) {
  const int P = (int)v.size();
  if (v[0]->r.kind == 0) {
    SphDescBuf db;
    make_sph_desc(v, db);
    fo_sph_batch* b = nullptr;
    int rc = fo::sph_batch_plan(s->ctx, &db.d, &b);
    if (rc) return rc;
    rc = fo::sph_batch_bytes(b) <= ar.cap
             ? FO_OK
             : fo::set_error(FO_ERR_NOMEM, "batch larger than its admitted footprint");
    if (!rc) rc = fo::sph_batch_bind(b, ar.p, stream);
    if (!rc) rc = fo_sph_batch_set_positions(b, db.pos1.data(), db.pos2.data());
    if (!rc) rc = fo_sph_batch_run(b, FO_STAGE_ALL, 0);
    std::vector<double> dist(P), R(3 * P), rmat(9 * P), X2(db.pos1.size());
    std::vector<int> perm(db.pos1.size() / 3), inv(P);
    const int no = fo_sph_batch_norient(b);
    std::vector<int> ind(3 * P * no);
    if (!rc) rc = fo_sph_batch_results(b, dist.data(), nullptr, X2.data(), perm.data(), R.data(),
                                       rmat.data(), inv.data());
    if (!rc) rc = fo_sph_batch_peaks(b, ind.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                     nullptr);
    fo_sph_batch_destroy(b);
    if (rc) return rc;
    size_t off = 0;
    for (int p = 0; p < P; ++p) {
      Req& q = *v[p];
      const int N = q.r.natoms;
      q.res.status = FO_OK;
      q.res.dist = dist[p];
      for (int k = 0; k < 3; ++k) q.res.R[k] = R[3 * p + k];
      for (int k = 0; k < 9; ++k) q.res.rmat[k] = rmat[9 * p + k];
      q.res.inverted = inv[p];
      for (int k = 0; k < 3 * no; ++k) q.res.ind[k] = ind[3 * p * no + k];
      q.X2.assign(X2.begin() + 3 * off, X2.begin() + 3 * (off + N));
      q.perm.assign(perm.begin() + off, perm.begin() + off + N);
      off += N;
    }
    return FO_OK;
  }
  PerDescBuf db;
  make_per_desc(v, db);
  fo_per_batch* b = nullptr;
  int rc = fo::per_batch_plan(s->ctx, &db.d, &b);
  if (rc) return rc;
  rc = fo::per_batch_bytes(b) <= ar.cap
           ? FO_OK
           : fo::set_error(FO_ERR_NOMEM, "batch larger than its admitted footprint");
  if (!rc) rc = fo::per_batch_bind(b, ar.p, stream);
  if (!rc) rc = fo_per_batch_set_positions(b, db.pos1.data(), db.pos2.data());
  if (!rc) rc = fo_per_batch_run(b, FO_PSTAGE_ALL, 0);
  std::vector<double> dist(P), disp(3 * P), X2(db.pos1.size());
  std::vector<int> perm(db.pos1.size() / 3), ind(3 * P);
  if (!rc) rc = fo_per_batch_results(b, dist.data(), nullptr, X2.data(), perm.data(), disp.data());
  if (!rc) rc = fo_per_batch_peaks(b, ind.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  fo_per_batch_destroy(b);
  if (rc) return rc;
  size_t off = 0;
  for (int p = 0; p < P; ++p) {
    Req& q = *v[p];
    const int N = q.r.natoms;
    q.res.status = FO_OK;
    q.res.dist = dist[p];
    for (int k = 0; k < 3; ++k) q.res.disp[k] = disp[3 * p + k];
    for (int k = 0; k < 3; ++k) q.res.ind[k] = ind[3 * p + k];
    q.X2.assign(X2.begin() + 3 * off, X2.begin() + 3 * (off + N));
    q.perm.assign(perm.begin() + off, perm.begin() + off + N);
    off += N;
  }
  return FO_OK;
}

void worker_main(fo_server* s) {
  (void)hipSetDevice(fo::ctx_device(s->ctx));
  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return;
  Arena ar;
  ar.ctx = s->ctx;
  const bool limited = s->cfg.hbm_budget != 0;
  const size_t budget = limited ? (size_t)s->cfg.hbm_budget : (size_t)-1;
  // hand the arena back (with a budget only: it is kept for reuse otherwise)
  auto release = [&](std::unique_lock<std::mutex>& lk) {
    if (!limited || !ar.cap) return;
    const size_t cap = ar.cap;
    lk.unlock();
    ar.release();
    lk.lock();
    s->committed -= cap;
    s->cv_work.notify_all();
  };
  for (;;) {
    std::vector<std::shared_ptr<Req>> batch;
    size_t need = 0;  // arena bytes the batch may use (sum of estimates)
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) break;
        if (s->pending.empty()) {
          release(lk);
          if (s->pending.empty() && !s->stop) s->cv_work.wait(lk);
          continue;
        }
        const std::shared_ptr<Req> head = s->pending.front();
        // bytes this worker may hold: the budget less the other arenas
        const size_t others = s->committed - ar.cap;
        const size_t avail = budget - std::min(budget, others);
        batch.clear();
        need = 0;
        for (const auto& q : s->pending) {
          if (!same_class(*q, *head)) continue;
          if (need + q->est > avail) break;
          batch.push_back(q);
          need += q->est;
          if ((int)batch.size() >= s->cfg.max_batch) break;
        }
        if (batch.empty()) {
          if (others == 0 && head->est > budget) {
            // can never fit, even alone: fail the head request
            head->res.status = FO_ERR_NOMEM;
            head->done = true;
            head->t_done = Clock::now();
            s->pending.pop_front();
            s->cv_done.notify_all();
            continue;
          }
          release(lk);  // wait for memory with nothing held
          s->cv_work.wait(lk);
          continue;
        }
        const double waited = ms_between(head->t_submit, Clock::now());
        if ((int)batch.size() < s->cfg.max_batch && s->running > 0 &&
            waited < s->cfg.max_wait_ms) {
          s->cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(
                                      s->cfg.max_wait_ms - waited));
          batch.clear();
          continue;
        }
        break;
      }
      if (s->stop && batch.empty()) break;
      const auto now = Clock::now();
      for (const auto& q : batch) {
        q->t_admit = now;
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), q));
      }
      s->running += 1;
      // the arena grows (if it must) to the admitted bytes: account first
      if (need > ar.cap) {
        s->committed += need - ar.cap;
        s->peak = std::max(s->peak, s->committed);
      }
    }
    int rc = FO_OK;
    if (need > ar.cap) {
      rc = ar.ensure(need);
      if (rc) {
        // committed counted `need` for this worker; a failed grow holds nothing
        std::lock_guard<std::mutex> lk(s->mu);
        s->committed -= need - ar.cap;
      }
    }
    if (!rc) rc = run_batch(s, batch, stream, ar);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      const auto now = Clock::now();
      for (const auto& q : batch) {
        if (rc) q->res.status = rc;
        q->res.batch_size = (int)batch.size();
        q->t_done = now;
        q->res.latency_ms = ms_between(q->t_submit, now);
        q->res.queue_ms = ms_between(q->t_submit, q->t_admit);
        q->done = true;
      }
      s->completed += (long long)batch.size();
      s->batches += 1;
      s->running -= 1;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->committed -= ar.cap;
  }
  ar.release();
  (void)hipStreamDestroy(stream);
}

}  // namespace

extern "C" {

int fo_server_create(fo_context* ctx, const fo_server_config* cfg, fo_server** out) {
  if (!ctx || !out) return fo::set_error(FO_ERR_INVALID, "NULL argument");
  auto* s = new fo_server();
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.max_batch <= 0) s->cfg.max_batch = 256;
  if (s->cfg.streams <= 0) s->cfg.streams = 2;
  if (s->cfg.max_wait_ms < 0) s->cfg.max_wait_ms = 0;
  for (int i = 0; i < s->cfg.streams; ++i) s->workers.emplace_back(worker_main, s);
  *out = s;
  return FO_OK;
}

int fo_server_destroy(fo_server* s) {
  if (!s) return FO_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& q : s->pending) {
      q->res.status = FO_ERR_CANCELLED;
      q->done = true;
    }
    s->pending.clear();
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
  return FO_OK;
}

int fo_server_submit(fo_server* s, const fo_request* r, long long* id) {
  if (!s || !r || !id) return fo::set_error(FO_ERR_INVALID, "NULL argument");
  if (r->natoms < 1 || !r->pos1 || !r->pos2) return fo::set_error(FO_ERR_INVALID, "bad request");
  auto q = std::make_shared<Req>();
  q->r = *r;
  q->pos1.assign(r->pos1, r->pos1 + 3 * (size_t)r->natoms);
  q->pos2.assign(r->pos2, r->pos2 + 3 * (size_t)r->natoms);
  int ga = 0;
  for (int g = 0; g < r->ngroups; ++g) {
    q->gsizes.push_back(r->group_sizes[g]);
    ga += r->group_sizes[g];
  }
  if (ga) q->gatoms.assign(r->group_atoms, r->group_atoms + ga);
  q->r.pos1 = q->pos1.data();
  q->r.pos2 = q->pos2.data();
  int rc;
  if (r->kind == 0) {
    if (r->L < 0) return fo::set_error(FO_ERR_INVALID, "L < 0");
    q->key_a = r->grid_n > 0 ? r->grid_n : next_pow2(2 * (r->L + 1));
    q->key_b = r->invert ? 1 : 0;
    std::vector<std::shared_ptr<Req>> one{q};
    SphDescBuf db;
    make_sph_desc(one, db);
    fo_sph_batch* b = nullptr;
    rc = fo::sph_batch_plan(s->ctx, &db.d, &b);
    if (rc) return rc;
    // every array of a batch is a sum over its pairs except the selection
    // scratch, sized per pair for the batch's largest degree: take the
    // largest degree the class's grid allows, so the sum of the requests'
    // estimates bounds any batch of them
    const long long n = q->key_a, Lc = std::min<long long>(n / 2 - 1, 63), L = r->L;
    const long long no = r->invert ? 2 : 1;
    const long long scr = (Lc + 1) * (2 * Lc + 1) + (Lc + 1) * n - (L + 1) * (2 * L + 1) -
                          (L + 1) * n;
    q->est = fo::sph_batch_bytes(b) + (size_t)(no * std::max(scr, 0LL)) * 16 + 256;
    fo_sph_batch_destroy(b);
  } else if (r->kind == 1) {
    const double V = r->box[0] * r->box[1] * r->box[2];
    if (!(q->r.scale > 0)) q->r.scale = std::cbrt(V / r->natoms) / 3.0;
    const int nk = r->nk > 0 ? r->nk : (int)std::ceil(1.3 * std::cbrt((double)r->natoms));
    q->key_a = nk;
    q->key_b = r->F > 0 ? r->F : next_fast_len(2 * (2 * nk + 1) + 1);
    std::vector<std::shared_ptr<Req>> one{q};
    PerDescBuf db;
    make_per_desc(one, db);
    fo_per_batch* b = nullptr;
    rc = fo::per_batch_plan(s->ctx, &db.d, &b);
    if (rc) return rc;
    q->est = fo::per_batch_bytes(b);
    fo_per_batch_destroy(b);
  } else {
    return fo::set_error(FO_ERR_INVALID, "kind must be 0 or 1");
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fo::set_error(FO_ERR_CANCELLED, "server stopped");
    q->id = s->next_id++;
    q->t_submit = Clock::now();
    s->reqs[q->id] = q;
    s->pending.push_back(q);
    *id = q->id;
  }
  s->cv_work.notify_one();
  return FO_OK;
}

int fo_server_wait(fo_server* s, long long id, double timeout_ms, fo_result* res, double* X2,
                   int* perm) {
  if (!s) return fo::set_error(FO_ERR_INVALID, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fo::set_error(FO_ERR_NOTFOUND, "unknown request id");
  std::shared_ptr<Req> q = it->second;
  auto ready = [&] { return q->done; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                  ready)) {
    return fo::set_error(FO_ERR_STATE, "timeout");
  }
  if (res) *res = q->res;
  if (X2 && !q->X2.empty()) std::copy(q->X2.begin(), q->X2.end(), X2);
  if (perm && !q->perm.empty()) std::copy(q->perm.begin(), q->perm.end(), perm);
  s->reqs.erase(it);
  return FO_OK;
}

int fo_server_stats(fo_server* s, long long* completed, long long* batches,
                    unsigned long long* peak_bytes) {
  if (!s) return fo::set_error(FO_ERR_INVALID, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (completed) *completed = s->completed;
  if (batches) *batches = s->batches;
  if (peak_bytes) *peak_bytes = s->peak;
  return FO_OK;
}

}  // extern "C"
