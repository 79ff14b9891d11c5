// Continuous-batching serving loop (see serving.h).
#include "serving.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <random>
#include <stdexcept>

#include "onionpir_gpu.h"

namespace onionpir {

std::vector<size_t> plan_batch_limits(size_t ndb, const std::function<size_t(size_t, size_t)> &ws, size_t fixed,
                                      size_t budget, size_t max_batch) {
  if (ndb == 0) throw std::invalid_argument("serving needs at least one database");
  if (max_batch == 0) throw std::invalid_argument("max_batch must be positive");
  std::vector<size_t> out(ndb, max_batch);
  if (budget == 0) return out;
  if (fixed >= budget) throw std::invalid_argument("HBM budget does not cover the resident client keys");
  const size_t share = (budget - fixed) / ndb;
  for (size_t d = 0; d < ndb; d++) {
    if (ws(d, 1) > share) throw std::invalid_argument("HBM budget share too small for one query's workspace");
    size_t lo = 1, hi = max_batch;  // largest B in [1, max_batch] with ws(d, B) <= share
    while (lo < hi) {
      const size_t mid = lo + (hi - lo + 1) / 2;
      if (ws(d, mid) <= share)
        lo = mid;
      else
        hi = mid - 1;
    }
    out[d] = lo;
  }
  return out;
}

ServingLoop::ServingLoop(std::vector<PirServer *> dbs, const ServeConfig &cfg) {
  t0_ = std::chrono::steady_clock::now();
  if (cfg.lanes == 0) throw std::invalid_argument("lanes must be positive");
  // one mutex per distinct context: a context serves one host thread at a time
  std::map<opir_ctx *, std::mutex *> by_ctx;
  size_t fixed = 0;
  for (PirServer *s : dbs) {
    if (!s) throw std::invalid_argument("null PirServer");
    opir_ctx *ctx = s->gpu_scheme()->ctx();
    if (!by_ctx.count(ctx)) {
      ctx_mus_.push_back(std::make_unique<std::mutex>());
      by_ctx[ctx] = ctx_mus_.back().get();
      if (cfg.keyed_clients) {  // planned key count at this context's key size
        size_t kb = 0;
        s->gpu_scheme()->check(opir_client_key_bytes(ctx, s->get_params().get_expan_height(), &kb));
        fixed += cfg.keyed_clients * kb;
      } else {  // the keys resident now
        size_t kb = 0;
        s->gpu_scheme()->check(opir_ctx_resident_key_bytes(ctx, &kb));
        fixed += kb;
      }
    }
    auto lane = std::make_unique<DbLane>();
    lane->srv = s;
    lane->lanes.resize(cfg.lanes);
    lane->lanes[0].ctx = ctx;
    lane->lanes[0].db = s->gpu_db();
    lane->lanes[0].ctx_mu = by_ctx[ctx];
    lanes_.push_back(std::move(lane));
  }
  auto release = [&]() {
    for (auto &d : lanes_)
      for (auto &l : d->lanes)
        if (l.clone) {
          if (l.db) opir_db_destroy(l.db);
          if (l.ctx) opir_ctx_destroy(l.ctx);
          l.db = nullptr;
          l.ctx = nullptr;
        }
  };
  try {
    // further lanes: a context clone (own streams, the parent's keys) and a
    // view of the same resident database (own workspace)
    for (auto &d : lanes_)
      for (size_t k = 1; k < d->lanes.size(); k++) {
        auto &l = d->lanes[k];
        const GpuScheme &gs = *d->srv->gpu_scheme();
        l.clone = true;
        gs.check(opir_ctx_clone(d->lanes[0].ctx, &l.ctx));
        ctx_mus_.push_back(std::make_unique<std::mutex>());
        l.ctx_mu = ctx_mus_.back().get();
        const int rc = opir_db_view(d->lanes[0].db, l.ctx, &l.db);
        if (rc != OPIR_OK) {
          const std::string msg = opir_last_error(l.ctx);
          throw std::runtime_error("database view: " + msg);
        }
      }
    // every lane of database d reserves a workspace of the admitted batch
    auto ws = [&](size_t d, size_t B) {
      size_t b = 0;
      lanes_[d]->srv->gpu_scheme()->check(opir_db_workspace_bytes(lanes_[d]->srv->gpu_db(), B, &b));
      return b * cfg.lanes;
    };
    if (!cfg.max_batch_db.empty() && cfg.max_batch_db.size() != lanes_.size())
      throw std::invalid_argument("max_batch_db needs one bound per database");
    caps_.assign(lanes_.size(), cfg.max_batch);
    for (size_t d = 0; d < cfg.max_batch_db.size(); d++) {
      if (cfg.max_batch_db[d] == 0) throw std::invalid_argument("max_batch_db bounds must be positive");
      caps_[d] = std::min(cfg.max_batch, cfg.max_batch_db[d]);
    }
    limits_ = plan_batch_limits(lanes_.size(), ws, fixed, cfg.hbm_budget, cfg.max_batch);
    for (size_t d = 0; d < lanes_.size(); d++) limits_[d] = std::min(limits_[d], caps_[d]);
    budget_ = cfg.hbm_budget;
    fixed_ = fixed;
    shrink_after_ = cfg.hbm_budget ? cfg.shrink_after_s : -1;  // no budget: every lane holds max_batch
    for (size_t d = 0; d < lanes_.size(); d++) {  // reserve every workspace before the first request
      lanes_[d]->peak_B = limits_[d];
      for (auto &l : lanes_[d]->lanes) {
        std::lock_guard<std::mutex> g(*l.ctx_mu);
        const int rc = opir_db_reserve(l.db, limits_[d]);
        if (rc != OPIR_OK) {
          const std::string msg = opir_last_error(l.ctx);
          if (rc == OPIR_EINVAL) throw std::invalid_argument(msg);
          throw std::runtime_error(msg);
        }
        l.res_B = limits_[d];
        l.res_bytes = lane_ws(d, l.res_B);
        reserved_total_ += l.res_bytes;
      }
    }
  } catch (...) {
    release();
    throw;
  }
  for (size_t d = 0; d < lanes_.size(); d++)
    for (size_t k = 0; k < lanes_[d]->lanes.size(); k++)
      lanes_[d]->lanes[k].th = std::thread(&ServingLoop::worker, this, d, k);
}

ServingLoop::~ServingLoop() { stop(); }

size_t ServingLoop::lane_ws(size_t d, size_t B) const {
  size_t b = 0;
  lanes_[d]->srv->gpu_scheme()->check(opir_db_workspace_bytes(lanes_[d]->srv->gpu_db(), B, &b));
  return b;
}

size_t ServingLoop::reserved_batch(size_t db, size_t lane) const {
  std::lock_guard<std::mutex> g(bmu_);
  return lanes_.at(db)->lanes.at(lane).res_B;
}
size_t ServingLoop::peak_batch(size_t db) const {
  std::lock_guard<std::mutex> g(bmu_);
  return lanes_.at(db)->peak_B;
}
size_t ServingLoop::grows() const {
  std::lock_guard<std::mutex> g(bmu_);
  return grows_;
}
size_t ServingLoop::trims() const {
  std::lock_guard<std::mutex> g(bmu_);
  return trims_;
}

// Grow lane li of database d toward `want` queries with budget no lane holds:
// the largest B in (res_B, want] whose extra workspace fits, and not past its
// share while another lane reclaims. The bytes are committed before the
// allocation and returned if it fails. A lane that stays below min(want, its
// share) flags a reclaim.
void ServingLoop::try_grow(size_t d, size_t li, size_t want) {
  DbLane::Lane &me = lanes_[d]->lanes[li];
  want = std::min(want, caps_[d]);
  size_t target = 0, bytes = 0, old_bytes = 0;
  {
    std::lock_guard<std::mutex> g(bmu_);
    const size_t cap = budget_ > fixed_ + reserved_total_ ? budget_ - fixed_ - reserved_total_ : 0;
    size_t lo = me.res_B, hi = want;
    if (others_reclaim(d, li)) hi = std::min(hi, std::max(me.res_B, limits_[d]));  // not past its share
    while (lo < hi) {  // ws is non-decreasing in B
      const size_t mid = lo + (hi - lo + 1) / 2;
      if (lane_ws(d, mid) - me.res_bytes <= cap)
        lo = mid;
      else
        hi = mid - 1;
    }
    me.reclaim = lo < std::min(want, limits_[d]);
    if (lo <= me.res_B) return;
    target = lo;
    bytes = lane_ws(d, target);
    old_bytes = me.res_bytes;
    reserved_total_ += bytes - old_bytes;
  }
  int rc;
  {
    std::lock_guard<std::mutex> g(*me.ctx_mu);
    rc = opir_db_reserve(me.db, target);
    if (rc != OPIR_OK) (void)opir_db_trim(me.db, me.res_B);  // back to what the plan holds
  }
  std::lock_guard<std::mutex> g(bmu_);
  if (rc != OPIR_OK) {
    reserved_total_ -= bytes - old_bytes;
    me.reclaim = me.res_B < std::min(want, limits_[d]);
    return;
  }
  me.res_B = target;
  me.res_bytes = bytes;
  lanes_[d]->peak_B = std::max(lanes_[d]->peak_B, target);
  grows_++;
}

// Trim lane li of database d to B queries' workspace (when it holds more).
void ServingLoop::trim_to(size_t d, size_t li, size_t B) {
  DbLane::Lane &me = lanes_[d]->lanes[li];
  {
    std::lock_guard<std::mutex> g(bmu_);
    if (me.res_B <= B) return;
  }
  int rc;
  {
    std::lock_guard<std::mutex> g(*me.ctx_mu);
    rc = opir_db_trim(me.db, B);
  }
  if (rc != OPIR_OK) return;  // the workspace stays as it was
  const size_t bytes = lane_ws(d, B);
  std::lock_guard<std::mutex> g(bmu_);
  reserved_total_ -= me.res_bytes - bytes;
  me.res_B = B;
  me.res_bytes = bytes;
  trims_++;
}

bool ServingLoop::others_reclaim(size_t d, size_t li) const {
  for (size_t e = 0; e < lanes_.size(); e++)
    for (size_t k = 0; k < lanes_[e]->lanes.size(); k++)
      if ((e != d || k != li) && lanes_[e]->lanes[k].reclaim) return true;
  return false;
}

double ServingLoop::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t ServingLoop::submit(size_t db, size_t client_id, std::vector<u64> query) {
  if (db >= lanes_.size()) throw std::out_of_range("unknown database index");
  const PirParams &p = lanes_[db]->srv->get_params();
  const size_t N = p.get_poly_degree(), K = p.K();
  if (query.size() != 2 * N * K) throw std::invalid_argument("query must hold [c0 K*N][c1 K*N]");
  // values must be reduced (limb by limb): a malformed query fails here, on
  // its own, rather than the batch it would have joined
  const auto &mods = p.get_rns_mods();
  for (size_t i = 0; i < query.size(); i++)
    if (query[i] >= mods[(i / N) % K]) throw std::invalid_argument("query value not reduced mod its modulus");
  std::lock_guard<std::mutex> g(mu_);
  if (stop_) throw std::runtime_error("serving loop is stopped");
  const uint64_t t = next_ticket_++;
  lanes_[db]->queue.push_back(Req{t, client_id, std::move(query), now()});
  submitted_++;
  cv_work_.notify_all();
  return t;
}

size_t ServingLoop::poll(std::vector<ServeResult> &out) {
  std::lock_guard<std::mutex> g(mu_);
  const size_t n = done_.size();
  for (auto &r : done_) out.push_back(std::move(r));
  done_.clear();
  return n;
}

void ServingLoop::wait_idle() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return completed_ == submitted_; });
}

void ServingLoop::stop() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
    cv_work_.notify_all();
  }
  for (auto &d : lanes_)
    for (auto &l : d->lanes)
      if (l.th.joinable()) l.th.join();
  for (auto &d : lanes_)
    for (auto &l : d->lanes)
      if (l.clone) {
        if (l.db) opir_db_destroy(l.db);
        if (l.ctx) opir_ctx_destroy(l.ctx);
        l.db = nullptr;
        l.ctx = nullptr;
      }
}

void ServingLoop::worker(size_t d, size_t li) {
  DbLane &lane = *lanes_[d];
  PirServer &srv = *lane.srv;
  DbLane::Lane &me = lane.lanes[li];
  const size_t N = srv.get_params().get_poly_degree(), QN = N * srv.get_params().K();  // response / query values per poly
  const bool replan = shrink_after_ >= 0;
  const auto idle_wait = std::chrono::duration<double>(std::max(shrink_after_, 0.0));
  std::vector<uint64_t> ids, q, r;
  for (;;) {
    std::vector<Req> batch;
    size_t want = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      const auto has_work = [&] { return stop_ || !lane.queue.empty(); };
      if (replan && reserved_batch(d, li) > 1) {
        if (!cv_work_.wait_for(lk, idle_wait, has_work)) {  // idle: give the workspace back
          lk.unlock();
          {
            std::lock_guard<std::mutex> g(bmu_);
            me.reclaim = false;
          }
          trim_to(d, li, 1);
          continue;
        }
      } else {
        cv_work_.wait(lk, has_work);
      }
      if (lane.queue.empty()) return;  // stopping, nothing left for this database
      want = lane.queue.size();
    }
    if (replan) {
      if (want > reserved_batch(d, li)) {
        try_grow(d, li, want);  // flags a reclaim if it stays below its share
      } else {
        std::lock_guard<std::mutex> g(bmu_);
        me.reclaim = false;
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      const size_t b = std::min(lane.queue.size(), reserved_batch(d, li));
      for (size_t i = 0; i < b; i++) {
        batch.push_back(std::move(lane.queue.front()));
        lane.queue.pop_front();
      }
    }
    if (batch.empty()) continue;  // another lane took them
    const size_t B = batch.size();
    ids.resize(B);
    q.resize(B * 2 * QN);
    r.resize(B * 2 * N);
    for (size_t i = 0; i < B; i++) {
      ids[i] = batch[i].client_id;
      std::memcpy(&q[i * 2 * QN], batch[i].query.data(), 2 * QN * sizeof(u64));
    }
    const double ts = now();
    int rc;
    std::string err;
    std::vector<int> rcs(B, OPIR_OK);
    std::vector<std::string> errs(B);
    {
      std::lock_guard<std::mutex> g(*me.ctx_mu);
      rc = opir_answer_batch(me.db, B, ids.data(), q.data(), r.data());
      if (rc != OPIR_OK) err = opir_last_error(me.ctx);
      if ((rc == OPIR_EINVAL || rc == OPIR_ERANGE) && B > 1) {
        // one request (an unknown client, say) failed the whole batch: answer
        // them one by one so that only the bad ones carry the error
        rc = OPIR_OK;
        for (size_t i = 0; i < B; i++) {
          rcs[i] = opir_answer_batch(me.db, 1, &ids[i], &q[i * 2 * QN], &r[i * 2 * N]);
          if (rcs[i] != OPIR_OK) errs[i] = opir_last_error(me.ctx);
        }
      } else if (rc != OPIR_OK) {
        std::fill(rcs.begin(), rcs.end(), rc);
        std::fill(errs.begin(), errs.end(), err);
      }
    }
    const double td = now();
    std::vector<ServeResult> res(B);
    for (size_t i = 0; i < B; i++) {
      ServeResult &x = res[i];
      x.ticket = batch[i].ticket;
      x.db = d;
      x.client_id = batch[i].client_id;
      x.t_submit = batch[i].t_submit;
      x.t_start = ts;
      x.t_done = td;
      x.batch = B;
      if (rcs[i] == OPIR_OK)
        x.resp.assign(r.begin() + i * 2 * N, r.begin() + (i + 1) * 2 * N);
      else
        x.error = errs[i];
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      for (auto &x : res) done_.push_back(std::move(x));
      completed_ += B;
      cv_done_.notify_all();
    }
    if (replan) {  // above its share while another lane waits below its own: give the extra back
      bool give = false;
      {
        std::lock_guard<std::mutex> g(bmu_);
        give = me.res_B > limits_[d] && others_reclaim(d, li);
      }
      if (give) trim_to(d, li, limits_[d]);
    }
  }
}

LoadRun run_poisson(ServingLoop &loop, const std::vector<PoolEntry> &pool, double lambda, double duration, u64 seed) {
  if (pool.empty()) throw std::invalid_argument("empty request pool");
  if (!(lambda > 0) || !(duration > 0)) throw std::invalid_argument("lambda and duration must be positive");
  std::mt19937_64 rng(seed);
  std::exponential_distribution<double> gap(lambda);
  std::uniform_int_distribution<size_t> pick(0, pool.size() - 1);
  std::map<uint64_t, size_t> origin;
  const auto start = std::chrono::steady_clock::now();
  double t = gap(rng);
  size_t sent = 0;
  while (t < duration) {
    std::this_thread::sleep_until(start + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                              std::chrono::duration<double>(t)));
    const size_t k = pick(rng);
    origin[loop.submit(pool[k].db, pool[k].client_id, pool[k].query)] = k;
    sent++;
    t += gap(rng);
  }
  loop.wait_idle();
  LoadRun run;
  std::vector<ServeResult> all;
  loop.poll(all);
  double first = 1e300, last = 0;
  for (auto &x : all) {
    auto it = origin.find(x.ticket);
    if (it == origin.end()) continue;  // a request submitted outside this run
    first = std::min(first, x.t_submit);
    last = std::max(last, x.t_done);
    run.pool_index.push_back(it->second);
    run.results.push_back(std::move(x));
  }
  run.offered_qps = sent / duration;
  run.span_s = run.results.empty() ? 0 : last - first;
  run.achieved_qps = run.span_s > 0 ? run.results.size() / run.span_s : 0;
  return run;
}

}  // namespace onionpir
